"""Per-lane VCD: the writer's format on synthetic lane records, read back
(no GPU); and, on the GPU, a traced lane's output values re-hashed from the
VCD give the oracle's hash of that stimulus in every form."""
import numpy as np
import pytest

import oracle
from rtlflow_amd import vcd

MAN = {"name": "t", "slots": {"C": 3, "S": 1, "I": 2, "Q": 0}, "lane_major_bytes": 0, "mems": [],
       "ports": [{"name": "clk", "dir": 2}],
       "vars": [{"name": "clk", "width": 1, "cls": 0, "words": 1, "loc": 0},
                {"name": "a", "width": 4, "cls": 0, "words": 1, "loc": 1},
                {"name": "__Vclklast__clk", "width": 1, "cls": 0, "words": 1, "loc": 2},
                {"name": "s", "width": 12, "cls": 1, "words": 1, "loc": 0},
                {"name": "w", "width": 40, "cls": 2, "words": 2, "loc": 0}]}


def _frame(clk, a, s, w):
    return bytes([clk, a, 0]) + int(s).to_bytes(2, "little") + int(w & 0xFFFFFFFF).to_bytes(4, "little") + \
        int(w >> 32).to_bytes(4, "little")


def test_vcd_format_round_trip():
    frames = [_frame(1, 3, 0x123, 0x12_3456_789A), _frame(1, 3, 0x124, 0x12_3456_789A),
              _frame(1, 9, 0x124, 0xFF_0000_0001)]
    text = vcd.write_vcd(MAN, frames, lane=5, stim=5)
    assert "$enddefinitions $end" in text and "$var wire 40" in text and "__Vclklast" not in text
    ch = vcd.read_vcd(text)
    assert vcd.values_at_cycles(ch, "a", 3) == [3, 3, 9]
    assert vcd.values_at_cycles(ch, "s", 3) == [0x123, 0x124, 0x124]
    assert vcd.values_at_cycles(ch, "w", 3) == [0x12_3456_789A, 0x12_3456_789A, 0xFF_0000_0001]
    # only changes after the first dump: a changes once
    assert len(ch["a"]) == 2
    # the clock falls mid-cycle and is high at each cycle's end
    assert (15, 0) in ch["clk"] and vcd.values_at_cycles(ch, "clk", 3) == [1, 1, 1]


@pytest.mark.gpu
def test_traced_lane_rehashes_to_the_oracle(gpu):
    from rtlflow_amd.designs import registry
    from rtlflow_amd.model import MODE_BITPLANE, MODE_FUSED, MODE_GRAPH
    name, n, cycles, lane = "net_tiny", 70, 20, 37
    seed = registry.SEED_CFG[2]
    cd = registry.get(name)
    want = oracle.netlist_run(registry.table_bytes(name), seed, nstim=1, cycles=cycles, stim0=lane)[0]
    texts = []
    for mode in (MODE_GRAPH, MODE_FUSED, MODE_BITPLANE):
        with cd.model(n) as m:
            m.begin(seed)
            frames = m.trace(lane, cycles, mode)
            assert (m.hashes()[lane] == want)
        text = vcd.write_vcd(cd.manifest, frames, lane=lane, stim=lane)
        po = vcd.values_at_cycles(vcd.read_vcd(text), "po", cycles)
        h = 0xCBF29CE484222325
        width = [p for p in cd.manifest["ports"] if p["name"] == "po"][0]["width"]
        nb = (width + 7) // 8
        nb = 1 if nb <= 1 else 2 if nb <= 2 else 4 if nb <= 4 else 8
        for v in po:
            h = oracle.fnv(int(v).to_bytes(nb, "little"), h)
        assert h == int(want), mode
        texts.append(text)
    assert texts[0] == texts[1] == texts[2]
