"""GPU: the multi-wavefront fused kernel (rf_config.fused_multiwave). One
workgroup serves 64 stimuli with several wavefronts, each running its share
of every mtask level; hashes and whole lane state equal the one-wavefront
fused kernel's and the oracle's, over ragged lane counts, runs split across
calls, done lanes and a hand-over between forms."""
import numpy as np
import pytest

import oracle
from rtlflow_amd.designs import registry, rv32i
from rtlflow_amd.model import MODE_FUSED, MODE_GRAPH


def _oracle(name, seed, n, cycles):
    if name == "alu":
        return oracle.alu_run(seed, n, cycles)
    if name == "systolic":
        return oracle.systolic_run(seed, n, cycles)
    return oracle.netlist_run(registry.table_bytes(name), seed, n, cycles)


@pytest.mark.gpu
@pytest.mark.parametrize("name,n,runs", [("alu", 300, (37, 23)), ("systolic", 200, (40, 1, 19)),
                                         ("net_small", 333, (25, 5)), ("mix", 130, (17,))])
def test_multiwave_matches_fused_and_oracle(gpu, name, n, runs):
    cd = registry.get(name)
    assert cd.manifest["mw_waves"] >= 2, cd.manifest["mw_waves"]
    seed = 0x3A7E + n
    print(f"multiwave {name} n={n} runs={runs} waves={cd.manifest['mw_waves']}", flush=True)
    with cd.model(n, fused_multiwave=1) as a, cd.model(n) as b:
        for m in (a, b):
            m.begin(seed)
            for c in runs:
                m.run_cycles(c, MODE_FUSED)
        got = a.hashes()
        assert (got == b.hashes()).all()
        for lane in (0, 63, 64, n - 1):
            assert a.lane_state(lane) == b.lane_state(lane), lane
    if name != "mix":
        want = _oracle(name, seed, n, sum(runs))
        assert (got == want).all(), f"{int((got != want).sum())} of {n} differ"


@pytest.mark.gpu
def test_multiwave_rv32i_images(gpu):
    seed = registry.SEED_CFG[3]
    n, cycles = 150, 700
    progs = rv32i.programs(seed, np.arange(n))
    ref = oracle.rv32i_run(progs, seed, [cycles] * n)
    cd = registry.get("rv32i")
    assert cd.manifest["mw_waves"] >= 2
    with cd.model(n, fused_multiwave=1) as m:
        m.mem_write("imem", progs)
        m.begin(seed)
        m.run_cycles(cycles, MODE_FUSED)
        got = m.hashes()
    assert (got == ref).all(), f"{int((got != ref).sum())} of {n} differ"


@pytest.mark.gpu
def test_multiwave_done_lanes_and_handover(gpu):
    cd = registry.get("alu")
    n, seed = 200, 0xD1
    done = np.zeros(n, np.uint8)
    done[[0, 5, 64, 199]] = 1
    with cd.model(n, fused_multiwave=1) as a, cd.model(n) as b:
        for m in (a, b):
            m.begin(seed)
            m.run_cycles(10, MODE_FUSED)
            m.set_done(done)
            m.run_cycles(7, MODE_FUSED)
            m.set_done(np.zeros(n, np.uint8))
            m.run_cycles(3, MODE_GRAPH)
            m.run_cycles(6, MODE_FUSED)
        assert (a.hashes() == b.hashes()).all()
        for lane in (0, 5, 6, 64, 199):
            assert a.lane_state(lane) == b.lane_state(lane), lane


@pytest.mark.gpu
def test_tuned_choice_matches_oracle(gpu):
    """The autotuner's choice for the systolic array (its library, launch
    length and, when chosen, the multi-wavefront kernel) against the oracle
    over a ragged lane count and runs split across launches."""
    from rtlflow_amd.autotune import choice
    if "systolic_tuned" not in registry.load_index():
        pytest.skip("systolic_tuned not built")
    t = choice("systolic")
    k = t["knobs"]
    cd = registry.get("systolic_tuned")
    n, seed = 520, 0x5C0
    print(f"tuned systolic knobs={k} waves={cd.manifest['mw_waves']}", flush=True)
    with cd.model(n, fused_chunk=k.get("chunk", 512), lane_pad=k.get("lane_pad", 0),
                  fused_multiwave=k.get("multiwave", 0)) as m:
        m.begin(seed)
        m.run_cycles(33, MODE_FUSED)
        m.run_cycles(40, MODE_FUSED)
        got = m.hashes()
    want = oracle.systolic_run(seed, n, 73)
    assert (got == want).all(), f"{int((got != want).sum())} of {n} differ"
