"""GPU: the bit-plane group rotation. When a design's 64-lane groups
outnumber the workgroups one launch runs, rf_run_cycles rotates the groups
through launches of one full round each; a group's runs before its last
write back only its state. With the window forced small (rf_config
bitplane_slots), every lane's hash and whole state (memories included) equal
a run without rotation and the oracle: ragged lane counts, cycle counts that
do not divide into the runs, consecutive calls, done lanes, a hand-over
between forms, synchronous memories and user input streams with capture."""
import numpy as np
import pytest

import ir_interp
import oracle
from rtlflow_amd.designs import registry
from rtlflow_amd.designs.mix import VERILOG as MIX
from rtlflow_amd.model import MODE_BITPLANE, MODE_FUSED, MODE_GRAPH
from rtlflow_amd.verilog import elaborate
from test_gpu_streams import _dev, _host

SEED = registry.SEED_CFG[2]


@pytest.mark.gpu
@pytest.mark.parametrize("name,lanes,slots,runs", [
    ("net_small", 453, 7, (5, 1, 1999)),  # 8 groups, the last one ragged
    ("net_small", 700, 10, (61, 1500)),
    ("net2", 1000, 15, (1333,)),
    ("net_tiny", 64 * 9, 8, (2000,)),
    ("net_tiny", 64 * 9, 4, (1001,)),
])
def test_rotation_matches_single_launch_and_oracle(gpu, name, lanes, slots, runs):
    cd = registry.get(name)
    print(f"rotation {name} lanes={lanes} slots={slots} runs={runs}", flush=True)
    with cd.model(lanes, bitplane_slots=slots) as a, cd.model(lanes) as b:
        for m in (a, b):
            m.begin(SEED)
            for c in runs:
                m.run_cycles(c, MODE_BITPLANE)
        got = a.hashes()
        assert (got == b.hashes()).all()
        assert a.checkpoint() == b.checkpoint()
        assert a.stats()["bitplane_launches"] > b.stats()["bitplane_launches"]
    want = oracle.netlist_run(registry.table_bytes(name), SEED, lanes, sum(runs))
    assert (got == want).all(), f"{int((got != want).sum())} of {lanes} differ"


@pytest.mark.gpu
def test_rotation_with_done_lanes_handover_and_memories(gpu):
    """Synchronous memory ports, lanes set done between runs, and runs in
    the fused and graph forms in between: state and hashes equal the same
    sequence without rotation."""
    cd = registry.get("rams")
    lanes = 330
    done = np.zeros(lanes, np.uint8)
    done[[1, 64, 200, 329]] = 1
    with cd.model(lanes, bitplane_slots=5) as a, cd.model(lanes) as b:
        for m in (a, b):
            m.begin(SEED)
            m.run_cycles(1200, MODE_BITPLANE)
            m.set_done(done)
            m.run_cycles(9, MODE_BITPLANE)
            m.set_done(np.zeros(lanes, np.uint8))
            m.run_cycles(4, MODE_GRAPH)
            m.run_cycles(3, MODE_FUSED)
            m.run_cycles(11, MODE_BITPLANE)
        assert (a.hashes() == b.hashes()).all()
        assert a.checkpoint() == b.checkpoint()


@pytest.mark.gpu
def test_rotation_with_input_stream_and_capture(gpu):
    """A user input stream and the output capture across rotated launches:
    hashes and every captured word equal a run without rotation, and the IR
    interpreter's for a slice of stimuli."""
    import torch
    cd = registry.get("mix")
    man = cd.manifest
    nin, nout = man["streams"]["in_words"], man["streams"]["out_words"]
    n, stim0, cyc = 200, 3, 600
    rng = np.random.default_rng(7)
    words = rng.integers(0, 2**64, size=(cyc, nin, n), dtype=np.uint64, endpoint=False)
    xin = _dev(words)
    caps, hs = [], []
    for slots in (3, 0):
        cap = torch.zeros((cyc, nout, n), dtype=torch.int64, device="cuda")
        with cd.model(n, bitplane_slots=slots, fused_chunk=cyc) as m:
            m.begin(SEED, stim0=stim0)
            m.set_input_stream(xin, stim0=stim0)
            m.set_output_capture(cap, stim0=stim0)
            m.run_cycles(cyc, MODE_BITPLANE)
            hs.append(m.hashes())
            assert (m.stats()["bitplane_launches"] > 1) == (slots != 0)
        caps.append(_host(cap))
    assert (hs[0] == hs[1]).all()
    assert (caps[0] == caps[1]).all()
    k0, k = 120, 8
    want = np.zeros((cyc, nout, k), np.uint64)
    ref = ir_interp.protocol(elaborate(MIX), SEED, k, cyc, stim0=stim0 + k0,
                             inputs=np.ascontiguousarray(words[:, :, k0:k0 + k]), capture=want)
    assert [int(v) for v in hs[0][k0:k0 + k]] == ref
    assert (caps[0][:, :, k0:k0 + k] == want).all()


@pytest.mark.gpu
def test_rotation_at_the_headline_shape(gpu):
    """net2 at 65,536 stimuli (1,024 groups against 768 workgroup slots at
    three per CU): the automatic rotation and one launch over all groups
    (window wider than the groups) give the same hash for every stimulus and
    the same lane state, and a strided sample matches the oracle."""
    cd = registry.get("net2")
    lanes, cycles = 65536, 300
    print(f"rotation net2 lanes={lanes} cycles={cycles}", flush=True)
    with cd.model(lanes, fused_chunk=cycles) as a, cd.model(lanes, fused_chunk=cycles, bitplane_slots=4096) as b:
        for m in (a, b):
            m.begin(SEED)
            m.run_cycles(cycles, MODE_BITPLANE)
        got = a.hashes()
        assert a.stats()["bitplane_launches"] > 1
        assert b.stats()["bitplane_launches"] == 1
        assert (got == b.hashes()).all()
        for lane in (0, 767 * 64 + 5, 768 * 64, lanes - 1):
            assert a.lane_state(lane) == b.lane_state(lane), lane
    tb = registry.table_bytes("net2")
    for g in (0, 511, 767, 768, 1023):
        want = oracle.netlist_run(tb, SEED, 64, cycles, stim0=g * 64)
        assert (got[g * 64:(g + 1) * 64] == want).all(), f"group {g}"
