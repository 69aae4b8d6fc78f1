"""GPU parity of the bit-plane form (64 stimuli per 64-bit word): hashes
equal the oracle's, lane state equals the per-lane forms', ragged lane
counts, done lanes, hand-over between forms and synchronous memories."""
import numpy as np
import pytest

import ir_interp
import oracle
from rtlflow_amd.designs import registry
from rtlflow_amd.designs.fuzz import verilog
from rtlflow_amd.model import MODE_BITPLANE, MODE_FUSED, MODE_GRAPH
from rtlflow_amd.verilog import elaborate


def _net(name, lanes, cycles, seed):
    return oracle.netlist_run(registry.table_bytes(name), seed, lanes, cycles)


@pytest.mark.gpu
@pytest.mark.parametrize("name,lanes,cycles", [("net_tiny", 1000, 40), ("net_small", 777, 30),
                                               ("net2", 2048, 24), ("net_ram", 333, 50)])
def test_bitplane_netlist(gpu, name, lanes, cycles):
    seed = 0xB17 + lanes
    cd = registry.get(name)
    assert cd.manifest["bitplane"]["supported"], cd.manifest["bitplane"]
    print(f"bitplane {name} lanes={lanes} cycles={cycles}", flush=True)
    with cd.model(lanes) as m:
        m.begin(seed)
        m.run_cycles(cycles, MODE_BITPLANE)
        got = m.hashes()
    want = _net(name, lanes, cycles, seed)
    assert (got == want).all(), f"{int((got != want).sum())} of {lanes} differ"


@pytest.mark.gpu
@pytest.mark.parametrize("threads", [128, 256, 1024])
@pytest.mark.parametrize("name", ["net_small", "net2", "net5s"])
def test_bitplane_workgroup_sizes(gpu, name, threads):
    """Every workgroup size (register-resident gate records or streamed)
    gives the oracle's hashes."""
    lanes, cycles, seed = 200, 21, 0x7EAD + threads
    cd = registry.get(name)
    print(f"bitplane {name} threads={threads}", flush=True)
    with cd.model(lanes, bitplane_threads=threads) as m:
        m.begin(seed)
        m.run_cycles(cycles, MODE_BITPLANE)
        got = m.hashes()
    want = _net(name, lanes, cycles, seed)
    assert (got == want).all(), f"{int((got != want).sum())} of {lanes} differ"


@pytest.mark.gpu
def test_bitplane_state_matches_fused(gpu):
    cd = registry.get("net_small")
    lanes, cycles, seed = 130, 17, 0x5EED
    with cd.model(lanes) as a, cd.model(lanes) as b:
        a.begin(seed)
        a.run_cycles(cycles, MODE_BITPLANE)
        b.begin(seed)
        b.run_cycles(cycles, MODE_FUSED)
        for lane in (0, 63, 64, 129):
            assert a.lane_state(lane) == b.lane_state(lane), lane
        assert (a.hashes() == b.hashes()).all()


@pytest.mark.gpu
def test_bitplane_done_lanes_and_handover(gpu):
    cd = registry.get("net_tiny")
    lanes, seed = 200, 0xD0
    done = np.zeros(lanes, np.uint8)
    done[[0, 5, 64, 199]] = 1
    with cd.model(lanes) as m:
        m.begin(seed)
        m.run_cycles(10, MODE_BITPLANE)
        m.set_done(done)
        m.run_cycles(7, MODE_BITPLANE)
        m.set_done(np.zeros(lanes, np.uint8))
        m.run_cycles(5, MODE_FUSED)
        m.run_cycles(4, MODE_GRAPH)
        m.run_cycles(3, MODE_BITPLANE)
        got = m.hashes()
    full = _net("net_tiny", lanes, 29, seed)
    assert (got[done == 0] == full[done == 0]).all()
    # done lanes sat out 7 cycles without advancing their cycle counter:
    # they equal a 22-cycle run of the same stimulus
    short = _net("net_tiny", lanes, 22, seed)
    assert (got[done == 1] == short[done == 1]).all()


@pytest.mark.gpu
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_bitplane_random_designs(gpu, seed):
    cd = registry.get(f"bfuzz{seed}")
    assert cd.manifest["bitplane"]["supported"], cd.manifest["bitplane"]
    n, cycles, stim_seed = 70, 9, 0xAB + seed
    ref = np.array(ir_interp.protocol(elaborate(verilog(seed, name=f"bfuzz{seed}", bitlevel=True, depth=4)),
                                      stim_seed, n, cycles), dtype=np.uint64)
    for mode in (MODE_BITPLANE, MODE_FUSED):
        with cd.model(n) as m:
            m.begin(stim_seed)
            m.run_cycles(cycles, mode)
            got = m.hashes()
        assert (got == ref).all(), f"mode {mode}: {int((got != ref).sum())} of {n} differ"


@pytest.mark.gpu
def test_bitplane_memory_ports_match_fused(gpu):
    """Synchronous memory ports (guarded writes, two writers of one memory,
    read of the element written at the same edge, lane-major and
    element-major memories): hashes and whole lane state, memories included,
    equal the fused form's, across done lanes and hand-over between forms."""
    cd = registry.get("rams")
    assert cd.manifest["bitplane"]["supported"], cd.manifest["bitplane"]
    lanes, seed = 150, 0x3A3
    done = np.zeros(lanes, np.uint8)
    done[[1, 64, 149]] = 1
    with cd.model(lanes) as a, cd.model(lanes) as b:
        for m, mode in ((a, MODE_BITPLANE), (b, MODE_FUSED)):
            m.begin(seed)
            m.run_cycles(23, mode)
            m.set_done(done)
            m.run_cycles(9, mode)
            m.set_done(np.zeros(lanes, np.uint8))
            m.run_cycles(4, MODE_GRAPH)
            m.run_cycles(11, mode)
        for lane in (0, 1, 63, 64, 65, 149):
            assert a.lane_state(lane) == b.lane_state(lane), lane
        assert (a.hashes() == b.hashes()).all()


@pytest.mark.gpu
@pytest.mark.parametrize("name,lanes", [("net_tiny", 256), ("net_ram", 130), ("net_ram", 64)])
def test_bitplane_batched_budgets(gpu, name, lanes):
    """The batched workload in the bit-plane form: uneven budgets (zero
    included), more stimuli than lanes (several rounds, the last one ragged),
    the queue order given or longest first; against the oracle and the
    per-lane batched queue."""
    seed = 0xB0D + lanes
    rng = np.random.default_rng(lanes)
    nstim = 300
    budgets = rng.integers(0, 90, size=nstim).astype(np.uint32)
    budgets[:3] = [0, 1, 89]
    want = oracle.netlist_run(registry.table_bytes(name), seed, nstim, budgets=budgets)
    cd = registry.get(name)
    print(f"bitplane queue {name} lanes={lanes} nstim={nstim}", flush=True)
    with cd.model(lanes) as m:
        got = m.submit_queue(seed, budgets, mode=MODE_BITPLANE)
        assert (got == want).all(), f"{int((got != want).sum())} of {nstim} differ"
        order = rng.permutation(nstim)
        got = m.submit_queue(seed, budgets, order=order, mode=MODE_BITPLANE)
        assert (got == want).all()
        assert (m.submit_queue(seed, budgets, mode=MODE_FUSED) == want).all()


@pytest.mark.gpu
def test_bitplane_refused_for_asynchronous_memory_reads(gpu):
    cd = registry.get("alu")
    assert not cd.manifest["bitplane"]["supported"]
    with cd.model(64) as m:
        m.begin(1)
        with pytest.raises(Exception, match="no bit-plane program"):
            m.run_cycles(1, MODE_BITPLANE)


@pytest.mark.gpu
@pytest.mark.parametrize("lanes,threads", [(49153, 512), (65536, 512)])
def test_bitplane_more_groups_than_workgroup_slots(gpu, lanes, threads):
    """net2 at more 64-lane groups than the chip holds at once (three
    workgroups of 512 threads per CU, 768 slots): groups of the second round
    and the ragged last group match the oracle, on a strided sample."""
    cd = registry.get("net2")
    cycles, seed = 100, registry.SEED_CFG[2]
    ng = (lanes + 63) // 64
    print(f"bitplane net2 lanes={lanes} groups={ng} cycles={cycles}", flush=True)
    with cd.model(lanes, bitplane_threads=threads, fused_chunk=cycles) as m:
        m.begin(seed)
        m.run_cycles(cycles, MODE_BITPLANE)
        got = m.hashes()
    tb = registry.table_bytes("net2")
    for g in sorted(x for x in {0, 383, 767, 768, 769, ng // 2, ng - 2, ng - 1} if x < ng):
        s0 = g * 64
        k = min(64, lanes - s0)
        want = oracle.netlist_run(tb, seed, k, cycles, stim0=s0)
        assert (got[s0:s0 + k] == want).all(), f"group {g}: {int((got[s0:s0 + k] != want).sum())} differ"
