"""GPU parity: generated gfx950 kernels vs the CPU oracle through the C ABI."""
import numpy as np
import pytest

import oracle
from rtlflow_amd.designs import registry
from rtlflow_amd.model import MODE_FUSED, MODE_GRAPH

SEED = registry.SEED_CFG[2]


@pytest.mark.gpu
@pytest.mark.parametrize("name,n,cycles", [("net_tiny", 300, 40), ("net_small", 1000, 30)])
@pytest.mark.parametrize("mode", [MODE_GRAPH, MODE_FUSED])
def test_netlist_protocol(gpu, name, n, cycles, mode):
    cd = registry.get(name)
    ref = oracle.netlist_run(registry.table_bytes(name), SEED, nstim=n, cycles=cycles)
    with cd.model(n) as m:
        m.begin(SEED)
        m.run_cycles(cycles, mode)
        got = m.hashes()
    assert (got == ref).all(), f"{int((got != ref).sum())} of {n} stimuli differ"


@pytest.mark.gpu
def test_netlist_split_runs_and_stim_offset(gpu):
    # running 10 + 15 cycles equals 25 cycles; stimulus ids offset by stim0
    cd = registry.get("net_tiny")
    ref = oracle.netlist_run(registry.table_bytes("net_tiny"), SEED, nstim=70, cycles=25, stim0=5)
    with cd.model(70) as m:
        m.begin(SEED, stim0=5)
        m.run_cycles(10, MODE_GRAPH)
        m.run_cycles(15, MODE_FUSED)
        got = m.hashes()
    assert (got == ref).all()


@pytest.mark.gpu
def test_netlist_batched_queue(gpu):
    cd = registry.get("net_small")
    rng = np.random.default_rng(4)
    budgets = rng.integers(0, 60, size=500).astype(np.uint32)
    ref = oracle.netlist_run(registry.table_bytes("net_small"), SEED, nstim=500, budgets=budgets)
    with cd.model(128) as m:
        got = m.submit_queue(SEED, budgets)
    assert (got == ref).all()


@pytest.mark.gpu
def test_netlist_eval_api_matches_protocol(gpu):
    """Per-lane set / eval / get through the C ABI (the Vtop surface) drives
    the same cycles as the on-device protocol."""
    cd = registry.get("net_tiny")
    n, cycles = 64, 12
    ref = oracle.netlist_run(registry.table_bytes("net_tiny"), SEED, nstim=n, cycles=cycles)
    with cd.model(n) as m:
        h = np.full(n, 0xcbf29ce484222325, dtype=np.uint64)
        for c in range(cycles):
            pi = np.array([oracle.stim(SEED, s, c, 0) & 0xffff for s in range(n)], np.uint16)
            m.set_port("pi", pi)
            m.set_port("clk", np.zeros(n, np.uint8))
            m.eval()
            m.set_port("clk", np.ones(n, np.uint8))
            m.eval()
            po = m.get_port("po")
            for s in range(n):
                h[s] = oracle.fnv(int(po[s]).to_bytes(2, "little"), int(h[s]))
    assert (h == ref).all()


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["net5s", "net5s_learned"])
def test_net5s_rams_batched_queue_and_graph(gpu, name):
    """Configuration 5's structure at one tenth of its gates and flip-flops:
    eight 1,024 x 64-bit synchronous RAMs and 128 inputs drawn in two chunks.
    Both partitions (instruction-count heuristic and learned) against the
    oracle: the batched queue with lognormal budgets, and the graph form."""
    from rtlflow_amd.designs import rv32i
    if name not in registry.load_index():
        pytest.skip(f"{name} is not built")
    seed = registry.SEED_CFG[5]
    tb = registry.table_bytes("net5s")
    budgets = rv32i.budgets(seed, 96, median=300, lo=0, hi=1500)
    ref = oracle.netlist_run(tb, seed, nstim=96, budgets=budgets)
    cd = registry.get(name)
    with cd.model(64) as m:
        got = m.submit_queue(seed, budgets)
    assert (got == ref).all(), f"{int((got != ref).sum())} of 96 stimuli differ"
    ref = oracle.netlist_run(tb, seed, nstim=128, cycles=60)
    with cd.model(128) as m:
        m.begin(seed)
        m.run_cycles(60, MODE_GRAPH)
        got = m.hashes()
    assert (got == ref).all()


@pytest.mark.gpu
def test_net5s_bitplane(gpu):
    """The same design in the bit-plane form (its RAMs as memory ports): the
    protocol and the batched workload against the oracle."""
    from rtlflow_amd.designs import rv32i
    from rtlflow_amd.model import MODE_BITPLANE
    seed = registry.SEED_CFG[5]
    tb = registry.table_bytes("net5s")
    cd = registry.get("net5s")
    assert cd.manifest["bitplane"]["supported"], cd.manifest["bitplane"]
    ref = oracle.netlist_run(tb, seed, nstim=200, cycles=80)
    with cd.model(200) as m:
        m.begin(seed)
        m.run_cycles(80, MODE_BITPLANE)
        assert (m.hashes() == ref).all()
    budgets = rv32i.budgets(seed, 150, median=300, lo=0, hi=1500)
    ref = oracle.netlist_run(tb, seed, nstim=150, budgets=budgets)
    with cd.model(128) as m:
        got = m.submit_queue(seed, budgets, mode=MODE_BITPLANE)
    assert (got == ref).all(), f"{int((got != ref).sum())} of 150 stimuli differ"
