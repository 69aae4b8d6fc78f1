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
