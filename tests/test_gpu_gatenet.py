"""GPU parity of the gate-net engine (include/rf_gatenet.h) against the C
oracle: a small netlist of configuration 5's shape in both forms with ragged
lane counts and uneven budgets, and configuration 5 at its stated size
(500,000 gates, 50,000 flip-flops, eight 1,024 x 64-bit RAMs, 128 inputs,
4,096 stimuli with lognormal budgets) on a sample of stimuli, with the
heuristic and the learned partitions."""
import numpy as np
import pytest

import oracle
from rtlflow_amd.designs import registry, rv32i
from rtlflow_amd.gatenet import GateNet

SEED = registry.SEED_CFG[5]


@pytest.mark.gpu
@pytest.mark.parametrize("lanes", [130, 1024])
@pytest.mark.parametrize("mode", [GateNet.GRAPH, GateNet.BITPLANE])
def test_gatenet_small_matches_oracle(gpu, lanes, mode):
    tb = registry.gatenet_table("net5_small")
    prog = registry.gatenet_program("net5_small")
    rng = np.random.default_rng(lanes + mode)
    budgets = rng.integers(0, 40, size=lanes).astype(np.uint32)
    print(f"gatenet net5_small lanes={lanes} mode={mode}", flush=True)
    with GateNet(prog, lanes) as g:
        got, _ = g.run(mode, SEED, budgets)
    want = oracle.netlist_run(tb, SEED, lanes, budgets=budgets)
    assert (got == want).all(), f"{int((got != want).sum())} of {lanes} differ"


def _net5_budgets(n, median):
    return rv32i.budgets(SEED, n, median=median, lo=1, hi=3 * median)


@pytest.mark.gpu
def test_config5_full_size_bitplane_matches_oracle(gpu):
    """4,096 stimuli, lognormal budgets (median 24): the first and last 16
    stimuli against the oracle."""
    n = 4096
    tb = registry.gatenet_table("net5")
    prog = registry.gatenet_program("net5")
    budgets = _net5_budgets(n, 24)
    print(f"gatenet net5 bitplane lanes={n} max budget={int(budgets.max())}", flush=True)
    with GateNet(prog, n) as g:
        got, ms = g.run(GateNet.BITPLANE, SEED, budgets)
    for s0 in (0, n - 16):
        want = oracle.netlist_run(tb, SEED, 16, stim0=s0, budgets=budgets[s0:s0 + 16])
        assert (got[s0:s0 + 16] == want).all(), s0


@pytest.mark.gpu
@pytest.mark.parametrize("schedule", ["heuristic", "learned"])
def test_config5_full_size_graph_form_matches_oracle(gpu, schedule):
    if schedule not in registry.gatenet_meta("net5")["schedules"]:
        pytest.skip("no learned partition chosen for net5")
    n = 4096
    tb = registry.gatenet_table("net5")
    prog = registry.gatenet_program("net5", schedule)
    budgets = _net5_budgets(n, 4)
    print(f"gatenet net5 graph {schedule} lanes={n} max budget={int(budgets.max())}", flush=True)
    with GateNet(prog, n) as g:
        got, ms = g.run(GateNet.GRAPH, SEED, budgets)
    for s0 in (0, n - 16):
        want = oracle.netlist_run(tb, SEED, 16, stim0=s0, budgets=budgets[s0:s0 + 16])
        assert (got[s0:s0 + 16] == want).all(), s0
