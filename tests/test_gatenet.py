"""The gate-net program (rtlflow_amd/gatenet.py, the host side of
include/rf_gatenet.h): levels, output cone, records, RAM tables and the mtask
partition checked on the CPU -- a NumPy model of the blob hashes as the C
oracle does, by levels (the bit-plane form's order, with the clk = 1 eval
limited to the output cone) and by mtasks (the per-lane graph form's)."""
import numpy as np
import pytest

import gatenet_sim
import oracle
from rtlflow_amd import gatenet as gn
from rtlflow_amd.designs.netlist import generate


@pytest.fixture(scope="module")
def small():
    _, tb = generate(n_gates=900, n_ff=90, n_pi=100, n_po=40, depth=9, seed=0x6A7E5, n_rams=2,
                     ram_depth=48, ram_width=20)
    return tb.to_bytes()


def test_table_round_trip(small):
    t = gn.parse_table(small)
    assert t.to_bytes() == small


def test_levels_and_cone(small):
    t = gn.parse_table(small)
    lvl = gn.gate_levels(t)
    first = t.first_gate
    full_lvl = np.concatenate([np.zeros(first, np.int32), lvl])
    ins = t.gate_in.astype(np.int64)
    for g in range(t.n_gates):
        k = 3 if t.gate_type[g] == 4 else 2
        assert lvl[g] == 1 + max(full_lvl[ins[g, :k]])
    cone = gn.output_cone(t, lvl)
    # brute force: gates reachable backwards from the outputs
    want = np.zeros(t.n_gates, bool)
    stack = [int(p) for p in t.po if p >= first]
    while stack:
        n = stack.pop()
        g = n - first
        if want[g]:
            continue
        want[g] = True
        k = 3 if t.gate_type[g] == 4 else 2
        stack += [int(i) for i in ins[g, :k] if i >= first]
    assert (cone == want).all() and 0 < cone.sum() < t.n_gates


@pytest.mark.parametrize("threads", [2, 8])
def test_program_matches_oracle_by_levels_and_by_mtasks(small, threads):
    t = gn.parse_table(small)
    prog = gn.program(t, threads=threads, chunk=64)
    seed = 0x52544C4600000005
    budgets = [9, 0, 4, 9, 1, 6]
    want = list(oracle.netlist_run(small, seed, len(budgets), budgets=budgets))
    assert gatenet_sim.run(prog, seed, budgets, "levels") == want
    assert gatenet_sim.run(prog, seed, budgets, "mtasks") == want


def test_partition_granularity_and_order(small):
    t = gn.parse_table(small)
    cg = gn.ChunkGraph(t, gn.gate_levels(t), 64)
    coarse, pc = gn.partition_mtasks(cg, 2)
    fine, pf = gn.partition_mtasks(cg, 32)
    assert len(coarse) < len(fine) <= len(cg.chunks)
    for mem, preds in ((coarse, pc), (fine, pf)):
        assert sorted(c for m in mem for c in m) == list(range(len(cg.chunks)))
        assert all(p < k for k, ps in enumerate(preds) for p in ps)
        # every chunk edge is respected by the mtask order
        where = {c: k for k, m in enumerate(mem) for c in m}
        assert all(where[a] <= where[b] for a, b in cg.edges)
    f = cg.features(coarse[0])
    assert len(f) == 6 and f[0] > 0 and f[3] >= f[1]
