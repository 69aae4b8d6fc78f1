"""Partitioning with a cost table (the learned cost path) changes the mtask
graph, never the results: host-build hashes stay equal to the oracle's."""
import numpy as np
import pytest

import oracle
from rtlflow_amd.compiler import Compiled, Options
from rtlflow_amd.costmodel import FeatureExtractor, TableCost, encode
from rtlflow_amd.designs.netlist import generate
from rtlflow_amd.hostsim import HostDesign
from rtlflow_amd.verilog import elaborate


def test_cost_table_changes_partition_not_results():
    src, table = generate(n_gates=600, n_ff=60, n_pi=16, n_po=16, depth=12, seed=21, name="cm")
    base = Compiled(elaborate(src), Options(threads=4, bitplane=False))
    rng = np.random.default_rng(0)
    costs = [int(c) for c in rng.integers(1, 400, len(base.sch.vertices))]
    learned = Compiled(elaborate(src), Options(threads=4, bitplane=False, cost_fn=TableCost(costs)))
    assert [v.cost for v in learned.sch.vertices] == costs
    assert [sorted(m.vids) for m in learned.mtasks] != [sorted(m.vids) for m in base.mtasks]
    want = oracle.netlist_run(table.to_bytes(), 0x77, 50, 20)
    for comp in (base, learned):
        got = HostDesign(comp).protocol(50, 0x77, 20)
        assert (got == want).all()


def test_cost_table_size_is_checked():
    src, _ = generate(n_gates=100, n_ff=10, n_pi=8, n_po=8, depth=6, seed=3, name="cm2")
    with pytest.raises(ValueError):
        Compiled(elaborate(src), Options(threads=2, bitplane=False, cost_fn=TableCost([1, 2, 3])))


def test_mtask_features_add_up():
    src, _ = generate(n_gates=300, n_ff=30, n_pi=8, n_po=8, depth=8, seed=5, name="cm3")
    comp = Compiled(elaborate(src), Options(threads=4, bitplane=False))
    fx = FeatureExtractor(comp)
    for k, mt in enumerate(comp.mtasks):
        f = fx.mtask(k)
        assert f[0] == sum(fx.vertex(v)[0] for v in mt.vids)   # instruction counts add
        assert f[4] == sum(fx.vertex(v)[4] for v in mt.vids)   # branches add
        assert f[1] <= sum(fx.vertex(v)[1] for v in mt.vids)   # fan-in shrinks on merging
    x = encode([fx.mtask(k) for k in range(len(comp.mtasks))])
    assert x.shape == (len(comp.mtasks), 6) and np.isfinite(x).all()


def test_committed_partition_choice_is_the_predicted_minimum():
    """The learned choice of partition granularity: the committed choice per
    target is the candidate with the least predicted eval time, a choice the
    instruction-count heuristic cannot make (every candidate partition of a
    design has the same total instruction count)."""
    import json
    from rtlflow_amd.designs.registry import COSTTABLES
    choice = json.loads((COSTTABLES / "partition_choice.json").read_text())
    for name, c in choice.items():
        pred = {int(t): v for t, v in c["predicted_us"].items()}
        assert c["threads"] == min(pred, key=pred.get)
    src, _ = generate(n_gates=800, n_ff=80, n_pi=16, n_po=16, depth=12, seed=4, name="eq")
    totals = set()
    for t in (2, 4, 8):
        comp = Compiled(elaborate(src), Options(threads=t, bitplane=False))
        fx = FeatureExtractor(comp)
        totals.add(sum(fx.mtask(k)[0] for k in range(len(comp.mtasks))))
    assert len(totals) == 1
