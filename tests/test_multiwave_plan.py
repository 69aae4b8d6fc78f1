"""The multi-wavefront plan (CPU): every mtask runs exactly once per eval,
after every predecessor (in an earlier level), and the wavefront count stays
within one workgroup."""
import pytest

from rtlflow_amd.compiler import Compiled, MW_MAX_WAVES, Options, mw_plan, mw_waves
from rtlflow_amd.designs import registry
from rtlflow_amd.verilog import elaborate


@pytest.mark.parametrize("name,threads", [("alu", 2), ("systolic", 2), ("systolic", 8), ("rv32i", 2),
                                          ("net_small", 4)])
def test_plan_covers_the_dag_in_order(name, threads):
    e = registry.ENTRIES[name]
    spec = e.make()
    comp = Compiled(elaborate(spec["verilog"], spec["top"]), Options(threads=threads, bitplane=False))
    W, levels = mw_plan(comp)
    assert 1 <= W <= MW_MAX_WAVES
    assert all(len(lv) == W for lv in levels)
    where = {}
    for li, lv in enumerate(levels):
        for ks in lv:
            for k in ks:
                assert k not in where
                where[k] = li
    assert sorted(where) == list(range(len(comp.mtasks)))
    idx = {mt.id: k for k, mt in enumerate(comp.mtasks)}
    for k, mt in enumerate(comp.mtasks):
        for p in mt.preds:
            assert where[idx[p]] < where[k]
    assert mw_waves(comp) in (0, W)
    if name == "systolic" and threads == 8:
        assert W >= 6      # 258 independent processing-element mtasks
