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


@pytest.mark.parametrize("name,threads,lds", [("alu", 2, True), ("systolic", 4, True), ("net_small", 4, True),
                                              ("net2", 4, False)])
def test_slots_in_lds_when_they_fit(name, threads, lds):
    """The kernel keeps the workgroup's 64 lanes of every slot in LDS when
    they fit (MW_LDS_MAX), with copies in and out of the lane-minor arrays
    covering every slot of every class; otherwise it works on the arrays."""
    from rtlflow_amd.compiler import CLASS_BYTES, MW_LDS_MAX, emit_fused_mw, mw_lds_slots
    e = registry.ENTRIES[name]
    spec = e.make()
    comp = Compiled(elaborate(spec["verilog"], spec["top"]), Options(threads=threads, bitplane=False))
    sl = mw_lds_slots(comp)
    total = sum(comp.layout.slots[c] * CLASS_BYTES[c] for c in "CSIQ") * 64
    assert (sl is not None) == (total <= MW_LDS_MAX) == lds
    if not mw_waves(comp):
        return
    src, W = emit_fused_mw(comp)
    for c in "CSIQ":
        n = comp.layout.slots[c]
        if lds and n:
            assert f"__shared__ {['uint8_t', 'uint16_t', 'uint32_t', 'uint64_t']['CSIQ'.index(c)]} lsoa{c}[{n * 64}];" in src
            # loaded before the cycle loop and written back after it
            assert src.count(f"k < {n * 64}u; k += {64 * W}u") == 2
    assert ("constexpr size_t NP = 64;" in src) == lds


@pytest.mark.parametrize("name,threads", [("alu", 2), ("rv32i", 2), ("net_small", 4)])
def test_falling_edge_eval_runs_the_cone_mtasks_only(name, threads):
    """After a launch's first cycle the clk = 0 eval runs only the mtasks
    holding falling-edge cone logic: every mtask appears once in the full
    eval and once more per cone mtask."""
    from rtlflow_amd.compiler import clk0_cone, emit_fused_mw
    e = registry.ENTRIES[name]
    spec = e.make()
    comp = Compiled(elaborate(spec["verilog"], spec["top"]), Options(threads=threads, bitplane=False))
    src, W = emit_fused_mw(comp)
    assert src and W >= 2
    cone = clk0_cone(comp)
    assert cone is not None
    inside = set(cone)
    n_cone = sum(any(v in inside for v in mt.vids) for mt in comp.mtasks)
    assert "if (ph == 0u && n != 0u) {" in src
    # one `if (live)` block per level that has work, in each of the two
    # evals (plus the one that loads the lane's hash, cycle and stimulus)
    _, levels = mw_plan(comp)
    full = sum(1 for lv in levels if any(lv))
    cone_lv = sum(1 for lv in levels if any(any(v in inside for v in comp.mtasks[k].vids)
                                            for ks in lv for k in ks))
    assert src.count("if (live) {") == 1 + full + cone_lv
    if name in ("alu", "rv32i"):
        assert n_cone == 0 and cone_lv == 0     # inputs are read by clocked logic only


@pytest.mark.parametrize("cap", [2, 5, 8])
def test_wavefront_cap(cap):
    """Options.mw_max_waves bounds the plan's wavefronts; the plan still
    covers every mtask once, after its predecessors."""
    e = registry.ENTRIES["systolic"]
    spec = e.make()
    comp = Compiled(elaborate(spec["verilog"], spec["top"]),
                    Options(threads=16, bitplane=False, split_reductions=32, mw_max_waves=cap))
    W, levels = mw_plan(comp)
    assert W == cap
    seen = sorted(k for lv in levels for ks in lv for k in ks)
    assert seen == list(range(len(comp.mtasks)))
