"""Emitter parity on the CPU: the one-lane host build of the generated code
(identical kernel bodies) against the C oracle, for every schedule form."""
import pytest

import oracle
from rtlflow_amd.compiler import Compiled, Options
from rtlflow_amd.designs.netlist import generate
from rtlflow_amd.hostsim import HostDesign
from rtlflow_amd.verilog import elaborate


@pytest.fixture(scope="module")
def tiny():
    src, table = generate(n_gates=300, n_ff=30, n_pi=16, n_po=16, depth=8, seed=7)
    return src, table.to_bytes()


@pytest.mark.parametrize("threads", [1, 3, 8])
def test_netlist_host_build_all_forms(tiny, threads):
    src, tb = tiny
    seed = 0x52544C4600000002
    ref = oracle.netlist_run(tb, seed, nstim=8, cycles=25)
    hd = HostDesign(Compiled(elaborate(src), Options(threads=threads)))
    assert (hd.protocol(8, seed, 25, mode=0) == ref).all()          # mtask table order
    assert (hd.protocol(8, seed, 25, mode=1) == ref).all()          # fused per-lane kernel
    for order_seed in (1, 2, 3):                                     # race test: any topo order
        assert (hd.protocol(8, seed, 25, mode=2, order_seed=order_seed) == ref).all()
    # batched work queue: uneven budgets, fewer lanes than stimuli
    budgets = [25, 3, 0, 11, 25, 7, 1, 25]
    refb = oracle.netlist_run(tb, seed, nstim=8, budgets=budgets)
    assert (hd.queue(3, seed, budgets) == refb).all()


def test_netlist_with_rams_and_wide_inputs_host_build():
    """Synchronous RAMs (registered read port, out-of-range addresses) and a
    70-bit input port drawn in two 64-bit chunks, in every schedule form."""
    src, table = generate(n_gates=400, n_ff=24, n_pi=70, n_po=20, depth=8, seed=31, n_rams=3,
                          ram_depth=12, ram_width=9)
    tb = table.to_bytes()
    seed = 0x52544C4600000005
    ref = oracle.netlist_run(tb, seed, nstim=6, cycles=60)
    hd = HostDesign(Compiled(elaborate(src), Options(threads=3)))
    for mode in (0, 1, 2):
        assert (hd.protocol(6, seed, 60, mode=mode, order_seed=2) == ref).all()
    budgets = [60, 0, 13, 60, 1, 44]
    assert (hd.queue(4, seed, budgets) == oracle.netlist_run(tb, seed, nstim=6, budgets=budgets)).all()


def test_calls_form_matches_inline_form(tiny):
    """Large designs call shared non-inlined mtask bodies; forcing that form
    on a small design must give the same hashes."""
    src, tb = tiny
    seed = 99
    ref = oracle.netlist_run(tb, seed, nstim=6, cycles=17)
    comp = Compiled(elaborate(src), Options(threads=4, inline_limit=1))
    assert comp.calls
    hd = HostDesign(comp)
    for mode in (0, 1, 2):
        assert (hd.protocol(6, seed, 17, mode=mode, order_seed=4) == ref).all()
    budgets = [17, 0, 5, 17, 2, 9]
    assert (hd.queue(2, seed, budgets) ==
            oracle.netlist_run(tb, seed, nstim=6, budgets=budgets)).all()


def test_two_design_libraries_in_one_process():
    """Every design library uses the same kernel names; they must stay
    library-local (a second library must not bind to the first one's kernels)."""
    seed = 5
    libs = []
    for gseed in (11, 12):
        src, table = generate(n_gates=150, n_ff=10, n_pi=8, n_po=8, depth=5, seed=gseed,
                              name="same_name")
        libs.append((HostDesign(Compiled(elaborate(src), Options(threads=2))), table.to_bytes()))
    for hd, tb in libs:
        assert (hd.protocol(4, seed, 9) == oracle.netlist_run(tb, seed, nstim=4, cycles=9)).all()


@pytest.mark.parametrize("threads", [1, 3])
def test_alu_host_build(threads):
    from rtlflow_amd.designs.alu import verilog
    seed = 0x52544C4600000001
    ref = oracle.alu_run(seed, 6, 150)
    hd = HostDesign(Compiled(elaborate(verilog()), Options(threads=threads)))
    for mode in (0, 1, 2):
        assert (hd.protocol(6, seed, 150, mode=mode, order_seed=7) == ref).all()
    budgets = [150, 0, 3, 77, 150, 1]
    want = [int(oracle.alu_run(seed, 1, b, stim0=i)[0]) for i, b in enumerate(budgets)]
    assert list(hd.queue(4, seed, budgets)) == want


def test_systolic_host_build():
    from rtlflow_amd.designs.systolic import verilog
    seed = 0x52544C4600000004
    ref = oracle.systolic_run(seed, 3, 530)
    hd = HostDesign(Compiled(elaborate(verilog()), Options(threads=2)))
    for mode in (0, 1):
        assert (hd.protocol(3, seed, 530, mode=mode) == ref).all()


def test_rv32i_host_build_and_iss():
    import ctypes as C
    import numpy as np
    from rtlflow_amd.designs import rv32i
    seed = 0x52544C4600000003
    n = 4
    progs = rv32i.programs(seed, np.arange(n))
    ref = oracle.rv32i_run(progs, seed, [2500] * n)
    for threads in (1, 2):
        hd = HostDesign(Compiled(elaborate(rv32i.verilog()), Options(threads=threads)))
        for mode in (0, 1, 2):
            assert (hd.protocol(n, seed, 2500, mode=mode, order_seed=9, images={"imem": progs}) == ref).all()
        b = [2500, 7, 0, 1200]
        assert (hd.queue(2, seed, b, images={"imem": progs}) == oracle.rv32i_run(progs, seed, b)).all()


def _circ_closed_form(seed, nstim, cycles, stim0=0):
    """circ: after each clock, o = a & 0x0f and ox = a (the settled x)."""
    out = []
    for s in range(stim0, stim0 + nstim):
        h = 0xcbf29ce484222325
        for c in range(cycles):
            a = oracle.stim(seed, s, c, 0) & 0xff
            h = oracle.fnv(bytes([a & 0x0f, a]), h)
        out.append(h)
    return out


def test_settle_loop_host_build():
    """A combinational cycle through two always blocks (UNOPTFLAT) settles
    by the change-detection loop in every schedule form and the batched
    queue; the closed form pins the result. A cycle that never settles
    reports non-convergence."""
    import ir_interp
    from rtlflow_amd.designs import settle
    seed = 0x5E77
    comp = Compiled(elaborate(settle.CIRC), Options(threads=2))
    assert [v.name for v in comp.sch.change_vars] == ["y"]
    want = _circ_closed_form(seed, 6, 20)
    assert ir_interp.protocol(elaborate(settle.CIRC), seed, 6, 20) == want
    hd = HostDesign(comp)
    for mode in (0, 1, 2):
        assert list(hd.protocol(6, seed, 20, mode=mode, order_seed=3)) == want
    budgets = [20, 0, 7, 20, 1, 13]
    got = list(hd.queue(3, seed, budgets))
    assert got == [_circ_closed_form(seed, 1, b, stim0=i)[0] for i, b in enumerate(budgets)]
    with pytest.raises(RuntimeError):
        HostDesign(Compiled(elaborate(settle.OSC), Options(threads=2))).protocol(2, seed, 3, mode=1)


def _smear_want(seed, nstim, cycles, budgets=None):
    from rtlflow_amd.designs import settle
    out = []
    for s in range(nstim):
        h = 0xcbf29ce484222325
        cyc = cycles if budgets is None else int(budgets[s])
        for b in settle.smear_closed_form(oracle.stim, seed, s, cyc):
            h = oracle.fnv(b, h)
        out.append(h)
    return out


def test_deep_settle_host_build():
    """smear needs up to 17 settle passes per eval, more than the graph
    form's two captured passes: the settle tail finishes them. Closed form,
    interpreter, the host loop, the fused kernel, the graph-form schedule
    (two passes + tail) and the batched queue agree."""
    import ir_interp
    from rtlflow_amd.designs import settle
    seed, n, cycles = 0x5EA1, 6, 40
    passes = [settle.smear_passes(oracle.stim, seed, s, cycles) for s in range(n)]
    assert max(max(p) for p in passes) >= 8 and sum(p > 4 for q in passes for p in q) >= 10
    want = _smear_want(seed, n, cycles)
    assert ir_interp.protocol(elaborate(settle.SMEAR), seed, n, cycles) == want
    comp = Compiled(elaborate(settle.SMEAR), Options(threads=2))
    assert comp.sch.change_vars
    hd = HostDesign(comp)
    for mode in (0, 1, 2, 3):
        assert list(hd.protocol(n, seed, cycles, mode=mode, order_seed=4)) == want, mode
    budgets = [40, 0, 9, 33, 1, 17]
    assert list(hd.queue(3, seed, budgets)) == _smear_want(seed, len(budgets), 0, budgets)


def test_graph_schedule_tail_reports_unsettled():
    """The graph-form schedule's settle tail reports a loop that never
    settles, and leaves designs without combinational cycles unchanged."""
    from rtlflow_amd.designs import settle
    seed = 0x5E77
    with pytest.raises(RuntimeError):
        HostDesign(Compiled(elaborate(settle.OSC), Options(threads=2))).protocol(2, seed, 3, mode=3)
    hd = HostDesign(Compiled(elaborate(settle.CIRC), Options(threads=2)))
    assert list(hd.protocol(6, seed, 20, mode=3)) == _circ_closed_form(seed, 6, 20)


def test_gated_clock_host_build():
    """A derived (gated) clock, clk & en: its logic runs before the block it
    clocks, whose edge test sees the new value; closed form, interpreter and
    every schedule form agree. A clock gated by a register is refused."""
    import ir_interp
    from rtlflow_amd.designs import settle
    from rtlflow_amd.order import CompileError
    seed, n, cycles = 0x6A7E, 5, 30
    want = []
    for s in range(n):
        h = 0xcbf29ce484222325
        for b in settle.gated_closed_form(oracle.stim, seed, s, cycles):
            h = oracle.fnv(b, h)
        want.append(h)
    assert ir_interp.protocol(elaborate(settle.GATED), seed, n, cycles) == want
    hd = HostDesign(Compiled(elaborate(settle.GATED), Options(threads=2)))
    for mode in (0, 1, 2):
        assert list(hd.protocol(n, seed, cycles, mode=mode, order_seed=5)) == want
    with pytest.raises(CompileError, match="clocked state"):
        Compiled(elaborate("""module g2(input clk, input en, output [3:0] o);
            reg e; reg [3:0] q; wire gc = clk & e;
            always @(posedge clk) e <= en;
            always @(posedge gc) q <= q + 4'd1;
            assign o = q; endmodule"""), Options(threads=1))
