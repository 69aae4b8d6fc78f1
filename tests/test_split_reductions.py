"""Reduction splitting (CPU): a combinational assignment that is one
associative operator over many operands becomes partial terms of at most
``max_terms`` operands plus a final assignment over them; the outputs stay
equal to the unsplit design's and to the IR interpreter's."""
import pytest

import ir_interp
import oracle
from rtlflow_amd.compiler import Compiled, Options, mw_plan
from rtlflow_amd.designs import registry
from rtlflow_amd.hostsim import HostDesign
from rtlflow_amd.ir import Assign, Op, Ref
from rtlflow_amd.order import _assoc_leaves, split_reductions
from rtlflow_amd.verilog import elaborate


def _design(n=70):
    L = ["module red(input clk, input [31:0] din, input [4:0] sel,",
         "           output [31:0] x, output [31:0] s, output [31:0] o, output [15:0] a);"]
    for i in range(n):
        L.append(f"  reg [31:0] r{i};")
    L.append("  always @(posedge clk) begin")
    for i in range(n):
        src = "din" if i == 0 else f"r{i - 1}"
        L.append(f"    r{i} <= {src} + {i * 2654435761 % 997}*sel;")
    L.append("  end")
    L.append("  assign x = " + " ^ ".join(f"r{i}" for i in range(n)) + ";")
    L.append("  assign s = " + " + ".join(f"r{i}" for i in range(n)) + ";")
    L.append("  assign o = " + " | ".join(f"(r{i} & 32'h{(i * 40503) & 0xffffffff:x})" for i in range(n)) + ";")
    L.append("  assign a = " + " & ".join(f"r{i}[{i % 16} +: 16]" for i in range(n)) + ";")
    L.append("endmodule")
    return "\n".join(L) + "\n"


@pytest.mark.parametrize("max_terms", [2, 7, 32])
def test_split_structure(max_terms):
    d = elaborate(_design())
    before = len(d.processes)
    added = split_reductions(d, max_terms)
    assert added > 0 and len(d.processes) > before
    for p in d.processes:
        if p.kind != "comb" or len(p.stmts) != 1 or not isinstance(p.stmts[0], Assign):
            continue
        st = p.stmts[0]
        if isinstance(st.rhs, Op) and st.rhs.op in ("xor", "or", "and", "add"):
            leaves = []
            _assoc_leaves(st.rhs, st.rhs.op, st.rhs.width, leaves)
            assert len(leaves) <= max_terms, (p.name, len(leaves))
    # the final assignments read only the partials
    parts = {v for v in d.vars if v.name.startswith("__Vpart__")}
    assert parts
    for out in ("x", "s", "o", "a"):
        (p,) = [p for p in d.processes if p.kind == "comb" and p.stmts[0].lhs.var.name == out]
        leaves = []
        _assoc_leaves(p.stmts[0].rhs, p.stmts[0].rhs.op, p.stmts[0].rhs.width, leaves)
        assert all(isinstance(e, Ref) and e.var in parts for e in leaves)


@pytest.mark.parametrize("max_terms,threads", [(8, 4), (32, 16)])
def test_split_design_matches_unsplit_and_interpreter(max_terms, threads):
    src = _design()
    seed, n, cycles = 0xA11CE, 6, 40
    want = ir_interp.protocol(elaborate(src), seed, n, cycles)
    base = HostDesign(Compiled(elaborate(src), Options(threads=threads, bitplane=False)))
    split = HostDesign(Compiled(elaborate(src), Options(threads=threads, bitplane=False,
                                                        split_reductions=max_terms)))
    for mode in (0, 1, 2):
        assert [int(h) for h in base.protocol(n, seed, cycles, mode=mode)] == want
        assert [int(h) for h in split.protocol(n, seed, cycles, mode=mode)] == want


def test_systolic_output_reduction_split_shortens_the_critical_level():
    """The systolic array's 256-term accumulator XOR is the costliest vertex
    of its eval; split into partials, the multi-wavefront plan's costliest
    wavefront per level drops and the design stays equal to the oracle."""
    spec = registry.ENTRIES["systolic"].make()

    def plan(**kw):
        comp = Compiled(elaborate(spec["verilog"], spec["top"]), Options(threads=16, bitplane=False, **kw))
        W, lv = mw_plan(comp)
        crit = sum(max(sum(max(1, comp.mtasks[k].cost) for k in ks) for ks in level) for level in lv)
        return comp, crit

    _, c0 = plan()
    comp, c1 = plan(split_reductions=32)
    assert c1 < 0.6 * c0, (c0, c1)
    seed = 0x5157
    got = HostDesign(comp).protocol(4, seed, 24, mode=1)
    assert (got == oracle.systolic_run(seed, 4, 24)).all()
