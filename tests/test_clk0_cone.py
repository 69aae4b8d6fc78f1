"""The falling-edge cone of the whole-cycle kernels (compiler.clk0_cone):
which combinational logic the clk = 0 eval of a cycle re-runs, checked
against a recomputation over the variables' reader/writer relation, and the
designs for which the whole eval must run instead. Parity of the kernels
that use it (fused, persistent) against the oracle is in test_hostsim.py
and the GPU tests; here a design with combinational logic both before and
after its clocked logic, and one with a negedge block, run through the host
build of those kernels against the IR interpreter."""
import numpy as np
import pytest

import ir_interp
from rtlflow_amd.compiler import Compiled, Options, clk0_carried, clk0_cone
from rtlflow_amd.designs import registry
from rtlflow_amd.designs.netlist import generate
from rtlflow_amd.hostsim import HostDesign
from rtlflow_amd.verilog import elaborate


def _comp(name, threads=2):
    spec = registry.ENTRIES[name].make()
    return Compiled(elaborate(spec["verilog"]), Options(threads=threads))


def _cone_by_fixpoint(comp):
    """Combinational vertices whose inputs can differ from the previous
    eval's: grown to a fixpoint over variables (no topological order), from
    the design's inputs and from every variable written after one of its
    readers in the eval order."""
    sch = comp.sch
    pos = {vid: k for k, vid in enumerate(sch.topo)}
    comb = [vx for vx in sch.vertices if vx.kind == "comb"]
    dirty = set(comp.design.inputs)
    for vx in comb:
        for v in vx.reads | vx.need_old:
            if any(pos[w.id] > pos[vx.id] for w in sch.vertices if v in w.writes):
                dirty.add(("stale", vx.id))
    picked = {vx.id for vx in comb if ("stale", vx.id) in dirty}
    while True:
        grew = False
        for vx in comb:
            if vx.id not in picked and (vx.reads | vx.need_old) & dirty:
                picked.add(vx.id)
                grew = True
            if vx.id in picked and not vx.writes <= dirty:
                dirty |= vx.writes
                grew = True
        if not grew:
            return picked


@pytest.mark.parametrize("name", ["net_tiny", "net_small", "alu", "rv32i", "systolic", "rams", "bfuzz1",
                                  "bfuzz2"])
def test_cone_matches_fixpoint(name):
    comp = _comp(name)
    cone = clk0_cone(comp)
    assert cone is not None
    assert set(cone) == _cone_by_fixpoint(comp)
    # in eval order
    pos = {vid: k for k, vid in enumerate(comp.sch.topo)}
    assert [pos[v] for v in cone] == sorted(pos[v] for v in cone)


def test_cone_of_designs_reading_inputs_only_in_clocked_logic_is_empty():
    # the clk = 0 eval reduces to the clock latch
    for name in ("alu", "rv32i", "systolic"):
        assert clk0_cone(_comp(name)) == [], name


def test_netlist_cone_and_carried_values():
    src, _ = generate(n_gates=300, n_ff=30, n_pi=16, n_po=16, depth=8, seed=7)
    comp = Compiled(elaborate(src), Options(threads=1))
    cone = clk0_cone(comp)
    comb = [vx.id for vx in comp.sch.vertices if vx.kind == "comb"]
    assert 0 < len(cone) < len(comb)
    carried = clk0_carried(comp, cone)
    inside = set(cone)
    # every carried value is read inside the cone and written outside it
    for v in carried:
        assert any(v in (comp.sch.vertices[c].reads | comp.sch.vertices[c].need_old) for c in cone)
        assert any(v in vx.writes for vx in comp.sch.vertices if vx.id not in inside)


@pytest.mark.parametrize("name", ["circ", "osc", "gated"])
def test_whole_eval_for_cycles_and_derived_clocks(name):
    assert clk0_cone(_comp(name)) is None


NEGEDGE = """
module neg(input clk, input [7:0] a, output [7:0] y);
  reg [7:0] p, n;
  always @(posedge clk) p <= p + a;
  always @(negedge clk) n <= n ^ p;
  assign y = p + n;
endmodule
"""

# combinational logic on both sides of the clocked logic: t reads the
# state and feeds it, u reads the state and an input
MIXED = """
module mix(input clk, input [7:0] a, input [7:0] b, output [7:0] y);
  reg [7:0] q, r;
  wire [7:0] t = q ^ r;
  wire [7:0] u = (q + a) & b;
  always @(posedge clk) begin
    q <= t + a;
    r <= u;
  end
  assign y = t | u;
endmodule
"""


@pytest.mark.parametrize("src,whole", [(NEGEDGE, True), (MIXED, False)])
def test_fused_and_batched_host_builds_match_interpreter(src, whole):
    d = elaborate(src)
    comp = Compiled(d, Options(threads=2))
    assert (clk0_cone(comp) is None) == whole
    seed, n, cycles = 0x51DE, 9, 23
    ref = np.array(ir_interp.protocol(elaborate(src), seed, n, cycles), dtype=np.uint64)
    hd = HostDesign(comp)
    assert (hd.protocol(n, seed, cycles, mode=1) == ref).all()     # fused
    assert (hd.protocol(n, seed, cycles, mode=0) == ref).all()     # mtask table order
    # persistent queue: uneven budgets on fewer lanes than stimuli; a
    # stimulus's first cycle runs the whole eval, later ones the cone
    budgets = [cycles, 1, 0, 7, cycles, 2, 11, cycles, 5]
    want = [ir_interp.protocol(elaborate(src), seed, 1, b, stim0=k)[0] for k, b in enumerate(budgets)]
    assert (hd.queue(4, seed, budgets) == np.array(want, dtype=np.uint64)).all()
