"""Bit-plane lowering: the encoded program, run by a NumPy model of the
kernel, reproduces the oracle's per-stimulus hashes; unsupported designs are
refused."""
import numpy as np
import pytest

import oracle
from bitplane_sim import Program
from rtlflow_amd import bitplane as bp
from rtlflow_amd.compiler import Compiled, Options
from rtlflow_amd.designs.netlist import generate
from rtlflow_amd.verilog import elaborate


def _prog(src, threads=2):
    comp = Compiled(elaborate(src), Options(threads=threads))
    return comp, bp.encode(bp.compile_program(comp))


@pytest.mark.parametrize("byte_form", [True, False])
@pytest.mark.parametrize("seed", [7, 8])
def test_netlist_program_matches_oracle(seed, byte_form, monkeypatch):
    """Both gate record forms: LDS byte offsets (programs of at most 8,192
    slots) and slot indices (larger programs, forced here)."""
    if not byte_form:
        monkeypatch.setattr(bp, "BYTE_FORM_MAX_SLOTS", 0)
    src, table = generate(n_gates=400, n_ff=40, n_pi=16, n_po=24, depth=10, seed=seed, name="nl")
    _, blob = _prog(src)
    assert bool(blob[1] >> 16) == byte_form
    got = Program(blob).run(0x55AA, 70, 25)
    want = oracle.netlist_run(table.to_bytes(), 0x55AA, 70, 25)
    assert (got == want).all(), int((got != want).sum())


def test_gates_are_folded_and_shared():
    n = bp.Net()
    a = n.leaf("pi", 0)
    b = n.leaf("pi", 1)
    assert n.AND(a, b) == n.AND(b, a)
    assert n.XOR(a ^ 1, b) == n.XOR(a, b) ^ 1
    assert n.MUX(a, b, b) == b and n.MUX(a ^ 1, b, 0) == n.AND(a ^ 1, b)
    assert n.OR(a, a ^ 1) == bp.T and n.AND(a, a ^ 1) == bp.F


def test_refuses_unsupported_memories_and_division():
    # an asynchronous read
    with pytest.raises(bp.Unsupported, match="asynchronous"):
        comp = Compiled(elaborate("""module m(input clk, input [3:0] a, output [7:0] o);
            reg [7:0] mem [0:15]; always @(posedge clk) mem[a] <= {4'd0, a};
            assign o = mem[a]; endmodule"""), Options(threads=1))
        bp.compile_program(comp)
    # read data feeding logic before the edge
    with pytest.raises(bp.Unsupported, match="before the edge"):
        comp = Compiled(elaborate("""module m(input clk, input [3:0] a, output reg [7:0] o);
            reg [7:0] mem [0:15]; always @(posedge clk) begin mem[a] <= mem[a] + 8'd1; o <= mem[a]; end
            endmodule"""), Options(threads=1))
        bp.compile_program(comp)
    with pytest.raises(bp.Unsupported):
        comp = Compiled(elaborate("""module m(input clk, input [7:0] a, input [7:0] b, output reg [7:0] o);
            always @(posedge clk) o <= a / b; endmodule"""), Options(threads=1))
        bp.compile_program(comp)


ARITH = """
module ar(input clk, input [7:0] a, input [7:0] b, input [2:0] sel, input [69:0] w,
          output [15:0] o, output [7:0] q, output [69:0] ow, output f);
    reg [7:0] acc; reg [15:0] cnt; reg [69:0] rw; reg [3:0] st;
    wire signed [7:0] sa = a;
    wire signed [7:0] sb = b;
    wire [7:0] sh = a << sel;
    wire [7:0] sr = $signed(b) >>> sel;
    always @(posedge clk) begin
        case (sel)
            3'd0: acc <= acc + a;
            3'd1: acc <= acc - b;
            3'd2: acc <= (sa < sb) ? a : b;
            3'd3: acc <= {acc[3:0], acc[7:4]} ^ sh;
            3'd4: acc[3:0] <= a[7:4];
            default: acc <= sr | (a > b ? 8'h80 : 8'h01);
        endcase
        if (a[0]) cnt <= cnt + {8'd0, b}; else if (b[1]) cnt <= cnt - 16'd3;
        rw <= {rw[68:0], rw[69]} ^ w;
        st[sel[1:0]] <= ^a;
    end
    assign o = cnt ^ {acc, acc} ^ {8'd0, a * b[3:0]};
    assign q = acc + (a == b ? 8'd7 : 8'd0) + {4'd0, st};
    assign ow = rw + w;
    assign f = &a | ~|b;
endmodule
"""


def test_arithmetic_program_matches_interpreter():
    import ir_interp
    comp, blob = _prog(ARITH, threads=2)
    got = Program(blob).run(0x77, 20, 30)
    want = np.array(ir_interp.protocol(elaborate(ARITH), 0x77, 20, 30), dtype=np.uint64)
    assert (got == want).all(), int((got != want).sum())


def test_supported_random_designs_match_interpreter():
    import ir_interp
    from rtlflow_amd.designs.fuzz import verilog
    done = 0
    for seed in range(1, 13):
        src = verilog(seed, name=f"f{seed}", bitlevel=True)
        comp = Compiled(elaborate(src), Options(threads=2))
        try:
            blob = bp.encode(bp.compile_program(comp))
        except bp.Unsupported:
            continue
        got = Program(blob).run(0x99 + seed, 9, 10)
        want = np.array(ir_interp.protocol(elaborate(src), 0x99 + seed, 9, 10), dtype=np.uint64)
        assert (got == want).all(), (seed, int((got != want).sum()))
        done += 1
    assert done >= 10


def test_write_back_chunks_restore_every_stored_value():
    """Phase F runs in chunks sized to the LDS budget of phases A and B; the
    written-back values equal the interpreter's settled state."""
    from bitplane_sim import run_state
    from lanediff_util import interp_lane_state
    src, _ = generate(n_gates=1500, n_ff=150, n_pi=32, n_po=32, depth=16, seed=11, name="wbc")
    comp, blob = _prog(src, threads=4)
    prog = bp.compile_program(comp)
    assert len(prog.fchunks) > 1
    got = run_state(Program(blob), 0x5EED, 70, 13, 66)
    want = interp_lane_state(elaborate(src), 0x5EED, 66, 13)
    man = comp.manifest()
    bad = [v["name"] for v in man["vars"] if v["name"] in want and got.get((v["cls"], v["loc"])) != want[v["name"]]]
    assert not bad, bad[:5]


def test_netlist_with_rams_matches_oracle():
    """Synchronous RAMs (registered read port, write port, out-of-range
    addresses of a depth that is not a power of two) and a 70-bit input port
    in two draws."""
    src, table = generate(n_gates=400, n_ff=30, n_pi=70, n_po=20, depth=8, seed=31, n_rams=3,
                          ram_depth=12, ram_width=9, name="nl")
    comp, blob = _prog(src)
    prog = Program(blob)
    assert len(prog.mems) == 3 and all(len(p) == 6 for p in prog.ports)
    got = prog.run(0x55AA, 70, 40)
    want = oracle.netlist_run(table.to_bytes(), 0x55AA, 70, 40)
    assert (got == want).all(), int((got != want).sum())




@pytest.mark.parametrize("lane_major_min_depth", [64, 1 << 30])
def test_memory_ports_match_host_build(lane_major_min_depth):
    """Guarded writes under if / case, two write ports on one memory (the
    later wins), reads of the address written at the same edge (old data),
    8-, 6- and 64-bit memories, element-major and lane-major layouts, and
    out-of-range addresses: the program against the per-lane host build."""
    from rtlflow_amd.designs.rams import verilog as rams
    from rtlflow_amd.hostsim import HostDesign
    comp = Compiled(elaborate(rams()), Options(threads=2, lane_major_min_depth=lane_major_min_depth))
    prog = bp.compile_program(comp)
    kinds = [r[0] for r in prog.ports[1]]
    assert kinds.count(0) == 4 and kinds.count(1) == 6
    got = Program(bp.encode(prog)).run(0x1234, 70, 60)
    want = HostDesign(comp).protocol(70, 0x1234, 60, mode=1)
    assert (got == want).all(), int((got != want).sum())
