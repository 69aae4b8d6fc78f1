"""Known answers of the Verilog expression rules (IEEE 1364-2005 5.4/5.5 and
the two-state conventions of the reference, include/rf_verilated.h), computed
by hand. Each expression drives its own output; the IR interpreter (the
oracle) and the one-lane host build of the emitted code must both produce the
listed values. A wrong width rule, sign extension, or shift/division corner
case fails here."""
import pytest

import ir_interp
from rtlflow_amd.compiler import Compiled, Options
from rtlflow_amd.hostsim import HostDesign
from rtlflow_amd.verilog import elaborate

CASES = [
    # (output width, expression, expected value)
    (16, "8'hFF + 8'h01", 0x0100),                  # context widens the add
    (8, "8'hFF + 8'h01", 0x00),
    (16, "(8'hFF + 8'h01) >> 1", 0x0080),            # carry survives the shift
    (4, "$signed(4'b1000) >>> 2", 0b1110),           # arithmetic shift
    (4, "4'b1000 >>> 2", 0b0010),                    # unsigned operand: logical
    (1, "4'sd15 < 4'd2", 0),                          # mixed signedness -> unsigned compare
    (1, "$signed(4'hF) < $signed(4'd2)", 1),          # -1 < 2
    (8, "-4'd1", 0xFF),                               # negation in 8-bit context
    (6, "{2{3'b101}}", 0b101101),
    (8, "8'hFF << 9", 0x00),                          # shift past the width
    (8, "8'h80 >> 8", 0x00),
    (8, "$signed(8'h80) >>> 12", 0xFF),               # all sign bits
    (8, "-8'sd7 / 8'sd2", 0xFD),                      # -3, truncating
    (8, "-8'sd7 % 8'sd2", 0xFF),                      # -1, sign of the dividend
    (8, "8'd7 / 8'd0", 0x00),                         # two-state: x / 0 = 0
    (8, "8'd7 % 8'd0", 0x00),
    (1, "^8'b1011_0000", 1),
    (1, "~|4'b0000", 1),
    (1, "&4'b1111", 1),
    (4, "4'b1010 ~^ 4'b0110", 0b0011),
    (8, "(4'd3 > 4'd2) + 8'd1", 2),                   # 1-bit compare zero-extends
    (12, "{4'hA, 8'h5C}", 0xA5C),
    (8, "16'hABCD >> 4", 0xBC),                       # truncation after shift
    (16, "$signed(8'hF0)", 0xFFF0),                   # sign extension to context
    (16, "8'hF0", 0x00F0),
    (32, "32'hFFFF_FFFF * 32'hFFFF_FFFF", 0x00000001),
    (64, "32'hFFFF_FFFF * 32'hFFFF_FFFF", 0xFFFFFFFE00000001),  # context widens operands
    (64, "64'hFFFF_FFFF * 64'hFFFF_FFFF", 0xFFFFFFFE00000001),
    (72, "72'h1_0000_0000_0000_0000 - 1", 0xFFFFFFFFFFFFFFFF),
    (100, "100'h1 << 99", 1 << 99),
    (100, "$signed(100'h8_0000_0000_0000_0000_0000_0000) >>> 98", (1 << 100) - 2),
    (1, "100'h1 < 100'h2_0000_0000_0000_0000", 1),
    (70, "70'h3F_FFFF_FFFF_FFFF_FFFF / 70'd3", 0x3FFFFFFFFFFFFFFFFF // 3),
]


def _design():
    lines = ["module known(input clk, input [7:0] dummy,"]
    lines.append(",\n".join(f"    output [{w - 1}:0] o{i}" for i, (w, _, _) in enumerate(CASES)) + ");")
    for i, (w, e, _) in enumerate(CASES):
        lines.append(f"    assign o{i} = {e};")
    lines.append("endmodule")
    return "\n".join(lines)


def test_known_answers_interpreter():
    d = elaborate(_design())
    lane = ir_interp.Lane(d)
    lane.eval()
    outs = [v for v in d.vars if v.kind == "output"]
    for (w, e, want), v in zip(CASES, outs):
        assert lane.v[v] == want, f"{e}: got {lane.v[v]:#x}, want {want:#x}"


def test_known_answers_emitted_code():
    d = elaborate(_design())
    want = ir_interp.protocol(elaborate(_design()), 3, 2, 2)
    # the interpreter's protocol hash over the known constants must match
    hd = HostDesign(Compiled(d, Options(threads=2)))
    for mode in (0, 1):
        assert list(hd.protocol(2, 3, 2, mode=mode)) == want


@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_random_designs_host_build(seed):
    from rtlflow_amd.designs import fuzz
    src = fuzz.verilog(seed)
    ref = ir_interp.protocol(elaborate(src), 5, 3, 10)
    hd = HostDesign(Compiled(elaborate(src), Options(threads=3)))
    for mode in (0, 1, 2):
        assert list(hd.protocol(3, 5, 10, mode=mode, order_seed=seed)) == ref


def test_nonblocking_swap_group():
    """Two clocked blocks that read each other's targets form one group; both
    must see the values from before the edge (a swap)."""
    src = """
    module swap(input clk, input [7:0] x, output [15:0] o);
        reg [7:0] a, b;
        reg [7:0] mem [0:3];
        always @(posedge clk) begin a <= b + x; mem[x[1:0]] <= a; end
        always @(posedge clk) b <= a ^ mem[x[1:0]];
        assign o = {a, b};
    endmodule"""
    ref = ir_interp.protocol(elaborate(src), 9, 4, 20)
    for thr in (1, 2):
        hd = HostDesign(Compiled(elaborate(src), Options(threads=thr)))
        for mode in (0, 1, 2):
            assert list(hd.protocol(4, 9, 20, mode=mode, order_seed=2)) == ref


def test_wide_slice_views_and_shift_in_idiom():
    """Word-aligned slices of wide values read in place (compare, reduce,
    concatenate), and ``v <= {v[W-1-k:0], x}`` shifted in place, including a
    rotate whose shifted-in bits read v itself: host build of every form
    against the IR interpreter."""
    import ir_interp
    from rtlflow_amd.compiler import Compiled, Options
    from rtlflow_amd.hostsim import HostDesign
    from rtlflow_amd.verilog import elaborate
    src = """module wv(input clk, input [63:0] a, input [63:0] b, output [7:0] f, output [95:0] o);
        reg [255:0] c; reg [199:0] d; reg [127:0] r; reg [95:0] q;
        always @(posedge clk) begin
            c <= {c[247:0], a[7:0] ^ b[7:0]};
            d <= {d[159:0], a[39:0]};
            r <= {r[126:0], r[127]};
            q <= {c[159:128], c[63:0]} ^ {r[31:0], d[191:160], r[127:96]};
            if (r == 128'd0) r <= {a, b};
        end
        assign f = {4'd0, c[255:128] < c[127:0], ^c[255:64], c[127:96] == d[127:96], c[95:64] > c[63:32]};
        assign o = q;
    endmodule"""
    seed, n, cycles = 0x31D3, 4, 90
    want = ir_interp.protocol(elaborate(src), seed, n, cycles)
    hd = HostDesign(Compiled(elaborate(src), Options(threads=2)))
    for mode in (0, 1, 3):
        assert list(hd.protocol(n, seed, cycles, mode=mode)) == want, mode
