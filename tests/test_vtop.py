"""The generated C++ model class and testbench compile against the C ABI
headers (host compiler only; no GPU)."""
import subprocess

from rtlflow_amd.compiler import Compiled, Options
from rtlflow_amd.emit_vtop import emit_testbench, emit_vtop_header
from rtlflow_amd.build import CSRC, INCLUDE
from rtlflow_amd.verilog import elaborate

SRC = """
module mix(input clk, input [6:0] a, input [99:0] w, input [40:0] q,
           output reg [15:0] o, output [99:0] ow, output [40:0] oq);
    reg [99:0] rw;
    always @(posedge clk) begin o <= o + {9'd0, a}; rw <= rw ^ w; end
    assign ow = rw;
    assign oq = q + 41'd1;
endmodule
"""


def test_header_and_testbench_compile(tmp_path):
    comp = Compiled(elaborate(SRC), Options(threads=1))
    man = comp.manifest()
    hdr = emit_vtop_header(man)
    assert "static constexpr rf_port_desc clk{" in hdr
    assert "static constexpr rf_port_desc w{" in hdr and ", 100u, 4u, 16u, 0u}" in hdr
    (tmp_path / "Vmix.h").write_text(hdr)
    (tmp_path / "Vmix_tb.cpp").write_text(emit_testbench(man))
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror", "-I", str(INCLUDE),
                        "-I", str(CSRC), "-I", str(tmp_path), str(tmp_path / "Vmix_tb.cpp")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
