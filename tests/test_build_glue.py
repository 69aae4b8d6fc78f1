"""Build glue: the emitted V<top>.mk + verilated.mk (the reference's
V3EmitMk output and include/verilated.mk.in rules, src/V3EmitMk.cpp:220-245,
include/verilated.mk.in:14-20) build, with make and hipcc for gfx950, the
same design library build.py does."""
import ctypes as C
import shutil
import subprocess

from rtlflow_amd import compile_design
from rtlflow_amd.build import PKG


class _DesignHead(C.Structure):
    _fields_ = [("abi", C.c_uint32), ("block", C.c_uint32), ("name", C.c_char_p),
                ("hash", C.c_uint64), ("slots", C.c_uint64 * 4), ("mem_row_bytes", C.c_uint64),
                ("nports", C.c_uint32), ("ports", C.c_void_p), ("nmems", C.c_uint32),
                ("mems", C.c_void_p), ("nmtasks", C.c_uint32)]


def _head(lib_path):
    lib = C.CDLL(str(lib_path), mode=C.RTLD_LOCAL)
    lib.rf_design_get.restype = C.POINTER(_DesignHead)
    h = lib.rf_design_get().contents
    return (h.abi, h.block, h.name, h.hash, tuple(h.slots), h.mem_row_bytes, h.nports, h.nmems,
            h.nmtasks)


def test_emitted_makefiles_build_the_same_library(tmp_path):
    src = """module glue(input clk, input [7:0] a, input [7:0] b, output [7:0] o, output [15:0] w);
        reg [7:0] r; reg [15:0] acc;
        always @(posedge clk) begin r <= a ^ b; acc <= acc + {8'd0, r}; end
        assign o = r; assign w = acc; endmodule"""
    cd = compile_design(src, name="glue", threads=2)
    d = cd.lib_path.parent
    for f in ("Vglue.hip", "Vglue.mk", "verilated.mk"):
        shutil.copy(d / f, tmp_path / f)
    r = subprocess.run(["make", "-f", "Vglue.mk", f"RF_ROOT={PKG}"], cwd=tmp_path,
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout[-3000:]
    assert "--offload-arch=gfx950" in r.stdout
    made = tmp_path / "libVglue.so"
    assert made.exists()
    assert _head(made) == _head(cd.lib_path)
