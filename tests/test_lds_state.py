"""Placement of variables and small memories in LDS by the whole-cycle
kernels (fused, persistent): which ones move, within what budget, and that
the generated code addresses one cell per lane. The GPU parity tests of the
ALU, RV32I and systolic designs run these kernels."""
import re

from rtlflow_amd import compiler as C
from rtlflow_amd.compiler import Compiled, Options, emit_source
from rtlflow_amd.verilog import elaborate


def _fused(src):
    i = src.index("rf_k_fused(rf_args a)")
    return src[i:src.index("\n}\n", i)]


def test_register_file_lives_in_lds():
    from rtlflow_amd.designs.alu import verilog
    comp = Compiled(elaborate(verilog()), Options(threads=2, block=128))
    lds = C.lds_memories(comp)
    assert [m.name for m in lds] == ["rf"]
    k = _fused(emit_source(comp))
    assert "__shared__ uint32_t lm" in k and "[32][128];" in k
    # every register-file access goes through this lane's column
    name = lds[next(iter(lds))]
    assert re.search(name + r"\[\(size_t\)\(.*\) \* 1 \+ 0\]\[threadIdx\.x\]", k)
    # copied in from and back to the lane-minor arrays
    assert k.count(f"for (uint32_t k = 0; k < 32u; ++k)") == 2


def test_large_state_moves_to_lds_within_budget():
    from rtlflow_amd.designs.systolic import verilog
    comp = Compiled(elaborate(verilog()), Options(threads=2, block=64))
    k = _fused(emit_source(comp))
    decl = re.findall(r"__shared__ (?:uint8_t|uint16_t|uint32_t|uint64_t) ls([CSIQW])\[(\d+)\]\[(\d+)\];", k)
    assert decl, "no LDS state"
    size = {"C": 1, "S": 2, "I": 4, "Q": 8, "W": 4}
    total = sum(size[c] * int(a) * int(b) for c, a, b in decl)
    assert total <= C.LDS_STATE_BUDGET
    # the wide row length is odd (distinct banks for one word of 64 lanes)
    w = [int(b) for c, a, b in decl if c == "W"]
    assert w and all(x % 2 == 1 for x in w)
    # at most REG_WORDS words of plain register locals remain
    regs = re.findall(r"\n    uint(32|64)_t v\d+ = 0;", k)
    assert sum(2 if r == "64" else 1 for r in regs) <= C.REG_WORDS


def test_no_lds_for_small_designs_host_build_or_calls_form():
    from rtlflow_amd.designs.alu import verilog
    comp = Compiled(elaborate(verilog()), Options(threads=2))
    assert "__shared__ uint32_t ls" not in _fused(emit_source(comp))   # 28 bytes of state
    comp.host = True
    assert C.lds_memories(comp) == {}
    comp.host = False
    comp2 = Compiled(elaborate(verilog()), Options(threads=2, inline_limit=1))
    assert comp2.calls and C.lds_memories(comp2) == {}
