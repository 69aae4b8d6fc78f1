"""Batched persistent kernel: uneven-length stimuli from a device work queue.

The reference runs every stimulus in its own lane for the whole launch and can
only mask finished lanes with ``done[]`` (src/V3EmitC.cpp:1654, :5361-5364):
a finished lane keeps occupying its thread. Here each lane runs a stimulus to
its cycle budget, writes the stimulus's hash, and immediately refills itself
with the next stimulus id from a device queue (one atomic per refill), so a
wavefront stays full until the queue drains. A refill zeroes the lane's state,
copies the stimulus's memory images, runs the initial blocks and the settle
loop -- the same steps a fresh model takes -- so every stimulus's hash is
identical to running it alone in a fresh model.

Queue order is chosen by the host (longest budget first by default, the LPT
rule, which bounds the tail by the longest single stimulus).
"""
from __future__ import annotations

from .compiler import (_eval_loop_body, _fused_locals, _hash_code, _initial_body, _prologue,
                       _stim_code, atomic, kernel_header)
from .emit_hip import Gen, is_wide
from .ir import CLASS_BYTES
from .layout import class_of


def emit_persistent(comp) -> str:
    d = comp.design
    gen = Gen(comp)
    gen.top_lines = []
    gen.local = _fused_locals(comp)
    for v in sorted(gen.local, key=lambda x: x.id):
        gen.declare(v)
    lay = comp.layout
    gen.emit("uint32_t q = lane;")
    gen.emit("while (q < a.nstim) {")
    gen.ind += 1
    gen.emit("const uint32_t sid = a.order ? a.order[q] : q;")
    gen.emit("const uint32_t budget = a.budget[sid];")
    # reset: registers and this lane's memories
    for v in sorted(gen.local, key=lambda x: x.id):
        if is_wide(v.width):
            gen.emit(f"rfw_zero({gen.name(v)}, {v.width});")
        else:
            gen.emit(f"{gen.name(v)} = 0;")
    gen.avail = set(gen.local)
    for m in [v for v in d.vars if v.is_mem and lay.has(v)]:
        if m in lay.lane_major:
            row = m.depth * (CLASS_BYTES[m.cls] if m.cls != "W" else 4 * m.words)
            gen.emit(f"{{ uint64_t* p = (uint64_t*)mb{m.id}; for (uint32_t k = 0; k < {row // 8}u; ++k) p[k] = 0ull; }}")
        else:
            c = class_of(m)
            arr = {"C": "sC", "S": "sS", "I": "sI", "Q": "sQ"}[c]
            gen.emit(f"for (uint32_t k = 0; k < {m.slots}u; ++k) {arr}[((size_t){lay.loc[m]} + k) * NP] = 0;")
    _initial_body(comp, gen)
    gen.emit("uint64_t h = RF_FNV_BASIS;")
    gen.emit("for (uint32_t cyc = 0u; cyc < budget; ++cyc) {")
    gen.ind += 1
    _stim_code(comp, gen, "(uint64_t)sid", "(uint64_t)cyc")
    gen.emit("#pragma unroll 1")
    gen.emit("for (uint32_t ph = 0u; ph < 2u; ++ph) {")
    gen.ind += 1
    for clk in d.clocks:
        gen.emit(f"{gen.name(clk)} = ph;")
        gen.dirty.add(clk)
    _eval_loop_body(comp, gen)
    gen.ind -= 1
    gen.emit("}")
    _hash_code(comp, gen, "h")
    gen.ind -= 1
    gen.emit("}")
    gen.emit("a.out_hash[sid] = h;")
    gen.emit(f"q = {atomic(comp, 'add', '&a.flags[RF_FLAG_QUEUE]', '1u')};")
    gen.ind -= 1
    gen.emit("}")
    lines = [kernel_header(comp, "rf_k_persistent")]
    _prologue(comp, gen, lines, check_done=False)
    lines += gen.top_lines + gen.lines + ["}"]
    return "\n".join(lines)
