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

from .compiler import (_eval_loop_body_split, _fused_locals, clk0_carried, clk0_cone, _hash_code, _initial_body, _lds_copy_lines,
                       _lds_decl_lines, _prologue, _stim_code, _stim_prefetch, atomic, kernel_header, lds_memories,
                       place_lds_state, plan_packing)
from .emit_hip import Gen, is_wide
from .ir import CLASS_BYTES
from .layout import class_of


def image_load_lines(comp, ind: str, lds=None) -> list:
    """Copy the stimulus's memory images (rf_set_mem_images) into the lane.
    a.images is a table of {image base pointer, (nstim << 32) | nelem} per
    memory of the design's memory table. ``lds``: memories the kernel holds
    in LDS (var -> array name), loaded there instead."""
    lds = lds or {}
    d = comp.design
    lay = comp.layout
    mems = [v for v in d.vars if v.is_mem and lay.has(v)]
    if not mems:
        return []
    cty_ = {"C": "uint8_t", "S": "uint16_t", "I": "uint32_t", "Q": "uint64_t"}
    arr = {"C": "a.c", "S": "a.s", "I": "a.i", "Q": "a.q"}
    out = [f"{ind}if (a.images) {{",
           f"{ind}    const uint64_t* im = (const uint64_t*)a.images;"]
    for j, m in enumerate(mems):
        c = class_of(m)
        T = cty_[c]
        n = m.words
        out.append(f"{ind}    {{ const {T}* src = (const {T}*)im[{2 * j}];")
        out.append(f"{ind}      const uint32_t ne = (uint32_t)im[{2 * j + 1}];")
        out.append(f"{ind}      if (src && sid < (uint32_t)(im[{2 * j + 1}] >> 32)) {{")
        out.append(f"{ind}        src += (size_t)sid * ne * {n}u;")
        if m in lds:
            out.append(f"{ind}        for (uint32_t k = 0; k < ne * {n}u; ++k) {lds[m]}[k][threadIdx.x] = src[k];")
        elif m in lay.lane_major:
            row = m.depth * (CLASS_BYTES[m.cls] if m.cls != "W" else 4 * m.words)
            out.append(f"{ind}        {T}* dst = ({T}*)(a.mem + (size_t){lay.lane_major[m]} * a.npad"
                       f" + (size_t)lane * {row}u);")
            out.append(f"{ind}        for (uint32_t k = 0; k < ne * {n}u; ++k) dst[k] = src[k];")
        else:
            out.append(f"{ind}        for (uint32_t k = 0; k < ne * {n}u; ++k) "
                       f"{arr[c]}[((size_t){lay.loc[m]} + k) * a.npad + lane] = src[k];")
        out.append(f"{ind}      }} }}")
    out.append(f"{ind}}}")
    return out


def emit_persistent(comp, io: bool = False) -> str:
    """The batched kernel; ``io``: the variant with the input-stream and
    output-capture code (rf_k_persistent_io), kept out of the plain one."""
    d = comp.design
    gen = Gen(comp)
    gen.lds = lds_memories(comp)
    gen.top_lines = []
    gen.local = _fused_locals(comp)
    cone = clk0_cone(comp)
    plan_packing(comp, gen, comp.stored | clk0_carried(comp, cone))
    lds_state = place_lds_state(comp, gen)
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
        elif v not in gen.packed:
            gen.emit(gen.set(v, "0"))
    for r in sorted(gen.pack_regs):
        gen.emit(f"{r} = 0u;")
    gen.avail = set(gen.local)
    for m in [v for v in d.vars if v.is_mem and lay.has(v)]:
        if m in gen.lds:
            gen.emit(f"for (uint32_t k = 0; k < {m.depth * m.words}u; ++k) {gen.lds[m]}[k][threadIdx.x] = 0;")
        elif m in lay.lane_major:
            row = m.depth * (CLASS_BYTES[m.cls] if m.cls != "W" else 4 * m.words)
            gen.emit(f"{{ uint64_t* p = (uint64_t*)mb{m.id}; for (uint32_t k = 0; k < {row // 8}u; ++k) p[k] = 0ull; }}")
        else:
            c = class_of(m)
            arr = {"C": "sC", "S": "sS", "I": "sI", "Q": "sQ"}[c]
            gen.emit(f"for (uint32_t k = 0; k < {m.slots}u; ++k) {arr}[((size_t){lay.loc[m]} + k) * NP] = 0;")
    for ln in image_load_lines(comp, "", gen.lds):
        gen.emit(ln)
    _initial_body(comp, gen)
    gen.emit("uint64_t h = RF_FNV_BASIS;")
    pf = io and _stim_prefetch(comp, gen, "(uint64_t)sid", "0ull", "budget != 0u", True)
    gen.emit("for (uint32_t cyc = 0u; cyc < budget; ++cyc) {")
    gen.ind += 1
    _stim_code(comp, gen, "(uint64_t)sid", "(uint64_t)cyc", pf, io)
    if pf:
        _stim_prefetch(comp, gen, "(uint64_t)sid", "(uint64_t)cyc + 1u", "cyc + 1u < budget", False)
    gen.emit("#pragma unroll 1")
    gen.emit("for (uint32_t ph = 0u; ph < 2u; ++ph) {")
    gen.ind += 1
    for clk in comp.in_clocks:
        gen.emit(gen.set(clk, "ph"))
        gen.dirty.add(clk)
    _eval_loop_body_split(comp, gen, "cyc == 0u", cone)
    gen.ind -= 1
    gen.emit("}")
    if io:
        _hash_code(comp, gen, "h", "(uint64_t)sid", "(uint64_t)cyc")
    else:
        _hash_code(comp, gen, "h")
    gen.ind -= 1
    gen.emit("}")
    gen.emit("a.out_hash[sid] = h;")
    gen.emit(f"q = {atomic(comp, 'add', '&a.flags[RF_FLAG_QUEUE]', '1u')};")
    gen.ind -= 1
    gen.emit("}")
    lines = [kernel_header(comp, "rf_k_persistent_io" if io else "rf_k_persistent")]
    lines += _lds_decl_lines(comp, gen.lds) + lds_state
    _prologue(comp, gen, lines, check_done=False)
    lines += gen.top_lines + gen.lines
    # the last stimulus's memories back to their storage (lanes that ran one)
    if gen.lds:
        lines.append("    if (lane < a.nstim) {")
        lines += ["    " + ln for ln in _lds_copy_lines(comp, gen.lds, to_lds=False)]
        lines.append("    }")
    lines.append("}")
    return "\n".join(lines)
