"""Compile a design IR into a gfx950 HIP design library source.

Pipeline (the reference's verilation stack, src/Verilator.cpp:116-537):
ordering (order.py: V3Order/V3Clock/V3Changed) -> cost (instrcount.py or the
learned model) -> mtask partition (partition.py: V3Partition) -> storage
assignment (which variables live in SoA memory) -> layout (layout.py) -> kernels
and descriptor (emit_hip.py + this module: V3EmitC).
"""
from __future__ import annotations

import hashlib
import json
from typing import Callable, Dict, List, Optional, Set, Tuple

from .emit_hip import ARR, CTY, Gen, cty, edge_cond, is_wide, lit
from .instrcount import process_cost
from .ir import CLASS_BYTES, Assign, Design, Ref, Var, words_of
from .layout import Layout, class_of
from .order import CompileError, Schedule, build_schedule, split_reductions, split_seq_processes
from .partition import MTask, partition, partition_stats

BLOCK = 256
CLS_ID = {"C": 0, "S": 1, "I": 2, "Q": 3}


class Options:
    def __init__(self, threads: int = 8, block: int = BLOCK, lane_major_min_depth: int = 64,
                 cost_fn: Optional[Callable] = None, merge_fn: Optional[Callable] = None,
                 settle_unroll: int = 2,
                 converge_limit: int = 100, state_in_regs: bool = True, keep=None,
                 mtask_stats: bool = False, inline_limit: int = 6000, split_always: bool = True,
                 mtask_waves: int = 4, bitplane: bool = True, split_reductions: int = 0,
                 mw_max_waves: int = 16):
        self.threads = threads            # --threads: partition granularity (cpLimit)
        self.block = block
        self.lane_major_min_depth = lane_major_min_depth
        self.cost_fn = cost_fn            # vertex -> cost; default V3InstrCount restatement
        self.merge_fn = merge_fn          # compiled design -> merge model of the contraction
        self.settle_unroll = settle_unroll
        self.converge_limit = converge_limit
        self.state_in_regs = state_in_regs
        self.keep = keep or set()
        # designs whose summed logic cost exceeds this use the "calls" form
        self.inline_limit = inline_limit
        # split clocked blocks into independent statement groups
        self.split_always = split_always
        # per-mtask kernels: minimum waves per SIMD (bounds their registers)
        self.mtask_waves = mtask_waves
        # also lower the design to the bit-plane form when it is in that subset
        self.bitplane = bitplane
        # split combinational reductions (one associative operator over more
        # operands than this) into partial terms; 0 = off
        self.split_reductions = split_reductions
        # most wavefronts per 64 stimuli of the multi-wavefront kernel (more
        # wavefronts leave each fewer registers)
        self.mw_max_waves = mw_max_waves

    def key(self) -> str:
        return f"t{self.threads}_b{self.block}_lm{self.lane_major_min_depth}"


class Compiled:
    host = False   # emit host functions (one-lane host build) instead of kernels

    def __init__(self, design: Design, opts: Optional[Options] = None):
        self.design = design
        self.opts = opts or Options()
        d = design
        if self.opts.split_always:
            split_seq_processes(d)
        if self.opts.split_reductions:
            split_reductions(d, self.opts.split_reductions)
        # synthetic clock-latch variables (__Vclklast__, src/V3Clock.cpp:64-92)
        self.clklast: Dict[Var, Var] = {}
        # the clocks the protocol drives (primary inputs); derived clocks are
        # combinational logic of them and get their own latch
        self.in_clocks: List[Var] = [c for c in d.clocks if c.kind == "input"]
        for clk in d.clocks:
            v = Var(f"__Vclklast__{clk.name}", 1, False, "reg")
            v.id = len(d.vars)
            d.vars.append(v)
            self.clklast[clk] = v
        self.sch: Schedule = build_schedule(d, self.opts.keep)
        sch = self.sch
        # change-detection shadows (src/V3Changed.cpp:55-70)
        self.chglast: Dict[Var, Var] = {}
        for v in sch.change_vars:
            s = Var(f"__Vchglast__{v.name}", v.width, False, "reg")
            s.id = len(d.vars)
            d.vars.append(s)
            self.chglast[v] = s
        # costs
        self.vcost: List[int] = []
        merged = self.opts.merge_fn(self) if self.opts.merge_fn is not None else None
        for vx in sch.vertices:
            if merged is not None:
                c = merged.vertex_cost(vx.id)
            elif self.opts.cost_fn is not None:
                c = int(self.opts.cost_fn(self, vx))
            else:
                c = sum(process_cost(p) for p in vx.procs)
            vx.cost = max(1, c)
            self.vcost.append(vx.cost)
        edges = [(vx.id, s) for vx in sch.vertices for s in vx.succs]
        self.mtasks: List[MTask] = partition(self.vcost, edges, self.opts.threads, order=sch.topo,
                                             merged=merged)
        self.vx2mt: Dict[int, int] = {}
        for k, mt in enumerate(self.mtasks):
            for vid in mt.vids:
                self.vx2mt[vid] = k
        self._storage()
        self._layout()
        # inline form (all logic in one function, state in registers) for
        # small designs; shared non-inlined mtask bodies for large ones (by
        # instruction count, whatever cost function the partitioner used)
        instr = sum(process_cost(p) for vx in sch.vertices for p in vx.procs)
        self.calls = instr > self.opts.inline_limit
        self._bitplane()

    def _bitplane(self):
        """Bit-plane program (rtlflow_amd.bitplane) or None with the reason."""
        self.bitplane_blob = None
        self.bitplane_info: dict = {"supported": False, "reason": "disabled"}
        if not self.opts.bitplane:
            return
        from . import bitplane as bp
        try:
            prog = bp.compile_program(self)
        except bp.Unsupported as e:
            self.bitplane_info = {"supported": False, "reason": str(e)}
            return
        self.bitplane_blob = bp.encode(prog)
        self.bitplane_info = {"supported": True, "slots": prog.nslots, "gates": prog.gates,
                              "levels": prog.nlevels, "words": len(self.bitplane_blob),
                              "state_bits": sum(x[3] for x in prog.state_load)}

    # ---------------------------------------------------------- storage
    def _storage(self):
        d = self.design
        sch = self.sch
        live_vars: Set[Var] = set()
        for vx in sch.vertices:
            live_vars |= vx.reads | vx.writes | vx.need_old
        P: Set[Var] = set()
        for v in d.vars:
            role = sch.roles.get(v)
            if v.kind in ("input", "output"):
                P.add(v)
            elif v.is_mem and v in live_vars:
                P.add(v)
            elif role == "state" and v in live_vars:
                P.add(v)
        P |= set(self.clklast.values())
        P |= set(sch.change_vars)
        P |= set(self.chglast.values())
        for vx in sch.vertices:
            if vx.kind == "seq":
                for v in vx.reads | vx.need_old:
                    if sch.roles.get(v) == "comb":
                        P.add(v)
            else:
                # latch-like comb vars (old value kept when not written)
                for v in vx.need_old & vx.writes:
                    P.add(v)
        # initial processes may write vars; keep those that are live
        for p in sch.initial:
            for v in sch.access[p.id].writes:
                if v in live_vars or v in P:
                    P.add(v)
        self.persistent = P
        T: Set[Var] = set()
        writer_mt: Dict[Var, Set[int]] = {}
        for vx in sch.vertices:
            for v in vx.writes:
                writer_mt.setdefault(v, set()).add(self.vx2mt[vx.id])
        for vx in sch.vertices:
            mt = self.vx2mt[vx.id]
            for v in vx.reads | vx.need_old:
                ws = writer_mt.get(v, set())
                if ws and (ws - {mt}):
                    T.add(v)
        for v, ws in writer_mt.items():
            if len(ws) > 1:
                T.add(v)
        self.transfer = T - P
        self.stored = P | T

    def _layout(self):
        d = self.design
        lay = Layout()
        order = []
        order += [v for v in d.vars if v.kind == "input"]
        order += [v for v in d.vars if v.kind == "output"]
        order += [v for v in d.vars if v in self.stored and self.sch.roles.get(v) == "state"]
        order += [v for v in d.vars if v in self.stored]
        for v in order:
            if v in self.stored or v.kind in ("input", "output"):
                lay.add(v, lane_major=(v.is_mem and v.depth >= self.opts.lane_major_min_depth))
        self.layout = lay

    # ------------------------------------------------------- manifest
    def manifest(self) -> dict:
        d = self.design
        lay = self.layout
        seq_bytes = 0
        for v in self.persistent:
            if v.is_mem or v.kind == "input":
                continue
            if self.sch.roles.get(v) == "state" or v in self.clklast.values():
                seq_bytes += v.slots * CLASS_BYTES[class_of(v)]
        in_bytes = sum(v.slots * CLASS_BYTES[class_of(v)] for v in d.inputs if v not in d.clocks)
        out_bytes = sum(v.slots * CLASS_BYTES[class_of(v)] for v in d.outputs)
        st = partition_stats(self.mtasks)
        return {
            "name": d.name,
            "slots": dict(lay.slots),
            "lane_major_bytes": lay.lane_major_bytes,
            "bytes_per_lane": lay.bytes_per_lane(),
            "bytes_seq": seq_bytes,
            "bytes_in": in_bytes,
            "bytes_out": out_bytes,
            "ops_per_cycle": 2 * sum(self.vcost),
            "vertices": len(self.sch.vertices),
            "mtasks": len(self.mtasks),
            "partition": st,
            "change_vars": [v.name for v in self.sch.change_vars],
            "ports": [{k: x for k, x in p.items() if k != "var"} for p in self.port_table()],
            "mems": self.mem_table(),
            "vars": self.var_table(),
            "threads": self.opts.threads,
            "block": self.opts.block,
            "form": "calls" if self.calls else "inline",
            "mw_waves": mw_waves(self),
            "bitplane": self.bitplane_info,
            "streams": self.stream_table(),
        }

    def stream_table(self) -> dict:
        """Word layout of the input stream and the output capture: one
        [port name, 64-bit chunk] pair per word, in word order."""
        ins, outs = [], []
        for p in self.port_table():
            k = (p["width"] + 63) // 64
            if p["dir"] == 0:
                ins += [[p["name"], j] for j in range(k)]
            elif p["hashed"]:
                outs += [[p["name"], j] for j in range(k)]
        return {"in_words": len(ins), "out_words": len(outs), "in": ins, "out": outs}

    def mem_table(self) -> List[dict]:
        out = []
        for m in self.design.vars:
            if not (m.is_mem and self.layout.has(m)):
                continue
            lm = m in self.layout.lane_major
            out.append({"name": m.name, "width": m.width, "cls": CLS_ID[class_of(m)],
                        "words": m.words, "depth": m.depth, "lane_major": lm,
                        "loc": self.layout.lane_major[m] if lm else self.layout.loc[m]})
        return out

    def var_table(self) -> List[dict]:
        """Stored scalar/wide variables and their slots (for lane state diffs)."""
        out = []
        for v in self.design.vars:
            if v.is_mem or v not in self.layout.loc:
                continue
            out.append({"name": v.name, "width": v.width, "cls": CLS_ID[class_of(v)],
                        "words": v.words, "loc": self.layout.loc[v],
                        "role": self.sch.roles.get(v, "internal")})
        return out

    def port_table(self) -> List[dict]:
        d = self.design
        out = []
        stim_idx = 0
        for v in d.vars:
            if v.kind not in ("input", "output"):
                continue
            if v.kind == "input":
                if v in d.clocks:
                    dirn = 2
                    si = 0xffffffff
                else:
                    dirn = 0
                    si = stim_idx
                    stim_idx += 1
            else:
                dirn = 1
                si = 0xffffffff
            out.append({"name": v.name, "dir": dirn, "width": v.width, "cls": CLS_ID[class_of(v)],
                        "words": v.slots, "loc": self.layout.loc[v], "stim_index": si,
                        "hashed": 1 if v.kind == "output" else 0, "var": v})
        return out


# =====================================================================
# Kernel emission
# =====================================================================

def kernel_header(comp: Compiled, name: str, occupancy: int = 0) -> str:
    """Device kernel, or for the host build a per-lane C function with the
    same body. ``occupancy``: minimum waves per SIMD the register allocation
    must leave room for (0 = compiler's choice)."""
    if comp.host:
        return f'extern "C" void {name}(rf_args a, uint32_t lane_in) {{'
    occ = f" __attribute__((amdgpu_waves_per_eu({occupancy})))" if occupancy else ""
    return (f'extern "C" __global__ void __launch_bounds__({comp.opts.block}){occ} '
            f'{name}(rf_args a) {{')


def lane_expr(comp: Compiled) -> str:
    return "lane_in" if comp.host else f"blockIdx.x * {comp.opts.block}u + threadIdx.x"


def atomic(comp: Compiled, op: str, ptr: str, val: str) -> str:
    if comp.host:
        return f"__atomic_fetch_{op}({ptr}, {val}, __ATOMIC_RELAXED)"
    return f"atomic{op.capitalize()}({ptr}, {val})"


def wave_or_flag(comp: Compiled, cond: str, ptr: str) -> str:
    """OR a per-lane condition into one device word: one atomic per wavefront."""
    if comp.host:
        return f"if ({cond}) __atomic_fetch_or({ptr}, 1u, __ATOMIC_RELAXED);"
    return f"if (__any({cond}) && (threadIdx.x & 63u) == 0u) atomicOr({ptr}, 1u);"


def _prologue(comp: Compiled, gen: Gen, lines: List[str], check_done=True, check_change=False):
    B = comp.opts.block
    lines.append(f"    const uint32_t lane = {lane_expr(comp)};")
    lines.append("    if (lane >= a.nlanes) return;")
    if check_done:
        lines.append("    if (a.done[lane]) return;")
    if check_change:
        lines.append("    if (!a.change[lane]) return;")
    lines.append("    size_t NP = a.npad;")
    lines.append("    uint8_t* __restrict__ sC = a.c + lane;")
    lines.append("    uint16_t* __restrict__ sS = a.s + lane;")
    lines.append("    uint32_t* __restrict__ sI = a.i + lane;")
    lines.append("    uint64_t* __restrict__ sQ = a.q + lane;")
    lay = comp.layout
    for m, off in lay.lane_major.items():
        c = class_of(m)
        row = m.depth * (CLASS_BYTES[m.cls] if m.cls != "W" else 4 * m.words)
        lines.append(f"    {CTY[c]}* __restrict__ mb{m.id} = ({CTY[c]}*)(a.mem + (size_t){off} * NP"
                     f" + (size_t)lane * {row}u);")
    lines.append("    (void)sC; (void)sS; (void)sI; (void)sQ;")


def _vertex_vars(comp: Compiled, vids: List[int]) -> Set[Var]:
    out: Set[Var] = set()
    for vid in vids:
        vx = comp.sch.vertices[vid]
        for p in vx.procs:
            acc = comp.sch.access[p.id]
            out |= acc.dep_reads | acc.need_old | acc.writes
    return {v for v in out if not v.is_mem}


def _nb_temps_needed(sch, procs) -> Set[Var]:
    """Non-blocking targets of a clocked group that need a temporary: those
    read (for their pre-edge value) after the group writes them. Processes
    run in list order; a process that is one assignment reads its right-hand
    side before it writes, any other process may read its own targets
    afterwards."""
    first: Dict[Var, int] = {}
    for k, p in enumerate(procs):
        for v in sch.access[p.id].nb_writes:
            first.setdefault(v, k)
    need: Set[Var] = set()
    for k, p in enumerate(procs):
        acc = sch.access[p.id]
        single = len(p.stmts) == 1 and isinstance(p.stmts[0], Assign)
        for v in acc.dep_reads | acc.need_old:
            if v in first and (first[v] < k or (first[v] == k and not single)):
                need.add(v)
    return need


def _emit_vertices(comp: Compiled, gen: Gen, vids: List[int], store_in_guard: bool):
    sch = comp.sch
    for vid in vids:
        vx = sch.vertices[vid]
        if vx.kind == "seq":
            edges = []
            for p in vx.procs:
                for e in p.edges:
                    if e not in edges:
                        edges.append(e)
            cond = edge_cond(gen, edges)
            gen.emit(f"if ({cond}) {{")
            gen.ind += 1
            saved = set(gen.avail)
            saved_dirty = set(gen.dirty)
            gen.nb_needed = _nb_temps_needed(sch, vx.procs)
            gen.seq_group(vx.procs, [sch.access[p.id] for p in vx.procs], f"{vx.procs[0].id}")
            gen.nb_needed = None
            if store_in_guard:
                written = set()
                for p in vx.procs:
                    written |= {v for v in sch.access[p.id].writes if not v.is_mem}
                for v in sorted(written, key=lambda x: x.id):
                    if v not in saved and v in comp.stored:
                        gen.store(v)
                        gen.dirty.discard(v)
            gen.ind -= 1
            gen.emit("}")
            gen.avail = saved
            if store_in_guard:
                gen.dirty = {v for v in gen.dirty if v in saved or v in saved_dirty}
        else:
            for p in vx.procs:
                gen.process(p, sch.access[p.id], f"{p.id}")


def _mtask_gen(comp: Compiled, k: int) -> Gen:
    mt = comp.mtasks[k]
    gen = Gen(comp)
    gen.top_lines = []
    gen.write_through = True
    gen.local = _vertex_vars(comp, mt.vids)
    # clocks and clock latches used by edge guards
    for vid in mt.vids:
        vx = comp.sch.vertices[vid]
        for p in vx.procs:
            for _, clk in p.edges:
                gen.local.add(clk)
                gen.local.add(comp.clklast[clk])
    _emit_vertices(comp, gen, mt.vids, store_in_guard=True)
    for v in sorted(gen.dirty, key=lambda x: x.id):
        if v in comp.stored:
            gen.store(v)
    return gen


def emit_mtask_kernel(comp: Compiled, k: int) -> str:
    # the graph-node kernel keeps its own inlined copy of the body (a call to
    # the shared non-inlined body would cost the call ABI's registers and
    # scratch in every launch)
    gen = _mtask_gen(comp, k)
    lines = [kernel_header(comp, f"rf_mtask_{k}", occupancy=comp.opts.mtask_waves)]
    _prologue(comp, gen, lines, check_done=True, check_change=bool(comp.sch.change_vars))
    lines += gen.top_lines
    lines += gen.lines
    lines.append("}")
    return "\n".join(lines)


# ---------------------------------------------------------------------------
# "calls" form for large designs: each mtask body is one non-inlined device
# function shared by its graph-node kernel and by the init / fused /
# persistent kernels, which call the bodies in topological order; values cross
# bodies through the SoA state (which the caches keep close).
# ---------------------------------------------------------------------------

def body_header(comp: Compiled, name: str, ret: str = "void", extra: str = "") -> str:
    q = "static" if comp.host else "__device__ __attribute__((noinline))"
    return f"{q} {ret} {name}(const rf_args& a, uint32_t lane{extra}) {{"


def _body_prologue(comp: Compiled, lines: List[str]):
    lines.append("    size_t NP = a.npad;")
    lines.append("    uint8_t* __restrict__ sC = a.c + lane;")
    lines.append("    uint16_t* __restrict__ sS = a.s + lane;")
    lines.append("    uint32_t* __restrict__ sI = a.i + lane;")
    lines.append("    uint64_t* __restrict__ sQ = a.q + lane;")
    for m, off in comp.layout.lane_major.items():
        c = class_of(m)
        row = m.depth * (CLASS_BYTES[m.cls] if m.cls != "W" else 4 * m.words)
        lines.append(f"    {CTY[c]}* __restrict__ mb{m.id} = ({CTY[c]}*)(a.mem + (size_t){off} * NP"
                     f" + (size_t)lane * {row}u);")
    lines.append("    (void)sC; (void)sS; (void)sI; (void)sQ; (void)NP;")


def _wrap_body(comp: Compiled, gen: Gen, name: str, ret: str = "void", extra: str = "",
               tail: Optional[List[str]] = None) -> str:
    lines = [body_header(comp, name, ret, extra)]
    _body_prologue(comp, lines)
    lines += gen.top_lines + gen.lines + (tail or []) + ["}"]
    return "\n".join(lines)


def emit_bodies(comp: Compiled) -> List[str]:
    d = comp.design
    out = []
    for k in range(len(comp.mtasks)):
        out.append(_wrap_body(comp, _mtask_gen(comp, k), f"rf_body_{k}"))
    # latch + change request -> returns the lane's change flag
    gen = Gen(comp)
    gen.top_lines = []
    gen.local = set(d.clocks) | set(comp.clklast.values()) | set(comp.sch.change_vars) | \
        set(comp.chglast.values())
    for clk, last in comp.clklast.items():
        gen.ensure(clk)
        gen.declare(last)
        gen.emit(gen.set(last, gen.get(clk)))
        gen.avail.add(last)
        gen.dirty.add(last)
    if comp.sch.change_vars:
        _change_code(comp, gen, "ch")
    else:
        gen.emit("const uint32_t ch = 0u;")
    for v in sorted(gen.dirty, key=lambda x: x.id):
        gen.store(v)
    out.append(_wrap_body(comp, gen, "rf_body_latch", "uint32_t", tail=["    return ch;"]))
    # stimulus for (sid, cyc) and clk = 0
    gen = Gen(comp)
    gen.top_lines = []
    gen.local = set(d.inputs)
    _stim_code(comp, gen, "sid", "cyc")
    for v in sorted(gen.dirty, key=lambda x: x.id):
        gen.store(v)
    out.append(_wrap_body(comp, gen, "rf_body_stim", extra=", uint64_t sid, uint64_t cyc"))
    # clock value
    gen = Gen(comp)
    gen.top_lines = []
    gen.local = set(comp.in_clocks)
    for clk in comp.in_clocks:
        gen.declare(clk)
        gen.emit(gen.set(clk, "v"))
        gen.store(clk)
    out.append(_wrap_body(comp, gen, "rf_body_setclk", extra=", uint32_t v"))
    # output hash
    gen = Gen(comp)
    gen.top_lines = []
    gen.local = set(d.outputs)
    _hash_code(comp, gen, "h", "sid", "cyc")
    out.append(_wrap_body(comp, gen, "rf_body_hash", "uint64_t", extra=", uint64_t h, uint64_t sid, uint64_t cyc",
                          tail=["    return h;"]))
    # initial blocks + clock latch
    gen = Gen(comp)
    gen.top_lines = []
    sch = comp.sch
    loc = set(d.clocks) | set(comp.clklast.values())
    for p in sch.initial:
        acc = sch.access[p.id]
        loc |= {v for v in acc.dep_reads | acc.need_old | acc.writes if not v.is_mem}
    gen.local = loc
    for p in sch.initial:
        gen.process(p, sch.access[p.id], f"i{p.id}")
    for clk, last in comp.clklast.items():
        gen.ensure(clk)
        gen.declare(last)
        gen.emit(gen.set(last, gen.get(clk)))
        gen.avail.add(last)
        gen.dirty.add(last)
    for v in sorted(gen.dirty, key=lambda x: x.id):
        if v in comp.stored:
            gen.store(v)
    out.append(_wrap_body(comp, gen, "rf_body_initial"))
    return out


def _calls_eval(comp: Compiled, ind: str) -> List[str]:
    """One eval through the bodies, with the settle loop."""
    calls = [f"{ind}    rf_body_{k}(a, lane);" for k in range(len(comp.mtasks))]
    lim = atomic(comp, "or", "&a.flags[RF_FLAG_UNCONVERGED]", "1u")
    if not comp.sch.change_vars:
        return calls + [f"{ind}    (void)rf_body_latch(a, lane);"]
    out = [f"{ind}for (uint32_t it = 0u;; ++it) {{"]
    out += [("    " + c) for c in calls]
    out.append(f"{ind}    if (!rf_body_latch(a, lane)) break;")
    out.append(f"{ind}    if (it + 1u >= a.converge_limit) {{ {lim}; break; }}")
    out.append(f"{ind}}}")
    return out


def emit_calls_kernels(comp: Compiled) -> List[str]:
    lay = comp.layout
    out = []
    B = comp.opts.block
    # init: initial blocks, latch, settle
    lines = [kernel_header(comp, "rf_k_init"),
             f"    const uint32_t lane = {lane_expr(comp)};",
             "    if (lane >= a.nlanes) return;",
             "    rf_body_initial(a, lane);"]
    lines += _calls_eval(comp, "    ")
    lines.append("}")
    out.append("\n".join(lines))
    # fused protocol kernel
    lines = [kernel_header(comp, "rf_k_fused"),
             f"    const uint32_t lane = {lane_expr(comp)};",
             "    if (lane >= a.nlanes) return;",
             "    if (a.done[lane]) return;",
             "    uint64_t h = a.hash[lane];",
             "    uint32_t cyc = a.cyc[lane];",
             "    const uint64_t sid = (uint64_t)a.stim[lane];",
             "    for (uint32_t n = 0u; n < a.ncycles; ++n, ++cyc) {",
             "        rf_body_stim(a, lane, sid, (uint64_t)cyc);",
             "        for (uint32_t ph = 0u; ph < 2u; ++ph) {",
             "            rf_body_setclk(a, lane, ph);"]
    lines += _calls_eval(comp, "            ")
    lines += ["        }",
              "        h = rf_body_hash(a, lane, h, sid, (uint64_t)cyc);",
              "    }",
              "    a.hash[lane] = h;",
              "    a.cyc[lane] = cyc;",
              "}"]
    out.append("\n".join(lines))
    # persistent batched kernel
    lines = [kernel_header(comp, "rf_k_persistent"),
             f"    const uint32_t lane = {lane_expr(comp)};",
             "    if (lane >= a.nlanes) return;",
             "    const size_t NP = a.npad;",
             "    uint32_t q = lane;",
             "    while (q < a.nstim) {",
             "        const uint32_t sid = a.order ? a.order[q] : q;",
             "        const uint32_t budget = a.budget[sid];"]
    for c, arr in (("C", "a.c"), ("S", "a.s"), ("I", "a.i"), ("Q", "a.q")):
        if lay.slots[c]:
            lines.append(f"        for (size_t k = 0; k < {lay.slots[c]}u; ++k) {arr}[k * NP + lane] = 0;")
    for m, off in lay.lane_major.items():
        row = m.depth * (CLASS_BYTES[m.cls] if m.cls != "W" else 4 * m.words)
        lines.append(f"        {{ uint64_t* p = (uint64_t*)(a.mem + (size_t){off} * NP + (size_t)lane * {row}u);")
        lines.append(f"          for (uint32_t k = 0; k < {row // 8}u; ++k) p[k] = 0ull; }}")
    from .persistent import image_load_lines
    lines += image_load_lines(comp, "        ")
    lines.append("        rf_body_initial(a, lane);")
    lines += _calls_eval(comp, "        ")
    lines += ["        uint64_t h = RF_FNV_BASIS;",
              "        for (uint32_t cyc = 0u; cyc < budget; ++cyc) {",
              "            rf_body_stim(a, lane, (uint64_t)sid, (uint64_t)cyc);",
              "            for (uint32_t ph = 0u; ph < 2u; ++ph) {",
              "                rf_body_setclk(a, lane, ph);"]
    lines += _calls_eval(comp, "                ")
    lines += ["            }",
              "            h = rf_body_hash(a, lane, h, (uint64_t)sid, (uint64_t)cyc);",
              "        }",
              "        a.out_hash[sid] = h;",
              f"        q = {atomic(comp, 'add', '&a.flags[RF_FLAG_QUEUE]', '1u')};",
              "    }",
              "}"]
    out.append("\n".join(lines))
    return out


def _change_code(comp: Compiled, gen: Gen, result: str):
    """change = OR (x != x_old); x_old = x (src/V3Changed.cpp:55-70, 313-316)."""
    gen.emit(f"uint32_t {result} = 0u;")
    for v in comp.sch.change_vars:
        old = comp.chglast[v]
        gen.ensure(v)
        gen.ensure(old)
        if is_wide(v.width):
            gen.emit(f"{result} |= rfw_eq({gen.name(v)}, {gen.name(old)}, {v.width}) ^ 1u;")
            gen.emit(f"rfw_copy({gen.name(old)}, {gen.name(v)}, {v.width});")
        else:
            gen.emit(f"{result} |= (uint32_t)({gen.get(v)} != {gen.get(old)});")
            gen.emit(gen.set(old, gen.get(v)))
        gen.dirty.add(old)


def _hash_code(comp: Compiled, gen: Gen, h: str, sid: Optional[str] = None, cyc: Optional[str] = None):
    """Fold the outputs into hash ``h``; with ``sid``/``cyc``, also store
    each output's 64-bit words into the output capture (rf_args.out)."""
    cap = []
    n = 0
    for port in comp.port_table():
        if not port["hashed"]:
            continue
        v = port["var"]
        gen.ensure(v)
        c = class_of(v)
        if is_wide(v.width):
            for w in range(v.words):
                gen.emit(f"{h} = rf_fnv_bytes({h}, (uint64_t){gen.name(v)}[{w}], 4);")
            for j in range((v.width + 63) // 64):
                hi = (f" | ((uint64_t){gen.name(v)}[{2 * j + 1}] << 32)" if 2 * j + 1 < v.words else "")
                cap.append(f"rf_capture(a, {sid}, {cyc}, {n}u, (uint64_t){gen.name(v)}[{2 * j}]{hi});")
                n += 1
        else:
            gen.emit(f"{h} = rf_fnv_bytes({h}, (uint64_t){gen.get(v)}, {CLASS_BYTES[c]});")
            cap.append(f"rf_capture(a, {sid}, {cyc}, {n}u, (uint64_t){gen.get(v)});")
            n += 1
    if sid is not None and cap:
        gen.emit("if (a.out.data) {")
        for ln in cap:
            gen.emit("    " + ln)
        gen.emit("}")


# cycle loops fetch the next cycle's input words one cycle ahead, into
# registers, when the design has at most this many
PREFETCH_WORDS = 8


def _input_words(comp: Compiled) -> List[Tuple[int, int]]:
    """(generator key, stream word) of every input word, in stream order."""
    out = []
    for port in comp.port_table():
        if port["dir"] == 0:
            for j in range((port["width"] + 63) // 64):
                out.append(((port["stim_index"] << 3) | j, len(out)))
    return out


def _stim_prefetch(comp: Compiled, gen: Gen, sid: str, cyc: str, live: str, declare: bool) -> bool:
    """Fetch the input words of cycle ``cyc`` into registers rfin<n> (the
    double buffer of a cycle loop: issued a cycle before their use, so the
    loads of a stream overlap the current cycle's logic). ``live``: the lane
    runs that cycle. Returns False (and emits nothing) for designs with more
    than PREFETCH_WORDS input words."""
    words = _input_words(comp)
    if not words or len(words) > PREFETCH_WORDS:
        return False
    for key, n in words:
        lhs = f"uint64_t rfin{n}" if declare else f"rfin{n}"
        gen.emit(f"{lhs} = rf_input(a, {sid}, {cyc}, {key}u, {n}u, {live});")
    return True


def _stim_code(comp: Compiled, gen: Gen, sid: str, cyc: str, prefetched: bool = False, io: bool = True):
    """Apply one cycle's inputs: from rf_input, or from the registers
    _stim_prefetch filled when ``prefetched``; without ``io`` from the
    generator only (no stream code in the kernel)."""
    n = 0
    for port in comp.port_table():
        if port["dir"] != 0:
            continue
        v = port["var"]
        gen.declare(v)
        p = port["stim_index"]
        if is_wide(v.width):
            for j in range((v.width + 63) // 64):
                src = f"rfin{n}" if prefetched else (f"rf_input(a, {sid}, {cyc}, {(p << 3) | j}u, {n}u)" if io
                                                     else f"rf_stim(a.seed, {sid}, {cyc}, {(p << 3) | j}u)")
                gen.emit(f"{{ const uint64_t r = {src};")
                n += 1
                gen.emit(f"  {gen.name(v)}[{2 * j}] = (uint32_t)r;")
                if 2 * j + 1 < v.words:
                    gen.emit(f"  {gen.name(v)}[{2 * j + 1}] = (uint32_t)(r >> 32); }}")
                else:
                    gen.emit("}")
            gen.emit(f"rfw_clean({gen.name(v)}, {v.width});")
        else:
            src = f"rfin{n}" if prefetched else (f"rf_input(a, {sid}, {cyc}, {p << 3}u, {n}u)" if io
                                                 else f"rf_stim(a.seed, {sid}, {cyc}, {p << 3}u)")
            gen.emit(gen.set(v, f"({cty(v.width)})({src} & {lit((1 << v.width) - 1, 64)})"))
            n += 1
        gen.avail.add(v)
        gen.dirty.add(v)


def emit_eval_end(comp: Compiled) -> str:
    d = comp.design
    gen = Gen(comp)
    gen.top_lines = []
    gen.local = set(d.clocks) | set(comp.clklast.values()) | set(comp.sch.change_vars) | \
        set(comp.chglast.values()) | set(d.outputs)
    body = gen
    # _last_assign
    for clk, last in comp.clklast.items():
        gen.ensure(clk)
        gen.declare(last)
        gen.emit(gen.set(last, gen.get(clk)))
        gen.avail.add(last)
        gen.dirty.add(last)
    if comp.sch.change_vars:
        _change_code(comp, gen, "ch")
        gen.emit("a.change[lane] = ch;")
        gen.emit(wave_or_flag(comp, "ch != 0u", "&a.flags[RF_FLAG_CHANGE]"))
        gen.emit(f"if (ch && a.phase != RF_PHASE_EVAL) {atomic(comp, 'or', '&a.flags[RF_FLAG_UNCONVERGED]', '1u')};")
        gen.emit("if (a.phase != RF_PHASE_EVAL) a.change[lane] = 1u;")
    gen.emit("if (a.phase == RF_PHASE_RISE) {")
    gen.ind += 1
    for clk in comp.in_clocks:
        gen.declare(clk)
        gen.emit(gen.set(clk, "1u"))
        gen.dirty.add(clk)
    gen.ind -= 1
    gen.emit("} else if (a.phase == RF_PHASE_HASH) {")
    gen.ind += 1
    gen.emit("uint64_t h = a.hash[lane];")
    _hash_code(comp, gen, "h", "(uint64_t)a.stim[lane]", "(uint64_t)a.cyc[lane]")
    gen.emit("a.hash[lane] = h;")
    gen.emit("a.cyc[lane] += 1u;")
    gen.ind -= 1
    gen.emit("}")
    for v in sorted(gen.dirty, key=lambda x: x.id):
        gen.store(v)
    lines = [kernel_header(comp, f"rf_k_eval_end")]
    _prologue(comp, gen, lines, check_done=True)
    lines += gen.top_lines + gen.lines + ["}"]
    return "\n".join(lines)


def emit_stim(comp: Compiled) -> str:
    d = comp.design
    gen = Gen(comp)
    gen.top_lines = []
    gen.local = set(d.inputs)
    _stim_code(comp, gen, "(uint64_t)a.stim[lane]", "(uint64_t)a.cyc[lane]")
    for clk in comp.in_clocks:
        gen.declare(clk)
        gen.emit(gen.set(clk, "0u"))
        gen.dirty.add(clk)
    if comp.sch.change_vars:
        gen.emit("a.change[lane] = 1u;")
    for v in sorted(gen.dirty, key=lambda x: x.id):
        gen.store(v)
    lines = [kernel_header(comp, f"rf_k_stim")]
    _prologue(comp, gen, lines, check_done=True)
    lines += gen.top_lines + gen.lines + ["}"]
    return "\n".join(lines)


def _all_comb_vids(comp: Compiled) -> List[int]:
    return [vid for vid in comp.sch.topo if comp.sch.vertices[vid].kind == "comb"]


def _initial_body(comp: Compiled, gen: Gen):
    """Initial blocks, clock latch, then the settle loop (src/V3Clock.cpp:228-270)."""
    sch = comp.sch
    for p in sch.initial:
        gen.process(p, sch.access[p.id], f"i{p.id}")
    for clk, last in comp.clklast.items():
        gen.ensure(clk)
        gen.declare(last)
        gen.emit(gen.set(last, gen.get(clk)))
        gen.avail.add(last)
        gen.dirty.add(last)
    comb = _all_comb_vids(comp)
    if comp.sch.change_vars:
        gen.emit("uint32_t it = 0u;")
        # preload every variable the loop body touches so the loop sees a
        # consistent register state on its back edge
        for v in sorted(_vertex_vars(comp, comb) | set(comp.chglast.values()) |
                        set(comp.sch.change_vars), key=lambda x: x.id):
            if v in comp.stored:
                gen.ensure(v)
            else:
                gen.declare(v)
                gen.avail.add(v)
        gen.emit("for (;;) {")
        gen.ind += 1
        _emit_vertices(comp, gen, comb, store_in_guard=False)
        _change_code(comp, gen, "ch")
        gen.emit(f"if (!ch) break;")
        gen.emit(f"if (++it >= a.converge_limit) {{ {atomic(comp, 'or', '&a.flags[RF_FLAG_UNCONVERGED]', '1u')}; break; }}")
        gen.ind -= 1
        gen.emit("}")
    else:
        _emit_vertices(comp, gen, comb, store_in_guard=False)


def emit_init(comp: Compiled) -> str:
    gen = Gen(comp)
    gen.top_lines = []
    sch = comp.sch
    loc = _vertex_vars(comp, list(range(len(sch.vertices))))
    for p in sch.initial:
        acc = sch.access[p.id]
        loc |= {v for v in acc.dep_reads | acc.need_old | acc.writes if not v.is_mem}
    loc |= set(comp.design.clocks) | set(comp.clklast.values()) | set(comp.chglast.values()) | \
        set(sch.change_vars)
    gen.local = loc
    plan_packing(comp, gen, comp.stored)
    _initial_body(comp, gen)
    for v in sorted(gen.dirty, key=lambda x: x.id):
        if v in comp.stored:
            gen.store(v)
    lines = [kernel_header(comp, f"rf_k_init")]
    _prologue(comp, gen, lines, check_done=False)
    lines += gen.top_lines + gen.lines + ["}"]
    return "\n".join(lines)


def emit_settle(comp: Compiled) -> Optional[str]:
    """Settle tail of the graph form (designs with change-detected variables
    only). The cycle graph holds ``settle_unroll`` passes of the mtask DAG;
    a lane whose change request is still set after them finishes the settle
    loop here, one thread per lane, pass after pass up to the converge limit.
    The reference replays the whole mtask graph until change[0] clears, up
    to 100 times (src/V3EmitC.cpp:5474-5491); lanes are independent and a
    settled lane's mtasks are skipped (src/V3EmitC.cpp:1654), so finishing
    each lane on its own gives the same values. ``a.ncycles`` carries the
    passes already run."""
    if not comp.sch.change_vars:
        return None
    if comp.calls:
        lim = atomic(comp, "or", "&a.flags[RF_FLAG_UNCONVERGED]", "1u")
        lines = [kernel_header(comp, "rf_k_settle"),
                 f"    const uint32_t lane = {lane_expr(comp)};",
                 "    if (lane >= a.nlanes) return;",
                 "    if (a.done[lane] || !a.change[lane]) return;",
                 "    for (uint32_t it = a.ncycles;; ++it) {",
                 f"        if (it >= a.converge_limit) {{ {lim}; break; }}"]
        lines += [f"        rf_body_{k}(a, lane);" for k in range(len(comp.mtasks))]
        lines += ["        if (!rf_body_latch(a, lane)) break;",
                  f"        if (it + 1u >= a.converge_limit) {{ {lim}; break; }}",
                  "    }",
                  "}"]
        return "\n".join(lines)
    gen = Gen(comp)
    gen.top_lines = []
    sch = comp.sch
    loc = _vertex_vars(comp, list(range(len(sch.vertices))))
    loc |= set(comp.design.clocks) | set(comp.clklast.values()) | set(comp.chglast.values()) | \
        set(sch.change_vars)
    gen.local = loc
    plan_packing(comp, gen, comp.stored)
    # every value the loop touches is in a register before the loop, so its
    # back edge sees one consistent register state
    for v in sorted(loc, key=lambda x: x.id):
        if v in comp.stored:
            gen.ensure(v)
        else:
            gen.declare(v)
            gen.avail.add(v)
    _eval_loop_body(comp, gen, start="a.ncycles")
    for v in sorted(gen.dirty, key=lambda x: x.id):
        if v in comp.stored:
            gen.store(v)
    lines = [kernel_header(comp, "rf_k_settle")]
    _prologue(comp, gen, lines, check_done=True, check_change=True)
    lines += gen.top_lines + gen.lines + ["}"]
    return "\n".join(lines)


def clk0_cone(comp: Compiled) -> Optional[List[int]]:
    """The vertices a cycle's clk = 0 eval must run once the design has
    settled: the combinational logic that reads the inputs (new this cycle)
    or the clock, directly or through other such logic. Nothing clocked
    fires on the falling clock and the state has not changed since the last
    settle, so every other value still holds. None when that does not hold
    (negedge blocks, derived clocks, combinational cycles)."""
    sch = comp.sch
    d = comp.design
    if sch.change_vars or any(c.kind != "input" for c in d.clocks):
        return None
    for vx in sch.vertices:
        if vx.kind == "seq" and any(e != "posedge" for p in vx.procs for e, _ in p.edges):
            return None
    # logic placed before a writer of what it reads (the combinational
    # inputs of clocked blocks that also read their state) last ran on the
    # old value
    pos = {vid: k for k, vid in enumerate(sch.topo)}
    last_write: Dict[Var, int] = {}
    for vid in sch.topo:
        for v in sch.vertices[vid].writes:
            last_write[v] = pos[vid]
    changed: Set[Var] = set(d.inputs)
    out = []
    for vid in sch.topo:
        vx = sch.vertices[vid]
        if vx.kind != "comb":
            continue
        reads = vx.reads | vx.need_old
        if reads & changed or any(last_write.get(v, -1) > pos[vid] for v in reads):
            out.append(vid)
            changed |= vx.writes
    return out


def clk0_carried(comp: Compiled, cone: Optional[List[int]]) -> Set[Var]:
    """Values the clk = 0 eval reads without computing them (written by
    logic outside ``cone``): they must keep their value from the previous
    eval, so they cannot share a register slot with shorter-lived values."""
    if cone is None:
        return set()
    sch = comp.sch
    inside = set(cone)
    read: Set[Var] = set()
    for vid in cone:
        read |= sch.vertices[vid].reads | sch.vertices[vid].need_old
    out: Set[Var] = set()
    for vx in sch.vertices:
        if vx.id not in inside:
            out |= vx.writes & read
    return out


def _eval_loop_body_split(comp: Compiled, gen: Gen, first: str, cone: Optional[List[int]]):
    """The two evals of a cycle (``ph`` = the clock): the clk = 0 eval runs
    only ``cone`` (clk0_cone()) except when ``first`` (the first cycle a
    kernel runs for a lane, whose combinational values it has not computed
    yet)."""
    if cone is None:
        _eval_loop_body(comp, gen)
        return
    gen.emit(f"if (ph == 0u && !({first})) {{")
    gen.ind += 1
    _emit_vertices(comp, gen, cone, store_in_guard=False)
    for clk, last in comp.clklast.items():
        gen.emit(gen.set(last, gen.get(clk)))
        gen.dirty.add(last)
    gen.ind -= 1
    gen.emit("} else {")
    gen.ind += 1
    _eval_loop_body(comp, gen)
    gen.ind -= 1
    gen.emit("}")


def _eval_loop_body(comp: Compiled, gen: Gen, start: str = "0u"):
    """One eval() with the settle loop on device (src/V3EmitC.cpp:5426-5496).
    ``start``: settle passes the eval has already run (the graph form's
    unrolled passes, for the settle-tail kernel)."""
    all_vids = list(comp.sch.topo)
    if comp.sch.change_vars:
        gen.emit(f"for (uint32_t it = {start};; ++it) {{")
        gen.ind += 1
        if start != "0u":
            gen.emit(f"if (it >= a.converge_limit) {{ {atomic(comp, 'or', '&a.flags[RF_FLAG_UNCONVERGED]', '1u')}; break; }}")
    _emit_vertices(comp, gen, all_vids, store_in_guard=False)
    for clk, last in comp.clklast.items():
        gen.emit(gen.set(last, gen.get(clk)))
        gen.dirty.add(last)
    if comp.sch.change_vars:
        _change_code(comp, gen, "ch")
        gen.emit("if (!ch) break;")
        gen.emit(f"if (it + 1u >= a.converge_limit) {{ {atomic(comp, 'or', '&a.flags[RF_FLAG_UNCONVERGED]', '1u')}; break; }}")
        gen.ind -= 1
        gen.emit("}")


def plan_packing(comp: Compiled, gen: Gen, carried: Set[Var]):
    """Pack the 1-bit register variables of an inline kernel into bit slots of
    32-bit registers (a 20,000-gate netlist keeps thousands of 1-bit values
    live, far more than the 512 registers a lane has as whole words).

    ``carried`` variables keep their value across evals (state, clock latches,
    ports, combinational values read by clocked logic) or are written back to
    the lane-minor arrays when the kernel ends (every stored variable, so the
    arrays hold the settled state): each gets a slot of its own. Every other 1-bit variable lives only inside one eval; its slot
    is taken at its defining vertex and freed after its last reader in the
    eval's topological order (linear-scan allocation)."""
    sch = comp.sch
    bits = [v for v in gen.local if v.width == 1]
    if len(bits) < 64:
        return
    pos = {vid: k for k, vid in enumerate(sch.topo)}
    first_def: Dict[Var, int] = {}
    last_use: Dict[Var, int] = {}
    for vid in sch.topo:
        vx = sch.vertices[vid]
        k = pos[vid]
        for v in vx.writes:
            if v not in first_def:
                first_def[v] = k
        for v in vx.reads | vx.need_old:
            last_use[v] = max(last_use.get(v, -1), k)
    dedicated = sorted((v for v in bits if v in carried or v not in first_def),
                       key=lambda x: x.id)
    transient = [v for v in bits if v not in carried and v in first_def]
    slot = 0
    packed: Dict[Var, Tuple[str, int]] = {}
    for v in dedicated:
        packed[v] = (f"pk{slot >> 5}", slot & 31)
        slot += 1
    base = (slot + 31) & ~31
    # linear scan over the eval order
    starts: Dict[int, List[Var]] = {}
    ends: Dict[int, List[Var]] = {}
    for v in transient:
        starts.setdefault(first_def[v], []).append(v)
        ends.setdefault(max(last_use.get(v, first_def[v]), first_def[v]), []).append(v)
    free: List[int] = []
    nxt = base
    import heapq
    for k in range(len(sch.topo)):
        for v in sorted(starts.get(k, []), key=lambda x: x.id):
            if free:
                s = heapq.heappop(free)
            else:
                s = nxt
                nxt += 1
            packed[v] = (f"pk{s >> 5}", s & 31)
        for v in ends.get(k, []):
            heapq.heappush(free, packed[v][0] and (int(packed[v][0][2:]) * 32 + packed[v][1]))
    gen.packed = packed
    comp.pack_stats = {"dedicated_bits": len(dedicated), "transient_bits": len(transient),
                       "registers": (nxt + 31) // 32}


def _fused_locals(comp: Compiled) -> Set[Var]:
    sch = comp.sch
    loc = _vertex_vars(comp, list(range(len(sch.vertices))))
    loc |= set(comp.design.clocks) | set(comp.clklast.values()) | set(comp.chglast.values()) | \
        set(sch.change_vars) | set(comp.design.inputs) | set(comp.design.outputs)
    return loc


# memories of at most this many bytes per lane are held in LDS by the
# whole-cycle kernels (fused and persistent), within this many bytes of LDS
# per workgroup
LDS_MEM_MAX_BYTES = 256
LDS_MEM_BUDGET = 32768


def lds_memories(comp: Compiled) -> Dict[Var, str]:
    """Small memories (a register file) that the whole-cycle kernels keep in
    LDS: one column per thread, so a lane's reads and writes cost an LDS
    access instead of a global one. Not in the host build or the calls form
    (whose mtask bodies also serve the per-mtask kernels)."""
    if comp.host or comp.calls:
        return {}
    out: Dict[Var, str] = {}
    total = 0
    for m in comp.design.vars:
        if not (m.is_mem and comp.layout.has(m)):
            continue
        b = m.depth * m.words * CLASS_BYTES[class_of(m)]
        if b <= LDS_MEM_MAX_BYTES and total + b * comp.opts.block <= LDS_MEM_BUDGET:
            out[m] = f"lm{m.id}"
            total += b * comp.opts.block
    return out


def _lds_decl_lines(comp: Compiled, lds: Dict[Var, str]) -> List[str]:
    return [f"    __shared__ {CTY[class_of(m)]} {nm}[{m.depth * m.words}][{comp.opts.block}];"
            for m, nm in lds.items()]


# whole-cycle kernels keep at most this many words of variables in
# registers; beyond it, variables move to LDS (within LDS_STATE_BUDGET bytes
# per workgroup, small memories included)
REG_WORDS = 160
LDS_STATE_BUDGET = 144 * 1024


def place_lds_state(comp: Compiled, gen: Gen) -> List[str]:
    """Move register variables of a whole-cycle kernel to LDS when they
    exceed REG_WORDS words (a design whose state would otherwise spill to
    scratch memory, which is global memory). Wide variables move first, then
    narrow ones in declaration order, while the workgroup's LDS budget
    lasts. Narrow variables live in one array per class, [slot][thread]
    (a wavefront's accesses to one variable are consecutive); wide ones in a
    thread-major array with an odd row length (a wavefront's accesses to one
    word fall in distinct banks). Returns the declarations."""
    if comp.host:
        return []
    B = comp.opts.block
    regs = [v for v in gen.local if v not in gen.packed]
    words = sum(v.words if is_wide(v.width) else (2 if v.width > 32 else 1) for v in regs)
    if words <= REG_WORDS:
        return []
    used = sum(m.depth * m.words * CLASS_BYTES[class_of(m)] for m in gen.lds) * B
    order = sorted((v for v in regs if is_wide(v.width)), key=lambda v: (-v.words, v.id)) + \
        sorted((v for v in regs if not is_wide(v.width)), key=lambda v: v.id)
    nslot = {"C": 0, "S": 0, "I": 0, "Q": 0}
    wide_off = 0
    moved_narrow, moved_wide = [], []
    for v in order:
        if words <= REG_WORDS:
            break
        if is_wide(v.width):
            cost = v.words * 4 * B
            if used + cost > LDS_STATE_BUDGET:
                continue
            moved_wide.append((v, wide_off))
            wide_off += v.words
            words -= v.words
        else:
            c = class_of(v)
            cost = CLASS_BYTES[c] * B
            if used + cost > LDS_STATE_BUDGET:
                continue
            moved_narrow.append((v, c, nslot[c]))
            nslot[c] += 1
            words -= 2 if v.width > 32 else 1
        used += cost
    lines = []
    for c in ("C", "S", "I", "Q"):
        if nslot[c]:
            lines.append(f"    __shared__ {CTY[c]} ls{c}[{nslot[c]}][{B}];")
    for v, c, k in moved_narrow:
        gen.lds_vars[v] = f"ls{c}[{k}][threadIdx.x]"
    if moved_wide:
        row = wide_off | 1
        lines.append(f"    __shared__ uint32_t lsW[{B}][{row}];")
        for v, off in moved_wide:
            gen.lds_vars[v] = f"&lsW[threadIdx.x][{off}]"
    return lines


def _lds_copy_lines(comp: Compiled, lds: Dict[Var, str], to_lds: bool) -> List[str]:
    """Copy this lane's small memories between their storage and LDS."""
    lay = comp.layout
    out = []
    for m, nm in lds.items():
        n = m.depth * m.words
        if m in lay.lane_major:
            src = f"mb{m.id}[k]"
        else:
            src = f"{ARR[class_of(m)]}[((size_t){lay.loc[m]} + k) * NP]"
        a, b = (f"{nm}[k][threadIdx.x]", src) if to_lds else (src, f"{nm}[k][threadIdx.x]")
        out.append(f"    for (uint32_t k = 0; k < {n}u; ++k) {a} = {b};")
    return out


def emit_fused(comp: Compiled, io: bool = False) -> str:
    """Protocol kernel: ``a.ncycles`` cycles per lane with all logic inlined.
    ``io``: the variant that reads an input stream (inputs fetched a cycle
    ahead) and fills an output capture (rf_k_fused_io); the plain variant has
    neither in its cycle loop."""
    d = comp.design
    gen = Gen(comp)
    gen.lds = lds_memories(comp)
    gen.top_lines = []
    gen.local = _fused_locals(comp)
    cone = clk0_cone(comp)
    plan_packing(comp, gen, comp.stored | clk0_carried(comp, cone))
    lds_state = place_lds_state(comp, gen)
    # registers that carry values across evals are loaded once
    for v in sorted(gen.local, key=lambda x: x.id):
        if v in comp.persistent:
            gen.ensure(v)
        else:
            gen.declare(v)
            gen.avail.add(v)
    gen.emit("uint64_t h = a.hash[lane];")
    gen.emit("uint32_t cyc = a.cyc[lane];")
    gen.emit("const uint64_t sid = (uint64_t)a.stim[lane];")
    pf = io and _stim_prefetch(comp, gen, "sid", "(uint64_t)cyc", "a.ncycles != 0u", True)
    gen.emit("for (uint32_t n = 0u; n < a.ncycles; ++n, ++cyc) {")
    gen.ind += 1
    _stim_code(comp, gen, "sid", "(uint64_t)cyc", pf, io)
    if pf:
        _stim_prefetch(comp, gen, "sid", "(uint64_t)cyc + 1u", "n + 1u < a.ncycles", False)
    gen.emit("#pragma unroll 1")
    gen.emit("for (uint32_t ph = 0u; ph < 2u; ++ph) {")
    gen.ind += 1
    for clk in comp.in_clocks:
        gen.emit(gen.set(clk, "ph"))
        gen.dirty.add(clk)
    _eval_loop_body_split(comp, gen, "n == 0u", cone)
    gen.ind -= 1
    gen.emit("}")
    if io:
        _hash_code(comp, gen, "h", "sid", "(uint64_t)cyc")
    else:
        _hash_code(comp, gen, "h")
    gen.ind -= 1
    gen.emit("}")
    gen.emit("a.hash[lane] = h;")
    gen.emit("a.cyc[lane] = cyc;")
    _opaque_bases(gen)
    for v in sorted(gen.dirty | (gen.local & comp.stored), key=lambda x: x.id):
        if v in comp.stored:
            gen.store(v)
    lines = [kernel_header(comp, "rf_k_fused_io" if io else "rf_k_fused")]
    lines += _lds_decl_lines(comp, gen.lds) + lds_state
    _prologue(comp, gen, lines, check_done=True)
    lines += _lds_copy_lines(comp, gen.lds, to_lds=True)
    lines += gen.top_lines + gen.lines
    lines += _lds_copy_lines(comp, gen.lds, to_lds=False) + ["}"]
    return "\n".join(lines)


# ---------------------------------------------------------------------------
# Multi-wavefront fused kernel: one workgroup serves 64 stimuli with W
# wavefronts. An eval runs the mtask DAG level by level (a level: the mtasks
# whose longest predecessor chain has the same length); each wavefront runs
# its share of a level's mtasks for the 64 lanes, values pass between
# mtasks through the lane-minor arrays (write-through stores), and a barrier
# closes every level. Wavefront 0 applies the inputs, sets the clock, runs
# the clock latch and hashes the outputs. This is the graph form's per-mtask
# code and dependence structure with the cycle loop inside one kernel, for
# designs whose per-stimulus logic is wider than one wavefront's issue.
# ---------------------------------------------------------------------------

MW_MAX_WAVES = 16


def mw_plan(comp: Compiled) -> Tuple[int, List[List[List[int]]]]:
    """(W, levels): the wavefront count (total mtask cost over the costliest
    dependence chain, at most MW_MAX_WAVES and the widest level) and per
    level per wavefront the mtask indices, assigned costliest first to the
    least loaded wavefront."""
    mts = comp.mtasks
    idx = {mt.id: k for k, mt in enumerate(mts)}
    lev: Dict[int, int] = {}
    cp: Dict[int, int] = {}

    def walk(k: int):
        if k not in lev:
            ps = [idx[p] for p in mts[k].preds if p in idx]
            for q in ps:
                walk(q)
            lev[k] = 1 + max((lev[q] for q in ps), default=-1)
            cp[k] = max(1, mts[k].cost) + max((cp[q] for q in ps), default=0)
    for k in range(len(mts)):
        walk(k)
    nl = 1 + max(lev.values(), default=-1)
    by_level = [[k for k in range(len(mts)) if lev[k] == L] for L in range(nl)]
    total = sum(max(1, mt.cost) for mt in mts)
    crit = max(cp.values(), default=1)
    cap = min(MW_MAX_WAVES, max(1, comp.opts.mw_max_waves))
    W = max(1, min(cap, -(-total // crit), max((len(b) for b in by_level), default=1)))
    out = []
    for ks in by_level:
        load = [0] * W
        waves: List[List[int]] = [[] for _ in range(W)]
        for k in sorted(ks, key=lambda k: -max(1, mts[k].cost)):
            w = min(range(W), key=lambda i: load[i])
            waves[w].append(k)
            load[w] += max(1, mts[k].cost)
        out.append([sorted(x) for x in waves])
    return W, out


def mw_waves(comp: Compiled) -> int:
    """Wavefronts of the design's rf_k_fused_mw, 0 when it has none."""
    if comp.host or comp.calls or not comp.mtasks or comp.sch.change_vars:
        return 0
    W = mw_plan(comp)[0]
    return W if W >= 2 else 0


# LDS of one multi-wavefront workgroup that may hold its lanes' slots
MW_LDS_MAX = 150 * 1024


def mw_lds_slots(comp: Compiled) -> Optional[Dict[str, int]]:
    """Slots per class of the design when the 64 lanes' worth of them fits
    the multi-wavefront kernel's LDS (the kernel then keeps them there for a
    launch), else None (slots stay in the lane-minor arrays)."""
    sl = comp.layout.slots
    total = sum(sl[c] * CLASS_BYTES[c] for c in ("C", "S", "I", "Q")) * 64
    return dict(sl) if total <= MW_LDS_MAX else None


def _mw_chunk(gen: Gen, ind: str) -> List[str]:
    """A generated code fragment as one scope."""
    return [ind + "{"] + [ind + "    " + ln.strip() for ln in gen.top_lines + gen.lines] + [ind + "}"]


def emit_fused_mw(comp: Compiled) -> Tuple[Optional[str], int]:
    """(source of rf_k_fused_mw, W), or (None, 0) where the form does not
    apply: the host build, the calls form, designs with change-detected
    variables (their settle loop stays with the one-wavefront forms) and
    designs with too little parallel logic for two wavefronts."""
    if not mw_waves(comp):
        return None, 0
    W, levels = mw_plan(comp)
    d = comp.design
    sl = mw_lds_slots(comp)
    L = [f'extern "C" __global__ void __launch_bounds__({64 * W}) rf_k_fused_mw(rf_args a) {{',
         "    const uint32_t wave = threadIdx.x >> 6;",
         "    const uint32_t t64 = threadIdx.x & 63u;",
         "    const uint32_t lane = blockIdx.x * 64u + t64;",
         "    const bool live = lane < a.nlanes && a.done[lane] == 0u;"]
    if sl is None:
        L += ["    size_t NP = a.npad;",
              "    uint8_t* __restrict__ sC = a.c + lane;",
              "    uint16_t* __restrict__ sS = a.s + lane;",
              "    uint32_t* __restrict__ sI = a.i + lane;",
              "    uint64_t* __restrict__ sQ = a.q + lane;"]
    else:
        # the workgroup's 64 lanes of every slot in LDS, [slot][lane]: the
        # mtask bodies' slot accesses become LDS accesses (stride 64), and
        # values cross wavefronts there; loaded at the start of a launch and
        # written back at its end
        L.append("    constexpr size_t NP = 64;")
        for c in ("C", "S", "I", "Q"):
            if sl[c]:
                L.append(f"    __shared__ {CTY[c]} lsoa{c}[{sl[c] * 64}];")
                L.append(f"    {CTY[c]}* __restrict__ s{c} = lsoa{c} + t64;")
            else:
                L.append(f"    {CTY[c]}* __restrict__ s{c} = nullptr;")
        L.append("    {")
        L.append("        const size_t g0 = (size_t)blockIdx.x * 64u;")
        for c in ("C", "S", "I", "Q"):
            if sl[c]:
                L.append(f"        for (uint32_t k = threadIdx.x; k < {sl[c] * 64}u; k += {64 * W}u)")
                L.append(f"            lsoa{c}[k] = a.{c.lower()}[(size_t)(k >> 6) * a.npad + g0 + (k & 63u)];")
        L.append("    }")
        L.append("    __syncthreads();")
    for m, off in comp.layout.lane_major.items():
        c = class_of(m)
        row = m.depth * (CLASS_BYTES[m.cls] if m.cls != "W" else 4 * m.words)
        L.append(f"    {CTY[c]}* __restrict__ mb{m.id} = ({CTY[c]}*)(a.mem + (size_t){off} * a.npad"
                 f" + (size_t)lane * {row}u);")
    L += ["    (void)sC; (void)sS; (void)sI; (void)sQ; (void)NP;",
          "    uint64_t h = 0ull, sid = 0ull;",
          "    uint32_t cyc = 0u;",
          "    if (live) {",
          "        h = a.hash[lane];",
          "        cyc = a.cyc[lane];",
          "        sid = (uint64_t)a.stim[lane];",
          "    }",
          "    for (uint32_t n = 0u; n < a.ncycles; ++n, ++cyc) {"]

    def wgen(local) -> Gen:
        g = Gen(comp)
        g.top_lines = []
        g.write_through = True
        g.local = set(local)
        return g

    g = wgen(d.inputs)
    _stim_code(comp, g, "sid", "(uint64_t)cyc", False, False)
    for v in sorted(g.dirty, key=lambda x: x.id):
        g.store(v)
    L.append("        if (wave == 0u && live)")
    L += _mw_chunk(g, "        ")
    L += ["        #pragma unroll 1",
          "        for (uint32_t ph = 0u; ph < 2u; ++ph) {"]
    g = wgen(comp.in_clocks)
    for clk in comp.in_clocks:
        g.declare(clk)
        g.emit(g.set(clk, "ph"))
        g.store(clk)
    L.append("            if (wave == 0u && live)")
    L += _mw_chunk(g, "            ")
    L.append("            __syncthreads();")

    def level_lines(lvs, ind):
        out = []
        for lv in lvs:
            busy = [(w, ks) for w, ks in enumerate(lv) if ks]
            if not busy:
                continue
            out.append(ind + "if (live) {")
            for i, (w, ks) in enumerate(busy):
                out.append(ind + f"    {'if' if i == 0 else '} else if'} (wave == {w}u) {{")
                for k in ks:
                    out.extend(_mw_chunk(_mtask_gen(comp, k), ind + "        "))
            out += [ind + "    }", ind + "}", ind + "__syncthreads();"]
        return out

    cone = clk0_cone(comp)
    if cone is None:
        L += level_lines(levels, "            ")
    else:
        # after the first cycle of a launch, the clk = 0 eval runs only the
        # mtasks holding falling-edge cone logic (clk0_cone()): every other
        # value still holds in the slots since the last settle
        inside = set(cone)
        in_cone = [any(v in inside for v in mt.vids) for mt in comp.mtasks]
        cone_levels = [[[k for k in ks if in_cone[k]] for ks in lv] for lv in levels]
        L.append("            if (ph == 0u && n != 0u) {")
        L += level_lines(cone_levels, "                ")
        L.append("            } else {")
        L += level_lines(levels, "                ")
        L.append("            }")
    g = wgen(set(d.clocks) | set(comp.clklast.values()))
    for clk, last in comp.clklast.items():
        g.ensure(clk)
        g.declare(last)
        g.emit(g.set(last, g.get(clk)))
        g.avail.add(last)
        g.store(last)
    L.append("            if (wave == 0u && live)")
    L += _mw_chunk(g, "            ")
    L += ["            __syncthreads();", "        }"]
    g = wgen(d.outputs)
    _hash_code(comp, g, "h")
    L.append("        if (wave == 0u && live)")
    L += _mw_chunk(g, "        ")
    L += ["    }",
          "    if (wave == 0u && live) {",
          "        a.hash[lane] = h;",
          "        a.cyc[lane] = cyc;",
          "    }"]
    if sl is not None:
        L += ["    __syncthreads();",
              "    {",
              "        const size_t g0 = (size_t)blockIdx.x * 64u;"]
        for c in ("C", "S", "I", "Q"):
            if sl[c]:
                L.append(f"        for (uint32_t k = threadIdx.x; k < {sl[c] * 64}u; k += {64 * W}u)")
                L.append(f"            a.{c.lower()}[(size_t)(k >> 6) * a.npad + g0 + (k & 63u)] = lsoa{c}[k];")
        L.append("    }")
    L.append("}")
    return "\n".join(L), W


def _opaque_bases(gen: Gen):
    """Before a kernel's final stores: the slot base pointers and stride
    through an optimisation barrier (RF_OPAQUE_*), so the stores recompute
    their addresses instead of reusing the prologue's across the loop."""
    gen.emit("RF_OPAQUE_S(NP);")
    for b in ("sC", "sS", "sI", "sQ"):
        gen.emit(f"RF_OPAQUE_V({b});")


def emit_reset(comp: Compiled) -> str:
    """Zero every slot of one lane (variable reset with --x-initial 0, the
    zeroing path of the reference's _ctor_var_reset, src/V3EmitC.cpp:2042-2047)."""
    lay = comp.layout
    lines = [kernel_header(comp, f"rf_k_reset")]
    lines.append(f"    const uint32_t lane = {lane_expr(comp)};")
    lines.append("    if (lane >= a.nlanes) return;")
    lines.append("    const size_t NP = a.npad;")
    for c, arr in (("C", "a.c"), ("S", "a.s"), ("I", "a.i"), ("Q", "a.q")):
        n = lay.slots[c]
        if n:
            lines.append(f"    for (size_t k = 0; k < {n}u; ++k) {arr}[k * NP + lane] = 0;")
    if lay.lane_major_bytes:
        lines.append(f"    uint64_t* row = (uint64_t*)(a.mem + (size_t)lane * 0);")
        for m, off in lay.lane_major.items():
            row = m.depth * (CLASS_BYTES[m.cls] if m.cls != "W" else 4 * m.words)
            lines.append(f"    {{ uint8_t* p = a.mem + (size_t){off} * NP + (size_t)lane * {row}u;")
            lines.append(f"      for (uint32_t k = 0; k < {row}u; ++k) p[k] = 0; }}")
    lines.append("}")
    return "\n".join(lines)


def emit_source(comp: Compiled, host: bool = False) -> str:
    """The design library source: gfx950 kernels, or (host=True) the same
    bodies as per-lane host functions for the one-lane host build."""
    comp.host = host
    try:
        return _emit_source(comp)
    finally:
        comp.host = False


def _emit_source(comp: Compiled) -> str:
    d = comp.design
    parts = [
        "// Generated by rtlflow_amd " + ("(host build)" if comp.host else "for gfx950") +
        ". Do not edit.",
        f"// design: {d.name}",
        "" if comp.host else '#include <hip/hip_runtime.h>',
        '#include "rf_design.h"',
        '#include "rf_ops.h"',
        '#include "rf_stim.h"',
        "",
    ]
    if comp.calls:
        for b in emit_bodies(comp):
            parts.append(b)
            parts.append("")
    for k in range(len(comp.mtasks)):
        parts.append(emit_mtask_kernel(comp, k))
        parts.append("")
    parts.append(emit_eval_end(comp))
    parts.append("")
    parts.append(emit_stim(comp))
    parts.append("")
    parts.append(emit_reset(comp))
    parts.append("")
    if comp.calls:
        for kx in emit_calls_kernels(comp):
            parts.append(kx)
            parts.append("")
    else:
        parts.append(emit_init(comp))
        parts.append("")
        from .persistent import emit_persistent
        for io in (False, True):
            parts.append(emit_fused(comp, io))
            parts.append("")
            parts.append(emit_persistent(comp, io))
            parts.append("")
        mw, _ = emit_fused_mw(comp)
        if mw:
            parts.append(mw)
            parts.append("")
    tail = emit_settle(comp)
    if tail:
        parts.append(tail)
        parts.append("")
    body = "\n".join(parts)
    h = int(hashlib.sha1(body.encode()).hexdigest()[:16], 16)
    parts = [body, emit_descriptor(comp, h)]
    return "\n".join(parts)


def emit_descriptor(comp: Compiled, h: int) -> str:
    d = comp.design
    man = comp.manifest()
    out = []
    ports = comp.port_table()
    out.append("static const rf_port rf_ports[] = {")
    for p in ports:
        out.append(f'    {{"{p["name"]}", {p["dir"]}u, {p["width"]}u, {p["cls"]}u, {p["words"]}u, '
                   f'{p["loc"]}ull, {p["stim_index"]}u, {p["hashed"]}u}},')
    if not ports:
        out.append("    {0}")
    out.append("};")
    mems = [v for v in d.vars if v.is_mem and comp.layout.has(v)]
    out.append("static const rf_mem rf_mems[] = {")
    for m in mems:
        lm = m in comp.layout.lane_major
        loc = comp.layout.lane_major[m] if lm else comp.layout.loc[m]
        row = m.depth * (CLASS_BYTES[m.cls] if m.cls != "W" else 4 * m.words)
        out.append(f'    {{"{m.name}", {m.width}u, {CLS_ID[class_of(m)]}u, {m.words}u, {m.depth}u, '
                   f'{1 if lm else 0}u, {loc}ull, {row}ull}},')
    if not mems:
        out.append("    {0}")
    out.append("};")
    # mtask predecessor lists (indices into the topologically ordered array)
    id2idx = {mt.id: k for k, mt in enumerate(comp.mtasks)}
    for k, mt in enumerate(comp.mtasks):
        preds = sorted(id2idx[p] for p in mt.preds)
        body = ", ".join(f"{x}u" for x in preds) if preds else "0u"
        out.append(f"static const uint32_t rf_preds_{k}[] = {{{body}}};")
    out.append("static const rf_mtask rf_mtasks[] = {")
    for k, mt in enumerate(comp.mtasks):
        out.append(f'    {{(const void*)&rf_mtask_{k}, "rf_mtask_{k}", {len(mt.preds)}u, rf_preds_{k}, '
                   f'{mt.cost}u}},')
    if not comp.mtasks:
        out.append("    {0}")
    out.append("};")
    clk_idx = 0xffffffff
    for i, p in enumerate(ports):
        if p["dir"] == 2:
            clk_idx = i
            break
    lay = comp.layout
    if comp.bitplane_blob:
        # bit-plane program (layout: rtlflow_amd/bitplane.py, encode)
        out.append("static const uint32_t rf_bitplane[] = {")
        blob = comp.bitplane_blob
        for i in range(0, len(blob), 16):
            out.append("    " + ", ".join(f"{x}u" for x in blob[i:i + 16]) + ",")
        out.append("};")
    out.append("static const rf_design rf_the_design = {")
    out.append(f"    RF_DESIGN_ABI, {comp.opts.block}u, \"{d.name}\", 0x{h:x}ull,")
    out.append(f"    {{{lay.slots['C']}ull, {lay.slots['S']}ull, {lay.slots['I']}ull, {lay.slots['Q']}ull}},")
    out.append(f"    {lay.lane_major_bytes}ull,")
    out.append(f"    {len(ports)}u, rf_ports, {len(mems)}u, rf_mems, {len(comp.mtasks)}u, rf_mtasks,")
    out.append("    (const void*)&rf_k_init, (const void*)&rf_k_eval_end, (const void*)&rf_k_stim,")
    out.append("    (const void*)&rf_k_fused, (const void*)&rf_k_persistent, (const void*)&rf_k_reset,")
    out.append("    (const void*)&rf_k_settle," if comp.sch.change_vars else "    0,")
    # the calls form's kernels read streams through the shared stimulus and
    # hash bodies: one kernel serves both entries
    if comp.calls:
        out.append("    (const void*)&rf_k_fused, (const void*)&rf_k_persistent,")
    else:
        out.append("    (const void*)&rf_k_fused_io, (const void*)&rf_k_persistent_io,")
    out.append(f"    {len(comp.sch.change_vars)}u, {clk_idx}u,")
    out.append(f"    {man['bytes_seq']}ull, 0ull, {man['bytes_in']}ull, {man['bytes_out']}ull,")
    out.append(f"    {man['ops_per_cycle']}ull,")
    if comp.bitplane_blob:
        out.append(f"    rf_bitplane, {len(comp.bitplane_blob)}ull,")
    else:
        out.append("    0, 0ull,")
    mw_w = mw_waves(comp)
    out.append(f"    (const void*)&rf_k_fused_mw, {mw_w}u" if mw_w else "    0, 0u")
    out.append("};")
    # the only exported symbol: kernels stay library-local so several design
    # libraries can live in one process
    out.append('extern "C" __attribute__((visibility("default"))) const rf_design* '
               'rf_design_get(void) { return &rf_the_design; }')
    # the same table under a design-specific name, for the generated C++ model
    # class (several designs may then be linked into one program)
    out.append(f'extern "C" __attribute__((visibility("default"))) const rf_design* '
               f'rf_design_get_{d.name}(void) {{ return &rf_the_design; }}')
    return "\n".join(out)
