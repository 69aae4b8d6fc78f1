"""Ordering of the design's logic into one eval's dependency DAG.

Restates what Verilator's V3Order / V3Clock / V3Changed decide for the
single-clock-domain designs this backend compiles:

* clocked (seq) logic runs first in an eval, guarded by its edge condition
  ``clk & ~__Vclklast__clk`` (src/V3Clock.cpp:372-389), and reads the values
  every variable had when the eval started (non-blocking semantics);
* combinational logic then runs in data-dependency order
  (src/V3Order.cpp:1827-1930 builds the mtask bodies in this order);
* ``__Vclklast__`` is latched at the end of the eval (``_last_assign``,
  src/V3Clock.cpp:64-92, 214-226);
* combinational cycles (Verilator's UNOPTFLAT logic) are broken and the
  variables on the broken edges are change-detected: the eval repeats while
  any of them changed (``_change_request``, src/V3Changed.cpp:55-70, 263-275).

The output is a list of logic vertices (one per comb process, one per group of
seq processes that must commit together) and the DAG edges between them, which
V3Partition-style merging (partition.py) coarsens into mtasks.
"""
from __future__ import annotations

from typing import Dict, List, Optional, Set, Tuple

from .ir import (Assign, Case, Design, Expr, If, Lhs, MemRd, Op, Process, Ref, Stmt, Var, expr_reads,
                 stmt_exprs, walk_stmts)


class CompileError(Exception):
    pass


# ----------------------------------------------------------------------------
# Per-process access analysis
# ----------------------------------------------------------------------------

class Access:
    """What a process reads and writes.

    dep_reads: variables whose value from before the process is used by an
        expression (dependency-relevant reads).
    need_old: variables whose pre-process value must be available to the
        emitted code: dep_reads, plus targets written partially or not on
        every path (their other bits / old value must be kept).
    writes: every variable written. def_writes: written in full on every path.
    """
    __slots__ = ("dep_reads", "need_old", "writes", "def_writes", "nb_writes", "b_writes",
                 "mem_reads")

    def __init__(self):
        self.dep_reads: Set[Var] = set()
        self.need_old: Set[Var] = set()
        self.writes: Set[Var] = set()
        self.def_writes: Set[Var] = set()
        self.nb_writes: Set[Var] = set()
        self.b_writes: Set[Var] = set()
        self.mem_reads: Set[Var] = set()


def _expr_vars(e, out: Set[Var]):
    expr_reads(e, out)


def analyze_process(p: Process) -> Access:
    acc = Access()

    def walk(stmts: List[Stmt], defined: Set[Var], nb_def: Set[Var]):
        for s in stmts:
            if isinstance(s, Assign):
                rd: Set[Var] = set()
                _expr_vars(s.rhs, rd)
                for e in s.lhs.exprs():
                    _expr_vars(e, rd)
                for v in rd:
                    if v.is_mem:
                        acc.mem_reads.add(v)
                        acc.dep_reads.add(v)
                    elif v not in defined:
                        acc.dep_reads.add(v)
                v = s.lhs.var
                acc.writes.add(v)
                (acc.nb_writes if s.nb else acc.b_writes).add(v)
                if s.lhs.whole:
                    if s.nb:
                        nb_def.add(v)
                    else:
                        defined.add(v)
                elif not v.is_mem:
                    if v not in defined and not s.nb:
                        acc.need_old.add(v)
                    if s.nb and v not in nb_def:
                        acc.need_old.add(v)
            elif isinstance(s, If):
                rd = set()
                _expr_vars(s.cond, rd)
                for v in rd:
                    if v.is_mem:
                        acc.mem_reads.add(v)
                        acc.dep_reads.add(v)
                    elif v not in defined:
                        acc.dep_reads.add(v)
                d1, n1 = set(defined), set(nb_def)
                d2, n2 = set(defined), set(nb_def)
                walk(s.then, d1, n1)
                walk(s.els, d2, n2)
                defined |= (d1 & d2)
                nb_def |= (n1 & n2)
            elif isinstance(s, Case):
                rd = set()
                _expr_vars(s.expr, rd)
                for v in rd:
                    if v.is_mem:
                        acc.mem_reads.add(v)
                        acc.dep_reads.add(v)
                    elif v not in defined:
                        acc.dep_reads.add(v)
                outs_d = []
                outs_n = []
                has_default = False
                for labels, body in s.items:
                    if labels is None:
                        has_default = True
                    d1, n1 = set(defined), set(nb_def)
                    walk(body, d1, n1)
                    outs_d.append(d1)
                    outs_n.append(n1)
                if not has_default:
                    outs_d.append(set(defined))
                    outs_n.append(set(nb_def))
                if outs_d:
                    inter_d = set.intersection(*outs_d)
                    inter_n = set.intersection(*outs_n)
                    defined |= inter_d
                    nb_def |= inter_n

    defined: Set[Var] = set()
    nb_def: Set[Var] = set()
    walk(p.stmts, defined, nb_def)
    acc.def_writes = {v for v in acc.writes if (v in defined or v in nb_def) and not v.is_mem}
    for v in acc.writes:
        if not v.is_mem and v not in acc.def_writes:
            acc.need_old.add(v)
    acc.need_old |= {v for v in acc.dep_reads if not v.is_mem}
    for _, clk in p.edges:
        acc.dep_reads.add(clk)
        acc.need_old.add(clk)
    return acc


def split_seq_processes(design: Design) -> int:
    """Split clocked blocks into groups of independent top-level statements
    (Verilator's V3Split role for always blocks): statements stay together
    when they write the same variable or memory, or when one reads a variable
    another one assigns with a blocking assignment. Reads of non-blocking
    targets across groups become anti-dependencies in the eval DAG, which
    preserves the read-old-value semantics. Returns the number of new
    processes."""
    out: List[Process] = []
    added = 0
    for p in design.processes:
        if p.kind != "seq" or len(p.stmts) < 2:
            out.append(p)
            continue
        info = [analyze_process(Process("seq", [s], p.edges)) for s in p.stmts]
        n = len(info)
        parent = list(range(n))

        def find(x):
            while parent[x] != x:
                parent[x] = parent[parent[x]]
                x = parent[x]
            return x

        def union(a, b):
            ra, rb = find(a), find(b)
            if ra != rb:
                parent[max(ra, rb)] = min(ra, rb)

        owner: Dict[Var, int] = {}
        for i, acc in enumerate(info):
            for v in acc.writes:
                if v in owner:
                    union(i, owner[v])
                else:
                    owner[v] = i
        bw: Dict[Var, int] = {}
        for i, acc in enumerate(info):
            for v in acc.b_writes:
                if v in bw:
                    union(i, bw[v])
                else:
                    bw[v] = i
        for i, acc in enumerate(info):
            for v in acc.dep_reads | acc.need_old:
                j = bw.get(v)
                if j is not None and j != i:
                    union(i, j)
        groups: Dict[int, List] = {}
        for i, s in enumerate(p.stmts):
            groups.setdefault(find(i), []).append(s)
        if len(groups) == 1:
            out.append(p)
            continue
        for k, (_, stmts) in enumerate(sorted(groups.items())):
            out.append(Process("seq", stmts, list(p.edges), name=f"{p.name}#{k}", line=p.line))
        added += len(groups) - 1
    design.processes = out
    design.finalize()
    return added


# associative, commutative operators of equal-width operands (modulo 2^width)
ASSOC_OPS = ("xor", "or", "and", "add")


def _assoc_leaves(e: Expr, op: str, width: int, out: List[Expr]):
    if isinstance(e, Op) and e.op == op and e.width == width and all(a.width == width for a in e.args):
        for a in e.args:
            _assoc_leaves(a, op, width, out)
    else:
        out.append(e)


def _assoc_chain(terms: List[Expr], op: str, width: int) -> Expr:
    """A balanced tree of ``op`` over ``terms``."""
    while len(terms) > 1:
        nxt = [Op(op, [terms[i], terms[i + 1]], width) for i in range(0, len(terms) - 1, 2)]
        if len(terms) % 2:
            nxt.append(terms[-1])
        terms = nxt
    return terms[0]


def split_reductions(design: Design, max_terms: int) -> int:
    """Split a combinational assignment whose value is one associative
    operator (xor, or, and, add) over more than ``max_terms`` operands into
    partial terms of at most ``max_terms`` operands, each its own
    combinational process writing a new internal variable, and a final
    assignment over the partials. Exact: the operators are associative and
    commutative modulo 2^width. The partials become separate vertices of the
    eval DAG, so the partitioner and the multi-wavefront plan can spread one
    long reduction (a 256-term output XOR) over several mtasks. Returns the
    number of new processes."""
    if max_terms < 2:
        return 0
    out: List[Process] = []
    added = 0
    for p in design.processes:
        st = p.stmts[0] if p.kind == "comb" and len(p.stmts) == 1 else None
        if not (isinstance(st, Assign) and not st.nb and st.lhs.idx is None and st.lhs.lsb is None
                and isinstance(st.rhs, Op) and st.rhs.op in ASSOC_OPS):
            out.append(p)
            continue
        w = st.rhs.width
        leaves: List[Expr] = []
        _assoc_leaves(st.rhs, st.rhs.op, w, leaves)
        reads: set = set()
        expr_reads(st.rhs, reads)
        # (an assignment that reads its own target stays whole)
        if len(leaves) <= max_terms or st.lhs.var in reads:
            out.append(p)
            continue
        # partial terms, level by level, until the final assignment has at
        # most max_terms operands
        terms, lvl = leaves, 0
        while len(terms) > max_terms:
            parts: List[Expr] = []
            for k in range(0, len(terms), max_terms):
                v = Var(f"__Vpart__{st.lhs.var.name}_{lvl}_{k // max_terms}", w, False, "wire")
                design.vars.append(v)
                out.append(Process("comb", [Assign(Lhs(v), _assoc_chain(terms[k:k + max_terms], st.rhs.op, w),
                                                   False)], name=f"{p.name}#part{lvl}.{k // max_terms}",
                                   line=p.line))
                parts.append(Ref(v))
                added += 1
            terms, lvl = parts, lvl + 1
        out.append(Process("comb", [Assign(st.lhs, _assoc_chain(terms, st.rhs.op, w), False)],
                           name=p.name, line=p.line))
    if added:
        design.processes = out
        design.finalize()
    return added


# ----------------------------------------------------------------------------
# Schedule
# ----------------------------------------------------------------------------

class Vertex:
    """A logic vertex of the eval DAG: one comb process, or a group of seq
    processes that share anti-dependencies and therefore commit together."""
    __slots__ = ("id", "kind", "procs", "cost", "succs", "preds", "reads", "writes", "need_old",
                 "edge_vars")

    def __init__(self, vid: int, kind: str, procs: List[Process]):
        self.id = vid
        self.kind = kind
        self.procs = procs
        self.cost = 0
        self.succs: Set[int] = set()
        self.preds: Set[int] = set()
        self.reads: Set[Var] = set()
        self.writes: Set[Var] = set()
        self.need_old: Set[Var] = set()
        self.edge_vars: Set[Var] = set()

    def __repr__(self):
        return f"V{self.id}({self.kind},{[p.id for p in self.procs]})"


class Schedule:
    def __init__(self, design: Design):
        self.design = design
        self.access: Dict[int, Access] = {}
        self.roles: Dict[Var, str] = {}
        self.initial: List[Process] = []
        self.vertices: List[Vertex] = []
        self.change_vars: List[Var] = []
        self.topo: List[int] = []
        self.observed: Set[Var] = set()
        self.live_procs: Set[int] = set()


def build_schedule(design: Design, keep: Optional[Set[str]] = None) -> Schedule:
    """Order the design. ``keep`` names extra variables to treat as observed
    (kept alive by dead-logic elimination)."""
    sch = Schedule(design)
    procs = design.processes
    for p in procs:
        sch.access[p.id] = analyze_process(p)

    # ---- roles
    writers: Dict[Var, List[Process]] = {}
    for p in procs:
        if p.kind == "initial":
            continue
        for v in sch.access[p.id].writes:
            writers.setdefault(v, []).append(p)
    clocks = set(design.clocks)
    for v in design.vars:
        ws = writers.get(v, [])
        kinds = {p.kind for p in ws}
        if v.kind == "input":
            if ws:
                raise CompileError(f"input {v.name} is driven inside the design")
            sch.roles[v] = "clock" if v in clocks else "input"
        elif not ws:
            sch.roles[v] = "const"
        elif kinds == {"seq"}:
            sch.roles[v] = "state"
        elif kinds == {"comb"}:
            sch.roles[v] = "comb"
        else:
            raise CompileError(f"{v.name} is driven by both clocked and combinational logic")
    for c in clocks:
        # a derived (gated) clock must be combinational logic of inputs and
        # state; one that a clocked block drives (a divider) is not modelled
        if c.kind != "input" and sch.roles.get(c) != "comb":
            raise CompileError(f"clock {c.name} must be a primary input or combinational "
                               "logic (clocks driven by clocked logic are not supported)")

    # ---- dead-logic elimination (V3Dead / V3Life role): keep what reaches an output
    observed = set(design.outputs)
    if keep:
        for v in design.vars:
            if v.name in keep:
                observed.add(v)
    sch.observed = observed
    live: Set[int] = set()
    work = [v for v in observed]
    seen: Set[Var] = set()
    while work:
        v = work.pop()
        if v in seen:
            continue
        seen.add(v)
        for p in writers.get(v, []):
            if p.id in live:
                continue
            live.add(p.id)
            acc = sch.access[p.id]
            for r in acc.dep_reads | acc.need_old | acc.mem_reads:
                if r not in seen:
                    work.append(r)
    sch.live_procs = live
    for p in procs:
        if p.kind == "initial":
            sch.initial.append(p)

    comb = [p for p in procs if p.kind == "comb" and p.id in live]
    seq = [p for p in procs if p.kind == "seq" and p.id in live]

    # ---- seq groups: merge SCCs of the reader->writer anti-dependency graph
    seq_writers: Dict[Var, List[Process]] = {}
    for p in seq:
        for v in sch.access[p.id].writes:
            seq_writers.setdefault(v, []).append(p)
    sadj: Dict[int, Set[int]] = {p.id: set() for p in seq}
    for p in seq:
        acc = sch.access[p.id]
        for v in acc.dep_reads | acc.need_old:
            for w in seq_writers.get(v, []):
                if w.id != p.id:
                    sadj[p.id].add(w.id)
        # several seq writers of one variable must also be serialized
    for v, ws in seq_writers.items():
        for a, b in zip(ws, ws[1:]):
            sadj[a.id].add(b.id)
    seq_sccs = _tarjan([p.id for p in seq], sadj)
    pid2proc = {p.id: p for p in procs}

    vertices: List[Vertex] = []
    proc2vx: Dict[int, int] = {}
    for comp in seq_sccs:
        comp = sorted(comp)
        vx = Vertex(len(vertices), "seq", [pid2proc[i] for i in comp])
        vertices.append(vx)
        for i in comp:
            proc2vx[i] = vx.id
    for p in comb:
        vx = Vertex(len(vertices), "comb", [p])
        vertices.append(vx)
        proc2vx[p.id] = vx.id
    for vx in vertices:
        for p in vx.procs:
            acc = sch.access[p.id]
            vx.reads |= acc.dep_reads
            vx.writes |= acc.writes
            vx.need_old |= acc.need_old
            for _, clk in p.edges:
                vx.edge_vars.add(clk)

    def add_edge(a: int, b: int):
        if a != b:
            vertices[a].succs.add(b)
            vertices[b].preds.add(a)

    # seq ordering inside the seq subgraph (condensed)
    for p in seq:
        for q in sadj[p.id]:
            add_edge(proc2vx[p.id], proc2vx[q])

    # comb dependency graph
    comb_writers: Dict[Var, List[Process]] = {}
    for p in comb:
        for v in sch.access[p.id].writes:
            comb_writers.setdefault(v, []).append(p)
    cadj: Dict[int, Set[int]] = {p.id: set() for p in comb}
    edge_vars: Dict[Tuple[int, int], Set[Var]] = {}
    for p in comb:
        acc = sch.access[p.id]
        for v in acc.dep_reads:
            for w in comb_writers.get(v, []):
                cadj[w.id].add(p.id)
                edge_vars.setdefault((w.id, p.id), set()).add(v)
    for v, ws in comb_writers.items():
        for a, b in zip(ws, ws[1:]):
            cadj[a.id].add(b.id)
    comb_sccs = _tarjan([p.id for p in comb], cadj)
    change: Set[Var] = set()
    pos: Dict[int, int] = {}
    for comp in comb_sccs:
        cyclic = len(comp) > 1 or any(i in cadj[i] for i in comp)
        comp = sorted(comp)
        for k, i in enumerate(comp):
            pos[i] = k
        if cyclic:
            cs = set(comp)
            for a in comp:
                for b in cadj[a]:
                    if b in cs and pos[b] <= pos[a]:
                        # broken (back or self) edge: change-detect its variables
                        change |= edge_vars.get((a, b), set())
    comp_of: Dict[int, int] = {}
    for ci, comp in enumerate(comb_sccs):
        for i in comp:
            comp_of[i] = ci
    for a in comb:
        for b in cadj[a.id]:
            if comp_of[a.id] == comp_of[b] and pos[b] <= pos[a.id]:
                continue
            add_edge(proc2vx[a.id], proc2vx[b])

    # seq -> comb edges: state written by seq feeds comb readers; comb
    # variables read by seq must not be overwritten before the seq logic ran.
    comb_readers: Dict[Var, List[Process]] = {}
    for c in comb:
        cacc = sch.access[c.id]
        for v in cacc.dep_reads | cacc.need_old:
            comb_readers.setdefault(v, []).append(c)
    # A derived clock (combinational logic of inputs, such as a gated clock)
    # is computed before the clocked blocks it triggers, so their edge test
    # sees its new value; its logic may not read clocked state (a clock that
    # state gates could glitch within the edge's time step).
    derived = {c for c in design.clocks if c.kind != "input"}
    cone: Set[int] = set()
    work = [w.id for c in derived for w in comb_writers.get(c, [])]
    while work:
        i = work.pop()
        if i in cone:
            continue
        cone.add(i)
        for v in sch.access[i].dep_reads | sch.access[i].need_old:
            if sch.roles.get(v) == "state":
                raise CompileError(f"derived clock logic reads clocked state {v.name} "
                                   "(gated by a register; not supported)")
            work += [w.id for w in comb_writers.get(v, [])]
    if any(len(comp) > 1 and cone & set(comp) for comp in comb_sccs):
        raise CompileError("derived clock logic in a combinational cycle")
    for s in seq:
        sacc = sch.access[s.id]
        eclk = {clk for _, clk in s.edges if clk in derived}
        for v in sacc.writes:
            for c in comb_readers.get(v, ()):
                add_edge(proc2vx[s.id], proc2vx[c.id])
        for v in (sacc.dep_reads | sacc.need_old) - eclk:
            for w in comb_writers.get(v, []):
                add_edge(proc2vx[s.id], proc2vx[w.id])
        for v in eclk:
            for w in comb_writers.get(v, []):
                add_edge(proc2vx[w.id], proc2vx[s.id])

    sch.vertices = vertices
    sch.change_vars = sorted(change, key=lambda v: v.id)
    sch.topo = _toposort(vertices)
    return sch


def _tarjan(nodes: List[int], adj: Dict[int, Set[int]]) -> List[List[int]]:
    """Strongly connected components (iterative Tarjan), in reverse topological
    order of the condensation; callers re-sort as needed."""
    index: Dict[int, int] = {}
    low: Dict[int, int] = {}
    onstack: Set[int] = set()
    stack: List[int] = []
    out: List[List[int]] = []
    counter = 0
    for root in nodes:
        if root in index:
            continue
        work = [(root, iter(sorted(adj.get(root, ()))))]
        index[root] = low[root] = counter
        counter += 1
        stack.append(root)
        onstack.add(root)
        while work:
            v, it = work[-1]
            advanced = False
            for w in it:
                if w not in index:
                    index[w] = low[w] = counter
                    counter += 1
                    stack.append(w)
                    onstack.add(w)
                    work.append((w, iter(sorted(adj.get(w, ())))))
                    advanced = True
                    break
                elif w in onstack:
                    low[v] = min(low[v], index[w])
            if advanced:
                continue
            work.pop()
            if work:
                u = work[-1][0]
                low[u] = min(low[u], low[v])
            if low[v] == index[v]:
                comp = []
                while True:
                    w = stack.pop()
                    onstack.discard(w)
                    comp.append(w)
                    if w == v:
                        break
                out.append(comp)
    return out


def _toposort(vertices: List[Vertex]) -> List[int]:
    import heapq
    indeg = {v.id: len(v.preds) for v in vertices}
    heap = [v.id for v in vertices if indeg[v.id] == 0]
    heapq.heapify(heap)
    out = []
    while heap:
        i = heapq.heappop(heap)
        out.append(i)
        for j in vertices[i].succs:
            indeg[j] -= 1
            if indeg[j] == 0:
                heapq.heappush(heap, j)
    if len(out) != len(vertices):
        raise CompileError("internal error: eval graph is cyclic after loop breaking")
    return out
