"""Coarsening of the eval DAG into mtasks (the V3Partition role).

``partition()`` restates V3Partition::go (src/V3Partition.cpp:2353-2479):

1. one LogicMTask per logic vertex, cost from the cost model (V3InstrCount,
   or the learned MLP) -- src/V3Partition.cpp:449-460;
2. critical path cost into / out of every mtask;
3. ``cpLimit = total * 3 / (5 * threads)`` -- src/V3Partition.cpp:2418-2425;
4. repeatedly merge the best-scoring candidate (an edge, or two siblings that
   share a parent or child) whose merged local critical path stays within the
   limit; edge score = new local critical path + 1, sibling score = new local
   critical path, so siblings win ties (src/V3Partition.cpp:1402-1438 and the
   rationale of selfTestX :1575-1605); a merge that would create a cycle is
   skipped;
5. transitive edges are removed.

The scoreboard here is a lazy min-heap: a popped candidate is rescored and
re-pushed if its score grew, which reproduces the reference's "rescore when the
score exceeds the limit before rescore" behaviour
(src/V3Partition.cpp:1132-1207) without its bookkeeping.

``stepCost`` rounds costs up on a log scale with 20 steps per e-fold (about
5%), as the reference does to make scores insensitive to small cost
differences (src/V3Partition.cpp:486-505). When no candidate scores within the
limit but more than 50 mtasks per thread remain, the limit is raised by 20%
and merging continues (src/V3Partition.cpp:1132-1158).
"""
from __future__ import annotations

import heapq
import math
from typing import Callable, Dict, List, Optional, Set, Tuple


def step_cost(cost: int) -> int:
    """Round up to the next of 20 log-steps per e-fold (about 5%)."""
    if cost <= 0:
        return 0
    logcost = math.log(cost)
    logcost = math.ceil(logcost * 20.0) / 20.0
    s = int(math.exp(logcost))
    if s < cost:
        s = cost
    return s


class MTask:
    __slots__ = ("id", "vids", "cost", "sc", "preds", "succs", "cp_fwd", "cp_rev", "alive", "gen")

    def __init__(self, mid: int, vids: List[int], cost: int):
        self.id = mid
        self.vids = vids
        self.cost = cost
        self.sc = step_cost(cost)
        self.preds: Set[int] = set()
        self.succs: Set[int] = set()
        self.cp_fwd = 0   # longest path cost of predecessors (excluding self)
        self.cp_rev = 0   # longest path cost of successors (excluding self)
        self.alive = True
        self.gen = 0

    def __repr__(self):
        return f"MTask{self.id}(cost={self.cost}, n={len(self.vids)})"


class Graph:
    def __init__(self, costs: List[int], edges: List[Tuple[int, int]]):
        self.m: Dict[int, MTask] = {}
        for i, c in enumerate(costs):
            self.m[i] = MTask(i, [i], c)
        for a, b in edges:
            if a != b:
                self.m[a].succs.add(b)
                self.m[b].preds.add(a)

    def alive(self) -> List[MTask]:
        return [t for t in self.m.values() if t.alive]

    # ------------------------------------------------------ critical paths
    def topo(self) -> List[int]:
        indeg = {i: len(t.preds) for i, t in self.m.items() if t.alive}
        ready = sorted(i for i, d in indeg.items() if d == 0)
        heapq.heapify(ready)
        out = []
        while ready:
            i = heapq.heappop(ready)
            out.append(i)
            for j in self.m[i].succs:
                indeg[j] -= 1
                if indeg[j] == 0:
                    heapq.heappush(ready, j)
        if len(out) != len(indeg):
            raise RuntimeError("mtask graph has a cycle")
        return out

    def init_cp(self):
        order = self.topo()
        for i in order:
            t = self.m[i]
            t.cp_fwd = max((self.m[p].cp_fwd + self.m[p].sc for p in t.preds), default=0)
        for i in reversed(order):
            t = self.m[i]
            t.cp_rev = max((self.m[s].cp_rev + self.m[s].sc for s in t.succs), default=0)

    # ---------------------------------------------------------- path check
    def has_path(self, a: int, b: int, skip_direct: bool) -> bool:
        """Is there a path a -> ... -> b (ignoring the direct edge a->b when
        skip_direct)? Bounded by critical-path order: only nodes whose
        forward cp lies between a's and b's can be on the path."""
        limit = self.m[b].cp_fwd
        stack = [s for s in self.m[a].succs if not (skip_direct and s == b)]
        seen = set()
        while stack:
            x = stack.pop()
            if x == b:
                return True
            if x in seen:
                continue
            seen.add(x)
            tx = self.m[x]
            if tx.cp_fwd >= limit:
                continue
            stack.extend(tx.succs)
        return False

    # --------------------------------------------------------------- merge
    def merge(self, keep: int, gone: int, cost: Optional[int] = None):
        k = self.m[keep]
        g = self.m[gone]
        k.vids.extend(g.vids)
        k.cost = k.cost + g.cost if cost is None else cost
        k.sc = step_cost(k.cost)
        for p in g.preds:
            if p == keep:
                continue
            pm = self.m[p]
            pm.succs.discard(gone)
            pm.succs.add(keep)
            k.preds.add(p)
        for s in g.succs:
            if s == keep:
                continue
            sm = self.m[s]
            sm.preds.discard(gone)
            sm.preds.add(keep)
            k.succs.add(s)
        k.preds.discard(gone)
        k.succs.discard(gone)
        g.alive = False
        g.preds = set()
        g.succs = set()
        k.gen += 1

    def update_cp_local(self, start: int):
        """Propagate critical-path changes from a merged node."""
        # forward
        stack = [start]
        while stack:
            i = stack.pop()
            t = self.m[i]
            m = self.m
            nf = max((m[p].cp_fwd + m[p].sc for p in t.preds), default=0)
            if nf != t.cp_fwd or i == start:
                t.cp_fwd = nf
                stack.extend(t.succs)
        stack = [start]
        while stack:
            i = stack.pop()
            t = self.m[i]
            m = self.m
            nr = max((m[s].cp_rev + m[s].sc for s in t.succs), default=0)
            if nr != t.cp_rev or i == start:
                t.cp_rev = nr
                stack.extend(t.preds)

    def remove_transitive(self):
        order = self.topo()
        pos = {i: k for k, i in enumerate(order)}
        for i in order:
            t = self.m[i]
            succs = sorted(t.succs, key=lambda x: pos[x])
            for s in succs:
                # s redundant if reachable from another successor
                for o in succs:
                    if o != s and pos[o] < pos[s] and s in self._reach(o, pos[s], pos):
                        t.succs.discard(s)
                        self.m[s].preds.discard(i)
                        break

    def _reach(self, a: int, maxpos: int, pos) -> Set[int]:
        seen = set()
        stack = [a]
        while stack:
            x = stack.pop()
            for y in self.m[x].succs:
                if y not in seen and pos[y] <= maxpos:
                    seen.add(y)
                    stack.append(y)
        return seen


def _fwd_without(g: Graph, tb: MTask, ta: MTask) -> int:
    """tb's forward critical path ignoring the edge from ta
    (critPathCostWithout, src/V3Partition.cpp:1402-1438)."""
    if ta.cp_fwd + ta.sc < tb.cp_fwd:
        return tb.cp_fwd
    m = g.m
    a = ta.id
    return max((m[p].cp_fwd + m[p].sc for p in tb.preds if p != a), default=0)


def _rev_without(g: Graph, ta: MTask, tb: MTask) -> int:
    if tb.cp_rev + tb.sc < ta.cp_rev:
        return ta.cp_rev
    m = g.m
    b = tb.id
    return max((m[s].cp_rev + m[s].sc for s in ta.succs if s != b), default=0)


def _union_cost(g: Graph, ta: MTask, tb: MTask, merged) -> int:
    """Cost of the mtask a merge of ta and tb would make: the sum of their
    costs (additive instruction counts), or the merge model's cost of the
    union when one is given."""
    return ta.cost + tb.cost if merged is None else merged.cost(ta, tb)


def edge_score(g: Graph, a: int, b: int, merged=None) -> int:
    """Local critical path through the merged node if edge a->b merges."""
    ta, tb = g.m[a], g.m[b]
    fwd = max(ta.cp_fwd, _fwd_without(g, tb, ta))
    rev = max(tb.cp_rev, _rev_without(g, ta, tb))
    return fwd + rev + step_cost(_union_cost(g, ta, tb, merged))


def sibling_score(g: Graph, a: int, b: int, merged=None) -> int:
    ta, tb = g.m[a], g.m[b]
    return (max(ta.cp_fwd, tb.cp_fwd) + max(ta.cp_rev, tb.cp_rev)
            + step_cost(_union_cost(g, ta, tb, merged)))


SIBLING_EDGE_LIMIT = 25


MAX_MTASKS_PER_THREAD = 50


def contract(g: Graph, cp_limit: int, max_mtasks: int = 1 << 30, merged=None):
    """PartContraction: merge the lowest-scoring candidates while the merged
    local critical path stays within ``cp_limit``. ``merged``: optional
    merge model (``cost(ta, tb)`` of a union, ``merge(keep, gone)`` told of
    every merge) whose non-additive costs replace the sum in the scores and
    in the merged mtask's cost."""
    alive_count = len(g.alive())
    g.init_cp()
    heap: List[Tuple[int, int, int, int, int, int]] = []  # score, kind, a, b, gen_a, gen_b
    seq = 0

    def push_edge(a, b):
        nonlocal seq
        heapq.heappush(heap, (edge_score(g, a, b, merged) + 1, 1, seq, a, b, 0))
        seq += 1

    def push_sib(a, b):
        nonlocal seq
        if a == b:
            return
        if a > b:
            a, b = b, a
        heapq.heappush(heap, (sibling_score(g, a, b, merged), 0, seq, a, b, 0))
        seq += 1

    def sibs_of(t: MTask):
        for rel in (t.preds, t.succs):
            rels = list(rel)[:SIBLING_EDGE_LIMIT]
            for r in rels:
                # a few shortest-cp children/parents of each relative
                kids = g.m[r].succs if rel is t.preds else g.m[r].preds
                kl = list(kids)
                if len(kl) > 3 * SIBLING_EDGE_LIMIT:
                    kl = kl[:3 * SIBLING_EDGE_LIMIT]
                cands = sorted(kl, key=lambda x: (g.m[x].cp_rev + g.m[x].sc, x))[:6]
                for k in cands:
                    if k != t.id:
                        push_sib(t.id, k)

    for t in g.alive():
        for s in t.succs:
            push_edge(t.id, s)
    for t in g.alive():
        ch = sorted(t.succs, key=lambda x: (g.m[x].cp_rev + g.m[x].sc, x))
        for i in range(0, min(len(ch), 2 * SIBLING_EDGE_LIMIT) - 1, 2):
            push_sib(ch[i], ch[i + 1])
        pa = sorted(t.preds, key=lambda x: (g.m[x].cp_fwd + g.m[x].sc, x))
        for i in range(0, min(len(pa), 2 * SIBLING_EDGE_LIMIT) - 1, 2):
            push_sib(pa[i], pa[i + 1])

    while heap:
        score, kind, _, a, b, _ = heapq.heappop(heap)
        ta, tb = g.m[a], g.m[b]
        if not (ta.alive and tb.alive):
            continue
        if kind == 1:
            if b not in ta.succs:
                continue
            actual = edge_score(g, a, b, merged) + 1
        else:
            actual = sibling_score(g, a, b, merged)
        if actual > score:
            heapq.heappush(heap, (actual, kind, seq, a, b, 0))
            seq += 1
            continue
        if actual > cp_limit:
            if alive_count > max_mtasks:
                cp_limit = (cp_limit * 120) // 100 + 1
                heapq.heappush(heap, (actual, kind, seq, a, b, 0))
                seq += 1
                continue
            break
        # cycle check: an edge merge needs no other path a->..->b; a sibling
        # merge needs no path between them in either direction
        if kind == 1:
            if g.has_path(a, b, skip_direct=True):
                continue
        else:
            if b in ta.succs or a in tb.succs:
                continue
            if ta.cp_fwd <= tb.cp_fwd:
                if g.has_path(a, b, skip_direct=False):
                    continue
            else:
                if g.has_path(b, a, skip_direct=False):
                    continue
            if ta.cp_fwd > tb.cp_fwd and g.has_path(a, b, skip_direct=False):
                continue
            if tb.cp_fwd > ta.cp_fwd and g.has_path(b, a, skip_direct=False):
                continue
        keep, gone = (a, b) if (len(ta.preds) + len(ta.succs)) >= (len(tb.preds) + len(tb.succs)) else (b, a)
        if merged is None:
            g.merge(keep, gone)
        else:
            c = merged.cost(g.m[keep], g.m[gone])
            merged.merge(keep, gone)
            g.merge(keep, gone, cost=c)
        alive_count -= 1
        g.update_cp_local(keep)
        tk = g.m[keep]
        for s in tk.succs:
            push_edge(keep, s)
        for p in tk.preds:
            push_edge(p, keep)
        sibs_of(tk)


def partition(costs: List[int], edges: List[Tuple[int, int]], threads: int,
              order: Optional[List[int]] = None, merged=None) -> List[MTask]:
    """Coarsen a vertex DAG. Returns mtasks in a topological order; each
    mtask's ``vids`` are sorted by ``order`` (the vertices' topological
    order) so its body is a valid sequential schedule. ``merged``: optional
    merge model for the contraction (see ``contract``); ``costs`` are then
    its costs of the single vertices."""
    g = Graph(costs, edges)
    total = sum(costs)
    threads = max(threads, 1)
    cp_limit = (total * 3) // (threads * 5)
    if threads <= 1:
        cp_limit = 1 << 62
    contract(g, cp_limit, MAX_MTASKS_PER_THREAD * threads if threads > 1 else 500, merged)
    g.remove_transitive()
    pos = {v: k for k, v in enumerate(order)} if order is not None else None
    out = []
    for i in g.topo():
        t = g.m[i]
        if pos is not None:
            t.vids.sort(key=lambda v: pos[v])
        out.append(t)
    return out


def partition_stats(mtasks: List[MTask]) -> Dict[str, int]:
    """Critical path and total cost (PartParallelismEst's report)."""
    idx = {t.id: t for t in mtasks}
    cp: Dict[int, int] = {}
    for t in mtasks:
        cp[t.id] = t.cost + max((cp[p] for p in t.preds), default=0)
    edges = sum(len(t.succs) for t in mtasks)
    return {"vertices": len(mtasks), "edges": edges, "total": sum(t.cost for t in mtasks),
            "critical_path": max(cp.values(), default=0)}
