"""Gate-level netlist engine, host side: the structural netlist of a seeded
random design (the gate table the netlist generator writes beside its
Verilog) lowered to the program blob of include/rf_gatenet.h.

This is the path for netlists beyond the per-design emitter's size
(configuration 5: 500,000 gates, 50,000 flip-flops, eight 1,024 x 64-bit
RAMs): the emitter would write hundreds of megabytes of per-lane code, and
the bit-plane LDS form holds at most 160 KB of planes per workgroup. Here
the same single-bit netlist is data: levels of gate records for the
bit-plane form, and mtasks of gate records for the per-lane graph form.

The mtask partition restates the reference's flow at a coarser grain: the
gates of each level are cut into chunks (the logic vertices), a chunk
depends on the chunks that produce its fan-ins, each chunk costs the
V3InstrCount estimate of its gates (src/V3InstrCount.cpp: an AND / OR / XOR
assign 6, NAND 7, a 2:1 mux 9 -- one per operator, two per variable load, one
per store, four per branch), and ``partition.partition`` contracts the chunk
DAG to the critical-path limit of ``--threads`` (src/V3Partition.cpp:2418-2425).
The learned schedule chooses ``threads`` by the MLP's predicted eval time,
as for the emitted designs.
"""
from __future__ import annotations

import struct
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from .designs.netlist import MUX, NAND, GateTable

MAGIC = 0x54474652          # 'RFGT'
VERSION = 1
HDR_WORDS = 32
GATE_COST = np.array([6, 6, 6, 7, 9], np.int64)   # AND OR XOR NAND MUX
CHUNK = 1024


def parse_table(buf: bytes) -> GateTable:
    """The generator's binary gate table (GateTable.to_bytes)."""
    if buf[:4] != b"RFNL":
        raise ValueError("not a gate table")
    h = struct.unpack_from("<9I", buf, 4)
    n_pi, n_ff, n_ram_bits, n_gates, n_po, n_rams, depth, width, abits = h
    a = np.frombuffer(buf, "<u4", offset=40)
    p = 0
    gt = a[p:p + n_gates].astype(np.uint8)
    p += n_gates
    gi = a[p:p + 3 * n_gates].reshape(n_gates, 3).astype(np.uint32)
    p += 3 * n_gates
    ffd = a[p:p + n_ff].astype(np.uint32)
    p += n_ff
    po = a[p:p + n_po].astype(np.uint32)
    p += n_po
    addr = np.zeros((n_rams, abits), np.uint32)
    we = np.zeros(n_rams, np.uint32)
    wd = np.zeros((n_rams, width), np.uint32)
    for r in range(n_rams):
        addr[r] = a[p:p + abits]
        we[r] = a[p + abits]
        wd[r] = a[p + abits + 1:p + abits + 1 + width]
        p += abits + 1 + width
    return GateTable(n_pi, n_ff, n_ram_bits, gt, gi, ffd, po, depth, width, addr, we, wd)


def gate_levels(t: GateTable) -> np.ndarray:
    """Level of every gate: 1 + the deepest fan-in (inputs, flip-flops and
    RAM read data are level 0). The generator writes gates after all their
    fan-ins, so a fixpoint over whole-array passes settles in depth passes."""
    first = t.first_gate
    n = t.n_gates
    lvl = np.zeros(first + n, np.int32)
    ins = t.gate_in.astype(np.int64)
    mux = t.gate_type == MUX
    while True:
        m = np.maximum(lvl[ins[:, 0]], lvl[ins[:, 1]])
        m = np.where(mux, np.maximum(m, lvl[ins[:, 2]]), m) + 1
        if np.array_equal(m, lvl[first:]):
            return m
        lvl[first:] = m


def output_cone(t: GateTable, lvl: np.ndarray) -> np.ndarray:
    """Gates the primary outputs depend on (through gates only); ``lvl``:
    the gates' levels."""
    first = t.first_gate
    mark = np.zeros(first + t.n_gates, bool)
    mark[t.po] = True
    ins = t.gate_in.astype(np.int64)
    mux = t.gate_type == MUX
    for level in range(int(lvl.max()) if len(lvl) else 0, 0, -1):
        g = np.nonzero((lvl == level) & mark[first:])[0]
        if not len(g):
            continue
        mark[ins[g, 0]] = True
        mark[ins[g, 1]] = True
        gm = g[mux[g]]
        mark[ins[gm, 2]] = True
    return mark[first:]


def records(t: GateTable, gates: np.ndarray) -> np.ndarray:
    """uint32 [n, 4] records {dst | type << 29, in0, in1, in2}."""
    first = t.first_gate
    ty = t.gate_type[gates].astype(np.uint32)
    rec = np.zeros((len(gates), 4), np.uint32)
    rec[:, 0] = (gates.astype(np.uint32) + first) | (ty << 29)
    rec[:, 1:] = t.gate_in[gates]
    mux = ty == MUX
    rec[~mux, 3] = rec[~mux, 1]      # unused third operand: any valid node
    return rec


def _by_level(lvl_of: np.ndarray, gates: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
    """Gates sorted by level (stable) and the level starts [nlev + 1]."""
    order = gates[np.argsort(lvl_of[gates], kind="stable")]
    levels = lvl_of[order]
    nlev = int(levels.max()) if len(levels) else 0
    starts = np.searchsorted(levels, np.arange(1, nlev + 2)).astype(np.uint32)
    return order, starts


# ------------------------------------------------------------------ partition
class ChunkGraph:
    """Gates of each level cut into chunks of at most ``chunk`` gates; a
    chunk depends on the chunks that produce its fan-ins."""

    def __init__(self, t: GateTable, lvl: np.ndarray, chunk: int = CHUNK):
        first = t.first_gate
        order, starts = _by_level(lvl, np.arange(t.n_gates))
        cid = np.zeros(t.n_gates, np.int64)
        chunks: List[np.ndarray] = []
        for level in range(len(starts) - 1):
            gs = order[starts[level]:starts[level + 1]]
            for k in range(0, len(gs), chunk):
                cid[gs[k:k + chunk]] = len(chunks)
                chunks.append(gs[k:k + chunk])
        self.chunks = chunks
        self.cid = cid
        node_c = np.full(first + t.n_gates, -1, np.int64)
        node_c[first:] = cid
        ins = t.gate_in.astype(np.int64)
        mux = t.gate_type == MUX
        src = [node_c[ins[:, 0]], node_c[ins[:, 1]], np.where(mux, node_c[ins[:, 2]], -1)]
        edges = set()
        for s in src:
            ok = s >= 0
            pairs = np.unique(np.stack([s[ok], cid[ok]], 1), axis=0)
            edges.update((int(a), int(b)) for a, b in pairs if a != b)
        self.edges = sorted(edges)
        self.cost = [int(GATE_COST[t.gate_type[c]].sum()) for c in chunks]
        self.t = t

    def features(self, members: Sequence[int]) -> List[float]:
        """The cost model's six features of an mtask made of chunks
        ``members`` (rtlflow_amd.costmodel.FEATURE_NAMES)."""
        t = self.t
        first = t.first_gate
        gs = np.concatenate([self.chunks[c] for c in members])
        ins = t.gate_in[gs].astype(np.int64)
        mux = t.gate_type[gs] == MUX
        read = np.unique(np.concatenate([ins[:, 0], ins[:, 1], ins[mux, 2]]))
        written = gs.astype(np.int64) + first
        inside = np.isin(read, written)
        fan_in = int((~inside).sum())
        # written nodes read by a gate outside the mtask, or by the edge / outputs
        outside = np.ones(t.n_gates, bool)
        outside[gs] = False
        ext_ins = t.gate_in[outside].astype(np.int64)
        ext_mux = t.gate_type[outside] == MUX
        ext_read = np.concatenate([ext_ins[:, 0], ext_ins[:, 1], ext_ins[ext_mux, 2], t.ff_d.astype(np.int64),
                                   t.po.astype(np.int64), t.ram_addr.reshape(-1).astype(np.int64),
                                   t.ram_we.astype(np.int64), t.ram_wdata.reshape(-1).astype(np.int64)])
        fan_out = int(np.isin(written, ext_read).sum())
        instr = float(GATE_COST[t.gate_type[gs]].sum())
        state = float(len(np.union1d(read, written)))
        return [instr, float(fan_in), float(fan_out), state, float(mux.sum()), 0.0]


def partition_mtasks(cg: ChunkGraph, threads: int) -> Tuple[List[List[int]], List[List[int]]]:
    """Contract the chunk DAG (V3Partition restatement) -> (mtasks as chunk
    lists, predecessors), mtasks in a topological order."""
    from .partition import partition
    mts = partition(cg.cost, cg.edges, threads, order=list(range(len(cg.chunks))))
    ids = {m.id: k for k, m in enumerate(mts)}
    members = [sorted(m.vids) for m in mts]
    preds = [sorted(ids[p] for p in m.preds) for m in mts]
    # topological order of the mtasks (the partitioner's is by id)
    n = len(mts)
    indeg = [len(p) for p in preds]
    succ: List[List[int]] = [[] for _ in range(n)]
    for k, ps in enumerate(preds):
        for p in ps:
            succ[p].append(k)
    ready = [k for k in range(n) if indeg[k] == 0]
    order: List[int] = []
    while ready:
        k = ready.pop(0)
        order.append(k)
        for s in succ[k]:
            indeg[s] -= 1
            if indeg[s] == 0:
                ready.append(s)
    if len(order) != n:
        raise RuntimeError("mtask graph has a cycle")
    pos = {k: i for i, k in enumerate(order)}
    return [members[k] for k in order], [sorted(pos[p] for p in preds[k]) for k in order]


# --------------------------------------------------------------------- blob
def program(t: GateTable, threads: Optional[int] = 4, chunk: int = CHUNK) -> np.ndarray:
    """The program blob (uint32) of include/rf_gatenet.h; ``threads`` None:
    no mtask partition (bit-plane form only)."""
    first = t.first_gate
    n_nodes = first + t.n_gates
    lvl = gate_levels(t)
    full, lev_full = _by_level(lvl, np.arange(t.n_gates))
    cone_mask = output_cone(t, lvl)
    cone, lev_cone = _by_level(lvl, np.nonzero(cone_mask)[0])
    rec_full = records(t, full)
    rec_cone = records(t, cone)
    if threads:
        cg = ChunkGraph(t, lvl, chunk)
        members, preds = partition_mtasks(cg, threads)
        mt_gates = []
        for mem in members:
            gs = np.concatenate([cg.chunks[c] for c in mem])
            mt_gates.append(gs[np.argsort(lvl[gs], kind="stable")])
    else:
        members, preds, mt_gates = [], [], []
    mt_tab = np.zeros((len(members), 4), np.uint32)
    pred_list: List[int] = []
    off = 0
    for k, gs in enumerate(mt_gates):
        mt_tab[k] = (off, len(gs), len(preds[k]), len(pred_list))
        pred_list += preds[k]
        off += len(gs)
    rec_mt = records(t, np.concatenate(mt_gates)) if mt_gates else np.zeros((0, 4), np.uint32)
    n_rams = len(t.ram_we)
    abits = t.ram_addr.shape[1] if n_rams else 0
    ram = np.concatenate([np.concatenate([t.ram_addr[r], [t.ram_we[r]], t.ram_wdata[r]])
                          for r in range(n_rams)]).astype(np.uint32) if n_rams else np.zeros(0, np.uint32)
    parts: Dict[str, np.ndarray] = {}
    words = [np.zeros(HDR_WORDS, np.uint32)]
    pos = HDR_WORDS

    def put(name, arr, align=1):
        nonlocal pos
        arr = np.asarray(arr, np.uint32).reshape(-1)
        pad = (-pos) % align
        if pad:
            words.append(np.zeros(pad, np.uint32))
            pos += pad
        parts[name] = pos
        words.append(arr)
        pos += len(arr)

    put("lev_full", lev_full)
    put("lev_cone", lev_cone)
    put("ff", t.ff_d)
    put("po", t.po)
    put("ram", ram)
    put("rec_full", rec_full, 4)
    put("rec_cone", rec_cone, 4)
    put("mt", mt_tab)
    put("preds", np.asarray(pred_list, np.uint32))
    put("rec_mt", rec_mt, 4)
    blob = np.concatenate(words)
    hdr = [MAGIC, VERSION, n_nodes, t.n_pi, t.n_ff, t.n_ram_bits, len(t.po), n_rams, t.ram_depth if n_rams else 0,
           t.ram_width if n_rams else 0, abits, len(lev_full) - 1, len(lev_cone) - 1, len(rec_full), len(rec_cone),
           len(members), parts["lev_full"], parts["lev_cone"], parts["ff"], parts["po"], parts["ram"],
           parts["rec_full"], parts["rec_cone"], parts["mt"], parts["preds"], parts["rec_mt"], len(rec_mt),
           len(blob)]
    blob[:len(hdr)] = np.asarray(hdr, np.uint32)
    return blob


def candidates(t: GateTable, threads_list=(2, 4, 8, 16, 32), chunk: int = CHUNK) -> Dict[str, list]:
    """Per candidate granularity, the cost-model features of its mtasks
    (for the MLP's choice of the partition, costmodel.choose_partitions)."""
    lvl = gate_levels(t)
    cg = ChunkGraph(t, lvl, chunk)
    out = {}
    for th in threads_list:
        members, _ = partition_mtasks(cg, th)
        out[str(th)] = [cg.features(m) for m in members]
    return out


# ------------------------------------------------------------ the C ABI side
class GateNet:
    """A loaded program on the current GPU (ctypes over rf_gatenet.h)."""

    GRAPH = 0
    BITPLANE = 1

    def __init__(self, prog: np.ndarray, n_lanes: int):
        import ctypes as C
        from .model import RFError, runtime
        self._C = C
        self._err = RFError
        rt = runtime()
        P, u32, u64 = C.c_void_p, C.c_uint32, C.c_uint64
        rt.rf_gn_create.argtypes = [P, C.c_size_t, u32, C.POINTER(P)]
        rt.rf_gn_run.argtypes = [P, u32, u64, P, P, C.POINTER(C.c_float)]
        rt.rf_gn_destroy.argtypes = [P]
        rt.rf_gn_mtasks.argtypes = [P]
        rt.rf_gn_mtasks.restype = u32
        rt.rf_gn_nodes_per_cycle.argtypes = [P, u32]
        rt.rf_gn_nodes_per_cycle.restype = u32
        self.rt = rt
        self.prog = np.ascontiguousarray(prog, np.uint32)
        self.n = int(n_lanes)
        h = C.c_void_p()
        self._ck(rt.rf_gn_create(self.prog.ctypes.data, len(self.prog), self.n, C.byref(h)))
        self.h = h

    def _ck(self, rc):
        if rc:
            raise self._err(f"rf error {rc}: {self.rt.rf_last_error().decode()}")

    def run(self, mode: int, seed: int, budgets) -> Tuple[np.ndarray, float]:
        C = self._C
        b = np.ascontiguousarray(budgets, np.uint32)
        if len(b) != self.n:
            raise ValueError(f"{len(b)} budgets for {self.n} lanes")
        out = np.zeros(self.n, np.uint64)
        ms = C.c_float(0)
        self._ck(self.rt.rf_gn_run(self.h, mode, seed & ((1 << 64) - 1), b.ctypes.data, out.ctypes.data,
                                   C.byref(ms)))
        return out, ms.value

    def mtasks(self) -> int:
        return self.rt.rf_gn_mtasks(self.h)

    def nodes_per_cycle(self, mode: int) -> int:
        return self.rt.rf_gn_nodes_per_cycle(self.h, mode)

    def close(self):
        if getattr(self, "h", None):
            self.rt.rf_gn_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
