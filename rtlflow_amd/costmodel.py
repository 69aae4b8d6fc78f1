"""Learned mtask cost model: features, the GPU-trained MLP, and its use as
the partitioner's cost function.

The reference costs every logic vertex with V3InstrCount
(src/V3Partition.cpp:456) and re-costs mtasks after partitioning
(finalizeCosts, src/V3Partition.cpp:2481-2545). Here a 3-layer MLP (6 -> 64 ->
64 -> 1, FP32, hand-written HIP kernels in csrc/rf_mlp.hip) maps six features
of a vertex or an mtask to its measured kernel time:

  0 instruction count (the V3InstrCount estimate)
  1 fan-in   (distinct variables read from outside the unit)
  2 fan-out  (distinct variables written that are read outside the unit)
  3 state bytes touched per lane (stored variables read or written)
  4 branch count (if / case / ?: nodes)
  5 wide-op count (operators on values wider than 64 bits)

Features enter as log1p, the target as log(microseconds), so the network
models multiplicative effects. Training data come from the simulator itself:
every mtask kernel of a set of synthetic designs is timed in isolation on the
GPU (rf_profile_mtasks). The heuristic remains the fallback and the reference
point: a learned schedule is kept only if its measured throughput is higher and its
hashes are identical.
"""
from __future__ import annotations

import ctypes as C
from typing import Dict, List, Optional, Sequence, Set

import numpy as np

from .instrcount import process_cost
from .ir import Case, Cond, If, Op, stmt_exprs, walk_expr, walk_stmts
from .layout import class_of

N_FEATURES = 6
FEATURE_NAMES = ["instr", "fan_in", "fan_out", "state_bytes", "branches", "wide_ops"]
CLS_BYTES = {"C": 1, "S": 2, "I": 4, "Q": 8, "W": 4}


def _proc_counts(p) -> Dict[str, int]:
    br = 0
    wide = 0
    for s in walk_stmts(p.stmts):
        if isinstance(s, (If, Case)):
            br += 1
        for e in stmt_exprs(s):
            for x in walk_expr(e):
                if isinstance(x, Cond):
                    br += 1
                elif isinstance(x, Op) and x.width > 64:
                    wide += 1
    return {"branches": br, "wide": wide}


class FeatureExtractor:
    """Features of vertices and vertex sets (mtasks) of a compiled design."""

    def __init__(self, comp):
        self.comp = comp
        sch = comp.sch
        self.readers: Dict[object, Set[int]] = {}
        for vx in sch.vertices:
            for v in vx.reads | vx.need_old:
                self.readers.setdefault(v, set()).add(vx.id)
        self.counts = {}
        for vx in sch.vertices:
            br = wide = 0
            for p in vx.procs:
                c = _proc_counts(p)
                br += c["branches"]
                wide += c["wide"]
            self.counts[vx.id] = (br, wide)

    def raw(self, vids: Sequence[int]) -> List[float]:
        comp = self.comp
        sch = comp.sch
        inside = set(vids)
        reads: Set = set()
        writes: Set = set()
        instr = br = wide = 0
        for vid in vids:
            vx = sch.vertices[vid]
            reads |= vx.reads | vx.need_old
            writes |= vx.writes
            instr += sum(process_cost(p) for p in vx.procs)
            b, w = self.counts[vid]
            br += b
            wide += w
        fan_in = len([v for v in reads if v not in writes])
        fan_out = len([v for v in writes if self.readers.get(v, set()) - inside])
        state = 0
        for v in reads | writes:
            if v in comp.stored and not v.is_mem:
                state += v.slots * CLS_BYTES[class_of(v)]
            elif v.is_mem:
                state += CLS_BYTES[class_of(v)] * v.words
        return [float(instr), float(fan_in), float(fan_out), float(state), float(br), float(wide)]

    def vertex(self, vid: int) -> List[float]:
        return self.raw([vid])

    def mtask(self, k: int) -> List[float]:
        return self.raw(self.comp.mtasks[k].vids)


def encode(raw: Sequence[Sequence[float]]) -> np.ndarray:
    return np.log1p(np.asarray(raw, dtype=np.float64)).astype(np.float32)


# ---------------------------------------------------------------- the MLP
_lib = None


def mlp_lib() -> C.CDLL:
    global _lib
    if _lib is None:
        from .model import runtime
        lib = runtime()
        P, u32, u64, f32 = C.c_void_p, C.c_uint32, C.c_uint64, C.c_float
        lib.rf_mlp_param_count.restype = u32
        lib.rf_mlp_param_count.argtypes = [u32]
        lib.rf_mlp_create.restype = C.c_int
        lib.rf_mlp_create.argtypes = [u32, P, u64, C.POINTER(P)]
        lib.rf_mlp_destroy.argtypes = [P]
        lib.rf_mlp_get_params.argtypes = [P, P]
        lib.rf_mlp_set_params.argtypes = [P, P]
        lib.rf_mlp_forward.argtypes = [P, P, u32, P]
        lib.rf_mlp_gradients.argtypes = [P, P, P, u32, P, P]
        lib.rf_mlp_train_step.argtypes = [P, P, P, u32, f32, P]
        lib.rf_mlp_last_error.restype = C.c_char_p
        _lib = lib
    return _lib


def _ck(rc):
    if rc != 0:
        raise RuntimeError("rf_mlp: " + mlp_lib().rf_mlp_last_error().decode())


class MLP:
    """Handle to the GPU MLP (rf_mlp.h)."""

    def __init__(self, params: Optional[np.ndarray] = None, seed: int = 1, n_features: int = N_FEATURES):
        lib = mlp_lib()
        self.F = n_features
        self.nparams = lib.rf_mlp_param_count(n_features)
        h = C.c_void_p()
        p = None
        if params is not None:
            params = np.ascontiguousarray(params, np.float32)
            assert params.size == self.nparams
            p = params.ctypes.data
        _ck(lib.rf_mlp_create(n_features, p, seed, C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            mlp_lib().rf_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self) -> np.ndarray:
        out = np.zeros(self.nparams, np.float32)
        _ck(mlp_lib().rf_mlp_get_params(self.h, out.ctypes.data))
        return out

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, np.float32)
        _ck(mlp_lib().rf_mlp_set_params(self.h, p.ctypes.data))

    def forward(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32)
        y = np.zeros(x.shape[0], np.float32)
        _ck(mlp_lib().rf_mlp_forward(self.h, x.ctypes.data, x.shape[0], y.ctypes.data))
        return y

    def gradients(self, x: np.ndarray, t: np.ndarray):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        g = np.zeros(self.nparams, np.float32)
        loss = C.c_float(0)
        _ck(mlp_lib().rf_mlp_gradients(self.h, x.ctypes.data, t.ctypes.data, x.shape[0],
                                       C.byref(loss), g.ctypes.data))
        return loss.value, g

    def train_step(self, x: np.ndarray, t: np.ndarray, lr: float) -> float:
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        loss = C.c_float(0)
        _ck(mlp_lib().rf_mlp_train_step(self.h, x.ctypes.data, t.ctypes.data, x.shape[0], lr,
                                        C.byref(loss)))
        return loss.value


def train(x: np.ndarray, y: np.ndarray, epochs: int = 400, lr: float = 3e-3, batch: int = 256,
          seed: int = 1, val_frac: float = 0.2) -> Dict:
    """Train on the GPU; returns params, normalisation and train/val loss."""
    x = np.asarray(x, np.float32)
    y = np.asarray(y, np.float32)
    rng = np.random.default_rng(seed)
    idx = rng.permutation(len(x))
    nval = max(1, int(len(x) * val_frac))
    va, tr = idx[:nval], idx[nval:]
    mu = x[tr].mean(0)
    sd = x[tr].std(0) + 1e-6
    ym, ys = float(y[tr].mean()), float(y[tr].std() + 1e-6)
    xn = (x - mu) / sd
    yn = (y - ym) / ys
    net = MLP(seed=seed)
    hist = []
    for ep in range(epochs):
        perm = rng.permutation(tr)
        for b in range(0, len(perm), batch):
            sel = perm[b:b + batch]
            net.train_step(xn[sel], yn[sel], lr)
        if ep % 20 == 0 or ep == epochs - 1:
            ptr = net.forward(xn[tr]) * ys + ym
            pva = net.forward(xn[va]) * ys + ym
            hist.append((ep, float(np.mean((ptr - y[tr]) ** 2)), float(np.mean((pva - y[va]) ** 2))))
    model = {"params": net.params().tolist(), "mu": mu.tolist(), "sd": sd.tolist(), "ym": ym, "ys": ys,
             "history": hist, "n_train": int(len(tr)), "n_val": int(nval)}
    net.close()
    return model


class LearnedCost:
    """Partitioner cost function backed by a trained model: predicted kernel
    microseconds per vertex, scaled to the instruction-count range."""

    def __init__(self, model: Dict, scale: float = 1000.0):
        self.model = model
        self.scale = scale
        self.mu = np.asarray(model["mu"], np.float32)
        self.sd = np.asarray(model["sd"], np.float32)
        self.net = MLP(np.asarray(model["params"], np.float32))
        self._cache: Dict[int, Dict[int, int]] = {}

    def predict_us(self, raw: Sequence[Sequence[float]]) -> np.ndarray:
        x = (encode(raw) - self.mu) / self.sd
        y = self.net.forward(x) * self.model["ys"] + self.model["ym"]
        return np.exp(y)

    def __call__(self, comp, vx) -> int:
        key = id(comp)
        if key not in self._cache:
            fx = FeatureExtractor(comp)
            raw = [fx.vertex(v.id) for v in comp.sch.vertices]
            us = self.predict_us(raw)
            self._cache[key] = {v.id: max(1, int(round(u * self.scale))) for v, u in
                                zip(comp.sch.vertices, us)}
        return self._cache[key][vx.id]


def forward_cpu(model: Dict, raw: Sequence[Sequence[float]]) -> np.ndarray:
    """The trained network's prediction (microseconds) on the host: the same
    F -> 64 -> 64 -> 1 ReLU network and parameter layout as the device
    kernels (csrc/rf_mlp.hip: W1 [64][F], b1, W2 [64][64], b2, w3, b3), in
    FP64, for use inside the compiler where no GPU is present."""
    p = np.asarray(model["params"], np.float64)
    F = N_FEATURES
    H = 64
    o = 0
    w1 = p[o:o + H * F].reshape(H, F); o += H * F
    b1 = p[o:o + H]; o += H
    w2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    w3 = p[o:o + H]; o += H
    b3 = p[o]
    x = (np.log1p(np.asarray(raw, np.float64)) - np.asarray(model["mu"], np.float64)) / \
        np.asarray(model["sd"], np.float64)
    h1 = np.maximum(x @ w1.T + b1, 0.0)
    h2 = np.maximum(h1 @ w2.T + b2, 0.0)
    y = (h2 @ w3 + b3) * model["ys"] + model["ym"]
    return np.exp(y)


class LearnedMerge:
    """Merge model for the contraction (``partition.contract``): the cost of
    an mtask is the trained network's predicted kernel time of its vertex
    set, less that of an empty kernel, in units of 1/scale us. The scores of
    the reference's contraction (src/V3Partition.cpp:1402-1438) add the two
    costs of a candidate merge; here the union is costed as a whole, so a
    merge of mtasks that read the same variables (fewer distinct inputs, less
    state touched) scores lower than the sum of its parts, which an additive
    instruction count cannot see.

    Per mtask it keeps the feature aggregates: instruction, branch and
    wide-op counts, and bit sets (Python integers over variable indices) of
    the variables read, written, and written-and-read-outside. A union's
    features follow from those: only variables one side writes and the other
    reads can turn internal, so only their readers are inspected."""

    def __init__(self, model: Dict, comp, scale: float = 100.0):
        self.scale = scale
        p = np.asarray(model["params"], np.float64)
        F, H = N_FEATURES, 64
        o = 0
        self.w1 = p[o:o + H * F].reshape(H, F); o += H * F
        self.b1 = p[o:o + H]; o += H
        self.w2 = p[o:o + H * H].reshape(H, H); o += H * H
        self.b2 = p[o:o + H]; o += H
        self.w3 = p[o:o + H]; o += H
        self.b3 = float(p[o])
        self.mu = np.asarray(model["mu"], np.float64)
        self.sd = np.asarray(model["sd"], np.float64)
        self.ym, self.ys = float(model["ym"]), float(model["ys"])
        fx = FeatureExtractor(comp)
        sch = comp.sch
        stored = self._stored_like(comp)
        idx: Dict[object, int] = {}
        for vx in sch.vertices:
            for v in vx.reads | vx.need_old | vx.writes:
                if v not in idx:
                    idx[v] = len(idx)
        self.nbytes = [0] * len(idx)
        for v, i in idx.items():
            if v in stored and not v.is_mem:
                self.nbytes[i] = v.slots * CLS_BYTES[class_of(v)]
            elif v.is_mem:
                self.nbytes[i] = CLS_BYTES[class_of(v)] * v.words
        # readers of each variable index (vertex ids)
        self.readers: List[List[int]] = [[] for _ in idx]
        for v, rs in fx.readers.items():
            if v in idx:
                self.readers[idx[v]] = sorted(rs)
        self.owner: List[int] = list(range(len(sch.vertices)))
        self.members: Dict[int, List[int]] = {}
        self.R: Dict[int, int] = {}
        self.W: Dict[int, int] = {}
        self.X: Dict[int, int] = {}     # written and read by a vertex outside
        self.cnt: Dict[int, tuple] = {}
        self.state: Dict[int, int] = {}
        for vx in sch.vertices:
            r = w = x = 0
            for v in vx.reads | vx.need_old:
                r |= 1 << idx[v]
            for v in vx.writes:
                w |= 1 << idx[v]
                if any(rd != vx.id for rd in self.readers[idx[v]]):
                    x |= 1 << idx[v]
            self.R[vx.id], self.W[vx.id], self.X[vx.id] = r, w, x
            b, wd = fx.counts[vx.id]
            self.cnt[vx.id] = (sum(process_cost(q) for q in vx.procs), b, wd)
            self.state[vx.id] = self._bytes(r | w)
            self.members[vx.id] = [vx.id]
        self.base = self._predict([0.0] * N_FEATURES)
        self._cache: Dict[tuple, tuple] = {}

    @staticmethod
    def _stored_like(comp) -> Set:
        """Variables the partition will keep in the SoA arrays, as far as it
        is known before partitioning: ports, memories, state, and values
        that more than one vertex touches (the rest are vertex-local
        temporaries)."""
        sch = comp.sch
        touched: Dict[object, int] = {}
        for vx in sch.vertices:
            for v in vx.reads | vx.need_old | vx.writes:
                touched[v] = touched.get(v, 0) + 1
        out = set()
        for v, n in touched.items():
            if v.kind in ("input", "output") or v.is_mem or sch.roles.get(v) == "state" or n > 1:
                out.add(v)
        return out

    def _bytes(self, bits: int) -> int:
        total = 0
        nb = self.nbytes
        while bits:
            low = bits & -bits
            total += nb[low.bit_length() - 1]
            bits ^= low
        return total

    def _predict(self, raw: Sequence[float]) -> float:
        x = (np.log1p(np.asarray(raw, np.float64)) - self.mu) / self.sd
        h1 = np.maximum(self.w1 @ x + self.b1, 0.0)
        h2 = np.maximum(self.w2 @ h1 + self.b2, 0.0)
        return float(np.exp((self.w3 @ h2 + self.b3) * self.ys + self.ym))

    def _x_union(self, a: int, b: int) -> int:
        """Written-and-read-outside set of the union of mtasks a and b."""
        xa, xb = self.X[a], self.X[b]
        out = xa | xb
        owner = self.owner
        for cand in ((xa & self.R[b]), (xb & self.R[a])):
            while cand:
                low = cand & -cand
                cand ^= low
                if all(owner[r] in (a, b) for r in self.readers[low.bit_length() - 1]):
                    out &= ~low
        return out

    def _union(self, a: int, b: int) -> tuple:
        r = self.R[a] | self.R[b]
        w = self.W[a] | self.W[b]
        x = self._x_union(a, b)
        ca, cb = self.cnt[a], self.cnt[b]
        state = self.state[a] + self.state[b] - self._bytes((self.R[a] | self.W[a]) & (self.R[b] | self.W[b]))
        feats = [float(ca[0] + cb[0]), float(bin(r & ~w).count("1")), float(bin(x).count("1")),
                 float(state), float(ca[1] + cb[1]), float(ca[2] + cb[2])]
        us = self._predict(feats)
        return max(1, int(round((us - self.base) * self.scale))), x, state

    def vertex_cost(self, vid: int) -> int:
        r, w, x = self.R[vid], self.W[vid], self.X[vid]
        c = self.cnt[vid]
        feats = [float(c[0]), float(bin(r & ~w).count("1")), float(bin(x).count("1")),
                 float(self.state[vid]), float(c[1]), float(c[2])]
        return max(1, int(round((self._predict(feats) - self.base) * self.scale)))

    def _get(self, ta, tb) -> tuple:
        key = (ta.id, ta.gen, tb.id, tb.gen) if ta.id < tb.id else (tb.id, tb.gen, ta.id, ta.gen)
        u = self._cache.get(key)
        if u is None:
            u = self._union(ta.id, tb.id)
            self._cache[key] = u
        return u

    def cost(self, ta, tb) -> int:
        return self._get(ta, tb)[0]

    def merge(self, keep: int, gone: int):
        _, x, state = self._union(keep, gone)
        self.X[keep] = x
        self.state[keep] = state
        self.R[keep] |= self.R.pop(gone)
        self.W[keep] |= self.W.pop(gone)
        self.X.pop(gone)
        self.state.pop(gone)
        ck, cg = self.cnt[keep], self.cnt.pop(gone)
        self.cnt[keep] = (ck[0] + cg[0], ck[1] + cg[1], ck[2] + cg[2])
        moved = self.members.pop(gone)
        for v in moved:
            self.owner[v] = keep
        self.members[keep].extend(moved)


def learned_merge(model: Optional[Dict] = None):
    """Options.merge_fn for a compile: the committed trained model's
    LearnedMerge of the design being compiled."""
    if model is None:
        import json
        from pathlib import Path
        model = json.loads((Path(__file__).parent / "costtables" / "model.json").read_text())

    def make(comp):
        return LearnedMerge(model, comp)
    return make


# =====================================================================
# Training data, training, prediction
# =====================================================================
#
# 1. build_variants() (CPU): partition variants of registered designs
#    (threads 2..16 give mtasks of every size); each library gets the
#    features of its mtasks, and the targets the features of their vertices.
# 2. profile_variants() (GPU): every mtask kernel timed alone at a fixed lane
#    count (rf_profile_mtasks).
# 3. fit() (GPU): the MLP learns log(us) from log1p(features).
# 4. vertex_costs(): predicted marginal cost of every vertex of a target,
#    written as a cost table that TableCost feeds to the partitioner.

TRAIN_DESIGNS = ("net_small", "net2", "alu", "systolic", "rv32i", "fuzz1", "fuzz2", "fuzz3", "fuzz4",
                 "bfuzz1", "bfuzz2", "bfuzz3")
TRAIN_THREADS = (2, 4, 8, 16)
TARGETS = ("net2", "systolic", "rv32i", "net5s")
PROFILE_LANES = 16384


def _cm_dir():
    from .build import BUILD
    return BUILD / "costmodel"


def build_variants(names=TRAIN_DESIGNS, threads=TRAIN_THREADS, jobs: int = 8, log=print) -> dict:
    import json
    from . import compile_design
    from .build import BUILD
    from .designs import registry
    index = {}
    for name in names:
        e = registry.ENTRIES[name]
        spec = e.make()
        for t in threads:
            opts = dict(e.opts)
            opts["threads"] = t
            opts["bitplane"] = False
            vname = f"{name}_t{t}"
            cd = compile_design(spec["verilog"], top=spec["top"], name=vname, jobs=jobs, **opts)
            comp = cd.compiled
            fx = FeatureExtractor(comp)
            rec = {"design": name, "threads": t, "lib": str(cd.lib_path.relative_to(BUILD)),
                   "mtask_features": [fx.mtask(k) for k in range(len(comp.mtasks))]}
            if name in TARGETS and t == e.opts.get("threads"):
                rec["vertex_features"] = [fx.vertex(v.id) for v in comp.sch.vertices]
            (cd.lib_path.parent / "features.json").write_text(json.dumps(rec))
            index[vname] = rec["lib"]
            if log:
                log(f"variant {vname}: {len(comp.mtasks)} mtasks")
    d = _cm_dir()
    d.mkdir(parents=True, exist_ok=True)
    (d / "index.json").write_text(json.dumps(index, indent=1))
    return index


def profile_variants(lanes: int = PROFILE_LANES, reps: int = 20, log=print) -> List[dict]:
    """GPU: time every mtask of every variant (mean of ``reps`` launches)."""
    import json
    from .build import BUILD
    from .model import MODE_GRAPH, DesignLib, Model
    index = json.loads((_cm_dir() / "index.json").read_text())
    samples = []
    for vname, lib in index.items():
        path = BUILD / lib
        rec = json.loads((path.parent / "features.json").read_text())
        dl = DesignLib(path)
        n = lanes
        with Model(dl, n) as m:
            m.begin(0x5EED)
            m.run_cycles(4, MODE_GRAPH)
            ms = m.profile_mtasks(reps)
        for k, f in enumerate(rec["mtask_features"]):
            samples.append({"variant": vname, "design": rec["design"], "threads": rec["threads"],
                            "mtask": k, "features": f, "us": float(ms[k]) * 1e3})
        if log:
            log(f"profiled {vname}: {len(rec['mtask_features'])} mtasks, "
                f"{float(ms.sum()) * 1e3:.1f} us per eval")
    return samples


def fit(samples: List[dict], epochs: int = 600, seed: int = 1) -> Dict:
    x = encode([s["features"] for s in samples])
    y = np.log(np.maximum([s["us"] for s in samples], 1e-3)).astype(np.float32)
    model = train(x, y, epochs=epochs, lr=2e-3, batch=128, seed=seed)
    model["features"] = FEATURE_NAMES
    model["lanes"] = PROFILE_LANES
    return model


def vertex_costs(model: Dict, vertex_features: List[List[float]], scale: float = 100.0) -> List[int]:
    """Predicted marginal cost of each vertex: predicted time of a kernel
    holding only that vertex, minus that of an empty kernel (the launch),
    in units of 1/scale us."""
    lc = LearnedCost(model)
    us = lc.predict_us(vertex_features)
    base = float(lc.predict_us([[0.0] * N_FEATURES])[0])
    lc.net.close()
    return [max(1, int(round((u - base) * scale))) for u in us]


class TableCost:
    """Partitioner cost function from a precomputed per-vertex cost table
    (vertex ids follow the deterministic schedule of the same source)."""

    def __init__(self, costs: List[int]):
        self.costs = list(costs)

    def __call__(self, comp, vx) -> int:
        if len(self.costs) != len(comp.sch.vertices):
            raise ValueError(f"cost table has {len(self.costs)} entries, design has "
                             f"{len(comp.sch.vertices)} vertices")
        return self.costs[vx.id]


CANDIDATE_THREADS = (2, 3, 4, 6, 8, 12, 16, 24, 32)


def partition_candidates(names=TARGETS, threads=CANDIDATE_THREADS, log=print) -> dict:
    """CPU: partition each target at several granularities (no HIP build) and
    record every candidate's mtask features, for the learned choice of the
    partition (``choose_partitions``)."""
    import json
    from .compiler import Compiled, Options
    from .designs import registry
    from .verilog import elaborate
    out = {}
    for name in names:
        e = registry.ENTRIES[name]
        spec = e.make()
        cands = {}
        for t in threads:
            opts = {k: v for k, v in e.opts.items() if k not in ("cost_table", "learned_threads")}
            opts.update(threads=t, bitplane=False)
            comp = Compiled(elaborate(spec["verilog"], spec["top"]), Options(**opts))
            fx = FeatureExtractor(comp)
            cands[str(t)] = [fx.mtask(k) for k in range(len(comp.mtasks))]
            if log:
                log(f"candidate {name} threads={t}: {len(comp.mtasks)} mtasks")
        out[name] = cands
    d = _cm_dir()
    d.mkdir(parents=True, exist_ok=True)
    path = d / "candidates.json"
    merged = json.loads(path.read_text()) if path.exists() else {}
    merged.update(out)
    path.write_text(json.dumps(merged))
    return out


def choose_partitions(model: Dict, candidates: dict) -> dict:
    """GPU: the MLP's predicted time of one eval (sum over the candidate's
    mtask kernels) for every candidate granularity; the choice is the
    fastest. The instruction-count heuristic cannot make this choice: every
    candidate has the same total instruction count."""
    lc = LearnedCost(model)
    res = {}
    for name, cands in candidates.items():
        pred = {t: float(lc.predict_us(f).sum()) for t, f in cands.items()}
        best = min(pred, key=pred.get)
        res[name] = {"threads": int(best), "predicted_us": pred}
    lc.net.close()
    return res


def main(argv=None):
    """GPU workflow. ``python -m rtlflow_amd.costmodel OUTDIR`` profiles the
    variants, trains, and predicts the targets' vertex costs and partition
    choices; ``python -m rtlflow_amd.costmodel OUTDIR --choose MODEL`` only
    predicts the partition choices with a trained model; ``--candidates
    PATH`` reads the candidates from PATH instead of the build folder."""
    import json
    import sys
    from pathlib import Path
    from .build import BUILD
    argv = argv if argv is not None else sys.argv[1:]
    out = Path(argv[0] if argv else "costmodel_out")
    out.mkdir(parents=True, exist_ok=True)
    cand_path = _cm_dir() / "candidates.json"
    if "--candidates" in argv:
        cand_path = Path(argv[argv.index("--candidates") + 1])
    if "--choose" in argv:
        model = json.loads(Path(argv[argv.index("--choose") + 1]).read_text())
    else:
        samples = profile_variants(log=lambda *a: print(*a, flush=True))
        (out / "samples.json").write_text(json.dumps(samples))
        designs = sorted({s["design"] for s in samples})
        model = fit(samples)
        print(f"trained on {len(samples)} mtasks from {len(designs)} designs: "
              f"val mse {model['history'][-1][2]:.4f}", flush=True)
        (out / "model.json").write_text(json.dumps(model))
        index = json.loads((_cm_dir() / "index.json").read_text())
        for vname, lib in index.items():
            rec = json.loads(((BUILD / lib).parent / "features.json").read_text())
            if "vertex_features" in rec:
                costs = vertex_costs(model, rec["vertex_features"])
                (out / f"costs_{rec['design']}.json").write_text(json.dumps(costs))
                print(f"cost table {rec['design']}: {len(costs)} vertices", flush=True)
    if cand_path.exists():
        choice = choose_partitions(model, json.loads(cand_path.read_text()))
        (out / "partition_choice.json").write_text(json.dumps(choice, indent=1))
        for name, c in choice.items():
            print(f"partition choice {name}: threads={c['threads']} "
                  f"(predicted {c['predicted_us'][str(c['threads'])]:.1f} us per eval)", flush=True)


if __name__ == "__main__":
    main()
