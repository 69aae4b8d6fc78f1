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
GPU (rf_profile_mtasks). The heuristic remains the fallback and the judge: a
learned schedule is kept only if its measured throughput is higher and its
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
