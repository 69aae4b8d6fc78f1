"""The heuristic schedule judges the learned one.

A design has up to three partitions: the instruction-count heuristic's
(V3InstrCount costs, src/V3Partition.cpp:456), the learned granularity (the
``--threads`` the MLP predicts fastest, ``*_learned`` libraries) and the
learned granularity with the contraction's merges scored by the MLP's cost of
each candidate union (``*_lmerge``). For each form a learned partition is
kept only if its measured throughput beats the heuristic's and its hashes
equal both the heuristic's and an independent checker's on the same stimuli;
otherwise the heuristic partition is used. The verdicts of a
GPU run are committed in ``costtables/judge.json``; ``scheduled()`` returns
the library a form should run.
"""
from __future__ import annotations

import json
import time
from pathlib import Path
from typing import Callable, Dict, Optional

import numpy as np

JUDGE = Path(__file__).resolve().parent / "costtables" / "judge.json"


def decide(heuristic_rate: float, learned_rate: float, heuristic_hashes, learned_hashes,
           expected=None, sample=None) -> Dict:
    """The verdict on measured rates and hashes: ``expected`` (the checker's
    hashes of the stimuli ``sample``, all stimuli when None) must equal both
    partitions' hashes, and the learned one must be faster."""
    hh = np.asarray(heuristic_hashes)
    lh = np.asarray(learned_hashes)
    same = hh.shape == lh.shape and bool((hh == lh).all())
    ok = same
    if expected is not None:
        idx = slice(None) if sample is None else np.asarray(sample)
        ok = ok and bool((lh[idx] == np.asarray(expected)).all())
    speedup = learned_rate / heuristic_rate if heuristic_rate > 0 else 0.0
    kept = bool(ok and speedup > 1.0)
    return {"heuristic": float(heuristic_rate), "learned": float(learned_rate),
            "speedup": float(speedup), "bit_exact": bool(ok), "kept": kept,
            "choice": "learned" if kept else "heuristic"}


def measure_graph(cd, lanes: int, cycles: int, seed: int):
    """(stimulus-cycles/s, hashes) of the per-mtask graph form; the model's
    set-up runs before the timed region."""
    from .model import MODE_GRAPH
    with cd.model(lanes) as m:
        m.begin(seed)
        m.run_cycles(0, MODE_GRAPH)
        m.run_cycles(2, MODE_GRAPH)      # graph instantiation and code load
        m.sync()
    with cd.model(lanes) as m:
        m.begin(seed)
        m.run_cycles(0, MODE_GRAPH)
        m.sync()
        t0 = time.perf_counter()
        m.run_cycles(cycles, MODE_GRAPH)
        m.sync()
        dt = time.perf_counter() - t0
        return lanes * cycles / dt, m.hashes()


def measure_batched(cd, lanes: int, budgets, seed: int):
    """(stimulus-cycles/s, hashes) of the batched persistent form."""
    b = np.asarray(budgets, np.uint32)
    with cd.model(lanes) as m:
        m.submit_queue(seed, b[: min(len(b), lanes)])   # code load
        m.sync()
        t0 = time.perf_counter()
        h = m.submit_queue(seed, b)
        dt = time.perf_counter() - t0
    return float(b.astype(np.int64).sum()) / dt, h


def judge(heuristic, learned, measure: Callable, checker: Optional[Callable] = None,
          sample=None) -> Dict:
    """Measure both partitions with ``measure(cd) -> (rate, hashes)`` and
    decide; ``checker(sample)`` gives the expected hashes of the stimuli
    ``sample``."""
    rh, hh = measure(heuristic)
    rl, hl = measure(learned)
    exp = checker(sample) if checker is not None else None
    v = decide(rh, rl, hh, hl, exp, sample)
    v["heuristic_mtasks"] = int(heuristic.manifest["mtasks"])
    v["learned_mtasks"] = int(learned.manifest["mtasks"])
    if heuristic.manifest.get("partition") == learned.manifest.get("partition"):
        # the MLP chose the heuristic's own granularity: nothing to judge
        v["identical"] = True
        v["kept"] = False
        v["choice"] = "heuristic"
    return v


def verdicts() -> Dict:
    return json.loads(JUDGE.read_text()) if JUDGE.exists() else {}


def record(name: str, form: str, verdict: Dict):
    d = verdicts()
    d.setdefault(name, {})[form] = verdict
    JUDGE.write_text(json.dumps(d, indent=1) + "\n")


# learned partitions a verdict may keep: the learned granularity alone
# (``<name>_learned``, verdict key ``form``) and with the learned merge
# scores (``<name>_lmerge``, verdict key ``form + MERGE``)
MERGE = "+merge"
VARIANTS = (("", "_learned"), (MERGE, "_lmerge"))


def scheduled(name: str, form: str):
    """The library a form of design ``name`` runs: of the learned partitions
    the committed verdicts kept, the fastest; the heuristic one when none
    was kept."""
    from .designs import registry
    index = registry.load_index()
    best, lib = 1.0, name
    for key, suffix in VARIANTS:
        v = verdicts().get(name, {}).get(form + key)
        if v and v.get("kept") and f"{name}{suffix}" in index and v["speedup"] > best:
            best, lib = v["speedup"], f"{name}{suffix}"
    return registry.get(lib)
