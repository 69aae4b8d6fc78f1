"""Autotuner: per design, the form and launch shape with the best measured
throughput among bit-exact candidates.

Knobs:

* compile time (each candidate is its own design library):
  - ``block``: stimuli per workgroup of the per-lane kernels (64 / 128 / 256);
  - ``threads``: the mtask fusion granularity (V3Partition's ``--threads``);
  - ``lane_major_min_depth``: the SoA layout of memories (lane-major region
    from this depth up, element-major below);
  - ``inline_limit``: all state of a fused kernel in registers, or the
    "calls" form whose mtask bodies work on the lane-minor arrays (for
    designs whose state does not fit in registers);
  - ``split_reductions``: long combinational reductions split into partial
    terms that the partitioner can place in separate mtasks;
* run time:
  - the form (bit-plane, fused, plain graph, batched queue);
  - bit-plane workgroup size (256 / 512 / 1,024 threads per 64 lanes);
  - fused launch length (cycles per launch);
  - the fused form's wavefronts per 64 stimuli: one, or the design's
    multi-wavefront kernel (rf_config.fused_multiwave, the mtask levels
    spread over several wavefronts with the lanes' slots in LDS);
  - batched form: lanes in the queue;
  - the SoA padding (extra lanes in every slot's row, rf_config.lane_pad)
    of the per-lane forms.

``build_candidates(name)`` compiles the compile-time candidates on the CPU;
``tune(name, checker)`` measures every candidate on the GPU on a bounded
workload, checks each one's hashes on a sample of stimuli against the
hashes ``checker`` gives for them (an independent model of the design: the
tests' CPU oracle, ``tests/autotune_check.py``), and writes the fastest
bit-exact one to ``rtlflow_amd/tuned/<name>.json`` together with the full
table; a tune in which no candidate matches fails. The cache records which
library build it measured.
"""
from __future__ import annotations

import json
import time
from pathlib import Path
from typing import Dict, List, Optional

import numpy as np

TUNED = Path(__file__).resolve().parent / "tuned"

# bounded tuning workloads (a fraction of each benchmark configuration)
WORKLOADS = {
    "net2": {"kind": "cycles", "lanes": 65536, "cycles": 200, "cfg": 2},
    "alu": {"kind": "cycles", "lanes": 1024, "cycles": 2000, "cfg": 1},
    "systolic": {"kind": "cycles", "lanes": 8192, "cycles": 500, "cfg": 4},
    "rv32i": {"kind": "queue", "nstim": 4096, "median": 2000, "cfg": 3},
    "net5s": {"kind": "cycles", "lanes": 4096, "cycles": 200, "cfg": 5},
}

COMPILE_SPACE = {
    "block": (64, 128, 256),
    # 0: every design in the "calls" form (mtask bodies as functions over the
    # lane-minor arrays) instead of all state in one kernel's registers
    "inline_limit": (6000, 0),
    # V3Partition's --threads: the critical-path limit of mtask merging
    "threads": (2, 4, 8, 16),
}
# SoA padding: extra lanes in every slot's row (rf_config.lane_pad)
LANE_PAD = (0, 64)
# operands per partial term of a split reduction (Options.split_reductions)
SPLIT_REDUCTIONS = 32
MEM_LAYOUT = (64, 1 << 30)


def _variant_name(name: str, opts: Dict) -> str:
    return name + "__" + "_".join(f"{k[0]}{v}" for k, v in sorted(opts.items()))


def compile_candidates(name: str) -> List[Dict]:
    """Compile-time option sets for a registered design. Designs with a
    bit-plane program are tuned at run time only (that form wins by an order
    of magnitude and does not depend on these options)."""
    from .designs import registry
    rec = registry.load_index().get(name)
    if rec is not None:
        man = json.loads((registry.BUILD / rec["lib"]).parent.joinpath("manifest.json").read_text())
        if man.get("bitplane", {}).get("supported"):
            return []
        has_mem = bool(man.get("mems"))
    else:
        has_mem = True
    base = {k: v for k, v in registry.ENTRIES[name].opts.items() if k in ("threads",)}
    out = []
    for block in COMPILE_SPACE["block"]:
        for lm in (MEM_LAYOUT if has_mem else (64,)):
            for il in COMPILE_SPACE["inline_limit"]:
                o = dict(base)
                o.update(block=block, lane_major_min_depth=lm, inline_limit=il)
                out.append(o)
    # the fusion granularity, one knob at a time from the default shape
    for t in COMPILE_SPACE["threads"]:
        if t != base.get("threads"):
            out.append({"threads": t, "block": 256, "lane_major_min_depth": 64, "inline_limit": 6000})
    # long combinational reductions split into partial terms, at a fine and
    # a coarse granularity (only designs that have such a reduction)
    # (with the multi-wavefront kernel's wavefronts capped at 8 for the finer
    # partitions, whose plans would otherwise take more wavefronts than their
    # registers allow)
    if _long_reductions(name, SPLIT_REDUCTIONS):
        for t, sr, mw in ((4, SPLIT_REDUCTIONS, 16), (16, SPLIT_REDUCTIONS, 16), (8, SPLIT_REDUCTIONS, 8),
                          (16, SPLIT_REDUCTIONS, 8), (4, SPLIT_REDUCTIONS // 2, 16)):
            o = {"threads": t, "block": 256, "lane_major_min_depth": 64, "inline_limit": 6000,
                 "split_reductions": sr}
            if mw != 16:
                o["mw_max_waves"] = mw
            out.append(o)
    return out


def _long_reductions(name: str, max_terms: int) -> bool:
    """Whether the design has a combinational reduction of more than
    ``max_terms`` operands (what the split_reductions option splits)."""
    from .designs import registry
    from .order import split_reductions
    from .verilog import elaborate
    spec = registry.ENTRIES[name].make()
    return split_reductions(elaborate(spec["verilog"], spec["top"]), max_terms) > 0


def build_candidates(name: str, jobs: int = 8, log=print) -> Dict[str, str]:
    """CPU: build every compile-time candidate of ``name``."""
    from . import compile_design
    from .build import BUILD
    from .designs import registry
    e = registry.ENTRIES[name]
    spec = e.make()
    built = {}
    for o in compile_candidates(name):
        opts = {k: v for k, v in e.opts.items() if k not in ("cost_table", "learned_threads")}
        opts.update(o)
        opts["bitplane"] = False
        vname = _variant_name(name, o)
        cd = compile_design(spec["verilog"], top=spec["top"], name=vname, jobs=jobs, **opts)
        built[vname] = str(cd.lib_path.relative_to(BUILD))
        if log:
            log(f"built {vname}")
    d = BUILD / "tune"
    d.mkdir(parents=True, exist_ok=True)
    (d / f"{name}.json").write_text(json.dumps({"variants": built, "opts": compile_candidates(name)}))
    return built


def _inputs(name: str, wl: Dict):
    from .designs import registry, rv32i
    seed = registry.SEED_CFG[wl["cfg"]]
    if name == "rv32i":
        progs = rv32i.programs(seed, np.arange(wl["nstim"]))
        budgets = rv32i.budgets(seed, wl["nstim"], median=wl["median"], lo=500, hi=20000)
        return seed, progs, budgets
    return seed, None, None


def _measure(cd, name: str, wl: Dict, form: str, knobs: Dict):
    """(stimulus-cycles/s, hashes) of one candidate."""
    from .model import MODE_BITPLANE, MODE_FUSED, MODE_GRAPH
    seed, progs, budgets = _inputs(name, wl)
    kw = {}
    if "bp_threads" in knobs:
        kw["bitplane_threads"] = knobs["bp_threads"]
    if "chunk" in knobs:
        kw["fused_chunk"] = knobs["chunk"]
    if "lane_pad" in knobs:
        kw["lane_pad"] = knobs["lane_pad"]
    if knobs.get("multiwave"):
        kw["fused_multiwave"] = 1
    if wl["kind"] == "queue":
        lanes = knobs["lanes"]
        with cd.model(lanes, **kw) as m:
            m.set_mem_images("imem", progs)
            m.sync()
            t0 = time.perf_counter()
            h = m.submit_queue(seed, budgets)
            dt = time.perf_counter() - t0
        return float(budgets.astype(np.int64).sum()) / dt, h
    mode = {"bitplane": MODE_BITPLANE, "fused": MODE_FUSED, "graph": MODE_GRAPH}[form]
    lanes, cycles = wl["lanes"], wl["cycles"]
    with cd.model(lanes, **kw) as m:            # warm-up: code load, first launches
        m.begin(seed)
        m.run_cycles(min(cycles, 50), mode)
        m.sync()
    with cd.model(lanes, **kw) as m:
        m.begin(seed)
        m.sync()
        t0 = time.perf_counter()
        m.run_cycles(cycles, mode)
        m.sync()
        dt = time.perf_counter() - t0
        h = m.hashes()
    return lanes * cycles / dt, h


def runtime_candidates(name: str, man: Dict, wl: Dict) -> List[tuple]:
    out = []
    if wl["kind"] == "queue":
        for lanes in (wl["nstim"] // 4, wl["nstim"] // 2, wl["nstim"]):
            for pad in LANE_PAD:
                out.append(("batched", {"lanes": lanes, "lane_pad": pad}))
        return out
    if man.get("bitplane", {}).get("supported"):
        for nt in (256, 512, 1024):
            out.append(("bitplane", {"bp_threads": nt}))
    for chunk in (64, 512):
        for pad in LANE_PAD:
            out.append(("fused", {"chunk": chunk, "lane_pad": pad}))
    if man.get("mw_waves", 0) >= 2:
        out.append(("fused", {"chunk": 512, "multiwave": 1}))
    out.append(("graph", {}))
    return out


def sample_indices(name: str, wl: Dict) -> np.ndarray:
    """Stimuli whose hashes are checked: the first and the last 64."""
    n = wl["nstim"] if wl["kind"] == "queue" else wl["lanes"]
    k = min(64, n)
    idx = list(range(k)) + (list(range(n - k, n)) if n > k else [])
    return np.array(idx)


def tune(name: str, checker, log=print) -> Dict:
    """GPU: measure every (library, form, knobs) candidate; write the cache.
    ``checker(name, workload, indices)`` returns the expected hashes of the
    workload's stimuli ``indices``."""
    from . import CompiledDesign
    from .build import BUILD
    from .designs import registry
    wl = WORKLOADS[name]
    libs = {name: registry.get(name)}
    copts = {name: {}}
    tf = BUILD / "tune" / f"{name}.json"
    if tf.exists():
        rec = json.loads(tf.read_text())
        for (vname, rel), o in zip(rec["variants"].items(), rec["opts"]):
            lp = BUILD / rel
            libs[vname] = CompiledDesign(lp, json.loads((lp.parent / "manifest.json").read_text()))
            copts[vname] = o
    rows = []
    idx = sample_indices(name, wl)
    want = np.asarray(checker(name, wl, idx), dtype=np.uint64)
    for vname, cd in libs.items():
        for form, knobs in runtime_candidates(name, cd.manifest, wl):
            rate, h = _measure(cd, name, wl, form, knobs)
            exact = bool(np.array_equal(h[idx], want))
            rows.append({"library": vname, "compile": copts[vname], "form": form, "knobs": knobs,
                         "rate": rate, "bit_exact": exact, "mtasks": cd.manifest["mtasks"],
                         "checked_stimuli": int(len(idx))})
            if log:
                log(f"{vname:40s} {form:9s} {json.dumps(knobs):24s} {rate:.4g} exact={exact}")
    exact_rows = [r for r in rows if r["bit_exact"]]
    if not exact_rows:
        raise RuntimeError(f"autotune {name}: no candidate matches the checker on the sampled stimuli")
    best = max(exact_rows, key=lambda r: r["rate"])
    res = {"design": name, "library": str(libs[name].lib_path.parent.name), "workload": wl,
           "best": best, "table": rows}
    TUNED.mkdir(parents=True, exist_ok=True)
    (TUNED / f"{name}.json").write_text(json.dumps(res, indent=1))
    return res


def choice(name: str) -> Optional[Dict]:
    """The cached best candidate for a design, or None."""
    p = TUNED / f"{name}.json"
    if not p.exists():
        return None
    return json.loads(p.read_text())["best"]
