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
* run time:
  - the form (bit-plane, fused, plain graph, batched queue);
  - bit-plane workgroup size (256 / 512 / 1,024 threads per 64 lanes);
  - fused launch length (cycles per launch);
  - batched form: lanes in the queue.

``build_candidates(name)`` compiles the compile-time candidates on the CPU;
``tune(name)`` measures every candidate on the GPU on a bounded workload,
checks each one's hashes against the first candidate's, and writes the
fastest to ``rtlflow_amd/tuned/<name>.json`` together with the full table.
The cache records which library build it measured.
``python -m rtlflow_amd.autotune NAME...`` runs ``tune`` (GPU).
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
}
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
    return out


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
            out.append(("batched", {"lanes": lanes}))
        return out
    if man.get("bitplane", {}).get("supported"):
        for nt in (256, 512, 1024):
            out.append(("bitplane", {"bp_threads": nt}))
    for chunk in (64, 512):
        out.append(("fused", {"chunk": chunk}))
    out.append(("graph", {}))
    return out


def tune(name: str, log=print) -> Dict:
    """GPU: measure every (library, form, knobs) candidate; write the cache."""
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
    ref = None
    for vname, cd in libs.items():
        for form, knobs in runtime_candidates(name, cd.manifest, wl):
            rate, h = _measure(cd, name, wl, form, knobs)
            exact = ref is None or bool(np.array_equal(h, ref))
            if ref is None:
                ref = h
            rows.append({"library": vname, "compile": copts[vname], "form": form, "knobs": knobs,
                         "rate": rate, "bit_exact": exact, "mtasks": cd.manifest["mtasks"]})
            if log:
                log(f"{vname:40s} {form:9s} {json.dumps(knobs):24s} {rate:.4g} exact={exact}")
    best = max((r for r in rows if r["bit_exact"]), key=lambda r: r["rate"])
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


def main(argv=None):
    import sys
    names = (argv if argv is not None else sys.argv[1:]) or list(WORKLOADS)
    for n in names:
        r = tune(n, log=lambda *a: print(*a, flush=True))
        b = r["best"]
        print(f"best {n}: {b['library']} {b['form']} {b['knobs']} {b['rate']:.4g} stimulus-cycles/s",
              flush=True)


if __name__ == "__main__":
    main()
