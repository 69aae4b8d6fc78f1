"""Named designs used by the tests, smoke check and benchmark.

``build_all()`` (called by __graft_entry__.build on the CPU) generates,
compiles and links every entry and records it in ``_build/designs/index.json``;
``get(name)`` on a GPU machine only reads that index and loads the prebuilt
library -- it never compiles. Each entry also keeps its generator inputs so
the CPU oracle can be driven with exactly the same design.
"""
from __future__ import annotations

import json
import time
from pathlib import Path
from typing import Callable, Dict, Optional

from ..build import BUILD

INDEX = BUILD / "designs" / "index.json"

# config seeds: "RTLF" + config number
SEED_CFG = {k: 0x52544C4600000000 + k for k in range(1, 6)}


class Entry:
    def __init__(self, name: str, make: Callable[[], dict], opts: dict):
        self.name = name
        self.make = make       # -> {"verilog": str, "top": str, "extra": {...}}
        self.opts = opts


def _netlist(**kw):
    def make():
        from .netlist import generate
        src, table = generate(**kw)
        return {"verilog": src, "top": kw.get("name", "netlist"), "table": table.to_bytes()}
    return make


def _alu(**kw):
    def make():
        from .alu import verilog
        return {"verilog": verilog(), "top": "alu"}
    return make


def _systolic(**kw):
    def make():
        from .systolic import verilog
        return {"verilog": verilog(**kw), "top": "systolic"}
    return make


def _rv32i(**kw):
    def make():
        from .rv32i import verilog
        return {"verilog": verilog(**kw), "top": "rv32i"}
    return make


ENTRIES: Dict[str, Entry] = {}


def register(name: str, make, **opts):
    ENTRIES[name] = Entry(name, make, opts)


# small netlists for parity tests
register("net_tiny", _netlist(n_gates=300, n_ff=30, n_pi=16, n_po=16, depth=8, seed=7,
                              name="net_tiny"), threads=4)
register("net_small", _netlist(n_gates=2000, n_ff=200, n_pi=64, n_po=64, depth=20,
                               seed=SEED_CFG[2] ^ 0xabc, name="net_small"), threads=4)
# configuration 2: 20,000 gates, 2,000 flip-flops, 64 PI / 64 PO, depth 40
register("net2", _netlist(seed=SEED_CFG[2], name="net2"), threads=4)
# configuration 5 at one tenth of its gate and flip-flop counts (50,000 gates,
# 5,000 flip-flops) with its full RAM complement: eight 1,024 x 64-bit
# synchronous RAMs, 128 primary inputs, 64 primary outputs
register("net5s", _netlist(n_gates=50000, n_ff=5000, n_pi=128, n_po=64, depth=40, seed=SEED_CFG[5],
                           n_rams=8, ram_depth=1024, ram_width=64, name="net5s"), threads=4)
# configuration 1: 4-stage pipelined ALU with a 32 x 32 register file
register("alu", _alu(), threads=2)
# configuration 4: 16 x 16 int8 systolic array + 4,096-bit shift/compare chain
register("systolic", _systolic(), threads=2)
# configuration 3: 5-stage RV32I pipeline, 64 KB instruction ROM, 16 KB data RAM
register("rv32i", _rv32i(), threads=2)


def _file_text(path: Path) -> dict:
    return {}


COSTTABLES = Path(__file__).resolve().parent.parent / "costtables"


def build_entry(name: str, jobs: int = 8, force: bool = False) -> dict:
    from .. import compile_design
    e = ENTRIES[name]
    t0 = time.time()
    spec = e.make()
    opts = dict(e.opts)
    learned = opts.pop("learned_threads", None)
    if learned:
        choice = json.loads((COSTTABLES / "partition_choice.json").read_text())
        opts["threads"] = int(choice[learned]["threads"])
    if opts.pop("learned_merge", False):
        # contraction scored by the learned cost of each candidate union
        from ..costmodel import learned_merge
        opts["merge_fn"] = learned_merge()
    table = opts.pop("cost_table", None)
    if table:
        # partition with the learned per-vertex costs instead of instruction counts
        from ..costmodel import TableCost
        opts["cost_fn"] = TableCost(json.loads((COSTTABLES / f"{table}.json").read_text()))
    cd = compile_design(spec["verilog"], top=spec["top"], name=name, jobs=jobs, force=force, **opts)
    rec = {"name": name, "lib": str(cd.lib_path.relative_to(BUILD)), "opts": e.opts,
           "build_s": round(time.time() - t0, 1)}
    from ..build import build_testbench
    exe = build_testbench(cd.lib_path, cd.manifest)
    rec["testbench"] = str(exe.relative_to(BUILD))
    if "table" in spec:
        tp = cd.lib_path.parent / "gates.bin"
        tp.write_bytes(spec["table"])
        rec["table"] = str(tp.relative_to(BUILD))
    return rec


def build_all(names=None, jobs: int = 8, log=print, force: bool = False, workers: int = 1) -> dict:
    """Build registered designs (all by default). ``force`` recompiles
    libraries whose generated source is unchanged; ``workers`` > 1 builds
    that many designs at once in separate processes, the slowest (by their
    last build time) first."""
    index = load_index()
    names = list(names or ENTRIES)

    def done(name, rec):
        index[name] = rec
        if log:
            log(f"built {name}: {rec['lib']} ({rec['build_s']} s)")
        INDEX.parent.mkdir(parents=True, exist_ok=True)
        INDEX.write_text(json.dumps(index, indent=1))
        _prune(name, Path(rec["lib"]).parent.name)

    if workers <= 1:
        for name in names:
            done(name, build_entry(name, jobs=jobs, force=force))
        return index
    from concurrent.futures import ProcessPoolExecutor, as_completed
    names.sort(key=lambda n: -index.get(n, {}).get("build_s", 0.0))
    with ProcessPoolExecutor(max_workers=workers) as ex:
        futs = {ex.submit(build_entry, n, 1, force): n for n in names}
        for f in as_completed(futs):
            done(futs[f], f.result())
    return index


def _prune(name: str, keep_dir: str):
    """Remove older builds of one design (they would otherwise travel with the
    tree to every GPU run)."""
    import shutil
    root = BUILD / "designs"
    for d in root.glob(f"{name}_*"):
        suffix = d.name[len(name) + 1:]
        if d.is_dir() and d.name != keep_dir and len(suffix) == 16 and \
                all(c in "0123456789abcdef" for c in suffix):
            shutil.rmtree(d)


def load_index() -> dict:
    if INDEX.exists():
        return json.loads(INDEX.read_text())
    return {}


def get(name: str):
    """Prebuilt design -> CompiledDesign (raises if it was not built)."""
    from .. import CompiledDesign
    index = load_index()
    if name not in index:
        raise RuntimeError(f"design {name!r} is not built; run __graft_entry__.build()")
    rec = index[name]
    lib = BUILD / rec["lib"]
    man = json.loads((lib.parent / "manifest.json").read_text())
    cd = CompiledDesign(lib, man)
    cd.record = rec
    return cd


def table_bytes(name: str) -> bytes:
    rec = load_index()[name]
    return (BUILD / rec["table"]).read_bytes()


def _fuzz(seed: int):
    def make():
        from .fuzz import verilog
        return {"verilog": verilog(seed, name=f"fuzz{seed}"), "top": f"fuzz{seed}"}
    return make


# the same designs partitioned by the learned cost model: the granularity
# whose predicted eval time (sum of the MLP's predicted mtask kernel times)
# is least ("learned_threads"); "cost_table" would also feed the MLP's
# per-vertex costs to the contraction (measured slower on net2)
register("net2_learned", _netlist(seed=SEED_CFG[2], name="net2"), learned_threads="net2",
         bitplane=False)
register("systolic_learned", _systolic(), learned_threads="systolic", bitplane=False)
register("net5s_learned", ENTRIES["net5s"].make, learned_threads="net5s", bitplane=False)
register("rv32i_learned", _rv32i(), learned_threads="rv32i", bitplane=False)
# the learned granularity with the contraction's merges scored by the
# learned cost of each candidate union (costmodel.LearnedMerge)
register("net2_lmerge", _netlist(seed=SEED_CFG[2], name="net2"), learned_threads="net2",
         learned_merge=True, bitplane=False)
register("systolic_lmerge", _systolic(), learned_threads="systolic", learned_merge=True, bitplane=False)
# net5s with learned merges is not registered: its graph form hit a GPU
# memory fault whose cause is not found (the host build of the same generated
# kernels is bit-exact and clean under AddressSanitizer)
register("rv32i_lmerge", _rv32i(), learned_threads="rv32i", learned_merge=True, bitplane=False)

# the autotuner's choice of compile-time options, where it measured one
# (rtlflow_amd/tuned/<name>.json, written by rtlflow_amd.autotune on the GPU)
def _register_tuned():
    tuned = Path(__file__).resolve().parent.parent / "tuned"
    for name in ("alu", "systolic", "rv32i"):
        p = tuned / f"{name}.json"
        if not p.exists():
            continue
        best = json.loads(p.read_text())["best"]
        if best.get("compile"):
            e = ENTRIES[name]
            opts = dict(e.opts)
            opts.update(best["compile"])
            register(f"{name}_tuned", e.make, **opts)


_register_tuned()

# random operator-coverage designs (compared with the IR interpreter)
for _s in (1, 2, 3, 4):
    register(f"fuzz{_s}", _fuzz(_s), threads=3)


def _bfuzz(seed: int):
    def make():
        from .fuzz import verilog
        return {"verilog": verilog(seed, name=f"bfuzz{seed}", bitlevel=True, depth=4),
                "top": f"bfuzz{seed}"}
    return make


# synchronous memories in the bit-plane form: a RAM netlist (lane-major
# 9-bit RAMs) and the memory-port coverage design
register("net_ram", _netlist(n_gates=2000, n_ff=200, n_pi=70, n_po=40, depth=12, seed=SEED_CFG[5] ^ 0x77,
                             n_rams=4, ram_depth=100, ram_width=9, name="net_ram"), threads=3)


def _rams():
    def make():
        from .rams import verilog
        return {"verilog": verilog(), "top": "rams"}
    return make


register("rams", _rams(), threads=2)


def _mix():
    def make():
        from .mix import VERILOG
        return {"verilog": VERILOG, "top": "mix"}
    return make


# every port shape of the input stream / output capture layouts
register("mix", _mix(), threads=2)

# combinational cycles: one that settles, one that never does
def _settle(which: str):
    def make():
        from . import settle
        return {"verilog": getattr(settle, which.upper()), "top": which}
    return make


register("circ", _settle("circ"), threads=2)
register("osc", _settle("osc"), threads=2)
register("smear", _settle("smear"), threads=2)
register("gated", _settle("gated"), threads=2)

# random designs inside the bit-plane subset (no memories, no division)
for _s in (1, 2, 3):
    register(f"bfuzz{_s}", _bfuzz(_s), threads=2)


# known-answer designs over the reference's arithmetic test vectors
def _kat(family: str):
    def make():
        from . import kat
        return {"verilog": kat.verilog(family), "top": f"kat_{family}"}
    return make


for _f in ("shift", "div", "signed", "red"):
    register(f"kat_{_f}", _kat(_f), threads=2)


# ---------------------------------------------------------------------------
# gate-level netlists run by the gate-net engine (include/rf_gatenet.h):
# configuration 5 at its stated size, and a small one of the same shape for
# the GPU parity tests. Each build writes the gate table and the program
# blobs of the heuristic partition (--threads 4) and, when the cost model
# has chosen one (costtables/partition_choice.json), the learned partition.
GATENETS = {
    "net5": dict(n_gates=500000, n_ff=50000, n_pi=128, n_po=64, depth=40, seed=SEED_CFG[5], n_rams=8,
                 ram_depth=1024, ram_width=64),
    "net5_small": dict(n_gates=6000, n_ff=600, n_pi=128, n_po=64, depth=40, seed=SEED_CFG[5] ^ 0x5A,
                       n_rams=8, ram_depth=1024, ram_width=64),
}
GATENET_THREADS = 4


def gatenet_dir(name: str) -> Path:
    return BUILD / "gatenet" / name


def build_gatenet(name: str, log=print) -> dict:
    from .. import gatenet as gn
    from . import netlist as nl
    import numpy as np
    t0 = time.time()
    kw = GATENETS[name]
    d = gatenet_dir(name)
    d.mkdir(parents=True, exist_ok=True)
    table = nl.generate(**kw)[1] if kw["n_gates"] <= 20000 else _table_only(kw)
    tb = table.to_bytes()
    (d / "gates.bin").write_bytes(tb)
    t = gn.parse_table(tb)
    rec = {"name": name, "gates": kw["n_gates"], "schedules": {}}
    choice = json.loads((COSTTABLES / "partition_choice.json").read_text()).get(name)
    scheds = {"heuristic": GATENET_THREADS}
    if choice:
        scheds["learned"] = int(choice["threads"])
    for tag, th in scheds.items():
        prog = gn.program(t, threads=th)
        (d / f"{tag}.bin").write_bytes(prog.astype("<u4").tobytes())
        rec["schedules"][tag] = {"threads": th, "mtasks": int(prog[15]), "words": int(len(prog))}
    rec["build_s"] = round(time.time() - t0, 1)
    (d / "meta.json").write_text(json.dumps(rec, indent=1))
    if log:
        log(f"built gate net {name}: {rec['schedules']} ({rec['build_s']} s)")
    return rec


def _table_only(kw):
    """The generator's gate table without its Verilog text (half a million
    assign statements the gate-net engine does not read)."""
    from . import netlist as nl
    saved = nl.to_verilog
    nl.to_verilog = lambda t, name="": ""
    try:
        return nl.generate(**kw)[1]
    finally:
        nl.to_verilog = saved


def build_gatenets(names=None, log=print) -> dict:
    return {n: build_gatenet(n, log) for n in (names or GATENETS)}


def gatenet_meta(name: str) -> dict:
    p = gatenet_dir(name) / "meta.json"
    if not p.exists():
        raise RuntimeError(f"gate net {name!r} is not built; run __graft_entry__.build()")
    return json.loads(p.read_text())


def gatenet_table(name: str) -> bytes:
    gatenet_meta(name)
    return (gatenet_dir(name) / "gates.bin").read_bytes()


def gatenet_program(name: str, schedule: str = "heuristic"):
    import numpy as np
    meta = gatenet_meta(name)
    if schedule not in meta["schedules"]:
        raise RuntimeError(f"gate net {name!r} has no {schedule} schedule")
    return np.frombuffer((gatenet_dir(name) / f"{schedule}.bin").read_bytes(), "<u4").copy()
