"""Build glue: hipcc rules for gfx950 (the reference's verilated.mk nvcc rules,
include/verilated.mk.in:14-20, :246-257, and the V3EmitMk-written Vtop.mk,
src/V3EmitMk.cpp:103-116).

* ``build_runtime()`` compiles ``librf_verilated.so`` (the runtime C ABI).
* ``build_design()`` compiles one generated design source into
  ``libV<name>.so`` exporting ``rf_design_get``; outputs are cached by the hash
  of the generated source and the compiler flags.
* ``emit_makefiles()`` writes ``V<top>.mk`` + a copy of ``verilated.mk`` for
  users who build with make instead of Python.

Everything is built in-tree under ``rtlflow_amd/_build``.
"""
from __future__ import annotations

import hashlib
import json
import os
import shutil
import subprocess
import threading
from pathlib import Path
from typing import Dict, List, Optional

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
INCLUDE = PKG.parent / "include"
BUILD = PKG / "_build"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# -packed-fp32-ops: keep hipcc from forming packed FP32 instructions
HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-variable", "-Wno-unused-but-set-variable", "-Wno-unused-function"]

_lock = threading.Lock()


class BuildError(RuntimeError):
    pass


def _run(cmd: List[str], cwd: Optional[Path] = None, log: Optional[Path] = None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if log is not None:
        log.write_text(" ".join(cmd) + "\n" + r.stdout)
    if r.returncode != 0:
        tail = "\n".join(r.stdout.splitlines()[-60:])
        raise BuildError(f"command failed ({r.returncode}): {' '.join(cmd[:6])} ...\n{tail}")
    return r.stdout


def runtime_path() -> Path:
    return BUILD / "librf_verilated.so"


def build_runtime(force: bool = False) -> Path:
    out = runtime_path()
    srcs = [CSRC / "rf_runtime.hip", CSRC / "rf_mlp.hip", CSRC / "rf_bitplane.hip", CSRC / "rf_gatenet.hip"]
    deps = srcs + [INCLUDE / "rf_verilated.h", INCLUDE / "rf_design.h", INCLUDE / "rf_mlp.h",
                   INCLUDE / "rf_gatenet.h", CSRC / "rf_bitplane.h", CSRC / "rf_stim.h",
                   CSRC / "rf_transpose.h"]
    if out.exists() and not force and all(out.stat().st_mtime >= p.stat().st_mtime for p in deps):
        return out
    BUILD.mkdir(parents=True, exist_ok=True)
    tmp = out.with_suffix(".so.tmp")
    _run([HIPCC, *HIP_FLAGS, "-shared", "-I", str(INCLUDE), "-I", str(CSRC), *map(str, srcs),
          "-o", str(tmp)], log=BUILD / "librf_verilated.log")
    os.replace(tmp, out)
    return out


def design_dir(name: str, key: str) -> Path:
    return BUILD / "designs" / f"{name}_{key}"


def build_design(name: str, sources: Dict[str, str], manifest: dict, jobs: int = 8,
                 force: bool = False) -> Path:
    """Compile generated HIP sources (file name -> text) into a design
    library. Returns the library path."""
    h = hashlib.sha1()
    for fn in sorted(sources):
        h.update(fn.encode())
        h.update(sources[fn].encode())
    h.update(" ".join(HIP_FLAGS).encode())
    for hdr in ("rf_ops.h", "rf_stim.h"):
        h.update((CSRC / hdr).read_bytes())
    h.update((INCLUDE / "rf_design.h").read_bytes())
    key = h.hexdigest()[:16]
    d = design_dir(name, key)
    lib = d / f"libV{name}.so"
    if lib.exists() and not force:
        return lib
    d.mkdir(parents=True, exist_ok=True)
    for fn, text in sources.items():
        (d / fn).write_text(text)
    (d / "manifest.json").write_text(json.dumps(manifest, indent=1, default=str))
    emit_makefiles(d, name, [fn for fn in sources if fn.endswith(".hip")])
    objs = []
    from concurrent.futures import ThreadPoolExecutor

    def cc(fn):
        obj = d / (Path(fn).stem + ".o")
        _run([HIPCC, *HIP_FLAGS, "-fvisibility=hidden", "-I", str(INCLUDE), "-I",
              str(CSRC), "-c", str(d / fn), "-o", str(obj)], log=d / (Path(fn).stem + ".log"))
        return obj

    hip_srcs = [fn for fn in sources if fn.endswith(".hip")]
    with ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(cc, hip_srcs))
    tmp = lib.with_suffix(".so.tmp")
    _run([HIPCC, *HIP_FLAGS, "-shared", "-Wl,-Bsymbolic", *map(str, objs), "-o", str(tmp)],
         log=d / "link.log")
    os.replace(tmp, lib)
    for o in objs:
        o.unlink()
    return lib


def build_testbench(lib: Path, manifest: dict) -> Path:
    """Write ``V<top>.h`` (the generated C++ model class) and the host-driven
    testbench ``V<top>_tb.cpp`` beside a built design library, and link the
    testbench against it and the runtime. Returns the executable."""
    from .emit_vtop import emit_testbench, emit_vtop_header
    d = lib.parent
    nm = manifest["name"]
    (d / f"V{nm}.h").write_text(emit_vtop_header(manifest))
    tb = d / f"V{nm}_tb.cpp"
    tb.write_text(emit_testbench(manifest))
    exe = d / f"V{nm}_tb"
    rt = build_runtime()
    rel = os.path.relpath(rt.parent, d)
    _run(["g++", "-O2", "-std=c++17", "-I", str(INCLUDE), "-I", str(CSRC), "-I", str(d), str(tb),
          "-o", str(exe), "-L", str(d), f"-l:{lib.name}", "-L", str(rt.parent), "-lrf_verilated",
          "-Wl,-rpath-link,/opt/rocm/lib",
          "-Wl,-rpath,$ORIGIN", "-Wl,-rpath,$ORIGIN/" + rel], log=d / "tb.log")
    return exe


def emit_makefiles(outdir: Path, top: str, sources: List[str]):
    """Write V<top>.mk (the V3EmitMk output) and verilated.mk (the include)."""
    outdir.mkdir(parents=True, exist_ok=True)
    (outdir / "verilated.mk").write_text(VERILATED_MK)
    objs = " ".join(Path(s).stem + ".o" for s in sources)
    (outdir / f"V{top}.mk").write_text(
        f"# Generated by rtlflow_amd: build V{top} for gfx950\n"
        f"RF_TOP = V{top}\n"
        f"RF_OBJS = {objs}\n"
        f"RF_ROOT ?= {PKG}\n"
        f"include verilated.mk\n")


VERILATED_MK = r"""# verilated.mk -- hipcc rules for generated models (gfx950)
HIPCC ?= /opt/rocm/bin/hipcc
RF_ARCH ?= gfx950
RF_FLAGS = --offload-arch=$(RF_ARCH) -O3 -std=c++17 -fPIC -fvisibility=hidden \
  -Xclang -target-feature -Xclang -packed-fp32-ops \
  -I$(RF_ROOT)/../include -I$(RF_ROOT)/csrc
%.o: %.hip
	$(HIPCC) $(RF_FLAGS) -c $< -o $@
lib$(RF_TOP).so: $(RF_OBJS)
	$(HIPCC) $(RF_FLAGS) -shared -Wl,-Bsymbolic $^ -o $@
# link a user testbench against the model and the runtime
$(RF_TOP)_tb: $(RF_TB) lib$(RF_TOP).so
	$(HIPCC) $(RF_FLAGS) $(RF_TB) -L. -l$(RF_TOP) -L$(RF_ROOT)/_build -lrf_verilated \
	  -Wl,-rpath,'$$ORIGIN' -Wl,-rpath,$(RF_ROOT)/_build -o $@
all: lib$(RF_TOP).so
.PHONY: all
"""
