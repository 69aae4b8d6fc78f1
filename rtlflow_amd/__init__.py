"""rtlflow_amd: MI355X-native batch-stimulus RTL simulation.

Front end (Verilog subset) -> ordering / mtask partition -> gfx950 HIP
kernels + design descriptor -> HIP runtime (librf_verilated.so).

    from rtlflow_amd import verilate
    design = verilate(open("alu.v").read(), threads=4)      # compile + build
    with design.model(n_lanes=1024) as m:                    # needs a GPU
        m.begin(seed=1); m.run_cycles(100); h = m.hashes()
"""
from __future__ import annotations

from pathlib import Path
from typing import Dict, Optional, Tuple

__all__ = ["verilate", "compile_design", "CompiledDesign"]


class CompiledDesign:
    """A built design library and its manifest (loadable without a GPU)."""

    def __init__(self, lib_path: Path, manifest: dict, compiled=None):
        self.lib_path = Path(lib_path)
        self.manifest = manifest
        self.compiled = compiled
        self._lib = None

    @property
    def name(self) -> str:
        return self.manifest["name"]

    def load(self):
        from .model import DesignLib
        if self._lib is None:
            self._lib = DesignLib(self.lib_path, self.manifest)
        return self._lib

    def model(self, n_lanes: int, **kw):
        from .model import Model
        return Model(self.load(), n_lanes, **kw)


def compile_design(src: str, top: Optional[str] = None, params: Optional[Dict[str, int]] = None,
                   name: Optional[str] = None, jobs: int = 8, force: bool = False,
                   **opts) -> CompiledDesign:
    """Verilog text -> built design library (the rtlflow compiler + make).
    A library whose generated source is unchanged is reused unless
    ``force``."""
    from .build import build_design
    from .compiler import Compiled, Options, emit_source
    from .verilog import elaborate

    design = elaborate(src, top, params)
    comp = Compiled(design, Options(**opts))
    text = emit_source(comp)
    man = comp.manifest()
    nm = name or design.name
    lib = build_design(nm, {f"V{nm}.hip": text}, man, jobs=jobs, force=force)
    return CompiledDesign(lib, man, comp)


def verilate(src: str, **kw) -> CompiledDesign:
    return compile_design(src, **kw)
