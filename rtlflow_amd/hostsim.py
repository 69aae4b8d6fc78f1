"""One-lane host build of a generated design (test utility).

Compiles the emitter's host-function variant of a design (identical kernel
bodies, one call per lane) with g++ together with csrc/rf_host.cpp, and runs
the clock-cycle protocol on the CPU. Used by CPU tests to check the emitter
against the oracle without a GPU, and for the mtask-order race test. It is not
a fallback for the GPU path.
"""
from __future__ import annotations

import ctypes as C
import hashlib
import os
import subprocess
from pathlib import Path
from typing import Optional

import numpy as np

from .build import BUILD, CSRC, INCLUDE, BuildError
from .compiler import Compiled, emit_source


class HostDesign:
    def __init__(self, comp: Compiled, opt: str = "-O1"):
        src = emit_source(comp, host=True)
        h = hashlib.sha1(src.encode() + opt.encode())
        for f in ("rf_host.cpp", "rf_ops.h", "rf_stim.h"):
            h.update((CSRC / f).read_bytes())
        h.update((INCLUDE / "rf_design.h").read_bytes())
        key = h.hexdigest()[:16]
        d = BUILD / "host" / f"{comp.design.name}_{key}"
        lib = d / f"libH{comp.design.name}.so"
        if not lib.exists():
            d.mkdir(parents=True, exist_ok=True)
            (d / "gen_host.cpp").write_text(src)
            cmd = ["g++", opt, "-std=c++17", "-fPIC", "-shared", "-w", "-fvisibility=hidden",
                   "-Wl,-Bsymbolic", "-I", str(INCLUDE), "-I",
                   str(CSRC), str(d / "gen_host.cpp"), str(CSRC / "rf_host.cpp"), "-o",
                   str(lib) + ".tmp"]
            r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
            if r.returncode != 0:
                raise BuildError("host build failed:\n" + r.stdout[-4000:])
            os.replace(str(lib) + ".tmp", lib)
        self.lib = C.CDLL(str(lib), mode=C.RTLD_LOCAL)
        self.lib.rf_design_get.restype = C.c_void_p
        self.desc = self.lib.rf_design_get()
        u32, u64, P = C.c_uint32, C.c_uint64, C.c_void_p
        self.lib.rf_host_protocol.restype = C.c_int
        self.lib.rf_host_protocol.argtypes = [P, u32, u64, u32, u32, u32, u64, P, P]
        self.lib.rf_host_queue.restype = C.c_int
        self.lib.rf_host_queue.argtypes = [P, u32, u64, P, u32, P, P, P]
        self.lib.rf_host_streams.restype = None
        self.lib.rf_host_streams.argtypes = [P, P]
        self._streams = None
        self.comp = comp
        self.mems = comp.mem_table()

    def set_streams(self, inputs=None, capture=None, in_stim0: int = 0, in_cycle0: int = 0,
                    cap_stim0: int = 0, cap_cycle0: int = 0):
        """Input stream / output capture (numpy uint64 [cycles, words,
        stimuli], host memory) for the runs that follow; None: generator / no
        capture."""
        from .model import rf_stream
        nin = self.comp.stream_table()["in_words"]
        nout = self.comp.stream_table()["out_words"]

        def mk(arr, s0, c0, words):
            if arr is None:
                return None
            assert arr.dtype == np.uint64 and arr.flags.c_contiguous and arr.shape[1] == words
            return rf_stream(arr.ctypes.data, s0, arr.shape[2], c0, arr.shape[0], words)

        si = mk(inputs, in_stim0, in_cycle0, nin)
        so = mk(capture, cap_stim0, cap_cycle0, nout)
        self._streams = (si, so, inputs, capture)
        self.lib.rf_host_streams(None if si is None else C.addressof(si),
                                 None if so is None else C.addressof(so))

    def _images(self, images):
        """{memory name: array [nstim, nelem]} -> (table, keep-alive list)."""
        if not images:
            return None, []
        dt = {0: np.uint8, 1: np.uint16, 2: np.uint32, 3: np.uint64}
        table = np.zeros(2 * max(1, len(self.mems)), np.uint64)
        keep = []
        for j, m in enumerate(self.mems):
            if m["name"] in images:
                arr = np.ascontiguousarray(images[m["name"]], dtype=dt[m["cls"]])
                keep.append(arr)
                table[2 * j] = arr.ctypes.data
                table[2 * j + 1] = (arr.shape[0] << 32) | arr.shape[1]
        keep.append(table)
        return table, keep

    def protocol(self, nlanes: int, seed: int, ncycles: int, mode: int = 0, stim0: int = 0,
                 order_seed: int = 0, images=None) -> np.ndarray:
        out = np.zeros(nlanes, np.uint64)
        table, keep = self._images(images)
        rc = self.lib.rf_host_protocol(self.desc, nlanes, seed & (2**64 - 1), stim0, ncycles, mode,
                                       order_seed, out.ctypes.data,
                                       None if table is None else table.ctypes.data)
        if rc:
            raise RuntimeError(f"host protocol failed: {rc}")
        return out

    def lane_state(self, lane: int, nlanes: int, seed: int, ncycles: int, mode: int = 0,
                   images=None) -> dict:
        """Run the protocol and return one lane's stored variables by name."""
        from .statediff import decode_state, state_bytes
        buf = np.zeros(state_bytes(self.comp.manifest()), np.uint8)
        self.lib.rf_host_capture.argtypes = [C.c_uint32, C.c_void_p]
        self.lib.rf_host_capture(lane, buf.ctypes.data)
        try:
            self.protocol(nlanes, seed, ncycles, mode=mode, images=images)
        finally:
            self.lib.rf_host_capture(0xffffffff, None)
        return decode_state(self.comp.manifest(), buf.tobytes())

    def queue(self, nlanes: int, seed: int, budgets, order=None, images=None) -> np.ndarray:
        b = np.ascontiguousarray(budgets, np.uint32)
        out = np.zeros(len(b), np.uint64)
        optr = None
        if order is not None:
            o = np.ascontiguousarray(order, np.uint32)
            optr = o.ctypes.data
        table, keep = self._images(images)
        rc = self.lib.rf_host_queue(self.desc, nlanes, seed & (2**64 - 1), b.ctypes.data, len(b), optr,
                                    out.ctypes.data, None if table is None else table.ctypes.data)
        if rc:
            raise RuntimeError(f"host queue failed: {rc}")
        return out
