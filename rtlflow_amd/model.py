"""Host-side model API over the C ABI (ctypes).

Mirrors the reference's generated model surface: ``RTLflow(size_t)`` + ``Vtop``
(src/V3EmitC.cpp:5246-5306, :3510-3792): construct a model for N stimuli, set
inputs per lane, ``eval()``, read outputs per lane; plus the clock-cycle
protocol with on-device stimulus/hashing and the batched work queue.

There is no fallback: if the runtime or the design library is missing, or no
GPU is present, constructing a model raises.
"""
from __future__ import annotations

import ctypes as C
import json
from pathlib import Path
from typing import Dict, List, Optional, Sequence, Tuple, Union

import numpy as np

from .build import runtime_path

MODE_GRAPH = 0
MODE_FUSED = 1
MODE_BITPLANE = 2

_DT = {0: np.uint8, 1: np.uint16, 2: np.uint32, 3: np.uint64}


class RFError(RuntimeError):
    pass


class rf_config(C.Structure):
    _fields_ = [("settle_unroll", C.c_uint32), ("converge_limit", C.c_uint32),
                ("fused_chunk", C.c_uint32), ("bitplane_threads", C.c_uint32),
                ("lane_pad", C.c_uint32), ("bitplane_slots", C.c_uint32),
                ("fused_multiwave", C.c_uint32)]


class rf_stats(C.Structure):
    _fields_ = [("fused_launches", C.c_uint64), ("fused_ms", C.c_double),
                ("graph_launches", C.c_uint64), ("graph_ms", C.c_double),
                ("bitplane_launches", C.c_uint64), ("bitplane_ms", C.c_double)]


class rf_stream(C.Structure):
    """A [cycle][word][stimulus] buffer of 64-bit words (input stream or
    output capture)."""
    _fields_ = [("data", C.c_void_p), ("stim0", C.c_uint64), ("nstim", C.c_uint32),
                ("cycle0", C.c_uint32), ("ncycles", C.c_uint32), ("words", C.c_uint32)]


_rt = None


def runtime() -> C.CDLL:
    """Load librf_verilated.so and declare its entry points."""
    global _rt
    if _rt is not None:
        return _rt
    path = runtime_path()
    if not path.exists():
        raise RFError(f"runtime library {path} is not built; run __graft_entry__.build()")
    lib = C.CDLL(str(path), mode=C.RTLD_GLOBAL)
    P = C.c_void_p
    u32, u64 = C.c_uint32, C.c_uint64
    sig = {
        "rf_abi_version": (u32, []),
        "rf_create": (C.c_int, [P, u32, C.POINTER(rf_config), C.POINTER(P)]),
        "rf_destroy": (C.c_int, [P]),
        "rf_set_port": (C.c_int, [P, u32, u32, u32, P]),
        "rf_get_port": (C.c_int, [P, u32, u32, u32, P]),
        "rf_mem_write": (C.c_int, [P, u32, u32, u32, u32, u32, P]),
        "rf_mem_read": (C.c_int, [P, u32, u32, u32, u32, u32, P]),
        "rf_set_done": (C.c_int, [P, u32, u32, P]),
        "rf_eval": (C.c_int, [P]),
        "rf_begin": (C.c_int, [P, u64, u32]),
        "rf_run_cycles": (C.c_int, [P, u32, u32]),
        "rf_hashes": (C.c_int, [P, u32, u32, P]),
        "rf_submit_queue": (C.c_int, [P, u64, P, u32, P, P]),
        "rf_submit_queue_bitplane": (C.c_int, [P, u64, P, u32, P, P]),
        "rf_submit_queue_static": (C.c_int, [P, u64, P, u32, P, P]),
        "rf_set_mem_images": (C.c_int, [P, u32, u32, u32, P]),
        "rf_stream_words": (C.c_int, [P, C.POINTER(u32), C.POINTER(u32)]),
        "rf_queue_open": (C.c_int, [P, u64, P, u32, P]),
        "rf_queue_step": (C.c_int, [P, u32, u32, C.POINTER(u32)]),
        "rf_queue_hashes": (C.c_int, [P, P, u32]),
        "rf_queue_save_bytes": (C.c_int, [P, C.POINTER(C.c_size_t)]),
        "rf_queue_save": (C.c_int, [P, P, C.c_size_t]),
        "rf_queue_load": (C.c_int, [P, P, C.c_size_t]),
        "rf_set_input_stream": (C.c_int, [P, P, u64, u32, u32, u32]),
        "rf_set_output_capture": (C.c_int, [P, P, u64, u32, u32, u32]),
        "rf_lane_state": (C.c_int, [P, u32, P, C.POINTER(C.c_size_t)]),
        "rf_trace_lane": (C.c_int, [P, u32, u32, u32, P, C.c_size_t]),
        "rf_checkpoint_bytes": (C.c_int, [P, C.POINTER(C.c_size_t)]),
        "rf_checkpoint_save": (C.c_int, [P, P, C.c_size_t]),
        "rf_checkpoint_load": (C.c_int, [P, P, C.c_size_t]),
        "rf_profile_mtasks": (C.c_int, [P, u32, P]),
        "rf_sync": (C.c_int, [P]),
        "rf_flags": (C.c_int, [P, P, u32]),
        "rf_get_stats": (C.c_int, [P, C.POINTER(rf_stats)]),
        "rf_reset_stats": (C.c_int, [P]),
        "rf_lanes": (u32, [P]),
        "rf_lane_stride": (u64, [P]),
        "rf_last_error": (C.c_char_p, []),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _rt = lib
    return lib


def _check(rc: int):
    if rc != 0:
        msg = runtime().rf_last_error().decode(errors="replace")
        raise RFError(f"rf error {rc}: {msg}")


class DesignLib:
    """A compiled design library (libV<top>.so) plus its manifest."""

    def __init__(self, lib_path: Union[str, Path], manifest: Optional[dict] = None):
        self.path = Path(lib_path)
        runtime()
        # RTLD_LOCAL: several design libraries may be loaded in one process
        self.lib = C.CDLL(str(self.path), mode=C.RTLD_LOCAL)
        self.lib.rf_design_get.restype = C.c_void_p
        self.lib.rf_design_get.argtypes = []
        self.desc = self.lib.rf_design_get()
        if manifest is None:
            mp = self.path.parent / "manifest.json"
            manifest = json.loads(mp.read_text())
        self.manifest = manifest
        self.ports = manifest["ports"]
        self.port_index = {p["name"]: i for i, p in enumerate(self.ports)}
        self.mems = manifest.get("mems", [])
        self.mem_index = {m["name"]: i for i, m in enumerate(self.mems)}

    @property
    def name(self) -> str:
        return self.manifest["name"]

    @property
    def nmtasks(self) -> int:
        return int(self.manifest["mtasks"])


class Model:
    """One design instantiated for ``n_lanes`` stimuli on the current GPU."""

    def __init__(self, design: DesignLib, n_lanes: int, settle_unroll: int = 2,
                 converge_limit: int = 100, fused_chunk: int = 512, bitplane_threads: int = 0,
                 lane_pad: int = 0, bitplane_slots: int = 0, fused_multiwave: int = 0):
        self.rt = runtime()
        self.design = design
        self.n = int(n_lanes)
        cfg = rf_config(settle_unroll, converge_limit, fused_chunk, bitplane_threads, lane_pad,
                        bitplane_slots, fused_multiwave)
        h = C.c_void_p()
        _check(self.rt.rf_create(design.desc, self.n, C.byref(cfg), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.rt.rf_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ------------------------------------------------------------ ports
    def _port(self, name: Union[str, int]):
        i = self.design.port_index[name] if isinstance(name, str) else int(name)
        return i, self.design.ports[i]

    def set_port(self, name, values, lane0: int = 0):
        i, p = self._port(name)
        arr = np.ascontiguousarray(values, dtype=_DT[p["cls"]])
        count = arr.shape[0]
        if arr.size != count * p["words"]:
            raise RFError(f"port {p['name']} expects {p['words']} words per lane")
        _check(self.rt.rf_set_port(self.h, i, lane0, count, arr.ctypes.data))

    def get_port(self, name, lane0: int = 0, count: Optional[int] = None) -> np.ndarray:
        i, p = self._port(name)
        count = self.n - lane0 if count is None else count
        out = np.zeros((count, p["words"]), dtype=_DT[p["cls"]])
        _check(self.rt.rf_get_port(self.h, i, lane0, count, out.ctypes.data))
        return out[:, 0] if p["words"] == 1 else out

    def mem_write(self, name, values: np.ndarray, lane0: int = 0, elem0: int = 0):
        mi = self.design.mem_index[name]
        m = self.design.mems[mi]
        arr = np.ascontiguousarray(values, dtype=_DT[m["cls"]])
        count, nelem = arr.shape[0], arr.shape[1]
        _check(self.rt.rf_mem_write(self.h, mi, lane0, count, elem0, nelem, arr.ctypes.data))

    def mem_read(self, name, lane0: int = 0, count: Optional[int] = None, elem0: int = 0,
                 nelem: Optional[int] = None) -> np.ndarray:
        mi = self.design.mem_index[name]
        m = self.design.mems[mi]
        count = self.n - lane0 if count is None else count
        nelem = m["depth"] - elem0 if nelem is None else nelem
        out = np.zeros((count, nelem, m["words"]), dtype=_DT[m["cls"]])
        _check(self.rt.rf_mem_read(self.h, mi, lane0, count, elem0, nelem, out.ctypes.data))
        return out[:, :, 0] if m["words"] == 1 else out

    def set_mem_images(self, name, images: Optional[np.ndarray]):
        mi = self.design.mem_index[name]
        m = self.design.mems[mi]
        if images is None:
            _check(self.rt.rf_set_mem_images(self.h, mi, 0, 0, None))
            return
        arr = np.ascontiguousarray(images, dtype=_DT[m["cls"]])
        _check(self.rt.rf_set_mem_images(self.h, mi, arr.shape[0], arr.shape[1], arr.ctypes.data))

    def set_done(self, done: np.ndarray, lane0: int = 0):
        arr = np.ascontiguousarray(done, dtype=np.uint8)
        _check(self.rt.rf_set_done(self.h, lane0, arr.shape[0], arr.ctypes.data))

    # ------------------------------------------------------- simulation
    def eval(self):
        _check(self.rt.rf_eval(self.h))

    def begin(self, seed: int, stim0: int = 0):
        _check(self.rt.rf_begin(self.h, seed & ((1 << 64) - 1), stim0))

    def run_cycles(self, n: int, mode: int = MODE_FUSED):
        _check(self.rt.rf_run_cycles(self.h, n, mode))

    def hashes(self) -> np.ndarray:
        out = np.zeros(self.n, dtype=np.uint64)
        _check(self.rt.rf_hashes(self.h, 0, self.n, out.ctypes.data))
        return out

    def submit_queue(self, seed: int, budgets: Sequence[int], order=None,
                     mode: int = MODE_FUSED) -> np.ndarray:
        """Per-stimulus hashes of a batched workload: ``mode`` MODE_FUSED runs
        the persistent per-lane kernel with its work queue, MODE_BITPLANE the
        bit-plane kernel over rounds of lanes."""
        b = np.ascontiguousarray(budgets, dtype=np.uint32)
        out = np.zeros(len(b), dtype=np.uint64)
        optr = None
        if order is not None:
            o = np.ascontiguousarray(order, dtype=np.uint32)
            optr = o.ctypes.data
        if mode == MODE_BITPLANE:
            fn = self.rt.rf_submit_queue_bitplane
        elif mode == MODE_FUSED:
            fn = self.rt.rf_submit_queue
        else:
            raise ValueError(f"no batched workload in mode {mode}")
        _check(fn(self.h, seed & ((1 << 64) - 1), b.ctypes.data, len(b), optr, out.ctypes.data))
        return out

    # ------------------------------------------------------- resumable queue
    def queue_open(self, seed: int, budgets: Sequence[int], order=None):
        b = np.ascontiguousarray(budgets, dtype=np.uint32)
        o = None if order is None else np.ascontiguousarray(order, dtype=np.uint32)
        _check(self.rt.rf_queue_open(self.h, seed & ((1 << 64) - 1), b.ctypes.data, len(b),
                                     None if o is None else o.ctypes.data))
        self._qn = len(b)

    def queue_step(self, max_stimuli: int, mode: int = MODE_FUSED) -> int:
        """Run the next ``max_stimuli`` queue positions; returns the position."""
        pos = C.c_uint32(0)
        _check(self.rt.rf_queue_step(self.h, max_stimuli, mode, C.byref(pos)))
        return pos.value

    def queue_hashes(self) -> np.ndarray:
        out = np.zeros(self._qn, dtype=np.uint64)
        _check(self.rt.rf_queue_hashes(self.h, out.ctypes.data, self._qn))
        return out

    def queue_save(self) -> bytes:
        n = C.c_size_t(0)
        _check(self.rt.rf_queue_save_bytes(self.h, C.byref(n)))
        buf = (C.c_uint8 * max(1, n.value))()
        _check(self.rt.rf_queue_save(self.h, buf, n.value))
        return bytes(buf)[:n.value]

    def queue_load(self, blob: bytes):
        buf = (C.c_uint8 * len(blob)).from_buffer_copy(blob)
        _check(self.rt.rf_queue_load(self.h, buf, len(blob)))
        self._qn = int(np.frombuffer(blob[32:36], np.uint32)[0])

    def lane_state(self, lane: int) -> bytes:
        n = C.c_size_t(0)
        _check(self.rt.rf_lane_state(self.h, lane, None, C.byref(n)))
        buf = (C.c_uint8 * n.value)()
        _check(self.rt.rf_lane_state(self.h, lane, buf, C.byref(n)))
        return bytes(buf)

    # ------------------------------------------------------- streams
    def stream_words(self) -> Tuple[int, int]:
        """64-bit words per stimulus-cycle of the input stream and of the
        output capture."""
        a, b = C.c_uint32(0), C.c_uint32(0)
        _check(self.rt.rf_stream_words(self.design.desc, C.byref(a), C.byref(b)))
        return a.value, b.value

    @staticmethod
    def _window(buf, words: int, what: str) -> Tuple[int, int, int]:
        """(pointer, ncycles, nstim) of a device buffer shaped [ncycles,
        words, nstim] (a torch tensor of 64-bit integers on this GPU)."""
        shape = tuple(buf.shape)
        if len(shape) != 3 or shape[1] != words:
            raise RFError(f"{what}: expected shape [cycles, {words}, stimuli], got {list(shape)}")
        if buf.element_size() != 8 or not buf.is_contiguous():
            raise RFError(f"{what}: expected a contiguous tensor of 64-bit words")
        # the model's stream does not wait for torch's: finish the work that
        # fills (or zeroes) the buffer before the model's kernels use it
        # (a buffer in host memory is the runtime's to refuse)
        if buf.is_cuda:
            import torch
            torch.cuda.current_stream(buf.device).synchronize()
        return buf.data_ptr(), shape[0], shape[2]

    def set_input_stream(self, buf, stim0: int = 0, cycle0: int = 0):
        """Inputs from ``buf`` [cycles, in_words, stimuli] (device memory)
        instead of the generator; None restores the generator."""
        if buf is None:
            _check(self.rt.rf_set_input_stream(self.h, None, 0, 0, 0, 0))
            self._in_buf = None
            return
        ptr, nc, ns = self._window(buf, self.stream_words()[0], "input stream")
        _check(self.rt.rf_set_input_stream(self.h, ptr, stim0, ns, cycle0, nc))
        self._in_buf = buf

    def set_output_capture(self, buf, stim0: int = 0, cycle0: int = 0):
        """Store every output word of the covered stimulus-cycles into ``buf``
        [cycles, out_words, stimuli] (device memory); None turns it off."""
        if buf is None:
            _check(self.rt.rf_set_output_capture(self.h, None, 0, 0, 0, 0))
            self._out_buf = None
            return
        ptr, nc, ns = self._window(buf, self.stream_words()[1], "output capture")
        _check(self.rt.rf_set_output_capture(self.h, ptr, stim0, ns, cycle0, nc))
        self._out_buf = buf

    def trace(self, lane: int, ncycles: int, mode: int = MODE_FUSED) -> List[bytes]:
        """Raw state of ``lane`` after each of the next ``ncycles`` cycles."""
        n = C.c_size_t(0)
        _check(self.rt.rf_lane_state(self.h, lane, None, C.byref(n)))
        sb = n.value
        buf = (C.c_uint8 * max(1, sb * ncycles))()
        _check(self.rt.rf_trace_lane(self.h, lane, ncycles, mode, buf, sb * ncycles))
        raw = bytes(buf)
        return [raw[k * sb:(k + 1) * sb] for k in range(ncycles)]

    # ------------------------------------------------------- checkpoints
    def checkpoint_bytes(self) -> int:
        n = C.c_size_t(0)
        _check(self.rt.rf_checkpoint_bytes(self.h, C.byref(n)))
        return n.value

    def checkpoint(self) -> bytes:
        """The model's whole simulation state as a host blob."""
        n = self.checkpoint_bytes()
        buf = (C.c_uint8 * n)()
        _check(self.rt.rf_checkpoint_save(self.h, buf, n))
        return bytes(buf)

    def restore(self, blob: bytes):
        buf = C.create_string_buffer(blob, len(blob))
        _check(self.rt.rf_checkpoint_load(self.h, buf, len(blob)))
        self.sync()

    def checkpoint_to(self, ptr: int, nbytes: int):
        """Save into caller-owned memory at ``ptr`` (host or device)."""
        _check(self.rt.rf_checkpoint_save(self.h, C.c_void_p(ptr), nbytes))

    def restore_from(self, ptr: int, nbytes: int):
        """Queue a restore from caller-owned memory at ``ptr`` (host or
        device) on the model's stream."""
        _check(self.rt.rf_checkpoint_load(self.h, C.c_void_p(ptr), nbytes))

    def profile_mtasks(self, reps: int = 20) -> np.ndarray:
        """Mean kernel time (ms) of every mtask of one eval, each timed alone."""
        out = np.zeros(max(1, self.design.nmtasks), dtype=np.float32)
        _check(self.rt.rf_profile_mtasks(self.h, reps, out.ctypes.data))
        return out[:self.design.nmtasks]

    def flags(self) -> np.ndarray:
        out = np.zeros(8, dtype=np.uint32)
        _check(self.rt.rf_flags(self.h, out.ctypes.data, 8))
        return out

    def stats(self) -> dict:
        s = rf_stats()
        _check(self.rt.rf_get_stats(self.h, C.byref(s)))
        return {"fused_launches": s.fused_launches, "fused_ms": s.fused_ms,
                "graph_launches": s.graph_launches, "graph_ms": s.graph_ms,
                "bitplane_launches": s.bitplane_launches, "bitplane_ms": s.bitplane_ms}

    def reset_stats(self):
        _check(self.rt.rf_reset_stats(self.h))

    def sync(self):
        _check(self.rt.rf_sync(self.h))
