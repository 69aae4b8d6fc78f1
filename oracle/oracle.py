"""ctypes loader for the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module, as the checker or the timed CPU baseline -- never as a product path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "liboracle.so"

_lib = None


def build(force: bool = False) -> Path:
    srcs = [HERE / "oracle_core.c", HERE / "oracle_designs.c", HERE / "oracle.h"]
    if LIB.exists() and not force and all(LIB.stat().st_mtime >= s.stat().st_mtime for s in srcs):
        return LIB
    r = subprocess.run(["make", "-C", str(HERE), "-B" if force else "all"], stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout[-3000:])
    return LIB


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        _lib = C.CDLL(str(LIB))
        u32, u64, P = C.c_uint32, C.c_uint64, C.c_void_p
        _lib.or_splitmix64.restype = u64
        _lib.or_splitmix64.argtypes = [u64]
        _lib.or_stim.restype = u64
        _lib.or_stim.argtypes = [u64, u64, u64, u64]
        _lib.or_fnv.restype = u64
        _lib.or_fnv.argtypes = [u64, P, C.c_size_t]
        _lib.or_netlist_run.restype = C.c_int
        _lib.or_netlist_run.argtypes = [P, C.c_size_t, u64, u32, u32, u32, P, P, C.c_int]
        _lib.or_alu_run.restype = C.c_int
        _lib.or_alu_run.argtypes = [u64, u32, u32, u32, P, C.c_int]
        _lib.or_systolic_run.restype = C.c_int
        _lib.or_systolic_run.argtypes = [u64, u32, u32, u32, P, C.c_int]
        _lib.or_rv32i_run.restype = C.c_int
        _lib.or_rv32i_run.argtypes = [P, u32, u64, u32, u32, P, P, C.c_int]
    return _lib


def splitmix64(x: int) -> int:
    return lib().or_splitmix64(x & (2**64 - 1))


def stim(seed: int, s: int, c: int, k: int) -> int:
    return lib().or_stim(seed & (2**64 - 1), s, c, k)


def fnv(data: bytes, h: int = 0xcbf29ce484222325) -> int:
    buf = C.create_string_buffer(data, len(data))
    return lib().or_fnv(h, buf, len(data))


def _threads(threads):
    return threads if threads else (os.cpu_count() or 1)


def netlist_run(table_bytes: bytes, seed: int, nstim: int, cycles: int = 0, stim0: int = 0,
                budgets=None, threads: int = 0) -> np.ndarray:
    out = np.zeros(nstim, np.uint64)
    buf = C.create_string_buffer(table_bytes, len(table_bytes))
    bptr = None
    if budgets is not None:
        b = np.ascontiguousarray(budgets, np.uint32)
        bptr = b.ctypes.data
    rc = lib().or_netlist_run(buf, len(table_bytes), seed & (2**64 - 1), stim0, nstim, cycles, bptr,
                              out.ctypes.data, _threads(threads))
    if rc:
        raise RuntimeError(f"or_netlist_run failed: {rc}")
    return out


def alu_run(seed: int, nstim: int, cycles: int, stim0: int = 0, threads: int = 0) -> np.ndarray:
    out = np.zeros(nstim, np.uint64)
    rc = lib().or_alu_run(seed & (2**64 - 1), stim0, nstim, cycles, out.ctypes.data, _threads(threads))
    if rc:
        raise RuntimeError(f"or_alu_run failed: {rc}")
    return out


def systolic_run(seed: int, nstim: int, cycles: int, stim0: int = 0, threads: int = 0) -> np.ndarray:
    out = np.zeros(nstim, np.uint64)
    rc = lib().or_systolic_run(seed & (2**64 - 1), stim0, nstim, cycles, out.ctypes.data,
                               _threads(threads))
    if rc:
        raise RuntimeError(f"or_systolic_run failed: {rc}")
    return out


def rv32i_run(programs: np.ndarray, seed: int, budgets, stim0: int = 0, threads: int = 0) -> np.ndarray:
    progs = np.ascontiguousarray(programs, np.uint32)
    b = np.ascontiguousarray(budgets, np.uint32)
    n = progs.shape[0]
    out = np.zeros(n, np.uint64)
    rc = lib().or_rv32i_run(progs.ctypes.data, progs.shape[1], seed & (2**64 - 1), stim0, n,
                            b.ctypes.data, out.ctypes.data, _threads(threads))
    if rc:
        raise RuntimeError(f"or_rv32i_run failed: {rc}")
    return out
