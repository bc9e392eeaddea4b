"""ctypes wrapper of liboracle_cpu.so (test infrastructure only; see oracle/__init__.py)."""

from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle_cpu.so"
_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        _lib = ctypes.CDLL(str(LIB))
        _lib.oracle_dfa_walk.restype = ctypes.c_int
        _lib.oracle_dfa_walk.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                         ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
        _lib.oracle_nfa_walk.restype = ctypes.c_int
        _lib.oracle_nfa_walk.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_uint64,
                                         ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                         ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
    return _lib


def default_threads() -> int:
    return max(1, min(16, os.cpu_count() or 1))


def dfa_walk(delta, accept, start, data, offsets, threads: int | None = None):
    delta = np.ascontiguousarray(delta, dtype=np.uint8)
    accept = np.ascontiguousarray(accept, dtype=np.uint8)
    data = np.ascontiguousarray(data, dtype=np.uint8)
    offsets = np.ascontiguousarray(offsets, dtype=np.int64)
    n = len(offsets) - 1
    final = np.empty(n, np.uint8)
    acc = np.empty(n, np.uint8)
    rc = lib().oracle_dfa_walk(delta.ctypes.data, accept.ctypes.data, delta.shape[0], int(start),
                               data.ctypes.data, offsets.ctypes.data, n, final.ctypes.data, acc.ctypes.data,
                               threads or default_threads())
    if rc:
        raise ValueError(f"oracle_dfa_walk rc={rc}")
    return final, acc


def nfa_walk(succ, byte_to_sym, start, accept, data, offsets, threads: int | None = None):
    succ = np.ascontiguousarray(succ, dtype=np.uint64)
    b2s = np.ascontiguousarray(byte_to_sym, dtype=np.int16)
    data = np.ascontiguousarray(data, dtype=np.uint8)
    offsets = np.ascontiguousarray(offsets, dtype=np.int64)
    n = len(offsets) - 1
    final = np.empty(n, np.uint64)
    acc = np.empty(n, np.uint8)
    rc = lib().oracle_nfa_walk(succ.ctypes.data, succ.shape[0], b2s.ctypes.data, int(start), int(accept),
                               data.ctypes.data, offsets.ctypes.data, n, final.ctypes.data, acc.ctypes.data,
                               threads or default_threads())
    if rc:
        raise ValueError(f"oracle_nfa_walk rc={rc}")
    return final, acc


def pnfa_forward(T, start, accept, symbols, col_scale=None, T_last=None, threads: int | None = None):
    """C restatement of ``oracle.prob.forward_loglik`` (FP64, threaded)."""
    T = np.asarray(T, dtype=np.float64)
    sigma = T.shape[0]
    Tt = np.ascontiguousarray(np.swapaxes(T, 1, 2))
    Tt_last = Tt if T_last is None else np.ascontiguousarray(np.swapaxes(np.asarray(T_last, np.float64), 1, 2))
    scale = None if col_scale is None else np.ascontiguousarray(col_scale, dtype=np.float64)
    s0 = np.asarray(start, dtype=np.float64)
    m = np.abs(s0).max()
    e0 = int(np.frexp(m)[1]) if m > 0 else 0
    s0 = np.ascontiguousarray(np.ldexp(s0, -e0))
    a = np.ascontiguousarray(accept, dtype=np.float64)
    sym = np.ascontiguousarray(symbols, dtype=np.uint8)
    n, L = sym.shape
    ll = np.empty(n, np.float64)
    lm = np.empty(n, np.float64)
    f = lib().oracle_pnfa_forward
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                          ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p,
                                          ctypes.c_void_p, ctypes.c_int]
    rc = f(Tt.ctypes.data, Tt_last.ctypes.data, scale.ctypes.data if scale is not None else None, sigma,
           s0.ctypes.data, e0, a.ctypes.data, sym.ctypes.data if sym.size else None, n, L, ll.ctypes.data,
           lm.ctypes.data, threads or default_threads())
    if rc:
        raise ValueError(f"oracle_pnfa_forward rc={rc}")
    return ll, lm


def gen_symbols(start: int, n: int, seed: int, lo: int = 0x40, span: int = 32) -> np.ndarray:
    """Bytes [start, start + n) of the device-generated synthetic input."""
    out = np.empty(n, np.uint8)
    f = lib().oracle_gen_symbols
    f.restype = None
    f.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32]
    if n:
        f(out.ctypes.data, int(start), int(n), int(seed), int(lo), int(span))
    return out
