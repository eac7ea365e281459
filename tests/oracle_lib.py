"""ctypes binding of the CPU oracle (oracle/liboracle_pcx.so) for the tests.

The oracle is the checker: only tests/, the smoke entry point and bench.py's
cpu_baseline leg load it.  Nothing under lala-pc_amd/ imports this module.
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "liboracle_pcx.so")

INF = 2**31 - 1
MINF = -(2**31)
OPS = {"ADD": 0, "MUL": 1, "TDIV": 2, "FDIV": 3, "CDIV": 4, "EDIV": 5,
       "MIN": 6, "MAX": 7, "EQ": 8, "LEQ": 9}


class Result(ctypes.Structure):
    _fields_ = [("has_changed", ctypes.c_int32), ("failed", ctypes.c_int32),
                ("sweeps", ctypes.c_int64), ("evals", ctypes.c_int64)]


_lib = None


def lib():
    """Load the oracle library, building it with its Makefile if absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_LIB):
            subprocess.check_call(["make", "-s", "-C", ORACLE_DIR])
        L = ctypes.CDLL(ORACLE_LIB)
        P32 = ctypes.POINTER(ctypes.c_int32)
        P64 = ctypes.POINTER(ctypes.c_int64)
        L.pcxo_fixpoint.argtypes = [ctypes.c_int64, P32, P32, P32, P32, P64, ctypes.c_int64,
                                    P32, P32, ctypes.c_int64, ctypes.c_int64, ctypes.POINTER(Result)]
        L.pcxo_fixpoint_batch.argtypes = [ctypes.c_int64, P32, P32, P32, P32, P64, P64, ctypes.c_int64,
                                          P32, P32, ctypes.c_int64, ctypes.c_int64, ctypes.c_int32,
                                          ctypes.POINTER(Result)]
        L.pcxo_sort_table.argtypes = [ctypes.c_int64, P32, P32, P32, P32]
        L.pcxo_tell_embed.argtypes = [ctypes.c_int64, P32, P32, P32, P32]
        L.pcxo_ask.argtypes = [ctypes.c_int32] * 4 + [P32, P32]
        L.pcxo_deduce.argtypes = [ctypes.c_int32] * 4 + [P32, P32, P32]
        for f in ("pcxo_check_shapes",):
            getattr(L, f).argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
        L.pcxo_check_exhaustive.argtypes = [ctypes.c_int] * 4 + [ctypes.c_char_p, ctypes.c_int]
        L.pcxo_check_singletons.argtypes = [ctypes.c_int] * 3 + [ctypes.c_char_p, ctypes.c_int]
        L.pcxo_check_mul_const.argtypes = [ctypes.c_int] * 2 + [ctypes.c_char_p, ctypes.c_int]
        L.pcxo_check_monotone.argtypes = [ctypes.c_int] * 3 + [ctypes.c_char_p, ctypes.c_int]
        PU64 = ctypes.POINTER(ctypes.c_uint64)
        L.pcxo_bits_tell_embed.argtypes = [ctypes.c_int64, P32, P32, PU64]
        L.pcxo_bits_sort.argtypes = [ctypes.c_int64, P32, P32, P32, P32]
        L.pcxo_bits_fixpoint.argtypes = [ctypes.c_int64, P32, P32, P32, P32, PU64, ctypes.c_int64,
                                         ctypes.c_int32, ctypes.c_int64, ctypes.POINTER(Result)]
        L.pcxo_bits_fixpoint_batch.argtypes = [ctypes.c_int64, P32, P32, P32, P32, PU64, ctypes.c_int64,
                                               ctypes.c_int64, ctypes.c_int32, ctypes.c_int64,
                                               ctypes.c_int32, ctypes.POINTER(Result)]
        L.pcxo_bits_check.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
        _lib = L
    return _lib


def p32(a):
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


def p64(a):
    if a is None:
        return None
    assert a.dtype == np.int64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))


def sort_table(op, x, y, z):
    """Stable sort by (op, y, x, z) in place (reference tell order)."""
    lib().pcxo_sort_table(len(op), p32(op), p32(x), p32(y), p32(z))


def tell_embed(op, x, lb, ub):
    return lib().pcxo_tell_embed(len(op), p32(op), p32(x), p32(lb), p32(ub))


def fixpoint(op, x, y, z, lb, ub, idx=None, max_sweeps=0):
    """Gauss-Seidel fixpoint; mutates lb/ub in place; returns Result."""
    r = Result()
    nidx = 0 if idx is None else len(idx)
    lib().pcxo_fixpoint(len(op), p32(op), p32(x), p32(y), p32(z), p64(idx), nidx,
                        p32(lb), p32(ub), len(lb), max_sweeps, ctypes.byref(r))
    return r


def fixpoint_batch(op, x, y, z, lb, ub, nvars, idx=None, off=None, threads=1, max_sweeps=0):
    """Independent instances over store slices of `nvars`; returns list of Result."""
    n_inst = len(lb) // nvars
    res = (Result * n_inst)()
    lib().pcxo_fixpoint_batch(len(op), p32(op), p32(x), p32(y), p32(z), p64(idx), p64(off),
                              n_inst, p32(lb), p32(ub), nvars, max_sweeps, threads, res)
    return list(res)


def ask(op, x, y, z, lb, ub):
    return bool(lib().pcxo_ask(int(op), int(x), int(y), int(z), p32(lb), p32(ub)))


BITS_OPS = {"LEQ": 0, "ADD": 1, "NEQ": 2, "EQR": 3}


def pu64(a):
    assert a.dtype == np.uint64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))


def bits_tell_embed(op, x, words):
    return lib().pcxo_bits_tell_embed(len(op), p32(op), p32(x), pu64(words))


def bits_sort(op, x, y, z):
    """Stable sort by bitset opcode (PC's formula-kind order)."""
    assert lib().pcxo_bits_sort(len(op), p32(op), p32(x), p32(y), p32(z)) == 0


def bits_fixpoint(op, x, y, z, words, faithful=False, max_sweeps=0):
    r = Result()
    lib().pcxo_bits_fixpoint(len(op), p32(op), p32(x), p32(y), p32(z), pu64(words), len(words),
                             int(faithful), max_sweeps, ctypes.byref(r))
    return r


def bits_fixpoint_batch(op, x, y, z, words, nvars, faithful=False, threads=1, max_sweeps=0):
    n_inst = len(words) // nvars
    res = (Result * n_inst)()
    lib().pcxo_bits_fixpoint_batch(len(op), p32(op), p32(x), p32(y), p32(z), pu64(words), nvars,
                                   n_inst, int(faithful), max_sweeps, threads, res)
    return list(res)


def word(values):
    w = 0
    for v in values:
        w |= 1 << v
    return np.uint64(w)


def values(w):
    w = int(w)
    return [v for v in range(64) if w >> v & 1]
