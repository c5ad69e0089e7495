"""CPU oracle for libmarshal_amd (test infrastructure only).

Loads oracle/liboracle.so (built from oracle/oracle.c by oracle/Makefile) and exposes its
restatements of the reference's algorithms on numpy arrays. Only the test suite, the smoke
check and the benchmark's CPU-baseline leg import this package; the library never does.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "liboracle.so")
_REF_SO = os.path.join(_HERE, "_ref", "libref_plan.so")
_lib = None


def build(ref=False):
    """Compiles the oracle (and, with ref=True and the reference tree present, oracle/_ref)."""
    subprocess.run(["make", "-s", "-C", _HERE, "all"], check=True)
    if ref:
        subprocess.run(["make", "-s", "-C", _HERE, "ref"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        L = ctypes.CDLL(_SO)
        i64, u64, vp, sz, i32 = (ctypes.c_int64, ctypes.c_uint64, ctypes.c_void_p,
                                 ctypes.c_size_t, ctypes.c_int)
        cp = ctypes.c_char_p
        L.orc_next.restype = u64; L.orc_next.argtypes = [u64, u64, u64]
        L.orc_num_cycles.restype = u64; L.orc_num_cycles.argtypes = [i64, i64]
        L.orc_longest_cycle.restype = u64; L.orc_longest_cycle.argtypes = [i64, i64]
        for f, n in (("orc_010", 3), ("orc_100", 3), ("orc_0100", 4)):
            getattr(L, f).restype = None
            getattr(L, f).argtypes = [vp, vp, sz] + [i64] * n
        L.orc_transpose.restype = None; L.orc_transpose.argtypes = [vp, vp, sz, i64, i64]
        L.orc_golden_transpose.restype = None
        L.orc_golden_transpose.argtypes = [vp, vp, sz, i64, i64]
        L.orc_inplace_010_bs.restype = i32; L.orc_inplace_010_bs.argtypes = [vp, sz, i64, i64, i64]
        L.orc_inplace_010_pttwac.restype = i32
        L.orc_inplace_010_pttwac.argtypes = [vp, sz, i64, i64, i64, i32, i32]
        L.orc_inplace_0100_pttwac.restype = i32
        L.orc_inplace_0100_pttwac.argtypes = [vp, sz, i64, i64, i64, i64]
        L.orc_inplace_100_pttwac.restype = i32
        L.orc_inplace_100_pttwac.argtypes = [vp, sz, i64, i64, i64]
        L.orc_inplace_transpose.restype = i32
        L.orc_inplace_transpose.argtypes = [vp, sz, i64, i64, i64, i64, i32, i32, i64, i32, i64,
                                            cp, sz]
        L.orc_padding.restype = None; L.orc_padding.argtypes = [vp, sz, i64, i64, i64]
        L.orc_unpadding.restype = None; L.orc_unpadding.argtypes = [vp, sz, i64, i64, i64]
        L.orc_tile_sizes.restype = i32
        L.orc_tile_sizes.argtypes = [i64, ctypes.POINTER(ctypes.c_int64), i32]
        L.orc_choose_tiles.restype = i32
        L.orc_choose_tiles.argtypes = [i64, i64, i64, i64] + [ctypes.POINTER(ctypes.c_int64)] * 4
        L.orc_par_010_bs.restype = i32; L.orc_par_010_bs.argtypes = [vp, sz, i64, i64, i64, i32]
        L.orc_par_0100_pttwac.restype = i32
        L.orc_par_0100_pttwac.argtypes = [vp, sz, i64, i64, i64, i64, i32]
        L.orc_par_transpose_3.restype = i32
        L.orc_par_transpose_3.argtypes = [vp, sz, i64, i64, i64, i64, i32]
        L.orc_num_cpus.restype = i32; L.orc_num_cpus.argtypes = []
        _lib = L
    return _lib


def ref_lib():
    """The reference's own planner header behind a C driver, or None where it is not built."""
    if not os.path.exists(_REF_SO):
        return None
    L = ctypes.CDLL(_REF_SO)
    L.ref_next.restype = ctypes.c_uint; L.ref_next.argtypes = [ctypes.c_uint, ctypes.c_int, ctypes.c_int]
    L.ref_num_cycles.restype = ctypes.c_uint; L.ref_num_cycles.argtypes = [ctypes.c_int, ctypes.c_int]
    L.ref_tile_sizes.restype = ctypes.c_int
    L.ref_tile_sizes.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.c_int]
    L.ref_feasible.restype = ctypes.c_int; L.ref_feasible.argtypes = [ctypes.c_int] * 5
    return L


def _p(a):
    assert a.flags.c_contiguous
    return ctypes.c_void_p(a.ctypes.data)


def _bytes(a):
    """Views any contiguous array as raw bytes (elements are opaque)."""
    return a.reshape(-1).view(np.uint8)


# ---- index map / statistics -------------------------------------------------------------
def next_index(i, m, n):
    return lib().orc_next(i, m, n)


def num_cycles(m, n):
    return lib().orc_num_cycles(m, n)


def longest_cycle(m, n):
    return lib().orc_longest_cycle(m, n)


# ---- out-of-place references (return new arrays with the input's dtype) --------------------
def op_010(x, A, a, B, es=None):
    es = x.itemsize if es is None else es
    y = np.empty_like(x)
    lib().orc_010(_p(x), _p(y), es, A, a, B)
    return y


def op_100(x, A, B, b, es=None):
    es = x.itemsize if es is None else es
    y = np.empty_like(x)
    lib().orc_100(_p(x), _p(y), es, A, B, b)
    return y


def op_0100(x, A, a, B, b, es=None):
    es = x.itemsize if es is None else es
    y = np.empty_like(x)
    lib().orc_0100(_p(x), _p(y), es, A, a, B, b)
    return y


def op_transpose(x, M, N, es=None):
    es = x.itemsize if es is None else es
    y = np.empty_like(x)
    lib().orc_transpose(_p(x), _p(y), es, M, N)
    return y


# ---- in-place restatements (modify x) ------------------------------------------------------
def inplace_010_bs(x, A, a, B, es=None):
    return lib().orc_inplace_010_bs(_p(x), es or x.itemsize, A, a, B)


def inplace_010_pttwac(x, A, a, B, R=1, P=1, es=None):
    return lib().orc_inplace_010_pttwac(_p(x), es or x.itemsize, A, a, B, R, P)


def inplace_0100_pttwac(x, A, a, B, b, es=None):
    return lib().orc_inplace_0100_pttwac(_p(x), es or x.itemsize, A, a, B, b)


def inplace_transpose(x, A, a, B, b, stages, R=1, max_mem=12288, sp=1, lds_bytes=48 * 1024,
                      es=None):
    route = ctypes.create_string_buffer(128)
    r = lib().orc_inplace_transpose(_p(x), es or x.itemsize, A, a, B, b, R, stages, max_mem, sp,
                                    lds_bytes, route, 128)
    return r, route.value.decode()


def padding(x, es, xs, ys, pad):
    lib().orc_padding(_p(x), es, xs, ys, pad)


def unpadding(x, es, xs, ys, pad):
    lib().orc_unpadding(_p(x), es, xs, ys, pad)


def tile_sizes(n):
    buf = (ctypes.c_int64 * 65536)()
    c = lib().orc_tile_sizes(n, buf, 65536)
    return list(buf[:min(c, 65536)])


def choose_tiles(h, w, min_limit, max_limit):
    vals = [ctypes.c_int64() for _ in range(4)]
    lib().orc_choose_tiles(h, w, min_limit, max_limit, *[ctypes.byref(v) for v in vals])
    return tuple(v.value for v in vals)


def golden_transpose(x, M, N, es=None):
    """In place, single-threaded: the reference test's two-call CPU golden (010 with unit tiles,
    then 100; test/cl_unittest.cc:547-551) through a temporary of the same size."""
    es = es or x.itemsize
    tmp = np.empty_like(x)
    lib().orc_golden_transpose(_p(x), _p(tmp), es, M, N)


def par_010_bs(x, A, a, B, nthreads=0, es=None):
    return lib().orc_par_010_bs(_p(x), es or x.itemsize, A, a, B, nthreads)


def par_0100_pttwac(x, A, a, B, b, nthreads=0, es=None):
    return lib().orc_par_0100_pttwac(_p(x), es or x.itemsize, A, a, B, b, nthreads)


def par_transpose_3(x, A, a, B, b, nthreads=0, es=None):
    return lib().orc_par_transpose_3(_p(x), es or x.itemsize, A, a, B, b, nthreads)


def num_cpus():
    return lib().orc_num_cpus()


# ---- inputs ---------------------------------------------------------------------------------
def splitmix64(seed, count):
    """Deterministic 64-bit words (splitmix64, Steele/Lea/Flood 2014) as a uint64 array."""
    mask = (1 << 64) - 1
    idx = np.arange(1, count + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = (np.uint64(seed) + idx * np.uint64(0x9E3779B97F4A7C15))
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    del mask
    return z


def tagged(count, es, seed=0):
    """`count` elements of `es` bytes as a flat uint8 array of count*es bytes.

    Elements of 4 or more bytes carry their index k in their first 4 (or 8) bytes and seeded
    noise in the rest, so every element is distinct and any misplacement is visible. 1- and
    2-byte elements are seeded noise: a misplaced element survives unnoticed only where its
    value happens to equal the right one, which a systematic index error cannot do everywhere.
    """
    words = (count * es + 7) // 8
    out = splitmix64(seed, words).view(np.uint8)[: count * es].reshape(count, es).copy()
    if es >= 4:
        w = min(es, 8)
        k = np.arange(count, dtype=np.uint64).view(np.uint8).reshape(count, 8)
        out[:, :w] = k[:, :w]
    return out.reshape(-1)
