"""Shared test helpers: the oracle (CPU restatement, ctypes), the product
package, the counter-based input generator in numpy, and error metrics."""
from __future__ import annotations

import ctypes
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import __graft_entry__ as graft  # noqa: E402

ORACLE_LIB = os.path.join(ROOT, "oracle", "lib", "liboracle.so")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def golden(name: str) -> dict:
    """A known-answer fixture from tests/golden (JSON data transcribed from
    the reference's tests and report; each file cites its source lines)."""
    with open(os.path.join(GOLDEN, name + ".json")) as f:
        return json.load(f)


_VP = ctypes.c_void_p
_I64 = ctypes.c_int64
_I64P = ctypes.POINTER(ctypes.c_int64)


class CacheLevel(ctypes.Structure):
    _fields_ = [("size", _I64), ("assoc", _I64), ("line", _I64)]


class CacheGeom(ctypes.Structure):
    _fields_ = [("l", CacheLevel * 3)]


class Blocking(ctypes.Structure):
    _fields_ = [(n, _I64) for n in
                "mr nr kc mc nc mtv_kb mtv_mb dot_sec1 dot_sec2 dot_sec3 dot_l3el".split()]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


# Cache geometry of the reference report's laptop (doc/intro.tex:182-210).
DOC_LAPTOP = tuple(tuple(lv) for lv in golden("blocking_doc_laptop")["levels"])


def ensure_built():
    if not graft.built():
        graft.build()


_oracle = None


def oracle():
    global _oracle
    if _oracle is None:
        ensure_built()
        o = ctypes.CDLL(ORACLE_LIB)
        o.oracle_ct_sqrt.restype = ctypes.c_double
        o.oracle_ct_sqrt.argtypes = [ctypes.c_double]
        o.oracle_blocking.restype = ctypes.c_int
        o.oracle_blocking.argtypes = [ctypes.POINTER(CacheGeom), ctypes.c_int, ctypes.c_int,
                                      ctypes.POINTER(Blocking)]
        o.oracle_cache_probe.argtypes = [ctypes.POINTER(CacheGeom)]
        o.oracle_pack_f32.argtypes = [_VP, _I64, _VP, _I64P, _I64, _I64]
        o.oracle_pack_trans_f32.argtypes = [_VP, _I64, _VP, _I64P, _I64, _I64]
        o.oracle_smtm.argtypes = [_VP, _I64P, _I64P, _VP, _I64P, _I64P, _VP, _I64P, _I64P,
                                  ctypes.c_int, ctypes.POINTER(Blocking), ctypes.c_int]
        o.oracle_smtv_first.argtypes = [_VP, _VP, _I64, _I64, _I64P, _VP,
                                        ctypes.POINTER(Blocking), ctypes.c_int]
        o.oracle_smtv_last.argtypes = [_VP, _VP, _I64, _I64, _I64P, _VP, ctypes.c_int]
        o.oracle_sdot.restype = ctypes.c_float
        o.oracle_sdot.argtypes = [_VP, _VP, _I64, ctypes.POINTER(Blocking), ctypes.c_int]
        o.omb_splitmix64_at.restype = ctypes.c_uint64
        o.omb_splitmix64_at.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
        o.omb_rng_u64.restype = ctypes.c_uint64
        o.omb_rng_u64.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64]
        o.omb_rng_fill_uniform.argtypes = [_VP, _I64, ctypes.c_uint64, ctypes.c_uint64, _I64]
        o.oracle_rand100_fill.argtypes = [_VP, _I64]
        o.oracle_srand.argtypes = [ctypes.c_uint]
        o.oracle_souter.argtypes = [_VP, _I64, _I64, ctypes.c_int, _VP, _VP, ctypes.c_int]
        o.oracle_transpose_block.restype = _I64
        o.oracle_transpose_block.argtypes = [_I64]
        o.oracle_stranspose_out.argtypes = [_VP, _I64P, _VP, _I64P, _I64P, _I64]
        o.oracle_stranspose_in.argtypes = [_VP, _I64, _I64]
        _oracle = o
    return _oracle


def pkg():
    ensure_built()
    return graft.load_package()


def _arr(*v):
    return (ctypes.c_int64 * len(v))(*v)


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_VP)


def _elem_strides(a: np.ndarray):
    return tuple(s // a.itemsize for s in a.strides)


def geom(levels) -> CacheGeom:
    g = CacheGeom()
    for i, (size, assoc, line) in enumerate(levels):
        g.l[i] = CacheLevel(size, assoc, line)
    return g


def blocking(levels=None, elem_size=4, last_order_c=0) -> Blocking:
    o = oracle()
    if levels is None:
        g = CacheGeom()
        o.oracle_cache_probe(ctypes.byref(g))
    else:
        g = geom(levels)
    b = Blocking()
    rc = o.oracle_blocking(ctypes.byref(g), elem_size, last_order_c, ctypes.byref(b))
    assert rc == 0
    return b


def host_blocking() -> Blocking:
    try:
        return blocking(None)
    except AssertionError:
        return blocking(DOC_LAPTOP)


# ------------------------------------------------------------ oracle calls --
def oracle_mtm(c: np.ndarray, a: np.ndarray, b: np.ndarray, blk: Blocking | None = None,
               nthreads: int = 0) -> None:
    """C += A*B with the reference algorithm; arrays may be C- or F-ordered."""
    M, K = a.shape
    N = b.shape[1]
    wc, wa, wb = _elem_strides(c), _elem_strides(a), _elem_strides(b)
    last_c = 1 if (wc[1] == 1 and (wc[0] != 1 or N <= 1)) else 0
    oracle().oracle_smtm(_ptr(c), _arr(M, N), _arr(*wc), _ptr(a), _arr(M, K), _arr(*wa),
                         _ptr(b), _arr(K, N), _arr(*wb), last_c,
                         ctypes.byref(blk) if blk is not None else None, nthreads)


def oracle_mtv(y: np.ndarray, a: np.ndarray, x: np.ndarray, blk: Blocking | None = None,
               nthreads: int = 0) -> None:
    """Column-major A: y += A x (first_order); row-major A: y = A x (last_order)."""
    M, N = a.shape
    wa = _elem_strides(a)
    if wa[0] == 1 and not (wa[1] == 1 and N > 1):
        oracle().oracle_smtv_first(_ptr(y), _ptr(a), M, N, _arr(*wa), _ptr(x),
                                   ctypes.byref(blk) if blk is not None else None, nthreads)
    else:
        oracle().oracle_smtv_last(_ptr(y), _ptr(a), M, N, _arr(*wa), _ptr(x), nthreads)


def oracle_vtm(y: np.ndarray, a: np.ndarray, x: np.ndarray, blk: Blocking | None = None,
               nthreads: int = 0) -> None:
    """y (=|+=) A^T x through the mtv helper of the other layout (include/mtv.hpp:
    164-236) on the true transposed view."""
    oracle_mtv(y, a.T, x, blk, nthreads)


def oracle_dot(a: np.ndarray, b: np.ndarray, blk: Blocking | None = None,
               nthreads: int = 0) -> float:
    return float(oracle().oracle_sdot(_ptr(a), _ptr(b), a.size,
                                      ctypes.byref(blk) if blk is not None else None, nthreads))


def oracle_outer(c: np.ndarray, a: np.ndarray, b: np.ndarray, nthreads: int = 0) -> None:
    """C += a b^T through the reference closure's path for C's storage order."""
    M, N = c.shape
    assert c.flags.c_contiguous or c.flags.f_contiguous
    last = 1 if c.flags.c_contiguous and not (c.flags.f_contiguous and M > 1) else 0
    oracle().oracle_souter(_ptr(c), M, N, last, _ptr(a), _ptr(b), nthreads)


def oracle_transpose(c: np.ndarray, a: np.ndarray, block: int = 0) -> None:
    """C = A^T out of place (any strides; c is N x M for an M x N a)."""
    oracle().oracle_stranspose_out(_ptr(c), _arr(*_elem_strides(c)), _ptr(a), _arr(*a.shape),
                                   _arr(*_elem_strides(a)), block)


def oracle_transpose_inplace(a: np.ndarray, block: int = 0) -> None:
    """Square column-major a transposed in place."""
    assert a.shape[0] == a.shape[1] and a.flags.f_contiguous
    oracle().oracle_stranspose_in(_ptr(a), a.shape[0], block)


def rand100(shape, order="C") -> np.ndarray:
    """rand() % 100 from glibc's process stream, filled in storage order."""
    n = int(np.prod(shape))
    flat = np.empty(n, np.float32)
    oracle().oracle_rand100_fill(_ptr(flat), n)
    return flat.reshape(shape, order=order)


# ------------------------------------------------------- input generator --
_G = 0x9E3779B97F4A7C15
_M64 = (1 << 64) - 1


def _mix_int(z: int) -> int:
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
    return z ^ (z >> 31)


def rng_key(seed: int, stream: int) -> int:
    return _mix_int((seed * _G + stream * 0xD1B54A32D192ED03 + 0x632BE59BD9B4E019) & _M64)


def rng_uniform(n: int, seed: int, stream: int, offset: int = 0) -> np.ndarray:
    """Host twin of omb_fill_uniform: U[-1,1), exact in FP32."""
    key = np.uint64(rng_key(seed, stream))
    out = np.empty(n, np.float32)
    chunk = 1 << 24
    with np.errstate(over="ignore"):
        for s in range(0, n, chunk):
            e = min(n, s + chunk)
            z = np.arange(offset + s + 1, offset + e + 1, dtype=np.uint64)
            z *= np.uint64(_G)
            z += key
            z ^= z >> np.uint64(30)
            z *= np.uint64(0xBF58476D1CE4E5B9)
            z ^= z >> np.uint64(27)
            z *= np.uint64(0x94D049BB133111EB)
            z ^= z >> np.uint64(31)
            out[s:e] = (z >> np.uint64(40)).astype(np.float32) * np.float32(2.0 ** -23) - np.float32(1.0)
    return out


def rng_pow2(n: int, seed: int, stream: int) -> np.ndarray:
    """sign * 2^u with u ~ U[-8, 8] from two generator streams, computed in
    FP64 on the host and rounded to FP32."""
    u = rng_uniform(n, seed, 2 * stream + 1000).astype(np.float64) * 8.0
    s = np.where(rng_uniform(n, seed, 2 * stream + 1001) < 0, -1.0, 1.0)
    return (s * np.exp2(u)).astype(np.float32)


# ------------------------------------------------------ launch pre-checks --
def assert_view_in_bounds(t, rows: int, cols: int, rs: int, cs: int) -> None:
    """Before a kernel reads or writes the view (i, j) -> base[i*rs + j*cs],
    check on the host that every element lies inside t's storage."""
    if rows == 0 or cols == 0:
        return
    storage = t.untyped_storage().nbytes() // t.element_size()
    first = t.storage_offset()
    last = first + (rows - 1) * rs + (cols - 1) * cs
    assert rs >= 1 and cs >= 1 and last < storage, (
        f"view {rows}x{cols} strides ({rs},{cs}) offset {first} exceeds storage {storage}")


# ---------------------------------------------------------------- errors --
def rel_errors(c, truth) -> tuple:
    """(max-relative, Frobenius-relative) of c against the FP64 truth."""
    d = np.asarray(c, dtype=np.float64) - truth
    return float(np.abs(d).max() / np.abs(truth).max()), float(
        np.linalg.norm(d) / np.linalg.norm(truth))
