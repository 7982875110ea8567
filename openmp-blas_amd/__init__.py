"""Host-side Python mirror of the reference's tensor-level BLAS calls, over
the gfx950 C-ABI library ``lib/libompblas.so`` (ctypes).

The reference's callers write ``amt::mtm(c, a, b, nthreads)()`` --
the call validates shapes and returns a closure that runs the product
(include/mtm.hpp:208-267, include/mtv.hpp:100-236, include/dot.hpp:176-224).
This module keeps those names, argument meanings and error behaviour
(``RuntimeError`` on shape mismatch when the closure is created) for torch
tensors that live on the GPU:

    mtm(c, a, b)()        C += A*B            (any row/column-major mix)
    mtv(c, a, b)()        column-major A: c += A*b ; row-major A: c = A*b
    vtm(c, a, b)()        c = A^T b (row-major A: accumulate, col-major: overwrite)
    dot_prod(c, a, b)()   c[0] = sum a*b
    outer_prod(c, a, b)() C += a b^T
    transpose(c, a)()     C = A^T   (transpose(a)() in place, square A)

A tensor's layout is read from its strides: row-major ("last_order") when
the column stride is 1, column-major ("first_order") when the row stride is 1.
Every call runs the HIP kernels; there is no CPU fallback -- loading fails
loudly when the library has not been built.
"""
from __future__ import annotations

import ctypes
import os
import re
from typing import Optional

_HERE = os.path.dirname(os.path.abspath(__file__))
# OMB_LIBRARY names another build of the same library (e.g. one whose host
# code is compiled with sanitizers); by default the in-tree build is loaded.
LIB_PATH = os.environ.get("OMB_LIBRARY") or os.path.join(_HERE, "lib", "libompblas.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "ompblas.h")

OK = 0
EINVAL = -1
EUNSUPPORTED = -2
ENOMEM = -3
EDEVICE = -4
EBUSY = -5

PLAIN_F32 = 0
SPLIT_BF16 = 1
SPLIT_BF16_P3 = 2
SPLIT_FP16 = 3       # rows of A and columns of B scaled by powers of two, 2 FP16 terms, 3 MFMA products

REQ_GEMM, REQ_GEMV, REQ_DOT = 0, 1, 2

_lib = None


class OmbError(RuntimeError):
    """A non-OK status from the C-ABI, with the library's message."""

    def __init__(self, status: int, msg: str):
        super().__init__(f"ompblas status {status}: {msg}")
        self.status = status


class GemmDesc(ctypes.Structure):
    _fields_ = [("M", ctypes.c_int64), ("N", ctypes.c_int64), ("K", ctypes.c_int64),
                ("A", ctypes.c_void_p), ("lda", ctypes.c_int64),
                ("B", ctypes.c_void_p), ("ldb", ctypes.c_int64),
                ("C", ctypes.c_void_p), ("ldc", ctypes.c_int64),
                ("alpha", ctypes.c_float), ("beta", ctypes.c_float)]


class GemvDesc(ctypes.Structure):
    _fields_ = [("M", ctypes.c_int64), ("N", ctypes.c_int64),
                ("A", ctypes.c_void_p), ("lda", ctypes.c_int64),
                ("x", ctypes.c_void_p), ("y", ctypes.c_void_p),
                ("alpha", ctypes.c_float), ("beta", ctypes.c_float)]


class Request(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("mode", ctypes.c_int32),
                ("m", ctypes.c_int64), ("n", ctypes.c_int64), ("k", ctypes.c_int64),
                ("arrival_s", ctypes.c_double),
                ("a", ctypes.c_void_p), ("b", ctypes.c_void_p), ("c", ctypes.c_void_p),
                ("ws_bytes", ctypes.c_int64)]


class ServeConfig(ctypes.Structure):
    _fields_ = [("slots", ctypes.c_int32), ("hbm_budget", ctypes.c_int64),
                ("hbm_reserve", ctypes.c_int64), ("spin_poll", ctypes.c_int32),
                ("no_batching", ctypes.c_int32)]


class RequestResult(ctypes.Structure):
    _fields_ = [("start_s", ctypes.c_double), ("done_s", ctypes.c_double),
                ("latency_s", ctypes.c_double), ("status", ctypes.c_int32),
                ("slot", ctypes.c_int32)]


class ServeStats(ctypes.Structure):
    _fields_ = [("wall_s", ctypes.c_double), ("p50_s", ctypes.c_double),
                ("p90_s", ctypes.c_double), ("p99_s", ctypes.c_double),
                ("max_s", ctypes.c_double), ("mean_s", ctypes.c_double),
                ("total_flop", ctypes.c_double), ("max_in_flight", ctypes.c_int32),
                ("admitted_late", ctypes.c_int32), ("batches", ctypes.c_int32),
                ("batched_requests", ctypes.c_int32)]


_I64, _F32, _VP, _I32, _U64, _F64 = (ctypes.c_int64, ctypes.c_float, ctypes.c_void_p,
                                     ctypes.c_int32, ctypes.c_uint64, ctypes.c_double)

_SIGS = {
    "omb_last_error": (ctypes.c_char_p, []),
    "omb_version": (ctypes.c_char_p, []),
    "omb_init": (_I32, []),
    "omb_sgemm": (_I32, [_I64, _I64, _I64, _F32, _VP, _I64, _I64, _VP, _I64, _I64, _F32, _VP,
                         _I64, _I64, _I32, _VP]),
    "omb_split_pack_bytes": (ctypes.c_size_t, [_I64, _I64]),
    "omb_split_pack": (_I32, [_I64, _I64, _VP, _I64, _I64, _VP, _VP]),
    "omb_sgemm_packed": (_I32, [_I64, _I64, _I64, _F32, _VP, _VP, _F32, _VP, _I64, _I64, _I32,
                                _VP]),
    "omb_sgemv": (_I32, [_I64, _I64, _F32, _VP, _I64, _I64, _VP, _F32, _VP, _VP]),
    "omb_sdot": (_I32, [_I64, _VP, _VP, _VP, _VP]),
    "omb_sdot_noalloc": (_I32, [_I64, _VP, _VP, _VP, _VP]),
    "omb_sger": (_I32, [_I64, _I64, _F32, _VP, _VP, _VP, _I64, _I64, _VP]),
    "omb_stranspose": (_I32, [_I64, _I64, _VP, _I64, _I64, _VP, _I64, _I64, _VP]),
    "omb_stranspose_inplace": (_I32, [_I64, _VP, _I64, _VP]),
    "omb_fill_uniform": (_I32, [_VP, _I64, _U64, _U64, _I64, _VP]),
    "omb_check_dgemm": (_I32, [_I64, _I64, _I64, _VP, _I64, _I64, _VP, _I64, _I64, _VP, _I64,
                               _I64, _VP, _VP]),
    "omb_check_dgemv": (_I32, [_I64, _I64, _VP, _I64, _I64, _VP, _VP, _VP, _VP]),
    "omb_check_ddot": (_I32, [_I64, _VP, _VP, _VP, _VP]),
    "omb_check_errnorm": (_I32, [_I64, _I64, _VP, _I64, _I64, _VP, _VP, _VP]),
    "omb_calib_fmac": (_I32, [_I64, _I64, _VP, _VP]),
    "omb_calib_mfma": (_I32, [_I64, _I64, _I32, _VP, _VP]),
    "omb_group_plan_create": (_I32, [ctypes.POINTER(GemmDesc), _I32, ctypes.POINTER(GemvDesc),
                                     _I32, _I32, _VP, ctypes.POINTER(_VP)]),
    "omb_group_plan_launch": (_I32, [_VP, _VP]),
    "omb_group_plan_launch_direct": (_I32, [_VP, _VP]),
    "omb_group_plan_describe": (_I32, [_VP, ctypes.c_char_p, ctypes.c_size_t]),
    "omb_group_plan_destroy": (_I32, [_VP]),
    "omb_autotune_group": (_I32, [ctypes.POINTER(GemmDesc), _I32, _I32, ctypes.c_char_p, _VP]),
    "omb_autotune_load": (_I32, [ctypes.c_char_p]),
    "omb_serve_request_bytes": (_I64, [ctypes.POINTER(Request)]),
    "omb_serve_run": (_I32, [ctypes.POINTER(Request), _I32, ctypes.POINTER(ServeConfig),
                             ctypes.POINTER(RequestResult), ctypes.POINTER(ServeStats)]),
    "omb_server_open": (_I32, [ctypes.POINTER(ServeConfig), _I64, ctypes.POINTER(_VP)]),
    "omb_server_submit": (_I32, [_VP, ctypes.POINTER(Request), ctypes.POINTER(_I64)]),
    "omb_server_poll": (_I32, [_VP, _I64, ctypes.POINTER(_I32), ctypes.POINTER(RequestResult)]),
    "omb_server_wait": (_I32, [_VP, _I64, ctypes.POINTER(RequestResult)]),
    "omb_server_close": (_I32, [_VP, ctypes.POINTER(ServeStats)]),
    "omb_is_device_pointer": (_I32, [_VP]),
    "omb_device_alloc": (_I32, [ctypes.POINTER(_VP), ctypes.c_size_t]),
    "omb_device_free": (_I32, [_VP]),
    "omb_memcpy_h2d": (_I32, [_VP, _VP, ctypes.c_size_t, _VP]),
    "omb_memcpy_d2h": (_I32, [_VP, _VP, ctypes.c_size_t, _VP]),
    "omb_stream_synchronize": (_I32, [_VP]),
}


def header_functions(path: str = HEADER_PATH) -> list:
    """Names of every function the public header declares."""
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(omb_[a-z0-9_]+)\s*\(", text)))


def lib() -> ctypes.CDLL:
    """The loaded C-ABI library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with "
                               "`python -c 'import __graft_entry__ as g; g.build()'`")
        l = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(l, name)
            f.restype = res
            f.argtypes = args
        _lib = l
    return _lib


def last_error() -> str:
    return lib().omb_last_error().decode()


def check(status: int) -> None:
    if status != OK:
        raise OmbError(status, last_error())


def _stream(stream) -> Optional[int]:
    if stream is None:
        return None
    return getattr(stream, "cuda_stream", stream)


# --------------------------------------------------------------- raw ABI --
def sgemm(M, N, K, alpha, A, rsA, csA, B, rsB, csB, beta, C, rsC, csC, mode=PLAIN_F32,
          stream=None) -> int:
    """Status of omb_sgemm on raw device pointers (ints)."""
    return lib().omb_sgemm(M, N, K, alpha, A, rsA, csA, B, rsB, csB, beta, C, rsC, csC, mode,
                           _stream(stream))


def sgemv(M, N, alpha, A, rsA, csA, x, beta, y, stream=None) -> int:
    return lib().omb_sgemv(M, N, alpha, A, rsA, csA, x, beta, y, _stream(stream))


def sdot(n, x, y, result, stream=None) -> int:
    return lib().omb_sdot(n, x, y, result, _stream(stream))


def sger(M, N, alpha, x, y, C, rsC, csC, stream=None) -> int:
    return lib().omb_sger(M, N, alpha, x, y, C, rsC, csC, _stream(stream))


def stranspose(M, N, A, rsA, csA, C, rsC, csC, stream=None) -> int:
    return lib().omb_stranspose(M, N, A, rsA, csA, C, rsC, csC, _stream(stream))


def stranspose_inplace(n, A, ld, stream=None) -> int:
    return lib().omb_stranspose_inplace(n, A, ld, _stream(stream))


# ------------------------------------------------------ torch conveniences --
def _ptr(t) -> int:
    return t.data_ptr()


def _strides(t):
    if t.dim() != 2:
        raise RuntimeError("expected a matrix")
    return t.stride(0), t.stride(1)


def fill_uniform(t, seed: int, stream_id: int, offset: int = 0, stream=None) -> None:
    """Fill a contiguous float32 device tensor with the counter-based U[-1,1)."""
    check(lib().omb_fill_uniform(_ptr(t), t.numel(), seed, stream_id, offset, _stream(stream)))


def is_matrix(t) -> bool:
    return t.dim() == 2


def is_vector(t) -> bool:
    return t.dim() == 1 or (t.dim() == 2 and (t.shape[0] == 1 or t.shape[1] == 1))


def mtm(c, a, b, num_threads=None, mode: int = PLAIN_F32, stream=None):
    """amt::mtm (include/mtm.hpp:208-267): validate, return a closure computing C += A*B.

    ``num_threads`` is accepted for signature parity and ignored (the GPU
    decides its own parallelism)."""
    if not (is_matrix(a) and is_matrix(b) and is_matrix(c)):
        raise RuntimeError("amt::mtm: a, b, and c must be the matrices")
    if not (a.shape[0] == c.shape[0] and a.shape[1] == b.shape[0] and c.shape[1] == b.shape[1]):
        raise RuntimeError("amt::mtm: dimension mismatch")
    M, K = a.shape
    N = b.shape[1]
    (rsa, csa), (rsb, csb), (rsc, csc) = _strides(a), _strides(b), _strides(c)

    def run():
        check(sgemm(M, N, K, 1.0, _ptr(a), rsa, csa, _ptr(b), rsb, csb, 1.0, _ptr(c), rsc, csc,
                    mode, stream))

    return run


def _vec_len(t) -> int:
    return t.numel()


def _is_first_order(a) -> bool:
    """Column-major (uBLAS first_order) when the row stride is 1."""
    rs, cs = _strides(a)
    if cs == 1 and rs == 1:
        return a.shape[1] <= 1
    return rs == 1


def mtv(c, a, b, num_threads=None, stream=None):
    """amt::mtv (include/mtv.hpp:100-162): c += A*b for column-major A
    (first_order helper, :15-70), c = A*b for row-major A (last_order, :72-98)."""
    if not (is_matrix(a) and is_vector(b) and is_vector(c)):
        raise RuntimeError("amt::mtv: c and b must be vector, and a must be a matrix")
    if a.shape[1] != _vec_len(b) or a.shape[0] != _vec_len(c):
        raise RuntimeError("amt::mtv: dimension mismatch")
    M, N = a.shape
    rs, cs = _strides(a)
    beta = 1.0 if _is_first_order(a) else 0.0

    def run():
        check(sgemv(M, N, 1.0, _ptr(a), rs, cs, _ptr(b), beta, _ptr(c), stream))

    return run


def vtm(c, a, b, num_threads=None, stream=None):
    """amt::vtm (include/mtv.hpp:164-236): c = A^T b.  The reference runs the
    helper of the other layout, so a column-major A overwrites c and a
    row-major A accumulates into c; the product itself is the true A^T b."""
    if not (is_matrix(a) and is_vector(b) and is_vector(c)):
        raise RuntimeError("amt::vtm: c and b must be vector, and a must be a matrix")
    if a.shape[1] != _vec_len(c) or a.shape[0] != _vec_len(b):
        raise RuntimeError("amt::vtm: dimension mismatch")
    M, N = a.shape
    rs, cs = _strides(a)
    beta = 0.0 if _is_first_order(a) else 1.0

    def run():
        # A^T is N x M with strides (cs, rs).
        check(sgemv(N, M, 1.0, _ptr(a), cs, rs, _ptr(b), beta, _ptr(c), stream))

    return run


def dot_prod(c, a, b, num_threads=None, stream=None):
    """amt::dot_prod (include/dot.hpp:176-224): c[0] = sum a*b (overwrite)."""
    if not (is_vector(a) and is_vector(b)):
        raise RuntimeError("amt::dot_prod: both tensor must be vector")
    if _vec_len(a) != _vec_len(b):
        raise RuntimeError("amt::dot_prod: both vector must be of same size")
    n = _vec_len(a)

    def run():
        check(sdot(n, _ptr(a), _ptr(b), _ptr(c), stream))

    return run


def dot_prod_tuning_param(c, a, b, b1, b2, b3, num_threads=None, stream=None):
    """amt::dot_prod_tuning_param (include/dot.hpp:91-140): dot_prod with the
    CPU cache block sizes b1, b2, b3, which the GPU kernel does not use (its
    grid depends only on the length)."""
    return dot_prod(c, a, b, num_threads, stream)


def outer_prod(c, a, b, num_threads=None, stream=None):
    """amt::outer_prod (include/outer.hpp:51-122): C += a b^T for vectors a
    (length M) and b (length N) and an M x N matrix C of either layout."""
    if not (is_vector(a) and is_vector(b)):
        raise RuntimeError("amt::outer_prod: both tensor must be vector")
    M, N = _vec_len(a), _vec_len(b)
    if not is_matrix(c) or c.shape[0] != M or c.shape[1] != N:
        raise RuntimeError("amt::outer_prod: dimension mismatch")
    rs, cs = _strides(c)

    def run():
        check(sger(M, N, 1.0, _ptr(a), _ptr(b), _ptr(c), rs, cs, stream))

    return run


def transpose(c, a=None, num_threads=None, stream=None):
    """amt::transpose (include/trans.hpp:82-166).  transpose(c, a)(): C = A^T
    out of place (C is N x M for an M x N A, any layouts); transpose(a)(): the
    square matrix a is transposed in place."""
    if a is None:
        a = c
        if not is_matrix(a):
            raise RuntimeError("amt::transpose: a must be a matrix")
        if a.shape[0] != a.shape[1]:
            raise RuntimeError("amt::transpose: in-place transpose needs a square matrix")
        rs, cs = _strides(a)
        if min(rs, cs) != 1:
            raise RuntimeError("amt::transpose: one stride must be 1")
        n, ld = a.shape[0], max(rs, cs)

        def run_inplace():
            check(stranspose_inplace(n, _ptr(a), ld, stream))

        return run_inplace
    if not (is_matrix(a) and is_matrix(c)):
        raise RuntimeError("amt::transpose: a and c must be the matrices")
    if a.shape[0] != c.shape[1] or a.shape[1] != c.shape[0]:
        raise RuntimeError("amt::transpose: dimension mismatch")
    M, N = a.shape
    (rsA, csA), (rsC, csC) = _strides(a), _strides(c)

    def run():
        check(stranspose(M, N, _ptr(a), rsA, csA, _ptr(c), rsC, csC, stream))

    return run


# ------------------------------------------------------------ FP64 checks --
def check_dgemm(a, b, c0=None, stream=None):
    """FP64 product of FP32 device matrices (c0 + A*B) as a new float64 tensor."""
    import torch
    M, K = a.shape
    N = b.shape[1]
    out = torch.empty((M, N), dtype=torch.float64, device=a.device)
    rsc, csc = _strides(c0) if c0 is not None else (N, 1)
    check(lib().omb_check_dgemm(M, N, K, _ptr(a), *_strides(a), _ptr(b), *_strides(b),
                                _ptr(c0) if c0 is not None else None, rsc, csc, _ptr(out),
                                _stream(stream)))
    return out


def errnorm(c, c64, stream=None):
    """(max|C-C64|, max|C64|, ||C-C64||_F^2, ||C64||_F^2) as Python floats."""
    import torch
    if c.dim() == 1:
        c = c.view(-1, 1)
        c64 = c64.view(-1, 1)
    out = torch.empty(4, dtype=torch.float64, device=c.device)
    M, N = c.shape
    check(lib().omb_check_errnorm(M, N, _ptr(c), *_strides(c), _ptr(c64), _ptr(out),
                                  _stream(stream)))
    check(lib().omb_stream_synchronize(_stream(stream)))
    return tuple(out.cpu().tolist())


def rel_errors(c, c64, stream=None):
    """(max-relative, Frobenius-relative) error of c against c64."""
    dmax, rmax, d2, r2 = errnorm(c, c64, stream)
    return dmax / rmax if rmax > 0 else dmax, (d2 / r2) ** 0.5 if r2 > 0 else d2 ** 0.5


# --------------------------------------------------------- grouped plans --
class GroupPlan:
    """A captured grouped launch (omb_group_plan_*): GEMMs given as tuples
    (M, N, K, A, lda, B, ldb, C, ldc, alpha, beta) and GEMVs as
    (M, N, A, lda, x, y, alpha, beta) of device pointers (row-major)."""

    def __init__(self, gemms, gemvs, mode: int = PLAIN_F32, stream=None):
        g = (GemmDesc * max(1, len(gemms)))(*[GemmDesc(*t) for t in gemms])
        v = (GemvDesc * max(1, len(gemvs)))(*[GemvDesc(*t) for t in gemvs])
        self._h = ctypes.c_void_p()
        check(lib().omb_group_plan_create(g, len(gemms), v, len(gemvs), mode, _stream(stream),
                                          ctypes.byref(self._h)))

    def launch(self, stream=None) -> None:
        check(lib().omb_group_plan_launch(self._h, _stream(stream)))

    def launch_direct(self, stream=None) -> None:
        check(lib().omb_group_plan_launch_direct(self._h, _stream(stream)))

    def describe(self) -> str:
        buf = ctypes.create_string_buffer(4096)
        check(lib().omb_group_plan_describe(self._h, buf, len(buf)))
        return buf.value.decode()

    def close(self) -> None:
        if self._h:
            lib().omb_group_plan_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def autotune_group(gemms, mode: int = PLAIN_F32, table_path: Optional[str] = None,
                   stream=None) -> None:
    """Time the tile classes on each (log2 M, log2 N, log2 K) bucket of
    `gemms` (tuples as for GroupPlan) and keep the fastest per bucket."""
    g = (GemmDesc * max(1, len(gemms)))(*[GemmDesc(*t) for t in gemms])
    check(lib().omb_autotune_group(g, len(gemms), mode,
                                   table_path.encode() if table_path else None, _stream(stream)))


def autotune_load(table_path: str) -> None:
    check(lib().omb_autotune_load(table_path.encode()))


# ---------------------------------------------------------- serving loop --
def serve_run(requests, slots: int = 8, hbm_budget: int = 0, hbm_reserve: int = 1 << 30,
              spin_poll: bool = True, batching: bool = True):
    """Replay `requests` (a list of Request) through omb_serve_run.

    Returns (results, stats): one dict per request (start_s, done_s,
    latency_s, status, slot, ws_bytes) and a dict of the run's statistics.
    A request whose workspace can never fit the budget comes back with
    status EBUSY; any other failure raises OmbError."""
    n = len(requests)
    reqs = (Request * max(1, n))(*requests)
    res = (RequestResult * max(1, n))()
    st = ServeStats()
    cfg = ServeConfig(slots, hbm_budget, hbm_reserve, 1 if spin_poll else 0,
                      0 if batching else 1)
    check(lib().omb_serve_run(reqs, n, ctypes.byref(cfg), res, ctypes.byref(st)))
    out = []
    for i in range(n):
        r = res[i]
        out.append(dict(start_s=r.start_s, done_s=r.done_s, latency_s=r.latency_s,
                        status=r.status, slot=r.slot, ws_bytes=reqs[i].ws_bytes))
    stats = {name: getattr(st, name) for name, _ in ServeStats._fields_}
    return out, stats


def _result_dict(r) -> dict:
    return dict(start_s=r.start_s, done_s=r.done_s, latency_s=r.latency_s, status=r.status,
                slot=r.slot)


class Server:
    """Live submission through omb_server_*: requests are admitted as they
    are submitted (the slots, arena and batching of serve_run) by a worker
    thread in the library.  submit() returns a ticket; wait(ticket) blocks
    for its result dict; close() waits for every request and returns the
    statistics.  Operands must be complete on the device at submission."""

    def __init__(self, slots: int = 8, max_request_bytes: int = 64 << 20, hbm_budget: int = 0,
                 hbm_reserve: int = 1 << 30, spin_poll: bool = True, batching: bool = True):
        cfg = ServeConfig(slots, hbm_budget, hbm_reserve, 1 if spin_poll else 0,
                          0 if batching else 1)
        self._h = ctypes.c_void_p()
        check(lib().omb_server_open(ctypes.byref(cfg), max_request_bytes, ctypes.byref(self._h)))

    def submit(self, request) -> int:
        t = _I64()
        check(lib().omb_server_submit(self._h, ctypes.byref(request), ctypes.byref(t)))
        return int(t.value)

    def poll(self, ticket: int) -> Optional[dict]:
        done, r = _I32(), RequestResult()
        check(lib().omb_server_poll(self._h, ticket, ctypes.byref(done), ctypes.byref(r)))
        return _result_dict(r) if done.value else None

    def wait(self, ticket: int) -> dict:
        r = RequestResult()
        check(lib().omb_server_wait(self._h, ticket, ctypes.byref(r)))
        return _result_dict(r)

    def close(self) -> dict:
        st = ServeStats()
        if self._h:
            h, self._h = self._h, ctypes.c_void_p()
            check(lib().omb_server_close(h, ctypes.byref(st)))
        return {name: getattr(st, name) for name, _ in ServeStats._fields_}

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def serve_request_bytes(request) -> int:
    """Workspace bytes omb_serve_run holds for `request` while it runs."""
    v = lib().omb_serve_request_bytes(ctypes.byref(request))
    if v < 0:
        raise OmbError(EINVAL, last_error())
    return int(v)


def log_uniform_ints(rng, lo: int, hi: int, size):
    """Integers log-uniform on [lo, hi]: floor(exp(U[ln lo, ln(hi + 1))))."""
    import numpy as np
    v = np.floor(np.exp(rng.uniform(np.log(lo), np.log(hi + 1), size))).astype(np.int64)
    return np.clip(v, lo, hi)


def serving_trace(count: int, rate_hz: float, seed: int = 4, mix=(0.5, 0.3, 0.2),
                  gemm=(32, 4096), gemv=(256, 16384), dot=(1 << 10, 1 << 24)):
    """Poisson arrivals at `rate_hz` with the GEMM / GEMV / dot mix: GEMM
    M, N, K, GEMV rows and cols, and dot length each log-uniform on the given
    ranges.  Returns a list of (kind, m, n, k, arrival_s)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate_hz, count)
    arrival = np.cumsum(gaps) - gaps[0]
    kinds = rng.choice(3, size=count, p=np.asarray(mix, dtype=np.float64) / sum(mix))
    g = log_uniform_ints(rng, gemm[0], gemm[1], (count, 3))
    v = log_uniform_ints(rng, gemv[0], gemv[1], (count, 2))
    d = log_uniform_ints(rng, dot[0], dot[1], count)
    out = []
    for i in range(count):
        k = int(kinds[i])
        if k == REQ_GEMM:
            out.append((k, int(g[i, 0]), int(g[i, 1]), int(g[i, 2]), float(arrival[i])))
        elif k == REQ_GEMV:
            out.append((k, int(v[i, 0]), int(v[i, 1]), 0, float(arrival[i])))
        else:
            out.append((k, 0, int(d[i]), 0, float(arrival[i])))
    return out


def request_flop(kind: int, m: int, n: int, k: int) -> float:
    return 2.0 * m * n * k if kind == REQ_GEMM else (2.0 * m * n if kind == REQ_GEMV else 2.0 * n)


def request_bytes(kind: int, m: int, n: int, k: int) -> float:
    if kind == REQ_GEMM:
        return 4.0 * (m * k + k * n + m * n)
    if kind == REQ_GEMV:
        return 4.0 * (m * n + m + n)
    return 8.0 * n
