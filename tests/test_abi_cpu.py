"""The C-ABI library loads, exports every declared symbol, and rejects bad
arguments before touching the device.  CPU only: no call here reaches a
kernel launch."""
import ctypes

import pytest

import helpers


def test_library_exports_every_header_symbol(ompblas):
    lib = ompblas.lib()
    names = ompblas.header_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert missing == []


def test_every_header_symbol_has_a_ctypes_signature(ompblas):
    assert sorted(ompblas.header_functions()) == sorted(ompblas._SIGS)


def test_version(ompblas):
    assert ompblas.lib().omb_version().decode().startswith("ompblas")


def _sgemm(ompblas, **kw):
    args = dict(M=4, N=4, K=4, alpha=1.0, A=16, rsA=4, csA=1, B=16, rsB=4, csB=1, beta=1.0,
                C=16, rsC=4, csC=1, mode=0)
    args.update(kw)
    return ompblas.sgemm(**args)


@pytest.mark.parametrize("kw,status", [
    (dict(M=-1), -1), (dict(N=-2), -1), (dict(K=-3), -1),
    (dict(C=None), -1), (dict(A=None), -1), (dict(B=None), -1),
    (dict(rsA=3), -1),                   # ld < K for a row-major A
    (dict(rsC=2, csC=2), -2),            # neither stride is 1
    (dict(rsB=0), -1),                   # stride 0
    (dict(mode=7), -1),
])
def test_sgemm_rejects_bad_arguments(ompblas, kw, status):
    assert _sgemm(ompblas, **kw) == status
    assert ompblas.last_error() != ""


def test_sgemm_empty_is_ok_without_device(ompblas):
    assert _sgemm(ompblas, M=0) == 0
    assert _sgemm(ompblas, N=0, C=None) == 0
    assert ompblas.last_error() == ""


def test_sgemv_and_dot_reject_bad_arguments(ompblas):
    assert ompblas.sgemv(-1, 4, 1.0, 16, 4, 1, 16, 0.0, 16) == -1
    assert ompblas.sgemv(4, 4, 1.0, None, 4, 1, 16, 0.0, 16) == -1
    assert ompblas.sgemv(4, 4, 1.0, 16, 4, 1, 16, 0.0, None) == -1
    assert ompblas.sgemv(4, 4, 1.0, 16, 2, 1, 16, 0.0, 16) == -1     # ld < N
    assert ompblas.sgemv(0, 4, 1.0, None, 4, 1, None, 0.0, None) == 0
    assert ompblas.sdot(-1, 16, 16, 16) == -1
    assert ompblas.sdot(4, 16, 16, None) == -1
    assert ompblas.sdot(4, None, 16, 16) == -1
    assert "omb_sdot" in ompblas.last_error()


def test_check_and_fill_reject_bad_arguments(ompblas):
    lib = ompblas.lib()
    assert lib.omb_fill_uniform(None, 10, 0, 0, 0, None) == -1
    assert lib.omb_fill_uniform(None, -1, 0, 0, 0, None) == -1
    assert lib.omb_check_dgemm(-1, 1, 1, 16, 1, 1, 16, 1, 1, None, 1, 1, 16, None) == -1
    assert lib.omb_check_errnorm(2, 2, 16, 2, 1, 16, None, None) == -1
    assert lib.omb_calib_fmac(0, 10, 16, None) == -1
    assert lib.omb_calib_fmac(4, 0, 16, None) == -1
    assert lib.omb_calib_fmac(4, 10, None, None) == -1
    assert "omb_calib_fmac" in ompblas.last_error()
    assert lib.omb_calib_mfma(0, 10, 1, 16, None) == -1
    assert lib.omb_calib_mfma(4, 0, 1, 16, None) == -1
    assert lib.omb_calib_mfma(4, 10, 0, None, None) == -1
    assert "omb_calib_mfma" in ompblas.last_error()


def test_torch_mirror_shape_errors(ompblas):
    """amt::mtm/mtv/dot_prod throw std::runtime_error at closure creation on
    non-matrix input or mismatched extents (include/mtm.hpp:234-250,
    include/mtv.hpp:126-144, include/dot.hpp:196-214)."""
    import torch
    a = torch.zeros(3, 4)
    b = torch.zeros(5, 2)
    c = torch.zeros(3, 2)
    with pytest.raises(RuntimeError, match="dimension mismatch"):
        ompblas.mtm(c, a, b)
    with pytest.raises(RuntimeError, match="must be the matrices"):
        ompblas.mtm(c, torch.zeros(3), b)
    with pytest.raises(RuntimeError, match="dimension mismatch"):
        ompblas.mtv(torch.zeros(3), a, torch.zeros(5))
    with pytest.raises(RuntimeError, match="same size"):
        ompblas.dot_prod(torch.zeros(1), torch.zeros(4), torch.zeros(5))


def test_oracle_is_not_linked_into_the_product():
    """The product library must not depend on the oracle."""
    import subprocess
    out = subprocess.run(["ldd", helpers.graft.load_package().LIB_PATH], capture_output=True,
                         text=True).stdout
    assert "oracle" not in out
    syms = subprocess.run(["nm", "-D", helpers.graft.load_package().LIB_PATH],
                          capture_output=True, text=True).stdout
    assert "oracle_" not in syms


def test_group_plan_rejects_bad_descriptors(ompblas):
    """Descriptor validation runs before any device call."""
    import ctypes
    lib = ompblas.lib()
    h = ctypes.c_void_p()
    bad = (ompblas.GemmDesc * 1)(ompblas.GemmDesc(4, 4, 4, 16, 2, 16, 4, 16, 4, 1.0, 0.0))  # lda < K
    assert lib.omb_group_plan_create(bad, 1, None, 0, 0, None, ctypes.byref(h)) == -1
    neg = (ompblas.GemmDesc * 1)(ompblas.GemmDesc(-1, 4, 4, 16, 4, 16, 4, 16, 4, 1.0, 0.0))
    assert lib.omb_group_plan_create(neg, 1, None, 0, 0, None, ctypes.byref(h)) == -1
    badv = (ompblas.GemvDesc * 1)(ompblas.GemvDesc(4, 8, 16, 4, 16, 16, 1.0, 0.0))       # lda < N
    assert lib.omb_group_plan_create(None, 0, badv, 1, 0, None, ctypes.byref(h)) == -1
    assert lib.omb_group_plan_create(None, 0, None, 0, 9, None, ctypes.byref(h)) == -1   # mode
    assert lib.omb_autotune_group(None, 0, 9, None, None) == -1                          # mode
    assert lib.omb_group_plan_create(None, -1, None, 0, 0, None, ctypes.byref(h)) == -1
    assert "group plan" in ompblas.last_error()
    assert lib.omb_autotune_load(b"/nonexistent/table.txt") == -1


def test_serve_run_rejects_bad_arguments(ompblas):
    """Trace and configuration validation runs before any device call."""
    import ctypes
    lib = ompblas.lib()
    R, Cfg = ompblas.Request, ompblas.ServeConfig
    res = (ompblas.RequestResult * 2)()
    st = ompblas.ServeStats()
    good = Cfg(2, 0, 0, 1)

    def run(reqs, cfg=good, n=None):
        arr = (R * max(1, len(reqs)))(*reqs)
        return lib.omb_serve_run(arr, len(reqs) if n is None else n, ctypes.byref(cfg), res,
                                 ctypes.byref(st))

    ok = R(ompblas.REQ_DOT, 0, 0, 8, 0, 0.0, 16, 16, 16, 0)
    assert run([ok], n=-1) == -1
    assert run([ok], cfg=Cfg(0, 0, 0, 1)) == -1                      # no slots
    assert run([ok], cfg=Cfg(2, -5, 0, 1)) == -1                     # negative budget
    assert run([R(7, 0, 1, 1, 1, 0.0, 16, 16, 16, 0)]) == -1         # unknown kind
    assert run([R(ompblas.REQ_GEMM, 9, 4, 4, 4, 0.0, 16, 16, 16, 0)]) == -1   # bad mode
    assert run([R(ompblas.REQ_GEMM, 1, 4, 4, 4, 0.0, 16, None, 16, 0)]) == -1  # missing B
    assert run([R(ompblas.REQ_GEMV, 0, 4, 4, 0, 0.0, 16, 16, None, 0)]) == -1  # missing y
    assert run([ok, R(ompblas.REQ_DOT, 0, 0, 8, 0, -1.0, 16, 16, 16, 0)]) == -1  # arrival < 0
    assert run([R(ompblas.REQ_DOT, 0, 0, 8, 0, float("nan"), 16, 16, 16, 0)]) == -1
    assert "omb_serve_run" in ompblas.last_error()


def test_live_server_rejects_bad_arguments(ompblas):
    """omb_server_*: configuration and handle validation runs before any
    device call (no server is opened here)."""
    lib = ompblas.lib()
    Cfg = ompblas.ServeConfig
    h = ctypes.c_void_p()
    assert lib.omb_server_open(None, 0, ctypes.byref(h)) == ompblas.EINVAL
    assert lib.omb_server_open(ctypes.byref(Cfg(0, 0, 0, 1)), 0, ctypes.byref(h)) == -1
    assert lib.omb_server_open(ctypes.byref(Cfg(65, 0, 0, 1)), 0, ctypes.byref(h)) == -1
    assert lib.omb_server_open(ctypes.byref(Cfg(2, -5, 0, 1)), 0, ctypes.byref(h)) == -1
    assert lib.omb_server_open(ctypes.byref(Cfg(2, 0, 0, 1)), -1, ctypes.byref(h)) == -1
    assert lib.omb_server_open(ctypes.byref(Cfg(2, 0, 0, 1)), 0, None) == -1
    assert "omb_server_open" in ompblas.last_error()
    r = ompblas.Request(ompblas.REQ_DOT, 0, 0, 8, 0, 0.0, 16, 16, 16, 0)
    t = ctypes.c_int64()
    d = ctypes.c_int32()
    assert lib.omb_server_submit(None, ctypes.byref(r), ctypes.byref(t)) == -1
    assert lib.omb_server_poll(None, 0, ctypes.byref(d), None) == -1
    assert lib.omb_server_wait(None, 0, None) == -1
    assert lib.omb_server_close(None, None) == -1


def test_serve_request_bytes(ompblas):
    """Workspace a request holds in the loop: split GEMMs their packed
    operands (rounded to 256 B), plain GEMMs with few tiles and K up to 512
    one partial plane per 128-k block, dots their partials; bad requests
    are refused."""
    R, lib = ompblas.Request, ompblas.lib()
    pack = lib.omb_split_pack_bytes(300, 260) + lib.omb_split_pack_bytes(200, 260)
    split = ompblas.serve_request_bytes(R(ompblas.REQ_GEMM, 1, 300, 200, 260, 0.0, 16, 16, 16, 0))
    assert split >= pack and split % 256 == 0
    # 3 x 2 tiles, K = 500: four single-block splits of 300 x 200 floats
    plain = ompblas.serve_request_bytes(R(ompblas.REQ_GEMM, 0, 300, 200, 500, 0.0, 16, 16, 16, 0))
    assert plain == 4 * 300 * 200 * 4
    assert ompblas.serve_request_bytes(R(ompblas.REQ_GEMM, 0, 300, 200, 100, 0.0, 16, 16, 16, 0)) == 0
    # scaled FP16: 2-plane packed operands, then exponents and maxima of every line
    f16 = ompblas.serve_request_bytes(R(ompblas.REQ_GEMM, 3, 300, 200, 260, 0.0, 16, 16, 16, 0))
    assert f16 % 256 == 0 and f16 >= 2 * (300 + 200) * 260 * 2 + (300 + 200) * 8
    assert f16 < split
    # the 3-product BF16 form packs only the two planes it reads
    p3 = ompblas.serve_request_bytes(R(ompblas.REQ_GEMM, 2, 300, 200, 260, 0.0, 16, 16, 16, 0))
    two_planes = 2 * (384 + 256) * 272 * 2          # rows padded to 128, k to 16
    assert p3 == (two_planes + 255) // 256 * 256 and p3 < split
    assert ompblas.serve_request_bytes(R(ompblas.REQ_GEMV, 0, 30, 20, 0, 0.0, 16, 16, 16, 0)) == 0
    assert ompblas.serve_request_bytes(R(ompblas.REQ_DOT, 0, 0, 5000, 0, 0.0, 16, 16, 16, 0)) > 0
    with pytest.raises(ompblas.OmbError):
        ompblas.serve_request_bytes(R(ompblas.REQ_GEMM, 9, 4, 4, 4, 0.0, 16, 16, 16, 0))


def test_serving_trace_mix_and_ranges(ompblas):
    """The Poisson trace: seeded, sorted arrivals, the 50/30/20 mix and the
    log-uniform size ranges."""
    import numpy as np
    t = ompblas.serving_trace(20000, 1000.0, seed=4)
    assert t == ompblas.serving_trace(20000, 1000.0, seed=4)
    kinds = np.array([x[0] for x in t])
    frac = [(kinds == k).mean() for k in range(3)]
    assert abs(frac[0] - 0.5) < 0.02 and abs(frac[1] - 0.3) < 0.02 and abs(frac[2] - 0.2) < 0.02
    arr = np.array([x[4] for x in t])
    assert arr[0] == 0.0 and np.all(np.diff(arr) >= 0)
    assert abs(np.diff(arr).mean() - 1e-3) < 5e-5                   # rate 1000/s
    g = np.array([x[1:4] for x in t if x[0] == 0])
    assert g.min() >= 32 and g.max() <= 4096
    # log-uniform: the median of each dimension sits near the geometric mean
    assert abs(np.median(g) - np.sqrt(32 * 4097)) / np.sqrt(32 * 4097) < 0.1
    v = np.array([x[1:3] for x in t if x[0] == 1])
    assert v.min() >= 256 and v.max() <= 16384
    d = np.array([x[2] for x in t if x[0] == 2])
    assert d.min() >= 1024 and d.max() <= 1 << 24


def test_cpp_dropin_program_is_built_against_the_library(ompblas):
    """build() compiles the C++ drop-in header and its test program with g++,
    linked to the product library (not to the oracle)."""
    import os
    import subprocess
    path = os.path.join(os.path.dirname(os.path.abspath(ompblas.__file__)), "lib", "test_dropin")
    assert os.path.exists(path)
    deps = subprocess.run(["ldd", path], capture_output=True, text=True).stdout
    assert "libompblas.so" in deps and "oracle" not in deps


def test_ger_and_transpose_reject_bad_arguments(ompblas):
    lib = ompblas.lib()
    assert lib.omb_sger(-1, 4, 1.0, 16, 16, 16, 4, 1, None) == -1
    assert lib.omb_sger(4, 4, 1.0, None, 16, 16, 4, 1, None) == -1
    assert lib.omb_sger(4, 4, 1.0, 16, 16, 16, 2, 1, None) == -1          # ld < N
    assert lib.omb_sger(4, 4, 1.0, 16, 16, 16, 8, 2, None) == -2          # no unit stride
    assert lib.omb_sger(0, 4, 1.0, None, None, None, 4, 1, None) == 0     # empty: no device call
    assert lib.omb_stranspose(4, 4, None, 4, 1, 16, 4, 1, None) == -1
    assert lib.omb_stranspose(4, 5, 16, 5, 1, 4096, 3, 1, None) == -1     # C is 5 x 4: ld 3 < 4
    assert lib.omb_stranspose(4, 4, 4096, 4, 1, 4096 + 8, 4, 1, None) == -1  # overlap
    assert "overlap" in ompblas.last_error()
    assert lib.omb_stranspose_inplace(8, 16, 4, None) == -1               # ld < n
    assert lib.omb_stranspose_inplace(0, None, 0, None) == 0


def test_outer_and_transpose_mirror_shape_errors(ompblas):
    """amt::outer_prod / amt::transpose throw at closure creation
    (include/outer.hpp:75-99, include/trans.hpp:97-115)."""
    import torch
    with pytest.raises(RuntimeError, match="dimension mismatch"):
        ompblas.outer_prod(torch.zeros(3, 4), torch.zeros(3), torch.zeros(5))
    with pytest.raises(RuntimeError, match="must be vector"):
        ompblas.outer_prod(torch.zeros(3, 4), torch.zeros(3, 2), torch.zeros(4))
    with pytest.raises(RuntimeError, match="dimension mismatch"):
        ompblas.transpose(torch.zeros(3, 4), torch.zeros(3, 4))
    with pytest.raises(RuntimeError, match="square"):
        ompblas.transpose(torch.zeros(3, 4))


@pytest.mark.parametrize("line", ["0 4 5 6 bf16x3_64x64", "1 4 5 6 f32_64x64",
                                  "3 4 5 6 f32_32x32", "9 4 5 6 f32_32x32",
                                  "0 4 5 6 f32_32x32 trailing"])
def test_autotune_load_refuses_class_of_another_mode(ompblas, tmp_path, line):
    """A plain-FP32 line naming the split tile, or a split line naming a plain
    tile, would make a plan compute another mode's arithmetic: refused with
    OMB_EINVAL before the table changes (no device call)."""
    path = tmp_path / "t.txt"
    path.write_text("# mode log2M log2N log2K class\n" + line + "\n")
    assert ompblas.lib().omb_autotune_load(str(path).encode()) == ompblas.EINVAL
    assert "omb_autotune_load" in ompblas.last_error()
