"""mtm (SGEMM) on the GPU through the C-ABI against the oracle.

Integer-exact cases mirror test/test.mtm.cpp:30-460 (8 (C,A,B) layout
combos, sizes 2..31, rand()%100, C = 0); floating-point cases compare the
max- and Frobenius-relative error against an FP64 product with the oracle's
(the reference algorithm's FP32 error on the same inputs): GPU <= 2x oracle.
"""
import numpy as np
import pytest
import torch

import helpers
from helpers import oracle_mtm, rand100, rel_errors

pytestmark = pytest.mark.gpu

LAYOUTS = ["FFF", "FFL", "FLF", "LFF", "FLL", "LFL", "LLF", "LLL"]
# plain FP32, split-BF16 (6 products), split-BF16 (3 products), scaled split-FP16 (3 products)
MODES = [0, 1, 2, 3]
ACCURATE = (0, 1, 3)  # the modes that keep the <= 2x error bound


def split16_bound(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Error bound of the scaled split-FP16 product, per element: FP32-level
    relative accuracy on |A||B| (2^-20 leaves room for the accumulation at
    these sizes) plus, for an element far below the largest of its row of A
    (column of B), FP16's subnormal step: an absolute 2^-39 of that largest,
    twice over for margin."""
    aa, bb = np.abs(a).astype(np.float64), np.abs(b).astype(np.float64)
    return (2.0 ** -20 * (aa @ bb)
            + 2.0 ** -38 * aa.max(axis=1, keepdims=True) * bb.sum(axis=0, keepdims=True)
            + 2.0 ** -38 * aa.sum(axis=1, keepdims=True) * bb.max(axis=0, keepdims=True))


def to_dev(a: np.ndarray, dev) -> torch.Tensor:
    """Device copy keeping the numpy array's storage order (strides)."""
    if a.flags.f_contiguous and not a.flags.c_contiguous:
        return torch.from_numpy(np.ascontiguousarray(a.T)).to(dev).t()
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def run_mtm(ompblas, c, a, b, mode, alpha=1.0, beta=1.0):
    M, K = a.shape
    N = b.shape[1]
    helpers.assert_view_in_bounds(a, M, K, a.stride(0), a.stride(1))
    helpers.assert_view_in_bounds(b, K, N, b.stride(0), b.stride(1))
    helpers.assert_view_in_bounds(c, M, N, c.stride(0), c.stride(1))
    st = ompblas.sgemm(M, N, K, alpha, a.data_ptr(), a.stride(0), a.stride(1), b.data_ptr(),
                       b.stride(0), b.stride(1), beta, c.data_ptr(), c.stride(0), c.stride(1),
                       mode)
    ompblas.check(st)


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("combo", LAYOUTS)
def test_mtm_integer_exact_reference_cases(dev, ompblas, combo, mode):
    """Values 0..99 are exact in one BF16 term and, scaled by a power of two,
    in one FP16 term, so every mode is exact here."""
    order = {"F": "F", "L": "C"}
    oc, oa, ob = (order[ch] for ch in combo)
    for sz in range(2, 32):
        a = rand100((sz, sz), oa)
        b = rand100((sz, sz), ob)
        c = np.zeros((sz, sz), np.float32, order=oc)
        oracle_mtm(c, a, b)
        tc = to_dev(np.zeros((sz, sz), np.float32, order=oc), dev)
        run_mtm(ompblas, tc, to_dev(a, dev), to_dev(b, dev), mode)
        got = tc.cpu().numpy()
        assert np.array_equal(got, c), (combo, sz, mode)


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("shape", [(1, 1, 1), (1, 300, 7), (300, 1, 7), (7, 5, 1),
                                   (129, 131, 17), (257, 100, 333), (64, 64, 2048),
                                   (1000, 37, 129)])
def test_mtm_ragged_integer_exact_accumulate(dev, ompblas, shape, mode):
    """C += A*B with C != 0 (the reference accumulates, simd_loop.hpp:169,187)."""
    M, N, K = shape
    a = rand100((M, K), "C") % 7
    b = rand100((K, N), "F") % 13
    c0 = rand100((M, N), "C")
    want = c0.astype(np.int64) + a.astype(np.int64) @ b.astype(np.int64)
    assert np.abs(want).max() < 2 ** 24
    tc = to_dev(c0, dev)
    run_mtm(ompblas, tc, to_dev(a, dev), to_dev(b, dev), mode)
    assert np.array_equal(tc.cpu().numpy().astype(np.int64), want)


@pytest.mark.parametrize("mode", MODES)
def test_mtm_leading_dimension_and_unaligned_views(dev, ompblas, mode):
    """Views with ld > extent, odd offsets (no 16-byte alignment) and odd ld."""
    big_a = torch.from_numpy(rand100((301, 203), "C")).to(dev)
    big_b = torch.from_numpy(rand100((157, 181), "C") % 9).to(dev).t()   # col-major 181x157
    big_c = torch.zeros((97, 389), device=dev)
    a = big_a[1:90, 3:120]           # 89 x 117, row stride 203, offset 206 floats
    b = big_b[3:120, 5:150]          # 117 x 145, col stride 181
    c = big_c[2:91, 7:152]           # 89 x 145, row stride 389
    run_mtm(ompblas, c, a, b, mode)
    want = a.cpu().numpy().astype(np.int64) @ b.cpu().numpy().astype(np.int64)
    assert np.array_equal(c.cpu().numpy().astype(np.int64), want)
    assert float(big_c[0].abs().sum()) == 0.0 and float(big_c[:, :7].abs().sum()) == 0.0


def test_mtm_alpha_beta_and_k_zero(dev, ompblas):
    a = torch.from_numpy(rand100((33, 20), "C")).to(dev)
    b = torch.from_numpy(rand100((20, 40), "C")).to(dev)
    c0 = torch.from_numpy(rand100((33, 40), "C")).to(dev)
    c = c0.clone()
    run_mtm(ompblas, c, a, b, 0, alpha=2.0, beta=0.5)
    want = 2.0 * (a.double() @ b.double()) + 0.5 * c0.double()
    assert torch.equal(c.double(), want)
    c = c0.clone()
    c[0, 0] = float("nan")
    run_mtm(ompblas, c, a, b, 0, alpha=1.0, beta=0.0)    # beta = 0 never reads C
    assert torch.equal(c.double(), a.double() @ b.double())
    c = c0.clone()
    ompblas.check(ompblas.sgemm(33, 40, 0, 1.0, a.data_ptr(), 20, 1, b.data_ptr(), 40, 1, 3.0,
                                c.data_ptr(), 40, 1, 0))
    assert torch.equal(c, 3.0 * c0)


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("combo", ["LLL", "FFF", "LFL", "FLF"])
def test_split_k_integer_exact(dev, ompblas, combo, mode):
    """Few output tiles and a long K: K is split over workgroups and the
    partial sums added in order.  Values 0..9 keep every sum exact (K * 81 <
    2^24), so each split count, layout and alpha/beta gives the exact product."""
    order = {"F": "F", "L": "C"}
    oc, oa, ob = (order[ch] for ch in combo)
    for m, n, k in [(200, 300, 3000), (1, 129, 4100), (257, 5, 1024), (128, 128, 20000)]:
        a = (rand100((m, k), oa) % 10).astype(np.float32)
        b = (rand100((k, n), ob) % 10).astype(np.float32)
        c0 = rand100((m, n), oc)
        want = 2.0 * (a.astype(np.float64) @ b.astype(np.float64)) + 0.5 * c0
        tc = to_dev(c0, dev)
        run_mtm(ompblas, tc, to_dev(a, dev), to_dev(b, dev), mode, alpha=2.0, beta=0.5)
        assert np.array_equal(tc.cpu().numpy().astype(np.float64), want), (combo, mode, m, n, k)


@pytest.mark.parametrize("mode", MODES)
def test_exact_k_splits_give_the_unsplit_bits(dev, ompblas, mode):
    """A product with few tiles splits K into single accumulation blocks
    whenever their partials fit (for any K here), and their in-order sum is
    the unsplit kernel's own fold: its rows equal, bit for bit, the same rows
    of a taller product that does not split (512 tiles: two workgroups per
    CU, one round, so a split cannot pay)."""
    N = 512
    big_m = 128 * 128                      # 128 x 4 tiles
    for m, k in [(200, 500), (1, 300), (129, 512), (77, 129), (200, 3000), (1, 4100),
                 (129, 1024), (77, 2000)]:
        a_big = torch.empty(big_m, k, device=dev)
        ompblas.fill_uniform(a_big, 11, 0)
        b = torch.empty(k, N, device=dev)
        ompblas.fill_uniform(b, 11, 1)
        c_big = torch.zeros(big_m, N, device=dev)
        run_mtm(ompblas, c_big, a_big, b, mode)
        c = torch.zeros(m, N, device=dev)
        run_mtm(ompblas, c, a_big[:m], b, mode)
        assert torch.equal(c, c_big[:m]), (mode, m, k)


@pytest.mark.parametrize("M,N,K", [(256, 192, 8192), (1024, 1024, 4096)])
def test_split_k_deterministic_and_within_error_bound(dev, ompblas, M, N, K):
    """Split-K products (U[-1,1)) repeat bit for bit and keep the <= 2x
    error bound against the oracle, in every mode that claims it: one with
    single-block splits (256 x 192 x 8192) and one whose partials would not
    fit as single blocks, which takes coarser splits (1024 x 1024 x 4096)."""
    a = helpers.rng_uniform(M * K, 7, 0).reshape(M, K)
    b = helpers.rng_uniform(K * N, 7, 1).reshape(K, N)
    truth = a.astype(np.float64) @ b.astype(np.float64)
    c_ref = np.zeros((M, N), np.float32)
    oracle_mtm(c_ref, a, b)
    ref = rel_errors(c_ref, truth)
    ta, tb = to_dev(a, dev), to_dev(b, dev)
    for mode in ACCURATE:
        outs = []
        for _ in range(2):
            tc = torch.zeros(M, N, device=dev)
            run_mtm(ompblas, tc, ta, tb, mode)
            outs.append(tc)
        assert torch.equal(outs[0], outs[1]), mode
        got = rel_errors(outs[0].cpu().numpy(), truth)
        assert got[0] <= 2 * ref[0] and got[1] <= 2 * ref[1], (mode, got, ref)


@pytest.mark.parametrize("orders", ["CC", "FF", "CF"])
def test_nonfinite_inputs(dev, ompblas, orders):
    """Plain FP32 follows IEEE exactly as the reference's FP32 loops do: an Inf
    or NaN input makes the outputs it feeds Inf or NaN, and nothing else.  The
    split modes turn every output an Inf or NaN input feeds into NaN (the
    split of an infinity is not finite).  Outputs no special value feeds are
    unchanged (integer inputs: exact)."""
    M, K, N = 40, 30, 35
    a = rand100((M, K), orders[0])
    b = rand100((K, N), orders[1])
    a[3, 7], a[10, 2], a[20, 5] = np.inf, np.nan, -np.inf
    b[4, 9] = np.inf
    # IEEE products in FP64, summed per element (Inf*0 and Inf-Inf give NaN)
    truth = (a.astype(np.float64)[:, :, None] * b.astype(np.float64)[None, :, :]).sum(axis=1)
    ta, tb = to_dev(a, dev), to_dev(b, dev)
    fed_rows, fed_col = [3, 10, 20], 9
    for mode in MODES:
        tc = torch.zeros(M, N, device=dev)
        run_mtm(ompblas, tc, ta, tb, mode)
        got = tc.cpu().numpy()
        if mode == 0:
            assert np.array_equal(got, truth.astype(np.float32), equal_nan=True)
            continue
        assert np.isnan(got[fed_rows, :]).all() and np.isnan(got[:, fed_col]).all(), mode
        keep = np.ones((M, N), bool)
        keep[fed_rows, :] = False
        keep[:, fed_col] = False
        assert np.array_equal(got[keep], truth.astype(np.float32)[keep]), mode


@pytest.mark.parametrize("orders", ["CC", "FF"])
def test_split_values_beyond_bf16_range_stay_finite(dev, ompblas, orders):
    """Finite |x| above BF16's largest finite value (about 3.3895e38; values
    above 3.3962e38 round to infinity in BF16) still split exactly: FLT_MAX-
    sized entries of A against tiny entries of B give finite C within FP32
    rounding of the FP64 product, in every mode."""
    M, K, N = 6, 6, 5
    a = rand100((M, K), orders[0])
    b = rand100((K, N), orders[1])
    fmax = np.finfo(np.float32).max
    a[:, 0] = np.array([fmax, -fmax, 3.3965e38, -3.397e38, 3.0e38, 1.0], np.float32)
    b[0, :] = b[0, :] * np.float32(2.0 ** -110)
    truth = a.astype(np.float64) @ b.astype(np.float64)
    ta, tb = to_dev(a, dev), to_dev(b, dev)
    for mode in MODES:
        tc = torch.zeros(M, N, device=dev)
        run_mtm(ompblas, tc, ta, tb, mode)
        got = tc.cpu().numpy().astype(np.float64)
        assert np.isfinite(got).all(), mode
        if mode == 3:
            # Rows scaled by their largest element: the small entries beside a
            # FLT_MAX-sized one keep only an absolute accuracy of 2^-39 of it.
            bound = split16_bound(a, b)
            assert np.all(np.abs(got - truth) <= bound), mode
            continue
        tol = 1e-6 if mode < 2 else 2e-5
        assert np.all(np.abs(got - truth) <= tol * np.abs(truth) + 1e-3), mode


@pytest.mark.parametrize("orders", ["CC", "FF", "CF"])
def test_split_fp16_rows_and_columns_of_any_scale(dev, ompblas, orders):
    """The scaled FP16 split normalises every row of A and column of B by a
    power of two, so rows and columns 2^+-60 apart (products up to 2^120) are
    as accurate as unscaled ones: componentwise within the split16 bound and
    within 2x plain FP32's error against FP64 per element class.  All-zero
    rows and columns (scale exponent 0) give exact zeros."""
    M, K, N = 96, 700, 80
    rng = np.random.default_rng(5)
    a = rng.uniform(-1, 1, (M, K)).astype(np.float32)
    b = rng.uniform(-1, 1, (K, N)).astype(np.float32)
    a *= (2.0 ** rng.choice([-60, 0, 60], size=(M, 1))).astype(np.float32)
    b *= (2.0 ** rng.choice([-60, 0, 60], size=(1, N))).astype(np.float32)
    a[7, :] = 0.0
    b[:, 11] = 0.0
    a = np.asarray(a, order=orders[0])
    b = np.asarray(b, order=orders[1])
    truth = a.astype(np.float64) @ b.astype(np.float64)
    ta, tb = to_dev(a, dev), to_dev(b, dev)
    got = {}
    for mode in (0, 3):
        tc = torch.zeros(M, N, device=dev)
        run_mtm(ompblas, tc, ta, tb, mode)
        got[mode] = tc.cpu().numpy().astype(np.float64)
    assert np.all(np.abs(got[3] - truth) <= split16_bound(a, b))
    assert (got[3][7, :] == 0).all() and (got[3][:, 11] == 0).all()
    # per element, relative to |A||B|: the scaled split is no worse than 2x plain
    absab = np.abs(a).astype(np.float64) @ np.abs(b).astype(np.float64)
    nz = absab > 0
    err = {m: (np.abs(got[m] - truth)[nz] / absab[nz]).max() for m in (0, 3)}
    assert err[3] <= 2 * err[0], err


def _fp_case(dev, ompblas, M, N, K, seed, c_order="C", a_order="C", b_order="C"):
    a = helpers.rng_uniform(M * K, seed, 0).reshape((M, K), order=a_order)
    b = helpers.rng_uniform(K * N, seed, 1).reshape((K, N), order=b_order)
    truth = a.astype(np.float64) @ b.astype(np.float64)
    c_ref = np.zeros((M, N), np.float32, order=c_order)
    oracle_mtm(c_ref, a, b)
    ref = rel_errors(c_ref, truth)
    ta, tb = to_dev(a, dev), to_dev(b, dev)
    out = {}
    for mode in MODES:
        tc = to_dev(np.zeros((M, N), np.float32, order=c_order), dev)
        run_mtm(ompblas, tc, ta, tb, mode)
        out[mode] = rel_errors(tc.cpu().numpy(), truth)
    return ref, out


def test_mtm_config1_error_vs_reference(dev, ompblas):
    """Config 1: row-major 1024^3, U[-1,1) from seed 0, C = 0.  Plain FP32 and
    split-BF16 (6 products): max- and Frobenius-relative error <= 2x the
    reference algorithm's (oracle, this host's KC)."""
    ref, out = _fp_case(dev, ompblas, 1024, 1024, 1024, 0)
    print("oracle", ref, "gpu", out)
    for mode in ACCURATE:
        assert out[mode][0] <= 2 * ref[0], (mode, out[mode], ref)
        assert out[mode][1] <= 2 * ref[1], (mode, out[mode], ref)
    assert out[2][1] < 1e-4   # 3-product variant: only a loose sanity bound


@pytest.mark.parametrize("orders", [("F", "F", "F"), ("C", "F", "C"), ("F", "C", "F")])
def test_mtm_mixed_layout_error(dev, ompblas, orders):
    ref, out = _fp_case(dev, ompblas, 777, 555, 1999, 13, *orders)
    for mode in ACCURATE:
        assert out[mode][1] <= 2 * ref[1], (mode, out[mode], ref)


def test_mtm_deterministic(dev, ompblas):
    M = N = K = 700
    a = torch.empty((M, K), device=dev)
    b = torch.empty((K, N), device=dev)
    ompblas.fill_uniform(a, 3, 0)
    ompblas.fill_uniform(b, 3, 1)
    for mode in MODES:
        outs = []
        for _ in range(2):
            c = torch.zeros((M, N), device=dev)
            run_mtm(ompblas, c, a, b, mode)
            outs.append(c)
        assert torch.equal(outs[0], outs[1])


@pytest.mark.parametrize("layout", ["row", "col"])
def test_mtm_8192_slab_error_vs_reference(dev, ompblas, layout):
    """Config 2 (8192^3, seed 1) in both layouts: error on a slab of 64 rows,
    GPU vs the oracle's on the same rows (the oracle's per-element result does
    not depend on which other rows are computed)."""
    n = 8192
    order = "C" if layout == "row" else "F"
    a = helpers.rng_uniform(n * n, 1, 0).reshape((n, n), order=order)
    b = helpers.rng_uniform(n * n, 1, 1).reshape((n, n), order=order)
    rows = slice(4000, 4064)
    truth = a[rows].astype(np.float64) @ b.astype(np.float64)
    c_ref = np.zeros((64, n), np.float32, order=order)
    oracle_mtm(c_ref, np.asarray(a[rows], order=order), b)
    ref = rel_errors(c_ref, truth)
    ta, tb = to_dev(a, dev), to_dev(b, dev)
    del a, b
    for mode in ACCURATE:
        tc = to_dev(np.zeros((n, n), np.float32, order=order), dev)
        run_mtm(ompblas, tc, ta, tb, mode)
        got = rel_errors(tc[rows].cpu().numpy(), truth)
        print(layout, mode, "gpu", got, "oracle", ref)
        assert got[0] <= 2 * ref[0] and got[1] <= 2 * ref[1], (mode, got, ref)


@pytest.mark.parametrize("dist", ["uniform", "pow2"])
def test_mtm_split_config_8192_both_distributions(dev, ompblas, dist):
    """The split-precision configuration: 8192^3 row-major from seed 2,
    U[-1,1) and sign * 2^u with u ~ U[-8, 8].  On a 64-row slab every
    accurate mode is within 2x the oracle's error on the same rows; on the
    whole matrix (FP64 check product) the split modes are within 2x plain
    FP32's error."""
    n = 8192
    if dist == "uniform":
        a = helpers.rng_uniform(n * n, 2, 0).reshape(n, n)
        b = helpers.rng_uniform(n * n, 2, 1).reshape(n, n)
    else:
        a = helpers.rng_pow2(n * n, 2, 0).reshape(n, n)
        b = helpers.rng_pow2(n * n, 2, 1).reshape(n, n)
    rows = slice(4096, 4160)
    truth = a[rows].astype(np.float64) @ b.astype(np.float64)
    c_ref = np.zeros((64, n), np.float32)
    oracle_mtm(c_ref, np.ascontiguousarray(a[rows]), b)
    ref = rel_errors(c_ref, truth)
    ta, tb = to_dev(a, dev), to_dev(b, dev)
    del a, b
    c64 = ompblas.check_dgemm(ta, tb)
    full = {}
    for mode in ACCURATE:
        tc = torch.zeros((n, n), device=dev)
        run_mtm(ompblas, tc, ta, tb, mode)
        got = rel_errors(tc[rows].cpu().numpy(), truth)
        full[mode] = ompblas.rel_errors(tc, c64)
        print(dist, mode, "slab gpu", got, "oracle", ref, "full", full[mode])
        assert got[0] <= 2 * ref[0] and got[1] <= 2 * ref[1], (dist, mode, got, ref)
        del tc
    for mode in (1, 3):
        assert full[mode][0] <= 2 * full[0][0] and full[mode][1] <= 2 * full[0][1], (
            dist, mode, full)


@pytest.mark.parametrize("nprod", [6, 3])
@pytest.mark.parametrize("shape", [(256, 384, 512), (130, 77, 1000), (1, 129, 17),
                                   (64, 64, 16), (300, 200, 33), (128, 128, 8192),
                                   (128, 128, 8200)])
def test_packed_split_path_matches_one_call(dev, ompblas, shape, nprod):
    """omb_split_pack + omb_sgemm_packed (operands split once, reusable) gives
    bitwise the result of omb_sgemm in the same split mode, for both layouts
    of C.  With 3 products the two run different loops: the packed operands
    hold 3 planes (single-buffered fragments), omb_sgemm stages the 2 it
    reads (fragments double-buffered in registers, slices in pairs), so this
    pins one loop against the other at 1, 2, 3 and an odd count of 16-k
    slices, and over K split across workgroups (one 128 x 128 tile with
    K = 8192 and 8200: 256-k splits, the last one a single slice)."""
    M, N, K = shape
    lib = ompblas.lib()
    a = torch.empty((M, K), device=dev)
    b = torch.empty((N, K), device=dev).t()       # column-major K x N
    ompblas.fill_uniform(a, 17, 0)
    ompblas.fill_uniform(b.t(), 17, 1)
    pa = torch.empty(lib.omb_split_pack_bytes(M, K), dtype=torch.uint8, device=dev)
    pb = torch.empty(lib.omb_split_pack_bytes(N, K), dtype=torch.uint8, device=dev)
    # B^T (N x K) of the column-major B (strides (1, K)) has strides (K, 1).
    bt_rs, bt_cs = b.stride(1), b.stride(0)
    helpers.assert_view_in_bounds(a, M, K, K, 1)
    helpers.assert_view_in_bounds(b, N, K, bt_rs, bt_cs)
    ompblas.check(lib.omb_split_pack(M, K, a.data_ptr(), K, 1, pa.data_ptr(), None))
    ompblas.check(lib.omb_split_pack(N, K, b.data_ptr(), bt_rs, bt_cs, pb.data_ptr(), None))
    mode = 1 if nprod == 6 else 2
    for c_order in ("C", "F"):
        want = to_dev(np.zeros((M, N), np.float32, order=c_order), dev)
        run_mtm(ompblas, want, a, b, mode)
        got = to_dev(np.zeros((M, N), np.float32, order=c_order), dev)
        ompblas.check(lib.omb_sgemm_packed(M, N, K, 1.0, pa.data_ptr(), pb.data_ptr(), 0.0,
                                           got.data_ptr(), got.stride(0), got.stride(1), nprod,
                                           None))
        assert torch.equal(got, want), (shape, nprod, c_order)


def test_concurrent_streams_equal_sequential(dev, ompblas):
    """Calls on different streams use their own workspaces (split-BF16 packed
    operands, split-K partials, column-GEMV and dot partials), so four
    streams running at once give the same bits as the same calls in sequence
    on the default stream."""
    jobs = []
    for s in range(4):
        m, n, k = 96 + 32 * s, 160 - 16 * s, 1024 + 512 * s       # few tiles: K is split
        a = torch.empty(m, k, device=dev)
        b = torch.empty(k, n, device=dev)
        ompblas.fill_uniform(a, 20 + s, 0)
        ompblas.fill_uniform(b, 20 + s, 1)
        v = torch.empty(3000 + 100 * s, 700, device=dev).t()   # column-major 700 x (3000 + 100s)
        ompblas.fill_uniform(v, 20 + s, 2)
        x = torch.empty(v.shape[1], device=dev)
        ompblas.fill_uniform(x, 20 + s, 3)
        jobs.append((m, n, k, a, b, v, x))
    torch.cuda.synchronize()

    def outputs(job):  # zeroed before any call is issued
        m, n, k, a, b, v, x = job
        return [torch.zeros(m, n, device=dev), torch.zeros(m, n, device=dev),
                torch.zeros(v.shape[0], device=dev), torch.zeros(1, device=dev)]

    def run(job, outs, stream):
        m, n, k, a, b, v, x = job
        sp = stream.cuda_stream if stream is not None else None
        for mode, c in zip((1, 0), outs[:2]):
            ompblas.check(ompblas.sgemm(m, n, k, 1.0, a.data_ptr(), k, 1, b.data_ptr(), n, 1, 0.0,
                                        c.data_ptr(), n, 1, mode, stream=sp))
        ompblas.check(ompblas.sgemv(v.shape[0], v.shape[1], 1.0, v.data_ptr(), v.stride(0),
                                    v.stride(1), x.data_ptr(), 1.0, outs[2].data_ptr(), stream=sp))
        ompblas.check(ompblas.sdot(x.numel(), x.data_ptr(), x.data_ptr(), outs[3].data_ptr(),
                                   stream=sp))
        return outs

    want = [outputs(j) for j in jobs]
    torch.cuda.synchronize()
    for j, o in zip(jobs, want):
        run(j, o, None)
    torch.cuda.synchronize()
    streams = [torch.cuda.Stream() for _ in jobs]
    for rep in range(2):
        got = [outputs(j) for j in jobs]
        torch.cuda.synchronize()
        for j, o, st in zip(jobs, got, streams):
            run(j, o, st)
        torch.cuda.synchronize()
        for w, g in zip(want, got):
            for i, (p, q) in enumerate(zip(w, g)):
                assert torch.equal(p, q), (rep, i)
