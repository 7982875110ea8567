"""The FP64 measurement kernels (omb_check_dgemm, omb_check_errnorm) against
numpy FP64 on the host.  Every GEMM error figure is measured against
omb_check_dgemm, so it is validated here on ragged shapes, both storage
orders of every operand, with and without the added c0, on both input
distributions of the split-precision configuration, and on 1024 sampled
entries of the full 8192^3 product.

Tolerance: both sides are FP64 dot products of FP32 inputs (exact
products), so each differs from the exact value by at most
gamma_K * sum_k |a_ik b_kj| with gamma_K = K u / (1 - K u), u = 2^-53
(Higham, Accuracy and Stability of Numerical Algorithms, eq. 3.5); the two
may differ by twice that.  A transposed operand, a wrong stride, a missing
c0 or a dropped k block moves entries by far more."""
import numpy as np
import pytest
import torch

import helpers

pytestmark = pytest.mark.gpu

U64 = 2.0 ** -53


def gamma(k: int) -> float:
    return k * U64 / (1 - k * U64)


def to_dev(a: np.ndarray, dev) -> torch.Tensor:
    if a.ndim == 2 and a.flags.f_contiguous and not a.flags.c_contiguous:
        return torch.from_numpy(np.ascontiguousarray(a.T)).to(dev).t()
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def inputs(kind: str, shape, seed: int, stream: int, order: str) -> np.ndarray:
    n = int(np.prod(shape))
    v = helpers.rng_uniform(n, seed, stream) if kind == "uniform" else \
        helpers.rng_pow2(n, seed, stream)
    return v.reshape(shape, order=order)


def check_dgemm(ompblas, ta, tb, c0=None):
    M, K = ta.shape
    N = tb.shape[1]
    helpers.assert_view_in_bounds(ta, M, K, ta.stride(0), ta.stride(1))
    helpers.assert_view_in_bounds(tb, K, N, tb.stride(0), tb.stride(1))
    if c0 is not None:
        helpers.assert_view_in_bounds(c0, M, N, c0.stride(0), c0.stride(1))
    out = ompblas.check_dgemm(ta, tb, c0)
    torch.cuda.synchronize()
    return out.cpu().numpy()


@pytest.mark.parametrize("kind", ["uniform", "pow2"])
@pytest.mark.parametrize("orders", [("C", "C", "C"), ("F", "F", "F"), ("C", "F", "F"),
                                    ("F", "C", "C")])
@pytest.mark.parametrize("with_c0", [False, True])
def test_check_dgemm_against_numpy(dev, ompblas, kind, orders, with_c0):
    """Ragged 1000 x 700 x 900 (no dimension a multiple of the 64-wide tile
    or the 16-wide k slice), every entry checked."""
    M, N, K = 1000, 700, 900
    a_order, b_order, c_order = orders
    a = inputs(kind, (M, K), 2, 0, a_order)
    b = inputs(kind, (K, N), 2, 1, b_order)
    c0 = inputs("uniform", (M, N), 2, 2, c_order) if with_c0 else None
    got = check_dgemm(ompblas, to_dev(a, dev), to_dev(b, dev),
                      to_dev(c0, dev) if with_c0 else None)
    a64, b64 = a.astype(np.float64), b.astype(np.float64)
    want = a64 @ b64
    scale = np.abs(a64) @ np.abs(b64)
    if with_c0:
        want += c0.astype(np.float64)
        scale += np.abs(c0.astype(np.float64))
    assert got.shape == (M, N) and got.flags.c_contiguous
    excess = np.abs(got - want) - 2 * gamma(K + 1) * scale
    assert excess.max() <= 0, (kind, orders, with_c0, float(excess.max()))
    # the check is sharp enough to see one dropped k (the smallest product
    # term is far above the tolerance): leaving out k = K-1 fails it
    short = want - np.outer(a64[:, K - 1], b64[K - 1, :])
    assert (np.abs(got - short) - 2 * gamma(K + 1) * scale).max() > 0


@pytest.mark.parametrize("kind", ["uniform", "pow2"])
def test_check_dgemm_8192_sampled(dev, ompblas, kind):
    """The split-precision configuration's inputs (8192^3, seed 2, both
    distributions): 1024 sampled entries of the FP64 check product against
    numpy FP64 dot products of the same rows and columns."""
    n = 8192
    a = torch.empty((n, n), device=dev)
    b = torch.empty((n, n), device=dev)
    if kind == "uniform":
        ompblas.fill_uniform(a, 2, 0)
        ompblas.fill_uniform(b, 2, 1)
    else:
        a.copy_(torch.from_numpy(helpers.rng_pow2(n * n, 2, 0)).view(n, n))
        b.copy_(torch.from_numpy(helpers.rng_pow2(n * n, 2, 1)).view(n, n))
    got = check_dgemm(ompblas, a, b)
    rng = np.random.default_rng(1024)
    rows = rng.integers(0, n, 32)
    cols = rng.integers(0, n, 32)
    ah = a[torch.from_numpy(rows).to(dev)].double().cpu().numpy()          # 32 x K
    bh = b[:, torch.from_numpy(cols).to(dev)].double().cpu().numpy()       # K x 32
    want = ah @ bh
    scale = np.abs(ah) @ np.abs(bh)
    sample = got[np.ix_(rows, cols)]
    assert (np.abs(sample - want) - 2 * gamma(n) * scale).max() <= 0
    # pow2 operands make products span 2^-16 .. 2^16: the check must still
    # resolve the smallest terms (relative agreement far below FP32's eps)
    assert np.abs(sample - want).max() <= 1e-12 * np.abs(want).max()


def test_check_errnorm_against_numpy(dev, ompblas):
    """omb_check_errnorm: max |C - C64|, max |C64|, sum (C - C64)^2 and
    sum C64^2 for a strided column-major C, against numpy."""
    M, N = 777, 1301
    c64 = inputs("pow2", (M, N), 3, 0, "C").astype(np.float64)
    c = (c64 + inputs("uniform", (M, N), 3, 1, "C").astype(np.float64) * 1e-3).astype(np.float32)
    big = torch.zeros((N, M + 5), device=dev)           # column-major C, ld = M + 5
    tc = big[:, :M].t()
    tc.copy_(torch.from_numpy(c))
    t64 = torch.from_numpy(c64).to(dev)
    helpers.assert_view_in_bounds(tc, M, N, tc.stride(0), tc.stride(1))
    got = ompblas.errnorm(tc, t64)
    d = c.astype(np.float64) - c64
    want = (np.abs(d).max(), np.abs(c64).max(), (d * d).sum(), (c64 * c64).sum())
    assert got[0] == want[0] and got[1] == want[1]
    assert got[2] == pytest.approx(want[2], rel=1e-12)
    assert got[3] == pytest.approx(want[3], rel=1e-12)
