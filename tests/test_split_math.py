"""The arithmetic the split GEMM modes rest on, restated in numpy and pinned
on the CPU (no GPU): the exactness of the 3-way BF16 split and the error
bound of the power-of-two scaled 2-way FP16 split.  The GPU tests check the
kernels' results; these check the claims those results rest on."""
import numpy as np
import pytest


BF16_MAX = float.fromhex("0x1.fep127")


def bf16_round(v: np.ndarray) -> np.ndarray:
    """Nearest BF16 value (ties to even) of float64 values that are exact
    FP32 values, in FP64 arithmetic: 8 significant bits, and BF16's
    subnormal step 2^-133 below 2^-126.  (FP64, not FP32 bit tricks: a
    process that loaded code built with -ffast-math flushes FP32 subnormals.)
    Values that round past BF16's largest become infinite."""
    v = np.asarray(v, np.float64)
    a = np.abs(v)
    e = np.floor(np.log2(np.where(a > 0, a, 1.0)))
    q = 2.0 ** (np.maximum(e, -126) - 7)
    r = np.where(a > 0, np.round(v / q) * q, v)
    return np.where(np.abs(r) > BF16_MAX, np.copysign(np.inf, v), r)


def split3_bf16(x: np.ndarray):
    """h0 = bf16(x) (truncated to 8 bits when x is finite but rounds past
    BF16's largest), h1 = bf16(x - h0), h2 = bf16(x - h0 - h1); the
    subtractions are exact, so FP64 gives the FP32 values."""
    x = np.asarray(x)
    x = x if x.dtype == np.float64 else x.astype(np.float64)   # values on FP32's grid
    h0 = bf16_round(x)
    over = np.isinf(h0) & np.isfinite(x)
    q = 2.0 ** (np.floor(np.log2(np.where(over, np.abs(x), 1.0))) - 7)
    h0 = np.where(over, np.trunc(x / q) * q, h0)
    r1 = x - h0
    h1 = bf16_round(r1)
    h2 = bf16_round(r1 - h1)
    return h0, h1, h2


def signed_pow2_samples(rng, e: int, n: int = 20000) -> np.ndarray:
    """FP32 values of magnitude in [2^e, 2^(e+1)), as float64 (rounded to
    FP32's grid in FP64, subnormals included)."""
    m = rng.uniform(1, 2, n) * np.where(rng.random(n) < 0.5, -1.0, 1.0)
    q = 2.0 ** (max(e, -126) - 23)                       # FP32 spacing there
    return np.round(m * 2.0 ** e / q) * q


@pytest.mark.parametrize("e", [127, 100, 0, -1, -60, -100, -110])
def test_bf16_split_is_exact_down_to_2_pow_minus_110(e):
    rng = np.random.default_rng(e + 1000)
    x = signed_pow2_samples(rng, e)
    h0, h1, h2 = split3_bf16(x)
    s = h0 + h1 + h2
    assert np.array_equal(s, x.astype(np.float64))
    # the pieces fall by at least 2^8 each
    nz = h1 != 0
    assert np.all(np.abs(h1[nz]) <= 2.0 ** -8 * np.abs(h0[nz]))


@pytest.mark.parametrize("e", [-111, -120, -126, -135, -145])
def test_bf16_split_below_2_pow_minus_110_is_within_2_pow_minus_134(e):
    """BF16's subnormal step is 2^-133: the last piece of a value below
    2^-110 rounds, by at most half of it."""
    rng = np.random.default_rng(-e)
    x = signed_pow2_samples(rng, e)
    h0, h1, h2 = split3_bf16(x)
    err = np.abs(h0 + h1 + h2 - x.astype(np.float64))
    assert err.max() <= 2.0 ** -134
    if e == -111:
        assert err.max() > 0        # the bound is reached: not exact below 2^-110


def test_bf16_split_of_values_beyond_bf16_range_is_exact():
    fmax = np.finfo(np.float32).max
    x = np.array([fmax, -fmax, 3.3965e38, -3.397e38, 3.39e38], np.float32)
    h0, h1, h2 = split3_bf16(x)
    assert np.isfinite(h0).all()
    assert np.array_equal(h0 + h1 + h2, x.astype(np.float64))


def scaled_split16(X: np.ndarray, axis: int):
    """Rows (axis 1) or columns (axis 0) scaled by 2^e, e = 14 - floor(log2
    amax) (0 for a zero line), then x 2^e = h0 + h1 + r, h0 = fp16(x 2^e),
    h1 = fp16(x 2^e - h0)."""
    amax = np.max(np.abs(X), axis=axis, keepdims=True).astype(np.float64)
    e = np.where(amax > 0, 14 - np.floor(np.log2(np.where(amax > 0, amax, 1.0))), 0)
    xs = (X.astype(np.float64) * 2.0 ** e).astype(np.float32)   # exact (power of two)
    h0 = xs.astype(np.float16)
    h1 = (xs - h0.astype(np.float32)).astype(np.float16)
    return h0.astype(np.float64), h1.astype(np.float64), e, amax


def test_fp16_scaled_split_carries_every_element_to_its_bound():
    """|x - x_hat| <= max(2^-22 |x|, 2^-39 amax) with x_hat = (h0 + h1) 2^-e,
    for lines whose entries span 2^40 (so both terms of the bound bind), and
    the scaled line maximum lies in [2^14, 2^15)."""
    rng = np.random.default_rng(7)
    X = (rng.uniform(-1, 1, (64, 512)) * 2.0 ** rng.uniform(-40, 0, (64, 512))).astype(np.float32)
    X *= (2.0 ** rng.integers(-100, 100, (64, 1))).astype(np.float32)
    h0, h1, e, amax = scaled_split16(X, 1)
    xhat = (h0 + h1) * 2.0 ** -e
    x = X.astype(np.float64)
    assert np.all(np.abs(x - xhat) <= np.maximum(2.0 ** -22 * np.abs(x), 2.0 ** -39 * amax))
    scaled_max = amax * 2.0 ** e
    assert np.all((scaled_max >= 2.0 ** 14) & (scaled_max < 2.0 ** 15))


@pytest.mark.parametrize("dist", ["uniform", "pow2"])
def test_fp16_three_products_are_below_fp32_rounding(dist):
    """Dropping a1 b1 and the residuals costs less than FP32 accumulation:
    with the three products summed exactly (FP64), the max-relative error
    against the FP64 product is under 1e-7 (plain FP32 at this size: about
    4e-7), for both bench input distributions."""
    rng = np.random.default_rng(11)
    M = N = 128
    K = 2048
    if dist == "uniform":
        A = rng.uniform(-1, 1, (M, K)).astype(np.float32)
        B = rng.uniform(-1, 1, (K, N)).astype(np.float32)
    else:
        A = (np.sign(rng.uniform(-1, 1, (M, K))) * 2.0 ** rng.uniform(-8, 8, (M, K))).astype(np.float32)
        B = (np.sign(rng.uniform(-1, 1, (K, N))) * 2.0 ** rng.uniform(-8, 8, (K, N))).astype(np.float32)
    a0, a1, ea, _ = scaled_split16(A, 1)
    b0, b1, eb, _ = scaled_split16(B, 0)
    C3 = (a0 @ b1 + a1 @ b0 + a0 @ b0) * 2.0 ** -(ea + eb)
    T = A.astype(np.float64) @ B.astype(np.float64)
    rel = np.abs(C3 - T).max() / np.abs(T).max()
    assert rel < 1e-7, rel
    # the BF16 split needs six products for the same: three are 2^-16-level short
    a = list(split3_bf16(A))
    b = [p.T for p in split3_bf16(B.T)]
    C3b = a[0] @ b[1] + a[1] @ b[0] + a[0] @ b[0]
    C6b = C3b + a[0] @ b[2] + a[1] @ b[1] + a[2] @ b[0]
    assert np.abs(C6b - T).max() / np.abs(T).max() < 1e-7
    assert np.abs(C3b - T).max() / np.abs(T).max() > 1e-6
