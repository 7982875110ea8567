"""Seeded random cases through the C-ABI: shapes, storage orders, leading
dimensions padded past the extent, operands at 4-byte offsets from 16-byte
alignment, and alpha / beta drawn together, for mtm (every mode), mtv (both
layouts), dot, the outer product and both transposes.  Operands are small
integers (0..9) and alpha / beta are powers of two or zero, so every
product result is exactly an FP64 one, and a transpose is a copy: each case
must match bit for bit.  This follows the reference's integer-exact and
known-answer tests (test/test.mtm.cpp:30-460, test/test.mtv.cpp:29-211,
test/test.dot_prod.cpp:10-28, test/test.outer_prod.cpp:29-182,
test/test.trans.cpp:11-74) with the layout mix drawn at random instead of
enumerated."""
import numpy as np
import pytest
import torch

import helpers

pytestmark = pytest.mark.gpu

MODES = [0, 1, 2, 3]
SCALARS = [0.0, 1.0, 2.0, -0.5, 0.25]


def log_int(rng, lo, hi):
    return int(np.exp(rng.uniform(np.log(lo), np.log(hi + 1))))


def placed(values: np.ndarray, order: str, pad: int, off: int, dev):
    """A device view holding `values` (rows x cols) in `order` ("C" row-major,
    "F" column-major), with leading dimension extent + pad, starting `off`
    floats into its buffer; returns (view, rs, cs)."""
    rows, cols = values.shape
    if order == "C":
        ld = max(cols, 1) + pad
        buf = torch.zeros(off + rows * ld, device=dev)
        view = buf[off:].view(rows, ld)[:, :cols]
        rs, cs = ld, 1
    else:
        ld = max(rows, 1) + pad
        buf = torch.zeros(off + cols * ld, device=dev)
        view = buf[off:].view(cols, ld)[:, :rows].t()
        rs, cs = 1, ld
    view.copy_(torch.from_numpy(values))
    helpers.assert_view_in_bounds(view, rows, cols, rs, cs)
    return view, rs, cs


@pytest.mark.parametrize("mode", MODES)
def test_mtm_random_layouts_integer_exact(dev, ompblas, mode):
    rng = np.random.default_rng(1000 + mode)
    for case in range(48):
        M, N, K = log_int(rng, 1, 300), log_int(rng, 1, 300), log_int(rng, 1, 600)
        oa, ob, oc = (rng.choice(["C", "F"]) for _ in range(3))
        pa, pb, pc = (int(rng.integers(0, 6)) for _ in range(3))
        fa, fb, fc = (int(rng.integers(0, 4)) for _ in range(3))
        alpha, beta = float(rng.choice(SCALARS)), float(rng.choice(SCALARS))
        a = rng.integers(0, 10, (M, K)).astype(np.float32)
        b = rng.integers(0, 10, (K, N)).astype(np.float32)
        c0 = rng.integers(0, 10, (M, N)).astype(np.float32)
        va, rsa, csa = placed(a, oa, pa, fa, dev)
        vb, rsb, csb = placed(b, ob, pb, fb, dev)
        vc, rsc, csc = placed(c0, oc, pc, fc, dev)
        ompblas.check(ompblas.sgemm(M, N, K, alpha, va.data_ptr(), rsa, csa, vb.data_ptr(), rsb,
                                    csb, beta, vc.data_ptr(), rsc, csc, mode))
        want = alpha * (a.astype(np.float64) @ b.astype(np.float64)) + beta * c0
        got = vc.cpu().numpy().astype(np.float64)
        assert np.array_equal(got, want), (mode, case, M, N, K, oa, ob, oc, alpha, beta)


def test_mtv_random_layouts_integer_exact(dev, ompblas):
    rng = np.random.default_rng(2000)
    for case in range(96):
        M, N = log_int(rng, 1, 20000), log_int(rng, 1, 4000)
        if M * N > 1 << 23:
            N = max(1, (1 << 23) // M)
        order = rng.choice(["C", "F"])
        pad, off, xoff = int(rng.integers(0, 6)), int(rng.integers(0, 4)), int(rng.integers(0, 4))
        alpha, beta = float(rng.choice(SCALARS)), float(rng.choice(SCALARS))
        a = rng.integers(0, 10, (M, N)).astype(np.float32)
        x = rng.integers(0, 10, N).astype(np.float32)
        y0 = rng.integers(0, 10, M).astype(np.float32)
        va, rs, cs = placed(a, order, pad, off, dev)
        xb = torch.zeros(xoff + N, device=dev)
        xb[xoff:].copy_(torch.from_numpy(x))
        y = torch.from_numpy(y0).to(dev)
        ompblas.check(ompblas.sgemv(M, N, alpha, va.data_ptr(), rs, cs, xb[xoff:].data_ptr(),
                                    beta, y.data_ptr()))
        want = alpha * (a.astype(np.float64) @ x.astype(np.float64)) + beta * y0
        assert np.array_equal(y.cpu().numpy().astype(np.float64), want), (case, M, N, order,
                                                                           alpha, beta)


def test_dot_random_lengths_integer_exact(dev, ompblas):
    rng = np.random.default_rng(3000)
    for case in range(64):
        n = log_int(rng, 1, 200000)
        off = int(rng.integers(0, 4))
        x = rng.integers(0, 10, n).astype(np.float32)
        y = rng.integers(0, 10, n).astype(np.float32)
        xb = torch.zeros(off + n, device=dev)
        yb = torch.zeros(n + 3, device=dev)
        xb[off:].copy_(torch.from_numpy(x))
        yb[3:].copy_(torch.from_numpy(y))
        r = torch.full((1,), -1.0, device=dev)
        ompblas.check(ompblas.sdot(n, xb[off:].data_ptr(), yb[3:].data_ptr(), r.data_ptr()))
        want = float(x.astype(np.float64) @ y.astype(np.float64))
        assert float(r.item()) == want, (case, n, off)


def test_ger_random_layouts_integer_exact(dev, ompblas):
    """C += alpha x y^T (include/outer.hpp:15-48 with alpha = 1 there)."""
    rng = np.random.default_rng(4000)
    for case in range(48):
        M, N = log_int(rng, 1, 3000), log_int(rng, 1, 3000)
        order = rng.choice(["C", "F"])
        pad, off = int(rng.integers(0, 6)), int(rng.integers(0, 4))
        alpha = float(rng.choice(SCALARS))
        x = rng.integers(0, 10, M).astype(np.float32)
        y = rng.integers(0, 10, N).astype(np.float32)
        c0 = rng.integers(0, 10, (M, N)).astype(np.float32)
        vc, rs, cs = placed(c0, order, pad, off, dev)
        xb = torch.zeros(1 + M, device=dev)
        xb[1:].copy_(torch.from_numpy(x))
        yd = torch.from_numpy(y).to(dev)
        ompblas.check(ompblas.sger(M, N, alpha, xb[1:].data_ptr(), yd.data_ptr(), vc.data_ptr(),
                                   rs, cs))
        want = c0.astype(np.float64) + alpha * np.outer(x.astype(np.float64), y)
        assert np.array_equal(vc.cpu().numpy().astype(np.float64), want), (case, M, N, order)


def test_transpose_random_layouts_bitwise(dev, ompblas):
    """C = A^T out of place (include/trans.hpp:23-45), any storage orders,
    padded and offset: a pure copy, so every bit of A must arrive."""
    rng = np.random.default_rng(5000)
    for case in range(48):
        M, N = log_int(rng, 1, 3000), log_int(rng, 1, 3000)
        oa, oc = rng.choice(["C", "F"]), rng.choice(["C", "F"])
        a = rng.standard_normal((M, N)).astype(np.float32)
        va, rsa, csa = placed(a, oa, int(rng.integers(0, 6)), int(rng.integers(0, 4)), dev)
        vc, rsc, csc = placed(np.zeros((N, M), np.float32), oc, int(rng.integers(0, 6)),
                              int(rng.integers(0, 4)), dev)
        ompblas.check(ompblas.stranspose(M, N, va.data_ptr(), rsa, csa, vc.data_ptr(), rsc, csc))
        assert np.array_equal(vc.cpu().numpy(), a.T), (case, M, N, oa, oc)


def test_transpose_inplace_random_sizes_bitwise(dev, ompblas):
    """Square A <- A^T in place (include/trans.hpp:47-71), rows ld apart;
    the padding between rows is left untouched."""
    rng = np.random.default_rng(6000)
    for case in range(32):
        n = log_int(rng, 1, 3000)
        ld = n + int(rng.integers(0, 6))
        off = int(rng.integers(0, 4))
        full = rng.standard_normal(off + n * ld).astype(np.float32)
        buf = torch.from_numpy(full).to(dev)
        ompblas.check(ompblas.stranspose_inplace(n, buf[off:].data_ptr(), ld))
        got = buf.cpu().numpy()
        body = full[off:].reshape(n, ld)
        want = full.copy()
        want_body = want[off:].reshape(n, ld)
        want_body[:, :n] = body[:, :n].T
        assert np.array_equal(got, want), (case, n, ld, off)
