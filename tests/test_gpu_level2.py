"""Outer product (SGER) and transpose on the GPU through the C-ABI against the
oracle: the integer-exact reference cases of test/test.outer_prod.cpp and
test/test.trans.cpp, every storage order, and size-independent properties at
16384^2 (a rank-1 update by a ones vector is one rounded add; transposing
twice is the identity)."""
import numpy as np
import pytest
import torch

import helpers
from helpers import rand100

pytestmark = pytest.mark.gpu


def to_dev(a, dev):
    if a.ndim == 2 and a.flags.f_contiguous and not a.flags.c_contiguous:
        return torch.from_numpy(np.ascontiguousarray(a.T)).to(dev).t()
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def to_host(t):
    return t.cpu().numpy()


@pytest.mark.parametrize("order", ["F", "C"])
def test_outer_integer_exact_reference_cases(dev, ompblas, order):
    """amt::outer_prod(c, a, b)(): C += a b^T; sizes 2..31 and rectangular."""
    shapes = [(s, s) for s in range(2, 32)] + [(3, 17), (40, 5), (1, 9), (300, 257), (1000, 3)]
    for m, n in shapes:
        a, b = rand100((m,)), rand100((n,))
        c = rand100((m, n), order)
        want = c.copy()
        helpers.oracle_outer(want, a, b)
        tc = to_dev(c, dev)
        ompblas.outer_prod(tc, torch.from_numpy(a).to(dev), torch.from_numpy(b).to(dev))()
        assert np.array_equal(to_host(tc), want), (order, m, n)


def test_outer_float_layout_independent_and_one_rounding(dev, ompblas):
    """Row- and column-major C give the same bits; each element is
    fma(alpha*x, y, c): within half an ulp of the exact value when alpha = 1."""
    m, n = 1031, 517
    x = torch.empty(m, device=dev)
    y = torch.empty(n, device=dev)
    c0 = torch.empty(m, n, device=dev)
    for t, sid in ((x, 0), (y, 1), (c0, 2)):
        ompblas.fill_uniform(t, 21, sid)
    outs = []
    for alpha in (1.0, -0.375):
        per = []
        for order in ("C", "F"):
            c = c0.clone() if order == "C" else c0.t().contiguous().t()
            ompblas.check(ompblas.sger(m, n, alpha, x.data_ptr(), y.data_ptr(), c.data_ptr(),
                                       c.stride(0), c.stride(1)))
            per.append(c.contiguous())
        assert torch.equal(per[0], per[1]), alpha
        outs.append(per[0])
    exact = (c0.double() + x.double()[:, None] * y.double()[None, :]).cpu().numpy()
    got = outs[0].cpu().numpy()
    ulp = np.spacing(np.abs(exact).astype(np.float32)).astype(np.float64)
    assert np.all(np.abs(got - exact) <= 0.5 * ulp + 1e-30)
    exact2 = (c0.double() + (-0.375 * x.double())[:, None] * y.double()[None, :]).cpu().numpy()
    assert np.allclose(outs[1].cpu().numpy(), exact2, rtol=0, atol=2e-6)


def test_outer_unaligned_views_and_ld(dev, ompblas):
    big = torch.zeros(403, 517, device=dev)
    c = big[3:301, 1:390]                 # 298 x 389, offsets break 16-byte alignment
    xa = torch.from_numpy(rand100((300,))).to(dev)[1:299]
    ya = torch.from_numpy(rand100((391,))).to(dev)[2:391]
    helpers.assert_view_in_bounds(c, 298, 389, c.stride(0), c.stride(1))
    ompblas.outer_prod(c, xa, ya)()
    want = np.outer(xa.cpu().numpy().astype(np.int64), ya.cpu().numpy().astype(np.int64))
    assert np.array_equal(c.cpu().numpy().astype(np.int64), want)
    assert float(big[:3].abs().sum()) == 0 and float(big[:, :1].abs().sum()) == 0
    assert float(big[301:].abs().sum()) == 0 and float(big[:, 390:].abs().sum()) == 0


def test_outer_16384_ones_vector_is_one_rounded_add(dev, ompblas):
    """At the bench size: x = 1 makes every element fma(1, y[j], c) = c + y[j]
    rounded once, which torch's add reproduces bit for bit."""
    n = 16384
    c = torch.empty(n, n, device=dev)
    ompblas.fill_uniform(c, 31, 0)
    y = torch.empty(n, device=dev)
    ompblas.fill_uniform(y, 31, 1)
    x = torch.ones(n, device=dev)
    want = c + y[None, :]
    ompblas.check(ompblas.sger(n, n, 1.0, x.data_ptr(), y.data_ptr(), c.data_ptr(), n, 1))
    assert torch.equal(c, want)


@pytest.mark.parametrize("ao,co", [("C", "C"), ("C", "F"), ("F", "C"), ("F", "F")])
def test_transpose_out_of_place_exact(dev, ompblas, ao, co):
    """amt::transpose(c, a)(): C = A^T, sizes 2..31 plus multi-tile shapes."""
    shapes = [(s, s) for s in range(2, 32)] + [(3, 70), (70, 3), (129, 65), (200, 333)]
    for m, n in shapes:
        a = rand100((m, n), ao)
        want = np.zeros((n, m), np.float32, order=co)
        helpers.oracle_transpose(want, a)
        ta = to_dev(a, dev)
        tc = to_dev(np.zeros((n, m), np.float32, order=co), dev)
        ompblas.transpose(tc, ta)()
        assert np.array_equal(to_host(tc), want), (ao, co, m, n)


def test_transpose_strided_views_and_overlap(dev, ompblas):
    big = torch.from_numpy(rand100((300, 400))).to(dev)
    a = big[5:205, 3:303]                                   # 200 x 300, ld 400
    out = torch.full((320, 260), -1.0, device=dev)
    c = out[7:307, 11:211]                                   # 300 x 200, ld 260
    helpers.assert_view_in_bounds(a, 200, 300, a.stride(0), a.stride(1))
    helpers.assert_view_in_bounds(c, 300, 200, c.stride(0), c.stride(1))
    ompblas.transpose(c, a)()
    assert torch.equal(c, a.t())
    assert float((out[:7] != -1).sum()) == 0 and float((out[:, 211:] != -1).sum()) == 0
    sq = torch.zeros(64, 64, device=dev)
    assert ompblas.stranspose(64, 64, sq.data_ptr(), 64, 1, sq.data_ptr(), 64, 1) == ompblas.EINVAL
    assert "overlap" in ompblas.last_error()


def test_transpose_in_place_exact(dev, ompblas):
    """amt::transpose(a)(): square, in place; sizes 2..31, tile edges, ld > n."""
    for n in list(range(2, 32)) + [63, 64, 65, 129, 1000]:
        a = rand100((n, n), "F")
        want = a.copy(order="F")
        helpers.oracle_transpose_inplace(want)
        ta = to_dev(a, dev)
        ompblas.transpose(ta)()
        assert np.array_equal(to_host(ta), want), n
    big = torch.from_numpy(rand100((150, 170))).to(dev)
    v = big[:130, 7:137]                                     # 130 x 130, ld 170
    want = v.t().clone()
    helpers.assert_view_in_bounds(v, 130, 130, v.stride(0), v.stride(1))
    ompblas.transpose(v)()
    assert torch.equal(v, want)


def test_transpose_16384_twice_is_identity(dev, ompblas):
    n = 16384
    a = torch.empty(n, n, device=dev)
    ompblas.fill_uniform(a, 41, 0)
    c = torch.empty(n, n, device=dev)
    ompblas.check(ompblas.stranspose(n, n, a.data_ptr(), n, 1, c.data_ptr(), n, 1))
    assert torch.equal(c[:5, :7], a[:7, :5].t()) and torch.equal(c[-3:, 100:104], a[100:104, -3:].t())
    back = a.clone()
    ompblas.check(ompblas.stranspose_inplace(n, c.data_ptr(), n))
    assert torch.equal(c, back)
