"""mtv / vtm (SGEMV) and dot (SDOT) on the GPU through the C-ABI against
the oracle: integer-exact reference cases (test/test.mtv.cpp, test.vtm.cpp,
test.dot_prod.cpp) and FP64-error bounds at config-2 sizes."""
import numpy as np
import pytest
import torch

import helpers
from helpers import oracle_dot, oracle_mtv, rand100

pytestmark = pytest.mark.gpu

MTV_SIZES = list(range(2, 64)) + list(range(64, 512, 11)) + [511]


def to_dev(a, dev):
    if a.ndim == 2 and a.flags.f_contiguous and not a.flags.c_contiguous:
        return torch.from_numpy(np.ascontiguousarray(a.T)).to(dev).t()
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


@pytest.mark.parametrize("order", ["F", "C"])
def test_mtv_integer_exact_reference_cases(dev, ompblas, order):
    """test/test.mtv.cpp:29-126 via the torch mirror amt.mtv(c, A, v)():
    first order accumulates into c, last order overwrites it."""
    for sz in MTV_SIZES:
        a = rand100((sz, sz), order)
        x = rand100((sz,))
        y0 = rand100((sz,))
        y = y0.copy()
        oracle_mtv(y, a, x)
        ty = torch.from_numpy(y0).to(dev)
        ompblas.mtv(ty, to_dev(a, dev), torch.from_numpy(x).to(dev))()
        assert np.array_equal(ty.cpu().numpy(), y), (order, sz)


@pytest.mark.parametrize("order", ["F", "C"])
def test_vtm_integer_exact_true_transpose(dev, ompblas, order):
    """test/test.vtm.cpp:29-125: c = A^T v; rectangular shapes as well."""
    shapes = [(s, s) for s in MTV_SIZES[::2]] + [(5, 9), (40, 3), (1, 17), (17, 1), (300, 1000)]
    for m, n in shapes:
        a = rand100((m, n), order)
        x = rand100((m,))
        ty = torch.zeros(n, device=dev)
        ompblas.vtm(ty, to_dev(a, dev), torch.from_numpy(x).to(dev))()
        want = a.astype(np.int64).T @ x.astype(np.int64)
        assert np.array_equal(ty.cpu().numpy().astype(np.int64), want), (order, m, n)


@pytest.mark.parametrize("order", ["F", "C"])
def test_gemv_unaligned_and_ld(dev, ompblas, order):
    big = torch.from_numpy(rand100((403, 517), order)).to(dev) if order == "C" else \
        to_dev(rand100((403, 517), "F"), dev)
    a = big[3:301, 1:390]                  # 298 x 389; offsets break 16-byte alignment
    x = torch.from_numpy(rand100((389,))).to(dev)
    y = torch.zeros(298, device=dev)
    helpers.assert_view_in_bounds(a, 298, 389, a.stride(0), a.stride(1))
    ompblas.check(ompblas.sgemv(298, 389, 1.0, a.data_ptr(), a.stride(0), a.stride(1),
                                x.data_ptr(), 0.0, y.data_ptr()))
    want = a.cpu().numpy().astype(np.int64) @ x.cpu().numpy().astype(np.int64)
    assert np.array_equal(y.cpu().numpy().astype(np.int64), want)


@pytest.mark.parametrize("order", ["F", "C"])
def test_gemv_config2_error_vs_reference(dev, ompblas, order):
    """Config 2: A 16384 x 16384 and x from seed 1, both layouts.  GPU error
    vs FP64 <= 2x the oracle's (reference algorithm, this host's blocking)."""
    n = 16384
    a = helpers.rng_uniform(n * n, 1, 0).reshape((n, n), order=order)
    x = helpers.rng_uniform(n, 1, 1)
    ta, tx = to_dev(a, dev), torch.from_numpy(x).to(dev)
    y64 = torch.empty(n, dtype=torch.float64, device=dev)
    ompblas.check(ompblas.lib().omb_check_dgemv(n, n, ta.data_ptr(), ta.stride(0), ta.stride(1),
                                                tx.data_ptr(), None, y64.data_ptr(), None))
    truth = y64.cpu().numpy()
    # validate the FP64 check kernel on 64 rows against numpy FP64
    rows = np.arange(0, n, 257)
    np.testing.assert_allclose(truth[rows], a[rows].astype(np.float64) @ x.astype(np.float64),
                               rtol=1e-12, atol=1e-12)
    y_ref = np.zeros(n, np.float32)
    oracle_mtv(y_ref, a, x)
    ref = helpers.rel_errors(y_ref, truth)
    ty = torch.zeros(n, device=dev)
    ompblas.check(ompblas.sgemv(n, n, 1.0, ta.data_ptr(), ta.stride(0), ta.stride(1),
                                tx.data_ptr(), 0.0, ty.data_ptr()))
    got = helpers.rel_errors(ty.cpu().numpy(), truth)
    print(order, "gpu", got, "oracle", ref)
    assert got[0] <= 2 * ref[0] and got[1] <= 2 * ref[1]


def test_gemv_deterministic_and_beta(dev, ompblas):
    M, N = 5000, 70000
    a = torch.empty((N, M), device=dev).t()     # column-major, split over column chunks
    ompblas.fill_uniform(a.t(), 8, 0)
    x = torch.empty(N, device=dev)
    ompblas.fill_uniform(x, 8, 1)
    y0 = torch.empty(M, device=dev)
    ompblas.fill_uniform(y0, 8, 2)
    helpers.assert_view_in_bounds(a, M, N, a.stride(0), a.stride(1))
    outs = []
    for _ in range(3):
        y = y0.clone()
        ompblas.check(ompblas.sgemv(M, N, 0.5, a.data_ptr(), a.stride(0), a.stride(1),
                                    x.data_ptr(), 2.0, y.data_ptr()))
        outs.append(y)
    assert torch.equal(outs[0], outs[1]) and torch.equal(outs[0], outs[2])
    want = 0.5 * (a.double() @ x.double()) + 2.0 * y0.double()
    assert torch.allclose(outs[0].double(), want, rtol=0, atol=1e-3)


def test_gemv_column_major_single_pass_rows(dev, ompblas):
    """Column-major A of whole 64-row blocks (at least one per CU) and whole
    256-column rounds runs one pass of 64-row workgroups; a ragged shape of
    the same size runs the column-chunked kernel.  An offset start, lda > M
    and beta != 0, on integer values: exact against int64, and the same bits
    at every alignment."""
    cus = torch.cuda.get_device_properties(dev).multi_processor_count
    for M, N in ((64 * cus, 256), (64 * cus + 64, 512), (64 * cus + 37, 301), (64 * cus, 257)):
        ys = []
        y0 = rand100((M,))
        x = torch.from_numpy(rand100((N,))).to(dev)
        a_np = rand100((M, N), "F")
        for off, pad in ((0, 0), (1, 3), (3, 1)):
            ld = M + pad
            buf = torch.zeros(off + N * ld, device=dev)
            view = buf[off:].view(N, ld)[:, :M]           # column-major M x N
            view.copy_(torch.from_numpy(np.ascontiguousarray(a_np.T)).to(dev))
            a = view.t()
            helpers.assert_view_in_bounds(a, M, N, a.stride(0), a.stride(1))
            y = torch.from_numpy(y0).to(dev)
            ompblas.check(ompblas.sgemv(M, N, 1.0, a.data_ptr(), a.stride(0), a.stride(1),
                                        x.data_ptr(), 3.0, y.data_ptr()))
            ys.append(y)
        want = a_np.astype(np.int64) @ x.cpu().numpy().astype(np.int64) + 3 * y0.astype(np.int64)
        assert np.array_equal(ys[0].cpu().numpy().astype(np.int64), want), (M, N)
        for y in ys[1:]:
            assert torch.equal(y, ys[0]), (M, N)


# ----------------------------------------------------------------- dot ----
DOT_SIZES = list(range(2, 200)) + list(range(200, (1 << 17) + 2, 1009)) + [(1 << 17) + 1]


def test_dot_closed_form(dev, ompblas):
    """test/test.dot_prod.cpp:10-28: a = 1..s, a.a = s(s+1)(2s+1)/6, within the
    summation bound (exact while the sum stays below 2^24)."""
    full = torch.arange(1, (1 << 17) + 2, dtype=torch.float32, device=dev)
    r = torch.zeros(1, device=dev)
    u = 2.0 ** -24
    for s in DOT_SIZES:
        a = full[:s]
        ompblas.dot_prod(r, a, a)()
        got = float(r.item())
        exact = s * (s + 1) * (2 * s + 1) // 6
        if exact < (1 << 24):
            assert got == exact, s
        else:
            assert abs(got - exact) <= (s * u / (1 - s * u)) * exact, s
        # src/dot_tuning.cpp's entry: block sizes accepted, same bits
        rt = torch.full((1,), -1.0, device=dev)
        ompblas.dot_prod_tuning_param(rt, a, a, 1024, 4096, 1 << 16)()
        assert torch.equal(rt, r), s


def test_dot_config2_error_vs_reference(dev, ompblas):
    """Config 2: two 2^28 vectors from seed 1 (FP64 cross-workgroup combine)."""
    n = 1 << 28
    x = torch.empty(n, device=dev)
    y = torch.empty(n, device=dev)
    ompblas.fill_uniform(x, 1, 0)
    ompblas.fill_uniform(y, 1, 1)
    r64 = torch.empty(1, dtype=torch.float64, device=dev)
    ompblas.check(ompblas.lib().omb_check_ddot(n, x.data_ptr(), y.data_ptr(), r64.data_ptr(), None))
    truth = float(r64.item())
    r = torch.zeros(1, device=dev)
    ompblas.dot_prod(r, x, y)()
    got = float(r.item())
    xs, ys = x.cpu().numpy(), y.cpu().numpy()
    ref = oracle_dot(xs, ys)
    t_np = float(np.dot(xs.astype(np.float64), ys.astype(np.float64)))
    assert abs(truth - t_np) <= 1e-9 * abs(t_np) + 1e-6
    absdot = float(np.dot(np.abs(xs).astype(np.float64), np.abs(ys).astype(np.float64)))
    print("dot gpu", got, "oracle", ref, "fp64", truth)
    # a single scalar's error can be small by chance: bound it by the larger
    # of 2x the oracle's and a fixed fraction of the summation bound
    bound = max(2 * abs(ref - truth), 1e-3 * (n * 2.0 ** -24) * absdot)
    assert abs(got - truth) <= bound
    ompblas.dot_prod(r, x, y)()
    assert float(r.item()) == got     # bitwise repeatable


def test_dot_edge_cases(dev, ompblas):
    r = torch.full((1,), 5.0, device=dev)
    ompblas.check(ompblas.sdot(0, 0, 0, r.data_ptr()))
    assert float(r.item()) == 0.0
    v = torch.arange(1, 40, dtype=torch.float32, device=dev)
    for off in range(1, 4):                      # unaligned starts
        a = v[off:off + 30]
        ompblas.dot_prod(r, a, a)()
        want = sum(float(t) ** 2 for t in range(off + 1, off + 31))
        assert float(r.item()) == want


@pytest.mark.parametrize("order", ["F", "C"])
def test_gemv_bits_do_not_depend_on_alignment(dev, ompblas, order):
    """The same values at 16-byte aligned and at 4/8/12-byte offset addresses,
    with lda a multiple of 4 or not, give bitwise the same y."""
    M, N = 1037, 2051
    src = torch.empty(M * N, device=dev)
    ompblas.fill_uniform(src, 9, 0)
    x0 = torch.empty(N, device=dev)
    ompblas.fill_uniform(x0, 9, 1)
    rows, cols = (M, N) if order == "C" else (N, M)
    ys = []
    for off, pad in ((0, 0), (1, 0), (2, 3), (3, 1)):
        ld = cols + pad
        buf = torch.zeros(off + rows * ld, device=dev)
        view = buf[off:].view(rows, ld)[:, :cols]
        view.copy_(src.view(rows, cols))
        a = view if order == "C" else view.t()          # M x N either way
        xb = torch.zeros(off + N, device=dev)
        xb[off:].copy_(x0)
        x = xb[off:]
        y = torch.zeros(M, device=dev)
        helpers.assert_view_in_bounds(a, M, N, a.stride(0), a.stride(1))
        ompblas.check(ompblas.sgemv(M, N, 1.0, a.data_ptr(), a.stride(0), a.stride(1),
                                    x.data_ptr(), 0.0, y.data_ptr()))
        ys.append(y)
    for y in ys[1:]:
        assert torch.equal(y, ys[0])
    a64 = src.view(rows, cols).double()
    want = (a64 if order == "C" else a64.t()) @ x0.double()
    assert torch.allclose(ys[0].double(), want, rtol=0, atol=1e-3)


def test_tall_row_gemv_bits_equal_the_grouped_item(dev, ompblas):
    """A row-major GEMV tall enough for 16-row wavefront groups (at least 16
    rows per wavefront on every CU) forms its addresses as scalar row bases
    plus 32-bit lane offsets, as the grouped kernel's GEMV item does; a
    short call (4- or 1-row groups) keeps 64-bit addresses.  Ragged N (tail
    lanes), lda not a multiple of 4 and 4-byte offset A and x: the tall call,
    the grouped item and the short call on the first rows give bitwise the
    same y, within rounding of an FP64 product."""
    M, N, ld = 16384 + 37, 2051, 2054
    buf = torch.empty(1 + M * ld, device=dev)
    ompblas.fill_uniform(buf, 21, 0)
    a = buf[1:].view(M, ld)[:, :N]
    xb = torch.empty(1 + N, device=dev)
    ompblas.fill_uniform(xb, 21, 1)
    x = xb[1:]
    helpers.assert_view_in_bounds(a, M, N, a.stride(0), a.stride(1))
    y_direct = torch.zeros(M, device=dev)
    ompblas.check(ompblas.sgemv(M, N, 1.5, a.data_ptr(), ld, 1, x.data_ptr(), 0.0,
                                y_direct.data_ptr()))
    y_group = torch.zeros(M, device=dev)
    plan = ompblas.GroupPlan([], [(M, N, a.data_ptr(), ld, x.data_ptr(), y_group.data_ptr(),
                                   1.5, 0.0)], ompblas.PLAIN_F32)
    plan.launch()
    torch.cuda.synchronize()
    plan.close()
    assert torch.equal(y_direct, y_group)
    for rows in (5000, 97):  # 4-row and 1-row groups on a 256-CU device
        y_short = torch.zeros(rows, device=dev)
        ompblas.check(ompblas.sgemv(rows, N, 1.5, a.data_ptr(), ld, 1, x.data_ptr(), 0.0,
                                    y_short.data_ptr()))
        assert torch.equal(y_short, y_direct[:rows]), rows
    want = 1.5 * (a.double() @ x.double())
    assert torch.allclose(y_direct.double(), want, rtol=0, atol=1e-3)


def test_dot_bits_do_not_depend_on_alignment(dev, ompblas):
    n = (1 << 20) + 7
    x0 = torch.empty(n, device=dev)
    y0 = torch.empty(n, device=dev)
    ompblas.fill_uniform(x0, 9, 2)
    ompblas.fill_uniform(y0, 9, 3)
    rs = []
    for ox, oy in ((0, 0), (1, 0), (0, 3), (2, 1), (3, 3)):
        xb = torch.zeros(ox + n, device=dev)
        yb = torch.zeros(oy + n, device=dev)
        xb[ox:].copy_(x0)
        yb[oy:].copy_(y0)
        r = torch.zeros(1, device=dev)
        ompblas.check(ompblas.sdot(n, xb[ox:].data_ptr(), yb[oy:].data_ptr(), r.data_ptr()))
        rs.append(float(r.item()))
    assert all(r == rs[0] for r in rs), rs
    want = float(torch.dot(x0.double(), y0.double()))
    assert abs(rs[0] - want) <= 1e-4 * float(torch.dot(x0.abs().double(), y0.abs().double()))
