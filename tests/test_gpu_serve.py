"""Serving loop (omb_serve_run, and the live server omb_server_*): mixed
GEMM / GEMV / dot requests admitted by HBM budget and overlapped on several
streams.  Every result must be bitwise
the same as the request's solo run through the direct entry points, whatever
ran beside it; a request whose workspace can never fit is refused."""
import numpy as np
import pytest
import torch

import helpers

pytestmark = pytest.mark.gpu


class Trace:
    """Device-resident operands for a serving trace (row-major, one
    allocation per operand so each request's buffers are independent)."""

    def __init__(self, dev, ompblas, trace, seed, mode):
        self.items = []
        for i, (kind, m, n, k, t) in enumerate(trace):
            if kind == ompblas.REQ_GEMM:
                sa, sb, sc = m * k, k * n, m * n
            elif kind == ompblas.REQ_GEMV:
                sa, sb, sc = m * n, n, m
            else:
                sa, sb, sc = n, n, 1
            a = torch.empty(max(sa, 1), device=dev)
            b = torch.empty(max(sb, 1), device=dev)
            c = torch.full((max(sc, 1),), float("nan"), device=dev)
            ompblas.fill_uniform(a, seed, 2 * i)
            ompblas.fill_uniform(b, seed, 2 * i + 1)
            self.items.append((kind, m, n, k, t, a, b, c))
        self.mode = mode
        self.ompblas = ompblas

    def requests(self, subset=None):
        R = self.ompblas.Request
        idx = range(len(self.items)) if subset is None else subset
        out = []
        for i in idx:
            kind, m, n, k, t, a, b, c = self.items[i]
            out.append(R(kind, self.mode, m, n, k, t if subset is None else 0.0,
                         a.data_ptr(), b.data_ptr(), c.data_ptr(), 0))
        return out

    def outputs(self):
        return [it[7].clone() for it in self.items]

    def clear(self):
        for it in self.items:
            it[7].fill_(float("nan"))

    def solo(self, i):
        """Request i through the direct entry point (C = A*B, y = A*x, r = x.y)."""
        p = self.ompblas
        kind, m, n, k, t, a, b, c = self.items[i]
        out = torch.full_like(c, float("nan"))
        if kind == p.REQ_GEMM:
            p.check(p.sgemm(m, n, k, 1.0, a.data_ptr(), k, 1, b.data_ptr(), n, 1, 0.0,
                            out.data_ptr(), n, 1, self.mode))
        elif kind == p.REQ_GEMV:
            p.check(p.sgemv(m, n, 1.0, a.data_ptr(), n, 1, b.data_ptr(), 0.0, out.data_ptr()))
        else:
            p.check(p.sdot(n, a.data_ptr(), b.data_ptr(), out.data_ptr()))
        return out


def small_trace(ompblas, count, rate, seed):
    return ompblas.serving_trace(count, rate, seed=seed, gemm=(32, 768), gemv=(256, 4096),
                                 dot=(1 << 10, 1 << 20))


@pytest.mark.parametrize("mode", [1, 0, 2, 3])
def test_mix_equals_solo_bitwise(dev, ompblas, mode):
    """High arrival rate, so requests overlap on the device; each output is
    bitwise equal to its solo run."""
    trace = small_trace(ompblas, 160, 2.0e5, seed=4)
    tr = Trace(dev, ompblas, trace, seed=4, mode=mode)
    res, st = ompblas.serve_run(tr.requests(), slots=8)
    torch.cuda.synchronize()
    assert all(r["status"] == 0 for r in res)
    assert st["max_in_flight"] > 1, st
    mixed = tr.outputs()
    for i in range(len(trace)):
        solo = tr.solo(i)
        assert not torch.isnan(mixed[i]).any(), i
        assert torch.equal(mixed[i], solo), (i, trace[i])
    # statistics are consistent with the per-request records
    lat = sorted(r["latency_s"] for r in res)
    assert st["p50_s"] <= st["p99_s"] <= st["max_s"] == lat[-1]
    assert all(r["done_s"] >= r["start_s"] >= 0 for r in res)
    assert all(r["start_s"] >= t[4] for r, t in zip(res, trace))
    want_flop = sum(ompblas.request_flop(k, m, n, kk) for k, m, n, kk, _ in trace)
    assert st["total_flop"] == pytest.approx(want_flop, rel=1e-12)


@pytest.mark.parametrize("mode", [1, 0, 3])
def test_batched_and_unbatched_runs_agree(dev, ompblas, mode):
    """Small requests eligible together run as grouped launches (the stats
    count them); the same trace with a stream per request gives bitwise the
    same outputs, and so does each solo call."""
    trace = small_trace(ompblas, 200, 4.0e5, seed=54)
    tr = Trace(dev, ompblas, trace, seed=54, mode=mode)
    res, st = ompblas.serve_run(tr.requests(), slots=4)
    torch.cuda.synchronize()
    assert all(r["status"] == 0 for r in res)
    assert st["batches"] > 0 and st["batched_requests"] >= 2 * st["batches"], st
    batched = tr.outputs()
    tr.clear()
    res2, st2 = ompblas.serve_run(tr.requests(), slots=4, batching=False)
    torch.cuda.synchronize()
    assert all(r["status"] == 0 for r in res2)
    assert st2["batches"] == 0 and st2["batched_requests"] == 0
    unbatched = tr.outputs()
    for i in range(len(trace)):
        assert not torch.isnan(batched[i]).any(), i
        assert torch.equal(batched[i], unbatched[i]), (i, trace[i])
        assert torch.equal(batched[i], tr.solo(i)), (i, trace[i])
    # every request of a batch shares its launch time and slot
    by_start = {}
    for r in res:
        by_start.setdefault((r["start_s"], r["slot"]), []).append(r)
    assert max(len(v) for v in by_start.values()) >= 2


def test_solo_run_through_the_loop_matches(dev, ompblas):
    """A request run alone through the loop equals the same request in the mix."""
    trace = small_trace(ompblas, 40, 1.0e5, seed=14)
    tr = Trace(dev, ompblas, trace, seed=14, mode=1)
    ompblas.serve_run(tr.requests(), slots=6)
    torch.cuda.synchronize()
    mixed = tr.outputs()
    tr.clear()
    for i in range(len(trace)):
        res, _ = ompblas.serve_run(tr.requests([i]), slots=1)
        assert res[0]["status"] == 0
    torch.cuda.synchronize()
    alone = tr.outputs()
    for i in range(len(trace)):
        assert torch.equal(mixed[i], alone[i]), (i, trace[i])


def test_small_budget_makes_requests_wait(dev, ompblas):
    """A budget of one split GEMM's workspace serialises the GEMMs; they
    wait for memory instead of failing, and results are unchanged.  (Without
    batching: batched split GEMMs split in the tile and hold no workspace.)"""
    trace = [(ompblas.REQ_GEMM, 300 + 7 * i, 200 + 5 * i, 260 + 3 * i, 0.0) for i in range(12)]
    trace += [(ompblas.REQ_DOT, 0, 5000 + i, 0, 0.0) for i in range(4)]
    tr = Trace(dev, ompblas, trace, seed=24, mode=1)
    need = max(ompblas.serve_request_bytes(r) for r in tr.requests()[:12])
    assert need >= max(ompblas.lib().omb_split_pack_bytes(m, k) +
                       ompblas.lib().omb_split_pack_bytes(n, k) for _, m, n, k, _ in trace[:12])
    res, st = ompblas.serve_run(tr.requests(), slots=8, hbm_budget=need + 4096, batching=False)
    torch.cuda.synchronize()
    assert all(r["status"] == 0 for r in res)
    assert st["admitted_late"] > 0
    mixed = tr.outputs()
    for i in range(len(trace)):
        assert torch.equal(mixed[i], tr.solo(i)), i


@pytest.mark.parametrize("mode", [0, 1, 2, 3])
def test_split_k_requests_equal_solo(dev, ompblas, mode):
    """GEMMs with few output tiles and long K take the split-K path, whose
    partial sums live in the request's workspace from the arena; overlapped
    in the loop they give the same bits as the direct entry point."""
    trace = [(ompblas.REQ_GEMM, 64 + 40 * i, 96 + 32 * i, 1024 + 700 * i, 0.0) for i in range(8)]
    trace += [(ompblas.REQ_DOT, 0, 70000 + i, 0, 0.0) for i in range(4)]
    tr = Trace(dev, ompblas, trace, seed=64, mode=mode)
    res, st = ompblas.serve_run(tr.requests(), slots=6)
    torch.cuda.synchronize()
    assert all(r["status"] == 0 for r in res)
    assert st["max_in_flight"] > 1, st
    mixed = tr.outputs()
    for i in range(len(trace)):
        assert torch.equal(mixed[i], tr.solo(i)), (i, trace[i])


def test_oversize_request_is_refused(dev, ompblas):
    """A request whose workspace exceeds the budget gets EBUSY; the rest run
    (without batching, where the split GEMM would need no workspace)."""
    trace = [(ompblas.REQ_GEMM, 512, 512, 512, 0.0), (ompblas.REQ_GEMV, 300, 400, 0, 0.0),
             (ompblas.REQ_DOT, 0, 1000, 0, 0.0)]
    tr = Trace(dev, ompblas, trace, seed=34, mode=1)
    res, st = ompblas.serve_run(tr.requests(), slots=2, hbm_budget=64 * 1024, batching=False)
    torch.cuda.synchronize()
    assert res[0]["status"] == ompblas.EBUSY
    assert res[0]["ws_bytes"] > 64 * 1024
    assert res[1]["status"] == 0 and res[2]["status"] == 0
    out = tr.outputs()
    assert torch.equal(out[1], tr.solo(1)) and torch.equal(out[2], tr.solo(2))
    assert torch.isnan(out[0]).all()


def test_split_gemm_requests_within_error_bound(dev, ompblas):
    """Split-BF16 requests keep the <= 2x plain FP32 error bound vs FP64."""
    trace = [(ompblas.REQ_GEMM, 777, 640, 1500, 0.0), (ompblas.REQ_GEMM, 64, 2048, 4096, 0.0)]
    tr = Trace(dev, ompblas, trace, seed=44, mode=1)
    ompblas.serve_run(tr.requests(), slots=2)
    torch.cuda.synchronize()
    for i, (_, m, n, k, _) in enumerate(trace):
        _, _, _, _, _, a, b, c = tr.items[i]
        A, B = a.view(m, k), b.view(k, n)
        c64 = ompblas.check_dgemm(A, B)
        split = ompblas.rel_errors(c.view(m, n), c64)
        plain = torch.zeros(m, n, device=dev)
        ompblas.mtm(plain, A, B)()
        ref = ompblas.rel_errors(plain, c64)
        assert split[0] <= 2 * ref[0] and split[1] <= 2 * ref[1], (split, ref)


def test_empty_and_degenerate_requests(dev, ompblas):
    trace = [(ompblas.REQ_GEMM, 5, 7, 0, 0.0), (ompblas.REQ_GEMV, 9, 0, 0, 0.0),
             (ompblas.REQ_DOT, 0, 0, 0, 0.0), (ompblas.REQ_GEMM, 0, 4, 4, 0.0)]
    tr = Trace(dev, ompblas, trace, seed=54, mode=1)
    res, _ = ompblas.serve_run(tr.requests(), slots=2)
    torch.cuda.synchronize()
    assert all(r["status"] == 0 for r in res)
    out = tr.outputs()
    assert torch.equal(out[0], torch.zeros(35, device=dev))
    assert torch.equal(out[1], torch.zeros(9, device=dev))
    assert torch.equal(out[2], torch.zeros(1, device=dev))
    assert torch.isnan(out[3]).all()      # M == 0: nothing written
    res, st = ompblas.serve_run([], slots=2)
    assert res == [] and st["p99_s"] == 0.0


def integer_operands(tr, lo, hi):
    """Replace every request's operands by integers: U[-1,1) scaled to
    [lo, hi] and rounded in place (GEMM and GEMV entries), or rounded to
    {-1, 0, 1} (dot entries, whose length reaches 2^24)."""
    p = tr.ompblas
    for kind, m, n, k, t, a, b, c in tr.items:
        for v in (a, b):
            if kind == p.REQ_DOT:
                v.round_()
            else:
                v.mul_((hi - lo) / 2.0).add_((hi + lo) / 2.0).round_()


@pytest.mark.parametrize("mode", [1, 0, 3])
def test_full_range_requests_give_exact_integer_answers(dev, ompblas, mode):
    """Requests drawn at the serving configuration's full ranges (GEMM M, N,
    K up to 4096, GEMV up to 16384 x 16384, dot up to 2^24) with integer
    operands: every output of the loop equals the exact integer result --
    |entries| <= 4 keeps every GEMM / GEMV partial sum below 2^24, and dot
    entries in {-1, 0, 1} keep its sums at most 2^24, so FP32 holds them
    exactly and the split modes carry small integers without rounding."""
    trace = ompblas.serving_trace(120, 4.0e4, seed=40 + mode)
    # the largest of each kind among the trace is included
    trace += [(ompblas.REQ_GEMM, 4096, 4096, 4096, trace[-1][4]),
              (ompblas.REQ_GEMV, 16384, 16384, 0, trace[-1][4]),
              (ompblas.REQ_DOT, 0, 1 << 24, 0, trace[-1][4])]
    tr = Trace(dev, ompblas, trace, seed=40 + mode, mode=mode)
    integer_operands(tr, -4, 4)
    torch.cuda.synchronize()
    res, st = ompblas.serve_run(tr.requests(), slots=8)
    torch.cuda.synchronize()
    assert all(r["status"] == 0 for r in res)
    assert st["max_in_flight"] > 1, st
    for i, (kind, m, n, k, t, a, b, c) in enumerate(tr.items):
        if kind == ompblas.REQ_GEMM:
            want = a.view(m, k).double() @ b.view(k, n).double()
            got = c.view(m, n).double()
        elif kind == ompblas.REQ_GEMV:
            want = a.view(m, n).double() @ b.double()
            got = c.double()
        else:
            want = torch.dot(a.double(), b.double()).view(1)
            got = c.double()
        assert torch.equal(got, want), (i, trace[i])


def test_serving_config_trace_matches_solo_and_oracle(dev, ompblas):
    """The serving configuration's trace as the benchmark draws it (seed 4,
    default ranges: GEMM 32-4096, GEMV 256-16384, dot 2^10-2^24) at 300
    requests and the stated 32,000 requests/s, split-BF16 GEMMs: every output
    is bitwise its solo run; the three largest GEMMs, GEMVs and dots are
    within 2x the oracle's error against an FP64 product on the host (the
    reference algorithm on the same inputs: include/mtm.hpp:116-206,
    include/mtv.hpp:72-98, include/dot.hpp:143-174)."""
    trace = ompblas.serving_trace(300, 3.2e4, seed=4)
    tr = Trace(dev, ompblas, trace, seed=4, mode=ompblas.SPLIT_BF16)
    res, st = ompblas.serve_run(tr.requests(), slots=8)
    torch.cuda.synchronize()
    assert all(r["status"] == 0 for r in res)
    print({k: st[k] for k in ("p50_s", "p99_s", "max_in_flight", "wall_s")})
    mixed = tr.outputs()
    for i in range(len(trace)):
        assert torch.equal(mixed[i], tr.solo(i)), (i, trace[i])

    def largest(kind, size):
        idx = [i for i, t in enumerate(trace) if t[0] == kind]
        return sorted(idx, key=lambda i: size(trace[i]), reverse=True)[:3]

    for i in largest(ompblas.REQ_GEMM, lambda t: t[1] * t[2] * t[3]):
        _, m, n, k, _, a, b, c = tr.items[i]
        ah, bh = a.view(m, k).cpu().numpy(), b.view(k, n).cpu().numpy()
        truth = ah.astype(np.float64) @ bh.astype(np.float64)
        c_ref = np.zeros((m, n), np.float32)
        helpers.oracle_mtm(c_ref, ah, bh)
        ref = helpers.rel_errors(c_ref, truth)
        got = helpers.rel_errors(mixed[i].view(m, n).cpu().numpy(), truth)
        print("gemm", (m, n, k), "gpu", got, "oracle", ref)
        assert got[0] <= 2 * ref[0] and got[1] <= 2 * ref[1], (i, trace[i], got, ref)
    for i in largest(ompblas.REQ_GEMV, lambda t: t[1] * t[2]):
        _, m, n, k, _, a, b, c = tr.items[i]
        ah, xh = a.view(m, n).cpu().numpy(), b.cpu().numpy()
        truth = ah.astype(np.float64) @ xh.astype(np.float64)
        y_ref = np.zeros(m, np.float32)
        helpers.oracle_mtv(y_ref, ah, xh)
        ref = helpers.rel_errors(y_ref, truth)
        got = helpers.rel_errors(mixed[i].cpu().numpy(), truth)
        print("gemv", (m, n), "gpu", got, "oracle", ref)
        assert got[0] <= 2 * ref[0] and got[1] <= 2 * ref[1], (i, trace[i], got, ref)
    for i in largest(ompblas.REQ_DOT, lambda t: t[2]):
        _, m, n, k, _, a, b, c = tr.items[i]
        xh, yh = a.cpu().numpy(), b.cpu().numpy()
        truth = float(np.dot(xh.astype(np.float64), yh.astype(np.float64)))
        absdot = float(np.dot(np.abs(xh).astype(np.float64), np.abs(yh).astype(np.float64)))
        ref = abs(helpers.oracle_dot(xh, yh) - truth)
        got = abs(float(mixed[i].item()) - truth)
        print("dot", n, "gpu", got, "oracle", ref)
        # one scalar's error can be small by chance: the larger of 2x the
        # oracle's and a fixed fraction of the summation bound n u sum|xy|
        assert got <= max(2 * ref, 1e-3 * n * 2.0 ** -24 * absdot), (i, got, ref)


@pytest.mark.parametrize("mode", [1, 0])
def test_live_server_matches_solo_bitwise(dev, ompblas, mode):
    """omb_server_*: requests submitted from two threads at once are admitted
    by the library's worker as they come (slots, arena and batching as in
    omb_serve_run); every output is bitwise equal to its solo run, each
    ticket is published once, and close() reports every request."""
    import threading
    trace = small_trace(ompblas, 160, 2.0e5, seed=24)
    tr = Trace(dev, ompblas, trace, seed=24, mode=mode)
    reqs = tr.requests()
    torch.cuda.synchronize()  # operands complete before they are submitted
    max_bytes = max(ompblas.serve_request_bytes(r) for r in reqs)
    srv = ompblas.Server(slots=8, max_request_bytes=max_bytes)
    tickets = [None] * len(reqs)

    def submit(part):
        for i in part:
            tickets[i] = srv.submit(reqs[i])

    threads = [threading.Thread(target=submit, args=(range(j, len(reqs), 2),))
               for j in range(2)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert sorted(tickets) == list(range(len(reqs)))
    results = [srv.wait(t) for t in tickets]
    assert srv.poll(tickets[0]) == results[0]
    st = srv.close()
    torch.cuda.synchronize()
    assert all(r["status"] == 0 for r in results), results
    assert all(r["done_s"] >= r["start_s"] >= 0 and r["latency_s"] >= 0 for r in results)
    mixed = tr.outputs()
    for i in range(len(reqs)):
        assert not torch.isnan(mixed[i]).any(), i
        assert torch.equal(mixed[i], tr.solo(i)), (i, trace[i])
    want_flop = sum(ompblas.request_flop(k, m, n, kk) for k, m, n, kk, _ in trace)
    assert st["total_flop"] == pytest.approx(want_flop, rel=1e-12)
    assert st["max_in_flight"] > 1, st
    assert st["p50_s"] <= st["p99_s"] <= st["max_s"]


def test_live_server_refuses_what_never_fits(dev, ompblas):
    """A request whose workspace exceeds the server's arena comes back with
    OMB_EBUSY; the requests around it run and match their solo runs."""
    trace = [(ompblas.REQ_GEMM, 512, 512, 512, 0.0), (ompblas.REQ_DOT, 0, 4096, 0, 0.0),
             (ompblas.REQ_GEMV, 300, 700, 0, 0.0)]
    tr = Trace(dev, ompblas, trace, seed=25, mode=1)
    reqs = tr.requests()
    torch.cuda.synchronize()
    assert ompblas.serve_request_bytes(reqs[0]) > 1 << 20
    with ompblas.Server(slots=1, max_request_bytes=1 << 20) as srv:
        tickets = [srv.submit(r) for r in reqs]
        results = [srv.wait(t) for t in tickets]
    torch.cuda.synchronize()
    assert results[0]["status"] == ompblas.EBUSY
    assert results[1]["status"] == 0 and results[2]["status"] == 0
    out = tr.outputs()
    assert torch.isnan(out[0]).all()
    for i in (1, 2):
        assert torch.equal(out[i], tr.solo(i)), i


def test_live_server_wakes_from_sleep(dev, ompblas):
    """Without spin_poll the idle worker sleeps until a submission arrives:
    requests submitted 2 ms apart (the worker idle in between) each complete
    and match their solo runs; polling an unfinished or finished ticket
    works from the submitting thread."""
    import time
    trace = small_trace(ompblas, 12, 1.0e3, seed=26)
    tr = Trace(dev, ompblas, trace, seed=26, mode=0)
    reqs = tr.requests()
    torch.cuda.synchronize()
    max_bytes = max(ompblas.serve_request_bytes(r) for r in reqs)
    with ompblas.Server(slots=2, max_request_bytes=max_bytes, spin_poll=False) as srv:
        results = []
        for q in reqs:
            t = srv.submit(q)
            r = srv.wait(t)
            assert srv.poll(t) == r
            results.append(r)
            time.sleep(0.002)
        st = srv.close()
    torch.cuda.synchronize()
    assert all(r["status"] == 0 for r in results)
    assert st["max_in_flight"] == 1
    out = tr.outputs()
    for i in range(len(reqs)):
        assert torch.equal(out[i], tr.solo(i)), (i, trace[i])
