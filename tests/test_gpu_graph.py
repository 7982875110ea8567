"""Direct calls captured in a hipGraph (through torch.cuda.graph) and
replayed: after one eager call on the capture stream (which sizes that
stream's workspace), omb_sgemm in every mode (split-K partials and the
scaled FP16 maxima included), omb_sgemv in both layouts and omb_sdot record
only kernels, and every replay gives the eager call's bits, with other
eager work between the replays."""
import pytest
import torch

pytestmark = pytest.mark.gpu


def _replays_equal(dev, make_call, out, reps=3):
    # the inputs were filled on the default stream; the capture stream is
    # not ordered after it
    torch.cuda.synchronize()
    s = torch.cuda.Stream()
    # eager reference on the capture stream (also sizes its workspace)
    with torch.cuda.stream(s):
        out.fill_(float("nan"))
    make_call(s)
    torch.cuda.synchronize()
    want = out.clone()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        make_call(s)
    torch.cuda.synchronize()
    for r in range(reps):
        out.fill_(float("nan"))              # eager work between replays
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        if not torch.equal(out, want):
            d = (out - want).abs().nan_to_num(float("inf"))
            bad = torch.nonzero(d > 0).flatten()
            raise AssertionError(f"replay {r}: {bad.numel()} of {out.numel()} differ, first "
                                 f"{bad[:8].tolist()}, nan {int(torch.isnan(out).sum())}, "
                                 f"got {out.flatten()[bad[:4]].tolist()} want "
                                 f"{want.flatten()[bad[:4]].tolist()}")
        scratch = torch.empty(1 << 20, device=dev).uniform_()   # unrelated eager kernels
        scratch.sum().item()
    return want


@pytest.mark.parametrize("mode", [0, 1, 2, 3])
@pytest.mark.parametrize("shape", [(300, 200, 3000), (512, 384, 1000), (129, 77, 31)])
def test_sgemm_graph_replays_equal_eager(dev, ompblas, mode, shape):
    M, N, K = shape
    a = torch.empty(M, K, device=dev)
    b = torch.empty(K, N, device=dev)
    ompblas.fill_uniform(a, 21, 0)
    ompblas.fill_uniform(b, 21, 1)
    c = torch.empty(M, N, device=dev)

    def call(s):
        ompblas.check(ompblas.sgemm(M, N, K, 1.0, a.data_ptr(), K, 1, b.data_ptr(), N, 1, 0.0,
                                    c.data_ptr(), N, 1, mode, stream=s))

    want = _replays_equal(dev, call, c)
    assert not torch.isnan(want).any()


@pytest.mark.parametrize("layout", ["row", "col"])
def test_sgemv_and_sdot_graph_replays_equal_eager(dev, ompblas, layout):
    M, N = 3000, 5000
    a = torch.empty(M * N, device=dev)
    x = torch.empty(N, device=dev)
    ompblas.fill_uniform(a, 22, 0)
    ompblas.fill_uniform(x, 22, 1)
    y = torch.empty(M, device=dev)
    rs, cs = (N, 1) if layout == "row" else (1, M)

    def gemv(s):
        ompblas.check(ompblas.sgemv(M, N, 1.0, a.data_ptr(), rs, cs, x.data_ptr(), 0.0,
                                    y.data_ptr(), stream=s))

    _replays_equal(dev, gemv, y)
    n = 1 << 27                                   # FP64 combine (n >= 2^26)
    u = torch.empty(n, device=dev)
    v = torch.empty(n, device=dev)
    ompblas.fill_uniform(u, 23, 0)
    ompblas.fill_uniform(v, 23, 1)
    r = torch.empty(1, device=dev)

    def dot(s):
        ompblas.check(ompblas.sdot(n, u.data_ptr(), v.data_ptr(), r.data_ptr(), stream=s))

    _replays_equal(dev, dot, r)
