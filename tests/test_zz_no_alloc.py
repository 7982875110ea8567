"""omb_sdot_noalloc (partials in the library's static device scratch, no
allocation) against omb_sdot through the C ABI: bitwise the same result on
consecutive calls, each later call smaller than the one before, so every
call must overwrite what it reads of the shared scratch (FP64 and FP32
combines, a single-workgroup length, the empty vector)."""
import pytest
import torch

pytestmark = pytest.mark.gpu


def test_noalloc_dot_matches_dot(dev, ompblas):
    lib = ompblas.lib()
    nmax = (1 << 26) + 12345                 # FP64 combine, 1024 partials
    x = torch.empty(nmax, device=dev)
    y = torch.empty(nmax, device=dev)
    ompblas.fill_uniform(x, 71, 0)
    ompblas.fill_uniform(y, 71, 1)
    torch.cuda.synchronize()
    for n in (nmax, 300007, 4099, 5, 0):     # each call smaller than the last
        print("omb_sdot_noalloc n =", n, flush=True)
        want = torch.full((1,), float("nan"), device=dev)
        got = torch.full((1,), float("nan"), device=dev)
        ompblas.check(lib.omb_sdot(n, x.data_ptr(), y.data_ptr(), want.data_ptr(), None))
        torch.cuda.synchronize()
        ompblas.check(lib.omb_sdot_noalloc(n, x.data_ptr(), y.data_ptr(), got.data_ptr(), None))
        torch.cuda.synchronize()
        assert torch.equal(got, want), (n, float(got.item()), float(want.item()))
        if n:
            ref = float(torch.dot(x[:n].double(), y[:n].double()))
            assert abs(float(got.item()) - ref) <= 1e-5 * float(
                torch.dot(x[:n].abs().double(), y[:n].abs().double()))
    assert lib.omb_sdot_noalloc(-1, x.data_ptr(), y.data_ptr(), got.data_ptr(), None) == ompblas.EINVAL


def test_noalloc_dot_on_a_side_stream(dev, ompblas):
    """On a non-default stream, after a larger call on the default stream:
    same bits as omb_sdot on that stream."""
    lib = ompblas.lib()
    n = 1 << 22
    x = torch.empty(n, device=dev)
    y = torch.empty(n, device=dev)
    ompblas.fill_uniform(x, 72, 0)
    ompblas.fill_uniform(y, 72, 1)
    s = torch.cuda.Stream()
    torch.cuda.synchronize()
    outs = []
    for fn in (lib.omb_sdot, lib.omb_sdot_noalloc):
        r = torch.full((1,), float("nan"), device=dev)
        ompblas.check(fn(n - 3, x.data_ptr(), y.data_ptr(), r.data_ptr(), s.cuda_stream))
        s.synchronize()
        outs.append(r)
    assert torch.equal(outs[0], outs[1])
