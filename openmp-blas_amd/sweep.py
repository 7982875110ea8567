"""Size sweeps of the GPU kernels in the reference's benchmark format.

The reference's drivers (src/mtm.cpp, src/mtv.cpp, src/dot.cpp,
src/outer.cpp, src/trans.cpp) run each routine over a range of sizes, time
MaxIter calls per size (include/benchmark.hpp:34-51) and collect GFLOP/s in
an amt::metric (include/metric.hpp) that prints min / max / average and peak
utilisation (str(), :435-464) and writes one CSV column per routine (csv(),
:478-500).  This module does the same on one MI355X: the same size ranges,
MaxIter and operation counts as the drivers' main(), operands resident in
HBM and timed with HIP events on the launch stream (one untimed call first).
Peaks are the MI355X rooflines expressed in each routine's unit (FP32 vector
peak for mtm; 8 TB/s times operations per byte for the bandwidth-bound ones).

    python openmp-blas_amd/sweep.py --op mtm [--csv out.csv] [--max 1024]
"""
from __future__ import annotations

import argparse
import os
import sys
from typing import Callable, Dict, List, Optional

HBM = 8.0e12
FP32_PEAK_GFLOPS = 157.3e3


def reference_range(start: float, end: float, stride: float, multiply: bool) -> List[float]:
    """amt::range(x, start, end, stride, fn) (include/range.hpp:29-44):
    start, fn(start, stride), ... while < end."""
    if start > end:
        raise RuntimeError("amt::range: start > end")
    out, v = [], start
    while v < end:
        out.append(v)
        v = v * stride if multiply else v + stride
    return out


# op -> (range args, MaxIter, operations per size, peak in the same unit, unit)
SWEEPS: Dict[str, dict] = {
    # src/mtm.cpp:372-375: range(32, 3*1024, 32, +), MaxIter 4; ops M N (2K - 1)
    "mtm": dict(range=(32.0, 3072.0, 32.0, False), iters=4,
                ops=lambda n: n * n * (2.0 * n - 1.0), peak=FP32_PEAK_GFLOPS, unit="GFlops"),
    # src/mtv.cpp:332-336: range(2, 2^16, 2, *), MaxIter 10; ops M (2N - 1)
    "mtv": dict(range=(2.0, 65536.0, 2.0, True), iters=10,
                ops=lambda n: n * (2.0 * n - 1.0), peak=HBM / 4.0 * 2.0 / 1e9, unit="GFlops"),
    # src/dot.cpp:198-202: range(2, 2^20, 1024, +), MaxIter 100; ops 2n
    "dot": dict(range=(2.0, float(1 << 20), 1024.0, False), iters=100,
                ops=lambda n: 2.0 * n, peak=HBM / 8.0 * 2.0 / 1e9, unit="GFlops"),
    # src/outer.cpp:320-322: range(32, 16382, 32, +), MaxIter 4; ops n^2
    "outer": dict(range=(32.0, 16382.0, 32.0, False), iters=4,
                  ops=lambda n: n * n, peak=HBM / 8.0 / 1e9, unit="GFlops"),
    # src/trans.cpp:438-441: range(32, 16 Ki, 32, +), MaxIter 4; ops M N sizeof(float)
    "trans": dict(range=(32.0, 16384.0, 32.0, False), iters=4,
                  ops=lambda n: n * n * 4.0, peak=HBM / 2.0 / 1e9, unit="GB/s (as GFlops)"),
}


class Metric:
    """amt::metric (include/metric.hpp:28-500): per routine name the series of
    GFLOP/s, its min, max and sum; str() and csv() in the reference's layout."""

    def __init__(self, total: int, peak: float, precision: str = "[Single-Precision]"):
        self.total, self.peak, self.precision = total, peak, precision
        self.data: Dict[str, dict] = {}

    def update(self, name: str, gflops: float) -> None:
        d = self.data.setdefault(name, {"plot": [], "min": self.peak, "max": 0.0, "agg": 0.0})
        d["plot"].append(gflops)
        d["min"] = min(d["min"], gflops)
        d["max"] = max(d["max"], gflops)
        d["agg"] += gflops

    def str(self, pattern: Optional[str] = None) -> str:
        pref = None
        if pattern:
            for k, v in self.data.items():
                if pattern in k:
                    pref = v
                    break
        out = [f"Peak Performance: {self.peak:g} GFlops"]
        for k, v in self.data.items():
            avg = v["agg"] / self.total
            out.append(f"Name: {k}")
            out.append(f"\tMin GFlops: {v['min']:g}")
            out.append(f"\tMax GFlops: {v['max']:g}")
            if pref is not None:
                pavg = pref["agg"] / self.total
                out.append(f"\tMax SpeedUp with respect to {pattern}: {pref['max'] / v['max']:g}")
                out.append(f"\tAvg SpeedUp with respect to {pattern}: {pavg / avg:g}")
            out.append(f"\tMax Peak Utilization in %: {v['max'] / self.peak * 100.0:g}")
            out.append(f"\tAvg GFlops: {avg:g}")
            out.append(f"\tAvg Peak Utilization in %: {avg / self.peak * 100.0:g}")
            out.append("")
        return "\n".join(out) + "\n"

    def csv(self, path: str) -> None:
        names = list(self.data)
        with open(path, "w") as f:
            f.write(",".join(f'"{k}"' for k in names) + "\n")
            for i in range(self.total):
                f.write(",".join(f"{self.data[k]['plot'][i]:g}" for k in names) + "\n")


def _load_package():
    here = os.path.dirname(os.path.abspath(__file__))
    sys.path.insert(0, os.path.dirname(here))
    import __graft_entry__ as graft  # noqa: E402
    return graft.load_package()


def run(op: str, sizes: List[float], iters: int, layout: str = "first",
        modes=("plain",), log: Callable[[str], None] = lambda s: None) -> Metric:
    """Time `op` at every size; returns the filled Metric."""
    import torch
    pkg = _load_package()
    lib = pkg.lib()
    pkg.check(lib.omb_init())
    cfg = SWEEPS[op]
    dev = torch.device("cuda:0")
    stream = torch.cuda.Stream()
    sp = stream.cuda_stream
    m = Metric(len(sizes), cfg["peak"])
    first = layout == "first"

    def mat(n):  # n x n in the chosen storage order, as (tensor, rs, cs)
        with torch.cuda.stream(stream):
            t = torch.empty(n * n, device=dev)
        pkg.fill_uniform(t, 1, n, stream=sp)
        return t, (1, n) if first else (n, 1)

    def timed(fn):
        fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(iters):
            fn()
        e1.record(stream)
        e1.synchronize()
        return e0.elapsed_time(e1) / 1e3 / iters

    for s in sizes:
        n = int(s)
        ops = cfg["ops"](float(n))
        if op == "mtm":
            a, (rs, cs) = mat(n)
            b, _ = mat(n)
            with torch.cuda.stream(stream):
                c = torch.zeros(n * n, device=dev)
            for mode in modes:
                md = {"plain": pkg.PLAIN_F32, "split": pkg.SPLIT_BF16}[mode]
                t = timed(lambda: lib.omb_sgemm(n, n, n, 1.0, a.data_ptr(), rs, cs, b.data_ptr(),
                                                rs, cs, 1.0, c.data_ptr(), rs, cs, md, sp))
                m.update(f"gfx950 mtm {mode}", ops / t / 1e9)
        elif op == "mtv":
            a, (rs, cs) = mat(n)
            with torch.cuda.stream(stream):
                x = torch.ones(n, device=dev)
                y = torch.zeros(n, device=dev)
            beta = 1.0 if first else 0.0  # first order accumulates, last order overwrites
            t = timed(lambda: lib.omb_sgemv(n, n, 1.0, a.data_ptr(), rs, cs, x.data_ptr(), beta,
                                            y.data_ptr(), sp))
            m.update("gfx950 mtv", ops / t / 1e9)
        elif op == "dot":
            with torch.cuda.stream(stream):
                x = torch.empty(n, device=dev)
                r = torch.zeros(1, device=dev)
            pkg.fill_uniform(x, 1, 0, stream=sp)
            t = timed(lambda: lib.omb_sdot(n, x.data_ptr(), x.data_ptr(), r.data_ptr(), sp))
            m.update("gfx950 dot", ops / t / 1e9)
        elif op == "outer":
            c, (rs, cs) = mat(n)
            with torch.cuda.stream(stream):
                x = torch.ones(n, device=dev)
            t = timed(lambda: lib.omb_sger(n, n, 1.0, x.data_ptr(), x.data_ptr(), c.data_ptr(),
                                           rs, cs, sp))
            m.update("gfx950 outer", ops / t / 1e9)
        elif op == "trans":
            a, (rs, cs) = mat(n)
            with torch.cuda.stream(stream):
                c = torch.empty(n * n, device=dev)
            t = timed(lambda: lib.omb_stranspose(n, n, a.data_ptr(), rs, cs, c.data_ptr(), rs, cs,
                                                 sp))
            m.update("gfx950 trans", ops / t / 1e9)
        else:
            raise ValueError(f"unknown op {op}")
        log(f"{op} n={n} done")
    return m


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--op", choices=sorted(SWEEPS), default="mtm")
    ap.add_argument("--layout", choices=("first", "last"), default="first",
                    help="storage order of the matrices (the drivers use first_order)")
    ap.add_argument("--min", type=float, default=None)
    ap.add_argument("--max", type=float, default=None)
    ap.add_argument("--step", type=float, default=None)
    ap.add_argument("--iters", type=int, default=None, help="MaxIter (default: the driver's)")
    ap.add_argument("--csv", default=None, help="write the per-size series here")
    ap.add_argument("--quiet", action="store_true")
    args = ap.parse_args(argv)
    cfg = SWEEPS[args.op]
    start, end, stride, mul = cfg["range"]
    sizes = reference_range(args.min or start, args.max or end, args.step or stride, mul)
    iters = args.iters or cfg["iters"]
    log = (lambda s: None) if args.quiet else (lambda s: print(s, file=sys.stderr, flush=True))
    modes = ("plain", "split") if args.op == "mtm" else ("plain",)
    m = run(args.op, sizes, iters, args.layout, modes, log)
    pattern = "split" if args.op == "mtm" else None
    print(m.str(pattern), end="")
    if args.csv:
        m.csv(args.csv)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
