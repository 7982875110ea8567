"""Benchmark of the gfx950 BLAS hot path; prints ONE JSON line.

Headline (`value`): split-BF16 SGEMM, M = N = K = 8192, row-major, operands
U[-1,1) from seed 2, C = 0: FP32-equivalent TFLOP/s = 2MNK / step time, where
one step is one omb_sgemm(mode = split-BF16) call (both split passes plus the
MFMA kernel), operands already resident in HBM.

Beside it: plain FP32 SGEMM 8192^3 (row- and column-major, seed 1), the
error of every GEMM variant against an FP64 product (both input
distributions), SGEMV 16384^2 and SDOT 2^28 in GB/s, the roofline of each
kernel from HIP events on the launch stream, and the reference algorithm
(the oracle, OpenMP on the host cores) timed on the same 8192^3 FP32 GEMM.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import __graft_entry__ as graft  # noqa: E402

METRIC = ("SGEMM TFLOP/s 8192^3 split-BF16 vs plain FP32 @≤2x FP32 err; "
          "GEMV GB/s; p99 ms")
PEAK_HBM = 8.0e12
PEAK_MFMA_BF16 = 2.5e15
PEAK_MFMA_FP16 = 2.5e15
PEAK_FP32 = 157.3e12


def profiled_traffic(kernel: str, workload: str):
    """HBM bytes per launch of `kernel` on `workload` as measured by the
    profiler's FETCH_SIZE / WRITE_SIZE passes (scaled by a copy of known
    size), read from the committed summary; None when it is absent."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_pmc.md")
    try:
        for line in open(path):
            parts = line.split()
            if (len(parts) == 4 and parts[0] == "traffic_bytes_per_launch"
                    and parts[1] == kernel and parts[2] == workload + ":"):
                return int(parts[3])
    except OSError:
        pass
    return None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


class Timer:
    """HIP events recorded on the launch stream."""

    def __init__(self, torch, stream):
        self.torch, self.stream = torch, stream

    def run(self, fn, iters):
        t = self.torch
        e0, e1 = t.cuda.Event(enable_timing=True), t.cuda.Event(enable_timing=True)
        e0.record(self.stream)
        for _ in range(iters):
            fn()
        e1.record(self.stream)
        e1.synchronize()
        return e0.elapsed_time(e1) / iters / 1e3  # seconds per call


_HOST_CORES = None


def host_cores() -> int:
    """CPUs in this process's affinity mask, as first read (the oracle's
    OpenMP team size)."""
    global _HOST_CORES
    if _HOST_CORES is None:
        _HOST_CORES = (len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity")
                       else os.cpu_count() or 1)
    return _HOST_CORES


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-errors", action="store_true", help="skip the FP64 error legs")
    ap.add_argument("--cpu-budget", type=float, default=25.0,
                    help="seconds of host time for the CPU baseline sample")
    ap.add_argument("--legs", default="all",
                    help="comma list of legs beside the headline: plain, errors, gemv, dot, "
                         "level2, grouped, serving, c1, cpu (default all); grouped_plans "
                         "times the grouped plans only (no per-instance calls)")
    ap.add_argument("--serve-rate", type=float, default=float(os.environ.get("OMB_SERVE_RATE", 32000)),
                    help="serving leg: Poisson arrival rate (requests/s)")
    ap.add_argument("--serve-count", type=int, default=int(os.environ.get("OMB_SERVE_COUNT", 3000)),
                    help="serving leg: requests in the trace")
    ap.add_argument("--serve-slots", type=int, default=int(os.environ.get("OMB_SERVE_SLOTS", 8)),
                    help="serving leg: requests in flight at most (one stream each)")
    ap.add_argument("--cpu-baseline-child", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--cpu-cores", type=int, default=0, help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.cpu_baseline_child:   # the CPU leg's own process (no GPU runtime)
        if not graft.built():
            graft.build()
        print(json.dumps(cpu_baseline(args.cpu_budget, args.cpu_cores or None)), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU sharding is a design, not a measured path)", file=sys.stderr)
        sys.exit(2)

    # The host CPUs this process may use, read before any import that starts
    # an OpenMP runtime: a runtime told to bind (OMP_PROC_BIND) pins the
    # main thread to one place, after which the affinity mask shows one CPU.
    host_cores()

    legs = {"plain", "errors", "gemv", "dot", "level2", "grouped", "serving", "c1", "cpu"}
    if args.legs != "all":
        legs = {x.strip() for x in args.legs.split(",") if x.strip()}
    if args.no_errors:
        legs.discard("errors")
    if args.no_cpu:
        legs.discard("cpu")

    # The CPU baseline runs first, in a child process, while this process has
    # loaded neither torch nor its OpenMP runtime.
    cpu = cpu_baseline_child(args.cpu_budget, host_cores()) if "cpu" in legs else None

    import torch

    if not graft.built():
        graft.build()
    pkg = graft.load_package()
    lib = pkg.lib()
    pkg.check(lib.omb_init())
    dev = torch.device("cuda:0")
    stream = torch.cuda.Stream()
    sp = stream.cuda_stream
    timer = Timer(torch, stream)
    out: dict = {}

    def sync():
        torch.cuda.synchronize()

    n = 8192
    flop = 2.0 * n * n * n
    with torch.cuda.stream(stream):
        A = torch.empty((n, n), device=dev)
        B = torch.empty((n, n), device=dev)
        C = torch.zeros((n, n), device=dev)
    pkg.fill_uniform(A, 2, 0, stream=sp)
    pkg.fill_uniform(B, 2, 1, stream=sp)
    sync()

    def gemm(mode, a=A, b=B, c=C, beta=0.0, layout="row"):
        if layout == "row":
            return lib.omb_sgemm(n, n, n, 1.0, a.data_ptr(), n, 1, b.data_ptr(), n, 1, beta,
                                 c.data_ptr(), n, 1, mode, sp)
        return lib.omb_sgemm(n, n, n, 1.0, a.data_ptr(), 1, n, b.data_ptr(), 1, n, beta,
                             c.data_ptr(), 1, n, mode, sp)

    # ------------------------------------------------ headline: split-BF16 --
    log("headline: split-BF16 sgemm 8192^3")
    for _ in range(args.warmup):
        pkg.check(gemm(pkg.SPLIT_BF16))
    sync()
    t0 = time.perf_counter()
    t_step = timer.run(lambda: gemm(pkg.SPLIT_BF16), args.steps)
    sync()
    wall = time.perf_counter() - t0
    value = flop / t_step / 1e12

    # a plain copy of known size (256 MiB read + 256 MiB written): the
    # reference point that scales the profiler's FETCH_SIZE / WRITE_SIZE
    with torch.cuda.stream(stream):
        torch.mul(A, 1.0, out=C)
    sync()

    # the MFMA kernel alone (pre-packed operands), for the roofline
    pbytes = lib.omb_split_pack_bytes(n, n)
    with torch.cuda.stream(stream):
        Ap = torch.empty(pbytes, dtype=torch.uint8, device=dev)
        Bp = torch.empty(pbytes, dtype=torch.uint8, device=dev)
    pkg.check(lib.omb_split_pack(n, n, A.data_ptr(), n, 1, Ap.data_ptr(), sp))
    pkg.check(lib.omb_split_pack(n, n, B.data_ptr(), 1, n, Bp.data_ptr(), sp))  # B^T view
    # the two packs of a step (A, then B^T), per pack
    def pack_both():
        lib.omb_split_pack(n, n, A.data_ptr(), n, 1, Ap.data_ptr(), sp)
        lib.omb_split_pack(n, n, B.data_ptr(), 1, n, Bp.data_ptr(), sp)
    t_pack = timer.run(pack_both, max(3, args.steps // 2)) / 2
    kernels = {}
    for nprod, name in ((6, "sgemm_bf16x3_kernel<6> (pre-packed operands)"),
                        (3, "sgemm_bf16x3_kernel<3> (pre-packed operands)")):
        fn = lambda: lib.omb_sgemm_packed(n, n, n, 1.0, Ap.data_ptr(), Bp.data_ptr(), 0.0,
                                          C.data_ptr(), n, 1, nprod, sp)
        for _ in range(2):
            fn()
        t_k = timer.run(fn, args.steps)
        kernels[name] = {"ms": t_k * 1e3, "mfma_tflops": nprod * flop / t_k / 1e12,
                         "fp32_equiv_tflops": flop / t_k / 1e12,
                         "frac_of_bf16_peak": nprod * flop / t_k / PEAK_MFMA_BF16}
    t_k6 = kernels["sgemm_bf16x3_kernel<6> (pre-packed operands)"]["ms"] / 1e3
    # The same kernel on all-zero operands: identical instructions and memory
    # traffic, no switching activity in the MFMA datapath.  The gap to the
    # random-operand time is the power-limited, data-dependent MFMA rate.
    with torch.cuda.stream(stream):
        Ap.zero_()
        Bp.zero_()
    sync()
    fn = lambda: lib.omb_sgemm_packed(n, n, n, 1.0, Ap.data_ptr(), Bp.data_ptr(), 0.0,
                                      C.data_ptr(), n, 1, 6, sp)
    fn()
    t_z = timer.run(fn, max(3, args.steps // 2))
    kernels["sgemm_bf16x3_kernel<6> (zero operands)"] = {
        "ms": t_z * 1e3, "mfma_tflops": 6 * flop / t_z / 1e12,
        "frac_of_bf16_peak": 6 * flop / t_z / PEAK_MFMA_BF16,
        "random_over_zero_time": t_k6 / t_z}
    # What any BF16 MFMA kernel can sustain here: MFMAs on register operands
    # only (no memory traffic), random against zero data, at the GEMM's
    # occupancy (2 workgroups of 256 threads per CU) and for the same MFMA
    # work (6 x 2 x 8192^3); and the vendor BF16 GEMM (torch.matmul on
    # hipBLASLt) at 8192^3 on random and zero BF16 data, as a reference.
    cus = torch.cuda.get_device_properties(dev).multi_processor_count
    with torch.cuda.stream(stream):
        sink = torch.zeros(1, device=dev)
    mblocks = 2 * cus
    miters = max(1, int(round(6 * flop / (mblocks * 4 * 64 * 32768.0))))
    mwork = mblocks * 4 * miters * 64 * 32768.0
    mceil = {}
    for rnd in (1, 0):
        fn = lambda: lib.omb_calib_mfma(mblocks, miters, rnd, sink.data_ptr(), sp)
        pkg.check(fn())
        mceil["random" if rnd else "zero"] = mwork / timer.run(fn, 3) / 1e12
    kernels["mfma_ceiling (register operands, no memory traffic)"] = {
        "random_tflops": mceil["random"], "zero_tflops": mceil["zero"],
        "random_over_zero_time": mceil["zero"] / mceil["random"],
        "work_flop": mwork, "kernel_frac_of_random_ceiling":
            6 * flop / t_k6 / 1e12 / mceil["random"]}
    vendor = {}
    with torch.cuda.stream(stream):
        Vb = torch.empty((n, n), dtype=torch.bfloat16, device=dev)
        Wb = torch.empty((n, n), dtype=torch.bfloat16, device=dev)
        Ob = torch.empty((n, n), dtype=torch.bfloat16, device=dev)
        for name, fill in (("random", True), ("zero", False)):
            if fill:
                Vb.copy_(A)
                Wb.copy_(B)
            else:
                Vb.zero_()
                Wb.zero_()
            torch.matmul(Vb, Wb, out=Ob)
            vendor[name + "_tflops"] = flop / timer.run(lambda: torch.matmul(Vb, Wb, out=Ob),
                                                        5) / 1e12
    del Vb, Wb, Ob
    kernels["vendor_bf16_gemm (torch.matmul 8192^3, hipBLASLt)"] = vendor

    kernels["split_pack_kernel (8192x8192)"] = {
        "ms": t_pack * 1e3, "gbps": (4 + 6) * n * n / t_pack / 1e9,
        "frac_of_hbm_peak": (4 + 6) * n * n / t_pack / PEAK_HBM}
    del Ap, Bp

    # split 3-product step time
    t_step3 = timer.run(lambda: gemm(pkg.SPLIT_BF16_P3), max(3, args.steps // 2))
    # scaled split-FP16 step (rows of A and columns of B scaled by powers of
    # two, 2 FP16 terms, 3 products on the FP16 MFMA)
    pkg.check(gemm(pkg.SPLIT_FP16))
    t_step16 = timer.run(lambda: gemm(pkg.SPLIT_FP16), max(3, args.steps // 2))

    # --------------------------------------------------------- plain FP32 --
    plain = {}
    for layout in (("row", "col") if "plain" in legs else ()):
        if layout == "row":
            log("plain fp32 sgemm 8192^3 (seed 1, row and column major)")
            pkg.fill_uniform(A, 1, 0, stream=sp)
            pkg.fill_uniform(B, 1, 1, stream=sp)
        fn = lambda: gemm(pkg.PLAIN_F32, layout=layout)
        fn()
        t_p = timer.run(fn, max(3, args.steps // 2))
        plain[layout] = {"ms": t_p * 1e3, "tflops": flop / t_p / 1e12,
                         "frac_of_fp32_peak": flop / t_p / PEAK_FP32}
    if plain:
        kernels["sgemm_f32_kernel (8192^3 row-major)"] = dict(plain["row"])
        # The FP32 FMA rate the CUs sustain with nothing but independent
        # v_fmac_f32 chains, at the plain kernel's occupancy (2 workgroups of
        # 256 threads per CU) and at 4: the practical ceiling of the VALU path.
        with torch.cuda.stream(stream):
            sink = torch.zeros(1, device=dev)
        cus = torch.cuda.get_device_properties(dev).multi_processor_count
        ceiling = {}
        for per_cu in (2, 4):
            blocks, iters = per_cu * cus, 8192
            fn = lambda: lib.omb_calib_fmac(blocks, iters, sink.data_ptr(), sp)
            pkg.check(fn())
            t_c = timer.run(fn, 5)
            ceiling[f"{per_cu}_workgroups_per_cu"] = blocks * 256 * iters * 128 / t_c / 1e12
        best = max(ceiling.values())
        kernels["valu_fmac_ceiling"] = {
            "tflops": ceiling, "frac_of_fp32_peak": best / (PEAK_FP32 / 1e12),
            "plain_row_frac_of_ceiling": plain["row"]["tflops"] / best}

    # ------------------------------------------------------------- errors --
    errors = {}
    if "errors" in legs:
        errors = error_leg(pkg, lib, torch, dev, stream, A, B, C)
    sync()

    # --------------------------------------------------------- GEMV / dot --
    del C
    gemv, dot, refs = {}, {}, {}
    if "gemv" in legs:
        log("sgemv 16384^2 (seed 1, row and column major)")
        m = 16384
        with torch.cuda.stream(stream):
            GA = torch.empty((m, m), device=dev)
            gx = torch.empty(m, device=dev)
            gy = torch.zeros(m, device=dev)
        pkg.fill_uniform(GA, 1, 0, stream=sp)
        pkg.fill_uniform(gx, 1, 1, stream=sp)
        for layout, (rs, cs, beta) in (("row", (m, 1, 0.0)), ("col", (1, m, 1.0))):
            fn = lambda: lib.omb_sgemv(m, m, 1.0, GA.data_ptr(), rs, cs, gx.data_ptr(), beta,
                                       gy.data_ptr(), sp)
            for _ in range(3):
                fn()
            t_g = timer.run(fn, max(10, args.steps))
            nbytes = 4.0 * (m * m + m + m) + (4.0 * m if beta != 0.0 else 0.0)
            gemv[layout] = {"us": t_g * 1e6, "gbps": nbytes / t_g / 1e9,
                            "frac_of_hbm_peak": nbytes / t_g / PEAK_HBM, "bytes": nbytes}
        kernels["sgemv_row_kernel (16384^2)"] = gemv["row"]
        kernels["sgemv_col_kernel (16384^2)"] = gemv["col"]
        # errors against an FP64 product (GPU check kernel), y = A x from y = 0
        with torch.cuda.stream(stream):
            y64 = torch.empty(m, dtype=torch.float64, device=dev)
        pkg.check(lib.omb_check_dgemv(m, m, GA.data_ptr(), m, 1, gx.data_ptr(), None,
                                      y64.data_ptr(), sp))
        sync()
        refs["y64"] = y64.cpu().numpy()
        for layout, (rs, cs) in (("row", (m, 1)), ("col", (1, m))):
            gy.zero_()
            sync()
            pkg.check(lib.omb_sgemv(m, m, 1.0, GA.data_ptr(), rs, cs, gx.data_ptr(), 0.0,
                                    gy.data_ptr(), sp))
            sync()
            # the column-major call reads A as its transpose: compare with A^T x
            want = refs["y64"] if layout == "row" else None
            if want is None:
                pkg.check(lib.omb_check_dgemv(m, m, GA.data_ptr(), 1, m, gx.data_ptr(), None,
                                              y64.data_ptr(), sp))
                sync()
                want = y64.cpu().numpy()
                refs["y64_col"] = want
            got = gy.cpu().numpy().astype(np.float64)
            gemv[layout]["max_rel"] = float(np.abs(got - want).max() / np.abs(want).max())
            gemv[layout]["frob_rel"] = float(np.linalg.norm(got - want) / np.linalg.norm(want))
        del GA
    if "dot" in legs:
        log("sdot 2^28 (seed 1)")
        nd = 1 << 28
        with torch.cuda.stream(stream):
            dx = torch.empty(nd, device=dev)
            dy = torch.empty(nd, device=dev)
            dr = torch.zeros(1, device=dev)
        pkg.fill_uniform(dx, 1, 0, stream=sp)
        pkg.fill_uniform(dy, 1, 1, stream=sp)
        fn = lambda: lib.omb_sdot(nd, dx.data_ptr(), dy.data_ptr(), dr.data_ptr(), sp)
        for _ in range(3):
            fn()
        t_d = timer.run(fn, max(10, args.steps))
        dot = {"us": t_d * 1e6, "gbps": 8.0 * nd / t_d / 1e9,
               "frac_of_hbm_peak": 8.0 * nd / t_d / PEAK_HBM}
        kernels["sdot_kernel (2^28)"] = dot
        # |s - s64| / |s64|, and |s - s64| / sum |x_i y_i| (independent of
        # cancellation; the classic gamma_n bound is vacuous here, n u = 16)
        with torch.cuda.stream(stream):
            r64 = torch.empty(1, dtype=torch.float64, device=dev)
            sabs = (dx.double().abs_().mul_(dy.double().abs_())).sum()
        pkg.check(lib.omb_check_ddot(nd, dx.data_ptr(), dy.data_ptr(), r64.data_ptr(), sp))
        sync()
        s64, sa = float(r64.item()), float(sabs.item())
        refs["dot64"], refs["dot_sum_abs"] = s64, sa
        err = abs(float(dr.item()) - s64)
        dot["rel_err"] = err / abs(s64)
        dot["err_over_sum_abs"] = err / sa
        del dx, dy
    sync()

    # ------------------------------------------- outer product / transpose --
    level2 = {}
    if "level2" in legs:
        log("sger and transpose 16384^2 (seed 1)")
        m = 16384
        with torch.cuda.stream(stream):
            LA = torch.empty((m, m), device=dev)
            LC = torch.empty((m, m), device=dev)
            lx = torch.empty(m, device=dev)
            ly = torch.empty(m, device=dev)
        for t, sid in ((LA, 0), (lx, 1), (ly, 2)):
            pkg.fill_uniform(t, 1, sid, stream=sp)
        fn = lambda: lib.omb_sger(m, m, 1.0, lx.data_ptr(), ly.data_ptr(), LA.data_ptr(), m, 1, sp)
        fn()
        t_g = timer.run(fn, max(10, args.steps))
        nbytes = 8.0 * m * m + 4.0 * (2 * m)
        level2["sger_row"] = {"us": t_g * 1e6, "gbps": nbytes / t_g / 1e9,
                              "frac_of_hbm_peak": nbytes / t_g / PEAK_HBM, "bytes": nbytes}
        fn = lambda: lib.omb_stranspose(m, m, LA.data_ptr(), m, 1, LC.data_ptr(), m, 1, sp)
        fn()
        t_t = timer.run(fn, max(10, args.steps))
        nbytes = 8.0 * m * m
        level2["transpose_out_of_place"] = {"us": t_t * 1e6, "gbps": nbytes / t_t / 1e9,
                                            "frac_of_hbm_peak": nbytes / t_t / PEAK_HBM,
                                            "bytes": nbytes}
        fn = lambda: lib.omb_stranspose_inplace(m, LC.data_ptr(), m, sp)
        fn()
        t_i = timer.run(fn, max(10, args.steps))
        level2["transpose_in_place"] = {"us": t_i * 1e6, "gbps": nbytes / t_i / 1e9,
                                        "frac_of_hbm_peak": nbytes / t_i / PEAK_HBM,
                                        "bytes": nbytes}
        kernels["ger_kernel (16384^2, row-major C)"] = level2["sger_row"]
        kernels["transpose_tiles_kernel (16384^2)"] = level2["transpose_out_of_place"]
        kernels["transpose_inplace_kernel (16384^2)"] = level2["transpose_in_place"]
        del LA, LC
        sync()

    # ------------------------------------------------ grouped and serving --
    grouped, serving = None, None
    if "grouped" in legs or "grouped_plans" in legs:
        grouped = grouped_leg(pkg, lib, torch, dev, stream, timer, args.steps, "cpu" in legs,
                              per_instance_calls="grouped" in legs)
        sync()
    if "serving" in legs:
        serving = serving_leg(pkg, torch, dev, stream, args.serve_rate, args.serve_count,
                              args.serve_slots)
        sync()

    c1 = None
    if "c1" in legs:
        c1 = c1_leg(pkg, lib, torch, dev, sp, timer)
        sync()

    # ------------------------------------------------------- CPU baseline --
    cpu_ref = None
    if "cpu" in legs:
        cpu_ref = cpu_reference_legs(pkg, torch, dev, stream, refs)

    res = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_step * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "arith": "FP32 operands split into 3 BF16 terms, 6 MFMA products, FP32 accumulation",
        "data": "synthetic: counter-based U[-1,1) from seed 2 (operands resident in HBM)",
        "config": {"workload": "sgemm split-bf16 M=N=K=8192 row-major, C = A*B",
                   "M": n, "N": n, "K": n, "layout": "row-major", "seed": 2},
        "roofline": {"bound": "mfma", "kernel": "sgemm_bf16x3_kernel<6>",
                     "achieved": round(6 * flop / t_k6 / 1e12, 2), "peak": 2500.0,
                     "unit": "TFLOP/s", "frac": round(6 * flop / t_k6 / PEAK_MFMA_BF16, 4),
                     "traffic": profiled_traffic("sgemm_bf16x3_kernel<6>", "8192^3"),
                     "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, "
                                     "scaled by a 256 MiB copy)",
                     "algorithmic_bytes": 2 * 3 * n * n * 2 + 4 * n * n,
                     "work_per_launch": "6 products x 2*8192^3 = 6.597e12 MFMA flop",
                     "sustained_mfma_ceiling_random_operands": round(
                         kernels["mfma_ceiling (register operands, no memory traffic)"]
                         ["random_tflops"], 2),
                     "frac_of_sustained_mfma_ceiling": round(
                         kernels["mfma_ceiling (register operands, no memory traffic)"]
                         ["kernel_frac_of_random_ceiling"], 4)},
        "cpu_baseline": cpu,
        "cpu_reference_path": cpu_ref,
        "config1_1024": c1,
        "wall_s_timed_region": round(wall, 4),
        "split_bf16_p3_step_ms": round(t_step3 * 1e3, 4),
        "split_fp16_scaled_p3": {
            "step_ms": round(t_step16 * 1e3, 4), "tflops": round(flop / t_step16 / 1e12, 3),
            "speedup_vs_plain": (round(plain["row"]["ms"] / (t_step16 * 1e3), 3)
                                 if plain else None),
            "arith": "rows of A / columns of B scaled by powers of two, 2 FP16 terms, "
                     "3 FP16 MFMA products, FP32 accumulation (errors: errors.*.split_fp16_p3)"},
        "plain_f32": plain,
        "speedup_split_vs_plain": (round(plain["row"]["ms"] / (t_step * 1e3), 3)
                                   if plain else None),
        "errors": errors,
        "gemv": gemv,
        "dot": dot,
        "outer_transpose": level2,
        "grouped": grouped,
        "serving": serving,
        "kernels": kernels,
        "device": torch.cuda.get_device_name(0),
    }
    print(json.dumps(res), flush=True)


def error_leg(pkg, lib, torch, dev, stream, A, B, C):
    """Max- and Frobenius-relative error vs an FP64 product (GPU check kernel)
    of plain FP32, split-BF16 (6 and 3 products) and the scaled split-FP16
    (3 products) at 8192^3, for U[-1,1)
    and sign*2^u (u ~ U[-8,8]) inputs from seed 2, and the componentwise
    error max |C - C64| / (|A| |B|).  The reference algorithm's (oracle's)
    error is measured on a 64-row slab, whose values do not depend on the
    other rows, and compared with the GPU's on the same slab."""
    sp = stream.cuda_stream
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import helpers

    n = 8192
    rows = slice(4096, 4160)
    out = {}
    C64 = torch.empty((n, n), dtype=torch.float64, device=dev)
    D64 = torch.empty((n, n), dtype=torch.float64, device=dev)
    for dist in ("uniform", "pow2"):
        log(f"errors vs FP64: {dist}")
        if dist == "uniform":
            pkg.fill_uniform(A, 2, 0, stream=sp)
            pkg.fill_uniform(B, 2, 1, stream=sp)
            torch.cuda.synchronize()
            a_h, b_h = A.cpu().numpy(), B.cpu().numpy()
        else:
            a_h = helpers.rng_pow2(n * n, 2, 0).reshape(n, n)
            b_h = helpers.rng_pow2(n * n, 2, 1).reshape(n, n)
            A.copy_(torch.from_numpy(a_h))
            B.copy_(torch.from_numpy(b_h))
            torch.cuda.synchronize()
        # |A| |B| in FP64 for the componentwise error max |C - C64| / (|A| |B|)
        with torch.cuda.stream(stream):
            Aabs, Babs = A.abs(), B.abs()
        pkg.check(lib.omb_check_dgemm(n, n, n, Aabs.data_ptr(), n, 1, Babs.data_ptr(), n, 1, None,
                                      n, 1, D64.data_ptr(), sp))
        pkg.check(lib.omb_check_dgemm(n, n, n, A.data_ptr(), n, 1, B.data_ptr(), n, 1, None, n, 1,
                                      C64.data_ptr(), sp))
        torch.cuda.synchronize()
        del Aabs, Babs
        truth = C64[rows].cpu().numpy()
        c_ref = np.zeros((64, n), np.float32)
        helpers.oracle_mtm(c_ref, np.ascontiguousarray(a_h[rows]), b_h)
        ref = helpers.rel_errors(c_ref, truth)
        comp_ref = float(np.max(np.abs(c_ref.astype(np.float64) - truth) / D64[rows].cpu().numpy()))
        d = {"oracle_slab": {"max_rel": ref[0], "frob_rel": ref[1], "componentwise": comp_ref}}
        for mode, name in ((pkg.PLAIN_F32, "plain_f32"), (pkg.SPLIT_BF16, "split_bf16_p6"),
                           (pkg.SPLIT_BF16_P3, "split_bf16_p3"),
                           (pkg.SPLIT_FP16, "split_fp16_p3")):
            pkg.check(lib.omb_sgemm(n, n, n, 1.0, A.data_ptr(), n, 1, B.data_ptr(), n, 1, 0.0,
                                    C.data_ptr(), n, 1, mode, sp))
            torch.cuda.synchronize()
            full = pkg.rel_errors(C, C64, stream=sp)
            slab = helpers.rel_errors(C[rows].cpu().numpy(), truth)
            with torch.cuda.stream(stream):
                comp = float(((C.double() - C64).abs_().div_(D64)).max())
            comp_slab = float(((C[rows].double() - C64[rows]).abs_().div_(D64[rows])).max())
            d[name] = {"max_rel": full[0], "frob_rel": full[1], "componentwise": comp,
                       "slab_max_rel": slab[0], "slab_frob_rel": slab[1],
                       "slab_ratio_vs_oracle": [slab[0] / ref[0], slab[1] / ref[1]],
                       "slab_componentwise_ratio_vs_oracle": comp_slab / comp_ref}
        p = d["plain_f32"]
        for name in ("split_bf16_p6", "split_bf16_p3", "split_fp16_p3"):
            d[name]["ratio_vs_plain_f32"] = [d[name]["max_rel"] / p["max_rel"],
                                             d[name]["frob_rel"] / p["frob_rel"]]
        out[dist] = d
    del C64, D64
    return out


def _pool(torch, dev, stream, sizes):
    """One device buffer holding consecutive 16-byte aligned slices."""
    sizes = np.asarray(sizes, dtype=np.int64)
    padded = (sizes + 3) // 4 * 4
    offs = np.concatenate([[0], np.cumsum(padded)[:-1]]).astype(np.int64)
    with torch.cuda.stream(stream):
        buf = torch.empty(max(int(padded.sum()), 4), device=dev)
    return buf, offs


def grouped_leg(pkg, lib, torch, dev, stream, timer, steps, cpu_sample=True,
                per_instance_calls=True):
    """4096 row-major GEMMs (M, N, K log-uniform on [16, 512]) and 4096 GEMVs
    (rows, cols log-uniform on [64, 4096]) from seed 3, operands U[-1,1),
    C = A*B and y = A*x, in one persistent grouped launch replayed from a
    hipGraph; tiles from the autotuner run on the same batch.  Compared with
    t_roof = sum over instances of max(flop / peak, bytes / 8 TB/s) and with
    the same work as one omb_sgemm / omb_sgemv call per instance."""
    log("grouped: 4096 GEMMs + 4096 GEMVs (seed 3)")
    sp = stream.cuda_stream
    rng = np.random.default_rng(3)
    gs = np.stack([pkg.log_uniform_ints(rng, 16, 512, 4096) for _ in range(3)], 1)
    vs = np.stack([pkg.log_uniform_ints(rng, 64, 4096, 4096) for _ in range(2)], 1)
    M, N, K = gs[:, 0], gs[:, 1], gs[:, 2]
    VM, VN = vs[:, 0], vs[:, 1]
    pools = [_pool(torch, dev, stream, x) for x in (M * K, K * N, M * N, VM * VN, VN, VM)]
    for i, (buf, _) in enumerate(pools):
        pkg.fill_uniform(buf, 3, i, stream=sp)
    ptr = lambda j, i: pools[j][0].data_ptr() + 4 * int(pools[j][1][i])
    gemms = [(int(M[i]), int(N[i]), int(K[i]), ptr(0, i), int(K[i]), ptr(1, i), int(N[i]),
              ptr(2, i), int(N[i]), 1.0, 0.0) for i in range(len(gs))]
    gemvs = [(int(VM[i]), int(VN[i]), ptr(3, i), int(VN[i]), ptr(4, i), ptr(5, i), 1.0, 0.0)
             for i in range(len(vs))]
    g_flop = float((2.0 * M * N * K).sum())
    g_bytes = 4.0 * (M * K + K * N + M * N)
    v_bytes = 4.0 * (VM * VN + VM + VN)
    v_flop = float((2.0 * VM * VN).sum())

    def roof(fp32_rate):
        tg = np.maximum(2.0 * M * N * K / fp32_rate, g_bytes / PEAK_HBM).sum()
        tv = np.maximum(2.0 * VM * VN / fp32_rate, v_bytes / PEAK_HBM).sum()
        return float(tg), float(tv)

    t0 = time.perf_counter()
    pkg.autotune_group(gemms, pkg.PLAIN_F32, None, stream=sp)
    t_tune = time.perf_counter() - t0

    def time_plan(g, v, mode):
        plan = pkg.GroupPlan(g, v, mode, stream=sp)
        for _ in range(2):
            plan.launch(sp)
        t = timer.run(lambda: plan.launch(sp), max(5, steps))
        desc = plan.describe()
        plan.close()
        return t, desc

    out = {"instances": {"gemm": len(gemms), "gemv": len(gemvs)},
           "gemm_flop": g_flop, "gemv_flop": v_flop,
           "gemm_bytes": float(g_bytes.sum()), "gemv_bytes": float(v_bytes.sum()),
           "autotune_s": round(t_tune, 3)}
    for mode, name, rate in ((pkg.PLAIN_F32, "plain_f32", PEAK_FP32),
                             (pkg.SPLIT_BF16, "split_bf16_p6", PEAK_MFMA_BF16 / 6),
                             (pkg.SPLIT_FP16, "split_fp16_p3", PEAK_MFMA_FP16 / 3)):
        t_all, desc = time_plan(gemms, gemvs, mode)
        t_g, _ = time_plan(gemms, [], mode)
        t_v, _ = time_plan([], gemvs, mode)
        tg, tv = roof(rate)
        classes = {}
        for tok in desc.split():
            if "=" in tok and tok.startswith("gemm:"):
                cls = tok.split("=", 1)[1].split("(")[0]
                cnt = int(tok.rsplit("(", 1)[1].rstrip(")")) if "(" in tok else 1
                classes[cls] = classes.get(cls, 0) + cnt
        out[name] = {"ms": t_all * 1e3, "t_roof_ms": (tg + tv) * 1e3,
                     "frac_of_roof": (tg + tv) / t_all,
                     "gemm_only_ms": t_g * 1e3, "gemm_tflops": g_flop / t_g / 1e12,
                     "gemm_frac_of_roof": tg / t_g,
                     "gemv_only_ms": t_v * 1e3, "gemv_gbps": float(v_bytes.sum()) / t_v / 1e9,
                     "gemv_frac_of_hbm_peak": float(v_bytes.sum()) / t_v / PEAK_HBM,
                     "gemm_tile_classes": classes,
                     "roof_rate_tflops": rate / 1e12}

    def per_instance():
        for d in gemms:
            m, n, k, a, lda, b, ldb, c, ldc, al, be = d
            lib.omb_sgemm(m, n, k, al, a, lda, 1, b, ldb, 1, be, c, ldc, 1, pkg.PLAIN_F32, sp)
        for d in gemvs:
            m, n, a, lda, x, y, al, be = d
            lib.omb_sgemv(m, n, al, a, lda, 1, x, be, y, sp)

    if per_instance_calls:
        per_instance()
        out["per_instance_calls_ms"] = timer.run(per_instance, 2) * 1e3
        out["speedup_vs_per_instance_calls"] = (out["per_instance_calls_ms"] /
                                                out["plain_f32"]["ms"])
    if cpu_sample:
        out["cpu_sequential_sample"] = _grouped_cpu_sample(torch, pools, M, N, K, VM, VN)
    del pools
    return out


def _grouped_cpu_sample(torch, pools, M, N, K, VM, VN, count=256):
    """The reference has no batched API: its path for the grouped batch is one
    call per instance.  The oracle (the reference algorithm on the host
    cores) runs the first `count` GEMMs and GEMVs of the batch in sequence,
    on the same inputs (copied from the device pools); the full-batch time
    is that sample scaled by FLOP (GEMMs) and bytes (GEMVs)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import helpers

    log(f"grouped: CPU oracle on the first {count} GEMMs and {count} GEMVs")
    cores = host_cores()
    blk = helpers.host_blocking()
    gi = range(min(count, len(M)))
    vi = range(min(count, len(VM)))
    # instances are laid out in order, so the sample is a prefix of each pool
    ends = [len(pools[j][0]) if (len(gi) if j < 3 else len(vi)) >= len(pools[j][1])
            else int(pools[j][1][len(gi) if j < 3 else len(vi)]) for j in range(6)]
    host = [pools[j][0][:ends[j]].cpu().numpy() for j in range(6)]
    view = lambda j, i, shape: host[j][int(pools[j][1][i]):int(pools[j][1][i]) + int(np.prod(shape))].reshape(shape)
    cs = [np.zeros((int(M[i]), int(N[i])), np.float32) for i in gi]
    ys = [np.zeros(int(VM[i]), np.float32) for i in vi]
    t0 = time.perf_counter()
    for i in gi:
        helpers.oracle_mtm(cs[i], view(0, i, (int(M[i]), int(K[i]))), view(1, i, (int(K[i]), int(N[i]))),
                           blk, cores)
    t_g = time.perf_counter() - t0
    t0 = time.perf_counter()
    for i in vi:
        helpers.oracle_mtv(ys[i], view(3, i, (int(VM[i]), int(VN[i]))), view(4, i, (int(VN[i]),)),
                           blk, cores)
    t_v = time.perf_counter() - t0
    fl = 2.0 * M * N * K
    vb = 4.0 * (VM * VN + VM + VN)
    full_ms = (t_g * fl.sum() / fl[:len(gi)].sum() + t_v * vb.sum() / vb[:len(vi)].sum()) * 1e3
    return {"cores": cores, "gemms": len(gi), "gemvs": len(vi),
            "gemm_ms": t_g * 1e3, "gemm_tflops": float(fl[:len(gi)].sum()) / t_g / 1e12,
            "gemv_ms": t_v * 1e3, "gemv_gbps": float(vb[:len(vi)].sum()) / t_v / 1e9,
            "full_batch_estimate_ms": full_ms}


def serving_leg(pkg, torch, dev, stream, rate, count, slots):
    """Serving loop: a Poisson trace (seed 4) of `count` requests at `rate`
    per second, 50% GEMM (M, N, K log-uniform on [32, 4096], split-BF16),
    30% GEMV (rows, cols log-uniform on [256, 16384]), 20% dot (n
    log-uniform on [2^10, 2^24]), operands U[-1,1) resident in HBM,
    admission by free HBM.  Reports arrival-to-result latency percentiles
    and aggregate TFLOP/s = total work / wall; the same trace replayed with
    one slot (FIFO, no overlap) is the comparison."""
    log(f"serving: {count} requests at {rate:.0f}/s, {slots} slots (seed 4)")
    sp = stream.cuda_stream
    trace = pkg.serving_trace(count, rate, seed=4)
    sz = []
    for k, m, n, kk, _ in trace:
        sz.append((m * kk, kk * n, m * n) if k == pkg.REQ_GEMM else
                  ((m * n, n, m) if k == pkg.REQ_GEMV else (n, n, 1)))
    sz = np.asarray(sz, dtype=np.int64)
    pools = [_pool(torch, dev, stream, sz[:, j]) for j in range(3)]
    pkg.fill_uniform(pools[0][0], 4, 0, stream=sp)
    pkg.fill_uniform(pools[1][0], 4, 1, stream=sp)
    torch.cuda.synchronize()

    def requests():
        return [pkg.Request(k, pkg.SPLIT_BF16, m, n, kk, t,
                            *[pools[j][0].data_ptr() + 4 * int(pools[j][1][i]) for j in range(3)], 0)
                for i, (k, m, n, kk, t) in enumerate(trace)]

    # Warm-up, untimed: every request once on one slot, so each kernel the
    # trace needs has been launched once before the timed runs (a first
    # launch costs 1.5-6 ms), then a short batched run for the grouped path.
    pkg.serve_run(requests(), slots=1, batching=False)
    pkg.serve_run(requests()[: min(64, count)], slots=slots)
    res, st = pkg.serve_run(requests(), slots=slots)
    kinds = np.array([t[0] for t in trace])
    lat = np.array([r["latency_s"] for r in res])
    svc = np.array([r["done_s"] - r["start_s"] for r in res])
    ok = np.array([r["status"] == 0 for r in res])
    busy_bytes = sum(pkg.request_bytes(k, m, n, kk) for k, m, n, kk, _ in trace)
    out = {"rate_per_s": rate, "requests": count, "slots": slots, "completed": int(ok.sum()),
           "p50_ms": st["p50_s"] * 1e3, "p90_ms": st["p90_s"] * 1e3, "p99_ms": st["p99_s"] * 1e3,
           "max_ms": st["max_s"] * 1e3, "mean_ms": st["mean_s"] * 1e3,
           "wall_s": st["wall_s"], "aggregate_tflops": st["total_flop"] / st["wall_s"] / 1e12,
           "offered_tflops": st["total_flop"] / (trace[-1][4] or 1.0) / 1e12,
           "operand_gbps": busy_bytes / st["wall_s"] / 1e9,
           "max_in_flight": st["max_in_flight"], "admitted_late": st["admitted_late"],
           "batches": st["batches"], "batched_requests": st["batched_requests"],
           "p99_ms_by_kind": {name: float(np.percentile(lat[kinds == k], 99)) * 1e3
                              for k, name in ((0, "gemm"), (1, "gemv"), (2, "dot"))
                              if (kinds == k).any()},
           "service_p99_ms": float(np.percentile(svc, 99)) * 1e3}
    def sweep_at(rates, batching):
        rows = []
        for r in rates:
            tr = pkg.serving_trace(count, r, seed=4)
            reqs = requests()
            for q, t in zip(reqs, tr):
                q.arrival_s = t[4]
            _, s2 = pkg.serve_run(reqs, slots=slots, batching=batching)
            rows.append({"rate_per_s": r, "p50_ms": s2["p50_s"] * 1e3,
                         "p99_ms": s2["p99_s"] * 1e3,
                         "aggregate_tflops": s2["total_flop"] / s2["wall_s"] / 1e12,
                         "batches": s2["batches"], "batched_requests": s2["batched_requests"]})
        return rows

    out["rate_sweep"] = sweep_at((8000.0, 16000.0, 48000.0, 64000.0, 96000.0), True)
    # the same trace with a stream per request (no grouped launches)
    out["rate_sweep_unbatched"] = sweep_at((32000.0, 64000.0, 96000.0), False)

    def live_at(r):
        """The same requests through the live server (omb_server_*), submitted
        from this thread at the trace's arrival times (busy-wait pacing);
        latency runs from each submission to its observed completion."""
        tr = pkg.serving_trace(count, r, seed=4)
        reqs = requests()
        max_bytes = max(pkg.serve_request_bytes(q) for q in reqs)
        srv = pkg.Server(slots=slots, max_request_bytes=max_bytes)
        lags, tickets = [], []
        t0 = time.perf_counter()
        for q, t in zip(reqs, tr):
            while time.perf_counter() - t0 < t[4]:
                pass
            lags.append(time.perf_counter() - t0 - t[4])
            tickets.append(srv.submit(q))
        done = [srv.wait(t) for t in tickets]
        s3 = srv.close()
        return {"rate_per_s": r, "completed": sum(1 for d in done if d["status"] == 0),
                "p50_ms": s3["p50_s"] * 1e3, "p99_ms": s3["p99_s"] * 1e3,
                "aggregate_tflops": s3["total_flop"] / s3["wall_s"] / 1e12,
                "batches": s3["batches"],
                "submit_lag_p99_ms": float(np.percentile(lags, 99)) * 1e3}

    out["live_submission"] = [live_at(r) for r in (rate, 64000.0)]
    res1, st1 = pkg.serve_run(requests(), slots=1, batching=False)
    out["fifo_one_slot"] = {"p50_ms": st1["p50_s"] * 1e3, "p99_ms": st1["p99_s"] * 1e3,
                            "aggregate_tflops": st1["total_flop"] / st1["wall_s"] / 1e12}
    del pools
    return out


def _mean_time(fn, iters):
    """Mean wall time of `iters` calls after one untimed call (the reference's
    amt::benchmark times MaxIter calls, include/benchmark.hpp:34-51)."""
    fn()
    times = []
    for _ in range(iters):
        t0 = time.perf_counter()
        fn()
        times.append(time.perf_counter() - t0)
    return float(np.mean(times))


def c1_leg(pkg, lib, torch, dev, sp, timer):
    """Config 1: row-major 1024^3, A and B U[-1,1) from seed 0, C = 0, on the
    oracle (the reference algorithm, all host cores, MaxIter 4 as
    src/mtm.cpp:373) and on the GPU kernels (20 timed calls), with the
    max- and Frobenius-relative errors against an FP64 product."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import helpers

    n = 1024
    log("config 1: mtm 1024^3 (seed 0), oracle and GPU")
    cores = host_cores()
    a = helpers.rng_uniform(n * n, 0, 0).reshape(n, n)
    b = helpers.rng_uniform(n * n, 0, 1).reshape(n, n)
    truth = a.astype(np.float64) @ b.astype(np.float64)
    blk = helpers.host_blocking()
    c = np.zeros((n, n), np.float32)

    def cpu_call():
        c[:] = 0
        helpers.oracle_mtm(c, a, b, blk, cores)

    t_cpu = _mean_time(cpu_call, 4)
    ref = helpers.rel_errors(c, truth)
    flop = 2.0 * n ** 3
    out = {"flop_2mnk": flop, "flop_reference_count": float(n * n * (2 * n - 1)),
           "oracle": {"ms": t_cpu * 1e3, "tflops": flop / t_cpu / 1e12, "cores": cores,
                      "max_rel": ref[0], "frob_rel": ref[1]}}
    A = torch.from_numpy(a).to(dev)
    B = torch.from_numpy(b).to(dev)
    C = torch.zeros((n, n), device=dev)
    torch.cuda.synchronize()
    for mode, name in ((pkg.PLAIN_F32, "gpu_plain_f32"), (pkg.SPLIT_BF16, "gpu_split_bf16_p6")):
        fn = lambda: lib.omb_sgemm(n, n, n, 1.0, A.data_ptr(), n, 1, B.data_ptr(), n, 1, 0.0,
                                   C.data_ptr(), n, 1, mode, sp)
        pkg.check(fn())
        torch.cuda.synchronize()
        err = helpers.rel_errors(C.cpu().numpy(), truth)
        t = timer.run(fn, 20)
        out[name] = {"us": t * 1e6, "tflops": flop / t / 1e12, "max_rel": err[0],
                     "frob_rel": err[1], "ratio_vs_oracle": [err[0] / ref[0], err[1] / ref[1]]}
    return out


def cpu_reference_legs(pkg, torch, dev, stream, refs=None):
    """The oracle (the reference algorithm, OpenMP on all host cores) on the
    bandwidth-bound configurations at full size: SGEMV 16384^2 in both
    layouts (MaxIter 10, src/mtv.cpp:332) and SDOT 2^28 (MaxIter 100,
    src/dot.cpp:198), inputs from seed 1 made by the device generator
    (bit-identical to the host twin) and copied to the host."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import helpers

    sp = stream.cuda_stream
    cores = host_cores()
    blk = helpers.host_blocking()
    out = {"cores": cores}
    n = 16384
    log("cpu reference path: oracle mtv 16384^2 (both layouts) and dot 2^28")
    with torch.cuda.stream(stream):
        Ad = torch.empty(n * n, device=dev)
        xd = torch.empty(n, device=dev)
    pkg.fill_uniform(Ad, 1, 0, stream=sp)
    pkg.fill_uniform(xd, 1, 1, stream=sp)
    torch.cuda.synchronize()
    a = Ad.cpu().numpy().reshape(n, n)
    x = xd.cpu().numpy()
    del Ad
    y = np.zeros(n, np.float32)
    nbytes = 4.0 * (n * n + 2 * n)
    refs = refs or {}

    def errs(yv, want):
        d = yv.astype(np.float64) - want
        return {"max_rel": float(np.abs(d).max() / np.abs(want).max()),
                "frob_rel": float(np.linalg.norm(d) / np.linalg.norm(want))}

    t = _mean_time(lambda: helpers.oracle_mtv(y, a, x, blk, cores), 10)
    out["mtv_row_major_16384"] = {"ms": t * 1e3, "gbps": nbytes / t / 1e9}
    if "y64" in refs:
        out["mtv_row_major_16384"].update(errs(y, refs["y64"]))      # y = A x overwrites
    t = _mean_time(lambda: helpers.oracle_mtv(y, a.T, x, blk, cores), 10)   # column-major view
    out["mtv_col_major_16384"] = {"ms": t * 1e3, "gbps": (nbytes + 4 * n) / t / 1e9}
    if "y64_col" in refs:
        y[:] = 0
        helpers.oracle_mtv(y, a.T, x, blk, cores)                     # y += A^T x from 0
        out["mtv_col_major_16384"].update(errs(y, refs["y64_col"]))
    del a
    nd = 1 << 28
    with torch.cuda.stream(stream):
        xd = torch.empty(nd, device=dev)
        yd = torch.empty(nd, device=dev)
    pkg.fill_uniform(xd, 1, 0, stream=sp)
    pkg.fill_uniform(yd, 1, 1, stream=sp)
    torch.cuda.synchronize()
    xh, yh = xd.cpu().numpy(), yd.cpu().numpy()
    del xd, yd
    t = _mean_time(lambda: helpers.oracle_dot(xh, yh, blk, cores), 100)
    out["dot_2^28"] = {"ms": t * 1e3, "gbps": 8.0 * nd / t / 1e9}
    if "dot64" in refs:
        err = abs(helpers.oracle_dot(xh, yh, blk, cores) - refs["dot64"])
        out["dot_2^28"]["rel_err"] = err / abs(refs["dot64"])
        out["dot_2^28"]["err_over_sum_abs"] = err / refs["dot_sum_abs"]
    return out


def _getconf(name: str) -> int:
    """One cache parameter as getconf(1) reports it (0 when unknown)."""
    import subprocess
    try:
        return int(subprocess.run(["getconf", name], capture_output=True, text=True).stdout)
    except (OSError, ValueError):
        return 0


def host_cache_keys(blk) -> dict:
    """Flat record of the host's cache geometry as sysconf reports it and of
    the blocking the oracle derives from it (MR/NR/KC/MC/NC, the mtv and dot
    blocks).  The reference's Linux probe puts the L3 associativity in the
    L2 slot (include/cache_manager.hpp:133-135); the blocking follows that."""
    out = {}
    for key, name in (("l1d", "LEVEL1_DCACHE"), ("l2", "LEVEL2_CACHE"), ("l3", "LEVEL3_CACHE")):
        out[f"{key}_bytes"] = _getconf(name + "_SIZE")
        out[f"{key}_assoc"] = _getconf(name + "_ASSOC")
        out[f"{key}_line"] = _getconf(name + "_LINESIZE")
    for k, v in blk.as_dict().items():
        out[f"blk_{k}"] = v
    return out


def cpu_baseline_child(budget_s: float, cores: int):
    """Run cpu_baseline in a child process that loads no GPU runtime and no
    other OpenMP runtime, with the oracle's team placed one thread per core
    (OMP_PLACES=cores, OMP_PROC_BIND=close, OMP_NUM_THREADS=cores).  Started
    before this process imports torch."""
    import subprocess
    env = dict(os.environ, OMP_PLACES="cores", OMP_PROC_BIND="close",
               OMP_NUM_THREADS=str(cores))
    log(f"cpu baseline: child process, {cores} threads, OMP_PLACES=cores OMP_PROC_BIND=close")
    out = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-baseline-child",
                          "--cpu-budget", str(budget_s), "--cpu-cores", str(cores)],
                         capture_output=True, text=True, env=env)
    sys.stderr.write(out.stderr)
    if out.returncode != 0:
        raise RuntimeError(f"cpu baseline child failed ({out.returncode})")
    return json.loads(out.stdout.strip().splitlines()[-1])


def cpu_baseline(budget_s: float, cores: int | None = None):
    """The oracle's mtm (the reference algorithm: BLIS 5-loop, MR=16, NR=6,
    KC/MC/NC from this host's caches, OpenMP over all visible cores) on the
    8192^3 FP32 product; mean of up to 4 timed calls after one warm-up
    (include/benchmark.hpp:34-51, MaxIter = 4 for mtm)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import helpers

    n = 8192
    cores = cores or host_cores()
    log(f"cpu baseline: oracle mtm 8192^3 on {cores} threads")
    a = helpers.rng_uniform(n * n, 1, 0).reshape(n, n)
    b = helpers.rng_uniform(n * n, 1, 1).reshape(n, n)
    c = np.zeros((n, n), np.float32)
    blk = helpers.host_blocking()
    t0 = time.perf_counter()
    helpers.oracle_mtm(c, a, b, blk, cores)
    t_first = time.perf_counter() - t0
    iters = int(max(1, min(4, (budget_s - t_first) // max(t_first, 1e-3))))
    times = []
    for _ in range(iters):
        c[:] = 0
        t0 = time.perf_counter()
        helpers.oracle_mtm(c, a, b, blk, cores)
        times.append(time.perf_counter() - t0)
    mean = float(np.mean(times))
    flop = 2.0 * n * n * n
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    res = {"value": round(flop / mean / 1e12, 4), "unit": "TFLOP/s", "cores": cores,
           "kind": "port",
           "sample": f"full 8192^3 FP32 mtm (C += A*B), mean of {iters} timed calls after "
                     f"1 warm-up ({t_first:.2f} s)",
           "median_s": float(np.median(times)), "mean_s": mean, "cpu_model": model,
           "omp_places": os.environ.get("OMP_PLACES"),
           "omp_proc_bind": os.environ.get("OMP_PROC_BIND"),
           "omp_num_threads": os.environ.get("OMP_NUM_THREADS"),
           "blocking": blk.as_dict(), "machine": platform.machine()}
    res.update(host_cache_keys(blk))
    return res


if __name__ == "__main__":
    main()
