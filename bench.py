"""Benchmark of the batched-FIR hot path on one MI355X.

Prints ONE JSON line.  The headline (``value``) is the long-tap workload in
BF16x3 on MFMA:

  longtap_k512   x [2048, 262144] f32 ~ N(0,1), 1% of samples scaled x1e3,
                 512-tap Kaiser(beta=8) band-pass 0.1-0.3 of Nyquist

A step is one pass of the FIR over the whole batch, inputs already resident
in HBM.  ``value`` = output samples / second over the K timed steps (device
time, HIP events on the launch stream).  Beside it the line carries, measured
after the timed region on the same inputs:
  compare.f32 / compare.fp16         the other arithmetic paths, Gsamples/s
  / compare.fft32                     (f32: FP32, on the FP32 matrix cores up to
                                      K = 513; fft32: FP32 FFT overlap-save,
                                      normwise error) and rel-L2 vs the FP64 oracle
  compare.fp32_direct_valu            the FP32 direct form on the vector units
                                      (fused multiply-adds; what F32 runs above
                                      K = 513)
  bandwidth_workload                  x [4096, 65536], 33-tap low-pass, FP32:
                                      GB/s against the 8 TB/s HBM peak
  roofline                            the dominant kernel of the headline:
                                      algorithmic TFLOP/s against 2.5 PFLOP/s,
                                      the issued MFMA rate (mfma_issued_frac),
                                      and from the committed rocprofv3
                                      counters: MFMA-busy fraction of SIMD
                                      cycles and the shader clock
  cpu_baseline                        the oracle's restatement of the
                                      reference CPU loop, one thread, on whole
                                      rows of the same workload (~10 s)

Other workloads (``--workload``): longtap_k2048, direct_k33, direct_k129,
longtap_k16384 (x [256, 262144], 16384-tap band-pass, FFT32 in two 8192-tap
blocks, compared with the FP32 direct form), ragged (10000 CSR signals), vjp_k255 (forward + dL/dx + dL/db), resample_2_3
(float64 polyphase resampler, x [4096, 65536], up 2, down 3, the reference's
31-tap filter), sosfilt_butter6 (float64 IIR cascade, x [4096, 65536], the
reference's sosfilt benchmark filter butter(6, 0.15): 3 sections; filtfilt with
--precision filtfilt, the reference loop's exact order with --precision
reference_order), copy_1gib (plain device copy: attainable bandwidth and
HBM-counter calibration).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0   # dense bf16/fp16 MFMA, MI355X
FP32_PEAK_TFLOPS = 157.3
F32_MFMA_MAX_K = 513  # the F32 mode's FP32-MFMA kernel takes K up to here

METRIC = "FIR Gsamples/s & %roofline: bf16x3 MFMA vs FP32 direct at rel L2 err<=1e-6"

DENSE = {
    "longtap_k512": dict(B=2048, N=262144, K=512, taps="bandpass", precision="bf16x3",
                         seed=0x5C120002, outliers=0.01),
    "longtap_k2048": dict(B=2048, N=262144, K=2048, taps="bandpass", precision="bf16x3",
                          seed=0x5C120002, outliers=0.01),
    "direct_k33": dict(B=4096, N=65536, K=33, taps="lowpass", precision="f32",
                       seed=0x5C120001, outliers=0.0),
    "direct_k129": dict(B=4096, N=65536, K=129, taps="lowpass", precision="f32",
                        seed=0x5C120001, outliers=0.0),
    # a filter far longer than the MFMA kernel takes: FFT32 in two 8192-tap blocks
    "longtap_k16384": dict(B=256, N=262144, K=16384, taps="bandpass", precision="fft32",
                           seed=0x5C120002, outliers=0.01),
}
# Longest filters the MFMA precisions take (LDS of one CU).
MFMA_MAX_K = {"bf16x3": 6689, "fp16": 23457}
HEADLINE = "longtap_k512"


def mfma_steps(K):
    return (K + 14) // 16 + 2


def make_taps(kind, K):
    from scir_amd import design
    if kind == "lowpass":
        return design.lowpass(K, 0.2)
    return design.bandpass_kaiser(K, (0.1, 0.3), 8.0)


def describe_taps(kind, K):
    if kind == "lowpass":
        return f"{K}-tap Hamming low-pass 0.2 of Nyquist, unit DC gain"
    return f"{K}-tap Kaiser(8) band-pass 0.1-0.3 of Nyquist"


def load_profile(tag, key):
    """A figure from a committed rocprofv3 --pmc summary under profiles/."""
    path = os.path.join(ROOT, "profiles", f"{tag}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get(key)


def load_traffic(tag):
    """HBM bytes per launch of the kernel (FETCH_SIZE and WRITE_SIZE runs)."""
    return load_profile(tag, "hbm_bytes_per_launch")


def timed(fn, steps, warmup, stream):
    """Run fn warmup times, then time `steps` back-to-back calls with one pair
    of HIP events on `stream` around all of them (no event packets between
    the launches).  Returns (ms_per_step, wall seconds); ms_per_step is the
    average launch duration including the gaps between launches."""
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    t_wall = time.perf_counter()
    start.record(stream)
    for _ in range(steps):
        fn()
    end.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_wall
    return start.elapsed_time(end) / steps, wall


def cpu_baseline(rows_np, b, budget_s):
    """Single-thread oracle (restatement of crates/scir-gpu/src/lib.rs:1134-1152)
    over whole rows of the workload until ~budget_s of CPU time is spent."""
    from oracle import oracle as orc
    taps = np.ascontiguousarray(b[::-1])
    done = 0
    t0 = time.perf_counter()
    while done < rows_np.shape[0]:
        orc.fir_ref_f32(rows_np[done:done + 1], taps)
        done += 1
        if time.perf_counter() - t0 >= budget_s:
            break
    dt = time.perf_counter() - t0
    n = rows_np.shape[1]
    return dict(value=round(done * n / dt / 1e9, 6), unit="Gsamples/s", cores=1, kind="port",
                host=host_cpu(),
                sample=f"{done} whole rows x {n} samples of the same workload, "
                       f"1 thread, {dt:.1f} s")


def host_cpu():
    """The host CPU model (/proc/cpuinfo) and the logical CPU counts."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None
    return dict(model=model, nproc=os.cpu_count(), usable=usable)


def cpu_all_cores(rows_np, b, budget_s):
    """The library's CPU arm (the reference loop's order, rows spread over
    host threads) on whole rows of the workload, all host threads the box
    grants (at most 16): the multi-core CPU figure beside the 1-thread
    baseline.  Not the oracle: a product path, reported, never the value."""
    import ctypes
    from scir_amd import _lib
    threads = min(16, len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity")
                  else (os.cpu_count() or 1))
    x = np.ascontiguousarray(rows_np, dtype=np.float32)
    bb = np.ascontiguousarray(b, dtype=np.float32)  # lfilter order, as the C ABI takes
    y = np.empty_like(x)
    rows, n = x.shape
    vp = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    t0 = time.perf_counter()
    reps = 0
    while True:
        _lib.check(_lib.lib().scir_hip_fir_f32_cpu(vp(x), n, vp(bb), bb.size, rows, n, vp(y), n,
                                                   threads))
        reps += 1
        if time.perf_counter() - t0 >= budget_s:
            break
    dt = time.perf_counter() - t0
    n = rows_np.size * reps
    return dict(value=round(n / dt / 1e9, 6), unit="Gsamples/s", cores=threads, host=host_cpu(),
                impl="library CPU arm (reference loop order, row-partitioned threads)",
                sample=f"{reps} passes over {rows_np.shape[0]} whole rows x {rows_np.shape[1]} "
                       f"samples, {threads} threads, {dt:.1f} s")


def host_tier(x, b_np, reps=3):
    """The drop-in host tier (scir_hip_fir_f32_host through the
    fir1d_batched_f32_hip mirror): pageable host arrays in and out, so every
    call moves 8 B per sample over PCIe.  Reported beside ``value``, never as
    it.  Returns Gsamples/s of the best of ``reps`` calls (wall clock)."""
    from scir_amd import gpu
    x_np = x.cpu().numpy()
    taps = np.ascontiguousarray(b_np[::-1])
    gpu.fir1d_batched_f32_hip(x_np, taps)  # grows the library's device buffers
    best = float("inf")
    for _ in range(reps):
        t0 = time.perf_counter()
        gpu.fir1d_batched_f32_hip(x_np, taps)
        best = min(best, time.perf_counter() - t0)
    n = x_np.size
    # The same entry point writing into one preallocated array: the mirror's
    # fresh output (like the reference's Vec) takes a page fault per page on
    # first touch, which this second figure leaves out.
    import ctypes
    from scir_amd import _lib
    L = _lib.lib()
    xa = np.ascontiguousarray(x_np, np.float32)
    ba = np.ascontiguousarray(taps, np.float32)
    ya = np.empty_like(xa)
    B, N = xa.shape
    vp = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    call = lambda: _lib.check(L.scir_hip_fir_f32_host(vp(xa), N, vp(ba), ba.size, B, N, vp(ya),
                                                       N, _lib.precision_code("auto")))
    call()
    best_pre = float("inf")
    for _ in range(reps):
        t0 = time.perf_counter()
        call()
        best_pre = min(best_pre, time.perf_counter() - t0)
    return dict(gsamples_s=round(n / best / 1e9, 3), ms_per_call=round(best * 1e3, 2),
                preallocated_out_gsamples_s=round(n / best_pre / 1e9, 3),
                preallocated_out_ms_per_call=round(best_pre * 1e3, 2),
                note="host arrays in and out (pageable), PCIe-inclusive, best of "
                     f"{reps} calls; gsamples_s allocates the result per call as the "
                     "reference's API does, preallocated_out_* writes into one array; "
                     "not the headline")


def reference_cpu_case(stream):
    """The reference's own CPU-runnable case, [64, 16384] f32 ~ N(0,1) with the
    33-tap Hamming low-pass: the oracle's 1-thread restatement of the reference
    loop over the whole case, the library's CPU arm on all usable threads, the
    device tier (inputs resident) and the drop-in host tier.  A parity case
    reported beside the headline, not a bench line of its own."""
    import torch
    from scir_amd import design, device
    from oracle import oracle as orc
    x_np = np.random.default_rng(0x5C120000).standard_normal((64, 16384)).astype(np.float32)
    b_np = design.lowpass(33, 0.2).astype(np.float32)
    n = x_np.size
    reps, t0 = 0, time.perf_counter()
    while reps < 3 or time.perf_counter() - t0 < 1.0:
        y_ref = orc.fir_ref_f32(x_np, b_np[::-1])
        reps += 1
    cpu1 = n * reps / (time.perf_counter() - t0) / 1e9
    x = torch.tensor(x_np, device="cuda")
    b = torch.tensor(b_np, device="cuda")
    y = torch.empty_like(x)
    ms, _ = timed(lambda: device.fir(x, b, "f32_exact", out=y), 200, 50, stream)
    exact = bool(np.array_equal(y.cpu().numpy(), y_ref))
    return dict(workload="[64, 16384] f32, 33-tap Hamming low-pass 0.2 of Nyquist",
                cpu_1thread_gsamples_s=round(cpu1, 4),
                cpu_all_cores=cpu_all_cores(x_np, b_np, 1.0),
                gpu_f32_exact_gsamples_s=round(n / (ms * 1e-3) / 1e9, 2),
                gpu_us_per_call=round(ms * 1e3, 1),
                gpu_bitwise_equal_reference_loop=exact,
                host_tier=host_tier(x, b_np, reps=5))


def dense_setup(name):
    import torch
    from scir_amd import device
    wl = DENSE[name]
    x = device.synth_normal(wl["B"], wl["N"], seed=wl["seed"], outlier_frac=wl["outliers"],
                            outlier_scale=1000.0)
    b_np = make_taps(wl["taps"], wl["K"])
    return wl, x, torch.tensor(b_np, device="cuda"), b_np


def rel_l2_rows(x, y, b_np, rows):
    from oracle import oracle as orc
    xs = x[rows].cpu().numpy()
    return orc.rel_l2(y[rows].cpu().numpy(), orc.lfilter_f64(xs, b_np))


def roofline_for(wl, name, prec, avg_launch_s):
    samples = wl["B"] * wl["N"]
    if prec in ("bf16x3", "fp16"):
        alg = 2.0 * wl["K"] * samples
        ach = alg / avg_launch_s / 1e12
        prods = 6 if prec == "bf16x3" else 1
        issued = 32.0 * mfma_steps(wl["K"]) * prods * samples
        # Gsamples/s with every MFMA pipe issuing back to back at the nominal
        # peak: the ceiling of this arithmetic at this K
        ceiling = BF16_PEAK_TFLOPS * 1e12 / (issued / samples) / 1e9
        return dict(bound="mfma", achieved=round(ach, 2), peak=BF16_PEAK_TFLOPS, unit="TFLOP/s",
                    frac=round(ach / BF16_PEAK_TFLOPS, 4),
                    traffic=load_traffic(f"fir_mfma_{prec}_{name}_pmc"),
                    mfma_issued_tflops=round(issued / avg_launch_s / 1e12, 1),
                    mfma_issued_frac=round(issued / avg_launch_s / 1e12 / BF16_PEAK_TFLOPS, 4),
                    mfma_busy_frac=load_profile(f"fir_mfma_{prec}_{name}_util", "mfma_busy_frac"),
                    sclk_ghz=load_profile(f"fir_mfma_{prec}_{name}_util", "sclk_ghz"),
                    ceiling_gsamples_s=round(ceiling, 1),
                    ceiling_note=f"{prods} MFMA products x {mfma_steps(wl['K'])} half-steps of "
                                 "32x32x16 per 32x32 output tile at 2.5 PFLOP/s")
    flops = 2.0 * wl["K"] * samples
    # F32 up to K = 513 runs on the FP32 matrix cores (a 16 x 16(NG) Toeplitz
    # block per 256 outputs, NG = ceil((K-1)/16) + 1); FP32 MFMA and FP32
    # vector share the 157.3 TFLOP/s peak on MI355X.
    f32_mfma = prec == "f32" and wl["K"] <= F32_MFMA_MAX_K
    extra = {}
    if f32_mfma:
        ng = (wl["K"] + 14) // 16 + 1
        issued = 32.0 * ng * samples / avg_launch_s / 1e12
        extra = dict(kernel="fir_direct_mfma_kernel (FP32 MFMA, v_mfma_f32_16x16x4_f32)",
                     mfma_issued_tflops=round(issued, 1),
                     mfma_issued_frac=round(issued / FP32_PEAK_TFLOPS, 4),
                     mfma_busy_frac=load_profile(f"fir_direct_{prec}_{name}_util",
                                                 "mfma_busy_frac"))
    if prec == "fft32" or flops / (8.0 * samples) < 78.6 / 8.0:
        # HBM-bound (the FFT path's work per sample is O(log K))
        alg = 8.0 * samples
        ach = alg / avg_launch_s / 1e9
        fam = "fft" if prec == "fft32" else "direct"
        return dict(bound="hbm", achieved=round(ach, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(ach / HBM_PEAK_GBS, 4),
                    traffic=load_traffic(f"fir_{fam}_{prec}_{name}_pmc"), **extra)
    ach = flops / avg_launch_s / 1e12
    return dict(bound="fp32", achieved=round(ach, 2), peak=FP32_PEAK_TFLOPS, unit="TFLOP/s",
                frac=round(ach / FP32_PEAK_TFLOPS, 4),
                traffic=load_traffic(f"fir_direct_{prec}_{name}_pmc"), **extra)


def run_dense(args):
    import torch
    from scir_amd import device
    name = args.workload
    wl, x, b, b_np = dense_setup(name)
    prec = args.precision or wl["precision"]
    B, N, K = wl["B"], wl["N"], wl["K"]
    y = torch.empty_like(x)
    stream = torch.cuda.current_stream()
    ms, wall = timed(lambda: device.fir(x, b, prec, out=y), args.steps, args.warmup, stream)
    value = B * N / (ms * 1e-3) / 1e9
    rows = [0, B // 2, B - 1] if K <= 4096 else [0, B - 1]
    out = dict(metric=METRIC, value=round(value, 2), unit="Gsamples/s", n_gpus=1,
               steps=args.steps, warmup=args.warmup, ms_per_step=round(ms, 4),
               higher_is_better=True, scaling="weak", vs_baseline=None, dtype=prec,
               data="synthetic",
               config=dict(workload=name, rows=B, samples_per_row=N, taps=K,
                           taps_design=describe_taps(wl["taps"], K),
                           outliers="1% of samples x1e3" if wl["outliers"] else "none",
                           precision=prec),
               rel_l2_vs_fp64=rel_l2_rows(x, y, b_np, rows), wall_s=round(wall, 3),
               roofline=roofline_for(wl, name, prec, ms * 1e-3))
    if not args.quick:
        comp = {}
        for other in ("f32", "fp16", "bf16x3", "fft32"):
            if other == prec or K > MFMA_MAX_K.get(other, K):
                continue
            # 20 untimed calls: the first FFT32 call on a device also builds its
            # twiddle tables, and the shader clock takes ~20 ms to settle after
            # the switch from another kernel (measured after the FP16 runs).
            ms_o, _ = timed(lambda: device.fir(x, b, other, out=y), max(5, args.steps // 2),
                            20, stream)
            comp[other] = dict(
                gsamples_s=round(B * N / (ms_o * 1e-3) / 1e9, 2), ms_per_step=round(ms_o, 4),
                rel_l2_vs_fp64=rel_l2_rows(x, y, b_np, rows))
            if other == "f32":
                comp[other]["kernel"] = ("fir_direct_mfma_kernel (FP32 MFMA)"
                                         if K <= F32_MFMA_MAX_K else "fir_direct_dma_kernel (FP32 VALU)")
        # The FP32 direct form on the VALU (fused multiply-adds, the LDS-DMA
        # loader), which F32 runs only above K = 513; below, F32 runs on the
        # FP32 matrix cores and this is the same arithmetic class timed on
        # the vector units.
        if K <= F32_MFMA_MAX_K:
            from scir_amd import _lib
            _lib.check(_lib.lib().scir_hip_tune_set(b"direct.loader", 2))
            try:
                ms_o, _ = timed(lambda: device.fir(x, b, "f32", out=y), max(5, args.steps // 2),
                                20, stream)
            finally:
                _lib.check(_lib.lib().scir_hip_tune_set(b"direct.loader", -1))
            comp["fp32_direct_valu"] = dict(
                gsamples_s=round(B * N / (ms_o * 1e-3) / 1e9, 2), ms_per_step=round(ms_o, 4),
                rel_l2_vs_fp64=rel_l2_rows(x, y, b_np, rows),
                kernel="fir_direct_dma_kernel (FP32 VALU)")
        out["compare"] = comp
        if name != "direct_k33":
            wl2, x2, b2, _ = dense_setup("direct_k33")
            y2 = torch.empty_like(x2)
            # 60 untimed launches (~26 ms): the clock settles after a load step.
            ms2, _ = timed(lambda: device.fir(x2, b2, "f32", out=y2), 20, 60, stream)
            r2 = roofline_for(wl2, "direct_k33", "f32", ms2 * 1e-3)
            out["bandwidth_workload"] = dict(
                workload="direct_k33", rows=wl2["B"], samples_per_row=wl2["N"], taps=33,
                precision="f32", gsamples_s=round(wl2["B"] * wl2["N"] / (ms2 * 1e-3) / 1e9, 2),
                ms_per_step=round(ms2, 4), roofline=r2,
                host_tier=host_tier(x2, make_taps(wl2["taps"], 33)))
            del x2, y2
        else:
            out["host_tier"] = host_tier(x, b_np)
    if not args.no_cpu_baseline:
        nrows = min(B, 256)
        out["cpu_baseline"] = cpu_baseline(x[:nrows].cpu().numpy(), b_np, args.cpu_budget)
        out["cpu_all_cores"] = cpu_all_cores(x[:min(B, 64)].cpu().numpy(), b_np,
                                             min(5.0, args.cpu_budget))
        if not args.quick:
            out["reference_cpu_case"] = reference_cpu_case(stream)
    if name == HEADLINE and not args.quick:
        del x, y
        torch.cuda.empty_cache()
        out["configs"] = other_configs(stream)
    return out


def other_configs(stream):
    """Every other configuration of the benchmark set, measured in the same
    run as the headline, each with its roofline (HBM traffic from the
    committed counter summaries where one exists):
      c1_direct_k129    x [4096, 65536], 129-tap low-pass, FP32 direct form
      c2_longtap_k2048  x [2048, 262144] with 1 % outliers x1e3, 2048-tap
                        Kaiser band-pass: BF16x3, FP32 direct, FFT32, FP16
      c3_ragged         10000 CSR signals (one persistent launch, its HIP-graph
                        replay and 10000 per-signal launches; F32 and BF16x3)
      c4_vjp_k255       forward + dL/dx + dL/db on x [1024, 65536], K = 255
    (c1 at K = 33 is ``bandwidth_workload``; c2 at K = 512 is the headline.)"""
    import torch
    from scir_amd import device
    from oracle import oracle as orc
    res = {}
    # c1, K = 129: FP32-vector bound
    wl, x1, b1, b1_np = dense_setup("direct_k129")
    y1 = torch.empty_like(x1)
    ms, _ = timed(lambda: device.fir(x1, b1, "f32", out=y1), 20, 30, stream)
    # What the drop-in's default (AUTO: the faster of F32 and BF16x3 MFMA by
    # K, same elementwise bound) runs on the same inputs; which one it took
    # is read off its output, bit for bit.
    ms_a, _ = timed(lambda: device.fir(x1, b1, "auto", out=y1), 20, 30, stream)
    ya = y1.clone()
    resolved = "other"
    for cand in ("f32", "bf16x3"):
        if torch.equal(ya, device.fir(x1, b1, cand)):
            resolved = cand
            break
    del ya
    auto = dict(resolves_to=resolved,
                gsamples_s=round(wl["B"] * wl["N"] / (ms_a * 1e-3) / 1e9, 2),
                ms_per_step=round(ms_a, 4),
                rel_l2_vs_fp64=rel_l2_rows(x1, y1, b1_np, [0, 2048, 4095]))
    ms, _ = timed(lambda: device.fir(x1, b1, "f32", out=y1), 20, 30, stream)
    res["c1_direct_k129"] = dict(
        rows=wl["B"], samples_per_row=wl["N"], taps=wl["K"], precision="f32",
        gsamples_s=round(wl["B"] * wl["N"] / (ms * 1e-3) / 1e9, 2), ms_per_step=round(ms, 4),
        rel_l2_vs_fp64=rel_l2_rows(x1, y1, b1_np, [0, 2048, 4095]),
        roofline=roofline_for(wl, "direct_k129", "f32", ms * 1e-3), auto=auto)
    del x1, y1
    # c2, K = 2048: every arithmetic on the same inputs
    wl, x2, b2, b2_np = dense_setup("longtap_k2048")
    y2 = torch.empty_like(x2)
    rows = [0, 1023, 2047]
    ref = orc.lfilter_f64(x2[rows].cpu().numpy(), b2_np)
    c2 = {}
    for prec, steps in (("bf16x3", 10), ("f32", 4), ("fft32", 10), ("fp16", 10)):
        ms, _ = timed(lambda: device.fir(x2, b2, prec, out=y2), steps, 5, stream)
        c2["fp32_direct" if prec == "f32" else prec] = dict(
            gsamples_s=round(wl["B"] * wl["N"] / (ms * 1e-3) / 1e9, 2), ms_per_step=round(ms, 4),
            rel_l2_vs_fp64=orc.rel_l2(y2[rows].cpu().numpy(), ref),
            roofline=roofline_for(wl, "longtap_k2048", prec, ms * 1e-3))
    res["c2_longtap_k2048"] = dict(rows=wl["B"], samples_per_row=wl["N"], taps=wl["K"],
                                   rel_l2_rows=rows, **c2)
    del x2, y2
    torch.cuda.empty_cache()
    res["c3_ragged"] = ragged_results(steps=5, warmup=2, precs=("bf16x3", "f32"), stream=stream)
    torch.cuda.empty_cache()
    res["c4_vjp_k255"] = vjp_results(steps=20, warmup=10, prec="bf16x3", stream=stream)
    return res


def ragged_rel_l2(x, y, offs, boffs, taps, n=8):
    """Relative L2 error vs the FP64 oracle over a seeded subset of signals
    (the shortest, the longest and n - 2 drawn at random)."""
    from oracle import oracle as orc
    lens = np.diff(offs)
    r = np.random.default_rng(0x5C12)
    pick = sorted(set([int(np.argmin(lens)), int(np.argmax(lens))] +
                      [int(v) for v in r.integers(0, lens.size, n - 2)]))
    ys, refs = [], []
    for s in pick:
        seg = slice(int(offs[s]), int(offs[s + 1]))
        ys.append(y[seg].cpu().numpy())
        refs.append(orc.lfilter_f64(x[seg].cpu().numpy(), taps[boffs[s]:boffs[s + 1]])[0])
    return orc.rel_l2(np.concatenate(ys), np.concatenate(refs)), len(pick)


def ragged_roofline(prec, macs, issued, ms):
    ach = 2.0 * macs / (ms * 1e-3) / 1e12
    if prec in ("bf16x3", "fp16"):
        prods = 6 if prec == "bf16x3" else 1
        iss = prods * issued / (ms * 1e-3) / 1e12
        return dict(bound="mfma", achieved=round(ach, 2), peak=BF16_PEAK_TFLOPS,
                    unit="TFLOP/s", frac=round(ach / BF16_PEAK_TFLOPS, 4),
                    traffic=load_traffic(f"fir_ragged_{prec}_ragged_pmc"),
                    mfma_issued_tflops=round(iss, 1),
                    mfma_issued_frac=round(iss / BF16_PEAK_TFLOPS, 4),
                    mfma_busy_frac=load_profile(f"fir_ragged_{prec}_ragged_util",
                                                "mfma_busy_frac"),
                    sclk_ghz=load_profile(f"fir_ragged_{prec}_ragged_util", "sclk_ghz"))
    return dict(bound="fp32", achieved=round(ach, 2), peak=FP32_PEAK_TFLOPS,
                unit="TFLOP/s", frac=round(ach / FP32_PEAK_TFLOPS, 4),
                traffic=load_traffic(f"fir_ragged_{prec}_ragged_pmc"))


def ragged_results(steps, warmup, precs, stream, baselines=True):
    """The ragged workload (10000 CSR signals, lengths log-uniform in
    [256, 2^20], K uniform in [15, 1023]): per precision one persistent
    launch (timed over `steps`), its HIP-graph replay, and the baseline of one
    dense launch per signal; rel-L2 vs FP64 on a seeded subset of signals.
    baselines=False leaves out the graph replay and the per-signal launches
    (10 000 dispatches, slow under a counter profiler)."""
    import torch
    from scir_amd import configs, device
    from scir_amd.ragged import Graph, RaggedPlan
    offs, boffs, taps = configs.ragged_layout()
    total = int(offs[-1])
    x = device.synth_normal(1, total, seed=configs.SEED_BASE + 3).view(-1)
    bt = torch.tensor(taps, device="cuda")
    y = torch.empty_like(x)
    lens, ks = np.diff(offs), np.diff(boffs)
    macs = float(np.sum(lens * ks))
    # MFMA flops issued: 32x32x16 blocks of 8192-output segments, per product.
    segs = (lens + 8191) // 8192
    issued = float(np.sum(segs * 8192 * 32.0 * np.array([mfma_steps(int(k)) for k in ks])))
    res = dict(signals=int(offs.size - 1), samples=total, lengths="log-uniform [256, 2^20]",
               taps="uniform [15, 1023], random Hamming low-pass")
    for prec in precs:
        plan = RaggedPlan(offs, boffs, prec)
        ms, _ = timed(lambda: plan.execute(x, bt, y), steps, warmup, stream)
        err, nsig = ragged_rel_l2(x, y, offs, boffs, taps)
        r = dict(gsamples_s=round(total / (ms * 1e-3) / 1e9, 2), ms_per_step=round(ms, 4),
                 rel_l2_vs_fp64=err, rel_l2_signals=nsig)
        if baselines:
            ms_s, _ = timed(lambda: plan.execute_per_signal(x, bt, y), 1, 1, stream)
            side = torch.cuda.Stream()
            torch.cuda.synchronize()
            with Graph(side) as g:
                plan.execute(x, bt, y)
            ms_g, _ = timed(g.launch, max(3, steps // 2), 1, side)
            g.close()
            r.update(graph_replay_ms=round(ms_g, 4), per_signal_launches_ms=round(ms_s, 3),
                     persistent_speedup=round(ms_s / ms, 2))
        plan.close()
        r["roofline"] = ragged_roofline(prec, macs, issued, ms)
        res[prec] = r
    del x, y
    return res


def run_ragged(args):
    from scir_amd import configs
    prec = args.precision or "f32"
    if prec not in ("f32", "f32_exact", "bf16x3", "fp16"):
        raise SystemExit("--precision for ragged: f32, f32_exact, bf16x3 or fp16")
    import torch
    stream = torch.cuda.current_stream()
    precs = (prec,) if args.quick else (prec,) + tuple(p for p in ("f32", "bf16x3") if p != prec)
    t0 = time.perf_counter()
    r = ragged_results(args.steps, args.warmup, precs, stream, baselines=not args.quick)
    wall = time.perf_counter() - t0
    main_r = r[prec]
    offs, _, _ = configs.ragged_layout()
    out = dict(metric=METRIC, value=main_r["gsamples_s"], unit="Gsamples/s",
               n_gpus=1, steps=args.steps, warmup=args.warmup, ms_per_step=main_r["ms_per_step"],
               higher_is_better=True, scaling="weak", vs_baseline=None, dtype=prec,
               data="synthetic",
               config=dict(workload="ragged", signals=offs.size - 1, samples=r["samples"],
                           lengths=r["lengths"], taps=r["taps"], precision=prec),
               rel_l2_vs_fp64=main_r["rel_l2_vs_fp64"], rel_l2_signals=main_r["rel_l2_signals"],
               wall_s=round(wall, 3), roofline=main_r["roofline"],
               compare=dict(per_signal_launches_ms=main_r.get("per_signal_launches_ms"),
                            graph_replay_ms=main_r.get("graph_replay_ms"),
                            persistent_speedup=main_r.get("persistent_speedup"),
                            **{p: r[p] for p in precs if p != prec}))
    return out


def vjp_results(steps, warmup, prec, stream):
    """Forward + dL/dx + dL/db for x [1024, 65536], K = 255 low-pass,
    L = sum(w * y): Gsamples/s of the whole step, and the dL/db kernel
    (FP32 MFMA) against the FP32 MFMA peak."""
    import torch
    from scir_amd import design, device
    B, N, K = 1024, 65536, 255
    x = device.synth_normal(B, N, seed=0x5C120004)
    w = device.synth_normal(B, N, seed=0x5C120005)
    b = torch.tensor(design.lowpass(K), device="cuda")
    y, dx = torch.empty_like(x), torch.empty_like(x)
    db = torch.empty(K, device="cuda")

    def step():
        device.fir(x, b, prec, out=y)
        device.vjp_dx(w, b, prec, out=dx)
        device.vjp_db(w, x, K, out=db)
    ms, wall = timed(step, steps, warmup, stream)
    ms_db, _ = timed(lambda: device.vjp_db(w, x, K, out=db), 5, 1, stream)
    flops_db = 2.0 * K * B * N
    ach = flops_db / (ms_db * 1e-3) / 1e12
    return dict(rows=B, samples_per_row=N, taps=K, step="forward + dL/dx + dL/db",
                precision=prec, gsamples_s=round(B * N / (ms * 1e-3) / 1e9, 2),
                ms_per_step=round(ms, 4), wall_s=round(wall, 3),
                db_ms=round(ms_db, 4),
                roofline=dict(bound="fp32", achieved=round(ach, 2), peak=FP32_PEAK_TFLOPS,
                              unit="TFLOP/s", frac=round(ach / FP32_PEAK_TFLOPS, 4),
                              kernel="fir_vjp_db_mfma_kernel (FP32 MFMA)",
                              traffic=load_traffic(f"fir_vjp_db_{prec}_vjp_k255_pmc"),
                              mfma_busy_frac=load_profile(f"fir_vjp_db_{prec}_vjp_k255_util",
                                                          "mfma_busy_frac")))


def run_vjp(args):
    import torch
    prec = args.precision or "bf16x3"
    r = vjp_results(args.steps, args.warmup, prec, torch.cuda.current_stream())
    return dict(metric=METRIC, value=r["gsamples_s"], unit="Gsamples/s",
                n_gpus=1, steps=args.steps, warmup=args.warmup, ms_per_step=r["ms_per_step"],
                higher_is_better=True, scaling="weak", vs_baseline=None, dtype=prec,
                data="synthetic",
                config=dict(workload="vjp_k255", rows=r["rows"], samples_per_row=r["samples_per_row"],
                            taps=r["taps"], step=r["step"], precision=prec),
                wall_s=r["wall_s"], roofline=r["roofline"])


def run_resample(args):
    """Batched float64 polyphase resampling (up 2, down 3, 31 taps): HBM
    bytes = 8 per input sample read + 8 per output sample written."""
    import ctypes
    import torch
    from scir_amd import _lib, signal
    B, N, up, down = 4096, 65536, 2, 3
    h = signal.RESAMPLE_H_2_3
    M = signal.resample_len(N, up, down, h.size)
    g = torch.Generator(device="cuda").manual_seed(0x5C120006)
    x = torch.randn((B, N), dtype=torch.float64, device="cuda", generator=g)
    hd = torch.tensor(h, device="cuda")
    y = torch.empty((B, M), dtype=torch.float64, device="cuda")
    stream = torch.cuda.current_stream()
    L = _lib.lib()
    vp = lambda t: ctypes.c_void_p(t.data_ptr())

    def step():
        _lib.check(L.scir_hip_resample_poly_f64_dev(vp(x), N, B, N, up, down, vp(hd), h.size,
                                                     vp(y), M, ctypes.c_void_p(stream.cuda_stream)))
    ms, wall = timed(step, args.steps, args.warmup, stream)
    alg = 8.0 * B * (N + M)
    gbs = alg / (ms * 1e-3) / 1e9
    out = dict(metric="polyphase resample Gsamples/s (output samples)",
               value=round(B * M / (ms * 1e-3) / 1e9, 2), unit="Gsamples/s", n_gpus=1,
               steps=args.steps, warmup=args.warmup, ms_per_step=round(ms, 4),
               higher_is_better=True, scaling="weak", vs_baseline=None, dtype="f64",
               data="synthetic",
               config=dict(workload="resample_2_3", rows=B, samples_per_row=N, up=up, down=down,
                           taps=int(h.size), outputs_per_row=M),
               wall_s=round(wall, 3),
               roofline=dict(bound="hbm", achieved=round(gbs, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                             frac=round(gbs / HBM_PEAK_GBS, 4),
                             traffic=load_traffic("resample_poly_f64_resample_2_3_pmc"),
                             kernel="resample_fixed_kernel<2,3,31> (lane windows in registers)"))
    rows = [0, B - 1]
    from oracle import oracle as orc
    xs = x[rows].cpu().numpy()
    err = max(float(np.max(np.abs(y[r].cpu().numpy() - orc.resample_poly_f64(xs[i], up, down, h))))
              for i, r in enumerate(rows))
    out["max_abs_err_vs_reference_order"] = err
    if not args.no_cpu_baseline:
        # the oracle's restatement of the reference loop, one thread, whole rows
        xh = x[:64].cpu().numpy()
        t0 = time.perf_counter()
        done = 0
        while time.perf_counter() - t0 < args.cpu_budget:  # cycles over the 64 rows
            orc.resample_poly_f64(xh[done % 64], up, down, h)
            done += 1
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = dict(value=round(done * M / dt / 1e9, 6), unit="Gsamples/s",
                                   cores=1, kind="port",
                                   sample=f"{done} whole rows (64 distinct) of the workload, "
                                          f"1 thread, {dt:.1f} s")
    return out


def run_sosfilt(args):
    """Batched float64 IIR cascade (the reference's sosfilt benchmark filter,
    crates/scir-signal/benches/signal.rs:8-18, butter(6, 0.15) = 3 sections) on
    [4096, 65536].  With 4096 rows each row is one wave's chunk: the call is a
    tiny setup launch and one filtering pass (sos_wave_kernel); its
    algorithmic HBM bytes are 8 per input sample read and 8 per output
    written."""
    import ctypes
    import torch
    from scipy import signal as ss
    from scir_amd import _lib
    B, N = 4096, 65536
    mode = args.precision or "chunked"
    if mode not in ("chunked", "reference_order", "filtfilt"):
        raise SystemExit(f"--precision for sosfilt_butter6: chunked, reference_order or "
                         f"filtfilt, not {mode}")
    sos_np = ss.butter(6, 0.15, output="sos")
    g = torch.Generator(device="cuda").manual_seed(0x5C120007)
    x = torch.randn((B, N), dtype=torch.float64, device="cuda", generator=g)
    sos = torch.tensor(sos_np, device="cuda")
    y = torch.empty_like(x)
    stream = torch.cuda.current_stream()
    L = _lib.lib()
    vp = lambda t: ctypes.c_void_p(t.data_ptr())
    fn = L.scir_hip_filtfilt_f64_dev if mode == "filtfilt" else L.scir_hip_sosfilt_f64_dev
    flags = 1 if mode == "reference_order" else 0

    def step():
        _lib.check(fn(vp(x), N, B, N, vp(sos), sos_np.shape[0], vp(y), N, flags,
                      ctypes.c_void_p(stream.cuda_stream)))
    ms, wall = timed(step, args.steps, args.warmup, stream)
    passes = 2 if mode == "filtfilt" else 1
    alg = 16.0 * B * N * passes
    gbs = alg / (ms * 1e-3) / 1e9
    tag = {"chunked": "sosfilt", "reference_order": "sosfilt_refo", "filtfilt": "filtfilt"}[mode]
    out = dict(metric="sosfilt Gsamples/s", value=round(B * N / (ms * 1e-3) / 1e9, 2),
               unit="Gsamples/s", n_gpus=1, steps=args.steps, warmup=args.warmup,
               ms_per_step=round(ms, 4), higher_is_better=True, scaling="weak",
               vs_baseline=None, dtype="f64", data="synthetic",
               config=dict(workload="sosfilt_butter6", rows=B, samples_per_row=N,
                           sections=int(sos_np.shape[0]), filter="butter(6, 0.15)", mode=mode),
               wall_s=round(wall, 3),
               roofline=dict(bound="hbm", achieved=round(gbs, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                             frac=round(gbs / HBM_PEAK_GBS, 4),
                             kernel="whole call (sos_setup_kernel + sos_wave_kernel)",
                             traffic=load_traffic(f"sos_f64_{tag}_sosfilt_butter6_pmc")))
    from oracle import oracle as orc
    rows = [0, B - 1]
    ref_fn = orc.filtfilt_f64 if mode == "filtfilt" else orc.sosfilt_f64
    xs = x[rows].cpu().numpy()
    ys = y[rows].cpu().numpy()
    ref = ref_fn(sos_np, xs)
    out["max_rel_err_vs_reference_order"] = float(np.max(np.abs(ys - ref)) / np.max(np.abs(ref)))
    if not args.no_cpu_baseline:
        # the oracle's restatement of the reference loop, one thread, whole rows
        xh = x[:64].cpu().numpy()
        t0 = time.perf_counter()
        done = 0
        while time.perf_counter() - t0 < args.cpu_budget:  # cycles over the 64 rows
            ref_fn(sos_np, xh[done % 64])
            done += 1
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = dict(value=round(done * N / dt / 1e9, 6), unit="Gsamples/s",
                                   cores=1, kind="port",
                                   sample=f"{done} whole rows (64 distinct) of the workload, "
                                          f"1 thread, {dt:.1f} s")
    return out


def run_copy(args):
    """Plain device copy of 1 GiB (2^28 floats): the attainable-bandwidth
    reference and the calibration kernel for the HBM byte counters."""
    import torch
    from scir_amd import device
    n = 1 << 28
    x = device.synth_normal(1, n, seed=1).view(-1)
    y = torch.empty_like(x)
    stream = torch.cuda.current_stream()
    ms, wall = timed(lambda: y.copy_(x), args.steps, args.warmup, stream)
    gbs = 8.0 * n / (ms * 1e-3) / 1e9
    return dict(metric="device copy GB/s (calibration)", value=round(gbs, 1), unit="GB/s",
                n_gpus=1, steps=args.steps, warmup=args.warmup, ms_per_step=round(ms, 4),
                higher_is_better=True, scaling="weak", vs_baseline=None, dtype="f32",
                data="synthetic", config=dict(workload="copy_1gib", bytes_read=4 * n,
                                              bytes_written=4 * n),
                roofline=dict(bound="hbm", achieved=round(gbs, 1), peak=HBM_PEAK_GBS,
                              unit="GB/s", frac=round(gbs / HBM_PEAK_GBS, 4), traffic=None))


def main():
    ap = argparse.ArgumentParser(description=__doc__,
                                 formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=None,
                    help="untimed steps before the timed ones (default: enough for about "
                         "25 ms of work, past the clock transient after a load step)")
    ap.add_argument("--workload", default=HEADLINE,
                    choices=sorted(DENSE) + ["ragged", "vjp_k255", "resample_2_3",
                                             "sosfilt_butter6", "copy_1gib"])
    ap.add_argument("--precision", default=None)
    ap.add_argument("--quick", action="store_true",
                    help="skip the side comparisons (and, for ragged, the per-signal launches)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=10.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (--gpus {args.gpus} requested). "
              "Rows and ragged signals shard across GPUs with no exchange; only the one-GPU "
              "figure is measured here.", file=sys.stderr)
        sys.exit(2)
    if args.warmup is None:
        # ~25 ms of untimed work per workload: the shader clock dips and
        # recovers over the first ~20 ms of a sustained load.
        args.warmup = {"direct_k33": 60, "direct_k129": 30, "copy_1gib": 60,
                       "ragged": 2}.get(args.workload, 20)
    import torch
    from scir_amd import _lib
    torch.cuda.set_device(0)
    _lib.check(_lib.lib().scir_hip_init())
    if args.workload == "ragged":
        out = run_ragged(args)
    elif args.workload == "vjp_k255":
        out = run_vjp(args)
    elif args.workload == "copy_1gib":
        out = run_copy(args)
    elif args.workload == "resample_2_3":
        out = run_resample(args)
    elif args.workload == "sosfilt_butter6":
        out = run_sosfilt(args)
    else:
        out = run_dense(args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
