"""Benchmark of the batched-FIR hot path on one MI355X.

Prints ONE JSON line (rank 0):
  metric/value/unit   Gsamples/s of the headline workload (output samples per
                      second, inputs resident in HBM, device time of K steps)
  roofline            dominant kernel: algorithmic bytes or flops per launch /
                      its average launch time (HIP events on the launch stream)
  cpu_baseline        the oracle's restatement of the reference CPU loop
                      (single thread) on a bounded sample of the same workload

A step = one pass of the FIR over the whole synthetic batch.  Workloads:
  direct_k33   x[4096, 65536] N(0,1), 33-tap Hamming low-pass 0.2, FP32 direct
  direct_k129  same x, 129-tap low-pass, FP32 direct
(``--workload`` selects; default is the headline workload.)
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
FP32_PEAK_TFLOPS = 157.3
BF16_PEAK_TFLOPS = 2500.0

WORKLOADS = {
    "direct_k33": dict(B=4096, N=65536, K=33, taps="lowpass", precision="f32",
                       bound="hbm", seed=0x5C120001, outliers=0.0),
    "direct_k129": dict(B=4096, N=65536, K=129, taps="lowpass", precision="f32",
                        bound="hbm", seed=0x5C120001, outliers=0.0),
}
HEADLINE = "direct_k33"


def make_taps(kind, K):
    from scir_amd import design
    if kind == "lowpass":
        return design.lowpass(K, 0.2)
    if kind == "bandpass":
        return design.bandpass_kaiser(K, (0.1, 0.3), 8.0)
    raise ValueError(kind)


def load_traffic(kernel_tag):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary, if any."""
    path = os.path.join(ROOT, "profiles", f"{kernel_tag}_pmc.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    return d.get("hbm_bytes_per_launch")


def cpu_baseline(x_rows, b, budget_s=10.0):
    """Single-thread oracle (restatement of crates/scir-gpu/src/lib.rs:1134-1152)
    over whole rows of the workload until ~budget_s of CPU time is spent."""
    from oracle import oracle as orc
    taps = np.ascontiguousarray(b[::-1])
    done = 0
    t0 = time.perf_counter()
    while done < x_rows.shape[0]:
        orc.fir_ref_f32(x_rows[done:done + 1], taps)
        done += 1
        if time.perf_counter() - t0 >= budget_s:
            break
    dt = time.perf_counter() - t0
    n = x_rows.shape[1]
    return dict(value=done * n / dt / 1e9, unit="Gsamples/s", cores=1, kind="port",
                sample=f"{done} full rows x {n} samples of the same workload, 1 thread, "
                       f"{dt:.1f} s")


def run(args):
    import torch
    from scir_amd import _lib, device
    from oracle import oracle as orc

    wl = WORKLOADS[args.workload]
    B, N, K = wl["B"], wl["N"], wl["K"]
    torch.cuda.set_device(0)
    _lib.check(_lib.lib().scir_hip_init())
    stream = torch.cuda.current_stream()
    x = device.synth_normal(B, N, seed=wl["seed"], outlier_frac=wl["outliers"],
                            outlier_scale=1000.0)
    b_np = make_taps(wl["taps"], K)
    b = torch.tensor(b_np, device="cuda")
    y = torch.empty_like(x)
    prec = args.precision or wl["precision"]

    for _ in range(args.warmup):
        device.fir(x, b, prec, out=y)
    torch.cuda.synchronize()

    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    t_wall = time.perf_counter()
    for i in range(args.steps):
        starts[i].record(stream)
        device.fir(x, b, prec, out=y)
        ends[i].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_wall
    per_launch_ms = [s.elapsed_time(e) for s, e in zip(starts, ends)]
    total_ms = starts[0].elapsed_time(ends[-1])
    ms_per_step = total_ms / args.steps
    samples = B * N
    value = samples / (ms_per_step * 1e-3) / 1e9
    avg_launch_s = float(np.mean(per_launch_ms)) * 1e-3

    # accuracy on a row subset vs the FP64 oracle
    rows = [0, B // 3, B - 1]
    xs = x[rows].cpu().numpy()
    y64 = orc.lfilter_f64(xs, b_np)
    rel = orc.rel_l2(y[rows].cpu().numpy(), y64)

    if wl["bound"] == "hbm":
        alg = 8.0 * samples  # 4 B read + 4 B written per output sample
        achieved = alg / avg_launch_s / 1e9
        roof = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 4),
                    traffic=load_traffic(f"fir_direct_{args.workload}"))
    else:
        alg = 2.0 * K * samples
        achieved = alg / avg_launch_s / 1e12
        roof = dict(bound="mfma", achieved=round(achieved, 2), peak=BF16_PEAK_TFLOPS,
                    unit="TFLOP/s", frac=round(achieved / BF16_PEAK_TFLOPS, 4),
                    traffic=load_traffic(f"fir_mfma_{args.workload}"))

    out = dict(
        metric="FIR Gsamples/s & %roofline: bf16x3 MFMA vs FP32 direct at rel L2 err<=1e-6",
        value=round(value, 2), unit="Gsamples/s", n_gpus=1, steps=args.steps,
        warmup=args.warmup, ms_per_step=round(ms_per_step, 4), higher_is_better=True,
        scaling="weak", vs_baseline=None, dtype="f32", data="synthetic",
        config=dict(workload=args.workload, rows=B, samples_per_row=N, taps=K,
                    taps_design=wl["taps"], precision=prec),
        rel_l2_vs_fp64=rel, wall_s=round(wall, 3), roofline=roof,
    )
    if not args.no_cpu_baseline:
        nrows = min(B, 512)
        out["cpu_baseline"] = cpu_baseline(x[:nrows].cpu().numpy(), b_np,
                                           budget_s=args.cpu_budget)
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__,
                                 formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default=HEADLINE, choices=sorted(WORKLOADS))
    ap.add_argument("--precision", default=None)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=10.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (--gpus {args.gpus} requested); "
              "rows shard across GPUs with no exchange, but only the one-GPU figure is "
              "measured here.", file=sys.stderr)
        sys.exit(2)
    run(args)


if __name__ == "__main__":
    main()
