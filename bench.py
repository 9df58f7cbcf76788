#!/usr/bin/env python3
"""bench.py -- GStencils/s of the tiled FFT-stencil path on one MI355X.

One bench "step" is one full pass of the hot path: T stencil steps over the whole synthetic grid
(ffs_run: floor(T/t) fused blocks + tail, replayed from a hipGraph), input already resident in HBM.
Default workload: Box2D25P, 8192 x 8192, T = 256, weights uniform[0,1) normalised to sum 1, grid
uniform[0,1), both from SplitMix64 seeded 0x5EED0002.  `value` is the fastest mixed-precision form
(FP16 hi/lo split operands, FP32 accumulation, FP64 state and spectrum power); the BF16 3-term form,
the plain FP64-MFMA form and the direct FP64 kernel are timed beside it, and all are checked against
the FP64 direct sweep (the oracle) on a bounded sample.

Printed: one JSON line (see the keys below).  --gpus N > 1 exits: the library runs on one GPU.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAKS = {  # MI355X published dense peaks
    "fp64": ("mfma", 78.6e12, "TFLOP/s"),
    "bf16": ("mfma", 2.5e15, "TFLOP/s"),
    "f32": ("mfma", 157.3e12, "TFLOP/s"),
    "bf16x2": ("mfma", 2.5e15, "TFLOP/s"),
    "f16": ("mfma", 2.5e15, "TFLOP/s"),
    "direct": ("valu", 78.6e12, "TFLOP/s"),  # FP64 vector peak
}
HBM_PEAK = 8.0e12

WORKLOADS = {
    "box2d25p": dict(shape="Box2D25P", sizes=(8192, 8192), T=256, seed=0x5EED0002,
                     forms=("f16", "bf16", "fp64", "direct")),
    "box2d9p": dict(shape="Box2D9P", sizes=(8192, 8192), T=128, seed=0x5EED0001,
                    forms=("f16", "bf16", "fp64", "direct")),
    "heat1d": dict(shape="Heat-1D", sizes=(1048576,), T=64, seed=0x5EED0000, forms=("fp64", "direct")),
    "box3d27p": dict(shape="Box3D27P", sizes=(512, 512, 512), T=32, seed=0x5EED0003,
                     forms=("bf16", "f32", "fp64", "direct"), init="sinusoids"),
    "batch96": dict(seed=0x5EED0004, forms=("f16", "bf16", "fp64")),
}


def sinusoid_field(ffs, seed, sizes):
    """Sum of 8 sinusoids (amplitude uniform[0,1)/8, phase uniform[0, 2 pi), integer wavevectors uniform
    in [1, 16]^3) plus 1e-3 uniform noise; all draws from SplitMix64(seed)."""
    u = ffs.fill_uniform(seed, 8 * 5)
    n0, n1, n2 = sizes
    x = np.zeros(sizes)
    i0 = np.arange(n0)[:, None, None] / n0
    i1 = np.arange(n1)[None, :, None] / n1
    i2 = np.arange(n2)[None, None, :] / n2
    for m in range(8):
        amp, ph = u[5 * m] / 8, 2 * np.pi * u[5 * m + 1]
        k = [1 + int(16 * u[5 * m + 2 + a]) for a in range(3)]
        x += amp * np.cos(2 * np.pi * (k[0] * i0 + k[1] * i1 + k[2] * i2) + ph)
    x += 1e-3 * ffs.fill_uniform(seed + 1, x.size).reshape(sizes)
    return x


def fft_baseline(torch, x_dev, w, radius, steps=4):
    """The reference's competitor: one full-grid FFT convolution per step (benchmarks/cufft-by-pytorch/
    rfft2D.py), here on rocFFT through torch.fft, with the kernel spectrum computed once."""
    d = x_dev.dim()
    K = 2 * radius + 1
    ker = torch.zeros_like(x_dev)
    ker[(slice(0, K),) * d] = torch.from_numpy(np.asarray(w).reshape((K,) * d)).to(x_dev.device)
    dims = tuple(range(d))
    kh = torch.fft.rfftn(ker, dim=dims)
    y = x_dev
    y = torch.fft.irfftn(torch.fft.rfftn(y, dim=dims) * kh, s=x_dev.shape, dim=dims)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        y = torch.fft.irfftn(torch.fft.rfftn(y, dim=dims) * kh, s=x_dev.shape, dim=dims)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    del ker, kh, y
    return {"value": x_dev.numel() / dt / 1e9, "unit": "GStencils/s", "ms_per_step": dt * 1e3,
            "what": "full-grid FP64 rocFFT convolution, one rfftn/irfftn pair per step (torch.fft)"}


def form_args(form):
    """Plan arguments of a bench form: a precision of the FFT path, or "direct" (FP64 stencil sum)."""
    return dict(precision="fp64", algorithm="direct") if form == "direct" else dict(precision=form)


def conv_step(torch, x, w, radius):
    """One periodic stencil step as a library convolution: circular padding of 2r cells on the low side
    of every axis, then torch's FP64 cross-correlation with the flipped weights, which is the reference
    convention out[p] = sum_a w[a] in[p - a] (src/2D/check_correct.hpp:64-100)."""
    d = x.dim()
    K = 2 * radius + 1
    wt = torch.as_tensor(np.asarray(w, dtype=np.float64).reshape((K,) * d)[(slice(None, None, -1),) * d].copy(),
                         device=x.device).reshape((1, 1) + (K,) * d)
    conv = {1: torch.nn.functional.conv1d, 2: torch.nn.functional.conv2d, 3: torch.nn.functional.conv3d}[d]
    pad = (2 * radius, 0) * d
    return lambda y: conv(torch.nn.functional.pad(y[None, None], pad, mode="circular"), wt)[0, 0]


def conv_baseline(torch, x_dev, w, radius, steps=3):
    """The reference's other competitor (benchmarks/cudnn/cudnn-test.cpp: a library FP64 convolution per
    step), here torch's FP64 convolution on ROCm with circular padding, one call per step."""
    step = conv_step(torch, x_dev, w, radius)
    y = step(x_dev)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        y = step(y)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    del y
    return {"value": x_dev.numel() / dt / 1e9, "unit": "GStencils/s", "ms_per_step": dt * 1e3,
            "what": "FP64 torch.nn.functional.conv%dd with circular padding, one call per step" % x_dev.dim()}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5, help="timed full runs")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="box2d25p", choices=sorted(WORKLOADS))
    ap.add_argument("--kernel-reps", type=int, default=10, help="timed single-block launches for the roofline")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="target length of the CPU-baseline sample")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and the oracle sample check")
    ap.add_argument("--forms", default=None, help="comma list overriding the workload's precision forms")
    ap.add_argument("--no-fft-baseline", action="store_true", help="skip the rocFFT and library-convolution baselines")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run 3 plain device copies of the grid (known bytes) to calibrate HBM counters")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: the library runs on one GPU only (no multi-GPU path); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch

    import flashfftstencil_amd as ffs

    torch.cuda.set_device(0)
    if args.workload == "batch96":
        return bench_batch(args, torch, ffs)
    wl = WORKLOADS[args.workload]
    forms = tuple(args.forms.split(",")) if args.forms else wl["forms"]
    _, radius, w, tag = ffs.named_stencil(wl["shape"], seed=wl["seed"])
    sizes, T = wl["sizes"], wl["T"]
    npts = int(np.prod(sizes))
    if wl.get("init") == "sinusoids":
        x = sinusoid_field(ffs, wl["seed"] + 1, sizes)
    else:
        x = ffs.fill_uniform(wl["seed"] + 1, npts).reshape(sizes)
    d_in = torch.from_numpy(x).cuda()
    stream = torch.cuda.current_stream()

    if args.copy_calibration:
        tmp = torch.empty_like(d_in)
        for _ in range(3):
            tmp.copy_(d_in)
        torch.cuda.synchronize()
        log(f"copy calibration: 3 copies of {d_in.numel() * 8} bytes")
        del tmp

    results = {}
    outputs = {}
    for form in forms:
        try:
            plan = ffs.Plan(len(sizes), sizes, w, radius, T, **form_args(form))
        except ffs.FFSError as e:
            log(f"form {form} unavailable: {e}")
            continue
        info = plan.info
        d_out = torch.empty_like(d_in)
        for _ in range(args.warmup):
            plan.run(d_in, d_out)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        e0.record(stream)
        for _ in range(args.steps):
            plan.run(d_in, d_out)
        e1.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        ms = e0.elapsed_time(e1) / args.steps
        gst = npts * T / (ms * 1e-3) / 1e9
        # dominant kernel alone: one fused block per launch, timed with events on the launch stream; the
        # launches ping-pong between two grids as inside a run (each reads the previous block's output)
        d_a, d_b = d_in.clone(), torch.empty_like(d_in)
        plan.run_block(d_a, d_b)
        torch.cuda.synchronize()
        k0, k1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        k0.record(stream)
        for i in range(args.kernel_reps):
            if i % 2 == 0:
                plan.run_block(d_b, d_a)
            else:
                plan.run_block(d_a, d_b)
        k1.record(stream)
        torch.cuda.synchronize()
        kms = k0.elapsed_time(k1) / args.kernel_reps
        del d_a, d_b
        bound, peak, unit = PEAKS[form]
        achieved = info["flops_per_block"] / (kms * 1e-3)
        results[form] = dict(
            value=gst, ms_per_step=ms, wall_s=wall, kernel_ms=kms, info=info,
            roofline={"bound": bound, "achieved": achieved / 1e12, "peak": peak / 1e12, "unit": unit,
                      "frac": achieved / peak, "traffic": None,
                      "hbm_algorithmic_GBps": info["bytes_per_block"] / (kms * 1e-3) / 1e9},
        )
        outputs[form] = d_out
        log(f"{form}: {gst:.1f} GStencil/s, {ms:.2f} ms/run, block kernel {kms:.3f} ms, "
            f"{achieved / 1e12:.2f} TFLOP/s ({achieved / peak:.1%} of {form} {bound} peak), P={info['tile']} "
            f"t={info['fuse']} tail={info['fuse_tail']} V={info['valid']} blocks={info['blocks']}")
        plan.close()

    if not results:
        print("bench.py: no precision form could be planned", file=sys.stderr)
        sys.exit(1)

    errors = {}
    if "fp64" in outputs:
        ref64 = outputs["fp64"]
        for form, d in outputs.items():
            if form == "fp64":
                continue
            diff = (d - ref64).abs()
            errors[f"{form}_vs_fp64_full_T"] = {
                "max_rel": float(diff.max() / ref64.abs().max()),
                "rel_l2": float(torch.linalg.vector_norm(diff) / torch.linalg.vector_norm(ref64)),
            }

    cpu = None
    if not args.no_cpu:
        from tests import _oracle

        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
        t1 = time.perf_counter()
        _oracle.sweep(x, radius, w, 1, threads=threads)
        one = time.perf_counter() - t1
        ts = int(max(1, min(T, math.floor(args.cpu_seconds / max(one, 1e-6)))))
        # small workloads: repeat the whole T-step sweep until the sample lasts about --cpu-seconds
        reps = int(max(1, min(1000, math.floor(args.cpu_seconds / max(one * ts, 1e-6))))) if ts == T else 1
        t1 = time.perf_counter()
        for _ in range(reps):
            ref = _oracle.sweep(x, radius, w, ts, threads=threads)
        cpu_s = time.perf_counter() - t1
        cpu = {"value": npts * ts * reps / cpu_s / 1e9, "unit": "GStencils/s", "cores": threads, "kind": "port",
               "sample": f"{wl['shape']} {'x'.join(map(str, sizes))}, {ts} of the {T} steps"
                         + (f" repeated {reps} times" if reps > 1 else "")
                         + f", FP64 direct periodic sweep (oracle/oracle.c, OpenMP {threads} threads), {cpu_s:.1f} s"}
        log(f"cpu baseline: {cpu['value']:.3f} GStencil/s on {threads} threads ({ts} steps, {cpu_s:.1f} s)")
        for form in outputs:
            plan = ffs.Plan(len(sizes), sizes, w, radius, ts, **form_args(form))
            d = torch.empty_like(d_in)
            plan.run(d_in, d)
            torch.cuda.synchronize()
            got = d.cpu().numpy()
            plan.close()
            _, rel, l2 = _oracle.errors(got, ref)
            errors[f"{form}_vs_oracle_T{ts}"] = {"max_rel": rel, "rel_l2": l2}
            log(f"{form} vs oracle after {ts} steps: max-rel {rel:.3e} rel-L2 {l2:.3e}")

    fftb = None
    if len(sizes) > 1 and not args.no_fft_baseline:
        try:
            fftb = fft_baseline(torch, d_in, w, radius)
            log(f"rocFFT per-step full-grid convolution: {fftb['value']:.2f} GStencil/s")
        except RuntimeError as e:  # out of memory on the largest grids
            log(f"fft baseline skipped: {e}")

    convb = None
    if not args.no_fft_baseline:
        try:
            convb = conv_baseline(torch, d_in, w, radius)
            log(f"library convolution per step: {convb['value']:.2f} GStencil/s")
        except RuntimeError as e:  # out of memory on the largest grids
            log(f"convolution baseline skipped: {e}")

    main_form = forms[0] if forms[0] in results else next(iter(results))
    r = results[main_form]
    for form, res in results.items():
        res["roofline"]["traffic"] = _traffic(args.workload, form)
    line = {
        "metric": "GStencils/s (points x steps / s)",
        "value": r["value"],
        "unit": "GStencils/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "f64", "bf16": "bf16x3-split/f32-acc (f64 state)", "bf16x2": "bf16x2-split/f32-acc (f64 state)",
                  "f32": "f32 (f64 state)", "f16": "f16x2-split/f32-acc (f64 state)",
                  "direct": "f64 (direct sum)"}[main_form],
        "data": "synthetic (SplitMix64 uniform[0,1) grid and weights)",
        "config": {
            "workload": f"{wl['shape']} {'x'.join(map(str, sizes))} T={T}",
            "stencil": wl["shape"], "grid": list(sizes), "T": T, "form": main_form,
            "tile": r["info"]["tile"], "fuse": r["info"]["fuse"], "fuse_tail": r["info"]["fuse_tail"],
            "valid": r["info"]["valid"], "blocks": r["info"]["blocks"],
        },
        "roofline": r["roofline"],
        "cpu_baseline": cpu,
        "forms": {f: {"value": v["value"], "ms_per_step": v["ms_per_step"], "kernel_ms": v["kernel_ms"],
                      "roofline": v["roofline"], "fuse": v["info"]["fuse"]} for f, v in results.items()},
        "errors": errors,
        "fft_baseline": fftb,
        "conv_baseline": convb,
    }
    if "fp64" in results and main_form != "fp64":
        line["speedup_vs_fp64"] = r["value"] / results["fp64"]["value"]
    print(json.dumps(line), flush=True)


def splitmix_ints(seed, n):
    out, s, m = [], seed, (1 << 64) - 1
    for _ in range(n):
        s = (s + 0x9E3779B97F4A7C15) & m
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & m
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & m
        out.append(z ^ (z >> 31))
    return out


def batch96_specs(ffs, seed):
    """96 grids: sides from {512, 640, ..., 4096}, stencil from {Heat-2D, Star2D9P, Box2D9P, Box2D25P},
    T from {16, 32, ..., 256}, weights and data sub-seeded; all draws from SplitMix64(seed)."""
    shapes = ["Heat-2D", "Star2D9P", "Box2D9P", "Box2D25P"]
    r = splitmix_ints(seed, 96 * 4)
    specs = []
    for g in range(96):
        n0 = 512 + 128 * (r[4 * g] % 29)
        n1 = 512 + 128 * (r[4 * g + 1] % 29)
        shape = shapes[r[4 * g + 2] % 4]
        T = 16 * (1 + r[4 * g + 3] % 16)
        _, radius, w, _ = ffs.named_stencil(shape, seed=seed + 1000 + g)
        specs.append(dict(shape=shape, sizes=(n0, n1), steps=T, radius=radius, weights=w, seed=seed + 2000 + g))
    return specs


def bench_batch(args, torch, ffs):
    specs = batch96_specs(ffs, WORKLOADS["batch96"]["seed"])
    form = (args.forms or WORKLOADS["batch96"]["forms"][0]).split(",")[0]
    pts = sum(s["sizes"][0] * s["sizes"][1] * s["steps"] for s in specs)
    d_in = [torch.from_numpy(ffs.fill_uniform(s["seed"], s["sizes"][0] * s["sizes"][1]).reshape(s["sizes"])).cuda()
            for s in specs]
    d_out = [torch.empty_like(d) for d in d_in]
    batch = ffs.Batch(specs, precision=form)
    stream = torch.cuda.current_stream()
    for _ in range(args.warmup):
        batch.run(d_in, d_out)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(args.steps):
        batch.run(d_in, d_out)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    gst = pts / (ms * 1e-3) / 1e9
    log(f"batch96 {form}: {gst:.1f} GStencil/s, {ms:.2f} ms/run, {pts:.3e} point-steps")
    # per-grid error against the FP64 form of the same grid
    worst = 0.0
    for i, s in enumerate(specs):
        plan = ffs.Plan(2, s["sizes"], s["weights"], s["radius"], s["steps"], precision="fp64")
        ref = torch.empty_like(d_in[i])
        plan.run(d_in[i], ref)
        torch.cuda.synchronize()
        plan.close()
        worst = max(worst, float((d_out[i] - ref).abs().max() / ref.abs().max()))
        del ref
    log(f"batch96: worst per-grid max-rel vs FP64 {worst:.3e}")
    info0 = batch.info(0)
    batch.close()
    line = {
        "metric": "GStencils/s (points x steps / s)", "value": gst, "unit": "GStencils/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": {"bf16": "bf16x3-split/f32-acc (f64 state)", "bf16x2": "bf16x2-split/f32-acc (f64 state)",
                  "f16": "f16x2-split/f32-acc (f64 state)", "fp64": "f64"}[form],
        "data": "synthetic (SplitMix64 uniform[0,1) grids and weights)",
        "config": {"workload": "batch of 96 2D grids, sides 512..4096, Heat-2D/Star2D9P/Box2D9P/Box2D25P, T 16..256",
                   "grids": 96, "point_steps": pts, "form": form, "tile": info0["tile"]},
        "roofline": None, "cpu_baseline": None,
        "errors": {"worst_grid_max_rel_vs_fp64": worst},
    }
    print(json.dumps(line), flush=True)


def _traffic(workload, form):
    """HBM bytes per launch of the dominant kernel from the committed rocprofv3 --pmc summary, if any."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            data = json.load(f)
        return data.get(f"{workload}/{form}")
    except (OSError, ValueError):
        return None


if __name__ == "__main__":
    main()
