#!/usr/bin/env python3
"""bench.py -- GStencils/s of the tiled FFT-stencil path on one MI355X, over the five workloads.

One bench "step" is one full pass of the hot path: T stencil steps over the whole synthetic grid
(ffs_run: floor(T/t) fused blocks + tail, replayed from a hipGraph), input already resident in HBM.

Workloads (seed 0x5EED0000 + index, SplitMix64; weights uniform[0,1) normalised to sum 1):
  C1 Heat-1D 1,048,576 x T=64           FP64 FFT form (+ direct FP64)
  C2 Box2D9P 8192^2 x T=128             FP64 FFT form (+ mixed forms, direct)
  C3 Box2D25P 8192^2 x T=256            mixed FP16 form vs FP64 (+ BF16, direct)  <- `value`
  C4 Box3D27P 512^3 x T=32              sinusoid field; mixed BF16 form vs FP64 (+ compensated FP32, direct)
  C5 96 2D grids, 512..4096 sides, 4 stencils, T 16..256: one persistent launch per round, fused depths
     and mapping from the batch autotuner (ffs_batch_autotune, timed outside the timed region; the
     planner's depths are timed beside it)

`value` is C3's mixed FP16 form (the metric's headline comparison); every workload's record sits under
`configs` with its forms' GStencils/s, the dominant kernel's roofline, max-rel / rel-L2 against the FP64
form (full T) and against the FP64 direct sweep (the oracle, on a bounded sample), and the oracle's CPU
time on the host's cores.

Roofline of the dominant kernel (one fused block = one launch), algorithmic counts only:
  flops  = the real GEMMs the kernel's formulation needs, each counted once (the Gauss 3-GEMM complex
           product, the inverse stages cut to the kept output block, never the split products of the
           mixed forms: executed MFMA flops / products per real GEMM, 1 FP64/FP32 (2 for the
           compensated FP32 form in 3D), 3 FP16, 6 BF16),
           against the matrix peak of the form's arithmetic (FP64 78.6 TF, FP32 157.3 TF, FP16/BF16
           2.5 PF); direct kernel: 2 |S| t per point against the FP64 vector peak;
  bytes  = 16 x points (FP64 state read once and written once per block) against 8 TB/s;
  frac   = the larger of the two fractions, `bound` names which.
Beside it: `matrix_frac_dense`, the dense count tiles x 8 d P^(d+1) per real tile (per-axis complex DFT
GEMMs with real-pair packing; 1D prime-factor window 240 = 15 x 16: 8 x 240 x (15 + 16)), and
`executed_mfma_tflops` (split products included).

Printed: one JSON line.  --gpus N > 1 exits: the library runs on one GPU.
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

MATRIX_PEAK = {"fp64": 78.6e12, "f32": 157.3e12, "bf16": 2.5e15, "f16": 2.5e15}
VECTOR_PEAK_FP64 = 78.6e12
HBM_PEAK = 8.0e12

WORKLOADS = {
    "heat1d": dict(cfg="C1", shape="Heat-1D", sizes=(1048576,), T=64, seed=0x5EED0000, forms=("fp64", "direct")),
    "box2d9p": dict(cfg="C2", shape="Box2D9P", sizes=(8192, 8192), T=128, seed=0x5EED0001,
                    forms=("fp64", "f16", "bf16", "direct")),
    "box2d25p": dict(cfg="C3", shape="Box2D25P", sizes=(8192, 8192), T=256, seed=0x5EED0002,
                     forms=("f16", "bf16", "fp64", "direct")),
    "box3d27p": dict(cfg="C4", shape="Box3D27P", sizes=(512, 512, 512), T=32, seed=0x5EED0003,
                     forms=("bf16", "fp64", "f32", "direct"), init="sinusoids"),
    "batch96": dict(cfg="C5", seed=0x5EED0004, forms=("f16", "bf16", "fp64")),
}
ORDER = ("heat1d", "box2d9p", "box2d25p", "box3d27p", "batch96")
DTYPES = {"fp64": "f64", "bf16": "bf16x3-split/f32-acc (f64 state)",
          "f32": "f32 (f64 state; 3D: compensated matrix, 2 products)",
          "f16": "f16x2-split/f32-acc (f64 state)", "direct": "f64 (direct sum)"}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


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


def form_args(form):
    """Plan arguments of a bench form: a precision of the FFT path, or "direct" (FP64 stencil sum)."""
    return dict(precision="fp64", algorithm="direct") if form == "direct" else dict(precision=form)


def tile_flops(info):
    """Algorithmic flops of one real tile: 8 d P^(d+1) for the dense per-axis complex DFT GEMMs with
    real-pair packing; the 1D prime-factor window (240 = 15 x 16) counts 8 x 240 x (15 + 16)."""
    d = info["dims"]
    if d == 1:
        return 8.0 * 240 * (15 + 16)
    P = info["tile"][0]
    return 8.0 * d * P ** (d + 1)


PRODUCTS = {"fp64": 1, "f32": 1, "f16": 3, "bf16": 6}  # MFMA products per real GEMM (split terms)


def products(form, info):
    """MFMA products per real GEMM; the 3D FP32 form adds the compensating low-order products."""
    return 2 if form == "f32" and info["dims"] == 3 else PRODUCTS[form]


def roofline(info, kernel_s, form, radius, taps):
    """Roofline object of one fused block (one launch) timed at kernel_s seconds."""
    pts = float(np.prod(info["sizes"][:info["dims"]]))
    bytes_alg = 16.0 * pts
    dense = None
    if form == "direct":
        flops = 2.0 * taps * pts * info["fuse"]
        peak, unit_peak = VECTOR_PEAK_FP64, "fp64 vector"
    else:
        flops = info["flops_per_block"] / products(form, info)
        dense = tile_flops(info) * info["tiles"]
        peak, unit_peak = MATRIX_PEAK[form], f"{form} matrix"
    f_frac = flops / kernel_s / peak
    b_frac = bytes_alg / kernel_s / HBM_PEAK
    out = {"executed_mfma_tflops": info["flops_per_block"] / kernel_s / 1e12 if form != "direct" else None,
           "flops_per_launch": flops, "dense_flops_per_launch": dense, "bytes_per_launch": bytes_alg,
           "kernel_ms": kernel_s * 1e3, "matrix_frac": f_frac,
           "matrix_frac_dense": dense / kernel_s / peak if dense else None, "hbm_frac": b_frac,
           "matrix_peak": unit_peak}
    if b_frac >= f_frac:
        out.update(bound="hbm", achieved=bytes_alg / kernel_s / 1e9, peak=HBM_PEAK / 1e9, unit="GB/s", frac=b_frac)
    else:
        out.update(bound="mfma" if form != "direct" else "valu", achieved=flops / kernel_s / 1e12,
                   peak=peak / 1e12, unit="TFLOP/s", frac=f_frac)
    return out


def traffic(workload, form):
    """HBM bytes per launch of the dominant kernel from the committed rocprofv3 --pmc summary, if any."""
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            return json.load(f).get(f"{workload}/{form}")
    except (OSError, ValueError):
        return None


def timed(torch, fn, warmup, steps, stream, min_ms=20.0):
    """ms per call of fn() over >= `steps` calls (more for very short calls, so the timed region lasts
    >= min_ms), with HIP events on `stream`, after `warmup` untimed calls."""
    for _ in range(max(warmup, 1)):
        fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    fn()
    e1.record(stream)
    torch.cuda.synchronize()
    one = max(e0.elapsed_time(e1), 1e-3)
    reps = max(steps, int(math.ceil(min_ms / one)))
    e0.record(stream)
    for _ in range(reps):
        fn()
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps, reps


def cpu_threads():
    return int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))


CPU_SWEEP = ("FP64 direct periodic sweep, register-blocked AVX2 FMAs (oracle/sweep_fast.c, equal to the "
             "oracle's strict sweep oracle/oracle.c to 1e-13)")


def oracle_sample(x, radius, w, T, seconds, label):
    """The oracle's fast sweep (FP64 direct periodic sweep, OpenMP on the host) on a bounded sample of the
    workload: the first ts of the T steps on the full grid, ts chosen so the sample lasts about `seconds`.
    Its result is also the reference for the vs_oracle errors."""
    from tests import _oracle

    threads = cpu_threads()

    def timed(k):
        t1 = time.perf_counter()
        _oracle.sweep_fast(x, radius, w, k, threads=threads)
        return time.perf_counter() - t1

    timed(1)  # first call: thread pool start-up
    # steps that fit the target (per step from an 8-step call, which also pays one allocation: an upper
    # bound); when all T steps fit, whole T-step runs are repeated until the target time has passed
    k = min(T, 8)
    per_step = timed(k) / k
    ts = int(max(1, min(T, math.floor(seconds / per_step))))
    reps = 0
    t1 = time.perf_counter()
    while True:
        ref = _oracle.sweep_fast(x, radius, w, ts, threads=threads)
        reps += 1
        if ts < T or time.perf_counter() - t1 >= seconds or reps >= 100000:
            break
    cpu_s = time.perf_counter() - t1
    cpu = {"value": x.size * ts * reps / cpu_s / 1e9, "unit": "GStencils/s", "cores": threads, "kind": "port",
           "sample": f"{label}, {ts} of the {T} steps" + (f" repeated {reps} times" if reps > 1 else "")
                     + f", {CPU_SWEEP}, OpenMP {threads} threads, {cpu_s:.1f} s"}
    return ts, ref, cpu


def fft_baseline(torch, x_dev, w, radius, steps=4):
    """The reference's competitor: one full-grid FFT convolution per step (benchmarks/cufft-by-pytorch/
    rfft2D.py), here on rocFFT through torch.fft, with the kernel spectrum computed once."""
    d = x_dev.dim()
    K = 2 * radius + 1
    ker = torch.zeros_like(x_dev)
    ker[(slice(0, K),) * d] = torch.from_numpy(np.asarray(w).reshape((K,) * d)).to(x_dev.device)
    dims = tuple(range(d))
    kh = torch.fft.rfftn(ker, dim=dims)
    y = torch.fft.irfftn(torch.fft.rfftn(x_dev, dim=dims) * kh, s=x_dev.shape, dim=dims)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        y = torch.fft.irfftn(torch.fft.rfftn(y, dim=dims) * kh, s=x_dev.shape, dim=dims)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    del ker, kh, y
    return {"value": x_dev.numel() / dt / 1e9, "unit": "GStencils/s", "ms_per_step": dt * 1e3,
            "what": "full-grid FP64 rocFFT convolution, one rfftn/irfftn pair per step (torch.fft)"}


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


def bench_grid(name, args, torch, ffs, forms=None):
    """One single-grid workload: every form timed (full runs and the dominant kernel alone), errors
    against the FP64 form and the oracle sample, the CPU baseline, the competitor baselines."""
    wl = WORKLOADS[name]
    forms = forms or wl["forms"]
    _, radius, w, tag = ffs.named_stencil(wl["shape"], seed=wl["seed"])
    taps = int(np.count_nonzero(w))
    sizes, T = wl["sizes"], wl["T"]
    npts = int(np.prod(sizes))
    label = f"{wl['shape']} {'x'.join(map(str, sizes))}"
    x = sinusoid_field(ffs, wl["seed"] + 1, sizes) if wl.get("init") == "sinusoids" else \
        ffs.fill_uniform(wl["seed"] + 1, npts).reshape(sizes)
    d_in = torch.from_numpy(x).cuda()
    stream = torch.cuda.current_stream()
    if args.copy_calibration:
        tmp = torch.empty_like(d_in)
        for _ in range(3):
            tmp.copy_(d_in)
        torch.cuda.synchronize()
        log(f"copy calibration: 3 copies of {d_in.numel() * 8} bytes")
        del tmp
    res, outputs = {}, {}
    for form in forms:
        try:
            plan = ffs.Plan(len(sizes), sizes, w, radius, T, **form_args(form))
        except ffs.FFSError as e:
            log(f"{name} {form} unavailable: {e}")
            continue
        info = plan.info
        d_out = torch.empty_like(d_in)
        t0 = time.perf_counter()
        ms, reps = timed(torch, lambda: plan.run(d_in, d_out), args.warmup, args.steps, stream)
        wall = time.perf_counter() - t0
        # dominant kernel alone: one fused block per launch on the launch stream, ping-ponging between two
        # grids as inside a run (each launch reads the previous block's output)
        d_a, d_b = d_in.clone(), torch.empty_like(d_in)
        flip = [0]

        def one_block():
            src, dst = (d_a, d_b) if flip[0] % 2 == 0 else (d_b, d_a)
            flip[0] += 1
            plan.run_block(src, dst)

        kms, _ = timed(torch, one_block, 1, args.kernel_reps, stream, min_ms=5.0)
        del d_a, d_b
        rl = roofline(info, kms * 1e-3, form, radius, taps)
        rl["traffic"] = traffic(name, form)
        if rl["traffic"] is not None:
            rl["traffic_from"] = "profiles/traffic.json (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, copy-calibrated)"
        gst = npts * T / (ms * 1e-3) / 1e9
        res[form] = dict(value=gst, ms_per_step=ms, timed_runs=reps, wall_s=wall, roofline=rl,
                         plan={k: info[k] for k in ("tile", "valid", "fuse", "fuse_tail", "blocks", "algorithm",
                                                    "mapping")})
        outputs[form] = d_out
        log(f"{name} {form}: {gst:.1f} GStencil/s, {ms:.3f} ms/run, block {kms:.3f} ms, roofline {rl['bound']} "
            f"{rl['frac']:.1%} (matrix {rl['matrix_frac']:.1%}, hbm {rl['hbm_frac']:.1%}), P={info['tile']} "
            f"t={info['fuse']} V={info['valid']} blocks={info['blocks']}")
        plan.close()
    if "fp64" in outputs:
        ref64 = outputs["fp64"]
        for form, d in outputs.items():
            if form == "fp64":
                continue
            diff = (d - ref64).abs()
            res[form]["vs_fp64_full_T"] = {
                "max_rel": float(diff.max() / ref64.abs().max()),
                "rel_l2": float(torch.linalg.vector_norm(diff) / torch.linalg.vector_norm(ref64))}
            del diff
    del outputs
    cpu = None
    if not args.no_cpu:
        ts, ref, cpu = oracle_sample(x, radius, w, T, args.cpu_seconds, label)
        log(f"{name} cpu baseline: {cpu['value']:.3f} GStencil/s on {cpu['cores']} threads ({cpu['sample']})")
        from tests import _oracle
        for form in res:
            plan = ffs.Plan(len(sizes), sizes, w, radius, ts, **form_args(form))
            d = torch.empty_like(d_in)
            plan.run(d_in, d)
            torch.cuda.synchronize()
            plan.close()
            _, rel, l2 = _oracle.errors(d.cpu().numpy(), ref)
            res[form]["vs_oracle"] = {"steps": ts, "max_rel": rel, "rel_l2": l2}
            log(f"{name} {form} vs oracle after {ts} steps: max-rel {rel:.3e} rel-L2 {l2:.3e}")
            del d
        del ref
    base = {}
    if not args.no_fft_baseline and len(sizes) > 1:
        for key, fn in (("fft_baseline", fft_baseline), ("conv_baseline", conv_baseline)):
            try:
                base[key] = fn(torch, d_in, w, radius)
                log(f"{name} {key}: {base[key]['value']:.2f} GStencil/s")
            except RuntimeError as e:  # out of memory on the largest grids
                log(f"{name} {key} skipped: {e}")
    del d_in
    torch.cuda.empty_cache()
    mixed = [f for f in res if f not in ("fp64", "direct")]
    rec = {"config": wl["cfg"], "workload": f"{label} T={T}", "stencil": wl["shape"], "grid": list(sizes), "T": T,
           "forms": res, "cpu_baseline": cpu, **base}
    if "fp64" in res:
        rec["fp64_value"] = res["fp64"]["value"]
        rec["fp64_roofline_frac"] = res["fp64"]["roofline"]["frac"]
        if mixed:
            rec["mixed_vs_fp64"] = {f: res[f]["value"] / res["fp64"]["value"] for f in mixed}
    return rec


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


def bench_batch(args, torch, ffs, forms=None):
    """C5: the 96-grid batch in each batched form; whole-run roofline (all rounds' launches), per-grid
    error against the FP64 form, oracle sample on the first grids."""
    wl = WORKLOADS["batch96"]
    forms = forms or wl["forms"]
    specs = batch96_specs(ffs, wl["seed"])
    pts = sum(s["sizes"][0] * s["sizes"][1] * s["steps"] for s in specs)
    d_in = [torch.from_numpy(ffs.fill_uniform(s["seed"], s["sizes"][0] * s["sizes"][1]).reshape(s["sizes"])).cuda()
            for s in specs]
    stream = torch.cuda.current_stream()
    res, outs = {}, {}
    for form in forms:
        d_out = [torch.empty_like(d) for d in d_in]
        # the planner's depths, then the batch autotuner's (whole-batch timings; cached per workload)
        batch = ffs.Batch(specs, precision=form)
        planned_ms, _ = timed(torch, lambda: batch.run(d_in, d_out), args.warmup, args.steps, stream)
        planned_fuse = [batch.info(i)["fuse"] for i in range(len(specs))]
        batch.close()
        t0 = time.perf_counter()
        cache = os.path.join(args.tune_dir, f"batch96_{form}.jsonl") if args.tune_dir else None
        tuned_specs, _ = ffs.Batch.autotune(specs, precision=form, cache=cache)
        tune_s = time.perf_counter() - t0
        batch = ffs.Batch(tuned_specs, precision=form)
        ms, reps = timed(torch, lambda: batch.run(d_in, d_out), args.warmup, args.steps, stream)
        gst = pts / (ms * 1e-3) / 1e9
        flops = bytes_alg = 0.0
        rounds = 0
        for i in range(len(specs)):
            inf = batch.info(i)
            rounds = max(rounds, inf["blocks"])
            flops += inf["flops_per_block"] / products(form, inf) * inf["blocks"]
            bytes_alg += 16.0 * inf["sizes"][0] * inf["sizes"][1] * inf["blocks"]
        s = ms * 1e-3
        f_frac, b_frac = flops / s / MATRIX_PEAK[form], bytes_alg / s / HBM_PEAK
        rl = {"scope": f"whole run ({rounds} launches, one per round)", "flops_per_run": flops,
              "bytes_per_run": bytes_alg, "matrix_frac": f_frac, "hbm_frac": b_frac, "traffic": None}
        if b_frac >= f_frac:
            rl.update(bound="hbm", achieved=bytes_alg / s / 1e9, peak=HBM_PEAK / 1e9, unit="GB/s", frac=b_frac)
        else:
            rl.update(bound="mfma", achieved=flops / s / 1e12, peak=MATRIX_PEAK[form] / 1e12, unit="TFLOP/s",
                      frac=f_frac)
        changed = sum(1 for i in range(len(specs)) if batch.info(i)["fuse"] != planned_fuse[i])
        res[form] = dict(value=gst, ms_per_step=ms, timed_runs=reps, roofline=rl, rounds=rounds,
                         autotune={"tune_s": tune_s, "planned_ms": planned_ms,
                                   "planned_value": pts / (planned_ms * 1e-3) / 1e9, "grids_retuned": changed,
                                   "mapping": tuned_specs[0]["mapping"]})
        outs[form] = d_out
        batch.close()
        log(f"batch96 {form}: {gst:.1f} GStencil/s, {ms:.2f} ms/run, {rounds} rounds, roofline {rl['bound']} "
            f"{rl['frac']:.1%}; planner depths {planned_ms:.2f} ms/run, autotune {tune_s:.1f} s changed {changed} grids")
    if "fp64" in outs:  # per-grid error of the mixed forms against the batched FP64 form
        for form in outs:
            if form == "fp64":
                continue
            worst = worst_l2 = 0.0
            for i in range(len(specs)):
                ref, got = outs["fp64"][i], outs[form][i]
                diff = (got - ref).abs()
                worst = max(worst, float(diff.max() / ref.abs().max()))
                worst_l2 = max(worst_l2, float(torch.linalg.vector_norm(diff) / torch.linalg.vector_norm(ref)))
            res[form]["vs_fp64_full_T"] = {"worst_grid_max_rel": worst, "worst_grid_rel_l2": worst_l2}
            log(f"batch96 {form}: worst per-grid max-rel vs FP64 {worst:.3e}")
    cpu = None
    if not args.no_cpu:  # oracle on the first grids of the batch, their full T, until ~cpu_seconds
        from tests import _oracle
        threads = cpu_threads()
        done, t_cpu, ptsc, worst = 0, 0.0, 0, {f: 0.0 for f in outs}
        while done < len(specs) and t_cpu < args.cpu_seconds:
            s = specs[done]
            x = d_in[done].cpu().numpy()
            t1 = time.perf_counter()
            ref = _oracle.sweep_fast(x, s["radius"], s["weights"], s["steps"], threads=threads)
            t_cpu += time.perf_counter() - t1
            ptsc += x.size * s["steps"]
            for f in outs:
                worst[f] = max(worst[f], _oracle.errors(outs[f][done].cpu().numpy(), ref)[1])
            done += 1
        cpu = {"value": ptsc / t_cpu / 1e9, "unit": "GStencils/s", "cores": threads, "kind": "port",
               "sample": f"the first {done} of the 96 grids, all their steps, {CPU_SWEEP}, "
                         f"OpenMP {threads} threads, {t_cpu:.1f} s"}
        for f in outs:
            res[f]["vs_oracle"] = {"grids": done, "worst_grid_max_rel": worst[f]}
        log(f"batch96 cpu baseline {cpu['value']:.3f} GStencil/s; oracle on {done} grids: {worst}")
    del outs, d_in
    torch.cuda.empty_cache()
    rec = {"config": wl["cfg"], "workload": "batch of 96 2D grids, sides 512..4096, Heat-2D/Star2D9P/Box2D9P/"
                                            "Box2D25P, T 16..256", "grids": 96, "point_steps": pts,
           "forms": res, "cpu_baseline": cpu}
    if "fp64" in res:
        rec["fp64_value"] = res["fp64"]["value"]
        rec["mixed_vs_fp64"] = {f: res[f]["value"] / res["fp64"]["value"] for f in res if f != "fp64"}
    return rec


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5, help="timed full runs per form (more for runs under 20 ms)")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="all", choices=("all",) + ORDER)
    ap.add_argument("--kernel-reps", type=int, default=10, help="timed single-block launches for the roofline")
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="target length of each CPU-baseline sample")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baselines and the oracle sample checks")
    ap.add_argument("--forms", default=None, help="comma list overriding the workloads' forms")
    ap.add_argument("--no-fft-baseline", action="store_true", help="skip the rocFFT and library-convolution baselines")
    ap.add_argument("--tune-dir", default=None, help="directory for the batch autotuner's cache files (default: none, "
                                                     "tune every run)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run 3 plain device copies of the grid (known bytes) to calibrate HBM counters")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: the library runs on one GPU only (no multi-GPU path); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch

    import flashfftstencil_amd as ffs

    torch.cuda.set_device(0)
    forms = tuple(args.forms.split(",")) if args.forms else None
    names = ORDER if args.workload == "all" else (args.workload,)
    configs = {}
    for name in names:
        t0 = time.perf_counter()
        rec = bench_batch(args, torch, ffs, forms) if name == "batch96" else bench_grid(name, args, torch, ffs, forms)
        rec["bench_s"] = time.perf_counter() - t0
        configs[WORKLOADS[name]["cfg"]] = rec
    head = configs.get("C3") or next(iter(configs.values()))
    main_form = next(iter(head["forms"]))
    r = head["forms"][main_form]
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
        "dtype": DTYPES[main_form],
        "data": "synthetic (SplitMix64 uniform[0,1) grids and weights; C4: seeded sinusoid field)",
        "config": {"workload": head["workload"], "form": main_form, **({k: head[k] for k in ("stencil", "grid", "T")}
                                                                        if "grid" in head else {}),
                   **r.get("plan", {})},
        "roofline": r["roofline"],
        "cpu_baseline": head["cpu_baseline"],
        "errors": {k: r[k] for k in ("vs_fp64_full_T", "vs_oracle") if k in r},
        "configs": configs,
    }
    if "fp64" in head["forms"] and main_form != "fp64":
        line["speedup_vs_fp64"] = r["value"] / head["forms"]["fp64"]["value"]
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
