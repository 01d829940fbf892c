#!/usr/bin/env python3
"""Benchmark of the batched Toeplitz SENSE normal operator on one MI355X.

Workload (a 4D-flow scan): 5 velocity encodings x 20 cardiac frames = 100 problems,
image 128^3 complex64, 16 coils, oversampled grid 256^3; kooshball readouts of 128
samples with per-frame spoke counts drawn from U{800..1500} in a seed-fixed golden-means
order; one Toeplitz kernel per frame (shared by its 5 encodings). A step is one batched
normal-operator apply over all 100 problems with inputs resident in HBM; the value is
problems * steps / seconds (normal-op applies/s).

Also reported:
- per-kernel durations from HIP events recorded on the launch stream over the timed
  region, and the roofline of the dominant kernel (algorithmic bytes per launch over
  its average duration);
- CG-SENSE recons/s (15 iterations on the l2-regularised system A + 0.1 lambda_max I)
  for the batched form (all 100 problems in one device-CG batch, eager and as a replayed
  HIP graph of one iteration) and for the plain per-frame drop-in (the explicit NUFFT
  SenseNormal per problem inside a host-synchronised CG, as the reference runs it),
  the plain path's p50/p99 latency, and the images of problem 0 against the FP64
  oracle's CG;
- the CPU baseline: the reference's complex64 ATen sequence for the Toeplitz SENSE
  normal operator on torch-CPU with every host thread.

python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic.json")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--coils", type=int, default=16)
    ap.add_argument("--frames", type=int, default=20)
    ap.add_argument("--encodings", type=int, default=5)
    ap.add_argument("--readout", type=int, default=128)
    ap.add_argument("--spokes-min", type=int, default=800)
    ap.add_argument("--spokes-max", type=int, default=1500)
    ap.add_argument("--coil-batch", type=int, default=None, help="coils per workspace (default: autotuned)")
    ap.add_argument("--problem-batch", type=int, default=None, help="problems per launch (default: autotuned)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="skip the autotuner (defaults: 16 coils and 4 problems per launch, library options as set)")
    ap.add_argument("--generic", action="store_true", help="use the generic five-pass path")
    ap.add_argument("--xy-variant", type=int, default=None,
                    help="plane kernels of the fused path (library option sense_xy_variant)")
    ap.add_argument("--zconv-variant", type=int, default=None,
                    help="z-convolution kernel at N0 = 128 (library option sense_zconv_variant, 0..3)")
    ap.add_argument("--zconv-batch", type=int, default=None,
                    help="problems per z-convolution launch at N0 = 128 (library option sense_zconv_batch, 1..16)")
    ap.add_argument("--cg-iters", type=int, default=15)
    ap.add_argument("--no-cg", action="store_true")
    ap.add_argument("--plain", type=int, default=3, help="problems solved by the plain per-frame path (0: skip)")
    ap.add_argument("--plain-tol", type=float, default=1e-6, help="NUFFT tolerance of the plain per-frame path")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="CPU baseline: repeat whole applies until this many seconds have passed")
    ap.add_argument("--cg-l2", type=float, default=0.1, help="l2 regularisation of the CG system, x lambda_max")
    ap.add_argument("--power-iters", type=int, default=20)
    ap.add_argument("--no-oracle-cg", action="store_true", help="skip the FP64 oracle CG of problem 0")
    ap.add_argument("--no-kernel-events", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device copy (the known-size kernel for counter calibration)")
    ap.add_argument("--mode", choices=["normal", "llr", "serve", "nufft", "llr-recon"], default="normal",
                    help="normal: the batched normal operator (default); llr: the LLR prox on 128^3 x 100 frames; "
                         "serve: the Poisson request stream through the continuous-batching server; "
                         "nufft: type-2 and type-1 NUFFT at 128^3 with one frame's trajectory (1150 x 128); "
                         "llr-recon: 20 FISTA iterations with the LLR prox at 128^3 x 100 frames x 16 coils")
    ap.add_argument("--llr-iters", type=int, default=20, help="llr-recon: shift sequence length")
    ap.add_argument("--nufft-locality", type=int, default=None,
                    help="nufft mode: point ordering of the plans (library option nufft_locality, 0..3)")
    ap.add_argument("--llr-frames", type=int, default=100)
    ap.add_argument("--llr-max-sweeps", type=int, default=None, help="cap on the LLR Jacobi sweeps (1..15)")
    ap.add_argument("--requests", type=int, default=200)
    ap.add_argument("--rate", type=float, default=2.0)
    ap.add_argument("--llr-fraction", type=float, default=0.25,
                    help="serve mode: share of multi-frame requests with the LLR prox")
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--set-option", action="append", default=[], metavar="NAME=VALUE",
                    help="library option (hasty_set_option) for the whole run; repeatable")
    return ap.parse_args()


MFMA_FP32_PEAK_TFLOPS = 157.3


def nufft_main(args, torch, native):
    """Type-2 and type-1 NUFFT at 128^3 (fine grid 256^3, tol 1e-6, ns = 7) on one
    4D-flow frame's kooshball (1150 spokes x 128 samples, M = 147 200). A step is one
    transform of each type; the value is transforms per second. Roofline: HBM, with the
    algorithmic bytes B_U = 20 M + 8 N^3 + 32 (2N)^3 per transform (coordinates and
    samples, the image, four fine-grid volumes)."""
    import numpy as np

    from hastycompute_amd import synthetic

    dev = torch.device("cuda:0")
    if args.nufft_locality is not None:
        from hastycompute_amd import _capi
        _capi.set_option("nufft_locality", args.nufft_locality)
    n, spokes, samp = args.n, 1150, 128
    co = synthetic.kooshball_torch(spokes, samp, 0, dev).contiguous()
    M = co.shape[1]
    g = torch.Generator(device=dev)
    g.manual_seed(args.seed)
    f = torch.randn((1, n, n, n), dtype=torch.complex64, device=dev, generator=g)
    c = torch.randn((1, M), dtype=torch.complex64, device=dev, generator=g)
    scale = 1.0 / math.sqrt(n ** 3)
    p2 = native.NufftPlan(co, [n, n, n], 2, -1, 1, 1e-6, scale)
    p1 = native.NufftPlan(co, [n, n, n], 1, +1, 1, 1e-6, scale)
    out2 = torch.empty((1, M), dtype=torch.complex64, device=dev)
    out1 = torch.empty((1, n, n, n), dtype=torch.complex64, device=dev)
    tune = None
    if not args.no_autotune:
        from hastycompute_amd import autotune

        tune = autotune.nufft(p1, c, out1)
    for _ in range(args.warmup):
        p2.execute(out2, f)
        p1.execute(c, out1)
    B = 20.0 * M + 8.0 * n ** 3 + 32.0 * (2 * n) ** 3
    res = {}
    for name, fn in (("type2", lambda: p2.execute(out2, f)), ("type1", lambda: p1.execute(c, out1))):
        torch.cuda.synchronize()
        native.timing_enable(True)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            fn()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / args.steps
        kt = native.timing_read()
        native.timing_enable(False)
        res[name] = {"ms": dt * 1e3, "achieved_GBs": B / dt / 1e9,
                     "kernels_ms": {k: v[0] / v[1] for k, v in kt.items()}}
    p1.close()
    p2.close()
    cpu = None
    if not args.no_cpu:
        # CPU: the oracle's ES-kernel NUFFT (C spreading/interpolation with OpenMP +
        # scipy.fft, FP64) at the same tolerance on the whole workload, one pair
        from oracle import oracle as O

        co_h = co.double().cpu().numpy()
        f_h = f[0].cpu().numpy().astype(np.complex128)
        c_h = c[0].cpu().numpy().astype(np.complex128)
        pairs = 0
        t = time.perf_counter()
        while pairs == 0 or time.perf_counter() - t < args.cpu_seconds:  # a bounded sample of whole pairs
            O.nufft_type2(co_h, f_h, -1, 1e-6)
            O.nufft_type1(co_h, c_h, [n, n, n], 1, 1e-6)
            pairs += 1
        cdt = time.perf_counter() - t
        cpu = {"value": pairs / cdt, "unit": "pairs/s", "cores": O.default_threads(), "kind": "port",
               "sample": f"{pairs} type-2 + type-1 pairs at {n}^3, M={M}, tol 1e-6: the FP64 oracle ES-kernel "
                         f"NUFFT (C spread/interp with OpenMP + scipy.fft, {O.default_threads()} threads)",
               "seconds": cdt}
    tr_s = 1e3 / sum(r["ms"] for r in res.values())
    dom = max(res, key=lambda k: res[k]["ms"])
    line = {"metric": f"NUFFT type-2 + type-1 pairs/s @{n}^3, M={M} (kooshball {spokes}x{samp}), tol 1e-6",
            "value": tr_s, "unit": "pairs/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": 1e3 / tr_s, "higher_is_better": True, "dtype": "c64 (fp32)",
            "data": "synthetic (seeded kooshball, random image and samples)",
            "config": {"workload": f"NUFFT {n}^3, fine {2 * n}^3, ns 7, M={M}", "M": M, "autotune": tune},
            "roofline": {"bound": "hbm", "kernel": dom, "achieved": res[dom]["achieved_GBs"], "peak": 8000.0,
                         "unit": "GB/s", "frac": res[dom]["achieved_GBs"] / 8000.0, "traffic": None,
                         "alg_bytes_per_transform": B},
            "transforms": res}
    if cpu is not None:
        line["cpu_baseline"] = cpu
    print(json.dumps(line), flush=True)


def llr_main(args, torch, native):
    """LLR prox on [F, 1, 128, 128, 128] (rank-5 temporal dynamics + noise): 16^3 blocks
    on the stride-16 grid (7^3 = 343 blocks, the last row skipped as in the reference),
    soft threshold lambda = 0.01 sigma_max of the first block. A step is one prox over the
    volume; the value is blocks per second. Roofline: the Gram and A W kernels on MFMA,
    8 M K^2 flops each per block (M = 4096 rows, K = F columns)."""
    import numpy as np

    dev = torch.device("cuda:0")
    F, n, b = args.llr_frames, args.n, 16
    if args.llr_max_sweeps is not None:
        from hastycompute_amd import _capi
        _capi.call("hasty_set_option", b"llr_max_sweeps", args.llr_max_sweeps)
    g = torch.Generator(device=dev)
    g.manual_seed(args.seed)
    U = torch.randn((5, n ** 3), dtype=torch.complex64, device=dev, generator=g)
    a = torch.randn((F, 5), dtype=torch.complex64, device=dev, generator=g)
    x = (a @ U).reshape(F, 1, n, n, n)
    x += 0.01 * torch.randn(x.shape, dtype=torch.complex64, device=dev, generator=g)
    del U
    blk = x[:, 0, :b, :b, :b].reshape(F, -1).T.cpu().numpy().astype(np.complex128)
    lam = 0.01 * float(np.linalg.svd(blk, compute_uv=False)[0])
    nblk = (n // b - 1) ** 3
    tune = None
    if not args.no_autotune:
        from hastycompute_amd import autotune

        tune = autotune.llr(list(x.shape), (b, b, b), (b, b, b))
    for _ in range(args.warmup):
        native.llr_normal_blocks(x, (b, b, b), (b, b, b), 1, lam, True)
    torch.cuda.synchronize()
    native.timing_enable(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        native.llr_normal_blocks(x, (b, b, b), (b, b, b), 1, lam, True)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    kt = native.timing_read()
    native.timing_enable(False)
    M = b ** 3
    flops = {"llr_gram": 8.0 * M * F * F * nblk, "llr_apply": 8.0 * M * F * F * nblk}
    kernels = {}
    for name, (ms, cnt) in kt.items():
        k = {"launches": cnt, "avg_ms": ms / cnt}
        if name in flops:
            k["achieved_TFLOPs"] = flops[name] / (ms / cnt / 1e3) / 1e12
            k["mfma_frac"] = k["achieved_TFLOPs"] / MFMA_FP32_PEAK_TFLOPS
        kernels[name] = k
    dom = max((k for k in kernels if k in flops), key=lambda k: kernels[k]["avg_ms"])
    line = {"metric": "LLR prox blocks/s @128^3x100 frames (16^3 blocks, Casorati 4096x100)",
            "value": nblk * args.steps / dt, "unit": "blocks/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": dt / args.steps * 1e3, "higher_is_better": True,
            "dtype": "c64 (fp32, MFMA fp32 accumulate)", "data": "synthetic rank-5 temporal phantom + noise",
            "config": {"workload": f"LLR prox, [{F},1,{n},{n},{n}], 16^3 blocks, stride 16, lambda=0.01 sigma_max",
                       "blocks": nblk, "lambda": lam, "autotune": tune},
            "roofline": {"bound": "mfma", "kernel": dom, "achieved": kernels[dom]["achieved_TFLOPs"],
                         "peak": MFMA_FP32_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": kernels[dom]["mfma_frac"],
                         "traffic": None, "flops_per_launch": flops[dom]},
            "kernels": kernels}
    if not args.no_cpu:
        line["cpu_baseline"] = llr_cpu_baseline(args, torch, x, lam, b)
    print(json.dumps(line), flush=True)


def llr_cpu_baseline(args, torch, x, lam, b, nblocks=8):
    """The reference's per-block SVT on host cores: svt_soft (lib/op/svt.cpp:117-144)
    as torch.linalg.svd of the block's Casorati matrix (complex64, 4096 x F), the soft
    threshold max(S - lam, 0) and U diag(S) Vh, on torch-CPU with every host thread;
    `nblocks` blocks of the benchmark volume (gather/scatter not timed)."""
    threads = cpu_threads()
    prev = torch.get_num_threads()
    torch.set_num_threads(threads)
    F = x.shape[0]
    mats = []
    for i in range(nblocks):
        z0 = b * i
        mats.append(x[:, 0, :b, :b, z0:z0 + b].reshape(F, -1).T.contiguous().cpu())
    done = 0
    t = time.perf_counter()
    with torch.inference_mode():
        # passes over the sample blocks until the bounded sample time is spent; every pass
        # thresholds the original blocks (a copy each, ~1.6 MB against a 4096 x F SVD)
        while done == 0 or time.perf_counter() - t < args.cpu_seconds:
            for A0 in mats:
                A = A0.clone()
                U, S, Vh = torch.linalg.svd(A, full_matrices=False)
                S = torch.clamp(S - lam, min=0.0)
                A.copy_((U * S.to(A.dtype)) @ Vh)
            done += len(mats)
    dt = time.perf_counter() - t
    torch.set_num_threads(prev)
    return {"value": done / dt, "unit": "blocks/s", "cores": threads, "kind": "port",
            "sample": f"{done} block SVTs ({nblocks} blocks of {b ** 3} x {F}, repeated): torch.linalg.svd + "
                      f"soft threshold + U diag(S) Vh (svt.cpp:117-144), complex64, torch-CPU ({threads} threads)",
            "seconds": dt}


def llr_recon_main(args, torch, native):
    """LLR-regularised 4D-flow reconstruction at the full size: [F, 1, 128, 128, 128]
    (F = 100 frames, one encoding), 16 coils, one kooshball trajectory per frame
    (800-1500 spokes x 128), 20 FISTA prox-gradient iterations (GradientDescent with
    accelerate, python/python/hastypy/base/solvers.py:85-162, inside FivePointLLR,
    recon.py:87-205): gradient T x - b through the batched Toeplitz normal operator (the
    subtraction fused into its last kernel), the step, then the LLR prox (16^3 blocks,
    the reference's shifted block grid, soft threshold lambda = 0.01 sigma_max). The
    data: a seeded rank-5 temporal phantom plus noise, b = T x_noisy per frame. A step is
    one prox-gradient iteration over all frames."""
    import numpy as np

    from hastycompute_amd import synthetic

    dev = torch.device("cuda:0")
    F, n, C, b = args.llr_frames, args.n, args.coils, 16
    rng = np.random.default_rng(args.seed)
    spokes = rng.integers(args.spokes_min, args.spokes_max + 1, size=F)
    offsets = np.concatenate([[0], np.cumsum(spokes)[:-1]])
    kern = []
    for f in range(F):
        co = synthetic.kooshball_torch(int(spokes[f]), args.readout, int(offsets[f]), dev)
        _, k = native.toeplitz_build(co, [n] * 3, tol=1e-8, diag=False)
        kern.append(k)
        del co
    K = torch.stack(kern)
    del kern
    smaps = synthetic.coil_maps_torch(C, n, dev, seed=args.seed)
    g = torch.Generator(device=dev)
    g.manual_seed(args.seed)
    ph = synthetic.ellipsoid_phantom_torch(n, dev, seed=args.seed).reshape(1, -1)
    comp = torch.cat([ph] + [ph * torch.randn((1, n ** 3), dtype=torch.complex64, device=dev, generator=g) * 0.2
                             for _ in range(4)])  # [5, N^3] spatial components
    temp = torch.randn((F, 5), dtype=torch.complex64, device=dev, generator=g) / math.sqrt(5)
    x_true = (temp @ comp).reshape(F, 1, n, n, n)
    del comp
    x_noisy = x_true + 0.01 * torch.randn(x_true.shape, dtype=torch.complex64, device=dev, generator=g)
    tune = None
    if not args.no_autotune:
        from hastycompute_amd import autotune

        tune = autotune.sense_normal([n] * 3, C, F)
        cb, pb = tune["coil_batch"], tune["problem_batch"]
    else:
        cb, pb = 16, 4
    op = native.SenseNormalToeplitz(smaps, K, coil_batch=cb, problem_batch=pb)
    kidx = list(range(F))
    rhs = torch.empty_like(x_true)
    op.apply(x_noisy.reshape(F, n, n, n), out=rhs.reshape(F, n, n, n), kernel_index=kidx)
    del x_noisy
    # step 1 / lambda_max of frame 0 (power iteration), threshold 0.01 sigma_max of a block
    v = rhs[0:1, 0].clone()
    w = torch.empty_like(v)
    for _ in range(args.power_iters):
        op.apply(v, out=w, kernel_index=[0])
        lmax = float(torch.linalg.vector_norm(w))
        v = w / lmax
    alpha = 1.0 / lmax
    blk = x_true[:, 0, 48:64, 48:64, 48:64].reshape(F, -1).T.cpu().numpy().astype(np.complex128)
    lam = 0.01 * float(np.linalg.svd(blk, compute_uv=False)[0])
    shifts = native.llr_shifts(args.llr_iters, (b, b, b))
    x = torch.zeros_like(x_true)
    z = torch.zeros_like(x_true)
    grad = torch.empty_like(x_true)
    x_old = torch.empty_like(x_true)
    from hastycompute_amd import _capi

    wsb = _capi.lib().hasty_llr_workspace_size(_capi.i64arr(list(x.shape)), _capi.i64arr((b, b, b)),
                                               _capi.i64arr((b, b, b)))
    ws = torch.empty(wsb, dtype=torch.uint8, device=dev)

    def iteration(it, t):
        x_old.copy_(x)
        # x = prox(z - alpha (T z - b)); the subtraction of b is fused into the operator
        op.apply(z.reshape(F, n, n, n), out=grad.reshape(F, n, n, n), kernel_index=kidx,
                 sub=rhs.reshape(F, n, n, n))
        torch.add(z, grad, alpha=-alpha, out=x)
        # iteration it uses the it-th shift of the reference's block-shift sequence
        sh = [[shifts[d][it % len(shifts[d])]] for d in range(3)]
        native.llr_normal_blocks(x, (b, b, b), (b, b, b), 1, lam, True, shifts=sh, workspace=ws)
        t_new = (1.0 + math.sqrt(1.0 + 4.0 * t * t)) / 2.0
        torch.sub(x, x_old, out=z)
        z.mul_((t - 1.0) / t_new).add_(x)
        return t_new

    t = 1.0
    for it in range(args.warmup):
        t = iteration(it, t)
    x.zero_()
    z.zero_()
    t = 1.0
    torch.cuda.synchronize()
    native.timing_enable(True)
    t0 = time.perf_counter()
    for it in range(args.steps):
        t = iteration(it, t)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    kt = native.timing_read()
    native.timing_enable(False)
    err = float(torch.linalg.vector_norm(x - x_true) / torch.linalg.vector_norm(x_true))
    grp = {"normal_op_ms": sum(v[0] for k, v in kt.items() if k.startswith("sense_")),
           "llr_prox_ms": sum(v[0] for k, v in kt.items() if k.startswith("llr_"))}
    line = {"metric": "LLR-regularised 4D-flow recon: prox-gradient iterations/s @128^3 x 100 frames x 16 coils",
            "value": args.steps / dt, "unit": "iterations/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": dt / args.steps * 1e3, "higher_is_better": True,
            "dtype": "c64 (fp32)", "data": "synthetic rank-5 temporal phantom + noise, b = T x_noisy",
            "config": {"workload": f"FISTA with the LLR prox, [{F},1,{n},{n},{n}], {C} coils, 16^3 blocks, "
                                   f"lambda=0.01 sigma_max, {args.steps} iterations",
                       "frames": F, "coils": C, "lambda": lam, "step": alpha, "autotune": tune},
            "recon_seconds": dt, "rel_error_vs_truth": err,
            "per_iteration_ms": {k: v / args.steps for k, v in grp.items()},
            "kernels": {k: {"launches": v[1], "avg_ms": v[0] / v[1]} for k, v in kt.items()}}
    print(json.dumps(line), flush=True)


def serve_main(args, torch):
    """The Poisson request stream (matrix 96^3/128^3/160^3, 8-32 coils, 1k-10k spokes x 2N
    samples, 15 CG iterations) through the continuous-batching server, in real time."""
    from hastycompute_amd import serving

    reqs = serving.poisson_requests(args.requests, args.rate, args.seed, iters=args.cg_iters,
                                    llr_fraction=args.llr_fraction)
    srv = serving.ReconServer(autotune=not args.no_autotune)

    def progress(msg):  # one line per finished request on stderr (the JSON line stays alone on stdout)
        if " done " in msg:
            print(msg, file=sys.stderr, flush=True)

    st = srv.run(reqs, realtime=True, log=progress)
    line = {"metric": "served CG-SENSE recons/s with p50/p99 latency (Poisson stream)",
            "value": st["recons_per_s"], "unit": "recons/s", "n_gpus": 1, "higher_is_better": True,
            "dtype": "c64 (fp32), FP64 CG scalars", "data": "synthetic (seeded phantoms, coil maps, kooshball)",
            "config": {"workload": f"{args.requests} requests, Poisson mean {args.rate}/s, N in {{96,128,160}}, "
                                   f"8-32 coils, 1k-10k spokes x 2N, {args.cg_iters} CG iterations, "
                                   f"{args.llr_fraction:.0%} multi-frame (4-8 frames) with the LLR prox",
                       "seed": args.seed, "llr_requests": sum(r.llr_thresh is not None for r in reqs),
                       "autotune": st["autotune"]},
            "latency_note": "from arrival to result; the k-space is produced by a scanner thread before arrival",
            **{k: st[k] for k in ("latency_p50_s", "latency_p99_s", "latency_mean_s", "makespan_s", "mean_batch",
                                  "max_active", "peak_hbm_bytes", "ticks", "late_scanner_data")}}
    print(json.dumps(line), flush=True)


def stored_traffic(key: str, units_per_launch: float):
    """HBM bytes per launch from the calibrated FETCH_SIZE / WRITE_SIZE figures in
    profiles/traffic.json (bytes per work unit x this run's units per launch), or None
    when there is no figure or the kernel's source has changed since it was measured."""
    from hastycompute_amd import profiling

    if not os.path.exists(TRAFFIC_FILE):
        return None, "no counter figure"
    ent = json.load(open(TRAFFIC_FILE)).get(key)
    if not ent or "hbm_bytes_per_unit" not in ent:
        return None, f"no counter figure for {key}"
    if ent.get("source_sha256") != profiling.source_sha256(ent["kernel_symbol"]):
        return None, f"counter figure for {key} was measured on another version of the kernel source"
    return ent["hbm_bytes_per_unit"] * units_per_launch, (
        f"calibrated FETCH_SIZE + WRITE_SIZE, {ent['hbm_bytes_per_unit']:.4g} B per problem x "
        f"{units_per_launch:g} problems per launch")


def normal_op_bytes(N: int, C: int) -> int:
    """Algorithmic HBM bytes of one normal-operator apply: x in + out (16 N^3), coil
    maps (8 C N^3), one complex64 kernel read (8 (2N)^3) and one oversampled-grid write
    + read per coil (16 C (2N)^3)."""
    return 16 * N ** 3 + 8 * C * N ** 3 + 8 * (2 * N) ** 3 + 16 * C * (2 * N) ** 3


def kernel_bytes(name: str, N: int, C: int, cb: int, P: int) -> int:
    """Algorithmic bytes the fused path's kernels move for P problems (all coils, coil
    batches of cb). The grid between kernels is half-padded: (2N)^2 x N complex64 per
    coil (32 N^3 B); the Toeplitz kernel is real float32 (2N)^3 (32 N^3 B)."""
    n3 = N ** 3
    nbat = -(-C // cb)
    if name == "sense_fwd_xy":   # x per coil batch + coil maps + grid write
        return P * (nbat * 8 * n3 + 8 * C * n3 + 32 * C * n3)
    if name == "sense_z_conv":   # grid read + write + kernel read per coil batch
        return P * (64 * C * n3 + nbat * 32 * n3)
    if name == "sense_inv_xy":   # grid read + coil maps + out write (+ out read when accumulating)
        return P * (32 * C * n3 + 8 * C * n3 + 8 * n3 + (nbat - 1) * 16 * n3)
    return 0


def make_workload(args, torch, native):
    import numpy as np

    from hastycompute_amd import synthetic

    dev = torch.device("cuda:0")
    rng = np.random.default_rng(args.seed)
    spokes = rng.integers(args.spokes_min, args.spokes_max + 1, size=args.frames)
    offsets = np.concatenate([[0], np.cumsum(spokes)[:-1]])
    kernels = []
    npts = []
    for f in range(args.frames):
        co = synthetic.kooshball_torch(int(spokes[f]), args.readout, int(offsets[f]), dev)
        _, k = native.toeplitz_build(co, [args.n] * 3, tol=1e-8, diag=False)
        kernels.append(k)
        npts.append(co.shape[1])
        del co
    K = torch.stack(kernels)
    del kernels
    smaps = synthetic.coil_maps_torch(args.coils, args.n, dev, seed=args.seed)
    P = args.frames * args.encodings
    g = torch.Generator(device=dev)
    g.manual_seed(args.seed)
    x = torch.randn((P, args.n, args.n, args.n), dtype=torch.complex64, device=dev, generator=g)
    kidx = [p // args.encodings for p in range(P)]
    torch.cuda.synchronize()
    return K, smaps, x, kidx, npts, spokes, offsets


def cpu_threads() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_baseline(args, torch, K, smaps, x):
    """The reference's CPU path for this operator: the ATen call sequence of
    CUDANormalNufftToeplitz::apply_inplace (lib/fft/nufft.cpp:1188-1195: fft_fftn_out
    with s = 2N, mul_ by the complex diagonal, fft_ifftn_out, copy of the cropped
    corner) inside the coil loop of sense_normal_apply (lib/mri/sense.ixx:533-577:
    imstore = in * S_c, the Toeplitz apply, imstore *= conj(S_c), out += imstore), run
    in complex64 on torch-CPU with every host thread. One problem at the benchmark size;
    applies are repeated until about args.cpu_seconds have passed (at least one)."""
    threads = cpu_threads()
    prev = torch.get_num_threads()
    torch.set_num_threads(threads)
    N = args.n
    L = 2 * N
    diag = (K[0].cpu() * float(N) ** 3).to(torch.complex64).contiguous()  # the reference's complex diagonal
    S = smaps.cpu()
    xin = x[0:1].cpu()
    out = torch.zeros_like(xin)
    imstore = torch.empty_like(xin)
    storage1 = torch.empty((1, L, L, L), dtype=torch.complex64)
    storage2 = torch.empty_like(storage1)
    dims = (1, 2, 3)
    napply = 0
    t = time.perf_counter()
    with torch.inference_mode():
        while True:
            out.zero_()
            for c in range(args.coils):
                torch.mul(xin, S[c], out=imstore)
                torch.fft.fftn(imstore, s=(L, L, L), dim=dims, out=storage1)
                storage1.mul_(diag)
                torch.fft.ifftn(storage1, dim=dims, out=storage2)
                imstore.copy_(storage2[:, :N, :N, :N])
                imstore.mul_(S[c].conj())
                out.add_(imstore)
            napply += 1
            if time.perf_counter() - t >= args.cpu_seconds:
                break
    dt = time.perf_counter() - t
    torch.set_num_threads(prev)
    return {"value": napply / dt, "unit": "applies/s", "cores": threads, "kind": "port",
            "sample": f"1 problem x {napply} applies at {N}^3 x {args.coils} coils: the reference's ATen sequence "
                      f"(nufft.cpp:1188-1195 in the coil loop of sense.ixx:533-577), complex64, torch-CPU "
                      f"({threads} threads) on {platform.processor() or platform.machine()}",
            "seconds": dt}


def oracle_cg_image(args, K, smaps, b, p, lam):
    """The FP64 oracle's CG (15 iterations) on problem p's system A + lam I, A the Toeplitz
    SENSE normal operator with the problem's GPU-built kernel (so the comparison pins the
    arithmetic of the operator and the CG, not the kernel build, which the tests pin)."""
    import numpy as np

    from oracle import oracle as O

    N = args.n
    diag = (K[p // args.encodings].double().cpu().numpy() * float(N) ** 3).astype(np.complex128)[None]
    S = smaps.cpu().numpy()
    coils = list(range(args.coils))
    w = O.default_threads()
    A = lambda v: O.sense_normal_toeplitz(diag, v, S, coils, workers=w) + lam * v
    bp = b[p:p + 1].cpu().numpy().astype(np.complex128)
    return O.conjugate_gradient(A, bp, np.zeros_like(bp), args.cg_iters)


def percentile(v, q):
    import numpy as np

    return float(np.percentile(np.asarray(v, dtype=np.float64), q))


def plain_recons(args, torch, spokes, offsets, smaps, b_all, x_batched, toeplitz_apply, lam=0.0):
    """The plain per-frame drop-in: for each sampled problem, the reference's explicit
    SENSE normal operator (the coil loop of sense_normal_apply, lib/mri/sense.ixx:533-577,
    over CUDANufftNormal, lib/fft/nufft.cpp:862-880, through HastyNufft.NufftNormal with
    type-2 / type-1 options at args.plain_tol) plus lam I, inside the host CG of
    ConjugateGradient::run (opalgs.cpp:50-96), one problem at a time. Also returns, per
    problem, the relative difference of one apply (explicit vs Toeplitz operator), the
    relative residuals |T x - b| / |b| of both CG results under the Toeplitz operator T,
    and the images."""
    from hastycompute_amd import synthetic
    from hastycompute_amd.ffi import hasty_nufft

    N = args.n
    coils = list(range(args.coils))
    lat, diffs, apply_diffs, res_plain, res_batched, images = [], [], [], [], [], []
    for j in range(args.plain):
        p = j * args.encodings  # encoding 0 of frames 0, 1, ...
        f = p // args.encodings
        co = synthetic.kooshball_torch(int(spokes[f]), args.readout, int(offsets[f]), smaps.device)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        op = hasty_nufft.NufftNormal(co, [1, N, N, N], hasty_nufft.NufftOptions(2, False, args.plain_tol),
                                     hasty_nufft.NufftOptions(1, True, args.plain_tol))
        im = torch.empty((1, N, N, N), dtype=torch.complex64, device=smaps.device)
        res = torch.empty_like(im)
        ks = torch.empty((1, co.shape[1]), dtype=torch.complex64, device=smaps.device)

        def A(v):  # l2_reg_linop(SenseNormal, lam) (python/python/hastypy/base/opalg.py:626-627)
            out = torch.zeros_like(v)
            for c in coils:
                torch.mul(v, smaps[c], out=im)
                op.apply(im, res, ks)
                out.addcmul_(res, smaps[c].conj())
            if lam:
                out.add_(v, alpha=lam)
            return out

        b = b_all[p:p + 1]
        x = torch.zeros_like(b)
        r = b - A(x)
        pv = r.clone()
        rz = torch.vdot(r.flatten(), r.flatten()).real.item()
        for _ in range(args.cg_iters):
            Ap = A(pv)
            pap = torch.vdot(pv.flatten(), Ap.flatten()).real.item()
            if pap <= 0:
                raise RuntimeError("A was not positive definite")
            alpha = rz / pap
            x.add_(pv, alpha=alpha)
            r.sub_(Ap, alpha=alpha)
            rz_new = torch.vdot(r.flatten(), r.flatten()).real.item()
            pv.mul_(rz_new / rz).add_(r)
            rz = rz_new
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - t0)
        nrm = torch.linalg.vector_norm
        ref = x_batched[p:p + 1]
        diffs.append(float(nrm(x - ref) / nrm(ref)))
        ta, tr = toeplitz_apply(b, p), toeplitz_apply(x, p)
        apply_diffs.append(float(nrm(A(b) - ta) / nrm(ta)))
        res_plain.append(float(nrm(tr - b) / nrm(b)))
        res_batched.append(float(nrm(toeplitz_apply(ref, p) - b) / nrm(b)))
        images.append(x)
        del op, co
    return lat, diffs, apply_diffs, res_plain, res_batched, images


def copy_calibration(torch):
    """The streaming ceiling: hasty_stream_probe over 1 GiB streams for the read:write
    mixes of the hot kernels (1:0 like the inverse plane kernel, 1:1 like the
    z-convolution, 1:4 like the forward plane kernel, and others), plain and
    non-temporal, best of 5 event-timed runs; then a 1 GiB torch copy, the known-size
    kernel the PMC byte counters are calibrated against."""
    import ctypes

    from hastycompute_amd import _capi as C

    GiB = 1 << 30
    bufs = [torch.full((GiB // 4,), 1.0, dtype=torch.float32, device="cuda") for _ in range(5)]
    ptr = [b.data_ptr() for b in bufs]
    stream = torch.cuda.current_stream()
    rows = []
    for nr, nw in ((1, 0), (2, 0), (0, 1), (1, 1), (2, 1), (1, 2), (1, 4), (4, 1)):
        for nt in (0, 1):
            src = (ctypes.c_void_p * 4)(*(ptr[:nr] + [None] * (4 - nr)))
            dst_list = ptr[nr:nr + nw] if nw else [ptr[4]]
            dst = (ctypes.c_void_p * 4)(*(dst_list + [None] * (4 - len(dst_list))))

            def run():
                C.call("hasty_stream_probe", nr, nw, nt, src, dst, GiB, C.stream_ptr())

            run()
            best = float("inf")
            for _ in range(5):
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                run()
                e1.record(stream)
                e1.synchronize()
                best = min(best, e0.elapsed_time(e1))
            gbs = (nr + nw) * GiB / (best / 1e3) / 1e9
            rows.append({"read_streams": nr, "write_streams": nw, "nontemporal": bool(nt), "ms": best,
                         "GB_s": gbs, "frac_of_8TBs": gbs / HBM_PEAK_GBS})
    n = 1 << 27  # 2^27 complex64 = 1 GiB
    a = torch.ones(n, dtype=torch.complex64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"stream_ceiling_GiB_per_stream": 1, "mixes": rows, "copy_calibration_bytes_read": n * 8,
                      "copy_calibration_bytes_written": n * 8}), flush=True)


def cg_section(args, torch, native, op, x, K, kidx, smaps, spokes, offsets):
    """CG-SENSE over the 100 problems: 15 iterations of ConjugateGradient::run
    (lib/op/opalgs.cpp:50-96) on l2_reg_linop(A, lam) = A + lam I
    (python/python/hastypy/base/opalg.py:626-627) with lam = 0.1 lambda_max, the
    regularisation under which the 15th iterate is well conditioned in its inputs (on the
    unregularised kooshball system the FP64 CG itself moves by ~1e-5 under a one-ulp
    change of b). Batched: one device-CG batch of all problems, eager and replaying a HIP
    graph of one iteration; plain: the per-frame drop-in. Images of problem 0 are
    compared with the FP64 oracle CG outside the timed regions."""
    from hastycompute_amd import synthetic

    P = x.shape[0]
    # right-hand sides b = T x_true with x_true a phantom whose global phase varies over
    # the problems (the velocity encodings and frames)
    ph = synthetic.ellipsoid_phantom_torch(args.n, x.device, seed=args.seed)
    phase = torch.exp(1j * torch.linspace(0.0, 3.0, P, device=x.device, dtype=torch.float64)).to(torch.complex64)
    b = torch.empty_like(x)
    op.apply((ph[None] * phase[:, None, None, None]).contiguous(), out=b, kernel_index=kidx)
    del ph
    # lambda_max of problem 0 by power iteration
    v = b[0:1].clone()
    w = torch.empty_like(v)
    for _ in range(args.power_iters):
        op.apply(v, out=w, kernel_index=kidx[:1])
        lmax = float(torch.linalg.vector_norm(w))
        v = w / lmax
    lam = args.cg_l2 * lmax
    A = lambda v_, o_: op.apply(v_, out=o_, kernel_index=kidx)
    res = {"iters": args.cg_iters, "system": f"A + lam I, lam = {args.cg_l2} lambda_max (l2_reg_linop)",
           "lambda": lam, "lambda_max": lmax}
    xs = {}
    for mode in ("eager", "graph"):
        x0 = torch.zeros_like(x)
        torch.cuda.synchronize()
        t = time.perf_counter()
        xs[mode], _ = native.cg_solve(A, b, x0, args.cg_iters, check_pd=False, l2=lam, graph=mode == "graph")
        torch.cuda.synchronize()
        dt = time.perf_counter() - t
        res[f"batched_{mode}_recons_per_s"] = P / dt
        res[f"batched_{mode}_makespan_s"] = dt
        del x0
    res["graph_equals_eager"] = bool(torch.equal(xs["eager"], xs["graph"]))
    best = max(("eager", "graph"), key=lambda m: res[f"batched_{m}_recons_per_s"])
    res["batched_recons_per_s"] = res[f"batched_{best}_recons_per_s"]
    res["batched_mode"] = best
    res["batched_note"] = ("all problems submitted together and finishing with the batch, so the makespan is every "
                           "problem's latency; latency distributions come from --mode serve")
    xb = xs["graph"]
    del xs
    images = []
    if args.plain > 0:
        def t_apply(v, p):
            o = torch.empty_like(v)
            op.apply(v, out=o, kernel_index=[kidx[p]])
            return o.add_(v, alpha=lam)

        lat, diffs, adiff, rp, rb, images = plain_recons(args, torch, spokes, offsets, smaps, b, xb, t_apply, lam)
        res.update({"plain_recons_per_s": len(lat) / sum(lat), "plain_latency_p50_s": percentile(lat, 50),
                    "plain_latency_p99_s": percentile(lat, 99), "plain_sampled_problems": len(lat),
                    "speedup_recons_per_s": res["batched_recons_per_s"] / (len(lat) / sum(lat)),
                    "plain_vs_batched_apply_rel_diff_max": max(adiff),
                    "plain_vs_batched_image_rel_diff_max": max(diffs),
                    "residual_rel_plain_max": max(rp), "residual_rel_batched_max": max(rb)})
    if not args.no_oracle_cg:
        from oracle import oracle as O

        t = time.perf_counter()
        ref = oracle_cg_image(args, K, smaps, b, 0, lam)
        res["oracle_cg_seconds"] = time.perf_counter() - t
        res["image_rel_l2_vs_fp64_oracle"] = {"batched": O.rel_l2(xb[0:1].cpu().numpy(), ref)}
        if images:
            res["image_rel_l2_vs_fp64_oracle"]["plain"] = O.rel_l2(images[0].cpu().numpy(), ref)
        res["image_rel_l2_vs_fp64_oracle"]["problem"] = 0
    return res


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "bench.py runs on one GPU; --gpus must be 1 (how the path would shard over "
                                   "several GPUs is described, not launched)", "n_gpus": args.gpus}))
        sys.exit(2)
    import torch

    from hastycompute_amd import native

    torch.cuda.set_device(0)
    if args.set_option:
        from hastycompute_amd import _capi
        for item in args.set_option:
            name, _, value = item.partition("=")
            _capi.set_option(name, int(value))
    if args.copy_calibration:
        copy_calibration(torch)
        return
    if args.mode == "llr":
        llr_main(args, torch, native)
        return
    if args.mode == "nufft":
        nufft_main(args, torch, native)
        return
    if args.mode == "serve":
        serve_main(args, torch)
        return
    if args.mode == "llr-recon":
        llr_recon_main(args, torch, native)
        return
    if args.xy_variant is not None:
        from hastycompute_amd import _capi
        _capi.call("hasty_set_option", b"sense_xy_variant", args.xy_variant)
    if args.zconv_batch is not None:
        from hastycompute_amd import _capi
        _capi.call("hasty_set_option", b"sense_zconv_batch", args.zconv_batch)
    if args.zconv_variant is not None:
        from hastycompute_amd import _capi
        _capi.call("hasty_set_option", b"sense_zconv_variant", args.zconv_variant)
    tune = None
    manual = any(v is not None for v in (args.xy_variant, args.zconv_batch, args.zconv_variant, args.coil_batch,
                                         args.problem_batch))
    if not (args.no_autotune or args.generic or manual):
        from hastycompute_amd import autotune

        # before any kernel is prepared: the tuner sets the plane-kernel option
        tune = autotune.sense_normal([args.n] * 3, args.coils, args.frames * args.encodings)
        args.coil_batch, args.problem_batch = tune["coil_batch"], tune["problem_batch"]
    args.coil_batch = 16 if args.coil_batch is None else args.coil_batch
    args.problem_batch = 4 if args.problem_batch is None else args.problem_batch
    K, smaps, x, kidx, npts, spokes, offsets = make_workload(args, torch, native)
    P = x.shape[0]
    op = native.SenseNormalToeplitz(smaps, K, coil_batch=args.coil_batch, problem_batch=args.problem_batch,
                                    fused=False if args.generic else None)
    out = torch.empty_like(x)
    stream = torch.cuda.current_stream()

    for _ in range(args.warmup):
        op.apply(x, out=out, kernel_index=kidx)
    torch.cuda.synchronize()
    events = not args.no_kernel_events
    if events:
        native.timing_enable(True)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        op.apply(x, out=out, kernel_index=kidx)
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ktimes = native.timing_read() if events else {}
    if events:
        native.timing_enable(False)
    ms_total = e0.elapsed_time(e1)
    ms_per_step = ms_total / args.steps
    applies_per_s = P * args.steps / (ms_total / 1e3)
    alg_bytes = normal_op_bytes(args.n, args.coils)
    op_achieved = alg_bytes * P / (ms_per_step / 1e3) / 1e9

    kernels = {}
    for name, (ms, cnt) in ktimes.items():
        kb = kernel_bytes(name, args.n, args.coils, min(args.coil_batch, args.coils), P * args.steps)
        kernels[name] = {"launches": cnt, "avg_ms": ms / cnt, "total_ms": ms,
                         "alg_bytes_per_launch": kb / cnt if kb else None,
                         "achieved_GBs": kb / (ms / 1e3) / 1e9 if kb else None}
    roofline = {"bound": "hbm", "peak": HBM_PEAK_GBS, "unit": "GB/s"}
    if kernels:
        dom = max(kernels, key=lambda k: kernels[k]["total_ms"])
        d = kernels[dom]
        traffic, traffic_note = stored_traffic(f"{dom}@{args.n}x{args.coils}", P * args.steps / d["launches"])
        roofline.update({"kernel": dom, "achieved": d["achieved_GBs"], "frac": (d["achieved_GBs"] or 0) / HBM_PEAK_GBS,
                         "traffic": traffic, "traffic_note": traffic_note,
                         "alg_bytes_per_launch": d["alg_bytes_per_launch"],
                         "avg_launch_ms": d["avg_ms"], "timing": "HIP events on the launch stream over the timed region"})
    else:
        roofline.update({"kernel": "sense_normal (whole pipeline)", "achieved": op_achieved,
                         "frac": op_achieved / HBM_PEAK_GBS, "traffic": None})

    line = {
        "metric": "SENSE normal-op applies/s @128^3x16coil; recons/s + p50/p99 latency vs plain",
        "value": applies_per_s,
        "unit": "applies/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 (fp32)",
        "data": "synthetic (seeded kooshball trajectories, Gaussian x linear-phase coil maps, random images)",
        "config": {"workload": f"4D-flow batched normal op: {args.encodings} enc x {args.frames} frames = {P} problems, "
                               f"{args.n}^3, {args.coils} coils, grid {2 * args.n}^3, "
                               f"{args.spokes_min}-{args.spokes_max} spokes x {args.readout} per frame",
                   "problems": P, "n": args.n, "coils": args.coils, "coil_batch": args.coil_batch,
                   "problem_batch": args.problem_batch, "path": "fused" if op.layout != 0 else "generic",
                   "kernel_layout": op.layout, "autotune": tune,
                   "points_per_frame_mean": float(sum(npts) / len(npts))},
        "roofline": roofline,
        "op_roofline": {"achieved": op_achieved, "frac": op_achieved / HBM_PEAK_GBS, "alg_bytes_per_apply": alg_bytes,
                        "what": "whole normal-op pipeline per problem (all kernels)"},
        "kernels": kernels,
        "wall_s": wall,
    }

    if not args.no_cg:
        line["cg"] = cg_section(args, torch, native, op, x, K, kidx, smaps, spokes, offsets)
        line["recons_per_s"] = line["cg"]["batched_recons_per_s"]
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(args, torch, K, smaps, x)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
