#!/usr/bin/env python3
"""Benchmark of the batched Toeplitz SENSE normal operator on one MI355X.

Workload (the 4D-flow scan): 5 velocity encodings x 20 cardiac frames = 100 problems,
image 128^3 complex64, 16 coils, oversampled grid 256^3; kooshball readouts of 128
samples with per-frame spoke counts drawn from U{800..1500} in a seed-fixed golden-means
order; one Toeplitz kernel per frame (shared by its 5 encodings). A step is one batched
normal-operator apply over all 100 problems with inputs resident in HBM; the value is
problems * steps / seconds (normal-op applies/s).

Also reported: CG-SENSE recons/s (15 iterations per problem, batched device CG), the
roofline of the apply, and the CPU baseline (the FP64 oracle's Toeplitz SENSE normal
operator on host cores, one problem x one apply).

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
    ap.add_argument("--coil-batch", type=int, default=16)
    ap.add_argument("--problem-batch", type=int, default=4)
    ap.add_argument("--generic", action="store_true", help="use the generic five-pass path")
    ap.add_argument("--cg-iters", type=int, default=15)
    ap.add_argument("--no-cg", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--seed", type=int, default=2024)
    return ap.parse_args()


def normal_op_bytes(N: int, C: int) -> int:
    """Algorithmic HBM bytes of one normal-operator apply: x in + out (16 N^3), coil
    maps (8 C N^3), one complex64 kernel read (8 (2N)^3) and one oversampled-grid write
    + read per coil (16 C (2N)^3)."""
    return 16 * N ** 3 + 8 * C * N ** 3 + 8 * (2 * N) ** 3 + 16 * C * (2 * N) ** 3


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
    return K, smaps, x, kidx, npts


def cpu_baseline(args, K, smaps, x):
    """The FP64 oracle's Toeplitz SENSE normal operator on host cores for one problem x
    one apply at the benchmark size (kernel values come from the GPU build; the timing
    does not depend on them)."""
    import numpy as np

    from oracle import oracle as O

    threads = O.default_threads()
    N = args.n
    # the oracle takes the reference's complex diagonal: D = kernel * N^d
    diag = (K[0].double().cpu().numpy() * float(N) ** 3).astype(np.complex128)[None]
    S = smaps.cpu().numpy().astype(np.complex128)
    x0 = x[0:1].cpu().numpy().astype(np.complex128)
    t = time.perf_counter()
    O.sense_normal_toeplitz(diag, x0, S, list(range(args.coils)), workers=threads)
    dt = time.perf_counter() - t
    return {"value": 1.0 / dt, "unit": "applies/s", "cores": threads, "kind": "port",
            "sample": f"1 problem x 1 apply at {N}^3 x {args.coils} coils, FP64 oracle "
                      f"(scipy.fft, {threads} threads) on {platform.processor() or platform.machine()}",
            "seconds": dt}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (multi-GPU sharding is described "
                                   "in DESIGN.md, not launched)", "n_gpus": args.gpus}))
        sys.exit(2)
    import torch

    from hastycompute_amd import native

    torch.cuda.set_device(0)
    K, smaps, x, kidx, npts = make_workload(args, torch, native)
    P = x.shape[0]
    op = native.SenseNormalToeplitz(smaps, K, coil_batch=args.coil_batch, problem_batch=args.problem_batch,
                                    fused=False if args.generic else None)
    out = torch.empty_like(x)
    stream = torch.cuda.current_stream()

    for _ in range(args.warmup):
        op.apply(x, out=out, kernel_index=kidx)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        op.apply(x, out=out, kernel_index=kidx)
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_total = e0.elapsed_time(e1)
    ms_per_step = ms_total / args.steps
    applies_per_s = P * args.steps / (ms_total / 1e3)
    alg_bytes = normal_op_bytes(args.n, args.coils)
    achieved = alg_bytes * P / (ms_per_step / 1e3) / 1e9

    extra = {}
    if not args.no_cg:
        # CG-SENSE: b = A x_true, 15 iterations from zero, all problems in one batch
        from hastycompute_amd.native import cg_solve

        b = out  # any right-hand side of the right shape
        x0 = torch.zeros_like(x)
        torch.cuda.synchronize()
        t = time.perf_counter()
        cg_solve(lambda v, o: op.apply(v, out=o, kernel_index=kidx), b, x0, args.cg_iters, check_pd=False)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t
        extra["recons_per_s"] = P / dt
        extra["cg_iters"] = args.cg_iters

    line = {
        "metric": "SENSE normal-op applies/s @128^3x16coil",
        "value": applies_per_s,
        "unit": "applies/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "complex64 (fp32)",
        "data": "synthetic (seeded kooshball trajectories, Gaussian x linear-phase coil maps, random images)",
        "config": {"workload": "4D-flow batched normal op: 5 enc x 20 frames = 100 problems, 128^3, 16 coils, "
                               "grid 256^3, 800-1500 spokes x 128 per frame",
                   "problems": P, "n": args.n, "coils": args.coils, "coil_batch": args.coil_batch,
                   "problem_batch": args.problem_batch, "path": "fused" if op.layout == 1 else "generic",
                   "points_per_frame_mean": float(sum(npts) / len(npts))},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "sense_normal_apply (whole Toeplitz pipeline per problem)",
                     "alg_bytes_per_problem": alg_bytes},
        "wall_s": wall,
    }
    line.update(extra)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(args, K, smaps, x)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
