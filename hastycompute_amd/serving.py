"""Continuous-batching CG-SENSE reconstruction server on one GPU.

Requests (a 3-D volume each: matrix N^3, C coils, a kooshball trajectory, k-space data)
arrive over time. The server keeps, per image size, one group of in-flight problems
whose coil maps, Toeplitz kernels and CG vectors live in HBM pools; every loop tick
runs ONE batched normal-operator launch sequence (hasty_sense_normal_toeplitz_ex, each
problem with its own maps and coil count) and ONE batched device-CG step per group.
New requests are admitted between ticks, in arrival order, while the HBM in use plus
the request's footprint stays within the budget; a finished request leaves its slot
(the last row moves into it), so the batch shrinks and grows continuously.

A request may span several time frames (one problem per frame, each with its own
trajectory); when it asks for it, the LLR prox (16^3 blocks, soft threshold,
OldNormal3DBlocksSVT) runs over its frames once their CG iterations are done, as the
reference's prox step does (python/python/hastypy/base/recon.py:184-190).

Each problem's arithmetic does not depend on which other problems share its launches
(per-problem workgroups, fixed-order reductions), so a request's result is the same
alone and in a mix; tests/test_gpu_serving.py checks that. In real-time runs a scanner
thread produces each request's k-space before it arrives, so latency covers serving
work only (A^H y, the Toeplitz kernel, CG, the optional prox).

The reference has no server: it reconstructs frames with BatchedSenseNormal /
BatchedConjugateGradient over host-thread contexts (lib/mri/batch_sense.cpp:349-450,
lib/op/opalgs.ixx:84-136, lib/threading/thread_pool.ixx:110-259).

    python -m hastycompute_amd.serving --requests 200 --rate 2 --seed 0
"""
from __future__ import annotations

import argparse
import ctypes
import json
import math
import queue as queue_mod
import threading
import time
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np
import torch

from . import _capi as C
from . import native, synthetic


@dataclass
class ReconRequest:
    rid: int
    n: int
    ncoils: int
    nspokes: int
    arrival: float = 0.0   # seconds after the server starts
    iters: int = 15
    seed: int = 0
    readout: Optional[int] = None  # samples per spoke (default 2 n)
    noise: float = 1e-3
    frames: int = 1        # time frames; each is a problem with its own trajectory
    llr_thresh: Optional[float] = None  # LLR prox on the frames after CG (16^3 blocks, soft)
    # (smaps, kernels [F, 8 N^3], A^H y [F, N, N, N]) computed earlier by ReconServer.setup;
    # None: set up on admission
    prepared: Optional[tuple] = None
    # the scanner's data (trajectories, coil maps, k-space), produced before the arrival
    scan: Optional[tuple] = None
    # filled in by the server
    t_admit: float = field(default=float("nan"))
    t_done: float = field(default=float("nan"))
    image: Optional[torch.Tensor] = None
    _out: dict = field(default_factory=dict, repr=False)

    @property
    def samples(self):
        return self.nspokes * (self.readout or 2 * self.n)

    @property
    def latency(self):
        return self.t_done - self.arrival


def resident_bytes(n: int, ncoils: int) -> int:
    """HBM a problem (one frame of a request) holds while in flight: coil maps, real
    Toeplitz kernel, five CG vectors (x, r, p, Ap, b)."""
    return 8 * ncoils * n ** 3 + 4 * (2 * n) ** 3 + 5 * 8 * n ** 3


def setup_bytes(n: int, ncoils: int, samples: int) -> int:
    """Transient HBM of a request's setup (frames are set up one after another): the FP64
    PSF NUFFT fine grid (4N)^3 and its 2N-mode output, one complex64 fine grid (2N)^3 for
    the data NUFFTs, the samples."""
    return 16 * (4 * n) ** 3 + 16 * (2 * n) ** 3 + 8 * (2 * n) ** 3 + samples * (12 + 8 * ncoils + 8)


def scan_bytes(req: "ReconRequest") -> int:
    """HBM of the scanner's data of a request until its setup has used it."""
    return req.frames * req.samples * (12 + 8 * req.ncoils) + 8 * req.ncoils * req.n ** 3


def acquire(req: ReconRequest, device, nufft_tol: float = 1e-6):
    """The scanner: per frame a kooshball trajectory (its own spoke range), the coil
    maps, and k-space y_c = N2(S_c x) + noise of a seeded phantom (complex64 on the GPU).
    Returns (coords list [3, M] per frame, smaps [C, N, N, N], kdata list [C, M] per frame)."""
    n, c = req.n, req.ncoils
    readout = req.readout or 2 * n
    g = torch.Generator(device=device)
    g.manual_seed(1000003 * req.seed + 17)
    smaps = synthetic.coil_maps_torch(c, n, device, seed=req.seed)
    x_true = synthetic.ellipsoid_phantom_torch(n, device, seed=req.seed)
    scale = 1.0 / math.sqrt(n ** 3)
    coords, kdata = [], []
    for f in range(req.frames):
        co = synthetic.kooshball_torch(req.nspokes, readout, (req.seed * 7919 + f * req.nspokes) % 1000003, device)
        fwd = native.NufftPlan(co, [n, n, n], 2, -1, 1, nufft_tol, scale)
        y = torch.empty((c, co.shape[1]), dtype=torch.complex64, device=device)
        for ci in range(c):
            fwd.execute(y[ci:ci + 1], (x_true * smaps[ci]).reshape(1, n, n, n).contiguous())
        y += req.noise * torch.randn(y.shape, dtype=torch.complex64, device=device, generator=g)
        fwd.close()
        coords.append(co)
        kdata.append(y)
    return coords, smaps, kdata


class _Group:
    """The in-flight problems (request frames) of one image size N; coil counts may
    differ. Slot i holds frame slot_frame[i] of request slot_req[i]."""

    def __init__(self, n: int, cmax: int, device, coil_batch: int, problem_batch: int, capacity: int = 4,
                 options: Optional[dict] = None):
        self.n, self.cmax, self.device = n, cmax, device
        self.n3 = [n, n, n]
        self.nvox = n ** 3
        self.coil_batch = min(coil_batch, cmax)
        self.problem_batch = problem_batch
        self.options = dict(options or {})  # library options this group's launches use
        self.slot_req: List[ReconRequest] = []
        self.slot_frame: List[int] = []
        self.ncoil_host: List[int] = []
        self.iters_done: List[int] = []
        self.layout = None  # the kernels' layout, fixed by the first prepared kernel
        self.coils_dev = torch.arange(cmax, dtype=torch.int32, device=device)
        ws = C.lib().hasty_sense_normal_workspace_size(C.i64arr(self.n3), self.coil_batch, problem_batch)
        self.ws = torch.empty(ws, dtype=torch.uint8, device=device)
        self.cap = 0
        self._alloc(capacity)

    def _use_options(self):
        for k, v in self.options.items():
            C.set_option(k, v)

    def _alloc(self, cap: int):
        n, dev = self.n, self.device
        old = None
        if self.cap:
            old = (self.smaps, self.kern, self.x, self.r, self.p, self.Ap, self.b, self.state, self.ncoil_dev)
        self.smaps = torch.zeros((cap * self.cmax, n, n, n), dtype=torch.complex64, device=dev)
        self.kern = torch.empty((cap, 8 * self.nvox), dtype=torch.float32, device=dev)
        self.x, self.r, self.p, self.Ap, self.b = (torch.zeros((cap, n, n, n), dtype=torch.complex64, device=dev)
                                                   for _ in range(5))
        self.state = torch.zeros((cap, 8), dtype=torch.float64, device=dev)
        self.ncoil_dev = torch.ones(cap, dtype=torch.int32, device=dev)
        self.cgws = torch.empty(C.lib().hasty_cg_workspace_size(cap, self.nvox), dtype=torch.uint8, device=dev)
        if old is not None:
            k = self.active
            for new_t, old_t, rows in zip((self.smaps, self.kern, self.x, self.r, self.p, self.Ap, self.b, self.state,
                                           self.ncoil_dev), old, [k * self.cmax] + [k] * 8):
                new_t[:rows].copy_(old_t[:rows])
        self.cap = cap

    @property
    def active(self):
        return len(self.slot_req)

    def add(self, req: ReconRequest, smaps: torch.Tensor, kernels_nat: torch.Tensor, b: torch.Tensor):
        """Adds the request's frames as problems: kernels_nat [F, 8 N^3], b [F, N, N, N]."""
        self._use_options()
        c = req.ncoils
        for f in range(req.frames):
            if self.active == self.cap:
                self._alloc(2 * self.cap)
            i = self.active
            self.smaps[i * self.cmax:i * self.cmax + c].copy_(smaps)
            layout = ctypes.c_int()
            C.call("hasty_sense_normal_prepare", kernels_nat[f].data_ptr(), self.kern[i].data_ptr(), 1,
                   C.i64arr(self.n3), ctypes.byref(layout), C.stream_ptr())
            if self.layout is None:
                self.layout = layout.value
            elif layout.value != self.layout:
                raise RuntimeError("kernel layout changed within a size group")
            self.b[i].copy_(b[f])
            self.x[i].zero_()
            self.Ap[i].zero_()
            self.ncoil_dev[i] = c
            C.call("hasty_cg_init", 1, self.nvox, self.b[i].data_ptr(), self.Ap[i].data_ptr(), self.r[i].data_ptr(),
                   self.p[i].data_ptr(), self.state[i].data_ptr(), self.cgws.data_ptr(), C.stream_ptr())
            self.slot_req.append(req)
            self.slot_frame.append(f)
            self.ncoil_host.append(c)
            self.iters_done.append(0)

    def step(self):
        P = self.active
        if P == 0:
            return
        self._use_options()
        C.call("hasty_sense_normal_toeplitz_ex", self.p.data_ptr(), self.Ap.data_ptr(), self.smaps.data_ptr(),
               self.cap * self.cmax, self.coils_dev.data_ptr(), self.cmax, self.cmax, C.i32arr(self.ncoil_host),
               self.ncoil_dev.data_ptr(), self.kern.data_ptr(), self.cap, self.layout, C.i32arr(list(range(P))),
               None, P, C.i64arr(self.n3), self.ws.data_ptr(), self.ws.numel(), self.coil_batch, C.stream_ptr())
        C.call("hasty_cg_iterate", P, self.nvox, self.x.data_ptr(), self.r.data_ptr(), self.p.data_ptr(),
               self.Ap.data_ptr(), self.state.data_ptr(), self.cgws.data_ptr(), 0.0, C.stream_ptr())
        for i in range(P):
            self.iters_done[i] += 1

    def pop_finished(self) -> List[ReconRequest]:
        """Frees the slots whose frames have run their iterations (keeping each frame's
        image) and returns the requests whose frames are all done."""
        done = []
        for i in reversed(range(self.active)):
            req = self.slot_req[i]
            if self.iters_done[i] < req.iters:
                continue
            req._out[self.slot_frame[i]] = self.x[i].clone()
            if len(req._out) == req.frames:
                done.append(req)
            last = self.active - 1
            if i != last:  # the last row takes the freed slot
                cm = self.cmax
                self.smaps[i * cm:(i + 1) * cm].copy_(self.smaps[last * cm:(last + 1) * cm])
                for t in (self.kern, self.x, self.r, self.p, self.Ap, self.b, self.state, self.ncoil_dev):
                    t[i].copy_(t[last])
                for lst in (self.slot_req, self.slot_frame, self.ncoil_host, self.iters_done):
                    lst[i] = lst[last]
            for lst in (self.slot_req, self.slot_frame, self.ncoil_host, self.iters_done):
                lst.pop()
        return done


class _Scanner(threading.Thread):
    """Produces each request's scanner data (acquire) on its own stream, in arrival
    order and up to `lookahead` requests ahead of the ones the server has taken, so that
    a request's data exists when it arrives and its latency covers serving work only."""

    def __init__(self, requests: List[ReconRequest], device, nufft_tol: float, lookahead: int = 6):
        super().__init__(daemon=True)
        self.requests, self.device, self.nufft_tol, self.lookahead = requests, device, nufft_tol, lookahead
        self.cv = threading.Condition()
        self.produced = 0
        self.taken = 0
        self.error: Optional[BaseException] = None
        self.stop = False

    def run(self):
        torch.cuda.set_device(self.device)
        stream = torch.cuda.Stream(device=self.device)
        try:
            for req in self.requests:
                with self.cv:
                    self.cv.wait_for(lambda: self.stop or self.produced - self.taken < self.lookahead)
                    if self.stop:
                        return
                if req.scan is None and req.prepared is None:
                    with torch.cuda.stream(stream):
                        req.scan = acquire(req, self.device, self.nufft_tol)
                    stream.synchronize()
                with self.cv:
                    self.produced += 1
                    self.cv.notify_all()
        except BaseException as e:  # surfaced by the server loop
            self.error = e
            with self.cv:
                self.produced = len(self.requests)
                self.cv.notify_all()

    def ready(self, k: int) -> bool:
        """True once the data of the first k requests (in arrival order) exists."""
        if self.error is not None:
            raise self.error
        return self.produced >= k

    def took(self):
        with self.cv:
            self.taken += 1
            self.cv.notify_all()

    def close(self):
        with self.cv:
            self.stop = True
            self.cv.notify_all()
        self.join()


class ReconServer:
    def __init__(self, device=None, hbm_budget: Optional[int] = None, coil_batch: int = 16, problem_batch: int = 4,
                 max_active: int = 64, cmax: int = 32, nufft_tol: float = 1e-6, kernel_tol: float = 1e-7,
                 autotune: bool = False):
        self.device = torch.device(device or "cuda:0")
        free, total = torch.cuda.mem_get_info(self.device)
        self.hbm_budget = hbm_budget if hbm_budget is not None else int(0.8 * total)
        self.coil_batch, self.problem_batch, self.max_active, self.cmax = coil_batch, problem_batch, max_active, cmax
        self.nufft_tol, self.kernel_tol = nufft_tol, kernel_tol
        self.autotune = autotune
        self.tuned = {}  # N -> autotuner choice
        self.groups = {}
        # (rid, HBM in use + reserved for admitted requests still in setup, footprint,
        # problems active before) per admission
        self.admissions = []
        self.reserved = 0

    # -- request setup (on arrival: A^H y and the trajectory's Toeplitz kernel per frame) --
    def setup(self, req: ReconRequest):
        dev = self.device
        n, c = req.n, req.ncoils
        coords, smaps, kdata = req.scan if req.scan is not None else acquire(req, dev, self.nufft_tol)
        scale = 1.0 / math.sqrt(n ** 3)
        kernels = torch.empty((req.frames, 8 * n ** 3), dtype=torch.float32, device=dev)
        b = torch.zeros((req.frames, n, n, n), dtype=torch.complex64, device=dev)
        im = torch.empty((1, n, n, n), dtype=torch.complex64, device=dev)
        for f in range(req.frames):
            adj = native.NufftPlan(coords[f], [n, n, n], 1, +1, 1, self.nufft_tol, scale)
            for ci in range(c):
                adj.execute(kdata[f][ci:ci + 1], im)
                b[f] += smaps[ci].conj() * im[0]
            adj.close()
            _, kern = native.toeplitz_build(coords[f], [n, n, n], tol=self.kernel_tol, diag=False)
            kernels[f].copy_(kern.reshape(-1))
            del kern
        return smaps, kernels, b

    def footprint(self, req: ReconRequest) -> int:
        return req.frames * resident_bytes(req.n, req.ncoils) + setup_bytes(req.n, req.ncoils, req.samples)

    def _group(self, n):
        if n not in self.groups:
            cb, pb, opts = self.coil_batch, self.problem_batch, {}
            if self.autotune:
                from . import autotune

                t = self.tuned.get(n) or autotune.sense_normal([n, n, n], self.cmax, 8)
                self.tuned[n] = t
                cb, pb = t["coil_batch"], t["problem_batch"]
                opts = {k: t[k] for k in ("sense_xy_variant", "sense_zconv_variant", "sense_zconv_batch")}
            self.groups[n] = _Group(n, self.cmax, self.device, cb, pb, options=opts)
        return self.groups[n]

    def _finish(self, req: ReconRequest, now_fn):
        """All frames done: the optional LLR prox over the frames, then the result."""
        frames = [req._out[f] for f in range(req.frames)]
        req._out = {}
        if req.llr_thresh is not None:
            x = torch.stack(frames).reshape(req.frames, 1, req.n, req.n, req.n).contiguous()
            native.llr_normal_blocks(x, (16, 16, 16), (16, 16, 16), 1, float(req.llr_thresh), True)
            frames = list(x[:, 0])
        req.image = frames[0].cpu() if req.frames == 1 else torch.stack(frames).cpu()
        req.t_done = now_fn()

    def _setup_worker(self, inq: "queue_mod.Queue", outq: "queue_mod.Queue"):
        """Request setup on its own stream (default priority) so admissions overlap the
        in-flight batches; results go back with an event the main stream waits on."""
        torch.cuda.set_device(self.device)
        stream = torch.cuda.Stream(device=self.device)
        while True:
            req = inq.get()
            if req is None:
                return
            try:
                with torch.cuda.stream(stream):
                    data = req.prepared if req.prepared is not None else self.setup(req)
                    req.scan = None  # the scanner's data is no longer needed
                    ev = torch.cuda.Event()
                    ev.record(stream)
                outq.put((req, data, ev))
            except Exception as e:  # surfaced by the main loop
                outq.put((req, e, None))

    def run(self, requests: List[ReconRequest], realtime: bool = True, log=None,
            async_setup: Optional[bool] = None, scanner: Optional[bool] = None) -> dict:
        """Serve `requests`. realtime: arrivals follow the wall clock (latency in seconds);
        otherwise one arrival joins per tick (latency in ticks). async_setup (default: with
        realtime) runs request setup on a worker thread and stream. scanner (default: with
        realtime) produces each request's k-space ahead of its arrival on another thread,
        so that latency measures serving work only; without it the setup also simulates
        the data."""
        for r in requests:
            if r.ncoils > self.cmax:
                raise RuntimeError(f"request {r.rid}: {r.ncoils} coils > server limit {self.cmax}")
            if r.frames < 1:
                raise RuntimeError(f"request {r.rid}: frames must be >= 1")
        if async_setup is None:
            async_setup = realtime
        if scanner is None:
            scanner = realtime
        for n in sorted({r.n for r in requests}):  # size groups (and their tuning) before the clock starts
            self._group(n)
        pending = sorted(requests, key=lambda r: (r.arrival, r.rid))
        scan = _Scanner(list(pending), self.device, self.nufft_tol) if scanner else None
        if scan is not None:
            scan.start()
            while not scan.ready(min(len(pending), scan.lookahead)):
                time.sleep(0.01)
        queue: List[ReconRequest] = []
        finished: List[ReconRequest] = []
        t0 = time.perf_counter()
        clock = (lambda: time.perf_counter() - t0) if realtime else None
        vclock = [0.0]  # without real time: arrivals are taken in order, one per tick at most
        now = clock if realtime else (lambda: vclock[0])
        ticks, batch_sum, max_seen = 0, 0, 0
        in_setup, taken, late = 0, 0, set()
        torch.cuda.reset_peak_memory_stats(self.device)
        main_stream = torch.cuda.current_stream(self.device)
        if async_setup:
            inq, outq = queue_mod.Queue(), queue_mod.Queue()
            worker = threading.Thread(target=self._setup_worker, args=(inq, outq), daemon=True)
            worker.start()

        def join_group(req, data):
            smaps, kern, b = data
            self._group(req.n).add(req, smaps, kern, b)

        try:
            while len(finished) < len(requests):
                t = now()
                while pending and (pending[0].arrival <= t or not realtime):
                    if scan is not None and not scan.ready(taken + 1):
                        if pending[0].arrival <= t:
                            late.add(pending[0].rid)  # arrived before the scanner had its data
                        break
                    queue.append(pending.pop(0))
                    taken += 1
                    if scan is not None:
                        scan.took()
                    if not realtime:
                        break
                if async_setup:  # finished setups join their groups
                    while True:
                        try:
                            req, data, ev = outq.get_nowait()
                        except queue_mod.Empty:
                            break
                        in_setup -= req.frames
                        self.reserved -= self.footprint(req)
                        if isinstance(data, Exception):
                            raise data
                        main_stream.wait_event(ev)
                        for tns in data:
                            tns.record_stream(main_stream)
                        join_group(req, data)
                        del data
                active = sum(g.active for g in self.groups.values()) + in_setup
                while queue and active < self.max_active:
                    req = queue[0]
                    used = torch.cuda.memory_allocated(self.device) + self.reserved
                    fp = self.footprint(req)
                    if used + fp > self.hbm_budget and active > 0:
                        break  # wait for memory to free up
                    queue.pop(0)
                    req.t_admit = now()
                    self.admissions.append((req.rid, used, fp, active))
                    if async_setup:
                        self.reserved += fp
                        inq.put(req)
                        in_setup += req.frames
                    else:
                        data = req.prepared if req.prepared is not None else self.setup(req)
                        req.scan = None
                        join_group(req, data)
                    active += req.frames
                    if log:
                        log(f"t={now():.2f}s admit {req.rid} N={req.n} C={req.ncoils} M={req.samples} "
                            f"F={req.frames} active={active}")
                if sum(g.active for g in self.groups.values()) == 0:
                    if async_setup and in_setup:
                        time.sleep(0.002)
                    elif realtime and pending:
                        time.sleep(max(0.0, min(0.05, pending[0].arrival - now())))
                    continue
                max_seen = max(max_seen, active)
                for g in self.groups.values():
                    batch_sum += g.active
                    g.step()
                ticks += 1
                if not realtime:
                    vclock[0] += 1.0
                for g in self.groups.values():
                    for req in g.pop_finished():
                        self._finish(req, now)
                        finished.append(req)
                        if log:
                            log(f"t={now():.2f}s done {req.rid} latency={req.latency:.3f}s")
        finally:
            if async_setup:
                inq.put(None)
                worker.join()
            if scan is not None:
                scan.close()
        torch.cuda.synchronize(self.device)
        makespan = now() - min(r.arrival for r in requests) if realtime else float(ticks)
        lat = np.array([r.latency for r in finished])
        return {"requests": len(finished), "recons_per_s": len(finished) / makespan if realtime else None,
                "makespan_s": makespan, "latency_p50_s": float(np.percentile(lat, 50)),
                "latency_p99_s": float(np.percentile(lat, 99)), "latency_mean_s": float(lat.mean()),
                "ticks": ticks, "mean_batch": batch_sum / max(1, ticks), "max_active": max_seen,
                "peak_hbm_bytes": int(torch.cuda.max_memory_allocated(self.device)),
                "late_scanner_data": len(late), "autotune": {str(k): v for k, v in self.tuned.items()}}


def poisson_requests(count: int, rate: float, seed: int = 0, sizes=(96, 128, 160), coils=(8, 32),
                     spokes=(1000, 10000), iters: int = 15, llr_fraction: float = 0.0, llr_frames=(4, 8),
                     llr_thresh: float = 0.05) -> List[ReconRequest]:
    """A seeded Poisson stream: exponential inter-arrival times with mean 1/rate; matrix
    size, coil count and spoke count drawn uniformly from the given sets / ranges. A
    fraction llr_fraction of the requests are multi-frame jobs (frames drawn from
    llr_frames, the spoke count split over the frames) with the LLR prox after CG."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(count):
        t += float(rng.exponential(1.0 / rate))
        r = ReconRequest(rid=i, n=int(rng.choice(sizes)), ncoils=int(rng.integers(coils[0], coils[1] + 1)),
                         nspokes=int(rng.integers(spokes[0], spokes[1] + 1)), arrival=t, iters=iters,
                         seed=seed * 100003 + i)
        if llr_fraction > 0 and rng.uniform() < llr_fraction:
            r.frames = int(rng.integers(llr_frames[0], llr_frames[1] + 1))
            r.nspokes = max(100, r.nspokes // r.frames)
            r.llr_thresh = llr_thresh
        out.append(r)
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m hastycompute_amd.serving")
    ap.add_argument("--requests", type=int, default=200)
    ap.add_argument("--rate", type=float, default=2.0, help="mean arrivals per second")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--iters", type=int, default=15)
    ap.add_argument("--budget-gb", type=float, default=None)
    ap.add_argument("--llr-fraction", type=float, default=0.0, help="share of multi-frame requests with the LLR prox")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--verbose", action="store_true")
    a = ap.parse_args(argv)
    torch.cuda.set_device(0)
    reqs = poisson_requests(a.requests, a.rate, a.seed, iters=a.iters, llr_fraction=a.llr_fraction)
    srv = ReconServer(hbm_budget=None if a.budget_gb is None else int(a.budget_gb * 2 ** 30),
                      autotune=not a.no_autotune)
    stats = srv.run(reqs, realtime=True, log=print if a.verbose else None)
    stats.update({"rate": a.rate, "seed": a.seed, "iters": a.iters})
    print(json.dumps(stats), flush=True)


if __name__ == "__main__":
    main()
