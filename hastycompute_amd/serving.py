"""Continuous-batching CG-SENSE reconstruction server on one GPU.

Requests (a 3-D volume each: matrix N^3, C coils, a kooshball trajectory, k-space data)
arrive over time. The server keeps, per image size, one group of in-flight problems
whose coil maps, Toeplitz kernels and CG vectors live in HBM pools; every loop tick
runs ONE batched normal-operator launch sequence (hasty_sense_normal_toeplitz_ex, each
problem with its own maps and coil count) and ONE batched device-CG step per group.
New requests are admitted between ticks, in arrival order, while the HBM in use plus
the request's footprint stays within the budget; a finished request leaves its slot
(the last row moves into it), so the batch shrinks and grows continuously.

Each problem's arithmetic does not depend on which other problems share its launches
(per-problem workgroups, fixed-order reductions), so a request's result is the same
alone and in a mix; tests/test_gpu_serving.py checks that.

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
    # (smaps, kernel, A^H y) computed earlier by ReconServer.setup; None: set up on admission
    prepared: Optional[tuple] = None
    # filled in by the server
    t_admit: float = field(default=float("nan"))
    t_done: float = field(default=float("nan"))
    image: Optional[torch.Tensor] = None

    @property
    def samples(self):
        return self.nspokes * (self.readout or 2 * self.n)

    @property
    def latency(self):
        return self.t_done - self.arrival


def resident_bytes(n: int, ncoils: int) -> int:
    """HBM a request holds while in flight: coil maps, real Toeplitz kernel, five CG
    vectors (x, r, p, Ap, b)."""
    return 8 * ncoils * n ** 3 + 4 * (2 * n) ** 3 + 5 * 8 * n ** 3


def setup_bytes(n: int, ncoils: int, samples: int) -> int:
    """Transient HBM of a request's setup: the FP64 PSF NUFFT fine grid (4N)^3 and its
    2N-mode output, one complex64 fine grid (2N)^3 for the data NUFFTs, the samples."""
    return 16 * (4 * n) ** 3 + 16 * (2 * n) ** 3 + 8 * (2 * n) ** 3 + samples * (12 + 8 * ncoils + 8)


class _Group:
    """The in-flight problems of one image size N (coil counts may differ)."""

    def __init__(self, n: int, cmax: int, device, coil_batch: int, problem_batch: int, capacity: int = 4):
        self.n, self.cmax, self.device = n, cmax, device
        self.n3 = [n, n, n]
        self.nvox = n ** 3
        self.coil_batch = min(coil_batch, cmax)
        self.problem_batch = problem_batch
        self.reqs: List[ReconRequest] = []
        self.ncoil_host: List[int] = []
        self.iters_done: List[int] = []
        self.layout = None  # the kernels' layout, fixed by the first prepared kernel
        self.coils_dev = torch.arange(cmax, dtype=torch.int32, device=device)
        ws = C.lib().hasty_sense_normal_workspace_size(C.i64arr(self.n3), self.coil_batch, problem_batch)
        self.ws = torch.empty(ws, dtype=torch.uint8, device=device)
        self.cap = 0
        self._alloc(capacity)

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
            k = len(self.reqs)
            for new_t, old_t, rows in zip((self.smaps, self.kern, self.x, self.r, self.p, self.Ap, self.b, self.state,
                                           self.ncoil_dev), old, [k * self.cmax] + [k] * 8):
                new_t[:rows].copy_(old_t[:rows])
        self.cap = cap

    @property
    def active(self):
        return len(self.reqs)

    def add(self, req: ReconRequest, smaps: torch.Tensor, kernel_nat: torch.Tensor, b: torch.Tensor):
        if self.active == self.cap:
            self._alloc(2 * self.cap)
        i = self.active
        c = req.ncoils
        self.smaps[i * self.cmax:i * self.cmax + c].copy_(smaps)
        layout = ctypes.c_int()
        C.call("hasty_sense_normal_prepare", kernel_nat.data_ptr(), self.kern[i].data_ptr(), 1, C.i64arr(self.n3),
               ctypes.byref(layout), C.stream_ptr())
        if self.layout is None:
            self.layout = layout.value
        elif layout.value != self.layout:
            raise RuntimeError("kernel layout changed within a size group")
        self.b[i].copy_(b.reshape(self.n, self.n, self.n))
        self.x[i].zero_()
        self.Ap[i].zero_()
        self.ncoil_dev[i] = c
        C.call("hasty_cg_init", 1, self.nvox, self.b[i].data_ptr(), self.Ap[i].data_ptr(), self.r[i].data_ptr(),
               self.p[i].data_ptr(), self.state[i].data_ptr(), self.cgws.data_ptr(), C.stream_ptr())
        self.reqs.append(req)
        self.ncoil_host.append(c)
        self.iters_done.append(0)

    def step(self):
        P = self.active
        if P == 0:
            return
        C.call("hasty_sense_normal_toeplitz_ex", self.p.data_ptr(), self.Ap.data_ptr(), self.smaps.data_ptr(),
               self.cap * self.cmax, self.coils_dev.data_ptr(), self.cmax, self.cmax, C.i32arr(self.ncoil_host),
               self.ncoil_dev.data_ptr(), self.kern.data_ptr(), self.cap, self.layout, C.i32arr(list(range(P))),
               None, P, C.i64arr(self.n3), self.ws.data_ptr(), self.ws.numel(), self.coil_batch, C.stream_ptr())
        C.call("hasty_cg_iterate", P, self.nvox, self.x.data_ptr(), self.r.data_ptr(), self.p.data_ptr(),
               self.Ap.data_ptr(), self.state.data_ptr(), self.cgws.data_ptr(), 0.0, C.stream_ptr())
        for i in range(P):
            self.iters_done[i] += 1

    def pop_finished(self, now_fn) -> List[ReconRequest]:
        done = []
        for i in reversed(range(self.active)):
            req = self.reqs[i]
            if self.iters_done[i] < req.iters:
                continue
            req.image = self.x[i].cpu()  # waits for this problem's last step
            req.t_done = now_fn()
            done.append(req)
            last = self.active - 1
            if i != last:  # the last row takes the freed slot
                cm = self.cmax
                self.smaps[i * cm:(i + 1) * cm].copy_(self.smaps[last * cm:(last + 1) * cm])
                for t in (self.kern, self.x, self.r, self.p, self.Ap, self.b, self.state, self.ncoil_dev):
                    t[i].copy_(t[last])
                self.reqs[i], self.ncoil_host[i], self.iters_done[i] = (self.reqs[last], self.ncoil_host[last],
                                                                       self.iters_done[last])
            self.reqs.pop()
            self.ncoil_host.pop()
            self.iters_done.pop()
        return done


class ReconServer:
    def __init__(self, device=None, hbm_budget: Optional[int] = None, coil_batch: int = 16, problem_batch: int = 4,
                 max_active: int = 64, cmax: int = 32, nufft_tol: float = 1e-6, kernel_tol: float = 1e-7):
        self.device = torch.device(device or "cuda:0")
        free, total = torch.cuda.mem_get_info(self.device)
        self.hbm_budget = hbm_budget if hbm_budget is not None else int(0.8 * total)
        self.coil_batch, self.problem_batch, self.max_active, self.cmax = coil_batch, problem_batch, max_active, cmax
        self.nufft_tol, self.kernel_tol = nufft_tol, kernel_tol
        self.groups = {}
        self.admissions = []  # (rid, HBM in use, footprint, active before) per admission

    # -- request setup (on arrival: the scanner's data, the trajectory's kernel, A^H y) --
    def setup(self, req: ReconRequest):
        dev = self.device
        n, c = req.n, req.ncoils
        readout = req.readout or 2 * n
        g = torch.Generator(device=dev)
        g.manual_seed(1000003 * req.seed + 17)
        coords = synthetic.kooshball_torch(req.nspokes, readout, (req.seed * 7919) % 1000003, dev)
        smaps = synthetic.coil_maps_torch(c, n, dev, seed=req.seed)
        x_true = synthetic.ellipsoid_phantom_torch(n, dev, seed=req.seed)
        scale = 1.0 / math.sqrt(n ** 3)
        fwd = native.NufftPlan(coords, [n, n, n], 2, -1, 1, self.nufft_tol, scale)
        adj = native.NufftPlan(coords, [n, n, n], 1, +1, 1, self.nufft_tol, scale)
        M = coords.shape[1]
        y = torch.empty((1, M), dtype=torch.complex64, device=dev)
        im = torch.empty((1, n, n, n), dtype=torch.complex64, device=dev)
        b = torch.zeros((n, n, n), dtype=torch.complex64, device=dev)
        for ci in range(c):
            fwd.execute(y, (x_true * smaps[ci]).reshape(1, n, n, n).contiguous())
            y += req.noise * torch.randn(y.shape, dtype=torch.complex64, device=dev, generator=g)
            adj.execute(y, im)
            b += smaps[ci].conj() * im[0]
        fwd.close()
        adj.close()
        _, kern = native.toeplitz_build(coords, [n, n, n], tol=self.kernel_tol, diag=False)
        return smaps, kern.reshape(-1), b

    def footprint(self, req: ReconRequest) -> int:
        return resident_bytes(req.n, req.ncoils) + setup_bytes(req.n, req.ncoils, req.samples)

    def _group(self, n):
        if n not in self.groups:
            self.groups[n] = _Group(n, self.cmax, self.device, self.coil_batch, self.problem_batch)
        return self.groups[n]

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
                    ev = torch.cuda.Event()
                    ev.record(stream)
                outq.put((req, data, ev))
            except Exception as e:  # surfaced by the main loop
                outq.put((req, e, None))

    def run(self, requests: List[ReconRequest], realtime: bool = True, log=None,
            async_setup: Optional[bool] = None) -> dict:
        """Serve `requests`. realtime: arrivals follow the wall clock (latency in seconds);
        otherwise one arrival joins per tick (latency in ticks). async_setup (default: with
        realtime) runs request setup on a worker thread and stream."""
        for r in requests:
            if r.ncoils > self.cmax:
                raise RuntimeError(f"request {r.rid}: {r.ncoils} coils > server limit {self.cmax}")
        if async_setup is None:
            async_setup = realtime
        pending = sorted(requests, key=lambda r: (r.arrival, r.rid))
        queue: List[ReconRequest] = []
        finished: List[ReconRequest] = []
        t0 = time.perf_counter()
        clock = (lambda: time.perf_counter() - t0) if realtime else None
        vclock = [0.0]  # without real time: arrivals are taken in order, one per tick at most
        now = clock if realtime else (lambda: vclock[0])
        ticks, batch_sum, max_seen = 0, 0, 0
        in_setup = 0
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
                    queue.append(pending.pop(0))
                    if not realtime:
                        break
                if async_setup:  # finished setups join their groups
                    while True:
                        try:
                            req, data, ev = outq.get_nowait()
                        except queue_mod.Empty:
                            break
                        in_setup -= 1
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
                    used = torch.cuda.memory_allocated(self.device)
                    if used + self.footprint(req) > self.hbm_budget and active > 0:
                        break  # wait for memory to free up
                    queue.pop(0)
                    req.t_admit = now()
                    self.admissions.append((req.rid, used, self.footprint(req), active))
                    if async_setup:
                        inq.put(req)
                        in_setup += 1
                    else:
                        join_group(req, req.prepared if req.prepared is not None else self.setup(req))
                    active += 1
                    if log:
                        log(f"t={now():.2f}s admit {req.rid} N={req.n} C={req.ncoils} M={req.samples} "
                            f"active={active}")
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
                    for req in g.pop_finished(now):
                        finished.append(req)
                        if log:
                            log(f"t={now():.2f}s done {req.rid} latency={req.latency:.3f}s")
        finally:
            if async_setup:
                inq.put(None)
                worker.join()
        torch.cuda.synchronize(self.device)
        makespan = now() - min(r.arrival for r in requests) if realtime else float(ticks)
        lat = np.array([r.latency for r in finished])
        return {"requests": len(finished), "recons_per_s": len(finished) / makespan if realtime else None,
                "makespan_s": makespan, "latency_p50_s": float(np.percentile(lat, 50)),
                "latency_p99_s": float(np.percentile(lat, 99)), "latency_mean_s": float(lat.mean()),
                "ticks": ticks, "mean_batch": batch_sum / max(1, ticks), "max_active": max_seen,
                "peak_hbm_bytes": int(torch.cuda.max_memory_allocated(self.device))}


def poisson_requests(count: int, rate: float, seed: int = 0, sizes=(96, 128, 160), coils=(8, 32),
                     spokes=(1000, 10000), iters: int = 15) -> List[ReconRequest]:
    """A seeded Poisson stream: exponential inter-arrival times with mean 1/rate; matrix
    size, coil count and spoke count drawn uniformly from the given sets / ranges."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(count):
        t += float(rng.exponential(1.0 / rate))
        out.append(ReconRequest(rid=i, n=int(rng.choice(sizes)), ncoils=int(rng.integers(coils[0], coils[1] + 1)),
                                nspokes=int(rng.integers(spokes[0], spokes[1] + 1)), arrival=t, iters=iters,
                                seed=seed * 100003 + i))
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m hastycompute_amd.serving")
    ap.add_argument("--requests", type=int, default=200)
    ap.add_argument("--rate", type=float, default=2.0, help="mean arrivals per second")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--iters", type=int, default=15)
    ap.add_argument("--budget-gb", type=float, default=None)
    ap.add_argument("--verbose", action="store_true")
    a = ap.parse_args(argv)
    torch.cuda.set_device(0)
    reqs = poisson_requests(a.requests, a.rate, a.seed, iters=a.iters)
    srv = ReconServer(hbm_budget=None if a.budget_gb is None else int(a.budget_gb * 2 ** 30))
    stats = srv.run(reqs, realtime=True, log=print if a.verbose else None)
    stats.update({"rate": a.rate, "seed": a.seed, "iters": a.iters})
    print(json.dumps(stats), flush=True)


if __name__ == "__main__":
    main()
