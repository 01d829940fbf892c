"""Torch-tensor wrappers over the C ABI (include/hasty_hip.h).

These are the thinnest typed layer over libhasty_hip.so: tensors must already be on
the GPU with the documented dtype and layout; results are written into caller- or
wrapper-allocated tensors on the current stream. The reference-named operator classes
(hasty::fft::CUDANufft, ...) are built on the same ABI in the C++ layer
(csrc/host) and exposed as torch.classes.Hasty*; these wrappers serve tests, the
batched driver and the benchmark.
"""
from __future__ import annotations

import ctypes
from typing import Optional, Sequence

import torch

from . import _capi as C


def _require(cond: bool, msg: str):
    if not cond:
        raise RuntimeError(msg)


def _cuda(t: torch.Tensor, name: str):
    _require(t.is_cuda, f"{name} must reside on a GPU")
    _require(t.is_contiguous(), f"{name} must be contiguous")


class NufftPlan:
    """A type-1 or type-2 NUFFT plan on the points `coords` ([ndim, M], float32 or
    float64 on the GPU; row d pairs with image dim d). `nmodes` are the image dims
    (slowest first). Scale multiplies every output (1/sqrt(prod nmodes) reproduces
    hasty::fft::CUDANufft::apply, lib/fft/nufft.cpp:487-506)."""

    def __init__(self, coords: torch.Tensor, nmodes: Sequence[int], nufft_type: int, isign: int,
                 ntransf: int = 1, tol: float = 1e-6, scale: float = 1.0):
        _cuda(coords, "coords")
        _require(coords.dtype in (torch.float32, torch.float64), "coords must be real")
        self.coords = coords  # borrowed by the plan
        self.ndim, self.M = int(coords.shape[0]), int(coords.shape[1])
        _require(len(nmodes) == self.ndim, "nmodes must match coords.size(0)")
        self.nmodes = [int(n) for n in nmodes]
        self.type = nufft_type
        self.ntransf = ntransf
        self.fp64 = coords.dtype == torch.float64
        self.cdtype = torch.complex128 if self.fp64 else torch.complex64
        h = ctypes.c_void_p()
        C.call("hasty_nufft_makeplan", nufft_type, self.ndim, C.i64arr(list(reversed(self.nmodes))), isign, ntransf,
               float(tol), int(self.fp64), ctypes.byref(h))
        self._h = h
        C.call("hasty_nufft_set_scale", h, float(scale))
        es = coords.element_size()
        base = coords.data_ptr()
        rows = [base + d * self.M * es for d in range(self.ndim)]
        xyz = [rows[self.ndim - 1 - i] if i < self.ndim else None for i in range(3)]
        C.call("hasty_nufft_setpts", h, self.M, xyz[0], xyz[1], xyz[2], C.stream_ptr())

    def info(self):
        nf = C.i64arr([0, 0, 0])
        ns = ctypes.c_int()
        beta = ctypes.c_double()
        C.call("hasty_nufft_info", self._h, nf, ctypes.byref(ns), ctypes.byref(beta))
        return list(reversed([nf[i] for i in range(self.ndim)])), ns.value, beta.value

    def execute(self, c: torch.Tensor, f: torch.Tensor):
        _cuda(c, "c")
        _cuda(f, "f")
        _require(c.dtype == self.cdtype and f.dtype == self.cdtype, "data dtype does not match the plan")
        _require(c.numel() == self.ntransf * self.M, "c must hold ntransf x M values")
        _require(f.numel() == self.ntransf * int(torch.tensor(self.nmodes).prod()), "f must hold ntransf x modes")
        C.call("hasty_nufft_execute", self._h, c.data_ptr(), f.data_ptr(), C.stream_ptr())

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            C.call("hasty_nufft_destroy", self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def nufft(coords: torch.Tensor, data: torch.Tensor, nmodes: Sequence[int], nufft_type: int, isign: int,
          tol: float = 1e-6, scale: float = 1.0) -> torch.Tensor:
    """One-shot NUFFT: type 1 maps [ntransf, M] -> [ntransf, *nmodes], type 2 the reverse."""
    ntransf = data.shape[0]
    plan = NufftPlan(coords, nmodes, nufft_type, isign, ntransf, tol, scale)
    if nufft_type == 1:
        out = torch.empty((ntransf, *nmodes), dtype=plan.cdtype, device=data.device)
        plan.execute(data, out)
    else:
        out = torch.empty((ntransf, plan.M), dtype=plan.cdtype, device=data.device)
        plan.execute(out, data)
    plan.close()
    return out


def toeplitz_build(coords: torch.Tensor, nmodes: Sequence[int], weights: Optional[torch.Tensor] = None,
                   tol: float = 1e-10, diag: bool = True, kernel: bool = True):
    """(diag_ref complex64 [2N...], kernel float32 [2N...]) of a trajectory; see
    hasty_toeplitz_build."""
    _cuda(coords, "coords")
    L = [2 * int(n) for n in nmodes]
    d = torch.empty(L, dtype=torch.complex64, device=coords.device) if diag else None
    k = torch.empty(L, dtype=torch.float32, device=coords.device) if kernel else None
    if weights is not None:
        _cuda(weights, "weights")
        _require(weights.dtype == torch.float32, "weights must be float32")
    C.call("hasty_toeplitz_build", coords.data_ptr(), int(coords.dtype == torch.float64), int(coords.shape[1]),
           len(nmodes), C.i64arr(nmodes), C.ptr(weights), C.ptr(d), C.ptr(k), float(tol), C.stream_ptr())
    return d, k


def toeplitz_apply(diag: torch.Tensor, x: torch.Tensor, mode: int = 0, out: Optional[torch.Tensor] = None,
                   mul: Optional[torch.Tensor] = None, storage1=None, storage2=None) -> torch.Tensor:
    """CUDANormalNufftToeplitz::apply family on [ntransf, N...] complex64 data."""
    nt, nm = x.shape[0], list(x.shape[1:])
    if out is None:
        out = x if mode == 3 else torch.empty_like(x)
    need = nt
    for n in nm:
        need *= 2 * n
    if storage1 is None:
        storage1 = torch.empty(need, dtype=torch.complex64, device=x.device)
    if storage2 is None:
        storage2 = torch.empty_like(storage1)
    C.call("hasty_toeplitz_apply", x.data_ptr(), out.data_ptr(), C.ptr(mul), diag.data_ptr(), nt, len(nm),
           C.i64arr(nm), mode, storage1.data_ptr(), storage2.data_ptr(), C.stream_ptr())
    return out


class SenseNormalToeplitz:
    """Batched Toeplitz SENSE normal operator (hasty_sense_normal_toeplitz):
    out[p] = sum_c conj(S_c) T_{k(p)}(S_c x[p]) - sub[p]."""

    def __init__(self, smaps: torch.Tensor, kernels: torch.Tensor, coils: Optional[Sequence[int]] = None,
                 coil_batch: int = 4, problem_batch: int = 1, fused: Optional[bool] = None):
        _cuda(smaps, "smaps")
        _cuda(kernels, "kernels")
        _require(smaps.dtype == torch.complex64 and smaps.dim() == 4, "smaps must be complex64 [C, N0, N1, N2]")
        if kernels.dim() == 3:
            kernels = kernels.unsqueeze(0)
        self.smaps = smaps
        self.n3 = list(smaps.shape[1:])
        _require(list(kernels.shape[1:]) == [2 * n for n in self.n3], "kernel must be [nk, 2N0, 2N1, 2N2]")
        coils = list(range(smaps.shape[0])) if coils is None else [int(c) for c in coils]
        _require(len(coils) >= 1 and all(0 <= c < smaps.shape[0] for c in coils), "coil index out of range")
        self.coils = torch.tensor(coils, dtype=torch.int32, device=smaps.device)
        self.ncoil = len(coils)
        self.coil_batch = max(1, min(coil_batch, self.ncoil))
        _require(kernels.dtype == torch.float32, "kernels must be float32")
        self.kernels = torch.empty_like(kernels)
        layout = ctypes.c_int()
        prev = C.get_option("sense_fused")
        if fused is not None:
            C.set_option("sense_fused", int(bool(fused)))
        try:
            C.call("hasty_sense_normal_prepare", kernels.data_ptr(), self.kernels.data_ptr(), kernels.shape[0],
                   C.i64arr(self.n3), ctypes.byref(layout), C.stream_ptr())
        finally:
            C.set_option("sense_fused", prev)
        self.layout = layout.value
        self.nvox = self.n3[0] * self.n3[1] * self.n3[2]
        self.problem_batch = max(1, problem_batch)
        ws = C.lib().hasty_sense_normal_workspace_size(C.i64arr(self.n3), self.coil_batch, self.problem_batch)
        self.workspace = torch.empty(ws, dtype=torch.uint8, device=smaps.device)
        self._ncoil_dev = {}

    def _check_image(self, t: torch.Tensor, name: str, P: int):
        _cuda(t, name)
        _require(t.dtype == torch.complex64, f"{name} must be complex64")
        _require(t.numel() == P * self.nvox, f"{name} must hold {P} images of {self.n3}")

    def apply(self, x: torch.Tensor, out: Optional[torch.Tensor] = None, kernel_index: Optional[Sequence[int]] = None,
              sub: Optional[torch.Tensor] = None, problem_ncoil: Optional[Sequence[int]] = None,
              coil_problem_stride: int = 0) -> torch.Tensor:
        """out[p] = sum_{i < ncoil_p} conj(S_{c_i + p s}) T_{k(p)}(S_{c_i + p s} x[p]) - sub[p] with
        c_i the operator's coil list, s = coil_problem_stride and ncoil_p = problem_ncoil[p]
        (default: every listed coil); hasty_sense_normal_toeplitz(_ex)."""
        _cuda(x, "x")
        _require(x.dtype == torch.complex64, "x must be complex64")
        _require(x.numel() % self.nvox == 0 and x.numel() > 0, f"x must hold whole images of {self.n3}")
        P = x.numel() // self.nvox
        if out is None:
            out = torch.empty_like(x)
        self._check_image(out, "out", P)
        if sub is not None:
            self._check_image(sub, "sub", P)
        kidx = [0] * P if kernel_index is None else [int(k) for k in kernel_index]
        _require(len(kidx) == P and all(0 <= k < self.kernels.shape[0] for k in kidx),
                 "kernel_index must give a valid kernel for every problem")
        if problem_ncoil is None and coil_problem_stride == 0:
            C.call("hasty_sense_normal_toeplitz", x.data_ptr(), out.data_ptr(), self.smaps.data_ptr(),
                   self.coils.data_ptr(), self.ncoil, self.kernels.data_ptr(), self.kernels.shape[0], self.layout,
                   C.i32arr(kidx), C.ptr(sub), P, C.i64arr(self.n3), self.workspace.data_ptr(),
                   self.workspace.numel(), self.coil_batch, C.stream_ptr())
            return out
        nc = [self.ncoil] * P if problem_ncoil is None else [int(c) for c in problem_ncoil]
        _require(len(nc) == P and all(1 <= c <= self.ncoil for c in nc), "problem_ncoil must be in 1..ncoil")
        key = tuple(nc)
        nc_dev = self._ncoil_dev.get(key)
        if nc_dev is None:  # kept, so repeated calls (and graph captures) allocate nothing
            nc_dev = self._ncoil_dev[key] = torch.tensor(nc, dtype=torch.int32, device=x.device)
        C.call("hasty_sense_normal_toeplitz_ex", x.data_ptr(), out.data_ptr(), self.smaps.data_ptr(),
               int(self.smaps.shape[0]), self.coils.data_ptr(), self.ncoil, int(coil_problem_stride), C.i32arr(nc),
               nc_dev.data_ptr(), self.kernels.data_ptr(), self.kernels.shape[0], self.layout, C.i32arr(kidx),
               C.ptr(sub), P, C.i64arr(self.n3), self.workspace.data_ptr(), self.workspace.numel(), self.coil_batch,
               C.stream_ptr())
        return out


class DeviceCG:
    """Batched device CG (hasty_cg_*) over `nproblems` vectors of n complex64 values."""

    def __init__(self, nproblems: int, n: int, device):
        self.P, self.n = nproblems, n
        self.state = torch.zeros((nproblems, 8), dtype=torch.float64, device=device)
        ws = C.lib().hasty_cg_workspace_size(nproblems, n)
        self.ws = torch.empty(ws, dtype=torch.uint8, device=device)

    def init(self, b, Ax, r, p):
        C.call("hasty_cg_init", self.P, self.n, b.data_ptr(), Ax.data_ptr(), r.data_ptr(), p.data_ptr(),
               self.state.data_ptr(), self.ws.data_ptr(), C.stream_ptr())

    def iterate(self, x, r, p, Ap, tol: float = 0.0, l2: float = 0.0):
        """One CG step given Ap = A p; with l2 > 0 the step is for A + l2 I
        (hasty_cg_iterate_l2)."""
        if l2:
            C.call("hasty_cg_iterate_l2", self.P, self.n, x.data_ptr(), r.data_ptr(), p.data_ptr(), Ap.data_ptr(),
                   self.state.data_ptr(), self.ws.data_ptr(), float(tol), float(l2), C.stream_ptr())
            return
        C.call("hasty_cg_iterate", self.P, self.n, x.data_ptr(), r.data_ptr(), p.data_ptr(), Ap.data_ptr(),
               self.state.data_ptr(), self.ws.data_ptr(), float(tol), C.stream_ptr())

    # preconditioned loop: init, z = P r, init_precond; per iteration iterate_xr, z = P r, iterate_pz
    def init_precond(self, r, z, p):
        C.call("hasty_cg_init_precond", self.P, self.n, r.data_ptr(), z.data_ptr(), p.data_ptr(),
               self.state.data_ptr(), self.ws.data_ptr(), C.stream_ptr())

    def iterate_xr(self, x, r, p, Ap, tol: float = 0.0):
        C.call("hasty_cg_iterate_xr", self.P, self.n, x.data_ptr(), r.data_ptr(), p.data_ptr(), Ap.data_ptr(),
               self.state.data_ptr(), self.ws.data_ptr(), float(tol), C.stream_ptr())

    def iterate_pz(self, r, z, p):
        C.call("hasty_cg_iterate_pz", self.P, self.n, r.data_ptr(), z.data_ptr(), p.data_ptr(),
               self.state.data_ptr(), self.ws.data_ptr(), C.stream_ptr())


class Graph:
    """A recorded sequence of library launches (hasty_graph_*): `record()` runs under a
    capture of `stream` (a created stream; it becomes the current stream meanwhile), and
    `replay()` launches the recorded graph on it. What `record` calls must not allocate
    or synchronise: run it once eagerly on the same buffers first."""

    def __init__(self, record, stream: Optional[torch.cuda.Stream] = None):
        self.stream = stream if stream is not None else torch.cuda.Stream()
        sp = self.stream.cuda_stream
        self._h = None
        with torch.cuda.stream(self.stream):
            C.call("hasty_graph_begin", sp)
            h = ctypes.c_void_p()
            try:
                record()
            except BaseException:
                if C.lib().hasty_graph_end(sp, ctypes.byref(h)) == 0 and h.value:
                    C.lib().hasty_graph_destroy(h)
                raise
            C.call("hasty_graph_end", sp, ctypes.byref(h))
        self._h = h

    @property
    def nodes(self) -> int:
        return int(C.lib().hasty_graph_nodes(self._h))

    def replay(self):
        C.call("hasty_graph_launch", self._h, self.stream.cuda_stream)

    def close(self):
        if self._h is not None and self._h.value:
            C.call("hasty_graph_destroy", self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def cg_solve(apply_A, b: torch.Tensor, x0: torch.Tensor, iters: int, tol: float = 0.0, check_pd: bool = True,
             apply_P=None, graph: bool = False, l2: float = 0.0):
    """ConjugateGradient::run (lib/op/opalgs.cpp:50-96) for a batch [P, ...] with the
    device CG; apply_A(v, out) writes A v into out; apply_P(v, out), if given, writes the
    preconditioned P v (the _P branch). With graph, the first iteration runs eagerly and
    the later ones replay a HIP graph of one iteration (apply_A + the CG step), so
    apply_A must not allocate. l2 > 0 solves with A + l2 I (l2_reg_linop,
    python/python/hastypy/base/opalg.py:626-627) without a pass for l2 p. A start x0 of
    all zeros skips the first apply (A 0 is exactly 0, so r = b to the bit). Returns
    (x, state)."""
    _require(l2 >= 0.0, "l2 must be >= 0")
    _require(not (l2 and apply_P is not None), "l2 covers the unpreconditioned loop")
    P = b.shape[0]
    n = b[0].numel()
    x = x0.clone()
    r = torch.empty_like(b)
    p = torch.empty_like(b)
    Ap = torch.empty_like(b)
    cg = DeviceCG(P, n, b.device)
    if iters < 1:
        return x, cg.state
    zero_start = int(torch.count_nonzero(torch.view_as_real(x0) if x0.is_complex() else x0)) == 0

    def residual_start():  # Ap = A x0 (+ l2 x0) for r = b - A x0
        if zero_start:
            Ap.zero_()
            return
        apply_A(x, Ap)
        if l2:
            Ap.add_(x, alpha=l2)

    if graph:
        _require(apply_P is None, "graph replay covers the unpreconditioned loop")
        main = torch.cuda.current_stream()
        side = torch.cuda.Stream()
        side.wait_stream(main)
        with torch.cuda.stream(side):
            residual_start()
            cg.init(b, Ap, r, p)
            apply_A(p, Ap)
            cg.iterate(x, r, p, Ap, tol, l2)

            def one_iteration():
                apply_A(p, Ap)
                cg.iterate(x, r, p, Ap, tol, l2)

            if iters > 1:
                g = Graph(one_iteration, side)
                for _ in range(iters - 1):
                    g.replay()
        main.wait_stream(side)
        if iters > 1:
            side.synchronize()
            g.close()
    elif apply_P is None:
        residual_start()
        cg.init(b, Ap, r, p)
        for _ in range(iters):
            apply_A(p, Ap)
            cg.iterate(x, r, p, Ap, tol, l2)
    else:
        residual_start()
        cg.init(b, Ap, r, p)
        z = torch.empty_like(b)
        apply_P(r, z)
        cg.init_precond(r, z, p)
        for _ in range(iters):
            apply_A(p, Ap)
            cg.iterate_xr(x, r, p, Ap, tol)
            apply_P(r, z)
            cg.iterate_pz(r, z, p)
    if check_pd:
        st = cg.state.cpu()
        if bool((st[:, 4] != 0).any()):
            raise RuntimeError("A was not positive definite")
    return x, cg.state


def circulant_precond_diagonal(smaps: torch.Tensor, coords: torch.Tensor, weights: Optional[torch.Tensor] = None,
                               lamda: float = 0.0, centered: bool = True, tol: float = 1e-10) -> torch.Tensor:
    """CirculantPreconditionerOp::build_diagonal (lib/mri/precond.cpp:65-117): complex64
    [N...] diagonal, centred order (the reference's) or FFT order; see
    hasty_circulant_precond_build."""
    _cuda(smaps, "smaps")
    _cuda(coords, "coords")
    _require(smaps.dtype == torch.complex64 and smaps.is_contiguous(), "smaps must be contiguous complex64")
    nm = list(smaps.shape[1:])
    if weights is not None:
        _cuda(weights, "weights")
        _require(weights.dtype == torch.float32, "weights must be float32")
        weights = weights.contiguous()
    coords = coords.contiguous()
    d = torch.empty(nm, dtype=torch.complex64, device=smaps.device)
    C.call("hasty_circulant_precond_build", smaps.data_ptr(), int(smaps.shape[0]), coords.data_ptr(),
           int(coords.dtype == torch.float64), int(coords.shape[1]), len(nm), C.i64arr(nm), C.ptr(weights),
           float(lamda), int(bool(centered)), d.data_ptr(), float(tol), C.stream_ptr())
    return d


def circulant_precond_apply(x: torch.Tensor, diag: torch.Tensor, out: Optional[torch.Tensor] = None,
                            diag_index: Optional[Sequence[int]] = None, centered: bool = True,
                            adjoint: bool = False) -> torch.Tensor:
    """CirculantPreconditionerOp::apply (precond.cpp:27-50) on x [P, N...] with diag [N...]
    or [nd, N...] (diag_index[p] picks problem p's diagonal; default p)."""
    _cuda(x, "x")
    _require(x.dtype == torch.complex64 and x.is_contiguous(), "x must be contiguous complex64")
    nm = list(x.shape[1:])
    d = diag if diag.dim() == len(nm) + 1 else diag.unsqueeze(0)
    _require(list(d.shape[1:]) == nm, "diag shape must match the image shape")
    d = d.contiguous()
    nd = int(d.shape[0])
    if out is None:
        out = torch.empty_like(x)
    idx = None
    if diag_index is not None:
        idx = (ctypes.c_int32 * len(diag_index))(*[int(i) for i in diag_index])
    elif nd == 1 and x.shape[0] > 1:
        idx = (ctypes.c_int32 * x.shape[0])(*([0] * x.shape[0]))
    wsb = C.lib().hasty_circulant_precond_workspace_size(nd, len(nm), C.i64arr(nm), int(bool(centered)),
                                                          int(bool(adjoint)))
    ws = torch.empty(max(int(wsb), 1), dtype=torch.uint8, device=x.device)
    C.call("hasty_circulant_precond_apply", x.data_ptr(), out.data_ptr(), d.data_ptr(), nd, int(x.shape[0]), idx,
           len(nm), C.i64arr(nm), int(bool(centered)), int(bool(adjoint)), ws.data_ptr(), int(wsb), C.stream_ptr())
    return out


def _vol_view(x: torch.Tensor, naxes: int):
    _cuda(x, "x")
    _require(x.dtype == torch.complex64 and x.is_contiguous(), "x must be contiguous complex64")
    _require(1 <= naxes <= 3 and x.dim() >= naxes, "1 to 3 transform axes")
    dims = [1] * (3 - naxes) + list(x.shape[x.dim() - naxes:])
    batch = x.numel() // max(1, int(torch.tensor(dims).prod().item()))
    return batch, dims


def dctn_(x: torch.Tensor, axes: Sequence[bool], inverse: bool = False) -> torch.Tensor:
    """In-place unnormalised DCT-II (or its inverse) over the last len(axes) dims of x
    where axes[i] is True (scipy.fft.dctn / idctn, norm=None); see hasty_dct3."""
    batch, dims = _vol_view(x, len(axes))
    mask = sum(1 << (3 - len(axes) + i) for i, a in enumerate(axes) if a)
    C.call("hasty_dct3", x.data_ptr(), batch, C.i64arr(dims), mask, int(bool(inverse)), C.stream_ptr())
    return x


def l1dct_prox_(x: torch.Tensor, axes: Sequence[bool], thresh: float) -> torch.Tensor:
    """ProximalL1DCT.tensor_apply in place: dctn, complex soft threshold, idctn."""
    batch, dims = _vol_view(x, len(axes))
    mask = sum(1 << (3 - len(axes) + i) for i, a in enumerate(axes) if a)
    C.call("hasty_l1dct_prox", x.data_ptr(), batch, C.i64arr(dims), mask, float(thresh), C.stream_ptr())
    return x


def llr_shifts(block_iter: int, block_shape: Sequence[int]):
    out = C.i32arr([0] * (3 * block_iter))
    C.call("hasty_llr_shifts", block_iter, C.i64arr(block_shape), out)
    return [[out[d * block_iter + b] for b in range(block_iter)] for d in range(3)]


def llr_normal_blocks(x: torch.Tensor, block_strides, block_shape, block_iter: int, thresh: float, soft: bool,
                      shifts=None, workspace: Optional[torch.Tensor] = None) -> torch.Tensor:
    """OldNormal3DBlocksSVT::apply on a [F, E, X, Y, Z] complex64 GPU tensor, in place."""
    _cuda(x, "input")
    _require(x.dtype == torch.complex64 and x.dim() == 5, "input must be complex64 [F, E, X, Y, Z]")
    dims = list(x.shape)
    ws_bytes = C.lib().hasty_llr_workspace_size(C.i64arr(dims), C.i64arr(block_strides), C.i64arr(block_shape))
    if workspace is None or workspace.numel() < ws_bytes:
        workspace = torch.empty(ws_bytes, dtype=torch.uint8, device=x.device)
    sh = None
    if shifts is not None:
        sh = C.i32arr([int(v) for row in shifts for v in row])
    C.call("hasty_llr_normal_blocks", x.data_ptr(), C.i64arr(dims), C.i64arr(block_strides), C.i64arr(block_shape),
           block_iter, sh, float(thresh), int(bool(soft)), workspace.data_ptr(), workspace.numel(), C.stream_ptr())
    return x


def llr_blocks(x: torch.Tensor, block_shape, corners, thresh: float, soft: bool) -> torch.Tensor:
    _cuda(x, "input")
    dims = list(x.shape)
    n = len(corners)
    ws_bytes = C.lib().hasty_llr_blocks_workspace_size(C.i64arr(dims), C.i64arr(block_shape), n)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=x.device)
    flat = C.i32arr([int(v) for c in corners for v in c])
    C.call("hasty_llr_blocks", x.data_ptr(), C.i64arr(dims), C.i64arr(block_shape), flat, n, float(thresh),
           int(bool(soft)), ws.data_ptr(), ws.numel(), C.stream_ptr())
    return x


def llr_grid_corners(dims, block_strides, block_shape, shift) -> list:
    """Corners [[x, y, z], ...] of the blocks one pass of llr_normal_blocks visits for one
    shift (the reference's strided grid and skip rule, svt.cpp:378-404); host only."""
    count = ctypes.c_int64(0)
    args = (C.i64arr(list(dims)), C.i64arr(list(block_strides)), C.i64arr(list(block_shape)),
            C.i32arr([int(v) for v in shift]))
    C.call("hasty_llr_grid_corners", *args, None, 0, ctypes.byref(count))
    out = C.i32arr([0] * (3 * max(1, count.value)))
    C.call("hasty_llr_grid_corners", *args, out, count.value, ctypes.byref(count))
    return [[out[3 * b + d] for d in range(3)] for b in range(count.value)]


def llr_blocks_workspace(x: torch.Tensor, block_shape, nblk: int) -> torch.Tensor:
    """Workspace for llr_blocks_dev on x with up to nblk blocks."""
    n = C.lib().hasty_llr_blocks_workspace_size(C.i64arr(list(x.shape)), C.i64arr(list(block_shape)), nblk)
    return torch.empty(n, dtype=torch.uint8, device=x.device)


def llr_blocks_dev(x: torch.Tensor, block_shape, corners: torch.Tensor, thresh: float, soft: bool,
                   workspace: torch.Tensor) -> torch.Tensor:
    """SVT of the disjoint blocks whose corners (int32 [nblk, 3]) are already on the GPU,
    in place on x [F, E, X, Y, Z]: no host copy and no synchronisation, so the call can
    be recorded in a Graph. Blocks outside the image are skipped."""
    _cuda(x, "input")
    _cuda(corners, "corners")
    _require(x.dtype == torch.complex64 and x.dim() == 5 and x.is_contiguous(),
             "input must be contiguous complex64 [F, E, X, Y, Z]")
    _require(corners.dtype == torch.int32 and corners.dim() == 2 and corners.shape[1] == 3 and corners.is_contiguous(),
             "corners must be contiguous int32 [nblk, 3]")
    C.call("hasty_llr_blocks_dev", x.data_ptr(), C.i64arr(list(x.shape)), C.i64arr(list(block_shape)),
           corners.data_ptr(), corners.shape[0], float(thresh), int(bool(soft)), workspace.data_ptr(),
           workspace.numel(), C.stream_ptr())
    return x


def svt_matrix(A: torch.Tensor, thresh: float, soft: bool) -> torch.Tensor:
    """svt_soft / svt_hard on a contiguous [R, F] complex64 GPU matrix, in place."""
    _cuda(A, "A")
    _require(A.dtype == torch.complex64 and A.dim() == 2, "A must be complex64 [R, F]")
    R, F = A.shape
    ws = torch.empty(C.lib().hasty_svt_workspace_size(R, F), dtype=torch.uint8, device=A.device)
    C.call("hasty_svt_matrix", A.data_ptr(), R, F, float(thresh), int(bool(soft)), ws.data_ptr(), ws.numel(),
           C.stream_ptr())
    return A


def timing_enable(on: bool = True) -> None:
    """Start (and clear) or stop per-kernel event timing inside the library."""
    C.call("hasty_timing_enable", 1 if on else 0)


def timing_read() -> dict:
    """{kernel name: (total_ms, launches)} of the launches timed since timing_enable."""
    L = C.lib()
    n = L.hasty_timing_entries()
    if n < 0:
        C.check(1)
    out = {}
    buf = ctypes.create_string_buffer(128)
    for i in range(n):
        ms = ctypes.c_double()
        cnt = ctypes.c_int64()
        C.call("hasty_timing_read", i, buf, 128, ctypes.byref(ms), ctypes.byref(cnt))
        out[buf.value.decode()] = (ms.value, cnt.value)
    return out
