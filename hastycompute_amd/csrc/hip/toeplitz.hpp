// Toeplitz-embedded normal operator: PSF build (FP64) and applies.
#pragma once

#include "common.hpp"

namespace hasty_hip {

// Build the doubled-grid PSF of a trajectory in FP64 and derive
//  - diag_ref: the reference's complex64 diagonal fftn(ifftshift(psf)) with psf scaled
//    by (2N)^-d (CUDANormalNufftToeplitz::build_diagonal, lib/fft/nufft.cpp:932-1053);
//  - kernel: the real float32 kernel used by the fused SENSE normal operator,
//    Re fftn(ifftshift(psf')) / ((2N)^d N^d), where psf' zeroes the k = -N planes that
//    never reach the cropped output (so the kernel is real); it already carries the
//    2^d that turns the reference's Toeplitz scale into SenseNormal's.
// coords: [ndim][M] (float or double), nmodes: image dims (slowest first), weights: [M]
// real or null. Either output may be null.
void toeplitz_build(const void* coords, int coord_fp64, int64_t M, int ndim, const int64_t* nmodes,
                    const float* weights, float2* diag_ref, float* kernel, double tol, hipStream_t s);

// CUDANormalNufftToeplitz::apply* (nufft.cpp:1163-1197) on a complex diagonal:
//   res = ifftn(fftn(in, s=2N) * diag)[:N]; mode 0: out = res, 1: out += res,
//   2: out += res * mul, 3: out = res with out == in allowed.
// in/out: [ntransf][N...]; diag: [ntransf][2N...]; storage: >= ntransf*N0*(2N1)*(2N2)
// complex each (the reference's storage1/storage2 are larger and suffice).
void toeplitz_apply_ref(const float2* in, float2* out, const float2* mul, const float2* diag, int ntransf, int ndim,
                        const int64_t* nmodes, int mode, float2* storage1, float2* storage2, hipStream_t s);

// Fused Toeplitz SENSE normal operator over a batch of problems:
//   out[p] = sum_{c in coils} conj(S_c) * crop(IFFT(K[kidx[p]] * FFT(pad(S_c * x[p])))) - (sub ? sub[p] : 0)
// x, out, sub: [P][N0][N1][N2]; smaps: [C][N0][N1][N2]; kernels: [nk][2N0][2N1][2N2] real.
// coils: device int array of ncoil coil indices. Workspace from
// sense_normal_workspace_bytes.
// Kernel layouts: 0 natural [kz][ky][kx] (generic five-pass path), 1 the fused path's
// coset order [kz][s][ky][m] (kx = P m + s, P = 2 N2 / 32) for the LDS and coset plane
// kernels, 2 the fused path's register-kernel order at N1 = N2 = 128. A prepared kernel
// keeps its layout, and the apply picks the plane kernels from it, so changing the
// sense_xy_variant option later cannot mismatch kernels prepared before.
int sense_kernel_layout(const int64_t* n3, bool allow_fused);
void sense_prepare_kernels(const float* natural, float* prepared, int64_t nk, const int64_t* n3, int layout,
                           hipStream_t s);
size_t sense_normal_workspace_bytes(const int64_t* n3, int coil_batch, int problem_batch);
// Problem p uses coil-map rows coils[i] + p * coil_pstride (0: all problems share them),
// for i < pcoils[p] when the per-problem counts are given (host copy for the host-driven
// generic path, device copy for the fused kernels), else for i < ncoil.
void sense_normal_toeplitz(const float2* x, float2* out, const float2* smaps, const int* coils_dev, int ncoil,
                           const float* kernels, int layout, const int* kidx_host, const float2* sub, int P,
                           const int64_t* n3, void* workspace, size_t ws_bytes, int coil_batch, hipStream_t s,
                           int coil_pstride = 0, const int* pcoils_host = nullptr, const int* pcoils_dev = nullptr);

// dct.hip
void dct3(float2* data, int64_t batch, const int64_t* dims, int axis_mask, int inverse, hipStream_t s);
void l1dct_prox(float2* data, int64_t batch, const int64_t* dims, int axis_mask, float thresh, hipStream_t s);

// precond.hip
void circulant_precond_build(const float2* smaps, int64_t ncoil, const void* coords, int coord_fp64, int64_t M,
                             int ndim, const int64_t* nmodes, const float* weights, double lambda, int centered,
                             float2* diag, double tol, hipStream_t s);
size_t circulant_precond_workspace_bytes(int64_t ndiag, int ndim, const int64_t* nmodes, int centered, int adjoint);
void circulant_precond_apply(const float2* in, float2* out, const float2* diag, int64_t ndiag, int64_t P,
                             const int32_t* dindex, int ndim, const int64_t* nmodes, int centered, int adjoint,
                             void* ws, size_t ws_bytes, hipStream_t s);

// sense_fused.hip
bool sense_fused_supported(const int64_t* n3);
// 1 (default): plane kernels with register four-step y transforms; 0: LDS Stockham passes.
void sense_fused_set_xy_variant(int v);
// problems per z-convolution launch of the fused operator at N0 = 128 (1..16)
void sense_fused_set_zconv_batch(int v);
int sense_fused_get_xy_variant();
int sense_fused_get_zconv_batch();
// z-convolution kernel at N0 = 128 (0..3, see sense_fused.hip)
void sense_fused_set_zconv_variant(int v);
int sense_fused_get_zconv_variant();
// true when the current sense_xy_variant gives the register kernels at this size
bool sense_fused_register_layout(const int64_t* n3);
void sense_fused_permute_kernel(const float* natural, float* permuted, const int64_t* n3, bool r16, hipStream_t s);
size_t sense_fused_workspace_bytes(const int64_t* n3, int pchunk, int ccount);
void sense_fused_apply(const float2* x, float2* out, const float2* smaps, const int* coils_dev, int ncoil,
                       const float* kernels, bool r16, const int* kidx, const float2* sub, int P, const int64_t* n3, void* ws,
                       size_t ws_bytes, int pchunk, int cchunk, hipStream_t s, int coil_pstride = 0,
                       const int* pcoils_dev = nullptr);

}  // namespace hasty_hip
