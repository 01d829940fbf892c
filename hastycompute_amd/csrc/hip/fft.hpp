// Batched, strided, pruned 1-D FFT passes along one axis of a multi-dimensional array.
//
// A "line" is one 1-D transform. Line l starts at
//     base(l) = map_or_id(l / inner) * outer + (l % inner) * unit
// and its elements are `stride` apart. This covers every axis of a row-major array,
// batches of arrays (outer = batch stride) and indirect batches (map, e.g. a coil list).
//
// Zero padding and cropping are folded into the loads and stores: a line of length L
// reads n_in <= L input elements and writes n_out <= L outputs. With `centered` the
// element n stands for frequency k = n - n/2 and lives at position k mod L (the
// FINUFFT modeord=0 layout and numpy's ifftshift for n == L).
#pragma once

#include "common.hpp"

namespace hasty_hip {

struct LineAddr {
    int64_t inner = 1;
    int64_t outer = 0;
    int64_t unit = 0;
    int64_t stride = 1;
    const int* map = nullptr;
};

template <typename R> struct AxisFft {
    const cplx<R>* in = nullptr;
    LineAddr ain;
    const cplx<R>* inmul = nullptr;  // optional per-element input multiplier
    LineAddr amul;
    int inmul_conj = 0;
    cplx<R>* out = nullptr;
    LineAddr aout;
    const cplx<R>* outmul = nullptr;  // optional per-element output multiplier
    LineAddr aomul;
    int outmul_conj = 0;
    int accumulate = 0;                // out += result
    const R* pre_scale = nullptr;      // length n_in, indexed by input element
    const R* post_scale = nullptr;     // length n_out, indexed by output element
    R scale = 1;
    int L = 0, n_in = 0, n_out = 0, in_centered = 0, out_centered = 0;
    int sign = -1;  // exp(sign * 2 pi i j k / L)
    int64_t nlines = 0;
    // Convolution mode: forward FFT, multiply by a kernel, inverse FFT.
    const R* conv_real = nullptr;
    const cplx<R>* conv_cplx = nullptr;
    LineAddr aconv;
};

template <typename R> void run_axis_fft(const AxisFft<R>& a, hipStream_t stream);

bool fft_size_supported(int L);
// 1 (default): complex64 lines of length 128, 192, 256 or 320 without coil multipliers,
// accumulation or convolution go through the register 16 x R four-step kernel; 0: the
// LDS Stockham one.
void fft_set_r16(int v);
int fft_get_r16();
// Lines per workgroup of a pass whose lines are strided (0 = auto, the same as a pass of
// contiguous lines).
void fft_set_strided_lines(int v);
int fft_get_strided_lines();
// Smallest even 2,3,5-smooth integer >= n.
int64_t next235even(int64_t n);

// Forward twiddle table exp(-2 pi i k / L), k < L, cached per device.
template <typename R> const cplx<R>* twiddles(int L);

}  // namespace hasty_hip
