// Autotuner of launch shapes (see tune.hip).
#pragma once

#include <functional>

#include "common.hpp"

namespace hasty_hip {

struct SenseChoice {
    int xy_variant = 2;     // plane kernels of the fused operator (library option sense_xy_variant)
    int zconv_batch = 16;   // problems per z-convolution launch at N0 = 128 (sense_zconv_batch)
    int zconv_variant = 2;  // z-convolution kernel at N0 = 128 (sense_zconv_variant)
    int problem_batch = 1;  // problems per launch (workspace sized for this many)
    int coil_batch = 1;     // coils sharing one oversampled-grid workspace
    double ms_per_problem = 0.0;
};

// Times the batched Toeplitz SENSE normal operator for `nproblems` problems (at most 8
// timed) of `ncoil` coils at image size n3 over the candidate plane kernels, problem
// batches, z-convolution kernels and batches and coil batches; sets the library options
// sense_xy_variant / sense_zconv_variant / sense_zconv_batch to the winners (kernels
// must be prepared after this call) and returns the choice. Cached per shape.
SenseChoice autotune_sense_normal(const int64_t* n3, int ncoil, int nproblems, bool allow_fused, hipStream_t s);

// Best of `reps` event-timed runs of fn on stream s after one untimed run, in ms.
double time_best_ms(const std::function<void()>& fn, hipStream_t s, int reps);
// x[i] uniform in [lo, hi) from a hash of (i, seed): deterministic synthetic data.
void fill_uniform(float* x, int64_t n, uint32_t seed, float lo, float hi, hipStream_t s);

// Streaming probe (probe.hip): nread streams read and nwrite streams written, `bytes`
// each, 16 bytes per lane, non-temporal if nt. A read-only mix writes nothing (dst[0]
// is a 16-byte sink it never touches for finite data).
void stream_probe(int nread, int nwrite, bool nt, const void* const* src, void* const* dst, int64_t bytes,
                  hipStream_t s);

}  // namespace hasty_hip
