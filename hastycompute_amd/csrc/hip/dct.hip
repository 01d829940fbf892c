// DCT-II / its inverse along chosen axes of batched 3-D complex64 volumes, and the L1
// soft-threshold prox in the DCT domain (ProximalL1DCT, python/python/hastypy/base/
// proximal.py:134-183, which calls cupyx.scipy.fft.dctn / idctn with norm=None):
//   dct:  y_k = 2 sum_n x_n cos(pi k (2n + 1) / (2N))
//   idct: x_n = (1 / 2N) (y_0 + 2 sum_{k>=1} y_k cos(pi k (2n + 1) / (2N)))
// Each axis pass is a real N x N matrix applied to every line (N <= 512): a workgroup
// stages 16 lines in LDS and each thread forms outputs from the matrix column held
// contiguously in global memory (read through L2). This prox runs once per iteration
// on a few images, next to normal-operator applies that move gigabytes.
#include <cmath>
#include <vector>

#include "common.hpp"
#include "toeplitz.hpp"

namespace hasty_hip {

namespace {

constexpr int kDctThreads = 256;
constexpr int kDctLines = 16;

// matT[n * N + k] = forward coefficient of x_n in y_k, or (inverse) of y_n in x_k
__global__ void dct_matrix_kernel(float* matT, int N, int inverse)
{
    const int i = blockIdx.x * kDctThreads + threadIdx.x;
    if (i >= N * N) return;
    const int n = i / N, k = i % N;
    double v;
    if (!inverse) {
        v = 2.0 * cos(M_PI * (double)k * (2.0 * n + 1.0) / (2.0 * N));
    } else {  // input index n = frequency, output index k = sample
        v = (n == 0 ? 1.0 : 2.0) * cos(M_PI * (double)n * (2.0 * k + 1.0) / (2.0 * N)) / (2.0 * N);
    }
    matT[i] = (float)v;
}

// Lines of length N along one axis: line l starts at (l / inner) * outer + l % inner,
// elements `stride` apart (in place).
__global__ void __launch_bounds__(kDctThreads) dct_axis_kernel(float2* data, const float* __restrict__ matT, int N,
                                                               int64_t nlines, int64_t inner, int64_t outer,
                                                               int64_t stride)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float2* buf = reinterpret_cast<float2*>(smem_raw);  // [kDctLines][N]
    const int t = threadIdx.x;
    const int64_t l0 = (int64_t)blockIdx.x * kDctLines;
    const int nl = (int)min((int64_t)kDctLines, nlines - l0);
    for (int e = t; e < nl * N; e += kDctThreads) {
        const int li = e / N, n = e % N;
        const int64_t l = l0 + li;
        buf[li * N + n] = data[(l / inner) * outer + (l % inner) + (int64_t)n * stride];
    }
    __syncthreads();
    for (int e = t; e < nl * N; e += kDctThreads) {
        const int li = e / N, k = e % N;
        float ax = 0.f, ay = 0.f;
        const float2* xl = buf + li * N;
        for (int n = 0; n < N; ++n) {
            const float c = matT[n * N + k];
            const float2 x = xl[n];
            ax = fmaf(c, x.x, ax);
            ay = fmaf(c, x.y, ay);
        }
        const int64_t l = l0 + li;
        data[(l / inner) * outer + (l % inner) + (int64_t)k * stride] = float2{ax, ay};
    }
}

__global__ void soft_kernel(float2* d, int64_t n, float thresh)
{
    const int64_t i = (int64_t)blockIdx.x * kDctThreads + threadIdx.x;
    if (i >= n) return;
    const float2 v = d[i];
    const float m = sqrtf(v.x * v.x + v.y * v.y);
    const float s = m > thresh ? (m - thresh) / m : 0.f;
    d[i] = float2{v.x * s, v.y * s};
}

}  // namespace

void dct3(float2* data, int64_t batch, const int64_t* dims, int axis_mask, int inverse, hipStream_t s)
{
    HASTY_REQUIRE(batch >= 0, "batch must be >= 0");
    for (int d = 0; d < 3; ++d) HASTY_REQUIRE(dims[d] >= 1, "dims must be >= 1");
    const int64_t D0 = dims[0], D1 = dims[1], D2 = dims[2];
    const int64_t vol = D0 * D1 * D2;
    if (batch == 0 || vol == 0) return;
    for (int ax = 0; ax < 3; ++ax) {
        if (!(axis_mask & (1 << ax)) || dims[ax] == 1) continue;
        const int N = (int)dims[ax];
        HASTY_REQUIRE(N <= 512, "DCT axis longer than 512");
        void* m = dev_alloc(sizeof(float) * (size_t)N * N);
        try {
            hipLaunchKernelGGL(dct_matrix_kernel, dim3((unsigned)ceil_div((int64_t)N * N, kDctThreads)),
                               dim3(kDctThreads), 0, s, (float*)m, N, inverse);
            check_launch("dct_matrix_kernel");
            int64_t inner, outer, stride, nlines;
            if (ax == 2) {
                inner = 1, outer = D2, stride = 1, nlines = batch * D0 * D1;
            } else if (ax == 1) {
                inner = D2, outer = D1 * D2, stride = D2, nlines = batch * D0 * D2;
            } else {
                inner = D1 * D2, outer = vol, stride = D1 * D2, nlines = batch * D1 * D2;
            }
            const size_t smem = sizeof(float2) * kDctLines * N;
            allow_lds(dct_axis_kernel, smem);
            hipLaunchKernelGGL(dct_axis_kernel, dim3((unsigned)ceil_div(nlines, kDctLines)), dim3(kDctThreads), smem,
                               s, data, (const float*)m, N, nlines, inner, outer, stride);
            check_launch("dct_axis_kernel");
            HASTY_HIP_CHECK(hipStreamSynchronize(s));
        } catch (...) {
            dev_free(m);
            throw;
        }
        dev_free(m);
    }
}

void l1dct_prox(float2* data, int64_t batch, const int64_t* dims, int axis_mask, float thresh, hipStream_t s)
{
    dct3(data, batch, dims, axis_mask, 0, s);
    const int64_t n = batch * dims[0] * dims[1] * dims[2];
    if (n > 0) {
        hipLaunchKernelGGL(soft_kernel, dim3((unsigned)ceil_div(n, kDctThreads)), dim3(kDctThreads), 0, s, data, n,
                           thresh);
        check_launch("soft_kernel");
    }
    dct3(data, batch, dims, axis_mask, 1, s);
}

}  // namespace hasty_hip
