// Device-side batched conjugate gradient (the body of ConjugateGradient::run,
// lib/op/opalgs.cpp:50-96) for P independent problems of n complex64 unknowns.
//
// All scalars (rzold, pAp, alpha, beta) live on the device in FP64; no step reads
// them back to the host, so an iteration is a fixed launch sequence that a hipGraph
// can capture. Inner products are reduced in FP64 in a fixed order (block partials,
// then one workgroup per problem), so results do not depend on scheduling.
//
// Per-problem state (doubles): [0] rzold, [1] alpha, [2] beta, [3] done, [4] not-PD,
// [5] last pAp, [6] iterations run.
#include "cg.hpp"

namespace hasty_hip {

constexpr int kCgState = 8;

namespace {

constexpr int kThreads = 256;
constexpr int kBlocksPerProblem = 240;

__device__ __forceinline__ double block_reduce(double v, double* sh)
{
    const int tid = threadIdx.x;
    sh[tid] = v;
    __syncthreads();
    for (int off = kThreads / 2; off > 0; off >>= 1) {
        if (tid < off) sh[tid] += sh[tid + off];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

// partial[p][blk] = sum Re(conj(a) * (b + shift a)) over this block's slice of problem p
// (b + shift a rounded to float, as cg_update_xr_kernel forms it)
__global__ void __launch_bounds__(kThreads) dot_re_kernel(const float2* a, const float2* b, int64_t n, double* partial,
                                                          float shift)
{
    __shared__ double sh[kThreads];
    const int p = blockIdx.y;
    const float2* ap = a + (int64_t)p * n;
    const float2* bp = b + (int64_t)p * n;
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
        const float2 u = ap[i];
        float2 v = bp[i];
        if (shift != 0.f) {
            v.x += shift * u.x;
            v.y += shift * u.y;
        }
        s += (double)u.x * (double)v.x + (double)u.y * (double)v.y;
    }
    s = block_reduce(s, sh);
    if (threadIdx.x == 0) partial[(int64_t)p * gridDim.x + blockIdx.x] = s;
}

__device__ __forceinline__ double sum_partials(const double* partial, int nblk, double* sh)
{
    double s = 0.0;
    for (int i = threadIdx.x; i < nblk; i += kThreads) s += partial[i];
    return block_reduce(s, sh);
}

// r = b - Ax; p = r; partial |r|^2
__global__ void __launch_bounds__(kThreads) cg_init_kernel(const float2* b, const float2* Ax, float2* r, float2* pv,
                                                           int64_t n, double* partial)
{
    __shared__ double sh[kThreads];
    const int p = blockIdx.y;
    const int64_t off = (int64_t)p * n;
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
        const float2 v = csub(b[off + i], Ax[off + i]);
        r[off + i] = v;
        pv[off + i] = v;
        s += (double)v.x * v.x + (double)v.y * v.y;
    }
    s = block_reduce(s, sh);
    if (threadIdx.x == 0) partial[(int64_t)p * gridDim.x + blockIdx.x] = s;
}

__global__ void __launch_bounds__(kThreads) cg_init_finalize(const double* partial, int nblk, double* state)
{
    __shared__ double sh[kThreads];
    const int p = blockIdx.x;
    const double s = sum_partials(partial + (int64_t)p * nblk, nblk, sh);
    if (threadIdx.x == 0) {
        double* st = state + (int64_t)p * kCgState;
        st[0] = s;
        st[1] = 0.0;
        st[2] = 0.0;
        st[3] = 0.0;
        st[4] = 0.0;
        st[5] = 0.0;
        st[6] = 0.0;
    }
}

// alpha = rzold / pAp, with the reference's early exit (resid < tol) and PD check.
__global__ void __launch_bounds__(kThreads) cg_alpha_kernel(const double* partial, int nblk, double* state, double tol)
{
    __shared__ double sh[kThreads];
    const int p = blockIdx.x;
    const double pAp = sum_partials(partial + (int64_t)p * nblk, nblk, sh);
    if (threadIdx.x == 0) {
        double* st = state + (int64_t)p * kCgState;
        st[5] = pAp;
        if (st[3] == 0.0 && sqrt(st[0]) < tol) st[3] = 1.0;
        if (st[3] == 0.0 && !(pAp > 0.0)) {
            st[3] = 1.0;
            st[4] = 1.0;
        }
        st[1] = st[3] != 0.0 ? 0.0 : st[0] / pAp;
        if (st[3] == 0.0) st[6] += 1.0;
    }
}

// x += alpha p; r -= alpha (Ap + shift p); partial |r|^2
__global__ void __launch_bounds__(kThreads) cg_update_xr_kernel(float2* x, float2* r, const float2* pv, const float2* Ap,
                                                                int64_t n, const double* state, double* partial,
                                                                float shift)
{
    __shared__ double sh[kThreads];
    const int p = blockIdx.y;
    const int64_t off = (int64_t)p * n;
    const double* st = state + (int64_t)p * kCgState;
    const float alpha = (float)st[1];
    const bool active = st[3] == 0.0;
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
        float2 rv = r[off + i];
        if (active) {
            const float2 pp = pv[off + i];
            float2 ap = Ap[off + i];
            if (shift != 0.f) {
                ap.x += shift * pp.x;
                ap.y += shift * pp.y;
            }
            float2 xv = x[off + i];
            xv.x += alpha * pp.x;
            xv.y += alpha * pp.y;
            x[off + i] = xv;
            rv.x -= alpha * ap.x;
            rv.y -= alpha * ap.y;
            r[off + i] = rv;
        }
        s += (double)rv.x * rv.x + (double)rv.y * rv.y;
    }
    s = block_reduce(s, sh);
    if (threadIdx.x == 0) partial[(int64_t)p * gridDim.x + blockIdx.x] = s;
}

__global__ void __launch_bounds__(kThreads) cg_beta_kernel(const double* partial, int nblk, double* state)
{
    __shared__ double sh[kThreads];
    const int p = blockIdx.x;
    const double rznew = sum_partials(partial + (int64_t)p * nblk, nblk, sh);
    if (threadIdx.x == 0) {
        double* st = state + (int64_t)p * kCgState;
        if (st[3] == 0.0) {
            st[2] = rznew / st[0];
            st[0] = rznew;
        } else {
            st[2] = 0.0;
        }
    }
}

// p = r + beta p
__global__ void __launch_bounds__(kThreads) cg_update_p_kernel(float2* pv, const float2* r, int64_t n, const double* state)
{
    const int p = blockIdx.y;
    const int64_t off = (int64_t)p * n;
    const double* st = state + (int64_t)p * kCgState;
    if (st[3] != 0.0) return;
    const float beta = (float)st[2];
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
        const float2 rv = r[off + i];
        float2 pp = pv[off + i];
        pp.x = rv.x + beta * pp.x;
        pp.y = rv.y + beta * pp.y;
        pv[off + i] = pp;
    }
}

// rzold = partial sum (Re <r, z>); p = z (preconditioned start)
__global__ void __launch_bounds__(kThreads) cg_rz_finalize(const double* partial, int nblk, double* state)
{
    __shared__ double sh[kThreads];
    const int p = blockIdx.x;
    const double s = sum_partials(partial + (int64_t)p * nblk, nblk, sh);
    if (threadIdx.x == 0) state[(int64_t)p * kCgState] = s;
}

__global__ void __launch_bounds__(kThreads) copy_kernel(float2* dst, const float2* src, int64_t n)
{
    const int p = blockIdx.y;
    const int64_t off = (int64_t)p * n;
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads)
        dst[off + i] = src[off + i];
}

int blocks_for(int64_t n)
{
    int64_t b = ceil_div(n, kThreads * 4);
    if (b < 1) b = 1;
    if (b > kBlocksPerProblem) b = kBlocksPerProblem;
    return (int)b;
}

}  // namespace

size_t cg_workspace_doubles(int64_t P, int64_t n) { return (size_t)P * kBlocksPerProblem; }

void cg_init(int64_t P, int64_t n, const float2* b, const float2* Ax, float2* r, float2* p, double* state, double* ws,
             hipStream_t s)
{
    HASTY_REQUIRE(P >= 1 && P < 65536 && n >= 1, "invalid CG problem sizes");
    const int nb = blocks_for(n);
    hipLaunchKernelGGL(cg_init_kernel, dim3(nb, (unsigned)P), dim3(kThreads), 0, s, b, Ax, r, p, n, ws);
    check_launch("cg_init_kernel");
    hipLaunchKernelGGL(cg_init_finalize, dim3((unsigned)P), dim3(kThreads), 0, s, (const double*)ws, nb, state);
    check_launch("cg_init_finalize");
}

void cg_dot_re(int64_t P, int64_t n, const float2* a, const float2* b, double* partial, int* nblk, hipStream_t s)
{
    const int nb = blocks_for(n);
    hipLaunchKernelGGL(dot_re_kernel, dim3(nb, (unsigned)P), dim3(kThreads), 0, s, a, b, n, partial, 0.f);
    check_launch("dot_re_kernel");
    *nblk = nb;
}

void cg_iterate(int64_t P, int64_t n, float2* x, float2* r, float2* p, const float2* Ap, double* state, double* ws,
                double tol, hipStream_t s, double shift)
{
    HASTY_REQUIRE(P >= 1 && P < 65536 && n >= 1, "invalid CG problem sizes");
    HASTY_REQUIRE(shift >= 0.0, "the l2 regularisation weight must be >= 0");
    const int nb = blocks_for(n);
    const float sh = (float)shift;
    hipLaunchKernelGGL(dot_re_kernel, dim3(nb, (unsigned)P), dim3(kThreads), 0, s, (const float2*)p, Ap, n, ws, sh);
    check_launch("dot_re_kernel");
    hipLaunchKernelGGL(cg_alpha_kernel, dim3((unsigned)P), dim3(kThreads), 0, s, (const double*)ws, nb, state, tol);
    check_launch("cg_alpha_kernel");
    hipLaunchKernelGGL(cg_update_xr_kernel, dim3(nb, (unsigned)P), dim3(kThreads), 0, s, x, r, (const float2*)p, Ap, n,
                       (const double*)state, ws, sh);
    check_launch("cg_update_xr_kernel");
    hipLaunchKernelGGL(cg_beta_kernel, dim3((unsigned)P), dim3(kThreads), 0, s, (const double*)ws, nb, state);
    check_launch("cg_beta_kernel");
    hipLaunchKernelGGL(cg_update_p_kernel, dim3(nb, (unsigned)P), dim3(kThreads), 0, s, p, (const float2*)r, n,
                       (const double*)state);
    check_launch("cg_update_p_kernel");
}

// Preconditioned variant (the _P branch of ConjugateGradient::run, opalgs.cpp:56-90):
// cg_init, then z = P r by the caller, then cg_init_precond; per iteration Ap = A p,
// cg_iterate_xr, z = P r, cg_iterate_pz.
void cg_init_precond(int64_t P, int64_t n, const float2* r, const float2* z, float2* p, double* state, double* ws,
                     hipStream_t s)
{
    HASTY_REQUIRE(P >= 1 && P < 65536 && n >= 1, "invalid CG problem sizes");
    const int nb = blocks_for(n);
    hipLaunchKernelGGL(dot_re_kernel, dim3(nb, (unsigned)P), dim3(kThreads), 0, s, r, z, n, ws, 0.f);
    check_launch("dot_re_kernel");
    hipLaunchKernelGGL(cg_rz_finalize, dim3((unsigned)P), dim3(kThreads), 0, s, (const double*)ws, nb, state);
    check_launch("cg_rz_finalize");
    hipLaunchKernelGGL(copy_kernel, dim3(nb, (unsigned)P), dim3(kThreads), 0, s, p, z, n);
    check_launch("copy_kernel");
}

void cg_iterate_xr(int64_t P, int64_t n, float2* x, float2* r, const float2* p, const float2* Ap, double* state,
                   double* ws, double tol, hipStream_t s)
{
    HASTY_REQUIRE(P >= 1 && P < 65536 && n >= 1, "invalid CG problem sizes");
    const int nb = blocks_for(n);
    hipLaunchKernelGGL(dot_re_kernel, dim3(nb, (unsigned)P), dim3(kThreads), 0, s, p, Ap, n, ws, 0.f);
    check_launch("dot_re_kernel");
    hipLaunchKernelGGL(cg_alpha_kernel, dim3((unsigned)P), dim3(kThreads), 0, s, (const double*)ws, nb, state, tol);
    check_launch("cg_alpha_kernel");
    hipLaunchKernelGGL(cg_update_xr_kernel, dim3(nb, (unsigned)P), dim3(kThreads), 0, s, x, r, p, Ap, n,
                       (const double*)state, ws, 0.f);
    check_launch("cg_update_xr_kernel");
}

void cg_iterate_pz(int64_t P, int64_t n, const float2* r, const float2* z, float2* p, double* state, double* ws,
                   hipStream_t s)
{
    HASTY_REQUIRE(P >= 1 && P < 65536 && n >= 1, "invalid CG problem sizes");
    const int nb = blocks_for(n);
    hipLaunchKernelGGL(dot_re_kernel, dim3(nb, (unsigned)P), dim3(kThreads), 0, s, r, z, n, ws, 0.f);
    check_launch("dot_re_kernel");
    hipLaunchKernelGGL(cg_beta_kernel, dim3((unsigned)P), dim3(kThreads), 0, s, (const double*)ws, nb, state);
    check_launch("cg_beta_kernel");
    hipLaunchKernelGGL(cg_update_p_kernel, dim3(nb, (unsigned)P), dim3(kThreads), 0, s, p, z, n, (const double*)state);
    check_launch("cg_update_p_kernel");
}

}  // namespace hasty_hip
