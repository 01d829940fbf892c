// Circulant preconditioner of the SENSE normal operator (CirculantPreconditionerOp,
// lib/mri/precond.cpp:27-117, which restates sigpy.mri.circulant_precond).
//
// The preconditioner is P = F^-1 diag(d) F with d = 1 / p, where p is the diagonal of
// F A^H A F^-1 on the image grid: the eigenvalues of the circulant matrix closest to the
// normal operator A^H A in Frobenius norm. For the SENSE normal operator with sample
// weights w, on the doubled grid (L = 2N per axis, Lg = prod L):
//   p[k] = scale / Lg^2 * DFT( R * psf )[2k]   (the frequencies of the N grid),
//   R    = IDFT_unnormalised( sum_c |DFT(pad(conj S_c))|^2 )   (coil autocorrelations),
//   psf  = the type-1 NUFFT of sqrt(w) onto the doubled modes, scaled by 1/sqrt(Lg),
//   scale = Lg^1.5 / prod(N)^2,
// the same quantity the reference forms coil by coil with centred orthonormal FFTs
// (the sum over coils moves inside the last FFT by linearity; |DFT|^2 does not depend on
// where the map sits in the padded grid). Then p += lambda, p == 0 -> 1, d = 1 / p.
// Everything is FP64 on the GPU; d is written as complex64.
//
// Applying P is a circular convolution, so centring (fftc / ifftc) only changes the
// order in which d is stored, not the data: with d in centred order the apply uses
// ifftshift(d). The adjoint fftn(conj(d) ifftn(x)) equals conj(P conj(x)).
#include <cmath>
#include <vector>

#include "fft.hpp"
#include "nufft.hpp"
#include "toeplitz.hpp"

namespace hasty_hip {

namespace {

constexpr int kThreads = 256;

template <typename K, typename... Args> void launch1d(K kernel, int64_t n, hipStream_t s, Args... args)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(kernel, dim3((unsigned)ceil_div(n, kThreads)), dim3(kThreads), 0, s, args...);
    check_launch("preconditioner elementwise kernel");
}

struct DevBuf {
    void* p = nullptr;
    explicit DevBuf(size_t bytes) { p = dev_alloc(bytes); }
    ~DevBuf() { dev_free(p); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
};

__global__ void to_double_kernel(const float* in, double* out, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i < n) out[i] = (double)in[i];
}

__global__ void sqrt_weights_kernel(const float* w, double2* c, int64_t M)
{
    const int64_t j = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (j < M) c[j] = double2{w ? sqrt((double)w[j]) : 1.0, 0.0};
}

// Zero-padded conj(S) on the doubled grid, map at the low corner of every axis.
__global__ void pad_conj_kernel(const float2* S, double2* g, int64_t N0, int64_t N1, int64_t N2, int64_t L0,
                                int64_t L1, int64_t L2)
{
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i >= L0 * L1 * L2) return;
    const int64_t i2 = i % L2, i1 = (i / L2) % L1, i0 = i / (L1 * L2);
    double2 v{0.0, 0.0};
    if (i0 < N0 && i1 < N1 && i2 < N2) {
        const float2 s = S[(i0 * N1 + i1) * N2 + i2];
        v = double2{(double)s.x, -(double)s.y};
    }
    g[i] = v;
}

// acc (+)= |g|^2
__global__ void abs2_acc_kernel(const double2* g, double2* acc, int64_t n, int first)
{
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    const double m = g[i].x * g[i].x + g[i].y * g[i].y;
    acc[i] = double2{first ? m : acc[i].x + m, 0.0};
}

// R (lag m, FFT order) times psf (modes in centred order: lag m at (m + N) mod L per axis)
__global__ void times_psf_kernel(double2* R, const double2* psf, int64_t L0, int64_t L1, int64_t L2)
{
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i >= L0 * L1 * L2) return;
    const int64_t i2 = i % L2, i1 = (i / L2) % L1, i0 = i / (L1 * L2);
    const int64_t j0 = (i0 + L0 / 2) % L0, j1 = (i1 + L1 / 2) % L1, j2 = (i2 + L2 / 2) % L2;
    const double2 a = R[i], b = psf[(j0 * L1 + j1) * L2 + j2];
    R[i] = double2{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}

// d at N-grid index n. The reference takes every second entry of the centred doubled
// grid: centred index i of the N grid reads doubled-grid frequency 2 i - N. In FFT order
// index n holds the centred entry (n + N/2) mod N (numpy's ifftshift).
__global__ void diag_kernel(const double2* D, float2* d, int64_t N0, int64_t N1, int64_t N2, int64_t L0, int64_t L1,
                            int64_t L2, double cst, double lambda, int centered)
{
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i >= N0 * N1 * N2) return;
    const int64_t n2 = i % N2, n1 = (i / N2) % N1, n0 = i / (N1 * N2);
    auto pos = [&](int64_t n, int64_t N, int64_t L) {
        const int64_t ic = centered ? n : (n + N / 2) % N;
        return ((2 * ic - N) % L + L) % L;
    };
    const int64_t k0 = pos(n0, N0, L0), k1 = pos(n1, N1, L1), k2 = pos(n2, N2, L2);
    const double2 v = D[(k0 * L1 + k1) * L2 + k2];
    double pr = v.x * cst + lambda, pi = v.y * cst;
    if (pr == 0.0 && pi == 0.0) pr = 1.0;
    const double m = pr * pr + pi * pi;
    d[i] = float2{(float)(pr / m), (float)(-pi / m)};
}

// Unnormalised in-place FFT over a [L0][L1][L2] double grid (axes with L == 1 skipped).
void fftn_d(double2* buf, const int64_t* L, int sign, hipStream_t s)
{
    const int64_t inner[3] = {L[1] * L[2], L[2], 1};
    const int64_t lines[3] = {L[1] * L[2], L[0] * L[2], L[0] * L[1]};
    for (int ax = 2; ax >= 0; --ax) {
        if (L[ax] <= 1) continue;
        AxisFft<double> a;
        a.in = buf;
        if (ax == 2)
            a.ain = LineAddr{1, L[2], 0, 1, nullptr};
        else if (ax == 1)
            a.ain = LineAddr{L[2], L[1] * L[2], 1, L[2], nullptr};
        else
            a.ain = LineAddr{L[1] * L[2], 0, 1, L[1] * L[2], nullptr};
        a.out = buf;
        a.aout = a.ain;
        a.L = (int)L[ax];
        a.n_in = (int)L[ax];
        a.n_out = (int)L[ax];
        a.sign = sign;
        a.nlines = lines[ax];
        (void)inner;
        run_axis_fft<double>(a, s);
    }
}

__global__ void ifftshift_copy_kernel(const float2* in, float2* out, int64_t N0, int64_t N1, int64_t N2)
{
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i >= N0 * N1 * N2) return;
    const int64_t n2 = i % N2, n1 = (i / N2) % N1, n0 = i / (N1 * N2);
    // FFT-order index n holds the centred entry at (n + N/2) mod N
    const int64_t c0 = (n0 + N0 / 2) % N0, c1 = (n1 + N1 / 2) % N1, c2 = (n2 + N2 / 2) % N2;
    out[i] = in[(c0 * N1 + c1) * N2 + c2];
}

__global__ void conj_copy_kernel(const float2* in, float2* out, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i < n) out[i] = float2{in[i].x, -in[i].y};
}

}  // namespace

void circulant_precond_build(const float2* smaps, int64_t ncoil, const void* coords, int coord_fp64, int64_t M,
                             int ndim, const int64_t* nmodes, const float* weights, double lambda, int centered,
                             float2* diag, double tol, hipStream_t s)
{
    HASTY_REQUIRE(ndim >= 1 && ndim <= 3, "Dimension must be 1,2 or 3");
    HASTY_REQUIRE(ncoil >= 1, "the preconditioner needs at least one coil map");
    HASTY_REQUIRE(M >= 0, "M must be >= 0");
    int64_t N[3] = {1, 1, 1}, L[3] = {1, 1, 1};
    for (int d = 0; d < ndim; ++d) {
        HASTY_REQUIRE(nmodes[d] >= 1, "nmodes must be >= 1");
        HASTY_REQUIRE(fft_size_supported((int)(2 * nmodes[d])), "2 N must be 2,3,5-smooth");
        N[3 - ndim + d] = nmodes[d];
        L[3 - ndim + d] = 2 * nmodes[d];
    }
    const int64_t nvox = N[0] * N[1] * N[2];
    const int64_t ngrid = L[0] * L[1] * L[2];

    // psf: type-1 (isign +1) FP64 NUFFT of sqrt(w) onto the doubled modes, / sqrt(Lg)
    DevBuf cd(sizeof(double) * std::max<int64_t>(1, (int64_t)ndim * M));
    const double* cdp = reinterpret_cast<const double*>(coords);
    if (!coord_fp64) {
        launch1d(to_double_kernel, (int64_t)ndim * M, s, (const float*)coords, (double*)cd.p, (int64_t)ndim * M);
        cdp = (const double*)cd.p;
    }
    DevBuf cw(sizeof(double2) * std::max<int64_t>(1, M));
    launch1d(sqrt_weights_kernel, M, s, weights, (double2*)cw.p, M);
    int64_t nm_flip[3];
    for (int d = 0; d < ndim; ++d) nm_flip[d] = 2 * nmodes[ndim - 1 - d];
    NufftPlan* plan = nufft_makeplan(1, ndim, nm_flip, 1, 1, tol, 1);
    DevBuf psf(sizeof(double2) * ngrid);
    try {
        const double* row[3] = {nullptr, nullptr, nullptr};
        for (int d = 0; d < ndim; ++d) row[d] = cdp + (int64_t)(ndim - 1 - d) * M;  // x = last row
        nufft_setpts(plan, M, row[0], row[1], row[2], s);
        nufft_execute(plan, cw.p, psf.p, s);
        HASTY_HIP_CHECK(hipStreamSynchronize(s));
    } catch (...) {
        nufft_destroy(plan);
        throw;
    }
    nufft_destroy(plan);

    // sum over coils of |DFT(pad conj S_c)|^2, then the autocorrelation R
    DevBuf g(sizeof(double2) * ngrid), acc(sizeof(double2) * ngrid);
    for (int64_t c = 0; c < ncoil; ++c) {
        launch1d(pad_conj_kernel, ngrid, s, smaps + c * nvox, (double2*)g.p, N[0], N[1], N[2], L[0], L[1], L[2]);
        fftn_d((double2*)g.p, L, -1, s);
        launch1d(abs2_acc_kernel, ngrid, s, (const double2*)g.p, (double2*)acc.p, ngrid, c == 0 ? 1 : 0);
    }
    fftn_d((double2*)acc.p, L, +1, s);
    launch1d(times_psf_kernel, ngrid, s, (double2*)acc.p, (const double2*)psf.p, L[0], L[1], L[2]);
    fftn_d((double2*)acc.p, L, -1, s);
    // scale / Lg^2 with scale = Lg^1.5 / nvox^2, and the psf's 1 / sqrt(Lg)
    const double Lg = (double)ngrid;
    const double cst = std::pow(Lg, 1.5) / ((double)nvox * (double)nvox) / (Lg * Lg) / std::sqrt(Lg);
    launch1d(diag_kernel, nvox, s, (const double2*)acc.p, diag, N[0], N[1], N[2], L[0], L[1], L[2], cst, lambda,
             centered);
    HASTY_HIP_CHECK(hipStreamSynchronize(s));
}

size_t circulant_precond_workspace_bytes(int64_t ndiag, int ndim, const int64_t* nmodes, int centered, int adjoint)
{
    int64_t nvox = 1;
    for (int d = 0; d < ndim; ++d) nvox *= nmodes[d];
    size_t b = 0;
    if (centered) b += sizeof(float2) * (size_t)(ndiag * nvox);
    if (adjoint) b += sizeof(float2) * (size_t)nvox;
    return b;
}

void circulant_precond_apply(const float2* in, float2* out, const float2* diag, int64_t ndiag, int64_t P,
                             const int32_t* dindex, int ndim, const int64_t* nmodes, int centered, int adjoint,
                             void* ws, size_t ws_bytes, hipStream_t s)
{
    HASTY_REQUIRE(ndim >= 1 && ndim <= 3, "Dimension must be 1,2 or 3");
    HASTY_REQUIRE(P >= 0 && ndiag >= 1, "invalid preconditioner batch");
    int64_t N[3] = {1, 1, 1};
    for (int d = 0; d < ndim; ++d) {
        HASTY_REQUIRE(nmodes[d] >= 1 && fft_size_supported((int)nmodes[d]), "N must be 2,3,5-smooth");
        N[3 - ndim + d] = nmodes[d];
    }
    const int64_t nvox = N[0] * N[1] * N[2];
    HASTY_REQUIRE(ws_bytes >= circulant_precond_workspace_bytes(ndiag, ndim, nmodes, centered, adjoint),
                  "preconditioner workspace too small");
    char* wp = reinterpret_cast<char*>(ws);
    const float2* d = diag;
    if (centered) {
        float2* dn = reinterpret_cast<float2*>(wp);
        for (int64_t k = 0; k < ndiag; ++k)
            launch1d(ifftshift_copy_kernel, nvox, s, diag + k * nvox, dn + k * nvox, N[0], N[1], N[2]);
        d = dn;
        wp += sizeof(float2) * ndiag * nvox;
    }
    float2* tmp = reinterpret_cast<float2*>(wp);
    for (int64_t p = 0; p < P; ++p) {
        const int64_t k = dindex ? dindex[p] : p;
        HASTY_REQUIRE(k >= 0 && k < ndiag, "diagonal index out of range");
        const float2* src = in + p * nvox;
        float2* dst = out + p * nvox;
        if (adjoint) {
            launch1d(conj_copy_kernel, nvox, s, src, tmp, nvox);
            src = tmp;
        }
        // ifftn(d * fftn(x)): forward passes on the two fast axes, the slow axis as one
        // convolution pass (forward, times d, inverse), inverse passes back
        const float2* cur = src;
        const int64_t inner[3] = {N[1] * N[2], N[2], 1};
        auto axis_addr = [&](int ax) {
            return ax == 2 ? LineAddr{1, N[2], 0, 1, nullptr}
                           : (ax == 1 ? LineAddr{N[2], N[1] * N[2], 1, N[2], nullptr}
                                      : LineAddr{N[1] * N[2], 0, 1, N[1] * N[2], nullptr});
        };
        (void)inner;
        int conv_ax = 0;
        while (conv_ax < 3 && N[conv_ax] == 1) ++conv_ax;
        for (int ax = 2; ax > conv_ax; --ax) {
            if (N[ax] == 1) continue;
            AxisFft<float> a;
            a.in = cur;
            a.ain = axis_addr(ax);
            a.out = dst;
            a.aout = a.ain;
            a.L = a.n_in = a.n_out = (int)N[ax];
            a.sign = -1;
            a.nlines = nvox / N[ax];
            run_axis_fft<float>(a, s);
            cur = dst;
        }
        {
            AxisFft<float> a;
            a.in = cur;
            a.ain = axis_addr(conv_ax);
            a.out = dst;
            a.aout = a.ain;
            a.L = a.n_in = a.n_out = (int)N[conv_ax];
            a.conv_cplx = d + k * nvox;
            a.aconv = a.ain;
            a.scale = (float)(1.0 / (double)nvox);
            a.nlines = nvox / N[conv_ax];
            run_axis_fft<float>(a, s);
            cur = dst;
        }
        for (int ax = conv_ax + 1; ax < 3; ++ax) {
            if (N[ax] == 1) continue;
            AxisFft<float> a;
            a.in = dst;
            a.ain = axis_addr(ax);
            a.out = dst;
            a.aout = a.ain;
            a.L = a.n_in = a.n_out = (int)N[ax];
            a.sign = 1;
            a.nlines = nvox / N[ax];
            run_axis_fft<float>(a, s);
        }
        if (adjoint) launch1d(conj_copy_kernel, nvox, s, (const float2*)dst, dst, nvox);
    }
}

}  // namespace hasty_hip
