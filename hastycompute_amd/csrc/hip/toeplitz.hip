// Toeplitz PSF build (FP64 on the GPU) and the Toeplitz / SENSE normal applies.
#include <cmath>
#include <vector>

#include "fft.hpp"
#include "nufft.hpp"
#include "toeplitz.hpp"

namespace hasty_hip {

namespace {

constexpr int kThreads = 256;

__global__ void coords_to_double_kernel(const float* in, double* out, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i < n) out[i] = (double)in[i];
}

__global__ void strengths_kernel(const float* w, double2* c, int64_t M)
{
    const int64_t j = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (j < M) c[j] = double2{w ? (double)w[j] : 1.0, 0.0};
}

// Copy with the k = -N entries (index 0 along any doubled dim) zeroed.
__global__ void zero_boundary_kernel(const double2* in, double2* out, int64_t L0, int64_t L1, int64_t L2, int ndim)
{
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    const int64_t n = L0 * L1 * L2;
    if (i >= n) return;
    const int64_t i2 = i % L2;
    const int64_t i1 = (i / L2) % L1;
    const int64_t i0 = i / (L1 * L2);
    bool z = (i2 == 0);
    if (ndim >= 2) z = z || (i1 == 0);
    if (ndim >= 3) z = z || (i0 == 0);
    out[i] = z ? double2{0.0, 0.0} : in[i];
}

__global__ void to_c64_kernel(const double2* in, float2* out, int64_t n, double scale)
{
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i < n) out[i] = float2{(float)(in[i].x * scale), (float)(in[i].y * scale)};
}

__global__ void real_to_f32_kernel(const double2* in, float* out, int64_t n, double scale)
{
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i < n) out[i] = (float)(in[i].x * scale);
}

__global__ void sub_kernel(float2* out, const float2* sub, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i < n) out[i] = csub(out[i], sub[i]);
}

template <typename K, typename... Args> void launch1d(K kernel, int64_t n, hipStream_t s, Args... args)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(kernel, dim3((unsigned)ceil_div(n, kThreads)), dim3(kThreads), 0, s, args...);
    check_launch("toeplitz elementwise kernel");
}

// In-place ifftshift + unnormalised forward fftn over a [L0][L1][L2] double grid
// (dims with L == 1 are skipped).
void shifted_fftn_d(double2* buf, const int64_t* L, hipStream_t s)
{
    const int64_t L0 = L[0], L1 = L[1], L2 = L[2];
    if (L2 > 1) {
        AxisFft<double> a;
        a.in = buf;
        a.ain = LineAddr{1, L2, 0, 1, nullptr};
        a.out = buf;
        a.aout = a.ain;
        a.L = (int)L2;
        a.n_in = (int)L2;
        a.n_out = (int)L2;
        a.in_centered = 1;
        a.sign = -1;
        a.nlines = L0 * L1;
        run_axis_fft<double>(a, s);
    }
    if (L1 > 1) {
        AxisFft<double> a;
        a.in = buf;
        a.ain = LineAddr{L2, L1 * L2, 1, L2, nullptr};
        a.out = buf;
        a.aout = a.ain;
        a.L = (int)L1;
        a.n_in = (int)L1;
        a.n_out = (int)L1;
        a.in_centered = 1;
        a.sign = -1;
        a.nlines = L0 * L2;
        run_axis_fft<double>(a, s);
    }
    if (L0 > 1) {
        AxisFft<double> a;
        a.in = buf;
        a.ain = LineAddr{L1 * L2, 0, 1, L1 * L2, nullptr};
        a.out = buf;
        a.aout = a.ain;
        a.L = (int)L0;
        a.n_in = (int)L0;
        a.n_out = (int)L0;
        a.in_centered = 1;
        a.sign = -1;
        a.nlines = L1 * L2;
        run_axis_fft<double>(a, s);
    }
}

struct DevBuf {
    void* p = nullptr;
    explicit DevBuf(size_t bytes) { p = dev_alloc(bytes); }
    ~DevBuf() { dev_free(p); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
};

}  // namespace

void toeplitz_build(const void* coords, int coord_fp64, int64_t M, int ndim, const int64_t* nmodes,
                    const float* weights, float2* diag_ref, float* kernel, double tol, hipStream_t s)
{
    HASTY_REQUIRE(ndim >= 1 && ndim <= 3, "Dimension must be 1,2 or 3");
    HASTY_REQUIRE(M >= 0, "M must be >= 0");
    int64_t L[3] = {1, 1, 1}, N[3] = {1, 1, 1};
    for (int d = 0; d < ndim; ++d) {
        HASTY_REQUIRE(nmodes[d] >= 1, "nmodes must be >= 1");
        N[3 - ndim + d] = nmodes[d];
        L[3 - ndim + d] = 2 * nmodes[d];
    }
    const int64_t ngrid = L[0] * L[1] * L[2];

    // FP64 coordinates and strengths.
    DevBuf cd(sizeof(double) * std::max<int64_t>(1, (int64_t)ndim * M));
    const double* cdp = reinterpret_cast<const double*>(coords);
    if (!coord_fp64) {
        launch1d(coords_to_double_kernel, (int64_t)ndim * M, s, (const float*)coords, (double*)cd.p, (int64_t)ndim * M);
        cdp = (const double*)cd.p;
    }
    DevBuf cw(sizeof(double2) * std::max<int64_t>(1, M));
    launch1d(strengths_kernel, M, s, weights, (double2*)cw.p, M);

    // Type-1 (isign +1) FP64 NUFFT onto the doubled modes: psf[n] = sum_j w_j exp(+i (n - N).x_j).
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

    const double inv_grid = 1.0 / (double)ngrid;
    DevBuf work(sizeof(double2) * ngrid);
    if (diag_ref) {
        HASTY_HIP_CHECK(hipMemcpyAsync(work.p, psf.p, sizeof(double2) * ngrid, hipMemcpyDeviceToDevice, s));
        shifted_fftn_d((double2*)work.p, L, s);
        launch1d(to_c64_kernel, ngrid, s, (const double2*)work.p, diag_ref, ngrid, inv_grid);
    }
    if (kernel) {
        double nvox = (double)(N[0] * N[1] * N[2]);
        launch1d(zero_boundary_kernel, ngrid, s, (const double2*)psf.p, (double2*)work.p, L[0], L[1], L[2], ndim);
        shifted_fftn_d((double2*)work.p, L, s);
        launch1d(real_to_f32_kernel, ngrid, s, (const double2*)work.p, kernel, ngrid, inv_grid / nvox);
    }
    HASTY_HIP_CHECK(hipStreamSynchronize(s));
}

void toeplitz_apply_ref(const float2* in, float2* out, const float2* mul, const float2* diag, int ntransf, int ndim,
                        const int64_t* nmodes, int mode, float2* st1, float2* st2, hipStream_t s)
{
    HASTY_REQUIRE(ndim >= 1 && ndim <= 3, "Dimension must be 1,2 or 3");
    HASTY_REQUIRE(ntransf >= 1, "ntransf must be >= 1");
    HASTY_REQUIRE(mode >= 0 && mode <= 3, "invalid Toeplitz apply mode");
    HASTY_REQUIRE(mode != 2 || mul != nullptr, "addcmul needs a multiplier");
    int64_t N[3] = {1, 1, 1}, L[3] = {1, 1, 1};
    for (int d = 0; d < ndim; ++d) {
        N[3 - ndim + d] = nmodes[d];
        L[3 - ndim + d] = 2 * nmodes[d];
    }
    const int64_t nt = ntransf;
    const int64_t nvox = N[0] * N[1] * N[2];
    const int64_t ngrid = L[0] * L[1] * L[2];
    const float scale = (float)(1.0 / (double)ngrid);
    // 1) dim 2 pad-FFT: in [nt][N0][N1][N2] -> st1 [nt][N0][N1][L2]
    {
        AxisFft<float> a;
        a.in = in;
        a.ain = LineAddr{1, N[2], 0, 1, nullptr};
        a.out = st1;
        a.aout = LineAddr{1, L[2], 0, 1, nullptr};
        a.L = (int)L[2];
        a.n_in = (int)N[2];
        a.n_out = (int)L[2];
        a.sign = -1;
        a.nlines = nt * N[0] * N[1];
        run_axis_fft<float>(a, s);
    }
    // 2) dim 1 pad-FFT: st1 -> st2 [nt][N0][L1][L2]
    {
        AxisFft<float> a;
        a.in = st1;
        a.ain = LineAddr{L[2], N[1] * L[2], 1, L[2], nullptr};
        a.out = st2;
        a.aout = LineAddr{L[2], L[1] * L[2], 1, L[2], nullptr};
        a.L = (int)L[1];
        a.n_in = (int)N[1];
        a.n_out = (int)L[1];
        a.sign = -1;
        a.nlines = nt * N[0] * L[2];
        run_axis_fft<float>(a, s);
    }
    // 3) dim 0: pad-FFT, multiply by diag, inverse FFT, crop (in place on st2)
    {
        AxisFft<float> a;
        a.in = st2;
        a.ain = LineAddr{L[1] * L[2], N[0] * L[1] * L[2], 1, L[1] * L[2], nullptr};
        a.out = st2;
        a.aout = a.ain;
        a.L = (int)L[0];
        a.n_in = (int)N[0];
        a.n_out = (int)N[0];
        a.conv_cplx = diag;
        a.aconv = LineAddr{L[1] * L[2], ngrid, 1, L[1] * L[2], nullptr};
        a.nlines = nt * L[1] * L[2];
        run_axis_fft<float>(a, s);
    }
    // 4) dim 1 inverse, crop: st2 [nt][N0][L1][L2] -> st1 [nt][N0][N1][L2]
    {
        AxisFft<float> a;
        a.in = st2;
        a.ain = LineAddr{L[2], L[1] * L[2], 1, L[2], nullptr};
        a.out = st1;
        a.aout = LineAddr{L[2], N[1] * L[2], 1, L[2], nullptr};
        a.L = (int)L[1];
        a.n_in = (int)L[1];
        a.n_out = (int)N[1];
        a.sign = 1;
        a.nlines = nt * N[0] * L[2];
        run_axis_fft<float>(a, s);
    }
    // 5) dim 2 inverse, crop, scale, output mode
    {
        AxisFft<float> a;
        a.in = st1;
        a.ain = LineAddr{1, L[2], 0, 1, nullptr};
        a.out = out;
        a.aout = LineAddr{1, N[2], 0, 1, nullptr};
        a.L = (int)L[2];
        a.n_in = (int)L[2];
        a.n_out = (int)N[2];
        a.sign = 1;
        a.scale = scale;
        a.accumulate = (mode == 1 || mode == 2) ? 1 : 0;
        if (mode == 2) {
            a.outmul = mul;
            a.aomul = a.aout;
        }
        a.nlines = nt * N[0] * N[1];
        run_axis_fft<float>(a, s);
    }
    (void)nvox;
}

int sense_kernel_layout(const int64_t* n3, bool allow_fused)
{
    if (!allow_fused || !sense_fused_supported(n3)) return 0;
    return sense_fused_register_layout(n3) ? 2 : 1;
}

void sense_prepare_kernels(const float* natural, float* prepared, int64_t nk, const int64_t* n3, int layout,
                           hipStream_t s)
{
    const int64_t kvox = 8 * n3[0] * n3[1] * n3[2];
    if (layout == 1 || layout == 2) {
        HASTY_REQUIRE(sense_fused_supported(n3), "fused kernel layout not available for these sizes");
        for (int64_t k = 0; k < nk; ++k)
            sense_fused_permute_kernel(natural + k * kvox, prepared + k * kvox, n3, layout == 2, s);
    } else {
        HASTY_REQUIRE(layout == 0, "unknown kernel layout");
        if (prepared != natural)
            HASTY_HIP_CHECK(hipMemcpyAsync(prepared, natural, sizeof(float) * kvox * nk, hipMemcpyDeviceToDevice, s));
    }
}

size_t sense_normal_workspace_bytes(const int64_t* n3, int coil_batch, int problem_batch)
{
    const int64_t N0 = n3[0], N1 = n3[1], N2 = n3[2];
    const int64_t L1 = 2 * N1, L2 = 2 * N2;
    const int64_t t0 = N0 * N1 * L2;
    const int64_t t1 = N0 * L1 * L2;
    size_t generic = sizeof(float2) * (size_t)coil_batch * (size_t)(t0 + t1);
    size_t fused = sense_fused_supported(n3) ? sense_fused_workspace_bytes(n3, std::max(1, problem_batch), coil_batch) : 0;
    return std::max(generic, fused);
}

void sense_normal_toeplitz(const float2* x, float2* out, const float2* smaps, const int* coils_dev, int ncoil,
                           const float* kernels, int layout, const int* kidx, const float2* sub, int P,
                           const int64_t* n3, void* workspace, size_t ws_bytes, int coil_batch, hipStream_t s,
                           int coil_pstride, const int* pcoils_host, const int* pcoils_dev)
{
    HASTY_REQUIRE(ncoil >= 1 && coil_batch >= 1, "need at least one coil");
    HASTY_REQUIRE(coil_pstride >= 0, "coil stride must be >= 0");
    if (pcoils_host)
        for (int p = 0; p < P; ++p)
            HASTY_REQUIRE(pcoils_host[p] >= 1 && pcoils_host[p] <= ncoil, "per-problem coil count out of range");
    if (layout == 1 || layout == 2) {
        HASTY_REQUIRE(sense_fused_supported(n3), "fused kernel layout not available for these sizes");
        HASTY_REQUIRE(!pcoils_host || pcoils_dev, "the fused layout needs the device copy of the coil counts");
        const int cb = std::min(coil_batch, ncoil);
        const size_t per = sense_fused_workspace_bytes(n3, 1, cb);
        const int pchunk = (int)std::max<size_t>(1, std::min<size_t>(64, ws_bytes / per));
        sense_fused_apply(x, out, smaps, coils_dev, ncoil, kernels, layout == 2, kidx, sub, P, n3, workspace, ws_bytes,
                          pchunk, cb, s, coil_pstride, pcoils_host ? pcoils_dev : nullptr);
        return;
    }
    HASTY_REQUIRE(layout == 0, "unknown kernel layout");
    HASTY_REQUIRE(ws_bytes >= sense_normal_workspace_bytes(n3, std::min(coil_batch, ncoil), 0),
                  "sense normal workspace too small");
    const int cb = std::min(coil_batch, ncoil);
    const int64_t N0 = n3[0], N1 = n3[1], N2 = n3[2];
    const int64_t L0 = 2 * N0, L1 = 2 * N1, L2 = 2 * N2;
    const int64_t nvox = N0 * N1 * N2;
    const int64_t kvox = L0 * L1 * L2;
    float2* T0 = reinterpret_cast<float2*>(workspace);
    float2* T1 = T0 + (int64_t)cb * N0 * N1 * L2;
    for (int p = 0; p < P; ++p) {
        const float2* xp = x + p * nvox;
        float2* op = out + p * nvox;
        const float* kp = kernels + (int64_t)kidx[p] * kvox;
        const float2* smaps_p = smaps + (int64_t)p * coil_pstride * nvox;
        const int ncoil_p = pcoils_host ? pcoils_host[p] : ncoil;
        for (int c0 = 0; c0 < ncoil_p; c0 += cb) {
            const int nc = std::min(cb, ncoil_p - c0);
            // K1: x * S_c, pad-FFT along dim 2 -> T0 [c][N0][N1][L2]
            {
                AxisFft<float> a;
                a.in = xp;
                a.ain = LineAddr{N0 * N1, 0, N2, 1, nullptr};
                a.inmul = smaps_p;
                a.amul = LineAddr{N0 * N1, nvox, N2, 1, coils_dev + c0};
                a.out = T0;
                a.aout = LineAddr{1, L2, 0, 1, nullptr};
                a.L = (int)L2;
                a.n_in = (int)N2;
                a.n_out = (int)L2;
                a.sign = -1;
                a.nlines = (int64_t)nc * N0 * N1;
                run_axis_fft<float>(a, s);
            }
            // K2: pad-FFT along dim 1 -> T1 [c][N0][L1][L2]
            {
                AxisFft<float> a;
                a.in = T0;
                a.ain = LineAddr{L2, N1 * L2, 1, L2, nullptr};
                a.out = T1;
                a.aout = LineAddr{L2, L1 * L2, 1, L2, nullptr};
                a.L = (int)L1;
                a.n_in = (int)N1;
                a.n_out = (int)L1;
                a.sign = -1;
                a.nlines = (int64_t)nc * N0 * L2;
                run_axis_fft<float>(a, s);
            }
            // K3: dim 0 pad-FFT, x kernel, inverse FFT, crop (in place)
            {
                AxisFft<float> a;
                a.in = T1;
                a.ain = LineAddr{L1 * L2, N0 * L1 * L2, 1, L1 * L2, nullptr};
                a.out = T1;
                a.aout = a.ain;
                a.L = (int)L0;
                a.n_in = (int)N0;
                a.n_out = (int)N0;
                a.conv_real = kp;
                a.aconv = LineAddr{L1 * L2, 0, 1, L1 * L2, nullptr};
                a.nlines = (int64_t)nc * L1 * L2;
                run_axis_fft<float>(a, s);
            }
            // K4: dim 1 inverse, crop -> T0 [c][N0][N1][L2]
            {
                AxisFft<float> a;
                a.in = T1;
                a.ain = LineAddr{L2, L1 * L2, 1, L2, nullptr};
                a.out = T0;
                a.aout = LineAddr{L2, N1 * L2, 1, L2, nullptr};
                a.L = (int)L1;
                a.n_in = (int)L1;
                a.n_out = (int)N1;
                a.sign = 1;
                a.nlines = (int64_t)nc * N0 * L2;
                run_axis_fft<float>(a, s);
            }
            // K5: dim 2 inverse, crop, * conj(S_c), accumulate into out (one launch per coil)
            for (int c = 0; c < nc; ++c) {
                AxisFft<float> a;
                a.in = T0 + (int64_t)c * N0 * N1 * L2;
                a.ain = LineAddr{1, L2, 0, 1, nullptr};
                a.out = op;
                a.aout = LineAddr{1, N2, 0, 1, nullptr};
                a.outmul = smaps_p;
                a.aomul = LineAddr{N0 * N1, nvox, N2, 1, coils_dev + c0 + c};
                a.outmul_conj = 1;
                a.accumulate = (c0 + c) > 0 ? 1 : 0;
                a.L = (int)L2;
                a.n_in = (int)L2;
                a.n_out = (int)N2;
                a.sign = 1;
                a.nlines = N0 * N1;
                run_axis_fft<float>(a, s);
            }
        }
        if (sub) launch1d(sub_kernel, nvox, s, op, sub + p * nvox, nvox);
    }
}

}  // namespace hasty_hip
