// LDS-staged mixed-radix (2,3,4,5,8) Stockham FFT passes for gfx950.
//
// One workgroup of 256 threads transforms a bundle of B lines held in LDS (two
// ping-pong buffers, line stride L+1 complex to spread banks, and the line's twiddles). Loads and stores go
// straight between HBM and LDS with the zero-pad / crop / centre mapping, optional
// per-element multipliers and per-element real scales, so pad+deapodise+FFT or
// FFT+crop+coil-multiply never need a separate elementwise pass.
//
// Radix passes follow the self-sorting Stockham formulation of Govindaraju et al.,
// "High performance discrete Fourier transforms on graphics processors" (SC08).
#include <atomic>
#include <map>
#include <mutex>
#include <type_traits>
#include <vector>
#include <cmath>

#include "fft.hpp"

namespace hasty_hip {

namespace {

constexpr int kFftThreads = 256;

__device__ __forceinline__ int64_t line_base(const LineAddr& a, int64_t l)
{
    int64_t q = l / a.inner;
    int64_t r = l - q * a.inner;
    if (a.map) q = a.map[q];
    return q * a.outer + r * a.unit;
}

// Multiply by i^s (s = +-1): s=+1 -> i z, s=-1 -> -i z.
template <int S, typename T> __device__ __forceinline__ T mul_is(T z)
{
    if (S > 0) return T{-z.y, z.x};
    return T{z.y, -z.x};
}

template <int RAD, int S, typename T, typename R> struct Dft;

template <int S, typename T, typename R> struct Dft<2, S, T, R> {
    __device__ __forceinline__ static void run(T* v)
    {
        T a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    }
};

template <int S, typename T, typename R> struct Dft<4, S, T, R> {
    __device__ __forceinline__ static void run(T* v)
    {
        T t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
        T t2 = cadd(v[1], v[3]), t3 = mul_is<S>(csub(v[1], v[3]));
        v[0] = cadd(t0, t2);
        v[2] = csub(t0, t2);
        v[1] = cadd(t1, t3);
        v[3] = csub(t1, t3);
    }
};

template <int S, typename T, typename R> struct Dft<8, S, T, R> {
    __device__ __forceinline__ static void run(T* v)
    {
        T e[4] = {v[0], v[2], v[4], v[6]};
        T o[4] = {v[1], v[3], v[5], v[7]};
        Dft<4, S, T, R>::run(e);
        Dft<4, S, T, R>::run(o);
        const R h = (R)0.70710678118654752440084436210485;
        // w8^k = exp(S 2 pi i k / 8)
        T w1 = T{h * (o[1].x - (R)S * o[1].y), h * (o[1].y + (R)S * o[1].x)};   // (1 + S i) h * o1
        T w2 = mul_is<S>(o[2]);
        T w3 = T{h * (-o[3].x - (R)S * o[3].y), h * (-o[3].y + (R)S * o[3].x)};  // (-1 + S i) h * o3
        v[0] = cadd(e[0], o[0]);
        v[4] = csub(e[0], o[0]);
        v[1] = cadd(e[1], w1);
        v[5] = csub(e[1], w1);
        v[2] = cadd(e[2], w2);
        v[6] = csub(e[2], w2);
        v[3] = cadd(e[3], w3);
        v[7] = csub(e[3], w3);
    }
};

template <int S, typename T, typename R> struct Dft<3, S, T, R> {
    __device__ __forceinline__ static void run(T* v)
    {
        const R c = (R)0.86602540378443864676372317075294;  // sqrt(3)/2
        T t = cadd(v[1], v[2]);
        T d = mul_is<S>(csub(v[1], v[2]));
        T m = T{v[0].x - (R)0.5 * t.x, v[0].y - (R)0.5 * t.y};
        v[0] = cadd(v[0], t);
        v[1] = T{m.x + c * d.x, m.y + c * d.y};
        v[2] = T{m.x - c * d.x, m.y - c * d.y};
    }
};

template <int S, typename T, typename R> struct Dft<5, S, T, R> {
    __device__ __forceinline__ static void run(T* v)
    {
        const R c1 = (R)0.30901699437494742410229341718282;   // cos(2pi/5)
        const R c2 = (R)-0.80901699437494742410229341718282;  // cos(4pi/5)
        const R s1 = (R)0.95105651629515357211643933337938;   // sin(2pi/5)
        const R s2 = (R)0.58778525229247312916870595463907;   // sin(4pi/5)
        T a1 = cadd(v[1], v[4]), b1 = csub(v[1], v[4]);
        T a2 = cadd(v[2], v[3]), b2 = csub(v[2], v[3]);
        T x0 = v[0];
        v[0] = cadd(x0, cadd(a1, a2));
        T p1 = T{x0.x + c1 * a1.x + c2 * a2.x, x0.y + c1 * a1.y + c2 * a2.y};
        T p2 = T{x0.x + c2 * a1.x + c1 * a2.x, x0.y + c2 * a1.y + c1 * a2.y};
        T q1 = mul_is<S>(T{s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y});
        T q2 = mul_is<S>(T{s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y});
        v[1] = cadd(p1, q1);
        v[4] = csub(p1, q1);
        v[2] = cadd(p2, q2);
        v[3] = csub(p2, q2);
    }
};

template <typename R, int RAD, int S>
__device__ __forceinline__ void stockham_pass(const cplx<R>* __restrict__ src, cplx<R>* __restrict__ dst, int L,
                                              int Lp, int Ns, int nb, const cplx<R>* __restrict__ tw, int tid)
{
    const int nbf = L / RAD;
    const int twstep = L / (Ns * RAD);
    for (int b = tid; b < nb * nbf; b += kFftThreads) {
        const int li = b / nbf;
        const int j = b - li * nbf;
        const int k = j % Ns;
        const cplx<R>* s = src + li * Lp + j;
        cplx<R> v[RAD];
        v[0] = s[0];
#pragma unroll
        for (int r = 1; r < RAD; ++r) {
            cplx<R> w = tw[r * k * twstep];
            if (S > 0) w.y = -w.y;
            v[r] = cmul(s[r * nbf], w);
        }
        Dft<RAD, S, cplx<R>, R>::run(v);
        cplx<R>* d = dst + li * Lp + (j / Ns) * Ns * RAD + k;
#pragma unroll
        for (int r = 0; r < RAD; ++r) d[r * Ns] = v[r];
    }
}

template <typename R, int S>
__device__ void fft_bundle(cplx<R>*& src, cplx<R>*& dst, int L, int Lp, int nb, uint64_t radcode, int nrad,
                           const cplx<R>* tw, int tid)
{
    int Ns = 1;
    for (int p = 0; p < nrad; ++p) {
        const int rad = (int)((radcode >> (4 * p)) & 15u);
        switch (rad) {
        case 2: stockham_pass<R, 2, S>(src, dst, L, Lp, Ns, nb, tw, tid); break;
        case 3: stockham_pass<R, 3, S>(src, dst, L, Lp, Ns, nb, tw, tid); break;
        case 4: stockham_pass<R, 4, S>(src, dst, L, Lp, Ns, nb, tw, tid); break;
        case 5: stockham_pass<R, 5, S>(src, dst, L, Lp, Ns, nb, tw, tid); break;
        default: stockham_pass<R, 8, S>(src, dst, L, Lp, Ns, nb, tw, tid); break;
        }
        __syncthreads();
        cplx<R>* t = src;
        src = dst;
        dst = t;
        Ns *= rad;
    }
}

struct RadixPlan {
    uint64_t code = 0;
    int n = 0;
};

// Input position i (0..L-1) -> input element n, or -1 for a padded zero.
__device__ __forceinline__ int in_element(int i, int L, int n_in, int centered)
{
    if (!centered) return i < n_in ? i : -1;
    const int h = n_in / 2;
    if (i < n_in - h) return i + h;
    if (i >= L - h) return i - (L - h);
    return -1;
}

template <typename R> struct AxisKernelArgs {
    AxisFft<R> a;
    const cplx<R>* tw;
    uint64_t radcode;
    int nrad;
    int lines_per_block;
    int stage_tw;  // twiddles staged in LDS
};

// MODE 0: plain FFT with sign a.sign; MODE 1: conv with real kernel; MODE 2: conv with
// complex kernel (forward FFT, multiply, inverse FFT).
template <typename R, int MODE>
__global__ void __launch_bounds__(kFftThreads) axis_fft_kernel(AxisKernelArgs<R> args)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const AxisFft<R>& a = args.a;
    const int L = a.L;
    const int Lp = L + 1;
    const int B = args.lines_per_block;
    const int tid = threadIdx.x;
    const int64_t l0 = (int64_t)blockIdx.x * B;
    const int nb = (int)min((int64_t)B, a.nlines - l0);

    cplx<R>* bufA = reinterpret_cast<cplx<R>*>(smem_raw);
    cplx<R>* bufB = bufA + B * Lp;
    // the L twiddles, staged with the data when they fit in LDS
    cplx<R>* tws = bufB + B * Lp;
    const cplx<R>* twl = args.stage_tw ? tws : args.tw;
    int64_t* bases = reinterpret_cast<int64_t*>(tws + (args.stage_tw ? L : 0));  // 5 * B entries
    int64_t* b_in = bases;
    int64_t* b_mul = bases + B;
    int64_t* b_out = bases + 2 * B;
    int64_t* b_omul = bases + 3 * B;
    int64_t* b_conv = bases + 4 * B;

    for (int li = tid; li < nb; li += kFftThreads) {
        const int64_t l = l0 + li;
        b_in[li] = line_base(a.ain, l);
        b_out[li] = line_base(a.aout, l);
        if (a.inmul) b_mul[li] = line_base(a.amul, l);
        if (a.outmul) b_omul[li] = line_base(a.aomul, l);
        if (MODE != 0) b_conv[li] = line_base(a.aconv, l);
    }
    __syncthreads();

    if (args.stage_tw)
        for (int k = tid; k < L; k += kFftThreads) tws[k] = args.tw[k];
    // Load (pad / centre / multiply / scale). Each thread issues kLoadUnroll independent
    // loads before its first LDS store, so a workgroup keeps kLoadUnroll x 256 element
    // loads in flight instead of 256.
    constexpr int kLoadUnroll = sizeof(R) == 4 ? 8 : 4;
    const bool in_fast = a.ain.stride != 1;
    const int total = nb * L;
    for (int e0 = tid; e0 < total; e0 += kFftThreads * kLoadUnroll) {
        cplx<R> v[kLoadUnroll];
        int slot[kLoadUnroll];
#pragma unroll
        for (int u = 0; u < kLoadUnroll; ++u) {
            const int e = e0 + u * kFftThreads;
            v[u] = cplx<R>{(R)0, (R)0};
            slot[u] = -1;
            if (e < total) {
                int li, i;
                if (in_fast) {
                    li = e % nb;
                    i = e / nb;
                } else {
                    li = e / L;
                    i = e - li * L;
                }
                slot[u] = li * Lp + i;
                const int n = in_element(i, L, a.n_in, a.in_centered);
                if (n >= 0) {
                    cplx<R> x = a.in[b_in[li] + (int64_t)n * a.ain.stride];
                    if (a.inmul) {
                        const cplx<R> m = a.inmul[b_mul[li] + (int64_t)n * a.amul.stride];
                        x = a.inmul_conj ? cmulc(m, x) : cmul(m, x);
                    }
                    if (a.pre_scale) x = cscale(x, a.pre_scale[n]);
                    v[u] = x;
                }
            }
        }
#pragma unroll
        for (int u = 0; u < kLoadUnroll; ++u)
            if (slot[u] >= 0) bufA[slot[u]] = v[u];
    }
    __syncthreads();

    cplx<R>* src = bufA;
    cplx<R>* dst = bufB;
    if (MODE == 0) {
        if (a.sign < 0)
            fft_bundle<R, -1>(src, dst, L, Lp, nb, args.radcode, args.nrad, twl, tid);
        else
            fft_bundle<R, 1>(src, dst, L, Lp, nb, args.radcode, args.nrad, twl, tid);
    } else {
        fft_bundle<R, -1>(src, dst, L, Lp, nb, args.radcode, args.nrad, twl, tid);
        for (int e = tid; e < nb * L; e += kFftThreads) {
            const int li = e / L;
            const int i = e - li * L;
            const int64_t kidx = b_conv[li] + (int64_t)i * a.aconv.stride;
            cplx<R> v = src[li * Lp + i];
            if (MODE == 1)
                v = cscale(v, a.conv_real[kidx]);
            else
                v = cmul(v, a.conv_cplx[kidx]);
            src[li * Lp + i] = v;
        }
        __syncthreads();
        fft_bundle<R, 1>(src, dst, L, Lp, nb, args.radcode, args.nrad, twl, tid);
    }

    // Store (crop / centre / scale / multiply / accumulate).
    const bool out_fast = a.aout.stride != 1;
    for (int e = tid; e < nb * a.n_out; e += kFftThreads) {
        int li, n;
        if (out_fast) {
            li = e % nb;
            n = e / nb;
        } else {
            li = e / a.n_out;
            n = e - li * a.n_out;
        }
        const int i = a.out_centered ? (n - a.n_out / 2 + L) % L : n;
        cplx<R> v = cscale(src[li * Lp + i], a.scale);
        if (a.post_scale) v = cscale(v, a.post_scale[n]);
        if (a.outmul) {
            const cplx<R> m = a.outmul[b_omul[li] + (int64_t)n * a.aomul.stride];
            v = a.outmul_conj ? cmulc(m, v) : cmul(m, v);
        }
        const int64_t addr = b_out[li] + (int64_t)n * a.aout.stride;
        if (a.accumulate) v = cadd(a.out[addr], v);
        a.out[addr] = v;
    }
}

// ---- complex64 lines of 16 R (R = 8, 12, 16, 20) as a 16 x R four-step FFT in registers -
//
// 16 lines per workgroup. With n = g + 16 m (g < 16, m < R) and k = k2 + R k1 (k2 < R,
// k1 < 16), X[k] = sum_g W16^(g k1) W_L^(g k2) sum_m x[g + 16 m] W_R^(m k2): thread g of a
// line holds x[g + 16 m], takes the R-point DFT over m and the twiddle W_L^(g k2), the
// line passes through LDS (row pitch 17), and thread k2 takes the 16-point DFT over g,
// holding X[k2 + R k1]. The workgroup has 16 max(16, R) threads: 16 per line in the first
// half, R per line in the second. Contiguous lines give a line's threads adjacent lanes
// (each load and store covers consecutive elements); strided lines give the 16 lines of
// one g or k2 adjacent lanes (each covers 16 adjacent lines). The pad / centre / scale /
// crop mapping is the general kernel's; lines with coil multipliers, accumulation or a
// convolution take the general kernel.
constexpr int kR16Lines = 16;

// DFT of size A B with sign S in registers, natural order in and out: n = A n2 + n1...
// written as n = B n1 + n2 (n1 < A, n2 < B), k = k1 + A k2: A-point DFTs over n1, the
// twiddles W_AB^(n2 k1), B-point DFTs over n2.
template <int A, int B, int S> __device__ __forceinline__ void dft_ab(float2* v)
{
    float2 t[B][A];
#pragma unroll
    for (int n2 = 0; n2 < B; ++n2) {
        float2 u[A];
#pragma unroll
        for (int n1 = 0; n1 < A; ++n1) u[n1] = v[B * n1 + n2];
        Dft<A, S, float2, float>::run(u);
#pragma unroll
        for (int k1 = 0; k1 < A; ++k1) {
            if (n2 * k1 == 0) {
                t[n2][k1] = u[k1];
            } else {
                const float ang = (float)S * 6.28318530717958648f * (float)(n2 * k1) / (float)(A * B);
                t[n2][k1] = cmul(u[k1], float2{cosf(ang), sinf(ang)});
            }
        }
    }
#pragma unroll
    for (int k1 = 0; k1 < A; ++k1) {
        float2 u[B];
#pragma unroll
        for (int n2 = 0; n2 < B; ++n2) u[n2] = t[n2][k1];
        Dft<B, S, float2, float>::run(u);
#pragma unroll
        for (int k2 = 0; k2 < B; ++k2) v[k1 + A * k2] = u[k2];
    }
}

template <int R, int S> __device__ __forceinline__ void dft_r(float2* v)
{
    if constexpr (R == 8) Dft<8, S, float2, float>::run(v);
    else if constexpr (R == 12) dft_ab<4, 3, S>(v);
    else if constexpr (R == 16) dft_ab<4, 4, S>(v);
    else dft_ab<4, 5, S>(v);  // 20
}

template <int R, int S>
__global__ void __launch_bounds__(kR16Lines * (R > 16 ? R : 16)) axis_fft_r16_kernel(AxisKernelArgs<float> args)
{
    constexpr int L = 16 * R;
    __shared__ float2 tr[kR16Lines * R * 17];
    __shared__ float2 tws[L];
    const AxisFft<float>& a = args.a;
    const int tid = threadIdx.x;
    // Transpose layout: element (line, g, k2) at (line R + k2) 17 + g when the lanes of a
    // wave hold one line's g (contiguous input), at (k2 16 + g) 17 + line when they hold 16
    // lines (strided input). With input and output both contiguous or both strided, a
    // half-wave's 8-byte accesses then fall on distinct banks (the strided layout's writes
    // share one bank pair); the old layout gave strided lines 8-way conflicts, a line step
    // being R 17 float2 = 32 banks mod 64.
    const bool lines_fast = a.ain.stride != 1;
    for (int k = tid; k < L; k += kR16Lines * (R > 16 ? R : 16)) tws[k] = args.tw[k];
    float2 v[R > 16 ? R : 16];
    if (tid < 256) {
        int li, g;
        if (a.ain.stride != 1) {
            g = tid >> 4;
            li = tid & 15;
        } else {
            li = tid >> 4;
            g = tid & 15;
        }
        const int64_t l = (int64_t)blockIdx.x * kR16Lines + li;
        const bool lv = l < a.nlines;
        const int64_t bin = lv ? line_base(a.ain, l) : 0;
#pragma unroll
        for (int m = 0; m < R; ++m) {
            const int n = in_element(g + 16 * m, L, a.n_in, a.in_centered);
            float2 x = float2{0.f, 0.f};
            if (lv && n >= 0) {
                x = a.in[bin + (int64_t)n * a.ain.stride];
                if (a.pre_scale) x = cscale(x, a.pre_scale[n]);
            }
            v[m] = x;
        }
        dft_r<R, S>(v);
    }
    __syncthreads();  // tws
    if (tid < 256) {
        int li, g;
        if (a.ain.stride != 1) {
            g = tid >> 4;
            li = tid & 15;
        } else {
            li = tid >> 4;
            g = tid & 15;
        }
#pragma unroll
        for (int k2 = 1; k2 < R; ++k2) {
            float2 w = tws[g * k2];
            if (S > 0) w.y = -w.y;
            v[k2] = cmul(v[k2], w);
        }
        if (lines_fast) {
#pragma unroll
            for (int k2 = 0; k2 < R; ++k2) tr[(k2 * 16 + g) * 17 + li] = v[k2];
        } else {
#pragma unroll
            for (int k2 = 0; k2 < R; ++k2) tr[(li * R + k2) * 17 + g] = v[k2];
        }
    }
    __syncthreads();
    if (tid >= 16 * R) return;
    int lo, k2;
    if (a.aout.stride != 1) {
        k2 = tid >> 4;
        lo = tid & 15;
    } else {
        lo = tid / R;
        k2 = tid - lo * R;
    }
    if (lines_fast) {
#pragma unroll
        for (int m = 0; m < 16; ++m) v[m] = tr[(k2 * 16 + m) * 17 + lo];
    } else {
#pragma unroll
        for (int m = 0; m < 16; ++m) v[m] = tr[(lo * R + k2) * 17 + m];
    }
    dft_r<16, S>(v);
    const int64_t l2 = (int64_t)blockIdx.x * kR16Lines + lo;
    if (l2 >= a.nlines) return;
    const int64_t bout = line_base(a.aout, l2);
    const int h = a.n_out / 2;
#pragma unroll
    for (int k1 = 0; k1 < 16; ++k1) {
        const int k = k2 + R * k1;
        int n = k;
        if (a.out_centered) {
            n = k + h;  // inverse of k = (n - h) mod L
            if (n >= L) n -= L;
        }
        if (n < a.n_out) {
            float2 y = cscale(v[k1], a.scale);
            if (a.post_scale) y = cscale(y, a.post_scale[n]);
            a.out[bout + (int64_t)n * a.aout.stride] = y;
        }
    }
}

template <int R> void launch_r16(const AxisKernelArgs<float>& args, int64_t blocks, hipStream_t stream)
{
    const unsigned threads = kR16Lines * (R > 16 ? R : 16);
    if (args.a.sign < 0)
        hipLaunchKernelGGL((axis_fft_r16_kernel<R, -1>), dim3((unsigned)blocks), dim3(threads), 0, stream, args);
    else
        hipLaunchKernelGGL((axis_fft_r16_kernel<R, 1>), dim3((unsigned)blocks), dim3(threads), 0, stream, args);
}

RadixPlan make_radix_plan(int L)
{
    RadixPlan p;
    int m = L;
    const int order[5] = {8, 4, 2, 3, 5};
    for (int oi = 0; oi < 5; ++oi) {
        const int r = order[oi];
        while (m % r == 0) {
            if (p.n >= 16) throw HastyError(kUnsupported, "FFT length has too many factors");
            p.code |= (uint64_t)r << (4 * p.n);
            ++p.n;
            m /= r;
        }
    }
    if (m != 1) throw HastyError(kUnsupported, "FFT length " + std::to_string(L) + " is not 2,3,5-smooth");
    return p;
}

std::atomic<int> g_strided_lines{0};  // lines per workgroup of a strided pass, 0 = auto
std::atomic<int> g_r16{1};            // L = 256 complex64 lines by the register four-step kernel

std::mutex g_tw_mutex;
std::map<std::tuple<int, int, int>, void*> g_tw_cache;  // (device, L, sizeof(R)) -> table

}  // namespace

void fft_set_r16(int v) { g_r16 = v; }
int fft_get_r16() { return g_r16; }
void fft_set_strided_lines(int v) { g_strided_lines = v; }
int fft_get_strided_lines() { return g_strided_lines; }

bool fft_size_supported(int L)
{
    if (L < 1 || L > 8192) return false;
    int m = L;
    for (int r : {2, 3, 5})
        while (m % r == 0) m /= r;
    return m == 1;
}

int64_t next235even(int64_t n)
{
    if (n <= 2) return 2;
    if (n % 2) ++n;
    for (;; n += 2) {
        int64_t m = n;
        for (int p : {2, 3, 5})
            while (m % p == 0) m /= p;
        if (m == 1) return n;
    }
}

template <typename R> const cplx<R>* twiddles(int L)
{
    const int dev = current_device();
    std::lock_guard<std::mutex> lock(g_tw_mutex);
    auto key = std::make_tuple(dev, L, (int)sizeof(R));
    auto it = g_tw_cache.find(key);
    if (it != g_tw_cache.end()) return reinterpret_cast<const cplx<R>*>(it->second);
    std::vector<cplx<R>> host(L);
    const long double twopi = 6.283185307179586476925286766559L;
    for (int k = 0; k < L; ++k) {
        const long double ang = -twopi * (long double)k / (long double)L;
        host[k] = cplx<R>{(R)std::cos(ang), (R)std::sin(ang)};
    }
    void* d = dev_alloc(sizeof(cplx<R>) * L);
    HASTY_HIP_CHECK(hipMemcpy(d, host.data(), sizeof(cplx<R>) * L, hipMemcpyHostToDevice));
    g_tw_cache[key] = d;
    return reinterpret_cast<const cplx<R>*>(d);
}

template <typename R> void run_axis_fft(const AxisFft<R>& a, hipStream_t stream)
{
    if (a.nlines <= 0 || a.n_out <= 0) return;
    HASTY_REQUIRE(a.L >= 1 && fft_size_supported(a.L), "unsupported FFT length " + std::to_string(a.L));
    HASTY_REQUIRE(a.n_in >= 0 && a.n_in <= a.L && a.n_out <= a.L, "FFT pad/crop sizes exceed the line length");
    const RadixPlan rp = make_radix_plan(a.L);
    AxisKernelArgs<R> args;
    args.a = a;
    args.tw = twiddles<R>(a.L);
    args.radcode = rp.code;
    args.nrad = rp.n;
    const int target = sizeof(R) == 4 ? 2048 : 1024;
    int B = target / a.L;
    B = B < 1 ? 1 : (B > 64 ? 64 : B);
    // Strided lines (adjacent lines adjacent in memory: a bundle of B lines reads B
    // consecutive elements of each row) may take another bundle size. Wider bundles read
    // whole 128-byte L2 lines, but at L = 256 they ran slower (66, 86 and 128 us per
    // 256^3 pass at 8, 16 and 32 lines): fewer workgroups fit a CU's LDS, and the pass is
    // latency-bound, not bound by the re-read halves of L2 lines.
    const int want = g_strided_lines.load();
    const bool strided = a.ain.stride != 1 || a.aout.stride != 1;
    if (want > 0 && strided) B = want;
    // complex128 strided lines of 512 / 640 (the Toeplitz kernel build) get 2 / 1 lines
    // from the rule above; 4 built the 128^3 and 160^3 kernels 3-4 % faster
    if (want == 0 && strided && sizeof(R) == 8 && B < 4) B = 4;
    auto lds_bytes = [&](int b, bool tw) {
        return sizeof(cplx<R>) * (2 * (size_t)b * (a.L + 1) + (tw ? a.L : 0)) + sizeof(int64_t) * 5 * b;
    };
    while (B > 1 && lds_bytes(B, false) > 160 * 1024) B /= 2;
    args.lines_per_block = B;
    args.stage_tw = lds_bytes(B, true) <= 160 * 1024 ? 1 : 0;
    const size_t smem = lds_bytes(B, args.stage_tw != 0);
    const int64_t blocks = ceil_div(a.nlines, B);
    HASTY_REQUIRE(blocks < (1ll << 31), "too many FFT lines for one launch");
    const int mode = a.conv_real ? 1 : (a.conv_cplx ? 2 : 0);
    KernelTimer t("fft_axis", stream);
    if constexpr (std::is_same<R, float>::value) {
        const bool reg = a.L == 128 || a.L == 192 || a.L == 256 || a.L == 320;
        if (g_r16.load() && reg && mode == 0 && !a.inmul && !a.outmul && !a.accumulate) {
            const int64_t rblocks = ceil_div(a.nlines, kR16Lines);
            HASTY_REQUIRE(rblocks < (1ll << 31), "too many FFT lines for one launch");
            switch (a.L / 16) {
            case 8: launch_r16<8>(args, rblocks, stream); break;
            case 12: launch_r16<12>(args, rblocks, stream); break;
            case 16: launch_r16<16>(args, rblocks, stream); break;
            default: launch_r16<20>(args, rblocks, stream); break;
            }
            check_launch("axis_fft_r16_kernel");
            return;
        }
    }
    if (mode == 0) {
        allow_lds(axis_fft_kernel<R, 0>, smem);
        hipLaunchKernelGGL((axis_fft_kernel<R, 0>), dim3((unsigned)blocks), dim3(kFftThreads), smem, stream, args);
    } else if (mode == 1) {
        allow_lds(axis_fft_kernel<R, 1>, smem);
        hipLaunchKernelGGL((axis_fft_kernel<R, 1>), dim3((unsigned)blocks), dim3(kFftThreads), smem, stream, args);
    } else {
        allow_lds(axis_fft_kernel<R, 2>, smem);
        hipLaunchKernelGGL((axis_fft_kernel<R, 2>), dim3((unsigned)blocks), dim3(kFftThreads), smem, stream, args);
    }
    check_launch("axis_fft_kernel");
}

template const cplx<float>* twiddles<float>(int);
template const cplx<double>* twiddles<double>(int);
template void run_axis_fft<float>(const AxisFft<float>&, hipStream_t);
template void run_axis_fft<double>(const AxisFft<double>&, hipStream_t);

}  // namespace hasty_hip
