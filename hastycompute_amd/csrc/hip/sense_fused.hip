// Fused Toeplitz SENSE normal operator for N1 = N2 = 32*NB (NB = 1..5), any N0 whose
// z line 2 N0 is 2,3,5-smooth and fits the z-convolution's LDS.
//
// For every problem p and coil c the operator is
//   out[p] += conj(S_c) * crop( IFFT3( K[p] * FFT3( pad(S_c * x[p]) ) ) ),
// a zero-padded (2N)^3 FFT pair around a real kernel. Three kernels move each coil's
// oversampled grid through HBM once (one half-size write+read twice), instead of the
// five strided passes of the generic path:
//
//  A  fwd_xy   one workgroup per (problem, z-plane): x[p] plane in registers; per coil
//              multiply by S_c, then the padded 2-D FFT of the N x N plane to L x L,
//              computed as P = L/32 cosets of the x-frequency: for coset s the 32
//              frequencies k = P m + s are a 32-point DFT of the x-row folded modulo 32
//              after the twiddle exp(-2 pi i s n / L) (one wavefront shuffle FFT per
//              row), followed by the L-point y FFT of the 32 columns in LDS. Writes
//              T[p][c][z][s][ky][m] (N0 planes of L*L, i.e. half the oversampled grid).
//  B  z_conv   one workgroup per bundle of 32 (ky, kx) columns: loads the kernel
//              column bundle once, then for each coil: z FFT (N0 -> L0 padded), times
//              the real kernel, inverse z FFT, crop to N0, in place in T.
//  C  inv_xy   one workgroup per (problem, z-plane): for each coil and coset, the
//              L-point inverse y FFT (cropped to N1 rows) in LDS, a 32-point inverse
//              wavefront FFT along m, the coset twiddle exp(+2 pi i s n / L), the
//              conj(S_c) multiply and the coil sum, all accumulated in registers; the
//              output plane is written once.
// The kernel K is stored in the same permuted (s, ky, m) order as T.
#include <atomic>
#include <vector>

#include "common.hpp"
#include "fft.hpp"

namespace hasty_hip {

namespace {

constexpr int kW = 32;           // coset width (x frequencies per coset)
constexpr int kXY = 1024;        // threads of the plane kernels
constexpr int kZ = 1024;         // threads of the z kernel

__device__ __forceinline__ int brev5(int b) { return (int)(__brev((unsigned)b) >> 27); }

// Workgroup barrier for LDS data exchange only: waits for this wave's LDS operations,
// not for its outstanding global stores. Threads of a workgroup here never read global
// data another thread of the same workgroup wrote, so stores of one coset (or coil) can
// drain while the next one is computed.
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// exp(sign 2 pi i k / n) from a forward table of length n (tw[k] = exp(-2 pi i k / n)).
template <int S> __device__ __forceinline__ float2 twid(const float2* tw, int k)
{
    float2 w = tw[k];
    if (S > 0) w.y = -w.y;
    return w;
}

__device__ __forceinline__ float2 shfl_xor2(float2 v, int m)
{
    return float2{__shfl_xor(v.x, m, 32), __shfl_xor(v.y, m, 32)};
}

// Forward 32-point DFT across the 32 lanes of a half-wave (lane b holds x[b]); radix-2
// decimation in frequency; on return lane b holds X[brev5(b)]. tw32: exp(-2 pi i k/32).
__device__ __forceinline__ float2 dft32_fwd(float2 v, int b, const float2* tw32)
{
#pragma unroll
    for (int h = 16; h >= 1; h >>= 1) {
        const float2 o = shfl_xor2(v, h);
        if (b & h) {
            // (x[b-h] - x[b]) * w_{2h}^{(b mod h)}, w_{2h} = exp(-2 pi i / 2h) = tw32[16/h]
            const float2 d = csub(o, v);
            v = cmul(d, tw32[(b & (h - 1)) * (16 / h)]);
        } else {
            v = cadd(v, o);
        }
    }
    return v;
}

// Inverse (sign +1) 32-point DFT across lanes: lane b holds X[brev5(b)] on entry
// (bit-reversed input), radix-2 decimation in time; on return lane b holds x[b].
__device__ __forceinline__ float2 dft32_inv_bitrev_in(float2 v, int b, const float2* tw32)
{
#pragma unroll
    for (int h = 1; h <= 16; h <<= 1) {
        const float2 o = shfl_xor2(v, h);
        // iterative radix-2 DIT on bit-reversed input: pair (g + j, g + j + h) of a 2h
        // group, u = a[g+j], t = w a[g+j+h], outputs u + t and u - t
        const int j = (b & (h - 1));
        float2 w = tw32[j * (16 / h)];
        w.y = -w.y;  // exp(+2 pi i j / 2h)
        if (b & h) {
            v = csub(o, cmul(w, v));
        } else {
            v = cadd(v, cmul(w, o));
        }
    }
    return v;
}

// Value of lane (lane ^ H) within the half-wave, without LDS traffic where the DPP
// forms reach: quad_perm for H = 1, 2, row_ror:8 for H = 8 (xor 8 inside a 16-lane row),
// ds_swizzle (bitmask mode, xor_mask = H) for H = 4, 16.
template <int H> __device__ __forceinline__ float xlane(float v)
{
    const int iv = __builtin_bit_cast(int, v);
    int r;
    if constexpr (H == 1) r = __builtin_amdgcn_update_dpp(0, iv, 0xB1, 0xF, 0xF, false);   // quad_perm [1,0,3,2]
    else if constexpr (H == 2) r = __builtin_amdgcn_update_dpp(0, iv, 0x4E, 0xF, 0xF, false);  // quad_perm [2,3,0,1]
    else if constexpr (H == 8) r = __builtin_amdgcn_update_dpp(0, iv, 0x128, 0xF, 0xF, false); // row_ror:8
    else r = __builtin_amdgcn_ds_swizzle(iv, 0x1F | (H << 10));
    return __builtin_bit_cast(float, r);
}
template <int H> __device__ __forceinline__ float2 xlane2(float2 v) { return float2{xlane<H>(v.x), xlane<H>(v.y)}; }

// Per-lane constants of the 32-point lane DFTs. A butterfly stage over lane pairs
// (b, b ^ h) is written without selects: each lane holds a sign, +1 on the lower lane
// of its pair and -1 on the upper one, and a twiddle that is 1 where the stage applies
// none, so every lane runs the same multiply-adds.
//  sgn[l]: sign of the stage with h = 1 << l;
//  fwd[i]: exp(-2 pi i (b mod h) / 2h) on the upper lanes (b & h), 1 on the lower ones,
//          for the forward stage with h = 16 >> i;
//  inv[i]: exp(+2 pi i (b mod h) / 2h) on the upper lanes, 1 on the lower ones, for the
//          inverse stage with h = 1 << i.
struct LaneTw32 {
    float sgn[5];
    float2 fwd[5];
    float2 inv[5];
};
__device__ __forceinline__ LaneTw32 lane_tw32(int b, const float2* tw32)
{
    LaneTw32 t;
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        t.sgn[i] = (b & (1 << i)) ? -1.f : 1.f;
        const int h = 16 >> i;
        t.fwd[i] = (b & h) ? tw32[(b & (h - 1)) * (16 / h)] : float2{1.f, 0.f};
        const int hi = 1 << i;
        const float2 w = tw32[(b & (hi - 1)) * (16 / hi)];
        t.inv[i] = (b & hi) ? float2{w.x, -w.y} : float2{1.f, 0.f};
    }
    return t;
}

__device__ __forceinline__ float2 sfma(float s, float2 v, float2 o)
{
    return float2{fmaf(s, v.x, o.x), fmaf(s, v.y, o.y)};
}

// Decimation-in-frequency 32-point DFT across the lanes of a half-wave, natural order in,
// bit-reversed order out. Stage h: t = o + v (lower lane) or o - v (upper lane), then
// times the lane's twiddle. The last stage (h = 1) has twiddle 1 on every lane.
__device__ __forceinline__ float2 dft32_fwd_r(float2 v, const LaneTw32& lt)
{
#define HASTY_DIF_STAGE(I, L)                                        \
    {                                                                \
        const float2 o = xlane2<(1 << L)>(v);                        \
        v = cmul(sfma(lt.sgn[L], v, o), lt.fwd[I]);                  \
    }
    HASTY_DIF_STAGE(0, 4)
    HASTY_DIF_STAGE(1, 3)
    HASTY_DIF_STAGE(2, 2)
    HASTY_DIF_STAGE(3, 1)
#undef HASTY_DIF_STAGE
    const float2 o = xlane2<1>(v);
    return sfma(lt.sgn[0], v, o);
}

// Decimation-in-time 32-point inverse DFT across a half-wave, bit-reversed order in,
// natural order out. Stage h: each lane first scales its value by its twiddle (1 on the
// lower lane), then v = y + o (lower) or o - y (upper), which is v + w o on the lower
// lane and v_lower - w v on the upper one. The first stage (h = 1) has twiddle 1.
__device__ __forceinline__ float2 dft32_inv_r(float2 v, const LaneTw32& lt)
{
    {
        const float2 o = xlane2<1>(v);
        v = sfma(lt.sgn[0], v, o);
    }
#define HASTY_DIT_STAGE(L)                                           \
    {                                                                \
        const float2 y = cmul(v, lt.inv[L]);                         \
        const float2 o = xlane2<(1 << L)>(y);                        \
        v = sfma(lt.sgn[L], y, o);                                   \
    }
    HASTY_DIT_STAGE(1)
    HASTY_DIT_STAGE(2)
    HASTY_DIT_STAGE(3)
    HASTY_DIT_STAGE(4)
#undef HASTY_DIT_STAGE
    return v;
}

template <int RAD, int S> __device__ __forceinline__ void dftR(float2* v);

template <int S> __device__ __forceinline__ float2 mis(float2 z)
{
    if (S > 0) return float2{-z.y, z.x};
    return float2{z.y, -z.x};
}
template <> __device__ __forceinline__ void dftR<2, -1>(float2* v) { float2 a = v[0]; v[0] = cadd(a, v[1]); v[1] = csub(a, v[1]); }
template <> __device__ __forceinline__ void dftR<2, 1>(float2* v) { float2 a = v[0]; v[0] = cadd(a, v[1]); v[1] = csub(a, v[1]); }
template <int S> __device__ __forceinline__ void dft4(float2* v)
{
    float2 t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
    float2 t2 = cadd(v[1], v[3]), t3 = mis<S>(csub(v[1], v[3]));
    v[0] = cadd(t0, t2);
    v[2] = csub(t0, t2);
    v[1] = cadd(t1, t3);
    v[3] = csub(t1, t3);
}
// dft4 of (v[0], v[1], 0, 0): the first stage of a DFT whose upper half of inputs is zero
template <int S> __device__ __forceinline__ void dft4_half(float2* v)
{
    const float2 a = v[0], b = v[1], mb = mis<S>(v[1]);
    v[0] = cadd(a, b);
    v[2] = csub(a, b);
    v[1] = cadd(a, mb);
    v[3] = csub(a, mb);
}
// with HALF the upper half of a DFT's inputs is zero and is not read
template <int S, bool HALF> __device__ __forceinline__ void dft4_first(float2* v)
{
    if constexpr (HALF) dft4_half<S>(v);
    else dft4<S>(v);
}
template <> __device__ __forceinline__ void dftR<4, -1>(float2* v) { dft4<-1>(v); }
template <> __device__ __forceinline__ void dftR<4, 1>(float2* v) { dft4<1>(v); }
template <int S, bool HALF = false> __device__ __forceinline__ void dft8(float2* v)
{
    float2 e[4] = {v[0], v[2], HALF ? float2{0.f, 0.f} : v[4], HALF ? float2{0.f, 0.f} : v[6]};
    float2 o[4] = {v[1], v[3], HALF ? float2{0.f, 0.f} : v[5], HALF ? float2{0.f, 0.f} : v[7]};
    dft4_first<S, HALF>(e);
    dft4_first<S, HALF>(o);
    const float h = 0.70710678118654752440f;
    float2 w1 = float2{h * (o[1].x - (float)S * o[1].y), h * (o[1].y + (float)S * o[1].x)};
    float2 w2 = mis<S>(o[2]);
    float2 w3 = float2{h * (-o[3].x - (float)S * o[3].y), h * (-o[3].y + (float)S * o[3].x)};
    v[0] = cadd(e[0], o[0]);
    v[4] = csub(e[0], o[0]);
    v[1] = cadd(e[1], w1);
    v[5] = csub(e[1], w1);
    v[2] = cadd(e[2], w2);
    v[6] = csub(e[2], w2);
    v[3] = cadd(e[3], w3);
    v[7] = csub(e[3], w3);
}
template <> __device__ __forceinline__ void dftR<8, -1>(float2* v) { dft8<-1>(v); }
template <> __device__ __forceinline__ void dftR<8, 1>(float2* v) { dft8<1>(v); }
template <int S> __device__ __forceinline__ void dft3(float2* v)
{
    const float c = 0.86602540378443864676f;
    float2 t = cadd(v[1], v[2]);
    float2 d = mis<S>(csub(v[1], v[2]));
    float2 m = float2{v[0].x - 0.5f * t.x, v[0].y - 0.5f * t.y};
    v[0] = cadd(v[0], t);
    v[1] = float2{m.x + c * d.x, m.y + c * d.y};
    v[2] = float2{m.x - c * d.x, m.y - c * d.y};
}
template <> __device__ __forceinline__ void dftR<3, -1>(float2* v) { dft3<-1>(v); }
template <> __device__ __forceinline__ void dftR<3, 1>(float2* v) { dft3<1>(v); }
template <int S> __device__ __forceinline__ void dft5(float2* v)
{
    const float c1 = 0.30901699437494742410f, c2 = -0.80901699437494742410f;
    const float s1 = 0.95105651629515357212f, s2 = 0.58778525229247312917f;
    float2 a1 = cadd(v[1], v[4]), b1 = csub(v[1], v[4]);
    float2 a2 = cadd(v[2], v[3]), b2 = csub(v[2], v[3]);
    float2 x0 = v[0];
    v[0] = cadd(x0, cadd(a1, a2));
    float2 p1 = float2{x0.x + c1 * a1.x + c2 * a2.x, x0.y + c1 * a1.y + c2 * a2.y};
    float2 p2 = float2{x0.x + c2 * a1.x + c1 * a2.x, x0.y + c2 * a1.y + c1 * a2.y};
    float2 q1 = mis<S>(float2{s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y});
    float2 q2 = mis<S>(float2{s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y});
    v[1] = cadd(p1, q1);
    v[4] = csub(p1, q1);
    v[2] = cadd(p2, q2);
    v[3] = csub(p2, q2);
}
template <> __device__ __forceinline__ void dftR<5, -1>(float2* v) { dft5<-1>(v); }
template <> __device__ __forceinline__ void dftR<5, 1>(float2* v) { dft5<1>(v); }

// One Stockham pass over NCOL interleaved columns of length L held in LDS as
// [row][NCOL] (column c, element e at src[e*NCOL + c]); src and dst distinct.
// `zero_from` (first pass only): input rows >= zero_from are known zeros.
template <int RAD, int S, int NCOL, int NT>
__device__ __forceinline__ void col_pass(const float2* __restrict__ src, float2* __restrict__ dst, int L, int Ns,
                                         const float2* __restrict__ twL, int zero_from)
{
    const int nbf = L / RAD;
    const int twstep = L / (Ns * RAD);
    for (int bf = threadIdx.x; bf < nbf * NCOL; bf += NT) {
        const int c = bf % NCOL;
        const int j = bf / NCOL;
        const int k = j % Ns;
        float2 v[RAD];
#pragma unroll
        for (int r = 0; r < RAD; ++r) {
            const int row = j + r * nbf;
            float2 x = row < zero_from ? src[row * NCOL + c] : float2{0.f, 0.f};
            if (r > 0) x = cmul(x, twid<S>(twL, r * k * twstep));
            v[r] = x;
        }
        dftR<RAD, S>(v);
        const int o = (j / Ns) * Ns * RAD + k;
#pragma unroll
        for (int r = 0; r < RAD; ++r) dst[(o + r * Ns) * NCOL + c] = v[r];
    }
}

// Full L-point FFT of NCOL columns, ping-pong between bufA and bufB; returns the buffer
// holding the result. Radices are encoded 4 bits each in `code`.
template <int S, int NCOL, int NT>
__device__ __forceinline__ float2* col_fft(float2* bufA, float2* bufB, int L, uint32_t code, int npass, const float2* twL, int zero_from)
{
    float2* src = bufA;
    float2* dst = bufB;
    int Ns = 1;
    for (int p = 0; p < npass; ++p) {
        const int rad = (code >> (4 * p)) & 15;
        const int zf = p == 0 ? zero_from : L;
        switch (rad) {
        case 2: col_pass<2, S, NCOL, NT>(src, dst, L, Ns, twL, zf); break;
        case 3: col_pass<3, S, NCOL, NT>(src, dst, L, Ns, twL, zf); break;
        case 4: col_pass<4, S, NCOL, NT>(src, dst, L, Ns, twL, zf); break;
        case 5: col_pass<5, S, NCOL, NT>(src, dst, L, Ns, twL, zf); break;
        default: col_pass<8, S, NCOL, NT>(src, dst, L, Ns, twL, zf); break;
        }
        lds_barrier();
        float2* t = src;
        src = dst;
        dst = t;
        Ns *= rad;
    }
    return src;
}

// Compile-time pass geometry (no integer division at run time). Ping-pong form.
template <int RAD, int S, int NCOL, int NT, int L, int NS>
__device__ __forceinline__ void col_pass_ct(const float2* __restrict__ src, float2* __restrict__ dst,
                                           const float2* __restrict__ twL, int zero_from)
{
    constexpr int nbf = L / RAD;
    constexpr int twstep = L / (NS * RAD);
    constexpr int total = nbf * NCOL;
    constexpr int iters = (total + NT - 1) / NT;
#pragma unroll
    for (int it = 0; it < iters; ++it) {
        const int bf = threadIdx.x + it * NT;
        if (total % NT != 0 && bf >= total) break;
        const int c = bf % NCOL;
        const int j = bf / NCOL;
        const int k = j % NS;
        float2 v[RAD];
#pragma unroll
        for (int r = 0; r < RAD; ++r) {
            const int row = j + r * nbf;
            float2 x = row < zero_from ? src[row * NCOL + c] : float2{0.f, 0.f};
            if (r > 0 && NS > 1) x = cmul(x, twid<S>(twL, r * k * twstep));
            v[r] = x;
        }
        dftR<RAD, S>(v);
        const int o = (j / NS) * NS * RAD + k;
#pragma unroll
        for (int r = 0; r < RAD; ++r) dst[(o + r * NS) * NCOL + c] = v[r];
    }
}

// Same pass in place: every thread first gathers all its butterflies into registers,
// the workgroup synchronises, then the results are written back (one LDS buffer).
template <int RAD, int S, int NCOL, int NT, int L, int NS>
__device__ __forceinline__ void col_pass_ip(float2* __restrict__ buf, const float2* __restrict__ twL, int zero_from)
{
    constexpr int nbf = L / RAD;
    constexpr int twstep = L / (NS * RAD);
    constexpr int total = nbf * NCOL;
    constexpr int BPT = (total + NT - 1) / NT;
    float2 v[BPT][RAD];
#pragma unroll
    for (int q = 0; q < BPT; ++q) {
        const int bf = threadIdx.x + q * NT;
        if (total % NT != 0 && bf >= total) break;
        const int c = bf % NCOL;
        const int j = bf / NCOL;
        const int k = j % NS;
#pragma unroll
        for (int r = 0; r < RAD; ++r) {
            const int row = j + r * nbf;
            float2 x = row < zero_from ? buf[row * NCOL + c] : float2{0.f, 0.f};
            if (r > 0 && NS > 1) x = cmul(x, twid<S>(twL, r * k * twstep));
            v[q][r] = x;
        }
        dftR<RAD, S>(v[q]);
    }
    lds_barrier();
#pragma unroll
    for (int q = 0; q < BPT; ++q) {
        const int bf = threadIdx.x + q * NT;
        if (total % NT != 0 && bf >= total) break;
        const int c = bf % NCOL;
        const int j = bf / NCOL;
        const int k = j % NS;
        const int o = (j / NS) * NS * RAD + k;
#pragma unroll
        for (int r = 0; r < RAD; ++r) buf[(o + r * NS) * NCOL + c] = v[q][r];
    }
    lds_barrier();
}

// L = 64 * NB as radix 8 x 8 (x NB): ping-pong form, returns the result buffer.
template <int S, int NCOL, int NT, int NB>
__device__ __forceinline__ float2* col_fft_ct(float2* a, float2* b, const float2* twL, int zero_from)
{
    constexpr int L = 64 * NB;
    col_pass_ct<8, S, NCOL, NT, L, 1>(a, b, twL, zero_from);
    lds_barrier();
    col_pass_ct<8, S, NCOL, NT, L, 8>(b, a, twL, L);
    lds_barrier();
    if constexpr (NB == 1) {
        return a;
    } else {
        col_pass_ct<NB, S, NCOL, NT, L, 64>(a, b, twL, L);
        lds_barrier();
        return b;
    }
}

// In-place form (result in buf).
template <int S, int NCOL, int NT, int NB>
__device__ __forceinline__ void col_fft_ip(float2* buf, const float2* twL, int zero_from)
{
    constexpr int L = 64 * NB;
    col_pass_ip<8, S, NCOL, NT, L, 1>(buf, twL, zero_from);
    col_pass_ip<8, S, NCOL, NT, L, 8>(buf, twL, L);
    if constexpr (NB > 1) col_pass_ip<NB, S, NCOL, NT, L, 64>(buf, twL, L);
}

struct FusedArgs {
    const float2* x;       // [P][N0][N][N]
    float2* out;           // [P][N0][N][N]
    const float2* sub;     // optional [P][N0][N][N]
    const float2* smaps;   // [Cmap][N0][N][N]
    const int* coils;      // [ncoil] device
    const float* kernels;  // [nk][L0][L*L] permuted
    float2* T;             // [pchunk][ccount][N0][L*L]
    const float2* twL;     // exp(-2 pi i k / L), L entries
    const float2* twL0;    // exp(-2 pi i k / L0)
    const float2* tw32;    // exp(-2 pi i k / 32)
    int N0, L0;
    int p0;                // first problem of the chunk
    int c0, ccount;        // coil range of this launch
    int accumulate;        // out += (later coil chunks)
    int coil_pstride;      // problem p uses coil-map rows coils[i] + p * coil_pstride
    const int* pcoils;     // optional [P] device: problem p uses its first pcoils[p] coils
    uint32_t codeL, codeL0;
    int npassL, npassL0;
};

// Coils problem p runs through in this launch (the coil chunk [c0, c0 + ccount) cut to
// the problem's own count). The grid workspace keeps the stride of ccount coils.
__device__ __forceinline__ int coil_count(const FusedArgs& a, int p)
{
    if (!a.pcoils) return a.ccount;
    const int n = a.pcoils[p] - a.c0;
    return n < 0 ? 0 : (n < a.ccount ? n : a.ccount);
}

template <int NB> __global__ void __launch_bounds__(kXY) fwd_xy_kernel(FusedArgs a)
{
    constexpr int N = 32 * NB, L = 64 * NB, P = 2 * NB;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float2* bufA = reinterpret_cast<float2*>(smem_raw);
    float2* bufB = bufA + L * kW;
    float2* twL = bufB + L * kW;
    float2* tw32 = twL + L;
    const int t = threadIdx.x;
    const int b = t & 31;
    const int yl = t >> 5;
    const int z = blockIdx.x;
    const int pl = blockIdx.y;
    const int p = a.p0 + pl;
    for (int i = t; i < L; i += kXY) twL[i] = a.twL[i];
    if (t < 32) tw32[t] = a.tw32[t];
    const int64_t plane = (int64_t)N * N;
    // the x plane is re-read per coil (from L2 after the first coil) to keep registers free
    const float2* xp = a.x + ((int64_t)p * a.N0 + z) * plane;
    lds_barrier();

    const int ncl = coil_count(a, a.p0 + pl);
    for (int ci = 0; ci < ncl; ++ci) {
        const int coil = a.coils[a.c0 + ci] + p * a.coil_pstride;
        const float2* sp = a.smaps + ((int64_t)coil * a.N0 + z) * plane;
        float2 v[NB][NB];
#pragma unroll
        for (int k = 0; k < NB; ++k)
#pragma unroll
            for (int j = 0; j < NB; ++j) {
                const int64_t idx = (int64_t)(yl + 32 * k) * N + b + 32 * j;
                v[k][j] = cmul(xp[idx], sp[idx]);
            }
        float2* Tp = a.T + (((int64_t)pl * a.ccount + ci) * a.N0 + z) * (int64_t)L * L;
        for (int s = 0; s < P; ++s) {
            const float2 w1 = twL[(s * b) % L];
            const float2 w32 = twL[(s * 32) % L];
#pragma unroll
            for (int k = 0; k < NB; ++k) {
                float2 acc = v[k][NB - 1];
#pragma unroll
                for (int j = NB - 2; j >= 0; --j) acc = cadd(cmul(acc, w32), v[k][j]);
                acc = cmul(acc, w1);
                acc = dft32_fwd(acc, b, tw32);
                bufA[(yl + 32 * k) * kW + brev5(b)] = acc;
            }
            lds_barrier();
            const float2* res = col_fft_ct<-1, kW, kXY, NB>(bufA, bufB, twL, N);
            float2* dst = Tp + (int64_t)s * L * kW;
            for (int e = t; e < L * kW; e += kXY) dst[e] = res[e];
            lds_barrier();
        }
    }
}

template <int NB> __global__ void __launch_bounds__(kXY) inv_xy_kernel(FusedArgs a)
{
    constexpr int N = 32 * NB, L = 64 * NB, P = 2 * NB;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float2* bufA = reinterpret_cast<float2*>(smem_raw);
    float2* bufB = bufA + L * kW;
    float2* twL = bufB + L * kW;
    float2* tw32 = twL + L;
    const int t = threadIdx.x;
    const int b = t & 31;
    const int yl = t >> 5;
    const int z = blockIdx.x;
    const int pl = blockIdx.y;
    const int p = a.p0 + pl;
    for (int i = t; i < L; i += kXY) twL[i] = a.twL[i];
    if (t < 32) tw32[t] = a.tw32[t];
    const int64_t plane = (int64_t)N * N;
    float2 acc[NB][NB];
#pragma unroll
    for (int k = 0; k < NB; ++k)
#pragma unroll
        for (int j = 0; j < NB; ++j) acc[k][j] = float2{0.f, 0.f};
    lds_barrier();

    const int ncl = coil_count(a, a.p0 + pl);
    for (int ci = 0; ci < ncl; ++ci) {
        const int coil = a.coils[a.c0 + ci] + p * a.coil_pstride;
        const float2* sp = a.smaps + ((int64_t)coil * a.N0 + z) * plane;
        float2 sv[NB][NB];
#pragma unroll
        for (int k = 0; k < NB; ++k)
#pragma unroll
            for (int j = 0; j < NB; ++j) {
                const float2 q = sp[(yl + 32 * k) * N + b + 32 * j];
                sv[k][j] = float2{q.x, -q.y};
            }
        const float2* Tp = a.T + (((int64_t)pl * a.ccount + ci) * a.N0 + z) * (int64_t)L * L;
        for (int s = 0; s < P; ++s) {
            const float2* src = Tp + (int64_t)s * L * kW;
            for (int e = t; e < L * kW; e += kXY) bufA[e] = src[e];
            lds_barrier();
            const float2* res = col_fft_ct<1, kW, kXY, NB>(bufA, bufB, twL, L);
            // rows y < N: inverse 32-point DFT over m (bit-reversed load), coset twiddle,
            // conj(S) multiply, accumulate
#pragma unroll
            for (int k = 0; k < NB; ++k) {
                const int y = yl + 32 * k;
                float2 u = res[y * kW + brev5(b)];
                u = dft32_inv_bitrev_in(u, b, tw32);
#pragma unroll
                for (int j = 0; j < NB; ++j) {
                    const int n = b + 32 * j;
                    const float2 w = twid<1>(twL, (s * n) % L);
                    acc[k][j] = cadd(acc[k][j], cmul(sv[k][j], cmul(w, u)));
                }
            }
            lds_barrier();
        }
    }
    float2* op = a.out + ((int64_t)p * a.N0 + z) * plane;
    const float2* sb = a.sub ? a.sub + ((int64_t)p * a.N0 + z) * plane : nullptr;
#pragma unroll
    for (int k = 0; k < NB; ++k)
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            const int64_t idx = (int64_t)(yl + 32 * k) * N + b + 32 * j;
            float2 r = acc[k][j];
            if (a.accumulate) r = cadd(r, op[idx]);
            else if (sb) r = csub(r, sb[idx]);
            op[idx] = r;
        }
}

// z-axis: FFT (N0 -> L0), multiply by the real kernel, inverse FFT, crop; in place.
// One launch per problem of the chunk (pl), with that problem's kernel K.
__global__ void __launch_bounds__(kZ) z_conv_kernel(FusedArgs a, int LL, int pl, const float* __restrict__ K)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int L0 = a.L0;
    float2* bufA = reinterpret_cast<float2*>(smem_raw);
    float2* bufB = bufA + L0 * kW;
    float2* twz = bufB + L0 * kW;
    const int t = threadIdx.x;
    const int pos0 = blockIdx.x * kW;  // 32 consecutive (s, ky, m) positions
    for (int i = t; i < L0; i += kZ) twz[i] = a.twL0[i];
    // kernel values of the elements this thread multiplies (fixed across coils)
    constexpr int kMaxPer = 16;  // L0 * 32 / 1024 <= 16 (L0 <= 512)
    float kv[kMaxPer];
    const int nel = L0 * kW;
#pragma unroll
    for (int i = 0; i < kMaxPer; ++i) {
        const int e = t + kZ * i;
        kv[i] = e < nel ? K[(int64_t)(e / kW) * LL + pos0 + (e % kW)] : 0.f;
    }
    lds_barrier();
    const int ncl = coil_count(a, a.p0 + pl);
    for (int ci = 0; ci < ncl; ++ci) {
        float2* Tc = a.T + ((int64_t)pl * a.ccount + ci) * a.N0 * (int64_t)LL + pos0;
        for (int e = t; e < a.N0 * kW; e += kZ) bufA[e] = Tc[(int64_t)(e / kW) * LL + (e % kW)];
        lds_barrier();
        float2* res = col_fft<-1, kW, kZ>(bufA, bufB, L0, a.codeL0, a.npassL0, twz, a.N0);
        float2* other = res == bufA ? bufB : bufA;
#pragma unroll
        for (int i = 0; i < kMaxPer; ++i) {
            const int e = t + kZ * i;
            if (e < nel) res[e] = cscale(res[e], kv[i]);
        }
        lds_barrier();
        const float2* res2 = col_fft<1, kW, kZ>(res, other, L0, a.codeL0, a.npassL0, twz, L0);
        for (int e = t; e < a.N0 * kW; e += kZ) Tc[(int64_t)(e / kW) * LL + (e % kW)] = res2[e];
        lds_barrier();
    }
}

// z-axis pass for L0 = 64 * NB0: 16-column bundles, 512 threads, one in-place LDS buffer
// (L0 * 16 complex), so several workgroups share a CU and overlap loads with FFT work.
constexpr int kZC = 16;
constexpr int kZT = 512;

template <int NB0>
__global__ void __launch_bounds__(kZT) z_conv_ct_kernel(FusedArgs a, int LL, int pl, const float* __restrict__ K)
{
    constexpr int L0 = 64 * NB0;
    constexpr int N0 = L0 / 2;
    constexpr int NEL = L0 * kZC;
    constexpr int PER = (NEL + kZT - 1) / kZT;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float2* buf = reinterpret_cast<float2*>(smem_raw);
    float2* twz = buf + NEL;
    const int t = threadIdx.x;
    const int pos0 = blockIdx.x * kZC;
    for (int i = t; i < L0; i += kZT) twz[i] = a.twL0[i];
    float kv[PER];
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int e = t + kZT * i;
        kv[i] = e < NEL ? K[(int64_t)(e / kZC) * LL + pos0 + (e % kZC)] : 0.f;
    }
    lds_barrier();
    // the next coil's bundle is prefetched into registers during the current transform
    constexpr int PFZ = (N0 * kZC + kZT - 1) / kZT;
    float2 pf[PFZ];
    {
        const float2* T0 = a.T + (int64_t)pl * a.ccount * N0 * (int64_t)LL + pos0;
#pragma unroll
        for (int i = 0; i < PFZ; ++i) {
            const int e = t + kZT * i;
            if (e < N0 * kZC) pf[i] = T0[(int64_t)(e / kZC) * LL + (e % kZC)];
        }
    }
    const int ncl = coil_count(a, a.p0 + pl);
    for (int ci = 0; ci < ncl; ++ci) {
        float2* Tc = a.T + ((int64_t)pl * a.ccount + ci) * N0 * (int64_t)LL + pos0;
#pragma unroll
        for (int i = 0; i < PFZ; ++i) {
            const int e = t + kZT * i;
            if (e < N0 * kZC) buf[e] = pf[i];
        }
        lds_barrier();
        if (ci + 1 < ncl) {
            const float2* Tn = Tc + (int64_t)N0 * LL;
#pragma unroll
            for (int i = 0; i < PFZ; ++i) {
                const int e = t + kZT * i;
                if (e < N0 * kZC) pf[i] = Tn[(int64_t)(e / kZC) * LL + (e % kZC)];
            }
        }
        col_fft_ip<-1, kZC, kZT, NB0>(buf, twz, N0);
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int e = t + kZT * i;
            if (e < NEL) buf[e] = cscale(buf[e], kv[i]);
        }
        lds_barrier();
        col_fft_ip<1, kZC, kZT, NB0>(buf, twz, L0);
#pragma unroll
        for (int i = 0; i < (N0 * kZC + kZT - 1) / kZT; ++i) {
            const int e = t + kZT * i;
            if (e < N0 * kZC) Tc[(int64_t)(e / kZC) * LL + (e % kZC)] = buf[e];
        }
        lds_barrier();
    }
}

// 16-point DFT in registers: n = 4a + b, k = c + 4d,
// X[c + 4d] = sum_b w16^{bc} w4^{bd} sum_a v[4a + b] w4^{ac}.
template <int S, bool HALF = false> __device__ __forceinline__ void dft16(float2* v)
{
    // w16^m = exp(S 2 pi i m / 16), m = 0..9
    const float C[10] = {1.f, 0.92387953251128674f, 0.70710678118654752f, 0.38268343236508977f, 0.f,
                         -0.38268343236508977f, -0.70710678118654752f, -0.92387953251128674f, -1.f,
                         -0.92387953251128674f};
    const float Sn[10] = {0.f, 0.38268343236508977f, 0.70710678118654752f, 0.92387953251128674f, 1.f,
                          0.92387953251128674f, 0.70710678118654752f, 0.38268343236508977f, 0.f,
                          -0.38268343236508977f};
    float2 t[16];
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        float2 q[4] = {v[b], v[4 + b], HALF ? float2{0.f, 0.f} : v[8 + b], HALF ? float2{0.f, 0.f} : v[12 + b]};
        dft4_first<S, HALF>(q);
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int m = b * c;
            t[4 * b + c] = m == 0 ? q[c] : cmul(q[c], float2{C[m], (float)S * Sn[m]});
        }
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        float2 q[4] = {t[c], t[4 + c], t[8 + c], t[12 + c]};
        dft4<S>(q);
#pragma unroll
        for (int d = 0; d < 4; ++d) v[c + 4 * d] = q[d];
    }
}

// 12-point DFT in registers: n = 4 n1 + n2, k = k1 + 3 k2,
// X[k1 + 3 k2] = sum_{n2} w12^{n2 k1} w4^{n2 k2} sum_{n1} v[4 n1 + n2] w3^{n1 k1}.
template <int S> __device__ __forceinline__ void dft12(float2* v)
{
    // w12^m = exp(S 2 pi i m / 12), m = 0..6
    const float C[7] = {1.f, 0.86602540378443865f, 0.5f, 0.f, -0.5f, -0.86602540378443865f, -1.f};
    const float Sn[7] = {0.f, 0.5f, 0.86602540378443865f, 1.f, 0.86602540378443865f, 0.5f, 0.f};
    float2 t[12];
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) {
        float2 q[3] = {v[n2], v[4 + n2], v[8 + n2]};
        dft3<S>(q);
#pragma unroll
        for (int k1 = 0; k1 < 3; ++k1) {
            const int mm = n2 * k1;
            t[3 * n2 + k1] = mm == 0 ? q[k1] : cmul(q[k1], float2{C[mm], (float)S * Sn[mm]});
        }
    }
#pragma unroll
    for (int k1 = 0; k1 < 3; ++k1) {
        float2 q[4] = {t[k1], t[3 + k1], t[6 + k1], t[9 + k1]};
        dft4<S>(q);
#pragma unroll
        for (int k2 = 0; k2 < 4; ++k2) v[k1 + 3 * k2] = q[k2];
    }
}

// 20-point DFT in registers, in place: n = 5 n1 + n2, k = k1 + 4 k2,
//   X[k1 + 4 k2] = sum_{n2} w20^{n2 k1} w5^{n2 k2} sum_{n1} v[5 n1 + n2] w4^{n1 k1};
// the 4-point DFTs over n1 write back into the slots they read (n2 + 5 k1), the
// 5-point DFTs over n2 then read and write slots 5 k1 + k2, so X[k] ends in slot
// dft_slot<20>(k) = 5 (k mod 4) + k / 4. With HALF the inputs n >= 10 (n1 >= 2) are
// zero and are not read.
template <int S, bool HALF = false> __device__ __forceinline__ void dft20_dr(float2* v)
{
    // w20^m = exp(S 2 pi i m / 20), m = 0..9 (n2 k1 <= 12; m >= 10 as -w20^{m-10})
    const float C[10] = {1.f, 0.95105651629515357f, 0.80901699437494742f, 0.58778525229247313f,
                         0.30901699437494742f, 0.f, -0.30901699437494742f, -0.58778525229247313f,
                         -0.80901699437494742f, -0.95105651629515357f};
    const float Sn[10] = {0.f, 0.30901699437494742f, 0.58778525229247313f, 0.80901699437494742f,
                          0.95105651629515357f, 1.f, 0.95105651629515357f, 0.80901699437494742f,
                          0.58778525229247313f, 0.30901699437494742f};
#pragma unroll
    for (int n2 = 0; n2 < 5; ++n2) {
        float2 q[4] = {v[n2], v[5 + n2], HALF ? float2{0.f, 0.f} : v[10 + n2], HALF ? float2{0.f, 0.f} : v[15 + n2]};
        dft4_first<S, HALF>(q);
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) {
            const int mm = n2 * k1;
            const float sg = mm >= 10 ? -1.f : 1.f;
            const int m10 = mm >= 10 ? mm - 10 : mm;
            v[n2 + 5 * k1] = mm == 0 ? q[k1] : cmul(q[k1], float2{sg * C[m10], sg * (float)S * Sn[m10]});
        }
    }
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) dft5<S>(v + 5 * k1);
}

// Register DFTs of length R: natural order for R in {4, 8, 12, 16}; for R = 20 the
// in-place form above, with output k in slot dft_slot<R>(k). With HALF (R = 20 only)
// the upper half of the inputs is zero and is not read.
template <int R> __host__ __device__ constexpr int dft_slot(int k) { return R == 20 ? 5 * (k % 4) + k / 4 : k; }

template <int R, int S, bool HALF = false> __device__ __forceinline__ void dftn(float2* v)
{
    static_assert(R == 4 || R == 8 || R == 12 || R == 16 || R == 20, "unsupported register DFT length");
    static_assert(!HALF || R == 20, "HALF is implemented for R = 20");
    if constexpr (R == 4) dft4<S>(v);
    else if constexpr (R == 8) dft8<S>(v);
    else if constexpr (R == 12) dft12<S>(v);
    else if constexpr (R == 16) dft16<S>(v);
    else dft20_dr<S, HALF>(v);
}

// Buffer loads and stores on one plane: a descriptor in SGPRs (base, size in bytes;
// accesses past the size read zero and write nothing), a per-lane 32-bit byte offset
// and a uniform one, so a thread needs one VGPR of address for all its accesses.
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_rsrc(const void* base, uint32_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, (int)bytes, 0x00020000);
}
// ST: streaming data read or written once (non-temporal: kept out of the caches' way of
// data that is used again, such as the image plane re-read for every coil).
template <bool ST = false> __device__ __forceinline__ float2 bload(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff)
{
    return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, ST ? 2 : 0));
}
template <bool ST = false>
__device__ __forceinline__ void bstore(__amdgpu_buffer_rsrc_t r, float2 v, uint32_t voff, uint32_t soff)
{
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, voff, soff, ST ? 2 : 0);
}

// Plane kernels with the L-point y transforms as register four-step FFTs,
// L = 16 * R2 (R2 = L / 16 = 4 NB): row y = y1 + 16 y2 and frequency ky = k1 + R2 k2, so
//   X[k1 + R2 k2] = sum_{y1} w16^{y1 k2} [ wL^{y1 k1} sum_{y2} x[y1 + 16 y2] wR2^{y2 k1} ].
// Each step is a DFT in registers; the two steps meet through one LDS transpose. Two
// cosets of 32 x-frequencies go through a workgroup together; each of the 512 threads
// owns one (column, y1 or k1) position of both cosets, and in the x step and the
// output step the rows yl + 16 k of lane b. 512 threads leave 256 VGPRs per thread for
// the plane (x S products, or the coil sum and conj(S)) plus two cosets of transforms.
// Compared with the Stockham passes in LDS this moves each element through LDS about
// three times instead of about ten.
// At N = 160 (L = 320 = 16 x 20) one coset block is 80 KB, so one coset goes through a
// step instead of two, and the y transforms over k1 (20 values) take threads r and r + 16.
constexpr int kXYW = 512;

// cosets per step of the coset kernels: two while both blocks fit in LDS (N <= 128)
__host__ __device__ constexpr int cosets_per_step(int NB) { return NB <= 4 ? 2 : 1; }

// LDS bytes of fwd_xy_w_kernel / inv_xy_w_kernel
constexpr size_t coset_kernel_lds(int NB)
{
    return sizeof(float2) * ((size_t)cosets_per_step(NB) * 64 * NB * kW + 64 * NB + 32);
}

template <int NB> __global__ void __launch_bounds__(kXYW) fwd_xy_w_kernel(FusedArgs a)
{
    constexpr int N = 32 * NB, L = 64 * NB, P = 2 * NB, R2 = 4 * NB, CS = cosets_per_step(NB);
    constexpr int CW = L * kW;  // one coset block: L rows x 32 columns
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float2* buf = reinterpret_cast<float2*>(smem_raw);  // [CS][CW]
    float2* twL = buf + CS * CW;
    float2* tw32 = twL + L;
    const int t = threadIdx.x;
    const int b = t & 31;   // x step: lane of the half-wave (x index b + 32 j)
    const int yl = t >> 5;  // x step: rows yl + 16 k
    const int m = t & 31;   // y steps: column (x frequency P m + s)
    const int r = t >> 5;   // y steps: y1 (step 1) or k1 (step 2)
    const int z = blockIdx.x;
    const int pl = blockIdx.y;
    const int p = a.p0 + pl;
    for (int i = t; i < L; i += kXYW) twL[i] = a.twL[i];
    if (t < 32) tw32[t] = a.tw32[t];
    const int64_t plane = (int64_t)N * N;
    // x and the coil maps through buffer descriptors: one VGPR of address per thread
    // (pixel (yl + 16 k, b + 32 j) at byte xoff + (16 k N + 32 j) * 8)
    const auto xr = plane_rsrc(a.x + ((int64_t)p * a.N0 + z) * plane, sizeof(float2) * N * N);
    const uint32_t xoff = (uint32_t)(yl * N + b) * 8u;
    const LaneTw32 lt = lane_tw32(b, a.tw32);
    lds_barrier();

    const int ncl = coil_count(a, a.p0 + pl);
    for (int ci = 0; ci < ncl; ++ci) {
        const int coil = a.coils[a.c0 + ci] + p * a.coil_pstride;
        const auto sr = plane_rsrc(a.smaps + ((int64_t)coil * a.N0 + z) * plane, sizeof(float2) * N * N);
        float2 v[2 * NB][NB];
#pragma unroll
        for (int k = 0; k < 2 * NB; ++k)
#pragma unroll
            for (int j = 0; j < NB; ++j) {
                const uint32_t so = (uint32_t)(16 * k * N + 32 * j) * 8u;
                v[k][j] = cmul(bload(xr, xoff, so), bload<true>(sr, xoff, so));
            }
        // this coil's T plane [s][ky][m]: element (s, ky, m) at byte (s CW + ky kW + m) * 8
        const auto tr = plane_rsrc(a.T + (((int64_t)pl * a.ccount + ci) * a.N0 + z) * (int64_t)L * L,
                                   sizeof(float2) * L * L);
        for (int s0 = 0; s0 < P; s0 += CS) {
            // x step: for coset s the 32 frequencies P m + s of each row, written as rows
            // y < N of the coset block (column m)
#pragma unroll
            for (int cc = 0; cc < CS; ++cc) {
                const int s = s0 + cc;
                const float2 w1 = twL[(s * b) % L];
                const float2 w32 = twL[(s * 32) % L];
#pragma unroll
                for (int k = 0; k < 2 * NB; ++k) {
                    float2 acc = v[k][NB - 1];
#pragma unroll
                    for (int j = NB - 2; j >= 0; --j) acc = cadd(cmul(acc, w32), v[k][j]);
                    acc = cmul(acc, w1);
                    acc = dft32_fwd_r(acc, lt);
                    buf[cc * CW + (yl + 16 * k) * kW + brev5(b)] = acc;
                }
            }
            lds_barrier();
            // y step 1 (thread: column m, y1 = r, both cosets): DFT over y2 (rows >= N
            // are zero), twiddle wL^{y1 k1}. A thread reads rows r + 16 y2 and writes rows
            // 16 k1 + r of its own column: the same rows (mod 16), touched by no other
            // thread, so the transform is in place without a barrier.
#pragma unroll 1
            for (int cs = 0; cs < CS; ++cs) {
                float2 u[R2];
#pragma unroll
                for (int y2 = 0; y2 < R2; ++y2)
                    u[y2] = y2 < R2 / 2 ? buf[cs * CW + (r + 16 * y2) * kW + m] : float2{0.f, 0.f};
                dftn<R2, -1, R2 == 20>(u);
#pragma unroll
                for (int k1 = 0; k1 < R2; ++k1) {
                    const float2 f = u[dft_slot<R2>(k1)];
                    buf[cs * CW + (k1 * 16 + r) * kW + m] = k1 == 0 ? f : cmul(f, twL[(r * k1) % L]);
                }
            }
            lds_barrier();
            // y step 2 (thread: column m, k1 = r, r + 16, ... < R2): DFT over y1 -> k2,
            // ky = k1 + R2 k2, straight to HBM
#pragma unroll 1
            for (int k1 = r; k1 < R2; k1 += 16) {
#pragma unroll 1
                for (int cs = 0; cs < CS; ++cs) {
                    float2 u[16];
#pragma unroll
                    for (int y1 = 0; y1 < 16; ++y1) u[y1] = buf[cs * CW + (k1 * 16 + y1) * kW + m];
                    dft16<-1>(u);
                    const uint32_t voff = (uint32_t)(k1 * kW + m) * 8u;
#pragma unroll
                    for (int k2 = 0; k2 < 16; ++k2)
                        bstore(tr, u[k2], voff, (uint32_t)((s0 + cs) * CW + R2 * k2 * kW) * 8u);
                }
            }
            lds_barrier();
        }
    }
}

// Row halves of the inverse coset kernel: at N = 160 a thread would hold 50 pixels of
// coil sum and 50 of coil image (200 VGPRs) besides the transforms, so two workgroups
// share a z-plane, each producing the rows y = y1 + 16 y2 with y2 in its half. Both run
// the whole step 1 (the same T reads); they are workgroups i and i + 8 of the launch,
// which dispatch sends to the same XCD at nearly the same time, so the second one's
// reads are L2 hits.
__host__ __device__ constexpr int inv_row_splits(int NB) { return NB == 5 ? 2 : 1; }

// grid.x of the inverse coset kernel for N0 z-planes
inline unsigned inv_w_grid_x(int NB, int N0)
{
    return inv_row_splits(NB) == 1 ? (unsigned)N0 : (unsigned)(16 * ((N0 + 7) / 8));
}

template <int NB> __global__ void __launch_bounds__(kXYW) inv_xy_w_kernel(FusedArgs a)
{
    constexpr int N = 32 * NB, L = 64 * NB, P = 2 * NB, R2 = 4 * NB, CS = cosets_per_step(NB);
    constexpr int HS = inv_row_splits(NB), KR = 2 * NB / HS;  // row halves; y2 values per thread
    constexpr int CW = L * kW;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float2* buf = reinterpret_cast<float2*>(smem_raw);  // [CS][16 y1][R2 k1][32 m]
    float2* twL = buf + CS * CW;
    float2* tw32 = twL + L;
    const int t = threadIdx.x;
    const int b = t & 31;   // step 2: lane (x index b + 32 j after the inverse DFT)
    const int yl = t >> 5;  // step 2: pixel rows yl + 16 (KR h + k), i.e. y1 = yl, y2 = KR h + k
    const int m = t & 31;   // step 1: column
    const int r = t >> 5;   // step 1: k1 = r, r + 16, ... (< R2)
    int z = blockIdx.x, h = 0;
    if constexpr (HS == 2) {
        z = 8 * (blockIdx.x / 16) + (blockIdx.x % 8);
        h = (blockIdx.x / 8) % 2;
        if (z >= a.N0) return;  // the whole workgroup: the last group of 8 planes may be short
    }
    const int y0 = KR * h;  // first y2 of this workgroup's rows
    const int pl = blockIdx.y;
    const int p = a.p0 + pl;
    for (int i = t; i < L; i += kXYW) twL[i] = a.twL[i];
    if (t < 32) tw32[t] = a.tw32[t];
    const int64_t plane = (int64_t)N * N;
    // coil maps, output and sub through buffer descriptors: pixel (yl + 16 y2, b + 32 j)
    // at byte xoff + (16 y2 N + 32 j) * 8
    const uint32_t xoff = (uint32_t)(yl * N + b) * 8u;
    const LaneTw32 lt = lane_tw32(b, a.tw32);
    float2 acc[KR][NB];
#pragma unroll
    for (int k = 0; k < KR; ++k)
#pragma unroll
        for (int j = 0; j < NB; ++j) acc[k][j] = float2{0.f, 0.f};
    lds_barrier();

    const int ncl = coil_count(a, a.p0 + pl);
    for (int ci = 0; ci < ncl; ++ci) {
        const int coil = a.coils[a.c0 + ci] + p * a.coil_pstride;
        const auto sr = plane_rsrc(a.smaps + ((int64_t)coil * a.N0 + z) * plane, sizeof(float2) * N * N);
        // this coil's image, summed over the cosets; times conj(S) once at the end
        float2 ic[KR][NB];
#pragma unroll
        for (int k = 0; k < KR; ++k)
#pragma unroll
            for (int j = 0; j < NB; ++j) ic[k][j] = float2{0.f, 0.f};
        // this coil's T plane [s][ky][m]: element (s, ky, m) at byte (s CW + ky kW + m) * 8
        const auto tr = plane_rsrc(a.T + (((int64_t)pl * a.ccount + ci) * a.N0 + z) * (int64_t)L * L,
                                   sizeof(float2) * L * L);
        for (int s0 = 0; s0 < P; s0 += CS) {
            // step 1 (thread: column m, k1 = r, r + 16, ..., each coset of the step):
            // inverse DFT over k2 of X[k1 + R2 k2] -> y1, twiddle wL^{-y1 k1}, into LDS as
            // [y1][k1][m]
#pragma unroll 1
            for (int k1 = r; k1 < R2; k1 += 16) {
#pragma unroll 1
                for (int cs = 0; cs < CS; ++cs) {
                    float2 u[16];
                    const uint32_t voff = (uint32_t)(k1 * kW + m) * 8u;
#pragma unroll
                    for (int k2 = 0; k2 < 16; ++k2)
                        u[k2] = bload(tr, voff, (uint32_t)((s0 + cs) * CW + R2 * k2 * kW) * 8u);
                    dft16<1>(u);
#pragma unroll
                    for (int q = 1; q < 16; ++q) u[q] = cmulc(twL[(q * k1) % L], u[q]);
#pragma unroll
                    for (int q = 0; q < 16; ++q) buf[cs * CW + (q * R2 + k1) * kW + m] = u[q];
                }
            }
            lds_barrier();
            // step 2 (thread: rows yl + 16 y2, lane b): inverse DFT over k1 -> y2 for
            // column m = brev5(b); inverse 32-point DFT across lanes; coset twiddle,
            // conj(S) multiply, coil sum
#pragma unroll 1
            for (int cc = 0; cc < CS; ++cc) {
                const int s = s0 + cc;
                float2 u[R2];
#pragma unroll
                for (int k1 = 0; k1 < R2; ++k1) u[k1] = buf[cc * CW + (yl * R2 + k1) * kW + brev5(b)];
                dftn<R2, 1>(u);
                float2 w[NB];
#pragma unroll
                for (int j = 0; j < NB; ++j) w[j] = twid<1>(twL, (s * (b + 32 * j)) % L);
#pragma unroll
                for (int k = 0; k < KR; ++k) {
                    // y2 = y0 + k; with row halves the half is chosen by a select, so the
                    // register array is indexed with constants only
                    const float2 xk = HS == 1 ? u[dft_slot<R2>(k)]
                                              : (h ? u[dft_slot<R2>(KR + k)] : u[dft_slot<R2>(k)]);
                    const float2 val = dft32_inv_r(xk, lt);
#pragma unroll
                    for (int j = 0; j < NB; ++j) ic[k][j] = cfma(ic[k][j], w[j], val);
                }
            }
            lds_barrier();
        }
#pragma unroll
        for (int k = 0; k < KR; ++k)
#pragma unroll
            for (int j = 0; j < NB; ++j)
                acc[k][j] = cfma_conj(acc[k][j], bload<true>(sr, xoff, (uint32_t)(16 * (y0 + k) * N + 32 * j) * 8u),
                                      ic[k][j]);
    }
    const auto orr = plane_rsrc(a.out + ((int64_t)p * a.N0 + z) * plane, sizeof(float2) * N * N);
    const auto sbr = plane_rsrc(a.sub ? a.sub + ((int64_t)p * a.N0 + z) * plane : a.out, sizeof(float2) * N * N);
#pragma unroll
    for (int k = 0; k < KR; ++k)
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            const uint32_t so = (uint32_t)(16 * (y0 + k) * N + 32 * j) * 8u;
            float2 rr = acc[k][j];
            if (a.accumulate) rr = cadd(rr, bload(orr, xoff, so));
            else if (a.sub) rr = csub(rr, bload(sbr, xoff, so));
            bstore(orr, rr, xoff, so);
        }
}

typedef __attribute__((address_space(3))) void lds_void_t;

// Copy one coset block (CW complex64, contiguous in HBM) into LDS with direct-to-LDS
// loads (global_load_lds_dwordx4): instruction i of wave w moves 1 KB, lane l the 16 B at
// byte (i * 512 + w * 64 + l) * 16 of both the block and the LDS buffer. Completion is
// counted by vmcnt; a wait on it and a barrier make the data visible to every wave.
template <int CW> __device__ __forceinline__ void coset_to_lds(const float2* __restrict__ src, float2* dst)
{
    static_assert((CW * 8) % (kXYW * 16) == 0, "coset block must be a whole number of 8 KB steps");
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const char* g = reinterpret_cast<const char*>(src);
    char* d = reinterpret_cast<char*>(dst);
#pragma unroll
    for (int i = 0; i < (CW * 8) / (kXYW * 16); ++i) {
        const int chunk = (i * kXYW + w * 64) * 16;
        __builtin_amdgcn_global_load_lds(g + chunk + l * 16, (lds_void_t*)(d + chunk), 16, 0, 0);
    }
}

// inv_xy_w_kernel with the HBM loads of the next coset in flight during the current
// coset's work: one coset per step, its raw block staged in LDS by direct-to-LDS loads
// issued right after the step-1 transform has consumed the previous one.
template <int NB> __global__ void __launch_bounds__(kXYW) inv_xy_d_kernel(FusedArgs a)
{
    constexpr int N = 32 * NB, L = 64 * NB, P = 2 * NB, R2 = 4 * NB;
    constexpr int CW = L * kW;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float2* stage = reinterpret_cast<float2*>(smem_raw);  // raw coset block [ky][m]
    float2* buf = stage + CW;                               // [y1][k1][m]
    float2* twL = buf + CW;
    const int t = threadIdx.x;
    const int b = t & 31;
    const int yl = t >> 5;
    const int m = t & 31;
    const int r = t >> 5;
    const int z = blockIdx.x;
    const int pl = blockIdx.y;
    const int p = a.p0 + pl;
    for (int i = t; i < L; i += kXYW) twL[i] = a.twL[i];
    const int64_t plane = (int64_t)N * N;
    const LaneTw32 lt = lane_tw32(b, a.tw32);
    float2 acc[2 * NB][NB];
#pragma unroll
    for (int k = 0; k < 2 * NB; ++k)
#pragma unroll
        for (int j = 0; j < NB; ++j) acc[k][j] = float2{0.f, 0.f};
    const int ncl = coil_count(a, p);
    auto block_of = [&](int ci, int s) {
        return a.T + (((int64_t)pl * a.ccount + ci) * a.N0 + z) * (int64_t)L * L + (int64_t)s * CW;
    };
    if (ncl > 0) coset_to_lds<CW>(block_of(0, 0), stage);

    for (int ci = 0; ci < ncl; ++ci) {
        const int coil = a.coils[a.c0 + ci] + p * a.coil_pstride;
        const float2* sp = a.smaps + ((int64_t)coil * a.N0 + z) * plane;
        float2 ic[2 * NB][NB];
#pragma unroll
        for (int k = 0; k < 2 * NB; ++k)
#pragma unroll
            for (int j = 0; j < NB; ++j) ic[k][j] = float2{0.f, 0.f};
        for (int s = 0; s < P; ++s) {
            __builtin_amdgcn_s_waitcnt(0);  // this wave's stage loads have landed
            lds_barrier();                  // ... and every other wave's
            // step 1 (thread: column m, k1 = r): inverse DFT over k2 -> y1, twiddle
            if (r < R2) {
                float2 u[16];
#pragma unroll
                for (int k2 = 0; k2 < 16; ++k2) u[k2] = stage[(r + R2 * k2) * kW + m];
                dft16<1>(u);
#pragma unroll
                for (int q = 1; q < 16; ++q) u[q] = cmulc(twL[(q * r) % L], u[q]);
#pragma unroll
                for (int q = 0; q < 16; ++q) buf[(q * R2 + r) * kW + m] = u[q];
            }
            lds_barrier();  // stage consumed, buf complete
            {
                const int nci = s + 1 < P ? ci : ci + 1;
                const int ns = s + 1 < P ? s + 1 : 0;
                if (nci < ncl) coset_to_lds<CW>(block_of(nci, ns), stage);
            }
            // step 2 (thread: rows yl + 16 k, lane b)
            float2 u[R2];
#pragma unroll
            for (int k1 = 0; k1 < R2; ++k1) u[k1] = buf[(yl * R2 + k1) * kW + brev5(b)];
            dftn<R2, 1>(u);
            float2 w[NB];
#pragma unroll
            for (int j = 0; j < NB; ++j) w[j] = twid<1>(twL, (s * (b + 32 * j)) % L);
#pragma unroll
            for (int k = 0; k < 2 * NB; ++k) {
                const float2 val = dft32_inv_r(u[k], lt);
#pragma unroll
                for (int j = 0; j < NB; ++j) ic[k][j] = cfma(ic[k][j], w[j], val);
            }
        }
#pragma unroll
        for (int k = 0; k < 2 * NB; ++k)
#pragma unroll
            for (int j = 0; j < NB; ++j)
                acc[k][j] = cfma_conj(acc[k][j], sp[(yl + 16 * k) * N + b + 32 * j], ic[k][j]);
    }
    float2* op = a.out + ((int64_t)p * a.N0 + z) * plane;
    const float2* sb = a.sub ? a.sub + ((int64_t)p * a.N0 + z) * plane : nullptr;
#pragma unroll
    for (int k = 0; k < 2 * NB; ++k)
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            const int64_t idx = (int64_t)(yl + 16 * k) * N + b + 32 * j;
            float2 rr = acc[k][j];
            if (a.accumulate) rr = cadd(rr, op[idx]);
            else if (sb) rr = csub(rr, sb[idx]);
            op[idx] = rr;
        }
}

// 32-point DFT in registers, radix 2 in time over two 16-point DFTs.
template <int S> __device__ __forceinline__ void dft32(float2* v)
{
    // w32^m = exp(S 2 pi i m / 32), m = 0..15
    const float C[16] = {1.f, 0.98078528040323044f, 0.92387953251128674f, 0.83146961230254524f,
                         0.70710678118654752f, 0.55557023301960222f, 0.38268343236508977f, 0.19509032201612827f,
                         0.f, -0.19509032201612827f, -0.38268343236508977f, -0.55557023301960222f,
                         -0.70710678118654752f, -0.83146961230254524f, -0.92387953251128674f, -0.98078528040323044f};
    float2 e[16], o[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        e[k] = v[2 * k];
        o[k] = v[2 * k + 1];
    }
    dft16<S>(e);
    dft16<S>(o);
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        // sin(2 pi k / 32) = cos(2 pi |8 - k| / 32)
        const float sn = k <= 8 ? C[8 - k] : C[k - 8];
        const float2 w = k == 0 ? o[0] : cmul(o[k], float2{C[k], (float)S * sn});
        v[k] = cadd(e[k], w);
        v[k + 16] = csub(e[k], w);
    }
}

// Plane kernels for N1 = N2 = 128 (L = 256) with every transform in registers. Thread t
// of 512 has q = t & 15 and g = t >> 4 and owns the rows y = g + 32 i (i < 4).
//  x: x = q + 16 x2 (x2 < 8 nonzero of 16), kx = k1 + 16 k2:
//       X[k1 + 16 k2] = sum_{x1} w16^{x1 k2} [ w256^{x1 k1} sum_{x2} v[x1 + 16 x2] w16^{x2 k1} ]
//     the inner DFT runs in the thread that holds x1 = q, the outer one in the thread
//     that holds k1 = q after a 16 x 16 transpose inside the 16 lanes of the row (LDS
//     private to the wavefront, no workgroup barrier).
//  y: y = g + 32 i (i < 4 nonzero of 8), ky = ka + 8 kb:
//       Y[ka + 8 kb] = sum_g w32^{g kb} [ w256^{g ka} sum_i v[g + 32 i] w8^{i ka} ]
//     the inner 8-point DFT in the row thread, the outer 32-point one in a thread that
//     owns one (column, ka) pair after an exchange through LDS, 64 columns at a time.
// The plane is stored at position ((h * 32 + kb) * 8 + ka) * 64 + q + 16 c for
// kx = q + 16 (4 h + c): a wavefront's 64 lanes write 512 contiguous bytes.
constexpr int kQT = 512;   // threads
constexpr int kXS = 272;   // x transpose tile of one 16-lane row: 16 x 17 (padded)
constexpr int kYS = 528;   // y exchange: stride of g, 8 ka x 64 columns + 16 (bank shift)
constexpr size_t kQLds = sizeof(float2) * (32 * kYS + 272 + 256);

__host__ __device__ __forceinline__ int64_t r16_pos(int ky, int kx)
{
    const int q = kx & 15, k2 = kx >> 4, h = k2 >> 2, c = k2 & 3;
    const int ka = ky & 7, kb = ky >> 3;
    return ((int64_t)(h * 32 + kb) * 8 + ka) * 64 + q + 16 * c;
}

// Order of LDS operations inside one wavefront (the hardware keeps it; this keeps the
// compiler from moving memory operations across).
// Both also keep the instruction scheduler from moving work across (it would hoist loads
// of later steps and run out of registers).
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    __builtin_amdgcn_sched_barrier(0);
}
__device__ __forceinline__ void phase_barrier()
{
    __builtin_amdgcn_sched_barrier(0);
    lds_barrier();
    __builtin_amdgcn_sched_barrier(0);
}

__global__ void __launch_bounds__(kQT) fwd_xy_r16_kernel(FusedArgs a)
{
    constexpr int N = 128, L = 256;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float2* yb = reinterpret_cast<float2*>(smem_raw);  // [32 g][kYS]
    float2* tq = yb + 32 * kYS;  // [16 q][17]: w256^{q k} (row stride 17: no bank conflicts)
    float2* tgt = tq + 272;      // [32 g][8]: w256^{g ka}
    const int t = threadIdx.x;
    const int q = t & 15, g = t >> 4;
    // x transpose tile of this 16-lane row (shares memory with yb; a workgroup barrier
    // separates the two uses)
    float2* xs = yb + g * kXS;
    const int z = blockIdx.x;
    const int pl = blockIdx.y;
    const int p = a.p0 + pl;
    if (t < 256) {
        tq[(t >> 4) * 17 + (t & 15)] = a.twL[((t >> 4) * (t & 15)) & (L - 1)];
        tgt[t] = a.twL[((t >> 3) * (t & 7)) & (L - 1)];
    }
    const int64_t plane = (int64_t)N * N;
    const auto xr = plane_rsrc(a.x + ((int64_t)p * a.N0 + z) * plane, sizeof(float2) * N * N);
    // byte offsets: pixel (g + 32 i, q + 16 x2) at xoff + i * 32768 + x2 * 128;
    // T position ((h * 32 + kb) * 8 + ka) * 64 + col at toff + (h * 32 + kb) * 4096
    const uint32_t xoff = (uint32_t)(g * N + q) * 8u;
    const uint32_t toff = (uint32_t)t * 8u;
    phase_barrier();

    const int ncl = coil_count(a, p);
    auto smap_rsrc = [&](int ci) {
        const int coil = a.coils[a.c0 + ci] + p * a.coil_pstride;
        return plane_rsrc(a.smaps + ((int64_t)coil * a.N0 + z) * plane, sizeof(float2) * N * N);
    };
    // x and S of the next row to transform, loaded one step ahead (the first row of the
    // next coil during the last y step of this one)
    float2 nx[8], ns[8];
    auto load_row = [&](const __amdgpu_buffer_rsrc_t& sr, int i) {
#pragma unroll
        for (int x2 = 0; x2 < 8; ++x2) {
            const uint32_t so = (uint32_t)(i * 32 * N + 16 * x2) * 8u;
            nx[x2] = bload(xr, xoff, so);
            ns[x2] = bload<true>(sr, xoff, so);
        }
    };
    if (ncl > 0) load_row(smap_rsrc(0), 0);
    for (int ci = 0; ci < ncl; ++ci) {
        const auto sr = smap_rsrc(ci);
        const auto tr = plane_rsrc(a.T + (((int64_t)pl * a.ccount + ci) * a.N0 + z) * (int64_t)L * L,
                                   sizeof(float2) * L * L);
        float2 X[4][16];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            float2 u[16];
#pragma unroll
            for (int x2 = 0; x2 < 8; ++x2) u[x2] = cmul(nx[x2], ns[x2]);
            if (i < 3) load_row(sr, i + 1);
#pragma unroll
            for (int x2 = 8; x2 < 16; ++x2) u[x2] = float2{0.f, 0.f};
            dft16<-1, true>(u);
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int k1 = 1; k1 < 16; ++k1) u[k1] = cmul(u[k1], tq[q * 17 + k1]);
            wave_lds_sync();
#pragma unroll
            for (int k1 = 0; k1 < 16; ++k1) xs[k1 * 17 + q] = u[k1];
            wave_lds_sync();
#pragma unroll
            for (int x1 = 0; x1 < 16; ++x1) X[i][x1] = xs[q * 17 + x1];
            dft16<-1>(X[i]);
            __builtin_amdgcn_sched_barrier(0);
        }
        phase_barrier();
#pragma unroll
        for (int h = 0; h < 4; ++h) {
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                float2 u[8];
#pragma unroll
                for (int i = 0; i < 4; ++i) u[i] = X[i][4 * h + c];
#pragma unroll
                for (int i = 4; i < 8; ++i) u[i] = float2{0.f, 0.f};
                dft8<-1, true>(u);
#pragma unroll
                for (int ka = 0; ka < 8; ++ka)
                    yb[g * kYS + ka * 64 + 16 * c + q] = ka ? cmul(u[ka], tgt[g * 8 + ka]) : u[0];
                __builtin_amdgcn_sched_barrier(0);
            }
            phase_barrier();
            if (h == 3 && ci + 1 < ncl) load_row(smap_rsrc(ci + 1), 0);
            {
                const int col = t & 63, ka = t >> 6;
                float2 u[32];
#pragma unroll
                for (int gg = 0; gg < 32; ++gg) u[gg] = yb[gg * kYS + ka * 64 + col];
                dft32<-1>(u);
#pragma unroll
                for (int kb = 0; kb < 32; ++kb) bstore<true>(tr, u[kb], toff, (uint32_t)(h * 32 + kb) * 4096u);
            }
            phase_barrier();
        }
    }
}

__global__ void __launch_bounds__(kQT) inv_xy_r16_kernel(FusedArgs a)
{
    constexpr int N = 128, L = 256;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float2* yb = reinterpret_cast<float2*>(smem_raw);
    float2* tq = yb + 32 * kYS;  // [16 q][17]: w256^{q k} (row stride 17: no bank conflicts)
    float2* tgt = tq + 272;      // [32 g][8]: w256^{g ka}
    const int t = threadIdx.x;
    const int q = t & 15, g = t >> 4;
    float2* xs = yb + g * kXS;
    const int z = blockIdx.x;
    const int pl = blockIdx.y;
    const int p = a.p0 + pl;
    if (t < 256) {
        tq[(t >> 4) * 17 + (t & 15)] = a.twL[((t >> 4) * (t & 15)) & (L - 1)];
        tgt[t] = a.twL[((t >> 3) * (t & 7)) & (L - 1)];
    }
    const int64_t plane = (int64_t)N * N;
    float2 acc[4][8];  // pixels (g + 32 i, q + 16 x2)
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int x2 = 0; x2 < 8; ++x2) acc[i][x2] = float2{0.f, 0.f};
    phase_barrier();

    const uint32_t xoff = (uint32_t)(g * N + q) * 8u;
    const uint32_t toff = (uint32_t)t * 8u;
    const int ncl = coil_count(a, p);
    for (int ci = 0; ci < ncl; ++ci) {
        const int coil = a.coils[a.c0 + ci] + p * a.coil_pstride;
        const auto sr = plane_rsrc(a.smaps + ((int64_t)coil * a.N0 + z) * plane, sizeof(float2) * N * N);
        const auto tr = plane_rsrc(a.T + (((int64_t)pl * a.ccount + ci) * a.N0 + z) * (int64_t)L * L,
                                   sizeof(float2) * L * L);
        float2 X[4][16];
        float2 sv[8];  // conj(S) row of the pixels the thread accumulates next
        auto load_srow = [&](float2 (&d)[8], int i) {
#pragma unroll
            for (int x2 = 0; x2 < 8; ++x2) d[x2] = bload<true>(sr, xoff, (uint32_t)(i * 32 * N + 16 * x2) * 8u);
        };
#pragma unroll
        for (int h = 0; h < 4; ++h) {
            {
                // inverse DFT over kb -> g for one (column, ka) pair per thread, as two
                // 16-point DFTs over the even and the odd kb (half the registers of one
                // 32-point DFT): yb row g' < 16 holds E[g'], row 16 + g' holds w32^{+g'} O[g'],
                // and g = g' + 16 s' is E[g'] + (-1)^{s'} w32^{g'} O[g'], formed by the reader.
                const int col = t & 63, ka = t >> 6;
                const float C[16] = {1.f, 0.98078528040323044f, 0.92387953251128674f, 0.83146961230254524f,
                                     0.70710678118654752f, 0.55557023301960222f, 0.38268343236508977f,
                                     0.19509032201612827f, 0.f, -0.19509032201612827f, -0.38268343236508977f,
                                     -0.55557023301960222f, -0.70710678118654752f, -0.83146961230254524f,
                                     -0.92387953251128674f, -0.98078528040323044f};
#pragma unroll
                for (int e = 0; e < 2; ++e) {
                    float2 u[16];
#pragma unroll
                    for (int m = 0; m < 16; ++m) u[m] = bload<true>(tr, toff, (uint32_t)(h * 32 + 2 * m + e) * 4096u);
                    dft16<1>(u);
#pragma unroll
                    for (int gp = 0; gp < 16; ++gp) {
                        const float sn = gp <= 8 ? C[8 - gp] : C[gp - 8];
                        const float2 v = (e && gp) ? cmul(u[gp], float2{C[gp], sn}) : u[gp];
                        yb[(16 * e + gp) * kYS + ka * 64 + col] = v;
                    }
                    __builtin_amdgcn_sched_barrier(0);
                }
            }
            phase_barrier();
            if (h == 3) load_srow(sv, 0);
            // twiddle w256^{-g ka}, inverse DFT over ka -> i, rows i < 4 kept
            const float sg = g < 16 ? 1.f : -1.f;
            const int gp = g & 15;
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                float2 u[8];
#pragma unroll
                for (int ka = 0; ka < 8; ++ka) {
                    const int o = ka * 64 + 16 * c + q;
                    u[ka] = sfma(sg, yb[(16 + gp) * kYS + o], yb[gp * kYS + o]);
                    if (ka) u[ka] = cmulc(tgt[g * 8 + ka], u[ka]);
                }
                dft8<1>(u);
#pragma unroll
                for (int i = 0; i < 4; ++i) X[i][4 * h + c] = u[i];
                __builtin_amdgcn_sched_barrier(0);
            }
            phase_barrier();
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            // inverse DFT over k2 -> x1 (thread k1 = q), transpose, twiddle w256^{-x1 k1},
            // inverse DFT over k1 -> x2 (thread x1 = q), pixels x2 < 8 kept
            // per-thread LDS offsets, opaque here so their address forms are built per row
            // instead of being held in registers across the column phase
            int xo = g * kXS, qo = q * 17;
            asm volatile("" : "+v"(xo), "+v"(qo));
            dft16<1>(X[i]);
            wave_lds_sync();
#pragma unroll
            for (int x1 = 0; x1 < 16; ++x1) yb[xo + x1 * 17 + q] = X[i][x1];
            wave_lds_sync();
            float2 u[16];
            u[0] = yb[xo + qo];
#pragma unroll
            for (int k1 = 1; k1 < 16; ++k1) u[k1] = cmulc(tq[qo + k1], yb[xo + qo + k1]);
            dft16<1>(u);
#pragma unroll
            for (int x2 = 0; x2 < 8; ++x2) acc[i][x2] = cfma_conj(acc[i][x2], sv[x2], u[x2]);
            // the next row's maps load during that row's transforms
            if (i < 3) load_srow(sv, i + 1);
            __builtin_amdgcn_sched_barrier(0);
        }
        phase_barrier();
    }
    float2* op = a.out + ((int64_t)p * a.N0 + z) * plane;
    const float2* sb = a.sub ? a.sub + ((int64_t)p * a.N0 + z) * plane : nullptr;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int x2 = 0; x2 < 8; ++x2) {
            const int64_t idx = (int64_t)(g + 32 * i) * N + q + 16 * x2;
            float2 rr = acc[i][x2];
            if (a.accumulate) rr = cadd(rr, op[idx]);
            else if (sb) rr = csub(rr, sb[idx]);
            op[idx] = rr;
        }
}

// z-axis pass for L0 = 256 (N0 = 128) as a 16 x 16 four-step FFT with the data in
// registers: each thread owns one (column, group) pair of a COLS-column bundle, does
// 16-point DFTs in registers and meets the others through one LDS transpose per
// direction; loads and stores go straight between HBM and registers. DB: two transpose
// buffers (forward and inverse), so a coil needs two workgroup barriers instead of four.
template <int COLS> struct ZConvShape {
    static constexpr int T = 16 * COLS;                  // threads
    static constexpr int S1 = COLS + 1;                  // LDS stride of the inner index (padded)
    static constexpr int S2 = COLS == 16 ? 280 : 152;    // LDS stride of the outer index (padded)
    static constexpr size_t lds_bytes(bool db) { return sizeof(float2) * (db ? 2 : 1) * 16 * S2; }
};

// Toeplitz kernels of the problems of one launch (problem pl0 + blockIdx.y uses k[blockIdx.y])
constexpr int kZMaxProblems = 16;
struct ZKernels {
    const float* k[kZMaxProblems];
};

template <int COLS, bool DB>
__global__ void __launch_bounds__(ZConvShape<COLS>::T) z_conv_r16_kernel(FusedArgs a, int LL, int pl0, ZKernels zk)
{
    using Sh = ZConvShape<COLS>;
    constexpr int S1 = Sh::S1, S2 = Sh::S2;
    const int pl = pl0 + (int)blockIdx.y;
    const float* __restrict__ K = zk.k[blockIdx.y];
    extern __shared__ float2 lds_all[];  // ZConvShape<COLS>::lds_bytes(DB)
    float2* ldsf = lds_all;
    float2* ldsi = lds_all + (DB ? 16 * S2 : 0);
    const int t = threadIdx.x;
    const int c = t % COLS;
    const int g = t / COLS;
    const int pos0 = blockIdx.x * COLS;
    float2 tw[16];  // w256^{g m}, forward sign
#pragma unroll
    for (int m = 0; m < 16; ++m) tw[m] = a.twL0[(g * m) & 255];
    float kv[16];   // kernel at k = g + 16 k2 of this column
#pragma unroll
    for (int k2 = 0; k2 < 16; ++k2) kv[k2] = K[(int64_t)(g + 16 * k2) * LL + pos0 + c];
    // one coil's half grid [128 z][LL] as a buffer: element (z, pos0 + c) of z = 16 n1 + g
    // at byte voff + n1 * 16 * LL * 8
    auto grid_rsrc = [&](int ci) {
        return plane_rsrc(a.T + ((int64_t)pl * a.ccount + ci) * 128 * (int64_t)LL, (uint32_t)(128 * 8 * (int64_t)LL));
    };
    const uint32_t voff = (uint32_t)(((int64_t)g * LL + pos0 + c) * 8);
    const uint32_t zstep = (uint32_t)(16 * 8 * (int64_t)LL);
    float2 pf[8];
    {
        const auto r0 = grid_rsrc(0);
#pragma unroll
        for (int n1 = 0; n1 < 8; ++n1) pf[n1] = bload<true>(r0, voff, n1 * zstep);
    }

    const int ncl = coil_count(a, a.p0 + pl);
    for (int ci = 0; ci < ncl; ++ci) {
        const auto rc = grid_rsrc(ci);
        float2 v[16];
#pragma unroll
        for (int n1 = 0; n1 < 8; ++n1) v[n1] = pf[n1];
#pragma unroll
        for (int n1 = 8; n1 < 16; ++n1) v[n1] = float2{0.f, 0.f};
        if (ci + 1 < ncl) {
            const auto rn = grid_rsrc(ci + 1);
#pragma unroll
            for (int n1 = 0; n1 < 8; ++n1) pf[n1] = bload<true>(rn, voff, n1 * zstep);
        }
        // forward, step A (n2 = g): DFT over n1 (n1 < 8 nonzero), twiddle w256^{g k1}, transpose
        dft16<-1, true>(v);
#pragma unroll
        for (int k1 = 1; k1 < 16; ++k1) v[k1] = cmul(v[k1], tw[k1]);
#pragma unroll
        for (int k1 = 0; k1 < 16; ++k1) ldsf[k1 * S2 + g * S1 + c] = v[k1];
        lds_barrier();
        // forward, step B (k1 = g): DFT over n2 -> X[g + 16 k2]; times the kernel
#pragma unroll
        for (int n2 = 0; n2 < 16; ++n2) v[n2] = ldsf[g * S2 + n2 * S1 + c];
        dft16<-1>(v);
#pragma unroll
        for (int k2 = 0; k2 < 16; ++k2) v[k2] = cscale(v[k2], kv[k2]);
        // inverse, step B' (k1 = g): IDFT over k2 -> n1, twiddle w256^{-g n1}
        dft16<1>(v);
#pragma unroll
        for (int n1 = 1; n1 < 16; ++n1) v[n1] = cmulc(tw[n1], v[n1]);
        // one buffer: all step-B reads are done before the transpose is overwritten.
        // Two buffers: the next coil's forward writes come after this coil's second
        // barrier, which every thread passes only after its step-B reads, and the next
        // inverse writes after the next first barrier, past this coil's step-A' reads.
        if (!DB) lds_barrier();
#pragma unroll
        for (int n1 = 0; n1 < 16; ++n1) ldsi[n1 * S2 + g * S1 + c] = v[n1];
        lds_barrier();
        // inverse, step A' (n1 = g): IDFT over k1 -> out[g + 16 n2], keep rows < 128
#pragma unroll
        for (int k1 = 0; k1 < 16; ++k1) v[k1] = ldsi[g * S2 + k1 * S1 + c];
        dft16<1>(v);
#pragma unroll
        for (int n2 = 0; n2 < 8; ++n2) bstore<true>(rc, v[n2], voff, n2 * zstep);
        if (!DB) lds_barrier();
    }
}

template <int COLS, bool DB> void launch_zconv(const FusedArgs& a, int LL, int q0, int nq, const ZKernels& zk, hipStream_t s)
{
    const size_t smem = ZConvShape<COLS>::lds_bytes(DB);
    allow_lds(z_conv_r16_kernel<COLS, DB>, smem);
    hipLaunchKernelGGL((z_conv_r16_kernel<COLS, DB>), dim3((unsigned)(LL / COLS), (unsigned)nq),
                       dim3(ZConvShape<COLS>::T), smem, s, a, LL, q0, zk);
}

// z-axis pass for L0 = 320 (N0 = 160) as a 20 x 16 four-step FFT in registers:
// z = 20 n1 + n2 (n1 < 8 nonzero), k = k1 + 16 k2. 16 columns per workgroup and 20 threads
// per column (320 threads): in the 16-point steps thread g of a column owns n2 = g, in
// the 20-point steps (over n2 -> k2 and back) the threads g < 16 own k1 = g and the
// other four wait at the barrier. One LDS transpose per direction; DB: separate forward
// and inverse buffers (two barriers per coil instead of four).
struct ZConv20 {
    static constexpr int COLS = 16, T = 20 * COLS;
    static constexpr int S1 = COLS + 1;  // stride of the inner index (padded)
    static constexpr int SF = 348;       // forward buffer [k1 < 16][n2 < 20 (S1)][c]
    static constexpr int SI = 280;       // inverse buffer [n2 < 20][k1 < 16 (S1)][c]
    // transpose buffer(s), then the w320 table
    static constexpr size_t lds_bytes(bool db)
    {
        return sizeof(float2) * ((db ? 16 * SF + 20 * SI : std::max(16 * SF, 20 * SI)) + 320);
    }
};

template <bool DB>
__global__ void __launch_bounds__(ZConv20::T) z_conv_r20_kernel(FusedArgs a, int LL, int pl0, ZKernels zk)
{
    constexpr int S1 = ZConv20::S1, SF = ZConv20::SF, SI = ZConv20::SI, COLS = ZConv20::COLS;
    const int pl = pl0 + (int)blockIdx.y;
    const float* __restrict__ K = zk.k[blockIdx.y];
    extern __shared__ float2 lds_all[];  // ZConv20::lds_bytes(DB)
    float2* ldsf = lds_all;
    float2* ldsi = lds_all + (DB ? 16 * SF : 0);
    float2* twz = lds_all + (DB ? 16 * SF + 20 * SI : std::max(16 * SF, 20 * SI));  // w320^k, forward sign
    const int t = threadIdx.x;
    const int c = t % COLS;
    const int g = t / COLS;  // 0..19
    const int pos0 = blockIdx.x * COLS;
    for (int i = t; i < 320; i += ZConv20::T) twz[i] = a.twL0[i];
    // the 20-point steps: kernel at k = g + 16 k2 of this column
    float kv[20];
#pragma unroll
    for (int k2 = 0; k2 < 20; ++k2) kv[k2] = g < 16 ? K[(int64_t)(g + 16 * k2) * LL + pos0 + c] : 0.f;
    // one coil's half grid [160 z][LL]: element (z, pos0 + c) of z = 20 n1 + g at byte
    // voff + n1 * 20 * LL * 8
    auto grid_rsrc = [&](int ci) {
        return plane_rsrc(a.T + ((int64_t)pl * a.ccount + ci) * 160 * (int64_t)LL, (uint32_t)(160 * 8 * (int64_t)LL));
    };
    const uint32_t voff = (uint32_t)(((int64_t)g * LL + pos0 + c) * 8);
    const uint32_t zstep = (uint32_t)(20 * 8 * (int64_t)LL);
    float2 pf[8];
    {
        const auto r0 = grid_rsrc(0);
#pragma unroll
        for (int n1 = 0; n1 < 8; ++n1) pf[n1] = bload<true>(r0, voff, n1 * zstep);
    }

    const int ncl = coil_count(a, a.p0 + pl);
    for (int ci = 0; ci < ncl; ++ci) {
        const auto rc = grid_rsrc(ci);
        float2 v[20];
#pragma unroll
        for (int n1 = 0; n1 < 8; ++n1) v[n1] = pf[n1];
#pragma unroll
        for (int n1 = 8; n1 < 16; ++n1) v[n1] = float2{0.f, 0.f};
        if (ci + 1 < ncl) {
            const auto rn = grid_rsrc(ci + 1);
#pragma unroll
            for (int n1 = 0; n1 < 8; ++n1) pf[n1] = bload<true>(rn, voff, n1 * zstep);
        }
        // forward, step A (n2 = g): DFT over n1 (n1 < 8 nonzero), twiddle w320^{g k1}, transpose
        dft16<-1, true>(v);
        if (ci == 0) lds_barrier();  // the twiddle table is written
#pragma unroll
        for (int k1 = 1; k1 < 16; ++k1) v[k1] = cmul(v[k1], twz[(g * k1) % 320]);
#pragma unroll
        for (int k1 = 0; k1 < 16; ++k1) ldsf[k1 * SF + g * S1 + c] = v[k1];
        lds_barrier();
        // forward, step B (k1 = g < 16): DFT over n2 -> X[g + 16 k2] (slot order), times
        // the kernel; inverse, step B': IDFT over k2 -> n2 (slot order). Waves 0-3 hold
        // g < 16 and wave 4 g >= 16, so the branches are uniform per wave; the barriers
        // stay outside them.
        float2 w[20];  // natural order k2 for the inverse
        if (g < 16) {
#pragma unroll
            for (int n2 = 0; n2 < 20; ++n2) v[n2] = ldsf[g * SF + n2 * S1 + c];
            dft20_dr<-1>(v);
#pragma unroll
            for (int k2 = 0; k2 < 20; ++k2) w[k2] = cscale(v[dft_slot<20>(k2)], kv[k2]);
            dft20_dr<1>(w);
        }
        // one buffer: all step-B reads are done before the transpose is overwritten
        if (!DB) lds_barrier();
        if (g < 16) {
            // twiddle w320^{-g n2}, transpose
#pragma unroll
            for (int n2 = 0; n2 < 20; ++n2) {
                const float2 f = w[dft_slot<20>(n2)];
                ldsi[n2 * SI + g * S1 + c] = n2 == 0 ? f : cmulc(twz[(g * n2) % 320], f);
            }
        }
        lds_barrier();
        // inverse, step A' (n2 = g): IDFT over k1 -> out[20 n1 + g], keep n1 < 8
#pragma unroll
        for (int k1 = 0; k1 < 16; ++k1) v[k1] = ldsi[g * SI + k1 * S1 + c];
        dft16<1>(v);
#pragma unroll
        for (int n1 = 0; n1 < 8; ++n1) bstore<true>(rc, v[n1], voff, n1 * zstep);
        if (!DB) lds_barrier();
    }
}

template <bool DB> void launch_zconv20(const FusedArgs& a, int LL, int q0, int nq, const ZKernels& zk, hipStream_t s)
{
    const size_t smem = ZConv20::lds_bytes(DB);
    allow_lds(z_conv_r20_kernel<DB>, smem);
    hipLaunchKernelGGL((z_conv_r20_kernel<DB>), dim3((unsigned)(LL / ZConv20::COLS), (unsigned)nq),
                       dim3(ZConv20::T), smem, s, a, LL, q0, zk);
}

uint32_t radix_code(int L, int* npass)
{
    uint32_t code = 0;
    int n = 0, m = L;
    for (int r : {8, 4, 2, 3, 5}) {
        while (m % r == 0) {
            code |= (uint32_t)r << (4 * n);
            ++n;
            m /= r;
        }
    }
    if (m != 1 || n > 8) throw HastyError(kUnsupported, "fused FFT length not 2,3,5-smooth");
    *npass = n;
    return code;
}

// Permute a natural-order kernel [L0][L][L] (kx fastest) into the plane order of the
// fused kernels: (s, ky, m) for the coset kernels, r16_pos for the L = 256 register ones.
__global__ void permute_kernel_kernel(const float* in, float* out, int L0, int L, int r16)
{
    const int64_t LL = (int64_t)L * L;
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (int64_t)L0 * LL) return;
    const int64_t kz = i / LL;
    const int64_t r = i % LL;
    const int ky = (int)(r / L), kx = (int)(r % L);
    if (r16) {
        out[kz * LL + r16_pos(ky, kx)] = in[i];
        return;
    }
    const int P = L / kW;
    const int s = kx % P, m = kx / P;
    out[kz * LL + ((int64_t)s * L + ky) * kW + m] = in[i];
}

}  // namespace

namespace {
// plane kernels: 0 LDS Stockham, 1 coset kernels with wavefront x DFTs, 2 (default) the
// register kernels where N1 = N2 = 128 and the coset kernels elsewhere, 3 coset kernels
// with the inverse fed by direct-to-LDS loads. The kernel permutation follows the
// variant, so it is set before kernels are prepared.
std::atomic<int> g_xy_variant{2};
std::atomic<int> g_zconv_batch{16};  // problems per z-convolution launch (N0 = 128), 1..16
// z-convolution at N0 = 128: 0 16 columns x 256 threads, one LDS buffer (four barriers
// per coil); 1 8 columns x 128 threads; 2 (default) 16 columns with two buffers (two
// barriers per coil); 3 8 columns with two buffers. At 128^3 x 16 coils, 4 problems per
// launch: 1.85, 3.58, 1.66 and 3.59 ms (8 columns leave 64-byte row segments per wave)
std::atomic<int> g_zconv_variant{2};

bool use_r16(int N) { return g_xy_variant == 2 && N == 128; }
}  // namespace

void sense_fused_set_xy_variant(int v) { g_xy_variant = v; }
void sense_fused_set_zconv_batch(int v) { g_zconv_batch = v; }
int sense_fused_get_xy_variant() { return g_xy_variant; }
int sense_fused_get_zconv_batch() { return g_zconv_batch; }
void sense_fused_set_zconv_variant(int v) { g_zconv_variant = v; }
int sense_fused_get_zconv_variant() { return g_zconv_variant; }

// The z-convolution has register / compact-LDS forms for 2 N0 = 64 k (k <= 5); other
// lengths use z_conv_kernel, which holds two 32-column blocks of the line in LDS.
static bool zconv_fits(int64_t L0)
{
    if (L0 % 64 == 0 && L0 <= 320) return true;
    return sizeof(float2) * (2 * (size_t)L0 * kW + (size_t)L0) <= 160 * 1024;
}

bool sense_fused_supported(const int64_t* n3)
{
    return n3[1] == n3[2] && n3[1] % 32 == 0 && n3[1] >= 32 && n3[1] <= 160 && fft_size_supported((int)(2 * n3[0])) &&
           zconv_fits(2 * n3[0]);
}

bool sense_fused_register_layout(const int64_t* n3) { return use_r16((int)n3[1]); }

void sense_fused_permute_kernel(const float* natural, float* permuted, const int64_t* n3, bool r16, hipStream_t s)
{
    const int L0 = (int)(2 * n3[0]), L = (int)(2 * n3[1]);
    const int64_t n = (int64_t)L0 * L * L;
    HASTY_REQUIRE(!r16 || L == 256, "the register-kernel layout needs N1 = N2 = 128");
    hipLaunchKernelGGL(permute_kernel_kernel, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, s, natural, permuted, L0,
                       L, r16 ? 1 : 0);
    check_launch("permute_kernel_kernel");
}

size_t sense_fused_workspace_bytes(const int64_t* n3, int pchunk, int ccount)
{
    const int64_t L = 2 * n3[1];
    return sizeof(float2) * (size_t)pchunk * ccount * n3[0] * L * L;
}

// kernels: permuted layout (sense_fused_permute_kernel with r16); kidx: host int [P].
void sense_fused_apply(const float2* x, float2* out, const float2* smaps, const int* coils_dev, int ncoil,
                       const float* kernels, bool r16, const int* kidx, const float2* sub, int P, const int64_t* n3,
                       void* ws, size_t ws_bytes, int pchunk, int cchunk, hipStream_t s, int coil_pstride,
                       const int* pcoils_dev)
{
    HASTY_REQUIRE(sense_fused_supported(n3), "fused SENSE normal operator needs N1 == N2 in {32, 64, 96, 128, 160}");
    const int N0 = (int)n3[0], N = (int)n3[1];
    const int NB = N / 32;
    const int L0 = 2 * N0, L = 2 * N;
    pchunk = std::max(1, std::min(std::min(pchunk, P), 64));
    cchunk = std::max(1, std::min(cchunk, ncoil));
    HASTY_REQUIRE(ws_bytes >= sense_fused_workspace_bytes(n3, pchunk, cchunk), "fused workspace too small");
    FusedArgs a;
    a.x = x;
    a.out = out;
    a.sub = sub;
    a.smaps = smaps;
    a.coils = coils_dev;
    a.coil_pstride = coil_pstride;
    a.pcoils = pcoils_dev;
    a.kernels = kernels;
    a.T = reinterpret_cast<float2*>(ws);
    a.twL = twiddles<float>(L);
    a.twL0 = twiddles<float>(L0);
    a.tw32 = twiddles<float>(32);
    a.N0 = N0;
    a.L0 = L0;
    a.codeL = radix_code(L, &a.npassL);
    a.codeL0 = radix_code(L0, &a.npassL0);
    size_t smem_xy = sizeof(float2) * (2 * (size_t)L * kW + L + 32);
    const size_t smem_z = sizeof(float2) * (2 * (size_t)L0 * kW + L0);
    const size_t smem_zct = sizeof(float2) * ((size_t)L0 * kZC + L0);
    decltype(&z_conv_ct_kernel<1>) zct = nullptr;
    if (L0 == 64) zct = z_conv_ct_kernel<1>;
    else if (L0 == 128) zct = z_conv_ct_kernel<2>;
    else if (L0 == 192) zct = z_conv_ct_kernel<3>;
    else if (L0 == 256) zct = z_conv_ct_kernel<4>;
    else if (L0 == 320) zct = z_conv_ct_kernel<5>;
    if (zct) allow_lds(zct, smem_zct);
    auto fwd = NB == 1 ? fwd_xy_kernel<1> : NB == 2 ? fwd_xy_kernel<2> : NB == 3 ? fwd_xy_kernel<3> : fwd_xy_kernel<4>;
    auto inv = NB == 1 ? inv_xy_kernel<1> : NB == 2 ? inv_xy_kernel<2> : NB == 3 ? inv_xy_kernel<3> : inv_xy_kernel<4>;
    int nxy = kXY;
    const int variant = g_xy_variant.load();
    // at N = 160 only the coset kernels (one coset per step) fit in LDS
    if (variant >= 1 || NB == 5) {
        fwd = NB == 1   ? fwd_xy_w_kernel<1>
              : NB == 2 ? fwd_xy_w_kernel<2>
              : NB == 3 ? fwd_xy_w_kernel<3>
              : NB == 4 ? fwd_xy_w_kernel<4>
                        : fwd_xy_w_kernel<5>;
        inv = NB == 1   ? inv_xy_w_kernel<1>
              : NB == 2 ? inv_xy_w_kernel<2>
              : NB == 3 ? inv_xy_w_kernel<3>
              : NB == 4 ? inv_xy_w_kernel<4>
                        : inv_xy_w_kernel<5>;
        nxy = kXYW;
        smem_xy = coset_kernel_lds(NB);
    }
    if (variant == 3 && NB <= 4)  // stage + transform buffers: the LDS of two coset blocks
        inv = NB == 1 ? inv_xy_d_kernel<1> : NB == 2 ? inv_xy_d_kernel<2> : NB == 3 ? inv_xy_d_kernel<3> : inv_xy_d_kernel<4>;
    HASTY_REQUIRE(!r16 || N == 128, "the register-kernel layout needs N1 = N2 = 128");
    if (r16) {
        fwd = fwd_xy_r16_kernel;
        inv = inv_xy_r16_kernel;
        nxy = kQT;
        smem_xy = kQLds;
    }
    allow_lds(fwd, smem_xy);
    allow_lds(inv, smem_xy);
    if (!zct && L0 != 256) allow_lds(z_conv_kernel, smem_z);
    // at N = 160 (always the coset kernels) the inverse runs two workgroups per z-plane
    const unsigned inv_grid_x = inv_w_grid_x(NB, N0);
    for (int p0 = 0; p0 < P; p0 += pchunk) {
        const int np = std::min(pchunk, P - p0);
        a.p0 = p0;
        for (int c0 = 0; c0 < ncoil; c0 += cchunk) {
            a.c0 = c0;
            a.ccount = std::min(cchunk, ncoil - c0);
            a.accumulate = c0 > 0 ? 1 : 0;
            {
                KernelTimer t("sense_fwd_xy", s);
                hipLaunchKernelGGL(fwd, dim3(N0, np), dim3(nxy), smem_xy, s, a);
                check_launch("fwd_xy_kernel");
            }
            const int zvar = g_zconv_variant.load();
            const bool z20 = L0 == 320 && (zvar == 0 || zvar == 2);  // else the LDS Stockham form
            if (L0 == 256 || z20) {
                // one launch per up to g_zconv_batch problems of the chunk
                const int zbatch = g_zconv_batch.load();
                for (int q0 = 0; q0 < np; q0 += zbatch) {
                    const int nq = std::min(zbatch, np - q0);
                    ZKernels zk{};
                    for (int j = 0; j < nq; ++j)
                        zk.k[j] = kernels + (int64_t)kidx[p0 + q0 + j] * L0 * (int64_t)L * L;
                    KernelTimer t("sense_z_conv", s);
                    if (z20) {
                        if (zvar == 2) launch_zconv20<true>(a, L * L, q0, nq, zk, s);
                        else launch_zconv20<false>(a, L * L, q0, nq, zk, s);
                    } else {
                        switch (zvar) {
                        case 1: launch_zconv<8, false>(a, L * L, q0, nq, zk, s); break;
                        case 2: launch_zconv<16, true>(a, L * L, q0, nq, zk, s); break;
                        case 3: launch_zconv<8, true>(a, L * L, q0, nq, zk, s); break;
                        default: launch_zconv<16, false>(a, L * L, q0, nq, zk, s); break;
                        }
                    }
                    check_launch("z_conv_kernel");
                }
            }
            for (int pl = 0; pl < np && L0 != 256 && !z20; ++pl) {
                const float* K = kernels + (int64_t)kidx[p0 + pl] * L0 * (int64_t)L * L;
                KernelTimer t("sense_z_conv", s);
                if (zct) {
                    hipLaunchKernelGGL(zct, dim3((unsigned)(L * L / kZC)), dim3(kZT), smem_zct, s, a, L * L, pl, K);
                } else {
                    hipLaunchKernelGGL(z_conv_kernel, dim3((unsigned)(L * L / kW)), dim3(kZ), smem_z, s, a, L * L, pl,
                                       K);
                }
                check_launch("z_conv_kernel");
            }
            KernelTimer t("sense_inv_xy", s);
            hipLaunchKernelGGL(inv, dim3(inv_grid_x, np), dim3(nxy), smem_xy, s, a);
            check_launch("inv_xy_kernel");
        }
    }
}

}  // namespace hasty_hip
