// Locally-low-rank singular value thresholding over blocks of a [F][E][X][Y][Z] image
// (the device replacement for svt_soft/svt_hard over extract_block/insert_block,
// lib/op/svt.cpp:37-169, 208-243, 354-489).
//
// For one block the Casorati matrix A (R = E*bx*by*bz rows (e,x,y,z), F columns) is
// thresholded through its Gram matrix: G = A^H A = V diag(s^2) V^H, and
//     U diag(s') V^H = A V diag(s'/s) V^H = A W,   W = V diag(g) V^H,
// with g = max(s - lambda, 0)/s (soft) or 1 for the `rank` largest s (hard).
//   1. gram:  MFMA 16x16x4 f32 (FP32 accumulate), upper triangle of G, rows split over
//             workgroups into partial Grams summed in a fixed order (deterministic);
//   2. eig:   cyclic complex Jacobi in FP32 on the upper triangle in LDS, parallel
//             (round-robin) ordering, rotations recorded; the eigenvectors that survive
//             the threshold are rebuilt from them, then W = V diag(g) V^H;
//   3. apply: MFMA A <- A W in place, 16-row slabs staged through LDS.
// Blocks are read and written directly in the image; no gather/scatter copies.
#include <atomic>
#include <algorithm>
#include <vector>

#include "llr.hpp"

namespace hasty_hip {

namespace {

typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int kThreads = 256;
// The eigensolver's workgroup: 256, 512 or 768 threads (option llr_eig_threads); every
// thread owns up to ceil(n (n + 2) / 8 / threads) upper 2x2 blocks of the n x n problem.
// A Jacobi round runs those block updates between two barriers, so more threads shorten
// it until the padded block slots and the waves meeting at each barrier cost more than
// they save. 100 frames, 343 blocks: 256 threads 1.16 ms, 512 1.04 ms, 1024 1.40 ms.
std::atomic<int> g_eig_threads{512};
constexpr int kMaxTilesPerWave = 9;  // F <= 128 -> 8x8 tiles -> 36 upper tiles / 4 waves
constexpr int kMaxSweeps = 15;       // Jacobi sweeps per block (4-6 reach FP32 rounding)
std::atomic<int> g_max_sweeps{kMaxSweeps};  // cap in force (option "llr_max_sweeps")

__device__ __forceinline__ int64_t row_addr(const LlrGeom& g, int3 c, int64_t r)
{
    const int iz = (int)(r % g.bz);
    int64_t t = r / g.bz;
    const int iy = (int)(t % g.by);
    t /= g.by;
    const int ix = (int)(t % g.bx);
    const int64_t e = t / g.bx;
    const int64_t xyz = g.X * g.Y * g.Z;
    return (e * xyz + (int64_t)(c.x + ix) * g.Y * g.Z + (int64_t)(c.y + iy) * g.Z + (c.z + iz)) * g.rstride;
}

// Partial Gram G = A^H A over rows [split * rows_per_split, ...) of each block.
//
// No LDS and no barriers: each wave streams every 16-row slab of the split and owns a
// fixed set of upper 16x16 tiles of G (tile t = wave + 4 k, as compile-time lists). A
// slab is loaded straight into the v_mfma_f32_16x16x4f32 operand layout: lane (m, kg)
// holds A[row][16 t + m] for the four rows 4 kg + ks (ks = 0..3) of frame group t; the
// same registers are the A operand (conjugated through the product below) of the tiles
// in row group t and the B operand of those in column group t. Row order inside a slab
// does not matter for A^H A, so each lane takes four consecutive rows (one row address).
// The four waves of a workgroup read the same slab (cache hits after the first); the
// next slab is loaded while the current one's MFMAs run. conj(a) b takes three real
// MFMAs: P1 = ar br, P2 = ai bi, P3 = (ar - ai)(br + bi); Re = P1 + P2, Im = P3 - P1 + P2.
// Frames past F read frame F - 1 (they only feed entries of G beyond F, which nothing
// reads); rows past the split's end are zero.
// row_addr for r < 2^31 (checked on the host) with 32-bit division.
__device__ __forceinline__ int64_t row_addr32(const LlrGeom& g, int3 c, uint32_t r)
{
    const uint32_t bz = (uint32_t)g.bz, by = (uint32_t)g.by, bx = (uint32_t)g.bx;
    const uint32_t t1 = r / bz, iz = r - t1 * bz;
    const uint32_t t2 = t1 / by, iy = t1 - t2 * by;
    const uint32_t e = t2 / bx, ix = t2 - e * bx;
    return ((int64_t)e * g.X * g.Y * g.Z + (int64_t)(c.x + (int)ix) * g.Y * g.Z + (int64_t)(c.y + (int)iy) * g.Z +
            (c.z + (int)iz)) * g.rstride;
}

__host__ __device__ constexpr int tile_ti(int t, int nt)
{
    int i = 0;
    while (t >= nt - i) {
        t -= nt - i;
        ++i;
    }
    return i;
}
__host__ __device__ constexpr int tile_tj(int t, int nt)
{
    int i = 0;
    while (t >= nt - i) {
        t -= nt - i;
        ++i;
    }
    return i + t;
}

// MFMAs of tiles K.. of a wave's list for one slab; the tile's frame groups are
// constant expressions, so the operands stay in registers.
template <int NT, int WAVE, int K, int NTW> struct GramTiles {
    __device__ __forceinline__ static void mfma(const float2 (&a)[NT][4], floatx4* p1, floatx4* p2, floatx4* p3)
    {
        constexpr int t = WAVE + 4 * K;
        constexpr int ti = tile_ti(t, NT), tj = tile_tj(t, NT);
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) {
            const float2 x = a[ti][ks], y = a[tj][ks];
            p1[K] = __builtin_amdgcn_mfma_f32_16x16x4f32(x.x, y.x, p1[K], 0, 0, 0);
            p2[K] = __builtin_amdgcn_mfma_f32_16x16x4f32(x.y, y.y, p2[K], 0, 0, 0);
            p3[K] = __builtin_amdgcn_mfma_f32_16x16x4f32(x.x - x.y, y.x + y.y, p3[K], 0, 0, 0);
        }
        GramTiles<NT, WAVE, K + 1, NTW>::mfma(a, p1, p2, p3);
    }
};
template <int NT, int WAVE, int NTW> struct GramTiles<NT, WAVE, NTW, NTW> {
    __device__ __forceinline__ static void mfma(const float2 (&)[NT][4], floatx4*, floatx4*, floatx4*) {}
};

// A block lies inside the image: corners handed over on the device are not seen by the
// host, so the kernels skip a block that does not (its Gram is zero, its rows untouched).
__device__ __forceinline__ bool block_inside(const LlrGeom& g, int3 c)
{
    return c.x >= 0 && c.y >= 0 && c.z >= 0 && c.x + g.bx <= g.X && c.y + g.by <= g.Y && c.z + g.bz <= g.Z;
}

template <int NT, int WAVE>
__device__ __forceinline__ void gram_wave(const float2* __restrict__ data, const LlrGeom& g, int3 c, int64_t r_begin,
                                          int64_t r_end, float2* __restrict__ G)
{
    constexpr int Fp = 16 * NT;
    constexpr int NTILE = NT * (NT + 1) / 2;
    constexpr int NTW = NTILE > WAVE ? (NTILE - WAVE + 3) / 4 : 0;
    if constexpr (NTW > 0) {
        const int lane = threadIdx.x & 63;
        const int m = lane & 15, kg = lane >> 4;
        const int F = (int)g.F;
        floatx4 p1[NTW], p2[NTW], p3[NTW];
#pragma unroll
        for (int k = 0; k < NTW; ++k) {
            p1[k] = floatx4{0.f, 0.f, 0.f, 0.f};
            p2[k] = floatx4{0.f, 0.f, 0.f, 0.f};
            p3[k] = floatx4{0.f, 0.f, 0.f, 0.f};
        }
        auto load_slab = [&](int64_t r0, float2 (&a)[NT][4]) {
            int64_t fs = g.fstride;
            int fm = F - 1;
            asm volatile("" : "+s"(fs), "+s"(fm));  // opaque per slab: nothing is hoisted out of the slab loop
            // the lane's rows rr0 .. rr0 + 3: one address when they lie in one z-line
            const int64_t rr0 = r0 + 4 * kg;
            int64_t addr[4];
            bool ok[4];
            if (rr0 + 3 < r_end && (uint32_t)rr0 % (uint32_t)g.bz + 3 < (uint32_t)g.bz) {
                const int64_t base = row_addr32(g, c, (uint32_t)rr0);
                if (g.rstride == 1 && ((base | fs) & 1) == 0) {
                    // four consecutive complex values per frame, 16-byte aligned: two 16-byte loads
#pragma unroll
                    for (int t = 0; t < NT; ++t) {
                        const float2* p = data + base + (int64_t)min(16 * t + m, fm) * fs;
                        const float4 v0 = *reinterpret_cast<const float4*>(p);
                        const float4 v1 = *reinterpret_cast<const float4*>(p + 2);
                        a[t][0] = float2{v0.x, v0.y};
                        a[t][1] = float2{v0.z, v0.w};
                        a[t][2] = float2{v1.x, v1.y};
                        a[t][3] = float2{v1.z, v1.w};
                    }
                    return;
                }
#pragma unroll
                for (int ks = 0; ks < 4; ++ks) {
                    addr[ks] = base + ks * g.rstride;
                    ok[ks] = true;
                }
            } else {
#pragma unroll
                for (int ks = 0; ks < 4; ++ks) {
                    ok[ks] = rr0 + ks < r_end;
                    addr[ks] = row_addr32(g, c, (uint32_t)(ok[ks] ? rr0 + ks : r_begin));
                }
            }
#pragma unroll
            for (int ks = 0; ks < 4; ++ks) {
#pragma unroll
                for (int t = 0; t < NT; ++t) {
                    const float2 v = data[addr[ks] + (int64_t)min(16 * t + m, fm) * fs];
                    a[t][ks] = ok[ks] ? v : float2{0.f, 0.f};
                }
            }
        };
        auto mfma_slab = [&](const float2 (&a)[NT][4]) { GramTiles<NT, WAVE, 0, NTW>::mfma(a, p1, p2, p3); };
        float2 a0[NT][4], a1[NT][4];
        int64_t r0 = r_begin;
        if (r0 < r_end) load_slab(r0, a0);
        while (r0 < r_end) {
            if (r0 + 16 < r_end) load_slab(r0 + 16, a1);
            mfma_slab(a0);
            r0 += 16;
            if (r0 >= r_end) break;
            if (r0 + 16 < r_end) load_slab(r0 + 16, a0);
            mfma_slab(a1);
            r0 += 16;
        }
#pragma unroll
        for (int k = 0; k < NTW; ++k) {
            const int t = WAVE + 4 * k;
            const int ti = tile_ti(t, NT), tj = tile_tj(t, NT);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int i = ti * 16 + 4 * kg + v;
                const int j = tj * 16 + m;
                const float re = p1[k][v] + p2[k][v];
                const float im = p3[k][v] - p1[k][v] + p2[k][v];
                G[(int64_t)i * Fp + j] = float2{re, im};
                if (ti != tj) G[(int64_t)j * Fp + i] = float2{re, -im};
            }
        }
    }
}

template <int NT>
__global__ void __launch_bounds__(kThreads) gram_kernel(const float2* __restrict__ data, LlrGeom g, const int3* corners,
                                                       int64_t rows_per_split, float2* __restrict__ gpart)
{
    constexpr int Fp = 16 * NT;
    const int blk = blockIdx.x;
    const int split = blockIdx.y;
    const int nsplit = gridDim.y;
    const int3 c = corners[blk];
    const int64_t R = g.E * g.bx * g.by * g.bz;
    const int64_t r_begin = split * rows_per_split;
    // a block outside the image gets an empty row range: its partial Gram is zero
    const int64_t r_end = block_inside(g, c) ? min(R, r_begin + rows_per_split) : r_begin;
    float2* G = gpart + ((int64_t)blk * nsplit + split) * Fp * Fp;
    switch (threadIdx.x >> 6) {
    case 0: gram_wave<NT, 0>(data, g, c, r_begin, r_end, G); break;
    case 1: gram_wave<NT, 1>(data, g, c, r_begin, r_end, G); break;
    case 2: gram_wave<NT, 2>(data, g, c, r_begin, r_end, G); break;
    default: gram_wave<NT, 3>(data, g, c, r_begin, r_end, G); break;
    }
}

template <int NT>
void launch_gram(const float2* data, const LlrGeom& g, const int3* corners, int64_t nblk, int nsplit, int64_t rps,
                 float2* gpart, hipStream_t s)
{
    hipLaunchKernelGGL(gram_kernel<NT>, dim3((unsigned)nblk, (unsigned)nsplit), dim3(kThreads), 0, s, data, g,
                       corners, rps, gpart);
    check_launch("gram_kernel");
}

// Upper-triangle block (k, l), k <= l, of an m x m block grid, row-major.
__device__ __forceinline__ int tri_idx(int k, int l, int m) { return (__mul24(k, 2 * m - k + 1) >> 1) + (l - k); }

// Orders the LDS accesses of one wave: what a lane wrote is visible to every lane after it.
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Sum over the workgroup, returned to every thread.
template <int NTH> __device__ double block_sum(double v, double* red)
{
    // the largest power of two <= NTH: the threads above it fold in first
    constexpr int P2 = NTH >= 512 ? 512 : (NTH >= 256 ? 256 : 128);
    static_assert(NTH >= P2 && NTH - P2 <= P2, "workgroup size out of range");
    const int tid = threadIdx.x;
    red[tid] = v;
    __syncthreads();
    if (NTH > P2) {
        if (tid < NTH - P2) red[tid] += red[tid + P2];
        __syncthreads();
    }
    for (int o = P2 / 2; o > 0; o >>= 1) {
        if (tid < o) red[tid] += red[tid + o];
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}

// Parallel (round-robin) Jacobi ordering on n = 2 h indices. In round rr pair x joins
//   a_x = x == 0 ? 0 : 1 + (x - 1 + rr) mod (n - 1),   b_x = 1 + (n - 2 - x + rr) mod (n - 1).
// The eigensolver keeps G in "slot" order: pair x always sits in slots (2x, 2x + 1) =
// (a_x, b_x), and between rounds the data moves by the fixed slot permutation below
// (slot 0 stays, the a side shifts down, the b side shifts up, the ends wrap).
__host__ __device__ inline int slot_next(int s, int h)
{
    if (h == 1 || s == 0) return s;
    const int x = s >> 1;
    if ((s & 1) == 0) return x == 1 ? 1 : 2 * (x - 1);
    return x == h - 1 ? 2 * (h - 1) : 2 * (x + 1) + 1;
}
// Index held by slot s in round 0, and its inverse.
__host__ __device__ inline int slot0_index(int s, int n) { return s == 0 ? 0 : ((s & 1) == 0 ? s >> 1 : n - 1 - (s >> 1)); }
__host__ __device__ inline int index_slot0(int i, int n) { return i < n / 2 ? 2 * i : 2 * (n - 1 - i) + 1; }

// Byte offsets of the eigensolver's LDS arrays for an n x n problem (n = 2 h):
//   Gb  2 x [h (h + 1) / 2 + 1][2] complex  G in slot order as upper 2x2 blocks (s, t) =
//                                   (2k + i, 2l + j), element 2 i + j (rows i = 0 and
//                                   i = 1 of all blocks in two arrays); in a diagonal
//                                   block element 2 (the mirror of 1) is not kept. The
//                                   extra block is a scratch target for threads with
//                                   fewer blocks than the others. Later: up to h
//                                   eigenvectors [h][n].
//   rc  [h] float4                  rotation of each pair: c, s, cos(phi), sin(phi)
//   gk  [n] float                   g of each eigenvalue
//   kl  [n] int                     kept columns
struct EigLayout {
    int rc, gk, kl, bytes;
    __host__ __device__ explicit EigLayout(int n)
    {
        const int h = n / 2;
        rc = 8 * 4 * (h * (h + 1) / 2 + 1);
        gk = rc + 16 * h;
        kl = gk + 4 * n;
        bytes = kl + 4 * n;
    }
};

// Upper 2x2 blocks (pairs k <= l) one thread updates per round, at most
// ceil(64 * 65 / 2 / threads) = 9 (n = 128, 256 threads); the kernel is instantiated per
// (threads, count).
constexpr int kMaxBlkPerThread = 9;
__host__ __device__ inline int eig_blk_per_thread(int n, int nth)
{
    const int half = n / 2;
    return (half * (half + 1) / 2 + nth - 1) / nth;
}

// Complex index of element el = 2 i + j of block b in Gb: the blocks' top rows (elements
// 0, 1) form one array of nb1 float4 and their bottom rows (2, 3) a second one, so the
// lanes of a wave that read consecutive blocks read consecutive 16-byte words.
__device__ __forceinline__ int gel(int b, int el, int nb1) { return (el >> 1) * 2 * nb1 + 2 * b + (el & 1); }

// Where element (s, t) of the slot-ordered matrix lands after the round's permutation:
// (its index in Gb) << 1, with bit 0 set when it lands below the diagonal and is stored
// conjugated.
__device__ __forceinline__ int moved_loc(int s, int t, int h, int nb1)
{
    const int s2 = slot_next(s, h), t2 = slot_next(t, h);
    const int lo = min(s2, t2), hi = max(s2, t2);
    const int idx = gel(tri_idx(lo >> 1, hi >> 1, h), ((lo & 1) << 1) + (hi & 1), nb1);
    return (idx << 1) | (s2 > t2 ? 1 : 0);
}

// Jacobi eigensolver + W = V diag(g) V^H, one workgroup per block.
//
// G (n = F rounded up to even) lives in LDS in slot order as upper 2x2 blocks; thread
// blocks are fixed for the whole solve. A round: every thread loads its blocks and wave 0
// forms the pair rotations from the diagonal blocks (written to `rot` as well); barrier;
// each block becomes J_k^H G_kl J_l and its elements are stored where the next round's
// slot order puts them (the destinations are the same every round); barrier. V is never
// formed: after convergence, the eigenvectors with g != 0 are rebuilt as
// v_k = J_1 J_2 ... J_R e_k by replaying the recorded rotations backwards, one wave per
// vector (its lanes are the pairs, so a round needs no workgroup barrier); that pass
// scales with the number of singular values that survive the threshold.
template <int NTH, int kBlkPerThread>
__global__ void __launch_bounds__(NTH) eig_w_kernel(const float2* __restrict__ gpart, int nsplit, int F, int Fp,
                                                        float thresh, int soft, int rank, int max_sweeps,
                                                        float4* rot_all, float2* __restrict__ Wout,
                                                        float* __restrict__ eigout)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int n = F + (F & 1);
    const int half = n / 2;
    const int nb1 = half * (half + 1) / 2 + 1;  // blocks of the triangle + the scratch block
    const EigLayout L(n);
    float2* Gs = reinterpret_cast<float2*>(smem_raw);
    float4* rcs = reinterpret_cast<float4*>(smem_raw + L.rc);
    float* gk = reinterpret_cast<float*>(smem_raw + L.gk);
    int* kl = reinterpret_cast<int*>(smem_raw + L.kl);
    __shared__ double red[NTH];
    __shared__ int nk_s;
    const int blk = blockIdx.x;
    const int tid = threadIdx.x;
    float4* rot = rot_all + (int64_t)blk * max_sweeps * (n - 1) * half;

    // G in slot order (sum of the partial Grams in split order; the partials are full
    // Hermitian matrices).
    const float2* Gb = gpart + (int64_t)blk * nsplit * Fp * Fp;
    for (int e = tid; e < n * n; e += NTH) {
        const int s = e / n, t = e - (e / n) * n;
        if (t < s) continue;
        const int i = slot0_index(s, n), j = slot0_index(t, n);
        float2 v = float2{0.f, 0.f};
        if (i < F && j < F) {
            for (int sp = 0; sp < nsplit; ++sp) {
                const float2 w = Gb[(int64_t)sp * Fp * Fp + (int64_t)i * Fp + j];
                v.x += w.x;
                v.y += w.y;
            }
            if (s == t) v.y = 0.f;
        }
        Gs[gel(tri_idx(s >> 1, t >> 1, half), ((s & 1) << 1) + (t & 1), nb1)] = v;
    }

    // This thread's blocks: e = tid + NTH i in row-major order over the upper
    // triangle of pairs (k, l), k <= l; and where their elements go after a round.
    int bk[kBlkPerThread], bl[kBlkPerThread], src[kBlkPerThread], dst[kBlkPerThread][4];
    {
        int k = 0, rem = tid;
#pragma unroll
        for (int i = 0; i < kBlkPerThread; ++i) {
            if (i > 0) rem += NTH;
            while (k < half && rem >= half - k) {
                rem -= half - k;
                ++k;
            }
            if (k < half) {
                bk[i] = k;
                bl[i] = k + rem;
                src[i] = gel(tri_idx(k, k + rem, half), 0, nb1);
#pragma unroll
                for (int el = 0; el < 4; ++el)
                    dst[i][el] = moved_loc(2 * k + (el >> 1), 2 * (k + rem) + (el & 1), half, nb1);
            } else {  // no block: read and write the scratch block, uniformly with the rest of the wave
                bk[i] = bl[i] = 0;
                src[i] = gel(nb1 - 1, 0, nb1);
#pragma unroll
                for (int el = 0; el < 4; ++el) dst[i][el] = gel(nb1 - 1, el, nb1) << 1;
            }
        }
    }
    __syncthreads();

    // |G|_F^2 and the off-diagonal mass from the kept elements (off-diagonal ones count twice)
    auto mass = [&](bool off_only) {
        double a = 0.0;
        for (int e = tid; e < n * n; e += NTH) {
            const int s = e / n, t = e - (e / n) * n;
            if (t < s || (off_only && s == t)) continue;
            const float2 v = Gs[gel(tri_idx(s >> 1, t >> 1, half), ((s & 1) << 1) + (t & 1), nb1)];
            a += (s == t ? 1.0 : 2.0) * ((double)v.x * v.x + (double)v.y * v.y);
        }
        return block_sum<NTH>(a, red);
    };
    const double total = mass(false);

    int sweeps = 0;
    for (; sweeps < max_sweeps; ++sweeps) {
        const double offn = mass(true);
        // off-diagonal mass below (1e-5 |G|_F)^2: at FP32 rounding (n eps |G| ~ 6e-6 |G|
        // for n = 100) further sweeps stop paying
        if (offn <= 1e-10 * total || total == 0.0) break;

        for (int round = 0; round < n - 1; ++round) {
            float4 va[kBlkPerThread], vb[kBlkPerThread];
#pragma unroll
            for (int i = 0; i < kBlkPerThread; ++i) {
                va[i] = *reinterpret_cast<const float4*>(Gs + src[i]);            // (2k, 2l), (2k, 2l+1)
                vb[i] = *reinterpret_cast<const float4*>(Gs + src[i] + 2 * nb1);  // (2k+1, 2l), (2k+1, 2l+1)
            }
            if (tid < half) {
                const int bd = tri_idx(tid, tid, half);
                const float4 d0 = *reinterpret_cast<const float4*>(Gs + gel(bd, 0, nb1));
                const float ap = d0.x, dq = Gs[gel(bd, 3, nb1)].x;
                const float2 bpq = float2{d0.z, d0.w};
                const float ab = hypotf(bpq.x, bpq.y);
                float c = 1.f, s = 0.f, cf = 1.f, sf = 0.f;
                if (ab > 1e-30f && ab > 1e-9f * (fabsf(ap) + fabsf(dq))) {
                    // hardware reciprocal / square root (1 ulp): this chain gates every round,
                    // and the rotation applied is the one recorded, with |e^{i phi}| and
                    // c^2 + s^2 equal to 1 to FP32 rounding (tau overflowing gives t = 0)
                    const float inv = __builtin_amdgcn_rcpf(ab);
                    cf = bpq.x * inv;
                    sf = bpq.y * inv;
                    const float tau = (dq - ap) * 0.5f * inv;
                    const float t = copysignf(
                        __builtin_amdgcn_rcpf(fabsf(tau) + __builtin_amdgcn_sqrtf(fmaf(tau, tau, 1.f))), tau);
                    c = __builtin_amdgcn_rsqf(fmaf(t, t, 1.f));
                    s = t * c;
                }
                const float4 r4 = make_float4(c, s, cf, sf);
                rcs[tid] = r4;
                rot[(int64_t)(sweeps * (n - 1) + round) * half + tid] = r4;
            }
            __syncthreads();
            // G <- J^H G J, block by block: the rows of pair k and the columns of pair l
            // form a 2x2 block that maps to J_k^H G_kl J_l (rows: r_p = c r_p - s e^{i phi} r_q,
            // r_q = s r_p + c e^{i phi} r_q; columns the same with e^{-i phi}). Each kept
            // element belongs to exactly one block k <= l, and the diagonal blocks end exactly
            // diagonal and real. Every block was read before the barrier, so the stores into
            // the next round's slot order cannot overwrite an unread element.
#pragma unroll
            for (int i = 0; i < kBlkPerThread; ++i) {
                const bool diag = bk[i] == bl[i];
                const float4 rk = rcs[bk[i]], rl = rcs[bl[i]];
                const float ck = rk.x, sk = rk.y, cl = rl.x, sl = rl.y;
                const float2 phk = float2{rk.z, rk.w};
                const float2 phl = float2{rl.z, -rl.w};
                const float2 a = float2{va[i].x, va[i].y}, bb = float2{va[i].z, va[i].w};
                const float2 cc = diag ? float2{bb.x, -bb.y} : float2{vb[i].x, vb[i].y};
                const float2 d = float2{vb[i].z, vb[i].w};
                const float2 ea = cmul(phk, cc), eb = cmul(phk, d);
                const float2 a1 = float2{ck * a.x - sk * ea.x, ck * a.y - sk * ea.y};
                const float2 b1 = float2{ck * bb.x - sk * eb.x, ck * bb.y - sk * eb.y};
                const float2 c1 = float2{sk * a.x + ck * ea.x, sk * a.y + ck * ea.y};
                const float2 d1 = float2{sk * bb.x + ck * eb.x, sk * bb.y + ck * eb.y};
                const float2 fb = cmul(phl, b1), fd = cmul(phl, d1);
                float2 nv[4];
                nv[0] = float2{cl * a1.x - sl * fb.x, cl * a1.y - sl * fb.y};
                nv[1] = float2{sl * a1.x + cl * fb.x, sl * a1.y + cl * fb.y};
                nv[2] = float2{cl * c1.x - sl * fd.x, cl * c1.y - sl * fd.y};
                nv[3] = float2{sl * c1.x + cl * fd.x, sl * c1.y + cl * fd.y};
                // a diagonal block is Hermitian: real diagonal, the rotated pair element
                // exactly zero, and element 2 (the mirror of 1, with the same destination
                // and the opposite conjugation) written as conj of element 1: the two
                // stores agree
                nv[0].y = diag ? 0.f : nv[0].y;
                nv[3].y = diag ? 0.f : nv[3].y;
                nv[1] = (diag && sk != 0.f) ? float2{0.f, 0.f} : nv[1];
                nv[2] = diag ? float2{nv[1].x, -nv[1].y} : nv[2];
#pragma unroll
                for (int el = 0; el < 4; ++el) {
                    const int loc = dst[i][el];
                    Gs[loc >> 1] = float2{nv[el].x, (loc & 1) ? -nv[el].y : nv[el].y};
                }
            }
            __syncthreads();
        }
    }

    // g_k from the eigenvalues: index k sits in slot index_slot0(k) again (whole sweeps)
    auto eig_of = [&](int k) {
        const int s = index_slot0(k, n);
        return Gs[gel(tri_idx(s >> 1, s >> 1, half), 3 * (s & 1), nb1)].x;
    };
    for (int k = tid; k < n; k += NTH) {
        const float lam = eig_of(k);
        float gv = 0.f;
        if (k < F) {
            if (soft) {
                const float sg = sqrtf(fmaxf(lam, 0.f));
                gv = sg > thresh ? (sg - thresh) / sg : 0.f;
            } else {
                int cnt = 0;
                for (int j = 0; j < F; ++j) {
                    const float lj = eig_of(j);
                    if (lj > lam || (lj == lam && j < k)) ++cnt;
                }
                gv = cnt < rank ? 1.f : 0.f;
            }
            if (eigout) eigout[(int64_t)blk * F + k] = lam;
        }
        gk[k] = gv;
    }
    __syncthreads();
    // kept columns, in increasing order
    const int lane = tid & 63, wave = tid >> 6;
    if (wave == 0) {
        int base = 0;
        for (int c = 0; c < n; c += 64) {
            const int x = c + lane;
            const bool keep = x < n && gk[x] != 0.f;
            const unsigned long long m = __ballot(keep);
            if (keep) kl[base + __popcll(m & ((1ull << lane) - 1ull))] = x;
            base += __popcll(m);
        }
        if (lane == 0) nk_s = base;
    }
    __syncthreads();
    const int nk = nk_s;

    // G is no longer needed: its storage holds up to n/2 eigenvectors X[v][n] at a time.
    float2* X = Gs;
    float2* W = Wout + (int64_t)blk * Fp * Fp;
    const int R = sweeps * (n - 1);
    const int cap = half;
    for (int c0 = 0; c0 < max(nk, 1); c0 += cap) {
        const int nkc = min(cap, nk - c0);
        for (int e = tid; e < nkc * n; e += NTH) {
            const int j = e / n, x = e - (e / n) * n;
            X[e] = float2{x == kl[c0 + j] ? 1.f : 0.f, 0.f};
        }
        __syncthreads();
        if (wave < nkc) {
            // lane m < n/2 applies the rotation of pair m = (a_m, b_m) to the wave's
            // vectors, round R-1 down to 0; rotations are prefetched 8 rounds ahead. The
            // residues of a_m - 1 and b_m - 1 step down by one per round.
            const bool act = lane < half;
            int A = act && lane > 0 ? (lane - 1 + n - 2) % (n - 1) : 0;
            int B = act ? (2 * n - 4 - lane) % (n - 1) : 0;
            float4 cur[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int r = R - 1 - j;
                cur[j] = act && r >= 0 ? rot[(int64_t)r * half + lane] : make_float4(1.f, 0.f, 1.f, 0.f);
            }
            for (int r0 = R - 1; r0 >= 0; r0 -= 8) {
                float4 nxt[8];
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const int r = r0 - 8 - j;
                    nxt[j] = act && r >= 0 ? rot[(int64_t)r * half + lane] : make_float4(1.f, 0.f, 1.f, 0.f);
                }
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    if (r0 - j < 0) break;
                    if (act) {
                        const int p = lane == 0 ? 0 : 1 + A;
                        const int q = 1 + B;
                        const float c = cur[j].x, s = cur[j].y;
                        const float2 eph = float2{cur[j].z, -cur[j].w};
                        for (int v = wave; v < nkc; v += NTH / 64) {
                            float2* xv = X + v * n;
                            const float2 xp = xv[p], xq = xv[q];
                            xv[p] = float2{c * xp.x + s * xq.x, c * xp.y + s * xq.y};
                            xv[q] = cmul(eph, float2{c * xq.x - s * xp.x, c * xq.y - s * xp.y});
                        }
                        A = A == 0 ? n - 2 : A - 1;
                        B = B == 0 ? n - 2 : B - 1;
                    }
                    wave_lds_sync();
                }
#pragma unroll
                for (int j = 0; j < 8; ++j) cur[j] = nxt[j];
            }
        }
        __syncthreads();
        // W (+)= sum_v g_v X_v X_v^H
        for (int e = tid; e < Fp * Fp; e += NTH) {
            const int i = e / Fp, j = e - (e / Fp) * Fp;
            float2 acc = c0 == 0 ? float2{0.f, 0.f} : W[e];
            if (i < F && j < F) {
                for (int v = 0; v < nkc; ++v) {
                    const float g = gk[kl[c0 + v]];
                    const float2 vi = X[v * n + i], vj = X[v * n + j];
                    // v[i] g conj(v[j])
                    acc.x += g * (vi.x * vj.x + vi.y * vj.y);
                    acc.y += g * (vi.y * vj.x - vi.x * vj.y);
                }
            }
            W[e] = acc;
        }
        __syncthreads();
    }
}

// A <- A W for rows [chunk*rows_per_wg, ...) of each block, in place in the image.
//
// Each wave owns 16-row slabs and forms out^T = W^T A^T with v_mfma_f32_16x16x4f32:
// the A^T operand (lane: row l & 15, frame 4 ks + l / 16) comes straight from the image,
// 16 lanes reading one 128-byte z-line of a frame, and the slab's output tile lands with
// the same lane layout, so loads and stores are whole lines. W^T comes from LDS (loaded
// once per workgroup). Complex products take three real MFMAs (Gauss):
//   P1 = Wr Ar, P2 = Wi Ai, P3 = (Wr + Wi)(Ar + Ai);  Re = P1 - P2, Im = P3 - P1 - P2.
// K runs over F rounded up to 4 only. A wave reads a slab whole before writing it and
// slabs are disjoint, so the update is in place; the next slab is loaded during the
// current one's MFMAs, half by half. The main loop has no workgroup barrier.
constexpr int kApplyThreads = 512;
constexpr int kApplySlabsPerWave = 4;

// LDS row stride of W (complex elements): an odd multiple of 16, so the four k rows one
// MFMA operand reads fall in disjoint bank groups.
__host__ __device__ inline int apply_stride(int Fp) { return ((Fp / 16) % 2 == 0) ? Fp + 16 : Fp; }

template <int NT>
__global__ void __launch_bounds__(kApplyThreads) apply_kernel(float2* __restrict__ data, LlrGeom g,
                                                              const int3* corners, int64_t rows_per_wg,
                                                              const float2* __restrict__ Wall)
{
    constexpr int Fp = 16 * NT;
    constexpr int KS = 4 * NT;  // k-steps for F = Fp
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int F = (int)g.F;
    const int K4 = (F + 3) / 4;
    const int S = apply_stride(Fp);
    float2* sW = reinterpret_cast<float2*>(smem_raw);  // W[k][j] at k * S + j
    const int blk = blockIdx.x;
    const int3 c = corners[blk];
    const int64_t R = g.E * g.bx * g.by * g.bz;
    const int64_t r_begin = (int64_t)blockIdx.y * rows_per_wg;
    const int64_t r_end = min(R, r_begin + rows_per_wg);
    if (r_begin >= r_end || !block_inside(g, c)) return;  // the whole workgroup
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int n = lane & 15;   // slab row of this lane
    const int kg = lane >> 4;  // k within a k-step

    const float2* W = Wall + (int64_t)blk * Fp * Fp;
    for (int e = threadIdx.x; e < 4 * K4 * Fp; e += kApplyThreads) {
        const int k = e / Fp, j = e - (e / Fp) * Fp;
        sW[k * S + j] = k < F ? W[(int64_t)k * Fp + j] : float2{0.f, 0.f};
    }
    __syncthreads();

    // one k-step per register: lane holds A[row r0 + n][frame 4 ks + kg]. Frames past F
    // (the last k-step) read frame F - 1 and rows past the end read the last row: both
    // meet zero rows of W or are not stored, so no per-lane load masks are needed.
    // k-steps [ks0, ks1) of one slab: lane holds A[row][frame 4 ks + kg]. Frames past F
    // (the last k-step) read frame F - 1 and rows past the end read the last row: both
    // meet zero rows of W or are not stored, so no per-lane load masks are needed.
    auto slab_row = [&](int64_t r0, int64_t& addr, bool& valid) {
        const int64_t r = r0 + n;
        valid = r < r_end;
        addr = row_addr(g, c, valid ? r : r_end - 1);
    };
    auto load_ks = [&](float2* a, int64_t addr, int ks0, int ks1) {
        int64_t fs = g.fstride;
        int fm = F - 1, k4 = K4;
        asm volatile("" : "+s"(fs), "+s"(fm), "+s"(k4));  // opaque per slab: nothing here is hoisted out of the slab loop
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
            if (ks < ks0 || ks >= ks1) continue;
            if (ks < k4) {
                const int k = min(4 * ks + kg, fm);
                a[ks] = data[addr + (int64_t)k * fs];
            } else {
                a[ks] = float2{0.f, 0.f};
            }
        }
    };

    // The next slab's operands are loaded in two halves into registers that the current
    // slab no longer needs: the first half after the current slab's first-half MFMAs (so
    // its latency hides behind the second half), the second half before the stores.
    constexpr int H = KS / 2;
    const int waves = kApplyThreads / 64;
    float2 a[KS];
    int64_t addr;
    bool valid;
    int64_t r0 = r_begin + 16 * wave;
    if (r0 >= r_end) return;  // no barrier follows
    slab_row(r0, addr, valid);
    load_ks(a, addr, 0, KS);
    for (; r0 < r_end; r0 += 16 * waves) {
        floatx4 p1[NT], p2[NT], p3[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            p1[t] = floatx4{0.f, 0.f, 0.f, 0.f};
            p2[t] = floatx4{0.f, 0.f, 0.f, 0.f};
            p3[t] = floatx4{0.f, 0.f, 0.f, 0.f};
        }
        int woff = kg * S + n, k4 = K4;
        asm volatile("" : "+v"(woff), "+s"(k4));  // per-k-step LDS bases and tests are formed here, not hoisted
        auto mfma_ks = [&](int ks0, int ks1) {
#pragma unroll
            for (int ks = 0; ks < KS; ++ks) {
                if (ks < ks0 || ks >= ks1) continue;
                if (ks < k4) {
                    const float ar = a[ks].x, ai = a[ks].y, as = ar + ai;
                    const float2* wrow = sW + woff + 4 * ks * S;
#pragma unroll
                    for (int t = 0; t < NT; ++t) {
                        const float2 w = wrow[16 * t];
                        p1[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(w.x, ar, p1[t], 0, 0, 0);
                        p2[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(w.y, ai, p2[t], 0, 0, 0);
                        p3[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(w.x + w.y, as, p3[t], 0, 0, 0);
                    }
                }
            }
        };
        const bool more = r0 + 16 * waves < r_end;
        int64_t addr_n = addr;
        bool valid_n = false;
        mfma_ks(0, H);
        if (more) {
            slab_row(r0 + 16 * waves, addr_n, valid_n);
            load_ks(a, addr_n, 0, H);
        }
        mfma_ks(H, KS);
        if (more) load_ks(a, addr_n, H, KS);
        // lane holds out[row n][col 16 t + 4 kg + v]; only the last tile can pass F
        if (valid) {
            int64_t fs = g.fstride;
            asm volatile("" : "+s"(fs));
            float2* q = data + addr + (int64_t)(4 * kg) * fs;
#pragma unroll
            for (int t = 0; t < NT; ++t) {
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    if (t < NT - 1 || 16 * t + 4 * kg + v < F) {
                        const float re = p1[t][v] - p2[t][v];
                        const float im = p3[t][v] - p1[t][v] - p2[t][v];
                        *q = float2{re, im};
                    }
                    q += v < 3 ? fs : 13 * fs;
                }
            }
        }
        addr = addr_n;
        valid = valid_n;
    }
}

size_t apply_smem(int64_t F)
{
    const int Fp = llr_padded(F);
    return sizeof(float) * 2 * (size_t)(4 * ((F + 3) / 4)) * apply_stride(Fp);
}

template <int NT>
void launch_apply(float2* data, const LlrGeom& g, const int3* corners, int64_t nblk, int64_t rows_per_wg,
                  int64_t nchunk, const float2* W, hipStream_t s)
{
    const size_t smem = apply_smem(g.F);
    allow_lds(apply_kernel<NT>, smem);
    hipLaunchKernelGGL(apply_kernel<NT>, dim3((unsigned)nblk, (unsigned)nchunk), dim3(kApplyThreads), smem, s, data,
                       g, corners, rows_per_wg, W);
    check_launch("apply_kernel");
}

template <int NTH, int B>
void launch_eig(const float2* gpart, int nsplit, const LlrGeom& g, int Fp, float thresh, int soft, int rank,
                float4* rot, float2* W, float* eig_out, int64_t nblk, hipStream_t s)
{
    const size_t smem = (size_t)EigLayout((int)(g.F + (g.F & 1))).bytes;
    allow_lds(eig_w_kernel<NTH, B>, smem);
    hipLaunchKernelGGL((eig_w_kernel<NTH, B>), dim3((unsigned)nblk), dim3(NTH), smem, s, gpart, nsplit, (int)g.F, Fp,
                       thresh, soft, rank, g_max_sweeps, rot, W, eig_out);
    check_launch("eig_w_kernel");
}

}  // namespace

int llr_padded(int64_t F) { return (int)(((F + 15) / 16) * 16); }

size_t llr_eig_smem(int64_t F) { return (size_t)EigLayout((int)(F + (F & 1))).bytes; }

// Recorded Jacobi rotations per block: kMaxSweeps sweeps of n - 1 rounds of n / 2 pairs.
static int64_t rot_per_block(int64_t F)
{
    const int64_t n = F + (F & 1);
    return (int64_t)kMaxSweeps * (n - 1) * (n / 2);
}

// Workspace: gpart [nblk][nsplit][Fp][Fp] + W [nblk][Fp][Fp] complex float, then the
// rotations [nblk][kMaxSweeps][n - 1][n / 2] float4.
size_t llr_workspace_bytes(int64_t nblk, int64_t F, int nsplit)
{
    const int64_t Fp = llr_padded(F);
    return sizeof(float2) * (size_t)nblk * (size_t)(nsplit + 1) * Fp * Fp +
           sizeof(float4) * (size_t)nblk * (size_t)rot_per_block(F);
}

void llr_set_max_sweeps(int sweeps)
{
    HASTY_REQUIRE(sweeps >= 1 && sweeps <= kMaxSweeps, "llr_max_sweeps must be in 1..15");
    g_max_sweeps = sweeps;
}

int llr_get_max_sweeps() { return g_max_sweeps; }

void llr_set_eig_threads(int threads)
{
    HASTY_REQUIRE(threads == 256 || threads == 512 || threads == 768, "llr_eig_threads must be 256, 512 or 768");
    g_eig_threads = threads;
}

int llr_get_eig_threads() { return g_eig_threads; }

void llr_check(int64_t F)
{
    HASTY_REQUIRE(F >= 1 && F <= 128, "LLR supports 1..128 columns (frames)");
    // dynamic LDS of the eigensolver plus its static reduction buffer (768 threads at most)
    HASTY_REQUIRE(llr_eig_smem(F) + sizeof(double) * 768 + 16 <= 160 * 1024, "LLR eigensolver tile exceeds LDS");
}

// Threshold the blocks at `corners_dev` (all disjoint) in place.
void llr_blocks(float2* data, const LlrGeom& g, const int3* corners_dev, int64_t nblk, float thresh, int soft, int rank,
                void* ws, size_t ws_bytes, int nsplit, float* eig_out, hipStream_t s)
{
    llr_check(g.F);
    if (nblk <= 0) return;
    HASTY_REQUIRE(nblk < 65536, "too many blocks in one LLR launch");
    HASTY_REQUIRE(g.E * g.bx * g.by * g.bz < (int64_t(1) << 31), "LLR block has too many rows");
    HASTY_REQUIRE(ws_bytes >= llr_workspace_bytes(nblk, g.F, nsplit), "LLR workspace too small");
    const int Fp = llr_padded(g.F);
    const int64_t R = g.E * g.bx * g.by * g.bz;
    float2* gpart = reinterpret_cast<float2*>(ws);
    float2* W = gpart + (int64_t)nblk * nsplit * Fp * Fp;
    float4* rot = reinterpret_cast<float4*>(W + (int64_t)nblk * Fp * Fp);
    int64_t rps = ceil_div(ceil_div(R, nsplit), 16) * 16;
    {
        KernelTimer t("llr_gram", s);
        const float2* d = data;
        switch (Fp / 16) {
        case 1: launch_gram<1>(d, g, corners_dev, nblk, nsplit, rps, gpart, s); break;
        case 2: launch_gram<2>(d, g, corners_dev, nblk, nsplit, rps, gpart, s); break;
        case 3: launch_gram<3>(d, g, corners_dev, nblk, nsplit, rps, gpart, s); break;
        case 4: launch_gram<4>(d, g, corners_dev, nblk, nsplit, rps, gpart, s); break;
        case 5: launch_gram<5>(d, g, corners_dev, nblk, nsplit, rps, gpart, s); break;
        case 6: launch_gram<6>(d, g, corners_dev, nblk, nsplit, rps, gpart, s); break;
        case 7: launch_gram<7>(d, g, corners_dev, nblk, nsplit, rps, gpart, s); break;
        default: launch_gram<8>(d, g, corners_dev, nblk, nsplit, rps, gpart, s); break;
        }
    }
    {
        KernelTimer t("llr_eig", s);
        const int n = (int)(g.F + (g.F & 1));
        const float2* gp = gpart;
        const int nth = g_eig_threads.load();
        const int bpt = eig_blk_per_thread(n, nth);
#define HASTY_EIG(NTH, B) launch_eig<NTH, B>(gp, nsplit, g, Fp, thresh, soft, rank, rot, W, eig_out, nblk, s)
        if (nth == 768) {
            switch (bpt) {
            case 1: HASTY_EIG(768, 1); break;
            case 2: HASTY_EIG(768, 2); break;
            default: HASTY_EIG(768, 3); break;
            }
        } else if (nth == 512) {
            switch (bpt) {
            case 1: HASTY_EIG(512, 1); break;
            case 2: HASTY_EIG(512, 2); break;
            case 3: HASTY_EIG(512, 3); break;
            case 4: HASTY_EIG(512, 4); break;
            default: HASTY_EIG(512, 5); break;
            }
        } else {
            switch (bpt) {
            case 1: HASTY_EIG(256, 1); break;
            case 2: HASTY_EIG(256, 2); break;
            case 3: HASTY_EIG(256, 3); break;
            case 4: HASTY_EIG(256, 4); break;
            case 5: HASTY_EIG(256, 5); break;
            case 6: HASTY_EIG(256, 6); break;
            case 7: HASTY_EIG(256, 7); break;
            case 8: HASTY_EIG(256, 8); break;
            default: HASTY_EIG(256, kMaxBlkPerThread); break;
            }
        }
#undef HASTY_EIG
    }
    {
        const int64_t rows_per_wg = 16 * (kApplyThreads / 64) * kApplySlabsPerWave;
        const int64_t nchunk = ceil_div(R, rows_per_wg);
        HASTY_REQUIRE(nchunk < 65536, "LLR block has too many rows for one launch");
        KernelTimer t("llr_apply", s);
        const float2* Wc = W;
        switch (Fp / 16) {
        case 1: launch_apply<1>(data, g, corners_dev, nblk, rows_per_wg, nchunk, Wc, s); break;
        case 2: launch_apply<2>(data, g, corners_dev, nblk, rows_per_wg, nchunk, Wc, s); break;
        case 3: launch_apply<3>(data, g, corners_dev, nblk, rows_per_wg, nchunk, Wc, s); break;
        case 4: launch_apply<4>(data, g, corners_dev, nblk, rows_per_wg, nchunk, Wc, s); break;
        case 5: launch_apply<5>(data, g, corners_dev, nblk, rows_per_wg, nchunk, Wc, s); break;
        case 6: launch_apply<6>(data, g, corners_dev, nblk, rows_per_wg, nchunk, Wc, s); break;
        case 7: launch_apply<7>(data, g, corners_dev, nblk, rows_per_wg, nchunk, Wc, s); break;
        default: launch_apply<8>(data, g, corners_dev, nblk, rows_per_wg, nchunk, Wc, s); break;
        }
    }
}

}  // namespace hasty_hip
