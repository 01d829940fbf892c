// ES-kernel NUFFT type-1 (spread -> FFT -> deconvolve/crop) and type-2
// (deconvolve/pad -> FFT -> interpolate) on gfx950.
//
// Method: Barnett, Magland, af Klinteberg, "A parallel non-uniform fast Fourier
// transform library based on an 'exponential of semicircle' kernel", SIAM J. Sci.
// Comput. 41(5), 2019: kernel phi(z) = exp(beta (sqrt(1 - (2z/ns)^2) - 1)), ns =
// ceil(-log10(tol/10)) fine-grid cells wide, fine grid 2x oversampled.
//
// setpts bins points into tiles of the fine grid with a device counting sort; the
// spreader gives each (tile, chunk of <= chunk_cap points) one workgroup that
// accumulates into an LDS copy of the tile plus its halo with LDS atomics and then
// flushes the tile to HBM with one global atomic per non-zero cell. In 3-D each lane
// walks the ns planes of its point's window from its own starting plane, and the LDS
// tile's row and plane pitches are padded so that those planes fall into different
// banks: points at one position (every spoke of a radial trajectory crosses k = 0)
// then update different words at any instant instead of serialising on one. Interpolation
// walks points in tile order so the w^3 gathers hit L2.
#include <atomic>
#include <algorithm>
#include <type_traits>
#include <cmath>
#include <vector>

#include "fft.hpp"
#include "nufft.hpp"

namespace hasty_hip {

namespace {

// points per spreader workgroup: complex64 chunks stay small so that a crowded tile
// (the k-space centre of a radial trajectory) spreads over many CUs; complex128 tiles
// (one workgroup per CU) are costlier to clear and flush, so they take more. On a 128^3
// kooshball (ns = 7, M = 147 200) complex64 spreading took 0.446 ms with 16^3 tiles and
// 256 points, 0.428 ms with 20^3 tiles and 128 points, 0.683 ms with 24^3 and 256.
inline int chunk_cap(bool fp64) { return fp64 ? 1024 : 128; }
std::atomic<int> g_tile{0};   // 3-D tile edge with halo of new plans (0: by precision)
std::atomic<int> g_chunk{0};  // points per subproblem of new plans (0: chunk_cap)
std::atomic<int> g_spread_variant{6};  // 3-D spreader, values as nufft_set_spread_variant
// point ordering of new plans: bit 0 super-tile bin order, bit 1 XCD-aware block order;
// 4 (default): 3 for type-1 plans (spreading), 0 for type-2 plans (interpolation)
std::atomic<int> g_locality{4};
constexpr int kThreads = 256;
constexpr double kTwoPi = 6.283185307179586476925286766559;

__device__ __forceinline__ double fine_pos(double x, int64_t nf)
{
    double u = x * (1.0 / kTwoPi);
    u -= floor(u);
    double t = u * (double)nf;
    if (t >= (double)nf) t -= (double)nf;
    return t;
}

// i mod n for the indices the kernels form (tile cells and kernel taps lie within a
// kernel width plus a tile of the grid, so one step usually suffices; the loops keep any
// argument correct without a 64-bit division)
__device__ __forceinline__ int64_t wrapi(int64_t i, int64_t n)
{
    while (i < 0) i += n;
    while (i >= n) i -= n;
    return i;
}

// Tile cell i = (i0 * T1 + i1) * T2 + i2 without integer division: (i + 1/2) / T lies at
// least 1/(2T) from an integer, far above the float rounding of the product for the
// cell counts of a tile (i < 2^20), so the truncated product is the exact quotient.
__device__ __forceinline__ void split_cell(int i, int T1, int T2, float r1, float r2, int& i0, int& i1, int& i2)
{
    const int q = (int)(((float)i + 0.5f) * r2);
    i2 = i - q * T2;
    i0 = (int)(((float)q + 0.5f) * r1);
    i1 = q - i0 * T1;
}

template <typename R> __device__ __forceinline__ R es_eval(R z, R beta, R c)
{
    const R t = (R)1 - c * z * z;
    return t > (R)0 ? exp(beta * (sqrt(t) - (R)1)) : (R)0;
}

// Bin numbering: bins are grouped into super-tiles of kSuper bins per active dimension
// (sup[d] super-tiles, g[d] = kSuper or 1 bins each), numbered super-tile by super-tile,
// so the points of neighbouring tiles sit next to each other in the sorted order and the
// fine-grid cells they share stay in L2 between them.
constexpr int kSuper = 4;
struct BinOrder {
    int g[3];
    int64_t sup[3];
};
__host__ __device__ __forceinline__ int64_t bin_index(const BinOrder& o, const int64_t* b)
{
    int64_t s = 0, l = 0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        s = s * o.sup[d] + b[d] / o.g[d];
        l = l * o.g[d] + b[d] % o.g[d];
    }
    return s * (o.g[0] * o.g[1] * o.g[2]) + l;
}
__host__ __device__ __forceinline__ void bin_coords(const BinOrder& o, int64_t idx, int64_t* b)
{
    const int64_t G = o.g[0] * o.g[1] * o.g[2];
    int64_t s = idx / G, l = idx - (idx / G) * G;
#pragma unroll
    for (int d = 2; d >= 0; --d) {
        b[d] = (s % o.sup[d]) * o.g[d] + l % o.g[d];
        s /= o.sup[d];
        l /= o.g[d];
    }
}
BinOrder bin_order(const int64_t* nbins, int super)
{
    BinOrder o;
    for (int d = 0; d < 3; ++d) {
        o.g[d] = nbins[d] > 1 ? super : 1;
        o.sup[d] = ceil_div(nbins[d], o.g[d]);
    }
    return o;
}
// Number of bin indices (bins of partial super-tiles included; they stay empty).
int64_t bin_range(const int64_t* nbins, int super)
{
    const BinOrder o = bin_order(nbins, super);
    return o.sup[0] * o.sup[1] * o.sup[2] * o.g[0] * o.g[1] * o.g[2];
}

// Workgroups are dispatched to the 8 XCDs in turn (workgroup i to XCD i mod 8), and each
// XCD has its own L2. For a grid padded to a multiple of 8, this maps workgroup i to the
// block the XCD's own contiguous eighth of the work, so neighbouring blocks (neighbouring
// tiles) share an L2. A bijection on [0, gridDim.x).
constexpr int kXcds = 8;
__device__ __forceinline__ int64_t xcd_block(int xcd)
{
    if (!xcd) return blockIdx.x;
    const int64_t per = gridDim.x / kXcds;
    return (int64_t)(blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
}
inline int64_t xcd_grid(int64_t nblocks, int xcd) { return xcd ? ceil_div(nblocks, kXcds) * kXcds : nblocks; }

template <typename R> struct PtsArgs {
    const R* pts[3];
    int64_t nf[3];
    int bs[3];
    int64_t nbins[3];
    BinOrder ord;
    int64_t M;
};

template <typename R> __device__ __forceinline__ int64_t point_bin(const PtsArgs<R>& a, int64_t j)
{
    int64_t b[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if (a.pts[d]) {
            const double t = fine_pos((double)a.pts[d][j], a.nf[d]);
            int64_t bi = (int64_t)(t / (double)a.bs[d]);
            if (bi >= a.nbins[d]) bi = a.nbins[d] - 1;
            b[d] = bi;
        } else {
            b[d] = 0;
        }
    }
    return bin_index(a.ord, b);
}

template <typename R> __global__ void bin_count_kernel(PtsArgs<R> a, int* counts, int* rank)
{
    const int64_t j = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (j >= a.M) return;
    rank[j] = atomicAdd(&counts[point_bin(a, j)], 1);
}

template <typename R> __global__ void bin_scatter_kernel(PtsArgs<R> a, const int* binstart, const int* rank, int* sortidx)
{
    const int64_t j = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (j >= a.M) return;
    sortidx[binstart[point_bin(a, j)] + rank[j]] = (int)j;
}

__global__ void fill_int_kernel(int* p, int64_t n, int v)
{
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void subcount_kernel(const int* counts, int* nsub, int64_t nbins, int cap)
{
    const int64_t b = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (b < nbins) nsub[b] = (counts[b] + cap - 1) / cap;
}

__global__ void subfill_kernel(const int* counts, const int* nsubstart, int* subbin, int* suboff, int64_t nbins,
                               int cap)
{
    const int64_t b = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (b >= nbins) return;
    const int s0 = nsubstart[b];
    const int ns = nsubstart[b + 1] - s0;
    for (int s = 0; s < ns; ++s) {
        subbin[s0 + s] = (int)b;
        suboff[s0 + s] = s * cap;
    }
}

// Exclusive scan of n ints into out[0..n] (out[n] = total), one workgroup.
// Exclusive prefix sum of n ints into out[0..n], out[n] the total, by one workgroup of
// 1024 threads walking tiles of 4096: each thread sums 4 consecutive elements (coalesced
// loads), the waves scan their sums with shuffles, wave totals go through LDS, and a
// running carry links the tiles.
__global__ void __launch_bounds__(1024) scan_kernel(const int* in, int* out, int64_t n)
{
    __shared__ int wsum[16];
    __shared__ int carry_s;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid == 0) carry_s = 0;
    __syncthreads();
    for (int64_t t0 = 0; t0 < n; t0 += 4096) {
        const int64_t b = t0 + 4 * (int64_t)tid;
        int v[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) v[k] = b + k < n ? in[b + k] : 0;
        const int mine = v[0] + v[1] + v[2] + v[3];
        int inc = mine;  // inclusive scan over the wave
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(inc, o);
            if (lane >= o) inc += y;
        }
        if (lane == 63) wsum[wave] = inc;
        const int carry = carry_s;
        __syncthreads();
        int wbase = 0, total = 0;
#pragma unroll
        for (int w = 0; w < 16; ++w) {
            wbase += w < wave ? wsum[w] : 0;
            total += wsum[w];
        }
        int run = carry + wbase + inc - mine;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (b + k < n) out[b + k] = run;
            run += v[k];
        }
        __syncthreads();  // every thread has read carry_s and wsum
        if (tid == 0) carry_s = carry + total;
        __syncthreads();
    }
    if (tid == 0) out[n] = carry_s;
}

template <typename R> __global__ void zero_kernel(cplx<R>* p, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i < n) p[i] = cplx<R>{(R)0, (R)0};
}

template <typename R> struct SpreadArgs {
    PtsArgs<R> pa;
    const int* sortidx;
    const int* binstart;
    const int* subbin;
    const int* suboff;
    const cplx<R>* c;  // strengths [M] (type 1 input) or values out (type 2)
    cplx<R>* cout;
    cplx<R>* fine;
    R beta;
    R cfac;  // 4 / ns^2
    int rp;  // LDS tile pitches (complex elements) between rows (last axis) and planes
    int pp;
    int maxsub;  // points per subproblem at most
    int64_t nsub;  // subproblems (with xcd the grid is padded to a multiple of the XCD count)
    int xcd;       // XCD-aware block order (xcd_block)
};

// LDS tile pitches for tile extents T and kernel width ns. With e complex elements
// spanning the 32 banks, the row pitch rp is odd modulo e and the plane pitch is
// ns * rp modulo e, so row k = ia * ns + ib of a point's window starts at bank group
// rp * k mod e and consecutive rows cycle through all e groups. Complex128 tiles are
// padded only for the window-row spreader (the lane-per-point one measured faster on
// the natural layout).
inline void tile_pitches(const int T[3], int ns, int esize, bool pad, int& rp, int& pp)
{
    if (!pad) {
        rp = T[2];
        pp = T[1] * T[2];
        return;
    }
    const int e = 128 / esize;
    rp = T[2] + ((T[2] % e) % 2 == 0 ? 1 : 0);
    const int plane = T[1] * rp;
    pp = plane + (((ns * rp - plane) % e) + e) % e;
}

template <typename R, int NS, int DIM>
__global__ void __launch_bounds__(kThreads) spread_kernel(SpreadArgs<R> a)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cplx<R>* tile = reinterpret_cast<cplx<R>*>(smem_raw);
    const PtsArgs<R>& pa = a.pa;
    const int64_t sp = xcd_block(a.xcd);
    if (sp >= a.nsub) return;
    const int bin = a.subbin[sp];
    const int start = a.binstart[bin] + a.suboff[sp];
    const int end = min(a.binstart[bin + 1], start + a.maxsub);

    int64_t bb[3];
    bin_coords(pa.ord, bin, bb);
    int T[3];
    int64_t o[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const bool act = d >= 3 - DIM;
        T[d] = act ? pa.bs[d] + NS + 1 : 1;
        o[d] = act ? bb[d] * pa.bs[d] - (NS + 1) / 2 : 0;
    }
    const int cells = T[0] * T[1] * T[2];
    const int rp = a.rp, pp = a.pp;
    for (int i = threadIdx.x; i < T[0] * pp; i += kThreads) tile[i] = cplx<R>{(R)0, (R)0};
    __syncthreads();

    for (int p = start + threadIdx.x; p < end; p += kThreads) {
        const int j = a.sortidx[p];
        const cplx<R> cj = a.c[j];
        R w[2][NS];
        int rel[3] = {0, 0, 0};
        double t0 = 0.0, l00 = 0.0;
#pragma unroll
        for (int d = 3 - DIM; d < 3; ++d) {
            const double t = fine_pos((double)pa.pts[d][j], pa.nf[d]);
            const double l0 = ceil(t - 0.5 * NS);
            rel[d] = (int)((int64_t)l0 - o[d]);
            if (d == 0) {
                t0 = t;
                l00 = l0;
            } else {
#pragma unroll
                for (int q = 0; q < NS; ++q) w[d - 1][q] = es_eval<R>((R)(l0 + q - t), a.beta, a.cfac);
            }
        }
        if (DIM == 3) {
            // planes ia of the window in lane-rotated order (complex64; complex128 tiles
            // measured faster in natural order)
            int ia = std::is_same<R, float>::value ? (int)(threadIdx.x & 63) % NS : 0;
            cplx<R>* base = tile + rel[0] * pp + rel[1] * rp + rel[2];
            for (int k = 0; k < NS; ++k) {
                const R wa = es_eval<R>((R)(l00 + ia - t0), a.beta, a.cfac);
#pragma unroll
                for (int ib = 0; ib < NS; ++ib) {
                    const R wab = wa * w[0][ib];
                    cplx<R>* row = base + ia * pp + ib * rp;
#pragma unroll
                    for (int q = 0; q < NS; ++q) {
                        const R ww = wab * w[1][q];
                        atomicAdd(&row[q].x, ww * cj.x);
                        atomicAdd(&row[q].y, ww * cj.y);
                    }
                }
                if (++ia == NS) ia = 0;
            }
        } else if (DIM == 2) {
#pragma unroll
            for (int ib = 0; ib < NS; ++ib) {
                const R wb = w[0][ib];
                cplx<R>* row = tile + (rel[1] + ib) * rp + rel[2];
#pragma unroll
                for (int q = 0; q < NS; ++q) {
                    const R ww = wb * w[1][q];
                    atomicAdd(&row[q].x, ww * cj.x);
                    atomicAdd(&row[q].y, ww * cj.y);
                }
            }
        } else {
            cplx<R>* row = tile + rel[2];
#pragma unroll
            for (int q = 0; q < NS; ++q) {
                atomicAdd(&row[q].x, w[1][q] * cj.x);
                atomicAdd(&row[q].y, w[1][q] * cj.y);
            }
        }
    }
    __syncthreads();

    const float r1 = 1.0f / (float)T[1], r2 = 1.0f / (float)T[2];
    for (int i = threadIdx.x; i < cells; i += kThreads) {
        int i0, i1, i2;
        split_cell(i, T[1], T[2], r1, r2, i0, i1, i2);
        const cplx<R> v = tile[i0 * pp + i1 * rp + i2];
        if (v.x == (R)0 && v.y == (R)0) continue;
        const int64_t g0 = wrapi(o[0] + i0, pa.nf[0]);
        const int64_t g1 = wrapi(o[1] + i1, pa.nf[1]);
        const int64_t g2 = wrapi(o[2] + i2, pa.nf[2]);
        cplx<R>* dst = a.fine + (g0 * pa.nf[1] + g1) * pa.nf[2] + g2;
        atomicAdd(&dst->x, v.x);
        atomicAdd(&dst->y, v.y);
    }
}

// 3-D spreading by window rows. A batch of kRowBatch points first
// stages its base offset, strength and 3 x NS kernel weights in LDS (one lane per
// point); then the NW waves of the group take the batch's points in turn, one point at a
// time per wave, lane ia * NS + ib updating the NS cells of window row (ia, ib). The
// lanes of one instruction touch NS * NS distinct rows, which the padded pitches place
// in distinct bank groups. With RMW the first wave alone accumulates, updating the tile
// with plain loads and stores (LDS executes one wave's operations in order), and the NW
// waves only clear and flush it; otherwise the waves share the points and the tile
// through LDS atomics.
constexpr int kRowBatch = 32;

template <typename R, int NS, int NW, bool RMW>
__global__ void __launch_bounds__(64 * NW) spread_rows_kernel(SpreadArgs<R> a)
{
    constexpr int S = 3 * NS + 2;  // staged values per point: strength, 3 x NS weights
    __shared__ R stage[kRowBatch * S];
    __shared__ int stage_off[kRowBatch];
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cplx<R>* tile = reinterpret_cast<cplx<R>*>(smem_raw);
    const PtsArgs<R>& pa = a.pa;
    const int64_t sp = xcd_block(a.xcd);
    if (sp >= a.nsub) return;
    const int bin = a.subbin[sp];
    const int start = a.binstart[bin] + a.suboff[sp];
    const int end = min(a.binstart[bin + 1], start + a.maxsub);
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;

    int64_t bb[3];
    bin_coords(pa.ord, bin, bb);
    int T[3];
    int64_t o[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        T[d] = pa.bs[d] + NS + 1;
        o[d] = bb[d] * pa.bs[d] - (NS + 1) / 2;
    }
    const int rp = a.rp, pp = a.pp;
    for (int i = tid; i < T[0] * pp; i += 64 * NW) tile[i] = cplx<R>{(R)0, (R)0};

    for (int b0 = start; b0 < end; b0 += kRowBatch) {
        const int np = min(kRowBatch, end - b0);
        __syncthreads();
        if (tid < np) {
            const int j = a.sortidx[b0 + tid];
            const cplx<R> cj = a.c[j];
            R* st = stage + tid * S;
            int off = 0;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const double t = fine_pos((double)pa.pts[d][j], pa.nf[d]);
                const double l0 = ceil(t - 0.5 * NS);
                const int rel = (int)((int64_t)l0 - o[d]);
                off += rel * (d == 0 ? pp : (d == 1 ? rp : 1));
#pragma unroll
                for (int q = 0; q < NS; ++q) st[2 + d * NS + q] = es_eval<R>((R)(l0 + q - t), a.beta, a.cfac);
            }
            stage_off[tid] = off;
            st[0] = cj.x;
            st[1] = cj.y;
        }
        __syncthreads();
        for (int pt = RMW ? 0 : wave; pt < (RMW && wave > 0 ? 0 : np); pt += RMW ? 1 : NW) {
            const R* st = stage + pt * S;
            const int off = stage_off[pt];
            const R cx = st[0], cy = st[1];
            // lane l updates rows l, l + 64, ... of the NS x NS window rows (ia, ib)
#pragma unroll
            for (int r0 = 0; r0 < NS * NS; r0 += 64) {
                const int r = r0 + lane;
                if (r < NS * NS) {
                    const int ia = r / NS;
                    const int ib = r - ia * NS;
                    const R wab = st[2 + ia] * st[2 + NS + ib];
                    cplx<R>* row = tile + off + ia * pp + ib * rp;
#pragma unroll
                    for (int q = 0; q < NS; ++q) {
                        const R ww = wab * st[2 + 2 * NS + q];
                        if (RMW) {
                            cplx<R> v = row[q];
                            v.x += ww * cx;
                            v.y += ww * cy;
                            row[q] = v;
                        } else {
                            atomicAdd(&row[q].x, ww * cx);
                            atomicAdd(&row[q].y, ww * cy);
                        }
                    }
                }
            }
        }
    }
    __syncthreads();

    const int cells = T[0] * T[1] * T[2];
    const float r1 = 1.0f / (float)T[1], r2 = 1.0f / (float)T[2];
    for (int i = tid; i < cells; i += 64 * NW) {
        int i0, i1, i2;
        split_cell(i, T[1], T[2], r1, r2, i0, i1, i2);
        const cplx<R> v = tile[i0 * pp + i1 * rp + i2];
        if (v.x == (R)0 && v.y == (R)0) continue;
        const int64_t g0 = wrapi(o[0] + i0, pa.nf[0]);
        const int64_t g1 = wrapi(o[1] + i1, pa.nf[1]);
        const int64_t g2 = wrapi(o[2] + i2, pa.nf[2]);
        cplx<R>* dst = a.fine + (g0 * pa.nf[1] + g1) * pa.nf[2] + g2;
        atomicAdd(&dst->x, v.x);
        atomicAdd(&dst->y, v.y);
    }
}

// 3-D spreading with the tile's planes dealt to the waves: wave w owns the tile planes
// i0 with i0 mod NW == w and applies every point of a staged batch to the planes of the
// point's window that it owns, lane r of a chunk of 64 updating cell (k, ib, q) of its
// k-th owned plane. No two waves touch one cell and one wave's LDS operations execute in
// order, so the read-modify-writes need no atomics while all NW waves accumulate at
// once (the window-row form above leaves the accumulation to one wave).
template <typename R, int NS, int NW>
__global__ void __launch_bounds__(64 * NW) spread_planes_kernel(SpreadArgs<R> a)
{
    constexpr int S = 3 * NS + 2;                 // staged values per point
    constexpr int KP = (NS + NW - 1) / NW;        // owned window planes per point, at most
    constexpr int NR = (KP * NS * NS + 63) / 64;  // 64-lane chunks of a point's owned cells
    __shared__ R stage[kRowBatch * S];
    __shared__ int stage_off[kRowBatch];
    __shared__ int stage_p0[kRowBatch];
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cplx<R>* tile = reinterpret_cast<cplx<R>*>(smem_raw);
    const PtsArgs<R>& pa = a.pa;
    const int64_t sp = xcd_block(a.xcd);
    if (sp >= a.nsub) return;
    const int bin = a.subbin[sp];
    const int start = a.binstart[bin] + a.suboff[sp];
    const int end = min(a.binstart[bin + 1], start + a.maxsub);
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;

    int64_t bb[3];
    bin_coords(pa.ord, bin, bb);
    int T[3];
    int64_t o[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        T[d] = pa.bs[d] + NS + 1;
        o[d] = bb[d] * pa.bs[d] - (NS + 1) / 2;
    }
    const int rp = a.rp, pp = a.pp;
    for (int i = tid; i < T[0] * pp; i += 64 * NW) tile[i] = cplx<R>{(R)0, (R)0};

    // this lane's cells of a point, chunk u: owned plane k, window row ib, column q
    int ck[NR], cib[NR], cq[NR];
#pragma unroll
    for (int u = 0; u < NR; ++u) {
        const int r = u * 64 + lane;
        ck[u] = r / (NS * NS);
        const int rem = r - ck[u] * NS * NS;
        cib[u] = rem / NS;
        cq[u] = rem - cib[u] * NS;
    }

    for (int b0 = start; b0 < end; b0 += kRowBatch) {
        const int np = min(kRowBatch, end - b0);
        __syncthreads();
        if (tid < np) {
            const int j = a.sortidx[b0 + tid];
            const cplx<R> cj = a.c[j];
            R* st = stage + tid * S;
            int off = 0;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const double t = fine_pos((double)pa.pts[d][j], pa.nf[d]);
                const double l0 = ceil(t - 0.5 * NS);
                const int rel = (int)((int64_t)l0 - o[d]);
                off += rel * (d == 0 ? pp : (d == 1 ? rp : 1));
                if (d == 0) stage_p0[tid] = rel;
#pragma unroll
                for (int q = 0; q < NS; ++q) st[2 + d * NS + q] = es_eval<R>((R)(l0 + q - t), a.beta, a.cfac);
            }
            stage_off[tid] = off;
            st[0] = cj.x;
            st[1] = cj.y;
        }
        __syncthreads();
        for (int pt = 0; pt < np; ++pt) {
            const R* st = stage + pt * S;
            const int off = stage_off[pt];
            // first window plane ia with (p0 + ia) mod NW == wave (p0 >= 0)
            const int ia0 = (wave - stage_p0[pt] % NW + NW) % NW;
            const R cx = st[0], cy = st[1];
            cplx<R> v[NR];
            int cell[NR];
            R ww[NR];
            bool on[NR];
            // all loads of the point first (distinct cells), then all stores: the next
            // point's loads follow them in program order, which one wave's LDS keeps
#pragma unroll
            for (int u = 0; u < NR; ++u) {
                const int ia = ia0 + ck[u] * NW;
                on[u] = ck[u] < KP && ia < NS;
                if (on[u]) {
                    cell[u] = off + ia * pp + cib[u] * rp + cq[u];
                    ww[u] = st[2 + ia] * st[2 + NS + cib[u]] * st[2 + 2 * NS + cq[u]];
                    v[u] = tile[cell[u]];
                }
            }
#pragma unroll
            for (int u = 0; u < NR; ++u) {
                if (on[u]) {
                    v[u].x += ww[u] * cx;
                    v[u].y += ww[u] * cy;
                    tile[cell[u]] = v[u];
                }
            }
        }
    }
    __syncthreads();

    const int cells = T[0] * T[1] * T[2];
    const float r1 = 1.0f / (float)T[1], r2 = 1.0f / (float)T[2];
    for (int i = tid; i < cells; i += 64 * NW) {
        int i0, i1, i2;
        split_cell(i, T[1], T[2], r1, r2, i0, i1, i2);
        const cplx<R> v = tile[i0 * pp + i1 * rp + i2];
        if (v.x == (R)0 && v.y == (R)0) continue;
        const int64_t g0 = wrapi(o[0] + i0, pa.nf[0]);
        const int64_t g1 = wrapi(o[1] + i1, pa.nf[1]);
        const int64_t g2 = wrapi(o[2] + i2, pa.nf[2]);
        cplx<R>* dst = a.fine + (g0 * pa.nf[1] + g1) * pa.nf[2] + g2;
        atomicAdd(&dst->x, v.x);
        atomicAdd(&dst->y, v.y);
    }
}

// ---- output-driven 3-D spreading ----------------------------------------------------------
//
// The fine grid is cut into cores of kGC^3 cells (kGC divides every fine-grid dim). A
// window of ns <= 16 cells overlaps one or two cores per dim, so each point belongs to at
// most 8 cores; setpts lists, per core, the points whose windows overlap it (the pairs,
// with the window corner relative to the core). One launch gives every core with at most
// kGatherPart such points one workgroup that spreads them into an LDS copy of the core
// and then stores the whole core, zeros included: each of its fine-grid cells is written
// once with plain, whole-line stores, and neither a zero pass nor global atomics are
// needed. A crowded core (the k-space centre of a radial trajectory) is zeroed by a small
// kernel beforehand and split into parts of kGatherPart points whose workgroups add their
// LDS cores to the grid with atomics. The crowded cores' parts are listed first, most
// crowded first, so the long ones start first.
constexpr int kGC = 16;
constexpr int kGatherWaves = 4;
// points per part (staged at once), and the core's LDS row pitch: complex64 cores take
// 128 points and a 19-cell pitch (three workgroups per CU at 51 KB for ns = 7), complex128
// cores 32 points and a 17-cell pitch (two workgroups per CU at 77 KB for ns = 11)
template <typename R> constexpr int gather_part() { return sizeof(R) == 4 ? 128 : 32; }
template <typename R> inline int gather_row_pitch(int ns)
{
    return sizeof(R) == 4 ? (ns < kGC ? kGC + 3 : kGC) : kGC + 1;
}

// ES kernel for complex64 staging with the hardware exp2 and square root (about 1 ulp
// each, far below the tolerances a complex64 plan is made for)
__device__ __forceinline__ float es_eval_fast(float z, float beta, float c)
{
    const float t = 1.f - c * z * z;
    return t > 0.f ? __builtin_amdgcn_exp2f(beta * 1.4426950408889634f * (__builtin_amdgcn_sqrtf(t) - 1.f)) : 0.f;
}

template <typename R> struct GatherBins {
    const R* pts[3];
    int64_t nf[3];
    int nb[3];
    int64_t M;
    int ns;
};

template <typename R> __device__ __forceinline__ int gather_point_bin(const GatherBins<R>& g, int64_t j, double* t)
{
    int b[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        t[d] = fine_pos((double)g.pts[d][j], g.nf[d]);
        int bi = (int)(t[d] * (1.0 / kGC));
        b[d] = bi < g.nb[d] ? bi : g.nb[d] - 1;
    }
    return (b[0] * g.nb[1] + b[1]) * g.nb[2] + b[2];
}

// One atomic per distinct bin of a wave: points along a trajectory come in runs that share
// a bin, and the k-space centre of a radial trajectory would otherwise take one atomic
// per spoke on a single counter.
template <typename R> __global__ void gather_count_kernel(GatherBins<R> g, int* counts, int* rank)
{
    const int64_t j = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    const int lane = threadIdx.x & 63;
    double t[3];
    const int bin = j < g.M ? gather_point_bin(g, j, t) : -1;
    bool pending = j < g.M;
    while (__ballot(pending)) {
        const unsigned long long act = __ballot(pending);
        const int leader = __ffsll((long long)act) - 1;
        const int lb = __shfl(bin, leader);
        const bool same = pending && bin == lb;
        const unsigned long long m = __ballot(same);
        int base = 0;
        if (lane == leader) base = atomicAdd(&counts[lb], __popcll(m));
        base = __shfl(base, leader);
        if (same) {
            rank[j] = base + __popcll(m & ((1ull << lane) - 1ull));
            pending = false;
        }
    }
}

// points in core-bin order: first window cell l0 and fine-grid position minus l0 per dim
template <typename R>
__global__ void gather_scatter_kernel(GatherBins<R> g, const int* binstart, const int* rank, int* l0, R* frac, int* idx)
{
    const int64_t j = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (j >= g.M) return;
    double t[3];
    const int64_t p = binstart[gather_point_bin(g, j, t)] + rank[j];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double l = ceil(t[d] - 0.5 * g.ns);
        l0[d * g.M + p] = (int)l;
        frac[d * g.M + p] = (R)(t[d] - l);
    }
    idx[p] = (int)j;
}

// (point, core) pairs of core ci with neighbour k of its 27 (periodic images included):
// the points of that neighbouring core bin whose windows overlap the core, with the
// window corner relative to the core packed 8 bits per dim (+16). One wave per
// (core, neighbour), no atomics: COUNT writes the count to pcount[27 ci + k], else the
// pairs go to pstart[27 ci + k] on in bin order (ballot compaction), so a core's pairs
// are contiguous and their order is fixed by the core-bin order of the points.
template <typename R, bool COUNT>
__global__ void __launch_bounds__(64) gather_pairs_kernel(GatherBins<R> g, const int* binstart, const int* l0,
                                                          int* pcount, const int* pstart, int* pairs, int* prel)
{
    const int ci = blockIdx.x / 27, k = blockIdx.x % 27;
    const int lane = threadIdx.x;
    const int cb[3] = {ci / (g.nb[2] * g.nb[1]), (ci / g.nb[2]) % g.nb[1], ci % g.nb[2]};
    const int dd[3] = {k / 9 - 1, (k / 3) % 3 - 1, k % 3 - 1};
    int bn[3], rel0[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        int x = cb[d] + dd[d];
        int sh = 0;
        if (x < 0) {
            x += g.nb[d];
            sh = -(int)g.nf[d];
        } else if (x >= g.nb[d]) {
            x -= g.nb[d];
            sh = (int)g.nf[d];
        }
        bn[d] = x;
        rel0[d] = sh - cb[d] * kGC;  // window cell + this = cell relative to the core
    }
    const int bin = (bn[0] * g.nb[1] + bn[1]) * g.nb[2] + bn[2];
    const int b0 = binstart[bin], V = binstart[bin + 1] - b0;
    int out = COUNT ? 0 : pstart[blockIdx.x];
    for (int v0 = 0; v0 < V; v0 += 64) {
        const int v = v0 + lane;
        bool acc = false;
        int p = 0, rel = 0;
        if (v < V) {
            p = b0 + v;
            acc = true;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const int r = l0[d * g.M + p] + rel0[d];
                acc = acc && r < kGC && r + g.ns > 0;
                rel |= (r + 16) << (8 * d);
            }
        }
        const unsigned long long m = __ballot(acc);
        if (!COUNT && acc) {
            const int e = out + __popcll(m & ((1ull << lane) - 1ull));
            pairs[e] = p;
            prel[e] = rel;
        }
        out += __popcll(m);
    }
    if (COUNT && lane == 0) pcount[blockIdx.x] = out;
}

// pairs per core from the scanned (core, neighbour) starts
__global__ void core_pairs_kernel(const int* pstart, int* cnt, int64_t ncore)
{
    const int64_t c = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (c < ncore) cnt[c] = pstart[27 * c + 27] - pstart[27 * c];
}

template <typename R> struct GatherArgs {
    const R* frac;       // [3][M], core-bin order
    const int* idx;      // [M]
    const int* pstart;   // 27 cores + 1: core ci's pairs are [pstart[27 ci], pstart[27 ci + 27])
    const int* pairs;    // sorted point of each pair
    const int* prel;     // window corner relative to the core, 8 bits per dim (+16)
    const int* list;     // (core, part) pairs of this launch
    const cplx<R>* c;
    cplx<R>* fine;
    int64_t M;
    int nb[3];
    int64_t nf[3];
    R beta, cfac;
    int rp;    // LDS row pitch of the core (float2), gather_row_pitch
    int natomic;  // list entries [0, natomic) (the crowded cores' parts) add with atomics
};

template <typename R> __device__ __forceinline__ R gather_es(R z, R beta, R c)
{
    if constexpr (sizeof(R) == 4)
        return es_eval_fast(z, beta, c);
    else
        return es_eval<R>(z, beta, c);
}

template <typename R, int NS>
__global__ void __launch_bounds__(64 * kGatherWaves) gather_spread_kernel(GatherArgs<R> a)
{
    using C2 = cplx<R>;
    constexpr int kGatherPart = gather_part<R>();
    constexpr int kGatherRest = gather_part<R>();  // points of the parts after the first
    constexpr int NW = kGatherWaves;
    constexpr int NT = 64 * NW;
    constexpr int S = 3 * NS + 2;
    constexpr int KP = (NS + NW - 1) / NW;   // owned window planes per point, at most
    constexpr int NQ = (NS * NS + 63) / 64;  // 64-lane chunks of a plane's NS x NS cells
    static_assert(kGatherPart <= NT, "one staging round per part");
    static_assert((NW & (NW - 1)) == 0, "plane ownership by a mask");
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int rp = a.rp;
    C2* core = reinterpret_cast<C2*>(smem_raw);
    R* stage = reinterpret_cast<R*>(core + kGC * kGC * rp);
    int* stage_p = reinterpret_cast<int*>(stage + kGatherPart * S);
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int ci = a.list[2 * blockIdx.x];
    const int part = a.list[2 * blockIdx.x + 1];
    // pairs of this part: the first part holds kGatherPart of the core's pairs, each
    // further part kGatherRest (staged kGatherPart at a time)
    const int pbase = a.pstart[27 * ci];
    const int cnt = a.pstart[27 * ci + 27] - pbase;
    const int lo = part == 0 ? 0 : kGatherPart + (part - 1) * kGatherRest;
    const int hi = min(cnt, part == 0 ? kGatherPart : lo + kGatherRest);

    for (int i = tid; i < kGC * kGC * rp; i += NT) core[i] = C2{(R)0, (R)0};
    // this lane's window rows and columns: chunk u holds (ib, q) = divmod(64 u + lane, NS)
    int cib[NQ], cq[NQ];
#pragma unroll
    for (int u = 0; u < NQ; ++u) {
        const int r = u * 64 + lane;
        cib[u] = r / NS;
        cq[u] = r - cib[u] * NS;
    }

    for (int r0 = lo; r0 < hi; r0 += kGatherPart) {
        const int np = min(kGatherPart, hi - r0);
        __syncthreads();  // the previous round is accumulated (and the core cleared)
        if (tid < np) {
            const int p = a.pairs[pbase + r0 + tid];
            const int rel = a.prel[pbase + r0 + tid];
            const C2 cj = a.c[a.idx[p]];
            R* st = stage + tid * S;
            st[0] = cj.x;
            st[1] = cj.y;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const R fr = a.frac[d * a.M + p];
#pragma unroll
                for (int q = 0; q < NS; ++q) st[2 + d * NS + q] = gather_es<R>((R)q - fr, a.beta, a.cfac);
                stage_p[3 * tid + d] = ((rel >> (8 * d)) & 255) - 16;
            }
        }
        __syncthreads();
        // Wave w owns the core planes i0 with i0 mod NW == w. Per point: each lane's
        // row-column cell offset and weight (strength folded in), then per owned plane
        // (a wave-uniform branch) the loads of its cells; while they are in flight, the
        // next point's row-column part; then the adds and stores. One wave's LDS
        // operations execute in order, so the next point's loads see these stores.
        int off[NQ];
        C2 wc[NQ];
        int p0 = 0;
        auto prep = [&](int e, int* o, C2* w, int& q0) {
            const R* st = stage + e * S;
            const int p1 = stage_p[3 * e + 1];
            const int p2 = stage_p[3 * e + 2];
            q0 = stage_p[3 * e];
            const R cx = st[0], cy = st[1];
#pragma unroll
            for (int u = 0; u < NQ; ++u) {
                const int i1 = p1 + cib[u], i2 = p2 + cq[u];
                const bool in = u * 64 + lane < NS * NS && (unsigned)i1 < (unsigned)kGC && (unsigned)i2 < (unsigned)kGC;
                o[u] = in ? i1 * rp + i2 : -1;
                const R w12 = in ? st[2 + NS + cib[u]] * st[2 + 2 * NS + cq[u]] : (R)0;
                w[u] = C2{w12 * cx, w12 * cy};
            }
        };
        if (np > 0) prep(0, off, wc, p0);
        for (int e = 0; e < np; ++e) {
            const R* st = stage + e * S;
            const int ia0 = (wave - p0) & (NW - 1);
            C2 val[KP][NQ];
            int cell0[KP];
            R w0[KP];
            bool pl[KP];
#pragma unroll
            for (int k = 0; k < KP; ++k) {
                const int ia = ia0 + k * NW;
                const int i0 = p0 + ia;
                pl[k] = ia < NS && (unsigned)i0 < (unsigned)kGC;  // wave-uniform
                cell0[k] = i0 * kGC * rp;
                w0[k] = pl[k] ? st[2 + ia] : (R)0;
                if (pl[k]) {
#pragma unroll
                    for (int u = 0; u < NQ; ++u)
                        if (off[u] >= 0) val[k][u] = core[cell0[k] + off[u]];
                }
            }
            int noff[NQ];
            C2 nwc[NQ];
            int np0 = 0;
            if (e + 1 < np) {
                prep(e + 1, noff, nwc, np0);
            } else {
#pragma unroll
                for (int u = 0; u < NQ; ++u) {
                    noff[u] = -1;
                    nwc[u] = C2{(R)0, (R)0};
                }
            }
#pragma unroll
            for (int k = 0; k < KP; ++k) {
                if (pl[k]) {
                    const R wa = w0[k];
#pragma unroll
                    for (int u = 0; u < NQ; ++u) {
                        if (off[u] >= 0) {
                            val[k][u].x += wa * wc[u].x;
                            val[k][u].y += wa * wc[u].y;
                            core[cell0[k] + off[u]] = val[k][u];
                        }
                    }
                }
            }
#pragma unroll
            for (int u = 0; u < NQ; ++u) {
                off[u] = noff[u];
                wc[u] = nwc[u];
            }
            p0 = np0;
        }
    }
    __syncthreads();

    int b[3];
    b[2] = ci % a.nb[2];
    b[1] = (ci / a.nb[2]) % a.nb[1];
    b[0] = ci / (a.nb[2] * a.nb[1]);
    const int64_t o0 = (int64_t)b[0] * kGC, o1 = (int64_t)b[1] * kGC, o2 = (int64_t)b[2] * kGC;
    for (int i = tid; i < kGC * kGC * kGC; i += NT) {
        const int i0 = i / (kGC * kGC), i1 = (i / kGC) % kGC, i2 = i % kGC;
        const C2 v = core[(i0 * kGC + i1) * rp + i2];
        C2* dst = a.fine + ((o0 + i0) * a.nf[1] + (o1 + i1)) * a.nf[2] + (o2 + i2);
        if ((int)blockIdx.x >= a.natomic) {
            *dst = v;
        } else if (v.x != (R)0 || v.y != (R)0) {
            atomicAdd(&dst->x, v.x);
            atomicAdd(&dst->y, v.y);
        }
    }
}

// zeros the cores whose parts all add with atomics (before the spreading launch)
template <typename R> __global__ void zero_cores_kernel(cplx<R>* fine, const int* cores, int nb1, int nb2, int64_t nf1,
                                                        int64_t nf2)
{
    const int ci = cores[blockIdx.x];
    const int64_t o0 = (int64_t)(ci / (nb2 * nb1)) * kGC, o1 = (int64_t)((ci / nb2) % nb1) * kGC,
                  o2 = (int64_t)(ci % nb2) * kGC;
    for (int i = threadIdx.x; i < kGC * kGC * kGC; i += kThreads) {
        const int i0 = i / (kGC * kGC), i1 = (i / kGC) % kGC, i2 = i % kGC;
        fine[((o0 + i0) * nf1 + (o1 + i1)) * nf2 + (o2 + i2)] = cplx<R>{(R)0, (R)0};
    }
}

// LDS bytes of a core and its stage. The complex64 pitch puts lane (ib, q) of a window
// plane at bank pair (3 ib + q) mod 16, at most 4 lanes per pair for ns <= 9; a window of
// 16 columns covers the banks by itself.
template <typename R> inline size_t gather_lds_bytes(int ns)
{
    return sizeof(cplx<R>) * kGC * kGC * gather_row_pitch<R>(ns) + sizeof(R) * gather_part<R>() * (3 * ns + 2) +
           sizeof(int) * gather_part<R>() * 3;
}

template <typename R, int NS> void launch_gather_ns(const GatherArgs<R>& a, int64_t blocks, hipStream_t s)
{
    const size_t smem = gather_lds_bytes<R>(NS);
    allow_lds(gather_spread_kernel<R, NS>, smem);
    hipLaunchKernelGGL((gather_spread_kernel<R, NS>), dim3((unsigned)blocks), dim3(64 * kGatherWaves), smem, s, a);
}

template <typename R> void launch_gather(int ns, const GatherArgs<R>& a, int64_t blocks, hipStream_t s)
{
    switch (ns) {
#define HASTY_GATHER_CASE(K) \
    case K: launch_gather_ns<R, K>(a, blocks, s); break;
        HASTY_GATHER_CASE(2)
        HASTY_GATHER_CASE(3)
        HASTY_GATHER_CASE(4)
        HASTY_GATHER_CASE(5)
        HASTY_GATHER_CASE(6)
        HASTY_GATHER_CASE(7)
        HASTY_GATHER_CASE(8)
        HASTY_GATHER_CASE(9)
        HASTY_GATHER_CASE(10)
        HASTY_GATHER_CASE(11)
        HASTY_GATHER_CASE(12)
        HASTY_GATHER_CASE(13)
        HASTY_GATHER_CASE(14)
        HASTY_GATHER_CASE(15)
        HASTY_GATHER_CASE(16)
#undef HASTY_GATHER_CASE
    default: throw HastyError(kUnsupported, "kernel width out of range");
    }
    check_launch("gather_spread_kernel");
}

template <typename R, int NS, int DIM>
__global__ void __launch_bounds__(kThreads) interp_kernel(SpreadArgs<R> a)
{
    const PtsArgs<R>& pa = a.pa;
    const int64_t p = xcd_block(a.xcd) * kThreads + threadIdx.x;
    if (p >= pa.M) return;
    const int j = a.sortidx[p];
    R w[2][NS];
    int64_t idx[2][NS];
    double t0 = 0.0, l00 = 0.0;
#pragma unroll
    for (int d = 3 - DIM; d < 3; ++d) {
        const double t = fine_pos((double)pa.pts[d][j], pa.nf[d]);
        const double l0 = ceil(t - 0.5 * NS);
        if (d == 0) {
            t0 = t;
            l00 = l0;
        } else {
            const int64_t il0 = (int64_t)l0;
#pragma unroll
            for (int q = 0; q < NS; ++q) {
                w[d - 1][q] = es_eval<R>((R)(l0 + q - t), a.beta, a.cfac);
                idx[d - 1][q] = wrapi(il0 + q, pa.nf[d]);
            }
        }
    }
    R sr = 0, si = 0;
    if (DIM == 3) {
        // complex64 rows that do not wrap: the window's cells from the even cell at or
        // below its first one, read as 16-byte pairs (W / 2 loads instead of NS) against
        // the weights shifted by that offset
        constexpr int W = (NS + 2) & ~1;
        R w8[W];
        bool pairs = false;
        int64_t a2 = 0;
        if constexpr (std::is_same<R, float>::value) {
            const int off = (int)(idx[1][0] & 1);
            a2 = idx[1][0] - off;
            pairs = a2 + W <= pa.nf[2] && idx[1][NS - 1] == idx[1][0] + NS - 1;
#pragma unroll
            for (int k = 0; k < W; ++k) {
                R v = (R)0;
#pragma unroll
                for (int q = 0; q < NS; ++q) v = (k == q + off) ? w[1][q] : v;
                w8[k] = v;
            }
        }
        for (int ia = 0; ia < NS; ++ia) {
            const R wa = es_eval<R>((R)(l00 + ia - t0), a.beta, a.cfac);
            const int64_t i0 = wrapi((int64_t)l00 + ia, pa.nf[0]);
#pragma unroll
            for (int ib = 0; ib < NS; ++ib) {
                const R wab = wa * w[0][ib];
                const cplx<R>* row = a.fine + (i0 * pa.nf[1] + idx[0][ib]) * pa.nf[2];
                R tr = 0, ti = 0;
                if (pairs) {
                    const float4* r4 = reinterpret_cast<const float4*>(row + a2);
#pragma unroll
                    for (int k = 0; k < W / 2; ++k) {
                        const float4 g = r4[k];
                        tr += w8[2 * k] * g.x + w8[2 * k + 1] * g.z;
                        ti += w8[2 * k] * g.y + w8[2 * k + 1] * g.w;
                    }
                } else {
#pragma unroll
                    for (int q = 0; q < NS; ++q) {
                        const cplx<R> g = row[idx[1][q]];
                        tr += w[1][q] * g.x;
                        ti += w[1][q] * g.y;
                    }
                }
                sr += wab * tr;
                si += wab * ti;
            }
        }
    } else if (DIM == 2) {
#pragma unroll
        for (int ib = 0; ib < NS; ++ib) {
            const cplx<R>* row = a.fine + idx[0][ib] * pa.nf[2];
            R tr = 0, ti = 0;
#pragma unroll
            for (int q = 0; q < NS; ++q) {
                const cplx<R> g = row[idx[1][q]];
                tr += w[1][q] * g.x;
                ti += w[1][q] * g.y;
            }
            sr += w[0][ib] * tr;
            si += w[0][ib] * ti;
        }
    } else {
#pragma unroll
        for (int q = 0; q < NS; ++q) {
            const cplx<R> g = a.fine[idx[1][q]];
            sr += w[1][q] * g.x;
            si += w[1][q] * g.y;
        }
    }
    a.cout[j] = cplx<R>{sr, si};
}

template <typename R, int NS, int DIM>
void launch_spread_ns(const SpreadArgs<R>& a, int64_t nsub, size_t smem, hipStream_t s, int variant)
{
    if constexpr (DIM == 3) {
        if (variant == 1) {
            allow_lds(spread_rows_kernel<R, NS, 1, true>, smem);
            KernelTimer t("nufft_spread", s);
            hipLaunchKernelGGL((spread_rows_kernel<R, NS, 1, true>), dim3((unsigned)xcd_grid(nsub, a.xcd)), dim3(64), smem, s,
                               a);
            return;
        }
        if (variant == 2) {
            allow_lds(spread_rows_kernel<R, NS, 4, false>, smem);
            KernelTimer t("nufft_spread", s);
            hipLaunchKernelGGL((spread_rows_kernel<R, NS, 4, false>), dim3((unsigned)xcd_grid(nsub, a.xcd)), dim3(256), smem,
                               s, a);
            return;
        }
        if (variant == 3) {
            allow_lds(spread_rows_kernel<R, NS, 4, true>, smem);
            KernelTimer t("nufft_spread", s);
            hipLaunchKernelGGL((spread_rows_kernel<R, NS, 4, true>), dim3((unsigned)xcd_grid(nsub, a.xcd)), dim3(256), smem,
                               s, a);
            return;
        }
        if (variant == 4) {
            allow_lds(spread_planes_kernel<R, NS, 4>, smem);
            KernelTimer t("nufft_spread", s);
            hipLaunchKernelGGL((spread_planes_kernel<R, NS, 4>), dim3((unsigned)xcd_grid(nsub, a.xcd)), dim3(256), smem, s,
                               a);
            return;
        }
        if (variant == 5) {
            allow_lds(spread_planes_kernel<R, NS, 8>, smem);
            KernelTimer t("nufft_spread", s);
            hipLaunchKernelGGL((spread_planes_kernel<R, NS, 8>), dim3((unsigned)xcd_grid(nsub, a.xcd)), dim3(512), smem, s,
                               a);
            return;
        }
    }
    allow_lds(spread_kernel<R, NS, DIM>, smem);
    KernelTimer t("nufft_spread", s);
    hipLaunchKernelGGL((spread_kernel<R, NS, DIM>), dim3((unsigned)xcd_grid(nsub, a.xcd)), dim3(kThreads), smem, s, a);
}
template <typename R, int NS, int DIM> void launch_interp_ns(const SpreadArgs<R>& a, int64_t M, hipStream_t s)
{
    KernelTimer t("nufft_interp", s);
    hipLaunchKernelGGL((interp_kernel<R, NS, DIM>), dim3((unsigned)xcd_grid(ceil_div(M, kThreads), a.xcd)), dim3(kThreads), 0,
                       s, a);
}

template <typename R, int DIM>
void dispatch_ns(bool spread, int ns, const SpreadArgs<R>& a, int64_t n, size_t smem, hipStream_t s, int variant)
{
#define HASTY_NS_CASE(K)                                              \
    case K:                                                           \
        if (spread)                                                   \
            launch_spread_ns<R, K, DIM>(a, n, smem, s, variant);      \
        else                                                          \
            launch_interp_ns<R, K, DIM>(a, n, s);                     \
        break;
    switch (ns) {
        HASTY_NS_CASE(2)
        HASTY_NS_CASE(3)
        HASTY_NS_CASE(4)
        HASTY_NS_CASE(5)
        HASTY_NS_CASE(6)
        HASTY_NS_CASE(7)
        HASTY_NS_CASE(8)
        HASTY_NS_CASE(9)
        HASTY_NS_CASE(10)
        HASTY_NS_CASE(11)
        HASTY_NS_CASE(12)
        HASTY_NS_CASE(13)
        HASTY_NS_CASE(14)
        HASTY_NS_CASE(15)
        HASTY_NS_CASE(16)
    default: throw HastyError(kUnsupported, "kernel width out of range");
    }
#undef HASTY_NS_CASE
    check_launch(spread ? "spread_kernel" : "interp_kernel");
}

// plans the output-driven spreader (variant 6) serves: complex64 type 1 in 3-D with every
// fine-grid dim a multiple of the core edge and at least three cores long (so the 27
// neighbours of a core are distinct cores)
bool gather_fits(const NufftPlan* p)
{
    if (p->type != 1 || p->dim != 3) return false;
    for (int d = 0; d < 3; ++d)
        if (p->nf[d] % kGC != 0 || p->nf[d] < 3 * kGC) return false;
    return true;
}

// the plan's 3-D spreader (nufft_autotune) or else the library option; variant 6 on a
// plan it does not serve runs variant 3
int spread_variant(const NufftPlan* p)
{
    const int v = p->spread_variant >= 0 ? p->spread_variant : g_spread_variant.load();
    return v == 6 && !gather_fits(p) ? 3 : v;
}

template <typename R> void dispatch_dim(bool spread, const NufftPlan* p, const SpreadArgs<R>& a, int64_t n, size_t smem,
                                        hipStream_t s)
{
    const int v = spread_variant(p);
    if (p->dim == 3)
        dispatch_ns<R, 3>(spread, p->ns, a, n, smem, s, v);
    else if (p->dim == 2)
        dispatch_ns<R, 2>(spread, p->ns, a, n, smem, s, v);
    else
        dispatch_ns<R, 1>(spread, p->ns, a, n, smem, s, v);
}

// Gauss-Legendre nodes/weights on [-1, 1] by Newton's method on P_n.
void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w)
{
    x.resize(n);
    w.resize(n);
    const double pi = 3.141592653589793238462643383279502884;
    for (int i = 0; i < n; ++i) {
        double z = std::cos(pi * (i + 0.75) / (n + 0.5));
        double dp = 1.0;
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = z;
            for (int k = 2; k <= n; ++k) {
                const double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            dp = n * (z * p1 - p0) / (z * z - 1.0);
            const double dz = p1 / dp;
            z -= dz;
            if (std::fabs(dz) < 1e-16) break;
        }
        x[i] = z;
        w[i] = 2.0 / ((1.0 - z * z) * dp * dp);
    }
}

// 1/Phi(k), k = m - n/2, Phi(k) = integral of the ES kernel times cos(2 pi k z / nf).
std::vector<double> es_deconv(int64_t nf, int ns, double beta, int64_t n)
{
    std::vector<double> xg, wg;
    gauss_legendre(200, xg, wg);
    std::vector<double> out(n);
    for (int64_t m = 0; m < n; ++m) {
        const double k = (double)(m - n / 2);
        double s = 0.0;
        for (size_t q = 0; q < xg.size(); ++q) {
            const double z = (xg[q] + 1.0) * (ns / 4.0);
            const double u = 2.0 * z / ns;
            const double phi = std::exp(beta * (std::sqrt(std::max(1.0 - u * u, 0.0)) - 1.0));
            s += wg[q] * (ns / 4.0) * phi * std::cos(kTwoPi * k * z / (double)nf);
        }
        out[m] = 1.0 / (2.0 * s);
    }
    return out;
}

template <typename R> void upload_real(void*& dst, const std::vector<double>& v)
{
    std::vector<R> h(v.begin(), v.end());
    dst = dev_alloc(sizeof(R) * h.size());
    HASTY_HIP_CHECK(hipMemcpy(dst, h.data(), sizeof(R) * h.size(), hipMemcpyHostToDevice));
}

void launch_fill(int* p, int64_t n, int v, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(fill_int_kernel, dim3((unsigned)ceil_div(n, kThreads)), dim3(kThreads), 0, s, p, n, v);
    check_launch("fill_int_kernel");
}

template <typename R> PtsArgs<R> pts_args(const NufftPlan* p)
{
    PtsArgs<R> a;
    for (int d = 0; d < 3; ++d) {
        a.pts[d] = reinterpret_cast<const R*>(p->pts[d]);
        a.nf[d] = p->nf[d];
        a.bs[d] = p->bs[d];
        a.nbins[d] = p->nbins[d];
    }
    a.ord = bin_order(p->nbins, (p->locality & 1) ? kSuper : 1);
    a.M = p->M;
    return a;
}

}  // namespace

void nufft_set_spread_variant(int v) { g_spread_variant = v; }
int nufft_get_spread_variant() { return g_spread_variant; }
void nufft_set_locality(int v) { g_locality = v; }
int nufft_get_locality() { return g_locality; }
void nufft_set_tile(int v) { g_tile = v; }
int nufft_get_tile() { return g_tile; }
void nufft_set_chunk(int v) { g_chunk = v; }
int nufft_get_chunk() { return g_chunk; }

void es_params(double tol, int* ns, double* beta)
{
    int w = (int)std::ceil(-std::log10(tol / 10.0));
    w = std::max(2, std::min(16, w));
    double c = 2.30;
    if (w == 2) c = 2.20;
    else if (w == 3) c = 2.26;
    else if (w == 4) c = 2.38;
    *ns = w;
    *beta = c * w;
}

template <typename R> void zero_fill(cplx<R>* p, int64_t n, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL((zero_kernel<R>), dim3((unsigned)ceil_div(n, kThreads)), dim3(kThreads), 0, s, p, n);
    check_launch("zero_kernel");
}
template void zero_fill<float>(cplx<float>*, int64_t, hipStream_t);
template void zero_fill<double>(cplx<double>*, int64_t, hipStream_t);

NufftPlan* nufft_makeplan(int type, int dim, const int64_t* nmodes, int iflag, int ntransf, double tol, int fp64)
{
    HASTY_REQUIRE(type == 1 || type == 2, "Only Nufft Type 1 and Type 2 is supported");
    HASTY_REQUIRE(dim >= 1 && dim <= 3, "Dimension must be 1,2 or 3");
    HASTY_REQUIRE(ntransf >= 1, "ntransf must be >= 1");
    HASTY_REQUIRE(tol > 0.0, "tol must be positive");
    auto* p = new NufftPlan();
    try {
        p->type = type;
        p->dim = dim;
        p->isign = iflag >= 0 ? 1 : -1;
        p->ntransf = ntransf;
        p->fp64 = fp64 ? 1 : 0;
        p->tol = tol;
        p->device = current_device();
        for (int d = 0; d < dim; ++d) {
            HASTY_REQUIRE(nmodes[d] >= 1, "nmodes must be >= 1");
            p->n[2 - d] = nmodes[d];
        }
        es_params(tol, &p->ns, &p->beta);
        const size_t csz = p->fp64 ? sizeof(double2) : sizeof(float2);
        for (int d = 3 - dim; d < 3; ++d) {
            p->nf[d] = next235even(std::max<int64_t>(2 * p->n[d], 2 * p->ns));
            HASTY_REQUIRE(fft_size_supported((int)p->nf[d]), "fine grid too large");
        }
        // Tile sizes in 3-D: 20^3 cells with the halo (64 KiB of LDS for complex64, 128 KiB
        // for complex128); the nufft_spread_tile option overrides the edge.
        if (dim == 3) {
            const double cells = 8000.0;
            int t = (int)std::floor(std::cbrt(cells) + 1e-9);
            if (g_tile.load() > 0) t = g_tile.load();
            int b = t - p->ns - 1;
            b = std::max(2, b);
            for (int d = 0; d < 3; ++d) p->bs[d] = b;
        } else if (dim == 2) {
            const int b = p->fp64 ? 24 : 32;
            p->bs[1] = p->bs[2] = b;
        } else {
            p->bs[2] = 512;
        }
        for (int d = 0; d < 3; ++d) p->nbins[d] = (d >= 3 - dim) ? ceil_div(p->nf[d], p->bs[d]) : 1;
        const int64_t nfine = p->nf[0] * p->nf[1] * p->nf[2];
        p->fine = dev_alloc(csz * nfine);
        // 3-D plans stage [f0][f1][n2] between the contiguous pass and the strided ones;
        // 1-D and 2-D plans stage [n0][n1][f2]
        const int64_t ntmp = dim == 3 ? p->nf[0] * p->nf[1] * p->n[2] : p->n[0] * p->n[1] * p->nf[2];
        p->tmp = dev_alloc(csz * std::max<int64_t>(1, ntmp));
        for (int d = 3 - dim; d < 3; ++d) {
            auto v = es_deconv(p->nf[d], p->ns, p->beta, p->n[d]);
            if (p->fp64)
                upload_real<double>(p->phiinv[d], v);
            else
                upload_real<float>(p->phiinv[d], v);
            if (p->fp64)
                twiddles<double>((int)p->nf[d]);
            else
                twiddles<float>((int)p->nf[d]);
        }
        p->chunk = g_chunk.load() > 0 ? g_chunk.load() : chunk_cap(p->fp64 != 0);
        const int loc = g_locality.load();
        p->locality = loc <= 3 ? loc : (type == 1 ? 3 : 0);
        const int64_t nb = bin_range(p->nbins, (p->locality & 1) ? kSuper : 1);
        p->counts = (int*)dev_alloc(sizeof(int) * nb);
        p->binstart = (int*)dev_alloc(sizeof(int) * (nb + 1));
        p->nsubbin = (int*)dev_alloc(sizeof(int) * (nb + 1));
    } catch (...) {
        nufft_destroy(p);
        throw;
    }
    return p;
}

void nufft_destroy(NufftPlan* p)
{
    if (!p) return;
    dev_free(p->fine);
    dev_free(p->tmp);
    for (int d = 0; d < 3; ++d) dev_free(p->phiinv[d]);
    dev_free(p->counts);
    dev_free(p->binstart);
    dev_free(p->nsubbin);
    dev_free(p->rank);
    dev_free(p->sortidx);
    dev_free(p->subbin);
    dev_free(p->suboff);
    dev_free(p->g_counts);
    dev_free(p->g_binstart);
    dev_free(p->g_l0);
    dev_free(p->g_frac);
    dev_free(p->g_idx);
    dev_free(p->g_rank);
    dev_free(p->g_pcount);
    dev_free(p->g_pstart);
    dev_free(p->g_pairs);
    dev_free(p->g_prel);
    dev_free(p->g_list);
    delete p;
}

namespace {

template <typename R> void setpts_typed(NufftPlan* p, hipStream_t s)
{
    const int64_t nb = bin_range(p->nbins, (p->locality & 1) ? kSuper : 1);
    PtsArgs<R> a = pts_args<R>(p);
    launch_fill(p->counts, nb, 0, s);
    if (p->M > 0) {
        hipLaunchKernelGGL((bin_count_kernel<R>), dim3((unsigned)ceil_div(p->M, kThreads)), dim3(kThreads), 0, s, a,
                           p->counts, p->rank);
        check_launch("bin_count_kernel");
    }
    hipLaunchKernelGGL(scan_kernel, dim3(1), dim3(1024), 0, s, (const int*)p->counts, p->binstart, nb);
    check_launch("scan_kernel");
    if (p->M > 0) {
        hipLaunchKernelGGL((bin_scatter_kernel<R>), dim3((unsigned)ceil_div(p->M, kThreads)), dim3(kThreads), 0, s, a,
                           (const int*)p->binstart, (const int*)p->rank, p->sortidx);
        check_launch("bin_scatter_kernel");
    }
    // Spreader subproblems: ceil(count / chunk_cap) per tile. The per-tile counts go into
    // `rank` (free again after the scatter) to avoid another buffer.
    int* nsub = p->rank;
    hipLaunchKernelGGL(subcount_kernel, dim3((unsigned)ceil_div(nb, kThreads)), dim3(kThreads), 0, s,
                       (const int*)p->counts, nsub, nb, p->chunk);
    check_launch("subcount_kernel");
    hipLaunchKernelGGL(scan_kernel, dim3(1), dim3(1024), 0, s, (const int*)nsub, p->nsubbin, nb);
    check_launch("scan_kernel");
    int total = 0;
    HASTY_HIP_CHECK(hipMemcpyAsync(&total, p->nsubbin + nb, sizeof(int), hipMemcpyDeviceToHost, s));
    HASTY_HIP_CHECK(hipStreamSynchronize(s));
    p->nsub = total;
    if (total > p->cap_sub) {
        dev_free(p->subbin);
        dev_free(p->suboff);
        p->subbin = nullptr;
        p->suboff = nullptr;
        p->subbin = (int*)dev_alloc(sizeof(int) * total);
        p->suboff = (int*)dev_alloc(sizeof(int) * total);
        p->cap_sub = total;
    }
    if (total > 0) {
        hipLaunchKernelGGL(subfill_kernel, dim3((unsigned)ceil_div(nb, kThreads)), dim3(kThreads), 0, s,
                           (const int*)p->counts, (const int*)p->nsubbin, p->subbin, p->suboff, nb,
                           p->chunk);
        check_launch("subfill_kernel");
    }
}

// padded LDS tiles: complex64 always, complex128 for the 3-D window-row spreader
bool pad_tile(const NufftPlan* p) { return !p->fp64 || (p->dim == 3 && spread_variant(p) != 0); }

template <typename R> SpreadArgs<R> spread_args(const NufftPlan* p)
{
    SpreadArgs<R> a;
    a.pa = pts_args<R>(p);
    a.sortidx = p->sortidx;
    a.binstart = p->binstart;
    a.subbin = p->subbin;
    a.suboff = p->suboff;
    a.fine = reinterpret_cast<cplx<R>*>(p->fine);
    a.beta = (R)p->beta;
    a.cfac = (R)(4.0 / ((double)p->ns * p->ns));
    a.c = nullptr;
    a.cout = nullptr;
    int T[3];
    for (int d = 0; d < 3; ++d) T[d] = (d >= 3 - p->dim) ? p->bs[d] + p->ns + 1 : 1;
    tile_pitches(T, p->ns, (int)sizeof(cplx<R>), pad_tile(p), a.rp, a.pp);
    a.maxsub = p->chunk;
    a.nsub = p->nsub;
    a.xcd = (p->locality & 2) ? 1 : 0;
    return a;
}

template <typename R> const R* phi(const NufftPlan* p, int d) { return reinterpret_cast<const R*>(p->phiinv[d]); }

// Core bins, (point, core) pairs and the two launch lists of the output-driven spreader,
// for the current points. The pair counts come back to the host (one synchronisation)
// to size the pair list, place each (core, neighbour) run, order the cores and split
// the crowded ones into parts.
template <typename R> void gather_build(NufftPlan* p, hipStream_t s)
{
    GatherBins<R> g;
    int nb[3];
    for (int d = 0; d < 3; ++d) {
        g.pts[d] = reinterpret_cast<const R*>(p->pts[d]);
        g.nf[d] = p->nf[d];
        nb[d] = g.nb[d] = (int)(p->nf[d] / kGC);
    }
    g.M = p->M;
    g.ns = p->ns;
    const int64_t ncore = (int64_t)nb[0] * nb[1] * nb[2];
    if (!p->g_counts) {
        p->g_counts = (int*)dev_alloc(sizeof(int) * ncore);
        p->g_binstart = (int*)dev_alloc(sizeof(int) * (ncore + 1));
        p->g_pcount = (int*)dev_alloc(sizeof(int) * 27 * ncore);
        p->g_pstart = (int*)dev_alloc(sizeof(int) * (27 * ncore + 1));
    }
    if (p->M > p->g_cap_M) {
        dev_free(p->g_l0);
        dev_free(p->g_frac);
        dev_free(p->g_idx);
        dev_free(p->g_rank);
        p->g_l0 = nullptr;
        p->g_frac = nullptr;
        p->g_idx = p->g_rank = nullptr;
        p->g_l0 = (int*)dev_alloc(sizeof(int) * 3 * p->M);
        p->g_frac = dev_alloc(sizeof(R) * 3 * p->M);
        p->g_idx = (int*)dev_alloc(sizeof(int) * p->M);
        p->g_rank = (int*)dev_alloc(sizeof(int) * p->M);
        p->g_cap_M = p->M;
    }
    const unsigned pblocks = (unsigned)ceil_div(std::max<int64_t>(p->M, 1), kThreads);
    launch_fill(p->g_counts, ncore, 0, s);
    if (p->M > 0) {
        hipLaunchKernelGGL((gather_count_kernel<R>), dim3(pblocks), dim3(kThreads), 0, s, g, p->g_counts, p->g_rank);
        check_launch("gather_count_kernel");
    }
    hipLaunchKernelGGL(scan_kernel, dim3(1), dim3(1024), 0, s, (const int*)p->g_counts, p->g_binstart, ncore);
    check_launch("scan_kernel");
    if (p->M > 0) {
        hipLaunchKernelGGL((gather_scatter_kernel<R>), dim3(pblocks), dim3(kThreads), 0, s, g,
                           (const int*)p->g_binstart, (const int*)p->g_rank, p->g_l0, (R*)p->g_frac, p->g_idx);
        check_launch("gather_scatter_kernel");
    }
    hipLaunchKernelGGL((gather_pairs_kernel<R, true>), dim3((unsigned)(27 * ncore)), dim3(64), 0, s, g,
                       (const int*)p->g_binstart, (const int*)p->g_l0, p->g_pcount, (const int*)nullptr, (int*)nullptr,
                       (int*)nullptr);
    check_launch("gather_pairs_kernel");
    hipLaunchKernelGGL(scan_kernel, dim3(1), dim3(1024), 0, s, (const int*)p->g_pcount, p->g_pstart, 27 * ncore);
    check_launch("scan_kernel");
    // per-core totals (into the bin counts, no longer needed) to the host
    hipLaunchKernelGGL(core_pairs_kernel, dim3((unsigned)ceil_div(ncore, kThreads)), dim3(kThreads), 0, s,
                       (const int*)p->g_pstart, p->g_counts, ncore);
    check_launch("core_pairs_kernel");
    std::vector<int> cnt((size_t)ncore);
    HASTY_HIP_CHECK(hipMemcpyAsync(cnt.data(), p->g_counts, sizeof(int) * ncore, hipMemcpyDeviceToHost, s));
    HASTY_HIP_CHECK(hipStreamSynchronize(s));
    int64_t npairs = 0;
    for (int v : cnt) npairs += v;
    HASTY_REQUIRE(npairs < (int64_t)1 << 31, "too many points for the output-driven spreader");
    if (npairs > p->g_cap_pairs) {
        dev_free(p->g_pairs);
        dev_free(p->g_prel);
        p->g_pairs = p->g_prel = nullptr;
        p->g_pairs = (int*)dev_alloc(sizeof(int) * npairs);
        p->g_prel = (int*)dev_alloc(sizeof(int) * npairs);
        p->g_cap_pairs = npairs;
    }
    if (npairs > 0) {
        hipLaunchKernelGGL((gather_pairs_kernel<R, false>), dim3((unsigned)(27 * ncore)), dim3(64), 0, s, g,
                           (const int*)p->g_binstart, (const int*)p->g_l0, (int*)nullptr, (const int*)p->g_pstart,
                           p->g_pairs, p->g_prel);
        check_launch("gather_pairs_kernel");
    }
    std::vector<int> order((size_t)ncore);
    for (int64_t i = 0; i < ncore; ++i) order[(size_t)i] = (int)i;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return cnt[x] > cnt[y]; });
    // one launch: first every part of the crowded cores (more pairs than one part holds;
    // their regions are zeroed beforehand and every part adds with atomics), most crowded
    // first, then one part per other core, stored with plain whole-row stores
    std::vector<int> list, crowded;
    list.reserve((size_t)ncore * 2 + 64);
    for (int ci : order) {
        const int parts = (int)std::max<int64_t>(1, ceil_div(cnt[ci], gather_part<R>()));
        if (parts < 2) continue;
        crowded.push_back(ci);
        for (int k = 0; k < parts; ++k) {
            list.push_back(ci);
            list.push_back(k);
        }
    }
    const int64_t natomic = (int64_t)list.size() / 2;
    for (int ci : order) {
        if (cnt[ci] > gather_part<R>()) continue;
        list.push_back(ci);
        list.push_back(0);
    }
    const size_t nlist = list.size();
    list.insert(list.end(), crowded.begin(), crowded.end());
    const int64_t entries = (int64_t)nlist / 2;
    if ((int64_t)list.size() > 2 * p->g_cap_list) {
        dev_free(p->g_list);
        p->g_list = nullptr;
        p->g_cap_list = 0;
        const int64_t cap = ((int64_t)list.size() + 1) / 2;
        p->g_list = (int*)dev_alloc(sizeof(int) * 2 * cap);
        p->g_cap_list = cap;
    }
    HASTY_HIP_CHECK(hipMemcpyAsync(p->g_list, list.data(), sizeof(int) * list.size(), hipMemcpyHostToDevice, s));
    HASTY_HIP_CHECK(hipStreamSynchronize(s));  // the host list goes out of scope
    p->g_first = entries;
    p->g_rest = natomic;
    p->g_nzero = (int64_t)crowded.size();
    p->gather_built = 1;
}

template <typename R> void gather_spread(NufftPlan* p, const cplx<R>* c, cplx<R>* fine, hipStream_t s)
{
    GatherArgs<R> a;
    a.frac = reinterpret_cast<const R*>(p->g_frac);
    a.idx = p->g_idx;
    a.pstart = p->g_pstart;
    a.pairs = p->g_pairs;
    a.prel = p->g_prel;
    a.list = p->g_list;
    a.c = c;
    a.fine = fine;
    a.M = p->M;
    for (int d = 0; d < 3; ++d) {
        a.nb[d] = (int)(p->nf[d] / kGC);
        a.nf[d] = p->nf[d];
    }
    a.beta = (R)p->beta;
    a.cfac = (R)(4.0 / ((double)p->ns * p->ns));
    a.rp = gather_row_pitch<R>(p->ns);
    a.natomic = (int)p->g_rest;
    KernelTimer t("nufft_spread", s);
    if (p->g_nzero > 0) {
        hipLaunchKernelGGL((zero_cores_kernel<R>), dim3((unsigned)p->g_nzero), dim3(kThreads), 0, s, fine,
                           (const int*)(p->g_list + 2 * p->g_first), a.nb[1], a.nb[2], a.nf[1], a.nf[2]);
        check_launch("zero_cores_kernel");
    }
    launch_gather<R>(p->ns, a, p->g_first, s);
}

template <typename R> void execute_type2(NufftPlan* p, cplx<R>* c, const cplx<R>* f, hipStream_t s)
{
    const int64_t n0 = p->n[0], n1 = p->n[1], n2 = p->n[2];
    const int64_t f0 = p->nf[0], f1 = p->nf[1], f2 = p->nf[2];
    cplx<R>* fine = reinterpret_cast<cplx<R>*>(p->fine);
    cplx<R>* tmp = reinterpret_cast<cplx<R>*>(p->tmp);
    const int64_t nvox = n0 * n1 * n2;
    for (int t = 0; t < p->ntransf; ++t) {
        const cplx<R>* ft = f + t * nvox;
        cplx<R>* ct = c + t * p->M;
        if (p->dim == 3) {
            // The strided passes pad the two small stages and the contiguous pass writes
            // the whole fine grid: 256-point lines read 1:2 along dim 2 run at 4+ TB/s,
            // strided ones near 3.
            {  // dim 0: image [n0][n1][n2] -> fine [f0][n1][n2]
                AxisFft<R> a;
                a.in = ft;
                a.ain = LineAddr{n1 * n2, 0, 1, n1 * n2, nullptr};
                a.out = fine;
                a.aout = a.ain;
                a.L = (int)f0;
                a.n_in = (int)n0;
                a.n_out = (int)f0;
                a.in_centered = 1;
                a.pre_scale = phi<R>(p, 0);
                a.scale = (R)p->scale;
                a.sign = p->isign;
                a.nlines = n1 * n2;
                run_axis_fft<R>(a, s);
            }
            {  // dim 1: fine [f0][n1][n2] -> tmp [f0][f1][n2]
                AxisFft<R> a;
                a.in = fine;
                a.ain = LineAddr{n2, n1 * n2, 1, n2, nullptr};
                a.out = tmp;
                a.aout = LineAddr{n2, f1 * n2, 1, n2, nullptr};
                a.L = (int)f1;
                a.n_in = (int)n1;
                a.n_out = (int)f1;
                a.in_centered = 1;
                a.pre_scale = phi<R>(p, 1);
                a.sign = p->isign;
                a.nlines = f0 * n2;
                run_axis_fft<R>(a, s);
            }
            {  // dim 2: tmp rows [f0][f1][n2] -> fine [f0][f1][f2]
                AxisFft<R> a;
                a.in = tmp;
                a.ain = LineAddr{1, n2, 0, 1, nullptr};
                a.out = fine;
                a.aout = LineAddr{1, f2, 0, 1, nullptr};
                a.L = (int)f2;
                a.n_in = (int)n2;
                a.n_out = (int)f2;
                a.in_centered = 1;
                a.pre_scale = phi<R>(p, 2);
                a.sign = p->isign;
                a.nlines = f0 * f1;
                run_axis_fft<R>(a, s);
            }
            if (p->M > 0) {
                SpreadArgs<R> sa = spread_args<R>(p);
                sa.cout = ct;
                dispatch_dim<R>(false, p, sa, p->M, 0, s);
            }
            continue;
        }
        // pass along dim 2 (contiguous): rows (i0, i1) -> [n0][n1][f2]
        {
            AxisFft<R> a;
            a.in = ft;
            a.ain = LineAddr{1, n2, 0, 1, nullptr};
            a.out = p->dim == 1 ? fine : tmp;
            a.aout = LineAddr{1, f2, 0, 1, nullptr};
            a.L = (int)f2;
            a.n_in = (int)n2;
            a.n_out = (int)f2;
            a.in_centered = 1;
            a.pre_scale = phi<R>(p, 2);
            a.scale = (R)p->scale;
            a.sign = p->isign;
            a.nlines = n0 * n1;
            run_axis_fft<R>(a, s);
        }
        if (p->dim >= 2) {  // dim 1: [n0][n1][f2] -> [n0][f1][f2] (first n0 planes of fine)
            AxisFft<R> a;
            a.in = tmp;
            a.ain = LineAddr{f2, n1 * f2, 1, f2, nullptr};
            a.out = fine;
            a.aout = LineAddr{f2, f1 * f2, 1, f2, nullptr};
            a.L = (int)f1;
            a.n_in = (int)n1;
            a.n_out = (int)f1;
            a.in_centered = 1;
            a.pre_scale = phi<R>(p, 1);
            a.sign = p->isign;
            a.nlines = n0 * f2;
            run_axis_fft<R>(a, s);
        }
        if (p->M > 0) {
            SpreadArgs<R> sa = spread_args<R>(p);
            sa.cout = ct;
            dispatch_dim<R>(false, p, sa, p->M, 0, s);
        }
    }
}

template <typename R> void execute_type1(NufftPlan* p, const cplx<R>* c, cplx<R>* f, hipStream_t s)
{
    const int64_t n0 = p->n[0], n1 = p->n[1], n2 = p->n[2];
    const int64_t f0 = p->nf[0], f1 = p->nf[1], f2 = p->nf[2];
    cplx<R>* fine = reinterpret_cast<cplx<R>*>(p->fine);
    cplx<R>* tmp = reinterpret_cast<cplx<R>*>(p->tmp);
    const int64_t nvox = n0 * n1 * n2;
    int T[3];
    for (int d = 0; d < 3; ++d) T[d] = (d >= 3 - p->dim) ? p->bs[d] + p->ns + 1 : 1;
    int rp = 0, pp = 0;
    tile_pitches(T, p->ns, (int)sizeof(cplx<R>), pad_tile(p), rp, pp);
    const size_t smem = sizeof(cplx<R>) * (size_t)T[0] * pp;
    const bool gather = spread_variant(p) == 6;
    HASTY_REQUIRE(gather || smem <= 160 * 1024, "spreading tile exceeds LDS");
    if (gather && !p->gather_built) {
        HASTY_REQUIRE(!stream_capturing(s),
                      "the output-driven spreader bins the points on its first execute; run one before capturing");
        gather_build<R>(p, s);
    }
    for (int t = 0; t < p->ntransf; ++t) {
        const cplx<R>* ct = c + t * p->M;
        cplx<R>* ft = f + t * nvox;
        if (gather) {
            // writes every fine-grid cell: no zero pass
            gather_spread<R>(p, ct, fine, s);
        } else {
            zero_fill<R>(fine, f0 * f1 * f2, s);
            if (p->nsub > 0) {
                SpreadArgs<R> sa = spread_args<R>(p);
                sa.c = ct;
                dispatch_dim<R>(true, p, sa, p->nsub, smem, s);
            }
        }
        if (p->dim == 3) {
            // The contiguous pass reads the whole fine grid and the strided passes crop
            // the two small stages (the reverse of type 2's order).
            {  // dim 2: fine rows [f0][f1][f2] -> tmp [f0][f1][n2]
                AxisFft<R> a;
                a.in = fine;
                a.ain = LineAddr{1, f2, 0, 1, nullptr};
                a.out = tmp;
                a.aout = LineAddr{1, n2, 0, 1, nullptr};
                a.L = (int)f2;
                a.n_in = (int)f2;
                a.n_out = (int)n2;
                a.out_centered = 1;
                a.post_scale = phi<R>(p, 2);
                a.scale = (R)p->scale;
                a.sign = p->isign;
                a.nlines = f0 * f1;
                run_axis_fft<R>(a, s);
            }
            {  // dim 1: tmp [f0][f1][n2] -> fine [f0][n1][n2]
                AxisFft<R> a;
                a.in = tmp;
                a.ain = LineAddr{n2, f1 * n2, 1, n2, nullptr};
                a.out = fine;
                a.aout = LineAddr{n2, n1 * n2, 1, n2, nullptr};
                a.L = (int)f1;
                a.n_in = (int)f1;
                a.n_out = (int)n1;
                a.out_centered = 1;
                a.post_scale = phi<R>(p, 1);
                a.sign = p->isign;
                a.nlines = f0 * n2;
                run_axis_fft<R>(a, s);
            }
            {  // dim 0: fine [f0][n1][n2] -> image [n0][n1][n2]
                AxisFft<R> a;
                a.in = fine;
                a.ain = LineAddr{n1 * n2, 0, 1, n1 * n2, nullptr};
                a.out = ft;
                a.aout = a.ain;
                a.L = (int)f0;
                a.n_in = (int)f0;
                a.n_out = (int)n0;
                a.out_centered = 1;
                a.post_scale = phi<R>(p, 0);
                a.sign = p->isign;
                a.nlines = n1 * n2;
                run_axis_fft<R>(a, s);
            }
            continue;
        }
        if (p->dim >= 2) {  // dim 1: [n0][f1][f2] -> [n0][n1][f2]
            AxisFft<R> a;
            a.in = fine;
            a.ain = LineAddr{f2, f1 * f2, 1, f2, nullptr};
            a.out = tmp;
            a.aout = LineAddr{f2, n1 * f2, 1, f2, nullptr};
            a.L = (int)f1;
            a.n_in = (int)f1;
            a.n_out = (int)n1;
            a.out_centered = 1;
            a.post_scale = phi<R>(p, 1);
            a.sign = p->isign;
            a.nlines = n0 * f2;
            run_axis_fft<R>(a, s);
        }
        {  // dim 2: rows -> image
            AxisFft<R> a;
            a.in = p->dim == 1 ? fine : tmp;
            a.ain = LineAddr{1, f2, 0, 1, nullptr};
            a.out = ft;
            a.aout = LineAddr{1, n2, 0, 1, nullptr};
            a.L = (int)f2;
            a.n_in = (int)f2;
            a.n_out = (int)n2;
            a.out_centered = 1;
            a.post_scale = phi<R>(p, 2);
            a.scale = (R)p->scale;
            a.sign = p->isign;
            a.nlines = n0 * n1;
            run_axis_fft<R>(a, s);
        }
    }
}

}  // namespace

int nufft_tile_edge(const NufftPlan* p) { return p->bs[2] + p->ns + 1; }

void nufft_retile(NufftPlan* p, int tile, int chunk, hipStream_t s)
{
    HASTY_REQUIRE(p != nullptr && p->dim == 3, "tiles are chosen for 3-D plans");
    HASTY_REQUIRE(tile >= 12 && tile <= 28 && chunk >= 32 && chunk <= 8192, "tile or chunk out of range");
    HASTY_HIP_CHECK(hipStreamSynchronize(s));  // the bin arrays may still be in use
    const int b = std::max(2, tile - p->ns - 1);
    for (int d = 0; d < 3; ++d) {
        p->bs[d] = b;
        p->nbins[d] = ceil_div(p->nf[d], b);
    }
    p->chunk = chunk;
    const int64_t nb = bin_range(p->nbins, (p->locality & 1) ? kSuper : 1);
    dev_free(p->counts);
    dev_free(p->binstart);
    dev_free(p->nsubbin);
    p->counts = p->binstart = p->nsubbin = nullptr;
    p->counts = (int*)dev_alloc(sizeof(int) * nb);
    p->binstart = (int*)dev_alloc(sizeof(int) * (nb + 1));
    p->nsubbin = (int*)dev_alloc(sizeof(int) * (nb + 1));
    nufft_setpts(p, p->M, p->pts[2], p->pts[1], p->pts[0], s);
}

void nufft_setpts(NufftPlan* p, int64_t M, const void* x, const void* y, const void* z, hipStream_t s)
{
    HASTY_REQUIRE(p != nullptr, "null plan");
    HASTY_REQUIRE(M >= 0, "M must be >= 0");
    const void* xyz[3] = {x, y, z};
    for (int d = 0; d < 3; ++d) p->pts[d] = nullptr;
    for (int d = 0; d < p->dim; ++d) {
        HASTY_REQUIRE(xyz[d] != nullptr || M == 0, "missing coordinate array");
        p->pts[2 - d] = xyz[d];
    }
    p->M = M;
    p->gather_built = 0;
    const int64_t nb = bin_range(p->nbins, (p->locality & 1) ? kSuper : 1);
    const int64_t need = std::max<int64_t>(M, nb);
    if (need > p->cap_M) {
        dev_free(p->rank);
        dev_free(p->sortidx);
        p->rank = nullptr;
        p->sortidx = nullptr;
        p->rank = (int*)dev_alloc(sizeof(int) * need);
        p->sortidx = (int*)dev_alloc(sizeof(int) * need);
        p->cap_M = need;
    }
    if (p->fp64)
        setpts_typed<double>(p, s);
    else
        setpts_typed<float>(p, s);
    if (spread_variant(p) == 6) {
        if (p->fp64)
            gather_build<double>(p, s);
        else
            gather_build<float>(p, s);
    }
}

bool nufft_gather_fits(const NufftPlan* p) { return gather_fits(p); }

void nufft_execute(NufftPlan* p, void* c, void* f, hipStream_t s)
{
    HASTY_REQUIRE(p != nullptr, "null plan");
    // c may be null only when there are no points (an empty tensor's pointer)
    HASTY_REQUIRE((c != nullptr || p->M == 0) && f != nullptr, "null data pointer");
    if (p->fp64) {
        if (p->type == 1)
            execute_type1<double>(p, (const double2*)c, (double2*)f, s);
        else
            execute_type2<double>(p, (double2*)c, (const double2*)f, s);
    } else {
        if (p->type == 1)
            execute_type1<float>(p, (const float2*)c, (float2*)f, s);
        else
            execute_type2<float>(p, (float2*)c, (const float2*)f, s);
    }
}

}  // namespace hasty_hip
