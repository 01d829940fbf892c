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
//   2. eig:   cyclic complex Jacobi in FP32 in LDS, parallel (round-robin) ordering,
//             then W = V diag(g) V^H;
//   3. apply: MFMA A <- A W in place, 16-row slabs staged through LDS.
// Blocks are read and written directly in the image; no gather/scatter copies.
#include <algorithm>
#include <vector>

#include "llr.hpp"

namespace hasty_hip {

namespace {

typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int kThreads = 256;
constexpr int kMaxTilesPerWave = 9;  // F <= 128 -> 8x8 tiles -> 36 upper tiles / 4 waves

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

// Upper-triangle tile index t -> (ti, tj), ti <= tj, row-major over ti.
__device__ __forceinline__ void tile_ij(int t, int nt, int& ti, int& tj)
{
    int i = 0;
    int rem = t;
    while (rem >= nt - i) {
        rem -= nt - i;
        ++i;
    }
    ti = i;
    tj = i + rem;
}

__global__ void __launch_bounds__(kThreads) gram_kernel(const float2* __restrict__ data, LlrGeom g, const int3* corners,
                                                       int Fp, int64_t rows_per_split, float2* __restrict__ gpart)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float* sAr = reinterpret_cast<float*>(smem_raw);
    float* sAi = sAr + 16 * Fp;
    const int blk = blockIdx.x;
    const int split = blockIdx.y;
    const int nsplit = gridDim.y;
    const int3 c = corners[blk];
    const int64_t R = g.E * g.bx * g.by * g.bz;
    const int64_t r_begin = split * rows_per_split;
    const int64_t r_end = min(R, r_begin + rows_per_split);
    const int F = (int)g.F;
    const int nt = Fp / 16;
    const int ntile = nt * (nt + 1) / 2;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;

    for (int i = threadIdx.x; i < 16 * Fp; i += kThreads) {
        sAr[i] = 0.f;
        sAi[i] = 0.f;
    }

    floatx4 accR[kMaxTilesPerWave], accI[kMaxTilesPerWave];
    int tI[kMaxTilesPerWave], tJ[kMaxTilesPerWave];
#pragma unroll
    for (int k = 0; k < kMaxTilesPerWave; ++k) {
        accR[k] = floatx4{0.f, 0.f, 0.f, 0.f};
        accI[k] = floatx4{0.f, 0.f, 0.f, 0.f};
        const int t = wave + 4 * k;
        if (t < ntile)
            tile_ij(t, nt, tI[k], tJ[k]);
        else
            tI[k] = tJ[k] = -1;
    }
    __syncthreads();

    for (int64_t r0 = r_begin; r0 < r_end; r0 += 16) {
        // stage 16 rows x F columns; 16 consecutive threads read 16 consecutive rows
        for (int e = threadIdx.x; e < 16 * F; e += kThreads) {
            const int kr = e & 15;
            const int f = e >> 4;
            const int64_t r = r0 + kr;
            float2 v = float2{0.f, 0.f};
            if (r < r_end) v = data[row_addr(g, c, r) + (int64_t)f * g.fstride];
            sAr[kr * Fp + f] = v.x;
            sAi[kr * Fp + f] = v.y;
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) {
            const int krow = ks * 4 + (lane >> 4);
            const int col = lane & 15;
#pragma unroll
            for (int k = 0; k < kMaxTilesPerWave; ++k) {
                if (tI[k] < 0) continue;
                const float ar = sAr[krow * Fp + tI[k] * 16 + col];
                const float ai = sAi[krow * Fp + tI[k] * 16 + col];
                const float br = sAr[krow * Fp + tJ[k] * 16 + col];
                const float bi = sAi[krow * Fp + tJ[k] * 16 + col];
                accR[k] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar, br, accR[k], 0, 0, 0);
                accR[k] = __builtin_amdgcn_mfma_f32_16x16x4f32(ai, bi, accR[k], 0, 0, 0);
                accI[k] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar, bi, accI[k], 0, 0, 0);
                accI[k] = __builtin_amdgcn_mfma_f32_16x16x4f32(-ai, br, accI[k], 0, 0, 0);
            }
        }
        __syncthreads();
    }

    float2* G = gpart + ((int64_t)blk * nsplit + split) * Fp * Fp;
#pragma unroll
    for (int k = 0; k < kMaxTilesPerWave; ++k) {
        if (tI[k] < 0) continue;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int i = tI[k] * 16 + 4 * (lane >> 4) + v;
            const int j = tJ[k] * 16 + (lane & 15);
            G[(int64_t)i * Fp + j] = float2{accR[k][v], accI[k][v]};
            if (tI[k] != tJ[k]) G[(int64_t)j * Fp + i] = float2{accR[k][v], -accI[k][v]};
        }
    }
}

// Jacobi eigensolver + W = V diag(g) V^H, one workgroup per block.
// LDS: G[n][n] and V[n][n] complex float (n = F rounded up to even).
__global__ void __launch_bounds__(kThreads) eig_w_kernel(const float2* __restrict__ gpart, int nsplit, int F, int Fp,
                                                        float thresh, int soft, int rank, int max_sweeps,
                                                        float2* __restrict__ Wout, float* __restrict__ eigout)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int n = F + (F & 1);
    float2* G = reinterpret_cast<float2*>(smem_raw);
    float2* V = G + n * n;
    float* rc = reinterpret_cast<float*>(V + n * n);  // per pair: c, s, cos(phi), sin(phi)
    int* pp = reinterpret_cast<int*>(rc + 4 * (n / 2));
    int* pq = pp + n / 2;
    __shared__ double red[kThreads];
    const int blk = blockIdx.x;
    const int tid = threadIdx.x;

    // Load G (sum of partial Grams in split order), Hermitian-symmetrised; V = I.
    const float2* Gb = gpart + (int64_t)blk * nsplit * Fp * Fp;
    for (int e = tid; e < n * n; e += kThreads) {
        const int i = e / n, j = e % n;
        float2 v = float2{0.f, 0.f};
        if (i < F && j < F) {
            const int ii = min(i, j), jj = max(i, j);
            for (int s = 0; s < nsplit; ++s) {
                const float2 t = Gb[(int64_t)s * Fp * Fp + (int64_t)ii * Fp + jj];
                v.x += t.x;
                v.y += t.y;
            }
            if (i > j) v.y = -v.y;
            if (i == j) v.y = 0.f;
        }
        G[e] = v;
        V[e] = float2{i == j ? 1.f : 0.f, 0.f};
    }
    __syncthreads();

    // ||G||_F^2 for the stopping rule
    double nrm = 0.0;
    for (int e = tid; e < n * n; e += kThreads) nrm += (double)G[e].x * G[e].x + (double)G[e].y * G[e].y;
    red[tid] = nrm;
    __syncthreads();
    for (int off = kThreads / 2; off > 0; off >>= 1) {
        if (tid < off) red[tid] += red[tid + off];
        __syncthreads();
    }
    const double total = red[0];
    __syncthreads();

    const int half = n / 2;
    for (int sweep = 0; sweep < max_sweeps; ++sweep) {
        double off = 0.0;
        for (int e = tid; e < n * n; e += kThreads) {
            const int i = e / n, j = e % n;
            if (i != j) off += (double)G[e].x * G[e].x + (double)G[e].y * G[e].y;
        }
        red[tid] = off;
        __syncthreads();
        for (int o = kThreads / 2; o > 0; o >>= 1) {
            if (tid < o) red[tid] += red[tid + o];
            __syncthreads();
        }
        const double offn = red[0];
        __syncthreads();
        // off-diagonal mass below (1e-5 |G|_F)^2: at FP32 rounding (n eps |G| ~ 6e-6 |G|
        // for n = 100) further sweeps stop paying
        if (offn <= 1e-10 * total || total == 0.0) break;

        for (int round = 0; round < n - 1; ++round) {
            // round-robin pairing: position 0 fixed, positions 1..n-1 rotate
            if (tid < half) {
                const int a = tid == 0 ? 0 : 1 + (tid - 1 + round) % (n - 1);
                const int b = 1 + (n - 1 - tid - 1 + round) % (n - 1);
                const int p = min(a, b), q = max(a, b);
                pp[tid] = p;
                pq[tid] = q;
                const float ap = G[p * n + p].x, dq = G[q * n + q].x;
                const float2 bpq = G[p * n + q];
                const float ab = hypotf(bpq.x, bpq.y);
                float c = 1.f, s = 0.f, cf = 1.f, sf = 0.f;
                if (ab > 1e-30f && ab > 1e-9f * (fabsf(ap) + fabsf(dq))) {
                    cf = bpq.x / ab;
                    sf = bpq.y / ab;
                    const float tau = (dq - ap) / (2.f * ab);
                    const float t = (tau >= 0.f ? 1.f : -1.f) / (fabsf(tau) + sqrtf(1.f + tau * tau));
                    c = 1.f / sqrtf(1.f + t * t);
                    s = t * c;
                }
                rc[4 * tid + 0] = c;
                rc[4 * tid + 1] = s;
                rc[4 * tid + 2] = cf;
                rc[4 * tid + 3] = sf;
            }
            __syncthreads();
            // G <- J^H G J, block by block: the rows of pair k and the columns of pair l
            // form a 2x2 block that maps to J_k^H G_kl J_l, read and written by one thread
            // (rows: r_p = c r_p - s e^{i phi} r_q, r_q = s r_p + c e^{i phi} r_q; columns the
            // same with e^{-i phi}). The diagonal blocks end exactly diagonal and real.
            for (int e = tid; e < half * half; e += kThreads) {
                const int k = e / half, l = e % half;
                const int pk = pp[k], qk = pq[k], pl = pp[l], ql = pq[l];
                const float ck = rc[4 * k], sk = rc[4 * k + 1];
                const float2 phk = float2{rc[4 * k + 2], rc[4 * k + 3]};
                const float cl = rc[4 * l], sl = rc[4 * l + 1];
                const float2 phl = float2{rc[4 * l + 2], -rc[4 * l + 3]};
                const float2 a = G[pk * n + pl], bb = G[pk * n + ql], cc = G[qk * n + pl], d = G[qk * n + ql];
                const float2 ea = cmul(phk, cc), eb = cmul(phk, d);
                const float2 a1 = float2{ck * a.x - sk * ea.x, ck * a.y - sk * ea.y};
                const float2 b1 = float2{ck * bb.x - sk * eb.x, ck * bb.y - sk * eb.y};
                const float2 c1 = float2{sk * a.x + ck * ea.x, sk * a.y + ck * ea.y};
                const float2 d1 = float2{sk * bb.x + ck * eb.x, sk * bb.y + ck * eb.y};
                const float2 fb = cmul(phl, b1), fd = cmul(phl, d1);
                float2 na = float2{cl * a1.x - sl * fb.x, cl * a1.y - sl * fb.y};
                float2 nb = float2{sl * a1.x + cl * fb.x, sl * a1.y + cl * fb.y};
                float2 nc = float2{cl * c1.x - sl * fd.x, cl * c1.y - sl * fd.y};
                float2 nd = float2{sl * c1.x + cl * fd.x, sl * c1.y + cl * fd.y};
                if (k == l) {
                    na.y = 0.f;
                    nd.y = 0.f;
                    if (sk != 0.f) nb = nc = float2{0.f, 0.f};
                }
                G[pk * n + pl] = na;
                G[pk * n + ql] = nb;
                G[qk * n + pl] = nc;
                G[qk * n + ql] = nd;
            }
            // V <- V J (the columns of each pair)
            for (int e = tid; e < half * n; e += kThreads) {
                const int k = e % half, i = e / half;
                const int p = pp[k], q = pq[k];
                const float c = rc[4 * k], s = rc[4 * k + 1];
                const float2 phc = float2{rc[4 * k + 2], -rc[4 * k + 3]};
                const float2 vp = V[i * n + p], vq = V[i * n + q];
                const float2 eq = cmul(phc, vq);
                V[i * n + p] = float2{c * vp.x - s * eq.x, c * vp.y - s * eq.y};
                V[i * n + q] = float2{s * vp.x + c * eq.x, s * vp.y + c * eq.y};
            }
            __syncthreads();
        }
    }

    // g_k from eigenvalues (stored in rc[0..n) after the rotation scratch is done)
    float* gk = rc;
    for (int k = tid; k < n; k += kThreads) {
        const float lam = G[k * n + k].x;
        float gv = 0.f;
        if (k < F) {
            if (soft) {
                const float sg = sqrtf(fmaxf(lam, 0.f));
                gv = sg > thresh ? (sg - thresh) / sg : 0.f;
            } else {
                int cnt = 0;
                for (int j = 0; j < F; ++j) {
                    const float lj = G[j * n + j].x;
                    if (lj > lam || (lj == lam && j < k)) ++cnt;
                }
                gv = cnt < rank ? 1.f : 0.f;
            }
            if (eigout) eigout[(int64_t)blk * F + k] = lam;
        }
        gk[k] = gv;
    }
    __syncthreads();

    float2* W = Wout + (int64_t)blk * Fp * Fp;
    for (int e = tid; e < Fp * Fp; e += kThreads) {
        const int i = e / Fp, j = e % Fp;
        float2 acc = float2{0.f, 0.f};
        if (i < F && j < F) {
            for (int k = 0; k < n; ++k) {
                const float g = gk[k];
                if (g == 0.f) continue;
                const float2 vi = V[i * n + k], vj = V[j * n + k];
                // V[i][k] g conj(V[j][k])
                acc.x += g * (vi.x * vj.x + vi.y * vj.y);
                acc.y += g * (vi.y * vj.x - vi.x * vj.y);
            }
        }
        W[e] = acc;
    }
}

// A <- A W for rows [chunk*rows_per_wg, ...) of each block, in place in the image.
__global__ void __launch_bounds__(kThreads) apply_kernel(float2* __restrict__ data, LlrGeom g, const int3* corners, int Fp,
                                                        int64_t rows_per_wg, const float2* __restrict__ Wall)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int Fs = Fp + 1;  // padded row stride of the A slab
    float* sWr = reinterpret_cast<float*>(smem_raw);
    float* sWi = sWr + Fp * Fp;
    float* sAr = sWi + Fp * Fp;
    float* sAi = sAr + 16 * Fs;
    const int blk = blockIdx.x;
    const int3 c = corners[blk];
    const int64_t R = g.E * g.bx * g.by * g.bz;
    const int64_t r_begin = (int64_t)blockIdx.y * rows_per_wg;
    const int64_t r_end = min(R, r_begin + rows_per_wg);
    if (r_begin >= r_end) return;
    const int F = (int)g.F;
    const int nt = Fp / 16;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;

    const float2* W = Wall + (int64_t)blk * Fp * Fp;
    for (int e = threadIdx.x; e < Fp * Fp; e += kThreads) {
        const float2 w = W[e];
        sWr[e] = w.x;
        sWi[e] = w.y;
    }
    for (int e = threadIdx.x; e < 16 * Fs; e += kThreads) {
        sAr[e] = 0.f;
        sAi[e] = 0.f;
    }
    __syncthreads();

    for (int64_t r0 = r_begin; r0 < r_end; r0 += 16) {
        for (int e = threadIdx.x; e < 16 * F; e += kThreads) {
            const int kr = e & 15;
            const int f = e >> 4;
            const int64_t r = r0 + kr;
            float2 v = float2{0.f, 0.f};
            if (r < r_end) v = data[row_addr(g, c, r) + (int64_t)f * g.fstride];
            sAr[kr * Fs + f] = v.x;
            sAi[kr * Fs + f] = v.y;
        }
        __syncthreads();
        floatx4 oR[2], oI[2];
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            oR[k] = floatx4{0.f, 0.f, 0.f, 0.f};
            oI[k] = floatx4{0.f, 0.f, 0.f, 0.f};
        }
        const int m = lane & 15;
        for (int ks = 0; ks < Fp / 4; ++ks) {
            const int kk = ks * 4 + (lane >> 4);
            const float ar = sAr[m * Fs + kk];
            const float ai = sAi[m * Fs + kk];
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                const int tj = wave + 4 * k;
                if (tj >= nt) continue;
                const float wr = sWr[kk * Fp + tj * 16 + m];
                const float wi = sWi[kk * Fp + tj * 16 + m];
                oR[k] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar, wr, oR[k], 0, 0, 0);
                oR[k] = __builtin_amdgcn_mfma_f32_16x16x4f32(-ai, wi, oR[k], 0, 0, 0);
                oI[k] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar, wi, oI[k], 0, 0, 0);
                oI[k] = __builtin_amdgcn_mfma_f32_16x16x4f32(ai, wr, oI[k], 0, 0, 0);
            }
        }
        __syncthreads();  // every wave is done reading the slab
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const int tj = wave + 4 * k;
            if (tj >= nt) continue;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = 4 * (lane >> 4) + v;
                const int col = tj * 16 + (lane & 15);
                sAr[row * Fs + col] = oR[k][v];
                sAi[row * Fs + col] = oI[k][v];
            }
        }
        __syncthreads();
        for (int e = threadIdx.x; e < 16 * F; e += kThreads) {
            const int kr = e & 15;
            const int f = e >> 4;
            const int64_t r = r0 + kr;
            if (r < r_end) data[row_addr(g, c, r) + (int64_t)f * g.fstride] = float2{sAr[kr * Fs + f], sAi[kr * Fs + f]};
        }
        __syncthreads();
    }
}

}  // namespace

int llr_padded(int64_t F) { return (int)(((F + 15) / 16) * 16); }

size_t llr_eig_smem(int64_t F)
{
    const int n = (int)(F + (F & 1));
    return sizeof(float2) * 2 * (size_t)n * n + sizeof(float) * 4 * (n / 2) + sizeof(int) * n;
}

// Workspace: gpart [nblk][nsplit][Fp][Fp] + W [nblk][Fp][Fp] complex float.
size_t llr_workspace_bytes(int64_t nblk, int64_t F, int nsplit)
{
    const int64_t Fp = llr_padded(F);
    return sizeof(float2) * (size_t)nblk * (size_t)(nsplit + 1) * Fp * Fp;
}

void llr_check(int64_t F)
{
    HASTY_REQUIRE(F >= 1 && F <= 128, "LLR supports 1..128 columns (frames)");
    // dynamic LDS of the eigensolver plus its 2 KiB static reduction buffer
    HASTY_REQUIRE(llr_eig_smem(F) + sizeof(double) * kThreads <= 160 * 1024, "LLR eigensolver tile exceeds LDS");
}

// Threshold the blocks at `corners_dev` (all disjoint) in place.
void llr_blocks(float2* data, const LlrGeom& g, const int3* corners_dev, int64_t nblk, float thresh, int soft, int rank,
                void* ws, size_t ws_bytes, int nsplit, float* eig_out, hipStream_t s)
{
    llr_check(g.F);
    if (nblk <= 0) return;
    HASTY_REQUIRE(nblk < 65536, "too many blocks in one LLR launch");
    HASTY_REQUIRE(ws_bytes >= llr_workspace_bytes(nblk, g.F, nsplit), "LLR workspace too small");
    const int Fp = llr_padded(g.F);
    const int64_t R = g.E * g.bx * g.by * g.bz;
    float2* gpart = reinterpret_cast<float2*>(ws);
    float2* W = gpart + (int64_t)nblk * nsplit * Fp * Fp;
    int64_t rps = ceil_div(ceil_div(R, nsplit), 16) * 16;
    {
        const size_t smem = sizeof(float) * 2 * 16 * Fp;
        KernelTimer t("llr_gram", s);
        hipLaunchKernelGGL(gram_kernel, dim3((unsigned)nblk, (unsigned)nsplit), dim3(kThreads), smem, s,
                           (const float2*)data, g, corners_dev, Fp, rps, gpart);
        check_launch("gram_kernel");
    }
    {
        const size_t smem = llr_eig_smem(g.F);
        allow_lds(eig_w_kernel, smem);
        KernelTimer t("llr_eig", s);
        hipLaunchKernelGGL(eig_w_kernel, dim3((unsigned)nblk), dim3(kThreads), smem, s, (const float2*)gpart, nsplit,
                           (int)g.F, Fp, thresh, soft, rank, 15, W, eig_out);
        check_launch("eig_w_kernel");
    }
    {
        const int64_t rows_per_wg = 256;
        const int64_t nchunk = ceil_div(R, rows_per_wg);
        const size_t smem = sizeof(float) * (2 * (size_t)Fp * Fp + 2 * 16 * (size_t)(Fp + 1));
        allow_lds(apply_kernel, smem);
        KernelTimer t("llr_apply", s);
        hipLaunchKernelGGL(apply_kernel, dim3((unsigned)nblk, (unsigned)nchunk), dim3(kThreads), smem, s, data, g,
                           corners_dev, Fp, rows_per_wg, (const float2*)W);
        check_launch("apply_kernel");
    }
}

}  // namespace hasty_hip
