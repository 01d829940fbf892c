// Autotuner of launch shapes. Each candidate runs on the caller's stream between HIP
// events, on synthetic data of the requested shape (the NUFFT: on the plan's own
// points); the fastest is kept, per shape, for the rest of the process. The search is a
// fixed, coordinate-wise order starting from the defaults, and a candidate replaces the
// incumbent only when it is at least 2 % faster, so near-ties keep the earlier choice.
#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "nufft.hpp"
#include "toeplitz.hpp"
#include "tune.hpp"

namespace hasty_hip {

namespace {

constexpr double kGain = 0.98;  // a challenger must take <= 98 % of the incumbent's time

// x[i] = lo + (hi - lo) u_i with u_i a hashed uniform in [0, 1) (deterministic)
__global__ void fill_uniform_kernel(float* x, int64_t n, uint32_t seed, float lo, float hi)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        uint32_t h = (uint32_t)i * 0x9E3779B1u ^ seed;
        h ^= h >> 16;
        h *= 0x85EBCA6Bu;
        h ^= h >> 13;
        h *= 0xC2B2AE35u;
        h ^= h >> 16;
        x[i] = lo + (hi - lo) * (float)(h >> 8) * (1.0f / 16777216.0f);
    }
}

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    explicit DevBuf(size_t b) : p(dev_alloc(b)), bytes(b) {}
    ~DevBuf() { dev_free(p); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
};

struct Events {
    hipEvent_t a = nullptr, b = nullptr;
    Events()
    {
        HASTY_HIP_CHECK(hipEventCreate(&a));
        HASTY_HIP_CHECK(hipEventCreate(&b));
    }
    ~Events()
    {
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
    }
};

// Best of `reps` timed runs after one untimed run, in ms.
template <typename Fn> double time_best(Fn&& fn, hipStream_t s, int reps)
{
    HASTY_REQUIRE(!stream_capturing(s), "the autotuner cannot run while the stream is being captured");
    fn();
    Events ev;
    double best = 1e300;
    for (int r = 0; r < reps; ++r) {
        HASTY_HIP_CHECK(hipEventRecord(ev.a, s));
        fn();
        HASTY_HIP_CHECK(hipEventRecord(ev.b, s));
        HASTY_HIP_CHECK(hipEventSynchronize(ev.b));
        float ms = 0.f;
        HASTY_HIP_CHECK(hipEventElapsedTime(&ms, ev.a, ev.b));
        best = std::min(best, (double)ms);
    }
    return best;
}

using SenseKey = std::tuple<int64_t, int64_t, int64_t, int, int, int>;
std::map<SenseKey, SenseChoice> g_sense;
std::mutex g_mu;

}  // namespace

void fill_uniform(float* x, int64_t n, uint32_t seed, float lo, float hi, hipStream_t s)
{
    const int64_t blocks = std::min<int64_t>(ceil_div(n, 256), 8192);
    hipLaunchKernelGGL(fill_uniform_kernel, dim3((unsigned)std::max<int64_t>(blocks, 1)), dim3(256), 0, s, x, n, seed,
                       lo, hi);
    check_launch("fill_uniform_kernel");
}

double time_best_ms(const std::function<void()>& fn, hipStream_t s, int reps) { return time_best(fn, s, reps); }

SenseChoice autotune_sense_normal(const int64_t* n3, int ncoil, int nproblems, bool allow_fused, hipStream_t s)
{
    HASTY_REQUIRE(ncoil >= 1 && nproblems >= 1, "need at least one coil and one problem");
    const int Pt = std::min(nproblems, 8);
    int layout = sense_kernel_layout(n3, allow_fused);
    const int fused = layout != 0 ? 1 : 0;
    std::lock_guard<std::mutex> lock(g_mu);
    const SenseKey key{n3[0], n3[1], n3[2], ncoil, Pt, fused};
    auto it = g_sense.find(key);
    if (it != g_sense.end()) {
        if (fused) {
            sense_fused_set_xy_variant(it->second.xy_variant);
            sense_fused_set_zconv_batch(it->second.zconv_batch);
            sense_fused_set_zconv_variant(it->second.zconv_variant);
        }
        return it->second;
    }
    const int64_t nvox = n3[0] * n3[1] * n3[2];
    const int64_t kvox = 8 * nvox;
    DevBuf smaps(sizeof(float2) * nvox * ncoil), x(sizeof(float2) * nvox * Pt), out(sizeof(float2) * nvox * Pt);
    DevBuf knat(sizeof(float) * kvox), kprep(sizeof(float) * kvox), coils(sizeof(int) * ncoil);
    fill_uniform((float*)smaps.p, 2 * nvox * ncoil, 1u, -1.f, 1.f, s);
    fill_uniform((float*)x.p, 2 * nvox * Pt, 2u, -1.f, 1.f, s);
    fill_uniform((float*)knat.p, kvox, 3u, 0.5f, 1.5f, s);
    std::vector<int> cl(ncoil);
    for (int c = 0; c < ncoil; ++c) cl[c] = c;
    HASTY_HIP_CHECK(hipMemcpyAsync(coils.p, cl.data(), sizeof(int) * ncoil, hipMemcpyHostToDevice, s));
    HASTY_HIP_CHECK(hipStreamSynchronize(s));
    std::vector<int> kidx(Pt, 0);
    size_t free_b = 0, total_b = 0;
    HASTY_HIP_CHECK(hipMemGetInfo(&free_b, &total_b));
    const size_t ws_cap = free_b / 2;
    DevBuf* ws = nullptr;
    auto run_with = [&](int cb, int pb) -> double {
        const size_t need = sense_normal_workspace_bytes(n3, cb, pb);
        if (need > ws_cap) return 1e300;
        if (!ws || ws->bytes < need) {
            delete ws;
            ws = nullptr;
            ws = new DevBuf(need);
        }
        auto fn = [&] {
            sense_normal_toeplitz((const float2*)x.p, (float2*)out.p, (const float2*)smaps.p, (const int*)coils.p,
                                  ncoil, (const float*)kprep.p, layout, kidx.data(), nullptr, Pt, n3, ws->p, need, cb,
                                  s);
        };
        return time_best(fn, s, 3);
    };
    SenseChoice best;
    try {
        const int cb0 = std::min(ncoil, 16);
        best.coil_batch = cb0;
        best.problem_batch = fused ? std::min(4, Pt) : 1;
        best.xy_variant = sense_fused_get_xy_variant();
        best.zconv_batch = sense_fused_get_zconv_batch();
        best.zconv_variant = sense_fused_get_zconv_variant();
        double tbest = 1e300;
        if (fused) {
            // at N = 160 every variant runs the coset kernels: one timing is enough
            std::vector<int> variants = n3[1] == 128   ? std::vector<int>{2, 1, 3, 0}
                                        : n3[1] == 160 ? std::vector<int>{best.xy_variant}
                                                       : std::vector<int>{1, 3, 0};
            for (int v : variants) {
                sense_fused_set_xy_variant(v);
                layout = sense_kernel_layout(n3, allow_fused);
                sense_prepare_kernels((const float*)knat.p, (float*)kprep.p, 1, n3, layout, s);
                const double t = run_with(best.coil_batch, best.problem_batch);
                if (t < kGain * tbest) {
                    tbest = t;
                    best.xy_variant = v;
                }
            }
            sense_fused_set_xy_variant(best.xy_variant);
            layout = sense_kernel_layout(n3, allow_fused);
            sense_prepare_kernels((const float*)knat.p, (float*)kprep.p, 1, n3, layout, s);
            for (int pb : {1, 2, 4, 8}) {
                if (pb > Pt || pb == best.problem_batch) continue;
                const double t = run_with(best.coil_batch, pb);
                if (t < kGain * tbest) {
                    tbest = t;
                    best.problem_batch = pb;
                }
            }
            if (2 * n3[0] == 256) {
                // the z-convolution kernel: the two 16-column ones (8 columns halve the
                // row segments a wave reads and run at half speed)
                for (int zv : {2, 0}) {
                    if (zv == best.zconv_variant) continue;
                    sense_fused_set_zconv_variant(zv);
                    const double t = run_with(best.coil_batch, best.problem_batch);
                    if (t < kGain * tbest) {
                        tbest = t;
                        best.zconv_variant = zv;
                    }
                }
                sense_fused_set_zconv_variant(best.zconv_variant);
                // problems per z-convolution launch: values at or above the problem batch
                // all mean one launch per chunk, so each effective value is timed once
                std::vector<int> seen = {std::min(best.zconv_batch, best.problem_batch)};
                for (int z : {16, 8, 4, 2, 1}) {
                    const int eff = std::min(z, best.problem_batch);
                    if (std::find(seen.begin(), seen.end(), eff) != seen.end()) continue;
                    seen.push_back(eff);
                    sense_fused_set_zconv_batch(z);
                    const double t = run_with(best.coil_batch, best.problem_batch);
                    if (t < kGain * tbest) {
                        tbest = t;
                        best.zconv_batch = z;
                    }
                }
                sense_fused_set_zconv_batch(best.zconv_batch);
            }
        } else {
            tbest = run_with(best.coil_batch, 1);
        }
        std::vector<int> cbs = {ncoil, (ncoil + 1) / 2, std::min(ncoil, 8), std::min(ncoil, 4)};
        for (int cb : cbs) {
            if (cb < 1 || cb == best.coil_batch) continue;
            const double t = run_with(cb, best.problem_batch);
            if (t < kGain * tbest) {
                tbest = t;
                best.coil_batch = cb;
            }
        }
        best.ms_per_problem = tbest / Pt;
    } catch (...) {
        delete ws;
        throw;
    }
    delete ws;
    g_sense[key] = best;
    return best;
}

int nufft_autotune(NufftPlan* p, void* c, void* f, hipStream_t s, double* ms)
{
    HASTY_REQUIRE(p != nullptr && c != nullptr && f != nullptr, "null argument");
    if (p->type != 1 || p->dim != 3) {
        if (ms) *ms = 0.0;
        return p->spread_variant >= 0 ? p->spread_variant : nufft_get_spread_variant();
    }
    const int prev = p->spread_variant;
    int best = nufft_get_spread_variant();
    double tbest = 1e300;
    try {
        std::vector<int> order = {best};
        for (int v : {6, 3, 5, 4, 1, 2, 0})
            if (v != best && (v != 6 || nufft_gather_fits(p))) order.push_back(v);
        for (int v : order) {
            p->spread_variant = v;
            const double t = time_best([&] { nufft_execute(p, c, f, s); }, s, 3);
            if (t < kGain * tbest) {
                tbest = t;
                best = v;
            }
        }
    } catch (...) {
        p->spread_variant = prev;
        throw;
    }
    p->spread_variant = best;
    // spreading tile (edge with halo) and points per workgroup of the tile spreaders,
    // complex64 plans: the points are re-binned for each candidate; the current shape is
    // timed first (the output-driven spreader has fixed cores)
    if (!p->fp64 && best != 6) {
        const int t0 = nufft_tile_edge(p), c0 = p->chunk;
        int bt = t0, bc = c0;
        const std::pair<int, int> cands[] = {{20, 128}, {16, 256}, {20, 256}, {16, 128}, {20, 64}};
        for (const auto& tc : cands) {
            if (tc.first == t0 && tc.second == c0) continue;
            nufft_retile(p, tc.first, tc.second, s);
            const double t = time_best([&] { nufft_execute(p, c, f, s); }, s, 3);
            if (t < kGain * tbest) {
                tbest = t;
                bt = tc.first;
                bc = tc.second;
            }
        }
        if (nufft_tile_edge(p) != bt || p->chunk != bc) nufft_retile(p, bt, bc, s);
    }
    if (ms) *ms = tbest;
    return best;
}

}  // namespace hasty_hip
