// Streaming-bandwidth probe: the HBM ceiling for a given mix of read and written
// streams, measured with the same kind of accesses the hot kernels make (16 bytes per
// lane, optionally non-temporal). A measurement hook for calibrating roofline figures;
// no counterpart in the reference.
#include <algorithm>

#include "common.hpp"
#include "tune.hpp"

namespace hasty_hip {

namespace {

constexpr int kProbeThreads = 256;
constexpr int kMaxStreams = 4;
typedef float f32x4 __attribute__((ext_vector_type(4)));

struct ProbeArgs {
    const f32x4* src[kMaxStreams];
    f32x4* dst[kMaxStreams];
    int64_t n;  // 16-byte elements per stream
    f32x4* sink;
};

template <int NR, int NW, bool NT> __global__ void __launch_bounds__(kProbeThreads) stream_probe_kernel(ProbeArgs a)
{
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    const int64_t stride = (int64_t)gridDim.x * kProbeThreads;
    for (int64_t i = (int64_t)blockIdx.x * kProbeThreads + threadIdx.x; i < a.n; i += stride) {
        f32x4 v = {1.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const f32x4 u = NT ? __builtin_nontemporal_load(a.src[r] + i) : a.src[r][i];
            v += u;
        }
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            f32x4 o = v;
            o.x += (float)w;
            if (NT)
                __builtin_nontemporal_store(o, a.dst[w] + i);
            else
                a.dst[w][i] = o;
        }
        if (NW == 0) acc += v;
    }
    // keeps the loads of a read-only mix alive; the condition never holds for finite data
    if (NW == 0 && acc.x == -1.2345e30f) a.sink[0] = acc;
}

template <int NR, int NW> void launch_probe(const ProbeArgs& a, bool nt, hipStream_t s)
{
    const int64_t blocks = std::min<int64_t>(ceil_div(a.n, kProbeThreads), 16384);
    if (nt)
        hipLaunchKernelGGL((stream_probe_kernel<NR, NW, true>), dim3((unsigned)blocks), dim3(kProbeThreads), 0, s, a);
    else
        hipLaunchKernelGGL((stream_probe_kernel<NR, NW, false>), dim3((unsigned)blocks), dim3(kProbeThreads), 0, s, a);
    check_launch("stream_probe_kernel");
}

}  // namespace

void stream_probe(int nread, int nwrite, bool nt, const void* const* src, void* const* dst, int64_t bytes,
                  hipStream_t s)
{
    HASTY_REQUIRE(nread >= 0 && nread <= kMaxStreams && nwrite >= 0 && nwrite <= kMaxStreams && nread + nwrite > 0,
                  "stream probe: 0..4 read and 0..4 written streams, at least one");
    HASTY_REQUIRE(bytes > 0 && bytes % 16 == 0, "stream probe: bytes per stream must be a positive multiple of 16");
    ProbeArgs a{};
    for (int r = 0; r < nread; ++r) {
        HASTY_REQUIRE(src && src[r], "stream probe: null source");
        a.src[r] = static_cast<const f32x4*>(src[r]);
    }
    for (int w = 0; w < nwrite; ++w) {
        HASTY_REQUIRE(dst && dst[w], "stream probe: null destination");
        a.dst[w] = static_cast<f32x4*>(dst[w]);
    }
    HASTY_REQUIRE(nwrite > 0 || (dst && dst[0]), "stream probe: a read-only mix needs dst[0] as a 16-byte sink");
    a.sink = nwrite > 0 ? a.dst[0] : static_cast<f32x4*>(dst[0]);
    a.n = bytes / 16;
    KernelTimer t("stream_probe", s);
#define HASTY_PROBE_CASE(R, W) \
    if (nread == R && nwrite == W) return launch_probe<R, W>(a, nt, s);
    HASTY_PROBE_CASE(1, 0)
    HASTY_PROBE_CASE(2, 0)
    HASTY_PROBE_CASE(0, 1)
    HASTY_PROBE_CASE(0, 2)
    HASTY_PROBE_CASE(1, 1)
    HASTY_PROBE_CASE(2, 1)
    HASTY_PROBE_CASE(1, 2)
    HASTY_PROBE_CASE(1, 4)
    HASTY_PROBE_CASE(4, 1)
#undef HASTY_PROBE_CASE
    throw HastyError(kUnsupported, "stream probe: mixes are 1:0, 2:0, 0:1, 0:2, 1:1, 2:1, 1:2, 1:4 and 4:1");
}

}  // namespace hasty_hip
