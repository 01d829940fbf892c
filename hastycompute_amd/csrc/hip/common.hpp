// Shared helpers for the gfx950 kernels behind libhasty_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

namespace hasty_hip {

template <typename R> struct cplx_of;
template <> struct cplx_of<float> { using type = float2; };
template <> struct cplx_of<double> { using type = double2; };
template <typename R> using cplx = typename cplx_of<R>::type;

template <typename T> __host__ __device__ __forceinline__ T cmul(T a, T b)
{
    return T{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// conj(a) * b
template <typename T> __host__ __device__ __forceinline__ T cmulc(T a, T b)
{
    return T{a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x};
}
template <typename T> __host__ __device__ __forceinline__ T cadd(T a, T b) { return T{a.x + b.x, a.y + b.y}; }
template <typename T> __host__ __device__ __forceinline__ T csub(T a, T b) { return T{a.x - b.x, a.y - b.y}; }
template <typename T, typename S> __host__ __device__ __forceinline__ T cscale(T a, S s)
{
    return T{a.x * s, a.y * s};
}
template <typename T> __host__ __device__ __forceinline__ T cconj(T a) { return T{a.x, -a.y}; }
// acc + a * b
template <typename T> __host__ __device__ __forceinline__ T cfma(T acc, T a, T b)
{
    return T{acc.x + a.x * b.x - a.y * b.y, acc.y + a.x * b.y + a.y * b.x};
}
// acc + conj(a) * b
template <typename T> __host__ __device__ __forceinline__ T cfma_conj(T acc, T a, T b)
{
    return T{acc.x + a.x * b.x + a.y * b.y, acc.y + a.x * b.y - a.y * b.x};
}

// Error type thrown inside the library; every extern "C" entry point converts it to a
// nonzero return code and a message readable through hasty_last_error().
struct HastyError : std::runtime_error {
    int code;
    HastyError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

enum ErrorCode : int {
    kOk = 0,
    kInvalidArgument = 1,
    kHipError = 2,
    kNotPositiveDefinite = 3,
    kUnsupported = 4,
    kOutOfMemory = 5,
};

#define HASTY_HIP_CHECK(expr)                                                                  \
    do {                                                                                       \
        hipError_t _e = (expr);                                                                \
        if (_e != hipSuccess)                                                                  \
            throw ::hasty_hip::HastyError(                                                     \
                _e == hipErrorOutOfMemory ? ::hasty_hip::kOutOfMemory : ::hasty_hip::kHipError, \
                std::string(#expr) + ": " + hipGetErrorString(_e));                           \
        } while (0)

#define HASTY_REQUIRE(cond, msg)                                                               \
    do {                                                                                       \
        if (!(cond)) throw ::hasty_hip::HastyError(::hasty_hip::kInvalidArgument, (msg));      \
    } while (0)

// Records the message for hasty_last_error() (thread-local).
void set_last_error(const std::string& msg);

// Kernel launch check (launch configuration errors surface here).
inline void check_launch(const char* what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw HastyError(kHipError, std::string(what) + ": " + hipGetErrorString(e));
}

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Raise a kernel's dynamic-LDS limit before a launch that needs more than 64 KiB
// (gfx950 has 160 KiB of LDS per CU).
template <typename K> inline void allow_lds(K kernel, size_t bytes)
{
    HASTY_REQUIRE(bytes <= 160 * 1024, "kernel needs more LDS than a gfx950 CU has");
    if (bytes > 64 * 1024)
        HASTY_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
}

// Device-memory helpers (synchronous allocation; never called inside hot loops).
void* dev_alloc(size_t bytes);
void dev_free(void* p);
int current_device();

// Per-kernel timing with HIP events recorded on the launch stream around each launch
// (off by default; hasty_timing_enable). Bracket a launch with a KernelTimer in scope.
bool timing_on();
hipEvent_t timing_begin(hipStream_t s);
void timing_end(const char* name, hipEvent_t start, hipStream_t s);  // name: a string literal
void timing_enable(bool on);                                        // also clears the statistics
int timing_entries();
bool timing_read(int index, std::string* name, double* total_ms, int64_t* count);

// True while `s` is being captured into a graph (its launches are recorded, not run).
inline bool stream_capturing(hipStream_t s)
{
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    return s != nullptr && hipStreamIsCapturing(s, &st) == hipSuccess && st == hipStreamCaptureStatusActive;
}

struct KernelTimer {
    KernelTimer(const char* n, hipStream_t st)
        : name(n), s(st), ev(timing_on() && !stream_capturing(st) ? timing_begin(st) : nullptr)
    {
    }
    ~KernelTimer()
    {
        if (ev) timing_end(name, ev, s);
    }
    KernelTimer(const KernelTimer&) = delete;
    KernelTimer& operator=(const KernelTimer&) = delete;
    const char* name;
    hipStream_t s;
    hipEvent_t ev;
};

}  // namespace hasty_hip
