// common.hpp -- runtime plumbing shared by every translation unit of
// libchimera_hip: status codes, the library stream, error text, staging of
// f2py-layout arguments that may live in host or device memory, and a few
// complex-arithmetic helpers on double2 / float2.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace chimera {

enum : int32_t {
    ST_OK = 0,
    ST_EINVAL = -1,   // bad argument (shape, size, null pointer)
    ST_EHIP = -2,     // HIP runtime error
    ST_ENOMEM = -3,   // device allocation failed
    ST_EUNSUP = -4,   // valid in the reference, not supported here (message says which)
};

// Per-thread last-error text (read through chimera_last_error()).
void set_error(const char* fmt, ...) __attribute__((format(printf, 1, 2)));
int32_t fail(int32_t code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
int32_t hip_fail(hipError_t e, const char* what, const char* file, int line);

// Stream of the calling thread's current context (chimera_context_bind).
hipStream_t stream();
void set_stream(hipStream_t s);
// True while the current context records a graph (chimera_capture_begin).
bool capturing();

#define CH_HIP(expr)                                                                  \
    do {                                                                              \
        hipError_t _e = (expr);                                                       \
        if (_e != hipSuccess) return ::chimera::hip_fail(_e, #expr, __FILE__, __LINE__); \
    } while (0)

#define CH_LAUNCH_CHECK(name)                                                             \
    do {                                                                                  \
        hipError_t _e = hipGetLastError();                                                \
        if (_e != hipSuccess) return ::chimera::hip_fail(_e, name, __FILE__, __LINE__);  \
    } while (0)

#define CH_TRY(expr)                         \
    do {                                     \
        int32_t _s = (expr);                 \
        if (_s != ::chimera::ST_OK) return _s; \
    } while (0)

// True when p points into device (or managed) memory that kernels may read.
bool is_device_ptr(const void* p);

// Staging of entry-point arguments.  Pointers to device memory pass through
// untouched; host pointers are copied into temporary device buffers on the
// library stream and, for written arguments, copied back by finish(), which
// then synchronises the stream.  A call whose arguments are all device
// pointers stays asynchronous.
class Stage {
  public:
    Stage() = default;
    Stage(const Stage&) = delete;
    Stage& operator=(const Stage&) = delete;
    ~Stage();

    template <class T>
    T* in(const T* p, size_t n) { return static_cast<T*>(put(const_cast<T*>(p), n * sizeof(T), true, false)); }
    template <class T>
    T* inout(T* p, size_t n) { return static_cast<T*>(put(p, n * sizeof(T), true, true)); }
    template <class T>
    T* out(T* p, size_t n) { return static_cast<T*>(put(p, n * sizeof(T), false, true)); }

    int32_t status() const { return st_; }
    bool staged() const { return any_host_; }
    int32_t finish();

  private:
    void* put(void* p, size_t bytes, bool copy_in, bool copy_out);
    struct Buf { void* host; void* dev; size_t bytes; bool copy_out; };
    std::vector<Buf> bufs_;
    int32_t st_ = ST_OK;
    bool any_host_ = false;
};

// Scratch device memory of the current context, grown on demand and reused
// across calls on its stream (the stream orders every use).  It cannot grow
// while the context captures a step or while graphs captured on it are alive.
void* scratch(int slot, size_t bytes);
// Byte fill of device memory by a kernel on the library stream (no hipMemsetAsync,
// so that a captured graph replays it correctly).
int32_t fill_bytes(void* p, int value, size_t bytes);
constexpr int SCRATCH_SLOTS = 24;

// Device copy of a small host table (job lists), cached by content in the
// current context and freed with it.
const void* device_table(const void* host, size_t bytes);
// Device copy of a constant table that any context may read at any time (FFT
// twiddles), cached by content for the life of the process.
const void* shared_table(const void* host, size_t bytes);

inline unsigned blocks_for(int64_t n, int threads) { return (unsigned)((n + threads - 1) / threads); }

// Kernel families timed by the optional in-library profiler.
enum KernelFamily : int {
    KF_PUSH_COORDS = 0,  // leap-frog position push
    KF_DEP_CURR = 1,     // current deposition (with the fused position push and charge deposition)
    KF_DEP_DENS = 2,     // charge deposition
    KF_GATHER_PUSH = 3,  // field gather + device + Boris (and the drop-in gather)
    KF_DHT = 4,          // Hankel-transform GEMMs of the FB transforms
    KF_FFT = 5,          // x FFTs of the FB transforms and the absorbing layer
    KF_FBOPS = 6,        // FB-space operator GEMMs (rot, grad, div, grad-div)
    KF_PSATD = 7,        // PSATD and elementwise field updates
    KF_SORT = 8,         // compaction, scans, sorting
    KF_OTHER = 9,
    KF_COUNT = 10
};

// While profiling is on, a KTimer brackets the launches in its scope with
// HIP events recorded on the library stream.
struct KTimer {
    explicit KTimer(int fam);
    ~KTimer();
    KTimer(const KTimer&) = delete;
    KTimer& operator=(const KTimer&) = delete;
    int fam;
    bool on;
    hipEvent_t a;
};

}  // namespace chimera

// ---------------------------------------------------------------- complex helpers
__host__ __device__ inline double2 cmk(double re, double im) { return make_double2(re, im); }
__host__ __device__ inline double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__host__ __device__ inline double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__host__ __device__ inline double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ inline double2 crs(double s, double2 a) { return make_double2(s * a.x, s * a.y); }
__host__ __device__ inline double2 cmuli(double2 a) { return make_double2(-a.y, a.x); }
__host__ __device__ inline double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
__host__ __device__ inline double2 cneg(double2 a) { return make_double2(-a.x, -a.y); }

// ---------------------------------------------------------------- global-memory access helpers
// A pointer read from a table in memory (rather than passed as a kernel
// argument) is not known to point to global memory, and the compiler emits
// FLAT loads for it.  FLAT accesses count against lgkmcnt as well as vmcnt, so
// the wait for an LDS read would also wait for them; these helpers access such
// pointers as global memory explicitly.
#define CH_GLOBAL __attribute__((address_space(1)))
typedef double ch_vd2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ double2 gld2(const double2* p, int64_t i) {
    const ch_vd2 v = ((const CH_GLOBAL ch_vd2*)p)[i];
    return make_double2(v.x, v.y);
}
__device__ __forceinline__ double gld1(const double* p, int64_t i) { return ((const CH_GLOBAL double*)p)[i]; }
__device__ __forceinline__ void gst2(double2* p, int64_t i, double2 v) {
    ch_vd2 w;
    w.x = v.x;
    w.y = v.y;
    ((CH_GLOBAL ch_vd2*)p)[i] = w;
}

// ---------------------------------------------------------------- single-precision complex (mixed precision)
__host__ __device__ inline float2 cmkf(float re, float im) { return make_float2(re, im); }
__host__ __device__ inline float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__host__ __device__ inline float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__host__ __device__ inline float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ inline float2 crs(float s, float2 a) { return make_float2(s * a.x, s * a.y); }
__host__ __device__ inline float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
__host__ __device__ inline float2 cneg(float2 a) { return make_float2(-a.x, -a.y); }

// Complex type of a real type: cplx_t<double> = double2, cplx_t<float> = float2.
template <typename R> struct CplxOf;
template <> struct CplxOf<double> { typedef double2 type; };
template <> struct CplxOf<float> { typedef float2 type; };
template <typename R> using cplx_t = typename CplxOf<R>::type;
template <typename R> __host__ __device__ inline cplx_t<R> cmake(R re, R im);
template <> __host__ __device__ inline double2 cmake<double>(double re, double im) { return make_double2(re, im); }
template <> __host__ __device__ inline float2 cmake<float>(float re, float im) { return make_float2(re, im); }
// complex conversion between precisions
__host__ __device__ inline float2 to_f(double2 a) { return make_float2((float)a.x, (float)a.y); }
template <typename R> __host__ __device__ inline cplx_t<R> cconv(double2 a);
template <> __host__ __device__ inline double2 cconv<double>(double2 a) { return a; }
template <> __host__ __device__ inline float2 cconv<float>(double2 a) { return to_f(a); }

typedef float ch_vf2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float2 gld2(const float2* p, int64_t i) {
    const ch_vf2 v = ((const CH_GLOBAL ch_vf2*)p)[i];
    return make_float2(v.x, v.y);
}
__device__ __forceinline__ float gld1(const float* p, int64_t i) { return ((const CH_GLOBAL float*)p)[i]; }
__device__ __forceinline__ void gst2(float2* p, int64_t i, float2 v) {
    ch_vf2 w;
    w.x = v.x;
    w.y = v.y;
    ((CH_GLOBAL ch_vf2*)p)[i] = w;
}
