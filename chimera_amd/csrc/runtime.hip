// runtime.hip -- library state: stream, error text, staging buffers, scratch
// memory and the cached device copies of small host tables.
#include <map>
#include <mutex>
#include <vector>

#include "common.hpp"
#include "../../include/chimera_hip.h"

namespace chimera {

namespace {
thread_local std::string g_err;
hipStream_t g_stream = nullptr;
std::mutex g_mu;

struct ScratchSlot { void* p = nullptr; size_t bytes = 0; };
ScratchSlot g_scratch[SCRATCH_SLOTS];

std::map<std::string, void*> g_tables;

bool g_prof = false;
struct TimerRec { hipEvent_t a, b; int fam; };
std::vector<TimerRec> g_recs;
std::vector<hipEvent_t> g_free_events;

hipEvent_t event_get() {
    if (!g_free_events.empty()) {
        hipEvent_t e = g_free_events.back();
        g_free_events.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}
}  // namespace

KTimer::KTimer(int f) : fam(f), on(g_prof), a(nullptr) {
    if (on) {
        a = event_get();
        if (a) (void)hipEventRecord(a, g_stream);
        else on = false;
    }
}

KTimer::~KTimer() {
    if (!on) return;
    hipEvent_t b = event_get();
    if (!b) return;
    (void)hipEventRecord(b, g_stream);
    g_recs.push_back({a, b, fam});
}

void set_error(const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
}

int32_t fail(int32_t code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int32_t hip_fail(hipError_t e, const char* what, const char* file, int line) {
    set_error("%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
    return e == hipErrorOutOfMemory ? ST_ENOMEM : ST_EHIP;
}

hipStream_t stream() { return g_stream; }
void set_stream(hipStream_t s) { g_stream = s; }

bool is_device_ptr(const void* p) {
    if (p == nullptr) return false;
    hipPointerAttribute_t attr;
    hipError_t e = hipPointerGetAttributes(&attr, p);
    if (e != hipSuccess) {
        (void)hipGetLastError();  // unregistered host memory reports an error: clear it
        return false;
    }
    return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged;
}

void* Stage::put(void* p, size_t bytes, bool copy_in, bool copy_out) {
    if (st_ != ST_OK) return nullptr;
    if (bytes == 0) return p;
    if (p == nullptr) {
        st_ = fail(ST_EINVAL, "null array argument");
        return nullptr;
    }
    if (is_device_ptr(p)) return p;
    any_host_ = true;
    void* d = nullptr;
    hipError_t e = hipMalloc(&d, bytes);
    if (e != hipSuccess) {
        st_ = hip_fail(e, "hipMalloc (staging)", __FILE__, __LINE__);
        return nullptr;
    }
    bufs_.push_back({p, d, bytes, copy_out});
    if (copy_in) {
        e = hipMemcpyAsync(d, p, bytes, hipMemcpyHostToDevice, g_stream);
        if (e != hipSuccess) {
            st_ = hip_fail(e, "hipMemcpyAsync H2D (staging)", __FILE__, __LINE__);
            return nullptr;
        }
    }
    return d;
}

int32_t Stage::finish() {
    if (st_ != ST_OK) return st_;
    if (!any_host_) return ST_OK;
    for (auto& b : bufs_)
        if (b.copy_out) CH_HIP(hipMemcpyAsync(b.host, b.dev, b.bytes, hipMemcpyDeviceToHost, g_stream));
    CH_HIP(hipStreamSynchronize(g_stream));
    for (auto& b : bufs_) {
        (void)hipFree(b.dev);
        b.dev = nullptr;
    }
    bufs_.clear();
    return ST_OK;
}

Stage::~Stage() {
    if (!bufs_.empty()) {
        (void)hipStreamSynchronize(g_stream);
        for (auto& b : bufs_)
            if (b.dev) (void)hipFree(b.dev);
    }
}

void* scratch(int slot, size_t bytes) {
    std::lock_guard<std::mutex> lk(g_mu);
    ScratchSlot& s = g_scratch[slot];
    if (s.bytes >= bytes && s.p) return s.p;
    if (s.p) {
        (void)hipStreamSynchronize(g_stream);
        (void)hipFree(s.p);
        s.p = nullptr;
        s.bytes = 0;
    }
    size_t want = bytes < 256 ? 256 : bytes;
    if (hipMalloc(&s.p, want) != hipSuccess) {
        s.p = nullptr;
        set_error("hipMalloc of %zu scratch bytes failed", want);
        return nullptr;
    }
    s.bytes = want;
    return s.p;
}

const void* device_table(const void* host, size_t bytes) {
    std::lock_guard<std::mutex> lk(g_mu);
    std::string key(static_cast<const char*>(host), bytes);
    auto it = g_tables.find(key);
    if (it != g_tables.end()) return it->second;
    if (g_tables.size() > 4096) {  // bound the cache: tables of a stale run are dropped
        (void)hipStreamSynchronize(g_stream);
        for (auto& kv : g_tables) (void)hipFree(kv.second);
        g_tables.clear();
    }
    void* d = nullptr;
    if (hipMalloc(&d, bytes) != hipSuccess) {
        set_error("hipMalloc of a %zu-byte table failed", bytes);
        return nullptr;
    }
    if (hipMemcpy(d, host, bytes, hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(d);
        set_error("upload of a %zu-byte table failed", bytes);
        return nullptr;
    }
    g_tables.emplace(std::move(key), d);
    return d;
}

__global__ void k_fill_u32(uint32_t* __restrict__ p, uint32_t v, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_fill_u8(uint8_t* __restrict__ p, uint8_t v, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_axpb(double* __restrict__ p, int64_t n, double a, double b) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = a * p[i] + b;
}

// Byte fill by a kernel (not hipMemsetAsync), so that a captured graph replays it correctly.
int32_t fill_bytes(void* p, int value, size_t bytes) {
    if (bytes == 0) return ST_OK;
    const uint8_t v8 = (uint8_t)value;
    if ((reinterpret_cast<uintptr_t>(p) & 3) == 0 && (bytes & 3) == 0) {
        const uint32_t v = v8 * 0x01010101u;
        const int64_t n = (int64_t)(bytes / 4);
        hipLaunchKernelGGL(k_fill_u32, blocks_for(n, 256), 256, 0, g_stream, static_cast<uint32_t*>(p), v, n);
    } else {
        hipLaunchKernelGGL(k_fill_u8, blocks_for((int64_t)bytes, 256), 256, 0, g_stream, static_cast<uint8_t*>(p), v8,
                           (int64_t)bytes);
    }
    CH_LAUNCH_CHECK("k_fill");
    return ST_OK;
}

}  // namespace chimera

using namespace chimera;

extern "C" {

int32_t chimera_malloc(void** ptr, int64_t bytes) {
    if (ptr == nullptr || bytes < 0) return fail(ST_EINVAL, "chimera_malloc: bad arguments");
    *ptr = nullptr;
    if (bytes == 0) bytes = 256;
    CH_HIP(hipMalloc(ptr, (size_t)bytes));
    CH_TRY(fill_bytes(*ptr, 0, (size_t)bytes));
    return ST_OK;
}

int32_t chimera_free(void* ptr) {
    if (ptr == nullptr) return ST_OK;
    CH_HIP(hipStreamSynchronize(g_stream));
    CH_HIP(hipFree(ptr));
    return ST_OK;
}

int32_t chimera_memcpy(void* dst, const void* src, int64_t bytes) {
    if (bytes < 0 || (bytes > 0 && (dst == nullptr || src == nullptr))) return fail(ST_EINVAL, "chimera_memcpy: bad arguments");
    if (bytes == 0) return ST_OK;
    const bool dd = is_device_ptr(dst), sd = is_device_ptr(src);
    if (dd && sd) {
        CH_HIP(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDeviceToDevice, g_stream));
        return ST_OK;
    }
    CH_HIP(hipStreamSynchronize(g_stream));
    CH_HIP(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyDefault));
    return ST_OK;
}

int32_t chimera_memset(void* ptr, int32_t value, int64_t bytes) {
    if (bytes < 0 || (bytes > 0 && ptr == nullptr)) return fail(ST_EINVAL, "chimera_memset: bad arguments");
    return fill_bytes(ptr, value, (size_t)bytes);
}

int32_t chimera_prof_enable(int32_t on) {
    g_prof = on != 0;
    return ST_OK;
}

int32_t chimera_prof_reset(void) {
    CH_HIP(hipStreamSynchronize(g_stream));
    for (auto& r : g_recs) {
        g_free_events.push_back(r.a);
        g_free_events.push_back(r.b);
    }
    g_recs.clear();
    return ST_OK;
}

int32_t chimera_prof_read(int32_t family, double* total_ms, int64_t* count) {
    if (total_ms == nullptr || count == nullptr || family < 0 || family >= KF_COUNT)
        return fail(ST_EINVAL, "chimera_prof_read: bad arguments");
    CH_HIP(hipStreamSynchronize(g_stream));
    double t = 0.0;
    int64_t n = 0;
    for (auto& r : g_recs) {
        if (r.fam != family) continue;
        float ms = 0.f;
        CH_HIP(hipEventElapsedTime(&ms, r.a, r.b));
        t += ms;
        n++;
    }
    *total_ms = t;
    *count = n;
    return ST_OK;
}

int32_t chimera_axpb(double* p, int64_t n, double a, double b) {
    if (n < 0 || (n > 0 && p == nullptr)) return fail(ST_EINVAL, "chimera_axpb: bad arguments");
    if (n == 0) return ST_OK;
    hipLaunchKernelGGL(k_axpb, blocks_for(n, 256), 256, 0, g_stream, p, n, a, b);
    CH_LAUNCH_CHECK("k_axpb");
    return ST_OK;
}


const char* chimera_version(void) { return "chimera_hip 0.1 (gfx950)"; }

const char* chimera_last_error(void) { return g_err.c_str(); }

int32_t chimera_set_stream(void* hip_stream) {
    set_stream(static_cast<hipStream_t>(hip_stream));
    return ST_OK;
}

void* chimera_get_stream(void) { return static_cast<void*>(stream()); }

int32_t chimera_synchronize(void) {
    CH_HIP(hipStreamSynchronize(stream()));
    return ST_OK;
}

}  // extern "C"
