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
std::mutex g_mu;

struct ScratchSlot { void* p = nullptr; size_t bytes = 0; };

// What one stream of work owns: the stream, the scratch memory its launches
// share, and the device copies of the small tables its launches read (kept
// until the context ends, so that a graph captured on the context can go on
// reading them).  The default context uses the stream set by
// chimera_set_stream and may drop its tables when the cache grows large.
struct Context {
    hipStream_t stream = nullptr;
    bool owns_stream = false;
    ScratchSlot scratch[SCRATCH_SLOTS];
    std::map<std::string, void*> tables;
    int graphs = 0;          // live graphs captured on this context
    bool capturing = false;
};
Context g_default;
thread_local Context* t_ctx = nullptr;

// A captured step: the instantiated graph and the context it was captured on
// (whose stream it is launched on and whose tables and scratch it reads).
struct Graph {
    hipGraphExec_t exec = nullptr;
    Context* ctx = nullptr;
};

Context& cur() { return t_ctx ? *t_ctx : g_default; }

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

void release_context(Context& c) {
    (void)hipStreamSynchronize(c.stream);
    for (auto& s : c.scratch) {
        if (s.p) (void)hipFree(s.p);
        s.p = nullptr;
        s.bytes = 0;
    }
    for (auto& kv : c.tables) (void)hipFree(kv.second);
    c.tables.clear();
}
}  // namespace

KTimer::KTimer(int f) : fam(f), on(g_prof && !cur().capturing), a(nullptr) {
    if (on) {
        a = event_get();
        if (a) (void)hipEventRecord(a, cur().stream);
        else on = false;
    }
}

KTimer::~KTimer() {
    if (!on) return;
    hipEvent_t b = event_get();
    if (!b) return;
    (void)hipEventRecord(b, cur().stream);
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

hipStream_t stream() { return cur().stream; }
void set_stream(hipStream_t s) { cur().stream = s; }
bool capturing() { return cur().capturing; }

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
        e = hipMemcpyAsync(d, p, bytes, hipMemcpyHostToDevice, cur().stream);
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
        if (b.copy_out) CH_HIP(hipMemcpyAsync(b.host, b.dev, b.bytes, hipMemcpyDeviceToHost, cur().stream));
    CH_HIP(hipStreamSynchronize(cur().stream));
    for (auto& b : bufs_) {
        (void)hipFree(b.dev);
        b.dev = nullptr;
    }
    bufs_.clear();
    return ST_OK;
}

Stage::~Stage() {
    if (!bufs_.empty()) {
        (void)hipStreamSynchronize(cur().stream);
        for (auto& b : bufs_)
            if (b.dev) (void)hipFree(b.dev);
    }
}

void* scratch(int slot, size_t bytes) {
    std::lock_guard<std::mutex> lk(g_mu);
    Context& c = cur();
    ScratchSlot& s = c.scratch[slot];
    if (s.bytes >= bytes && s.p) return s.p;
    if (c.capturing) {
        set_error("scratch slot %d must grow to %zu bytes inside a captured step: run the step once before "
                  "capturing it", slot, bytes);
        return nullptr;
    }
    if (s.p) {
        if (c.graphs > 0) {
            set_error("scratch slot %d must grow to %zu bytes while graphs captured on this context still use it",
                      slot, bytes);
            return nullptr;
        }
        (void)hipStreamSynchronize(c.stream);
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
    Context& c = cur();
    std::string key(static_cast<const char*>(host), bytes);
    auto it = c.tables.find(key);
    if (it != c.tables.end()) return it->second;
    if (c.capturing) {
        set_error("a %zu-byte table first needed inside a captured step: run the step once before capturing it",
                  bytes);
        return nullptr;
    }
    if (c.tables.size() > 4096 && c.graphs == 0) {  // bound the cache: tables of a stale run are dropped
        (void)hipStreamSynchronize(c.stream);
        for (auto& kv : c.tables) (void)hipFree(kv.second);
        c.tables.clear();
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
    c.tables.emplace(std::move(key), d);
    return d;
}

const void* shared_table(const void* host, size_t bytes) {
    static std::map<std::string, void*> tables;    // process lifetime: never freed
    std::lock_guard<std::mutex> lk(g_mu);
    std::string key(static_cast<const char*>(host), bytes);
    auto it = tables.find(key);
    if (it != tables.end()) return it->second;
    if (cur().capturing) {
        set_error("a %zu-byte shared table first needed inside a captured step: run the step once before "
                  "capturing it", bytes);
        return nullptr;
    }
    void* d = nullptr;
    if (hipMalloc(&d, bytes) != hipSuccess) {
        set_error("hipMalloc of a %zu-byte shared table failed", bytes);
        return nullptr;
    }
    if (hipMemcpy(d, host, bytes, hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(d);
        set_error("upload of a %zu-byte shared table failed", bytes);
        return nullptr;
    }
    tables.emplace(std::move(key), d);
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
        hipLaunchKernelGGL(k_fill_u32, blocks_for(n, 256), 256, 0, cur().stream, static_cast<uint32_t*>(p), v, n);
    } else {
        hipLaunchKernelGGL(k_fill_u8, blocks_for((int64_t)bytes, 256), 256, 0, cur().stream, static_cast<uint8_t*>(p), v8,
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
    CH_HIP(hipStreamSynchronize(cur().stream));
    CH_HIP(hipFree(ptr));
    return ST_OK;
}

int32_t chimera_memcpy(void* dst, const void* src, int64_t bytes) {
    if (bytes < 0 || (bytes > 0 && (dst == nullptr || src == nullptr))) return fail(ST_EINVAL, "chimera_memcpy: bad arguments");
    if (bytes == 0) return ST_OK;
    const bool dd = is_device_ptr(dst), sd = is_device_ptr(src);
    if (dd && sd) {
        CH_HIP(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDeviceToDevice, cur().stream));
        return ST_OK;
    }
    CH_HIP(hipStreamSynchronize(cur().stream));
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
    CH_HIP(hipDeviceSynchronize());
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
    CH_HIP(hipDeviceSynchronize());
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
    hipLaunchKernelGGL(k_axpb, blocks_for(n, 256), 256, 0, cur().stream, p, n, a, b);
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

// ---------------------------------------------------------------- contexts and captured steps

int32_t chimera_context_create(void** ctx) {
    if (ctx == nullptr) return fail(ST_EINVAL, "chimera_context_create: null handle");
    *ctx = nullptr;
    Context* c = new Context();
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return hip_fail(e, "hipStreamCreateWithFlags", __FILE__, __LINE__);
    }
    c->owns_stream = true;
    *ctx = c;
    return ST_OK;
}

int32_t chimera_context_destroy(void* ctx) {
    if (ctx == nullptr) return ST_OK;
    Context* c = static_cast<Context*>(ctx);
    if (c->capturing) return fail(ST_EINVAL, "chimera_context_destroy: the context is capturing");
    if (c->graphs > 0) return fail(ST_EINVAL, "chimera_context_destroy: %d graphs captured on it are alive", c->graphs);
    if (t_ctx == c) t_ctx = nullptr;
    release_context(*c);
    if (c->owns_stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return ST_OK;
}

int32_t chimera_context_bind(void* ctx) {
    if (cur().capturing) return fail(ST_EINVAL, "chimera_context_bind: the current context is capturing");
    t_ctx = static_cast<Context*>(ctx);
    return ST_OK;
}

void* chimera_context_current(void) { return t_ctx; }

int32_t chimera_context_query(void* ctx, int32_t* done) {
    if (done == nullptr) return fail(ST_EINVAL, "chimera_context_query: null output");
    Context& c = ctx ? *static_cast<Context*>(ctx) : g_default;
    hipError_t e = hipStreamQuery(c.stream);
    if (e == hipSuccess) {
        *done = 1;
        return ST_OK;
    }
    if (e == hipErrorNotReady) {
        (void)hipGetLastError();
        *done = 0;
        return ST_OK;
    }
    return hip_fail(e, "hipStreamQuery", __FILE__, __LINE__);
}

int32_t chimera_capture_begin(void) {
    Context& c = cur();
    if (c.capturing) return fail(ST_EINVAL, "chimera_capture_begin: already capturing");
    if (c.stream == nullptr) return fail(ST_EINVAL, "chimera_capture_begin: the null stream cannot be captured");
    CH_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
    c.capturing = true;
    return ST_OK;
}

int32_t chimera_capture_end(void** graph) {
    Context& c = cur();
    if (graph == nullptr) return fail(ST_EINVAL, "chimera_capture_end: null handle");
    *graph = nullptr;
    if (!c.capturing) return fail(ST_EINVAL, "chimera_capture_end: not capturing");
    c.capturing = false;
    hipGraph_t g = nullptr;
    CH_HIP(hipStreamEndCapture(c.stream, &g));
    hipGraphExec_t x = nullptr;
    hipError_t e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate", __FILE__, __LINE__);
    Graph* h = new Graph();
    h->exec = x;
    h->ctx = &c;
    c.graphs++;
    *graph = h;
    return ST_OK;
}

int32_t chimera_capture_abort(void) {
    Context& c = cur();
    if (!c.capturing) return ST_OK;
    c.capturing = false;
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(c.stream, &g);
    if (g) (void)hipGraphDestroy(g);
    (void)hipGetLastError();
    return ST_OK;
}

int32_t chimera_graph_launch(void* graph) {
    if (graph == nullptr) return fail(ST_EINVAL, "chimera_graph_launch: null graph");
    Graph* h = static_cast<Graph*>(graph);
    if (h->ctx->capturing) return fail(ST_EINVAL, "chimera_graph_launch: its context is capturing");
    CH_HIP(hipGraphLaunch(h->exec, h->ctx->stream));
    return ST_OK;
}

int32_t chimera_graph_destroy(void* graph) {
    if (graph == nullptr) return ST_OK;
    Graph* h = static_cast<Graph*>(graph);
    CH_HIP(hipStreamSynchronize(h->ctx->stream));
    CH_HIP(hipGraphExecDestroy(h->exec));
    if (h->ctx->graphs > 0) h->ctx->graphs--;
    delete h;
    return ST_OK;
}

int32_t chimera_device_synchronize(void) {
    CH_HIP(hipDeviceSynchronize());
    return ST_OK;
}

int32_t chimera_mem_info(int64_t* free_bytes, int64_t* total_bytes) {
    if (free_bytes == nullptr || total_bytes == nullptr) return fail(ST_EINVAL, "chimera_mem_info: null output");
    size_t f = 0, t = 0;
    CH_HIP(hipMemGetInfo(&f, &t));
    *free_bytes = (int64_t)f;
    *total_bytes = (int64_t)t;
    return ST_OK;
}

}  // extern "C"
