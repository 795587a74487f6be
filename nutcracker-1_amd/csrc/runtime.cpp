// Device runtime: error reporting, the per-process stream, a caching device
// allocator and host <-> device copies.
#include "nc_internal.hpp"

#include <map>
#include <mutex>

namespace nc {

static thread_local std::string g_last_error;

void set_last_error(std::string const& m) { g_last_error = m; }
std::string const& last_error() { return g_last_error; }

void check_hip(hipError_t e, const char* what, const char* file, int line) {
    if (e != hipSuccess) {
        char buf[512];
        snprintf(buf, sizeof buf, "HIP error %d (%s) in %s at %s:%d", (int)e, hipGetErrorString(e),
                 what, file, line);
        throw Error(buf, -100);
    }
}

Context& Context::get() {
    static Context* ctx = [] {
        auto* c = new Context();
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
            throw Error("no HIP device is visible: the MI355X core needs a GPU", -101);
        NC_HIP(hipGetDevice(&c->device));
        NC_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        NC_HIP(hipEventCreate(&c->ev_a));
        NC_HIP(hipEventCreate(&c->ev_b));
        return c;
    }();
    return *ctx;
}

void Context::sync() { NC_HIP(hipStreamSynchronize(stream)); }

// ----------------------------------------------------------------------------
// Caching allocator
// ----------------------------------------------------------------------------
namespace {
struct Pool {
    std::mutex mu;
    std::map<size_t, std::vector<void*>> free_lists;
    size_t in_use = 0, cached = 0;
};
Pool& pool() {
    static Pool* p = new Pool();
    return *p;
}
size_t round_size(size_t b) {
    if (b == 0) b = 1;
    if (b <= (size_t(64) << 20)) return (b + 4095) & ~size_t(4095);
    // above 64 MiB: round up to a quarter of the next power of two
    size_t p2 = 1;
    while (p2 < b) p2 <<= 1;
    size_t q = p2 / 4;
    return ((b + q - 1) / q) * q;
}
}  // namespace

void* dev_alloc(size_t bytes) {
    Context::get();
    size_t r = round_size(bytes);
    Pool& p = pool();
    {
        std::lock_guard<std::mutex> g(p.mu);
        auto it = p.free_lists.find(r);
        if (it != p.free_lists.end() && !it->second.empty()) {
            void* ptr = it->second.back();
            it->second.pop_back();
            p.cached -= r;
            p.in_use += r;
            return ptr;
        }
    }
    void* ptr = nullptr;
    hipError_t e = hipMalloc(&ptr, r);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        dev_release_cache();
        NC_HIP(hipMalloc(&ptr, r));
    }
    std::lock_guard<std::mutex> g(p.mu);
    p.in_use += r;
    return ptr;
}

void dev_free(void* ptr, size_t bytes) {
    if (!ptr) return;
    size_t r = round_size(bytes);
    Pool& p = pool();
    std::lock_guard<std::mutex> g(p.mu);
    p.free_lists[r].push_back(ptr);
    p.in_use -= r;
    p.cached += r;
}

size_t dev_bytes_in_use() { return pool().in_use; }

void dev_release_cache() {
    Pool& p = pool();
    std::lock_guard<std::mutex> g(p.mu);
    // the stream may still use freed blocks: drain it before returning memory
    (void)hipStreamSynchronize(Context::get().stream);
    for (auto& kv : p.free_lists)
        for (void* q : kv.second) (void)hipFree(q);
    p.free_lists.clear();
    p.cached = 0;
}

void h2d(void* dst, void const* src, size_t bytes) {
    if (!bytes) return;
    auto& c = Context::get();
    NC_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c.stream));
    NC_HIP(hipStreamSynchronize(c.stream));
}
void d2h(void* dst, void const* src, size_t bytes) {
    if (!bytes) return;
    auto& c = Context::get();
    NC_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c.stream));
    NC_HIP(hipStreamSynchronize(c.stream));
}
void d2d(void* dst, void const* src, size_t bytes) {
    if (!bytes) return;
    NC_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, Context::get().stream));
}
void dzero(void* dst, size_t bytes) {
    if (!bytes) return;
    NC_HIP(hipMemsetAsync(dst, 0, bytes, Context::get().stream));
}

int GemmPlan::max_k() const {
    int best = 0;
    for (auto const& b : batches) {
        int k = 0;
        for (int s = b.seg_begin; s < b.seg_end; ++s) k += segs[s].K;
        best = std::max(best, k);
    }
    return best;
}

}  // namespace nc
