// Device runtime: error reporting, the per-process stream, a caching device
// allocator and host <-> device copies.
#include "nc_internal.hpp"

#include <atomic>
#include <condition_variable>
#include <cstdlib>
#include <exception>
#include <map>
#include <mutex>
#include <thread>

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

// ---------------------------------------------------------------------------
// Host thread pool for independent per-problem host work (small eigen-problems
// of a batch).  Workers = min(hardware threads, OMP_NUM_THREADS, 16) - 1; the
// calling thread takes part.
// ---------------------------------------------------------------------------
namespace {
struct HostJob {
    std::function<void(int)> const* fn = nullptr;
    int n = 0;
    std::atomic<int> next{0}, finished{0};
    std::mutex mu;
    std::condition_variable cv;
    std::exception_ptr err;
    // claims indices until none is left; a worker that arrives late finds none
    void work() {
        for (;;) {
            const int i = next.fetch_add(1);
            if (i >= n) break;
            try {
                (*fn)(i);
            } catch (...) {
                std::lock_guard<std::mutex> g(mu);
                if (!err) err = std::current_exception();
            }
            if (finished.fetch_add(1) + 1 == n) {
                std::lock_guard<std::mutex> g(mu);
                cv.notify_all();
            }
        }
    }
};
struct HostPool {
    std::mutex mu;
    std::condition_variable cv;
    std::shared_ptr<HostJob> job;
    uint64_t gen = 0;
    explicit HostPool(int workers) {
        for (int w = 0; w < workers; ++w) std::thread([this] { loop(); }).detach();
    }
    void loop() {
        uint64_t seen = 0;
        for (;;) {
            std::shared_ptr<HostJob> j;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv.wait(lk, [&] { return gen != seen; });
                seen = gen;
                j = job;
            }
            if (j) j->work();
        }
    }
    void run(int n, std::function<void(int)> const& f) {
        auto j = std::make_shared<HostJob>();
        j->fn = &f;
        j->n = n;
        {
            std::lock_guard<std::mutex> g(mu);
            job = j;
            ++gen;
        }
        cv.notify_all();
        j->work();
        {
            std::unique_lock<std::mutex> lk(j->mu);
            j->cv.wait(lk, [&] { return j->finished.load() == n; });
        }
        {
            std::lock_guard<std::mutex> g(mu);
            job.reset();
        }
        if (j->err) std::rethrow_exception(j->err);
    }
};
HostPool* host_pool() {
    static HostPool* p = [] {
        int hw = (int)std::thread::hardware_concurrency();
        if (hw <= 0) hw = 1;
        if (const char* e = std::getenv("OMP_NUM_THREADS")) {
            const int v = std::atoi(e);
            if (v > 0) hw = std::min(hw, v);
        }
        hw = std::min(hw, 16);
        return hw > 1 ? new HostPool(hw - 1) : nullptr;
    }();
    return p;
}
std::mutex g_pool_mu;  // one parallel region at a time
}  // namespace

void host_parallel_for(int n, std::function<void(int)> const& fn) {
    HostPool* p = n >= 4 ? host_pool() : nullptr;
    std::unique_lock<std::mutex> lk(g_pool_mu, std::defer_lock);
    if (p && lk.try_lock()) {
        p->run(n, fn);
        return;
    }
    for (int i = 0; i < n; ++i) fn(i);
}

}  // namespace nc
