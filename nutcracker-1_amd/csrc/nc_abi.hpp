// Helpers shared by the C-ABI entry points: host inputs staged into device
// buffers, zeroed outputs on failure, and the exception guard that turns a
// failure into a recorded message and an error code (the reference's Fortran
// core stops the process instead).
#pragma once

#include "nc_internal.hpp"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <memory>

namespace nc {

// Host-pointer residency (residency.cpp): with a byte budget set, inputs whose
// address, size and fingerprint match a device copy made earlier are not
// uploaded again, and outputs keep a device copy.
struct ResidencyStats {
    int64_t hits = 0, misses = 0, bytes_uploaded = 0, bytes_reused = 0, outputs_kept = 0, entries = 0,
            bytes_held = 0;
};
bool residency_on();
// the device copy of a host input (uploaded if absent or changed); null when
// residency is off or the array is larger than the budget
std::shared_ptr<DevBuf<char>> resident_input(const void* host, size_t count);
// d2h of an output, and its device copy kept for the host address
void resident_output(void* host, const double2* dev, size_t count);
int64_t residency_budget(int64_t bytes);
void residency_forget(const void* host);
ResidencyStats residency_stats();

// A host input on the device: the resident copy when there is one (read only),
// else an upload into a buffer of its own.
struct HostIn {
    ZBuf buf;
    std::shared_ptr<DevBuf<char>> res;
    HostIn(const void* host, size_t count) {
        if (count) res = resident_input(host, count);
        if (!res) {
            buf.resize(std::max<size_t>(1, count));
            if (count) h2d(buf.ptr, host, count * sizeof(double2));
        }
    }
    double2* ptr() { return res ? reinterpret_cast<double2*>(res->ptr) : buf.ptr; }
};

// an output written to the host (and kept on the device with residency on)
inline void host_out(void* host, const double2* dev, size_t count) { resident_output(host, dev, count); }

inline void fill_zero(void* p, size_t bytes) {
    if (p && bytes) memset(p, 0, bytes);
}

template <class F>
int guarded(const char* name, F&& f) {
    try {
        return f();
    } catch (Error const& e) {
        set_last_error(std::string(name) + ": " + e.what());
        fprintf(stderr, "[nutcracker-hip] %s: %s\n", name, e.what());
        return e.code != 0 ? e.code : -1;
    } catch (std::exception const& e) {
        set_last_error(std::string(name) + ": " + e.what());
        fprintf(stderr, "[nutcracker-hip] %s: %s\n", name, e.what());
        return -1;
    }
}

}  // namespace nc
