// Helpers shared by the C-ABI entry points: host inputs staged into device
// buffers, zeroed outputs on failure, and the exception guard that turns a
// failure into a recorded message and an error code (the reference's Fortran
// core stops the process instead).
#pragma once

#include "nc_internal.hpp"

#include <algorithm>
#include <cstdio>
#include <cstring>

namespace nc {

struct HostIn {
    ZBuf buf;
    HostIn(const void* host, size_t count) : buf(std::max<size_t>(1, count)) {
        if (count) h2d(buf.ptr, host, count * sizeof(double2));
    }
    double2* ptr() { return buf.ptr; }
};

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
