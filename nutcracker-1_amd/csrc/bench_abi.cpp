// Timing entry points for the H_eff matvec with every operand resident in HBM.
// Timing uses HIP events recorded on the library's stream (the stream every
// kernel of the matvec is launched on).
#include "nc_internal.hpp"
#include "../../include/nutcracker_core_abi.h"

#include <algorithm>
#include <type_traits>

using namespace nc;

namespace {

template <typename CT>
struct BenchPlan;
template <>
struct BenchPlan<double2> {
    typedef GemmPlan type;
    static void run(GemmPlan const& p, hipStream_t s) { zgemm_run(p, s); }
};
template <>
struct BenchPlan<float2> {
    typedef CGemmPlan type;
    static void run(CGemmPlan const& p, hipStream_t s) { cgemm_run(p, s); }
};

// Operands arrive as complex128 on the host; CT = float2 narrows them on the
// device before timing, and widens y back to complex128.
template <typename CT>
int bench_heff(const char* name, uint32_t bl, uint32_t br, uint32_t cl, uint32_t cr, uint32_t d, uint32_t nmat,
               const uint32_t* idx, const nc_zcomplex* mats, const nc_zcomplex* L, const nc_zcomplex* R,
               const nc_zcomplex* x, nc_zcomplex* y, uint32_t warmup, uint32_t reps, double* total_ms,
               double* phase_ms) {
    try {
        auto& c = Context::get();
        hipStream_t s = c.stream;
        MpoSite op;
        op.build((int)d, (int)cl, (int)cr, (int)nmat, idx, reinterpret_cast<const zd*>(mats));
        const size_t nL = (size_t)bl * bl * cl, nR = (size_t)br * br * cr, nx = (size_t)br * bl * d;
        DevBuf<CT> dL(nL), dR(nR), dx(nx), dy(nx);
        {
            ZBuf tL(nL), tR(nR), tx(nx);
            h2d(tL.ptr, L, nL * sizeof(double2));
            h2d(tR.ptr, R, nR * sizeof(double2));
            h2d(tx.ptr, x, nx * sizeof(double2));
            if constexpr (std::is_same<CT, double2>::value) {
                d2d(dL.ptr, tL.ptr, nL * sizeof(double2));
                d2d(dR.ptr, tR.ptr, nR * sizeof(double2));
                d2d(dx.ptr, tx.ptr, nx * sizeof(double2));
            } else {
                to_c64(tL.ptr, dL.ptr, (int64_t)nL, s);
                to_c64(tR.ptr, dR.ptr, (int64_t)nR, s);
                to_c64(tx.ptr, dx.ptr, (int64_t)nx, s);
            }
            NC_HIP(hipStreamSynchronize(s));
        }
        Workspace ws;
        for (uint32_t i = 0; i < warmup; ++i) heff_matvec((int)bl, (int)br, op, dL.ptr, dR.ptr, dx.ptr, dy.ptr, ws, s);
        NC_HIP(hipStreamSynchronize(s));
        hipEvent_t e0, e1;
        NC_HIP(hipEventCreate(&e0));
        NC_HIP(hipEventCreate(&e1));
        NC_HIP(hipEventRecord(e0, s));
        for (uint32_t i = 0; i < reps; ++i) heff_matvec((int)bl, (int)br, op, dL.ptr, dR.ptr, dx.ptr, dy.ptr, ws, s);
        NC_HIP(hipEventRecord(e1, s));
        NC_HIP(hipEventSynchronize(e1));
        float ms = 0;
        NC_HIP(hipEventElapsedTime(&ms, e0, e1));
        *total_ms = ms;
        if (phase_ms) {
            // per-phase averages: the two GEMM launches (step A, step C) and the mix kernel
            phase_ms[0] = phase_ms[1] = phase_ms[2] = 0;
            hipEvent_t ev[4];
            for (auto& e : ev) NC_HIP(hipEventCreate(&e));
            const uint32_t preps = reps < 3 ? reps : 3;
            const int64_t rs = (int64_t)br * bl;
            ws.ensure(rs * d * op.froms.size(), rs * d * op.tos.size(), (CT*)nullptr);
            CT* P = ws.p((CT*)nullptr);
            CT* Q = ws.q((CT*)nullptr);
            for (uint32_t i = 0; i < preps; ++i) {
                typename BenchPlan<CT>::type pa;
                pa.M = (int)br; pa.N = (int)bl;
                pa.key = TuneKey{'A', (int)bl, (int)br, (int)d, (int)cl, (int)cr, 1};
                for (size_t fi = 0; fi < op.froms.size(); ++fi)
                    for (uint32_t a = 0; a < d; ++a) {
                        pa.begin_batch(P + ((int64_t)fi * d + a) * rs, br);
                        pa.add_seg(dx.ptr + a * rs, br, dL.ptr + (int64_t)op.froms[fi] * bl * bl, bl, (int)bl);
                    }
                typename BenchPlan<CT>::type pc;
                pc.M = (int)br; pc.N = (int)(bl * d);
                pc.key = TuneKey{'C', (int)bl, (int)br, (int)d, (int)cl, (int)cr, 1};
                pc.begin_batch(dy.ptr, br);
                for (size_t ti = 0; ti < op.tos.size(); ++ti)
                    pc.add_seg(dR.ptr + (int64_t)op.tos[ti] * br * br, br, Q + (int64_t)ti * d * rs, br, (int)br);
                NC_HIP(hipEventRecord(ev[0], s));
                BenchPlan<CT>::run(pa, s);
                NC_HIP(hipEventRecord(ev[1], s));
                mix_blocks(op, true, P, Q, rs, s);
                NC_HIP(hipEventRecord(ev[2], s));
                BenchPlan<CT>::run(pc, s);
                NC_HIP(hipEventRecord(ev[3], s));
                NC_HIP(hipEventSynchronize(ev[3]));
                float t;
                NC_HIP(hipEventElapsedTime(&t, ev[0], ev[1])); phase_ms[0] += t / preps;
                NC_HIP(hipEventElapsedTime(&t, ev[1], ev[2])); phase_ms[1] += t / preps;
                NC_HIP(hipEventElapsedTime(&t, ev[2], ev[3])); phase_ms[2] += t / preps;
            }
            for (auto& e : ev) NC_HIP(hipEventDestroy(e));
        }
        NC_HIP(hipEventDestroy(e0));
        NC_HIP(hipEventDestroy(e1));
        if constexpr (std::is_same<CT, double2>::value) {
            d2h(y, dy.ptr, nx * sizeof(double2));
        } else {
            ZBuf ty(nx);
            to_c128(dy.ptr, ty.ptr, (int64_t)nx, s);
            NC_HIP(hipStreamSynchronize(s));
            d2h(y, ty.ptr, nx * sizeof(double2));
        }
        return 0;
    } catch (std::exception const& e) {
        set_last_error(std::string(name) + ": " + e.what());
        fprintf(stderr, "[nutcracker-hip] %s: %s\n", name, e.what());
        return -1;
    }
}

// Plain 16-byte-per-lane copy: the calibration kernel for HBM byte counters
// (reads and writes exactly `n16` * 16 bytes).
__global__ __launch_bounds__(256) void nc_copy_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                      int64_t n16) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) dst[i] = src[i];
}

}  // namespace

extern "C" int nc_bench_copy(uint64_t bytes, uint32_t reps, double* avg_ms) {
    try {
        if (bytes % 16 != 0 || bytes == 0) throw Error("copy size must be a positive multiple of 16 bytes", -2);
        auto& c = Context::get();
        hipStream_t s = c.stream;
        const int64_t n16 = (int64_t)(bytes / 16);
        DevBuf<uint4> a((size_t)n16), b((size_t)n16);
        NC_HIP(hipMemsetAsync(a.ptr, 1, bytes, s));
        const unsigned grid = (unsigned)std::min<int64_t>(8192, (n16 + 255) / 256);
        hipLaunchKernelGGL(nc_copy_kernel, dim3(grid), dim3(256), 0, s, a.ptr, b.ptr, n16);  // warm-up
        NC_HIP(hipGetLastError());
        hipEvent_t e0, e1;
        NC_HIP(hipEventCreate(&e0));
        NC_HIP(hipEventCreate(&e1));
        NC_HIP(hipEventRecord(e0, s));
        for (uint32_t r = 0; r < reps; ++r)
            hipLaunchKernelGGL(nc_copy_kernel, dim3(grid), dim3(256), 0, s, a.ptr, b.ptr, n16);
        NC_HIP(hipGetLastError());
        NC_HIP(hipEventRecord(e1, s));
        NC_HIP(hipEventSynchronize(e1));
        float ms = 0;
        NC_HIP(hipEventElapsedTime(&ms, e0, e1));
        *avg_ms = reps ? ms / reps : 0.0;
        NC_HIP(hipEventDestroy(e0));
        NC_HIP(hipEventDestroy(e1));
        return 0;
    } catch (std::exception const& e) {
        set_last_error(std::string("nc_bench_copy: ") + e.what());
        fprintf(stderr, "[nutcracker-hip] nc_bench_copy: %s\n", e.what());
        return -1;
    }
}

extern "C" int nc_bench_heff(uint32_t bl, uint32_t br, uint32_t cl, uint32_t cr, uint32_t d, uint32_t nmat,
                             const uint32_t* idx, const nc_zcomplex* mats, const nc_zcomplex* L,
                             const nc_zcomplex* R, const nc_zcomplex* x, nc_zcomplex* y, uint32_t warmup,
                             uint32_t reps, double* total_ms, double* phase_ms) {
    return bench_heff<double2>("nc_bench_heff", bl, br, cl, cr, d, nmat, idx, mats, L, R, x, y, warmup, reps,
                               total_ms, phase_ms);
}

extern "C" int nc_bench_heff_c64(uint32_t bl, uint32_t br, uint32_t cl, uint32_t cr, uint32_t d, uint32_t nmat,
                                 const uint32_t* idx, const nc_zcomplex* mats, const nc_zcomplex* L,
                                 const nc_zcomplex* R, const nc_zcomplex* x, nc_zcomplex* y, uint32_t warmup,
                                 uint32_t reps, double* total_ms, double* phase_ms) {
    return bench_heff<float2>("nc_bench_heff_c64", bl, br, cl, cr, d, nmat, idx, mats, L, R, x, y, warmup, reps,
                              total_ms, phase_ms);
}

// The complex64 matvec on 16-bit planes (the mixed form's single-site matvec,
// in the plane form nc_set_plane_form selects): L and R are split once outside
// the timed region; every timed matvec splits x (with its scale reduction for
// fp16 planes), runs step A, the mixing into Q's planes, step C and the FP64
// sum of the K slices.  phase_ms (5 entries, optional): split of x, step A,
// mix, step C, slice sum.
extern "C" int nc_bench_heff_planes(uint32_t bl, uint32_t br, uint32_t cl, uint32_t cr, uint32_t d, uint32_t nmat,
                                    const uint32_t* idx, const nc_zcomplex* mats, const nc_zcomplex* L,
                                    const nc_zcomplex* R, const nc_zcomplex* x, nc_zcomplex* y, uint32_t warmup,
                                    uint32_t reps, double* total_ms, double* phase_ms) {
    try {
        auto& c = Context::get();
        hipStream_t s = c.stream;
        MpoSite op;
        op.build((int)d, (int)cl, (int)cr, (int)nmat, idx, reinterpret_cast<const zd*>(mats));
        const size_t nL = (size_t)bl * bl * cl, nR = (size_t)br * br * cr, nx = (size_t)br * bl * d;
        ZBuf dL(nL), dR(nR), dx(nx), dy(nx);
        h2d(dL.ptr, L, nL * sizeof(double2));
        h2d(dR.ptr, R, nR * sizeof(double2));
        h2d(dx.ptr, x, nx * sizeof(double2));
        PlaneHeff ph((int)bl, (int)br, op, dL.ptr, dR.ptr, s);
        for (uint32_t i = 0; i < warmup; ++i) ph.apply(dx.ptr, dy.ptr, s);
        NC_HIP(hipStreamSynchronize(s));
        hipEvent_t ev[6];
        for (auto& e : ev) NC_HIP(hipEventCreate(&e));
        NC_HIP(hipEventRecord(ev[0], s));
        for (uint32_t i = 0; i < reps; ++i) ph.apply(dx.ptr, dy.ptr, s);
        NC_HIP(hipEventRecord(ev[1], s));
        NC_HIP(hipEventSynchronize(ev[1]));
        float ms = 0;
        NC_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
        *total_ms = ms;
        if (phase_ms) {
            for (int k = 0; k < 5; ++k) phase_ms[k] = 0;
            const uint32_t preps = reps < 3 ? (reps ? reps : 1) : 3;
            for (uint32_t i = 0; i < preps; ++i) {
                NC_HIP(hipEventRecord(ev[0], s));
                ph.split_x(dx.ptr, s);
                NC_HIP(hipEventRecord(ev[1], s));
                ph.step_a(s);
                NC_HIP(hipEventRecord(ev[2], s));
                ph.mix(s);
                NC_HIP(hipEventRecord(ev[3], s));
                ph.step_c(s);
                NC_HIP(hipEventRecord(ev[4], s));
                ph.sum(dy.ptr, s);
                NC_HIP(hipEventRecord(ev[5], s));
                NC_HIP(hipEventSynchronize(ev[5]));
                for (int k = 0; k < 5; ++k) {
                    float t;
                    NC_HIP(hipEventElapsedTime(&t, ev[k], ev[k + 1]));
                    phase_ms[k] += t / preps;
                }
            }
        }
        for (auto& e : ev) NC_HIP(hipEventDestroy(e));
        NC_HIP(hipStreamSynchronize(s));
        d2h(y, dy.ptr, nx * sizeof(double2));
        return 0;
    } catch (std::exception const& e) {
        set_last_error(std::string("nc_bench_heff_planes: ") + e.what());
        fprintf(stderr, "[nutcracker-hip] nc_bench_heff_planes: %s\n", e.what());
        return -1;
    }
}

// K slices of the two plane GEMMs of a matvec: out[0] step A, out[1] step C
extern "C" int nc_planes_splits(uint32_t bl, uint32_t br, uint32_t d, uint32_t nfrom, uint32_t nto, int32_t* out) {
    const int64_t tiles_a = (int64_t)((br + kPlaneTile - 1) / kPlaneTile) * ((bl + kPlaneTile - 1) / kPlaneTile) *
                            nfrom * d;
    const int64_t tiles_c = (int64_t)((br + kPlaneTile - 1) / kPlaneTile) * ((bl * d + kPlaneTile - 1) / kPlaneTile);
    try {
        out[0] = plane_split_count(tiles_a, (int)((bl + 31) / 32));
        out[1] = plane_split_count(tiles_c, (int)(nto * ((br + 31) / 32)));
    } catch (std::exception const& e) {
        set_last_error(std::string("nc_planes_splits: ") + e.what());
        return -1;
    }
    return 0;
}

// Site-move statistics: polar splits, Newton-Schulz iterations, SVD fallbacks.
extern "C" int nc_polar_stats(int64_t* out) {
    nc::PolarStats p = nc::polar_stats();
    out[0] = p.calls;
    out[1] = p.iterations;
    out[2] = p.svd_fallbacks;
    return 0;
}

extern "C" int nc_set_polar_max_steps(int steps) {
    if (steps < 0 || steps > 1000) return -1;
    const int prev = nc::polar_max_steps();
    nc::polar_max_steps() = steps;
    return prev;
}
