// The drop-in boundary: extern "C" entry points with the reference's Fortran-core
// names and calling convention (every scalar by pointer, complex128 column-major
// arrays, 1-based block indices), as declared in sources/common/core.cpp.  Each
// call copies its host inputs to the device, runs the MI355X kernels and copies
// the outputs back; nothing is computed on the host CPU except the small
// Hessenberg / <= 4 x 4 Hermitian eigen-problems of the eigen-solver.
//
// Failures never abort the process: the entry point records a message
// (nc_last_error) and returns an error code, or leaves zeroed outputs for the
// reference's void routines.
#include "nc_abi.hpp"
#include "../../include/nutcracker_core_abi.h"

#include <algorithm>
#include <cstring>

using namespace nc;

namespace {

Workspace& ws() {
    static thread_local Workspace w;
    return w;
}

void build_op(MpoSite& op, uint32_t d, uint32_t cl, uint32_t cr, uint32_t nmat, const uint32_t* idx, const nc_zcomplex* mats) {
    op.build((int)d, (int)cl, (int)cr, (int)nmat, idx, reinterpret_cast<const zd*>(mats));
}

}  // namespace

extern "C" {

const char* nc_last_error(void) { return last_error().c_str(); }

int nc_set_optimizer_ncv(int ncv) {
    if (ncv < 3 || ncv > kMaxNcv) return -1;
    const int prev = optimizer_ncv();
    optimizer_ncv() = ncv;
    return prev;
}

int nc_set_mixed_planes(int mode) {
    if (mode < 0 || mode > 2) return -1;
    const int prev = mixed_planes_mode();
    mixed_planes_mode() = mode;
    return prev;
}

int nc_set_plane_form(int form) {
    if (form != kPlaneBF16 && form != kPlaneF16) return -1;
    const int prev = plane_form();
    plane_form() = form;
    return prev;
}

int64_t nc_set_host_residency(int64_t max_bytes) { return residency_budget(max_bytes); }

int nc_host_residency_forget(const void* host) {
    residency_forget(host);
    return 0;
}

int nc_host_residency_stats(int64_t* out) {
    const ResidencyStats st = residency_stats();
    out[0] = st.hits;
    out[1] = st.misses;
    out[2] = st.bytes_uploaded;
    out[3] = st.bytes_reused;
    out[4] = st.outputs_kept;
    out[5] = st.entries;
    out[6] = st.bytes_held;
    return 0;
}

int nc_optimizer_stats(int64_t* out) {
    const OptimizerStats st = optimizer_stats();
    out[0] = st.solves;
    out[1] = st.first_run_failures;
    out[2] = st.ncv_retries;
    out[3] = st.dense_fallbacks;
    out[4] = st.unconverged;
    return 0;
}

int nc_device_info(char* name, int name_len, int* cu_count, double* clock_ghz, size_t* total_mem) {
    return guarded("nc_device_info", [&] {
        auto& c = Context::get();
        hipDeviceProp_t p;
        NC_HIP(hipGetDeviceProperties(&p, c.device));
        if (name && name_len > 0) {
            strncpy(name, p.gcnArchName, (size_t)name_len - 1);
            name[name_len - 1] = 0;
        }
        if (cu_count) *cu_count = p.multiProcessorCount;
        if (clock_ghz) *clock_ghz = p.clockRate / 1e6;
        if (total_mem) *total_mem = p.totalGlobalMem;
        return 0;
    });
}

// ---------------------------------------------------------------------------
// Environment contractors
// ---------------------------------------------------------------------------
void contract_sos_left_(const uint32_t* bl, const uint32_t* br, const uint32_t* cl, const uint32_t* cr,
                        const uint32_t* d, const nc_zcomplex* left_environment, const uint32_t* number_of_matrices,
                        const uint32_t* sparse_operator_indices, const nc_zcomplex* sparse_operator_matrices,
                        const nc_zcomplex* state_site_tensor, nc_zcomplex* new_left_environment) {
    const size_t nout = (size_t)*br * *br * *cr;
    int rc = guarded("contract_sos_left_", [&] {
        auto s = Context::get().stream;
        MpoSite op;
        build_op(op, *d, *cl, *cr, *number_of_matrices, sparse_operator_indices, sparse_operator_matrices);
        HostIn L(left_environment, (size_t)*bl * *bl * *cl);
        HostIn psi(state_site_tensor, (size_t)*br * *bl * *d);
        ZBuf out(nout);
        env_left((int)*bl, (int)*br, op, L.ptr(), psi.ptr(), out.ptr, ws(), s);
        host_out(new_left_environment, out.ptr, nout);
        return 0;
    });
    if (rc) fill_zero(new_left_environment, nout * sizeof(nc_zcomplex));
}

void contract_sos_right_(const uint32_t* bl, const uint32_t* br, const uint32_t* cl, const uint32_t* cr,
                         const uint32_t* d, const nc_zcomplex* right_environment, const uint32_t* number_of_matrices,
                         const uint32_t* sparse_operator_indices, const nc_zcomplex* sparse_operator_matrices,
                         const nc_zcomplex* state_site_tensor, nc_zcomplex* new_right_environment) {
    const size_t nout = (size_t)*bl * *bl * *cl;
    int rc = guarded("contract_sos_right_", [&] {
        auto s = Context::get().stream;
        MpoSite op;
        build_op(op, *d, *cl, *cr, *number_of_matrices, sparse_operator_indices, sparse_operator_matrices);
        HostIn R(right_environment, (size_t)*br * *br * *cr);
        HostIn psi(state_site_tensor, (size_t)*br * *bl * *d);
        ZBuf out(nout);
        env_right((int)*bl, (int)*br, op, R.ptr(), psi.ptr(), out.ptr, ws(), s);
        host_out(new_right_environment, out.ptr, nout);
        return 0;
    });
    if (rc) fill_zero(new_right_environment, nout * sizeof(nc_zcomplex));
}

void compute_expectation_(const uint32_t* bl, const uint32_t* br, const uint32_t* cl, const uint32_t* cr,
                          const uint32_t* d, const nc_zcomplex* left_environment,
                          const nc_zcomplex* state_site_tensor, const uint32_t* number_of_matrices,
                          const uint32_t* sparse_operator_indices, const nc_zcomplex* sparse_operator_matrices,
                          const nc_zcomplex* right_environment, nc_zcomplex* expectation) {
    int rc = guarded("compute_expectation_", [&] {
        auto s = Context::get().stream;
        MpoSite op;
        build_op(op, *d, *cl, *cr, *number_of_matrices, sparse_operator_indices, sparse_operator_matrices);
        HostIn L(left_environment, (size_t)*bl * *bl * *cl);
        HostIn R(right_environment, (size_t)*br * *br * *cr);
        HostIn psi(state_site_tensor, (size_t)*br * *bl * *d);
        zd e = nc::expectation((int)*bl, (int)*br, op, L.ptr(), psi.ptr(), R.ptr(), ws(), s);
        expectation->re = e.real();
        expectation->im = e.imag();
        return 0;
    });
    if (rc) { expectation->re = 0; expectation->im = 0; }
}

void contract_expectation_boundaries_(const uint32_t* b, const uint32_t* c, const nc_zcomplex* left_boundary,
                                      const nc_zcomplex* right_boundary, nc_zcomplex* expectation) {
    int rc = guarded("contract_expectation_boundaries_", [&] {
        auto s = Context::get().stream;
        const size_t n = (size_t)*b * *b * *c;
        HostIn L(left_boundary, n), R(right_boundary, n);
        zd e = trace_pair(R.ptr(), L.ptr(), *b, *c, s);  // sum L(j,i,k) R(i,j,k)
        expectation->re = e.real();
        expectation->im = e.imag();
        return 0;
    });
    if (rc) { expectation->re = 0; expectation->im = 0; }
}

// ---------------------------------------------------------------------------
// The three matvec stages of the reference, for parity checks of the grouped
// kernels (core.f95:532-621).  Stage 1 materialises the dense tensor exactly as
// the reference does; the eigen-solver itself never builds it.
// ---------------------------------------------------------------------------
__global__ void stage1_kernel(const double2* L, const double2* W, const int* from0, const int* to0, int nmat,
                              int64_t bl, int64_t d, int64_t cr, double2* I) {
    // I(i, a, k2, j, b) += W_n(a, b) L(i, j, from_n) over blocks with to_n = k2
    const int64_t total = bl * d * cr * bl * d;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        int64_t r = e;
        const int64_t i = r % bl; r /= bl;
        const int64_t a = r % d; r /= d;
        const int64_t k2 = r % cr; r /= cr;
        const int64_t j = r % bl; r /= bl;
        const int64_t b = r;
        double2 acc = make_double2(0, 0);
        for (int n = 0; n < nmat; ++n) {
            if (to0[n] != k2) continue;
            const double2 w = W[a + b * d + (int64_t)n * d * d];
            const double2 l = L[i + j * bl + (int64_t)from0[n] * bl * bl];
            acc.x += w.x * l.x - w.y * l.y;
            acc.y += w.x * l.y + w.y * l.x;
        }
        I[e] = acc;
    }
}

void iteration_stage_1_(const uint32_t* bl, const uint32_t* cl, const uint32_t* cr, const uint32_t* d,
                        const nc_zcomplex* left_environment, const uint32_t* number_of_matrices,
                        const uint32_t* sparse_operator_indices, const nc_zcomplex* sparse_operator_matrices,
                        nc_zcomplex* iteration_stage_1_tensor) {
    const size_t nout = (size_t)*bl * *d * *cr * *bl * *d;
    int rc = guarded("iteration_stage_1_", [&] {
        auto s = Context::get().stream;
        const int nmat = (int)*number_of_matrices;
        std::vector<int> f(nmat), t(nmat);
        for (int n = 0; n < nmat; ++n) {
            f[n] = (int)sparse_operator_indices[2 * n] - 1;
            t[n] = (int)sparse_operator_indices[2 * n + 1] - 1;
            if (f[n] < 0 || f[n] >= (int)*cl || t[n] < 0 || t[n] >= (int)*cr) throw Error("block index out of range", -3);
        }
        DevBuf<int> df(nmat), dt(nmat);
        h2d(df.ptr, f.data(), nmat * sizeof(int));
        h2d(dt.ptr, t.data(), nmat * sizeof(int));
        HostIn L(left_environment, (size_t)*bl * *bl * *cl);
        HostIn W(sparse_operator_matrices, (size_t)*d * *d * nmat);
        ZBuf out(nout);
        unsigned g = (unsigned)std::min<size_t>(4096, (nout + 255) / 256);
        hipLaunchKernelGGL(stage1_kernel, dim3(g), dim3(256), 0, s, L.ptr(), W.ptr(), df.ptr, dt.ptr, nmat,
                           (int64_t)*bl, (int64_t)*d, (int64_t)*cr, out.ptr);
        NC_HIP(hipGetLastError());
        d2h(iteration_stage_1_tensor, out.ptr, nout * sizeof(double2));
        return 0;
    });
    if (rc) fill_zero(iteration_stage_1_tensor, nout * sizeof(nc_zcomplex));
}

void iteration_stage_2_(const uint32_t* bl, const uint32_t* br, const uint32_t* cr, const uint32_t* d,
                        const nc_zcomplex* iteration_stage_1_tensor, const nc_zcomplex* state_site_tensor,
                        nc_zcomplex* iteration_stage_2_tensor) {
    const size_t nout = (size_t)*br * *cr * *bl * *d;
    int rc = guarded("iteration_stage_2_", [&] {
        auto s = Context::get().stream;
        const int64_t K = (int64_t)*bl * *d, N = (int64_t)*cr * *bl * *d;
        HostIn I(iteration_stage_1_tensor, (size_t)K * N);
        HostIn x(state_site_tensor, (size_t)*br * K);
        ZBuf out(nout);
        GemmPlan p;
        p.M = (int)*br; p.N = (int)N;
        p.begin_batch(out.ptr, *br);
        p.add_seg(x.ptr(), *br, I.ptr(), K, (int)K);
        zgemm_run(p, s);
        d2h(iteration_stage_2_tensor, out.ptr, nout * sizeof(double2));
        return 0;
    });
    if (rc) fill_zero(iteration_stage_2_tensor, nout * sizeof(nc_zcomplex));
}

void iteration_stage_3_(const uint32_t* bl, const uint32_t* br, const uint32_t* cr, const uint32_t* d,
                        const nc_zcomplex* iteration_stage_2_tensor, const nc_zcomplex* right_environment,
                        nc_zcomplex* output_state_site_tensor) {
    const size_t nout = (size_t)*br * *bl * *d;
    int rc = guarded("iteration_stage_3_", [&] {
        auto s = Context::get().stream;
        const int64_t K = (int64_t)*br * *cr, N = (int64_t)*bl * *d;
        HostIn T(iteration_stage_2_tensor, (size_t)K * N);
        HostIn R(right_environment, (size_t)*br * K);
        ZBuf out(nout);
        GemmPlan p;
        p.M = (int)*br; p.N = (int)N;
        p.begin_batch(out.ptr, *br);
        p.add_seg(R.ptr(), *br, T.ptr(), K, (int)K);
        zgemm_run(p, s);
        d2h(output_state_site_tensor, out.ptr, nout * sizeof(double2));
        return 0;
    });
    if (rc) fill_zero(output_state_site_tensor, nout * sizeof(nc_zcomplex));
}

// ---------------------------------------------------------------------------
// Dense optimisation matrix (core.f95:1090-1140), built column by column with
// the grouped matvec applied to unit vectors.
// ---------------------------------------------------------------------------
void compute_optimization_matrix_(const uint32_t* bl, const uint32_t* br, const uint32_t* cl, const uint32_t* cr,
                                  const uint32_t* d, const nc_zcomplex* left_environment,
                                  const uint32_t* number_of_matrices, const uint32_t* sparse_operator_indices,
                                  const nc_zcomplex* sparse_operator_matrices, const nc_zcomplex* right_environment,
                                  nc_zcomplex* optimization_matrix) {
    const size_t n = (size_t)*br * *bl * *d;
    int rc = guarded("compute_optimization_matrix_", [&] {
        auto s = Context::get().stream;
        MpoSite op;
        build_op(op, *d, *cl, *cr, *number_of_matrices, sparse_operator_indices, sparse_operator_matrices);
        HostIn L(left_environment, (size_t)*bl * *bl * *cl);
        HostIn R(right_environment, (size_t)*br * *br * *cr);
        ZBuf e(n), M(n * n);
        for (size_t j = 0; j < n; ++j) {
            dzero(e.ptr, n * sizeof(double2));
            double2 one = make_double2(1, 0);
            NC_HIP(hipMemcpyAsync(e.ptr + j, &one, sizeof(double2), hipMemcpyHostToDevice, s));
            heff_matvec((int)*bl, (int)*br, op, L.ptr(), R.ptr(), e.ptr, M.ptr + j * n, ws(), s);
            NC_HIP(hipStreamSynchronize(s));
        }
        d2h(optimization_matrix, M.ptr, n * n * sizeof(double2));
        return 0;
    });
    if (rc) fill_zero(optimization_matrix, n * n * sizeof(nc_zcomplex));
}

// ---------------------------------------------------------------------------
// optimize_ (core.f95:1911-2047)
// ---------------------------------------------------------------------------
uint32_t optimize_(const uint32_t* bl, const uint32_t* br, const uint32_t* cl, const uint32_t* cr, const uint32_t* d,
                   const nc_zcomplex* left_environment, const uint32_t* number_of_matrices,
                   const uint32_t* sparse_operator_indices, const nc_zcomplex* sparse_operator_matrices,
                   const nc_zcomplex* right_environment, const uint32_t* number_of_projectors,
                   const uint32_t* number_of_reflectors, const uint32_t* orthogonal_subspace_dimension,
                   const nc_zcomplex* reflectors, const nc_zcomplex* coefficients, const uint32_t* swaps,
                   const char* which, const double* tol, uint32_t* number_of_iterations,
                   const nc_zcomplex* guess, nc_zcomplex* result, nc_zcomplex* eigenvalue, double* normal) {
    const size_t n = (size_t)*br * *bl * *d;
    int rc = guarded("optimize_", [&] {
        auto s = Context::get().stream;
        MpoSite op;
        build_op(op, *d, *cl, *cr, *number_of_matrices, sparse_operator_indices, sparse_operator_matrices);
        HostIn L(left_environment, (size_t)*bl * *bl * *cl);
        HostIn R(right_environment, (size_t)*br * *br * *cr);
        HostIn g(guess, n);
        ZBuf res(n);
        Projector proj;
        const int nproj = (int)*number_of_projectors;
        if (nproj > 0) {
            proj.set((int)n, nproj, (int)*number_of_reflectors, (int)*orthogonal_subspace_dimension,
                     reinterpret_cast<const zd*>(reflectors), reinterpret_cast<const zd*>(coefficients), swaps);
        } else {
            proj.n = (int)n;
            proj.orth = (int)*orthogonal_subspace_dimension;
        }
        // core.f95:1962-1983: info 5 when the orthogonal space is larger than the
        // full space, 10 when it is empty, whatever the projector count
        if (nproj == 0 && (size_t)*orthogonal_subspace_dimension > n) return 5;
        if (nproj == 0 && *orthogonal_subspace_dimension == 0) return 10;
        OptimizeArgs a{(int)*bl, (int)*br, (int)*cl, (int)*cr, (int)*d, L.ptr(), R.ptr(), &op, &proj, which, *tol,
                       (int)*number_of_iterations, g.ptr(), res.ptr};
        a.ncv = optimizer_ncv();
        OptimizeOut o = optimize_site(a, ws(), s);
        if (o.info == 5 || o.info == 10 || o.info == 11) return o.info;
        host_out(result, res.ptr, n);
        eigenvalue->re = o.eigenvalue.real();
        eigenvalue->im = o.eigenvalue.imag();
        *normal = o.normal;
        *number_of_iterations = (uint32_t)o.iterations;
        return o.info;
    });
    return (uint32_t)rc;
}

// ---------------------------------------------------------------------------
// Site move (core.f95:2708-2878).  C-side argument names as in core.cpp.
// ---------------------------------------------------------------------------
int32_t norm_denorm_going_left_(const uint32_t* bll, const uint32_t* bl, const uint32_t* br, const uint32_t* dl,
                                const uint32_t* d, const nc_zcomplex* site_tensor_to_denormalize,
                                const nc_zcomplex* site_tensor_to_normalize, nc_zcomplex* denormalized_site_tensor,
                                nc_zcomplex* normalized_site_tensor) {
    // Fortran (bl, bm, br, dl, dr) = C (bll, bl, br, dl, d)
    const size_t nd = (size_t)*bl * *bll * *dl, nn = (size_t)*br * *bl * *d;
    int rc = guarded("norm_denorm_going_left_", [&] {
        auto s = Context::get().stream;
        HostIn td(site_tensor_to_denormalize, nd), tn(site_tensor_to_normalize, nn);
        ZBuf od(nd), on(nn);
        int info = norm_denorm_left((int)*bll, (int)*bl, (int)*br, (int)*dl, (int)*d, td.ptr(), tn.ptr(), od.ptr, on.ptr, s);
        host_out(denormalized_site_tensor, od.ptr, nd);
        host_out(normalized_site_tensor, on.ptr, nn);
        return info;
    });
    return rc;
}

int32_t norm_denorm_going_right_(const uint32_t* bl, const uint32_t* br, const uint32_t* brr, const uint32_t* d,
                                 const uint32_t* dr, const nc_zcomplex* site_tensor_to_normalize,
                                 const nc_zcomplex* site_tensor_to_denormalize, nc_zcomplex* normalized_site_tensor,
                                 nc_zcomplex* denormalized_site_tensor) {
    // Fortran (bl, bm, br, dl, dr) = C (bl, br, brr, d, dr)
    const size_t nn = (size_t)*br * *bl * *d, nd = (size_t)*brr * *br * *dr;
    int rc = guarded("norm_denorm_going_right_", [&] {
        auto s = Context::get().stream;
        HostIn tn(site_tensor_to_normalize, nn), td(site_tensor_to_denormalize, nd);
        ZBuf on(nn), od(nd);
        int info = norm_denorm_right((int)*bl, (int)*br, (int)*brr, (int)*d, (int)*dr, tn.ptr(), td.ptr(), on.ptr, od.ptr, s);
        host_out(normalized_site_tensor, on.ptr, nn);
        host_out(denormalized_site_tensor, od.ptr, nd);
        return info;
    });
    return rc;
}

// ---------------------------------------------------------------------------
// Projector helpers (core.f95:1266-1339, 1866-1909)
// ---------------------------------------------------------------------------
void compute_overlap_with_projectors_(const uint32_t* number_of_projectors, const uint32_t* number_of_reflectors,
                                      const nc_zcomplex* reflectors, const nc_zcomplex* coefficients,
                                      const uint32_t* swaps, const uint32_t* vector_size, const nc_zcomplex* vector,
                                      const nc_zcomplex* overlap) {
    // the Fortran routine writes a double precision value into the caller's slot
    double* out = reinterpret_cast<double*>(const_cast<nc_zcomplex*>(overlap));
    int rc = guarded("compute_overlap_with_projectors_", [&] {
        auto s = Context::get().stream;
        if (*number_of_projectors == 0) { *out = 0.0; return 0; }
        const size_t n = *vector_size;
        Projector p;
        p.set((int)n, (int)*number_of_projectors, (int)*number_of_reflectors, (int)n,
              reinterpret_cast<const zd*>(reflectors), reinterpret_cast<const zd*>(coefficients), swaps);
        HostIn v(vector, n);
        ZBuf tmp(n);
        *out = p.overlap(v.ptr(), tmp.ptr, s);
        return 0;
    });
    if (rc) *out = 0.0;
}

void filter_components_outside_orthog_(const uint32_t* full_space_dimension, const uint32_t* number_of_projectors,
                                       const uint32_t* number_of_reflectors,
                                       const uint32_t* orthogonal_subspace_dimension, const nc_zcomplex* reflectors,
                                       const nc_zcomplex* coefficients, const uint32_t* swaps,
                                       const nc_zcomplex* input, const nc_zcomplex* output) {
    nc_zcomplex* outp = const_cast<nc_zcomplex*>(output);
    const size_t n = *full_space_dimension;
    int rc = guarded("filter_components_outside_orthog_", [&] {
        auto s = Context::get().stream;
        if (*number_of_projectors == 0) { memcpy(outp, input, n * sizeof(nc_zcomplex)); return 0; }
        Projector p;
        p.set((int)n, (int)*number_of_projectors, (int)*number_of_reflectors, (int)*orthogonal_subspace_dimension,
              reinterpret_cast<const zd*>(reflectors), reinterpret_cast<const zd*>(coefficients), swaps);
        HostIn v(input, n);
        ZBuf pr(std::max<size_t>(1, p.orth)), tmp(n), out(n);
        p.project(v.ptr(), pr.ptr, tmp.ptr, s);
        p.unproject(pr.ptr, out.ptr, s);
        d2h(outp, out.ptr, n * sizeof(double2));
        return 0;
    });
    if (rc) fill_zero(outp, n * sizeof(nc_zcomplex));
}

void nc_project_into_orthogonal_space(uint32_t n, uint32_t nproj, uint32_t nrefl, uint32_t orth,
                                      const nc_zcomplex* reflectors, const nc_zcomplex* coefficients,
                                      const uint32_t* swaps, const nc_zcomplex* in, nc_zcomplex* out) {
    int rc = guarded("nc_project_into_orthogonal_space", [&] {
        auto s = Context::get().stream;
        Projector p;
        if (nproj) p.set((int)n, (int)nproj, (int)nrefl, (int)orth, reinterpret_cast<const zd*>(reflectors),
                         reinterpret_cast<const zd*>(coefficients), swaps);
        else { p.n = (int)n; p.orth = (int)n; }
        HostIn v(in, n);
        ZBuf pr(std::max<size_t>(1, orth)), tmp(n);
        p.project(v.ptr(), pr.ptr, tmp.ptr, s);
        d2h(out, pr.ptr, (size_t)(nproj ? orth : n) * sizeof(double2));
        return 0;
    });
    if (rc) fill_zero(out, (size_t)orth * sizeof(nc_zcomplex));
}

void nc_unproject_from_orthogonal_space(uint32_t n, uint32_t nproj, uint32_t nrefl, uint32_t orth,
                                        const nc_zcomplex* reflectors, const nc_zcomplex* coefficients,
                                        const uint32_t* swaps, const nc_zcomplex* in, nc_zcomplex* out) {
    int rc = guarded("nc_unproject_from_orthogonal_space", [&] {
        auto s = Context::get().stream;
        Projector p;
        if (nproj) p.set((int)n, (int)nproj, (int)nrefl, (int)orth, reinterpret_cast<const zd*>(reflectors),
                         reinterpret_cast<const zd*>(coefficients), swaps);
        else { p.n = (int)n; p.orth = (int)n; }
        HostIn v(in, nproj ? orth : n);
        ZBuf full(n);
        p.unproject(v.ptr(), full.ptr, s);
        d2h(out, full.ptr, (size_t)n * sizeof(double2));
        return 0;
    });
    if (rc) fill_zero(out, (size_t)n * sizeof(nc_zcomplex));
}

// ---------------------------------------------------------------------------
// Overlap (projector) boundaries and vectors (core.f95:850-1005, 3299-3373)
// ---------------------------------------------------------------------------
void contract_vs_left_(const uint32_t* b_left_old, const uint32_t* b_right_old, const uint32_t* b_left_new,
                       const uint32_t* b_right_new, const uint32_t* d, const nc_zcomplex* left_environment,
                       const nc_zcomplex* normalized_projector_site_tensor,
                       const nc_zcomplex* normalized_state_site_tensor, nc_zcomplex* new_left_environment) {
    const int64_t blo = *b_left_old, bro = *b_right_old, bln = *b_left_new, brn = *b_right_new, dd = *d;
    const size_t nout = (size_t)(brn * bro);
    int rc = guarded("contract_vs_left_", [&] {
        auto s = Context::get().stream;
        HostIn L(left_environment, (size_t)(bln * blo));
        HostIn P(normalized_projector_site_tensor, (size_t)(blo * dd * bro));
        HostIn S(normalized_state_site_tensor, (size_t)(brn * bln * dd));
        ZBuf out(nout);
        vs_left(blo, bro, bln, brn, dd, L.ptr(), P.ptr(), S.ptr(), out.ptr, s);
        d2h(new_left_environment, out.ptr, nout * sizeof(double2));
        return 0;
    });
    if (rc) fill_zero(new_left_environment, nout * sizeof(nc_zcomplex));
}

void contract_vs_right_(const uint32_t* b_left_old, const uint32_t* b_right_old, const uint32_t* b_left_new,
                        const uint32_t* b_right_new, const uint32_t* d, const nc_zcomplex* right_environment,
                        const nc_zcomplex* normalized_projector_site_tensor,
                        const nc_zcomplex* normalized_state_site_tensor, nc_zcomplex* new_right_environment) {
    const int64_t blo = *b_left_old, bro = *b_right_old, bln = *b_left_new, brn = *b_right_new, dd = *d;
    const size_t nout = (size_t)(blo * bln);
    int rc = guarded("contract_vs_right_", [&] {
        auto s = Context::get().stream;
        HostIn R(right_environment, (size_t)(bro * brn));
        HostIn P(normalized_projector_site_tensor, (size_t)(blo * dd * bro));
        HostIn S(normalized_state_site_tensor, (size_t)(brn * bln * dd));
        ZBuf out(nout);
        vs_right(blo, bro, bln, brn, dd, R.ptr(), P.ptr(), S.ptr(), out.ptr, s);
        d2h(new_right_environment, out.ptr, nout * sizeof(double2));
        return 0;
    });
    if (rc) fill_zero(new_right_environment, nout * sizeof(nc_zcomplex));
}

void form_overlap_vector_(const uint32_t* b_left_old, const uint32_t* b_right_old, const uint32_t* b_left_new,
                          const uint32_t* b_right_new, const uint32_t* d, const nc_zcomplex* left_environment,
                          const nc_zcomplex* right_environment, const nc_zcomplex* unnormalized_projector_site_tensor,
                          nc_zcomplex* overlap_vector) {
    const int64_t blo = *b_left_old, bro = *b_right_old, bln = *b_left_new, brn = *b_right_new, dd = *d;
    const size_t nout = (size_t)(brn * bln * dd);
    int rc = guarded("form_overlap_vector_", [&] {
        auto s = Context::get().stream;
        HostIn L(left_environment, (size_t)(bln * blo));
        HostIn R(right_environment, (size_t)(bro * brn));
        HostIn P(unnormalized_projector_site_tensor, (size_t)(blo * dd * bro));
        ZBuf out(nout);
        nc::overlap_vector(blo, bro, bln, brn, dd, L.ptr(), R.ptr(), P.ptr(), out.ptr, s);
        d2h(overlap_vector, out.ptr, nout * sizeof(double2));
        return 0;
    });
    if (rc) fill_zero(overlap_vector, nout * sizeof(nc_zcomplex));
}

void form_overlap_site_tensor_(const uint32_t* br, const uint32_t* bl, const uint32_t* d,
                               const nc_zcomplex* state_site_tensor, nc_zcomplex* overlap_site_tensor) {
    const size_t n = (size_t)*br * *bl * *d;
    int rc = guarded("form_overlap_site_tensor_", [&] {
        auto s = Context::get().stream;
        HostIn S(state_site_tensor, n);
        ZBuf out(std::max<size_t>(1, n));
        overlap_site(*br, *bl, *d, S.ptr(), out.ptr, s);
        d2h(overlap_site_tensor, out.ptr, n * sizeof(double2));
        return 0;
    });
    if (rc) fill_zero(overlap_site_tensor, n * sizeof(nc_zcomplex));
}

void form_norm_overlap_tensors_(const uint32_t* bl, const uint32_t* bm, const uint32_t* br, const uint32_t* dl,
                                const uint32_t* dr, const nc_zcomplex* unnormalized_state_tensor_1,
                                const nc_zcomplex* right_norm_state_tensor_2, nc_zcomplex* left_norm_overlap_tensor_1,
                                nc_zcomplex* unnormalized_overlap_tensor_1, nc_zcomplex* unnormalized_state_tensor_2,
                                nc_zcomplex* right_norm_overlap_tensor_2) {
    const size_t n1 = (size_t)*bm * *bl * *dl, n2 = (size_t)*br * *bm * *dr;
    int rc = guarded("form_norm_overlap_tensors_", [&] {
        auto s = Context::get().stream;
        HostIn U1(unnormalized_state_tensor_1, n1);
        HostIn R2(right_norm_state_tensor_2, n2);
        ZBuf lo1(n1), uo1(n1), u2(n2), ro2(n2);
        int info = norm_overlap_tensors((int)*bl, (int)*bm, (int)*br, (int)*dl, (int)*dr, U1.ptr(), R2.ptr(), lo1.ptr,
                                        uo1.ptr, u2.ptr, ro2.ptr, s);
        // the Fortran routine stops the process here ("Unable to normalize tensor.")
        if (info != 0) throw Error("Unable to normalize tensor.", -55);
        d2h(left_norm_overlap_tensor_1, lo1.ptr, n1 * sizeof(double2));
        d2h(unnormalized_overlap_tensor_1, uo1.ptr, n1 * sizeof(double2));
        d2h(unnormalized_state_tensor_2, u2.ptr, n2 * sizeof(double2));
        d2h(right_norm_overlap_tensor_2, ro2.ptr, n2 * sizeof(double2));
        return 0;
    });
    if (rc) {
        fill_zero(left_norm_overlap_tensor_1, n1 * sizeof(nc_zcomplex));
        fill_zero(unnormalized_overlap_tensor_1, n1 * sizeof(nc_zcomplex));
        fill_zero(unnormalized_state_tensor_2, n2 * sizeof(nc_zcomplex));
        fill_zero(right_norm_overlap_tensor_2, n2 * sizeof(nc_zcomplex));
    }
}

uint32_t convert_vectors_to_reflectors_(const uint32_t* n, const uint32_t* m, nc_zcomplex* vectors, uint32_t* rank,
                                        nc_zcomplex* coefficients, uint32_t* swaps) {
    // n: vector length (rows), m: number of vectors (the Fortran routine's m, n)
    const int64_t rows = *n;
    const int cols = (int)*m;
    const int k = (int)std::min<int64_t>(rows, cols);
    int rc = guarded("convert_vectors_to_reflectors_", [&] {
        if (cols == 0) { *rank = 0; return 0; }
        auto s = Context::get().stream;
        HostIn V(vectors, (size_t)(rows * cols));
        std::vector<zd> tau(k);
        std::vector<uint32_t> sw(k);
        int r = reflectors_from_vectors(rows, cols, V.ptr(), tau.data(), sw.data(), s);
        d2h(vectors, V.ptr(), (size_t)(rows * cols) * sizeof(double2));
        for (int i = 0; i < k; ++i) {
            coefficients[i].re = tau[i].real();
            coefficients[i].im = tau[i].imag();
            swaps[i] = sw[i];
        }
        *rank = (uint32_t)r;
        return 0;
    });
    if (rc) *rank = 0;
    return *rank;
}

}  // extern "C"
