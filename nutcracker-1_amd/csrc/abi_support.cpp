// The rest of the Fortran-core entry points that sources/common/core.cpp binds:
// random state sites and projectors, single-site normalisation, bond-dimension
// growth between two sites and around one site, expectation-boundary
// construction, and the two small contractions of the flat-vector and matrix
// helpers.  Same calling convention as abi.cpp: every scalar by pointer,
// column-major complex128 arrays, host pointers in and out, all arithmetic on
// the device.
//
// The reference draws its random tensors from gfortran's global rand() stream
// (core.f95:2617-2632, 2673-2706, 2962-2983), which has no counterpart off
// gfortran.  Here every call draws from a counter-based hash on the device,
// keyed by a per-process call counter: the stream differs, the distribution
// family (independent uniform real and imaginary parts) and every property the
// callers rely on (normalisation, orthonormal columns, rank) are the same.
#include "nc_abi.hpp"
#include "../../include/nutcracker_core_abi.h"

#include <atomic>

using namespace nc;

namespace {

// one key per random draw of this process
uint64_t next_draw_key(uint64_t salt) {
    static std::atomic<uint64_t> draws{0};
    const uint64_t k = draws.fetch_add(1) + 1;
    return 0xD1B54A32D192ED03ull ^ (k * 0x9E3779B97F4A7C15ull) ^ (salt * 0xBF58476D1CE4E5B9ull);
}

// A (rows x cols, rows >= cols) with orthonormal columns, from random entries:
// the polar factor of the transposed random matrix, transposed back.  The
// reference orthogonalises by a QR (orthogonalize_matrix_in_place,
// core.f95:196-209); any orthonormal basis of a random span is an equally valid
// draw.
void random_isometry(int64_t rows, int cols, double2* out, uint64_t salt, hipStream_t s) {
    if (rows < cols) throw Error("Not enough degrees of freedom to normalize.", -54);
    ZBuf W((size_t)cols * rows), Up((size_t)cols * rows), Pc((size_t)cols * cols);
    fill_random(W.ptr, (int64_t)cols * rows, next_draw_key(salt), s);
    const int info = polar_split(cols, rows, W.ptr, Up.ptr, Pc.ptr, s);
    if (info != 0) throw Error("random isometry: bad rank", -78);
    permute_102(Up.ptr, out, cols, rows, 1, s);  // (cols x rows) -> (rows x cols)
}

// out(i, j, k) = x(i) y(j) w(k) with conj applied to x (cx) or y (cy)
__global__ void outer3_kernel(const double2* x, const double2* y, const double2* w, int64_t b, int64_t c, int cx,
                              int cy, double2* out) {
    const int64_t total = b * b * c;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = e % b, j = (e / b) % b, k = e / (b * b);
        double2 a = x[i], bb = y[j];
        if (cx) a.y = -a.y;
        if (cy) bb.y = -bb.y;
        const double2 ab = make_double2(a.x * bb.x - a.y * bb.y, a.x * bb.y + a.y * bb.x);
        const double2 ww = w[k];
        out[e] = make_double2(ab.x * ww.x - ab.y * ww.y, ab.x * ww.y + ab.y * ww.x);
    }
}

void outer3(const nc_zcomplex* sb, const nc_zcomplex* ob, uint32_t b, uint32_t c, bool right, nc_zcomplex* out) {
    auto s = Context::get().stream;
    HostIn S(sb, b), O(ob, c);
    const size_t n = (size_t)b * b * c;
    ZBuf D(std::max<size_t>(1, n));
    const unsigned g = (unsigned)std::min<size_t>(4096, (n + 255) / 256 + 1);
    hipLaunchKernelGGL(outer3_kernel, dim3(g), dim3(256), 0, s, S.ptr(), S.ptr(), O.ptr(), (int64_t)b, (int64_t)c,
                       right ? 1 : 0, right ? 0 : 1, D.ptr);
    NC_HIP(hipGetLastError());
    d2h(out, D.ptr, n * sizeof(double2));
}

}  // namespace

extern "C" {

// ---------------------------------------------------------------------------
// Random state sites (core.f95:2634-2671)
// ---------------------------------------------------------------------------
void rand_norm_state_site_tensor_(const uint32_t* br, const uint32_t* bl, const uint32_t* d,
                                  nc_zcomplex* state_site_tensor) {
    const size_t n = (size_t)*br * *bl * *d;
    int rc = guarded("rand_norm_state_site_tensor_", [&] {
        // The Fortran routine gives workspace(br d, bl) orthonormal columns, i.e.
        // sum_{r,s} psi(r,l,s) conj(psi(r,l',s)) = delta(l,l').  That is a
        // (bl x br d) matrix A(l, (r, s)) = psi(r, l, s) with orthonormal rows:
        // the polar factor of a random one, permuted back to (br, bl, d).
        if ((int64_t)*br * *d < (int64_t)*bl) throw Error("Not enough degrees of freedom to normalize.", -54);
        auto s = Context::get().stream;
        const int64_t ncol = (int64_t)*br * *d;
        ZBuf W(std::max<size_t>(1, n)), Up(std::max<size_t>(1, n)), out(std::max<size_t>(1, n)),
            Pc((size_t)*bl * *bl);
        fill_random(W.ptr, (int64_t)n, next_draw_key(1), s);
        const int info = polar_split((int)*bl, ncol, W.ptr, Up.ptr, Pc.ptr, s);
        if (info != 0) throw Error("rand_norm_state_site_tensor: bad rank", -78);
        permute_102(Up.ptr, out.ptr, *bl, *br, *d, s);  // (bl, br, d) -> (br, bl, d)
        d2h(state_site_tensor, out.ptr, n * sizeof(double2));
        return 0;
    });
    if (rc) fill_zero(state_site_tensor, n * sizeof(nc_zcomplex));
}

void rand_unnorm_state_site_tensor_(const uint32_t* br, const uint32_t* bl, const uint32_t* d,
                                    nc_zcomplex* state_site_tensor) {
    const size_t n = (size_t)*br * *bl * *d;
    int rc = guarded("rand_unnorm_state_site_tensor_", [&] {
        auto s = Context::get().stream;
        ZBuf W(std::max<size_t>(1, n));
        fill_random(W.ptr, (int64_t)n, next_draw_key(2), s);
        const double nr = dznrm2(W.ptr, (int64_t)n, s);
        if (!(nr > 0)) throw Error("random state site of zero norm", -56);
        zscal(W.ptr, zd(1.0 / nr, 0), (int64_t)n, s);
        d2h(state_site_tensor, W.ptr, n * sizeof(double2));
        return 0;
    });
    if (rc) fill_zero(state_site_tensor, n * sizeof(nc_zcomplex));
}

// random_projector_matrix (core.f95:2673-2706): random vectors converted to
// Householder reflectors
void random_projector_matrix_(const uint32_t* projector_length, const uint32_t* number_of_projectors,
                              uint32_t* rank, nc_zcomplex* reflectors, nc_zcomplex* coefficients, uint32_t* swaps) {
    const int64_t rows = *projector_length;
    const int cols = (int)*number_of_projectors;
    const int k = (int)std::min<int64_t>(rows, cols);
    int rc = guarded("random_projector_matrix_", [&] {
        if (cols == 0) return 0;
        auto s = Context::get().stream;
        ZBuf V((size_t)(rows * cols));
        fill_random(V.ptr, rows * cols, next_draw_key(3), s);
        std::vector<zd> tau(k);
        std::vector<uint32_t> sw(k);
        const int r = reflectors_from_vectors(rows, cols, V.ptr, tau.data(), sw.data(), s);
        d2h(reflectors, V.ptr, (size_t)(rows * cols) * sizeof(double2));
        for (int i = 0; i < k; ++i) {
            coefficients[i].re = tau[i].real();
            coefficients[i].im = tau[i].imag();
            swaps[i] = sw[i];
        }
        *rank = (uint32_t)r;
        return 0;
    });
    if (rc && rank) *rank = 0;
}

// ---------------------------------------------------------------------------
// Single-site normalisation (core.f95:2880-2960): the unitary polar factor of
// the site as a (br x bl d) matrix (left) or, transposed, as (bl x br d) (right)
// ---------------------------------------------------------------------------
int32_t norm_for_left_(const uint32_t* br, const uint32_t* bl, const uint32_t* d, const nc_zcomplex* site_tensor,
                       nc_zcomplex* normalized_site_tensor) {
    const size_t n = (size_t)*br * *bl * *d;
    int rc = guarded("norm_for_left_", [&] {
        if ((int64_t)*bl * *d < (int64_t)*br) throw Error("Not enough degrees of freedom to normalize.", -54);
        auto s = Context::get().stream;
        HostIn S(site_tensor, n);
        ZBuf out(std::max<size_t>(1, n)), Pc((size_t)*br * *br);
        const int info = polar_split((int)*br, (int64_t)*bl * *d, S.ptr(), out.ptr, Pc.ptr, s);
        d2h(normalized_site_tensor, out.ptr, n * sizeof(double2));
        return info;
    });
    return rc;
}

int32_t norm_for_right_(const uint32_t* br, const uint32_t* bl, const uint32_t* d, const nc_zcomplex* site_tensor,
                        nc_zcomplex* normalized_site_tensor) {
    const size_t n = (size_t)*br * *bl * *d;
    int rc = guarded("norm_for_right_", [&] {
        if ((int64_t)*br * *d < (int64_t)*bl) throw Error("Not enough degrees of freedom to normalize.", -54);
        auto s = Context::get().stream;
        HostIn S(site_tensor, n);
        ZBuf T(std::max<size_t>(1, n)), U(std::max<size_t>(1, n)), out(std::max<size_t>(1, n)),
            Pc((size_t)*bl * *bl);
        permute_102(S.ptr(), T.ptr, *br, *bl, *d, s);  // (br, bl, d) -> (bl, br, d)
        const int info = polar_split((int)*bl, (int64_t)*br * *d, T.ptr, U.ptr, Pc.ptr, s);
        permute_102(U.ptr, out.ptr, *bl, *br, *d, s);
        d2h(normalized_site_tensor, out.ptr, n * sizeof(double2));
        return info;
    });
    return rc;
}

// ---------------------------------------------------------------------------
// Bond growth between two sites (core.f95:3090-3144): a random new_bm x bm
// isometry M widens the bond (the left site absorbs M, the right site M^H),
// then norm_denorm_going_left re-normalises the right site.  Argument roles as
// in core.cpp:548-578: site_tensor_to_normalize is the left site (bm, bl, dl),
// site_tensor_to_denormalize the right site (br, bm, dr); the left output
// (new_bm, bl, dl) goes to normalized_site_tensor and the right one
// (br, new_bm, dr) to denormalized_site_tensor, positionally as the Fortran
// routine writes them.
// ---------------------------------------------------------------------------
int32_t increase_bandwidth_between_(const uint32_t* bl, const uint32_t* bm, const uint32_t* br, const uint32_t* dl,
                                    const uint32_t* dr, const uint32_t* new_bm,
                                    const nc_zcomplex* site_tensor_to_normalize,
                                    const nc_zcomplex* site_tensor_to_denormalize,
                                    nc_zcomplex* normalized_site_tensor, nc_zcomplex* denormalized_site_tensor) {
    const int64_t BL = *bl, BM = *bm, BR = *br, DL = *dl, DR = *dr, NB = *new_bm;
    const size_t n1 = (size_t)(NB * BL * DL), n2 = (size_t)(BR * NB * DR);
    int rc = guarded("increase_bandwidth_between_", [&] {
        if (NB < BM) throw Error("the new bandwidth dimension is smaller than the old one", -77);
        auto s = Context::get().stream;
        HostIn A(site_tensor_to_normalize, (size_t)(BM * BL * DL));
        HostIn B(site_tensor_to_denormalize, (size_t)(BR * BM * DR));
        ZBuf M((size_t)(NB * BM)), A1(std::max<size_t>(1, n1)), B1(std::max<size_t>(1, n2));
        random_isometry(NB, (int)BM, M.ptr, 4, s);
        // enlarged left site (new_bm, bl, dl) = M A
        GemmPlan g1;
        g1.M = (int)NB; g1.N = (int)(BL * DL);
        g1.begin_batch(A1.ptr, NB);
        g1.add_seg(M.ptr, NB, A.ptr(), BM, (int)BM);
        zgemm_run(g1, s);
        // enlarged right site (br, new_bm, dr): B_s M^H for each physical index
        GemmPlan g2;
        g2.M = (int)BR; g2.N = (int)NB; g2.opA = OpN; g2.opB = OpC;
        for (int64_t a = 0; a < DR; ++a) {
            g2.begin_batch(B1.ptr + a * BR * NB, BR);
            g2.add_seg(B.ptr() + a * BR * BM, BR, M.ptr, NB, (int)BM);
        }
        zgemm_run(g2, s);
        ZBuf L2(std::max<size_t>(1, n1)), R2(std::max<size_t>(1, n2));
        const int info = norm_denorm_left((int)BL, (int)NB, (int)BR, (int)DL, (int)DR, A1.ptr, B1.ptr, L2.ptr, R2.ptr, s);
        d2h(normalized_site_tensor, L2.ptr, n1 * sizeof(double2));
        d2h(denormalized_site_tensor, R2.ptr, n2 * sizeof(double2));
        return info;
    });
    if (rc < 0) {
        fill_zero(normalized_site_tensor, n1 * sizeof(nc_zcomplex));
        fill_zero(denormalized_site_tensor, n2 * sizeof(nc_zcomplex));
    }
    return rc;
}

// Bond growth around one site with both environments (core.f95:3246-3297), for
// a random new_b x b isometry M:
//   new_state_s = conj(M) psi_s M^H,  new_left_k = M L_k M^H,  new_right_k = conj(M) R_k M^T
void increase_bandwidth_with_environment_(const uint32_t* b, const uint32_t* c, const uint32_t* d,
                                          const uint32_t* new_b, const nc_zcomplex* old_state_site_tensor,
                                          const nc_zcomplex* old_left_exp_environment,
                                          const nc_zcomplex* old_right_exp_environment,
                                          nc_zcomplex* new_state_site_tensor, nc_zcomplex* new_left_exp_environment,
                                          nc_zcomplex* new_right_exp_environment) {
    const int64_t B = *b, C = *c, D = *d, NB = *new_b;
    const size_t ns = (size_t)(NB * NB * D), ne = (size_t)(NB * NB * C);
    int rc = guarded("increase_bandwidth_with_environment_", [&] {
        if (NB < B) throw Error("the new bandwidth dimension is smaller than the old one", -77);
        auto s = Context::get().stream;
        HostIn S(old_state_site_tensor, (size_t)(B * B * D));
        HostIn L(old_left_exp_environment, (size_t)(B * B * C));
        HostIn R(old_right_exp_environment, (size_t)(B * B * C));
        ZBuf M((size_t)(NB * B)), Mc((size_t)(NB * B)), T((size_t)(B * NB * std::max(C, D)));
        ZBuf So(std::max<size_t>(1, ns)), Lo(std::max<size_t>(1, ne)), Ro(std::max<size_t>(1, ne));
        random_isometry(NB, (int)B, M.ptr, 5, s);
        zconj_copy(Mc.ptr, M.ptr, NB * B, s);
        // slab-wise X_k op(M): T_k (b x new_b)
        auto right_factor = [&](const double2* X, int64_t slabs, int opb) {
            GemmPlan g;
            g.M = (int)B; g.N = (int)NB; g.opA = OpN; g.opB = opb;
            for (int64_t k = 0; k < slabs; ++k) {
                g.begin_batch(T.ptr + k * B * NB, B);
                g.add_seg(X + k * B * B, B, M.ptr, NB, (int)B);
            }
            zgemm_run(g, s);
        };
        // slab-wise F T_k: out_k (new_b x new_b)
        auto left_factor = [&](const double2* F, int64_t slabs, double2* out) {
            GemmPlan g;
            g.M = (int)NB; g.N = (int)NB;
            for (int64_t k = 0; k < slabs; ++k) {
                g.begin_batch(out + k * NB * NB, NB);
                g.add_seg(F, NB, T.ptr + k * B * NB, B, (int)B);
            }
            zgemm_run(g, s);
        };
        right_factor(S.ptr(), D, OpC);  // psi_s M^H
        left_factor(Mc.ptr, D, So.ptr);  // conj(M) psi_s M^H
        right_factor(L.ptr(), C, OpC);  // L_k M^H
        left_factor(M.ptr, C, Lo.ptr);   // M L_k M^H
        right_factor(R.ptr(), C, OpT);  // R_k M^T
        left_factor(Mc.ptr, C, Ro.ptr);  // conj(M) R_k M^T
        d2h(new_state_site_tensor, So.ptr, ns * sizeof(double2));
        d2h(new_left_exp_environment, Lo.ptr, ne * sizeof(double2));
        d2h(new_right_exp_environment, Ro.ptr, ne * sizeof(double2));
        return 0;
    });
    if (rc) {
        fill_zero(new_state_site_tensor, ns * sizeof(nc_zcomplex));
        fill_zero(new_left_exp_environment, ne * sizeof(nc_zcomplex));
        fill_zero(new_right_exp_environment, ne * sizeof(nc_zcomplex));
    }
}

// ---------------------------------------------------------------------------
// Expectation boundaries (core.f95:3402-3436)
// ---------------------------------------------------------------------------
void construct_left_exp_boundary_(const uint32_t* b, const uint32_t* c, const nc_zcomplex* state_boundary,
                                  const nc_zcomplex* operator_boundary, nc_zcomplex* left_expectation_boundary) {
    // L(i, j, k) = s(i) conj(s(j)) o(k)
    int rc = guarded("construct_left_exp_boundary_", [&] {
        outer3(state_boundary, operator_boundary, *b, *c, false, left_expectation_boundary);
        return 0;
    });
    if (rc) fill_zero(left_expectation_boundary, (size_t)*b * *b * *c * sizeof(nc_zcomplex));
}

void construct_right_exp_boundary_(const uint32_t* b, const uint32_t* c, const nc_zcomplex* state_boundary,
                                   const nc_zcomplex* operator_boundary, nc_zcomplex* right_expectation_boundary) {
    // R(i, j, k) = conj(s(i)) s(j) o(k)
    int rc = guarded("construct_right_exp_boundary_", [&] {
        outer3(state_boundary, operator_boundary, *b, *c, true, right_expectation_boundary);
        return 0;
    });
    if (rc) fill_zero(right_expectation_boundary, (size_t)*b * *b * *c * sizeof(nc_zcomplex));
}

// ---------------------------------------------------------------------------
// contract_matrix_left (core.f95:1195-1234): new_L = X L X^H, X (br x bl)
// ---------------------------------------------------------------------------
void contract_matrix_left_(const uint32_t* bl, const uint32_t* br, const nc_zcomplex* left_environment,
                           const nc_zcomplex* matrix, nc_zcomplex* new_left_environment) {
    const int64_t BL = *bl, BR = *br;
    const size_t nout = (size_t)(BR * BR);
    int rc = guarded("contract_matrix_left_", [&] {
        auto s = Context::get().stream;
        HostIn L(left_environment, (size_t)(BL * BL)), X(matrix, (size_t)(BR * BL));
        ZBuf T((size_t)std::max<int64_t>(1, BR * BL)), out(std::max<size_t>(1, nout));
        GemmPlan g1;
        g1.M = (int)BR; g1.N = (int)BL;
        g1.begin_batch(T.ptr, BR);
        g1.add_seg(X.ptr(), BR, L.ptr(), BL, (int)BL);
        zgemm_run(g1, s);
        GemmPlan g2;
        g2.M = (int)BR; g2.N = (int)BR; g2.opA = OpN; g2.opB = OpC;
        g2.begin_batch(out.ptr, BR);
        g2.add_seg(T.ptr, BR, X.ptr(), BR, (int)BL);
        zgemm_run(g2, s);
        d2h(new_left_environment, out.ptr, nout * sizeof(double2));
        return 0;
    });
    if (rc) fill_zero(new_left_environment, nout * sizeof(nc_zcomplex));
}

// ---------------------------------------------------------------------------
// extend_state_vector_fragment (core.f95:1163-1193):
//   new(r, s, l) = sum_m site(r, m, s) old(m, l)
// ---------------------------------------------------------------------------
void extend_state_vector_fragment_(const uint32_t* bm, const uint32_t* br, const uint32_t* dl, const uint32_t* dr,
                                   const nc_zcomplex* old_state_vector_fragment, const nc_zcomplex* site_tensor,
                                   nc_zcomplex* new_state_vector_fragment) {
    const int64_t BM = *bm, BR = *br, DL = *dl, DR = *dr;
    const size_t nout = (size_t)(BR * DR * DL);
    int rc = guarded("extend_state_vector_fragment_", [&] {
        auto s = Context::get().stream;
        HostIn O(old_state_vector_fragment, (size_t)(BM * DL)), S(site_tensor, (size_t)(BR * BM * DR));
        ZBuf out(std::max<size_t>(1, nout));
        GemmPlan g;
        g.M = (int)BR; g.N = (int)DL;
        for (int64_t a = 0; a < DR; ++a) {
            g.begin_batch(out.ptr + a * BR, BR * DR);
            g.add_seg(S.ptr() + a * BR * BM, BR, O.ptr(), BM, (int)BM);
        }
        zgemm_run(g, s);
        d2h(new_state_vector_fragment, out.ptr, nout * sizeof(double2));
        return 0;
    });
    if (rc) fill_zero(new_state_vector_fragment, nout * sizeof(nc_zcomplex));
}

}  // extern "C"
