// Site move: the polar split behind norm_denorm_going_left/right
// (core.f95:1-55, 2708-2878).  For A (m x n, m <= n) with thin SVD A = U S Vh:
//   normalized   = U Vh                      (rows orthonormal)
//   denormalized = conj(U) S U^T X = conj(U S U^H) X
// so the state across the bond is unchanged.  The SVD runs on the device
// (rocSOLVER gesvd); the three products are this library's MFMA GEMM.
#include "nc_internal.hpp"

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <cmath>
#include <mutex>

namespace nc {

namespace {
rocblas_handle solver_handle() {
    static rocblas_handle h = [] {
        rocblas_handle hh;
        if (rocblas_create_handle(&hh) != rocblas_status_success) throw Error("rocblas_create_handle failed", -50);
        return hh;
    }();
    return h;
}
}  // namespace

__global__ void scale_cols_kernel(double2* X, const double* sv, int64_t m, int64_t k) {
    const int64_t total = m * k;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const double f = sv[e / m];
        double2 v = X[e];
        X[e] = make_double2(v.x * f, v.y * f);
    }
}

int polar_split(int m, int64_t ncol, const double2* A, double2* Up, double2* Pconj, hipStream_t s) {
    if (m <= 0 || ncol < m) throw Error("polar_split: need 0 < m <= n", -51);
    rocblas_handle h = solver_handle();
    if (rocblas_set_stream(h, s) != rocblas_status_success) throw Error("rocblas_set_stream failed", -52);
    // the SVD is only given finite input: a non-finite matrix is reported as a
    // failed normalisation (NormalizationError) instead
    const double anorm = dznrm2(A, (int64_t)m * ncol, s);
    if (!std::isfinite(anorm)) return 1000;
    ZBuf Aw((size_t)m * ncol), U((size_t)m * m), Vt((size_t)m * ncol), Us((size_t)m * m);
    DevBuf<double> S((size_t)m), E((size_t)std::max(1, m));
    DevBuf<int> info(1);
    d2d(Aw.ptr, A, (size_t)m * ncol * sizeof(double2));
    rocblas_status st = rocsolver_zgesvd(h, rocblas_svect_singular, rocblas_svect_singular, m, (rocblas_int)ncol,
                                         (rocblas_double_complex*)Aw.ptr, m, S.ptr, (rocblas_double_complex*)U.ptr, m,
                                         (rocblas_double_complex*)Vt.ptr, m, E.ptr, rocblas_outofplace, info.ptr);
    if (st != rocblas_status_success) throw Error("rocsolver_zgesvd failed", -53);
    int hinfo = 0;
    d2h(&hinfo, info.ptr, sizeof(int));
    // U_p = U Vh
    {
        GemmPlan p;
        p.M = m; p.N = (int)ncol; p.opA = OpN; p.opB = OpN;
        p.begin_batch(Up, m);
        p.add_seg(U.ptr, m, Vt.ptr, m, m);
        zgemm_run(p, s);
    }
    // conj(P) = conj(U S U^H)
    d2d(Us.ptr, U.ptr, (size_t)m * m * sizeof(double2));
    unsigned g = (unsigned)std::min<int64_t>(4096, ((int64_t)m * m + 255) / 256);
    hipLaunchKernelGGL(scale_cols_kernel, dim3(g), dim3(256), 0, s, Us.ptr, S.ptr, (int64_t)m, (int64_t)m);
    NC_HIP(hipGetLastError());
    {
        GemmPlan p;
        p.M = m; p.N = m; p.opA = OpN; p.opB = OpC;
        p.begin_batch(Pconj, m);
        p.add_seg(Us.ptr, m, U.ptr, m, m);
        zgemm_run(p, s);
    }
    conj_inplace(Pconj, (int64_t)m * m, s);
    return hinfo;
}

int norm_denorm_right(int bl, int bm, int br, int dl, int dr, const double2* to_norm, const double2* to_denorm,
                      double2* norm_out, double2* denorm_out, hipStream_t s) {
    if ((int64_t)bl * dl < bm) throw Error("Not enough degrees of freedom to normalize.", -54);
    const int64_t nn = (int64_t)bl * dl, nd = (int64_t)br * dr;
    ZBuf Pc((size_t)bm * bm), W1((size_t)bm * nd), W2((size_t)bm * nd);
    int info = polar_split(bm, nn, to_norm, norm_out, Pc.ptr, s);
    permute_102(to_denorm, W1.ptr, br, bm, dr, s);  // (br, bm, dr) -> (bm, br, dr)
    GemmPlan p;
    p.M = bm; p.N = (int)nd; p.opA = OpN; p.opB = OpN;
    p.begin_batch(W2.ptr, bm);
    p.add_seg(Pc.ptr, bm, W1.ptr, bm, bm);
    zgemm_run(p, s);
    permute_102(W2.ptr, denorm_out, bm, br, dr, s);  // back to (br, bm, dr)
    return info;
}

int norm_denorm_left(int bl, int bm, int br, int dl, int dr, const double2* to_denorm, const double2* to_norm,
                     double2* denorm_out, double2* norm_out, hipStream_t s) {
    if ((int64_t)br * dr < bm) throw Error("Not enough degrees of freedom to normalize.", -54);
    const int64_t nn = (int64_t)br * dr, nd = (int64_t)bl * dl;
    ZBuf Pc((size_t)bm * bm), W1((size_t)bm * nn), W2((size_t)bm * nn);
    permute_102(to_norm, W1.ptr, br, bm, dr, s);  // (br, bm, dr) -> (bm, br, dr)
    int info = polar_split(bm, nn, W1.ptr, W2.ptr, Pc.ptr, s);
    permute_102(W2.ptr, norm_out, bm, br, dr, s);
    GemmPlan p;
    p.M = bm; p.N = (int)nd; p.opA = OpN; p.opB = OpN;
    p.begin_batch(denorm_out, bm);
    p.add_seg(Pc.ptr, bm, to_denorm, bm, bm);
    zgemm_run(p, s);
    return info;
}

}  // namespace nc
