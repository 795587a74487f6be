// Site move: the polar split behind norm_denorm_going_left/right
// (core.f95:1-55, 2708-2878).  For A (m x n, m <= n) with thin SVD A = U S Vh:
//   normalized   = U Vh                      (rows orthonormal)
//   denormalized = conj(U) S U^T X = conj(U S U^H) X
// so the state across the bond is unchanged.  The SVD runs on the device
// (rocSOLVER gesvd); the three products are this library's MFMA GEMM.
#include "nc_internal.hpp"

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
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

// out(i,j) = conj of the Hermitian part of P: 0.5 (conj(P(i,j)) + P(j,i)); one
// thread per unordered pair, so the update is in place
__global__ void conj_hermitian_part_kernel(double2* P, int64_t m) {
    const int64_t total = m * m;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = e % m, j = e / m;
        if (i > j) continue;
        const double2 a = P[i + j * m], b = P[j + i * m];
        const double2 nij = make_double2(0.5 * (a.x + b.x), 0.5 * (b.y - a.y));
        const double2 nji = make_double2(0.5 * (b.x + a.x), 0.5 * (a.y - b.y));
        P[i + j * m] = nij;
        P[j + i * m] = nji;
    }
}

// X(i, j) += scale * u(i, j) with u uniform in [-1, 1) from a counter hash of
// the element index (splitmix64), so the perturbation is deterministic
__global__ void add_hash_noise_kernel(double2* X, int64_t total, double scale) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        uint64_t z = (uint64_t)e * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        const double ur = (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
        z = z * 0x9E3779B97F4A7C15ull + 0x2545F4914F6CDD1Dull;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        const double ui = (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
        double2 v = X[e];
        X[e] = make_double2(v.x + scale * ur, v.y + scale * ui);
    }
}

namespace {
PolarStats g_polar_stats;

// Scaled Newton-Schulz iteration for the unitary polar factor of a wide
// matrix X (m x n, m <= n):  X <- (alpha/2) (3 I - alpha^2 X X^H) X  for a lower
// bound l of the smallest singular value (X is first divided by its Frobenius
// norm, so every singular value lies in (0, 1]).  The cubic
// p(x) = (alpha/2) x (3 - alpha^2 x^2) peaks at p(1/alpha) = 1 and maps [l, 1]
// onto [min(p(l), p(1)), 1].  The equioscillation scaling
// alpha = sqrt(3 / (1 + l + l^2)) makes p(l) = p(1) ~ 2.6 l, which for small l
// sends the largest singular values down to the round-off level and loses
// their singular vectors; alpha is therefore capped so that p(1) >= 0.1
// (alpha = 1.698, growth 2.55 per step for small singular values instead of
// 2.60), which keeps every direction at least a tenth of ||X|| in size.  The
// bound l is also raised to sqrt(1 - ||X X^H - I||_F) whenever that is larger,
// so iterates near convergence (and nearly unitary input) take plain alpha = 1
// steps; a scaled step there would map sigma = 1 down to p(1).
// A site after optimisation is often numerically rank deficient (Schmidt
// values far below round-off); its polar factor is then not determined in
// those directions, and the iteration would stall there.  A deterministic
// perturbation of Frobenius norm 1e-14 ||A|| floors the singular values near
// 1e-16 ||A||, the schedule starts from l = 1e-16 and converges in about 45
// steps for any input; the two-site state changes by ~1e-14 relative
// (P = A U_p^H below, kept Hermitian).
// The Gram matrix of each step gives ||X X^H - I||_F for free.  Only GEMMs of
// this library are used.
bool polar_newton_schulz(int m, int64_t n, const double2* A, double2* Up, hipStream_t s, int* iters_out) {
    const double nrm = dznrm2(A, (int64_t)m * n, s);
    if (!(nrm > 0) || !std::isfinite(nrm)) return false;
    ZBuf X0((size_t)m * n), X1((size_t)m * n), G((size_t)m * m);
    d2d(X0.ptr, A, (size_t)m * n * sizeof(double2));
    const int64_t total = (int64_t)m * n;
    // uniform entries in [-1,1) + i[-1,1) have RMS sqrt(2/3)
    const double pscale = 1e-14 * nrm / std::sqrt((double)total) * std::sqrt(1.5);
    {
        unsigned gsz = (unsigned)std::min<int64_t>(4096, (total + 255) / 256);
        hipLaunchKernelGGL(add_hash_noise_kernel, dim3(gsz), dim3(256), 0, s, X0.ptr, total, pscale);
        NC_HIP(hipGetLastError());
    }
    const double nrm2 = dznrm2(X0.ptr, total, s);
    zscal(X0.ptr, zd(1.0 / nrm2, 0), total, s);
    double l = 1e-16;
    const int max_iters = 64;
    // alpha_cap solves (a/2)(3 - a^2) = 0.1 on [1, sqrt(3)]
    static const double alpha_cap = [] {
        double lo = 1.0, hi = std::sqrt(3.0);
        for (int k = 0; k < 80; ++k) {
            const double mid = 0.5 * (lo + hi);
            if (0.5 * mid * (3.0 - mid * mid) > 0.1) lo = mid;
            else hi = mid;
        }
        return lo;
    }();
    bool last = false;
    for (int it = 0; it < max_iters; ++it) {
        GemmPlan g;  // G = X X^H
        g.M = m; g.N = m; g.opA = OpN; g.opB = OpC;
        g.begin_batch(G.ptr, m);
        g.add_seg(X0.ptr, m, X0.ptr, m, (int)n);
        zgemm_run(g, s);
        add_identity(G.ptr, m, zd(-1, 0), s);              // E = X X^H - I
        const double err = dznrm2(G.ptr, (int64_t)m * m, s);
        if (!std::isfinite(err)) return false;
        if (last || err < 1e-15 * m) {
            d2d(Up, X0.ptr, (size_t)m * n * sizeof(double2));
            *iters_out = it;
            return true;
        }
        // every sigma lies in [l, 1], and sigma^2 >= 1 - ||X X^H - I||_F: the
        // measured error raises the bound, so a nearly unitary iterate takes
        // plain (alpha = 1) steps, which keep sigma = 1 fixed
        if (err < 1.0) l = std::max(l, std::sqrt(1.0 - err));
        double alpha = 1.0;
        if (l < 1.0 - 1e-3) alpha = std::min(alpha_cap, std::sqrt(3.0 / (1.0 + l + l * l)));
        // H = (alpha/2)(3 - alpha^2) I - (alpha^3/2) E
        zscal(G.ptr, zd(-0.5 * alpha * alpha * alpha, 0), (int64_t)m * m, s);
        add_identity(G.ptr, m, zd(0.5 * alpha * (3.0 - alpha * alpha), 0), s);
        GemmPlan u;  // X1 = H X0
        u.M = m; u.N = (int)n; u.opA = OpN; u.opB = OpN;
        u.begin_batch(X1.ptr, m);
        u.add_seg(G.ptr, m, X0.ptr, m, m);
        zgemm_run(u, s);
        std::swap(X0.ptr, X1.ptr);
        const double p1 = 0.5 * alpha * (3.0 - alpha * alpha);
        l = std::min({1.0, 0.5 * alpha * l * (3.0 - alpha * alpha * l * l), p1});
        // cubic convergence: once every |sigma^2 - 1| < 1e-7 one more step reaches round-off
        if (err < 1e-7) last = true;
    }
    return false;
}
}  // namespace

PolarStats polar_stats() { return g_polar_stats; }

int polar_split(int m, int64_t ncol, const double2* A, double2* Up, double2* Pconj, hipStream_t s) {
    if (m <= 0 || ncol < m) throw Error("polar_split: need 0 < m <= n", -51);
    // the SVD is only given finite input: a non-finite matrix is reported as a
    // failed normalisation (NormalizationError) instead
    const double anorm = dznrm2(A, (int64_t)m * ncol, s);
    if (!std::isfinite(anorm)) return 1000;
    ++g_polar_stats.calls;
    int iters = 0;
    if (polar_newton_schulz(m, ncol, A, Up, s, &iters)) {
        g_polar_stats.iterations += iters;
        // P = A U_p^H is Hermitian positive semi-definite; keep its Hermitian part
        GemmPlan p;
        p.M = m; p.N = m; p.opA = OpN; p.opB = OpC;
        p.begin_batch(Pconj, m);
        p.add_seg(A, m, Up, m, (int)ncol);
        zgemm_run(p, s);
        unsigned gsz = (unsigned)std::min<int64_t>(4096, ((int64_t)m * m + 255) / 256);
        hipLaunchKernelGGL(conj_hermitian_part_kernel, dim3(gsz), dim3(256), 0, s, Pconj, (int64_t)m);
        NC_HIP(hipGetLastError());
        return 0;
    }
    // rank-deficient (or numerically so) input: the SVD completes the polar
    // factor with an orthonormal basis of the null directions, as zgesdd does
    ++g_polar_stats.svd_fallbacks;
    return polar_split_svd(m, ncol, A, Up, Pconj, s);
}

int polar_split_svd(int m, int64_t ncol, const double2* A, double2* Up, double2* Pconj, hipStream_t s) {
    rocblas_handle h = solver_handle();
    if (rocblas_set_stream(h, s) != rocblas_status_success) throw Error("rocblas_set_stream failed", -52);
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
