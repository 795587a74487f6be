// Implicitly restarted Arnoldi (Lehoucq-Sorensen), nev = 1, exact shifts, with
// the Krylov basis resident on the device.  This is the algorithm of ARPACK's
// znaupd/zneupd as the reference drives it (core.f95:2465-2527: bmat = 'I',
// mode 1, iparam(1) = 1, info = 1 so the start vector is the caller's guess):
//   * Arnoldi steps with classical Gram-Schmidt and one DGKS correction
//     (accept when ||w|| > 0.717 ||A v||, as znaitr does);
//   * Ritz pairs of the small Hessenberg H on the host; convergence when the
//     Ritz estimate |beta e_m^T s| <= tol max(eps^(2/3), |theta|) (znconv);
//   * restart by ncv - 1 shifted QR steps with the unwanted Ritz values.
// The restart count returned plays the role of ARPACK's iparam(3).
#include "nc_internal.hpp"

#include <algorithm>
#include <cmath>
#include <limits>

namespace nc {

namespace {

int wanted_index(std::vector<zd> const& th, const char* which) {
    int best = 0;
    for (int i = 1; i < (int)th.size(); ++i) {
        if (which[0] == 'S' && which[1] == 'R') {
            if (th[i].real() < th[best].real()) best = i;
        } else if (which[0] == 'L' && which[1] == 'R') {
            if (th[i].real() > th[best].real()) best = i;
        } else {  // LM
            if (std::abs(th[i]) > std::abs(th[best])) best = i;
        }
    }
    return best;
}

// indices of the Ritz values from most to least wanted
std::vector<int> wanted_order(std::vector<zd> const& th, const char* which) {
    std::vector<int> o(th.size());
    for (size_t i = 0; i < o.size(); ++i) o[i] = (int)i;
    std::stable_sort(o.begin(), o.end(), [&](int a, int b) {
        if (which[0] == 'S' && which[1] == 'R') return th[a].real() < th[b].real();
        if (which[0] == 'L' && which[1] == 'R') return th[a].real() > th[b].real();
        return std::abs(th[a]) > std::abs(th[b]);
    });
    return o;
}

// one explicit shifted QR step on the m x m Hessenberg H, accumulating Q
void shifted_qr_step(int m, std::vector<zd>& H, zd mu, std::vector<zd>& Q) {
    auto h = [&](int i, int j) -> zd& { return H[(size_t)i + (size_t)j * m]; };
    auto q = [&](int i, int j) -> zd& { return Q[(size_t)i + (size_t)j * m]; };
    std::vector<double> cs(m);
    std::vector<zd> ss(m);
    for (int k = 0; k < m; ++k) h(k, k) -= mu;
    for (int k = 0; k < m - 1; ++k) {
        zd x = h(k, k), y = h(k + 1, k);
        double c;
        zd s;
        const double ay = std::abs(y), ax = std::abs(x);
        if (ay == 0.0) { c = 1; s = 0; }
        else if (ax == 0.0) { c = 0; s = std::conj(y) / ay; }
        else { double r = std::hypot(ax, ay); c = ax / r; s = (x / ax) * std::conj(y) / r; }
        cs[k] = c; ss[k] = s;
        for (int j = k; j < m; ++j) {
            zd a = h(k, j), b = h(k + 1, j);
            h(k, j) = c * a + s * b;
            h(k + 1, j) = -std::conj(s) * a + c * b;
        }
    }
    for (int k = 0; k < m - 1; ++k) {
        const double c = cs[k];
        const zd s = ss[k];
        for (int i = 0; i <= std::min(k + 1, m - 1); ++i) {
            zd a = h(i, k), b = h(i, k + 1);
            h(i, k) = c * a + std::conj(s) * b;
            h(i, k + 1) = -s * a + c * b;
        }
        for (int i = 0; i < m; ++i) {
            zd a = q(i, k), b = q(i, k + 1);
            q(i, k) = c * a + std::conj(s) * b;
            q(i, k + 1) = -s * a + c * b;
        }
    }
    for (int k = 0; k < m; ++k) h(k, k) += mu;
    for (int j = 0; j < m; ++j)
        for (int i = j + 2; i < m; ++i) h(i, j) = 0.0;
}

}  // namespace

IramResult iram(Operator& op, int ncv, const char* which, double tol, int max_restarts,
                const double2* guess, double2* result, hipStream_t s) {
    IramResult out;
    const int64_t n = op.n;
    if (ncv > n || ncv < 2 || n < 1) { out.info = -3; return out; }
    if (max_restarts < 1) max_restarts = 1;
    const double eps = std::numeric_limits<double>::epsilon();
    const double eps23 = std::pow(eps, 2.0 / 3.0);
    const double tol_eff = tol > 0 ? tol : eps;

    ZBuf V((size_t)n * (ncv + 1));   // columns 0..ncv-1: basis, column ncv: residual f
    ZBuf W((size_t)n * 2), Vk;
    double2* f = V.ptr + (size_t)n * ncv;
    double2* w = W.ptr;
    double2* tmp = W.ptr + n;

    // start vector: the caller's guess, normalised
    d2d(V.ptr, guess, n * sizeof(double2));
    double nrm = dznrm2(V.ptr, n, s);
    if (!(nrm > 0) || !std::isfinite(nrm)) { out.info = -9; return out; }
    zscal(V.ptr, zd(1.0 / nrm, 0), n, s);

    std::vector<zd> H((size_t)ncv * ncv, zd(0, 0));
    auto h = [&](int i, int j) -> zd& { return H[(size_t)i + (size_t)j * ncv]; };
    int k = 0;          // current factorization length
    double beta = 0.0;  // ||f||
    std::vector<zd> coef(ncv + 1), corr(ncv + 1);

    for (int iter = 1;; ++iter) {
        // ---- extend the Arnoldi factorization from length k to ncv
        for (int j = k; j < ncv; ++j) {
            double2* vj = V.ptr + (size_t)n * j;
            if (j > 0) {
                d2d(vj, f, n * sizeof(double2));
                zscal(vj, zd(1.0 / beta, 0), n, s);
                h(j, j - 1) = beta;
            }
            op.apply(vj, w, s);
            ++out.matvecs;
            const double wnorm = dznrm2(w, n, s);
            zdotc_multi(V.ptr, n, j + 1, w, n, coef.data(), s);
            zgemv_sub(V.ptr, n, j + 1, coef.data(), w, n, s);
            double rnorm = dznrm2(w, n, s);
            double ref = wnorm;
            for (int pass = 0; pass < 2 && !(rnorm > 0.717 * ref); ++pass) {
                zdotc_multi(V.ptr, n, j + 1, w, n, corr.data(), s);
                zgemv_sub(V.ptr, n, j + 1, corr.data(), w, n, s);
                for (int i = 0; i <= j; ++i) coef[i] += corr[i];
                ref = rnorm;
                rnorm = dznrm2(w, n, s);
            }
            if (!(rnorm > 0.717 * ref) && rnorm < 1e-14 * wnorm) rnorm = 0.0;  // invariant subspace
            for (int i = 0; i <= j; ++i) h(i, j) = coef[i];
            d2d(f, w, n * sizeof(double2));
            beta = rnorm;
            if (beta == 0.0 && j + 1 < ncv) {
                // exact invariant subspace found early: the Ritz values of the
                // leading block are exact, finish with a shortened factorization
                std::vector<zd> Hs((size_t)(j + 1) * (j + 1));
                for (int c = 0; c <= j; ++c)
                    for (int r = 0; r <= j; ++r) Hs[(size_t)r + (size_t)c * (j + 1)] = h(r, c);
                std::vector<zd> th, Y;
                hessenberg_eig(j + 1, Hs, th, Y);
                int wi = wanted_index(th, which);
                std::vector<zd> yv(Y.begin() + (size_t)wi * (j + 1), Y.begin() + (size_t)(wi + 1) * (j + 1));
                zgemv_n(V.ptr, n, j + 1, yv.data(), result, n, s);
                double rn = dznrm2(result, n, s);
                zscal(result, zd(1.0 / rn, 0), n, s);
                out.eigenvalue = th[wi];
                out.restarts = iter;
                out.info = 0;
                return out;
            }
        }
        // ---- Ritz values and estimates
        std::vector<zd> th, Y;
        hessenberg_eig(ncv, H, th, Y);
        const int wi = wanted_index(th, which);
        const double est = beta * std::abs(Y[(size_t)(ncv - 1) + (size_t)wi * ncv]);
        const bool converged = est <= tol_eff * std::max(eps23, std::abs(th[wi]));
        if (converged || iter >= max_restarts) {
            out.restarts = iter;
            if (!converged) { out.info = -14; return out; }
            std::vector<zd> yv(Y.begin() + (size_t)wi * ncv, Y.begin() + (size_t)(wi + 1) * ncv);
            zgemv_n(V.ptr, n, ncv, yv.data(), result, n, s);
            double rn = dznrm2(result, n, s);
            zscal(result, zd(1.0 / rn, 0), n, s);
            out.eigenvalue = th[wi];
            out.info = 0;
            return out;
        }
        // ---- implicit restart.  ARPACK keeps nev (= 1) vectors and applies the
        // ncv - 1 unwanted Ritz values as exact shifts; a larger basis keeps half
        // of its vectors (thick restart) and shifts out the least wanted half.
        const int keep = ncv <= 3 ? 1 : ncv / 2;
        std::vector<int> order = wanted_order(th, which);
        std::vector<zd> Q((size_t)ncv * ncv, zd(0, 0));
        for (int i = 0; i < ncv; ++i) Q[(size_t)i + (size_t)i * ncv] = 1.0;
        for (int t = keep; t < ncv; ++t) shifted_qr_step(ncv, H, th[order[t]], Q);
        // V_k <- V Q(:, 0:keep) ; f <- V q_keep H(keep, keep-1) + f Q(ncv-1, keep-1)
        const zd hk = h(keep, keep - 1);
        const zd qm = Q[(size_t)(ncv - 1) + (size_t)(keep - 1) * ncv];
        zgemv_n(V.ptr, n, ncv, &Q[(size_t)keep * ncv], tmp, n, s);   // V q_keep
        zaxpby(tmp, qm, f, hk, n, s);                                  // tmp = hk V q_keep + qm f
        if (Vk.n < (size_t)n * keep) Vk.resize((size_t)n * keep);
        if (keep == 1) {
            zgemv_n(V.ptr, n, ncv, Q.data(), Vk.ptr, n, s);
        } else {
            ZBuf Qd((size_t)ncv * keep);
            h2d(Qd.ptr, Q.data(), (size_t)ncv * keep * sizeof(double2));
            GemmPlan g;
            g.M = (int)n; g.N = keep; g.opA = OpN; g.opB = OpN;
            g.begin_batch(Vk.ptr, n);
            g.add_seg(V.ptr, n, Qd.ptr, ncv, ncv);
            zgemm_run(g, s);
            NC_HIP(hipStreamSynchronize(s));
        }
        d2d(V.ptr, Vk.ptr, (size_t)n * keep * sizeof(double2));
        d2d(f, tmp, n * sizeof(double2));
        std::vector<zd> Hk((size_t)keep * keep);
        for (int c = 0; c < keep; ++c)
            for (int r = 0; r < keep; ++r) Hk[(size_t)r + (size_t)c * keep] = (r <= c + 1) ? h(r, c) : zd(0, 0);
        std::fill(H.begin(), H.end(), zd(0, 0));
        for (int c = 0; c < keep; ++c)
            for (int r = 0; r < keep; ++r) h(r, c) = Hk[(size_t)r + (size_t)c * keep];
        beta = dznrm2(f, n, s);
        if (keep == 1) {
            // keep v_0 exactly unit norm against round-off drift
            double n0 = dznrm2(V.ptr, n, s);
            zscal(V.ptr, zd(1.0 / n0, 0), n, s);
        }
        k = keep;
        if (beta == 0.0) {
            // the restart landed on an invariant subspace: its Ritz pairs are exact
            std::vector<zd> th2, Y2;
            hessenberg_eig(keep, Hk, th2, Y2);
            const int w2 = wanted_index(th2, which);
            std::vector<zd> yv(Y2.begin() + (size_t)w2 * keep, Y2.begin() + (size_t)(w2 + 1) * keep);
            zgemv_n(V.ptr, n, keep, yv.data(), result, n, s);
            double rn = dznrm2(result, n, s);
            zscal(result, zd(1.0 / rn, 0), n, s);
            out.eigenvalue = th2[w2];
            out.restarts = iter;
            out.info = 0;
            return out;
        }
    }
}

}  // namespace nc
