// Small dense eigen-problems on the host: the Arnoldi projection H (a few
// columns), and the <= 4 x 4 projected H_eff of the dense strategy.  These are
// the reference's zlahqr/ztrevc (inside ARPACK) and zheevr (core.f95:211-462)
// steps; the sizes here never exceed a few dozen.
#include "nc_internal.hpp"

#include <algorithm>
#include <cmath>
#include <limits>

namespace nc {

namespace {

// Givens rotation G = [[c, s], [-conj(s), c]] with G [x; y] = [r; 0], c real.
void givens(zd x, zd y, double& c, zd& s) {
    const double ay = std::abs(y);
    if (ay == 0.0) { c = 1.0; s = 0.0; return; }
    const double ax = std::abs(x);
    const double r = std::hypot(ax, ay);
    if (ax == 0.0) { c = 0.0; s = std::conj(y) / ay; return; }
    c = ax / r;
    s = (x / ax) * std::conj(y) / r;
}

}  // namespace

void hessenberg_eig(int m, std::vector<zd> const& H, std::vector<zd>& w, std::vector<zd>& V) {
    // column-major m x m; T <- Schur form, Z accumulates the unitary
    std::vector<zd> T(H), Z((size_t)m * m, zd(0, 0));
    auto t = [&](int i, int j) -> zd& { return T[(size_t)i + (size_t)j * m]; };
    auto z = [&](int i, int j) -> zd& { return Z[(size_t)i + (size_t)j * m]; };
    for (int i = 0; i < m; ++i) z(i, i) = 1.0;
    // only the Hessenberg part is meaningful
    for (int j = 0; j < m; ++j)
        for (int i = j + 2; i < m; ++i) t(i, j) = 0.0;
    const double eps = std::numeric_limits<double>::epsilon();
    int hi = m - 1;
    int iter = 0, total = 0;
    std::vector<double> cs(m);
    std::vector<zd> ss(m);
    while (hi > 0) {
        int l = hi;
        while (l > 0) {
            double sc = std::abs(t(l - 1, l - 1)) + std::abs(t(l, l));
            if (sc == 0.0) sc = 1.0;
            if (std::abs(t(l, l - 1)) <= eps * sc) break;
            --l;
        }
        if (l > 0) t(l, l - 1) = 0.0;
        if (l == hi) { --hi; iter = 0; continue; }
        if (++total > 100 * m) throw Error("hessenberg_eig: QR iteration did not converge", -20);
        // Wilkinson shift from the trailing 2 x 2 of the active block
        zd a = t(hi - 1, hi - 1), b = t(hi - 1, hi), c = t(hi, hi - 1), d = t(hi, hi);
        zd mu;
        if (iter > 0 && iter % 10 == 0) {
            mu = d + std::abs(c) * zd(0.75, 0.25);  // exceptional shift
        } else {
            zd half = 0.5 * (a - d);
            zd disc = std::sqrt(half * half + b * c);
            zd m1 = d - b * c / (half + disc), m2 = d - b * c / (half - disc);
            if (std::abs(half + disc) == 0.0) m1 = d;
            if (std::abs(half - disc) == 0.0) m2 = d;
            mu = (std::abs(m1 - d) < std::abs(m2 - d)) ? m1 : m2;
        }
        ++iter;
        for (int k = l; k <= hi; ++k) t(k, k) -= mu;
        for (int k = l; k < hi; ++k) {
            givens(t(k, k), t(k + 1, k), cs[k], ss[k]);
            const double cc = cs[k];
            const zd sv = ss[k];
            for (int j = k; j < m; ++j) {
                zd x = t(k, j), y = t(k + 1, j);
                t(k, j) = cc * x + sv * y;
                t(k + 1, j) = -std::conj(sv) * x + cc * y;
            }
        }
        for (int k = l; k < hi; ++k) {
            const double cc = cs[k];
            const zd sv = ss[k];
            const int rmax = std::min(k + 2, hi);
            for (int i = 0; i <= rmax; ++i) {
                zd x = t(i, k), y = t(i, k + 1);
                t(i, k) = cc * x + std::conj(sv) * y;
                t(i, k + 1) = -sv * x + cc * y;
            }
            for (int i = 0; i < m; ++i) {
                zd x = z(i, k), y = z(i, k + 1);
                z(i, k) = cc * x + std::conj(sv) * y;
                z(i, k + 1) = -sv * x + cc * y;
            }
        }
        for (int k = l; k <= hi; ++k) t(k, k) += mu;
    }
    w.resize(m);
    for (int i = 0; i < m; ++i) w[i] = t(i, i);
    // eigenvectors of the triangular T, then back to the original basis
    double tnorm = 0;
    for (auto const& v : T) tnorm = std::max(tnorm, std::abs(v));
    const double small = std::max(tnorm * eps, std::numeric_limits<double>::min());
    V.assign((size_t)m * m, zd(0, 0));
    std::vector<zd> yv(m);
    for (int k = 0; k < m; ++k) {
        std::fill(yv.begin(), yv.end(), zd(0, 0));
        yv[k] = 1.0;
        for (int j = k - 1; j >= 0; --j) {
            zd acc = 0;
            for (int i = j + 1; i <= k; ++i) acc += t(j, i) * yv[i];
            zd den = t(j, j) - w[k];
            if (std::abs(den) < small) den = small;
            yv[j] = -acc / den;
        }
        double nrm = 0;
        for (int i = 0; i < m; ++i) {
            zd acc = 0;
            for (int j = 0; j <= k; ++j) acc += z(i, j) * yv[j];
            V[(size_t)i + (size_t)k * m] = acc;
            nrm += std::norm(acc);
        }
        nrm = std::sqrt(nrm);
        for (int i = 0; i < m; ++i) V[(size_t)i + (size_t)k * m] /= nrm;
    }
}

void hermitian_eig(int m, std::vector<zd> const& Ain, std::vector<double>& w, std::vector<zd>& V) {
    // the Hermitian matrix defined by the upper triangle (zheevr with uplo = 'U')
    std::vector<zd> A((size_t)m * m);
    auto a = [&](int i, int j) -> zd& { return A[(size_t)i + (size_t)j * m]; };
    for (int j = 0; j < m; ++j)
        for (int i = 0; i <= j; ++i) {
            zd v = Ain[(size_t)i + (size_t)j * m];
            if (i == j) v = zd(v.real(), 0.0);
            a(i, j) = v;
            a(j, i) = std::conj(v);
        }
    V.assign((size_t)m * m, zd(0, 0));
    auto vv = [&](int i, int j) -> zd& { return V[(size_t)i + (size_t)j * m]; };
    for (int i = 0; i < m; ++i) vv(i, i) = 1.0;
    const double eps = std::numeric_limits<double>::epsilon();
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0, tot = 0;
        for (int j = 0; j < m; ++j)
            for (int i = 0; i < m; ++i) {
                tot += std::norm(a(i, j));
                if (i != j) off += std::norm(a(i, j));
            }
        if (off <= eps * eps * tot || off == 0.0) break;
        for (int p = 0; p < m - 1; ++p)
            for (int q = p + 1; q < m; ++q) {
                const zd apq = a(p, q);
                const double mag = std::abs(apq);
                if (mag == 0.0) continue;
                const zd ph = apq / mag;             // e^{i phi}
                const double app = a(p, p).real(), aqq = a(q, q).real();
                const double tau = (aqq - app) / (2.0 * mag);
                const double tt = (tau >= 0 ? 1.0 : -1.0) / (std::abs(tau) + std::sqrt(1.0 + tau * tau));
                const double c = 1.0 / std::sqrt(1.0 + tt * tt), s = tt * c;
                const zd em = std::conj(ph);        // e^{-i phi}
                // columns: A <- A G, G = D R with D = diag(1, e^{-i phi})
                for (int k = 0; k < m; ++k) {
                    zd xp = a(k, p), xq = a(k, q);
                    a(k, p) = c * xp - s * em * xq;
                    a(k, q) = s * xp + c * em * xq;
                    zd vp = vv(k, p), vq = vv(k, q);
                    vv(k, p) = c * vp - s * em * vq;
                    vv(k, q) = s * vp + c * em * vq;
                }
                // rows: A <- G^H A
                for (int k = 0; k < m; ++k) {
                    zd xp = a(p, k), xq = a(q, k);
                    a(p, k) = c * xp - s * ph * xq;
                    a(q, k) = s * xp + c * ph * xq;
                }
                a(p, q) = 0.0;
                a(q, p) = 0.0;
            }
    }
    std::vector<int> order(m);
    for (int i = 0; i < m; ++i) order[i] = i;
    std::sort(order.begin(), order.end(), [&](int x, int y) { return a(x, x).real() < a(y, y).real(); });
    std::vector<zd> Vs((size_t)m * m);
    w.resize(m);
    for (int k = 0; k < m; ++k) {
        w[k] = a(order[k], order[k]).real();
        for (int i = 0; i < m; ++i) Vs[(size_t)i + (size_t)k * m] = vv(i, order[k]);
    }
    V.swap(Vs);
}

}  // namespace nc
