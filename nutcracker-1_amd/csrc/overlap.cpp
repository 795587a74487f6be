// Overlap (projector) contractions and projector formation on the device
// (core.f95:850-1005, 1341-1435, 3299-3373).
//
// An excited level is optimised in the orthogonal complement of the lower
// levels.  For each lower level p the chain keeps overlap boundaries
// (the lower state's sites contracted with the current state's, left and
// right of the active site) and, at the active site, the overlap vector
//     v_p(r, l, s) = sum  L_p(l, l') O_p(l', s, r') R_p(r', r)
// whose bilinear product with the site tensor is <lower level | state>.  The
// vectors of all lower levels are turned into Householder reflectors by a
// column-pivoted QR (the algorithm of LAPACK's zgeqp3 / zlaqp2 / zlarfg); the
// optimizer then works in the space they annihilate.
#include "nc_internal.hpp"

#include <algorithm>
#include <cmath>

namespace nc {

// new L(brn, bro) = S(brn, bln*d) [L(bln, blo) P(blo, d*bro)]  (core.f95:850-896)
void vs_left(int64_t blo, int64_t bro, int64_t bln, int64_t brn, int64_t d, const double2* L, const double2* P,
             const double2* S, double2* out, hipStream_t s) {
    ZBuf inter((size_t)(bln * d * bro));
    GemmPlan p1;
    p1.M = (int)bln; p1.N = (int)(d * bro);
    p1.begin_batch(inter.ptr, bln);
    p1.add_seg(L, bln, P, blo, (int)blo);
    zgemm_run(p1, s);
    GemmPlan p2;
    p2.M = (int)brn; p2.N = (int)bro;
    p2.begin_batch(out, brn);
    p2.add_seg(S, brn, inter.ptr, bln * d, (int)(bln * d));
    zgemm_run(p2, s);
}

// new R(blo, bln) = [P(blo*d, bro) R(bro, brn)] viewed (blo, d*brn) times
// S^T(d*brn, bln)  (core.f95:898-952)
void vs_right(int64_t blo, int64_t bro, int64_t bln, int64_t brn, int64_t d, const double2* R, const double2* P,
              const double2* S, double2* out, hipStream_t s) {
    ZBuf inter((size_t)(blo * d * brn)), st((size_t)(d * brn * bln));
    GemmPlan p1;
    p1.M = (int)(blo * d); p1.N = (int)brn;
    p1.begin_batch(inter.ptr, blo * d);
    p1.add_seg(P, blo * d, R, bro, (int)bro);
    zgemm_run(p1, s);
    // st(s, r, l) = S(r, l, s): the transpose of S viewed as a (brn*bln) x d matrix
    permute_102(S, st.ptr, brn * bln, d, 1, s);
    GemmPlan p2;
    p2.M = (int)blo; p2.N = (int)bln;
    p2.begin_batch(out, blo);
    p2.add_seg(inter.ptr, blo, st.ptr, d * brn, (int)(d * brn));
    zgemm_run(p2, s);
}

// v(brn, bln, d) from L(bln, blo), R(bro, brn) and the unnormalised overlap
// site P(blo, d, bro)  (core.f95:954-1005)
void overlap_vector(int64_t blo, int64_t bro, int64_t bln, int64_t brn, int64_t d, const double2* L,
                    const double2* R, const double2* P, double2* out, hipStream_t s) {
    ZBuf inter((size_t)(blo * d * brn)), t((size_t)(bln * d * brn));
    GemmPlan p1;  // inter(blo*d, brn) = P(blo*d, bro) R(bro, brn)
    p1.M = (int)(blo * d); p1.N = (int)brn;
    p1.begin_batch(inter.ptr, blo * d);
    p1.add_seg(P, blo * d, R, bro, (int)bro);
    zgemm_run(p1, s);
    GemmPlan p2;  // t(bln, d*brn) = L(bln, blo) inter(blo, d*brn)
    p2.M = (int)bln; p2.N = (int)(d * brn);
    p2.begin_batch(t.ptr, bln);
    p2.add_seg(L, bln, inter.ptr, blo, (int)blo);
    zgemm_run(p2, s);
    // out(r, l, s) = t(l, s, r): transpose of t viewed as (bln*d) x brn
    permute_102(t.ptr, out, bln * d, brn, 1, s);
}

// O(bl, d, br) = conj(S(br, bl, d))  (core.f95:3299-3313)
void overlap_site(int64_t br, int64_t bl, int64_t d, const double2* S, double2* out, hipStream_t s) {
    permute_102(S, out, br, bl * d, 1, s);
    conj_inplace(out, br * bl * d, s);
}

// core.f95:3315-3373: left-normalise site 1 into site 2 and form the overlap
// sites of the normalised, unnormalised and right-normalised tensors.
int norm_overlap_tensors(int bl, int bm, int br, int dl, int dr, const double2* unnorm1, const double2* rightnorm2,
                         double2* left_ov1, double2* unnorm_ov1, double2* unnorm2, double2* right_ov2,
                         hipStream_t s) {
    ZBuf leftnorm1((size_t)bm * bl * dl);
    int info = norm_denorm_right(bl, bm, br, dl, dr, unnorm1, rightnorm2, leftnorm1.ptr, unnorm2, s);
    if (info != 0) return info;
    overlap_site(bm, bl, dl, unnorm1, unnorm_ov1, s);
    overlap_site(bm, bl, dl, leftnorm1.ptr, left_ov1, s);
    overlap_site(br, bm, dr, rightnorm2, right_ov2, s);
    return 0;
}

namespace {
zd read1(const double2* p) {
    double2 h;
    d2h(&h, p, sizeof(double2));
    return zd(h.x, h.y);
}
void write1(double2* p, zd v) {
    double2 h = make_double2(v.real(), v.imag());
    h2d(p, &h, sizeof(double2));
}
// zladiv-free complex reciprocal of (a - b) with b real, as in zlarfg
zd recip(zd z) { return zd(1.0, 0.0) / z; }
}  // namespace

// core.f95:1341-1435.  V (m x n, column-major, device) is overwritten by the
// Householder QR with column pivoting: reflector i stored below the diagonal
// of column i (unit leading entry implied), R on and above the diagonal.
// tau (min(m,n)) and swaps (min(m,n), 1-based) are host arrays; returns rank.
//
// Pivoting follows zlaqp2: the column of largest partial norm is swapped in
// (first on ties), partial norms are downdated and recomputed when the
// downdate loses more than half the digits; reflectors follow zlarfg
// (beta = -sign(Re alpha) |(alpha, x)|, tau = (beta - alpha) / beta) and are
// applied as H^H from the left.
int reflectors_from_vectors(int64_t m, int n, double2* V, zd* tau, uint32_t* swaps, hipStream_t s) {
    if (n <= 0) return 0;
    const int k = (int)std::min<int64_t>(m, n);
    std::vector<double> vn1(n), vn2(n);
    std::vector<int> jpvt(n);
    for (int j = 0; j < n; ++j) {
        vn1[j] = vn2[j] = dznrm2(V + (size_t)j * m, m, s);
        jpvt[j] = j;
    }
    const double tol3z = std::sqrt(1.1102230246251565e-16);  // sqrt(dlamch('E'))
    ZBuf colbuf((size_t)m);
    for (int i = 0; i < k; ++i) {
        // pivot
        int pvt = i;
        for (int j = i + 1; j < n; ++j)
            if (vn1[j] > vn1[pvt]) pvt = j;
        if (pvt != i) {
            d2d(colbuf.ptr, V + (size_t)pvt * m, (size_t)m * sizeof(double2));
            d2d(V + (size_t)pvt * m, V + (size_t)i * m, (size_t)m * sizeof(double2));
            d2d(V + (size_t)i * m, colbuf.ptr, (size_t)m * sizeof(double2));
            std::swap(jpvt[pvt], jpvt[i]);
            vn1[pvt] = vn1[i];
            vn2[pvt] = vn2[i];
        }
        // reflector for rows i..m-1 of column i
        double2* col = V + (size_t)i * m;
        const zd alpha = read1(col + i);
        const int64_t ntail = m - i - 1;
        const double xnorm = ntail > 0 ? dznrm2(col + i + 1, ntail, s) : 0.0;
        zd t(0, 0);
        if (!(xnorm == 0.0 && alpha.imag() == 0.0)) {
            const double a = std::hypot(std::hypot(alpha.real(), alpha.imag()), xnorm);
            const double beta = alpha.real() >= 0 ? -a : a;
            t = zd((beta - alpha.real()) / beta, -alpha.imag() / beta);
            if (ntail > 0) zscal(col + i + 1, recip(alpha - zd(beta, 0)), ntail, s);
            write1(col + i, zd(beta, 0));
        }
        tau[i] = t;
        // apply H^H = I - conj(tau) v v^H to columns i+1..n-1 (rows i..m-1)
        if (i + 1 < n && t != zd(0, 0)) {
            const int nrest = n - i - 1;
            std::vector<zd> h(nrest);
            if (ntail > 0) zdotc_multi(V + (size_t)(i + 1) * m + i + 1, m, nrest, col + i + 1, ntail, h.data(), s);
            for (int j = 0; j < nrest; ++j) {
                double2* cj = V + (size_t)(i + 1 + j) * m;
                // w = v^H c = c_i + sum conj(v_k) c_k = c_i + conj(h_j)
                const zd ci = read1(cj + i);
                const zd w = ci + (ntail > 0 ? std::conj(h[j]) : zd(0, 0));
                const zd f = -std::conj(t) * w;
                write1(cj + i, ci + f);
                if (ntail > 0) zaxpy(cj + i + 1, f, col + i + 1, ntail, s);
            }
        }
        // partial norm downdate (zlaqp2)
        for (int j = i + 1; j < n; ++j) {
            if (vn1[j] == 0.0) continue;
            const double rij = std::abs(read1(V + (size_t)j * m + i));
            double temp = 1.0 - (rij / vn1[j]) * (rij / vn1[j]);
            temp = std::max(temp, 0.0);
            const double temp2 = temp * (vn1[j] / vn2[j]) * (vn1[j] / vn2[j]);
            if (temp2 <= tol3z) {
                vn1[j] = (m - i - 1 > 0) ? dznrm2(V + (size_t)j * m + i + 1, m - i - 1, s) : 0.0;
                vn2[j] = vn1[j];
            } else {
                vn1[j] *= std::sqrt(temp);
            }
        }
    }
    // null reflectors (R rows of norm below the threshold) are swapped to the
    // end (core.f95:1417-1432); the threshold is the single-precision literal 1e-12
    const double thresh = (double)1e-12f;
    std::vector<double2> R((size_t)k * n);
    for (int j = 0; j < n; ++j) d2h(R.data() + (size_t)j * k, V + (size_t)j * m, (size_t)k * sizeof(double2));
    std::vector<double> norms(k);
    for (int i = 0; i < k; ++i) {
        double acc = 0;
        for (int j = i; j < n; ++j) {
            const double2 v = R[(size_t)i + (size_t)j * k];
            acc += v.x * v.x + v.y * v.y;
        }
        norms[i] = std::sqrt(acc);
        swaps[i] = (uint32_t)(i + 1);
    }
    int i = 0;
    for (; i < k; ++i) {
        if (norms[i] < thresh) {
            int j = i;
            for (;;) {
                ++j;
                if (j == k) break;
                if (norms[j] > thresh) break;
            }
            if (j == k) break;
            std::swap(norms[i], norms[j]);
            swaps[i] = (uint32_t)(j + 1);
        }
    }
    return i;
}

}  // namespace nc
