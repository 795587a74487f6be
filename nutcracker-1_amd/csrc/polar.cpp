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
namespace {
PolarStats g_polar_stats;

// Lower bound of the normalized singular values that the scaled iteration's
// schedule starts from: a guess (truncated MPS spectra and random sites lie
// above it), checked against the error at every step; the guaranteed bound
// 1e-16 (the perturbation below keeps every singular value above it) is
// tracked beside it and takes over when the guess is contradicted.
constexpr double kNsGuess = 1e-6;

// alpha_cap solves (a/2)(3 - a^2) = 0.1 on [1, sqrt(3)]
double ns_alpha_cap() {
    static const double cap = [] {
        double lo = 1.0, hi = std::sqrt(3.0);
        for (int k = 0; k < 80; ++k) {
            const double mid = 0.5 * (lo + hi);
            if (0.5 * mid * (3.0 - mid * mid) > 0.1) lo = mid;
            else hi = mid;
        }
        return lo;
    }();
    return cap;
}

}  // namespace

PolarStats polar_stats() { return g_polar_stats; }

int& polar_max_steps() {
    static int n = 64;
    return n;
}

int polar_split(int m, int64_t ncol, const double2* A, double2* Up, double2* Pconj, hipStream_t s) {
    // the batched iteration with one job (its schedule runs on the device)
    std::vector<PolarJob> jobs(1);
    jobs[0].m = m;
    jobs[0].n = ncol;
    jobs[0].A = A;
    jobs[0].Up = Up;
    jobs[0].Pconj = Pconj;
    polar_split_batch(jobs, s);
    return jobs[0].info;
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

// ---------------------------------------------------------------------------
// Batched forms for many chains: descriptor lists staged without a host
// synchronisation, one launch for all items.
// ---------------------------------------------------------------------------
namespace {
struct BItem {
    double2* x;
    int64_t n;   // elements (or m for the m x m kernels)
    double a, b;
};

__global__ void bsumsq_kernel(const BItem* __restrict__ it, double* __restrict__ partial) {
    const BItem d = it[blockIdx.y];
    double acc = 0;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < d.n; e += (int64_t)gridDim.x * blockDim.x) {
        const double2 v = d.x[e];
        acc += v.x * v.x + v.y * v.y;
    }
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    __shared__ double sh[4];
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) partial[(int64_t)blockIdx.y * gridDim.x + blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

__global__ void bscale_kernel(const BItem* __restrict__ it) {
    const BItem d = it[blockIdx.y];
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < d.n; e += (int64_t)gridDim.x * blockDim.x) {
        double2 v = d.x[e];
        d.x[e] = make_double2(v.x * d.a, v.y * d.a);
    }
}

// X (n x n) <- a X + b I
__global__ void bidentity_kernel(const BItem* __restrict__ it) {
    const BItem d = it[blockIdx.y];
    const int64_t total = d.n * d.n;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        double2 v = d.x[e];
        v.x *= d.a;
        v.y *= d.a;
        if (e % d.n == e / d.n) v.x += d.b;
        d.x[e] = v;
    }
}

__global__ void bnoise_kernel(const BItem* __restrict__ it) {
    const BItem d = it[blockIdx.y];
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < d.n; e += (int64_t)gridDim.x * blockDim.x) {
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
        double2 v = d.x[e];
        d.x[e] = make_double2(v.x + d.a * ur, v.y + d.a * ui);
    }
}

// X(i, i) += a for i < m, X an m x n matrix (n >= m, leading dimension m)
__global__ void bdiag_kernel(const BItem* __restrict__ it) {
    const BItem d = it[blockIdx.y];
    const int64_t m = d.n;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
        double2 v = d.x[i + i * m];
        d.x[i + i * m] = make_double2(v.x + d.a, v.y);
    }
}

__global__ void bherm_kernel(const BItem* __restrict__ it) {
    const BItem d = it[blockIdx.y];
    const int64_t m = d.n, total = m * m;
    double2* P = d.x;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = e % m, j = e / m;
        if (i > j) continue;
        const double2 a = P[i + j * m], b = P[j + i * m];
        P[i + j * m] = make_double2(0.5 * (a.x + b.x), 0.5 * (b.y - a.y));
        P[j + i * m] = make_double2(0.5 * (b.x + a.x), 0.5 * (a.y - b.y));
    }
}

// ---- device-driven batched iteration: per-job schedule state on the device
struct NsState {
    double l;     // lower bound the schedule follows (a guess until contradicted)
    double ls;    // guaranteed lower bound (from 1e-16 through the same polynomials)
    double a, b;  // this step's H = a E + b I
    int phase;    // 0 running, 1 one more step, 2 converged, 3 failed (non-finite)
    int iters;
    int live;     // 1 while the job iterates
    int just;     // 1 in the iteration the job converged (its result is copied out)
    int m;
    int _pad;
};

struct NsItem {
    double2* G;      // m x m Gram matrix / E / H
    const double2* X[2];  // the iterate in the two parity buffers
    double2* Up;     // result
    int64_t n;       // elements of X (m * ncols)
};

// E = G - I for the live jobs.  The Gram GEMM computes only the tiles on and
// above the diagonal; the elements below it are filled from the Hermitian
// symmetry (each reads an element above the diagonal, which no thread writes).
__global__ void ns_eye_kernel(const NsItem* __restrict__ it, const NsState* __restrict__ st) {
    const int q = blockIdx.y;
    if (!st[q].live) return;
    const NsItem d = it[q];
    const int64_t m = st[q].m, total = m * m;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = e % m, j = e / m;
        if (i == j) {
            d.G[e].x -= 1.0;
        } else if (i > j) {
            const double2 t = d.G[j + i * m];
            d.G[e] = make_double2(t.x, -t.y);
        }
    }
}

// partial sums of |E|^2 per job (fixed block order)
__global__ void ns_sumsq_kernel(const NsItem* __restrict__ it, const NsState* __restrict__ st,
                                double* __restrict__ partial) {
    const int q = blockIdx.y;
    double acc = 0;
    if (st[q].live) {
        const NsItem d = it[q];
        const int64_t total = (int64_t)st[q].m * st[q].m;
        for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
            const double2 v = d.G[e];
            acc += v.x * v.x + v.y * v.y;
        }
    }
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    __shared__ double sh[4];
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) partial[(int64_t)q * gridDim.x + blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

// one step of the schedule above (error, lower bound, alpha) for every live job
// (one 64-lane block per job: the partial sums are added lane-strided, then
// across the lanes, in a fixed order)
__global__ __launch_bounds__(64) void ns_decide_kernel(NsState* __restrict__ st, const double* __restrict__ partial,
                                                       int nb, double cap) {
    const int q = blockIdx.x;
    if (!st[q].live) {
        if (threadIdx.x == 0) st[q].just = 0;
        return;
    }
    double e2 = 0;
    for (int b = threadIdx.x; b < nb; b += 64) e2 += partial[(int64_t)q * nb + b];
    for (int off = 32; off > 0; off >>= 1) e2 += __shfl_xor(e2, off, 64);
    if (threadIdx.x != 0) return;
    NsState s = st[q];
    s.just = 0;
    const double err = sqrt(e2);
    if (!isfinite(err)) {
        s.phase = 3;
        s.live = 0;
    } else if (s.phase == 1 || err < 1e-15 * s.m) {
        s.phase = 2;
        s.live = 0;
        s.just = 1;
    } else {
        double l = s.l, ls = s.ls;
        if (err < 1.0) {
            l = fmax(l, sqrt(1.0 - err));
            ls = fmax(ls, sqrt(1.0 - err));
        }
        // singular values all in [l, 1] give |E|_F <= sqrt(m) (1 - l^2); a larger
        // error contradicts the guess, and the schedule continues from the
        // guaranteed bound
        if (l > ls && err > sqrt((double)s.m) * (1.0 - l * l) * (1.0 + 1e-9)) l = ls;
        double alpha = 1.0;
        if (l < 1.0 - 1e-3) alpha = fmin(cap, sqrt(3.0 / (1.0 + l + l * l)));
        const double p1 = 0.5 * alpha * (3.0 - alpha * alpha);
        s.l = fmin(fmin(1.0, 0.5 * alpha * l * (3.0 - alpha * alpha * l * l)), p1);
        // the polynomial increases up to 1 / alpha, so a bound below it maps to
        // a bound (and the largest singular value, 1, to p1)
        s.ls = fmin(fmin(1.0, 0.5 * alpha * ls * (3.0 - alpha * alpha * ls * ls)), p1);
        s.ls = fmin(s.ls, s.l);
        // H = (alpha/2)(3 - alpha^2) I - (alpha^3/2) E
        s.a = -0.5 * alpha * alpha * alpha;
        s.b = 0.5 * alpha * (3.0 - alpha * alpha);
        ++s.iters;
        if (err < 1e-7) s.phase = 1;
    }
    st[q] = s;
}

// H = a E + b I for the live jobs
__global__ void ns_h_kernel(const NsItem* __restrict__ it, const NsState* __restrict__ st) {
    const int q = blockIdx.y;
    if (!st[q].live) return;
    const NsItem d = it[q];
    const double a = st[q].a, b = st[q].b;
    const int64_t m = st[q].m, total = m * m;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        double2 v = d.G[e];
        v.x *= a;
        v.y *= a;
        if (e % m == e / m) v.x += b;
        d.G[e] = v;
    }
}

// Up = X (current parity) for the jobs that converged in this iteration
__global__ void ns_copy_kernel(const NsItem* __restrict__ it, const NsState* __restrict__ st, int parity) {
    const int q = blockIdx.y;
    if (!st[q].just) return;
    const NsItem d = it[q];
    const double2* __restrict__ X = d.X[parity];
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < d.n; e += (int64_t)gridDim.x * blockDim.x)
        d.Up[e] = X[e];
}

template <typename K>
void launch_items(K kernel, std::vector<BItem> const& items, int64_t max_elems, hipStream_t s) {
    if (items.empty() || max_elems <= 0) return;
    int slot;
    auto* d = (const BItem*)stage_descriptors(items.data(), items.size() * sizeof(BItem), s, &slot);
    const unsigned g = (unsigned)std::max<int64_t>(
        1, std::min<int64_t>(std::max<int64_t>(1, 4096 / (int64_t)items.size()), (max_elems + 255) / 256));
    hipLaunchKernelGGL(kernel, dim3(g, (unsigned)items.size()), dim3(256), 0, s, d);
    NC_HIP(hipGetLastError());
    release_descriptors(slot, s);
}
}  // namespace

std::vector<double> sumsq_batch(std::vector<const double2*> const& x, std::vector<int64_t> const& n, hipStream_t s) {
    const size_t P = x.size();
    std::vector<double> out(P, 0.0);
    if (P == 0) return out;
    std::vector<BItem> items(P);
    int64_t maxn = 1;
    for (size_t p = 0; p < P; ++p) {
        items[p] = BItem{const_cast<double2*>(x[p]), n[p], 0, 0};
        maxn = std::max(maxn, n[p]);
    }
    const unsigned g = (unsigned)std::max<int64_t>(1, std::min<int64_t>(std::max<int64_t>(1, 2048 / (int64_t)P),
                                                                         (maxn + 255) / 256));
    DevBuf<double> partial((size_t)P * g);
    int slot;
    auto* d = (const BItem*)stage_descriptors(items.data(), items.size() * sizeof(BItem), s, &slot);
    hipLaunchKernelGGL(bsumsq_kernel, dim3(g, (unsigned)P), dim3(256), 0, s, d, partial.ptr);
    NC_HIP(hipGetLastError());
    release_descriptors(slot, s);
    std::vector<double> h((size_t)P * g);
    d2h(h.data(), partial.ptr, h.size() * sizeof(double));
    for (size_t p = 0; p < P; ++p)
        for (unsigned b = 0; b < g; ++b) out[p] += h[p * g + b];
    return out;
}

void scale_batch(std::vector<double2*> const& x, std::vector<int64_t> const& n, std::vector<double> const& f,
                 hipStream_t s) {
    std::vector<BItem> items;
    int64_t maxn = 0;
    for (size_t p = 0; p < x.size(); ++p) {
        items.push_back(BItem{x[p], n[p], f[p], 0});
        maxn = std::max(maxn, n[p]);
    }
    launch_items(bscale_kernel, items, maxn, s);
}

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
// values far below round-off), and right after a bond grows it is exactly so;
// its polar factor is then not determined in those directions, and the
// iteration would stall there.  A perturbation of Frobenius norm 1e-14 ||A||
// along the leading diagonal, E = eps [I 0], floors the singular values near
// 1e-16 ||A|| and completes the null directions with the projections of the
// low-index unit vectors, as LAPACK's zgesdd completes them with the first
// basis vectors its reflectors leave (the reference's mysvd, core.f95:1-55):
// after a bond grows, the new directions are low-index basis states, not
// random ones.  A hashed perturbation a hundred times smaller keeps the floor
// where the unit vectors have no component in the null space.  The schedule
// follows the guessed bound kNsGuess and, beside it, the guaranteed bound from
// l = 1e-16; every step checks the guess against the error (all singular
// values in [l, 1] give ||X X^H - I||_F <= sqrt(m) (1 - l^2)) and falls back to
// the guaranteed bound when it is contradicted.  A well-conditioned site (a
// random one, as when a chain is reset) converges in about 17 steps, a site
// with singular values near the floor in about 45; the two-site state changes
// by ~1e-14 relative (P = A U_p^H below, kept Hermitian).
// The Gram matrix of each step gives ||X X^H - I||_F for free; being
// Hermitian, only its upper tiles are computed (a step costs 1.5 GEMMs of
// m x m x n instead of 2).  Only GEMMs of
// this library are used.
// Non-finite input is refused (info 1000, reported as a normalisation error);
// a job that does not converge in polar_max_steps() steps (64 unless set; 0
// sends every job there) or has zero norm is completed by
// the SVD (rocsolver_zgesvd), which gives the null directions an orthonormal
// basis as zgesdd does.  P = A U_p^H is Hermitian positive semi-definite; its
// Hermitian part is kept (conjugated, as the neighbour absorbs it).
void polar_split_batch(std::vector<PolarJob>& jobs, hipStream_t s) {
    const size_t J = jobs.size();
    struct St {
        ZBuf X0, X1, G;
        double l = 1e-16;
        bool last = false, done = false, ok = false;
        int iters = 0;
    };
    std::vector<St> st(J);
    std::vector<const double2*> xs;
    std::vector<int64_t> ns;
    for (auto const& j : jobs) {
        if (j.m <= 0 || j.n < j.m) throw Error("polar_split: need 0 < m <= n", -51);
        xs.push_back(j.A);
        ns.push_back((int64_t)j.m * j.n);
    }
    std::vector<double> a2 = sumsq_batch(xs, ns, s);
    std::vector<BItem> noise, diag;
    int64_t maxn = 0, maxrows = 0;
    for (size_t q = 0; q < J; ++q) {
        const double nrm = std::sqrt(a2[q]);
        jobs[q].info = 0;
        if (!std::isfinite(nrm)) { jobs[q].info = 1000; st[q].done = true; continue; }
        ++g_polar_stats.calls;
        if (!(nrm > 0)) { st[q].done = true; continue; }  // the SVD below
        const int64_t total = ns[q];
        st[q].X0.resize(total);
        st[q].X1.resize(total);
        st[q].G.resize((size_t)jobs[q].m * jobs[q].m);
        d2d(st[q].X0.ptr, jobs[q].A, (size_t)total * sizeof(double2));
        noise.push_back(BItem{st[q].X0.ptr, total, 1e-16 * nrm / std::sqrt((double)total) * std::sqrt(1.5), 0});
        diag.push_back(BItem{st[q].X0.ptr, (int64_t)jobs[q].m, 1e-14 * nrm / std::sqrt((double)jobs[q].m), 0});
        maxn = std::max(maxn, total);
        maxrows = std::max(maxrows, (int64_t)jobs[q].m);
    }
    launch_items(bdiag_kernel, diag, maxrows, s);
    launch_items(bnoise_kernel, noise, maxn, s);
    {
        std::vector<const double2*> x0;
        std::vector<int64_t> n0;
        std::vector<size_t> ids;
        for (size_t q = 0; q < J; ++q)
            if (!st[q].done) { x0.push_back(st[q].X0.ptr); n0.push_back(ns[q]); ids.push_back(q); }
        std::vector<double> s2 = sumsq_batch(x0, n0, s);
        std::vector<double2*> xv;
        std::vector<double> f;
        for (size_t k = 0; k < ids.size(); ++k) { xv.push_back(st[ids[k]].X0.ptr); f.push_back(1.0 / std::sqrt(s2[k])); }
        scale_batch(xv, n0, f, s);
    }
    // The iteration runs on the device: every job's schedule (error, alpha, the
    // lower bound l, convergence) lives in device memory and gates the job's
    // launches, and the host reads the phases only every few iterations.
    std::vector<size_t> act;
    for (size_t q = 0; q < J; ++q)
        if (!st[q].done) act.push_back(q);
    if (!act.empty()) {
        const size_t A = act.size();
        std::vector<NsItem> items(A);
        std::vector<NsState> hs(A);
        int64_t maxmm = 1, maxn = 1;
        GemmPlan gram[2], upd[2];
        DevBuf<NsState> dst(A);
        for (size_t k = 0; k < A; ++k) {
            const size_t q = act[k];
            const int m = jobs[q].m;
            items[k] = NsItem{st[q].G.ptr, {st[q].X0.ptr, st[q].X1.ptr}, jobs[q].Up, ns[q]};
            hs[k] = NsState{kNsGuess, 1e-16, 0, 0, 0, 0, 1, 0, m, 0};
            maxmm = std::max<int64_t>(maxmm, (int64_t)m * m);
            maxn = std::max<int64_t>(maxn, ns[q]);
        }
        h2d(dst.ptr, hs.data(), A * sizeof(NsState));
        const int* live0 = &dst.ptr[0].live;
        // tuning keys: the largest site of the batch and the job count
        TuneKey kg;
        kg.tag = 'g';
        kg.count = (int)A;
        for (size_t q : act) {
            kg.bl = std::max(kg.bl, jobs[q].m);
            kg.br = std::max<int>(kg.br, (int)jobs[q].n);
        }
        TuneKey ku = kg;
        ku.tag = 'u';
        for (int p = 0; p < 2; ++p) {
            gram[p].M = gram[p].N = 1; gram[p].opA = OpN; gram[p].opB = OpC;
            upd[p].M = upd[p].N = 1; upd[p].opA = OpN; upd[p].opB = OpN;
            gram[p].key = kg;
            upd[p].key = ku;
            for (size_t k = 0; k < A; ++k) {
                const size_t q = act[k];
                const int m = jobs[q].m;
                const int* lv = reinterpret_cast<const int*>(reinterpret_cast<const char*>(live0) + k * sizeof(NsState));
                double2* Xp = p == 0 ? st[q].X0.ptr : st[q].X1.ptr;
                double2* Xo = p == 0 ? st[q].X1.ptr : st[q].X0.ptr;
                gram[p].begin_batch(st[q].G.ptr, m, m, m, lv);
                gram[p].batches.back().upper = 1;  // X X^H is Hermitian
                gram[p].add_seg(Xp, m, Xp, m, (int)jobs[q].n);
                upd[p].begin_batch(Xo, m, m, (int)jobs[q].n, lv);
                upd[p].add_seg(st[q].G.ptr, m, Xp, m, m);
            }
        }
        // (a first prepare may tune the plan by running it: gram plans write only
        // G, upd[0] only X1, both scratch here; upd[1], which writes X0, then
        // finds the choice made for upd[0] -- same key -- and does not run)
        std::unique_ptr<PreparedLaunch> pg[2] = {prepare_gemm(gram[0]), prepare_gemm(gram[1])};
        std::unique_ptr<PreparedLaunch> pu[2] = {prepare_gemm(upd[0]), prepare_gemm(upd[1])};
        DevBuf<NsItem> ditems(A);
        h2d(ditems.ptr, items.data(), A * sizeof(NsItem));
        const unsigned gm = (unsigned)std::max<int64_t>(
            1, std::min<int64_t>(std::max<int64_t>(1, 4096 / (int64_t)A), (maxmm + 255) / 256));
        const unsigned gs = (unsigned)std::max<int64_t>(
            1, std::min<int64_t>(std::max<int64_t>(1, 2048 / (int64_t)A), (maxmm + 255) / 256));
        const unsigned gc = (unsigned)std::max<int64_t>(
            1, std::min<int64_t>(std::max<int64_t>(1, 4096 / (int64_t)A), (maxn + 255) / 256));
        DevBuf<double> partial((size_t)A * gs);
        const double cap = ns_alpha_cap();
        constexpr int kCheck = 6;  // iterations between host reads of the phases
        const int max_steps = polar_max_steps();
        for (int it = 0; it < max_steps; ++it) {
            const int p = it & 1;
            pg[p]->launch(s);
            hipLaunchKernelGGL(ns_eye_kernel, dim3(gm, (unsigned)A), dim3(256), 0, s, ditems.ptr, dst.ptr);
            hipLaunchKernelGGL(ns_sumsq_kernel, dim3(gs, (unsigned)A), dim3(256), 0, s, ditems.ptr, dst.ptr,
                               partial.ptr);
            hipLaunchKernelGGL(ns_decide_kernel, dim3((unsigned)A), dim3(64), 0, s, dst.ptr, partial.ptr, (int)gs,
                               cap);
            hipLaunchKernelGGL(ns_copy_kernel, dim3(gc, (unsigned)A), dim3(256), 0, s, ditems.ptr, dst.ptr, p);
            hipLaunchKernelGGL(ns_h_kernel, dim3(gm, (unsigned)A), dim3(256), 0, s, ditems.ptr, dst.ptr);
            NC_HIP(hipGetLastError());
            pu[p]->launch(s);
            if (it % kCheck == kCheck - 1 || it == max_steps - 1) {
                d2h(hs.data(), dst.ptr, A * sizeof(NsState));
                bool running = false;
                for (auto const& h : hs) running |= h.live != 0;
                if (!running) break;
            }
        }
        d2h(hs.data(), dst.ptr, A * sizeof(NsState));
        for (size_t k = 0; k < A; ++k) {
            const size_t q = act[k];
            st[q].done = true;
            st[q].ok = hs[k].phase == 2;
            st[q].iters = hs[k].iters;
        }
    }
    // P = A U_p^H (Hermitian part) for the converged ones; the SVD for the rest
    GemmPlan pg;
    pg.M = pg.N = 1; pg.opA = OpN; pg.opB = OpC;
    std::vector<BItem> herm;
    int64_t maxm = 0;
    for (size_t q = 0; q < J; ++q) {
        if (jobs[q].info == 1000) continue;
        if (!st[q].ok) {
            ++g_polar_stats.svd_fallbacks;
            jobs[q].info = polar_split_svd(jobs[q].m, jobs[q].n, jobs[q].A, jobs[q].Up, jobs[q].Pconj, s);
            continue;
        }
        g_polar_stats.iterations += st[q].iters;
        const int m = jobs[q].m;
        pg.begin_batch(jobs[q].Pconj, m, m, m);
        pg.add_seg(jobs[q].A, m, jobs[q].Up, m, (int)jobs[q].n);
        herm.push_back(BItem{jobs[q].Pconj, m, 0, 0});
        maxm = std::max<int64_t>(maxm, (int64_t)m * m);
    }
    if (!pg.batches.empty()) zgemm_run(pg, s);
    launch_items(bherm_kernel, herm, maxm, s);
}

}  // namespace nc
