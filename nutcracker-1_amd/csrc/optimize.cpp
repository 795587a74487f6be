// Site optimizer: the reference's optimize dispatcher and its three strategies
// (core.f95:1911-2609), with every vector operation on the device, plus the
// Householder projector algebra used for excited states (core.f95:464-504,
// 1266-1339, 1513-1664).
#include "nc_internal.hpp"

#include <atomic>
#include <chrono>

#include <algorithm>
#include <cmath>
#include <cstring>

namespace nc {

// ----------------------------------------------------------------------------
// Projectors
// ----------------------------------------------------------------------------
__global__ void swap_entries_kernel(double2* x, const uint32_t* swaps, int nsw, int forward) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    if (forward) {
        for (int i = 0; i < nsw; ++i) {
            int s = (int)swaps[i] - 1;
            if (s != i) { double2 t = x[i]; x[i] = x[s]; x[s] = t; }
        }
    } else {
        for (int i = nsw - 1; i >= 0; --i) {
            int s = (int)swaps[i] - 1;
            if (s != i) { double2 t = x[i]; x[i] = x[s]; x[s] = t; }
        }
    }
}

// explicit reflector vectors: v_i (zeros above i, 1 at i, stored part below) and conj(v_i)
__global__ void build_reflectors_kernel(const double2* refl, double2* V, double2* Vc, int64_t n, int nrefl) {
    const int64_t total = n * nrefl;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = e % n;
        const int i = (int)(e / n);
        double2 v;
        if (j < i) v = make_double2(0, 0);
        else if (j == i) v = make_double2(1, 0);
        else v = refl[e];
        V[e] = v;
        Vc[e] = make_double2(v.x, -v.y);
    }
}

struct ProjectorDevice {
    ZBuf V, Vc;           // nrefl explicit reflectors
    ZBuf partial;         // per-block dot products of one reflector application
    DevBuf<uint32_t> swaps;
    std::vector<zd> Rtop;  // nrefl x nrefl upper block of the stored reflectors
};

static ProjectorDevice& proj_dev(Projector const& p) {
    if (!p.dev) throw Error("projector used before set()", -30);
    return *p.dev;
}

void Projector::set(int n_, int nproj_, int nrefl_, int orth_, const zd* refl, const zd* coeffs_,
                    const uint32_t* swaps_) {
    n = n_; nproj = nproj_; nrefl = nrefl_; orth = orth_;
    if (nrefl > nproj || nrefl > n) throw Error("projector: more reflectors than vectors", -31);
    swaps.assign(swaps_, swaps_ + nrefl);
    coeffs.assign(coeffs_, coeffs_ + nrefl);
    for (auto sw : swaps)
        if (sw < 1 || (int)sw > nrefl) throw Error("projector: swap index out of range", -32);
    d_refl.resize((size_t)n * std::max(1, nproj));
    if (nproj) h2d(d_refl.ptr, refl, (size_t)n * nproj * sizeof(double2));
    std::vector<zd> rtop((size_t)nrefl * nrefl);
    for (int j = 0; j < nrefl; ++j)
        for (int i = 0; i < nrefl; ++i) rtop[(size_t)i + (size_t)j * nrefl] = refl[(size_t)i + (size_t)j * n];
    finish(rtop, Context::get().stream);
}

void Projector::set_device(int n_, int nproj_, int nrefl_, int orth_, const double2* d_refl_in, const zd* coeffs_,
                           const uint32_t* swaps_) {
    n = n_; nproj = nproj_; nrefl = nrefl_; orth = orth_;
    if (nrefl > nproj || nrefl > n) throw Error("projector: more reflectors than vectors", -31);
    swaps.assign(swaps_, swaps_ + nrefl);
    coeffs.assign(coeffs_, coeffs_ + nrefl);
    for (auto sw : swaps)
        if (sw < 1 || (int)sw > nrefl) throw Error("projector: swap index out of range", -32);
    d_refl.resize((size_t)n * std::max(1, nproj));
    if (nproj) d2d(d_refl.ptr, d_refl_in, (size_t)n * nproj * sizeof(double2));
    std::vector<zd> rtop((size_t)nrefl * nrefl);
    for (int j = 0; j < nrefl; ++j) {
        std::vector<double2> col(nrefl);
        d2h(col.data(), d_refl.ptr + (size_t)j * n, (size_t)nrefl * sizeof(double2));
        for (int i = 0; i < nrefl; ++i) rtop[(size_t)i + (size_t)j * nrefl] = zd(col[i].x, col[i].y);
    }
    finish(rtop, Context::get().stream);
}

void Projector::finish(const std::vector<zd>& rtop, hipStream_t s) {
    dev = std::make_shared<ProjectorDevice>();
    if (!nrefl) return;
    dev->V.resize((size_t)n * nrefl);
    dev->Vc.resize((size_t)n * nrefl);
    dev->partial.resize(refl_partials());
    unsigned g = (unsigned)std::min<int64_t>(4096, ((int64_t)n * nrefl + 255) / 256);
    hipLaunchKernelGGL(build_reflectors_kernel, dim3(g), dim3(256), 0, s, d_refl.ptr, dev->V.ptr, dev->Vc.ptr,
                       (int64_t)n, nrefl);
    NC_HIP(hipGetLastError());
    dev->swaps.resize(nrefl);
    h2d(dev->swaps.ptr, swaps.data(), nrefl * sizeof(uint32_t));
    dev->Rtop = rtop;
}

// x <- x H_0 H_1 ... (forward, zunmqr 'R','N') or x <- x H_{k-1}^H ... H_0^H ('R','C');
// two launches per reflector and no host read, so a projected matvec stays
// stream-ordered
static void apply_reflectors_row(Projector const& p, double2* x, bool conj_trans, hipStream_t s) {
    ProjectorDevice& dv = proj_dev(p);
    const int64_t n = p.n;
    for (int t = 0; t < p.nrefl; ++t) {
        const int i = conj_trans ? p.nrefl - 1 - t : t;
        const zd tau = conj_trans ? std::conj(p.coeffs[i]) : p.coeffs[i];
        if (tau == zd(0, 0)) continue;
        // x H = x - tau (x v) v^H
        zrefl_apply(x, dv.V.ptr + (size_t)i * n, dv.Vc.ptr + (size_t)i * n, tau, n, dv.partial.ptr, s);
    }
}

void Projector::project(const double2* full, double2* out, double2* tmp, hipStream_t s) const {
    if (!active()) { d2d(out, full, (size_t)n * sizeof(double2)); return; }
    d2d(tmp, full, (size_t)n * sizeof(double2));
    apply_reflectors_row(*this, tmp, false, s);
    if (nrefl) {
        hipLaunchKernelGGL(swap_entries_kernel, dim3(1), dim3(1), 0, s, tmp, proj_dev(*this).swaps.ptr, nrefl, 1);
        NC_HIP(hipGetLastError());
    }
    d2d(out, tmp + (n - orth), (size_t)orth * sizeof(double2));
}

void Projector::unproject(const double2* in, double2* full, hipStream_t s) const {
    if (!active()) { d2d(full, in, (size_t)n * sizeof(double2)); return; }
    if (n - orth > 0) dzero(full, (size_t)(n - orth) * sizeof(double2));
    d2d(full + (n - orth), in, (size_t)orth * sizeof(double2));
    if (nrefl) {
        hipLaunchKernelGGL(swap_entries_kernel, dim3(1), dim3(1), 0, s, full, proj_dev(*this).swaps.ptr, nrefl, 0);
        NC_HIP(hipGetLastError());
    }
    apply_reflectors_row(*this, full, true, s);
}

double Projector::overlap(const double2* v, double2* tmp, hipStream_t s) const {
    if (!active()) return 0.0;
    // compute_overlap_with_projectors (core.f95:1266-1339)
    d2d(tmp, v, (size_t)n * sizeof(double2));
    apply_reflectors_row(*this, tmp, false, s);
    if (nrefl) {
        auto& dv = proj_dev(*this);
        hipLaunchKernelGGL(swap_entries_kernel, dim3(1), dim3(1), 0, s, tmp, dv.swaps.ptr, nrefl, 1);
        hipLaunchKernelGGL(swap_entries_kernel, dim3(1), dim3(1), 0, s, tmp, dv.swaps.ptr, nrefl, 0);
        NC_HIP(hipGetLastError());
    }
    std::vector<zd> head(std::max(1, nrefl));
    if (nrefl) {
        std::vector<double2> h(nrefl);
        d2h(h.data(), tmp, nrefl * sizeof(double2));
        for (int i = 0; i < nrefl; ++i) head[i] = zd(h[i].x, h[i].y);
    }
    double acc = 0;
    if (nrefl) {
        auto& R = proj_dev(*this).Rtop;
        for (int j = 0; j < nrefl; ++j) {  // w_j = sum_{i <= j} R(i, j) head_i  (ztrmv 'U','T')
            zd wj = 0;
            for (int i = 0; i <= j; ++i) wj += R[(size_t)i + (size_t)j * nrefl] * head[i];
            acc += std::norm(wj);
        }
    }
    for (int j = nrefl; j < nproj; ++j) {
        zd wj = zdotu(d_refl.ptr + (size_t)j * n, tmp, n, s);
        acc += std::norm(wj);
    }
    return std::sqrt(acc);
}

// ----------------------------------------------------------------------------
// operate_on (core.f95:2529-2556): unproject -> H_eff -> project
// ----------------------------------------------------------------------------
namespace {
struct HeffOperator : Operator {
    OptimizeArgs const& a;
    Workspace& ws;
    ZBuf full_in, full_out, tmp;
    HeffOperator(OptimizeArgs const& a_, Workspace& ws_) : a(a_), ws(ws_) {
        const int64_t nfull = (int64_t)a.d * a.bl * a.br;
        n = a.proj->active() ? a.proj->orth : nfull;
        if (a.proj->active()) {
            full_in.resize(nfull);
            full_out.resize(nfull);
            tmp.resize(nfull);
        }
    }
    void apply(const double2* x, double2* y, hipStream_t s) override {
        if (!a.proj->active()) {
            heff_matvec(a.bl, a.br, *a.op, a.L, a.R, x, y, ws, s);
            return;
        }
        a.proj->unproject(x, full_in.ptr, s);
        heff_matvec(a.bl, a.br, *a.op, a.L, a.R, full_in.ptr, full_out.ptr, ws, s);
        a.proj->project(full_out.ptr, y, tmp.ptr, s);
    }
};

// The same operator with the contraction in complex64 (bf16 split products or
// FP32 matrix cores); the vectors handed to the eigen-solver (and the
// projectors) stay FP64.  With `planes` the matvec runs on operands split once
// into bf16 planes (L and R per site solve, x per application).
struct Heff32Operator : Operator {
    OptimizeArgs const& a;
    Workspace& ws;
    PlaneHeff* planes;
    ZBuf full_in, full_out, tmp;
    DevBuf<float2> x32, y32;
    Heff32Operator(OptimizeArgs const& a_, Workspace& ws_, PlaneHeff* planes_ = nullptr)
        : a(a_), ws(ws_), planes(planes_) {
        const int64_t nfull = (int64_t)a.d * a.bl * a.br;
        n = a.proj->active() ? a.proj->orth : nfull;
        if (!planes) {
            x32.resize(nfull);
            y32.resize(nfull);
        }
        if (a.proj->active()) {
            full_in.resize(nfull);
            full_out.resize(nfull);
            tmp.resize(nfull);
        }
    }
    void apply(const double2* x, double2* y, hipStream_t s) override {
        const int64_t nfull = (int64_t)a.d * a.bl * a.br;
        const double2* xin = x;
        if (a.proj->active()) {
            a.proj->unproject(x, full_in.ptr, s);
            xin = full_in.ptr;
        }
        if (planes) {
            double2* yo = a.proj->active() ? full_out.ptr : y;
            planes->apply(xin, yo, s);
            if (a.proj->active()) a.proj->project(full_out.ptr, y, tmp.ptr, s);
            return;
        }
        to_c64(xin, x32.ptr, nfull, s);
        heff_matvec(a.bl, a.br, *a.op, a.L32, a.R32, x32.ptr, y32.ptr, ws, s);
        if (a.proj->active()) {
            to_c128(y32.ptr, full_out.ptr, nfull, s);
            a.proj->project(full_out.ptr, y, tmp.ptr, s);
        } else {
            to_c128(y32.ptr, y, nfull, s);
        }
    }
};
}  // namespace

// strategy 1 (core.f95:2049-2129): dense projected H_eff and a Hermitian eigen-solve
static void strategy_1(OptimizeArgs const& a, HeffOperator& op, OptimizeOut& out, hipStream_t s) {
    const int m = (int)op.n;
    ZBuf e((size_t)m), col((size_t)m), pv((size_t)m);
    std::vector<zd> H((size_t)m * m);
    std::vector<double2> hcol(m);
    for (int j = 0; j < m; ++j) {
        dzero(e.ptr, (size_t)m * sizeof(double2));
        double2 one = make_double2(1, 0);
        NC_HIP(hipMemcpyAsync(e.ptr + j, &one, sizeof(double2), hipMemcpyHostToDevice, s));
        op.apply(e.ptr, col.ptr, s);
        ++out.matvecs;
        d2h(hcol.data(), col.ptr, (size_t)m * sizeof(double2));
        for (int i = 0; i < m; ++i) H[(size_t)i + (size_t)j * m] = zd(hcol[i].x, hcol[i].y);
    }
    std::vector<double> w;
    std::vector<zd> V;
    const std::vector<zd> Hfull = H;
    hermitian_eig(m, H, w, V);
    int k;
    const char* which = a.which;
    if (which[0] == 'S' && which[1] == 'R') k = 0;
    else if (which[0] == 'L' && which[1] == 'R') k = m - 1;
    else if (which[0] == 'L' && which[1] == 'M') k = (std::abs(w[m - 1]) > std::abs(w[0])) ? m - 1 : 0;
    else throw Error(std::string("Invalid eigenvalue requested: ") + std::string(which, 2), -40);
    std::vector<double2> hv(m);
    for (int i = 0; i < m; ++i) hv[i] = make_double2(V[(size_t)i + (size_t)k * m].real(), V[(size_t)i + (size_t)k * m].imag());
    h2d(pv.ptr, hv.data(), (size_t)m * sizeof(double2));
    a.proj->unproject(pv.ptr, a.result, s);
    // the eigen-solve reads one triangle; the Rayleigh quotient of the chosen
    // vector against the whole matrix equals w[k] to round-off when H_eff is
    // Hermitian, and stays consistent with the expectation value when it is only
    // Hermitian to FP32 round-off (complex64 environments)
    zd rq = 0;
    for (int i = 0; i < m; ++i) {
        zd row = 0;
        for (int j = 0; j < m; ++j) row += Hfull[(size_t)i + (size_t)j * m] * V[(size_t)j + (size_t)k * m];
        rq += std::conj(V[(size_t)i + (size_t)k * m]) * row;
    }
    out.eigenvalue = rq;
    out.info = 0;
    out.iterations = 0;
}

namespace {
std::atomic<int64_t> g_opt_stats[5];
}

int& optimizer_ncv() {
    static int ncv = 3;
    return ncv;
}

OptimizerStats optimizer_stats() {
    OptimizerStats st;
    st.solves = g_opt_stats[0].load();
    st.first_run_failures = g_opt_stats[1].load();
    st.ncv_retries = g_opt_stats[2].load();
    st.dense_fallbacks = g_opt_stats[3].load();
    st.unconverged = g_opt_stats[4].load();
    return st;
}

OptimizeOut optimize_site(OptimizeArgs const& a, Workspace& ws, hipStream_t s) {
    OptimizeOut out;
    g_opt_stats[0].fetch_add(1);
    const int64_t nfull = (int64_t)a.d * a.bl * a.br;
    const int orth = a.proj->active() ? a.proj->orth : (int)nfull;
    out.iterations = a.max_iterations;
    if (orth > nfull) { out.info = 5; return out; }
    double overlap = 0.0;
    if (a.proj->active()) {
        ZBuf tmp((size_t)nfull);
        overlap = a.proj->overlap(a.guess, tmp.ptr, s);
    }
    if (orth <= 0) { out.info = 10; return out; }
    if (orth < nfull && overlap > 1e-7) { out.info = 11; return out; }

    HeffOperator op(a, ws);
    if (orth <= 4) {
        out.strategy = 1;
        strategy_1(a, op, out, s);
    } else {
        out.strategy = (a.bl * a.br < a.cl * a.cr) ? 2 : 3;
        ZBuf pguess((size_t)orth), pres((size_t)orth), tmp((size_t)nfull);
        a.proj->project(a.guess, pguess.ptr, tmp.ptr, s);
        const bool mixed = a.mixed && a.L32 && a.R32;
        Heff32Operator* op32 = nullptr;
        std::unique_ptr<Heff32Operator> op32_owner;
        std::unique_ptr<PlaneHeff> planes;
        if (mixed) {
            if (use_planes(a.bl, a.br)) {
                planes.reset(new PlaneHeff(a.bl, a.br, *a.op, a.L, a.R, s));
                // the Krylov solver applies the operator only to its basis
                // vectors: unit norm, or at most 1 / 0.8 for a vector whose second
                // Gram-Schmidt update is deferred (iram.cpp, kupdate2_kernel);
                // projector unprojection keeps the norm
                planes->assume_input_norm(1.3, s);
            }
            op32_owner.reset(new Heff32Operator(a, ws, planes.get()));
            op32 = op32_owner.get();
        }
        Operator& kop = mixed ? static_cast<Operator&>(*op32) : static_cast<Operator&>(op);
        const double ktol = mixed ? a.tol32 : a.tol;
        auto count = [&](IramResult const& rr) {
            if (mixed) out.matvecs32 += rr.matvecs;
            else out.matvecs += rr.matvecs;
        };
        // a Krylov basis cannot be larger than the (projected) space
        int ncv = std::min(a.ncv, std::min(orth, kMaxNcv));
        int iterations = a.max_iterations;
        IramResult r = iram(kop, ncv, a.which, ktol, iterations, pguess.ptr, pres.ptr, s);
        count(r);
        iterations = r.restarts;
        bool done = false;
        if (r.info == -14) {
            g_opt_stats[1].fetch_add(1);
            if (nfull < iterations) {
                g_opt_stats[3].fetch_add(1);
                strategy_1(a, op, out, s);
                done = true;
            } else {
                // core.f95:2588-2602: grow ncv by 3/2 while it does not exceed
                // the restart count.  The basis is capped at the projected space
                // and at the solver's largest basis; once the cap is reached
                // without convergence the site keeps info -14.
                const int cap = std::min(orth, kMaxNcv);
                while (ncv <= iterations && ncv < cap) {
                    ncv = std::min(std::max(ncv * 3 / 2, ncv + 1), cap);
                    g_opt_stats[2].fetch_add(1);
                    r = iram(kop, ncv, a.which, ktol, iterations, pguess.ptr, pres.ptr, s);
                    count(r);
                    iterations = r.restarts;
                    if (r.info != -14) break;
                }
            }
        }
        if (!done && mixed && r.info == 0) {
            if (a.polish) {
                // continue in FP64 from the complex64 solution
                ZBuf start((size_t)orth);
                d2d(start.ptr, pres.ptr, (size_t)orth * sizeof(double2));
                IramResult r2 = iram(op, std::min(a.ncv, std::min(orth, kMaxNcv)), a.which, a.tol, a.max_iterations,
                                     start.ptr, pres.ptr, s);
                out.matvecs += r2.matvecs;
                if (r2.info == 0) r = r2;
            } else {
                // FP64 Rayleigh quotient of the complex64 solution (one FP64 matvec)
                ZBuf y((size_t)orth);
                op.apply(pres.ptr, y.ptr, s);
                ++out.matvecs;
                zd num;
                zdotc_multi(pres.ptr, orth, 1, y.ptr, orth, &num, s);
                const double nn = dznrm2(pres.ptr, orth, s);
                r.eigenvalue = num / (nn * nn);
            }
        }
        if (!done) {
            if (r.info == -14) g_opt_stats[4].fetch_add(1);
            out.info = r.info;
            out.iterations = iterations;
            out.eigenvalue = r.eigenvalue;
            if (r.info == 0) a.proj->unproject(pres.ptr, a.result, s);
            else dzero(a.result, (size_t)nfull * sizeof(double2));
        }
    }
    out.normal = dznrm2(a.result, nfull, s);
    return out;
}

}  // namespace nc

namespace nc {

// ----------------------------------------------------------------------------
// Site solves of several chains in lockstep.  Every Arnoldi step applies all
// active H_eff with one grouped GEMM launch for step A (every chain's (from, a)
// products, unequal sizes), one mix launch and one grouped launch for step C;
// the orthogonalisation of all chains shares its launches too (iram_batch).
// Sites with projectors or a dense-solve size (orth <= 4) are solved one by one.
// ----------------------------------------------------------------------------
namespace {
// shape key of a grouped plan: the largest bond dimensions and the chain count
TuneKey group_key(int tag, std::vector<OptimizeArgs const*> const& as) {
    TuneKey k;
    k.tag = tag;
    k.count = (int)as.size();
    for (auto const* a : as) {
        k.bl = std::max(k.bl, a->bl);
        k.br = std::max(k.br, a->br);
        k.d = std::max(k.d, a->d);
        k.cl = std::max(k.cl, a->cl);
        k.cr = std::max(k.cr, a->cr);
    }
    return k;
}

// Grouped H_eff matvec plans of several chains, y_q = H_q x_q (complex128):
// step A (X_a L_f), the block mixing, step C (sum_t R_t Q_t).  live[q], when
// given, gates chain q at launch time.
void heff_plans64(std::vector<OptimizeArgs const*> const& as, std::vector<Workspace*> const& ws,
                  std::vector<const double2*> const& xs, std::vector<double2*> const& ys, const int* live,
                  GemmPlan& A, std::vector<MixItem<double2>>& mix, GemmPlan& Cp) {
    A.M = A.N = Cp.M = Cp.N = 1;
    A.key = group_key('a', as);
    Cp.key = group_key('c', as);
    for (size_t q = 0; q < as.size(); ++q) {
        OptimizeArgs const& a = *as[q];
        Workspace& w = *ws[q];
        const int64_t rs = (int64_t)a.br * a.bl;
        MpoSite const& op = *a.op;
        const int* lv = live ? live + q : nullptr;
        w.ensure(rs * a.d * op.froms.size(), rs * a.d * op.tos.size());
        for (size_t fi = 0; fi < op.froms.size(); ++fi)
            for (int aa = 0; aa < op.d; ++aa) {
                A.begin_batch(w.P.ptr + ((int64_t)fi * op.d + aa) * rs, a.br, a.br, a.bl, lv);
                A.add_seg(xs[q] + aa * rs, a.br, a.L + (int64_t)op.froms[fi] * a.bl * a.bl, a.bl, a.bl);
            }
        MixItem<double2> mi{&op, w.P.ptr, w.Q.ptr, rs};
        mi.live = lv;
        mix.push_back(mi);
        Cp.begin_batch(ys[q], a.br, a.br, a.bl * op.d, lv);
        for (size_t ti = 0; ti < op.tos.size(); ++ti)
            Cp.add_seg(a.R + (int64_t)op.tos[ti] * a.br * a.br, a.br, w.Q.ptr + (int64_t)ti * op.d * rs, a.br, a.br);
    }
}

// the same with complex64 operands (L32, R32 and the complex64 workspaces)
void heff_plans32(std::vector<OptimizeArgs const*> const& as, std::vector<Workspace*> const& ws,
                  std::vector<const float2*> const& xs, std::vector<float2*> const& ys, const int* live,
                  CGemmPlan& A, std::vector<MixItem<float2>>& mix, CGemmPlan& Cp) {
    A.M = A.N = Cp.M = Cp.N = 1;
    A.key = group_key('a', as);
    Cp.key = group_key('c', as);
    for (size_t q = 0; q < as.size(); ++q) {
        OptimizeArgs const& a = *as[q];
        Workspace& w = *ws[q];
        const int64_t rs = (int64_t)a.br * a.bl;
        MpoSite const& op = *a.op;
        const int* lv = live ? live + q : nullptr;
        w.ensure(rs * a.d * op.froms.size(), rs * a.d * op.tos.size(), (float2*)nullptr);
        for (size_t fi = 0; fi < op.froms.size(); ++fi)
            for (int aa = 0; aa < op.d; ++aa) {
                A.begin_batch(w.P32.ptr + ((int64_t)fi * op.d + aa) * rs, a.br, a.br, a.bl, lv);
                A.add_seg(xs[q] + aa * rs, a.br, a.L32 + (int64_t)op.froms[fi] * a.bl * a.bl, a.bl, a.bl);
            }
        MixItem<float2> mi{&op, w.P32.ptr, w.Q32.ptr, rs};
        mi.live = lv;
        mix.push_back(mi);
        Cp.begin_batch(ys[q], a.br, a.br, a.bl * op.d, lv);
        for (size_t ti = 0; ti < op.tos.size(); ++ti)
            Cp.add_seg(a.R32 + (int64_t)op.tos[ti] * a.br * a.br, a.br, w.Q32.ptr + (int64_t)ti * op.d * rs, a.br,
                       a.br);
    }
}
}  // namespace

void optimize_sites_batched(std::vector<OptimizeArgs const*> const& args, std::vector<Workspace*> const& wss,
                            std::vector<OptimizeOut>& outs, hipStream_t s, double* krylov_ms) {
    const size_t C = args.size();
    outs.assign(C, OptimizeOut());
    std::vector<size_t> idx;  // the chains solved together
    for (size_t c = 0; c < C; ++c) {
        OptimizeArgs const& a = *args[c];
        const int64_t nfull = (int64_t)a.d * a.bl * a.br;
        if (a.proj->active() || nfull <= 4) outs[c] = optimize_site(a, *wss[c], s);
        else idx.push_back(c);
    }
    if (idx.empty()) return;
    // group by Krylov size
    std::vector<int> ncvs;
    for (size_t c : idx) {
        OptimizeArgs const& a = *args[c];
        const int nc = std::min<int64_t>(a.ncv, (int64_t)a.d * a.bl * a.br);
        if (std::find(ncvs.begin(), ncvs.end(), nc) == ncvs.end()) ncvs.push_back(nc);
    }
    for (int ncv : ncvs) {
        std::vector<size_t> grp;
        for (size_t c : idx) {
            OptimizeArgs const& a = *args[c];
            if (std::min<int64_t>(a.ncv, (int64_t)a.d * a.bl * a.br) == ncv) grp.push_back(c);
        }
        const bool mixed = args[grp[0]]->mixed;
        std::vector<std::unique_ptr<HeffOperator>> op64;
        std::vector<std::unique_ptr<Heff32Operator>> op32;
        std::vector<IramTask> tasks(grp.size());
        for (size_t q = 0; q < grp.size(); ++q) {
            OptimizeArgs const& a = *args[grp[q]];
            if (a.mixed != mixed) throw Error("optimize_sites_batched: mixed and FP64 chains in one group", -33);
            op64.emplace_back(new HeffOperator(a, *wss[grp[q]]));
            if (mixed) op32.emplace_back(new Heff32Operator(a, *wss[grp[q]]));
            const int64_t rs = (int64_t)a.br * a.bl;
            if (mixed) wss[grp[q]]->ensure(rs * a.d * a.op->froms.size(), rs * a.d * a.op->tos.size(), (float2*)nullptr);
            else wss[grp[q]]->ensure(rs * a.d * a.op->froms.size(), rs * a.d * a.op->tos.size());
            IramTask& t = tasks[q];
            t.op = mixed ? static_cast<Operator*>(op32.back().get()) : static_cast<Operator*>(op64.back().get());
            t.ncv = ncv;
            t.which = a.which;
            t.tol = mixed ? a.tol32 : a.tol;
            t.max_restarts = a.max_iterations;
            t.guess = a.guess;
            t.result = a.result;
        }
        // Persistent grouped operator: every chain's matvec reads a fixed input
        // vector and writes a fixed output vector, so the grouped launches are
        // prepared once for the whole solve and gated per chain by the solver's
        // live flags.
        BatchApply apply;
        std::vector<ZBuf> xin(mixed ? 0 : grp.size()), yout(mixed ? 0 : grp.size());
        std::vector<std::unique_ptr<PreparedLaunch>> prepared;
        apply.single = mixed;
        for (size_t q = 0; q < grp.size(); ++q) {
            if (mixed) {
                apply.in.push_back(op32[q]->x32.ptr);
                apply.out.push_back(op32[q]->y32.ptr);
            } else {
                const int64_t nfull = op64[q]->n;
                xin[q].resize((size_t)nfull);
                yout[q].resize((size_t)nfull);
                apply.in.push_back(xin[q].ptr);
                apply.out.push_back(yout[q].ptr);
            }
        }
        std::vector<OptimizeArgs const*> gargs;
        std::vector<Workspace*> gws;
        for (size_t q = 0; q < grp.size(); ++q) {
            gargs.push_back(args[grp[q]]);
            gws.push_back(wss[grp[q]]);
        }
        apply.bind = [&](const int* live, hipStream_t) {
            if (!mixed) {
                GemmPlan A, Cp;
                std::vector<MixItem<double2>> mix;
                std::vector<const double2*> xs;
                std::vector<double2*> ys;
                for (size_t q = 0; q < grp.size(); ++q) {
                    xs.push_back(xin[q].ptr);
                    ys.push_back(yout[q].ptr);
                }
                heff_plans64(gargs, gws, xs, ys, live, A, mix, Cp);
                prepared.push_back(prepare_gemm(A));
                prepared.push_back(prepare_mix(mix));
                prepared.push_back(prepare_gemm(Cp));
            } else {
                CGemmPlan A, Cp;
                std::vector<MixItem<float2>> mix;
                std::vector<const float2*> xs;
                std::vector<float2*> ys;
                for (size_t q = 0; q < grp.size(); ++q) {
                    xs.push_back(op32[q]->x32.ptr);
                    ys.push_back(op32[q]->y32.ptr);
                }
                heff_plans32(gargs, gws, xs, ys, live, A, mix, Cp);
                prepared.push_back(prepare_gemm(A));
                prepared.push_back(prepare_mix(mix));
                prepared.push_back(prepare_gemm(Cp));
            }
        };
        apply.launch = [&](hipStream_t st) {
            for (auto const& pl : prepared) pl->launch(st);
        };
        const auto k0 = std::chrono::steady_clock::now();
        iram_batch(tasks, apply, s);
        if (krylov_ms)
            *krylov_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - k0).count();
        // Post-processing of the whole group: the mixed form's FP64 Rayleigh
        // quotient as one grouped FP64 matvec, and every dot product and norm in
        // one batched reduction (one host read).
        std::vector<size_t> rq, nrm;  // positions in grp: Rayleigh quotient; norm of the result
        std::vector<IramResult> rs(grp.size());
        for (size_t q = 0; q < grp.size(); ++q) {
            const size_t c = grp[q];
            OptimizeArgs const& a = *args[c];
            const int64_t nfull = (int64_t)a.d * a.bl * a.br;
            IramResult& r = rs[q];
            r = tasks[q].out;
            OptimizeOut& o = outs[c];
            if (r.info == -14) {  // the retries of optimize_site, one chain at a time
                o = optimize_site(a, *wss[c], s);
                if (mixed) o.matvecs32 += r.matvecs;
                else o.matvecs += r.matvecs;
                continue;
            }
            o.strategy = (a.bl * a.br < a.cl * a.cr) ? 2 : 3;
            if (mixed) o.matvecs32 += r.matvecs;
            else o.matvecs += r.matvecs;
            if (mixed && r.info == 0) {
                if (a.polish) {
                    ZBuf start((size_t)nfull);
                    d2d(start.ptr, a.result, (size_t)nfull * sizeof(double2));
                    IramResult r2 = iram(*op64[q], std::min<int64_t>(a.ncv, std::min<int64_t>(nfull, kMaxNcv)), a.which,
                                         a.tol, a.max_iterations, start.ptr, a.result, s);
                    o.matvecs += r2.matvecs;
                    if (r2.info == 0) r = r2;
                } else {
                    rq.push_back(q);
                    ++o.matvecs;
                }
            }
            if (r.info != 0) dzero(a.result, (size_t)nfull * sizeof(double2));
            else nrm.push_back(q);
        }
        std::vector<ZBuf> hy(rq.size());
        if (!rq.empty()) {
            std::vector<OptimizeArgs const*> ra;
            std::vector<Workspace*> rw;
            std::vector<const double2*> xs;
            std::vector<double2*> ys;
            for (size_t k = 0; k < rq.size(); ++k) {
                OptimizeArgs const& a = *args[grp[rq[k]]];
                hy[k].resize((size_t)a.d * a.bl * a.br);
                ra.push_back(&a);
                rw.push_back(wss[grp[rq[k]]]);
                xs.push_back(a.result);
                ys.push_back(hy[k].ptr);
            }
            GemmPlan A, Cp;
            std::vector<MixItem<double2>> mix;
            heff_plans64(ra, rw, xs, ys, nullptr, A, mix, Cp);
            zgemm_run(A, s);
            mix_blocks_batch(mix, s);
            zgemm_run(Cp, s);
        }
        std::vector<RedItem> items;
        for (size_t k = 0; k < rq.size(); ++k) {
            OptimizeArgs const& a = *args[grp[rq[k]]];
            items.push_back(RedItem{a.result, hy[k].ptr, (int64_t)a.d * a.bl * a.br});
        }
        for (size_t q : nrm) {
            OptimizeArgs const& a = *args[grp[q]];
            items.push_back(RedItem{a.result, a.result, (int64_t)a.d * a.bl * a.br});
        }
        const std::vector<zd> red = reduce_batch(items, s);
        std::vector<double> norm_of(grp.size(), 0.0);
        for (size_t k = 0; k < nrm.size(); ++k) norm_of[nrm[k]] = std::sqrt(std::max(0.0, red[rq.size() + k].real()));
        for (size_t k = 0; k < rq.size(); ++k) {
            const double nn = norm_of[rq[k]];
            rs[rq[k]].eigenvalue = red[k] / (nn * nn);
        }
        for (size_t q = 0; q < grp.size(); ++q) {
            if (tasks[q].out.info == -14) continue;
            OptimizeOut& o = outs[grp[q]];
            o.info = rs[q].info;
            o.iterations = rs[q].restarts;
            o.eigenvalue = rs[q].eigenvalue;
            o.normal = norm_of[q];
        }
    }
}

}  // namespace nc
