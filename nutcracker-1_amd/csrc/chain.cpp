// Device-resident sweep driver: the C++ counterpart of the reference's Chain /
// BaseChain (sources/common/chain.cpp, base_chain.cpp; includes/nutcracker/
// chain.hpp:248-315) for a fixed bond dimension.  State sites, environments and
// workspaces stay in HBM for the whole run; the host keeps only dimensions,
// energies and the small eigen-solver matrices.
//
//  * reset (chain.cpp:175-237): sites 2..N right-normalised from caller-supplied
//    raw tensors, site 1 divided by its norm, right environments built by
//    contract_sos_right, energy = <psi|H|psi> at site 1.
//  * move<side> / absorb<side> (chain.hpp:248-315, states.cpp:94-164): polar
//    site move, renormalisation of the new middle site, environment update.
//  * optimizeSite (base_chain.cpp:81-110) with the post-checks of
//    optimizeStateSite (optimizer.cpp:201-315) and the acceptance rule.
//  * performOptimizationSweep (chain.cpp:147-163), sweepUntilConverged
//    (base_chain.cpp:112-123).
#include "nc_internal.hpp"
#include "nc_chain.hpp"

#include <chrono>
#include <cmath>

namespace nc {

namespace {
double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

bool outside_tolerance(double a, double b, double t) {
    // utilities.hpp:418-420
    return ((std::abs(a) + std::abs(b)) / 2 > t) && (std::abs(a - b) / (std::abs(a) + std::abs(b) + t) > t);
}

// regression checks of the optimizer modes (optimizer.cpp:190-200)
bool regression(const char* which, double from, double to, double tol) {
    if (which[0] == 'S') return to > from && outside_tolerance(from, to, tol);
    if (which[0] == 'L' && which[1] == 'R') return to < from && outside_tolerance(from, to, tol);
    return std::abs(to) < std::abs(from) && outside_tolerance(std::abs(from), std::abs(to), tol);
}
}  // namespace

Chain::Chain(int n_sites, std::vector<std::unique_ptr<MpoSite>> ops_, std::vector<int> bonds_)
    : N(n_sites), ops(std::move(ops_)), bonds(std::move(bonds_)) {
    if ((int)ops.size() != N || (int)bonds.size() != N + 1) throw Error("chain: inconsistent sizes", -60);
    sites.resize(N);
    for (int i = 0; i < N; ++i) sites[i].resize(site_size(i));
    trivial.resize(1);
    double2 one = make_double2(1, 0);
    h2d(trivial.ptr, &one, sizeof(double2));
}

size_t Chain::site_size(int i) const { return (size_t)bonds[i + 1] * bonds[i] * ops[i]->d; }

void Chain::set_raw_site(int i, const double2* host) {
    h2d(sites[i].ptr, host, site_size(i) * sizeof(double2));
}

void Chain::get_site(int i, double2* host) const {
    // sites[i] is current only for the middle site and the stored neighbours
    d2h(host, sites[i].ptr, site_size(i) * sizeof(double2));
}

void Chain::normalize_right_site(int i, hipStream_t s) {
    // rows l of A(l; r, s) orthonormal: polar factor of the (bl x br d) matrix
    const int bl = bonds[i], br = bonds[i + 1], d = ops[i]->d;
    if ((int64_t)br * d < bl) throw Error("Not enough degrees of freedom to normalize.", -61);
    ZBuf W((size_t)bl * br * d), Up((size_t)bl * br * d), Pc((size_t)bl * bl);
    permute_102(sites[i].ptr, W.ptr, br, bl, d, s);
    polar_split(bl, (int64_t)br * d, W.ptr, Up.ptr, Pc.ptr, s);
    permute_102(Up.ptr, sites[i].ptr, bl, br, d, s);
}

namespace {
void grow(DevBuf<float2>& b, size_t n) {
    if (b.n < n) b.resize(n);
}
}  // namespace

void Chain::update_left_env(int i, hipStream_t s) {
    // absorb<Left>: L_{i+1} = contract_sos_left(L_i, site i)
    MpoSite const& op = *ops[i];
    const int bl = bonds[i], br = bonds[i + 1];
    const size_t nout = (size_t)br * br * op.cr;
    left_envs[i + 1].resize(nout);
    if (!(opts.mixed && opts.env32)) {
        env_left(bl, br, op, left_envs[i].ptr, sites[i].ptr, left_envs[i + 1].ptr, ws, s);
        return;
    }
    const size_t nin = (size_t)bl * bl * op.cl, npsi = site_size(i);
    grow(t32a, nin);
    grow(t32b, npsi);
    grow(t32c, nout);
    to_c64(left_envs[i].ptr, t32a.ptr, nin, s);
    to_c64(sites[i].ptr, t32b.ptr, npsi, s);
    env_left(bl, br, op, t32a.ptr, t32b.ptr, t32c.ptr, ws, s);
    to_c128(t32c.ptr, left_envs[i + 1].ptr, nout, s);
}

void Chain::update_right_env(int i, hipStream_t s) {
    // absorb<Right>: R_i = contract_sos_right(R_{i+1}, site i)
    MpoSite const& op = *ops[i];
    const int bl = bonds[i], br = bonds[i + 1];
    const size_t nout = (size_t)bl * bl * op.cl;
    right_envs[i].resize(nout);
    if (!(opts.mixed && opts.env32)) {
        env_right(bl, br, op, right_envs[i + 1].ptr, sites[i].ptr, right_envs[i].ptr, ws, s);
        return;
    }
    const size_t nin = (size_t)br * br * op.cr, npsi = site_size(i);
    grow(t32a, nin);
    grow(t32b, npsi);
    grow(t32c, nout);
    to_c64(right_envs[i + 1].ptr, t32a.ptr, nin, s);
    to_c64(sites[i].ptr, t32b.ptr, npsi, s);
    env_right(bl, br, op, t32a.ptr, t32b.ptr, t32c.ptr, ws, s);
    to_c128(t32c.ptr, right_envs[i].ptr, nout, s);
}

void Chain::reset() {
    hipStream_t s = Context::get().stream;
    double t0 = now_ms();
    // right environments: R_N = trivial, R_i = contract_sos_right(R_{i+1}, site i)
    right_envs.clear();
    left_envs.clear();
    right_envs.resize(N + 1);
    left_envs.resize(N + 1);
    right_envs[N].resize(1);
    d2d(right_envs[N].ptr, trivial.ptr, sizeof(double2));
    for (int i = N - 1; i >= 1; --i) {
        normalize_right_site(i, s);
        update_right_env(i, s);
    }
    // first site: divided by its norm (core.f95:2661-2671)
    double nrm = dznrm2(sites[0].ptr, site_size(0), s);
    zscal(sites[0].ptr, zd(1.0 / nrm, 0), site_size(0), s);
    left_envs[0].resize(1);
    d2d(left_envs[0].ptr, trivial.ptr, sizeof(double2));
    current = 0;
    optimized = false;
    zd e = site_expectation();
    const double imag_tol = opts.mixed && opts.env32 ? 1e-6 : 1e-10;  // complex64 environments: FP32 round-off
    if (std::abs(e.imag()) / std::abs(e) > imag_tol) throw Error("InitialChainEnergyNotRealError", -62);
    energy = e.real();
    stats.t_reset_ms += now_ms() - t0;
}

zd Chain::site_expectation() {
    hipStream_t s = Context::get().stream;
    const int i = current;
    return expectation(bonds[i], bonds[i + 1], *ops[i], left_envs[i].ptr, sites[i].ptr, right_envs[i + 1].ptr, ws, s);
}

void Chain::move_right() {
    hipStream_t s = Context::get().stream;
    const int i = current;
    if (i + 1 >= N) throw Error("move_right past the last site", -63);
    double t0 = now_ms();
    const int bl = bonds[i], bm = bonds[i + 1], br = bonds[i + 2];
    const int dl = ops[i]->d, dr = ops[i + 1]->d;
    ZBuf nl(site_size(i)), dn(site_size(i + 1));
    int info = norm_denorm_right(bl, bm, br, dl, dr, sites[i].ptr, sites[i + 1].ptr, nl.ptr, dn.ptr, s);
    if (info != 0) throw Error("NormalizationError", info);
    double nrm = dznrm2(dn.ptr, site_size(i + 1), s);
    zscal(dn.ptr, zd(1.0 / nrm, 0), site_size(i + 1), s);
    d2d(sites[i].ptr, nl.ptr, site_size(i) * sizeof(double2));
    d2d(sites[i + 1].ptr, dn.ptr, site_size(i + 1) * sizeof(double2));
    NC_HIP(hipStreamSynchronize(s));
    double t1 = now_ms();
    update_left_env(i, s);
    NC_HIP(hipStreamSynchronize(s));
    double t2 = now_ms();
    stats.t_move_ms += t1 - t0;
    stats.t_env_ms += t2 - t1;
    ++stats.moves;
    current = i + 1;
    optimized = false;
}

void Chain::move_left() {
    hipStream_t s = Context::get().stream;
    const int i = current;
    if (i == 0) throw Error("move_left past the first site", -64);
    double t0 = now_ms();
    const int bll = bonds[i - 1], bm = bonds[i], br = bonds[i + 1];
    const int dl = ops[i - 1]->d, d = ops[i]->d;
    ZBuf dn(site_size(i - 1)), nr(site_size(i));
    int info = norm_denorm_left(bll, bm, br, dl, d, sites[i - 1].ptr, sites[i].ptr, dn.ptr, nr.ptr, s);
    if (info != 0) throw Error("NormalizationError", info);
    double nrm = dznrm2(dn.ptr, site_size(i - 1), s);
    zscal(dn.ptr, zd(1.0 / nrm, 0), site_size(i - 1), s);
    d2d(sites[i - 1].ptr, dn.ptr, site_size(i - 1) * sizeof(double2));
    d2d(sites[i].ptr, nr.ptr, site_size(i) * sizeof(double2));
    NC_HIP(hipStreamSynchronize(s));
    double t1 = now_ms();
    update_right_env(i, s);
    NC_HIP(hipStreamSynchronize(s));
    double t2 = now_ms();
    stats.t_move_ms += t1 - t0;
    stats.t_env_ms += t2 - t1;
    ++stats.moves;
    current = i - 1;
    optimized = false;
}

int Chain::optimize_site() {
    hipStream_t s = Context::get().stream;
    const int i = current;
    double t0 = now_ms();
    const int bl = bonds[i], br = bonds[i + 1];
    MpoSite const& op = *ops[i];
    ZBuf result(site_size(i));
    Projector none;
    none.n = (int)site_size(i);
    none.orth = (int)site_size(i);
    OptimizeArgs a{bl, br, op.cl, op.cr, op.d, left_envs[i].ptr, right_envs[i + 1].ptr, &op, &none,
                   opts.which, opts.site_tol, opts.max_iterations, sites[i].ptr, result.ptr, opts.ncv};
    DevBuf<float2> L32, R32;
    if (opts.mixed) {
        const size_t nl = (size_t)bl * bl * op.cl, nr = (size_t)br * br * op.cr;
        L32.resize(nl);
        R32.resize(nr);
        to_c64(left_envs[i].ptr, L32.ptr, nl, s);
        to_c64(right_envs[i + 1].ptr, R32.ptr, nr, s);
        a.L32 = L32.ptr;
        a.R32 = R32.ptr;
        a.mixed = true;
        a.tol32 = opts.tol32;
        a.polish = opts.polish;
    }
    OptimizeOut o = nc::optimize_site(a, ws, s);
    stats.matvecs += o.matvecs;
    stats.matvecs32 += o.matvecs32;
    ++stats.site_solves;
    int status = o.info;
    // complex64 environments are Hermitian only to FP32 round-off and carry its
    // noise into the energy: the complex-eigenvalue and regression checks and the
    // sweep convergence test use that floor
    const double floor_tol = opts.mixed && opts.env32 ? std::max(opts.sanity, opts.env32_floor) : opts.sanity;
    if (status == 0) {
        // optimizeStateSite post-checks (optimizer.cpp:265-300)
        zd ev = o.eigenvalue;
        zd ex = expectation(bl, br, op, left_envs[i].ptr, result.ptr, right_envs[i + 1].ptr, ws, s);
        const double san = opts.sanity;
        if ((std::abs(ev) + std::abs(ex)) / 2 > san && std::abs(ev - ex) / (std::abs(ev) + std::abs(ex) + san) > san)
            status = 101;  // OptimizerObtainedEigenvalueDifferentFromExpectationValue
        else if (std::abs(ev) > san && std::abs(ev.imag()) / std::abs(ev) > floor_tol)
            status = 102;  // OptimizerObtainedComplexEigenvalue
        else if (o.normal < 1 - san)
            status = 103;  // OptimizerObtainedVanishingEigenvector
    }
    if (status == 0) {
        const double e = o.eigenvalue.real();
        if (regression(opts.which, energy, e, floor_tol)) {
            status = 104;  // OptimizerObtainedRegressiveEigenvalue
        } else {
            if ((energy >= 0 && e >= 0) || (energy <= 0 && e <= 0) ||
                outside_tolerance(std::abs(energy), std::abs(e), opts.sanity)) {
                energy = e;
                d2d(sites[i].ptr, result.ptr, site_size(i) * sizeof(double2));
            }
            optimized = true;
        }
    }
    NC_HIP(hipStreamSynchronize(s));
    stats.t_opt_ms += now_ms() - t0;
    if (status != 0) ++stats.failures;
    last_status = status;
    return status;
}

void Chain::sweep() {
    // chain.cpp:147-163
    const int start = current;
    auto opt = [&] {
        int st = optimize_site();
        if (st != 0 && opts.rethrow) throw Error("optimizer failure at site " + std::to_string(current), st);
    };
    if (!optimized) opt();
    while (current + 1 < N) { move_right(); opt(); }
    while (current > 0) { move_left(); opt(); }
    while (current < start) { move_right(); opt(); }
    ++stats.sweeps;
}

void Chain::sweep_until_converged() {
    // base_chain.cpp:112-123 (Chain's convergence energy is always available, so
    // the optional initial sweep never runs)
    double prev = energy;
    sweep();
    double cur = energy;
    const double tol = opts.mixed && opts.env32 ? std::max(opts.sweep_tol, opts.env32_floor) : opts.sweep_tol;
    while (outside_tolerance(prev, cur, tol)) {
        sweep();
        prev = cur;
        cur = energy;
    }
}

}  // namespace nc
