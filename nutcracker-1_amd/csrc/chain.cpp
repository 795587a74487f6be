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

#include <algorithm>
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

std::vector<int> bandwidth_sequence(int requested, std::vector<int> const& dims) {
    // computeBandwidthDimensionSequence (utilities.cpp:32-91): 1, d1, d1 d2, ...
    // capped at `requested`, from both ends
    if (requested <= 0) throw Error("requested bandwidth must be greater than zero", -75);
    const int n = (int)dims.size();
    const int middle = (n + 1) / 2;
    auto side = [&](bool reverse) {
        std::vector<int> out{1};
        uint64_t p = 1;
        for (int k = 0; k < middle; ++k) {
            p *= (uint64_t)dims[reverse ? n - 1 - k : k];
            if (p >= (uint64_t)requested) break;
            out.push_back((int)p);
        }
        return out;
    };
    std::vector<int> fwd = side(false), rev = side(true);
    if ((int)fwd.size() == middle + 1 || (int)rev.size() == middle + 1)
        throw Error("RequestedBandwidthDimensionTooLargeError", -76);
    const int fill = n + 1 - (int)fwd.size() - (int)rev.size();
    for (int k = 0; k < fill; ++k) fwd.push_back(requested);
    for (int k = (int)rev.size() - 1; k >= 0; --k) fwd.push_back(rev[k]);
    return fwd;
}

int maximum_bandwidth(std::vector<int> const& dims) {
    // maximumBandwidthDimension (utilities.hpp:264-283)
    const int n = (int)dims.size();
    if (n == 1) return 1;
    const int middle = (n + 1) / 2;
    uint64_t a = 1, b = 1;
    const uint64_t cap = 0xFFFFFFFFull;
    for (int k = 0; k < middle; ++k) {
        a = std::min(cap, a * (uint64_t)dims[k]);
        b = std::min(cap, b * (uint64_t)dims[n - 1 - k]);
    }
    return (int)std::min<uint64_t>({a, b, (uint64_t)0x7FFFFFFF});
}

int minimum_bandwidth_for_projectors(std::vector<int> const& dims, int nproj) {
    // minimumBandwidthDimensionForProjectorCount (projectors.cpp:153-172)
    if (nproj == 0) return 1;
    const int maxb = maximum_bandwidth(dims);
    for (int b = 1; b <= maxb; ++b) {
        std::vector<int> seq = bandwidth_sequence(b, dims);
        for (size_t i = 0; i < dims.size(); ++i)
            if ((int64_t)dims[i] * seq[i] * seq[i + 1] > nproj) return b;
    }
    return maxb;
}

Chain::Chain(int n_sites, std::vector<std::unique_ptr<MpoSite>> ops_, std::vector<int> bonds_)
    : N(n_sites), ops(std::move(ops_)), bonds(std::move(bonds_)) {
    if ((int)ops.size() != N || (int)bonds.size() != N + 1) throw Error("chain: inconsistent sizes", -60);
    chi = *std::max_element(bonds.begin(), bonds.end());
    initial_chi = chi;
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
    const size_t np = projectors.size();
    left_ovl.clear();
    right_ovl.clear();
    left_ovl.resize(np);
    right_ovl.resize(np);
    for (size_t p = 0; p < np; ++p) {
        left_ovl[p].resize(N + 1);
        right_ovl[p].resize(N + 1);
        left_ovl[p][0].resize(1);
        d2d(left_ovl[p][0].ptr, trivial.ptr, sizeof(double2));
        right_ovl[p][N].resize(1);
        d2d(right_ovl[p][N].ptr, trivial.ptr, sizeof(double2));
    }
    for (int i = N - 1; i >= 1; --i) {
        normalize_right_site(i, s);
        update_right_env(i, s);
        update_right_overlaps(i, s);
    }
    // first site: divided by its norm (core.f95:2661-2671)
    double nrm = dznrm2(sites[0].ptr, site_size(0), s);
    zscal(sites[0].ptr, zd(1.0 / nrm, 0), site_size(0), s);
    left_envs[0].resize(1);
    d2d(left_envs[0].ptr, trivial.ptr, sizeof(double2));
    current = 0;
    optimized = false;
    reset_projector_matrix();
    if (np > 0) {
        // chain.cpp:218-225: find a site with room outside the lower levels,
        // filter the state there, and come back to the first site
        while (site_projector.orth == 0) {
            if (current + 1 >= N) throw Error("no site has room outside the lower levels at this bond dimension", -70);
            move_right();
        }
        const int64_t n = (int64_t)site_size(current);
        ZBuf proj_x((size_t)site_projector.orth), tmp((size_t)n), full((size_t)n);
        site_projector.project(sites[current].ptr, proj_x.ptr, tmp.ptr, s);
        site_projector.unproject(proj_x.ptr, full.ptr, s);
        const double fn = dznrm2(full.ptr, n, s);
        if (!(fn > 1e-7)) throw Error("the initial state lies in the space of the lower levels", -71);
        zscal(full.ptr, zd(1.0 / fn, 0), n, s);
        d2d(sites[current].ptr, full.ptr, (size_t)n * sizeof(double2));
        while (current > 0) move_left();
        optimized = false;
    }
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
    update_left_overlaps(i, s);
    NC_HIP(hipStreamSynchronize(s));
    double t2 = now_ms();
    stats.t_move_ms += t1 - t0;
    stats.t_env_ms += t2 - t1;
    ++stats.moves;
    current = i + 1;
    optimized = false;
    reset_projector_matrix();
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
    update_right_overlaps(i, s);
    NC_HIP(hipStreamSynchronize(s));
    double t2 = now_ms();
    stats.t_move_ms += t1 - t0;
    stats.t_env_ms += t2 - t1;
    ++stats.moves;
    current = i - 1;
    optimized = false;
    reset_projector_matrix();
}

void Chain::prepare_site(SitePrep& p) {
    hipStream_t s = Context::get().stream;
    const int i = current;
    p.t0 = now_ms();
    const int bl = bonds[i], br = bonds[i + 1];
    MpoSite const& op = *ops[i];
    p.result.resize(site_size(i));
    p.none.n = (int)site_size(i);
    p.none.orth = (int)site_size(i);
    p.proj = projectors.empty() ? &p.none : &site_projector;
    p.a = OptimizeArgs{bl, br, op.cl, op.cr, op.d, left_envs[i].ptr, right_envs[i + 1].ptr, &op, p.proj,
                       opts.which, opts.site_tol, opts.max_iterations, sites[i].ptr, p.result.ptr, opts.ncv};
    if (opts.mixed) {
        const size_t nl = (size_t)bl * bl * op.cl, nr = (size_t)br * br * op.cr;
        p.L32.resize(nl);
        p.R32.resize(nr);
        to_c64(left_envs[i].ptr, p.L32.ptr, nl, s);
        to_c64(right_envs[i + 1].ptr, p.R32.ptr, nr, s);
        p.a.L32 = p.L32.ptr;
        p.a.R32 = p.R32.ptr;
        p.a.mixed = true;
        p.a.tol32 = opts.tol32;
        p.a.polish = opts.polish;
    }
}

int Chain::check_site(SitePrep& p, OptimizeOut const& o) {
    // optimizeStateSite post-checks (optimizer.cpp:265-310) and the regression
    // check of BaseChain::optimizeSite (base_chain.cpp:85-100)
    hipStream_t s = Context::get().stream;
    const int i = current;
    const int bl = bonds[i], br = bonds[i + 1];
    MpoSite const& op = *ops[i];
    int status = o.info;
    // complex64 environments are Hermitian only to FP32 round-off and carry its
    // noise into the energy: the complex-eigenvalue and regression checks and the
    // sweep convergence test use that floor
    const double floor_tol = opts.mixed && opts.env32 ? std::max(opts.sanity, opts.env32_floor) : opts.sanity;
    if (status == 0) {
        zd ev = o.eigenvalue;
        const zd ex = p.have_ex ? p.ex : expectation(bl, br, op, left_envs[i].ptr, p.result.ptr, right_envs[i + 1].ptr,
                                                     ws, s);
        p.have_ex = false;  // a refined result is checked against its own value
        const double san = opts.sanity;
        if ((std::abs(ev) + std::abs(ex)) / 2 > san && std::abs(ev - ex) / (std::abs(ev) + std::abs(ex) + san) > san)
            status = 101;  // OptimizerObtainedEigenvalueDifferentFromExpectationValue
        else if (std::abs(ev) > san && std::abs(ev.imag()) / std::abs(ev) > floor_tol)
            status = 102;  // OptimizerObtainedComplexEigenvalue
        else if (o.normal < 1 - san)
            status = 103;  // OptimizerObtainedVanishingEigenvector
        else if (p.proj->active()) {
            ZBuf tmp(site_size(i));
            if (p.proj->overlap(p.result.ptr, tmp.ptr, s) > san)
                status = 105;  // OptimizerObtainedEigenvectorInProjectorSpace
        }
    }
    if (status == 0 && regression(opts.which, energy, o.eigenvalue.real(), floor_tol))
        status = 104;  // OptimizerObtainedRegressiveEigenvalue
    return status;
}

int Chain::finish_site(SitePrep& p, OptimizeOut const& o_in) {
    hipStream_t s = Context::get().stream;
    const int i = current;
    OptimizeOut o = o_in;
    stats.matvecs += o.matvecs;
    stats.matvecs32 += o.matvecs32;
    ++stats.site_solves;
    int status = check_site(p, o);
    if (status == 104 && p.a.mixed) {
        // a complex64 Krylov vector converged to 1e-7 can sit measurably above
        // the previous energy when the site's lowest pair is nearly degenerate:
        // finish that site in FP64, starting from it
        OptimizeArgs a2 = p.a;
        a2.mixed = false;
        a2.polish = false;
        ZBuf guess(site_size(i));
        d2d(guess.ptr, p.result.ptr, site_size(i) * sizeof(double2));
        a2.guess = guess.ptr;
        o = nc::optimize_site(a2, ws, s);
        stats.matvecs += o.matvecs;
        ++stats.fp64_refinements;
        status = check_site(p, o);
    }
    if (status == 0) {
        // acceptance rule of BaseChain::optimizeSite (base_chain.cpp:101-105)
        const double e = o.eigenvalue.real();
        if ((energy >= 0 && e >= 0) || (energy <= 0 && e <= 0) ||
            outside_tolerance(std::abs(energy), std::abs(e), opts.sanity)) {
            energy = e;
            d2d(sites[i].ptr, p.result.ptr, site_size(i) * sizeof(double2));
        }
        optimized = true;
    }
    NC_HIP(hipStreamSynchronize(s));
    stats.t_opt_ms += now_ms() - p.t0;
    if (status != 0) ++stats.failures;
    last_status = status;
    return status;
}

int Chain::optimize_site() {
    SitePrep p;
    prepare_site(p);
    OptimizeOut o = nc::optimize_site(p.a, ws, Context::get().stream);
    return finish_site(p, o);
}

// ---------------------------------------------------------------------------
// Excited levels (chain.cpp:60-73, 240-263, 301-310; projectors.hpp:324-485)
// ---------------------------------------------------------------------------
void Chain::update_left_overlaps(int i, hipStream_t s) {
    // absorb<Left> (chain.hpp:248-279): the lower level's left-canonical site
    for (size_t p = 0; p < projectors.size(); ++p) {
        LevelProjector const& pr = projectors[p];
        const int blo = pr.bonds[i], bro = pr.bonds[i + 1], bln = bonds[i], brn = bonds[i + 1];
        left_ovl[p][i + 1].resize((size_t)brn * bro);
        vs_left(blo, bro, bln, brn, ops[i]->d, left_ovl[p][i].ptr, pr.left[i].ptr, sites[i].ptr,
                left_ovl[p][i + 1].ptr, s);
    }
}

void Chain::update_right_overlaps(int i, hipStream_t s) {
    // absorb<Right>: the lower level's right-canonical site
    for (size_t p = 0; p < projectors.size(); ++p) {
        LevelProjector const& pr = projectors[p];
        const int blo = pr.bonds[i], bro = pr.bonds[i + 1], bln = bonds[i], brn = bonds[i + 1];
        right_ovl[p][i].resize((size_t)blo * bln);
        vs_right(blo, bro, bln, brn, ops[i]->d, right_ovl[p][i + 1].ptr, pr.right[i].ptr, sites[i].ptr,
                 right_ovl[p][i].ptr, s);
    }
}

void Chain::reset_projector_matrix() {
    // formProjectorMatrix (projectors.hpp:402-485): one overlap vector per lower
    // level at the current site, turned into Householder reflectors
    const size_t np = projectors.size();
    if (np == 0) return;
    hipStream_t s = Context::get().stream;
    const int i = current;
    const int64_t n = (int64_t)site_size(i);
    ZBuf V((size_t)n * np);
    for (size_t p = 0; p < np; ++p) {
        LevelProjector const& pr = projectors[p];
        overlap_vector(pr.bonds[i], pr.bonds[i + 1], bonds[i], bonds[i + 1], ops[i]->d, left_ovl[p][i].ptr,
                       right_ovl[p][i + 1].ptr, pr.middle[i].ptr, V.ptr + (size_t)p * n, s);
    }
    const int k = (int)std::min<int64_t>(n, (int64_t)np);
    std::vector<zd> tau(k);
    std::vector<uint32_t> sw(k);
    const int rank = reflectors_from_vectors(n, (int)np, V.ptr, tau.data(), sw.data(), s);
    site_projector.set_device((int)n, (int)np, k, (int)(n - rank), V.ptr, tau.data(), sw.data());
}

void Chain::add_projector_from_state() {
    // constructAndAddProjectorFromState (chain.cpp:63-73) via
    // computeProjectorFromStateSites (projectors.hpp:324-359)
    if (current != 0) throw Error("ChainNotAtFirstSiteError", -72);
    hipStream_t s = Context::get().stream;
    LevelProjector pr;
    pr.bonds = bonds;
    pr.left.resize(N);
    pr.middle.resize(N);
    pr.right.resize(N);
    ZBuf cur(site_size(0));
    d2d(cur.ptr, sites[0].ptr, site_size(0) * sizeof(double2));
    for (int i = 0; i + 1 < N; ++i) {
        const int bl = bonds[i], bm = bonds[i + 1], br = bonds[i + 2];
        pr.left[i].resize(site_size(i));
        pr.middle[i].resize(site_size(i));
        pr.right[i + 1].resize(site_size(i + 1));
        ZBuf next(site_size(i + 1));
        int info = norm_overlap_tensors(bl, bm, br, ops[i]->d, ops[i + 1]->d, cur.ptr, sites[i + 1].ptr,
                                        pr.left[i].ptr, pr.middle[i].ptr, next.ptr, pr.right[i + 1].ptr, s);
        if (info != 0) throw Error("NormalizationError", info);
        cur = std::move(next);
    }
    pr.middle[N - 1].resize(site_size(N - 1));
    overlap_site(bonds[N], bonds[N - 1], ops[N - 1]->d, cur.ptr, pr.middle[N - 1].ptr, s);
    NC_HIP(hipStreamSynchronize(s));
    projectors.push_back(std::move(pr));
    optimized = false;
}

void Chain::clear_projectors() {
    projectors.clear();
    left_ovl.clear();
    right_ovl.clear();
}

double Chain::projector_overlap() {
    // computeProjectorOverlapAtCurrentSite (chain.cpp:60-62)
    if (projectors.empty()) return 0.0;
    hipStream_t s = Context::get().stream;
    ZBuf tmp(site_size(current));
    return site_projector.overlap(sites[current].ptr, tmp.ptr, s);
}

void Chain::randomize(uint64_t seed) {
    hipStream_t s = Context::get().stream;
    for (int i = 0; i < N; ++i) {
        const int64_t n = (int64_t)site_size(i);
        const uint64_t key = (seed + 1) * 0xD1B54A32D192ED03ull ^ ((uint64_t)(i + 1) * 0x8CB92BA72F3D8DD7ull);
        fill_random(sites[i].ptr, n, key, s);
    }
    NC_HIP(hipStreamSynchronize(s));
}

std::vector<int> Chain::dims() const {
    std::vector<int> d(N);
    for (int i = 0; i < N; ++i) d[i] = ops[i]->d;
    return d;
}

void Chain::set_bandwidth(int new_chi) {
    chi = new_chi;
    bonds = bandwidth_sequence(new_chi, dims());
    for (int i = 0; i < N; ++i) sites[i].resize(site_size(i));
}

void Chain::increase_bandwidth(int new_chi) {
    // increaseBandwidthDimension (chain.cpp:75-121) with increase_bandwidth_between
    // (core.f95:3090-3144): from the right end, each bond that grows from b to b'
    // is widened by a random b' x b isometry M -- the left site absorbs M, the
    // right site M^H -- and the right site is re-normalised by a polar move
    // whose factor the left site absorbs.  The state is unchanged.
    optimized = false;
    if (new_chi == chi) return;
    if (new_chi < chi) throw Error("the bandwidth dimension can only grow", -77);
    if (current != 0) throw Error("ChainNotAtFirstSiteError", -72);
    hipStream_t s = Context::get().stream;
    std::vector<int> nb = bandwidth_sequence(new_chi, dims());
    std::vector<int> cur = bonds;  // bonds of the tensors as they are being widened
    for (int i = N - 1; i >= 1; --i) {
        const int old = cur[i], nw = nb[i];
        const int dl = ops[i - 1]->d, dr = ops[i]->d;
        const int bll = cur[i - 1], br = cur[i + 1];
        ZBuf A1, B1;
        const double2* a_in = sites[i - 1].ptr;
        const double2* b_in = sites[i].ptr;
        if (nw != old) {
            // M = transpose of the polar factor of a random old x nw matrix
            ZBuf W((size_t)old * nw), Up((size_t)old * nw), Pc((size_t)old * old), M((size_t)nw * old);
            fill_random(W.ptr, (int64_t)old * nw, 0xA24BAED4963EE407ull ^ ((uint64_t)new_chi << 32) ^ (uint64_t)i, s);
            int info = polar_split(old, nw, W.ptr, Up.ptr, Pc.ptr, s);
            if (info != 0) throw Error("create_bandwidth_increase_matrix: bad rank", -78);
            permute_102(Up.ptr, M.ptr, old, nw, 1, s);
            // left site (old, bll, dl) -> (nw, bll, dl): M A
            A1.resize((size_t)nw * bll * dl);
            GemmPlan g1;
            g1.M = nw; g1.N = bll * dl;
            g1.begin_batch(A1.ptr, nw);
            g1.add_seg(M.ptr, nw, sites[i - 1].ptr, old, old);
            zgemm_run(g1, s);
            // right site (br, old, dr) -> (br, nw, dr): B_s M^H for each s
            B1.resize((size_t)br * nw * dr);
            GemmPlan g2;
            g2.M = br; g2.N = nw; g2.opA = OpN; g2.opB = OpC;
            for (int a = 0; a < dr; ++a) {
                g2.begin_batch(B1.ptr + (size_t)a * br * nw, br);
                g2.add_seg(sites[i].ptr + (size_t)a * br * old, br, M.ptr, nw, old);
            }
            zgemm_run(g2, s);
            a_in = A1.ptr;
            b_in = B1.ptr;
        }
        // right-normalise the right site (it may hold the factor of the bond to
        // its right), the left one absorbs the factor
        ZBuf L2((size_t)nw * bll * dl), R2((size_t)br * nw * dr);
        int info = norm_denorm_left(bll, nw, br, dl, dr, a_in, b_in, L2.ptr, R2.ptr, s);
        if (info != 0) throw Error("NormalizationError", info);
        sites[i - 1] = std::move(L2);
        sites[i] = std::move(R2);
        cur[i] = nw;
        // absorb<Right>: environment and overlap boundaries of site i
        bonds = cur;
        update_right_env(i, s);
        update_right_overlaps(i, s);
    }
    bonds = nb;
    chi = new_chi;
    NC_HIP(hipStreamSynchronize(s));
    reset_projector_matrix();
}

void Chain::optimize_chain() {
    // optimizeChain (base_chain.cpp:64-79); with fixed chi only the first step
    sweep_until_converged();
    if (!opts.grow) return;
    const int maxb = maximum_bandwidth(dims());
    double prev = energy;
    increase_bandwidth(std::min(chi + 1, maxb));
    sweep_until_converged();
    double cur = energy;
    while (outside_tolerance(prev, cur, opts.chain_tol) && chi < maxb) {
        increase_bandwidth(std::min(chi + 1, maxb));
        sweep_until_converged();
        prev = cur;
        cur = energy;
    }
}

void Chain::restart_level(uint64_t seed) {
    // reset (chain.cpp:175-191): the initial bond dimension, raised to the
    // minimum that leaves room outside the lower levels, then a new random state
    if (opts.grow || chi < minimum_bandwidth_for_projectors(dims(), (int)projectors.size())) {
        const int maxb = maximum_bandwidth(dims());
        const int want = std::min(maxb, std::max(opts.grow ? initial_chi : chi,
                                                 minimum_bandwidth_for_projectors(dims(), (int)projectors.size())));
        if (want != chi || opts.grow) set_bandwidth(want);
    }
    randomize(seed);
    reset();
}

std::vector<double> Chain::solve_for_multiple_levels(int levels, uint64_t seed,
                                                     std::function<void(int)> const& after_level) {
    // solveForMultipleLevels (chain.cpp:301-310): converge (optimizeChain), keep
    // the state as a projector, restart from a new random state orthogonal to
    // the lower levels
    std::vector<double> energies;
    for (int k = 0; k + 1 < levels; ++k) {
        optimize_chain();
        energies.push_back(energy);
        if (after_level) after_level(k);
        add_projector_from_state();
        restart_level(seed + (uint64_t)k + 1);
    }
    if (levels > 0) {
        optimize_chain();
        energies.push_back(energy);
        if (after_level) after_level(levels - 1);
    }
    return energies;
}

zd mps_overlap(int N, std::vector<int> const& d, std::vector<int> const& bonds_a,
               std::vector<const double2*> const& a, std::vector<int> const& bonds_b,
               std::vector<const double2*> const& b) {
    // contracted from the left with the overlap sites (conjugates) of a
    hipStream_t s = Context::get().stream;
    ZBuf L(1), ov, next;
    double2 one = make_double2(1, 0);
    h2d(L.ptr, &one, sizeof(double2));
    for (int i = 0; i < N; ++i) {
        const int blo = bonds_a[i], bro = bonds_a[i + 1], bln = bonds_b[i], brn = bonds_b[i + 1];
        ov.resize((size_t)blo * d[i] * bro);
        overlap_site(bro, blo, d[i], a[i], ov.ptr, s);
        next.resize((size_t)brn * bro);
        vs_left(blo, bro, bln, brn, d[i], L.ptr, ov.ptr, b[i], next.ptr, s);
        std::swap(L, next);
    }
    double2 r;
    d2h(&r, L.ptr, sizeof(double2));
    return zd(r.x, r.y);
}

zd mps_expectation(std::vector<MpoSite const*> const& ops, std::vector<int> const& bonds,
                   std::vector<const double2*> const& sites) {
    hipStream_t s = Context::get().stream;
    const int N = (int)ops.size();
    if (ops.front()->cl != 1 || ops.back()->cr != 1) throw Error("expectation: outer operator bonds must be 1", -66);
    if (bonds.front() != 1 || bonds.back() != 1) throw Error("expectation: outer state bonds must be 1", -67);
    Workspace ws;
    ZBuf L(1), next;
    double2 one = make_double2(1, 0);
    h2d(L.ptr, &one, sizeof(double2));
    for (int i = 0; i < N; ++i) {
        if (i + 1 < N && ops[i]->cr != ops[i + 1]->cl) throw Error("expectation: operator bond mismatch", -65);
        next.resize((size_t)bonds[i + 1] * bonds[i + 1] * ops[i]->cr);
        env_left(bonds[i], bonds[i + 1], *ops[i], L.ptr, sites[i], next.ptr, ws, s);
        std::swap(L, next);
    }
    double2 r;
    d2h(&r, L.ptr, sizeof(double2));
    return zd(r.x, r.y);
}

zd state_overlap(Chain const& a, Chain const& b) {
    // computeStateOverlap (states.hpp:349-382): <a|b>; every site of either chain
    // holds its part of the state whatever the position of its middle site
    if (a.N != b.N) throw Error("state overlap: chains of different lengths", -73);
    for (int i = 0; i < a.N; ++i)
        if (b.ops[i]->d != a.ops[i]->d) throw Error("state overlap: physical dimensions differ", -74);
    std::vector<const double2*> pa, pb;
    for (int i = 0; i < a.N; ++i) {
        pa.push_back(a.sites[i].ptr);
        pb.push_back(b.sites[i].ptr);
    }
    return mps_overlap(a.N, a.dims(), a.bonds, pa, b.bonds, pb);
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
