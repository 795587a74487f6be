// Device-resident sweep driver (see chain.cpp).
#pragma once

#include "nc_internal.hpp"

namespace nc {

struct ChainOptions {
    // chain_options.cpp:21-29 defaults
    double site_tol = 1e-12;
    double sweep_tol = 1e-12;
    double sanity = 1e-12;
    int max_iterations = 10000;
    char which[3] = {'S', 'R', 0};
    int ncv = 3;
    bool rethrow = false;  // raise on optimizer failures instead of skipping the site
    // mixed-precision form: matvecs in complex64 on FP32 matrix cores (and, with
    // env32, the environment updates too; environments stay complex128 in HBM),
    // Krylov vectors, orthogonalisation, Rayleigh quotient and site move in FP64
    bool mixed = false;
    bool env32 = false;
    double env32_floor = 1e-6;  // relative energy noise floor of complex64 environments
    double tol32 = 1e-7;
    bool polish = false;
    // optimizeChain (base_chain.cpp:64-79): grow the bond dimension by one
    // (computeNewBandwidthDimension = chi + 1, chain_options.cpp:28) after each
    // convergence until the energy changes by less than chain_tol; off = fixed chi
    bool grow = false;
    double chain_tol = 1e-12;
};

struct ChainStats {
    int64_t matvecs = 0, site_solves = 0, moves = 0, sweeps = 0, failures = 0, matvecs32 = 0;
    int64_t fp64_refinements = 0;  // mixed form: sites finished in FP64 after a regressive complex64 solve
    double t_opt_ms = 0, t_move_ms = 0, t_env_ms = 0, t_reset_ms = 0;
};

// A lower level kept for excited-state solves (projectors.hpp:324-359): the
// overlap sites (conjugated, (bl, d, br)) of the level's state in left-canonical
// (left[i], i < N-1), centred (middle[i]) and right-canonical (right[i], i > 0)
// form.
struct LevelProjector {
    std::vector<ZBuf> left, middle, right;
    std::vector<int> bonds;
};

// one site solve in preparation: the optimizer's arguments and their buffers
struct SitePrep {
    OptimizeArgs a{};
    ZBuf result;
    DevBuf<float2> L32, R32;
    Projector none;
    Projector const* proj = nullptr;
    double t0 = 0;
    // expectation value of the result, when a batch already computed it
    bool have_ex = false;
    zd ex{0, 0};
};

struct Chain {
    int N;
    std::vector<std::unique_ptr<MpoSite>> ops;
    std::vector<int> bonds;  // N + 1 bond dimensions
    std::vector<ZBuf> sites;
    std::vector<ZBuf> left_envs, right_envs;  // left_envs[i]: sites < i; right_envs[i]: sites >= i
    ZBuf trivial;
    Workspace ws;
    ChainOptions opts;
    ChainStats stats;
    int current = 0;
    bool optimized = false;
    double energy = 0;
    int last_status = 0;

    Chain(int n_sites, std::vector<std::unique_ptr<MpoSite>> ops, std::vector<int> bonds);
    size_t site_size(int i) const;
    void set_raw_site(int i, const double2* host);
    void get_site(int i, double2* host) const;
    void normalize_right_site(int i, hipStream_t s);
    // environment updates (FP64, or complex64 arithmetic in the mixed form)
    void update_left_env(int i, hipStream_t s);
    void update_right_env(int i, hipStream_t s);
    DevBuf<float2> t32a, t32b, t32c;  // complex64 staging for the mixed form
    // excited levels: lower levels, their overlap boundaries with the current
    // state (left_ovl[p][i]: sites < i, right_ovl[p][i]: sites >= i) and the
    // projector matrix of the current site (chain.cpp:240-263)
    std::vector<LevelProjector> projectors;
    std::vector<std::vector<ZBuf>> left_ovl, right_ovl;
    Projector site_projector;
    void update_left_overlaps(int i, hipStream_t s);
    void update_right_overlaps(int i, hipStream_t s);
    void reset_projector_matrix();
    void add_projector_from_state();
    void clear_projectors();
    double projector_overlap();
    void randomize(uint64_t seed);
    // after_level(k), when given, runs once level k has converged (the chain at
    // its first site, before the level becomes a projector)
    std::vector<double> solve_for_multiple_levels(int levels, uint64_t seed,
                                                  std::function<void(int)> const& after_level = nullptr);
    // bond dimensions (utilities.cpp:32-91, utilities.hpp:264-283)
    int chi = 1, initial_chi = 1;
    std::vector<int> dims() const;
    void set_bandwidth(int new_chi);            // new bond sequence, sites reallocated (contents undefined)
    void increase_bandwidth(int new_chi);       // increaseBandwidthDimension (chain.cpp:75-121)
    void optimize_chain();                      // optimizeChain (base_chain.cpp:64-79)
    void restart_level(uint64_t seed);          // new random state at the initial bond dimension, reset
    void reset();
    zd site_expectation();
    void move_right();
    void move_left();
    int optimize_site();
    void prepare_site(SitePrep& p);
    int finish_site(SitePrep& p, OptimizeOut const& o);
    int check_site(SitePrep& p, OptimizeOut const& o);
    bool converged_flag = false;  // used by the lockstep driver
    void sweep();
    void sweep_until_converged();
};

// Independent chains (same length) sweeping in lockstep with batched site
// solves; the chains are owned by the caller.
struct ChainBatch {
    std::vector<Chain*> chains;
    int64_t lockstep_rounds = 0;
    explicit ChainBatch(std::vector<Chain*> cs);
    double t_moves_ms = 0, t_solves_ms = 0;
    // parts of the site solves: per-chain preparation, the batched solve, per-chain post-checks
    double t_prepare_ms = 0, t_batched_ms = 0, t_finish_ms = 0, t_krylov_ms = 0;
    void optimize_sites(std::vector<Chain*> const& active);
    void move_all(std::vector<Chain*> const& cs, int dir);
    void env_left_batch(std::vector<Chain*> const& cs, hipStream_t s);
    void env_right_batch(std::vector<Chain*> const& cs, hipStream_t s);
    void sweep(std::vector<Chain*> const& active);
    void sweep_all();
    void sweep_until_converged();
    // Chain::reset of every chain with grouped launches (the chains hold raw
    // sites, as after nc_chain_set_raw_site)
    void reset_all();
};

std::vector<int> bandwidth_sequence(int requested, std::vector<int> const& dims);
int maximum_bandwidth(std::vector<int> const& dims);
int minimum_bandwidth_for_projectors(std::vector<int> const& dims, int nproj);

// <a|b> of two chains' states (computeStateOverlap, states.hpp:349-382)
zd state_overlap(Chain const& a, Chain const& b);
// <a|b> of two matrix product states given as device sites (br, bl, d) with
// their bond dimensions (N + 1 each) and physical dimensions
zd mps_overlap(int N, std::vector<int> const& d, std::vector<int> const& bonds_a,
               std::vector<const double2*> const& a, std::vector<int> const& bonds_b,
               std::vector<const double2*> const& b);
// <psi|H|psi> of a matrix product state with an MPO (contract_sos_left from the
// trivial boundary, core.f95:623-677, as computeExpectationValue does)
zd mps_expectation(std::vector<MpoSite const*> const& ops, std::vector<int> const& bonds,
                   std::vector<const double2*> const& sites);

}  // namespace nc
