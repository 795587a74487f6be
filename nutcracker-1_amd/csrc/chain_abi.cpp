// C-ABI of the device-resident sweep driver (no counterpart in the reference's
// Fortran core: it replaces the host loop of chain.cpp / base_chain.cpp that
// would otherwise cross the ABI once per contraction).
#include "nc_chain.hpp"
#include "../../include/nutcracker_core_abi.h"

#include <algorithm>
#include <cstring>

using namespace nc;

struct nc_chain {
    std::unique_ptr<Chain> chain;
};

struct nc_chain_batch {
    std::unique_ptr<ChainBatch> batch;
};

namespace {
template <class F>
int guarded_chain(const char* name, F&& f) {
    try {
        return f();
    } catch (Error const& e) {
        set_last_error(std::string(name) + ": " + e.what());
        fprintf(stderr, "[nutcracker-hip] %s: %s\n", name, e.what());
        return e.code != 0 ? e.code : -1;
    } catch (std::exception const& e) {
        set_last_error(std::string(name) + ": " + e.what());
        fprintf(stderr, "[nutcracker-hip] %s: %s\n", name, e.what());
        return -1;
    }
}
}  // namespace

extern "C" {

nc_chain* nc_chain_create(uint32_t n_sites, const uint32_t* d, const uint32_t* cl, const uint32_t* cr,
                          const uint32_t* nmat, const uint32_t* idx_all, const nc_zcomplex* mats_all,
                          const uint32_t* bonds) {
    nc_chain* out = nullptr;
    guarded_chain("nc_chain_create", [&] {
        std::vector<std::unique_ptr<MpoSite>> ops;
        size_t io = 0, mo = 0;
        for (uint32_t i = 0; i < n_sites; ++i) {
            auto op = std::make_unique<MpoSite>();
            op->build((int)d[i], (int)cl[i], (int)cr[i], (int)nmat[i], idx_all + io,
                      reinterpret_cast<const zd*>(mats_all) + mo);
            io += 2 * (size_t)nmat[i];
            mo += (size_t)d[i] * d[i] * nmat[i];
            ops.push_back(std::move(op));
        }
        for (uint32_t i = 0; i + 1 < n_sites; ++i)
            if (cr[i] != cl[i + 1]) throw Error("operator bond mismatch between sites", -65);
        if (cl[0] != 1 || cr[n_sites - 1] != 1) throw Error("the chain's outer operator bonds must be 1", -66);
        std::vector<int> b(bonds, bonds + n_sites + 1);
        if (b.front() != 1 || b.back() != 1) throw Error("the chain's outer state bonds must be 1", -67);
        auto* c = new nc_chain();
        c->chain = std::make_unique<Chain>((int)n_sites, std::move(ops), std::move(b));
        out = c;
        return 0;
    });
    return out;
}

void nc_chain_destroy(nc_chain* c) { delete c; }

int nc_chain_set_options(nc_chain* c, double site_tol, double sweep_tol, double sanity, uint32_t max_iterations,
                         const char* which, uint32_t ncv, int rethrow) {
    return guarded_chain("nc_chain_set_options", [&] {
        auto& o = c->chain->opts;
        o.site_tol = site_tol;
        o.sweep_tol = sweep_tol;
        o.sanity = sanity;
        o.max_iterations = (int)max_iterations;
        if (!(which[0] == 'S' && which[1] == 'R') && !(which[0] == 'L' && (which[1] == 'R' || which[1] == 'M')))
            throw Error("which must be SR, LR or LM", -68);
        o.which[0] = which[0];
        o.which[1] = which[1];
        o.ncv = (int)(ncv < 3 ? 3 : ncv);
        o.rethrow = rethrow != 0;
        return 0;
    });
}

int nc_chain_set_mixed(nc_chain* c, int mixed, int env32, double tol32, int polish) {
    return guarded_chain("nc_chain_set_mixed", [&] {
        auto& o = c->chain->opts;
        o.mixed = mixed != 0;
        o.env32 = env32 != 0;
        o.tol32 = tol32;
        o.polish = polish != 0;
        return 0;
    });
}

int nc_chain_set_raw_site(nc_chain* c, uint32_t i, const nc_zcomplex* data) {
    return guarded_chain("nc_chain_set_raw_site", [&] {
        if ((int)i >= c->chain->N) throw Error("site index out of range", -69);
        c->chain->set_raw_site((int)i, reinterpret_cast<const double2*>(data));
        return 0;
    });
}

int nc_chain_get_site(nc_chain* c, uint32_t i, nc_zcomplex* out) {
    return guarded_chain("nc_chain_get_site", [&] {
        if ((int)i >= c->chain->N) throw Error("site index out of range", -69);
        c->chain->get_site((int)i, reinterpret_cast<double2*>(out));
        return 0;
    });
}

int nc_chain_get_environment(nc_chain* c, uint32_t i, int side, nc_zcomplex* out) {
    return guarded_chain("nc_chain_get_environment", [&] {
        Chain& ch = *c->chain;
        if ((int)i >= ch.N) throw Error("site index out of range", -69);
        if (side == 0) {
            const size_t n = (size_t)ch.bonds[i] * ch.bonds[i] * ch.ops[i]->cl;
            d2h(out, ch.left_envs[i].ptr, n * sizeof(double2));
        } else {
            const size_t n = (size_t)ch.bonds[i + 1] * ch.bonds[i + 1] * ch.ops[i]->cr;
            d2h(out, ch.right_envs[i + 1].ptr, n * sizeof(double2));
        }
        return 0;
    });
}

int nc_chain_reset(nc_chain* c) {
    return guarded_chain("nc_chain_reset", [&] { c->chain->reset(); return 0; });
}

int nc_chain_optimize_site(nc_chain* c) {
    return guarded_chain("nc_chain_optimize_site", [&] { return c->chain->optimize_site(); });
}

int nc_chain_move(nc_chain* c, int direction) {
    return guarded_chain("nc_chain_move", [&] {
        if (direction > 0) c->chain->move_right();
        else c->chain->move_left();
        return 0;
    });
}

int nc_chain_sweep(nc_chain* c) {
    return guarded_chain("nc_chain_sweep", [&] { c->chain->sweep(); return 0; });
}

int nc_chain_sweep_until_converged(nc_chain* c) {
    return guarded_chain("nc_chain_sweep_until_converged", [&] { c->chain->sweep_until_converged(); return 0; });
}

int nc_chain_energy(nc_chain* c, double* energy, int* current_site) {
    return guarded_chain("nc_chain_energy", [&] {
        *energy = c->chain->energy;
        if (current_site) *current_site = c->chain->current;
        return 0;
    });
}

int nc_chain_expectation(nc_chain* c, nc_zcomplex* out) {
    return guarded_chain("nc_chain_expectation", [&] {
        zd e = c->chain->site_expectation();
        out->re = e.real();
        out->im = e.imag();
        return 0;
    });
}

int nc_chain_set_growth(nc_chain* c, int grow, double chain_tol) {
    return guarded_chain("nc_chain_set_growth", [&] {
        c->chain->opts.grow = grow != 0;
        c->chain->opts.chain_tol = chain_tol;
        return 0;
    });
}

int nc_chain_increase_bandwidth(nc_chain* c, uint32_t new_chi) {
    return guarded_chain("nc_chain_increase_bandwidth", [&] { c->chain->increase_bandwidth((int)new_chi); return 0; });
}

int nc_chain_optimize_chain(nc_chain* c) {
    return guarded_chain("nc_chain_optimize_chain", [&] { c->chain->optimize_chain(); return 0; });
}

int nc_chain_get_bonds(nc_chain* c, uint32_t* bonds) {
    return guarded_chain("nc_chain_get_bonds", [&] {
        for (int i = 0; i <= c->chain->N; ++i) bonds[i] = (uint32_t)c->chain->bonds[i];
        return 0;
    });
}

int nc_chain_add_projector(nc_chain* c) {
    return guarded_chain("nc_chain_add_projector", [&] { c->chain->add_projector_from_state(); return 0; });
}

int nc_chain_clear_projectors(nc_chain* c) {
    return guarded_chain("nc_chain_clear_projectors", [&] { c->chain->clear_projectors(); return 0; });
}

int nc_chain_projector_overlap(nc_chain* c, double* overlap) {
    return guarded_chain("nc_chain_projector_overlap", [&] { *overlap = c->chain->projector_overlap(); return 0; });
}

int nc_chain_randomize(nc_chain* c, uint64_t seed) {
    return guarded_chain("nc_chain_randomize", [&] { c->chain->randomize(seed); return 0; });
}

int nc_chain_restart_level(nc_chain* c, uint64_t seed) {
    return guarded_chain("nc_chain_restart_level", [&] { c->chain->restart_level(seed); return 0; });
}

int nc_chain_solve_for_multiple_levels(nc_chain* c, uint32_t levels, uint64_t seed, double* energies) {
    return guarded_chain("nc_chain_solve_for_multiple_levels", [&] {
        std::vector<double> e = c->chain->solve_for_multiple_levels((int)levels, seed);
        for (size_t k = 0; k < e.size(); ++k) energies[k] = e[k];
        return 0;
    });
}

int nc_chain_state_overlap(nc_chain* a, nc_chain* b, nc_zcomplex* out) {
    return guarded_chain("nc_chain_state_overlap", [&] {
        zd v = state_overlap(*a->chain, *b->chain);
        out->re = v.real();
        out->im = v.imag();
        return 0;
    });
}

nc_chain_batch* nc_chain_batch_create(nc_chain* const* chains, uint32_t n) {
    nc_chain_batch* out = nullptr;
    guarded_chain("nc_chain_batch_create", [&] {
        std::vector<Chain*> cs;
        for (uint32_t i = 0; i < n; ++i) cs.push_back(chains[i]->chain.get());
        auto* b = new nc_chain_batch();
        b->batch = std::make_unique<ChainBatch>(std::move(cs));
        out = b;
        return 0;
    });
    return out;
}

void nc_chain_batch_destroy(nc_chain_batch* b) { delete b; }

int nc_chain_batch_times(nc_chain_batch* b, double* times_ms) {
    return guarded_chain("nc_chain_batch_times", [&] {
        times_ms[0] = b->batch->t_moves_ms;
        times_ms[1] = b->batch->t_solves_ms;
        times_ms[2] = b->batch->t_prepare_ms;
        times_ms[3] = b->batch->t_batched_ms;
        times_ms[4] = b->batch->t_finish_ms;
        times_ms[5] = b->batch->t_krylov_ms;
        return 0;
    });
}

int nc_chain_batch_reset(nc_chain_batch* b) {
    return guarded_chain("nc_chain_batch_reset", [&] {
        b->batch->reset_all();
        return 0;
    });
}

int nc_chain_batch_sweep(nc_chain_batch* b) {
    return guarded_chain("nc_chain_batch_sweep", [&] { b->batch->sweep_all(); return 0; });
}

int nc_chain_batch_sweep_until_converged(nc_chain_batch* b) {
    return guarded_chain("nc_chain_batch_sweep_until_converged", [&] { b->batch->sweep_until_converged(); return 0; });
}

int nc_chain_stats(nc_chain* c, int64_t* counts, double* times_ms) {
    return guarded_chain("nc_chain_stats", [&] {
        auto const& s = c->chain->stats;
        counts[0] = s.matvecs;
        counts[1] = s.site_solves;
        counts[2] = s.moves;
        counts[3] = s.sweeps;
        counts[4] = s.failures;
        counts[5] = s.matvecs32;
        counts[6] = s.fp64_refinements;
        times_ms[0] = s.t_opt_ms;
        times_ms[1] = s.t_move_ms;
        times_ms[2] = s.t_env_ms;
        times_ms[3] = s.t_reset_ms;
        return 0;
    });
}

int nc_set_autotune(int on) {
    const int prev = nc::gemm_autotune() ? 1 : 0;
    nc::gemm_autotune() = on != 0;
    return prev;
}

int nc_gemm_force_choice(int bk, int nsplit, int m3) {
    if (bk != 0 && bk != 16 && bk != 32 && bk != -16 && bk != 6 && bk != 8 && bk != 9) return -1;
    if (nsplit < 0 || nsplit > 8 || (nsplit & (nsplit - 1))) return -1;
    nc::gemm_force_choice(bk, nsplit, m3 < 0 ? -1 : (m3 ? 1 : 0));
    return 0;
}

int64_t nc_autotune_table(char* buf, int64_t len) {
    const std::string t = nc::gemm_tune_table();
    if (buf && len > 0) {
        const size_t n = std::min<size_t>((size_t)len - 1, t.size());
        std::memcpy(buf, t.data(), n);
        buf[n] = 0;
    }
    return (int64_t)t.size() + 1;
}

int nc_set_krylov_graphs(int on) {
    const int prev = nc::krylov_graphs() ? 1 : 0;
    nc::krylov_graphs() = on != 0;
    return prev;
}

}  // extern "C"
