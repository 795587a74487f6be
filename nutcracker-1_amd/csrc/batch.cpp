// Many independent chains advancing through their sweeps in lockstep (a
// parameter sweep of couplings, each chain with its own bond dimension).
//
// The chains keep their own sites, environments and workspaces; at every step
// of the sweep (chain.cpp:147-163) all of them sit on the same site, and their
// site solves run together (optimize_sites_batched: one grouped launch per
// GEMM step for all chains, one orthogonalisation launch per Arnoldi kernel).
// Moves stay per chain.  A chain whose energy has converged
// (sweepUntilConverged, base_chain.cpp:112-123) leaves the lockstep; the others
// go on, so every chain performs exactly the sweeps it would perform alone.
#include "nc_chain.hpp"

namespace nc {

namespace {
bool outside_tol(double a, double b, double t) {
    // utilities.hpp:418-420
    return ((std::abs(a) + std::abs(b)) / 2 > t) && (std::abs(a - b) / (std::abs(a) + std::abs(b) + t) > t);
}
}  // namespace

ChainBatch::ChainBatch(std::vector<Chain*> cs) : chains(std::move(cs)) {
    if (chains.empty()) throw Error("chain batch: no chains", -80);
    for (Chain* c : chains) {
        if (c->N != chains[0]->N) throw Error("chain batch: chains of different lengths", -81);
        if (!c->projectors.empty()) throw Error("chain batch: chains with lower levels are not batched", -82);
    }
}

void ChainBatch::optimize_sites(std::vector<Chain*> const& active) {
    hipStream_t s = Context::get().stream;
    std::vector<std::unique_ptr<SitePrep>> preps;
    std::vector<OptimizeArgs const*> args;
    std::vector<Workspace*> wss;
    for (Chain* c : active) {
        preps.emplace_back(new SitePrep());
        c->prepare_site(*preps.back());
        args.push_back(&preps.back()->a);
        wss.push_back(&c->ws);
    }
    std::vector<OptimizeOut> outs;
    optimize_sites_batched(args, wss, outs, s);
    for (size_t q = 0; q < active.size(); ++q) {
        Chain* c = active[q];
        int st = c->finish_site(*preps[q], outs[q]);
        if (st != 0 && c->opts.rethrow)
            throw Error("optimizer failure at site " + std::to_string(c->current) + " of a batched chain", st);
    }
}

void ChainBatch::sweep(std::vector<Chain*> const& active) {
    // performOptimizationSweep (chain.cpp:147-163) for every active chain at once
    if (active.empty()) return;
    const int N = active[0]->N;
    for (Chain* c : active)
        if (c->current != active[0]->current) throw Error("chain batch: chains are not on the same site", -83);
    const int start = active[0]->current;
    std::vector<Chain*> todo;
    for (Chain* c : active)
        if (!c->optimized) todo.push_back(c);
    optimize_sites(todo);
    auto step = [&](int dir) {
        for (Chain* c : active) {
            if (dir > 0) c->move_right();
            else c->move_left();
        }
        optimize_sites(active);
    };
    while (active[0]->current + 1 < N) step(+1);
    while (active[0]->current > 0) step(-1);
    while (active[0]->current < start) step(+1);
    for (Chain* c : active) ++c->stats.sweeps;
}

void ChainBatch::sweep_all() { sweep(chains); }

void ChainBatch::sweep_until_converged() {
    // sweepUntilConverged (base_chain.cpp:112-123) per chain
    std::vector<double> prev(chains.size()), cur(chains.size());
    for (size_t c = 0; c < chains.size(); ++c) prev[c] = chains[c]->energy;
    sweep(chains);
    for (size_t c = 0; c < chains.size(); ++c) cur[c] = chains[c]->energy;
    for (;;) {
        std::vector<Chain*> active;
        std::vector<size_t> ids;
        for (size_t c = 0; c < chains.size(); ++c) {
            Chain* ch = chains[c];
            const double tol = ch->opts.mixed && ch->opts.env32 ? std::max(ch->opts.sweep_tol, ch->opts.env32_floor)
                                                                : ch->opts.sweep_tol;
            if (outside_tol(prev[c], cur[c], tol)) {
                active.push_back(ch);
                ids.push_back(c);
            }
        }
        if (active.empty()) break;
        sweep(active);
        for (size_t q = 0; q < ids.size(); ++q) {
            prev[ids[q]] = cur[ids[q]];
            cur[ids[q]] = chains[ids[q]]->energy;
        }
        ++lockstep_rounds;
    }
}

}  // namespace nc
