// Many independent chains advancing through their sweeps in lockstep (a
// parameter sweep of couplings, each chain with its own bond dimension).
//
// The chains keep their own sites, environments and workspaces; at every step
// of the sweep (chain.cpp:147-163) all of them sit on the same site, and their
// site solves run together (optimize_sites_batched: one grouped launch per
// GEMM step for all chains, one orthogonalisation launch per Arnoldi kernel),
// and so do their moves (one batched Newton-Schulz polar split, grouped GEMMs
// for the neighbour and the environment update).  A chain whose energy has converged
// (sweepUntilConverged, base_chain.cpp:112-123) leaves the lockstep; the others
// go on, so every chain performs exactly the sweeps it would perform alone.
#include "nc_chain.hpp"

#include <chrono>

namespace nc {

namespace {
double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

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

namespace {
// contract_sos_right for several chains in grouped launches: out = R' from R
// (br x br x cr) and the site psi (br, bl, d), out bl x bl x cl
struct EnvRightItem {
    MpoSite const* op;
    int bl, br;
    const double2* R;
    const double2* psi;
    double2* out;
    Workspace* ws;
};
void grouped_env_right(std::vector<EnvRightItem> const& items, hipStream_t s) {
    GemmPlan A, Cp;
    A.M = A.N = Cp.M = Cp.N = 1;
    Cp.opA = OpC; Cp.opB = OpN;
    A.key.tag = 'p';
    Cp.key.tag = 'q';
    A.key.count = Cp.key.count = (int)items.size();
    for (auto const& it : items) {
        for (TuneKey* k : {&A.key, &Cp.key}) {
            k->bl = std::max(k->bl, it.bl);
            k->br = std::max(k->br, it.br);
            k->d = std::max(k->d, it.op->d);
            k->cl = std::max(k->cl, it.op->cl);
            k->cr = std::max(k->cr, it.op->cr);
        }
    }
    std::vector<MixItem<double2>> mix;
    for (auto const& it : items) {
        MpoSite const& op = *it.op;
        const int bl = it.bl, br = it.br, d = op.d;
        const int64_t rs = (int64_t)br * bl;
        Workspace& w = *it.ws;
        w.ensure(rs * d * op.tos.size(), rs * d * op.froms.size());
        for (size_t ti = 0; ti < op.tos.size(); ++ti)
            for (int a = 0; a < d; ++a) {
                A.begin_batch(w.P.ptr + ((int64_t)ti * d + a) * rs, br, br, bl);
                A.add_seg(it.R + (int64_t)op.tos[ti] * br * br, br, it.psi + a * rs, br, br);
            }
        mix.push_back(MixItem<double2>{&op, w.P.ptr, w.Q.ptr, rs, false});
        if ((int)op.froms.size() < op.cl) dzero(it.out, (size_t)bl * bl * op.cl * sizeof(double2));
        for (size_t fi = 0; fi < op.froms.size(); ++fi) {
            Cp.begin_batch(it.out + (int64_t)op.froms[fi] * bl * bl, bl, bl, bl);
            for (int b = 0; b < d; ++b)
                Cp.add_seg(it.psi + b * rs, br, w.Q.ptr + ((int64_t)fi * d + b) * rs, br, br);
        }
    }
    zgemm_run(A, s);
    mix_blocks_batch(mix, s);
    zgemm_run(Cp, s);
}
}  // namespace

void ChainBatch::optimize_sites(std::vector<Chain*> const& active) {
    hipStream_t s = Context::get().stream;
    std::vector<std::unique_ptr<SitePrep>> preps;
    std::vector<OptimizeArgs const*> args;
    std::vector<Workspace*> wss;
    double t0 = now_ms();
    for (Chain* c : active) {
        preps.emplace_back(new SitePrep());
        c->prepare_site(*preps.back());
        args.push_back(&preps.back()->a);
        wss.push_back(&c->ws);
    }
    double t1 = now_ms();
    t_prepare_ms += t1 - t0;
    std::vector<OptimizeOut> outs;
    optimize_sites_batched(args, wss, outs, s, &t_krylov_ms);
    t0 = now_ms();
    t_batched_ms += t0 - t1;
    // the expectation values of the post-checks (compute_expectation, core.f95:
    // 1045-1088) of every solved chain: one grouped environment contraction and
    // one batched trace
    {
        std::vector<EnvRightItem> items;
        std::vector<ZBuf> Ns;
        std::vector<size_t> who;
        for (size_t q = 0; q < active.size(); ++q) {
            if (outs[q].info != 0) continue;
            Chain* c = active[q];
            const int i = c->current;
            MpoSite const& op = *c->ops[i];
            const int bl = c->bonds[i], br = c->bonds[i + 1];
            Ns.emplace_back((size_t)bl * bl * op.cl);
            items.push_back(EnvRightItem{&op, bl, br, c->right_envs[i + 1].ptr, preps[q]->result.ptr, Ns.back().ptr,
                                         &c->ws});
            who.push_back(q);
        }
        grouped_env_right(items, s);
        std::vector<RedItem> tr;
        for (size_t k = 0; k < who.size(); ++k) {
            Chain* c = active[who[k]];
            const int i = c->current;
            tr.push_back(RedItem{Ns[k].ptr, c->left_envs[i].ptr, (int64_t)c->bonds[i], (int64_t)c->ops[i]->cl, true});
        }
        const std::vector<zd> ex = reduce_batch(tr, s);
        for (size_t k = 0; k < who.size(); ++k) {
            preps[who[k]]->have_ex = true;
            preps[who[k]]->ex = ex[k];
        }
    }
    for (size_t q = 0; q < active.size(); ++q) {
        Chain* c = active[q];
        int st = c->finish_site(*preps[q], outs[q]);
        if (st != 0 && c->opts.rethrow)
            throw Error("optimizer failure at site " + std::to_string(c->current) + " of a batched chain", st);
    }
    t_finish_ms += now_ms() - t0;
}

// ---------------------------------------------------------------------------
// Moves of all chains together: one batched polar split, grouped GEMMs for the
// neighbour and for the environment update (states.cpp:94-164 via
// norm_denorm_going_*, boundaries.cpp:145-189 via contract_sos_*).
// ---------------------------------------------------------------------------
void ChainBatch::env_left_batch(std::vector<Chain*> const& cs, hipStream_t s) {
    // L_{i+1} = contract_sos_left(L_i, site i), i = current - 1 (after the move)
    GemmPlan A, Cp;
    A.M = A.N = Cp.M = Cp.N = 1;
    Cp.opA = OpN; Cp.opB = OpC;
    A.key.tag = 'a';
    Cp.key.tag = 'l';
    A.key.count = Cp.key.count = (int)cs.size();
    for (Chain* c : cs) {
        const int i = c->current - 1;
        for (TuneKey* k : {&A.key, &Cp.key}) {
            k->bl = std::max(k->bl, c->bonds[i]);
            k->br = std::max(k->br, c->bonds[i + 1]);
            k->d = std::max(k->d, c->ops[i]->d);
            k->cl = std::max(k->cl, c->ops[i]->cl);
            k->cr = std::max(k->cr, c->ops[i]->cr);
        }
    }
    std::vector<MixItem<double2>> mix;
    for (Chain* c : cs) {
        const int i = c->current - 1;
        MpoSite const& op = *c->ops[i];
        const int bl = c->bonds[i], br = c->bonds[i + 1], d = op.d;
        const int64_t rs = (int64_t)br * bl;
        c->ws.ensure(rs * d * op.froms.size(), rs * d * op.tos.size());
        c->left_envs[i + 1].resize((size_t)br * br * op.cr);
        const double2* psi = c->sites[i].ptr;
        for (size_t fi = 0; fi < op.froms.size(); ++fi)
            for (int a = 0; a < d; ++a) {
                A.begin_batch(c->ws.P.ptr + ((int64_t)fi * d + a) * rs, br, br, bl);
                A.add_seg(psi + a * rs, br, c->left_envs[i].ptr + (int64_t)op.froms[fi] * bl * bl, bl, bl);
            }
        mix.push_back(MixItem<double2>{&op, c->ws.P.ptr, c->ws.Q.ptr, rs, true});
        if ((int)op.tos.size() < op.cr) dzero(c->left_envs[i + 1].ptr, (size_t)br * br * op.cr * sizeof(double2));
        for (size_t ti = 0; ti < op.tos.size(); ++ti) {
            Cp.begin_batch(c->left_envs[i + 1].ptr + (int64_t)op.tos[ti] * br * br, br, br, br);
            Cp.add_seg(c->ws.Q.ptr + (int64_t)ti * d * rs, br, psi, br, bl * d);
        }
    }
    zgemm_run(A, s);
    mix_blocks_batch(mix, s);
    zgemm_run(Cp, s);
}


// reset (chain.cpp:175-237) of every chain at once: each site i >= 1 is replaced
// by its right-normalised polar factor (independently of the others, as
// Chain::reset does), so all of them -- every site of every chain -- are one
// batched polar split; the right environments then follow site by site with
// grouped contractions, the first sites are normalised and the energies read
// with batched reductions.
void ChainBatch::reset_all() {
    hipStream_t s = Context::get().stream;
    bool plain = true;  // complex64 environments take the per-chain path
    for (Chain* c : chains)
        if (c->opts.mixed && c->opts.env32) plain = false;
    if (!plain) {
        for (Chain* c : chains) c->reset();
        return;
    }
    const double t0 = now_ms();
    const int N = chains[0]->N;
    for (Chain* c : chains) {
        c->right_envs.clear();
        c->left_envs.clear();
        c->right_envs.resize(N + 1);
        c->left_envs.resize(N + 1);
        c->right_envs[N].resize(1);
        d2d(c->right_envs[N].ptr, c->trivial.ptr, sizeof(double2));
        c->left_envs[0].resize(1);
        d2d(c->left_envs[0].ptr, c->trivial.ptr, sizeof(double2));
        c->left_ovl.clear();
        c->right_ovl.clear();
    }
    // every right normalisation in one batched polar split
    std::vector<PolarJob> jobs;
    std::vector<ZBuf> W, Up, Pc;
    struct Where { Chain* c; int i; };
    std::vector<Where> where;
    for (Chain* c : chains)
        for (int i = N - 1; i >= 1; --i) {
            const int bl = c->bonds[i], br = c->bonds[i + 1], d = c->ops[i]->d;
            if ((int64_t)br * d < bl) throw Error("Not enough degrees of freedom to normalize.", -61);
            where.push_back(Where{c, i});
        }
    W.resize(where.size());
    Up.resize(where.size());
    Pc.resize(where.size());
    for (size_t k = 0; k < where.size(); ++k) {
        Chain* c = where[k].c;
        const int i = where[k].i;
        const int bl = c->bonds[i], br = c->bonds[i + 1], d = c->ops[i]->d;
        const size_t n = (size_t)bl * br * d;
        W[k].resize(n);
        Up[k].resize(n);
        Pc[k].resize((size_t)bl * bl);
        permute_102(c->sites[i].ptr, W[k].ptr, br, bl, d, s);
        jobs.push_back(PolarJob{bl, (int64_t)br * d, W[k].ptr, Up[k].ptr, Pc[k].ptr, 0});
    }
    polar_split_batch(jobs, s);
    for (size_t k = 0; k < where.size(); ++k) {
        if (jobs[k].info != 0) throw Error("NormalizationError", jobs[k].info);
        Chain* c = where[k].c;
        const int i = where[k].i;
        permute_102(Up[k].ptr, c->sites[i].ptr, c->bonds[i], c->bonds[i + 1], c->ops[i]->d, s);
    }
    W.clear();
    Up.clear();
    Pc.clear();
    // right environments, site by site for all chains
    for (int i = N - 1; i >= 1; --i) {
        std::vector<EnvRightItem> items;
        for (Chain* c : chains) {
            MpoSite const& op = *c->ops[i];
            c->right_envs[i].resize((size_t)c->bonds[i] * c->bonds[i] * op.cl);
            items.push_back(EnvRightItem{&op, c->bonds[i], c->bonds[i + 1], c->right_envs[i + 1].ptr,
                                         c->sites[i].ptr, c->right_envs[i].ptr, &c->ws});
        }
        grouped_env_right(items, s);
    }
    // first sites divided by their norms (core.f95:2661-2671)
    std::vector<const double2*> x0;
    std::vector<double2*> x0w;
    std::vector<int64_t> n0;
    for (Chain* c : chains) {
        x0.push_back(c->sites[0].ptr);
        x0w.push_back(c->sites[0].ptr);
        n0.push_back((int64_t)c->site_size(0));
    }
    const std::vector<double> ss = sumsq_batch(x0, n0, s);
    std::vector<double> f;
    for (double v : ss) f.push_back(1.0 / std::sqrt(v));
    scale_batch(x0w, n0, f, s);
    // energies: the expectation value at the first site
    std::vector<EnvRightItem> items;
    std::vector<ZBuf> Ns(chains.size());
    for (size_t q = 0; q < chains.size(); ++q) {
        Chain* c = chains[q];
        MpoSite const& op = *c->ops[0];
        Ns[q].resize((size_t)c->bonds[0] * c->bonds[0] * op.cl);
        items.push_back(EnvRightItem{&op, c->bonds[0], c->bonds[1], c->right_envs[1].ptr, c->sites[0].ptr,
                                     Ns[q].ptr, &c->ws});
    }
    grouped_env_right(items, s);
    std::vector<RedItem> tr;
    for (size_t q = 0; q < chains.size(); ++q) {
        Chain* c = chains[q];
        tr.push_back(RedItem{Ns[q].ptr, c->left_envs[0].ptr, (int64_t)c->bonds[0], (int64_t)c->ops[0]->cl, true});
    }
    const std::vector<zd> e = reduce_batch(tr, s);
    const double share = (now_ms() - t0) / chains.size();
    for (size_t q = 0; q < chains.size(); ++q) {
        Chain* c = chains[q];
        c->current = 0;
        c->optimized = false;
        const double imag_tol = c->opts.mixed && c->opts.env32 ? 1e-6 : 1e-10;
        if (std::abs(e[q].imag()) / std::abs(e[q]) > imag_tol) throw Error("InitialChainEnergyNotRealError", -62);
        c->energy = e[q].real();
        c->stats.t_reset_ms += share;
    }
}

void ChainBatch::env_right_batch(std::vector<Chain*> const& cs, hipStream_t s) {
    // R_i = contract_sos_right(R_{i+1}, site i), i = current + 1 (after the move)
    std::vector<EnvRightItem> items;
    for (Chain* c : cs) {
        const int i = c->current + 1;
        MpoSite const& op = *c->ops[i];
        c->right_envs[i].resize((size_t)c->bonds[i] * c->bonds[i] * op.cl);
        items.push_back(EnvRightItem{&op, c->bonds[i], c->bonds[i + 1], c->right_envs[i + 1].ptr, c->sites[i].ptr,
                                     c->right_envs[i].ptr, &c->ws});
    }
    grouped_env_right(items, s);
}

void ChainBatch::move_all(std::vector<Chain*> const& cs, int dir) {
    hipStream_t s = Context::get().stream;
    bool plain = true;  // complex64 environments and lower levels take the per-chain path
    for (Chain* c : cs)
        if ((c->opts.mixed && c->opts.env32) || !c->projectors.empty()) plain = false;
    if (!plain) {
        for (Chain* c : cs) {
            if (dir > 0) c->move_right();
            else c->move_left();
        }
        return;
    }
    for (Chain* c : cs) {
        if (dir > 0 && c->current + 1 >= c->N) throw Error("move_right past the last site", -63);
        if (dir < 0 && c->current == 0) throw Error("move_left past the first site", -64);
    }
    std::vector<PolarJob> jobs(cs.size());
    std::vector<ZBuf> Up(cs.size()), Pc(cs.size()), W1(cs.size()), W2(cs.size()), moved(cs.size());
    for (size_t q = 0; q < cs.size(); ++q) {
        Chain* c = cs[q];
        const int i = c->current;
        if (dir > 0) {
            // the site (bm, bl, dl) is the wide bm x (bl dl) matrix to normalise
            const int bl = c->bonds[i], bm = c->bonds[i + 1];
            const int dl = c->ops[i]->d;
            if ((int64_t)bl * dl < bm) throw Error("Not enough degrees of freedom to normalize.", -54);
            Up[q].resize(c->site_size(i));
            Pc[q].resize((size_t)bm * bm);
            jobs[q] = PolarJob{bm, (int64_t)bl * dl, c->sites[i].ptr, Up[q].ptr, Pc[q].ptr, 0};
        } else {
            // the site (br, bm, dr) as bm x (br dr)
            const int bm = c->bonds[i], br = c->bonds[i + 1];
            const int dr = c->ops[i]->d;
            if ((int64_t)br * dr < bm) throw Error("Not enough degrees of freedom to normalize.", -54);
            W1[q].resize(c->site_size(i));
            W2[q].resize(c->site_size(i));
            Pc[q].resize((size_t)bm * bm);
            permute_102(c->sites[i].ptr, W1[q].ptr, br, bm, dr, s);
            jobs[q] = PolarJob{bm, (int64_t)br * dr, W1[q].ptr, W2[q].ptr, Pc[q].ptr, 0};
        }
    }
    polar_split_batch(jobs, s);
    for (size_t q = 0; q < cs.size(); ++q)
        if (jobs[q].info != 0) throw Error("NormalizationError", jobs[q].info);
    // the neighbour absorbs conj(P)
    GemmPlan g;
    g.M = g.N = 1;
    std::vector<ZBuf> T(cs.size());
    for (size_t q = 0; q < cs.size(); ++q) {
        Chain* c = cs[q];
        const int i = c->current;
        if (dir > 0) {
            const int bm = c->bonds[i + 1], br = c->bonds[i + 2], dr = c->ops[i + 1]->d;
            T[q].resize(c->site_size(i + 1));
            W2[q].resize(c->site_size(i + 1));
            permute_102(c->sites[i + 1].ptr, T[q].ptr, br, bm, dr, s);  // (br, bm, dr) -> (bm, br, dr)
            g.begin_batch(W2[q].ptr, bm, bm, br * dr);
            g.add_seg(Pc[q].ptr, bm, T[q].ptr, bm, bm);
        } else {
            const int bll = c->bonds[i - 1], bm = c->bonds[i], dl = c->ops[i - 1]->d;
            moved[q].resize(c->site_size(i - 1));
            g.begin_batch(moved[q].ptr, bm, bm, bll * dl);
            g.add_seg(Pc[q].ptr, bm, c->sites[i - 1].ptr, bm, bm);
        }
    }
    zgemm_run(g, s);
    std::vector<double2*> mid;
    std::vector<const double2*> midc;
    std::vector<int64_t> mn;
    for (size_t q = 0; q < cs.size(); ++q) {
        Chain* c = cs[q];
        const int i = c->current;
        if (dir > 0) {
            const int bm = c->bonds[i + 1], br = c->bonds[i + 2], dr = c->ops[i + 1]->d;
            moved[q].resize(c->site_size(i + 1));
            permute_102(W2[q].ptr, moved[q].ptr, bm, br, dr, s);  // back to (br, bm, dr)
            std::swap(c->sites[i], Up[q]);                         // left-normalised site i
            mn.push_back((int64_t)c->site_size(i + 1));
        } else {
            const int bm = c->bonds[i], br = c->bonds[i + 1], dr = c->ops[i]->d;
            permute_102(W2[q].ptr, c->sites[i].ptr, bm, br, dr, s);  // right-normalised site i
            mn.push_back((int64_t)c->site_size(i - 1));
        }
        mid.push_back(moved[q].ptr);
        midc.push_back(moved[q].ptr);
    }
    // states.cpp:119-122 / 155-158: the new middle site is renormalised
    std::vector<double> nn = sumsq_batch(midc, mn, s);
    std::vector<double> f(nn.size());
    for (size_t q = 0; q < nn.size(); ++q) f[q] = 1.0 / std::sqrt(nn[q]);
    scale_batch(mid, mn, f, s);
    for (size_t q = 0; q < cs.size(); ++q) {
        Chain* c = cs[q];
        if (dir > 0) std::swap(c->sites[c->current + 1], moved[q]);
        else std::swap(c->sites[c->current - 1], moved[q]);
        c->current += dir;
        c->optimized = false;
        ++c->stats.moves;
    }
    if (dir > 0) env_left_batch(cs, s);
    else env_right_batch(cs, s);
    NC_HIP(hipStreamSynchronize(s));
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
        const double t0 = now_ms();
        move_all(active, dir);
        t_moves_ms += now_ms() - t0;
        const double t1 = now_ms();
        optimize_sites(active);
        t_solves_ms += now_ms() - t1;
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
