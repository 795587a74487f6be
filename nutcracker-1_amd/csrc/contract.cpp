// Grouped sparse-MPO contractions on the device (complex128 on FP64 matrix
// cores, or complex64 on FP32 ones for the mixed-precision form).
//
// The reference builds a dense stage-1 tensor I(bl,d,cr,bl,d) and runs two
// zgemms per matvec (core.f95:532-621).  Here the block list is contracted as
//   A: P_{f,a} = X_a L_f            one GEMM per (distinct from f, physical a)
//   B: Q_t(.,b) = sum_{n: to_n = t} sum_a W_n(a,b) P_{from_n,a}    (mix kernel)
//   C: y = sum_t R_t Q_t            one K-segmented GEMM, K = br * n_to
// which needs 8 d bl br (n_from bl + n_to br) + 8 nblk bl br d^2 flops instead
// of 8 w d chi^3 (d+1).  The environment updates use the same three steps with
// conj(psi) in place of R (left) or with R first and conj(psi) last (right).
#include "nc_internal.hpp"

#include <algorithm>

namespace nc {

void mpo_mix_tables(int d, int nmat, const uint32_t* idx1, const zd* mats, int cl, int cr, MpoTables& t) {
    if (d <= 0 || cl <= 0 || cr <= 0 || nmat <= 0) throw Error("operator site: bad dimensions", -2);
    t.froms.clear();
    t.tos.clear();
    for (int n = 0; n < nmat; ++n) {
        const int f = (int)idx1[2 * n] - 1, to = (int)idx1[2 * n + 1] - 1;
        if (f < 0 || f >= cl || to < 0 || to >= cr) throw Error("operator site: block index out of range", -3);
        t.froms.push_back(f);
        t.tos.push_back(to);
    }
    std::sort(t.froms.begin(), t.froms.end());
    t.froms.erase(std::unique(t.froms.begin(), t.froms.end()), t.froms.end());
    std::sort(t.tos.begin(), t.tos.end());
    t.tos.erase(std::unique(t.tos.begin(), t.tos.end()), t.tos.end());
    t.blk_from.resize(nmat);
    t.blk_to.resize(nmat);
    for (int n = 0; n < nmat; ++n) {
        const int f = (int)idx1[2 * n] - 1, to = (int)idx1[2 * n + 1] - 1;
        t.blk_from[n] = (int)(std::lower_bound(t.froms.begin(), t.froms.end(), f) - t.froms.begin());
        t.blk_to[n] = (int)(std::lower_bound(t.tos.begin(), t.tos.end(), to) - t.tos.begin());
    }
    for (int dir = 0; dir < 2; ++dir) {
        const bool src_is_from = dir == 0;
        const int ndst = (int)(src_is_from ? t.tos.size() : t.froms.size());
        t.ptr[dir].assign(ndst + 1, 0);
        t.src[dir].clear();
        t.W[dir].clear();
        for (int dst = 0; dst < ndst; ++dst) {
            for (int n = 0; n < nmat; ++n) {
                const int bd = src_is_from ? t.blk_to[n] : t.blk_from[n];
                if (bd != dst) continue;
                t.src[dir].push_back(src_is_from ? t.blk_from[n] : t.blk_to[n]);
                for (int k = 0; k < d * d; ++k) t.W[dir].push_back(mats[(size_t)n * d * d + k]);
            }
            t.ptr[dir][dst + 1] = (int)t.src[dir].size();
        }
    }
}

void MpoSite::build(int d_, int cl_, int cr_, int nmat_, const uint32_t* idx1, const zd* mats_) {
    if (d_ <= 0 || cl_ <= 0 || cr_ <= 0 || nmat_ <= 0) throw Error("operator site: bad dimensions", -2);
    d = d_; cl = cl_; cr = cr_; nmat = nmat_;
    idx.assign(idx1, idx1 + 2 * (size_t)nmat);
    mats.assign(mats_, mats_ + (size_t)d * d * nmat);
    MpoTables t;
    mpo_mix_tables(d, nmat, idx1, mats_, cl, cr, t);
    froms = t.froms;
    tos = t.tos;
    blk_from = t.blk_from;
    blk_to = t.blk_to;
    // mix tables, grouped by destination
    for (int dir = 0; dir < 2; ++dir) {
        std::vector<int> const& ptr = t.ptr[dir];
        std::vector<int> const& src = t.src[dir];
        std::vector<zd> const& w = t.W[dir];
        std::vector<float2> w32(w.size());
        for (size_t i = 0; i < w.size(); ++i) w32[i] = make_float2((float)w[i].real(), (float)w[i].imag());
        mix_ptr[dir].resize(ptr.size());
        mix_src[dir].resize(src.size());
        mix_W[dir].resize(w.size());
        mix_W32[dir].resize(w.size());
        h2d(mix_ptr[dir].ptr, ptr.data(), ptr.size() * sizeof(int));
        h2d(mix_src[dir].ptr, src.data(), src.size() * sizeof(int));
        h2d(mix_W[dir].ptr, w.data(), w.size() * sizeof(zd));
        h2d(mix_W32[dir].ptr, w32.data(), w32.size() * sizeof(float2));
    }
}

namespace {

template <typename CT>
struct Plan;
template <>
struct Plan<double2> {
    typedef GemmPlan type;
    static void run(GemmPlan const& p, hipStream_t s) { zgemm_run(p, s); }
};
template <>
struct Plan<float2> {
    typedef CGemmPlan type;
    static void run(CGemmPlan const& p, hipStream_t s) { cgemm_run(p, s); }
};

// Step A of the matvec / left environment: P_{f,a} = X_a L_f
template <typename CT>
void step_a_left(int bl, int br, MpoSite const& op, const CT* L, const CT* X, CT* P, hipStream_t s) {
    const int64_t rs = (int64_t)br * bl;
    typename Plan<CT>::type plan;
    plan.M = br; plan.N = bl; plan.opA = OpN; plan.opB = OpN; plan.accumulate = false;
    plan.key = TuneKey{'A', bl, br, op.d, op.cl, op.cr, 1};
    for (size_t fi = 0; fi < op.froms.size(); ++fi)
        for (int a = 0; a < op.d; ++a) {
            plan.begin_batch(P + ((int64_t)fi * op.d + a) * rs, br);
            plan.add_seg(X + a * rs, br, L + (int64_t)op.froms[fi] * bl * bl, bl, bl);
        }
    Plan<CT>::run(plan, s);
}

template <typename CT>
void heff_matvec_t(int bl, int br, MpoSite const& op, const CT* L, const CT* R, const CT* x, CT* y, Workspace& ws,
                   hipStream_t s) {
    const int64_t rs = (int64_t)br * bl;
    const int d = op.d;
    ws.ensure(rs * d * op.froms.size(), rs * d * op.tos.size(), (CT*)nullptr);
    CT* P = ws.p((CT*)nullptr);
    CT* Q = ws.q((CT*)nullptr);
    step_a_left<CT>(bl, br, op, L, x, P, s);
    mix_blocks(op, true, P, Q, rs, s);
    typename Plan<CT>::type plan;
    plan.M = br; plan.N = bl * d; plan.opA = OpN; plan.opB = OpN; plan.accumulate = false;
    plan.key = TuneKey{'C', bl, br, d, op.cl, op.cr, 1};
    plan.begin_batch(y, br);
    for (size_t ti = 0; ti < op.tos.size(); ++ti)
        plan.add_seg(R + (int64_t)op.tos[ti] * br * br, br, Q + (int64_t)ti * d * rs, br, br);
    Plan<CT>::run(plan, s);
}

template <typename CT>
void env_left_t(int bl, int br, MpoSite const& op, const CT* L, const CT* psi, CT* newL, Workspace& ws,
                hipStream_t s) {
    const int64_t rs = (int64_t)br * bl;
    const int d = op.d;
    ws.ensure(rs * d * op.froms.size(), rs * d * op.tos.size(), (CT*)nullptr);
    CT* P = ws.p((CT*)nullptr);
    CT* Q = ws.q((CT*)nullptr);
    step_a_left<CT>(bl, br, op, L, psi, P, s);
    mix_blocks(op, true, P, Q, rs, s);
    if ((int)op.tos.size() < op.cr) dzero(newL, (size_t)br * br * op.cr * sizeof(CT));
    typename Plan<CT>::type plan;
    plan.M = br; plan.N = br; plan.opA = OpN; plan.opB = OpC; plan.accumulate = false;
    plan.key = TuneKey{'L', bl, br, d, op.cl, op.cr, 1};
    for (size_t ti = 0; ti < op.tos.size(); ++ti) {
        plan.begin_batch(newL + (int64_t)op.tos[ti] * br * br, br);
        plan.add_seg(Q + (int64_t)ti * d * rs, br, psi, br, bl * d);
    }
    Plan<CT>::run(plan, s);
}

template <typename CT>
void env_right_t(int bl, int br, MpoSite const& op, const CT* R, const CT* psi, CT* newR, Workspace& ws,
                 hipStream_t s) {
    const int64_t rs = (int64_t)br * bl;
    const int d = op.d;
    ws.ensure(rs * d * op.tos.size(), rs * d * op.froms.size(), (CT*)nullptr);
    CT* P = ws.p((CT*)nullptr);
    CT* Q = ws.q((CT*)nullptr);
    {   // P'_{t,a} = R_t Psi_a
        typename Plan<CT>::type plan;
        plan.M = br; plan.N = bl; plan.opA = OpN; plan.opB = OpN; plan.accumulate = false;
        plan.key = TuneKey{'r', bl, br, d, op.cl, op.cr, 1};
        for (size_t ti = 0; ti < op.tos.size(); ++ti)
            for (int a = 0; a < d; ++a) {
                plan.begin_batch(P + ((int64_t)ti * d + a) * rs, br);
                plan.add_seg(R + (int64_t)op.tos[ti] * br * br, br, psi + a * rs, br, br);
            }
        Plan<CT>::run(plan, s);
    }
    mix_blocks(op, false, P, Q, rs, s);
    if ((int)op.froms.size() < op.cl) dzero(newR, (size_t)bl * bl * op.cl * sizeof(CT));
    typename Plan<CT>::type plan;  // N_f = sum_b Psi_b^H Q'_{f,b}
    plan.M = bl; plan.N = bl; plan.opA = OpC; plan.opB = OpN; plan.accumulate = false;
    plan.key = TuneKey{'R', bl, br, d, op.cl, op.cr, 1};
    for (size_t fi = 0; fi < op.froms.size(); ++fi) {
        plan.begin_batch(newR + (int64_t)op.froms[fi] * bl * bl, bl);
        for (int b = 0; b < d; ++b)
            plan.add_seg(psi + b * rs, br, Q + ((int64_t)fi * d + b) * rs, br, br);
    }
    Plan<CT>::run(plan, s);
}

}  // namespace

void heff_matvec(int bl, int br, MpoSite const& op, const double2* L, const double2* R, const double2* x, double2* y,
                 Workspace& ws, hipStream_t s) {
    heff_matvec_t<double2>(bl, br, op, L, R, x, y, ws, s);
}
void heff_matvec(int bl, int br, MpoSite const& op, const float2* L, const float2* R, const float2* x, float2* y,
                 Workspace& ws, hipStream_t s) {
    heff_matvec_t<float2>(bl, br, op, L, R, x, y, ws, s);
}
void env_left(int bl, int br, MpoSite const& op, const double2* L, const double2* psi, double2* newL, Workspace& ws,
              hipStream_t s) {
    env_left_t<double2>(bl, br, op, L, psi, newL, ws, s);
}
void env_left(int bl, int br, MpoSite const& op, const float2* L, const float2* psi, float2* newL, Workspace& ws,
              hipStream_t s) {
    env_left_t<float2>(bl, br, op, L, psi, newL, ws, s);
}
void env_right(int bl, int br, MpoSite const& op, const double2* R, const double2* psi, double2* newR, Workspace& ws,
               hipStream_t s) {
    env_right_t<double2>(bl, br, op, R, psi, newR, ws, s);
}
void env_right(int bl, int br, MpoSite const& op, const float2* R, const float2* psi, float2* newR, Workspace& ws,
               hipStream_t s) {
    env_right_t<float2>(bl, br, op, R, psi, newR, ws, s);
}

zd expectation(int bl, int br, MpoSite const& op, const double2* L, const double2* psi, const double2* R,
               Workspace& ws, hipStream_t s) {
    ZBuf N((size_t)bl * bl * op.cl);
    env_right(bl, br, op, R, psi, N.ptr, ws, s);
    return trace_pair(N.ptr, L, bl, op.cl, s);
}

}  // namespace nc
