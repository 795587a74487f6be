// H_eff in complex64 on the bf16 matrix cores with operands split once.
//
// The complex64 matvec y = H_eff x (steps A, mix, C of the grouped
// contraction) spends its time in products whose operands are split into
// three bf16 parts each (hi + mid + lo, six partial products per real
// product, Gauss's 3M form for the complex one).  Splitting inside the GEMM
// repeats the work for every output tile that reads an operand and costs more
// vector instructions than the matrix-core instructions it feeds.  Here the
// operands are split once, outside the GEMM:
//   * the environments L_f and R_t when the operator is built (once per site
//     solve: they are fixed across its hundreds of matvecs),
//   * x once per matvec (split_planes),
//   * step A's output by the mix kernel, which writes Q_t's planes directly
//     (mix_split_kernel),
// and the GEMM (cgemm_planes_kernel) only moves planes into LDS and runs MFMAs.
//
// Plane layout.  A complex matrix Z with rows r < R and inner (K) index k < K
// is held as nine bf16 planes, p = 3 c + j for the component c (re, im,
// re + im) and the part j (hi, mid, lo).  Each plane is an array of 16-byte
// chunks, chunk (kb, r) = the 8 bf16 of k = 8 kb .. 8 kb + 7 in row r, at
// chunk index kb * Rp + r; Rp = R rounded up to 128, Kp = K rounded up to 32,
// Kb = Kp / 8 chunk columns, plane stride Kb * Rp chunks.  Chunks of
// k >= K are zero; rows R .. Rp - 1 are never written and only reach output
// rows or columns that are not stored.  One MFMA operand fragment (lane l: row
// l % 16, k = 8 (l / 16) .. + 7) is then a single 16-byte chunk, 16 lanes read
// 256 contiguous bytes, and a tile's stage is a set of contiguous 1 KB runs.
//
// GEMM: 128 x 128 output tile per 512-thread workgroup (8 waves in 2 x 4, each
// wave 4 x 2 MFMA blocks of 16 x 16 for each of the three real products), K
// in stages of 32, one component (3 parts of A and of B, 48 KB) per pipeline
// step.  The three components cycle through three LDS buffers, so buffer q
// always holds component q; each step's planes are copied global -> LDS by
// global_load_lds_dwordx4 two steps ahead of their use, and a step waits only
// for its own copies (vmcnt) before the workgroup barrier.  K can be split
// over several workgroups per tile (slices summed by the consumer).
#include "nc_bf16.hpp"
#include "nc_internal.hpp"
#include "nc_lds_dma.hpp"

#include <algorithm>
#include <cmath>

namespace nc {

void PlaneSet::shape(int R_, int K_, int count_) {
    if (R_ <= 0 || K_ <= 0 || count_ <= 0) throw Error("plane set: bad shape", -2);
    R = R_;
    K = K_;
    count = count_;
    Rp = (R + kPlaneTile - 1) / kPlaneTile * kPlaneTile;
    Kb = (K + 31) / 32 * 4;
    ps = (int64_t)Kb * Rp;
    buf.resize((size_t)(9 * ps * count));
}

namespace {

constexpr int kPT = kPlaneTile;

typedef const __attribute__((address_space(1))) void* GlobalPtr;

// nine planes of the 8 values re[], im[] at chunk d (plane stride ps)
__device__ __forceinline__ void store_planes(u32x4* __restrict__ d, int64_t ps, const float (&re)[8],
                                             const float (&im)[8]) {
    float sm[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) sm[e] = re[e] + im[e];
    u32x4 h, m, l;
    split3(re, h, m, l);
    d[0] = h; d[ps] = m; d[2 * ps] = l;
    split3(im, h, m, l);
    d[3 * ps] = h; d[4 * ps] = m; d[5 * ps] = l;
    split3(sm, h, m, l);
    d[6 * ps] = h; d[7 * ps] = m; d[8 * ps] = l;
}

// Planes of matrix z: element (r, k) at src[z * sz + r * sr + k * sk].  With
// KCONTIG (sk = 1) a wave covers 16 rows x 4 chunks (4 x 128 contiguous source
// bytes per row, 256 contiguous destination bytes per chunk column), otherwise
// consecutive lanes take consecutive rows.
template <bool KCONTIG>
__global__ __launch_bounds__(256) void split_planes_kernel(const double2* __restrict__ src, int64_t sr, int64_t sk,
                                                           int64_t sz, u32x4* __restrict__ dst, int64_t ps,
                                                           int64_t dz, int R, int K, int Rp, int Kb) {
    const int z = blockIdx.y;
    int r, kb;
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (KCONTIG) {
        const int64_t w = gid >> 6;
        const int lane = threadIdx.x & 63;
        const int nrg = (R + 15) >> 4;
        r = (int)(w % nrg) * 16 + (lane >> 2);
        kb = (int)(w / nrg) * 4 + (lane & 3);
    } else {
        r = (int)(gid % R);
        kb = (int)(gid / R);
    }
    if (r >= R || kb >= Kb) return;
    const double2* s = src + z * sz + (int64_t)r * sr;
    float re[8], im[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const int k = 8 * kb + e;
        re[e] = 0.f;
        im[e] = 0.f;
        if (k < K) {
            const double2 v = s[(int64_t)k * sk];
            re[e] = (float)v.x;
            im[e] = (float)v.y;
        }
    }
    store_planes(dst + z * dz + (int64_t)kb * Rp + r, ps, re, im);
}

// three bf16 parts of the two values (x0, x1): one packed pair per part
__device__ __forceinline__ void split_pair(float x0, float x1, unsigned& h, unsigned& m, unsigned& l) {
    h = bf16_pair(x0, x1);
    x0 -= bf16_lo(h);
    x1 -= bf16_hi(h);
    m = bf16_pair(x0, x1);
    x0 -= bf16_lo(m);
    x1 -= bf16_hi(m);
    l = bf16_pair(x0, x1);
}

// Step C's B operand planes from step A's output.  For destination t (an MPO
// 'to' index), column n = l + bl b and inner index k < br,
//   Q_t(k, n) = sum_{blocks j: to = t} sum_a W_j(a, b) sum_c P_c[(src_j d + a) rs + k + br l],
// with the nslice K-split slices of P summed on the fly.  A wave takes one
// chunk column kb and 16 consecutive l; lane (l, i) holds k = 8 kb + 2 i and
// k + 1, so each source block is one 16-byte load per lane (64 contiguous
// bytes per l) and each plane one 4-byte store per lane (256 contiguous bytes
// per store).  EVEN (br even): the pair is 16-byte aligned and inside the
// matrix whenever its first value is.  D > 0 fixes the physical dimension;
// D = 0 takes any d, four outputs b per pass over the blocks.
template <int D, bool EVEN>
__global__ __launch_bounds__(256) void mix_split_kernel(const float2* __restrict__ P, int64_t pslice, int nslice,
                                                        const int* __restrict__ dst_ptr,
                                                        const int* __restrict__ blk_src,
                                                        const float2* __restrict__ W, int d_rt, int br, int bl,
                                                        int64_t rs, unsigned* __restrict__ Q, int64_t ps, int64_t qz,
                                                        int Np, int Kb) {
    constexpr int NB = D > 0 ? D : 4;
    const int d = D > 0 ? D : d_rt;
    const int t = blockIdx.y;
    const int nlg = (bl + 15) >> 4;
    const int64_t w = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    const int kb = (int)(w / nlg);
    const int l = (int)(w % nlg) * 16 + (lane >> 2);
    const int pi = lane & 3;
    if (kb >= Kb || l >= bl) return;
    const int k = 8 * kb + 2 * pi;
    const bool in0 = k < br, in1 = k + 1 < br;
    const int64_t e = (int64_t)br * l + k;
    const int jb = dst_ptr[t], je = dst_ptr[t + 1];
    for (int b0 = 0; b0 < d; b0 += NB) {
        float r0[NB], i0[NB], r1[NB], i1[NB];
#pragma unroll
        for (int b = 0; b < NB; ++b) r0[b] = i0[b] = r1[b] = i1[b] = 0.f;
        for (int j = jb; j < je; ++j) {
            const float2* Pj = P + (int64_t)blk_src[j] * d * rs + e;
            const float2* Wj = W + (int64_t)j * d * d;
#pragma unroll(D > 0 ? D : 1)
            for (int a = 0; a < d; ++a) {
                const float2* pp = Pj + (int64_t)a * rs;
                float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
                if (EVEN) {
                    if (in0) {
                        v = *reinterpret_cast<const float4*>(pp);
                        for (int c = 1; c < nslice; ++c) {
                            const float4 u = *reinterpret_cast<const float4*>(pp + c * pslice);
                            v.x += u.x; v.y += u.y; v.z += u.z; v.w += u.w;
                        }
                    }
                } else {
                    for (int c = 0; c < nslice; ++c) {
                        if (in0) { const float2 u = pp[c * pslice]; v.x += u.x; v.y += u.y; }
                        if (in1) { const float2 u = pp[c * pslice + 1]; v.z += u.x; v.w += u.y; }
                    }
                }
#pragma unroll
                for (int b = 0; b < NB; ++b) {
                    if (b0 + b < d) {
                        const float2 wv = Wj[a + (b0 + b) * d];
                        r0[b] += wv.x * v.x - wv.y * v.y;
                        i0[b] += wv.x * v.y + wv.y * v.x;
                        r1[b] += wv.x * v.z - wv.y * v.w;
                        i1[b] += wv.x * v.w + wv.y * v.z;
                    }
                }
            }
        }
        // nine planes, dword pi of chunk (kb, n); dword stride of a plane 4 ps
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            if (b0 + b < d) {
                unsigned* q = Q + 4 * (t * qz + (int64_t)kb * Np + l + (int64_t)bl * (b0 + b)) + pi;
                const int64_t s4 = 4 * ps;
                unsigned h, m, lo;
                split_pair(r0[b], r1[b], h, m, lo);
                q[0] = h; q[s4] = m; q[2 * s4] = lo;
                split_pair(i0[b], i1[b], h, m, lo);
                q[3 * s4] = h; q[4 * s4] = m; q[5 * s4] = lo;
                split_pair(r0[b] + i0[b], r1[b] + i1[b], h, m, lo);
                q[6 * s4] = h; q[7 * s4] = m; q[8 * s4] = lo;
            }
        }
    }
}

// The same mix for every destination at once: a thread loads each source
// block's pair (k, k + 1) of row l once (every K-split slice summed), and forms
// all Q_t(k, l + bl b) from registers, instead of one grid row per destination
// t reading the sources of its own blocks (the source slabs of a block list
// with several blocks per source are then read once, not once per block).
// Up to MAXF distinct sources and MAXT destinations; br even (16-byte pairs).
template <int D, int MAXF, int MAXT>
__global__ __launch_bounds__(256) void mix_split_all_kernel(const float2* __restrict__ P, int64_t pslice, int nslice,
                                                            int nsrc, int ndst, const int* __restrict__ dst_ptr,
                                                            const int* __restrict__ blk_src,
                                                            const float2* __restrict__ W, int br, int bl, int64_t rs,
                                                            unsigned* __restrict__ Q, int64_t ps, int64_t qz, int Np,
                                                            int Kb) {
    const int nlg = (bl + 15) >> 4;
    const int64_t w = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    const int kb = (int)(w / nlg);
    const int l = (int)(w % nlg) * 16 + (lane >> 2);
    const int pi = lane & 3;
    if (kb >= Kb || l >= bl) return;
    const int k = 8 * kb + 2 * pi;
    const bool in0 = k < br;
    const int64_t e = (int64_t)br * l + k;
    // every source slab's pair, slices summed
    float4 v[MAXF][D];
#pragma unroll
    for (int f = 0; f < MAXF; ++f)
#pragma unroll
        for (int a = 0; a < D; ++a) {
            float4 x = make_float4(0.f, 0.f, 0.f, 0.f);
            if (f < nsrc && in0) {
                const float2* pp = P + ((int64_t)f * D + a) * rs + e;
                x = *reinterpret_cast<const float4*>(pp);
                for (int c = 1; c < nslice; ++c) {
                    const float4 u = *reinterpret_cast<const float4*>(pp + c * pslice);
                    x.x += u.x; x.y += u.y; x.z += u.z; x.w += u.w;
                }
            }
            v[f][a] = x;
        }
    const int64_t s4 = 4 * ps;
#pragma unroll
    for (int t = 0; t < MAXT; ++t) {
        if (t >= ndst) break;
        float r0[D], i0[D], r1[D], i1[D];
#pragma unroll
        for (int b = 0; b < D; ++b) r0[b] = i0[b] = r1[b] = i1[b] = 0.f;
        const int jb = dst_ptr[t], je = dst_ptr[t + 1];
        for (int j = jb; j < je; ++j) {
            const int f = blk_src[j];
            const float2* Wj = W + (int64_t)j * D * D;
#pragma unroll
            for (int a = 0; a < D; ++a) {
                // the source's pair, picked from registers
                float4 x = v[0][a];
#pragma unroll
                for (int q = 1; q < MAXF; ++q)
                    if (f == q) x = v[q][a];
#pragma unroll
                for (int b = 0; b < D; ++b) {
                    const float2 wv = Wj[a + b * D];
                    r0[b] += wv.x * x.x - wv.y * x.y;
                    i0[b] += wv.x * x.y + wv.y * x.x;
                    r1[b] += wv.x * x.z - wv.y * x.w;
                    i1[b] += wv.x * x.w + wv.y * x.z;
                }
            }
        }
#pragma unroll
        for (int b = 0; b < D; ++b) {
            unsigned* q = Q + 4 * (t * qz + (int64_t)kb * Np + l + (int64_t)bl * b) + pi;
            unsigned h, m, lo;
            split_pair(r0[b], r1[b], h, m, lo);
            q[0] = h; q[s4] = m; q[2 * s4] = lo;
            split_pair(i0[b], i1[b], h, m, lo);
            q[3 * s4] = h; q[4 * s4] = m; q[5 * s4] = lo;
            split_pair(r0[b] + i0[b], r1[b] + i1[b], h, m, lo);
            q[6 * s4] = h; q[7 * s4] = m; q[8 * s4] = lo;
        }
    }
}

// out[e] = sum_c Y[c slice + e] (in FP64), e < n
__global__ __launch_bounds__(256) void reduce_slices_kernel(const float2* __restrict__ Y, int64_t slice, int ns,
                                                            int64_t n, double2* __restrict__ out) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        double sx = 0, sy = 0;
        for (int c = 0; c < ns; ++c) {
            const float2 v = Y[c * slice + e];
            sx += v.x;
            sy += v.y;
        }
        out[e] = make_double2(sx, sy);
    }
}

// ---- the GEMM on planes
struct PStage {
    u32x4 a[3][4][kPT];  // [part][chunk column][row] of the A tile's current component
    u32x4 b[3][4][kPT];
};

// 16 bytes per lane from global memory into LDS (lane i's chunk at g + voff_i
// lands at LDS address l + 16 i); awaited only by the vmcnt of step_barrier.
__device__ __forceinline__ void dma16(const u32x4* g, unsigned voff, const u32x4* l) { lds_dma16(g, voff, l); }

// One pipeline step's copies: component Q of the K stage whose chunk column 0
// is at pa (A, offset to the tile's first row) and pb (B).  The 24 runs of 64
// chunks of each operand (t = 8 part + 2 chunk column + half) are shared out
// 24 / NW per wave; off.a[Q][u] is the byte offset of this lane's chunk of run
// t = (24 / NW) wave + u from pa (likewise for B), fixed for the whole launch.
template <int NW>
struct IssueOffsets {
    static constexpr int kRuns = 24 / NW;
    unsigned a[3][kRuns], b[3][kRuns];
};
template <int Q, int NW>
__device__ __forceinline__ void planes_issue(PStage& S, const u32x4* pa, const u32x4* pb, IssueOffsets<NW> const& off,
                                             int wave) {
    constexpr int R = IssueOffsets<NW>::kRuns;
#pragma unroll
    for (int u = 0; u < R; ++u) {
        const int t = wave * R + u;
        const int p = t >> 3, kb = (t >> 1) & 3, h = (t & 1) * 64;
        dma16(pa, off.a[Q][u], &S.a[p][kb][h]);
        dma16(pb, off.b[Q][u], &S.b[p][kb][h]);
    }
}

// The six partial products of one component for the wave's 4 x JB blocks,
// smallest first (lo.hi, mid.mid, hi.lo, mid.hi, hi.mid, hi.hi).  The
// scheduling barrier keeps each step's MFMAs inside it, so the accumulators of
// the three components keep their registers across the loop.
template <int JB>
__device__ __forceinline__ void planes_compute(PStage const& S, f32x4 (&acc)[4][JB], int wm, int wn, int lane) {
    __builtin_amdgcn_sched_barrier(0);
    const int r = lane & 15, kb = lane >> 4;
    bf16x8 b[JB][3];
#pragma unroll
    for (int j = 0; j < JB; ++j)
#pragma unroll
        for (int p = 0; p < 3; ++p) b[j][p] = __builtin_bit_cast(bf16x8, S.b[p][kb][wn + 16 * j + r]);
    constexpr int pa[6] = {2, 1, 0, 1, 0, 0};
    constexpr int pb[6] = {0, 1, 2, 0, 1, 0};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        bf16x8 a[3];
#pragma unroll
        for (int p = 0; p < 3; ++p) a[p] = __builtin_bit_cast(bf16x8, S.a[p][kb][wm + 16 * i + r]);
#pragma unroll
        for (int t = 0; t < 6; ++t)
#pragma unroll
            for (int j = 0; j < JB; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[pa[t]], b[j][pb[t]], acc[i][j], 0, 0, 0);
    }
}

// Workgroup barrier after this wave's copies up to the last N issued have
// landed in LDS and its LDS reads are done.  __syncthreads() would also wait
// for the copies still in flight for later steps (its fence covers every LDS
// write, DMA included); the memory clobber keeps the compiler from moving LDS
// accesses across the barrier.
template <int N>
__device__ __forceinline__ void step_barrier() {
    static_assert(N == 0 || N == 6 || N == 12, "vmcnt of a pipeline step");
    if (N == 0) asm volatile("s_waitcnt vmcnt(0)\n\ts_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
    else if (N == 6) asm volatile("s_waitcnt vmcnt(6)\n\ts_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
    else asm volatile("s_waitcnt vmcnt(12)\n\ts_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// NW waves per workgroup on the 128 x 128 tile: NW = 4, each wave 64 x 64 (one
// wave per SIMD); NW = 8, each wave 64 x 32 (two waves per SIMD: while one
// waits at a barrier or for its LDS reads the other keeps the matrix pipe
// busy; 96 accumulator registers per wave instead of 192).
template <int NW>
__global__ __launch_bounds__(64 * NW) void cgemm_planes_kernel(const PlaneBatch* __restrict__ batches, int nsplit) {
    constexpr int JB = 16 / NW;                         // 16 x 16 blocks per wave along N
    constexpr int kIssuesPerStep = 2 * IssueOffsets<NW>::kRuns;  // copies per wave per step
    __shared__ PStage S0, S1, S2;
    const PlaneBatch bt = batches[blockIdx.y];
    const int tiles_m = (bt.M + kPT - 1) / kPT, tiles_n = (bt.N + kPT - 1) / kPT;
    const int nitems = tiles_m * tiles_n * nsplit;
    int w = blockIdx.x;
    if (w >= nitems) return;
    // each XCD (workgroups w = x mod 8) takes a contiguous range of items
    if ((nitems & 7) == 0 && (int)gridDim.x == nitems) w = (w & 7) * (nitems >> 3) + (w >> 3);
    const int tile = w / nsplit, split = w % nsplit;
    const int m0 = (tile / tiles_n) * kPT, n0 = (tile % tiles_n) * kPT;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wm = (wave & 1) * 64, wn = (wave >> 1) * (16 * JB);
    IssueOffsets<NW> off;
#pragma unroll
    for (int q = 0; q < 3; ++q)
#pragma unroll
        for (int u = 0; u < IssueOffsets<NW>::kRuns; ++u) {
            const int t = wave * IssueOffsets<NW>::kRuns + u;
            const int p = t >> 3, kb = (t >> 1) & 3, h = (t & 1) * 64;
            off.a[q][u] = 16u * (unsigned)((3 * q + p) * bt.a_ps + (int64_t)kb * bt.a_rows + h + lane);
            off.b[q][u] = 16u * (unsigned)((3 * q + p) * bt.b_ps + (int64_t)kb * bt.b_rows + h + lane);
        }

    f32x4 acc[3][4][JB];
#pragma unroll
    for (int q = 0; q < 3; ++q)
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < JB; ++j) acc[q][i][j] = f32x4{0, 0, 0, 0};

    const int nks = bt.nseg * bt.nks;
    const int ks0 = (int)((int64_t)nks * split / nsplit), ks1 = (int)((int64_t)nks * (split + 1) / nsplit);
    if (ks0 < ks1) {
        // chunk column 0 of K stage ks (segment ks / nks, stage ks % nks in it)
        // and of the next stage, each advanced by a select
        const u32x4* A = reinterpret_cast<const u32x4*>(bt.A) + m0;
        const u32x4* B = reinterpret_cast<const u32x4*>(bt.B) + n0;
        const int64_t a_st = 4 * (int64_t)bt.a_rows, b_st = 4 * (int64_t)bt.b_rows;
        int seg = ks0 / bt.nks, kl = ks0 - seg * bt.nks;
        const u32x4* pa = A + seg * bt.a_seg + kl * a_st;
        const u32x4* pb = B + seg * bt.b_seg + kl * b_st;
        auto advance = [&](const u32x4*& qa, const u32x4*& qb) {
            const bool wrap = ++kl == bt.nks;
            qa += wrap ? bt.a_seg - (bt.nks - 1) * a_st : a_st;
            qb += wrap ? bt.b_seg - (bt.nks - 1) * b_st : b_st;
            kl = wrap ? 0 : kl;
        };
        const u32x4* na = pa;
        const u32x4* nb = pb;
        if (ks0 + 1 < ks1) advance(na, nb);
        planes_issue<0, NW>(S0, pa, pb, off, wave);
        planes_issue<1, NW>(S1, pa, pb, off, wave);
        // One body for every K stage: the last one copies its own stage again
        // into the buffers of the following steps (never read), so the loop has
        // no tail and all copies are awaited before the epilogue.
        for (int ks = ks0; ks < ks1; ++ks) {
            // component re: buffer 0 (its copies were issued two steps ago)
            step_barrier<kIssuesPerStep>();
            planes_issue<2, NW>(S2, pa, pb, off, wave);
            planes_compute<JB>(S0, acc[0], wm, wn, lane);
            // component im: buffer 1
            step_barrier<kIssuesPerStep>();
            planes_issue<0, NW>(S0, na, nb, off, wave);
            planes_compute<JB>(S1, acc[1], wm, wn, lane);
            // component re + im: buffer 2
            step_barrier<kIssuesPerStep>();
            planes_issue<1, NW>(S1, na, nb, off, wave);
            planes_compute<JB>(S2, acc[2], wm, wn, lane);
            pa = na;
            pb = nb;
            if (ks + 2 < ks1) advance(na, nb);
        }
        step_barrier<0>();  // no copy may still write LDS when the workgroup ends
    }

    float2* __restrict__ C = bt.C + (int64_t)split * bt.slice;
    const int lr = lane & 15, lk = lane >> 4;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < JB; ++j) {
            const int col = n0 + wn + 16 * j + lr;
            if (col >= bt.N) continue;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = m0 + wm + 16 * i + 4 * lk + v;
                if (row < bt.M) {
                    const float t1 = acc[0][i][j][v], t2 = acc[1][i][j][v], t3 = acc[2][i][j][v];
                    C[row + (int64_t)col * bt.ldc] = make_float2(t1 - t2, t3 - t1 - t2);
                }
            }
        }
}

int cu_count() {
    static int n = [] {
        int dev = 0, v = 0;
        NC_HIP(hipGetDevice(&dev));
        NC_HIP(hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev));
        return v > 0 ? v : 1;
    }();
    return n;
}

}  // namespace

void split_planes(const double2* src, int64_t sr, int64_t sk, int64_t sz, int nz, PlaneSet& dst, int z0,
                  hipStream_t s) {
    if (nz <= 0) return;
    if (z0 < 0 || z0 + nz > dst.count) throw Error("split_planes: matrix index out of range", -2);
    u32x4* d = reinterpret_cast<u32x4*>(dst.at(z0));
    const int64_t dz = 9 * dst.ps;
    if (sk == 1) {
        const int64_t waves = (int64_t)((dst.R + 15) / 16) * ((dst.Kb + 3) / 4);
        dim3 grid((unsigned)((waves * 64 + 255) / 256), (unsigned)nz);
        hipLaunchKernelGGL(split_planes_kernel<true>, grid, dim3(256), 0, s, src, sr, sk, sz, d, dst.ps, dz, dst.R,
                           dst.K, dst.Rp, dst.Kb);
    } else {
        const int64_t n = (int64_t)dst.R * dst.Kb;
        dim3 grid((unsigned)((n + 255) / 256), (unsigned)nz);
        hipLaunchKernelGGL(split_planes_kernel<false>, grid, dim3(256), 0, s, src, sr, sk, sz, d, dst.ps, dz, dst.R,
                           dst.K, dst.Rp, dst.Kb);
    }
    NC_HIP(hipGetLastError());
}

void mix_split_planes(MpoSite const& op, const float2* P, int64_t pslice, int nslice, int br, int bl, PlaneSet& Q,
                      hipStream_t s) {
    const int ndst = (int)op.tos.size();
    if (Q.count != ndst || Q.R != bl * op.d || Q.K != br) throw Error("mix_split_planes: plane shape", -2);
    if (nslice < 1) throw Error("mix_split_planes: no slices", -2);
    const int64_t rs = (int64_t)br * bl;
    const int64_t waves = (int64_t)((bl + 15) / 16) * Q.Kb;
    dim3 grid((unsigned)((waves + 3) / 4), (unsigned)ndst);
    unsigned* q = reinterpret_cast<unsigned*>(Q.buf.ptr);
    const int* ptr = op.mix_ptr[0].ptr;
    const int* src = op.mix_src[0].ptr;
    const float2* W = op.mix_W32[0].ptr;
    const int64_t qz = 9 * Q.ps;
    // 16-byte pair loads need an even br (then every pair offset is even) and
    // an even slice stride
    const bool even = br % 2 == 0 && pslice % 2 == 0;
    const int nsrc = (int)op.froms.size();
    // every destination from one read of each source, where that still gives
    // the chip enough waves (measured at chi = 1024: 0.072 against 0.089 ms;
    // at chi = 256 a launch of 512 waves is slower than the per-destination one)
    if (even && (op.d == 2 || op.d == 3) && nsrc <= 8 && ndst <= 8 && waves >= 4096) {
        dim3 g1((unsigned)((waves + 3) / 4));
        if (op.d == 2)
            hipLaunchKernelGGL((mix_split_all_kernel<2, 8, 8>), g1, dim3(256), 0, s, P, pslice, nslice, nsrc, ndst, ptr,
                               src, W, br, bl, rs, q, Q.ps, qz, Q.Rp, Q.Kb);
        else
            hipLaunchKernelGGL((mix_split_all_kernel<3, 8, 8>), g1, dim3(256), 0, s, P, pslice, nslice, nsrc, ndst, ptr,
                               src, W, br, bl, rs, q, Q.ps, qz, Q.Rp, Q.Kb);
        NC_HIP(hipGetLastError());
        return;
    }
#define NC_MIX_SPLIT(DD, EV)                                                                                      \
    hipLaunchKernelGGL((mix_split_kernel<DD, EV>), grid, dim3(256), 0, s, P, pslice, nslice, ptr, src, W, op.d, br, \
                       bl, rs, q, Q.ps, qz, Q.Rp, Q.Kb)
    switch (op.d) {
        case 1: if (even) NC_MIX_SPLIT(1, true); else NC_MIX_SPLIT(1, false); break;
        case 2: if (even) NC_MIX_SPLIT(2, true); else NC_MIX_SPLIT(2, false); break;
        case 3: if (even) NC_MIX_SPLIT(3, true); else NC_MIX_SPLIT(3, false); break;
        case 4: if (even) NC_MIX_SPLIT(4, true); else NC_MIX_SPLIT(4, false); break;
        default: if (even) NC_MIX_SPLIT(0, true); else NC_MIX_SPLIT(0, false); break;
    }
#undef NC_MIX_SPLIT
    NC_HIP(hipGetLastError());
}

void reduce_slices(const float2* Y, int64_t slice, int ns, int64_t n, double2* out, hipStream_t s) {
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(reduce_slices_kernel, dim3((unsigned)std::max<int64_t>(blocks, 1)), dim3(256), 0, s, Y, slice,
                       ns, n, out);
    NC_HIP(hipGetLastError());
}

int plane_split_count(int64_t tiles, int nks) {
    const int cus = cu_count();
    int best = 1;
    double bc = 1e300;
    for (int ns = 1; ns <= 8 && ns <= nks; ++ns) {
        // rounds of one workgroup per CU, each 1/ns of a tile's work, plus a
        // small charge per extra slice (its output traffic)
        const double rounds = std::ceil((double)(tiles * ns) / cus);
        const double c = rounds / ns * (1.0 + 0.03 * (ns - 1));
        if (c < bc - 1e-12) { bc = c; best = ns; }
    }
    return best;
}

void PlaneGemm::prepare(std::vector<PlaneBatch> const& bs, int nsplit_) {
    if (bs.empty()) throw Error("plane GEMM: no batches", -2);
    nsplit = std::max(1, nsplit_);
    nbatch = (int)bs.size();
    items = 0;
    for (auto const& b : bs) {
        if (b.M <= 0 || b.N <= 0 || b.nseg <= 0 || b.nks <= 0) throw Error("plane GEMM: bad batch", -2);
        if (b.a_rows < (b.M + kPT - 1) / kPT * kPT || b.b_rows < (b.N + kPT - 1) / kPT * kPT || b.a_rows % kPT ||
            b.b_rows % kPT || b.a_ps < 4 * (int64_t)b.nks * b.a_rows || b.b_ps < 4 * (int64_t)b.nks * b.b_rows)
            throw Error("plane GEMM: planes do not cover the batch's tiles", -2);
        if (nsplit > b.nseg * b.nks) throw Error("plane GEMM: more K slices than K stages", -2);
        // a copy's byte offset within one matrix's nine planes is 32-bit
        if (144 * b.a_ps >= (int64_t)1 << 32 || 144 * b.b_ps >= (int64_t)1 << 32)
            throw Error("plane GEMM: plane set above 4 GiB per matrix", -2);
        const int64_t it = (int64_t)((b.M + kPT - 1) / kPT) * ((b.N + kPT - 1) / kPT) * nsplit;
        items = std::max(items, it);
    }
    batches.resize(bs.size());
    h2d(batches.ptr, bs.data(), bs.size() * sizeof(PlaneBatch));
}

// eight waves per workgroup: measured 0.80 against 0.84 ms per chi = 1024 plane
// matvec with four (the two waves of a SIMD overlap each other's LDS waits)
constexpr int kPlaneWaves = 8;

void PlaneGemm::launch(hipStream_t s) const {
    hipLaunchKernelGGL(cgemm_planes_kernel<kPlaneWaves>, dim3((unsigned)items, (unsigned)nbatch),
                       dim3(64 * kPlaneWaves), 0, s, batches.ptr, nsplit);
    NC_HIP(hipGetLastError());
}

// ----------------------------------------------------------------------------
int& mixed_planes_mode() {
    static int mode = 1;
    return mode;
}

// Measured at the XXZ bulk site (d = 2): the plane matvec is slower than the
// complex64 GEMM path from chi = 192 to 384 (its 128 x 128 tiles leave CUs
// idle) and faster from 512 on (0.158 against 0.188 ms at 512, 0.82 against
// 1.01 ms at 1024).
constexpr int64_t kPlaneMinSites = 512 * 512;

bool use_planes(int bl, int br) {
    const int m = mixed_planes_mode();
    return m == 2 || (m == 1 && (int64_t)bl * br >= kPlaneMinSites);
}

PlaneHeff::PlaneHeff(int bl_, int br_, MpoSite const& op_, const double2* L, const double2* R, hipStream_t s)
    : bl(bl_), br(br_), op(op_) {
    const int d = op.d;
    const int nf = (int)op.froms.size(), nt = (int)op.tos.size();
    if (bl <= 0 || br <= 0 || d <= 0 || nf == 0 || nt == 0) throw Error("plane H_eff: bad shape", -2);
    const int64_t rs = (int64_t)br * bl;
    Lp.shape(bl, bl, nf);
    Rp.shape(br, br, nt);
    Xp.shape(br, bl, d);
    Qp.shape(bl * d, br, nt);
    for (int fi = 0; fi < nf; ++fi)  // L_f(k, n) = L[f bl^2 + k + n bl]: rows n, K index k
        split_planes(L + (int64_t)op.froms[fi] * bl * bl, bl, 1, 0, 1, Lp, fi, s);
    for (int ti = 0; ti < nt; ++ti)  // R_t(m, k) = R[t br^2 + m + k br]
        split_planes(R + (int64_t)op.tos[ti] * br * br, 1, br, 0, 1, Rp, ti, s);

    // step A: P_{f,a} (br x bl) = X_a L_f, one batch per (f, a)
    const int64_t tiles_a = (int64_t)((br + kPT - 1) / kPT) * ((bl + kPT - 1) / kPT) * nf * d;
    nsa = plane_split_count(tiles_a, Xp.Kb / 4);
    pslice = rs * d * nf;
    P.resize((size_t)(pslice * nsa));
    std::vector<PlaneBatch> bs;
    for (int fi = 0; fi < nf; ++fi)
        for (int a = 0; a < d; ++a) {
            PlaneBatch b{};
            b.C = P.ptr + ((int64_t)fi * d + a) * rs;
            b.ldc = br;
            b.slice = pslice;
            b.A = Xp.at(a);
            b.B = Lp.at(fi);
            b.a_ps = Xp.ps;
            b.b_ps = Lp.ps;
            b.a_rows = Xp.Rp;
            b.b_rows = Lp.Rp;
            b.M = br;
            b.N = bl;
            b.nseg = 1;
            b.nks = Xp.Kb / 4;
            bs.push_back(b);
        }
    gemm_a.prepare(bs, nsa);

    // step C: y (br x bl d) = sum_t R_t Q_t, K = br per segment (one segment
    // per 't' index: consecutive matrices of the R and Q plane sets)
    const int64_t tiles_c = (int64_t)((br + kPT - 1) / kPT) * ((bl * d + kPT - 1) / kPT);
    nsc = plane_split_count(tiles_c, nt * (Rp.Kb / 4));
    yslice = rs * d;
    Y.resize((size_t)(yslice * nsc));
    bs.clear();
    PlaneBatch b{};
    b.C = Y.ptr;
    b.ldc = br;
    b.slice = yslice;
    b.A = Rp.at(0);
    b.B = Qp.at(0);
    b.a_ps = Rp.ps;
    b.b_ps = Qp.ps;
    b.a_seg = 9 * Rp.ps;
    b.b_seg = 9 * Qp.ps;
    b.a_rows = Rp.Rp;
    b.b_rows = Qp.Rp;
    b.M = br;
    b.N = bl * d;
    b.nseg = nt;
    b.nks = Rp.Kb / 4;
    bs.push_back(b);
    gemm_c.prepare(bs, nsc);
}

void PlaneHeff::apply(const double2* x, double2* y, hipStream_t s) {
    const int64_t rs = (int64_t)br * bl;
    split_planes(x, 1, br, rs, op.d, Xp, 0, s);  // X_a(m, k) = x[a rs + m + k br]
    gemm_a.launch(s);
    mix_split_planes(op, P.ptr, pslice, nsa, br, bl, Qp, s);
    gemm_c.launch(s);
    reduce_slices(Y.ptr, yslice, nsc, yslice, y, s);
}

}  // namespace nc
