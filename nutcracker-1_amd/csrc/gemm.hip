// Complex GEMM for gfx950 on MFMA, FP64 (v_mfma_f64_16x16x4_f64) and FP32
// (v_mfma_f32_16x16x4_f32), batched and K-segmented:
//     C_b (M x N) = sum_s opA(A_s) opB(B_s)  (+ C_b when accumulating)
// Output tile 64 x 64 per 256-thread workgroup (4 waves in 2 x 2, each wave 2 x 2
// MFMA tiles of 16 x 16); K is staged through LDS in steps of BK, with the next
// step's global loads issued before the current step's MFMAs (register
// prefetch).  Complex products use either four real MFMAs per 16x16x4 step
// (4M: re*re - im*im, re*im + im*re) or Gauss's three (3M: re*re, im*im,
// (re+im)*(re+im), recombined at the end), which is normwise stable and saves
// a quarter of the matrix-core work.
//
// Lane layouts (measured on the MI355X): A[i][k] in lane i + 16k, B[k][j] in
// lane j + 16k; the f64 accumulator holds rows (lane / 16) + 4 v of column
// lane % 16, the f32 one rows 4 (lane / 16) + v.
//
// Workgroups are dispatched round-robin over the 8 XCDs; the linear workgroup
// id is remapped so that each XCD works on a contiguous block of output tiles
// (neighbouring tiles share rows of A or columns of B in that XCD's L2).
#include "nc_bf16.hpp"
#include "nc_internal.hpp"
#include "nc_lds_dma.hpp"

#include <algorithm>
#include <cstddef>
#include <cstdlib>
#include <cstring>
#include <chrono>
#include <map>
#include <mutex>

namespace nc {

namespace {

template <typename T>
struct Cplx;
template <>
struct Cplx<double> {
    typedef double2 type;
    typedef double v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 mfma(double a, double b, v4 c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lk, int v) { return lk + 4 * v; }
};
template <>
struct Cplx<float> {
    typedef float2 type;
    typedef float v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 mfma(float a, float b, v4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lk, int v) { return 4 * lk + v; }
};

constexpr int TBM = 64, TBN = 64, TPAD = 1;

template <typename T>
struct SegT {
    const typename Cplx<T>::type* A;
    const typename Cplx<T>::type* B;
    int64_t lda, ldb;
    int32_t K, _pad;
};
template <typename T>
struct BatchT {
    typename Cplx<T>::type* C;
    int64_t ldc;
    int32_t seg_begin, seg_end;
    int32_t M, N;      // output size of this batch (filled in on the host)
    const int* live;   // null, or a device flag: 0 = skip this batch in this launch
    int32_t upper = 0;  // 1: only the tiles on and above the diagonal (Hermitian C)
    int32_t _pad = 0;
};

template <typename T, int OPA, int OPB, int BK>
__device__ __forceinline__ void load_tiles(const SegT<T>& sg, int k0, int m0, int n0, int M, int N, int tid,
                                           typename Cplx<T>::type (&ra)[BK / 4], typename Cplx<T>::type (&rb)[BK / 4]) {
    typedef typename Cplx<T>::type CT;
    // global memory (no flat addressing), loaded as a native two-element vector
    typedef T V2 __attribute__((ext_vector_type(2)));
    typedef const V2 __attribute__((address_space(1)))* GP;
    if (m0 + 64 <= M && n0 + 64 <= N && k0 + BK <= sg.K) {
        // interior tile: every element in range, no per-element checks
        const GP A = (GP)sg.A;
        const GP B = (GP)sg.B;
#pragma unroll
        for (int it = 0; it < BK / 4; ++it) {
            const int id = tid + 256 * it;
            V2 v;
            if (OPA == OpN) v = A[(int64_t)(m0 + (id & 63)) + (int64_t)(k0 + (id >> 6)) * sg.lda];
            else v = A[(int64_t)(k0 + id % BK) + (int64_t)(m0 + id / BK) * sg.lda];
            ra[it].x = v.x;
            ra[it].y = OPA == OpC ? -v.y : v.y;
        }
#pragma unroll
        for (int it = 0; it < BK / 4; ++it) {
            const int id = tid + 256 * it;
            V2 v;
            if (OPB == OpN) v = B[(int64_t)(k0 + id % BK) + (int64_t)(n0 + id / BK) * sg.ldb];
            else v = B[(int64_t)(n0 + (id & 63)) + (int64_t)(k0 + (id >> 6)) * sg.ldb];
            rb[it].x = v.x;
            rb[it].y = OPB == OpC ? -v.y : v.y;
        }
        return;
    }
#pragma unroll
    for (int it = 0; it < BK / 4; ++it) {
        const int id = tid + 256 * it;
        int mm, kk;
        if (OPA == OpN) { mm = id & 63; kk = id >> 6; }
        else            { kk = id % BK; mm = id / BK; }
        const int gm = m0 + mm, gk = k0 + kk;
        CT v;
        v.x = 0; v.y = 0;
        if (gm < M && gk < sg.K) {
            if (OPA == OpN) v = sg.A[(int64_t)gm + (int64_t)gk * sg.lda];
            else {
                v = sg.A[(int64_t)gk + (int64_t)gm * sg.lda];
                if (OPA == OpC) v.y = -v.y;
            }
        }
        ra[it] = v;
    }
#pragma unroll
    for (int it = 0; it < BK / 4; ++it) {
        const int id = tid + 256 * it;
        int nn, kk;
        if (OPB == OpN) { kk = id % BK; nn = id / BK; }
        else            { nn = id & 63; kk = id >> 6; }
        const int gn = n0 + nn, gk = k0 + kk;
        CT v;
        v.x = 0; v.y = 0;
        if (gn < N && gk < sg.K) {
            if (OPB == OpN) v = sg.B[(int64_t)gk + (int64_t)gn * sg.ldb];
            else {
                v = sg.B[(int64_t)gn + (int64_t)gk * sg.ldb];
                if (OPB == OpC) v.y = -v.y;
            }
        }
        rb[it] = v;
    }
}

template <typename T, int OPA, int OPB, int BK>
__device__ __forceinline__ void store_tiles(int tid, typename Cplx<T>::type (*As)[TBM + TPAD],
                                            typename Cplx<T>::type (*Bs)[TBN + TPAD],
                                            const typename Cplx<T>::type (&ra)[BK / 4],
                                            const typename Cplx<T>::type (&rb)[BK / 4]) {
#pragma unroll
    for (int it = 0; it < BK / 4; ++it) {
        const int id = tid + 256 * it;
        int mm, kk;
        if (OPA == OpN) { mm = id & 63; kk = id >> 6; }
        else            { kk = id % BK; mm = id / BK; }
        As[kk][mm] = ra[it];
    }
#pragma unroll
    for (int it = 0; it < BK / 4; ++it) {
        const int id = tid + 256 * it;
        int nn, kk;
        if (OPB == OpN) { kk = id % BK; nn = id / BK; }
        else            { nn = id & 63; kk = id >> 6; }
        Bs[kk][nn] = rb[it];
    }
}

// the MFMAs of one staged K step: the wave's 2 x 2 blocks of 16 x 16
template <typename T, bool M3, int BK>
__device__ __forceinline__ void mma_stage(const typename Cplx<T>::type (*As)[TBM + TPAD],
                                          const typename Cplx<T>::type (*Bs)[TBN + TPAD],
                                          typename Cplx<T>::v4 (&acc)[M3 ? 3 : 2][2][2], int wm, int wn, int lrow,
                                          int lk) {
    typedef typename Cplx<T>::type CT;
#pragma unroll
    for (int kq = 0; kq < BK / 4; ++kq) {
        const int k = kq * 4 + lk;
        CT a[2], b[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) a[i] = As[k][wm + i * 16 + lrow];
#pragma unroll
        for (int j = 0; j < 2; ++j) b[j] = Bs[k][wn + j * 16 + lrow];
        if constexpr (M3) {
            T as[2], bs[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) as[i] = a[i].x + a[i].y;
#pragma unroll
            for (int j = 0; j < 2; ++j) bs[j] = b[j].x + b[j].y;
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    acc[0][i][j] = Cplx<T>::mfma(a[i].x, b[j].x, acc[0][i][j]);
                    acc[1][i][j] = Cplx<T>::mfma(a[i].y, b[j].y, acc[1][i][j]);
                    acc[2][i][j] = Cplx<T>::mfma(as[i], bs[j], acc[2][i][j]);
                }
        } else {
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                const T nai = -a[i].y;
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    acc[0][i][j] = Cplx<T>::mfma(a[i].x, b[j].x, acc[0][i][j]);
                    acc[0][i][j] = Cplx<T>::mfma(nai, b[j].y, acc[0][i][j]);
                    acc[1][i][j] = Cplx<T>::mfma(a[i].x, b[j].y, acc[1][i][j]);
                    acc[1][i][j] = Cplx<T>::mfma(a[i].y, b[j].x, acc[1][i][j]);
                }
            }
        }
    }
}

// NBUF = 1: one LDS stage, two barriers per K step.  NBUF = 2: two LDS stages;
// the next step is stored into the other stage while this one is consumed, one
// barrier per K step.
// The complex128 3-multiply form holds three FP64 accumulator sets (96
// registers); left to itself the compiler also keeps a whole K stage of LDS
// operands live and ends at one wave per SIMD, so it is asked for two.
template <typename T, bool M3, int BK>
constexpr int cgemm_min_waves() { return (sizeof(T) == sizeof(double) && M3 && BK == 16) ? 2 : 1; }

template <typename T, int OPA, int OPB, bool ACC, bool M3, int BK, int NBUF>
__global__ __launch_bounds__(256, (cgemm_min_waves<T, M3, BK>())) void cgemm_kernel(const BatchT<T>* __restrict__ batches,
                                                    const SegT<T>* __restrict__ segs) {
    typedef typename Cplx<T>::type CT;
    typedef typename Cplx<T>::v4 V4;
    __shared__ CT As[NBUF][BK][TBM + TPAD];
    __shared__ CT Bs[NBUF][BK][TBN + TPAD];

    // every batch has its own M x N (grouped launch of unequal problems): the
    // grid covers the largest, surplus workgroups of smaller batches exit
    const BatchT<T> bt = batches[blockIdx.y];
    if (bt.live && *bt.live == 0) return;
    const int M = bt.M, N = bt.N;
    const int tiles_m = (M + TBM - 1) / TBM, tiles_n = (N + TBN - 1) / TBN;
    const int ntiles = tiles_m * tiles_n;
    const int w = blockIdx.x;
    if (w >= ntiles) return;
    // XCD-aware tile order: workgroup w runs on XCD w % 8; give each XCD a
    // contiguous range of tiles (row-major over the tile grid)
    // (not for an upper-triangle launch: its working tiles crowd the first rows,
    // and the plain order spreads them over the XCDs)
    int tile = w;
    if (!bt.upper && (ntiles & 7) == 0 && (int)gridDim.x == ntiles) tile = (w & 7) * (ntiles >> 3) + (w >> 3);
    const int tm = tile / tiles_n, tn = tile % tiles_n;
    if (bt.upper && tm > tn) return;

    const int m0 = tm * TBM, n0 = tn * TBN;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm = (wave & 1) * 32;
    const int wn = (wave >> 1) * 32;
    const int lrow = lane & 15;
    const int lk = lane >> 4;

    constexpr int NACC = M3 ? 3 : 2;
    V4 acc[NACC][2][2];
#pragma unroll
    for (int q = 0; q < NACC; ++q)
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) acc[q][i][j] = V4{0, 0, 0, 0};

    CT ra[BK / 4], rb[BK / 4];
    int sidx = bt.seg_begin;
    int k0 = 0;
    // the current segment stays in registers; it is re-read only when K moves
    // on to the next one
    SegT<T> sg;
    if (sidx < bt.seg_end) {
        sg = segs[sidx];
        load_tiles<T, OPA, OPB, BK>(sg, k0, m0, n0, M, N, tid, ra, rb);
    }
    if (NBUF == 1) {
        while (sidx < bt.seg_end) {
            store_tiles<T, OPA, OPB, BK>(tid, As[0], Bs[0], ra, rb);
            __syncthreads();
            // advance and prefetch the next stage
            k0 += BK;
            if (k0 >= sg.K) {
                ++sidx;
                k0 = 0;
                if (sidx < bt.seg_end) sg = segs[sidx];
            }
            if (sidx < bt.seg_end) load_tiles<T, OPA, OPB, BK>(sg, k0, m0, n0, M, N, tid, ra, rb);
            mma_stage<T, M3, BK>(As[0], Bs[0], acc, wm, wn, lrow, lk);
            __syncthreads();
        }
    } else {
        int cur = 0;
        if (sidx < bt.seg_end) store_tiles<T, OPA, OPB, BK>(tid, As[0], Bs[0], ra, rb);
        __syncthreads();
        while (sidx < bt.seg_end) {
            k0 += BK;
            if (k0 >= sg.K) {
                ++sidx;
                k0 = 0;
                if (sidx < bt.seg_end) sg = segs[sidx];
            }
            const bool more = sidx < bt.seg_end;
            if (more) load_tiles<T, OPA, OPB, BK>(sg, k0, m0, n0, M, N, tid, ra, rb);
            mma_stage<T, M3, BK>(As[cur], Bs[cur], acc, wm, wn, lrow, lk);
            // the other stage was last read before the previous barrier
            if (more) store_tiles<T, OPA, OPB, BK>(tid, As[cur ^ 1], Bs[cur ^ 1], ra, rb);
            __syncthreads();
            cur ^= 1;
        }
    }

    CT* __restrict__ C = bt.C;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int col = n0 + wn + j * 16 + lrow;
            if (col >= N) continue;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = m0 + wm + i * 16 + Cplx<T>::row(lk, v);
                if (row < M) {
                    CT o;
                    if (M3) {
                        const T t1 = acc[0][i][j][v], t2 = acc[1][i][j][v], t3 = acc[2][i][j][v];
                        o.x = t1 - t2;
                        o.y = t3 - t1 - t2;
                    } else {
                        o.x = acc[0][i][j][v];
                        o.y = acc[1][i][j][v];
                    }
                    CT* p = C + (int64_t)row + (int64_t)col * bt.ldc;
                    if (ACC) {
                        CT c = *p;
                        o.x += c.x;
                        o.y += c.y;
                    }
                    *p = o;
                }
            }
        }
}

// ---- complex128 GEMM, both operands untransposed, fed by global -> LDS copies
//
// For plans whose batches all have M and N multiples of 64 and whose segments
// all have K a multiple of 8 (the bulk sites of a chain once its bonds are
// multiples of 64).  Each K stage of 8 is copied straight into LDS by
// global_load_lds_dwordx4 (one complex128 per lane and copy), NST - 1 stages
// ahead of its use, into a ring of NST LDS stages.  No prefetch registers are
// held, so the 3-multiply form fits three waves per SIMD (three workgroups per
// CU at NST = 3) instead of two, and a stage's copies have NST - 1 stages of
// matrix-core work to land in.  Output tile, wave layout, MFMAs and epilogue
// are those of cgemm_kernel.
//
// LDS stage (16 KB): a[k][m] as A lies in memory (64 consecutive m of one k
// are one 1 KB run, one copy instruction), and b[n][k'] with the k slot of
// column n permuted, k' = k ^ ((n / 2) mod 8): B's columns are K-contiguous,
// so one copy instruction takes 8 columns x 8 k (128 contiguous bytes per
// column), and without the permutation the 16 lanes that read one MFMA B
// fragment (16 consecutive n at one k, 128 bytes apart) would fall on two
// 16-byte bank groups; with it they cover all sixteen.
// Copies per stage: 8 runs of A (one per k) and 8 of B (8 columns each); wave
// w issues A runs 2w, 2w + 1 and B runs 2w, 2w + 1, four per wave, so a wave
// waiting for stage s leaves 4 (NST - 2) newer copies in flight.
constexpr int kDmaF64 = 8;   // launch choice codes of this kernel (GemmChoice::bk): NST = 3
constexpr int kDmaF64x4 = 9;  // NST = 4
constexpr int DBK = 8;
struct DmaStage {
    double2 a[DBK][64];
    double2 b[64 * DBK];
};

// the k slot of (column n, k) in a stage's b
__device__ __forceinline__ int dma_bslot(int n, int k) { return n * DBK + (k ^ ((n >> 1) & 7)); }

struct DmaOffsets {
    unsigned a0, a1, b0, b1;
};
// one segment's per-lane copy offsets (bytes from the stage's A / B base).
// BC: op(B) = B^H with B stored N x K (the Gram matrix X X^H), so op(B)'s
// rows at one k are 64 consecutive elements of B's column k, copied as A's
// runs are, into b viewed as [k][n]; the conjugate is taken when the
// fragments are read.
template <bool BC>
__device__ __forceinline__ DmaOffsets dma_offsets(int64_t lda, int64_t ldb, int wave, int lane) {
    DmaOffsets o;
    o.a0 = 16u * (unsigned)(lane + (2 * wave) * lda);
    o.a1 = 16u * (unsigned)(lane + (2 * wave + 1) * lda);
    if constexpr (BC) {
        o.b0 = 16u * (unsigned)(lane + (2 * wave) * ldb);
        o.b1 = 16u * (unsigned)(lane + (2 * wave + 1) * ldb);
    } else {
        const int n0 = 8 * (2 * wave) + (lane >> 3), n1 = n0 + 8, ks = lane & 7;
        o.b0 = 16u * (unsigned)((ks ^ ((n0 >> 1) & 7)) + n0 * ldb);
        o.b1 = 16u * (unsigned)((ks ^ ((n1 >> 1) & 7)) + n1 * ldb);
    }
    return o;
}

// this wave's copies of every stage but the NST - 2 newest have landed, its LDS
// reads are done, then the workgroup barrier; the memory clobber keeps the
// compiler from moving LDS reads across it
template <int NST>
__device__ __forceinline__ void dma_stage_barrier() {
    static_assert(NST == 3 || NST == 4, "stage ring of 3 or 4");
    if constexpr (NST == 3) asm volatile("s_waitcnt vmcnt(4)\n\ts_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
    else asm volatile("s_waitcnt vmcnt(8)\n\ts_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

template <bool M3, bool BC>
__device__ __forceinline__ void dma_compute(const DmaStage& S, Cplx<double>::v4 (&acc)[M3 ? 3 : 2][2][2], int wm,
                                            int wn, int lrow, int lk) {
    const double2(*bt)[64] = reinterpret_cast<const double2(*)[64]>(S.b);
#pragma unroll
    for (int q = 0; q < DBK / 4; ++q) {
        const int k = 4 * q + lk;
        double2 a[2], b[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) a[i] = S.a[k][wm + 16 * i + lrow];
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            if constexpr (BC) {
                b[j] = bt[k][wn + 16 * j + lrow];
                b[j].y = -b[j].y;
            } else {
                b[j] = S.b[dma_bslot(wn + 16 * j + lrow, k)];
            }
        }
        if constexpr (M3) {
            double as[2], bs[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) as[i] = a[i].x + a[i].y;
#pragma unroll
            for (int j = 0; j < 2; ++j) bs[j] = b[j].x + b[j].y;
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    acc[0][i][j] = Cplx<double>::mfma(a[i].x, b[j].x, acc[0][i][j]);
                    acc[1][i][j] = Cplx<double>::mfma(a[i].y, b[j].y, acc[1][i][j]);
                    acc[2][i][j] = Cplx<double>::mfma(as[i], bs[j], acc[2][i][j]);
                }
        } else {
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                const double nai = -a[i].y;
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    acc[0][i][j] = Cplx<double>::mfma(a[i].x, b[j].x, acc[0][i][j]);
                    acc[0][i][j] = Cplx<double>::mfma(nai, b[j].y, acc[0][i][j]);
                    acc[1][i][j] = Cplx<double>::mfma(a[i].x, b[j].y, acc[1][i][j]);
                    acc[1][i][j] = Cplx<double>::mfma(a[i].y, b[j].x, acc[1][i][j]);
                }
            }
        }
    }
}

// Tile order over the 8 XCDs (workgroups are dispatched round-robin over them,
// so linear workgroup id l runs on XCD l mod 8), chosen on the host for plans
// whose batches share one output size and whose tile count per batch is a
// multiple of 8:
//   xmap 1 (several batches): each XCD takes one eighth of the (batch, tile)
//     range, batch-major, so a batch's tiles run on one or two XCDs and its
//     operand panels are read into their L2 only;
//   xmap 2 (one batch): the tile grid is cut into xa x xb = 8 blocks, one per
//     XCD, so that an XCD reads tiles_m / xa row panels of A and tiles_n / xb
//     column panels of B instead of all of them;
//   xmap 0: each XCD takes a contiguous row-major range of every batch's tiles.
// Within an XCD's share the tiles run in row-major order.
template <bool ACC, bool M3, int NST, bool BC>
__global__ __launch_bounds__(256, (NST == 3 && M3) ? 3 : 2) void cgemm_dma_kernel(
    const BatchT<double>* __restrict__ batches, const SegT<double>* __restrict__ segs, int xmap, int xa, int xb) {
    typedef Cplx<double>::v4 V4;
    __shared__ DmaStage S[NST];

    int bidx = blockIdx.y, w = blockIdx.x;
    if (xmap == 1) {
        const int lin = blockIdx.y * gridDim.x + blockIdx.x;
        const int item = (lin & 7) * ((gridDim.y * gridDim.x) >> 3) + (lin >> 3);
        bidx = item / gridDim.x;
        w = item - bidx * gridDim.x;
    }
    const BatchT<double> bt = batches[bidx];
    if (bt.live && *bt.live == 0) return;
    const int M = bt.M, N = bt.N;
    const int tiles_m = M / TBM, tiles_n = N / TBN;
    const int ntiles = tiles_m * tiles_n;
    if (w >= ntiles) return;
    int tm, tn;
    if (xmap == 2) {
        const int bm = tiles_m / xa, bn = tiles_n / xb, x = w & 7, slot = w >> 3;
        tm = (x / xb) * bm + slot / bn;
        tn = (x % xb) * bn + slot % bn;
    } else {
        int tile = w;
        if (xmap == 0 && !bt.upper && (ntiles & 7) == 0 && (int)gridDim.x == ntiles)
            tile = (w & 7) * (ntiles >> 3) + (w >> 3);
        tm = tile / tiles_n;
        tn = tile % tiles_n;
    }
    if (bt.upper && tm > tn) return;

    const int m0 = tm * TBM, n0 = tn * TBN;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm = (wave & 1) * 32;
    const int wn = (wave >> 1) * 32;
    const int lrow = lane & 15;
    const int lk = lane >> 4;

    constexpr int NACC = M3 ? 3 : 2;
    V4 acc[NACC][2][2];
#pragma unroll
    for (int q = 0; q < NACC; ++q)
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) acc[q][i][j] = V4{0, 0, 0, 0};

    int sidx = bt.seg_begin;
    if (sidx < bt.seg_end) {
        // issue cursor: the segment and K offset of the next stage to copy; past
        // the last stage it stays there, so every step issues the same number of
        // copies (the extra ones land in stages that are never read)
        SegT<double> sg = segs[sidx];
        int k0 = 0;
        DmaOffsets off = dma_offsets<BC>(sg.lda, sg.ldb, wave, lane);
        int nks = 0;
        for (int i = bt.seg_begin; i < bt.seg_end; ++i) nks += segs[i].K / DBK;
        auto issue = [&](DmaStage& st) {
            const double2* pa = sg.A + m0 + (int64_t)k0 * sg.lda;
            const double2* pb = BC ? sg.B + n0 + (int64_t)k0 * sg.ldb : sg.B + k0 + (int64_t)n0 * sg.ldb;
            lds_dma16(pa, off.a0, &st.a[2 * wave][0]);
            lds_dma16(pa, off.a1, &st.a[2 * wave + 1][0]);
            lds_dma16(pb, off.b0, &st.b[64 * (2 * wave)]);
            lds_dma16(pb, off.b1, &st.b[64 * (2 * wave + 1)]);
            if (k0 + DBK < sg.K) {
                k0 += DBK;
            } else if (sidx + 1 < bt.seg_end) {
                ++sidx;
                sg = segs[sidx];
                k0 = 0;
                off = dma_offsets<BC>(sg.lda, sg.ldb, wave, lane);
            }
        };
#pragma unroll
        for (int s = 0; s < NST - 1; ++s) issue(S[s]);
        int cur = 0;
        for (int ks = 0; ks < nks; ++ks) {
            dma_stage_barrier<NST>();
            // the stage consumed in the previous step is free for the copies of
            // the step NST - 1 ahead
            issue(S[cur == 0 ? NST - 1 : cur - 1]);
            dma_compute<M3, BC>(S[cur], acc, wm, wn, lrow, lk);
            cur = cur == NST - 1 ? 0 : cur + 1;
        }
        // no copy may still write LDS when the workgroup ends
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    }

    double2* __restrict__ C = bt.C;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int col = n0 + wn + j * 16 + lrow;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = m0 + wm + i * 16 + Cplx<double>::row(lk, v);
                double2 o;
                if (M3) {
                    const double t1 = acc[0][i][j][v], t2 = acc[1][i][j][v], t3 = acc[2][i][j][v];
                    o.x = t1 - t2;
                    o.y = t3 - t1 - t2;
                } else {
                    o.x = acc[0][i][j][v];
                    o.y = acc[1][i][j][v];
                }
                double2* p = C + (int64_t)row + (int64_t)col * bt.ldc;
                if (ACC) {
                    const double2 c = *p;
                    o.x += c.x;
                    o.y += c.y;
                }
                *p = o;
            }
        }
}

// ---- complex64 GEMM on the bf16 matrix cores (both operands untransposed)
//
// Every FP32 operand value v is split into three bf16 parts, v = hi + mid + lo
// + O(2^-27 |v|) (each part the round-to-nearest bf16 of what the previous
// ones leave; the residuals are exact in FP32), and a real product x y is
// formed from the six partial products whose size is above 2^-24 |x y|:
// hi.hi, hi.mid, mid.hi, hi.lo, mid.mid, lo.hi, all accumulated in FP32 by
// v_mfma_f32_16x16x32_bf16.  The dropped terms are below 3 2^-27 |x y|, so the
// result has FP32 accuracy, at 6/16 of the FP32 matrix-core time per product
// (the bf16 rate is 16x the FP32 one).  The complex product is Gauss's 3M
// form, so the three real matrices re, im, re + im of each operand are split.
//
// Layout: 64 x 64 output tile per 256-thread workgroup, K staged 32 deep.  The
// staging thread (row r = lane, K group g = wave) loads 8 consecutive K values
// of row r, splits them and writes one 16-byte chunk per (matrix, part) into
// LDS planes [9][64 rows][4 chunks of 8 bf16], chunk g stored at a position
// permuted by row (b6_chunk) so that the 16-byte reads of consecutive rows hit
// distinct banks.
// MFMA operands (measured lane layout): lane l holds row l % 16 of the block
// and K values 8 (l / 16) .. 8 (l / 16) + 7; the accumulator holds rows
// 4 (l / 16) + v of column l % 16, as the FP32 16x16x4 one.
constexpr int B6K = 32;
constexpr int kBf16x6 = 6;  // launch choice code of this kernel (GemmChoice::bk)

// the 8 complex values of row r (M index for A, N index for B) at K positions
// k0 + 8 g .. + 7 of the current segment; zero outside the matrix
template <bool IS_A>
__device__ __forceinline__ void b6_load(const SegT<float>& sg, int k0, int r0, int R, int r, int g, float2 (&v)[8]) {
    // global memory (not flat: flat loads would share the LDS wait counter)
    typedef float V2 __attribute__((ext_vector_type(2)));
    typedef const V2 __attribute__((address_space(1)))* GP;
    const int row = r0 + r;
    const int kb = k0 + 8 * g;
    const GP P = (GP)(IS_A ? sg.A : sg.B);
    const int64_t ld = IS_A ? sg.lda : sg.ldb;
    // A(m, k) at A[m + k lda] (lanes consecutive in m); B(k, n) at B[k + n ldb]
    // (the 8 values of a lane contiguous)
    const int64_t base = IS_A ? (int64_t)row + (int64_t)kb * ld : (int64_t)kb + (int64_t)row * ld;
    const int64_t step = IS_A ? ld : 1;
    if (row < R && kb + 8 <= sg.K) {
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const V2 t = P[base + e * step];
            v[e] = make_float2(t.x, t.y);
        }
    } else {
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            v[e] = make_float2(0.f, 0.f);
            if (row < R && kb + e < sg.K) {
                const V2 t = P[base + e * step];
                v[e] = make_float2(t.x, t.y);
            }
        }
    }
}

// LDS chunk of (row r, K chunk g): rows are 64 bytes (four 16-byte chunks), so
// a 16-byte access of consecutive rows meets every bank once only if the chunk
// is permuted by row.  The permutation f(r) takes four distinct values on the
// even and on the odd rows of every 8 consecutive rows (conflict-free when 8
// lanes are served per cycle, 32 banks) and on each residue mod 4 of every 16
// rows (16 lanes per cycle, 64 banks); the earlier f(r) = (r >> 2) & 3 met the
// second condition only, and PMC counted half the LDS cycles as bank conflicts.
__device__ __forceinline__ int b6_chunk(int r, int g) { return g ^ ((r >> 2) & 3) ^ (((r >> 1) & 1) << 1); }

// splits re, im and re + im of the 8 values into the 9 planes of row r, chunk g
__device__ __forceinline__ void b6_store(bf16x8 (*S)[64][4], int r, int g, const float2 (&v)[8]) {
    const int c = b6_chunk(r, g);
    float x[8];
    bf16x8 h, m, l;
#pragma unroll
    for (int e = 0; e < 8; ++e) x[e] = v[e].x;
    split3(x, h, m, l);
    S[0][r][c] = h; S[1][r][c] = m; S[2][r][c] = l;
#pragma unroll
    for (int e = 0; e < 8; ++e) x[e] = v[e].y;
    split3(x, h, m, l);
    S[3][r][c] = h; S[4][r][c] = m; S[5][r][c] = l;
#pragma unroll
    for (int e = 0; e < 8; ++e) x[e] = v[e].x + v[e].y;
    split3(x, h, m, l);
    S[6][r][c] = h; S[7][r][c] = m; S[8][r][c] = l;
}

template <bool ACC>
__global__ __launch_bounds__(256) void cgemm_b6_kernel(const BatchT<float>* __restrict__ batches,
                                                       const SegT<float>* __restrict__ segs) {
    __shared__ bf16x8 As[9][64][4];
    __shared__ bf16x8 Bs[9][64][4];

    const BatchT<float> bt = batches[blockIdx.y];
    if (bt.live && *bt.live == 0) return;
    const int M = bt.M, N = bt.N;
    const int tiles_m = (M + TBM - 1) / TBM, tiles_n = (N + TBN - 1) / TBN;
    const int ntiles = tiles_m * tiles_n;
    const int w = blockIdx.x;
    if (w >= ntiles) return;
    int tile = w;
    if (!bt.upper && (ntiles & 7) == 0 && (int)gridDim.x == ntiles) tile = (w & 7) * (ntiles >> 3) + (w >> 3);
    const int tm = tile / tiles_n, tn = tile % tiles_n;
    if (bt.upper && tm > tn) return;
    const int m0 = tm * TBM, n0 = tn * TBN;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm = (wave & 1) * 32;
    const int wn = (wave >> 1) * 32;
    const int lrow = lane & 15;
    const int lk = lane >> 4;

    f32x4 acc[3][2][2];
#pragma unroll
    for (int q = 0; q < 3; ++q)
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) acc[q][i][j] = f32x4{0, 0, 0, 0};

    // K stages of this batch (each segment in stages of 32, its tail zero-filled)
    int nst = 0;
    for (int i = bt.seg_begin; i < bt.seg_end; ++i) nst += segs[i].K > 0 ? (segs[i].K + B6K - 1) / B6K : 0;
    // loads run two stages ahead (two register sets): one stage's MFMAs are too
    // short to hide the load latency at the two workgroups per CU that the LDS
    // planes allow
    // (segments with K = 0 have no stage and are stepped over)
    int lsidx = bt.seg_begin, lk0 = 0;
    SegT<float> lsg;
    auto next_seg = [&]() {
        while (lsidx < bt.seg_end && (lsg = segs[lsidx]).K <= 0) ++lsidx;
    };
    next_seg();
    auto load_next = [&](float2 (&va)[8], float2 (&vb)[8]) {
        b6_load<true>(lsg, lk0, m0, M, lane, wave, va);
        b6_load<false>(lsg, lk0, n0, N, lane, wave, vb);
        lk0 += B6K;
        if (lk0 >= lsg.K) {
            ++lsidx;
            lk0 = 0;
            next_seg();
        }
    };
    auto stage = [&](float2 (&va)[8], float2 (&vb)[8], int st) {
        b6_store(As, lane, wave, va);
        b6_store(Bs, lane, wave, vb);
        __syncthreads();
        if (st + 2 < nst) load_next(va, vb);
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            bf16x8 a[2][3], b[2][3];
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                const int r = wm + i * 16 + lrow;
                const int c = b6_chunk(r, lk);
#pragma unroll
                for (int p = 0; p < 3; ++p) a[i][p] = As[3 * q + p][r][c];
            }
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const int r = wn + j * 16 + lrow;
                const int c = b6_chunk(r, lk);
#pragma unroll
                for (int p = 0; p < 3; ++p) b[j][p] = Bs[3 * q + p][r][c];
            }
            // partial products from the smallest up (lo.hi, mid.mid, hi.lo,
            // mid.hi, hi.mid, hi.hi), the four accumulators interleaved
            constexpr int pa[6] = {2, 1, 0, 1, 0, 0};
            constexpr int pb[6] = {0, 1, 2, 0, 1, 0};
#pragma unroll
            for (int t = 0; t < 6; ++t)
#pragma unroll
                for (int i = 0; i < 2; ++i)
#pragma unroll
                    for (int j = 0; j < 2; ++j)
                        acc[q][i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i][pa[t]], b[j][pb[t]], acc[q][i][j],
                                                                               0, 0, 0);
        }
        __syncthreads();
    };
    float2 va0[8], vb0[8], va1[8], vb1[8];
    if (nst > 0) load_next(va0, vb0);
    if (nst > 1) load_next(va1, vb1);
    for (int st = 0; st < nst; st += 2) {
        stage(va0, vb0, st);
        if (st + 1 < nst) stage(va1, vb1, st + 1);
    }

    float2* __restrict__ C = bt.C;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int col = n0 + wn + j * 16 + lrow;
            if (col >= N) continue;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = m0 + wm + i * 16 + 4 * lk + v;
                if (row < M) {
                    const float t1 = acc[0][i][j][v], t2 = acc[1][i][j][v], t3 = acc[2][i][j][v];
                    float2 o = make_float2(t1 - t2, t3 - t1 - t2);
                    float2* p = C + (int64_t)row + (int64_t)col * bt.ldc;
                    if (ACC) {
                        const float2 c = *p;
                        o.x += c.x;
                        o.y += c.y;
                    }
                    *p = o;
                }
            }
        }
}

// C_b = sum_c W_{b,c} (+ C_b) for the split-K partial products
template <typename CT>
__global__ void splitk_reduce_kernel(const CT* __restrict__ W, int64_t slice, int nsplit, CT* const* __restrict__ Cptr,
                                     const int64_t* __restrict__ ldc, const int* const* __restrict__ live, int M, int N,
                                     int accumulate) {
    const int b = blockIdx.y;
    if (live[b] && *live[b] == 0) return;
    const int64_t total = (int64_t)M * N;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = e % M, c = e / M;
        CT s;
        s.x = 0; s.y = 0;
        for (int q = 0; q < nsplit; ++q) {
            const CT v = W[((int64_t)b * nsplit + q) * slice + e];
            s.x += v.x;
            s.y += v.y;
        }
        CT* p = Cptr[b] + r + c * ldc[b];
        if (accumulate) {
            const CT o = *p;
            s.x += o.x;
            s.y += o.y;
        }
        *p = s;
    }
}

// ---------------------------------------------------------------------------
// Host side: descriptor staging ring and launch
// ---------------------------------------------------------------------------
struct Arena {
    static constexpr size_t kSlot = 128 * 1024;
    static constexpr int kSlots = 64;
    char* dev = nullptr;
    char* host = nullptr;
    hipEvent_t done[kSlots];
    int next = 0;
    Arena() {
        NC_HIP(hipMalloc(&dev, kSlot * kSlots));
        NC_HIP(hipHostMalloc((void**)&host, kSlot * kSlots, hipHostMallocDefault));
        for (int i = 0; i < kSlots; ++i) NC_HIP(hipEventCreateWithFlags(&done[i], hipEventDisableTiming));
        for (int i = 0; i < kSlots; ++i) NC_HIP(hipEventRecord(done[i], Context::get().stream));
    }
    // copies `bytes` of host data into the next slot; returns the device address
    char* stage(const void* const* parts, const size_t* sizes, int nparts, size_t* offsets, hipStream_t s, int* slot_out) {
        size_t off = 0;
        for (int i = 0; i < nparts; ++i) {
            offsets[i] = off;
            off = (off + sizes[i] + 255) & ~size_t(255);
        }
        if (off > kSlot) {
            // a list larger than a ring slot: one device buffer of its own, written
            // by a stream-ordered copy (earlier launches reading it have finished
            // before the copy runs)
            std::vector<char> h(off);
            for (int i = 0; i < nparts; ++i)
                if (sizes[i]) std::memcpy(h.data() + offsets[i], parts[i], sizes[i]);
            if (big.n < off) big.resize(off);
            NC_HIP(hipMemcpyAsync(big.ptr, h.data(), off, hipMemcpyHostToDevice, s));
            NC_HIP(hipStreamSynchronize(s));
            *slot_out = -1;
            return big.ptr;
        }
        int slot = next;
        next = (next + 1) % kSlots;
        NC_HIP(hipEventSynchronize(done[slot]));
        char* h = host + slot * kSlot;
        for (int i = 0; i < nparts; ++i)
            if (sizes[i]) std::memcpy(h + offsets[i], parts[i], sizes[i]);
        char* d = dev + slot * kSlot;
        NC_HIP(hipMemcpyAsync(d, h, off, hipMemcpyHostToDevice, s));
        *slot_out = slot;
        return d;
    }
    void release(int slot, hipStream_t s) {
        if (slot >= 0) NC_HIP(hipEventRecord(done[slot], s));
    }
    DevBuf<char> big;  // lists larger than a slot
};
Arena& arena() {
    static Arena* a = new Arena();
    return *a;
}

template <typename T, int OPA, int OPB, int BK, int NBUF>
void launch_bk(bool acc, bool m3, dim3 grid, const BatchT<T>* b, const SegT<T>* s, hipStream_t st) {
#define NC_L(ACCV, M3V) \
    hipLaunchKernelGGL((cgemm_kernel<T, OPA, OPB, ACCV, M3V, BK, NBUF>), grid, dim3(256), 0, st, b, s)
    if (acc) { if (m3) NC_L(true, true); else NC_L(true, false); }
    else { if (m3) NC_L(false, true); else NC_L(false, false); }
#undef NC_L
}

// bk: 16 or 32 (one LDS stage), or -16 (16 with two LDS stages)
template <typename T, int OPA, int OPB>
void launch_typed(int bk, bool acc, bool m3, dim3 grid, const BatchT<T>* b, const SegT<T>* s, hipStream_t st) {
    if (bk == 32) launch_bk<T, OPA, OPB, 32, 1>(acc, m3, grid, b, s, st);
    else if (bk == -16) launch_bk<T, OPA, OPB, 16, 2>(acc, m3, grid, b, s, st);
    else launch_bk<T, OPA, OPB, 16, 1>(acc, m3, grid, b, s, st);
}

template <int NST, bool BC>
void launch_dma(bool acc, bool m3, dim3 grid, const BatchT<double>* b, const SegT<double>* s, int xmap, int xa,
                int xb, hipStream_t st) {
#define NC_L(ACCV, M3V) \
    hipLaunchKernelGGL((cgemm_dma_kernel<ACCV, M3V, NST, BC>), grid, dim3(256), 0, st, b, s, xmap, xa, xb)
    if (acc) { if (m3) NC_L(true, true); else NC_L(true, false); }
    else { if (m3) NC_L(false, true); else NC_L(false, false); }
#undef NC_L
}

// tile order of an LDS-copy launch (see cgemm_dma_kernel)
struct XcdMap {
    int mode = 0, a = 1, b = 1;
};

template <typename T>
void launch_any(int opA, int opB, int bk, bool acc, bool m3, dim3 grid, const BatchT<T>* b, const SegT<T>* s,
                hipStream_t st, XcdMap xm = XcdMap()) {
    if (bk == kDmaF64 || bk == kDmaF64x4) {
        // build_plan keeps these codes only for eligible complex128 N-N and N-C plans
        if constexpr (sizeof(T) == sizeof(double)) {
            if (opA == OpN && (opB == OpN || opB == OpC)) {
                const auto* bd = reinterpret_cast<const BatchT<double>*>(b);
                const auto* sd = reinterpret_cast<const SegT<double>*>(s);
                if (opB == OpN) {
                    if (bk == kDmaF64) launch_dma<3, false>(acc, m3, grid, bd, sd, xm.mode, xm.a, xm.b, st);
                    else launch_dma<4, false>(acc, m3, grid, bd, sd, xm.mode, xm.a, xm.b, st);
                } else {
                    if (bk == kDmaF64) launch_dma<3, true>(acc, m3, grid, bd, sd, xm.mode, xm.a, xm.b, st);
                    else launch_dma<4, true>(acc, m3, grid, bd, sd, xm.mode, xm.a, xm.b, st);
                }
                return;
            }
        }
        bk = 16;
    }
    if (bk == kBf16x6) {
        if constexpr (sizeof(T) == sizeof(float)) {
            if (opA == OpN && opB == OpN) {
                if (acc) hipLaunchKernelGGL(cgemm_b6_kernel<true>, grid, dim3(256), 0, st, b, s);
                else hipLaunchKernelGGL(cgemm_b6_kernel<false>, grid, dim3(256), 0, st, b, s);
                return;
            }
        }
        bk = 16;  // complex128 and transposed operands: the FP MFMA kernel
    }
#define NC_O(a, bb) \
    if (opA == a && opB == bb) { launch_typed<T, a, bb>(bk, acc, m3, grid, b, s, st); return; }
    NC_O(OpN, OpN) NC_O(OpN, OpT) NC_O(OpN, OpC) NC_O(OpT, OpN) NC_O(OpC, OpN) NC_O(OpC, OpC) NC_O(OpT, OpT)
#undef NC_O
    throw Error("GEMM: unsupported operand combination", -71);
}

GemmTuning g_tuning;

// splits every batch's K range into `nsplit` nearly equal chunks (segment
// boundaries respected by cutting segments), each chunk its own batch
template <typename Seg>
void split_segments(std::vector<Seg> const& segs, int sb, int se, int nsplit, int opA, int opB,
                    std::vector<std::vector<Seg>>& out) {
    int64_t total = 0;
    for (int i = sb; i < se; ++i) total += segs[i].K;
    out.assign(nsplit, {});
    int64_t per = (total + nsplit - 1) / nsplit;
    per = ((per + 15) / 16) * 16;
    int chunk = 0;
    int64_t filled = 0;
    for (int i = sb; i < se; ++i) {
        Seg s = segs[i];
        int64_t k = 0;
        while (k < s.K) {
            if (chunk >= nsplit) chunk = nsplit - 1;
            int64_t take = std::min<int64_t>(s.K - k, per - filled);
            if (chunk == nsplit - 1) take = s.K - k;
            Seg p = s;
            p.K = (int32_t)take;
            // advance the operand pointers to column / row k of op(A), op(B)
            p.A = s.A + (opA == OpN ? k * s.lda : k);
            p.B = s.B + (opB == OpN ? k : k * s.ldb);
            out[chunk].push_back(p);
            k += take;
            filled += take;
            if (filled >= per && chunk < nsplit - 1) { ++chunk; filled = 0; }
        }
    }
}

// A plan's launch shape and final descriptor lists (split-K expanded)
template <typename T>
struct Built {
    typedef typename Cplx<T>::type CT;
    std::vector<BatchT<T>> b;
    std::vector<SegT<T>> sg;
    std::vector<CT*> cptr;          // split-K: the real outputs, one per original batch
    std::vector<int64_t> ldcs;
    std::vector<const int*> lives;
    int opA = OpN, opB = OpN;
    bool acc = false, m3 = false;
    int bk = 16;
    int nsplit = 1;
    int64_t slice = 0, maxtiles = 0;
    int M = 0, N = 0;
    CT* W = nullptr;                // split-K workspace
    XcdMap xmap;                    // tile order of an LDS-copy launch
    bool empty() const { return b.empty() || maxtiles <= 0; }
};

// Launch choices of one plan: the kernel and its K step per LDS round (16, 32,
// -16 for 16 with two LDS stages, kBf16x6, kDmaF64 / kDmaF64x4; 0 = the
// default: the LDS-copy pipeline where the plan allows it, else 16), K split
// (0 = the heuristic below), complex product form (-1 = the precision's
// default).
constexpr int kDefaultKernel = 0;
struct GemmChoice {
    int bk = kDefaultKernel;
    int nsplit = 0;
    int m3 = -1;
};
// a choice forced on every plan (tests of each kernel variant); bk = 0: none
GemmChoice g_forced{0, 0, -1};

// whether a built plan can run on cgemm_dma_kernel: complex128, A untransposed
// and B untransposed or conjugate-transposed, every batch's M and N multiples
// of the 64 x 64 tile, every segment's K a positive multiple of the 8-deep
// stage, and the per-lane copy offsets (at most 7 columns of A, and 63 columns
// of B or 7 for a conjugate-transposed B, past a stage's base) within 32 bits
template <typename T>
bool dma_shapes_ok(int opA, int opB, std::vector<BatchT<T>> const& batches, std::vector<SegT<T>> const& segs) {
    if (sizeof(T) != sizeof(double) || opA != OpN || (opB != OpN && opB != OpC)) return false;
    const bool bc = opB == OpC;
    for (auto const& b : batches) {
        if (b.M <= 0 || b.N <= 0 || b.M % TBM || b.N % TBN) return false;
        for (int i = b.seg_begin; i < b.seg_end; ++i) {
            auto const& sg = segs[i];
            if (sg.K <= 0 || sg.K % DBK || sg.lda < b.M || sg.ldb < (bc ? b.N : sg.K)) return false;
            if (16 * (7 * sg.lda + 64) >= (int64_t(1) << 31)) return false;
            if (16 * (bc ? 7 * sg.ldb + 64 : 63 * sg.ldb + 8) >= (int64_t(1) << 31)) return false;
        }
    }
    return true;
}
inline bool is_dma(int bk) { return bk == kDmaF64 || bk == kDmaF64x4; }
// the final kernel of a built plan (an LDS-copy choice the split lists do not
// allow falls back to the register-staged kernel) and its XCD tile order
template <typename T>
void settle_kernel(Built<T>& bp) {
    if (is_dma(bp.bk) && !dma_shapes_ok(bp.opA, bp.opB, bp.b, bp.sg)) bp.bk = 16;
    bp.xmap = XcdMap();
    if (!is_dma(bp.bk) || !g_tuning.dma_xcd_blocks) return;
    bool uniform = true;
    for (auto const& b : bp.b) uniform = uniform && b.M == bp.b[0].M && b.N == bp.b[0].N && !b.upper;
    const int tm = bp.b[0].M / TBM, tn = bp.b[0].N / TBN, nb = (int)bp.b.size();
    if (!uniform || (tm * tn) % 8 || bp.maxtiles != (int64_t)tm * tn) return;
    if (nb > 1) {
        bp.xmap.mode = 1;
        return;
    }
    // the 8 blocks with the fewest panels per XCD (tiles_m / a + tiles_n / b)
    int best = 1 << 30;
    for (int a : {1, 2, 4, 8}) {
        const int b = 8 / a;
        if (tm % a || tn % b) continue;
        if (tm / a + tn / b < best) {
            best = tm / a + tn / b;
            bp.xmap = XcdMap{2, a, b};
        }
    }
}

// Fills per-batch sizes, chooses split-K and expands it.  `work(bytes)` gives the
// split-K workspace.
template <typename T, typename WorkFn>
void build_plan(int opA, int opB, bool accumulate, int M, int N, std::vector<BatchT<T>>& batches,
                std::vector<SegT<T>> const& segs, Built<T>& out, WorkFn work, GemmChoice ch = GemmChoice()) {
    typedef typename Cplx<T>::type CT;
    out = Built<T>();
    out.opA = opA;
    out.opB = opB;
    out.acc = accumulate;
    if (batches.empty()) return;
    // per-batch sizes (0 = the plan's); a grouped launch of unequal batches
    bool uniform = true;
    int64_t maxtiles = 0, tiles = 0;
    for (auto& b : batches) {
        if (b.M == 0) b.M = M;
        if (b.N == 0) b.N = N;
        if (b.M != batches[0].M || b.N != batches[0].N) uniform = false;
        const int64_t t = (int64_t)((b.M + TBM - 1) / TBM) * ((b.N + TBN - 1) / TBN);
        maxtiles = std::max(maxtiles, t);
        tiles += t;
    }
    out.maxtiles = maxtiles;
    if (maxtiles <= 0) return;
    M = batches[0].M;
    N = batches[0].N;
    out.M = M;
    out.N = N;
    int64_t maxk = 0;
    for (auto const& b : batches) {
        int64_t k = 0;
        for (int i = b.seg_begin; i < b.seg_end; ++i) k += segs[i].K;
        maxk = std::max(maxk, k);
    }
    out.bk = (ch.bk == 32 || ch.bk == -16 || ch.bk == kBf16x6 || is_dma(ch.bk)) ? ch.bk : 16;
    if (ch.bk == kDefaultKernel) out.bk = dma_shapes_ok<T>(opA, opB, batches, segs) ? kDmaF64 : 16;
    // split K when the launch would leave most of the chip idle (only plans
    // whose batches share one output size): fewer tiles than the workgroups the
    // chip runs at once (three per CU for the LDS-copy kernel)
    const int64_t few = is_dma(out.bk) ? g_tuning.split_k_tiles_dma : g_tuning.split_k_tiles;
    int nsplit = 1;
    if (ch.nsplit > 0) {
        if (uniform) nsplit = ch.nsplit;
    } else if (uniform && g_tuning.split_k && tiles < few) {
        while (nsplit < 8 && tiles * nsplit * 2 <= few && maxk / (nsplit * 2) >= 256) nsplit *= 2;
    }
    out.nsplit = nsplit;
    out.m3 = ch.m3 >= 0 ? ch.m3 != 0 : (sizeof(T) == sizeof(double) ? g_tuning.use_3m_f64 : g_tuning.use_3m_f32);
    if (nsplit == 1) {
        out.b = batches;
        out.sg = segs;
        if (!uniform) {
            // largest outputs first: workgroups are dispatched in batch order, so
            // the tail of the launch is made of the small batches' tiles
            std::stable_sort(out.b.begin(), out.b.end(), [](BatchT<T> const& x, BatchT<T> const& y) {
                return (int64_t)x.M * x.N > (int64_t)y.M * y.N;
            });
        }
        settle_kernel(out);
        return;
    }
    // split-K: partial products into a workspace, then a reduction into C
    const int64_t slice = (int64_t)M * N;
    out.slice = slice;
    CT* W = (CT*)work((size_t)slice * nsplit * batches.size() * sizeof(CT));
    out.W = W;
    for (size_t b = 0; b < batches.size(); ++b) {
        std::vector<std::vector<SegT<T>>> chunks;
        split_segments(segs, batches[b].seg_begin, batches[b].seg_end, nsplit, opA, opB, chunks);
        for (int q = 0; q < nsplit; ++q) {
            BatchT<T> x;
            x.C = W + ((int64_t)b * nsplit + q) * slice;
            x.ldc = M;
            x.M = M;
            x.N = N;
            x.live = batches[b].live;
            x.upper = batches[b].upper;
            x.seg_begin = (int32_t)out.sg.size();
            for (auto const& sg : chunks[q]) out.sg.push_back(sg);
            x.seg_end = (int32_t)out.sg.size();
            out.b.push_back(x);
        }
        out.cptr.push_back(batches[b].C);
        out.ldcs.push_back(batches[b].ldc);
        out.lives.push_back(batches[b].live);
    }
    settle_kernel(out);
}

// device copies of a built plan's lists
template <typename T>
struct BuiltDev {
    const BatchT<T>* b;
    const SegT<T>* sg;
    typename Cplx<T>::type* const* cptr;
    const int64_t* ldcs;
    const int* const* lives;
};

template <typename T>
void parts_of(Built<T> const& bp, const void* parts[5], size_t sizes[5]) {
    parts[0] = bp.b.data();     sizes[0] = bp.b.size() * sizeof(BatchT<T>);
    parts[1] = bp.sg.data();    sizes[1] = bp.sg.size() * sizeof(SegT<T>);
    parts[2] = bp.cptr.data();  sizes[2] = bp.cptr.size() * sizeof(void*);
    parts[3] = bp.ldcs.data();  sizes[3] = bp.ldcs.size() * sizeof(int64_t);
    parts[4] = bp.lives.data(); sizes[4] = bp.lives.size() * sizeof(void*);
}

template <typename T>
BuiltDev<T> dev_of(char* d, const size_t offs[5]) {
    typedef typename Cplx<T>::type CT;
    return BuiltDev<T>{(const BatchT<T>*)(d + offs[0]), (const SegT<T>*)(d + offs[1]), (CT* const*)(d + offs[2]),
                       (const int64_t*)(d + offs[3]), (const int* const*)(d + offs[4])};
}

template <typename T>
void launch_built(Built<T> const& bp, BuiltDev<T> const& dv, hipStream_t s) {
    typedef typename Cplx<T>::type CT;
    dim3 grid((unsigned)bp.maxtiles, (unsigned)bp.b.size());
    launch_any<T>(bp.opA, bp.opB, bp.bk, bp.nsplit == 1 ? bp.acc : false, bp.m3, grid, dv.b, dv.sg, s, bp.xmap);
    NC_HIP(hipGetLastError());
    if (bp.nsplit == 1) return;
    unsigned g = (unsigned)std::min<int64_t>(2048, (bp.slice + 255) / 256);
    hipLaunchKernelGGL(splitk_reduce_kernel<CT>, dim3(g, (unsigned)bp.cptr.size()), dim3(256), 0, s, bp.W, bp.slice,
                       bp.nsplit, dv.cptr, dv.ldcs, dv.lives, bp.M, bp.N, bp.acc ? 1 : 0);
    NC_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Autotuner.  The first time a plan with a given shape key runs, every
// candidate launch choice (K step 16 or 32, K split 1-8 where the batches share
// one size, 3- or 4-multiplication complex product) is timed with HIP events on
// the plan's own operands (after a warm-up, two interleaved rounds, the minimum
// per candidate); the fastest is kept for the key when it beats the default
// choice by more than 3 %.  Accumulating plans and plans of less than 2e8 flops
// are never tuned (the first are not repeatable, the second launch-bound).
// ---------------------------------------------------------------------------
struct TuneEntry {
    GemmChoice choice;
    int kernel = 16;  // the kernel the choice resolves to for this plan
    double default_us = 0, best_us = 0;
    int candidates = 0;
};
std::mutex g_tune_mu;
std::map<std::vector<int>, TuneEntry> g_tune;
double g_tune_seconds = 0;  // host time spent timing candidates (under g_tune_mu)

std::vector<int> tune_id(int prec, int opA, int opB, TuneKey const& k) {
    return {prec, opA, opB, k.tag, k.bl, k.br, k.d, k.cl, k.cr, k.count};
}

template <typename T>
GemmChoice choose(int opA, int opB, bool accumulate, int M, int N, std::vector<BatchT<T>> const& batches,
                  std::vector<SegT<T>> const& segs, TuneKey const& key, hipStream_t s) {
    if (g_forced.bk != 0) return g_forced;
    if (!key.tag || accumulate || !gemm_autotune() || batches.empty()) return GemmChoice();
    const std::vector<int> id = tune_id(sizeof(T) == sizeof(double) ? 64 : 32, opA, opB, key);
    {
        std::lock_guard<std::mutex> g(g_tune_mu);
        auto it = g_tune.find(id);
        if (it != g_tune.end()) return it->second.choice;
    }
    const auto t_begin = std::chrono::steady_clock::now();
    // candidates; every batch runs during tuning (live flags ignored)
    std::vector<BatchT<T>> bs = batches;
    bool uniform = true;
    int64_t maxk = 0;
    for (auto& b : bs) {
        b.live = nullptr;
        if (b.M == 0) b.M = M;
        if (b.N == 0) b.N = N;
        if (b.M != bs[0].M || b.N != bs[0].N) uniform = false;
        int64_t k = 0;
        for (int i = b.seg_begin; i < b.seg_end; ++i) k += segs[i].K;
        maxk = std::max(maxk, k);
    }
    // small plans (a few microseconds of matrix-core work) keep the default:
    // their time is launch overhead, which no choice changes
    double flops = 0;
    for (auto const& b : bs) {
        int64_t k = 0;
        for (int i = b.seg_begin; i < b.seg_end; ++i) k += segs[i].K;
        flops += 8.0 * b.M * b.N * k;
    }
    if (flops < 2e8) return GemmChoice();
    std::vector<GemmChoice> cands;
    cands.push_back(GemmChoice());
    for (int bk : {16, 32, -16})
        for (int m3 : {0, 1})
            for (int ns : {1, 2, 4, 8}) {
                if (ns > 1 && (!uniform || maxk / ns < 64)) continue;
                cands.push_back(GemmChoice{bk, ns, m3});
            }
    // complex128 untransposed operands: the LDS-copy pipeline (dropped below
    // where the plan's shapes do not allow it)
    if (sizeof(T) == sizeof(double) && opA == OpN && (opB == OpN || opB == OpC))
        for (int bk : {kDmaF64, kDmaF64x4})
            for (int m3 : {0, 1})
                for (int ns : {1, 2, 4, 8}) {
                    if (ns > 1 && (!uniform || maxk / ns < 64)) continue;
                    cands.push_back(GemmChoice{bk, ns, m3});
                }
    // complex64 untransposed operands: the bf16 split kernel (always 3M)
    if (sizeof(T) == sizeof(float) && opA == OpN && opB == OpN && g_tuning.bf16x6)
        for (int ns : {1, 2, 4, 8}) {
            if (ns > 1 && (!uniform || maxk / ns < 64)) continue;
            cands.push_back(GemmChoice{kBf16x6, ns, 1});
        }
    // every candidate built once (own descriptors; the split-K workspace is
    // shared, candidates run one after another)
    struct Cand {
        Built<T> bp;
        DevBuf<char> desc;
        BuiltDev<T> dv{};
        bool skip = false;
    };
    std::vector<Cand> built(cands.size());
    size_t work_need = 0;
    for (size_t c = 0; c < cands.size(); ++c) {
        std::vector<BatchT<T>> b2 = bs;
        build_plan<T>(opA, opB, false, M, N, b2, segs, built[c].bp, [&](size_t bytes) {
            work_need = std::max(work_need, bytes);
            return nullptr;  // fixed up below
        }, cands[c]);
    }
    DevBuf<char> work(std::max<size_t>(work_need, 1));
    for (size_t c = 0; c < cands.size(); ++c) {
        std::vector<BatchT<T>> b2 = bs;
        Cand& cd = built[c];
        build_plan<T>(opA, opB, false, M, N, b2, segs, cd.bp, [&](size_t) { return (void*)work.ptr; }, cands[c]);
        if (cd.bp.empty()) return GemmChoice();
        // a kernel the plan's shapes do not allow: not a candidate
        cd.skip = is_dma(cands[c].bk) && cd.bp.bk != cands[c].bk;
        // an explicit choice that builds the default's launch: timed as the default
        if (c > 0 && cd.bp.bk == built[0].bp.bk && cd.bp.nsplit == built[0].bp.nsplit &&
            cd.bp.m3 == built[0].bp.m3)
            cd.skip = true;
        if (cd.skip) continue;
        const void* parts[5];
        size_t sizes[5], offs[5], off = 0;
        parts_of(cd.bp, parts, sizes);
        for (int i = 0; i < 5; ++i) {
            offs[i] = off;
            off = (off + sizes[i] + 255) & ~size_t(255);
        }
        std::vector<char> host(off);
        for (int i = 0; i < 5; ++i)
            if (sizes[i]) std::memcpy(host.data() + offs[i], parts[i], sizes[i]);
        cd.desc.resize(off);
        h2d(cd.desc.ptr, host.data(), off);
        cd.dv = dev_of<T>(cd.desc.ptr, offs);
    }
    hipEvent_t e0, e1;
    NC_HIP(hipEventCreate(&e0));
    NC_HIP(hipEventCreate(&e1));
    // clocks up first, then two interleaved rounds; the minimum per candidate
    for (int w = 0; w < 3; ++w) launch_built(built[0].bp, built[0].dv, s);
    std::vector<double> us(cands.size(), 1e30);
    const int reps = 3;
    for (int round = 0; round < 2; ++round)
        for (size_t c = 0; c < cands.size(); ++c) {
            if (built[c].skip) continue;
            launch_built(built[c].bp, built[c].dv, s);
            NC_HIP(hipEventRecord(e0, s));
            for (int r = 0; r < reps; ++r) launch_built(built[c].bp, built[c].dv, s);
            NC_HIP(hipEventRecord(e1, s));
            NC_HIP(hipEventSynchronize(e1));
            float ms = 0;
            NC_HIP(hipEventElapsedTime(&ms, e0, e1));
            us[c] = std::min(us[c], 1e3 * ms / reps);
        }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    size_t best = 0;
    int ncand = 0;
    for (size_t c = 0; c < cands.size(); ++c) {
        if (built[c].skip) continue;
        ++ncand;
        if (us[c] < us[best]) best = c;
    }
    TuneEntry e;
    e.default_us = us[0];
    e.best_us = us[best];
    e.candidates = ncand;
    const size_t pick = us[best] < 0.97 * us[0] ? best : 0;
    e.choice = pick ? cands[pick] : GemmChoice();
    e.kernel = built[pick].bp.bk;
    std::lock_guard<std::mutex> g(g_tune_mu);
    g_tune[id] = e;
    g_tune_seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t_begin).count();
    return e.choice;
}

template <typename T>
void run_plan(int opA, int opB, bool accumulate, int M, int N, std::vector<BatchT<T>>& batches,
              std::vector<SegT<T>> const& segs, hipStream_t s, TuneKey const& key = TuneKey()) {
    static thread_local DevBuf<char> work;
    const GemmChoice ch = choose<T>(opA, opB, accumulate, M, N, batches, segs, key, s);
    Built<T> bp;
    build_plan<T>(opA, opB, accumulate, M, N, batches, segs, bp, [&](size_t bytes) {
        if (work.n < bytes) work.resize(bytes);
        return (void*)work.ptr;
    }, ch);
    if (bp.empty()) return;
    const void* parts[5];
    size_t sizes[5], offs[5];
    parts_of(bp, parts, sizes);
    int slot;
    char* d = arena().stage(parts, sizes, 5, offs, s, &slot);
    launch_built(bp, dev_of<T>(d, offs), s);
    arena().release(slot, s);
}

// a plan uploaded once into device memory of its own
template <typename T>
struct PreparedGemmT final : PreparedLaunch {
    Built<T> bp;
    DevBuf<char> desc, work;
    BuiltDev<T> dv{};
    void launch(hipStream_t s) const override {
        if (!bp.empty()) launch_built(bp, dv, s);
    }
};

template <typename T>
std::unique_ptr<PreparedLaunch> prepare_plan(int opA, int opB, bool accumulate, int M, int N,
                                             std::vector<BatchT<T>>& batches, std::vector<SegT<T>> const& segs,
                                             TuneKey const& key) {
    std::unique_ptr<PreparedGemmT<T>> p(new PreparedGemmT<T>());
    PreparedGemmT<T>* q = p.get();
    const GemmChoice ch = choose<T>(opA, opB, accumulate, M, N, batches, segs, key, Context::get().stream);
    build_plan<T>(opA, opB, accumulate, M, N, batches, segs, q->bp, [q](size_t bytes) {
        q->work.resize(bytes);
        return (void*)q->work.ptr;
    }, ch);
    if (q->bp.empty()) return p;
    const void* parts[5];
    size_t sizes[5], offs[5], off = 0;
    parts_of(q->bp, parts, sizes);
    for (int i = 0; i < 5; ++i) {
        offs[i] = off;
        off = (off + sizes[i] + 255) & ~size_t(255);
    }
    std::vector<char> host(off);
    for (int i = 0; i < 5; ++i)
        if (sizes[i]) std::memcpy(host.data() + offs[i], parts[i], sizes[i]);
    q->desc.resize(off);
    h2d(q->desc.ptr, host.data(), off);
    q->dv = dev_of<T>(q->desc.ptr, offs);
    return p;
}

}  // namespace

GemmTuning& gemm_tuning() { return g_tuning; }

bool& gemm_autotune() {
    static bool on = [] {
        const char* e = std::getenv("NC_AUTOTUNE");
        return !(e && e[0] == '0');
    }();
    return on;
}

void gemm_force_choice(int bk, int nsplit, int m3) { g_forced = GemmChoice{bk, nsplit, m3}; }

// the kernel of a launch choice as the tuning table prints it
std::string kernel_name(int bk) {
    switch (bk) {
        case -16: return "bk 16 2-stage";
        case kBf16x6: return "bk 32 bf16x6";
        case kDmaF64: return "bk 8 lds-copy 3-stage";
        case kDmaF64x4: return "bk 8 lds-copy 4-stage";
        default: return "bk " + std::to_string(bk);
    }
}

double gemm_tune_seconds() {
    std::lock_guard<std::mutex> g(g_tune_mu);
    return g_tune_seconds;
}

std::string gemm_tune_table() {
    std::lock_guard<std::mutex> g(g_tune_mu);
    std::string out;
    char line[256];
    for (auto const& kv : g_tune) {
        auto const& k = kv.first;
        auto const& e = kv.second;
        snprintf(line, sizeof line,
                 "fp%d op%d%d step %c chi_l %d chi_r %d d %d w %d/%d chains %d: %s split %d %s (%.1f us, default %.1f us, %d candidates)\n",
                 k[0], k[1], k[2], (char)k[3], k[4], k[5], k[6], k[7], k[8], k[9],
                 kernel_name(e.kernel).c_str(),
                 e.choice.nsplit,
                 e.choice.m3 < 0 ? "3M default" : (e.choice.m3 ? "3M" : "4M"), e.best_us, e.default_us,
                 e.candidates);
        out += line;
    }
    return out;
}

void* stage_descriptors(const void* host, size_t bytes, hipStream_t s, int* slot) {
    const void* parts[1] = {host};
    size_t sizes[1] = {bytes};
    size_t offs[1];
    return arena().stage(parts, sizes, 1, offs, s, slot) + offs[0];
}
void release_descriptors(int slot, hipStream_t s) { arena().release(slot, s); }

void zgemm_run(GemmPlan const& plan, hipStream_t s) {
    static_assert(sizeof(BatchT<double>) == sizeof(GemmBatch), "descriptor layout");
    static_assert(sizeof(SegT<double>) == sizeof(GemmSeg), "descriptor layout");
    static_assert(offsetof(BatchT<double>, upper) == offsetof(GemmBatch, upper), "descriptor layout");
    std::vector<BatchT<double>> b(plan.batches.size());
    std::vector<SegT<double>> sg(plan.segs.size());
    std::memcpy(b.data(), plan.batches.data(), b.size() * sizeof(GemmBatch));
    std::memcpy(sg.data(), plan.segs.data(), sg.size() * sizeof(GemmSeg));
    run_plan<double>(plan.opA, plan.opB, plan.accumulate, plan.M, plan.N, b, sg, s, plan.key);
}

namespace {
// the plan built with the default launch choice and no K split
Built<double> fixed_plan(GemmPlan const& plan) {
    std::vector<BatchT<double>> b(plan.batches.size());
    std::vector<SegT<double>> sg(plan.segs.size());
    std::memcpy(b.data(), plan.batches.data(), b.size() * sizeof(GemmBatch));
    std::memcpy(sg.data(), plan.segs.data(), sg.size() * sizeof(GemmSeg));
    Built<double> bp;
    build_plan<double>(plan.opA, plan.opB, plan.accumulate, plan.M, plan.N, b, sg, bp,
                       [](size_t) -> void* { throw Error("GEMM: a fixed launch takes no K split", -72); },
                       GemmChoice{16, 1, -1});
    return bp;
}
void fixed_layout(Built<double> const& bp, const void* parts[5], size_t sizes[5], size_t offs[5], size_t& total) {
    parts_of(bp, parts, sizes);
    total = 0;
    for (int i = 0; i < 5; ++i) {
        offs[i] = total;
        total = (total + sizes[i] + 255) & ~size_t(255);
    }
}
}  // namespace

size_t zgemm_desc_bytes(GemmPlan const& plan) {
    Built<double> bp = fixed_plan(plan);
    const void* parts[5];
    size_t sizes[5], offs[5], total;
    fixed_layout(bp, parts, sizes, offs, total);
    return total;
}

void zgemm_run_at(GemmPlan const& plan, void* desc, size_t cap, hipStream_t s) {
    Built<double> bp = fixed_plan(plan);
    if (bp.empty()) return;
    const void* parts[5];
    size_t sizes[5], offs[5], total;
    fixed_layout(bp, parts, sizes, offs, total);
    if (total > cap) throw Error("GEMM: descriptor lists larger than the space given", -73);
    for (int i = 0; i < 5; ++i)
        if (sizes[i]) store_bytes(static_cast<char*>(desc) + offs[i], parts[i], sizes[i], s);
    launch_built(bp, dev_of<double>(static_cast<char*>(desc), offs), s);
}

static void to_float_lists(CGemmPlan const& plan, std::vector<BatchT<float>>& b, std::vector<SegT<float>>& sg) {
    for (auto const& x : plan.batches)
        b.push_back(BatchT<float>{x.C, x.ldc, x.seg_begin, x.seg_end, x.M, x.N, x.live, x.upper, 0});
    for (auto const& x : plan.segs) sg.push_back(SegT<float>{x.A, x.B, x.lda, x.ldb, x.K, 0});
}

void cgemm_run(CGemmPlan const& plan, hipStream_t s) {
    std::vector<BatchT<float>> b;
    std::vector<SegT<float>> sg;
    to_float_lists(plan, b, sg);
    run_plan<float>(plan.opA, plan.opB, plan.accumulate, plan.M, plan.N, b, sg, s, plan.key);
}

std::unique_ptr<PreparedLaunch> prepare_gemm(GemmPlan const& plan) {
    std::vector<BatchT<double>> b(plan.batches.size());
    std::vector<SegT<double>> sg(plan.segs.size());
    std::memcpy(b.data(), plan.batches.data(), b.size() * sizeof(GemmBatch));
    std::memcpy(sg.data(), plan.segs.data(), sg.size() * sizeof(GemmSeg));
    return prepare_plan<double>(plan.opA, plan.opB, plan.accumulate, plan.M, plan.N, b, sg, plan.key);
}

std::unique_ptr<PreparedLaunch> prepare_gemm(CGemmPlan const& plan) {
    std::vector<BatchT<float>> b;
    std::vector<SegT<float>> sg;
    to_float_lists(plan, b, sg);
    return prepare_plan<float>(plan.opA, plan.opB, plan.accumulate, plan.M, plan.N, b, sg, plan.key);
}

}  // namespace nc
