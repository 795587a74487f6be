// H_eff in complex64 on the matrix cores with operands split once.
//
// The complex64 matvec y = H_eff x (steps A, mix, C of the grouped
// contraction) spends its time in products whose operands are split into
// 16-bit parts (Gauss's 3M form for the complex product).  Splitting inside
// the GEMM repeats the work for every output tile that reads an operand and
// costs more vector instructions than the matrix-core instructions it feeds.
// Here the operands are split once, outside the GEMM:
//   * the environments L_f and R_t when the operator is built (once per site
//     solve: they are fixed across its hundreds of matvecs),
//   * x once per matvec (split_planes),
//   * step A's output by the mix kernel, which writes Q_t's planes directly
//     (mix_split_kernel),
// and the GEMM (cgemm_planes_kernel) only moves planes into LDS and runs MFMAs.
//
// Two part forms:
//   * fp16 (default): each matrix is multiplied by a power of two that puts its
//     largest |re|, |im|, |re + im| at most 2^14 (exact; the GEMM divides it
//     out of its result), and each value is split into two fp16 parts
//     (hi + lo, 22 significant bits).  Three partial products per real product
//     (lo.hi, hi.lo, hi.hi) on v_mfma_f32_16x16x32_f16.  The scale of the
//     environments comes from a device reduction over their values; that of
//     x from the caller's bound on ||x|| (the Krylov solver's basis vectors)
//     or, without one, from a reduction per matvec; the scale of Q_t (written
//     by the mix kernel before the values exist) from a bound:
//     |Q_t| <= max|x| * max_b sum_{blocks n -> t} sum_a |W_n(a, b)|
//     * (largest column 1-norm of L_{from n}).  Each XCD takes a block of
//     output tiles, so that the workgroups sharing its L2 share row panels.
//   * bf16: three bf16 parts (hi + mid + lo), no scale, six partial products
//     (lo.hi, mid.mid, hi.lo, mid.hi, hi.mid, hi.hi) on
//     v_mfma_f32_16x16x32_bf16.
// Both keep FP32 accuracy in the products; the fp16 form runs half the
// matrix-core instructions and moves two thirds of the bytes.
//
// Plane layout.  A complex matrix Z with rows r < R and inner (K) index k < K
// is held as 3 * parts planes, p = parts * c + j for the component c (re, im,
// re + im) and the part j.  Each plane is an array of 16-byte chunks, chunk
// (kb, r) = the 8 values of k = 8 kb .. 8 kb + 7 in row r, at chunk index
// kb * Rp + r; Rp = R rounded up to 128, Kp = K rounded up to 32, Kb = Kp / 8
// chunk columns, plane stride Kb * Rp chunks.  Chunks of k >= K are zero; rows
// R .. Rp - 1 are never written and only reach output rows or columns that are
// not stored.  One MFMA operand fragment (lane l: row l % 16, k = 8 (l / 16)
// .. + 7) is then a single 16-byte chunk, 16 lanes read 256 contiguous bytes,
// and a tile's stage is a set of contiguous 1 KB runs.
//
// GEMM: 128 x 128 output tile per 512-thread workgroup (8 waves in 2 x 4, each
// wave 4 x 2 MFMA blocks of 16 x 16 for each of the three real products), K
// in stages of 32, one component (all parts of A and of B: 32 KB fp16, 48 KB
// bf16) per pipeline step.  The three components cycle through three LDS
// buffers, so buffer q always holds component q; each step's planes are copied
// global -> LDS by global_load_lds_dwordx4 two steps ahead of their use, and a
// step waits only for its own copies (vmcnt) before the workgroup barrier.  K
// can be split over several workgroups per tile (slices summed by the
// consumer).
#include "nc_bf16.hpp"
#include "nc_internal.hpp"
#include "nc_lds_dma.hpp"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>

namespace nc {

void PlaneSet::shape(int R_, int K_, int count_, int form_) {
    if (R_ <= 0 || K_ <= 0 || count_ <= 0) throw Error("plane set: bad shape", -2);
    if (form_ != kPlaneBF16 && form_ != kPlaneF16) throw Error("plane set: unknown part form", -2);
    R = R_;
    K = K_;
    count = count_;
    form = form_;
    np = form == kPlaneF16 ? 6 : 9;
    Rp = (R + kPlaneTile - 1) / kPlaneTile * kPlaneTile;
    Kb = (K + 31) / 32 * 4;
    ps = (int64_t)Kb * Rp;
    buf.resize((size_t)(np * ps * count));
    scale.resize((size_t)count);
    if (form == kPlaneBF16) {
        std::vector<float> one((size_t)count, 1.0f);
        h2d(scale.ptr, one.data(), one.size() * sizeof(float));
    }
}

namespace {
constexpr int kRedBlocks = 512;  // blocks of a scale reduction
}

namespace {

constexpr int kPT = kPlaneTile;

// ---- fp16 scales

// 2^k with k = 14 - e for m = f 2^e (f in [0.5, 1)), so that m 2^k <= 2^14; 1
// for m = 0, NaN for a non-finite m.  k is kept within [-60, 60], so products
// and quotients of two scales stay normal FP32 numbers.
__device__ __forceinline__ float scale_for(float m) {
    if (!(m < INFINITY)) return NAN;
    if (!(m > 0.0f)) return 1.0f;
    int e;
    (void)frexpf(m, &e);
    const int k = std::min(60, std::max(-60, 14 - e));
    return ldexpf(1.0f, k);
}

// The largest |re|, |im|, |re + im| (giving the planes' scale) and the largest
// modulus of n complex128 values: every block's maxima, rounded up to FP32, go
// into slot[0] and slot[1] by atomicMax on their bit patterns (FP32 bit
// patterns of non-negative values order as unsigned integers).  The slots must
// be zero before; maxima do not depend on the order of the updates.
__global__ __launch_bounds__(256) void plane_amax_kernel(const double2* __restrict__ src, int64_t n,
                                                         unsigned* __restrict__ slot) {
    double mc = 0.0, mm = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double2 v = src[i];
        mc = fmax(mc, fmax(fabs(v.x + v.y), fmax(fabs(v.x), fabs(v.y))));
        mm = fmax(mm, v.x * v.x + v.y * v.y);
    }
    for (int off = 32; off > 0; off >>= 1) {
        mc = fmax(mc, __shfl_xor(mc, off, 64));
        mm = fmax(mm, __shfl_xor(mm, off, 64));
    }
    __shared__ double sh[2][4];
    if ((threadIdx.x & 63) == 0) {
        sh[0][threadIdx.x >> 6] = mc;
        sh[1][threadIdx.x >> 6] = mm;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const double a = fmax(fmax(sh[0][0], sh[0][1]), fmax(sh[0][2], sh[0][3]));
        const double b = sqrt(fmax(fmax(sh[1][0], sh[1][1]), fmax(sh[1][2], sh[1][3])));
        float af = (float)a, bf = (float)b;
        if ((double)af < a) af = nextafterf(af, INFINITY);
        if ((double)bf < b) bf = nextafterf(bf, INFINITY);
        atomicMax(slot, __float_as_uint(af));
        atomicMax(slot + 1, __float_as_uint(bf));
    }
}

// *out = max(*out, largest column 1-norm (sum of moduli) of the R x C
// column-major complex128 matrix src), one wave per column; FP32 bit patterns
// of non-negative values order as unsigned integers.
__global__ __launch_bounds__(256) void colnorm1_max_kernel(const double2* __restrict__ src, int R, int C,
                                                           unsigned* __restrict__ out) {
    const int col = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (col >= C) return;
    const double2* c = src + (int64_t)col * R;
    double acc = 0.0;
    for (int r = lane; r < R; r += 64) acc += hypot(c[r].x, c[r].y);
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if (lane == 0) {
        const float f = (float)(acc * (1.0 + 1e-6));  // a bound: rounded up
        atomicMax(out, __float_as_uint(f));
    }
}

// Scale of Q_t's fp16 planes: S / sR[t], with S = 2^E the largest power of two
// for which S / sR[u] * sqrt(2) * bound(|Q_u|) <= 2^14 for every destination u,
// bound(|Q_u|) = amod_x * max_b sum_{blocks j -> u} sum_a |W_j(a, b)| n1[src_j]
// (|P_f(r, l)| <= amod_x * column 1-norm of L_f).  q_scale_exponent gives E
// (INT_MAX when every bound is 0, i.e. Q = 0; INT_MIN for a non-finite bound);
// every thread forms the same value.  q_plane_scale: 1 for Q = 0, NaN when the
// exponent leaves the FP32 range, so that the result fails loudly.
__device__ int q_scale_exponent(int ndst, int d, const int* __restrict__ dst_ptr, const int* __restrict__ blk_src,
                                const float2* __restrict__ W, PlaneQScale const& qs) {
    const float ax = *qs.amod_x;
    int E = INT_MAX;
    for (int u = 0; u < ndst; ++u) {
        float C = 0.0f;
        for (int b = 0; b < d; ++b) {
            float cb = 0.0f;
            for (int j = dst_ptr[u]; j < dst_ptr[u + 1]; ++j) {
                const float n1 = qs.n1[blk_src[j]];
                for (int a = 0; a < d; ++a) {
                    const float2 w = W[(int64_t)j * d * d + a + b * d];
                    cb += (fabsf(w.x) + fabsf(w.y)) * n1;
                }
            }
            C = fmaxf(C, cb);
        }
        const float bound = 1.4142136f * 1.001f * ax * C;
        if (bound > 0.0f) {
            if (!(bound < INFINITY)) return INT_MIN;
            int eb, er;
            (void)frexpf(bound, &eb);     // bound <= 2^eb
            (void)frexpf(qs.sR[u], &er);  // sR[u] = 2^(er - 1)
            E = std::min(E, er - 1 + 14 - eb);
        }
    }
    return E;
}
__device__ __forceinline__ float q_plane_scale(int E, float sRt) {
    if (E == INT_MAX) return 1.0f;
    if (E == INT_MIN) return NAN;
    int et;
    (void)frexpf(sRt, &et);
    const int k = E - (et - 1);
    if (k < -126 || k > 127) return NAN;
    return ldexpf(1.0f, k);
}

// ---- splitting

// planes of the 8 values re[], im[] at chunk d (plane stride ps): nine bf16
// planes, or six fp16 planes of values already scaled
template <bool F16>
__device__ __forceinline__ void store_planes(u32x4* __restrict__ d, int64_t ps, const float (&re)[8],
                                             const float (&im)[8]) {
    float sm[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) sm[e] = re[e] + im[e];
    if (F16) {
        u32x4 h, l;
        split2(re, h, l);
        d[0] = h; d[ps] = l;
        split2(im, h, l);
        d[2 * ps] = h; d[3 * ps] = l;
        split2(sm, h, l);
        d[4 * ps] = h; d[5 * ps] = l;
    } else {
        u32x4 h, m, l;
        split3(re, h, m, l);
        d[0] = h; d[ps] = m; d[2 * ps] = l;
        split3(im, h, m, l);
        d[3 * ps] = h; d[4 * ps] = m; d[5 * ps] = l;
        split3(sm, h, m, l);
        d[6 * ps] = h; d[7 * ps] = m; d[8 * ps] = l;
    }
}

// Planes of matrix z: element (r, k) at src[z * sz + r * sr + k * sk], times
// (F16) the scale of the largest value in mslot[0] (plane_amax_kernel), which
// thread 0 of block (0, 0) records in sc[0 .. nz - 1].  With KCONTIG (sk = 1) a wave covers 16 rows x 4 chunks (4 x 128
// contiguous source bytes per row, 256 contiguous destination bytes per chunk
// column), otherwise consecutive lanes take consecutive rows.
template <bool KCONTIG, bool F16>
__global__ __launch_bounds__(256) void split_planes_kernel(const double2* __restrict__ src, int64_t sr, int64_t sk,
                                                           int64_t sz, u32x4* __restrict__ dst, int64_t ps,
                                                           int64_t dz, int R, int K, int Rp, int Kb,
                                                           const unsigned* __restrict__ mslot,
                                                           float* __restrict__ sc) {
    const int z = blockIdx.y;
    const float fs = F16 ? scale_for(__uint_as_float(mslot[0])) : 1.0f;
    if (F16 && blockIdx.x == 0 && z == 0 && threadIdx.x == 0)
        for (int u = 0; u < (int)gridDim.y; ++u) sc[u] = fs;
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
    const double s = fs;
    const double2* p = src + z * sz + (int64_t)r * sr;
    float re[8], im[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const int k = 8 * kb + e;
        re[e] = 0.f;
        im[e] = 0.f;
        if (k < K) {
            const double2 v = p[(int64_t)k * sk];
            re[e] = (float)(v.x * s);
            im[e] = (float)(v.y * s);
        }
    }
    store_planes<F16>(dst + z * dz + (int64_t)kb * Rp + r, ps, re, im);
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

// the planes of the pairs (r0, r1) (re) and (i0, i1) (im) at dword q (dword
// stride of a plane s4): nine bf16 planes, or six fp16 planes of the values
// times sq
template <bool F16>
__device__ __forceinline__ void store_pair_planes(unsigned* __restrict__ q, int64_t s4, float r0, float r1,
                                                  float i0, float i1, float sq) {
    if (F16) {
        r0 *= sq; r1 *= sq; i0 *= sq; i1 *= sq;
        unsigned h, l;
        split2_pair(r0, r1, h, l);
        q[0] = h; q[s4] = l;
        split2_pair(i0, i1, h, l);
        q[2 * s4] = h; q[3 * s4] = l;
        split2_pair(r0 + i0, r1 + i1, h, l);
        q[4 * s4] = h; q[5 * s4] = l;
    } else {
        unsigned h, m, lo;
        split_pair(r0, r1, h, m, lo);
        q[0] = h; q[s4] = m; q[2 * s4] = lo;
        split_pair(i0, i1, h, m, lo);
        q[3 * s4] = h; q[4 * s4] = m; q[5 * s4] = lo;
        split_pair(r0 + i0, r1 + i1, h, m, lo);
        q[6 * s4] = h; q[7 * s4] = m; q[8 * s4] = lo;
    }
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
// D = 0 takes any d, four outputs b per pass over the blocks.  F16: the values
// are multiplied by Q_t's scale (q_plane_scale), which block (0, t) records.
template <int D, bool EVEN, bool F16>
__global__ __launch_bounds__(256) void mix_split_kernel(const float2* __restrict__ P, int64_t pslice, int nslice,
                                                        const int* __restrict__ dst_ptr,
                                                        const int* __restrict__ blk_src,
                                                        const float2* __restrict__ W, int d_rt, int br, int bl,
                                                        int64_t rs, unsigned* __restrict__ Q, int64_t ps, int64_t qz,
                                                        int Np, int Kb, int ndst, PlaneQScale qs,
                                                        float* __restrict__ qscale) {
    constexpr int NB = D > 0 ? D : 4;
    const int d = D > 0 ? D : d_rt;
    const int t = blockIdx.y;
    const float sq = F16 ? q_plane_scale(q_scale_exponent(ndst, d, dst_ptr, blk_src, W, qs), qs.sR[t]) : 1.0f;
    if (F16 && blockIdx.x == 0 && threadIdx.x == 0) qscale[t] = sq;
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
        // the planes, dword pi of chunk (kb, n); dword stride of a plane 4 ps
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            if (b0 + b < d) {
                unsigned* q = Q + 4 * (t * qz + (int64_t)kb * Np + l + (int64_t)bl * (b0 + b)) + pi;
                store_pair_planes<F16>(q, 4 * ps, r0[b], r1[b], i0[b], i1[b], sq);
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
template <int D, int MAXF, int MAXT, bool F16>
__global__ __launch_bounds__(256) void mix_split_all_kernel(const float2* __restrict__ P, int64_t pslice, int nslice,
                                                            int nsrc, int ndst, const int* __restrict__ dst_ptr,
                                                            const int* __restrict__ blk_src,
                                                            const float2* __restrict__ W, int br, int bl, int64_t rs,
                                                            unsigned* __restrict__ Q, int64_t ps, int64_t qz, int Np,
                                                            int Kb, PlaneQScale qs, float* __restrict__ qscale) {
    const int nlg = (bl + 15) >> 4;
    const int64_t w = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    const int kb = (int)(w / nlg);
    const int l = (int)(w % nlg) * 16 + (lane >> 2);
    const int pi = lane & 3;
    const int E = F16 ? q_scale_exponent(ndst, D, dst_ptr, blk_src, W, qs) : 0;
    if (F16 && blockIdx.x == 0 && threadIdx.x == 0)
        for (int t = 0; t < ndst; ++t) qscale[t] = q_plane_scale(E, qs.sR[t]);
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
        const float sq = F16 ? q_plane_scale(E, qs.sR[t]) : 1.0f;
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
            store_pair_planes<F16>(q, s4, r0[b], r1[b], i0[b], i1[b], sq);
        }
    }
}

// out[e] = sum_c Y[c slice + e] (in FP64), e < n; reset[0 .. 1] (optional) set
// to zero for the next matvec's maxima
__global__ __launch_bounds__(256) void reduce_slices_kernel(const float2* __restrict__ Y, int64_t slice, int ns,
                                                            int64_t n, double2* __restrict__ out,
                                                            unsigned* __restrict__ reset) {
    if (reset && blockIdx.x == 0 && threadIdx.x < 2) reset[threadIdx.x] = 0u;
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
template <int NP>
struct PStage {
    u32x4 a[NP][4][kPT];  // [part][chunk column][row] of the A tile's current component
    u32x4 b[NP][4][kPT];
};

// 16 bytes per lane from global memory into LDS (lane i's chunk at g + voff_i
// lands at LDS address l + 16 i); awaited only by the vmcnt of step_barrier.
__device__ __forceinline__ void dma16(const u32x4* g, unsigned voff, const u32x4* l) { lds_dma16(g, voff, l); }

// One pipeline step's copies: component Q of the K stage whose chunk column 0
// is at pa (A, offset to the tile's first row) and pb (B).  The 8 NP runs of 64
// chunks of each operand (t = 8 part + 2 chunk column + half) are shared out
// 8 NP / NW per wave; off.a[Q][u] is the byte offset of this lane's chunk of
// run t = (8 NP / NW) wave + u from pa (likewise for B), fixed for the whole
// launch.
template <int NW, int NP>
struct IssueOffsets {
    static constexpr int kRuns = 8 * NP / NW;
    unsigned a[3][kRuns], b[3][kRuns];
};
template <int Q, int NW, int NP>
__device__ __forceinline__ void planes_issue(PStage<NP>& S, const u32x4* pa, const u32x4* pb,
                                             IssueOffsets<NW, NP> const& off, int wave) {
    constexpr int R = IssueOffsets<NW, NP>::kRuns;
#pragma unroll
    for (int u = 0; u < R; ++u) {
        const int t = wave * R + u;
        const int p = t >> 3, kb = (t >> 1) & 3, h = (t & 1) * 64;
        dma16(pa, off.a[Q][u], &S.a[p][kb][h]);
        dma16(pb, off.b[Q][u], &S.b[p][kb][h]);
    }
}

// The partial products of one component for the wave's 4 x JB blocks,
// smallest first: fp16 (lo.hi, hi.lo, hi.hi); bf16 (lo.hi, mid.mid, hi.lo,
// mid.hi, hi.mid, hi.hi).  The scheduling barrier keeps each step's MFMAs
// inside it, so the accumulators of the three components keep their registers
// across the loop.
template <int JB, bool F16>
__device__ __forceinline__ void planes_compute(PStage<F16 ? 2 : 3> const& S, f32x4 (&acc)[4][JB], int wm, int wn,
                                               int lane) {
    constexpr int NP = F16 ? 2 : 3;
    constexpr int NPROD = F16 ? 3 : 6;
    constexpr int pa[6] = {F16 ? 1 : 2, F16 ? 0 : 1, 0, 1, 0, 0};
    constexpr int pb[6] = {0, F16 ? 1 : 1, F16 ? 0 : 2, 0, 1, 0};
    __builtin_amdgcn_sched_barrier(0);
    const int r = lane & 15, kb = lane >> 4;
    u32x4 b[JB][NP];
#pragma unroll
    for (int j = 0; j < JB; ++j)
#pragma unroll
        for (int p = 0; p < NP; ++p) b[j][p] = S.b[p][kb][wn + 16 * j + r];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        u32x4 a[NP];
#pragma unroll
        for (int p = 0; p < NP; ++p) a[p] = S.a[p][kb][wm + 16 * i + r];
#pragma unroll
        for (int t = 0; t < NPROD; ++t)
#pragma unroll
            for (int j = 0; j < JB; ++j) {
                if (F16)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8, a[pa[t]]),
                                                                       __builtin_bit_cast(f16x8, b[j][pb[t]]),
                                                                       acc[i][j], 0, 0, 0);
                else
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a[pa[t]]),
                                                                        __builtin_bit_cast(bf16x8, b[j][pb[t]]),
                                                                        acc[i][j], 0, 0, 0);
            }
    }
}

// Workgroup barrier after this wave's copies up to the last N issued have
// landed in LDS and its LDS reads are done.  __syncthreads() would also wait
// for the copies still in flight for later steps (its fence covers every LDS
// write, DMA included); the memory clobber keeps the compiler from moving LDS
// accesses across the barrier.
template <int N>
__device__ __forceinline__ void step_barrier() {
    static_assert(N == 0 || N == 4 || N == 6 || N == 8 || N == 12, "vmcnt of a pipeline step");
    if (N == 0) asm volatile("s_waitcnt vmcnt(0)\n\ts_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
    else if (N == 4) asm volatile("s_waitcnt vmcnt(4)\n\ts_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
    else if (N == 6) asm volatile("s_waitcnt vmcnt(6)\n\ts_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
    else if (N == 8) asm volatile("s_waitcnt vmcnt(8)\n\ts_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
    else asm volatile("s_waitcnt vmcnt(12)\n\ts_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// NW waves per workgroup on the 128 x 128 tile: NW = 4, each wave 64 x 64 (one
// wave per SIMD); NW = 8, each wave 64 x 32 (two waves per SIMD: while one
// waits at a barrier or for its LDS reads the other keeps the matrix pipe
// busy; 96 accumulator registers per wave instead of 192).
template <int NW, bool F16>
__global__ __launch_bounds__(64 * NW) void cgemm_planes_kernel(const PlaneBatch* __restrict__ batches, int nsplit) {
    constexpr int NP = F16 ? 2 : 3;
    constexpr int JB = 16 / NW;                         // 16 x 16 blocks per wave along N
    constexpr int kIssuesPerStep = 2 * IssueOffsets<NW, NP>::kRuns;  // copies per wave per step
    __shared__ PStage<NP> S0, S1, S2;
    const PlaneBatch bt = batches[blockIdx.y];
    const int tiles_m = (bt.M + kPT - 1) / kPT, tiles_n = (bt.N + kPT - 1) / kPT;
    const int nitems = tiles_m * tiles_n * nsplit;
    int w = blockIdx.x;
    if (w >= nitems) return;
    int tile, split;
    if (bt.xbm > 0 && (int)gridDim.x == nitems) {
        // XCD x (workgroups w = x mod 8) takes a block of xbm x xbn tiles, all
        // their K slices: the workgroups running together on one XCD (one L2)
        // read xbm row panels of A and xbn of B, not one of A and all of B
        const int x = w & 7, li = w >> 3;
        const int nbx = tiles_n / bt.xbn;
        split = li % nsplit;
        const int lt = li / nsplit;
        tile = ((x / nbx) * bt.xbm + lt / bt.xbn) * tiles_n + (x % nbx) * bt.xbn + lt % bt.xbn;
    } else {
        // each XCD takes a contiguous range of items
        if ((nitems & 7) == 0 && (int)gridDim.x == nitems) w = (w & 7) * (nitems >> 3) + (w >> 3);
        tile = w / nsplit;
        split = w % nsplit;
    }
    const int m0 = (tile / tiles_n) * kPT, n0 = (tile % tiles_n) * kPT;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wm = (wave & 1) * 64, wn = (wave >> 1) * (16 * JB);
    IssueOffsets<NW, NP> off;
#pragma unroll
    for (int q = 0; q < 3; ++q)
#pragma unroll
        for (int u = 0; u < IssueOffsets<NW, NP>::kRuns; ++u) {
            const int t = wave * IssueOffsets<NW, NP>::kRuns + u;
            const int p = t >> 3, kb = (t >> 1) & 3, h = (t & 1) * 64;
            off.a[q][u] = 16u * (unsigned)((NP * q + p) * bt.a_ps + (int64_t)kb * bt.a_rows + h + lane);
            off.b[q][u] = 16u * (unsigned)((NP * q + p) * bt.b_ps + (int64_t)kb * bt.b_rows + h + lane);
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
        planes_issue<0, NW, NP>(S0, pa, pb, off, wave);
        planes_issue<1, NW, NP>(S1, pa, pb, off, wave);
        // One body for every K stage: the last one copies its own stage again
        // into the buffers of the following steps (never read), so the loop has
        // no tail and all copies are awaited before the epilogue.
        for (int ks = ks0; ks < ks1; ++ks) {
            // component re: buffer 0 (its copies were issued two steps ago)
            step_barrier<kIssuesPerStep>();
            planes_issue<2, NW, NP>(S2, pa, pb, off, wave);
            planes_compute<JB, F16>(S0, acc[0], wm, wn, lane);
            // component im: buffer 1
            step_barrier<kIssuesPerStep>();
            planes_issue<0, NW, NP>(S0, na, nb, off, wave);
            planes_compute<JB, F16>(S1, acc[1], wm, wn, lane);
            // component re + im: buffer 2
            step_barrier<kIssuesPerStep>();
            planes_issue<1, NW, NP>(S1, na, nb, off, wave);
            planes_compute<JB, F16>(S2, acc[2], wm, wn, lane);
            pa = na;
            pb = nb;
            if (ks + 2 < ks1) advance(na, nb);
        }
        step_barrier<0>();  // no copy may still write LDS when the workgroup ends
    }

    // fp16 planes carry the scales of A and B (powers of two): divided out here
    const float inv = F16 ? 1.0f / (bt.a_sc[0] * bt.b_sc[0]) : 1.0f;
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
                    C[row + (int64_t)col * bt.ldc] = make_float2((t1 - t2) * inv, (t3 - t1 - t2) * inv);
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

void plane_amax(const double2* src, int64_t n, unsigned* slot, hipStream_t s) {
    if (n <= 0) throw Error("plane_amax: empty input", -2);
    const unsigned g = (unsigned)std::max<int64_t>(1, std::min<int64_t>(kRedBlocks, (n + 1023) / 1024));
    hipLaunchKernelGGL(plane_amax_kernel, dim3(g), dim3(256), 0, s, src, n, slot);
    NC_HIP(hipGetLastError());
}

void split_planes(const double2* src, int64_t sr, int64_t sk, int64_t sz, int nz, PlaneSet& dst, int z0,
                  const unsigned* mslot, hipStream_t s) {
    if (nz <= 0) return;
    if (z0 < 0 || z0 + nz > dst.count) throw Error("split_planes: matrix index out of range", -2);
    if (dst.form == kPlaneF16 && mslot == nullptr) throw Error("split_planes: fp16 planes need their maximum", -2);
    u32x4* d = reinterpret_cast<u32x4*>(dst.at(z0));
    const int64_t dz = dst.np * dst.ps;
    float* sc = dst.scale.ptr + z0;
    const bool f16 = dst.form == kPlaneF16;
    if (sk == 1) {
        const int64_t waves = (int64_t)((dst.R + 15) / 16) * ((dst.Kb + 3) / 4);
        dim3 grid((unsigned)((waves * 64 + 255) / 256), (unsigned)nz);
        if (f16)
            hipLaunchKernelGGL((split_planes_kernel<true, true>), grid, dim3(256), 0, s, src, sr, sk, sz, d, dst.ps,
                               dz, dst.R, dst.K, dst.Rp, dst.Kb, mslot, sc);
        else
            hipLaunchKernelGGL((split_planes_kernel<true, false>), grid, dim3(256), 0, s, src, sr, sk, sz, d, dst.ps,
                               dz, dst.R, dst.K, dst.Rp, dst.Kb, mslot, sc);
    } else {
        const int64_t n = (int64_t)dst.R * dst.Kb;
        dim3 grid((unsigned)((n + 255) / 256), (unsigned)nz);
        if (f16)
            hipLaunchKernelGGL((split_planes_kernel<false, true>), grid, dim3(256), 0, s, src, sr, sk, sz, d, dst.ps,
                               dz, dst.R, dst.K, dst.Rp, dst.Kb, mslot, sc);
        else
            hipLaunchKernelGGL((split_planes_kernel<false, false>), grid, dim3(256), 0, s, src, sr, sk, sz, d, dst.ps,
                               dz, dst.R, dst.K, dst.Rp, dst.Kb, mslot, sc);
    }
    NC_HIP(hipGetLastError());
}

void mix_split_planes(MpoSite const& op, const float2* P, int64_t pslice, int nslice, int br, int bl, PlaneSet& Q,
                      PlaneQScale const* qs, hipStream_t s) {
    const int ndst = (int)op.tos.size();
    if (Q.count != ndst || Q.R != bl * op.d || Q.K != br) throw Error("mix_split_planes: plane shape", -2);
    if (nslice < 1) throw Error("mix_split_planes: no slices", -2);
    const bool f16 = Q.form == kPlaneF16;
    if (f16 && (qs == nullptr || !qs->amod_x || !qs->n1 || !qs->sR))
        throw Error("mix_split_planes: fp16 planes need the scale bounds", -2);
    const PlaneQScale q16 = f16 ? *qs : PlaneQScale{};
    const int64_t rs = (int64_t)br * bl;
    const int64_t waves = (int64_t)((bl + 15) / 16) * Q.Kb;
    dim3 grid((unsigned)((waves + 3) / 4), (unsigned)ndst);
    unsigned* q = reinterpret_cast<unsigned*>(Q.buf.ptr);
    const int* ptr = op.mix_ptr[0].ptr;
    const int* src = op.mix_src[0].ptr;
    const float2* W = op.mix_W32[0].ptr;
    const int64_t qz = Q.np * Q.ps;
    // 16-byte pair loads need an even br (then every pair offset is even) and
    // an even slice stride
    const bool even = br % 2 == 0 && pslice % 2 == 0;
    const int nsrc = (int)op.froms.size();
    // every destination from one read of each source, where that still gives
    // the chip enough waves (measured at chi = 1024: 0.072 against 0.089 ms;
    // at chi = 256 a launch of 512 waves is slower than the per-destination one)
    if (even && (op.d == 2 || op.d == 3) && nsrc <= 8 && ndst <= 8 && waves >= 4096) {
        dim3 g1((unsigned)((waves + 3) / 4));
#define NC_MIX_ALL(DD, F)                                                                                          \
    hipLaunchKernelGGL((mix_split_all_kernel<DD, 8, 8, F>), g1, dim3(256), 0, s, P, pslice, nslice, nsrc, ndst, ptr, \
                       src, W, br, bl, rs, q, Q.ps, qz, Q.Rp, Q.Kb, q16, Q.scale.ptr)
        if (op.d == 2) { if (f16) NC_MIX_ALL(2, true); else NC_MIX_ALL(2, false); }
        else { if (f16) NC_MIX_ALL(3, true); else NC_MIX_ALL(3, false); }
#undef NC_MIX_ALL
        NC_HIP(hipGetLastError());
        return;
    }
#define NC_MIX_SPLIT(DD, EV, F)                                                                                   \
    hipLaunchKernelGGL((mix_split_kernel<DD, EV, F>), grid, dim3(256), 0, s, P, pslice, nslice, ptr, src, W, op.d, \
                       br, bl, rs, q, Q.ps, qz, Q.Rp, Q.Kb, ndst, q16, Q.scale.ptr)
#define NC_MIX_SPLIT_E(DD, F) \
    if (even) NC_MIX_SPLIT(DD, true, F); else NC_MIX_SPLIT(DD, false, F)
#define NC_MIX_SPLIT_D(DD) \
    if (f16) { NC_MIX_SPLIT_E(DD, true); } else { NC_MIX_SPLIT_E(DD, false); }
    switch (op.d) {
        case 1: NC_MIX_SPLIT_D(1); break;
        case 2: NC_MIX_SPLIT_D(2); break;
        case 3: NC_MIX_SPLIT_D(3); break;
        case 4: NC_MIX_SPLIT_D(4); break;
        default: NC_MIX_SPLIT_D(0); break;
    }
#undef NC_MIX_SPLIT_D
#undef NC_MIX_SPLIT_E
#undef NC_MIX_SPLIT
    NC_HIP(hipGetLastError());
}

void reduce_slices(const float2* Y, int64_t slice, int ns, int64_t n, double2* out, unsigned* reset,
                   hipStream_t s) {
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(reduce_slices_kernel, dim3((unsigned)std::max<int64_t>(blocks, 1)), dim3(256), 0, s, Y, slice,
                       ns, n, out, reset);
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

// The XCD tile block of a batch (xbm x xbn tiles per XCD, eight blocks tiling
// the batch's tiles_m x tiles_n grid), the squarest that divides the grid:
// the workgroups of one XCD then share the fewest row panels.  0 x 0 (one
// contiguous range of items per XCD) when none exists.
static void xcd_blocks(PlaneBatch& b) {
    const int tm = (b.M + kPT - 1) / kPT, tn = (b.N + kPT - 1) / kPT;
    b.xbm = b.xbn = 0;
    int best = 1 << 30;
    for (int pm = 1; pm <= 8; pm *= 2) {
        const int pn = 8 / pm;  // blocks along M and N
        if (tm % pm || tn % pn) continue;
        const int bm = tm / pm, bn = tn / pn;
        if (bm + bn <= best) {  // ties: the smaller A block (batches sharing B run together)
            best = bm + bn;
            b.xbm = bm;
            b.xbn = bn;
        }
    }
}

void PlaneGemm::prepare(std::vector<PlaneBatch> const& bs, int nsplit_, int form_) {
    if (bs.empty()) throw Error("plane GEMM: no batches", -2);
    if (form_ != kPlaneBF16 && form_ != kPlaneF16) throw Error("plane GEMM: unknown part form", -2);
    form = form_;
    const int np = form == kPlaneF16 ? 6 : 9;
    nsplit = std::max(1, nsplit_);
    nbatch = (int)bs.size();
    items = 0;
    for (auto const& b : bs) {
        if (b.M <= 0 || b.N <= 0 || b.nseg <= 0 || b.nks <= 0) throw Error("plane GEMM: bad batch", -2);
        if (b.a_rows < (b.M + kPT - 1) / kPT * kPT || b.b_rows < (b.N + kPT - 1) / kPT * kPT || b.a_rows % kPT ||
            b.b_rows % kPT || b.a_ps < 4 * (int64_t)b.nks * b.a_rows || b.b_ps < 4 * (int64_t)b.nks * b.b_rows)
            throw Error("plane GEMM: planes do not cover the batch's tiles", -2);
        if (nsplit > b.nseg * b.nks) throw Error("plane GEMM: more K slices than K stages", -2);
        if (form == kPlaneF16 && (b.a_sc == nullptr || b.b_sc == nullptr))
            throw Error("plane GEMM: fp16 planes without their scales", -2);
        // a copy's byte offset within one matrix's planes is 32-bit
        if (16 * np * b.a_ps >= (int64_t)1 << 32 || 16 * np * b.b_ps >= (int64_t)1 << 32)
            throw Error("plane GEMM: plane set above 4 GiB per matrix", -2);
        const int64_t it = (int64_t)((b.M + kPT - 1) / kPT) * ((b.N + kPT - 1) / kPT) * nsplit;
        items = std::max(items, it);
    }
    std::vector<PlaneBatch> hb(bs);
    for (auto& b : hb) xcd_blocks(b);
    batches.resize(hb.size());
    h2d(batches.ptr, hb.data(), hb.size() * sizeof(PlaneBatch));
}

// eight waves per workgroup: measured 0.80 against 0.84 ms per chi = 1024 plane
// matvec with four (the two waves of a SIMD overlap each other's LDS waits);
// fp16 planes: 0.555 against 0.583 ms
constexpr int kPlaneWaves = 8;

void PlaneGemm::launch(hipStream_t s) const {
    if (form == kPlaneF16)
        hipLaunchKernelGGL((cgemm_planes_kernel<kPlaneWaves, true>), dim3((unsigned)items, (unsigned)nbatch),
                           dim3(64 * kPlaneWaves), 0, s, batches.ptr, nsplit);
    else
        hipLaunchKernelGGL((cgemm_planes_kernel<kPlaneWaves, false>), dim3((unsigned)items, (unsigned)nbatch),
                           dim3(64 * kPlaneWaves), 0, s, batches.ptr, nsplit);
    NC_HIP(hipGetLastError());
}

// ----------------------------------------------------------------------------
int& mixed_planes_mode() {
    static int mode = 1;
    return mode;
}

int& plane_form() {
    static int form = kPlaneF16;
    return form;
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
    : bl(bl_), br(br_), form(plane_form()), op(op_) {
    const int d = op.d;
    const int nf = (int)op.froms.size(), nt = (int)op.tos.size();
    if (bl <= 0 || br <= 0 || d <= 0 || nf == 0 || nt == 0) throw Error("plane H_eff: bad shape", -2);
    const int64_t rs = (int64_t)br * bl;
    Lp.shape(bl, bl, nf, form);
    Rp.shape(br, br, nt, form);
    Xp.shape(br, bl, d, form);
    Qp.shape(bl * d, br, nt, form);
    // fp16 planes: maxima slots (x: largest component and modulus; one per
    // L_f and R_t) and the column 1-norms of the L_f, all zero before their
    // atomic maxima
    const unsigned* lslot = nullptr;
    const unsigned* rslot = nullptr;
    if (form == kPlaneF16) {
        slots.resize((size_t)(2 + 2 * nf + 2 * nt));
        n1.resize((size_t)nf);
        NC_HIP(hipMemsetAsync(slots.ptr, 0, slots.bytes(), s));
        NC_HIP(hipMemsetAsync(n1.ptr, 0, n1.bytes(), s));
        for (int fi = 0; fi < nf; ++fi) {
            const double2* Lf = L + (int64_t)op.froms[fi] * bl * bl;
            plane_amax(Lf, (int64_t)bl * bl, slots.ptr + 2 + 2 * fi, s);
            hipLaunchKernelGGL(colnorm1_max_kernel, dim3((unsigned)((bl + 3) / 4)), dim3(256), 0, s, Lf, bl, bl,
                               reinterpret_cast<unsigned*>(n1.ptr + fi));
            NC_HIP(hipGetLastError());
        }
        for (int ti = 0; ti < nt; ++ti)
            plane_amax(R + (int64_t)op.tos[ti] * br * br, (int64_t)br * br, slots.ptr + 2 + 2 * nf + 2 * ti, s);
        qs = PlaneQScale{reinterpret_cast<const float*>(slots.ptr + 1), n1.ptr, Rp.scale.ptr};
        lslot = slots.ptr + 2;
        rslot = slots.ptr + 2 + 2 * nf;
    }
    for (int fi = 0; fi < nf; ++fi)  // L_f(k, n) = L[f bl^2 + k + n bl]: rows n, K index k
        split_planes(L + (int64_t)op.froms[fi] * bl * bl, bl, 1, 0, 1, Lp, fi, lslot ? lslot + 2 * fi : nullptr, s);
    for (int ti = 0; ti < nt; ++ti)  // R_t(m, k) = R[t br^2 + m + k br]
        split_planes(R + (int64_t)op.tos[ti] * br * br, 1, br, 0, 1, Rp, ti, rslot ? rslot + 2 * ti : nullptr, s);

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
            b.a_sc = Xp.scale.ptr + a;
            b.b_sc = Lp.scale.ptr + fi;
            bs.push_back(b);
        }
    gemm_a.prepare(bs, nsa, form);

    // step C: y (br x bl d) = sum_t R_t Q_t, K = br per segment (one segment
    // per 't' index: consecutive matrices of the R and Q plane sets); with fp16
    // planes every segment's two scales have the product of segment 0's
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
    b.a_seg = Rp.np * Rp.ps;
    b.b_seg = Qp.np * Qp.ps;
    b.a_rows = Rp.Rp;
    b.b_rows = Qp.Rp;
    b.M = br;
    b.N = bl * d;
    b.nseg = nt;
    b.nks = Rp.Kb / 4;
    b.a_sc = Rp.scale.ptr;
    b.b_sc = Qp.scale.ptr;
    bs.push_back(b);
    gemm_c.prepare(bs, nsc, form);
}

// fp16 planes: the bound on ||x|| in slots 0 and 1 (largest |re|, |im|,
// |re + im| <= sqrt(2) bound, largest modulus <= bound), which then stand for
// every x and are not reset
void PlaneHeff::assume_input_norm(double bound, hipStream_t s) {
    if (form != kPlaneF16) return;
    if (!(bound > 0) || !std::isfinite(bound)) throw Error("plane H_eff: bad input norm bound", -2);
    const float m[2] = {(float)(std::sqrt(2.0) * bound * (1 + 1e-6)), (float)(bound * (1 + 1e-6))};
    unsigned u[2];
    std::memcpy(u, m, sizeof u);
    NC_HIP(hipMemcpyAsync(slots.ptr, u, sizeof u, hipMemcpyHostToDevice, s));
    NC_HIP(hipStreamSynchronize(s));
    norm_bound = bound;
}

// fp16 planes: x's maxima go into slots 0 and 1, which sum() zeroes again
// for the next matvec (unless a norm bound stands for them)
void PlaneHeff::split_x(const double2* x, hipStream_t s) {
    const int64_t rs = (int64_t)br * bl;
    if (form == kPlaneF16 && norm_bound == 0) plane_amax(x, rs * op.d, slots.ptr, s);
    // X_a(m, k) = x[a rs + m + k br]
    split_planes(x, 1, br, rs, op.d, Xp, 0, form == kPlaneF16 ? slots.ptr : nullptr, s);
}

void PlaneHeff::mix(hipStream_t s) {
    mix_split_planes(op, P.ptr, pslice, nsa, br, bl, Qp, form == kPlaneF16 ? &qs : nullptr, s);
}

void PlaneHeff::sum(double2* y, hipStream_t s) {
    reduce_slices(Y.ptr, yslice, nsc, yslice, y, form == kPlaneF16 && norm_bound == 0 ? slots.ptr : nullptr, s);
}

void PlaneHeff::apply(const double2* x, double2* y, hipStream_t s) {
    split_x(x, s);
    step_a(s);
    mix(s);
    step_c(s);
    sum(y, s);
}

}  // namespace nc
