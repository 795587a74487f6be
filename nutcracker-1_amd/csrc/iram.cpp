// Implicitly restarted Arnoldi (Lehoucq-Sorensen), nev = 1, exact shifts, with
// the Krylov basis resident on the device.  This is the algorithm of ARPACK's
// znaupd/zneupd as the reference drives it (core.f95:2465-2527: bmat = 'I',
// mode 1, iparam(1) = 1, info = 1 so the start vector is the caller's guess):
//   * Arnoldi steps with classical Gram-Schmidt applied twice (CGS2; znaitr
//     applies the second pass when the first loses more than 0.717 of the
//     norm -- here it always runs, which is never less orthogonal), a third
//     pass when the second loses as much, and znaitr's rule for the residual:
//     zero (an invariant subspace) only when the last pass taken also lost
//     that much, otherwise kept as the next direction however small.  The
//     first pass projects on the last two basis vectors only: for the
//     Hermitian effective Hamiltonian the other components are round-off,
//     which the full second pass removes (a non-Hermitian operator leaves
//     more, and the second and third passes then act as znaitr's two);
//   * Ritz pairs of the small Hessenberg H on the host; convergence when the
//     Ritz estimate |beta e_m^T s| <= tol max(eps^(2/3), |theta|) (znconv);
//   * restart by ncv - 1 shifted QR steps with the unwanted Ritz values
//     (ncv <= 3, as ARPACK with nev = 1), or keeping ncv / 2 Ritz vectors
//     (thick restart) for larger ncv.
// The restart count returned plays the role of ARPACK's iparam(3).
//
// The Arnoldi extension from k to ncv vectors runs without a host round trip:
// each step's Gram-Schmidt coefficients, norm and Hessenberg column stay on the
// device (two-pass deterministic reductions), the new basis vector is written
// in place, and the host reads the Hessenberg matrices once per extension.
// Several independent problems (one per chain) extend in lockstep, so one
// kernel launch serves all of them.
#include "nc_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <limits>

namespace nc {

namespace {

constexpr int KT = 256;  // threads per block
constexpr int KC = 8;    // basis columns per register group

struct KDesc {
    double2* V;   // n x (ncv + 1), column-major: basis columns 0..ncv-1, residual f in column ncv
    int64_t n;
    int kstart;   // first step of the current extension; ncv = idle
    int _pad;
};

__device__ __forceinline__ double kwarp_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Vectors normalised on the device, with no host read: dst = src / ||src|| (0
// when the norm is 0), the norm also written to *beta when beta is set.  Partial
// sums per block and their total are formed in a fixed order.
struct NormItem {
    const double2* src;
    double2* dst;
    int64_t n;
    double* beta;
};
__global__ __launch_bounds__(256) void knorm_partial_kernel(const NormItem* __restrict__ it,
                                                            double* __restrict__ partial) {
    const NormItem d = it[blockIdx.y];
    double acc = 0;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < d.n; e += (int64_t)gridDim.x * blockDim.x) {
        const double2 v = d.src[e];
        acc += v.x * v.x + v.y * v.y;
    }
    __shared__ double sh[4];
    acc = kwarp_sum(acc);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) partial[(int64_t)blockIdx.y * gridDim.x + blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}
__global__ __launch_bounds__(64) void knorm_final_kernel(const NormItem* __restrict__ it,
                                                         const double* __restrict__ partial, int nb,
                                                         double* __restrict__ scale) {
    const int p = blockIdx.x;
    double s2 = 0;
    for (int b = threadIdx.x; b < nb; b += 64) s2 += partial[(int64_t)p * nb + b];
    s2 = kwarp_sum(s2);
    if (threadIdx.x != 0) return;
    const double beta = sqrt(s2);
    if (it[p].beta) *it[p].beta = beta;
    scale[p] = beta > 0.0 ? 1.0 / beta : 0.0;
}
__global__ __launch_bounds__(256) void knorm_apply_kernel(const NormItem* __restrict__ it,
                                                          const double* __restrict__ scale) {
    const NormItem d = it[blockIdx.y];
    const double f = scale[blockIdx.y];
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < d.n; e += (int64_t)gridDim.x * blockDim.x) {
        const double2 v = d.src[e];
        d.dst[e] = make_double2(v.x * f, v.y * f);
    }
}

// work space of normalize_on_device for up to `items` vectors
struct NormSpace {
    DevBuf<double> partial, scale;
    int cap = 0;
    static unsigned blocks(size_t items, int64_t maxn) {
        return (unsigned)std::max<int64_t>(1, std::min<int64_t>(std::max<int64_t>(1, 2048 / (int64_t)items),
                                                                 (maxn + 255) / 256));
    }
    void reserve(int items) {
        if (items <= cap) return;
        partial.resize((size_t)items * blocks(1, 1 << 30));
        scale.resize((size_t)items);
        cap = items;
    }
};
void normalize_on_device(std::vector<NormItem> const& items, NormSpace& ws, hipStream_t s) {
    if (items.empty()) return;
    if ((int)items.size() > ws.cap) throw Error("normalize_on_device: work space too small", -8);
    int64_t maxn = 1;
    for (auto const& it : items) maxn = std::max(maxn, it.n);
    const unsigned g = NormSpace::blocks(items.size(), maxn);
    int slot;
    auto* d = (const NormItem*)stage_descriptors(items.data(), items.size() * sizeof(NormItem), s, &slot);
    const unsigned P = (unsigned)items.size();
    hipLaunchKernelGGL(knorm_partial_kernel, dim3(g, P), dim3(256), 0, s, d, ws.partial.ptr);
    hipLaunchKernelGGL(knorm_final_kernel, dim3(P), dim3(64), 0, s, d, ws.partial.ptr, (int)g, ws.scale.ptr);
    hipLaunchKernelGGL(knorm_apply_kernel, dim3(g, P), dim3(256), 0, s, d, ws.scale.ptr);
    NC_HIP(hipGetLastError());
    release_descriptors(slot, s);
}

// partial dots of columns 0..j+1 of V against w = column j+1 (the last one is
// ||w||^2): partial[((p * ldp) + c) * nblk + b].  gate (optional): problems
// whose flag is 0 skip the step.
__global__ __launch_bounds__(KT) void kdot_kernel(const KDesc* __restrict__ desc, int j, double2* __restrict__ partial,
                                                  int ldp, const int* __restrict__ gate) {
    const int p = blockIdx.y;
    const KDesc d = desc[p];
    if (j < d.kstart) return;
    if (gate && gate[p] == 0) return;
    const int m = j + 2;
    const int64_t n = d.n;
    const double2* __restrict__ w = d.V + (int64_t)(j + 1) * n;
    __shared__ double sh[KT / 64][2 * KC];
    const int nb = gridDim.x;
    for (int c0 = 0; c0 < m; c0 += KC) {
        double re[KC], im[KC];
#pragma unroll
        for (int q = 0; q < KC; ++q) { re[q] = 0; im[q] = 0; }
        for (int64_t i = (int64_t)blockIdx.x * KT + threadIdx.x; i < n; i += (int64_t)nb * KT) {
            const double2 wv = w[i];
#pragma unroll
            for (int q = 0; q < KC; ++q) {
                if (c0 + q < m) {
                    const double2 a = d.V[i + (int64_t)(c0 + q) * n];
                    re[q] += a.x * wv.x + a.y * wv.y;
                    im[q] += a.x * wv.y - a.y * wv.x;
                }
            }
        }
#pragma unroll
        for (int q = 0; q < KC; ++q) {
            re[q] = kwarp_sum(re[q]);
            im[q] = kwarp_sum(im[q]);
        }
        if ((threadIdx.x & 63) == 0) {
#pragma unroll
            for (int q = 0; q < KC; ++q) {
                sh[threadIdx.x >> 6][2 * q] = re[q];
                sh[threadIdx.x >> 6][2 * q + 1] = im[q];
            }
        }
        __syncthreads();
        if (threadIdx.x < KC && c0 + (int)threadIdx.x < m) {
            const int q = threadIdx.x;
            double tr = 0, ti = 0;
            for (int wv = 0; wv < KT / 64; ++wv) { tr += sh[wv][2 * q]; ti += sh[wv][2 * q + 1]; }
            partial[((int64_t)p * ldp + c0 + q) * nb + blockIdx.x] = make_double2(tr, ti);
        }
        __syncthreads();
    }
}

// The same dots for m = j + 2 <= MM columns in one pass over w (the general
// kernel above reads w once per group of KC columns).  The per-block partial
// sums are formed in the same order as there, so the results are the same.
template <int MM>
__global__ __launch_bounds__(KT) void kdot1_kernel(const KDesc* __restrict__ desc, int j,
                                                   double2* __restrict__ partial, int ldp,
                                                   const int* __restrict__ gate) {
    const int p = blockIdx.y;
    const KDesc d = desc[p];
    if (j < d.kstart) return;
    if (gate && gate[p] == 0) return;
    const int m = j + 2;
    const int64_t n = d.n;
    const double2* __restrict__ w = d.V + (int64_t)(j + 1) * n;
    const int nb = gridDim.x;
    double re[MM], im[MM];
#pragma unroll
    for (int q = 0; q < MM; ++q) { re[q] = 0; im[q] = 0; }
    for (int64_t i = (int64_t)blockIdx.x * KT + threadIdx.x; i < n; i += (int64_t)nb * KT) {
        const double2 wv = w[i];
        double2 a[MM];
#pragma unroll
        for (int q = 0; q < MM; ++q)
            if (q < m) a[q] = d.V[i + (int64_t)q * n];
#pragma unroll
        for (int q = 0; q < MM; ++q) {
            if (q < m) {
                re[q] += a[q].x * wv.x + a[q].y * wv.y;
                im[q] += a[q].x * wv.y - a[q].y * wv.x;
            }
        }
    }
    __shared__ double sh[KT / 64][2 * MM];
#pragma unroll
    for (int q = 0; q < MM; ++q) {
        re[q] = kwarp_sum(re[q]);
        im[q] = kwarp_sum(im[q]);
    }
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int q = 0; q < MM; ++q) {
            sh[threadIdx.x >> 6][2 * q] = re[q];
            sh[threadIdx.x >> 6][2 * q + 1] = im[q];
        }
    }
    __syncthreads();
    if (threadIdx.x < MM && (int)threadIdx.x < m) {
        const int q = threadIdx.x;
        double tr = 0, ti = 0;
        for (int wv = 0; wv < KT / 64; ++wv) { tr += sh[wv][2 * q]; ti += sh[wv][2 * q + 1]; }
        partial[((int64_t)p * ldp + q) * nb + blockIdx.x] = make_double2(tr, ti);
    }
}

// The first Gram-Schmidt pass against the last two basis vectors only: the
// partial dots of V_{j-1}, V_j against w and ||w||^2 (entry j+1), in the layout
// of kdot1_kernel, with zeros for the older columns.  For a Hermitian operator
// A v_j lies in span(v_{j-1}, v_j, v_{j+1}) (the Hessenberg matrix is
// tridiagonal), so the full second pass only removes round-off along the older
// vectors; a non-Hermitian operator leaves more there, which the second pass
// removes and the gated third pass refines.
__global__ __launch_bounds__(KT) void kdotl_kernel(const KDesc* __restrict__ desc, int j,
                                                   double2* __restrict__ partial, int ldp) {
    const int p = blockIdx.y;
    const KDesc d = desc[p];
    if (j < d.kstart) return;
    const int64_t n = d.n;
    const double2* __restrict__ w = d.V + (int64_t)(j + 1) * n;
    const double2* __restrict__ vj = d.V + (int64_t)j * n;
    const double2* __restrict__ vm = d.V + (int64_t)(j > 0 ? j - 1 : 0) * n;
    const int nb = gridDim.x;
    double r0 = 0, i0 = 0, r1 = 0, i1 = 0, nw = 0;
    for (int64_t i = (int64_t)blockIdx.x * KT + threadIdx.x; i < n; i += (int64_t)nb * KT) {
        const double2 wv = w[i], a = vj[i];
        const double2 b = j > 0 ? vm[i] : make_double2(0.0, 0.0);
        r1 += a.x * wv.x + a.y * wv.y;
        i1 += a.x * wv.y - a.y * wv.x;
        r0 += b.x * wv.x + b.y * wv.y;
        i0 += b.x * wv.y - b.y * wv.x;
        nw += wv.x * wv.x + wv.y * wv.y;
    }
    r0 = kwarp_sum(r0);
    i0 = kwarp_sum(i0);
    r1 = kwarp_sum(r1);
    i1 = kwarp_sum(i1);
    nw = kwarp_sum(nw);
    __shared__ double sh[KT / 64][5];
    if ((threadIdx.x & 63) == 0) {
        sh[threadIdx.x >> 6][0] = r0;
        sh[threadIdx.x >> 6][1] = i0;
        sh[threadIdx.x >> 6][2] = r1;
        sh[threadIdx.x >> 6][3] = i1;
        sh[threadIdx.x >> 6][4] = nw;
    }
    __syncthreads();
    for (int c = threadIdx.x; c <= j + 1; c += KT) {
        double tr = 0, ti = 0;
        if (c == j - 1) {
            for (int wv = 0; wv < KT / 64; ++wv) { tr += sh[wv][0]; ti += sh[wv][1]; }
        } else if (c == j) {
            for (int wv = 0; wv < KT / 64; ++wv) { tr += sh[wv][2]; ti += sh[wv][3]; }
        } else if (c == j + 1) {
            for (int wv = 0; wv < KT / 64; ++wv) tr += sh[wv][4];
        }
        partial[((int64_t)p * ldp + c) * nb + blockIdx.x] = make_double2(tr, ti);
    }
}

// out[p * ldp + c] = sum_b partial[((p * ldp) + c) * nblk + b], c = 0..j+1
__global__ __launch_bounds__(KT) void kreduce_kernel(const KDesc* __restrict__ desc, int j,
                                                     const double2* __restrict__ partial, int ldp, int nblk,
                                                     double2* __restrict__ out, const int* __restrict__ gate) {
    const int p = blockIdx.y, c = blockIdx.x;
    if (j < desc[p].kstart || c > j + 1) return;
    if (gate && gate[p] == 0) return;
    __shared__ double sre[KT / 64], sim[KT / 64];
    double re = 0, im = 0;
    for (int b = threadIdx.x; b < nblk; b += KT) {
        const double2 v = partial[((int64_t)p * ldp + c) * nblk + b];
        re += v.x;
        im += v.y;
    }
    re = kwarp_sum(re);
    im = kwarp_sum(im);
    if ((threadIdx.x & 63) == 0) { sre[threadIdx.x >> 6] = re; sim[threadIdx.x >> 6] = im; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double tr = 0, ti = 0;
        for (int k = 0; k < KT / 64; ++k) { tr += sre[k]; ti += sim[k]; }
        out[(int64_t)p * ldp + c] = make_double2(tr, ti);
    }
}

// w -= sum_{c <= j} V_c coef_c, and partial ||w||^2 per block
__global__ __launch_bounds__(KT) void kupdate_kernel(const KDesc* __restrict__ desc, int j,
                                                     const double2* __restrict__ coef, int ldp,
                                                     double* __restrict__ pnorm, const int* __restrict__ gate) {
    const int p = blockIdx.y;
    const KDesc d = desc[p];
    if (j < d.kstart) return;
    if (gate && gate[p] == 0) return;
    __shared__ double2 sc[64];
    __shared__ double sn[KT / 64];
    const int m = j + 1;
    for (int c = threadIdx.x; c < m; c += KT) sc[c] = coef[(int64_t)p * ldp + c];
    __syncthreads();
    const int64_t n = d.n;
    double2* __restrict__ w = d.V + (int64_t)(j + 1) * n;
    double acc = 0;
    for (int64_t i = (int64_t)blockIdx.x * KT + threadIdx.x; i < n; i += (int64_t)gridDim.x * KT) {
        double2 v = w[i];
        for (int c = 0; c < m; ++c) {
            const double2 a = d.V[i + (int64_t)c * n];
            const double2 h = sc[c];
            v.x -= a.x * h.x - a.y * h.y;
            v.y -= a.x * h.y + a.y * h.x;
        }
        w[i] = v;
        acc += v.x * v.x + v.y * v.y;
    }
    acc = kwarp_sum(acc);
    if ((threadIdx.x & 63) == 0) sn[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0;
        for (int k = 0; k < KT / 64; ++k) t += sn[k];
        pnorm[(int64_t)p * gridDim.x + blockIdx.x] = t;
    }
}

// The second update, with znaitr's test taken before it from the norms
// (||w2||^2 = ||w1||^2 - ||c2||^2 for the second pass's coefficients c2).  When
// the update keeps more than 0.8 of the norm (no third pass follows) and
// `allow` (step j + 1 exists and its fused kernel takes the correction), the
// update is deferred: column j + 1 keeps w1, which the finishing kernels scale
// by 1 / ||w2||, and step j + 1's fused kernel subtracts V_{0..j} corr,
// corr = c2 / ||w2||, from it.  For the Hermitian effective Hamiltonian c2 is
// round-off (or the complex64 operator's departure from Hermitian), so the
// operator of step j + 1 sees a vector within that much of the final one; the
// Hessenberg column of step j + 1 accounts for the difference (kfinish).
// flag[p] = 1 marks a deferred update (and b2s[p] = ||w2||).
__global__ __launch_bounds__(KT) void kupdate2_kernel(const KDesc* __restrict__ desc, int j,
                                                      const double2* __restrict__ coef, int ldp,
                                                      double* __restrict__ pnorm, int allow, int* __restrict__ flag,
                                                      double2* __restrict__ corr, double* __restrict__ b2s) {
    const int p = blockIdx.y;
    const KDesc d = desc[p];
    if (j < d.kstart) return;
    __shared__ double2 sc[64];
    __shared__ double sn[KT / 64];
    __shared__ int sdelay;
    __shared__ double sb2;
    const int m = j + 1;
    for (int c = threadIdx.x; c < m; c += KT) sc[c] = coef[(int64_t)p * ldp + c];
    __syncthreads();
    if (threadIdx.x == 0) {
        double s2 = 0;
        for (int c = 0; c < m; ++c) s2 += sc[c].x * sc[c].x + sc[c].y * sc[c].y;
        const double b1sq = coef[(int64_t)p * ldp + m].x;
        const double b2sq = b1sq - s2;
        const int delay = (allow && b2sq > 0.64 * b1sq && b2sq > 0.0) ? 1 : 0;
        sdelay = delay;
        sb2 = delay ? sqrt(b2sq) : 0.0;
    }
    __syncthreads();
    if (sdelay) {
        if (blockIdx.x == 0) {
            const double inv = 1.0 / sb2;
            for (int c = threadIdx.x; c < m; c += KT)
                corr[(int64_t)p * ldp + c] = make_double2(sc[c].x * inv, sc[c].y * inv);
            if (threadIdx.x == 0) {
                flag[p] = 1;
                b2s[p] = sb2;
            }
        }
        return;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) flag[p] = 0;
    const int64_t n = d.n;
    double2* __restrict__ w = d.V + (int64_t)(j + 1) * n;
    double acc = 0;
    for (int64_t i = (int64_t)blockIdx.x * KT + threadIdx.x; i < n; i += (int64_t)gridDim.x * KT) {
        double2 v = w[i];
        for (int c = 0; c < m; ++c) {
            const double2 a = d.V[i + (int64_t)c * n];
            const double2 h = sc[c];
            v.x -= a.x * h.x - a.y * h.y;
            v.y -= a.x * h.y + a.y * h.x;
        }
        w[i] = v;
        acc += v.x * v.x + v.y * v.y;
    }
    acc = kwarp_sum(acc);
    if ((threadIdx.x & 63) == 0) sn[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0;
        for (int k = 0; k < KT / 64; ++k) t += sn[k];
        pnorm[(int64_t)p * gridDim.x + blockIdx.x] = t;
    }
}

// The first update fused with the second pass's dot products: w -= sum_{c<=j}
// V_c coef_c, then the partial dots of V_0..V_j against the updated w and its
// squared norm (entry j+1), in the layout of kdot_kernel.  Each row's basis
// entries are loaded once, into registers, and serve both.
// When step j - 1 deferred its second update (pend[p]), column j holds
// v' = w1 / beta and is first corrected to v_j = v' - sum_{c<j} V_c corr_c, in
// registers and written back.
template <int MC>
__global__ __launch_bounds__(KT) void kupdot_kernel(const KDesc* __restrict__ desc, int j,
                                                    const double2* __restrict__ coef, int ldp,
                                                    double2* __restrict__ partial, const int* __restrict__ pend,
                                                    const double2* __restrict__ corr) {
    const int p = blockIdx.y;
    const KDesc d = desc[p];
    if (j < d.kstart) return;
    __shared__ double2 sc[MC], scr[MC];
    __shared__ double sh[KT / 64][2 * MC + 1];
    const int m = j + 1;
    const bool fix = pend != nullptr && j - 1 >= d.kstart && pend[p] != 0;
    for (int c = threadIdx.x; c < m; c += KT) {
        sc[c] = coef[(int64_t)p * ldp + c];
        scr[c] = (fix && c < m - 1) ? corr[(int64_t)p * ldp + c] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    const int64_t n = d.n;
    double2* __restrict__ w = d.V + (int64_t)(j + 1) * n;
    double2* __restrict__ vlast = d.V + (int64_t)j * n;
    double re[MC], im[MC], nrm = 0;
#pragma unroll
    for (int c = 0; c < MC; ++c) { re[c] = 0; im[c] = 0; }
    const int nb = gridDim.x;
    for (int64_t i = (int64_t)blockIdx.x * KT + threadIdx.x; i < n; i += (int64_t)nb * KT) {
        // The coefficients are read from LDS row by row: hoisted out of the
        // loop they would hold 2 MC more double2 registers (255 VGPRs and one
        // wave per SIMD at MC = 16, against 158 and three with this barrier).
        asm volatile("" ::: "memory");
        // the row's basis entries, all loads issued before the first use
        double2 a[MC];
#pragma unroll
        for (int c = 0; c < MC; ++c)
            if (c < m) a[c] = d.V[i + (int64_t)c * n];
        double2 v = w[i];
        if (fix) {
            double cx = 0, cy = 0;
#pragma unroll
            for (int c = 0; c < MC; ++c) {
                if (c < m - 1) {
                    const double2 h = scr[c];
                    cx += a[c].x * h.x - a[c].y * h.y;
                    cy += a[c].x * h.y + a[c].y * h.x;
                }
            }
#pragma unroll
            for (int c = 0; c < MC; ++c) {
                if (c == m - 1) {
                    a[c].x -= cx;
                    a[c].y -= cy;
                    vlast[i] = a[c];
                }
            }
        }
#pragma unroll
        for (int c = 0; c < MC; ++c) {
            if (c < m) {
                const double2 h = sc[c];
                v.x -= a[c].x * h.x - a[c].y * h.y;
                v.y -= a[c].x * h.y + a[c].y * h.x;
            }
        }
        w[i] = v;
        nrm += v.x * v.x + v.y * v.y;
#pragma unroll
        for (int c = 0; c < MC; ++c) {
            if (c < m) {
                re[c] += a[c].x * v.x + a[c].y * v.y;
                im[c] += a[c].x * v.y - a[c].y * v.x;
            }
        }
    }
#pragma unroll
    for (int c = 0; c < MC; ++c) {
        re[c] = kwarp_sum(re[c]);
        im[c] = kwarp_sum(im[c]);
    }
    nrm = kwarp_sum(nrm);
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int c = 0; c < MC; ++c) {
            sh[threadIdx.x >> 6][2 * c] = re[c];
            sh[threadIdx.x >> 6][2 * c + 1] = im[c];
        }
        sh[threadIdx.x >> 6][2 * MC] = nrm;
    }
    __syncthreads();
    for (int c = threadIdx.x; c <= m; c += KT) {
        double tr = 0, ti = 0;
        if (c < m) {
            for (int wv = 0; wv < KT / 64; ++wv) { tr += sh[wv][2 * c]; ti += sh[wv][2 * c + 1]; }
        } else {
            for (int wv = 0; wv < KT / 64; ++wv) tr += sh[wv][2 * MC];
        }
        partial[((int64_t)p * ldp + c) * nb + blockIdx.x] = make_double2(tr, ti);
    }
}

// v + its value on the other lane of the pair (lanes 2r, 2r + 1): both lanes
// get the same sum (the addition is commutative).
__device__ __forceinline__ double kpair_sum(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)b, 0xB1, 0xF, 0xF, false);  // quad_perm 1, 0, 3, 2
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), 0xB1, 0xF, 0xF, false);
    const double o = __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
    return v + o;
}
// Sum over the 32 lanes of one parity (xor offsets 32 .. 2).
__device__ __forceinline__ double kparity_sum(double v) {
    for (int off = 32; off > 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// kupdot_kernel with each row shared by a lane pair, for 9 to 24 columns: lane
// parity h takes the columns c = 2 k + h (k < MH), so a lane holds MH basis
// entries and MH dot accumulators instead of 2 MH (149 VGPRs and three waves
// per SIMD at 24 columns, against 223 and two; 108 and four at 16, against 158
// and three).  The pair adds its partial correction and update sums across the
// two lanes; lane 0 of the pair writes the row.
template <int MH>
__global__ __launch_bounds__(KT) void kupdot_pair_kernel(const KDesc* __restrict__ desc, int j,
                                                         const double2* __restrict__ coef, int ldp,
                                                         double2* __restrict__ partial, const int* __restrict__ pend,
                                                         const double2* __restrict__ corr) {
    constexpr int MC = 2 * MH;
    const int p = blockIdx.y;
    const KDesc d = desc[p];
    if (j < d.kstart) return;
    __shared__ double2 sc[MC], scr[MC];
    __shared__ double sh[KT / 64][2 * MC + 1];
    const int m = j + 1;
    const bool fix = pend != nullptr && j - 1 >= d.kstart && pend[p] != 0;
    for (int c = threadIdx.x; c < m; c += KT) {
        sc[c] = coef[(int64_t)p * ldp + c];
        scr[c] = (fix && c < m - 1) ? corr[(int64_t)p * ldp + c] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    const int64_t n = d.n;
    double2* __restrict__ w = d.V + (int64_t)(j + 1) * n;
    double2* __restrict__ vlast = d.V + (int64_t)j * n;
    const int h = threadIdx.x & 1;
    double re[MH], im[MH], nrm = 0;
#pragma unroll
    for (int k = 0; k < MH; ++k) { re[k] = 0; im[k] = 0; }
    const int nb = gridDim.x;
    for (int64_t i = ((int64_t)blockIdx.x * KT + threadIdx.x) >> 1; i < n; i += (int64_t)nb * (KT / 2)) {
        asm volatile("" ::: "memory");  // coefficients stay in LDS (see kupdot_kernel)
        double2 a[MH];
#pragma unroll
        for (int k = 0; k < MH; ++k)
            if (2 * k + h < m) a[k] = d.V[i + (int64_t)(2 * k + h) * n];
        double2 v = w[i];
        if (fix) {
            double cx = 0, cy = 0;
#pragma unroll
            for (int k = 0; k < MH; ++k) {
                if (2 * k + h < m - 1) {
                    const double2 q = scr[2 * k + h];
                    cx += a[k].x * q.x - a[k].y * q.y;
                    cy += a[k].x * q.y + a[k].y * q.x;
                }
            }
            cx = kpair_sum(cx);
            cy = kpair_sum(cy);
#pragma unroll
            for (int k = 0; k < MH; ++k) {
                if (2 * k + h == m - 1) {
                    a[k].x -= cx;
                    a[k].y -= cy;
                    vlast[i] = a[k];
                }
            }
        }
        double ux = 0, uy = 0;
#pragma unroll
        for (int k = 0; k < MH; ++k) {
            if (2 * k + h < m) {
                const double2 q = sc[2 * k + h];
                ux += a[k].x * q.x - a[k].y * q.y;
                uy += a[k].x * q.y + a[k].y * q.x;
            }
        }
        ux = kpair_sum(ux);
        uy = kpair_sum(uy);
        v.x -= ux;
        v.y -= uy;
        if (h == 0) {
            w[i] = v;
            nrm += v.x * v.x + v.y * v.y;
        }
#pragma unroll
        for (int k = 0; k < MH; ++k) {
            if (2 * k + h < m) {
                re[k] += a[k].x * v.x + a[k].y * v.y;
                im[k] += a[k].x * v.y - a[k].y * v.x;
            }
        }
    }
#pragma unroll
    for (int k = 0; k < MH; ++k) {
        re[k] = kparity_sum(re[k]);
        im[k] = kparity_sum(im[k]);
    }
    nrm = kparity_sum(nrm);
    const int lane = threadIdx.x & 63;
    if (lane < 2) {
#pragma unroll
        for (int k = 0; k < MH; ++k) {
            sh[threadIdx.x >> 6][2 * (2 * k + lane)] = re[k];
            sh[threadIdx.x >> 6][2 * (2 * k + lane) + 1] = im[k];
        }
        if (lane == 0) sh[threadIdx.x >> 6][2 * MC] = nrm;
    }
    __syncthreads();
    for (int c = threadIdx.x; c <= m; c += KT) {
        double tr = 0, ti = 0;
        if (c < m) {
            for (int wv = 0; wv < KT / 64; ++wv) { tr += sh[wv][2 * c]; ti += sh[wv][2 * c + 1]; }
        } else {
            for (int wv = 0; wv < KT / 64; ++wv) tr += sh[wv][2 * MC];
        }
        partial[((int64_t)p * ldp + c) * nb + blockIdx.x] = make_double2(tr, ti);
    }
}

// Hessenberg column j = coef1 + coef2, beta_j = ||w||; scale[p] = 1 / beta (or 0).
// The new vector lies in the span of the basis (beta = 0) when the second
// Gram-Schmidt pass removes more than 1 - 0.717 of what the first pass left
// (znaitr's test, steps 4-5: ARPACK then tries a third pass and zeroes the
// residual when that one loses as much).  A residual that survives the second
// pass is kept as the next direction however small it is, as znaitr keeps it:
// when the start vector is an exact eigenvector, that round-off residual is the
// next Krylov direction.
// One 64-lane block per problem: the partial norms are added lane-strided, then
// across the lanes, in a fixed order.
// After the second pass: ||w|| (beta2, kept in b2s) and whether a third pass
// is needed (the second one removed more than 1 - 0.717 of what the first left;
// znaitr step 5 then refines once more).
__global__ __launch_bounds__(64) void kcheck_kernel(const KDesc* __restrict__ desc, int j,
                                                    const double2* __restrict__ coef2, int ldp,
                                                    const double* __restrict__ pnorm, int nblk,
                                                    int* __restrict__ need3, double* __restrict__ b2s,
                                                    const int* __restrict__ deferred) {
    const int p = blockIdx.x;
    if (j < desc[p].kstart) return;
    if (deferred && deferred[p]) {
        if (threadIdx.x == 0) need3[p] = 0;
        return;
    }
    double b2 = 0;
    for (int b = threadIdx.x; b < nblk; b += 64) b2 += pnorm[(int64_t)p * nblk + b];
    for (int off = 32; off > 0; off >>= 1) b2 += __shfl_xor(b2, off, 64);
    if (threadIdx.x != 0) return;
    const double beta2 = sqrt(b2);
    const double beta1 = sqrt(fmax(0.0, coef2[(int64_t)p * ldp + j + 1].x));
    b2s[p] = beta2;
    need3[p] = (beta2 > 0.0 && !(beta2 > 0.717 * beta1)) ? 1 : 0;
}

__global__ __launch_bounds__(64) void kfinish_kernel(const KDesc* __restrict__ desc, int j, int ncv,
                                                     const double2* __restrict__ coef1,
                                                     const double2* __restrict__ coef2,
                                                     const double2* __restrict__ coef3, int ldp,
                                                     const double* __restrict__ pnorm, int nblk,
                                                     const int* __restrict__ need3, const double* __restrict__ b2s,
                                                     double2* __restrict__ H, double* __restrict__ betas,
                                                     double* __restrict__ scale, const int* __restrict__ pend,
                                                     const int* __restrict__ deferred,
                                                     const double2* __restrict__ corr,
                                                     const double2* __restrict__ gin, double2* __restrict__ gout) {
    const int p = blockIdx.x;
    const int kstart = desc[p].kstart;
    if (j < kstart) return;
    const bool third = need3[p] != 0;
    double b2 = 0;
    if (third) {
        for (int b = threadIdx.x; b < nblk; b += 64) b2 += pnorm[(int64_t)p * nblk + b];
        for (int off = 32; off > 0; off >>= 1) b2 += __shfl_xor(b2, off, 64);
    }
    double2* Hp = H + (int64_t)p * (ncv + 1) * ncv;
    __shared__ double2 hcol[64];  // this step's column, for the g term below
    __shared__ double sbeta;
    // the operator of step j ran on v' = v_j + V corr (a deferred update of
    // step j - 1), so A v_j = A v' - V g with g = H corr, g formed by the
    // previous step's kfinish.  One lane per row of the column.
    const bool fixed = pend != nullptr && j - 1 >= kstart && pend[p] != 0;
    for (int r = threadIdx.x; r <= j; r += 64) {
        const double2 a = coef1[(int64_t)p * ldp + r], c = coef2[(int64_t)p * ldp + r];
        double2 h = make_double2(a.x + c.x, a.y + c.y);
        if (third) {
            const double2 e = coef3[(int64_t)p * ldp + r];
            h.x += e.x;
            h.y += e.y;
        }
        if (fixed) {
            const double2 gg = gin[(int64_t)p * ldp + r];
            h.x -= gg.x;
            h.y -= gg.y;
        }
        Hp[r + (int64_t)j * (ncv + 1)] = h;
        hcol[r] = h;
    }
    if (threadIdx.x == 0) {
        // beta = ||w|| after the last pass taken; zero when that pass also removed
        // more than 1 - 0.717 of its input (the residual lies in the basis span)
        double beta = b2s[p];
        if (third) {
            const double beta3 = sqrt(b2);
            beta = (beta3 > 0.717 * beta) ? beta3 : 0.0;
        }
        Hp[(j + 1) + (int64_t)j * (ncv + 1)] = make_double2(beta, 0.0);
        betas[(int64_t)p * ncv + j] = beta;
        scale[p] = beta > 0 ? 1.0 / beta : 0.0;
        sbeta = beta;
    }
    if (deferred == nullptr || deferred[p] == 0) return;
    __syncthreads();
    // this step's update was deferred: g = H(0..j+1, 0..j) corr for step j + 1.
    // Subdiagonal entries from this extension (and the one the restart left at
    // kstart - 1) are in betas, the kept block of a restart in H.
    for (int r = threadIdx.x; r <= j + 1; r += 64) {
        double gx = 0, gy = 0;
        for (int c = (r > 0 ? r - 1 : 0); c <= j; ++c) {
            double2 h;
            if (c == j) h = (r == j + 1) ? make_double2(sbeta, 0.0) : hcol[r];
            else if (r == c + 1 && c >= kstart - 1) h = make_double2(betas[(int64_t)p * ncv + c], 0.0);
            else h = Hp[r + (int64_t)c * (ncv + 1)];
            const double2 q = corr[(int64_t)p * ldp + c];
            gx += h.x * q.x - h.y * q.y;
            gy += h.x * q.y + h.y * q.x;
        }
        gout[(int64_t)p * ldp + r] = make_double2(gx, gy);
    }
}

// the kept Hessenberg block of each restarting problem, in the device copy the
// deferred updates read: slot q = {(p, keep), H(0..keep-1, 0..keep-1)}
__global__ __launch_bounds__(64) void khk_kernel(const double2* __restrict__ src, int slot, int ncv,
                                                 double2* __restrict__ H) {
    const double2* sq = src + (int64_t)blockIdx.x * slot;
    const int p = (int)sq[0].x, keep = (int)sq[0].y;
    double2* Hp = H + (int64_t)p * (ncv + 1) * ncv;
    for (int e = threadIdx.x; e < keep * keep; e += 64) Hp[(e % keep) + (int64_t)(e / keep) * (ncv + 1)] = sq[1 + e];
}

// V(:, j+1) *= scale (the next basis vector), unless it is the residual f
__global__ __launch_bounds__(KT) void kscale_kernel(const KDesc* __restrict__ desc, int j, int ncv,
                                                    const double* __restrict__ scale) {
    const int p = blockIdx.y;
    const KDesc d = desc[p];
    if (j < d.kstart || j + 1 >= ncv) return;
    const double sc = scale[p];
    double2* w = d.V + (int64_t)(j + 1) * d.n;
    for (int64_t i = (int64_t)blockIdx.x * KT + threadIdx.x; i < d.n; i += (int64_t)gridDim.x * KT) {
        double2 v = w[i];
        v.x *= sc;
        v.y *= sc;
        w[i] = v;
    }
}

// The restart of every restarting problem in one pass over its basis:
// W(:, c) = sum_k V(:, k) Q(k, c) for c < nout (= keep + 1), and the new
// residual W(:, ncv) = qm f + hk W(:, keep) with f = V(:, ncv).  Each row's
// basis entries are loaded once into registers; Q (ncv x nout, column-major)
// is read from LDS row by row (the barrier keeps the compiler from holding it
// in registers).  For MV >= ncv and MO >= nout.
struct CombDesc {
    const double2* V;
    double2* W;
    const double2* Q;
    int64_t n;
    int ncv, nout;
    double2 qm, hk;
};
template <int MV, int MO>
__global__ __launch_bounds__(KT) void kcombine_kernel(const CombDesc* __restrict__ desc) {
    const CombDesc d = desc[blockIdx.y];
    __shared__ double2 q[MV * MO];
    for (int e = threadIdx.x; e < d.ncv * d.nout; e += KT) q[(e % d.ncv) + (e / d.ncv) * MV] = d.Q[e];
    __syncthreads();
    const int64_t n = d.n;
    const int keep = d.nout - 1;
    for (int64_t i = (int64_t)blockIdx.x * KT + threadIdx.x; i < n; i += (int64_t)gridDim.x * KT) {
        asm volatile("" ::: "memory");
        double2 a[MV];
#pragma unroll
        for (int k = 0; k < MV; ++k)
            if (k < d.ncv) a[k] = d.V[i + (int64_t)k * n];
        const double2 f = d.V[i + (int64_t)d.ncv * n];
        double2 fo = make_double2(d.qm.x * f.x - d.qm.y * f.y, d.qm.x * f.y + d.qm.y * f.x);
#pragma unroll
        for (int c = 0; c < MO; ++c) {
            if (c < d.nout) {
                double ax = 0, ay = 0;
#pragma unroll
                for (int k = 0; k < MV; ++k) {
                    if (k < d.ncv) {
                        const double2 h = q[k + c * MV];
                        ax += a[k].x * h.x - a[k].y * h.y;
                        ay += a[k].x * h.y + a[k].y * h.x;
                    }
                }
                d.W[i + (int64_t)c * n] = make_double2(ax, ay);
                if (c == keep) {
                    fo.x += d.hk.x * ax - d.hk.y * ay;
                    fo.y += d.hk.x * ay + d.hk.y * ax;
                }
            }
        }
        d.W[i + (int64_t)d.ncv * n] = fo;
    }
}

// operator vectors of a persistent batch
struct IoDesc {
    void* in;
    void* out;
};

// in[p] = V(:, j) (narrowed to complex64 when SINGLE) for every problem that
// takes step j; live[p] tells the operator's launches which problems do
template <bool SINGLE>
__global__ __launch_bounds__(KT) void kgather_kernel(const KDesc* __restrict__ desc, const IoDesc* __restrict__ io,
                                                     int* __restrict__ live, int j) {
    const int p = blockIdx.y;
    const KDesc d = desc[p];
    const bool on = j >= d.kstart;
    if (blockIdx.x == 0 && threadIdx.x == 0) live[p] = on ? 1 : 0;
    if (!on) return;
    const double2* __restrict__ src = d.V + (int64_t)j * d.n;
    for (int64_t i = (int64_t)blockIdx.x * KT + threadIdx.x; i < d.n; i += (int64_t)gridDim.x * KT) {
        const double2 v = src[i];
        if (SINGLE) static_cast<float2*>(io[p].in)[i] = make_float2((float)v.x, (float)v.y);
        else static_cast<double2*>(io[p].in)[i] = v;
    }
}

// V(:, j+1) = out[p] (widened from complex64 when SINGLE)
template <bool SINGLE>
__global__ __launch_bounds__(KT) void kscatter_kernel(const KDesc* __restrict__ desc, const IoDesc* __restrict__ io,
                                                      int j) {
    const int p = blockIdx.y;
    const KDesc d = desc[p];
    if (j < d.kstart) return;
    double2* __restrict__ dst = d.V + (int64_t)(j + 1) * d.n;
    for (int64_t i = (int64_t)blockIdx.x * KT + threadIdx.x; i < d.n; i += (int64_t)gridDim.x * KT) {
        if (SINGLE) {
            const float2 v = static_cast<const float2*>(io[p].out)[i];
            dst[i] = make_double2((double)v.x, (double)v.y);
        } else {
            dst[i] = static_cast<const double2*>(io[p].out)[i];
        }
    }
}

// page-locked host array (stream-ordered copies without a host wait)
template <typename T>
struct Pinned {
    T* ptr = nullptr;
    size_t n = 0;
    Pinned() = default;
    Pinned(Pinned const&) = delete;
    Pinned& operator=(Pinned const&) = delete;
    ~Pinned() {
        if (ptr) (void)hipHostFree(ptr);
    }
    void reserve(size_t count) {
        if (count <= n) return;
        if (ptr) NC_HIP(hipHostFree(ptr));
        ptr = nullptr;
        NC_HIP(hipHostMalloc((void**)&ptr, count * sizeof(T), hipHostMallocDefault));
        n = count;
    }
    T& operator[](size_t i) { return ptr[i]; }
    T const& operator[](size_t i) const { return ptr[i]; }
};

// Instantiated graphs of one solve, destroyed with it.  The stream they run on
// is drained first (an exception may leave a replay in flight), and each
// captured graph lives as long as its executable.
struct GraphCache {
    struct Entry {
        int key;
        hipGraph_t graph;
        hipGraphExec_t exec;
    };
    std::vector<Entry> entries;
    hipStream_t stream = nullptr;
    ~GraphCache() {
        if (entries.empty()) return;
        (void)hipStreamSynchronize(stream);
        for (auto& e : entries) {
            (void)hipGraphExecDestroy(e.exec);
            (void)hipGraphDestroy(e.graph);
        }
    }
    hipGraphExec_t find(int key) const {
        for (auto const& e : entries)
            if (e.key == key) return e.exec;
        return nullptr;
    }
};

int wanted_index(std::vector<zd> const& th, const char* which) {
    int best = 0;
    for (int i = 1; i < (int)th.size(); ++i) {
        if (which[0] == 'S' && which[1] == 'R') {
            if (th[i].real() < th[best].real()) best = i;
        } else if (which[0] == 'L' && which[1] == 'R') {
            if (th[i].real() > th[best].real()) best = i;
        } else {  // LM
            if (std::abs(th[i]) > std::abs(th[best])) best = i;
        }
    }
    return best;
}

// indices of the Ritz values from most to least wanted
std::vector<int> wanted_order(std::vector<zd> const& th, const char* which) {
    std::vector<int> o(th.size());
    for (size_t i = 0; i < o.size(); ++i) o[i] = (int)i;
    std::stable_sort(o.begin(), o.end(), [&](int a, int b) {
        if (which[0] == 'S' && which[1] == 'R') return th[a].real() < th[b].real();
        if (which[0] == 'L' && which[1] == 'R') return th[a].real() > th[b].real();
        return std::abs(th[a]) > std::abs(th[b]);
    });
    return o;
}

// one explicit shifted QR step on the m x m Hessenberg H, accumulating Q
void shifted_qr_step(int m, std::vector<zd>& H, zd mu, std::vector<zd>& Q) {
    auto h = [&](int i, int j) -> zd& { return H[(size_t)i + (size_t)j * m]; };
    auto q = [&](int i, int j) -> zd& { return Q[(size_t)i + (size_t)j * m]; };
    std::vector<double> cs(m);
    std::vector<zd> ss(m);
    for (int k = 0; k < m; ++k) h(k, k) -= mu;
    for (int k = 0; k < m - 1; ++k) {
        zd x = h(k, k), y = h(k + 1, k);
        double c;
        zd s;
        const double ay = std::abs(y), ax = std::abs(x);
        if (ay == 0.0) { c = 1; s = 0; }
        else if (ax == 0.0) { c = 0; s = std::conj(y) / ay; }
        else { double r = std::hypot(ax, ay); c = ax / r; s = (x / ax) * std::conj(y) / r; }
        cs[k] = c; ss[k] = s;
        for (int j = k; j < m; ++j) {
            zd a = h(k, j), b = h(k + 1, j);
            h(k, j) = c * a + s * b;
            h(k + 1, j) = -std::conj(s) * a + c * b;
        }
    }
    for (int k = 0; k < m - 1; ++k) {
        const double c = cs[k];
        const zd s = ss[k];
        for (int i = 0; i <= std::min(k + 1, m - 1); ++i) {
            zd a = h(i, k), b = h(i, k + 1);
            h(i, k) = c * a + std::conj(s) * b;
            h(i, k + 1) = -s * a + c * b;
        }
        for (int i = 0; i < m; ++i) {
            zd a = q(i, k), b = q(i, k + 1);
            q(i, k) = c * a + std::conj(s) * b;
            q(i, k + 1) = -s * a + c * b;
        }
    }
    for (int k = 0; k < m; ++k) h(k, k) += mu;
    for (int j = 0; j < m; ++j)
        for (int i = j + 2; i < m; ++i) h(i, j) = 0.0;
}

// per-problem solver state
struct KState {
    IramTask* task;
    int64_t n = 0;
    ZBuf V, Vk;
    std::vector<zd> H;  // ncv x ncv, host copy
    int k = 0;
    double beta = 0.0;
    bool done = false;
    int iter = 0;       // restarts performed (ARPACK iparam(3))
    int fresh = 0;      // random vectors drawn so far
    // a restart normalised v_k on the device; ||f|| = H(k, k-1) is read back with
    // the Hessenberg columns of the next extension
    bool restart_pending = false;
};

// ARPACK's getv0 after an invariant subspace (znaitr step 1, znaupd): the
// factorization continues from a random vector orthogonal to the first k basis
// vectors, with a zero subdiagonal H(k, k-1).  False when the basis already
// spans the space.
bool continue_with_random(KState& x, int p, int k, hipStream_t s) {
    if (k >= x.n) return false;
    double2* v = x.V.ptr + (size_t)k * x.n;
    const uint64_t key = 0x51ED270B27D3A6C1ull ^ ((uint64_t)(p + 1) * 0x9E3779B97F4A7C15ull) ^
                         ((uint64_t)(++x.fresh) * 0xC2B2AE3D27D4EB4Full);
    fill_random(v, x.n, key, s);
    std::vector<zd> c(k);
    for (int pass = 0; pass < 2; ++pass) {
        zdotc_multi(x.V.ptr, x.n, k, v, x.n, c.data(), s);
        zgemv_sub(x.V.ptr, x.n, k, c.data(), v, x.n, s);
    }
    const double nr = dznrm2(v, x.n, s);
    if (!(nr > 0)) return false;
    zscal(v, zd(1.0 / nr, 0), x.n, s);
    return true;
}

// the Ritz vector of the leading m x m block with coefficients y
void finish_with(KState& st, int m, std::vector<zd> const& y, zd theta, int iter, hipStream_t s) {
    IramTask& t = *st.task;
    zgemv_n(st.V.ptr, st.n, m, y.data(), t.result, st.n, s);
    const double rn = dznrm2(t.result, st.n, s);
    zscal(t.result, zd(1.0 / rn, 0), st.n, s);
    t.out.eigenvalue = theta;
    t.out.restarts = iter;
    t.out.info = 0;
    st.done = true;
}

}  // namespace

void iram_batch(std::vector<IramTask>& tasks, BatchApply const& apply, hipStream_t s) {
    const int P = (int)tasks.size();
    if (P == 0) return;
    const int ncv = tasks[0].ncv;
    const double eps = std::numeric_limits<double>::epsilon();
    const double eps23 = std::pow(eps, 2.0 / 3.0);
    if (ncv + 2 > 64) throw Error("iram: ncv above 62 is not supported", -8);
    std::vector<KState> st(P);
    int64_t nmax = 1;
    for (int p = 0; p < P; ++p) {
        IramTask& t = tasks[p];
        if (t.ncv != ncv) throw Error("iram_batch: every problem of a batch needs the same ncv", -8);
        t.out = IramResult();
        KState& x = st[p];
        x.task = &t;
        x.n = t.op->n;
        if (ncv > x.n || ncv < 2 || x.n < 1) { t.out.info = -3; x.done = true; continue; }
        if (t.max_restarts < 1) t.max_restarts = 1;
        x.V.resize((size_t)x.n * (ncv + 1));
        x.H.assign((size_t)ncv * ncv, zd(0, 0));
        d2d(x.V.ptr, t.guess, (size_t)x.n * sizeof(double2));
        const double nrm = dznrm2(x.V.ptr, x.n, s);
        if (!(nrm > 0) || !std::isfinite(nrm)) { t.out.info = -9; x.done = true; continue; }
        zscal(x.V.ptr, zd(1.0 / nrm, 0), x.n, s);
        nmax = std::max(nmax, x.n);
    }
    // reduction geometry: about 2048 blocks over the batch, deterministic per (P, nmax)
    const int nblk = (int)std::max<int64_t>(1, std::min<int64_t>((nmax + 4 * KT - 1) / (4 * KT),
                                                                   std::max(1, 2048 / P)));
    const int ldp = ncv + 2;
    ZBuf partial((size_t)P * ldp * nblk), coef1((size_t)P * ldp), coef2((size_t)P * ldp), coef3((size_t)P * ldp);
    DevBuf<int> need3(P);
    DevBuf<double> b2s(P);
    NormSpace nspace;
    nspace.reserve(2 * P);
    // Hessenberg columns and subdiagonals of all problems in one buffer, read
    // back with one copy per extension
    const size_t nH = (size_t)P * (ncv + 1) * ncv, nB = (size_t)P * ncv;
    ZBuf HB(nH + (nB + 1) / 2);
    double2* const Hd_ptr = HB.ptr;
    double* const betas_ptr = reinterpret_cast<double*>(HB.ptr + nH);
    DevBuf<double> pnorm((size_t)P * nblk), scale(P);
    DevBuf<KDesc> ddesc(P);
    ZBuf coefs;  // Ritz / restart coefficients of one round, uploaded together
    Pinned<KDesc> hdesc;
    hdesc.reserve(P);
    Pinned<double2> hHB;
    hHB.reserve(HB.n);
    const double2* const Hh = hHB.ptr;
    const double* const bh = reinterpret_cast<const double*>(hHB.ptr + nH);
    Pinned<double2> hcoef;
    // persistent operator: fixed in/out vectors, launches prepared once
    const bool pers = apply.persistent();
    DevBuf<int> live;
    DevBuf<IoDesc> iod;
    GraphCache graphs;
    if (pers) {
        if ((int)apply.in.size() != P || (int)apply.out.size() != P)
            throw Error("iram_batch: persistent operator needs one input and one output per problem", -8);
        std::vector<IoDesc> hio(P);
        for (int p = 0; p < P; ++p) hio[p] = IoDesc{apply.in[p], apply.out[p]};
        iod.resize(P);
        h2d(iod.ptr, hio.data(), (size_t)P * sizeof(IoDesc));
        std::vector<int> zeros(P, 0);
        live.resize(P);
        h2d(live.ptr, zeros.data(), (size_t)P * sizeof(int));
        apply.bind(live.ptr, s);
    }
    const dim3 g(nblk, P), gr(ldp, P);
    // one Arnoldi step of every problem active at step j, persistent form
    auto step = [&](int j) {
        if (apply.single) hipLaunchKernelGGL(kgather_kernel<true>, g, dim3(KT), 0, s, ddesc.ptr, iod.ptr, live.ptr, j);
        else hipLaunchKernelGGL(kgather_kernel<false>, g, dim3(KT), 0, s, ddesc.ptr, iod.ptr, live.ptr, j);
        NC_HIP(hipGetLastError());
        apply.launch(s);
        if (apply.single) hipLaunchKernelGGL(kscatter_kernel<true>, g, dim3(KT), 0, s, ddesc.ptr, iod.ptr, j);
        else hipLaunchKernelGGL(kscatter_kernel<false>, g, dim3(KT), 0, s, ddesc.ptr, iod.ptr, j);
        NC_HIP(hipGetLastError());
    };
    // the Gram-Schmidt kernels of step j (classical Gram-Schmidt twice; the
    // first update and the second pass's dots share one read of the basis)
    // the dots of step j against columns 0..j+1, in one pass over w where the
    // column count allows it
    auto dots = [&](int j, const int* gate, dim3 gd) {
        const int m = j + 2;
        if (m <= 8) hipLaunchKernelGGL(kdot1_kernel<8>, gd, dim3(KT), 0, s, ddesc.ptr, j, partial.ptr, ldp, gate);
        else if (m <= 12) hipLaunchKernelGGL(kdot1_kernel<12>, gd, dim3(KT), 0, s, ddesc.ptr, j, partial.ptr, ldp, gate);
        else if (m <= 18) hipLaunchKernelGGL(kdot1_kernel<18>, gd, dim3(KT), 0, s, ddesc.ptr, j, partial.ptr, ldp, gate);
        else hipLaunchKernelGGL(kdot_kernel, gd, dim3(KT), 0, s, ddesc.ptr, j, partial.ptr, ldp, gate);
    };
    // znaitr's third pass runs for few steps, but its gated launches are made
    // at every step: a grid of at most 256 blocks per problem keeps the
    // launches that find no problem to refine short
    const int nblk3 = std::min(nblk, 256);
    const dim3 g3(nblk3, P);
    // its dots, when they run, read the whole basis: four blocks per CU (at one
    // block per CU they streamed at about 2.2 TB/s)
    const int nblk3d = std::min(nblk, 1024);
    const dim3 g3d(nblk3d, P);
    const bool local_first = gs_local_first();
    const bool defer = gs_defer_second();
    // deferred second updates: flags per step parity, the correction
    // coefficients, and the Hessenberg term g per step parity
    DevBuf<int> dflag((size_t)2 * P);
    ZBuf corr((size_t)P * ldp), gbuf((size_t)2 * P * ldp);
    dzero(dflag.ptr, dflag.bytes());
    auto orthogonalize = [&](int j) {
        const int* ng = nullptr;
        int* const fcur = dflag.ptr + (size_t)(j & 1) * P;           // step j's deferral
        const int* const fprev = dflag.ptr + (size_t)((j + 1) & 1) * P;  // step j - 1's
        // step j + 1 exists and its fused kernel (at most 24 columns) applies the correction
        const int allow = (defer && j + 1 < ncv && j + 2 <= 24) ? 1 : 0;
        if (local_first) hipLaunchKernelGGL(kdotl_kernel, g, dim3(KT), 0, s, ddesc.ptr, j, partial.ptr, ldp);
        else dots(j, ng, g);
        hipLaunchKernelGGL(kreduce_kernel, gr, dim3(KT), 0, s, ddesc.ptr, j, partial.ptr, ldp, nblk, coef1.ptr, ng);
        const int m = j + 1;
        if (m <= 4)
            hipLaunchKernelGGL(kupdot_kernel<4>, g, dim3(KT), 0, s, ddesc.ptr, j, coef1.ptr, ldp, partial.ptr, fprev,
                               (const double2*)corr.ptr);
        else if (m <= 8)
            hipLaunchKernelGGL(kupdot_kernel<8>, g, dim3(KT), 0, s, ddesc.ptr, j, coef1.ptr, ldp, partial.ptr, fprev,
                               (const double2*)corr.ptr);
        else if (m <= 16)
            hipLaunchKernelGGL(kupdot_pair_kernel<8>, g, dim3(KT), 0, s, ddesc.ptr, j, coef1.ptr, ldp, partial.ptr,
                               fprev, (const double2*)corr.ptr);
        else if (m <= 24)
            hipLaunchKernelGGL(kupdot_pair_kernel<12>, g, dim3(KT), 0, s, ddesc.ptr, j, coef1.ptr, ldp, partial.ptr,
                               fprev, (const double2*)corr.ptr);
        else {
            hipLaunchKernelGGL(kupdate_kernel, g, dim3(KT), 0, s, ddesc.ptr, j, coef1.ptr, ldp, pnorm.ptr, ng);
            dots(j, ng, g);
        }
        hipLaunchKernelGGL(kreduce_kernel, gr, dim3(KT), 0, s, ddesc.ptr, j, partial.ptr, ldp, nblk, coef2.ptr, ng);
        hipLaunchKernelGGL(kupdate2_kernel, g, dim3(KT), 0, s, ddesc.ptr, j, coef2.ptr, ldp, pnorm.ptr, allow, fcur,
                           corr.ptr, b2s.ptr);
        // znaitr's third pass, only for the problems whose second pass lost too much
        hipLaunchKernelGGL(kcheck_kernel, dim3(P), dim3(64), 0, s, ddesc.ptr, j, coef2.ptr, ldp, pnorm.ptr, nblk,
                           need3.ptr, b2s.ptr, (const int*)fcur);
        dots(j, need3.ptr, g3d);
        hipLaunchKernelGGL(kreduce_kernel, gr, dim3(KT), 0, s, ddesc.ptr, j, partial.ptr, ldp, nblk3d, coef3.ptr,
                           need3.ptr);
        hipLaunchKernelGGL(kupdate_kernel, g3, dim3(KT), 0, s, ddesc.ptr, j, coef3.ptr, ldp, pnorm.ptr, need3.ptr);
        hipLaunchKernelGGL(kfinish_kernel, dim3(P), dim3(64), 0, s, ddesc.ptr, j, ncv, coef1.ptr, coef2.ptr,
                           coef3.ptr, ldp, pnorm.ptr, nblk3, need3.ptr, b2s.ptr, Hd_ptr, betas_ptr, scale.ptr, fprev,
                           (const int*)fcur, (const double2*)corr.ptr,
                           (const double2*)(gbuf.ptr + (size_t)(j & 1) * P * ldp),
                           gbuf.ptr + (size_t)((j + 1) & 1) * P * ldp);
        hipLaunchKernelGGL(kscale_kernel, g, dim3(KT), 0, s, ddesc.ptr, j, ncv, scale.ptr);
        NC_HIP(hipGetLastError());
    };

    for (;;) {
        int jmin = ncv;
        for (int p = 0; p < P; ++p) {
            hdesc[p].V = st[p].V.ptr;
            hdesc[p].n = st[p].n;
            hdesc[p].kstart = st[p].done ? ncv : st[p].k;
            hdesc[p]._pad = 0;
            if (!st[p].done) jmin = std::min(jmin, st[p].k);
        }
        if (jmin >= ncv) break;  // every problem finished
        // stream-ordered: the host array is rewritten only after the next read-back
        NC_HIP(hipMemcpyAsync(ddesc.ptr, hdesc.ptr, (size_t)P * sizeof(KDesc), hipMemcpyHostToDevice, s));
        // ---- extend every active factorization from its k to ncv
        for (int j = jmin; j < ncv; ++j)
            for (int p = 0; p < P; ++p)
                if (!st[p].done && st[p].k <= j) ++tasks[p].out.matvecs;
        if (pers) {
            // The extension from jmin depends only on device data (the problem
            // table above, the bases, the prepared launches), so it is recorded
            // once per jmin and replayed on later restarts.
            auto extend = [&]() {
                for (int j = jmin; j < ncv; ++j) {
                    step(j);
                    orthogonalize(j);
                }
            };
            if (krylov_graphs()) {
                hipGraphExec_t ge = graphs.find(jmin);
                if (!ge) {
                    hipGraph_t gph = nullptr;
                    NC_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                    try {
                        extend();
                    } catch (...) {
                        (void)hipStreamEndCapture(s, &gph);
                        if (gph) (void)hipGraphDestroy(gph);
                        throw;
                    }
                    NC_HIP(hipStreamEndCapture(s, &gph));
                    const hipError_t e = hipGraphInstantiate(&ge, gph, nullptr, nullptr, 0);
                    if (e != hipSuccess) (void)hipGraphDestroy(gph);
                    NC_HIP(e);
                    graphs.stream = s;
                    graphs.entries.push_back(GraphCache::Entry{jmin, gph, ge});
                }
                NC_HIP(hipGraphLaunch(ge, s));
            } else {
                extend();
            }
        } else {
            std::vector<int> who;
            std::vector<const double2*> xs;
            std::vector<double2*> ys;
            for (int j = jmin; j < ncv; ++j) {
                who.clear(); xs.clear(); ys.clear();
                for (int p = 0; p < P; ++p) {
                    if (st[p].done || st[p].k > j) continue;
                    who.push_back(p);
                    xs.push_back(st[p].V.ptr + (size_t)j * st[p].n);
                    ys.push_back(st[p].V.ptr + (size_t)(j + 1) * st[p].n);
                }
                apply(who, xs, ys, s);
                orthogonalize(j);
            }
        }
        NC_HIP(hipMemcpyAsync(hHB.ptr, HB.ptr, HB.n * sizeof(double2), hipMemcpyDeviceToHost, s));
        NC_HIP(hipStreamSynchronize(s));
        // ---- per problem (host): Ritz values, convergence, restart shifts.  The
        // device work of the finishing and restarting problems is then done for
        // all of them at once: one grouped GEMM, batched vector updates, one
        // host read of their norms.
        struct Fin { int p, m, iter; std::vector<zd> y; zd theta; };
        struct Rst { int p, keep; std::vector<zd> Q; zd hk, qm; std::vector<zd> Hk; };
        // Per problem, on host threads (no device work): merge the new
        // Hessenberg columns, Ritz values, convergence test, restart shifts.
        enum { D_NONE, D_STOP, D_FAIL, D_FIN, D_RST, D_RSTZERO };
        struct Dec {
            int kind = D_NONE;
            int stop = -1;
            Fin f;
            Rst r;
        };
        std::vector<Dec> dec(P);
        host_parallel_for(P, [&](int p) {
            KState& x = st[p];
            if (x.done) return;
            Dec& D = dec[p];
            IramTask& t = *x.task;
            const int iter = ++x.iter;
            auto h = [&](int i, int jj) -> zd& { return x.H[(size_t)i + (size_t)jj * ncv]; };
            const double2* Hp = Hh + (size_t)p * (ncv + 1) * ncv;
            if (x.restart_pending) {
                // the residual norm of the last restart, normalised on the device
                x.restart_pending = false;
                const double rb = bh[(size_t)p * ncv + x.k - 1];
                if (rb == 0.0) {
                    // the restart landed on an invariant subspace: the extension
                    // just run started from a zero vector and is dropped
                    --x.iter;
                    D.kind = D_RSTZERO;
                    return;
                }
                h(x.k, x.k - 1) = rb;
            }
            int stop = -1;  // first step whose new vector vanished (invariant subspace)
            for (int c = x.k; c < ncv; ++c) {
                for (int r = 0; r <= c; ++r) h(r, c) = zd(Hp[r + (size_t)c * (ncv + 1)].x, Hp[r + (size_t)c * (ncv + 1)].y);
                const double bc = bh[(size_t)p * ncv + c];
                if (c + 1 < ncv) h(c + 1, c) = bc;
                if (bc == 0.0 && c + 1 < ncv && stop < 0) stop = c;
                if (stop >= 0) break;
            }
            if (stop >= 0) {
                D.kind = D_STOP;
                D.stop = stop;
                return;
            }
            x.beta = bh[(size_t)p * ncv + ncv - 1];
            std::vector<zd> th, Y;
            hessenberg_eig(ncv, x.H, th, Y);
            const int wi = wanted_index(th, t.which);
            const double est = x.beta * std::abs(Y[(size_t)(ncv - 1) + (size_t)wi * ncv]);
            const double tol_eff = t.tol > 0 ? t.tol : eps;
            const bool converged = est <= tol_eff * std::max(eps23, std::abs(th[wi]));
            if (converged || iter >= t.max_restarts) {
                if (!converged) {
                    D.kind = D_FAIL;
                    return;
                }
                D.kind = D_FIN;
                D.f = Fin{p, ncv, iter, std::vector<zd>(Y.begin() + (size_t)wi * ncv, Y.begin() + (size_t)(wi + 1) * ncv),
                          th[wi]};
                return;
            }
            // ---- implicit restart.  ARPACK keeps nev (= 1) vectors and applies
            // the ncv - 1 unwanted Ritz values as exact shifts; a larger basis
            // keeps half of its vectors (thick restart) and shifts out the rest.
            D.kind = D_RST;
            Rst& r = D.r;
            r.p = p;
            r.keep = ncv <= 3 ? 1 : std::max(1, std::min(ncv - 2, (ncv * thick_keep_percent() + 50) / 100));
            const int keep = r.keep;
            std::vector<int> order = wanted_order(th, t.which);
            r.Q.assign((size_t)ncv * ncv, zd(0, 0));
            for (int i = 0; i < ncv; ++i) r.Q[(size_t)i + (size_t)i * ncv] = 1.0;
            for (int q = keep; q < ncv; ++q) shifted_qr_step(ncv, x.H, th[order[q]], r.Q);
            r.hk = h(keep, keep - 1);
            r.qm = r.Q[(size_t)(ncv - 1) + (size_t)(keep - 1) * ncv];
            r.Hk.assign((size_t)keep * keep, zd(0, 0));
            for (int c = 0; c < keep; ++c)
                for (int rr = 0; rr < keep; ++rr) r.Hk[(size_t)rr + (size_t)c * keep] = (rr <= c + 1) ? h(rr, c) : zd(0, 0);
        });
        std::vector<Fin> fins;
        std::vector<Rst> rsts;
        for (int p = 0; p < P; ++p) {
            Dec& D = dec[p];
            KState& x = st[p];
            IramTask& t = *x.task;
            if (D.kind == D_FIN) {
                fins.push_back(std::move(D.f));
            } else if (D.kind == D_RST) {
                rsts.push_back(std::move(D.r));
            } else if (D.kind == D_FAIL) {
                t.out.restarts = x.iter;
                t.out.info = -14;
                x.done = true;
            } else if (D.kind == D_RSTZERO) {
                auto h = [&](int i, int jj) -> zd& { return x.H[(size_t)i + (size_t)jj * ncv]; };
                const int keep = x.k;
                for (int c = keep; c < ncv; ++c)
                    for (int r = 0; r < ncv; ++r) h(r, c) = 0.0;
                t.out.matvecs -= ncv - keep;
                // continue from a random vector orthogonal to the kept ones
                // (H(keep, keep - 1) = 0), as ARPACK's getv0 does
                if (continue_with_random(x, p, keep, s)) {
                    h(keep, keep - 1) = 0.0;
                    continue;
                }
                // the kept vectors span the whole space: their Ritz pairs are exact
                std::vector<zd> Hk((size_t)keep * keep), th2, Y2;
                for (int c = 0; c < keep; ++c)
                    for (int r = 0; r < keep; ++r) Hk[(size_t)r + (size_t)c * keep] = h(r, c);
                hessenberg_eig(keep, Hk, th2, Y2);
                const int w2 = wanted_index(th2, t.which);
                std::vector<zd> yv(Y2.begin() + (size_t)w2 * keep, Y2.begin() + (size_t)(w2 + 1) * keep);
                finish_with(x, keep, yv, th2[w2], x.iter, s);
            } else if (D.kind == D_STOP) {
                auto h = [&](int i, int jj) -> zd& { return x.H[(size_t)i + (size_t)jj * ncv]; };
                const int stop = D.stop;
                // invariant subspace found early: continue from a random vector
                // orthogonal to it, as ARPACK does (the extension does not count
                // as a restart)
                if (continue_with_random(x, p, stop + 1, s)) {
                    h(stop + 1, stop) = 0.0;
                    for (int c = stop + 1; c < ncv; ++c)
                        for (int r = 0; r < ncv; ++r) h(r, c) = 0.0;
                    x.k = stop + 1;
                    --x.iter;
                    continue;
                }
                // the basis spans the whole space: its Ritz values are exact
                const int m = stop + 1;
                std::vector<zd> Hs((size_t)m * m);
                for (int c = 0; c < m; ++c)
                    for (int r = 0; r < m; ++r) Hs[(size_t)r + (size_t)c * m] = h(r, c);
                std::vector<zd> th, Y;
                hessenberg_eig(m, Hs, th, Y);
                const int wi = wanted_index(th, t.which);
                fins.push_back(Fin{p, m, x.iter, std::vector<zd>(Y.begin() + (size_t)wi * m,
                                                                  Y.begin() + (size_t)(wi + 1) * m), th[wi]});
            }
        }
        if (fins.empty() && rsts.empty()) continue;
        // coefficients of all of them in one upload
        std::vector<double2> hc;
        std::vector<size_t> off_f(fins.size()), off_r(rsts.size());
        for (size_t q = 0; q < fins.size(); ++q) {
            off_f[q] = hc.size();
            for (zd v : fins[q].y) hc.push_back(make_double2(v.real(), v.imag()));
        }
        for (size_t q = 0; q < rsts.size(); ++q) {
            off_r[q] = hc.size();
            for (size_t e = 0; e < (size_t)ncv * (rsts[q].keep + 1); ++e)
                hc.push_back(make_double2(rsts[q].Q[e].real(), rsts[q].Q[e].imag()));
        }
        // the kept Hessenberg blocks, for the device copy that deferred updates read
        const int hk_slot = 1 + ncv * ncv;
        const size_t off_hk = hc.size();
        if (defer) {
            for (auto const& r : rsts) {
                hc.push_back(make_double2((double)r.p, (double)r.keep));
                for (int e = 0; e < ncv * ncv; ++e)
                    hc.push_back(e < r.keep * r.keep ? make_double2(r.Hk[e].real(), r.Hk[e].imag())
                                                     : make_double2(0.0, 0.0));
            }
        }
        if (coefs.n < hc.size()) coefs.resize(hc.size());
        // stream-ordered upload; the host array is rewritten only after a later
        // host wait (the norms below)
        hcoef.reserve(hc.size());
        std::copy(hc.begin(), hc.end(), hcoef.ptr);
        NC_HIP(hipMemcpyAsync(coefs.ptr, hcoef.ptr, hc.size() * sizeof(double2), hipMemcpyHostToDevice, s));
        if (defer && !rsts.empty()) {
            hipLaunchKernelGGL(khk_kernel, dim3((unsigned)rsts.size()), dim3(64), 0, s,
                               (const double2*)(coefs.ptr + off_hk), hk_slot, ncv, Hd_ptr);
            NC_HIP(hipGetLastError());
        }
        // result = V y (finishing); W(:, 0:keep) = V Q(:, 0:keep) (restarting)
        // and f <- hk V q_keep + qm f, written to W(:, ncv); then W becomes the
        // basis.  Up to 24 basis vectors and 18 kept ones the restarts take one
        // pass over each basis (kcombine_kernel), otherwise a GEMM and two axpbys.
        int max_keep = 0;
        for (auto const& r : rsts) max_keep = std::max(max_keep, r.keep);
        // kept vectors of the variant for ncv
        const int comb_mo = ncv <= 4 ? 3 : (ncv <= 8 ? 7 : (ncv <= 16 ? 13 : 19));
        const bool fused = ncv <= 24 && max_keep + 1 <= comb_mo;
        GemmPlan gp;
        gp.M = gp.N = 1;
        for (size_t q = 0; q < fins.size(); ++q) {
            KState& x = st[fins[q].p];
            gp.begin_batch(x.task->result, x.n, (int)x.n, 1);
            gp.add_seg(x.V.ptr, x.n, coefs.ptr + off_f[q], fins[q].m, fins[q].m);
        }
        std::vector<CombDesc> cds;
        int64_t cmax = 1;
        for (size_t q = 0; q < rsts.size(); ++q) {
            KState& x = st[rsts[q].p];
            if (x.Vk.n < (size_t)x.n * (ncv + 1)) x.Vk.resize((size_t)x.n * (ncv + 1));
            if (fused) {
                const zd qm = rsts[q].qm, hk = rsts[q].hk;
                cds.push_back(CombDesc{x.V.ptr, x.Vk.ptr, coefs.ptr + off_r[q], x.n, ncv, rsts[q].keep + 1,
                                       make_double2(qm.real(), qm.imag()), make_double2(hk.real(), hk.imag())});
                cmax = std::max(cmax, x.n);
            } else {
                gp.begin_batch(x.Vk.ptr, x.n, (int)x.n, rsts[q].keep + 1);
                gp.add_seg(x.V.ptr, x.n, coefs.ptr + off_r[q], ncv, ncv);
            }
        }
        if (!gp.batches.empty()) zgemm_run(gp, s);
        if (!cds.empty()) {
            int slot;
            auto* dd = (const CombDesc*)stage_descriptors(cds.data(), cds.size() * sizeof(CombDesc), s, &slot);
            const dim3 gc((unsigned)std::max<int64_t>(1, std::min<int64_t>((cmax + KT - 1) / KT,
                                                                            std::max(1, 4096 / (int)cds.size()))),
                          (unsigned)cds.size());
            if (ncv <= 4)
                hipLaunchKernelGGL((kcombine_kernel<4, 3>), gc, dim3(KT), 0, s, dd);
            else if (ncv <= 8)
                hipLaunchKernelGGL((kcombine_kernel<8, 7>), gc, dim3(KT), 0, s, dd);
            else if (ncv <= 16)
                hipLaunchKernelGGL((kcombine_kernel<16, 13>), gc, dim3(KT), 0, s, dd);
            else
                hipLaunchKernelGGL((kcombine_kernel<24, 19>), gc, dim3(KT), 0, s, dd);
            NC_HIP(hipGetLastError());
            release_descriptors(slot, s);
            for (auto const& r : rsts) std::swap(st[r.p].V, st[r.p].Vk);
        } else {
            std::vector<double2*> y;
            std::vector<const double2*> xv;
            std::vector<zd> a, b;
            std::vector<int64_t> nn;
            for (auto const& r : rsts) {
                KState& x = st[r.p];
                y.push_back(x.Vk.ptr + (size_t)x.n * ncv);
                xv.push_back(x.V.ptr + (size_t)x.n * ncv);
                a.push_back(r.qm);
                b.push_back(zd(0, 0));
                nn.push_back(x.n);
            }
            axpby_batch(y, a, xv, b, nn, s);
            xv.clear(); a.clear(); b.clear();
            for (auto const& r : rsts) {
                KState& x = st[r.p];
                xv.push_back(x.Vk.ptr + (size_t)x.n * r.keep);
                a.push_back(r.hk);
                b.push_back(zd(1, 0));
            }
            axpby_batch(y, a, xv, b, nn, s);
            for (auto const& r : rsts) std::swap(st[r.p].V, st[r.p].Vk);
        }
        // Normalised on the device, with no host read: the results; for every
        // restart v_keep = f / ||f|| (||f|| = H(keep, keep - 1) lands in the
        // Hessenberg read-back buffer and is taken up with the next extension's
        // columns); with keep = 1 also v_0 (kept exactly unit norm against
        // round-off drift).  The host goes straight on to the next extension.
        std::vector<NormItem> ni;
        for (auto const& f : fins) {
            KState& x = st[f.p];
            ni.push_back(NormItem{x.task->result, x.task->result, x.n, nullptr});
            x.task->out.eigenvalue = f.theta;
            x.task->out.restarts = f.iter;
            x.task->out.info = 0;
            x.done = true;
        }
        for (auto const& r : rsts) {
            KState& x = st[r.p];
            const int keep = r.keep;
            auto h = [&](int i, int jj) -> zd& { return x.H[(size_t)i + (size_t)jj * ncv]; };
            std::fill(x.H.begin(), x.H.end(), zd(0, 0));
            for (int c = 0; c < keep; ++c)
                for (int rr = 0; rr < keep; ++rr) h(rr, c) = r.Hk[(size_t)rr + (size_t)c * keep];
            x.k = keep;
            x.restart_pending = true;
            ni.push_back(NormItem{x.V.ptr + (size_t)x.n * ncv, x.V.ptr + (size_t)x.n * keep, x.n,
                                  betas_ptr + (size_t)r.p * ncv + keep - 1});
            if (keep == 1) ni.push_back(NormItem{x.V.ptr, x.V.ptr, x.n, nullptr});
        }
        normalize_on_device(ni, nspace, s);
    }
}

IramResult iram(Operator& op, int ncv, const char* which, double tol, int max_restarts,
                const double2* guess, double2* result, hipStream_t s) {
    std::vector<IramTask> t(1);
    t[0].op = &op;
    t[0].ncv = ncv;
    t[0].which = which;
    t[0].tol = tol;
    t[0].max_restarts = max_restarts;
    t[0].guess = guess;
    t[0].result = result;
    iram_batch(t, BatchApply::each(t), s);
    return t[0].out;
}

bool& gs_local_first() {
    // NC_GS_LOCAL=0 in the environment makes the first pass a full one
    static bool on = [] {
        const char* e = std::getenv("NC_GS_LOCAL");
        return !(e && e[0] == '0');
    }();
    return on;
}

int& thick_keep_percent() {
    // NC_THICK_KEEP=<percent> in the environment keeps that share of the basis
    // at a thick restart (50 unless set)
    static int pct = [] {
        const char* e = std::getenv("NC_THICK_KEEP");
        const int v = e ? std::atoi(e) : 50;
        return (v >= 10 && v <= 90) ? v : 50;
    }();
    return pct;
}

bool& gs_defer_second() {
    // NC_GS_DEFER=0 in the environment runs every second update in its own step
    static bool on = [] {
        const char* e = std::getenv("NC_GS_DEFER");
        return !(e && e[0] == '0');
    }();
    return on;
}

bool& krylov_graphs() {
    // NC_KRYLOV_GRAPHS=0 in the environment sets the default to direct launches
    static bool on = [] {
        const char* e = std::getenv("NC_KRYLOV_GRAPHS");
        return !(e && e[0] == '0');
    }();
    return on;
}

BatchApply BatchApply::each(std::vector<IramTask> const& tasks) {
    BatchApply a;
    a.fn = [&tasks](std::vector<int> const& who, std::vector<const double2*> const& xs,
                    std::vector<double2*> const& ys, hipStream_t s) {
        for (size_t q = 0; q < who.size(); ++q) tasks[who[q]].op->apply(xs[q], ys[q], s);
    };
    return a;
}

}  // namespace nc
