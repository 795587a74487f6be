// gfx950 device kernels of the variational-MPS core (FP64 complex path).
//
//  (the complex GEMM lives in gemm.hip)
//  * mix: the sparse MPO block mixing Q(., b, dst) = sum W_n(a,b) P(., a, src).
//  * vector kernels: reductions, axpy and friends for the Arnoldi solver and
//    the Householder projectors.
#include "nc_internal.hpp"

#include <algorithm>
#include <cstring>

namespace nc {

typedef double f64x4 __attribute__((ext_vector_type(4)));

// ----------------------------------------------------------------------------
// Mix kernel.  Blocks are grouped by destination (CSR); each thread owns one
// element e of the slab and produces all d outputs of one destination.
// ----------------------------------------------------------------------------
struct MixDesc {
    int d, ndst;
    int64_t rs;
};

// One thread, one slab element e, outputs b0 .. b0 + NB - 1 of destination dst
// (those below d).  D > 0 fixes the physical dimension at compile time.
template <typename CT, int D, int NB>
__device__ __forceinline__ void mix_element(const CT* __restrict__ P, CT* __restrict__ Q,
                                            const int* __restrict__ dst_ptr, const int* __restrict__ blk_src,
                                            const CT* __restrict__ W, int d_rt, int64_t rs, int64_t e, int dst,
                                            int b0) {
    const int d = (D > 0) ? D : d_rt;
    CT q[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) { q[b].x = 0; q[b].y = 0; }
    for (int t = dst_ptr[dst]; t < dst_ptr[dst + 1]; ++t) {
        const int src = blk_src[t];
        const CT* Wn = W + (int64_t)t * d * d;
        for (int a = 0; a < d; ++a) {
            const CT p = P[((int64_t)src * d + a) * rs + e];
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                if (b0 + b < d) {
                    const CT w = Wn[a + (b0 + b) * d];
                    q[b].x += w.x * p.x - w.y * p.y;
                    q[b].y += w.x * p.y + w.y * p.x;
                }
            }
        }
    }
#pragma unroll
    for (int b = 0; b < NB; ++b)
        if (b0 + b < d) Q[((int64_t)dst * d + b0 + b) * rs + e] = q[b];
}

// D = 0: any physical dimension, four outputs per pass over the blocks
template <typename CT, int D>
__global__ __launch_bounds__(256) void mix_kernel(const CT* __restrict__ P, CT* __restrict__ Q,
                                                  const int* __restrict__ dst_ptr,
                                                  const int* __restrict__ blk_src,
                                                  const CT* __restrict__ W, MixDesc md) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int dst = blockIdx.y;
    if (e >= md.rs) return;
    if (D > 0) {
        mix_element<CT, D, (D > 0 ? D : 1)>(P, Q, dst_ptr, blk_src, W, D, md.rs, e, dst, 0);
    } else {
        for (int b0 = 0; b0 < md.d; b0 += 4) mix_element<CT, 0, 4>(P, Q, dst_ptr, blk_src, W, md.d, md.rs, e, dst, b0);
    }
}

// The same mix for every destination at once: one thread, one slab element e,
// each source block's value P(e, a, src) loaded once (up to MAXF sources) and
// every Q(e, b, dst) formed from registers, instead of one grid row per
// destination reading the sources of its own blocks.
template <typename CT, int D, int MAXF, int MAXT>
__global__ __launch_bounds__(256) void mix_all_kernel(const CT* __restrict__ P, CT* __restrict__ Q,
                                                      const int* __restrict__ dst_ptr,
                                                      const int* __restrict__ blk_src, const CT* __restrict__ W,
                                                      int nsrc, int ndst, int64_t rs) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= rs) return;
    CT v[MAXF][D];
#pragma unroll
    for (int f = 0; f < MAXF; ++f)
#pragma unroll
        for (int a = 0; a < D; ++a) {
            CT x;
            x.x = 0; x.y = 0;
            if (f < nsrc) x = P[((int64_t)f * D + a) * rs + e];
            v[f][a] = x;
        }
#pragma unroll
    for (int t = 0; t < MAXT; ++t) {
        if (t >= ndst) break;
        CT q[D];
#pragma unroll
        for (int b = 0; b < D; ++b) { q[b].x = 0; q[b].y = 0; }
        for (int j = dst_ptr[t]; j < dst_ptr[t + 1]; ++j) {
            const int f = blk_src[j];
            const CT* Wn = W + (int64_t)j * D * D;
#pragma unroll
            for (int a = 0; a < D; ++a) {
                CT p = v[0][a];
#pragma unroll
                for (int r = 1; r < MAXF; ++r)
                    if (f == r) p = v[r][a];
#pragma unroll
                for (int b = 0; b < D; ++b) {
                    const CT w = Wn[a + b * D];
                    q[b].x += w.x * p.x - w.y * p.y;
                    q[b].y += w.x * p.y + w.y * p.x;
                }
            }
        }
#pragma unroll
        for (int b = 0; b < D; ++b) Q[((int64_t)t * D + b) * rs + e] = q[b];
    }
}

template <typename CT>
static void mix_blocks_t(MpoSite const& op, bool src_is_from, const CT* P, CT* Q, int64_t rs, const CT* W,
                         hipStream_t s) {
    if (op.d < 1) throw Error("mix_blocks: physical dimension below 1", -4);
    const int dir = src_is_from ? 0 : 1;
    const int ndst = (int)(src_is_from ? op.tos.size() : op.froms.size());
    if (ndst == 0 || rs == 0) return;
    MixDesc md{op.d, ndst, rs};
    dim3 grid((unsigned)((rs + 255) / 256), (unsigned)ndst);
    const int* ptr = op.mix_ptr[dir].ptr;
    const int* srcs = op.mix_src[dir].ptr;
    // every destination from one read of each source slab, where the slab alone
    // still gives the chip enough waves (4096)
    const int nsrc = (int)(src_is_from ? op.froms.size() : op.tos.size());
    if ((op.d == 2 || op.d == 3) && nsrc <= 8 && ndst <= 8 && rs >= 4096 * 64) {
        dim3 g1((unsigned)((rs + 255) / 256));
        if (op.d == 2)
            hipLaunchKernelGGL((mix_all_kernel<CT, 2, 8, 8>), g1, dim3(256), 0, s, P, Q, ptr, srcs, W, nsrc, ndst, rs);
        else
            hipLaunchKernelGGL((mix_all_kernel<CT, 3, 8, 8>), g1, dim3(256), 0, s, P, Q, ptr, srcs, W, nsrc, ndst, rs);
        NC_HIP(hipGetLastError());
        return;
    }
    switch (op.d) {
        case 2: hipLaunchKernelGGL((mix_kernel<CT, 2>), grid, dim3(256), 0, s, P, Q, ptr, srcs, W, md); break;
        case 3: hipLaunchKernelGGL((mix_kernel<CT, 3>), grid, dim3(256), 0, s, P, Q, ptr, srcs, W, md); break;
        case 4: hipLaunchKernelGGL((mix_kernel<CT, 4>), grid, dim3(256), 0, s, P, Q, ptr, srcs, W, md); break;
        default: hipLaunchKernelGGL((mix_kernel<CT, 0>), grid, dim3(256), 0, s, P, Q, ptr, srcs, W, md); break;
    }
    NC_HIP(hipGetLastError());
}

void mix_blocks(MpoSite const& op, bool src_is_from, const double2* P, double2* Q, int64_t rs, hipStream_t s) {
    mix_blocks_t<double2>(op, src_is_from, P, Q, rs, op.mix_W[src_is_from ? 0 : 1].ptr, s);
}
void mix_blocks(MpoSite const& op, bool src_is_from, const float2* P, float2* Q, int64_t rs, hipStream_t s) {
    mix_blocks_t<float2>(op, src_is_from, P, Q, rs, op.mix_W32[src_is_from ? 0 : 1].ptr, s);
}

void mix_blocks_at(int d, int ndst, const int* dst_ptr, const int* blk_src, const double2* W, const double2* P,
                   double2* Q, int64_t rs, hipStream_t s) {
    if (ndst == 0 || rs == 0) return;
    MixDesc md{d, ndst, rs};
    dim3 grid((unsigned)((rs + 255) / 256), (unsigned)ndst);
    switch (d) {
        case 2: hipLaunchKernelGGL((mix_kernel<double2, 2>), grid, dim3(256), 0, s, P, Q, dst_ptr, blk_src, W, md); break;
        case 3: hipLaunchKernelGGL((mix_kernel<double2, 3>), grid, dim3(256), 0, s, P, Q, dst_ptr, blk_src, W, md); break;
        case 4: hipLaunchKernelGGL((mix_kernel<double2, 4>), grid, dim3(256), 0, s, P, Q, dst_ptr, blk_src, W, md); break;
        default: hipLaunchKernelGGL((mix_kernel<double2, 0>), grid, dim3(256), 0, s, P, Q, dst_ptr, blk_src, W, md); break;
    }
    NC_HIP(hipGetLastError());
}

// ----------------------------------------------------------------------------
// Host bytes into device memory through kernel arguments: the launch copies the
// argument block, so the host buffer may change as soon as the call returns,
// and nothing is allocated or staged in pinned memory.
// ----------------------------------------------------------------------------
namespace {
constexpr int kArgBlob = 2048;
struct ArgBlob {
    unsigned char bytes[kArgBlob];
};
__global__ __launch_bounds__(256) void store_blob_kernel(ArgBlob b, unsigned char* __restrict__ dst, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) dst[i] = b.bytes[i];
}
}  // namespace

void store_bytes(void* dst, const void* src, size_t n, hipStream_t s) {
    const unsigned char* p = static_cast<const unsigned char*>(src);
    unsigned char* q = static_cast<unsigned char*>(dst);
    for (size_t off = 0; off < n; off += kArgBlob) {
        ArgBlob b;
        const int len = (int)std::min<size_t>(kArgBlob, n - off);
        std::memcpy(b.bytes, p + off, (size_t)len);
        hipLaunchKernelGGL(store_blob_kernel, dim3(1), dim3(256), 0, s, b, q + off, len);
        NC_HIP(hipGetLastError());
    }
}

namespace {
__global__ __launch_bounds__(256) void zero_fill_kernel(double2* __restrict__ x, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        x[i] = make_double2(0.0, 0.0);
}
}  // namespace

void zero_fill(double2* x, int64_t n, hipStream_t s) {
    if (n <= 0) return;
    const unsigned g = (unsigned)std::min<int64_t>(4096, (n + 255) / 256);
    hipLaunchKernelGGL(zero_fill_kernel, dim3(g), dim3(256), 0, s, x, n);
    NC_HIP(hipGetLastError());
}

// the same mixing for several chains: blockIdx.z selects the chain
template <typename CT>
struct MixBatchDesc {
    const CT* P;
    CT* Q;
    const int* dst_ptr;
    const int* blk_src;
    const CT* W;
    int64_t rs;
    int ndst, d;
    const int* live;  // null, or a device flag: 0 = skip this chain
};

template <typename CT, int D>
__global__ __launch_bounds__(256) void mix_batch_kernel(const MixBatchDesc<CT>* __restrict__ descs) {
    const MixBatchDesc<CT> md = descs[blockIdx.z];
    if (md.live && *md.live == 0) return;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int dst = blockIdx.y;
    if (dst >= md.ndst || e >= md.rs) return;
    if (D > 0) {
        mix_element<CT, D, (D > 0 ? D : 1)>(md.P, md.Q, md.dst_ptr, md.blk_src, md.W, D, md.rs, e, dst, 0);
    } else {
        for (int b0 = 0; b0 < md.d; b0 += 4)
            mix_element<CT, 0, 4>(md.P, md.Q, md.dst_ptr, md.blk_src, md.W, md.d, md.rs, e, dst, b0);
    }
}

template <typename CT>
static bool mix_descs(std::vector<MixItem<CT>> const& items, const CT* (*W)(MpoSite const&, int),
                      std::vector<MixBatchDesc<CT>>& ds, dim3& grid, int& d) {
    if (items.empty()) return false;
    d = items[0].op->d;
    int64_t maxrs = 0;
    int maxdst = 0;
    for (auto const& it : items) {
        if (it.op->d != d) throw Error("mix_blocks_batch: chains with different physical dimensions", -4);
        const int dir = it.src_is_from ? 0 : 1;
        MixBatchDesc<CT> m;
        m.P = it.P;
        m.Q = it.Q;
        m.dst_ptr = it.op->mix_ptr[dir].ptr;
        m.blk_src = it.op->mix_src[dir].ptr;
        m.W = W(*it.op, dir);
        m.rs = it.rs;
        m.ndst = (int)(it.src_is_from ? it.op->tos.size() : it.op->froms.size());
        m.d = d;
        m.live = it.live;
        ds.push_back(m);
        maxrs = std::max(maxrs, it.rs);
        maxdst = std::max(maxdst, m.ndst);
    }
    if (maxrs == 0 || maxdst == 0) return false;
    if (d < 1) throw Error("mix_blocks_batch: physical dimension below 1", -4);
    grid = dim3((unsigned)((maxrs + 255) / 256), (unsigned)maxdst, (unsigned)ds.size());
    return true;
}

template <typename CT>
static void mix_launch(const MixBatchDesc<CT>* dd, dim3 grid, int d, hipStream_t s) {
    switch (d) {
        case 2: hipLaunchKernelGGL((mix_batch_kernel<CT, 2>), grid, dim3(256), 0, s, dd); break;
        case 3: hipLaunchKernelGGL((mix_batch_kernel<CT, 3>), grid, dim3(256), 0, s, dd); break;
        case 4: hipLaunchKernelGGL((mix_batch_kernel<CT, 4>), grid, dim3(256), 0, s, dd); break;
        default: hipLaunchKernelGGL((mix_batch_kernel<CT, 0>), grid, dim3(256), 0, s, dd); break;
    }
    NC_HIP(hipGetLastError());
}

template <typename CT>
static void mix_batch_t(std::vector<MixItem<CT>> const& items, const CT* (*W)(MpoSite const&, int), hipStream_t s) {
    std::vector<MixBatchDesc<CT>> ds;
    dim3 grid;
    int d;
    if (!mix_descs(items, W, ds, grid, d)) return;
    int slot;
    auto* dd = (const MixBatchDesc<CT>*)stage_descriptors(ds.data(), ds.size() * sizeof(MixBatchDesc<CT>), s, &slot);
    mix_launch(dd, grid, d, s);
    release_descriptors(slot, s);
}

template <typename CT>
struct PreparedMix final : PreparedLaunch {
    DevBuf<MixBatchDesc<CT>> desc;
    dim3 grid;
    int d = 0;
    bool any = false;
    void launch(hipStream_t s) const override {
        if (any) mix_launch(desc.ptr, grid, d, s);
    }
};

template <typename CT>
static std::unique_ptr<PreparedLaunch> prepare_mix_t(std::vector<MixItem<CT>> const& items,
                                                     const CT* (*W)(MpoSite const&, int)) {
    std::unique_ptr<PreparedMix<CT>> p(new PreparedMix<CT>());
    std::vector<MixBatchDesc<CT>> ds;
    p->any = mix_descs(items, W, ds, p->grid, p->d);
    if (p->any) {
        p->desc.resize(ds.size());
        h2d(p->desc.ptr, ds.data(), ds.size() * sizeof(MixBatchDesc<CT>));
    }
    return p;
}

static const double2* mixw64(MpoSite const& op, int dir) { return op.mix_W[dir].ptr; }
static const float2* mixw32(MpoSite const& op, int dir) { return op.mix_W32[dir].ptr; }
void mix_blocks_batch(std::vector<MixItem<double2>> const& items, hipStream_t s) { mix_batch_t(items, mixw64, s); }
void mix_blocks_batch(std::vector<MixItem<float2>> const& items, hipStream_t s) { mix_batch_t(items, mixw32, s); }
std::unique_ptr<PreparedLaunch> prepare_mix(std::vector<MixItem<double2>> const& items) {
    return prepare_mix_t(items, mixw64);
}
std::unique_ptr<PreparedLaunch> prepare_mix(std::vector<MixItem<float2>> const& items) {
    return prepare_mix_t(items, mixw32);
}


__global__ void to_c64_kernel(const double2* __restrict__ a, float2* __restrict__ b, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        b[i] = make_float2((float)a[i].x, (float)a[i].y);
}
__global__ void to_c128_kernel(const float2* __restrict__ a, double2* __restrict__ b, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        b[i] = make_double2((double)a[i].x, (double)a[i].y);
}
void to_c64(const double2* src, float2* dst, int64_t n, hipStream_t s) {
    if (n <= 0) return;
    unsigned g = (unsigned)std::min<int64_t>(8192, (n + 255) / 256);
    hipLaunchKernelGGL(to_c64_kernel, dim3(g), dim3(256), 0, s, src, dst, n);
    NC_HIP(hipGetLastError());
}
void to_c128(const float2* src, double2* dst, int64_t n, hipStream_t s) {
    if (n <= 0) return;
    unsigned g = (unsigned)std::min<int64_t>(8192, (n + 255) / 256);
    hipLaunchKernelGGL(to_c128_kernel, dim3(g), dim3(256), 0, s, src, dst, n);
    NC_HIP(hipGetLastError());
}

// ----------------------------------------------------------------------------
// Vector kernels
// ----------------------------------------------------------------------------
constexpr int RED_THREADS = 256;
constexpr int RED_BLOCKS = 512;

__device__ __forceinline__ double warp_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// partial[j * gridDim.x + block] = sum over the block's grid-stride range of conj(V_j) w
__global__ __launch_bounds__(RED_THREADS) void zdotc_multi_kernel(const double2* __restrict__ V, int64_t ldv,
                                                                  int m, const double2* __restrict__ w,
                                                                  int64_t n, double2* __restrict__ partial) {
    __shared__ double sre[RED_THREADS / 64], sim[RED_THREADS / 64];
    for (int j = 0; j < m; ++j) {
        double re = 0, im = 0;
        const double2* Vj = V + (int64_t)j * ldv;
        for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
            const double2 a = Vj[i], b = w[i];
            re += a.x * b.x + a.y * b.y;
            im += a.x * b.y - a.y * b.x;
        }
        re = warp_sum(re);
        im = warp_sum(im);
        if ((threadIdx.x & 63) == 0) { sre[threadIdx.x >> 6] = re; sim[threadIdx.x >> 6] = im; }
        __syncthreads();
        if (threadIdx.x == 0) {
            double tr = 0, ti = 0;
            for (int k = 0; k < RED_THREADS / 64; ++k) { tr += sre[k]; ti += sim[k]; }
            partial[(int64_t)j * gridDim.x + blockIdx.x] = make_double2(tr, ti);
        }
        __syncthreads();
    }
}

// final reduction: out[j] = sum_b partial[j * nb + b]; one block per j
__global__ __launch_bounds__(RED_THREADS) void reduce_partials_kernel(const double2* __restrict__ partial, int nb,
                                                                      double2* __restrict__ out) {
    __shared__ double sre[RED_THREADS / 64], sim[RED_THREADS / 64];
    const int j = blockIdx.x;
    double re = 0, im = 0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) {
        re += partial[(int64_t)j * nb + b].x;
        im += partial[(int64_t)j * nb + b].y;
    }
    re = warp_sum(re);
    im = warp_sum(im);
    if ((threadIdx.x & 63) == 0) { sre[threadIdx.x >> 6] = re; sim[threadIdx.x >> 6] = im; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double tr = 0, ti = 0;
        for (int k = 0; k < RED_THREADS / 64; ++k) { tr += sre[k]; ti += sim[k]; }
        out[j] = make_double2(tr, ti);
    }
}

__global__ __launch_bounds__(RED_THREADS) void zdotu_kernel(const double2* __restrict__ x, const double2* __restrict__ y,
                                                            int64_t n, double2* __restrict__ partial) {
    __shared__ double sre[RED_THREADS / 64], sim[RED_THREADS / 64];
    double re = 0, im = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double2 a = x[i], b = y[i];
        re += a.x * b.x - a.y * b.y;
        im += a.x * b.y + a.y * b.x;
    }
    re = warp_sum(re);
    im = warp_sum(im);
    if ((threadIdx.x & 63) == 0) { sre[threadIdx.x >> 6] = re; sim[threadIdx.x >> 6] = im; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double tr = 0, ti = 0;
        for (int k = 0; k < RED_THREADS / 64; ++k) { tr += sre[k]; ti += sim[k]; }
        partial[blockIdx.x] = make_double2(tr, ti);
    }
}

// x <- x - tau (x . v) vc for one Householder reflector, with the dot product
// kept on the device: every block adds up the nb partial sums zdotu_kernel
// left (nb <= RED_BLOCKS) before it updates its share of x.
__global__ __launch_bounds__(RED_THREADS) void refl_update_kernel(double2* __restrict__ x, const double2* __restrict__ vc,
                                                                  int64_t n, const double2* __restrict__ partial, int nb,
                                                                  double2 tau) {
    __shared__ double sre[RED_THREADS / 64], sim[RED_THREADS / 64];
    double re = 0, im = 0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) {
        re += partial[b].x;
        im += partial[b].y;
    }
    re = warp_sum(re);
    im = warp_sum(im);
    if ((threadIdx.x & 63) == 0) { sre[threadIdx.x >> 6] = re; sim[threadIdx.x >> 6] = im; }
    __syncthreads();
    double tr = 0, ti = 0;
    for (int k = 0; k < RED_THREADS / 64; ++k) { tr += sre[k]; ti += sim[k]; }
    // c = -tau * dot
    const double cr = -(tau.x * tr - tau.y * ti), ci = -(tau.x * ti + tau.y * tr);
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double2 v = vc[i];
        double2 a = x[i];
        a.x += cr * v.x - ci * v.y;
        a.y += cr * v.y + ci * v.x;
        x[i] = a;
    }
}

__global__ __launch_bounds__(RED_THREADS) void trace_pair_kernel(const double2* __restrict__ A, const double2* __restrict__ B,
                                                                 int64_t b, int64_t c, double2* __restrict__ partial) {
    __shared__ double sre[RED_THREADS / 64], sim[RED_THREADS / 64];
    double re = 0, im = 0;
    const int64_t total = b * b * c;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = e % b, j = (e / b) % b, k = e / (b * b);
        const double2 x = A[e], y = B[j + i * b + k * b * b];
        re += x.x * y.x - x.y * y.y;
        im += x.x * y.y + x.y * y.x;
    }
    re = warp_sum(re);
    im = warp_sum(im);
    if ((threadIdx.x & 63) == 0) { sre[threadIdx.x >> 6] = re; sim[threadIdx.x >> 6] = im; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double tr = 0, ti = 0;
        for (int k = 0; k < RED_THREADS / 64; ++k) { tr += sre[k]; ti += sim[k]; }
        partial[blockIdx.x] = make_double2(tr, ti);
    }
}

__global__ void zgemv_sub_kernel(const double2* __restrict__ V, int64_t ldv, int m, const double2* __restrict__ h,
                                 double2* __restrict__ w, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double2 acc = w[i];
        for (int j = 0; j < m; ++j) {
            const double2 v = V[(int64_t)j * ldv + i], c = h[j];
            acc.x -= v.x * c.x - v.y * c.y;
            acc.y -= v.x * c.y + v.y * c.x;
        }
        w[i] = acc;
    }
}

__global__ void zgemv_n_kernel(const double2* __restrict__ V, int64_t ldv, int m, const double2* __restrict__ c,
                               double2* __restrict__ out, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double2 acc = make_double2(0, 0);
        for (int j = 0; j < m; ++j) {
            const double2 v = V[(int64_t)j * ldv + i], k = c[j];
            acc.x += v.x * k.x - v.y * k.y;
            acc.y += v.x * k.y + v.y * k.x;
        }
        out[i] = acc;
    }
}

__global__ void zaxpby_kernel(double2* __restrict__ y, double2 a, const double2* __restrict__ x, double2 b, int64_t n,
                              int has_x) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double2 yv = y[i];
        double2 o;
        o.x = b.x * yv.x - b.y * yv.y;
        o.y = b.x * yv.y + b.y * yv.x;
        if (has_x) {
            const double2 xv = x[i];
            o.x += a.x * xv.x - a.y * xv.y;
            o.y += a.x * xv.y + a.y * xv.x;
        }
        y[i] = o;
    }
}

__global__ void conj_copy_kernel(double2* __restrict__ dst, const double2* __restrict__ src, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double2 v = src[i];
        dst[i] = make_double2(v.x, -v.y);
    }
}

__global__ void permute_102_kernel(const double2* __restrict__ src, double2* __restrict__ dst, int64_t n0, int64_t n1,
                                   int64_t n2) {
    const int64_t total = n0 * n1 * n2;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        // dst(j, i, k) = src(i, j, k), e enumerates dst
        const int64_t j = e % n1, i = (e / n1) % n0, k = e / (n0 * n1);
        dst[e] = src[i + j * n0 + k * n0 * n1];
    }
}

__global__ void scale_rows_kernel(double2* __restrict__ X, const double* __restrict__ sv, int64_t m, int64_t ncol) {
    const int64_t total = m * ncol;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const double f = sv[e % m];
        double2 v = X[e];
        X[e] = make_double2(v.x * f, v.y * f);
    }
}


static unsigned grid_for(int64_t n) {
    int64_t g = (n + 255) / 256;
    if (g > 4096) g = 4096;
    if (g < 1) g = 1;
    return (unsigned)g;
}

namespace {
// pinned result buffer for reductions
struct RedScratch {
    DevBuf<double2> partial, out;
    double2* host = nullptr;
    RedScratch() {
        partial.resize((size_t)RED_BLOCKS * 64);
        out.resize(64);
        NC_HIP(hipHostMalloc((void**)&host, 64 * sizeof(double2), hipHostMallocDefault));
    }
};
RedScratch& red() {
    static RedScratch* r = new RedScratch();
    return *r;
}
int red_blocks(int64_t n) {
    int64_t b = (n + RED_THREADS - 1) / RED_THREADS;
    if (b > RED_BLOCKS) b = RED_BLOCKS;
    if (b < 1) b = 1;
    return (int)b;
}
}  // namespace

void zdotc_multi(const double2* V, int64_t ldv, int m, const double2* w, int64_t n, zd* h_host, hipStream_t s) {
    if (m <= 0) return;
    auto& r = red();
    int done = 0;
    while (done < m) {
        int mm = std::min(64, m - done);
        int nb = red_blocks(n);
        hipLaunchKernelGGL(zdotc_multi_kernel, dim3(nb), dim3(RED_THREADS), 0, s, V + (int64_t)done * ldv, ldv, mm, w, n,
                           r.partial.ptr);
        hipLaunchKernelGGL(reduce_partials_kernel, dim3(mm), dim3(RED_THREADS), 0, s, r.partial.ptr, nb, r.out.ptr);
        NC_HIP(hipGetLastError());
        NC_HIP(hipMemcpyAsync(r.host, r.out.ptr, mm * sizeof(double2), hipMemcpyDeviceToHost, s));
        NC_HIP(hipStreamSynchronize(s));
        for (int j = 0; j < mm; ++j) h_host[done + j] = zd(r.host[j].x, r.host[j].y);
        done += mm;
    }
}

double dznrm2(const double2* x, int64_t n, hipStream_t s) {
    zd h;
    zdotc_multi(x, n, 1, x, n, &h, s);
    return std::sqrt(std::max(0.0, h.real()));
}

zd zdotu(const double2* x, const double2* y, int64_t n, hipStream_t s) {
    auto& r = red();
    int nb = red_blocks(n);
    hipLaunchKernelGGL(zdotu_kernel, dim3(nb), dim3(RED_THREADS), 0, s, x, y, n, r.partial.ptr);
    hipLaunchKernelGGL(reduce_partials_kernel, dim3(1), dim3(RED_THREADS), 0, s, r.partial.ptr, nb, r.out.ptr);
    NC_HIP(hipGetLastError());
    NC_HIP(hipMemcpyAsync(r.host, r.out.ptr, sizeof(double2), hipMemcpyDeviceToHost, s));
    NC_HIP(hipStreamSynchronize(s));
    return zd(r.host[0].x, r.host[0].y);
}

int refl_partials() { return RED_BLOCKS; }

void zrefl_apply(double2* x, const double2* v, const double2* vc, zd tau, int64_t n, double2* partial,
                 hipStream_t s) {
    const int nb = red_blocks(n);
    hipLaunchKernelGGL(zdotu_kernel, dim3(nb), dim3(RED_THREADS), 0, s, x, v, n, partial);
    const unsigned g = (unsigned)std::min<int64_t>(1024, (n + RED_THREADS - 1) / RED_THREADS);
    hipLaunchKernelGGL(refl_update_kernel, dim3(std::max(1u, g)), dim3(RED_THREADS), 0, s, x, vc, n, partial, nb,
                       make_double2(tau.real(), tau.imag()));
    NC_HIP(hipGetLastError());
}

zd trace_pair(const double2* A, const double2* B, int64_t b, int64_t c, hipStream_t s) {
    auto& r = red();
    int nb = red_blocks(b * b * c);
    hipLaunchKernelGGL(trace_pair_kernel, dim3(nb), dim3(RED_THREADS), 0, s, A, B, b, c, r.partial.ptr);
    hipLaunchKernelGGL(reduce_partials_kernel, dim3(1), dim3(RED_THREADS), 0, s, r.partial.ptr, nb, r.out.ptr);
    NC_HIP(hipGetLastError());
    NC_HIP(hipMemcpyAsync(r.host, r.out.ptr, sizeof(double2), hipMemcpyDeviceToHost, s));
    NC_HIP(hipStreamSynchronize(s));
    return zd(r.host[0].x, r.host[0].y);
}

namespace {
// small coefficient vectors are staged through a pinned ring
struct CoefRing {
    static constexpr int kSlots = 64, kLen = 128;
    double2* host = nullptr;
    DevBuf<double2> dev;
    hipEvent_t done[kSlots];
    int next = 0;
    CoefRing() {
        NC_HIP(hipHostMalloc((void**)&host, kSlots * kLen * sizeof(double2), hipHostMallocDefault));
        dev.resize(kSlots * kLen);
        for (int i = 0; i < kSlots; ++i) {
            NC_HIP(hipEventCreateWithFlags(&done[i], hipEventDisableTiming));
            NC_HIP(hipEventRecord(done[i], Context::get().stream));
        }
    }
    const double2* stage(const zd* c, int m, hipStream_t s) {
        if (m > kLen) throw Error("coefficient vector too long");
        int slot = next;
        next = (next + 1) % kSlots;
        NC_HIP(hipEventSynchronize(done[slot]));
        double2* h = host + slot * kLen;
        for (int j = 0; j < m; ++j) h[j] = make_double2(c[j].real(), c[j].imag());
        double2* d = dev.ptr + slot * kLen;
        NC_HIP(hipMemcpyAsync(d, h, m * sizeof(double2), hipMemcpyHostToDevice, s));
        NC_HIP(hipEventRecord(done[slot], s));
        return d;
    }
};
CoefRing& coef() {
    static CoefRing* c = new CoefRing();
    return *c;
}
}  // namespace

void zgemv_sub(const double2* V, int64_t ldv, int m, const zd* h, double2* w, int64_t n, hipStream_t s) {
    if (m <= 0) return;
    const double2* dh = coef().stage(h, m, s);
    hipLaunchKernelGGL(zgemv_sub_kernel, dim3(grid_for(n)), dim3(256), 0, s, V, ldv, m, dh, w, n);
    NC_HIP(hipGetLastError());
}

void zgemv_n(const double2* V, int64_t ldv, int m, const zd* c, double2* out, int64_t n, hipStream_t s) {
    const double2* dc = coef().stage(c, m, s);
    hipLaunchKernelGGL(zgemv_n_kernel, dim3(grid_for(n)), dim3(256), 0, s, V, ldv, m, dc, out, n);
    NC_HIP(hipGetLastError());
}

void zaxpby(double2* y, zd a, const double2* x, zd b, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(zaxpby_kernel, dim3(grid_for(n)), dim3(256), 0, s, y, make_double2(a.real(), a.imag()), x,
                       make_double2(b.real(), b.imag()), n, x != nullptr ? 1 : 0);
    NC_HIP(hipGetLastError());
}
void zscal(double2* x, zd a, int64_t n, hipStream_t s) { zaxpby(x, zd(0, 0), nullptr, a, n, s); }
void zaxpy(double2* y, zd a, const double2* x, int64_t n, hipStream_t s) { zaxpby(y, a, x, zd(1, 0), n, s); }

void zconj_copy(double2* dst, const double2* src, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(conj_copy_kernel, dim3(grid_for(n)), dim3(256), 0, s, dst, src, n);
    NC_HIP(hipGetLastError());
}
void conj_inplace(double2* X, int64_t n, hipStream_t s) { zconj_copy(X, X, n, s); }

void permute_102(const double2* src, double2* dst, int64_t n0, int64_t n1, int64_t n2, hipStream_t s) {
    hipLaunchKernelGGL(permute_102_kernel, dim3(grid_for(n0 * n1 * n2)), dim3(256), 0, s, src, dst, n0, n1, n2);
    NC_HIP(hipGetLastError());
}

void scale_rows(double2* X, const double* sv, int64_t m, int64_t ncol, hipStream_t s) {
    hipLaunchKernelGGL(scale_rows_kernel, dim3(grid_for(m * ncol)), dim3(256), 0, s, X, sv, m, ncol);
    NC_HIP(hipGetLastError());
}


// uniform entries in [-1, 1) + i [-1, 1) from a counter-based hash (splitmix64)
__global__ void fill_random_kernel(double2* x, int64_t n, uint64_t key) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        uint64_t z = key ^ ((uint64_t)e * 0x9E3779B97F4A7C15ull);
        double v[2];
        for (int h = 0; h < 2; ++h) {
            z += 0x9E3779B97F4A7C15ull;
            uint64_t t = z;
            t = (t ^ (t >> 30)) * 0xBF58476D1CE4E5B9ull;
            t = (t ^ (t >> 27)) * 0x94D049BB133111EBull;
            t ^= t >> 31;
            v[h] = (double)(t >> 11) * (2.0 / 9007199254740992.0) - 1.0;
        }
        x[e] = make_double2(v[0], v[1]);
    }
}

void fill_random(double2* x, int64_t n, uint64_t key, hipStream_t s) {
    if (n <= 0) return;
    const unsigned g = (unsigned)std::min<int64_t>(4096, (n + 255) / 256);
    hipLaunchKernelGGL(fill_random_kernel, dim3(g), dim3(256), 0, s, x, n, key);
    NC_HIP(hipGetLastError());
}

// y = a x + b y for several vectors (b = 0: y = a x, y not read)
struct AxpbyDesc {
    double2* y;
    const double2* x;
    double2 a, b;
    int64_t n;
};
__global__ void axpby_batch_kernel(const AxpbyDesc* __restrict__ ds) {
    const AxpbyDesc d = ds[blockIdx.y];
    const bool usey = d.b.x != 0.0 || d.b.y != 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < d.n; i += (int64_t)gridDim.x * blockDim.x) {
        double2 o = make_double2(0.0, 0.0);
        if (usey) {
            const double2 yv = d.y[i];
            o.x = d.b.x * yv.x - d.b.y * yv.y;
            o.y = d.b.x * yv.y + d.b.y * yv.x;
        }
        if (d.x) {
            const double2 xv = d.x[i];
            o.x += d.a.x * xv.x - d.a.y * xv.y;
            o.y += d.a.x * xv.y + d.a.y * xv.x;
        }
        d.y[i] = o;
    }
}

void axpby_batch(std::vector<double2*> const& y, std::vector<zd> const& a, std::vector<const double2*> const& x,
                 std::vector<zd> const& b, std::vector<int64_t> const& n, hipStream_t s) {
    if (y.empty()) return;
    std::vector<AxpbyDesc> ds(y.size());
    int64_t maxn = 0;
    for (size_t q = 0; q < y.size(); ++q) {
        ds[q] = AxpbyDesc{y[q], x[q], make_double2(a[q].real(), a[q].imag()), make_double2(b[q].real(), b[q].imag()),
                          n[q]};
        maxn = std::max(maxn, n[q]);
    }
    if (maxn == 0) return;
    int slot;
    auto* dd = (const AxpbyDesc*)stage_descriptors(ds.data(), ds.size() * sizeof(AxpbyDesc), s, &slot);
    const unsigned g = (unsigned)std::max<int64_t>(
        1, std::min<int64_t>(std::max<int64_t>(1, 4096 / (int64_t)y.size()), (maxn + 255) / 256));
    hipLaunchKernelGGL(axpby_batch_kernel, dim3(g, (unsigned)y.size()), dim3(256), 0, s, dd);
    NC_HIP(hipGetLastError());
    release_descriptors(slot, s);
}

// ---------------------------------------------------------------------------
// Several reductions in one launch: blockIdx.y selects the item, each block
// writes one partial sum, a second kernel sums the partials of every item in a
// fixed order (deterministic), and the host reads all results at once.
// ---------------------------------------------------------------------------
struct RedItemDev {
    const double2* x;
    const double2* y;
    int64_t n;     // elements (dot) or b (trace)
    int64_t c;     // trace: third extent
    int kind, _pad;
};

__global__ __launch_bounds__(RED_THREADS) void reduce_batch_kernel(const RedItemDev* __restrict__ items,
                                                                   double2* __restrict__ partial) {
    __shared__ double sre[RED_THREADS / 64], sim[RED_THREADS / 64];
    const RedItemDev it = items[blockIdx.y];
    double re = 0, im = 0;
    if (it.kind == 0) {  // sum conj(x) y
        for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < it.n; e += (int64_t)gridDim.x * blockDim.x) {
            const double2 a = it.x[e], b = it.y[e];
            re += a.x * b.x + a.y * b.y;
            im += a.x * b.y - a.y * b.x;
        }
    } else {  // sum_{i,j,k} x(i,j,k) y(j,i,k), b x b x c
        const int64_t b = it.n, total = it.n * it.n * it.c;
        for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
            const int64_t i = e % b, j = (e / b) % b, k = e / (b * b);
            const double2 a = it.x[e], v = it.y[j + i * b + k * b * b];
            re += a.x * v.x - a.y * v.y;
            im += a.x * v.y + a.y * v.x;
        }
    }
    re = warp_sum(re);
    im = warp_sum(im);
    if ((threadIdx.x & 63) == 0) { sre[threadIdx.x >> 6] = re; sim[threadIdx.x >> 6] = im; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double tr = 0, ti = 0;
        for (int k = 0; k < RED_THREADS / 64; ++k) { tr += sre[k]; ti += sim[k]; }
        partial[(int64_t)blockIdx.y * gridDim.x + blockIdx.x] = make_double2(tr, ti);
    }
}

std::vector<zd> reduce_batch(std::vector<RedItem> const& items, hipStream_t s) {
    std::vector<zd> out(items.size(), zd(0, 0));
    if (items.empty()) return out;
    std::vector<RedItemDev> dv(items.size());
    int64_t maxn = 1;
    for (size_t q = 0; q < items.size(); ++q) {
        RedItem const& it = items[q];
        dv[q] = RedItemDev{it.x, it.y, it.n, it.c, it.trace ? 1 : 0, 0};
        maxn = std::max(maxn, it.trace ? it.n * it.n * it.c : it.n);
    }
    const int nb = (int)std::max<int64_t>(1, std::min<int64_t>((maxn + RED_THREADS - 1) / RED_THREADS,
                                                                std::max<int64_t>(1, 4096 / (int64_t)items.size())));
    static thread_local DevBuf<double2> partial, res;
    static thread_local std::vector<double2> host;
    if (partial.n < (size_t)nb * items.size()) partial.resize((size_t)nb * items.size());
    if (res.n < items.size()) res.resize(items.size());
    int slot;
    auto* d = (const RedItemDev*)stage_descriptors(dv.data(), dv.size() * sizeof(RedItemDev), s, &slot);
    hipLaunchKernelGGL(reduce_batch_kernel, dim3(nb, (unsigned)items.size()), dim3(RED_THREADS), 0, s, d, partial.ptr);
    hipLaunchKernelGGL(reduce_partials_kernel, dim3((unsigned)items.size()), dim3(RED_THREADS), 0, s, partial.ptr, nb,
                       res.ptr);
    NC_HIP(hipGetLastError());
    release_descriptors(slot, s);
    host.resize(items.size());
    d2h(host.data(), res.ptr, items.size() * sizeof(double2));
    for (size_t q = 0; q < items.size(); ++q) out[q] = zd(host[q].x, host[q].y);
    return out;
}

}  // namespace nc
