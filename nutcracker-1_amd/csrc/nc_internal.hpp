// Internal declarations of the MI355X (gfx950) variational-MPS core.
//
// Data layout (identical to the reference's Fortran core, core.f95): every tensor
// is complex128 stored as interleaved (re, im) doubles in column-major order.
//   state site    psi(r, l, s)   shape (br, bl, d)
//   left env      L(i, j, k)     shape (bl, bl, cl)
//   right env     R(r', r, k)    shape (br, br, cr)
//   MPO blocks    idx(2, n) 1-based (from, to), mats(d, d, n) = W_n(a, b)
#pragma once

#include <hip/hip_runtime.h>

#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <stdexcept>
#include <string>
#include <functional>
#include <vector>

namespace nc {

using zd = std::complex<double>;

// ----------------------------------------------------------------------------
// Errors
// ----------------------------------------------------------------------------
struct Error : std::runtime_error {
    int code;
    Error(std::string const& m, int c = -1) : std::runtime_error(m), code(c) {}
};

void check_hip(hipError_t e, const char* what, const char* file, int line);
#define NC_HIP(x) ::nc::check_hip((x), #x, __FILE__, __LINE__)

void set_last_error(std::string const& m);
std::string const& last_error();

// ----------------------------------------------------------------------------
// Device context: one stream per process on the current device, a caching
// device allocator, and pinned staging for host <-> device copies.
// ----------------------------------------------------------------------------
struct Context {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev_a = nullptr, ev_b = nullptr;
    static Context& get();
    void sync();
};

// Caching device allocator: blocks are rounded up to 4 KiB (or to the next
// power-of-two quarter above 64 MiB) and kept on per-size free lists.
void* dev_alloc(size_t bytes);
void dev_free(void* p, size_t bytes);
size_t dev_bytes_in_use();
void dev_release_cache();

template <class T>
struct DevBuf {
    T* ptr = nullptr;
    size_t n = 0;
    DevBuf() = default;
    explicit DevBuf(size_t count) { resize(count); }
    DevBuf(DevBuf const&) = delete;
    DevBuf& operator=(DevBuf const&) = delete;
    DevBuf(DevBuf&& o) noexcept : ptr(o.ptr), n(o.n) { o.ptr = nullptr; o.n = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) { release(); ptr = o.ptr; n = o.n; o.ptr = nullptr; o.n = 0; }
        return *this;
    }
    ~DevBuf() { release(); }
    void release() {
        if (ptr) dev_free(ptr, n * sizeof(T));
        ptr = nullptr; n = 0;
    }
    void resize(size_t count) {
        if (count == n) return;
        release();
        if (count) ptr = static_cast<T*>(dev_alloc(count * sizeof(T)));
        n = count;
    }
    size_t bytes() const { return n * sizeof(T); }
};

using ZBuf = DevBuf<double2>;

void h2d(void* dst, void const* src, size_t bytes);
void d2h(void* dst, void const* src, size_t bytes);
void d2d(void* dst, void const* src, size_t bytes);
void dzero(void* dst, size_t bytes);

// ----------------------------------------------------------------------------
// Hand-written complex GEMM (gfx950 MFMA).  C(M x N) = sum_s opA(A_s) opB(B_s)
// (+ C when accumulate), batched over independent C's that share M and N.
// op: 0 = N, 1 = T, 2 = C (conjugate transpose).
// ----------------------------------------------------------------------------
struct GemmSeg {
    const double2* A;
    const double2* B;
    int64_t lda, ldb;
    int32_t K, _pad;
};
struct GemmBatch {
    double2* C;
    int64_t ldc;
    int32_t seg_begin, seg_end;
    int32_t M, N;             // this batch's output size; 0 = the plan's M, N
    const int* live = nullptr;  // device flag read at launch: 0 = skip this batch
    int32_t upper = 0;          // 1: C is Hermitian, only its tiles on and above the
    int32_t _pad = 0;           //    diagonal are computed; the elements below them are
                                //    undefined afterwards (a split-K launch overwrites them)
};

enum Op : int { OpN = 0, OpT = 1, OpC = 2 };

// Launch-time choices of the GEMM (see gemm.hip).
struct GemmTuning {
    // Gauss's three-multiplication complex product (3/4 of the matrix-core
    // work): the default in both precisions; the autotuner also times 4M
    bool use_3m_f64 = true;
    bool use_3m_f32 = true;
    // complex64 plans may also be tuned onto the bf16 split kernel (FP32
    // accuracy from six bf16 partial products); NC_BF16X6=0 takes it out
    bool bf16x6 = [] {
        const char* e = std::getenv("NC_BF16X6");
        return !(e && e[0] == '0');
    }();
    bool split_k = true;            // split K when a launch has few output tiles
    int64_t split_k_tiles = 1024;   // ... fewer than this many (tiles x batches)
    int64_t split_k_tiles_dma = 768;  // ... for the LDS-copy kernel (3 workgroups per CU)
    // XCD-blocked tile order of the LDS-copy kernel (NC_DMA_XCD_BLOCKS=0 keeps
    // the per-batch row-major ranges)
    bool dma_xcd_blocks = [] {
        const char* e = std::getenv("NC_DMA_XCD_BLOCKS");
        return !(e && e[0] == '0');
    }();
};
GemmTuning& gemm_tuning();

// Shape key of a GEMM plan for the autotuner: the site's bond dimensions
// (chi_left, chi_right), physical dimension, MPO bonds, and which contraction
// step the plan is (tag; 0 = not tuned).  Grouped plans of several chains use
// the largest bond dimensions and the number of chains.
struct TuneKey {
    int tag = 0;
    int bl = 0, br = 0, d = 0, cl = 0, cr = 0, count = 1;
};
// Autotuner switch (default on; NC_AUTOTUNE=0 in the environment turns the
// default off) and a text table of the choices made so far.
bool& gemm_autotune();
std::string gemm_tune_table();
// force one launch choice on every GEMM plan (bk 16, 32 or -16 = 16 with two LDS
// stages, 6 = complex64 on the bf16 split kernel, 8 / 9 = complex128 LDS-copy
// pipeline of 3 / 4 stages, where the plan allows them; nsplit 0 = heuristic;
// m3 -1 = default, 0 = 4M, 1 = 3M); bk = 0 clears
void gemm_force_choice(int bk, int nsplit, int m3);

// Host-side builder: keeps descriptor lists and uploads them per launch into a
// small reusable device arena.
struct GemmPlan {
    std::vector<GemmBatch> batches;
    std::vector<GemmSeg> segs;
    int M = 0, N = 0, opA = OpN, opB = OpN;
    bool accumulate = false;
    TuneKey key;
    void clear() { batches.clear(); segs.clear(); }
    void begin_batch(double2* C, int64_t ldc, int M_ = 0, int N_ = 0, const int* live = nullptr) {
        batches.push_back(GemmBatch{C, ldc, (int32_t)segs.size(), (int32_t)segs.size(), M_, N_, live});
    }
    void add_seg(const double2* A, int64_t lda, const double2* B, int64_t ldb, int K) {
        segs.push_back(GemmSeg{A, B, lda, ldb, K, 0});
        batches.back().seg_end = (int32_t)segs.size();
    }
    int max_k() const;
};

// Upload-and-launch (descriptor arena is stream ordered).
void zgemm_run(GemmPlan const& plan, hipStream_t s);
// The same plan with its descriptor lists written into caller-provided device
// memory (`cap` bytes at `desc`, zgemm_desc_bytes of them needed) by
// stream-ordered store kernels; the default launch choice, no autotuning and
// no K split, so nothing is allocated.
size_t zgemm_desc_bytes(GemmPlan const& plan);
void zgemm_run_at(GemmPlan const& plan, void* desc, size_t cap, hipStream_t s);

// n host bytes into device memory at dst through kernel arguments (stream
// ordered; the host buffer may be reused as soon as the call returns)
void store_bytes(void* dst, const void* src, size_t n, hipStream_t s);
// x[0 .. n) = 0 with a fill kernel
void zero_fill(double2* x, int64_t n, hipStream_t s);

// The same for complex64 operands (FP32 MFMA), used by the mixed-precision form.
struct CGemmSeg {
    const float2* A;
    const float2* B;
    int64_t lda, ldb;
    int32_t K, _pad;
};
struct CGemmBatch {
    float2* C;
    int64_t ldc;
    int32_t seg_begin, seg_end;
    int32_t M, N;               // 0 = the plan's M, N
    const int* live = nullptr;  // device flag read at launch: 0 = skip this batch
    int32_t upper = 0;          // as in GemmBatch
    int32_t _pad = 0;
};
struct CGemmPlan {
    std::vector<CGemmBatch> batches;
    std::vector<CGemmSeg> segs;
    int M = 0, N = 0, opA = OpN, opB = OpN;
    bool accumulate = false;
    TuneKey key;
    void begin_batch(float2* C, int64_t ldc, int M_ = 0, int N_ = 0, const int* live = nullptr) {
        batches.push_back(CGemmBatch{C, ldc, (int32_t)segs.size(), (int32_t)segs.size(), M_, N_, live});
    }
    void add_seg(const float2* A, int64_t lda, const float2* B, int64_t ldb, int K) {
        segs.push_back(CGemmSeg{A, B, lda, ldb, K, 0});
        batches.back().seg_end = (int32_t)segs.size();
    }
};
void cgemm_run(CGemmPlan const& plan, hipStream_t s);

// A launch whose descriptors were uploaded once into device memory of its own:
// launch() does no host work beyond the kernel launches, so it can run many
// times (and be recorded by a stream capture).  The pointers it was built with
// must stay valid while it is used.
struct PreparedLaunch {
    virtual ~PreparedLaunch() = default;
    virtual void launch(hipStream_t s) const = 0;
};
std::unique_ptr<PreparedLaunch> prepare_gemm(GemmPlan const& plan);
std::unique_ptr<PreparedLaunch> prepare_gemm(CGemmPlan const& plan);

// ----------------------------------------------------------------------------
// Sparse MPO block list, prepared for grouped contraction.
// ----------------------------------------------------------------------------
struct MpoSite {
    int d = 0, cl = 0, cr = 0, nmat = 0;
    std::vector<uint32_t> idx;   // 2 * nmat, 1-based (from, to)
    std::vector<zd> mats;        // d * d * nmat, W_n(a, b) column-major
    // distinct from / to indices (0-based, sorted) and each block's position in them
    std::vector<int> froms, tos;
    std::vector<int> blk_from, blk_to;  // compacted index per block
    // destination-grouped block tables for the mix kernel, per direction
    // (dir 0: source = from, destination = to; dir 1: source = to, destination = from)
    DevBuf<int> mix_ptr[2], mix_src[2];
    DevBuf<double2> mix_W[2];
    DevBuf<float2> mix_W32[2];  // the same blocks rounded to complex64
    void build(int d, int cl, int cr, int nmat, const uint32_t* idx1, const zd* mats);
};

// The host side of MpoSite::build: distinct from / to indices, each block's
// position in them, and per direction the destination-grouped mix tables.
struct MpoTables {
    std::vector<int> froms, tos, blk_from, blk_to;
    std::vector<int> ptr[2], src[2];
    std::vector<zd> W[2];
};
void mpo_mix_tables(int d, int nmat, const uint32_t* idx1, const zd* mats, int cl, int cr, MpoTables& t);

// Mix kernel: Q(e, b, dst) = sum_{n : dst(n) = dst} sum_a W_n(a, b) P(e, a, src(n))
// where e runs over an element slab of size rs.  P has d * nsrc slabs, Q d * ndst.
// Small descriptor lists for a launch: copied through a pinned ring with a
// stream-ordered copy (no host synchronisation); release after the launch.
void* stage_descriptors(const void* host, size_t bytes, hipStream_t s, int* slot);
void release_descriptors(int slot, hipStream_t s);

// The mix step of several chains' matvecs in one launch (src = from blocks).
template <typename CT>
struct MixItem {
    MpoSite const* op;
    const CT* P;
    CT* Q;
    int64_t rs;
    bool src_is_from = true;  // false: the right-environment direction (src = to)
    const int* live = nullptr;  // device flag read at launch: 0 = skip this chain
};
void mix_blocks_batch(std::vector<MixItem<double2>> const& items, hipStream_t s);
void mix_blocks_batch(std::vector<MixItem<float2>> const& items, hipStream_t s);
std::unique_ptr<PreparedLaunch> prepare_mix(std::vector<MixItem<double2>> const& items);
std::unique_ptr<PreparedLaunch> prepare_mix(std::vector<MixItem<float2>> const& items);
// y_q = a_q x_q + b_q y_q for several vectors in one launch (x_q may be null;
// b_q = 0 does not read y_q)
void axpby_batch(std::vector<double2*> const& y, std::vector<zd> const& a, std::vector<const double2*> const& x,
                 std::vector<zd> const& b, std::vector<int64_t> const& n, hipStream_t s);

void mix_blocks(MpoSite const& op, bool src_is_from, const double2* P, double2* Q,
                int64_t rs, hipStream_t s);
// the same with the destination-grouped tables (ndst + 1 offsets, source slab
// per block, blocks W) given as device pointers
void mix_blocks_at(int d, int ndst, const int* dst_ptr, const int* blk_src, const double2* W, const double2* P,
                   double2* Q, int64_t rs, hipStream_t s);
void mix_blocks(MpoSite const& op, bool src_is_from, const float2* P, float2* Q,
                int64_t rs, hipStream_t s);

// complex128 <-> complex64 conversion
void to_c64(const double2* src, float2* dst, int64_t n, hipStream_t s);
void to_c128(const float2* src, double2* dst, int64_t n, hipStream_t s);

// ----------------------------------------------------------------------------
// Operands split once into 16-bit planes for the complex64 matvec on the
// matrix cores.  A plane set holds `count` complex matrices (rows r < R, inner
// index k < K), each as its three real components (re, im, re + im) split into
// parts:
//   * kPlaneBF16: three bf16 parts (hi, mid, lo), nine planes, six partial
//     products per real product;
//   * kPlaneF16: the matrix multiplied by a power of two (scale) that puts its
//     largest |re|, |im|, |re + im| at most 2^14, two fp16 parts (hi, lo), six
//     planes, three partial products.
// A plane is an array of 16-byte chunks, chunk (kb, r) = 8 values of k = 8 kb
// .. 8 kb + 7 in row r at index kb * Rp + r (Rp = R rounded up to kPlaneTile,
// Kb = K rounded up to 32, over 8).  Chunks of k >= K are zero.
// ----------------------------------------------------------------------------
constexpr int kPlaneTile = 128;
constexpr int kPlaneBF16 = 0, kPlaneF16 = 1;
struct PlaneSet {
    DevBuf<uint4> buf;
    DevBuf<float> scale;  // per matrix: the power of two its values were multiplied by (kPlaneF16)
    int R = 0, K = 0, Rp = 0, Kb = 0, count = 0, form = kPlaneBF16, np = 9;
    int64_t ps = 0;  // plane stride in chunks
    void shape(int R, int K, int count, int form);
    uint4* at(int i) const { return buf.ptr + (int64_t)i * np * ps; }
};
// slot[0] = max(slot[0], largest |re|, |im|, |re + im|), slot[1] = max(slot[1],
// largest modulus) of the n contiguous complex128 values at src, as FP32 bit
// patterns rounded up (the slots hold zero or earlier maxima)
void plane_amax(const double2* src, int64_t n, unsigned* slot, hipStream_t s);
// planes of matrices z0 .. z0 + nz - 1 of dst: element (r, k) of matrix z0 + z
// at src[z sz + r sr + k sk] (complex128, rounded to FP32, and split).
// kPlaneF16: multiplied by the scale of the largest value in mslot[0] (one
// scale for the nz matrices, recorded in dst.scale[z0 ..])
void split_planes(const double2* src, int64_t sr, int64_t sk, int64_t sz, int nz, PlaneSet& dst, int z0,
                  const unsigned* mslot, hipStream_t s);
// Q's planes (one matrix per MPO 'to' index t, rows n = l + bl b, K index
// k < br) from step A's output P_{f,a} (nslice K slices pslice apart, summed).
// kPlaneF16 (qs): Q_t is multiplied by S / sR[t] with S the power of two that
// keeps the bound amod_x * C_t of |Q_t| (C_t from the largest column 1-norms n1
// of the L_f and the MPO weights) below 2^14 / sqrt(2) for every t; S / sR[t]
// is written to Q.scale[t].
struct PlaneQScale {
    const float* amod_x;  // largest modulus of x
    const float* n1;      // per distinct 'from' index: largest column 1-norm of L_f
    const float* sR;      // per distinct 'to' index: scale of R_t's planes
};
void mix_split_planes(MpoSite const& op, const float2* P, int64_t pslice, int nslice, int br, int bl, PlaneSet& Q,
                      PlaneQScale const* qs, hipStream_t s);
// out[e] = sum_c Y[c slice + e] in FP64, e < n; reset[0 .. 1] (optional, device)
// zeroed for the next matvec's maxima
void reduce_slices(const float2* Y, int64_t slice, int ns, int64_t n, double2* out, unsigned* reset,
                   hipStream_t s);

// The GEMM on planes: C_b (M x N, complex64, column-major ldc) = sum over the
// batch's nseg segments of A_s B_s, with A's planes in row role (rows = M
// index) and B's in row role too (rows = N index).  Segment s's planes start
// s * a_seg (b_seg) chunks after A (B), and every segment has nks K stages of
// 32.  K is split over nsplit workgroups per tile; slice c is written at
// C + c * slice.  kPlaneF16: the result is divided by a_sc[0] * b_sc[0] (every
// segment's scales have that product).
struct PlaneBatch {
    float2* C;
    int64_t ldc, slice;
    const uint4* A;
    const uint4* B;
    int64_t a_ps, b_ps;      // plane strides (chunks)
    int64_t a_seg, b_seg;    // segment strides (chunks)
    int32_t a_rows, b_rows;  // Rp of the two plane sets
    int32_t M, N;
    int32_t nseg, nks;
    const float* a_sc;
    const float* b_sc;
    int32_t xbm, xbn;  // tiles per XCD block (set by PlaneGemm::prepare; 0: contiguous ranges)
};
struct PlaneGemm {
    DevBuf<PlaneBatch> batches;
    int nsplit = 1, nbatch = 0, form = kPlaneBF16;
    int64_t items = 0;
    void prepare(std::vector<PlaneBatch> const& bs, int nsplit, int form);
    void launch(hipStream_t s) const;
};
// K slices per tile that fill the CUs best for `tiles` tiles of nks K stages
int plane_split_count(int64_t tiles, int nks);

// H_eff x in complex64 on planes (no projectors): L and R split when built,
// x, Q split per application; the result is summed into FP64.  apply() runs
// the five phases below in order.
struct PlaneHeff {
    int bl, br, form;
    MpoSite const& op;
    PlaneSet Lp, Rp, Xp, Qp;
    DevBuf<float2> P, Y;
    DevBuf<unsigned> slots;  // fp16: maxima of x (2), of each L_f (2 each), of each R_t (2 each)
    DevBuf<float> n1;        // fp16: largest column 1-norm of each L_f
    PlaneQScale qs{};
    int nsa = 1, nsc = 1;
    int64_t pslice = 0, yslice = 0;
    PlaneGemm gemm_a, gemm_c;
    double norm_bound = 0;   // fp16: caller's bound on ||x|| (0: x's maxima reduced per matvec)
    PlaneHeff(int bl, int br, MpoSite const& op, const double2* L, const double2* R, hipStream_t s);
    // fp16 planes: every later x has ||x||_2 <= bound, so its scale follows
    // from the bound and no reduction over x runs per matvec
    void assume_input_norm(double bound, hipStream_t s);
    void split_x(const double2* x, hipStream_t s);
    void step_a(hipStream_t s) { gemm_a.launch(s); }
    void mix(hipStream_t s);
    void step_c(hipStream_t s) { gemm_c.launch(s); }
    void sum(double2* y, hipStream_t s);
    void apply(const double2* x, double2* y, hipStream_t s);
};
// Plane form of new PlaneHeff objects: kPlaneF16 (default) or kPlaneBF16.
int& plane_form();
// Mixed-form single-site matvecs: 0 on the complex64 GEMM, 1 on planes when
// bl * br >= 512 * 512 (default), 2 always on planes.
int& mixed_planes_mode();
bool use_planes(int bl, int br);

// ----------------------------------------------------------------------------
// Contractions (FP64, grouped).  All pointers are device pointers.
// ----------------------------------------------------------------------------
struct Workspace {
    ZBuf P, Q;
    DevBuf<float2> P32, Q32;
    void ensure(size_t nP, size_t nQ) {
        if (P.n < nP) P.resize(nP);
        if (Q.n < nQ) Q.resize(nQ);
    }
    void ensure(size_t nP, size_t nQ, double2*) { ensure(nP, nQ); }
    void ensure(size_t nP, size_t nQ, float2*) {
        if (P32.n < nP) P32.resize(nP);
        if (Q32.n < nQ) Q32.resize(nQ);
    }
    double2* p(double2*) { return P.ptr; }
    double2* q(double2*) { return Q.ptr; }
    float2* p(float2*) { return P32.ptr; }
    float2* q(float2*) { return Q32.ptr; }
};

// The grouped contractions in complex64 on FP32 matrix cores (mixed precision
// form): the same three steps as the FP64 functions below.
void heff_matvec(int bl, int br, MpoSite const& op, const float2* L, const float2* R,
                 const float2* x, float2* y, Workspace& ws, hipStream_t s);
void env_left(int bl, int br, MpoSite const& op, const float2* L, const float2* psi,
              float2* newL, Workspace& ws, hipStream_t s);
void env_right(int bl, int br, MpoSite const& op, const float2* R, const float2* psi,
               float2* newR, Workspace& ws, hipStream_t s);

// y = H_eff x (no projectors), the product of core.f95's iteration_stage_1/2/3.
void heff_matvec(int bl, int br, MpoSite const& op, const double2* L, const double2* R,
                 const double2* x, double2* y, Workspace& ws, hipStream_t s);
// New left environment (br, br, cr) from L (bl, bl, cl) and psi (core.f95:623-677).
void env_left(int bl, int br, MpoSite const& op, const double2* L, const double2* psi,
              double2* newL, Workspace& ws, hipStream_t s);
// New right environment (bl, bl, cl) from R (br, br, cr) and psi (core.f95:679-848).
void env_right(int bl, int br, MpoSite const& op, const double2* R, const double2* psi,
               double2* newR, Workspace& ws, hipStream_t s);
// <psi|H|psi> at a site (core.f95:1045-1088).
zd expectation(int bl, int br, MpoSite const& op, const double2* L, const double2* psi,
               const double2* R, Workspace& ws, hipStream_t s);

// ----------------------------------------------------------------------------
// Vector kernels (Arnoldi / projectors), device pointers, length n.
// ----------------------------------------------------------------------------
// h[j] = sum_i conj(V[j*ldv + i]) w[i], j < m   (result to host)
void zdotc_multi(const double2* V, int64_t ldv, int m, const double2* w, int64_t n, zd* h_host,
                 hipStream_t s);
// w -= sum_j V_j h[j]
void zgemv_sub(const double2* V, int64_t ldv, int m, const zd* h, double2* w, int64_t n,
               hipStream_t s);
double dznrm2(const double2* x, int64_t n, hipStream_t s);
zd zdotu(const double2* x, const double2* y, int64_t n, hipStream_t s);  // sum x*y (no conj)
// x -= tau (sum_i x_i v_i) vc, stream-ordered with no host read; partial holds
// refl_partials() entries and must not be shared with concurrent work
void zrefl_apply(double2* x, const double2* v, const double2* vc, zd tau, int64_t n, double2* partial,
                 hipStream_t s);
int refl_partials();
void zscal(double2* x, zd a, int64_t n, hipStream_t s);
void zaxpy(double2* y, zd a, const double2* x, int64_t n, hipStream_t s);  // y += a x
// y = a x + b y  (general combination)
void zaxpby(double2* y, zd a, const double2* x, zd b, int64_t n, hipStream_t s);
// out = V c (m columns)
void zgemv_n(const double2* V, int64_t ldv, int m, const zd* c, double2* out, int64_t n,
             hipStream_t s);
void zconj_copy(double2* dst, const double2* src, int64_t n, hipStream_t s);
// permute a (n0, n1, n2) tensor to (n1, n0, n2)
void permute_102(const double2* src, double2* dst, int64_t n0, int64_t n1, int64_t n2,
                 hipStream_t s);
// scale rows: X(i, j) *= sv[i] for an (m x ncol) matrix
void scale_rows(double2* X, const double* sv, int64_t m, int64_t ncol, hipStream_t s);
void conj_inplace(double2* X, int64_t n, hipStream_t s);
// deterministic pseudo-random entries, uniform in [-1, 1) + i [-1, 1), from `key`
void fill_random(double2* x, int64_t n, uint64_t key, hipStream_t s);
// add alpha * I to an (m x m) matrix
// sum_k sum_ij A(i, j, k) B(j, i, k)
zd trace_pair(const double2* A, const double2* B, int64_t b, int64_t c, hipStream_t s);

// ----------------------------------------------------------------------------
// Projectors: Householder reflectors (LAPACK storage) applied to vectors.
// ----------------------------------------------------------------------------
struct ProjectorDevice;
struct Projector {
    int n = 0, nproj = 0, nrefl = 0, orth = 0;
    std::vector<uint32_t> swaps;  // nrefl, 1-based
    std::vector<zd> coeffs;       // nrefl
    DevBuf<double2> d_refl;       // n * nproj
    std::shared_ptr<ProjectorDevice> dev;  // explicit reflector vectors
    bool active() const { return nproj > 0; }
    // reflectors from the host (refl) or already on the device (d_refl_in)
    void set(int n, int nproj, int nrefl, int orth, const zd* refl, const zd* coeffs,
             const uint32_t* swaps);
    void set_device(int n, int nproj, int nrefl, int orth, const double2* d_refl_in, const zd* coeffs,
                    const uint32_t* swaps);
    // full (n) -> orthogonal (orth), and back
    void project(const double2* full, double2* out, double2* tmp, hipStream_t s) const;
    void unproject(const double2* in, double2* full, hipStream_t s) const;
    double overlap(const double2* v, double2* tmp, hipStream_t s) const;

  private:
    void finish(const std::vector<zd>& rtop, hipStream_t s);
};

// Overlap contractions and projector formation (overlap.cpp)
void vs_left(int64_t blo, int64_t bro, int64_t bln, int64_t brn, int64_t d, const double2* L, const double2* P,
             const double2* S, double2* out, hipStream_t s);
void vs_right(int64_t blo, int64_t bro, int64_t bln, int64_t brn, int64_t d, const double2* R, const double2* P,
              const double2* S, double2* out, hipStream_t s);
void overlap_vector(int64_t blo, int64_t bro, int64_t bln, int64_t brn, int64_t d, const double2* L,
                    const double2* R, const double2* P, double2* out, hipStream_t s);
void overlap_site(int64_t br, int64_t bl, int64_t d, const double2* S, double2* out, hipStream_t s);
int norm_overlap_tensors(int bl, int bm, int br, int dl, int dr, const double2* unnorm1, const double2* rightnorm2,
                         double2* left_ov1, double2* unnorm_ov1, double2* unnorm2, double2* right_ov2,
                         hipStream_t s);
int reflectors_from_vectors(int64_t m, int n, double2* V, zd* tau, uint32_t* swaps, hipStream_t s);

// ----------------------------------------------------------------------------
// Eigen-solver (implicitly restarted Arnoldi, nev = 1) with device vectors.
// ----------------------------------------------------------------------------
struct Operator {
    int64_t n = 0;
    virtual void apply(const double2* x, double2* y, hipStream_t s) = 0;
    virtual ~Operator() = default;
};

struct IramResult {
    int info = 0;        // 0 ok, -14 no converged Ritz value, other < 0 failures
    int restarts = 0;    // ARPACK iparam(3)
    int64_t matvecs = 0;
    zd eigenvalue{0, 0};
};

// Largest Krylov basis the solver takes (its per-step coefficient tables hold
// ncv + 2 entries in 64 lanes).
constexpr int kMaxNcv = 62;

// which: "SR", "LR", "LM".  guess / result are device vectors of length op.n.
IramResult iram(Operator& op, int ncv, const char* which, double tol, int max_restarts,
                const double2* guess, double2* result, hipStream_t s);

// Several independent eigen-problems (one per chain) solved in lockstep: each
// Arnoldi step applies every active operator (through `apply`, which may batch
// them into grouped launches) and orthogonalises all of them with the same
// kernel launches.  All tasks share ncv.
struct IramTask {
    Operator* op = nullptr;
    int ncv = 3;
    const char* which = "SR";
    double tol = 0;
    int max_restarts = 1;
    const double2* guess = nullptr;
    double2* result = nullptr;
    IramResult out;
};
struct BatchApply {
    // y_q = A_{who[q]} x_q for every listed problem
    std::function<void(std::vector<int> const&, std::vector<const double2*> const&, std::vector<double2*> const&,
                       hipStream_t)>
        fn;
    void operator()(std::vector<int> const& who, std::vector<const double2*> const& xs,
                    std::vector<double2*> const& ys, hipStream_t s) const {
        fn(who, xs, ys, s);
    }
    // one operator application after another
    static BatchApply each(std::vector<IramTask> const& tasks);

    // Persistent form (optional; used when `launch` is set).  Problem p's operator
    // reads the fixed vector in[p] and writes out[p] (complex64 when `single`,
    // complex128 otherwise).  bind(live, s) prepares launches for every problem,
    // each skipped while the device flag live[p] is 0; launch(s) runs them for
    // all problems with no host work, so the Krylov solver can record whole
    // Arnoldi extensions in a graph.
    std::vector<void*> in, out;
    bool single = false;
    std::function<void(const int* live, hipStream_t)> bind;
    std::function<void(hipStream_t)> launch;
    bool persistent() const { return (bool)launch; }
};
// Launch options of the Krylov solver: record each Arnoldi extension of a
// persistent batch in a HIP graph and replay it (default on; the environment
// variable NC_KRYLOV_GRAPHS=0 makes direct launches the default).
bool& krylov_graphs();
// first Gram-Schmidt pass against the last two basis vectors only (default on;
// NC_GS_LOCAL=0 in the environment makes it a full pass)
bool& gs_local_first();
// second Gram-Schmidt update deferred into the next step's fused kernel where
// the second pass's coefficients are small (default on; NC_GS_DEFER=0 in the
// environment turns it off)
bool& gs_defer_second();
// share of the basis (percent) a thick restart keeps for ncv > 3
int& thick_keep_percent();
void iram_batch(std::vector<IramTask>& tasks, BatchApply const& apply, hipStream_t s);

// Several reductions with one launch pair and one host read: an item is
// sum conj(x) y over n elements, or (trace) sum_{i,j,k} x(i,j,k) y(j,i,k) with
// x, y of shape n x n x c.  Deterministic (fixed summation order).
struct RedItem {
    const double2* x;
    const double2* y;
    int64_t n;
    int64_t c = 1;
    bool trace = false;
};
std::vector<zd> reduce_batch(std::vector<RedItem> const& items, hipStream_t s);

// Runs fn(i) for i in [0, n) on a small pool of host threads (the caller
// included; serially for small n or when another parallel region is running).
// fn must not call HIP.
void host_parallel_for(int n, std::function<void(int)> const& fn);

// Small dense host eigen-problems.
// Eigenvalues and right eigenvectors of a general complex upper Hessenberg matrix.
void hessenberg_eig(int m, std::vector<zd> const& H, std::vector<zd>& w, std::vector<zd>& V);
// Hermitian (upper triangle) eigen-decomposition, ascending eigenvalues.
void hermitian_eig(int m, std::vector<zd> const& A, std::vector<double>& w, std::vector<zd>& V);

// ----------------------------------------------------------------------------
// optimize_ (core.f95:1911-2609) on the device.
// ----------------------------------------------------------------------------
struct OptimizeArgs {
    int bl, br, cl, cr, d;
    const double2* L;  // device
    const double2* R;  // device
    MpoSite const* op;
    Projector const* proj;  // may be inactive
    const char* which;
    double tol;
    int max_iterations;
    const double2* guess;  // device, full space
    double2* result;       // device, full space
    int ncv = 3;
    // mixed-precision form: the Krylov iteration applies H_eff in complex64 on
    // FP32 matrix cores with these environment copies, to tolerance tol32; the
    // eigenvalue is then the FP64 Rayleigh quotient of the result, and with
    // `polish` the FP64 iteration continues from it to `tol`
    const float2* L32 = nullptr;
    const float2* R32 = nullptr;
    bool mixed = false;
    double tol32 = 1e-7;
    bool polish = false;
};
struct OptimizeOut {
    int info = 0;
    int iterations = 0;
    int strategy = 0;
    int64_t matvecs = 0;    // FP64 operator applications
    int64_t matvecs32 = 0;  // complex64 operator applications (mixed form)
    zd eigenvalue{0, 0};
    double normal = 0;
};
// Site solves of several independent chains in lockstep (grouped launches).
// krylov_ms, when given, accumulates the wall time of the Krylov iterations.
void optimize_sites_batched(std::vector<OptimizeArgs const*> const& args, std::vector<Workspace*> const& wss,
                            std::vector<OptimizeOut>& outs, hipStream_t s, double* krylov_ms = nullptr);
OptimizeOut optimize_site(OptimizeArgs const& a, Workspace& ws, hipStream_t s);
// Counters of optimize_site since load: solves, first Krylov runs that ended
// with -14, ncv retries run (core.f95:2588-2602), dense strategy-1 fallbacks
// (2569-2587), solves returned with -14.
struct OptimizerStats {
    int64_t solves = 0, first_run_failures = 0, ncv_retries = 0, dense_fallbacks = 0, unconverged = 0;
};
OptimizerStats optimizer_stats();
// Krylov basis size of the drop-in optimize_ (core.f95:2567 hard-wires 3)
int& optimizer_ncv();

// ----------------------------------------------------------------------------
// Site move: polar split of an (m x ncol) matrix, m <= ncol.
//   U_p = U Vh (rows orthonormal), P = U S U^H, so that A = P U_p.
// Returns the SVD info code (0 on success).  ``Pconj`` receives conj(P) (m x m).
// ----------------------------------------------------------------------------
int polar_split(int m, int64_t ncol, const double2* A, double2* Up, double2* Pconj, hipStream_t s);
// Several polar splits (one per chain) advanced together: every Newton-Schulz
// step of all of them is two grouped GEMM launches and one host read.
struct PolarJob {
    int m = 0;
    int64_t n = 0;
    const double2* A = nullptr;
    double2* Up = nullptr;
    double2* Pconj = nullptr;
    int info = 0;
};
void polar_split_batch(std::vector<PolarJob>& jobs, hipStream_t s);
// sum |x|^2 of several vectors, and x *= f (real) for several vectors
std::vector<double> sumsq_batch(std::vector<const double2*> const& x, std::vector<int64_t> const& n, hipStream_t s);
void scale_batch(std::vector<double2*> const& x, std::vector<int64_t> const& n, std::vector<double> const& f,
                 hipStream_t s);
// the same through rocSOLVER's SVD (used when the iteration cannot converge)
int polar_split_svd(int m, int64_t ncol, const double2* A, double2* Up, double2* Pconj, hipStream_t s);
struct PolarStats {
    int64_t calls = 0, iterations = 0, svd_fallbacks = 0;
};
PolarStats polar_stats();
// Newton-Schulz steps before a polar split is completed by the SVD (default 64)
int& polar_max_steps();

// norm_denorm_going_left / right on device tensors (core.f95:2708-2878).
int norm_denorm_left(int bl, int bm, int br, int dl, int dr, const double2* to_denorm,
                     const double2* to_norm, double2* denorm_out, double2* norm_out, hipStream_t s);
int norm_denorm_right(int bl, int bm, int br, int dl, int dr, const double2* to_norm,
                      const double2* to_denorm, double2* norm_out, double2* denorm_out,
                      hipStream_t s);

}  // namespace nc
