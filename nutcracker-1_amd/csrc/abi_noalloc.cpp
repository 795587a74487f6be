// contract_sos_left_ for callers that cannot allocate device memory (code under
// HIP graph capture, tight serving loops): the environment, site tensor and
// output are the caller's device buffers, and the work space of the grouped
// contraction -- the MPO mix tables, the GEMM descriptor lists and the two
// slabs P and Q (core.f95:623-677 restated as step A, the block mix and step
// C) -- lives in one file-scope device array of fixed size.  Nothing is
// allocated: the small host tables reach the device through kernel arguments,
// the GEMMs take the default launch choice with no K split and no autotuning,
// and unused output slices are cleared with a fill kernel.
//
// Each GPU holds its own copy of the array for the life of the process; every
// call on a GPU uses that GPU's copy, so calls that overlap on one GPU (two
// streams, two threads) race and must be ordered by the caller.
#include "nc_abi.hpp"
#include "../../include/nutcracker_core_abi.h"

#include <mutex>

using namespace nc;

namespace {

constexpr size_t kScratchBytes = size_t(64) << 20;  // per GPU
__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch[kMaxDevices];

// The device address of the current GPU's scratch array (the host-side name is
// not a device address).  Looked up once per GPU, with that GPU current.
unsigned char* scratch_on_current_device() {
    int dev = 0;
    NC_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) throw Error("device index out of range", -91);
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr)
            throw Error(std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e), -92);
        g_scratch[dev] = p;
    }
    return static_cast<unsigned char*>(g_scratch[dev]);
}

size_t up256(size_t b) { return (b + 255) & ~size_t(255); }

}  // namespace

extern "C" {

int nc_contract_sos_left_noalloc(uint32_t bl, uint32_t br, uint32_t cl, uint32_t cr, uint32_t d,
                                 const nc_zcomplex* left_environment, uint32_t number_of_matrices,
                                 const uint32_t* sparse_operator_indices, const nc_zcomplex* sparse_operator_matrices,
                                 const nc_zcomplex* state_site_tensor, nc_zcomplex* new_left_environment,
                                 void* stream) {
    return guarded("nc_contract_sos_left_noalloc", [&] {
        const hipStream_t s = static_cast<hipStream_t>(stream);
        if (bl == 0 || br == 0) throw Error("bond dimensions must be positive", -2);
        MpoTables t;
        mpo_mix_tables((int)d, (int)number_of_matrices, sparse_operator_indices,
                       reinterpret_cast<const zd*>(sparse_operator_matrices), (int)cl, (int)cr, t);
        const int64_t rs = (int64_t)br * bl;
        const int nf = (int)t.froms.size(), nt = (int)t.tos.size();
        const double2* L = reinterpret_cast<const double2*>(left_environment);
        const double2* psi = reinterpret_cast<const double2*>(state_site_tensor);
        double2* newL = reinterpret_cast<double2*>(new_left_environment);
        // layout of the scratch array
        const size_t b_ptr = (t.ptr[0].size()) * sizeof(int), b_src = t.src[0].size() * sizeof(int),
                     b_w = t.W[0].size() * sizeof(double2);
        const size_t o_ptr = 0, o_src = up256(o_ptr + b_ptr), o_w = up256(o_src + b_src);
        const size_t o_desc_a = up256(o_w + b_w);
        unsigned char* base = scratch_on_current_device();
        // step A: P_{f,a} = psi_a L_f; the mix: Q_t = sum_n W_n P_{from_n}; step C:
        // newL_t = Q_t psi^H -- the three steps of env_left
        auto plans = [&](double2* P, double2* Q, GemmPlan& A, GemmPlan& Cp) {
            A = GemmPlan();
            A.M = (int)br; A.N = (int)bl; A.opA = OpN; A.opB = OpN;
            for (int fi = 0; fi < nf; ++fi)
                for (uint32_t a = 0; a < d; ++a) {
                    A.begin_batch(P + ((int64_t)fi * d + a) * rs, br);
                    A.add_seg(psi + a * rs, br, L + (int64_t)t.froms[fi] * bl * bl, bl, (int)bl);
                }
            Cp = GemmPlan();
            Cp.M = (int)br; Cp.N = (int)br; Cp.opA = OpN; Cp.opB = OpC;
            for (int ti = 0; ti < nt; ++ti) {
                Cp.begin_batch(newL + (int64_t)t.tos[ti] * br * br, br);
                Cp.add_seg(Q + (int64_t)ti * d * rs, br, psi, br, (int)(bl * d));
            }
        };
        // the descriptor sizes do not depend on the slab addresses
        GemmPlan A, Cp;
        double2* const any = reinterpret_cast<double2*>(base);
        plans(any, any, A, Cp);
        const size_t b_desc_a = zgemm_desc_bytes(A), b_desc_c = zgemm_desc_bytes(Cp);
        const size_t o_desc_c = up256(o_desc_a + b_desc_a);
        const size_t o_p = up256(o_desc_c + b_desc_c);
        const size_t b_p = (size_t)rs * d * nf * sizeof(double2), b_q = (size_t)rs * d * nt * sizeof(double2);
        const size_t o_q = up256(o_p + b_p);
        const size_t need = o_q + b_q;
        if (need > kScratchBytes) {
            char msg[256];
            snprintf(msg, sizeof msg, "work space of %zu bytes exceeds the %zu-byte scratch array (bl %u, br %u, d %u, "
                     "%d from / %d to blocks)", need, kScratchBytes, bl, br, d, nf, nt);
            throw Error(msg, -90);
        }
        int* d_ptr = reinterpret_cast<int*>(base + o_ptr);
        int* d_src = reinterpret_cast<int*>(base + o_src);
        double2* d_w = reinterpret_cast<double2*>(base + o_w);
        double2* P = reinterpret_cast<double2*>(base + o_p);
        double2* Q = reinterpret_cast<double2*>(base + o_q);
        plans(P, Q, A, Cp);
        store_bytes(d_ptr, t.ptr[0].data(), b_ptr, s);
        store_bytes(d_src, t.src[0].data(), b_src, s);
        store_bytes(d_w, t.W[0].data(), b_w, s);
        zgemm_run_at(A, base + o_desc_a, b_desc_a, s);
        mix_blocks_at((int)d, nt, d_ptr, d_src, d_w, P, Q, rs, s);
        // slices of the new environment that no block reaches are zero
        for (uint32_t k = 0, ti = 0; k < cr; ++k) {
            if (ti < (uint32_t)nt && t.tos[ti] == (int)k) { ++ti; continue; }
            zero_fill(newL + (int64_t)k * br * br, (int64_t)br * br, s);
        }
        zgemm_run_at(Cp, base + o_desc_c, b_desc_c, s);
        return 0;
    });
}

}  // extern "C"
