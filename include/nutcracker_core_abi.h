/*
 * nutcracker_core_abi.h -- the C-ABI of the MI355X variational-MPS core.
 *
 * The first group re-declares, with the same names and argument meaning, the
 * gfortran-convention entry points that the reference's C++ layer binds in
 * sources/common/core.cpp (every scalar passed by pointer, complex128 arrays in
 * column-major order, 1-based MPO block indices and swap indices).  Linking the
 * reference's host code against libnutcracker_hip.so instead of its Fortran core
 * therefore needs no source change.  Each declaration cites the reference
 * declaration it replaces and the Fortran routine whose behaviour it follows.
 *
 * The second group (nc_*) has no counterpart in the reference: device
 * information, the last error message, the projector helpers that core.f95 keeps
 * internal, and the device-resident sweep driver.
 *
 * Complex numbers are passed as nc_zcomplex, which has the layout of
 * std::complex<double> and Fortran double complex.
 */
#ifndef NUTCRACKER_CORE_ABI_H
#define NUTCRACKER_CORE_ABI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct nc_zcomplex {
    double re, im;
} nc_zcomplex;

/* ---- environment contractors ------------------------------------------- */

/* core.cpp:235-245; core.f95:623-677.  new_left_environment(br,br,cr) from
   left_environment(bl,bl,cl) and state_site_tensor(br,bl,d). */
void contract_sos_left_(const uint32_t* bl, const uint32_t* br, const uint32_t* cl, const uint32_t* cr,
                        const uint32_t* d, const nc_zcomplex* left_environment,
                        const uint32_t* number_of_matrices, const uint32_t* sparse_operator_indices,
                        const nc_zcomplex* sparse_operator_matrices, const nc_zcomplex* state_site_tensor,
                        nc_zcomplex* new_left_environment);

/* core.cpp:272-282; core.f95:679-848.  new_right_environment(bl,bl,cl). */
void contract_sos_right_(const uint32_t* bl, const uint32_t* br, const uint32_t* cl, const uint32_t* cr,
                         const uint32_t* d, const nc_zcomplex* right_environment,
                         const uint32_t* number_of_matrices, const uint32_t* sparse_operator_indices,
                         const nc_zcomplex* sparse_operator_matrices, const nc_zcomplex* state_site_tensor,
                         nc_zcomplex* new_right_environment);

/* core.cpp:33-44; core.f95:1045-1088. */
void compute_expectation_(const uint32_t* bl, const uint32_t* br, const uint32_t* cl, const uint32_t* cr,
                          const uint32_t* d, const nc_zcomplex* left_environment,
                          const nc_zcomplex* state_site_tensor, const uint32_t* number_of_matrices,
                          const uint32_t* sparse_operator_indices, const nc_zcomplex* sparse_operator_matrices,
                          const nc_zcomplex* right_environment, nc_zcomplex* expectation);

/* core.cpp:179-193; core.f95:1142-1161. */
void contract_expectation_boundaries_(const uint32_t* b, const uint32_t* c, const nc_zcomplex* left_boundary,
                                      const nc_zcomplex* right_boundary, nc_zcomplex* expectation);

/* core.cpp:72-84; core.f95:1090-1140.  optimization_matrix is (n, n), n = br*bl*d. */
void compute_optimization_matrix_(const uint32_t* bl, const uint32_t* br, const uint32_t* cl, const uint32_t* cr,
                                  const uint32_t* d, const nc_zcomplex* left_environment,
                                  const uint32_t* number_of_matrices, const uint32_t* sparse_operator_indices,
                                  const nc_zcomplex* sparse_operator_matrices, const nc_zcomplex* right_environment,
                                  nc_zcomplex* optimization_matrix);

/* ---- matvec stages (internal Fortran routines core.f95:532-621, bound by the
   reference's f2py tests, tests/python/test.py:245-328) -------------------- */

void iteration_stage_1_(const uint32_t* bl, const uint32_t* cl, const uint32_t* cr, const uint32_t* d,
                        const nc_zcomplex* left_environment, const uint32_t* number_of_matrices,
                        const uint32_t* sparse_operator_indices, const nc_zcomplex* sparse_operator_matrices,
                        nc_zcomplex* iteration_stage_1_tensor);
void iteration_stage_2_(const uint32_t* bl, const uint32_t* br, const uint32_t* cr, const uint32_t* d,
                        const nc_zcomplex* iteration_stage_1_tensor, const nc_zcomplex* state_site_tensor,
                        nc_zcomplex* iteration_stage_2_tensor);
void iteration_stage_3_(const uint32_t* bl, const uint32_t* br, const uint32_t* cr, const uint32_t* d,
                        const nc_zcomplex* iteration_stage_2_tensor, const nc_zcomplex* right_environment,
                        nc_zcomplex* output_state_site_tensor);

/* ---- site optimizer ------------------------------------------------------ */

/* core.cpp:676-693; core.f95:1911-2609.  Returns info: 0 success, -14 no
   converged eigenvalue, 5 / 10 / 11 projector-space errors, other negative
   values for device failures.  `which` is two characters ("SR", "LR", "LM")
   with no hidden length.  number_of_iterations: in = maximum restarts,
   out = restarts performed (ARPACK iparam(3)). */
uint32_t optimize_(const uint32_t* bl, const uint32_t* br, const uint32_t* cl, const uint32_t* cr, const uint32_t* d,
                   const nc_zcomplex* left_environment, const uint32_t* number_of_matrices,
                   const uint32_t* sparse_operator_indices, const nc_zcomplex* sparse_operator_matrices,
                   const nc_zcomplex* right_environment, const uint32_t* number_of_projectors,
                   const uint32_t* number_of_reflectors, const uint32_t* orthogonal_subspace_dimension,
                   const nc_zcomplex* reflectors, const nc_zcomplex* coefficients, const uint32_t* swaps,
                   const char* which, const double* tol, uint32_t* number_of_iterations,
                   const nc_zcomplex* guess, nc_zcomplex* result, nc_zcomplex* eigenvalue, double* normal);

/* ---- site move ------------------------------------------------------------ */

/* core.cpp:580-587; core.f95:2708-2788.  Returns the SVD info (0 on success). */
int32_t norm_denorm_going_left_(const uint32_t* bll, const uint32_t* bl, const uint32_t* br, const uint32_t* dl,
                                const uint32_t* d, const nc_zcomplex* site_tensor_to_denormalize,
                                const nc_zcomplex* site_tensor_to_normalize, nc_zcomplex* denormalized_site_tensor,
                                nc_zcomplex* normalized_site_tensor);

/* core.cpp:609-616; core.f95:2790-2878. */
int32_t norm_denorm_going_right_(const uint32_t* bl, const uint32_t* br, const uint32_t* brr, const uint32_t* d,
                                 const uint32_t* dr, const nc_zcomplex* site_tensor_to_normalize,
                                 const nc_zcomplex* site_tensor_to_denormalize, nc_zcomplex* normalized_site_tensor,
                                 nc_zcomplex* denormalized_site_tensor);

/* ---- projectors ---------------------------------------------------------- */

/* core.cpp:109-118; core.f95:1266-1339.  Writes a double into the first eight
   bytes of *overlap, as the Fortran routine does. */
void compute_overlap_with_projectors_(const uint32_t* number_of_projectors, const uint32_t* number_of_reflectors,
                                      const nc_zcomplex* reflectors, const nc_zcomplex* coefficients,
                                      const uint32_t* swaps, const uint32_t* vector_size, const nc_zcomplex* vector,
                                      const nc_zcomplex* overlap);

/* core.cpp:431-441; core.f95:1866-1909. */
void filter_components_outside_orthog_(const uint32_t* full_space_dimension, const uint32_t* number_of_projectors,
                                       const uint32_t* number_of_reflectors,
                                       const uint32_t* orthogonal_subspace_dimension, const nc_zcomplex* reflectors,
                                       const nc_zcomplex* coefficients, const uint32_t* swaps,
                                       const nc_zcomplex* input, const nc_zcomplex* output);

/* core.cpp:309-324; core.f95:850-896. */
void contract_vs_left_(const uint32_t* b_left_old, const uint32_t* b_right_old, const uint32_t* b_left_new,
                       const uint32_t* b_right_new, const uint32_t* d, const nc_zcomplex* left_environment,
                       const nc_zcomplex* normalized_projector_site_tensor,
                       const nc_zcomplex* normalized_state_site_tensor, nc_zcomplex* new_left_environment);

/* core.cpp:338-348; core.f95:898-952. */
void contract_vs_right_(const uint32_t* b_left_old, const uint32_t* b_right_old, const uint32_t* b_left_new,
                        const uint32_t* b_right_new, const uint32_t* d, const nc_zcomplex* right_environment,
                        const nc_zcomplex* normalized_projector_site_tensor,
                        const nc_zcomplex* normalized_state_site_tensor, nc_zcomplex* new_right_environment);

/* core.cpp:517-525; core.f95:954-1005.  overlap_vector(b_right_new, b_left_new, d)
   from left_environment(b_left_new, b_left_old), right_environment(b_right_old,
   b_right_new) and the unnormalised overlap site (b_left_old, d, b_right_old). */
void form_overlap_vector_(const uint32_t* b_left_old, const uint32_t* b_right_old, const uint32_t* b_left_new,
                          const uint32_t* b_right_new, const uint32_t* d, const nc_zcomplex* left_environment,
                          const nc_zcomplex* right_environment, const nc_zcomplex* unnormalized_projector_site_tensor,
                          nc_zcomplex* overlap_vector);

/* core.cpp:371-378; core.f95:1341-1435.  n = vector length, m = number of
   vectors; vectors(n, m) is overwritten by its column-pivoted Householder QR,
   coefficients and swaps have min(n, m) entries, *rank = number of non-null
   reflectors (also returned, as the reference's declaration has it). */
uint32_t convert_vectors_to_reflectors_(const uint32_t* n, const uint32_t* m, nc_zcomplex* vectors, uint32_t* rank,
                                        nc_zcomplex* coefficients, uint32_t* swaps);

/* core.cpp:498-504; core.f95:3299-3313.  overlap(bl, d, br) = conj(state(br, bl, d)). */
void form_overlap_site_tensor_(const uint32_t* br, const uint32_t* bl, const uint32_t* d,
                               const nc_zcomplex* state_site_tensor, nc_zcomplex* overlap_site_tensor);

/* core.cpp:463-472; core.f95:3315-3373.  Left-normalises state 1 (bm, bl, dl) into
   state 2 (br, bm, dr) and forms the three overlap sites.  (The reference's C++
   declaration marks the four outputs const; they are written.) */
void form_norm_overlap_tensors_(const uint32_t* bl, const uint32_t* bm, const uint32_t* br, const uint32_t* dl,
                                const uint32_t* dr, const nc_zcomplex* unnormalized_state_tensor_1,
                                const nc_zcomplex* right_norm_state_tensor_2, nc_zcomplex* left_norm_overlap_tensor_1,
                                nc_zcomplex* unnormalized_overlap_tensor_1, nc_zcomplex* unnormalized_state_tensor_2,
                                nc_zcomplex* right_norm_overlap_tensor_2);

/* ---- random tensors, normalisation, bond growth, boundaries --------------
   The random routines draw from a device hash keyed by a per-process call
   counter instead of gfortran's rand() stream (core.f95:2617-2632): the
   numbers differ, the properties (normalisation, orthonormality, rank) hold. */

/* core.cpp:735-750; core.f95:2634-2659.  Random site (br, bl, d) with
   sum_{r,s} psi(r,l,s) conj(psi(r,l',s)) = delta(l,l'); needs br*d >= bl. */
void rand_norm_state_site_tensor_(const uint32_t* br, const uint32_t* bl, const uint32_t* d,
                                  nc_zcomplex* state_site_tensor);

/* core.cpp:752-767; core.f95:2661-2671.  Random site of unit Frobenius norm. */
void rand_unnorm_state_site_tensor_(const uint32_t* br, const uint32_t* bl, const uint32_t* d,
                                    nc_zcomplex* state_site_tensor);

/* core.cpp:769-783; core.f95:2673-2706.  Random vectors (length
   projector_length, number_of_projectors of them) converted to Householder
   reflectors as convert_vectors_to_reflectors_ does. */
void random_projector_matrix_(const uint32_t* projector_length, const uint32_t* number_of_projectors,
                              uint32_t* rank, nc_zcomplex* reflectors, nc_zcomplex* coefficients, uint32_t* swaps);

/* core.cpp:638-655; core.f95:2880-2917.  The unitary polar factor of the site
   as a (br x bl*d) matrix.  Returns the SVD info (0 on success). */
int32_t norm_for_left_(const uint32_t* br, const uint32_t* bl, const uint32_t* d, const nc_zcomplex* site_tensor,
                       nc_zcomplex* normalized_site_tensor);

/* core.cpp:657-674; core.f95:2919-2960.  The same for the (bl x br*d) matrix
   of the transposed site. */
int32_t norm_for_right_(const uint32_t* br, const uint32_t* bl, const uint32_t* d, const nc_zcomplex* site_tensor,
                        nc_zcomplex* normalized_site_tensor);

/* core.cpp:548-578; core.f95:3090-3144.  Widens the bond between a left site
   (bm, bl, dl) and a right site (br, bm, dr) to new_bm with a random isometry
   and re-normalises the right site (norm_denorm_going_left).  Outputs, in
   argument order: the left site (new_bm, bl, dl), the right site
   (br, new_bm, dr).  Returns the SVD info. */
int32_t increase_bandwidth_between_(const uint32_t* bl, const uint32_t* bm, const uint32_t* br, const uint32_t* dl,
                                    const uint32_t* dr, const uint32_t* new_bm,
                                    const nc_zcomplex* site_tensor_to_normalize,
                                    const nc_zcomplex* site_tensor_to_denormalize,
                                    nc_zcomplex* normalized_site_tensor, nc_zcomplex* denormalized_site_tensor);

/* core.cpp:785-819; core.f95:3246-3297.  For a random new_b x b isometry M:
   state_s -> conj(M) state_s M^H, left_k -> M left_k M^H,
   right_k -> conj(M) right_k M^T. */
void increase_bandwidth_with_environment_(const uint32_t* b, const uint32_t* c, const uint32_t* d,
                                          const uint32_t* new_b, const nc_zcomplex* old_state_site_tensor,
                                          const nc_zcomplex* old_left_exp_environment,
                                          const nc_zcomplex* old_right_exp_environment,
                                          nc_zcomplex* new_state_site_tensor, nc_zcomplex* new_left_exp_environment,
                                          nc_zcomplex* new_right_exp_environment);

/* core.cpp:139-159; core.f95:3402-3418.  L(i,j,k) = s(i) conj(s(j)) o(k). */
void construct_left_exp_boundary_(const uint32_t* b, const uint32_t* c, const nc_zcomplex* state_boundary,
                                  const nc_zcomplex* operator_boundary, nc_zcomplex* left_expectation_boundary);

/* core.cpp:161-181; core.f95:3420-3436.  R(i,j,k) = conj(s(i)) s(j) o(k). */
void construct_right_exp_boundary_(const uint32_t* b, const uint32_t* c, const nc_zcomplex* state_boundary,
                                   const nc_zcomplex* operator_boundary, nc_zcomplex* right_expectation_boundary);

/* core.cpp:210-233; core.f95:1195-1234.  new_left(br, br) = X L X^H for
   X = matrix (br, bl), L = left_environment (bl, bl). */
void contract_matrix_left_(const uint32_t* bl, const uint32_t* br, const nc_zcomplex* left_environment,
                           const nc_zcomplex* matrix, nc_zcomplex* new_left_environment);

/* core.cpp:400-429; core.f95:1163-1193.
   new(r, s, l) = sum_m site(r, m, s) old(m, l), new (br, dr, dl). */
void extend_state_vector_fragment_(const uint32_t* bm, const uint32_t* br, const uint32_t* dl, const uint32_t* dr,
                                   const nc_zcomplex* old_state_vector_fragment, const nc_zcomplex* site_tensor,
                                   nc_zcomplex* new_state_vector_fragment);

/* ---- entry points with no counterpart in the reference --------------------- */

/* Message of the last failed call on this thread (the reference's C API keeps
   a thread-local error string, sources/c-interface/errors.cpp:7-25). */
const char* nc_last_error(void);

/* contract_sos_left_ for callers that cannot allocate device memory (code under
   HIP graph capture, serving loops that forbid hipMalloc).  No counterpart in
   the reference.  left_environment (bl, bl, cl), state_site_tensor (br, bl, d)
   and new_left_environment (br, br, cr) are DEVICE pointers; the block list
   (indices, matrices) is read from host memory during the call.  The work is
   queued on `stream` (a hipStream_t; NULL is HIP's null stream) and the call
   returns without waiting.  The work space lives in a fixed 64 MiB device array
   per GPU (about br*bl*d*(n_from + n_to)*16 bytes are needed: bl = br = 448 at
   d = 2 with 5 + 5 distinct MPO bonds); larger inputs return -90 with a
   message in nc_last_error.  No hipMalloc, hipFree or allocating library call
   is made.  Every call on one GPU uses that GPU's array: calls that overlap on
   one GPU race.  Returns 0 or an error code. */
int nc_contract_sos_left_noalloc(uint32_t bl, uint32_t br, uint32_t cl, uint32_t cr, uint32_t d,
                                 const nc_zcomplex* left_environment, uint32_t number_of_matrices,
                                 const uint32_t* sparse_operator_indices, const nc_zcomplex* sparse_operator_matrices,
                                 const nc_zcomplex* state_site_tensor, nc_zcomplex* new_left_environment,
                                 void* stream);

/* Krylov basis size used by optimize_ (the reference hard-wires 3 at
   core.f95:2567; 3 is the default here too).  Accepts 3..62; returns the
   previous value, or -1 for a value outside that range. */
int nc_set_optimizer_ncv(int ncv);

/* Host-pointer residency of the C-ABI entry points above (no counterpart in the
   reference; opt-in, for callers such as the reference's C++ layer, whose host
   arrays keep their address and are handed back unchanged): with a byte budget
   > 0 the library keeps a device copy of the arrays it reads and writes, keyed by
   host address, and does not upload an input again when its address, size and
   fingerprint (every word of arrays up to 32 KiB; 4096 words spread over larger
   ones) match a copy.  Outputs are still written to the host.  An in-place edit
   that touches no sampled word is not seen: forget such an array first.  Copies
   are evicted least recently used first beyond the budget; 0 turns residency off
   and frees every copy.  Not for concurrent calls from several threads.
   nc_set_host_residency returns the previous budget; nc_host_residency_stats
   fills out[0] hits, out[1] misses, out[2] bytes uploaded, out[3] bytes served
   from copies, out[4] outputs kept, out[5] copies held, out[6] bytes held. */
int64_t nc_set_host_residency(int64_t max_bytes);
int nc_host_residency_forget(const void* host);
int nc_host_residency_stats(int64_t* out);

/* How mixed-precision single-site solves apply H_eff in complex64: 0 on the
   complex64 GEMM (operands split inside it), 1 on operands split once into
   16-bit planes when bl * br >= 512 * 512 (the default), 2 always on planes.
   Returns the previous mode, or -1 for a mode outside 0..2.  No counterpart in
   the reference (its matvec is complex128 only). */
int nc_set_mixed_planes(int mode);

/* Part form of the planes the mixed form splits its operands into: 1 (the
   default) two fp16 parts of each value scaled by a power of two per matrix,
   three partial products per real product; 0 three bf16 parts, six partial
   products.  Both keep FP32 accuracy.  Takes effect for site solves started
   afterwards; returns the previous form, or -1 for another value.  No
   counterpart in the reference (its matvec is complex128 only). */
int nc_set_plane_form(int form);

/* Counters of the site optimizer since load (5 values): site solves, first
   Krylov runs that ended with -14, ncv x 3/2 retries run (core.f95:2588-2602),
   dense strategy-1 fallbacks (core.f95:2569-2587), solves returned with -14. */
int nc_optimizer_stats(int64_t* out);

/* Architecture name, compute-unit count, clock and memory of the current device. */
int nc_device_info(char* name, int name_len, int* cu_count, double* clock_ghz, size_t* total_mem);

/* project_into_orthogonal_space / unproject_from_orthogonal_space
   (core.f95:1513-1664), internal to the reference's Fortran core. */
void nc_project_into_orthogonal_space(uint32_t n, uint32_t nproj, uint32_t nrefl, uint32_t orth,
                                      const nc_zcomplex* reflectors, const nc_zcomplex* coefficients,
                                      const uint32_t* swaps, const nc_zcomplex* in, nc_zcomplex* out);
void nc_unproject_from_orthogonal_space(uint32_t n, uint32_t nproj, uint32_t nrefl, uint32_t orth,
                                        const nc_zcomplex* reflectors, const nc_zcomplex* coefficients,
                                        const uint32_t* swaps, const nc_zcomplex* in, nc_zcomplex* out);

/* ---- device-resident sweep driver ------------------------------------------
   The host loop of the reference's Chain / BaseChain (chain.cpp:147-237,
   base_chain.cpp:81-123, chain.hpp:248-315) for a fixed bond dimension, with
   every state site and environment kept in HBM.  Operator sites are passed
   concatenated: idx_all holds 2*nmat[i] entries per site, mats_all d[i]^2*nmat[i];
   bonds holds n_sites+1 bond dimensions.  Functions return 0 or an error code
   (optimizer failures use the reference's info codes, or 101-104 for the
   post-checks of optimizer.cpp:281-310 and base_chain.cpp:97-99). */
/* Timing of `reps` H_eff matvecs (after `warmup` untimed ones) with every operand
   resident in HBM; y receives the last result.  total_ms from HIP events on the
   library's stream; phase_ms (optional, 3 entries) the average durations of the
   step-A GEMM launch, the MPO mix kernel and the step-C GEMM launch. */
int nc_bench_heff(uint32_t bl, uint32_t br, uint32_t cl, uint32_t cr, uint32_t d, uint32_t nmat,
                  const uint32_t* idx, const nc_zcomplex* mats, const nc_zcomplex* L, const nc_zcomplex* R,
                  const nc_zcomplex* x, nc_zcomplex* y, uint32_t warmup, uint32_t reps, double* total_ms,
                  double* phase_ms);

/* The same matvec in complex64 on FP32 matrix cores (the mixed-precision form's
   matvec): operands are narrowed on the device before timing and y is widened
   back to complex128. */
int nc_bench_heff_c64(uint32_t bl, uint32_t br, uint32_t cl, uint32_t cr, uint32_t d, uint32_t nmat,
                      const uint32_t* idx, const nc_zcomplex* mats, const nc_zcomplex* L, const nc_zcomplex* R,
                      const nc_zcomplex* x, nc_zcomplex* y, uint32_t warmup, uint32_t reps, double* total_ms,
                      double* phase_ms);

/* The same matvec in complex64 on the bf16 matrix cores with its operands
   split once into bf16 planes (the mixed form's single-site matvec): L and R
   are split before timing; each timed matvec splits x, runs step A, mixes into
   Q's planes, runs step C and sums the K slices into complex128 y.  phase_ms,
   when not NULL, receives 5 averages: split of x, step A, mix, step C, sum.
   Benchmark entry point, no counterpart in the reference. */
int nc_bench_heff_planes(uint32_t bl, uint32_t br, uint32_t cl, uint32_t cr, uint32_t d, uint32_t nmat,
                         const uint32_t* idx, const nc_zcomplex* mats, const nc_zcomplex* L, const nc_zcomplex* R,
                         const nc_zcomplex* x, nc_zcomplex* y, uint32_t warmup, uint32_t reps, double* total_ms,
                         double* phase_ms);

/* K slices the plane matvec's two GEMM launches use (out[0] step A, out[1]
   step C) for a site of bonds bl, br, physical dimension d and nfrom / nto
   distinct MPO block indices.  Benchmark entry point. */
int nc_planes_splits(uint32_t bl, uint32_t br, uint32_t d, uint32_t nfrom, uint32_t nto, int32_t* out);

/* Calibration kernel for HBM byte counters: copies `bytes` (a multiple of 16)
   between two device buffers `reps` times; *avg_ms is the mean launch time. */
int nc_bench_copy(uint64_t bytes, uint32_t reps, double* avg_ms);

/* Site-move statistics since load: out[0] polar splits, out[1] Newton-Schulz
   iterations, out[2] splits completed by the SVD (zero input, or no convergence
   within the step limit). */
int nc_polar_stats(int64_t* out);
/* Newton-Schulz steps a polar split may take before rocSOLVER's SVD completes it
   (no counterpart in the reference; default 64, 0 sends every split to the SVD).
   Returns the previous limit, or -1 outside [0, 1000]. */
int nc_set_polar_max_steps(int steps);

typedef struct nc_chain nc_chain;

nc_chain* nc_chain_create(uint32_t n_sites, const uint32_t* d, const uint32_t* cl, const uint32_t* cr,
                          const uint32_t* nmat, const uint32_t* idx_all, const nc_zcomplex* mats_all,
                          const uint32_t* bonds);
void nc_chain_destroy(nc_chain* chain);
int nc_chain_set_options(nc_chain* chain, double site_tol, double sweep_tol, double sanity,
                         uint32_t max_iterations, const char* which, uint32_t ncv, int rethrow);
/* Mixed-precision form: H_eff matvecs (mixed != 0) and environment updates
   (env32 != 0 as well) in complex64 on FP32 matrix cores; Krylov vectors,
   orthogonalisation, the Rayleigh quotient and the site move in FP64.  The Krylov
   iteration stops at tol32; the eigenvalue is the FP64 Rayleigh quotient of the
   result, or (polish != 0) the FP64 iteration continues from it to the site
   tolerance.  Call before nc_chain_reset for env32 to cover the first build. */
int nc_chain_set_mixed(nc_chain* chain, int mixed, int env32, double tol32, int polish);
/* raw (unnormalised) site tensor (br, bl, d); nc_chain_reset canonicalises */
int nc_chain_set_raw_site(nc_chain* chain, uint32_t site, const nc_zcomplex* data);
int nc_chain_get_site(nc_chain* chain, uint32_t site, nc_zcomplex* out);
/* The environment of the current site: side 0 the left one (bl, bl, cl), side 1
   the right one (br, br, cr). */
int nc_chain_get_environment(nc_chain* chain, uint32_t site, int side, nc_zcomplex* out);
int nc_chain_reset(nc_chain* chain);
int nc_chain_optimize_site(nc_chain* chain);
int nc_chain_move(nc_chain* chain, int direction);
int nc_chain_sweep(nc_chain* chain);
int nc_chain_sweep_until_converged(nc_chain* chain);
int nc_chain_energy(nc_chain* chain, double* energy, int* current_site);
int nc_chain_expectation(nc_chain* chain, nc_zcomplex* out);
/* Bond dimension growth.  set_growth: optimizeChain (base_chain.cpp:64-79)
   grows chi by one after each convergence until the energy changes by less than
   chain_tol (grow = 0: fixed chi).  increase_bandwidth: increaseBandwidthDimension
   (chain.cpp:75-121) at the first site, the state unchanged.  get_bonds writes
   the N + 1 current bond dimensions. */
int nc_chain_set_growth(nc_chain* chain, int grow, double chain_tol);
int nc_chain_increase_bandwidth(nc_chain* chain, uint32_t new_chi);
int nc_chain_optimize_chain(nc_chain* chain);
int nc_chain_get_bonds(nc_chain* chain, uint32_t* bonds);
/* Excited levels (chain.cpp:60-73, 240-263, 301-310 of the reference).  The
   chain must be at its first site.  add_projector keeps the current state as a
   lower level; after nc_chain_reset every site solve works in the space
   orthogonal to the lower levels.  projector_overlap is the norm of the current
   site's overlap with them.  randomize draws new raw site tensors from `seed`
   (call nc_chain_reset afterwards).  solve_for_multiple_levels converges
   `levels` levels in turn at the chain's bond dimension (new raw states from
   seed + 1, seed + 2, ...) and writes their energies. */
int nc_chain_add_projector(nc_chain* chain);
int nc_chain_clear_projectors(nc_chain* chain);
int nc_chain_projector_overlap(nc_chain* chain, double* overlap);
int nc_chain_randomize(nc_chain* chain, uint64_t seed);
/* reset (chain.cpp:175-237) for the next level: the bond dimension goes back to
   the initial one (with growth) raised to the minimum that leaves room outside
   the lower levels, then new raw sites from `seed` are canonicalised. */
int nc_chain_restart_level(nc_chain* chain, uint64_t seed);
int nc_chain_solve_for_multiple_levels(nc_chain* chain, uint32_t levels, uint64_t seed, double* energies);
/* <a|b> of two chains' states (computeStateOverlap, states.hpp:349-382): the
   chains need the same length and physical dimensions, not the same bonds. */
int nc_chain_state_overlap(nc_chain* a, nc_chain* b, nc_zcomplex* overlap);
/* Independent chains of the same length (a parameter sweep: each its own
   couplings and bond dimension) sweeping in lockstep: all chains sit on the same
   site, and their site solves share grouped kernel launches.  A chain whose
   energy has converged leaves the lockstep, so each performs exactly the sweeps
   it would perform alone.  The chains stay owned by the caller (destroy the
   batch first); every chain must be at its first site and have no lower levels. */
typedef struct nc_chain_batch nc_chain_batch;
nc_chain_batch* nc_chain_batch_create(nc_chain* const* chains, uint32_t n);
void nc_chain_batch_destroy(nc_chain_batch* batch);
int nc_chain_batch_sweep(nc_chain_batch* batch);
/* nc_chain_reset of every chain of the batch with grouped launches (all right
   normalisations in one batched polar split); the chains hold raw sites. */
int nc_chain_batch_reset(nc_chain_batch* batch);
int nc_chain_batch_sweep_until_converged(nc_chain_batch* batch);
/* wall time (6 values, ms) spent in lockstep moves [0] and site solves [1]; the
   site solves split into per-chain preparation [2], the batched solves [3] and
   per-chain post-checks [4]; [5] is the part of [3] inside the Krylov iteration */
int nc_chain_batch_times(nc_chain_batch* batch, double* times_ms);
/* Launch option of the batched Krylov solver (no counterpart in the reference):
   on = 1 (default) records each Arnoldi extension of a lockstep site solve in a
   HIP graph and replays it on later restarts; on = 0 launches the same kernels
   one by one.  Results are identical either way.  Returns the previous setting. */
int nc_set_krylov_graphs(int on);
/* GEMM autotuner (no counterpart in the reference).  On (default; NC_AUTOTUNE=0
   in the environment turns the default off), the first launch of every
   contraction plan with a new shape key (chi_left, chi_right, d, MPO bonds,
   contraction step; grouped plans: the largest of each and the chain count)
   times the candidate launch choices -- K step per LDS round 16 or 32, K split
   1/2/4/8, 3- or 4-multiplication complex product -- and keeps the fastest.
   nc_set_autotune returns the previous setting; nc_autotune_table writes the
   choices made so far as text (one line per key) and returns the length needed
   including the terminating zero. */
int nc_set_autotune(int on);
int64_t nc_autotune_table(char* buf, int64_t len);
/* Forces one launch choice on every GEMM plan, for tests of each kernel variant
   (no counterpart in the reference): bk 16 or 32 (K step per LDS stage) or -16
   (16 with two LDS stages), or 6 (complex64 untransposed operands on the bf16
   matrix cores, each FP32 value split into three bf16 parts; other plans take
   bk 16), 8 or 9 (complex128 untransposed operands whose M and N are multiples
   of 64 and K of 8, copied global -> LDS through a ring of 3 or 4 eight-deep
   stages; other plans take bk 16), nsplit 0 (heuristic) or 1/2/4/8, m3 -1 (the
   precision's default), 0
   (4 real products) or 1 (3); bk = 0 ends the forcing.
   Returns -1 for values outside these. */
int nc_gemm_force_choice(int bk, int nsplit, int m3);
/* counts (7): FP64 matvecs, site solves, moves, sweeps, failures, complex64 matvecs,
   sites of the mixed form finished in FP64 (a complex64 solve that came out above
   the previous energy);
   times_ms (4): optimize, site move, environment update, reset */
int nc_chain_stats(nc_chain* chain, int64_t* counts, double* times_ms);

#ifdef __cplusplus
}
#endif

#endif /* NUTCRACKER_CORE_ABI_H */
