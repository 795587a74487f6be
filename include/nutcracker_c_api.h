/* The reference's C API front door (includes/nutcracker.h) for operators and
 * states in its protobuf wire format, the simple least-eigenvalue solvers, and
 * state expectation values and overlaps -- implemented on the MI355X core in
 * libnutcracker_hip.so.  Names, arguments and error behaviour follow the
 * reference: a failing call records a thread-local message
 * (Nutcracker_getError) and returns NULL, or leaves its outputs untouched.
 *
 * Scope: operators enter through Nutcracker_Operator_deserialize (an
 * OperatorBuffer message, sources/common/nutcracker.proto:26-35); the operator
 * and state builders (the MPO compiler) are not part of this library.
 *
 * Complex results are nc_zcomplex (two doubles, the layout of C's
 * "complex double" and C++'s std::complex<double>).
 */
#ifndef NUTCRACKER_C_API_H
#define NUTCRACKER_C_API_H

#include <stdint.h>

#include "nutcracker_core_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct NutcrackerOperator NutcrackerOperator;
typedef struct NutcrackerSerialization NutcrackerSerialization;
typedef struct NutcrackerState NutcrackerState;

/* errors.cpp:12-24: the calling thread's last error message, or NULL */
void Nutcracker_clearError(void);
char const* Nutcracker_getError(void);

/* serialization.cpp:6-14: the bytes of a serialized message */
void Nutcracker_Serialization_free(NutcrackerSerialization* s);
uint32_t Nutcracker_Serialization_getSize(NutcrackerSerialization* s);
void Nutcracker_Serialization_write(NutcrackerSerialization* s, void* buffer);

/* operator.cpp:10-22: an operator from / to an OperatorBuffer message (the
   distinct sites, and the 0-based sequence of them along the chain) */
NutcrackerOperator* Nutcracker_Operator_deserialize(uint32_t size, void const* data);
void Nutcracker_Operator_free(NutcrackerOperator* op);
NutcrackerSerialization* Nutcracker_Operator_serialize(NutcrackerOperator const* op);

/* operator.cpp:23-52: the lowest number_of_levels eigenvalues (as float) of
   the operator, from a chain with initial bond dimension 2 grown by one until
   the energy settles (optimizeChain), sanity threshold 1e-12, site 1e-9, sweep
   1e-8, chain 1e-8, each level orthogonal to the lower ones.  The variant with
   eigenvectors also returns each level's state (free with
   Nutcracker_State_free). */
void Nutcracker_Operator_simpleSolveForLeastEigenvalues(NutcrackerOperator const* op, uint32_t number_of_levels,
                                                        float* eigenvalues);
void Nutcracker_Operator_simpleSolveForLeastEigenvaluesWithEigenvectors(NutcrackerOperator const* op,
                                                                        uint32_t number_of_levels,
                                                                        float* eigenvalues,
                                                                        NutcrackerState** eigenvectors);

/* state.cpp:9-27: states from / to a StateBuffer message (first site MIDDLE,
   the others RIGHT-normalised), <psi|H|psi> (computeExpectationValue) and
   <psi1|psi2> (computeStateOverlap) on the device */
NutcrackerState* Nutcracker_State_deserialize(uint32_t size, void const* data);
void Nutcracker_State_free(NutcrackerState* state);
NutcrackerSerialization* Nutcracker_State_serialize(NutcrackerState const* state);
void Nutcracker_State_computeExpectation(NutcrackerState const* state, NutcrackerOperator const* op,
                                         nc_zcomplex* result);
void Nutcracker_State_computeOverlap(NutcrackerState const* state1, NutcrackerState const* state2,
                                     nc_zcomplex* result);

#ifdef __cplusplus
}
#endif

#endif
