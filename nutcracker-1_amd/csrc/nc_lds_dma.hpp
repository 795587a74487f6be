// Global -> LDS copies without register staging (gfx950
// global_load_lds_dwordx4), shared by the GEMM kernels that feed their matrix
// cores from a ring of LDS stages.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace nc {

typedef __attribute__((address_space(3))) void* LdsAddr;

// 16 bytes per lane from global memory into LDS: lane i's 16 bytes at byte
// offset voff_i from g land at LDS address lds + 16 i.  g and lds must be the
// same in every lane (they go through scalar registers).  Written out (rather
// than through the builtin) so that the compiler does not add waits for these
// copies before LDS reads of other buffers; it has no register result, and its
// completion is awaited only by an explicit vmcnt wait of the caller.  One
// wait state between the M0 write and its use.
// (M0 is clobbered: clang warns that it is a reserved register, which matters
// only to code that keeps a value in M0 across the statement; no kernel that
// uses this helper does.)
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Winline-asm"
__device__ __forceinline__ void lds_dma16(const void* g, unsigned voff, const void* lds) {
    const unsigned m0 = (unsigned)(uintptr_t)(LdsAddr)(lds);
    asm volatile("s_mov_b32 m0, %2\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %0, %1" ::"v"(voff), "s"(g), "s"(m0)
                 : "m0");
}
#pragma clang diagnostic pop

}  // namespace nc
