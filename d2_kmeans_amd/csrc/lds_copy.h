// One wave's copy from global memory straight into LDS (global_load_lds),
// issued by hand: the compiler's own form of the instruction makes every
// later LDS read wait for all copies in flight (vmcnt(0)), which would drain
// a pipeline of copies each step.  4 or 16 bytes per lane: lane l's bytes go
// to lds + 4 l or lds + 16 l, lds being the same for every lane (M0 holds it).
// One wait state between the M0 write and the copy that reads it.  No
// register output; the address registers are read at issue.  The callers
// count these copies in their own s_waitcnt vmcnt and wait for all of them
// before any other vector memory access is awaited.  (M0 is a reserved
// register: the compiler notes its clobber without tracking it, and nothing
// else in the kernels that use these copies uses M0.)
#pragma once

#include <hip/hip_runtime.h>

#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Winline-asm"
__device__ __forceinline__ unsigned int lds_m0(void *lds) {
  return __builtin_amdgcn_readfirstlane(
      (unsigned int)(size_t)(__attribute__((address_space(3))) char *)lds);
}
__device__ __forceinline__ void lds_copy16(const void *gsrc, void *lds) {
  asm volatile("s_mov_b32 m0, %1\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %0, off"
               :
               : "v"(gsrc), "s"(lds_m0(lds))
               : "memory", "m0");
}
__device__ __forceinline__ void lds_copy4(const void *gsrc, void *lds) {
  asm volatile("s_mov_b32 m0, %1\n\ts_nop 0\n\tglobal_load_lds_dword %0, off"
               :
               : "v"(gsrc), "s"(lds_m0(lds))
               : "memory", "m0");
}
#pragma clang diagnostic pop
