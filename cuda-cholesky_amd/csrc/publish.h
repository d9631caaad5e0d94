// Publishing a workgroup's global stores to other workgroups (any XCD) with
// one L2 write-back per workgroup.
//
// An agent-scope release fence compiles to `buffer_wbl2 sc1` (write this
// XCD's L2 back) + `s_waitcnt vmcnt(0)`, and the wait covers only the
// issuing wave's stores; __syncthreads() itself waits for LDS traffic only.
// So before the barrier every wave waits for its own global stores to be
// acknowledged by the L2 (publish_wait_own), and after it one wave writes
// the L2 back once (publish_release) and then sets the flag or counter with
// a relaxed agent-scope atomic.  A fence in every wave instead would issue
// one L2 write-back per wave.
#ifndef HC_PUBLISH_H
#define HC_PUBLISH_H

#include <hip/hip_runtime.h>

namespace hc {

// s_waitcnt vmcnt(0): vmcnt field zero (bits 3:0 and 15:14), expcnt (6:4)
// and lgkmcnt (11:8) at their maximum, i.e. not waited for.
constexpr int kWaitVmcnt0 = 0x0F70;

__device__ __forceinline__ void publish_wait_own() { __builtin_amdgcn_s_waitcnt(kWaitVmcnt0); }

// After the barrier, in the one wave that then stores the flag.  The explicit
// wait after the fence keeps the write-back complete before the flag store.
__device__ __forceinline__ void publish_release() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
}

}  // namespace hc

#endif
