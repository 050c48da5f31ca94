// Device-side synchronisation used by the persistent solver kernels.
//
// grid_barrier: one arrival counter and one generation word. Only the block leader performs
// the agent-scope fences: every wave first waits for its own memory operations to complete
// (s_waitcnt 0), the block synchronises, then the leader writes back / invalidates the L2
// once for the whole block. Fencing in every wave costs one L2 write-back per wave and
// measured ~0.1 us per block on MI355X; one per block divides that by the waves per block.
// Every wait is bounded: after AFX_SPIN_LIMIT polls it gives up and sets err[0], so a logic
// error can never leave the GPU spinning (results are then flagged, not trusted).
#pragma once
#include <hip/hip_runtime.h>

namespace afx {

#ifndef AFX_SPIN_LIMIT
#define AFX_SPIN_LIMIT (1u << 26)
#endif

struct GridBar {
  unsigned *cnt;  // arrival counter (own 128-byte line)
  unsigned *gen;  // generation word (own 128-byte line)
  int *err;       // set when a wait gave up
};

// wait until this wave's outstanding vector / LDS / scalar memory operations completed
__device__ __forceinline__ void wave_drain() { __builtin_amdgcn_s_waitcnt(0); }

// All threads of all blocks must call it the same number of times. `phase` is the block's
// view of the generation word (kept by thread 0).
__device__ __forceinline__ void grid_barrier(const GridBar &b, unsigned &phase) {
  wave_drain();
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();  // release the block's writes at agent scope
    const unsigned G = gridDim.x;
    const unsigned target = phase + 1;
    const unsigned old = __hip_atomic_fetch_add(b.cnt, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (old == G - 1) {
      __hip_atomic_store(b.cnt, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(b.gen, target, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
    unsigned spins = 0;
    while ((int)(__hip_atomic_load(b.gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) - target) < 0) {
      __builtin_amdgcn_s_sleep(1);
      if (++spins > AFX_SPIN_LIMIT) {
        atomicOr(b.err, 1);
        break;
      }
    }
    __threadfence();  // acquire: drop stale cache lines before the block reads others' data
    phase = target;
  }
  __syncthreads();
}

}  // namespace afx
