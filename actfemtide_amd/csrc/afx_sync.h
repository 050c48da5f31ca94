// Device-side synchronisation used by the persistent solver kernels.
//
// grid_barrier: a two-level arrival tree and one generation word. Block b arrives at group
// counter b % AFX_BAR_GROUPS; the last arrival of a group arrives at the top counter; the last
// top arrival publishes the new generation. With one counter, the 256 leaders' atomics
// serialise on one address (~9.5 us per barrier measured); the tree spreads them over 16
// addresses (measured: 256 blocks, 9.4 -> 8.7 us; the release-only / acquire-only fences
// below bring it to 5.1 us, of which 1.8 us is the arrival tree itself). Counters only grow (after barrier m a group counter holds m * group size and the
// top counter m * groups), so no counter is ever reset. Only the block leader performs the
// agent-scope fences: every wave first waits for its own memory operations to complete
// (s_waitcnt 0), the block synchronises, then the leader writes back / invalidates the L2
// once for the whole block. Fencing in every wave costs one L2 write-back per wave and
// measured ~0.1 us per block on MI355X; one per block divides that by the waves per block.
// Every wait is bounded: after AFX_SPIN_LIMIT polls it gives up and sets err[0], so a logic
// error can never leave the GPU spinning (results are then flagged, not trusted).
#pragma once
#include <hip/hip_runtime.h>

namespace afx {

#ifndef AFX_BAR_GROUPS
#define AFX_BAR_GROUPS 16
#endif

#ifndef AFX_SPIN_LIMIT
#define AFX_SPIN_LIMIT (1u << 26)
#endif

struct GridBar {
  unsigned *cnt;  // top-level arrival counter (own 128-byte line)
  unsigned *gen;  // generation word (own 128-byte line)
  unsigned *grp;  // AFX_BAR_GROUPS group arrival counters, 32 words (128 bytes) apart
  int *err;       // set when a wait gave up
};

// wait until this wave's outstanding vector / LDS / scalar memory operations completed
__device__ __forceinline__ void wave_drain() { __builtin_amdgcn_s_waitcnt(0); }

// All threads of all blocks must call it the same number of times. `phase` is the block's
// view of the generation word (kept by thread 0). Returns true in every thread of the block when
// err[0] was found set (by any block's wait that gave up, or by the host before the launch): the
// caller then leaves its loop, and the waits of blocks still in barriers give up within 1024
// polls instead of AFX_SPIN_LIMIT, so a launch whose blocks disagree ends promptly. The flag is
// read after the wait when `poll` is set (once per solver iteration) or when the wait was long;
// a fast barrier without `poll` does not pay for the extra load.
__device__ __forceinline__ bool grid_barrier(const GridBar &b, unsigned &phase, bool poll = false) {
  __shared__ int s_abort;
  wave_drain();
  __syncthreads();
  if (threadIdx.x == 0) {
    // release the block's writes at agent scope (L2 write-back); a full __threadfence would also
    // invalidate, and its second half costs ~1.8 us per barrier on MI355X for nothing
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    const unsigned G = gridDim.x;
    const unsigned NG = G < AFX_BAR_GROUPS ? G : AFX_BAR_GROUPS;
    const unsigned g = blockIdx.x % NG;
    const unsigned gsize = G / NG + (g < G % NG ? 1u : 0u);
    const unsigned target = phase + 1;
    const unsigned old = __hip_atomic_fetch_add(b.grp + 32 * g, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (old + 1u == target * gsize) {
      const unsigned top = __hip_atomic_fetch_add(b.cnt, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (top + 1u == target * NG) __hip_atomic_store(b.gen, target, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
    unsigned spins = 0;
    bool long_wait = false;
    while ((int)(__hip_atomic_load(b.gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) - target) < 0) {
      __builtin_amdgcn_s_sleep(1);
      if (++spins > AFX_SPIN_LIMIT) {
        atomicOr(b.err, 1);
        break;
      }
      if ((spins & 1023u) == 0) {
        long_wait = true;
        if (__hip_atomic_load(b.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) break;
      }
    }
    // acquire: drop stale L2 lines before the block reads others' data
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    phase = target;
    s_abort = (poll || long_wait) ? (__hip_atomic_load(b.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) : 0;
  }
  __syncthreads();
  return s_abort != 0;
}

}  // namespace afx
