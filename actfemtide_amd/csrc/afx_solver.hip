// COCG solve kernels for gfx950 (MI355X).
//
// The whole preconditioned COCG iteration of COCG13_single (src/solver/COCG13_single.f90:181-323)
// runs inside ONE persistent cooperative launch per chunk of iterations:
//   z = M^-1 r  -- pipelined forward (strict lower) and backward (strict upper) sweeps: block b
//                  owns the contiguous rows [rstart[b], rstart[b+1]) and walks them in steps of
//                  increasing dependency level, keeping its own z in LDS; z values another block
//                  needs are posted to mailbox slots that the reader polls (point-to-point, no
//                  grid barrier per level, no progress counters);
//   rho = r^T z -- fused into the backward sweep, reduced deterministically across blocks;
//   p = z + beta p;  q = A p (CSR, `lpr` lanes per row) fused with p^T q;
//   x += alpha p, r -= alpha q fused with ||r||^2 and the per-column convergence test.
// Bilinear forms are unconjugated (COCG); SpMV, reductions, alpha, beta, rho and all vectors but
// z are FP64. The preconditioner (sweep values, DD, z) is complex FP64 or complex FP32 (TP).
//
// Several right-hand sides (columns) share each matrix read: vectors are stored row-interleaved
// [n][k], a lane owns one (row, column) pair, and each column runs its own recurrence with its
// own convergence flag. A column may carry its own angular frequency w: the matrix is held as
// (K, M) pairs and the entry used is K + i*w*M, so several tidal constituents share one launch.
#include <hip/hip_cooperative_groups.h>

#include <algorithm>
#include <cstdlib>
#include <type_traits>
#include <vector>

#include "afx_internal.h"
#include "afx_sync.h"

namespace cg = cooperative_groups;

namespace afx {

template <typename T>
__device__ __forceinline__ cx<T> operator+(cx<T> a, cx<T> b) {
  return {a.r + b.r, a.i + b.i};
}
template <typename T>
__device__ __forceinline__ cx<T> operator-(cx<T> a, cx<T> b) {
  return {a.r - b.r, a.i - b.i};
}
template <typename T>
__device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b) {
  return {a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r};
}
// acc + a*b
template <typename T>
__device__ __forceinline__ cx<T> cfma(cx<T> a, cx<T> b, cx<T> acc) {
  acc.r = fma(a.r, b.r, acc.r);
  acc.r = fma(-a.i, b.i, acc.r);
  acc.i = fma(a.r, b.i, acc.i);
  acc.i = fma(a.i, b.r, acc.i);
  return acc;
}
// matrix entry (K, M) at frequency w
template <typename T>
__device__ __forceinline__ cx<T> mval(cx<T> km, T w) {
  return {km.r, w * km.i};
}
template <typename T>
__device__ __forceinline__ cx<double> up(cx<T> a) {
  return {(double)a.r, (double)a.i};
}
template <typename T>
__device__ __forceinline__ cx<T> down(cx<double> a) {
  return {(T)a.r, (T)a.i};
}
// a / b as (a * conj(b)) / |b|^2 with b scaled by an exact power of two (C99 Annex G form,
// the complex division the reference's compiled Fortran performs): the scaling avoids
// overflow without changing the rounding of the textbook formula.
// bfloat16 <-> float: round to nearest even (the coefficients are finite)
__device__ __forceinline__ unsigned int f2bf16(float f) {
  const unsigned int u = __float_as_uint(f);
  return (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
}
// coefficient q (0..3) of a 16-byte group of node-block coefficients, as complex FP64
__device__ __forceinline__ cx<double> nb_coef(uint4 g, int q) {
  const unsigned int w = q == 0 ? g.x : q == 1 ? g.y : q == 2 ? g.z : g.w;
  return {(double)__uint_as_float(w << 16), (double)__uint_as_float(w & 0xffff0000u)};
}

__device__ __forceinline__ cx<double> cdivd(cx<double> a, cx<double> b) {
  double c = b.r, d = b.i;
  const double m = fmax(fabs(c), fabs(d));
  int e = 0;
  if (m > 0.0 && isfinite(m)) {
    e = ilogb(m);
    c = scalbn(c, -e);
    d = scalbn(d, -e);
  }
  const double den = __dmul_rn(c, c) + __dmul_rn(d, d);
  cx<double> r;
  r.r = scalbn((__dmul_rn(a.r, c) + __dmul_rn(a.i, d)) / den, -e);
  r.i = scalbn((__dmul_rn(a.i, c) - __dmul_rn(a.r, d)) / den, -e);
  return r;
}
// Lane exchange inside aligned groups of 2..16 lanes through DPP (no LDS round trip).
// (bound_ctrl set: no old value to merge; every source lane of these patterns is in the group)
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)b, CTRL, 0xF, 0xF, true);
  const unsigned hi = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)(b >> 32), CTRL, 0xF, 0xF, true);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
template <int CTRL>
__device__ __forceinline__ float dpp_mov(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, true));
}

// Sum over an aligned group of LPR lanes, every lane of the group getting the same bits:
// quad_perm [1,0,3,2] and [2,3,0,1] (xor 1, xor 2), then row_half_mirror (lane i <- 7-i: the
// other quad of 8) and row_mirror (lane i <- 15-i: the other half of 16). Every group's lanes
// must execute it together.
template <int LPR, typename T>
__device__ __forceinline__ cx<T> group_sum(cx<T> v) {
  static_assert(LPR == 1 || LPR == 2 || LPR == 4 || LPR == 8 || LPR == 16, "group of 1..16 lanes");
  if constexpr (LPR >= 2) {
    v.r += dpp_mov<0xB1>(v.r);
    v.i += dpp_mov<0xB1>(v.i);
  }
  if constexpr (LPR >= 4) {
    v.r += dpp_mov<0x4E>(v.r);
    v.i += dpp_mov<0x4E>(v.i);
  }
  if constexpr (LPR >= 8) {
    v.r += dpp_mov<0x141>(v.r);
    v.i += dpp_mov<0x141>(v.i);
  }
  if constexpr (LPR >= 16) {
    v.r += dpp_mov<0x140>(v.r);
    v.i += dpp_mov<0x140>(v.i);
  }
  return v;
}

// Mailbox exchange of z between blocks. An exported z value travels through a mailbox slot that
// holds MB_EMPTY (the all-ones bit pattern: a NaN that no arithmetic produces) until the owner
// writes the value; a reader polls the slot itself, so the value is its own ready flag and one
// hand-off costs a write-through store plus the reader's load -- no progress counter and no wait
// for store completion before raising one. Slots are accessed with agent-scope relaxed 64-bit
// atomics (they bypass the per-XCD non-coherent caches). The owner writes MB_EMPTY back once every
// reader is past the sweep (after the next grid barrier). A poll gives up after AFX_SPIN_LIMIT
// tries, or soon after another wait gave up, and flags err: results are then not trusted.
constexpr unsigned long long MB_EMPTY = ~0ull;

__device__ __forceinline__ unsigned long long mb_ld(const unsigned long long *q) {
  return __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void mb_st(unsigned long long *q, unsigned long long v) {
  __hip_atomic_store(q, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ bool mb_give_up(unsigned spins, int *err) {
  if (spins > AFX_SPIN_LIMIT) {
    atomicOr(err, 2);
    return true;
  }
  return (spins & 1023u) == 0 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
}
__device__ __forceinline__ cx<double> mb_get(const cx<double> *p, int *err) {
  const unsigned long long *q = (const unsigned long long *)p;
  unsigned long long a = mb_ld(q), b = mb_ld(q + 1);
  unsigned spins = 0;
  while (a == MB_EMPTY || b == MB_EMPTY) {
    if (mb_give_up(++spins, err)) break;
    __builtin_amdgcn_s_sleep(1);
    if (a == MB_EMPTY) a = mb_ld(q);
    if (b == MB_EMPTY) b = mb_ld(q + 1);
  }
  return {__longlong_as_double((long long)a), __longlong_as_double((long long)b)};
}
__device__ __forceinline__ cx<float> mb_get(const cx<float> *p, int *err) {
  const unsigned long long *q = (const unsigned long long *)p;
  unsigned long long a = mb_ld(q);
  unsigned spins = 0;
  while (a == MB_EMPTY) {
    if (mb_give_up(++spins, err)) break;
    __builtin_amdgcn_s_sleep(1);
    a = mb_ld(q);
  }
  cx<float> r;
  __builtin_memcpy(&r, &a, 8);
  return r;
}
__device__ __forceinline__ void mb_put(cx<double> *p, cx<double> v) {
  unsigned long long *q = (unsigned long long *)p;
  mb_st(q, (unsigned long long)__double_as_longlong(v.r));
  mb_st(q + 1, (unsigned long long)__double_as_longlong(v.i));
}
__device__ __forceinline__ void mb_put(cx<float> *p, cx<float> v) {
  unsigned long long a;
  __builtin_memcpy(&a, &v, 8);
  mb_st((unsigned long long *)p, a);
}
__device__ __forceinline__ void mb_clear(cx<double> *p) {
  mb_st((unsigned long long *)p, MB_EMPTY);
  mb_st((unsigned long long *)p + 1, MB_EMPTY);
}
__device__ __forceinline__ void mb_clear(cx<float> *p) { mb_st((unsigned long long *)p, MB_EMPTY); }

// empty the block's own exported slots of both mailboxes (rows [rs, re), k columns); the flags
// of 8 slots per thread are loaded before the first is tested
template <typename TP>
__device__ void mb_clear_own2(cx<TP> *mba, const unsigned char *expa, cx<TP> *mbb, const unsigned char *expb, int rs,
                              int re, int k) {
  const long long lo = (long long)rs * k, hi = (long long)re * k;
  for (long long i0 = lo + threadIdx.x; i0 < hi; i0 += 8 * 256) {
    unsigned char fa[8], fb[8];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const long long i = (i0 + u * 256 < hi ? i0 + u * 256 : i0) / k;
      fa[u] = expa[i];
      fb[u] = expb[i];
    }
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const long long i = i0 + u * 256;
      if (i < hi && fa[u]) mb_clear(&mba[i]);
      if (i < hi && fb[u]) mb_clear(&mbb[i]);
    }
  }
}

__device__ __forceinline__ double zero_of(double) { return 0.0; }
__device__ __forceinline__ cx<double> zero_of(cx<double>) { return {0.0, 0.0}; }
__device__ __forceinline__ double addv(double a, double b) { return a + b; }
__device__ __forceinline__ cx<double> addv(cx<double> a, cx<double> b) { return {a.r + b.r, a.i + b.i}; }

// Sum v over the threads of the block that share a column (threads t, t+kp, t+2kp, ...);
// every thread returns the sum of its own column. Fixed tree order: deterministic.
template <typename V>
__device__ V block_cols(V v, V *lds, int kp) {
  const int t = threadIdx.x;
  lds[t] = v;
  __syncthreads();
  for (int s = (256 / kp) >> 1; s >= 1; s >>= 1) {
    if (t / kp < s) lds[t] = addv(lds[t], lds[t + s * kp]);
    __syncthreads();
  }
  V r = lds[t % kp];
  __syncthreads();
  return r;
}

// Grid-wide per-column reduction with identical results in every block that sums the same block
// range [g0, g1) (all blocks, or the blocks of one stacked system; !dist only).
template <typename V>
__device__ V grid_cols(V v, V *part, V *fin, V *lds, int k, int kp, int c, bool dist, const GridBar &bar,
                       unsigned &phase, int g0, int g1, bool &abort, bool poll = false) {
  const int t = threadIdx.x;
  const int G = gridDim.x;
  V bs = block_cols(v, lds, kp);
  if (t < kp && t < k) part[(size_t)blockIdx.x * k + t] = bs;
  abort |= grid_barrier(bar, phase, poll && !dist);
  if (!dist) {
    V s = zero_of(v);
    if (c < k)
      for (int g = g0 + t / kp; g < g1; g += 256 / kp) s = addv(s, part[(size_t)g * k + c]);
    return block_cols(s, lds, kp);
  }
  for (int cc = blockIdx.x; cc < k; cc += G) {
    V s = zero_of(v);
    for (int g = t; g < G; g += 256) s = addv(s, part[(size_t)g * k + cc]);
    s = block_cols(s, lds, 1);
    if (t == 0) fin[cc] = s;
  }
  abort |= grid_barrier(bar, phase, poll);
  return (c < k) ? fin[c] : zero_of(v);
}

// Nodal gradient-space correction (AFX_PC_GRAD). g = G^T r at interior node a: the incident
// edges' r with +1 where the edge ends at a and -1 where it starts there; zn = g / (i w mnd_a),
// i.e. the Jacobi step on the nodal operator G^T A G, whose diagonal is i w diag(G^T M G) because
// the curl-curl part annihilates gradients.
__device__ __forceinline__ cx<double> nodal_g(const int *__restrict__ nptr, const int *__restrict__ nadj,
                                              const cx<double> *R, int a) {
  const int q0 = nptr[a], q1 = nptr[a + 1];
  cx<double> g = {0.0, 0.0};
#pragma unroll 4
  for (int q = q0; q < q1; q++) {
    const int v = nadj[q];
    const cx<double> r = R[(v > 0 ? v : -v) - 1];
    if (v > 0) {
      g.r += r.r;
      g.i += r.i;
    } else {
      g.r -= r.r;
      g.i -= r.i;
    }
  }
  return g;
}
__device__ __forceinline__ cx<double> nodal_z(cx<double> g, double w, double mnd) {
  const double f = 1.0 / (w * mnd);
  return {g.i * f, -g.r * f};  // g / (i w mnd)
}

// all columns finished? identical in every block because the states are
__device__ bool block_all(bool mine) { return __syncthreads_and(mine ? 1 : 0) != 0; }

// One pipelined triangular sweep (FWD: z = D~^-1 (r - L z); else z -= D~^-1 U z) over the CSR
// side layout, any column count. Own rows are read from LDS (zs, single column) or Z; rows of
// other blocks are polled from the mailbox mb; exported rows are posted to mb.
template <typename TP, int LPR, bool MULTI, bool FWD>
__device__ void sweep(const CocgParams<TP> &p, const SweepSide &sd, cx<TP> *zs, cx<TP> *mb, int c, bool cvalid,
                      TP w, int f, cx<double> &rho_loc) {
  const int b = blockIdx.x;
  const int t = threadIdx.x;
  const int k = p.k;
  const int nf = p.nf;
  const int rs = p.rstart[b], re = p.rstart[b + 1];
  const bool lds = !MULTI && p.lds_rows > 0;
  const cx<TP> *val = (const cx<TP> *)sd.val;
  const int kp = MULTI ? p.kp : 1;
  const int slot = MULTI ? t / kp : t / LPR;
  const int nslot = MULTI ? 256 / kp : 256 / LPR;
  const int lane = MULTI ? 0 : t % LPR;
  const int s0 = sd.bstep[b], s1 = sd.bstep[b + 1];
  __syncthreads();
  for (int s = s0; s < s1; ++s) {
    const int lo = sd.spos[s], hi = sd.spos[s + 1];
    for (int pos = lo + slot; pos < hi; pos += nslot) {
      const int row = sd.row[pos];
      const int eb = sd.ptr[pos], ee = sd.ptr[pos + 1];
      if (MULTI) {
        if (cvalid) {
          cx<TP> acc = {TP(0), TP(0)};
          for (int q = eb; q < ee; ++q) {
            const int cl = sd.col[q];
            const size_t cix = (size_t)cl * k + c;
            const cx<TP> zc = (cl >= rs && cl < re) ? p.Z[cix] : mb_get(&mb[cix], p.err);
            acc = cfma(mval(val[q], w), zc, acc);
          }
          const size_t ix = (size_t)row * k + c;
          cx<TP> z;
          if (FWD) {
            z = cmul(down<TP>(p.R[ix]) - acc, p.DD[(size_t)row * nf + f]);
          } else {
            z = p.Z[ix] - cmul(p.DD[(size_t)row * nf + f], acc);
            rho_loc = addv(rho_loc, cmul(p.R[ix], up(z)));
          }
          p.Z[ix] = z;
          if (sd.exp[row]) mb_put(&mb[ix], z);
        }
      } else {
        cx<TP> acc = {TP(0), TP(0)};
        for (int q = eb + lane; q < ee; q += LPR) {
          const int cl = sd.col[q];
          const bool own = cl >= rs && cl < re;
          const cx<TP> zc = own ? (lds ? zs[cl - rs] : p.Z[cl]) : mb_get(&mb[cl], p.err);
          acc = cfma(mval(val[q], w), zc, acc);
        }
        acc = group_sum<LPR>(acc);
        if (lane == 0) {
          cx<TP> z;
          if (FWD) {
            z = cmul(down<TP>(p.R[row]) - acc, p.DD[row]);
          } else {
            const cx<TP> zo = lds ? zs[row - rs] : p.Z[row];
            z = zo - cmul(p.DD[row], acc);
            rho_loc = addv(rho_loc, cmul(p.R[row], up(z)));
          }
          if (lds)
            zs[row - rs] = z;
          else
            p.Z[row] = z;
          if (sd.exp[row]) mb_put(&mb[row], z);
        }
      }
    }
    __syncthreads();
  }
}

// raw (non-polling) mailbox loads of the remote entries of a prefetched row
__device__ __forceinline__ void mb_raw(const cx<double> *p, cx<double> &v) {
  const unsigned long long *q = (const unsigned long long *)p;
  v.r = __longlong_as_double((long long)mb_ld(q));
  v.i = __longlong_as_double((long long)mb_ld(q + 1));
}
__device__ __forceinline__ void mb_raw(const cx<float> *p, cx<float> &v) {
  const unsigned long long a = mb_ld((const unsigned long long *)p);
  __builtin_memcpy(&v, &a, 8);
}
__device__ __forceinline__ bool mb_empty(cx<double> v) {
  return (unsigned long long)__double_as_longlong(v.r) == MB_EMPTY ||
         (unsigned long long)__double_as_longlong(v.i) == MB_EMPTY;
}
__device__ __forceinline__ bool mb_empty(cx<float> v) {
  unsigned long long a;
  __builtin_memcpy(&a, &v, 8);
  return a == MB_EMPTY;
}

// Wait until none of the N slots marked `rem` holds MB_EMPTY: every pending slot is reloaded in
// the same round (one memory round trip per retry for all of them, not one per slot).
template <typename TP, int N>
__device__ __forceinline__ void mb_poll(const cx<TP> *mb, const size_t (&idx)[N], const bool (&rem)[N],
                                        cx<TP> (&z)[N], int *err, unsigned &nfallback) {
  bool any = false;
#pragma unroll
  for (int i = 0; i < N; i++) any |= rem[i] && mb_empty(z[i]);
  if (!any) return;
  ++nfallback;
  unsigned spins = 0;
  for (;;) {
#pragma unroll
    for (int i = 0; i < N; i++)
      if (rem[i] && mb_empty(z[i])) mb_raw(&mb[idx[i]], z[i]);
    any = false;
#pragma unroll
    for (int i = 0; i < N; i++) any |= rem[i] && mb_empty(z[i]);
    if (!any || mb_give_up(++spins, err)) break;
    __builtin_amdgcn_s_sleep(1);
  }
}

// Single-column sweep with the block's z (and, on the forward pass, r) staged in LDS.
// zs[slot] holds r_i until row i is processed (forward) and z_i afterwards.
//
// Every step holds at most AFX_STEP_ROWS rows; row slot t/8 of the block takes one of them (or
// the dummy position) with 8 lanes, lane l holding the entry pair (2l, 2l+1) of the 16-wide row.
// The sweep is a software pipeline in which no wait targets a memory operation issued fewer than
// PMB steps earlier (vmcnt retires in issue order, so a wait for a young load would also wait
// for every older load and store). Step s
//   * reads the z values of its rows' own-block entries from LDS,
//   * requests the matrix data of step s+PMAT (positions from the step bounds staged in LDS) and
//     the mailbox slots of step s+PMB (from entry sources that arrived PMAT-PMB steps ago) into
//     the register sets step s-1 released -- issuing these overlaps the LDS reads,
//   * computes its rows from the data requested at steps s-PMAT and s-PMB (a mailbox value
//     still empty is polled: it waits for the producer's post), posts z to LDS and, when another
//     block reads it, to the mailbox, and ends with the block barrier.
// Each register set is named by a compile-time index of a 6-way unrolled loop, so no set is
// ever copied (a copy would wait for the set's loads).
namespace pipe {
constexpr int L = 8;             // lanes per row
constexpr int NS = 256 / L;      // row slots per block (== AFX_STEP_ROWS)
constexpr int PMAT = 5;          // steps between requesting a row's matrix data and using it
constexpr int PMB = 1;           // steps between requesting its mailbox slots and using them
static_assert(NS == AFX_STEP_ROWS, "one row slot per row of a step");
}  // namespace pipe

template <typename TP>
struct PSet {
  int2 rw;     // {row | exported << 31, LDS slot}
  int src[2];  // entry sources (>= 0 LDS slot, < 0 mailbox slot ~src)
  cx<TP> val[2];
  cx<TP> dd;
};

// requests of one row slot: the position lo + slot of the step [lo, hi), or the dummy position
template <typename TP>
__device__ __forceinline__ void pset_load(const SweepSide &sd, int lo, int hi, int slot, int l8, PSet<TP> &r) {
  const int p0 = lo + slot;
  const int pos = p0 < hi ? p0 : sd.npos;
  const size_t e0 = (size_t)pos * AFX_PIPE_W + 2 * l8;
  r.rw = ((const int2 *)sd.prow)[pos];
  const int2 c = *(const int2 *)(sd.psrc + e0);
  r.src[0] = c.x;
  r.src[1] = c.y;
  const cx<TP> *v = (const cx<TP> *)sd.pval + e0;
  r.val[0] = v[0];
  r.val[1] = v[1];
  r.dd = ((const cx<TP> *)sd.dpos)[pos];
}

// whether any lane of the wave holds an entry of another block in set r (wave-uniform)
template <typename TP>
__device__ __forceinline__ bool pset_remote(const PSet<TP> &r) {
  return __builtin_amdgcn_ballot_w64((r.src[0] | r.src[1]) < 0) != 0;
}

// raw mailbox loads for the entries of a loaded set; an entry that needs none loads slot src
// (src <= n: any valid slot, the value is not used). Issued by every wave: a conditional issue
// would make the compiler count these loads as absent and wait for younger loads after them.
template <typename TP>
__device__ __forceinline__ void pset_mb(const PSet<TP> &r, const cx<TP> *mb, cx<TP> (&z)[2]) {
#pragma unroll
  for (int e = 0; e < 2; e++) {
    const int sc = r.src[e];
    mb_raw(&mb[sc < 0 ? ~sc : sc], z[e]);
  }
}

template <typename TP, bool FWD>
__device__ void sweep_pipe(const CocgParams<TP> &p, const SweepSide &sd, const int *sb, cx<TP> *zs, cx<TP> *mb,
                           TP w, unsigned &nfallback) {
  using namespace pipe;
  const int b = blockIdx.x;
  const int t = threadIdx.x;
  const int slot = t / L, l8 = t % L;
  const int s0 = sd.bstep[b], s1 = sd.bstep[b + 1];
  // first position of step s (the block's end past its last step): staged in LDS
  auto bnd = [&](int s) { return sb[min(s, s1) - s0]; };

  PSet<TP> R0, R1, R2, R3, R4, R5;
  cx<TP> M0[2], M1[2];
  pset_load(sd, bnd(s0), bnd(s0 + 1), slot, l8, R0);
  pset_load(sd, bnd(s0 + 1), bnd(s0 + 2), slot, l8, R1);
  pset_load(sd, bnd(s0 + 2), bnd(s0 + 3), slot, l8, R2);
  pset_load(sd, bnd(s0 + 3), bnd(s0 + 4), slot, l8, R3);
  pset_load(sd, bnd(s0 + 4), bnd(s0 + 5), slot, l8, R4);
  pset_mb(R0, mb, M0);
  int lo = bnd(s0 + 5);  // first position of the next step to request
  int hi = bnd(s0 + 6);

  // Step s uses r (requested at step s-5) and m (mailbox slots requested at step s-1); it requests
  // the mailbox slots of step s+1 into mn from rm, then the matrix data of step s+5 into rn (both
  // sets were released by step s-1). The mailbox requests go first: a wait for them then never
  // covers the matrix requests issued in the same step.
  auto step = [&](int s, const PSet<TP> &r, const cx<TP>(&m)[2], PSet<TP> &rn, cx<TP>(&mn)[2],
                  const PSet<TP> &rm) {
    cx<TP> zl[2];
#pragma unroll
    for (int e = 0; e < 2; e++) zl[e] = zs[max(r.src[e], 0)];
    const cx<TP> zo = zs[r.rw.y];
    cx<TP> zv[2] = {zl[0], zl[1]};
    const int hn = bnd(s + PMAT + 2);  // upper bound of the step requested next
    if (pset_remote(r)) {
      // entries of other blocks: the mailbox values requested PMB steps ago; an empty one is
      // polled before this step's requests are issued, so that a poll's wait (vmcnt retires in
      // issue order) does not include loads issued in this step
      bool rem[2];
      size_t idx[2];
#pragma unroll
      for (int e = 0; e < 2; e++) {
        const int sc = r.src[e];
        rem[e] = sc < 0;
        idx[e] = rem[e] ? (size_t)~sc : 0;
        zv[e] = rem[e] ? m[e] : zl[e];
      }
      mb_poll(mb, idx, rem, zv, p.err, nfallback);
    }
    pset_mb(rm, mb, mn);
    pset_load(sd, lo, hi, slot, l8, rn);
    lo = hi;
    hi = hn;
    cx<TP> acc = cfma(mval(r.val[0], w), zv[0], cx<TP>{TP(0), TP(0)});
    acc = cfma(mval(r.val[1], w), zv[1], acc);
    acc = group_sum<L>(acc);
    if (l8 == 0) {
      const cx<TP> z = FWD ? cmul(zo - acc, r.dd) : zo - cmul(r.dd, acc);
      zs[r.rw.y] = z;
      if (r.rw.x < 0) mb_put(&mb[r.rw.x & 0x7fffffff], z);
    }
    __syncthreads();
  };
  for (int s = s0; s < s1; s += 6) {
    step(s, R0, M0, R5, M1, R1);
    if (s + 1 >= s1) break;
    step(s + 1, R1, M1, R0, M0, R2);
    if (s + 2 >= s1) break;
    step(s + 2, R2, M0, R1, M1, R3);
    if (s + 3 >= s1) break;
    step(s + 3, R3, M1, R2, M0, R4);
    if (s + 4 >= s1) break;
    step(s + 4, R4, M0, R3, M1, R5);
    if (s + 5 >= s1) break;
    step(s + 5, R5, M1, R4, M0, R0);
  }
}

// Several columns (2 <= k <= 128) share each sweep step: RL = LPR * k lanes per row, lane (c, l)
// takes column c and entries l, l + LPR, ... of the row. The z of the block's last ring_w
// positions (all columns) live in an LDS ring indexed by position; an entry whose source is a
// row of the same block less than ring_w - 3*smax positions earlier reads the ring (its slot is
// rewritten at least two steps later), an older one reads Z in global memory (ring_w >= 5*smax +
// 64, so it was written at least two steps earlier), a row of another block reads the mailbox.
// The same three-stage register pipeline as the single-column sweep hides the loads.
template <typename TP, int LPR, int E>
struct MRegs {
  int row;
  int exp;
  int col[E];
  int epos[E];
  cx<TP> val[E];
};

template <typename TP, int LPR, int E>
__device__ __forceinline__ void load_mstep(const SweepSide &sd, int posc, int lane, MRegs<TP, LPR, E> &r) {
  r.row = sd.row[posc];
  r.exp = sd.expp[posc];
  const size_t base = (size_t)posc * sd.ew;
#pragma unroll
  for (int e = 0; e < E; e++) {
    const int q = min(lane + e * LPR, sd.ew - 1);
    r.col[e] = sd.ecol[base + q];
    r.epos[e] = sd.epos[base + q];
    r.val[e] = ((const cx<TP> *)sd.eval)[base + q];
  }
}

// stage-1 loads of a loaded row: remote z (mailbox, raw), far own z (Z), the row's input value
// and its DD; every load unconditional (unused ones read index 0)
template <typename TP, int LPR, int E, bool FWD>
__device__ __forceinline__ void mfetch(const CocgParams<TP> &p, const MRegs<TP, LPR, E> &r, bool valid,
                                       const bool (&qok)[E], int q, int near_lim, int c, int f, const cx<TP> *mb,
                                       cx<TP> (&zx)[E], cx<TP> &zin, cx<TP> &dd) {
  const int k = p.k;
#pragma unroll
  for (int e = 0; e < E; e++) {
    const int cl = r.col[e];
    const bool use = valid && qok[e] && cl >= 0;
    const bool rem = use && r.epos[e] < 0;
    const bool far = use && r.epos[e] >= 0 && q - r.epos[e] >= near_lim;
    const size_t ix = (size_t)(use ? cl : 0) * k + c;
    if (rem)
      mb_raw(&mb[ix], zx[e]);
    else
      zx[e] = p.Z[far ? ix : 0];
  }
  const size_t rx = (size_t)(valid ? r.row : 0) * k + c;
  zin = FWD ? down<TP>(p.R[rx]) : p.Z[rx];
  dd = p.DD[(size_t)(valid ? r.row : 0) * p.nf + f];
}

template <typename TP, int LPR, int E, bool FWD>
__device__ void sweep_ring(const CocgParams<TP> &p, const SweepSide &sd, cx<TP> *ring, cx<TP> *mb,
                           unsigned &nfallback) {
  const int b = blockIdx.x;
  const int t = threadIdx.x;
  const int k = p.k;
  const int RL = LPR * k;
  const int nslot = 256 / RL;
  const int slot = t / RL;
  const bool active = slot < nslot;
  const int c = active ? (t % RL) / LPR : 0;  // this lane's column, its frequency and DD set
  const int lane = t % LPR;
  const TP w = (TP)p.wcol[c];
  const int f = p.col2f[c];
  const int s0 = sd.bstep[b], s1 = sd.bstep[b + 1];
  const int W = p.ring_w;
  const int near_lim = W - 3 * sd.smax;
  bool qok[E];
#pragma unroll
  for (int e = 0; e < E; e++) qok[e] = lane + e * LPR < sd.ew;
  auto sp = [&](int s) { return sd.spos[__builtin_amdgcn_readfirstlane(s < s1 ? s : s1)]; };
  auto clampp = [&](int lo, int hi) { return max(0, min(lo + slot, hi - 1)); };
  int l0 = sp(s0), l1 = sp(s0 + 1), l2 = sp(s0 + 2), l3 = sp(s0 + 3);
  MRegs<TP, LPR, E> R0, R1, R2;
  cx<TP> X0[E], X1[E], X2[E], I0, I1, I2, D0, D1, D2;
  load_mstep(sd, clampp(l0, l1), lane, R0);
  load_mstep(sd, clampp(l1, l2), lane, R1);
  mfetch<TP, LPR, E, FWD>(p, R0, active && l0 + slot < l1, qok, l0 + slot, near_lim, c, f, mb, X0, I0, D0);
  __syncthreads();
  auto do_row = [&](int q, const MRegs<TP, LPR, E> &r, const cx<TP>(&zx)[E], cx<TP> zin, cx<TP> dd) {
    cx<TP> acc = {TP(0), TP(0)};
    cx<TP> zv[E];
    bool rem[E];
    size_t idx[E];
#pragma unroll
    for (int e = 0; e < E; e++) {
      const int cl = r.col[e];
      rem[e] = qok[e] && cl >= 0 && r.epos[e] < 0;
      idx[e] = rem[e] ? (size_t)cl * k + c : 0;
      zv[e] = cx<TP>{TP(0), TP(0)};
      if (qok[e] && cl >= 0) {
        const int ps = r.epos[e];
        cx<TP> zc;
        if (ps >= 0 && q - ps < near_lim) {
          zc = ring[(size_t)(ps & (W - 1)) * k + c];
        } else if (ps >= 0) {
          zc = zx[e];
        } else {
          zc = zx[e];
        }
        zv[e] = zc;
      }
    }
    mb_poll(mb, idx, rem, zv, p.err, nfallback);
#pragma unroll
    for (int e = 0; e < E; e++)
      if (qok[e] && r.col[e] >= 0) acc = cfma(mval(r.val[e], w), zv[e], acc);
    acc = group_sum<LPR>(acc);
    if (lane == 0) {
      const cx<TP> z = FWD ? cmul(zin - acc, dd) : zin - cmul(dd, acc);
      const size_t ix = (size_t)r.row * k + c;
      ring[(size_t)(q & (W - 1)) * k + c] = z;
      p.Z[ix] = z;
      if (r.exp) mb_put(&mb[ix], z);
    }
  };
  auto step = [&](const MRegs<TP, LPR, E> &cur, const MRegs<TP, LPR, E> &nx1, MRegs<TP, LPR, E> &nx2,
                  const cx<TP>(&zx)[E], cx<TP> zin, cx<TP> dd, cx<TP>(&zx1)[E], cx<TP> &zin1, cx<TP> &dd1, int l4) {
    mfetch<TP, LPR, E, FWD>(p, nx1, active && l1 + slot < l2, qok, l1 + slot, near_lim, c, f, mb, zx1, zin1, dd1);
    load_mstep(sd, clampp(l2, l3), lane, nx2);
    if (active && l0 + slot < l1) do_row(l0 + slot, cur, zx, zin, dd);
    if (active)
      for (int pos = l0 + slot + nslot; pos < l1; pos += nslot) {
        MRegs<TP, LPR, E> ex;
        cx<TP> xx[E], xi, xd;
        load_mstep(sd, pos, lane, ex);
        mfetch<TP, LPR, E, FWD>(p, ex, true, qok, pos, near_lim, c, f, mb, xx, xi, xd);
        do_row(pos, ex, xx, xi, xd);
      }
    __syncthreads();
    l0 = l1;
    l1 = l2;
    l2 = l3;
    l3 = l4;
  };
  for (int s = s0; s < s1; s += 3) {
    step(R0, R1, R2, X0, I0, D0, X1, I1, D1, sp(s + 4));
    if (s + 1 >= s1) break;
    step(R1, R2, R0, X1, I1, D1, X2, I2, D2, sp(s + 5));
    if (s + 2 >= s1) break;
    step(R2, R0, R1, X2, I2, D2, X0, I0, D0, sp(s + 6));
  }
}

// q = A p over the rows sslot, sslot + snslot, ... below vend, LPR lanes per row, and
// c1 += p_r q_r (row order within a row slot). Rows are taken in rounds of SU rows per row slot.
// A software pipeline over rounds: while round g is computed, the P gathers of round g, the
// values and columns of round g+1 and the row bounds of round g+2 are in flight; each register
// set is named by a compile-time index of a 2-way unrolled loop (no copies). Entries of a row are
// summed in the same order as a row-at-a-time loop. SE covers rows of up to 20 entries (the
// edge-element rows have at most 19); longer rows finish in the loop after the unrolled entries.
template <int LPR>
__device__ __forceinline__ void spmv_pipe(const int *__restrict__ rowptr, const int *__restrict__ col,
                                          const cx<double> *__restrict__ val, double w, const cx<double> *P,
                                          cx<double> *Q, long long n, long long sslot, long long snslot,
                                          long long vend, int lane, cx<double> &c1_loc) {
  constexpr int SU = LPR >= 8 ? 4 : 2, SE = LPR >= 4 ? (20 + LPR - 1) / LPR : 6;
  struct RB {  // row bounds of one round
    int qb[SU], qe[SU];
  };
  struct VC {  // values and columns of one round (column -1: no entry)
    int qb[SU], qe[SU];
    int cl[SU][SE];
    cx<double> vv[SU][SE];
  };
  const long long rstep = (long long)SU * snslot;
  auto load_rb = [&](long long row0, RB &b) {
#pragma unroll
    for (int u = 0; u < SU; u++) {
      const long long r = row0 + u * snslot;
      const long long rc = r < n ? r : n - 1;
      b.qb[u] = rowptr[rc];
      b.qe[u] = rowptr[rc + 1];
      if (r >= vend) b.qe[u] = b.qb[u];
    }
  };
  auto load_vc = [&](const RB &b, VC &v) {
#pragma unroll
    for (int u = 0; u < SU; u++) {
      v.qb[u] = b.qb[u];
      v.qe[u] = b.qe[u];
#pragma unroll
      for (int e = 0; e < SE; e++) {
        const int q = b.qb[u] + lane + e * LPR;
        const bool in = q < b.qe[u];
        const int qc = in ? q : 0;
        v.vv[u][e] = val[qc];
        const int cq = col[qc];  // loaded unconditionally (a fixed count of loads per round)
        v.cl[u][e] = in ? cq : -1;
      }
    }
  };
  auto round = [&](long long row0, const VC &v, VC &vn, const RB &bn, RB &bnn) {
    cx<double> pv[SU][SE], pr[SU];
#pragma unroll
    for (int u = 0; u < SU; u++) {
#pragma unroll
      for (int e = 0; e < SE; e++) pv[u][e] = P[max(v.cl[u][e], 0)];
      const long long r = row0 + u * snslot;
      pr[u] = P[r < n ? r : 0];
    }
    load_vc(bn, vn);
    load_rb(row0 + 2 * rstep, bnn);
#pragma unroll
    for (int u = 0; u < SU; u++) {
      cx<double> acc = {0.0, 0.0};
#pragma unroll
      for (int e = 0; e < SE; e++)
        if (v.cl[u][e] >= 0) acc = cfma(mval(v.vv[u][e], w), pv[u][e], acc);
      for (int q = v.qb[u] + lane + SE * LPR; q < v.qe[u]; q += LPR) acc = cfma(mval(val[q], w), P[col[q]], acc);
      const long long r = row0 + u * snslot;
      const cx<double> a = group_sum<LPR>(acc);
      if (lane == 0 && r < vend) {
        Q[r] = a;
        c1_loc = addv(c1_loc, cmul(pr[u], a));
      }
    }
  };
  RB b0, b1;
  VC v0, v1;
  load_rb(sslot, b0);
  load_rb(sslot + rstep, b1);
  load_vc(b0, v0);
  for (long long row0 = sslot; row0 < vend; row0 += 2 * rstep) {
    round(row0, v0, v1, b1, b0);
    if (row0 + rstep >= vend) break;
    round(row0 + rstep, v1, v0, b0, b1);
  }
}

// dynamic LDS of the single-column path: z of the block's rows, then the step bounds of the
// forward and backward sweeps (16-byte aligned)
template <typename TP>
__host__ __device__ constexpr size_t lds_bounds_offset(int lds_rows) {
  return ((size_t)lds_rows * sizeof(cx<TP>) + 15) & ~(size_t)15;
}

// The persistent COCG iteration (COCG13_single.f90:181-323). SWEEP = false compiles the
// preconditioner sweeps out (the DIAG / NONE kernel).
template <typename TP, int LPR, bool MULTI, bool PROF, bool SWEEP>
__device__ __forceinline__ void cocg_body(const CocgParams<TP> &p) {
  __shared__ cx<double> lds_c[256];
  __shared__ double lds_d[256];
  extern __shared__ unsigned char dyn[];
  cx<TP> *zs = (cx<TP> *)dyn;  // [lds_rows]
  const int t = threadIdx.x;
  const long long gtid = (long long)blockIdx.x * 256 + t;
  const long long nthr = (long long)gridDim.x * 256;
  const int k = p.k;
  const int kp = MULTI ? p.kp : 1;
  const int c = MULTI ? (t % kp) : 0;
  const bool cvalid = c < k;
  const long long slot = MULTI ? gtid / kp : gtid / LPR;
  const long long nslot = MULTI ? nthr / kp : nthr / LPR;
  const int lane = MULTI ? 0 : (int)(gtid % LPR);
  const double w = cvalid ? p.wcol[c] : 0.0;
  const TP wT = (TP)w;
  const int f = cvalid ? p.col2f[c] : 0;
  const bool sweep_pc = SWEEP && (p.pc == AFX_PC_IC || p.pc == AFX_PC_SSOR || p.pc == AFX_PC_SSOR_BJ);
  const int pb = blockIdx.x;
  // the row range of the block (sweep partition; without sweeps only stacked systems use it, and
  // the grid may then be larger than the partition)
  const bool use_rows = SWEEP || (!MULTI && p.nsys > 1);
  const int rs = use_rows ? p.rstart[pb] : 0, re = use_rows ? p.rstart[pb + 1] : 0;
  // stacked systems: this block's system (its state index), the blocks its reductions sum, and
  // the rows its vector passes cover (its own rows; otherwise all rows, grid-strided)
  const bool sys = !MULTI && p.nsys > 1;
  const int cst = sys ? p.bsys[pb] : c;
  const int g0 = sys ? p.sblk[cst] : 0, g1 = sys ? p.sblk[cst + 1] : (int)gridDim.x;
  const long long vlo = sys ? rs + t : gtid, vstep = sys ? 256 : nthr, vend = sys ? re : p.n;
  const bool dist = !sys && p.dist_reduce != 0;
  // single-column sweeps: the block's step bounds of both directions are staged in LDS once
  const bool pipe_sweep = SWEEP && !MULTI && p.lds_rows > 0 && p.fw.psrc != nullptr && p.bw.psrc != nullptr;
  int *sbf = (int *)(dyn + lds_bounds_offset<TP>(p.lds_rows));
  int *sbb = sbf + p.fw.maxsteps + 1;
  if (pipe_sweep) {
    const int f0 = p.fw.bstep[pb], f1 = p.fw.bstep[pb + 1];
    const int b0 = p.bw.bstep[pb], b1 = p.bw.bstep[pb + 1];
    for (int i = t; i <= f1 - f0; i += 256) sbf[i] = p.fw.spos[f0 + i];
    for (int i = t; i <= b1 - b0; i += 256) sbb[i] = p.bw.spos[b0 + i];
  }
  unsigned phase = 0;
  if (t == 0) phase = __hip_atomic_load(p.bar.gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();

  int st = cvalid ? p.conv[cst] : 1;
  const double bn = cvalid ? p.bnrm2[cst] : 1.0;
  cx<double> rho1 = cvalid ? p.rho1[cst] : cx<double>{0.0, 0.0};
  double res = cvalid ? p.resid[cst] : 0.0;
  int its = cvalid ? p.iters[cst] : 0;
  int it = p.status[0];
  const int it_end = min(it + p.chunk, p.maxit);
  bool alldone;
  if (sys) {  // every stacked system finished (any count of systems)
    int running = 0;
    for (int q = t; q < p.nsys; q += 256) running |= (p.conv[q] == 0) ? 1 : 0;
    alldone = __syncthreads_or(running) == 0;
  } else {
    alldone = block_all(st != 0);
  }
  // optional per-block phase timing (wall clock, 100 MHz) and mailbox poll count
  constexpr bool prof = PROF;
  unsigned long long tfw = 0, tbw = 0, tbar = 0, trest = 0, titer = 0, trho = 0, tpost = 0, tpu = 0, tsp = 0, tclr = 0;
  unsigned nfb = 0;

  // blocks of the lower half empty the next iteration's mailboxes between their forward and
  // backward sweeps (the backward sweep reaches them last), the upper half after the backward
  // sweep (while they wait for rho): the time each block waits anyway
  const bool clear_after_fw = blockIdx.x < gridDim.x / 2;
  // set when a grid barrier found the error flag raised: every block leaves after the iteration
  // (the flag never clears during a launch, so every block sees it by its next barrier)
  bool abort = false;
  while (!alldone && it < it_end && !abort) {
    ++it;
    const unsigned long long ci0 = prof ? wall_clock64() : 0;
    // mailboxes of this iteration, and the other set, which no block touches in this iteration
    // (its last reader finished before the previous rho reduction) and which must be empty at
    // the next one (the grid barriers in between order the clear before any use)
    cx<TP> *const mbf = (it & 1) ? p.mbf2 : p.mbf, *const mbb = (it & 1) ? p.mbb2 : p.mbb;
    cx<TP> *const mbf_nx = (it & 1) ? p.mbf : p.mbf2, *const mbb_nx = (it & 1) ? p.mbb : p.mbb2;
    // ---- z = M^-1 r (COCG13_single.f90:186-241) and rho = r^T z (:246-250) ----
    cx<double> rho_loc = {0.0, 0.0};
    // nodal part of z = M^-1 r (AFX_PC_GRAD): zn on the interior nodes, and its share of
    // rho = r^T z, r^T G zn = (G^T r)^T zn; the edges add G zn in the p update, after the rho
    // barrier has published every block's zn
    if (!MULTI && p.grad) {
      // stacked systems: the nodes of the block's own system, split over that system's blocks (its
      // rho sums only its own blocks' partials)
      const long long a0 = sys ? p.noff[cst] + (long long)(pb - g0) * 256 + t : gtid;
      const long long a1 = sys ? p.noff[cst + 1] : p.nn;
      const long long astep = sys ? (long long)(g1 - g0) * 256 : nthr;
      for (long long a = a0; a < a1; a += astep) {
        const cx<double> g = nodal_g(p.nptr, p.nadj, p.R, (int)a);
        const cx<double> zn = nodal_z(g, w, p.mnd[a]);
        p.Zn[a] = zn;
        rho_loc = addv(rho_loc, cmul(g, zn));
      }
    }
    if (MULTI && p.grad && cvalid) {  // one lane per (node, column), zn at the column's omega
      for (long long a = slot; a < p.nn; a += nslot) {
        const int q0 = p.nptr[a], q1 = p.nptr[a + 1];
        cx<double> g = {0.0, 0.0};
        for (int q = q0; q < q1; q++) {
          const int v = p.nadj[q];
          const cx<double> r = p.R[(size_t)((v > 0 ? v : -v) - 1) * k + c];
          g.r += v > 0 ? r.r : -r.r;
          g.i += v > 0 ? r.i : -r.i;
        }
        const cx<double> zn = nodal_z(g, w, p.mnd[a]);
        p.Zn[(size_t)a * k + c] = zn;
        rho_loc = addv(rho_loc, cmul(g, zn));
      }
    }
    if (sweep_pc) {
      // forward then backward sweep with no grid barrier between them: a block's backward sweep
      // reads its own rows' forward z (LDS or Z) and other blocks' backward z (its own mailbox),
      // never another block's forward result.
      if (pipe_sweep) {
        // the block's r staged in LDS, 8 loads in flight per thread
        for (int r0 = rs + t; r0 < re; r0 += 8 * 256) {
          cx<double> v[8];
#pragma unroll
          for (int u = 0; u < 8; u++) v[u] = p.R[min(r0 + u * 256, re - 1)];
#pragma unroll
          for (int u = 0; u < 8; u++)
            if (r0 + u * 256 < re) zs[r0 + u * 256 - rs] = down<TP>(v[u]);
        }
        if (t == 0) zs[p.lds_rows - 1] = cx<TP>{TP(0), TP(0)};  // the slot padding entries read
        const unsigned long long c0 = prof ? wall_clock64() : 0;
        __syncthreads();
        sweep_pipe<TP, true>(p, p.fw, sbf, zs, mbf, wT, nfb);
        const unsigned long long c1 = prof ? wall_clock64() : 0;
        if (clear_after_fw) mb_clear_own2(mbf_nx, p.fw.exp, mbb_nx, p.bw.exp, rs, re, 1);
        const unsigned long long c2 = prof ? wall_clock64() : 0;
        sweep_pipe<TP, false>(p, p.bw, sbb, zs, mbb, wT, nfb);
        if (prof) {
          const unsigned long long c3 = wall_clock64();
          tfw += c1 - c0;
          tbar += c2 - c1;
          tbw += c3 - c2;
        }
        // z written back and rho = r^T z accumulated in row order, 8 loads in flight per thread
        for (int r0 = rs + t; r0 < re; r0 += 8 * 256) {
          cx<double> rv[8];
#pragma unroll
          for (int u = 0; u < 8; u++) rv[u] = p.R[min(r0 + u * 256, re - 1)];
#pragma unroll
          for (int u = 0; u < 8; u++) {
            const int r = r0 + u * 256;
            if (r < re) {
              const cx<TP> z = zs[r - rs];
              p.Z[r] = z;
              rho_loc = addv(rho_loc, cmul(rv[u], up(z)));
            }
          }
        }
      } else if (MULTI && p.ring_w > 0) {
        const unsigned long long c0 = prof ? wall_clock64() : 0;
        sweep_ring<TP, 2, 8, true>(p, p.fw, zs, mbf, nfb);
        const unsigned long long c1 = prof ? wall_clock64() : 0;
        if (clear_after_fw) mb_clear_own2(mbf_nx, p.fw.exp, mbb_nx, p.bw.exp, rs, re, k);
        const unsigned long long c2 = prof ? wall_clock64() : 0;
        sweep_ring<TP, 2, 8, false>(p, p.bw, zs, mbb, nfb);
        if (prof) {
          const unsigned long long c3 = wall_clock64();
          tfw += c1 - c0;
          tbar += c2 - c1;
          tbw += c3 - c2;
        }
        __syncthreads();
        if (cvalid)
          for (long long row = rs + t / kp; row < re; row += 256 / kp) {
            const size_t ix = (size_t)row * k + c;
            rho_loc = addv(rho_loc, cmul(p.R[ix], up(p.Z[ix])));
          }
      } else {
        sweep<TP, LPR, MULTI, true>(p, p.fw, zs, mbf, c, cvalid, wT, f, rho_loc);
        if (clear_after_fw) mb_clear_own2(mbf_nx, p.fw.exp, mbb_nx, p.bw.exp, rs, re, MULTI ? k : 1);
        sweep<TP, LPR, MULTI, false>(p, p.bw, zs, mbb, c, cvalid, wT, f, rho_loc);
        if (!MULTI && p.lds_rows > 0)
          for (int r = rs + t; r < re; r += 256) p.Z[r] = zs[r - rs];
      }
    } else {
      if (MULTI) {
        if (cvalid)
          for (long long row = slot; row < p.n; row += nslot) {
            const size_t ix = (size_t)row * k + c;
            const cx<double> r = p.R[ix];
            const cx<TP> z = (p.pc == AFX_PC_DIAG) ? cmul(down<TP>(r), p.DD[(size_t)row * p.nf + f]) : down<TP>(r);
            p.Z[ix] = z;
            rho_loc = addv(rho_loc, cmul(r, up(z)));
          }
      } else {
        for (long long row = vlo; row < vend; row += vstep) {
          const cx<double> r = p.R[row];
          cx<TP> z;
          if (p.pc == AFX_PC_NODEBLK) {  // node blocks: the row's 8 coefficients, r on its block's rows
            const int m = p.bm[row], loc = m & 15, sz = (m >> 4) + 1;
            const uint4 *cf = (const uint4 *)p.bi + (size_t)row * (AFX_NB_W / 4);
            const uint4 g0 = cf[0], g1 = cf[1];
            cx<double> acc = {0.0, 0.0};
#pragma unroll
            for (int j = 0; j < AFX_NB_W; j++) acc = cfma(nb_coef(j < 4 ? g0 : g1, j & 3), p.R[row - loc + min(j, sz - 1)], acc);
            z = down<TP>(acc);
          } else {
            z = (p.pc == AFX_PC_DIAG) ? cmul(down<TP>(r), p.DD[row]) : down<TP>(r);
          }
          p.Z[row] = z;
          rho_loc = addv(rho_loc, cmul(r, up(z)));
        }
      }
    }
    if (sweep_pc && !clear_after_fw) {
      const unsigned long long cc0 = prof ? wall_clock64() : 0;
      mb_clear_own2(mbf_nx, p.fw.exp, mbb_nx, p.bw.exp, rs, re, MULTI ? k : 1);
      if (prof) tclr += wall_clock64() - cc0;
    }
    const unsigned long long cr0 = prof ? wall_clock64() : 0;
    if (prof && !sweep_pc) tfw += cr0 - ci0;  // without sweeps: the z pass (edge and nodal parts)
    const cx<double> rho = grid_cols(rho_loc, p.part_rho, p.fin_rho, lds_c, k, kp, c, dist, p.bar, phase, g0, g1, abort);
    const unsigned long long cr1 = prof ? wall_clock64() : 0;
    if (prof) trho += cr1 - cr0;

    // ---- p = z (first iteration) or z + beta p (:258-268) ----
    const bool first = (it == 1);
    const cx<double> beta = first ? cx<double>{0.0, 0.0} : cdivd(rho, rho1);
    if (MULTI) {
      if (cvalid)
        for (long long row = slot; row < p.n; row += nslot) {
          const size_t ix = (size_t)row * k + c;
          cx<double> z = up(p.Z[ix]);
          if (p.grad) {  // + G zn
            const int2 e = ((const int2 *)p.en)[row];
            const cx<double> za = p.Zn[(size_t)e.x * k + c], zb = p.Zn[(size_t)e.y * k + c];
            z.r += zb.r - za.r;
            z.i += zb.i - za.i;
          }
          p.P[ix] = first ? z : cfma(beta, p.P[ix], z);
        }
    } else {
      // 8 rows per thread in flight (one wave per SIMD: the loads must overlap)
      if (p.grad) {
        // z = z_edge + G zn: the edge's end-node correction minus its start-node correction
        for (long long r0 = vlo; r0 < vend; r0 += 8 * vstep) {
          cx<TP> zv[8];
          cx<double> pv[8], za[8], zb[8];
          int2 ev[8];
#pragma unroll
          for (int u = 0; u < 8; u++) {
            const long long r = r0 + u * vstep < vend ? r0 + u * vstep : r0;
            zv[u] = p.Z[r];
            pv[u] = first ? cx<double>{0.0, 0.0} : p.P[r];
            ev[u] = ((const int2 *)p.en)[r];
          }
#pragma unroll
          for (int u = 0; u < 8; u++) {
            za[u] = p.Zn[ev[u].x];
            zb[u] = p.Zn[ev[u].y];
          }
#pragma unroll
          for (int u = 0; u < 8; u++) {
            const long long r = r0 + u * vstep;
            const cx<double> z0 = up(zv[u]);
            const cx<double> z = {z0.r + (zb[u].r - za[u].r), z0.i + (zb[u].i - za[u].i)};
            if (r < vend) p.P[r] = first ? z : cfma(beta, pv[u], z);
          }
        }
      } else {
        for (long long r0 = vlo; r0 < vend; r0 += 8 * vstep) {
          cx<TP> zv[8];
          cx<double> pv[8];
#pragma unroll
          for (int u = 0; u < 8; u++) {
            const long long r = r0 + u * vstep < vend ? r0 + u * vstep : r0;
            zv[u] = p.Z[r];
            pv[u] = first ? cx<double>{0.0, 0.0} : p.P[r];
          }
#pragma unroll
          for (int u = 0; u < 8; u++) {
            const long long r = r0 + u * vstep;
            if (r < vend) p.P[r] = first ? up(zv[u]) : cfma(beta, pv[u], up(zv[u]));
          }
        }
      }
    }
    abort |= grid_barrier(p.bar, phase);
    const unsigned long long cp1 = prof ? wall_clock64() : 0;
    if (prof) tpu += cp1 - cr1;

    // ---- q = A p (:273-288) and c1 = p^T q (:293-296) ----
    cx<double> c1_loc = {0.0, 0.0};
    if (MULTI) {
      // one lane per (row, column): the row's first ME values and columns are loaded before the
      // first gather, and all its gathers before the first product (entries summed in row order)
      constexpr int ME = 20;
      for (long long row = slot; row < p.n; row += nslot) {
        const int qb = p.rowptr[row], qe = p.rowptr[row + 1];
        if (cvalid) {
          cx<double> vv[ME], pv[ME];
          int cl[ME];
#pragma unroll
          for (int e = 0; e < ME; e++) {
            const int q = qb + e;
            const bool in = q < qe;
            const int qc = in ? q : qb;
            vv[e] = p.val[qc];
            const int cq = p.col[qc];
            cl[e] = in ? cq : -1;
          }
#pragma unroll
          for (int e = 0; e < ME; e++) pv[e] = p.P[(size_t)max(cl[e], 0) * k + c];
          const size_t ix = (size_t)row * k + c;
          const cx<double> prow = p.P[ix];
          cx<double> acc = {0.0, 0.0};
#pragma unroll
          for (int e = 0; e < ME; e++)
            if (cl[e] >= 0) acc = cfma(mval(vv[e], w), pv[e], acc);
          for (int q = qb + ME; q < qe; ++q) acc = cfma(mval(p.val[q], w), p.P[(size_t)p.col[q] * k + c], acc);
          p.Q[ix] = acc;
          c1_loc = addv(c1_loc, cmul(prow, acc));
        }
      }
    } else {
      // stacked systems: the block's own rows, LPR lanes per row slot
      const long long sslot = sys ? rs + t / LPR : slot, snslot = sys ? 256 / LPR : nslot;
      spmv_pipe<LPR>(p.rowptr, p.col, p.val, w, p.P, p.Q, p.n, sslot, snslot, vend, lane, c1_loc);
    }
    const cx<double> c1 = grid_cols(c1_loc, p.part_c1, p.fin_c1, lds_c, k, kp, c, dist, p.bar, phase, g0, g1, abort);
    if (prof) tsp += wall_clock64() - cp1;
    const cx<double> alpha = cdivd(rho, c1);
    const bool run = (st == 0);
    if (run && (!(c1.r != 0.0 || c1.i != 0.0) || !isfinite(alpha.r) || !isfinite(alpha.i) ||
                !(rho.r != 0.0 || rho.i != 0.0))) {
      st = -200;  // breakdown
      its = it;
    }
    const bool upd = (st == 0);
    const cx<double> malpha = {-alpha.r, -alpha.i};

    // ---- x += alpha p, r -= alpha q (:304-307), ||r||^2 (:310-313) ----
    double rr_loc = 0.0;
    if (MULTI) {
      if (cvalid && upd)
        for (long long row = slot; row < p.n; row += nslot) {
          const size_t ix = (size_t)row * k + c;
          p.X[ix] = cfma(alpha, p.P[ix], p.X[ix]);
          const cx<double> r = cfma(malpha, p.Q[ix], p.R[ix]);
          p.R[ix] = r;
          rr_loc += r.r * r.r + r.i * r.i;
        }
    } else if (upd) {
      // 4 rows per thread in flight; |r|^2 accumulated in row order
      for (long long r0 = vlo; r0 < vend; r0 += 4 * vstep) {
        cx<double> xv[4], pv[4], qv[4], rv[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
          const long long r = r0 + u * vstep < vend ? r0 + u * vstep : r0;
          xv[u] = p.X[r];
          pv[u] = p.P[r];
          qv[u] = p.Q[r];
          rv[u] = p.R[r];
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
          const long long row = r0 + u * vstep;
          if (row < vend) {
            p.X[row] = cfma(alpha, pv[u], xv[u]);
            const cx<double> r = cfma(malpha, qv[u], rv[u]);
            p.R[row] = r;
            rr_loc += r.r * r.r + r.i * r.i;
          }
        }
      }
    }
    const double rr = grid_cols(rr_loc, p.part_rr, p.fin_rr, lds_d, k, kp, c, dist, p.bar, phase, g0, g1, abort, true);
    if (upd) {
      res = sqrt(rr / bn);  // :315 RESID = sqrt(DNRM2/BNRM2)
      if (res <= p.tol) {
        st = 1;
        its = it;
      } else if (it == p.maxit) {
        st = -100;
        its = it + 1;  // the Fortran DO variable ends at MAXIT + 1
      }
    }
    rho1 = rho;
    if (sys) {
      // every system's state reaches every block through one more grid barrier (all blocks run
      // the same iterations, a finished system's x and r no longer change)
      if (t == 0) p.run[pb] = (st == 0) ? 1 : 0;
      abort |= grid_barrier(p.bar, phase);
      int any = 0;
      for (int g = t; g < (int)gridDim.x; g += 256) any |= p.run[g];
      alldone = __syncthreads_or(any) == 0;
    } else {
      alldone = block_all(st != 0);
    }
    if (prof) {
      const unsigned long long ce = wall_clock64();
      titer += ce - ci0;
      tpost += ce - cr1;
    }
  }
  if (prof) {
    trest = titer - tfw - tbw - tbar;
    if (t == 0) {
      unsigned long long *d = p.dbg + (size_t)blockIdx.x * AFX_DBG_SLOTS;
      d[0] += tfw;
      d[1] += tbar;
      d[2] += tbw;
      d[3] += trest;
      d[4] += titer;
      d[6] += trho;
      d[7] += tpost;
      d[8] += tpu;
      d[9] += tsp;
      d[10] += tclr;
    }
    atomicAdd(p.dbg + (size_t)blockIdx.x * AFX_DBG_SLOTS + 5, (unsigned long long)nfb);
  }
  // the state of each column (or of each stacked system, from its first block)
  if (sys ? pb == g0 : pb == 0) {
    if (cvalid && t < kp) {
      p.conv[cst] = st;
      p.rho1[cst] = rho1;
      p.resid[cst] = res;
      p.iters[cst] = its;
    }
  }
  if (pb == 0 && t == 0) {
    p.status[0] = it;
    p.status[1] = alldone ? 1 : 0;
  }
}

// One block of 256 threads per CU (the cooperative grid never exceeds the CU count), so each
// SIMD runs a single wave of this kernel: let it use the whole register file instead of
// spilling the software-pipelined sweep registers to scratch. PROF selects the instantiation
// that records per-phase timing (its timers would otherwise hold scalar registers that the
// sweep loops need).
template <typename TP, int LPR, bool MULTI, bool PROF>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1))) k_cocg(CocgParams<TP> p) {
  cocg_body<TP, LPR, MULTI, PROF, true>(p);
}

// The same iteration without the sweeps (DIAG or NONE edge preconditioner, with or without the
// nodal correction, one column, FP64): every pass is a bandwidth-bound stream with no dependency
// chain, so the launch runs two waves per SIMD (two blocks per CU) to keep more loads in flight.
template <int LPR, bool PROF>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) k_cocg_flat(CocgParams<double> p) {
  cocg_body<double, LPR, false, PROF, false>(p);
}

// several columns without sweeps (one lane per row and column)
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1))) k_cocg_flat_multi(CocgParams<double> p) {
  cocg_body<double, 1, true, false, false>(p);
}

// ---------------------------------------------------------------------------------------
// Several columns (k <= K) of the iteration without sweeps, the columns of a row kept by the
// same lanes: the SpMV reads each matrix entry once for all k columns (K + i w_c M per column c)
// and the vector passes stream whole rows of the [n][k] vectors. Each column keeps its own
// alpha, beta, stopping test and state, as k independent COCG13_single iterations (same
// recurrence and unconjugated forms). Reductions: per-wave XOR-shuffle sums, per-block sums in
// LDS, per-block partials [G][k] added in block order by every block (deterministic).

// g = G^T r at interior node a for column c of the [n][k] vector R, incidences added in list order;
// the first NI incidences and their r are all requested before the first use.
template <int NI>
__device__ __forceinline__ cx<double> nodal_g_cols(const int *__restrict__ nptr, const int *__restrict__ nadj,
                                                   const cx<double> *R, int k, int c, int a) {
  const int q0 = nptr[a], q1 = nptr[a + 1];
  const int ql = q1 > 0 ? q1 - 1 : 0;
  int v[NI];
  cx<double> rr[NI];
#pragma unroll
  for (int j = 0; j < NI; j++) {
    const int q = q0 + j;
    const int vq = nadj[q < q1 ? q : ql];  // loaded unconditionally: a fixed count of loads
    v[j] = q < q1 ? vq : 0;
  }
#pragma unroll
  for (int j = 0; j < NI; j++) {
    const int e = v[j] > 0 ? v[j] - 1 : (v[j] < 0 ? -v[j] - 1 : 0);
    rr[j] = R[(size_t)e * k + c];
  }
  cx<double> g = {0.0, 0.0};
#pragma unroll
  for (int j = 0; j < NI; j++) {
    if (v[j] > 0) {
      g.r += rr[j].r;
      g.i += rr[j].i;
    } else if (v[j] < 0) {
      g.r -= rr[j].r;
      g.i -= rr[j].i;
    }
  }
  for (int q = q0 + NI; q < q1; q++) {
    const int vq = nadj[q];
    const cx<double> r = R[(size_t)((vq > 0 ? vq : -vq) - 1) * k + c];
    g.r += vq > 0 ? r.r : -r.r;
    g.i += vq > 0 ? r.i : -r.i;
  }
  return g;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
  return v;
}

// v[c] summed over the grid for c < k; every thread returns the totals. lds: 4*K complex.
template <int K>
__device__ __forceinline__ void grid_sum_blk(cx<double> (&v)[K], int k, cx<double> *part, cx<double> *lds,
                                             const GridBar &bar, unsigned &phase, bool &abort, bool poll) {
  const int t = threadIdx.x, wid = t >> 6, lane = t & 63;
#pragma unroll
  for (int c = 0; c < K; c++) {
    v[c].r = wave_sum(v[c].r);
    v[c].i = wave_sum(v[c].i);
  }
  if (lane == 0)
#pragma unroll
    for (int c = 0; c < K; c++) lds[wid * K + c] = v[c];
  __syncthreads();
  if (t < k) {
    cx<double> b = lds[t];
    for (int w = 1; w < 4; w++) b = addv(b, lds[w * K + t]);
    part[(size_t)blockIdx.x * k + t] = b;
  }
  abort |= grid_barrier(bar, phase, poll);
  // totals: thread t sums column t % K over the blocks t / K, t / K + 256 / K, ... in order
  constexpr int J = 256 / K;
  const int c = t % K, j = t / K;
  cx<double> sum = {0.0, 0.0};
  if (c < k)
    for (int g = j; g < (int)gridDim.x; g += J) sum = addv(sum, part[(size_t)g * k + c]);
  __shared__ cx<double> red[256];
  red[t] = sum;
  __syncthreads();
  for (int h = J >> 1; h >= 1; h >>= 1) {
    if (j < h) red[t] = addv(red[t], red[t + h * K]);
    __syncthreads();
  }
#pragma unroll
  for (int cc = 0; cc < K; cc++) v[cc] = red[cc];
  __syncthreads();
}

// v (this thread's column c) added into acc[c] without indexing a register array at run time
template <int K, typename V>
__device__ __forceinline__ void add_col(V (&acc)[K], int c, V v) {
#pragma unroll
  for (int cc = 0; cc < K; cc++)
    if (cc == c) acc[cc] = addv(acc[cc], v);
}
template <int K, typename V>
__device__ __forceinline__ V pick_col(const V (&a)[K], int c) {
  V r = a[0];
#pragma unroll
  for (int cc = 1; cc < K; cc++)
    if (cc == c) r = a[cc];
  return r;
}

template <int K>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1))) k_cocg_blk(CocgParams<double> p) {
  constexpr int LPR = 4, SE = 5;
  __shared__ cx<double> lds_k[4 * K];
  const int t = threadIdx.x;
  const int gtid = blockIdx.x * 256 + t;
  const int nthr = gridDim.x * 256;
  const int n = p.n, k = p.k;
  const bool diag = p.pc == AFX_PC_DIAG;
  const bool nblk = p.pc == AFX_PC_NODEBLK;
  // element passes over the [n][k] vectors: element i = row * k + c, a stride that is a multiple
  // of k so every thread keeps one column (consecutive threads, consecutive 16-byte elements)
  const int estride = (nthr / k) * k;
  const int ec = gtid % k;
  const bool eact = gtid < estride;
  const long long nk = (long long)n * k, nnk = (long long)p.nn * k;
  unsigned phase = 0;
  if (t == 0) phase = __hip_atomic_load(p.bar.gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  double w[K], bn[K], res[K];
  int st[K], its[K];
  cx<double> rho1[K];
#pragma unroll
  for (int c = 0; c < K; c++) {
    const bool v = c < k;
    w[c] = v ? p.wcol[c] : 0.0;
    st[c] = v ? p.conv[c] : 1;
    bn[c] = v ? p.bnrm2[c] : 1.0;
    rho1[c] = v ? p.rho1[c] : cx<double>{0.0, 0.0};
    res[c] = v ? p.resid[c] : 0.0;
    its[c] = v ? p.iters[c] : 0;
  }
  const double we = p.wcol[ec];
  const int fe = p.col2f[ec];
  int it = p.status[0];
  const int it_end = min(it + p.chunk, p.maxit);
  bool alldone = true;
#pragma unroll
  for (int c = 0; c < K; c++) alldone = alldone && st[c] != 0;
  bool abort = false;
  // optional per-block phase timing (wall clock, 100 MHz): z pass, rho sum, p update, SpMV, c1 sum,
  // x/r pass, rr sum
  unsigned long long tph[7] = {0, 0, 0, 0, 0, 0, 0}, tmark = 0;
  const bool prof = p.dbg != nullptr;
#define AFX_BLK_MARK(q)                                 \
  if (prof) {                                           \
    const unsigned long long now_ = wall_clock64();     \
    tph[q] += now_ - tmark;                             \
    tmark = now_;                                       \
  }
  while (!alldone && it < it_end && !abort) {
    ++it;
    if (prof) tmark = wall_clock64();
    // ---- z = M^-1 r and rho = r^T z per column ----
    cx<double> acc[K];
#pragma unroll
    for (int c = 0; c < K; c++) acc[c] = cx<double>{0.0, 0.0};
    cx<double> mine = {0.0, 0.0};
    if (p.grad && eact)
      for (long long i = gtid; i < nnk; i += estride) {  // (node, column) elements
        const int a = (int)(i / k);
        const cx<double> g = nodal_g_cols<16>(p.nptr, p.nadj, p.R, k, ec, a);
        const cx<double> zn = nodal_z(g, we, p.mnd[a]);
        p.Zn[i] = zn;
        mine = addv(mine, cmul(g, zn));
      }
    if (eact && nblk)
      // node blocks: z of (row, column) from the row's 8 coefficients (the column's frequency
      // slice) and the column's r on the rows of the row's block (at most 8 rows, stride k)
      for (long long i0 = gtid; i0 < nk; i0 += 2LL * estride) {
        cx<double> rb[2][AFX_NB_W], rv[2];
        uint4 cv[2][AFX_NB_W / 4];
#pragma unroll
        for (int u = 0; u < 2; u++) {
          const long long i = i0 + u * (long long)estride < nk ? i0 + u * (long long)estride : i0;
          const long long row = i / k;
          const int m = p.bm[row], loc = m & 15, sz = (m >> 4) + 1;
          const uint4 *cf = (const uint4 *)p.bi + ((size_t)fe * n + row) * (AFX_NB_W / 4);
#pragma unroll
          for (int j = 0; j < AFX_NB_W / 4; j++) cv[u][j] = cf[j];
#pragma unroll
          for (int j = 0; j < AFX_NB_W; j++) rb[u][j] = p.R[(row - loc + min(j, sz - 1)) * k + ec];
          rv[u] = p.R[i];
        }
#pragma unroll
        for (int u = 0; u < 2; u++) {
          const long long i = i0 + u * (long long)estride;
          if (i < nk) {
            cx<double> z = {0.0, 0.0};
#pragma unroll
            for (int j = 0; j < AFX_NB_W; j++) z = cfma(nb_coef(cv[u][j >> 2], j & 3), rb[u][j], z);
            p.Z[i] = z;
            mine = addv(mine, cmul(rv[u], z));
          }
        }
      }
    else if (eact)
      for (long long i0 = gtid; i0 < nk; i0 += 4LL * estride) {
        cx<double> rv[4], dv[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
          const long long i = i0 + u * (long long)estride < nk ? i0 + u * (long long)estride : i0;
          rv[u] = p.R[i];
          dv[u] = diag ? p.DD[(size_t)(i / k) * p.nf + fe] : cx<double>{1.0, 0.0};
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
          const long long i = i0 + u * (long long)estride;
          if (i < nk) {
            const cx<double> z = cmul(rv[u], dv[u]);
            p.Z[i] = z;
            mine = addv(mine, cmul(rv[u], z));
          }
        }
      }
    add_col<K>(acc, ec, mine);
    AFX_BLK_MARK(0)
    grid_sum_blk<K>(acc, k, p.part_rho, lds_k, p.bar, phase, abort, false);
    AFX_BLK_MARK(1)
    cx<double> rho[K], beta[K];
    const bool first = (it == 1);
#pragma unroll
    for (int c = 0; c < K; c++) {
      rho[c] = acc[c];
      beta[c] = first ? cx<double>{0.0, 0.0} : cdivd(rho[c], rho1[c]);
    }
    // ---- p = z + G zn (+ beta p) ----
    {
      const cx<double> be = pick_col<K>(beta, ec);
      if (eact)
        for (long long i0 = gtid; i0 < nk; i0 += 4LL * estride) {
          cx<double> zv[4], pv[4], za[4], zb[4];
#pragma unroll
          for (int u = 0; u < 4; u++) {
            const long long i = i0 + u * (long long)estride < nk ? i0 + u * (long long)estride : i0;
            zv[u] = p.Z[i];
            pv[u] = first ? cx<double>{0.0, 0.0} : p.P[i];
            if (p.grad) {
              const int2 e = ((const int2 *)p.en)[i / k];
              za[u] = p.Zn[(size_t)e.x * k + ec];
              zb[u] = p.Zn[(size_t)e.y * k + ec];
            } else {
              za[u] = zb[u] = cx<double>{0.0, 0.0};
            }
          }
#pragma unroll
          for (int u = 0; u < 4; u++) {
            const long long i = i0 + u * (long long)estride;
            const cx<double> z = {zv[u].r + (zb[u].r - za[u].r), zv[u].i + (zb[u].i - za[u].i)};
            if (i < nk) p.P[i] = first ? z : cfma(be, pv[u], z);
          }
        }
    }
    abort |= grid_barrier(p.bar, phase);
    AFX_BLK_MARK(2)
    // ---- q = A p, c1 = p^T q: LPR lanes per row, a row's entries read once for all columns;
    // every gather of a row (its SE entries x all K columns, a column block being one contiguous
    // run of the [n][k] vector) is requested before the first product. The row's first entry is
    // its diagonal, so lane 0's first gathers are the row's own p, which c1 needs. (32 lanes per
    // row, a lane per (entry slot, column), gathered each column block with one instruction but
    // kept 8x fewer rows in flight: 4x slower at 64x64x48.) ----
#pragma unroll
    for (int c = 0; c < K; c++) acc[c] = cx<double>{0.0, 0.0};
    {
      const int lane = t % LPR;
      for (int row = gtid / LPR; row < n; row += nthr / LPR) {
        const int qb = p.rowptr[row], qe = p.rowptr[row + 1];
        cx<double> vv[SE], pv[SE][K];
        int cl[SE];
#pragma unroll
        for (int e = 0; e < SE; e++) {
          const int q = qb + lane + e * LPR;
          const bool in = q < qe;
          const int qc = in ? q : qb;
          vv[e] = p.val[qc];
          const int cq = p.col[qc];
          cl[e] = in ? cq : -1;
        }
#pragma unroll
        for (int e = 0; e < SE; e++)
#pragma unroll
          for (int c = 0; c < K; c++) pv[e][c] = p.P[(size_t)max(cl[e], 0) * k + (c < k ? c : k - 1)];
        cx<double> sum[K];
#pragma unroll
        for (int c = 0; c < K; c++) {
          sum[c] = cx<double>{0.0, 0.0};
#pragma unroll
          for (int e = 0; e < SE; e++)
            if (cl[e] >= 0) sum[c] = cfma(mval(vv[e], w[c]), pv[e][c], sum[c]);
        }
        for (int q = qb + lane + SE * LPR; q < qe; q += LPR)
#pragma unroll
          for (int c = 0; c < K; c++)
            if (c < k) sum[c] = cfma(mval(p.val[q], w[c]), p.P[(size_t)p.col[q] * k + c], sum[c]);
#pragma unroll
        for (int c = 0; c < K; c++) {
          sum[c] = group_sum<LPR>(sum[c]);
          if (lane == 0 && c < k) {
            p.Q[(size_t)row * k + c] = sum[c];
            acc[c] = addv(acc[c], cmul(pv[0][c], sum[c]));
          }
        }
      }
    }
    AFX_BLK_MARK(3)
    grid_sum_blk<K>(acc, k, p.part_c1, lds_k, p.bar, phase, abort, false);
    AFX_BLK_MARK(4)
    cx<double> alpha[K];
    bool upd[K];
#pragma unroll
    for (int c = 0; c < K; c++) {
      const cx<double> c1 = acc[c];
      alpha[c] = cdivd(rho[c], c1);
      if (st[c] == 0 && (!(c1.r != 0.0 || c1.i != 0.0) || !isfinite(alpha[c].r) || !isfinite(alpha[c].i) ||
                         !(rho[c].r != 0.0 || rho[c].i != 0.0))) {
        st[c] = -200;  // breakdown
        its[c] = it;
      }
      upd[c] = st[c] == 0;
    }
    // ---- x += alpha p, r -= alpha q, ||r||^2 per column ----
    cx<double> rr[K];
#pragma unroll
    for (int c = 0; c < K; c++) rr[c] = cx<double>{0.0, 0.0};
    {
      const cx<double> al = pick_col<K>(alpha, ec);
      const bool ue = pick_col<K>(upd, ec);
      double rre = 0.0;
      if (eact && ue)
        for (long long i0 = gtid; i0 < nk; i0 += 4LL * estride) {
          cx<double> xv[4], pv[4], qv[4], rv[4];
#pragma unroll
          for (int u = 0; u < 4; u++) {
            const long long i = i0 + u * (long long)estride < nk ? i0 + u * (long long)estride : i0;
            xv[u] = p.X[i];
            pv[u] = p.P[i];
            qv[u] = p.Q[i];
            rv[u] = p.R[i];
          }
#pragma unroll
          for (int u = 0; u < 4; u++) {
            const long long i = i0 + u * (long long)estride;
            if (i < nk) {
              p.X[i] = cfma(al, pv[u], xv[u]);
              const cx<double> r = cfma(cx<double>{-al.r, -al.i}, qv[u], rv[u]);
              p.R[i] = r;
              rre += r.r * r.r + r.i * r.i;
            }
          }
        }
      add_col<K>(rr, ec, cx<double>{rre, 0.0});
    }
    AFX_BLK_MARK(5)
    grid_sum_blk<K>(rr, k, (cx<double> *)p.part_rr, lds_k, p.bar, phase, abort, true);
    AFX_BLK_MARK(6)
    alldone = true;
#pragma unroll
    for (int c = 0; c < K; c++) {
      if (upd[c]) {
        res[c] = sqrt(rr[c].r / bn[c]);  // :315 RESID = sqrt(DNRM2/BNRM2)
        if (res[c] <= p.tol) {
          st[c] = 1;
          its[c] = it;
        } else if (it == p.maxit) {
          st[c] = -100;
          its[c] = it + 1;
        }
      }
      rho1[c] = rho[c];
      alldone = alldone && st[c] != 0;
    }
  }
#undef AFX_BLK_MARK
  if (prof && t == 0)
    for (int q = 0; q < 7; q++) p.dbg[(size_t)blockIdx.x * AFX_DBG_SLOTS + q] += tph[q];
  if (blockIdx.x == 0 && t == 0) {
#pragma unroll
    for (int c = 0; c < K; c++)
      if (c < k) {
        p.conv[c] = st[c];
        p.rho1[c] = rho1[c];
        p.resid[c] = res[c];
        p.iters[c] = its[c];
      }
    p.status[0] = it;
    p.status[1] = alldone ? 1 : 0;
  }
}

// ---------------------------------------------------------------------------------------
// One column without sweeps (edge preconditioner DIAG or NONE, with or without the nodal
// correction), FP64: one block of 256 threads per CU, three grid synchronisations per iteration
// instead of four (with node blocks and the nodal correction a fourth, after the tiles, lets the
// nodal loop gather the r' they wrote). The recurrence is COCG13_single's (:181-323); the
// preconditioner apply of iteration it+1 moves into the x, r pass of iteration it:
//   P   p = z + beta p (:258-268), z = D r + G zn formed on the fly (z is never stored);
//   S   q = A p (:273-288) and c1 = p^T q (:293-296)                      -> reduction of c1;
//   XN  x += alpha p, r' = r - alpha q (:304-307) into the other residual buffer, |r'|^2
//       (:310-313) and the edge part r'^T D r' of the next rho (:246-250); on the interior
//       nodes g' = G^T r', each r'_e recomputed from r and q by the same operation as the edge
//       pass (so g' is G^T of the stored r' bit for bit), zn' = g' / (i w mnd) and the nodal
//       part g'^T zn' of the next rho                                     -> one reduction of
//       (|r'|^2, rho').
// The residual alternates between R (before odd iterations) and Z (before even ones). A launch
// that starts the solve computes z and rho of the initial residual first; a launch that resumes
// takes rho from fin_rho. All reductions add per-block partials in block order in every block
// (identical scalars everywhere, deterministic). The c1 partials and the (rho, |r|^2) partials
// live in separate halves of part4: a block may write the next reduction's partials while a
// slower block still reads the last one's, and only the c1 -> x, r step has no other grid
// barrier in between.
template <int NV>
__device__ __forceinline__ void block_sum_n(double (&v)[NV], double *lds) {
  constexpr int NW = 256 / 64;
#pragma unroll
  for (int i = 0; i < NV; i++) v[i] = wave_sum(v[i]);
  const int wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int i = 0; i < NV; i++) lds[wv * NV + i] = v[i];
  __syncthreads();
#pragma unroll
  for (int i = 0; i < NV; i++) {
    double s = lds[i];
#pragma unroll
    for (int q = 1; q < NW; q++) s += lds[q * NV + i];
    v[i] = s;
  }
  __syncthreads();
}

template <int NV>
__device__ __forceinline__ void grid_sum_n(double (&v)[NV], double *part, double *lds, const GridBar &bar,
                                           unsigned &phase, bool &abort) {
  block_sum_n<NV>(v, lds);
  if (threadIdx.x == 0)
#pragma unroll
    for (int i = 0; i < NV; i++) part[(size_t)blockIdx.x * NV + i] = v[i];
  abort |= grid_barrier(bar, phase);
  double s[NV];
#pragma unroll
  for (int i = 0; i < NV; i++) s[i] = 0.0;
  for (int g = threadIdx.x; g < (int)gridDim.x; g += 256)
#pragma unroll
    for (int i = 0; i < NV; i++) s[i] += part[(size_t)g * NV + i];
  block_sum_n<NV>(s, lds);
#pragma unroll
  for (int i = 0; i < NV; i++) v[i] = s[i];
}

// g = G^T r' at interior node a with r' = r - alpha q recomputed per incident edge (the x, r
// pass's operation), incidences added in list order. The first NI incidences, their r and their
// q are all requested before the first use (a node of a tetrahedral mesh has about 14 edges).
template <int NI>
__device__ __forceinline__ cx<double> nodal_g_next(const int *__restrict__ nptr, const int *__restrict__ nadj,
                                                   const cx<double> *R, const cx<double> *Q, cx<double> malpha,
                                                   long long a) {
  const int q0 = nptr[a], q1 = nptr[a + 1];
  const int ql = q1 > 0 ? q1 - 1 : 0;
  int v[NI];
  cx<double> rr[NI], qq[NI];
#pragma unroll
  for (int j = 0; j < NI; j++) {
    const int q = q0 + j;
    const int vq = nadj[q < q1 ? q : ql];  // loaded unconditionally: a fixed count of loads
    v[j] = q < q1 ? vq : 0;
  }
#pragma unroll
  for (int j = 0; j < NI; j++) {
    const int e = v[j] > 0 ? v[j] - 1 : (v[j] < 0 ? -v[j] - 1 : 0);
    rr[j] = R[e];
    qq[j] = Q[e];
  }
  cx<double> g = {0.0, 0.0};
#pragma unroll
  for (int j = 0; j < NI; j++) {
    const cx<double> rn = cfma(malpha, qq[j], rr[j]);
    if (v[j] > 0) {
      g.r += rn.r;
      g.i += rn.i;
    } else if (v[j] < 0) {
      g.r -= rn.r;
      g.i -= rn.i;
    }
  }
  for (int q = q0 + NI; q < q1; q++) {
    const int vq = nadj[q];
    const int e = (vq > 0 ? vq : -vq) - 1;
    const cx<double> rn = cfma(malpha, Q[e], R[e]);
    if (vq > 0) {
      g.r += rn.r;
      g.i += rn.i;
    } else {
      g.r -= rn.r;
      g.i -= rn.i;
    }
  }
  return g;
}

// One block per CU, so each SIMD runs a single wave: the whole register file holds the
// software-pipelined SpMV rounds and 8 (p pass) or 4 (x, r pass) rows per thread in flight.
template <int LPR, bool PROF>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1))) k_cocg1(CocgParams<double> p) {
  __shared__ double lds[4 * 4];
  __shared__ cx<double> tile_r[AFX_NB_UT][256];  // node-block pass: r' of the round's tiles
  constexpr int UP = 8, UX = 4;  // rows per thread in flight in the p pass and the x, r pass
  const int t = threadIdx.x;
  const long long gtid = (long long)blockIdx.x * 256 + t;
  const long long nthr = (long long)gridDim.x * 256;
  const long long n = p.n;
  const double w = p.wcol[0];
  const bool diag = (p.pc == AFX_PC_DIAG);
  const bool nblk = (p.pc == AFX_PC_NODEBLK);
  const bool grad = p.grad != 0;
  const cx<double> *__restrict__ DD = p.DD;
  cx<double> *const buf[2] = {p.R, p.Z};
  double *const part_c1 = p.part4, *const part_xn = p.part4 + 2 * (size_t)gridDim.x;  // [G][2], [G][3]
  unsigned phase = 0;
  if (t == 0) phase = __hip_atomic_load(p.bar.gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  int st = p.conv[0];
  const double bn = p.bnrm2[0];
  cx<double> rho1 = p.rho1[0], rho = p.fin_rho[0];
  double res = p.resid[0];
  int its = p.iters[0];
  int it = p.status[0];
  const int it_end = min(it + p.chunk, p.maxit);
  bool abort = false;
  unsigned long long tp = 0, ts = 0, tx = 0, ttot = 0;

  // edge part of rho = r^T z for the residual r (z_e = D r, or r without a preconditioner)
  auto edge_rho = [&](cx<double> r, cx<double> d, double (&acc)[3]) {
    const cx<double> z = diag ? cmul(r, d) : r;
    const cx<double> rz = cmul(r, z);
    acc[0] += rz.r;
    acc[1] += rz.i;
  };
  // nodal part: zn = g / (i w mnd) stored, g^T zn added
  auto nodal_put = [&](long long a, cx<double> g, double (&acc)[3]) {
    const cx<double> zn = nodal_z(g, w, p.mnd[a]);
    p.Zn[a] = zn;
    const cx<double> gz = cmul(g, zn);
    acc[0] += gz.r;
    acc[1] += gz.i;
  };

  // node-block pass: over tiles of whole blocks (UT tiles per round, every load of a round issued
  // first), x += al p when Xx is given, r' = r - alpha q (q null: r' = r) into Rn when Rn is given,
  // |r'|^2 into acc[2], Zb = Bi r' with r' staged in LDS, and the edge part r'^T Zb of rho into
  // acc[0..1]
  auto nblk_pass = [&](const cx<double> *Rr, const cx<double> *Qq, cx<double> ma, cx<double> *Rn, cx<double> *Xx,
                       cx<double> al, double (&acc)[3]) {
    constexpr int UT = AFX_NB_UT;
    for (long long t0 = blockIdx.x; t0 < p.ntile; t0 += UT * (long long)gridDim.x) {
      cx<double> rv[UT], qv[UT], xv[UT], pv[UT];
      uint4 cv[UT][AFX_NB_W / 4];
      int row[UT], m[UT];
      bool ok[UT];
#pragma unroll
      for (int u = 0; u < UT; u++) {
        const long long tl = t0 + u * (long long)gridDim.x;
        const int rs = tl < p.ntile ? p.tptr[tl] : 0, re = tl < p.ntile ? p.tptr[tl + 1] : 0;
        row[u] = rs + t;
        ok[u] = row[u] < re;
        const int rc = ok[u] ? row[u] : 0;
        rv[u] = Rr[rc];
        qv[u] = Qq ? Qq[rc] : cx<double>{0.0, 0.0};
        xv[u] = Xx ? Xx[rc] : cx<double>{0.0, 0.0};
        pv[u] = Xx ? p.P[rc] : cx<double>{0.0, 0.0};
        m[u] = ok[u] ? (int)p.bm[rc] : 0;
#pragma unroll
        for (int j = 0; j < AFX_NB_W / 4; j++) cv[u][j] = ((const uint4 *)p.bi)[(size_t)rc * (AFX_NB_W / 4) + j];
      }
#pragma unroll
      for (int u = 0; u < UT; u++) {
        const cx<double> rn = ok[u] ? cfma(ma, qv[u], rv[u]) : cx<double>{0.0, 0.0};
        if (ok[u]) {
          if (Xx) Xx[row[u]] = cfma(al, pv[u], xv[u]);
          if (Rn) Rn[row[u]] = rn;
          acc[2] += rn.r * rn.r + rn.i * rn.i;
        }
        rv[u] = rn;
        tile_r[u][t] = rn;
      }
      __syncthreads();
#pragma unroll
      for (int u = 0; u < UT; u++) {
        const int i = m[u] & 15, sz = (m[u] >> 4) + 1;
        cx<double> z = {0.0, 0.0};
#pragma unroll
        for (int j = 0; j < AFX_NB_W; j++)
          z = cfma(nb_coef(cv[u][j >> 2], j & 3), tile_r[u][t - i + min(j, sz - 1)], z);
        if (ok[u]) {
          p.Zb[row[u]] = z;
          const cx<double> rz = cmul(rv[u], z);
          acc[0] += rz.r;
          acc[1] += rz.i;
        }
      }
      __syncthreads();
    }
  };

  if (st == 0 && it == 0) {  // prologue: z and rho of the initial residual (in R)
    double acc[3] = {0.0, 0.0, 0.0};
    const cx<double> *R = p.R;
    if (nblk)
      nblk_pass(R, nullptr, cx<double>{0.0, 0.0}, nullptr, nullptr, cx<double>{0.0, 0.0}, acc);
    else
      for (long long r0 = gtid; r0 < n; r0 += nthr) edge_rho(R[r0], diag ? DD[r0] : cx<double>{0.0, 0.0}, acc);
    if (grad)
      for (long long a = gtid; a < p.nn; a += nthr) nodal_put(a, nodal_g(p.nptr, p.nadj, R, (int)a), acc);
    double v[2] = {acc[0], acc[1]};
    grid_sum_n<2>(v, part_xn, lds, p.bar, phase, abort);
    rho = {v[0], v[1]};
  }
  bool alldone = (st != 0);
  while (!alldone && it < it_end && !abort) {
    ++it;
    const unsigned long long c0 = PROF ? wall_clock64() : 0;
    const cx<double> *R = buf[(it - 1) & 1];
    cx<double> *Rn = buf[it & 1];
    const bool first = (it == 1);
    const cx<double> beta = first ? cx<double>{0.0, 0.0} : cdivd(rho, rho1);

    // ---- P: p = z + beta p ----
    for (long long r0 = gtid; r0 < n; r0 += UP * nthr) {
      cx<double> rv[UP], dv[UP], pv[UP], za[UP], zb[UP];
      int2 ev[UP];
#pragma unroll
      for (int u = 0; u < UP; u++) {
        const long long r = r0 + u * nthr < n ? r0 + u * nthr : r0;
        rv[u] = nblk ? p.Zb[r] : R[r];  // node blocks: Bi r, formed by the last r pass
        dv[u] = diag ? DD[r] : cx<double>{0.0, 0.0};
        pv[u] = first ? cx<double>{0.0, 0.0} : p.P[r];
        ev[u] = grad ? ((const int2 *)p.en)[r] : int2{0, 0};
      }
#pragma unroll
      for (int u = 0; u < UP; u++) {
        za[u] = grad ? p.Zn[ev[u].x] : cx<double>{0.0, 0.0};
        zb[u] = grad ? p.Zn[ev[u].y] : cx<double>{0.0, 0.0};
      }
#pragma unroll
      for (int u = 0; u < UP; u++) {
        const long long r = r0 + u * nthr;
        const cx<double> ze = diag ? cmul(rv[u], dv[u]) : rv[u];
        const cx<double> z = {ze.r + (zb[u].r - za[u].r), ze.i + (zb[u].i - za[u].i)};
        if (r < n) p.P[r] = first ? z : cfma(beta, pv[u], z);
      }
    }
    abort |= grid_barrier(p.bar, phase);
    const unsigned long long c1t = PROF ? wall_clock64() : 0;

    // ---- S: q = A p, c1 = p^T q ----
    cx<double> c1l = {0.0, 0.0};
    spmv_pipe<LPR>(p.rowptr, p.col, p.val, w, p.P, p.Q, n, gtid / LPR, nthr / LPR, n, (int)(gtid % LPR), c1l);
    double cacc[2] = {c1l.r, c1l.i};
    grid_sum_n<2>(cacc, part_c1, lds, p.bar, phase, abort);
    const unsigned long long c2t = PROF ? wall_clock64() : 0;
    const cx<double> c1 = {cacc[0], cacc[1]};
    const cx<double> alpha = cdivd(rho, c1);
    if (!(c1.r != 0.0 || c1.i != 0.0) || !isfinite(alpha.r) || !isfinite(alpha.i) || !(rho.r != 0.0 || rho.i != 0.0)) {
      st = -200;  // breakdown: x and r stay as they are
      its = it;
      alldone = true;
      if (PROF) ts += c2t - c1t;
      break;
    }
    const cx<double> malpha = {-alpha.r, -alpha.i};

    // ---- XN: x += alpha p, r' = r - alpha q, |r'|^2, and z, rho of r' ----
    double acc[3] = {0.0, 0.0, 0.0};  // rho' (re, im), |r'|^2
    if (nblk) {
      nblk_pass(R, p.Q, malpha, Rn, p.X, alpha, acc);
    } else
    for (long long r0 = gtid; r0 < n; r0 += UX * nthr) {
      cx<double> xv[UX], pv[UX], qv[UX], rv[UX], dv[UX];
#pragma unroll
      for (int u = 0; u < UX; u++) {
        const long long r = r0 + u * nthr < n ? r0 + u * nthr : r0;
        xv[u] = p.X[r];
        pv[u] = p.P[r];
        qv[u] = p.Q[r];
        rv[u] = R[r];
        dv[u] = diag ? DD[r] : cx<double>{0.0, 0.0};
      }
#pragma unroll
      for (int u = 0; u < UX; u++) {
        const long long r = r0 + u * nthr;
        if (r < n) {
          p.X[r] = cfma(alpha, pv[u], xv[u]);
          const cx<double> rn = cfma(malpha, qv[u], rv[u]);
          Rn[r] = rn;
          acc[2] += rn.r * rn.r + rn.i * rn.i;
          edge_rho(rn, dv[u], acc);
        }
      }
    }
    if (grad && nblk && p.nadj16) {
      // r' is in Rn once every block's tiles are done. Each node's 16 incidence slots load as four
      // 16-byte vectors one node ahead, and its r' values are gathered (16 B per incidence, not r
      // and q at 32 B): the same values added in the same order as nodal_g_next.
      abort |= grid_barrier(p.bar, phase);
      const int4 *A16 = (const int4 *)p.nadj16;
      int4 cur[4];
      if (gtid < p.nn)
#pragma unroll
        for (int j = 0; j < 4; j++) cur[j] = A16[gtid * 4 + j];
      for (long long a = gtid; a < p.nn; a += nthr) {
        const long long an = a + nthr < p.nn ? a + nthr : a;
        int4 nxt[4];
#pragma unroll
        for (int j = 0; j < 4; j++) nxt[j] = A16[an * 4 + j];
        const int q0 = p.nptr[a], q1 = p.nptr[a + 1];
        int v[16];
#pragma unroll
        for (int j = 0; j < 4; j++) {
          v[4 * j] = cur[j].x;
          v[4 * j + 1] = cur[j].y;
          v[4 * j + 2] = cur[j].z;
          v[4 * j + 3] = cur[j].w;
        }
        cx<double> rr[16];
#pragma unroll
        for (int j = 0; j < 16; j++) rr[j] = Rn[v[j] > 0 ? v[j] - 1 : (v[j] < 0 ? -v[j] - 1 : 0)];
        cx<double> g = {0.0, 0.0};
#pragma unroll
        for (int j = 0; j < 16; j++) {
          if (v[j] > 0) {
            g.r += rr[j].r;
            g.i += rr[j].i;
          } else if (v[j] < 0) {
            g.r -= rr[j].r;
            g.i -= rr[j].i;
          }
        }
        for (int q = q0 + 16; q < q1; q++) {  // incidences beyond the 16 slots
          const int vq = p.nadj[q];
          const cx<double> rn = Rn[(vq > 0 ? vq : -vq) - 1];
          g.r += vq > 0 ? rn.r : -rn.r;
          g.i += vq > 0 ? rn.i : -rn.i;
        }
        nodal_put(a, g, acc);
#pragma unroll
        for (int j = 0; j < 4; j++) cur[j] = nxt[j];
      }
    } else if (grad)
      for (long long a = gtid; a < p.nn; a += nthr) nodal_put(a, nodal_g_next<16>(p.nptr, p.nadj, R, p.Q, malpha, a), acc);
    grid_sum_n<3>(acc, part_xn, lds, p.bar, phase, abort);
    const double rr = acc[2];
    res = sqrt(rr / bn);  // :315 RESID = sqrt(DNRM2/BNRM2)
    if (res <= p.tol) {
      st = 1;
      its = it;
    } else if (it == p.maxit) {
      st = -100;
      its = it + 1;  // the Fortran DO variable ends at MAXIT + 1
    }
    rho1 = rho;
    rho = {acc[0], acc[1]};
    alldone = (st != 0);
    if (PROF) {
      const unsigned long long c3t = wall_clock64();
      tp += c1t - c0;
      ts += c2t - c1t;
      tx += c3t - c2t;
      ttot += c3t - c0;
    }
  }
  if (PROF && t == 0) {
    unsigned long long *d = p.dbg + (size_t)blockIdx.x * AFX_DBG_SLOTS;
    d[0] += tp;
    d[2] += ts;
    d[3] += tx;
    d[4] += ttot;
  }
  if (blockIdx.x == 0 && t == 0) {
    p.conv[0] = st;
    p.rho1[0] = rho1;
    p.fin_rho[0] = rho;
    p.resid[0] = res;
    p.iters[0] = its;
    p.status[0] = it;
    p.status[1] = alldone ? 1 : 0;
  }
}

template <bool PROF>
static const void *cocg1_kernel(int lpr) {
  switch (lpr) {
    case 1:
    case 2: return (const void *)k_cocg1<2, PROF>;
    case 4: return (const void *)k_cocg1<4, PROF>;
    case 16: return (const void *)k_cocg1<16, PROF>;
    default: return (const void *)k_cocg1<8, PROF>;
  }
}

// Blocks of the fused single-column kernel that are co-resident, at most one per CU.
int max_coop_grid_cocg1(int lpr) {
  int dev = 0, ncu = 0, nb = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, cocg1_kernel<false>(lpr), 256, 0) != hipSuccess) return 0;
  return std::min(nb, 1) * ncu;
}

hipError_t launch_cocg1(const CocgParams<double> &p, int lpr, int grid, hipStream_t s) {
  if (p.k != 1 || p.nsys > 1 || (p.pc != AFX_PC_DIAG && p.pc != AFX_PC_NONE && p.pc != AFX_PC_NODEBLK) || !p.part4 ||
      (p.pc == AFX_PC_NODEBLK && (!p.Zb || !p.bi || !p.tptr || !p.bm)))
    return hipErrorInvalidValue;
  void *args[] = {(void *)&p};
  const void *f = p.dbg ? cocg1_kernel<true>(lpr) : cocg1_kernel<false>(lpr);
  return hipLaunchCooperativeKernel(f, dim3(grid), dim3(256), args, 0, s);
}

hipError_t launch_cocg_blk(const CocgParams<double> &p, int grid, hipStream_t s) {
  void *args[] = {(void *)&p};
  const void *f = p.k <= 2 ? (const void *)k_cocg_blk<2> : p.k <= 4 ? (const void *)k_cocg_blk<4>
                                                                    : (const void *)k_cocg_blk<8>;
  if (p.k > 8) return hipErrorInvalidValue;
  return hipLaunchCooperativeKernel(f, dim3(grid), dim3(256), args, 0, s);
}

template <int LPR>
static const void *flat_kernel(bool prof) {
  return prof ? (const void *)k_cocg_flat<LPR, true> : (const void *)k_cocg_flat<LPR, false>;
}
static const void *flat_kernel_for(int lpr, bool prof = false) {
  switch (lpr) {
    case 1: return flat_kernel<1>(prof);
    case 2: return flat_kernel<2>(prof);
    case 4: return flat_kernel<4>(prof);
    case 16: return flat_kernel<16>(prof);
    default: return flat_kernel<8>(prof);
  }
}

// Largest co-resident grid of the flat kernel (blocks), at most `per_cu` blocks per CU.
int max_coop_grid_diag(int lpr, int per_cu) {
  int dev = 0, ncu = 0, nb = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, flat_kernel_for(lpr), 256, 0) != hipSuccess) return 0;
  return std::min(nb, per_cu) * ncu;
}

hipError_t launch_cocg_diag(const CocgParams<double> &p, int lpr, int grid, hipStream_t s) {
  void *args[] = {(void *)&p};
  const void *f = p.k > 1 ? (const void *)k_cocg_flat_multi : flat_kernel_for(lpr, p.dbg != nullptr);
  return hipLaunchCooperativeKernel(f, dim3(grid), dim3(256), args, 0, s);
}

template <typename TP>
size_t cocg_dyn_lds(const CocgParams<TP> &p, int grid) {
  (void)grid;
  if (p.ring_w > 0) return (size_t)p.ring_w * p.k * sizeof(cx<TP>);
  if (p.lds_rows <= 0) return 0;
  return lds_bounds_offset<TP>(p.lds_rows) + (size_t)(p.fw.maxsteps + p.bw.maxsteps + 2) * sizeof(int);
}

template <typename TP, int LPR, bool MULTI, bool PROF>
static hipError_t launch_one(const CocgParams<TP> &p, int grid, hipStream_t s) {
  void *args[] = {(void *)&p};
  const size_t dyn = cocg_dyn_lds(p, grid);
  const void *f = (const void *)k_cocg<TP, LPR, MULTI, PROF>;
  hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn);
  if (e != hipSuccess) return e;
  return hipLaunchCooperativeKernel(f, dim3(grid), dim3(256), args, dyn, s);
}

template <typename TP, bool PROF>
static hipError_t launch_p(const CocgParams<TP> &p, int lpr, int grid, hipStream_t s) {
  if (p.k > 1) return launch_one<TP, 1, true, PROF>(p, grid, s);
  switch (lpr) {
    case 1: return launch_one<TP, 1, false, PROF>(p, grid, s);
    case 2: return launch_one<TP, 2, false, PROF>(p, grid, s);
    case 4: return launch_one<TP, 4, false, PROF>(p, grid, s);
    case 8: return launch_one<TP, 8, false, PROF>(p, grid, s);
    case 16: return launch_one<TP, 16, false, PROF>(p, grid, s);
    default: return hipErrorInvalidValue;
  }
}

template <typename TP>
static hipError_t launch_t(const CocgParams<TP> &p, int lpr, int grid, hipStream_t s) {
  if (grid != p.Gs) return hipErrorInvalidValue;
  return p.dbg ? launch_p<TP, true>(p, lpr, grid, s) : launch_p<TP, false>(p, lpr, grid, s);
}

hipError_t launch_cocg64(const CocgParams<double> &p, int lpr, int grid, hipStream_t s) { return launch_t(p, lpr, grid, s); }
hipError_t launch_cocg32(const CocgParams<float> &p, int lpr, int grid, hipStream_t s) { return launch_t(p, lpr, grid, s); }

template <typename TP, int LPR, bool MULTI>
static int occ_one(size_t dyn) {
  int nb = 0, nbp = 0;
  const void *f = (const void *)k_cocg<TP, LPR, MULTI, false>, *fp = (const void *)k_cocg<TP, LPR, MULTI, true>;
  if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn) != hipSuccess ||
      hipFuncSetAttribute(fp, hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn) != hipSuccess)
    return 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, 256, dyn) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&nbp, fp, 256, dyn) != hipSuccess)
    return 0;
  return std::min(nb, nbp);
}

// Largest co-resident grid (blocks) for the cooperative launch with `dyn` bytes of LDS.
int max_coop_grid(int lpr, bool multi, bool fp64, size_t dyn) {
  int dev = 0, ncu = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  int nb = 0;
  if (fp64) {
    if (multi) nb = occ_one<double, 1, true>(dyn);
    else nb = (lpr == 1) ? occ_one<double, 1, false>(dyn) : (lpr == 2) ? occ_one<double, 2, false>(dyn)
            : (lpr == 4) ? occ_one<double, 4, false>(dyn) : (lpr == 8) ? occ_one<double, 8, false>(dyn)
                                                          : occ_one<double, 16, false>(dyn);
  } else {
    if (multi) nb = occ_one<float, 1, true>(dyn);
    else nb = (lpr == 1) ? occ_one<float, 1, false>(dyn) : (lpr == 2) ? occ_one<float, 2, false>(dyn)
            : (lpr == 4) ? occ_one<float, 4, false>(dyn) : (lpr == 8) ? occ_one<float, 8, false>(dyn)
                                                         : occ_one<float, 16, false>(dyn);
  }
  return nb * ncu;
}

// ---------------------------------------------------------------------------------------
// Plain (non-persistent) FP64 kernels: SpMV / residual, norms, value gathers, casts.

// Y = A X or Y = B - A X, one lane per (row, column) for k > 1, 8 lanes per row for k == 1;
// optional partial |Y|^2 and |B|^2 per block and column.
__global__ void __launch_bounds__(256) k_spmv(SpmvParams p) {
  __shared__ double lds[256];
  const int t = threadIdx.x;
  const int k = p.k;
  const bool multi = k > 1;
  const int kp = multi ? p.kp : 1;
  const int c = multi ? t % kp : 0;
  const long long gtid = (long long)blockIdx.x * 256 + t;
  const long long nthr = (long long)gridDim.x * 256;
  const int LPR = 8;
  const long long slot = multi ? gtid / kp : gtid / LPR;
  const long long nslot = multi ? nthr / kp : nthr / LPR;
  const int lane = multi ? 0 : (int)(gtid % LPR);
  double rr = 0.0, bb = 0.0;
  const bool cvalid = c < k;
  const double w = cvalid ? p.wcol[c] : 0.0;
  for (long long row = slot; row < p.n; row += nslot) {
    const int qb = p.rowptr[row], qe = p.rowptr[row + 1];
    cx<double> acc = {0.0, 0.0};
    if (multi) {
      if (cvalid)
        for (int q = qb; q < qe; ++q) acc = cfma(mval(p.val[q], w), p.X[(size_t)p.col[q] * k + c], acc);
    } else {
      for (int q = qb + lane; q < qe; q += LPR) acc = cfma(mval(p.val[q], w), p.X[p.col[q]], acc);
      acc = group_sum<8>(acc);
    }
    if (cvalid && lane == 0) {
      const size_t ix = (size_t)row * k + c;
      cx<double> y = acc;
      if (p.B) {
        const cx<double> b = p.B[ix];
        y = b - acc;
        bb += b.r * b.r + b.i * b.i;
      }
      if (p.Y) p.Y[ix] = y;
      rr += y.r * y.r + y.i * y.i;
    }
  }
  if (p.part_rr) {
    const double s = block_cols(rr, lds, kp);
    const double s2 = block_cols(bb, lds, kp);
    if (t < kp && t < k) {
      p.part_rr[(size_t)blockIdx.x * k + t] = s;
      p.part_bb[(size_t)blockIdx.x * k + t] = s2;
    }
  }
}

__global__ void k_norm_finish(int G, int k, const double *part_rr, const double *part_bb, double *res) {
  __shared__ double lds[256];
  const int c = blockIdx.x;
  double a = 0.0, b = 0.0;
  for (int g = threadIdx.x; g < G; g += 256) {
    a += part_rr[(size_t)g * k + c];
    b += part_bb[(size_t)g * k + c];
  }
  a = block_cols(a, lds, 1);
  b = block_cols(b, lds, 1);
  if (threadIdx.x == 0) {
    res[2 * c] = a;      // ||b - A x||^2
    res[2 * c + 1] = b;  // ||b||^2
  }
}

hipError_t launch_spmv(const SpmvParams &p, hipStream_t s) {
  long long rows_per_pass = (p.k > 1) ? (256 / p.kp) : 32;
  long long g = (p.n + rows_per_pass - 1) / rows_per_pass;
  if (g > 65536) g = 65536;
  if (g < 1) g = 1;
  SpmvParams q = p;
  q.part_rr = nullptr;
  q.part_bb = nullptr;
  hipLaunchKernelGGL(k_spmv, dim3((unsigned)g), dim3(256), 0, s, q);
  return hipGetLastError();
}

// Y = A(w) X for one column as a kernel of its own: the q = A p of COCG13_single.f90:270-290.
// A full CSR row holds the diagonal, the upper and the lower entries in that order; LPR lanes
// read a row, lane j taking entries qb + j + LPR e (e < SE; longer rows finish in a loop), and
// the lane sums are added by a fixed DPP tree. A wave takes SU sets of 64/LPR consecutive rows,
// so each value and column request of the wave is one contiguous stream; every load of a set
// is issued before the first wait (row bounds, then values and columns, then the X gathers).
// Blocks take consecutive row ranges in launch order. Measured on the 1,407,152-edge system
// (505.9 MB of compulsory bytes): 100 us, 5.0 TB/s; a read-only stream of 1 GiB reaches
// 6.0 TB/s on the same GPU. Numbering the blocks XCD-major (each XCD a contiguous share of the
// rows), interleaving row bands inside a share, non-temporal value loads, several rounds per
// wave and 8 waves per SIMD were each measured and were no faster: the first two cut the L2
// misses from 608 to 495 MB per launch and still took 5-11% longer.
template <int LPR, int SU, int SE>
__global__ void __launch_bounds__(256) k_spmv_stream(int n, const int *__restrict__ rowptr,
                                                     const int *__restrict__ col,
                                                     const cx<double> *__restrict__ val, double w,
                                                     const cx<double> *__restrict__ X, cx<double> *__restrict__ Y) {
  constexpr int RPS = 64 / LPR;  // rows per set
  const int lane64 = threadIdx.x & 63;
  const int lane = lane64 % LPR;
  const long long row0 = ((long long)blockIdx.x * 4 + (threadIdx.x >> 6)) * (RPS * SU) + lane64 / LPR;
  if (row0 - lane64 / LPR >= n) return;  // wave-uniform
  int qb[SU], qe[SU], cl[SU][SE];
  cx<double> vv[SU][SE], xv[SU][SE];
#pragma unroll
  for (int u = 0; u < SU; u++) {
    const long long r = row0 + u * RPS;
    const long long rc = r < n ? r : n - 1;
    qb[u] = rowptr[rc];
    qe[u] = rowptr[rc + 1];  // loaded unconditionally: a conditional load makes later waits conservative
    if (r >= n) qe[u] = qb[u];
  }
#pragma unroll
  for (int u = 0; u < SU; u++)
#pragma unroll
    for (int e = 0; e < SE; e++) {
      const int q = qb[u] + lane + e * LPR;
      const bool in = q < qe[u];
      const int qc = in ? q : 0;  // entry 0 exists (nnz >= 1 is checked by the caller)
      vv[u][e] = val[qc];
      const int cq = col[qc];
      cl[u][e] = in ? cq : -1;
    }
#pragma unroll
  for (int u = 0; u < SU; u++)
#pragma unroll
    for (int e = 0; e < SE; e++) xv[u][e] = X[max(cl[u][e], 0)];
#pragma unroll
  for (int u = 0; u < SU; u++) {
    cx<double> acc = {0.0, 0.0};
#pragma unroll
    for (int e = 0; e < SE; e++)
      if (cl[u][e] >= 0) acc = cfma(mval(vv[u][e], w), xv[u][e], acc);
    for (int q = qb[u] + lane + SE * LPR; q < qe[u]; q += LPR) acc = cfma(mval(val[q], w), X[col[q]], acc);
    const cx<double> a = group_sum<LPR>(acc);
    const long long r = row0 + u * RPS;
    if (lane == 0 && r < n) Y[r] = a;
  }
}

template <int LPR, int SU, int SE>
hipError_t launch_spmv_stream_t(int n, const int *rowptr, const int *col, const cx<double> *val, double w,
                                const cx<double> *X, cx<double> *Y, hipStream_t s) {
  constexpr long long rows_per_block = 4LL * (64 / LPR) * SU;
  const long long nb = (n + rows_per_block - 1) / rows_per_block;
  if (nb > 0x7fffffffLL) return hipErrorInvalidValue;
  hipLaunchKernelGGL((k_spmv_stream<LPR, SU, SE>), dim3((unsigned)nb), dim3(256), 0, s, n, rowptr, col, val, w, X,
                     Y);
  return hipGetLastError();
}

// Lanes per row from the longest row: rows of up to 20 entries (the edge-element systems) take
// 4 lanes x 5 entries, up to 24 take 8 x 3, longer rows 16 x 2 and the tail loop.
hipError_t launch_spmv_stream(int n, int maxrow, const int *rowptr, const int *col, const cx<double> *val, double w,
                              const cx<double> *X, cx<double> *Y, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  if (maxrow <= 20) return launch_spmv_stream_t<4, 2, 5>(n, rowptr, col, val, w, X, Y, s);
  if (maxrow <= 24) return launch_spmv_stream_t<8, 4, 3>(n, rowptr, col, val, w, X, Y, s);
  return launch_spmv_stream_t<16, 2, 2>(n, rowptr, col, val, w, X, Y, s);
}

// Per-block partial sums for the minimal-residual scaling of an initial guess, per column c:
// part[(g*k + c)*4 + {0,1,2,3}] = sum over rows of Re conj(y) b, Im conj(y) b, |y|^2, |b|^2.
__global__ void __launch_bounds__(256) k_warm_dots(long long n, int k, int kp, const cx<double> *Y,
                                                   const cx<double> *B, double *part) {
  __shared__ double lds[256];
  const int t = threadIdx.x;
  const int c = t % kp;
  const bool cvalid = c < k;
  const long long slot = ((long long)blockIdx.x * 256 + t) / kp;
  const long long nslot = (long long)gridDim.x * 256 / kp;
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  if (cvalid)
    for (long long row = slot; row < n; row += nslot) {
      const cx<double> y = Y[row * k + c], b = B[row * k + c];
      v[0] += y.r * b.r + y.i * b.i;
      v[1] += y.r * b.i - y.i * b.r;
      v[2] += y.r * y.r + y.i * y.i;
      v[3] += b.r * b.r + b.i * b.i;
    }
  for (int q = 0; q < 4; q++) {
    const double sum = block_cols(v[q], lds, kp);
    if (t < kp && cvalid) part[((size_t)blockIdx.x * k + t) * 4 + q] = sum;
  }
}
hipError_t launch_warm_dots(long long n, int k, int kp, int grid, const cx<double> *Y, const cx<double> *B,
                            double *part, hipStream_t s) {
  hipLaunchKernelGGL(k_warm_dots, dim3(grid), dim3(256), 0, s, n, k, kp, Y, B, part);
  return hipGetLastError();
}

// X[row][c] *= g[c]
__global__ void k_scale_cols_cx(long long n, int k, const cx<double> *g, cx<double> *X) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i < n * k) X[i] = cmul(g[i % k], X[i]);
}
hipError_t launch_scale_cols_cx(long long n, int k, const cx<double> *g, cx<double> *X, hipStream_t s) {
  hipLaunchKernelGGL(k_scale_cols_cx, dim3((unsigned)((n * k + 255) / 256)), dim3(256), 0, s, n, k, g, X);
  return hipGetLastError();
}

// One column of a row-interleaved (n, k) array to and from a contiguous n-vector.
__global__ void k_col_get(long long n, int k, int c, const cx<double> *A, cx<double> *v) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i < n) v[i] = A[i * k + c];
}
__global__ void k_col_put(long long n, int k, int c, const cx<double> *v, cx<double> *A) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i < n) A[i * k + c] = v[i];
}
hipError_t launch_col_get(long long n, int k, int c, const cx<double> *A, cx<double> *v, hipStream_t s) {
  hipLaunchKernelGGL(k_col_get, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, k, c, A, v);
  return hipGetLastError();
}
hipError_t launch_col_put(long long n, int k, int c, const cx<double> *v, cx<double> *A, hipStream_t s) {
  hipLaunchKernelGGL(k_col_put, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, k, c, v, A);
  return hipGetLastError();
}
// columns c0 .. c0+g-1 of A [n][k] to / from a [n][g] block
__global__ void k_cols_get(long long n, int k, int c0, int g, const cx<double> *A, cx<double> *B) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i < n * g) B[i] = A[(i / g) * k + c0 + i % g];
}
__global__ void k_cols_put(long long n, int k, int c0, int g, const cx<double> *B, cx<double> *A) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i < n * g) A[(i / g) * k + c0 + i % g] = B[i];
}
hipError_t launch_cols_get(long long n, int k, int c0, int g, const cx<double> *A, cx<double> *B, hipStream_t s) {
  hipLaunchKernelGGL(k_cols_get, dim3((unsigned)((n * g + 255) / 256)), dim3(256), 0, s, n, k, c0, g, A, B);
  return hipGetLastError();
}
hipError_t launch_cols_put(long long n, int k, int c0, int g, const cx<double> *B, cx<double> *A, hipStream_t s) {
  hipLaunchKernelGGL(k_cols_put, dim3((unsigned)((n * g + 255) / 256)), dim3(256), 0, s, n, k, c0, g, B, A);
  return hipGetLastError();
}

hipError_t launch_residual_norms(const SpmvParams &p, int grid, double *res_out_dev, hipStream_t s) {
  hipLaunchKernelGGL(k_spmv, dim3(grid), dim3(256), 0, s, p);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_norm_finish, dim3(p.k), dim3(256), 0, s, grid, p.k, p.part_rr, p.part_bb, res_out_dev);
  return hipGetLastError();
}

// DD for one level set; `nf` frequencies per row, lane = (row slot, frequency).
// kind IC: W = sd*D_i - sum_{j<i} a_ij (a_ij + sig*sum_U(j)) DD_j, DD_i = 1/W (COCG13_single.f90:107-124)
// kind SSOR/DIAG: DD_i = 1/(sd*D_i) (:126-130)
__global__ void __launch_bounds__(256) k_precond_ic(int n, int nf, int nfp, double sd, double sig,
                                                     const cx<double> *Dkm, int nlevL, const int *lvlL,
                                                     const int *rowL, const int *Lptr, const int *Lcol,
                                                     const cx<double> *Lval, const int *INU,
                                                     const cx<double> *AUkm, const double *wf, cx<double> *DD) {
  cg::grid_group grid = cg::this_grid();
  const int t = threadIdx.x;
  const long long gtid = (long long)blockIdx.x * 256 + t;
  const long long nthr = (long long)gridDim.x * 256;
  const int f = t % nfp;
  const long long slot = gtid / nfp, nslot = nthr / nfp;
  const bool fvalid = f < nf;
  const double w = fvalid ? wf[f] : 0.0;
  for (int lv = 0; lv < nlevL; ++lv) {
    const int lo = lvlL[lv], hi = lvlL[lv + 1];
    if (fvalid)
      for (long long pos = lo + slot; pos < hi; pos += nslot) {
        const int row = rowL[pos];
        const cx<double> dv = mval(Dkm[row], w);
        cx<double> W = {sd * dv.r, sd * dv.i};
        for (int q = Lptr[pos]; q < Lptr[pos + 1]; ++q) {
          const cx<double> a = mval(Lval[q], w);
          const int j = Lcol[q];
          cx<double> ss = a;
          if (sig != 0.0)
            for (int u = INU[j]; u < INU[j + 1]; ++u) {
              const cx<double> au = mval(AUkm[u], w);
              ss.r += au.r * sig;
              ss.i += au.i * sig;
            }
          W = W - cmul(cmul(a, ss), DD[(size_t)j * nf + f]);
        }
        DD[(size_t)row * nf + f] = cdivd(cx<double>{1.0, 0.0}, W);
      }
    grid.sync();
  }
}

__global__ void k_precond_diag(int n, int nf, double sd, const cx<double> *Dkm, const double *wf, cx<double> *DD) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= (long long)n * nf) return;
  const int row = (int)(i / nf), f = (int)(i % nf);
  const cx<double> dv = mval(Dkm[row], wf[f]);
  DD[i] = cdivd(cx<double>{1.0, 0.0}, cx<double>{sd * dv.r, sd * dv.i});
}

hipError_t launch_precond_setup(int n, int nf, int kind, double sd, double sig, const cx<double> *Dkm,
                                const int *, const cx<double> *, const int *, int nlevL, const int *lvlL,
                                const int *rowL, const int *Lptr, const int *Lcol, const cx<double> *Lval,
                                const int *INU, const cx<double> *AUkm, const double *wf, cx<double> *DD,
                                int grid, hipStream_t s) {
  if (kind == AFX_PC_IC) {
    int nfp = 1;
    while (nfp < nf) nfp <<= 1;
    if (nfp > 256) return hipErrorInvalidValue;
    void *args[] = {&n, &nf, &nfp, &sd, &sig, (void *)&Dkm, &nlevL, (void *)&lvlL, (void *)&rowL, (void *)&Lptr,
                    (void *)&Lcol, (void *)&Lval, (void *)&INU, (void *)&AUkm, (void *)&wf, (void *)&DD};
    return hipLaunchCooperativeKernel((const void *)k_precond_ic, dim3(grid), dim3(256), args, 0, s);
  }
  if (kind == AFX_PC_SSOR || kind == AFX_PC_DIAG || kind == AFX_PC_SSOR_BJ) {
    long long m = (long long)n * nf;
    hipLaunchKernelGGL(k_precond_diag, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, n, nf, sd, Dkm, wf, DD);
    return hipGetLastError();
  }
  return hipSuccess;
}

int max_coop_grid_ic() {
  int nb = 0, dev = 0, ncu = 0;
  hipGetDevice(&dev);
  hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, (const void *)k_precond_ic, 256, 0);
  return nb * ncu;
}

// One level of the preconditioner sweeps with ordinary launches (parity hook, single column).
__global__ void k_level_fwd(int lo, int hi, const int *rowL, const int *Lptr, const int *Lcol, const cx<double> *Lval,
                            const cx<double> *DD, const double *w, const cx<double> *r, cx<double> *z) {
  const int pos = lo + blockIdx.x * 256 + threadIdx.x;
  if (pos >= hi) return;
  const int row = rowL[pos];
  cx<double> acc = {0.0, 0.0};
  for (int q = Lptr[pos]; q < Lptr[pos + 1]; ++q) acc = cfma(mval(Lval[q], w[0]), z[Lcol[q]], acc);
  z[row] = cmul(r[row] - acc, DD[row]);
}
__global__ void k_level_bwd(int lo, int hi, const int *rowU, const int *Uptr, const int *Ucol, const cx<double> *Uval,
                            const cx<double> *DD, const double *w, cx<double> *z) {
  const int pos = lo + blockIdx.x * 256 + threadIdx.x;
  if (pos >= hi) return;
  const int row = rowU[pos];
  cx<double> acc = {0.0, 0.0};
  for (int q = Uptr[pos]; q < Uptr[pos + 1]; ++q) acc = cfma(mval(Uval[q], w[0]), z[Ucol[q]], acc);
  z[row] = z[row] - cmul(DD[row], acc);
}
__global__ void k_diag_apply(int n, int kind, const cx<double> *DD, const cx<double> *r, cx<double> *z) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) z[i] = (kind == AFX_PC_DIAG) ? cmul(r[i], DD[i]) : r[i];
}

hipError_t launch_precond_apply_levels(int n, int kind, int nlevL, const int *lvlL, const int *rowL,
                                       const int *Lptr, const int *Lcol, const cx<double> *Lval, int nlevU,
                                       const int *lvlU, const int *rowU, const int *Uptr, const int *Ucol,
                                       const cx<double> *Uval, const cx<double> *DD, const double *w,
                                       const cx<double> *r, cx<double> *z, hipStream_t s) {
  if (kind != AFX_PC_IC && kind != AFX_PC_SSOR) {
    hipLaunchKernelGGL(k_diag_apply, dim3((n + 255) / 256), dim3(256), 0, s, n, kind, DD, r, z);
    return hipGetLastError();
  }
  std::vector<int> hl(nlevL + 1), hu(nlevU + 1);
  hipError_t e = hipMemcpy(hl.data(), lvlL, (nlevL + 1) * sizeof(int), hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(hu.data(), lvlU, (nlevU + 1) * sizeof(int), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return e;
  for (int l = 0; l < nlevL; l++) {
    const int m = hl[l + 1] - hl[l];
    hipLaunchKernelGGL(k_level_fwd, dim3((m + 255) / 256), dim3(256), 0, s, hl[l], hl[l + 1], rowL, Lptr, Lcol, Lval,
                       DD, w, r, z);
  }
  for (int l = 0; l < nlevU; l++) {
    const int m = hu[l + 1] - hu[l];
    hipLaunchKernelGGL(k_level_bwd, dim3((m + 255) / 256), dim3(256), 0, s, hu[l], hu[l + 1], rowU, Uptr, Ucol, Uval,
                       DD, w, z);
  }
  return hipGetLastError();
}

// mnd[a] = diag(G^T M G)[a]: for each edge e at node a (sign s_ae) the entries M_ef of row e whose
// column f is also an edge at a (sign s_af), from the full CSR (K, M) pairs. One thread per node.
__global__ void k_nodal_diag(int nn, const int *nptr, const int *nadj, const int *en, const int *rowptr,
                             const int *col, const cx<double> *val, double *mnd) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= nn) return;
  double acc = 0.0;
  for (int q = nptr[a]; q < nptr[a + 1]; q++) {
    const int v = nadj[q];
    const int e = (v > 0 ? v : -v) - 1;
    const double se = v > 0 ? 1.0 : -1.0;
    for (int j = rowptr[e]; j < rowptr[e + 1]; j++) {
      const int f = col[j];
      const int f0 = en[2 * f], f1 = en[2 * f + 1];
      const double sf = (f1 == a) ? 1.0 : (f0 == a) ? -1.0 : 0.0;
      acc += se * sf * val[j].i;
    }
  }
  mnd[a] = acc;
}

// host-side apply (parity hook): zn = g / (i w mnd) on the nodes, then z[e] += zn[end] - zn[start]
__global__ void k_grad_nodes(int nn, const int *nptr, const int *nadj, const double *mnd, double w,
                             const cx<double> *r, cx<double> *zn) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a > nn) return;
  zn[a] = (a == nn) ? cx<double>{0.0, 0.0} : nodal_z(nodal_g(nptr, nadj, r, a), w, mnd[a]);
}
__global__ void k_grad_edges(int n, const int *en, const cx<double> *zn, cx<double> *z) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const cx<double> a = zn[en[2 * e]], b = zn[en[2 * e + 1]];
  z[e].r += b.r - a.r;
  z[e].i += b.i - a.i;
}

// Node-block Jacobi setup (AFX_PC_NODEBLK): the inverse of A's diagonal block over each node block
// (bm[e] = i | (size - 1) << 4, rows e - i .. e - i + size - 1), A = K + i w M from the full CSR
// (K, M) pairs. Eight lanes per block, lane i holding row i of the block padded to 8 x 8 with
// identity rows; in-place Gauss-Jordan without pivoting in FP64 (every principal submatrix of
// K + i w M with K positive semidefinite and w M positive definite is nonsingular; the identity
// padding leaves the block's inverse unchanged), the pivot row taken from its lane by shuffles;
// then (X + X^T) / 2 (the transpose through LDS) rounded to FP32 and then to bfloat16 (both to
// nearest even), so the stored operator is exactly symmetric. Row i of the block is stored as
// AFX_NB_W coefficients at bi[(e0 + i) * AFX_NB_W], zero beyond the block's size.
__global__ void __launch_bounds__(256) k_nodeblk_setup(int nblk, const int *__restrict__ bstart,
                                                       const unsigned char *__restrict__ bm,
                                                       const int *__restrict__ rowptr, const int *__restrict__ col,
                                                       const cx<double> *__restrict__ val, double w,
                                                       nbcoef *__restrict__ bi) {
  constexpr int W = AFX_NB_W;
  __shared__ cx<double> tr[256 / W][W][W];
  const int li = threadIdx.x & (W - 1), grp = threadIdx.x / W;
  const int g = blockIdx.x * (256 / W) + grp;
  const bool okg = g < nblk;
  const int e0 = okg ? bstart[g] : 0;
  const int sz = okg ? (bm[e0] >> 4) + 1 : 1;
  const bool live = okg && li < sz;
  cx<double> a[W];
#pragma unroll
  for (int j = 0; j < W; j++) a[j] = {j == li ? 1.0 : 0.0, 0.0};
  if (live) {
#pragma unroll
    for (int j = 0; j < W; j++) a[j] = {0.0, 0.0};
    for (int q = rowptr[e0 + li]; q < rowptr[e0 + li + 1]; q++) {
      const int c = col[q] - e0;
      const cx<double> v = mval(val[q], w);
#pragma unroll
      for (int j = 0; j < W; j++)
        if (c == j && j < sz) a[j] = v;
    }
  }
  const int base = threadIdx.x & ~(W - 1);
#pragma unroll
  for (int k = 0; k < W; k++) {
    cx<double> rk[W];
#pragma unroll
    for (int j = 0; j < W; j++) {
      rk[j].r = __shfl(a[j].r, base + k);
      rk[j].i = __shfl(a[j].i, base + k);
    }
    const cx<double> piv = cdivd(cx<double>{1.0, 0.0}, rk[k]);
    if (li == k) {
#pragma unroll
      for (int j = 0; j < W; j++) a[j] = (j == k) ? piv : cmul(rk[j], piv);
    } else {
      const cx<double> f = a[k];
#pragma unroll
      for (int j = 0; j < W; j++)
        a[j] = (j == k) ? cmul(cx<double>{-f.r, -f.i}, piv) : cfma(cx<double>{-f.r, -f.i}, cmul(rk[j], piv), a[j]);
    }
  }
#pragma unroll
  for (int j = 0; j < W; j++) tr[grp][li][j] = a[j];
  __syncthreads();
  if (!live) return;
#pragma unroll
  for (int j = 0; j < W; j++) {
    const cx<double> t = tr[grp][j][li];
    const cx<double> v = (j < sz) ? cx<double>{0.5 * (a[j].r + t.r), 0.5 * (a[j].i + t.i)} : cx<double>{0.0, 0.0};
    bi[(size_t)(e0 + li) * W + j] = f2bf16((float)v.r) | (f2bf16((float)v.i) << 16);
  }
}

// z = Bi r for the parity hook: one block per tile of whole node blocks, r staged in LDS.
__global__ void __launch_bounds__(256) k_nodeblk_apply(const int *__restrict__ tptr, const unsigned char *__restrict__ bm,
                                                       const nbcoef *__restrict__ bi, const cx<double> *__restrict__ r,
                                                       cx<double> *__restrict__ z) {
  __shared__ cx<double> tr[256];
  const int t = threadIdx.x;
  const int rs = tptr[blockIdx.x], re = tptr[blockIdx.x + 1];
  const int row = rs + t;
  const bool ok = row < re;
  tr[t] = ok ? r[row] : cx<double>{0.0, 0.0};
  __syncthreads();
  if (!ok) return;
  const int m = bm[row], i = m & 15, sz = (m >> 4) + 1;
  cx<double> acc = {0.0, 0.0};
#pragma unroll
  for (int j = 0; j < AFX_NB_W; j++) {
    const uint4 g = ((const uint4 *)bi)[(size_t)row * (AFX_NB_W / 4) + (j >> 2)];
    acc = cfma(nb_coef(g, j & 3), tr[t - i + min(j, sz - 1)], acc);
  }
  z[row] = acc;
}

hipError_t launch_nodeblk_setup(int nblk, const int *bstart, const unsigned char *bm, const int *rowptr,
                                const int *col, const cx<double> *val, double w, nbcoef *bi, hipStream_t s) {
  if (nblk <= 0) return hipSuccess;
  constexpr int per = 256 / AFX_NB_W;
  hipLaunchKernelGGL(k_nodeblk_setup, dim3((nblk + per - 1) / per), dim3(256), 0, s, nblk, bstart, bm, rowptr, col,
                     val, w, bi);
  return hipGetLastError();
}
hipError_t launch_nodeblk_apply(int ntile, const int *tptr, const unsigned char *bm, const nbcoef *bi,
                                const cx<double> *r, cx<double> *z, hipStream_t s) {
  if (ntile <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_nodeblk_apply, dim3(ntile), dim3(256), 0, s, tptr, bm, bi, r, z);
  return hipGetLastError();
}

hipError_t launch_nodal_diag(int nn, const int *nptr, const int *nadj, const int *en, const int *rowptr,
                             const int *col, const cx<double> *val, double *mnd, hipStream_t s) {
  if (nn <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_nodal_diag, dim3((nn + 255) / 256), dim3(256), 0, s, nn, nptr, nadj, en, rowptr, col, val, mnd);
  return hipGetLastError();
}
hipError_t launch_grad_apply(int n, int nn, const int *nptr, const int *nadj, const int *en, const double *mnd,
                             double w, const cx<double> *r, cx<double> *zn, cx<double> *z, hipStream_t s) {
  hipLaunchKernelGGL(k_grad_nodes, dim3((nn + 1 + 255) / 256), dim3(256), 0, s, nn, nptr, nadj, mnd, w, r, zn);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_grad_edges, dim3((n + 255) / 256), dim3(256), 0, s, n, en, zn, z);
  return hipGetLastError();
}

__global__ void k_gather(int m, const int *map, const cx<double> *canon, cx<double> *out, cx<float> *out32) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= m) return;
  const int src = map[i];
  const cx<double> v = (src >= 0) ? canon[src] : cx<double>{0.0, 0.0};
  if (out) out[i] = v;
  if (out32) out32[i] = cx<float>{(float)v.r, (float)v.i};
}

hipError_t launch_gather_vals(int m, const int *map, const cx<double> *canon, cx<double> *out, cx<float> *out32,
                              hipStream_t s) {
  if (m <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_gather, dim3((m + 255) / 256), dim3(256), 0, s, m, map, canon, out, out32);
  return hipGetLastError();
}

// position-ordered copy of the preconditioner diagonal (single frequency)
__global__ void k_dpos(int n, const int *row, const cx<double> *DD, cx<double> *d64, cx<float> *d32) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const cx<double> v = DD[row[i]];
  d64[i] = v;
  d32[i] = cx<float>{(float)v.r, (float)v.i};
}
hipError_t launch_dpos(int n, const int *row, const cx<double> *DD, cx<double> *d64, cx<float> *d32, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_dpos, dim3((n + 255) / 256), dim3(256), 0, s, n, row, DD, d64, d32);
  return hipGetLastError();
}

__global__ void k_cast(long long m, const cx<double> *in, cx<float> *out) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i < m) out[i] = cx<float>{(float)in[i].r, (float)in[i].i};
}
hipError_t launch_cast_vec(long long m, const cx<double> *in, cx<float> *out, hipStream_t s) {
  if (m <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_cast, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, m, in, out);
  return hipGetLastError();
}

// every mailbox slot to MB_EMPTY (a fill kernel, so the solve can sit in a captured graph)
__global__ void k_mb_fill(long long words, unsigned long long *mb) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i < words) mb[i] = MB_EMPTY;
}
hipError_t launch_mb_fill(long long words, unsigned long long *mb, hipStream_t s) {
  if (words <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_mb_fill, dim3((unsigned)((words + 255) / 256)), dim3(256), 0, s, words, mb);
  return hipGetLastError();
}

// x += d (FP32 correction into the FP64 iterate)
__global__ void k_axpy32(long long m, const cx<float> *d, cx<double> *x) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i < m) {
    x[i].r += (double)d[i].r;
    x[i].i += (double)d[i].i;
  }
}
hipError_t launch_axpy_f32_to_f64(long long m, const cx<float> *d, cx<double> *x, hipStream_t s) {
  if (m <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_axpy32, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, m, d, x);
  return hipGetLastError();
}

// out[i][c] = in[i][c] * scale[c] (FP64 -> FP32), the normalised residual of a refinement step
__global__ void k_scale_cols(long long n, int k, const cx<double> *in, const double *scale, cx<float> *out) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n * k) return;
  const double s = scale[i % k];
  out[i] = cx<float>{(float)(in[i].r * s), (float)(in[i].i * s)};
}
hipError_t launch_scale_cols(long long n, int k, const cx<double> *in, const double *scale, cx<float> *out,
                             hipStream_t s) {
  const long long m = n * k;
  if (m <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_scale_cols, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, n, k, in, scale, out);
  return hipGetLastError();
}

// x[i][c] += d[i][c] * scale[c] (undo the normalisation, accumulate in FP64)
__global__ void k_axpy_scaled(long long n, int k, const cx<float> *d, const double *scale, cx<double> *x) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n * k) return;
  const double s = scale[i % k];
  x[i].r += (double)d[i].r * s;
  x[i].i += (double)d[i].i * s;
}
hipError_t launch_axpy_scaled(long long n, int k, const cx<float> *d, const double *scale, cx<double> *x,
                              hipStream_t s) {
  const long long m = n * k;
  if (m <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_axpy_scaled, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, n, k, d, scale, x);
  return hipGetLastError();
}

}  // namespace afx
