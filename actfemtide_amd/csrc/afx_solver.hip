// COCG solve kernels for gfx950 (MI355X).
//
// The whole preconditioned COCG iteration of COCG13_single (src/solver/COCG13_single.f90:181-323)
// runs inside ONE persistent cooperative launch per chunk of iterations:
//   z = M^-1 r  -- level-scheduled forward (strict lower) and backward (strict upper) sweeps,
//                  one grid barrier per dependency level, rows of a level stored contiguously;
//   rho = r^T z -- fused into the backward sweep, reduced deterministically across blocks;
//   p = z + beta p;  q = A p (CSR, `lpr` lanes per row) fused with p^T q;
//   x += alpha p, r -= alpha q fused with ||r||^2 and the per-column convergence test.
// Bilinear forms are unconjugated (COCG); every reduction and every scalar (alpha, beta, rho)
// is FP64, also when vectors and matrix are complex FP32 (the mixed-precision inner solve).
//
// Several right-hand sides (columns) share each matrix read: vectors are stored row-interleaved
// [n][k], a lane owns one (row, column) pair, and each column runs its own recurrence with its
// own convergence flag. A column may carry its own angular frequency w: the matrix is held as
// (K, M) pairs and the entry used is K + i*w*M, so several tidal constituents share one launch.
#include <hip/hip_cooperative_groups.h>

#include <vector>

#include "afx_internal.h"

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
__device__ __forceinline__ cx<double> cdivd(cx<double> a, cx<double> b) {
  // Smith's algorithm
  cx<double> r;
  if (fabs(b.r) >= fabs(b.i)) {
    double t = b.i / b.r, d = b.r + b.i * t;
    r.r = (a.r + a.i * t) / d;
    r.i = (a.i - a.r * t) / d;
  } else {
    double t = b.r / b.i, d = b.r * t + b.i;
    r.r = (a.r * t + a.i) / d;
    r.i = (a.i * t - a.r) / d;
  }
  return r;
}
__device__ __forceinline__ double shfl_xor_d(double v, int m) { return __shfl_xor(v, m, 64); }
__device__ __forceinline__ float shfl_xor_d(float v, int m) { return __shfl_xor(v, m, 64); }

template <int LPR, typename T>
__device__ __forceinline__ cx<T> group_sum(cx<T> v) {
#pragma unroll
  for (int m = LPR / 2; m >= 1; m >>= 1) {
    v.r += shfl_xor_d(v.r, m);
    v.i += shfl_xor_d(v.i, m);
  }
  return v;
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

// Grid-wide per-column reduction with identical results in every block.
template <typename V>
__device__ V grid_cols(V v, V *part, V *fin, V *lds, int k, int kp, int c, bool dist,
                       cg::grid_group &grid) {
  const int t = threadIdx.x;
  const int G = gridDim.x;
  V bs = block_cols(v, lds, kp);
  if (t < kp && t < k) part[(size_t)blockIdx.x * k + t] = bs;
  grid.sync();
  if (!dist) {
    V s = zero_of(v);
    if (c < k)
      for (int g = t / kp; g < G; g += 256 / kp) s = addv(s, part[(size_t)g * k + c]);
    return block_cols(s, lds, kp);
  }
  for (int cc = blockIdx.x; cc < k; cc += G) {
    V s = zero_of(v);
    for (int g = t; g < G; g += 256) s = addv(s, part[(size_t)g * k + cc]);
    s = block_cols(s, lds, 1);
    if (t == 0) fin[cc] = s;
  }
  grid.sync();
  return (c < k) ? fin[c] : zero_of(v);
}

// all columns finished? identical in every block because the states are
__device__ bool block_all(bool mine) { return __syncthreads_and(mine ? 1 : 0) != 0; }

template <typename T, int LPR, bool MULTI>
__global__ void __launch_bounds__(256) k_cocg(CocgParams<T> p) {
  __shared__ cx<double> lds_c[256];
  __shared__ double lds_d[256];
  cg::grid_group grid = cg::this_grid();
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
  const T w = cvalid ? p.wcol[c] : T(0);
  const int f = cvalid ? p.col2f[c] : 0;
  const int nf = p.nf;
  const bool sweep = (p.pc == AFX_PC_IC || p.pc == AFX_PC_SSOR);
  const bool dist = p.dist_reduce != 0;

  int st = cvalid ? p.conv[c] : 1;
  const double bn = cvalid ? p.bnrm2[c] : 1.0;
  cx<double> rho1 = cvalid ? p.rho1[c] : cx<double>{0.0, 0.0};
  double res = cvalid ? p.resid[c] : 0.0;
  int its = cvalid ? p.iters[c] : 0;
  int it = p.status[0];
  const int it_end = min(it + p.chunk, p.maxit);
  bool alldone = block_all(st != 0);

  while (!alldone && it < it_end) {
    ++it;
    // ---- z = M^-1 r (COCG13_single.f90:186-241) and rho = r^T z (:246-250) ----
    cx<double> rho_loc = {0.0, 0.0};
    if (sweep) {
      for (int lv = 0; lv < p.nlevL; ++lv) {
        const int lo = p.lvlL[lv], hi = p.lvlL[lv + 1];
        for (long long pos = lo + slot; pos < hi; pos += nslot) {
          const int row = p.rowL[pos];
          const int qb = p.Lptr[pos], qe = p.Lptr[pos + 1];
          if (MULTI) {
            if (cvalid) {
              cx<T> acc = {T(0), T(0)};
              for (int q = qb; q < qe; ++q) acc = cfma(mval(p.Lval[q], w), p.Z[(size_t)p.Lcol[q] * k + c], acc);
              const size_t ix = (size_t)row * k + c;
              p.Z[ix] = cmul(p.R[ix] - acc, p.DD[(size_t)row * nf + f]);
            }
          } else {
            cx<T> acc = {T(0), T(0)};
            for (int q = qb + lane; q < qe; q += LPR) acc = cfma(mval(p.Lval[q], w), p.Z[p.Lcol[q]], acc);
            acc = group_sum<LPR>(acc);
            if (lane == 0) p.Z[row] = cmul(p.R[row] - acc, p.DD[row]);
          }
        }
        grid.sync();
      }
      for (int lv = 0; lv < p.nlevU; ++lv) {
        const int lo = p.lvlU[lv], hi = p.lvlU[lv + 1];
        for (long long pos = lo + slot; pos < hi; pos += nslot) {
          const int row = p.rowU[pos];
          const int qb = p.Uptr[pos], qe = p.Uptr[pos + 1];
          if (MULTI) {
            if (cvalid) {
              cx<T> acc = {T(0), T(0)};
              for (int q = qb; q < qe; ++q) acc = cfma(mval(p.Uval[q], w), p.Z[(size_t)p.Ucol[q] * k + c], acc);
              const size_t ix = (size_t)row * k + c;
              const cx<T> z = p.Z[ix] - cmul(p.DD[(size_t)row * nf + f], acc);
              p.Z[ix] = z;
              rho_loc = addv(rho_loc, cmul(up(p.R[ix]), up(z)));
            }
          } else {
            cx<T> acc = {T(0), T(0)};
            for (int q = qb + lane; q < qe; q += LPR) acc = cfma(mval(p.Uval[q], w), p.Z[p.Ucol[q]], acc);
            acc = group_sum<LPR>(acc);
            if (lane == 0) {
              const cx<T> z = p.Z[row] - cmul(p.DD[row], acc);
              p.Z[row] = z;
              rho_loc = addv(rho_loc, cmul(up(p.R[row]), up(z)));
            }
          }
        }
        grid.sync();
      }
    } else {
      if (MULTI) {
        if (cvalid)
          for (long long row = slot; row < p.n; row += nslot) {
            const size_t ix = (size_t)row * k + c;
            const cx<T> r = p.R[ix];
            const cx<T> z = (p.pc == AFX_PC_DIAG) ? cmul(r, p.DD[(size_t)row * nf + f]) : r;
            p.Z[ix] = z;
            rho_loc = addv(rho_loc, cmul(up(r), up(z)));
          }
      } else {
        for (long long row = gtid; row < p.n; row += nthr) {
          const cx<T> r = p.R[row];
          const cx<T> z = (p.pc == AFX_PC_DIAG) ? cmul(r, p.DD[row]) : r;
          p.Z[row] = z;
          rho_loc = addv(rho_loc, cmul(up(r), up(z)));
        }
      }
    }
    const cx<double> rho = grid_cols(rho_loc, p.part_rho, p.fin_rho, lds_c, k, kp, c, dist, grid);

    // ---- p = z (first iteration) or z + beta p (:258-268) ----
    const bool first = (it == 1);
    const cx<double> beta = first ? cx<double>{0.0, 0.0} : cdivd(rho, rho1);
    const cx<T> betaT = down<T>(beta);
    if (MULTI) {
      if (cvalid)
        for (long long row = slot; row < p.n; row += nslot) {
          const size_t ix = (size_t)row * k + c;
          p.P[ix] = first ? p.Z[ix] : cfma(betaT, p.P[ix], p.Z[ix]);
        }
    } else {
      for (long long row = gtid; row < p.n; row += nthr) p.P[row] = first ? p.Z[row] : cfma(betaT, p.P[row], p.Z[row]);
    }
    grid.sync();

    // ---- q = A p (:273-288) and c1 = p^T q (:293-296) ----
    cx<double> c1_loc = {0.0, 0.0};
    for (long long row = slot; row < p.n; row += nslot) {
      const int qb = p.rowptr[row], qe = p.rowptr[row + 1];
      if (MULTI) {
        if (cvalid) {
          cx<T> acc = {T(0), T(0)};
          for (int q = qb; q < qe; ++q) acc = cfma(mval(p.val[q], w), p.P[(size_t)p.col[q] * k + c], acc);
          const size_t ix = (size_t)row * k + c;
          p.Q[ix] = acc;
          c1_loc = addv(c1_loc, cmul(up(p.P[ix]), up(acc)));
        }
      } else {
        cx<T> acc = {T(0), T(0)};
        for (int q = qb + lane; q < qe; q += LPR) acc = cfma(mval(p.val[q], w), p.P[p.col[q]], acc);
        acc = group_sum<LPR>(acc);
        if (lane == 0) {
          p.Q[row] = acc;
          c1_loc = addv(c1_loc, cmul(up(p.P[row]), up(acc)));
        }
      }
    }
    const cx<double> c1 = grid_cols(c1_loc, p.part_c1, p.fin_c1, lds_c, k, kp, c, dist, grid);
    const cx<double> alpha = cdivd(rho, c1);
    const bool run = (st == 0);
    if (run && (!(c1.r != 0.0 || c1.i != 0.0) || !isfinite(alpha.r) || !isfinite(alpha.i) ||
                !(rho.r != 0.0 || rho.i != 0.0))) {
      st = -200;  // breakdown
      its = it;
    }
    const bool upd = (st == 0);
    const cx<T> alphaT = down<T>(alpha);
    const cx<T> malphaT = {-alphaT.r, -alphaT.i};

    // ---- x += alpha p, r -= alpha q (:304-307), ||r||^2 (:310-313) ----
    double rr_loc = 0.0;
    if (MULTI) {
      if (cvalid && upd)
        for (long long row = slot; row < p.n; row += nslot) {
          const size_t ix = (size_t)row * k + c;
          p.X[ix] = cfma(alphaT, p.P[ix], p.X[ix]);
          const cx<T> r = cfma(malphaT, p.Q[ix], p.R[ix]);
          p.R[ix] = r;
          rr_loc += (double)r.r * (double)r.r + (double)r.i * (double)r.i;
        }
    } else if (upd) {
      for (long long row = gtid; row < p.n; row += nthr) {
        p.X[row] = cfma(alphaT, p.P[row], p.X[row]);
        const cx<T> r = cfma(malphaT, p.Q[row], p.R[row]);
        p.R[row] = r;
        rr_loc += (double)r.r * (double)r.r + (double)r.i * (double)r.i;
      }
    }
    const double rr = grid_cols(rr_loc, p.part_rr, p.fin_rr, lds_d, k, kp, c, dist, grid);
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
    alldone = block_all(st != 0);
  }
  if (blockIdx.x == 0) {
    if (cvalid && t < kp) {
      p.conv[c] = st;
      p.rho1[c] = rho1;
      p.resid[c] = res;
      p.iters[c] = its;
    }
    if (t == 0) {
      p.status[0] = it;
      p.status[1] = alldone ? 1 : 0;
    }
  }
}

template <typename T, int LPR, bool MULTI>
static hipError_t launch_one(const CocgParams<T> &p, int grid, hipStream_t s) {
  void *args[] = {(void *)&p};
  return hipLaunchCooperativeKernel((const void *)k_cocg<T, LPR, MULTI>, dim3(grid), dim3(256), args, 0, s);
}

template <typename T>
static hipError_t launch_t(const CocgParams<T> &p, int lpr, int grid, hipStream_t s) {
  if (p.k > 1) return launch_one<T, 1, true>(p, grid, s);
  switch (lpr) {
    case 1: return launch_one<T, 1, false>(p, grid, s);
    case 2: return launch_one<T, 2, false>(p, grid, s);
    case 4: return launch_one<T, 4, false>(p, grid, s);
    case 8: return launch_one<T, 8, false>(p, grid, s);
    case 16: return launch_one<T, 16, false>(p, grid, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_cocg64(const CocgParams<double> &p, int lpr, int grid, hipStream_t s) { return launch_t(p, lpr, grid, s); }
hipError_t launch_cocg32(const CocgParams<float> &p, int lpr, int grid, hipStream_t s) { return launch_t(p, lpr, grid, s); }

template <typename T, int LPR, bool MULTI>
static int occ_one() {
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, (const void *)k_cocg<T, LPR, MULTI>, 256, 0) != hipSuccess)
    return 0;
  return nb;
}

// Largest co-resident grid (blocks) for the cooperative launch.
int max_coop_grid(int lpr, bool multi, bool fp64) {
  int dev = 0, ncu = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  int nb = 0;
  if (fp64) {
    if (multi) nb = occ_one<double, 1, true>();
    else nb = (lpr == 1) ? occ_one<double, 1, false>() : (lpr == 2) ? occ_one<double, 2, false>()
            : (lpr == 4) ? occ_one<double, 4, false>() : (lpr == 8) ? occ_one<double, 8, false>()
                                                       : occ_one<double, 16, false>();
  } else {
    if (multi) nb = occ_one<float, 1, true>();
    else nb = (lpr == 1) ? occ_one<float, 1, false>() : (lpr == 2) ? occ_one<float, 2, false>()
            : (lpr == 4) ? occ_one<float, 4, false>() : (lpr == 8) ? occ_one<float, 8, false>()
                                                      : occ_one<float, 16, false>();
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
      p.Y[ix] = y;
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
  if (kind == AFX_PC_SSOR || kind == AFX_PC_DIAG) {
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

__global__ void k_gather(int m, const int *map, const cx<double> *canon, cx<double> *out, cx<float> *out32) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= m) return;
  const cx<double> v = canon[map[i]];
  if (out) out[i] = v;
  if (out32) out32[i] = cx<float>{(float)v.r, (float)v.i};
}

hipError_t launch_gather_vals(int m, const int *map, const cx<double> *canon, cx<double> *out, cx<float> *out32,
                              hipStream_t s) {
  if (m <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_gather, dim3((m + 255) / 256), dim3(256), 0, s, m, map, canon, out, out32);
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
