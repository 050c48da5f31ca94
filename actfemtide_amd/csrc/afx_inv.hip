// Gauss-Newton data-space model update on the device: the real Jacobian rows of genjacobian1's
// amplitude / phase derivatives (src/src_inv_ap/m_jacobian_ap.f90:338-345), the normal matrix
// C = J Cm J^T + alpha Cd and the update of getnewmodel (:623-787) with solvebeta's C beta = X.
//
// The J Cm J^T product is a real dense contraction (ndat x nmodel by nmodel x ndat) and runs on
// FP64 MFMA (v_mfma_f64_16x16x4f64); C is factorised on the device by LU with partial pivoting
// (solvebeta's PARDISO factorisation is the general real one: Cm of the minimum-support-gradient
// stabiliser is not symmetric) and the two triangular solves run in one block.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "afx_internal.h"

using afx::cx;

namespace {

typedef double double4v __attribute__((ext_vector_type(4)));

// C[m][n] = sum_k A[m][k] B[n][k] (+ C when accumulate), FP64. One wave computes a 32x32 tile as
// 2x2 MFMA 16x16 tiles; 4 waves per block cover 64x64. Operand layout of the 16x16x4 MFMA: lane l
// supplies A[row l%16][k l/16] and B[k l/16][col l%16]; it holds D[l/16 + 4*i][l%16], i = 0..3 (the
// FP64 variant interleaves rows, unlike the 4*(l/16)+i of the FP32 16x16 forms).
__global__ void __launch_bounds__(256) k_dgemm_nt(const double *__restrict__ A, int lda, const double *__restrict__ B,
                                                  int ldb, int M, int N, int K, double *__restrict__ C, int ldc,
                                                  int accumulate) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int m0 = blockIdx.x * 64 + (wave >> 1) * 32;
  const int n0 = blockIdx.y * 64 + (wave & 1) * 32;
  const int r = lane & 15, kk = lane >> 4;
  double4v acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) acc[i][j] = double4v{0.0, 0.0, 0.0, 0.0};
  const int ma0 = m0 + r, ma1 = m0 + 16 + r, nb0 = n0 + r, nb1 = n0 + 16 + r;
  const double *pa0 = A + (size_t)min(ma0, M - 1) * lda, *pa1 = A + (size_t)min(ma1, M - 1) * lda;
  const double *pb0 = B + (size_t)min(nb0, N - 1) * ldb, *pb1 = B + (size_t)min(nb1, N - 1) * ldb;
  for (int k = 0; k < K; k += 4) {
    const int kc = k + kk;
    const bool kin = kc < K;
    const double a0 = (kin && ma0 < M) ? pa0[kc] : 0.0, a1 = (kin && ma1 < M) ? pa1[kc] : 0.0;
    const double b0 = (kin && nb0 < N) ? pb0[kc] : 0.0, b1 = (kin && nb1 < N) ? pb1[kc] : 0.0;
    acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
    acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
    acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
    acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
  }
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int m = m0 + 16 * i + kk + 4 * q, n = n0 + 16 * j + r;
        if (m < M && n < N) {
          double *c = C + (size_t)m * ldc + n;
          *c = accumulate ? *c + acc[i][j][q] : acc[i][j][q];
        }
      }
}

// G[d][i] = sum_j J[d][j] Cm[i][j]: G = (Cm J^T)^T, the transpose of getnewmodel's CMJT (Cm CSR)
__global__ void k_spmm_rows(int ndat, int nmodel, const double *J, const int *ptr, const int *col, const double *val,
                            double *G) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)ndat * nmodel) return;
  const int d = (int)(t / nmodel), i = (int)(t % nmodel);
  const double *jr = J + (size_t)d * nmodel;
  double s = 0.0;
  for (int q = ptr[i]; q < ptr[i + 1]; ++q) s += val[q] * jr[col[q]];
  G[t] = s;
}

__global__ void k_add_diag(int n, double *C, int ldc, const double *d, double alpha) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) C[(size_t)i * ldc + i] += alpha * d[i];
}

// y[d] = sum_i A[d][i] x[i] (one block per row)
__global__ void k_gemv_rows(int nrow, int ncol, const double *A, const double *x, double *y) {
  __shared__ double red[256];
  const int d = blockIdx.x;
  double s = 0.0;
  for (int i = threadIdx.x; i < ncol; i += 256) s += A[(size_t)d * ncol + i] * x[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) y[d] = red[0];
}

// y[i] = sum_d A[d][i] x[d] (threads over i, coalesced rows)
__global__ void k_gemv_cols(int nrow, int ncol, const double *A, const double *x, double *y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ncol) return;
  double s = 0.0;
  for (int d = 0; d < nrow; ++d) s += A[(size_t)d * ncol + i] * x[d];
  y[i] = s;
}

// solvebeta (src/src_inv/solvebeta.f90) factors C with PARDISO as a general real matrix
// (mtype 11): here LU with partial pivoting of the row-major n x n A, one column per step.
// Step 1 (one block): the pivot row p = argmax_{i >= j} |A[i][j]| (the lowest such i), rows j
// and p swapped, piv[j] = p; a zero pivot sets *bad.
__global__ void __launch_bounds__(1024) k_lu_pivot(int n, int j, double *A, int *piv, int *bad) {
  __shared__ double bv[1024];
  __shared__ int bi[1024];
  const int t = threadIdx.x;
  double best = -1.0;
  int bidx = j;
  for (int i = j + t; i < n; i += 1024) {
    const double v = fabs(A[(size_t)i * n + j]);
    if (v > best) {
      best = v;
      bidx = i;
    }
  }
  bv[t] = best;
  bi[t] = bidx;
  __syncthreads();
  for (int w = 512; w > 0; w >>= 1) {
    if (t < w && (bv[t + w] > bv[t] || (bv[t + w] == bv[t] && bi[t + w] < bi[t]))) {
      bv[t] = bv[t + w];
      bi[t] = bi[t + w];
    }
    __syncthreads();
  }
  const int p = bi[0];
  if (t == 0) {
    piv[j] = p;
    if (!(bv[0] > 0.0)) *bad = j + 1;
  }
  if (p != j)
    for (int k = t; k < n; k += 1024) {
      const double a = A[(size_t)j * n + k];
      A[(size_t)j * n + k] = A[(size_t)p * n + k];
      A[(size_t)p * n + k] = a;
    }
}
// step 2: the multipliers A[i][j] /= A[j][j] (i > j)
__global__ void k_lu_scale(int n, int j, double *A) {
  const int i = j + 1 + blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) A[(size_t)i * n + j] /= A[(size_t)j * n + j];
}
// step 3: A[i][k] -= A[i][j] A[j][k] (i, k > j; rows over blockIdx.y, columns over threads)
__global__ void k_lu_update(int n, int j, double *A) {
  const int i = j + 1 + blockIdx.y;
  const int k = j + 1 + blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && k < n) A[(size_t)i * n + k] -= A[(size_t)i * n + j] * A[(size_t)j * n + k];
}

// A x = b from the factors: y = P b, L y' = y (unit lower), U x = y'; one block of 1024 threads
__global__ void __launch_bounds__(1024) k_lu_solve(int n, const double *LU, const int *piv, const double *b, double *x,
                                                   double *y) {
  __shared__ double red[1024];
  const int t = threadIdx.x;
  if (t == 0) {
    for (int i = 0; i < n; i++) y[i] = b[i];
    for (int j = 0; j < n; j++) {
      const double a = y[j];
      y[j] = y[piv[j]];
      y[piv[j]] = a;
    }
  }
  __syncthreads();
  for (int i = 1; i < n; ++i) {
    double s = 0.0;
    for (int k = t; k < i; k += 1024) s += LU[(size_t)i * n + k] * y[k];
    red[t] = s;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
      if (t < w) red[t] += red[t + w];
      __syncthreads();
    }
    if (t == 0) y[i] -= red[0];
    __syncthreads();
  }
  for (int i = n - 1; i >= 0; --i) {
    double s = 0.0;
    for (int k = i + 1 + t; k < n; k += 1024) s += LU[(size_t)i * n + k] * x[k];
    red[t] = s;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
      if (t < w) red[t] += red[t + w];
      __syncthreads();
    }
    if (t == 0) x[i] = (y[i] - red[0]) / LU[(size_t)i * n + i];
    __syncthreads();
  }
}

// genjacobian1 :338-345: rows o of dJ (npar, nobs complex, o fastest) divided by be[o]:
// amp row o: Re(dbe/be)/ln10, phase row o: Im(dbe/be)*180/pi -> out [2][nobs][npar]
__global__ void k_ampphase(int npar, int nobs, const cx<double> *J, const cx<double> *be, double *out) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)npar * nobs) return;
  const int m = (int)(t / nobs), o = (int)(t % nobs);
  const cx<double> a = J[t], b = be[o];
  const double den = b.r * b.r + b.i * b.i;
  const double qr = (a.r * b.r + a.i * b.i) / den, qi = (a.i * b.r - a.r * b.i) / den;
  out[(size_t)o * npar + m] = qr / log(10.0);
  out[((size_t)nobs + o) * npar + m] = qi * 180.0 / M_PI;
}

thread_local std::string g_inv_err;

int fail(int code, const char *m) {
  g_inv_err = m;
  return code;
}

template <typename T>
int dalloc(T **p, size_t n) {
  if (hipMalloc((void **)p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) return fail(AFX_ERR_HIP, "hipMalloc");
  return AFX_OK;
}
template <typename T>
int upload(T **p, const T *h, size_t n) {
  int rc = dalloc(p, n);
  if (rc) return rc;
  if (n && hipMemcpy(*p, h, n * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) return fail(AFX_ERR_HIP, "copy");
  return AFX_OK;
}
template <typename T>
void dfree(T *&p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

}  // namespace

extern "C" {

const char *afx_inv_last_error(void) { return g_inv_err.c_str(); }

int afx_dgemm_nt_device(int M, int N, int K, const double *dA, int lda, const double *dB, int ldb, double *dC, int ldc,
                        int accumulate) {
  if (M <= 0 || N <= 0 || K <= 0 || !dA || !dB || !dC || lda < K || ldb < K || ldc < N)
    return fail(AFX_ERR_ARG, "dgemm: bad arguments");
  hipLaunchKernelGGL(k_dgemm_nt, dim3((M + 63) / 64, (N + 63) / 64), dim3(256), 0, 0, dA, lda, dB, ldb, M, N, K, dC,
                     ldc, accumulate);
  if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) return fail(AFX_ERR_HIP, "dgemm");
  return AFX_OK;
}

int afx_jacobian_ampphase_device(int npar, int nobs, const void *dJ, const void *dbe, double *dout) {
  if (npar <= 0 || nobs <= 0 || !dJ || !dbe || !dout) return fail(AFX_ERR_ARG, "ampphase: bad arguments");
  const long long m = (long long)npar * nobs;
  hipLaunchKernelGGL(k_ampphase, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, 0, npar, nobs,
                     (const cx<double> *)dJ, (const cx<double> *)dbe, dout);
  if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) return fail(AFX_ERR_HIP, "ampphase");
  return AFX_OK;
}

int afx_gn_update_device(int ndat, int nmodel, const double *dJ, const int *cm_ptr, const int *cm_col,
                         const double *cm_val, const double *cd, double alpha, const double *dobs, const double *dcal,
                         const double *model, const double *model_ref, double *model_new, double *beta_out) {
  if (ndat <= 0 || nmodel <= 0 || !dJ || !cm_ptr || !cm_col || !cm_val || !cd || !dobs || !dcal || !model ||
      !model_ref || !model_new)
    return fail(AFX_ERR_ARG, "gn_update: bad arguments");
  if (ndat > 65535) return fail(AFX_ERR_ARG, "gn_update: ndat above 65535 (the LU update grid's row extent)");
  const int nnz = cm_ptr[nmodel];
  for (int q = 0; q < nnz; q++)
    if (cm_col[q] < 0 || cm_col[q] >= nmodel) return fail(AFX_ERR_ARG, "gn_update: Cm column out of range");
  int *dptr = nullptr, *dcol = nullptr, *dbad = nullptr, *dpiv = nullptr;
  double *dval = nullptr, *dG = nullptr, *dC = nullptr, *dcd = nullptr, *ddm = nullptr, *dJM = nullptr;
  double *dX = nullptr, *dy = nullptr, *dbeta = nullptr, *dmod = nullptr;
  int rc;
  auto cleanup = [&]() {
    dfree(dptr); dfree(dcol); dfree(dbad); dfree(dpiv); dfree(dval); dfree(dG); dfree(dC); dfree(dcd); dfree(ddm);
    dfree(dJM); dfree(dX); dfree(dy); dfree(dbeta); dfree(dmod);
  };
  std::vector<double> dm(nmodel);
  for (int i = 0; i < nmodel; i++) dm[i] = model[i] - model_ref[i];  // getnewmodel :712
  if ((rc = upload(&dptr, cm_ptr, (size_t)nmodel + 1)) || (rc = upload(&dcol, cm_col, (size_t)nnz)) ||
      (rc = upload(&dval, cm_val, (size_t)nnz)) || (rc = dalloc(&dG, (size_t)ndat * nmodel)) ||
      (rc = dalloc(&dC, (size_t)ndat * ndat)) || (rc = upload(&dcd, cd, (size_t)ndat)) ||
      (rc = upload(&ddm, dm.data(), (size_t)nmodel)) || (rc = dalloc(&dJM, (size_t)ndat)) ||
      (rc = dalloc(&dX, (size_t)ndat)) || (rc = dalloc(&dy, (size_t)ndat)) || (rc = dalloc(&dbeta, (size_t)ndat)) ||
      (rc = dalloc(&dmod, (size_t)nmodel)) || (rc = dalloc(&dbad, 1)) || (rc = dalloc(&dpiv, (size_t)ndat))) {
    cleanup();
    return rc;
  }
  bool ok = hipMemset(dbad, 0, sizeof(int)) == hipSuccess;
  // G = (Cm J^T)^T (getnewmodel's CMJT, :699-701); C = J CMJT + alpha Cd = J G^T + alpha Cd (:704-711)
  const long long ng = (long long)ndat * nmodel;
  hipLaunchKernelGGL(k_spmm_rows, dim3((unsigned)((ng + 255) / 256)), dim3(256), 0, 0, ndat, nmodel, dJ, dptr, dcol,
                     dval, dG);
  hipLaunchKernelGGL(k_dgemm_nt, dim3((ndat + 63) / 64, (ndat + 63) / 64), dim3(256), 0, 0, dJ, nmodel, dG, nmodel,
                     ndat, ndat, nmodel, dC, ndat, 0);
  hipLaunchKernelGGL(k_add_diag, dim3((ndat + 255) / 256), dim3(256), 0, 0, ndat, dC, ndat, dcd, alpha);
  // X = dobs - dcal + J (m - m_ref) (:713-715)
  hipLaunchKernelGGL(k_gemv_rows, dim3(ndat), dim3(256), 0, 0, ndat, nmodel, dJ, ddm, dJM);
  std::vector<double> JM(ndat), X(ndat);
  ok = ok && hipGetLastError() == hipSuccess && hipMemcpy(JM.data(), dJM, ndat * 8, hipMemcpyDeviceToHost) == hipSuccess;
  for (int d = 0; d < ndat; d++) X[d] = dobs[d] - dcal[d] + JM[d];
  ok = ok && hipMemcpy(dX, X.data(), ndat * 8, hipMemcpyHostToDevice) == hipSuccess;
  // C beta = X (solvebeta): LU with partial pivoting in place, then the two triangular solves
  for (int j = 0; ok && j < ndat; j++) {
    hipLaunchKernelGGL(k_lu_pivot, dim3(1), dim3(1024), 0, 0, ndat, j, dC, dpiv, dbad);
    const int rest = ndat - j - 1;
    if (rest > 0) {
      hipLaunchKernelGGL(k_lu_scale, dim3((rest + 255) / 256), dim3(256), 0, 0, ndat, j, dC);
      hipLaunchKernelGGL(k_lu_update, dim3((rest + 255) / 256, rest), dim3(256), 0, 0, ndat, j, dC);
    }
  }
  hipLaunchKernelGGL(k_lu_solve, dim3(1), dim3(1024), 0, 0, ndat, dC, dpiv, dX, dbeta, dy);
  // m_new = m_ref + Cm J^T beta = m_ref + G^T beta (:724-733)
  hipLaunchKernelGGL(k_gemv_cols, dim3((nmodel + 255) / 256), dim3(256), 0, 0, ndat, nmodel, dG, dbeta, dmod);
  int bad = 0;
  ok = ok && hipGetLastError() == hipSuccess && hipDeviceSynchronize() == hipSuccess &&
       hipMemcpy(&bad, dbad, sizeof(int), hipMemcpyDeviceToHost) == hipSuccess &&
       hipMemcpy(model_new, dmod, (size_t)nmodel * 8, hipMemcpyDeviceToHost) == hipSuccess &&
       (!beta_out || hipMemcpy(beta_out, dbeta, (size_t)ndat * 8, hipMemcpyDeviceToHost) == hipSuccess);
  cleanup();
  if (!ok) return fail(AFX_ERR_HIP, "gn_update: HIP error");
  if (bad) return fail(AFX_ERR_BREAKDOWN, "gn_update: C is singular (zero pivot)");
  for (int i = 0; i < nmodel; i++) model_new[i] += model_ref[i];
  return AFX_OK;
}

// Extreme eigenvalues of the symmetric tridiagonal matrix (diagonal a[0..m), off-diagonal
// b[0..m-1)) by bisection on Sturm sequence counts (the number of eigenvalues below x is the
// number of negative pivots of T - x I), from the Gershgorin interval down to adjacent doubles.
static int sturm_below(const std::vector<double> &a, const std::vector<double> &b, double x) {
  int cnt = 0;
  double d = 1.0;
  for (size_t i = 0; i < a.size(); i++) {
    const double bb = i ? b[i - 1] * b[i - 1] : 0.0;
    d = a[i] - x - (i ? bb / d : 0.0);
    if (d == 0.0) d = -1e-300;  // a zero pivot counts as negative (x is an eigenvalue: move below it)
    if (d < 0.0) cnt++;
  }
  return cnt;
}
static double tridiag_eig(const std::vector<double> &a, const std::vector<double> &b, int k) {
  // k-th smallest eigenvalue (0-based)
  const size_t m = a.size();
  double lo = a[0], hi = a[0];
  for (size_t i = 0; i < m; i++) {
    const double r = (i ? std::fabs(b[i - 1]) : 0.0) + (i + 1 < m ? std::fabs(b[i]) : 0.0);
    lo = std::min(lo, a[i] - r);
    hi = std::max(hi, a[i] + r);
  }
  for (int it = 0; it < 2200; it++) {
    const double mid = 0.5 * (lo + hi);
    if (!(mid > lo && mid < hi)) break;
    if (sturm_below(a, b, mid) > k) hi = mid;
    else lo = mid;
  }
  return 0.5 * (lo + hi);
}

int afx_spectral_alpha_device(int ndat, int nmodel, const double *dJ, const int *cm_ptr, const int *cm_col,
                              const double *cm_val, const double *cd, int kmax, double *alpha, double *emin,
                              double *emax) {
  if (ndat <= 0 || nmodel <= 0 || kmax < 1 || !dJ || !cm_ptr || !cm_col || !cm_val || !cd || !alpha)
    return fail(AFX_ERR_ARG, "spectral_alpha: bad arguments");
  const int nnz = cm_ptr[nmodel];
  for (int q = 0; q < nnz; q++)
    if (cm_col[q] < 0 || cm_col[q] >= nmodel) return fail(AFX_ERR_ARG, "spectral_alpha: Cm column out of range");
  int *dptr = nullptr, *dcol = nullptr;
  double *dval = nullptr, *dG = nullptr, *dC = nullptr, *dq = nullptr, *dr = nullptr;
  auto cleanup = [&]() { dfree(dptr); dfree(dcol); dfree(dval); dfree(dG); dfree(dC); dfree(dq); dfree(dr); };
  int rc;
  if ((rc = upload(&dptr, cm_ptr, (size_t)nmodel + 1)) || (rc = upload(&dcol, cm_col, (size_t)nnz)) ||
      (rc = upload(&dval, cm_val, (size_t)nnz)) || (rc = dalloc(&dG, (size_t)ndat * nmodel)) ||
      (rc = dalloc(&dC, (size_t)ndat * ndat)) || (rc = dalloc(&dq, (size_t)ndat)) || (rc = dalloc(&dr, (size_t)ndat))) {
    cleanup();
    return rc;
  }
  // JCMJT = J (Cm J^T) (m_spectral.f90:77-80), as G = (Cm J^T)^T and C = J G^T on FP64 MFMA
  const long long ng = (long long)ndat * nmodel;
  hipLaunchKernelGGL(k_spmm_rows, dim3((unsigned)((ng + 255) / 256)), dim3(256), 0, 0, ndat, nmodel, dJ, dptr, dcol,
                     dval, dG);
  hipLaunchKernelGGL(k_dgemm_nt, dim3((ndat + 63) / 64, (ndat + 63) / 64), dim3(256), 0, 0, dJ, nmodel, dG, nmodel,
                     ndat, ndat, nmodel, dC, ndat, 0);
  bool ok = hipGetLastError() == hipSuccess;
  // LANCZOSMINMAX (:102-179): q1 = ones / |ones|; res = C q; alpha_j = q.res; res -= beta_{j-1} v + alpha_j q;
  // beta_j = |res|; stop early when beta_j = 0 (the Krylov space is invariant)
  const size_t N = (size_t)ndat;
  std::vector<double> q(N, 1.0), v(N), res(N), al, be;
  {
    double qq = 0.0;
    for (size_t i = 0; i < N; i++) qq += q[i] * q[i];
    const double sq = std::sqrt(qq);
    for (size_t i = 0; i < N; i++) q[i] = q[i] / sq;
  }
  for (int j = 0; ok && j < kmax; j++) {
    if (j > 0) {
      v = q;
      for (size_t i = 0; i < N; i++) q[i] = res[i] / be[j - 1];
    }
    ok = hipMemcpy(dq, q.data(), N * 8, hipMemcpyHostToDevice) == hipSuccess;
    hipLaunchKernelGGL(k_gemv_rows, dim3(ndat), dim3(256), 0, 0, ndat, ndat, dC, dq, dr);
    ok = ok && hipGetLastError() == hipSuccess && hipMemcpy(res.data(), dr, N * 8, hipMemcpyDeviceToHost) == hipSuccess;
    if (!ok) break;
    if (j > 0)
      for (size_t i = 0; i < N; i++) res[i] = res[i] - be[j - 1] * v[i];
    double a = 0.0;
    for (size_t i = 0; i < N; i++) a += q[i] * res[i];
    for (size_t i = 0; i < N; i++) res[i] = res[i] - a * q[i];
    double rr = 0.0;
    for (size_t i = 0; i < N; i++) rr += res[i] * res[i];
    al.push_back(a);
    be.push_back(std::sqrt(rr));
    if (be.back() == 0.0) break;
  }
  cleanup();
  if (!ok) return fail(AFX_ERR_HIP, "spectral_alpha: HIP error");
  // eigenvalues of the tridiagonal (alpha_1..alpha_m; beta_1..beta_{m-1}) (:153-170)
  std::vector<double> off(be.begin(), be.end() - 1);
  const double lo = tridiag_eig(al, off, 0), hi = tridiag_eig(al, off, (int)al.size() - 1);
  const double numer = std::max(std::fabs(hi), std::fabs(lo));
  double denom = 0.0;  // (:84-88)
  for (int i = 0; i < ndat; i++) denom = std::max(denom, cd[i]);
  if (!(denom > 0.0)) return fail(AFX_ERR_ARG, "spectral_alpha: Cd has no positive entry");
  *alpha = std::fabs(numer / denom);
  if (emin) *emin = lo;
  if (emax) *emax = hi;
  return AFX_OK;
}

int afx_model_roughening(int ntet, const int *n4, int nphys1, const int *ele2model, int nmodel, const int *iactive,
                         int *nrow, long long *nnz, int *ptr, int *col, double *val) {
  if (ntet <= 0 || !n4 || nphys1 < 0 || nphys1 >= ntet || !ele2model || nmodel <= 0 || !nrow || !nnz)
    return fail(AFX_ERR_ARG, "model_roughening: bad arguments");
  const int nphys2 = ntet - nphys1;
  for (int i = 0; i < nphys2; i++)
    if (ele2model[i] < 1 || ele2model[i] > nmodel) return fail(AFX_ERR_ARG, "model_roughening: ele2model out of range");
  // faces of the land elements as sorted node triples; two land elements listing the same face
  // are neighbours (GENDQ :652-675 through face2ele)
  struct Face {
    int a, b, c, e;
  };
  static const int FN[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};
  std::vector<Face> faces((size_t)nphys2 * 4);
  for (int i = 0; i < nphys2; i++) {
    const int *t = &n4[(size_t)(nphys1 + i) * 4];
    for (int f = 0; f < 4; f++) {
      int v[3] = {t[FN[f][0]], t[FN[f][1]], t[FN[f][2]]};
      std::sort(v, v + 3);
      faces[(size_t)i * 4 + f] = {v[0], v[1], v[2], i};
    }
  }
  std::sort(faces.begin(), faces.end(), [](const Face &x, const Face &y) {
    return x.a != y.a ? x.a < y.a : x.b != y.b ? x.b < y.b : x.c != y.c ? x.c < y.c : x.e < y.e;
  });
  std::vector<std::pair<int, int>> adj;  // (model, neighbour model), both directions, 0-based
  for (size_t q = 0; q + 1 < faces.size(); q++) {
    const Face &x = faces[q], &y = faces[q + 1];
    if (x.a == y.a && x.b == y.b && x.c == y.c) {
      const int m1 = ele2model[x.e] - 1, m2 = ele2model[y.e] - 1;
      if (m1 != m2) {
        adj.push_back({m1, m2});
        adj.push_back({m2, m1});
      }
    }
  }
  std::sort(adj.begin(), adj.end());
  adj.erase(std::unique(adj.begin(), adj.end()), adj.end());
  // R = Q^T D Q with every value replaced: -1 off the diagonal, (row length - 1) on it (:770-790);
  // rows and columns of inactive models dropped, the values kept (:795-830)
  std::vector<int> newid(nmodel, -1);
  int na = 0;
  for (int m = 0; m < nmodel; m++)
    if (!iactive || iactive[m]) newid[m] = na++;
  std::vector<int> deg(nmodel, 0);
  for (const auto &pr : adj) deg[pr.first]++;
  for (int m = 0; m < nmodel; m++)
    if (deg[m] == 0 && newid[m] >= 0) return fail(AFX_ERR_ARG, "model_roughening: a model has no neighbour model");
  long long cnt = 0;
  for (int m = 0; m < nmodel; m++)
    if (newid[m] >= 0) cnt++;  // diagonal
  for (const auto &pr : adj)
    if (newid[pr.first] >= 0 && newid[pr.second] >= 0) cnt++;
  *nrow = na;
  *nnz = cnt;
  if (!ptr) return AFX_OK;
  if (!col || !val) return fail(AFX_ERR_ARG, "model_roughening: null output");
  size_t a = 0;
  long long pos = 0;
  for (int m = 0; m < nmodel; m++) {
    while (a < adj.size() && adj[a].first < m) a++;
    if (newid[m] < 0) continue;
    ptr[newid[m]] = (int)pos;
    bool diag = false;
    for (size_t b = a; b <= adj.size(); b++) {
      const bool end = b == adj.size() || adj[b].first != m;
      const int nb = end ? nmodel : adj[b].second;
      if (!diag && nb > m) {  // columns ascending, the diagonal in place
        col[pos] = newid[m];
        val[pos++] = (double)deg[m];
        diag = true;
      }
      if (end) break;
      if (newid[nb] >= 0) {
        col[pos] = newid[nb];
        val[pos++] = -1.0;
      }
    }
  }
  ptr[na] = (int)pos;
  return AFX_OK;
}

}  // extern "C"
