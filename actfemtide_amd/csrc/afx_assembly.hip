// Element assembly, boundary conditions, observation operator and adjoint Jacobian kernels
// for gfx950 (MI355X).
//
// Assembly (GENMAT, src/solver_mpi/forward_bxyz_mpi.f90:99-311) is split in two launches:
//   k_element    -- one thread per tetrahedron: geometry, curl-curl entries S, mass entries
//                   Y*mu0*sigma*L0^2 and the tidal v x F right-hand side, written to
//                   entry-major arrays [entry][tet] so that neighbouring threads store
//                   neighbouring words;
//   k_slot_gather -- one thread per matrix slot: sums the contributions of the tets that
//                   touch the slot in ascending tet order (the order sup_iccg adds them,
//                   m_iccg_var_takuto.f90:69-98), through a slot -> (entry, tet) map built
//                   once per mesh. Deterministic, no atomics.
#include "afx_internal.h"

namespace afx {

namespace {
__constant__ int KL[6][2] = {{0, 1}, {1, 2}, {0, 2}, {0, 3}, {1, 3}, {2, 3}};
__constant__ int LMN[4][3] = {{1, 2, 3}, {0, 3, 2}, {0, 1, 3}, {0, 2, 1}};

__device__ __forceinline__ double inner3(const double *a, const double *b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
__device__ __forceinline__ void outer3(const double *a, const double *b, double *c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}
__device__ __forceinline__ double intv(int k, int m, double v) { return (k != m) ? v / 20.0 : v / 10.0; }

// edge vectors (calxmn, m_fem_util.f90:42-53), gradients and volume (gradnodebasisfun :12-40)
__device__ bool geometry(const double x[4][3], double xx[6][3], double gn[4][3], double &v) {
  for (int d = 0; d < 3; d++) {
    xx[0][d] = x[3][d] - x[2][d];
    xx[1][d] = x[3][d] - x[0][d];
    xx[2][d] = x[1][d] - x[3][d];
    xx[3][d] = x[2][d] - x[1][d];
    xx[4][d] = x[0][d] - x[2][d];
    xx[5][d] = x[1][d] - x[0][d];
  }
  double c[3], m4[3] = {-xx[3][0], -xx[3][1], -xx[3][2]};
  outer3(xx[0], xx[4], c);
  v = 1.0 / 6.0 * inner3(c, m4);
  if (!(v > 0.0)) return false;
  const double s = 1.0 / 6.0 / v;
  double a[3], b[3];
  for (int d = 0; d < 3; d++) { a[d] = x[3][d] - x[1][d]; b[d] = x[2][d] - x[1][d]; }
  outer3(a, b, c);
  for (int d = 0; d < 3; d++) gn[0][d] = s * c[d];
  for (int d = 0; d < 3; d++) { a[d] = x[2][d] - x[0][d]; b[d] = x[3][d] - x[0][d]; }
  outer3(a, b, c);
  for (int d = 0; d < 3; d++) gn[1][d] = s * c[d];
  for (int d = 0; d < 3; d++) { a[d] = x[3][d] - x[0][d]; b[d] = x[1][d] - x[0][d]; }
  outer3(a, b, c);
  for (int d = 0; d < 3; d++) gn[2][d] = s * c[d];
  for (int d = 0; d < 3; d++) { a[d] = x[1][d] - x[0][d]; b[d] = x[2][d] - x[0][d]; }
  outer3(a, b, c);
  for (int d = 0; d < 3; d++) gn[3][d] = s * c[d];
  return true;
}

// mass-term integral of GENMAT (forward_bxyz_mpi.f90:223-232) for local edges i, j
__device__ __forceinline__ double mass_yy(const double gn[4][3], double v, int i, int j) {
  const int k = KL[i][0], l = KL[i][1], m = KL[j][0], n = KL[j][1];
  return intv(k, m, v) * inner3(gn[l], gn[n]) - intv(k, n, v) * inner3(gn[l], gn[m]) -
         intv(l, m, v) * inner3(gn[k], gn[n]) + intv(l, n, v) * inner3(gn[k], gn[m]);
}

__device__ __forceinline__ void load_tet(const double *xyz, const int *n4, int t, double x[4][3]) {
  for (int j = 0; j < 4; j++) {
    const int nd = n4[(size_t)t * 4 + j] - 1;
    x[j][0] = xyz[(size_t)nd * 3];
    x[j][1] = xyz[(size_t)nd * 3 + 1];
    x[j][2] = xyz[(size_t)nd * 3 + 2];
  }
}
}  // namespace

__global__ void __launch_bounds__(256) k_element(ElemParams p) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= p.ntet) return;
  double x[4][3], xx[6][3], gn[4][3], v;
  load_tet(p.xyz, p.n4, t, x);
  if (!geometry(x, xx, gn, v)) {
    atomicOr(p.err, 1);
    return;
  }
  double dir[6];
  for (int j = 0; j < 6; j++) dir[j] = (p.n6line[(size_t)t * 6 + j] < 0) ? -1.0 : 1.0;
  const double AA = 1.0 / 9.0 / v;
  const double ms = (4.0 * 3.14159265358979323846 * 1.0e-7) * p.sigma[t] * 1.0e6;  // mu0*sigma*L0^2
  const size_t nt = (size_t)p.ntet;
  // entry e: a == b -> e = a; a > b -> e = 6 + a(a-1)/2 + b  (element-lower entry, sup_iccg :91-95)
  for (int a = 0; a < 6; a++) {
    for (int b = 0; b <= a; b++) {
      const int e = (a == b) ? a : 6 + a * (a - 1) / 2 + b;
      const double S = inner3(xx[a], xx[b]) * dir[a] * dir[b] * AA;
      const double Y = mass_yy(gn, v, a, b) * dir[a] * dir[b];
      p.eK[e * nt + t] = S;
      p.eM[e * nt + t] = Y * ms;
    }
  }
  // tidal source (forward_bxyz_mpi.f90:267-299): S2_i = mu0*sigma*sum_j [intv(k,j) gn_l - intv(l,j) gn_k].vF_j * d_i * 1e3
  cx<double> S2[6];
  for (int i = 0; i < 6; i++) S2[i] = {0.0, 0.0};
  if (p.group[t] == 2) {
    cx<double> vF[4][3];
    for (int j = 0; j < 4; j++) {
      const int nd = p.n4[(size_t)t * 4 + j] - 1;
      const cx<double> *vv = &p.vxyz[(size_t)nd * 3];
      const double *F = &p.fxyz[(size_t)nd * 3];
      vF[j][0] = {vv[1].r * F[2] - vv[2].r * F[1], vv[1].i * F[2] - vv[2].i * F[1]};
      vF[j][1] = {vv[2].r * F[0] - vv[0].r * F[2], vv[2].i * F[0] - vv[0].i * F[2]};
      vF[j][2] = {vv[0].r * F[1] - vv[1].r * F[0], vv[0].i * F[1] - vv[1].i * F[0]};
    }
    const double pre = (4.0 * 3.14159265358979323846 * 1.0e-7) * p.sigma[t];
    for (int i = 0; i < 6; i++) {
      const int k = KL[i][0], l = KL[i][1];
      cx<double> acc = {0.0, 0.0};
      for (int j = 0; j < 4; j++) {
        const cx<double> il = {gn[l][0] * vF[j][0].r + gn[l][1] * vF[j][1].r + gn[l][2] * vF[j][2].r,
                               gn[l][0] * vF[j][0].i + gn[l][1] * vF[j][1].i + gn[l][2] * vF[j][2].i};
        const cx<double> ik = {gn[k][0] * vF[j][0].r + gn[k][1] * vF[j][1].r + gn[k][2] * vF[j][2].r,
                               gn[k][0] * vF[j][0].i + gn[k][1] * vF[j][1].i + gn[k][2] * vF[j][2].i};
        const double a1 = intv(k, j, v), a2 = intv(l, j, v);
        acc.r += pre * (a1 * il.r - a2 * ik.r) * dir[i] * 1.0e3;
        acc.i += pre * (a1 * il.i - a2 * ik.i) * dir[i] * 1.0e3;
      }
      S2[i] = acc;
    }
  }
  for (int i = 0; i < 6; i++) p.eB[i * nt + t] = S2[i];
}

hipError_t launch_element(const ElemParams &p, hipStream_t s) {
  if (p.ntet <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_element, dim3((p.ntet + 255) / 256), dim3(256), 0, s, p);
  return hipGetLastError();
}

// out[s] = (sum K, sum M) over the slot's contributions in ascending tet order
__global__ void k_slot_gather(int nslot, const int *cptr, const int *cidx, const double *eK, const double *eM,
                              cx<double> *out) {
  const int s = blockIdx.x * 256 + threadIdx.x;
  if (s >= nslot) return;
  double k = 0.0, m = 0.0;
  for (int q = cptr[s]; q < cptr[s + 1]; ++q) {
    const int ix = cidx[q];
    k += eK[ix];
    m += eM[ix];
  }
  out[s] = cx<double>{k, m};
}

hipError_t launch_slot_gather(int nslot, const int *cptr, const int *cidx, const double *eK, const double *eM,
                              cx<double> *out, hipStream_t s) {
  if (nslot <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_slot_gather, dim3((nslot + 255) / 256), dim3(256), 0, s, nslot, cptr, cidx, eK, eM, out);
  return hipGetLastError();
}

__global__ void k_rhs_gather(int n, const int *cptr, const int *cidx, const cx<double> *eB, cx<double> *b) {
  const int s = blockIdx.x * 256 + threadIdx.x;
  if (s >= n) return;
  cx<double> acc = {0.0, 0.0};
  for (int q = cptr[s]; q < cptr[s + 1]; ++q) {
    const cx<double> e = eB[cidx[q]];
    acc.r += e.r;
    acc.i += e.i;
  }
  b[s] = acc;
}

hipError_t launch_rhs_gather(int n, const int *cptr, const int *cidx, const cx<double> *eB, double,
                             cx<double> *b, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_rhs_gather, dim3((n + 255) / 256), dim3(256), 0, s, n, cptr, cidx, eB, b);
  return hipGetLastError();
}

// COPY_UL_ICCG12 (m_iccg_var_takuto.f90:107-167): AL(pos) = AU(mirror(pos))
__global__ void k_mirror(long long nL, const int *mirror, const cx<double> *AU, cx<double> *AL) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i < nL) AL[i] = AU[mirror[i]];
}
hipError_t launch_mirror(long long nL, const int *mirror, const cx<double> *AU, cx<double> *AL, hipStream_t s) {
  if (nL <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_mirror, dim3((unsigned)((nL + 255) / 256)), dim3(256), 0, s, nL, mirror, AU, AL);
  return hipGetLastError();
}

// SET_BC_ICCG matrix part (forward_bxyz_mpi.f90:607-634) with zero Dirichlet values
__global__ void k_bc_values(int n, const int *INU, const int *IAU, const int *INL, const int *IAL,
                            const unsigned char *dir, cx<double> *D, cx<double> *AU, cx<double> *AL) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const cx<double> z = {0.0, 0.0};
  if (dir[i]) {
    D[i] = cx<double>{1.0, 0.0};
    for (int q = INL[i]; q < INL[i + 1]; ++q) AL[q] = z;
    for (int q = INU[i]; q < INU[i + 1]; ++q) AU[q] = z;
  } else {
    for (int q = INL[i]; q < INL[i + 1]; ++q)
      if (dir[IAL[q] - 1]) AL[q] = z;
    for (int q = INU[i]; q < INU[i + 1]; ++q)
      if (dir[IAU[q] - 1]) AU[q] = z;
  }
}
hipError_t launch_bc_values(int n, const int *INU, const int *IAU, const int *INL, const int *IAL,
                            const unsigned char *dir, cx<double> *D, cx<double> *AU, cx<double> *AL,
                            hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_bc_values, dim3((n + 255) / 256), dim3(256), 0, s, n, INU, IAU, INL, IAL, dir, D, AU, AL);
  return hipGetLastError();
}

// GENBCCSEM (forward_bxyz_mpi.f90:467-519)
__global__ void k_genbc(int nline, const int *line, const double *xyz, double zmin, double zmax, double xout,
                        double yout, unsigned char *bc) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= nline) return;
  const double xo = xout - 1.0, yo = yout - 1.0, z0 = zmin + 1.0, z1 = zmax - 1.0;
  const int a = line[2 * i] - 1, b = line[2 * i + 1] - 1;
  const double x1 = xyz[a * 3], y1 = xyz[a * 3 + 1], zz1 = xyz[a * 3 + 2];
  const double x2 = xyz[b * 3], y2 = xyz[b * 3 + 1], zz2 = xyz[b * 3 + 2];
  const bool g = (x1 <= -xo && x2 <= -xo) || (x1 >= xo && x2 >= xo) || (y1 <= -yo && y2 <= -yo) ||
                 (y1 >= yo && y2 >= yo) || (zz1 <= z0 && zz2 <= z0) || (zz1 >= z1 && zz2 >= z1);
  bc[i] = g ? 1 : 0;
}
hipError_t launch_genbc(int nline, const int *line, const double *xyz, double zmin, double zmax, double xout,
                        double yout, unsigned char *bc, hipStream_t s) {
  if (nline <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_genbc, dim3((nline + 255) / 256), dim3(256), 0, s, nline, line, xyz, zmin, zmax, xout, yout, bc);
  return hipGetLastError();
}

// nodebasisfun (m_fem_util.f90:56-83) at P for tet t and FINDELEMENT0's test (m_mesh_type.f90:715-719):
// all four barycentric coordinates >= 0. Evaluated without contraction, in the reference's
// operation order (helpers included), so the decision on a shared face, edge or node is the
// reference's.
__device__ __forceinline__ double inner3_nc(const double *a, const double *b) {
#pragma clang fp contract(off)
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
__device__ __forceinline__ void outer3_nc(const double *a, const double *b, double *c) {
#pragma clang fp contract(off)
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}
__device__ bool tet_contains(const double *xyz, const int *n4, int t, const double *P) {
#pragma clang fp contract(off)
  double x[4][3];
  load_tet(xyz, n4, t, x);
  double v0[3], v1[3], x12[3], x13[3], x14[3], x24[3], x23[3], c[3];
  for (int d = 0; d < 3; d++) {
    v0[d] = -(x[0][d] - P[d]);
    v1[d] = -(x[1][d] - P[d]);
    x12[d] = x[1][d] - x[0][d];
    x13[d] = x[2][d] - x[0][d];
    x14[d] = x[3][d] - x[0][d];
    x24[d] = x[3][d] - x[1][d];
    x23[d] = x[2][d] - x[1][d];
  }
  outer3_nc(x12, x13, c);
  const double a5 = 1.0 / 6.0 * inner3_nc(c, x14);
  const double a4 = 1.0 / 6.0 * inner3_nc(c, v0) / a5;
  outer3_nc(x14, x12, c);
  const double a3 = 1.0 / 6.0 * inner3_nc(c, v0) / a5;
  outer3_nc(x13, x14, c);
  const double a2 = 1.0 / 6.0 * inner3_nc(c, v0) / a5;
  outer3_nc(x24, x23, c);
  const double a1 = 1.0 / 6.0 * inner3_nc(c, v1) / a5;
  return a1 >= 0.0 && a2 >= 0.0 && a3 >= 0.0 && a4 >= 0.0;
}

// FINDELEMENT0 (m_mesh_type.f90:697-734) over a bucket grid (the reference's own search,
// findelewithgrid :204-277, lists tets by the bucket of their centroid). One wave per point: it
// tests every tet whose centroid bucket lies within the largest centroid-to-vertex extent of
// the point (plus one bucket for rounding), so every tet containing the point is tested, and
// keeps the lowest containing id -- the tet FINDELEMENT0's scan in id order returns. Tets of a
// run of buckets along x are contiguous in `btet`, so the lanes stride over each run.
__global__ void __launch_bounds__(256) k_findelement_grid(int nobs, const double *pts, PointGrid g,
                                                          const int *bptr, const int *btet, const double *xyz,
                                                          const int *n4, int *first) {
  const int o = (int)(blockIdx.x * 4 + (threadIdx.x >> 6));
  const int lane = threadIdx.x & 63;
  if (o >= nobs) return;  // wave-uniform
  const double P[3] = {pts[o * 3], pts[o * 3 + 1], pts[o * 3 + 2]};
  int lo[3], hi[3];
  for (int d = 0; d < 3; d++) {
    const double a = floor((P[d] - g.R[d] - g.lo[d]) / g.h[d]) - 1.0;
    const double b = floor((P[d] + g.R[d] - g.lo[d]) / g.h[d]) + 1.0;
    lo[d] = (int)fmax(0.0, fmin(a, (double)(g.n[d] - 1)));
    hi[d] = (int)fmax(0.0, fmin(b, (double)(g.n[d] - 1)));
  }
  int best = 0x7fffffff;
  for (int iz = lo[2]; iz <= hi[2]; iz++)
    for (int iy = lo[1]; iy <= hi[1]; iy++) {
      const int b0 = (iz * g.n[1] + iy) * g.n[0];
      for (int q = bptr[b0 + lo[0]] + lane; q < bptr[b0 + hi[0] + 1]; q += 64) {
        const int t = btet[q];
        if (t < best && tet_contains(xyz, n4, t, P)) best = t;
      }
    }
  for (int m = 32; m >= 1; m >>= 1) best = min(best, __shfl_xor(best, m, 64));
  if (lane == 0) first[o] = best;
}
hipError_t launch_findelement_grid(int nobs, const double *pts, const PointGrid &g, const int *bptr, const int *btet,
                                   const double *xyz, const int *n4, int *first, hipStream_t s) {
  if (nobs <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_findelement_grid, dim3((unsigned)((nobs + 3) / 4)), dim3(256), 0, s, nobs, pts, g, bptr, btet,
                     xyz, n4, first);
  return hipGetLastError();
}

// mul_matcrs_cv over 6-entry rows: out[o][c] = sum_j val[o][j] * X[item[o][j]-1][c]
__global__ void k_obs_apply(int nobs, const int *item, const double *val, int ncol, const cx<double> *X,
                            cx<double> *out) {
  const long long gid = (long long)blockIdx.x * 256 + threadIdx.x;
  if (gid >= (long long)nobs * ncol) return;
  const int o = (int)(gid / ncol), c = (int)(gid % ncol);
  cx<double> s = {0.0, 0.0};
  for (int j = 0; j < 6; j++) {
    const double a = val[o * 6 + j];
    const cx<double> xv = X[(size_t)(item[o * 6 + j] - 1) * ncol + c];
    s.r += a * xv.r;
    s.i += a * xv.i;
  }
  out[gid] = s;
}
// IXYHOUT (src/solver_mpi/n_ebfem_bxyz_mpi.f90:192-301): the seafloor current at observation
// point o. h = -z (0 on land, :227-231); vF = v x F at the nodes of row o of coeff_vF (cr_outer,
// src/common/m_outerinnerproduct.f90:45-53) interpolated by that row (mul_matcrs_cv order,
// :244-251) as is v (x, y); then, in the reference's order of operations (:261-268),
//   ixyh_e = sigma h E,  ixyh_vF = sigma h vF 1e-3,  ixyh = sigma h (E + vF 1e-3),  vh = h v.
// out[o] = {ixyh_e (x, y), ixyh_vF (x, y), ixyh (x, y), vh (x, y)}; contraction off, so every
// product and sum is rounded as written.
__global__ void k_ixyh(int nobs, const double *zobs, const cx<double> *exy, const int *ptr, const int *item,
                       const double *val, const cx<double> *v, const double *F, double sigma, cx<double> *out,
                       double *hout) {
#pragma clang fp contract(off)
  const int o = blockIdx.x * 256 + threadIdx.x;
  if (o >= nobs) return;
  double h = -zobs[o];
  if (h <= 0.0) h = 0.0;
  cx<double> vf[2] = {{0.0, 0.0}, {0.0, 0.0}}, vh[2] = {{0.0, 0.0}, {0.0, 0.0}};
  for (int q = ptr[o]; q < ptr[o + 1]; ++q) {
    const int nd = item[q] - 1;
    const double a = val[q];
    const cx<double> vx = v[3 * (size_t)nd], vy = v[3 * (size_t)nd + 1], vz = v[3 * (size_t)nd + 2];
    const double Fx = F[3 * (size_t)nd], Fy = F[3 * (size_t)nd + 1], Fz = F[3 * (size_t)nd + 2];
    const cx<double> c0 = {vy.r * Fz - vz.r * Fy, vy.i * Fz - vz.i * Fy};
    const cx<double> c1 = {vz.r * Fx - vx.r * Fz, vz.i * Fx - vx.i * Fz};
    vf[0].r = vf[0].r + a * c0.r;
    vf[0].i = vf[0].i + a * c0.i;
    vf[1].r = vf[1].r + a * c1.r;
    vf[1].i = vf[1].i + a * c1.i;
    vh[0].r = vh[0].r + a * vx.r;
    vh[0].i = vh[0].i + a * vx.i;
    vh[1].r = vh[1].r + a * vy.r;
    vh[1].i = vh[1].i + a * vy.i;
  }
  const double sh = sigma * h;
  cx<double> *r = out + 8 * (size_t)o;
  for (int c = 0; c < 2; c++) {
    const cx<double> e = exy[2 * (size_t)o + c];
    r[c] = {sh * e.r, sh * e.i};
    r[2 + c] = {sh * vf[c].r * 1.0e-3, sh * vf[c].i * 1.0e-3};
    r[4 + c] = {sh * (e.r + vf[c].r * 1.0e-3), sh * (e.i + vf[c].i * 1.0e-3)};
    r[6 + c] = {h * vh[c].r, h * vh[c].i};
  }
  hout[o] = h;
}
hipError_t launch_ixyh(int nobs, const double *zobs, const cx<double> *exy, const int *ptr, const int *item,
                       const double *val, const cx<double> *v, const double *F, double sigma, cx<double> *out,
                       double *hout, hipStream_t s) {
  if (nobs <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_ixyh, dim3((nobs + 255) / 256), dim3(256), 0, s, nobs, zobs, exy, ptr, item, val, v, F, sigma,
                     out, hout);
  return hipGetLastError();
}
hipError_t launch_obs_apply(int nobs, const int *item, const double *val, int ncol, const cx<double> *X,
                            cx<double> *out, hipStream_t s) {
  const long long m = (long long)nobs * ncol;
  if (m <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_obs_apply, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, nobs, item, val, ncol, X, out);
  return hipGetLastError();
}

// Adjoint right-hand sides of solvePARDISOinv (src/src_inv/solvePARDISOinv.f90:176-187): column o
// of B is row o of the real CRS observation matrix (stack 0-based offsets, item 1-based edges)
// scattered into a dense n-vector (assignment, as b1(P%item(j), i) = P%val(j)); B row-interleaved.
__global__ void k_adjoint_rhs(int nobs, const int *stack, const int *item, const double *val, cx<double> *B) {
  const int o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= nobs) return;
  for (int q = stack[o]; q < stack[o + 1]; ++q) B[(size_t)(item[q] - 1) * nobs + o] = cx<double>{val[q], 0.0};
}
__global__ void k_zero_cx(long long m, cx<double> *B) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i < m) B[i] = cx<double>{0.0, 0.0};
}
hipError_t launch_adjoint_rhs(int n, int nobs, const int *stack, const int *item, const double *val, cx<double> *B,
                              hipStream_t s) {
  const long long m = (long long)n * nobs;
  if (m <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_zero_cx, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, m, B);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_adjoint_rhs, dim3((nobs + 255) / 256), dim3(256), 0, s, nobs, stack, item, val, B);
  return hipGetLastError();
}

// Jacobian, step 1 (genjacobian1 :292-322): per selected tet q,
//   w_q[i] = -(dBB/dm) * sum_j Y(i,j) x[e_j],  dBB/dm = i*omega*mu0*sigma*L0^2*dlnsig
__global__ void k_jac_w(int nsel, const int *tets, const double *xyz, const int *n4, const int *n6line,
                        const double *sigma, double omega, double dlnsig, const cx<double> *x, cx<double> *w,
                        int *edges) {
  const int q = blockIdx.x * 256 + threadIdx.x;
  if (q >= nsel) return;
  const int t = tets[q];
  double xt[4][3], xx[6][3], gn[4][3], v;
  load_tet(xyz, n4, t, xt);
  double dir[6];
  int ed[6];
  for (int j = 0; j < 6; j++) {
    const int s = n6line[(size_t)t * 6 + j];
    dir[j] = (s < 0) ? -1.0 : 1.0;
    ed[j] = (s < 0 ? -s : s) - 1;
    edges[(size_t)q * 6 + j] = ed[j];
  }
  if (!geometry(xt, xx, gn, v)) {
    for (int i = 0; i < 6; i++) w[(size_t)q * 6 + i] = cx<double>{0.0, 0.0};
    return;
  }
  // dBB/dm is purely imaginary: b = omega*mu0*sigma*L0^2*dlnsig
  const double b = omega * (4.0 * 3.14159265358979323846 * 1.0e-7) * sigma[t] * 1.0e6 * dlnsig;
  cx<double> xe[6];
  for (int j = 0; j < 6; j++) xe[j] = x[ed[j]];
  for (int i = 0; i < 6; i++) {
    cx<double> s = {0.0, 0.0};
    for (int j = 0; j < 6; j++) {
      const double y = mass_yy(gn, v, i, j) * dir[i] * dir[j];
      s.r += y * xe[j].r;
      s.i += y * xe[j].i;
    }
    // -(i b) * s = (b*s.i, -b*s.r)
    w[(size_t)q * 6 + i] = cx<double>{b * s.i, -b * s.r};
  }
}
hipError_t launch_jac_w(int nsel, const int *tets, const double *xyz, const int *n4, const int *n6line,
                        const double *sigma, double omega, double dlnsig, const cx<double> *x, cx<double> *w,
                        int *edges, hipStream_t s) {
  if (nsel <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_jac_w, dim3((nsel + 255) / 256), dim3(256), 0, s, nsel, tets, xyz, n4, n6line, sigma, omega,
                     dlnsig, x, w, edges);
  return hipGetLastError();
}

// Jacobian, step 2 (genjacobian1 :324-331): J[m][o] = sum_{q in m} sum_j U[e_qj][o] * w_q[j].
// One block per parameter, one thread per observation column (the block is nobs rounded up to
// whole waves, at most 1024 threads; wider U in steps of the block): U rows are read
// contiguously. A parameter's (edge, weight) pairs -- 36 for a cell of 6 tets, about 19 distinct
// edges -- are merged first: the first wave marks each edge's first occurrence, adds the weights
// of its repeats in pair order and compacts the distinct edges (their first-occurrence order) into
// LDS. Every U row is then read once per block, and the JAC_BATCH loads of a batch are in flight
// before the first product (a batch of 24 -- a whole cell at once -- ran 1.8x slower: fewer waves
// per CU). Parameters of more than 64 pairs take the pair-by-pair loop.
constexpr int JAC_PAIRS = 64, JAC_BATCH = 8;
__global__ void __launch_bounds__(1024) k_jac_acc(int npar, const int *par_ptr, int nobs, const cx<double> *U,
                                                  const cx<double> *w, const int *edges, cx<double> *J) {
  __shared__ int pe[JAC_PAIRS], se[JAC_PAIRS];
  __shared__ cx<double> pw[JAC_PAIRS], sw[JAC_PAIRS];
  __shared__ int s_nd;
  const int m = blockIdx.x;
  if (m >= npar) return;
  const int t = threadIdx.x;
  const int qb = par_ptr[m], qe = par_ptr[m + 1];
  const int np = (qe - qb) * 6;  // (edge, weight) pairs of the parameter, pair i = (q, j) = (qb + i / 6, i % 6)
  const bool merged = np <= JAC_PAIRS;
  if (merged && t < 64) {
    pe[t] = t < np ? edges[(size_t)qb * 6 + t] : -1;
    pw[t] = t < np ? w[(size_t)qb * 6 + t] : cx<double>{0.0, 0.0};
  }
  __syncthreads();
  if (merged && t < 64) {
    const int e = pe[t];
    bool first = t < np;
    for (int i = 0; i < t; i++)
      if (pe[i] == e) first = false;
    const unsigned long long mask = __ballot(first);
    if (first) {
      cx<double> ws = {0.0, 0.0};
      for (int i = t; i < np; i++)
        if (pe[i] == e) {
          ws.r += pw[i].r;
          ws.i += pw[i].i;
        }
      const int pos = __popcll(mask & ((1ull << t) - 1ull));
      se[pos] = e;
      sw[pos] = ws;
    }
    if (t == 0) s_nd = __popcll(mask);
  }
  __syncthreads();
  for (int o = t; o < nobs; o += blockDim.x) {
    cx<double> acc = {0.0, 0.0};
    if (merged) {
      const int nd = s_nd;
      for (int d0 = 0; d0 < nd; d0 += JAC_BATCH) {
        cx<double> u[JAC_BATCH];
#pragma unroll
        for (int b = 0; b < JAC_BATCH; b++) {
          const int d = d0 + b < nd ? d0 + b : d0;
          u[b] = U[(size_t)se[d] * nobs + o];
        }
#pragma unroll
        for (int b = 0; b < JAC_BATCH; b++)
          if (d0 + b < nd) {
            const cx<double> ww = sw[d0 + b];
            acc.r = fma(u[b].r, ww.r, acc.r);
            acc.r = fma(-u[b].i, ww.i, acc.r);
            acc.i = fma(u[b].r, ww.i, acc.i);
            acc.i = fma(u[b].i, ww.r, acc.i);
          }
      }
    } else {
      for (int i = 0; i < np; i++) {
        const cx<double> ww = w[(size_t)qb * 6 + i];
        const cx<double> u = U[(size_t)edges[(size_t)qb * 6 + i] * nobs + o];
        acc.r = fma(u.r, ww.r, acc.r);
        acc.r = fma(-u.i, ww.i, acc.r);
        acc.i = fma(u.r, ww.i, acc.i);
        acc.i = fma(u.i, ww.r, acc.i);
      }
    }
    J[(size_t)m * nobs + o] = acc;
  }
}
hipError_t launch_jac_acc(int npar, const int *par_ptr, int nobs, const cx<double> *U, const cx<double> *w,
                          const int *edges, cx<double> *J, hipStream_t s) {
  if (npar <= 0 || nobs <= 0) return hipSuccess;
  const int bs = std::min(1024, ((nobs + 63) / 64) * 64);
  hipLaunchKernelGGL(k_jac_acc, dim3(npar), dim3(bs), 0, s, npar, par_ptr, nobs, U, w, edges, J);
  return hipGetLastError();
}

}  // namespace afx
