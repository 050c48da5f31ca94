/*
 * CPU oracle for the tidal edge-FEM hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this
 * library, and only as the checker (or the timed CPU baseline). The product path
 * (actfemtide_amd + libafx.so) never calls it.
 *
 * Plain C restatement of the reference's serial Fortran, in the same operation order,
 * with 0-based C arrays and the reference's 1-based index values where the reference
 * stores them (n4, n6line, IAU/IAL column ids, INU/INL offsets starting at 0).
 * Complex numbers are (re, im) pairs with the Fortran product rule spelled out, and the
 * library is built with -ffp-contract=off so that no FMA changes the rounding.
 *
 * Pinning of each function (the checks live in tests/):
 *   orc_gradnodebasis / element matrices : closed forms (unit tet, Jin's textbook
 *       element matrices), S*G = 0 for the discrete gradient, sum of grad lambda = 0,
 *       and the flang-built reference GENMAT where available (oracle/_ref).
 *   orc_pattern : brute-force pair set; reference set_iccg_var7_dofn via oracle/_ref.
 *   orc_genmat / orc_copy_ul / orc_genbc / orc_set_bc : dense assembly on tiny meshes,
 *       Dirichlet-face closed form, reference GENMAT/COPY_UL/GENBCCSEM/SET_BC via oracle/_ref.
 *   orc_cocg13 : identical iterate sequence to the reference COCG13_single via oracle/_ref,
 *       a direct solve (scipy SuperLU) on the same system, diag(M) = diag(A) for D-ILU.
 *   orc_obs_* : Whitney-form exactness (constant E, linear A), reference PREPOBSCOEFF via
 *       oracle/_ref.
 *   orc_jacobian : central finite differences and a direct-sensitivity solve.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct { double re, im; } cplx;

static inline cplx cadd(cplx a, cplx b) { cplx r = {a.re + b.re, a.im + b.im}; return r; }
static inline cplx csub(cplx a, cplx b) { cplx r = {a.re - b.re, a.im - b.im}; return r; }
static inline cplx cmul(cplx a, cplx b) {
  cplx r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
  return r;
}
/* real * complex, as Fortran promotes the real to (x, 0) */
static inline cplx rmul(double x, cplx b) {
  cplx a = {x, 0.0};
  return cmul(a, b);
}
static inline cplx cdiv(cplx a, cplx b) {
  /* (a*conj(b))/|b|^2 with power-of-two scaling of b (C99 Annex G / compiler-rt __divdc3,
     which flang calls for complex(8) division): the scaling is exact, so the rounding is
     that of the textbook formula. */
  double c = b.re, d = b.im;
  int ilogbw = 0;
  double logbw = logb(fmax(fabs(c), fabs(d)));
  if (isfinite(logbw)) {
    ilogbw = (int)logbw;
    c = scalbn(c, -ilogbw);
    d = scalbn(d, -ilogbw);
  }
  double denom = c * c + d * d;
  cplx r;
  r.re = scalbn((a.re * c + a.im * d) / denom, -ilogbw);
  r.im = scalbn((a.im * c - a.re * d) / denom, -ilogbw);
  return r;
}

/* kl: local edges (m_fem_edge_util.f90:141-147), 0-based */
static const int KL[6][2] = {{0, 1}, {1, 2}, {0, 2}, {0, 3}, {1, 3}, {2, 3}};
/* lmn: local faces (m_fem_edge_util.f90:149-153), 0-based */
static const int LMN[4][3] = {{1, 2, 3}, {0, 3, 2}, {0, 1, 3}, {0, 2, 1}};
/* C: face-to-edge rotation matrix (m_fem_edge_util.f90:155-159) */
static const int CROT[4][6] = {{0, 1, 0, 0, -1, 1}, {0, 0, -1, 1, 0, -1},
                               {1, 0, 0, -1, 1, 0}, {-1, -1, 1, 0, 0, 0}};

#define DMU (4.0 * 3.14159265358979323846 * 1.0e-7) /* m_constants.f90:324 */

static inline double inner3(const double *a, const double *b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
static inline void outer3(const double *a, const double *b, double *c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}
static inline void sub3(const double *a, const double *b, double *c) {
  c[0] = a[0] - b[0]; c[1] = a[1] - b[1]; c[2] = a[2] - b[2];
}
/* volume(x1,x2,x3) = 1/6 (x1 x x2).x3  (m_fem_util.f90:101-109) */
static inline double volume3(const double *x1, const double *x2, const double *x3) {
  double c[3];
  outer3(x1, x2, c);
  return 1.0 / 6.0 * inner3(c, x3);
}

/* calxmn (m_fem_util.f90:42-53): x = {x4-x3, x4-x1, x2-x4, x3-x2, x1-x3, x2-x1} */
void orc_calxmn(const double x[4][3], double xx[6][3]) {
  sub3(x[3], x[2], xx[0]);
  sub3(x[3], x[0], xx[1]);
  sub3(x[1], x[3], xx[2]);
  sub3(x[2], x[1], xx[3]);
  sub3(x[0], x[2], xx[4]);
  sub3(x[1], x[0], xx[5]);
}

/* gradnodebasisfun (m_fem_util.f90:12-40). Returns -1 when v <= 0 (the reference stops). */
int orc_gradnodebasis(const double x[4][3], double gn[4][3], double *v_out) {
  double xx[6][3], m4[3], a[3], b[3], c[3];
  orc_calxmn(x, xx);
  m4[0] = -xx[3][0]; m4[1] = -xx[3][1]; m4[2] = -xx[3][2];
  double v = volume3(xx[0], xx[4], m4);
  *v_out = v;
  if (v <= 0) return -1;
  double s = 1.0 / 6.0 / v;
  sub3(x[3], x[1], a); sub3(x[2], x[1], b); outer3(a, b, c);
  for (int d = 0; d < 3; d++) gn[0][d] = s * c[d];
  sub3(x[2], x[0], a); sub3(x[3], x[0], b); outer3(a, b, c);
  for (int d = 0; d < 3; d++) gn[1][d] = s * c[d];
  sub3(x[3], x[0], a); sub3(x[1], x[0], b); outer3(a, b, c);
  for (int d = 0; d < 3; d++) gn[2][d] = s * c[d];
  sub3(x[1], x[0], a); sub3(x[2], x[0], b); outer3(a, b, c);
  for (int d = 0; d < 3; d++) gn[3][d] = s * c[d];
  return 0;
}

/* intv (m_fem_util.f90:86-96) */
static inline double intv(int k, int m, double v) { return (k != m) ? v / 20.0 : v / 10.0; }

/* nodebasisfun (m_fem_util.f90:56-83) */
void orc_nodebasis(const double x[4][3], const double *p, double a[4]) {
  double v[4][3], x12[3], x13[3], x14[3], x24[3], x23[3], mv[3];
  for (int j = 0; j < 4; j++) sub3(x[j], p, v[j]);
  sub3(x[1], x[0], x12); sub3(x[2], x[0], x13); sub3(x[3], x[0], x14);
  sub3(x[3], x[1], x24); sub3(x[2], x[1], x23);
  double a5 = volume3(x12, x13, x14);
  mv[0] = -v[0][0]; mv[1] = -v[0][1]; mv[2] = -v[0][2];
  double a4 = volume3(x12, x13, mv) / a5;
  double a3 = volume3(x14, x12, mv) / a5;
  double a2 = volume3(x13, x14, mv) / a5;
  mv[0] = -v[1][0]; mv[1] = -v[1][1]; mv[2] = -v[1][2];
  double a1 = volume3(x24, x23, mv) / a5;
  a[0] = a1; a[1] = a2; a[2] = a3; a[3] = a4;
}

/*
 * Element matrices of GENMAT (forward_bxyz_mpi.f90:159-235) for one tetrahedron:
 *   S(j,k)  = inner(xx_j, xx_k) * d_j * d_k * (1/9/v)                 (real curl-curl)
 *   yy(i,j) = mass-term integral; S1 = yy * d_i * d_j * BB, BB = i*omega*mu0*sigma*L0^2
 * Outputs S[36] (row-major i*6+j), Y[36] = yy*d_i*d_j (real), and v.
 */
int orc_element(const double x[4][3], const int dir[6], double S[36], double Y[36], double gn_out[4][3],
                double *v_out) {
  double xx[6][3], gn[4][3], v;
  orc_calxmn(x, xx);
  if (orc_gradnodebasis(x, gn, &v) != 0) { *v_out = v; return -1; }
  double AA = 1.0 / 9.0 / v;
  for (int j = 0; j < 6; j++)
    for (int k = 0; k < 6; k++)
      S[j * 6 + k] = inner3(xx[j], xx[k]) * (double)dir[j] * (double)dir[k] * AA;
  for (int i = 0; i < 6; i++) {
    for (int j = 0; j < 6; j++) {
      int k = KL[i][0], l = KL[i][1], m = KL[j][0], n = KL[j][1];
      double yy = intv(k, m, v) * inner3(gn[l], gn[n]) - intv(k, n, v) * inner3(gn[l], gn[m]) -
                  intv(l, m, v) * inner3(gn[k], gn[n]) + intv(l, n, v) * inner3(gn[k], gn[m]);
      Y[i * 6 + j] = yy * (double)dir[i] * (double)dir[j];
    }
  }
  if (gn_out) memcpy(gn_out, gn, sizeof(gn));
  *v_out = v;
  return 0;
}

/* locate column `col` (1-based) in upper row `row` (1-based); -1 if missing */
static long find_u(const int *INU, const int *IAU, int row, int col) {
  for (long p = INU[row - 1]; p < INU[row]; p++)
    if (IAU[p] == col) return p;
  return -1;
}

/*
 * Sparsity pattern, C_set_dof2 restated (m_iccg_var_takuto.f90:361-485) with dofn = 1:
 * per row, sorted unique strict-upper and strict-lower column ids of edges sharing a tet.
 * Pass IAU/IAL = NULL to only count (INU/INL filled); then call again with arrays of
 * size INU[n] / INL[n]. Returns 0, or -1 on a bad edge id.
 */
int orc_pattern(int nline, int ntet, const int *n6line, int *INU, int *IAU, int *INL, int *IAL) {
  /* per-row dynamic sorted lists */
  int cap = 32;
  int *cu = (int *)calloc((size_t)nline, sizeof(int));
  int *cl = (int *)calloc((size_t)nline, sizeof(int));
  int *lu = (int *)malloc((size_t)nline * cap * sizeof(int));
  int *ll = (int *)malloc((size_t)nline * cap * sizeof(int));
  if (!cu || !cl || !lu || !ll) { free(cu); free(cl); free(lu); free(ll); return -2; }
  int rc = 0;
  for (int t = 0; t < ntet && rc == 0; t++) {
    int dof[6];
    for (int a = 0; a < 6; a++) {
      int d = abs(n6line[t * 6 + a]);
      if (d < 1 || d > nline) { rc = -1; break; }
      dof[a] = d;
    }
    if (rc) break;
    for (int a = 0; a < 6; a++) {
      for (int b = 0; b < 6; b++) {
        int i = dof[a], j = dof[b];
        if (i == j) continue;
        int *cnt = (i < j) ? &cu[i - 1] : &cl[i - 1];
        int *lst = (i < j) ? &lu[(size_t)(i - 1) * cap] : &ll[(size_t)(i - 1) * cap];
        int pos = *cnt, exist = 0;
        for (int q = 0; q < *cnt; q++) {
          if (lst[q] == j) { exist = 1; break; }
          if (j < lst[q]) { pos = q; break; }
        }
        if (exist) continue;
        if (*cnt >= cap) { rc = -3; break; }
        for (int q = *cnt; q > pos; q--) lst[q] = lst[q - 1];
        lst[pos] = j;
        (*cnt)++;
      }
      if (rc) break;
    }
  }
  if (rc == 0) {
    INU[0] = 0; INL[0] = 0;
    for (int i = 0; i < nline; i++) { INU[i + 1] = INU[i] + cu[i]; INL[i + 1] = INL[i] + cl[i]; }
    if (IAU && IAL) {
      for (int i = 0; i < nline; i++) {
        memcpy(&IAU[INU[i]], &lu[(size_t)i * cap], (size_t)cu[i] * sizeof(int));
        memcpy(&IAL[INL[i]], &ll[(size_t)i * cap], (size_t)cl[i] * sizeof(int));
      }
    }
  }
  free(cu); free(cl); free(lu); free(ll);
  return rc;
}

/*
 * GENMAT (forward_bxyz_mpi.f90:99-311) + sup_iccg (m_iccg_var_takuto.f90:43-100).
 *   xyz[node*3], n4[ntet*4] (1-based), group[ntet] (n4flag(:,2)), n6line[ntet*6],
 *   sigma[ntet] (per-element sigma already selected), vxyz[node*3] complex, fxyz[node*3].
 *   Outputs accumulate into D[nline], AU[nU], b[nline] (caller zeroes them: INITIALIZE).
 * Returns 0; -1 for a non-positive volume (reference stops); -2 for a missing slot.
 */
int orc_genmat(int node, int ntet, const double *xyz, const int *n4, const int *group,
               const int *n6line, int nline, const double *sigma, double omega,
               const cplx *vxyz, const double *fxyz, const int *INU, const int *IAU,
               cplx *D, cplx *AU, cplx *b) {
  (void)node; (void)nline;
  const double L0 = 1.0e3;
  for (int t = 0; t < ntet; t++) {
    int dir[6], tdof[6];
    double x[4][3], S[36], Y[36], gn[4][3], v;
    for (int j = 0; j < 6; j++) {
      dir[j] = (n6line[t * 6 + j] < 0) ? -1 : 1;
      tdof[j] = n6line[t * 6 + j] * dir[j];
    }
    for (int j = 0; j < 4; j++)
      for (int d = 0; d < 3; d++) x[j][d] = xyz[(size_t)(n4[t * 4 + j] - 1) * 3 + d];
    if (orc_element(x, dir, S, Y, gn, &v) != 0) return -1;
    /* BB = iunit*omega*dmu*sigma*L0**2 evaluated left to right */
    cplx iu = {0.0, 1.0};
    cplx BB = rmul(omega, iu);
    BB = cmul(BB, (cplx){DMU, 0.0});
    BB = cmul(BB, (cplx){sigma[t], 0.0});
    BB = cmul(BB, (cplx){L0 * L0, 0.0});
    cplx ek[36];
    for (int i = 0; i < 36; i++) {
      cplx s1 = rmul(Y[i], BB);
      ek[i].re = S[i] + s1.re;
      ek[i].im = 0.0 + s1.im;
    }
    /* sup_iccg: diagonal and upper, taking the element-lower entry */
    for (int ie = 0; ie < 6; ie++) {
      int ig = tdof[ie];
      for (int je = 0; je < 6; je++) {
        int jg = tdof[je];
        if (ig == jg) {
          D[ig - 1] = cadd(D[ig - 1], ek[ie * 6 + je]);
        } else if (ig < jg) {
          long p = find_u(INU, IAU, ig, jg);
          if (p < 0) return -2;
          cplx val = (ie > je) ? ek[ie * 6 + je] : ek[je * 6 + ie];
          AU[p] = cadd(AU[p], val);
        }
      }
    }
    /* tidal right-hand side for sea elements (forward_bxyz_mpi.f90:267-299) */
    cplx S2[6] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}};
    if (group[t] == 2) {
      cplx vF[4][3];
      for (int i = 0; i < 4; i++) {
        int no = n4[t * 4 + i] - 1;
        const cplx *vv = &vxyz[(size_t)no * 3];
        const double *F = &fxyz[(size_t)no * 3];
        /* cr_outer(v, F) (m_outerinnerproduct.f90:281-289) */
        vF[i][0] = csub(rmul(F[2], vv[1]), rmul(F[1], vv[2]));
        vF[i][1] = csub(rmul(F[0], vv[2]), rmul(F[2], vv[0]));
        vF[i][2] = csub(rmul(F[1], vv[0]), rmul(F[0], vv[1]));
        /* note: Fortran evaluates v1(2)*v2(3) as complex*real; rmul(x, c) gives the same
           rounding because the product rule is symmetric for a zero imaginary part */
      }
      for (int i = 0; i < 6; i++) {
        int k = KL[i][0], l = KL[i][1];
        for (int j = 0; j < 4; j++) {
          /* rc_inner(gn_l, vF_j) */
          cplx il = rmul(gn[l][0], vF[j][0]);
          il = cadd(il, rmul(gn[l][1], vF[j][1]));
          il = cadd(il, rmul(gn[l][2], vF[j][2]));
          cplx ik = rmul(gn[k][0], vF[j][0]);
          ik = cadd(ik, rmul(gn[k][1], vF[j][1]));
          ik = cadd(ik, rmul(gn[k][2], vF[j][2]));
          cplx term = csub(rmul(intv(k, j, v), il), rmul(intv(l, j, v), ik));
          double pre = DMU * sigma[t];
          cplx c = rmul(pre, term);
          c = rmul((double)dir[i], c);
          c = rmul(1.0e3, c);
          S2[i] = cadd(S2[i], c);
        }
      }
    }
    for (int i = 0; i < 6; i++) b[tdof[i] - 1] = cadd(b[tdof[i] - 1], S2[i]);
  }
  return 0;
}

/* Mass-only and curl-only assembly (A = K + i*omega*M split): K gets S, M gets
   Y*mu0*sigma*L0^2 (so that BB*Y = i*omega*M). Same slot rule as sup_iccg. */
int orc_genmat_split(int ntet, const double *xyz, const int *n4, const int *n6line,
                     const double *sigma, const int *INU, const int *IAU, double *Kd,
                     double *Ku, double *Md, double *Mu) {
  const double L0 = 1.0e3;
  for (int t = 0; t < ntet; t++) {
    int dir[6], tdof[6];
    double x[4][3], S[36], Y[36], v;
    for (int j = 0; j < 6; j++) {
      dir[j] = (n6line[t * 6 + j] < 0) ? -1 : 1;
      tdof[j] = n6line[t * 6 + j] * dir[j];
    }
    for (int j = 0; j < 4; j++)
      for (int d = 0; d < 3; d++) x[j][d] = xyz[(size_t)(n4[t * 4 + j] - 1) * 3 + d];
    if (orc_element(x, dir, S, Y, NULL, &v) != 0) return -1;
    double ms = DMU * sigma[t] * (L0 * L0);
    for (int ie = 0; ie < 6; ie++) {
      int ig = tdof[ie];
      for (int je = 0; je < 6; je++) {
        int jg = tdof[je];
        int e = (ie > je) ? ie * 6 + je : je * 6 + ie;
        if (ig == jg) {
          Kd[ig - 1] += S[e];
          Md[ig - 1] += Y[e] * ms;
        } else if (ig < jg) {
          long p = find_u(INU, IAU, ig, jg);
          if (p < 0) return -2;
          Ku[p] += S[e];
          Mu[p] += Y[e] * ms;
        }
      }
    }
  }
  return 0;
}

/* COPY_UL_ICCG12 (m_iccg_var_takuto.f90:107-167). Returns -1 on a missing mirror slot. */
int orc_copy_ul(int n, const int *INU, const int *IAU, const int *INL, const int *IAL,
                const cplx *AU, cplx *AL) {
  long nl = INL[n];
  for (long p = 0; p < nl; p++) { AL[p].re = 0; AL[p].im = 0; }
  for (int i = 1; i <= n; i++) {
    for (long num = INU[i - 1]; num < INU[i]; num++) {
      int j = IAU[num];
      long pos = -1;
      for (long q = INL[j - 1]; q < INL[j]; q++)
        if (IAL[q] == i) { pos = q; break; }
      if (pos < 0) return -1;
      AL[pos] = AU[num];
    }
  }
  return 0;
}

/* GENBCCSEM (forward_bxyz_mpi.f90:467-519): line_bc[i] = 1 when both end points lie within
   1 km of the same outer face; line_group gets the face id (1 top .. 6 bottom). */
void orc_genbc(double zmin, double zmax, double xout, double yout, int nline, const int *line,
               const double *xyz, unsigned char *line_bc, int *line_group) {
  double xout1 = xout - 1.0, yout1 = yout - 1.0, zmin1 = zmin + 1.0, zmax1 = zmax - 1.0;
  for (int i = 0; i < nline; i++) {
    int n1 = line[2 * i] - 1, n2 = line[2 * i + 1] - 1;
    double x1 = xyz[n1 * 3], y1 = xyz[n1 * 3 + 1], z1 = xyz[n1 * 3 + 2];
    double x2 = xyz[n2 * 3], y2 = xyz[n2 * 3 + 1], z2 = xyz[n2 * 3 + 2];
    int g = 0;
    if (x1 <= -xout1 && x2 <= -xout1) g = 3;
    else if (x1 >= xout1 && x2 >= xout1) g = 5;
    else if (y1 <= -yout1 && y2 <= -yout1) g = 4;
    else if (y1 >= yout1 && y2 >= yout1) g = 2;
    else if (z1 <= zmin1 && z2 <= zmin1) g = 6;
    else if (z1 >= zmax1 && z2 >= zmax1) g = 1;
    line_bc[i] = (g != 0);
    if (line_group) line_group[i] = g;
  }
}

/* SET_BC_ICCG (forward_bxyz_mpi.f90:567-636); rf is (n, nsr) column-major. */
void orc_set_bc(int n, int nsr, const int *INU, const int *IAU, const int *INL, const int *IAL,
                cplx *D, cplx *AU, cplx *AL, cplx *rf, const cplx *dval, const unsigned char *dir) {
  for (int i = 0; i < n; i++) {
    if (dir[i]) {
      for (int j = 0; j < nsr; j++) rf[(size_t)j * n + i] = dval[i];
    } else {
      cplx temp = cmul(D[i], dval[i]);
      for (long p = INL[i]; p < INL[i + 1]; p++) temp = cadd(temp, cmul(AL[p], dval[IAL[p] - 1]));
      for (long p = INU[i]; p < INU[i + 1]; p++) temp = cadd(temp, cmul(AU[p], dval[IAU[p] - 1]));
      for (int j = 0; j < nsr; j++) rf[(size_t)j * n + i] = csub(rf[(size_t)j * n + i], temp);
    }
  }
  for (int i = 0; i < n; i++) {
    if (dir[i]) {
      D[i].re = 1.0; D[i].im = 0.0;
      for (long p = INL[i]; p < INL[i + 1]; p++) { AL[p].re = 0; AL[p].im = 0; }
      for (long p = INU[i]; p < INU[i + 1]; p++) { AU[p].re = 0; AU[p].im = 0; }
    } else {
      for (long p = INL[i]; p < INL[i + 1]; p++)
        if (dir[IAL[p] - 1]) { AL[p].re = 0; AL[p].im = 0; }
      for (long p = INU[i]; p < INU[i + 1]; p++)
        if (dir[IAU[p] - 1]) { AU[p].re = 0; AU[p].im = 0; }
    }
  }
}

/* y = A x with the split storage, row order D, U, L (COCG13_single.f90:273-288) */
void orc_spmv(int n, const cplx *D, const int *INU, const int *IAU, const cplx *AU,
              const int *INL, const int *IAL, const cplx *AL, const cplx *x, cplx *y) {
  for (int j = 0; j < n; j++) {
    cplx w = cmul(D[j], x[j]);
    for (long p = INU[j]; p < INU[j + 1]; p++) w = cadd(w, cmul(AU[p], x[IAU[p] - 1]));
    for (long p = INL[j]; p < INL[j + 1]; p++) w = cadd(w, cmul(AL[p], x[IAL[p] - 1]));
    y[j] = w;
  }
}

/* Preconditioner setup (COCG13_single.f90:105-130). kind: 1 IC/ILU, 2 SSOR, 3 DIAG, 0 none */
void orc_precond_setup(int n, int kind, double sigma_diag, double sigma, const cplx *D,
                       const int *INU, const cplx *AU, const int *INL, const int *IAL,
                       const cplx *AL, cplx *DD) {
  for (int i = 0; i < n; i++) { DD[i].re = 0; DD[i].im = 0; }
  if (kind == 1) {
    for (int i = 0; i < n; i++) {
      cplx W = rmul(sigma_diag, D[i]);
      for (long k = INL[i]; k < INL[i + 1]; k++) {
        cplx SS = AL[k];
        int id = IAL[k];
        for (long kk = INU[id - 1]; kk < INU[id]; kk++) SS = cadd(SS, rmul(sigma, AU[kk]));
        W = csub(W, cmul(cmul(AL[k], SS), DD[id - 1]));
      }
      DD[i] = cdiv((cplx){1.0, 0.0}, W);
    }
  } else if (kind == 2 || kind == 3) {
    for (int i = 0; i < n; i++) DD[i] = cdiv((cplx){1.0, 0.0}, rmul(sigma_diag, D[i]));
  }
}

/* z = M^-1 r (COCG13_single.f90:186-241); kinds 1,2 sweep, 3 diag, 0 copy */
void orc_precond_apply(int n, int kind, const int *INU, const int *IAU, const cplx *AU,
                       const int *INL, const int *IAL, const cplx *AL, const cplx *DD,
                       const cplx *r, cplx *z) {
  for (int i = 0; i < n; i++) z[i] = r[i];
  if (kind == 1 || kind == 2) {
    for (int i = 0; i < n; i++) {
      cplx w = r[i];
      for (long j = INL[i]; j < INL[i + 1]; j++) w = csub(w, cmul(AL[j], z[IAL[j] - 1]));
      z[i] = cmul(w, DD[i]);
    }
    for (int i = n - 1; i >= 0; i--) {
      cplx sw = {0.0, 0.0};
      for (long j = INU[i]; j < INU[i + 1]; j++) sw = cadd(sw, cmul(AU[j], z[IAU[j] - 1]));
      z[i] = csub(z[i], cmul(DD[i], sw));
    }
  } else if (kind == 3) {
    for (int i = 0; i < n; i++) z[i] = cmul(r[i], DD[i]);
  }
}

static int precond_kind(const char *p) {
  if (strncmp(p, "IC", 2) == 0 || strncmp(p, "ILU", 3) == 0) return 1;
  if (strncmp(p, "SSOR", 4) == 0) return 2;
  if (strncmp(p, "DIAG", 4) == 0) return 3;
  return 0;
}

/*
 * COCG13_single (COCG13_single.f90:9-341), same argument meaning:
 *   D, AU, AL modified in place for NSET = 2 (scaling); B scaled and restored; X in/out;
 *   *resid: tolerance in, achieved recurrence residual out; *iter: max in, count out
 *   (on exhaustion the reference's DO variable ends at MAXIT + 1); *error: 0 / -100 / 101.
 *   DD_state and scale_state: caller-owned length-n arrays that play the role of the SAVEd DD and
 *   SCALE (:39-40); *have_state plays IFLAG (:44,49-62): on the first call SCALE is set to 1, an
 *   NSET = 2 call overwrites it, NSET = 1 scales B by it, NSET = 0 leaves B and DD alone and still
 *   divides B by it at the end (:98-100,326-329). resid_hist (optional, length >= maxit) receives
 *   ||r||/||b|| per iteration.
 */
void orc_cocg13(int np, int npl, int npu, cplx *D, cplx *AL, const int *INL, const int *IAL,
                cplx *AU, const int *INU, const int *IAU, cplx *B, cplx *X, const char *precond,
                double sigma_diag, double sigma, double *resid, int *iter, int *error, int nset,
                cplx *DD_state, double *scale_state, int *have_state, double *resid_hist) {
  (void)npl; (void)npu;
  int n = np;
  *error = 0;
  if (!*have_state && nset == 0) { *error = 101; return; }
  double *scale = scale_state;
  cplx *R = (cplx *)malloc((size_t)n * sizeof(cplx));
  cplx *Zq = (cplx *)malloc((size_t)n * sizeof(cplx));
  cplx *P = (cplx *)malloc((size_t)n * sizeof(cplx));
  if (!*have_state) {
    for (int i = 0; i < n; i++) scale[i] = 1.0;
    *have_state = 1;
  }
  int maxit = *iter;
  double tol = *resid;
  int kind = precond_kind(precond);
  if (nset >= 1) {
    if (nset == 2) {
      for (int i = 0; i < n; i++) scale[i] = 1.0 / sqrt(hypot(D[i].re, D[i].im));
      for (int j = 0; j < n; j++) {
        double ad = hypot(D[j].re, D[j].im);
        D[j] = cdiv((cplx){ad, 0.0}, D[j]);
        for (long i = INU[j]; i < INU[j + 1]; i++) {
          int c = IAU[i] - 1;
          AU[i] = rmul(scale[j], rmul(scale[c], AU[i]));
        }
        for (long i = INL[j]; i < INL[j + 1]; i++) {
          int c = IAL[i] - 1;
          AL[i] = rmul(scale[j], rmul(scale[c], AL[i]));
        }
      }
    }
    for (int i = 0; i < n; i++) B[i] = rmul(scale[i], B[i]);
    orc_precond_setup(n, kind, sigma_diag, sigma, D, INU, AU, INL, IAL, AL, DD_state);
  }
  /* r0 = b - A x (order: D, U, L) */
  for (int j = 0; j < n; j++) {
    cplx w = csub(B[j], cmul(D[j], X[j]));
    for (long i = INU[j]; i < INU[j + 1]; i++) w = csub(w, cmul(AU[i], X[IAU[i] - 1]));
    for (long i = INL[j]; i < INL[j + 1]; i++) w = csub(w, cmul(AL[i], X[IAL[i] - 1]));
    R[j] = w;
  }
  double bnrm2 = 0.0;
  for (int i = 0; i < n; i++) bnrm2 = bnrm2 + (B[i].re * B[i].re + B[i].im * B[i].im);
  if (bnrm2 == 0.0) {
    for (int i = 0; i < n; i++) { X[i].re = 0; X[i].im = 0; }
    *resid = 0.0;
    *iter = 0;
    free(R); free(Zq); free(P);
    return;
  }
  cplx rho = {0, 0}, rho1 = {0, 0};
  int it;
  double res = 0.0;
  for (it = 1; it <= maxit; it++) {
    orc_precond_apply(n, kind, INU, IAU, AU, INL, IAL, AL, DD_state, R, Zq);
    cplx rho0 = {0, 0};
    for (int i = 0; i < n; i++) rho0 = cadd(rho0, cmul(R[i], Zq[i]));
    rho = rho0;
    if (it == 1) {
      for (int i = 0; i < n; i++) P[i] = Zq[i];
    } else {
      cplx beta = cdiv(rho, rho1);
      for (int i = 0; i < n; i++) P[i] = cadd(Zq[i], cmul(beta, P[i]));
    }
    /* q = A p (q shares storage with z) */
    for (int j = 0; j < n; j++) {
      cplx w = cmul(D[j], P[j]);
      for (long i = INU[j]; i < INU[j + 1]; i++) w = cadd(w, cmul(AU[i], P[IAU[i] - 1]));
      for (long i = INL[j]; i < INL[j + 1]; i++) w = cadd(w, cmul(AL[i], P[IAL[i] - 1]));
      Zq[j] = w;
    }
    cplx c1 = {0, 0};
    for (int i = 0; i < n; i++) c1 = cadd(c1, cmul(P[i], Zq[i]));
    cplx alpha = cdiv(rho, c1);
    for (int i = 0; i < n; i++) {
      X[i] = cadd(X[i], cmul(alpha, P[i]));
      R[i] = csub(R[i], cmul(alpha, Zq[i]));
    }
    double dnrm2 = 0.0;
    for (int i = 0; i < n; i++) dnrm2 = dnrm2 + (R[i].re * R[i].re + R[i].im * R[i].im);
    res = sqrt(dnrm2 / bnrm2);
    if (resid_hist) resid_hist[it - 1] = res;
    if (res <= tol) break;
    if (it == maxit) *error = -100;
    rho1 = rho;
  }
  *resid = res;
  *iter = it;  /* MAXIT + 1 after an exhausted loop, as the Fortran DO variable */
  for (int i = 0; i < n; i++) {
    X[i] = rmul(scale[i], X[i]);
    B[i].re = B[i].re / scale[i];  /* B(i) = B(i) / SCALE(i) */
    B[i].im = B[i].im / scale[i];
  }
  free(R); free(Zq); free(P);
}

/* ---------------- observation operator ---------------- */

/* FINDELEMENT0 (m_mesh_type.f90:697-734): first tet whose 4 barycentric coords are >= 0.
   Returns the 0-based element or -1. */
int orc_findelement0(int ntet, const double *xyz, const int *n4, const double *p, double a[4]) {
  for (int t = 0; t < ntet; t++) {
    double x[4][3];
    for (int j = 0; j < 4; j++)
      for (int d = 0; d < 3; d++) x[j][d] = xyz[(size_t)(n4[t * 4 + j] - 1) * 3 + d];
    orc_nodebasis(x, p, a);
    if (a[0] >= 0.0 && a[1] >= 0.0 && a[2] >= 0.0 && a[3] >= 0.0) return t;
  }
  return -1;
}

/*
 * PREPOBSCOEFF (n_ebfem_bxyz_mpi.f90:653-726) for nobs points xyzobs[nobs*3]:
 * rows of 6 entries sorted by edge id; item[nobs*6] (1-based edge ids),
 * val_e[3][nobs*6] (edge basis, coeffobs(1,:)), val_b[3][nobs*6] (face basis, coeffobs(2,:)).
 * Returns 0, or -(i+1) when observation i lies outside the mesh.
 */
int orc_prepobscoeff(int nobs, const double *xyzobs, int ntet, const double *xyz, const int *n4,
                     const int *n6line, int *item, double *val_e, double *val_b, int *iele_out) {
  for (int i = 0; i < nobs; i++) {
    double a4[4];
    int t = orc_findelement0(ntet, xyz, n4, &xyzobs[i * 3], a4);
    if (t < 0) return -(i + 1);
    if (iele_out) iele_out[i] = t;
    double x[4][3], gn[4][3], v, lam[4], w[6][3], wf[4][3], wfe[6][3];
    for (int j = 0; j < 4; j++)
      for (int d = 0; d < 3; d++) x[j][d] = xyz[(size_t)(n4[t * 4 + j] - 1) * 3 + d];
    orc_gradnodebasis(x, gn, &v);
    orc_nodebasis(x, &xyzobs[i * 3], lam);
    for (int e = 0; e < 6; e++) {
      int k = KL[e][0], l = KL[e][1];
      for (int d = 0; d < 3; d++) w[e][d] = lam[k] * gn[l][d] - lam[l] * gn[k][d];
    }
    for (int f = 0; f < 4; f++) {
      int l = LMN[f][0], m = LMN[f][1], n = LMN[f][2];
      double c1[3], c2[3], c3[3];
      outer3(gn[m], gn[n], c1);
      outer3(gn[l], gn[m], c2);
      outer3(gn[n], gn[l], c3);
      for (int d = 0; d < 3; d++)
        wf[f][d] = 2.0 * (lam[l] * c1[d] + lam[n] * c2[d] + lam[m] * c3[d]);
    }
    /* wfe = wf(3,4) * C(4,6) (mul_ab) */
    for (int e = 0; e < 6; e++)
      for (int d = 0; d < 3; d++) {
        double s = 0.0;
        for (int f = 0; f < 4; f++) s = s + wf[f][d] * (double)CROT[f][e];
        wfe[e][d] = s;
      }
    int order[6] = {0, 1, 2, 3, 4, 5};
    int ids[6];
    for (int e = 0; e < 6; e++) ids[e] = abs(n6line[t * 6 + e]);
    for (int a = 1; a < 6; a++) {  /* ids are distinct: any sort gives the same order */
      int o = order[a], q = a;
      while (q > 0 && ids[order[q - 1]] > ids[o]) { order[q] = order[q - 1]; q--; }
      order[q] = o;
    }
    for (int j = 0; j < 6; j++) {
      int e = order[j];
      int sg = (n6line[t * 6 + e] < 0) ? -1 : 1;
      item[i * 6 + j] = ids[e];
      for (int c = 0; c < 3; c++) {
        val_e[(size_t)c * nobs * 6 + i * 6 + j] = w[e][c] * (double)sg;
        val_b[(size_t)c * nobs * 6 + i * 6 + j] = wfe[e][c] * (double)sg;
      }
    }
  }
  return 0;
}

/* mul_matcrs_cv (m_matrix.f90:83-102) over the 6-entry observation rows */
void orc_obs_apply(int nobs, const int *item, const double *val, const cplx *x, cplx *out) {
  for (int i = 0; i < nobs; i++) {
    cplx s = {0, 0};
    for (int j = 0; j < 6; j++) s = cadd(s, rmul(val[i * 6 + j], x[item[i * 6 + j] - 1]));
    out[i] = s;
  }
}

/*
 * Adjoint Jacobian accumulation, genjacobian1 (src/src_inv_ap/m_jacobian_ap.f90:174-386)
 * restated for one observed component and one forward solution, without thresholding:
 *   dbe[o, m] = sum_{t in m} sum_j u_o[e_j] * ( -(S1_t(dBB/dm) x_t)_j )
 * with S1_t = Y_t * dBB/dm and dBB/dm = i*omega*mu0*sigma_t*L0^2 * dsig_factor,
 * where dsig_factor = d(ln sigma)/dm (ln 10 for m = log10 sigma; -ln 10 for log10 rho).
 * u is (nobs, nline) row-major: u[o*nline + e] (the adjoint solutions A^-1 P_o^T).
 * par_ptr/par_tet: CSR map parameter -> element ids (0-based), sigma[ntet].
 */
int orc_jacobian(int nobs, int nline, const cplx *u, const cplx *x, int npar, const int *par_ptr,
                 const int *par_tet, const double *xyz, const int *n4, const int *n6line,
                 const double *sigma, double omega, double dsig_factor, cplx *dbe) {
  const double L0 = 1.0e3;
  (void)nline;
  for (int m = 0; m < npar; m++) {
    for (int o = 0; o < nobs; o++) { dbe[(size_t)o * npar + m].re = 0; dbe[(size_t)o * npar + m].im = 0; }
    for (int q = par_ptr[m]; q < par_ptr[m + 1]; q++) {
      int t = par_tet[q];
      int dir[6], ed[6];
      double x4[4][3], S[36], Y[36], v;
      for (int j = 0; j < 6; j++) {
        dir[j] = (n6line[t * 6 + j] < 0) ? -1 : 1;
        ed[j] = n6line[t * 6 + j] * dir[j] - 1;
      }
      for (int j = 0; j < 4; j++)
        for (int d = 0; d < 3; d++) x4[j][d] = xyz[(size_t)(n4[t * 4 + j] - 1) * 3 + d];
      if (orc_element(x4, dir, S, Y, NULL, &v) != 0) return -1;
      cplx iu = {0.0, 1.0};
      cplx dBB = rmul(omega, iu);
      dBB = rmul(DMU, dBB);
      dBB = rmul(sigma[t], dBB);
      dBB = rmul(L0 * L0, dBB);
      dBB = rmul(dsig_factor, dBB);
      cplx aal[6];
      for (int i = 0; i < 6; i++) {
        aal[i].re = 0; aal[i].im = 0;
        for (int j = 0; j < 6; j++) aal[i] = cadd(aal[i], cmul(rmul(Y[i * 6 + j], dBB), x[ed[j]]));
      }
      for (int o = 0; o < nobs; o++) {
        cplx s = dbe[(size_t)o * npar + m];
        for (int j = 0; j < 6; j++) {
          cplx na = {-aal[j].re, -aal[j].im};
          s = cadd(s, cmul(u[(size_t)o * nline + ed[j]], na));
        }
        dbe[(size_t)o * npar + m] = s;
      }
    }
  }
  return 0;
}

/* ---------------- CPU baseline: the parallel phases of one COCG13_single iteration ----------------
 * For timing only (the CPU baseline's all-cores figure): one iteration's SpMV (:273-288), its
 * p^T q (:293-296), the x, r update with ||r||^2 (:304-313), r^T z (:246-250) and the p update
 * (:258-268), each loop split over `nthreads` OpenMP threads (rows in static chunks, the
 * reductions combined per thread). The SSOR / D-ILU sweeps are sequential in the reference's row
 * order and are not part of it. Returns p^T q + ||r||^2 + r^T z (so nothing is optimised away). */
double orc_cocg_parallel_phases(int n, const cplx *D, const int *INU, const int *IAU, const cplx *AU,
                                const int *INL, const int *IAL, const cplx *AL, cplx *X, cplx *R, const cplx *Z,
                                cplx *P, cplx *Q, int nthreads) {
  double c1r = 0.0, rr = 0.0, rzr = 0.0;
  const cplx alpha = {1e-3, 2e-3}, beta = {0.5, -0.25};
#pragma omp parallel for num_threads(nthreads) schedule(static) reduction(+ : rzr)
  for (int i = 0; i < n; i++) {
    rzr += cmul(R[i], Z[i]).re;
    P[i] = cadd(Z[i], cmul(beta, P[i]));
  }
#pragma omp parallel for num_threads(nthreads) schedule(static) reduction(+ : c1r)
  for (int j = 0; j < n; j++) {
    cplx w = cmul(D[j], P[j]);
    for (long p = INU[j]; p < INU[j + 1]; p++) w = cadd(w, cmul(AU[p], P[IAU[p] - 1]));
    for (long p = INL[j]; p < INL[j + 1]; p++) w = cadd(w, cmul(AL[p], P[IAL[p] - 1]));
    Q[j] = w;
    c1r += cmul(P[j], w).re;
  }
#pragma omp parallel for num_threads(nthreads) schedule(static) reduction(+ : rr)
  for (int i = 0; i < n; i++) {
    X[i] = cadd(X[i], cmul(alpha, P[i]));
    R[i] = csub(R[i], cmul(alpha, Q[i]));
    rr += R[i].re * R[i].re + R[i].im * R[i].im;
  }
  return c1r + rr + rzr;
}
