// Host side of the C ABI declared in include/afx.h: pattern analysis (level sets, CSR
// layouts, assembly slot maps), device residency, and the solve drivers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <string>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>
#include <vector>

#include "afx_internal.h"

using afx::cx;

namespace {
thread_local std::string g_err;

int fail(int code, const char *msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(expr)                                                    \
  do {                                                                  \
    hipError_t e_ = (expr);                                             \
    if (e_ != hipSuccess) {                                             \
      g_err = std::string(#expr) + ": " + hipGetErrorString(e_);        \
      return AFX_ERR_HIP;                                               \
    }                                                                   \
  } while (0)

template <typename T>
int dalloc(T **p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  hipError_t e = hipMalloc((void **)p, count * sizeof(T));
  if (e != hipSuccess) {
    g_err = std::string("hipMalloc: ") + hipGetErrorString(e);
    return AFX_ERR_HIP;
  }
  return AFX_OK;
}
template <typename T>
void dfree(T *&p) {
  if (p) (void)hipFree((void *)p);
  p = nullptr;
}
template <typename T>
int upload(T **p, const T *h, size_t count) {
  int rc = dalloc(p, count);
  if (rc) return rc;
  if (count) HIPCHK(hipMemcpy(*p, h, count * sizeof(T), hipMemcpyHostToDevice));
  return AFX_OK;
}

int next_pow2(int k) {
  int p = 1;
  while (p < k) p <<= 1;
  return p;
}
}  // namespace

struct DevSide {
  int *bstep = nullptr, *spos = nullptr, *row = nullptr, *ptr = nullptr, *col = nullptr, *map = nullptr;
  unsigned char *exp = nullptr, *expp = nullptr;
  int *rowx = nullptr;
  int maxsteps = 0;
  int npos = 0;  // positions (the dummy position of the single-column layout is npos)
  int *psrc = nullptr, *pmap = nullptr, *prow = nullptr;
  cx<double> *pval64 = nullptr;
  cx<float> *pval32 = nullptr;
  cx<double> *val64 = nullptr;
  cx<float> *val32 = nullptr;
  long long nnz = 0;
  int ew = 0;
  int *ecol = nullptr, *emap = nullptr, *epos = nullptr;
  int smax = 0;
  cx<double> *eval64 = nullptr, *dpos64 = nullptr;
  cx<float> *eval32 = nullptr, *dpos32 = nullptr;
};

struct afx_system_s {
  int n = 0;
  long long nU = 0, nL = 0, nnz = 0;
  int maxrow = 0;  // longest full-CSR row
  // right-hand-side block width per column count k (afx_autotune): 1 = the k-column kernel,
  // 0 = columns one at a time, -1 = not tuned (one at a time)
  std::vector<int> batch_multi = std::vector<int>(257, -1);
  std::vector<int> INU, IAU, INL, IAL;  // reference layout (IAU/IAL 1-based)
  int nlevL = 0, nlevU = 0;
  hipStream_t stream = nullptr;
  // device pattern
  int *d_INU = nullptr, *d_IAU = nullptr, *d_INL = nullptr, *d_IAL = nullptr;
  int *d_rowptr = nullptr, *d_col = nullptr, *d_map_full = nullptr;
  int *d_lvlL = nullptr, *d_rowL = nullptr, *d_Lptr = nullptr, *d_Lcol = nullptr, *d_mapL = nullptr;
  int *d_lvlU = nullptr, *d_rowU = nullptr, *d_Uptr = nullptr, *d_Ucol = nullptr, *d_mapU = nullptr;
  // canonical values (K, M): [0,n) D, [n, n+nU) AU, [n+nU, n+nU+nL) AL
  cx<double> *d_canon = nullptr;
  bool have_values = false;
  double omega = 1.0;  // effective entry K + i*omega*M
  // derived layouts
  cx<double> *d_val64 = nullptr, *d_Lval64 = nullptr, *d_Uval64 = nullptr;
  cx<float> *d_val32 = nullptr, *d_Lval32 = nullptr, *d_Uval32 = nullptr;
  bool derived_valid = false;
  // preconditioner
  int pc_kind = -1;
  double pc_sd = 1.0, pc_sig = 0.0;
  std::vector<double> pc_w;
  cx<double> *d_DD = nullptr;
  cx<float> *d_DD32 = nullptr;
  bool pc_valid = false;
  // mesh
  int node = 0, ntet = 0;
  std::vector<int> n6line;
  double *d_xyz = nullptr, *d_fxyz = nullptr;
  int *d_n4 = nullptr, *d_n6 = nullptr, *d_group = nullptr;
  cx<double> *d_vxyz = nullptr;
  std::vector<double> h_xyz;
  std::vector<int> h_n4;
  bool have_mesh = false;
  // bucket grid of the tets for point location (built on the first afx_prepobscoeff)
  afx::PointGrid pgrid{};
  int *d_bptr = nullptr, *d_btet = nullptr;
  bool have_pgrid = false;
  int *d_cptr = nullptr, *d_cidx = nullptr, *d_rptr = nullptr, *d_ridx = nullptr, *d_mirror = nullptr;
  bool have_slots = false;
  cx<double> *d_rhs_unit = nullptr;  // tidal RHS of the last assembly
  // pipelined sweeps
  int Gs = 0, rows_per_block = 0;
  int *d_rstart = nullptr;
  // stacked independent systems (afx_create_stacked): row offsets [nsys + 1]; each block's rows
  // lie in one system: d_bsys [Gs] its system, d_sblk [nsys + 1] the first block of each system
  int nsys = 1;
  std::vector<int> sys_off;
  int *d_sys_off = nullptr, *d_bsys = nullptr, *d_sblk = nullptr;
  // right-hand-side block width "stack": k columns solved as k stacked copies of this system
  // (each at its column's omega, per-system scalars); the copies' handle, the omegas its values
  // were taken at and the values version they copy (val_ver grows with every value change)
  afx_system stack = nullptr;
  std::vector<double> stack_w;
  long long stack_ver = -1, val_ver = 0;
  DevSide fw, bw, bjfw, bjbw;
  unsigned *d_bar = nullptr;
  int *d_err = nullptr;
  // launch configuration (autotuned)
  int lpr64 = 8, lpr32 = 8;
  // launch timing of the persistent COCG kernel since the last afx_kernel_timing reset
  long long prof_launches = 0, prof_iters = 0;
  double prof_ms = 0.0;
  int ncu = 0;
  // nodal gradient-space correction (afx_set_nodes, AFX_PC_GRAD): interior nodes, incidence,
  // end nodes of each edge, and diag(G^T M G) of the current values
  int nn = -1;
  int *d_nptr = nullptr, *d_nadj = nullptr, *d_en = nullptr;
  int *d_nadj16 = nullptr;  // the incidence lists at 16 slots per node (the first 16 of each)
  int *d_noff = nullptr;  // stacked systems: first interior node of each system [nsys + 1]
  double *d_mnd = nullptr;
  // node-block Jacobi (AFX_PC_NODEBLK, afx_set_nodes): blocks of consecutive rows whose edges
  // start at the same node (at most AFX_NB_W rows; longer runs are split), packed whole into
  // tiles of at most 256 rows; per row its place in its block (i | (size - 1) << 4), and the
  // block inverses as AFX_NB_W complex-bfloat16 coefficients per row
  int ntile = 0, nblk = 0;
  int *d_tptr = nullptr, *d_bstart = nullptr;  // tile bounds [ntile + 1], first row of each block [nblk]
  unsigned char *d_bm = nullptr;
  afx::nbcoef *d_bi = nullptr;
};

namespace {

// Level sets of a strict triangle given as per-row dependency lists.
// forward: level(i) = 1 + max level(j), j in L(i) (j < i); backward over U with j > i.
void level_sets(int n, const std::vector<int> &ptr, const std::vector<int> &colm1, bool forward,
                std::vector<int> &lvl_ptr, std::vector<int> &rows) {
  std::vector<int> level(n, 0);
  int maxl = 0;
  if (forward) {
    for (int i = 0; i < n; i++) {
      int l = 0;
      for (int q = ptr[i]; q < ptr[i + 1]; q++) l = std::max(l, level[colm1[q]] + 1);
      level[i] = l;
      maxl = std::max(maxl, l);
    }
  } else {
    for (int i = n - 1; i >= 0; i--) {
      int l = 0;
      for (int q = ptr[i]; q < ptr[i + 1]; q++) l = std::max(l, level[colm1[q]] + 1);
      level[i] = l;
      maxl = std::max(maxl, l);
    }
  }
  const int nlev = (n > 0) ? maxl + 1 : 0;
  lvl_ptr.assign(nlev + 1, 0);
  for (int i = 0; i < n; i++) lvl_ptr[level[i] + 1]++;
  for (int l = 0; l < nlev; l++) lvl_ptr[l + 1] += lvl_ptr[l];
  rows.assign(n, 0);
  std::vector<int> fill(lvl_ptr.begin(), lvl_ptr.end() - (nlev > 0 ? 1 : 0));
  for (int i = 0; i < n; i++) rows[fill[level[i]]++] = i;  // ascending row id inside a level
}

int validate_pattern(int n, const int *INU, const int *IAU, const int *INL, const int *IAL) {
  if (n <= 0) return fail(AFX_ERR_ARG, "n must be positive");
  if (INU[0] != 0 || INL[0] != 0) return fail(AFX_ERR_ARG, "INU(0)/INL(0) must be 0");
  for (int i = 0; i < n; i++) {
    if (INU[i + 1] < INU[i] || INL[i + 1] < INL[i]) return fail(AFX_ERR_ARG, "INU/INL not monotone");
    for (int q = INU[i]; q < INU[i + 1]; q++)
      if (IAU[q] <= i + 1 || IAU[q] > n) return fail(AFX_ERR_ARG, "IAU entry not strictly upper");
    for (int q = INL[i]; q < INL[i + 1]; q++)
      if (IAL[q] < 1 || IAL[q] >= i + 1) return fail(AFX_ERR_ARG, "IAL entry not strictly lower");
  }
  return AFX_OK;
}

int build_device_pattern(afx_system s) {
  const int n = s->n;
  const long long nU = s->nU, nL = s->nL;
  // full CSR rows: diagonal, upper, lower (the order of COCG13_single's products, :273-288)
  std::vector<int> rowptr(n + 1), col(s->nnz), map_full(s->nnz);
  long long pos = 0;
  for (int i = 0; i < n; i++) {
    rowptr[i] = (int)pos;
    col[pos] = i;
    map_full[pos++] = i;
    for (int q = s->INU[i]; q < s->INU[i + 1]; q++) {
      col[pos] = s->IAU[q] - 1;
      map_full[pos++] = n + q;
    }
    for (int q = s->INL[i]; q < s->INL[i + 1]; q++) {
      col[pos] = s->IAL[q] - 1;
      map_full[pos++] = (int)(n + nU + q);
    }
  }
  rowptr[n] = (int)pos;
  s->maxrow = 0;
  for (int i = 0; i < n; i++) s->maxrow = std::max(s->maxrow, rowptr[i + 1] - rowptr[i]);
  // sweeps
  std::vector<int> Lc(nL), Uc(nU);
  for (long long q = 0; q < nL; q++) Lc[q] = s->IAL[q] - 1;
  for (long long q = 0; q < nU; q++) Uc[q] = s->IAU[q] - 1;
  std::vector<int> lvlL, rowL, lvlU, rowU;
  level_sets(n, s->INL, Lc, true, lvlL, rowL);
  level_sets(n, s->INU, Uc, false, lvlU, rowU);
  s->nlevL = (int)lvlL.size() - 1;
  s->nlevU = (int)lvlU.size() - 1;
  std::vector<int> Lptr(n + 1), Lcol(nL), mapL(nL), Uptr(n + 1), Ucol(nU), mapU(nU);
  pos = 0;
  for (int p = 0; p < n; p++) {
    const int i = rowL[p];
    Lptr[p] = (int)pos;
    for (int q = s->INL[i]; q < s->INL[i + 1]; q++) {
      Lcol[pos] = Lc[q];
      mapL[pos++] = (int)(n + nU + q);
    }
  }
  Lptr[n] = (int)pos;
  pos = 0;
  for (int p = 0; p < n; p++) {
    const int i = rowU[p];
    Uptr[p] = (int)pos;
    for (int q = s->INU[i]; q < s->INU[i + 1]; q++) {
      Ucol[pos] = Uc[q];
      mapU[pos++] = n + q;
    }
  }
  Uptr[n] = (int)pos;
  int rc;
  if ((rc = upload(&s->d_INU, s->INU.data(), n + 1))) return rc;
  if ((rc = upload(&s->d_INL, s->INL.data(), n + 1))) return rc;
  if ((rc = upload(&s->d_IAU, s->IAU.data(), nU))) return rc;
  if ((rc = upload(&s->d_IAL, s->IAL.data(), nL))) return rc;
  if ((rc = upload(&s->d_rowptr, rowptr.data(), n + 1))) return rc;
  if ((rc = upload(&s->d_col, col.data(), s->nnz))) return rc;
  if ((rc = upload(&s->d_map_full, map_full.data(), s->nnz))) return rc;
  if ((rc = upload(&s->d_lvlL, lvlL.data(), lvlL.size()))) return rc;
  if ((rc = upload(&s->d_rowL, rowL.data(), n))) return rc;
  if ((rc = upload(&s->d_Lptr, Lptr.data(), n + 1))) return rc;
  if ((rc = upload(&s->d_Lcol, Lcol.data(), nL))) return rc;
  if ((rc = upload(&s->d_mapL, mapL.data(), nL))) return rc;
  if ((rc = upload(&s->d_lvlU, lvlU.data(), lvlU.size()))) return rc;
  if ((rc = upload(&s->d_rowU, rowU.data(), n))) return rc;
  if ((rc = upload(&s->d_Uptr, Uptr.data(), n + 1))) return rc;
  if ((rc = upload(&s->d_Ucol, Ucol.data(), nU))) return rc;
  if ((rc = upload(&s->d_mapU, mapU.data(), nU))) return rc;
  const size_t ncanon = (size_t)n + nU + nL;
  if ((rc = dalloc(&s->d_canon, ncanon))) return rc;
  HIPCHK(hipMemset(s->d_canon, 0, ncanon * sizeof(cx<double>)));
  if ((rc = dalloc(&s->d_val64, s->nnz))) return rc;
  if ((rc = dalloc(&s->d_val32, s->nnz))) return rc;
  if ((rc = dalloc(&s->d_Lval64, nL))) return rc;
  if ((rc = dalloc(&s->d_Lval32, nL))) return rc;
  if ((rc = dalloc(&s->d_Uval64, nU))) return rc;
  if ((rc = dalloc(&s->d_Uval32, nU))) return rc;
  // mirror map for COPY_UL: lower (j, i) <- upper (i, j)
  std::vector<int> mirror(nL, -1);
  for (int i = 0; i < n; i++)
    for (int q = s->INU[i]; q < s->INU[i + 1]; q++) {
      const int j = s->IAU[q] - 1;
      const int *b = &s->IAL[s->INL[j]], *e = &s->IAL[s->INL[j + 1]];
      const int *f = std::lower_bound(b, e, i + 1);
      if (f == e || *f != i + 1) return fail(AFX_ERR_PATTERN, "pattern is not structurally symmetric");
      mirror[s->INL[j] + (f - b)] = q;
    }
  for (long long q = 0; q < nL; q++)
    if (mirror[q] < 0) return fail(AFX_ERR_PATTERN, "lower entry without an upper mirror");
  if ((rc = upload(&s->d_mirror, mirror.data(), nL))) return rc;
  // launch configuration
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  HIPCHK(hipDeviceGetAttribute(&s->ncu, hipDeviceAttributeMultiprocessorCount, dev));
  return AFX_OK;
}

int ensure_derived(afx_system s) {
  if (s->derived_valid) return AFX_OK;
  HIPCHK(afx::launch_gather_vals((int)s->nnz, s->d_map_full, s->d_canon, s->d_val64, s->d_val32, s->stream));
  HIPCHK(afx::launch_gather_vals((int)s->nL, s->d_mapL, s->d_canon, s->d_Lval64, s->d_Lval32, s->stream));
  HIPCHK(afx::launch_gather_vals((int)s->nU, s->d_mapU, s->d_canon, s->d_Uval64, s->d_Uval32, s->stream));
  HIPCHK(afx::launch_gather_vals((int)s->fw.nnz, s->fw.map, s->d_canon, s->fw.val64, s->fw.val32, s->stream));
  HIPCHK(afx::launch_gather_vals((int)s->bw.nnz, s->bw.map, s->d_canon, s->bw.val64, s->bw.val32, s->stream));
  HIPCHK(afx::launch_gather_vals(s->n * s->fw.ew, s->fw.emap, s->d_canon, s->fw.eval64, s->fw.eval32, s->stream));
  HIPCHK(afx::launch_gather_vals(s->n * s->bw.ew, s->bw.emap, s->d_canon, s->bw.eval64, s->bw.eval32, s->stream));
  for (DevSide *d : {&s->bjfw, &s->bjbw}) {
    HIPCHK(afx::launch_gather_vals((int)d->nnz, d->map, s->d_canon, d->val64, d->val32, s->stream));
    HIPCHK(afx::launch_gather_vals(s->n * d->ew, d->emap, s->d_canon, d->eval64, d->eval32, s->stream));
  }
  for (DevSide *d : {&s->fw, &s->bw, &s->bjfw, &s->bjbw})
    if (d->pmap)
      HIPCHK(afx::launch_gather_vals((s->n + 1) * AFX_PIPE_W, d->pmap, s->d_canon, d->pval64, d->pval32, s->stream));
  s->derived_valid = true;
  return AFX_OK;
}

// Preconditioner diagonal for the distinct column frequencies `wf`.
int ensure_precond(afx_system s, int kind_full, double sd, double sig, const std::vector<double> &wf) {
  if (s->pc_valid && s->pc_kind == kind_full && s->pc_sd == sd && s->pc_sig == sig && s->pc_w == wf) return AFX_OK;
  const bool grad = (kind_full & AFX_PC_GRAD) != 0;
  const int kind = kind_full & 0xff;
  if (grad && (kind != AFX_PC_SSOR && kind != AFX_PC_SSOR_BJ && kind != AFX_PC_DIAG && kind != AFX_PC_NODEBLK))
    return fail(AFX_ERR_BADOPT, "AFX_PC_GRAD goes with SSOR, SSOR_BJ, DIAG or NODEBLK");
  if (grad && s->nn < 0) return fail(AFX_ERR_ARG, "AFX_PC_GRAD needs afx_set_nodes");
  if (kind == AFX_PC_NODEBLK && s->ntile <= 0) return fail(AFX_ERR_ARG, "AFX_PC_NODEBLK needs afx_set_nodes");
  int rc = ensure_derived(s);
  if (rc) return rc;
  if (grad)
    HIPCHK(afx::launch_nodal_diag(s->nn, s->d_nptr, s->d_nadj, s->d_en, s->d_rowptr, s->d_col, s->d_val64, s->d_mnd,
                                  s->stream));
  const int nf = (int)wf.size();
  dfree(s->d_DD);
  dfree(s->d_DD32);
  if ((rc = dalloc(&s->d_DD, (size_t)s->n * nf))) return rc;
  if ((rc = dalloc(&s->d_DD32, (size_t)s->n * nf))) return rc;
  HIPCHK(hipMemsetAsync(s->d_DD, 0, (size_t)s->n * nf * sizeof(cx<double>), s->stream));
  double *d_wf = nullptr;
  if ((rc = upload(&d_wf, wf.data(), nf))) return rc;
  int grid = 0;
  if (kind == AFX_PC_IC) {
    grid = std::min(afx::max_coop_grid_ic(), std::max(8, s->ncu / 8));
    if (grid <= 0) return fail(AFX_ERR_HIP, "no cooperative grid for the D-ILU setup");
  }
  // node blocks: the diagonal as for DIAG (the k-column and parity paths read it) and the block
  // inverses at the one frequency
  const int dkind = (kind == AFX_PC_NODEBLK) ? AFX_PC_DIAG : kind;
  hipError_t e = afx::launch_precond_setup(s->n, nf, dkind, sd, sig, s->d_canon, nullptr, nullptr, nullptr, s->nlevL,
                                           s->d_lvlL, s->d_rowL, s->d_Lptr, s->d_Lcol, s->d_Lval64, s->d_INU,
                                           s->d_canon + s->n, d_wf, s->d_DD, grid, s->stream);
  if (e == hipSuccess && kind == AFX_PC_NODEBLK) {
    // one slice of coefficients per distinct frequency
    dfree(s->d_bi);
    if (dalloc(&s->d_bi, (size_t)s->n * AFX_NB_W * nf) != AFX_OK) e = hipErrorOutOfMemory;
    for (int f = 0; f < nf && e == hipSuccess; f++)
      e = afx::launch_nodeblk_setup(s->nblk, s->d_bstart, s->d_bm, s->d_rowptr, s->d_col, s->d_val64, wf[f],
                                    s->d_bi + (size_t)f * s->n * AFX_NB_W, s->stream);
  }
  if (e == hipSuccess) e = afx::launch_cast_vec((long long)s->n * nf, s->d_DD, s->d_DD32, s->stream);
  if (e == hipSuccess && nf == 1) e = afx::launch_dpos(s->n, s->fw.row, s->d_DD, s->fw.dpos64, s->fw.dpos32, s->stream);
  if (e == hipSuccess && nf == 1) e = afx::launch_dpos(s->n, s->bw.row, s->d_DD, s->bw.dpos64, s->bw.dpos32, s->stream);
  if (e == hipSuccess && nf == 1)
    e = afx::launch_dpos(s->n, s->bjfw.row, s->d_DD, s->bjfw.dpos64, s->bjfw.dpos32, s->stream);
  if (e == hipSuccess && nf == 1)
    e = afx::launch_dpos(s->n, s->bjbw.row, s->d_DD, s->bjbw.dpos64, s->bjbw.dpos32, s->stream);
  hipError_t e2 = hipStreamSynchronize(s->stream);
  dfree(d_wf);
  if (e != hipSuccess || e2 != hipSuccess) {
    g_err = std::string("preconditioner setup: ") + hipGetErrorString(e != hipSuccess ? e : e2);
    return AFX_ERR_HIP;
  }
  s->pc_kind = kind_full;
  s->pc_sd = sd;
  s->pc_sig = sig;
  s->pc_w = wf;
  s->pc_valid = true;
  return AFX_OK;
}

void invalidate_values(afx_system s) {
  s->derived_valid = false;
  s->pc_valid = false;
  s->val_ver++;
}

// column frequencies -> distinct list + map
void column_freqs(afx_system s, int k, const double *scale, std::vector<double> &wcol, std::vector<double> &wf,
                  std::vector<int> &col2f) {
  wcol.resize(k);
  col2f.resize(k);
  wf.clear();
  for (int c = 0; c < k; c++) {
    const double w = s->omega * (scale ? scale[c] : 1.0);
    wcol[c] = w;
    int f = -1;
    for (size_t q = 0; q < wf.size(); q++)
      if (wf[q] == w) f = (int)q;
    if (f < 0) {
      f = (int)wf.size();
      wf.push_back(w);
    }
    col2f[c] = f;
  }
}

// Pipelined-sweep layout of one triangle (see afx_solver.hip): rows of each block sorted by
// dependency level into steps, and the export flag of every row another block reads.
struct SideHost {
  std::vector<int> bstep, spos, row, ptr, col, map;
  std::vector<unsigned char> exp, expp;
  std::vector<int> rowx;  // row | (exported << 31) by position
  int maxsteps = 0;
  // single-column sweep layout, positions 0..n-1 plus a dummy position n (see SweepSide)
  bool pipe = false;
  std::vector<int> psrc, pmap, prow;
  // fixed-width (padded) entries; epos: position of a same-block column row, else -1
  int ew = 0, smax = 0;
  std::vector<int> ecol, emap, epos;
};

void build_side(int n, int G, const std::vector<int> &rstart, const std::vector<int> &blk,
                const std::vector<int> &P0, const std::vector<int> &C0, long long canon_off, bool fwd, SideHost &h,
                bool local_only = false) {
  // local_only: drop the couplings between blocks (block-Jacobi), keeping canonical indices
  std::vector<int> P, C1, qmap;
  if (local_only) {
    P.assign(n + 1, 0);
    for (int i = 0; i < n; i++) {
      for (int q = P0[i]; q < P0[i + 1]; q++)
        if (blk[C0[q] - 1] == blk[i]) {
          C1.push_back(C0[q]);
          qmap.push_back(q);
        }
      P[i + 1] = (int)C1.size();
    }
  } else {
    P = P0;
    C1 = C0;
    qmap.resize(C0.size());
    for (size_t q = 0; q < C0.size(); q++) qmap[q] = (int)q;
  }
  // Schedule key of each row: the longest dependency chain ending at it, where a dependency on
  // a row of the same block counts 1 and one on another block's row counts `rw` (a mailbox hop
  // costs several block steps). Rows of a block run in increasing key, so a row whose input
  // comes from another block is placed after the block's independent work instead of stalling
  // the step it would share with that work. Keys increase along every dependency, so rows of
  // one key are independent and the smallest pending key of any block can always proceed.
  int rw = 1;
  if (const char *e = getenv("AFX_HOP_WEIGHT")) rw = std::max(1, std::min(atoi(e), 64));
  std::vector<int> lev(n, 0);
  auto wdep = [&](int i, int j) { return blk[i] == blk[j] ? 1 : rw; };
  if (fwd) {
    for (int i = 0; i < n; i++)
      for (int q = P[i]; q < P[i + 1]; q++) lev[i] = std::max(lev[i], lev[C1[q] - 1] + wdep(i, C1[q] - 1));
  } else {
    for (int i = n - 1; i >= 0; i--)
      for (int q = P[i]; q < P[i + 1]; q++) lev[i] = std::max(lev[i], lev[C1[q] - 1] + wdep(i, C1[q] - 1));
  }
  h.exp.assign(n, 0);
  for (int i = 0; i < n; i++)
    for (int q = P[i]; q < P[i + 1]; q++) {
      const int j = C1[q] - 1;
      if (blk[j] != blk[i]) h.exp[j] = 1;
    }
  std::vector<int> lstep(n, 0);
  h.bstep.assign(G + 1, 0);
  h.spos.clear();
  h.row.clear();
  h.row.reserve(n);
  std::vector<int> step_blk;
  for (int b = 0; b < G; b++) {
    h.bstep[b] = (int)h.spos.size();
    std::vector<int> rows;
    for (int i = rstart[b]; i < rstart[b + 1]; i++) rows.push_back(i);
    std::stable_sort(rows.begin(), rows.end(), [&](int x, int y) { return lev[x] < lev[y]; });
    // a step holds rows of one level, at most AFX_STEP_ROWS of them (the single-column sweep
    // gives each row of a step its own 8-lane group of the block)
    int prev = -1, ls = -1, cnt = 0;
    for (int i : rows) {
      if (lev[i] != prev || cnt == AFX_STEP_ROWS) {
        cnt = 0;
        h.spos.push_back((int)h.row.size());
        step_blk.push_back(b);
        prev = lev[i];
        ++ls;
      }
      lstep[i] = ls;
      ++cnt;
      h.row.push_back(i);
    }
    h.maxsteps = std::max(h.maxsteps, ls + 1);
  }
  const int nstep = (int)h.spos.size();
  h.bstep[G] = nstep;
  h.spos.push_back(n);
  h.ptr.assign(n + 1, 0);
  h.col.clear();
  h.map.clear();
  for (int pos = 0; pos < n; pos++) {
    const int i = h.row[pos];
    h.ptr[pos] = (int)h.col.size();
    for (int q = P[i]; q < P[i + 1]; q++) {
      h.col.push_back(C1[q] - 1);
      h.map.push_back((int)(canon_off + qmap[q]));
    }
  }
  h.ptr[n] = (int)h.col.size();
  h.expp.resize(n);
  for (int pos = 0; pos < n; pos++) h.expp[pos] = h.exp[h.row[pos]];
  h.rowx.resize(n);
  for (int pos = 0; pos < n; pos++) h.rowx[pos] = (int)((unsigned)h.row[pos] | ((unsigned)h.expp[pos] << 31));
  // even width (>= 2): a lane of the single-column sweep loads an aligned pair of entries
  h.ew = 0;
  for (int pos = 0; pos < n; pos++) h.ew = std::max(h.ew, h.ptr[pos + 1] - h.ptr[pos]);
  h.ew = std::max(2, (h.ew + 1) & ~1);
  h.ecol.assign((size_t)n * h.ew, -1);
  h.emap.assign((size_t)n * h.ew, -1);
  h.epos.assign((size_t)n * h.ew, -1);
  std::vector<int> rowpos(n);
  for (int pos = 0; pos < n; pos++) rowpos[h.row[pos]] = pos;
  h.smax = 0;
  for (int st = 0; st < nstep; st++) h.smax = std::max(h.smax, h.spos[st + 1] - h.spos[st]);
  for (int pos = 0; pos < n; pos++)
    for (int q = h.ptr[pos]; q < h.ptr[pos + 1]; q++) {
      h.ecol[(size_t)pos * h.ew + (q - h.ptr[pos])] = h.col[q];
      h.emap[(size_t)pos * h.ew + (q - h.ptr[pos])] = h.map[q];
      if (blk[h.col[q]] == blk[h.row[pos]]) h.epos[(size_t)pos * h.ew + (q - h.ptr[pos])] = rowpos[h.col[q]];
    }
  // Single-column sweep layout: AFX_PIPE_W entries per position, entry source src >= 0 the LDS
  // slot of a row of the same block, src < 0 the mailbox slot ~src of another block's row;
  // padding reads the zero dummy slot (index = largest block) with value 0. prow = {row |
  // exported << 31, LDS slot}. Position n is a dummy row (all padding, DD 0, not exported) that
  // fills the unused row slots of a step, so the kernel needs no validity tests.
  h.pipe = h.ew <= AFX_PIPE_W;
  if (h.pipe) {
    int dslot = 0;
    for (int b = 0; b < G; b++) dslot = std::max(dslot, rstart[b + 1] - rstart[b]);
    const size_t W = AFX_PIPE_W;
    h.psrc.assign((size_t)(n + 1) * W, dslot);
    h.pmap.assign((size_t)(n + 1) * W, -1);
    h.prow.assign((size_t)(n + 1) * 2, 0);
    for (int pos = 0; pos < n; pos++) {
      const int i = h.row[pos], bi = blk[i];
      h.prow[2 * (size_t)pos] = h.rowx[pos];
      h.prow[2 * (size_t)pos + 1] = i - rstart[bi];
      for (int q = h.ptr[pos]; q < h.ptr[pos + 1]; q++) {
        const int j = h.col[q];
        const size_t e = (size_t)pos * W + (q - h.ptr[pos]);
        h.psrc[e] = (blk[j] == bi) ? j - rstart[bi] : ~j;
        h.pmap[e] = h.map[q];
      }
    }
    h.prow[2 * (size_t)n] = 0;
    h.prow[2 * (size_t)n + 1] = dslot;
  }
}

int upload_side(const SideHost &h, DevSide &d) {
  int rc;
  if ((rc = upload(&d.bstep, h.bstep.data(), h.bstep.size())) || (rc = upload(&d.spos, h.spos.data(), h.spos.size())) ||
      (rc = upload(&d.row, h.row.data(), h.row.size())) || (rc = upload(&d.ptr, h.ptr.data(), h.ptr.size())) ||
      (rc = upload(&d.col, h.col.data(), h.col.size())) || (rc = upload(&d.map, h.map.data(), h.map.size())) ||
      (rc = upload(&d.exp, h.exp.data(), h.exp.size())) || (rc = upload(&d.expp, h.expp.data(), h.expp.size())) ||
      (rc = upload(&d.rowx, h.rowx.data(), h.rowx.size())) || (rc = dalloc(&d.val64, h.col.size())) ||
      (rc = dalloc(&d.val32, h.col.size())))
    return rc;
  d.nnz = (long long)h.col.size();
  d.ew = h.ew;
  d.maxsteps = h.maxsteps;
  d.npos = (int)h.row.size();
  const size_t ne = h.ecol.size();
  d.smax = h.smax;
  if ((rc = upload(&d.ecol, h.ecol.data(), ne)) || (rc = upload(&d.emap, h.emap.data(), ne)) ||
      (rc = upload(&d.epos, h.epos.data(), ne)) ||
      (rc = dalloc(&d.eval64, ne)) || (rc = dalloc(&d.eval32, ne)) || (rc = dalloc(&d.dpos64, h.row.size() + 1)) ||
      (rc = dalloc(&d.dpos32, h.row.size() + 1)))
    return rc;
  // the dummy position's DD stays zero
  HIPCHK(hipMemset(d.dpos64, 0, (h.row.size() + 1) * sizeof(cx<double>)));
  HIPCHK(hipMemset(d.dpos32, 0, (h.row.size() + 1) * sizeof(cx<float>)));
  if (h.pipe) {
    const size_t np = h.psrc.size();
    if ((rc = upload(&d.psrc, h.psrc.data(), np)) || (rc = upload(&d.pmap, h.pmap.data(), np)) ||
        (rc = upload(&d.prow, h.prow.data(), h.prow.size())) || (rc = dalloc(&d.pval64, np)) ||
        (rc = dalloc(&d.pval32, np)))
      return rc;
  }
  return AFX_OK;
}

void free_side(DevSide &d) {
  dfree(d.bstep); dfree(d.spos); dfree(d.row); dfree(d.ptr); dfree(d.col); dfree(d.map);
  dfree(d.exp); dfree(d.expp); dfree(d.rowx); dfree(d.val64); dfree(d.val32);
  dfree(d.ecol); dfree(d.emap); dfree(d.epos); dfree(d.eval64); dfree(d.eval32); dfree(d.dpos64); dfree(d.dpos32);
  dfree(d.psrc); dfree(d.pmap); dfree(d.prow); dfree(d.pval64); dfree(d.pval32);
}

afx::SweepSide side_view(const DevSide &d, bool fp64) {
  afx::SweepSide v{};
  v.bstep = d.bstep; v.spos = d.spos; v.row = d.row; v.ptr = d.ptr; v.col = d.col;
  v.val = fp64 ? (const void *)d.val64 : (const void *)d.val32;
  v.exp = d.exp;
  v.expp = d.expp;
  v.rowx = d.rowx;
  v.maxsteps = d.maxsteps;
  v.npos = d.npos;
  v.psrc = getenv("AFX_NO_PIPE") ? nullptr : d.psrc;
  v.prow = d.prow;
  v.pval = fp64 ? (const void *)d.pval64 : (const void *)d.pval32;
  // the fixed-width path needs at most 16 entries per position (the E * lanes of the kernel)
  const bool ell = d.ew <= 16 && !getenv("AFX_NO_ELL");
  v.ew = d.ew;
  v.epos = d.epos;
  v.smax = d.smax;
  v.ecol = ell ? d.ecol : nullptr;
  v.eval = fp64 ? (const void *)d.eval64 : (const void *)d.eval32;
  v.dpos = fp64 ? (const void *)d.dpos64 : (const void *)d.dpos32;
  return v;
}

int build_sweeps(afx_system s) {
  const int n = s->n;
  int G = std::max(1, std::min(s->ncu, (n + 63) / 64));
  if (const char *e = getenv("AFX_SWEEP_BLOCKS")) G = std::max(1, std::min(atoi(e), n));
  s->Gs = G;
  // the sweeps may use fewer row slabs than the launch has blocks (a slab boundary costs a
  // mailbox hand-off on the sweep's critical path); blocks past the last slab own no rows
  int Gp = G;
  if (const char *e = getenv("AFX_SWEEP_PARTS")) Gp = std::max(1, std::min(atoi(e), G));
  std::vector<int> rstart(G + 1), blk(n);
  for (int b = 0; b <= G; b++) rstart[b] = (int)((long long)std::min(b, Gp) * n / Gp);
  if (s->nsys > 1) {
    // stacked systems: every slab inside one system, ceil(rows * G / n) blocks per system (at
    // most its rows), then the largest counts lowered until they add up to at most G; blocks
    // left over own no rows and count with the last system
    const int m = s->nsys;
    if (m > G) return fail(AFX_ERR_ARG, ("more stacked systems than blocks (" + std::to_string(G) + ")").c_str());
    std::vector<int> nb(m);
    long long tot = 0;
    for (int q = 0; q < m; q++) {
      const long long rows = s->sys_off[q + 1] - s->sys_off[q];
      nb[q] = (int)std::max(1LL, std::min(rows, (rows * G + n - 1) / n));
      tot += nb[q];
    }
    while (tot > G) {
      int big = 0;
      for (int q = 1; q < m; q++)
        if (nb[q] > nb[big]) big = q;
      nb[big]--;
      tot--;
    }
    std::vector<int> bsys(G, m - 1), sblk(m + 1, G);
    int b = 0;
    for (int q = 0; q < m; q++) {
      sblk[q] = b;
      const long long o = s->sys_off[q], rows = s->sys_off[q + 1] - o;
      for (int j = 0; j < nb[q]; j++, b++) {
        rstart[b] = (int)(o + rows * j / nb[q]);
        bsys[b] = q;
      }
    }
    for (; b <= G; b++) rstart[b] = n;
    int rc;
    if ((rc = upload(&s->d_bsys, bsys.data(), bsys.size())) || (rc = upload(&s->d_sblk, sblk.data(), sblk.size())) ||
        (rc = upload(&s->d_sys_off, s->sys_off.data(), s->sys_off.size())))
      return rc;
  }
  for (int b = 0; b < G; b++)
    for (int i = rstart[b]; i < rstart[b + 1]; i++) blk[i] = b;
  SideHost fw, bw;
  build_side(n, G, rstart, blk, s->INL, s->IAL, (long long)n + s->nU, true, fw);
  build_side(n, G, rstart, blk, s->INU, s->IAU, n, false, bw);
  SideHost bjfw, bjbw;
  build_side(n, G, rstart, blk, s->INL, s->IAL, (long long)n + s->nU, true, bjfw, true);
  build_side(n, G, rstart, blk, s->INU, s->IAU, n, false, bjbw, true);
  s->rows_per_block = 0;
  for (int b = 0; b < G; b++) s->rows_per_block = std::max(s->rows_per_block, rstart[b + 1] - rstart[b]);
  int rc;
  if ((rc = upload(&s->d_rstart, rstart.data(), rstart.size()))) return rc;
  if ((rc = upload_side(fw, s->fw)) || (rc = upload_side(bw, s->bw)) || (rc = upload_side(bjfw, s->bjfw)) ||
      (rc = upload_side(bjbw, s->bjbw)))
    return rc;
  if ((rc = dalloc(&s->d_bar, AFX_BAR_WORDS)) || (rc = dalloc(&s->d_err, 1))) return rc;
  HIPCHK(hipMemset(s->d_bar, 0, AFX_BAR_WORDS * sizeof(unsigned)));
  HIPCHK(hipMemset(s->d_err, 0, sizeof(int)));
  return AFX_OK;
}

// Workspace of one Krylov solve; TP is the precision of the preconditioner vector z.
template <typename TP>
struct Work {
  cx<double> *R = nullptr, *P = nullptr, *Q = nullptr;
  cx<TP> *Z = nullptr, *mbf = nullptr, *mbb = nullptr;
  cx<double> *part_rho = nullptr, *part_c1 = nullptr, *fin_rho = nullptr, *fin_c1 = nullptr, *rho1 = nullptr;
  double *part_rr = nullptr, *fin_rr = nullptr, *bnrm2 = nullptr, *resid = nullptr, *norms = nullptr;
  double *part_a = nullptr, *part_b = nullptr, *wcol = nullptr, *part4 = nullptr;
  int *conv = nullptr, *iters = nullptr, *status = nullptr, *col2f = nullptr, *run = nullptr;
  int *h_status = nullptr;
  cx<double> *Zn = nullptr, *Zb = nullptr;
  ~Work() {
    dfree(R); dfree(Z); dfree(P); dfree(Q); dfree(mbf); dfree(mbb); dfree(Zn); dfree(Zb);
    dfree(part_rho); dfree(part_c1); dfree(fin_rho); dfree(fin_c1); dfree(rho1);
    dfree(part_rr); dfree(fin_rr); dfree(bnrm2); dfree(resid); dfree(norms); dfree(part_a); dfree(part_b);
    dfree(part4);
    dfree(conv); dfree(iters); dfree(status); dfree(col2f); dfree(wcol); dfree(run);
    if (h_status) (void)hipHostFree(h_status);
  }
  // ks: entries of the per-column state (the number of stacked systems for a stacked solve)
  int alloc(long long n, int k, int G, int ks) {
    int rc;
    const size_t nk = (size_t)n * k;
    if ((rc = dalloc(&R, nk)) || (rc = dalloc(&Z, nk)) || (rc = dalloc(&P, nk)) || (rc = dalloc(&Q, nk))) return rc;
    const size_t gk = (size_t)G * k;
    // part_rr holds complex partials in the k-column kernel without sweeps: twice the doubles
    if ((rc = dalloc(&part_rho, gk)) || (rc = dalloc(&part_c1, gk)) || (rc = dalloc(&part_rr, 2 * gk))) return rc;
    if ((rc = dalloc(&part_a, gk)) || (rc = dalloc(&part_b, gk)) || (rc = dalloc(&part4, 5 * (size_t)G))) return rc;
    if ((rc = dalloc(&fin_rho, k)) || (rc = dalloc(&fin_c1, k)) || (rc = dalloc(&fin_rr, k))) return rc;
    if ((rc = dalloc(&rho1, ks)) || (rc = dalloc(&bnrm2, ks)) || (rc = dalloc(&resid, ks))) return rc;
    if ((rc = dalloc(&norms, 2 * (size_t)ks)) || (rc = dalloc(&conv, ks)) || (rc = dalloc(&iters, ks))) return rc;
    if ((rc = dalloc(&status, 2)) || (rc = dalloc(&col2f, k)) || (rc = dalloc(&wcol, k))) return rc;
    if ((rc = dalloc(&run, G))) return rc;
    if (hipHostMalloc((void **)&h_status, 3 * sizeof(int), hipHostMallocDefault) != hipSuccess)
      return fail(AFX_ERR_HIP, "hipHostMalloc");
    return AFX_OK;
  }
};

__global__ void k_init_state(int k, const double *norms, double *bnrm2, int *conv, int *iters, double *resid,
                             cx<double> *rho1, int *status) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c == 0) {
    status[0] = 0;
    status[1] = 0;
  }
  if (c >= k) return;
  const double bb = norms[2 * c + 1];
  bnrm2[c] = bb;
  conv[c] = (bb == 0.0) ? 2 : 0;  // COCG13_single.f90:171-177: zero rhs -> X = 0, RESID = 0, ITER = 0
  iters[c] = 0;
  resid[c] = 0.0;
  rho1[c] = cx<double>{0.0, 0.0};
}

// per stacked system s (one block each): norms[2s] = |R_s|^2, norms[2s+1] = |B_s|^2 over rows
// [off[s], off[s+1]), in a fixed order
__global__ void __launch_bounds__(256) k_seg_norms(int m, const int *off, const cx<double> *R, const cx<double> *B,
                                                   double *norms) {
  __shared__ double lr[256], lb[256];
  const int q = blockIdx.x, t = threadIdx.x;
  if (q >= m) return;
  double rr = 0.0, bb = 0.0;
  for (long long i = off[q] + t; i < off[q + 1]; i += 256) {
    const cx<double> r = R[i], b = B[i];
    rr += r.r * r.r + r.i * r.i;
    bb += b.r * b.r + b.i * b.i;
  }
  lr[t] = rr;
  lb[t] = bb;
  __syncthreads();
  for (int h = 128; h >= 1; h >>= 1) {
    if (t < h) {
      lr[t] += lr[t + h];
      lb[t] += lb[t + h];
    }
    __syncthreads();
  }
  if (t == 0) {
    norms[2 * q] = lr[0];
    norms[2 * q + 1] = lb[0];
  }
}

// X = 0 on the rows of stacked systems whose right-hand side is zero (conv == 2)
__global__ void k_zero_sys(int m, const int *off, const int *conv, cx<double> *X) {
  const int q = blockIdx.x;
  if (q >= m || conv[q] != 2) return;
  for (long long i = off[q] + threadIdx.x; i < off[q + 1]; i += blockDim.x) X[i] = cx<double>{0.0, 0.0};
}

__global__ void k_zero_cols(long long n, int k, const int *conv, cx<double> *X) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n * k) return;
  if (conv[i % k] == 2) X[i] = cx<double>{0.0, 0.0};
}

template <typename TP>
int lds_rows_for(afx_system s, bool multi) {
  if (multi) return 0;
  int dev = 0, maxlds = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) return 0;
  // z of the block's rows, the staged step bounds of both sweep directions, the static LDS
  const int ms = std::max(s->fw.maxsteps + s->bw.maxsteps, s->bjfw.maxsteps + s->bjbw.maxsteps) + 2;
  const size_t need =
      (((size_t)(s->rows_per_block + 1) * sizeof(cx<TP>) + 15) & ~(size_t)15) + (size_t)ms * sizeof(int) + 6144;
  if (getenv("AFX_NO_LDS")) return 0;
  // one slot past the largest block: the zero slot that padding entries read
  return (need <= (size_t)maxlds) ? s->rows_per_block + 1 : 0;
}

// LDS ring (positions) of the several-column sweep: the largest power of two whose k columns
// fit in 128 KB, at least 5*smax + 64 positions (the kernel reads the ring only for sources less
// than W - 3*smax positions back, and loads older ones from global memory one step ahead, which
// needs them two steps back); 0 = not used.
template <typename TP>
int ring_w_for(afx_system s, int k, int pc) {
  pc &= 0xff;
  if (k <= 1 || getenv("AFX_NO_RING")) return 0;
  if (pc != AFX_PC_IC && pc != AFX_PC_SSOR && pc != AFX_PC_SSOR_BJ) return 0;
  const bool bj = (pc == AFX_PC_SSOR_BJ);
  const DevSide &fw = bj ? s->bjfw : s->fw, &bw = bj ? s->bjbw : s->bw;
  if (fw.ew > 16 || bw.ew > 16 || k > AFX_RING_KC) return 0;
  int W = 1;
  while ((size_t)(2 * W) * k * sizeof(cx<TP>) <= 131072) W *= 2;
  const int smax = std::max(fw.smax, bw.smax);
  return (W >= 5 * smax + 64) ? W : 0;
}

// One COCG solve on device vectors B (const) and X (in/out, the initial guess), [n][k] layout,
// with the preconditioner in precision TP.
template <typename TP>
int krylov(afx_system s, int k, const cx<double> *B, cx<double> *X, const std::vector<double> &wcol,
           const std::vector<int> &col2f, int nf, int pc, double tol, int maxit, int chunk, afx_stats *stats,
           int *host_conv, bool stacked = false) {
  const bool fp64 = std::is_same<TP, double>::value;
  const bool multi = k > 1;
  const bool grad = (pc & AFX_PC_GRAD) != 0;
  pc &= 0xff;
  if (grad && (s->nn < 0 || (stacked && s->nsys > 1 && !s->d_noff)))
    return fail(AFX_ERR_ARG, "AFX_PC_GRAD needs afx_set_nodes (on a stacked system, nodes that do not couple two systems)");
  // stacked systems: one column, per-system scalars and convergence; stats and host_conv then
  // hold one entry per system
  const bool sys = stacked && s->nsys > 1;
  if (sys && k != 1) return fail(AFX_ERR_ARG, "stacked systems take one right-hand side");
  const int ks = sys ? s->nsys : k;
  const int kp = multi ? next_pow2(k) : 1;
  if (kp > 256) return fail(AFX_ERR_ARG, "at most 256 right-hand sides per launch");
  // the streaming kernel (no sweep: DIAG or NONE edge preconditioner, one column, FP64) runs
  // at several blocks per CU; every other case runs k_cocg on the sweep partition
  const bool nblk = (pc == AFX_PC_NODEBLK);
  if (nblk && (k > 8 || !fp64))
    return fail(AFX_ERR_ARG, "AFX_PC_NODEBLK runs FP64, at most 8 columns per launch");
  const bool stream = fp64 && !(stacked && s->nsys > 1) && (pc == AFX_PC_DIAG || pc == AFX_PC_NONE || nblk) &&
                      (!multi || k <= 8) && (nblk || !getenv("AFX_NO_STREAM"));
  int per_cu = 1;
  if (const char *e = getenv("AFX_STREAM_PER_CU")) per_cu = std::max(1, std::min(8, atoi(e)));
  // one column: the fused kernel (three grid synchronisations per iteration); AFX_FUSED=0 selects
  // the four-pass kernel
  const char *fe = getenv("AFX_FUSED");
  const bool fused = stream && !multi && (nblk || !(fe && fe[0] == '0'));
  const int G = fused ? afx::max_coop_grid_cocg1(s->lpr64)
                : stream ? afx::max_coop_grid_diag(s->lpr64, per_cu) : s->Gs;
  if (G <= 0) return fail(AFX_ERR_HIP, "no co-resident grid for the COCG kernel");
  Work<TP> w;
  int rc = w.alloc(s->n, k, std::max(G, s->Gs), ks);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(w.col2f, col2f.data(), k * sizeof(int), hipMemcpyHostToDevice, s->stream));
  HIPCHK(hipMemcpyAsync(w.wcol, wcol.data(), k * sizeof(double), hipMemcpyHostToDevice, s->stream));
  // the error flag of the bounded device waits starts clear; AFX_TEST_RAISE_ERR raises it before
  // the first launch instead (a test of the abort path: every block then leaves at its first grid
  // barrier and the solve returns AFX_ERR_HIP)
  {
    static const int raised = 4;
    if (getenv("AFX_TEST_RAISE_ERR"))
      HIPCHK(hipMemcpyAsync(s->d_err, &raised, sizeof(int), hipMemcpyHostToDevice, s->stream));
    else
      HIPCHK(hipMemsetAsync(s->d_err, 0, sizeof(int), s->stream));
  }
  // initial residual r0 = b - A x and ||b||^2 (COCG13_single.f90:139-169)
  afx::SpmvParams sp{};
  sp.n = s->n;
  sp.k = k;
  sp.kp = kp;
  sp.rowptr = s->d_rowptr;
  sp.col = s->d_col;
  sp.val = s->d_val64;
  sp.wcol = w.wcol;
  sp.X = X;
  sp.B = B;
  sp.Y = w.R;
  sp.part_rr = w.part_a;
  sp.part_bb = w.part_b;
  HIPCHK(afx::launch_residual_norms(sp, G, w.norms, s->stream));
  if (sys) {
    hipLaunchKernelGGL(k_seg_norms, dim3(ks), dim3(256), 0, s->stream, ks, s->d_sys_off, w.R, B, w.norms);
    HIPCHK(hipGetLastError());
  }
  hipLaunchKernelGGL(k_init_state, dim3((ks + 255) / 256), dim3(256), 0, s->stream, ks, w.norms, w.bnrm2, w.conv,
                     w.iters, w.resid, w.rho1, w.status);
  HIPCHK(hipGetLastError());
  if (sys)
    hipLaunchKernelGGL(k_zero_sys, dim3(ks), dim3(256), 0, s->stream, ks, s->d_sys_off, w.conv, X);
  else
    hipLaunchKernelGGL(k_zero_cols, dim3((unsigned)(((long long)s->n * k + 255) / 256)), dim3(256), 0, s->stream,
                       (long long)s->n, k, w.conv, X);
  HIPCHK(hipGetLastError());
  afx::CocgParams<TP> p{};
  p.n = s->n;
  p.k = k;
  p.kp = kp;
  p.nf = nf;
  p.pc = pc;
  p.dist_reduce = (!sys && (long long)G * k > 4096) ? 1 : 0;
  p.nsys = sys ? ks : 1;
  p.bsys = sys ? s->d_bsys : nullptr;
  p.sblk = sys ? s->d_sblk : nullptr;
  p.run = w.run;
  p.rowptr = s->d_rowptr;
  p.col = s->d_col;
  p.val = s->d_val64;
  p.Gs = G;
  p.lds_rows = lds_rows_for<TP>(s, multi);
  p.ring_w = ring_w_for<TP>(s, k, pc);
  p.rstart = s->d_rstart;
  p.grad = grad ? 1 : 0;
  p.nn = grad ? s->nn : 0;
  p.nptr = s->d_nptr;
  p.nadj = s->d_nadj;
  p.nadj16 = s->d_nadj16;
  p.en = s->d_en;
  p.noff = s->d_noff;
  p.mnd = s->d_mnd;
  if (grad) {  // [nn + 1][k]: the last row (boundary nodes) stays zero
    if ((rc = dalloc(&w.Zn, ((size_t)s->nn + 1) * k))) return rc;
    HIPCHK(hipMemsetAsync(w.Zn, 0, ((size_t)s->nn + 1) * k * sizeof(cx<double>), s->stream));
  }
  p.Zn = w.Zn;
  p.part4 = w.part4;
  if (nblk && (rc = dalloc(&w.Zb, (size_t)s->n))) return rc;
  p.ntile = s->ntile;
  p.tptr = s->d_tptr;
  p.bm = s->d_bm;
  p.bi = s->d_bi;
  p.Zb = w.Zb;
  const bool bj = (pc == AFX_PC_SSOR_BJ);
  p.fw = side_view(bj ? s->bjfw : s->fw, fp64);
  p.bw = side_view(bj ? s->bjbw : s->bw, fp64);
  // sweep mailboxes: two sets per direction (iteration it uses set it & 1 and empties the other
  // for the next iteration), every slot empty at the start of the solve; one row more than n:
  // the single-column sweep's unused loads may read slot n
  if (pc == AFX_PC_IC || pc == AFX_PC_SSOR || pc == AFX_PC_SSOR_BJ) {
    const size_t slots = ((size_t)s->n + 1) * k;
    if ((rc = dalloc(&w.mbf, 2 * slots)) || (rc = dalloc(&w.mbb, 2 * slots))) return rc;
    const long long words = 2 * (long long)slots * (long long)sizeof(cx<TP>) / 8;
    HIPCHK(afx::launch_mb_fill(words, (unsigned long long *)w.mbf, s->stream));
    HIPCHK(afx::launch_mb_fill(words, (unsigned long long *)w.mbb, s->stream));
    p.mbf2 = w.mbf + slots;
    p.mbb2 = w.mbb + slots;
  } else {
    p.mbf2 = nullptr;
    p.mbb2 = nullptr;
  }
  p.mbf = w.mbf;
  p.mbb = w.mbb;
  p.DD = fp64 ? (const cx<TP> *)s->d_DD : (const cx<TP> *)s->d_DD32;
  p.col2f = w.col2f;
  p.wcol = w.wcol;
  p.B = B;
  p.X = X;
  p.R = w.R;
  p.Z = w.Z;
  p.P = w.P;
  p.Q = w.Q;
  p.part_rho = w.part_rho;
  p.part_c1 = w.part_c1;
  p.part_rr = w.part_rr;
  p.fin_rho = w.fin_rho;
  p.fin_c1 = w.fin_c1;
  p.fin_rr = w.fin_rr;
  p.rho1 = w.rho1;
  p.bnrm2 = w.bnrm2;
  p.conv = w.conv;
  p.iters = w.iters;
  p.resid = w.resid;
  p.status = w.status;
  p.bar.cnt = s->d_bar;
  p.bar.gen = s->d_bar + 32;
  p.bar.grp = s->d_bar + 64;
  // counters and generation start from zero for every solve (a launch that gave up on a wait
  // may have left them inconsistent)
  HIPCHK(hipMemsetAsync(s->d_bar, 0, AFX_BAR_WORDS * sizeof(unsigned), s->stream));
  p.bar.err = s->d_err;
  p.err = s->d_err;
  p.tol = tol;
  p.maxit = maxit;
  p.chunk = chunk > 0 ? chunk : 200;
  const int lpr = fp64 ? s->lpr64 : s->lpr32;
  const bool verbose = getenv("AFX_VERBOSE") != nullptr;
  // AFX_PROFILE: per-block phase timing of the persistent kernel, summarised on stderr
  unsigned long long *dbg = nullptr;
  struct DbgFree {
    unsigned long long *&q;
    ~DbgFree() { dfree(q); }
  } dbg_free{dbg};
  if (getenv("AFX_PROFILE")) {
    if ((rc = dalloc(&dbg, (size_t)G * AFX_DBG_SLOTS))) return rc;
    HIPCHK(hipMemsetAsync(dbg, 0, (size_t)G * AFX_DBG_SLOTS * sizeof(unsigned long long), s->stream));
  }
  p.dbg = dbg;
  const auto t_start = std::chrono::steady_clock::now();
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  HIPCHK(hipEventCreate(&ev0));
  HIPCHK(hipEventCreate(&ev1));
  int it_before = 0;
  for (;;) {
    HIPCHK(hipEventRecord(ev0, s->stream));
    hipError_t e = fused ? afx::launch_cocg1(*(afx::CocgParams<double> *)&p, lpr, G, s->stream)
                   : (stream && multi) ? afx::launch_cocg_blk(*(afx::CocgParams<double> *)&p, G, s->stream)
                   : stream ? afx::launch_cocg_diag(*(afx::CocgParams<double> *)&p, lpr, G, s->stream)
                   : fp64 ? afx::launch_cocg64(*(afx::CocgParams<double> *)&p, lpr, G, s->stream)
                          : afx::launch_cocg32(*(afx::CocgParams<float> *)&p, lpr, G, s->stream);
    if (e != hipSuccess) {
      (void)hipEventDestroy(ev0);
      (void)hipEventDestroy(ev1);
      g_err = std::string("cooperative COCG launch: ") + hipGetErrorString(e);
      return AFX_ERR_HIP;
    }
    HIPCHK(hipEventRecord(ev1, s->stream));
    HIPCHK(hipMemcpyAsync(w.h_status, w.status, 2 * sizeof(int), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipMemcpyAsync(w.h_status + 2, s->d_err, sizeof(int), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, ev0, ev1));
    s->prof_launches += 1;
    s->prof_ms += ms;
    s->prof_iters += w.h_status[0] - it_before;
    it_before = w.h_status[0];
    if (verbose) {
      const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
      fprintf(stderr, "afx: %s k=%d G=%d lpr=%d lds=%d ring=%d it=%d done=%d %.3fs (%.1f us/it)\n",
              fused ? "fp64-fused" : stream ? "fp64-stream" : fp64 ? "fp64" : "fp32-pc", k, G, lpr, p.lds_rows, p.ring_w, w.h_status[0],
              w.h_status[1], dt,
              w.h_status[0] ? 1e6 * dt / w.h_status[0] : 0.0);
    }
    // a wait that gave up (or a raised flag) ends the solve: results are not trusted
    if (w.h_status[2] || w.h_status[1] || w.h_status[0] >= maxit) break;
  }
  (void)hipEventDestroy(ev0);
  (void)hipEventDestroy(ev1);
  if (dbg) {
    std::vector<unsigned long long> h((size_t)G * AFX_DBG_SLOTS);
    HIPCHK(hipMemcpy(h.data(), dbg, h.size() * 8, hipMemcpyDeviceToHost));
    const double its = std::max(1, w.h_status[0]);
    double mean[11] = {0}, mx[11] = {0};
    for (int b = 0; b < G; b++)
      for (int q = 0; q < 11; q++) {
        const double v = (double)h[(size_t)b * AFX_DBG_SLOTS + q] / its * (q != 5 || (stream && multi) ? 0.01 : 1.0);  // 100 MHz -> us
        mean[q] += v / G;
        mx[q] = std::max(mx[q], v);
      }
    if (fused)
      fprintf(stderr,
              "afx profile fused (us/iteration, mean/max over %d blocks): p pass + barrier %.1f/%.1f  spmv + c1 sum "
              "%.1f/%.1f  x,r,z pass + sums %.1f/%.1f  total %.1f/%.1f\n",
              G, mean[0], mx[0], mean[2], mx[2], mean[3], mx[3], mean[4], mx[4]);
    else if (stream && multi)
      fprintf(stderr,
              "afx profile k-column (us/iteration, mean/max over %d blocks): z pass %.1f/%.1f  rho sum %.1f/%.1f  "
              "p update %.1f/%.1f  spmv %.1f/%.1f  c1 sum %.1f/%.1f  x,r pass %.1f/%.1f  rr sum %.1f/%.1f\n",
              G, mean[0], mx[0], mean[1], mx[1], mean[2], mx[2], mean[3], mx[3], mean[4], mx[4], mean[5], mx[5],
              mean[6], mx[6]);
    else
    fprintf(stderr,
            "afx profile (us/iteration, mean/max over %d blocks): fw sweep %.1f/%.1f  clear after fw %.1f/%.1f  "
            "bw sweep %.1f/%.1f  clear after bw %.1f/%.1f  rest %.1f/%.1f (rho wait %.1f/%.1f, after rho %.1f/%.1f: "
            "p update %.1f, spmv+c1 %.1f)  total %.1f/%.1f  mailbox polls/iteration/block %.1f/%.0f\n",
            G, mean[0], mx[0], mean[1], mx[1], mean[2], mx[2], mean[10], mx[10], mean[3], mx[3], mean[6], mx[6],
            mean[7], mx[7], mean[8], mean[9], mean[4], mx[4], mean[5], mx[5]);
  }
  int herr = 0;
  HIPCHK(hipMemcpy(&herr, s->d_err, sizeof(int), hipMemcpyDeviceToHost));
  if (herr) return fail(AFX_ERR_HIP, "a bounded device wait gave up (synchronisation error)");
  std::vector<int> conv(ks), iters(ks);
  std::vector<double> resid(ks);
  HIPCHK(hipMemcpy(conv.data(), w.conv, ks * sizeof(int), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(iters.data(), w.iters, ks * sizeof(int), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(resid.data(), w.resid, ks * sizeof(double), hipMemcpyDeviceToHost));
  int ret = AFX_OK;
  for (int c = 0; c < ks; c++) {
    int err = 0;
    if (conv[c] == -100) err = AFX_ERR_MAXIT;
    if (conv[c] == -200) err = AFX_ERR_BREAKDOWN;
    if (conv[c] == 0) {  // still running when maxit was reached (cannot normally happen)
      err = AFX_ERR_MAXIT;
      iters[c] = maxit + 1;
    }
    if (stats) {
      stats[c].iter = iters[c];
      stats[c].error = err;
      stats[c].resid = resid[c];
    }
    if (host_conv) host_conv[c] = conv[c];
    if (err == AFX_ERR_BREAKDOWN) ret = AFX_ERR_BREAKDOWN;
    else if (err && ret == AFX_OK) ret = err;
  }
  return ret;
}

}  // namespace

// =========================================================================================
// C ABI
// =========================================================================================
namespace {

int create_common(afx_system s) {
  s->nU = s->INU[s->n];
  s->nL = s->INL[s->n];
  s->nnz = (long long)s->n + s->nU + s->nL;
  if (s->nnz > 0x7fffffffLL) return fail(AFX_ERR_ARG, "more than 2^31 stored entries");
  HIPCHK(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
  int rc = build_device_pattern(s);
  if (rc) return rc;
  return build_sweeps(s);
}

void release(afx_system s) {
  dfree(s->d_INU); dfree(s->d_IAU); dfree(s->d_INL); dfree(s->d_IAL);
  dfree(s->d_rowptr); dfree(s->d_col); dfree(s->d_map_full);
  dfree(s->d_lvlL); dfree(s->d_rowL); dfree(s->d_Lptr); dfree(s->d_Lcol); dfree(s->d_mapL);
  dfree(s->d_lvlU); dfree(s->d_rowU); dfree(s->d_Uptr); dfree(s->d_Ucol); dfree(s->d_mapU);
  dfree(s->d_canon); dfree(s->d_val64); dfree(s->d_Lval64); dfree(s->d_Uval64);
  dfree(s->d_val32); dfree(s->d_Lval32); dfree(s->d_Uval32);
  dfree(s->d_DD); dfree(s->d_DD32);
  dfree(s->d_xyz); dfree(s->d_fxyz); dfree(s->d_n4); dfree(s->d_n6); dfree(s->d_group); dfree(s->d_vxyz);
  dfree(s->d_bptr); dfree(s->d_btet);
  dfree(s->d_cptr); dfree(s->d_cidx); dfree(s->d_rptr); dfree(s->d_ridx); dfree(s->d_mirror);
  dfree(s->d_rhs_unit);
  free_side(s->fw);
  free_side(s->bw);
  free_side(s->bjfw);
  free_side(s->bjbw);
  dfree(s->d_rstart); dfree(s->d_bar); dfree(s->d_err);
  dfree(s->d_sys_off); dfree(s->d_bsys); dfree(s->d_sblk);
  dfree(s->d_nptr); dfree(s->d_nadj); dfree(s->d_nadj16); dfree(s->d_en); dfree(s->d_mnd); dfree(s->d_noff);
  dfree(s->d_tptr); dfree(s->d_bm); dfree(s->d_bi); dfree(s->d_bstart);
  if (s->stack) afx_destroy(s->stack);
  s->stack = nullptr;
  if (s->stream) (void)hipStreamDestroy(s->stream);
  s->stream = nullptr;
}

int pc_from_string(const char *pc) {
  // COCG13_single.f90:107,126,202,236 compare the leading characters
  if (!pc) return AFX_PC_NONE;
  if (strncmp(pc, "IC", 2) == 0 || strncmp(pc, "ILU", 3) == 0) return AFX_PC_IC;
  if (strncmp(pc, "SSOR", 4) == 0) return AFX_PC_SSOR;
  if (strncmp(pc, "DIAG", 4) == 0) return AFX_PC_DIAG;
  return AFX_PC_NONE;
}

// Assembly slot maps from n6line: canonical slot (D or AU) -> (entry, tet) contributions in
// ascending tet order, and edge -> (local edge, tet) for the right-hand side.
int build_slot_maps(afx_system s) {
  const int n = s->n, ntet = s->ntet;
  const long long nslot = (long long)n + s->nU;
  std::vector<int> slot_of((size_t)ntet * 21);
  std::vector<int> cnt(nslot + 1, 0);
  for (int t = 0; t < ntet; t++) {
    int g[6];
    for (int j = 0; j < 6; j++) g[j] = std::abs(s->n6line[(size_t)t * 6 + j]);
    for (int a = 0; a < 6; a++)
      for (int b = 0; b <= a; b++) {
        const int e = (a == b) ? a : 6 + a * (a - 1) / 2 + b;
        long long slot;
        if (a == b) {
          slot = g[a] - 1;
        } else {
          const int lo = std::min(g[a], g[b]), hi = std::max(g[a], g[b]);
          if (lo == hi) return fail(AFX_ERR_PATTERN, "a tetrahedron repeats an edge");
          const int *bb = &s->IAU[s->INU[lo - 1]], *ee = &s->IAU[s->INU[lo]];
          const int *f = std::lower_bound(bb, ee, hi);
          if (f == ee || *f != hi) return fail(AFX_ERR_PATTERN, "sup_iccg: dof not found");
          slot = n + s->INU[lo - 1] + (f - bb);
        }
        slot_of[(size_t)t * 21 + e] = (int)slot;
        cnt[slot + 1]++;
      }
  }
  for (long long q = 0; q < nslot; q++) cnt[q + 1] += cnt[q];
  std::vector<int> cidx((size_t)ntet * 21), fill(cnt.begin(), cnt.end() - 1);
  for (int t = 0; t < ntet; t++)
    for (int e = 0; e < 21; e++) cidx[fill[slot_of[(size_t)t * 21 + e]]++] = e * ntet + t;
  std::vector<int> rcnt(n + 1, 0), ridx((size_t)ntet * 6);
  for (int t = 0; t < ntet; t++)
    for (int j = 0; j < 6; j++) rcnt[std::abs(s->n6line[(size_t)t * 6 + j])]++;
  for (int i = 0; i < n; i++) rcnt[i + 1] += rcnt[i];
  std::vector<int> rfill(rcnt.begin(), rcnt.end() - 1);
  for (int t = 0; t < ntet; t++)
    for (int j = 0; j < 6; j++) ridx[rfill[std::abs(s->n6line[(size_t)t * 6 + j]) - 1]++] = j * ntet + t;
  int rc;
  dfree(s->d_cptr); dfree(s->d_cidx); dfree(s->d_rptr); dfree(s->d_ridx);
  if ((rc = upload(&s->d_cptr, cnt.data(), cnt.size()))) return rc;
  if ((rc = upload(&s->d_cidx, cidx.data(), cidx.size()))) return rc;
  if ((rc = upload(&s->d_rptr, rcnt.data(), rcnt.size()))) return rc;
  if ((rc = upload(&s->d_ridx, ridx.data(), ridx.size()))) return rc;
  s->have_slots = true;
  return AFX_OK;
}

__global__ void k_inv_or_zero(int k, const double *norms, double tol, double *scale, double *inv) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= k) return;
  const double rr = norms[2 * c], bb = norms[2 * c + 1];
  const double nr = sqrt(rr);
  const bool done = (bb == 0.0) || (nr <= tol * sqrt(bb));
  scale[c] = done ? 0.0 : 1.0 / nr;
  inv[c] = done ? 0.0 : nr;
}

int solve_device_impl(afx_system s, int k, const cx<double> *dB, cx<double> *dX, const double *omega_scale,
                      const afx_opts *o, afx_stats *stats, bool stacked = false, int force_width = -1);

int ensure_stack(afx_system s, int k, const std::vector<double> &wcol);
int get_values_at(afx_system s, double w, double *D, double *AU, double *AL);

// Copies per stacked launch for k columns: as many as let a block's share of the copies' rows
// fit in LDS (about 150 KB of complex doubles per block), in groups of equal size; 0 when fewer
// than two copies fit.
int stack_group(afx_system s, int k) {
  const double per_copy = (double)s->n / std::max(1, s->ncu) * sizeof(cx<double>);
  int gmax = (int)std::min<double>(256.0, std::floor(150000.0 / std::max(per_copy, 1.0)));
  if (const char *e = getenv("AFX_STACK_GROUP")) gmax = std::min(gmax, std::max(2, atoi(e)));  // a smaller cap
  if (gmax < 2 || k < 2) return 0;
  const int groups = (k + gmax - 1) / gmax;
  return (k + groups - 1) / groups;
}

// Right-hand-side block width "stack": the k columns solved as independent systems stacked
// block-diagonally (copy c holds K + i*omega_c*M) in single-column launches with per-system
// scalars, stack_group(k) copies per launch; the last group is padded with zero right-hand
// sides (solved as x = 0 at once), so every launch reuses the kept copies' pattern.
int solve_stacked_columns(afx_system s, int k, const cx<double> *dB, cx<double> *dX, const double *omega_scale,
                          const afx_opts *o, afx_stats *stats) {
  std::vector<double> wcol, wf;
  std::vector<int> col2f;
  column_freqs(s, k, omega_scale, wcol, wf, col2f);
  int g = stack_group(s, k);
  if (g == 0) g = k;  // asked for explicitly (AFX_BATCH=stack): one launch, whatever fits in LDS
  const long long n = s->n;
  cx<double> *b1 = nullptr, *x1 = nullptr;
  int rc;
  if ((rc = dalloc(&b1, (size_t)n * g)) || (rc = dalloc(&x1, (size_t)n * g))) {
    dfree(b1);
    return rc;
  }
  int ret = AFX_OK;
  std::vector<afx_stats> st(g);
  for (int c0 = 0; c0 < k; c0 += g) {
    const int c1 = std::min(k, c0 + g);
    std::vector<double> wg(g, wcol[c1 - 1]);
    for (int c = c0; c < c1; c++) wg[c - c0] = wcol[c];
    if ((rc = ensure_stack(s, g, wg))) {
      ret = rc;
      break;
    }
    hipError_t e = hipMemsetAsync(b1, 0, (size_t)n * g * sizeof(cx<double>), s->stream);
    if (e == hipSuccess) e = hipMemsetAsync(x1, 0, (size_t)n * g * sizeof(cx<double>), s->stream);
    for (int c = c0; c < c1 && e == hipSuccess; c++) {
      e = afx::launch_col_get(n, k, c, dB, b1 + (size_t)(c - c0) * n, s->stream);
      if (e == hipSuccess) e = afx::launch_col_get(n, k, c, dX, x1 + (size_t)(c - c0) * n, s->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) {
      ret = fail(AFX_ERR_HIP, hipGetErrorString(e));
      break;
    }
    afx_opts o2 = *o;
    o2.nset = 1;
    const int r = solve_device_impl(s->stack, 1, b1, x1, nullptr, &o2, st.data(), true);
    if (r != AFX_OK && r != AFX_ERR_MAXIT && r != AFX_ERR_BREAKDOWN) {
      ret = r;
      break;
    }
    if (r == AFX_ERR_BREAKDOWN || (r == AFX_ERR_MAXIT && ret == AFX_OK)) ret = r;
    e = hipStreamSynchronize(s->stack->stream);
    for (int c = c0; c < c1 && e == hipSuccess; c++) {
      if (stats) stats[c] = st[c - c0];
      e = afx::launch_col_put(n, k, c, x1 + (size_t)(c - c0) * n, dX, s->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) {
      ret = fail(AFX_ERR_HIP, hipGetErrorString(e));
      break;
    }
  }
  dfree(b1);
  dfree(x1);
  return ret;
}

// Blocks of g columns (the last one smaller) solved one block after another, each by one k-column
// launch at the columns' own frequencies.
int solve_groups(afx_system s, int k, int g, const cx<double> *dB, cx<double> *dX, const double *omega_scale,
                 const afx_opts *o, afx_stats *stats) {
  cx<double> *bg = nullptr, *xg = nullptr;
  int rc;
  if ((rc = dalloc(&bg, (size_t)s->n * g)) || (rc = dalloc(&xg, (size_t)s->n * g))) {
    dfree(bg);
    return rc;
  }
  int ret = AFX_OK;
  std::vector<double> sc(g);
  for (int c0 = 0; c0 < k && rc == AFX_OK; c0 += g) {
    const int gg = std::min(g, k - c0);
    for (int j = 0; j < gg; j++) sc[j] = omega_scale ? omega_scale[c0 + j] : 1.0;
    hipError_t e = afx::launch_cols_get(s->n, k, c0, gg, dB, bg, s->stream);
    if (e == hipSuccess) e = afx::launch_cols_get(s->n, k, c0, gg, dX, xg, s->stream);
    if (e != hipSuccess) {
      rc = fail(AFX_ERR_HIP, hipGetErrorString(e));
      break;
    }
    const int r = solve_device_impl(s, gg, bg, xg, sc.data(), o, stats ? stats + c0 : nullptr, false, gg > 1 ? 1 : 0);
    if (r != AFX_OK && r != AFX_ERR_MAXIT && r != AFX_ERR_BREAKDOWN) {
      rc = r;
      break;
    }
    if (r == AFX_ERR_BREAKDOWN || (r == AFX_ERR_MAXIT && ret == AFX_OK)) ret = r;
    e = afx::launch_cols_put(s->n, k, c0, gg, xg, dX, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) rc = fail(AFX_ERR_HIP, hipGetErrorString(e));
  }
  dfree(bg);
  dfree(xg);
  return rc != AFX_OK ? rc : ret;
}

// Right-hand-side block width 1: the k columns solved one after another by the single-column
// kernel, each at its own frequency. Its pipelined sweep keeps a block's whole row slab in LDS,
// which k columns do not fit; measured per column, one at a time is 1.9x (32x32x24) and 2.3x
// (64x64x48) faster than the k-column kernel.
int solve_columns(afx_system s, int k, const cx<double> *dB, cx<double> *dX, const double *omega_scale,
                  const afx_opts *o, afx_stats *stats) {
  cx<double> *b1 = nullptr, *x1 = nullptr;
  int rc;
  if ((rc = dalloc(&b1, s->n)) || (rc = dalloc(&x1, s->n))) {
    dfree(b1);
    return rc;
  }
  int ret = AFX_OK;
  for (int c = 0; c < k && rc == AFX_OK; c++) {
    const double sc = omega_scale ? omega_scale[c] : 1.0;
    afx_stats st1{};
    hipError_t e = afx::launch_col_get(s->n, k, c, dB, b1, s->stream);
    if (e == hipSuccess) e = afx::launch_col_get(s->n, k, c, dX, x1, s->stream);
    if (e != hipSuccess) {
      rc = fail(AFX_ERR_HIP, hipGetErrorString(e));
      break;
    }
    const int r = solve_device_impl(s, 1, b1, x1, &sc, o, &st1);
    if (r != AFX_OK && r != AFX_ERR_MAXIT && r != AFX_ERR_BREAKDOWN) {
      rc = r;
      break;
    }
    if (r == AFX_ERR_BREAKDOWN || (r == AFX_ERR_MAXIT && ret == AFX_OK)) ret = r;
    if (stats) stats[c] = st1;
    e = afx::launch_col_put(s->n, k, c, x1, dX, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) rc = fail(AFX_ERR_HIP, hipGetErrorString(e));
  }
  dfree(b1);
  dfree(x1);
  return rc != AFX_OK ? rc : ret;
}

// FP64 true residual |B_s - A X_s| / |B_s| of every stacked system (0 for a zero B_s)
int true_residual_stacked(afx_system s, const cx<double> *dB, const cx<double> *dX, double *res) {
  std::vector<double> wcol, wf;
  std::vector<int> col2f;
  column_freqs(s, 1, nullptr, wcol, wf, col2f);
  const int G = 2 * s->ncu, m = s->nsys;
  cx<double> *dY = nullptr;
  double *dw = nullptr, *pa = nullptr, *pb = nullptr, *norms = nullptr;
  int rc;
  if ((rc = dalloc(&dY, (size_t)s->n)) || (rc = upload(&dw, wcol.data(), 1)) || (rc = dalloc(&pa, (size_t)G)) ||
      (rc = dalloc(&pb, (size_t)G)) || (rc = dalloc(&norms, 2 * (size_t)std::max(m, 1)))) {
    dfree(dY); dfree(dw); dfree(pa); dfree(pb); dfree(norms);
    return rc;
  }
  afx::SpmvParams sp{};
  sp.n = s->n;
  sp.k = 1;
  sp.kp = 1;
  sp.rowptr = s->d_rowptr;
  sp.col = s->d_col;
  sp.val = s->d_val64;
  sp.wcol = dw;
  sp.X = dX;
  sp.B = dB;
  sp.Y = dY;
  sp.part_rr = pa;
  sp.part_bb = pb;
  hipError_t e = afx::launch_residual_norms(sp, G, norms, s->stream);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_seg_norms, dim3(m), dim3(256), 0, s->stream, m, s->d_sys_off, dY, dB, norms);
    e = hipGetLastError();
  }
  std::vector<double> h(2 * (size_t)m);
  if (e == hipSuccess) e = hipMemcpyAsync(h.data(), norms, h.size() * 8, hipMemcpyDeviceToHost, s->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  dfree(dY); dfree(dw); dfree(pa); dfree(pb); dfree(norms);
  if (e != hipSuccess) return fail(AFX_ERR_HIP, hipGetErrorString(e));
  for (int q = 0; q < m; q++) res[q] = h[2 * q + 1] > 0.0 ? std::sqrt(h[2 * q] / h[2 * q + 1]) : 0.0;
  return AFX_OK;
}

int solve_device_impl(afx_system s, int k, const cx<double> *dB, cx<double> *dX, const double *omega_scale,
                      const afx_opts *o, afx_stats *stats, bool stacked, int force_width) {
  if (!s->have_values) return fail(AFX_ERR_ARG, "no matrix values");
  if (k > 256) return fail(AFX_ERR_ARG, "at most 256 right-hand sides per call");
  int rc = ensure_derived(s);
  if (rc) return rc;
  // stacked systems: one column; the states, stats and true residuals are per system
  const bool sys = stacked && s->nsys > 1;
  if (sys && (k != 1 || omega_scale)) return fail(AFX_ERR_ARG, "stacked systems take one right-hand side");
  // block width: columns one at a time unless the k-column kernel was measured faster for this k
  // (afx_autotune), is asked for (AFX_BATCH=multi; AFX_BATCH=columns forces one at a time), or
  // the caller reuses a preconditioner set up for exactly these k frequencies (nset = 0)
  const char *bm = getenv("AFX_BATCH");
  const int width = force_width >= 0 ? force_width
                    : bm ? (std::string(bm) == "multi" ? 1 : std::string(bm) == "stack" ? 2 : 0) : s->batch_multi[k];
  // the node blocks, like the nodal correction, have no stacked-copies form: one column at a time
  // unless the k-column kernel was chosen
  const bool grad = ((o->nset == 0 ? s->pc_kind : o->precond) & AFX_PC_GRAD) != 0 ||
                    ((o->nset == 0 ? s->pc_kind : o->precond) & 0xff) == AFX_PC_NODEBLK;
  const bool multi = width == 1;
  if (k > 1 && o->nset != 0 && width == 2 && !stacked && !grad)
    return solve_stacked_columns(s, k, dB, dX, omega_scale, o, stats);
  if (k > 1 && (o->nset != 0 || grad) && !multi) return solve_columns(s, k, dB, dX, omega_scale, o, stats);
  // the k-column kernel without sweeps takes at most 8 columns: larger blocks in groups of 8
  const int base_pc = (o->nset == 0 ? s->pc_kind : o->precond) & 0xff;
  if (multi && k > 8 && o->precision == AFX_FP64 &&
      (base_pc == AFX_PC_DIAG || base_pc == AFX_PC_NONE || base_pc == AFX_PC_NODEBLK) &&
      !getenv("AFX_NO_STREAM"))
    return solve_groups(s, k, 8, dB, dX, omega_scale, o, stats);
  std::vector<double> wcol, wf;
  std::vector<int> col2f;
  column_freqs(s, k, omega_scale, wcol, wf, col2f);
  const int pc = o->precond;
  if (o->nset == 0) {
    // reuse the existing setup (the SAVEd DD of COCG13_single.f90:39,57-62)
    if (!s->pc_valid || s->pc_w != wf) return AFX_ERR_BADOPT;
  } else if (pc != AFX_PC_NONE) {
    if ((rc = ensure_precond(s, pc, o->sigma_diag, o->sigma, wf))) return rc;
  }
  const int kind = (o->nset == 0) ? s->pc_kind : pc;
  const int nf = (int)wf.size();
  const int ks = sys ? s->nsys : k;
  std::vector<afx_stats> st(ks);
  for (auto &x : st) x = afx_stats{0, 0, 0.0, -1.0, 0};
  int ret = AFX_OK;
  // COCG with an FP64 (AFX_FP64) or complex-FP32 (AFX_MIXED) preconditioner; the recurrence,
  // SpMV and scalars are FP64. With check_true, the FP64 true residual is recomputed and the
  // solve restarts from the current iterate (an FP64 refinement step) while it exceeds tol.
  const int outer_max = o->check_true ? std::max(1, o->max_outer) : 1;
  std::vector<int> total(ks, 0);
  for (int outer = 0; outer < outer_max; ++outer) {
    std::vector<afx_stats> ist(ks);
    int budget = o->maxit;
    for (int c = 0; c < ks; c++) budget = std::min(budget, std::max(1, o->maxit - total[c]));
    ret = (o->precision == AFX_FP64)
              ? krylov<double>(s, k, dB, dX, wcol, col2f, nf, kind, o->tol, budget, o->chunk, ist.data(), nullptr, sys)
              : krylov<float>(s, k, dB, dX, wcol, col2f, nf, kind, o->tol, budget, o->chunk, ist.data(), nullptr, sys);
    if (ret != AFX_OK && ret != AFX_ERR_MAXIT && ret != AFX_ERR_BREAKDOWN) return ret;
    for (int c = 0; c < ks; c++) {
      total[c] += (outer == 0 || ist[c].iter > 0) ? std::min(ist[c].iter, budget) : 0;
      st[c] = ist[c];
      st[c].iter = (outer == 0) ? ist[c].iter : total[c];
      st[c].outer = outer;
    }
    if (!o->check_true) break;
    std::vector<double> tr(ks);
    if ((rc = sys ? true_residual_stacked(s, dB, dX, tr.data())
                  : afx_true_residual_device(s, k, dB, dX, omega_scale, tr.data())))
      return rc;
    bool ok = true;
    for (int c = 0; c < ks; c++) {
      st[c].true_resid = tr[c];
      if (!(tr[c] <= o->tol)) ok = false;
    }
    if (ok || ret == AFX_ERR_BREAKDOWN) break;
    bool budget_left = false;
    for (int c = 0; c < ks; c++)
      if (total[c] < o->maxit) budget_left = true;
    if (!budget_left) break;
  }
  if (stats)
    for (int c = 0; c < ks; c++) stats[c] = st[c];
  return ret;
}

}  // namespace

extern "C" {

void afx_default_opts(afx_opts *o) {
  if (!o) return;
  o->precond = AFX_PC_SSOR;
  o->sigma_diag = 1.0;
  o->sigma = 0.0;
  o->tol = 1e-8;
  o->maxit = 100000;
  o->nset = 1;
  o->precision = AFX_FP64;
  o->inner_tol = 1e-4;
  o->max_outer = 20;
  o->chunk = 0;
  o->check_true = 1;
}

const char *afx_last_error(void) { return g_err.c_str(); }

int afx_create(int n, const int *INU, const int *IAU, const int *INL, const int *IAL, afx_system *out) {
  if (!out || !INU || !INL) return fail(AFX_ERR_ARG, "null argument");
  *out = nullptr;
  int rc = validate_pattern(n, INU, IAU, INL, IAL);
  if (rc) return rc;
  afx_system s = new afx_system_s();
  s->n = n;
  s->INU.assign(INU, INU + n + 1);
  s->INL.assign(INL, INL + n + 1);
  s->IAU.assign(IAU, IAU + INU[n]);
  s->IAL.assign(IAL, IAL + INL[n]);
  rc = create_common(s);
  if (rc) {
    release(s);
    delete s;
    return rc;
  }
  *out = s;
  return AFX_OK;
}

int afx_create_stacked(int m, const int *row_off, int n, const int *INU, const int *IAU, const int *INL,
                       const int *IAL, afx_system *out) {
  if (!out || !row_off || !INU || !INL || m < 1) return fail(AFX_ERR_ARG, "bad arguments");
  *out = nullptr;
  if (row_off[0] != 0 || row_off[m] != n) return fail(AFX_ERR_ARG, "row offsets must run from 0 to n");
  for (int q = 0; q < m; q++)
    if (row_off[q + 1] <= row_off[q]) return fail(AFX_ERR_ARG, "every stacked system needs at least one row");
  int rc = validate_pattern(n, INU, IAU, INL, IAL);
  if (rc) return rc;
  // no entry may couple two systems
  for (int q = 0; q < m; q++)
    for (int i = row_off[q]; i < row_off[q + 1]; i++) {
      for (int j = INU[i]; j < INU[i + 1]; j++)
        if (IAU[j] - 1 < row_off[q] || IAU[j] - 1 >= row_off[q + 1])
          return fail(AFX_ERR_ARG, "an upper entry couples two stacked systems");
      for (int j = INL[i]; j < INL[i + 1]; j++)
        if (IAL[j] - 1 < row_off[q] || IAL[j] - 1 >= row_off[q + 1])
          return fail(AFX_ERR_ARG, "a lower entry couples two stacked systems");
    }
  afx_system s = new afx_system_s();
  s->n = n;
  s->INU.assign(INU, INU + n + 1);
  s->INL.assign(INL, INL + n + 1);
  s->IAU.assign(IAU, IAU + INU[n]);
  s->IAL.assign(IAL, IAL + INL[n]);
  s->nsys = m;
  s->sys_off.assign(row_off, row_off + m + 1);
  rc = create_common(s);
  if (rc) {
    release(s);
    delete s;
    return rc;
  }
  *out = s;
  return AFX_OK;
}

int afx_create_from_n6line(int nline, int ntet, const int *n6line, afx_system *out) {
  if (!out || !n6line || nline <= 0 || ntet <= 0) return fail(AFX_ERR_ARG, "bad arguments");
  *out = nullptr;
  // C_set_dof2 (m_iccg_var_takuto.f90:361-485): per row, the sorted set of edges sharing a tet
  std::vector<long long> up_keys, lo_keys;
  up_keys.reserve((size_t)ntet * 15);
  lo_keys.reserve((size_t)ntet * 15);
  for (int t = 0; t < ntet; t++) {
    int g[6];
    for (int j = 0; j < 6; j++) {
      g[j] = std::abs(n6line[(size_t)t * 6 + j]);
      if (g[j] < 1 || g[j] > nline) return fail(AFX_ERR_ARG, "n6line entry out of range");
    }
    for (int a = 0; a < 6; a++)
      for (int b = 0; b < 6; b++) {
        if (g[a] < g[b]) up_keys.push_back((long long)(g[a] - 1) * nline + (g[b] - 1));
        else if (g[a] > g[b]) lo_keys.push_back((long long)(g[a] - 1) * nline + (g[b] - 1));
      }
  }
  auto finish = [&](std::vector<long long> &keys, std::vector<int> &ptr, std::vector<int> &idx) {
    std::sort(keys.begin(), keys.end());
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    ptr.assign(nline + 1, 0);
    idx.resize(keys.size());
    for (size_t q = 0; q < keys.size(); q++) {
      ptr[keys[q] / nline + 1]++;
      idx[q] = (int)(keys[q] % nline) + 1;
    }
    for (int i = 0; i < nline; i++) ptr[i + 1] += ptr[i];
  };
  afx_system s = new afx_system_s();
  s->n = nline;
  finish(up_keys, s->INU, s->IAU);
  finish(lo_keys, s->INL, s->IAL);
  s->ntet = ntet;
  s->n6line.assign(n6line, n6line + (size_t)ntet * 6);
  int rc = create_common(s);
  if (rc) {
    release(s);
    delete s;
    return rc;
  }
  *out = s;
  return AFX_OK;
}

int afx_destroy(afx_system s) {
  if (!s) return AFX_OK;
  release(s);
  delete s;
  return AFX_OK;
}

int afx_get_sizes(afx_system s, int *n, int *npu, int *npl, int *nlev_l, int *nlev_u) {
  if (!s) return fail(AFX_ERR_ARG, "null system");
  if (n) *n = s->n;
  if (npu) *npu = (int)s->nU;
  if (npl) *npl = (int)s->nL;
  if (nlev_l) *nlev_l = s->nlevL;
  if (nlev_u) *nlev_u = s->nlevU;
  return AFX_OK;
}

int afx_get_pattern(afx_system s, int *INU, int *IAU, int *INL, int *IAL) {
  if (!s) return fail(AFX_ERR_ARG, "null system");
  if (INU) std::copy(s->INU.begin(), s->INU.end(), INU);
  if (IAU) std::copy(s->IAU.begin(), s->IAU.end(), IAU);
  if (INL) std::copy(s->INL.begin(), s->INL.end(), INL);
  if (IAL) std::copy(s->IAL.begin(), s->IAL.end(), IAL);
  return AFX_OK;
}

int afx_set_values(afx_system s, const double *D, const double *AU, const double *AL) {
  if (!s || !D) return fail(AFX_ERR_ARG, "null argument");
  const size_t n = s->n, nU = s->nU, nL = s->nL;
  HIPCHK(hipMemcpy(s->d_canon, D, n * sizeof(cx<double>), hipMemcpyHostToDevice));
  if (nU) HIPCHK(hipMemcpy(s->d_canon + n, AU, nU * sizeof(cx<double>), hipMemcpyHostToDevice));
  if (nL) HIPCHK(hipMemcpy(s->d_canon + n + nU, AL, nL * sizeof(cx<double>), hipMemcpyHostToDevice));
  s->omega = 1.0;  // (Re, Im) stored as (K, M): entry = K + i*1*M
  s->have_values = true;
  invalidate_values(s);
  return AFX_OK;
}

int afx_get_values(afx_system s, double *D, double *AU, double *AL) {
  if (!s) return fail(AFX_ERR_ARG, "null system");
  return get_values_at(s, s->omega, D, AU, AL);
}

}  // extern "C"

namespace {

// the split values D, AU, AL (interleaved complex, host) at angular frequency w
int get_values_at(afx_system s, double w, double *D, double *AU, double *AL) {
  const size_t n = s->n, nU = s->nU, nL = s->nL;
  std::vector<cx<double>> h(n + nU + nL);
  HIPCHK(hipStreamSynchronize(s->stream));
  HIPCHK(hipMemcpy(h.data(), s->d_canon, h.size() * sizeof(cx<double>), hipMemcpyDeviceToHost));
  auto put = [&](double *dst, size_t off, size_t m) {
    if (!dst) return;
    for (size_t q = 0; q < m; q++) {
      dst[2 * q] = h[off + q].r;
      dst[2 * q + 1] = w * h[off + q].i;
    }
  };
  put(D, 0, n);
  put(AU, n, nU);
  put(AL, n + nU, nL);
  return AFX_OK;
}

// the k stacked copies of block width "stack" for the column omegas wcol (pattern once per k,
// values again whenever the omegas or the system's values changed)
int ensure_stack(afx_system s, int k, const std::vector<double> &wcol) {
  if (s->stack && s->stack->nsys == k && s->stack_w == wcol && s->stack_ver == s->val_ver) return AFX_OK;
  const long long n = s->n, nU = s->nU, nL = s->nL;
  int rc;
  if (!s->stack || s->stack->nsys != k) {
    if (s->stack) afx_destroy(s->stack);
    s->stack = nullptr;
    if (n * k > 0x7fffffffLL || s->nnz * k > 0x7fffffffLL)
      return fail(AFX_ERR_ARG, "the stacked copies exceed 32-bit indices");
    std::vector<int> off(k + 1), INU(n * k + 1), INL(n * k + 1), IAU(nU * k), IAL(nL * k);
    INU[0] = INL[0] = 0;
    for (int c = 0; c < k; c++) {
      off[c] = (int)(c * n);
      for (long long i = 0; i < n; i++) {
        INU[c * n + i + 1] = (int)(c * nU + s->INU[i + 1]);
        INL[c * n + i + 1] = (int)(c * nL + s->INL[i + 1]);
      }
      for (long long j = 0; j < nU; j++) IAU[c * nU + j] = (int)(s->IAU[j] + c * n);
      for (long long j = 0; j < nL; j++) IAL[c * nL + j] = (int)(s->IAL[j] + c * n);
    }
    off[k] = (int)(n * k);
    afx_system t = nullptr;
    if ((rc = afx_create_stacked(k, off.data(), (int)(n * k), INU.data(), IAU.data(), INL.data(), IAL.data(), &t)))
      return rc;
    t->lpr64 = s->lpr64;  // the same rows: the lane mapping tuned for one copy
    t->lpr32 = s->lpr32;
    s->stack = t;
  }
  std::vector<cx<double>> h(n + nU + nL);
  HIPCHK(hipStreamSynchronize(s->stream));
  HIPCHK(hipMemcpy(h.data(), s->d_canon, h.size() * sizeof(cx<double>), hipMemcpyDeviceToHost));
  std::vector<double> D(2 * n * k), AU(2 * nU * k), AL(2 * nL * k);
  for (int c = 0; c < k; c++) {
    const double w = wcol[c];
    auto put = [&](double *dst, size_t off, size_t m) {
      for (size_t q = 0; q < m; q++) {
        dst[2 * q] = h[off + q].r;
        dst[2 * q + 1] = w * h[off + q].i;
      }
    };
    put(D.data() + 2 * n * c, 0, n);
    put(AU.data() + 2 * nU * c, n, nU);
    put(AL.data() + 2 * nL * c, n + nU, nL);
  }
  if ((rc = afx_set_values(s->stack, D.data(), AU.data(), AL.data()))) return rc;
  s->stack_w = wcol;
  s->stack_ver = s->val_ver;
  return AFX_OK;
}

}  // namespace

extern "C" {

int afx_set_mesh(afx_system s, int node, int ntet, const double *xyz, const int *n4, const int *group,
                 const double *vxyz, const double *fxyz) {
  if (!s || !xyz || !n4 || !group) return fail(AFX_ERR_ARG, "null argument");
  if (s->n6line.empty() || ntet != s->ntet) return fail(AFX_ERR_ARG, "system was not created from this n6line");
  for (long long q = 0; q < (long long)ntet * 4; q++)
    if (n4[q] < 1 || n4[q] > node) return fail(AFX_ERR_ARG, "n4 entry out of range");
  dfree(s->d_xyz); dfree(s->d_fxyz); dfree(s->d_n4); dfree(s->d_n6); dfree(s->d_group); dfree(s->d_vxyz);
  dfree(s->d_bptr); dfree(s->d_btet);
  s->have_pgrid = false;
  int rc;
  std::vector<double> zero3((size_t)node * 3, 0.0), zero6((size_t)node * 6, 0.0);
  if ((rc = upload(&s->d_xyz, xyz, (size_t)node * 3))) return rc;
  if ((rc = upload(&s->d_n4, n4, (size_t)ntet * 4))) return rc;
  if ((rc = upload(&s->d_n6, s->n6line.data(), (size_t)ntet * 6))) return rc;
  if ((rc = upload(&s->d_group, group, (size_t)ntet))) return rc;
  if ((rc = upload(&s->d_fxyz, fxyz ? fxyz : zero3.data(), (size_t)node * 3))) return rc;
  if ((rc = upload((double **)&s->d_vxyz, vxyz ? vxyz : zero6.data(), (size_t)node * 6))) return rc;
  s->node = node;
  s->h_xyz.assign(xyz, xyz + (size_t)node * 3);
  s->h_n4.assign(n4, n4 + (size_t)ntet * 4);
  s->have_mesh = true;
  if (!s->have_slots && (rc = build_slot_maps(s))) return rc;
  return AFX_OK;
}

int afx_assemble(afx_system s, const double *sigma_tet, double omega, double *b_out) {
  if (!s || !sigma_tet) return fail(AFX_ERR_ARG, "null argument");
  if (!s->have_mesh) return fail(AFX_ERR_ARG, "afx_set_mesh first");
  const int ntet = s->ntet, n = s->n;
  double *d_sig = nullptr, *eK = nullptr, *eM = nullptr;
  cx<double> *eB = nullptr;
  int *d_err = nullptr;
  int rc;
  if ((rc = upload(&d_sig, sigma_tet, ntet)) || (rc = dalloc(&eK, (size_t)ntet * 21)) ||
      (rc = dalloc(&eM, (size_t)ntet * 21)) || (rc = dalloc(&eB, (size_t)ntet * 6)) || (rc = dalloc(&d_err, 1))) {
    dfree(d_sig); dfree(eK); dfree(eM); dfree(eB); dfree(d_err);
    return rc;
  }
  if (!s->d_rhs_unit && (rc = dalloc(&s->d_rhs_unit, n))) return rc;
  hipError_t e = hipMemsetAsync(d_err, 0, sizeof(int), s->stream);
  afx::ElemParams ep{};
  ep.ntet = ntet;
  ep.xyz = s->d_xyz;
  ep.n4 = s->d_n4;
  ep.n6line = s->d_n6;
  ep.group = s->d_group;
  ep.sigma = d_sig;
  ep.vxyz = s->d_vxyz;
  ep.fxyz = s->d_fxyz;
  ep.eK = eK;
  ep.eM = eM;
  ep.eB = eB;
  ep.err = d_err;
  if (e == hipSuccess) e = afx::launch_element(ep, s->stream);
  if (e == hipSuccess)
    e = afx::launch_slot_gather((int)(n + s->nU), s->d_cptr, s->d_cidx, eK, eM, s->d_canon, s->stream);
  if (e == hipSuccess) e = afx::launch_mirror(s->nL, s->d_mirror, s->d_canon + n, s->d_canon + n + s->nU, s->stream);
  if (e == hipSuccess) e = afx::launch_rhs_gather(n, s->d_rptr, s->d_ridx, eB, 0.0, s->d_rhs_unit, s->stream);
  int herr = 0;
  if (e == hipSuccess) e = hipMemcpyAsync(&herr, d_err, sizeof(int), hipMemcpyDeviceToHost, s->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  dfree(d_sig); dfree(eK); dfree(eM); dfree(eB); dfree(d_err);
  if (e != hipSuccess) {
    g_err = std::string("assembly: ") + hipGetErrorString(e);
    return AFX_ERR_HIP;
  }
  if (herr) return fail(AFX_ERR_GEOM, "GEGEGE! volume is less than 0 (gradnodebasisfun)");
  s->omega = omega;
  s->have_values = true;
  invalidate_values(s);
  if (b_out) HIPCHK(hipMemcpy(b_out, s->d_rhs_unit, (size_t)n * sizeof(cx<double>), hipMemcpyDeviceToHost));
  return AFX_OK;
}

int afx_set_omega(afx_system s, double omega) {
  if (!s) return fail(AFX_ERR_ARG, "null system");
  if (s->omega != omega) {
    s->omega = omega;
    s->pc_valid = false;
  }
  return AFX_OK;
}

int afx_genbc(afx_system s, const int *line, double zmin, double zmax, double xout, double yout,
              unsigned char *line_bc) {
  if (!s || !line || !line_bc) return fail(AFX_ERR_ARG, "null argument");
  if (!s->have_mesh) return fail(AFX_ERR_ARG, "afx_set_mesh first");
  int *d_line = nullptr;
  unsigned char *d_bc = nullptr;
  int rc;
  if ((rc = upload(&d_line, line, (size_t)s->n * 2)) || (rc = dalloc(&d_bc, s->n))) {
    dfree(d_line);
    return rc;
  }
  hipError_t e = afx::launch_genbc(s->n, d_line, s->d_xyz, zmin, zmax, xout, yout, d_bc, s->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(line_bc, d_bc, s->n, hipMemcpyDeviceToHost, s->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  dfree(d_line);
  dfree(d_bc);
  if (e != hipSuccess) return fail(AFX_ERR_HIP, hipGetErrorString(e));
  return AFX_OK;
}

int afx_set_nodes(afx_system s, int node, const int *line, const unsigned char *dirichlet) {
  if (!s || !line || node <= 0) return fail(AFX_ERR_ARG, "bad arguments");
  const int n = s->n;
  // boundary nodes: the end nodes of Dirichlet edges
  std::vector<unsigned char> bnd(node, 0);
  for (int e = 0; e < n; e++) {
    const int a = line[2 * e] - 1, b = line[2 * e + 1] - 1;
    if (a < 0 || a >= node || b < 0 || b >= node || a == b) return fail(AFX_ERR_ARG, "line: bad node id");
    if (dirichlet && dirichlet[e]) bnd[a] = bnd[b] = 1;
  }
  std::vector<int> id(node, -1);
  int nn = 0;
  for (int v = 0; v < node; v++)
    if (!bnd[v]) id[v] = nn++;
  std::vector<int> en(2 * (size_t)n), cnt(nn + 1, 0);
  for (int e = 0; e < n; e++) {
    const int a = id[line[2 * e] - 1], b = id[line[2 * e + 1] - 1];
    en[2 * e] = a < 0 ? nn : a;
    en[2 * e + 1] = b < 0 ? nn : b;
    if (a >= 0) cnt[a + 1]++;
    if (b >= 0) cnt[b + 1]++;
  }
  for (int a = 0; a < nn; a++) cnt[a + 1] += cnt[a];
  std::vector<int> adj(cnt[nn]), fill(cnt.begin(), cnt.end() - 1);
  for (int e = 0; e < n; e++) {  // ascending edge ids per node
    if (en[2 * e] < nn) adj[fill[en[2 * e]]++] = -(e + 1);
    if (en[2 * e + 1] < nn) adj[fill[en[2 * e + 1]]++] = e + 1;
  }
  // stacked systems: each system's interior nodes must form one contiguous range that no other
  // system's edges touch (the parts' node ids concatenated in system order)
  std::vector<int> noff;
  if (s->nsys > 1) {
    noff.assign(s->nsys + 1, 0);
    std::vector<int> lo(s->nsys, nn), hi(s->nsys, -1);
    for (int q = 0; q < s->nsys; q++)
      for (int e = s->sys_off[q]; e < s->sys_off[q + 1]; e++)
        for (int j = 0; j < 2; j++)
          if (en[2 * e + j] < nn) {
            lo[q] = std::min(lo[q], en[2 * e + j]);
            hi[q] = std::max(hi[q], en[2 * e + j]);
          }
    int next = 0;
    for (int q = 0; q < s->nsys; q++) {
      if (hi[q] < 0) {  // no interior node
        noff[q] = next;
        continue;
      }
      if (lo[q] != next) return fail(AFX_ERR_ARG, "afx_set_nodes: a stacked system's nodes are not contiguous");
      noff[q] = lo[q];
      next = hi[q] + 1;
    }
    noff[s->nsys] = next;
    if (next != nn) return fail(AFX_ERR_ARG, "afx_set_nodes: interior nodes outside the stacked systems");
  }
  // node blocks: maximal runs of consecutive rows with the same first node (MKLINE numbers edges
  // by their first node, so a run is the node's edges towards higher-numbered nodes), cut into
  // pieces of at most AFX_NB_W rows; tiles of whole blocks, at most 256 rows each
  std::vector<unsigned char> bm(n);
  std::vector<int> tptr(1, 0), bstart;
  for (int e0 = 0; e0 < n;) {
    int e1 = e0 + 1;
    while (e1 < n && e1 - e0 < AFX_NB_W && line[2 * e1] == line[2 * e0]) e1++;
    bstart.push_back(e0);
    if (e1 - tptr.back() > 256) tptr.push_back(e0);
    for (int e = e0; e < e1; e++) bm[e] = (unsigned char)((e - e0) | ((e1 - e0 - 1) << 4));
    e0 = e1;
  }
  tptr.push_back(n);
  dfree(s->d_nptr); dfree(s->d_nadj); dfree(s->d_nadj16); dfree(s->d_en); dfree(s->d_mnd); dfree(s->d_noff);
  dfree(s->d_tptr); dfree(s->d_bm); dfree(s->d_bi); dfree(s->d_bstart);
  s->nn = -1;
  s->ntile = s->nblk = 0;
  int rc;
  if (!noff.empty() && (rc = upload(&s->d_noff, noff.data(), noff.size()))) return rc;
  // the lists again at a fixed width of 16 slots (0 = unused), loaded as four 16-byte vectors
  std::vector<int> adj16((size_t)nn * 16, 0);
  for (int a = 0; a < nn; a++)
    for (int q = cnt[a]; q < cnt[a + 1] && q - cnt[a] < 16; q++) adj16[(size_t)a * 16 + (q - cnt[a])] = adj[q];
  if ((rc = upload(&s->d_nptr, cnt.data(), cnt.size())) || (rc = upload(&s->d_nadj, adj.data(), adj.size())) ||
      (nn > 0 && (rc = upload(&s->d_nadj16, adj16.data(), adj16.size()))) ||
      (rc = upload(&s->d_en, en.data(), en.size())) || (rc = dalloc(&s->d_mnd, (size_t)nn + 1)) ||
      (rc = upload(&s->d_tptr, tptr.data(), tptr.size())) || (rc = upload(&s->d_bm, bm.data(), bm.size())) ||
      (rc = upload(&s->d_bstart, bstart.data(), bstart.size())))
    return rc;
  s->nn = nn;
  s->ntile = (int)tptr.size() - 1;
  s->nblk = (int)bstart.size();
  if (s->pc_kind & AFX_PC_GRAD || (s->pc_kind & 0xff) == AFX_PC_NODEBLK) s->pc_valid = false;
  return AFX_OK;
}

int afx_set_bc(afx_system s, const unsigned char *dirichlet, int nrhs, double *rhs, int ldr) {
  if (!s || !dirichlet) return fail(AFX_ERR_ARG, "null argument");
  unsigned char *d_dir = nullptr;
  int rc;
  if ((rc = upload(&d_dir, dirichlet, s->n))) return rc;
  hipError_t e = afx::launch_bc_values(s->n, s->d_INU, s->d_IAU, s->d_INL, s->d_IAL, d_dir, s->d_canon,
                                       s->d_canon + s->n, s->d_canon + s->n + s->nU, s->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  dfree(d_dir);
  if (e != hipSuccess) return fail(AFX_ERR_HIP, hipGetErrorString(e));
  // right-hand side: rf(i) = 0 on Dirichlet rows; non-Dirichlet rows subtract K*0 = 0
  if (rhs)
    for (int c = 0; c < nrhs; c++)
      for (int i = 0; i < s->n; i++)
        if (dirichlet[i]) {
          rhs[2 * ((size_t)c * ldr + i)] = 0.0;
          rhs[2 * ((size_t)c * ldr + i) + 1] = 0.0;
        }
  invalidate_values(s);
  return AFX_OK;
}

int afx_precond_setup(afx_system s, int kind, double sigma_diag, double sigma) {
  if (!s) return fail(AFX_ERR_ARG, "null system");
  if (!s->have_values) return fail(AFX_ERR_ARG, "no matrix values");
  std::vector<double> wf(1, s->omega);
  return ensure_precond(s, kind, sigma_diag, sigma, wf);
}

int afx_get_precond_diag(afx_system s, double *DD) {
  if (!s || !DD) return fail(AFX_ERR_ARG, "null argument");
  if (!s->pc_valid) return fail(AFX_ERR_ARG, "no preconditioner set up");
  const int nf = (int)s->pc_w.size();
  std::vector<cx<double>> h((size_t)s->n * nf);
  HIPCHK(hipMemcpy(h.data(), s->d_DD, h.size() * sizeof(cx<double>), hipMemcpyDeviceToHost));
  for (int i = 0; i < s->n; i++) {
    DD[2 * i] = h[(size_t)i * nf].r;
    DD[2 * i + 1] = h[(size_t)i * nf].i;
  }
  return AFX_OK;
}

int afx_spmv(afx_system s, const double *x, double *y) {
  if (!s || !x || !y) return fail(AFX_ERR_ARG, "null argument");
  int rc = ensure_derived(s);
  if (rc) return rc;
  cx<double> *dx = nullptr, *dy = nullptr;
  if ((rc = upload(&dx, (const cx<double> *)x, s->n)) || (rc = dalloc(&dy, s->n))) {
    dfree(dx); dfree(dy);
    return rc;
  }
  hipError_t e = afx::launch_spmv_stream(s->n, s->maxrow, s->d_rowptr, s->d_col, s->d_val64, s->omega, dx, dy, s->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(y, dy, (size_t)s->n * 16, hipMemcpyDeviceToHost, s->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  dfree(dx); dfree(dy);
  if (e != hipSuccess) return fail(AFX_ERR_HIP, hipGetErrorString(e));
  return AFX_OK;
}

int afx_spmv_device(afx_system s, const void *dX, void *dY, double *kernel_ms) {
  if (!s || !dX || !dY) return fail(AFX_ERR_ARG, "null argument");
  if (dX == dY) return fail(AFX_ERR_ARG, "x and y must not overlap");
  int rc = ensure_derived(s);
  if (rc) return rc;
  if (s->nnz < 1) return fail(AFX_ERR_ARG, "empty matrix");
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  if (kernel_ms) {
    HIPCHK(hipEventCreate(&ev0));
    if (hipEventCreate(&ev1) != hipSuccess) {
      (void)hipEventDestroy(ev0);
      return fail(AFX_ERR_HIP, "hipEventCreate");
    }
  }
  hipError_t e = kernel_ms ? hipEventRecord(ev0, s->stream) : hipSuccess;
  if (e == hipSuccess)
    e = afx::launch_spmv_stream(s->n, s->maxrow, s->d_rowptr, s->d_col, s->d_val64, s->omega, (const cx<double> *)dX,
                                (cx<double> *)dY, s->stream);
  if (kernel_ms) {
    if (e == hipSuccess) e = hipEventRecord(ev1, s->stream);
    if (e == hipSuccess) e = hipEventSynchronize(ev1);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, ev0, ev1);
    *kernel_ms = ms;
    (void)hipEventDestroy(ev0);
    (void)hipEventDestroy(ev1);
  }
  if (e != hipSuccess) return fail(AFX_ERR_HIP, hipGetErrorString(e));
  return AFX_OK;
}

int afx_precond_apply(afx_system s, const double *r, double *z) {
  if (!s || !r || !z) return fail(AFX_ERR_ARG, "null argument");
  if (!s->pc_valid || s->pc_w.size() != 1) return fail(AFX_ERR_ARG, "no single-frequency preconditioner set up");
  cx<double> *dr = nullptr, *dz = nullptr;
  double *dw = nullptr;
  int rc;
  const double w = s->pc_w[0];
  if ((rc = upload(&dr, (const cx<double> *)r, s->n)) || (rc = dalloc(&dz, s->n)) || (rc = upload(&dw, &w, 1))) {
    dfree(dr); dfree(dz); dfree(dw);
    return rc;
  }
  hipError_t e = (s->pc_kind & 0xff) == AFX_PC_NODEBLK
                     ? afx::launch_nodeblk_apply(s->ntile, s->d_tptr, s->d_bm, s->d_bi, dr, dz, s->stream)
                     : afx::launch_precond_apply_levels(s->n, s->pc_kind & 0xff, s->nlevL, s->d_lvlL, s->d_rowL,
                                                        s->d_Lptr, s->d_Lcol, s->d_Lval64, s->nlevU, s->d_lvlU,
                                                        s->d_rowU, s->d_Uptr, s->d_Ucol, s->d_Uval64, s->d_DD, dw, dr,
                                                        dz, s->stream);
  cx<double> *dzn = nullptr;
  if (e == hipSuccess && (s->pc_kind & AFX_PC_GRAD)) {
    if (dalloc(&dzn, (size_t)s->nn + 1) != AFX_OK) e = hipErrorOutOfMemory;
    if (e == hipSuccess)
      e = afx::launch_grad_apply(s->n, s->nn, s->d_nptr, s->d_nadj, s->d_en, s->d_mnd, w, dr, dzn, dz, s->stream);
  }
  if (e == hipSuccess) e = hipMemcpyAsync(z, dz, (size_t)s->n * 16, hipMemcpyDeviceToHost, s->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  dfree(dr); dfree(dz); dfree(dw); dfree(dzn);
  if (e != hipSuccess) return fail(AFX_ERR_HIP, hipGetErrorString(e));
  return AFX_OK;
}

int afx_solve_device(afx_system s, int nrhs, const void *dB, void *dX, const double *omega_scale,
                     const afx_opts *opts, afx_stats *stats) {
  if (!s || !dB || !dX || nrhs <= 0) return fail(AFX_ERR_ARG, "bad arguments");
  afx_opts o;
  if (opts) o = *opts;
  else afx_default_opts(&o);
  return solve_device_impl(s, nrhs, (const cx<double> *)dB, (cx<double> *)dX, omega_scale, &o, stats);
}

int afx_solve_stacked_device(afx_system s, const void *dB, void *dX, const afx_opts *opts, afx_stats *stats) {
  if (!s || !dB || !dX) return fail(AFX_ERR_ARG, "bad arguments");
  afx_opts o;
  if (opts) o = *opts;
  else afx_default_opts(&o);
  return solve_device_impl(s, 1, (const cx<double> *)dB, (cx<double> *)dX, nullptr, &o, stats, true);
}

int afx_stacked_count(afx_system s, int *nsys) {
  if (!s || !nsys) return fail(AFX_ERR_ARG, "bad arguments");
  *nsys = s->nsys;
  return AFX_OK;
}

int afx_solve(afx_system s, int nrhs, const double *B, int ldb, double *X, int ldx, const afx_opts *opts,
              afx_stats *stats) {
  if (!s || !B || !X || nrhs <= 0 || ldb < s->n || ldx < s->n) return fail(AFX_ERR_ARG, "bad arguments");
  const size_t n = s->n;
  std::vector<cx<double>> hb(n * nrhs), hx(n * nrhs);
  for (int c = 0; c < nrhs; c++)
    for (size_t i = 0; i < n; i++) {
      hb[i * nrhs + c] = cx<double>{B[2 * ((size_t)c * ldb + i)], B[2 * ((size_t)c * ldb + i) + 1]};
      hx[i * nrhs + c] = cx<double>{X[2 * ((size_t)c * ldx + i)], X[2 * ((size_t)c * ldx + i) + 1]};
    }
  cx<double> *dB = nullptr, *dX = nullptr;
  int rc;
  if ((rc = upload(&dB, hb.data(), hb.size())) || (rc = upload(&dX, hx.data(), hx.size()))) {
    dfree(dB);
    dfree(dX);
    return rc;
  }
  afx_opts o;
  if (opts) o = *opts;
  else afx_default_opts(&o);
  rc = solve_device_impl(s, nrhs, dB, dX, nullptr, &o, stats);
  if (rc == AFX_OK || rc == AFX_ERR_MAXIT || rc == AFX_ERR_BREAKDOWN) {
    if (hipMemcpy(hx.data(), dX, hx.size() * sizeof(cx<double>), hipMemcpyDeviceToHost) != hipSuccess) rc = AFX_ERR_HIP;
    for (int c = 0; c < nrhs; c++)
      for (size_t i = 0; i < n; i++) {
        X[2 * ((size_t)c * ldx + i)] = hx[i * nrhs + c].r;
        X[2 * ((size_t)c * ldx + i) + 1] = hx[i * nrhs + c].i;
      }
  }
  dfree(dB);
  dfree(dX);
  return rc;
}

}  // extern "C"

// ---- the residual norms without allocation --------------------------------------------------
// Device scratch of afx_true_residual_device_noalloc: the column omegas [AFX_NOALLOC_MAX_COLS],
// then the per-block partial sums of ||b - A x||^2 and ||b||^2 [AFX_NOALLOC_MAX_GRID][nrhs] each,
// then the finished norms [2 nrhs]. One copy per device, held until the process ends.
#define AFX_NOALLOC_MAX_COLS 64
#define AFX_NOALLOC_MAX_GRID 1024
#define AFX_NOALLOC_SCRATCH_DOUBLES \
  (AFX_NOALLOC_MAX_COLS + 2 * AFX_NOALLOC_MAX_GRID * AFX_NOALLOC_MAX_COLS + 2 * AFX_NOALLOC_MAX_COLS)
static __device__ __align__(256) double scratch[AFX_NOALLOC_SCRATCH_DOUBLES];

namespace {
// the device address of `scratch` on each device, looked up once per device (the symbol's host
// name is a stand-in, not a device address)
std::mutex g_scratch_mu;
double *g_scratch_dev[64] = {nullptr};

int scratch_for_current_device(double **out) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return fail(AFX_ERR_ARG, "device id beyond 63");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void *ptr = nullptr;
    HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    g_scratch_dev[dev] = (double *)ptr;
  }
  *out = g_scratch_dev[dev];
  return AFX_OK;
}

struct OmegaArg {
  double w[AFX_NOALLOC_MAX_COLS];
};
__global__ void k_put_omegas(int k, OmegaArg a, double *out) {
  const int c = threadIdx.x;
  if (c < k) out[c] = a.w[c];
}
}  // namespace

extern "C" {

int afx_true_residual_device_noalloc(afx_system s, int nrhs, const void *dB, const void *dX,
                                     const double *omega_scale, double *res) {
  if (!s || !dB || !dX || !res || nrhs <= 0) return fail(AFX_ERR_ARG, "bad arguments");
  if (nrhs > AFX_NOALLOC_MAX_COLS) return fail(AFX_ERR_ARG, "afx_true_residual_device_noalloc: more than 64 columns");
  if (!s->have_values) return fail(AFX_ERR_ARG, "no matrix values");
  const int G = std::min(2 * s->ncu, AFX_NOALLOC_MAX_GRID);
  if (G <= 0) return fail(AFX_ERR_ARG, "no compute units");
  double *sc = nullptr;
  int rc = scratch_for_current_device(&sc);
  if (rc) return rc;
  if ((rc = ensure_derived(s))) return rc;  // value gathers only (kernels, no allocation)
  OmegaArg a{};
  for (int c = 0; c < nrhs; c++) a.w[c] = s->omega * (omega_scale ? omega_scale[c] : 1.0);
  double *dw = sc, *pa = sc + AFX_NOALLOC_MAX_COLS, *pb = pa + (size_t)G * nrhs, *norms = pb + (size_t)G * nrhs;
  hipLaunchKernelGGL(k_put_omegas, dim3(1), dim3(AFX_NOALLOC_MAX_COLS), 0, s->stream, nrhs, a, dw);
  HIPCHK(hipGetLastError());
  afx::SpmvParams sp{};
  sp.n = s->n;
  sp.k = nrhs;
  sp.kp = nrhs > 1 ? next_pow2(nrhs) : 1;
  sp.rowptr = s->d_rowptr;
  sp.col = s->d_col;
  sp.val = s->d_val64;
  sp.wcol = dw;
  sp.X = (const cx<double> *)dX;
  sp.B = (const cx<double> *)dB;
  sp.Y = nullptr;  // only the norms are kept
  sp.part_rr = pa;
  sp.part_bb = pb;
  HIPCHK(afx::launch_residual_norms(sp, G, norms, s->stream));
  double h[2 * AFX_NOALLOC_MAX_COLS];
  HIPCHK(hipMemcpyAsync(h, norms, 2 * (size_t)nrhs * sizeof(double), hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  for (int c = 0; c < nrhs; c++) res[c] = (h[2 * c + 1] > 0) ? std::sqrt(h[2 * c] / h[2 * c + 1]) : 0.0;
  return AFX_OK;
}

int afx_true_residual_device(afx_system s, int nrhs, const void *dB, const void *dX, const double *omega_scale,
                             double *res) {
  if (!s || !dB || !dX || !res || nrhs <= 0) return fail(AFX_ERR_ARG, "bad arguments");
  int rc = ensure_derived(s);
  if (rc) return rc;
  std::vector<double> wcol, wf;
  std::vector<int> col2f;
  column_freqs(s, nrhs, omega_scale, wcol, wf, col2f);
  const int G = 2 * s->ncu;
  cx<double> *dY = nullptr;
  double *dw = nullptr, *pa = nullptr, *pb = nullptr, *norms = nullptr;
  if ((rc = dalloc(&dY, (size_t)s->n * nrhs)) || (rc = upload(&dw, wcol.data(), nrhs)) ||
      (rc = dalloc(&pa, (size_t)G * nrhs)) || (rc = dalloc(&pb, (size_t)G * nrhs)) ||
      (rc = dalloc(&norms, 2 * (size_t)nrhs))) {
    dfree(dY); dfree(dw); dfree(pa); dfree(pb); dfree(norms);
    return rc;
  }
  afx::SpmvParams sp{};
  sp.n = s->n;
  sp.k = nrhs;
  sp.kp = nrhs > 1 ? next_pow2(nrhs) : 1;
  sp.rowptr = s->d_rowptr;
  sp.col = s->d_col;
  sp.val = s->d_val64;
  sp.wcol = dw;
  sp.X = (const cx<double> *)dX;
  sp.B = (const cx<double> *)dB;
  sp.Y = dY;
  sp.part_rr = pa;
  sp.part_bb = pb;
  hipError_t e = (sp.kp > 256) ? hipErrorInvalidValue : afx::launch_residual_norms(sp, G, norms, s->stream);
  std::vector<double> h(2 * (size_t)nrhs);
  if (e == hipSuccess) e = hipMemcpyAsync(h.data(), norms, h.size() * 8, hipMemcpyDeviceToHost, s->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  dfree(dY); dfree(dw); dfree(pa); dfree(pb); dfree(norms);
  if (e != hipSuccess) return fail(AFX_ERR_HIP, hipGetErrorString(e));
  for (int c = 0; c < nrhs; c++) res[c] = (h[2 * c + 1] > 0) ? std::sqrt(h[2 * c] / h[2 * c + 1]) : 0.0;
  return AFX_OK;
}

int afx_scale_initial_guess_device(afx_system s, int nrhs, const void *dB, void *dX, const double *omega_scale,
                                   double *gamma, double *res_given, double *res_scaled) {
  if (!s || !dB || !dX || nrhs <= 0) return fail(AFX_ERR_ARG, "bad arguments");
  int rc = ensure_derived(s);
  if (rc) return rc;
  const int kp = nrhs > 1 ? next_pow2(nrhs) : 1;
  if (kp > 256) return fail(AFX_ERR_ARG, "at most 256 right-hand sides");
  std::vector<double> wcol, wf;
  std::vector<int> col2f;
  column_freqs(s, nrhs, omega_scale, wcol, wf, col2f);
  const int G = 2 * s->ncu;
  cx<double> *dY = nullptr, *dg = nullptr;
  double *dw = nullptr, *part = nullptr;
  if ((rc = dalloc(&dY, (size_t)s->n * nrhs)) || (rc = upload(&dw, wcol.data(), nrhs)) ||
      (rc = dalloc(&part, (size_t)G * nrhs * 4)) || (rc = dalloc(&dg, (size_t)nrhs))) {
    dfree(dY); dfree(dw); dfree(part); dfree(dg);
    return rc;
  }
  afx::SpmvParams sp{};
  sp.n = s->n;
  sp.k = nrhs;
  sp.kp = kp;
  sp.rowptr = s->d_rowptr;
  sp.col = s->d_col;
  sp.val = s->d_val64;
  sp.wcol = dw;
  sp.X = (const cx<double> *)dX;
  sp.Y = dY;
  std::vector<double> h((size_t)G * nrhs * 4);
  hipError_t e = afx::launch_spmv(sp, s->stream);
  if (e == hipSuccess) e = afx::launch_warm_dots(s->n, nrhs, kp, G, dY, (const cx<double> *)dB, part, s->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(h.data(), part, h.size() * 8, hipMemcpyDeviceToHost, s->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  std::vector<cx<double>> g(nrhs);
  if (e == hipSuccess) {
    for (int c = 0; c < nrhs; c++) {
      double a = 0, bi = 0, yy = 0, bb = 0;  // fixed summation order over blocks
      for (int q = 0; q < G; q++) {
        const double *v = &h[((size_t)q * nrhs + c) * 4];
        a += v[0];
        bi += v[1];
        yy += v[2];
        bb += v[3];
      }
      // ||b - y||^2 = bb - 2 Re(y^H b) + yy; with x -> gamma x, gamma = y^H b / y^H y:
      // ||b - gamma y||^2 = bb - |y^H b|^2 / yy
      const double r0 = bb > 0 ? std::sqrt(std::max(0.0, bb - 2.0 * a + yy) / bb) : 0.0;
      double r1 = 1.0;
      cx<double> gc{0.0, 0.0};
      if (bb > 0 && yy > 0) {
        const double rs = std::sqrt(std::max(0.0, bb - (a * a + bi * bi) / yy) / bb);
        if (rs < 1.0) {  // lowers the zero guess's relative residual: keep the scaled guess
          gc = cx<double>{a / yy, bi / yy};
          r1 = rs;
        }
      }
      if (bb == 0) r1 = 0.0;
      g[c] = gc;
      if (gamma) {
        gamma[2 * c] = gc.r;
        gamma[2 * c + 1] = gc.i;
      }
      if (res_given) res_given[c] = r0;
      if (res_scaled) res_scaled[c] = r1;
    }
    e = hipMemcpyAsync(dg, g.data(), nrhs * sizeof(cx<double>), hipMemcpyHostToDevice, s->stream);
  }
  if (e == hipSuccess) e = afx::launch_scale_cols_cx(s->n, nrhs, dg, (cx<double> *)dX, s->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  dfree(dY); dfree(dw); dfree(part); dfree(dg);
  if (e != hipSuccess) return fail(AFX_ERR_HIP, hipGetErrorString(e));
  return AFX_OK;
}

}  // extern "C"

// -----------------------------------------------------------------------------------------
namespace {
const int KLH[6][2] = {{0, 1}, {1, 2}, {0, 2}, {0, 3}, {1, 3}, {2, 3}};
const int LMNH[4][3] = {{1, 2, 3}, {0, 3, 2}, {0, 1, 3}, {0, 2, 1}};
const int CROTH[4][6] = {{0, 1, 0, 0, -1, 1}, {0, 0, -1, 1, 0, -1}, {1, 0, 0, -1, 1, 0}, {-1, -1, 1, 0, 0, 0}};

inline double in3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline void out3(const double *a, const double *b, double *c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}
inline double vol3(const double *a, const double *b, const double *c) {
  double t[3];
  out3(a, b, t);
  return 1.0 / 6.0 * in3(t, c);
}
// gradnodebasisfun (m_fem_util.f90:12-40)
bool grad_h(const double x[4][3], double gn[4][3], double &v) {
  double x34[3], x13[3], x23[3], a[3], b[3], c[3];
  for (int d = 0; d < 3; d++) {
    x34[d] = x[3][d] - x[2][d];
    x13[d] = x[0][d] - x[2][d];
    x23[d] = -(x[2][d] - x[1][d]);
  }
  v = vol3(x34, x13, x23);
  if (!(v > 0)) return false;
  const double s = 1.0 / 6.0 / v;
  const int P[4][4] = {{3, 1, 2, 1}, {2, 0, 3, 0}, {3, 0, 1, 0}, {1, 0, 2, 0}};
  for (int k = 0; k < 4; k++) {
    for (int d = 0; d < 3; d++) {
      a[d] = x[P[k][0]][d] - x[P[k][1]][d];
      b[d] = x[P[k][2]][d] - x[P[k][3]][d];
    }
    out3(a, b, c);
    for (int d = 0; d < 3; d++) gn[k][d] = s * c[d];
  }
  return true;
}
// nodebasisfun (m_fem_util.f90:56-83)
void lambda_h(const double x[4][3], const double *p, double a[4]) {
  double v0[3], v1[3], x12[3], x13[3], x14[3], x24[3], x23[3];
  for (int d = 0; d < 3; d++) {
    v0[d] = -(x[0][d] - p[d]);
    v1[d] = -(x[1][d] - p[d]);
    x12[d] = x[1][d] - x[0][d];
    x13[d] = x[2][d] - x[0][d];
    x14[d] = x[3][d] - x[0][d];
    x24[d] = x[3][d] - x[1][d];
    x23[d] = x[2][d] - x[1][d];
  }
  const double a5 = vol3(x12, x13, x14);
  a[3] = vol3(x12, x13, v0) / a5;
  a[2] = vol3(x14, x12, v0) / a5;
  a[1] = vol3(x13, x14, v0) / a5;
  a[0] = vol3(x24, x23, v1) / a5;
}
}  // namespace

extern "C" {

int afx_kernel_timing(afx_system s, int reset, long long *launches, long long *iterations, double *kernel_ms) {
  if (!s) return fail(AFX_ERR_ARG, "null system");
  if (launches) *launches = s->prof_launches;
  if (iterations) *iterations = s->prof_iters;
  if (kernel_ms) *kernel_ms = s->prof_ms;
  if (reset) {
    s->prof_launches = 0;
    s->prof_iters = 0;
    s->prof_ms = 0.0;
  }
  return AFX_OK;
}

// Right-hand-side block width for k columns: 20 FP64 iterations of the k-column kernel against
// k times 20 iterations of one column with the single-column kernel; the faster is kept.
static int autotune_width(afx_system s, int k, int kind, int verbose) {
  const size_t nk = (size_t)s->n * k;
  std::vector<cx<double>> hb(nk, cx<double>{1.0, 0.5});
  cx<double> *dB = nullptr, *dX = nullptr;
  int rc;
  if ((rc = upload(&dB, hb.data(), nk)) || (rc = dalloc(&dX, nk))) {
    dfree(dB);
    return rc;
  }
  afx_opts o;
  afx_default_opts(&o);
  o.precond = kind;
  o.sigma_diag = s->pc_sd;
  o.sigma = s->pc_sig;
  o.tol = 0.0;
  o.maxit = 20;
  o.check_true = 0;
  // [0] one column (times k), [1] k columns, [2] k stacked copies (only where the copies' rows
  // of a block fit in LDS; built before timing)
  double t[3] = {1e300, 1e300, 1e300};
  std::vector<double> wcol, wf;
  std::vector<int> col2f;
  column_freqs(s, k, nullptr, wcol, wf, col2f);
  // the copies are only tried when a block's share of a group's rows can fit in LDS (checked
  // exactly once built): a sweep without it is several times slower
  const std::string keep_err = g_err;
  const int grp = stack_group(s, k);
  // stacked copies carry no node incidence: not a width for the nodal correction
  const bool can_stack = grp >= 2 && !(kind & AFX_PC_GRAD) && (kind & 0xff) != AFX_PC_NODEBLK &&
                         ensure_stack(s, grp, std::vector<double>(wcol.begin(), wcol.begin() + grp)) == AFX_OK &&
                         lds_rows_for<double>(s->stack, false) > 0;
  g_err = keep_err;
  static const char *const names[3] = {"one column x k", "k-column kernel", "k stacked copies"};
  for (int mode = 0; mode < 3; mode++) {
    if (mode == 2 && !can_stack) continue;
    s->batch_multi[k] = mode;
    if (hipMemset(dX, 0, nk * sizeof(cx<double>)) != hipSuccess) break;
    (void)hipDeviceSynchronize();
    auto t0 = std::chrono::steady_clock::now();
    std::vector<afx_stats> st(k);
    const int r2 = solve_device_impl(s, mode == 0 ? 1 : k, dB, dX, nullptr, &o, st.data());
    (void)hipDeviceSynchronize();
    const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (r2 == AFX_OK || r2 == AFX_ERR_MAXIT) t[mode] = mode == 0 ? dt * k : dt;
    if (verbose) fprintf(stderr, "afx_autotune k=%d %s: %.3f ms / 20 it\n", k, names[mode], t[mode] * 1e3);
  }
  int best = 0;
  for (int mode = 1; mode < 3; mode++)
    if (t[mode] < t[best]) best = mode;
  s->batch_multi[k] = best;
  dfree(dB);
  dfree(dX);
  return AFX_OK;
}

int afx_autotune(afx_system s, int nrhs, int verbose) {
  // Rows per wavefront (lanes per row `lpr`) of the single-column kernel, per preconditioner
  // precision: time 20 iterations of each candidate on this system. The grid is the sweep
  // partition (one block per row range), so only the lane mapping is tuned.
  if (!s || nrhs <= 0 || nrhs > 256) return fail(AFX_ERR_ARG, "bad arguments");
  if (!s->have_values) return fail(AFX_ERR_ARG, "no matrix values");
  const int kind = s->pc_valid ? s->pc_kind : AFX_PC_SSOR;
  if (nrhs > 1) return autotune_width(s, nrhs, kind, verbose);
  const size_t nk = (size_t)s->n;
  std::vector<cx<double>> hb(nk, cx<double>{1.0, 0.5});
  cx<double> *dB = nullptr, *dX = nullptr;
  int rc;
  if ((rc = upload(&dB, hb.data(), nk)) || (rc = dalloc(&dX, nk))) {
    dfree(dB);
    return rc;
  }
  afx_opts o;
  afx_default_opts(&o);
  o.precond = kind;
  o.sigma_diag = s->pc_sd;
  o.sigma = s->pc_sig;
  o.tol = 0.0;
  o.maxit = 20;
  o.check_true = 0;
  for (int prec = 0; prec < 2; prec++) {
    const bool fp64 = prec == 0;
    double best = 1e300;
    int best_lpr = 8;
    for (int lpr : {2, 4, 8, 16}) {
      (fp64 ? s->lpr64 : s->lpr32) = lpr;
      if (hipMemset(dX, 0, nk * sizeof(cx<double>)) != hipSuccess) break;
      o.precision = fp64 ? AFX_FP64 : AFX_MIXED;
      (void)hipDeviceSynchronize();
      auto t0 = std::chrono::steady_clock::now();
      std::vector<afx_stats> st(1);
      int r2 = solve_device_impl(s, 1, dB, dX, nullptr, &o, st.data());
      (void)hipDeviceSynchronize();
      const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      if (r2 != AFX_OK && r2 != AFX_ERR_MAXIT) continue;
      if (verbose) fprintf(stderr, "afx_autotune %s lpr=%d: %.3f ms / 20 it\n", fp64 ? "fp64" : "fp32-pc", lpr, dt * 1e3);
      if (dt < best) {
        best = dt;
        best_lpr = lpr;
      }
    }
    (fp64 ? s->lpr64 : s->lpr32) = best_lpr;
  }
  dfree(dB);
  dfree(dX);
  return AFX_OK;
}

// The drop-in keeps the system of its last call: a caller that solves one frequency after
// another on the same mesh (forward_bxyz's solveCG) repeats the pattern analysis only when the
// index arrays change. Like the reference's SAVEd DD, SCALE and IFLAG (COCG13_single.f90:39-44,
// 57-62), the kept system's preconditioner and the scaling vector survive between calls: NSET = 0
// reuses both, NSET = 1 reuses the scaling of the last NSET = 2 call. Calls are serialised, as the
// reference's SAVEd routine requires.
static std::mutex g_drop_mu;
static afx_system g_drop_sys = nullptr;
static bool g_drop_iflag = false;       // IFLAG: a call with NSET >= 1 has run (DD and SCALE exist)
static std::vector<double> g_drop_scale;  // SCALE (NP entries, 1 until an NSET = 2 call)
static int g_drop_pc = -1;              // PRECOND of the setup that DD holds

static bool same_pattern(afx_system s, int n, const int *INU, const int *IAU, const int *INL, const int *IAL) {
  if (!s || s->n != n || s->nU != INU[n] || s->nL != INL[n]) return false;
  return std::memcmp(s->INU.data(), INU, sizeof(int) * ((size_t)n + 1)) == 0 &&
         std::memcmp(s->INL.data(), INL, sizeof(int) * ((size_t)n + 1)) == 0 &&
         std::memcmp(s->IAU.data(), IAU, sizeof(int) * (size_t)s->nU) == 0 &&
         std::memcmp(s->IAL.data(), IAL, sizeof(int) * (size_t)s->nL) == 0;
}

int afx_cocg13_single_release(void) {
  std::lock_guard<std::mutex> lk(g_drop_mu);
  int rc = g_drop_sys ? afx_destroy(g_drop_sys) : AFX_OK;
  g_drop_sys = nullptr;
  g_drop_iflag = false;
  g_drop_scale.clear();
  g_drop_pc = -1;
  return rc;
}

int afx_cocg13_single(int NP, int NPL, int NPU, double *D, double *AL, const int *INL, const int *IAL, double *AU,
                      const int *INU, const int *IAU, double *B, double *X, const char *PRECOND, double SIGMA_DIAG,
                      double SIGMA, double *RESID, int *ITER, int *ERROR, int NSET) {
  if (!D || !B || !X || !RESID || !ITER || !ERROR || !INU || !INL) return fail(AFX_ERR_ARG, "null argument");
  if (NP <= 0 || NSET < 0 || NSET > 2) return fail(AFX_ERR_ARG, "bad NP or NSET");
  *ERROR = 0;
  if (INU[NP] != NPU || INL[NP] != NPL) return fail(AFX_ERR_ARG, "NPU/NPL do not match INU/INL");
  std::lock_guard<std::mutex> lk(g_drop_mu);
  // the reference allocates DD and SCALE once, at its first call, and cannot change NP after it;
  // a different NP here starts the saved state afresh
  if (g_drop_iflag && (int)g_drop_scale.size() != NP) {
    g_drop_iflag = false;
    g_drop_pc = -1;
  }
  const int pc = pc_from_string(PRECOND);
  if (NSET == 0) {
    // COCG13_single.f90:49-52: nothing saved yet. Here the kept preconditioner must also belong to
    // the same pattern and preconditioner kind (the reference would apply its DD to any matrix)
    if (!g_drop_iflag || g_drop_pc != pc || !same_pattern(g_drop_sys, NP, INU, IAU, INL, IAL) ||
        !g_drop_sys->pc_valid) {
      *ERROR = AFX_ERR_BADOPT;
      return AFX_ERR_BADOPT;
    }
  }
  if (!g_drop_iflag) {  // :57-62
    g_drop_scale.assign(NP, 1.0);
    g_drop_iflag = true;
  }
  std::vector<double> &scale = g_drop_scale;
  if (NSET == 2) {  // COCG13_single.f90:77-97, in place on the caller's arrays
    for (int i = 0; i < NP; i++) scale[i] = 1.0 / std::sqrt(std::hypot(D[2 * i], D[2 * i + 1]));
    for (int j = 0; j < NP; j++) {
      const double ad = std::hypot(D[2 * j], D[2 * j + 1]);
      const double dr = D[2 * j], di = D[2 * j + 1], den = dr * dr + di * di;
      D[2 * j] = ad * dr / den;
      D[2 * j + 1] = -ad * di / den;
      for (int q = INU[j]; q < INU[j + 1]; q++) {
        const double f = scale[IAU[q] - 1] * scale[j];
        AU[2 * q] *= f;
        AU[2 * q + 1] *= f;
      }
      for (int q = INL[j]; q < INL[j + 1]; q++) {
        const double f = scale[IAL[q] - 1] * scale[j];
        AL[2 * q] *= f;
        AL[2 * q + 1] *= f;
      }
    }
  }
  // :98-100 B scaled only when NSET >= 1 (the saved SCALE, 1 unless an NSET = 2 call set it)
  const bool scaled_b = NSET >= 1;
  if (scaled_b)
    for (int i = 0; i < NP; i++) {
      B[2 * i] *= scale[i];
      B[2 * i + 1] *= scale[i];
    }
  auto unscale_b = [&]() {
    if (scaled_b)
      for (int i = 0; i < NP; i++) {
        B[2 * i] /= scale[i];
        B[2 * i + 1] /= scale[i];
      }
  };
  bool fresh = false;
  if (!same_pattern(g_drop_sys, NP, INU, IAU, INL, IAL)) {
    if (g_drop_sys) afx_destroy(g_drop_sys);
    g_drop_sys = nullptr;
    g_drop_pc = -1;
    int rc0 = afx_create(NP, INU, IAU, INL, IAL, &g_drop_sys);
    if (rc0) {
      g_drop_sys = nullptr;
      unscale_b();
      return rc0;
    }
    fresh = true;
  }
  afx_system s = g_drop_sys;
  int rc = AFX_OK;
  if (NSET == 0) {
    // the current values with the saved DD: the sweeps' DD (dpos) and the preconditioner state
    // are left as they are, only the matrix values are replaced
    rc = afx_set_values(s, D, AU, AL);
    if (rc == AFX_OK) rc = ensure_derived(s);
    if (rc == AFX_OK) s->pc_valid = true;
  } else {
    rc = afx_set_values(s, D, AU, AL);
    if (rc == AFX_OK && fresh) rc = afx_autotune(s, 1, 0);  // lane width for this pattern, once
  }
  afx_opts o;
  afx_default_opts(&o);
  o.precond = pc;
  o.sigma_diag = SIGMA_DIAG;
  o.sigma = SIGMA;
  o.tol = *RESID;
  o.maxit = *ITER;
  o.nset = NSET == 0 ? 0 : 1;
  o.check_true = 0;
  afx_stats st{};
  if (rc == AFX_OK) rc = afx_solve(s, 1, B, NP, X, NP, &o, &st);
  if (rc != AFX_OK && rc != AFX_ERR_MAXIT && rc != AFX_ERR_BREAKDOWN) {
    unscale_b();
    if (rc == AFX_ERR_BADOPT) *ERROR = AFX_ERR_BADOPT;
    return rc;
  }
  if (NSET >= 1) g_drop_pc = pc;
  for (int i = 0; i < NP; i++) {  // :326-329, also for NSET = 0 (B divided by the saved SCALE)
    X[2 * i] *= scale[i];
    X[2 * i + 1] *= scale[i];
    B[2 * i] /= scale[i];
    B[2 * i + 1] /= scale[i];
  }
  *RESID = st.resid;
  *ITER = st.iter;
  *ERROR = st.error;
  return rc;
}

// ---- solvePARDISO / solvePARDISOinv shaped entries -------------------------------------------
// The system of the last call is kept and reused while the pattern is the same (a per-frequency
// caller repeats the pattern analysis only when the mesh changes); calls are serialised.
static std::mutex g_pds_mu;
static afx_system g_pds_sys = nullptr;

// The solvePARDISO-shaped entries promise the recomputed FP64 true residual tol: a solve whose
// recurrence converged but whose true residual stayed above tol after the restarts (a right-hand
// side beyond what FP64 can resolve) reports AFX_ERR_MAXIT, as one that ran out of iterations.
static int pardiso_status(int rc, const afx_stats *st, int k, double tol) {
  if (rc != AFX_OK) return rc;
  for (int c = 0; c < k; c++)
    if (!(st[c].true_resid <= tol)) return AFX_ERR_MAXIT;
  return AFX_OK;
}

// Pattern, values, node incidence and preconditioner of the kept system; *pc receives the
// preconditioner: node blocks with the nodal correction when the edges' end nodes are given,
// else SSOR.
static int pardiso_system(int doftot, const int *INU, const int *IAU, const int *INL, const int *IAL, const double *D,
                          const double *AU, const double *AL, int node, const int *line,
                          const unsigned char *line_bc, int *pc) {
  if (!INU || !INL || !D || doftot <= 0) return fail(AFX_ERR_ARG, "bad arguments");
  bool fresh = false;
  if (!same_pattern(g_pds_sys, doftot, INU, IAU, INL, IAL)) {
    if (g_pds_sys) afx_destroy(g_pds_sys);
    g_pds_sys = nullptr;
    int rc = afx_create(doftot, INU, IAU, INL, IAL, &g_pds_sys);
    if (rc) {
      g_pds_sys = nullptr;
      return rc;
    }
    fresh = true;
  }
  afx_system s = g_pds_sys;
  int rc = afx_set_values(s, D, AU, AL);
  if (rc) return rc;
  *pc = AFX_PC_SSOR;
  if (line && node > 0) {
    if ((rc = afx_set_nodes(s, node, line, line_bc))) return rc;
    *pc = AFX_PC_NODEBLK | AFX_PC_GRAD;
  }
  if ((rc = afx_precond_setup(s, *pc, 1.0, 0.0))) return rc;
  if (fresh) rc = afx_autotune(s, 1, 0);
  return rc;
}

int afx_solve_pardiso_release(void) {
  std::lock_guard<std::mutex> lk(g_pds_mu);
  int rc = g_pds_sys ? afx_destroy(g_pds_sys) : AFX_OK;
  g_pds_sys = nullptr;
  return rc;
}

int afx_solve_pardiso(int doftot, int nsr, const int *INU, const int *IAU, const int *INL, const int *IAL,
                      const double *D, const double *AU, const double *AL, int node, const int *line,
                      const unsigned char *line_bc, const double *rhs, double *xout, double tol) {
  if (nsr <= 0 || !rhs || !xout || !(tol > 0.0)) return fail(AFX_ERR_ARG, "bad arguments");
  std::lock_guard<std::mutex> lk(g_pds_mu);
  int pc = 0;
  int rc = pardiso_system(doftot, INU, IAU, INL, IAL, D, AU, AL, node, line, line_bc, &pc);
  if (rc) return rc;
  afx_opts o;
  afx_default_opts(&o);
  o.precond = pc;
  o.tol = tol;
  o.maxit = 400000;
  o.check_true = 1;
  o.max_outer = 4;
  std::fill(xout, xout + 2 * (size_t)doftot * nsr, 0.0);
  std::vector<afx_stats> st(nsr);
  return pardiso_status(afx_solve(g_pds_sys, nsr, rhs, doftot, xout, doftot, &o, st.data()), st.data(), nsr, tol);
}

// Adjoint solutions of the observation rows (solvePARDISOinv.f90:176-196): for every component
// c < ncomp with flags[c] != 0 (all when flags is null), u_o = A^-1 P_o^T for the nobs rows of
// PTR(c) in blocks of at most 64 columns, kept where |u| > 1e-10 as the CRS of u^T
// (convcomp_full2crs and its transposition; m_matrix.f90:405-451). *ret becomes AFX_ERR_MAXIT when
// some column missed the tolerance; a component needing more than ut_cap entries is reported
// after all are counted.
static int adjoint_components(afx_system s, const afx_opts &o, int doftot, int ncomp, int nobs, const int *flags,
                              const int *ptr_stack, const int *ptr_base, const int *ptr_item, const double *ptr_val,
                              int ut_cap, int *ut_stack, int *ut_ntot, int *ut_item, double *ut_val, int *ret) {
  for (int c = 0; c < ncomp; c++) ut_ntot[c] = 0;
  if (nobs == 0) return AFX_OK;
  if (!ptr_stack || !ptr_base || !ptr_item || !ptr_val || !ut_stack || (ut_cap > 0 && (!ut_item || !ut_val)))
    return fail(AFX_ERR_ARG, "bad arguments");
  const double threshold = 1.0e-10;
  const int blk = std::min(nobs, 64);
  cx<double> *dB = nullptr, *dX = nullptr;
  int rc;
  if ((rc = dalloc(&dB, (size_t)doftot * blk)) || (rc = dalloc(&dX, (size_t)doftot * blk))) {
    dfree(dB);
    return rc;
  }
  std::vector<cx<double>> U((size_t)doftot * blk);
  bool short_cap = false;
  for (int c = 0; c < ncomp && rc == AFX_OK; c++) {
    if (flags && !flags[c]) continue;
    const int *stk = ptr_stack + (size_t)c * (nobs + 1);
    int *ust = ut_stack + (size_t)c * (nobs + 1);
    int *uit = ut_item + (size_t)c * ut_cap;
    double *uva = ut_val + 2 * (size_t)c * ut_cap;
    long long cnt = 0;
    ust[0] = 0;
    for (int o0 = 0; o0 < nobs && rc == AFX_OK; o0 += blk) {
      const int k = std::min(blk, nobs - o0);
      // the rows o0 .. o0 + k of PTR(c) as a CRS of their own
      std::vector<int> st2(k + 1);
      for (int q = 0; q <= k; q++) st2[q] = stk[o0 + q] - stk[o0];
      const int *it2 = ptr_item + ptr_base[c] + stk[o0];
      const double *va2 = ptr_val + ptr_base[c] + stk[o0];
      if ((rc = afx_adjoint_rhs_device(s, k, st2.data(), it2, va2, dB))) break;
      if (hipMemsetAsync(dX, 0, (size_t)doftot * k * sizeof(cx<double>), s->stream) != hipSuccess) {
        rc = fail(AFX_ERR_HIP, "hipMemsetAsync");
        break;
      }
      std::vector<afx_stats> st(k);
      const int r = pardiso_status(afx_solve_device(s, k, dB, dX, nullptr, &o, st.data()), st.data(), k, o.tol);
      if (r != AFX_OK && r != AFX_ERR_MAXIT) {
        rc = r;
        break;
      }
      if (r == AFX_ERR_MAXIT) *ret = AFX_ERR_MAXIT;
      if (hipMemcpy(U.data(), dX, (size_t)doftot * k * sizeof(cx<double>), hipMemcpyDeviceToHost) != hipSuccess) {
        rc = fail(AFX_ERR_HIP, "hipMemcpy");
        break;
      }
      for (int q = 0; q < k; q++) {  // row o0 + q of u^T: edges in ascending order
        for (int e = 0; e < doftot; e++) {
          const cx<double> v = U[(size_t)e * k + q];
          if (std::hypot(v.r, v.i) > threshold) {
            if (cnt < ut_cap) {
              uit[cnt] = e + 1;
              uva[2 * cnt] = v.r;
              uva[2 * cnt + 1] = v.i;
            }
            cnt++;
          }
        }
        ust[o0 + q + 1] = (int)std::min<long long>(cnt, 0x7fffffffLL);
      }
    }
    if (cnt > 0x7fffffffLL) rc = fail(AFX_ERR_ARG, "ut: more than 2^31 entries in one component");
    ut_ntot[c] = (int)std::min<long long>(cnt, 0x7fffffffLL);
    if (cnt > ut_cap) short_cap = true;
  }
  dfree(dB);
  dfree(dX);
  if (rc) return rc;
  if (short_cap) return fail(AFX_ERR_ARG, "ut_cap too small: ut_ntot holds the entries each component needs");
  return AFX_OK;
}

// options of the solvePARDISO-shaped entries: FP64 COCG to the recomputed true residual tol
static afx_opts pardiso_opts(int pc, double tol) {
  afx_opts o;
  afx_default_opts(&o);
  o.precond = pc;
  o.tol = tol;
  o.maxit = 400000;
  o.check_true = 1;
  o.max_outer = 4;
  return o;
}

int afx_solve_pardiso_inv(int doftot, int nsr, const int *INU, const int *IAU, const int *INL, const int *IAL,
                          const double *D, const double *AU, const double *AL, int node, const int *line,
                          const unsigned char *line_bc, const double *rhs, double *xout, int nobs,
                          const int *iflag_comp, const int *ptr_stack, const int *ptr_base, const int *ptr_item,
                          const double *ptr_val, int ut_cap, int *ut_stack, int *ut_ntot, int *ut_item, double *ut_val,
                          double tol) {
  if (nsr < 0 || nobs < 0 || !iflag_comp || !ut_ntot || !(tol > 0.0) || ut_cap < 0)
    return fail(AFX_ERR_ARG, "bad arguments");
  std::lock_guard<std::mutex> lk(g_pds_mu);
  int pc = 0;
  int rc = pardiso_system(doftot, INU, IAU, INL, IAL, D, AU, AL, node, line, line_bc, &pc);
  if (rc) return rc;
  afx_system s = g_pds_sys;
  const afx_opts o = pardiso_opts(pc, tol);
  int ret = AFX_OK;
  // the forward solve (solvePARDISOinv.f90:157-164)
  if (nsr > 0) {
    if (!rhs || !xout) return fail(AFX_ERR_ARG, "bad arguments");
    std::fill(xout, xout + 2 * (size_t)doftot * nsr, 0.0);
    std::vector<afx_stats> st(nsr);
    rc = pardiso_status(afx_solve(s, nsr, rhs, doftot, xout, doftot, &o, st.data()), st.data(), nsr, tol);
    if (rc != AFX_OK && rc != AFX_ERR_MAXIT) return rc;
    ret = rc;
  }
  rc = adjoint_components(s, o, doftot, 5, nobs, iflag_comp, ptr_stack, ptr_base, ptr_item, ptr_val, ut_cap, ut_stack,
                          ut_ntot, ut_item, ut_val, &ret);
  return rc ? rc : ret;
}

int afx_solve_pardiso_jointinv(int act, int mt, int doftot, int nsr, const int *INU, const int *IAU, const int *INL,
                               const int *IAL, const double *D, const double *AU, const double *AL, int node,
                               const int *line, const unsigned char *line_bc, const double *rhs, const double *rhs_mt,
                               double *xout, double *xout_mt, int nobs, const int *iflag_comp, const int *ptr_stack,
                               const int *ptr_base, const int *ptr_item, const double *ptr_val, int nobs_mt,
                               const int *ptr_mt_stack, const int *ptr_mt_base, const int *ptr_mt_item,
                               const double *ptr_mt_val, int ut_cap, int *ut_stack, int *ut_ntot, int *ut_item,
                               double *ut_val, int ut_mt_cap, int *ut_mt_stack, int *ut_mt_ntot, int *ut_mt_item,
                               double *ut_mt_val, double tol) {
  if (nsr < 0 || nobs < 0 || nobs_mt < 0 || !iflag_comp || !ut_ntot || !ut_mt_ntot || !(tol > 0.0) || ut_cap < 0 ||
      ut_mt_cap < 0 || (act && nsr > 0 && (!rhs || !xout)) || (mt && (!rhs_mt || !xout_mt)))
    return fail(AFX_ERR_ARG, "bad arguments");
  for (int c = 0; c < 5; c++) ut_ntot[c] = 0;
  for (int c = 0; c < 4; c++) ut_mt_ntot[c] = 0;
  // no flagged component with rows: the reference leaves before the factorisation (:56-64)
  if (nobs == 0) return AFX_OK;
  std::lock_guard<std::mutex> lk(g_pds_mu);
  int pc = 0;
  int rc = pardiso_system(doftot, INU, IAU, INL, IAL, D, AU, AL, node, line, line_bc, &pc);
  if (rc) return rc;
  afx_system s = g_pds_sys;
  const afx_opts o = pardiso_opts(pc, tol);
  int ret = AFX_OK;
  // the active-source and the two MT forward solves (:173-197) share one launch: nsr + 2 columns
  const int na = act ? nsr : 0, nm = mt ? 2 : 0, nc = na + nm;
  if (nc > 0) {
    std::vector<double> B(2 * (size_t)doftot * nc), X(2 * (size_t)doftot * nc, 0.0);
    if (na) std::copy(rhs, rhs + 2 * (size_t)doftot * na, B.begin());
    if (nm) std::copy(rhs_mt, rhs_mt + 2 * (size_t)doftot * nm, B.begin() + 2 * (size_t)doftot * na);
    std::vector<afx_stats> st(nc);
    rc = pardiso_status(afx_solve(s, nc, B.data(), doftot, X.data(), doftot, &o, st.data()), st.data(), nc, tol);
    if (rc != AFX_OK && rc != AFX_ERR_MAXIT) return rc;
    ret = rc;
    if (na) std::copy(X.begin(), X.begin() + 2 * (size_t)doftot * na, xout);
    if (nm) std::copy(X.begin() + 2 * (size_t)doftot * na, X.end(), xout_mt);
  }
  // the active-source components (:212-236), then the 4 MT components bx, by, ex, ey (:246-290)
  if (act && (rc = adjoint_components(s, o, doftot, 5, nobs, iflag_comp, ptr_stack, ptr_base, ptr_item, ptr_val,
                                      ut_cap, ut_stack, ut_ntot, ut_item, ut_val, &ret)))
    return rc;
  if (mt && (rc = adjoint_components(s, o, doftot, 4, nobs_mt, nullptr, ptr_mt_stack, ptr_mt_base, ptr_mt_item,
                                     ptr_mt_val, ut_mt_cap, ut_mt_stack, ut_mt_ntot, ut_mt_item, ut_mt_val, &ret)))
    return rc;
  return ret;
}

int afx_solve_cg(int doftot, int NPL, int NPU, const double *D, const double *AL, const int *INL, const int *IAL,
                 const double *AU, const int *INU, const int *IAU, const double *rhs, double *X) {
  // solveCG (forward_bxyz_mpi.f90:70-94)
  std::vector<double> d(D, D + 2 * (size_t)doftot), al(AL, AL + 2 * (size_t)NPL), au(AU, AU + 2 * (size_t)NPU),
      b(rhs, rhs + 2 * (size_t)doftot);
  for (int i = 0; i < 2 * doftot; i++) X[i] = 0.0;
  double resid = 1.e-4f;  // RESID=1.e-4 is a default-real literal
  int iter = 3000, error = 0;
  char pc[21] = "SSOR                ";
  int rc = afx_cocg13_single(doftot, NPL, NPU, d.data(), al.data(), INL, IAL, au.data(), INU, IAU, b.data(), X, pc,
                             1.0, 0.0, &resid, &iter, &error, 1);
  return rc;
}

// The point-location grid of the mesh: buckets of about 8 tets (by centroid) over the node
// bounding box, each bucket's tets in increasing id, and the largest centroid-to-vertex extent
// per axis (the search radius).
static int build_point_grid(afx_system s) {
  const int ntet = s->ntet, node = s->node;
  afx::PointGrid g{};
  double lo[3], hi[3];
  for (int d = 0; d < 3; d++) lo[d] = hi[d] = s->h_xyz[d];
  for (int v = 0; v < node; v++)
    for (int d = 0; d < 3; d++) {
      lo[d] = std::min(lo[d], s->h_xyz[(size_t)v * 3 + d]);
      hi[d] = std::max(hi[d], s->h_xyz[(size_t)v * 3 + d]);
    }
  double ext[3], vol = 1.0;
  for (int d = 0; d < 3; d++) {
    ext[d] = hi[d] - lo[d];
    if (!(ext[d] > 0.0)) ext[d] = 1.0;
    vol *= ext[d];
  }
  const double nb = std::max(1.0, ntet / 8.0);
  const double h0 = std::cbrt(vol / nb);
  for (int d = 0; d < 3; d++) {
    g.n[d] = (int)std::max(1.0, std::min(1024.0, std::ceil(ext[d] / h0)));
    g.h[d] = ext[d] / g.n[d];
    g.lo[d] = lo[d];
    g.R[d] = 0.0;
  }
  const long long nbk = (long long)g.n[0] * g.n[1] * g.n[2];
  std::vector<int> key(ntet), bptr(nbk + 1, 0), btet(ntet);
  for (int t = 0; t < ntet; t++) {
    double x[4][3], c[3] = {0.0, 0.0, 0.0};
    for (int j = 0; j < 4; j++)
      for (int d = 0; d < 3; d++) {
        x[j][d] = s->h_xyz[(size_t)(s->h_n4[(size_t)t * 4 + j] - 1) * 3 + d];
        c[d] += 0.25 * x[j][d];
      }
    int ib[3];
    for (int d = 0; d < 3; d++) {
      for (int j = 0; j < 4; j++) g.R[d] = std::max(g.R[d], std::fabs(x[j][d] - c[d]));
      ib[d] = (int)std::max(0.0, std::min(std::floor((c[d] - g.lo[d]) / g.h[d]), (double)(g.n[d] - 1)));
    }
    key[t] = (ib[2] * g.n[1] + ib[1]) * g.n[0] + ib[0];
    bptr[key[t] + 1]++;
  }
  for (long long b = 0; b < nbk; b++) bptr[b + 1] += bptr[b];
  std::vector<int> fill(bptr.begin(), bptr.end() - 1);
  for (int t = 0; t < ntet; t++) btet[fill[key[t]]++] = t;  // increasing id within a bucket
  dfree(s->d_bptr);
  dfree(s->d_btet);
  int rc;
  if ((rc = upload(&s->d_bptr, bptr.data(), (size_t)nbk + 1)) || (rc = upload(&s->d_btet, btet.data(), (size_t)ntet)))
    return rc;
  s->pgrid = g;
  s->have_pgrid = true;
  return AFX_OK;
}

int afx_prepobscoeff(afx_system s, int nobs, const double *xyzobs, int *item, double *val_e, double *val_b,
                     int *iele) {
  if (!s || !xyzobs || !item || !val_e || !val_b) return fail(AFX_ERR_ARG, "null argument");
  if (!s->have_mesh) return fail(AFX_ERR_ARG, "afx_set_mesh first");
  if (nobs <= 0) return AFX_OK;
  int rc;
  if (!s->have_pgrid && (rc = build_point_grid(s))) return rc;
  double *dp = nullptr;
  int *dfirst = nullptr;
  if ((rc = upload(&dp, xyzobs, (size_t)nobs * 3)) || (rc = dalloc(&dfirst, nobs))) {
    dfree(dp);
    return rc;
  }
  std::vector<int> first(nobs, 0x7fffffff);
  hipError_t e = afx::launch_findelement_grid(nobs, dp, s->pgrid, s->d_bptr, s->d_btet, s->d_xyz, s->d_n4, dfirst,
                                              s->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(first.data(), dfirst, nobs * sizeof(int), hipMemcpyDeviceToHost, s->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  dfree(dp);
  dfree(dfirst);
  if (e != hipSuccess) return fail(AFX_ERR_HIP, hipGetErrorString(e));
  for (int o = 0; o < nobs; o++) {
    const int t = first[o];
    if (t == 0x7fffffff) return fail(AFX_ERR_OUTSIDE, "GEGEGE! No corresponding element was found");
    if (iele) iele[o] = t;
    double x[4][3], gn[4][3], v, lam[4], w[6][3], wf[4][3], wfe[6][3];
    for (int j = 0; j < 4; j++)
      for (int d = 0; d < 3; d++) x[j][d] = s->h_xyz[(size_t)(s->h_n4[(size_t)t * 4 + j] - 1) * 3 + d];
    grad_h(x, gn, v);
    lambda_h(x, &xyzobs[o * 3], lam);
    for (int ed = 0; ed < 6; ed++) {  // EDGEBASISFUN (m_fem_edge_util.f90:198-218)
      const int k = KLH[ed][0], l = KLH[ed][1];
      for (int d = 0; d < 3; d++) w[ed][d] = lam[k] * gn[l][d] - lam[l] * gn[k][d];
    }
    for (int f = 0; f < 4; f++) {  // FACEBASISFUN (:163-195)
      const int l = LMNH[f][0], m = LMNH[f][1], nn = LMNH[f][2];
      double c1[3], c2[3], c3[3];
      out3(gn[m], gn[nn], c1);
      out3(gn[l], gn[m], c2);
      out3(gn[nn], gn[l], c3);
      for (int d = 0; d < 3; d++) wf[f][d] = 2.0 * (lam[l] * c1[d] + lam[nn] * c2[d] + lam[m] * c3[d]);
    }
    for (int ed = 0; ed < 6; ed++)
      for (int d = 0; d < 3; d++) {
        double acc = 0.0;
        for (int f = 0; f < 4; f++) acc += wf[f][d] * (double)CROTH[f][ed];
        wfe[ed][d] = acc;
      }
    int order[6] = {0, 1, 2, 3, 4, 5}, ids[6];
    for (int ed = 0; ed < 6; ed++) ids[ed] = std::abs(s->n6line[(size_t)t * 6 + ed]);
    std::sort(order, order + 6, [&](int a, int b) { return ids[a] < ids[b]; });  // SORT_INDEX by edge id
    for (int j = 0; j < 6; j++) {
      const int ed = order[j];
      const double sg = (s->n6line[(size_t)t * 6 + ed] < 0) ? -1.0 : 1.0;
      item[o * 6 + j] = ids[ed];
      for (int c = 0; c < 3; c++) {
        val_e[(size_t)c * nobs * 6 + o * 6 + j] = w[ed][c] * sg;
        val_b[(size_t)c * nobs * 6 + o * 6 + j] = wfe[ed][c] * sg;
      }
    }
  }
  return AFX_OK;
}

int afx_ixyh(int nobs, const double *xyzobs, const double *exy, const int *stack, const int *item, const double *val,
             int node, const double *vxyz, const double *fxyz, double sigma, double *out, double *h) {
  if (nobs <= 0 || node <= 0 || !xyzobs || !exy || !stack || !item || !val || !vxyz || !fxyz || !out || !h)
    return fail(AFX_ERR_ARG, "ixyh: bad arguments");
  if (stack[0] != 0) return fail(AFX_ERR_ARG, "ixyh: stack must start at 0");
  for (int o = 0; o < nobs; o++)
    if (stack[o + 1] < stack[o]) return fail(AFX_ERR_ARG, "ixyh: stack must be non-decreasing");
  const int nnz = stack[nobs];
  for (int q = 0; q < nnz; q++)
    if (item[q] < 1 || item[q] > node) return fail(AFX_ERR_ARG, "ixyh: item out of range");
  std::vector<double> z(nobs);
  for (int o = 0; o < nobs; o++) z[o] = xyzobs[3 * (size_t)o + 2];
  double *dz = nullptr, *dval = nullptr, *dF = nullptr, *dh = nullptr;
  int *dptr = nullptr, *ditem = nullptr;
  cx<double> *dexy = nullptr, *dv = nullptr, *dout = nullptr;
  auto cleanup = [&]() {
    dfree(dz); dfree(dval); dfree(dF); dfree(dh); dfree(dptr); dfree(ditem); dfree(dexy); dfree(dv); dfree(dout);
  };
  int rc;
  if ((rc = upload(&dz, z.data(), (size_t)nobs)) || (rc = upload(&dptr, stack, (size_t)nobs + 1)) ||
      (rc = upload(&ditem, item, (size_t)nnz)) || (rc = upload(&dval, val, (size_t)nnz)) ||
      (rc = upload(&dexy, (const cx<double> *)exy, 2 * (size_t)nobs)) ||
      (rc = upload(&dv, (const cx<double> *)vxyz, 3 * (size_t)node)) || (rc = upload(&dF, fxyz, 3 * (size_t)node)) ||
      (rc = dalloc(&dout, 8 * (size_t)nobs)) || (rc = dalloc(&dh, (size_t)nobs))) {
    cleanup();
    return rc;
  }
  hipError_t e = afx::launch_ixyh(nobs, dz, dexy, dptr, ditem, dval, dv, dF, sigma, dout, dh, 0);
  if (e == hipSuccess) e = hipMemcpy(out, dout, 8 * (size_t)nobs * sizeof(cx<double>), hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(h, dh, (size_t)nobs * sizeof(double), hipMemcpyDeviceToHost);
  cleanup();
  if (e != hipSuccess) return fail(AFX_ERR_HIP, hipGetErrorString(e));
  return AFX_OK;
}

int afx_obs_apply_device(afx_system s, int nobs, const int *item, const double *val, int ncol, const void *dX,
                         double *out) {
  if (!s || !item || !val || !dX || !out) return fail(AFX_ERR_ARG, "null argument");
  for (int q = 0; q < nobs * 6; q++)
    if (item[q] < 1 || item[q] > s->n) return fail(AFX_ERR_ARG, "item out of range");
  int *di = nullptr;
  double *dv = nullptr;
  cx<double> *dout = nullptr;
  int rc;
  if ((rc = upload(&di, item, (size_t)nobs * 6)) || (rc = upload(&dv, val, (size_t)nobs * 6)) ||
      (rc = dalloc(&dout, (size_t)nobs * ncol))) {
    dfree(di); dfree(dv);
    return rc;
  }
  hipError_t e = afx::launch_obs_apply(nobs, di, dv, ncol, (const cx<double> *)dX, dout, s->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(out, dout, (size_t)nobs * ncol * 16, hipMemcpyDeviceToHost, s->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  dfree(di); dfree(dv); dfree(dout);
  if (e != hipSuccess) return fail(AFX_ERR_HIP, hipGetErrorString(e));
  return AFX_OK;
}

int afx_adjoint_rhs_device(afx_system s, int nobs, const int *stack, const int *item, const double *val, void *dB) {
  if (!s || !stack || !item || !val || !dB || nobs <= 0) return fail(AFX_ERR_ARG, "bad arguments");
  if (stack[0] != 0) return fail(AFX_ERR_ARG, "stack must start at 0");
  for (int o = 0; o < nobs; o++)
    if (stack[o + 1] < stack[o]) return fail(AFX_ERR_ARG, "stack must be non-decreasing");
  const int nnz = stack[nobs];
  for (int q = 0; q < nnz; q++)
    if (item[q] < 1 || item[q] > s->n) return fail(AFX_ERR_ARG, "item out of range");
  int *dst = nullptr, *dit = nullptr;
  double *dv = nullptr;
  int rc;
  if ((rc = upload(&dst, stack, (size_t)nobs + 1)) || (rc = upload(&dit, item, (size_t)std::max(nnz, 1))) ||
      (rc = upload(&dv, val, (size_t)std::max(nnz, 1)))) {
    dfree(dst); dfree(dit); dfree(dv);
    return rc;
  }
  hipError_t e = afx::launch_adjoint_rhs(s->n, nobs, dst, dit, dv, (cx<double> *)dB, s->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  dfree(dst); dfree(dit); dfree(dv);
  if (e != hipSuccess) return fail(AFX_ERR_HIP, hipGetErrorString(e));
  return AFX_OK;
}

int afx_jacobian_device(afx_system s, int nobs, const void *dU, const void *dx, int npar, const int *par_ptr,
                        const int *par_tet, const double *sigma_tet, double omega, double dlnsig_dm, void *dJ) {
  if (!s || !dU || !dx || !par_ptr || !par_tet || !sigma_tet || !dJ) return fail(AFX_ERR_ARG, "null argument");
  if (!s->have_mesh) return fail(AFX_ERR_ARG, "afx_set_mesh first");
  const int nsel = par_ptr[npar];
  for (int q = 0; q < nsel; q++)
    if (par_tet[q] < 0 || par_tet[q] >= s->ntet) return fail(AFX_ERR_ARG, "par_tet out of range");
  int *dptr = nullptr, *dtet = nullptr, *dedges = nullptr;
  double *dsig = nullptr;
  cx<double> *dw = nullptr;
  int rc;
  if ((rc = upload(&dptr, par_ptr, (size_t)npar + 1)) || (rc = upload(&dtet, par_tet, (size_t)nsel)) ||
      (rc = upload(&dsig, sigma_tet, (size_t)s->ntet)) || (rc = dalloc(&dw, (size_t)nsel * 6)) ||
      (rc = dalloc(&dedges, (size_t)nsel * 6))) {
    dfree(dptr); dfree(dtet); dfree(dsig); dfree(dw);
    return rc;
  }
  hipError_t e = afx::launch_jac_w(nsel, dtet, s->d_xyz, s->d_n4, s->d_n6, dsig, omega, dlnsig_dm,
                                   (const cx<double> *)dx, dw, dedges, s->stream);
  if (e == hipSuccess)
    e = afx::launch_jac_acc(npar, dptr, nobs, (const cx<double> *)dU, dw, dedges, (cx<double> *)dJ, s->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  dfree(dptr); dfree(dtet); dfree(dsig); dfree(dw); dfree(dedges);
  if (e != hipSuccess) return fail(AFX_ERR_HIP, hipGetErrorString(e));
  return AFX_OK;
}

}  // extern "C"
