// Internal declarations shared by the afx HIP sources (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/afx.h"
#include "afx_sync.h"

namespace afx {

// most rows in one step of a sweep (the single-column sweep gives each its own 8-lane group)
#define AFX_STEP_ROWS 32
// entries per position of the single-column sweep layout (8 lanes x 2)
#define AFX_PIPE_W 16
// rows of one node block of AFX_PC_NODEBLK (the coefficients kept per row)
#define AFX_NB_W 8
// tiles of the node-block pass in flight per round (each thread one row of each)
#ifndef AFX_NB_UT
#define AFX_NB_UT 2
#endif
// a node-block coefficient: a complex number as two bfloat16 halves (real in the low 16 bits)
typedef unsigned int nbcoef;
// words of the grid barrier: top counter, generation, AFX_BAR_GROUPS group counters, each on
// its own 128-byte line
#define AFX_BAR_WORDS (32 * (2 + AFX_BAR_GROUPS))

// per-block profile counters of the persistent solve (AFX_PROFILE)
#define AFX_DBG_SLOTS 16

// most columns the LDS-ring sweep takes in one launch (2 lanes per row and column)
#ifndef AFX_RING_KC
#define AFX_RING_KC 128
#endif

// complex number as a (re, im) pair; for matrix values the same pair holds (K, M) with the
// effective entry K + i*w*M for a column whose angular-frequency factor is w.
template <typename T>
struct alignas(2 * sizeof(T)) cx {
  T r, i;
};

// One direction of the pipelined sweep (forward over the strict lower triangle, backward over
// the strict upper), laid out in block-step order.
struct SweepSide {
  const int *bstep;  // [G+1] first step of each block
  const int *spos;   // [nstep+1] first position of each step
  const int *row;    // [n] row at each position
  const int *ptr;    // [n+1] entry range at each position
  const int *col;    // [nnz_side] column (0-based) of each entry
  const void *val;   // [nnz_side] cx<TP> (K, M) values
  const unsigned char *exp;  // [n] by row id: 1: another block reads this row's z
  const unsigned char *expp; // [n] the same flag by position
  const int *rowx;           // [n] row | (exported << 31) by position
  int maxsteps;              // most steps of one block
  // fixed-width copy for the single-column LDS path: entry e of position pos at pos*ew + e
  int ew;                    // entries per position (padding: column -1, value 0)
  const int *ecol;
  const void *eval;          // cx<TP>
  const void *dpos;          // [n] cx<TP>: DD in position order
  const int *epos;           // [n * ew] position of the entry's column row when it is a row of the
                             // same block, -1 otherwise (ring-buffer sweep of several columns)
  int smax;                  // largest number of positions in one step
  // single-column sweep layout (null when a row has more than AFX_PIPE_W entries): per position
  // AFX_PIPE_W entry sources (>= 0: LDS slot of a same-block row; < 0: mailbox slot ~src;
  // padding: the zero LDS slot) and values, and prow = {row | exported << 31, LDS slot}; position
  // npos is a dummy row that writes only the zero slot
  int npos;
  const int *psrc;
  const int *prow;
  const void *pval;          // cx<TP>
};

template <typename TP>
struct CocgParams {
  int n, k, kp, nf;  // rows, columns, padded columns (lanes per row slot), distinct frequencies
  int pc;            // AFX_PC_*
  int dist_reduce;   // 1: block c reduces column c then grid sync; 0: every block reduces
  // full CSR, FP64 (row order: diagonal, strict upper, strict lower)
  const int *rowptr, *col;
  const cx<double> *val;
  // pipelined sweeps
  int Gs;            // blocks of the launch == sweep blocks
  int lds_rows;      // > 0: own-block z staged in LDS (single column)
  int ring_w;        // > 0: several columns, z of the last ring_w positions kept in an LDS ring
  const int *rstart; // [G+1] row range of each block
  SweepSide fw, bw;
  cx<TP> *mbf, *mbb;   // [n][k] mailboxes of the forward / backward sweep (MB_EMPTY when idle)
  cx<TP> *mbf2, *mbb2; // a second set: iteration it uses set (it & 1), the other one is emptied meanwhile
  const cx<TP> *DD;    // [n][nf]
  const int *col2f;    // [k]
  const double *wcol;  // [k] omega of each column
  // vectors [n][k]
  const cx<double> *B;
  cx<double> *X, *R, *P, *Q;
  cx<TP> *Z;
  // reductions [G][k] and finals [k]
  cx<double> *part_rho, *part_c1;
  double *part_rr;
  cx<double> *fin_rho, *fin_c1;
  double *fin_rr;
  // per-column state
  cx<double> *rho1;
  double *bnrm2;
  int *conv;   // 0 running, 1 converged, 2 zero rhs, -100 maxit, -200 breakdown
  int *iters;
  double *resid;
  int *status;  // [0] iterations done so far, [1] all columns finished
  GridBar bar;
  int *err;     // bounded waits that gave up
  double tol;
  int maxit, chunk;
  unsigned long long *dbg;  // optional [G][AFX_DBG_SLOTS] phase timing (wall clock ticks) and poll counts
  // stacked independent systems (nsys > 1, one column): block b's rows lie in system bsys[b],
  // whose blocks are sblk[s] .. sblk[s+1]-1; rho1, bnrm2, conv, iters and resid are per system
  // and every vector pass runs over the block's own rows
  int nsys;
  const int *bsys, *sblk;
  int *run;  // [G] the block's system still running after the last iteration
  // nodal gradient-space correction (AFX_PC_GRAD, single column): nn interior nodes, their
  // incident edges nadj[nptr[a] .. nptr[a+1]) as +(e+1) when the edge ends at a and -(e+1) when it
  // starts there, each edge's end nodes en[2e], en[2e+1] (interior index, or nn for a boundary
  // node), mnd = diag(G^T M G) (> 0), Zn [nn+1] the nodal correction (slot nn stays 0)
  int grad;
  int nn;
  const int *nptr, *nadj, *en;
  // the same lists at a fixed width of 16 slots per node (0 = unused slot); a node with more
  // incidences holds its first 16 here and the rest only in nadj
  const int *nadj16;
  const int *noff;  // stacked systems: system q's interior nodes are [noff[q], noff[q+1])
  const double *mnd;
  cx<double> *Zn;
  double *part4;  // [G][5] per-block partials of the fused single-column kernel (c1: 2, rho and |r|^2: 3)
  // node-block Jacobi (pc == AFX_PC_NODEBLK, fused single-column kernel): tiles of whole blocks
  // [ntile + 1], each row's place in its block, the coefficients (AFX_NB_W bfloat16 complex per
  // row), and Zb = Bi r of the residual
  int ntile;
  const int *tptr;
  const unsigned char *bm;
  const nbcoef *bi;
  cx<double> *Zb;
};

struct SpmvParams {
  int n, k, kp;  // kp: k rounded up to a power of two (lanes per row slot); 1 when k == 1
  const int *rowptr, *col;
  const cx<double> *val;
  const double *wcol;
  const cx<double> *X;
  const cx<double> *B;  // optional: computes B - A X when non-null
  cx<double> *Y;
  double *part_rr, *part_bb;  // optional [G][k] partial |Y|^2 and |B|^2
};

// Bucket grid of the tets for point location: n[d] buckets of width h[d] from lo[d]; a tet is
// listed in the bucket of its centroid; R[d] is the largest |vertex - centroid| along axis d.
struct PointGrid {
  double lo[3], h[3], R[3];
  int n[3];
};

// launchers (afx_solver.hip)
hipError_t launch_cocg64(const CocgParams<double> &p, int lpr, int grid, hipStream_t s);
hipError_t launch_cocg32(const CocgParams<float> &p, int lpr, int grid, hipStream_t s);
int max_coop_grid(int lpr, bool multi, bool fp64, size_t dyn);
// the COCG iteration without sweeps (DIAG / NONE edge preconditioner, optional nodal
// correction, one column, FP64) at two waves per SIMD
int max_coop_grid_diag(int lpr, int per_cu);
hipError_t launch_cocg_diag(const CocgParams<double> &p, int lpr, int grid, hipStream_t s);
// one column, DIAG / NONE (+ nodal correction), FP64: the fused kernel, one block of 256 threads per CU
int max_coop_grid_cocg1(int lpr);
// node-block Jacobi (AFX_PC_NODEBLK): block inverses from the CSR values at omega w, and z = Bi r
hipError_t launch_nodeblk_setup(int nblk, const int *bstart, const unsigned char *bm, const int *rowptr,
                                const int *col, const cx<double> *val, double w, nbcoef *bi, hipStream_t s);
hipError_t launch_nodeblk_apply(int ntile, const int *tptr, const unsigned char *bm, const nbcoef *bi,
                                const cx<double> *r, cx<double> *z, hipStream_t s);
hipError_t launch_cocg1(const CocgParams<double> &p, int lpr, int grid, hipStream_t s);
int max_coop_grid_ic();
hipError_t launch_spmv(const SpmvParams &p, hipStream_t s);
hipError_t launch_spmv_stream(int n, int maxrow, const int *rowptr, const int *col, const cx<double> *val, double w,
                              const cx<double> *X, cx<double> *Y, hipStream_t s);
hipError_t launch_residual_norms(const SpmvParams &p, int grid, double *res_out_dev, hipStream_t s);
hipError_t launch_precond_setup(int n, int nf, int kind, double sd, double sig, const cx<double> *Dkm,
                                const int *Dpos_full, const cx<double> *full_val, const int *rowptr,
                                int nlevL, const int *lvlL, const int *rowL, const int *Lptr,
                                const int *Lcol, const cx<double> *Lval, const int *Uptr_rowmajor,
                                const cx<double> *AUkm, const double *wf, cx<double> *DD, int grid,
                                hipStream_t s);
hipError_t launch_precond_apply_levels(int n, int kind, int nlevL, const int *lvlL, const int *rowL,
                                       const int *Lptr, const int *Lcol, const cx<double> *Lval, int nlevU,
                                       const int *lvlU, const int *rowU, const int *Uptr, const int *Ucol,
                                       const cx<double> *Uval, const cx<double> *DD, const double *w,
                                       const cx<double> *r, cx<double> *z, hipStream_t s);
// nodal gradient-space correction (AFX_PC_GRAD): mnd[a] = sum_{e,f at a} s_ae s_af M_ef over the
// full CSR (K, M) values; the host-side apply z += G (g / (i w mnd)), g = G^T r
hipError_t launch_nodal_diag(int nn, const int *nptr, const int *nadj, const int *en, const int *rowptr,
                             const int *col, const cx<double> *val, double *mnd, hipStream_t s);
hipError_t launch_grad_apply(int n, int nn, const int *nptr, const int *nadj, const int *en, const double *mnd,
                             double w, const cx<double> *r, cx<double> *zn, cx<double> *z, hipStream_t s);
hipError_t launch_gather_vals(int m, const int *map, const cx<double> *canon, cx<double> *out,
                              cx<float> *out32, hipStream_t s);
hipError_t launch_dpos(int n, const int *row, const cx<double> *DD, cx<double> *d64, cx<float> *d32, hipStream_t s);
hipError_t launch_cast_vec(long long m, const cx<double> *in, cx<float> *out, hipStream_t s);
hipError_t launch_mb_fill(long long words, unsigned long long *mb, hipStream_t s);
hipError_t launch_warm_dots(long long n, int k, int kp, int grid, const cx<double> *Y, const cx<double> *B,
                            double *part, hipStream_t s);
hipError_t launch_scale_cols_cx(long long n, int k, const cx<double> *g, cx<double> *X, hipStream_t s);
hipError_t launch_col_get(long long n, int k, int c, const cx<double> *A, cx<double> *v, hipStream_t s);
hipError_t launch_col_put(long long n, int k, int c, const cx<double> *v, cx<double> *A, hipStream_t s);
hipError_t launch_cols_get(long long n, int k, int c0, int g, const cx<double> *A, cx<double> *B, hipStream_t s);
hipError_t launch_cols_put(long long n, int k, int c0, int g, const cx<double> *B, cx<double> *A, hipStream_t s);
// k <= 8 columns of the iteration without sweeps, a row's matrix entries read once for all columns
hipError_t launch_cocg_blk(const CocgParams<double> &p, int grid, hipStream_t s);
hipError_t launch_axpy_f32_to_f64(long long m, const cx<float> *d, cx<double> *x, hipStream_t s);
hipError_t launch_scale_cols(long long n, int k, const cx<double> *in, const double *scale,
                             cx<float> *out, hipStream_t s);
hipError_t launch_axpy_scaled(long long n, int k, const cx<float> *d, const double *scale,
                              cx<double> *x, hipStream_t s);

// assembly and post-processing (afx_assembly.hip)
struct ElemParams {
  int ntet;
  const double *xyz;
  const int *n4, *n6line, *group;
  const double *sigma;
  const cx<double> *vxyz;
  const double *fxyz;
  double *eK, *eM;   // [21][ntet] element entries (6 diagonal, 15 strictly upper pairs)
  cx<double> *eB;    // [6][ntet] tidal right-hand side per local edge (without omega)
  int *err;          // set to 1 when a volume is <= 0
};
hipError_t launch_element(const ElemParams &p, hipStream_t s);
hipError_t launch_slot_gather(int nslot, const int *cptr, const int *cidx, const double *eK, const double *eM,
                              cx<double> *out, hipStream_t s);
hipError_t launch_rhs_gather(int n, const int *cptr, const int *cidx, const cx<double> *eB, double omega_unused,
                             cx<double> *b, hipStream_t s);
hipError_t launch_mirror(long long nL, const int *mirror, const cx<double> *AU, cx<double> *AL, hipStream_t s);
hipError_t launch_bc_values(int n, const int *INU, const int *IAU, const int *INL, const int *IAL,
                            const unsigned char *dir, cx<double> *D, cx<double> *AU, cx<double> *AL,
                            hipStream_t s);
hipError_t launch_genbc(int nline, const int *line, const double *xyz, double zmin, double zmax, double xout,
                        double yout, unsigned char *bc, hipStream_t s);
hipError_t launch_findelement_grid(int nobs, const double *pts, const PointGrid &g, const int *bptr, const int *btet,
                                   const double *xyz, const int *n4, int *first, hipStream_t s);
hipError_t launch_ixyh(int nobs, const double *zobs, const cx<double> *exy, const int *ptr, const int *item,
                       const double *val, const cx<double> *v, const double *F, double sigma, cx<double> *out,
                       double *hout, hipStream_t s);
hipError_t launch_obs_apply(int nobs, const int *item, const double *val, int ncol, const cx<double> *X,
                            cx<double> *out, hipStream_t s);
hipError_t launch_adjoint_rhs(int n, int nobs, const int *stack, const int *item, const double *val, cx<double> *B,
                              hipStream_t s);
hipError_t launch_jac_w(int ntet_sel, const int *tets, const double *xyz, const int *n4, const int *n6line,
                        const double *sigma, double omega, double dlnsig, const cx<double> *x, cx<double> *w,
                        int *edges, hipStream_t s);
hipError_t launch_jac_acc(int npar, const int *par_ptr, int nobs, const cx<double> *U, const cx<double> *w,
                          const int *edges, cx<double> *J, hipStream_t s);

}  // namespace afx
