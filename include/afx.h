/*
 * afx.h -- C ABI of the MI355X (gfx950) tidal edge-FEM hot path.
 *
 * Conventions shared by every entry point:
 *   - plain pointers and sizes only; complex numbers are interleaved (re, im) doubles,
 *     exactly the memory layout of Fortran complex(8) arrays;
 *   - index arrays use the reference's layout: INU/INL are offsets (0:n) starting at 0,
 *     IAU/IAL/n4/n6line/line hold 1-based ids;
 *   - host-pointer entry points copy in and out; *_device entry points take device
 *     pointers that already live in HBM (same layouts);
 *   - return codes mirror COCG13_single (0 ok, -100 max iterations, 101 bad options) and
 *     add AFX_ERR_* for conditions where the reference calls `stop`;
 *   - one afx_system per host thread; calls on distinct systems are independent.
 *     The only global state is the drop-in afx_cocg13_single's saved system and scaling,
 *     the counterpart of the reference routine's SAVEd DD and SCALE.
 */
#ifndef AFX_H
#define AFX_H

#ifdef __cplusplus
extern "C" {
#endif

#define AFX_OK 0
#define AFX_ERR_MAXIT (-100)     /* COCG13_single.f90:320 ERROR = -100 */
#define AFX_ERR_BADOPT 101       /* COCG13_single.f90:49-52 ERROR = 101 */
#define AFX_ERR_BREAKDOWN (-200) /* rho or p^T q became 0 / non-finite (no reference code) */
#define AFX_ERR_HIP (-300)       /* a HIP runtime call failed */
#define AFX_ERR_ARG (-400)       /* invalid argument or size */
#define AFX_ERR_GEOM (-500)      /* tet volume <= 0: gradnodebasisfun stops (m_fem_util.f90:23,31-39) */
#define AFX_ERR_PATTERN (-600)   /* CSR slot missing: sup_iccg stops (m_iccg_var_takuto.f90:85-88) */
#define AFX_ERR_OUTSIDE (-700)   /* point outside the mesh: FINDELEMENT0 stops (m_mesh_type.f90:731-733) */

/* preconditioner kinds, PRECOND of COCG13_single.f90:107,126,202,236 */
#define AFX_PC_NONE 0
#define AFX_PC_IC 1   /* 'IC' / 'ILU': diagonal-modified ILU (D-ILU) */
#define AFX_PC_SSOR 2 /* 'SSOR' */
#define AFX_PC_DIAG 3 /* 'DIAG' */
#define AFX_PC_SSOR_BJ 4 /* SSOR on the diagonal blocks of the device row partition (no reference
                            counterpart: changes the preconditioner, keeps the COCG iteration) */
#define AFX_PC_NODEBLK 5 /* node-block Jacobi (no reference counterpart): the inverse of each
                            diagonal block of A over a run of consecutive rows whose edges start at
                            the same node (the edge numbering of MKLINE puts a node's edges towards
                            higher-numbered nodes next to each other; runs of more than 8 rows are
                            cut), symmetrised and kept as bfloat16 complex (8 per row, 32 B), applied
                            in FP64. Sweep-free like DIAG; needs
                            afx_set_nodes; FP64; one coefficient set per distinct column frequency;
                            single columns, k-column launches (k <= 8) and stacked systems */
/* Flag OR-ed into AFX_PC_SSOR, AFX_PC_SSOR_BJ, AFX_PC_DIAG or AFX_PC_NODEBLK (no reference counterpart): adds
   the nodal gradient-space correction G diag(G^T A G)^-1 G^T to the edge preconditioner, G the
   discrete gradient of the interior nodes (afx_set_nodes). The curl-curl part annihilates
   gradients (K G = 0), so A's gradient components are set by the tiny i*omega*mu0*sigma mass
   alone, and in the air they are what makes edge-only SSOR-COCG need tens of thousands of
   iterations; the correction solves for them on the nodes (diag(G^T A G) = i*omega*diag(G^T M G)).
   The preconditioner stays complex symmetric, so COCG is unchanged. Single columns, k-column
   launches and stacked systems (each system's interior nodes contiguous, afx_set_nodes on the
   stack); the stacked-copies block width is not used with it. */
#define AFX_PC_GRAD 0x100

/* arithmetic of the Krylov loop */
#define AFX_FP64 0  /* all-FP64 COCG */
#define AFX_MIXED 1 /* complex-FP32 preconditioner (DD, sweep values and z in FP32); the SpMV,
                       the recurrence and its scalars stay FP64. With check_true, the FP64 true
                       residual is recomputed after each COCG run and the run restarts from the
                       current iterate while it exceeds tol, at most max_outer times */

typedef struct afx_system_s *afx_system;

typedef struct {
  int precond;       /* AFX_PC_* */
  double sigma_diag; /* SIGMA_DIAG */
  double sigma;      /* SIGMA */
  double tol;        /* RESID in: relative recurrence residual ||r||/||b|| */
  int maxit;         /* ITER in */
  int nset;          /* NSET: 1 preconditioned, 2 scaled + preconditioned, 0 reuse setup */
  int precision;     /* AFX_FP64 or AFX_MIXED */
  double inner_tol;  /* unused (kept for the layout of the struct) */
  int max_outer;     /* with check_true: most COCG runs restarted from the iterate while the true
                        residual exceeds tol */
  int chunk;         /* iterations per persistent launch (0: default) */
  int check_true;    /* 1: recompute the FP64 true residual at the end */
} afx_opts;

typedef struct {
  int iter;          /* COCG iterations (reference ITER out; MAXIT + 1 when exhausted) */
  int error;         /* 0, -100, -200 */
  double resid;      /* recurrence residual ||r_k||/||b|| (reference RESID out) */
  double true_resid; /* FP64 ||b - A x||/||b|| (when check_true) */
  int outer;         /* refinement steps (AFX_MIXED) */
} afx_stats;

void afx_default_opts(afx_opts *o);
const char *afx_last_error(void);

/* ---- system lifecycle -------------------------------------------------------------- */

/* Pattern from the split CSR of type(global_matrix) (src/solver/m_iccg_var_takuto.f90:9-18):
   analyses level sets, builds the SpMV and sweep layouts on the device. */
int afx_create(int n, const int *INU, const int *IAU, const int *INL, const int *IAL, afx_system *out);

/* Pattern from the tet->edge map, replacing set_iccg_var7_dofn + C_set_dof2
   (src/solver/m_iccg_var_takuto.f90:277-485, called at src/solver_mpi/n_ebfem_bxyz_mpi.f90:104-107). */
int afx_create_from_n6line(int nline, int ntet, const int *n6line, afx_system *out);

/* m independent systems stacked into one block-diagonal system (no reference counterpart: the
   reference solves one mesh per forward_bxyz call, src/solver_mpi/forward_bxyz_mpi.f90:3-67).
   System q owns rows [row_off[q], row_off[q+1]) (0-based, row_off[0] = 0, row_off[m] = n, every
   system at least one row); the split CSR arrays are afx_create's, and no entry may couple two
   systems. The sweep blocks are cut at the system boundaries, so afx_solve_stacked_device can
   keep alpha, beta and convergence per system. m is at most the number of sweep blocks (one
   per compute unit on a large system). */
int afx_create_stacked(int m, const int *row_off, int n, const int *INU, const int *IAU, const int *INL,
                       const int *IAL, afx_system *out);

/* Number of stacked systems of a handle (1 unless made by afx_create_stacked). */
int afx_stacked_count(afx_system s, int *nsys);

int afx_destroy(afx_system s);
int afx_get_sizes(afx_system s, int *n, int *npu, int *npl, int *nlev_l, int *nlev_u);
int afx_get_pattern(afx_system s, int *INU, int *IAU, int *INL, int *IAL);

/* ---- matrix values ----------------------------------------------------------------- */

/* Set D(n), AU(npu), AL(npl) (complex, host) -- the arrays COCG13_single receives
   (COCG13_single.f90:25-29). */
int afx_set_values(afx_system s, const double *D, const double *AU, const double *AL);
int afx_get_values(afx_system s, double *D, double *AU, double *AL);

/* Mesh for on-device assembly: xyz(node,3) [km], n4(ntet,4), group(ntet) = n4flag(:,2),
   vxyz(node,3) complex [m/s], fxyz(node,3) [nT], all row-major (C order); n6line is the
   one given at afx_create_from_n6line. Mirrors the data GENMAT reads
   (src/solver_mpi/forward_bxyz_mpi.f90:99-155). */
int afx_set_mesh(afx_system s, int node, int ntet, const double *xyz, const int *n4, const int *group,
                 const double *vxyz, const double *fxyz);

/* GENMAT + sup_iccg + COPY_UL_ICCG12 (forward_bxyz_mpi.f90:99-311; m_iccg_var_takuto.f90:43-167)
   on the device: one thread per tet, then a deterministic slot-ordered reduction.
   sigma_tet(ntet) [S/m] is the per-element conductivity GENMAT selects (:196-208).
   The assembled matrix is K + i*omega*M; afx_assemble stores K and M separately so that
   afx_set_omega can switch frequency without re-assembly. b_out (nline complex, host or NULL)
   receives the tidal right-hand side for this omega. */
int afx_assemble(afx_system s, const double *sigma_tet, double omega, double *b_out);
int afx_set_omega(afx_system s, double omega);

/* GENBCCSEM (forward_bxyz_mpi.f90:467-519): Dirichlet mask of the edges lying within 1 km of
   one outer face; line(nline,2) 1-based. line_bc out (nline bytes). */
int afx_genbc(afx_system s, const int *line, double zmin, double zmax, double xout, double yout,
              unsigned char *line_bc);

/* Edge-to-node incidence for AFX_PC_GRAD: line(nline, 2) the 1-based end nodes of every edge
   (READLINE's line, m_line_type.f90:28-48; the edge points from line(i,1) to line(i,2), the
   orientation n6line's signs refer to), node the node count, dirichlet (nline bytes, 1 = Dirichlet
   edge, as GENBCCSEM marks them). Nodes on a Dirichlet edge are boundary nodes (their gradients
   have tangential components on the boundary) and are left out of G. */
int afx_set_nodes(afx_system s, int node, const int *line, const unsigned char *dirichlet);

/* SET_BC_ICCG (forward_bxyz_mpi.f90:567-636) with zero Dirichlet values: zeroes Dirichlet rows
   and columns in place (pattern kept), D = 1, and zeroes rhs(n, nrhs) (host, may be NULL)
   at Dirichlet rows. */
int afx_set_bc(afx_system s, const unsigned char *dirichlet, int nrhs, double *rhs, int ldr);

/* ---- preconditioner and solves ----------------------------------------------------- */

/* Preconditioner setup of COCG13_single.f90:103-130 on the device. */
int afx_precond_setup(afx_system s, int kind, double sigma_diag, double sigma);

/* Solve A X = B for nrhs right-hand sides with COCG (COCG13_single.f90:181-323 per column,
   all columns sharing each matrix read). B(n,nrhs), X(n,nrhs) column-major host arrays
   (leading dimensions ldb, ldx); X holds the initial guess on entry (warm start) and the
   solution on exit. stats: nrhs entries or NULL. Returns 0, -100 (some column hit maxit),
   -200, or an AFX_ERR_*. */
int afx_solve(afx_system s, int nrhs, const double *B, int ldb, double *X, int ldx, const afx_opts *opts,
              afx_stats *stats);

/* Same as afx_solve with B and X already in device memory, row-interleaved: element (i, c)
   at index i*nrhs + c (the layout the batched kernels use). omega_scale (nrhs doubles, host,
   or NULL for all ones) gives each column its own frequency: column c solves
   (K + i*omega_c*M) x = b with omega_c = omega * omega_scale[c] (needs afx_assemble values). */
int afx_solve_device(afx_system s, int nrhs, const void *dB, void *dX, const double *omega_scale,
                     const afx_opts *opts, afx_stats *stats);

/* One persistent COCG launch over all systems of an afx_create_stacked handle: one device
   column B, X (n complex doubles; X holds the initial guess), each system with its own alpha,
   beta, ||b_s||, stopping test ||r_s|| / ||b_s|| <= tol and FP64 true residual check (the
   COCG13_single semantics per system, COCG13_single.f90:164-169,315); finished systems stay
   fixed while the others iterate. stats: one entry per system (afx_stacked_count) or NULL.
   Returns 0, -100 (some system hit maxit), -200, or an AFX_ERR_*. */
int afx_solve_stacked_device(afx_system s, const void *dB, void *dX, const afx_opts *opts, afx_stats *stats);

/* FP64 true residual ||b - A x|| / ||b|| per column of device, row-interleaved B and X. */
int afx_true_residual_device(afx_system s, int nrhs, const void *dB, const void *dX,
                             const double *omega_scale, double *res);

/* afx_true_residual_device without any allocation (no counterpart in the reference: for callers
   that cannot allocate device memory -- code under HIP graph capture, tight serving loops,
   harnesses that forbid hipMalloc). Same arguments and results; the per-block partial sums and
   the column omegas live in a static device array of the library (about 1 MB, one per device,
   kept until the process ends). At most 64 columns (AFX_ERR_ARG beyond). Every call on a device
   shares that array: calls that overlap on one device race, so serialise them (one stream). */
int afx_true_residual_device_noalloc(afx_system s, int nrhs, const void *dB, const void *dX,
                                     const double *omega_scale, double *res);

/* Minimal-residual scaling of device initial guesses X (row-interleaved, nrhs columns), for the
   learned warm start (no reference counterpart; the reference always starts COCG13_single from
   the X it is given, COCG13_single.f90:139-145). Per column: y = A x, gamma = y^H b / y^H y;
   X <- gamma X when that lowers ||b - A X|| / ||b|| below 1 (the zero guess's value), else
   X <- 0 (fallback). gamma (2*nrhs, re/im), res_given and res_scaled (nrhs, relative residual
   of X as given and as returned) may be NULL. */
int afx_scale_initial_guess_device(afx_system s, int nrhs, const void *dB, void *dX,
                                   const double *omega_scale, double *gamma, double *res_given,
                                   double *res_scaled);

/* Launch timing (HIP events on the system's stream) of the persistent COCG kernel accumulated
   since the last reset: launches, COCG iterations they ran, and total kernel milliseconds. */
int afx_kernel_timing(afx_system s, int reset, long long *launches, long long *iterations, double *kernel_ms);

/* Autotune for nrhs columns on this system. nrhs = 1: the rows per wavefront of the persistent
   kernel. nrhs > 1: the right-hand-side block width of afx_solve / afx_solve_device, timing 20
   iterations of each: one column at a time, the nrhs-column kernel, and nrhs stacked copies of
   the system (copy c at column c's omega) in one launch with per-system scalars (only where the
   copies' rows of a sweep block fit in LDS; the copies are kept by the handle and rebuilt when
   the values or the column omegas change). The environment variable AFX_BATCH=columns|multi|stack
   overrides the choice. */
int afx_autotune(afx_system s, int nrhs, int verbose);

/* y = A x (complex, host arrays) with the device SpMV -- parity hook. */
int afx_spmv(afx_system s, const double *x, double *y);

/* q = A p on the device, one column (COCG13_single.f90:270-290: the D, upper and lower entries of
   each row, partial sums of lane-strided entries added in a fixed tree): dX, dY device arrays of
   n complex values at the system's frequency, not overlapping. Runs on the system's stream. With kernel_ms non-null it records HIP
   events around the launch on that stream, waits for them and stores the launch's duration. */
int afx_spmv_device(afx_system s, const void *dX, void *dY, double *kernel_ms);

/* z = M^-1 r with the device preconditioner (after afx_precond_setup) -- parity hook. */
int afx_precond_apply(afx_system s, const double *r, double *z);

/* Copy out DD (complex n): the preconditioner diagonal of COCG13_single.f90:105-130. */
int afx_get_precond_diag(afx_system s, double *DD);

/* ---- drop-in mirrors of the reference solver calls ---------------------------------- */

/* Same arguments and meaning as COCG13_single (src/solver/COCG13_single.f90:9-36), by
   pointer for the Fortran caller (scalars by value where intent(in)): D, AU, AL modified for
   NSET = 2; B scaled and restored; RESID, ITER in/out; ERROR out. PRECOND is 20 characters.
   Saved state, as the reference's SAVEd DD / SCALE / IFLAG (:39-44,57-62): the device system of
   the last call (pattern analysis, device arrays, preconditioner) is kept and reused while
   INU/IAU/INL/IAL are the same, and so is SCALE (all ones until an NSET = 2 call sets it; NSET = 1
   then scales B by it, as the reference does). NSET = 0 solves with the current D/AU/AL and the
   preconditioner diagonal of the last NSET >= 1 call, B unscaled and divided by SCALE on exit
   (:98-100,326-329); it returns ERROR = 101 when there is no such call, and also when the kept
   system has another pattern or the last setup another PRECOND (where the reference would apply
   a DD that belongs to another matrix). A different NP starts the saved state afresh (the
   reference cannot change NP). On an AFX_ERR_* return B is restored. Calls are serialised. */
int afx_cocg13_single(int NP, int NPL, int NPU, double *D, double *AL, const int *INL, const int *IAL,
                      double *AU, const int *INU, const int *IAU, double *B, double *X, const char *PRECOND,
                      double SIGMA_DIAG, double SIGMA, double *RESID, int *ITER, int *ERROR, int NSET);

/* Frees the system that afx_cocg13_single keeps between calls (no reference counterpart: the
   reference frees nothing either; for callers that want the device memory back). */
int afx_cocg13_single_release(void);

/* solvePARDISO(doftot, nsr, A, rhs, xout, ip) (src/solver/solvePARDISO.f90:6-16): A's split arrays
   (type(global_matrix), m_iccg_var_takuto.f90:9-18), rhs(doftot, nsr) and xout(doftot, nsr)
   column-major complex. The direct factorisation is replaced by FP64 COCG to the relative FP64
   true residual tol (||b - A x|| / ||b|| per column, recomputed; PARDISO's own accuracy is about
   1e-12); with the edges' end nodes line(2, nline) (READLINE's g_line%line, 1-based) and node count,
   and the Dirichlet mask line_bc (GENBCCSEM's line_bc as one byte per edge), the preconditioner is
   AFX_PC_NODEBLK + AFX_PC_GRAD, otherwise SSOR (line = NULL). Returns 0, -100 (tol not reached: the
   iterations ran out, or the recomputed true residual is still above tol after 4 restarts, as for a
   right-hand side whose solution FP64 cannot resolve) or an AFX_ERR_* where the reference stops. The system of the last call is kept while the pattern is
   the same (afx_solve_pardiso_release frees it); calls are serialised. ip is not needed. */
int afx_solve_pardiso(int doftot, int nsr, const int *INU, const int *IAU, const int *INL, const int *IAL,
                      const double *D, const double *AU, const double *AL, int node, const int *line,
                      const unsigned char *line_bc, const double *rhs, double *xout, double tol);

/* solvePARDISOinv(doftot, nsr, A, rhs, xout, PTR, ut, iflag_comp, ip, np)
   (src/src_inv/solvePARDISOinv.f90:5-23): the forward solve as afx_solve_pardiso (nsr may be 0),
   then for every component c < 5 with iflag_comp[c] != 0 the adjoint solutions u_o = A^-1 P_o^T of
   the nobs rows of PTR(c) (:176-191), returned as ut(c), the CRS of u^T [nobs x doftot] keeping
   |u| > 1e-10 (:43,194-196; convcomp_full2crs, m_matrix.f90:405-451). Every flagged component has
   nobs rows (as the reference assumes). PTR(c): stack ptr_stack[c * (nobs + 1) + i] (0-based
   offsets), item (1-based edge ids) and val at ptr_item / ptr_val + ptr_base[c]. ut(c): stack
   ut_stack[c * (nobs + 1) + i], item (1-based) and complex val at ut_item + c * ut_cap and
   ut_val + 2 * c * ut_cap, entry count ut_ntot[c]. When some component needs more than ut_cap
   entries it returns AFX_ERR_ARG with ut_ntot holding the counts needed (call again with a larger
   ut_cap). Same kept system and serialisation as afx_solve_pardiso. */
int afx_solve_pardiso_inv(int doftot, int nsr, const int *INU, const int *IAU, const int *INL, const int *IAL,
                          const double *D, const double *AU, const double *AL, int node, const int *line,
                          const unsigned char *line_bc, const double *rhs, double *xout, int nobs,
                          const int *iflag_comp, const int *ptr_stack, const int *ptr_base, const int *ptr_item,
                          const double *ptr_val, int ut_cap, int *ut_stack, int *ut_ntot, int *ut_item, double *ut_val,
                          double tol);

/* solvePARDISOjointinv(ACT, MT, nline, nsr, A, rhs, rhs_mt, xout, xout_mt, PTR, PTR_mt, ut, ut_mt,
   iflag_comp, ip, np) (src/src_inv_joint/solvePARDISOjointinv.f90:5-26), the joint tidal + MT
   caller's twin of afx_solve_pardiso_inv on one kept system (forward_joint.f90:93):
   - act != 0: the nsr active-source columns rhs -> xout and, for every flagged component of PTR(5)
     (nobs rows each), ut(c), all laid out as in afx_solve_pardiso_inv;
   - mt != 0: the two MT columns rhs_mt(doftot, 2) -> xout_mt(doftot, 2) and the adjoint solutions of
     all 4 MT components of PTR_mt (bx, by, ex, ey; nobs_mt rows each), returned as ut_mt(c) with
     ptr_mt_* / ut_mt_* in the layout of ptr_* / ut_* (stride nobs_mt + 1, capacity ut_mt_cap).
   The forward columns of both run in one launch. As the reference (:56-64), nothing is solved when
   nobs = 0 (no flagged active component with rows), whatever mt says. The PTR_mt_*.dat files the
   reference writes when ip = 0 (:248-261) are not written. Returns as afx_solve_pardiso_inv. */
int afx_solve_pardiso_jointinv(int act, int mt, int doftot, int nsr, const int *INU, const int *IAU, const int *INL,
                               const int *IAL, const double *D, const double *AU, const double *AL, int node,
                               const int *line, const unsigned char *line_bc, const double *rhs, const double *rhs_mt,
                               double *xout, double *xout_mt, int nobs, const int *iflag_comp, const int *ptr_stack,
                               const int *ptr_base, const int *ptr_item, const double *ptr_val, int nobs_mt,
                               const int *ptr_mt_stack, const int *ptr_mt_base, const int *ptr_mt_item,
                               const double *ptr_mt_val, int ut_cap, int *ut_stack, int *ut_ntot, int *ut_item,
                               double *ut_val, int ut_mt_cap, int *ut_mt_stack, int *ut_mt_ntot, int *ut_mt_item,
                               double *ut_mt_val, double tol);

/* Frees the system afx_solve_pardiso / afx_solve_pardiso_inv keep between calls (no reference
   counterpart: PARDISO frees its factors at the end of each call, solvePARDISO.f90:165-166). */
int afx_solve_pardiso_release(void);

/* solveCG (src/solver_mpi/forward_bxyz_mpi.f90:70-94): X = 0, 'SSOR', RESID 1e-4, ITER 3000. */
int afx_solve_cg(int doftot, int NPL, int NPU, const double *D, const double *AL, const int *INL,
                 const int *IAL, const double *AU, const int *INU, const int *IAU, const double *rhs, double *X);

/* ---- observation operator ---------------------------------------------------------- */

/* PREPOBSCOEFF (src/solver_mpi/n_ebfem_bxyz_mpi.f90:653-726): for each site, FINDELEMENT0's
   first containing tet (device search), and the 6-entry rows of edge-basis (E) and face-basis
   (B = curl A) coefficients sorted by edge id. item(nobs,6) 1-based, val_e/val_b (3,nobs,6),
   iele(nobs) 0-based. Needs afx_set_mesh. */
int afx_prepobscoeff(afx_system s, int nobs, const double *xyzobs, int *item, double *val_e, double *val_b,
                     int *iele);

/* f(nobs, ncol) = sum_j val(o,j) * x(item(o,j), c): mul_matcrs_cv (src/common/m_matrix.f90:83-102)
   over device, row-interleaved X (n, ncol); out is host (nobs, ncol) complex row-major. */
int afx_obs_apply_device(afx_system s, int nobs, const int *item, const double *val, int ncol, const void *dX,
                         double *out);

/* IXYHOUT (src/solver_mpi/n_ebfem_bxyz_mpi.f90:192-301): seafloor current at nobs points.
   xyzobs (nobs, 3) [km]; exy (nobs, 2) complex E at the points (the reference rebuilds it from
   CALOBSRESP's amplitude and phase, :219-225); coeff_vF as CSR (stack[nobs + 1] from 0, item
   1-based node ids, val); vxyz (node, 3) complex [m/s] and fxyz (node, 3) [nT] (readvxyz /
   readfxyz, m_readfvxyz.f90); sigma the ocean conductivity (the reference's literal 3.3, a
   single-precision value, :221). Out (host): out (nobs, 8) complex = ixyh_e, ixyh_vF, ixyh and
   vh, each (x, y); h (nobs) [km]. Does not need a system handle. */
int afx_ixyh(int nobs, const double *xyzobs, const double *exy, const int *stack, const int *item, const double *val,
             int node, const double *vxyz, const double *fxyz, double sigma, double *out, double *h);

/* ---- adjoint Jacobian -------------------------------------------------------------- */

/* Adjoint right-hand sides of solvePARDISOinv (src/src_inv/solvePARDISOinv.f90:176-187): the
   observation matrix PTR (real CRS, nobs rows over the n edges: stack(nobs+1) 0-based offsets,
   item 1-based edge ids, val) transposed into nobs dense device columns, row-interleaved
   (element (i, o) at i*nobs + o, the afx_solve_device layout). Solving with afx_solve_device
   gives the adjoint fields u_o = A^-1 P_o^T that solvePARDISOinv returns as ut. */
int afx_adjoint_rhs_device(afx_system s, int nobs, const int *stack, const int *item, const double *val, void *dB);

/* genjacobian1 (src/src_inv_ap/m_jacobian_ap.f90:174-386) without thresholding:
     J[m, o] = sum_{t in m} sum_j U[e_j(t), o] * w_t[j],
     w_t = -(dBB_t/dm) * Y_t x_t,  dBB_t/dm = i*omega*mu0*sigma_t*L0^2 * dlnsig_dm
   U: device, row-interleaved (nline, nobs) adjoint solutions A^-1 P^T (the afx_solve_device
   layout); x: device forward solution (nline complex); parameters as CSR par_ptr(npar+1),
   par_tet (0-based tets). J: device (npar, nobs) complex, o fastest. Needs afx_set_mesh and
   sigma_tet (ntet) as used in assembly. */
int afx_jacobian_device(afx_system s, int nobs, const void *dU, const void *dx, int npar, const int *par_ptr,
                        const int *par_tet, const double *sigma_tet, double omega, double dlnsig_dm, void *dJ);

/* ---- Gauss-Newton model update (inversion step after the Jacobian) --------------------- */

/* genjacobian1's conversion of the adjoint accumulation into real Jacobian rows
   (src/src_inv_ap/m_jacobian_ap.f90:338-345): for dJ (npar, nobs) complex device (o fastest,
   the afx_jacobian_device output) and the observed fields dbe (nobs complex, device),
   dout[0][o][m] = Re(dJ[m][o] / be[o]) / ln 10 and dout[1][o][m] = Im(dJ[m][o] / be[o]) * 180/pi
   (device, [2][nobs][npar] doubles): the d log10|be| / dm and d phase[deg] / dm rows that
   genjacobian2 stacks into JJ (:393-482). */
int afx_jacobian_ampphase_device(int npar, int nobs, const void *dJ, const void *dbe, double *dout);

/* getnewmodel (src/src_inv_ap/m_jacobian_ap.f90:623-787) with solvebeta: dJ the dense real
   Jacobian (ndat, nmodel) row-major on the device; Cm (nmodel x nmodel, symmetric) as CSR
   (cm_ptr 0-based offsets, cm_col 0-based), cd (ndat) the diagonal of Cd, all host;
     C = J Cm J^T + alpha Cd (FP64 MFMA), X = dobs - dcal + J (model - model_ref),
     C beta = X (device LU with partial pivoting: solvebeta's PARDISO solve is the general real
     one, mtype 11, src/src_inv/solvebeta.f90:78), model_new = model_ref + Cm J^T beta.
   beta_out (ndat, host) may be NULL. Returns AFX_ERR_BREAKDOWN when C is singular (a zero pivot). */
int afx_gn_update_device(int ndat, int nmodel, const double *dJ, const int *cm_ptr, const int *cm_col,
                         const double *cm_val, const double *cd, double alpha, const double *dobs, const double *dcal,
                         const double *model, const double *model_ref, double *model_new, double *beta_out);

/* alphaspectralradius_v2 (src/src_inv_ap/m_spectral.f90:63-99) with LANCZOSMINMAX (:102-179):
   alpha = rho(J Cm J^T) / max_i cd_i, rho the largest |eigenvalue| of the tridiagonal that kmax
   Lanczos steps on J Cm J^T build from the normalized ones vector (fewer when a step's beta is 0).
   dJ (ndat, nmodel) row-major on the device; Cm CSR (0-based) and cd host. emin/emax (may be NULL)
   receive the extreme Ritz values. */
int afx_spectral_alpha_device(int ndat, int nmodel, const double *dJ, const int *cm_ptr, const int *cm_col,
                              const double *cm_val, const double *cd, int kmax, double *alpha, double *emin,
                              double *emax);

/* GENDQ (src/src_inv_ap/m_modelroughness.f90:575-862): the model roughening matrix R of the MSG
   and SM roughness (R^T R for SM). Land elements are tets nphys1+1..ntet (1-based n4, ntet x 4),
   ele2model their 1-based models (ntet - nphys1 entries). Two models are neighbours when land
   elements of each share a tet face; R(m, m') = -1 for neighbours, R(m, m) = the number of
   neighbour models (GENDQ's Q^T D Q with every value replaced, :770-790). iactive (nmodel, 0/1,
   NULL = all): rows and columns of inactive models are dropped and the rest renumbered in
   order, the values kept (:795-830; the reference keeps the column ids, the same when the
   inactive models are the last ones). CSR, 0-based, columns ascending. Call with ptr == NULL to
   get *nrow and *nnz, then with ptr (nrow + 1), col and val (nnz). AFX_ERR_ARG when an active
   model has no neighbour model (the reference stops, :786-789). Host code: no GPU needed. */
int afx_model_roughening(int ntet, const int *n4, int nphys1, const int *ele2model, int nmodel, const int *iactive,
                         int *nrow, long long *nnz, int *ptr, int *col, double *val);

/* C (M, N) = A (M, K) B (N, K)^T (+ C if accumulate), FP64 on MFMA, device row-major arrays --
   the dense contraction of the update, exposed for its parity test. */
int afx_dgemm_nt_device(int M, int N, int K, const double *dA, int lda, const double *dB, int ldb, double *dC, int ldc,
                        int accumulate);

/* message of the last failure of the three calls above on the calling thread */
const char *afx_inv_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
