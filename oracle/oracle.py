"""ctypes front-end of the CPU oracle (liborc.so) -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this module,
as the checker; the product path never does. Every function mirrors one reference
routine; see afx_oracle.c for the file:line each one follows.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def build(force: bool = False) -> str:
    path = os.path.join(_HERE, "liborc.so")
    if force or not os.path.exists(path):
        subprocess.run(["make", "-s", "-C", _HERE, "liborc.so"], check=True)
    return path


def lib():
    """The oracle library; ORC_LIB names another build of it (the sanitizer build of
    `make -C oracle sanitize`)."""
    global _LIB
    if _LIB is None:
        _LIB = C.CDLL(os.environ.get("ORC_LIB") or build())
    return _LIB


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _c128(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


def pattern(nline: int, n6line: np.ndarray):
    """C_set_dof2 (m_iccg_var_takuto.f90:361-485): returns INU, IAU, INL, IAL (IAU/IAL 1-based)."""
    n6 = _i32(n6line)
    INU = np.zeros(nline + 1, np.int32)
    INL = np.zeros(nline + 1, np.int32)
    rc = lib().orc_pattern(C.c_int(nline), C.c_int(n6.shape[0]), _p(n6), _p(INU), None, _p(INL), None)
    if rc:
        raise RuntimeError(f"orc_pattern rc={rc}")
    IAU = np.zeros(INU[-1], np.int32)
    IAL = np.zeros(INL[-1], np.int32)
    rc = lib().orc_pattern(C.c_int(nline), C.c_int(n6.shape[0]), _p(n6), _p(INU), _p(IAU), _p(INL), _p(IAL))
    if rc:
        raise RuntimeError(f"orc_pattern rc={rc}")
    return INU, IAU, INL, IAL


def genmat(case, omega: float, INU, IAU, sigma=None):
    """GENMAT + sup_iccg (forward_bxyz_mpi.f90:99-311): returns D, AU, b (complex128)."""
    sig = _f64(case.sigma_tet() if sigma is None else sigma)
    n = case.nline
    D = np.zeros(n, np.complex128)
    AU = np.zeros(int(INU[-1]), np.complex128)
    b = np.zeros(n, np.complex128)
    xyz, n4, grp, n6 = _f64(case.xyz), _i32(case.n4), _i32(case.group), _i32(case.n6line)
    v, f = _c128(case.vxyz), _f64(case.fxyz)
    INU, IAU = _i32(INU), _i32(IAU)
    rc = lib().orc_genmat(C.c_int(case.node), C.c_int(case.ntet), _p(xyz), _p(n4), _p(grp), _p(n6),
                          C.c_int(n), _p(sig), C.c_double(omega), _p(v), _p(f), _p(INU), _p(IAU),
                          _p(D), _p(AU), _p(b))
    if rc:
        raise RuntimeError(f"orc_genmat rc={rc}")
    return D, AU, b


def genmat_split(case, INU, IAU, sigma=None):
    """Curl-curl K and conductivity mass M (A = K + i*omega*M) with the sup_iccg slot rule."""
    sig = _f64(case.sigma_tet() if sigma is None else sigma)
    n = case.nline
    Kd = np.zeros(n); Ku = np.zeros(int(INU[-1])); Md = np.zeros(n); Mu = np.zeros(int(INU[-1]))
    xyz, n4, n6 = _f64(case.xyz), _i32(case.n4), _i32(case.n6line)
    rc = lib().orc_genmat_split(C.c_int(case.ntet), _p(xyz), _p(n4), _p(n6), _p(sig), _p(_i32(INU)),
                                _p(_i32(IAU)), _p(Kd), _p(Ku), _p(Md), _p(Mu))
    if rc:
        raise RuntimeError(f"orc_genmat_split rc={rc}")
    return Kd, Ku, Md, Mu


def copy_ul(INU, IAU, INL, IAL, AU):
    n = len(INU) - 1
    AL = np.zeros(int(INL[-1]), np.complex128)
    rc = lib().orc_copy_ul(C.c_int(n), _p(_i32(INU)), _p(_i32(IAU)), _p(_i32(INL)), _p(_i32(IAL)),
                           _p(_c128(AU)), _p(AL))
    if rc:
        raise RuntimeError("orc_copy_ul: missing mirror slot")
    return AL


def genbc(case):
    """GENBCCSEM (forward_bxyz_mpi.f90:467-519): (line_bc bool, line_group int)."""
    xout, yout, zmin, zmax = case.bounds
    bc = np.zeros(case.nline, np.uint8)
    grp = np.zeros(case.nline, np.int32)
    lib().orc_genbc(C.c_double(zmin), C.c_double(zmax), C.c_double(xout), C.c_double(yout),
                    C.c_int(case.nline), _p(_i32(case.line)), _p(_f64(case.xyz)), _p(bc), _p(grp))
    return bc.astype(bool), grp


def set_bc(INU, IAU, INL, IAL, D, AU, AL, rf, dirichlet, dval=None):
    """SET_BC_ICCG (forward_bxyz_mpi.f90:567-636), in place on copies; returns D, AU, AL, rf."""
    n = len(D)
    D, AU, AL = _c128(D).copy(), _c128(AU).copy(), _c128(AL).copy()
    rf = np.asfortranarray(np.asarray(rf, np.complex128).reshape(n, -1)).copy(order="F")
    nsr = rf.shape[1]
    dv = np.zeros(n, np.complex128) if dval is None else _c128(dval)
    dm = np.ascontiguousarray(dirichlet, dtype=np.uint8)
    lib().orc_set_bc(C.c_int(n), C.c_int(nsr), _p(_i32(INU)), _p(_i32(IAU)), _p(_i32(INL)), _p(_i32(IAL)),
                     _p(D), _p(AU), _p(AL), rf.ctypes.data_as(C.c_void_p), _p(dv), _p(dm))
    return D, AU, AL, rf


def spmv(INU, IAU, INL, IAL, D, AU, AL, x):
    n = len(D)
    y = np.zeros(n, np.complex128)
    lib().orc_spmv(C.c_int(n), _p(_c128(D)), _p(_i32(INU)), _p(_i32(IAU)), _p(_c128(AU)), _p(_i32(INL)),
                   _p(_i32(IAL)), _p(_c128(AL)), _p(_c128(x)), _p(y))
    return y


PRECOND_KIND = {"IC": 1, "ILU": 1, "SSOR": 2, "DIAG": 3, "NONE": 0}


def precond_setup(kind, D, INU, AU, INL, IAL, AL, sigma_diag=1.0, sigma=0.0):
    n = len(D)
    DD = np.zeros(n, np.complex128)
    lib().orc_precond_setup(C.c_int(n), C.c_int(PRECOND_KIND[kind]), C.c_double(sigma_diag),
                            C.c_double(sigma), _p(_c128(D)), _p(_i32(INU)), _p(_c128(AU)), _p(_i32(INL)),
                            _p(_i32(IAL)), _p(_c128(AL)), _p(DD))
    return DD


def precond_apply(kind, INU, IAU, AU, INL, IAL, AL, DD, r):
    n = len(DD)
    z = np.zeros(n, np.complex128)
    lib().orc_precond_apply(C.c_int(n), C.c_int(PRECOND_KIND[kind]), _p(_i32(INU)), _p(_i32(IAU)),
                            _p(_c128(AU)), _p(_i32(INL)), _p(_i32(IAL)), _p(_c128(AL)), _p(_c128(DD)),
                            _p(_c128(r)), _p(z))
    return z


class Cocg13State:
    """The SAVEd DD and SCALE and the IFLAG of COCG13_single (COCG13_single.f90:39-44), kept
    between calls of cocg13(..., state=...) as the reference keeps them between its calls."""

    def __init__(self, n):
        self.DD = np.zeros(n, np.complex128)
        self.scale = np.ones(n, np.float64)
        self.have = C.c_int(0)


def cocg13(INU, IAU, INL, IAL, D, AU, AL, B, X=None, precond="IC", resid=1e-8, maxit=3000, nset=1,
           sigma_diag=1.0, sigma=0.0, history=False, state=None):
    """COCG13_single (COCG13_single.f90:9-341) on copies of the inputs.

    Returns dict(x, resid, iter, error, hist[, D, AU, AL, B]) -- the modified D/AU/AL/B are
    returned too because the reference modifies them in place for NSET = 2. Without `state` each
    call is the routine's first call in a fresh program; with a Cocg13State the saved state
    carries over, as between successive calls of the reference."""
    n = len(D)
    D, AU, AL, B = _c128(D).copy(), _c128(AU).copy(), _c128(AL).copy(), _c128(B).copy()
    X = np.zeros(n, np.complex128) if X is None else _c128(X).copy()
    st = Cocg13State(n) if state is None else state
    DD, scale, have = st.DD, st.scale, st.have
    r = C.c_double(resid)
    it = C.c_int(maxit)
    err = C.c_int(0)
    hist = np.zeros(max(maxit, 1), np.float64)
    pc = C.create_string_buffer(precond.ljust(20).encode()[:20], 21)
    lib().orc_cocg13(C.c_int(n), C.c_int(int(INL[-1])), C.c_int(int(INU[-1])), _p(D), _p(AL), _p(_i32(INL)),
                     _p(_i32(IAL)), _p(AU), _p(_i32(INU)), _p(_i32(IAU)), _p(B), _p(X), pc,
                     C.c_double(sigma_diag), C.c_double(sigma), C.byref(r), C.byref(it), C.byref(err),
                     C.c_int(nset), _p(DD), _p(scale), C.byref(have), _p(hist))
    out = dict(x=X, resid=r.value, iter=it.value, error=err.value, D=D, AU=AU, AL=AL, B=B)
    if history:
        out["hist"] = hist[: max(0, min(it.value, maxit))].copy()
    return out


def prepobscoeff(case, xyzobs):
    """PREPOBSCOEFF (n_ebfem_bxyz_mpi.f90:653-726): (item (nobs,6), val_e (3,nobs,6), val_b (3,nobs,6), iele)."""
    xo = _f64(xyzobs).reshape(-1, 3)
    nobs = xo.shape[0]
    item = np.zeros((nobs, 6), np.int32)
    ve = np.zeros((3, nobs, 6))
    vb = np.zeros((3, nobs, 6))
    iele = np.zeros(nobs, np.int32)
    rc = lib().orc_prepobscoeff(C.c_int(nobs), _p(xo), C.c_int(case.ntet), _p(_f64(case.xyz)),
                                _p(_i32(case.n4)), _p(_i32(case.n6line)), _p(item), _p(ve), _p(vb), _p(iele))
    if rc:
        raise RuntimeError(f"observation {-rc - 1} outside the mesh")
    return item, ve, vb, iele


def obs_apply(item, val, x):
    nobs = item.shape[0]
    out = np.zeros(nobs, np.complex128)
    lib().orc_obs_apply(C.c_int(nobs), _p(_i32(item)), _p(_f64(val)), _p(_c128(x)), _p(out))
    return out


def ixyhout(xyzobs, exy, stack, item, val, vxyz, fxyz, sigma):
    """IXYHOUT (src/solver_mpi/n_ebfem_bxyz_mpi.f90:213-268), restated in numpy in the reference's
    order of operations: h = -z, 0 on land (:227-231); vF = cr_outer(v, F) at the nodes
    (m_outerinnerproduct.f90:45-53, :240-242); vF_obs, v_obs = mul_matcrs_cv rows in item order
    (m_matrix.f90:83-102, :244-251); ixyh_e = (sigma h) E; ixyh_vF = ((sigma h) vF) 1e-3;
    ixyh = (sigma h) (E + vF 1e-3); vh = h v (:261-268). Real and imaginary parts are computed
    separately (a real times a complex). Returns (ixyh_e, ixyh_vF, ixyh, vh) each (nobs, 2)
    complex, and h (nobs,)."""
    xo = np.asarray(xyzobs, np.float64).reshape(-1, 3)
    nobs = xo.shape[0]
    exy = np.asarray(exy, np.complex128).reshape(nobs, 2)
    v = np.asarray(vxyz, np.complex128).reshape(-1, 3)
    F = np.asarray(fxyz, np.float64).reshape(-1, 3)
    out = [np.zeros((nobs, 2), np.complex128) for _ in range(4)]
    hs = np.zeros(nobs)
    for o in range(nobs):
        h = -xo[o, 2]
        if h <= 0.0:
            h = 0.0
        vf = [[0.0, 0.0], [0.0, 0.0]]
        vh = [[0.0, 0.0], [0.0, 0.0]]
        for q in range(stack[o], stack[o + 1]):
            nd = item[q] - 1
            a = float(val[q])
            vx, vy, vz = v[nd]
            Fx, Fy, Fz = (float(t) for t in F[nd])
            c0 = (vy.real * Fz - vz.real * Fy, vy.imag * Fz - vz.imag * Fy)
            c1 = (vz.real * Fx - vx.real * Fz, vz.imag * Fx - vx.imag * Fz)
            for c, cc in ((0, c0), (1, c1)):
                vf[c][0] = vf[c][0] + a * cc[0]
                vf[c][1] = vf[c][1] + a * cc[1]
            for c, w in ((0, vx), (1, vy)):
                vh[c][0] = vh[c][0] + a * w.real
                vh[c][1] = vh[c][1] + a * w.imag
        sh = sigma * h
        for c in range(2):
            e = exy[o, c]
            out[0][o, c] = complex(sh * e.real, sh * e.imag)
            out[1][o, c] = complex(sh * vf[c][0] * 1.0e-3, sh * vf[c][1] * 1.0e-3)
            out[2][o, c] = complex(sh * (e.real + vf[c][0] * 1.0e-3), sh * (e.imag + vf[c][1] * 1.0e-3))
            out[3][o, c] = complex(h * vh[c][0], h * vh[c][1])
        hs[o] = h
    return out[0], out[1], out[2], out[3], hs


def jacobian(u, x, par_ptr, par_tet, case, omega, dsig_factor, sigma=None):
    """genjacobian1 restated (m_jacobian_ap.f90:174-386): dbe (nobs, npar) complex."""
    u = _c128(u)
    nobs, nline = u.shape
    npar = len(par_ptr) - 1
    sig = _f64(case.sigma_tet() if sigma is None else sigma)
    out = np.zeros((nobs, npar), np.complex128)
    rc = lib().orc_jacobian(C.c_int(nobs), C.c_int(nline), _p(u), _p(_c128(x)), C.c_int(npar),
                            _p(_i32(par_ptr)), _p(_i32(par_tet)), _p(_f64(case.xyz)), _p(_i32(case.n4)),
                            _p(_i32(case.n6line)), _p(sig), C.c_double(omega), C.c_double(dsig_factor),
                            _p(out))
    if rc:
        raise RuntimeError("orc_jacobian failed")
    return out


def forward_system(case, omega, sigma=None):
    """forward_bxyz steps INITIALIZE..SET_BC_ICCG (forward_bxyz_mpi.f90:43-58).

    Returns dict(INU, IAU, INL, IAL, D, AU, AL, b, bc)."""
    INU, IAU, INL, IAL = pattern(case.nline, case.n6line)
    D, AU, b = genmat(case, omega, INU, IAU, sigma)
    AL = copy_ul(INU, IAU, INL, IAL, AU)
    bc, _ = genbc(case)
    D, AU, AL, rf = set_bc(INU, IAU, INL, IAL, D, AU, AL, b, bc)
    return dict(INU=INU, IAU=IAU, INL=INL, IAL=IAL, D=D, AU=AU, AL=AL, b=rf[:, 0].copy(), bc=bc)


def to_scipy(INU, IAU, INL, IAL, D, AU, AL):
    """Full matrix as scipy CSR (for direct solves and dense checks)."""
    import scipy.sparse as sp
    n = len(D)
    ru = np.repeat(np.arange(n), np.diff(INU))
    rl = np.repeat(np.arange(n), np.diff(INL))
    rows = np.concatenate([np.arange(n), ru, rl])
    cols = np.concatenate([np.arange(n), np.asarray(IAU) - 1, np.asarray(IAL) - 1])
    vals = np.concatenate([D, AU, AL])
    return sp.csr_matrix((vals, (rows, cols)), shape=(n, n))


def lanczos_minmax(A, kmax):
    """LANCZOSMINMAX (src/src_inv_ap/m_spectral.f90:102-179) on a dense symmetric A: kmax Lanczos
    steps from the normalized ones vector (stopping when a beta is 0), then the extreme eigenvalues
    of the tridiagonal (the reference's TQLI; here LAPACK's symmetric eigensolver). Returns
    (emin, emax, alphas, betas)."""
    A = np.asarray(A, np.float64)
    n = A.shape[0]
    q = np.ones(n)
    q = q / np.sqrt(np.dot(q, q))
    res = A @ q                                   # :135
    al = [float(np.dot(q, res))]                  # :136
    res = res - al[0] * q                         # :137
    be = [float(np.sqrt(np.dot(res, res)))]       # :138
    for j in range(1, kmax):                      # :140-151
        if be[-1] == 0.0:
            break
        v = q
        q = res / be[j - 1]
        res = A @ q
        res = res - be[j - 1] * v
        al.append(float(np.dot(q, res)))
        res = res - al[j] * q
        be.append(float(np.sqrt(np.dot(res, res))))
    m = len(al)
    T = np.diag(al) + np.diag(be[:m - 1], 1) + np.diag(be[:m - 1], -1)
    ev = np.linalg.eigvalsh(T)
    return float(ev[0]), float(ev[-1]), np.array(al), np.array(be)


def spectral_alpha_v2(J, Cm, cd, kmax):
    """alphaspectralradius_v2 (m_spectral.f90:63-99): alpha = rho(J Cm J^T) / max(Cd)."""
    JCMJT = J @ (Cm @ J.T)
    emin, emax, _, _ = lanczos_minmax(JCMJT, kmax)
    numer = max(abs(emax), abs(emin))
    denom = max(0.0, float(np.max(cd)))
    return abs(numer / denom), emin, emax


def model_roughening(n4, nphys1, ele2model, nmodel, iactive=None):
    """GENDQ (src/src_inv_ap/m_modelroughness.f90:575-862) step by step: the land-element
    connection matrix D (diagonal = number of land neighbours, -1 per neighbour, :634-676), the
    element-to-model map Q (:732-743), Q^T D Q as a structural product (the reference multiplies
    with MKL, which keeps every structural entry), every value replaced by -1 and the diagonal by
    (row length - 1) (:770-790), then the active rows and columns (:795-830). Neighbours are
    found with a dictionary of the faces' node sets. Returns a dense (nactive, nactive) array."""
    n4 = np.asarray(n4).reshape(-1, 4)
    ntet = n4.shape[0]
    nphys2 = ntet - nphys1
    owner = {}
    nbrs = [[] for _ in range(nphys2)]
    for i in range(nphys2):
        t = n4[nphys1 + i]
        for f in range(4):
            key = frozenset(int(t[k]) for k in range(4) if k != f)
            if key in owner:
                j = owner[key]
                nbrs[i].append(j)
                nbrs[j].append(i)
            else:
                owner[key] = i
    D = np.zeros((nphys2, nphys2))
    for i in range(nphys2):
        D[i, i] = len(nbrs[i])
        for j in nbrs[i]:
            D[i, j] = -1.0
    Q = np.zeros((nphys2, nmodel))
    Q[np.arange(nphys2), np.asarray(ele2model) - 1] = 1.0
    S = (Q.T @ (np.abs(D) > 0) @ Q) > 0          # structure of Q^T D Q
    R = np.where(S, -1.0, 0.0)
    for m in range(nmodel):
        j1 = int(S[m].sum())
        if j1 == 1:
            raise ValueError(f"model {m + 1} has no neighbour model")
        R[m, m] = j1 - 1
    act = np.ones(nmodel, bool) if iactive is None else np.asarray(iactive, bool)
    return R[np.ix_(act, act)]


# ---------------------------------------------------------------------------------------------
# Nodal gradient-space correction of the edge preconditioner (no reference counterpart: the
# reference's COCG13_single offers IC / SSOR / DIAG only, COCG13_single.f90:107-130). Written with
# scipy sparse products, a route independent of the library's per-node loops.

def discrete_gradient(line, node, dirichlet):
    """G (nline x nn): column a is the gradient of interior node a's hat function, +1 on edges that
    end at the node and -1 on edges that start there (an edge points from line(i,1) to
    line(i,2)). Nodes on a Dirichlet edge are left out. Returns (G csr, interior node ids)."""
    import scipy.sparse as sp
    line = np.asarray(line).reshape(-1, 2) - 1
    n = line.shape[0]
    bnd = np.zeros(node, bool)
    d = np.asarray(dirichlet, bool)
    bnd[line[d, 0]] = True
    bnd[line[d, 1]] = True
    inner = np.nonzero(~bnd)[0]
    Gfull = sp.csr_matrix((np.r_[-np.ones(n), np.ones(n)], (np.r_[np.arange(n), np.arange(n)],
                                                          np.r_[line[:, 0], line[:, 1]])), shape=(n, node))
    return Gfull[:, inner].tocsr(), inner


def grad_correction(A, G, omega_eff=1.0):
    """The correction operator r -> G (g / (i w mnd)), g = G^T r, mnd = diag(G^T Im(A) G) / w for A
    given at angular frequency w (the imaginary part of A is w M). Jacobi on G^T A G, whose real
    part G^T K G vanishes for a curl-curl K."""
    mnd = np.asarray((G.T @ (A.imag @ G)).diagonal())
    dn = 1.0 / (1j * mnd)
    return lambda r: G @ (dn * (G.T @ r)), mnd


def node_blocks(line, width=8):
    """Node blocks of AFX_PC_NODEBLK (no reference counterpart): maximal runs of consecutive edges
    with the same first node (line(1, e)), cut into pieces of at most `width` edges from the
    start of the run. Returns the list of 0-based edge index arrays."""
    first = np.asarray(line).reshape(-1, 2)[:, 0]
    n = first.size
    starts = np.r_[0, np.nonzero(np.diff(first))[0] + 1, n]
    out = []
    for a, c in zip(starts[:-1], starts[1:]):
        for s0 in range(a, c, width):
            out.append(np.arange(s0, min(c, s0 + width)))
    return out


def _bf16(x32):
    """float32 array rounded to bfloat16 (nearest even on the float32 bits), returned as float32."""
    u = np.ascontiguousarray(x32, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def node_block_inverse(A, blocks):
    """The node-block Jacobi operator: each diagonal block of A inverted (numpy LU with pivoting,
    batched over blocks padded to 8 x 8 with the identity), symmetrised as (X + X^T) / 2, rounded
    to complex FP32 and then each part to bfloat16 (nearest even), as a scipy CSR matrix. The
    blocks' entries are scattered from A's CSR entries whose row and column lie in the same
    block."""
    import scipy.sparse as sp
    A = sp.csr_matrix(A)
    n = A.shape[0]
    W = 8
    nb = len(blocks)
    blk = np.empty(n, np.int64)
    loc = np.empty(n, np.int64)
    size = np.array([len(idx) for idx in blocks])
    if size.max() > W:
        raise ValueError("node blocks of more than 8 rows")
    starts = np.array([idx[0] for idx in blocks])
    blk[:] = np.repeat(np.arange(nb), size)
    loc[:] = np.arange(n) - np.repeat(starts, size)
    rows = np.repeat(np.arange(n), np.diff(A.indptr))
    cols = A.indices
    m = blk[rows] == blk[cols]
    D = np.zeros((nb, W, W), np.complex128)
    # the padding of a block smaller than 8 x 8 is the identity (it keeps the batch invertible)
    D[:, np.arange(W), np.arange(W)] = np.where(np.arange(W)[None, :] < size[:, None], 0.0, 1.0)
    D[blk[rows[m]], loc[rows[m]], loc[cols[m]]] = A.data[m]
    X = np.linalg.inv(D)
    X = (0.5 * (X + np.transpose(X, (0, 2, 1)))).astype(np.complex64)
    X = _bf16(X.real).astype(np.float64) + 1j * _bf16(X.imag).astype(np.float64)
    # row r's stored entries: its block's columns start .. start + size - 1
    ii = np.repeat(np.arange(n), size[blk])
    off = np.arange(W)[None, :]
    jj = (starts[blk][:, None] + off)[off < size[blk][:, None]]
    vals = X[blk[ii], loc[ii], jj - starts[blk[ii]]]
    return sp.csr_matrix((vals, (ii, jj)), shape=(n, n))


def cocg_precond(A, b, M, tol=1e-8, maxit=100000, x0=None):
    """COCG of COCG13_single.f90:181-323 (same recurrence, unconjugated forms, stop on the
    recurrence residual over ||b||) with an arbitrary complex-symmetric preconditioner M (a
    function r -> z). Returns (x, iterations, recurrence residual)."""
    x = np.zeros_like(b) if x0 is None else np.array(x0, np.complex128)
    r = b - A @ x
    bn = np.linalg.norm(b)
    rho1 = p = None
    res = np.linalg.norm(r) / bn
    for it in range(1, maxit + 1):
        z = M(r)
        rho = r @ z
        p = z if it == 1 else z + (rho / rho1) * p
        q = A @ p
        al = rho / (p @ q)
        x += al * p
        r -= al * q
        res = np.linalg.norm(r) / bn
        if res <= tol:
            return x, it, res
        rho1 = rho
    return x, maxit + 1, res


def cocg_parallel_phases(o, nthreads, reps=5):
    """Seconds per call of orc_cocg_parallel_phases (one COCG13_single iteration's SpMV, dot
    products and vector updates split over nthreads OpenMP threads) on the split system o; the
    median of reps calls after one warm-up call."""
    import time
    n = len(o["D"])
    rng = np.random.default_rng(0)
    vec = lambda: (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex128)  # noqa: E731
    X, R, Z, P, Q = vec(), vec(), vec(), vec(), np.zeros(n, np.complex128)
    D, AU, AL = _c128(o["D"]), _c128(o["AU"]), _c128(o["AL"])
    INU, IAU, INL, IAL = _i32(o["INU"]), _i32(o["IAU"]), _i32(o["INL"]), _i32(o["IAL"])
    f = lib().orc_cocg_parallel_phases
    f.restype = C.c_double
    args = (C.c_int(n), _p(D), _p(INU), _p(IAU), _p(AU), _p(INL), _p(IAL), _p(AL), _p(X), _p(R), _p(Z), _p(P),
            _p(Q), C.c_int(nthreads))
    f(*args)
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        f(*args)
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts))
