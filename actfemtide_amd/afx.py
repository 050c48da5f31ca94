"""Python binding of libafx.so (include/afx.h) and the host-side mirror of the reference's
solver call surface.

The reference entry points this module mirrors (same argument meaning and error codes):
  COCG13_single        src/solver/COCG13_single.f90:9-341        -> cocg13_single()
  solveCG              src/solver_mpi/forward_bxyz_mpi.f90:70-94 -> solve_cg()
  forward_bxyz         src/solver_mpi/forward_bxyz_mpi.f90:3-67  -> forward_bxyz()
  PREPOBSCOEFF         src/solver_mpi/n_ebfem_bxyz_mpi.f90:653-726 -> System.prepobscoeff()
  CALOBSEBCOMP/RESP    src/solver_mpi/n_ebfem_bxyz_mpi.f90:397-423,570-609 -> calobsebcomp()
  solvePARDISOinv      src/src_inv/solvePARDISOinv.f90:5-244     -> solve_pardiso_inv()
  solvePARDISOjointinv src/src_inv_joint/solvePARDISOjointinv.f90:5-333 -> solve_pardiso_jointinv_host()
  genjacobian1         src/src_inv_ap/m_jacobian_ap.f90:174-386  -> System.jacobian_device(),
                                                                    jacobian_ampphase_device()
  getnewmodel          src/src_inv_ap/m_jacobian_ap.f90:623-787  -> gn_update()
There is no CPU fallback: every call goes through the HIP library, and loading fails loudly
when libafx.so is missing.
"""
from __future__ import annotations

import atexit
import ctypes as C
import dataclasses
import os
import sys
import weakref

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libafx.so")

AFX_OK = 0
AFX_ERR_MAXIT = -100
AFX_ERR_BADOPT = 101
AFX_ERR_BREAKDOWN = -200
AFX_ERR_HIP = -300
AFX_ERR_ARG = -400
AFX_ERR_GEOM = -500
AFX_ERR_PATTERN = -600
AFX_ERR_OUTSIDE = -700

AFX_PC_GRAD = 0x100
PC = {"NONE": 0, "IC": 1, "ILU": 1, "SSOR": 2, "DIAG": 3, "SSOR_BJ": 4,
      # edge preconditioner + the nodal gradient-space correction (needs System.set_nodes)
      "SSOR+GRAD": 2 | AFX_PC_GRAD, "SSOR_BJ+GRAD": 4 | AFX_PC_GRAD, "DIAG+GRAD": 3 | AFX_PC_GRAD,
      # node-block Jacobi (needs System.set_nodes), alone or with the nodal correction
      "NODEBLK": 5, "NODEBLK+GRAD": 5 | AFX_PC_GRAD}
FP64, MIXED = 0, 1


class AfxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"afx error {code}: {msg}")
        self.code = code


class afx_opts(C.Structure):
    _fields_ = [("precond", C.c_int), ("sigma_diag", C.c_double), ("sigma", C.c_double), ("tol", C.c_double),
                ("maxit", C.c_int), ("nset", C.c_int), ("precision", C.c_int), ("inner_tol", C.c_double),
                ("max_outer", C.c_int), ("chunk", C.c_int), ("check_true", C.c_int)]


class afx_stats(C.Structure):
    _fields_ = [("iter", C.c_int), ("error", C.c_int), ("resid", C.c_double), ("true_resid", C.c_double),
                ("outer", C.c_int)]


# exported symbols and their prototypes (argtypes, restype)
_VP, _I, _D = C.c_void_p, C.c_int, C.c_double
SIGNATURES = {
    "afx_default_opts": ([C.POINTER(afx_opts)], None),
    "afx_last_error": ([], C.c_char_p),
    "afx_create": ([_I, _VP, _VP, _VP, _VP, C.POINTER(_VP)], _I),
    "afx_create_stacked": ([_I, _VP, _I, _VP, _VP, _VP, _VP, C.POINTER(_VP)], _I),
    "afx_stacked_count": ([_VP, C.POINTER(_I)], _I),
    "afx_create_from_n6line": ([_I, _I, _VP, C.POINTER(_VP)], _I),
    "afx_destroy": ([_VP], _I),
    "afx_get_sizes": ([_VP, _VP, _VP, _VP, _VP, _VP], _I),
    "afx_get_pattern": ([_VP, _VP, _VP, _VP, _VP], _I),
    "afx_set_values": ([_VP, _VP, _VP, _VP], _I),
    "afx_get_values": ([_VP, _VP, _VP, _VP], _I),
    "afx_set_mesh": ([_VP, _I, _I, _VP, _VP, _VP, _VP, _VP], _I),
    "afx_assemble": ([_VP, _VP, _D, _VP], _I),
    "afx_set_omega": ([_VP, _D], _I),
    "afx_genbc": ([_VP, _VP, _D, _D, _D, _D, _VP], _I),
    "afx_set_bc": ([_VP, _VP, _I, _VP, _I], _I),
    "afx_set_nodes": ([_VP, _I, _VP, _VP], _I),
    "afx_precond_setup": ([_VP, _I, _D, _D], _I),
    "afx_solve": ([_VP, _I, _VP, _I, _VP, _I, C.POINTER(afx_opts), _VP], _I),
    "afx_solve_device": ([_VP, _I, _VP, _VP, _VP, C.POINTER(afx_opts), _VP], _I),
    "afx_solve_stacked_device": ([_VP, _VP, _VP, C.POINTER(afx_opts), _VP], _I),
    "afx_true_residual_device": ([_VP, _I, _VP, _VP, _VP, _VP], _I),
    "afx_true_residual_device_noalloc": ([_VP, _I, _VP, _VP, _VP, _VP], _I),
    "afx_scale_initial_guess_device": ([_VP, _I, _VP, _VP, _VP, _VP, _VP, _VP], _I),
    "afx_autotune": ([_VP, _I, _I], _I),
    "afx_kernel_timing": ([_VP, _I, _VP, _VP, _VP], _I),
    "afx_spmv": ([_VP, _VP, _VP], _I),
    "afx_spmv_device": ([_VP, _VP, _VP, _VP], _I),
    "afx_precond_apply": ([_VP, _VP, _VP], _I),
    "afx_get_precond_diag": ([_VP, _VP], _I),
    "afx_cocg13_single": ([_I, _I, _I, _VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP, C.c_char_p, _D, _D,
                           C.POINTER(_D), C.POINTER(_I), C.POINTER(_I), _I], _I),
    "afx_cocg13_single_release": ([], _I),
    "afx_solve_cg": ([_I, _I, _I, _VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP], _I),
    "afx_solve_pardiso": ([_I, _I, _VP, _VP, _VP, _VP, _VP, _VP, _VP, _I, _VP, _VP, _VP, _VP, _D], _I),
    "afx_solve_pardiso_inv": ([_I, _I, _VP, _VP, _VP, _VP, _VP, _VP, _VP, _I, _VP, _VP, _VP, _VP, _I, _VP, _VP, _VP,
                               _VP, _VP, _I, _VP, _VP, _VP, _VP, _D], _I),
    "afx_solve_pardiso_jointinv": ([_I, _I, _I, _I, _VP, _VP, _VP, _VP, _VP, _VP, _VP, _I, _VP, _VP, _VP, _VP, _VP,
                                    _VP, _I, _VP, _VP, _VP, _VP, _VP, _I, _VP, _VP, _VP, _VP, _I, _VP, _VP, _VP, _VP,
                                    _I, _VP, _VP, _VP, _VP, _D], _I),
    "afx_solve_pardiso_release": ([], _I),
    "afx_prepobscoeff": ([_VP, _I, _VP, _VP, _VP, _VP, _VP], _I),
    "afx_obs_apply_device": ([_VP, _I, _VP, _VP, _I, _VP, _VP], _I),
    "afx_ixyh": ([_I, _VP, _VP, _VP, _VP, _VP, _I, _VP, _VP, _D, _VP, _VP], _I),
    "afx_adjoint_rhs_device": ([_VP, _I, _VP, _VP, _VP, _VP], _I),
    "afx_jacobian_device": ([_VP, _I, _VP, _VP, _I, _VP, _VP, _VP, _D, _D, _VP], _I),
    "afx_jacobian_ampphase_device": ([_I, _I, _VP, _VP, _VP], _I),
    "afx_gn_update_device": ([_I, _I, _VP, _VP, _VP, _VP, _VP, _D, _VP, _VP, _VP, _VP, _VP, _VP], _I),
    "afx_spectral_alpha_device": ([_I, _I, _VP, _VP, _VP, _VP, _VP, _I, _VP, _VP, _VP], _I),
    "afx_model_roughening": ([_I, _VP, _I, _VP, _I, _VP, _VP, _VP, _VP, _VP, _VP], _I),
    "afx_dgemm_nt_device": ([_I, _I, _I, _VP, _I, _VP, _I, _VP, _I, _I], _I),
    "afx_inv_last_error": ([], C.c_char_p),
}

_lib = None


def lib():
    """Load libafx.so (raises when it is missing: there is no fallback path)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m actfemtide_amd.build`")
        # PyTorch carries its own HIP runtime under the same soname (libamdhip64.so.7). Loading it
        # first makes libafx.so bind to that one instance; loaded the other way round, the process
        # holds two HIP runtimes and the second to initialise finds no GPU.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
        # the drop-in's kept system is freed while the HIP runtime is still up (an exit handler
        # of the library itself would run after the runtime has been torn down)
        atexit.register(_release_at_exit)
    return _lib


# every open System, so that the exit handler can free them before the runtime goes away
_open_systems = weakref.WeakSet()


def _release_at_exit():
    if _lib is None:
        return
    for s in list(_open_systems):
        try:
            s.close()
        except Exception:
            pass
    _lib.afx_cocg13_single_release()
    _lib.afx_solve_pardiso_release()


def _err(code):
    if code not in (AFX_OK,):
        raise AfxError(code, lib().afx_last_error().decode())


def _p(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _c128(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


def default_opts(**kw) -> afx_opts:
    o = afx_opts()
    lib().afx_default_opts(C.byref(o))
    for k, v in kw.items():
        if k == "precond" and isinstance(v, str):
            v = PC[v.strip().upper()]
        setattr(o, k, v)
    return o


@dataclasses.dataclass
class Stats:
    iter: int
    error: int
    resid: float
    true_resid: float
    outer: int


def _stats(arr, k):
    return [Stats(arr[c].iter, arr[c].error, arr[c].resid, arr[c].true_resid, arr[c].outer) for c in range(k)]


class System:
    """One assembled linear system on the device (an afx_system handle)."""

    def __init__(self, handle):
        self.h = C.c_void_p(handle)
        n, npu, npl, ll, lu = (C.c_int() for _ in range(5))
        _err(lib().afx_get_sizes(self.h, C.byref(n), C.byref(npu), C.byref(npl), C.byref(ll), C.byref(lu)))
        self.n, self.npu, self.npl, self.nlev_l, self.nlev_u = n.value, npu.value, npl.value, ll.value, lu.value
        self.omega = None  # angular frequency of the values (afx_assemble / afx_set_omega)
        self.tuned_widths = set()  # column counts afx_autotune has chosen a block width for
        _open_systems.add(self)

    @classmethod
    def from_n6line(cls, nline: int, n6line: np.ndarray) -> "System":
        h = C.c_void_p()
        n6 = _i32(n6line)
        _err(lib().afx_create_from_n6line(nline, n6.shape[0], _p(n6), C.byref(h)))
        return cls(h.value)

    @classmethod
    def from_pattern(cls, INU, IAU, INL, IAL) -> "System":
        h = C.c_void_p()
        INU, IAU, INL, IAL = _i32(INU), _i32(IAU), _i32(INL), _i32(IAL)
        _err(lib().afx_create(len(INU) - 1, _p(INU), _p(IAU), _p(INL), _p(IAL), C.byref(h)))
        return cls(h.value)

    @classmethod
    def from_stacked(cls, row_off, INU, IAU, INL, IAL) -> "System":
        """Several independent systems stacked block-diagonally (afx_create_stacked); system q
        owns rows [row_off[q], row_off[q+1])."""
        h = C.c_void_p()
        ro = _i32(row_off)
        INU, IAU, INL, IAL = _i32(INU), _i32(IAU), _i32(INL), _i32(IAL)
        _err(lib().afx_create_stacked(len(ro) - 1, _p(ro), len(INU) - 1, _p(INU), _p(IAU), _p(INL), _p(IAL),
                                      C.byref(h)))
        return cls(h.value)

    @property
    def nsys(self) -> int:
        m = C.c_int()
        _err(lib().afx_stacked_count(self.h, C.byref(m)))
        return m.value

    def solve_stacked_device(self, dB, dX, **opts):
        """afx_solve_stacked_device: dB, dX torch complex128 GPU tensors of n elements; returns
        (one Stats per stacked system, return code)."""
        import torch
        for t in (dB, dX):
            if t.dtype != torch.complex128 or not t.is_cuda or not t.is_contiguous() or t.numel() != self.n:
                raise ValueError("dB, dX: contiguous complex128 GPU tensors of n elements")
        m = self.nsys
        o = default_opts(**opts)
        st = (afx_stats * m)()
        rc = lib().afx_solve_stacked_device(self.h, C.c_void_p(dB.data_ptr()), C.c_void_p(dX.data_ptr()),
                                            C.byref(o), st)
        if rc not in (AFX_OK, AFX_ERR_MAXIT, AFX_ERR_BREAKDOWN):
            _err(rc)
        return _stats(st, m), rc

    def close(self):
        if self.h:
            lib().afx_destroy(self.h)
            self.h = None

    def __del__(self, _finalizing=sys.is_finalizing):
        # during interpreter shutdown the HIP runtime (torch's) may already be gone: leave the
        # device memory to the process exit instead of calling into it (the module's globals,
        # `sys` among them, may already be cleared then, hence the bound default)
        if _finalizing():
            return
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def pattern(self):
        INU = np.zeros(self.n + 1, np.int32)
        INL = np.zeros(self.n + 1, np.int32)
        IAU = np.zeros(self.npu, np.int32)
        IAL = np.zeros(self.npl, np.int32)
        _err(lib().afx_get_pattern(self.h, _p(INU), _p(IAU), _p(INL), _p(IAL)))
        return INU, IAU, INL, IAL

    def set_values(self, D, AU, AL):
        D, AU, AL = _c128(D), _c128(AU), _c128(AL)
        _err(lib().afx_set_values(self.h, _p(D), _p(AU), _p(AL)))

    def get_values(self):
        D = np.zeros(self.n, np.complex128)
        AU = np.zeros(self.npu, np.complex128)
        AL = np.zeros(self.npl, np.complex128)
        _err(lib().afx_get_values(self.h, _p(D), _p(AU), _p(AL)))
        return D, AU, AL

    def set_mesh(self, case):
        self._xyz = _f64(case.xyz)
        self._n4 = _i32(case.n4)
        self._group = _i32(case.group)
        v = _c128(case.vxyz)
        f = _f64(case.fxyz)
        _err(lib().afx_set_mesh(self.h, case.node, case.ntet, _p(self._xyz), _p(self._n4), _p(self._group),
                                _p(v), _p(f)))

    def assemble(self, sigma_tet, omega: float) -> np.ndarray:
        """GENMAT + COPY_UL on the device; returns the tidal right-hand side b."""
        b = np.zeros(self.n, np.complex128)
        sig = _f64(sigma_tet)
        _err(lib().afx_assemble(self.h, _p(sig), omega, _p(b)))
        self.omega = omega
        return b

    def set_omega(self, omega: float):
        _err(lib().afx_set_omega(self.h, omega))
        self.omega = omega

    def genbc(self, line, zmin, zmax, xout, yout) -> np.ndarray:
        line = _i32(line)
        bc = np.zeros(self.n, np.uint8)
        _err(lib().afx_genbc(self.h, _p(line), zmin, zmax, xout, yout, _p(bc)))
        return bc.astype(bool)

    def set_nodes(self, node, line, dirichlet):
        """Edge-to-node incidence for the nodal gradient-space correction (afx_set_nodes): line
        (nline, 2) 1-based end nodes, dirichlet (nline,) mask."""
        ln = _i32(line).reshape(-1, 2)
        d = np.ascontiguousarray(dirichlet, dtype=np.uint8)
        _err(lib().afx_set_nodes(self.h, int(node), _p(ln), _p(d)))

    def set_bc(self, dirichlet, rhs=None):
        d = np.ascontiguousarray(dirichlet, dtype=np.uint8)
        if rhs is None:
            _err(lib().afx_set_bc(self.h, _p(d), 0, None, self.n))
            return None
        r = np.array(rhs, dtype=np.complex128, order="F", copy=True).reshape(self.n, -1, order="F")
        _err(lib().afx_set_bc(self.h, _p(d), r.shape[1], C.c_void_p(r.ctypes.data), self.n))
        return r[:, 0] if np.ndim(rhs) == 1 else r

    def precond_setup(self, kind="SSOR", sigma_diag=1.0, sigma=0.0):
        _err(lib().afx_precond_setup(self.h, PC[kind] if isinstance(kind, str) else kind, sigma_diag, sigma))

    def precond_diag(self):
        DD = np.zeros(self.n, np.complex128)
        _err(lib().afx_get_precond_diag(self.h, _p(DD)))
        return DD

    def precond_apply(self, r):
        r = _c128(r)
        z = np.zeros(self.n, np.complex128)
        _err(lib().afx_precond_apply(self.h, _p(r), _p(z)))
        return z

    def spmv(self, x):
        x = _c128(x)
        y = np.zeros(self.n, np.complex128)
        _err(lib().afx_spmv(self.h, _p(x), _p(y)))
        return y

    def spmv_device(self, dX, dY, timed=False):
        """dY = A dX on the device (torch complex128 tensors of n elements on the GPU). Returns the
        launch's duration in ms (HIP events on the library's stream) when timed, else None."""
        import torch
        for t in (dX, dY):
            if t.dtype != torch.complex128 or not t.is_cuda or not t.is_contiguous() or t.numel() != self.n:
                raise ValueError("dX, dY: contiguous complex128 GPU tensors of n elements")
        ms = C.c_double()
        _err(lib().afx_spmv_device(self.h, C.c_void_p(dX.data_ptr()), C.c_void_p(dY.data_ptr()),
                                   C.byref(ms) if timed else None))
        return ms.value if timed else None

    def autotune(self, nrhs=1, verbose=False):
        _err(lib().afx_autotune(self.h, nrhs, 1 if verbose else 0))

    def kernel_timing(self, reset=True):
        """(launches, iterations, kernel_ms) of the persistent COCG kernel since the last reset."""
        la, it, ms = C.c_longlong(), C.c_longlong(), C.c_double()
        _err(lib().afx_kernel_timing(self.h, 1 if reset else 0, C.byref(la), C.byref(it), C.byref(ms)))
        return la.value, it.value, ms.value

    def solve(self, B, X0=None, **opts):
        """Solve A X = B (host arrays, columns = right-hand sides); X0 is the initial guess."""
        B2 = np.asfortranarray(np.asarray(B, np.complex128).reshape(self.n, -1, order="F"))
        k = B2.shape[1]
        X = (np.zeros((self.n, k), np.complex128, order="F") if X0 is None
             else np.array(np.asarray(X0, np.complex128).reshape(self.n, -1, order="F"), order="F", copy=True))
        o = default_opts(**opts)
        st = (afx_stats * k)()
        rc = lib().afx_solve(self.h, k, C.c_void_p(B2.ctypes.data), self.n, C.c_void_p(X.ctypes.data), self.n,
                             C.byref(o), st)
        if rc not in (AFX_OK, AFX_ERR_MAXIT, AFX_ERR_BREAKDOWN):
            _err(rc)
        out = X[:, 0] if np.ndim(B) == 1 else X
        return out, _stats(st, k)

    def solve_device(self, dB, dX, omega_scale=None, **opts):
        """dB, dX: torch complex128 tensors on the GPU, shape (n, k) row-major (element (i, c) at i*k + c)."""
        k = 1 if dB.dim() == 1 else dB.shape[1]
        o = default_opts(**opts)
        st = (afx_stats * k)()
        sc = None if omega_scale is None else _f64(omega_scale)
        rc = lib().afx_solve_device(self.h, k, C.c_void_p(dB.data_ptr()), C.c_void_p(dX.data_ptr()), _p(sc),
                                    C.byref(o), st)
        if rc not in (AFX_OK, AFX_ERR_MAXIT, AFX_ERR_BREAKDOWN):
            _err(rc)
        return _stats(st, k), rc

    def true_residual_device(self, dB, dX, omega_scale=None):
        k = 1 if dB.dim() == 1 else dB.shape[1]
        res = np.zeros(k)
        sc = None if omega_scale is None else _f64(omega_scale)
        _err(lib().afx_true_residual_device(self.h, k, C.c_void_p(dB.data_ptr()), C.c_void_p(dX.data_ptr()),
                                            _p(sc), _p(res)))
        return res

    def true_residual_device_noalloc(self, dB, dX, omega_scale=None):
        """afx_true_residual_device_noalloc: the same norms with the library's static scratch."""
        k = 1 if dB.dim() == 1 else dB.shape[1]
        res = np.zeros(k)
        sc = None if omega_scale is None else _f64(omega_scale)
        _err(lib().afx_true_residual_device_noalloc(self.h, k, C.c_void_p(dB.data_ptr()), C.c_void_p(dX.data_ptr()),
                                                    _p(sc), _p(res)))
        return res

    def scale_initial_guess_device(self, dB, dX, omega_scale=None):
        """Minimal-residual scaling of the initial guesses dX in place (zero where scaling does
        not lower the residual below the zero guess's); returns (gamma, res_given, res_scaled)."""
        k = 1 if dB.dim() == 1 else dB.shape[1]
        g = np.zeros(2 * k)
        r0 = np.zeros(k)
        r1 = np.zeros(k)
        sc = None if omega_scale is None else _f64(omega_scale)
        _err(lib().afx_scale_initial_guess_device(self.h, k, C.c_void_p(dB.data_ptr()), C.c_void_p(dX.data_ptr()),
                                                  _p(sc), _p(g), _p(r0), _p(r1)))
        return g[0::2] + 1j * g[1::2], r0, r1

    def prepobscoeff(self, xyzobs):
        """PREPOBSCOEFF: (item (nobs,6) 1-based, val_e (3,nobs,6), val_b (3,nobs,6), iele (nobs,))."""
        xo = _f64(xyzobs).reshape(-1, 3)
        nobs = xo.shape[0]
        item = np.zeros((nobs, 6), np.int32)
        ve = np.zeros((3, nobs, 6))
        vb = np.zeros((3, nobs, 6))
        iele = np.zeros(nobs, np.int32)
        _err(lib().afx_prepobscoeff(self.h, nobs, _p(xo), _p(item), _p(ve), _p(vb), _p(iele)))
        return item, ve, vb, iele

    def obs_apply_device(self, item, val, dX):
        """f[o, c] = sum_j val[o, j] * X[item[o, j] - 1, c] for device X (n, ncol) complex128."""
        item, val = _i32(item), _f64(val)
        nobs = item.shape[0]
        ncol = 1 if dX.dim() == 1 else dX.shape[1]
        out = np.zeros((nobs, ncol), np.complex128)
        _err(lib().afx_obs_apply_device(self.h, nobs, _p(item), _p(val), ncol, C.c_void_p(dX.data_ptr()), _p(out)))
        return out

    def adjoint_rhs_device(self, stack, item, val, dB):
        """P^T columns of solvePARDISOinv into dB (n, nobs) complex128 device tensor."""
        st, it, v = _i32(stack), _i32(item), _f64(val)
        _err(lib().afx_adjoint_rhs_device(self.h, len(st) - 1, _p(st), _p(it), _p(v), C.c_void_p(dB.data_ptr())))

    def jacobian_device(self, dU, dx, par_ptr, par_tet, sigma_tet, omega, dlnsig_dm, dJ):
        """genjacobian1 accumulation: dJ (npar, nobs) complex128 device tensor is overwritten."""
        nobs = dU.shape[1]
        pp, pt, sg = _i32(par_ptr), _i32(par_tet), _f64(sigma_tet)
        _err(lib().afx_jacobian_device(self.h, nobs, C.c_void_p(dU.data_ptr()), C.c_void_p(dx.data_ptr()),
                                       len(pp) - 1, _p(pp), _p(pt), _p(sg), omega, dlnsig_dm,
                                       C.c_void_p(dJ.data_ptr())))


# ---------------------------------------------------------------------------------------
# mirrors of the reference's solver call surface

def cocg13_single(D, AL, INL, IAL, AU, INU, IAU, B, X, precond="IC", sigma_diag=1.0, sigma=0.0,
                  resid=1e-8, iter=3000, nset=1):
    """COCG13_single (src/solver/COCG13_single.f90:9-341) through afx_cocg13_single.

    D, AL, AU, B, X are complex128 numpy arrays updated in place exactly as the reference does
    (D/AU/AL scaled for nset=2, B scaled and restored, X = solution). Returns (resid, iter, error)."""
    for a in (D, AL, AU, B, X):
        if a.dtype != np.complex128 or not a.flags.c_contiguous:
            raise TypeError("D, AL, AU, B, X must be contiguous complex128 arrays")
    INL, IAL, INU, IAU = _i32(INL), _i32(IAL), _i32(INU), _i32(IAU)
    r = C.c_double(resid)
    it = C.c_int(iter)
    err = C.c_int(0)
    pc = precond.ljust(20).encode()[:20]
    rc = lib().afx_cocg13_single(len(D), len(AL), len(AU), _p(D), _p(AL), _p(INL), _p(IAL), _p(AU), _p(INU),
                                 _p(IAU), _p(B), _p(X), pc, sigma_diag, sigma, C.byref(r), C.byref(it),
                                 C.byref(err), nset)
    if rc not in (AFX_OK, AFX_ERR_MAXIT, AFX_ERR_BREAKDOWN, AFX_ERR_BADOPT):
        _err(rc)
    return r.value, it.value, err.value


def cocg13_single_release():
    """Free the device system that afx_cocg13_single keeps between calls on one pattern."""
    _err(lib().afx_cocg13_single_release())


def solve_cg(D, AL, INL, IAL, AU, INU, IAU, rhs):
    """solveCG (src/solver_mpi/forward_bxyz_mpi.f90:70-94): 'SSOR', RESID 1e-4, ITER 3000, X0 = 0."""
    X = np.zeros(len(D), np.complex128)
    D, AL, AU, rhs = _c128(D), _c128(AL), _c128(AU), _c128(rhs)
    INL, IAL, INU, IAU = _i32(INL), _i32(IAL), _i32(INU), _i32(IAU)
    rc = lib().afx_solve_cg(len(D), len(AL), len(AU), _p(D), _p(AL), _p(INL), _p(IAL), _p(AU), _p(INU), _p(IAU),
                            _p(rhs), _p(X))
    if rc not in (AFX_OK, AFX_ERR_MAXIT):
        _err(rc)
    return X


def solve_pardiso(D, AU, AL, INU, IAU, INL, IAL, rhs, node=0, line=None, line_bc=None, tol=1e-12):
    """solvePARDISO (src/solver/solvePARDISO.f90:6-16) through afx_solve_pardiso on host arrays:
    rhs (n, nsr) or (n,) complex; returns xout of the same shape. With the edges' end nodes
    (line (nline, 2) 1-based), the node count and the Dirichlet mask, DIAG + the nodal correction,
    else SSOR; FP64 COCG to the true residual tol. Raises on AFX_ERR_*; -100 leaves xout as reached
    and is returned as the second value."""
    D, AU, AL = _c128(D), _c128(AU), _c128(AL)
    INU, IAU, INL, IAL = _i32(INU), _i32(IAU), _i32(INL), _i32(IAL)
    r = np.asfortranarray(np.asarray(rhs, np.complex128).reshape(len(D), -1, order="F"))
    x = np.zeros_like(r, order="F")
    ln = None if line is None else _i32(line).reshape(-1, 2)
    bc = None if line_bc is None else np.ascontiguousarray(line_bc, dtype=np.uint8)
    rc = lib().afx_solve_pardiso(len(D), r.shape[1], _p(INU), _p(IAU), _p(INL), _p(IAL), _p(D), _p(AU), _p(AL),
                                 int(node), _p(ln), _p(bc), C.c_void_p(r.ctypes.data), C.c_void_p(x.ctypes.data),
                                 float(tol))
    if rc not in (AFX_OK, AFX_ERR_MAXIT):
        _err(rc)
    return (x[:, 0] if np.ndim(rhs) == 1 else x), rc


def _pack_ptr(PTR, ncomp, flags=None):
    """Observation matrices as the C ABI takes them: a list of up to ncomp real CRS matrices
    (stack 0-based, item 1-based, val) or None, the flagged ones with the same row count.
    Returns (flags, nobs, stacks (ncomp, nobs+1), base, item, val); unflagged components are
    left out."""
    Ps = list(PTR) + [None] * (ncomp - len(PTR))
    fl = np.array([0 if (P is None or (flags is not None and not flags[c])) else 1 for c, P in enumerate(Ps)],
                  np.int32)
    Ps = [P if fl[c] else None for c, P in enumerate(Ps)]  # the row count is the flagged components'
    nobs = max((len(P[0]) - 1 for P in Ps if P is not None), default=0)
    stacks = np.zeros((ncomp, nobs + 1), np.int32)
    base = np.zeros(ncomp, np.int32)
    items, vals, off = [], [], 0
    for c, P in enumerate(Ps):
        if P is None:
            continue
        stacks[c] = _i32(P[0])
        base[c] = off
        items.append(_i32(P[1]))
        vals.append(_f64(P[2]))
        off += len(P[1])
    item = _i32(np.concatenate(items)) if items else np.zeros(1, np.int32)
    val = _f64(np.concatenate(vals)) if vals else np.zeros(1)
    return fl, nobs, stacks, base, item, val


def _unpack_ut(flags, nobs, n, cap, ust, ntot, uit, uva):
    import scipy.sparse as sp
    ut = []
    for c in range(len(flags)):
        if not flags[c]:
            ut.append(0)
            continue
        k = int(ntot[c])
        ut.append(sp.csr_matrix((uva[c * cap:c * cap + k], uit[c * cap:c * cap + k] - 1, ust[c]), shape=(nobs, n)))
    return ut


def solve_pardiso_inv_host(D, AU, AL, INU, IAU, INL, IAL, rhs, PTR, iflag_comp=None, node=0, line=None,
                           line_bc=None, tol=1e-12, ut_cap=None):
    """solvePARDISOinv (src/src_inv/solvePARDISOinv.f90:5-23) through afx_solve_pardiso_inv on host
    arrays: PTR a list of 5 observation matrices (stack 0-based, item 1-based, val) or None, all
    with the same row count. Returns (xout, ut, rc) with ut[c] a scipy CSR [nobs x n] (0 when not
    flagged)."""
    D, AU, AL = _c128(D), _c128(AU), _c128(AL)
    INU, IAU, INL, IAL = _i32(INU), _i32(IAU), _i32(INL), _i32(IAL)
    n = len(D)
    r = np.asfortranarray(np.asarray(rhs, np.complex128).reshape(n, -1, order="F"))
    x = np.zeros_like(r, order="F")
    flags, nobs, stacks, base, item, val = _pack_ptr(PTR, 5, iflag_comp)
    ln = None if line is None else _i32(line).reshape(-1, 2)
    bc = None if line_bc is None else np.ascontiguousarray(line_bc, dtype=np.uint8)
    cap = int(ut_cap) if ut_cap is not None else nobs * n
    ust = np.zeros((5, nobs + 1), np.int32)
    ntot = np.zeros(5, np.int32)
    uit = np.zeros(5 * max(cap, 1), np.int32)
    uva = np.zeros(5 * max(cap, 1), np.complex128)
    rc = lib().afx_solve_pardiso_inv(n, r.shape[1], _p(INU), _p(IAU), _p(INL), _p(IAL), _p(D), _p(AU), _p(AL),
                                     int(node), _p(ln), _p(bc), C.c_void_p(r.ctypes.data), C.c_void_p(x.ctypes.data),
                                     nobs, _p(flags), _p(stacks), _p(base), _p(item), _p(val), cap, _p(ust), _p(ntot),
                                     _p(uit), _p(uva), float(tol))
    if rc not in (AFX_OK, AFX_ERR_MAXIT):
        _err(rc)
    return x, _unpack_ut(flags, nobs, n, cap, ust, ntot, uit, uva), rc


def solve_pardiso_jointinv_host(D, AU, AL, INU, IAU, INL, IAL, rhs, rhs_mt, PTR, PTR_mt, iflag_comp=None,
                                act=True, mt=True, node=0, line=None, line_bc=None, tol=1e-12):
    """solvePARDISOjointinv (src/src_inv_joint/solvePARDISOjointinv.f90:5-26) through
    afx_solve_pardiso_jointinv: rhs (n, nsr) active-source columns, rhs_mt (n, 2) MT columns,
    PTR up to 5 and PTR_mt 4 observation matrices as in solve_pardiso_inv_host. Returns
    (xout, xout_mt, ut, ut_mt, rc); xout_mt and ut_mt stay zero / 0 when mt is off, xout and ut
    when act is off, and everything when no flagged PTR component has rows (the reference's early
    exit)."""
    D, AU, AL = _c128(D), _c128(AU), _c128(AL)
    INU, IAU, INL, IAL = _i32(INU), _i32(IAU), _i32(INL), _i32(IAL)
    n = len(D)
    r = np.asfortranarray(np.asarray(rhs, np.complex128).reshape(n, -1, order="F"))
    rm = np.asfortranarray(np.asarray(rhs_mt, np.complex128).reshape(n, 2, order="F"))
    x = np.zeros_like(r, order="F")
    xm = np.zeros_like(rm, order="F")
    flags, nobs, stacks, base, item, val = _pack_ptr(PTR, 5, iflag_comp)
    fl_mt, nobs_mt, st_mt, base_mt, item_mt, val_mt = _pack_ptr(PTR_mt, 4)
    ln = None if line is None else _i32(line).reshape(-1, 2)
    bc = None if line_bc is None else np.ascontiguousarray(line_bc, dtype=np.uint8)
    cap, cap_mt = nobs * n, nobs_mt * n
    ust, ntot = np.zeros((5, nobs + 1), np.int32), np.zeros(5, np.int32)
    uit, uva = np.zeros(5 * max(cap, 1), np.int32), np.zeros(5 * max(cap, 1), np.complex128)
    ust_mt, ntot_mt = np.zeros((4, nobs_mt + 1), np.int32), np.zeros(4, np.int32)
    uit_mt, uva_mt = np.zeros(4 * max(cap_mt, 1), np.int32), np.zeros(4 * max(cap_mt, 1), np.complex128)
    rc = lib().afx_solve_pardiso_jointinv(
        1 if act else 0, 1 if mt else 0, n, r.shape[1], _p(INU), _p(IAU), _p(INL), _p(IAL), _p(D), _p(AU), _p(AL),
        int(node), _p(ln), _p(bc), C.c_void_p(r.ctypes.data), C.c_void_p(rm.ctypes.data), C.c_void_p(x.ctypes.data),
        C.c_void_p(xm.ctypes.data), nobs, _p(flags), _p(stacks), _p(base), _p(item), _p(val), nobs_mt, _p(st_mt),
        _p(base_mt), _p(item_mt), _p(val_mt), cap, _p(ust), _p(ntot), _p(uit), _p(uva), cap_mt, _p(ust_mt),
        _p(ntot_mt), _p(uit_mt), _p(uva_mt), float(tol))
    if rc not in (AFX_OK, AFX_ERR_MAXIT):
        _err(rc)
    ut = _unpack_ut(flags if act else np.zeros(5, np.int32), nobs, n, cap, ust, ntot, uit, uva)
    ut_mt = _unpack_ut(fl_mt if mt and nobs else np.zeros(4, np.int32), nobs_mt, n, cap_mt, ust_mt, ntot_mt, uit_mt,
                       uva_mt)
    return x, xm, ut, ut_mt, rc


def solve_pardiso_release():
    _err(lib().afx_solve_pardiso_release())


def solve_pardiso_inv(s: System, rhs, PTR, iflag_comp=None, threshold=1e-10, max_cols=256, autotune=False,
                      **opts):
    """solvePARDISOinv (src/src_inv/solvePARDISOinv.f90:5-244) on the device system `s`.

    rhs (n, nsr) complex; PTR: list of up to 5 observation matrices, each a tuple
    (stack (nobs+1,) 0-based, item 1-based, val) = real CRS [nobs x n]; iflag_comp selects the
    components (default: all given). Returns (xout (n, nsr), ut) with ut[c] the adjoint solutions
    A^-1 P_o^T as a scipy CSR matrix [nobs x n] keeping |u| > threshold (convcomp_full2crs,
    m_matrix.f90:405-451; 0 for unselected components), and the solver stats of every column.
    The forward and adjoint solves use COCG to `tol` (default 1e-8) instead of a factorisation;
    adjoint columns run in batches of at most `max_cols` per call. With autotune, afx_autotune
    picks the block width of each batch size first (one column at a time, the k-column kernel, or
    stacked copies of the system in one launch, which is 3x faster for 12 columns at 32x32x24)."""
    import scipy.sparse as sp
    import torch
    dev = torch.device("cuda")
    rhs = np.asarray(rhs, np.complex128)
    if rhs.ndim == 1:
        rhs = rhs[:, None]
    n, nsr = rhs.shape
    if n != s.n:
        raise AfxError(AFX_ERR_ARG, "rhs rows != system size")
    opts = dict(dict(precond="SSOR", tol=1e-8, maxit=400000, chunk=2000), **opts)
    dB = torch.from_numpy(np.ascontiguousarray(rhs)).to(dev)
    dX = torch.zeros_like(dB)
    st_f, rc = s.solve_device(dB, dX, **opts)
    xout = dX.cpu().numpy()
    flags = list(iflag_comp) if iflag_comp is not None else [1] * len(PTR)
    ut, stats = [], [st_f]
    for c, P in enumerate(PTR):
        if not flags[c] or P is None:
            ut.append(0)
            continue
        stack, item, val = (np.asarray(a) for a in P)
        nobs = len(stack) - 1
        U = np.zeros((n, nobs), np.complex128)
        for o0 in range(0, nobs, max_cols):
            o1 = min(nobs, o0 + max_cols)
            sub = (stack[o0:o1 + 1] - stack[o0], item[stack[o0]:stack[o1]], val[stack[o0]:stack[o1]])
            dRHS = torch.empty((n, o1 - o0), dtype=torch.complex128, device=dev)
            s.adjoint_rhs_device(*sub, dRHS)
            dU = torch.zeros_like(dRHS)
            if autotune and o1 - o0 > 1 and (o1 - o0) not in s.tuned_widths:
                s.autotune(o1 - o0)
                s.tuned_widths.add(o1 - o0)
            st_a, _ = s.solve_device(dRHS, dU, **opts)
            stats.append(st_a)
            U[:, o0:o1] = dU.cpu().numpy()
        keep = np.abs(U) > threshold
        rows, cols = np.nonzero(keep.T)
        ut.append(sp.csr_matrix((U.T[rows, cols], (rows, cols)), shape=(nobs, n)))
    return xout, ut, stats


def _inv_err(code):
    if code != AFX_OK:
        raise AfxError(code, lib().afx_inv_last_error().decode())


def ixyh(xyzobs, exy, stack, item, val, vxyz, fxyz, sigma):
    """IXYHOUT (src/solver_mpi/n_ebfem_bxyz_mpi.f90:192-301) on the device: returns (ixyh_e,
    ixyh_vF, ixyh, vh), each (nobs, 2) complex, and h (nobs,). exy (nobs, 2) complex E at the
    points; coeff_vF as CSR (stack from 0, item 1-based); vxyz (node, 3) complex; fxyz (node, 3)."""
    xo = _f64(xyzobs).reshape(-1, 3)
    nobs = xo.shape[0]
    e = _c128(exy).reshape(nobs, 2)
    st, it, v = _i32(stack), _i32(item), _f64(val)
    vx = _c128(vxyz).reshape(-1, 3)
    fx = _f64(fxyz).reshape(-1, 3)
    out = np.zeros((nobs, 8), np.complex128)
    h = np.zeros(nobs)
    _err(lib().afx_ixyh(nobs, _p(xo), _p(e), _p(st), _p(it), _p(v), vx.shape[0], _p(vx), _p(fx), float(sigma),
                        _p(out), _p(h)))
    return out[:, 0:2], out[:, 2:4], out[:, 4:6], out[:, 6:8], h


def jacobian_ampphase_device(dJ, dbe, dout):
    """genjacobian1's real rows (m_jacobian_ap.f90:338-345): dJ (npar, nobs) complex128, dbe (nobs,)
    complex128, dout (2, nobs, npar) float64 -- all torch device tensors."""
    npar, nobs = dJ.shape
    _inv_err(lib().afx_jacobian_ampphase_device(npar, nobs, C.c_void_p(dJ.data_ptr()), C.c_void_p(dbe.data_ptr()),
                                                C.c_void_p(dout.data_ptr())))


def dgemm_nt_device(dA, dB, dC, accumulate=False):
    M, K = dA.shape
    N = dB.shape[0]
    _inv_err(lib().afx_dgemm_nt_device(M, N, K, C.c_void_p(dA.data_ptr()), K, C.c_void_p(dB.data_ptr()), K,
                                       C.c_void_p(dC.data_ptr()), N, 1 if accumulate else 0))


def gn_update(dJ, Cm, cd, alpha, dobs, dcal, model, model_ref):
    """getnewmodel + solvebeta (src/src_inv_ap/m_jacobian_ap.f90:623-787) on the device: dJ a
    (ndat, nmodel) float64 torch device tensor, Cm a scipy sparse symmetric (nmodel, nmodel)
    matrix, the rest host arrays. Returns (model_new, beta)."""
    import scipy.sparse as sp
    ndat, nmodel = dJ.shape
    Cm = sp.csr_matrix(Cm)
    ptr, col, val = _i32(Cm.indptr), _i32(Cm.indices), _f64(Cm.data)
    cd, dobs, dcal = _f64(cd), _f64(dobs), _f64(dcal)
    model, model_ref = _f64(model), _f64(model_ref)
    out = np.zeros(nmodel)
    beta = np.zeros(ndat)
    _inv_err(lib().afx_gn_update_device(ndat, nmodel, C.c_void_p(dJ.data_ptr()), _p(ptr), _p(col), _p(val), _p(cd),
                                        float(alpha), _p(dobs), _p(dcal), _p(model), _p(model_ref), _p(out),
                                        _p(beta)))
    return out, beta


def spectral_alpha(dJ, Cm, cd, kmax=30):
    """alphaspectralradius_v2 + LANCZOSMINMAX (src/src_inv_ap/m_spectral.f90:63-179) on the device:
    alpha = rho(J Cm J^T) / max(cd) from kmax Lanczos steps. dJ a (ndat, nmodel) float64 torch
    device tensor, Cm scipy sparse symmetric, cd host. Returns (alpha, emin, emax)."""
    import scipy.sparse as sp
    ndat, nmodel = dJ.shape
    Cm = sp.csr_matrix(Cm)
    al, lo, hi = C.c_double(), C.c_double(), C.c_double()
    _inv_err(lib().afx_spectral_alpha_device(ndat, nmodel, C.c_void_p(dJ.data_ptr()), _p(_i32(Cm.indptr)),
                                             _p(_i32(Cm.indices)), _p(_f64(Cm.data)), _p(_f64(cd)), int(kmax),
                                             C.byref(al), C.byref(lo), C.byref(hi)))
    return al.value, lo.value, hi.value


def forward_system(case, omega: float, sigma_tet=None) -> tuple:
    """forward_bxyz steps INITIALIZE .. SET_BC_ICCG (forward_bxyz_mpi.f90:43-58) on the device:
    pattern from n6line, GENMAT + COPY_UL, GENBCCSEM, SET_BC_ICCG. Returns (System, b, dirichlet)."""
    s = System.from_n6line(case.nline, case.n6line)
    s.set_mesh(case)
    b = s.assemble(case.sigma_tet() if sigma_tet is None else sigma_tet, omega)
    xout, yout, zmin, zmax = case.bounds
    bc = s.genbc(case.line, zmin, zmax, xout, yout)
    b = s.set_bc(bc, b)
    s.set_nodes(case.node, case.line, bc)
    return s, b, bc


def forward_bxyz(case, omega: float, precond="SSOR", tol=1e-8, maxit=200000, **opts):
    """forward_bxyz (forward_bxyz_mpi.f90:3-67) with the COCG solve on the device."""
    s, b, bc = forward_system(case, omega)
    x, st = s.solve(b, precond=precond, tol=tol, maxit=maxit, **opts)
    return s, b, x, st[0]


def amp_phase(c):
    """amp / phase [deg] of CALOBSRESP (n_ebfem_bxyz_mpi.f90:612-626)."""
    return np.sqrt(c.real ** 2 + c.imag ** 2), np.degrees(np.arctan2(c.imag, c.real))


def calobsebcomp(s: System, item, val_e, val_b, dX, omega):
    """CALOBSEBCOMP (n_ebfem_bxyz_mpi.f90:397-423): fields [Bx, By, Bz, Ex, Ey] (5, nobs, ncol),
    B = face rows . A, E = -i*omega * edge rows . A."""
    out = []
    for c in range(3):
        out.append(s.obs_apply_device(item, val_b[c], dX))
    for c in range(2):
        out.append(-1j * omega * s.obs_apply_device(item, val_e[c], dX))
    return np.stack(out)
