"""The Fortran iso_c_binding module (include/afx_mod.f90) compiles with flang, links against
libafx.so, and -- on the GPU -- runs the COCG13_single drop-in from a Fortran program on the
reference's own assembled system (the tiny golden fixture)."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FLANG = os.environ.get("FC", "/opt/rocm/llvm/bin/flang")
LIBDIR = os.path.join(ROOT, "actfemtide_amd")

PROG_OPTS = """
program t
  use afx
  implicit none
  type(afx_opts) :: o
  call afx_default_opts(o)
  print '(i0,1x,es12.5,1x,i0,1x,i0)', o%precond, o%tol, o%maxit, o%nset
end program
"""

PROG_COCG = """
program t
  use afx
  implicit none
  integer(4) :: NP, NPL, NPU, ITER, ERROR, u
  complex(8), allocatable :: D(:), AU(:), AL(:), B(:), X(:)
  integer(4), allocatable :: INU(:), INL(:), IAU(:), IAL(:)
  real(8) :: RESID
  open(newunit=u, file='sys.bin', access='stream', form='unformatted', status='old')
  read(u) NP, NPL, NPU
  allocate(D(NP), AU(NPU), AL(NPU), B(NP), X(NP), INU(0:NP), INL(0:NP), IAU(NPU), IAL(NPL))
  read(u) INU, INL, IAU, IAL, D, AU, AL, B
  close(u)
  X = (0.d0, 0.d0)
  RESID = 1.d-8
  ITER = 20000
  call afx_cocg13_single_f(NP, NPL, NPU, D, AL, INL, IAL, AU, INU, IAU, B, X, 'SSOR', &
                           1.d0, 0.d0, RESID, ITER, ERROR, 1)
  open(newunit=u, file='x.bin', access='stream', form='unformatted', status='replace')
  write(u) X
  close(u)
  print '(i0,1x,i0,1x,es12.5)', ITER, ERROR, RESID
end program
"""


def _build(tmp, src):
    if not os.path.exists(FLANG):
        pytest.skip("flang not available")
    with open(os.path.join(tmp, "t.f90"), "w") as f:
        f.write(src)
    cmd = [FLANG, "-J", tmp, os.path.join(ROOT, "include", "afx_mod.f90"), os.path.join(tmp, "t.f90"),
           f"-L{LIBDIR}", "-lafx", f"-Wl,-rpath,{LIBDIR}", "-o", os.path.join(tmp, "t")]
    r = subprocess.run(cmd, capture_output=True, text=True, cwd=tmp)
    assert r.returncode == 0, r.stdout + r.stderr
    return os.path.join(tmp, "t")


def test_fortran_module_links_and_reads_defaults(tmp_path):
    exe = _build(str(tmp_path), PROG_OPTS)
    r = subprocess.run([exe], capture_output=True, text=True, cwd=str(tmp_path), timeout=60)
    assert r.returncode == 0, r.stderr
    pc, tol, maxit, nset = r.stdout.split()
    assert int(pc) == 2 and abs(float(tol) - 1e-8) < 1e-20 and int(nset) == 1 and int(maxit) > 0


@pytest.mark.gpu
def test_fortran_dropin_cocg13_single(tmp_path):
    from tests.test_golden import load
    g = load("tiny_ssor")
    exe = _build(str(tmp_path), PROG_COCG)
    NP, NPU = len(g["D"]), len(g["AU"])
    NPL = len(g["IAL"])
    with open(os.path.join(str(tmp_path), "sys.bin"), "wb") as f:
        np.array([NP, NPL, NPU], np.int32).tofile(f)
        for k in ("INU", "INL", "IAU", "IAL"):
            np.asarray(g[k], np.int32).tofile(f)
        for k in ("D", "AU", "AL", "b"):
            np.asarray(g[k], np.complex128).tofile(f)
    r = subprocess.run([exe], capture_output=True, text=True, cwd=str(tmp_path), timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    it, err, resid = r.stdout.split()
    x = np.fromfile(os.path.join(str(tmp_path), "x.bin"), np.complex128)
    it_ref = int(g["iter"][0])
    print("fortran drop-in: iter", it, "error", err, "resid", resid, "reference iter", it_ref)
    assert int(err) == 0 and float(resid) <= 1e-8
    assert abs(int(it) - it_ref) <= max(3, 0.08 * it_ref)
    assert np.linalg.norm(x - g["x"]) <= 1e-5 * np.linalg.norm(g["x"])
