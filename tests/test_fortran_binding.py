"""The Fortran iso_c_binding module (include/afx_mod.f90) compiles with flang, links against
libafx.so, and -- on the GPU -- runs the COCG13_single drop-in from a Fortran program on the
reference's own assembled system (the tiny golden fixture).

The Fortran programs run as child processes. A child started after this test process has used
the GPU itself was seen to find no device (its first HIP call fails), so this file's GPU tests
run before any other GPU test (pytest's default order, by file name, does that)."""
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

PROG_IXYH = """
program t
  use afx
  implicit none
  integer(4), parameter :: nobs = 3, node = 4
  real(8) :: xyzobs(3,nobs), fxyz(3,node), val(6), h(nobs)
  complex(8) :: exy(2,nobs), vxyz(3,node), out(2,4,nobs)
  integer(4) :: stack(0:nobs), item(6), i, rc
  real(4) :: sig4
  sig4 = 3.3
  do i = 1, nobs
    xyzobs(:,i) = (/ 1.d0*i, -2.d0*i, -0.5d0*i /)
    exy(:,i) = (/ (1.d0, 2.d0)*i, (-3.d0, 0.5d0) /)
  end do
  xyzobs(3,3) = 0.25d0
  do i = 1, node
    vxyz(:,i) = (/ (0.2d0, -0.1d0), (0.05d0, 0.3d0), (0.d0, 0.d0) /)
    fxyz(:,i) = (/ 20000.d0, -5000.d0, 40000.d0 /)
  end do
  stack = (/ 0, 2, 4, 6 /)
  item = (/ 1, 2, 2, 3, 3, 4 /)
  val = 0.5d0
  rc = afx_ixyh(nobs, xyzobs, exy, stack, item, val, node, vxyz, fxyz, dble(sig4), out, h)
  open(newunit=i, file='ixyh.bin', access='stream', form='unformatted', status='replace')
  write(i) out, h
  close(i)
  print '(i0)', rc
end program
"""


PROG_PARDISO = """
program t
  use afx
  implicit none
  integer(4) :: NP, NPL, NPU, node, ierr, u, i
  complex(8), allocatable :: D(:), AU(:), AL(:), rhs(:,:), xout(:,:)
  integer(4), allocatable :: INU(:), INL(:), IAU(:), IAL(:), line(:,:), ibc(:)
  logical, allocatable :: line_bc(:)
  open(newunit=u, file='sys.bin', access='stream', form='unformatted', status='old')
  read(u) NP, NPL, NPU, node
  allocate(D(NP), AU(NPU), AL(NPL), rhs(NP,2), xout(NP,2), INU(0:NP), INL(0:NP), IAU(NPU), IAL(NPL))
  allocate(line(2,NP), ibc(NP), line_bc(NP))
  read(u) INU, INL, IAU, IAL, line, ibc, D, AU, AL, rhs
  close(u)
  do i = 1, NP
    line_bc(i) = ibc(i) /= 0
  end do
  call afx_solvepardiso_f(NP, 2, D, AU, AL, INU, IAU, INL, IAL, node, line, line_bc, rhs, xout, ierr)
  open(newunit=u, file='x.bin', access='stream', form='unformatted', status='replace')
  write(u) xout
  close(u)
  print '(i0)', ierr
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
    assert int(err) == 0, r.stdout + r.stderr[-3000:]
    x = np.fromfile(os.path.join(str(tmp_path), "x.bin"), np.complex128)
    it_ref = int(g["iter"][0])
    print("fortran drop-in: iter", it, "error", err, "resid", resid, "reference iter", it_ref)
    assert int(err) == 0 and float(resid) <= 1e-8
    assert abs(int(it) - it_ref) <= max(3, 0.08 * it_ref)
    assert np.linalg.norm(x - g["x"]) <= 1e-5 * np.linalg.norm(g["x"])


@pytest.mark.gpu
def test_fortran_ixyh_reference_shapes(tmp_path):
    """afx_ixyh called from Fortran with IXYHOUT's own array shapes: uniform v and F make
    ixyh_vF = sigma h (v x F) 1e-3 and vh = h v; the third point is on land (h = 0)."""
    exe = _build(str(tmp_path), PROG_IXYH)
    r = subprocess.run([exe], capture_output=True, text=True, cwd=str(tmp_path), timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert int(r.stdout.split()[0]) == 0
    raw = np.fromfile(os.path.join(str(tmp_path), "ixyh.bin"), np.uint8)
    out = raw[:3 * 8 * 16].view(np.complex128).reshape(3, 4, 2)
    h = raw[3 * 8 * 16:].view(np.float64)
    sig = float(np.float32(3.3))
    v = np.array([0.2 - 0.1j, 0.05 + 0.3j, 0.0])
    F = np.array([20000.0, -5000.0, 40000.0])
    vf = np.cross(v, F)[:2] * 1e-3
    assert np.array_equal(h, [0.5, 1.0, 0.0])
    for o in range(3):
        e = np.array([(1 + 2j) * (o + 1), -3 + 0.5j])
        assert np.allclose(out[o, 0], sig * h[o] * e, rtol=1e-15, atol=0)
        assert np.allclose(out[o, 1], sig * h[o] * vf, rtol=1e-14, atol=0)
        assert np.allclose(out[o, 2], out[o, 0] + out[o, 1], rtol=1e-15, atol=0)
        assert np.allclose(out[o, 3], h[o] * v[:2], rtol=1e-15, atol=0)


@pytest.mark.gpu
def test_fortran_solvepardiso_two_right_hand_sides(tmp_path):
    """solvePARDISO's replacement called from Fortran with its own argument shapes (rhs(doftot,2),
    xout(doftot,2), g_line%line(2,nline), GENBCCSEM's logical mask) on the reference's own
    assembled system: both columns at a FP64 true residual <= 1e-11 (a direct-solve-class
    residual), and the fields at the fixture's sites within 1e-9 of a direct solve (x itself
    differs from it in the air's weakly fixed gradient components)."""
    from actfemtide_amd import mesh
    from oracle import oracle as O
    from tests.test_golden import load
    g = load("tiny_ssor")
    c = mesh.tide_case(4, 4, 6, seed=11)
    exe = _build(str(tmp_path), PROG_PARDISO)
    NP, NPU, NPL = len(g["D"]), len(g["AU"]), len(g["IAL"])
    b = np.asarray(g["b"], np.complex128)
    rhs = np.stack([b, 2j * b], axis=1)
    with open(os.path.join(str(tmp_path), "sys.bin"), "wb") as f:
        np.array([NP, NPL, NPU, c.node], np.int32).tofile(f)
        for k in ("INU", "INL", "IAU", "IAL"):
            np.asarray(g[k], np.int32).tofile(f)
        np.asarray(g["mk_line"], np.int32).tofile(f)      # line(2, nline): MKLINE_V2's edges
        np.asarray(g["line_bc"], np.int32).tofile(f)
        for k in ("D", "AU", "AL"):
            np.asarray(g[k], np.complex128).tofile(f)
        np.asfortranarray(rhs).T.tofile(f)                # column-major (doftot, 2)
    r = subprocess.run([exe], capture_output=True, text=True, cwd=str(tmp_path), timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr[-3000:]
    assert int(r.stdout.split()[0]) == 0, r.stdout + r.stderr[-3000:]
    x = np.fromfile(os.path.join(str(tmp_path), "x.bin"), np.complex128).reshape(2, NP).T
    A = O.to_scipy(g["INU"], g["IAU"], g["INL"], g["IAL"], g["D"], g["AU"], g["AL"])
    for k in range(2):
        tr = np.linalg.norm(rhs[:, k] - A @ x[:, k]) / np.linalg.norm(rhs[:, k])
        print(f"fortran solvePARDISO column {k}: true residual {tr:.2e}")
        assert tr <= 1e-11
    import scipy.sparse.linalg as spla
    item, ve, vb, _ = O.prepobscoeff(c, g["sites"])
    for k in range(2):
        xd = spla.spsolve(A.tocsc(), rhs[:, k])
        for val in list(vb) + list(ve):
            f, fd = O.obs_apply(item, val, x[:, k]), O.obs_apply(item, val, xd)
            assert np.abs(f - fd).max() <= 1e-9 * np.abs(fd).max()
