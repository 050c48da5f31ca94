"""Generate the golden vectors of tests/golden/*.npz from the reference's own Fortran.

Runs oracle/_ref/ref_harness (the reference's READMESH_TOTAL, READLINE, READCOND,
set_iccg_var7_dofn, GENMAT, COPY_UL_ICCG12, GENBCCSEM, SET_BC_ICCG, COCG13_single, FINDELEMENT0,
EDGEBASISFUN, FACEBASISFUN, MKLINE_V2, MKN6, readvxyz, readfxyz -- compiled by
oracle/build_ref.sh) on synthetic cases written in the reference's file formats, and stores the
inputs' identifiers and the outputs as .npz fixtures. Run it in the build container, where
/root/reference exists:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from actfemtide_amd import formats, mesh  # noqa: E402

HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
OMEGA_M2 = 2.0 * np.pi / mesh.PERIODS["M2"]

# (name, nx, ny, nz, seed, precond, tol, maxit, nset, n_sites, full arrays?)
CASES = [
    ("tiny_ssor", 4, 4, 6, 11, "SSOR", 1e-8, 20000, 1, 4, True),
    ("tiny_ic", 4, 4, 6, 11, "IC", 1e-8, 300, 1, 0, False),
    ("tiny_diag_nset2", 4, 4, 6, 11, "DIAG", 1e-8, 20000, 2, 0, False),
    ("small_ssor", 8, 8, 6, 1001, "SSOR", 1e-8, 20000, 1, 8, False),
    ("cfg1_ssor", 16, 16, 12, 1001, "SSOR", 1e-8, 20000, 1, 16, False),
    # successive calls on the same arrays: the SAVEd DD / SCALE / IFLAG of COCG13_single.f90:39-62
    # (NSET = 0 reuses DD; NSET = 2 scales D/AU/AL in place and saves SCALE; a later NSET = 1 call
    # scales B by the saved SCALE; NSET = 0 divides B by it on exit)
    ("tiny_nset_seq", 4, 4, 6, 11, "SSOR", 1e-8, 20000, 1, 0, False,
     [(0, "SSOR", 0.5), (2, "SSOR", 1.0), (1, "SSOR", 1.0), (0, "SSOR", 2.0)]),
]


def sites_for(case, n, seed):
    """Seeded seafloor sites 1 m below the sea bottom plane, off every face of the Kuhn split."""
    rng = np.random.default_rng(seed)
    h = case.h
    xout, yout = case.nx * h / 2, case.ny * h / 2
    zb = -h  # sea layer is [-h, 0]
    pts = []
    for _ in range(n):
        i = rng.integers(1, case.nx - 1)
        j = rng.integers(1, case.ny - 1)
        fx, fy = rng.uniform(0.13, 0.87, size=2)
        if abs(fx - fy) < 0.05:
            fy = min(0.87, fy + 0.1)
        pts.append([-xout + (i + fx) * h, -yout + (j + fy) * h, zb - 0.001])
    return np.array(pts)


def run_case(name, nx, ny, nz, seed, precond, tol, maxit, nset, n_sites, full, seq=()):
    case = mesh.tide_case(nx, ny, nz, seed=seed)
    sites = sites_for(case, n_sites, seed + 17)
    with tempfile.TemporaryDirectory() as d:
        formats.write_case(case, d, sites=sites)
        xout, yout, zmin, zmax = case.bounds
        drv = "\n".join([
            "em3d.msh", "lineinfo.dat", "cond.msh", "vxyz_mesh", "fxyz_mesh",
            str(case.nodes_ocean), repr(float(OMEGA_M2)), f"{float(zmin)!r} {float(zmax)!r} {float(xout)!r} {float(yout)!r}",
            repr(float(case.sigma_air)), precond, f"{float(tol)!r} {maxit} {nset} 1.0 0.0 1", str(len(sites)),
        ] + [f"{float(p[0])!r} {float(p[1])!r} {float(p[2])!r}" for p in sites]
            + [str(len(seq))] + [f"{k} {pc} {float(f)!r}" for k, pc, f in seq]) + "\n"
        r = subprocess.run([HARNESS], input=drv, text=True, capture_output=True, cwd=d)
        if r.returncode != 0:
            raise RuntimeError(f"{name}: harness failed\n{r.stdout[-2000:]}\n{r.stderr[-2000:]}")

        def rd(nm, dt):
            return np.fromfile(os.path.join(d, nm + ".bin"), dtype=dt)

        hist = np.loadtxt(os.path.join(d, "fort.12")) if os.path.exists(os.path.join(d, "fort.12")) else np.zeros((0, 2))
        out = dict(nx=nx, ny=ny, nz=nz, seed=seed, precond=precond, tol=tol, maxit=maxit, nset=nset,
                   omega=OMEGA_M2, sites=sites,
                   x=rd("x", np.complex128), resid=rd("resid", np.float64), iter=rd("iter", np.int32),
                   hist_iter=hist[:, 0].astype(np.int32) if hist.size else np.zeros(0, np.int32),
                   hist_resid=hist[:, 1] if hist.size else np.zeros(0))
        if full:
            for nm in ("INU", "INL", "IAU", "IAL", "line_bc", "mk_line", "mk_n6line"):
                out[nm] = rd(nm, np.int32)
            for nm in ("D_genmat", "AU_genmat", "b_genmat", "AL_copy", "D", "AU", "AL", "b", "vxyz"):
                out[nm] = rd(nm, np.complex128)
            out["fxyz"] = rd("fxyz", np.float64)
        else:
            # pattern and edge numbering as checksums (exact integers), values as samples
            for nm in ("INU", "INL", "IAU", "IAL", "line_bc", "mk_line", "mk_n6line"):
                a = rd(nm, np.int32).astype(np.int64)
                out[nm + "_len"] = a.size
                out[nm + "_sum"] = int(a.sum())
                out[nm + "_wsum"] = int((a * (np.arange(a.size) % 9973 + 1)).sum())
            rng = np.random.default_rng(5)
            for nm in ("D", "AU", "AL", "b"):
                a = rd(nm, np.complex128)
                idx = np.sort(rng.choice(a.size, size=min(a.size, 400), replace=False))
                out[nm + "_idx"] = idx
                out[nm + "_val"] = a[idx]
                out[nm + "_norm"] = np.linalg.norm(a)
        if seq:
            out["seq_nset"] = np.array([k for k, _, _ in seq], np.int32)
            out["seq_precond"] = np.array([pc for _, pc, _ in seq])
            out["seq_bfac"] = np.array([f for _, _, f in seq])
            for q in range(1, len(seq) + 1):
                out[f"seq_x{q}"] = rd(f"seq_x{q}", np.complex128)
                out[f"seq_b{q}"] = rd(f"seq_b{q}", np.complex128)
                out[f"seq_resid{q}"] = rd(f"seq_resid{q}", np.float64)
                out[f"seq_iter{q}"] = rd(f"seq_iter{q}", np.int32)
        if n_sites:
            out["obs_iele"] = rd("obs_iele", np.int32)
            out["obs_w"] = rd("obs_w", np.float64).reshape(n_sites, 6, 3)
            out["obs_wf"] = rd("obs_wf", np.float64).reshape(n_sites, 6, 3)
            out["obs_a"] = rd("obs_a", np.float64).reshape(n_sites, 4)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
    print(name, "iter", out["iter"], "resid", out["resid"])


def main():
    if not os.path.exists(HARNESS):
        subprocess.run(["bash", os.path.join(ROOT, "oracle", "build_ref.sh")], check=True)
    only = sys.argv[1:]
    for c in CASES:
        if not only or c[0] in only:
            run_case(*c)


if __name__ == "__main__":
    main()
