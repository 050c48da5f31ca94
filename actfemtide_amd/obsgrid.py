"""Observation grids, the seafloor current and the result files of the tidal forward run.

  PREPOBSCOEFF_XY       an nx x ny grid of points at one depth
                        (src/solver_mpi/n_ebfem_bxyz_mpi.f90:732-820)
  PREPOBSCOEFF_XYSBTM   points 1 m above the centroid of every seafloor triangle, with the
                        node-average rows coeff_vF (n_ebfem_bxyz_mpi.f90:1030-1080)
  IXYHOUT               the seafloor current sigma h (E + v x F) (n_ebfem_bxyz_mpi.f90:192-301),
                        computed on the device (`afx_ixyh`)
  OUTFREQFILES2         per-frequency grid response files (src/ana/m_outresp.f90:181-214)
  OUTOBSFILESFWD        per-site response files (src/ana/m_outresp.f90:216-257)

The grids' field coefficients come from the library's point location and basis functions
(`System.prepobscoeff`, CALCOEFF_LINE n_ebfem_bxyz_mpi.f90:838-932), and the fields from
`afx.calobsebcomp`, exactly as for the observatories.
"""
from __future__ import annotations

import math
import os

import numpy as np

from . import afx

# IXYHOUT's ocean conductivity: the literal 3.3 is a default (single precision) real
# stored into a real(8) (n_ebfem_bxyz_mpi.f90:221).
SIGMA_OCEAN_IXYH = float(np.float32(3.3))


# ------------------------------------------------------------------------------ grids
def xy_grid(xmin, xmax, ymin, ymax, nx, ny, z):
    """PREPOBSCOEFF_XY's points (n_ebfem_bxyz_mpi.f90:775-784): cell centres of an nx x ny grid
    over [xmin, xmax] x [ymin, ymax] at height z [km]; point (j-1)*ny + i for x index j and y
    index i (1-based), so y runs fastest."""
    dx = (xmax - xmin) / float(nx)
    dy = (ymax - ymin) / float(ny)
    pts = np.empty((nx * ny, 3))
    for j in range(nx):
        for i in range(ny):
            pts[j * ny + i] = (j * dx + xmin + dx / 2.0, i * dy + ymin + dy / 2.0, z)
    return pts


def seafloor_mesh(case):
    """The horizontal triangle mesh of a synthetic box case and its node map, in the roles of the
    reference's ki_mesh (polygon_ki.msh) and ki23dptr (READKI23DPTR, n_ebfem_bxyz_mpi.f90:305-330):
    n3 (ntri, 3) 1-based horizontal node ids, counter-clockwise, two triangles per horizontal cell
    split along the same diagonal as the cells' tetrahedra; ki23dptr (2, nodek) 1-based 3-D node
    ids of each horizontal node at the sea surface (row 0) and at the seafloor (row 1)."""
    nx, ny = case.nx, case.ny
    nodek = (nx + 1) * (ny + 1)
    if case.nodes_ocean != 2 * nodek:
        raise ValueError("seafloor_mesh expects one sea layer (two ocean node planes)")
    # the ocean nodes come first: the seafloor plane, then the sea-surface plane, i fastest
    k2 = np.arange(nodek)
    ki23dptr = np.stack([nodek + k2 + 1, k2 + 1]).astype(np.int32)
    n3 = np.empty((2 * nx * ny, 3), np.int32)
    t = 0
    for j in range(ny):
        for i in range(nx):
            a, b = j * (nx + 1) + i + 1, j * (nx + 1) + i + 2
            c, d = (j + 1) * (nx + 1) + i + 2, (j + 1) * (nx + 1) + i + 1
            n3[t] = (a, b, c)
            n3[t + 1] = (a, c, d)
            t += 2
    return n3, ki23dptr


def xysbtm_points(xyz, n3, ki23dptr):
    """PREPOBSCOEFF_XYSBTM (n_ebfem_bxyz_mpi.f90:1063-1075): one point per triangle, 1 m above the
    centroid of its three seafloor nodes, and coeff_vF, the rows averaging those nodes (1/3 each).
    Returns (xyz_obs (ntri, 3), (stack, item 1-based, val))."""
    ntri = n3.shape[0]
    pts = np.zeros((ntri, 3))
    item = np.empty(3 * ntri, np.int32)
    val = np.full(3 * ntri, 1.0 / 3.0)
    up = np.array([0.0, 0.0, 0.001]) / 3.0
    for i in range(ntri):
        x3 = np.zeros(3)
        for j in range(3):
            k = ki23dptr[1, n3[i, j] - 1]
            x3 = x3 + xyz[k - 1] / 3.0 + up
            item[3 * i + j] = k
        pts[i] = x3
    stack = np.arange(0, 3 * ntri + 1, 3, dtype=np.int32)
    return pts, (stack, item, val)


# --------------------------------------------------------------------------- IXYHOUT
def ixyhout(xyz_obs, fields, coeff_vf, vxyz, fxyz, sigma=SIGMA_OCEAN_IXYH):
    """IXYHOUT's arrays (n_ebfem_bxyz_mpi.f90:213-268) for one source: fields (5, nobs) complex
    [Bx, By, Bz, Ex, Ey] at the points (CALOBSEBCOMP). Returns a dict of ixyh_e, ixyh_vF, ixyh and
    vh (each (nobs, 2) complex), bxyz (nobs, 3) complex and h (nobs,)."""
    stack, item, val = coeff_vf
    fields = np.asarray(fields)
    e, v, tot, vh, h = afx.ixyh(xyz_obs, fields[3:5].T, stack, item, val, vxyz, fxyz, sigma)
    return dict(ixyh_e=e, ixyh_vF=v, ixyh=tot, vh=vh, bxyz=fields[0:3].T.copy(), h=h)


# ----------------------------------------------------------------- Fortran number edits
def fortran_e(x, w=15, d=7):
    """The Ew.d edit of a Fortran compiler: [-]0.ddddddd E+ee, right-justified."""
    x = float(x)
    if not math.isfinite(x):
        return ("NaN" if math.isnan(x) else ("-Infinity" if x < 0 else "Infinity")).rjust(w)
    sign = "-" if math.copysign(1.0, x) < 0 else ""
    if x == 0.0:
        digits, ex = "0" * d, 0
    else:
        m = f"{abs(x):.{d - 1}e}"
        digits = m[0] + m[2:2 + d - 1]
        ex = int(m[m.index("e") + 1:]) + 1
    exs = f"E{ex:+03d}" if abs(ex) <= 99 else f"{ex:+04d}"
    return f"{sign}0.{digits}{exs}".rjust(w)


def fortran_g(x, w=15, d=7):
    """The Gw.d edit: F(w-4).(d-s) and 4 blanks when 0.1 <= |x| < 10**d after rounding to d
    significant digits (10**(s-1) <= |x| < 10**s), the Ew.d edit otherwise."""
    x = float(x)
    if math.isfinite(x):
        if x == 0.0:
            return f"{x:.{d - 1}f}".rjust(w - 4) + " " * 4
        m = f"{abs(x):.{d - 1}e}"
        s = int(m[m.index("e") + 1:]) + 1
        if 0 <= s <= d:
            return f"{x:.{d - s}f}".rjust(w - 4) + " " * 4
    return fortran_e(x, w, d)


def fortran_freq_tag(freq):
    """`write(num,'(e9.2)') freq`: the frequency tag of the file names (e.g. ' 0.22E-04')."""
    return fortran_e(freq, 9, 2)


def _amp_phase_pairs(c):
    amp, ph = afx.amp_phase(np.asarray(c))
    out = np.empty(amp.shape + (2,))
    out[..., 0] = amp
    out[..., 1] = ph
    return out


# ---------------------------------------------------------------------------- writers
def write_coord(path, xyz_obs):
    """IXYHOUT's coordinate file (n_ebfem_bxyz_mpi.f90:271-279): nobs, then '(3g15.7)' rows."""
    with open(path, "w") as f:
        f.write(f"{xyz_obs.shape[0]:12d}\n")
        for p in xyz_obs:
            f.write("".join(fortran_g(v) for v in p) + "\n")


def ixyh_path(outputfolder, freq):
    """'ixyh_xy2D' // trim(adjustl(num)) // '.dat' (n_ebfem_bxyz_mpi.f90:282-284)."""
    return os.path.join(outputfolder, "ixyh_xy2D" + fortran_freq_tag(freq).strip() + ".dat")


def write_ixyh(path, xyz_obs, r):
    """IXYHOUT's '(26e15.7)' rows (n_ebfem_bxyz_mpi.f90:285-288): xyz, ixyh_e, ixyh_vF, ixyh, vh
    (complex pairs as (re, im)), bxyz, h."""
    with open(path, "w") as f:
        for o in range(xyz_obs.shape[0]):
            vals = list(xyz_obs[o])
            for key in ("ixyh_e", "ixyh_vF", "ixyh", "vh"):
                for c in r[key][o]:
                    vals += [c.real, c.imag]
            for c in r["bxyz"][o]:
                vals += [c.real, c.imag]
            vals.append(r["h"][o])
            f.write("".join(fortran_e(v) for v in vals) + "\n")


def outfreqfiles2_path(outputfolder, xyfilehead, freq, k):
    """header // xyheader // num(2:9) // '_S' // k // '.dat' (m_outresp.f90:199-201)."""
    return outputfolder + xyfilehead + fortran_freq_tag(freq)[1:9] + f"_S{k}.dat"


def write_outfreqfiles2(outputfolder, xyfilehead, freq, xyz_obs, fields_by_source):
    """OUTFREQFILES2 (m_outresp.f90:181-214): one file per source, '(12e15.7)' rows of x, y and
    (amplitude, phase [deg]) of Bx, By, Bz, Ex, Ey. fields_by_source: (nsr, 5, nobs) complex.
    Returns the paths written."""
    paths = []
    for k, fields in enumerate(np.asarray(fields_by_source), start=1):
        path = outfreqfiles2_path(outputfolder, xyfilehead, freq, k)
        ap = _amp_phase_pairs(fields)  # (5, nobs, 2)
        with open(path, "w") as f:
            for o in range(xyz_obs.shape[0]):
                vals = [xyz_obs[o, 0], xyz_obs[o, 1]] + [ap[j, o, q] for j in range(5) for q in range(2)]
                f.write("".join(fortran_e(v) for v in vals) + "\n")
        paths.append(path)
    return paths


def write_outobsfilesfwd(outputfolder, sites, sources, freqs, resp):
    """OUTOBSFILESFWD (m_outresp.f90:216-257): for every site and source, head // site // '_' //
    source // '.dat' with one '(11g15.7)' row per frequency: freq and (amplitude, phase [deg]) of
    Bx, By, Bz, Ex, Ey. resp: (nfreq, nsr, 5, nobs) complex. Returns the paths written."""
    resp = np.asarray(resp)
    ap = _amp_phase_pairs(resp)  # (nfreq, nsr, 5, nobs, 2)
    paths = []
    for l, site in enumerate(sites):
        for k, sour in enumerate(sources):
            path = outputfolder + site.strip() + "_" + sour.strip() + ".dat"
            with open(path, "w") as f:
                for i, fr in enumerate(freqs):
                    vals = [fr] + [ap[i, k, j, l, q] for j in range(5) for q in range(2)]
                    f.write("".join(fortran_g(v) for v in vals) + "\n")
            paths.append(path)
    return paths
