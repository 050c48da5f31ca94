"""Readers and writers of the reference's on-disk formats.

  em3d.msh     gmsh 2.2 ASCII, tetrahedra as `id 4 2 tag1 group n1 n2 n3 n4`; read by readmsh2
               (src/common/m_mesh_type.f90:379-468), written by OUTEM3D (tmtgem/mesh/combine3d.f90:96-122)
  lineinfo.dat nline, (2I10) line pairs, ntet, (6I10) n6line rows; READLINE
               (src/common/m_line_type.f90:28-48), written by tmtgem/mesh/mkline.f90:25-30
  cond.msh     11 header lines, nphys2, `element-id rho` pairs; READCOND (src/common/m_param.f90:449-485),
               written by OUTCOND (src/src_post2/change_model2cond.f90:244-296)
  vxyz_mesh    `lat lon alt (ur,ui) (vr,vi) idx` per ocean node; readvxyz (src/src_tide/m_readfvxyz.f90:34-62)
  fxyz_mesh    `lat lon alt Fe Fn Fu F` per ocean node [nT]; readfxyz (m_readfvxyz.f90:7-30)
  *.ctl        20-character label column, `##` comment lines (readcontrolfile, m_param.f90:939-999)
"""
from __future__ import annotations

import math
import re

import numpy as np

KM_PER_DEG = 111.19492664455873  # 6371 km * pi / 180


# ---------------------------------------------------------------------------------------- msh
def write_em3d_msh(path, xyz, n4, group, n4flag1=None, points=None):
    """Write the 3-D mesh. `points` (optional) are point elements `id 15 2 0 1 node`."""
    ntet = n4.shape[0]
    t1 = np.zeros(ntet, np.int64) if n4flag1 is None else np.asarray(n4flag1)
    pts = [] if points is None else list(points)
    with open(path, "w") as f:
        f.write("$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\n")
        f.write(f"{xyz.shape[0]:10d}\n")
        for i, (x, y, z) in enumerate(xyz, 1):
            f.write(f"{i:10d} {float(x)!r} {float(y)!r} {float(z)!r}\n")
        f.write("$EndNodes\n$Elements\n")
        f.write(f"{len(pts) + ntet:10d}\n")
        eid = 0
        for p in pts:
            eid += 1
            f.write(f"{eid:10d}  15   2   0   1{int(p):10d}\n")
        for t in range(ntet):
            eid += 1
            a, b, c, d = n4[t]
            f.write(f"{eid:10d}  4   2{int(t1[t]):4d}{int(group[t]):10d}{a:10d}{b:10d}{c:10d}{d:10d}\n")
        f.write("$EndElements\n")


_NNODES = {15: 1, 1: 2, 2: 3, 4: 4}


def read_em3d_msh(path):
    """Return (xyz (node,3), n4 (ntet,4) 1-based, n4flag (ntet,2)) as readmsh2 fills them."""
    with open(path) as f:
        lines = f.read().split("\n")
    p = 4
    node = int(lines[p].split()[0])
    p += 1
    xyz = np.array([[float(v) for v in lines[p + i].split()[1:4]] for i in range(node)])
    p += node + 2
    nele = int(lines[p].split()[0])
    p += 1
    n4, flag = [], []
    for i in range(nele):
        tok = lines[p + i].split()
        itype, ntag = int(tok[1]), int(tok[2])
        tags = [int(v) for v in tok[3:3 + ntag]]
        nodes = [int(v) for v in tok[3 + ntag:3 + ntag + _NNODES[itype]]]
        if itype == 4:
            n4.append(nodes)
            flag.append(tags[:2] if len(tags) >= 2 else tags + [0])
    return xyz, np.array(n4, np.int32).reshape(-1, 4), np.array(flag, np.int32).reshape(-1, 2)


# ----------------------------------------------------------------------------------- lineinfo
def write_lineinfo(path, line, n6line):
    with open(path, "w") as f:
        f.write(f"{line.shape[0]:12d}\n")
        for a, b in line:
            f.write(f"{a:10d}{b:10d}\n")
        f.write(f"{n6line.shape[0]:12d}\n")
        for r in n6line:
            f.write("".join(f"{v:10d}" for v in r) + "\n")


def read_lineinfo(path):
    """READLINE: fixed I10 columns for the line pairs and n6line rows."""
    with open(path) as f:
        lines = f.read().split("\n")
    nline = int(lines[0].split()[0])
    line = np.array([[int(l[0:10]), int(l[10:20])] for l in lines[1:1 + nline]], np.int32)
    ntet = int(lines[1 + nline].split()[0])
    rows = lines[2 + nline:2 + nline + ntet]
    n6 = np.array([[int(l[10 * j:10 * j + 10]) for j in range(6)] for l in rows], np.int32)
    return line, n6


# --------------------------------------------------------------------------------------- cond
def write_cond(path, rho, nphys1, shift=0):
    """OUTCOND layout: element ids are shift + nphys1 + i (the crust tets are the last ones)."""
    with open(path, "w") as f:
        f.write('$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$ElementData\n1\n"A rho model view"\n1\n0.0\n3\n0\n1\n')
        f.write(f"{len(rho):10d}\n")
        for i, r in enumerate(rho, 1):
            f.write(f"{shift + nphys1 + i:10d} {float(r)!r}\n")
        f.write("$EndElementData\n")


def read_cond(path):
    """READCOND: skip 11 lines, nphys2, then `index rho`; sigma = 1/rho (-9999 when |rho| <= 1e-10).

    The index column is read and ignored, as the reference does (m_param.f90:467-479)."""
    with open(path) as f:
        lines = f.read().split("\n")
    nphys2 = int(lines[11].split()[0])
    vals = [lines[12 + i].split() for i in range(nphys2)]
    index = np.array([int(v[0]) for v in vals], np.int64)
    rho = np.array([float(v[1].replace("D", "E").replace("d", "e")) for v in vals])
    sigma = np.where(np.abs(rho) > 1e-10, 1.0 / np.where(rho == 0, 1.0, rho), float(np.float32(-9999.0)))
    return index, rho, sigma


# -------------------------------------------------------------------------------- vxyz / fxyz
def _latlon(xyz):
    lat = xyz[:, 1] / KM_PER_DEG
    lon = xyz[:, 0] / KM_PER_DEG
    return lat, lon


def write_vxyz(path, xyz, v):
    lat, lon = _latlon(xyz)
    with open(path, "w") as f:
        for i in range(xyz.shape[0]):
            f.write(f"  {float(lat[i])!r}  {float(lon[i])!r}  {float(xyz[i, 2])!r}   ({float(v[i, 0].real)!r},{float(v[i, 0].imag)!r})"
                    f"   ({float(v[i, 1].real)!r},{float(v[i, 1].imag)!r})  {i + 1:10d}\n")


def write_fxyz(path, xyz, F):
    lat, lon = _latlon(xyz)
    with open(path, "w") as f:
        for i in range(xyz.shape[0]):
            tot = math.sqrt(float(np.dot(F[i], F[i])))
            f.write(f"  {float(lat[i])!r}  {float(lon[i])!r}  {float(xyz[i, 2])!r}  {float(F[i, 0])!r}  {float(F[i, 1])!r}  {float(F[i, 2])!r}  {tot!r}\n")


_TOK = re.compile(r"\([^)]*\)|[^\s,]+")


def _num(s):
    return float(s.replace("D", "E").replace("d", "e"))


def read_vxyz(path, nodes, node):
    """readvxyz: the first `nodes` records fill vxyz(1:2, i); vxyz(3, :) = 0; rows past nodes are 0."""
    v = np.zeros((node, 3), np.complex128)
    with open(path) as f:
        for i in range(nodes):
            tok = _TOK.findall(f.readline())
            for c in range(2):
                re_, im_ = tok[3 + c].strip("()").split(",")
                v[i, c] = complex(_num(re_), _num(im_))
    return v


def read_fxyz(path, nodes, node):
    """readfxyz: columns 4-6 of the first `nodes` records."""
    F = np.zeros((node, 3))
    with open(path) as f:
        for i in range(nodes):
            tok = _TOK.findall(f.readline())
            F[i] = [_num(t) for t in tok[3:6]]
    return F


# ---------------------------------------------------------------------------------------- ctl
def read_ctl_lines(path):
    """readcontrolfile + subtractcommentout: drop `##` lines, cut the 20-character label column."""
    out = []
    with open(path) as f:
        for l in f.read().split("\n"):
            if l[:2] == "##":
                continue
            out.append(l[20:].rstrip())
    return out


def write_forward_ctl(path, meshfile, linefile, oceanmesh, vxyzfile, fxyzfile, condfile, freqs, xout, yout,
                      zmin, zmax, sites, sigma_air=1e-8, lenout=0.0):
    """A tide_fwd.ctl in the reference's sequential layout (Kikai/fwd/tide_fwd.ctl; read order of
    readparam, m_param.f90:145-446): no topography, xy sites, one dummy wire source of 0 A,
    conductivity from file (condflag 1)."""
    def row(label, value):
        return f"{label:<19.19}!{value}\n"

    w = ["## forward control file\n", row("itopofile 0 or 1", 0), row("mesh file", meshfile),
         row("2d triangle z file", meshfile), row("local line file", linefile), row("Ocean mesh file.", oceanmesh),
         row("mesh ctl file", "mesh.ctl"), row("Vxyzfile", vxyzfile), row("Fxyzfile", fxyzfile),
         row("output folder", "./result/"), row("header2d  (a50)", "tide2d"), row("header3d  (a50)", "tide3d"),
         row("# of frequency", len(freqs))]
    w += [row("Frequency [Hz]", repr(float(fq))) for fq in freqs]
    # xbound(4) = -(xbound(2) - lenout) = xout, so the west bound is lenout - xout
    w += [row("west bound", repr(lenout - xout)), row("east bound", repr(xout)), row("south bound", repr(lenout - yout)),
          row("north bound", repr(yout)), row("lenout [km]", repr(lenout)), row("upz in [km]  (>0)", repr(zmax)),
          row("downz in [km](<0)", repr(zmin)), row("zmax   [km]", repr(zmax)), row("zmin   [km]", repr(zmin))]
    w += [row(lab, "0.0") for lab in ("sizein [km]", "sizebo [km]", "sigma_obs [km]", "A_obs     [km]",
                                      "dlen_source [km]", "sigma_src [km]", "A_src     [km]")]
    w += [row("# of observatory", len(sites)), row("lonlat(1),xyz (2)", 2), row("dummy", ""),
          row("lonlatorigin", "0.0 0.0")]
    for i, s in enumerate(sites, 1):
        w += [row("  Name", f"S{i}"), row("  xyz", f"{float(s[0])!r} {float(s[1])!r} {float(s[2])!r}")]
    w += [row("0:no,1:nxny,2:ntrik", 0), row("# of sources", 1), row("Source Name", "S1"),
          row("source start point", "0.0 0.0 0.0"), row("source end   point", "0.0 0.0 0.0"),
          row("Elcetric current[A]", "0.0"), row("sigma_air    [S/m]", repr(sigma_air)),
          row("condflag 0:home,1:", 1), row("condfile", condfile)]
    with open(path, "w") as f:
        f.writelines(w)


def read_forward_ctl(path):
    """Parse the subset of readparam this path uses (no topography, xy sites, condflag 1)."""
    L = read_ctl_lines(path)
    it = iter(L)
    nxt = lambda: next(it)  # noqa: E731
    out = {}
    if int(nxt().split()[0]) != 0:
        raise ValueError("only itopoflag = 0 is supported")
    out["meshfile"] = nxt().strip()
    out["zmeshfile"] = nxt().strip()
    out["linefile"] = nxt().strip()
    out["oceanmesh"] = nxt().strip()
    out["meshctl"] = nxt().strip()
    out["vxyz"] = nxt().strip()
    out["fxyz"] = nxt().strip()
    for _ in range(3):
        nxt()
    nf = int(nxt().split()[0])
    out["freqs"] = [_num(nxt().split()[0]) for _ in range(nf)]
    xb2, xb3, yb2, yb3 = (_num(nxt().split()[0]) for _ in range(4))
    lenout, upzin, downzin, zmax, zmin = (_num(nxt().split()[0]) for _ in range(5))
    for _ in range(7):
        nxt()
    out["zbound"] = (zmin, downzin, upzin, zmax)
    out["xout"] = -(xb2 - lenout)
    out["yout"] = -(yb2 - lenout)
    nobs = int(nxt().split()[0])
    flag = int(nxt().split()[0])
    if flag != 2:
        raise ValueError("only xyz sites (lonlatflag 2) are supported")
    nxt()
    nxt()
    sites = []
    for _ in range(nobs):
        nxt()
        sites.append([_num(v) for v in nxt().split()[:3]])
    out["sites"] = np.array(sites)
    if int(nxt().split()[0]) != 0:
        raise ValueError("only ixyflag = 0 is supported")
    nsrc = int(nxt().split()[0])
    for _ in range(3 * nsrc):
        nxt()
    nxt()
    out["sigma_air"] = _num(nxt().split()[0])
    out["condflag"] = int(nxt().split()[0])
    out["condfile"] = nxt().strip()
    return out


def write_case(case, directory, freqs=(1.0 / 44714.0,), sites=()):
    """Write a synthetic case in the reference layout; returns the file paths."""
    import os
    os.makedirs(directory, exist_ok=True)
    p = {k: os.path.join(directory, v) for k, v in dict(mesh="em3d.msh", line="lineinfo.dat", cond="cond.msh",
                                                           vxyz="vxyz_mesh", fxyz="fxyz_mesh", ctl="tide_fwd.ctl").items()}
    write_em3d_msh(p["mesh"], case.xyz, case.n4, case.group)
    write_lineinfo(p["line"], case.line, case.n6line)
    write_cond(p["cond"], case.rho_crust if case.rho_crust is not None else 1.0 / case.sigma_crust, case.nphys1)
    nodes = case.nodes_ocean
    write_vxyz(p["vxyz"], case.xyz[:nodes], case.vxyz[:nodes])
    write_fxyz(p["fxyz"], case.xyz[:nodes], case.fxyz[:nodes])
    xout, yout, zmin, zmax = case.bounds
    write_forward_ctl(p["ctl"], "em3d.msh", "lineinfo.dat", "ocean.msh", "vxyz_mesh", "fxyz_mesh", "cond.msh",
                      freqs, xout, yout, zmin, zmax, sites, sigma_air=case.sigma_air)
    return p
