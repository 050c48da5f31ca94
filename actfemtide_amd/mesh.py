"""Synthetic tidal edge-FEM cases: a box of hexahedral cells, each split into 6
tetrahedra (Kuhn split around the shared body diagonal), with the node and element
ordering the reference's mesh pipeline produces.

Conventions followed from the reference:
  * coordinates in km, x east, y north, z up (doc/Manual_ActFEMtide.md:50);
  * the box is centred at x = y = 0, because GENBCCSEM tests +-xout, +-yout
    (src/solver_mpi/forward_bxyz_mpi.f90:483-507);
  * tetrahedral node order is right handed: gradnodebasisfun stops when the signed
    volume 1/6((x4-x3)x(x1-x3)).(x2-x3) is not positive (src/solver/m_fem_util.f90:22-23);
  * ocean-layer nodes come first, so that the first `nodes` lines of the vxyz/fxyz files
    are the ocean nodes (src/src_tide/m_readfvxyz.f90:24-27,51-57; tmtgem/mesh/combine3d.f90:96-122);
  * elements: sea (group 2), then air (group 1), then crust (group >= 3) last, because the
    crust conductivity is indexed sigma(iele - nphys1) (forward_bxyz_mpi.f90:196-208).
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

from . import lines as _lines

# Tidal constituent periods in seconds.
PERIODS = {"M2": 44714.0, "S2": 43200.0, "N2": 45570.0, "K1": 86164.0, "O1": 92950.0}

# Sea conductivity the reference's GENMAT hard-codes: `sigma=3.33` is a default-real
# (single precision) literal stored into a real(8) (forward_bxyz_mpi.f90:198).
SIGMA_SEA_REFERENCE = float(np.float32(3.33))
SIGMA_AIR_DEFAULT = 1.0e-8

GROUP_AIR = 1
GROUP_SEA = 2
GROUP_CRUST = 3


@dataclasses.dataclass
class TideCase:
    """One synthetic forward problem in the reference's data model.

    Integer arrays are 1-based where the reference is (n4, line, n6line)."""

    nx: int
    ny: int
    nz: int
    h: float
    xyz: np.ndarray          # (node, 3) float64 [km]
    n4: np.ndarray           # (ntet, 4) int32, 1-based node ids
    group: np.ndarray        # (ntet,) int32, n4flag(:,2): 1 air, 2 sea, >=3 crust
    nodes_ocean: int         # number of leading ocean nodes (vxyz/fxyz lines)
    line: np.ndarray         # (nline, 2) int32, 1-based node ids, line(1,i) < line(2,i)
    n6line: np.ndarray       # (ntet, 6) int32, signed 1-based edge ids
    sigma_crust: np.ndarray  # (nphys2,) float64 [S/m], one per crust tet, in element order
    vxyz: np.ndarray         # (node, 3) complex128 [m/s]; zero beyond nodes_ocean, vz = 0
    fxyz: np.ndarray         # (node, 3) float64 [nT]; zero beyond nodes_ocean
    sigma_air: float = SIGMA_AIR_DEFAULT
    sigma_sea: float = SIGMA_SEA_REFERENCE
    cell_of_tet: np.ndarray | None = None   # (ntet,) int32 cell id (k*ny + j)*nx + i
    log10_sigma_cell: np.ndarray | None = None  # (nz, ny, nx) crust log10 sigma (nan elsewhere)
    rho_crust: np.ndarray | None = None  # (nphys2,) resistivity written to cond.msh; sigma = 1/rho

    @property
    def node(self) -> int:
        return self.xyz.shape[0]

    @property
    def ntet(self) -> int:
        return self.n4.shape[0]

    @property
    def nline(self) -> int:
        return self.line.shape[0]

    @property
    def nphys2(self) -> int:
        return int(np.count_nonzero(self.group >= GROUP_CRUST))

    @property
    def nphys1(self) -> int:
        return self.ntet - self.nphys2

    @property
    def bounds(self):
        """(xout, yout, zmin, zmax): the values forward_bxyz passes to GENBCCSEM
        (forward_bxyz_mpi.f90:34-37)."""
        return (self.nx * self.h / 2.0, self.ny * self.h / 2.0,
                float(self.xyz[:, 2].min()), float(self.xyz[:, 2].max()))

    def sigma_tet(self) -> np.ndarray:
        """Per-element conductivity exactly as GENMAT selects it for condflag = 1
        (forward_bxyz_mpi.f90:196-208): air -> sigma_air, sea -> sigma_sea,
        crust -> sigma(iele - nphys1)."""
        s = np.empty(self.ntet, dtype=np.float64)
        s[self.group == GROUP_AIR] = self.sigma_air
        s[self.group == GROUP_SEA] = self.sigma_sea
        crust = self.group >= GROUP_CRUST
        idx = np.nonzero(crust)[0]
        if idx.size:
            if idx[0] != self.nphys1 or idx[-1] != self.ntet - 1:
                raise ValueError("crust elements must be the last nphys2 elements")
            s[crust] = self.sigma_crust
        return s


def _kuhn_tets():
    """Corner offsets of the 6 tetrahedra of a unit cell, each tet a path
    (0,0,0) -> e_a -> e_a + e_b -> (1,1,1) for a permutation (a, b, c) of the axes,
    with the node order fixed so that the reference's signed volume is positive."""
    e = np.eye(3, dtype=np.int64)
    tets = []
    for perm in ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)):
        a, b, _ = perm
        p = [np.zeros(3, np.int64), e[a], e[a] + e[b], np.ones(3, np.int64)]
        x = np.array(p, dtype=np.float64)
        v = np.dot(np.cross(x[3] - x[2], x[0] - x[2]), x[1] - x[2]) / 6.0
        if v < 0:
            p[1], p[2] = p[2], p[1]
        tets.append(p)
    return np.array(tets)  # (6, 4, 3)


def box_mesh(nx: int, ny: int, nz: int, h: float = 2.5, n_air: int = 3, n_sea: int = 1):
    """Nodes, tets and groups of an nx x ny x nz box; returns
    (xyz, n4 (1-based), group, nodes_ocean, cell_of_tet)."""
    if nz < n_air + n_sea + 1:
        raise ValueError("need at least one crust layer")
    n_crust = nz - n_air - n_sea
    # node grid index (i, j, k), k from the bottom; sea surface at z = 0
    ii, jj, kk = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    ii = ii.transpose(2, 1, 0).ravel()  # (k, j, i) lexicographic, i fastest
    jj = jj.transpose(2, 1, 0).ravel()
    kk = kk.transpose(2, 1, 0).ravel()
    x = -nx * h / 2.0 + ii * h
    y = -ny * h / 2.0 + jj * h
    z = (kk - (n_crust + n_sea)) * h
    # ocean nodes: the node planes bounding the sea layers
    k_sea_lo, k_sea_hi = n_crust, n_crust + n_sea
    is_ocean = (kk >= k_sea_lo) & (kk <= k_sea_hi)
    order = np.concatenate([np.nonzero(is_ocean)[0], np.nonzero(~is_ocean)[0]])
    new_id = np.empty(order.size, np.int64)
    new_id[order] = np.arange(order.size)
    xyz = np.stack([x[order], y[order], z[order]], axis=1).astype(np.float64)
    nodes_ocean = int(is_ocean.sum())

    def nid(i, j, k):
        return new_id[(k * (ny + 1) + j) * (nx + 1) + i]

    kt = _kuhn_tets()  # (6,4,3)
    ci, cj, ck = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    ci = ci.transpose(2, 1, 0).ravel()
    cj = cj.transpose(2, 1, 0).ravel()
    ck = ck.transpose(2, 1, 0).ravel()
    ncell = ci.size
    n4 = np.empty((ncell, 6, 4), np.int64)
    for t in range(6):
        for a in range(4):
            off = kt[t, a]
            n4[:, t, a] = nid(ci + off[0], cj + off[1], ck + off[2])
    cell_group = np.where(ck >= n_crust + n_sea, GROUP_AIR,
                          np.where(ck >= n_crust, GROUP_SEA, GROUP_CRUST))
    cell_id = np.arange(ncell)
    cells = np.concatenate([cell_id[cell_group == GROUP_SEA], cell_id[cell_group == GROUP_AIR],
                            cell_id[cell_group == GROUP_CRUST]])
    n4 = n4[cells].reshape(-1, 4) + 1
    group = np.repeat(cell_group[cells], 6).astype(np.int32)
    cell_of_tet = np.repeat(cells, 6).astype(np.int32)
    return xyz, n4.astype(np.int32), group, nodes_ocean, cell_of_tet


def fourier_velocity(xyz: np.ndarray, seed: int, nmodes: int = 4, amp: float = 0.05,
                     lx: float = 1.0, ly: float = 1.0) -> np.ndarray:
    """Horizontal tidal velocity [m/s] at the given points: a seeded sum of `nmodes`
    random horizontal Fourier modes of amplitude `amp`, each with a random direction,
    wave vector and complex phase; v_z = 0 (m_readfvxyz.f90:55 forces it)."""
    rng = np.random.default_rng(seed)
    v = np.zeros((xyz.shape[0], 3), np.complex128)
    for _ in range(nmodes):
        p, q = rng.integers(1, 4, size=2)
        theta = rng.uniform(0.0, 2.0 * math.pi)
        phi = rng.uniform(0.0, 2.0 * math.pi)
        psi = rng.uniform(0.0, 2.0 * math.pi)
        kx, ky = 2.0 * math.pi * p / lx, 2.0 * math.pi * q / ly
        shape = np.cos(kx * xyz[:, 0] + ky * xyz[:, 1] + psi)
        c = amp * complex(math.cos(phi), math.sin(phi))
        v[:, 0] += c * math.cos(theta) * shape
        v[:, 1] += c * math.sin(theta) * shape
    return v


def tide_case(nx: int, ny: int, nz: int, seed: int, h: float = 2.5, *, spheres: int = 0,
              sigma_sea: float = SIGMA_SEA_REFERENCE, sigma_air: float = SIGMA_AIR_DEFAULT,
              crust_mean: float = -2.0, crust_std: float = 0.5,
              log10_sigma_field: np.ndarray | None = None,
              F=(0.0, 3.0e4, -3.5e4), like: TideCase | None = None) -> TideCase:
    """Build a complete synthetic case.

    Crust log10(sigma) is drawn per cell from N(crust_mean, crust_std) with
    numpy.random.default_rng(seed) (all 6 tets of a cell share it), unless a full
    (nz, ny, nx) `log10_sigma_field` is given. `spheres` adds that many 1 S/m balls of
    radius 3 cells with seeded centres inside the crust. `like`: a case of the same box whose
    mesh arrays (nodes, tets, groups, edges) are shared instead of built again."""
    if like is not None:
        if (like.nx, like.ny, like.nz, like.h) != (nx, ny, nz, h):
            raise ValueError("`like` is a case of another box")
        xyz, n4, group, nodes_ocean, cell_of_tet = like.xyz, like.n4, like.group, like.nodes_ocean, like.cell_of_tet
        line, n6line = like.line, like.n6line
    else:
        xyz, n4, group, nodes_ocean, cell_of_tet = box_mesh(nx, ny, nz, h)
        line, n6line = _lines.mkline_mkn6(xyz.shape[0], n4)
    rng = np.random.default_rng(seed)
    n_crust = nz - 4
    if log10_sigma_field is None:
        field = np.full((nz, ny, nx), np.nan)
        field[:n_crust] = rng.normal(crust_mean, crust_std, size=(n_crust, ny, nx))
    else:
        field = np.array(log10_sigma_field, dtype=np.float64, copy=True)
        field[n_crust:] = np.nan
    if spheres:
        kc, jc, ic = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
        r = 3.0
        for _ in range(spheres):
            c = np.array([rng.uniform(r, nx - r), rng.uniform(r, ny - r),
                          rng.uniform(r, max(r + 1e-3, n_crust - r))])
            inside = ((ic + 0.5 - c[0]) ** 2 + (jc + 0.5 - c[1]) ** 2 + (kc + 0.5 - c[2]) ** 2) <= r * r
            inside &= kc < n_crust
            field[inside] = 0.0
    crust = group >= GROUP_CRUST
    cell = cell_of_tet[crust]
    kk = cell // (nx * ny)
    jj = (cell // nx) % ny
    ii = cell % nx
    # resistivity is what cond.msh stores; the reference reads sigma = 1/rho (m_param.f90:476-479)
    rho_crust = 10.0 ** (-field[kk, jj, ii])
    sigma_crust = 1.0 / rho_crust
    vel = np.zeros((xyz.shape[0], 3), np.complex128)
    fxyz = np.zeros((xyz.shape[0], 3), np.float64)
    lx, ly = nx * h, ny * h
    vel[:nodes_ocean] = fourier_velocity(xyz[:nodes_ocean], seed + 7919, lx=lx, ly=ly)
    fxyz[:nodes_ocean] = np.asarray(F, np.float64)
    return TideCase(nx=nx, ny=ny, nz=nz, h=h, xyz=xyz, n4=n4, group=group,
                    nodes_ocean=nodes_ocean, line=line, n6line=n6line,
                    sigma_crust=sigma_crust.astype(np.float64), rho_crust=rho_crust, vxyz=vel, fxyz=fxyz,
                    sigma_air=sigma_air, sigma_sea=sigma_sea, cell_of_tet=cell_of_tet,
                    log10_sigma_cell=field)


def refine_cells(case: TideCase, cells) -> TideCase:
    """The case with each of the given crust cells split into 12 tets: a new node at the cell's
    centre joined to the 12 triangles that cover the cell's faces (the faces its 6 tets do not
    share with each other), so the neighbours' face triangulation is kept and the mesh stays
    conforming. The new nodes take the highest ids and the new tets go last, keeping the crust
    tets last; each new tet keeps its cell's conductivity. Refining the 8 cells around a grid
    node gives that node 22 edges, and the corner nodes' runs of edges towards higher-numbered
    nodes grow past 8, as on the reference's unstructured meshes."""
    cells = np.unique(np.asarray(cells, np.int64))
    crust = case.group >= GROUP_CRUST
    cot = case.cell_of_tet.astype(np.int64)
    sel = np.isin(cot, cells)
    if not np.all(crust[sel]) or np.count_nonzero(sel) != 6 * cells.size:
        raise ValueError("refine_cells: crust cells of a mesh with 6 tets per cell")
    sig_tet = case.sigma_tet()
    xyz = [case.xyz]
    new_n4, new_sig, new_cell = [], [], []
    nid = case.node
    for c in cells:
        tets = case.n4[cot == c] - 1
        faces = {}
        for t in tets:
            for drop in range(4):
                f = tuple(sorted(np.delete(t, drop)))
                faces[f] = faces.get(f, 0) + 1
        outer = [f for f, k in faces.items() if k == 1]
        corners = np.unique(tets)
        if len(outer) != 12 or corners.size != 8:
            raise ValueError("refine_cells: cell is not a 6-tet cube")
        xyz.append(case.xyz[corners].mean(axis=0, keepdims=True))
        s = sig_tet[np.nonzero(cot == c)[0][0]]
        for f in outer:
            q = [f[0], f[1], f[2], nid]
            x = np.vstack([case.xyz[list(f)], xyz[-1]])
            v = np.dot(np.cross(x[3] - x[2], x[0] - x[2]), x[1] - x[2])
            if v < 0:  # the reference's signed volume must be positive
                q[0], q[1] = q[1], q[0]
            new_n4.append(q)
            new_sig.append(s)
            new_cell.append(c)
        nid += 1
    keep = ~sel
    n4 = np.vstack([case.n4[keep], np.asarray(new_n4, np.int64) + 1]).astype(np.int32)
    group = np.concatenate([case.group[keep], np.full(len(new_n4), GROUP_CRUST, np.int32)]).astype(np.int32)
    sigma_crust = np.concatenate([sig_tet[keep & crust], np.asarray(new_sig)])
    cell_of_tet = np.concatenate([case.cell_of_tet[keep], np.asarray(new_cell)]).astype(np.int32)
    xyz = np.vstack(xyz)
    add = xyz.shape[0] - case.node
    vxyz = np.vstack([case.vxyz, np.zeros((add, 3), np.complex128)])
    fxyz = np.vstack([case.fxyz, np.zeros((add, 3), np.float64)])
    line, n6line = _lines.mkline_mkn6(xyz.shape[0], n4)
    return dataclasses.replace(case, xyz=xyz, n4=n4, group=group, line=line, n6line=n6line,
                               sigma_crust=sigma_crust, rho_crust=1.0 / sigma_crust, vxyz=vxyz, fxyz=fxyz,
                               cell_of_tet=cell_of_tet)


def gaussian_random_field(nz: int, ny: int, nx: int, seed: int, mean: float = -2.0,
                          std: float = 0.7, corr_cells: float = 4.0) -> np.ndarray:
    """Stationary Gaussian random field on the cell grid (spectral synthesis with a
    Gaussian covariance of correlation length `corr_cells`), rescaled to the given
    mean and standard deviation."""
    rng = np.random.default_rng(seed)
    w = rng.standard_normal((nz, ny, nx))
    kz = np.fft.fftfreq(nz)[:, None, None]
    ky = np.fft.fftfreq(ny)[None, :, None]
    kx = np.fft.fftfreq(nx)[None, None, :]
    spec = np.exp(-0.5 * (2 * math.pi * corr_cells) ** 2 * (kx ** 2 + ky ** 2 + kz ** 2) / 2.0)
    f = np.real(np.fft.ifftn(np.fft.fftn(w) * spec))
    f = (f - f.mean()) / (f.std() + 1e-300)
    return mean + std * f
