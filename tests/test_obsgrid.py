"""Observation grids, the IXYHOUT restatement and the result-file writers on the CPU.

What pins them:
- PREPOBSCOEFF_XY's point order and positions follow n_ebfem_bxyz_mpi.f90:775-784 (checked
  against the formula for every point: y runs fastest, cell centres);
- the seafloor mesh covers the box exactly (total area, orientation) and its node map lands on
  the seafloor and sea-surface planes at the right (x, y);
- XYSBTM points are the node centroids plus 1 m (n_ebfem_bxyz_mpi.f90:1066-1073);
- the oracle's IXYHOUT reduces to closed forms: uniform v and F give v x F, land points give 0,
  sigma is the single-precision 3.3;
- the number edits reproduce a Fortran compiler's E15.7 / G15.7 / E9.2 output (fixture printed by
  AMD flang), including the file name the reference's own plot script reads
  (Kikai/fwd/plot_ixyh.sh: ixyh_xy2D0.22E-04.dat);
- the writers' files parse back to the values written (7 significant digits).
"""
import os

import numpy as np
import pytest

from actfemtide_amd import obsgrid as G


def test_xy_grid_order_and_centres():
    p = G.xy_grid(-10.0, 30.0, -5.0, 15.0, 4, 5, -0.5)
    assert p.shape == (20, 3)
    for j in range(4):
        for i in range(5):
            assert p[j * 5 + i, 0] == pytest.approx(-10.0 + 10.0 * j + 5.0)
            assert p[j * 5 + i, 1] == pytest.approx(-5.0 + 4.0 * i + 2.0)
    assert np.all(p[:, 2] == -0.5)


@pytest.fixture(scope="module")
def small_case():
    from actfemtide_amd import mesh
    return mesh.tide_case(6, 5, 6, seed=3)


def test_seafloor_mesh_covers_box(small_case):
    c = small_case
    n3, k23 = G.seafloor_mesh(c)
    assert n3.shape == (2 * c.nx * c.ny, 3)
    assert k23.shape == (2, (c.nx + 1) * (c.ny + 1))
    bottom = c.xyz[k23[1] - 1]
    top = c.xyz[k23[0] - 1]
    zf = np.unique(bottom[:, 2])
    assert zf.size == 1 and np.all(top[:, 2] == 0.0) and zf[0] == -c.h
    assert np.array_equal(bottom[:, :2], top[:, :2])
    xy = bottom[:, :2]
    tri = xy[n3 - 1]
    area = 0.5 * ((tri[:, 1, 0] - tri[:, 0, 0]) * (tri[:, 2, 1] - tri[:, 0, 1])
                  - (tri[:, 2, 0] - tri[:, 0, 0]) * (tri[:, 1, 1] - tri[:, 0, 1]))
    assert np.all(area > 0)
    assert area.sum() == pytest.approx(c.nx * c.ny * c.h * c.h)


def test_xysbtm_points_are_centroids_1m_up(small_case):
    c = small_case
    n3, k23 = G.seafloor_mesh(c)
    pts, (stack, item, val) = G.xysbtm_points(c.xyz, n3, k23)
    nodes = k23[1][n3 - 1]
    ref = c.xyz[nodes - 1].mean(axis=1) + np.array([0, 0, 0.001])
    assert np.allclose(pts, ref, rtol=0, atol=1e-12)
    assert np.array_equal(item.reshape(-1, 3), nodes)
    assert np.array_equal(stack, 3 * np.arange(n3.shape[0] + 1))
    assert np.all(val == 1.0 / 3.0)


def test_ixyh_oracle_closed_forms():
    from oracle import oracle as O
    rng = np.random.default_rng(0)
    node, nobs = 7, 5
    v0 = np.array([0.3 + 0.1j, -0.2 + 0.4j, 0.0])
    F0 = np.array([11000.0, -3000.0, 42000.0])
    vxyz = np.tile(v0, (node, 1))
    fxyz = np.tile(F0, (node, 1))
    stack = np.arange(0, 3 * nobs + 1, 3)
    item = rng.integers(1, node + 1, 3 * nobs).astype(np.int32)
    val = np.full(3 * nobs, 1.0 / 3.0)
    xyz = np.column_stack([rng.standard_normal(nobs), rng.standard_normal(nobs), -rng.uniform(0.1, 4, nobs)])
    xyz[4, 2] = 0.2  # on land: h = 0
    exy = rng.standard_normal((nobs, 2)) + 1j * rng.standard_normal((nobs, 2))
    sig = G.SIGMA_OCEAN_IXYH
    assert sig == float(np.float32(3.3)) and sig != 3.3
    e, vf, tot, vh, h = O.ixyhout(xyz, exy, stack, item, val, vxyz, fxyz, sig)
    cross = np.cross(v0, F0)
    for o in range(nobs):
        hh = max(-xyz[o, 2], 0.0)
        assert h[o] == hh
        assert np.allclose(vf[o], sig * hh * cross[:2] * 1e-3, rtol=1e-14, atol=0)
        assert np.allclose(e[o], sig * hh * exy[o], rtol=1e-15, atol=0)
        assert np.allclose(tot[o], e[o] + vf[o], rtol=1e-14, atol=1e-300)
        assert np.allclose(vh[o], hh * v0[:2], rtol=1e-14, atol=0)
    assert np.all(e[4] == 0) and np.all(tot[4] == 0) and np.all(vh[4] == 0)


EDIT_VALUES = [1.0, -2.5e-5, 0.0, 0.5, 12345.678, 9999999.6, 0.099999996, -180.0, 1.0 / 44714.0]


def test_fortran_edits_match_compiler():
    """tests/golden/fortran_edits.txt holds what AMD flang prints for these values with (e15.7)
    and (g15.7), and the (e9.2) tag of 1/44714 Hz (made by tests/golden/make_fortran_edits.sh)."""
    path = os.path.join(os.path.dirname(__file__), "golden", "fortran_edits.txt")
    lines = [line.rstrip("\n")[1:-1] for line in open(path)]
    assert len(lines) == 2 * len(EDIT_VALUES) + 1
    for i, x in enumerate(EDIT_VALUES):
        assert G.fortran_e(x) == lines[2 * i], x
        assert G.fortran_g(x) == lines[2 * i + 1], x
    assert G.fortran_freq_tag(1.0 / 44714.0) == lines[-1]


def test_file_names_follow_the_reference():
    # Kikai/fwd/plot_ixyh.sh reads ./result/ixyh_xy2D0.22E-04.dat for the M2 frequency 1/44714 Hz
    assert G.fortran_freq_tag(1.0 / 44714.0) == " 0.22E-04"
    assert os.path.basename(G.ixyh_path("./result/", 1.0 / 44714.0)) == "ixyh_xy2D0.22E-04.dat"
    assert G.outfreqfiles2_path("out/", "xy", 1.0 / 44714.0, 1) == "out/xy0.22E-04_S1.dat"


def test_writers_parse_back(tmp_path):
    rng = np.random.default_rng(1)
    nobs = 6
    xyz = rng.standard_normal((nobs, 3)) * 50
    r = {k: rng.standard_normal((nobs, 2)) + 1j * rng.standard_normal((nobs, 2))
         for k in ("ixyh_e", "ixyh_vF", "ixyh", "vh")}
    r["bxyz"] = rng.standard_normal((nobs, 3)) + 1j * rng.standard_normal((nobs, 3))
    r["h"] = rng.uniform(0, 5, nobs)
    p = tmp_path / "ixyh.dat"
    G.write_ixyh(str(p), xyz, r)
    rows = [[float(t) for t in line.split()] for line in open(p)]
    assert len(rows) == nobs and all(len(row) == 26 for row in rows)
    for o in range(nobs):
        want = list(xyz[o])
        for k in ("ixyh_e", "ixyh_vF", "ixyh", "vh", "bxyz"):
            for c in r[k][o]:
                want += [c.real, c.imag]
        want.append(r["h"][o])
        assert np.allclose(rows[o], want, rtol=5e-7, atol=0)
    fields = rng.standard_normal((2, 5, nobs)) + 1j * rng.standard_normal((2, 5, nobs))
    paths = G.write_outfreqfiles2(str(tmp_path) + "/", "xy", 2.2e-5, xyz, fields)
    assert [os.path.basename(q) for q in paths] == ["xy0.22E-04_S1.dat", "xy0.22E-04_S2.dat"]
    rows = np.loadtxt(paths[1])
    assert rows.shape == (nobs, 12)
    assert np.allclose(rows[:, 2], np.abs(fields[1, 0]), rtol=5e-7)
    assert np.allclose(rows[:, 3], np.degrees(np.angle(fields[1, 0])), rtol=5e-7, atol=1e-5)
    resp = rng.standard_normal((3, 1, 5, 2)) + 1j * rng.standard_normal((3, 1, 5, 2))
    paths = G.write_outobsfilesfwd(str(tmp_path) + "/", ["SITE1", "SITE2"], ["S1"], [1e-5, 2e-5, 3e-5], resp)
    rows = np.loadtxt(paths[1])
    assert rows.shape == (3, 11)
    assert np.allclose(rows[:, 0], [1e-5, 2e-5, 3e-5], rtol=5e-7)
    assert np.allclose(rows[:, 1], np.abs(resp[:, 0, 0, 1]), rtol=5e-7)
