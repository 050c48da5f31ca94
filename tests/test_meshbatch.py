"""The stacked (block-diagonal) system of several meshes, on the CPU.

What pins it:
- block_diag_split's arrays, read back as a matrix, equal scipy.sparse.block_diag of the parts'
  matrices, and keep the split-triangle invariants (IAU ascending and > row, AL mirrors AU);
- the tolerance rule: the oracle's COCG13_single (SSOR) on the stacked system, with unit-norm
  right-hand sides and the joint tolerance tol / sqrt(m), leaves every part's FP64 true residual
  <= tol, and a part's solution, scaled back, is as close to the direct solution (scipy SuperLU)
  as the part's own solve to the same tolerance allows;
- zero right-hand sides and size mismatches;
- afx_create_stacked refuses offsets that do not cover the rows and entries that couple two
  systems (checked on the host before any device work).
"""
import math

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from actfemtide_amd import mesh
from actfemtide_amd import meshbatch as MB
from oracle import oracle as O

OMEGAS = (2 * math.pi / 44714.0, 2 * math.pi / 43200.0, 2 * math.pi / 86164.0)


@pytest.fixture(scope="module")
def parts():
    out = []
    for i, dims in enumerate([(5, 4, 6), (6, 6, 5), (8, 7, 6)]):
        c = mesh.tide_case(*dims, seed=40 + i)
        out.append(O.forward_system(c, OMEGAS[i]))
    return out


def _arrays(o):
    return (o["INU"], o["IAU"], o["INL"], o["IAL"]), (o["D"], o["AU"], o["AL"])


def test_block_diag_split_equals_scipy(parts):
    pats, vals = zip(*[_arrays(o) for o in parts])
    INU, IAU, INL, IAL, D, AU, AL = MB.block_diag_split(pats, vals)
    A = O.to_scipy(INU, IAU, INL, IAL, D, AU, AL)
    ref = sp.block_diag([O.to_scipy(*p, *v) for p, v in zip(pats, vals)], format="csr")
    assert A.shape == ref.shape
    assert (A != ref).nnz == 0
    n = len(D)
    for i in range(n):
        u = IAU[INU[i]:INU[i + 1]]
        assert np.all(np.diff(u) > 0) and np.all(u > i + 1)
    # AL(j, i) == AU(i, j)
    ru = np.repeat(np.arange(n), np.diff(INU))
    rl = np.repeat(np.arange(n), np.diff(INL))
    U = sp.coo_matrix((AU, (IAU - 1, ru)), shape=(n, n)).tocsr()
    L = sp.coo_matrix((AL, (rl, IAL - 1)), shape=(n, n)).tocsr()
    assert (U != L).nnz == 0


def test_joint_tolerance_bounds_every_part(parts):
    tol = 1e-8
    pats, vals = zip(*[_arrays(o) for o in parts])
    sizes = [len(o["D"]) for o in parts]
    INU, IAU, INL, IAL, D, AU, AL = MB.block_diag_split(pats, vals)
    # right-hand sides of very different size: the scaling is what makes the rule hold
    bs = [o["b"] * f for o, f in zip(parts, (1.0, 1e4, 1e-3))]
    b, norms = MB.stack_rhs(sizes, bs)
    assert np.allclose(norms, [np.linalg.norm(x) for x in bs], rtol=1e-15)
    jt = MB.joint_tol(tol, norms)
    assert jt == pytest.approx(tol / math.sqrt(3), rel=1e-15)
    r = O.cocg13(INU, IAU, INL, IAL, D, AU, AL, b, precond="SSOR", resid=jt, maxit=100000)
    assert r["error"] == 0
    off = np.concatenate([[0], np.cumsum(sizes)])
    for s, (p, v) in enumerate(zip(pats, vals)):
        A = O.to_scipy(*p, *v)
        xs = r["x"][off[s]:off[s + 1]] * norms[s]
        tr = np.linalg.norm(bs[s] - A @ xs) / norms[s]
        own = O.cocg13(*p, *v, bs[s], precond="SSOR", resid=tol, maxit=100000)
        x_dir = spla.spsolve(A.tocsc(), bs[s])
        e_joint = np.linalg.norm(xs - x_dir) / np.linalg.norm(x_dir)
        e_own = np.linalg.norm(own["x"] - x_dir) / np.linalg.norm(x_dir)
        print(f"part {s} n={sizes[s]}: true residual {tr:.2e}, error joint {e_joint:.1e} own {e_own:.1e}, "
              f"iterations joint {r['iter']} own {own['iter']}")
        assert tr <= tol
        assert e_joint <= max(10 * e_own, 1e-12)


def test_zero_and_mismatched_right_hand_sides(parts):
    sizes = [len(o["D"]) for o in parts]
    bs = [parts[0]["b"], np.zeros(sizes[1]), parts[2]["b"]]
    b, norms = MB.stack_rhs(sizes, bs)
    assert norms[1] == 0.0 and not np.any(b[sizes[0]:sizes[0] + sizes[1]])
    assert MB.joint_tol(1e-8, norms) == pytest.approx(1e-8 / math.sqrt(2), rel=1e-15)
    with pytest.raises(ValueError):
        MB.stack_rhs(sizes, bs[:2])
    with pytest.raises(ValueError):
        MB.stack_rhs(sizes, [bs[0], bs[1][:-1], bs[2]])
    pats, vals = zip(*[_arrays(o) for o in parts])
    with pytest.raises(ValueError):
        MB.block_diag_split(pats, vals[:2])
    with pytest.raises(ValueError):
        MB.block_diag_split(pats[:1], [(vals[0][0][:-1], vals[0][1], vals[0][2])])


def test_create_stacked_rejects_bad_input(parts):
    from actfemtide_amd import afx
    pats, vals = zip(*[_arrays(o) for o in parts])
    INU, IAU, INL, IAL, D, AU, AL = MB.block_diag_split(pats, vals)
    sizes = [len(o["D"]) for o in parts]
    off = np.concatenate([[0], np.cumsum(sizes)])
    bad_offsets = [off[:-1], np.r_[1, off[1:]], np.r_[off[:2], off[1:]]]  # short of n, not from 0, empty system
    for ro in bad_offsets:
        with pytest.raises(afx.AfxError) as e:
            afx.System.from_stacked(ro, INU, IAU, INL, IAL)
        assert e.value.code == afx.AFX_ERR_ARG
    # the first two systems as one claimed to be two: their rows' entries cross the cut
    cut = np.r_[0, sizes[0] // 2, off[1:]]
    with pytest.raises(afx.AfxError) as e:
        afx.System.from_stacked(cut, INU, IAU, INL, IAL)
    assert e.value.code == afx.AFX_ERR_ARG and "couples" in str(e.value)
