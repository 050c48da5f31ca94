"""Several independent systems -- meshes of different sizes, each with its own conductivity
model, frequency and right-hand side -- solved by one persistent COCG launch.

The reference solves one mesh per call (forward_bxyz, src/solver_mpi/forward_bxyz_mpi.f90:3-67,
one COCG13_single or solvePARDISO call per frequency). Here m assembled systems A_s x_s = b_s are
stacked into the block-diagonal system

    diag(A_1, ..., A_m) [x_1; ...; x_m] = [b_1/|b_1|; ...; b_m/|b_m|]

whose split-triangle arrays are the parts' arrays with row and column ids offset. COCG on it is
COCG on every part at once: the SpMV and the SSOR sweeps never couple two parts (the sweeps'
level sets run the parts' rows side by side, so a launch costs the longest part's dependency
chain, not the sum), and the scalars alpha and beta are shared. With every right-hand side
scaled to unit norm, |b| = sqrt(m), so stopping the joint solve at a relative residual of
tol / sqrt(m) gives |r|^2 = sum_s |r_s|^2 <= tol^2 and hence |r_s| / |b_s| <= tol for every
part. The FP64 true residual is recomputed at the end and the solve restarts from the current
iterate while it is above that bound, so the guarantee holds for the true residual too.
"""
from __future__ import annotations

import math

import numpy as np

from . import afx


def block_diag_split(patterns, values):
    """Split-triangle arrays (INU, IAU, INL, IAL 1-based, D, AU, AL) of diag(A_1, ..., A_m) from
    the parts' patterns [(INU, IAU, INL, IAL)] and values [(D, AU, AL)]: rows, upper and lower
    entries of part s follow those of part s - 1, with column ids offset by the rows before it."""
    if len(patterns) != len(values) or not patterns:
        raise ValueError("one (pattern, values) pair per part, at least one part")
    INU, INL = [np.zeros(1, np.int64)], [np.zeros(1, np.int64)]
    IAU, IAL, D, AU, AL = [], [], [], [], []
    off = 0
    for (inu, iau, inl, ial), (d, au, al) in zip(patterns, values):
        inu, inl = np.asarray(inu, np.int64), np.asarray(inl, np.int64)
        n = len(inu) - 1
        if len(d) != n or len(inl) != n + 1 or len(au) != inu[-1] or len(al) != inl[-1]:
            raise ValueError("a part's values do not match its pattern")
        INU.append(INU[-1][-1] + inu[1:])
        INL.append(INL[-1][-1] + inl[1:])
        IAU.append(np.asarray(iau, np.int64) + off)
        IAL.append(np.asarray(ial, np.int64) + off)
        D.append(np.asarray(d, np.complex128))
        AU.append(np.asarray(au, np.complex128))
        AL.append(np.asarray(al, np.complex128))
        off += n
    INU, INL = np.concatenate(INU), np.concatenate(INL)
    if INU[-1] >= 2 ** 31 or INL[-1] >= 2 ** 31 or off >= 2 ** 31:
        raise ValueError("the stacked system exceeds 32-bit indices")
    return (INU.astype(np.int32), np.concatenate(IAU).astype(np.int32), INL.astype(np.int32),
            np.concatenate(IAL).astype(np.int32), np.concatenate(D), np.concatenate(AU), np.concatenate(AL))


def stack_rhs(sizes, bs):
    """Each b_s scaled to unit norm and stacked; returns (b, norms). A zero b_s stays zero (its
    part's solution is 0, as COCG13_single returns for a zero right-hand side)."""
    if len(bs) != len(sizes):
        raise ValueError("one right-hand side per part")
    offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    norms = np.zeros(len(sizes))
    b = np.zeros(int(offsets[-1]), np.complex128)
    for s, bb in enumerate(bs):
        bb = np.asarray(bb, np.complex128).ravel()
        if bb.size != sizes[s]:
            raise ValueError(f"part {s}: right-hand side of {bb.size} rows, system of {sizes[s]}")
        norms[s] = float(np.linalg.norm(bb))
        if norms[s] > 0.0:
            b[offsets[s]:offsets[s + 1]] = bb / norms[s]
    return b, norms


def joint_tol(tol: float, norms) -> float:
    """Relative tolerance of the joint solve that bounds every part's relative residual by tol:
    with unit-norm parts, |b| = sqrt(number of nonzero parts)."""
    nz = int(np.count_nonzero(np.asarray(norms) > 0.0))
    return tol / math.sqrt(max(nz, 1))


class MeshBatch:
    """The block-diagonal system of several assembled afx.System parts (boundary conditions
    already applied), on the device as one afx.System.

    The parts' values are copied at construction; the parts themselves are only used again to
    report each part's true residual."""

    def __init__(self, parts):
        parts = list(parts)
        if not parts:
            raise ValueError("no parts")
        self.parts = parts
        self.sizes = [p.n for p in parts]
        self.offsets = np.concatenate([[0], np.cumsum(self.sizes)]).astype(np.int64)
        arrays = block_diag_split([p.pattern() for p in parts], [p.get_values() for p in parts])
        self.system = afx.System.from_pattern(*arrays[:4])
        self.system.set_values(*arrays[4:])
        self.n = self.system.n

    @property
    def m(self) -> int:
        return len(self.parts)

    def close(self):
        self.system.close()

    def joint_rhs(self, bs):
        """stack_rhs over this batch's parts."""
        return stack_rhs(self.sizes, bs)

    def solve_device(self, dB, dX, tol=1e-8, norms=None, **opts):
        """Joint solve of the stacked unit-norm right-hand side dB (from joint_rhs) into dX, both
        torch complex128 tensors of n elements on the GPU; dX holds the initial guess. Returns
        (stats of the joint solve, return code)."""
        nrm = np.ones(self.m) if norms is None else norms
        opts.setdefault("check_true", 1)
        return self.system.solve_device(dB, dX, tol=joint_tol(tol, nrm), **opts)

    def part(self, v, s: int):
        """Part s of a stacked vector (a view)."""
        return v[int(self.offsets[s]):int(self.offsets[s + 1])]

    def solve(self, bs, tol=1e-8, **opts):
        """Solve every part to tol; bs are host arrays. Returns (xs, joint stats, per-part FP64
        true residuals |b_s - A_s x_s| / |b_s| recomputed on each part's own system)."""
        import torch
        b, norms = self.joint_rhs(bs)
        dB = torch.from_numpy(b).cuda()
        dX = torch.zeros_like(dB)
        st, rc = self.solve_device(dB, dX, tol=tol, norms=norms, **opts)
        x = dX.cpu().numpy()
        xs, tr = [], []
        for s in range(self.m):
            xp = self.part(x, s) * norms[s]
            xs.append(xp)
            bp = np.asarray(bs[s], np.complex128).ravel()
            if norms[s] == 0.0:
                tr.append(0.0)
                continue
            r = self.parts[s].true_residual_device(torch.from_numpy(bp).cuda(), torch.from_numpy(xp).cuda())
            tr.append(float(r[0]))
        return xs, st[0], np.array(tr)
