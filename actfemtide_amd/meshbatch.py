"""Several independent systems -- meshes of different sizes, each with its own conductivity
model, frequency and right-hand side -- solved by one persistent COCG launch.

The reference solves one mesh per call (forward_bxyz, src/solver_mpi/forward_bxyz_mpi.f90:3-67,
one COCG13_single or solvePARDISO call per frequency). Here m assembled systems A_s x_s = b_s are
stacked into the block-diagonal system diag(A_1, ..., A_m) [x_1; ...; x_m] = [b_1; ...; b_m],
whose split-triangle arrays are the parts' arrays with row and column ids offset. The SpMV and
the SSOR sweeps never couple two parts, and the sweeps' level sets run the parts' rows side by
side, so a launch costs the longest part's dependency chain, not the sum.

Two ways to run COCG on the stack:
- per system (default; afx_create_stacked + afx_solve_stacked_device): the sweep blocks are cut at
  the part boundaries and every part keeps its own alpha, beta, stopping test and true residual
  check, so each part follows its own COCG13_single iteration;
- shared scalars (afx_create + afx_solve_device on the stacked arrays): one COCG over the whole
  stack. Every right-hand side is then scaled to unit norm, so |b| = sqrt(m), and stopping at a
  relative residual of tol / sqrt(m) gives sum_s |r_s|^2 <= tol^2, hence |r_s| / |b_s| <= tol for
  every part; the FP64 true residual is checked at the end and the solve restarts from its
  iterate while it is above that bound.
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
    """The block-diagonal stack of several assembled afx.System parts (boundary conditions
    already applied), on the device as one afx.System.

    per_system=True builds the stacked handle whose solve keeps the scalars per part; False
    builds a plain handle for the shared-scalar COCG. The parts' values are copied at
    construction; the parts themselves are only used again to report each part's true residual
    in the shared-scalar solve."""

    def __init__(self, parts, per_system: bool = True, *, patterns=None, values=None):
        parts = list(parts)
        if patterns is None:
            if not parts:
                raise ValueError("no parts")
            patterns = [p.pattern() for p in parts]
            values = [p.get_values() for p in parts]
        self.parts = parts
        self.per_system = per_system
        self.sizes = [len(pt[0]) - 1 for pt in patterns]
        self.offsets = np.concatenate([[0], np.cumsum(self.sizes)]).astype(np.int64)
        arrays = block_diag_split(patterns, values)
        if per_system:
            self.system = afx.System.from_stacked(self.offsets, *arrays[:4])
        else:
            self.system = afx.System.from_pattern(*arrays[:4])
        self.system.set_values(*arrays[4:])
        self.n = self.system.n

    @classmethod
    def frequency_stack(cls, system, omegas) -> "MeshBatch":
        """One assembled system (afx_assemble values, boundary conditions applied) stacked once
        per angular frequency in omegas, copy c holding K + i*omegas[c]*M, for the per-system
        solve: the right-hand sides of several tidal constituents, or several right-hand sides
        at one frequency, solved as independent systems in one launch. The system's own
        frequency is restored afterwards."""
        if system.omega is None:
            raise ValueError("the system has no assembled frequency (afx_assemble first)")
        pat = system.pattern()
        omega0 = system.omega
        vals = []
        try:
            for w in omegas:
                system.set_omega(float(w))
                vals.append(system.get_values())
        finally:
            system.set_omega(omega0)
        return cls([], True, patterns=[pat] * len(vals), values=vals)

    def set_nodes(self, node_counts, lines, dirichlets):
        """Edge-to-node incidence of the stack for the nodal gradient-space correction
        (afx_set_nodes): each part's line (nline_s, 2) 1-based and Dirichlet mask, its node ids
        offset by the nodes of the parts before it, so the stacked G is block diagonal too."""
        if not (len(node_counts) == len(lines) == len(dirichlets) == self.m):
            raise ValueError("one node count, line array and Dirichlet mask per part")
        off, ls = 0, []
        for nc, ln in zip(node_counts, lines):
            ls.append(np.asarray(ln, np.int64).reshape(-1, 2) + off)
            off += int(nc)
        self.system.set_nodes(off, np.concatenate(ls), np.concatenate([np.asarray(d, bool) for d in dirichlets]))

    @property
    def m(self) -> int:
        return len(self.sizes)

    def close(self):
        self.system.close()

    def joint_rhs(self, bs):
        """The stacked right-hand side and the parts' norms: as given for the per-system solve,
        each part scaled to unit norm for the shared-scalar solve (stack_rhs)."""
        b, norms = stack_rhs(self.sizes, bs)
        if self.per_system:
            b = np.concatenate([np.asarray(x, np.complex128).ravel() for x in bs])
        return b, norms

    def solve_device(self, dB, dX, tol=1e-8, norms=None, **opts):
        """Solve the stacked right-hand side dB (from joint_rhs) into dX, both torch complex128
        tensors of n elements on the GPU; dX holds the initial guess. Returns (stats, return
        code): one Stats per part for the per-system solve, one for the shared-scalar solve."""
        opts.setdefault("check_true", 1)
        if self.per_system:
            return self.system.solve_stacked_device(dB, dX, tol=tol, **opts)
        nrm = np.ones(self.m) if norms is None else norms
        return self.system.solve_device(dB, dX, tol=joint_tol(tol, nrm), **opts)

    def part(self, v, s: int):
        """Part s of a stacked vector (a view)."""
        return v[int(self.offsets[s]):int(self.offsets[s + 1])]

    def solve(self, bs, tol=1e-8, **opts):
        """Solve every part to tol; bs are host arrays. Returns (xs, stats, per-part FP64 true
        residuals |b_s - A_s x_s| / |b_s|)."""
        import torch
        b, norms = self.joint_rhs(bs)
        dB = torch.from_numpy(b).cuda()
        dX = torch.zeros_like(dB)
        st, rc = self.solve_device(dB, dX, tol=tol, norms=norms, **opts)
        x = dX.cpu().numpy()
        if self.per_system:
            xs = [self.part(x, s).copy() for s in range(self.m)]
            return xs, st, np.array([t.true_resid for t in st])
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
        return xs, st, np.array(tr)
