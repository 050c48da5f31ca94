"""Gauss-Newton model updates of the tidal inversion on the device, with the forward and adjoint
solutions of one update kept as the initial guesses of the next.

One update is the body of the reference's inversion iteration (src/src_inv_ap/n_inv_ap.f90:258-430)
for one process holding every constituent:
  1. model2cond (:259): sigma of each crust cell = 10**m of its parameter;
  2. per constituent, forward_apinv (:280, forward_apinv.f90:64-71): assembly, boundary
     conditions, the forward solve A x = b and the adjoint solves A U = P^T of the observed field
     components (solvePARDISOinv); here every constituent's forward solve is one column of a
     k-column launch and each constituent's adjoint columns are one call;
  3. CALOBSEBCOMP (:284): the site fields F = P x (B = face rows . x);
  4. genjacobian1 (:303, m_jacobian_ap.f90:174-386): dF/dm accumulated from U and x, turned into
     the rows of log10 amplitude and phase in degrees (:338-345);
  5. GENDVEC_AP / CALRMS_AP (:343-345): the data vector and nRMS = sqrt(sum r^2 / Cd / ndat);
  6. alpha (:398-420): at the first update gamma * alphaspectralradius_v2 (times beta^2 for the
     minimum-support stabiliser), afterwards the cooling alpha * 10**(-1/3) whenever nRMS fell by
     less than 10%;
  7. getnewmodel (:424, m_jacobian_ap.f90:623-787) with Cm = GENBMI_MS of the model change.
The outer control of that program (L-curves over alpha, replacing the reference model, its output
files, MPI) is not restated. The phases' misfit is wrapped into (-180, 180] degrees.

The initial guess of every solve of update k + 1 is the solution of update k when `warm` is set
(solves are otherwise started from zero): the model changes little between updates, so the
previous fields are close to the new ones in every component, unlike a predicted field.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import afx, mesh
from . import regularization as reg


def crust_parameters(case):
    """One parameter per crust cell (its 6 tets), numbered by increasing cell id as
    regularization.crust_cell_models numbers the models: (par_ptr, par_tet 0-based, cell ids)."""
    crust = np.nonzero(case.group >= mesh.GROUP_CRUST)[0]
    cells, inv = np.unique(case.cell_of_tet[crust], return_inverse=True)
    order = np.argsort(inv, kind="stable")
    ptr = np.concatenate([[0], np.cumsum(np.bincount(inv, minlength=cells.size))])
    return ptr.astype(np.int32), crust[order].astype(np.int32), cells


def sigma_of_model(case, par_ptr, par_tet, m):
    """Per-tet conductivity with crust cell j at 10**m[j] (model2cond); air and sea as in the case."""
    s = case.sigma_tet()
    s[par_tet] = np.repeat(10.0 ** np.asarray(m, np.float64), np.diff(par_ptr))
    return s


def central_sites(case, n, seed, half=3):
    """Seeded synthetic sites 1 m above the sea bottom over the central (2 half)^2 cells, each
    strictly inside a tet (away from the cell's diagonal faces)."""
    rng = np.random.default_rng(seed)
    h = case.h
    pts = []
    for _ in range(n):
        i = case.nx // 2 + rng.integers(-half, half)
        j = case.ny // 2 + rng.integers(-half, half)
        fx, fy = rng.uniform(0.13, 0.87, size=2)
        if abs(fx - fy) < 0.05:
            fy = min(0.87, fy + 0.1)
        pts.append([-case.nx * h / 2 + (i + fx) * h, -case.ny * h / 2 + (j + fy) * h, -h + 0.001])
    return np.array(pts)


def block_model(case, cells, value, background, half=2, layers=3):
    """A synthetic model: log10 sigma `value` in the crust cells within `half` cells of the centre
    horizontally and in the top `layers` crust layers, `background` elsewhere. Returns (model,
    mask of the block)."""
    nx, ny = case.nx, case.ny
    i, j, k = cells % nx, (cells // nx) % ny, cells // (nx * ny)
    inside = (abs(i - nx // 2) <= half) & (abs(j - ny // 2) <= half) & (k > k.max() - layers)
    m = np.full(cells.size, float(background))
    m[inside] = value
    return m, inside


def wrap_degrees(d):
    return (np.asarray(d) + 180.0) % 360.0 - 180.0


@dataclass
class Update:
    """What one update measured."""
    nrms: float                     # of the model the update started from
    alpha: float
    forward_iters: list             # per constituent
    adjoint_iters: list             # per constituent: iterations of each adjoint column
    worst_true_resid: float         # largest FP64 true residual over every solve of the update
    solve_seconds: float            # forward + adjoint solves (device work, synchronised)
    seconds: float                  # the whole update
    model: np.ndarray = field(repr=False, default=None)   # the new model


class GaussNewton:
    """Gauss-Newton updates on one box case.

    sites: (nsite, 3) observation points; omegas: angular frequencies of the constituents; the
    observed components are Bx, By and Bz at every site, as log10 amplitude and phase. `d_obs`
    is the observed data vector (see `data`); `m_ref` the reference (and starting) model, one
    log10 sigma per crust cell. err_amp / err_pha: data errors of log10 amplitude and of phase in
    degrees (Cd = err^2; err_pha defaults to the phase error equal to err_amp in relative terms);
    beta: the minimum-support stabiliser's beta; gamma: alpha's factor on the spectral radius."""

    def __init__(self, case, sites, omegas, m_ref, *, err_amp=0.01, err_pha=None, beta=10.0, gamma=0.1,
                 precond="NODEBLK+GRAD", tol=1e-8, maxit=40000, warm=True, autotune=True):
        import torch
        self.case = case
        self.omegas = np.asarray(omegas, np.float64)
        self.nf = len(self.omegas)
        self.par_ptr, self.par_tet, self.cells = crust_parameters(case)
        self.nmodel = len(self.par_ptr) - 1
        self.m_ref = np.asarray(m_ref, np.float64).copy()
        if self.m_ref.shape != (self.nmodel,):
            raise ValueError(f"m_ref needs one value per crust cell ({self.nmodel})")
        self.beta, self.gamma = float(beta), float(gamma)
        self.precond, self.tol, self.maxit, self.warm = precond, float(tol), int(maxit), bool(warm)
        w0 = float(self.omegas[0])
        self.s, b, self.bc = afx.forward_system(case, w0, sigma_tet=sigma_of_model(case, self.par_ptr, self.par_tet,
                                                                                  self.m_ref))
        self.n = self.s.n
        self.item, _, self.vb, _ = self.s.prepobscoeff(np.asarray(sites, np.float64))
        self.nsite = self.item.shape[0]
        self.nobs = 3 * self.nsite
        # P^T: the B rows of every site, component-major (column c * nsite + o), kept as contiguous
        # blocks of equal width (a solve call takes at most 256 columns)
        stack = np.arange(0, 6 * self.nsite + 1, 6, dtype=np.int32)
        cols = []
        for c in range(3):
            d = torch.zeros((self.n, self.nsite), dtype=torch.complex128, device="cuda")
            self.s.adjoint_rhs_device(stack, self.item, self.vb[c], d)
            cols.append(d)
        dP = torch.cat(cols, dim=1)
        del cols
        nb = -(-self.nobs // 256)
        w = -(-self.nobs // nb)
        self.blocks = [(a, min(a + w, self.nobs)) for a in range(0, self.nobs, w)]
        self.dP = [dP[:, a:b].contiguous() for a, b in self.blocks]
        del dP
        self.dB = torch.from_numpy(np.ascontiguousarray(np.tile(b[:, None], (1, self.nf)))).cuda()
        self.scale_f = self.omegas / w0
        # the phase error that matches err_amp in relative terms: d log10|F| = Re(dF/F) / ln 10 and
        # d phase = Im(dF/F) * 180 / pi
        if err_pha is None:
            err_pha = err_amp * math.log(10.0) * 180.0 / math.pi
        self.cd = np.tile(np.r_[np.full(self.nobs, err_amp ** 2), np.full(self.nobs, err_pha ** 2)], self.nf)
        self.ndat = self.cd.size
        self.dX = torch.zeros_like(self.dB)                               # forward fields, (n, nf)
        self.dU = [[torch.zeros_like(p) for p in self.dP] for _ in range(self.nf)]  # adjoint blocks per constituent
        self.alpha = None
        self.nrms_prev = None
        self.d_obs = None
        self.printer = None   # optional callable: a progress line per solve call
        if autotune:
            self.s.precond_setup(self.precond)
            self.s.autotune(1)
            if self.nf > 1:
                self.s.autotune(self.nf)
            for k in sorted({b - a for a, b in self.blocks}):
                self.s.autotune(k)

    def close(self):
        self.s.close()

    # -- forward model ------------------------------------------------------------------------
    def _assemble(self, m):
        b = self.s.assemble(sigma_of_model(self.case, self.par_ptr, self.par_tet, m), float(self.omegas[0]))
        self.s.set_bc(self.bc, b)
        self.s.precond_setup(self.precond)

    def _solve(self, dB, dX, scale):
        import torch
        if not self.warm:
            dX.zero_()
        # up to two FP64 restarts from the iterate when the true residual is still above tol
        st, rc = self.s.solve_device(dB, dX, omega_scale=scale, precond=self.precond, tol=self.tol, maxit=self.maxit,
                                     max_outer=3)
        tr = self.s.true_residual_device(dB, dX, omega_scale=scale)
        if rc != 0 or not np.all(tr <= self.tol * 1.05):
            raise afx.AfxError(rc if rc != 0 else afx.AFX_ERR_MAXIT,
                               f"solve did not reach {self.tol:g}: true residual {tr.max():.2e}")
        torch.cuda.synchronize()
        return [x.iter for x in st], float(tr.max())

    def fields(self):
        """Site fields of the current forward solutions: (nf, nobs) complex (component-major)."""
        out = np.zeros((self.nf, self.nobs), np.complex128)
        for f in range(self.nf):
            x = self.dX[:, f].contiguous()
            for c in range(3):
                out[f, c * self.nsite:(c + 1) * self.nsite] = self.s.obs_apply_device(self.item, self.vb[c], x)[:, 0]
        return out

    @staticmethod
    def data(F):
        """Data vector of site fields F (nf, nobs): per constituent [log10 |F|, phase in degrees]."""
        amp, pha = afx.amp_phase(F)
        return np.concatenate([np.r_[np.log10(amp[f]), pha[f]] for f in range(F.shape[0])])

    def forward(self, m):
        """Solve every constituent at model m; returns (site fields, iterations, worst residual)."""
        self._assemble(m)
        its, tr = self._solve(self.dB, self.dX, self.scale_f if self.nf > 1 else None)
        return self.fields(), its, tr

    # -- one update ---------------------------------------------------------------------------
    def update(self, m):
        """One Gauss-Newton update from model m (needs d_obs). Returns an Update with the new model."""
        import torch
        if self.d_obs is None:
            raise ValueError("set d_obs first")
        t0 = time.perf_counter()
        m = np.asarray(m, np.float64)
        self._assemble(m)
        ts = time.perf_counter()
        fits, worst = self._solve(self.dB, self.dX, self.scale_f if self.nf > 1 else None)
        if self.printer:
            self.printer(f"forward solves: iterations {fits} ({time.perf_counter() - ts:.1f}s)")
        aits = []
        for f in range(self.nf):
            it = []
            for dP, dU in zip(self.dP, self.dU[f]):
                sc = np.full(dP.shape[1], self.scale_f[f]) if self.nf > 1 else None
                itb, tr = self._solve(dP, dU, sc)
                it += itb
                worst = max(worst, tr)
            aits.append(it)
            if self.printer:
                self.printer(f"constituent {f}: {self.nobs} adjoint solves, {sum(it)} iterations "
                             f"({time.perf_counter() - ts:.1f}s)")
        solve_s = time.perf_counter() - ts
        F = self.fields()
        d_cal = self.data(F)
        sigma = sigma_of_model(self.case, self.par_ptr, self.par_tet, m)
        dJr = torch.empty((self.ndat, self.nmodel), dtype=torch.float64, device="cuda")
        dJ = torch.empty((self.nmodel, self.nobs), dtype=torch.complex128, device="cuda")
        for f in range(self.nf):
            xf = self.dX[:, f].contiguous()
            for (a, b), dU in zip(self.blocks, self.dU[f]):
                dJb = dJ if len(self.blocks) == 1 else torch.empty((self.nmodel, b - a), dtype=torch.complex128,
                                                                    device="cuda")
                self.s.jacobian_device(dU, xf, self.par_ptr, self.par_tet, sigma, float(self.omegas[f]),
                                       math.log(10.0), dJb)
                if len(self.blocks) > 1:
                    dJ[:, a:b] = dJb
            rows = dJr[2 * self.nobs * f:2 * self.nobs * (f + 1)].view(2, self.nobs, self.nmodel)
            afx.jacobian_ampphase_device(dJ, torch.from_numpy(F[f]).cuda(), rows)
        # misfit with the phases wrapped; the update sees d_obs through d_cal + that misfit
        res = self.d_obs - d_cal
        for f in range(self.nf):
            k0 = 2 * self.nobs * f + self.nobs
            res[k0:k0 + self.nobs] = wrap_degrees(res[k0:k0 + self.nobs])
        nrms = float(np.sqrt(np.sum(res * res / self.cd) / self.ndat))
        dm = m - self.m_ref
        Cm = reg.bmi_ms(dm, self.beta)
        if self.alpha is None:
            a, _, _ = afx.spectral_alpha(dJr, Cm, self.cd, 30)
            self.alpha = self.gamma * a * self.beta ** 2
        elif nrms / self.nrms_prev > 0.9:
            self.alpha *= 10.0 ** (-1.0 / 3.0)
        self.nrms_prev = nrms
        m_new, _ = afx.gn_update(dJr, Cm, self.cd, self.alpha, d_cal + res, d_cal, m, self.m_ref)
        torch.cuda.synchronize()
        return Update(nrms=nrms, alpha=self.alpha, forward_iters=fits, adjoint_iters=aits, worst_true_resid=worst,
                      solve_seconds=solve_s, seconds=time.perf_counter() - t0, model=m_new)

    def run(self, m0, updates=3):
        """`updates` Gauss-Newton updates from m0; returns (list of Update, final model, its nRMS)."""
        out, m = [], np.asarray(m0, np.float64)
        for _ in range(updates):
            u = self.update(m)
            out.append(u)
            m = u.model
        F, _, _ = self.forward(m)
        res = self.d_obs - self.data(F)
        for f in range(self.nf):
            k0 = 2 * self.nobs * f + self.nobs
            res[k0:k0 + self.nobs] = wrap_degrees(res[k0:k0 + self.nobs])
        return out, m, float(np.sqrt(np.sum(res * res / self.cd) / self.ndat))


if __name__ == "__main__":
    import argparse
    import json

    ap = argparse.ArgumentParser(description="time warm-started Gauss-Newton updates on a synthetic block model")
    ap.add_argument("--grid", type=int, nargs=3, default=(64, 64, 48), metavar=("NX", "NY", "NZ"))
    ap.add_argument("--sites", type=int, default=96)
    ap.add_argument("--constituents", type=int, default=5)
    ap.add_argument("--updates", type=int, default=3)
    ap.add_argument("--block", type=float, default=-1.5, help="log10 sigma of the block (background -2)")
    ap.add_argument("--cold", action="store_true", help="start every solve from zero instead")
    a = ap.parse_args()
    om = [2 * math.pi / mesh.PERIODS[c] for c in ("M2", "S2", "N2", "K1", "O1")[:a.constituents]]
    t0 = time.perf_counter()

    def say(m):
        print(f"[gn {time.perf_counter() - t0:7.1f}s] {m}", flush=True)

    case = mesh.tide_case(*a.grid, seed=1003, sigma_sea=3.3)
    sites = central_sites(case, a.sites, 5, half=max(2, a.grid[0] // 8))
    ncell = len(crust_parameters(case)[2])
    gn = GaussNewton(case, sites, om, np.full(ncell, -2.0), beta=10.0, gamma=0.1, warm=not a.cold)
    say(f"{a.grid} cells, {gn.n} edges, {ncell} parameters, {gn.nobs} adjoint columns x {gn.nf} constituents")
    m_true, _ = block_model(case, gn.cells, a.block, -2.0, half=max(2, a.grid[0] // 10))
    F, its, _ = gn.forward(m_true)
    gn.d_obs = gn.data(F)
    gn.dX.zero_()
    say(f"synthetic data from the block model (forward iterations {its})")
    gn.printer = say
    m = gn.m_ref
    for k in range(a.updates):
        u = gn.update(m)
        m = u.model
        print(json.dumps(dict(update=k + 1, nrms=u.nrms, alpha=u.alpha, forward_iters=u.forward_iters,
                              adjoint_iters=[int(sum(x)) for x in u.adjoint_iters], solve_seconds=round(u.solve_seconds, 2),
                              update_seconds=round(u.seconds, 2), worst_true_resid=u.worst_true_resid)), flush=True)
    gn.close()
