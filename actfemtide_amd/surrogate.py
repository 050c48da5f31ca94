"""Learned warm start for the tidal forward solve: data generation with the method itself,
training on the surrogate kernels, and warm-started COCG with fallback.

There is no reference counterpart (the reference has no learned code). The pipeline only
changes the initial guess X of the COCG solve (COCG13_single.f90:9-36, X in/out):
  1. conductivity models: crust log10 sigma as a Gaussian random field (mesh.gaussian_random_field)
     with the same air / sea layers, each with its own seeded tidal source;
  2. each model is assembled and solved on the GPU (this library's COCG) and the edge solution
     is projected to per-cell constant fields (the training targets);
  3. the encoder-decoder (nn.Surrogate) is trained with Adam on those pairs;
  4. for a new model the prediction is averaged onto edges (Dirichlet edges zero), scaled by the
     minimal-residual factor (afx_scale_initial_guess_device, which falls back to X = 0 when the
     prediction does not lower ||b - A X|| / ||b|| below 1), and COCG finishes the solve to the
     same tolerance as a cold start.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import afx, mesh, nn


def grf_case(grid, seed, std=0.7, corr_cells=4.0, sigma_sea=3.3, like=None):
    """A box case whose crust log10 sigma is a Gaussian random field (mean -2); `like` (a case of
    the same grid) shares its mesh arrays instead of building them again."""
    nx, ny, nz = grid
    f = mesh.gaussian_random_field(nz, ny, nx, seed, mean=-2.0, std=std, corr_cells=corr_cells)
    return mesh.tide_case(nx, ny, nz, seed=seed, log10_sigma_field=f, sigma_sea=sigma_sea, like=like)


@dataclass
class Sample:
    seed: int
    feat: object          # bf16 [ncell][cin] device tensor
    target: object        # f32 [ncell][6] device tensor (scaled cell fields)
    b: object             # complex128 [n] device right-hand side (after boundary conditions)
    x: object             # complex128 [n] device solution
    iters: int
    seconds: float


class Dataset:
    """Solved models on one grid. All models share the mesh, the pattern and one device System;
    only conductivities and sources change."""

    def __init__(self, grid, omega, device="cuda", cin=8):
        import torch
        self.grid = tuple(grid)
        self.omega = omega
        self.cin = cin
        self.dev = torch.device(device)
        base = grf_case(grid, 0)
        self.base = base
        self.s = afx.System.from_n6line(base.nline, base.n6line)
        self.s.set_mesh(base)
        xout, yout, zmin, zmax = base.bounds
        self.s.assemble(base.sigma_tet(), omega)
        self.bc = self.s.genbc(base.line, zmin, zmax, xout, yout)
        self.s.set_nodes(base.node, base.line, self.bc)  # the nodal correction of DIAG+GRAD
        self.bc_dev = torch.from_numpy(np.ascontiguousarray(self.bc, np.uint8)).to(self.dev)
        self.maps = nn.GridMaps(base, device=device)
        self.scale = None
        self.samples: list[Sample] = []

    def system_for(self, case):
        """Assemble and constrain the device system for `case`; returns the host rhs."""
        self.s.set_mesh(case)
        b = self.s.assemble(case.sigma_tet(), self.omega)
        return self.s.set_bc(self.bc, b)

    def features(self, case):
        import torch
        f = nn.cell_features(case, self.cin)
        return torch.from_numpy(f).to(self.dev).to(torch.bfloat16)

    def case(self, seed):
        """The model of `seed` on this grid (grf_case on the shared mesh)."""
        return grf_case(self.grid, seed, like=self.base)

    def add(self, seed, tol=1e-6, maxit=200000, precond="DIAG+GRAD"):
        import torch
        case = self.case(seed)
        b = self.system_for(case)
        self.s.precond_setup(precond)
        db = torch.from_numpy(np.ascontiguousarray(b, np.complex128)).to(self.dev)
        dx = torch.zeros_like(db)
        t0 = time.perf_counter()
        st, _ = self.s.solve_device(db, dx, precond=precond, tol=tol, maxit=maxit, chunk=2000)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        if self.scale is None:  # RMS of the first model's cell fields
            F1 = self.maps.edges_to_cells(dx, 1.0).cpu().numpy().astype(np.float64)
            self.scale = float(np.sqrt(np.mean(F1 ** 2))) or 1.0
        tgt = self.maps.edges_to_cells(dx, self.scale)
        smp = Sample(seed, self.features(case), tgt, db, dx, st[0].iter, dt)
        self.samples.append(smp)
        return smp


@dataclass
class TrainLog:
    losses: list = field(default_factory=list)
    seconds: float = 0.0
    steps: int = 0


def train(net: nn.Surrogate, data: Dataset, idx, steps=300, batch=4, lr=2e-3, lr_min=1e-4, seed=0, log_every=50,
          printer=None):
    """Adam on the masked-free MSE between the network output and the scaled cell fields."""
    import torch
    rng = np.random.default_rng(seed)
    nz, ny, nx = data.grid[2], data.grid[1], data.grid[0]
    V1 = nx * ny * nz
    log = TrainLog()
    t0 = time.perf_counter()
    idx = list(idx)
    for step in range(steps):
        pick = rng.choice(len(idx), size=min(batch, len(idx)), replace=False)
        xs = torch.cat([data.samples[idx[i]].feat for i in pick], 0)
        ts = torch.cat([data.samples[idx[i]].target for i in pick], 0)
        nb = len(pick)
        out = net.forward(xs, (nb, nz, ny, nx))
        loss, grad = net.loss(out, ts)
        net.backward(grad)
        frac = step / max(steps - 1, 1)
        net.step(lr=lr_min + 0.5 * (lr - lr_min) * (1 + math.cos(math.pi * frac)))
        if step % log_every == 0 or step == steps - 1:
            l = float(loss.item())
            log.losses.append((step, l))
            if printer:
                printer(f"step {step} loss {l:.4e}")
    torch.cuda.synchronize()
    log.seconds = time.perf_counter() - t0
    log.steps = steps
    _ = V1
    return log


def predict_x0(net: nn.Surrogate, data: Dataset, feat, air=False):
    """Network prediction averaged onto edges, complex128 [n]; zero on Dirichlet edges and,
    unless `air`, on edges that touch only air cells."""
    nz, ny, nx = data.grid[2], data.grid[1], data.grid[0]
    out = net.forward(feat, (1, nz, ny, nx))
    zero = data.bc_dev if air else (data.bc_dev | data.maps.air_only)
    return data.maps.cells_to_edges(out, data.scale, ldf=32, zero=zero)


def warm_solve(net, data: Dataset, case, tol=1e-8, precond="DIAG+GRAD", maxit=400000, accept=0.5, air=False):
    """Solve `case` cold (X = 0) and warm (surrogate X, minimal-residual scaled) and return a dict
    with iterations, times, initial residuals and the FP64 true residuals of both. The warm start
    is used only when its scaled relative residual is at most `accept`; otherwise the solve
    starts from X = 0 (the fallback), so a poor prediction costs one SpMV, not iterations."""
    import torch
    b = data.system_for(case)
    data.s.precond_setup(precond)
    db = torch.from_numpy(np.ascontiguousarray(b, np.complex128)).to(data.dev)
    x_cold = torch.zeros_like(db)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st_c, _ = data.s.solve_device(db, x_cold, precond=precond, tol=tol, maxit=maxit, chunk=2000)
    torch.cuda.synchronize()
    t_cold = time.perf_counter() - t0
    feat = data.features(case)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    x_warm = predict_x0(net, data, feat, air=air)
    g, r_given, r_scaled = data.s.scale_initial_guess_device(db, x_warm)
    fallback = bool(g[0] == 0) or float(r_scaled[0]) > accept
    if fallback:
        x_warm.zero_()
    st_w, _ = data.s.solve_device(db, x_warm, precond=precond, tol=tol, maxit=maxit, chunk=2000)
    torch.cuda.synchronize()
    t_warm = time.perf_counter() - t0
    res = data.s.true_residual_device(db, x_warm)
    res_c = data.s.true_residual_device(db, x_cold)
    return dict(iters_cold=st_c[0].iter, iters_warm=st_w[0].iter, t_cold=t_cold, t_warm=t_warm,
                r0_given=float(r_given[0]), r0_scaled=float(r_scaled[0]), gamma=complex(g[0]),
                true_cold=float(res_c[0]), true_warm=float(res[0]), fallback=fallback)


def config4(ntrain=2000, ntest=200, grid=(32, 32, 24), base=32, steps=4000, batch=4, lr=2e-3, tol_train=1e-6,
            accept=0.5, seed0=10000, test_seed0=900000, printer=print):
    """The learned warm start at its configuration's own size: ntrain Gaussian-random-field models
    (std 0.7, correlation 4 cells, M2) on the grid, solved by this library (DIAG+GRAD to tol_train)
    as training data; the 4-level encoder-decoder with `base` channels trained on them (Adam,
    `steps` steps of `batch` models); ntest held-out models each solved cold (X = 0) and
    warm-started, both to a 1e-8 FP64 true residual. Returns the measurements as a dict."""
    import torch
    t0 = time.perf_counter()
    d = Dataset(grid, 2 * math.pi / mesh.PERIODS["M2"])
    its = []
    for i in range(ntrain):
        its.append(d.add(seed0 + i, tol=tol_train).iters)
        if (i + 1) % 100 == 0:
            printer(f"[config4] {i + 1}/{ntrain} training models solved ({time.perf_counter() - t0:.0f}s)")
    torch.cuda.synchronize()
    t_data = time.perf_counter() - t0
    net = nn.Surrogate(base=base, cin=8, seed=0)
    log = train(net, d, range(ntrain), steps=steps, batch=batch, lr=lr, log_every=max(1, steps // 10),
                printer=lambda m: printer(f"[config4] {m}"))
    rows, kept = [], []
    for k in range(ntest):
        case = d.case(test_seed0 + k)
        rows.append(warm_solve(net, d, case, tol=1e-8, accept=accept))
        # the same model with the scaled prediction always kept (no fallback)
        kept.append(warm_solve(net, d, case, tol=1e-8, accept=math.inf))
        if (k + 1) % 25 == 0:
            printer(f"[config4] {k + 1}/{ntest} held-out models solved cold and warm")
    ic = np.array([r["iters_cold"] for r in rows])
    iw = np.array([r["iters_warm"] for r in rows])
    ik = np.array([r["iters_warm"] for r in kept])
    tc = sum(r["t_cold"] for r in rows)
    tw = sum(r["t_warm"] for r in rows)
    r0 = np.array([r["r0_scaled"] for r in rows])
    return dict(
        no_fallback=dict(iters_mean=round(float(ik.mean()), 1), iteration_gain=round(float(1.0 - ik.sum() / ic.sum()), 4),
                         models_fewer_iters=int(np.sum(ik < ic)), models_more_iters=int(np.sum(ik > ic)),
                         solves_per_s=round(ntest / sum(r["t_warm"] for r in kept), 3),
                         true_resid_max=max(r["true_warm"] for r in kept)),
        grid="x".join(map(str, grid)), base_channels=base, levels=4, train_models=ntrain, heldout_models=ntest,
        data_seconds=round(t_data, 1), train_iters_mean=round(float(np.mean(its)), 1), train_tol=tol_train,
        train_steps=steps, batch=batch, train_seconds=round(log.seconds, 2),
        ms_per_train_step=round(1e3 * log.seconds / max(log.steps, 1), 3),
        loss_first=log.losses[0][1], loss_last=log.losses[-1][1],
        heldout_iters_cold_mean=round(float(ic.mean()), 1), heldout_iters_warm_mean=round(float(iw.mean()), 1),
        iteration_gain=round(float(1.0 - iw.sum() / ic.sum()), 4),
        solves_per_s_cold=round(ntest / tc, 3), solves_per_s_warm=round(ntest / tw, 3),
        fallbacks=int(sum(r["fallback"] for r in rows)),
        r0_scaled_min=round(float(r0.min()), 4), r0_scaled_median=round(float(np.median(r0)), 4),
        r0_given_median=round(float(np.median([r["r0_given"] for r in rows])), 2),
        true_resid_max=max(max(r["true_cold"], r["true_warm"]) for r in rows))


def time_training_step(grid=(32, 32, 24), base=32, batch=4, steps=10, warmup=3, seed=0, device="cuda"):
    """Wall time of one Adam step (forward, MSE loss, backward, update) of nn.Surrogate on
    synthetic features and targets of the grid's shape (the kernels' work does not depend on
    the values). Returns ms per step, the GEMM calls of one step as (M, N, K, splits) (splits None:
    chosen by nn.gemm_nt) and their flops."""
    import torch
    nx, ny, nz = grid
    net = nn.Surrogate(base=base, cin=8, seed=seed)
    g = torch.Generator(device=device).manual_seed(seed)
    x = torch.randn(batch * nx * ny * nz, 8, device=device, generator=g).to(torch.bfloat16)
    tgt = torch.randn(batch * nx * ny * nz, 6, device=device, generator=g)

    def step():
        out = net.forward(x, (batch, nz, ny, nx))
        _, grad = net.loss(out, tgt)
        net.backward(grad)
        net.step(lr=1e-3)

    for _ in range(warmup):
        step()
    calls = []
    gemm = nn.gemm_nt

    def recording(A, B, M, N, K, *a, **kw):
        calls.append((M, N, K, kw.get("splits")))
        return gemm(A, B, M, N, K, *a, **kw)

    nn.gemm_nt = recording
    try:
        step()
    finally:
        nn.gemm_nt = gemm
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    ms = 1e3 * (time.perf_counter() - t0) / steps
    return ms, calls, sum(2.0 * M * N * K for M, N, K, _ in calls)


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser(description="time surrogate training steps")
    ap.add_argument("--grid", type=int, nargs=3, default=(32, 32, 24), metavar=("NX", "NY", "NZ"))
    ap.add_argument("--base", type=int, default=32)
    ap.add_argument("--batch", type=int, default=4)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--config4", type=int, nargs=2, default=None, metavar=("NTRAIN", "NTEST"),
                    help="train on NTRAIN solved models and compare warm and cold solves of NTEST held-out ones")
    ap.add_argument("--train-steps", type=int, default=4000)
    ap.add_argument("--gemm-shapes", type=int, default=0, metavar="N",
                    help="also time the step's N largest GEMM shapes on their own (10 launches each)")
    a = ap.parse_args()
    if a.config4:
        import json
        out = config4(a.config4[0], a.config4[1], tuple(a.grid), a.base, steps=a.train_steps, batch=a.batch,
                      printer=lambda m: print(m, flush=True))
        print(json.dumps(out), flush=True)
        raise SystemExit(0)
    ms, calls, flops = time_training_step(tuple(a.grid), a.base, a.batch, a.steps)
    print(f"{len(calls)} GEMM calls per step, {flops / 1e9:.1f} GFLOP; {ms:.2f} ms per training step, "
          f"{flops / ms / 1e9:.1f} TFLOP/s of GEMM work over the whole step", flush=True)
    if a.gemm_shapes:
        import torch
        for M, N, K, splits in sorted(set(calls), key=lambda c: -c[0] * c[1] * c[2])[:a.gemm_shapes]:
            A = torch.randn(M, K, device="cuda").to(torch.bfloat16).view(torch.int16)
            B = torch.randn(N, K, device="cuda").to(torch.bfloat16).view(torch.int16)
            C = torch.empty(M, N, device="cuda")
            for _ in range(3):
                nn.gemm_nt(A, B, M, N, K, C_out=C, splits=splits)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(10):
                nn.gemm_nt(A, B, M, N, K, C_out=C, splits=splits)
            e1.record()
            torch.cuda.synchronize()
            us = 1e3 * e0.elapsed_time(e1) / 10
            io = 2.0 * (M * K + N * K) + 4.0 * M * N
            print(f"  M={M} N={N} K={K} (splits {splits or 'auto'}): {us:.1f} us, {2.0 * M * N * K / us / 1e6:.1f} TFLOP/s, "
                  f"{io / us / 1e3:.0f} GB/s of operand + output bytes", flush=True)
