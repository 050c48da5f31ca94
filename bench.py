"""Benchmark: edge-FEM solves/s to a 1e-8 relative residual on one MI355X.

Workload (the 1.4M-edge scale-up case): 64x64x48 cells of 2.5 km (1,179,648 tets, 1,407,152
edges), crust log10 sigma ~ N(-2, 0.5) per cell plus 3 spheres of 1 S/m (seed 1002), sea
3.3 S/m, air 1e-8 S/m, tidal v x F source at M2, tangential E = 0, one right-hand side.
Mesh, pattern, assembly, boundary conditions and the preconditioner setup happen before the
timed region; inputs are resident in HBM.

A step is one solve from x0 = 0 to an FP64 true residual ||b - A x||/||b|| <= 1e-8 (recomputed
on the device at the end) with FP64 COCG (the reference's recurrence, COCG13_single.f90:181-323)
preconditioned by --precond, by default NODEBLK+GRAD: node-block Jacobi (the inverse of A's
diagonal block over each node's consecutive edges, afx.h AFX_PC_NODEBLK) plus the nodal
gradient-space correction G diag(G^T A G)^-1 G^T (AFX_PC_GRAD). The whole iteration runs in one
persistent launch per 2000 iterations. The timed loop runs --steps solves or stops after the
step that passes --time-budget seconds; `steps` reports the solves timed.

Beside it, the comparator: plain FP64 SSOR-COCG in the reference's row order (the preconditioner
the reference's solveCG passes; its 'IC' stagnates on this system), one solve, and the other
corrected forms (SSOR+GRAD, SSOR_BJ+GRAD, DIAG+GRAD), one solve each; "variants" holds their solves/s, the
speedup of the measured form over plain SSOR, and how far the two solutions' seafloor site fields
differ.

After the solves, the SpMV q = A p runs as a kernel of its own (k_spmv_stream, 50 timed launches
on the solution vector); the line's "spmv" object gives its algorithmic GB/s against the HBM peak.

By default the line also carries the batched forms (skipped with --no-variants): the 5 tidal
constituents of this mesh (one column at a time, the k-column launch, and a warm chain where each
constituent starts from the previous one's solution); solvePARDISOinv's 48 adjoint columns
(16 seafloor sites x 3 electric components) on the 32x32x24-cell mesh, one column at a time
against the autotuned block width; the 16-mesh uneven batch one mesh at a time against one
per-system launch; 3 Gauss-Newton updates on the 32x32x24-cell mesh (5 constituents, 16 sites x
3 magnetic components: 5 forward and 240 adjoint solves, the Jacobian of every crust cell and the
model update per update), with every solve warm-started from the previous update's solution
against every solve started from zero (--gn-updates N with --no-variants runs only that).

Variants, off by default (their results are documented):
  --mixed    complex-FP32 preconditioner inside FP64 refinement (same tolerance);
  --batched  the 5 tidal constituents (M2, S2, N2, K1, O1) in one call, per-column omega (the
             autotuner picks one column at a time or the 5-column kernel);
  --rhs-stack  the same 5 constituents as 5 copies of the system (K + i omega_c M each) stacked
             block-diagonally and solved with per-system scalars in one launch, against the
             autotuned k-column call;
  --adjoint-batch SITES  SITES seeded seafloor sites x 3 electric components as adjoint
             right-hand sides (solvePARDISOinv's P^T columns) at M2: one column at a time against
             the autotuned block width (stacked copies in groups where they fit LDS);
  --mesh-batch  (also in the default line, per-system only) 16 independent meshes from 8x8x6 to
             40x40x30 cells (1,830,720 edges in all), each
             with its own seeded crust, at M2: one solve per mesh against the stacked
             block-diagonal system in one launch (actfemtide_amd.meshbatch), with per-system and
             with shared COCG scalars, every mesh to a 1e-8 true residual each way.

`python bench.py [--gpus 1] [--steps K] [--warmup W] [--size cfg2|cfg3|small] [--mixed] [--batched]
[--mesh-batch]`
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
SIZES = {"cfg2": (64, 64, 48), "cfg3": (32, 32, 24), "small": (16, 16, 12)}
CONSTITUENTS = ["M2", "S2", "N2", "K1", "O1"]
# the uneven batch: 16 boxes whose side grows from 8 to 40 cells, depth 3/4 of the side
MESH_BATCH = [(8 + round(32 * i / 15), 8 + round(32 * i / 15), round(0.75 * (8 + round(32 * i / 15))))
              for i in range(16)]


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def bytes_per_iteration(n, nU, nL, k, z_bytes):
    """Compulsory HBM bytes of one COCG iteration, every array touched once:
    SpMV: (K, M) values 16 B + column 4 B per entry, row pointers, p read, q written (FP64);
    two sweeps: values 16 B + column 4 B per strict-triangle entry, position metadata, r read,
    DD read, z written (z in the preconditioner precision);
    vector phase: p update (z, p read, p written), x/r update (x, p, r, q read; x, r written)."""
    nnz = n + nU + nL
    spmv = nnz * 20 + 4 * (n + 1) + 2 * 16 * n * k
    sweeps = (nU + nL) * 20 + 2 * 8 * n + 2 * (16 * n * k + 16 * n + z_bytes * n * k)
    vec = (z_bytes + 16 + 16) * n * k + 6 * 16 * n * k
    return spmv + sweeps + vec


def spmv_bytes(n, nnz):
    """Compulsory HBM bytes of one SpMV over the full CSR: (K, M) values 16 B and a column 4 B per
    entry, the row pointers, x read and y written (16 B per row each)."""
    return nnz * 20 + 4 * (n + 1) + 2 * 16 * n


def measured_traffic(workload, iterations_per_launch):
    """HBM bytes per launch of k_cocg from the committed rocprofv3 --pmc summary of this workload
    (FETCH_SIZE and WRITE_SIZE per COCG iteration, scaled by the counters' reading for a copy of
    known size), or None when no summary exists."""
    path = os.path.join(ROOT, "profiles", f"k_cocg_{workload}_pmc.json")
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        m = json.load(f)
    return int(m["hbm_bytes_per_iteration"] * iterations_per_launch), os.path.relpath(path, ROOT)


def bytes_per_iteration_grad(n, nnz, nn, nadj, edge_pc, nU=0, nL=0):
    """Compulsory HBM bytes of one COCG iteration with the nodal gradient-space correction, every
    array touched once per pass:
    SpMV: (K, M) values 16 B + column 4 B per entry, row pointers, p read, q written;
    edge part of z: DIAG reads r and DD and writes z; SSOR / SSOR_BJ as in bytes_per_iteration;
    nodal part: the incidence list (4 B per entry), r read, diag(G^T M G) read, zn written;
    p update: z, p read, p written, the edges' end nodes (8 B per edge), zn read;
    x, r update: x, p, q, r read, x and r written."""
    spmv = nnz * 20 + 4 * (n + 1) + 2 * 16 * n
    if edge_pc == "DIAG":
        edge = 3 * 16 * n
    else:
        edge = (nU + nL) * 20 + 2 * 8 * n + 2 * (16 * n + 16 * n + 16 * n)
    nodal = 4 * nadj + 16 * n + 8 * nn + 16 * nn
    pupd = 3 * 16 * n + 8 * n + 16 * nn
    xr = 6 * 16 * n
    return spmv + edge + nodal + pupd + xr


def bytes_per_iteration_fused(n, nnz, nn, nadj, edge_pc, ntile=0):
    """Compulsory HBM bytes of one iteration of the fused single-column kernel (k_cocg1: DIAG,
    NONE or NODEBLK edge preconditioner with the nodal correction), every array touched once per
    pass:
    p pass: r (or, NODEBLK, Bi r), D (DIAG), p read, p written, the edges' end nodes (8 B per
    edge), zn read;
    SpMV: (K, M) values 16 B + column 4 B per entry, row pointers, p read, q written;
    x, r, z pass: x, p, q, r, D (DIAG) read, x and the next r written; NODEBLK: 8 complex-bfloat16
    coefficients (32 B) and 1 B of block place per row, the tile bounds, Bi r written; on the
    nodes the incidence list (4 B per entry; NODEBLK: its 16-slot copy, 64 B per node), the row
    pointers, diag(G^T M G) read and zn written (the r and q, or NODEBLK's r', gathered there are
    arrays the edge part reads or writes in the same pass)."""
    d = 16 * n if edge_pc == "DIAG" else 0
    ppass = 3 * 16 * n + d + 8 * n + 16 * nn
    spmv = nnz * 20 + 4 * (n + 1) + 2 * 16 * n
    lists = 64 * nn if edge_pc == "NODEBLK" else 4 * nadj
    xn = 6 * 16 * n + d + lists + 4 * (nn + 1) + 8 * nn + 16 * nn
    if edge_pc == "NODEBLK":
        xn += 32 * n + n + 4 * (ntile + 1) + 16 * n
    return ppass + spmv + xn


def timed_solves(s, dB, dX, opts, steps, budget, label):
    """Solve from x0 = 0 `steps` times (stopping after the step that passes `budget` seconds);
    returns (steps, seconds, last stats, (launches, iterations, kernel ms))."""
    import torch
    s.kernel_timing(reset=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    done, st = 0, None
    for _ in range(steps):
        dX.zero_()
        st, rc = s.solve_device(dB, dX, **opts)
        done += 1
        log(f"{label} step {done}: {st[0].iter} iterations, rc {rc}, {time.perf_counter() - t0:.2f}s")
        if time.perf_counter() - t0 > budget:
            break
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    return done, dt, st[0], s.kernel_timing(reset=True)


def site_fields(s, item, ve, vb, dX, omega):
    from actfemtide_amd import afx
    return afx.calobsebcomp(s, item, ve, vb, dX, omega)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--size", default="cfg2", choices=sorted(SIZES))
    ap.add_argument("--precond", default="NODEBLK+GRAD",
                    help="preconditioner of the measured solve (afx.PC); the plain FP64 SSOR-COCG of the "
                         "reference's order is always timed beside it as the comparator")
    ap.add_argument("--plain-steps", type=int, default=1, help="timed solves of the plain SSOR comparator (0: skip)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--mixed", action="store_true", help="also time the complex-FP32-preconditioner solve")
    ap.add_argument("--batched", action="store_true", help="also time the 5 constituents in one launch")
    ap.add_argument("--jacobian", type=int, default=0, metavar="NOBS",
                    help="also time genjacobian1's accumulation over every crust cell for NOBS adjoint columns")
    ap.add_argument("--no-variants", action="store_true",
                    help="skip the batched forms (5 constituents; 48 adjoint columns on 32x32x24 cells; the "
                         "16-mesh batch)")
    ap.add_argument("--adjoint-batch", type=int, default=0, metavar="SITES",
                    help="also time SITES seafloor sites x 3 electric adjoint right-hand sides, one column at a "
                         "time against the autotuned block width")
    ap.add_argument("--rhs-stack", action="store_true",
                    help="also time the 5 constituents stacked as 5 systems in one launch")
    ap.add_argument("--mesh-batch", action="store_true",
                    help="also time 16 meshes of 8x8x6 to 40x40x30 cells one by one and in one launch")
    ap.add_argument("--gn-updates", type=int, default=0, metavar="N",
                    help="with --no-variants: time N warm- and cold-started Gauss-Newton updates on 32x32x24")
    ap.add_argument("--maxit", type=int, default=400000, help="iteration cap of each solve")
    ap.add_argument("--time-budget", type=float, default=60.0,
                    help="seconds of timed solves: the timed loop stops after the first step that ends past it "
                         "(at least one step); `steps` in the line reports the steps actually timed")
    ap.add_argument("--warmup-iters", type=int, default=50,
                    help="iterations of each warmup solve (one persistent launch; a warmup need not converge)")
    ap.add_argument("--spmv-launches", type=int, default=50, help="timed launches of the standalone SpMV")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="after the solves, copy this many bytes device-to-device (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: each system is solved on a single MI355X; several GPUs would "
              "each own whole constituents / right-hand-side blocks (no figure is measured for N > 1)",
              file=sys.stderr)
        sys.exit(2)
    os.environ.setdefault("AFX_VERBOSE", "1")
    t_start = time.time()

    import torch

    from actfemtide_amd import afx, mesh

    dev = torch.device("cuda:0")
    nx, ny, nz = SIZES[args.size]
    t0 = time.time()
    case = mesh.tide_case(nx, ny, nz, seed=1002, spheres=3, sigma_sea=3.3)
    log(f"mesh {nx}x{ny}x{nz}: ntet={case.ntet} nline={case.nline} ({time.time() - t0:.1f}s)")
    omegas = np.array([2 * math.pi / mesh.PERIODS[c] for c in CONSTITUENTS])
    w0 = omegas[0]
    t0 = time.time()
    s, b, bc = afx.forward_system(case, w0)
    t_system = time.time() - t0
    log(f"system: n={s.n} nU={s.npu} levels L/U={s.nlev_l}/{s.nlev_u} ({t_system:.1f}s)")
    n = s.n
    nnz = n + s.npu + s.npl
    dB = torch.from_numpy(b.astype(np.complex128)).to(dev)
    pc = args.precond
    opts = dict(precond=pc, tol=1e-8, maxit=args.maxit, chunk=2000, precision=afx.FP64)
    s.precond_setup(pc)
    s.autotune(1)
    torch.cuda.synchronize()
    dX = torch.zeros(n, dtype=torch.complex128, device=dev)
    # warmup: short solves of one persistent launch each (the kernels, the value gathers and the
    # work buffers are exercised; a warmup need not converge)
    for i in range(args.warmup):
        dX.zero_()
        s.solve_device(dB, dX, **dict(opts, maxit=args.warmup_iters))
    steps, t_run, st_main, (launches, iters, kms) = timed_solves(s, dB, dX, opts, args.steps, args.time_budget, pc)
    true_res = float(s.true_residual_device(dB, dX)[0])
    value = steps / t_run
    its = st_main.iter
    log(f"{pc}: {steps} solve(s) in {t_run:.3f}s, {its} iterations, true residual {true_res:.2e}; "
        f"k_cocg {launches} launches {kms:.1f} ms")

    # roofline of the persistent COCG kernel: algorithmic bytes of the iterations each launch ran
    # over the launches' measured durations (HIP events on the library's stream)
    base = pc.split("+")[0]
    if pc.endswith("+GRAD"):
        ln = np.asarray(case.line) - 1
        bnd = np.zeros(case.node, bool)
        bnd[ln[bc].ravel()] = True
        nadj = int(np.count_nonzero(~bnd[ln]))  # (edge, interior end node) incidences
        nn = int(np.count_nonzero(~bnd))
        if base in ("DIAG", "NONE", "NODEBLK"):
            ntile = -(-n // 249) if base == "NODEBLK" else 0  # tiles hold at least 249 rows
            by = bytes_per_iteration_fused(n, nnz, nn, nadj, base, ntile)
        else:
            by = bytes_per_iteration_grad(n, nnz, nn, nadj, base, s.npu, s.npl)
    else:
        by = bytes_per_iteration(n, s.npu, s.npl, 1, 16)
    achieved = by * iters / (kms / 1e3) / 1e9
    ipl = iters / max(launches, 1)
    traffic, tsrc = measured_traffic(f"{nx}x{ny}x{nz}_{pc.replace('+', '_')}", ipl)
    roof = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 4), traffic=traffic,
                kernel=(f"k_cocg1 (fused persistent COCG without sweeps, FP64, 1 column, {pc})"
                        if base in ("DIAG", "NONE", "NODEBLK") else f"k_cocg (persistent COCG, FP64, 1 column, {pc})"),
                bytes_per_iteration=by, avg_launch_ms=round(kms / max(launches, 1), 3),
                iterations_per_launch=round(ipl, 1), us_per_iteration=round(1e3 * kms / max(iters, 1), 1),
                traffic_source=tsrc)

    # observation-site fields of the measured solution (96 seafloor sites, B and E)
    item, ve, vb, _ = s.prepobscoeff(seafloor_sites(case, 96, 31))
    f_main = site_fields(s, item, ve, vb, dX, w0)

    # the comparator: plain FP64 SSOR-COCG in the reference's order (solveCG's preconditioner)
    variants = {}
    plain = None
    if args.plain_steps > 0 and pc != "SSOR":
        s.precond_setup("SSOR")
        s.autotune(1)
        dXp = torch.zeros_like(dX)
        popts = dict(opts, precond="SSOR")
        psteps, pt, pst, (pla, pit, pms) = timed_solves(s, dB, dXp, popts, args.plain_steps, 1.0, "plain SSOR")
        f_plain = site_fields(s, item, ve, vb, dXp, w0)
        ferr = float(max(np.abs(f_main[k] - f_plain[k]).max() / np.abs(f_plain[k]).max() for k in range(5)))
        plain = dict(solves_per_s=round(psteps / pt, 6), seconds_per_solve=round(pt / psteps, 3), iters=pst.iter,
                     true_resid=float(s.true_residual_device(dB, dXp)[0]),
                     us_per_iteration=round(1e3 * pms / max(pit, 1), 1),
                     note="FP64 SSOR-COCG in the reference's row order (COCG13_single with solveCG's 'SSOR')")
        variants["plain_fp64_ssor"] = plain
        variants["speedup_vs_plain"] = round(value / plain["solves_per_s"], 2)
        variants["site_fields_vs_plain"] = dict(worst_rel_diff=ferr, sites=96,
                                                note="B (3 components) and E (x, y) at seafloor sites of the "
                                                     "measured solution against the plain solution, both at a "
                                                     "1e-8 true residual")
        log(f"plain SSOR: {pt / psteps:.2f}s per solve, {pst.iter} iterations; speedup "
            f"{variants['speedup_vs_plain']}x; site fields differ by {ferr:.1e}")
        del dXp
    # the other forms of the nodal correction, one solve each
    for other in ("SSOR+GRAD", "SSOR_BJ+GRAD", "DIAG+GRAD", "NODEBLK+GRAD"):
        if other == pc or time.time() - t_start > 300:
            continue
        s.precond_setup(other)
        s.autotune(1)
        dXo = torch.zeros_like(dX)
        ostep, ot, ost, (ola, oit, oms) = timed_solves(s, dB, dXo, dict(opts, precond=other), 1, 1.0, other)
        variants[other] = dict(solves_per_s=round(ostep / ot, 4), seconds=round(ot, 4), iters=ost.iter,
                               us_per_iteration=round(1e3 * oms / max(oit, 1), 1),
                               true_resid=float(s.true_residual_device(dB, dXo)[0]))
        del dXo
    s.precond_setup(pc)
    s.autotune(1)

    # the SpMV (q = A p) as a kernel of its own, on the solution vector: algorithmic bytes of one
    # product over the average launch duration (HIP events on the library's stream)
    dQ = torch.empty_like(dX)
    for _ in range(5):
        s.spmv_device(dX, dQ, timed=True)
    sp_ms = [s.spmv_device(dX, dQ, timed=True) for _ in range(args.spmv_launches)]
    sp_bytes = spmv_bytes(n, nnz)
    sp_avg = sum(sp_ms) / len(sp_ms)
    spmv = dict(kernel="k_spmv_stream", bound="hbm", achieved=round(sp_bytes / (sp_avg / 1e3) / 1e9, 1),
                peak=HBM_PEAK_GBS, unit="GB/s", frac=round(sp_bytes / (sp_avg / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
                bytes_per_launch=sp_bytes, avg_launch_us=round(1e3 * sp_avg, 2),
                min_launch_us=round(1e3 * min(sp_ms), 2), launches=len(sp_ms))
    log(f"SpMV: {1e3 * sp_avg:.1f} us per launch, {spmv['achieved']:.0f} GB/s ({100 * spmv['frac']:.1f}% of peak)")

    if args.mixed:
        dXm = torch.zeros(n, dtype=torch.complex128, device=dev)
        torch.cuda.synchronize()
        tm = time.perf_counter()
        stm, rcm = s.solve_device(dB, dXm, **dict(opts, precision=afx.MIXED, max_outer=8))
        torch.cuda.synchronize()
        tm = time.perf_counter() - tm
        variants["mixed_fp32_preconditioner"] = dict(
            solves_per_s=round(1.0 / tm, 6), seconds=round(tm, 2), iters=stm[0].iter, outer=stm[0].outer,
            true_resid=float(s.true_residual_device(dB, dXm)[0]), rc=rcm)
        log(f"mixed: {tm:.2f}s, {stm[0].iter} iterations")
    if args.batched:
        k = len(CONSTITUENTS)
        s.autotune(k)
        dBk = dB.reshape(n, 1).repeat(1, k).contiguous()
        dXk = torch.zeros((n, k), dtype=torch.complex128, device=dev)
        torch.cuda.synchronize()
        tb = time.perf_counter()
        stb, rcb = s.solve_device(dBk, dXk, omega_scale=omegas / w0, **opts)
        torch.cuda.synchronize()
        tb = time.perf_counter() - tb
        variants["batched_5_constituents"] = dict(
            solves_per_s=round(k / tb, 6), seconds=round(tb, 2), iters=[x.iter for x in stb],
            true_resid_max=float(s.true_residual_device(dBk, dXk, omega_scale=omegas / w0).max()))
        log(f"batched: {k} solves in {tb:.2f}s, iterations {[x.iter for x in stb]}")

    if args.jacobian:
        variants[f"jacobian_{args.jacobian}_columns"] = jacobian_variant(s, case, w0, args.jacobian)
    if args.adjoint_batch:
        variants[f"adjoint_{3 * args.adjoint_batch}_columns"] = adjoint_batch_variant(s, case, opts, args.adjoint_batch)
    if not args.no_variants and time.time() - t_start < 300:
        # the batched forms (a failure here is recorded, it does not end the run):
        # the 5 constituents of this mesh; solvePARDISOinv's 48 adjoint columns (16 seafloor sites
        # x 3 electric components) on the 32x32x24-cell mesh; the 16-mesh uneven batch
        for name, fn in (("constituents_5", lambda: constituents_variant(s, b, omegas, w0, opts)),
                         ("adjoint_48_columns_32x32x24", lambda: adjoint32(pc, opts, w0)),
                         ("mesh_batch_16", lambda: mesh_batch_variant(opts, w0, grad_only=True)),
                         ("gauss_newton_warm_starts_32x32x24", lambda: gn_variant())):
            try:
                variants[name] = fn()
            except Exception as e:  # noqa: BLE001
                variants[name] = {"error": f"{type(e).__name__}: {e}"}
            os.environ.pop("AFX_BATCH", None)
    if args.gn_updates and args.no_variants:
        variants["gauss_newton_warm_starts_32x32x24"] = gn_variant(updates=args.gn_updates)
    if args.rhs_stack:
        variants["stacked_5_constituents"] = rhs_stack_variant(s, b, omegas, w0, opts)
    if args.mesh_batch:
        variants["mesh_batch_16_all_modes"] = mesh_batch_variant(opts, w0)

    if args.copy_bytes:
        src = torch.empty(args.copy_bytes // 8, dtype=torch.float64, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"calibration copy of {args.copy_bytes} bytes done")

    # end to end for one more frequency (S2) on the same pattern: GENMAT + COPY_UL, GENBCCSEM +
    # SET_BC_ICCG, the preconditioner setup and the value gathers (a 1-iteration solve builds
    # them), i.e. the per-frequency work forward_bxyz repeats, plus the measured solve time
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b2 = s.assemble(case.sigma_tet(), omegas[1])
    b2 = s.set_bc(bc, b2)
    s.precond_setup(pc)
    dB2 = torch.from_numpy(b2).to(dev)
    dX2 = torch.zeros_like(dB2)
    s.solve_device(dB2, dX2, **dict(opts, maxit=1, tol=0.0))
    torch.cuda.synchronize()
    t_freq = time.perf_counter() - t0
    end_to_end = dict(solves_per_s=round(1.0 / (t_freq + t_run / steps), 6),
                      setup_per_frequency_s=round(t_freq, 4), first_system_s=round(t_system, 2),
                      note="per frequency: assembly, boundary conditions, preconditioner setup and value "
                           "gathers on the device, added to the measured solve; first_system_s (pattern "
                           "analysis, mesh upload, first assembly) is paid once per mesh")
    log(f"end to end: {t_freq * 1e3:.1f} ms setup per frequency, {end_to_end['solves_per_s']} solves/s")
    s.close()

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(case, w0, plain["iters"] if plain else its, bc=bc,
                           grad_iters=its if pc in ("DIAG+GRAD", "NODEBLK+GRAD") else 0, grad_edge=pc.split("+")[0])

    out = {
        "metric": "edge-FEM solves/s to 1e-8 rel. residual, warm-start mixed vs plain COCG, 1 GPU",
        "value": round(value, 6),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * t_run / steps, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128",
        "data": "synthetic",
        "config": {"workload": f"{nx}x{ny}x{nz} cells ({case.ntet} tets, {case.nline} edges), one M2 right-hand "
                               f"side, FP64 COCG preconditioned by {pc} from x0 = 0 to a 1e-8 FP64 true residual",
                   "edges": case.nline, "tets": case.ntet, "rhs_per_step": 1, "precond": pc,
                   "parallelism": "1 GPU"},
        "iters": its,
        "true_residual": true_res,
        "roofline": roof,
        "spmv": spmv,
        "end_to_end": end_to_end,
        "cpu_baseline": cpu,
        "variants": variants or None,
    }
    print(json.dumps(out), flush=True)


def jacobian_variant(s, case, omega, nobs, reps=3):
    """genjacobian1's accumulation (afx_jacobian_device: k_jac_w, then k_jac_acc, one block per
    parameter) with one parameter per crust cell (6 tets) and nobs adjoint columns of synthetic
    values (the kernel's work does not depend on them). Wall time per call includes the upload of
    the parameter map; the kernel's own duration comes from a rocprofv3 trace."""
    import torch

    crust = np.nonzero(case.group >= 3)[0]
    cells, inv = np.unique(case.cell_of_tet[crust], return_inverse=True)
    order = np.argsort(inv, kind="stable")
    tets = crust[order].astype(np.int32)
    ptr = np.concatenate([[0], np.cumsum(np.bincount(inv))]).astype(np.int32)
    npar = len(cells)
    g = torch.Generator(device="cuda").manual_seed(5)
    dU = torch.randn((s.n, nobs), dtype=torch.complex128, device="cuda", generator=g)
    dx = torch.randn((s.n,), dtype=torch.complex128, device="cuda", generator=g)
    dJ = torch.empty((npar, nobs), dtype=torch.complex128, device="cuda")
    sig = case.sigma_tet()
    edges = np.unique(np.abs(np.asarray(case.n6line)[tets]).ravel())
    # compulsory bytes of k_jac_acc: the U rows of the cells' edges once, J written once, the
    # per-tet weights (6 complex) and edge ids (6 int) once
    acc_bytes = len(edges) * nobs * 16 + npar * nobs * 16 + len(tets) * 6 * (16 + 4) + (npar + 1) * 4
    times = []
    for _ in range(reps + 1):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        s.jacobian_device(dU, dx, ptr, tets, sig, omega, -math.log(10.0), dJ)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    best = min(times[1:])
    log(f"jacobian: {npar} parameters x {nobs} columns, {1e3 * best:.1f} ms per call (wall), "
        f"k_jac_acc compulsory bytes {acc_bytes / 1e9:.2f} GB")
    return dict(parameters=npar, columns=nobs, tets=len(tets), edges=len(edges), wall_ms=round(1e3 * best, 2),
                acc_bytes=acc_bytes, note="kernel durations: profiles/k_jac_acc_64x64x48.md")


def seafloor_sites(case, n, seed):
    """Seeded sites 1 m below the sea-bottom plane (z = -h), strictly inside a tet."""
    rng = np.random.default_rng(seed)
    h = case.h
    xout, yout = case.nx * h / 2, case.ny * h / 2
    pts = []
    for _ in range(n):
        i = rng.integers(1, case.nx - 1)
        j = rng.integers(1, case.ny - 1)
        fx, fy = rng.uniform(0.13, 0.87, size=2)
        if abs(fx - fy) < 0.05:
            fy = min(0.87, fy + 0.1)
        pts.append([-xout + (i + fx) * h, -yout + (j + fy) * h, -h - 0.001])
    return np.array(pts)


def gn_variant(nsite=16, updates=3, grid=None):
    """Gauss-Newton updates (actfemtide_amd.inversion) on the 32x32x24-cell mesh with the 5 tidal
    constituents: per update 5 forward solves and 5 x 3 nsite adjoint solves (Bx, By, Bz rows at
    nsite sites 1 m above the sea bottom), the Jacobian of every crust cell, the model update.
    Run twice on the same synthetic data (a conductive block): every solve warm-started from the
    previous update's solution, and every solve from zero. Reports per update the solve seconds
    (synchronised), iterations and nRMS, and the warm run's speedup on the solves of updates 2..N."""
    import torch

    from actfemtide_amd import inversion, mesh
    case = mesh.tide_case(*(grid or SIZES["cfg3"]), seed=1003, sigma_sea=3.3)
    sites = inversion.central_sites(case, nsite, 5, half=4)
    omegas = [2 * math.pi / mesh.PERIODS[c] for c in CONSTITUENTS]
    ncell = len(inversion.crust_parameters(case)[2])
    out = dict(mesh="x".join(map(str, grid or SIZES["cfg3"])), sites=nsite, constituents=len(omegas),
               adjoint_columns_per_update=3 * nsite * len(omegas), parameters=ncell, updates=updates)
    for warm in (True, False):
        gn = inversion.GaussNewton(case, sites, omegas, np.full(ncell, -2.0), beta=10.0, gamma=0.1, warm=warm)
        m_true, _ = inversion.block_model(case, gn.cells, -1.5, -2.0, half=3)
        F, _, _ = gn.forward(m_true)
        gn.d_obs = gn.data(F)
        gn.dX.zero_()
        ups, _, nr = gn.run(gn.m_ref, updates=updates)
        torch.cuda.synchronize()
        out["warm" if warm else "cold"] = dict(
            nrms=[round(u.nrms, 4) for u in ups] + [round(nr, 4)],
            solve_seconds=[round(u.solve_seconds, 3) for u in ups],
            update_seconds=[round(u.seconds, 3) for u in ups],
            forward_iters=[sum(u.forward_iters) for u in ups],
            adjoint_iters=[int(sum(map(sum, u.adjoint_iters))) for u in ups],
            worst_true_resid=max(u.worst_true_resid for u in ups))
        gn.close()
        log(f"gauss-newton ({'warm' if warm else 'cold'}): {out['warm' if warm else 'cold']}")
    w, c = out["warm"], out["cold"]
    out["speedup_solves_after_first"] = round(sum(c["solve_seconds"][1:]) / sum(w["solve_seconds"][1:]), 3)
    out["note"] = ("synthetic inversion of a conductive block (log10 sigma -1.5 in -2, 7x7x3 cells) from Bx, By, "
                   "Bz amplitude and phase; warm: each "
                   "update's solves start from the previous update's solutions; cold: from zero; both reach "
                   "a 1e-8 FP64 true residual in every solve")
    return out


def adjoint32(pc, opts, omega):
    """adjoint_batch_variant with 16 sites (48 columns) on the 32x32x24-cell mesh."""
    from actfemtide_amd import afx, mesh
    c3 = mesh.tide_case(*SIZES["cfg3"], seed=1002, spheres=3, sigma_sea=3.3)
    s3, _, _ = afx.forward_system(c3, omega)
    s3.precond_setup(pc)
    s3.autotune(1)
    out = adjoint_batch_variant(s3, c3, opts, 16)
    s3.close()
    return out


def adjoint_batch_variant(s, case, opts, nsites):
    """solvePARDISOinv's adjoint right-hand sides (P^T columns of the 3 electric components at
    nsites seafloor sites, seed 21) solved one column at a time and with the autotuned block
    width. Solves only, inputs on the device."""
    import torch

    item, ve, vb, _ = s.prepobscoeff(seafloor_sites(case, nsites, 21))
    stack = np.arange(0, 6 * nsites + 1, 6, dtype=np.int32)
    cols = []
    for comp in range(3):
        d = torch.zeros((s.n, nsites), dtype=torch.complex128, device="cuda")
        s.adjoint_rhs_device(stack, item, ve[comp], d)
        cols.append(d)
    dB = torch.cat(cols, dim=1).contiguous()
    k = dB.shape[1]
    out = {}
    for width in ("columns", "autotuned"):
        if width == "columns":
            os.environ["AFX_BATCH"] = "columns"
        else:
            os.environ.pop("AFX_BATCH", None)
            s.autotune(k)
        dX = torch.zeros_like(dB)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st, rc = s.solve_device(dB, dX, **opts)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        # the solver's own FP64 check of each column (the stopping guarantee), and the same residual
        # recomputed over the k-column layout (another summation order: within ~1e-11 of it)
        tr = max(x.true_resid for x in st)
        tr2 = float(s.true_residual_device(dB, dX).max())
        log(f"{k} adjoint columns, {width}: {dt:.3f}s, iterations {min(x.iter for x in st)}..{max(x.iter for x in st)}, worst true residual "
            f"{tr:.3e} (recomputed {tr2:.3e})")
        its = [x.iter for x in st]
        out[width] = dict(solves_per_s=round(k / dt, 4), seconds=round(dt, 3), iters_min=min(its), iters_max=max(its),
                          true_resid_max=tr, true_resid_max_recomputed=tr2, rc=rc)
    out["speedup"] = round(out["columns"]["seconds"] / out["autotuned"]["seconds"], 3)
    return out


def constituents_variant(s, b, omegas, w0, opts):
    """The 5 tidal constituents (M2, S2, N2, K1, O1) of one mesh, each column at its own omega:
    one column at a time from x0 = 0; the k-column launch (AFX_BATCH=multi); and one at a time with
    each constituent warm-started from the previous constituent's solution (the right-hand side is
    the same source, so the fields of neighbouring frequencies are close). Solves only, inputs on
    the device; every column to a 1e-8 FP64 true residual."""
    import torch

    k = len(omegas)
    sc = omegas / w0
    dB = torch.from_numpy(np.ascontiguousarray(np.tile(b[:, None], (1, k)))).cuda()
    out = {}
    for width in ("columns", "multi"):
        os.environ["AFX_BATCH"] = width
        wopts = opts
        dX = torch.zeros_like(dB)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st, rc = s.solve_device(dB, dX, omega_scale=sc, **wopts)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        tr = float(s.true_residual_device(dB, dX, omega_scale=sc).max())
        out[width] = dict(solves_per_s=round(k / dt, 4), seconds=round(dt, 4), iters=[x.iter for x in st],
                          true_resid_max=tr, rc=rc, precond=wopts["precond"])
        log(f"5 constituents, {width}: {dt:.3f}s, iterations {[x.iter for x in st]}, worst {tr:.2e}")
    os.environ.pop("AFX_BATCH", None)
    # warm chain: column c starts from column c - 1's solution
    dX = torch.zeros_like(dB)
    its = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for c in range(k):
        x = (dX[:, c - 1] if c else torch.zeros_like(dB[:, 0])).contiguous()
        bc_ = dB[:, c].contiguous()
        st, rc = s.solve_device(bc_, x, omega_scale=np.array([sc[c]]), **opts)
        dX[:, c] = x
        its.append(st[0].iter)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    tr = float(s.true_residual_device(dB, dX.contiguous(), omega_scale=sc).max())
    out["warm_chain"] = dict(solves_per_s=round(k / dt, 4), seconds=round(dt, 4), iters=its, true_resid_max=tr)
    log(f"5 constituents, warm chain: {dt:.3f}s, iterations {its}, worst {tr:.2e}")
    best = min(out[m]["seconds"] for m in ("multi", "warm_chain"))
    out["speedup_vs_columns"] = round(out["columns"]["seconds"] / best, 3)
    return out


def rhs_stack_variant(s, b, omegas, w0, opts):
    """The 5 constituents of one mesh: the autotuned 5-column call (one column at a time or the
    5-column kernel) against the system stacked once per constituent (MeshBatch.frequency_stack)
    in one launch with per-system scalars. Solves only, inputs on the device."""
    import torch

    from actfemtide_amd.meshbatch import MeshBatch

    k, n = len(omegas), s.n
    s.autotune(k)
    dBk = torch.from_numpy(np.ascontiguousarray(np.tile(b[:, None], (1, k)))).cuda()
    dXk = torch.zeros_like(dBk)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st, rc = s.solve_device(dBk, dXk, omega_scale=omegas / w0, **opts)
    torch.cuda.synchronize()
    t_cols = time.perf_counter() - t0
    tr_cols = float(s.true_residual_device(dBk, dXk, omega_scale=omegas / w0).max())
    log(f"5 constituents, autotuned call: {t_cols:.2f}s, iterations {[x.iter for x in st]}")
    t0 = time.time()
    mb = MeshBatch.frequency_stack(s, omegas)
    mb.system.precond_setup("SSOR")
    mb.system.autotune(1)
    log(f"stacked constituents: n={mb.n} ({time.time() - t0:.1f}s)")
    dB = torch.from_numpy(np.tile(b, k)).cuda()
    dX = torch.zeros_like(dB)
    mb.system.kernel_timing(reset=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st2, rc2 = mb.solve_device(dB, dX, **opts)
    torch.cuda.synchronize()
    t_stack = time.perf_counter() - t0
    launches, iters, kms = mb.system.kernel_timing(reset=True)
    log(f"5 constituents stacked: {t_stack:.2f}s, iterations {[x.iter for x in st2]}, "
        f"{1e3 * kms / max(iters, 1):.1f} us/iteration")
    out = dict(columns=dict(solves_per_s=round(k / t_cols, 4), seconds=round(t_cols, 3), iters=[x.iter for x in st],
                            true_resid_max=tr_cols, rc=rc),
               stacked=dict(solves_per_s=round(k / t_stack, 4), seconds=round(t_stack, 3),
                            iters=[x.iter for x in st2], us_per_iteration=round(1e3 * kms / max(iters, 1), 1),
                            true_resid_max=max(x.true_resid for x in st2), rc=rc2),
               speedup=round(t_cols / t_stack, 3))
    mb.close()
    return out


def mesh_batch_variant(opts, omega, grad_only=False):
    """The 16 meshes of MESH_BATCH (seeds 3000 + i, crust N(-2, 0.5) per cell, M2): each solved
    on its own (pattern, preconditioner setup and autotuning before the timed loop), then all
    together as one block-diagonal system (per-system scalars; and shared scalars unless
    grad_only). The preconditioner is opts["precond"]; with the nodal correction the stack carries
    the parts' node incidence. Both timings cover the solves only, with inputs on the device."""
    import torch

    from actfemtide_amd import afx, mesh
    from actfemtide_amd.meshbatch import MeshBatch, joint_tol

    pc = opts["precond"]
    parts, bs, cases, bcs = [], [], [], []
    t0 = time.time()
    for i, dims in enumerate(MESH_BATCH):
        c = mesh.tide_case(*dims, seed=3000 + i)
        s, b, bc = afx.forward_system(c, omega)
        s.precond_setup(pc)
        s.autotune(1)
        parts.append(s)
        bs.append(b)
        cases.append(c)
        bcs.append(bc)
    log(f"mesh batch: {len(parts)} systems, {sum(p.n for p in parts)} edges ({time.time() - t0:.1f}s)")
    dBs = [torch.from_numpy(b).cuda() for b in bs]
    dXs = [torch.zeros_like(d) for d in dBs]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    own = []
    for s, dB, dX in zip(parts, dBs, dXs):
        st, _ = s.solve_device(dB, dX, **opts)
        own.append(st[0].iter)
    torch.cuda.synchronize()
    t_own = time.perf_counter() - t0
    own_res = [float(s.true_residual_device(dB, dX)[0]) for s, dB, dX in zip(parts, dBs, dXs)]
    log(f"mesh batch, one solve per mesh: {t_own:.2f}s, iterations {own}")

    modes = {}
    for mode in ("per_system",) if grad_only else ("per_system", "shared"):
        t0 = time.time()
        mb = MeshBatch(parts, per_system=(mode == "per_system"))
        if pc.endswith("+GRAD"):
            mb.set_nodes([c.node for c in cases], [c.line for c in cases], bcs)
        mb.system.precond_setup(pc)
        mb.system.autotune(1)
        log(f"stacked system ({mode}): n={mb.n} levels L/U={mb.system.nlev_l}/{mb.system.nlev_u} "
            f"({time.time() - t0:.1f}s)")
        b, norms = mb.joint_rhs(bs)
        dB = torch.from_numpy(b).cuda()
        dX = torch.zeros_like(dB)
        mb.system.kernel_timing(reset=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st, rc = mb.solve_device(dB, dX, norms=norms, **opts)
        torch.cuda.synchronize()
        t_joint = time.perf_counter() - t0
        launches, iters, kms = mb.system.kernel_timing(reset=True)
        res = []
        for i, s in enumerate(parts):
            xs = (mb.part(dX, i) * (1.0 if mb.per_system else float(norms[i]))).contiguous()
            res.append(float(s.true_residual_device(dBs[i], xs)[0]))
        its = [x.iter for x in st]
        log(f"mesh batch, one launch ({mode}): {t_joint:.2f}s, iterations {its}, "
            f"{1e3 * kms / max(iters, 1):.1f} us/iteration, worst part residual {max(res):.2e}")
        modes[mode] = dict(solves_per_s=round(len(parts) / t_joint, 4), seconds=round(t_joint, 3),
                           iters=its if mb.per_system else its[0], us_per_iteration=round(1e3 * kms / max(iters, 1), 1),
                           rc=rc, true_resid_max=max(res), speedup=round(t_own / t_joint, 3))
        if not mb.per_system:
            modes[mode]["joint_tol"] = joint_tol(1e-8, norms)
        mb.close()
    out = dict(meshes=[f"{a}x{b_}x{c}" for a, b_, c in MESH_BATCH], edges=[p.n for p in parts],
               one_by_one=dict(solves_per_s=round(len(parts) / t_own, 4), seconds=round(t_own, 3), iters=own,
                               true_resid_max=max(own_res)),
               one_launch_per_system_scalars=modes["per_system"], one_launch_shared_scalars=modes.get("shared"))
    for p in parts:
        p.close()
    return out


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(case, omega, plain_iters, bc=None, grad_iters=0, grad_edge="DIAG"):
    """The reference's CPU solver, restated: the oracle's serial COCG13_single (C, SSOR, the
    reference's row order) on the host cores of this machine.
    - value: one core on the same 1.4M-edge system; the per-iteration time is the median of 5
      samples of 6 iterations (after a 1-iteration warm-up), scaled to the iterations the plain
      solve needs (plain_iters, measured on the GPU for the same system) -- extrapolated;
    - full_solve_16x16x12: one whole SSOR solve (not extrapolated) at 16x16x12 cells, median of 3;
    - all_cores: the iteration's parallel phases (SpMV, dot products, vector updates) on every core
      with OpenMP, the sequential sweeps on one, as an all-core COCG13 iteration would run;
    - superlu_16x16x12: a direct complex LU solve (scipy SuperLU, symmetric-mode minimum degree
      ordering; the PARDISO-class reference) on one core, median of 3;
    - same_algorithm (when grad_iters): the oracle's restatement of the measured form (COCG with
      the grad_edge edge preconditioner -- the diagonal or the node blocks -- and the nodal
      correction, scipy, one core) on the same system, the median of 3 samples of 5 iterations
      scaled to grad_iters -- extrapolated."""
    import scipy.sparse.linalg as sla

    from actfemtide_amd import mesh
    from oracle import oracle as O
    ncores = os.cpu_count() or 1
    if os.environ.get("OMP_NUM_THREADS", "").isdigit():  # the process's share of the host
        ncores = min(ncores, int(os.environ["OMP_NUM_THREADS"]))
    t0 = time.time()
    o = O.forward_system(case, omega)
    t_asm = time.time() - t0
    args = (o["INU"], o["IAU"], o["INL"], o["IAL"], o["D"], o["AU"], o["AL"], o["b"])
    O.cocg13(*args, precond="SSOR", resid=0.0, maxit=1)
    per_it = []
    for _ in range(5):
        t0 = time.perf_counter()
        O.cocg13(*args, precond="SSOR", resid=0.0, maxit=6)
        per_it.append((time.perf_counter() - t0) / 6)
    t_it = float(np.median(per_it))
    per_solve = t_it * plain_iters
    # sequential sweeps of one iteration (the SSOR apply) and the parallel phases on all cores
    DD = O.precond_setup("SSOR", o["D"], o["INU"], o["AU"], o["INL"], o["IAL"], o["AL"])
    r = np.asarray(o["b"])
    sw = []
    for _ in range(5):
        t0 = time.perf_counter()
        O.precond_apply("SSOR", o["INU"], o["IAU"], o["AU"], o["INL"], o["IAL"], o["AL"], DD, r)
        sw.append(time.perf_counter() - t0)
    t_sweep = float(np.median(sw))
    t_par1 = O.cocg_parallel_phases(o, 1)
    t_parn = O.cocg_parallel_phases(o, ncores)
    t_it_n = t_sweep + t_parn
    # one full solve at 16x16x12 cells and its direct solve
    c1 = mesh.tide_case(16, 16, 12, seed=1002, spheres=3, sigma_sea=3.3)
    o1 = O.forward_system(c1, omega)
    args1 = (o1["INU"], o1["IAU"], o1["INL"], o1["IAL"], o1["D"], o1["AU"], o1["AL"], o1["b"])
    full = []
    for _ in range(3):
        t0 = time.perf_counter()
        r1 = O.cocg13(*args1, precond="SSOR", resid=1e-8, maxit=100000)
        full.append(time.perf_counter() - t0)
    A1 = O.to_scipy(*args1[:7]).tocsc()
    lu_t = []
    for _ in range(3):
        t0 = time.perf_counter()
        lu = sla.splu(A1, permc_spec="MMD_AT_PLUS_A", options=dict(SymmetricMode=True))
        x1 = lu.solve(np.asarray(o1["b"]))
        lu_t.append(time.perf_counter() - t0)
    lu_res = float(np.linalg.norm(o1["b"] - A1 @ x1) / np.linalg.norm(o1["b"]))
    same = None
    if grad_iters and bc is not None:
        A = O.to_scipy(o["INU"], o["IAU"], o["INL"], o["IAL"], o["D"], o["AU"], o["AL"]).tocsr()
        G, _ = O.discrete_gradient(case.line, case.node, bc)
        corr, _ = O.grad_correction(A, G)
        if grad_edge == "NODEBLK":
            Bi = O.node_block_inverse(A, O.node_blocks(case.line))
            M = lambda r: Bi @ r + corr(r)  # noqa: E731
        else:
            DDg = O.precond_setup("DIAG", o["D"], o["INU"], o["AU"], o["INL"], o["IAL"], o["AL"])
            M = lambda r: DDg * r + corr(r)  # noqa: E731
        bb = np.asarray(o["b"])
        O.cocg_precond(A, bb, M, tol=0.0, maxit=1)
        ts = []
        for _ in range(3):
            t0 = time.perf_counter()
            O.cocg_precond(A, bb, M, tol=0.0, maxit=5)
            ts.append((time.perf_counter() - t0) / 5)
        t_g = float(np.median(ts))
        same = dict(value=round(1.0 / (t_g * grad_iters), 6), unit="solves/s", cores=1, kind="port",
                    extrapolated=True, ms_per_iteration=round(1e3 * t_g, 2), iters=grad_iters,
                    note=f"the oracle's scipy restatement of the measured form (COCG, "
                         f"{'node-block' if grad_edge == 'NODEBLK' else 'diagonal'} edge preconditioner + nodal "
                         f"gradient-space correction) on the same system, median of 3 samples of 5 iterations "
                         f"scaled to the GPU solve's iteration count")
        log(f"cpu baseline, same algorithm: {1e3 * t_g:.1f} ms per iteration, {same['value']} solves/s")
    log(f"cpu baseline: {1e3 * t_it:.1f} ms per iteration on 1 core ({1e3 * t_sweep:.1f} ms sweeps, parallel "
        f"phases {1e3 * t_par1:.1f} ms on 1 core, {1e3 * t_parn:.1f} ms on {ncores}); 16x16x12 full solve "
        f"{np.median(full):.2f}s ({r1['iter']} iterations), SuperLU {np.median(lu_t):.2f}s")
    return dict(value=round(1.0 / per_solve, 8), unit="solves/s", cores=1, kind="port",
                extrapolated=True,
                sample=f"the C oracle's COCG13_single restatement (SSOR, reference row order) on the same "
                       f"{case.nline}-edge system: median of 5 samples of 6 iterations ({1e3 * t_it:.1f} ms per "
                       f"iteration) scaled to the {plain_iters} iterations of the plain solve; oracle assembly "
                       f"({t_asm:.1f}s) not included",
                host=dict(nproc=os.cpu_count(), threads_used=ncores, cpu_model=_cpu_model()),
                all_cores=dict(cores=ncores, value=round(1.0 / (t_it_n * plain_iters), 8), unit="solves/s",
                               extrapolated=True, ms_per_iteration=round(1e3 * t_it_n, 2),
                               sweeps_ms=round(1e3 * t_sweep, 2), parallel_phases_ms_1core=round(1e3 * t_par1, 2),
                               parallel_phases_ms_all=round(1e3 * t_parn, 2),
                               note="SpMV, dot products and vector updates with OpenMP on every core, the SSOR "
                                    "sweeps sequential in the reference's row order (medians of 5)"),
                full_solve_16x16x12=dict(seconds=round(float(np.median(full)), 3), iters=int(r1["iter"]),
                                         cores=1, edges=c1.nline, extrapolated=False,
                                         solves_per_s=round(1.0 / float(np.median(full)), 4)),
                superlu_16x16x12=dict(seconds=round(float(np.median(lu_t)), 3), cores=1, edges=c1.nline,
                                      residual=lu_res, note="scipy SuperLU complex128 factor + solve, "
                                                            "MMD_AT_PLUS_A ordering, symmetric mode"),
                same_algorithm=same)


if __name__ == "__main__":
    main()
