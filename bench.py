#!/usr/bin/env python3
"""BEM step time on one MI355X: FP32 + FP64 iterative refinement vs plain FP64.

One step = the hot path of BEMProblem::solve (source/bem_problem.cc:969-987) on the
synthetic closed Wigley half-box: assemble_system (dense G, H and the solid angles) +
solve_system (rhs, constraints, preconditioner, GMRES, scatter) with the boundary data
already resident in HBM.  Both precisions run back to back on the same mesh and boundary
data; the FP32-IR solution is compared with the FP64 one.

Prints ONE JSON line (rank 0).  value = FP32-IR milliseconds per step.  Each precision's
run starts after the same idle pause (--settle), so neither inherits the other's heat.
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

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
FP64_VALU_PEAK_TFLOPS = 78.6
ASM_FLOPS_PER_ROW_QP = 36  # algorithmic flops per (collocation row, quadrature point)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def run_precision(mesh, bc, precision, steps, warmup, tol, ir_inner=1e-6, tile=16, pblock=64,
                  inner_format=2, inner16=3e-4, coarse=128):
    import torch

    from wavebem_amd.bem import BEMProblem, Config

    n = mesh.n_dofs
    cfg = Config(precision=precision, solver_tolerance=tol, solver_max_steps=5000,
                 ir_inner_reduction=ir_inner, assembly_tile=tile, precond_block=pblock,
                 ir_inner_format=inner_format, ir_inner16_reduction=inner16,
                 precond_coarse=coarse)
    p = BEMProblem(mesh, cfg)
    p.compute_constraints(bc)
    bc_d = torch.tensor(bc, dtype=torch.float64, device="cuda")
    phi_d = torch.tensor(bc * mesh.surface_nodes, dtype=torch.float64, device="cuda")
    dphi_d = torch.tensor(bc * mesh.other_nodes, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    for w in range(warmup):
        p.assemble_system()
        st = p.solve_system_dev(phi_d.data_ptr(), dphi_d.data_ptr(), bc_d.data_ptr())
        log(f"[{precision}] warmup {w}: {st['iterations']} it, sweeps {st['ir_sweeps']}, "
            f"true rel res {st['true_rel_residual']:.2e}")
    torch.cuda.synchronize()
    p.set_profiling(True)
    t_asm = 0.0
    stats = []
    t0 = time.perf_counter()
    for s in range(steps):
        ta = time.perf_counter()
        p.assemble_system()
        t_asm += time.perf_counter() - ta
        stats.append(p.solve_system_dev(phi_d.data_ptr(), dphi_d.data_ptr(), bc_d.data_ptr()))
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ks = p.kernel_stats()
    p.set_profiling(False)
    ms = (t1 - t0) * 1e3 / steps
    log(f"[{precision}] {ms:.1f} ms/step (assembly {t_asm * 1e3 / steps:.1f} ms), "
        f"iterations {stats[-1]['iterations']}, sweeps {stats[-1]['ir_sweeps']}")
    out = dict(ms_per_step=ms, assembly_ms=t_asm * 1e3 / steps, stats=stats[-1], kernels=ks,
               phi=phi_d.cpu().numpy(), dphi=dphi_d.cpu().numpy(), device_bytes=p.device_bytes())
    if precision == "fp32_ir" and inner_format == 2:
        out["inner_info"] = p.inner_operator_info()
    p.close()
    del bc_d, phi_d, dphi_d
    torch.cuda.empty_cache()
    return out


def kernel_avg_ms(ks, name):
    k = ks[name]
    return k["total_ms"] / k["launches"] if k["launches"] else float("nan")


def run_adjoint(workload, steps):
    """Hull-drag gradient with respect to 500 free-surface elevations and 200 hull node
    positions (the adjoint configuration): forward solve, transposed solve, backward of
    the assembly; FP64.  Returns a dict of timings (ms, mean over `steps`)."""
    import torch

    from wavebem_amd import mesh as M
    from wavebem_amd.adjoint import bem_solve, hull_drag
    from wavebem_amd.bem import BEMProblem, Config

    m = M.wigley_preset(workload)
    ptr, idx = M.double_nodes_csr(m.support_points)
    bc, U = M.wigley_bc(m, seed=0)
    rng = np.random.default_rng(7)
    rep = np.array([idx[ptr[d]:ptr[d + 1]].min() for d in range(m.n_dofs)])
    fs = np.unique(rep[np.unique(m.cell_dofs[m.cell_material == M.FREE_SURFACE])])
    hull_cells = m.cell_material == M.HULL
    hn = np.unique(rep[np.unique(m.cell_dofs[hull_cells])])
    fs_sel = rng.choice(fs, 500, replace=False)
    hull_sel = rng.choice(hn, 200, replace=False)
    X0 = m.support_points.copy()
    for d in fs_sel:
        X0[idx[ptr[d]:ptr[d + 1]], 2] += rng.uniform(-5e-3, 5e-3)
    phi_t = torch.zeros(m.n_dofs, dtype=torch.float64)
    times = {"forward": [], "adjoint_solve": [], "backward_assembly": [], "total": []}
    with BEMProblem(m, Config(solver_tolerance=1e-10), double_nodes=(ptr, idx)) as p:
        p.compute_constraints(bc)
        cells, orient = m.cell_dofs[hull_cells], m.cell_orient[hull_cells]
        for s in range(steps + 1):
            X = torch.tensor(X0, dtype=torch.float64, requires_grad=True)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            phi, dphi = bem_solve(X, p, bc)
            t1 = time.perf_counter()
            J = hull_drag(X, cells, orient, phi, dphi, phi_t, U)
            # the backward pieces, timed one by one
            g_phi, g_dphi = torch.autograd.grad(J, (phi, dphi), retain_graph=True)
            g_u = g_phi.numpy() * p.other_nodes + g_dphi.numpy() * p.surface_nodes
            t2 = time.perf_counter()
            lam = p.solve_adjoint(g_u)
            t3 = time.perf_counter()
            p.residual_vjp(lam, phi.detach().numpy(), dphi.detach().numpy())
            t4 = time.perf_counter()
            # one tangent (Newton JVP) of the residual along a seeded node motion
            dxv = rng.standard_normal((m.n_dofs, 3)) * 1e-3
            t5 = time.perf_counter()
            p.residual_jvp(dxv, phi.detach().numpy(), dphi.detach().numpy())
            t6 = time.perf_counter()
            if s == 0:
                continue  # warmup
            times["forward"].append((t1 - t0) * 1e3)
            times["adjoint_solve"].append((t3 - t2) * 1e3)
            times["backward_assembly"].append((t4 - t3) * 1e3)
            times["total"].append((t4 - t0) * 1e3)
            times.setdefault("tangent_residual_jvp", []).append((t6 - t5) * 1e3)
    out = {k: round(float(np.mean(v)), 3) for k, v in times.items()}
    out["n_dofs"] = m.n_dofs
    out["gradient_entries"] = 500 + 3 * 200
    return out


def run_multi(workload, k, steps):
    """k seeded Dirichlet data sets solved together in FP32-IR (batched matrix-core
    products), against k single FP32-IR solves and k single FP64 solves of the same
    systems; ms per solve phase (assembly excluded) and the worst error against FP64."""
    from wavebem_amd import mesh as M
    from wavebem_amd.bem import BEMProblem, Config

    m = M.wigley_preset(workload)
    ptr, idx = M.double_nodes_csr(m.support_points)
    tmp = np.stack([M.wigley_bc(m, seed=s)[0] for s in range(k)])
    N = m.other_nodes > 0
    S = ~N
    out = {"n_dofs": m.n_dofs, "k": k}
    sols = {}
    for prec in ("fp32_ir", "fp64"):
        with BEMProblem(m, Config(precision=prec), double_nodes=(ptr, idx)) as p:
            p.compute_constraints(tmp[0])
            t = time.perf_counter()
            p.assemble_system()
            p.solve_system(tmp[0] * m.surface_nodes, tmp[0] * m.other_nodes, tmp[0])
            out[f"{prec}_assembly_and_first_solve_ms"] = round((time.perf_counter() - t) * 1e3, 2)
            ts = []
            for _ in range(steps):
                phi_s, dphi_s = [], []
                t = time.perf_counter()
                for v in range(k):
                    phi = tmp[v] * m.surface_nodes
                    dphi = tmp[v] * m.other_nodes
                    p.solve_system(phi, dphi, tmp[v])
                    phi_s.append(phi)
                    dphi_s.append(dphi)
                ts.append(time.perf_counter() - t)
            out[f"{prec}_single_solves_ms"] = round(float(np.mean(ts)) * 1e3, 2)
            sols[prec] = (np.array(phi_s), np.array(dphi_s))
            if prec == "fp32_ir":
                tm = []
                for _ in range(steps):
                    phi = tmp * m.surface_nodes
                    dphi = tmp * m.other_nodes
                    t = time.perf_counter()
                    st = p.solve_system_multi(phi, dphi, tmp)
                    tm.append(time.perf_counter() - t)
                out["fp32_ir_multi_ms"] = round(float(np.mean(tm)) * 1e3, 2)
                out["fp32_ir_multi_iterations"] = st["iterations"]
                out["fp32_ir_multi_sweeps"] = st["ir_sweeps"]
                sols["multi"] = (phi, dphi)
    p64, d64 = sols["fp64"]
    for key in ("multi", "fp32_ir"):
        ph, dp = sols[key]
        out[f"{key}_max_err_vs_fp64"] = float(max(
            np.abs(ph[:, N] - p64[:, N]).max() / np.abs(p64[:, N]).max(),
            np.abs(dp[:, S] - d64[:, S]).max() / np.abs(d64[:, S]).max()))
    return out


def run_bdf(workload, steps, dt):
    """The free-surface time-loop stand-in (wavebem_amd.timeloop): fixed-step BDF2 with
    Newton iterations whose J v comes from wb_residual_jvp + wb_solve_rhs, every residual
    re-assembling G and H at the displaced mesh; FP32-IR, then FP64, on the same problem.
    Returns per-step times, Newton/GMRES counts and the hull-force drift of FP32-IR against
    FP64."""
    from wavebem_amd import mesh as M
    from wavebem_amd import timeloop as TL
    from wavebem_amd.bem import BEMProblem, Config

    m = M.wigley_preset(workload)
    ptr, idx = M.double_nodes_csr(m.support_points)
    bc, U = M.wigley_bc(m, seed=0)
    # the hull oscillates in surge with 1% of the Froude-0.3 speed (period 0.5 s, ramped over
    # 10 steps), so the waves stay small against the cells, as linear theory assumes
    eps, period = 1e-2, 0.5
    bc = np.where(m.other_nodes > 0, eps * bc, bc)
    U = eps * U

    def forcing(t):
        return TL.smooth_ramp(t, 10 * dt) * math.cos(2 * math.pi * t / period)

    out = {"n_dofs": m.n_dofs, "steps": steps, "dt": dt,
           "forcing": f"surge oscillation, {eps} x the Froude-0.3 speed, period {period} s"}
    res = {}
    for prec in ("fp32_ir", "fp64"):
        # data 100 times below the default scale: absolute tolerances 100 times smaller
        with BEMProblem(m, Config(precision=prec, solver_tolerance=1e-12),
                        double_nodes=(ptr, idx)) as p:
            p.compute_constraints(bc)
            dae = TL.FreeSurfaceStandIn(p, m, (ptr, idx), bc, forcing=forcing)
            it = TL.BDFIntegrator(dae, np.zeros(dae.size), dt, atol=1e-11)
            phis, forces, times, states = [np.zeros(m.n_dofs)], [], [], []
            for s in range(steps):
                t0 = time.perf_counter()
                y = it.step()
                phi, _, _ = dae.bem_state(it.t, y)
                times.append((time.perf_counter() - t0) * 1e3)
                phis = (phis + [phi.copy()])[-3:]
                forces.append(float(dae.hull_force(it.t, y, it.derivative(phis), U,
                                                   it.last_yp[:dae.n_nodes])[0]))
                states.append(y.copy())
                tr = " ".join(f"|F| {a:.2e} d {b:.2e} k {c}" for a, b, c in it.last_trace)
                log(f"bdf {prec} step {s + 1}/{steps}: {times[-1]:.0f} ms, force "
                    f"{forces[-1]:.6e}; newton: {tr}")
            if prec == "fp64":
                v = np.random.default_rng(1).standard_normal(dae.size)
                fr, geo = dae.jvp_parts(it.t, y, v)
                # size of the term the reference's frozen-G/H product leaves out, for a
                # seeded direction at the last state
                out["geometric_term_rel_at_last_state"] = float(
                    np.abs(geo).max() / max(np.abs(fr + geo).max(), 1e-300))
            c = dae.counters
            res[prec] = dict(ms=times, forces=np.array(forces), states=np.array(states))
            out[f"{prec}_ms_per_step"] = round(float(np.mean(times)), 2)
            out[f"{prec}_newton_per_step"] = round(it.stats["newton"] / steps, 2)
            out[f"{prec}_gmres_per_step"] = round(it.stats["krylov"] / steps, 2)
            out[f"{prec}_steps_by_newton_count"] = {
                str(k): int(v) for k, v in zip(*np.unique(it.stats["newton_per_step"],
                                                          return_counts=True))}
            out[f"{prec}_bem_solves_per_step"] = round(
                (c["solves"] + c["jvp_solves"]) / steps, 2)
            out[f"{prec}_assemblies_per_step"] = round(c["assemblies"] / steps, 2)
    f32, f64 = res["fp32_ir"]["forces"], res["fp64"]["forces"]
    out["force_rel_drift_max"] = float(np.abs(f32 - f64).max() / np.abs(f64).max())
    s32, s64 = res["fp32_ir"]["states"], res["fp64"]["states"]
    out["state_rel_drift_max"] = float(np.abs(s32 - s64).max() / np.abs(s64).max())
    out["force_fp64_first_last"] = [float(f64[0]), float(f64[-1])]
    out["speedup_vs_fp64"] = round(out["fp64_ms_per_step"] / out["fp32_ir_ms_per_step"], 3)
    return out


def run_moving(workload, steps):
    """Prescribed free-surface motion (no DAE integrator): every step moves the free-surface
    nodes to a travelling wave of 2e-3 L amplitude, re-assembles G and H and solves, in
    FP32-IR and in FP64 side by side, and evaluates the hull drag.  Returns per-step times
    and the drag drift of FP32-IR against FP64."""
    import torch

    from wavebem_amd import mesh as M
    from wavebem_amd.adjoint import hull_drag
    from wavebem_amd.bem import BEMProblem, Config

    m = M.wigley_preset(workload)
    ptr, idx = M.double_nodes_csr(m.support_points)
    bc, U = M.wigley_bc(m, seed=0)
    fs = np.unique(m.cell_dofs[m.cell_material == M.FREE_SURFACE])
    moving = np.unique(np.concatenate([idx[ptr[d]:ptr[d + 1]] for d in fs]))
    X0 = m.support_points.copy()
    hull = m.cell_material == M.HULL
    cells, orient = m.cell_dofs[hull], m.cell_orient[hull]
    phi_t = torch.zeros(m.n_dofs, dtype=torch.float64)
    dt, lam = 0.01, 0.5
    out = {}
    drags = {}
    probs = {prec: BEMProblem(m, Config(precision=prec), double_nodes=(ptr, idx))
             for prec in ("fp32_ir", "fp64")}
    try:
        for p in probs.values():
            p.compute_constraints(bc)
        times = {k: [] for k in probs}
        for s in range(steps + 1):
            t = s * dt
            X = X0.copy()
            xm = X0[moving]
            X[moving, 2] += 2e-3 * np.sin(2 * np.pi * (xm[:, 0] - U * t) / lam)
            tmp = bc * (1.0 + 0.1 * np.sin(2 * np.pi * t))
            for prec, p in probs.items():
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                p.set_support_points(X)
                p.assemble_system()
                phi = tmp * m.surface_nodes
                dphi = tmp * m.other_nodes
                p.solve_system(phi, dphi, tmp)
                t1 = time.perf_counter()
                if s:
                    times[prec].append((t1 - t0) * 1e3)
                d = float(hull_drag(torch.from_numpy(X), cells, orient, torch.from_numpy(phi),
                                    torch.from_numpy(dphi), phi_t, U))
                drags.setdefault(prec, []).append(d)
        for prec in probs:
            out[f"{prec}_ms_per_step"] = round(float(np.mean(times[prec])), 2)
    finally:
        for p in probs.values():
            p.close()
    d32, d64 = np.array(drags["fp32_ir"]), np.array(drags["fp64"])
    out["drag_rel_drift_max"] = float(np.abs(d32 - d64).max() / np.abs(d64).max())
    out["drag_fp64_first_last"] = [float(d64[0]), float(d64[-1])]
    out["n_dofs"] = m.n_dofs
    out["steps"] = steps
    return out


def profiled_traffic(workload, kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter summary of this
    workload (FETCH_SIZE + WRITE_SIZE, each scaled by a plain copy kernel), or None."""
    path = os.path.join(ROOT, "profiles", f"bem_step_{workload}.md")
    if not os.path.exists(path):
        return None
    key = f"traffic_bytes_per_launch {kernel} = "
    with open(path) as f:
        for ln in f:
            if ln.startswith(key):
                return int(float(ln[len(key):]))
    return None


def host_cpu():
    """CPU model and core counts of the host the baseline runs on."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(),
            "affinity_cores": len(os.sched_getaffinity(0))}


def cpu_leg(mesh, ptr, idx, iters64, threads, rows_sample, gemv_n):
    """One timed leg of the oracle dense path: assemble_system on `rows_sample` rows (all
    rows when it is n), scaled by n / rows, plus FullMatrix::vmult timed at n = gemv_n and
    scaled by (n / gemv_n)^2, twice per GMRES step (the reference reads G and H)."""
    import oracle as O

    n = mesh.n_dofs
    rows = min(rows_sample, n)
    r0 = (n - rows) // 2
    t0 = time.perf_counter()
    O.assemble_rows(mesh.support_points, mesh.cell_dofs, mesh.cell_orient, ptr, idx, r0,
                    r0 + rows, n_threads=threads)
    t_rows = time.perf_counter() - t0
    t_asm = t_rows * n / rows
    gn = min(gemv_n, n)
    t_gemv = O.gemv_time(gn, reps=3, n_threads=threads) * (n / gn) ** 2
    t_step = t_asm + iters64 * 2 * t_gemv
    scaled = "unscaled" if rows == n else f"scaled by n/rows"
    gscaled = "" if gn == n else f" (scaled by (n/{gn})^2)"
    return dict(value=t_step * 1e3, unit="ms", cores=threads, kind="port",
                sample=(f"oracle assemble_system rows [{r0},{r0 + rows}) of {n} "
                        f"({t_rows:.1f} s, {scaled}) + FullMatrix::vmult timed at n={gn}"
                        f"{gscaled} x 2 per GMRES step x {iters64} steps (the FP64 GPU "
                        f"iteration count); preconditioner not counted"),
                assembly_ms=t_asm * 1e3, gemv_ms=t_gemv * 1e3)


def cpu_baseline(mesh, iters64, threads, rows_sample, gemv_n, fma=True, serial_rows=2000):
    """Oracle (CPU restatement of the reference dense path) on bounded samples: all the
    cores this process may use, and one thread (the reference is serial)."""
    from wavebem_amd.mesh import double_nodes_csr

    ptr, idx = double_nodes_csr(mesh.support_points)
    out = cpu_leg(mesh, ptr, idx, iters64, threads, rows_sample, gemv_n)
    out["host"] = host_cpu()
    out["serial"] = cpu_leg(mesh, ptr, idx, iters64, 1, serial_rows, min(gemv_n, 4096))
    if fma:
        out["bemfma"] = bemfma_baseline(mesh, ptr, idx, iters64)
    return out


def bemfma_baseline(mesh, ptr, idx, iters64, trunc_order=4, levels=10):
    """The reference's other solution method, BEMFMA, restated on the host (oracle/bem_fma.c,
    serial like the reference) at the reference's default parameters (FMA Truncation Order
    4, Number of Octree Levels 10): octree + near-field + element-expansion set-up once per
    geometry, then one FMM product per GMRES step (plus alpha and the rhs)."""
    import oracle as O

    rng = np.random.default_rng(0)
    phi, psi = rng.standard_normal(mesh.n_dofs), rng.standard_normal(mesh.n_dofs)
    t0 = time.perf_counter()
    f = O.FMA(mesh.support_points, mesh.cell_dofs, mesh.cell_orient, ptr, idx,
              trunc_order=trunc_order, num_levels=levels)
    t_setup = time.perf_counter() - t0
    t0 = time.perf_counter()
    f.products(phi, psi)
    t_prod = time.perf_counter() - t0
    info = f.info()
    f.close()
    step = t_setup + (iters64 + 2) * t_prod
    return dict(value=step * 1e3, unit="ms", cores=1, kind="port",
                sample=(f"oracle BEMFMA at n={mesh.n_dofs}: set-up {t_setup:.1f} s once, one "
                        f"product {t_prod:.2f} s, x ({iters64} GMRES steps + alpha + rhs)"),
                setup_s=round(t_setup, 2), product_ms=round(t_prod * 1e3, 1),
                trunc_order=trunc_order, octree_levels=levels, blocks=info["blocks"],
                near_field_entries=info["near_entries"])


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="40k", help="Wigley preset: 2.5k 10k 20k 40k 60k")
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--cpu-rows", type=int, default=0,
                    help="rows of the all-cores assembly sample (0: all rows up to 21,000 "
                         "dofs, else 24,000)")
    ap.add_argument("--cpu-gemv-n", type=int, default=8192)
    ap.add_argument("--only", choices=["fp64", "fp32_ir"], default=None)
    ap.add_argument("--ir-inner16", type=float, default=3e-4,
                    help="FP32-IR with the FP16 inner operator: inner reduction per sweep")
    ap.add_argument("--ir-inner-format", type=int, default=2, choices=(0, 1, 2),
                    help="FP32-IR inner Krylov operator: 0 FP32 (hi), 1 row-scaled FP16 copy, "
                         "2 compressed FP16 copy (dense near blocks, low-rank far blocks)")
    ap.add_argument("--ir-inner", type=float, default=1e-7,
                    help="FP32-IR inner GMRES relative residual reduction per sweep")
    ap.add_argument("--tile", type=int, default=16, help="max columns per assembly tile (1..16)")
    ap.add_argument("--no-fma-baseline", action="store_true",
                    help="skip the host BEMFMA timing inside the CPU baseline")
    ap.add_argument("--precond-block", type=int, default=64,
                    help="block-Jacobi block size (rows of the device ordering, <= 128)")
    ap.add_argument("--settle", type=float, default=8.0, metavar="S",
                    help="idle seconds before each precision's run, so that both start from "
                         "the same state (the second run otherwise inherits the first one's "
                         "heat: about 2%% on its GEMVs); not part of any timed region")
    ap.add_argument("--coarse", type=int, default=128,
                    help="coarse-space unknowns of the preconditioner (<= 192, 0 = block Jacobi "
                         "only)")
    ap.add_argument("--multi", type=int, default=0, metavar="K",
                    help="only time K right-hand sides solved together (default workload "
                         "60k) and exit")
    ap.add_argument("--bdf", type=int, default=0, metavar="STEPS",
                    help="free-surface time-loop stand-in: STEPS BDF steps with Newton J v")
    ap.add_argument("--dt", type=float, default=0.01, help="time step of --bdf")
    ap.add_argument("--moving", type=int, default=0, metavar="STEPS",
                    help="only run STEPS re-assembled steps of a prescribed free-surface "
                         "motion in FP32-IR and FP64 (drag drift, ms per step) and exit")
    ap.add_argument("--adjoint", action="store_true",
                    help="only time the drag-gradient adjoint (default workload 10k) and exit")
    ap.add_argument("--copy-calibrate", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (4 launches) and exit: the "
                         "reference kernel for scaling FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: measures one GPU only; the path would shard by rows of G and H "
              "across GPUs, which is not implemented", file=sys.stderr)
        sys.exit(2)

    import torch

    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)
    if args.bdf:
        r = run_bdf(args.workload, args.bdf, args.dt)
        print(json.dumps({"metric": "free-surface BDF step time (Newton-Krylov, re-assembly "
                          "per residual)", "unit": "ms",
                          "value": r["fp32_ir_ms_per_step"], "higher_is_better": False,
                          "dtype": "fp16+fp32/bf16+fp64-ir", "data": "synthetic",
                          "config": {"workload": f"wigley_closed_halfbox_{args.workload}",
                                     "n_dofs": r["n_dofs"], "steps": r["steps"],
                                     "dt": r["dt"], "time_loop": "stand-in (linear free "
                                     "surface, fixed-step BDF2, full-tangent Newton)"},
                          "detail": r}), flush=True)
        return
    if args.moving:
        r = run_moving(args.workload, args.moving)
        print(json.dumps({"metric": "moving-mesh step time (re-assembly + solve)", "unit": "ms",
                          "value": r["fp32_ir_ms_per_step"], "higher_is_better": False,
                          "dtype": "fp16+fp32/bf16+fp64-ir", "data": "synthetic",
                          "config": {"workload": f"wigley_closed_halfbox_{args.workload}",
                                     "n_dofs": r["n_dofs"], "steps": r["steps"]},
                          "detail": r}), flush=True)
        return
    if args.multi:
        wl = args.workload if args.workload != "40k" else "60k"
        r = run_multi(wl, args.multi, args.steps)
        print(json.dumps({"metric": "several right-hand sides: solve time", "unit": "ms",
                          "value": r["fp32_ir_multi_ms"], "higher_is_better": False,
                          "dtype": "fp32+fp64-ir", "data": "synthetic",
                          "config": {"workload": f"wigley_closed_halfbox_{wl}",
                                     "n_dofs": r["n_dofs"], "rhs": args.multi},
                          "detail": r}), flush=True)
        return
    if args.adjoint:
        wl = args.workload if args.workload != "40k" else "10k"
        r = run_adjoint(wl, args.steps)
        print(json.dumps({"metric": "hull-drag gradient (adjoint) time", "unit": "ms",
                          "value": r["total"], "higher_is_better": False, "dtype": "f64",
                          "data": "synthetic", "config": {"workload": f"wigley_closed_halfbox_{wl}",
                                                          "n_dofs": r["n_dofs"],
                                                          "gradient_entries": r["gradient_entries"]},
                          "phases_ms": r}), flush=True)
        return
    if args.copy_calibrate:
        nb = args.copy_calibrate << 20
        src = torch.ones(nb // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(4):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read_per_launch": nb, "copy_bytes_written_per_launch": nb,
                          "launches": 4}), flush=True)
        return
    from wavebem_amd import mesh as M

    t = time.perf_counter()
    mesh = M.wigley_preset(args.workload)
    bc, U = M.wigley_bc(mesh, froude=0.3, seed=0)
    log(f"mesh {args.workload}: n={mesh.n_dofs} cells={mesh.n_cells} "
        f"({time.perf_counter() - t:.1f} s)")

    res = {}
    for prec in ("fp64", "fp32_ir"):
        if args.only and prec != args.only:
            continue
        if args.settle > 0:
            time.sleep(args.settle)  # same idle time before each precision's run
        res[prec] = run_precision(mesh, bc, prec, args.steps, args.warmup, args.tol,
                                  args.ir_inner, args.tile, args.precond_block,
                                  args.ir_inner_format, args.ir_inner16, args.coarse)

    n = mesh.n_dofs
    N = mesh.other_nodes > 0
    S = ~N
    head = res.get("fp32_ir") or res["fp64"]
    line = {
        "metric": "BEM step time (assembly+GMRES) at 40k dofs: FP32-IR vs FP64, at 1e-8 rel err",
        "value": round(head["ms_per_step"], 3),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms_per_step"], 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": (("fp32+fp64-ir" if args.ir_inner_format == 0 else "fp16+fp32/bf16+fp64-ir")
                  if "fp32_ir" in res else "f64"),
        "data": "synthetic",
        "config": {"workload": f"wigley_closed_halfbox_{args.workload}", "n_dofs": n,
                   "n_cells": mesh.n_cells, "quadrature": "QGauss 4x4 + QGaussOneOverR(5)",
                   "gmres_tol_abs_preconditioned": args.tol, "gmres_restart": 98,
                   "preconditioner": f"block_jacobi_{args.precond_block}"
                                     + (f"+coarse_{args.coarse}" if args.coarse else ""),
                   "froude": 0.3, "settle_s_before_each_precision": args.settle,
                   "fp32_ir_inner_operator": {0: "fp32", 1: "fp16_row_scaled",
                                              2: "compressed_fp16_lowrank"}[args.ir_inner_format]},
    }
    if "fp64" in res:
        line["fp64_ms_per_step"] = round(res["fp64"]["ms_per_step"], 3)
        line["fp64_iterations"] = res["fp64"]["stats"]["iterations"]
        line["fp64_true_rel_residual"] = res["fp64"]["stats"]["true_rel_residual"]
    if "fp32_ir" in res and "inner_info" in res["fp32_ir"]:
        info = res["fp32_ir"]["inner_info"]
        line["fp32_ir_inner_operator"] = dict(
            info, fp16_copy_bytes=2 * n * n,
            compression_vs_fp16=round(2 * n * n / max(info["bytes_per_product"], 1), 2))
    if "fp32_ir" in res:
        st = res["fp32_ir"]["stats"]
        line["fp32_ir_inner_iterations"] = st["iterations"]
        line["fp32_ir_sweeps"] = st["ir_sweeps"]
        k = min(st["ir_sweeps"], 8)
        line["fp32_ir_sweep_iterations"] = st["sweep_iterations"][:k]
        line["fp32_ir_sweep_residual"] = st["sweep_residual"][:k]
        line["fp32_ir_true_rel_residual"] = st["true_rel_residual"]
    if "fp64" in res and "fp32_ir" in res:
        a, b = res["fp32_ir"], res["fp64"]
        line["speedup_vs_fp64"] = round(b["ms_per_step"] / a["ms_per_step"], 3)
        line["phi_rel_err_vs_fp64"] = float(np.abs(a["phi"][N] - b["phi"][N]).max()
                                            / np.abs(b["phi"][N]).max())
        line["dphidn_rel_err_vs_fp64"] = float(np.abs(a["dphi"][S] - b["dphi"][S]).max()
                                               / np.abs(b["dphi"][S]).max())
        # hull pressure drag of both solutions (linearised Bernoulli, steady: phi_t = 0),
        # evaluated after the timed steps
        import torch

        from wavebem_amd.adjoint import hull_drag
        hull = mesh.cell_material == M.HULL
        Xh = torch.from_numpy(mesh.support_points)
        z = torch.zeros(n, dtype=torch.float64)
        dr = {k: float(hull_drag(Xh, mesh.cell_dofs[hull], mesh.cell_orient[hull],
                                 torch.from_numpy(v["phi"]), torch.from_numpy(v["dphi"]), z, U))
              for k, v in (("fp32_ir", a), ("fp64", b))}
        line["drag_fp64"] = dr["fp64"]
        line["drag_rel_err_vs_fp64"] = abs(dr["fp32_ir"] - dr["fp64"]) / abs(dr["fp64"])
    line["phase_ms"] = {k: {"assembly": round(v["assembly_ms"], 3),
                            "solve": round(v["ms_per_step"] - v["assembly_ms"], 3)}
                        for k, v in res.items()}

    # roofline of the kernel with the most time per step of the headline path: the
    # regular-cell assembly (FP64 vector ALU), then the inner-solve GEMV (HBM)
    ks = head["kernels"]
    hp = "fp32_ir" if "fp32_ir" in res else "fp64"
    a = ks["assemble"]
    asm_ms = kernel_avg_ms(ks, "assemble")
    asm_flops = ASM_FLOPS_PER_ROW_QP * n * 16 * mesh.n_cells
    asm_tf = asm_flops / (asm_ms * 1e-3) / 1e12
    asm_kernel = "k_assemble_strips<4, 1>" if hp == "fp32_ir" else "k_assemble_strips<4, 0>"
    line["roofline"] = {"kernel": asm_kernel, "bound": "mfma",
                        "pipe": "fp64 vector ALU (no matrix-core work; MI355X's FP64 vector "
                                "and matrix peaks are both 78.6 TFLOP/s)",
                        "achieved": round(asm_tf, 2), "peak": FP64_VALU_PEAK_TFLOPS,
                        "unit": "TFLOP/s", "frac": round(asm_tf / FP64_VALU_PEAK_TFLOPS, 4),
                        "traffic": profiled_traffic(args.workload, asm_kernel),
                        "avg_launch_ms": round(asm_ms, 4), "launches": a["launches"],
                        "algorithmic_flops_per_launch": int(asm_flops),
                        "flops_per_row_point": ASM_FLOPS_PER_ROW_QP,
                        "share_of_step": round(asm_ms / head["ms_per_step"], 3)}
    if "fp32_ir" not in res:
        gname, esz, xsz, kname = "gemv64", 8, 8, "k_gemv_f64"
    elif args.ir_inner_format == 2:
        gname, esz, xsz, kname = "gemvc", 0, 4, "k_hm_gemv + k_hm_finish"
    elif args.ir_inner_format == 1:
        gname, esz, xsz, kname = "gemv16", 2, 4, "k_gemv_h16"
    else:
        gname, esz, xsz, kname = "gemv32", 4, 4, "k_gemv_mixed<1>"
    g = ks[gname]
    avg_ms = kernel_avg_ms(ks, gname)
    rows_per_launch = g["rows"] / max(g["launches"], 1)
    if gname == "gemvc":
        # the compressed operator's dense blocks and low-rank factors, x read, y written
        bytes_per_launch = head["inner_info"]["bytes_per_product"] + 2 * n * xsz
    else:
        # matrix rows + x read + y written (the FP16 rows also read one 4-byte row scale)
        bytes_per_launch = rows_per_launch * n * esz + 2 * n * xsz
        if esz == 2:
            bytes_per_launch += rows_per_launch * 4
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    # launches per step: the per-GMRES-step products are timed on one launch in 8 (the
    # library's event pairs idle the GPU); calls counts every launch
    per_step = g["calls"] / max(args.steps, 1)
    line["roofline_gemv"] = {"kernel": kname, "bound": "hbm",
                             "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                             "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                             "traffic": profiled_traffic(args.workload, kname),
                             "avg_launch_ms": round(avg_ms, 4), "launches": g["launches"],
                             "calls": g["calls"], "timed_one_launch_in": 8,
                             "algorithmic_bytes_per_launch": int(bytes_per_launch),
                             "share_of_step": round(avg_ms * per_step / head["ms_per_step"], 3)}
    kern = {}
    for prec, r in res.items():
        for name, k in r["kernels"].items():
            if k["launches"]:
                kern[f"{prec}:{name}"] = {"launches": k["launches"], "calls": k["calls"],
                                          "avg_ms": round(k["total_ms"] / k["launches"], 4)}
        a = r["kernels"]["assemble"]
        if a["launches"]:
            t_a = a["total_ms"] / a["launches"] * 1e-3
            flops = ASM_FLOPS_PER_ROW_QP * n * 16 * mesh.n_cells
            kern[f"{prec}:assemble"]["tflops_algorithmic"] = round(flops / t_a / 1e12, 2)
            kern[f"{prec}:assemble"]["valu_frac_fp64_peak"] = round(
                flops / t_a / 1e12 / FP64_VALU_PEAK_TFLOPS, 4)
        for gk, es in (("gemv64", 8), ("gemvhl", 6), ("gemv32", 4), ("gemv16", 2)):
            k = r["kernels"][gk]
            if k["launches"]:
                rpl = k["rows"] / k["launches"]
                kern[f"{prec}:{gk}"]["GBps_algorithmic"] = round(
                    rpl * n * es / (k["total_ms"] / k["launches"] * 1e-3) / 1e9, 1)
    line["kernels"] = kern

    if not args.no_cpu_baseline:
        threads = args.cpu_threads or min(16, len(os.sched_getaffinity(0)))
        iters = res["fp64"]["stats"]["iterations"] if "fp64" in res else 100
        rows = args.cpu_rows or (n if n <= 21000 else 24000)
        t = time.perf_counter()
        line["cpu_baseline"] = cpu_baseline(mesh, iters, threads, rows, args.cpu_gemv_n,
                                            fma=not args.no_fma_baseline)
        log(f"cpu baseline {time.perf_counter() - t:.1f} s")
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
