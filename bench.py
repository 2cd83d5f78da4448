#!/usr/bin/env python3
"""TDGL+Poisson throughput of the MI355X FerroX hot path on one GPU.

Workload: 256x256x128-cell polycrystalline MFIM (dx = dy = 0.5 nm, dz =
0.1 nm; 4 nm Al2O3 under an 8 nm HZO-like layer and a 0.8 nm DE spacer; the
layer tiled by 400 seeded Voronoi grains, each with a Haar-random rotation of
its polar axis, 70 % ferroelectric and 30 % paraelectric (alpha = +1e9 m/F);
0.5 V bias; random initial Pz), FP64 throughout, forward Euler.  A step is
one pass of main.cpp:261-623: Poisson RHS, Poisson solve (zero initial guess,
answer within the MLMG stopping rule ||r||_inf <= 1e-10 max(||b||, ||r0||);
this stack is laterally periodic with a z-layered permittivity, so the
library solves it directly by lateral Hartley transforms and one tridiagonal
solve per mode), TDGL update, steady-state bookkeeping.

Prints one JSON line.  `value` is steps/s with every field resident in HBM;
`roofline` is measured live on the dominant kernel (the TDGL update) with HIP
events on the stream it runs on; `kernels` gives the same for every pass of
the step; `multigrid` repeats the workload on the multigrid (FP64 and FP32
V-cycles inside FP64 refinement); `cpu_baseline` times the FP64 CPU oracle
(the build's restatement of the reference algorithm, OpenMP) on a bounded
sample of the same workload on this host.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cpu-steps", type=int, default=6, help="steps of the CPU oracle sample (0 = skip)")
    ap.add_argument("--cpu-single-steps", type=int, default=1,
                    help="steps of the one-thread CPU oracle sample (0 = skip)")
    ap.add_argument("--no-kernel-timing", action="store_true")
    ap.add_argument("--multigrid", action="store_true", help="the headline leg on the multigrid (poisson_direct = 0)")
    ap.add_argument("--no-fp32", action="store_true", help="skip the FP32 legs (direct and multigrid)")
    ap.add_argument("--no-mg", action="store_true", help="skip the multigrid legs")
    ap.add_argument("--mg-steps", type=int, default=0, help="steps of the multigrid legs (0: --steps, which the "
                    "comparison with the direct run needs)")
    ap.add_argument("--no-configs", action="store_true", help="skip the 64x64x36 MFIM and 128x128x64 MFISM legs")
    ap.add_argument("--mfism-steps", type=int, default=2000, help="steps of the 128x128x64 MFISM leg (its 0 -> 1 V ramp)")
    ap.add_argument("--no-diff", action="store_true", help="skip the differentiation-form leg")
    ap.add_argument("--no-diff-large", action="store_true", help="skip the 128x128x64 repeat of the differentiation form")
    ap.add_argument("--no-batch", action="store_true", help="skip the batching-form leg")
    ap.add_argument("--batch-steps", type=int, default=1000)
    ap.add_argument("--fused", type=int, default=1, help="multigrid legs: one red+black sweep per pass (1), two sweeps per pass (2), or per-colour kernels everywhere (0)")
    return ap.parse_args()


def measured_traffic(name="traffic_tdgl_poly256.md"):
    """HBM bytes per launch of a kernel from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE runs (copy-kernel calibrated; the counters cannot
    be read inside this process)."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                return int(line.split(":")[1])
    except OSError:
        pass
    return None


def diff_leg(large: bool) -> dict:
    """Differentiation form: the P-V curve of a 64x64x36 MFIM under a
    triangular +-3 V gate sweep over 400 steps, loss = mean squared misfit to
    the curve of theta* = theta * (1 + U(-0.1, 0.1)) (seeded); adjoint with
    checkpoints every 20 steps, checked against central differences
    (h = 1e-6 relative) on alpha and epsilonX_fe."""
    import numpy as np
    import ferrox_amd as fx
    from ferrox_amd import configs
    from ferrox_amd.diff import central_fd, triangular_sweep

    def case(nx, ny, nz, lx, ly, lz, de_top, fe_top, precision=0):
        text = configs.mfim(nx=nx, ny=ny, nz=nz, lx=lx, ly=ly, lz=lz, de_top=de_top, fe_top=fe_top, nsteps=400)
        sim = fx.FerroX(fx.from_text(text, mg_tol_rel=1e-13, precision=precision))
        return sim, sim.get("Pz")

    nsteps, K = 400, 20
    V = triangular_sweep(nsteps, 3.0)
    sim, P0 = case(64, 64, 36, 32e-9, 32e-9, 9e-9, 4e-9, 9e-9)
    theta = sim.diff_get_params()
    rng = np.random.default_rng(2024)
    theta_star = theta * (1 + rng.uniform(-0.1, 0.1, 6))

    def curve(t):
        sim.diff_set_params(t)
        sim.set("Pz", P0)
        return sim.diff_forward(V, K)

    target = curve(theta_star)
    loss = lambda m: float(np.mean((m - target) ** 2))
    sim.sync()
    t0 = time.perf_counter()
    m = curve(theta)
    sim.sync()
    t1 = time.perf_counter()
    gt, gv = sim.diff_backward(2 * (m - target) / nsteps)
    t2 = time.perf_counter()
    fd = {}
    for q, name in ((0, "alpha"), (5, "epsilonX_fe")):
        g_fd = central_fd(lambda t: loss(curve(t)), theta, q)
        fd[name] = abs(gt[q] - g_fd) / abs(g_fd)
    # the precision form's gradient: the same curve and adjoint with FP32
    # solves inside FP64 refinement (same stopping rule)
    s32, P32 = case(64, 64, 36, 32e-9, 32e-9, 9e-9, 4e-9, 9e-9, precision=1)
    s32.diff_set_params(theta)
    s32.set("Pz", P0)
    m32 = s32.diff_forward(V, K)
    g32, _ = s32.diff_backward(2 * (m32 - target) / nsteps)
    s32.close()
    sel = np.abs(gt) > 0
    grad32_err = float(np.max(np.abs(g32[sel] - gt[sel]) / np.abs(gt[sel])))
    out = {"workload": "64x64x36 MFIM, triangular +-3 V sweep, 400 Euler steps, checkpoints every 20 steps, "
                       "solves to 1e-13",
           "fp32_refine_grad_max_rel_err": grad32_err,
           "loss": loss(m), "forward_s": round(t1 - t0, 3), "backward_s": round(t2 - t1, 3),
           "grad_over_forward": round((t2 - t1) / (t1 - t0), 2),
           "checkpoint_MB": round((nsteps // K + 1 + 2 * K) * P0.size * 8 / 2**20, 1),
           "grad_theta": [float(v) for v in gt], "fd_rel_err": fd}
    print(f"# differentiation: {out}", file=sys.stderr, flush=True)
    sim.close()
    if large:
        sim, P0 = case(128, 128, 64, 64e-9, 64e-9, 16e-9, 7e-9, 16e-9)
        sim.set("Pz", P0)
        t0 = time.perf_counter()
        m = sim.diff_forward(V, K)
        t1 = time.perf_counter()
        sim.diff_backward(np.ones(nsteps) / nsteps)
        t2 = time.perf_counter()
        out["large"] = {"workload": "128x128x64 MFIM, same sweep", "forward_s": round(t1 - t0, 3),
                        "backward_s": round(t2 - t1, 3),
                        "checkpoint_MB": round((nsteps // K + 1 + 2 * K) * P0.size * 8 / 2**20, 1)}
        sim.close()
    return out


def config_legs(mfism_steps: int) -> dict:
    """The two single-device FP64 configurations besides the throughput
    workload: 64x64x36 MFIM (Al2O3 under HZO, 1000 Euler steps at 0 V) and
    128x128x64 MFISM (p-Si, Boltzmann carriers, the nonlinear outer loop of
    ComputePhi_Rho, gate ramped 0 -> 1 V over 2000 steps; the first
    mfism_steps of the ramp are timed), both with the reference's Poisson
    semantics (zero initial guess, 1e-10)."""
    import ferrox_amd as fx
    from ferrox_amd import configs
    out = {}
    for name, text, n in (("mfim_64", configs.mfim_64(nsteps=1000), 1000),
                          ("mfism_128", configs.mfism_128(nsteps=2000), mfism_steps)):
        g = fx.FerroX(fx.from_text(text))
        g.sync()
        t0 = time.perf_counter()
        st = g.step(n)
        g.sync()
        secs = time.perf_counter() - t0
        N = g.n[0] * g.n[1] * g.n[2]
        out[name] = {"steps": n, "steps_per_s": round(n / secs, 2), "cell_updates_per_s": round(N * n / secs, 1),
                     "vcycles_per_step": round(st["total_vcycles"] / max(1, st["steps_done"]), 3),
                     "outer_iters_per_step": round(st["total_outer_iters"] / max(1, st["steps_done"]), 3),
                     "gate_V_end": round(st["Phi_Bc_hi"], 4)}
        g.close()
    print(f"# configs: {out}", file=sys.stderr, flush=True)
    return out


def batch_leg(nsteps: int) -> dict:
    """Batching form: 48 independent MFIM devices (lateral extents from
    {32,48,64,96,128}^2 cells, 3-10 nm FE over 4 nm Al2O3 at dz = 0.25 nm,
    bias U(-1, 1) V, per-device seeds), FP32 solve with one FP64 correction,
    autotuned launch shapes, captured step graphs; the same batch in FP64 for
    the error."""
    import numpy as np
    import ferrox_amd as fx
    from ferrox_amd import configs
    from ferrox_amd.batch import Batch

    texts = configs.batch_devices(n=48, nsteps=nsteps)
    out = {"workload": "48 MFIM devices, lateral {32..128}^2 x 0.5 nm, FE 3-10 nm / 4 nm Al2O3 at dz 0.25 nm, "
                       "bias U(-1,1) V", "devices": 48}
    res = {}
    for prec, name in ((1, "fp32_refine"), (0, "fp64")):
        b = Batch([fx.from_text(t, precision=prec) for t in texts])
        tune = b.autotune()
        b.step(10)
        wall = b.step(nsteps)
        N = b.cells
        out["cells"] = N
        out[name] = {"steps": nsteps, "ms_per_step": round(1e3 * wall / nsteps, 4),
                     "batch_steps_per_s": round(nsteps / wall, 2),
                     "device_steps_per_s": round(48 * nsteps / wall, 1),
                     "cell_updates_per_s": round(N * nsteps / wall, 1), "autotune": tune["chosen"]}
        res[name] = [(b.get(d, "Pz"), b.get(d, "Phi")) for d in range(48)]
        b.close()
    eP = max(float(np.abs(p32 - p64).max() / np.abs(p64).max()) for (p32, _), (p64, _) in zip(res["fp32_refine"], res["fp64"]))
    eF = max(float(np.abs(f32 - f64).max() / np.abs(f64).max()) for (_, f32), (_, f64) in zip(res["fp32_refine"], res["fp64"]))
    out["max_rel_err_fp32_vs_fp64"] = {"Pz": eP, "Phi": eF}
    print(f"# batching: {out}", file=sys.stderr, flush=True)
    return out


def timed_run(text, grains, steps, warmup, **kw):
    """steps/s of `steps` steps after `warmup` untimed ones, with the stats"""
    import ferrox_amd as fx
    from ferrox_amd import configs
    g = fx.FerroX(fx.from_text(text, **kw))
    configs.apply_fields(g, grains)
    if warmup:
        g.step(warmup)
    g.sync()
    t0 = time.perf_counter()
    st = g.step(steps)
    g.sync()
    secs = time.perf_counter() - t0
    return g, st, secs


def mg_legs(text, grains, steps, warmup, fused, P64, F64, fp32: bool) -> dict:
    """The same workload on the multigrid (poisson_direct = 0): FP64, and FP32
    V-cycles inside FP64 refinement (the precision form), with the roofline
    of the finest-level sweep."""
    import numpy as np
    out = {}
    for prec, name in ((0, "fp64"), (1, "fp32_refine")):
        if prec and not fp32:
            continue
        g, st, secs = timed_run(text, grains, steps, warmup, poisson_direct=0, precision=prec, mg_fused_sweep=fused)
        n_v = st["total_vcycles"] / max(1, st["steps_done"])
        leg = {"steps": steps, "steps_per_s": round(steps / secs, 3), "vcycles_per_step": round(n_v, 3)}
        if prec == 0:
            t_s = g.bench_kernel("smooth0", 50)
            N = g.n[0] * g.n[1] * g.n[2]
            ach = 24.0 * N / (t_s * 1e-3) / 1e9
            leg["smooth0"] = {"kernel": "k_mg_gsrb_stream (finest level, red+black in one pass)", "avg_launch_ms": t_s,
                              "achieved_GBs": round(ach, 1), "frac": round(ach / HBM_PEAK_GBS, 4),
                              "traffic": measured_traffic("traffic_mg_sweep_poly256.md")}
        if P64 is not None:
            leg["max_rel_err_Pz_vs_direct_fp64"] = float(np.abs(g.get("Pz") - P64).max() / np.abs(P64).max())
            leg["max_rel_err_phi_vs_direct_fp64"] = float(np.abs(g.get("Phi") - F64).max() / np.abs(F64).max())
        g.close()
        out[name] = leg
    if "fp32_refine" in out:
        out["fp32_refine"]["speedup_vs_fp64"] = round(out["fp32_refine"]["steps_per_s"] / out["fp64"]["steps_per_s"], 3)
    print(f"# multigrid: {out}", file=sys.stderr, flush=True)
    return out


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU only; --gpus {a.gpus} is not supported "
              "(multi-GPU slab decomposition is a design note, not built)", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import ferrox_amd as fx
    from ferrox_amd import configs

    text = configs.mfim_poly_256(nsteps=a.steps + a.warmup)
    inp = fx.from_text(text)
    nx, ny, nz = inp.n_cell
    N = nx * ny * nz
    grains = configs.polycrystal_fields(inp, n_grains=400, fe_fraction=0.7, alpha_pe=1.0e9, seed=7)
    main_kw = {"poisson_direct": 0, "mg_fused_sweep": a.fused} if a.multigrid else {}
    g, st, secs = timed_run(text, grains, a.steps, a.warmup, **main_kw)
    print(f"# levels {g.levels()}", file=sys.stderr, flush=True)
    steps_per_s = a.steps / secs
    n_v = st["total_vcycles"] / max(1, st["steps_done"])
    n_outer = st["total_outer_iters"] / max(1, st["steps_done"])
    mask = g.get("mask")
    N_FE = int((mask == 0).sum())
    fe_planes = np.flatnonzero((mask == 0).any(axis=(1, 2)))   # mask is (nz, ny, nx)
    k0, k1 = int(fe_planes[0]), int(fe_planes[-1])
    print(f"# {a.steps} steps in {secs:.4f} s, {n_v:.2f} V-cycles/step, {n_outer:.2f} outer/step, {N_FE} FE cells",
          file=sys.stderr, flush=True)

    # algorithmic bytes per step (FP64, scalar P, 1-byte material id, 4-byte
    # grain-table index): the RHS pass reads P, the material id and the grain
    # index and writes the RHS of each FE cell (21 B); the TDGL update reads
    # P, phi, the material id and the grain index and writes P (29 B per FE
    # cell).  The direct solve: the forward x and y transforms read and write
    # the planes with a nonzero right-hand side (the FE planes: 16 B per cell
    # each), the z solve reads those, writes the planes the inverse
    # transforms read and one more either side (where it continues when phi
    # is read elsewhere) and reads the pivot table of those planes once
    # ((n0/2+1)(n1/2+1) doubles a plane), the inverse y and x transforms read
    # and write the planes the update reads (the FE planes and one either
    # side: 16 B per cell each).  V-cycles (154 B per cell) where the
    # multigrid runs instead.
    direct = n_v == 0
    plane = nx * ny
    n_in = (k1 - k0 + 1) * plane
    n_out = (min(k1 + 1, nz - 1) - max(k0 - 1, 0) + 1) * plane
    z_planes = min(nz, n_out // plane + 2)
    n_piv = (nx // 2 + 1) * (ny // 2 + 1) * nz
    B_direct = 32.0 * n_in + 8.0 * n_in + 8.0 * z_planes * plane + 8.0 * (nx // 2 + 1) * (ny // 2 + 1) * z_planes \
        + 32.0 * n_out
    B_step = N_FE * (21 + 29) + (B_direct if direct else 0.0) + N * (n_v * 154 + (n_outer - 1) * 58)

    roof = None
    kernels = None
    if not a.no_kernel_timing:
        reps = 50
        kernels = {}
        # (name, algorithmic bytes per launch, description)
        table = [("smooth0", 24.0 * N, "k_mg_gsrb_stream: one red+black sweep of the finest multigrid level")] if not direct else []
        table += [("tdgl", 29.0 * N_FE, "TDGL update: k_tdgl_stream over the FE cells (interior and next to a z face "
                                        "only), then over any plane holding other FE cells next to a face"),
                 ("rhs", 21.0 * N_FE, "Poisson RHS: k_rhs_stream over the FE cells (interior and next to a z face "
                                      "only), then over any plane holding other FE cells next to a face"),
                 ("fft_x", 16.0 * N, "k_fft_rows forward (Hartley along x)"),
                 ("fft_y", 16.0 * N, "k_fft_cols (Hartley along y, in place)"),
                 ("fft_tri", 16.0 * N + 8.0 * n_piv, "k_fft_tri2 (the z systems of every plane solved from both ends "
                                                      "with tabulated pivots)"),
                 ("fft_xinv", 16.0 * N, "k_fft_rows inverse"),
                 ("fft_solve", 80.0 * N + 8.0 * n_piv, "the whole direct solve on every plane (five launches)"),
                 ("copy", 16.0 * N, "plain copy (device bandwidth reference)")]
        if not direct:
            table = [t for t in table if not t[0].startswith("fft")]
        for name, alg, desc in table:
            t = g.bench_kernel(name, reps)
            kernels[name] = {"what": desc, "avg_launch_ms": round(t, 5), "alg_bytes": alg,
                             "achieved_GBs": round(alg / (t * 1e-3) / 1e9, 1),
                             "frac": round(alg / (t * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
        t = kernels["tdgl"]["avg_launch_ms"]
        ach = kernels["tdgl"]["achieved_GBs"]
        roof = {"bound": "hbm", "kernel": "TDGL update (k_tdgl_stream over every FE cell; no face-plane launch on "
                                          "this workload), the largest launch of the step",
                "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                "traffic": measured_traffic(), "alg_bytes_per_launch": 29.0 * N_FE, "avg_launch_ms": t,
                "alg_bytes_per_FE_cell": 29}
    P64, F64 = g.get("Pz"), g.get("Phi")
    g.close()

    # precision form on the direct path: FP32 transforms and tridiagonal
    # solves inside FP64 refinement to the same stopping rule
    fp32 = None
    if not a.no_fp32:
        g2, st2, s2 = timed_run(text, grains, a.steps, a.warmup, precision=1)
        eP = float(np.abs(g2.get("Pz") - P64).max() / np.abs(P64).max())
        eF = float(np.abs(g2.get("Phi") - F64).max() / np.abs(F64).max())
        g2.close()
        fp32 = {"steps_per_s": round(a.steps / s2, 3), "speedup_vs_fp64": round((a.steps / s2) / steps_per_s, 3),
                "fp32_solves_per_step": round(st2["total_vcycles"] / max(1, st2["steps_done"]), 3),
                "max_rel_err_Pz": eP, "max_rel_err_phi": eF}
        print(f"# fp32+refine (direct): {fp32}", file=sys.stderr, flush=True)

    mg = None if a.no_mg else mg_legs(text, grains, a.mg_steps or a.steps, a.warmup, a.fused,
                                      P64 if not a.mg_steps or a.mg_steps == a.steps else None, F64, not a.no_fp32)
    other = None if a.no_configs else config_legs(a.mfism_steps)
    diff = None if a.no_diff else diff_leg(not a.no_diff_large)
    batch = None if a.no_batch else batch_leg(a.batch_steps)

    cpu = None
    if a.cpu_steps > 0:
        from oracle.oracle import Oracle, threads, set_threads
        o = Oracle(fx.from_text(text))
        configs.apply_fields(o, grains)
        c0 = time.perf_counter()
        o.step(a.cpu_steps)
        c1 = time.perf_counter()
        o.close()
        cpu = {"value": round(a.cpu_steps / (c1 - c0), 5), "unit": "steps/s", "cores": threads(),
               "kind": "port",
               "sample": f"{a.cpu_steps} step(s) of the same 256x256x128 polycrystal workload, FP64 oracle (OpenMP, "
                         f"multigrid to 1e-10), {c1 - c0:.1f} s"}
        if a.cpu_single_steps > 0:
            # the same oracle on one host thread
            n_all = threads()
            set_threads(1)
            o = Oracle(fx.from_text(text))
            configs.apply_fields(o, grains)
            c0 = time.perf_counter()
            o.step(a.cpu_single_steps)
            c1 = time.perf_counter()
            o.close()
            set_threads(n_all)
            cpu["single_core"] = {"value": round(a.cpu_single_steps / (c1 - c0), 5), "unit": "steps/s", "cores": 1,
                                  "sample": f"{a.cpu_single_steps} step(s), {c1 - c0:.1f} s"}

    out = {
        "metric": "TDGL+Poisson steps/s at 256x256x128 (FP64)",
        "value": round(steps_per_s, 3),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(1e3 * secs / a.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded random initial Pz, seeded Voronoi grains and rotations)",
        "config": {"workload": "poly_256: 256x256x128 polycrystalline MFIM, 400 Voronoi grains (70% FE, "
                               "30% PE alpha=+1e9), 4 nm Al2O3 / 8 nm FE / 0.8 nm DE spacer, 0.5 V, "
                               "zero-guess Poisson within the 1e-10 MLMG rule"
                               + (" (direct solve)" if direct else ""),
                   "cells": N, "fe_cells": N_FE, "poisson": "direct" if direct else "multigrid",
                   "vcycles_per_step": round(n_v, 3), "outer_iters_per_step": round(n_outer, 3)},
        "cell_updates_per_s": round(N * steps_per_s, 1),
        "step_bytes_algorithmic": B_step,
        "direct_solve_planes": {"rhs_nonzero": [k0, k1], "phi_formed": [max(k0 - 1, 0), min(k1 + 1, nz - 1)],
                                "of": nz} if direct else None,
        "step_hbm_GBs_algorithmic": round(B_step * steps_per_s / 1e9, 1),
        "step_roofline_frac": round(B_step * steps_per_s / 1e9 / HBM_PEAK_GBS, 4),
        "roofline": roof,
        "kernels": kernels,
        "fp32_refine": fp32,
        "multigrid": mg,
        "other_configs": other,
        "differentiation": diff,
        "batching": batch,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
