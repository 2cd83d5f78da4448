#!/usr/bin/env python3
"""TDGL+Poisson throughput of the MI355X FerroX hot path on one GPU.

Workload: 256x256x128-cell polycrystalline MFIM (dx = dy = 0.5 nm, dz =
0.1 nm; 4 nm Al2O3 under an 8 nm HZO-like layer and a 0.8 nm DE spacer; the
layer tiled by 400 seeded Voronoi grains, each with a Haar-random rotation of
its polar axis, 70 % ferroelectric and 30 % paraelectric (alpha = +1e9 m/F);
0.5 V bias; random initial Pz), FP64 throughout, forward Euler, the
reference's Poisson semantics (zero initial guess each step, MLMG stopping rule
||r||_inf <= 1e-10 max(||b||, ||r0||)).
A step is one pass of main.cpp:261-623: Newton RHS, multigrid solve, fused
TDGL update, steady-state reduction.

Prints one JSON line.  `value` is steps/s with every field resident in HBM;
`roofline` is measured live on the dominant kernel with HIP events on the
stream it runs on; `cpu_baseline` times the FP64 CPU oracle (the build's
restatement of the reference algorithm, OpenMP) on a bounded sample of the
same workload on this host.
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
    ap.add_argument("--warm-start", action="store_true", help="previous phi as the multigrid initial guess")
    ap.add_argument("--no-kernel-timing", action="store_true")
    ap.add_argument("--no-fp32", action="store_true", help="skip the FP32+FP64-refinement leg")
    ap.add_argument("--no-warm", action="store_true", help="skip the warm-start leg")
    ap.add_argument("--no-configs", action="store_true", help="skip the 64x64x36 MFIM and 128x128x64 MFISM legs")
    ap.add_argument("--mfism-steps", type=int, default=200, help="steps of the 128x128x64 MFISM leg (of its 2000-step ramp)")
    ap.add_argument("--no-diff", action="store_true", help="skip the differentiation-form leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the batching-form leg")
    ap.add_argument("--batch-steps", type=int, default=1000)
    ap.add_argument("--diff-large", action="store_true", help="also time the adjoint at 128x128x64")
    ap.add_argument("--fused", type=int, default=1, help="streamed levels: one red+black sweep per pass (1), two sweeps per pass (2), or per-colour kernels everywhere (0)")
    return ap.parse_args()


def measured_traffic():
    """HBM bytes per launch of the finest-level sweep from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE runs (copy-kernel calibrated; the
    counters cannot be read inside this process)."""
    path = os.path.join(ROOT, "profiles", "traffic_mg_sweep_poly256.md")
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

    def case(nx, ny, nz, lx, ly, lz, de_top, fe_top):
        text = configs.mfim(nx=nx, ny=ny, nz=nz, lx=lx, ly=ly, lz=lz, de_top=de_top, fe_top=fe_top, nsteps=400)
        sim = fx.FerroX(fx.from_text(text, mg_tol_rel=1e-13))
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
    out = {"workload": "64x64x36 MFIM, triangular +-3 V sweep, 400 Euler steps, checkpoints every 20 steps, "
                       "solves to 1e-13",
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


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU only; --gpus {a.gpus} is not supported "
              "(multi-GPU slab decomposition is a design note, not built)", file=sys.stderr)
        sys.exit(2)
    import ferrox_amd as fx
    from ferrox_amd import configs

    text = configs.mfim_poly_256(nsteps=a.steps + a.warmup)
    inp = fx.from_text(text, mg_warm_start=int(a.warm_start), mg_fused_sweep=a.fused)
    nx, ny, nz = inp.n_cell
    N = nx * ny * nz
    grains = configs.polycrystal_fields(inp, n_grains=400, fe_fraction=0.7, alpha_pe=1.0e9, seed=7)
    g = fx.FerroX(inp)
    configs.apply_fields(g, grains)
    print(f"# levels {g.levels()}", file=sys.stderr, flush=True)
    if a.warmup:
        g.step(a.warmup)
    g.sync()
    t0 = time.perf_counter()
    st = g.step(a.steps)
    g.sync()
    t1 = time.perf_counter()
    secs = t1 - t0
    steps_per_s = a.steps / secs
    n_v = st["total_vcycles"] / max(1, st["steps_done"])
    n_outer = st["total_outer_iters"] / max(1, st["steps_done"])
    print(f"# {a.steps} steps in {secs:.3f} s, {n_v:.2f} V-cycles/step, {n_outer:.2f} outer/step",
          file=sys.stderr, flush=True)

    # algorithmic bytes per step (FP64, scalar P, 1-byte material id, 4-byte
    # grain-table index read by the RHS and TDGL kernels)
    B_step = N * (25 + 17 + 8 + n_v * 154 + n_outer * 58)
    roof = None
    if not a.no_kernel_timing:
        reps = 50
        t_s = g.bench_kernel("smooth0", reps)         # ms per launch of the finest-level smoother
        t_t = g.bench_kernel("tdgl", reps)
        t_c = g.bench_kernel("copy", reps)
        # algorithmic bytes of one red-black sweep of the finest level: read phi and rhs,
        # write phi = 24 B/cell (no Helmholtz term in this workload, so neither the
        # material id nor the A-coefficient is read); one launch of k_mg_smooth is one
        # colour, i.e. half a sweep
        fused = bool(inp.mg_fused_sweep)
        alg = (24.0 if fused else 12.0) * N
        ach = alg / (t_s * 1e-3) / 1e9
        kname = "k_mg_gsrb_stream (finest level, red+black in one pass)" if fused else "k_mg_smooth (finest level, one colour)"
        roof = {"bound": "hbm", "kernel": kname,
                "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": measured_traffic() if fused else None,
                "alg_bytes_per_launch": alg, "avg_launch_ms": t_s,
                # TDGL update: read Pz, phi, the material id and the grain index, write Pz
                "tdgl_kernel_GBs": round(29.0 * N / (t_t * 1e-3) / 1e9, 1), "tdgl_avg_launch_ms": t_t,
                "copy_GBs": round(16.0 * N / (t_c * 1e-3) / 1e9, 1)}
    P64, F64 = g.get("Pz"), g.get("Phi")
    g.close()

    # precision form: FP32 V-cycles inside FP64 iterative refinement, same
    # workload and same number of steps, compared with the FP64 run
    fp32 = None
    if not a.no_fp32:
        g2 = fx.FerroX(fx.from_text(text, mg_warm_start=int(a.warm_start), precision=1, mg_fused_sweep=a.fused))
        configs.apply_fields(g2, grains)
        if a.warmup:
            g2.step(a.warmup)
        g2.sync()
        t0 = time.perf_counter()
        st2 = g2.step(a.steps)
        g2.sync()
        s2 = time.perf_counter() - t0
        import numpy as np
        eP = float(np.abs(g2.get("Pz") - P64).max() / np.abs(P64).max())
        eF = float(np.abs(g2.get("Phi") - F64).max() / np.abs(F64).max())
        g2.close()
        fp32 = {"steps_per_s": round(a.steps / s2, 3), "speedup_vs_fp64": round((a.steps / s2) / steps_per_s, 3),
                "vcycles_per_step": round(st2["total_vcycles"] / max(1, st2["steps_done"]), 3),
                "max_rel_err_Pz": eP, "max_rel_err_phi": eF}
        print(f"# fp32+refine: {fp32}", file=sys.stderr, flush=True)

    # the same FP64 workload with the previous step's phi as the initial guess
    # (an option of this build: the reference zeroes phi before every solve);
    # the stopping rule is unchanged, so the answer meets the same residual bound
    warm = None
    if not a.no_warm and not a.warm_start:
        g3 = fx.FerroX(fx.from_text(text, mg_warm_start=1, mg_fused_sweep=a.fused))
        configs.apply_fields(g3, grains)
        if a.warmup:
            g3.step(a.warmup)
        g3.sync()
        t0 = time.perf_counter()
        st3 = g3.step(a.steps)
        g3.sync()
        s3 = time.perf_counter() - t0
        import numpy as np
        eP = float(np.abs(g3.get("Pz") - P64).max() / np.abs(P64).max())
        eF = float(np.abs(g3.get("Phi") - F64).max() / np.abs(F64).max())
        g3.close()
        warm = {"steps_per_s": round(a.steps / s3, 3), "speedup_vs_zero_guess": round((a.steps / s3) / steps_per_s, 3),
                "vcycles_per_step": round(st3["total_vcycles"] / max(1, st3["steps_done"]), 3),
                "max_rel_err_Pz": eP, "max_rel_err_phi": eF}
        print(f"# warm start: {warm}", file=sys.stderr, flush=True)

    other = None if a.no_configs else config_legs(a.mfism_steps)
    diff = None if a.no_diff else diff_leg(a.diff_large)
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
               "sample": f"{a.cpu_steps} step(s) of the same 256x256x128 polycrystal workload, FP64 oracle (OpenMP), "
                         f"{c1 - c0:.1f} s"}
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
                               "zero-guess Poisson to 1e-10" + (" (warm start)" if a.warm_start else ""),
                   "cells": N, "vcycles_per_step": round(n_v, 3), "outer_iters_per_step": round(n_outer, 3)},
        "cell_updates_per_s": round(N * steps_per_s, 1),
        "step_bytes_algorithmic": B_step,
        "step_hbm_GBs_algorithmic": round(B_step * steps_per_s / 1e9, 1),
        "step_roofline_frac": round(B_step * steps_per_s / 1e9 / HBM_PEAK_GBS, 4),
        "roofline": roof,
        "fp32_refine": fp32,
        "warm_start": warm,
        "other_configs": other,
        "differentiation": diff,
        "batching": batch,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
