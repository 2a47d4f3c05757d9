#!/usr/bin/env python3
"""Throughput of the MI355X model step on the 1024 x 1024 x 80, 500 m workload.

One "step" is RegionalModel.encStep (model.js:1442-1476): CI map, eddy viscosity,
velocity and moisture advection, forces + Thompson microphysics, divergence, the padded
1072 x 1072 x 80 pressure solve and the projection. Inputs are resident in HBM.

Prints one JSON line: sim-s per wall-s (``value``), cell-updates/s, the per-kernel HIP-event
times, the roofline of the dominant kernel, and the CPU leg: a CPU baseline (the oracle's FP32
step on a bounded sample, rank 0 only) and the error of one step against the FP64 oracle of the
same step.
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
FP32_MFMA_PEAK_TFS = 157.3
BF16_MFMA_PEAK_TFS = 2500.0
# README.md:47 of the reference: 32x real time at 1024x1024x80, dt 3 s on an RTX 5090 (93.8 ms/step);
# at the same cell-update rate and dt 2 s that is 21.3 sim-s per wall-s
RTX5090_SIM_S_PER_WALL_S_DT2 = 2.0 / 0.09375
# the mixed Thomas sweep solves a mode in FP64 where its condition estimate 4 / (|lx + ly| dz^2) exceeds this bound
# (at 1024x1024x80 the solve's error against FP64 is 6.469e-7 with bounds 16 to 1024 alike; 64 keeps ~23k modes in
# FP64)
FP64_COND = 64.0

# compulsory bytes per model cell of each stencil kernel: each field read once + written once
# (float4 texel 16 B, aux 8 B, rhs 4 B); halo re-reads excluded
# (a fused kernel counts only its own inputs and outputs: m.adv, nut and both advections in one launch, reads V,
# Q1, Q2 and writes V, aux, Q1, Q2; the hurricane step's separate advection kernels likewise)
CELL_BYTES = {"m.nut": 40, "m.advV1": 32, "m.advV2": 48, "m.adv": 104, "m.advQ1": 80, "m.advQ2": 112,
              "m.phys": 112,
              "m.div": 20, "m.proj": 36}


def kernel_roofline(name, ms, nx, ny, nz, nxp, nyp, gemm):
    cells = nx * ny * nz
    cells_p = nxp * nyp * nz
    if name in CELL_BYTES:
        b = CELL_BYTES[name] * cells
        return {"bound": "hbm", "achieved": b / (ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "algorithmic_bytes": b}
    if name in ("ps.xf", "ps.xi", "ps.yf", "ps.yi"):
        n = nxp if name in ("ps.xf", "ps.xi") else nyp
        fl = cells_p * n  # folded DST: two (M x n/2 x n/2) GEMMs per pass = cells_p * n flops
        # FP32 MFMA, or BF16 MFMA shared by the 3 (bf16x3) or 6 (bf16x6) piece products of each flop
        peak = {"fp32": FP32_MFMA_PEAK_TFS, "bf16x3": BF16_MFMA_PEAK_TFS / 3.0, "bf16x6": BF16_MFMA_PEAK_TFS / 6.0}[gemm]
        return {"bound": "mfma", "achieved": fl / (ms * 1e-3) / 1e12, "peak": peak, "unit": "TFLOP/s",
                "algorithmic_flops": fl}
    if name == "ps.fold" and gemm in ("bf16x3", "bf16x6"):
        pieces = 2 if gemm == "bf16x3" else 3
        b = (4 + 2 * pieces) * cells_p  # rhs read once, its folded pieces written once
        return {"bound": "hbm", "achieved": b / (ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "algorithmic_bytes": b}
    if name == "m.divf" and gemm in ("bf16x3", "bf16x6"):
        pieces = 2 if gemm == "bf16x3" else 3
        # V read once and rhs written once per interior cell, the folded pieces written once per padded cell
        b = 20 * cells + 2 * pieces * cells_p
        return {"bound": "hbm", "achieved": b / (ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "algorithmic_bytes": b}
    if name == "ps.tri":
        b = 8 * cells_p
        return {"bound": "hbm", "achieved": b / (ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "algorithmic_bytes": b}
    return None


def measured_traffic(kernel: str, size: str):
    """HBM bytes per launch of a kernel from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary
    (profiles/hbm_traffic_<size>.md, copy-calibrated), or None."""
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_{size}.md")
    try:
        txt = open(path).read()
        blk = txt.split("```json", 1)[1].split("```", 1)[0]
        return json.loads(blk)[kernel]["traffic_bytes"]
    except (OSError, IndexError, KeyError, ValueError):
        return None


def cpu_baseline(threads: int, dt: float, full_cells: int):
    """The oracle's FP32 step (C, OpenMP) on a 256 x 256 x 80 slice of the workload at 500 m, plus the 128 x 128 x
    40 run of config 1 and the 1024 x 1024 x 80 solve alone, each on a bounded sample."""
    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from _cases import host_model, uniforms  # noqa: E402
    from _oracle import Oracle, step_f32  # noqa: E402
    from outbreaklab_amd.configs import wk_text_for_cape  # noqa: E402

    nx = ny = 256
    nz = 80
    m = host_model(nx, ny, nz, 500.0, scheme="thompson", text=wk_text_for_cape(3000.0, 25.0))
    o = Oracle(uniforms(m, 1.0), m.B, threads=threads)
    A, B, ph = o.lsf(np.zeros(8 * 148, np.float32))
    V, Q1, Q2 = o.init(A, B, 1)
    V = o.bubble(V, 0.5 * m.Lx, 0.5 * m.Ly, 1500.0, 10000.0, 1500.0, 2.0)
    f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)  # noqa: E731
    V2, Q12, Q22 = f32(np.stack([V, V])), f32(np.stack([Q1, Q1])), f32(np.stack([Q2, Q2]))
    Vt, Q1t, Q2t = np.zeros_like(V2[0]), np.zeros_like(Q12[0]), np.zeros_like(Q12[0])
    aux = np.zeros((nz, ny, nx, 2), np.float32)
    ciMap = np.zeros((ny, nx, 4), np.float32)
    rhs = np.zeros((nz, m.nyp, m.nxp), np.float32)
    phi = np.zeros_like(rhs)
    A, B, ph = f32(A), f32(B), f32(ph)
    bd = np.zeros(8 * 148, np.float32)
    par = 0
    times = []
    for s in range(3):
        m.time = s * dt
        U = uniforms(m, dt)
        t0 = time.perf_counter()
        step_f32(U, m.B, bd, None, s == 0, par, V2, Q12, Q22, Vt, Q1t, Q2t, aux, ph, A, B, ciMap, rhs, phi,
                 threads=threads)
        times.append(time.perf_counter() - t0)
        par = 1 - par
    t = sorted(times[1:])[len(times[1:]) // 2]
    rate = nx * ny * nz / t
    return {"value": rate * dt / full_cells, "unit": "sim-s/wall-s", "cores": threads,
            "host_cpus": os.cpu_count(), "kind": "port", "cell_updates_per_s": rate,
            "sample": f"oracle FP32 step (C/OpenMP, {threads} threads, dense-GEMM solve as the reference) on "
                      f"256x256x80 at 500 m (pad 24), median of 2 steps after 1 warm-up ({t:.2f} s/step), scaled "
                      f"by cells to 1024x1024x80; the reference's own WGSL on a CPU WebGPU adapter is unavailable "
                      f"here",
            "config1": cpu_config1(threads), "solve_alone": cpu_solve_alone(threads)}


def cpu_config1(threads: int, steps: int = 40):
    """The oracle's FP32 step on config 1 (128 x 128 x 40 at 2 km, Lin, one 2 K bubble, dt 6 s): ms per step over
    `steps` steps from the bubble's start, and the 600-step run it implies."""
    import numpy as np

    from _cases import host_model, uniforms  # noqa: E402
    from _oracle import Oracle, step_f32  # noqa: E402

    nx = ny = 128
    nz, dt = 40, 6.0
    m = host_model(nx, ny, nz, 2000.0, scheme="lin")
    o = Oracle(uniforms(m, 1.0), m.B, threads=threads)
    A, B, ph = o.lsf(np.zeros(8 * 148, np.float32))
    V, Q1, Q2 = o.init(A, B, 0)
    V = o.bubble(V, 0.5 * m.Lx, 0.5 * m.Ly, 1500.0, 10000.0, 1500.0, 2.0)
    f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)  # noqa: E731
    V2, Q12, Q22 = f32(np.stack([V, V])), f32(np.stack([Q1, Q1])), f32(np.stack([Q2, Q2]))
    Vt, Q1t, Q2t = np.zeros_like(V2[0]), np.zeros_like(Q12[0]), np.zeros_like(Q12[0])
    aux = np.zeros((nz, ny, nx, 2), np.float32)
    ciMap = np.zeros((ny, nx, 4), np.float32)
    rhs = np.zeros((nz, m.nyp, m.nxp), np.float32)
    phi = np.zeros_like(rhs)
    A, B, ph = f32(A), f32(B), f32(ph)
    bd = np.zeros(8 * 148, np.float32)
    par = 0
    t0 = time.perf_counter()
    for s in range(steps):
        m.time = s * dt
        step_f32(uniforms(m, dt), m.B, bd, None, s % 10 == 0, par, V2, Q12, Q22, Vt, Q1t, Q2t, aux, ph, A, B, ciMap,
                 rhs, phi, threads=threads)
        par = 1 - par
    ms = (time.perf_counter() - t0) * 1e3 / steps
    return {"ms_per_step": round(ms, 2), "sim_s_per_wall_s": round(dt / (ms * 1e-3), 1),
            "run_600_steps_s": round(600 * ms * 1e-3, 1), "cores": threads,
            "sample": f"{steps} oracle FP32 steps of 128x128x40 at 2 km (pad {m.pad}) from the bubble's start"}


def cpu_solve_alone(threads: int):
    """The oracle's FP32 solve with the reference's arithmetic (dense sine-matrix GEMMs, FP32 Thomas) on the
    1024 x 1024 x 80 solve-alone workload: seconds per solve (one solve)."""
    import numpy as np

    from _oracle import solve as oracle_solve  # noqa: E402
    from outbreaklab_amd.configs import solve_rhs_1024
    from outbreaklab_amd.sounding import base_state, parse_sharppy, wk_sounding_text

    nx = ny = 1024
    nz, dx = 80, 500.0
    dz = 20000.0 / nz
    Bs = base_state(parse_sharppy(wk_sounding_text()), nz, dz)["B"]
    rc, rw = Bs[2:8 * nz:8].astype(np.float64)[:nz], Bs[3::8].astype(np.float64)[: nz + 1]
    rhs = solve_rhs_1024(nx, ny, nz, seed=1).astype(np.float32)
    t0 = time.perf_counter()
    oracle_solve(nx, ny, nz, dx, dz, rc, rw, rhs, "f32ref", threads=threads)
    s = time.perf_counter() - t0
    return {"s_per_solve": round(s, 2), "cores": threads,
            "sample": "one oracle solve in the reference's FP32 arithmetic (solver.js:12-103: dense sine-matrix "
                      "GEMMs, FP32 Thomas) of the 1024x1024x80 k^-5/3 RHS"}


def error_vs_fp64(m, dt, forms, threads: int):
    """Error of one GPU step against the FP64 oracle of the same step, from the model's current state, for each
    solver form: every prognostic field, theta' (K RMS), phi and the post-projection divergence (relative to the
    divergence the projection removed)."""
    from _fp64 import fp64_report  # noqa: E402

    t0 = time.perf_counter()
    r = fp64_report(m, dt, forms, threads=threads)
    out = {form: {k: float(f"{v:.4g}") for k, v in e.items()} for form, e in r.items()}
    out["state"] = f"after {m.steps} steps ({m.time:.0f} s simulated)"
    out["method"] = ("one GPU step per form from the same state; FP64 step = the oracle's FP64 kernels around an "
                     "FP64 scipy DST solve of the same rhs (tests/_fp64.py)")
    out["check_s"] = round(time.perf_counter() - t0, 1)
    return out


def timed(m, dt, steps, warm=2):
    for _ in range(warm):
        m.step(dt)
    m.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        m.step(dt)
    m.sync()
    return (time.perf_counter() - t0) / steps * 1e3


def kernel_times(m, dt, steps=3):
    m.profile(True)
    for _ in range(steps):
        m.step(dt)
    prof = m.profile_read()
    m.profile(False)
    return {k: round(v[0], 4) for k, v in prof.items()}


def _cloud(m):
    """Cells holding more than 1e-9 kg/kg of condensate, total condensate (sum of mixing ratios), and the column
    condensate map (sum over levels) of the model's current state."""
    import numpy as np

    q1, q2 = m.Q1[0].download(), m.Q2[0].download()
    c = q1[..., 1] + q1[..., 2] + q1[..., 3] + q2[..., 0] + q2[..., 1]
    return int((c > 1e-9).sum()), float(c.sum(dtype=np.float64)), c.sum(axis=0, dtype=np.float64)


def free_run(m, dt, net, steps):
    """From the model's current state, `steps` plain steps and `steps` steps with the surrogate attached (the state
    restored between them): cloudy cells, total condensate and the relative L2 difference of the column-condensate
    maps at the end. Unlike the judge, which compares each step with the plain block from the same inputs, this
    lets the surrogate's own errors feed back into the storm."""
    import numpy as np

    fields = [m.V[0], m.V[1], m.Q1[0], m.Q1[1], m.Q2[0], m.Q2[1], m.Vt, m.Q1t, m.Q2t, m.aux, m.phT, m.lsfA, m.lsfB,
              m.ciMap, m.rhs, m.phi]
    saved = [f.download() for f in fields]
    clock = (m.time, m.steps, m.parity, m.lsfT, m.phiScale)
    start = _cloud(m)
    out = {}
    for mode in ("plain", "surrogate"):
        for f, a in zip(fields, saved):
            f.upload(a)
        m.time, m.steps, m.parity, m.lsfT, m.phiScale = clock
        m.set_surrogate(net if mode == "surrogate" else None)
        for _ in range(steps):
            m.step(dt)
        m.sync()
        out[mode] = _cloud(m)
    m.set_surrogate(None)
    cp, cs = out["plain"][2], out["surrogate"][2]
    return {"steps": steps, "start": {"cloudy_cells": start[0], "condensate_sum": round(start[1], 2)},
            "plain": {"cloudy_cells": out["plain"][0], "condensate_sum": round(out["plain"][1], 2)},
            "surrogate": {"cloudy_cells": out["surrogate"][0], "condensate_sum": round(out["surrogate"][1], 2)},
            "condensate_ratio": round(out["surrogate"][1] / out["plain"][1], 4),
            "column_condensate_rel_l2": float(np.linalg.norm(cs - cp) / np.linalg.norm(cp))}


def combined_form(m, dt, steps, spin, teach_steps, p_active, train_steps, plain_gemm="bf16x6", judge_steps=600,
                  hand_min=1e-4):
    """The combined form on the same grid: the run continued until storms hold condensate, teacher
    pairs from `teach_steps` plain steps, GPU training, then the plain step (transforms `plain_gemm`, and the
    FP32-MFMA transforms) and the combined form (surrogate + split-BF16 transforms + mixed FP32/FP64 Thomas)
    timed from the same developed state, one judged step (the surrogate against the plain block from the same
    inputs), then `judge_steps` judged steps: the run continues on the surrogate while each step also reruns the
    plain block from the same inputs, and the statistics (mixing ratios, column condensate, composite reflectivity,
    surface precipitation accumulated over the run) cover the whole run."""
    from outbreaklab_amd.surrogate import Surrogate, fit, teach

    t0 = time.perf_counter()
    for _ in range(spin):
        m.step(dt)
    t_spin = time.perf_counter() - t0
    print(f"#   spun {spin} steps; sampling the plain block over {teach_steps} steps ...", file=sys.stderr, flush=True)
    rows = teach(m, dt, teach_steps, capacity=24_000_000, p_active=p_active)
    print(f"#   {rows.shape[0]} teacher rows; training {train_steps} steps ...", file=sys.stderr, flush=True)
    net = Surrogate(seed=3)
    losses = fit(net, rows, steps=train_steps, batch=8192, lr=2e-3)
    net.set_hand_min(hand_min)
    prep = time.perf_counter() - t0
    m.set_solver_mode("fp32", "mixed")
    ms_fp32 = timed(m, dt, steps)
    m.set_solver_mode(plain_gemm, "mixed")
    ms_plain = timed(m, dt, steps)
    k_plain = kernel_times(m, dt)
    m.set_surrogate(net)
    m.set_solver_mode("bf16x3", "mixed")
    ms_comb = timed(m, dt, steps)
    k_comb = kernel_times(m, dt)
    m.surrogate_stats(reset=True)
    m.step(dt, judge=True)
    st = m.surrogate_stats(reset=True)
    print(f"#   judged run of {judge_steps} steps ...", file=sys.stderr, flush=True)
    t1 = time.perf_counter()
    for _ in range(judge_steps):
        m.step(dt, judge=True)
    run = m.surrogate_stats(reset=True)
    t_run = time.perf_counter() - t1
    m.set_surrogate(None)
    m.set_solver_mode(plain_gemm, "mixed")
    print(f"#   free runs of {judge_steps} steps, plain and surrogate ...", file=sys.stderr, flush=True)
    free = free_run(m, dt, net, judge_steps)
    m.set_solver_mode(plain_gemm, "mixed")
    net.destroy()
    return {
        "ms_per_step_plain": ms_plain, "plain_gemm": plain_gemm, "ms_per_step_plain_fp32_mfma": ms_fp32,
        "ms_per_step_combined": ms_comb,
        "sim_s_per_wall_s_combined": dt / (ms_comb * 1e-3), "speedup_vs_plain": ms_plain / ms_comb,
        "speedup_vs_plain_fp32_mfma": ms_fp32 / ms_comb,
        "state": f"after {m.steps} steps ({m.time:.0f} s simulated)",
        "surrogate_points_per_step": st["points"] / max(st["steps"], 1), "fallbacks": st["fallbacks"],
        "teacher_rows": int(rows.shape[0]),
        "teacher": f"{teach_steps} plain steps after {spin} from the start, each cloudy point sampled with "
                   f"probability {p_active}",
        "training": f"{train_steps} Adam steps x 8192 rows, loss {losses[0]:.4f} -> {losses[-1]:.4f}",
        "prep_s": {"spin": round(t_spin, 1), "teach_and_train": round(prep - t_spin, 1)},
        "judge_one_step": {"rel_l2": {k: round(v, 6) for k, v in st["rel_l2"].items()},
                           "column_condensate_rel_l2": st["column_condensate_rel_l2"],
                           "dbz_rms": st["dbz_rms"], "dbz_max_abs_diff": st["dbz_max_abs_diff"],
                           "precip_rel_l2": st["precip_rel_l2"]},
        "judged_run": {"steps": run["judged"], "wall_s": round(t_run, 1),
                       "state_end": f"after {m.steps} steps ({m.time:.0f} s simulated)",
                       "rel_l2": {k: round(v, 6) for k, v in run["rel_l2"].items()},
                       "column_condensate_rel_l2": run["column_condensate_rel_l2"],
                       "dbz_rms": run["dbz_rms"], "dbz_max_abs_diff": run["dbz_max_abs_diff"],
                       "precip_accumulated_rel_l2": run["precip_rel_l2"],
                       "precip_total_surrogate_kg": run["precip_total_surrogate"],
                       "precip_total_plain_kg": run["precip_total_plain"],
                       "points_per_step": run["points"] / max(run["steps"], 1), "fallbacks": run["fallbacks"]},
        "cost_ms": {"surrogate": round(k_comb.get("m.surr", 0.0) + k_comb["m.phys"] - k_plain["m.phys"], 4),
                    "inference": k_comb.get("m.surr", 0.0),
                    "point_lists_in_phys": round(k_comb["m.phys"] - k_plain["m.phys"], 4),
                    "note": "the plain block runs only at cloudy or saturated points (clear air takes its exact "
                            "identity branch), so the surrogate replaces only those points' block; at this GPU's "
                            "rates its network plus the point lists cost more than that block"},
        "free_run": free,
        "handing": f"points with at least {hand_min:g} kg/kg of condensate inside the training envelope; thinner "
                   "cloud runs the plain block",
        "kernels_plain_ms": k_plain, "kernels_combined_ms": k_comb,
    }


def plain_128_run(steps=600):
    """The 128 x 128 x 40, 2 km plain run (Lin, one 2 K bubble, dt 6 s) timed over all its steps: ms per step,
    sim-s per wall-s, and the strongest updraft at the end."""
    import numpy as np
    from outbreaklab_amd import RegionalModel
    from outbreaklab_amd.configs import setup_plain_128

    m = RegionalModel()
    try:
        dt = setup_plain_128(m, seed=0, scheme="lin")
        m.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            m.step(dt)
        m.sync()
        ms = (time.perf_counter() - t0) * 1e3 / steps
        V, Q1, _ = m.state()
        cells = m.nx * m.ny * m.nz
        return {"workload": f"plain 128x128x40 at 2000 m, Lin, dt {dt:.0f} s, one 2 K bubble", "steps": steps,
                "ms_per_step": round(ms, 4), "sim_s_per_wall_s": round(dt / (ms * 1e-3), 1),
                "cell_updates_per_s": round(cells / (ms * 1e-3), 1), "w_max": round(float(V[..., 2].max()), 2),
                "state_finite": bool(np.isfinite(V).all() and np.isfinite(Q1).all()),
                "note": "whole run timed, including its first steps; 6.6e5 cells per step, so launch-bound"}
    finally:
        m.destroy()


# the reference's published speeds on its outbreak grids, no radars, RTX 50-series (README.md:23-28), x real time
README_SPEEDS = {"quick": 2900, "standard": 870, "large": 520, "fine": 290, "m500": 190, "m250": 70, "m100": 22}


def presets_sweep(steps=20, warmup=5):
    """The plain step on each of the reference's outbreak grids at dt 6 dx_km (app.js:1666), next to the speed the
    reference's README quotes for that grid."""
    from outbreaklab_amd import RegionalModel
    from outbreaklab_amd.configs import setup_outbreak_preset

    out = {}
    for key, ref in README_SPEEDS.items():
        m = RegionalModel()
        try:
            dt = setup_outbreak_preset(m, key)
            for _ in range(warmup):
                m.step(dt)
            m.sync()
            t0 = time.perf_counter()
            for _ in range(steps):
                m.step(dt)
            m.sync()
            ms = (time.perf_counter() - t0) * 1e3 / steps
            out[key] = {"grid": f"{m.nx}x{m.ny}x{m.nz} at {m.dx:.0f} m", "dt": dt, "ms_per_step": round(ms, 3),
                        "x_real_time": round(dt / (ms * 1e-3), 1), "readme_x_real_time_rtx50": ref,
                        "ratio": round(dt / (ms * 1e-3) / ref, 2)}
            print(f"# {key}: {out[key]['grid']}, {ms:.3f} ms/step, {out[key]['x_real_time']}x real time "
                  f"(README ~{ref}x)", file=sys.stderr, flush=True)
        finally:
            m.destroy()
    return out


def hurricane_step(key="h2", steps=20, warmup=8):
    """The hurricane-mode step (filter every 4th step, PREP, midpoint trajectories, sea fluxes) on the reference's
    hurricane grid ``key``: ms per step over ``steps`` steps after ``warmup``, sim-s per wall-s, per-kernel times."""
    from outbreaklab_amd import RegionalModel
    from outbreaklab_amd.configs import setup_hurricane_preset

    m = RegionalModel()
    try:
        dt = setup_hurricane_preset(m, key)
        for _ in range(warmup):
            m.step(dt)
        m.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            m.step(dt)
        m.sync()
        ms = (time.perf_counter() - t0) * 1e3 / steps
        m.profile(True)
        for _ in range(4):
            m.step(dt)
        prof = m.profile_read()
        m.profile(False)
        cells = m.nx * m.ny * m.nz
        return {"workload": f"hurricane {key} {m.nx}x{m.ny}x{m.nz} at {m.dx:.0f} m, dt {dt:.0f} s, 35 m/s vortex, "
                            f"302 K sea", "ms_per_step": round(ms, 3), "sim_s_per_wall_s": round(dt / (ms * 1e-3), 1),
                "cell_updates_per_s": round(cells / (ms * 1e-3), 1),
                "kernels_ms": {k: round(v[0], 4) for k, v in prof.items()},
                "note": "filter launches every 4th step (its per-launch time is shown); the hurricane advections "
                        "walk z over LDS rings, one block per 32x8 column tile"}
    finally:
        m.destroy()


def radar_volume(m, volumes=2):
    """One VCP 212 volume (17 cuts, six moments) of a radar 60 km west / 40 km south of the domain
    centre, sampled from the current state: the radar kernels' HIP-event time per volume."""
    from outbreaklab_amd.radar import Geo, RadarSim

    geo = Geo(35.9, -97.0, m.Lx, m.Ly, 0.0)
    sim = RadarSim(m, geo)
    try:
        lat, lon = geo.fromModel(m.Lx / 2 - 60e3, m.Ly / 2 - 40e3)
        R = sim.addRadar({"id": "KTST", "lat": lat, "lon": lon, "h": 0.0}, {"vcp": 212, "seed": 1})
        R["phase"] = 0.0
        dur = R["vcp"]["dur"]
        for d in sim.encDue(None, dur - 1.0):
            sim.collect(d)
        m.profile(True)
        t0 = time.perf_counter()
        for v in range(volumes):
            for d in sim.encDue(None, dur * (v + 1) + dur - 1.0):
                sim.collect(d)
        wall = (time.perf_counter() - t0) * 1e3 / volumes
        prof = m.profile_read()
        m.profile(False)
        ms, cnt = prof["r.cut"]
        vol_ms = ms * cnt / volumes
        gates = sum(c["nrad"] * R["ngRef"] for c in R["cuts"])
        return {"volume_kernel_ms": round(vol_ms, 3), "volume_wall_ms": round(wall, 3), "cuts": len(R["cuts"]),
                "gates_per_volume": gates, "ref_gates": R["ngRef"], "dop_gates": R["ngDop"], "bytes": R["bytes"],
                "gates_per_s": round(gates / (vol_ms * 1e-3), 1),
                "note": "virtual WSR-88D (radar.js:41-418), one work-item per gate; wall includes the pinned "
                        "readback and the host"}
    finally:
        sim.destroy()


def solve_alone(nx=1024, ny=1024, nz=80, dx=500.0):
    """The pressure solve alone on 1024 x 1024 x 80 (dz 250 m, no padding): a seeded k^-5/3 RHS,
    FP32-MFMA, split hi/lo BF16 and three-piece BF16 transforms with mixed Thomas, device time per solve and the
    relative error against an FP64 separable solve (scipy DST-II, independent of this library)."""
    import numpy as np
    import scipy.fft

    from outbreaklab_amd import PoissonDirect
    from outbreaklab_amd.configs import solve_rhs_1024
    from outbreaklab_amd.sounding import base_state, parse_sharppy, wk_sounding_text

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from _ref_solve import solve_scipy  # noqa: E402

    dz = 20000.0 / nz
    B = base_state(parse_sharppy(wk_sounding_text()), nz, dz)["B"]
    rc, rw = B[2:8 * nz:8].astype(np.float64)[:nz], B[3::8].astype(np.float64)[: nz + 1]
    rhs = solve_rhs_1024(nx, ny, nz, seed=1)
    with scipy.fft.set_workers(16):
        ref = solve_scipy(rhs, dx, dz, rc, rw)
    ps = PoissonDirect(nx, ny, nz, dx, dz)
    ps.setDensity(rc, rw)
    out = {"grid": f"{nx}x{ny}x{nz}, dx {dx:.0f} m, dz {dz:.0f} m", "thomas": f"mixed (FP64 where the condition estimate exceeds {FP64_COND:g})"}
    r32 = rhs.astype(np.float32)
    for gemm in ("fp32", "bf16x3", "bf16x6"):
        phi = ps.solve(r32, gemm=gemm, thomas="mixed")
        ps.profile(True)
        for _ in range(3):
            ps.solve(r32, gemm=gemm, thomas="mixed")
        prof = ps.profile_read()
        ps.profile(False)
        ms = sum(prof.values())
        err = float(np.sqrt(np.sum((phi - ref) ** 2) / np.sum(ref ** 2)))
        out[gemm] = {"ms_per_solve": round(ms, 4), "passes_ms": {k: round(v, 4) for k, v in prof.items()},
                     "rel_err_vs_fp64": err}
    ps.destroy()
    return out


def calibrate_copy(mib: int):
    """Plain device-to-device copies of a known size: FETCH_SIZE / WRITE_SIZE of their dispatches
    give the scale from the HBM counters to bytes."""
    import torch

    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"calibrate_copy_bytes": n * 4, "copies": 5}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--gemm", default="bf16x6", choices=["fp32", "bf16x3", "bf16x6"],
                    help="sine-transform GEMM form of the plain step: bf16x6 (three BF16 pieces, FP32 accuracy; the "
                         "default), fp32 (FP32 MFMA) or bf16x3 (split hi/lo, about 16 bits)")
    ap.add_argument("--thomas", default="mixed", choices=["fp32", "fp64", "mixed"])
    ap.add_argument("--size", default="1024,1024,80")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--profile-steps", type=int, default=5)
    ap.add_argument("--form", default="plain", choices=["plain", "combined", "c5", "hurricane", "presets"],
                    help="plain: the plain step at 1024x1024x80 (and the combined form on the same grid); "
                         "combined: only the combined form at 1024x1024x80; c5: the combined form at 1024x1024x100, "
                         "250 m; hurricane: only the hurricane-mode step "
                         "(--preset); presets: the plain step on each of the reference's outbreak grids")
    ap.add_argument("--preset", default="h2", choices=["h2", "h3", "h15"],
                    help="the hurricane grid of --form hurricane (model.js:1151-1153)")
    ap.add_argument("--no-combined", action="store_true", help="skip the combined-form measurement")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run 5 device-to-device copies of MIB MiB (HBM-counter calibration) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported (the multi-GPU "
              "decomposition is a design only)", file=sys.stderr)
        sys.exit(2)

    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.form == "c5":
        bench_c5(args)
        return
    if args.form == "combined":
        bench_combined(args)
        return
    if args.form == "presets":
        print(json.dumps({"presets": presets_sweep(steps=args.steps, warmup=args.warmup)}))
        return
    if args.form == "hurricane":
        print(json.dumps({"hurricane": hurricane_step(args.preset, steps=args.steps, warmup=args.warmup)}))
        return

    from outbreaklab_amd import RegionalModel
    from outbreaklab_amd.configs import setup_plain_1024

    nx, ny, nz = (int(v) for v in args.size.split(","))
    m = RegionalModel()
    dt = setup_plain_1024(m, seed=1, scheme="thompson", n=(nx, ny, nz))
    m.set_solver_mode(args.gemm, args.thomas, FP64_COND)
    print(f"# workload {nx}x{ny}x{nz} dx 500 m, pad {m.pad}, Thompson, dt {dt} s; warmup {args.warmup}", file=sys.stderr,
          flush=True)
    for _ in range(args.warmup):
        m.step(dt)
    m.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        m.step(dt)
    m.sync()
    wall = time.perf_counter() - t0
    ms_step = wall / args.steps * 1e3
    cells = nx * ny * nz
    print(f"# timed {args.steps} steps: {ms_step:.3f} ms/step", file=sys.stderr, flush=True)
    # the same step with the FP32-MFMA transforms, for comparison (not the headline form unless --gemm fp32)
    ms_fp32 = None
    if args.gemm != "fp32":
        m.set_solver_mode("fp32", args.thomas, FP64_COND)
        ms_fp32 = timed(m, dt, args.steps, warm=1)
        m.set_solver_mode(args.gemm, args.thomas, FP64_COND)
        print(f"# FP32-MFMA transforms: {ms_fp32:.3f} ms/step", file=sys.stderr, flush=True)

    # per-kernel HIP-event times on the model's stream (separate, shorter run)
    m.profile(True)
    for _ in range(args.profile_steps):
        m.step(dt)
    prof = m.profile_read()
    m.profile(False)
    kern = {}
    for name, (ms, cnt) in prof.items():
        r = kernel_roofline(name, ms, nx, ny, nz, m.nxp, m.nyp, args.gemm)
        kern[name] = {"ms": round(ms, 4), "launches": cnt}
        if r:
            r["frac"] = r["achieved"] / r["peak"]
            kern[name].update({k: (round(v, 4) if isinstance(v, float) else v) for k, v in r.items()})
    dom = max((k for k in kern if "bound" in kern[k]), key=lambda k: kern[k]["ms"])
    d = kern[dom]
    roof = {"kernel": dom, "bound": d["bound"], "achieved": d["achieved"], "peak": d["peak"], "unit": d["unit"],
            "frac": d["frac"], "traffic": measured_traffic(dom, f"{nx}x{ny}x{nz}"),
            "algorithmic": d.get("algorithmic_bytes", d.get("algorithmic_flops"))}
    V, Q1, _ = m.state()
    import numpy as np

    finite = bool(np.isfinite(V).all() and np.isfinite(Q1).all())
    # requestMaps of the current state (the step's display consumer): the map kernel's HIP-event time
    m.profile(True)
    for _ in range(3):
        m.requestMaps()()
    mprof = m.profile_read()
    m.profile(False)
    maps = None
    if "m.map" in mprof:
        ms_map = mprof["m.map"][0]
        alg = 48 * cells  # V, Q1, Q2 once per cell (a clear column's brightness temperature visits every level)
        maps = {"kernel_ms": round(ms_map, 4), "algorithmic_bytes": alg, "achieved_GBps": round(alg / (ms_map * 1e-3) / 1e9, 1),
                "note": "2-D maps (model.js:1028-1100), one work-item per column; host readback not included"}
    print("# radar volume ...", file=sys.stderr, flush=True)
    try:
        radar = radar_volume(m)
    except Exception as e:  # reported, not fatal to the plain measurement
        radar = {"error": f"{type(e).__name__}: {e}"}
    c1 = None
    if not args.no_combined:
        print("# plain 128x128x40 run (600 steps) ...", file=sys.stderr, flush=True)
        try:
            c1 = plain_128_run()
        except Exception as e:  # reported, not fatal to the plain measurement
            c1 = {"error": f"{type(e).__name__}: {e}"}
    hur = None
    if not args.no_combined:
        print("# hurricane step (h2) ...", file=sys.stderr, flush=True)
        try:
            hur = hurricane_step()
        except Exception as e:  # reported, not fatal to the plain measurement
            hur = {"error": f"{type(e).__name__}: {e}"}
    comb = None
    if not args.no_combined:
        print("# combined form (surrogate + split-BF16) on the same grid ...", file=sys.stderr, flush=True)
        try:
            comb = combined_form(m, dt, args.steps, spin=600, teach_steps=150, p_active=0.1, train_steps=20000,
                                 plain_gemm=args.gemm)
        except Exception as e:  # reported, not fatal to the plain measurement
            comb = {"error": f"{type(e).__name__}: {e}"}
    solve = None
    if not args.no_combined:
        print("# pressure solve alone (FP32 / split BF16 vs FP64) ...", file=sys.stderr, flush=True)
        try:
            solve = solve_alone()
        except Exception as e:  # reported, not fatal to the plain measurement
            solve = {"error": f"{type(e).__name__}: {e}"}
    cpu = fp64 = None
    if not args.no_cpu:
        # the CPU leg: the oracle as the baseline and as the checker of one step against FP64
        print("# CPU baseline (oracle) ...", file=sys.stderr, flush=True)
        cpu = cpu_baseline(args.cpu_threads, dt, cells)
        print("# error of one step against the FP64 oracle of the same step ...", file=sys.stderr, flush=True)
        try:
            fp64 = error_vs_fp64(m, dt, (args.gemm, "fp32") if args.gemm != "fp32" else ("fp32",), args.cpu_threads)
        except Exception as e:  # reported, not fatal to the plain measurement
            fp64 = {"error": f"{type(e).__name__}: {e}"}
    out = {
        "metric": "sim-s per wall-s at 1024x1024x80 (plain step)",
        "value": dt / (ms_step * 1e-3),
        "unit": "sim-s/wall-s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "cell_updates_per_s": cells / (ms_step * 1e-3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": (dt / (ms_step * 1e-3)) / RTX5090_SIM_S_PER_WALL_S_DT2,
        "dtype": "f32",
        "data": "synthetic (WK sounding tuned to 3000 J/kg SBCAPE, 25 m/s shear, 8 seeded 2 K bubbles)",
        "config": {"workload": f"plain step {nx}x{ny}x{nz} at 500 m, Thompson, dt {dt} s, solve "
                               f"{m.nxp}x{m.nyp}x{nz}", "gemm": args.gemm, "thomas": args.thomas},
        "ms_per_step_fp32_mfma_transforms": ms_fp32,
        "state_finite": finite,
        "roofline": roof,
        "kernels": kern,
        "cpu_baseline": cpu,
        "error_vs_fp64": fp64,
        "combined_form": comb,
        "solve_alone": solve,
        "maps": maps,
        "radar": radar,
        "plain_128": c1,
        "hurricane": hur,
        "vs_baseline_basis": "reference's published RTX 5090 step at 1024x1024x80 (README.md:47: 32x real time "
                             "at dt 3 s = 8.95e8 cell-updates/s), i.e. 21.3 sim-s/wall-s at dt 2 s",
    }
    print(json.dumps(out))
    m.destroy()


def bench_combined(args):
    """Only the combined form at 1024 x 1024 x 80, 500 m (the plain workload's storms)."""
    from outbreaklab_amd import RegionalModel
    from outbreaklab_amd.configs import setup_plain_1024

    m = RegionalModel()
    dt = setup_plain_1024(m, seed=1, scheme="thompson")
    print(f"# workload {m.nx}x{m.ny}x{m.nz} dx 500 m, pad {m.pad}, Thompson, dt {dt} s", file=sys.stderr, flush=True)
    c = combined_form(m, dt, args.steps, spin=600, teach_steps=150, p_active=0.1, train_steps=20000,
                      plain_gemm=args.gemm)
    cells = m.nx * m.ny * m.nz
    out = {"metric": "sim-s per wall-s at 1024x1024x80 (surrogate + split-BF16)",
           "value": c["sim_s_per_wall_s_combined"], "unit": "sim-s/wall-s", "n_gpus": 1, "steps": args.steps,
           "warmup": 2, "ms_per_step": c["ms_per_step_combined"],
           "cell_updates_per_s": cells / (c["ms_per_step_combined"] * 1e-3), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32 state, bf16x3 transforms, bf16 network",
           "data": "synthetic (WK sounding tuned to 3000 J/kg SBCAPE, 25 m/s shear, 8 seeded 2 K bubbles)",
           "config": {"workload": f"combined form {m.nx}x{m.ny}x{m.nz} at 500 m, Thompson, dt {dt} s"},
           "combined_form": c}
    print(json.dumps(out))
    m.destroy()


def bench_c5(args):
    """The combined form at 1024 x 1024 x 100, 250 m (one outflow boundary, dt 1 s)."""
    from outbreaklab_amd import RegionalModel
    from outbreaklab_amd.configs import setup_combined_1024

    m = RegionalModel()
    dt = setup_combined_1024(m, seed=1)
    nx, ny, nz = m.nx, m.ny, m.nz
    print(f"# workload {nx}x{ny}x{nz} dx 250 m, pad {m.pad}, Thompson, dt {dt} s", file=sys.stderr, flush=True)
    c = combined_form(m, dt, args.steps, spin=900, teach_steps=100, p_active=0.04, train_steps=20000,
                      plain_gemm=args.gemm)
    cells = nx * ny * nz
    out = {
        "metric": "sim-s per wall-s at 1024x1024x100 (surrogate + split-BF16)",
        "value": c["sim_s_per_wall_s_combined"], "unit": "sim-s/wall-s", "n_gpus": 1, "steps": args.steps,
        "warmup": 2, "ms_per_step": c["ms_per_step_combined"],
        "cell_updates_per_s": cells / (c["ms_per_step_combined"] * 1e-3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32 state, bf16x3 transforms, bf16 network",
        "data": "synthetic (WK sounding at 4000 J/kg, one seeded outflow boundary dT -6 K 1.5 km deep)",
        "config": {"workload": f"combined form {nx}x{ny}x{nz} at 250 m, Thompson, dt {dt} s, solve "
                               f"{m.nxp}x{m.nyp}x{nz}, surrogate microphysics, split-BF16 transforms, mixed Thomas"},
        "combined_form": c,
    }
    print(json.dumps(out))
    m.destroy()


if __name__ == "__main__":
    main()
