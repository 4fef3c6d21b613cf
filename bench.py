"""Benchmark of the radial BEM hot path on one MI355X.

Workload (one "step"): the 5-electrode electrostatic mirror meshed into N = 20,000 cubic
line elements, voltages U[-2, 2] kV from seed 1: dense fill (FP64 + FP32 copy) + singular
diagonal + LU of the FP32 copy on MFMA + FP64 double-double iterative refinement, i.e. the
BEM fill+solve of the metric.  Reported beside it in the same JSON line:

* the FP64 path (FP64 LU) time and the charge difference of the two paths;
* the field (potential + E) on a 1000 x 1000 (r, z) grid (10^6 points) from the charges;
* batched tracing: the 4-electrode lens (N = 8,000) with a -1.1 kV reflecting electrode,
  10^6 electrons (lognormal energy, median 1 keV, sigma 0.5; disc of 0.2 mm diameter;
  5 mrad angles) through the FieldRadialAxial interpolation, transmitted and reflected
  landings in one launch; landing errors against the FP64 oracle's direct BEM traces (golden
  fixtures from the reference's C backend) and the GPU's direct FP64 BEM tracer, fate
  mismatches counted as failures; the tracer's roofline;
* the magnetic + electrostatic lens (N = 12,000, 10 keV) and its learned surrogate, also on an
  off-axis beam against the oracle;
* roofline of the dominant kernel (HIP events on the launch stream);
* cpu_baseline: the reference's own C backend (oracle/_ref, compiled from its sources)
  on bounded samples of the same work, on this host's cores.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-trace] [--no-field] [--trace-only]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"fp64": 78.6e12, "fp32": 157.3e12, "bf16": 2500e12, "hbm": 8.0e12}
# MI355X: 256 CUs of 4 SIMDs, 16 lanes each, at 2.4 GHz (the FP64 vector peak above is
# 256 * 4 * 16 * 2 FLOP * 2.4e9)
PEAK_CUS, PEAK_CLOCK_HZ = 256, 2.4e9
# VALU instructions per potential pair of the fill's branch-free loop (gfx950 assembly)
FILL_VALU_PER_PAIR = 82
# HBM bytes per launch of the LU's bulk-update GEMM (k_gemm_x6) in this workload
# (N = 20,000): rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE averaged over its 59 launches in one
# factorisation, each scaled by what the counter reads for a 256 MiB device-to-device copy
# (--copy-calibration 256): fetch x2.0, write x1.0.  Counters cannot be read inside this
# process, so the figure is the one measured for this kernel and launch schedule; other N
# report null.
GEMM_X6_TRAFFIC_N20000 = 920.9e6
# algorithmic FLOP per unit (reference formulas, +-*/sqrt = 1, FMA = 2, log = 20)
C_POT, C_FIELD, C_FLUX = 71, 215, 219


def log(*a):
    print(*a, file=sys.stderr, flush=True)


class Timer:
    """HIP events recorded on the current (launch) stream."""

    def __init__(self):
        import torch
        self.s = torch.cuda.Event(enable_timing=True)
        self.e = torch.cuda.Event(enable_timing=True)

    def __enter__(self):
        self.s.record()
        return self

    def __exit__(self, *a):
        import torch
        self.e.record()
        torch.cuda.synchronize()
        self.ms = self.s.elapsed_time(self.e)


def mirror_problem(N):
    from traceon_amd import configs
    mesh, volts = configs.mirror(N, seed=1)
    return configs.lines_types_values(mesh, volts)


def fill_solve(lines, types, values, mode, gemm_timing=None):
    """One fill+solve step on the device; returns (charges tensor, stats, times dict).
    gemm_timing (device.KernelTiming): HIP events around the LU's MFMA GEMM launches."""
    import torch
    from traceon_amd import device
    t = {}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    E = device.Elements(lines, types, values)
    with Timer() as tf:
        M, M32 = device.assemble(E, fp64=True, fp32=(mode == "mixed"))
    F = np.where(np.isin(types, [1, 2, 5]), values, 0.0)
    with Timer() as ts:
        X, stats, lu = device.solve_system(M, F, M32=M32, mode=mode, gemm_timing=gemm_timing)
    torch.cuda.synchronize()
    t["total_s"] = time.perf_counter() - t0
    t["fill_ms"] = tf.ms
    t["solve_ms"] = ts.ms
    del M, M32, lu
    return X, stats, t


def fill_kernel_roofline(lines, types, values, reps=3):
    """Time the fill kernel alone (HIP events on its stream) for the roofline entry."""
    import torch
    from traceon_amd import device
    E = device.Elements(lines, types, values)
    N = E.N
    M = torch.empty((N, N), dtype=torch.float64, device="cuda")
    L = device._L
    L.tb_fill_matrix(device.ptr(M), N, None, 0, device.ptr(E.rows), device.ptr(E.src), N, 0, N - 1, 0, device.stream_ptr())
    with Timer() as t:
        for _ in range(reps):
            L.tb_fill_matrix(device.ptr(M), N, None, 0, device.ptr(E.rows), device.ptr(E.src), N, 0, N - 1, 0, device.stream_ptr())
    ms = t.ms / reps
    npot = int(np.sum(np.isin(types, [1, 2, 5])))
    flops = (npot * C_POT + (N - npot) * C_FLUX) * N * 16.0
    del M
    return ms, flops


def field_grid_leg(X, lines, types, values, n_side=1000):
    import torch
    from traceon_amd import device
    E = device.Elements(lines, types, values)
    S = device.Sources(X[0], E.jac, E.pos)
    r = np.linspace(0.0, 4.0e-3, n_side)
    z = np.linspace(-0.2e-3, 10.2e-3, n_side)
    R, Z = np.meshgrid(r, z, indexing="ij")
    pts = device.to_dev(np.column_stack([R.ravel(), np.zeros(R.size), Z.ravel()]))
    S.field_points(pts[:1000])   # warm-up
    with Timer() as t:
        P, Fd = S.field_points(pts)
    pairs = float(pts.shape[0]) * S.n
    return t.ms, pairs, P, Fd, pts


GOLDEN = os.path.join(ROOT, "tests", "golden")
C_RK, C_AXEVAL = 530, 125   # RKF45 step overhead; one paraxial series evaluation (E or H)


def golden(name):
    """A golden fixture (FP64 oracle results written by tests/golden/make_config_fixtures.py), or None."""
    path = os.path.join(GOLDEN, name)
    return np.load(path, allow_pickle=False) if os.path.isfile(path) else None


def traces(tracer, P, V, exit_plane, back_plane=None, timer=False, atol=1e-8):
    """One batched launch: landing (exit plane, else back plane), plane index (0 exit, 1 back,
    -1 none), accepted and attempted steps; optionally timed with HIP events."""
    import torch
    torch.cuda.synchronize()
    shape = TRACE_SHAPE or "auto"
    if timer:
        with Timer() as t:
            r = tracer.trace_batch(P, V, plane=exit_plane, plane2=back_plane, shape=shape, atol=atol)
    else:
        r, t = tracer.trace_batch(P, V, plane=exit_plane, plane2=back_plane, shape=shape, atol=atol), None
    land = r.landing.cpu().numpy()
    plane = np.where(np.isfinite(land[:, 0]), 0, -1)
    if back_plane is not None:
        l2 = r.landing2.cpu().numpy()
        back = (plane < 0) & np.isfinite(l2[:, 0])
        land[back] = l2[back]
        plane[back] = 1
    out = {"landing": land, "plane": plane, "n_accepted": r.n_accepted.cpu().numpy(), "n_attempted": r.n_attempted.cpu().numpy(),
           "parked": r.n_parked}
    if t is not None:
        out["ms"] = t.ms
    return out


def landing_error(a, b):
    """Landing statistics of trace set a against reference b (same particles): RMS and max over
    the particles with the same fate; a particle whose fate differs counts as a failure of the
    10 nm bar (fate_mismatches)."""
    n = min(len(a["plane"]), len(b["plane"]))
    same = a["plane"][:n] == b["plane"][:n]
    ok = same & (b["plane"][:n] >= 0)
    e = np.linalg.norm(a["landing"][:n][ok, :3] - b["landing"][:n][ok, :3], axis=1)
    rms = float(np.sqrt(np.mean(e ** 2)) * 1e9) if ok.any() else None
    return {"n": int(n), "fate_mismatches": int((~same).sum()), "landing_rms_nm": rms,
            "landing_max_nm": float(e.max() * 1e9) if ok.any() else None,
            "within_10nm": bool(ok.any() and (~same).sum() == 0 and rms is not None and rms <= 10.0)}


def oracle_traces(f, prefix):
    return {"landing": f[prefix + "landing"], "plane": f[prefix + "plane"]}


def shape_str(s):
    return f"{'ordered' if s.ordered else 'unordered'},batch={s.batch},occ={s.occupancy}"


def parse_shape(text):
    from traceon_amd.tracing import LaunchShape
    parts = text.split(",")
    kv = dict(p.split("=") for p in parts[1:])
    return LaunchShape(ordered=parts[0] == "ordered", batch=int(kv.get("batch", 0)), occupancy=int(kv.get("occ", 2)))


# a fixed launch shape for the tracing legs (--trace-shape), else each leg autotunes
TRACE_SHAPE = None


def tuned(tracer, P, V, **planes):
    from traceon_amd import autotune
    if TRACE_SHAPE is not None:
        return TRACE_SHAPE, {}
    return autotune.tune(tracer, P, V, **planes)


def roofline_trace(steps_attempted, ms, series_per_eval):
    """Tracer roofline: algorithmic FLOP per attempted RKF45 step = C_rk + 6 evaluations of
    series_per_eval paraxial series (C_axeval each), against the FP64 vector peak."""
    flop = float(steps_attempted) * (C_RK + 6 * series_per_eval * C_AXEVAL)
    achieved = flop / (ms * 1e-3) / 1e12
    return {"kernel": "k_trace_axial (persistent RKF45, lane per particle)", "bound": "fp64-valu",
            "flop_per_attempted_step": C_RK + 6 * series_per_eval * C_AXEVAL, "attempted_steps": int(steps_attempted),
            "achieved": achieved, "peak": PEAK["fp64"] / 1e12, "unit": "TFLOP/s", "frac": achieved / (PEAK["fp64"] / 1e12)}


def tracing_leg(n_particles, n_sample, n_axial=7000):
    """Batched tracing (C2): the reflecting 4-electrode lens (N = 8000) and 10^6 electrons
    (lognormal 1 keV, 0.2 mm disc, 5 mrad) through FieldRadialAxial; transmitted electrons land
    on z = 9.5 mm, reflected ones back through z = -4.5 mm, both recorded in the same launch."""
    import torch
    from traceon_amd import configs, device, solver
    from traceon_amd.excitation import Excitation
    from traceon_amd.field import FieldRadialAxial
    mm = 1e-3
    mesh, volts = configs.reflecting_lens(8000)
    exc = Excitation(mesh)
    exc.add_voltage(**volts)
    field = solver.solve_direct(exc)
    t0 = time.perf_counter()
    # The interpolation covers the tracing bounds plus the longest step (hmax |v| = 1 cm,
    # tracing.c:68-70) on both sides: RK stage points of a step can lie that far ahead, and
    # FieldRadialAxial is zero outside its range (radial.c:313-384).  5 um sample spacing.
    axial = FieldRadialAxial(field, -15.0 * mm, 20.0 * mm, N=n_axial)
    torch.cuda.synchronize()
    t_axial_build = time.perf_counter() - t0
    # the first n_particles of the 10^6-electron workload beam (the oracle fixture holds the
    # traces of its first particles, whatever n_particles a profiling run asks for)
    P, V, bounds, exit_plane, back_plane = configs.reflecting_beam(max(n_particles, 1_000_000))
    P, V = P[:n_particles], V[:n_particles]
    tr_ax = axial.get_tracer(bounds)
    Pd, Vd = device.to_dev(P), device.to_dev(V)    # the beam is resident in HBM before timing
    traces(tr_ax, Pd[:4096], Vd[:4096], exit_plane, back_plane)   # warm-up
    shape, tuning = tuned(tr_ax, Pd, Vd, plane=exit_plane, plane2=back_plane)
    res = traces(tr_ax, Pd, Vd, exit_plane, back_plane, timer=True)
    nacc = res["n_accepted"]
    out = {"n": n_particles, "ms": res["ms"], "electrons_per_s": n_particles / (res["ms"] * 1e-3),
           "steps_attempted": int(res["n_attempted"].sum()), "steps_accepted": int(nacc.sum()),
           "reflected_fraction": float(np.mean(res["plane"] == 1)), "lost_fraction": float(np.mean(res["plane"] < 0)),
           "len_ratio_max_min": float(nacc.max() / max(nacc.min(), 1)),
           "len_ratio_p999_p001": float(np.percentile(nacc, 99.9) / max(np.percentile(nacc, 0.1), 1)),
           "n_axial": len(axial.z), "axial_build_s": t_axial_build, "N_elements": len(field.electrostatic_point_charges),
           "launch_shape": shape_str(shape), "tuning_ms": {shape_str(k): round(v, 3) for k, v in tuning.items()}}
    out["roofline"] = roofline_trace(out["steps_attempted"], res["ms"], 1)
    # The same workload as one captured HIP graph per beam (tracing.BeamPipeline): the start
    # states drawn on the device (tb_sample_beam: the same distributions, another generator)
    # and traced by the no-alloc tracer in the autotuned shape; one replay timed with HIP
    # events on the stream it runs on, the sampling included.
    if 0 < n_particles <= 1 << 21:
        from traceon_amd.tracing import BeamPipeline
        pipe = BeamPipeline(tr_ax, n_particles, dict(z0=-5.0 * mm, r_max=0.1 * mm, e_median=1000.0, e_sigma=0.5,
                                                     angle_sigma=5e-3), atol=1e-8, plane=exit_plane, plane2=back_plane,
                            shape=shape)
        pipe(1)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        r = pipe(2)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1)
        exit_ok = torch.isfinite(r.landing[:, 0])
        back_only = torch.isfinite(r.landing2[:, 0]) & ~exit_ok
        out["graph_pipeline"] = {"n": n_particles, "ms": ms, "electrons_per_s": n_particles / (ms * 1e-3),
                                 "reflected_fraction": float(back_only.double().mean().item()),
                                 "steps_attempted": int(r.n_attempted.sum().item())}
        del pipe, r
    # accuracy: against the FP64 oracle's direct BEM traces (reference C backend, converged
    # diagonal) of the beam's first particles, and against the GPU's direct FP64 BEM tracer
    # (itself within 1e-9 m of the oracle) on a larger sample
    f = golden("config_c2.npz")
    if f is not None:
        out["vs_oracle_bem"] = landing_error(res, oracle_traces(f, "bem_"))
        out["vs_oracle_axial"] = landing_error(res, oracle_traces(f, "axial_"))
    tr_bem = field.get_tracer(bounds)
    ref = traces(tr_bem, P[:n_sample], V[:n_sample], exit_plane, back_plane, timer=True)
    out["vs_gpu_bem"] = landing_error(res, ref)
    out["electrons_per_s_direct_bem_fp64"] = n_sample / (ref["ms"] * 1e-3)
    out["n_sample"] = int(n_sample)
    ctx = {"field": field, "axial": axial, "P": P, "V": V, "bounds": bounds, "plane": exit_plane}
    return out, ctx


def magnetic_problem():
    from traceon_amd import configs
    from traceon_amd.excitation import Excitation
    mesh, volts, perms, currents = configs.magnetic_lens(12000)
    exc = Excitation(mesh)
    exc.add_voltage(**volts)
    exc.add_magnetizable(**perms)
    exc.add_current(**currents)
    return exc


def c3_leg(n_particles, n_sample, n_axial=7000):
    """Magnetic + electrostatic lens (C3): coil 500 A-turn, two mu_r = 1000 pole pieces (bore 2 mm,
    gap 3 mm) and a three-tube einzel lens, N = 12,000; electrostatic and magnetostatic solves,
    10 keV electrons traced through both axial series with the Lorentz force."""
    import torch
    from traceon_amd import configs, device, solver
    from traceon_amd.field import FieldRadialAxial
    mm = 1e-3
    exc = magnetic_problem()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    field = solver.solve_direct(exc)
    torch.cuda.synchronize()
    t_solve = time.perf_counter() - t0
    t0 = time.perf_counter()
    axial = FieldRadialAxial(field, -24.0 * mm, 22.0 * mm, N=int(n_axial * 46 / 35))
    torch.cuda.synchronize()
    t_axial = time.perf_counter() - t0
    P, V, bounds, plane = configs.magnetic_beam(max(n_particles, 1_000_000))
    P, V = P[:n_particles], V[:n_particles]
    tr_ax = axial.get_tracer(bounds)
    Pd, Vd = device.to_dev(P), device.to_dev(V)
    traces(tr_ax, Pd[:4096], Vd[:4096], plane)
    shape, tuning = tuned(tr_ax, Pd, Vd, plane=plane)
    res = traces(tr_ax, Pd, Vd, plane, timer=True)
    out = {"N_electrostatic": len(field.electrostatic_point_charges), "N_magnetostatic": len(field.magnetostatic_point_charges),
           "N_coil_rings": len(field.current_point_charges), "solve_s": t_solve, "axial_build_s": t_axial, "n": n_particles,
           "ms": res["ms"], "electrons_per_s": n_particles / (res["ms"] * 1e-3), "steps_attempted": int(res["n_attempted"].sum()),
           "landed_fraction": float(np.mean(res["plane"] == 0)),
           "launch_shape": shape_str(shape), "tuning_ms": {shape_str(k): round(v, 3) for k, v in tuning.items()}}
    out["roofline"] = roofline_trace(out["steps_attempted"], res["ms"], 2)
    f = golden("config_c3.npz")
    if f is not None:
        out["vs_oracle_bem"] = landing_error(res, oracle_traces(f, "bem_"))
        out["vs_oracle_axial"] = landing_error(res, oracle_traces(f, "axial_"))
    ref = traces(field.get_tracer(bounds), P[:n_sample], V[:n_sample], plane, timer=True)
    out["vs_gpu_bem"] = landing_error(res, ref)
    out["electrons_per_s_direct_bem_fp64"] = n_sample / (ref["ms"] * 1e-3)
    out["n_sample"] = int(n_sample)
    ctx = {"field": field, "axial": axial, "P": Pd, "V": Vd, "bounds": bounds, "plane": plane, "ref": ref, "res": res}
    return out, ctx


def c4_leg(ctx, n_samples=1 << 22, steps=8000):
    """Learned surrogate (C4) in the C3 geometry: the MLP (2-128x4-5, SiLU, BF16 MFMA, FP32
    master weights) learns BEM - axial from n_samples BEM-labelled (r, z) points over the lens
    region; the tracer adds it to the axial field inside r <= 0.6 mm and uses the direct BEM
    sum outside.  Reported: build times, field error per radial band with and without the
    correction, tracing of the C3 beam (10^6 electrons) against direct FP64 BEM, and the
    off-axis beam (start radius 0.3-0.5 mm) against the FP64 oracle's direct BEM traces, beside
    the axial series alone."""
    import torch
    from traceon_amd import configs
    from traceon_amd.surrogate import FieldRadialAxialLearned, FieldSurrogate
    mm = 1e-3
    field, axial, P, V = ctx["field"], ctx["axial"], ctx["P"], ctx["V"]
    bounds, plane = ctx["bounds"], ctx["plane"]
    sur = FieldSurrogate(field, axial, r_max=0.6 * mm, z_min=bounds[2, 0], z_max=bounds[2, 1])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sur.generate(n_samples)
    torch.cuda.synchronize()
    t_gen = time.perf_counter() - t0
    hist = sur.train(steps=steps, batch=16384, log_every=max(steps // 4, 1))
    val = sur.validate()
    learned = FieldRadialAxialLearned(sur)
    tr = learned.get_tracer(bounds)
    traces(tr, P[:4096], V[:4096], plane)
    shape, tuning = tuned(tr, P, V, plane=plane)
    res = traces(tr, P, V, plane, timer=True)
    bands = {k: {"axial_E_rms": v["axial"]["E_rms"], "learned_E_rms": v["learned"]["E_rms"], "axial_H_rms": v["axial"]["H_rms"],
                 "learned_H_rms": v["learned"]["H_rms"]} for k, v in val.items()}
    out = {"model": "MLP 2-128-128-128-128-5 SiLU, BF16 MFMA hidden layers, FP32 master weights", "samples": n_samples,
           "train_steps": steps, "batch": 16384, "label_s": t_gen, "train_s": sur.train_seconds, "loss_history": hist,
           "field_error_by_radius": bands, "n": len(P), "ms": res["ms"], "electrons_per_s": len(P) / (res["ms"] * 1e-3),
           "vs_gpu_bem": landing_error(res, ctx["ref"]), "launch_shape": shape_str(shape),
           "parked": res["parked"],   # finished on the stage-synchronous tracer (ring-source sums)
           "tuning_ms": {shape_str(k): round(v, 3) for k, v in tuning.items()}}
    # Off-axis beam: RK45's own error at the workload's atol = 1e-8 is about 400 nm here (the
    # oracle's landings at 1e-8 and 1e-10 differ by that much), so the fields are compared at
    # atol = 1e-11, against the GPU's direct FP64 BEM tracer, which equals the oracle's direct
    # BEM traces (at 1e-8 and 1e-10) to 1e-14 m on these electrons.
    f = golden("config_c4.npz")
    Po, Vo, bo, po = configs.magnetic_beam(64, seed=4, r_range=(0.3 * mm, 0.5 * mm))
    fine = 1e-11
    off = {"n": 64, "start_radius_mm": [0.3, 0.5], "atol": fine}
    if f is not None:
        off["gpu_bem_vs_oracle_bem_atol_1e-8"] = landing_error(traces(field.get_tracer(bo), Po, Vo, po), oracle_traces(f, "bem_"))
    ref = traces(field.get_tracer(bo), Po, Vo, po, atol=fine)
    strict = FieldRadialAxialLearned(sur, error_tolerance=3e-7)
    for name, fld in (("axial", axial), ("learned", learned), ("learned_check_3e-7", strict)):
        off[name] = landing_error(traces(fld.get_tracer(bo), Po, Vo, po, atol=fine), ref)
    out["off_axis_beam_vs_direct_bem"] = off
    return out


def cpu_baseline(lines, types, values, M64, F, trace_info, threads):
    """The reference's C backend (oracle/_ref) on bounded samples, on `threads` host cores:
    fill of a contiguous block of rows split over threads (util.split_collect, extrapolated to
    all rows); scipy QUADPACK diagonal on a sample of elements (extrapolated); numpy.linalg.solve
    (LAPACK dgesv, the reference's call) on the full N x N matrix of the workload, copied from the
    device (equal to the reference's fill to 1e-13), timed whole.  Result: seconds for the N =
    20,000 fill+solve."""
    from oracle.oracle import reference, restatement
    lib = reference()
    kind = "reference"
    if lib is None:
        lib, kind = restatement(), "port"
    N = len(lines)
    jac, pos = lib.fill_jacobian_buffer_radial(lines)
    rows = max(threads * 8, 256)
    M = np.zeros((rows, N))
    import threading
    chunks = [c for c in np.array_split(np.arange(rows), threads) if len(c)]

    def work(c):
        lib.lib.fill_matrix_radial(M, lines, types, values, jac, pos, N, N, int(c[0]), int(c[-1]))

    t0 = time.perf_counter()
    ts = [threading.Thread(target=work, args=(c,)) for c in chunks]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    t_fill = (time.perf_counter() - t0) * N / rows
    nd = 200
    t0 = time.perf_counter()
    for i in np.linspace(0, N - 1, nd).astype(int):
        lib.self_potential_quad(lines[i])
    t_diag = (time.perf_counter() - t0) * N / nd
    A = M64.cpu().numpy()
    t0 = time.perf_counter()
    np.linalg.solve(A, F)
    t_solve = time.perf_counter() - t0
    del A
    out = {"fill_solve_s": t_fill + t_diag + t_solve, "fill_s": t_fill, "diag_s": t_diag, "solve_s": t_solve, "kind": kind,
           "sample": f"fill {rows} of {N} rows over {threads} threads (extrapolated), QUADPACK diagonal {nd} of {N} elements "
                     f"(1 thread, as solver.py:239-247; extrapolated), numpy.linalg.solve of the full {N} x {N} matrix "
                     f"(OpenBLAS, timed whole)"}
    if trace_info is not None:
        # tracing: the reference C tracer through its axial field function, one particle per call
        ax = trace_info["axial"]
        P, V = trace_info["P"], trace_info["V"]
        ns = 64
        t0 = time.perf_counter()

        def tw(idx):
            for i in idx:
                lib.trace_axial(P[i], V[i], -1.602176634e-19 / 9.1093837015e-31, ax.z, ax.electrostatic_coeffs,
                                ax.magnetostatic_coeffs, trace_info["bounds"])

        ts = [threading.Thread(target=tw, args=(c,)) for c in np.array_split(np.arange(ns), threads) if len(c)]
        for t in ts:
            t.start()
        for t in ts:
            t.join()
        out["electrons_per_s_axial"] = ns / (time.perf_counter() - t0)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--N", type=int, default=20000)
    ap.add_argument("--particles", type=int, default=1_000_000)
    ap.add_argument("--sample", type=int, default=1024)
    ap.add_argument("--n-axial", type=int, default=7000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-trace", action="store_true")
    ap.add_argument("--no-field", action="store_true")
    ap.add_argument("--no-c3", action="store_true")
    ap.add_argument("--no-c4", action="store_true")
    ap.add_argument("--trace-only", action="store_true",
                    help="run only the C2 and C3 tracing legs (for kernel profiles); prints their JSON")
    ap.add_argument("--trace-shape", default=None, metavar="SHAPE",
                    help="fixed tracer launch shape instead of autotuning, e.g. ordered,batch=1,occ=2 (for kernel profiles)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first copy MIB MiB device-to-device three times (a known-size kernel for HBM counter scaling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the fill, the factorisation and the tracers here use one device; "
              "there is no multi-GPU path", file=sys.stderr)
        sys.exit(2)
    import torch
    assert torch.cuda.is_available(), "bench.py needs the GPU"
    if args.copy_calibration > 0:
        src = torch.ones(args.copy_calibration << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"bench: copy calibration, 3 x {args.copy_calibration} MiB device-to-device")
        del src, dst
    global TRACE_SHAPE
    if args.trace_shape:
        TRACE_SHAPE = parse_shape(args.trace_shape)
    if args.trace_only:
        c2, _ = tracing_leg(args.particles, args.sample, args.n_axial)
        c3, _ = c3_leg(args.particles, args.sample, args.n_axial)
        print(json.dumps({"tracing": c2, "c3_magnetic_lens": c3}))
        return
    lines, types, values = mirror_problem(args.N)
    N = len(lines)
    log(f"bench: mirror N={N}, warmup {args.warmup}, steps {args.steps}")
    for _ in range(args.warmup):
        fill_solve(lines, types, values, "mixed")
    from traceon_amd.device import KernelTiming
    times, tfs = [], []
    for k in range(args.steps):
        torch.cuda.synchronize()
        X, stats, t = fill_solve(lines, types, values, "mixed")
        times.append(t["total_s"])
        tfs.append(t)
        log(f"bench: step {k}: fill+solve {t['total_s']:.3f} s (fill {t['fill_ms']:.1f} ms, solve {t['solve_ms']:.1f} ms) {stats}")
    # one more step with HIP events around every GEMM launch of the factorisation (the events
    # add work between launches, so this step is not one of the timed ones)
    gt = KernelTiming()
    _, _, tg = fill_solve(lines, types, values, "mixed", gemm_timing=gt)
    log(f"bench: instrumented step {tg['total_s']:.3f} s, {gt.launches} GEMM launches, {gt.ms:.1f} ms in GEMMs")
    X64, stats64, t64 = fill_solve(lines, types, values, "fp64")
    log(f"bench: fp64 path {t64['total_s']:.3f} s {stats64}")
    dq = float((X - X64).abs().max().item() / X64.abs().max().item())
    fill_ms, fill_flops = fill_kernel_roofline(lines, types, values)
    log(f"bench: fill kernel {fill_ms:.2f} ms, {fill_flops / fill_ms / 1e9:.1f} TFLOP/s")
    result = {
        "metric": "electrons traced/s @<=10nm landing error; BEM fill+solve time at N=20k",
        "value": float(np.median(times)), "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": float(np.median(times) * 1e3), "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 (fill, residuals) + f32 (LU: split-BF16 MFMA bulk updates)", "data": "synthetic geometry (seeded), no external data",
        "config": {"workload": f"electrostatic mirror, 5 electrodes, N={N} line elements: fill + FP32-LU + FP64 refinement",
                   "N": N, "parallelism": "single GPU"},
        "fill_solve_fp64_s": t64["total_s"], "fill_ms": float(np.median([t["fill_ms"] for t in tfs])),
        "solve_ms": float(np.median([t["solve_ms"] for t in tfs])), "refine_iterations": stats.iterations,
        "charge_rel_diff_mixed_vs_fp64": dq,
    }
    # dominant kernel of the step: the bulk-update GEMM of the FP32 factorisation (panel groups
    # of 4 and 2, K = 512 and 256), timed with HIP events around each of its launches on the
    # stream it runs on.  Its FP32 operands are split into three BF16 parts and multiplied on
    # BF16 MFMA, six products per FP32 product: the algorithm's unit, one FP32 multiply-add of
    # the update, is 12 BF16 FLOP.  Per HBM byte (A and B once, C read and written) that is
    # 6 x 2K / (8 + 8K/n) BF16 FLOP, about 690 on average here, above the BF16 ridge point of
    # 2.5 PFLOP/s / 8 TB/s = 312: the bound is the MFMA rate (below the ridge, HBM).
    gemm_s_avg = gt.ms / max(gt.launches, 1) * 1e-3
    gemm_flop_avg = gt.flops / max(gt.launches, 1)
    gemm_bytes_avg = gt.bytes / max(gt.launches, 1)
    bf16_flop_avg = 6 * gemm_flop_avg
    mfma_bound = bf16_flop_avg / max(gemm_bytes_avg, 1.0) > PEAK["bf16"] / PEAK["hbm"]
    roof = {"kernel": "k_gemm_x6 (LU bulk updates: FP32 operands split into three BF16 parts, "
                      "six v_mfma_f32_16x16x32_bf16 per 16x16x32 block)",
            "bound": "mfma" if mfma_bound else "hbm",
            "achieved": bf16_flop_avg / gemm_s_avg / 1e12 if mfma_bound else gemm_bytes_avg / gemm_s_avg / 1e9,
            "peak": PEAK["bf16"] / 1e12 if mfma_bound else PEAK["hbm"] / 1e9, "unit": "TFLOP/s" if mfma_bound else "GB/s",
            "traffic": GEMM_X6_TRAFFIC_N20000 if N == 20000 else None, "traffic_unit": "bytes per launch",
            "algorithmic_bytes_per_launch": gemm_bytes_avg, "launches": gt.launches,
            "flop_per_launch_fp32": gemm_flop_avg, "flop_per_launch_bf16": bf16_flop_avg,
            "launch_ms": gemm_s_avg * 1e3, "hbm_algorithmic_gbps": gemm_bytes_avg / gemm_s_avg / 1e9,
            "fp32_equivalent_tflops": gemm_flop_avg / gemm_s_avg / 1e12, "fp32_mfma_peak_tflops": PEAK["fp32"] / 1e12,
            "share_of_step": gt.ms / max(tg["fill_ms"] + tg["solve_ms"], 1e-9)}
    roof["frac"] = roof["achieved"] / roof["peak"]
    result["roofline"] = roof
    rf = {"kernel": "k_fill (dense FP64 fill)", "bound": "fp64-valu", "achieved": fill_flops / (fill_ms * 1e-3) / 1e12,
          "peak": PEAK["fp64"] / 1e12, "unit": "TFLOP/s", "flop_per_launch": fill_flops, "launch_ms": fill_ms}
    rf["frac"] = rf["achieved"] / rf["peak"]
    # the fill issues more VALU instructions than FLOPs: 82 per potential pair in the gfx950
    # assembly of its branch-free loop (register copies, conversions, compares besides the FP64
    # arithmetic), each wave64 instruction 4 cycles on a 16-lane SIMD
    if fill_flops == float(N) * N * 16.0 * C_POT:   # potential rows only (the mirror)
        issue_peak = PEAK_CUS * 4 * 16 * PEAK_CLOCK_HZ
        rf["valu_instr_per_pair"] = FILL_VALU_PER_PAIR
        rf["valu_issue_frac"] = FILL_VALU_PER_PAIR * N * N * 16.0 / (fill_ms * 1e-3) / issue_peak
    result["roofline_fill"] = rf
    if not args.no_field:
        ms, pairs, P, Fd, pts = field_grid_leg(X, lines, types, values)
        result["field_grid"] = {"points": int(pts.shape[0]), "ms": ms, "points_per_s": pts.shape[0] / (ms * 1e-3),
                                "tflops": pairs * (C_POT + C_FIELD) / (ms * 1e-3) / 1e12}
        log(f"bench: field grid {pts.shape[0]} points in {ms:.1f} ms")
    trace_info = None
    if not args.no_trace:
        c2, trace_info = tracing_leg(args.particles, args.sample, args.n_axial)
        result["tracing"] = c2
        result["electrons_per_s"] = c2["electrons_per_s"]
        log(f"bench: C2 traced {c2['n']} electrons in {c2['ms']:.1f} ms ({c2['electrons_per_s']:.3e}/s), reflected "
            f"{c2['reflected_fraction']:.3f}, lengths max/min {c2['len_ratio_max_min']:.1f}; vs oracle BEM "
            f"{c2.get('vs_oracle_bem')}; vs GPU BEM {c2['vs_gpu_bem']}; roofline {c2['roofline']['frac']:.3f}")
    if not args.no_c3:
        c3, ctx = c3_leg(args.particles, args.sample, args.n_axial)
        result["c3_magnetic_lens"] = c3
        log(f"bench: C3 solve {c3['solve_s']:.3f} s, traced {c3['n']} electrons in {c3['ms']:.1f} ms "
            f"({c3['electrons_per_s']:.3e}/s); vs oracle BEM {c3.get('vs_oracle_bem')}; vs GPU BEM {c3['vs_gpu_bem']}")
        if not args.no_c4:
            c4 = c4_leg(ctx)
            c4["speedup_vs_direct_bem_fp64"] = c4["electrons_per_s"] / c3["electrons_per_s_direct_bem_fp64"]
            result["c4_surrogate"] = c4
            log(f"bench: C4 surrogate labels {c4['label_s']:.2f} s, training {c4['train_s']:.2f} s, traced {c4['n']} electrons "
                f"in {c4['ms']:.1f} ms ({c4['electrons_per_s']:.3e}/s); vs GPU BEM {c4['vs_gpu_bem']}; off-axis beam "
                f"{c4.get('off_axis_beam_vs_direct_bem')}")
    if not args.no_cpu:
        from traceon_amd import device
        threads = min(16, len(os.sched_getaffinity(0)))
        E = device.Elements(lines, types, values)
        M64, _ = device.assemble(E)
        F = np.where(np.isin(types, [1, 2, 5]), values, 0.0)
        cb = cpu_baseline(lines, types, values, M64, F, trace_info, threads)
        del M64, E
        result["cpu_baseline"] = {"value": cb["fill_solve_s"], "unit": "s", "cores": threads, "kind": cb["kind"],
                                  "sample": cb["sample"], "fill_s": cb["fill_s"], "diag_s": cb["diag_s"], "solve_s": cb["solve_s"]}
        if "electrons_per_s_axial" in cb:
            result["cpu_baseline"]["electrons_per_s_axial"] = cb["electrons_per_s_axial"]
        result["speedup_vs_cpu"] = cb["fill_solve_s"] / result["value"]
        log(f"bench: cpu baseline {cb}")
    print(json.dumps(result))


if __name__ == "__main__":
    main()
