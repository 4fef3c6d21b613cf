"""Benchmark of the radial BEM hot path on one MI355X.

Workload (one "step"): the 5-electrode electrostatic mirror meshed into N = 20,000 cubic
line elements, voltages U[-2, 2] kV from seed 1: dense fill (FP64 + FP32 copy) + singular
diagonal + LU of the FP32 copy on MFMA + FP64 double-double iterative refinement, i.e. the
BEM fill+solve of the metric.  Reported beside it in the same JSON line:

* the FP64 path (FP64 LU) time and the charge difference of the two paths;
* the field (potential + E) on a 1000 x 1000 (r, z) grid (10^6 points) from the charges;
* batched tracing: the 4-electrode lens (N = 8,000) with a -1.1 kV reflecting electrode,
  10^6 electrons (lognormal energy, median 1 keV, sigma 0.5; disc of 0.2 mm diameter;
  5 mrad angles) through the FieldRadialAxial interpolation, landing RMS error against
  the direct FP64 BEM tracer on a seeded sample;
* roofline of the dominant kernel (HIP events on the launch stream);
* cpu_baseline: the reference's own C backend (oracle/_ref, compiled from its sources)
  on bounded samples of the same work, on this host's cores.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-trace] [--no-field]
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
# HBM bytes per launch of the LU's bulk-update GEMM (k_gemm_x6) in this workload
# (N = 20,000): rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE averaged over its 59 launches in one
# factorisation, each scaled by what the counter reads for a 256 MiB device-to-device copy
# (--copy-calibration 256): fetch x2.0, write x1.0.  Counters cannot be read inside this
# process, so the figure is the one measured for this kernel and launch schedule; other N
# report null.
GEMM_X6_TRAFFIC_N20000 = 920.3e6
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


def tracing_leg(n_particles, n_sample, n_axial=7000, seed=2):
    """Batched tracing through the 4-electrode lens (field from a GPU solve)."""
    import torch
    from traceon_amd import configs, solver, tracing
    from traceon_amd.excitation import Excitation
    from traceon_amd.field import FieldRadialAxial
    mm = 1e-3
    mesh, volts = configs.einzel_lens(8000, voltages=(0.0, 1800.0, 0.0), fourth=(4.0 * mm, 1.0 * mm, -1100.0))
    exc = Excitation(mesh)
    exc.add_voltage(**volts)
    field = solver.solve_direct(exc)
    bounds = np.array([[-1.0 * mm, 1.0 * mm], [-1.0 * mm, 1.0 * mm], [-5.0 * mm, 10.0 * mm]])
    t0 = time.perf_counter()
    # The interpolation covers the tracing bounds plus the longest step (hmax |v| = 1 cm,
    # tracing.c:68-70) on both sides: RK stage points of a step can lie that far ahead, and
    # FieldRadialAxial is zero outside its range (radial.c:313-384).  5 um sample spacing.
    axial = FieldRadialAxial(field, -15.0 * mm, 20.0 * mm, N=n_axial)
    torch.cuda.synchronize()
    t_axial_build = time.perf_counter() - t0
    rng = np.random.default_rng(seed)
    n = n_particles
    energy = 1000.0 * np.exp(0.5 * rng.standard_normal(n))
    rad = 0.1 * mm * np.sqrt(rng.random(n))       # uniform in a disc of 0.2 mm diameter
    phi = 2 * np.pi * rng.random(n)
    tx, ty = rng.normal(0, 5e-3, n), rng.normal(0, 5e-3, n)
    d = np.column_stack([np.tan(tx), np.tan(ty), np.ones(n)])
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    speed = np.sqrt(2 * energy * tracing.E_CHARGE / tracing.M_E)
    P = np.column_stack([rad * np.cos(phi), rad * np.sin(phi), np.full(n, -5.0 * mm)])
    V = d * speed[:, None]
    plane = (np.array([0.0, 0.0, 9.5 * mm]), np.array([0.0, 0.0, 1.0]))
    tr_ax = axial.get_tracer(bounds)
    warm = tr_ax.trace_batch(P[:4096], V[:4096], plane=plane)
    torch.cuda.synchronize()
    with Timer() as t:
        res = tr_ax.trace_batch(P, V, plane=plane)
    steps = int(res.n_attempted.sum().item())
    land = res.landing.cpu().numpy()
    reflected = float(np.mean(res.final[:, 2].cpu().numpy() < -4.999 * mm))
    nacc = res.n_accepted.cpu().numpy()
    # error reference: direct FP64 BEM field, same particles (sample)
    tr_bem = field.get_tracer(bounds)
    with Timer() as tb:
        ref = tr_bem.trace_batch(P[:n_sample], V[:n_sample], plane=plane)
    land_ref = ref.landing.cpu().numpy()
    ok = np.isfinite(land_ref[:, 0]) & np.isfinite(land[:n_sample, 0])
    err = np.linalg.norm(land[:n_sample][ok, :2] - land_ref[ok, :2], axis=1)
    steps_bem = int(ref.n_attempted.sum().item())
    return {
        "n": n, "ms": t.ms, "steps": steps, "reflected_fraction": reflected,
        "len_ratio": float(np.percentile(nacc, 99.9) / max(np.percentile(nacc, 0.1), 1)),
        "landing_rms_nm": float(np.sqrt(np.mean(err ** 2)) * 1e9) if ok.any() else None,
        "landing_max_nm": float(np.max(err) * 1e9) if ok.any() else None,
        "n_sample": int(n_sample), "bem_ms": tb.ms, "bem_steps": steps_bem, "n_axial": len(axial.z),
        "axial_build_s": t_axial_build, "field": field, "axial": axial, "bounds": bounds, "P": P, "V": V, "plane": plane,
        "N_elements": len(field.electrostatic_point_charges),
    }


def c3_leg(n_particles, n_sample, n_axial=7000, seed=3):
    """Magnetic + electrostatic lens (C3): coil 500 A-turn, two mu_r = 1000 pole pieces (bore 2 mm,
    gap 3 mm) and a three-tube einzel lens, N = 12,000; electrostatic and magnetostatic solves,
    10 keV electrons traced through the combined field with the Lorentz force."""
    import torch
    from traceon_amd import configs, solver, tracing
    from traceon_amd.excitation import Excitation
    from traceon_amd.field import FieldRadialAxial
    mm = 1e-3
    mesh, volts, perms, currents = configs.magnetic_lens(12000)
    exc = Excitation(mesh)
    exc.add_voltage(**volts)
    exc.add_magnetizable(**perms)
    exc.add_current(**currents)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    field = solver.solve_direct(exc)
    torch.cuda.synchronize()
    t_solve = time.perf_counter() - t0
    t0 = time.perf_counter()
    axial = FieldRadialAxial(field, -24.0 * mm, 22.0 * mm, N=int(n_axial * 46 / 35))
    torch.cuda.synchronize()
    t_axial = time.perf_counter() - t0
    rng = np.random.default_rng(seed)
    n = n_particles
    rad = 0.1 * mm * np.sqrt(rng.random(n))
    phi = 2 * np.pi * rng.random(n)
    tx, ty = rng.normal(0, 5e-3, n), rng.normal(0, 5e-3, n)
    d = np.column_stack([np.tan(tx), np.tan(ty), np.ones(n)])
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    speed = np.sqrt(2 * 10000.0 * tracing.E_CHARGE / tracing.M_E)
    P = np.column_stack([rad * np.cos(phi), rad * np.sin(phi), np.full(n, -14.0 * mm)])
    V = d * speed
    bounds = np.array([[-1.0 * mm, 1.0 * mm], [-1.0 * mm, 1.0 * mm], [-14.0 * mm, 12.0 * mm]])
    plane = (np.array([0.0, 0.0, 11.0 * mm]), np.array([0.0, 0.0, 1.0]))
    tr_ax = axial.get_tracer(bounds)
    tr_ax.trace_batch(P[:4096], V[:4096], plane=plane)
    torch.cuda.synchronize()
    with Timer() as t:
        res = tr_ax.trace_batch(P, V, plane=plane)
    land = res.landing.cpu().numpy()
    with Timer() as tb:
        ref = field.get_tracer(bounds).trace_batch(P[:n_sample], V[:n_sample], plane=plane)
    land_ref = ref.landing.cpu().numpy()
    ok = np.isfinite(land_ref[:, 0]) & np.isfinite(land[:n_sample, 0])
    err = np.linalg.norm(land[:n_sample][ok, :2] - land_ref[ok, :2], axis=1)
    out = {"N_electrostatic": len(field.electrostatic_point_charges), "N_magnetostatic": len(field.magnetostatic_point_charges),
           "N_coil_rings": len(field.current_point_charges), "solve_s": t_solve, "axial_build_s": t_axial, "n": n,
           "ms": t.ms, "electrons_per_s": n / (t.ms * 1e-3), "steps": int(res.n_attempted.sum().item()),
           "landed_fraction": float(np.mean(np.isfinite(land[:, 0]))),
           "landing_rms_nm": float(np.sqrt(np.mean(err ** 2)) * 1e9) if ok.any() else None,
           "landing_max_nm": float(np.max(err) * 1e9) if ok.any() else None, "n_sample": int(n_sample),
           "electrons_per_s_direct_bem_fp64": n_sample / (tb.ms * 1e-3)}
    ctx = {"field": field, "axial": axial, "P": P, "V": V, "bounds": bounds, "plane": plane, "land_ref": land_ref}
    return out, ctx


def c4_leg(ctx, n_samples=1 << 22, steps=8000):
    """Learned surrogate (C4) in the C3 geometry: the MLP (2-128x4-5, SiLU, BF16 MFMA, FP32
    master weights) learns BEM - axial from n_samples BEM-labelled (r, z) points over the lens
    region; the tracer adds it to the axial field inside r <= 0.6 mm and uses the direct BEM
    sum outside.  Reported: build times, field error per radial band with and without the
    correction, and 10 keV tracing of the same electrons as C3 against direct FP64 BEM."""
    import torch
    from traceon_amd.surrogate import FieldRadialAxialLearned, FieldSurrogate
    mm = 1e-3
    field, axial, P, V = ctx["field"], ctx["axial"], ctx["P"], ctx["V"]
    bounds, plane, land_ref = ctx["bounds"], ctx["plane"], ctx["land_ref"]
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
    tr.trace_batch(P[:4096], V[:4096], plane=plane)
    torch.cuda.synchronize()
    with Timer() as t:
        res = tr.trace_batch(P, V, plane=plane)
    land = res.landing.cpu().numpy()
    ns = len(land_ref)
    ok = np.isfinite(land_ref[:, 0]) & np.isfinite(land[:ns, 0])
    err = np.linalg.norm(land[:ns][ok, :2] - land_ref[ok, :2], axis=1)
    bands = {k: {"axial_E_rms": v["axial"]["E_rms"], "learned_E_rms": v["learned"]["E_rms"], "axial_H_rms": v["axial"]["H_rms"],
                 "learned_H_rms": v["learned"]["H_rms"]} for k, v in val.items()}
    return {"model": "MLP 2-128-128-128-128-5 SiLU, BF16 MFMA hidden layers, FP32 master weights", "samples": n_samples,
            "train_steps": steps, "batch": 16384, "label_s": t_gen, "train_s": sur.train_seconds, "loss_history": hist,
            "field_error_by_radius": bands, "n": len(P), "ms": t.ms, "electrons_per_s": len(P) / (t.ms * 1e-3),
            "landing_rms_nm": float(np.sqrt(np.mean(err ** 2)) * 1e9) if ok.any() else None,
            "landing_max_nm": float(np.max(err) * 1e9) if ok.any() else None}


def cpu_baseline(lines, types, values, trace_info, threads):
    """The reference's C backend (oracle/_ref) on bounded samples, on `threads` host cores.
    Fill: a contiguous block of rows split over threads (util.split_collect); diagonal:
    scipy QUADPACK on a sample of elements; solve: numpy.linalg.solve at a smaller N
    scaled by N^3.  Result: estimated seconds for the full N = 20,000 fill+solve."""
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
    Ns = 6000
    A = np.random.default_rng(0).standard_normal((Ns, Ns)) + Ns * np.eye(Ns)
    b = np.ones(Ns)
    t0 = time.perf_counter()
    np.linalg.solve(A, b)
    t_solve = (time.perf_counter() - t0) * (N / Ns) ** 3
    out = {"fill_solve_s": t_fill + t_diag + t_solve, "fill_s": t_fill, "diag_s": t_diag, "solve_s": t_solve, "kind": kind,
           "sample": f"fill {rows} of {N} rows over {threads} threads, QUADPACK diagonal {nd} of {N} elements (1 thread, as "
                     f"solver.py:239-247), numpy.linalg.solve at N={Ns} scaled by (N/{Ns})^3"}
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
    result["roofline_fill"] = rf
    if not args.no_field:
        ms, pairs, P, Fd, pts = field_grid_leg(X, lines, types, values)
        result["field_grid"] = {"points": int(pts.shape[0]), "ms": ms, "points_per_s": pts.shape[0] / (ms * 1e-3),
                                "tflops": pairs * (C_POT + C_FIELD) / (ms * 1e-3) / 1e12}
        log(f"bench: field grid {pts.shape[0]} points in {ms:.1f} ms")
    trace_info = None
    if not args.no_trace:
        trace_info = tracing_leg(args.particles, args.sample, args.n_axial)
        eps = trace_info["n"] / (trace_info["ms"] * 1e-3)
        result["tracing"] = {k: trace_info[k] for k in ("n", "ms", "steps", "reflected_fraction", "len_ratio", "landing_rms_nm",
                                                        "landing_max_nm", "n_sample", "n_axial", "N_elements")}
        result["tracing"]["electrons_per_s"] = eps
        result["tracing"]["electrons_per_s_direct_bem_fp64"] = trace_info["n_sample"] / (trace_info["bem_ms"] * 1e-3)
        result["electrons_per_s"] = eps
        log(f"bench: traced {trace_info['n']} electrons in {trace_info['ms']:.1f} ms ({eps:.3e}/s), "
            f"landing RMS {trace_info['landing_rms_nm']} nm vs direct BEM")
    if not args.no_c3:
        c3, ctx = c3_leg(args.particles, args.sample, args.n_axial)
        result["c3_magnetic_lens"] = c3
        log(f"bench: C3 solve {c3['solve_s']:.3f} s, traced {c3['n']} electrons in {c3['ms']:.1f} ms "
            f"({c3['electrons_per_s']:.3e}/s), landing RMS {c3['landing_rms_nm']} nm vs direct BEM")
        if not args.no_c4:
            c4 = c4_leg(ctx)
            c4["speedup_vs_direct_bem_fp64"] = c4["electrons_per_s"] / c3["electrons_per_s_direct_bem_fp64"]
            result["c4_surrogate"] = c4
            log(f"bench: C4 surrogate labels {c4['label_s']:.2f} s, training {c4['train_s']:.2f} s, traced {c4['n']} electrons "
                f"in {c4['ms']:.1f} ms ({c4['electrons_per_s']:.3e}/s), landing RMS {c4['landing_rms_nm']} nm vs direct BEM")
    if not args.no_cpu:
        threads = min(16, len(os.sched_getaffinity(0)))
        cb = cpu_baseline(lines, types, values, trace_info, threads)
        result["cpu_baseline"] = {"value": cb["fill_solve_s"], "unit": "s", "cores": threads, "kind": cb["kind"],
                                  "sample": cb["sample"], "fill_s": cb["fill_s"], "diag_s": cb["diag_s"], "solve_s": cb["solve_s"]}
        if "electrons_per_s_axial" in cb:
            result["cpu_baseline"]["electrons_per_s_axial"] = cb["electrons_per_s_axial"]
        result["speedup_vs_cpu"] = cb["fill_solve_s"] / result["value"]
        log(f"bench: cpu baseline {cb}")
    print(json.dumps(result))


if __name__ == "__main__":
    main()
