"""Launch-shape autotuner of the persistent tracer (tb_trace).

The tracer's speed depends on how its queue is drained, never its results: whether the
particles are handed out in the order tb_trace_order sorts them (the lanes of a wave then
travel together and share the axial coefficients they load), whether a wave refills lanes
one by one or takes 64 particles at once, and how many waves per SIMD the axial-only kernel
is compiled for (fewer registers per wave against spills).  ``tune`` times each candidate
on the caller's own particles with HIP events on the launch stream and keeps the fastest
per (field kind, particle-count bucket, atol); ``Tracer.trace_batch(shape="auto")`` then
uses it.
"""
from __future__ import annotations

import math

import numpy as np

from .tracing import DEFAULT_SHAPE, E_CHARGE, M_E, LaunchShape, batch_state, run_trace

_cache: dict = {}


def _kind(program) -> str:
    return "learned" if program.has_mlp else ("sources" if program.has_sources else "axial")


def _key(program, n: int, atol: float):
    return (_kind(program), int(math.log2(max(n, 1))), float(atol))


def candidates(program) -> list[LaunchShape]:
    """The shapes worth timing for this field: occupancy applies to the axial-only kernel."""
    occ = (1, 2, 3, 4) if _kind(program) == "axial" else (2,)
    out = [LaunchShape(ordered=True, batch=b, occupancy=o) for o in occ for b in (0, 1)]
    out += [LaunchShape(ordered=False, batch=0, occupancy=o) for o in occ]
    return out


def cached_shape(program, n: int, atol: float) -> LaunchShape:
    return _cache.get(_key(program, n, atol), DEFAULT_SHAPE)


def tune(tracer, positions, velocities, mass=M_E, charge=-E_CHARGE, atol=1e-8, plane=None, plane2=None,
         max_attempts=1 << 62, shapes=None):
    """Time every candidate shape (or `shapes`) on these particles; cache and return the
    fastest with the timings {shape: ms}."""
    import torch
    from . import device
    S0 = batch_state(positions, velocities)
    if not isinstance(S0, torch.Tensor):
        S0 = device.to_dev(S0)
    qm = np.asarray(charge, dtype=np.float64) / np.asarray(mass, dtype=np.float64)
    if qm.size > 1:
        qm = np.broadcast_to(qm, (len(S0),))
    prog = tracer.program()
    times = {}
    for sh in shapes or candidates(prog):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        r = run_trace(prog, S0, qm, tracer.bounds, atol, plane=plane, plane2=plane2, max_attempts=max_attempts, shape=sh)
        b.record()
        b.synchronize()
        del r
        times[sh] = a.elapsed_time(b)
    best = min(times, key=times.get)
    _cache[_key(prog, len(S0), atol)] = best
    return best, times
