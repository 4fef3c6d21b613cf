"""Particle tracing — counterpart of traceon/tracing.py.

``Tracer.__call__`` (= ``trace_single``) keeps the reference signature and returns the full
trajectory ``(times (T,), states (T,6))``.  For FieldRadialBEM and FieldRadialAxial it goes
through the reference-named ``trace_particle`` entry point with the field's C function and
argument struct, exactly as traceon/tracing.py:141-184 does; for a FieldSuperposition (whose
reference path is a Python callback per RK stage) it runs the GPU tracer on a device field
program.  Additions: ``Tracer.trace_batch`` traces many particles in one persistent launch
and returns final states, landing states on a plane and step counts.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from math import sqrt, sin, cos

import numpy as np

from . import backend

M_E = 9.1093837015e-31
E_CHARGE = 1.602176634e-19


def _convert_velocity_to_SI(velocity, mass):
    speed_eV = np.linalg.norm(velocity)
    speed = sqrt(2 * speed_eV * E_CHARGE / mass)
    return speed * np.array(velocity) / speed_eV


def velocity_vec(eV: float, direction, mass: float = M_E):
    """tracing.py:32-56: velocity (m/s) of a particle with kinetic energy eV along direction."""
    assert eV > 0.0, "Please provide a positive energy in eV"
    d = np.array(direction, dtype=np.float64)
    assert d.shape == (3,)
    return _convert_velocity_to_SI(eV * d / np.linalg.norm(d), mass)


def velocity_vec_spherical(eV: float, theta: float, phi: float):
    return velocity_vec(eV, [sin(theta) * cos(phi), sin(theta) * sin(phi), cos(theta)])


def velocity_vec_xz_plane(eV: float, angle: float, downward: bool = True):
    sign = -1 if downward else 1
    return velocity_vec(eV, [sin(angle), 0.0, sign * cos(angle)])


# ---------------------------------------------------------------------------------------
# ctypes mirrors of the Tier B tracer structs (include/traceon_hip.h)
# ---------------------------------------------------------------------------------------
class SourcesT(C.Structure):
    _fields_ = [("w", C.c_void_p), ("r", C.c_void_p), ("z", C.c_void_p), ("n", C.c_int), ("pad", C.c_int)]


class FieldCompT(C.Structure):
    _fields_ = [("kind", C.c_int), ("transform", C.c_int), ("bounds_mode", C.c_int), ("pad", C.c_int), ("factor", C.c_double),
                ("inv", C.c_double * 12), ("basis", C.c_double * 9), ("bounds", C.c_double * 6),
                ("elec", SourcesT), ("mag", SourcesT), ("cur", SourcesT),
                ("zs", C.c_void_p), ("ce", C.c_void_p), ("cm", C.c_void_p), ("nz", C.c_int), ("axial_parts", C.c_int),
                ("uE", C.c_double * 3), ("uH", C.c_double * 3),
                ("mlp", C.c_void_p), ("mlp_zs", C.c_void_p), ("mlp_nz", C.c_int), ("pad3", C.c_int), ("mlp_geom", C.c_double * 6),
                ("mlp_wbf", C.c_void_p)]


class FieldProgramT(C.Structure):
    _fields_ = [("comp", C.c_void_p), ("ncomp", C.c_int), ("bounds_mode", C.c_int), ("flags", C.c_int), ("pad", C.c_int),
                ("bounds", C.c_double * 6)]


PROGRAM_HAS_MLP, PROGRAM_NO_SOURCES, PROGRAM_LEARNED_ONLY = 1, 2, 4


class TraceParamsT(C.Structure):
    _fields_ = [("field", FieldProgramT), ("state0", C.c_void_p), ("t0", C.c_void_p), ("h0", C.c_void_p), ("qm", C.c_void_p),
                ("qm_stride", C.c_int), ("n", C.c_int), ("counter", C.c_void_p), ("bounds", C.c_double * 6), ("atol", C.c_double),
                ("max_attempts", C.c_longlong), ("plane_p", C.c_double * 3), ("plane_n", C.c_double * 3), ("plane_nn", C.c_double),
                ("final_state", C.c_void_p), ("landing", C.c_void_p), ("status", C.c_void_p), ("n_accepted", C.c_void_p),
                ("n_attempted", C.c_void_p), ("t_final", C.c_void_p), ("h_final", C.c_void_p), ("traj", C.c_void_p),
                ("times", C.c_void_p), ("traj_cap", C.c_longlong), ("traj_len", C.c_void_p),
                ("plane2_p", C.c_double * 3), ("plane2_n", C.c_double * 3), ("plane2_nn", C.c_double), ("landing2", C.c_void_p),
                ("order", C.c_void_p), ("batch", C.c_int), ("occupancy", C.c_int),
                ("hmax0", C.c_void_p), ("hmax_final", C.c_void_p), ("park", C.c_int)]


COMP_KIND = {"bem": 0, "axial": 1, "uniform": 2, "learned": 1}
BOUNDS_NONE, BOUNDS_STRICT, BOUNDS_INCLUSIVE = 0, 1, 2
STATUS = {0: "left bounds", 1: "block full", 2: "step cap"}
STATUS_PARKED = 3   # tb_trace with park = 1 only; run_trace resumes such particles, so callers never see it


def _set_sources(dst: SourcesT, charges):
    if charges is None or len(charges) == 0:
        dst.w = dst.r = dst.z = None
        dst.n = 0
        return None
    s = charges.device_sources()
    dst.w, dst.r, dst.z, dst.n = s.w.data_ptr(), s.r.data_ptr(), s.z.data_ptr(), s.n
    return s


def make_component(field, factor: float, kind: str, in_superposition: bool):
    """One GPU field component.  A field traced on its own follows its C adaptor
    (strict bounds, no transform: tracing.c:125-174, field.py:814-816); inside a
    superposition it follows the Python Field API (affine map, inclusive bounds, basis
    rotation: field.py:272-329)."""
    c = FieldCompT()
    c.kind = COMP_KIND[kind]
    c.factor = float(factor)
    keep = []
    if in_superposition:
        c.transform = int(field.has_transform())
        inv = field._inverse_transformation_matrix[:3, :].ravel()
        for i in range(12):
            c.inv[i] = inv[i]
        for i in range(9):
            c.basis[i] = field._basis.ravel()[i]
        mode = BOUNDS_INCLUSIVE
    else:
        c.transform = 0
        mode = BOUNDS_STRICT if kind == "bem" else BOUNDS_NONE
    if field.field_bounds is not None and (in_superposition or kind == "bem"):
        c.bounds_mode = mode
        for i, v in enumerate(np.asarray(field.field_bounds, dtype=np.float64).ravel()):
            c.bounds[i] = v
    else:
        c.bounds_mode = BOUNDS_NONE
    if kind == "bem":
        keep.append(_set_sources(c.elec, field.electrostatic_point_charges))
        keep.append(_set_sources(c.mag, field.magnetostatic_point_charges))
        keep.append(_set_sources(c.cur, field.current_point_charges))
    elif kind in ("axial", "learned"):
        z, ce, cm = field.device_arrays()
        c.zs, c.ce, c.cm, c.nz = z.data_ptr(), ce.data_ptr(), cm.data_ptr(), len(field.z)
        # series that are identically zero (no charges of that kind) are not evaluated
        c.axial_parts = (1 if np.any(field.electrostatic_coeffs) else 0) | (2 if np.any(field.magnetostatic_coeffs) else 0)
        if c.axial_parts == 0:
            c.axial_parts = 3
        keep += [z, ce, cm]
        if kind == "learned":
            # the learned correction, and the ring sources for outside its trust region
            sur = field.surrogate
            wts, zs, geom = sur.tracer_arrays(field.error_tolerance)
            c.mlp, c.mlp_zs, c.mlp_nz = wts.data_ptr(), zs.data_ptr(), zs.shape[1]
            for i, v in enumerate(geom[:5]):
                c.mlp_geom[i] = v
            c.mlp_geom[5] = field.trust_radius
            c.mlp_wbf = sur.Wbf.data_ptr()
            keep += [wts, zs, sur.Wbf]
            keep.append(_set_sources(c.elec, sur.field.electrostatic_point_charges))
            keep.append(_set_sources(c.mag, sur.field.magnetostatic_point_charges))
            keep.append(_set_sources(c.cur, sur.field.current_point_charges))
    else:
        for i in range(3):
            c.uE[i] = field.E[i]
            c.uH[i] = field.H[i]
    return c, keep


class DeviceFieldProgram:
    """A field (or superposition) laid out for the GPU tracer; keeps its buffers alive."""

    def __init__(self, field):
        import torch
        from .field import FieldSuperposition
        comps = field.device_components(1.0)
        self.keep = [k for _, k in comps]
        # ring-source sums make a field evaluation cost ~10^5 pairs
        self.n_sources = int(sum(c.elec.n + c.mag.n + c.cur.n for c, _ in comps))
        self.has_sources = self.n_sources > 0
        self.has_mlp = any(c.mlp for c, _ in comps)
        n = len(comps)
        raw = (FieldCompT * n)(*[c for c, _ in comps])
        host = torch.frombuffer(bytearray(bytes(raw)), dtype=torch.uint8)
        self.comp_buf = host.to("cuda")
        self.prog = FieldProgramT()
        self.prog.comp = self.comp_buf.data_ptr()
        self.prog.ncomp = n
        # a program without ring sources runs the tracer variant without source-sum code
        self.prog.flags = (PROGRAM_HAS_MLP if self.has_mlp else 0) | (0 if self.has_sources else PROGRAM_NO_SOURCES)
        c0 = comps[0][0] if len(comps) == 1 else None
        if c0 is not None and c0.mlp and c0.mlp_wbf and not c0.transform:
            # one learned component: its sources serve only the fallback, which parks
            self.prog.flags |= PROGRAM_LEARNED_ONLY
        # the superposition's own bounds (field.py:285-290 on the top-level object)
        if isinstance(field, FieldSuperposition) and field.field_bounds is not None:
            self.prog.bounds_mode = BOUNDS_INCLUSIVE
            for i, v in enumerate(np.asarray(field.field_bounds).ravel()):
                self.prog.bounds[i] = v
        else:
            self.prog.bounds_mode = BOUNDS_NONE

    def evaluate(self, points):
        """(E, H) at points (M,3) through the field program."""
        import torch
        from . import device
        P = device.to_dev(np.atleast_2d(points))
        M = P.shape[0]
        E = torch.zeros((M, 3), dtype=torch.float64, device=P.device)
        H = torch.zeros_like(E)
        backend.check(backend.backend_lib.tb_eval_program(C.byref(self.prog), P.data_ptr(), M, E.data_ptr(), H.data_ptr(),
                                                          device.stream_ptr()), "eval_program")
        return E.cpu().numpy(), H.cpu().numpy()


class BatchResult:
    def __init__(self, final, landing, status, n_accepted, n_attempted, t_final, landing2=None):
        self.final = final
        self.landing = landing
        self.landing2 = landing2
        self.status = status
        self.n_accepted = n_accepted
        self.n_attempted = n_attempted
        self.t_final = t_final
        self.n_parked = 0   # particles the persistent tracer handed to the stage-synchronous one


# Below this many particles a field with ring sources is traced stage-synchronously
# (tb_trace_split: every stage's source sums spread over the chip); above it the
# persistent lane-per-particle kernel (tb_trace) has enough lanes to fill the GPU.
SPLIT_MAX_PARTICLES = 16384


@dataclass(frozen=True)
class LaunchShape:
    """How the persistent tracer drains its particle queue; no choice changes a result.
    ordered: hand particles out sorted by start z and start v_z (tb_trace_order), so that
    the lanes of a wave travel together; batch: a wave takes 64 particles at once when all
    its lanes are free; occupancy: waves per SIMD of the axial-only kernel (1-4);
    waves: workgroups of the launch (0: as many as the CUs hold at that occupancy)."""
    ordered: bool = True
    batch: int = 0
    occupancy: int = 2
    waves: int = 0


DEFAULT_SHAPE = LaunchShape()
# below this many particles the queue is not reordered (fewer particles than lanes)
ORDER_MIN_PARTICLES = 4096


def queue_order(S0):
    """tb_trace_order over a device state tensor (n, 6): the int32 queue order."""
    import torch
    from . import device
    n = S0.shape[0]
    L = backend.backend_lib
    order = torch.empty(n, dtype=torch.int32, device=S0.device)
    ws = torch.empty(L.tb_trace_order_workspace_bytes(n), dtype=torch.uint8, device=S0.device)
    backend.check(L.tb_trace_order(S0.data_ptr(), n, order.data_ptr(), ws.data_ptr(), device.stream_ptr()), "trace_order")
    return order


def run_trace(program: DeviceFieldProgram, state0, qm, bounds, atol=1e-8, plane=None, traj_cap=0, t0=None,
              max_attempts=1 << 62, waves=0, split=None, plane2=None, shape: LaunchShape | None = None,
              park: bool | None = None):
    """GPU tracer over many particles (all arrays numpy or device tensors).
    plane, plane2: (p0, normal); the state at each one's last crossing (NaN if none).
    split: None picks tb_trace_split for fields with ring sources (a learned field's are its
    fallback) and at most SPLIT_MAX_PARTICLES particles, else the persistent tb_trace.  shape: launch shape of
    the persistent tracer (None: DEFAULT_SHAPE); waves > 0 overrides shape.waves.  On the persistent
    tracer, particles of a learned field whose step needs the ring sources are parked there and
    finished on tb_trace_split (result.n_parked of them); park=False sums the sources in the
    persistent kernel instead (None: park where it applies)."""
    import torch
    from . import device
    P, out = trace_params(program, state0, qm, bounds, atol, plane, traj_cap, t0, max_attempts, plane2)
    S0, n, dev = out._keep[0], P.n, out.final.device
    L = backend.backend_lib
    order = None
    if split is None:
        split = program.has_sources and n <= SPLIT_MAX_PARTICLES
    if split:
        _trace_split(program, P, n, dev)
    else:
        shape = shape or DEFAULT_SHAPE
        if shape.ordered and n >= ORDER_MIN_PARTICLES:
            order = queue_order(S0)
            P.order = order.data_ptr()
        P.batch = int(shape.batch)
        P.occupancy = int(shape.occupancy)
        # a learned field's fallback sums: the persistent kernel parks the particles that need
        # one and the stage-synchronous tracer finishes them (not for recorded trajectories or
        # a step cap, which a resumed trace would count from zero)
        can_park = program.has_mlp and program.has_sources and not traj_cap and max_attempts >= (1 << 62)
        park = can_park if park is None else (park and can_park)
        if park:
            hfin = torch.empty(n, dtype=torch.float64, device=dev)
            hmaxf = torch.empty(n, dtype=torch.float64, device=dev)
            P.h_final, P.hmax_final, P.park = hfin.data_ptr(), hmaxf.data_ptr(), 1
        backend.check(L.tb_trace(C.byref(P), int(waves or shape.waves), device.stream_ptr()), "trace")
        if park:
            _resume_parked(program, out, hfin, hmaxf, qm, bounds, atol, plane, plane2, max_attempts)
    out._keep = out._keep + (order,)
    return out


class BeamT(C.Structure):
    """tb_beam_t: a seeded beam drawn on the device (tb_sample_beam)."""
    _fields_ = [(name, C.c_double) for name in ("z0", "r_max", "e_median", "e_sigma", "angle_sigma", "mass", "charge")]


def sample_beam(n: int, seed, z0: float, r_max: float, e_median: float, e_sigma: float = 0.0, angle_sigma: float = 0.0,
                mass: float = M_E, charge: float = E_CHARGE, out=None):
    """(n, 6) start states drawn on the device (tb_sample_beam): (x, y) uniform in the disc of
    radius r_max on z = z0, kinetic energy e_median exp(e_sigma g) eV, the direction
    (tan ax, tan ay, 1) normalised with ax, ay ~ N(0, angle_sigma), non-relativistic speed.
    seed: an int or a one-element uint64/int64 device tensor (read when the launch runs)."""
    import torch
    from . import device
    dev = device.device()
    if not torch.is_tensor(seed):
        seed = torch.tensor([int(seed) & ((1 << 63) - 1)], dtype=torch.int64, device=dev)
    if out is None:
        out = torch.empty((n, 6), dtype=torch.float64, device=dev)
    assert out.shape == (n, 6) and out.dtype == torch.float64 and out.is_contiguous()
    spec = BeamT(z0, r_max, e_median, e_sigma, angle_sigma, mass, abs(charge))
    backend.check(backend.backend_lib.tb_sample_beam(C.byref(spec), seed.data_ptr(), n, out.data_ptr(), device.stream_ptr()),
                  "sample_beam")
    return out


class BeamPipeline:
    """Sample -> trace captured once as a HIP graph and replayed per seed: tb_sample_beam
    writes the start states, tb_trace_noalloc (the persistent tracer with its counter, queue
    order and sort workspace in the library's static scratch array, so nothing is allocated
    inside the capture) traces them.  Each call writes a new seed into device memory, replays
    the graph on the current stream and returns the same result buffers (copy what must be
    kept).  One pipeline at a time per GPU: every tb_trace_noalloc shares the scratch array.
    Particles of a learned field whose step needs ring sources are summed in the persistent
    kernel (no parking), as tb_trace with park = 0."""

    def __init__(self, tracer, n: int, beam: dict, atol: float = 1e-8, plane=None, plane2=None,
                 charge: float = -E_CHARGE, mass: float = M_E, shape: LaunchShape | None = None):
        import torch
        from . import device
        L = backend.backend_lib
        if n > L.tb_trace_noalloc_max_particles():
            raise ValueError(f"{n} particles exceed the static maximum of {L.tb_trace_noalloc_max_particles()}")
        dev = device.device()
        self.n = n
        self.beam = dict(beam)
        self.beam.setdefault("mass", mass)
        self.beam.setdefault("charge", abs(charge))
        self.seed = torch.zeros(1, dtype=torch.int64, device=dev)
        self.state0 = torch.empty((n, 6), dtype=torch.float64, device=dev)
        program = tracer.program()
        self.P, self.out = trace_params(program, self.state0, charge / mass, tracer.bounds, atol, plane, 0, None, 1 << 62, plane2)
        shape = shape or DEFAULT_SHAPE
        self.P.batch, self.P.occupancy, self.P.park = int(shape.batch), int(shape.occupancy), 0
        occ = shape.occupancy if 1 <= shape.occupancy <= 4 else 2
        cus = torch.cuda.get_device_properties(dev).multi_processor_count
        self.waves = int(shape.waves) or cus * (4 * occ if not program.has_sources else 8)
        self.ordered = int(bool(shape.ordered))
        self._program = program
        self.graph = None
        self._run_eager()   # first launches outside the capture (the scratch array's set-up)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self._launch()
        self.graph = g

    def _launch(self):
        from . import device
        sample_beam(self.n, self.seed, out=self.state0, **self.beam)
        backend.check(backend.backend_lib.tb_trace_noalloc(C.byref(self.P), self.waves, self.ordered, device.stream_ptr()),
                      "trace_noalloc")

    def _run_eager(self):
        self._launch()

    def __call__(self, seed: int, graph: bool = True) -> BatchResult:
        """Trace the beam of this seed; graph=False launches the same kernels one by one."""
        self.seed.fill_(int(seed) & ((1 << 63) - 1))
        if graph:
            self.graph.replay()
        else:
            self._run_eager()
        return self.out


def _trace_split(program: DeviceFieldProgram, P, n, dev):
    import torch
    from . import device
    L = backend.backend_lib
    nchunk = L.tb_trace_split_chunks(n, program.n_sources)
    ws = torch.empty(L.tb_trace_split_workspace_bytes(n, nchunk), dtype=torch.uint8, device=dev)
    backend.check(L.tb_trace_split(C.byref(P), nchunk, ws.data_ptr(), device.stream_ptr()), "trace_split")
    return ws


def _resume_parked(program, out, hfin, hmaxf, qm, bounds, atol, plane, plane2, max_attempts):
    """Finish the particles that tb_trace parked (status STATUS_PARKED: state, t, h and counts
    at the start of the step that needed ring sources) on the stage-synchronous tracer, from
    that state with the same step size and step cap, in batches of SPLIT_MAX_PARTICLES; their
    counts add up and a plane crossed after the resume replaces one crossed before it."""
    import torch
    idx = torch.nonzero(out.status == STATUS_PARKED).flatten()
    out.n_parked = int(idx.numel())
    qm_arr = np.atleast_1d(np.asarray(qm, dtype=np.float64))
    for c0 in range(0, out.n_parked, SPLIT_MAX_PARTICLES):
        sel = idx[c0:c0 + SPLIT_MAX_PARTICLES]
        m = int(sel.numel())
        qm_sel = qm_arr[sel.cpu().numpy()] if qm_arr.size > 1 else qm_arr
        S, H, HM = out.final[sel].contiguous(), hfin[sel].contiguous(), hmaxf[sel].contiguous()
        P2, o2 = trace_params(program, S, qm_sel, bounds, atol, plane, 0, out.t_final[sel].cpu().numpy(), max_attempts, plane2)
        P2.h0, P2.hmax0 = H.data_ptr(), HM.data_ptr()
        ws = _trace_split(program, P2, m, S.device)
        out.final[sel] = o2.final
        out.status[sel] = o2.status
        out.t_final[sel] = o2.t_final
        out.n_accepted[sel] += o2.n_accepted
        out.n_attempted[sel] += o2.n_attempted
        for mine, theirs in ((out.landing, o2.landing), (out.landing2, o2.landing2)):
            if mine is not None:
                mine[sel] = torch.where(torch.isfinite(theirs[:, :1]), theirs, mine[sel])
        del ws, S, H, HM, o2


def trace_params(program: DeviceFieldProgram, state0, qm, bounds, atol=1e-8, plane=None, traj_cap=0, t0=None,
                 max_attempts=1 << 62, plane2=None):
    """The tb_trace_params_t of a trace (a zeroed work counter, no queue order, launch shape
    0) and the BatchResult whose device tensors it points to (result._keep holds the inputs)."""
    import torch
    from . import device
    S0 = state0 if isinstance(state0, torch.Tensor) else device.to_dev(np.atleast_2d(state0))
    n = S0.shape[0]
    qm_arr = np.atleast_1d(np.asarray(qm, dtype=np.float64))
    QM = device.to_dev(qm_arr)
    dev = S0.device
    P = TraceParamsT()
    P.field = program.prog
    P.state0 = S0.data_ptr()
    T0 = device.to_dev(np.asarray(t0, dtype=np.float64)) if t0 is not None else None
    P.t0 = T0.data_ptr() if T0 is not None else None
    P.h0 = None
    P.qm = QM.data_ptr()
    P.qm_stride = 1 if qm_arr.size > 1 else 0
    P.n = n
    counter = torch.zeros(1, dtype=torch.int32, device=dev)
    P.counter = counter.data_ptr()
    for i, v in enumerate(np.asarray(bounds, dtype=np.float64).ravel()):
        P.bounds[i] = v
    P.atol = float(atol)
    P.max_attempts = int(max_attempts)
    if plane is not None:
        p0, nrm = plane
        for i in range(3):
            P.plane_p[i] = p0[i]
            P.plane_n[i] = nrm[i]
        P.plane_nn = float(np.linalg.norm(nrm))
    else:
        P.plane_n[2] = 1.0
        P.plane_nn = 1.0
    if plane2 is not None:
        p0, nrm = plane2
        for i in range(3):
            P.plane2_p[i] = p0[i]
            P.plane2_n[i] = nrm[i]
        P.plane2_nn = float(np.linalg.norm(nrm))
    final = torch.empty((n, 6), dtype=torch.float64, device=dev)
    landing = torch.empty((n, 6), dtype=torch.float64, device=dev) if plane is not None else None
    landing2 = torch.empty((n, 6), dtype=torch.float64, device=dev) if plane2 is not None else None
    P.landing2 = landing2.data_ptr() if landing2 is not None else None
    status = torch.empty(n, dtype=torch.int32, device=dev)
    nacc = torch.empty(n, dtype=torch.int64, device=dev)
    natt = torch.empty(n, dtype=torch.int64, device=dev)
    tfin = torch.empty(n, dtype=torch.float64, device=dev)
    P.final_state = final.data_ptr()
    P.landing = landing.data_ptr() if landing is not None else None
    P.status = status.data_ptr()
    P.n_accepted = nacc.data_ptr()
    P.n_attempted = natt.data_ptr()
    P.t_final = tfin.data_ptr()
    P.h_final = None
    traj = times = tlen = None
    if traj_cap:
        traj = torch.empty((n, traj_cap, 6), dtype=torch.float64, device=dev)
        times = torch.empty((n, traj_cap), dtype=torch.float64, device=dev)
        tlen = torch.empty(n, dtype=torch.int64, device=dev)
        P.traj, P.times, P.traj_cap, P.traj_len = traj.data_ptr(), times.data_ptr(), int(traj_cap), tlen.data_ptr()
    out = BatchResult(final, landing, status, nacc, natt, tfin, landing2)
    out.traj, out.times, out.traj_len = traj, times, tlen
    out._keep = (S0, QM, T0, counter)
    return P, out


class Tracer:
    """tracing.py:104-245."""

    def __init__(self, field, bounds):
        self.field = field
        bounds = np.array(bounds, dtype=np.float64)
        assert bounds.shape == (3, 2)
        self.bounds = bounds
        self._program = None
        from .field import FieldRadialBEM, FieldRadialAxial
        if type(field) in (FieldRadialBEM, FieldRadialAxial):
            self.trace_fun, self.low_level_args = field.get_low_level_trace_function()
            self.trace_args = C.cast(C.pointer(self.low_level_args), C.c_void_p)
        else:
            self.trace_fun = None
            self.low_level_args = None
            self.trace_args = None

    def program(self) -> DeviceFieldProgram:
        if self._program is None:
            self._program = DeviceFieldProgram(self.field)
        return self._program

    def __call__(self, *args, **kwargs):
        return self.trace_single(*args, **kwargs)

    def trace_single(self, position, velocity, mass: float = M_E, charge: float = -E_CHARGE, atol: float = 1e-8):
        position = np.array(position, dtype=np.float64)
        velocity = np.array(velocity, dtype=np.float64)
        qm = charge / mass
        if self.trace_fun is not None:
            return backend.trace_particle(position, velocity, qm, self.trace_fun, self.bounds, atol, self.trace_args)
        return self._trace_program(position, velocity, qm, atol)

    def _trace_program(self, position, velocity, qm, atol):
        """The reference's block loop (backend/__init__.py:382-420) over the GPU tracer."""
        prog = self.program()
        cap = backend.TRACING_BLOCK_SIZE
        state = np.concatenate([position, velocity])[None, :]
        t0 = 0.0
        tb, sb = [], []
        while True:
            r = run_trace(prog, state, qm, self.bounds, atol, traj_cap=cap, t0=np.array([t0]), waves=1)
            n = int(r.traj_len[0].item())
            s = r.traj[0, :n].cpu().numpy()
            t = r.times[0, :n].cpu().numpy()
            sb.append(s[1:] if sb else s)
            tb.append(t[1:] if tb else t)
            if n != cap:
                break
            state = sb[-1][-1][None, :].copy()
            t0 = tb[-1][-1]
        if len(sb) == 1:
            return tb[0], sb[0]
        return np.concatenate(tb), np.concatenate(sb)

    def trace_multiple(self, position, velocity, mass=M_E, charge=-E_CHARGE, atol: float = 1e-8):
        """tracing.py:210-245: one trajectory per particle (numpy broadcasting)."""
        positions, velocities = np.broadcast_arrays(np.atleast_2d(position), np.atleast_2d(velocity))
        N = len(positions)
        masses = np.broadcast_to(np.asarray(mass, dtype=np.float64), (N,))
        charges = np.broadcast_to(np.asarray(charge, dtype=np.float64), (N,))
        return [self.trace_single(p, v, mass=float(m), charge=float(c), atol=atol)
                for p, v, m, c in zip(positions, velocities, masses, charges)]

    def trace_batch(self, positions, velocities, mass=M_E, charge=-E_CHARGE, atol: float = 1e-8, plane=None,
                    max_attempts=1 << 62, waves=0, plane2=None, shape="auto") -> BatchResult:
        """Trace many particles in one persistent launch; no trajectories are stored.
        plane = (p0, normal) records the state where each trajectory last crosses it
        (what xy_plane_intersection returns on the full trajectory), NaN if never; plane2 a
        second such plane (result.landing2), e.g. the entrance plane for reflected particles.
        shape: a LaunchShape, or "auto" for the one autotune.tune picked for this field and
        particle count (DEFAULT_SHAPE before any tuning)."""
        state0 = batch_state(positions, velocities)
        qm = np.asarray(charge, dtype=np.float64) / np.asarray(mass, dtype=np.float64)
        if qm.size > 1:
            qm = np.broadcast_to(qm, (len(state0),))
        if isinstance(shape, str):
            from . import autotune
            shape = autotune.cached_shape(self.program(), len(state0), atol)
        return run_trace(self.program(), state0, qm, self.bounds, atol, plane=plane, max_attempts=max_attempts, waves=waves,
                         plane2=plane2, shape=shape)


def batch_state(positions, velocities):
    """Start states (n, 6) from positions and velocities: numpy arrays (broadcast against each
    other), or device tensors (n, 3) that stay on the device."""
    import torch
    if isinstance(positions, torch.Tensor) or isinstance(velocities, torch.Tensor):
        p = torch.as_tensor(positions, dtype=torch.float64, device="cuda")
        v = torch.as_tensor(velocities, dtype=torch.float64, device="cuda")
        p, v = torch.broadcast_tensors(torch.atleast_2d(p), torch.atleast_2d(v))
        return torch.cat([p, v], dim=1).contiguous()
    positions = np.atleast_2d(np.asarray(positions, dtype=np.float64))
    velocities = np.atleast_2d(np.asarray(velocities, dtype=np.float64))
    positions, velocities = np.broadcast_arrays(positions, velocities)
    return np.concatenate([positions, velocities], axis=1)


def plane_intersection(positions, p0, normal):
    positions = np.array(positions, dtype=np.float64)
    assert positions.shape == (len(positions), 6), "The positions array should have shape (N, 6)"
    return backend.plane_intersection(positions, np.array(p0, dtype=np.float64), np.array(normal, dtype=np.float64))


def xy_plane_intersection(positions, z: float):
    return plane_intersection(positions, [0, 0, z], [0, 0, 1])


def xz_plane_intersection(positions, y: float):
    return plane_intersection(positions, [0, y, 0], [0, 1, 0])


def yz_plane_intersection(positions, x: float):
    return plane_intersection(positions, [x, 0, 0], [1, 0, 0])


def axis_intersection(positions) -> float:
    return yz_plane_intersection(positions, 0)[2]
