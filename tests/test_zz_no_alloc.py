"""tb_trace_noalloc (the tracer with its work counter, queue order and sort workspace in a
static device array) against tb_trace through the C-ABI: bit-identical final states,
landings, statuses and step counts, on two calls in a row, the second with fewer particles
(the array keeps the first call's contents), and the size limit's error."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

QM = -1.602176634e-19 / 9.1093837015e-31


@pytest.fixture(scope="module")
def axial_program():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from traceon_amd import configs, solver, tracing
    from traceon_amd.excitation import Excitation
    from traceon_amd.field import FieldRadialAxial
    mesh, volts, diel = configs.small_geometry(240)
    exc = Excitation(mesh)
    exc.add_voltage(**volts)
    exc.add_dielectric(**diel)
    field = solver.solve_direct(exc)
    ax = FieldRadialAxial(field, -2.4e-3, 2.4e-3, N=500)
    return tracing.DeviceFieldProgram(ax)


def beam(n, seed):
    from traceon_amd import tracing
    rng = np.random.default_rng(seed)
    P = np.column_stack([rng.uniform(0, 0.3e-3, n), np.zeros(n), np.full(n, -2.3e-3)])
    d = np.column_stack([rng.normal(0, 0.02, n), rng.normal(0, 0.02, n), np.ones(n)])
    V = np.array([tracing.velocity_vec(e, dd) for e, dd in zip(rng.uniform(20, 80, n), d)])
    return np.hstack([P, V])


def results(out):
    return [x.cpu().numpy() for x in (out.final, out.landing, out.landing2, out.status, out.n_accepted, out.n_attempted)]


def test_noalloc_matches_tb_trace(axial_program):
    import torch
    from traceon_amd import backend, device, tracing
    L = backend.backend_lib
    bounds = [[-0.9e-3, 0.9e-3], [-0.9e-3, 0.9e-3], [-2.4e-3, 2.4e-3]]
    plane = (np.array([0, 0, 2.0e-3]), np.array([0, 0, 1.0]))
    plane2 = (np.array([0, 0, -2.2e-3]), np.array([0, 0, 1.0]))
    for n, seed in ((30_000, 1), (7_000, 2)):
        S = beam(n, seed)
        ref = tracing.run_trace(axial_program, S, QM, bounds, plane=plane, plane2=plane2,
                                shape=tracing.LaunchShape(ordered=True, batch=1, occupancy=2))
        for ordered in (1, 0):
            print("particles", n, "ordered", ordered, flush=True)
            P, out = tracing.trace_params(axial_program, S, QM, bounds, plane=plane, plane2=plane2)
            P.counter = None           # the entry point uses its own
            P.batch, P.occupancy = 1, 2
            rc = L.tb_trace_noalloc(C.byref(P), 0, ordered, device.stream_ptr())
            assert rc == 0, backend.last_error()
            torch.cuda.synchronize()
            for a, b in zip(results(out), results(ref)):
                assert np.array_equal(a, b, equal_nan=True)
            assert np.mean(np.isfinite(results(out)[1][:, 0])) > 0.5


def test_noalloc_rejects_too_many_particles(axial_program):
    from traceon_amd import backend, device, tracing
    L = backend.backend_lib
    nmax = L.tb_trace_noalloc_max_particles()
    assert nmax == 1 << 21
    assert L.tb_trace_noalloc_scratch_bytes() >= 8 * nmax
    P, out = tracing.trace_params(axial_program, beam(16, 3), QM, [[-1e-3, 1e-3]] * 2 + [[-2.4e-3, 2.4e-3]])
    P.n = nmax + 1                  # rejected on the host before any launch
    assert L.tb_trace_noalloc(C.byref(P), 0, 1, device.stream_ptr()) == -1
    assert "exceed the static maximum" in backend.last_error()
