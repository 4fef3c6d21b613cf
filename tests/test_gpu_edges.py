"""Edge cases of the HIP path against the oracle (plain C restatement of the reference's
backend): the ring kernels on a grid that reaches the near-singular pair (source a few
nanometres from the target), the axis (r0 = 0 and r0 < MIN_DISTANCE_AXIS = 1e-10,
three_d.c:2) and far pairs; current rings on and off the axis; empty inputs; a particle that
starts outside the bounds (tracing.c:79: no step is taken); solves at sizes below, at and
just above the 128-row panel."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
MM = 1e-3


@pytest.fixture(scope="module")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return True


def ring_grid():
    """(r0, z0, dr, dz): targets on and near the axis and off it; offsets from 1e-12 m to
    1 cm in four directions, keeping the source ring radius r0 + dr positive."""
    out = []
    for r0 in (0.0, 1e-11, 1e-6, 0.5 * MM, 2 * MM):
        for d in (1e-12, 1e-9, 1e-6, 1e-4, 1e-3, 1e-2):
            for ur, uz in ((1.0, 0.0), (0.0, 1.0), (0.6, 0.8), (-0.6, 0.8)):
                dr, dz = d * ur, d * uz
                if r0 + dr <= 0.0:
                    continue
                out.append((r0, 0.3 * MM, dr, dz))
    return out


def test_ring_kernels_on_a_grid_vs_oracle(gpu, oracle):
    """potential_radial_ring, dr1_ and dz1_potential_radial_ring (radial_ring.c:21-53) through
    the reference-named entry points.  Bar: 1e-12 relative, or 1e-13 of the largest value of
    the same kernel at that target where a component passes through zero (dz = 0 gives
    d/dz = 0; on the axis d/dr = 0)."""
    from traceon_amd import backend as B
    grid = ring_grid()
    worst = {}
    for name in ("potential_radial_ring", "dr1_potential_radial_ring", "dz1_potential_radial_ring"):
        g = np.array([getattr(B, name)(*a) for a in grid])
        o = np.array([getattr(oracle.lib, name)(*a, None) for a in grid])
        r0 = np.array([a[0] for a in grid])
        scale = np.array([np.max(np.abs(o[r0 == x])) for x in r0])
        err = np.where(g == o, 0.0, np.abs(g - o) / np.maximum(np.abs(o), 1e-1 * scale + 1e-300))
        assert np.all(np.isfinite(err)), name
        worst[name] = float(err.max())
        bad = np.nonzero(err > 1e-12)[0]
        assert len(bad) == 0, (name, [(grid[i], g[i], o[i]) for i in bad[:5]])
        if name == "dr1_potential_radial_ring":
            assert np.all(g[r0 < 1e-10] == 0.0)   # MIN_DISTANCE_AXIS (radial_ring.c:23-25)
    print("ring kernels, worst relative error over", len(grid), "pairs:", worst, flush=True)


def test_current_ring_on_and_off_the_axis_vs_oracle(gpu, oracle):
    """current_field_radial_ring (radial_ring.c:78-107) and its axial potential
    (radial_ring.c:73-76): on the axis (H_r = 0), next to the ring, far from it."""
    from traceon_amd import backend as B
    worst = 0.0
    for x0 in (0.0, 1e-11, 0.5 * MM, 0.999 * MM, 1.001 * MM, 3 * MM):
        for y0 in (-2 * MM, -1e-6, 0.0, 1e-6, 2 * MM):
            if abs(x0 - 1 * MM) < 1e-9 and abs(y0) < 1e-9:
                continue
            g = np.asarray(B.current_field_radial_ring(x0, y0, 1 * MM, 0.0))
            out = np.zeros(2)
            oracle.lib.current_field_radial_ring(x0, y0, 1 * MM, 0.0, out)
            e = np.max(np.abs(g - out)) / max(np.max(np.abs(out)), 1e-300)
            worst = max(worst, e)
            assert e < 1e-12, (x0, y0, g, out)
            if x0 < 1e-10:
                assert g[0] == 0.0
    for z0 in (-3 * MM, -1e-6, 0.0, 1e-6, 3 * MM):
        g, o = B.current_potential_axial_radial_ring(z0, 1 * MM, 0.0), oracle.lib.current_potential_axial_radial_ring(z0, 1 * MM, 0.0)
        assert abs(g - o) <= 1e-12 * max(abs(o), 1.0)
    print("current ring, worst relative error", worst, flush=True)


def test_empty_inputs(gpu):
    """No sources, no points, no particles: zeros or empty results, no error."""
    import torch
    from traceon_amd import configs, device, solver, tracing
    from traceon_amd.excitation import Excitation
    from traceon_amd.field import FieldRadialAxial
    S0 = device.Sources(np.zeros(0), np.zeros((0, 16)), np.zeros((0, 16, 2)))
    P, F = S0.field_points(device.to_dev(np.array([[1e-4, 0.0, 0.0]])))
    assert float(P.abs().max()) == 0.0 and float(F.abs().max()) == 0.0
    mesh, volts = configs.einzel_lens(300)
    exc = Excitation(mesh)
    exc.add_voltage(**volts)
    field = solver.solve_direct(exc)
    S = field.electrostatic_point_charges.device_sources()
    P, F = S.field_points(torch.zeros((0, 3), dtype=torch.float64, device="cuda"))
    assert P.shape == (0,) and F.shape == (0, 3)
    axial = FieldRadialAxial(field, -5 * MM, 5 * MM, N=400)
    bounds = np.array([[-0.9 * MM, 0.9 * MM], [-0.9 * MM, 0.9 * MM], [-4 * MM, 4 * MM]])
    for fld in (field, axial):
        r = fld.get_tracer(bounds).trace_batch(np.zeros((0, 3)), np.zeros((0, 3)), plane=(np.zeros(3), np.array([0.0, 0, 1])))
        assert r.final.shape[0] == 0 and r.landing.shape[0] == 0


def test_particle_starting_outside_the_bounds_takes_no_step(gpu):
    """tracing.c:79: the step loop is not entered, so the trajectory is the start state alone;
    both tracers retire such a particle at once (status 0, no attempted step, no landing),
    beside particles that do trace."""
    from traceon_amd import configs, solver, tracing
    from traceon_amd.excitation import Excitation
    from traceon_amd.field import FieldRadialAxial
    mesh, volts = configs.einzel_lens(300)
    exc = Excitation(mesh)
    exc.add_voltage(**volts)
    field = solver.solve_direct(exc)
    axial = FieldRadialAxial(field, -5 * MM, 5 * MM, N=400)
    bounds = np.array([[-0.9 * MM, 0.9 * MM], [-0.9 * MM, 0.9 * MM], [-4 * MM, 4 * MM]])
    P = np.array([[0.1 * MM, 0.0, -3.9 * MM], [0.1 * MM, 0.0, -4.5 * MM], [0.95 * MM, 0.0, 0.0], [0.0, 0.0, -3.9 * MM]])
    V = np.tile(tracing.velocity_vec(500.0, [0, 0, 1]), (4, 1))
    plane = (np.array([0.0, 0.0, 3.5 * MM]), np.array([0.0, 0.0, 1.0]))
    QM = np.full(4, -tracing.E_CHARGE / tracing.M_E)
    for fld in (field, axial):
        prog = fld.get_tracer(bounds).program()
        for split in ((True, False) if prog.has_sources else (False,)):
            r = tracing.run_trace(prog, np.hstack([P, V]), QM, bounds, plane=plane, split=split)
            natt, st = r.n_attempted.cpu().numpy(), r.status.cpu().numpy()
            fin, land = r.final.cpu().numpy(), r.landing.cpu().numpy()
            out = [1, 2]
            assert np.all(natt[out] == 0) and np.all(st[out] == 0), (natt, st)
            assert np.array_equal(fin[out], np.hstack([P, V])[out])
            assert np.all(np.isnan(land[out]))
            assert np.all(natt[[0, 3]] > 0) and np.all(np.isfinite(land[[0, 3], 0]))
    # the reference-named single-particle call: the start state alone
    t, s = field.get_tracer(bounds)(P[1], V[1])
    assert s.shape == (1, 6) and np.array_equal(s[0], np.hstack([P[1], V[1]]))


@pytest.mark.parametrize("N", [1, 2, 3, 127, 128, 129, 257])
def test_solves_at_small_and_ragged_sizes(gpu, N):
    """FP64 and mixed-precision solves against numpy.linalg.solve on diagonally dominant random
    matrices whose size is below, at and just above the 128-row panel (single-panel
    factorisation, a last panel of 1 column, a partial triangular-solve block)."""
    import torch
    from traceon_amd import device
    rng = np.random.default_rng(N)
    A = rng.uniform(-1, 1, (N, N)) + 2.0 * N * np.eye(N)
    b = rng.standard_normal((1, N))
    ref = np.linalg.solve(A, b[0])
    M64 = torch.tensor(A, device="cuda")
    X, st, _ = device.solve_system(M64, b, mode="fp64")
    e64 = float(np.max(np.abs(X.cpu().numpy()[0] - ref)) / np.max(np.abs(ref)))
    X2, st2, _ = device.solve_system(M64, b, M32=M64.float(), mode="mixed")
    e32 = float(np.max(np.abs(X2.cpu().numpy()[0] - ref)) / np.max(np.abs(ref)))
    print(f"N = {N}: fp64 {e64:.1e} ({st.mode}), mixed {e32:.1e} ({st2.mode}, {st2.iterations} steps)", flush=True)
    assert e64 < 1e-13 and e32 < 1e-13
    assert st2.mode == "mixed"
