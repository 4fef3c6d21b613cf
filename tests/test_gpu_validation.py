"""Known-answer tests of the reference's validation scripts and of its coil fixture, through
the HIP path (solve_direct, FieldRadialAxial, Tracer, field point API):

- validation/edwards2007.py:58: potential 6.69099430708 at (12, 0, 4) (the paper's value);
- validation/simple_mirror.py:53: landing radius 0.16327355811 of a 100 eV electron reflected
  by a flat mirror at -110 V, traced through FieldRadialAxial(0.02, 4);
- validation/two_current_coils.py:50: H_r = -91.94907464785867 A/m at (10 mm, 0, 0) of two
  opposed coils beside a mu_r = 25 block inside a magnetostatic boundary;
- validation/einzel_lens.py:61: focal length 3.915970140918643 of a three-aperture einzel lens;
- tests/test_radial.py:94-114: axial H of a rectangular coil against the published curve
  (tests/golden/axial_magnetic_field.txt) to 2 %.

The meshes are this library's own (mesh.py): uniform elements, three times finer at each of
three levels.  Electrode corners make the surface charge singular, so uniform meshes converge
slowly (errors fall about as h^1 to h^1.5, and as h^0.7 next to the mu_r = 25 block); each test
checks that the error falls at every refinement and that the finest level (or, for the coils,
the Richardson extrapolation of the three levels) reaches the known answer to the stated bar."""


def converges(values, want, bar, name, extrapolate=False):
    errs = [abs(v - want) / abs(want) for v in values]
    for v, e in zip(values, errs):
        print(f"{name}: {v:.12f} (rel err {e:.2e})", flush=True)
    assert all(b < a for a, b in zip(errs, errs[1:])), errs
    if extrapolate:
        v1, v2, v3 = values
        r = (v2 - v1) / (v3 - v2)
        v = v3 + (v3 - v2) / (r - 1.0)
        print(f"{name}: Richardson {v:.10f} (rel err {abs(v - want) / abs(want):.2e})", flush=True)
        assert abs(v - want) / abs(want) < bar
    else:
        assert errs[-1] < bar
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return True


def edwards(n):
    from traceon_amd import solver
    from traceon_amd.excitation import Excitation
    from traceon_amd.mesh import path_mesh
    inner = path_mesh([[0, 0, 5], [12, 0, 5], [12, 0, 15], [0, 0, 15]], "inner", n_elements=n)
    boundary = path_mesh([[0, 0, 0], [20, 0, 0], [20, 0, 20], [0, 0, 20]], "boundary", n_elements=2 * n)
    exc = Excitation(inner + boundary)
    exc.add_voltage(boundary=0.0, inner=10.0)
    return solver.solve_direct(exc).potential_at_point([12, 0.0, 4])


def test_edwards2007_potential(gpu):
    converges([edwards(n) for n in (500, 1500, 4500)], 6.69099430708, 1e-4, "Edwards 2007")


def mirror(n):
    from traceon_amd import solver, tracing as T
    from traceon_amd.excitation import Excitation
    from traceon_amd.field import FieldRadialAxial
    from traceon_amd.mesh import path_mesh
    boundary = path_mesh([[0, 0, -1], [2, 0, -1], [2, 0, 1], [0.3, 0, 1]], "boundary", n_elements=n)
    m = path_mesh([[0.0, 0, 0.0], [1.0, 0, 0.0]], "mirror", n_elements=max(n // 5, 1))
    exc = Excitation(boundary + m)
    exc.add_voltage(mirror=-110.0, boundary=0.0)
    field = solver.solve_direct(exc)
    axial = FieldRadialAxial(field, 0.02, 4)
    tracer = axial.get_tracer(((-0.22, 0.22), (-0.22, 0.22), (0.02, 11)))
    _, pos = tracer([0.0, 0.0, 10.0], T.velocity_vec_xz_plane(100, 1e-3))
    p = T.xy_plane_intersection(pos, 10)
    return float(np.linalg.norm(p[:1]))


def test_simple_mirror_landing(gpu):
    converges([mirror(n) for n in (1000, 3000, 9000)], 1.6327355811e-01, 1e-4, "simple mirror")


def coils(n_line, rings, sectors):
    from traceon_amd import solver
    from traceon_amd.excitation import Excitation
    from traceon_amd.mesh import disk_surface_mesh, path_mesh, rectangle_path_mesh
    mm = 1e-3
    boundary = path_mesh([[0, 0, 50 * mm], [100 * mm, 0, 50 * mm], [100 * mm, 0, -50 * mm], [0, 0, -50 * mm]], "boundary",
                         n_elements=n_line)
    block = rectangle_path_mesh(5 * mm, 15 * mm, -1 * mm, 1 * mm, "block", n_elements=n_line // 2)
    c1 = disk_surface_mesh(10 * mm, 5 * mm, 1 * mm, "coil1", rings, sectors)
    c2 = disk_surface_mesh(10 * mm, -5 * mm, 1 * mm, "coil2", rings, sectors)
    exc = Excitation(c1 + c2 + boundary + block)
    exc.add_current(coil1=10, coil2=-10)
    exc.add_magnetizable(block=25)
    exc.add_magnetostatic_boundary("boundary")
    return solver.solve_direct(exc).magnetostatic_field_at_point([10 * mm, 0.0, 0.0])[0]


def test_two_current_coils_field(gpu):
    # the coil triangles matter little (2304 and 9216 triangles agree to 2e-7); the lines of
    # the block and the boundary carry the error
    converges([coils(n, 12, 96) for n in (800, 2400, 7200)], -91.94907464785867, 1e-3, "two current coils",
              extrapolate=True)


def einzel(n):
    from traceon_amd import solver, tracing as T
    from traceon_amd.excitation import Excitation
    from traceon_amd.field import FieldRadialAxial
    from traceon_amd.mesh import path_mesh
    TH, SP, R = 0.5, 0.5, 0.15
    extent = 2.0 - 0.1

    def aperture(z, name):
        return path_mesh([[extent, 0, z - TH / 2], [R, 0, z - TH / 2], [R, 0, z + TH / 2], [extent, 0, z + TH / 2]], name,
                         n_elements=n)
    mesh = path_mesh([[0, 0, 1.75], [2.0, 0, 1.75], [2.0, 0, -1.75], [0, 0, -1.75]], "boundary", n_elements=2 * n)
    mesh = mesh + aperture(-TH - SP, "ground") + aperture(0.0, "lens") + aperture(TH + SP, "ground")
    exc = Excitation(mesh)
    exc.add_voltage(ground=0.0, lens=1000)
    exc.add_electrostatic_boundary("boundary")
    field = solver.solve_direct(exc)
    field.set_bounds(((-R, R), (-R, R), (-1.5, 1.5)))
    axial = FieldRadialAxial(field, -1.5, 1.5, 600)
    tracer = axial.get_tracer(((-R, R), (-R, R), (-5, 3.5)))
    _, pos = tracer(np.array([R / 3, 0.0, 3]), T.velocity_vec_xz_plane(1000, 0))
    return -T.axis_intersection(pos)


def test_einzel_lens_focal_length(gpu):
    converges([einzel(n) for n in (400, 1200, 3600)], 3.915970140918643, 1e-4, "einzel lens")


def test_rectangular_coil_axial_field(gpu):
    """tests/test_radial.py:94-114: a 1 x 1 coil section at (1..2, 1..2), 5 A-turn, meshed at
    0.1 (here 10 x 10 cells, 200 triangles), H_z on the axis against the published curve."""
    import os
    from scipy.interpolate import CubicSpline
    from traceon_amd import solver
    from traceon_amd.excitation import Excitation
    from traceon_amd.mesh import rectangle_surface_mesh
    mesh = rectangle_surface_mesh(1.0, 2.0, 1.0, 2.0, "coil", 10, 10)
    exc = Excitation(mesh)
    exc.add_current(coil=5)
    field = solver.solve_direct(exc)
    z = np.linspace(-0.5, 3.5, 300)
    Hz = np.array([field.current_field_at_point([0.0, 0.0, zz])[2] for zz in z])
    data = np.loadtxt(os.path.join(os.path.dirname(__file__), "golden", "axial_magnetic_field.txt"), delimiter=",")
    ref = CubicSpline(data[:, 0], data[:, 1] * 1e-3)(z)
    print("max rel deviation from the published curve", np.max(np.abs(Hz - ref) / np.abs(ref)), flush=True)
    assert np.allclose(ref, Hz, atol=0.0, rtol=2e-2)
