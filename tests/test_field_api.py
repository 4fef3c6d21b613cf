"""Field API: coordinate transforms, bounds and superposition typing (tests/test_field.py:15-164).

The transform, bounds and typing rules are host logic and run on the CPU with fields built
from synthetic point charges (nothing is evaluated there).  The evaluations through a
transform need the GPU and run on the solved two-rectangle geometry of the reference's
test: a ring at 1 V and a ring at magnetostatic potential -1."""
import numpy as np
import pytest


def _charges(n, seed):
    from traceon_amd.field import EffectivePointCharges
    rng = np.random.default_rng(seed)
    pos = np.stack([rng.uniform(0.1, 1.0, (n, 16)), rng.uniform(-1.5, 1.5, (n, 16))], axis=2)
    return EffectivePointCharges(rng.normal(size=n), rng.uniform(0.01, 0.1, (n, 16)), pos)


@pytest.fixture(scope="module")
def synthetic_field():
    from traceon_amd.field import FieldRadialBEM
    return FieldRadialBEM(_charges(6, 1), _charges(4, 2))


def test_map_points(synthetic_field):
    f = synthetic_field
    moved = f.move(dx=1)
    assert np.allclose(moved.get_origin(), [1.0, 0.0, 0.0])
    assert np.allclose(moved.get_basis(), f.get_basis())
    T = np.eye(4)
    T[0, 3] = 1.0
    assert np.allclose(np.linalg.inv(moved._inverse_transformation_matrix), T)
    # a rotation after the translation carries the origin along: x -> -z, z -> x
    turned = moved.rotate(Ry=np.pi / 2)
    assert np.allclose(turned.get_origin(), [0, 0, -1])
    assert np.allclose(turned.get_basis(), [[0, 0, -1], [0, 1, 0], [1, 0, 0]])
    assert np.allclose(np.linalg.inv(turned._inverse_transformation_matrix),
                       [[0, 0, -1, 0], [0, 1, 0, 0], [1, 0, 0, -1], [0, 0, 0, 1]])
    # the untouched field keeps the identity frame
    assert not f.has_transform() and turned.has_transform()


def test_map_points_to_local(synthetic_field):
    f = synthetic_field
    moved = f.move(dx=1)
    p = np.array([1.0, 1.0, 1.0])
    assert np.allclose(moved.map_points_to_local(p + [1, 0, 0]), p)
    turned = moved.rotate(Ry=np.pi / 2)
    assert np.allclose(turned.map_points_to_local(turned.get_origin()), f.get_origin())
    # the frame matrix holds the images of the unit vectors as its rows (field.py:156-178), so
    # a turn about z by pi/2 (basis rows y, -x, z) maps global (0, 2, 0.5) to local (-2, 0, 0.5)
    r = f.rotate_around_axis([0, 0, 1], np.pi / 2)
    assert np.allclose(r.get_basis(), [[0, 1, 0], [-1, 0, 0], [0, 0, 1]])
    assert np.allclose(r.map_points_to_local([0.0, 2.0, 0.5]), [-2.0, 0.0, 0.5])


def test_field_bounds_global_and_local(synthetic_field):
    f = synthetic_field.copy()
    mirrored = f.mirror_xy()
    mirrored.set_bounds([[0, 1], [0, 1], [0, 1]])
    f.set_bounds([[0, 1], [0, 1], [0, 1]])
    assert np.allclose(mirrored.field_bounds, f.field_bounds)   # bounds are local by default
    mirrored.set_bounds([[0, 1], [0, 1], [0, 1]], global_coordinates=True)
    assert np.allclose(mirrored.field_bounds, [[0, 1], [0, 1], [-1, 0]])
    # outside the bounds the global evaluation is zero without touching the charges
    assert mirrored.electrostatic_potential_at_point([0.5, 0.5, -0.5]) == 0.0
    assert np.array_equal(mirrored.electrostatic_field_at_point([0.5, 0.5, -0.5]), np.zeros(3))


def test_superposition_typing(synthetic_field):
    from traceon_amd.field import FieldRadialBEM, FieldSuperposition, UniformField
    f = synthetic_field
    scaled = 2 * f
    moved = f.move(dx=1)
    magnetic_only = FieldRadialBEM(magnetostatic_point_charges=f.magnetostatic_point_charges)
    # same geometry: the charges add and the type stays
    assert isinstance(f + scaled, FieldRadialBEM)
    assert np.allclose((f + scaled).electrostatic_point_charges.charges, 3 * f.electrostatic_point_charges.charges)
    s = f + magnetic_only
    assert isinstance(s, FieldRadialBEM)
    assert np.allclose(s.magnetostatic_point_charges.charges, 2 * f.magnetostatic_point_charges.charges)
    assert np.allclose(s.electrostatic_point_charges.charges, f.electrostatic_point_charges.charges)
    # another frame or another type: a superposition
    assert isinstance(f + moved, FieldSuperposition)
    assert isinstance(f + UniformField(E=(0, 0, 1)), FieldSuperposition)
    sup = FieldSuperposition([f, moved])
    assert len(sup + moved) == 3 and len(sup + UniformField()) == 3
    assert isinstance(f - f, FieldRadialBEM)
    assert np.allclose((f - f).electrostatic_point_charges.charges, 0.0)


@pytest.fixture(scope="module")
def solved():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from traceon_amd import solver
    from traceon_amd.excitation import Excitation
    from traceon_amd.field import FieldRadialAxial
    from traceon_amd.mesh import rectangle_path_mesh
    mesh = rectangle_path_mesh(0.1, 1.0, 1.0, 1.5, "pos", n_elements=40) + rectangle_path_mesh(0.1, 1.0, -1.5, -1.0, "neg",
                                                                                              n_elements=40)
    exc = Excitation(mesh)
    exc.add_voltage(pos=1.0)
    exc.add_magnetostatic_potential(neg=-1.0)
    field = solver.solve_direct(exc)
    return field, FieldRadialAxial(field, -2, 2, 100)


@pytest.mark.gpu
def test_axial_field_inherits_the_frame(solved):
    from traceon_amd.field import FieldRadialAxial
    field, axial = solved
    a1 = axial.move(dz=1).rotate(Rz=np.pi)
    a2 = FieldRadialAxial(field.move(dz=1).rotate(Rz=np.pi), -2, 2, 100)
    assert np.allclose(a1._inverse_transformation_matrix, a2._inverse_transformation_matrix)


@pytest.mark.gpu
def test_evaluation_through_a_transform(solved):
    field, axial = solved
    pg = np.array([1.0, 1.0, 1.0])
    for f in (field.move(dx=1), axial.move(dx=1)):
        pl = f.map_points_to_local(pg)
        assert np.isclose(f.electrostatic_potential_at_point(pg), f.electrostatic_potential_at_local_point(pl), rtol=1e-14)
        assert np.isclose(f.magnetostatic_potential_at_point(pg), f.magnetostatic_potential_at_local_point(pl), rtol=1e-14)
        assert np.allclose(f.electrostatic_field_at_point(pg), f.electrostatic_field_at_local_point(pl), rtol=1e-14)
        assert np.allclose(f.magnetostatic_field_at_point(pg), f.magnetostatic_field_at_local_point(pl), rtol=1e-14)
    # a rotation turns the local field vector with the basis: about z by pi/2, local x -> global y
    r = field.rotate(Rz=np.pi / 2)
    Eg = r.electrostatic_field_at_point([0.0, 0.4, 0.6])   # between the rings, outside the conductors
    El = field.electrostatic_field_at_point([0.4, 0.0, 0.6])
    assert abs(El[0]) > 1e-2 * abs(El).max()
    assert np.allclose(Eg, [-El[1], El[0], El[2]], rtol=1e-12, atol=1e-12 * abs(El).max())


@pytest.mark.gpu
def test_bounds_zero_only_outside(solved):
    field, _ = solved
    m = field.mirror_xy()
    m.set_bounds([[0, 1], [0, 1], [0, 1]], global_coordinates=True)
    point = [0.5, 0.5, -0.5]
    assert m.electrostatic_potential_at_point(point) == 0.0
    assert m.electrostatic_potential_at_local_point(point) != 0.0


@pytest.mark.gpu
def test_superposition_typing_with_axial(solved):
    from traceon_amd.field import FieldRadialAxial, FieldRadialBEM, FieldSuperposition
    field, axial = solved
    assert isinstance(axial + 2 * axial, FieldRadialAxial)
    assert isinstance(field + FieldRadialBEM(field.magnetostatic_point_charges), FieldRadialBEM)
    assert isinstance(axial + field.move(dx=1), FieldSuperposition)
    assert isinstance(field + axial, FieldSuperposition)
    assert isinstance(axial + FieldRadialAxial(field, -2, 2, 200), FieldSuperposition)
    sup = FieldSuperposition([field, field.move(dx=1)])
    assert len(sup + field.move(dx=1)) == 3 and len(sup + axial) == 3
    # the sum of equal-frame axial fields evaluates as the doubled field
    z = np.array([0.0, 0.0, 0.3])
    assert np.isclose((axial + axial).electrostatic_potential_at_point(z), 2 * axial.electrostatic_potential_at_point(z),
                      rtol=1e-13)
