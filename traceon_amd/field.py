"""Fields — counterpart of traceon/field.py (EffectivePointCharges, Field, FieldSuperposition,
FieldBEM, FieldRadialBEM, FieldAxial, FieldRadialAxial).

The public methods keep the reference's names and semantics (global -> local affine map,
inclusive field bounds, basis rotation of the result, superposition arithmetic).  Every
evaluation runs in the HIP kernels: single points go through the reference-named C entry
points (as the reference's ctypes calls do), and the batched additions
``*_at_points`` evaluate whole point sets in one launch on device-resident charges.
"""
from __future__ import annotations

import copy
from itertools import product

import numpy as np

from . import backend
from .excitation import Symmetry

FACTOR_AXIAL_DERIV_SAMPLING_2D = 0.2   # field.py:819


def _is_numeric(x):
    return isinstance(x, (int, float, np.generic))


class GeometricObject:
    """move / rotate / mirror on top of ``map_points`` (mesher.py:25-170)."""

    def map_points(self, fun):
        raise NotImplementedError

    def move(self, dx: float = 0.0, dy: float = 0.0, dz: float = 0.0):
        return self.map_points(lambda p: np.asarray(p) + np.array([dx, dy, dz]))

    def rotate(self, Rx: float = 0.0, Ry: float = 0.0, Rz: float = 0.0, origin=(0, 0, 0)):
        assert sum([Rx == 0.0, Ry == 0.0, Rz == 0.0]) >= 2, "Only supply one axis of rotation"
        if Rx != 0.0:
            return self.rotate_around_axis([1, 0, 0], angle=Rx, origin=origin)
        if Ry != 0.0:
            return self.rotate_around_axis([0, 1, 0], angle=Ry, origin=origin)
        if Rz != 0.0:
            return self.rotate_around_axis([0, 0, 1], angle=Rz, origin=origin)
        return self.map_points(lambda x: np.array(x))

    def rotate_around_axis(self, axis=(0, 0, 1), angle: float = 0.0, origin=(0, 0, 0)):
        origin = np.asarray(origin, dtype=float)
        a = np.asarray(axis, dtype=float)
        a = a / np.linalg.norm(a)
        if angle == 0:
            return self.map_points(lambda x: np.array(x))
        # Rodrigues: R = I + sin(t) K + (1 - cos(t)) K^2 with K the cross-product matrix of the axis
        K = np.array([[0.0, -a[2], a[1]], [a[2], 0.0, -a[0]], [-a[1], a[0], 0.0]])
        R = np.eye(3) + np.sin(angle) * K + (1 - np.cos(angle)) * (K @ K)
        return self.map_points(lambda p: origin + R @ (np.asarray(p) - origin))

    def mirror_xz(self):
        return self.map_points(lambda p: np.array([p[0], -p[1], p[2]]))

    def mirror_yz(self):
        return self.map_points(lambda p: np.array([-p[0], p[1], p[2]]))

    def mirror_xy(self):
        return self.map_points(lambda p: np.array([p[0], p[1], -p[2]]))


class EffectivePointCharges:
    """Charges (or currents) with their Gauss weights*jacobians and positions (field.py:46-131)."""

    def __init__(self, charges, jacobians, positions, directions=None):
        self.charges = np.array(charges, dtype=np.float64)
        self.jacobians = np.array(jacobians, dtype=np.float64)
        self.positions = np.array(positions, dtype=np.float64)
        self.directions = np.array(directions, dtype=np.float64) if directions is not None else None
        N = len(self.charges)
        nq = self.jacobians.shape[1] if self.jacobians.ndim == 2 else 0
        assert self.charges.shape == (N,) and self.jacobians.shape == (N, nq)
        assert self.positions.shape in ((N, nq, 3), (N, nq, 2))
        self._dev = None

    @staticmethod
    def empty_2d():
        return EffectivePointCharges(np.empty((0,)), np.empty((0, 16)), np.empty((0, 16, 2)))

    @staticmethod
    def empty_3d():
        return EffectivePointCharges(np.empty((0,)), np.empty((0, 12)), np.empty((0, 12, 3)))

    def is_2d(self):
        return self.jacobians.shape[1] == backend.N_QUAD_2D

    def is_3d(self):
        return self.jacobians.shape[1] == backend.N_TRIANGLE_QUAD

    def _matches_geometry(self, other):
        return (self.positions.shape == other.positions.shape and np.allclose(self.positions, other.positions)
                and self.jacobians.shape == other.jacobians.shape and np.allclose(self.jacobians, other.jacobians))

    def __len__(self):
        return len(self.charges)

    def __add__(self, other):
        if not isinstance(other, EffectivePointCharges) or self.is_2d() != other.is_2d():
            return NotImplemented
        if self._matches_geometry(other):
            return EffectivePointCharges(self.charges + other.charges, self.jacobians, self.positions)
        return EffectivePointCharges(np.concatenate([self.charges, other.charges]), np.concatenate([self.jacobians, other.jacobians]),
                                     np.concatenate([self.positions, other.positions]))

    def __radd__(self, other):
        return self.__add__(other)

    def __mul__(self, other):
        if _is_numeric(other):
            return EffectivePointCharges(other * self.charges, self.jacobians, self.positions)
        return NotImplemented

    def __rmul__(self, other):
        return self.__mul__(other)

    def __neg__(self):
        return -1 * self

    def __sub__(self, other):
        if isinstance(other, EffectivePointCharges):
            return self.__add__(-other)
        return NotImplemented

    def device_sources(self):
        """Flattened ring sources on the GPU (cached)."""
        if self._dev is None:
            from . import device
            self._dev = device.Sources(self.charges, self.jacobians, self.positions, "line" if self.is_2d() else "ring")
        return self._dev


class Field(GeometricObject):
    def __init__(self):
        self._origin = np.array([0.0, 0.0, 0.0])
        self._basis = np.eye(3)
        self._update_inverse_transformation_matrix()
        self.field_bounds = None

    def get_origin(self):
        return self._origin.copy()

    def get_basis(self):
        return self._basis.copy()

    def _update_inverse_transformation_matrix(self):
        T = np.eye(4)
        T[:3, :3] = self._basis
        T[:3, 3] = self._origin
        assert np.linalg.det(T) != 0, "Transformations of field have resulted in a two-dimensional coordinate system."
        self._inverse_transformation_matrix = np.linalg.inv(T)

    def has_transform(self) -> bool:
        return not (np.array_equal(self._origin, np.zeros(3)) and np.array_equal(self._basis, np.eye(3)))

    def copy(self):
        return copy.copy(self)

    def map_points(self, fun):
        """field.py:168-178: the origin and the images of the basis vectors define the new frame."""
        fc = self.copy()
        fc._origin = np.asarray(fun(self._origin), dtype=np.float64)
        fc._basis = np.array([np.asarray(fun(b + self._origin)) - fc._origin for b in self._basis])
        fc._update_inverse_transformation_matrix()
        return fc

    def map_points_to_local(self, point):
        h = self._inverse_transformation_matrix @ np.array([*point, 1.0])
        return h[:3]

    def set_bounds(self, bounds, global_coordinates: bool = False):
        bounds = np.array(bounds, dtype=np.float64)
        assert bounds.shape == (3, 2)
        if global_coordinates:
            corners = np.array([self.map_points_to_local(c) for c in product(*bounds)])
            bounds = np.column_stack((corners.min(axis=0), corners.max(axis=0)))
        self.field_bounds = bounds

    def _within_field_bounds(self, point) -> bool:
        return bool(self.field_bounds is None or np.all((self.field_bounds[:, 0] <= point) & (point <= self.field_bounds[:, 1])))

    def field_at_point(self, point):
        e, m = self.is_electrostatic(), self.is_magnetostatic()
        if e and not m:
            return self.electrostatic_field_at_point(point)
        if m and not e:
            return self.magnetostatic_field_at_point(point)
        raise RuntimeError("Cannot use field_at_point when both electric and magnetic fields are present")

    def potential_at_point(self, point):
        e, m = self.is_electrostatic(), self.is_magnetostatic()
        if e and not m:
            return self.electrostatic_potential_at_point(point)
        if m and not e:
            return self.magnetostatic_potential_at_point(point)
        raise RuntimeError("Cannot use potential_at_point when both electric and magnetic fields are present")

    def _global(self, point, local_fun, vector=True):
        lp = self.map_points_to_local(point)
        if self._within_field_bounds(lp):
            v = local_fun(lp)
            return self._basis @ v if vector else v
        return np.zeros(3) if vector else 0.0

    def electrostatic_field_at_point(self, point):
        return self._global(point, self.electrostatic_field_at_local_point)

    def magnetostatic_field_at_point(self, point):
        return self._global(point, self.magnetostatic_field_at_local_point)

    def current_field_at_point(self, point):
        return self._global(point, self.current_field_at_local_point)

    def electrostatic_potential_at_point(self, point):
        return self._global(point, self.electrostatic_potential_at_local_point, vector=False)

    def magnetostatic_potential_at_point(self, point):
        return self._global(point, self.magnetostatic_potential_at_local_point, vector=False)

    def is_electrostatic(self):
        return False

    def is_magnetostatic(self):
        return False

    def electrostatic_field_at_local_point(self, point):
        return np.zeros(3)

    def magnetostatic_field_at_local_point(self, point):
        return np.zeros(3)

    def current_field_at_local_point(self, point):
        return np.zeros(3)

    def electrostatic_potential_at_local_point(self, point):
        return 0.0

    def magnetostatic_potential_at_local_point(self, point):
        return 0.0

    def __add__(self, other):
        if isinstance(other, Field) and not isinstance(other, FieldSuperposition):
            return FieldSuperposition([self, other])
        return NotImplemented

    def __mul__(self, other):
        if _is_numeric(other):
            return FieldSuperposition([self], [other])
        return NotImplemented

    def __rmul__(self, other):
        return self.__mul__(other)

    def __neg__(self):
        return -1 * self

    def __sub__(self, other):
        if isinstance(other, Field):
            return self + (-other)
        return NotImplemented

    def get_tracer(self, bounds):
        from .tracing import Tracer
        return Tracer(self, bounds)

    # -- GPU field program (the device counterpart of get_low_level_trace_function) --------
    def device_components(self, factor: float = 1.0):
        """List of (tb_field_comp_t, keep-alive) describing this field for the GPU tracer."""
        raise TypeError(f"{type(self).__name__} has no GPU field description; only FieldRadialBEM, FieldRadialAxial, "
                        "UniformField and superpositions of them can be traced on the GPU")


class FieldSuperposition(Field):
    """Linear combination of fields (field.py:424-507)."""

    def __init__(self, fields, factors=None):
        super().__init__()
        assert all(isinstance(f, Field) for f in fields)
        self.fields = list(fields)
        self.factors = np.ones(len(self.fields)) if factors is None else np.array(factors, dtype=np.float64)

    def map_points(self, fun):
        return FieldSuperposition([f.map_points(fun) for f in self.fields], self.factors)

    def electrostatic_field_at_local_point(self, point):
        return np.sum([fa * f.electrostatic_field_at_point(point) for fa, f in zip(self.factors, self.fields)], axis=0)

    def magnetostatic_field_at_local_point(self, point):
        return np.sum([fa * f.magnetostatic_field_at_point(point) for fa, f in zip(self.factors, self.fields)], axis=0)

    def current_field_at_local_point(self, point):
        return np.sum([fa * f.current_field_at_point(point) for fa, f in zip(self.factors, self.fields)], axis=0)

    def electrostatic_potential_at_local_point(self, point):
        return sum(fa.item() * f.electrostatic_potential_at_point(point) for fa, f in zip(self.factors, self.fields))

    def magnetostatic_potential_at_local_point(self, point):
        return sum(fa.item() * f.magnetostatic_potential_at_point(point) for fa, f in zip(self.factors, self.fields))

    def is_electrostatic(self):
        return any(f.is_electrostatic() for f in self.fields)

    def is_magnetostatic(self):
        return any(f.is_magnetostatic() for f in self.fields)

    def __add__(self, other):
        if isinstance(other, FieldSuperposition):
            return FieldSuperposition(self.fields + other.fields, np.concatenate((self.factors, other.factors)))
        return FieldSuperposition(self.fields + [other], np.concatenate((self.factors, [1.0])))

    def __radd__(self, other):
        return FieldSuperposition([other] + self.fields, [1.0] + list(self.factors))

    def __iadd__(self, other):
        self.fields = (self + other).fields
        return self

    def __mul__(self, other):
        if _is_numeric(other):
            return FieldSuperposition(self.fields, other * self.factors)
        return NotImplemented

    def __rmul__(self, other):
        return self.__mul__(other)

    def __getitem__(self, index):
        if isinstance(index, slice):
            return FieldSuperposition(self.fields[index], self.factors[index])
        return self.factors[index] * self.fields[index]

    def __len__(self):
        return len(self.fields)

    def __iter__(self):
        for fa, f in zip(self.factors, self.fields):
            yield f * fa.item()

    def device_components(self, factor: float = 1.0):
        comps = []
        for fa, f in zip(self.factors, self.fields):
            if isinstance(f, FieldSuperposition):
                if f.field_bounds is not None or f.has_transform():
                    raise TypeError("nested superpositions with their own bounds or transform are not traced on the GPU")
                comps += f.device_components(factor * float(fa))
            else:
                comps += f.device_components(factor * float(fa), in_superposition=True)
        return comps


class UniformField(Field):
    """Constant E and H everywhere (inside the bounds): the analytic fields the reference's
    tracer tests use through Python callbacks (tests/test_tracing.py:26-133)."""

    def __init__(self, E=(0.0, 0.0, 0.0), H=(0.0, 0.0, 0.0)):
        super().__init__()
        self.E = np.array(E, dtype=np.float64)
        self.H = np.array(H, dtype=np.float64)

    def is_electrostatic(self):
        return bool(np.any(self.E != 0))

    def is_magnetostatic(self):
        return bool(np.any(self.H != 0))

    def electrostatic_field_at_local_point(self, point):
        return self.E.copy()

    def magnetostatic_field_at_local_point(self, point):
        return self.H.copy()

    def device_components(self, factor: float = 1.0, in_superposition: bool = False):
        from . import tracing
        return [tracing.make_component(self, factor, kind="uniform", in_superposition=in_superposition)]


class FieldBEM(Field):
    def __init__(self, electrostatic_point_charges, magnetostatic_point_charges, current_point_charges):
        super().__init__()
        self.electrostatic_point_charges = electrostatic_point_charges
        self.magnetostatic_point_charges = magnetostatic_point_charges
        self.current_point_charges = current_point_charges

    def is_electrostatic(self):
        return len(self.electrostatic_point_charges) > 0

    def is_magnetostatic(self):
        return len(self.magnetostatic_point_charges) > 0 or len(self.current_point_charges) > 0

    def _matches_geometry(self, other):
        return self.__class__ == other.__class__ and np.allclose(self._origin, other._origin) and np.allclose(self._basis, other._basis)

    def __add__(self, other):
        if self._matches_geometry(other):
            fc = self.copy()
            fc.electrostatic_point_charges = self.electrostatic_point_charges + other.electrostatic_point_charges
            fc.magnetostatic_point_charges = self.magnetostatic_point_charges + other.magnetostatic_point_charges
            fc.current_point_charges = self.current_point_charges + other.current_point_charges
            return fc
        return super().__add__(other)

    def __sub__(self, other):
        if isinstance(other, Field):
            return self.__add__(-other)
        return NotImplemented

    def __radd__(self, other):
        return self.__add__(other)

    def __mul__(self, other):
        if _is_numeric(other):
            fc = self.copy()
            fc.electrostatic_point_charges = self.electrostatic_point_charges * other
            fc.magnetostatic_point_charges = self.magnetostatic_point_charges * other
            fc.current_point_charges = self.current_point_charges * other
            return fc
        return super().__mul__(other)

    def __neg__(self):
        return self.__class__(-self.electrostatic_point_charges, -self.magnetostatic_point_charges, -self.current_point_charges)

    def __rmul__(self, other):
        return self.__mul__(other)

    def area_of_elements(self, indices):
        return sum(self.area_of_element(i) for i in indices)

    def charge_on_element(self, i):
        return self.area_of_element(i) * self.electrostatic_point_charges.charges[i]

    def charge_on_elements(self, indices):
        return sum(self.charge_on_element(i) for i in indices)


class FieldRadialBEM(FieldBEM):
    """Radially symmetric field of BEM charges (field.py:620-816)."""

    def __init__(self, electrostatic_point_charges=None, magnetostatic_point_charges=None, current_point_charges=None):
        e = electrostatic_point_charges if electrostatic_point_charges is not None else EffectivePointCharges.empty_2d()
        m = magnetostatic_point_charges if magnetostatic_point_charges is not None else EffectivePointCharges.empty_2d()
        c = current_point_charges if current_point_charges is not None else EffectivePointCharges.empty_3d()
        self.symmetry = Symmetry.RADIAL
        super().__init__(e, m, c)

    # ---- single points: the reference-named C entry points ----
    def current_field_at_local_point(self, point):
        c = self.current_point_charges
        return backend.current_field_radial(np.asarray(point, dtype=np.float64), c.charges, c.jacobians, c.positions)

    def electrostatic_field_at_local_point(self, point):
        e = self.electrostatic_point_charges
        return backend.field_radial(np.asarray(point, dtype=np.float64), e.charges, e.jacobians, e.positions)

    def magnetostatic_field_at_local_point(self, point):
        # field.py:672-695: the coil part goes through current_field_at_point on the local point
        current = self.current_field_at_point(point)
        m = self.magnetostatic_point_charges
        return current + backend.field_radial(np.asarray(point, dtype=np.float64), m.charges, m.jacobians, m.positions)

    def electrostatic_potential_at_local_point(self, point):
        e = self.electrostatic_point_charges
        return backend.potential_radial(np.asarray(point, dtype=np.float64), e.charges, e.jacobians, e.positions)

    def magnetostatic_potential_at_local_point(self, point):
        m = self.magnetostatic_point_charges
        return backend.potential_radial(np.asarray(point, dtype=np.float64), m.charges, m.jacobians, m.positions)

    def current_potential_axial(self, z: float) -> float:
        c = self.current_point_charges
        return backend.current_potential_axial(float(z), c.charges, c.jacobians, c.positions)

    def get_electrostatic_axial_potential_derivatives(self, z):
        z = np.asarray(z, dtype=np.float64)
        return self._axial_derivatives(z, self.electrostatic_point_charges)

    def get_magnetostatic_axial_potential_derivatives(self, z):
        z = np.asarray(z, dtype=np.float64)
        return self._axial_derivatives(z, self.magnetostatic_point_charges) + self.get_current_axial_potential_derivatives(z)

    def get_current_axial_potential_derivatives(self, z):
        z = np.asarray(z, dtype=np.float64)
        return self._axial_derivatives(z, self.current_point_charges)

    def _axial_derivatives(self, z, charges):
        if len(charges) == 0 or len(z) == 0:
            return np.zeros((len(z), 9))
        from . import device
        zt = device.to_dev(z)
        return charges.device_sources().axial_derivatives(zt).cpu().numpy()

    def area_of_element(self, i):
        j = self.electrostatic_point_charges.jacobians
        p = self.electrostatic_point_charges.positions
        return 2 * np.pi * np.sum(j[i] * p[i, :, 0])

    # ---- batched evaluation on the GPU (additions) ----
    def _points_local(self, points):
        P = np.atleast_2d(np.asarray(points, dtype=np.float64))
        if self.has_transform():
            H = np.concatenate([P, np.ones((len(P), 1))], axis=1)
            P = (H @ self._inverse_transformation_matrix.T)[:, :3]
        inside = np.ones(len(P), dtype=bool)
        if self.field_bounds is not None:
            b = self.field_bounds
            inside = np.all((b[:, 0] <= P) & (P <= b[:, 1]), axis=1)
        return P, inside

    def _batch(self, points, charges, want_pot, want_field):
        from . import device
        P, inside = self._points_local(points)
        pts = device.to_dev(P)
        pot, fld = charges.device_sources().field_points(pts, pot=want_pot, field=want_field)
        pot = pot.cpu().numpy() if pot is not None else None
        fld = fld.cpu().numpy() if fld is not None else None
        if fld is not None:
            fld = fld @ self._basis.T
            fld[~inside] = 0.0
        if pot is not None:
            pot[~inside] = 0.0
        return pot, fld

    def electrostatic_potential_at_points(self, points):
        return self._batch(points, self.electrostatic_point_charges, True, False)[0]

    def electrostatic_field_at_points(self, points):
        return self._batch(points, self.electrostatic_point_charges, False, True)[1]

    def electrostatic_potential_and_field_at_points(self, points):
        return self._batch(points, self.electrostatic_point_charges, True, True)

    def magnetostatic_field_at_points(self, points):
        from . import device
        _, f = self._batch(points, self.magnetostatic_point_charges, False, True)
        if len(self.current_point_charges):
            P, inside = self._points_local(points)
            _, fc = self.current_point_charges.device_sources().field_points(device.to_dev(P), pot=False, field=True)
            fc = fc.cpu().numpy() @ self._basis.T
            fc[~inside] = 0.0
            f = f + fc
        return f

    def get_low_level_trace_function(self):
        """field.py:814-816: the C field function and its argument struct."""
        args = backend.FieldEvaluationArgsRadial(self.electrostatic_point_charges, self.magnetostatic_point_charges,
                                                 self.current_point_charges, self.field_bounds)
        return backend.traceable("field_radial_traceable"), args

    def device_components(self, factor: float = 1.0, in_superposition: bool = False):
        from . import tracing
        return [tracing.make_component(self, factor, kind="bem", in_superposition=in_superposition)]

    def __str__(self):
        return (f"<FieldRadialBEM elec={len(self.electrostatic_point_charges)} mag={len(self.magnetostatic_point_charges)} "
                f"rings={len(self.current_point_charges)}>")


class FieldAxial(Field):
    def __init__(self, field, z, electrostatic_coeffs=None, magnetostatic_coeffs=None):
        super().__init__()
        self.field = field
        self._origin = field._origin
        self._basis = field._basis
        self._update_inverse_transformation_matrix()
        N = len(z)
        assert z.shape == (N,)
        assert electrostatic_coeffs is not None or magnetostatic_coeffs is not None
        assert z[0] < z[-1], "z values in axial interpolation should be ascending"
        self.z = z
        self.electrostatic_coeffs = electrostatic_coeffs if electrostatic_coeffs is not None else np.zeros_like(magnetostatic_coeffs)
        self.magnetostatic_coeffs = magnetostatic_coeffs if magnetostatic_coeffs is not None else np.zeros_like(electrostatic_coeffs)
        self.has_electrostatic = bool(np.any(self.electrostatic_coeffs != 0.0))
        self.has_magnetostatic = bool(np.any(self.magnetostatic_coeffs != 0.0))
        self._dev = None

    def is_electrostatic(self):
        return self.has_electrostatic

    def is_magnetostatic(self):
        return self.has_magnetostatic

    def _matches_geometry(self, other):
        return (self.__class__ == other.__class__ and np.allclose(self._origin, other._origin)
                and np.allclose(self._basis, other._basis) and self.z.shape == other.z.shape and np.allclose(self.z, other.z))

    def __add__(self, other):
        if self._matches_geometry(other):
            fc = self.copy()
            fc.electrostatic_coeffs = self.electrostatic_coeffs + other.electrostatic_coeffs
            fc.magnetostatic_coeffs = self.magnetostatic_coeffs + other.magnetostatic_coeffs
            fc._dev = None
            return fc
        return super().__add__(other)

    def __sub__(self, other):
        if isinstance(other, Field):
            return self.__add__(-other)
        return NotImplemented

    def __radd__(self, other):
        return self.__add__(other)

    def __mul__(self, other):
        if _is_numeric(other):
            fc = self.copy()
            fc.electrostatic_coeffs = other * self.electrostatic_coeffs
            # field.py:895 scales the electrostatic coefficients into the magnetostatic slot;
            # kept as the reference does it
            fc.magnetostatic_coeffs = other * self.electrostatic_coeffs
            fc._dev = None
            return fc
        return super().__mul__(other)

    def __neg__(self):
        return -1 * self

    def __rmul__(self, other):
        return self.__mul__(other)


class FieldRadialAxial(FieldAxial):
    """Paraxial interpolation of a FieldRadialBEM (field.py:940-1066)."""

    def __init__(self, field: FieldRadialBEM, zmin: float, zmax: float, N: int | None = None):
        assert isinstance(field, FieldRadialBEM)
        z, ce, cm = FieldRadialAxial._get_interpolation_coefficients(field, zmin, zmax, N=N)
        super().__init__(field, z, ce, cm)
        assert self.electrostatic_coeffs.shape == (len(z) - 1, backend.DERIV_2D_MAX, 6)

    @staticmethod
    def _get_interpolation_coefficients(field, zmin, zmax, N=None):
        """Derivatives sampled on the axis and their quintic/cubic interpolation, both on the GPU."""
        from . import device
        assert zmax > zmin, "zmax should be bigger than zmin"
        Nc = max(len(field.electrostatic_point_charges.charges), len(field.magnetostatic_point_charges.charges))
        N = N if N is not None else int(FACTOR_AXIAL_DERIV_SAMPLING_2D * Nc)
        z = np.linspace(zmin, zmax, N)
        zt = device.to_dev(z)
        import torch
        de = torch.zeros((N, 9), dtype=torch.float64, device=zt.device)
        dm = torch.zeros((N, 9), dtype=torch.float64, device=zt.device)
        if len(field.electrostatic_point_charges):
            field.electrostatic_point_charges.device_sources().axial_derivatives(zt, out=de)
        if len(field.magnetostatic_point_charges):
            field.magnetostatic_point_charges.device_sources().axial_derivatives(zt, out=dm)
        if len(field.current_point_charges):
            field.current_point_charges.device_sources().axial_derivatives(zt, out=dm)
        ce = device.axial_coefficients(de, zt).cpu().numpy()
        cm = device.axial_coefficients(dm, zt).cpu().numpy()
        return z, ce, cm

    def electrostatic_field_at_local_point(self, point):
        return backend.field_radial_derivs(np.asarray(point, dtype=np.float64), self.z, self.electrostatic_coeffs)

    def magnetostatic_field_at_local_point(self, point):
        return backend.field_radial_derivs(np.asarray(point, dtype=np.float64), self.z, self.magnetostatic_coeffs)

    def electrostatic_potential_at_local_point(self, point):
        return backend.potential_radial_derivs(np.asarray(point, dtype=np.float64), self.z, self.electrostatic_coeffs)

    def magnetostatic_potential_at_local_point(self, point):
        return backend.potential_radial_derivs(np.asarray(point, dtype=np.float64), self.z, self.magnetostatic_coeffs)

    def device_arrays(self):
        if self._dev is None:
            from . import device
            self._dev = (device.to_dev(self.z), device.to_dev(self.electrostatic_coeffs), device.to_dev(self.magnetostatic_coeffs))
        return self._dev

    def _axial_batch(self, points, which):
        from . import device
        P = np.atleast_2d(np.asarray(points, dtype=np.float64))
        if self.has_transform():
            H = np.concatenate([P, np.ones((len(P), 1))], axis=1)
            P = (H @ self._inverse_transformation_matrix.T)[:, :3]
        z, ce, cm = self.device_arrays()
        pot, fld = device.axial_points(device.to_dev(P), z, ce if which == "e" else cm)
        fld = fld.cpu().numpy() @ self._basis.T
        pot = pot.cpu().numpy()
        if self.field_bounds is not None:
            b = self.field_bounds
            out = ~np.all((b[:, 0] <= P) & (P <= b[:, 1]), axis=1)
            fld[out] = 0.0
            pot[out] = 0.0
        return pot, fld

    def electrostatic_field_at_points(self, points):
        return self._axial_batch(points, "e")[1]

    def electrostatic_potential_at_points(self, points):
        return self._axial_batch(points, "e")[0]

    def magnetostatic_field_at_points(self, points):
        return self._axial_batch(points, "m")[1]

    def get_low_level_trace_function(self):
        args = backend.FieldDerivsArgs(self.z, self.electrostatic_coeffs, self.magnetostatic_coeffs)
        return backend.traceable("field_radial_derivs_traceable"), args

    def device_components(self, factor: float = 1.0, in_superposition: bool = False):
        from . import tracing
        return [tracing.make_component(self, factor, kind="axial", in_superposition=in_superposition)]
