"""ctypes front-end of libtraceon_hip.so — the counterpart of traceon/backend/__init__.py.

Same loader role (traceon/backend/__init__.py:23-55), same constants read with
``in_dll`` (:58-66), same numpy-level wrappers with the same names, shapes and
accumulate semantics (:355-780).  Every numeric call runs on the GPU; the library must
be built (``python -m traceon_amd.build``) and there is no host fallback.

The reference has no error channel for most of these functions; this binding clears
the library's error slot before each call and raises ``RuntimeError`` if the call set it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np
from numpy.ctypeslib import ndpointer

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtraceon_hip.so")

if not os.path.isfile(LIB_PATH):
    raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m traceon_amd.build` (hipcc, gfx950)")

# PyTorch ships its own HIP and HSA runtimes.  The library's calls take torch's streams and
# allocations, so it must bind to those same runtimes: loaded after torch, its HIP dependency
# resolves to torch's copy; loaded first, the process would hold two runtimes and the second
# one to open the device finds none.
import torch  # noqa: E402,F401

backend_lib = C.CDLL(LIB_PATH)

TRACING_BLOCK_SIZE = C.c_size_t.in_dll(backend_lib, "TRACING_BLOCK_SIZE").value
DERIV_2D_MAX = C.c_int.in_dll(backend_lib, "DERIV_2D_MAX_SYM").value
N_QUAD_2D = C.c_int.in_dll(backend_lib, "N_QUAD_2D_SYM").value
N_TRIANGLE_QUAD = C.c_int.in_dll(backend_lib, "N_TRIANGLE_QUAD_SYM").value
NU_MAX = C.c_int.in_dll(backend_lib, "NU_MAX_SYM").value
M_MAX = C.c_int.in_dll(backend_lib, "M_MAX_SYM").value


def arr(ndim=None, shape=None, dtype=np.float64):
    return ndpointer(ndim=ndim, shape=shape, dtype=dtype, flags=("C_CONTIGUOUS", "ALIGNED"))


dbl = C.c_double
dbl_p = C.POINTER(dbl)
vp = C.c_void_p
sz = C.c_size_t
i32 = C.c_int
v2 = arr(shape=(2,))
v3 = arr(shape=(3,))

field_fun = C.CFUNCTYPE(None, C.POINTER(dbl), C.POINTER(dbl), vp, C.POINTER(dbl), C.POINTER(dbl))


def ensure_contiguous_aligned(a):
    assert isinstance(a, np.ndarray)
    return np.require(a, requirements=("C_CONTIGUOUS", "ALIGNED"))


_SIG = {
    # Tier R: the reference's own FFI (traceon/backend/__init__.py:287-328)
    "ellipkm1": (dbl, [dbl]), "ellipk": (dbl, [dbl]), "ellipem1": (dbl, [dbl]), "ellipe": (dbl, [dbl]),
    "normal_2d": (None, [arr(ndim=1), arr(ndim=1), v2]),
    "higher_order_normal_radial": (None, [dbl, v3, v3, v3, v3, v2]),
    "position_and_jacobian_radial": (None, [dbl, v3, v3, v3, v3, v2, dbl_p]),
    "delta_position_and_jacobian_radial": (None, [dbl, v3, v3, v3, v3, v2, dbl_p]),
    "trace_particle": (sz, [arr(shape=(TRACING_BLOCK_SIZE,)), arr(shape=(TRACING_BLOCK_SIZE, 6)), dbl, field_fun,
                            arr(shape=(3, 2)), dbl, vp]),
    "potential_radial_ring": (dbl, [dbl, dbl, dbl, dbl, vp]),
    "dr1_potential_radial_ring": (dbl, [dbl, dbl, dbl, dbl, vp]),
    "dz1_potential_radial_ring": (dbl, [dbl, dbl, dbl, dbl, vp]),
    "axial_derivatives_radial": (None, [arr(ndim=2), arr(ndim=1), arr(ndim=2), arr(ndim=3), sz, arr(ndim=1), sz]),
    "potential_radial": (dbl, [v3, arr(ndim=1), arr(ndim=2), arr(ndim=3), sz]),
    "potential_radial_derivs": (dbl, [v3, arr(ndim=1), arr(ndim=3), sz]),
    "flux_density_to_charge_factor": (dbl, [dbl]),
    "charge_radial": (dbl, [arr(ndim=2), dbl]),
    "field_radial": (None, [v3, v3, arr(ndim=1), arr(ndim=2), arr(ndim=3), sz]),
    "field_radial_derivs": (None, [v3, v3, arr(ndim=1), arr(ndim=3), sz]),
    "current_potential_axial_radial_ring": (dbl, [dbl, dbl, dbl]),
    "current_potential_axial": (dbl, [dbl, arr(ndim=1), arr(ndim=2), arr(ndim=3), sz]),
    "current_field_radial_ring": (None, [dbl, dbl, dbl, dbl, v2]),
    "current_field_radial": (None, [v3, v3, arr(ndim=1), arr(ndim=2), arr(ndim=3), sz]),
    "current_axial_derivatives_radial": (None, [arr(ndim=2), arr(ndim=1), arr(ndim=2), arr(ndim=3), sz, arr(ndim=1), sz]),
    "fill_jacobian_buffer_radial": (None, [arr(ndim=2), arr(ndim=3), arr(ndim=3), sz]),
    "self_potential_radial": (dbl, [dbl, vp]),
    "self_field_dot_normal_radial": (dbl, [dbl, vp]),
    "normal_3d": (None, [arr(shape=(3, 3)), v3]),
    "position_and_jacobian_3d": (None, [dbl, dbl, arr(ndim=2), v3, dbl_p]),
    "triangle_areas": (None, [arr(ndim=3), arr(ndim=1), sz]),
    "self_potential_radial_integral": (dbl, [arr(shape=(4, 3))]),
    "self_field_dot_normal_radial_integral": (dbl, [arr(shape=(4, 3)), dbl]),
    "fill_matrix_radial": (None, [arr(ndim=2), arr(ndim=3), arr(dtype=np.uint8, ndim=1), arr(ndim=1), arr(ndim=2), arr(ndim=3),
                                  sz, sz, i32, i32]),
    "fill_jacobian_buffer_3d": (None, [arr(ndim=2), arr(ndim=3), arr(ndim=3), sz]),
    "plane_intersection": (C.c_bool, [v3, v3, arr(ndim=2), sz, arr(shape=(6,))]),
    # Tier B (device pointers)
    "tb_last_error": (C.c_char_p, []),
    "tb_clear_error": (None, []),
    "tb_version": (i32, []),
    "tb_struct_sizes": (None, [arr(dtype=np.uint64, shape=(4,))]),
    "tb_synchronize": (i32, [vp]),
    "tb_rowinfo_bytes": (sz, []),
    "tb_prepare_elements": (i32, [vp, vp, vp, i32, vp, vp, vp, vp, vp, vp]),
    "tb_fill_matrix": (i32, [vp, sz, vp, sz, vp, vp, i32, i32, i32, i32, vp]),
    "tb_fill_diagonal": (i32, [vp, sz, vp, sz, vp, vp, i32, i32, i32, vp, vp]),
    "tb_lu_workspace_bytes": (sz, [i32, i32]),
    "tb_lu_factor": (i32, [vp, sz, i32, vp, vp, vp, vp]),
    "tb_lu_factor_f32": (i32, [vp, sz, i32, vp, vp, vp, vp]),
    "tb_lu_factor_timed": (i32, [vp, sz, i32, vp, vp, vp, i32, vp, vp]),
    "tb_lu_solve": (i32, [vp, sz, i32, vp, vp, i32, vp, vp]),
    "tb_lu_solve_f32": (i32, [vp, sz, i32, vp, vp, i32, vp, vp]),
    "tb_residual": (i32, [vp, sz, i32, vp, vp, vp, i32, vp, vp]),
    "tb_axpy": (i32, [vp, vp, sz, vp]),
    "tb_lu_solve_status": (i32, [vp, i32, i32, vp, vp]),
    "tb_absmax": (i32, [vp, sz, vp, vp]),
    "tb_convert_f64_f32": (i32, [vp, sz, vp, sz, i32, i32, vp]),
    "tb_gemm_f64": (i32, [i32, i32, i32, dbl, vp, sz, vp, sz, dbl, vp, sz, vp]),
    "tb_gemm_f32": (i32, [i32, i32, i32, C.c_float, vp, sz, vp, sz, C.c_float, vp, sz, vp]),
    "tb_sources": (i32, [vp, vp, vp, i32, i32, i32, i32, vp, vp, vp, vp]),
    "tb_field_points": (i32, [vp, i32, vp, vp, vp, i32, vp, vp, i32, vp]),
    "tb_current_points": (i32, [vp, i32, vp, vp, vp, i32, vp, i32, vp]),
    "tb_current_axial_potential": (i32, [vp, i32, vp, vp, vp, i32, vp, vp]),
    "tb_axial_derivatives": (i32, [vp, vp, i32, vp, vp, vp, i32, i32, vp]),
    "tb_axial_coefficients": (i32, [vp, vp, i32, vp, vp, vp]),
    "tb_axial_points": (i32, [vp, i32, vp, i32, vp, vp, vp, vp]),
    "tb_jacobian_3d": (i32, [vp, i32, vp, vp, vp]),
    "tb_charge_radial": (i32, [vp, vp, i32, vp, vp]),
    "tb_trace": (i32, [vp, i32, vp]),
    "tb_trace_order_workspace_bytes": (sz, [i32]),
    "tb_trace_order": (i32, [vp, i32, vp, vp, vp]),
    "tb_trace_noalloc": (i32, [vp, i32, i32, vp]),
    "tb_focus_workspace_bytes": (sz, []),
    "tb_focus_position": (i32, [vp, i32, C.c_longlong, vp, vp, vp]),
    "tb_sample_beam": (i32, [vp, vp, i32, vp, vp]),
    "tb_trace_noalloc_max_particles": (i32, []),
    "tb_trace_noalloc_scratch_bytes": (sz, []),
    "tb_trace_split_chunks": (i32, [i32, C.c_longlong]),
    "tb_trace_split_workspace_bytes": (sz, [i32, i32]),
    "tb_trace_split": (i32, [vp, i32, vp, vp]),
    "tb_mlp_nparams": (i32, []),
    "tb_mlp_workspace_bytes": (sz, [i32]),
    "tb_gemm_bf16": (i32, [i32, i32, i32, vp, C.c_long, C.c_long, vp, C.c_long, C.c_long, vp, C.c_long, i32, i32, vp]),
    "tb_mlp_forward": (i32, [vp, vp, vp, i32, vp, vp, vp]),
    "tb_mlp_grad": (i32, [vp, vp, vp, vp, i32, vp, vp, vp, vp]),
    "tb_mlp_adam": (i32, [vp, vp, vp, vp, vp, C.c_float, C.c_float, C.c_float, C.c_float, i32, vp]),
    "tb_mlp_to_bf16": (i32, [vp, vp, vp]),
    "tb_mlp_gather": (i32, [i32, vp, vp, vp, vp, vp, vp]),
    "tb_eval_program": (i32, [vp, vp, i32, vp, vp, vp]),
}

TIER_R = [n for n in _SIG if not n.startswith("tb_")]

for _name, (_res, _args) in _SIG.items():
    _f = getattr(backend_lib, _name)
    _f.restype = _res
    _f.argtypes = _args


def last_error() -> str:
    return backend_lib.tb_last_error().decode()


def _call(name, *args):
    """Call a Tier R function; raise if the library recorded an error."""
    backend_lib.tb_clear_error()
    f = getattr(backend_lib, name)
    new_args = [ensure_contiguous_aligned(a) if isinstance(a, np.ndarray) else a for a in args]
    out = f(*new_args)
    err = last_error()
    if err:
        raise RuntimeError(f"{name}: {err}")
    return out


def check(status: int, name: str = "") -> None:
    """Raise if a Tier B call returned non-zero."""
    if status != 0:
        raise RuntimeError(f"{name}: {last_error()}")


# ---------------------------------------------------------------------------------------
# Reference-named wrappers (traceon/backend/__init__.py:355-780)
# ---------------------------------------------------------------------------------------
ellipkm1 = np.vectorize(lambda p: _call("ellipkm1", float(p)), otypes=[np.float64])
ellipk = np.vectorize(lambda k: _call("ellipk", float(k)), otypes=[np.float64])
ellipem1 = np.vectorize(lambda p: _call("ellipem1", float(p)), otypes=[np.float64])
ellipe = np.vectorize(lambda k: _call("ellipe", float(k)), otypes=[np.float64])


def _v3(v):
    v = np.asarray(v, dtype=np.float64)
    if v.shape == (2,):
        # a 2D (x, z) node: the C side reads coordinates 0 and 2
        return np.array([v[0], 0.0, v[1]])
    return np.ascontiguousarray(v)


def higher_order_normal_radial(alpha: float, vertices) -> np.ndarray:
    vertices = np.asarray(vertices, dtype=np.float64)
    assert vertices.shape == (4, 3)
    normal = np.zeros(2)
    _call("higher_order_normal_radial", float(alpha), _v3(vertices[0]), _v3(vertices[2]), _v3(vertices[3]), _v3(vertices[1]), normal)
    assert np.isclose(np.linalg.norm(normal), 1.0)
    return normal


def normal_2d(p1, p2) -> np.ndarray:
    normal = np.zeros(2)
    _call("normal_2d", np.asarray(p1, dtype=np.float64), np.asarray(p2, dtype=np.float64), normal)
    return normal


def position_and_jacobian_radial(alpha, v1, v2, v3, v4):
    """Returns (jac, pos) with v1..v4 the nodes in parameter order (backend/__init__.py:434-451)."""
    pos = np.zeros(2)
    jac = dbl(0.0)
    _call("position_and_jacobian_radial", float(alpha), _v3(v1), _v3(v2), _v3(v3), _v3(v4), pos, C.pointer(jac))
    return jac.value, pos


def delta_position_and_jacobian_radial(alpha, v1, v2, v3, v4):
    pos = np.zeros(2)
    jac = dbl(0.0)
    _call("delta_position_and_jacobian_radial", float(alpha), _v3(v1), _v3(v2), _v3(v3), _v3(v4), pos, C.pointer(jac))
    return jac.value, pos


def potential_radial_ring(r0, z0, delta_r, delta_z):
    return _call("potential_radial_ring", float(r0), float(z0), float(delta_r), float(delta_z), None)


def dr1_potential_radial_ring(r0, z0, delta_r, delta_z):
    return _call("dr1_potential_radial_ring", float(r0), float(z0), float(delta_r), float(delta_z), None)


def dz1_potential_radial_ring(r0, z0, delta_r, delta_z):
    return _call("dz1_potential_radial_ring", float(r0), float(z0), float(delta_r), float(delta_z), None)


def flux_density_to_charge_factor(K: float) -> float:
    return _call("flux_density_to_charge_factor", float(K))


def current_potential_axial_radial_ring(z0, r, z):
    return _call("current_potential_axial_radial_ring", float(z0), float(r), float(z))


def current_field_radial_ring(x0, y0, x, y):
    res = np.zeros(2)
    _call("current_field_radial_ring", float(x0), float(y0), float(x), float(y), res)
    return res


def axial_derivatives_radial(z, charges, jac_buffer, pos_buffer):
    z = np.ascontiguousarray(z, dtype=np.float64)
    derivs = np.zeros((z.size, DERIV_2D_MAX))
    assert jac_buffer.shape == (len(charges), N_QUAD_2D)
    assert pos_buffer.shape == (len(charges), N_QUAD_2D, 2)
    _call("axial_derivatives_radial", derivs, np.asarray(charges, dtype=np.float64), jac_buffer, pos_buffer, len(charges), z, len(z))
    return derivs


def potential_radial(point, charges, jac_buffer, pos_buffer) -> float:
    point = np.asarray(point, dtype=np.float64)
    assert point.shape == (3,)
    assert jac_buffer.shape == (len(charges), N_QUAD_2D)
    assert pos_buffer.shape == (len(charges), N_QUAD_2D, 2)
    return _call("potential_radial", point, np.asarray(charges, dtype=np.float64), jac_buffer, pos_buffer, len(charges))


def potential_radial_derivs(point, z, coeffs) -> float:
    assert coeffs.shape == (len(z) - 1, DERIV_2D_MAX, 6)
    return _call("potential_radial_derivs", np.asarray(point, dtype=np.float64), z, coeffs, len(z))


def charge_radial(vertices, charge: float) -> float:
    vertices = np.asarray(vertices, dtype=np.float64)
    assert vertices.shape == (len(vertices), 3)
    return _call("charge_radial", vertices, float(charge))


def field_radial(point, charges, jac_buffer, pos_buffer):
    point = np.asarray(point, dtype=np.float64)
    assert point.shape == (3,)
    assert jac_buffer.shape == (len(charges), N_QUAD_2D)
    assert pos_buffer.shape == (len(charges), N_QUAD_2D, 2)
    field = np.zeros(3)
    _call("field_radial", point, field, np.asarray(charges, dtype=np.float64), jac_buffer, pos_buffer, len(charges))
    return field


def field_radial_derivs(point, z, coeffs):
    point = np.asarray(point, dtype=np.float64)
    assert coeffs.shape == (len(z) - 1, DERIV_2D_MAX, 6)
    field = np.zeros(3)
    _call("field_radial_derivs", point, field, z, coeffs, len(z))
    return field


def current_potential_axial(z, currents, jac_buffer, pos_buffer) -> float:
    N = len(currents)
    assert jac_buffer.shape == (N, N_TRIANGLE_QUAD) and pos_buffer.shape == (N, N_TRIANGLE_QUAD, 3)
    return _call("current_potential_axial", float(z), np.asarray(currents, dtype=np.float64), jac_buffer, pos_buffer, N)


def current_field_radial(p0, currents, jac_buffer, pos_buffer):
    p0 = np.asarray(p0, dtype=np.float64)
    N = len(currents)
    assert jac_buffer.shape == (N, N_TRIANGLE_QUAD) and pos_buffer.shape == (N, N_TRIANGLE_QUAD, 3)
    result = np.zeros(3)
    _call("current_field_radial", p0, result, np.asarray(currents, dtype=np.float64), jac_buffer, pos_buffer, N)
    return result


def current_axial_derivatives_radial(z, currents, jac_buffer, pos_buffer):
    z = np.ascontiguousarray(z, dtype=np.float64)
    N = len(currents)
    derivs = np.zeros((z.size, DERIV_2D_MAX))
    _call("current_axial_derivatives_radial", derivs, np.asarray(currents, dtype=np.float64), jac_buffer, pos_buffer, N, z, len(z))
    return derivs


def fill_jacobian_buffer_radial(vertices):
    vertices = np.ascontiguousarray(vertices, dtype=np.float64)
    assert vertices.shape == (len(vertices), 4, 3)
    N = len(vertices)
    jac_buffer = np.zeros((N, N_QUAD_2D))
    pos_buffer = np.zeros((N, N_QUAD_2D, 2))
    if N:
        _call("fill_jacobian_buffer_radial", jac_buffer, pos_buffer, vertices, N)
    return jac_buffer, pos_buffer


def fill_jacobian_buffer_3d(vertices):
    vertices = np.ascontiguousarray(vertices, dtype=np.float64)
    N = len(vertices)
    assert vertices.shape == (N, 3, 3)
    jac_buffer = np.zeros((N, N_TRIANGLE_QUAD))
    pos_buffer = np.zeros((N, N_TRIANGLE_QUAD, 3))
    if N:
        _call("fill_jacobian_buffer_3d", jac_buffer, pos_buffer, vertices, N)
    return jac_buffer, pos_buffer


def normal_3d(tri) -> np.ndarray:
    """Unit normal of a flat triangle (backend/__init__.py:377-380)."""
    tri = np.ascontiguousarray(tri, dtype=np.float64)
    assert tri.shape == (3, 3)
    normal = np.zeros(3)
    _call("normal_3d", tri, normal)
    return normal


def position_and_jacobian_3d(alpha: float, beta: float, triangle):
    """(jac, pos) of a flat triangle's affine map (backend/__init__.py:424-432)."""
    triangle = np.ascontiguousarray(triangle, dtype=np.float64)
    assert triangle.shape == (3, 3)
    pos = np.zeros(3)
    jac = dbl(0.0)
    _call("position_and_jacobian_3d", float(alpha), float(beta), triangle, pos, C.pointer(jac))
    return jac.value, pos


def triangle_areas(triangles) -> np.ndarray:
    """Areas of flat triangles (backend/__init__.py:776-780)."""
    triangles = np.ascontiguousarray(triangles, dtype=np.float64)
    assert triangles.shape == (len(triangles), 3, 3)
    out = np.zeros(len(triangles))
    if len(triangles):
        _call("triangle_areas", triangles, out, len(triangles))
    return out


def self_potential_radial(vertices) -> float:
    """The singular self potential of one element (the integral the reference hands to
    scipy.integrate.quad at backend/__init__.py:688-692), computed on the GPU."""
    vertices = np.ascontiguousarray(vertices, dtype=np.float64)
    assert vertices.shape == (4, 3)
    return _call("self_potential_radial_integral", vertices)


def self_field_dot_normal_radial(vertices, K: float) -> float:
    vertices = np.ascontiguousarray(vertices, dtype=np.float64)
    assert vertices.shape == (4, 3)
    return _call("self_field_dot_normal_radial_integral", vertices, float(K))


def fill_matrix_radial(matrix, lines, excitation_types, excitation_values, jac_buffer, pos_buffer, start_index, end_index):
    N = len(lines)
    assert np.all(lines[:, :, 1] == 0.0)
    assert matrix.shape == (N, N)
    assert lines.shape == (N, 4, 3)
    assert excitation_types.shape == (N,)
    assert excitation_values.shape == (N,)
    assert jac_buffer.shape == (N, N_QUAD_2D)
    assert pos_buffer.shape == (N, N_QUAD_2D, 2)
    assert 0 <= start_index < N and 0 <= end_index < N and start_index <= end_index, (start_index, end_index, N)
    assert matrix.flags["C_CONTIGUOUS"], "the matrix is filled in place"
    _call("fill_matrix_radial", matrix, lines, excitation_types, excitation_values, jac_buffer, pos_buffer, N, matrix.shape[0],
          int(start_index), int(end_index))


def plane_intersection(positions, p0, normal):
    p0 = np.asarray(p0, dtype=np.float64)
    normal = np.asarray(normal, dtype=np.float64)
    assert p0.shape == (3,) and normal.shape == (3,)
    assert positions.shape == (len(positions), 6)
    result = np.zeros(6)
    found = _call("plane_intersection", p0, normal, np.ascontiguousarray(positions, dtype=np.float64), len(positions), result)
    if not found:
        raise ValueError("Plane intersection not found. Does the trajectory actually cross the plane?")
    return result


# ---- tracing through the reference's argument structs ------------------------------------
class EffectivePointCharges2D(C.Structure):
    """struct effective_point_charges_2d (radial.c:10-15), backend/__init__.py:107-128."""
    _fields_ = [("charges", dbl_p), ("jacobians", dbl_p), ("positions", dbl_p), ("N", C.c_size_t)]

    def __init__(self, eff, *args, **kwargs):
        super().__init__(*args, **kwargs)
        self.charges_arr = ensure_contiguous_aligned(np.asarray(eff.charges, dtype=np.float64))
        self.jacobians_arr = ensure_contiguous_aligned(np.asarray(eff.jacobians, dtype=np.float64))
        self.positions_arr = ensure_contiguous_aligned(np.asarray(eff.positions, dtype=np.float64))
        self.charges = self.charges_arr.ctypes.data_as(dbl_p)
        self.jacobians = self.jacobians_arr.ctypes.data_as(dbl_p)
        self.positions = self.positions_arr.ctypes.data_as(dbl_p)
        self.N = len(eff)


EffectivePointCharges3D = EffectivePointCharges2D


class FieldEvaluationArgsRadial(C.Structure):
    """struct field_evaluation_args (radial.c:175-180), backend/__init__.py:190-223."""
    _fields_ = [("elec_charges", vp), ("mag_charges", vp), ("current_charges", vp), ("bounds", dbl_p)]

    def __init__(self, elec, mag, current, bounds, *args, **kwargs):
        super().__init__(*args, **kwargs)
        assert bounds is None or np.asarray(bounds).shape == (3, 2)
        self.eff_elec = EffectivePointCharges2D(elec)
        self.eff_mag = EffectivePointCharges2D(mag)
        self.eff_current = EffectivePointCharges3D(current)
        self.elec_charges = C.cast(C.pointer(self.eff_elec), vp)
        self.mag_charges = C.cast(C.pointer(self.eff_mag), vp)
        self.current_charges = C.cast(C.pointer(self.eff_current), vp)
        if bounds is None:
            self.bounds = None
        else:
            self.bounds_arr = ensure_contiguous_aligned(np.asarray(bounds, dtype=np.float64))
            self.bounds = self.bounds_arr.ctypes.data_as(dbl_p)


class FieldDerivsArgs(C.Structure):
    """struct field_derivs_args (three_d.c:26-31), backend/__init__.py:259-280."""
    _fields_ = [("z_interpolation", dbl_p), ("electrostatic_axial_coeffs", dbl_p), ("magnetostatic_axial_coeffs", dbl_p),
                ("N_z", C.c_size_t)]

    def __init__(self, z, elec, mag, *args, **kwargs):
        super().__init__(*args, **kwargs)
        assert z.shape == (len(z),) and elec.shape[0] == len(z) - 1 and mag.shape[0] == len(z) - 1
        self.z_arr = ensure_contiguous_aligned(np.asarray(z, dtype=np.float64))
        self.elec_arr = ensure_contiguous_aligned(np.asarray(elec, dtype=np.float64))
        self.mag_arr = ensure_contiguous_aligned(np.asarray(mag, dtype=np.float64))
        self.z_interpolation = self.z_arr.ctypes.data_as(dbl_p)
        self.electrostatic_axial_coeffs = self.elec_arr.ctypes.data_as(dbl_p)
        self.magnetostatic_axial_coeffs = self.mag_arr.ctypes.data_as(dbl_p)
        self.N_z = len(z)


def traceable(name: str):
    """Function pointer of one of the library's traceable field functions
    (as ``backend.field_fun((name, backend.backend_lib))`` in the reference)."""
    return field_fun((name, backend_lib))


def trace_particle_wrapper(position_, velocity_, fill_positions_fun):
    """The reference's 1e5-step block loop (backend/__init__.py:382-420)."""
    position = np.array(position_, dtype=np.float64)
    velocity = np.array(velocity_, dtype=np.float64)
    assert position.shape == (3,) and velocity.shape == (3,)
    pos_blocks, times_blocks = [], []
    times = np.zeros(TRACING_BLOCK_SIZE)
    positions = np.zeros((TRACING_BLOCK_SIZE, 6))
    positions[0] = np.concatenate((position, velocity))
    while True:
        N = fill_positions_fun(times, positions)
        pos_blocks.append(positions[1:N] if len(pos_blocks) > 0 else positions[:N])
        times_blocks.append(times[1:N] if len(times_blocks) > 0 else times[:N])
        if N != TRACING_BLOCK_SIZE:
            break
        times = np.zeros(TRACING_BLOCK_SIZE)
        positions = np.zeros((TRACING_BLOCK_SIZE, 6))
        positions[0] = pos_blocks[-1][-1]
        times[0] = times_blocks[-1][-1]
    if len(pos_blocks) == 1:
        return times_blocks[0], pos_blocks[0]
    return np.concatenate(times_blocks), np.concatenate(pos_blocks)


def trace_particle(position, velocity, charge_over_mass, field, bounds, atol, args=None):
    bounds = np.ascontiguousarray(bounds, dtype=np.float64)

    def fill(T, P):
        n = _call("trace_particle", T, P, float(charge_over_mass), field, bounds, float(atol), args)
        if n == 0:
            raise RuntimeError("trace_particle: " + (last_error() or "no state written"))
        return n

    return trace_particle_wrapper(position, velocity, fill)
