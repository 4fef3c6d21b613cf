"""ORACLE — test infrastructure only.

ctypes front-end over the C restatement (``liboracle.so``, built from ``oracle.c``)
and, where it was built, over the reference's own C backend (``_ref/traceon_backend.so``).
It restates the Python-side steps of the reference that sit on the path:

* the log-singular diagonal through ``scipy.integrate.quad`` with the reference's
  tolerances (traceon/backend/__init__.py:688-708),
* the threaded row split of the fill (traceon/util.py:34-77, traceon/solver.py:206-253),
* the dense solve through ``numpy.linalg.solve`` (traceon/solver.py:173),
* the quintic-Hermite / not-a-knot axial coefficients (traceon/field.py:906-937),
* the block loop around ``trace_particle`` (traceon/backend/__init__.py:382-420).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg import
this module; the product path (``traceon_amd``) never does.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

import numpy as np
from numpy.ctypeslib import ndpointer

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "traceon_backend.so")

N_QUAD_2D = 16
N_TRIANGLE_QUAD = 12
DERIV_2D_MAX = 9
TRACING_BLOCK_SIZE = 100000
MU_0 = 1.25663706212e-06

_dbl = C.c_double
_sz = C.c_size_t
_vp = C.c_void_p


def _arr(dtype=np.float64):
    return ndpointer(dtype=dtype, flags=("C_CONTIGUOUS", "ALIGNED"))


field_fun = C.CFUNCTYPE(None, C.POINTER(_dbl), C.POINTER(_dbl), _vp, C.POINTER(_dbl), C.POINTER(_dbl))


class Charges2D(C.Structure):
    """struct effective_point_charges_2d (radial.c:10-15)."""
    _fields_ = [("charges", C.POINTER(_dbl)), ("jacobians", C.POINTER(_dbl)), ("positions", C.POINTER(_dbl)), ("N", _sz)]

    def __init__(self, q, jac, pos):
        super().__init__()
        self._keep = [np.ascontiguousarray(a, dtype=np.float64) for a in (q, jac, pos)]
        self.charges, self.jacobians, self.positions = (a.ctypes.data_as(C.POINTER(_dbl)) for a in self._keep)
        self.N = len(self._keep[0])


class FieldEvaluationArgs(C.Structure):
    """struct field_evaluation_args (radial.c:175-180)."""
    _fields_ = [("elec", _vp), ("mag", _vp), ("cur", _vp), ("bounds", C.POINTER(_dbl))]

    def __init__(self, elec, mag, cur, bounds=None):
        super().__init__()
        self._e, self._m, self._c = Charges2D(*elec), Charges2D(*mag), Charges2D(*cur)
        self.elec = C.cast(C.pointer(self._e), _vp)
        self.mag = C.cast(C.pointer(self._m), _vp)
        self.cur = C.cast(C.pointer(self._c), _vp)
        if bounds is None:
            self.bounds = None
        else:
            self._b = np.ascontiguousarray(bounds, dtype=np.float64).reshape(3, 2)
            self.bounds = self._b.ctypes.data_as(C.POINTER(_dbl))


class FieldDerivsArgs(C.Structure):
    """struct field_derivs_args (three_d.c:26-31)."""
    _fields_ = [("z", C.POINTER(_dbl)), ("elec", C.POINTER(_dbl)), ("mag", C.POINTER(_dbl)), ("N_z", _sz)]

    def __init__(self, z, elec, mag):
        super().__init__()
        self._keep = [np.ascontiguousarray(a, dtype=np.float64) for a in (z, elec, mag)]
        self.z, self.elec, self.mag = (a.ctypes.data_as(C.POINTER(_dbl)) for a in self._keep)
        self.N_z = len(z)


class SelfFdnArgs(C.Structure):
    """struct self_field_dot_normal_radial_args (radial.c:198-201)."""
    _fields_ = [("line", C.POINTER(_dbl)), ("K", _dbl)]


class OracleLib:
    """One of the two CPU libraries (restatement or reference build) with the
    reference's ctypes signatures (traceon/backend/__init__.py:287-328)."""

    def __init__(self, path: str):
        self.path = path
        self.lib = C.CDLL(path)
        L = self.lib
        a = _arr()
        sig = {
            "ellipkm1": (_dbl, [_dbl]), "ellipem1": (_dbl, [_dbl]), "ellipk": (_dbl, [_dbl]), "ellipe": (_dbl, [_dbl]),
            "potential_radial_ring": (_dbl, [_dbl, _dbl, _dbl, _dbl, _vp]),
            "dr1_potential_radial_ring": (_dbl, [_dbl, _dbl, _dbl, _dbl, _vp]),
            "dz1_potential_radial_ring": (_dbl, [_dbl, _dbl, _dbl, _dbl, _vp]),
            "current_potential_axial_radial_ring": (_dbl, [_dbl, _dbl, _dbl]),
            "current_field_radial_ring": (None, [_dbl, _dbl, _dbl, _dbl, a]),
            "axial_derivatives_radial_ring": (None, [_dbl, _dbl, _dbl, a]),
            "current_axial_derivatives_radial_ring": (None, [_dbl, _dbl, _dbl, a]),
            "flux_density_to_charge_factor": (_dbl, [_dbl]),
            "position_and_jacobian_radial": (None, [_dbl, a, a, a, a, a, C.POINTER(_dbl)]),
            "delta_position_and_jacobian_radial": (None, [_dbl, a, a, a, a, a, C.POINTER(_dbl)]),
            "higher_order_normal_radial": (None, [_dbl, a, a, a, a, a]),
            "normal_2d": (None, [a, a, a]),
            "fill_jacobian_buffer_radial": (None, [a, a, a, _sz]),
            "fill_jacobian_buffer_3d": (None, [a, a, a, _sz]),
            "charge_radial": (_dbl, [a, _dbl]),
            "fill_matrix_radial": (None, [a, a, _arr(np.uint8), a, a, a, _sz, _sz, C.c_int, C.c_int]),
            "self_potential_radial": (_dbl, [_dbl, _vp]),
            "self_field_dot_normal_radial": (_dbl, [_dbl, _vp]),
            "potential_radial": (_dbl, [a, a, a, a, _sz]),
            "field_radial": (None, [a, a, a, a, a, _sz]),
            "current_potential_axial": (_dbl, [_dbl, a, a, a, _sz]),
            "current_field_radial": (None, [a, a, a, a, a, _sz]),
            "axial_derivatives_radial": (None, [a, a, a, a, _sz, a, _sz]),
            "current_axial_derivatives_radial": (None, [a, a, a, a, _sz, a, _sz]),
            "potential_radial_derivs": (_dbl, [a, a, a, _sz]),
            "field_radial_derivs": (None, [a, a, a, a, _sz]),
            "trace_particle": (_sz, [a, a, _dbl, field_fun, a, _dbl, _vp]),
            "plane_intersection": (C.c_bool, [a, a, a, _sz, a]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        # fill_matrix_radial returns int in the restatement and void in the reference
        self.is_reference = os.path.basename(path).startswith("traceon_backend")
        self.field_radial_traceable = C.cast(L.field_radial_traceable, _vp).value
        self.field_radial_derivs_traceable = C.cast(L.field_radial_derivs_traceable, _vp).value

    # -- element helpers -----------------------------------------------------------------
    @staticmethod
    def _nodes(line):
        line = np.ascontiguousarray(line, dtype=np.float64)
        return line[0], line[2], line[3], line[1]

    def position_and_jacobian_radial(self, alpha, line):
        pos = np.zeros(2)
        jac = _dbl(0.0)
        self.lib.position_and_jacobian_radial(alpha, *self._nodes(line), pos, C.byref(jac))
        return jac.value, pos

    def delta_position_and_jacobian_radial(self, alpha, line):
        pos = np.zeros(2)
        jac = _dbl(0.0)
        self.lib.delta_position_and_jacobian_radial(alpha, *self._nodes(line), pos, C.byref(jac))
        return jac.value, pos

    def higher_order_normal_radial(self, alpha, line):
        n = np.zeros(2)
        self.lib.higher_order_normal_radial(alpha, *self._nodes(line), n)
        return n

    def current_field_radial_ring(self, x0, y0, x, y):
        out = np.zeros(2)
        self.lib.current_field_radial_ring(x0, y0, x, y, out)
        return out

    def fill_jacobian_buffer_radial(self, lines):
        lines = np.ascontiguousarray(lines, dtype=np.float64)
        N = len(lines)
        jac, pos = np.zeros((N, N_QUAD_2D)), np.zeros((N, N_QUAD_2D, 2))
        self.lib.fill_jacobian_buffer_radial(jac, pos, lines, N)
        return jac, pos

    def fill_jacobian_buffer_3d(self, tris):
        tris = np.ascontiguousarray(tris, dtype=np.float64)
        N = len(tris)
        jac, pos = np.zeros((N, N_TRIANGLE_QUAD)), np.zeros((N, N_TRIANGLE_QUAD, 3))
        self.lib.fill_jacobian_buffer_3d(jac, pos, tris, N)
        return jac, pos

    def charge_radial(self, line, q):
        return self.lib.charge_radial(np.ascontiguousarray(line, dtype=np.float64), q)

    # -- fill --------------------------------------------------------------------------------
    def fill_matrix_rows(self, M, lines, types, values, jac, pos, start, end):
        self.lib.fill_matrix_radial(M, lines, types, values, jac, pos, len(lines), M.shape[1], int(start), int(end))

    def fill_matrix(self, lines, types, values, jac=None, pos=None, threads=None):
        """Off-diagonal fill with the reference's row split over threads (util.py:57-77)."""
        lines = np.ascontiguousarray(lines, dtype=np.float64)
        types = np.ascontiguousarray(types, dtype=np.uint8)
        values = np.ascontiguousarray(values, dtype=np.float64)
        if jac is None:
            jac, pos = self.fill_jacobian_buffer_radial(lines)
        N = len(lines)
        M = np.zeros((N, N))
        threads = threads or max(1, (os.cpu_count() or 2))
        chunks = [c for c in np.array_split(np.arange(N), threads) if len(c)]
        ts = [threading.Thread(target=self.fill_matrix_rows, args=(M, lines, types, values, jac, pos, c[0], c[-1]))
              for c in chunks]
        for t in ts:
            t.start()
        for t in ts:
            t.join()
        return M

    # -- singular diagonal through QUADPACK (backend/__init__.py:688-708) ----------------------
    def self_potential_quad(self, line, epsabs=1e-9, epsrel=1e-9, limit=250):
        from scipy import LowLevelCallable
        from scipy.integrate import quad
        line = np.ascontiguousarray(line, dtype=np.float64)
        ll = LowLevelCallable(self.lib.self_potential_radial, user_data=line.ctypes.data_as(_vp))
        return quad(ll, -1, 1, points=(0,), epsabs=epsabs, epsrel=epsrel, limit=limit)[0]

    def self_field_dot_normal_quad(self, line, K, epsabs=1e-9, epsrel=1e-9, limit=250):
        from scipy import LowLevelCallable
        from scipy.integrate import quad
        line = np.ascontiguousarray(line, dtype=np.float64)
        args = SelfFdnArgs()
        args.K = float(K)
        args.line = line.ctypes.data_as(C.POINTER(_dbl))
        ll = LowLevelCallable(self.lib.self_field_dot_normal_radial, user_data=C.cast(C.pointer(args), _vp))
        return quad(ll, -1, 1, points=(0,), epsabs=epsabs, epsrel=epsrel, limit=limit)[0]

    def diagonal(self, lines, types, values, **kw):
        """solver.py:239-247: self potential, or self flux - 1 for dielectric/magnetizable rows."""
        d = np.zeros(len(lines))
        for i, (line, t, v) in enumerate(zip(lines, types, values)):
            if t in (3, 6):
                d[i] = self.self_field_dot_normal_quad(line, v, **kw) - 1.0
            else:
                d[i] = self.self_potential_quad(line, **kw)
        return d

    def matrix(self, lines, types, values, threads=None):
        M = self.fill_matrix(lines, types, values, threads=threads)
        M[np.diag_indices(len(lines))] = self.diagonal(lines, types, values)
        return M

    # -- field sums -------------------------------------------------------------------------
    def potential_radial(self, p, q, jac, pos):
        return self.lib.potential_radial(np.asarray(p, dtype=np.float64), q, jac, pos, len(q))

    def field_radial(self, p, q, jac, pos):
        out = np.zeros(3)
        self.lib.field_radial(np.asarray(p, dtype=np.float64), out, q, jac, pos, len(q))
        return out

    def current_field_radial(self, p, I, jac, pos):
        out = np.zeros(3)
        self.lib.current_field_radial(np.asarray(p, dtype=np.float64), out, I, jac, pos, len(I))
        return out

    def current_potential_axial(self, z, I, jac, pos):
        return self.lib.current_potential_axial(float(z), I, jac, pos, len(I))

    def axial_derivatives_radial(self, z, q, jac, pos):
        z = np.ascontiguousarray(z, dtype=np.float64)
        d = np.zeros((len(z), DERIV_2D_MAX))
        self.lib.axial_derivatives_radial(d, q, jac, pos, len(q), z, len(z))
        return d

    def current_axial_derivatives_radial(self, z, I, jac, pos):
        z = np.ascontiguousarray(z, dtype=np.float64)
        d = np.zeros((len(z), DERIV_2D_MAX))
        self.lib.current_axial_derivatives_radial(d, I, jac, pos, len(I), z, len(z))
        return d

    def potential_radial_derivs(self, p, z, coeffs):
        return self.lib.potential_radial_derivs(np.asarray(p, dtype=np.float64), z, coeffs, len(z))

    def field_radial_derivs(self, p, z, coeffs):
        out = np.zeros(3)
        self.lib.field_radial_derivs(np.asarray(p, dtype=np.float64), out, z, coeffs, len(z))
        return out

    # -- tracing ----------------------------------------------------------------------------
    def trace(self, position, velocity, qm, fn_addr, args_struct, bounds, atol):
        """trace_particle with the reference's block loop (backend/__init__.py:382-420)."""
        bounds = np.ascontiguousarray(bounds, dtype=np.float64).reshape(3, 2)
        fn = field_fun(fn_addr)
        argp = C.cast(C.pointer(args_struct), _vp) if args_struct is not None else None
        times = np.zeros(TRACING_BLOCK_SIZE)
        states = np.zeros((TRACING_BLOCK_SIZE, 6))
        states[0, :3] = position
        states[0, 3:] = velocity
        tb, sb = [], []
        while True:
            n = self.lib.trace_particle(times, states, qm, fn, bounds, atol, argp)
            sb.append(states[1:n].copy() if sb else states[:n].copy())
            tb.append(times[1:n].copy() if tb else times[:n].copy())
            if n != TRACING_BLOCK_SIZE:
                break
            last_s, last_t = sb[-1][-1].copy(), tb[-1][-1]
            times = np.zeros(TRACING_BLOCK_SIZE)
            states = np.zeros((TRACING_BLOCK_SIZE, 6))
            states[0], times[0] = last_s, last_t
        return np.concatenate(tb), np.concatenate(sb)

    def trace_bem(self, position, velocity, qm, elec, mag=None, cur=None, bounds=None, field_bounds=None, atol=1e-8):
        empty2 = (np.zeros(0), np.zeros((0, 16)), np.zeros((0, 16, 2)))
        empty3 = (np.zeros(0), np.zeros((0, 12)), np.zeros((0, 12, 3)))
        args = FieldEvaluationArgs(elec if elec is not None else empty2, mag if mag is not None else empty2,
                                   cur if cur is not None else empty3, field_bounds)
        return self.trace(position, velocity, qm, self.field_radial_traceable, args, bounds, atol)

    def trace_axial(self, position, velocity, qm, z, elec_coeffs, mag_coeffs, bounds, atol=1e-8):
        args = FieldDerivsArgs(z, elec_coeffs, mag_coeffs)
        return self.trace(position, velocity, qm, self.field_radial_derivs_traceable, args, bounds, atol)

    def plane_intersection(self, states, p0, normal):
        out = np.zeros(6)
        states = np.ascontiguousarray(states, dtype=np.float64)
        ok = self.lib.plane_intersection(np.asarray(p0, dtype=np.float64), np.asarray(normal, dtype=np.float64),
                                         states, len(states), out)
        if not ok:
            raise ValueError("Plane intersection not found. Does the trajectory actually cross the plane?")
        return out


def quintic_spline_coefficients(z, derivs):
    """field.py:906-937: per derivative order i<=6 a quintic Hermite (BPoly.from_derivatives
    with value, first and second derivative at each knot), for i=7,8 a not-a-knot CubicSpline;
    coefficients (Nz-1, 9, 6) highest power first, cubic ones in slots 2..5."""
    from scipy.interpolate import BPoly, PPoly, CubicSpline
    z = np.asarray(z, dtype=np.float64)
    derivs = np.asarray(derivs)           # (9, Nz)
    c = np.zeros((z.size - 1, 9, 6))
    for i, d in enumerate(derivs):
        if i + 2 < len(derivs):
            bp = BPoly.from_derivatives(z, np.array([d, derivs[i + 1], derivs[i + 2]]).T)
            pp = PPoly.from_bernstein_basis(bp)
            c[:, i, 0:] = pp.c.T
        else:
            pp = CubicSpline(z, d)
            c[:, i, 2:] = pp.c.T
    return c


# -- whole-workload drivers (the reference's Python steps around the C calls) --------------
def _threads(threads=None):
    return threads or max(1, len(os.sched_getaffinity(0)))


def run_threaded(fn, items, threads=None):
    """fn(item) over items on a thread pool (the C calls release the GIL), results in order."""
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(_threads(threads)) as ex:
        return list(ex.map(fn, items))


def axial_derivatives_threaded(lib, z, q, jac, pos, current=False, threads=None):
    """field.py:983: the on-axis derivatives of order 0..8 split over threads by z."""
    z = np.ascontiguousarray(z, dtype=np.float64)
    chunks = [c for c in np.array_split(np.arange(len(z)), 4 * _threads(threads)) if len(c)]
    f = lib.current_axial_derivatives_radial if current else lib.axial_derivatives_radial
    parts = run_threaded(lambda c: f(z[c], q, jac, pos), chunks, threads)
    return np.concatenate(parts, axis=0)


def centres_and_normals(lib, lines):
    """Element centres (alpha = 0) and normals there, solver.py:100-109, 142-147."""
    c = np.array([lib.position_and_jacobian_radial(0.0, ln)[1] for ln in lines])
    n = np.array([lib.higher_order_normal_radial(0.0, ln) for ln in lines])
    return c, n


def coil_currents(lib, triangles, total_current):
    """solver.py:317-348: ring sources of a coil's triangles with density I / area."""
    cj, cp = lib.fill_jacobian_buffer_3d(triangles)
    return np.full(len(cj), total_current / np.sum(cj)), cj, cp


def magnetizable_rhs(lib, lines, K, I, cj, cp):
    """solver.py:118-150 for magnetizable rows: -factor(mu_r) (H_pre(centre) . n), H_pre the
    coil field (current_field_radial)."""
    c, n = centres_and_normals(lib, lines)
    H = np.array(run_threaded(lambda p: lib.current_field_radial(np.array([p[0], 0.0, p[1]]), I, cj, cp), list(c)))
    fac = 2.0 * (K - 1.0) / (1.0 + K)
    return -fac * (H[:, 0] * n[:, 0] + H[:, 2] * n[:, 1])


def refine_long_double(M, F, x, iters=3, block=1024):
    """Iterative refinement of x against M with residuals in long double (80-bit): returns the
    solution of the assembled system to about the rounding of FP64, and the factor's pivots."""
    import scipy.linalg as sla
    lu = sla.lu_factor(M)
    xl = np.asarray(x, dtype=np.longdouble).copy()
    Fl = np.asarray(F, dtype=np.longdouble)
    for _ in range(iters):
        r = np.empty(len(F), dtype=np.longdouble)
        for r0 in range(0, len(F), block):
            Ml = M[r0:r0 + block].astype(np.longdouble)
            r[r0:r0 + block] = Fl[r0:r0 + block] - Ml @ xl
        xl += sla.lu_solve(lu, r.astype(np.float64))
    return xl.astype(np.float64)


def landing(lib, states, exit_plane, back_plane=None):
    """The landing state: where the trajectory last crosses exit_plane, else back_plane
    (plane_intersection, tracing.c:184-223); NaN row if neither is crossed."""
    for pl in (exit_plane, back_plane):
        if pl is None:
            continue
        try:
            return lib.plane_intersection(states, pl[0], pl[1]), pl
        except ValueError:
            pass
    return np.full(6, np.nan), None


_cache = {}


def restatement() -> OracleLib:
    if "o" not in _cache:
        if not os.path.isfile(ORACLE_SO):
            raise RuntimeError(f"{ORACLE_SO} not built: run `make -C oracle`")
        _cache["o"] = OracleLib(ORACLE_SO)
    return _cache["o"]


def reference() -> OracleLib | None:
    if "r" not in _cache:
        _cache["r"] = OracleLib(REF_SO) if os.path.isfile(REF_SO) else None
    return _cache["r"]


def focus_position(positions):
    """focus.py:9-43 restated: numpy.linalg.lstsq of the 2N x 3 system in (z, x, y) built from
    the last state of every trajectory."""
    last = np.array([p[-1] for p in positions], dtype=np.float64)
    ax, ay = last[:, 3] / last[:, 5], last[:, 4] / last[:, 5]
    x, y, z = last[:, 0], last[:, 1], last[:, 2]
    n = len(last)
    A = np.zeros((2 * n, 3))
    A[:n, 0], A[n:, 0] = -ax, -ay
    A[:n, 1] = 1.0
    A[n:, 2] = 1.0
    rhs = np.concatenate([x - z * ax, y - z * ay])
    zf, xf, yf = np.linalg.lstsq(A, rhs, rcond=None)[0]
    return np.array([xf, yf, zf])
