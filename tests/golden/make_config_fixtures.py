"""Writes the golden fixtures of the five benchmark workloads (config_c0.npz ... config_c4.npz)
from the reference's own C backend (oracle/_ref, compiled from its sources by oracle/Makefile;
the oracle's restatement where that build is absent), with the reference's Python steps
restated in oracle/oracle.py: QUADPACK diagonal (converged, see TIGHT, and at the reference's
tolerance), numpy.linalg.solve, scipy axial coefficients, the trace block loop and
plane_intersection.  The GPU tests in
tests/test_gpu_configs.py compare the HIP path against these files.

    python tests/golden/make_config_fixtures.py c0 c1 c2 c3 c4     (all: no argument)

Geometries and beams come from traceon_amd/configs.py (host-side numpy only); each file
records a hash of the element array so that a changed geometry is detected.  What each file
holds is listed in its `contents` entry.  Runtimes on 8 cores: c0 1 min, c1 6 min, c2 3 min,
c3 3 min, c4 10 min."""
import hashlib
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from traceon_amd import configs  # noqa: E402
from traceon_amd.excitation import Excitation  # noqa: E402

MM = 1e-3
DERIV_STRIDE = 8   # axial derivatives are stored at every 8th sample
QM = -configs.E_CHARGE / configs.M_ELECTRON


def lib():
    r = O.reference()
    return (r, "reference C backend (oracle/_ref)") if r is not None else (O.restatement(), "oracle restatement")


def line_hash(lines):
    return hashlib.sha256(np.ascontiguousarray(lines, dtype=np.float64).tobytes()).hexdigest()


def log(*a):
    print(*a, flush=True)


# QUADPACK at the reference's tolerance (epsabs = epsrel = 1e-9, backend/__init__.py:692, 708)
# accepts, on some elements of fine meshes, results 1e-5 from the integral: the integrals are
# ~6e-6 there and the absolute tolerance admits that.  The oracle's charges, fields and traces
# therefore come from the same integrands integrated to convergence; the reference-tolerance
# diagonal and its solution are kept beside them to check that the GPU solves the reference's
# own assembled system.
TIGHT = dict(epsabs=1e-30, epsrel=1e-13, limit=1000)
TIGHT_ATOL = 1e-10


def solve(L, lines, types, values, F=None, reference_tolerance=False):
    """Fill, QUADPACK diagonal (converged; with reference_tolerance also the reference's),
    numpy.linalg.solve.  Returns M (converged diagonal), d, F, q and, if asked, the
    reference-tolerance diagonal and charges."""
    t = time.time()
    M = L.fill_matrix(lines, types, values)
    d = L.diagonal(lines, types, values, **TIGHT)
    M[np.diag_indices(len(lines))] = d
    F = np.where(np.isin(types, [1, 2, 5]), values, 0.0) if F is None else F
    q = np.linalg.solve(M, F)
    extra = {}
    if reference_tolerance:
        d_ref = L.diagonal(lines, types, values)
        Mr = M.copy()
        Mr[np.diag_indices(len(lines))] = d_ref
        q_ref = np.linalg.solve(Mr, F)
        extra = {"diag_ref": d_ref, "q_ref": q_ref, "q_ref_refined": O.refine_long_double(Mr, F, q_ref, iters=2)}
        del Mr
        log(f"  reference-tolerance diagonal: worst {np.max(np.abs(d_ref - d) / np.abs(d)):.2e} from the converged one")
    log(f"  solve N={len(lines)}: {time.time() - t:.1f} s")
    return M, d, F, q, extra


def mv_vectors(seed, N):
    """The two vectors of the matrix checks: standard normal from default_rng(seed)."""
    return np.random.default_rng(seed).standard_normal((2, N))


def matrix_checks(M, d, seed, rows):
    """Off-diagonal matrix-vector products with two seeded vectors (they pin every entry through
    linearity), plus a few whole rows."""
    V = mv_vectors(seed, len(d))
    Moff = M @ V.T - d[:, None] * V.T
    return {"mv_seed": np.array(seed), "mv_offdiag": Moff.T, "rows_index": np.asarray(rows), "rows": M[rows].copy()}


def field_points(L, pts, q, jac, pos):
    pot = np.array(O.run_threaded(lambda p: L.potential_radial(p, q, jac, pos), list(pts)))
    fld = np.array(O.run_threaded(lambda p: L.field_radial(p, q, jac, pos), list(pts)))
    return pot, fld


def trace_set(L, P, V, fn, exit_plane, back_plane=None):
    """Trace each particle (thread pool) and keep its landing, accepted steps and final state."""
    t = time.time()
    res = O.run_threaded(lambda i: fn(P[i], V[i]), range(len(P)))
    land, steps, final, which = [], [], [], []
    for times, states in res:
        x, pl = O.landing(L, states, exit_plane, back_plane)
        land.append(x)
        steps.append(len(states) - 1)
        final.append(states[-1])
        which.append(0 if pl is exit_plane else (1 if pl is not None else -1))
    log(f"  traced {len(P)} particles in {time.time() - t:.1f} s")
    return {"landing": np.array(land), "steps": np.array(steps), "final": np.array(final), "plane": np.array(which)}


def axial(L, z, sources):
    d = np.zeros((len(z), 9))
    for (q, jac, pos, current) in sources:
        d += O.axial_derivatives_threaded(L, z, q, jac, pos, current=current)
    return d, O.quintic_spline_coefficients(z, d.T)


def save(name, meta, **arrays):
    path = os.path.join(HERE, name)
    np.savez_compressed(path, **arrays, meta=np.array(repr(meta)))
    log(f"wrote {path} ({os.path.getsize(path) / 1024:.0f} KiB)")


def c0():
    """Plain FP64 einzel lens, N = 3000: matrix, diagonal, charges, fields, 100 direct-BEM traces."""
    L, kind = lib()
    mesh, volts = configs.einzel_lens(3000)
    lines, types, values = configs.lines_types_values(mesh, volts)
    M, d, F, q, ref = solve(L, lines, types, values, reference_tolerance=True)
    qr = O.refine_long_double(M, F, q)
    jac, pos = L.fill_jacobian_buffer_radial(lines)
    mc = matrix_checks(M, d, 100, [0, 777, 1500, 2999])
    pts = np.array([[r, 0.0, z] for r in np.linspace(0.0, 0.9 * MM, 10) for z in np.linspace(-4.5 * MM, 9.5 * MM, 20)])
    pot, fld = field_points(L, pts, q, jac, pos)
    P, V, bounds, plane = configs.einzel_beam(100)
    tr = trace_set(L, P, V, lambda p, v: L.trace_bem(p, v, QM, (q, jac, pos), bounds=bounds), plane)
    save("config_c0.npz", {"kind": kind, "lines_sha256": line_hash(lines), "N": len(lines),
                           "contents": "off-diagonal M v for two seeded vectors, four rows, QUADPACK diagonal converged "
                                       "(diag) and at the reference's tolerance (diag_ref), LAPACK charges of both (q_lapack, "
                                       "q_ref) and both refined in long double, potential and field at 200 points, 100 "
                                       "direct-BEM traces (landing on z = 10 mm, accepted steps)"},
         diag=d, q_lapack=q, q_refined=qr, **ref, pts=pts, pot=pot, field=fld, P=P, V=V, bounds=bounds, **mc,
         **{f"trace_{k}": v for k, v in tr.items()})


def c1():
    """Mirror, N = 20,000: matrix checks, diagonal, charges (LAPACK and refined), fields at 256
    points of the 1000 x 1000 grid."""
    L, kind = lib()
    mesh, volts = configs.mirror(20000, seed=1)
    lines, types, values = configs.lines_types_values(mesh, volts)
    M, d, F, q, ref = solve(L, lines, types, values, reference_tolerance=True)
    t = time.time()
    qr = O.refine_long_double(M, F, q, iters=2)
    log(f"  refinement {time.time() - t:.1f} s")
    jac, pos = L.fill_jacobian_buffer_radial(lines)
    mc = matrix_checks(M, d, 101, [0, 9999, 19999])
    rng = np.random.default_rng(101)
    rng.standard_normal((2, len(lines)))   # the matrix-check vectors come first in this stream
    mc["rows"] = mc["rows"][:, ::10].copy()        # every 10th column
    del M
    r = np.linspace(0.0, 4.0e-3, 1000)
    z = np.linspace(-0.2e-3, 10.2e-3, 1000)
    ii = rng.integers(0, 1000, (256, 2))
    pts = np.column_stack([r[ii[:, 0]], np.zeros(256), z[ii[:, 1]]])
    pot, fld = field_points(L, pts, q, jac, pos)
    save("config_c1.npz", {"kind": kind, "lines_sha256": line_hash(lines), "N": len(lines),
                           "contents": "off-diagonal M v for two seeded vectors, rows 0, 9999, 19999 (every 10th column), "
                                       "QUADPACK diagonal converged (diag) and at the reference's tolerance (diag_ref), LAPACK "
                                       "charges of both and both refined in long double, potential and field at 256 grid points"},
         diag=d, q_lapack=q, q_refined=qr, **ref, pts=pts, pot=pot, field=fld, **mc)


def c2():
    """Reflecting lens, N = 8000: oracle solve, FieldRadialAxial over [-15, 20] mm with 7000
    samples, 512 axial traces and 16 direct-BEM traces of the batched beam's first particles."""
    L, kind = lib()
    mesh, volts = configs.reflecting_lens(8000)
    lines, types, values = configs.lines_types_values(mesh, volts)
    M, d, F, q, _ = solve(L, lines, types, values)
    del M
    jac, pos = L.fill_jacobian_buffer_radial(lines)
    z = np.linspace(-15.0 * MM, 20.0 * MM, 7000)
    t = time.time()
    dz, co = axial(L, z, [(q, jac, pos, False)])
    log(f"  axial series {time.time() - t:.1f} s")
    P, V, bounds, exit_plane, back_plane = configs.reflecting_beam(1_000_000)
    n_ax, n_bem = 512, 16
    zero = np.zeros_like(co)
    ta = trace_set(L, P[:n_ax], V[:n_ax], lambda p, v: L.trace_axial(p, v, QM, z, co, zero, bounds), exit_plane, back_plane)
    tb = trace_set(L, P[:n_bem], V[:n_bem], lambda p, v: L.trace_bem(p, v, QM, (q, jac, pos), bounds=bounds), exit_plane,
                   back_plane)
    save("config_c2.npz", {"kind": kind, "lines_sha256": line_hash(lines), "N": len(lines),
                           "contents": "LAPACK charges, axial samples, derivatives at every 8th sample, 512 axial and 16 direct-BEM traces of "
                                       "the first particles of reflecting_beam(10^6) (landing on z = 9.5 mm, else back "
                                       "through z = -4.5 mm; accepted steps; final state)"},
         q=q, z=z, derivs=dz[::DERIV_STRIDE], P=P[:n_ax], V=V[:n_ax], bounds=bounds, **{f"axial_{k}": v for k, v in ta.items()},
         **{f"bem_{k}": v for k, v in tb.items()})


def magnetic_problem(L):
    mesh, volts, perms, currents = configs.magnetic_lens(12000)
    exc = Excitation(mesh)
    exc.add_voltage(**volts)
    exc.add_magnetizable(**perms)
    exc.add_current(**currents)
    le, ne = exc.get_electrostatic_active_elements()
    lm, nm = exc.get_magnetostatic_active_elements()
    te = np.ones(len(le), dtype=np.uint8)
    ve = np.zeros(len(le))
    for n, idx in ne.items():
        ve[idx] = volts[n]
    tm = np.full(len(lm), 6, dtype=np.uint8)
    vm = np.zeros(len(lm))
    for n, idx in nm.items():
        vm[idx] = perms[n]
    tri = mesh.points[mesh.triangles.astype(np.int64)]
    I, cj, cp = O.coil_currents(L, tri, currents["coil"])
    return (le, te, ve), (lm, tm, vm), (I, cj, cp)


def c3_c4(which):
    """Magnetic + electrostatic lens, N = 12,000: the two oracle solves (coil RHS), the axial
    series of both fields over [-24, 22] mm, and traces: c3 = 512 axial and 16 direct-BEM traces
    of the 10 keV beam's first particles; c4 = 64 direct-BEM and axial traces of the off-axis
    beam (start radius 0.3-0.5 mm) that the learned surrogate serves."""
    L, kind = lib()
    (le, te, ve), (lm, tm, vm), (I, cj, cp) = magnetic_problem(L)
    _, _, _, qe, _ = solve(L, le, te, ve)
    F = O.magnetizable_rhs(L, lm, vm, I, cj, cp)
    _, _, _, qm, _ = solve(L, lm, tm, vm, F=F)
    je, pe = L.fill_jacobian_buffer_radial(le)
    jm, pm = L.fill_jacobian_buffer_radial(lm)
    z = np.linspace(-24.0 * MM, 22.0 * MM, int(7000 * 46 / 35))
    t = time.time()
    de, ce = axial(L, z, [(qe, je, pe, False)])
    dm, cm = axial(L, z, [(qm, jm, pm, False), (I, cj, cp, True)])
    log(f"  axial series {time.time() - t:.1f} s")
    if which == "c3":
        P, V, bounds, plane = configs.magnetic_beam(1_000_000)
        n_ax, n_bem = 512, 16
    else:
        P, V, bounds, plane = configs.magnetic_beam(64, seed=4, r_range=(0.3 * MM, 0.5 * MM))
        n_ax, n_bem = 64, 64
    ta = trace_set(L, P[:n_ax], V[:n_ax], lambda p, v: L.trace_axial(p, v, QM, z, ce, cm, bounds), plane)
    tb = trace_set(L, P[:n_bem], V[:n_bem],
                   lambda p, v: L.trace_bem(p, v, QM, (qe, je, pe), mag=(qm, jm, pm), cur=(I, cj, cp), bounds=bounds), plane)
    extra = {}
    if which == "c3":
        extra = dict(qe=qe, qm=qm, z=z, derivs_e=de[::DERIV_STRIDE], derivs_m=dm[::DERIV_STRIDE])
    else:
        # the same beam at atol = 1e-10: there the RK45 truncation error is far below 10 nm, so
        # landing differences measure field error rather than a diverging step sequence
        tight = dict(atol=TIGHT_ATOL)
        ta2 = trace_set(L, P[:n_ax], V[:n_ax], lambda p, v: L.trace_axial(p, v, QM, z, ce, cm, bounds, **tight), plane)
        tb2 = trace_set(L, P[:n_bem], V[:n_bem], lambda p, v: L.trace_bem(p, v, QM, (qe, je, pe), mag=(qm, jm, pm), cur=(I, cj, cp),
                                                                         bounds=bounds, **tight), plane)
        extra = {**{f"axial_tight_{k}": v for k, v in ta2.items()}, **{f"bem_tight_{k}": v for k, v in tb2.items()},
                 "tight_atol": TIGHT_ATOL}
    save(f"config_{which}.npz", {"kind": kind, "lines_sha256_e": line_hash(le), "lines_sha256_m": line_hash(lm),
                                 "N_e": len(le), "N_m": len(lm), "rings": len(I),
                                 "contents": ("LAPACK charges of both solves, axial derivatives of both series at every 8th sample, " if which == "c3"
                                              else "") + f"{n_ax} axial and {n_bem} direct-BEM traces (landing on z = 11 mm, "
                                                         "accepted steps, final state)" +
                                             ("; both again at atol = 1e-10 (axial_tight_*, bem_tight_*)" if which == "c4" else "")},
         P=P[:max(n_ax, n_bem)], V=V[:max(n_ax, n_bem)], bounds=bounds, **extra,
         **{f"axial_{k}": v for k, v in ta.items()}, **{f"bem_{k}": v for k, v in tb.items()})


def main(argv):
    todo = argv or ["c0", "c1", "c2", "c3", "c4"]
    for w in todo:
        log(f"== {w}")
        t = time.time()
        {"c0": c0, "c1": c1, "c2": c2, "c3": lambda: c3_c4("c3"), "c4": lambda: c3_c4("c4")}[w]()
        log(f"== {w} done in {time.time() - t:.0f} s")


if __name__ == "__main__":
    main(sys.argv[1:])
