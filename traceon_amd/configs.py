"""Synthetic geometries of the benchmark workloads (SI units, metres).

* ``einzel_lens``      three coaxial tube electrodes (inner radius 1 mm, lengths 2/1/2 mm,
                       0.5 mm gaps) at 0/1800/0 V; optional fourth electrode (the
                       reflecting lens of the batched-tracing workload).
* ``mirror``           five nested electrodes (disc, tubes, annuli) with voltages drawn
                       uniformly from [-2, 2] kV — the mixed-precision solve workload.
* ``magnetic_lens``    two iron pole pieces (mu_r = 1000, bore 2 mm, gap 3 mm) around a
                       500 A-turn rectangular coil, plus a three-tube einzel lens.

Tubes are closed rectangles in (r, z) with wall thickness ``WALL``; element counts are
split over the electrodes in proportion to their perimeters.
"""
from __future__ import annotations

import numpy as np

from .mesh import Mesh, path_mesh, rectangle_path_mesh, rectangle_surface_mesh, line_mesh

MM = 1e-3
WALL = 0.25 * MM


def _split_counts(perims, n_total):
    perims = np.asarray(perims, dtype=np.float64)
    c = np.maximum(4, np.round(perims / perims.sum() * n_total).astype(int))
    return c


def einzel_lens(n_elements: int = 3000, voltages=(0.0, 1800.0, 0.0), fourth: tuple | None = None):
    """Returns (mesh, {name: voltage}).  fourth = (z_start, length, voltage) adds a tube."""
    R = 1.0 * MM
    tubes = [("t1", -3.0 * MM, -1.0 * MM, voltages[0]), ("t2", -0.5 * MM, 0.5 * MM, voltages[1]),
             ("t3", 1.0 * MM, 3.0 * MM, voltages[2])]
    if fourth is not None:
        z0, L, V = fourth
        tubes.append(("t4", z0, z0 + L, V))
    perims = [2 * (z1 - z0) + 2 * WALL for _, z0, z1, _ in tubes]
    counts = _split_counts(perims, n_elements)
    mesh = Mesh()
    volts = {}
    for (name, z0, z1, V), n in zip(tubes, counts):
        mesh = mesh + rectangle_path_mesh(R, R + WALL, z0, z1, name, n_elements=int(n))
        volts[name] = float(V)
    return mesh, volts


def mirror(n_elements: int = 20000, seed: int = 1):
    """Five nested electrodes; voltages U[-2000, 2000] V from the seed."""
    rng = np.random.default_rng(seed)
    V = rng.uniform(-2000.0, 2000.0, 5)
    shapes = [
        # solid disc: its outline without the segment on the axis (a ring of radius 0 carries no charge)
        ("disc", [[0.0, 0.0, -0.5 * MM], [3.0 * MM, 0.0, -0.5 * MM], [3.0 * MM, 0.0, -0.25 * MM], [0.0, 0.0, -0.25 * MM]], False),
        ("tube_in", [[3.5 * MM, 0.0, 0.0], [3.75 * MM, 0.0, 0.0], [3.75 * MM, 0.0, 5.0 * MM], [3.5 * MM, 0.0, 5.0 * MM]], True),
        ("annulus_in", [[1.0 * MM, 0.0, 5.5 * MM], [3.75 * MM, 0.0, 5.5 * MM], [3.75 * MM, 0.0, 5.75 * MM], [1.0 * MM, 0.0, 5.75 * MM]], True),
        ("tube_out", [[4.5 * MM, 0.0, -1.0 * MM], [4.75 * MM, 0.0, -1.0 * MM], [4.75 * MM, 0.0, 10.0 * MM], [4.5 * MM, 0.0, 10.0 * MM]], True),
        ("annulus_out", [[1.5 * MM, 0.0, 10.5 * MM], [4.75 * MM, 0.0, 10.5 * MM], [4.75 * MM, 0.0, 10.75 * MM],
                         [1.5 * MM, 0.0, 10.75 * MM]], True),
    ]
    perims = [np.sum(np.linalg.norm(np.diff(np.vstack([v, v[:1]]) if cl else np.asarray(v), axis=0), axis=1)) for _, v, cl in shapes]
    counts = _split_counts(perims, n_elements)
    counts[-1] += n_elements - counts.sum()
    mesh = Mesh()
    volts = {}
    for (name, verts, closed), n, v in zip(shapes, counts, V):
        mesh = mesh + path_mesh(verts, name, n_elements=int(n), closed=closed)
        volts[name] = float(v)
    return mesh, volts


def magnetic_lens(n_elements: int = 12000, coil_cells: tuple = (10, 10)):
    """Returns (mesh, voltages, permeabilities, currents)."""
    # einzel part: three tubes far from the magnetic gap
    n_e = n_elements // 2
    n_m = n_elements - n_e
    R = 1.0 * MM
    tubes = [("e1", -12.0 * MM, -10.0 * MM, 0.0), ("e2", -9.5 * MM, -8.5 * MM, 1500.0), ("e3", -8.0 * MM, -6.0 * MM, 0.0)]
    perims = [2 * (z1 - z0) + 2 * WALL for _, z0, z1, _ in tubes]
    mesh = Mesh()
    volts = {}
    for (name, z0, z1, V), n in zip(tubes, _split_counts(perims, n_e)):
        mesh = mesh + rectangle_path_mesh(R, R + WALL, z0, z1, name, n_elements=int(n))
        volts[name] = V
    # pole pieces: bore radius 1 mm (2 mm bore), gap 3 mm between the two pole faces at z = +-1.5 mm
    bore = 1.0 * MM
    pole_top = [[bore, 0.0, 1.5 * MM], [8.0 * MM, 0.0, 1.5 * MM], [8.0 * MM, 0.0, 6.0 * MM], [bore, 0.0, 6.0 * MM]]
    pole_bot = [[bore, 0.0, -6.0 * MM], [8.0 * MM, 0.0, -6.0 * MM], [8.0 * MM, 0.0, -1.5 * MM], [bore, 0.0, -1.5 * MM]]
    per = [np.sum(np.linalg.norm(np.diff(np.vstack([v, v[:1]]), axis=0), axis=1)) for v in (pole_top, pole_bot)]
    c = _split_counts(per, n_m)
    mesh = mesh + path_mesh(pole_top, "pole_top", n_elements=int(c[0]), closed=True)
    mesh = mesh + path_mesh(pole_bot, "pole_bot", n_elements=int(c[1]), closed=True)
    # coil between the poles: r in [3, 6] mm, z in [-1.2, 1.2] mm, 500 A-turn
    mesh = mesh + rectangle_surface_mesh(3.0 * MM, 6.0 * MM, -1.2 * MM, 1.2 * MM, "coil", *coil_cells)
    return mesh, volts, {"pole_top": 1000.0, "pole_bot": 1000.0}, {"coil": 500.0}


def small_geometry(n_elements: int = 300, with_flux: bool = True):
    """A compact test geometry: two tubes at +-100 V and (optionally) a dielectric ring with K = 3."""
    mesh = rectangle_path_mesh(1.0 * MM, 1.25 * MM, -2.0 * MM, -0.25 * MM, "a", n_elements=n_elements // 3)
    mesh = mesh + rectangle_path_mesh(1.0 * MM, 1.25 * MM, 0.25 * MM, 2.0 * MM, "b", n_elements=n_elements // 3)
    volts = {"a": -100.0, "b": 100.0}
    diel = {}
    if with_flux:
        mesh = mesh + rectangle_path_mesh(1.5 * MM, 1.75 * MM, -0.5 * MM, 0.5 * MM, "d", n_elements=n_elements - 2 * (n_elements // 3))
        diel["d"] = 3.0
    return mesh, volts, diel


def lines_types_values(mesh: Mesh, volts: dict, flux: dict | None = None, flux_type: int = 3, pot_type: int = 1):
    """Flatten to (lines (N,4,3), types (N,), values (N,)) in mesh line order."""
    flux = flux or {}
    N = len(mesh.lines)
    types = np.zeros(N, dtype=np.uint8)
    values = np.zeros(N)
    active = np.zeros(N, dtype=bool)
    for name, v in volts.items():
        idx = mesh.physical_to_lines[name]
        types[idx] = pot_type
        values[idx] = v
        active[idx] = True
    for name, K in flux.items():
        idx = mesh.physical_to_lines[name]
        types[idx] = flux_type
        values[idx] = K
        active[idx] = True
    lines = mesh.points[mesh.lines.astype(np.int64)]
    return lines[active], types[active], values[active]


# ---------------------------------------------------------------------------------------
# Beams of the tracing workloads: (positions (n,3), velocities (n,3) m/s), tracer bounds
# (3,2) and landing planes (point, normal).  Electrons; seeded, so every consumer draws the
# same particles.
# ---------------------------------------------------------------------------------------
E_CHARGE = 1.602176634e-19
M_ELECTRON = 9.1093837015e-31


def _speed(eV):
    return np.sqrt(2.0 * np.asarray(eV, dtype=np.float64) * E_CHARGE / M_ELECTRON)


def _directions(rng, n, sigma):
    tx, ty = rng.normal(0.0, sigma, n), rng.normal(0.0, sigma, n)
    d = np.column_stack([np.tan(tx), np.tan(ty), np.ones(n)])
    return d / np.linalg.norm(d, axis=1, keepdims=True)


def einzel_beam(n: int = 100, seed: int = 0):
    """Plain einzel workload: electrons at 1 keV from z = -5 mm with radial offset uniform in
    [0, 0.1] mm (along x) and velocity along +z, traced to z = +10 mm; landing on z = +10 mm."""
    rng = np.random.default_rng(seed)
    r0 = rng.uniform(0.0, 0.1 * MM, n)
    P = np.column_stack([r0, np.zeros(n), np.full(n, -5.0 * MM)])
    V = np.zeros((n, 3))
    V[:, 2] = _speed(1000.0)
    bounds = np.array([[-1.0 * MM, 1.0 * MM], [-1.0 * MM, 1.0 * MM], [-5.0 * MM, 10.0 * MM]])
    plane = (np.array([0.0, 0.0, 10.0 * MM]), np.array([0.0, 0.0, 1.0]))
    return P, V, bounds, plane


# The batched-tracing lens: the einzel lens plus a -1.1 kV tube of 0.08 mm at z = 4 mm, whose
# saddle potential reflects about a fifth (22%) of the lognormal 1 keV beam below.
REFLECTOR = (4.0 * MM, 0.08 * MM, -1100.0)


def reflecting_lens(n_elements: int = 8000):
    return einzel_lens(n_elements, voltages=(0.0, 1800.0, 0.0), fourth=REFLECTOR)


def reflecting_beam(n: int, seed: int = 2):
    """Batched-tracing workload: kinetic energy 1000 eV * exp(0.5 N(0,1)), start points uniform
    in a disc of 0.2 mm diameter at z = -5 mm, angles N(0, 5 mrad).  Transmitted electrons land
    on z = +9.5 mm, reflected ones on their way back through z = -4.5 mm (the last crossing of
    that plane, as plane_intersection scans backwards, tracing.c:184-223)."""
    rng = np.random.default_rng(seed)
    energy = 1000.0 * np.exp(0.5 * rng.standard_normal(n))
    rad = 0.1 * MM * np.sqrt(rng.random(n))
    phi = 2 * np.pi * rng.random(n)
    d = _directions(rng, n, 5e-3)
    P = np.column_stack([rad * np.cos(phi), rad * np.sin(phi), np.full(n, -5.0 * MM)])
    V = d * _speed(energy)[:, None]
    bounds = np.array([[-1.0 * MM, 1.0 * MM], [-1.0 * MM, 1.0 * MM], [-5.0 * MM, 10.0 * MM]])
    exit_plane = (np.array([0.0, 0.0, 9.5 * MM]), np.array([0.0, 0.0, 1.0]))
    back_plane = (np.array([0.0, 0.0, -4.5 * MM]), np.array([0.0, 0.0, 1.0]))
    return P, V, bounds, exit_plane, back_plane


def magnetic_beam(n: int, seed: int = 3, r_range=(0.0, 0.1 * MM)):
    """Magnetic + electrostatic workload: 10 keV electrons from z = -14 mm, start radius with
    r^2 uniform over r_range (a disc for r_range = (0, 0.1 mm)), angles N(0, 5 mrad); landing on
    z = +11 mm.  r_range = (0.3 mm, 0.5 mm) is the off-axis beam of the learned-surrogate
    workload, where the paraxial series loses accuracy."""
    rng = np.random.default_rng(seed)
    lo, hi = r_range
    rad = np.sqrt(lo * lo + (hi * hi - lo * lo) * rng.random(n))
    phi = 2 * np.pi * rng.random(n)
    d = _directions(rng, n, 5e-3)
    P = np.column_stack([rad * np.cos(phi), rad * np.sin(phi), np.full(n, -14.0 * MM)])
    V = d * _speed(10000.0)
    bounds = np.array([[-1.0 * MM, 1.0 * MM], [-1.0 * MM, 1.0 * MM], [-14.0 * MM, 12.0 * MM]])
    plane = (np.array([0.0, 0.0, 11.0 * MM]), np.array([0.0, 0.0, 1.0]))
    return P, V, bounds, plane
