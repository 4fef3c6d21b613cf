"""Radial mesh producer: line paths and rectangles meshed into cubic line elements, and
rectangles meshed into triangles for coils.

The reference builds these through traceon/geometry.py and traceon/mesher.py.  ``Mesh`` keeps
the reference's arrays -- ``points`` (P,3), ``lines`` (N,2) or (N,4) in gmsh line4 order
[start, end, 1/3, 2/3] (mesher.py:545-567), ``triangles`` (T,3) and the ``physical_to_lines`` /
``physical_to_triangles`` name maps -- in the x-z plane (y = 0), and its constructor does what
the reference's does (mesher.py:186-226): it drops degenerate triangles, merges points that
agree once the low 16 bits of each coordinate's mantissa are zeroed (the kept points are the
quantised ones), and orients the lines of every electrode outward, connected path by connected
path.  Straight paths are meshed into two-point lines and raised to cubic elements by adding
the 1/3 and 2/3 points, as the solver does with such meshes (solver.py:356-362, mesher.py:
569-585), so the element vertices are those the reference produces from the same points.

Orientation follows the reference's default (mesher.py:602-742, 760-852): within a connected
path the lines are made to follow each other, then the path is flipped if the sum over its
lines of mid_r * n_r * length is not positive (normal (dz, -dr) / length); boundaries are
flipped inward by ``ensure_inward_normals``.
"""
from __future__ import annotations

import numpy as np

MANTISSA_MASK = np.uint64(0xFFFFFFFFFFFF0000)


def quantise(points) -> np.ndarray:
    """The coordinates with the low 16 bits of their mantissas zeroed (mesher.py:270-271)."""
    q = np.array(points, dtype=np.float64)
    q.view(np.uint64)[...] &= MANTISSA_MASK
    return q


def deduplicate_points(points, lines, triangles):
    """mesher.py:262-294: quantise, sort (z, then y, then x: numpy.lexsort of the columns),
    keep the first point of every run of equal ones, renumber the elements."""
    if not len(points):
        return points, lines, triangles
    q = quantise(points)
    order = np.lexsort(q.T)
    qs = q[order]
    keep = np.ones(len(qs), dtype=bool)
    keep[1:] = np.any(qs[1:] != qs[:-1], axis=1)
    renumber = np.empty(len(q), dtype=np.uint64)
    renumber[order] = np.cumsum(keep) - 1
    lines = renumber[lines] if len(lines) else lines
    triangles = renumber[triangles] if len(triangles) else triangles
    return qs[keep], lines, triangles


def triangle_areas(tri) -> np.ndarray:
    """Areas of triangles (T, 3, 3)."""
    tri = np.asarray(tri, dtype=np.float64)
    return 0.5 * np.linalg.norm(np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0]), axis=1)


def _flip(line):
    return line[[1, 0, 3, 2]] if len(line) == 4 else line[[1, 0]]


def _connected_paths(lines) -> list:
    """Index sets of the lines that share vertices, in the order the reference labels them
    (mesher.py:678-712: first unlabelled element starts a component)."""
    touching = {}
    for i, el in enumerate(lines):
        for v in el:
            touching.setdefault(int(v), []).append(i)
    label = np.full(len(lines), -1)
    comps = []
    for i in range(len(lines)):
        if label[i] >= 0:
            continue
        c = len(comps)
        label[i] = c
        stack = [i]
        while stack:
            e = stack.pop()
            for v in lines[e]:
                for n in touching[int(v)]:
                    if label[n] < 0:
                        label[n] = c
                        stack.append(n)
        comps.append(np.where(label == c)[0])
    return comps


def _make_consecutive(lines) -> None:
    """Flip lines in place so that neighbours sharing an end point follow each other,
    spreading from line 0 (mesher.py:808-852: a neighbour is kept when one line ends where the
    other starts)."""
    ends = {}
    for i, el in enumerate(lines[:, :2]):
        for v in el:
            ends.setdefault(int(v), []).append(i)
    done = np.zeros(len(lines), dtype=bool)   # line 0 is not marked: it meets itself first
    stack = [0]
    while stack:
        t = stack.pop()
        neighbours = [n for v in lines[t].tolist() for n in ends.get(int(v), ())]
        for n in neighbours:
            if done[n]:
                continue
            a0, a1 = lines[t, :2]
            b0, b1 = lines[n, :2]
            if not (a1 == b0 or b1 == a0):
                lines[n] = _flip(lines[n])
            done[n] = True
            stack.append(n)


def _outward(lines, points) -> bool:
    """mesher.py:733-742: sum of mid_r * n_r * length > 0, with n = (dz, -dr) / length."""
    v = points[lines[:, :2].astype(np.int64)]
    d = v[:, 1] - v[:, 0]
    length = np.linalg.norm(d, axis=1)
    nr = d[:, 2] / length
    mid = (v[:, 0, 0] + v[:, 1, 0]) / 2.0
    return float(np.sum(mid * nr * length)) > 0.0


def _concat(a, b):
    """mesher.py:170-178: an empty side gives the other unchanged."""
    if not len(a):
        return b.copy()
    if not len(b):
        return a.copy()
    assert a.shape[1:] == b.shape[1:], "cannot add a higher-order mesh to a simple one"
    return np.concatenate([a, b])


def _higher_order_arrays(mesh):
    """Constructor arguments of mesh with its two-point lines raised to cubic elements."""
    v0, v1 = mesh.lines[:, 0].astype(np.int64), mesh.lines[:, 1].astype(np.int64)
    P = mesh.points
    p2 = P[v0] + (P[v1] - P[v0]) * 1 / 3
    p3 = P[v0] + (P[v1] - P[v0]) * 2 / 3
    n, m = len(P), len(mesh.lines)
    lines = np.stack([mesh.lines[:, 0], mesh.lines[:, 1], np.arange(n, n + m, dtype=np.uint64),
                      np.arange(n + m, n + 2 * m, dtype=np.uint64)], axis=1)
    return np.concatenate([P, p2, p3]), lines, mesh.physical_to_lines, mesh.triangles, mesh.physical_to_triangles


class Mesh:
    def __init__(self, points=None, lines=None, physical_to_lines=None, triangles=None, physical_to_triangles=None,
                 ensure_outward_normals: bool = True):
        self.points = np.zeros((0, 3)) if points is None or len(points) == 0 else np.array(points, dtype=np.float64)
        self.lines = np.zeros((0, 2), dtype=np.uint64) if lines is None or len(lines) == 0 else np.array(lines, dtype=np.uint64)
        self.triangles = (np.zeros((0, 3), dtype=np.uint64) if triangles is None or len(triangles) == 0
                          else np.array(triangles, dtype=np.uint64))
        self.physical_to_lines = {k: np.asarray(v, dtype=np.int64) for k, v in (physical_to_lines or {}).items()}
        self.physical_to_triangles = {k: np.asarray(v, dtype=np.int64) for k, v in (physical_to_triangles or {}).items()}
        self._remove_degenerate_triangles()
        self.points, self.lines, self.triangles = deduplicate_points(self.points, self.lines, self.triangles)
        if ensure_outward_normals:
            for name in self.get_electrodes():
                self.ensure_outward_normals(name)
        assert np.all(self.lines < len(self.points)) and np.all(self.triangles < len(self.points))
        assert not len(self.lines) or self.lines.shape[1] in (2, 4), "lines hold 2 or 4 points"

    def _remove_degenerate_triangles(self) -> None:
        """mesher.py:246-260: triangles of area below 1e-12 are dropped, the groups renumbered."""
        if not len(self.triangles):
            return
        bad = triangle_areas(self.points[self.triangles[:, :3].astype(np.int64)]) < 1e-12
        if not bad.any():
            return
        renumber = np.arange(len(self.triangles)) - np.cumsum(bad)
        self.triangles = self.triangles[~bad]
        for k, v in self.physical_to_triangles.items():
            self.physical_to_triangles[k] = renumber[v[~bad[v]]]

    # mesher.py:229-235
    def is_higher_order(self) -> bool:
        return len(self.lines) > 0 and self.lines.shape[1] == 4

    def is_2d(self) -> bool:
        return True

    def is_3d(self) -> bool:
        return False

    def get_electrodes(self):
        return list(self.physical_to_lines.keys()) + list(self.physical_to_triangles.keys())

    def __add__(self, other: "Mesh") -> "Mesh":
        """Concatenate two meshes (mesher.py:308-337): a new Mesh, so shared points merge."""
        off = np.uint64(len(self.points))
        lines = _concat(self.lines, other.lines + off)
        tris = _concat(self.triangles, other.triangles + off)
        p2l = {k: v.copy() for k, v in self.physical_to_lines.items()}
        for k, v in other.physical_to_lines.items():
            v = v + len(self.lines)
            p2l[k] = np.concatenate([p2l[k], v]) if k in p2l else v
        p2t = {k: v.copy() for k, v in self.physical_to_triangles.items()}
        for k, v in other.physical_to_triangles.items():
            v = v + len(self.triangles)
            p2t[k] = np.concatenate([p2t[k], v]) if k in p2t else v
        return Mesh(np.concatenate([self.points, other.points]), lines, p2l, tris, p2t)

    def to_higher_order(self) -> "Mesh":
        """Two-point lines raised to cubic elements (mesher.py:545-585): the points
        p0 + (p1 - p0) * 1 / 3 and p0 + (p1 - p0) * 2 / 3 are appended and the lines become
        [p0, p1, 1/3, 2/3]; a new Mesh (quantised, merged, oriented) of the result."""
        if not len(self.lines) or self.lines.shape[1] == 4:
            return self
        return Mesh(*_higher_order_arrays(self))

    def line_vertices(self, name: str | None = None) -> np.ndarray:
        idx = self.physical_to_lines[name] if name is not None else np.arange(len(self.lines))
        return self.points[self.lines[idx].astype(np.int64)]

    def _orient(self, name: str, outward: bool) -> None:
        """mesher.py:619-636, 844-852: every connected path of the electrode is made
        consecutive, then flipped if its normals point the wrong way."""
        idx = self.physical_to_lines[name]
        el = self.lines[idx]
        if not len(el):
            return
        for comp in _connected_paths(el):
            sub = el[comp]
            _make_consecutive(sub)
            if _outward(sub, self.points) != outward:
                sub = np.array([_flip(l) for l in sub], dtype=np.uint64)
            el[comp] = sub
        self.lines[idx] = el

    def _orient_triangles(self, name: str) -> None:
        """mesher.py:602-617, 760-795 for triangles in the x-z plane (the coil sections of the
        radial path): within each connected set the triangles take the winding of its first
        one (for coplanar edge neighbours triangle_orientation_is_equal compares the signs of
        the y normals, triangle.c:6-71); the outward test (mesher.py:715-731) then sums
        mid_x * n_x * area = 0 for such triangles, is false, and flips the set -- after the
        first triangle was flipped against itself, so it ends as given."""
        idx = self.physical_to_triangles[name]
        tr = self.triangles[idx]
        if not len(tr):
            return
        v = self.points[tr.astype(np.int64)]
        ny = np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0])[:, 1]
        for comp in _connected_paths(tr):
            wrong = np.sign(ny[comp]) != np.sign(ny[comp[0]])
            flip = comp[wrong]
            tr[flip] = tr[flip][:, [0, 2, 1]]
        self.triangles[idx] = tr

    def ensure_outward_normals(self, name: str) -> None:
        if name in self.physical_to_triangles:
            self._orient_triangles(name)
        if name in self.physical_to_lines:
            self._orient(name, True)

    def ensure_inward_normals(self, name: str) -> None:
        if name in self.physical_to_triangles:
            self._orient_triangles(name)
        if name in self.physical_to_lines:
            self._orient(name, False)

    def _normals_outward(self, name: str) -> bool:
        """Whether every connected path of the electrode has outward normals."""
        el = self.lines[self.physical_to_lines[name]]
        return all(_outward(el[c], self.points) for c in _connected_paths(el))


def _segment_points(p0, p1, n):
    t = np.linspace(0.0, 1.0, n + 1)
    return np.asarray(p0)[None, :] + t[:, None] * (np.asarray(p1) - np.asarray(p0))[None, :]


def path_mesh(vertices, name: str, n_elements: int | None = None, mesh_size: float | None = None, closed: bool = False,
              ensure_outward_normals: bool = True) -> Mesh:
    """Mesh a polyline (x, 0, z) vertices into cubic line elements.  The element count of each
    straight segment is proportional to its length (at least one).  The two-point lines go
    through Mesh (quantised, merged, oriented outward unless asked otherwise, mesher.py:
    186-226) and are then raised to cubic elements (Mesh.to_higher_order)."""
    V = np.asarray(vertices, dtype=np.float64)
    if closed:
        V = np.concatenate([V, V[:1]])
    seg = np.linalg.norm(np.diff(V, axis=0), axis=1)
    total = seg.sum()
    if n_elements is not None:
        counts = np.maximum(1, np.round(seg / total * n_elements).astype(int))
        counts[np.argmax(seg)] += n_elements - counts.sum()   # exact total
    else:
        counts = np.maximum(1, np.ceil(seg / mesh_size).astype(int))
    pts = [V[0][None, :]]
    for a, b, n in zip(V[:-1], V[1:], counts):
        pts.append(_segment_points(a, b, n)[1:])
    P = np.concatenate(pts)
    if closed:
        P = P[:-1]
    nseg = len(P) if closed else len(P) - 1
    starts = np.arange(nseg)
    lines = np.stack([starts, (starts + 1) % len(P)], axis=1)
    mesh = Mesh(P, lines, {name: np.arange(nseg)}, ensure_outward_normals=ensure_outward_normals)
    if not ensure_outward_normals:
        return Mesh(*_higher_order_arrays(mesh), ensure_outward_normals=False)
    return mesh.to_higher_order()


def line_mesh(p0, p1, name, n_elements=None, mesh_size=None) -> Mesh:
    return path_mesh([p0, p1], name, n_elements=n_elements, mesh_size=mesh_size)


def rectangle_path_mesh(rmin, rmax, zmin, zmax, name, n_elements=None, mesh_size=None) -> Mesh:
    """Closed rectangle in (r, z), counter-clockwise (outward normals)."""
    V = [[rmin, 0.0, zmin], [rmax, 0.0, zmin], [rmax, 0.0, zmax], [rmin, 0.0, zmax]]
    return path_mesh(V, name, n_elements=n_elements, mesh_size=mesh_size, closed=True)


def rectangle_surface_mesh(rmin, rmax, zmin, zmax, name, nr, nz) -> Mesh:
    """Rectangle in the x-z plane split into 2*nr*nz triangles (coil cross sections)."""
    r = np.linspace(rmin, rmax, nr + 1)
    z = np.linspace(zmin, zmax, nz + 1)
    R, Z = np.meshgrid(r, z, indexing="ij")
    points = np.stack([R.ravel(), np.zeros(R.size), Z.ravel()], axis=1)
    idx = np.arange(R.size).reshape(nr + 1, nz + 1)
    a, b, c, d = idx[:-1, :-1].ravel(), idx[1:, :-1].ravel(), idx[1:, 1:].ravel(), idx[:-1, 1:].ravel()
    tris = np.concatenate([np.stack([a, b, c], 1), np.stack([a, c, d], 1)])
    return Mesh(points, triangles=tris, physical_to_triangles={name: np.arange(len(tris))})


def circle_mesh(x0: float, z0: float, radius: float, name: str, n_elements: int) -> Mesh:
    """Closed circle around (x0, z0) in the xz plane, counter-clockwise from the positive x
    direction as Path.circle_xz (geometry.py:293-313), so its normals point outwards.
    Curved cubic elements: all four nodes of every element lie on the circle (the
    reference meshes such paths with higher_order=True)."""
    n = int(n_elements)
    th = 2 * np.pi * np.arange(n) / n
    dth = 2 * np.pi / n

    def at(t):
        return np.stack([x0 + radius * np.cos(t), np.zeros_like(t), z0 + radius * np.sin(t)], axis=1)

    points = np.concatenate([at(th), at(th + dth / 3.0), at(th + 2.0 * dth / 3.0)])
    starts = np.arange(n)
    lines = np.stack([starts, (starts + 1) % n, n + starts, 2 * n + starts], axis=1)
    mesh = Mesh(points, lines, {name: np.arange(n)})
    mesh.ensure_outward_normals(name)
    return mesh


def disk_surface_mesh(x0: float, z0: float, radius: float, name: str, n_rings: int, n_sectors: int) -> Mesh:
    """A disk in the x-z plane (centre (x0, z0), as Surface.disk_xz, geometry.py:1164-1180)
    split into triangles: a fan of n_sectors around the centre and n_rings - 1 rings of
    quadrilaterals (two triangles each) between concentric circles of n_sectors points.
    Every boundary point lies on the circle (coil cross sections)."""
    th = 2 * np.pi * np.arange(n_sectors) / n_sectors
    rad = radius * np.arange(1, n_rings + 1) / n_rings
    ring = np.stack([x0 + rad[:, None] * np.cos(th)[None, :], np.zeros((n_rings, n_sectors)),
                     z0 + rad[:, None] * np.sin(th)[None, :]], axis=2).reshape(-1, 3)
    points = np.concatenate([[[x0, 0.0, z0]], ring])
    idx = 1 + np.arange(n_rings * n_sectors).reshape(n_rings, n_sectors)
    nxt = np.roll(idx, -1, axis=1)
    tris = [np.stack([np.zeros(n_sectors, dtype=np.int64), idx[0], nxt[0]], axis=1)]
    for k in range(n_rings - 1):
        a, b, c, d = idx[k], idx[k + 1], nxt[k + 1], nxt[k]
        tris += [np.stack([a, b, c], axis=1), np.stack([a, c, d], axis=1)]
    tris = np.concatenate(tris)
    return Mesh(points, triangles=tris, physical_to_triangles={name: np.arange(len(tris))})
