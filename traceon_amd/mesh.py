"""Minimal radial mesh producer: straight line paths and rectangles meshed into cubic
line elements, and rectangles meshed into triangles for coils.

The reference builds these through traceon/geometry.py and traceon/mesher.py (parametric
paths, gmsh-free meshing, deduplication, outward normals).  Only what the radial BEM path
consumes is produced here: a ``Mesh`` with the reference's arrays —
``points`` (P,3), ``lines`` (N,4) in gmsh line4 order [start, end, 1/3, 2/3]
(mesher.py:545-567), ``triangles`` (T,3) and the ``physical_to_lines`` /
``physical_to_triangles`` name maps — in the x-z plane (y = 0).

Orientation follows the reference's default (mesher.py:194-220, 733-742): closed paths are
traversed counter-clockwise in (r, z) so that normals (dz, -dr) point outward; boundaries
are flipped inward by ``ensure_inward_normals``.
"""
from __future__ import annotations

import numpy as np


class Mesh:
    def __init__(self, points=None, lines=None, physical_to_lines=None, triangles=None, physical_to_triangles=None):
        self.points = np.zeros((0, 3)) if points is None else np.asarray(points, dtype=np.float64)
        self.lines = np.zeros((0, 4), dtype=np.uint64) if lines is None else np.asarray(lines, dtype=np.uint64)
        self.triangles = np.zeros((0, 3), dtype=np.uint64) if triangles is None else np.asarray(triangles, dtype=np.uint64)
        self.physical_to_lines = {k: np.asarray(v, dtype=np.int64) for k, v in (physical_to_lines or {}).items()}
        self.physical_to_triangles = {k: np.asarray(v, dtype=np.int64) for k, v in (physical_to_triangles or {}).items()}

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
        """Concatenate two meshes (mesher.py:308-337, without point deduplication)."""
        off = len(self.points)
        lines = np.concatenate([self.lines, other.lines + np.uint64(off)]) if len(other.lines) else self.lines
        tris = np.concatenate([self.triangles, other.triangles + np.uint64(off)]) if len(other.triangles) else self.triangles
        p2l = {k: v.copy() for k, v in self.physical_to_lines.items()}
        for k, v in other.physical_to_lines.items():
            v = v + len(self.lines)
            p2l[k] = np.concatenate([p2l[k], v]) if k in p2l else v
        p2t = {k: v.copy() for k, v in self.physical_to_triangles.items()}
        for k, v in other.physical_to_triangles.items():
            v = v + len(self.triangles)
            p2t[k] = np.concatenate([p2t[k], v]) if k in p2t else v
        return Mesh(np.concatenate([self.points, other.points]), lines, p2l, tris, p2t)

    def line_vertices(self, name: str | None = None) -> np.ndarray:
        idx = self.physical_to_lines[name] if name is not None else np.arange(len(self.lines))
        return self.points[self.lines[idx].astype(np.int64)]

    def _flip_lines(self, name: str) -> None:
        idx = self.physical_to_lines[name]
        l = self.lines[idx]
        # reverse direction: start<->end and 1/3<->2/3
        self.lines[idx] = l[:, [1, 0, 3, 2]]

    def _normals_outward(self, name: str) -> bool:
        # mesher.py:733-742: sum of mid_r * n_r * length > 0
        v = self.points[self.lines[self.physical_to_lines[name]][:, :2].astype(np.int64)]
        d = v[:, 1] - v[:, 0]
        length = np.linalg.norm(d, axis=1)
        nr = d[:, 2] / np.where(length > 0, length, 1.0)
        mid = 0.5 * (v[:, 0, 0] + v[:, 1, 0])
        return float(np.sum(mid * nr * length)) > 0.0

    def ensure_outward_normals(self, name: str) -> None:
        if name in self.physical_to_lines and not self._normals_outward(name):
            self._flip_lines(name)

    def ensure_inward_normals(self, name: str) -> None:
        if name in self.physical_to_lines and self._normals_outward(name):
            self._flip_lines(name)


def _segment_points(p0, p1, n):
    t = np.linspace(0.0, 1.0, n + 1)
    return np.asarray(p0)[None, :] + t[:, None] * (np.asarray(p1) - np.asarray(p0))[None, :]


def path_mesh(vertices, name: str, n_elements: int | None = None, mesh_size: float | None = None, closed: bool = False,
              ensure_outward_normals: bool = True) -> Mesh:
    """Mesh a polyline (x, 0, z) vertices into cubic line elements.  The element count of each
    straight segment is proportional to its length (at least one).  As the reference's Mesh
    constructor (mesher.py:217-219), the element direction is flipped where needed so that
    the normals point outward."""
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
    ends = (starts + 1) % len(P)
    p0, p1 = P[starts], P[ends]
    third = p0 + (p1 - p0) / 3.0
    two_third = p0 + (p1 - p0) * 2.0 / 3.0
    points = np.concatenate([P, third, two_third])
    n = len(P)
    lines = np.stack([starts, ends, n + np.arange(nseg), n + nseg + np.arange(nseg)], axis=1)
    mesh = Mesh(points, lines, {name: np.arange(nseg)})
    if ensure_outward_normals:
        mesh.ensure_outward_normals(name)
    return mesh


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
