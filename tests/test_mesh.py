"""The mesh producer (traceon_amd/mesh.py) against the reference's mesher semantics
(mesher.py:170-337, 545-585, 602-852), CPU only: the vectors of the reference's own
tests/test_mesher.py cases as data, and what the mathematics fixes -- the quantum of the
16-bit mantissa truncation, lexicographic order, idempotence, and the outward-normal sum."""
import numpy as np
import pytest

from traceon_amd import mesh as M
from traceon_amd.mesh import Mesh


def test_dedup_no_duplicates():
    pts = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]])
    m = Mesh(pts, triangles=[[0, 1, 2]], ensure_outward_normals=False)
    # lexsort by (z, y, x): (0,0,0), (1,0,0), (0,1,0) is already that order
    assert np.array_equal(m.points, pts) and np.array_equal(m.triangles, [[0, 1, 2]])


def test_dedup_with_duplicates():
    pts = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, 0.0], [0.0, 1.0, 0.0]])
    m = Mesh(pts, triangles=[[0, 1, 3], [2, 1, 3]], ensure_outward_normals=False)
    assert np.array_equal(m.points, [[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]])
    assert np.array_equal(m.triangles, [[0, 1, 2], [0, 1, 2]])


def test_dedup_all_duplicates_removes_the_degenerate_triangle():
    m = Mesh(np.ones((3, 3)), triangles=[[0, 1, 2]])
    assert np.array_equal(m.points, [[1.0, 1.0, 1.0]]) and m.triangles.shape == (0, 3)


def test_dedup_close_points():
    pts = np.array([[0.0, 2.0, 3.0], [0.0, 0.0, 1.0], [0.0, 0.0, 1 + 1e-12], [1.0, 0.0, 0.0]])
    m = Mesh(pts, triangles=[[0, 1, 2]], ensure_outward_normals=False)
    assert np.allclose(m.points, [[1.0, 0.0, 0.0], [0.0, 0.0, 1.0], [0.0, 2.0, 3.0]])
    assert np.array_equal(m.triangles, [[2, 1, 1]])


def test_empty_meshes():
    assert len(Mesh(np.array([]), triangles=np.array([])).points) == 0
    m = Mesh(np.random.default_rng(0).random((10, 3)), triangles=np.array([]))
    assert m.points.shape == (10, 3)


def test_quantum_of_the_mantissa_truncation():
    """Zeroing 16 of 52 mantissa bits moves a value towards zero by less than 2^-36 of it, and
    values that agree in all but those bits merge."""
    rng = np.random.default_rng(3)
    x = rng.uniform(-1e3, 1e3, 10_000)
    q = M.quantise(x)
    assert np.all(np.abs(q) <= np.abs(x)) and np.all(np.abs(x - q) < 2.0 ** -36 * np.abs(x))
    assert np.array_equal(M.quantise(q), q)                      # idempotent
    ulp = np.spacing(1.0)
    assert M.quantise(1.0 + 65535 * ulp) == 1.0 and M.quantise(1.0 + 65536 * ulp) > 1.0
    pts = np.array([[1.0, 0.0, 2.0], [1.0 + 100 * ulp, 0.0, 2.0], [1.0 + 70000 * ulp, 0.0, 2.0]])
    p, lines, _ = M.deduplicate_points(pts, np.array([[0, 2], [1, 2]], dtype=np.uint64), np.zeros((0, 3), np.uint64))
    assert len(p) == 2 and lines[0, 0] == lines[1, 0]


def test_dedup_large_random_mesh_keeps_geometry():
    """reference tests/test_mesher.py test_complex_large_mesh: zeros, exact duplicates and
    points 1e-12 apart; element geometry survives up to the quantum, points come out sorted by
    (z, y, x) and unique."""
    rng = np.random.default_rng(7)
    pts = rng.random((1000, 3))
    pts[:100] = 0.0
    pts[rng.choice(1000, 100, replace=False)] = pts[rng.choice(1000, 100, replace=False)]
    pts[rng.choice(1000, 100, replace=False)] += rng.random((100, 3)) * 1e-12
    tri = rng.permutation(999).reshape(333, 3).astype(np.uint64)
    p, _, t = M.deduplicate_points(pts, np.zeros((0, 2), np.uint64), tri)
    assert np.allclose(pts[tri.astype(int)], p[t.astype(int)], rtol=0, atol=1e-11)
    assert np.array_equal(np.lexsort(p.T), np.arange(len(p)))
    assert len(np.unique(p, axis=0)) == len(p)


def test_connected_paths():
    """reference tests/test_mesher.py TestConnectedElements."""
    cases = [([[1, 2]], [[0]]), ([[1, 2], [2, 3]], [[0, 1]]), ([[1, 2], [3, 4]], [[0], [1]]),
             ([[1, 2], [2, 3], [3, 4], [5, 6]], [[0, 1, 2], [3]]), ([[1, 2, 3], [3, 4, 5], [6, 7, 8]], [[0, 1], [2]])]
    for el, want in cases:
        got = M._connected_paths(np.array(el, dtype=np.uint64))
        assert [list(g) for g in got] == want


def outward_sum(m, idx):
    v = m.points[m.lines[idx][:, :2].astype(int)]
    d = v[:, 1] - v[:, 0]
    return np.sum((v[:, 0, 0] + v[:, 1, 0]) / 2 * d[:, 2])


def test_every_connected_path_is_oriented_outward_and_consecutive():
    """Two closed loops in one electrode, one given clockwise: each loop ends consecutive
    (each line starts where the previous ends) with sum(mid_r n_r length) > 0."""
    a = M.path_mesh([[1, 0, 0], [2, 0, 0], [2, 0, 1], [1, 0, 1]], "e", n_elements=40, closed=True, ensure_outward_normals=False)
    b = M.path_mesh([[3, 0, 0], [3, 0, 1], [4, 0, 1], [4, 0, 0]], "e", n_elements=40, closed=True, ensure_outward_normals=False)
    assert outward_sum(b, b.physical_to_lines["e"]) < 0            # b as given: inward
    m = a + b                                                        # a new Mesh orients outward
    comps = M._connected_paths(m.lines[m.physical_to_lines["e"]])
    assert len(comps) == 2
    for c in comps:
        assert outward_sum(m, c) > 0
        L = m.lines[c][:, :2]
        assert np.all(np.isin(L[:, 1], L[:, 0]))                     # closed and consecutive
    m.ensure_inward_normals("e")
    for c in comps:
        assert outward_sum(m, c) < 0


def test_higher_order_conversion():
    """mesher.py:545-585: p0 + (p1 - p0) * 1 / 3 and * 2 / 3 from the quantised end points,
    quantised in turn; the lines become [p0, p1, 1/3, 2/3]."""
    pts = np.array([[0.1, 0.0, 0.3], [0.7, 0.0, 0.35], [0.9, 0.0, 1.1]])
    m = Mesh(pts, [[0, 1], [1, 2]], {"p": [0, 1]}, ensure_outward_normals=False)
    h = m.to_higher_order()
    assert h.is_higher_order() and len(h.lines) == 2
    v = h.points[h.lines.astype(int)]
    q = M.quantise(pts)
    for k, (i, j) in enumerate(((0, 1), (1, 2))):
        assert np.array_equal(v[k, 0], q[i]) and np.array_equal(v[k, 1], q[j])
        assert np.array_equal(v[k, 2], M.quantise(q[i] + (q[j] - q[i]) * 1 / 3))
        assert np.array_equal(v[k, 3], M.quantise(q[i] + (q[j] - q[i]) * 2 / 3))
    assert h.to_higher_order() is h


def test_solver_raises_two_point_meshes():
    from traceon_amd import solver
    from traceon_amd.excitation import Excitation
    m = Mesh([[1.0, 0, 0], [2.0, 0, 0], [2.0, 0, 1.0]], [[0, 1], [1, 2]], {"a": [0, 1]})
    exc = Excitation(m)
    exc.add_voltage(a=1.0)
    up = solver._excitation_to_higher_order(exc)
    assert up.mesh.is_higher_order() and not exc.mesh.is_higher_order()


def test_coil_triangles_keep_a_consistent_winding():
    m = M.rectangle_surface_mesh(1.0, 2.0, 0.0, 1.0, "coil", 3, 2)
    t = m.triangles.copy()
    t[1] = t[1][[0, 2, 1]]
    m2 = Mesh(m.points, triangles=t, physical_to_triangles={"coil": np.arange(len(t))})
    v = m2.points[m2.triangles.astype(int)]
    ny = np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0])[:, 1]
    assert np.all(np.sign(ny) == np.sign(ny[0]))
    assert np.array_equal(m2.triangles[0], m.triangles[0])
