"""Beam focus (focus.py:9-43).  CPU: the oracle's restatement (numpy.linalg.lstsq of the
reference's system) against the reference's own tests/test_focus.py cases and against rays
built through a known point.  GPU: tb_focus_position (closed form, device reduction) against
the oracle, to 1e-13 of the rays' extent on the reference's cases and 1e-10 on the beams, on those cases and on 10^6 rays."""
import numpy as np
import pytest

from oracle.oracle import focus_position as oracle_focus


def reference_cases():
    from traceon_amd import tracing as T
    v1 = T.velocity_vec(10, [-1e-3, 0.0, -1])
    v2 = T.velocity_vec(10, [1e-3, 0.0, -1])
    p1 = np.concatenate((3 * v1, v1))[None, :]
    p2 = np.concatenate((3 * v2, v2))[None, :]
    cases = [([p1.copy(), p2.copy()], [0.0, 0.0, 0.0])]
    p1[0, 0] += 1
    p2[0, 0] += 1
    cases.append(([p1.copy(), p2.copy()], [1.0, 0.0, 0.0]))
    for p in (p1, p2):
        p[0, 1] += 1
        p[0, 2] += 1
    cases.append(([p1.copy(), p2.copy()], [1.0, 1.0, 1.0]))
    w = [T.velocity_vec_spherical(1, 0, 0), T.velocity_vec_spherical(5, 1 / 30, 1 / 30),
         T.velocity_vec_spherical(10, 1 / 30, np.pi / 2)]
    cases.append(([np.concatenate((v, v))[None, :] for v in w], [0.0, 0.0, 0.0]))
    return cases


def rays_through(point, n, seed):
    """Final states on rays through `point`: angles of a few mrad, at distances 1-5 mm past it."""
    rng = np.random.default_rng(seed)
    a, b = rng.normal(0, 3e-3, n), rng.normal(0, 3e-3, n)
    t = rng.uniform(1e-3, 5e-3, n)
    vz = rng.uniform(1e7, 2e7, n)
    S = np.empty((n, 6))
    S[:, 0], S[:, 1], S[:, 2] = point[0] + a * t, point[1] + b * t, point[2] + t
    S[:, 3], S[:, 4], S[:, 5] = a * vz, b * vz, vz
    return S


def test_oracle_focus_reference_cases():
    for traj, want in reference_cases():
        assert np.allclose(oracle_focus(traj), want, atol=1e-9)


def test_oracle_focus_of_rays_through_a_point():
    point = np.array([1.2e-4, -3.0e-5, 2.5e-3])
    S = rays_through(point, 2000, 1)
    got = oracle_focus([s[None, :] for s in S])
    assert np.max(np.abs(got - point)) < 1e-13


@pytest.mark.gpu
def test_device_focus_vs_oracle():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from traceon_amd import focus
    for traj, want in reference_cases():
        g = focus.focus_position(traj)
        o = oracle_focus(traj)
        scale = max(np.max(np.abs(np.array([t[-1, :3] for t in traj]))), np.max(np.abs(o)))   # the rays' extent
        print("case", want, g, o, flush=True)
        assert np.allclose(g, o, rtol=0, atol=1e-13 * scale)
    point = np.array([1.2e-4, -3.0e-5, 2.5e-3])
    S = rays_through(point, 1_000_000, 2)
    g = focus.focus_position_batch(torch.tensor(S, device="cuda"))
    extent = 5e-3
    print("1e6 rays through a point:", g - point, flush=True)
    assert np.max(np.abs(g - point)) < 1e-10 * extent
    # a beam without a common point: against numpy.linalg.lstsq of the reference's system
    S[:, :2] += np.random.default_rng(3).normal(0, 1e-7, (len(S), 2))
    o = oracle_focus([s[None, :] for s in S[:20000]])
    g = focus.focus_position_batch(torch.tensor(S[:20000], device="cuda"))
    assert np.max(np.abs(g - o)) < 1e-10 * extent
    # strided rows: the landing columns of a wider array
    W = torch.zeros((20000, 8), dtype=torch.float64, device="cuda")
    W[:, :6] = torch.tensor(S[:20000], device="cuda")
    assert np.array_equal(focus.focus_position_batch(W), g)
