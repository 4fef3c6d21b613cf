"""The device beam sampler's generator and transforms, restated in numpy (no reference
counterpart: the reference takes start states from its caller).  SplitMix64 is pinned by the
published known-answer sequence of the generator seeded with 1234567; the GPU tests compare
tb_sample_beam with beam_states() below and check the drawn distributions."""
import numpy as np

M64 = (1 << 64) - 1
GOLDEN = 0x9E3779B97F4A7C15


def splitmix64_at(seed: int, j: int) -> int:
    """Output j (j = 1, 2, ...) of SplitMix64 seeded with `seed` (Steele, Lea and Flood 2014)."""
    z = (seed + j * GOLDEN) & M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def splitmix64_vec(seed: int, j: np.ndarray) -> np.ndarray:
    """splitmix64_at over an array of indices (uint64 arithmetic wraps modulo 2^64)."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + j.astype(np.uint64) * np.uint64(GOLDEN)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def beam_states(n, seed, z0, r_max, e_median, e_sigma=0.0, angle_sigma=0.0, mass=9.1093837015e-31,
                charge=1.602176634e-19):
    """(n, 6) start states as tb_sample_beam draws them: draw k of particle i is output
    8 i + k + 1; u = top 53 bits / 2^53; r = r_max sqrt(u0), phi = 2 pi u1; Box-Muller pairs
    (u2, u3) for the two angles and (u4, u5) for the energy's normal deviate."""
    i = np.arange(n, dtype=np.uint64)

    def u(k):
        return (splitmix64_vec(seed, np.uint64(8) * i + np.uint64(k + 1)) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53

    two_pi = 2 * np.pi
    r = r_max * np.sqrt(u(0))
    phi = two_pi * u(1)
    m1, a1 = np.sqrt(-2 * np.log(1 - u(2))), two_pi * u(3)
    m2, a2 = np.sqrt(-2 * np.log(1 - u(4))), two_pi * u(5)
    tx, ty = np.tan(angle_sigma * m1 * np.cos(a1)), np.tan(angle_sigma * m1 * np.sin(a1))
    energy = e_median * np.exp(e_sigma * m2 * np.cos(a2))
    speed = np.sqrt(2 * energy * charge / mass)
    inv = speed / np.sqrt(tx * tx + ty * ty + 1)
    return np.column_stack([r * np.cos(phi), r * np.sin(phi), np.full(n, z0), tx * inv, ty * inv, inv])


def test_splitmix64_known_answers():
    # the generator's published test sequence for seed 1234567
    want = [6457827717110365317, 3203168211198807973, 9817491932198370423, 4593380528125082431,
            16408922859458223821]
    assert [splitmix64_at(1234567, j) for j in range(1, 6)] == want
    assert [int(x) for x in splitmix64_vec(1234567, np.arange(1, 6))] == want


def test_beam_states_shapes_and_ranges():
    S = beam_states(20000, 7, z0=-5e-3, r_max=1e-4, e_median=1000.0, e_sigma=0.5, angle_sigma=5e-3)
    assert S.shape == (20000, 6) and np.all(np.isfinite(S))
    assert np.all(np.hypot(S[:, 0], S[:, 1]) <= 1e-4) and np.all(S[:, 2] == -5e-3) and np.all(S[:, 5] > 0)
    # r^2 / r_max^2 uniform, log energy N(log 1000, 0.5), slopes N(0, 5 mrad)
    r2 = np.sort((S[:, 0] ** 2 + S[:, 1] ** 2) / 1e-8)
    assert np.max(np.abs(r2 - (np.arange(1, 20001) - 0.5) / 20000)) < 0.015
    E = 0.5 * 9.1093837015e-31 * np.sum(S[:, 3:] ** 2, axis=1) / 1.602176634e-19
    assert abs(np.mean(np.log(E / 1000.0))) < 0.02 and abs(np.std(np.log(E)) - 0.5) < 0.01
    assert abs(np.std(np.arctan(S[:, 3] / S[:, 5])) - 5e-3) < 1e-4


def test_sample_beam_rejects_bad_arguments_without_a_gpu():
    """tb_sample_beam checks its arguments on the host before any launch; BeamPipeline refuses
    more particles than the no-alloc tracer's static maximum before touching the device."""
    import ctypes as C
    import pytest
    from traceon_amd import backend, tracing
    L = backend.backend_lib
    seed = C.c_ulonglong(1)
    good = tracing.BeamT(-5e-3, 1e-4, 1000.0, 0.5, 5e-3, 9.1093837015e-31, 1.602176634e-19)
    for field, value in (("mass", 0.0), ("charge", -1.0), ("e_median", 0.0), ("r_max", -1.0), ("e_sigma", float("nan"))):
        bad = tracing.BeamT(*[getattr(good, f) for f, _ in tracing.BeamT._fields_])
        setattr(bad, field, value)
        assert L.tb_sample_beam(C.byref(bad), C.byref(seed), 4, C.c_void_p(16), None) == -1
        assert b"tb_sample_beam" in L.tb_last_error()
    assert L.tb_sample_beam(None, C.byref(seed), 4, C.c_void_p(16), None) == -1
    assert L.tb_sample_beam(C.byref(good), C.byref(seed), -1, C.c_void_p(16), None) == -1
    with pytest.raises(ValueError, match="static maximum"):
        tracing.BeamPipeline(None, L.tb_trace_noalloc_max_particles() + 1, {})
