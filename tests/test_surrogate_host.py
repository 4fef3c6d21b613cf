"""Host-side pieces of the learned surrogate that run without a GPU."""
import numpy as np
import torch


def test_bin_sums_match_bincount():
    """The per-bin sums behind the surrogate's z scales: equal to numpy's bincount with weights
    (empty bins included), and the same bits when repeated."""
    from traceon_amd.surrogate import _bin_sums
    rng = np.random.default_rng(3)
    b = rng.integers(1, 12, 5000)          # bins 0 and 12..15 stay empty
    v = rng.normal(size=(5000, 5)) * np.exp(rng.normal(size=(5000, 1)) * 5)
    got = _bin_sums(torch.tensor(b), torch.tensor(v), 16).numpy()
    ref = np.stack([np.bincount(b, weights=v[:, k], minlength=16) for k in range(5)], axis=1)
    assert got.shape == (16, 5)
    assert np.all(got[0] == 0) and np.all(got[12:] == 0)
    assert np.allclose(got, ref, rtol=1e-12, atol=1e-12 * np.abs(v).sum())
    again = _bin_sums(torch.tensor(b), torch.tensor(v), 16).numpy()
    assert np.array_equal(got, again)
