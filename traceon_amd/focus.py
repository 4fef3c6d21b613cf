"""Beam focus -- counterpart of traceon/focus.py.

``focus_position(positions)`` keeps the reference signature (focus.py:9-43): a list of
trajectories (T_i, 6), of which the last state of each is used.  ``focus_position_batch``
takes the final states of a batched trace as they lie on the device ((n, 6), e.g.
``BatchResult.final`` or ``.landing``) and never copies them to the host.  Both run
tb_focus_position: the least-squares point of the rays extended backwards, in the closed form
of the reference's numpy.linalg.lstsq system, reduced on the GPU in a fixed order.
"""
from __future__ import annotations

import numpy as np

from . import backend


def focus_position_batch(states) -> np.ndarray:
    """(x, y, z) of the focus of the rays given by device states (n, 6) (a torch tensor with
    unit stride along a row, any row stride)."""
    import torch
    from . import device
    S = states if isinstance(states, torch.Tensor) else device.to_dev(np.asarray(states, dtype=np.float64))
    assert S.dtype == torch.float64 and S.dim() == 2 and S.shape[1] >= 6 and S.stride(1) == 1
    L = backend.backend_lib
    ws = torch.empty(L.tb_focus_workspace_bytes(), dtype=torch.uint8, device=S.device)
    out = torch.empty(3, dtype=torch.float64, device=S.device)
    backend.check(L.tb_focus_position(S.data_ptr(), S.shape[0], S.stride(0), out.data_ptr(), ws.data_ptr(),
                                      device.stream_ptr()), "focus_position")
    return out.cpu().numpy()


def focus_position(positions) -> np.ndarray:
    """focus.py:9-43: the focus of the trajectories `positions` (iterable of (T, 6) arrays)."""
    assert all(p.shape == (len(p), 6) for p in positions)
    last = np.array([p[-1] for p in positions], dtype=np.float64)
    return focus_position_batch(last)
