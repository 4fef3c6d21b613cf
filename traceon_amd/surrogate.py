"""Learned field surrogate: an MLP that corrects the paraxial (axial-derivative) field of
FieldRadialAxial towards the direct BEM field away from the axis.

The paraxial series of field.py:1-15 is truncated after the r^8 term (radial.c:313-384), so
its error grows quickly with the distance from the axis (in the C2 lens: 1e-9 relative at
r < 0.1 mm, 1e-5 at 0.3 mm, 1e-3 at 0.5 mm). The surrogate learns that residual,

    D(r, z) = F_BEM(r, z) - F_axial(r, z),  F = (Phi, E_r, E_z, H_r, H_z),

from samples labelled by the GPU's own BEM sums. The trained network is a
2-128-128-128-128-5 SiLU MLP, with hidden layers in BF16 on MFMA and FP32 master weights
(csrc/mlp.hip). The residual spans many decades: it grows like r^10 off the axis and varies
by orders of magnitude along z between electrode gaps and quiet regions. The targets are
therefore D_k / (w(r) s_k(z)), with w(r) = (r/r_max)^p and s_k(z) the RMS of D_k / w in
z bins, so that they are O(1) everywhere and the fit error scales with the local residual.
Below r_lo the axial series is already accurate, so there is no correction there.

In the tracer the field is F_axial + correction inside the trust region, which is r <= r_max
within the axial interpolation range. The correction applies for r >= r_lo within the z range
it was trained on. Outside the trust region the tracer falls back to the direct BEM sum, and
so it does wherever the surrogate's embedded error estimate exceeds a tolerance: after
training, ``calibrate`` measures the fit error of the scaled targets per z bin on fresh
BEM-labelled samples, so that w(r) sigma_k(z) predicts the error of the correction at (r, z)
relative to the field's scale; the tracer compares that prediction with the tolerance before it
evaluates the network. The surrogate has no counterpart in the reference.
"""
from __future__ import annotations

import math
import time

import numpy as np
import torch

from . import backend
from . import device
from .field import FieldRadialAxial, FieldRadialBEM

NOUT = 5   # Phi, E_r, E_z, H_r, H_z


def _L():
    return backend.backend_lib


def _bin_sums(b: torch.Tensor, v: torch.Tensor, nb: int) -> torch.Tensor:
    """Sum of the rows of v (M, k) per bin b (M,) in [0, nb): (nb, k), the same bits on every
    run (an atomic scatter-add would add in whatever order the hardware takes).  The rows are
    stably sorted by bin and scattered into a dense (nb, largest bin, k) array, one slot
    each, which a plain reduction then sums."""
    order = torch.argsort(b, stable=True)
    bs = b[order]
    cnt = torch.bincount(b, minlength=nb)
    start = torch.cumsum(cnt, 0) - cnt
    slot = torch.arange(b.shape[0], device=b.device) - start[bs]
    dense = torch.zeros((nb, int(cnt.max().item()), v.shape[1]), dtype=v.dtype, device=v.device)
    dense[bs, slot] = v[order]
    return dense.sum(dim=1)


def _silu_he_init(rng, n_param):
    """He-style initialisation in the FP32 parameter layout of csrc/mlp_kernels.h."""
    W = 128
    P = np.zeros(n_param, dtype=np.float32)
    o = 0
    for size, fan_in, is_w in ((W * 2, 2, True), (W, 0, False), (W * W, W, True), (W, 0, False), (W * W, W, True), (W, 0, False),
                               (W * W, W, True), (W, 0, False), (NOUT * W, W, True), (NOUT, 0, False)):
        if is_w:
            P[o:o + size] = rng.standard_normal(size) * math.sqrt(2.0 / fan_in)
        o += size
    assert o == n_param
    return P


class FieldSurrogate:
    """Trains and evaluates the residual MLP for one (FieldRadialBEM, FieldRadialAxial) pair."""

    def __init__(self, field: FieldRadialBEM, axial: FieldRadialAxial, r_max: float, z_min: float | None = None,
                 z_max: float | None = None, r_lo: float | None = None, weight_power: float = 8.0, z_bins: int = 128,
                 seed: int = 0):
        self.field = field
        self.axial = axial
        self.r_max = float(r_max)
        self.r_lo = float(0.25 * r_max if r_lo is None else r_lo)
        self.z_min = float(axial.z[0] if z_min is None else z_min)
        self.z_max = float(axial.z[-1] if z_max is None else z_max)
        self.p = float(weight_power)
        self.nz = int(z_bins)
        self.rng = np.random.default_rng(seed)
        self.seed = int(seed)
        L = _L()
        self.nparam = L.tb_mlp_nparams()
        dev = device.device()
        self.P = torch.tensor(_silu_he_init(self.rng, self.nparam), device=dev)
        self.G = torch.zeros(self.nparam, dtype=torch.float32, device=dev)
        self.M = torch.zeros_like(self.P)
        self.V = torch.zeros_like(self.P)
        self.Wbf = torch.empty(3 * 128 * 128, dtype=torch.int16, device=dev)
        backend.check(L.tb_mlp_to_bf16(device.ptr(self.P), device.ptr(self.Wbf), device.stream_ptr()), "mlp_to_bf16")
        self.zscale = np.ones((NOUT, self.nz))   # s_k(z) at the bin centres
        self.zerr = None                         # sigma_k(z): predicted error per unit weight, relative to the field scale
        self.active = np.ones(NOUT, dtype=bool)  # components with a nonzero residual
        self.step = 0
        self.X = self.T = None
        self.history = []

    # ---- sampling and labels -----------------------------------------------------------
    def _exact_and_axial(self, pts: torch.Tensor):
        """(F_BEM, F_axial) as (M, 5) device tensors at points (r, 0, z)."""
        M = pts.shape[0]
        dev = pts.device
        Fb = torch.zeros((M, NOUT), dtype=torch.float64, device=dev)
        Fa = torch.zeros_like(Fb)
        f = self.field
        if len(f.electrostatic_point_charges):
            P, E = f.electrostatic_point_charges.device_sources().field_points(pts, pot=True, field=True)
            Fb[:, 0], Fb[:, 1], Fb[:, 2] = P, E[:, 0], E[:, 2]
        H = torch.zeros((M, 3), dtype=torch.float64, device=dev)
        if len(f.magnetostatic_point_charges):
            _, Hm = f.magnetostatic_point_charges.device_sources().field_points(pts, pot=False, field=True)
            H += Hm
        if len(f.current_point_charges):
            _, Hc = f.current_point_charges.device_sources().field_points(pts, pot=False, field=True)
            H += Hc
        Fb[:, 3], Fb[:, 4] = H[:, 0], H[:, 2]
        z, ce, cm = self.axial.device_arrays()
        pa, ea = device.axial_points(pts, z, ce, pot=True, field=True)
        _, ha = device.axial_points(pts, z, cm, pot=False, field=True)
        Fa[:, 0], Fa[:, 1], Fa[:, 2], Fa[:, 3], Fa[:, 4] = pa, ea[:, 0], ea[:, 2], ha[:, 0], ha[:, 2]
        return Fb, Fa

    def _weight(self, r):
        return torch.where(r >= self.r_lo, (r / self.r_max) ** self.p, torch.zeros_like(r))

    def _inputs(self, r, z):
        return torch.stack([2.0 * (r - self.r_lo) / (self.r_max - self.r_lo) - 1.0,
                            2.0 * (z - self.z_min) / (self.z_max - self.z_min) - 1.0], dim=1)

    def _ztable(self, table, z):
        """A (5, nz) bin-centre table, piecewise linear between the z-bin centres, at (M,) points: (M, 5)."""
        t = (z - self.z_min) / (self.z_max - self.z_min) * self.nz - 0.5
        t = torch.clamp(t, 0.0, self.nz - 1.0)
        i0 = torch.clamp(torch.floor(t).long(), 0, self.nz - 2) if self.nz > 1 else torch.zeros_like(t, dtype=torch.long)
        f = (t - i0.double())[:, None]
        zs = torch.tensor(table.T, device=z.device)   # (nz, 5)
        return zs[i0] * (1.0 - f) + zs[torch.clamp(i0 + 1, max=self.nz - 1)] * f

    def _zscale(self, z):
        """s_k(z), piecewise linear between the z-bin centres, (M, 5)."""
        return self._ztable(self.zscale, z)

    def generate(self, n_samples: int, chunk: int = 1 << 19):
        """n_samples (r, z) uniform in [r_lo, r_max] x [z_min, z_max], labelled by BEM - axial."""
        dev = device.device()
        Rs, Zs, Ds = [], [], []
        for s in range(0, n_samples, chunk):
            m = min(chunk, n_samples - s)
            r = torch.tensor(self.rng.uniform(self.r_lo, self.r_max, m), dtype=torch.float64, device=dev)
            z = torch.tensor(self.rng.uniform(self.z_min, self.z_max, m), dtype=torch.float64, device=dev)
            pts = torch.stack([r, torch.zeros_like(r), z], dim=1).contiguous()
            Fb, Fa = self._exact_and_axial(pts)
            Rs.append(r)
            Zs.append(z)
            Ds.append(Fb - Fa)
        r, z, D = torch.cat(Rs), torch.cat(Zs), torch.cat(Ds)
        Y = D / self._weight(r)[:, None]
        # s_k(z): RMS of D_k / w per z bin, spread to the neighbours (max over 3 bins) and
        # floored at 1e-6 of the largest, so that no bin divides by (almost) zero
        b = torch.clamp(((z - self.z_min) / (self.z_max - self.z_min) * self.nz).long(), 0, self.nz - 1)
        ss = _bin_sums(b, Y * Y, self.nz)
        cnt = torch.bincount(b, minlength=self.nz).double().clamp(min=1.0)
        rms = torch.sqrt(ss / cnt[:, None]).cpu().numpy()          # (nz, 5)
        pad = np.pad(rms, ((1, 1), (0, 0)), mode="edge")
        rms = np.maximum(np.maximum(pad[:-2], pad[1:-1]), pad[2:])
        top = rms.max(axis=0)
        # a component with no residual at all (e.g. H of a purely electrostatic field) is not
        # learned: unit scale for its (zero) targets, and its correction is zero
        self.active = top > 0
        rms = np.where(top[None, :] > 0, np.maximum(rms, 1e-6 * top[None, :]), 1.0)
        self.zscale = rms.T.copy()
        Y = Y / self._zscale(z)
        self.X = self._inputs(r, z).float().contiguous()
        self.T = Y.float().contiguous()
        return self

    # ---- training ------------------------------------------------------------------------
    def train(self, steps: int = 4000, batch: int = 16384, lr: float = 2e-3, lr_min: float = 2e-5, log_every: int = 0):
        """Adam with a cosine learning-rate schedule on random mini-batches; returns the loss
        history (mean squared error of the scaled targets, every log_every steps)."""
        assert self.X is not None, "generate() first"
        L = _L()
        dev = self.P.device
        n = self.X.shape[0]
        ws = torch.empty(L.tb_mlp_workspace_bytes(batch), dtype=torch.uint8, device=dev)
        Xb = torch.empty((batch, 2), dtype=torch.float32, device=dev)
        Tb = torch.empty((batch, NOUT), dtype=torch.float32, device=dev)
        loss = torch.zeros(1, dtype=torch.float32, device=dev)
        st = device.stream_ptr()
        p = device.ptr
        gen = torch.Generator(device=dev)
        gen.manual_seed(self.seed + self.step)   # mini-batch order follows the surrogate's seed
        t0 = time.time()
        for it in range(steps):
            idx = torch.randint(0, n, (batch,), dtype=torch.int32, device=dev, generator=gen)
            backend.check(L.tb_mlp_gather(batch, p(idx), p(self.X), p(self.T), p(Xb), p(Tb), st), "mlp_gather")
            backend.check(L.tb_mlp_grad(p(self.P), p(self.Wbf), p(Xb), p(Tb), batch, p(self.G), p(loss), p(ws), st), "mlp_grad")
            self.step += 1
            a = lr_min + 0.5 * (lr - lr_min) * (1.0 + math.cos(math.pi * it / max(steps - 1, 1)))
            backend.check(L.tb_mlp_adam(p(self.P), p(self.G), p(self.M), p(self.V), p(self.Wbf), a, 0.9, 0.999, 1e-8, self.step, st),
                          "mlp_adam")
            if log_every and (it % log_every == 0 or it == steps - 1):
                self.history.append((it, float(loss.item()) / (batch * NOUT)))
        torch.cuda.synchronize()
        self.train_seconds = time.time() - t0
        return self.history

    # ---- the embedded error estimate ----------------------------------------------------------
    def calibrate(self, n_samples: int = 1 << 20, chunk: int = 1 << 19, safety: float = 3.0, seed: int = 12345):
        """Fit error of the trained network per z bin on fresh samples (r uniform in
        [r_lo, r_max], z uniform in the trained range), labelled by BEM - axial like the training
        set: sigma_k(z) = safety * RMS(prediction - label) / w(r), per component, spread to the
        neighbouring bins (max over 3) and divided by the largest |E| (components 1, 2) or |H|
        (3, 4) of the samples.  w(r) sigma_k(z) then bounds the correction's error at (r, z)
        relative to the field's scale."""
        dev = device.device()
        rng = np.random.default_rng(seed)
        ss = torch.zeros((self.nz, NOUT), dtype=torch.float64, device=dev)
        cnt = torch.zeros(self.nz, dtype=torch.float64, device=dev)
        sE = sH = 0.0
        for s0 in range(0, n_samples, chunk):
            m = min(chunk, n_samples - s0)
            r = torch.tensor(rng.uniform(self.r_lo, self.r_max, m), dtype=torch.float64, device=dev)
            z = torch.tensor(rng.uniform(self.z_min, self.z_max, m), dtype=torch.float64, device=dev)
            pts = torch.stack([r, torch.zeros_like(r), z], dim=1).contiguous()
            Fb, Fa = self._exact_and_axial(pts)
            sE = max(sE, float(Fb[:, 1:3].abs().max().item()))
            sH = max(sH, float(Fb[:, 3:5].abs().max().item()))
            err = (self.correction(r, z) - (Fb - Fa)) / self._weight(r)[:, None]
            b = torch.clamp(((z - self.z_min) / (self.z_max - self.z_min) * self.nz).long(), 0, self.nz - 1)
            ss += _bin_sums(b, err * err, self.nz)
            cnt += torch.bincount(b, minlength=self.nz).double()
        rms = torch.sqrt(ss / cnt.clamp(min=1.0)[:, None]).cpu().numpy()
        pad = np.pad(rms, ((1, 1), (0, 0)), mode="edge")
        rms = np.maximum(np.maximum(pad[:-2], pad[1:-1]), pad[2:])
        scale = np.array([1.0, sE or 1.0, sE or 1.0, sH or 1.0, sH or 1.0])
        self.zerr = (safety * rms / scale[None, :]).T.copy() * self.active[:, None]
        self.zerr[0] = 0.0   # the potential is not part of the traced force
        return self

    def predicted_error(self, r: torch.Tensor, z: torch.Tensor) -> torch.Tensor:
        """max_k w(r) sigma_k(z) at (r, z) (device tensors): the predicted error of the correction
        relative to the field's scale; 0 where no correction applies."""
        if self.zerr is None:
            self.calibrate()
        e = self._ztable(self.zerr, z).max(dim=1).values * self._weight(r)
        inside = (r >= self.r_lo) & (z >= self.z_min) & (z <= self.z_max)
        return torch.where(inside, e, torch.zeros_like(e))

    # ---- evaluation ----------------------------------------------------------------------
    def correction(self, r: torch.Tensor, z: torch.Tensor) -> torch.Tensor:
        """The learned residual (M, 5) in physical units at (r, z) (device tensors)."""
        L = _L()
        M = r.shape[0]
        X = self._inputs(r, z).float().contiguous()
        Y = torch.empty((M, NOUT), dtype=torch.float32, device=r.device)
        ws = torch.empty(L.tb_mlp_workspace_bytes(M), dtype=torch.uint8, device=r.device)
        backend.check(L.tb_mlp_forward(device.ptr(self.P), device.ptr(self.Wbf), device.ptr(X), M, device.ptr(Y), device.ptr(ws),
                                       device.stream_ptr()), "mlp_forward")
        act = torch.tensor(self.active, dtype=torch.float64, device=r.device)
        return Y.double() * self._zscale(z) * self._weight(r)[:, None] * act[None, :]

    def tracer_arrays(self, error_tolerance: float = float("inf")):
        """Device arrays of the tracer's learned-correction component (csrc/trace_kernels.h):
        FP32 weights with W2..W4 taken at their BF16 values (the function the network was
        trained as) and transposed to (in, out), the tables (10, nz): s_k(z), then
        sigma_k(z) / error_tolerance (a point whose w(r) max_k of these exceeds 1 goes to the
        BEM sum), and the geometry (r_lo, r_max, z_min, z_max, p, 0)."""
        P = self.P.detach().clone()
        o = 384
        for l in range(3):
            # the BF16 values the MFMA forward pass was trained with (Wbf), transposed
            w = self.Wbf[l * 128 * 128:(l + 1) * 128 * 128].view(torch.int16).to(torch.int32) & 0xFFFF
            w = (w << 16).view(torch.float32).view(128, 128)
            P[o:o + 128 * 128] = w.t().contiguous().view(-1)
            o += 128 * 128 + 128
        if self.zerr is None:
            self.calibrate()
        with np.errstate(divide="ignore", invalid="ignore"):
            if error_tolerance <= 0:
                zerr = np.where(self.active[:, None], 1e300, 0.0) * np.ones_like(self.zerr)
            else:
                zerr = self.zerr / error_tolerance
        zerr[0] = 0.0
        tables = np.concatenate([self.zscale * self.active[:, None], zerr], axis=0)
        zs = torch.tensor(tables, dtype=torch.float64, device=P.device).contiguous()
        geom = [self.r_lo, self.r_max, self.z_min, self.z_max, self.p, 0.0]
        return P.contiguous(), zs, geom

    def validate(self, n: int = 20000, bins=(0.1, 0.2, 0.3, 0.5, 0.7, 1.0), seed: int = 99):
        """Max and RMS error of E (and H) relative to the largest |E| (|H|) of the samples, for
        the axial field alone and with the correction, per radial bin (fractions of r_max)."""
        rng = np.random.default_rng(seed)
        dev = self.P.device
        r = torch.tensor(rng.uniform(0.0, self.r_max, n), dtype=torch.float64, device=dev)
        z = torch.tensor(rng.uniform(self.z_min, self.z_max, n), dtype=torch.float64, device=dev)
        pts = torch.stack([r, torch.zeros_like(r), z], dim=1).contiguous()
        Fb, Fa = self._exact_and_axial(pts)
        Fc = Fa + self.correction(r, z)
        out = {}
        Fb_, Fa_, Fc_, r_ = Fb.cpu().numpy(), Fa.cpu().numpy(), Fc.cpu().numpy(), r.cpu().numpy() / self.r_max
        # errors relative to the largest field of the samples (absolute where that field is zero)
        sE = np.max(np.abs(Fb_[:, 1:3])) or 1.0
        sH = np.max(np.abs(Fb_[:, 3:5])) or 1.0
        lo = 0.0
        for hi in bins:
            m = (r_ >= lo) & (r_ < hi)
            if m.any():
                row = {}
                for name, F in (("axial", Fa_), ("learned", Fc_)):
                    eE = np.abs(F[m, 1:3] - Fb_[m, 1:3]) / sE
                    eH = np.abs(F[m, 3:5] - Fb_[m, 3:5]) / sH
                    row[name] = {"E_max": float(eE.max()), "E_rms": float(np.sqrt(np.mean(eE ** 2))),
                                 "H_max": float(eH.max()), "H_rms": float(np.sqrt(np.mean(eH ** 2)))}
                out[f"r/r_max in [{lo:.2f},{hi:.2f})"] = row
            lo = hi
        return out


class FieldRadialAxialLearned(FieldRadialAxial):
    """FieldRadialAxial plus the learned off-axis correction, with direct-BEM fallback outside
    the surrogate's trust region.  Field evaluation and tracing go through the GPU; the
    reference-named single-point trace functions do not apply (trace_batch / Tracer)."""

    def __init__(self, surrogate: FieldSurrogate, trust_radius: float | None = None, error_tolerance: float = 1e-5):
        """trust_radius: the axial field (+ correction) serves r <= trust_radius inside the axial
        range; the direct BEM sum serves everything else.  Default: the surrogate's r_max.
        A trust radius of 0 sends every off-axis evaluation to the BEM sum.
        error_tolerance: inside the trust region, points where the surrogate's predicted error
        (FieldSurrogate.predicted_error, relative to the field's scale) exceeds it also go to the
        BEM sum; inf disables the check, 0 sends every corrected point to the BEM sum."""
        ax = surrogate.axial
        FieldRadialAxial.__bases__[0].__init__(self, ax.field, ax.z, ax.electrostatic_coeffs, ax.magnetostatic_coeffs)
        self.surrogate = surrogate
        self.trust_radius = float(surrogate.r_max if trust_radius is None else trust_radius)
        self.error_tolerance = float(error_tolerance)
        self._origin, self._basis = ax._origin, ax._basis
        self._update_inverse_transformation_matrix()

    def _corrected(self, points):
        """Trust region (as the tracer's): r <= trust_radius inside the axial range; the
        correction where r_lo <= r <= r_max inside the trained z range."""
        P = np.atleast_2d(np.asarray(points, dtype=np.float64))
        s = self.surrogate
        r = np.hypot(P[:, 0], P[:, 1])
        inside = (r <= self.trust_radius) & (P[:, 2] >= self.z[0]) & (P[:, 2] <= self.z[-1])
        dev = device.device()
        rt, zt = torch.tensor(r, device=dev), torch.tensor(P[:, 2], device=dev)
        pe = s.predicted_error(rt, zt).cpu().numpy()
        checked = (r >= s.r_lo) & (P[:, 2] >= s.z_min) & (P[:, 2] <= s.z_max)
        inside &= ~(checked & ((pe > self.error_tolerance) if self.error_tolerance > 0 else True))
        C = s.correction(rt, zt).cpu().numpy()
        C[~((P[:, 2] >= s.z_min) & (P[:, 2] <= s.z_max) & inside)] = 0.0
        return P, r, inside, C

    def electrostatic_field_at_points(self, points):
        P, r, inside, C = self._corrected(points)
        E = super().electrostatic_field_at_points(P)
        with np.errstate(invalid="ignore", divide="ignore"):
            cx = np.where(r > 0, P[:, 0] / r, 0.0)
            cy = np.where(r > 0, P[:, 1] / r, 0.0)
        E[:, 0] += C[:, 1] * cx
        E[:, 1] += C[:, 1] * cy
        E[:, 2] += C[:, 2]
        if (~inside).any():
            E[~inside] = self.field.electrostatic_field_at_points(P[~inside])
        return E

    def magnetostatic_field_at_points(self, points):
        P, r, inside, C = self._corrected(points)
        H = super().magnetostatic_field_at_points(P)
        with np.errstate(invalid="ignore", divide="ignore"):
            cx = np.where(r > 0, P[:, 0] / r, 0.0)
            cy = np.where(r > 0, P[:, 1] / r, 0.0)
        H[:, 0] += C[:, 3] * cx
        H[:, 1] += C[:, 3] * cy
        H[:, 2] += C[:, 4]
        if (~inside).any():
            H[~inside] = self.field.magnetostatic_field_at_points(P[~inside])
        return H

    def device_components(self, factor: float = 1.0, in_superposition: bool = False):
        from . import tracing
        return [tracing.make_component(self, factor, kind="learned", in_superposition=in_superposition)]
