"""BEM solver — counterpart of traceon/solver.py (solve_direct, solve_direct_superposition,
ElectrostaticSolverRadial, MagnetostaticSolverRadial).

The assembly (solver.py:206-253), right-hand side (118-150) and dense solve (160-180) keep
the reference's semantics; each runs on the GPU:

* element buffers, normals and the dense fill + singular diagonal in HIP kernels on a
  device-resident matrix (no host copy of the N x N matrix);
* the solve by the library's LU with partial pivoting.  ``precision="fp64"`` (default)
  factors the FP64 matrix; ``precision="mixed"`` factors an FP32 copy on MFMA.  Both refine
  the charges with double-double residuals against the FP64 matrix, so they converge to
  the correctly rounded solution of the assembled system.
"""
from __future__ import annotations

import copy
import time

import numpy as np

from . import backend
from . import excitation as E
from .field import EffectivePointCharges, FieldRadialBEM


def _log(msg):
    import os
    if os.environ.get("TRACEON_LOG_LEVEL", "").upper() in ("INFO", "DEBUG"):
        print(msg)


class SolverRadial:
    """solver.py:39-253 (radial)."""

    def __init__(self, excitation, precision: str = "fp64"):
        from . import device
        self.excitation = excitation
        self.precision = precision
        vertices, names = self.get_active_elements()
        self.vertices = np.ascontiguousarray(vertices, dtype=np.float64)
        self.names = names
        N = len(vertices)
        types = np.zeros(N, dtype=np.uint8)
        values = np.zeros(N)
        for n, idx in names.items():
            t, v = excitation.excitation_types[n]
            types[idx] = int(t)
            if t != E.ExcitationType.VOLTAGE_FUN:
                values[idx] = v
        self.excitation_types = types
        self.excitation_values = values
        self.elements = device.Elements(self.vertices, types, values) if N else None
        if N:
            import torch

            def row_info():
                # RowInfo per element: r0, z0, nr, nz, flux factor, (kind, pad)
                return self.elements.rows.view(dtype=torch.float64).reshape(N, -1).cpu().numpy()

            info = row_info()
            self._targets = info[:, 0:2].copy()
            # VOLTAGE_FUN values at the element centres (solver.py:57-61, 100-109)
            if np.any(types == int(E.ExcitationType.VOLTAGE_FUN)):
                centres = self.centres()
                for n, idx in names.items():
                    t, f = excitation.excitation_types[n]
                    if t == E.ExcitationType.VOLTAGE_FUN:
                        values[idx] = [f(*centres[i]) for i in idx]
                self.elements = device.Elements(self.vertices, types, values)
                info = row_info()
            self.jac_buffer = self.elements.jac.cpu().numpy()
            self.pos_buffer = self.elements.pos.cpu().numpy()
            self.normals = info[:, 2:4].copy()
        else:
            self.jac_buffer = np.zeros((0, 16))
            self.pos_buffer = np.zeros((0, 16, 2))
            self.normals = np.zeros((0, 2))
            self._targets = np.zeros((0, 2))

    def get_number_of_matrix_elements(self) -> int:
        return len(self.vertices)

    def get_normal_vectors(self):
        return self.normals

    def centres(self):
        """Element centres at alpha = 0 as 3D points (solver.py:100-109), from the device's
        per-row data."""
        t = self._targets
        return np.stack([t[:, 0], np.zeros(len(t)), t[:, 1]], axis=1)

    def get_center_of_element(self, index):
        return self.centres()[index]

    def get_preexisting_field_at_points(self, points):
        return np.zeros((len(points), 3))

    def get_right_hand_side(self):
        """solver.py:118-150."""
        st = time.time()
        N = self.get_number_of_matrix_elements()
        F = np.zeros(N)
        t = self.excitation_types
        pot = np.isin(t, [1, 2, 5])
        F[pot] = self.excitation_values[pot]
        mag = t == int(E.ExcitationType.MAGNETIZABLE)
        if np.any(mag):
            idx = np.where(mag)[0]
            centres = self.centres()[idx]
            H = self.get_preexisting_field_at_points(centres)
            n3 = np.stack([self.normals[idx, 0], np.zeros(len(idx)), self.normals[idx, 1]], axis=1)
            K = self.excitation_values[idx]
            F[idx] = -(2.0 * (K - 1) / (1 + K)) * np.sum(H * n3, axis=1)
        assert np.all(np.isfinite(F))
        _log(f"Computing right hand side of linear system took {(time.time() - st) * 1000:.0f} ms")
        return F

    def get_matrix_device(self, fp32: bool = False):
        from . import device
        return device.assemble(self.elements, fp64=True, fp32=fp32)

    def get_matrix(self):
        """The assembled matrix as a numpy array (solver.py:206-253)."""
        M, _ = self.get_matrix_device()
        M = M.cpu().numpy()
        assert np.all(np.isfinite(M)), "Matrix contains non-finite values."
        return M

    def solve_matrix(self, right_hand_side=None):
        """solver.py:160-180: charges for each right-hand side, wrapped as fields."""
        from . import device
        F = np.array([self.get_right_hand_side()]) if right_hand_side is None else np.atleast_2d(right_hand_side)
        N = self.get_number_of_matrix_elements()
        if N == 0:
            return [self.charges_to_field(EffectivePointCharges.empty_2d()) for _ in range(len(F))]
        assert all(f.shape == (N,) for f in F)
        st = time.time()
        mixed = self.precision == "mixed"
        M64, M32 = self.get_matrix_device(fp32=mixed)
        import torch
        assert bool(torch.isfinite(M64).all().item()), "Matrix contains non-finite values."
        X, stats, _ = device.solve_system(M64, F, M32=M32, mode="mixed" if mixed else "fp64")
        charges = X.cpu().numpy()
        self.last_solve_stats = stats
        _log(f"Time for building and solving matrix: {(time.time() - st) * 1000:.0f} ms ({stats})")
        assert np.all(np.isfinite(charges)) and charges.shape == F.shape
        return [self.charges_to_field(EffectivePointCharges(c, self.jac_buffer, self.pos_buffer)) for c in charges]


class ElectrostaticSolverRadial(SolverRadial):
    def get_active_elements(self):
        return self.excitation.get_electrostatic_active_elements()

    def charges_to_field(self, charges):
        return FieldRadialBEM(electrostatic_point_charges=charges)


class MagnetostaticSolverRadial(SolverRadial):
    def __init__(self, excitation, precision: str = "fp64"):
        super().__init__(excitation, precision)
        self.preexisting_field = FieldRadialBEM(magnetostatic_point_charges=self.get_permanent_magnet_field().magnetostatic_point_charges,
                                                current_point_charges=self.get_current_field().current_point_charges)

    def get_active_elements(self):
        return self.excitation.get_magnetostatic_active_elements()

    def get_preexisting_field_at_points(self, points):
        return self.preexisting_field.magnetostatic_field_at_points(points)

    def get_permanent_magnet_field(self):
        """solver.py:279-315: magnetic surface charge n . M on permanent-magnet elements."""
        mesh = self.excitation.mesh
        if not len(mesh.lines) or not self.excitation.has_permanent_magnet():
            return FieldRadialBEM(magnetostatic_point_charges=EffectivePointCharges.empty_2d())
        import torch
        from . import device
        all_vertices = np.ascontiguousarray(mesh.points[mesh.lines.astype(np.int64)], dtype=np.float64)
        # jacobian buffers and normals at alpha = 0 (higher_order_normal_radial) from one device pass
        els = device.Elements(all_vertices, np.ones(len(all_vertices), dtype=np.uint8), np.zeros(len(all_vertices)))
        jac, pos = els.jac.cpu().numpy(), els.pos.cpu().numpy()
        normals = els.rows.view(dtype=torch.float64).reshape(len(all_vertices), -1).cpu().numpy()[:, 2:4]
        charges, jacs, poss = [], [], []
        for name, (t, vec) in self.excitation.excitation_types.items():
            if t != E.ExcitationType.PERMANENT_MAGNET or name not in mesh.physical_to_lines:
                continue
            idx = mesh.physical_to_lines[name]
            if not len(idx):
                continue
            n = normals[idx]
            charges.extend(n[:, 0] * vec[0] + n[:, 1] * vec[2])
            jacs.extend(jac[idx])
            poss.extend(pos[idx])
        if not charges:
            return FieldRadialBEM(magnetostatic_point_charges=EffectivePointCharges.empty_2d())
        return FieldRadialBEM(magnetostatic_point_charges=EffectivePointCharges(np.array(charges), np.array(jacs), np.array(poss)))

    def get_current_field(self):
        """solver.py:317-348: coil triangles -> current rings with density I / area."""
        mesh = self.excitation.mesh
        if not len(mesh.triangles) or not self.excitation.has_current():
            return FieldRadialBEM(current_point_charges=EffectivePointCharges.empty_3d())
        jac, pos = backend.fill_jacobian_buffer_3d(mesh.points[mesh.triangles.astype(np.int64)])
        currents, jacs, poss = [], [], []
        for name, (t, I) in self.excitation.excitation_types.items():
            if t != E.ExcitationType.CURRENT or name not in mesh.physical_to_triangles:
                continue
            idx = mesh.physical_to_triangles[name]
            if not len(idx):
                continue
            area = np.sum(jac[idx])
            currents.extend(np.full(len(idx), I / area))
            jacs.extend(jac[idx])
            poss.extend(pos[idx])
        if not currents:
            return FieldRadialBEM(current_point_charges=EffectivePointCharges.empty_3d())
        return FieldRadialBEM(current_point_charges=EffectivePointCharges(np.array(currents), np.array(jacs), np.array(poss)))

    def charges_to_field(self, charges):
        return FieldRadialBEM(magnetostatic_point_charges=self.preexisting_field.magnetostatic_point_charges + charges,
                              current_point_charges=self.preexisting_field.current_point_charges)


def _excitation_to_higher_order(excitation):
    """solver.py:356-362: a mesh of two-point lines is raised to cubic elements before solving."""
    mesh = excitation.mesh
    if mesh.is_2d() and len(mesh.lines) and not mesh.is_higher_order():
        excitation = copy.copy(excitation)
        excitation.mesh = mesh.to_higher_order()
    return excitation


def solve_direct_superposition(excitation, precision: str = "fp64"):
    """solver.py:364-400: one factorisation, one right-hand side per electrode."""
    excitation = _excitation_to_higher_order(excitation)
    elec, mag = excitation._split_for_superposition()
    out = {}
    if elec:
        names = list(elec.keys())
        F = np.array([ElectrostaticSolverRadial(elec[n], precision).get_right_hand_side() for n in names])
        out.update(dict(zip(names, ElectrostaticSolverRadial(excitation, precision).solve_matrix(F))))
    if mag:
        names = list(mag.keys())
        F = np.array([MagnetostaticSolverRadial(mag[n], precision).get_right_hand_side() for n in names])
        out.update(dict(zip(names, MagnetostaticSolverRadial(excitation, precision).solve_matrix(F))))
    return out


def solve_direct(excitation, precision: str = "fp64") -> FieldRadialBEM:
    """solver.py:402-431."""
    excitation = _excitation_to_higher_order(excitation)
    mag, elec = excitation.is_magnetostatic(), excitation.is_electrostatic()
    if mag and elec:
        e = ElectrostaticSolverRadial(excitation, precision).solve_matrix()[0]
        m = MagnetostaticSolverRadial(excitation, precision).solve_matrix()[0]
        return e + m
    if elec:
        return ElectrostaticSolverRadial(excitation, precision).solve_matrix()[0]
    if mag:
        return MagnetostaticSolverRadial(excitation, precision).solve_matrix()[0]
    raise RuntimeError("Solving for an empty excitation")
