"""Excitation specification — counterpart of traceon/excitation.py.

Same enum values (excitation.py:28-87, definitions.c:34-40), same ``add_*`` methods,
same superposition split (excitation.py:318-346) and active-element selection
(excitation.py:348-404), so solver code and user scripts read as with the reference.
Only the radial symmetry is implemented.
"""
from __future__ import annotations

from enum import IntEnum

import numpy as np

MU_0 = 1.25663706212e-06


class Symmetry(IntEnum):
    RADIAL = 0
    THREE_D = 2

    def __str__(self):
        return "radial" if self == Symmetry.RADIAL else "3d"

    def is_2d(self):
        return self == Symmetry.RADIAL

    def is_3d(self):
        return self == Symmetry.THREE_D


class ExcitationType(IntEnum):
    VOLTAGE_FIXED = 1
    VOLTAGE_FUN = 2
    DIELECTRIC = 3
    CURRENT = 4
    MAGNETOSTATIC_POT = 5
    MAGNETIZABLE = 6
    PERMANENT_MAGNET = 7

    def is_electrostatic(self) -> bool:
        return self in (ExcitationType.VOLTAGE_FIXED, ExcitationType.VOLTAGE_FUN, ExcitationType.DIELECTRIC)

    def is_magnetostatic(self) -> bool:
        return self in (ExcitationType.MAGNETOSTATIC_POT, ExcitationType.MAGNETIZABLE, ExcitationType.CURRENT,
                        ExcitationType.PERMANENT_MAGNET)


# types whose excitation produces a field by itself (excitation.py:311-316)
_SUPERPOSED = (ExcitationType.VOLTAGE_FIXED, ExcitationType.VOLTAGE_FUN, ExcitationType.CURRENT, ExcitationType.PERMANENT_MAGNET)


class Excitation:
    def __init__(self, mesh, symmetry: Symmetry = Symmetry.RADIAL) -> None:
        if symmetry != Symmetry.RADIAL:
            raise NotImplementedError("only radial symmetry is implemented")
        self.mesh = mesh
        self.electrodes = mesh.get_electrodes()
        self.excitation_types: dict[str, tuple[ExcitationType, object]] = {}
        self.symmetry = symmetry
        assert mesh.points.shape[1] == 2 or np.all(mesh.points[:, 1] == 0.0), \
            "When symmetry is RADIAL, the geometry should lie in the XZ plane"

    def __str__(self):
        return "<Excitation " + ", ".join(f"{n}={v} ({t.name})" for n, (t, v) in self.excitation_types.items()) + ">"

    def _lines(self, what: str, name: str) -> None:
        assert name in self.electrodes, f"Electrode '{name}' is not present in the mesh"
        assert name in self.mesh.physical_to_lines, f"{what} on '{name}' needs line elements in radial symmetry"

    def _triangles(self, what: str, name: str) -> None:
        assert name in self.electrodes, f"Electrode '{name}' is not present in the mesh"
        assert name in self.mesh.physical_to_triangles, f"{what} on '{name}' needs triangles in radial symmetry"

    def add_voltage(self, **kwargs) -> None:
        for name, voltage in kwargs.items():
            self._lines("voltage", name)
            if isinstance(voltage, (int, float, np.floating, np.integer)):
                self.excitation_types[name] = (ExcitationType.VOLTAGE_FIXED, float(voltage))
            elif callable(voltage):
                self.excitation_types[name] = (ExcitationType.VOLTAGE_FUN, voltage)
            else:
                raise NotImplementedError("Unrecognized voltage value")

    def add_current(self, **kwargs) -> None:
        for name, current in kwargs.items():
            self._triangles("current", name)
            self.excitation_types[name] = (ExcitationType.CURRENT, float(current))

    def add_magnetostatic_potential(self, **kwargs) -> None:
        for name, pot in kwargs.items():
            self._lines("magnetostatic potential", name)
            self.excitation_types[name] = (ExcitationType.MAGNETOSTATIC_POT, float(pot))

    def add_magnetizable(self, **kwargs) -> None:
        for name, mu in kwargs.items():
            self._lines("magnetizable", name)
            self.excitation_types[name] = (ExcitationType.MAGNETIZABLE, float(mu))

    def add_permanent_magnet(self, **kwargs) -> None:
        for name, vec in kwargs.items():
            v = np.array(vec, dtype=np.float64) / MU_0   # Tesla -> A/m (excitation.py:229)
            self._lines("permanent magnet", name)
            assert v.shape == (3,) and v[0] == 0.0 and v[1] == 0.0, \
                "in radial symmetry the magnetization must be [0, 0, B]"
            self.excitation_types[name] = (ExcitationType.PERMANENT_MAGNET, v)

    def add_dielectric(self, **kwargs) -> None:
        for name, K in kwargs.items():
            self._lines("dielectric", name)
            self.excitation_types[name] = (ExcitationType.DIELECTRIC, float(K))

    def add_electrostatic_boundary(self, *args: str, ensure_inward_normals: bool = True) -> None:
        """A boundary is a dielectric with K = 0 (excitation.py:263-285)."""
        if ensure_inward_normals:
            for e in args:
                self.mesh.ensure_inward_normals(e)
        self.add_dielectric(**{a: 0.0 for a in args})

    def add_magnetostatic_boundary(self, *args: str, ensure_inward_normals: bool = True) -> None:
        if ensure_inward_normals:
            for e in args:
                self.mesh.ensure_inward_normals(e)
        self.add_magnetizable(**{a: 0.0 for a in args})

    def has_permanent_magnet(self) -> bool:
        return any(t == ExcitationType.PERMANENT_MAGNET for t, _ in self.excitation_types.values())

    def has_current(self) -> bool:
        return any(t == ExcitationType.CURRENT for t, _ in self.excitation_types.values())

    def is_electrostatic(self) -> bool:
        return any(t in (ExcitationType.VOLTAGE_FIXED, ExcitationType.VOLTAGE_FUN) for t, _ in self.excitation_types.values())

    def is_magnetostatic(self) -> bool:
        return any(t in (ExcitationType.MAGNETOSTATIC_POT, ExcitationType.PERMANENT_MAGNET, ExcitationType.CURRENT)
                   for t, _ in self.excitation_types.values())

    def _split_for_superposition(self):
        """One excitation per field-producing electrode, the others zeroed (excitation.py:318-346)."""
        elec, mag = {}, {}
        for name, (t, v) in self.excitation_types.items():
            if t not in _SUPERPOSED:
                continue
            types = {}
            for n, (t2, v2) in self.excitation_types.items():
                if n == name or t2 not in _SUPERPOSED:
                    types[n] = (t2, v2)
                elif t2 == ExcitationType.VOLTAGE_FUN:
                    types[n] = (t2, lambda *_: 0.0)
                else:
                    types[n] = (t2, np.zeros_like(v2) if isinstance(v2, np.ndarray) else 0.0)
            exc = Excitation(self.mesh, self.symmetry)
            exc.excitation_types = types
            (elec if t.is_electrostatic() else mag)[name] = exc
        return elec, mag

    def _get_active_elements(self, kind: str):
        """Lines with an excitation of the kind, in mesh order (excitation.py:348-373)."""
        if kind == "electrostatic":
            check = lambda t: t.is_electrostatic()
        else:
            check = lambda t: t in (ExcitationType.MAGNETIZABLE, ExcitationType.MAGNETOSTATIC_POT)
        elements = self.mesh.lines
        physicals = self.mesh.physical_to_lines
        inactive = np.full(len(elements), True)
        for name, (t, _) in self.excitation_types.items():
            if check(t) and name in physicals:
                inactive[physicals[name]] = False
        map_index = np.arange(len(elements)) - np.cumsum(inactive)
        names = {n: map_index[i] for n, i in physicals.items() if n in self.excitation_types and check(self.excitation_types[n][0])}
        return self.mesh.points[elements[~inactive].astype(np.int64)], names

    def get_electrostatic_active_elements(self):
        return self._get_active_elements("electrostatic")

    def get_magnetostatic_active_elements(self):
        return self._get_active_elements("magnetostatic")
