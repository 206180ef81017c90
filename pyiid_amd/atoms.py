"""Minimal stand-in for ``ase.Atoms`` carrying exactly what the scattering
path touches: positions, atomic numbers, per-atom arrays, ``info``, masses,
momenta and a calculator hook (the reference reads ``get_positions``,
``numbers``/``get_atomic_numbers``, ``arrays``, ``set_array``/``get_array`` and
``info['exp'/'scatter_atoms']``: elasticscatter/__init__.py:110-151,
flat_multi_cpu_wrap.py:9-19).  Real ``ase.Atoms`` objects work as well,
wherever ASE is installed.
"""
import re
from copy import deepcopy

import numpy as np

from .formfactors import MASSES, SYMBOL_TO_Z


def _parse_symbols(symbols):
    if isinstance(symbols, str):
        out = []
        for sym, cnt in re.findall(r"([A-Z][a-z]?)(\d*)", symbols):
            out += [SYMBOL_TO_Z[sym]] * (int(cnt) if cnt else 1)
        return np.array(out, dtype=np.int64)
    vals = list(symbols)
    if vals and isinstance(vals[0], str):
        return np.array([SYMBOL_TO_Z[s] for s in vals], dtype=np.int64)
    return np.array(vals, dtype=np.int64)


class Atoms:
    """A structure: ``Atoms('Au4', positions)`` or ``Atoms(numbers=..., positions=...)``."""

    def __init__(self, symbols=None, positions=None, numbers=None, masses=None):
        if numbers is None:
            numbers = _parse_symbols(symbols if symbols is not None else [])
        numbers = np.asarray(numbers, dtype=np.int64)
        n = len(numbers)
        pos = np.zeros((n, 3)) if positions is None else np.array(positions, dtype=np.float64)
        if pos.shape != (n, 3):
            raise ValueError("positions must have shape (%d, 3)" % n)
        self.arrays = {"numbers": numbers.copy(), "positions": pos}
        self.info = {}
        self._masses = None if masses is None else np.asarray(masses, dtype=np.float64)
        self._momenta = np.zeros((n, 3))
        self.calc = None

    # --- geometry ---------------------------------------------------------
    def __len__(self):
        return len(self.arrays["numbers"])

    @property
    def numbers(self):
        return self.arrays["numbers"]

    @property
    def positions(self):
        return self.arrays["positions"]

    @positions.setter
    def positions(self, value):
        self.arrays["positions"][:] = value

    def get_positions(self):
        return self.arrays["positions"].copy()

    def set_positions(self, value):
        self.arrays["positions"] = np.array(value, dtype=np.float64).reshape(len(self), 3)

    def get_atomic_numbers(self):
        return self.arrays["numbers"].copy()

    def get_array(self, name, copy=True):
        a = self.arrays[name]
        return a.copy() if copy else a

    def set_array(self, name, a):
        if a is None:
            self.arrays.pop(name, None)
            return
        a = np.array(a)
        if a.shape[0] != len(self):
            raise ValueError("array %r has %d rows for %d atoms" % (name, a.shape[0], len(self)))
        self.arrays[name] = a

    def get_masses(self):
        if self._masses is not None:
            return self._masses.copy()
        return np.array([MASSES[int(z)] for z in self.numbers], dtype=np.float64)

    def get_center_of_mass(self):
        m = self.get_masses()
        return (m[:, None] * self.positions).sum(0) / m.sum()

    def center(self):
        """Translate the bounding box centre to the origin (distances unchanged)."""
        if len(self):
            p = self.arrays["positions"]
            p -= (p.min(0) + p.max(0)) / 2.0

    # --- dynamics ---------------------------------------------------------
    def get_momenta(self):
        return self._momenta.copy()

    def set_momenta(self, p):
        self._momenta = np.array(p, dtype=np.float64).reshape(len(self), 3)

    def get_velocities(self):
        return self._momenta / self.get_masses()[:, None]

    def set_velocities(self, v):
        self.set_momenta(np.asarray(v) * self.get_masses()[:, None])

    def get_kinetic_energy(self):
        return 0.5 * float(np.sum(self._momenta ** 2 / self.get_masses()[:, None]))

    # --- calculator -------------------------------------------------------
    def set_calculator(self, calc):
        self.calc = calc

    def get_potential_energy(self):
        if self.calc is None:
            raise RuntimeError("no calculator attached")
        return self.calc.get_potential_energy(self)

    def get_forces(self):
        if self.calc is None:
            raise RuntimeError("no calculator attached")
        return self.calc.get_forces(self)

    def get_total_energy(self):
        return self.get_potential_energy() + self.get_kinetic_energy()

    # --- editing ----------------------------------------------------------
    def copy(self):
        new = Atoms(numbers=self.numbers, positions=self.positions, masses=self._masses)
        for k, v in self.arrays.items():
            if k not in ("numbers", "positions"):
                new.arrays[k] = v.copy()
        new.info = deepcopy(self.info)
        new._momenta = self._momenta.copy()
        new.calc = self.calc
        return new

    def __add__(self, other):
        if not isinstance(other, Atoms):
            other = Atoms(numbers=[other[0]], positions=[other[1]])
        return Atoms(numbers=np.concatenate([self.numbers, other.numbers]),
                     positions=np.concatenate([self.positions, other.positions]))

    def __delitem__(self, idx):
        keep = np.ones(len(self), dtype=bool)
        keep[idx] = False
        for k in list(self.arrays):
            self.arrays[k] = self.arrays[k][keep]
        self._momenta = self._momenta[keep]
        if self._masses is not None:
            self._masses = self._masses[keep]
