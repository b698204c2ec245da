"""Atoms and molecules in bohr, mirroring ``calchem_model::{Atom, Molecule}``.

Reference: crates/calchem-model/src/lib.rs:4 (ANGSTROM_TO_BOHR), :32-56
(Atom::from_angstrom / from_bohr) and :113-118 (Molecule).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Sequence

import numpy as np

ANGSTROM_TO_BOHR = 1.8897261245650618
BOHR_TO_ANGSTROM = 1.0 / ANGSTROM_TO_BOHR

_SYMBOLS = [
    "H", "He", "Li", "Be", "B", "C", "N", "O", "F", "Ne", "Na", "Mg", "Al", "Si", "P", "S",
    "Cl", "Ar", "K", "Ca",
]


def atomic_number(symbol: str) -> int:
    s = symbol.strip()
    s = s[:1].upper() + s[1:].lower()
    try:
        return _SYMBOLS.index(s) + 1
    except ValueError as exc:
        raise ValueError(f"unsupported element symbol {symbol!r}") from exc


@dataclass
class Atom:
    symbol: str
    z: int
    position_bohr: np.ndarray

    @classmethod
    def from_angstrom(cls, symbol: str, pos: Sequence[float]) -> "Atom":
        p = np.asarray(pos, dtype=np.float64)
        if p.shape != (3,) or not np.all(np.isfinite(p)):
            raise ValueError(f"atom {symbol!r} has invalid coordinates {pos!r}")
        return cls(symbol.strip(), atomic_number(symbol), p * ANGSTROM_TO_BOHR)

    @classmethod
    def from_bohr(cls, symbol: str, pos: Sequence[float]) -> "Atom":
        p = np.asarray(pos, dtype=np.float64)
        if p.shape != (3,) or not np.all(np.isfinite(p)):
            raise ValueError(f"atom {symbol!r} has invalid coordinates {pos!r}")
        return cls(symbol.strip(), atomic_number(symbol), p.copy())


@dataclass
class Molecule:
    atoms: List[Atom] = field(default_factory=list)
    charge: int = 0
    multiplicity: int = 1

    @property
    def n_atoms(self) -> int:
        return len(self.atoms)

    def z_array(self) -> np.ndarray:
        return np.array([a.z for a in self.atoms], dtype=np.int32)

    def coords_bohr(self) -> np.ndarray:
        return np.array([a.position_bohr for a in self.atoms], dtype=np.float64).reshape(-1, 3)

    def n_electrons(self) -> int:
        return int(sum(a.z for a in self.atoms)) - self.charge
