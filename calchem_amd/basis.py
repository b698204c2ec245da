"""Cartesian Gaussian basis sets, mirroring ``calchem_basis::BasisSet``.

Reference semantics:
  * JSON layout {name, source, elements: {sym: [{shell_type, exponents,
    coefficients}]}}: crates/calchem-basis/src/lib.rs:80-90, data/*.json.
  * primitive_norm, normalized_contraction: crates/calchem-basis/src/normalization.rs:19-97.
  * Cartesian component order per shell type: crates/calchem-basis/src/lib.rs:135-160.
  * BasisSet::build (contraction normalised for lmn = (L,0,0), primitive norm
    of the leading component, AO offsets in atom/shell order): lib.rs:462-520.
  * Per-shell quadrature data (r2 cutoff (40+4l)/alpha_min, weighted norm
    c~ N(alpha; L00), component scale N(alpha_0; lmn)/N(alpha_0; L00)):
    crates/calchem-dft/src/ao_grid.rs:24-71.

The flattened arrays use the reference GPU ABI records AoShellGpu / AoPrimGpu /
AoCompGpu (crates/calchem-cuda/src/lib.rs:46-75) so they can be handed to the
C ABI in include/calchem_hip.h unchanged.
"""
from __future__ import annotations

import json
import math
import os
from dataclasses import dataclass
from typing import Dict, List, Tuple

import numpy as np

from .molecule import Molecule

_DATA_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "basis")

_NAME_TO_FILE = {
    "sto-3g": "sto-3g.json",
    "6-31g": "6-31g.json",
    "6-31g(d)": "6-31gd.json",
    "6-31g*": "6-31gd.json",
    "6-31gd": "6-31gd.json",
    "6-31g(d,p)": "6-31gdp.json",
    "6-31g**": "6-31gdp.json",
    "6-31gdp": "6-31gdp.json",
}

SHELL_L = {"s": 0, "p": 1, "d": 2, "f": 3}

CART_COMPONENTS: Dict[int, List[Tuple[int, int, int]]] = {
    0: [(0, 0, 0)],
    1: [(1, 0, 0), (0, 1, 0), (0, 0, 1)],
    2: [(2, 0, 0), (1, 1, 0), (1, 0, 1), (0, 2, 0), (0, 1, 1), (0, 0, 2)],
    3: [(3, 0, 0), (2, 1, 0), (2, 0, 1), (1, 2, 0), (1, 1, 1), (1, 0, 2), (0, 3, 0),
        (0, 2, 1), (0, 1, 2), (0, 0, 3)],
}

AO_SHELL_DTYPE = np.dtype(
    [("center", "<f8", (3,)), ("r2_cutoff", "<f8"), ("ao_offset", "<i4"), ("n_funcs", "<i4"),
     ("prim_start", "<i4"), ("n_prims", "<i4"), ("comp_start", "<i4"), ("atom_index", "<i4")],
    align=True)
AO_PRIM_DTYPE = np.dtype([("exponent", "<f8"), ("weighted_norm", "<f8")], align=True)
AO_COMP_DTYPE = np.dtype([("scale", "<f8"), ("lx", "<i4"), ("ly", "<i4"), ("lz", "<i4"),
                          ("pad", "<i4")], align=True)
assert AO_SHELL_DTYPE.itemsize == 56 and AO_PRIM_DTYPE.itemsize == 16
assert AO_COMP_DTYPE.itemsize == 24


def _dfact_odd(k: int) -> float:
    acc = 1.0
    while k > 0:
        acc *= k
        k -= 2
    return acc


def primitive_norm(alpha: float, lmn: Tuple[int, int, int]) -> float:
    """N(alpha; l,m,n) = (2a/pi)^(3/4) sqrt((4a)^L / ((2l-1)!!(2m-1)!!(2n-1)!!))."""
    l, m, n = lmn
    radial = (2.0 * alpha / math.pi) ** 0.75
    num = (4.0 * alpha) ** (l + m + n)
    den = _dfact_odd(2 * l - 1) * _dfact_odd(2 * m - 1) * _dfact_odd(2 * n - 1)
    return radial * math.sqrt(num / den)


def normalized_contraction(exps, coefs, L: int) -> np.ndarray:
    """Scale raw coefficients so that <phi|phi> = 1 (normalization.rs:59-97)."""
    e = np.asarray(exps, dtype=np.float64)
    c = np.asarray(coefs, dtype=np.float64)
    if e.size == 0 or e.size != c.size:
        raise ValueError("invalid primitive exponents/coefficients length")
    if np.any(~np.isfinite(e)) or np.any(e <= 0.0) or np.any(~np.isfinite(c)):
        raise ValueError("primitive exponents must be finite and positive")
    # Concentric overlap of normalised primitives (l,0,0): the angular factors
    # cancel and only (2 sqrt(a_i a_j) / (a_i + a_j))^(L + 3/2) remains.
    g = e[:, None] + e[None, :]
    s_ij = (2.0 * np.sqrt(e[:, None] * e[None, :]) / g) ** (L + 1.5)
    self_ov = float(c @ s_ij @ c)
    if not (self_ov > 0.0 and math.isfinite(self_ov)):
        raise ValueError("contracted shell self-overlap is non-positive")
    return c / math.sqrt(self_ov)


@dataclass
class Shell:
    atom_index: int
    center: np.ndarray
    l: int
    exponents: np.ndarray
    coeffs: np.ndarray       # contraction-normalised c~
    raw_coeffs: np.ndarray
    norms: np.ndarray        # N(alpha; L,0,0)
    ao_offset: int

    @property
    def n_functions(self) -> int:
        return (self.l + 1) * (self.l + 2) // 2


class BasisSet:
    """Basis set on a molecule (calchem-basis/src/lib.rs:248-256,462-520)."""

    def __init__(self, name: str, shells: List[Shell], n_ao: int):
        self.name = name
        self.shells = shells
        self.n_ao = n_ao

    @staticmethod
    def load_raw(name: str) -> dict:
        key = name.strip().lower()
        if key not in _NAME_TO_FILE:
            raise ValueError(f"unsupported basis {name!r}; available: {sorted(_NAME_TO_FILE)}")
        with open(os.path.join(_DATA_DIR, _NAME_TO_FILE[key]), "r", encoding="utf-8") as fh:
            return json.load(fh)

    @classmethod
    def from_name(cls, name: str, mol: Molecule) -> "BasisSet":
        raw = cls.load_raw(name)
        if mol.n_atoms == 0:
            raise ValueError("cannot build basis set for empty molecule")
        shells: List[Shell] = []
        ao = 0
        for ai, atom in enumerate(mol.atoms):
            elem = raw["elements"].get(atom.symbol)
            if not elem:
                raise ValueError(f"element {atom.symbol!r} not in basis {raw.get('name')!r}")
            for rs in elem:
                L = SHELL_L[rs["shell_type"].strip().lower()]
                e = np.asarray(rs["exponents"], dtype=np.float64)
                c = np.asarray(rs["coefficients"], dtype=np.float64)
                cn = normalized_contraction(e, c, L)
                norms = np.array([primitive_norm(a, (L, 0, 0)) for a in e])
                sh = Shell(ai, atom.position_bohr.copy(), L, e, cn, c, norms, ao)
                shells.append(sh)
                ao += sh.n_functions
        return cls(raw.get("name", name), shells, ao)

    # ------------------------------------------------------------------ #
    def raw_arrays(self) -> dict:
        """Unnormalised shell data (what the JSON file holds) as flat arrays."""
        n_prim = np.array([s.exponents.size for s in self.shells], dtype=np.int32)
        prim_start = np.zeros(len(self.shells), dtype=np.int32)
        if len(self.shells) > 1:
            prim_start[1:] = np.cumsum(n_prim)[:-1]
        return {
            "atom": np.array([s.atom_index for s in self.shells], dtype=np.int32),
            "center": np.array([s.center for s in self.shells], dtype=np.float64).reshape(-1, 3),
            "l": np.array([s.l for s in self.shells], dtype=np.int32),
            "n_prim": n_prim,
            "prim_start": prim_start,
            "exps": np.concatenate([s.exponents for s in self.shells]).astype(np.float64),
            "coefs": np.concatenate([s.raw_coeffs for s in self.shells]).astype(np.float64),
        }

    def ao_records(self) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
        """AoShellGpu / AoPrimGpu / AoCompGpu records (ao_grid.rs:24-71)."""
        shells = np.zeros(len(self.shells), dtype=AO_SHELL_DTYPE)
        n_p = sum(s.exponents.size for s in self.shells)
        n_c = sum(s.n_functions for s in self.shells)
        prims = np.zeros(n_p, dtype=AO_PRIM_DTYPE)
        comps = np.zeros(n_c, dtype=AO_COMP_DTYPE)
        ip = ic = 0
        for k, s in enumerate(self.shells):
            amin = max(float(np.min(s.exponents)), 1e-12)
            shells[k]["center"] = s.center
            shells[k]["r2_cutoff"] = (40.0 + 4.0 * s.l) / amin
            shells[k]["ao_offset"] = s.ao_offset
            shells[k]["n_funcs"] = s.n_functions
            shells[k]["prim_start"] = ip
            shells[k]["n_prims"] = s.exponents.size
            shells[k]["comp_start"] = ic
            shells[k]["atom_index"] = s.atom_index
            for a, cn, nm in zip(s.exponents, s.coeffs, s.norms):
                prims[ip]["exponent"] = a
                prims[ip]["weighted_norm"] = cn * nm
                ip += 1
            lead = primitive_norm(float(s.exponents[0]), (s.l, 0, 0))
            for lmn in CART_COMPONENTS[s.l]:
                comps[ic]["scale"] = primitive_norm(float(s.exponents[0]), lmn) / lead
                comps[ic]["lx"], comps[ic]["ly"], comps[ic]["lz"] = lmn
                ic += 1
        return shells, prims, comps

    def min_exponent(self) -> float:
        return float(min(np.min(s.exponents) for s in self.shells))
