"""Seeded synthetic workloads for the Vxc build.

Random-number specification (fixed so every run and every machine builds the
same inputs): SplitMix64 (Steele, Lea, Flood 2014) produces 64-bit words;
uniform doubles are (word >> 11) * 2^-53; standard normals come in Box-Muller
pairs z0 = r cos(2 pi u2), z1 = r sin(2 pi u2), r = sqrt(-2 ln(1 - u1)).

Workloads:
  water()            H2O, O (0,0,0.1173), H (0,+-0.7572,-0.4692) Angstrom.
  alkane_c8h18()     all-trans n-octane, C-C 1.54, C-H 1.09 Angstrom,
                     tetrahedral angles.
  water_cluster(n)   n rigid waters, uniform random orientation (normalised
                     Gaussian quaternion), centre uniform in a 12 Angstrom
                     cube, rejected if any O-O distance is below 2.8 Angstrom.
  random_density()   C_occ ~ N(0,1) (n_ao x n_occ, row-major draw order),
                     Löwdin S-orthonormalised C <- C (C^T S C)^(-1/2)
                     (crates/calchem-scf/src/guess.rs:146-158), P = 2 C C^T.
"""
from __future__ import annotations

import math
from typing import List, Tuple

import numpy as np

from .basis import BasisSet
from .molecule import Atom, Molecule
from .overlap import overlap_matrix

_MASK = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed: int):
        self.state = seed & _MASK
        self._spare = None

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & _MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK
        return z ^ (z >> 31)

    def uniform(self) -> float:
        return (self.next_u64() >> 11) * (1.0 / 9007199254740992.0)

    def normal(self) -> float:
        if self._spare is not None:
            v, self._spare = self._spare, None
            return v
        u1 = self.uniform()
        u2 = self.uniform()
        r = math.sqrt(-2.0 * math.log(1.0 - u1))
        self._spare = r * math.sin(2.0 * math.pi * u2)
        return r * math.cos(2.0 * math.pi * u2)

    def normals(self, n: int) -> np.ndarray:
        return np.array([self.normal() for _ in range(n)], dtype=np.float64)


WATER_ANGSTROM = [("O", (0.0, 0.0, 0.1173)), ("H", (0.0, 0.7572, -0.4692)),
                  ("H", (0.0, -0.7572, -0.4692))]


def water() -> Molecule:
    return Molecule([Atom.from_angstrom(s, p) for s, p in WATER_ANGSTROM])


def alkane_c8h18(n_carbon: int = 8) -> Molecule:
    """All-trans CnH(2n+2) zig-zag in the xy plane with ideal tetrahedral angles."""
    cc, ch = 1.54, 1.09
    theta = math.acos(-1.0 / 3.0)  # 109.4712 degrees
    half = theta / 2.0
    dx, dy = cc * math.sin(half), cc * math.cos(half)
    def pos(k: int) -> np.ndarray:
        return np.array([k * dx, (k % 2) * dy, 0.0])

    carbons = [pos(i) for i in range(n_carbon)]
    atoms = [("C", c) for c in carbons]
    hyd: List[np.ndarray] = []
    for i, c in enumerate(carbons):
        nbrs = []
        if i > 0:
            nbrs.append(carbons[i - 1] - c)
        if i < n_carbon - 1:
            nbrs.append(carbons[i + 1] - c)
        if len(nbrs) == 1:
            # Terminal carbon: one H sits where the next zig-zag carbon would
            # be (trans to the chain); that direction also fixes the other two.
            virt = pos(i - 1 if i == 0 else i + 1) - c
            hyd.append(c + ch * virt / np.linalg.norm(virt))
            nbrs.append(virt)
        b1 = nbrs[0] / np.linalg.norm(nbrs[0])
        b2 = nbrs[1] / np.linalg.norm(nbrs[1])
        u = -(b1 + b2)
        u /= np.linalg.norm(u)
        nrm = np.cross(b1, b2)
        nrm /= np.linalg.norm(nrm)
        for s in (1.0, -1.0):
            d = u * math.cos(half) + s * nrm * math.sin(half)
            hyd.append(c + ch * d)
    atoms += [("H", h) for h in hyd]
    return Molecule([Atom.from_angstrom(s, p) for s, p in atoms])


def _rotation_from_quaternion(q: np.ndarray) -> np.ndarray:
    w, x, y, z = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


def water_cluster(n_waters: int, seed: int, box: float = 12.0, min_oo: float = 2.8,
                  max_tries: int = 200000) -> Molecule:
    rng = SplitMix64(seed)
    ref = np.array([p for _, p in WATER_ANGSTROM])
    centres: List[np.ndarray] = []
    atoms: List[Tuple[str, np.ndarray]] = []
    tries = 0
    while len(centres) < n_waters:
        tries += 1
        if tries > max_tries:
            raise RuntimeError(f"could not place {n_waters} waters in a {box} A cube")
        q = rng.normals(4)
        q /= np.linalg.norm(q)
        c = np.array([rng.uniform() * box for _ in range(3)])
        if any(np.linalg.norm(c - o) < min_oo for o in centres):
            continue
        R = _rotation_from_quaternion(q)
        xyz = (R @ (ref - ref[0]).T).T + c
        centres.append(c)
        for (s, _), p in zip(WATER_ANGSTROM, xyz):
            atoms.append((s, p))
    return Molecule([Atom.from_angstrom(s, p) for s, p in atoms])


def batching_cluster_sizes(n_clusters: int = 64, seed: int = 3, lo: int = 2, hi: int = 32) -> List[int]:
    rng = SplitMix64(seed)
    return [lo + int(rng.uniform() * (hi - lo + 1)) for _ in range(n_clusters)]


def batching_clusters(n_clusters: int = 64, seed: int = 3) -> List[Molecule]:
    """The many-molecule workload: clusters of uneven sizes (2-32 waters, drawn
    with `seed`), cluster k placed by the water-cluster rule with seed
    1000 * seed + k."""
    sizes = batching_cluster_sizes(n_clusters, seed)
    return [water_cluster(n, 1000 * seed + k) for k, n in enumerate(sizes)]


def random_density(basis: BasisSet, n_occ: int, seed: int, S: np.ndarray = None) -> np.ndarray:
    n = basis.n_ao
    if n_occ <= 0 or n_occ > n:
        raise ValueError(f"n_occ={n_occ} out of range for n_ao={n}")
    rng = SplitMix64(seed)
    C = rng.normals(n * n_occ).reshape(n, n_occ)
    if S is None:
        S = overlap_matrix(basis)
    M = C.T @ S @ C
    ev, U = np.linalg.eigh(M)
    if np.any(ev <= 1e-8):
        raise RuntimeError("occupied metric is not positive definite")
    C = C @ (U @ np.diag(1.0 / np.sqrt(ev)) @ U.T)
    P = 2.0 * C @ C.T
    return 0.5 * (P + P.T)


# Named workloads (molecule builder, basis, grid (n_rad, n_ang), n_occ, density seed).
WORKLOADS = {
    "h2o": dict(build=water, basis="6-31g(d,p)", n_rad=75, n_ang=302, n_occ=5, seed=0,
                funcs=("b3lyp",)),
    "c8h18": dict(build=alkane_c8h18, basis="6-31g(d,p)", n_rad=75, n_ang=302, n_occ=33, seed=1,
                  funcs=("pbe", "b3lyp")),
    "water20": dict(build=lambda: water_cluster(20, 2), basis="6-31g(d,p)", n_rad=75, n_ang=302,
                    n_occ=100, seed=2, funcs=("b3lyp",)),
}
