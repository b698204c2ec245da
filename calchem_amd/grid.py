"""Molecular DFT grid, mirroring ``calchem_dft::grid::MolecularGrid``.

MolecularGrid.build / build_for_basis follow crates/calchem-dft/src/grid.rs:614-634
(GridLevel radial counts and angular pruning, Fine's 99-shell floor,
basis-aware radial minimum grid.rs:648-661); MolecularGrid.explicit builds an
unpruned (n_rad x n_ang) grid with the same radial map and Becke partition.
Becke weights are computed on the GPU by libcalchem_hip.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import List, Optional, Tuple

import numpy as np

from . import _native
from .molecule import Molecule

LEVELS = {"coarse": 0, "medium": 1, "fine": 2, "ultrafine": 3}
EXPLICIT = 4


def level_from_name(name: str) -> str:
    n = name.strip().lower()
    if n in ("coarse", "small", "1"):
        return "coarse"
    if n in ("fine", "3"):
        return "fine"
    if n in ("ultrafine", "ultra_fine", "4"):
        return "ultrafine"
    return "medium"


def radial_min_for_basis(basis) -> int:
    """Extra radial shells for diffuse exponents (grid.rs:648-661)."""
    alpha = basis.min_exponent()
    extra = 0
    if alpha < 1e-3:
        extra = int(math.ceil(180.0 * (1e-4 / max(alpha, 1e-12)) ** 0.25))
    return min(extra, 4096)


@dataclass
class MolecularGrid:
    points: np.ndarray  # [n, 6] rows (x, y, z, weight, raw_weight, host_atom); empty when streaming
    level: str
    # streaming grids (grid.rs:636-729): (molecule, level id, n_rad, n_ang, radial_min), points made per task
    streaming: Optional[tuple] = None

    @property
    def n_points(self) -> int:
        return int(self.points.shape[0])

    @property
    def coords(self) -> np.ndarray:
        return self.points[:, :3]

    @property
    def weights(self) -> np.ndarray:
        return self.points[:, 3]

    def xyzw(self) -> np.ndarray:
        return np.ascontiguousarray(self.points[:, :4])

    @staticmethod
    def _build(mol: Molecule, level: int, n_rad: int, n_ang: int, radial_min: int, name: str):
        L = _native.require_device()
        z = np.ascontiguousarray(mol.z_array(), dtype=np.int32)
        xyz = np.ascontiguousarray(mol.coords_bohr().ravel(), dtype=np.float64)
        cnt = L.calchem_hip_grid_build(len(z), _native.ptr(z), _native.ptr(xyz), level, n_rad, n_ang,
                                       radial_min, None, 0)
        _native.check(cnt, "calchem_hip_grid_build (count)")
        out = np.zeros((max(cnt, 1), 6))
        got = L.calchem_hip_grid_build(len(z), _native.ptr(z), _native.ptr(xyz), level, n_rad, n_ang,
                                       radial_min, _native.ptr(out), cnt)
        _native.check(got, "calchem_hip_grid_build")
        return MolecularGrid(out[:got].copy(), name)

    @classmethod
    def build(cls, mol: Molecule, level: str = "medium") -> "MolecularGrid":
        lv = level_from_name(level)
        return cls._build(mol, LEVELS[lv], 0, 0, 0, lv)

    @classmethod
    def build_for_basis(cls, mol: Molecule, basis, level: str = "medium") -> "MolecularGrid":
        lv = level_from_name(level)
        return cls._build(mol, LEVELS[lv], 0, 0, radial_min_for_basis(basis), lv)

    @classmethod
    def explicit(cls, mol: Molecule, n_rad: int, n_ang: int) -> "MolecularGrid":
        if n_ang not in (26, 50, 110, 194, 302, 590):
            raise ValueError(f"no Lebedev rule with {n_ang} points")
        if n_rad <= 0:
            raise ValueError("n_rad must be positive")
        return cls._build(mol, EXPLICIT, n_rad, n_ang, 0, f"{n_rad}x{n_ang}")

    # ---- streaming grids: points generated per (atom, radial range) task on the device ----
    @classmethod
    def build_streaming(cls, mol: Molecule, level: str = "medium") -> "MolecularGrid":
        """MolecularGrid::build_streaming (grid.rs:636-639): no resident points."""
        lv = level_from_name(level)
        return cls(np.zeros((0, 6)), lv, (mol, LEVELS[lv], 0, 0, 0))

    @classmethod
    def build_streaming_for_basis(cls, mol: Molecule, basis, level: str = "medium") -> "MolecularGrid":
        lv = level_from_name(level)
        return cls(np.zeros((0, 6)), lv, (mol, LEVELS[lv], 0, 0, radial_min_for_basis(basis)))

    def is_streaming(self) -> bool:
        return self.streaming is not None and self.points.shape[0] == 0

    def streaming_tasks(self) -> List[Tuple[int, int, int]]:
        """(atom, first radial shell, last radial shell) in blocks of 8 shells (grid.rs:685-704)."""
        if self.streaming is None:
            return []
        mol, lv, n_rad, _, rmin = self.streaming
        L = _native.require_device()
        z = np.ascontiguousarray(mol.z_array(), dtype=np.int32)
        cnt = _native.check(L.calchem_hip_grid_streaming_tasks(len(z), _native.ptr(z), lv, n_rad, rmin, None, 0),
                            "calchem_hip_grid_streaming_tasks (count)")
        out = np.zeros((max(cnt, 1), 3), dtype=np.int32)
        _native.check(L.calchem_hip_grid_streaming_tasks(len(z), _native.ptr(z), lv, n_rad, rmin, _native.ptr(out),
                                                         cnt), "calchem_hip_grid_streaming_tasks")
        return [tuple(int(v) for v in row) for row in out[:cnt]]

    def task_points(self, task: Tuple[int, int, int]) -> np.ndarray:
        """generate_task_points (grid.rs:706-782) on the device: rows as in `points`."""
        mol, lv, n_rad, n_ang, rmin = self.streaming
        L = _native.require_device()
        z = np.ascontiguousarray(mol.z_array(), dtype=np.int32)
        xyz = np.ascontiguousarray(mol.coords_bohr().ravel(), dtype=np.float64)
        a, s0, s1 = task
        cnt = _native.check(L.calchem_hip_grid_task_points(len(z), _native.ptr(z), _native.ptr(xyz), lv, n_rad, n_ang,
                                                           rmin, a, s0, s1, None, 0), "calchem_hip_grid_task_points")
        out = np.zeros((max(cnt, 1), 6))
        got = _native.check(L.calchem_hip_grid_task_points(len(z), _native.ptr(z), _native.ptr(xyz), lv, n_rad, n_ang,
                                                           rmin, a, s0, s1, _native.ptr(out), cnt),
                            "calchem_hip_grid_task_points")
        return out[:got].copy()
