"""Pulay DIIS, restating crates/calchem-scf/src/diis.rs:174-410 (in-memory path).

The error vectors' dot matrix is normalised by its largest diagonal element
and by its largest entry before the bordered system is solved; a solution is
accepted when its weights sum to 1 within 1e-6 and none exceeds 50 in
magnitude, otherwise the oldest vectors are dropped one by one, and with no
acceptable solution the newest Fock matrix is returned unchanged.
"""
from __future__ import annotations

from typing import List, Optional

import numpy as np


def solve_linear_system(a: np.ndarray, b: np.ndarray) -> Optional[np.ndarray]:
    """Gaussian elimination with partial pivoting; None when a pivot is below
    1e-14 (crates/calchem-tensor/src/lib.rs:426-470)."""
    n = a.shape[0]
    aug = np.column_stack([a.astype(np.float64), b.astype(np.float64)])
    for k in range(n):
        r = k + int(np.argmax(np.abs(aug[k:, k])))
        if not np.isfinite(aug[r, k]) or abs(aug[r, k]) < 1e-14:
            return None
        if r != k:
            aug[[k, r]] = aug[[r, k]]
        aug[k + 1:] -= np.outer(aug[k + 1:, k] / aug[k, k], aug[k])
    x = np.zeros(n)
    for k in range(n - 1, -1, -1):
        x[k] = (aug[k, n] - aug[k, k + 1:n] @ x[k + 1:]) / aug[k, k]
    return x if np.all(np.isfinite(x)) else None


def solve_pulay_coeffs(dots: np.ndarray) -> Optional[np.ndarray]:
    m = dots.shape[0]
    if m < 2:
        return None
    max_diag = float(np.max(np.diag(dots)))
    if not np.isfinite(max_diag) or max_diag <= 1e-30:
        return None
    B = np.zeros((m + 1, m + 1))
    B[:m, :m] = dots / max_diag
    scale = float(np.max(np.abs(B[:m, :m])))
    if not np.isfinite(scale) or scale <= np.finfo(np.float64).tiny:
        return None
    B[:m, :m] /= scale
    B[:m, m] = B[m, :m] = -1.0
    rhs = np.zeros(m + 1)
    rhs[m] = -1.0
    c = solve_linear_system(B, rhs)
    if c is None:
        return None
    w = c[:m]
    if not np.isfinite(w.sum()) or abs(w.sum() - 1.0) > 1e-6 or np.any(~np.isfinite(w)) or np.any(np.abs(w) > 50.0):
        return None
    return w


class DiisState:
    def __init__(self, max_vecs: int = 6):
        self.max_vecs = max(2, int(max_vecs))
        self.focks: List[np.ndarray] = []
        self.errs: List[np.ndarray] = []

    def push_and_extrapolate(self, fock: np.ndarray, err: np.ndarray) -> np.ndarray:
        while len(self.focks) >= self.max_vecs:
            self.focks.pop(0)
            self.errs.pop(0)
        self.focks.append(fock.copy())
        self.errs.append(err.copy())
        if len(self.focks) < 2:
            return fock.copy()
        E = np.array([e.ravel() for e in self.errs])
        dots = E @ E.T
        total = len(self.focks)
        for start in range(total - 1):
            w = solve_pulay_coeffs(dots[start:, start:])
            if w is not None:
                out = np.zeros_like(fock)
                for k, ck in enumerate(w):
                    out += ck * self.focks[start + k]
                return out
        return fock.copy()
