"""Device-resident integrals of one geometry, over the C ABI in include/calchem_hip.h.

Mirrors what the reference SCF driver takes from calchem-int for this path:
compute_overlap / compute_kinetic / compute_nuclear
(crates/calchem-int/src/one_electron.rs:115-168) and the two-electron part of
the Fock matrix (RksFockBuilder, crates/calchem-scf/src/fock_builder.rs:105-250).
J and K come from Cholesky vectors of the repulsion integrals built and kept
on the GPU (threshold `tau`, 1e-10 by default); the reference contracts exact
integrals, so results agree to the decomposition error.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native
from .basis import BasisSet
from .molecule import Molecule


class IntegralEngine:
    def __init__(self, mol: Molecule, basis: BasisSet):
        self._lib = _native.require_device()
        self.n_ao = basis.n_ao
        self.n_shells = len(basis.shells)
        self.handle = 0
        self.rank = 0
        self.tau = None
        shells, prims, comps = basis.ao_records()
        shells, prims, comps = (np.ascontiguousarray(x) for x in (shells, prims, comps))
        atoms = np.ascontiguousarray(np.column_stack([mol.coords_bohr(), mol.z_array().astype(np.float64)]))
        h = _native.new_handle()
        rc = self._lib.calchem_hip_int_create(_native.ptr(shells), len(shells), _native.ptr(prims), len(prims),
                                              _native.ptr(comps), len(comps), self.n_ao, mol.n_atoms,
                                              _native.ptr(atoms), h)
        _native.check(rc, "calchem_hip_int_create")
        self.handle = h

    def close(self):
        if self.handle:
            self._lib.calchem_hip_int_release(self.handle)
            self.handle = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def one_electron(self):
        """(S, T, V): overlap, kinetic and nuclear attraction, n x n."""
        n = self.n_ao
        S, T, V = (np.zeros((n, n)) for _ in range(3))
        rc = self._lib.calchem_hip_int_one_electron(self.handle, _native.ptr(S), _native.ptr(T), _native.ptr(V))
        _native.check(rc, "calchem_hip_int_one_electron")
        return S, T, V

    def eri_diag(self) -> np.ndarray:
        out = np.zeros(self.n_ao * (self.n_ao + 1) // 2)
        _native.check(self._lib.calchem_hip_int_eri_diag(self.handle, _native.ptr(out)), "calchem_hip_int_eri_diag")
        return out

    def eri_columns(self, shell_pairs, n_cols: int) -> np.ndarray:
        """Exact (mu nu|lam sig) columns of the given shell pairs over the packed pairs mu >= nu."""
        a = np.ascontiguousarray([p[0] for p in shell_pairs], dtype=np.int32)
        b = np.ascontiguousarray([p[1] for p in shell_pairs], dtype=np.int32)
        out = np.zeros((n_cols, self.n_ao * (self.n_ao + 1) // 2))
        rc = self._lib.calchem_hip_int_eri_columns(self.handle, len(a), _native.ptr(a), _native.ptr(b),
                                                   _native.ptr(out))
        _native.check(rc, "calchem_hip_int_eri_columns")
        return out

    def eri_full(self, basis: BasisSet) -> np.ndarray:
        """Dense (ij|kl) for small bases (the atomic blocks of the SAD guess)."""
        n = self.n_ao
        pairs = [(i, j) for i in range(len(basis.shells)) for j in range(i + 1)]
        ncol = sum(basis.shells[i].n_functions * basis.shells[j].n_functions for i, j in pairs)
        cols = self.eri_columns(pairs, ncol)
        iu, ju = np.tril_indices(n)
        full = np.zeros((n, n, n, n))
        c = 0
        for i, j in pairs:
            si, sj = basis.shells[i], basis.shells[j]
            for x in range(si.n_functions):
                for y in range(sj.n_functions):
                    k, l = si.ao_offset + x, sj.ao_offset + y
                    col = cols[c]
                    full[iu, ju, k, l] = col
                    full[ju, iu, k, l] = col
                    full[iu, ju, l, k] = col
                    full[ju, iu, l, k] = col
                    c += 1
        return full

    def cholesky(self, tau: float = 1e-10, max_rank: int = 0) -> int:
        r = C.c_int(0)
        rc = self._lib.calchem_hip_int_cholesky(self.handle, float(tau), int(max_rank), C.byref(r))
        _native.check(rc, "calchem_hip_int_cholesky")
        self.rank, self.tau = r.value, tau
        return self.rank

    def jk_factors(self, c_pos: np.ndarray, c_neg=None, want_j=True, want_k=True):
        """J and K of P = c_pos c_pos^T - c_neg c_neg^T (factors n x k)."""
        n = self.n_ao
        cp = np.ascontiguousarray(c_pos, dtype=np.float64).reshape(n, -1)
        cn = None if c_neg is None else np.ascontiguousarray(c_neg, dtype=np.float64).reshape(n, -1)
        J = np.zeros((n, n)) if want_j else None
        K = np.zeros((n, n)) if want_k else None
        rc = self._lib.calchem_hip_int_jk(self.handle, cp.shape[1], _native.ptr(cp),
                                          0 if cn is None else cn.shape[1], None if cn is None else _native.ptr(cn),
                                          None if J is None else _native.ptr(J), None if K is None else _native.ptr(K))
        _native.check(rc, "calchem_hip_int_jk")
        return J, K

    def jk(self, P: np.ndarray, want_j=True, want_k=True):
        """J and K of a symmetric density P, split into its positive and negative parts."""
        w, U = np.linalg.eigh(0.5 * (P + P.T))
        scale = max(float(np.abs(w).max(initial=0.0)), 1e-300)
        pos, neg = w > 1e-14 * scale, w < -1e-14 * scale
        cp = U[:, pos] * np.sqrt(w[pos])
        cn = U[:, neg] * np.sqrt(-w[neg]) if neg.any() else None
        if cp.shape[1] == 0:
            cp = np.zeros((self.n_ao, 1))
        return self.jk_factors(cp, cn, want_j, want_k)
