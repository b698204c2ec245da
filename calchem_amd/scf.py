"""Closed-shell SCF (RHF and RKS) around the device Vxc build and J / K,
restating crates/calchem-scf/src/rks.rs:131-615 (with rhf.rs as the
exchange-only case):

* h = T + V_ne; G = J - (c_x / 2) K; F = h + G + V_xc;
  E = sum P (h + G / 2) + E_xc + E_nuc (rks.rs:14-44, 253-280);
* orthogonalisation X = S^{-1/2} (eigenvalues below 1e-10 refused), guess:
  projected orbitals, else SAD, else the core Hamiltonian (rks.rs:205-231);
* DIIS on the commutator FPS - SPF (skipped on the first SAD iteration and
  used from its second vector), level shift 0.2 for n >= 40 while the error
  exceeds 0.08, damping 0.3 when the error grows above 0.15
  (rks.rs:282-349);
* convergence on |dE|, the RMS density change and the commutator, with the
  tight-plateau rule (rks.rs:357-377), then the fixed-point polish with FP64
  builds (rks.rs:392-578) and the final commutator check (rks.rs:580-600).

Mixed precision (precision="mixed"): the Vxc builds run on the mixed
(FP32-contraction) kernels until the DIIS error falls below `mixed_switch`
(1e-5), and on the FP64 kernels from then on, polish included.  A learned
initial density (calchem_amd.learned_guess) enters as `guess_density`.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import xc as xcmod
from .basis import BasisSet
from .diis import DiisState
from .grid import MolecularGrid
from .guess import project_occupied_density, sad_guess
from .integrals import IntegralEngine
from .molecule import Molecule


class ScfNotConverged(RuntimeError):
    pass


@dataclass
class ScfOptions:
    """calchem_model::ScfOptions defaults (crates/calchem-model/src/lib.rs:262-300) and
    this build's own: Cholesky threshold of the repulsion integrals, Vxc precision."""
    max_iter: int = 100
    conv_energy: float = 1e-8
    conv_density: float = 1e-8
    use_diis: bool = True
    diis_max_vecs: int = 6
    damping: float = 0.0
    cholesky_tau: float = 1e-10
    precision: str = "fp64"      # "fp64" or "mixed"
    mixed_switch: float = 1e-5   # DIIS error below which a mixed run switches to FP64 builds


@dataclass
class ScfResult:
    converged: bool
    iterations: int
    total_energy: float
    electronic_energy: float
    nuclear_repulsion: float
    e_xc: float
    mo_energies: np.ndarray
    mo_coefficients: np.ndarray
    density: np.ndarray          # total P = 2 C_occ C_occ^T
    fock: np.ndarray
    overlap: np.ndarray
    guess: str
    energy_history: List[float] = field(default_factory=list)
    rms_density_history: List[float] = field(default_factory=list)
    diis_error_history: List[float] = field(default_factory=list)
    precision_history: List[str] = field(default_factory=list)
    timings: dict = field(default_factory=dict)


def nuclear_repulsion_energy(mol: Molecule) -> float:
    R = mol.coords_bohr()
    z = mol.z_array().astype(np.float64)
    d = np.linalg.norm(R[:, None, :] - R[None, :, :], axis=-1)
    iu = np.triu_indices(len(z), 1)
    return float(np.sum(z[iu[0]] * z[iu[1]] / d[iu]))


def _rms_diff(a, b) -> float:
    return float(np.sqrt(np.mean((a - b) ** 2)))


def _diagonalize(F, X, n_occ, cp_prev=None, level_shift=0.0):
    """diagonalize_fock_with_shift (rks.rs:46-93): F' = X^T F X, shifted virtuals."""
    fp = X.T @ F @ X
    if level_shift > 0.0 and cp_prev is not None:
        co = cp_prev[:, :n_occ]
        fp = fp + level_shift * (np.eye(fp.shape[0]) - co @ co.T)
    eps, cp = np.linalg.eigh(fp)
    c = X @ cp
    co = c[:, :n_occ]
    return eps, c, 2.0 * co @ co.T, cp


class RksSolver:
    """One geometry: integrals, Cholesky vectors and the Vxc plan stay resident
    for every SCF run on it (several guesses, precisions)."""

    def __init__(self, mol: Molecule, basis: BasisSet, xc: Optional[str], grid: Optional[MolecularGrid] = None,
                 opts: Optional[ScfOptions] = None):
        self.mol, self.basis = mol, basis
        self.xc = None if xc is None else xcmod.canonical(xc)
        self.opts = opts or ScfOptions()
        self.cx = 1.0 if self.xc is None else xcmod.EXACT_EXCHANGE[self.xc]
        n_el = mol.n_electrons()
        if n_el % 2 or mol.multiplicity != 1:
            raise ValueError("RKS requires a closed-shell singlet state (n_alpha == n_beta)")
        self.n_occ = n_el // 2
        self.n = basis.n_ao
        if self.n == 0 or self.n_occ > self.n:
            raise ValueError(f"Insufficient AO basis functions ({self.n}) for {self.n_occ} occupied orbitals")
        t0 = time.perf_counter()
        self.ints = IntegralEngine(mol, basis)
        S, T, V = self.ints.one_electron()
        self.S, self.h = S, T + V
        t1 = time.perf_counter()
        self.rank = self.ints.cholesky(self.opts.cholesky_tau)
        t2 = time.perf_counter()
        self.vxc = None
        if self.xc is not None:
            from .vxc import RksVxcEngine
            if grid is None:
                grid = MolecularGrid.build_for_basis(mol, basis, "medium")
            self.grid = grid
            self.vxc = RksVxcEngine(basis, grid)
        t3 = time.perf_counter()
        self.e_nuc = nuclear_repulsion_energy(mol)
        w, U = np.linalg.eigh(S)
        if np.any(w <= 1e-10):
            raise np.linalg.LinAlgError(
                f"Overlap matrix is singular or too ill-conditioned for full-rank SCF (eigenvalue {w.min():.3e})")
        self.X = U @ np.diag(w ** -0.5) @ U.T
        self.setup_s = {"one_electron": t1 - t0, "cholesky": t2 - t1, "vxc_plan": t3 - t2}

    def close(self):
        self.ints.close()
        if self.vxc is not None:
            self.vxc.close()

    # ------------------------------------------------------------------ #
    def sad(self) -> np.ndarray:
        return sad_guess(self.mol, self.basis, self.S, self.n_occ, IntegralEngine)

    def fock(self, P, precision="fp64", C_occ=None):
        """(F, E_total, E_xc, G) at density P; C_occ (P = 2 C C^T) spares the factorisation."""
        if C_occ is not None:
            J, K = self.ints.jk_factors(np.sqrt(2.0) * C_occ, None, True, self.cx != 0.0)
        else:
            J, K = self.ints.jk(P, True, self.cx != 0.0)
        G = J if self.cx == 0.0 else J - 0.5 * self.cx * K
        F = self.h + G
        e = float(np.sum(P * (self.h + 0.5 * G))) + self.e_nuc
        e_xc = 0.0
        if self.vxc is not None:
            e_xc, Vxc, _ = self.vxc.evaluate_full(self.xc, P, precision)
            F = F + Vxc
            e += e_xc
        return F, e, e_xc

    def _comm(self, F, P):
        return F @ P @ self.S - self.S @ P @ F

    def run(self, guess_density: Optional[np.ndarray] = None, guess_orbitals: Optional[np.ndarray] = None,
            guess_name: Optional[str] = None, superposition_guess: bool = True) -> ScfResult:
        """One SCF.  A density guess (learned) replaces SAD; like the SAD density
        it is not idempotent, so with superposition_guess its first Fock matrix
        stays out of DIIS and level shifting, as SAD's does (rks.rs:296-349)."""
        o = self.opts
        n, n_occ, X = self.n, self.n_occ, self.X
        t_start = time.perf_counter()
        cp_prev = None
        used_sad = False
        P = None
        if guess_orbitals is not None:
            P = project_occupied_density(guess_orbitals, n_occ, self.S, 2.0)
            gname = "projected"
        if P is None and guess_density is not None:
            P = np.array(guess_density, dtype=np.float64)
            gname = guess_name or "density"
            used_sad = superposition_guess
        if P is None:
            try:
                P = self.sad()
                used_sad, gname = True, "sad"
            except ValueError:
                _, _, P, cp_prev = _diagonalize(self.h, X, n_occ)
                gname = "core"
        mixed = o.precision == "mixed"
        switched = not mixed
        diis = DiisState(o.diis_max_vecs)
        diis_pushed = 0
        hist_e, hist_d, hist_err, hist_prec = [], [], [], []
        e_prev, converged, used = 0.0, False, 0
        last_de, last_rms, prev_err = np.inf, np.inf, np.inf
        C_occ = None
        t_2e = t_diag = 0.0
        for it in range(1, o.max_iter + 1):
            used = it
            prec = "fp64" if switched else "fp32"
            hist_prec.append(prec)
            t0 = time.perf_counter()
            F, e_tot, _ = self.fock(P, prec, C_occ)
            t_2e += time.perf_counter() - t0
            hist_e.append(e_tot)
            err = self._comm(F, P)
            diis_err = float(np.linalg.norm(err))
            hist_err.append(diis_err)
            if not switched and diis_err < o.mixed_switch:
                switched = True
            skip = used_sad and it == 1
            f_ex = F
            if o.use_diis and not skip:
                diis_pushed += 1
                ex = diis.push_and_extrapolate(F, err)
                f_ex = ex if diis_pushed >= 2 else F
            shift = 0.2 if (n >= 40 and not skip and diis_err > 0.08) else 0.0
            t0 = time.perf_counter()
            _, c_new, p_new, cp_new = _diagonalize(f_ex, X, n_occ, cp_prev, shift)
            t_diag += time.perf_counter() - t0
            cp_prev = cp_new
            if 0.0 < o.damping < 1.0:
                damp = o.damping
            elif not skip and diis_err > 0.15 and diis_err > prev_err * 1.2:
                damp = 0.3
            else:
                damp = 0.0
            if not skip:
                prev_err = diis_err
            if damp > 0.0:
                p_new = (1.0 - damp) * p_new + damp * P
                C_occ = None
            else:
                C_occ = c_new[:, :n_occ]
            delta_e = np.inf if it == 1 else abs(e_tot - e_prev)
            rms_d = _rms_diff(p_new, P)
            hist_d.append(rms_d)
            last_de, last_rms = delta_e, rms_d
            P, e_prev = p_new, e_tot
            grad_tol = max(o.conv_density * 100.0, 1e-6)
            e_plateau = max(1e-11 * max(abs(e_tot) / 100.0, 1.0), o.conv_energy * 0.1)
            d_plateau = min(max(o.conv_density * 0.5, 1e-10), max(o.conv_density, 1e-8))
            tight = it > 8 and delta_e < e_plateau and rms_d < d_plateau and diis_err < grad_tol
            met = it > 1 and ((delta_e < o.conv_energy and rms_d < o.conv_density and diis_err < grad_tol) or tight)
            # a mixed run only stops on an iteration built with the FP64 kernels
            if met and prec == "fp64":
                converged = True
                break
        if not converged:
            raise ScfNotConverged(f"SCF not converged in {used} iterations (dE={last_de:.2e}, rmsD={last_rms:.2e})")

        # fixed-point polish with FP64 builds (rks.rs:392-578)
        fp_d_tol = o.conv_density
        fp_c_tol = max(o.conv_density * 100.0, 1e-6)
        final_p = P
        final_f, e_final, e_xc = self.fock(final_p, "fp64", C_occ)
        comm0 = self._comm(final_f, final_p)
        eps_c, c_c, p_c, _ = _diagonalize(final_f, X, n_occ)
        final_eps, final_c = eps_c, c_c
        if not (np.linalg.norm(comm0) <= fp_c_tol and _rms_diff(p_c, final_p) < fp_d_tol):
            fp_diis = DiisState(max(o.diis_max_vecs, 6)) if o.use_diis else None
            if fp_diis is not None:
                fp_diis.push_and_extrapolate(final_f, comm0)
            final_p = p_c
            final_f, e_final, e_xc = self.fock(final_p, "fp64", c_c[:, :n_occ])
            for step in range(1, 50):
                comm = self._comm(final_f, final_p)
                if np.linalg.norm(comm) <= fp_c_tol and step == 1:
                    eps_x, c_x, p_x, _ = _diagonalize(final_f, X, n_occ)
                    if _rms_diff(p_x, final_p) < fp_d_tol:
                        final_eps, final_c = eps_x, c_x
                        break
                if fp_diis is not None:
                    ex = fp_diis.push_and_extrapolate(final_f, comm)
                    f_step = ex if step >= 2 else final_f
                else:
                    f_step = final_f
                eps_n, c_n, p_n, _ = _diagonalize(f_step, X, n_occ)
                if not o.use_diis:
                    damp = o.damping if 0.0 < o.damping < 1.0 else 0.2
                    p_n = (1.0 - damp) * p_n + damp * final_p
                diff = _rms_diff(p_n, final_p)
                final_p, final_eps, final_c = p_n, eps_n, c_n
                final_f, e_final, e_xc = self.fock(final_p, "fp64", None if not o.use_diis else c_n[:, :n_occ])
                comm2 = self._comm(final_f, final_p)
                if diff < fp_d_tol and np.linalg.norm(comm2) < fp_c_tol:
                    if step < 2:
                        break
                    eps_x, c_x, p_x, _ = _diagonalize(final_f, X, n_occ)
                    final_eps, final_c, final_p = eps_x, c_x, p_x
                    final_f, e_final, e_xc = self.fock(final_p, "fp64", c_x[:, :n_occ])
                    if np.linalg.norm(self._comm(final_f, final_p)) <= fp_c_tol:
                        break
        cn = float(np.linalg.norm(self._comm(final_f, final_p)))
        if not np.isfinite(cn) or cn > fp_c_tol:
            raise ScfNotConverged(f"SCF polish failed: commutator {cn:.2e}")
        t_total = time.perf_counter() - t_start
        return ScfResult(True, used, e_final, e_final - self.e_nuc, self.e_nuc, e_xc, final_eps, final_c, final_p,
                         final_f, self.S, gname, hist_e, hist_d, hist_err, hist_prec,
                         {"scf_total": t_total, "scf_2e_vxc": t_2e, "scf_diag": t_diag, **self.setup_s})


def rks(mol: Molecule, basis: BasisSet, xc: str, grid: Optional[MolecularGrid] = None,
        opts: Optional[ScfOptions] = None, guess_density=None, guess_orbitals=None) -> ScfResult:
    """rks / rks_with_grid / rks_with_grid_and_guess (rks.rs:96-129)."""
    s = RksSolver(mol, basis, xc, grid, opts)
    try:
        return s.run(guess_density, guess_orbitals)
    finally:
        s.close()


def rhf(mol: Molecule, basis: BasisSet, opts: Optional[ScfOptions] = None) -> ScfResult:
    """Closed-shell Hartree-Fock: the same driver with full exchange and no grid (rhf.rs)."""
    s = RksSolver(mol, basis, None, None, opts)
    try:
        return s.run()
    finally:
        s.close()
