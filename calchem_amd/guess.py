"""Initial densities for the SCF, restating crates/calchem-scf/src/guess.rs.

* project_occupied_density (guess.rs:116-186): occupied orbitals of a nearby
  geometry, Loewdin-orthonormalised in the new overlap metric.
* sad_guess (guess.rs:374-770): superposition of atomic densities, each from a
  spherically averaged atomic UHF in the atom's own basis subset, placed
  block-diagonally and scaled so that Tr(P S) = 2 n_occ.  The atomic
  integrals come from the same device integral engine as the molecule's.
"""
from __future__ import annotations

from typing import Dict, Optional, Tuple

import numpy as np

from .basis import BasisSet, Shell
from .diis import DiisState
from .molecule import Atom, Molecule


def project_occupied_density(c_old: np.ndarray, n_occ: int, s_new: np.ndarray, scale: float = 2.0):
    n = s_new.shape[0]
    if n_occ <= 0 or n_occ > n or c_old.shape[0] != n or not np.all(np.isfinite(c_old)):
        return None
    c = c_old[:, :n_occ]
    m = c.T @ s_new @ c
    w, U = np.linalg.eigh(m)
    if np.any(~np.isfinite(w)) or np.any(w <= 1e-8):
        return None
    cp = c @ (U @ np.diag(w ** -0.5) @ U.T)
    P = scale * cp @ cp.T
    if not np.all(np.isfinite(P)) or abs(np.sum(P * s_new) - scale * n_occ) > 1e-6:
        return None
    return P


def atomic_valence_spin_counts(z: int, n_core: int = 0) -> Optional[Tuple[int, int]]:
    """Hund's-rule (n_alpha, n_beta) of the neutral atom (guess.rs:323-372)."""
    z_eff = z - n_core
    if z_eff <= 0:
        return None
    fill = [(1, 2), (1, 2), (3, 6), (1, 2), (3, 6), (1, 2), (5, 10), (3, 6), (1, 2), (5, 10), (3, 6), (1, 2),
            (7, 14), (5, 10), (3, 6), (1, 2), (7, 14), (5, 10), (3, 6)]
    rem, unpaired = z, 0
    for n_orb, cap in fill:
        if rem == 0:
            break
        take = min(rem, cap)
        if take < cap:
            alpha = min(take, n_orb)
            unpaired += alpha - (take - alpha)
        rem -= take
    if unpaired > z_eff or (z_eff - unpaired) % 2:
        unpaired = z_eff % 2
    nb = (z_eff - unpaired) // 2
    return nb + unpaired, nb


def spherical_density_from_fock(F, X, n_elec):
    """Occupy degenerate clusters (within 1e-5 Eh) fractionally (guess.rs:374-422)."""
    m = F.shape[0]
    if n_elec == 0:
        return np.zeros((m, m))
    eps, cp = np.linalg.eigh(X.T @ F @ X)
    c = X @ cp
    P = np.zeros((m, m))
    remaining = float(n_elec)
    start = 0
    while start < m and remaining > 1e-12:
        end = start + 1
        while end < m and abs(eps[end] - eps[start]) < 1e-5:
            end += 1
        deg = end - start
        take = min(remaining, float(deg))
        cc = c[:, start:end]
        P += (take / deg) * cc @ cc.T
        remaining -= take
        start = end
    return None if remaining > 1e-8 else P


def _rms(a, b):
    return float(np.sqrt(np.mean((a - b) ** 2)))


def solve_atomic_sad_block(symbol: str, shells, engine_factory) -> Optional[np.ndarray]:
    """Spherically averaged atomic UHF in the atom's basis subset (guess.rs:424-604)."""
    atom = Atom.from_bohr(symbol, (0.0, 0.0, 0.0))
    counts = atomic_valence_spin_counts(atom.z)
    if counts is None or not shells:
        return None
    na, nb = counts
    local = []
    ao = 0
    for sh in shells:
        local.append(Shell(0, np.zeros(3), sh.l, sh.exponents, sh.coeffs, sh.raw_coeffs, sh.norms, ao))
        ao += sh.n_functions
    m = ao
    if m == 0 or m < na or m < nb:
        return None
    basis = BasisSet("atomic_sad", local, m)
    mol = Molecule([atom])
    eng = engine_factory(mol, basis)
    try:
        S, T, V = eng.one_electron()
        if m == 1:
            return None if S[0, 0] <= 1e-12 else np.array([[(na + nb) / S[0, 0]]])
        h = T + V
        w, U = np.linalg.eigh(S)
        if np.any(w <= 1e-10):
            return None
        X = U @ np.diag(w ** -0.5) @ U.T
        pa = spherical_density_from_fock(h, X, na)
        pb = spherical_density_from_fock(h, X, nb)
        if pa is None or pb is None:
            return None
        eri = eng.eri_full(basis)
    finally:
        eng.close()
    diis = DiisState(6)
    e_prev = 0.0
    for it in range(1, 61):
        pt = pa + pb
        J = np.einsum("ijkl,kl->ij", eri, pt)
        fa = h + J - np.einsum("ikjl,kl->ij", eri, pa)
        fb = h + J - (np.einsum("ikjl,kl->ij", eri, pb) if nb > 0 else 0.0)
        e_elec = 0.5 * float(np.sum(pt * h + pa * fa + pb * fb))
        err = np.vstack([fa @ pa @ S - S @ pa @ fa, fb @ pb @ S - S @ pb @ fb])
        err_norm = float(np.linalg.norm(err))
        if it >= 2:
            ex = diis.push_and_extrapolate(np.vstack([fa, fb]), err)
            fa_s, fb_s = ex[:m], ex[m:]
        else:
            fa_s, fb_s = fa, fb
        pa_new = spherical_density_from_fock(fa_s, X, na)
        pb_new = spherical_density_from_fock(fb_s, X, nb)
        if pa_new is None or pb_new is None:
            return None
        if it > 10 or (it > 2 and err_norm > 0.5):
            pa_new = 0.8 * pa_new + 0.2 * pa
            pb_new = 0.8 * pb_new + 0.2 * pb
        delta_e = np.inf if it == 1 else abs(e_elec - e_prev)
        rms_d = max(_rms(pa_new, pa), _rms(pb_new, pb))
        pa, pb, e_prev = pa_new, pb_new, e_elec
        if it > 1 and delta_e < 1e-8 and rms_d < 1e-6:
            P = pa + pb
            return P if np.all(np.isfinite(P)) else None
    return None


_SAD_CACHE: Dict[tuple, Optional[np.ndarray]] = {}


def _shell_key(shells):
    return tuple((sh.l, tuple(np.round(sh.exponents, 14)), tuple(np.round(sh.raw_coeffs, 14))) for sh in shells)


def _aufbau_fallback(z: int, shells) -> np.ndarray:
    """Diagonal occupations when the atomic solve fails (guess.rs:188-321,628-680)."""
    order = [(0, 0, 2), (0, 1, 2), (1, 0, 6), (0, 2, 2), (1, 1, 6), (0, 3, 2), (2, 0, 10), (1, 2, 6), (0, 4, 2)]
    occ = {0: [0.0] * 7, 1: [0.0] * 6, 2: [0.0] * 4, 3: [0.0] * 2}
    rem = float(z)
    for t, idx, cap in order:
        if rem <= 0:
            break
        take = min(rem, cap)
        occ[t][idx] = take
        rem -= take
    m = sum(sh.n_functions for sh in shells)
    P = np.zeros((m, m))
    off = 0
    by_l: Dict[int, list] = {}
    offs = []
    for sh in shells:
        offs.append(off)
        by_l.setdefault(sh.l, []).append(len(offs) - 1)
        off += sh.n_functions
    for l, idxs in by_l.items():
        active = [v for v in occ.get(l, []) if v > 0]
        alloc = [0.0] * len(idxs)
        if active:
            if len(idxs) == len(active):
                alloc = active
            elif len(idxs) > len(active):
                alloc[:len(active) - 1] = active[:-1]
                per = active[-1] / (len(idxs) - len(active) + 1)
                for i in range(len(active) - 1, len(idxs)):
                    alloc[i] = per
            else:
                alloc = active[len(active) - len(idxs):]
        for k, si in enumerate(idxs):
            nf = shells[si].n_functions
            for c in range(nf):
                P[offs[si] + c, offs[si] + c] = alloc[k] / nf
    return P


def sad_guess(mol: Molecule, basis: BasisSet, S: np.ndarray, n_occ: int, engine_factory) -> np.ndarray:
    """SAD total density scaled to Tr(P S) = 2 n_occ (guess.rs:682-770)."""
    if n_occ <= 0:
        raise ValueError("RHF/RKS SAD guess requires at least one occupied orbital")
    n = basis.n_ao
    P = np.zeros((n, n))
    for ai, atom in enumerate(mol.atoms):
        shells = [sh for sh in basis.shells if sh.atom_index == ai]
        key = (atom.symbol, _shell_key(shells))
        if key not in _SAD_CACHE:
            _SAD_CACHE[key] = solve_atomic_sad_block(atom.symbol, shells, engine_factory)
        block = _SAD_CACHE[key]
        if block is None:
            block = _aufbau_fallback(atom.z, shells)
        idx = np.concatenate([np.arange(sh.ao_offset, sh.ao_offset + sh.n_functions) for sh in shells])
        P[np.ix_(idx, idx)] = block
    tr = float(np.sum(P * S))
    if tr <= 1e-12:
        raise ValueError("SAD guess trace is degenerate or zero")
    return P * (2.0 * n_occ / tr)
