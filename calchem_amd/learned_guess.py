"""Learned SCF starting density for water clusters: an atom-pair network in
place of the SAD guess (sad_guess at crates/calchem-scf/src/rks.rs:221-231,
which stays the fallback for anything the network does not cover).

What the network predicts.  For every pair of atoms (A, B) whose waters' O-O
distance is within CUTOFF, the block P_AB of the converged total density,
expressed in local frames: each water carries an orthonormal frame (origin
O, z along the H-H bisector, x along H1 - H2 made orthogonal to z), and the
block is rotated into the frames of A's and B's waters,
    P_loc = T_A^{-1} P_AB T_B^{-T},
with T the Cartesian AO rotation matrices of the atoms' shells (s: 1, p: the
rotation, d: the 6 x 6 action on normalised Cartesian monomials).  The local
block does not change when the whole cluster is rotated or translated, and
the features (pair type, roles in the waters, the geometry of B's water in
A's frame, their relative orientation, point-charge electric fields and
radial neighbour counts) do not either.  The prediction is rotated back,
assembled into a symmetric P and scaled so that Tr(P S) = 2 n_occ, as the
SAD density is (guess.rs:726-752).

The networks (calchem_amd/csrc/mlp.hip), one per pair type: 80 -> 128 -> 128 -> 225,
SiLU, FP32 weights, trained on the GPU with the project's own forward, backward and Adam
kernels on densities that the FP64 SCF produces for perturbed 8-16-water
clusters (B3LYP/6-31G(d), Medium grid).  Run

    python -m calchem_amd.learned_guess train --n-train 2000 --n-eval 100

to generate the data, train and evaluate (iterations and final energy of SCF
runs from the learned start against SAD starts on held-out 16-water clusters).
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import os
import sys
import time
from typing import List, Optional, Tuple

import numpy as np

from . import _native
from .basis import CART_COMPONENTS, BasisSet, primitive_norm
from .configs import SplitMix64, water_cluster
from .molecule import ANGSTROM_TO_BOHR, Atom, Molecule

D_IN, HIDDEN, D_OUT = 132, 128, 225
D_IN_BASE = 80  # the pair features; D_IN adds the environment of both waters (env_features)
TYPES = ["onsite_O", "onsite_H", "intra_OH", "intra_HH", "inter_OO", "inter_OH", "inter_HO", "inter_HH"]
CUTOFF_ANGSTROM = 7.0          # O-O distance below which the waters' pair blocks are predicted
CHARGES = {8: -0.834, 1: 0.417}  # point charges (e) of the field features
ENV_WIDTHS = (2.5, 4.0, 6.0)     # Angstrom, radial neighbour descriptors
BASIS_NAME = "6-31g(d)"
XC = "b3lyp"
DEFAULT_MODEL = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "learned_guess_b3lyp_631gd.npz")


# ---------------------------------------------------------------------------
# Geometry: clusters, water frames, AO rotation matrices
# ---------------------------------------------------------------------------
def perturbed_cluster(n_waters: int, seed: int, amplitude: float = 0.1) -> Molecule:
    """A water cluster (configs.water_cluster placement) with every Cartesian
    coordinate moved by a seeded uniform offset in [-amplitude, amplitude] Angstrom."""
    mol = water_cluster(n_waters, seed)
    rng = SplitMix64(seed * 7919 + 17)
    atoms = []
    for a in mol.atoms:
        d = np.array([(2.0 * rng.uniform() - 1.0) * amplitude for _ in range(3)]) * ANGSTROM_TO_BOHR
        atoms.append(Atom(a.symbol, a.z, a.position_bohr + d))
    return Molecule(atoms)


def water_topology(mol: Molecule) -> Optional[np.ndarray]:
    """(n_waters, 3) atom indices (O, H1, H2) when the molecule is a list of
    waters in that order with O-H below 1.3 Angstrom, else None."""
    z = mol.z_array()
    if len(z) == 0 or len(z) % 3:
        return None
    W = np.arange(len(z)).reshape(-1, 3)
    R = mol.coords_bohr()
    for o, h1, h2 in W:
        if z[o] != 8 or z[h1] != 1 or z[h2] != 1:
            return None
        if max(np.linalg.norm(R[h1] - R[o]), np.linalg.norm(R[h2] - R[o])) > 1.3 * ANGSTROM_TO_BOHR:
            return None
    return W


def water_frames(R: np.ndarray, W: np.ndarray) -> np.ndarray:
    """(n_waters, 3, 3): columns x, y, z of each water's frame."""
    o, h1, h2 = R[W[:, 0]], R[W[:, 1]], R[W[:, 2]]
    zax = 0.5 * (h1 + h2) - o
    zax /= np.linalg.norm(zax, axis=1, keepdims=True)
    xax = h1 - h2
    xax -= np.sum(xax * zax, axis=1, keepdims=True) * zax
    xax /= np.linalg.norm(xax, axis=1, keepdims=True)
    yax = np.cross(zax, xax)
    return np.stack([xax, yax, zax], axis=2)


def _comp_scale(lmn) -> float:
    L = sum(lmn)
    return primitive_norm(1.0, tuple(lmn)) / primitive_norm(1.0, (L, 0, 0))


def ao_rotation(l: int, R: np.ndarray) -> np.ndarray:
    """T with phi_c(R^T r) = sum_c' T[c', c] phi_c'(r) for the normalised
    Cartesian components of a shell of angular momentum l (calchem order)."""
    comps = CART_COMPONENTS[l]
    n = len(comps)
    if l == 0:
        return np.ones((1, 1))
    index = {c: i for i, c in enumerate(comps)}
    T = np.zeros((n, n))
    for ci, c in enumerate(comps):
        axes = [k for k in range(3) for _ in range(c[k])]
        # (R^T r)_a1 ... (R^T r)_al expanded over global monomials
        terms = {(): 1.0}
        for a in axes:
            nxt = {}
            for mono, coef in terms.items():
                for m in range(3):
                    key = tuple(sorted(mono + (m,)))
                    nxt[key] = nxt.get(key, 0.0) + coef * R[m, a]
            terms = nxt
        for mono, coef in terms.items():
            e = [0, 0, 0]
            for m in mono:
                e[m] += 1
            cp = index[tuple(e)]
            T[cp, ci] += coef * _comp_scale(c) / _comp_scale(tuple(e))
    return T


def atom_rotation(basis: BasisSet, atom: int, R: np.ndarray) -> np.ndarray:
    blocks = [ao_rotation(sh.l, R) for sh in basis.shells if sh.atom_index == atom]
    n = sum(b.shape[0] for b in blocks)
    T = np.zeros((n, n))
    o = 0
    for b in blocks:
        T[o:o + b.shape[0], o:o + b.shape[0]] = b
        o += b.shape[0]
    return T


def atom_aos(basis: BasisSet, atom: int) -> np.ndarray:
    return np.concatenate([np.arange(sh.ao_offset, sh.ao_offset + sh.n_functions)
                           for sh in basis.shells if sh.atom_index == atom])


# ---------------------------------------------------------------------------
# Pairs and features
# ---------------------------------------------------------------------------
def _field(points: np.ndarray, src: np.ndarray, q: np.ndarray) -> np.ndarray:
    d = points[:, None, :] - src[None, :, :]
    r = np.linalg.norm(d, axis=-1)
    r = np.where(r < 1e-6, np.inf, r)
    return np.sum(q[None, :, None] * d / r[..., None] ** 3, axis=1)


def pair_list(mol: Molecule, W: np.ndarray) -> np.ndarray:
    """(n_pairs, 5): A, B (A <= B), type, water of A, water of B."""
    R = mol.coords_bohr()
    O = R[W[:, 0]]
    cut = CUTOFF_ANGSTROM * ANGSTROM_TO_BOHR
    water_of = np.repeat(np.arange(len(W)), 3)
    rows = []
    for i in range(len(W)):
        for j in range(i, len(W)):
            if i != j and np.linalg.norm(O[i] - O[j]) > cut:
                continue
            for a in W[i]:
                for b in W[j]:
                    if i == j and b < a:
                        continue
                    za, zb = mol.atoms[a].z, mol.atoms[b].z
                    if a == b:
                        t = 0 if za == 8 else 1
                    elif i == j:
                        t = 2 if (za == 8 or zb == 8) else 3
                    else:
                        t = {(8, 8): 4, (8, 1): 5, (1, 8): 6, (1, 1): 7}[(za, zb)]
                    rows.append((a, b, t, water_of[a], water_of[b]))
    return np.array(rows, dtype=np.int64).reshape(-1, 5)


def _env_terms(points: np.ndarray, src: np.ndarray, q: np.ndarray):
    """Potential, field and field gradient at points from point charges q at src."""
    d = points[:, None, :] - src[None, :, :]
    r = np.linalg.norm(d, axis=-1)
    r = np.where(r < 1e-6, np.inf, r)
    pot = np.sum(q[None, :] / r, axis=1)
    E = np.sum(q[None, :, None] * d / r[..., None] ** 3, axis=1)
    G = np.einsum("k,pka,pkb->pab", q, d, d * (3.0 / r[..., None] ** 5)) - \
        np.eye(3)[None] * np.sum(q[None, :] / r ** 3, axis=1)[:, None, None]
    return pot, E, G


def env_features(mol: Molecule, W: np.ndarray, pairs: np.ndarray) -> np.ndarray:
    """The environment of the pair's two waters i and j (52 per pair), from the
    point charges of the other waters, in the frame of water i: electric
    fields at the three atoms of i and of j (18), potentials there (6), the
    field gradient at both oxygens (xx, yy, xy, xz, yz: 10), and Gaussian
    neighbour counts by element at three widths around the three atoms of i
    (18)."""
    R = mol.coords_bohr() / ANGSTROM_TO_BOHR  # Angstrom
    F = water_frames(R, W)
    z = mol.z_array()
    q = np.array([CHARGES[int(v)] for v in z])
    water_of = np.repeat(np.arange(len(W)), 3)
    out = np.zeros((len(pairs), D_IN - D_IN_BASE))
    cache = {}
    for k, (_, _, _, i, j) in enumerate(pairs):
        if (i, j) not in cache:
            v = np.zeros(D_IN - D_IN_BASE)
            other = (water_of != i) & (water_of != j)
            if other.any():
                Fi = F[i]
                src, qs = R[other], q[other]
                pot, E, G = _env_terms(np.concatenate([R[W[i]], R[W[j]]]), src, qs)
                v[0:18] = (E @ Fi).ravel()
                v[18:24] = pot
                for s_, p_ in enumerate((0, 3)):
                    Gl = Fi.T @ G[p_] @ Fi
                    v[24 + 5 * s_:29 + 5 * s_] = [Gl[0, 0], Gl[1, 1], Gl[0, 1], Gl[0, 2], Gl[1, 2]]
                zs = z[other]
                for s_, c in enumerate(W[i]):
                    d = np.linalg.norm(src - R[c], axis=1)
                    for e, el in enumerate((8, 1)):
                        for w, sig in enumerate(ENV_WIDTHS):
                            v[34 + 6 * s_ + 3 * e + w] = np.sum(np.exp(-(d[zs == el] / sig) ** 2))
            cache[(i, j)] = v
        out[k] = cache[(i, j)]
    return out


def pair_features(mol: Molecule, W: np.ndarray, pairs: np.ndarray) -> np.ndarray:
    R = mol.coords_bohr() / ANGSTROM_TO_BOHR  # Angstrom
    F = water_frames(R, W)
    z = mol.z_array()
    q = np.array([CHARGES[int(v)] for v in z])
    water_of = np.repeat(np.arange(len(W)), 3)
    role = np.tile(np.arange(3), len(W))
    X = np.zeros((len(pairs), D_IN), dtype=np.float64)
    X[:, D_IN_BASE:] = env_features(mol, W, pairs)
    for k, (a, b, t, i, j) in enumerate(pairs):
        Fi = F[i]
        oi = R[W[i, 0]]
        x = X[k]
        x[t] = 1.0
        x[8 + role[a]] = 1.0
        x[11 + role[b]] = 1.0
        x[14:20] = ((R[W[i, 1:]] - oi) @ Fi).ravel()
        x[20:29] = ((R[W[j]] - oi) @ Fi).ravel()
        x[29:38] = (Fi.T @ F[j]).ravel()
        dab = np.linalg.norm(R[b] - R[a])
        x[38] = dab
        x[39] = math.exp(-dab / 2.0)
        other = (water_of != i) & (water_of != j)
        if other.any():
            src, qs = R[other], q[other]
            pts = np.stack([R[W[i, 0]], R[W[j, 0]], R[a], R[b]])
            E = _field(pts, src, qs) @ Fi
            x[40:43], x[43:46], x[58:61], x[61:64] = E[0], E[1], E[2], E[3]
            for s, c in enumerate((a, b)):
                d = np.linalg.norm(src - R[c], axis=1)
                zs = z[other]
                for e, el in enumerate((8, 1)):
                    for w, sig in enumerate(ENV_WIDTHS):
                        x[46 + 6 * s + 3 * e + w] = np.sum(np.exp(-(d[zs == el] / sig) ** 2))
        # the pair's own geometry: direction of B from A in A's frame, decaying distance functions
        if dab > 1e-8:
            x[64:67] = ((R[b] - R[a]) / dab) @ Fi
        for w, sig in enumerate((0.5, 1.0, 2.0, 4.0)):
            x[67 + w] = math.exp(-dab / sig)
        x[71] = 1.0 / (1.0 + dab)
        doo = np.linalg.norm(R[W[j, 0]] - oi)
        x[72] = doo
        x[73] = math.exp(-doo / 2.0)
    return X


def local_blocks(mol: Molecule, basis: BasisSet, W: np.ndarray, pairs: np.ndarray, P: np.ndarray):
    """Targets: the local-frame blocks (n_pairs, 225) and their masks."""
    R = mol.coords_bohr()
    F = water_frames(R, W)
    water_of = np.repeat(np.arange(len(W)), 3)
    Tinv = [atom_rotation(basis, a, F[water_of[a]].T) for a in range(mol.n_atoms)]  # T(R)^-1 = T(R^T)
    aos = [atom_aos(basis, a) for a in range(mol.n_atoms)]
    Y = np.zeros((len(pairs), D_OUT))
    M = np.zeros((len(pairs), D_OUT))
    for k, (a, b, _, _, _) in enumerate(pairs):
        blk = Tinv[a] @ P[np.ix_(aos[a], aos[b])] @ Tinv[b].T
        Y[k, :blk.size] = blk.ravel()
        M[k, :blk.size] = 1.0
    return Y, M


def assemble_density(mol: Molecule, basis: BasisSet, W: np.ndarray, pairs: np.ndarray, Yloc: np.ndarray) -> np.ndarray:
    R = mol.coords_bohr()
    F = water_frames(R, W)
    water_of = np.repeat(np.arange(len(W)), 3)
    T = [atom_rotation(basis, a, F[water_of[a]]) for a in range(mol.n_atoms)]
    aos = [atom_aos(basis, a) for a in range(mol.n_atoms)]
    P = np.zeros((basis.n_ao, basis.n_ao))
    for k, (a, b, _, _, _) in enumerate(pairs):
        na, nb = len(aos[a]), len(aos[b])
        blk = T[a] @ Yloc[k, :na * nb].reshape(na, nb) @ T[b].T
        if a == b:
            blk = 0.5 * (blk + blk.T)
        P[np.ix_(aos[a], aos[b])] = blk
        if a != b:
            P[np.ix_(aos[b], aos[a])] = blk.T
    return P


# ---------------------------------------------------------------------------
# The network on the device
# ---------------------------------------------------------------------------
def n_params(d_in=D_IN, h=HIDDEN, d_out=D_OUT) -> int:
    return h * d_in + h + h * h + h + d_out * h + d_out


def init_params(seed: int, d_in=D_IN, h=HIDDEN, d_out=D_OUT) -> np.ndarray:
    """Uniform +-sqrt(6 / (fan_in + fan_out)) weights (Glorot), zero biases, SplitMix64 stream."""
    rng = SplitMix64(seed)
    out = []
    for fi, fo in ((d_in, h), (h, h), (h, d_out)):
        lim = math.sqrt(6.0 / (fi + fo))
        out.append(np.array([(2.0 * rng.uniform() - 1.0) * lim for _ in range(fi * fo)]))
        out.append(np.zeros(fo))
    return np.concatenate(out).astype(np.float32)


class DeviceMlp:
    """The network's parameters and training state on the GPU (C ABI in include/calchem_hip.h)."""

    def __init__(self, params: np.ndarray, d_in=D_IN, h=HIDDEN, d_out=D_OUT):
        self._lib = _native.require_device()
        self.d_in, self.h, self.d_out = d_in, h, d_out
        p = np.ascontiguousarray(params, dtype=np.float32)
        if p.size != n_params(d_in, h, d_out):
            raise ValueError("parameter count does not match the layer sizes")
        hd = _native.new_handle()
        _native.check(self._lib.calchem_hip_mlp_create(d_in, h, d_out, _native.ptr(p), hd), "calchem_hip_mlp_create")
        self.handle = hd

    def close(self):
        if self.handle:
            self._lib.calchem_hip_mlp_release(self.handle)
            self.handle = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self) -> np.ndarray:
        out = np.zeros(n_params(self.d_in, self.h, self.d_out), dtype=np.float32)
        _native.check(self._lib.calchem_hip_mlp_get_params(self.handle, _native.ptr(out)), "calchem_hip_mlp_get_params")
        return out

    def forward(self, X: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(X, dtype=np.float32)
        y = np.zeros((x.shape[0], self.d_out), dtype=np.float32)
        _native.check(self._lib.calchem_hip_mlp_forward(self.handle, x.shape[0], _native.ptr(x), _native.ptr(y)),
                      "calchem_hip_mlp_forward")
        return y

    def loss_grad(self, X, T, M):
        x, t, m = (np.ascontiguousarray(a, dtype=np.float32) for a in (X, T, M))
        loss = C.c_float(0.0)
        g = np.zeros(n_params(self.d_in, self.h, self.d_out), dtype=np.float32)
        _native.check(self._lib.calchem_hip_mlp_loss_grad(self.handle, x.shape[0], _native.ptr(x), _native.ptr(t),
                                                          _native.ptr(m), C.byref(loss), _native.ptr(g)),
                      "calchem_hip_mlp_loss_grad")
        return float(loss.value), g

    def train(self, X, T, M, epochs: int, batch: int, lr: float, seed: int) -> np.ndarray:
        x, t, m = (np.ascontiguousarray(a, dtype=np.float32) for a in (X, T, M))
        _native.check(self._lib.calchem_hip_mlp_set_data(self.handle, x.shape[0], _native.ptr(x), _native.ptr(t),
                                                         _native.ptr(m)), "calchem_hip_mlp_set_data")
        hist = np.zeros(epochs, dtype=np.float32)
        _native.check(self._lib.calchem_hip_mlp_train(self.handle, epochs, batch, float(lr), seed, _native.ptr(hist)),
                      "calchem_hip_mlp_train")
        return hist


# ---------------------------------------------------------------------------
# The guess
# ---------------------------------------------------------------------------
class LearnedGuess:
    """A trained model: one network per pair type (params [type][n_params]) and
    the feature / target standardisation."""

    def __init__(self, params, x_mean, x_std, y_mean, y_std, meta=None):
        self.params = np.asarray(params, dtype=np.float32).reshape(len(TYPES), -1)
        self.x_mean, self.x_std = np.asarray(x_mean), np.asarray(x_std)
        # inputs of the networks: D_IN, or the D_IN_BASE pair features alone for
        # a model trained without the environment features
        self.d_in = int(self.x_mean.shape[0])
        if self.d_in not in (D_IN, D_IN_BASE) or self.params.shape[1] != n_params(self.d_in):
            raise ValueError("model layer sizes do not match the feature set")
        self.y_mean, self.y_std = np.asarray(y_mean), np.asarray(y_std)  # [type][225]
        self.meta = meta or {}
        self._nets = None

    @classmethod
    def load(cls, path: str = DEFAULT_MODEL) -> "LearnedGuess":
        z = np.load(path, allow_pickle=False)
        meta = json.loads(str(z["meta"])) if "meta" in z else {}
        return cls(z["params"], z["x_mean"], z["x_std"], z["y_mean"], z["y_std"], meta)

    def save(self, path: str):
        np.savez(path, params=self.params, x_mean=self.x_mean, x_std=self.x_std, y_mean=self.y_mean,
                 y_std=self.y_std, meta=json.dumps(self.meta))

    def nets(self):
        if self._nets is None:
            self._nets = [DeviceMlp(p, d_in=self.d_in) for p in self.params]
        return self._nets

    def close(self):
        for n in self._nets or []:
            n.close()
        self._nets = None

    def predict(self, X: np.ndarray, types: np.ndarray) -> np.ndarray:
        """Local-frame blocks (n, 225) for raw features X of pairs of the given types."""
        Xs = (X[:, :self.d_in] - self.x_mean) / self.x_std
        Y = np.zeros((len(X), D_OUT))
        for t, net in enumerate(self.nets()):
            sel = np.nonzero(types == t)[0]
            if len(sel):
                Y[sel] = net.forward(Xs[sel]).astype(np.float64) * self.y_std[t] + self.y_mean[t]
        return Y

    def covers(self, mol: Molecule, basis: BasisSet) -> bool:
        return basis.name.lower().replace("*", "(d)") in ("6-31g(d)", "6-31gd") and water_topology(mol) is not None

    def density(self, mol: Molecule, basis: BasisSet, S: np.ndarray, n_occ: int,
                purify: int = 2) -> Optional[np.ndarray]:
        """The predicted total density scaled to Tr(P S) = 2 n_occ, or None when
        the model does not cover the molecule / basis (the caller then uses SAD).
        purify > 0: that many McWeeny steps D <- 3 DSD - 2 DSDSD on D = P / 2
        (McWeeny, Rev. Mod. Phys. 32, 335 (1960)) before the trace is restored."""
        if not self.covers(mol, basis):
            return None
        W = water_topology(mol)
        pairs = pair_list(mol, W)
        Y = self.predict(pair_features(mol, W, pairs), pairs[:, 2])
        P = assemble_density(mol, basis, W, pairs, Y)
        for _ in range(purify):
            D = 0.5 * P
            DS = D @ S
            D = 3.0 * DS @ D - 2.0 * DS @ DS @ D
            P = D + D.T  # 2 D, symmetrised
        tr = float(np.sum(P * S))
        if not np.isfinite(tr) or tr <= 1e-12:
            return None
        return P * (2.0 * n_occ / tr)


# ---------------------------------------------------------------------------
# Data, training, evaluation
# ---------------------------------------------------------------------------
def _log(msg):
    print(f"[learned_guess] {msg}", flush=True)


def converged_cluster(mol, grid_level="medium", tau=1e-8):
    """FP64 B3LYP/6-31G(d) SCF from SAD: (basis, solver, result); the solver stays open."""
    from .grid import MolecularGrid
    from .scf import RksSolver, ScfOptions
    basis = BasisSet.from_name(BASIS_NAME, mol)
    grid = MolecularGrid.build_for_basis(mol, basis, grid_level)
    s = RksSolver(mol, basis, XC, grid, ScfOptions(cholesky_tau=tau))
    return basis, s, s.run()


def cluster_plan(n_clusters: int, seed: int, lo: int = 8, hi: int = 16):
    """(index, waters, placement seed) of each cluster: sizes uniform in [lo, hi] from a SplitMix64 stream."""
    rng = SplitMix64(seed)
    return [(k, lo + int(rng.uniform() * (hi - lo + 1)), 100000 * seed + k) for k in range(n_clusters)]


def block_mask(types: np.ndarray) -> np.ndarray:
    """Valid outputs of each pair type (6-31G(d): O 15 AOs, H 2)."""
    size = {0: 225, 1: 4, 2: 30, 3: 4, 4: 225, 5: 30, 6: 30, 7: 4}
    M = np.zeros((len(types), D_OUT), dtype=np.float32)
    for t, n in size.items():
        M[types == t, :n] = 1.0
    return M


def _limit_threads():
    try:
        from threadpoolctl import threadpool_limits
        threadpool_limits(2)
    except Exception:
        pass


def _gen_chunk(plan):
    _limit_threads()
    Xs, Ys, Ts = [], [], []
    t0 = time.time()
    for i, (k, n_w, cseed) in enumerate(plan):
        mol = perturbed_cluster(n_w, cseed)
        basis, s, res = converged_cluster(mol)
        s.close()
        W = water_topology(mol)
        pairs = pair_list(mol, W)
        Xs.append(pair_features(mol, W, pairs).astype(np.float32))
        Y, _ = local_blocks(mol, basis, W, pairs, res.density)
        Ys.append(Y.astype(np.float32))
        Ts.append(pairs[:, 2].astype(np.int8))
        if (i + 1) % 25 == 0:
            _log(f"worker at cluster {k}: {i + 1}/{len(plan)} done, {time.time() - t0:.0f}s")
    return np.vstack(Xs), np.vstack(Ys), np.concatenate(Ts)


def _chunks(items, workers):
    return [items[w::workers] for w in range(workers)]


def _pool(workers):
    import multiprocessing as mp
    return mp.get_context("spawn").Pool(workers)


def generate(n_clusters: int, seed: int, lo: int = 8, hi: int = 16, pool=None, workers: int = 1):
    """Features X, local-frame target blocks Y and pair types of the converged
    densities of n_clusters perturbed clusters, in cluster order whatever the
    number of worker processes (each its own HIP context on the same GPU)."""
    plan = cluster_plan(n_clusters, seed, lo, hi)
    t0 = time.time()
    if pool is None:
        parts = [_gen_chunk(plan)]
    else:
        # contiguous ranges, so that concatenating the workers' parts keeps cluster order
        step = (len(plan) + workers - 1) // workers
        parts = pool.map(_gen_chunk, [plan[i:i + step] for i in range(0, len(plan), step)])
    X = np.vstack([p[0] for p in parts])
    Y = np.vstack([p[1] for p in parts])
    T = np.concatenate([p[2] for p in parts])
    _log(f"data: {n_clusters} clusters, {len(T)} pairs in {time.time() - t0:.0f}s")
    return X, Y, block_mask(T), T


def fit(X, Y, M, types, steps=12000, batch=4096, lr=2e-3, seed=5):
    """One network per pair type, each for about `steps` Adam steps.  Features
    standardised over all pairs; targets minus their per-type mean block over
    one scale per type, so each type's loss weighs absolute density errors."""
    x_mean = X.mean(axis=0, dtype=np.float64)
    x_std = X.std(axis=0, dtype=np.float64)
    x_std[x_std < 1e-6] = 1.0
    Xs = ((X - x_mean) / x_std).astype(np.float32)
    y_mean = np.zeros((len(TYPES), D_OUT))
    y_std = np.ones((len(TYPES), D_OUT))
    params, hists = [], []
    t0 = time.time()
    for t in range(len(TYPES)):
        sel = np.nonzero(types == t)[0]
        p0 = init_params(seed + 101 * t)
        if len(sel) == 0:
            params.append(p0)
            hists.append([])
            continue
        y_mean[t] = Y[sel].mean(axis=0, dtype=np.float64)
        R = (Y[sel] - y_mean[t]) * M[sel]
        y_std[t] = max(float(np.sqrt(np.sum(R.astype(np.float64) ** 2) / M[sel].sum())), 1e-12)
        bsz = int(min(batch, max(256, len(sel) // 8)))
        epochs = int(min(1500, max(40, round(steps * bsz / len(sel)))))
        net = DeviceMlp(p0)
        h = net.train(Xs[sel], (R / y_std[t]).astype(np.float32), M[sel], epochs, bsz, lr, seed + t)
        params.append(net.params())
        net.close()
        hists.append(h.tolist())
        _log(f"type {TYPES[t]}: {len(sel)} pairs, {epochs} epochs of batch {bsz}, loss {h[0]:.3e} -> {h[-1]:.3e} "
             f"(scale {y_std[t][0]:.3e}), {time.time() - t0:.0f}s")
    return LearnedGuess(np.stack(params), x_mean, x_std, y_mean, y_std), hists


def _eval_chunk(args):
    _limit_threads()
    model_path, plan = args
    from .grid import MolecularGrid
    from .scf import RksSolver, ScfOptions
    model = LearnedGuess.load(model_path)
    rows = []
    for k, n_w, cseed in plan:
        mol = perturbed_cluster(n_w, cseed)
        basis = BasisSet.from_name(BASIS_NAME, mol)
        grid = MolecularGrid.build_for_basis(mol, basis, "medium")
        s = RksSolver(mol, basis, XC, grid, ScfOptions(cholesky_tau=1e-8))
        r_sad = s.run()
        t0 = time.perf_counter()
        P0 = model.density(mol, basis, s.S, s.n_occ, purify=0)
        t_inf = time.perf_counter() - t0
        r_ml = s.run(guess_density=P0, guess_name="learned")
        P1 = model.density(mol, basis, s.S, s.n_occ, purify=2)
        r_mp = s.run(guess_density=P1, guess_name="learned+purified")
        P_sad = s.sad()
        s.close()
        rows.append({"cluster": k, "waters": n_w, "iters_sad": r_sad.iterations, "iters_learned": r_ml.iterations,
                     "iters_learned_purified": r_mp.iterations,
                     "e_sad": r_sad.total_energy, "e_learned": r_ml.total_energy,
                     "dE_purified": r_mp.total_energy - r_sad.total_energy,
                     "guess_err_purified": float(np.abs(P1 - r_sad.density).max()),
                     "dE": r_ml.total_energy - r_sad.total_energy,
                     "guess_err_learned": float(np.abs(P0 - r_sad.density).max()),
                     "guess_err_sad": float(np.abs(P_sad - r_sad.density).max()),
                     "guess_rms_learned": float(np.sqrt(np.mean((P0 - r_sad.density) ** 2))),
                     "guess_rms_sad": float(np.sqrt(np.mean((P_sad - r_sad.density) ** 2))),
                     "inference_s": t_inf})
    model.close()
    return rows


def evaluate(model_path: str, n_clusters: int, seed: int, n_waters: int = 16, pool=None, workers: int = 1):
    """SCF from SAD and from the learned density on held-out clusters (both FP64, same solver)."""
    plan = [(k, n_waters, 100000 * seed + k) for k in range(n_clusters)]
    t0 = time.time()
    if pool is None:
        rows = _eval_chunk((model_path, plan))
    else:
        parts = pool.map(_eval_chunk, [(model_path, c) for c in _chunks(plan, workers)])
        rows = sorted([r for p in parts for r in p], key=lambda r: r["cluster"])
    _log(f"eval: {n_clusters} clusters in {time.time() - t0:.0f}s")
    return rows


def diagnose(model_path: str, n_clusters: int, seed: int, n_waters: int = 16) -> dict:
    """Which pair types' errors cost SCF iterations: starts built from the
    learned blocks with the blocks of some pair types replaced by those of
    the converged density (FP64 SCF from SAD on the same solver), each
    purified as the default guess is, against SAD."""
    from .grid import MolecularGrid
    from .scf import RksSolver, ScfOptions
    model = LearnedGuess.load(model_path)
    cases = {"learned": lambda t: np.zeros(len(t), bool), "exact_onsite": lambda t: t < 2,
             "exact_intra": lambda t: t < 4, "exact_inter": lambda t: t >= 4,
             "exact_all": lambda t: np.ones(len(t), bool)}
    its = {"sad": [], **{k: [] for k in cases}}
    for k in range(n_clusters):
        mol = perturbed_cluster(n_waters, 100000 * seed + k)
        basis = BasisSet.from_name(BASIS_NAME, mol)
        grid = MolecularGrid.build_for_basis(mol, basis, "medium")
        s = RksSolver(mol, basis, XC, grid, ScfOptions(cholesky_tau=1e-8))
        r0 = s.run()
        its["sad"].append(r0.iterations)
        W = water_topology(mol)
        pairs = pair_list(mol, W)
        Y = model.predict(pair_features(mol, W, pairs), pairs[:, 2])
        Yx, _ = local_blocks(mol, basis, W, pairs, r0.density)
        for name, sel in cases.items():
            P = assemble_density(mol, basis, W, pairs, np.where(sel(pairs[:, 2])[:, None], Yx, Y))
            for _ in range(2):
                D = 0.5 * P
                DS = D @ s.S
                D = 3.0 * DS @ D - 2.0 * DS @ DS @ D
                P = D + D.T
            P = P * (2.0 * s.n_occ / float(np.sum(P * s.S)))
            its[name].append(s.run(guess_density=P, guess_name=name).iterations)
        s.close()
        _log(f"cluster {k}: " + ", ".join(f"{a} {b[-1]}" for a, b in its.items()))
    model.close()
    return {"clusters": n_clusters, "waters": n_waters, "seed": seed,
            "mean_iterations": {a: float(np.mean(b)) for a, b in its.items()}, "iterations": its}


def main(argv=None):
    ap = argparse.ArgumentParser(description="train and evaluate the learned SCF guess on the GPU")
    sub = ap.add_subparsers(dest="cmd", required=True)
    dg = sub.add_parser("diagnose", help="SCF iterations from learned starts with some pair types' blocks exact")
    dg.add_argument("--model", default=DEFAULT_MODEL)
    dg.add_argument("--clusters", type=int, default=8)
    dg.add_argument("--seed", type=int, default=1004)
    dg.add_argument("--report", default=None)
    tr = sub.add_parser("train")
    tr.add_argument("--n-train", type=int, default=2000)
    tr.add_argument("--n-eval", type=int, default=100)
    tr.add_argument("--steps", type=int, default=12000, help="Adam steps per pair-type network")
    tr.add_argument("--batch", type=int, default=4096)
    tr.add_argument("--lr", type=float, default=2e-3)
    tr.add_argument("--seed", type=int, default=4)
    tr.add_argument("--workers", type=int, default=6)
    tr.add_argument("--out", default=DEFAULT_MODEL)
    tr.add_argument("--report", default=None)
    a = ap.parse_args(argv)
    if a.cmd == "diagnose":
        rep = diagnose(a.model, a.clusters, a.seed)
        _log("mean iterations: " + ", ".join(f"{k} {v:.3f}" for k, v in rep["mean_iterations"].items()))
        if a.report:
            with open(a.report, "w") as fh:
                json.dump(rep, fh, indent=1)
        return 0
    t0 = time.time()
    # the worker processes start before this process touches the GPU and serve
    # both the data generation and the evaluation
    pool = _pool(a.workers) if a.workers > 1 else None
    X, Y, M, types = generate(a.n_train, a.seed, pool=pool, workers=a.workers)
    model, hist = fit(X, Y, M, types, a.steps, a.batch, a.lr, a.seed + 1)
    model.meta = {"basis": BASIS_NAME, "xc": XC, "n_train_clusters": a.n_train, "n_pairs": int(len(X)),
                  "steps_per_type": a.steps, "batch": a.batch, "lr": a.lr, "seed": a.seed, "loss_history": hist,
                  "cutoff_angstrom": CUTOFF_ANGSTROM}
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    model.save(a.out)
    _log(f"saved {a.out}")
    rows = evaluate(a.out, a.n_eval, a.seed + 1000, pool=pool, workers=a.workers)
    if pool is not None:
        pool.close()
        pool.join()
    it_s = np.array([r["iters_sad"] for r in rows])
    it_l = np.array([r["iters_learned"] for r in rows])
    it_p = np.array([r["iters_learned_purified"] for r in rows])
    report = {"mean_iters_learned_purified": float(it_p.mean()),
              "iteration_reduction_purified": float(1.0 - it_p.mean() / it_s.mean()),"n_train_clusters": a.n_train, "n_eval_clusters": a.n_eval, "pairs": int(len(X)),
              "mean_iters_sad": float(it_s.mean()), "mean_iters_learned": float(it_l.mean()),
              "iteration_reduction": float(1.0 - it_l.mean() / it_s.mean()),
              "max_abs_dE": float(max(max(abs(r["dE"]), abs(r["dE_purified"])) for r in rows)),
              "mean_guess_err_sad": float(np.mean([r["guess_err_sad"] for r in rows])),
              "mean_guess_err_learned": float(np.mean([r["guess_err_learned"] for r in rows])),
              "loss_history": hist, "seconds": time.time() - t0, "clusters": rows}
    _log(f"SAD {it_s.mean():.2f} vs learned {it_l.mean():.2f} iterations (purified {it_p.mean():.2f}) "
         f"({100 * report['iteration_reduction']:.1f}% fewer), max |dE| {report['max_abs_dE']:.2e}, "
         f"max |dP| of the guess {report['mean_guess_err_sad']:.3f} (SAD) vs {report['mean_guess_err_learned']:.3f}")
    if a.report:
        with open(a.report, "w") as fh:
            json.dump(report, fh, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
