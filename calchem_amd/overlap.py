"""Analytic overlap matrix S of a Cartesian Gaussian basis.

Only used to build S-orthonormal test densities (P = 2 C C^T with C^T S C = 1,
the Löwdin projection of crates/calchem-scf/src/guess.rs:146-158) and to check
N = Tr(P S).  The 1D factors use the closed binomial form of the Gaussian
product theorem (Helgaker, Jorgensen, Olsen, ch. 9):
    S_x(i, j) = sqrt(pi/p) sum_{a<=i, b<=j, a+b even} C(i,a) C(j,b)
                PA^(i-a) PB^(j-b) (a+b-1)!! / (2p)^((a+b)/2),
evaluated for all primitive pairs at once.
"""
from __future__ import annotations

import math

import numpy as np

from .basis import BasisSet


def _dfact_odd(k: int) -> float:
    acc = 1.0
    while k > 0:
        acc *= k
        k -= 2
    return acc


def _s1d(i: np.ndarray, j: np.ndarray, PA: np.ndarray, PB: np.ndarray, p: np.ndarray) -> np.ndarray:
    out = np.zeros_like(p)
    lmax = int(max(i.max(initial=0), j.max(initial=0)))
    for a in range(lmax + 1):
        for b in range(lmax + 1):
            if (a + b) % 2:
                continue
            mask = (a <= i) & (b <= j)
            if not mask.any():
                continue
            ci = np.array([math.comb(int(x), a) if a <= x else 0 for x in range(lmax + 1)])[i]
            cj = np.array([math.comb(int(x), b) if b <= x else 0 for x in range(lmax + 1)])[j]
            term = (ci * cj * _dfact_odd(a + b - 1) * PA ** np.maximum(i - a, 0) * PB ** np.maximum(j - b, 0)
                    / (2.0 * p) ** ((a + b) // 2))
            out += np.where(mask, term, 0.0)
    return out * np.sqrt(np.pi / p)


def primitive_expansion(basis: BasisSet):
    """Per primitive Cartesian function: (ao, centre, exponent, coefficient, lmn)."""
    shells, prims, comps = basis.ao_records()
    ao, cen, alp, coef, lmn = [], [], [], [], []
    for sh in shells:
        pr = prims[sh["prim_start"]:sh["prim_start"] + sh["n_prims"]]
        for c in range(sh["n_funcs"]):
            cp = comps[sh["comp_start"] + c]
            for q in pr:
                ao.append(sh["ao_offset"] + c)
                cen.append(sh["center"])
                alp.append(q["exponent"])
                coef.append(q["weighted_norm"] * cp["scale"])
                lmn.append((cp["lx"], cp["ly"], cp["lz"]))
    return (np.array(ao), np.array(cen).reshape(-1, 3), np.array(alp), np.array(coef),
            np.array(lmn, dtype=np.int64).reshape(-1, 3))


def overlap_matrix(basis: BasisSet) -> np.ndarray:
    ao, cen, alp, coef, lmn = primitive_expansion(basis)
    n = basis.n_ao
    S = np.zeros((n, n))
    # Rows in chunks to bound memory (pairs are np x np).
    npr = len(ao)
    step = max(1, 4_000_000 // max(npr, 1))
    for r0 in range(0, npr, step):
        r1 = min(npr, r0 + step)
        a = alp[r0:r1, None]
        b = alp[None, :]
        p = a + b
        AB = cen[r0:r1, None, :] - cen[None, :, :]
        val = np.exp(-a * b / p * np.sum(AB * AB, axis=-1)) * coef[r0:r1, None] * coef[None, :]
        for q in range(3):
            Pq = (a * cen[r0:r1, None, q] + b * cen[None, :, q]) / p
            val = val * _s1d(np.broadcast_to(lmn[r0:r1, None, q], p.shape),
                             np.broadcast_to(lmn[None, :, q], p.shape),
                             Pq - cen[r0:r1, None, q], Pq - cen[None, :, q], p)
        blk = np.zeros((r1 - r0, n))
        np.add.at(blk, (np.repeat(np.arange(r1 - r0), npr), np.tile(ao, r1 - r0)), val.ravel())
        np.add.at(S, ao[r0:r1], blk)
    return S
