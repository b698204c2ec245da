"""Analytic overlap matrix S of a Cartesian Gaussian basis.

Only used to build S-orthonormal test densities (P = 2 C C^T with C^T S C = 1,
the Löwdin projection of crates/calchem-scf/src/guess.rs:146-158) and to check
N = Tr(P S).  The 1D factors use the closed binomial form
    S_x(i, j) = sum_{a<=i, b<=j, a+b even} C(i,a) C(j,b) PA^(i-a) PB^(j-b)
                (a+b-1)!! / (2p)^((a+b)/2) * sqrt(pi/p)
of the Gaussian product theorem (Helgaker, Jorgensen, Olsen, ch. 9).
"""
from __future__ import annotations

import math

import numpy as np

from .basis import CART_COMPONENTS, BasisSet


def _dfact_odd(k: int) -> float:
    acc = 1.0
    while k > 0:
        acc *= k
        k -= 2
    return acc


def _s1d(i: int, j: int, PA: np.ndarray, PB: np.ndarray, p: np.ndarray) -> np.ndarray:
    out = np.zeros_like(p)
    for a in range(i + 1):
        for b in range(j + 1):
            if (a + b) % 2:
                continue
            out += (math.comb(i, a) * math.comb(j, b) * _dfact_odd(a + b - 1)
                    * PA ** (i - a) * PB ** (j - b) / (2.0 * p) ** ((a + b) // 2))
    return out * np.sqrt(np.pi / p)


def overlap_matrix(basis: BasisSet) -> np.ndarray:
    shells, prims, comps = basis.ao_records()
    n = basis.n_ao
    S = np.zeros((n, n))
    for A in shells:
        pa = prims[A["prim_start"]:A["prim_start"] + A["n_prims"]]
        ca = comps[A["comp_start"]:A["comp_start"] + A["n_funcs"]]
        for B in shells:
            pb = prims[B["prim_start"]:B["prim_start"] + B["n_prims"]]
            cb = comps[B["comp_start"]:B["comp_start"] + B["n_funcs"]]
            a = pa["exponent"][:, None]
            b = pb["exponent"][None, :]
            p = a + b
            AB = A["center"] - B["center"]
            pref = np.exp(-a * b / p * float(AB @ AB)) * (pa["weighted_norm"][:, None]
                                                          * pb["weighted_norm"][None, :])
            Pc = [(a * A["center"][q] + b * B["center"][q]) / p for q in range(3)]
            blk = np.zeros((A["n_funcs"], B["n_funcs"]))
            for ia, c1 in enumerate(ca):
                la = (c1["lx"], c1["ly"], c1["lz"])
                for ib, c2 in enumerate(cb):
                    lb = (c2["lx"], c2["ly"], c2["lz"])
                    v = pref.copy()
                    for q in range(3):
                        v = v * _s1d(la[q], lb[q], Pc[q] - A["center"][q], Pc[q] - B["center"][q], p)
                    blk[ia, ib] = v.sum() * c1["scale"] * c2["scale"]
            o1, o2 = A["ao_offset"], B["ao_offset"]
            S[o1:o1 + A["n_funcs"], o2:o2 + B["n_funcs"]] = blk
    return S
