"""Precision study of the density contraction rho = phi^T P phi on the 20-water
workload (test infrastructure, CPU only; uses the oracle's AO values).

For each way of rounding the inputs of T = P_sub Phi it reports the change of
E_xc against the FP64 build, to first order,
    dE_xc = sum_g w_g (v_rho d rho + 2 v_gamma grad rho . d grad rho),
with v_rho, v_gamma of the B3LYP functional at the FP64 density (the
perturbations are below 1e-6 relative, so second-order terms are far below
the 1e-7 Eh gate).  Variants:

  fp32          Phi and P rounded to FP32, FP32 products and accumulation
  fp32_phi      Phi rounded, P exact
  fp32_p        P rounded, Phi exact
  split2_fp32   Phi and P each as two FP32 terms, three products
  hybrid_tau    rows with max |phi| > tau in a block kept FP64, the rest FP32
  bf16x2        Phi and P each as hi + lo BF16 terms, three BF16 products
                (hi hi + hi lo + lo hi) accumulated in FP32: what a BF16 MFMA
                split computes
  bf16x3        three BF16 terms each, six products (all with i + j <= 2)
  hybrid_bf16x3 hybrid rows FP64, the rest bf16x3
  cache_tau     hybrid_tau, and the epilogue (rho = sum phi T, grad rho) also
                takes phi and grad phi of the non-leading rows rounded to FP32:
                what a per-geometry AO cache holding those rows in FP32 gives

    python oracle/precision_study.py [--stride 4] [--cluster N SEED --grid NR NA --occ N SEED --func F]

(--stride k keeps every k-th 128-point batch of the grid; dE_xc is scaled by
k, an estimate of the whole grid's.)
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "tests"))

import oracle_lib as O  # noqa: E402
from calchem_amd import configs  # noqa: E402
from calchem_amd.basis import BasisSet  # noqa: E402


def bf16(x: np.ndarray) -> np.ndarray:
    """Round float32 values to bfloat16 (round to nearest even), returned as float32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return r.view(np.float32)


def split(x: np.ndarray, terms: int, kind: str):
    out, rest = [], np.asarray(x, dtype=np.float64)
    for _ in range(terms):
        t = bf16(rest.astype(np.float32)) if kind == "bf16" else rest.astype(np.float32)
        out.append(t)
        rest = rest - t.astype(np.float64)
    return out


def contract(P, Phi, variant, lead=None):
    """T = P Phi (m x n) under a rounding variant; lead marks FP64 rows (hybrid)."""
    if variant == "fp64":
        return P @ Phi
    if variant == "fp32":
        return (P.astype(np.float32) @ Phi.astype(np.float32)).astype(np.float64)
    if variant == "fp32_phi":
        return P @ Phi.astype(np.float32).astype(np.float64)
    if variant == "fp32_p":
        return P.astype(np.float32).astype(np.float64) @ Phi
    if variant in ("split2_fp32", "bf16x2", "bf16x3"):
        kind = "fp32" if variant == "split2_fp32" else "bf16"
        n = 3 if variant == "bf16x3" else 2
        ps, fs = split(P, n, kind), split(Phi, n, kind)
        acc = np.zeros((P.shape[0], Phi.shape[1]), dtype=np.float32)
        for i in range(n):
            for j in range(n):
                if i + j < n:
                    acc += ps[i] @ fs[j]
        return acc.astype(np.float64)
    if variant.startswith("hybrid"):
        # as the mixed rho kernel: rows of T for the leading AOs on FP64 (all of
        # their columns), the other rows in the base precision
        base = "bf16x3" if variant.endswith("bf16x3") else "fp32"
        T = contract(P, Phi, base)
        if lead is not None and lead.any():
            T[lead] = P[lead] @ Phi
        return T
    raise ValueError(variant)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--stride", type=int, default=4)
    ap.add_argument("--tau", type=float, nargs="*", default=[0.25, 2.0])
    ap.add_argument("--cluster", type=int, nargs=2, default=None, metavar=("N_WATERS", "SEED"),
                    help="a seeded water cluster instead of the 20-water workload")
    ap.add_argument("--grid", type=int, nargs=2, default=None, metavar=("N_RAD", "N_ANG"))
    ap.add_argument("--occ", type=int, nargs=2, default=None, metavar=("N_OCC", "SEED"),
                    help="occupied orbitals and seed of the random density")
    ap.add_argument("--func", default="b3lyp")
    ap.add_argument("--variants", nargs="*", default=None)
    a = ap.parse_args()
    w = configs.WORKLOADS["water20"]
    mol = configs.water_cluster(*a.cluster) if a.cluster else w["build"]()
    basis = BasisSet.from_name(w["basis"], mol)
    n_occ, seed = a.occ if a.occ else (w["n_occ"], w["seed"])
    P = configs.random_density(basis, n_occ, seed)
    n_rad, n_ang = a.grid if a.grid else (w["n_rad"], w["n_ang"])
    pts = O.build_grid(mol, "explicit", n_rad, n_ang)
    ob = O.OracleBasis(basis)
    variants = a.variants or (["fp32", "fp32_phi", "fp32_p", "split2_fp32", "bf16x2", "bf16x3", "hybrid_bf16x3"] +
                              [f"hybrid_{t:g}" for t in a.tau] + [f"cache_{t:g}" for t in a.tau])
    dE = {v: 0.0 for v in variants}
    t0 = time.time()
    nb = (len(pts) + 127) // 128
    for b in range(0, nb, a.stride):
        sl = pts[128 * b:128 * (b + 1)]
        val, grad = O.ao_dense(ob, sl)
        sig = (np.abs(val) > 1e-16).any(axis=0) | (np.abs(grad) > 1e-16).any(axis=(0, 2))
        idx = np.nonzero(sig)[0]
        if len(idx) == 0:
            continue
        phi = val[:, idx].T                      # m x n
        dphi = grad[:, idx, :].transpose(2, 1, 0)  # 3 x m x n
        Ps = P[np.ix_(idx, idx)]
        T = Ps @ phi
        rho = np.sum(phi * T, axis=0)
        grho = 2.0 * np.sum(dphi * T[None], axis=1)  # 3 x n
        ww = sl[:, 3]
        vr = np.zeros(len(sl))
        vg = np.zeros(len(sl))
        for g in range(len(sl)):
            if rho[g] >= 1e-18:
                gam = 0.25 * float(grho[:, g] @ grho[:, g])
                o = O.eval_xc(a.func, 0.5 * rho[g], 0.5 * rho[g], gam, gam, gam)
                vr[g] = 0.5 * (o[1] + o[2])                 # d f / d rho
                vg[g] = 0.25 * (o[3] + o[4] + o[5])         # d f / d |grad rho|^2
        for v in variants:
            lead = None
            if v.startswith("hybrid") or v.startswith("cache"):
                tau = 0.25 if v.endswith("bf16x3") else float(v.split("_")[1])
                lead = np.abs(phi).max(axis=1) > tau
            Tv = contract(Ps, phi, "hybrid_" + v.split("_")[1] if v.startswith("cache") else v, lead)
            ph, dph = phi, dphi
            if v.startswith("cache"):
                # the non-leading rows' AO values and gradients as FP32
                ph = np.where(lead[:, None], phi, phi.astype(np.float32).astype(np.float64))
                dph = np.where(lead[None, :, None], dphi, dphi.astype(np.float32).astype(np.float64))
            # the epilogue forms rho and grad rho with these AO values
            rho_v = np.sum(ph * Tv, axis=0)
            grho_v = 2.0 * np.sum(dph * Tv[None], axis=1)
            d = ww * (vr * (rho_v - rho) + 2.0 * vg * np.sum(grho * (grho_v - grho), axis=0))
            dE[v] += float(np.sum(d))
        if b % (50 * a.stride) == 0:
            print(f"batch {b}/{nb} ({time.time() - t0:.0f}s)", flush=True)
    print(f"{mol.n_atoms // 3} waters, {a.func}, {n_rad}x{n_ang}, {nb} batches, every {a.stride}: "
          f"estimated dE_xc of the whole grid")
    for v in variants:
        print(f"  {v:15s} {a.stride * dE[v]: .2e} Eh")


if __name__ == "__main__":
    main()
