"""The device integral engine against the oracle (through the C ABI).

One-electron integrals and exact repulsion columns: the oracle's Obara-Saika /
HGP integrals (themselves pinned by the reference's golden values,
tests/test_oracle_integrals.py) to 1e-12.  Cholesky vectors: the
reconstruction error is bounded by the threshold; J and K of random
densities against the oracle's exact contraction.
"""
import numpy as np
import pytest

import oracle_lib as O
from calchem_amd import configs
from calchem_amd.basis import BasisSet
from calchem_amd.integrals import IntegralEngine
from calchem_amd.molecule import Atom, Molecule

pytestmark = pytest.mark.gpu


def h2o_bohr():
    return Molecule([Atom.from_bohr("O", (0, 0, 0)), Atom.from_bohr("H", (0, 1.432, 0)),
                     Atom.from_bohr("H", (1.112, -0.716, 0))])


def dimer():
    return configs.water_cluster(2, 7)


CASES = [("h2o", "sto-3g"), ("h2o", "6-31g(d,p)"), ("dimer", "6-31g(d)")]


def _mol(name):
    return h2o_bohr() if name == "h2o" else dimer()


@pytest.mark.parametrize("name,bname", CASES)
def test_one_electron_matches_oracle(name, bname):
    mol = _mol(name)
    basis = BasisSet.from_name(bname, mol)
    eng = IntegralEngine(mol, basis)
    S, T, V = eng.one_electron()
    So, To, Vo = O.one_electron(O.OracleBasis(basis), mol)
    d = [float(np.abs(a - b).max()) for a, b in ((S, So), (T, To), (V, Vo))]
    print(f"{name}/{bname}: n={basis.n_ao} max|dS|={d[0]:.1e} max|dT|={d[1]:.1e} max|dV|={d[2]:.1e}")
    assert max(d) < 1e-12
    for M in (S, T, V):
        assert np.array_equal(M, M.T)
    eng.close()


def _packed_oracle(E):
    n = E.shape[0]
    iu, ju = np.tril_indices(n)
    return E[iu, ju][:, iu, ju]  # [pair, pair]


@pytest.mark.parametrize("name,bname", CASES)
def test_repulsion_columns_and_diagonal_match_oracle(name, bname):
    mol = _mol(name)
    basis = BasisSet.from_name(bname, mol)
    eng = IntegralEngine(mol, basis)
    ns = len(basis.shells)
    pairs = [(i, j) for i in range(ns) for j in range(ns)]  # both orientations
    ncol = sum(basis.shells[i].n_functions * basis.shells[j].n_functions for i, j in pairs)
    cols = eng.eri_columns(pairs, ncol)
    E = O.eri_full(O.OracleBasis(basis))
    n = basis.n_ao
    iu, ju = np.tril_indices(n)
    c = 0
    worst = 0.0
    for i, j in pairs:
        si, sj = basis.shells[i], basis.shells[j]
        for x in range(si.n_functions):
            for y in range(sj.n_functions):
                ref = E[iu, ju, si.ao_offset + x, sj.ao_offset + y]
                worst = max(worst, float(np.abs(cols[c] - ref).max()))
                c += 1
    diag = eng.eri_diag()
    dd = float(np.abs(diag - E[iu, ju, iu, ju]).max())
    print(f"{name}/{bname}: {ncol} columns, max|d(ij|kl)|={worst:.1e}, diagonal {dd:.1e}")
    assert worst < 1e-12 and dd < 1e-12
    eng.close()


@pytest.mark.parametrize("tau", [1e-12, 1e-8])
def test_cholesky_jk_against_exact_contraction(tau):
    mol = dimer()
    basis = BasisSet.from_name("6-31g(d)", mol)
    eng = IntegralEngine(mol, basis)
    r = eng.cholesky(tau)
    E = O.eri_full(O.OracleBasis(basis))
    n = basis.n_ao
    rng = np.random.default_rng(4)
    C = rng.normal(size=(n, 6))
    Cn = rng.normal(size=(n, 2)) * 0.3
    P = C @ C.T - Cn @ Cn.T
    J, K = eng.jk_factors(C, Cn)
    Jr = np.einsum("ijkl,kl->ij", E, P)
    Kr = np.einsum("ikjl,kl->ij", E, P)
    scale = np.abs(P).max()
    dj, dk = float(np.abs(J - Jr).max()), float(np.abs(K - Kr).max())
    print(f"tau={tau:g}: rank {r} of {n * (n + 1) // 2}, max|dJ|={dj:.1e} max|dK|={dk:.1e} (|P|max {scale:.1f})")
    assert np.array_equal(J, J.T) and np.array_equal(K, K.T)
    bound = 50.0 * np.sqrt(tau) * np.abs(P).sum() if tau > 1e-10 else 1e-9 * scale
    assert dj < bound and dk < bound
    # J / K of a density given as a matrix (split into positive / negative parts)
    J2, K2 = eng.jk(P)
    assert np.abs(J2 - J).max() < 1e-10 * scale * n and np.abs(K2 - K).max() < 1e-10 * scale * n
    eng.close()
