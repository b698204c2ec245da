"""Oracle Vxc / E_xc / N against exact identities.

Restates the reference's own checks (crates/calchem-dft/src/vxc.rs):
  V symmetric and V = dE_xc/dP by finite differences (vxc.rs:2388-2434),
  UKS(P/2, P/2) == RKS (vxc.rs:2436-2448), per-spin UKS derivatives
  (vxc.rs:2450-2495), per-point == cached BLAS-3 path (vxc.rs:2604-2649),
  zero-weight grid gives zero (vxc.rs:2745-2783);
and N = sum w rho = Tr(P S) to quadrature accuracy.
"""
import numpy as np
import pytest

import oracle_lib as O
from calchem_amd import configs
from calchem_amd.basis import BasisSet
from calchem_amd.molecule import Atom, Molecule
from calchem_amd.overlap import overlap_matrix


def distorted_water():
    return Molecule([Atom.from_angstrom("O", (0.02, -0.04, 0.11)), Atom.from_angstrom("H", (0.81, 0.59, -0.37)),
                     Atom.from_angstrom("H", (-0.74, 0.63, -0.42))])


def positive_density(n, shift):
    # P = A A^T + 0.05 I with sin-seeded A (vxc.rs:2364-2385)
    mu = np.arange(n)[:, None]
    k = np.arange(4)[None, :]
    A = 0.22 * np.sin((mu * 7 + k * 13 + shift * 5 + 3) * 0.37)
    return A @ A.T + 0.05 * np.eye(n)


@pytest.fixture(scope="module")
def setup():
    mol = distorted_water()
    b = BasisSet.from_name("6-31g(d)", mol)
    g = O.build_grid(mol, "coarse")
    return mol, b, O.OracleBasis(b), g


@pytest.mark.parametrize("func", ["svwn5", "pbe", "b3lyp"])
def test_vxc_is_exact_density_derivative(setup, func):
    _, b, ob, g = setup
    n = b.n_ao
    P = positive_density(n, 1)
    e0, V, _ = O.vxc_rks(ob, g, func, P, "perpoint")
    assert np.abs(V - V.T).max() < 1e-14
    eps = 1e-6
    for mu, nu in [(0, 0), (0, 3), (2, 7), (5, 11), (8, 14), (12, 12)]:
        Pp, Pm = P.copy(), P.copy()
        if mu == nu:
            Pp[mu, mu] += eps
            Pm[mu, mu] -= eps
            den = 2 * eps
        else:
            for a, c in ((mu, nu), (nu, mu)):
                Pp[a, c] += eps
                Pm[a, c] -= eps
            den = 4 * eps
        fd = (O.vxc_rks(ob, g, func, Pp, "perpoint")[0] - O.vxc_rks(ob, g, func, Pm, "perpoint")[0]) / den
        assert abs(V[mu, nu] - fd) < 1e-9, (func, mu, nu, V[mu, nu], fd)


@pytest.mark.parametrize("func", ["svwn5", "blyp", "pbe", "b3lyp", "b3lyp5", "pbe0"])
def test_perpoint_equals_cached_and_uks_half(setup, func):
    _, b, ob, g = setup
    P = positive_density(b.n_ao, 1)
    e1, V1, n1 = O.vxc_rks(ob, g, func, P, "perpoint")
    e2, V2, n2 = O.vxc_rks(ob, g, func, P, "cached")
    assert abs(e1 - e2) < 1e-12 and np.abs(V1 - V2).max() < 1e-12 and abs(n1 - n2) < 1e-12
    eu, Va, Vb, nu_ = O.vxc_uks(ob, g, func, 0.5 * P, 0.5 * P)
    assert abs(eu - e1) < 1e-13
    assert np.abs(Va - V1).max() < 1e-13 and np.abs(Vb - V1).max() < 1e-13
    assert abs(nu_ - n1) < 1e-12


@pytest.mark.parametrize("func", ["blyp", "b3lyp", "pbe0"])
def test_uks_per_spin_derivatives(setup, func):
    _, b, ob, g = setup
    pa, pb = positive_density(b.n_ao, 2), positive_density(b.n_ao, 5)
    _, Va, Vb, _ = O.vxc_uks(ob, g, func, pa, pb)
    eps = 1e-6
    for mu, nu in [(1, 1), (1, 6), (4, 10), (9, 15)]:
        d = np.zeros_like(pa)
        if mu == nu:
            d[mu, mu] = eps
            den = 2 * eps
        else:
            d[mu, nu] = d[nu, mu] = eps
            den = 4 * eps
        fa = (O.vxc_uks(ob, g, func, pa + d, pb)[0] - O.vxc_uks(ob, g, func, pa - d, pb)[0]) / den
        fb = (O.vxc_uks(ob, g, func, pa, pb + d)[0] - O.vxc_uks(ob, g, func, pa, pb - d)[0]) / den
        assert abs(Va[mu, nu] - fa) < 1e-9 and abs(Vb[mu, nu] - fb) < 1e-9


def test_electron_count_matches_trace_ps():
    mol = configs.water()
    b = BasisSet.from_name("6-31g(d,p)", mol)
    S = overlap_matrix(b)
    P = configs.random_density(b, 5, 0, S)
    g = O.build_grid(mol, "explicit", 75, 302)
    _, _, ne = O.vxc_rks(O.OracleBasis(b), g, "b3lyp", P)
    assert abs(ne - np.trace(P @ S)) < 1e-5
    assert abs(np.trace(P @ S) - 10.0) < 1e-10


def test_zero_weight_and_empty_grids(setup):
    _, b, ob, g = setup
    P = positive_density(b.n_ao, 1)
    gz = g.copy()
    gz[:, 3] = 0.0
    e, V, ne = O.vxc_rks(ob, gz, "b3lyp", P)
    assert e == 0.0 and ne == 0.0 and np.all(V == 0.0)
    e, V, ne = O.vxc_rks(ob, g[:0], "b3lyp", P, "perpoint")
    assert e == 0.0 and np.all(V == 0.0)


def test_density_screen_drops_functional_only(setup):
    # Points with rho < 1e-18 contribute nothing to E_xc and V (vxc.rs:220,894).
    _, b, ob, g = setup
    far = np.array([[1e4, 0, 0, 1.0, 1.0, 0.0]])
    P = positive_density(b.n_ao, 1)
    e, V, ne = O.vxc_rks(ob, far, "pbe", P)
    assert e == 0.0 and np.all(V == 0.0) and ne == 0.0
