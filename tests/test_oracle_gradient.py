"""The oracle's XC nuclear gradient against central finite differences of its
own E_xc (basis functions and grid moved with the atoms, density matrix held
fixed) and translational invariance.

Reference checks followed: vxc.rs:2512-2546 (finite differences of E_xc with
the grid rebuilt, and sum over atoms of the gradient ~ 0).  The finite
differences use h = 1e-4 bohr central differences, whose truncation error is
~1e-9 here; the bar is 1e-8.
"""
import numpy as np
import pytest

import oracle_lib as O
from calchem_amd import configs
from calchem_amd.basis import BasisSet
from calchem_amd.molecule import Atom, Molecule

A2B = 1.8897261245650618
SYMS = ["O", "H", "H"]
BASE = np.array([[0.02, -0.04, 0.11], [0.81, 0.59, -0.37], [-0.74, 0.63, -0.42]])  # vxc.rs:2345-2362


def _mol(xyz_ang):
    return Molecule([Atom.from_angstrom(s, p) for s, p in zip(SYMS, xyz_ang)])


def _energy(xyz_ang, func, Pa, Pb, n_rad=50, n_ang=194):
    mol = _mol(xyz_ang)
    b = BasisSet.from_name("6-31g(d)", mol)
    g = O.build_grid(mol, "explicit", n_rad, n_ang)
    e, _, _, _ = O.vxc_uks(O.OracleBasis(b), g, func, Pa, Pb)
    return e


def _fd(func, Pa, Pb, comps, h=1e-4):
    out = []
    for a, q in comps:
        xp, xm = BASE.copy(), BASE.copy()
        xp[a, q] += h / A2B
        xm[a, q] -= h / A2B
        out.append((_energy(xp, func, Pa, Pb) - _energy(xm, func, Pa, Pb)) / (2 * h))
    return np.array(out)


@pytest.mark.parametrize("func", ["pbe", "b3lyp", "svwn5"])
def test_rks_gradient_matches_finite_differences(func):
    mol = _mol(BASE)
    b = BasisSet.from_name("6-31g(d)", mol)
    P = configs.random_density(b, 5, 0)
    g = O.build_grid(mol, "explicit", 50, 194)
    G = O.xc_gradient(O.OracleBasis(b), mol, g, func, 0.5 * P, 0.5 * P)
    comps = [(0, 0), (0, 2), (1, 1), (2, 0), (2, 2)]
    fd = _fd(func, 0.5 * P, 0.5 * P, comps)
    an = np.array([G[a, q] for a, q in comps])
    print(func, "analytic", an, "fd", fd, "diff", an - fd)
    assert np.abs(an - fd).max() < 1e-8
    assert np.abs(G.sum(axis=0)).max() < 1e-12


def test_uks_gradient_matches_finite_differences():
    mol = _mol(BASE)
    b = BasisSet.from_name("6-31g(d)", mol)
    Pa = 0.5 * configs.random_density(b, 5, 1)
    Pb = 0.5 * configs.random_density(b, 4, 2)
    g = O.build_grid(mol, "explicit", 50, 194)
    G = O.xc_gradient(O.OracleBasis(b), mol, g, "b3lyp", Pa, Pb)
    comps = [(0, 1), (1, 0), (2, 2)]
    fd = _fd("b3lyp", Pa, Pb, comps)
    an = np.array([G[a, q] for a, q in comps])
    print("uks analytic", an, "fd", fd)
    assert np.abs(an - fd).max() < 1e-8
    assert np.abs(G.sum(axis=0)).max() < 1e-12


def test_rks_gradient_equals_uks_with_half_densities():
    mol = _mol(BASE)
    b = BasisSet.from_name("6-31g(d)", mol)
    P = configs.random_density(b, 5, 0)
    g = O.build_grid(mol, "explicit", 30, 110)
    ob = O.OracleBasis(b)
    G1 = O.xc_gradient(ob, mol, g, "pbe", 0.5 * P, 0.5 * P)
    # the same density split unevenly between the spins is a different state
    G2 = O.xc_gradient(ob, mol, g, "pbe", 0.3 * P, 0.7 * P)
    assert np.abs(G1 - G2).max() > 1e-6
    # and a single atom has no XC force
    one = Molecule([Atom.from_angstrom("O", (0.0, 0.0, 0.0))])
    b1 = BasisSet.from_name("6-31g(d)", one)
    P1 = configs.random_density(b1, 4, 3)
    g1 = O.build_grid(one, "explicit", 30, 110)
    assert np.abs(O.xc_gradient(O.OracleBasis(b1), one, g1, "pbe", 0.5 * P1, 0.5 * P1)).max() == 0.0
