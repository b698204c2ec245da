"""XC nuclear gradient on the GPU against the oracle's (which is pinned by
finite differences of E_xc and translational invariance in
test_oracle_gradient.py), through the C ABI.

Bar: max |dG| < 1e-9 Eh/bohr against the oracle, sum over atoms < 1e-10.
"""
import numpy as np
import pytest

import oracle_lib as O
from calchem_amd import configs
from calchem_amd.basis import BasisSet
from calchem_amd.grid import MolecularGrid
from calchem_amd.molecule import Atom, Molecule
from calchem_amd.vxc import RksVxcEngine, UksVxcEngine

pytestmark = pytest.mark.gpu

G_TOL, T_TOL = 1e-9, 1e-10


def _water():
    # crates/calchem-dft/src/vxc.rs:2345-2362
    return Molecule([Atom.from_angstrom("O", (0.02, -0.04, 0.11)), Atom.from_angstrom("H", (0.81, 0.59, -0.37)),
                     Atom.from_angstrom("H", (-0.74, 0.63, -0.42))])


def _check(mol, basis, grid, func, Pa, Pb, eng):
    eng.set_atoms(mol, grid)
    if Pb is None:
        G = eng.xc_gradient(func, Pa)
        ref = O.xc_gradient(O.OracleBasis(basis), mol, grid.points, func, 0.5 * Pa, 0.5 * Pa)
    else:
        G = eng.xc_gradient(func, Pa, Pb)
        ref = O.xc_gradient(O.OracleBasis(basis), mol, grid.points, func, Pa, Pb)
    d = np.abs(G - ref).max()
    print(f"{func}: max|dG|={d:.2e} |sum G|={np.abs(G.sum(axis=0)).max():.2e} max|G|={np.abs(ref).max():.3e}")
    assert d < G_TOL
    assert np.abs(G.sum(axis=0)).max() < T_TOL
    return G


@pytest.mark.parametrize("func", ["pbe", "b3lyp", "svwn5", "blyp"])
def test_rks_gradient_h2o(func):
    mol = _water()
    basis = BasisSet.from_name("6-31g(d)", mol)
    grid = MolecularGrid.explicit(mol, 50, 194)
    P = configs.random_density(basis, 5, 0)
    _check(mol, basis, grid, func, P, None, RksVxcEngine(basis, grid))


def test_uks_gradient_h2o():
    mol = _water()
    basis = BasisSet.from_name("6-31g(d,p)", mol)
    grid = MolecularGrid.build(mol, "medium")
    Pa = 0.5 * configs.random_density(basis, 5, 1)
    Pb = 0.5 * configs.random_density(basis, 4, 2)
    _check(mol, basis, grid, "b3lyp", Pa, Pb, UksVxcEngine(basis, grid))


def test_rks_gradient_c8h18_pbe():
    mol = configs.alkane_c8h18()
    basis = BasisSet.from_name("6-31g(d,p)", mol)
    grid = MolecularGrid.explicit(mol, 40, 110)
    P = configs.random_density(basis, 33, 1)
    _check(mol, basis, grid, "pbe", P, None, RksVxcEngine(basis, grid))


def test_rks_gradient_more_atoms_than_a_wavefront():
    """72 atoms: the Becke derivative kernel's lanes hold two atoms each."""
    mol = configs.water_cluster(24, 31)
    basis = BasisSet.from_name("sto-3g", mol)
    grid = MolecularGrid.explicit(mol, 20, 50)
    P = configs.random_density(basis, 5 * 24, 31)
    _check(mol, basis, grid, "pbe", P, None, RksVxcEngine(basis, grid))
