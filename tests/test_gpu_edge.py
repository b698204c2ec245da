"""Edge cases the reference's own engine tests hold, on the GPU through the C ABI,
and invariances of the build paths.

Reference checks followed: a grid of one zero-weight point gives E_xc = 0 and
V = 0 exactly for both the restricted and the unrestricted engine, and finite
symmetric results for an ordinary coarse grid
(crates/calchem-dft/src/vxc.rs:2745-2783).  Also: a zero density matrix
(every point below the density screen) gives E_xc = N = 0 and V = 0 exactly.
"""
import numpy as np
import pytest

from calchem_amd import configs
from calchem_amd.basis import BasisSet
from calchem_amd.grid import MolecularGrid
from calchem_amd.molecule import Atom, Molecule
from calchem_amd.vxc import RksVxcEngine, UksVxcEngine

pytestmark = pytest.mark.gpu


def _water():
    # crates/calchem-dft/src/vxc.rs:2345-2362
    return Molecule([Atom.from_angstrom("O", (0.02, -0.04, 0.11)), Atom.from_angstrom("H", (0.81, 0.59, -0.37)),
                     Atom.from_angstrom("H", (-0.74, 0.63, -0.42))])


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_zero_weight_grid_gives_exact_zero(prec):
    mol = _water()
    basis = BasisSet.from_name("sto-3g", mol)
    Pa = 0.5 * configs.random_density(basis, 3, 1)
    Pb = 0.5 * configs.random_density(basis, 2, 2)
    grid = MolecularGrid(np.array([[0.0, 0.0, 0.0, 0.0, 0.0, 0.0]]), "coarse")
    e, V, ne = RksVxcEngine(basis, grid).evaluate_full("b3lyp", Pa + Pb, prec)
    assert e == 0.0 and ne == 0.0 and np.all(V == 0.0)
    eu, Va, Vb, nu = UksVxcEngine(basis, grid).evaluate_full("b3lyp", Pa, Pb, prec)
    assert eu == 0.0 and nu == 0.0 and np.all(Va == 0.0) and np.all(Vb == 0.0)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_coarse_grid_finite_and_symmetric(prec):
    mol = _water()
    basis = BasisSet.from_name("sto-3g", mol)
    Pa = 0.5 * configs.random_density(basis, 3, 1)
    Pb = 0.5 * configs.random_density(basis, 2, 2)
    grid = MolecularGrid.build(mol, "coarse")
    e, V, ne = RksVxcEngine(basis, grid).evaluate_full("b3lyp", Pa + Pb, prec)
    eu, Va, Vb, nu = UksVxcEngine(basis, grid).evaluate_full("b3lyp", Pa, Pb, prec)
    assert np.isfinite(e) and np.isfinite(eu) and np.isfinite(ne) and np.isfinite(nu)
    for M in (V, Va, Vb):
        assert np.all(np.isfinite(M)) and np.abs(M - M.T).max() == 0.0


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_zero_density_gives_exact_zero(prec):
    mol = _water()
    basis = BasisSet.from_name("6-31g(d)", mol)
    grid = MolecularGrid.build(mol, "coarse")
    Z = np.zeros((basis.n_ao, basis.n_ao))
    e, V, ne = RksVxcEngine(basis, grid).evaluate_full("pbe", Z, prec)
    assert e == 0.0 and ne == 0.0 and np.all(V == 0.0)
    eu, Va, Vb, nu = UksVxcEngine(basis, grid).evaluate_full("pbe", Z, Z, prec)
    assert eu == 0.0 and nu == 0.0 and np.all(Va == 0.0) and np.all(Vb == 0.0)


@pytest.mark.parametrize("prec,datasheet", [("fp64", 78.6), ("fp32", 157.3)])
def test_measured_mfma_ceiling_is_plausible(prec, datasheet):
    # the probe bench.py reports beside the datasheet peak: within (40 %, 105 %] of it
    import bench
    t = bench.measured_mfma_peak(prec, n_wg=1024, iters=500)
    assert 0.4 * datasheet < t <= 1.05 * datasheet, t


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_profiled_builds_equal_plain_builds(prec):
    # With profiling marks on, a build gathers all density blocks in one launch
    # before the rho launches; without, each rho launch class gathers its own
    # blocks on its side stream.  Same kernels on the same blocks: results are
    # bit-identical (restricted and unrestricted).
    w = configs.WORKLOADS["c8h18"]
    mol = w["build"]()
    basis = BasisSet.from_name(w["basis"], mol)
    grid = MolecularGrid.explicit(mol, 50, 194)
    P = configs.random_density(basis, w["n_occ"], 3)
    eng = RksVxcEngine(basis, grid)
    assert eng.stats()["max_m_rho"] > 64  # blocks in more than one 64-row launch class
    e0, V0, n0 = eng.evaluate_full("b3lyp", P, prec)
    eng.set_profiling(True)
    e1, V1, n1 = eng.evaluate_full("b3lyp", P, prec)
    eng.set_profiling(False)
    assert e0 == e1 and n0 == n1 and np.array_equal(V0, V1)
    ue = UksVxcEngine(basis, grid)
    Pa, Pb = 0.6 * P, 0.4 * P
    r0 = ue.evaluate_full("pbe", Pa, Pb, prec)
    ue.set_profiling(True)
    r1 = ue.evaluate_full("pbe", Pa, Pb, prec)
    ue.set_profiling(False)
    assert r0[0] == r1[0] and r0[3] == r1[3] and np.array_equal(r0[1], r1[1]) and np.array_equal(r0[2], r1[2])
