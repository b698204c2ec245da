"""The SCF row on the GPU: RHF / RKS with device integrals, Cholesky J / K,
the device Vxc build, SAD and DIIS (calchem_amd/scf.py, restating
crates/calchem-scf/src/rks.rs:131-615).

Anchors: the reference's converged energies (RHF: rhf_sto3g.rs, rhf_631g.rs,
polarization_and_ecp.rs; B3LYP/6-31G(d) on the Fine grid and Medium vs
Fine: kohn_sham_dft.rs:64-79) and the oracle's plain SCF on the same grid to
1e-8 Eh; the mixed-then-FP64 SCF (mixed Vxc kernels until the DIIS error
falls below 1e-5) against the all-FP64 SCF within 1e-8 Eh.
"""
import math

import numpy as np
import pytest

import oracle_lib as O
import oracle_scf
from calchem_amd import configs, guess
from calchem_amd.basis import BasisSet
from calchem_amd.grid import MolecularGrid
from calchem_amd.integrals import IntegralEngine
from calchem_amd.molecule import Atom, Molecule
from calchem_amd.scf import RksSolver, ScfOptions, rhf

pytestmark = pytest.mark.gpu


def h2o_bohr():
    return Molecule([Atom.from_bohr("O", (0, 0, 0)), Atom.from_bohr("H", (0, 1.432, 0)),
                     Atom.from_bohr("H", (1.112, -0.716, 0))])


def h2():
    return Molecule([Atom.from_bohr("H", (0, 0, 0)), Atom.from_bohr("H", (0, 0, 1.4))])


def water_standard():
    r, half = 0.9572, math.radians(104.52 / 2.0)
    y, z = r * math.sin(half), -r * math.cos(half)
    return Molecule([Atom.from_angstrom("O", (0, 0, 0)), Atom.from_angstrom("H", (0, y, z)),
                     Atom.from_angstrom("H", (0, -y, z))])


TIGHT = ScfOptions(cholesky_tau=1e-12, conv_energy=1e-11, conv_density=1e-10)


def test_sad_from_device_integrals_equals_oracle_sad():
    mol = configs.water_cluster(2, 7)
    basis = BasisSet.from_name("6-31g(d)", mol)
    S, _, _ = O.one_electron(O.OracleBasis(basis), mol)
    guess._SAD_CACHE.clear()
    P_gpu = guess.sad_guess(mol, basis, S, 10, IntegralEngine)
    guess._SAD_CACHE.clear()
    P_ref = guess.sad_guess(mol, basis, S, 10, oracle_scf.OracleIntegrals)
    guess._SAD_CACHE.clear()
    assert np.abs(P_gpu - P_ref).max() < 1e-9


@pytest.mark.parametrize("mol,bname,ref,tol", [
    (h2o_bohr(), "sto-3g", -74.69531911423618, 1e-10),
    (h2o_bohr(), "6-31g", -75.800277509588, 1e-9),
    (h2(), "6-31g", -1.126742704452, 1e-9),
    (water_standard(), "6-31g(d)", -76.01075, 5e-4),
])
def test_rhf_energies(mol, bname, ref, tol):
    res = rhf(mol, BasisSet.from_name(bname, mol), TIGHT)
    print(f"RHF {bname}: E={res.total_energy:.12f} ref={ref} ({res.iterations} iterations, guess {res.guess})")
    assert res.converged and abs(res.total_energy - ref) < tol


def test_b3lyp_water_fine_medium_and_oracle():
    mol = water_standard()
    basis = BasisSet.from_name("6-31g(d)", mol)
    out = {}
    for level in ("fine", "medium"):
        grid = MolecularGrid.build_for_basis(mol, basis, level)
        s = RksSolver(mol, basis, "b3lyp", grid, TIGHT)
        res = s.run()
        s.close()
        e_ref, _, _ = oracle_scf.scf(mol, basis, "b3lyp", grid.points, conv_e=1e-11, conv_d=1e-9)
        print(f"B3LYP {level}: E={res.total_energy:.10f} oracle={e_ref:.10f} ({res.iterations} iterations)")
        assert abs(res.total_energy - e_ref) < 1e-8
        out[level] = res.total_energy
    assert abs(out["fine"] - (-76.40895)) < 1.5e-3
    assert abs(out["medium"] - out["fine"]) < 5e-5


def test_mixed_then_fp64_scf_equals_fp64_scf():
    mol = configs.water_cluster(4, 2)
    basis = BasisSet.from_name("6-31g(d,p)", mol)
    grid = MolecularGrid.explicit(mol, 50, 194)
    res = {}
    for prec in ("fp64", "mixed"):
        s = RksSolver(mol, basis, "b3lyp", grid, ScfOptions(precision=prec))
        r = s.run()
        s.close()
        res[prec] = r
        n32 = r.precision_history.count("fp32")
        print(f"{prec}: E={r.total_energy:.10f} iterations={r.iterations} (mixed builds {n32}) "
              f"guess={r.guess} setup={ {k: round(v, 3) for k, v in r.timings.items()} }")
    assert res["mixed"].precision_history[0] == "fp32" and res["mixed"].precision_history[-1] == "fp64"
    assert abs(res["mixed"].total_energy - res["fp64"].total_energy) < 1e-8
