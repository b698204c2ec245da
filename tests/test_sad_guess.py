"""The SAD guess (calchem_amd/guess.py, restating crates/calchem-scf/src/guess.rs:323-770)
on the oracle's integrals: electron counts, the H-atom UHF energy in 6-31G,
spherical averaging, block structure and Tr(P S)."""
import numpy as np

import oracle_lib as O
from calchem_amd import guess
from calchem_amd.basis import BasisSet
from calchem_amd.molecule import Atom, Molecule
from oracle_scf import OracleIntegrals


def test_spin_counts():
    assert guess.atomic_valence_spin_counts(1) == (1, 0)
    assert guess.atomic_valence_spin_counts(6) == (4, 2)
    assert guess.atomic_valence_spin_counts(7) == (5, 2)
    assert guess.atomic_valence_spin_counts(8) == (5, 3)


def test_hydrogen_atom_block_is_the_ground_state():
    mol = Molecule([Atom.from_bohr("H", (0, 0, 0))])
    basis = BasisSet.from_name("6-31g", mol)
    P = guess.solve_atomic_sad_block("H", basis.shells, OracleIntegrals)
    S, T, V = O.one_electron(O.OracleBasis(basis), mol)
    assert abs(np.sum(P * S) - 1.0) < 1e-12
    # one electron: E = Tr(P h) = the lowest eigenvalue of h in the S metric (-0.4982329 Eh in 6-31G)
    e = float(np.sum(P * (T + V)))
    w = np.linalg.eigvals(np.linalg.solve(S, T + V)).real.min()
    assert abs(e - w) < 1e-10 and abs(e - (-0.4982329)) < 1e-6, e


def test_oxygen_block_is_spherical_and_holds_eight_electrons():
    mol = Molecule([Atom.from_bohr("O", (0, 0, 0))])
    basis = BasisSet.from_name("6-31g(d)", mol)
    P = guess.solve_atomic_sad_block("O", basis.shells, OracleIntegrals)
    S, _, _ = O.one_electron(O.OracleBasis(basis), mol)
    assert abs(np.sum(P * S) - 8.0) < 1e-8
    assert np.abs(P - P.T).max() < 1e-12
    # the three components of each p shell carry the same population (spherical average)
    for sh in basis.shells:
        if sh.l == 1:
            d = np.diag(P)[sh.ao_offset:sh.ao_offset + 3]
            assert np.ptp(d) < 1e-8, d


def test_water_sad_density():
    mol = Molecule([Atom.from_bohr("O", (0, 0, 0)), Atom.from_bohr("H", (0, 1.432, 0)),
                    Atom.from_bohr("H", (1.112, -0.716, 0))])
    basis = BasisSet.from_name("6-31g(d,p)", mol)
    S, _, _ = O.one_electron(O.OracleBasis(basis), mol)
    P = guess.sad_guess(mol, basis, S, 5, OracleIntegrals)
    assert abs(np.sum(P * S) - 10.0) < 1e-10
    atom = np.concatenate([[sh.atom_index] * sh.n_functions for sh in basis.shells])
    assert np.all(P[atom[:, None] != atom[None, :]] == 0.0)
