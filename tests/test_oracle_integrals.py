"""The oracle's integrals and SCF energy against fixed values.

- Boys function: tests/golden/boys_points.json (mpmath quadrature of the
  definition, tests/golden/gen_boys_fixtures.py) and the reference's own
  values (crates/calchem-int/tests/boys.rs:23-49).
- One-electron and repulsion integrals: the reference's H2 and H2O STO-3G
  values (calchem-int/tests/one_electron.rs:10-126, eri_reference.rs:5-60,
  from mpmath closed forms and an independent Python code) and its Cartesian
  d-shell blocks (one_electron.rs:128-215).
- Whole-Hamiltonian anchors: converged RHF energies of H2O/STO-3G
  (calchem-scf/tests/rhf_sto3g.rs:60-75, 1e-10), H2O and H2 6-31G
  (rhf_631g.rs:71-77, rhf_sto3g.rs:78-104, 1e-9), RHF/6-31G(d)
  (polarization_and_ecp.rs:59-63, 5e-4) and B3LYP/6-31G(d) on the Fine grid
  (calchem-scf/tests/kohn_sham_dft.rs:64-79: -76.40895 within 1.5e-3, Medium
  vs Fine within 5e-5).
"""
import json
import math
import os

import numpy as np
import pytest

import oracle_lib as O
import oracle_scf
from calchem_amd.basis import BasisSet, Shell, normalized_contraction, primitive_norm
from calchem_amd.molecule import Atom, Molecule

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "boys_points.json")


def h2():
    return Molecule([Atom.from_bohr("H", (0, 0, 0)), Atom.from_bohr("H", (0, 0, 1.4))])


def h2o_bohr():
    return Molecule([Atom.from_bohr("O", (0, 0, 0)), Atom.from_bohr("H", (0, 1.432, 0)),
                     Atom.from_bohr("H", (1.112, -0.716, 0))])


def water_standard():
    r, half = 0.9572, math.radians(104.52 / 2.0)
    y, z = r * math.sin(half), -r * math.cos(half)
    return Molecule([Atom.from_angstrom("O", (0, 0, 0)), Atom.from_angstrom("H", (0, y, z)),
                     Atom.from_angstrom("H", (0, -y, z))])


def test_boys_against_mpmath_and_reference_values():
    worst = 0.0
    for r in json.load(open(GOLDEN)):
        F = O.boys(r["m"], r["T"])[r["m"]]
        ref = float(r["F"])
        worst = max(worst, abs(F - ref) / ref)
    assert worst < 2e-15, worst
    for m, t, v in [(0, 0.0, 1.0), (1, 0.0, 1 / 3), (4, 0.0, 1 / 9), (0, 1e-8, 0.9999999966666667),
                    (0, 1.0, 0.7468241328124271), (1, 1.0, 0.18947234582049235), (2, 1.0, 0.10026879814501737),
                    (0, 25.0, 0.17724538509027909508), (4, 25.0, 2.977722148749626937e-6),
                    (0, 29.34071349, 0.16360990987453259531), (0, 50.0, 0.12533141373155003)]:
        assert abs(O.boys(m, t)[m] - v) < 1e-14, (m, t)


def test_h2_sto3g_integrals():
    mol = h2()
    ob = O.OracleBasis(BasisSet.from_name("sto-3g", mol))
    S, T, V = O.one_electron(ob, mol)
    E = O.eri_full(ob)
    assert abs(S[0, 0] - 1) < 1e-14 and abs(S[0, 1] - 0.659318206134864) < 1e-12
    assert abs(T[0, 0] - 0.760031883566609) < 1e-12 and abs(T[0, 1] - 0.2364546559796740044) < 1e-12
    assert abs(V[0, 0] + 1.8804408924734455503) < 1e-12 and abs(V[0, 1] + 1.1948346203692893879) < 1e-12
    for idx, v in [((0, 0, 0, 0), 0.7746059439198977), ((1, 1, 1, 1), 0.7746059439198977),
                   ((0, 0, 1, 1), 0.5696759256037445), ((0, 1, 0, 1), 0.29702854027693154),
                   ((0, 0, 0, 1), 0.44410765803196034)]:
        assert abs(E[idx] - v) < 1e-12, idx


def test_h2o_sto3g_integrals():
    mol = h2o_bohr()
    ob = O.OracleBasis(BasisSet.from_name("sto-3g", mol))
    S, T, V = O.one_electron(ob, mol)
    E = O.eri_full(ob)
    np.testing.assert_allclose(S[0], [1.0, 0.236703926750692, 0, 0, 0, 0.085690824055376, 0.097488180540617],
                               rtol=0, atol=1e-12)
    np.testing.assert_allclose(np.diag(T), [29.003204305138183, 0.808127955339426, 2.528731199156927,
                                            2.528731199156927, 2.528731199156928, 0.760031883566609,
                                            0.760031883566609], rtol=0, atol=1e-11)
    np.testing.assert_allclose(np.diag(V), [-62.072885242882819, -10.430686918351427, -10.375796985370741,
                                            -10.480786072935665, -10.219140223130537, -6.779615288341706,
                                            -7.085094157869087], rtol=0, atol=1e-11)
    assert abs(V[2, 5] + 0.084699603336947) < 1e-12 and abs(V[3, 5] + 3.643299760159863) < 1e-12
    for idx, v in [((0, 0, 0, 0), 4.785065765661066), ((1, 1, 1, 1), 0.817206321588127),
                   ((2, 2, 2, 2), 0.880159093300264), ((2, 3, 2, 3), 0.047444445101798),
                   ((1, 2, 5, 5), 0.0), ((2, 5, 3, 6), -0.004937516143074), ((5, 6, 5, 6), 0.072470787680740)]:
        assert abs(E[idx] - v) < 1e-12, idx
    for M in (S, T, V):
        assert np.abs(M - M.T).max() < 1e-13


def _single(l, alpha, centre, atom, offset):
    c = normalized_contraction([alpha], [1.0], l)
    return Shell(atom, np.array(centre, dtype=float), l, np.array([alpha]), c, np.array([1.0]),
                 np.array([primitive_norm(alpha, (l, 0, 0))]), offset)


def test_cartesian_d_shell_blocks():
    shells = [_single(2, 0.8, (0, 0, 0), 0, 0), _single(0, 0.5, (0, 0, 1.2), 1, 6)]
    basis = BasisSet("d+s", shells, 7)
    mol = Molecule([Atom.from_bohr("O", (0, 0, 0))])  # the nuclear block is over an O nucleus at the origin
    S, T, V = O.one_electron(O.OracleBasis(basis), mol)
    np.testing.assert_allclose(np.diag(S)[:6], 1.0, rtol=0, atol=1e-14)
    assert abs(S[0, 3] - 1 / 3) < 1e-14
    np.testing.assert_allclose(np.diag(T)[:6], [1.733333333333334, 2.8000000000000001, 2.8000000000000001,
                                                1.733333333333334, 2.8000000000000001, 1.733333333333334],
                               rtol=0, atol=1e-13)
    np.testing.assert_allclose(S[:6, 6], [0.437890071945269, 0, 0, 0.437890071945269, 0, 0.680413804099572],
                               rtol=0, atol=1e-13)
    np.testing.assert_allclose(T[:6, 6], [0.116390662909950, 0, 0, 0.116390662909950, 0, 0.572622289693949],
                               rtol=0, atol=1e-13)
    np.testing.assert_allclose(V[:6, 6], [-2.847606831141998, 0, 0, -2.847606831141998, 0, -4.128554081538142],
                               rtol=0, atol=1e-12)


def test_eri_independent_of_shell_order_and_positive():
    """Reversing the atom order makes the code take other VRR / HRR routes to
    the same integrals; the pair matrix (ij|kl) over ij, kl is a Gram matrix."""
    mol = h2o_bohr()
    rev = Molecule(list(reversed(mol.atoms)))
    b1, b2 = BasisSet.from_name("6-31g(d,p)", mol), BasisSet.from_name("6-31g(d,p)", rev)
    E1, E2 = O.eri_full(O.OracleBasis(b1)), O.eri_full(O.OracleBasis(b2))
    # AO map: atom blocks swapped
    perm = []
    for s2 in b2.shells:
        same = [s for s in b1.shells if s.atom_index == 2 - s2.atom_index]
        k = [i for i, s in enumerate(b2.shells) if s.atom_index == s2.atom_index]
        k = k.index(next(i for i, s in enumerate(b2.shells) if s is s2))
        perm += list(range(same[k].ao_offset, same[k].ao_offset + s2.n_functions))
    p = np.array(perm)
    assert np.abs(E2 - E1[np.ix_(p, p, p, p)]).max() < 1e-13
    n = b1.n_ao
    w = np.linalg.eigvalsh(E1.reshape(n * n, n * n))
    assert w.min() > -1e-12


@pytest.mark.parametrize("mol,basis,ref,tol", [
    (h2o_bohr(), "sto-3g", -74.69531911423618, 1e-10),
    (h2o_bohr(), "6-31g", -75.800277509588, 1e-9),
    (h2(), "6-31g", -1.126742704452, 1e-9),
    (water_standard(), "6-31g(d)", -76.01075, 5e-4),
])
def test_rhf_energies(mol, basis, ref, tol):
    e, _, _ = oracle_scf.scf(mol, BasisSet.from_name(basis, mol))
    assert abs(e - ref) < tol, (basis, e, ref)


def test_b3lyp_water_fine_and_medium_grids():
    mol = water_standard()
    basis = BasisSet.from_name("6-31g(d)", mol)
    e_f, P, _ = oracle_scf.scf(mol, basis, "b3lyp", O.build_grid(mol, "fine"), conv_e=1e-9, conv_d=1e-7)
    e_m, _, _ = oracle_scf.scf(mol, basis, "b3lyp", O.build_grid(mol, "medium"), P0=P, conv_e=1e-9, conv_d=1e-7)
    assert abs(e_f - (-76.40895)) < 1.5e-3, e_f
    assert abs(e_m - e_f) < 5e-5, (e_m, e_f)
