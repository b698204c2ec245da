"""Oracle grid pieces against known answers.

Lebedev rules: scipy.integrate.lebedev_rule point sets, monomial exactness and
the exp benchmark (crates/calchem-dft/src/lebedev.rs:279-325).  Molecular grid:
closed-form Gaussian integrals (grid.rs:838-888), Becke partition of unity.
"""
import math

import numpy as np
import pytest
from scipy.integrate import lebedev_rule

import oracle_lib as O
from calchem_amd.molecule import Atom, Molecule

RULES = [(26, 7), (50, 11), (110, 17), (194, 23), (302, 29), (590, 41)]


@pytest.mark.parametrize("n,order", RULES)
def test_lebedev_matches_scipy(n, order):
    L = O.lebedev(n)
    x, w = lebedev_rule(order)
    x = x.T
    d = np.linalg.norm(L[:, None, :3] - x[None], axis=-1)
    j = d.argmin(axis=1)
    assert len(set(j.tolist())) == n
    assert d.min(axis=1).max() < 1e-14
    np.testing.assert_allclose(L[:, 3], w[j], rtol=1e-14, atol=0)


def _dfact(k):
    return 1.0 if k <= 0 else float(np.prod(np.arange(k, 0, -2, dtype=np.float64)))


@pytest.mark.parametrize("n,order", RULES)
def test_lebedev_monomials_exact(n, order):
    L = O.lebedev(n)
    u, w = L[:, :3], L[:, 3]
    assert np.abs(np.sum(u * u, axis=1) - 1.0).max() < 1e-14
    for p in range(order + 1):
        for q in range(order + 1 - p):
            for r in range(order + 1 - p - q):
                num = np.sum(w * u[:, 0] ** p * u[:, 1] ** q * u[:, 2] ** r)
                if p % 2 or q % 2 or r % 2:
                    exact = 0.0
                else:
                    exact = 4 * math.pi * _dfact(p - 1) * _dfact(q - 1) * _dfact(r - 1) / _dfact(p + q + r + 1)
                assert abs(num - exact) < 1e-12, (n, p, q, r)


def test_lebedev_exp_benchmark():
    L = O.lebedev(302)
    assert abs(np.sum(L[:, 3] * np.exp(L[:, 0])) - 4 * math.pi * math.sinh(1.0)) < 1e-13


def test_bragg_slater_and_radial_map():
    assert O.lib().or_bragg_slater_bohr(1) == pytest.approx(0.35 * 1.8897261245650618, rel=1e-15)
    assert O.lib().or_bragg_slater_bohr(8) == pytest.approx(0.60 * 1.8897261245650618, rel=1e-15)
    # Single atom: weights are the bare radial x angular products, sum_t w_rad = integral of r^2 e^{-r^2}.
    mol = Molecule([Atom.from_bohr("H", (0.0, 0.0, 0.0))])
    g = O.build_grid(mol, "explicit", 75, 302)
    r2 = np.sum(g[:, :3] ** 2, axis=1)
    np.testing.assert_allclose(g[:, 3], g[:, 4], rtol=0, atol=0)
    assert abs(np.sum(g[:, 3] * np.exp(-r2)) - math.pi ** 1.5) / math.pi ** 1.5 < 1e-9


def test_single_center_gaussian_integral():
    # grid.rs:838-857
    mol = Molecule([Atom.from_bohr("C", (0.0, 0.0, 0.0))])
    g = O.build_grid(mol, "medium")
    r2 = np.sum(g[:, :3] ** 2, axis=1)
    for a in (0.15, 1.0, 8.5):
        exact = (math.pi / a) ** 1.5
        assert abs(np.sum(g[:, 3] * np.exp(-a * r2)) - exact) / exact < 1e-6


def distorted_water_grid():
    return Molecule([Atom.from_angstrom("O", (0.02, -0.04, 0.12)), Atom.from_angstrom("H", (0.08, 0.78, -0.45)),
                     Atom.from_angstrom("H", (-0.05, -0.74, -0.49))])


def test_multicenter_gaussian_sum():
    # grid.rs:860-888
    mol = distorted_water_grid()
    g = O.build_grid(mol, "medium")
    val = np.zeros(g.shape[0])
    for a in mol.atoms:
        val += np.exp(-np.sum((g[:, :3] - a.position_bohr) ** 2, axis=1))
    exact = 3.0 * math.pi ** 1.5
    assert abs(np.sum(g[:, 3] * val) - exact) / exact < 2e-5


def test_becke_partition_of_unity():
    mol = distorted_water_grid()
    rng = np.random.default_rng(7)
    for _ in range(200):
        r = rng.normal(scale=2.0, size=3)
        s = sum(O.becke_weight(mol, h, r) for h in range(3))
        # the screened fast partition drops cells below 1e-20 products only
        assert abs(s - 1.0) < 1e-12


def test_grid_levels_point_counts_and_filter():
    mol = distorted_water_grid()
    for lv in ("coarse", "medium", "fine"):
        g = O.build_grid(mol, lv)
        assert np.all(g[:, 3] > 1e-18)
        assert np.all(np.diff(g[:, 5]) >= 0)  # atom-major order
    # Fine uses at least 99 radial shells (grid.rs:745-752)
    g = O.build_grid(Molecule([Atom.from_bohr("H", (0, 0, 0))]), "fine")
    r = np.sort(np.linalg.norm(g[:, :3], axis=1))
    assert 1 + np.count_nonzero(np.diff(r) > 1e-9 * r[1:]) == 99


def test_explicit_grid_raw_count_and_weights():
    mol = distorted_water_grid()
    g = O.build_grid(mol, "explicit", 75, 302)
    assert g.shape[0] <= 3 * 75 * 302
    assert np.all(g[:, 3] <= g[:, 4] * (1 + 1e-15))


# ---------------------------------------------------------------------------
# The Becke partition and its radii, pinned independently of the oracle's own
# code: Becke, J. Chem. Phys. 88, 2547 (1988), eqs. 11-13, 19 and Appendix
# (k = 3 iterations of p(x) = 3x/2 - x^3/2, atomic-size adjustment
# a_ij = u_ij / (u_ij^2 - 1), u_ij = (chi - 1)/(chi + 1), chi = R_i / R_j,
# |a_ij| <= 1/2), with the Bragg-Slater radii of Slater, J. Chem. Phys. 41,
# 3199 (1964) (C 0.70, O 0.60 Angstrom) and Becke's 0.35 Angstrom for H.
# The reference restates the same construction at grid.rs:58-115,321-419.
# ---------------------------------------------------------------------------
ANG = 1.8897261245650618
SLATER_1964_ANGSTROM = {1: 0.35, 6: 0.70, 8: 0.60}


def becke1988_weights(z, xyz, r):
    """Unscreened Becke cell weights w_A(r) for every atom A (numpy, from the paper)."""
    n = len(z)
    rad = np.array([SLATER_1964_ANGSTROM[int(q)] * ANG for q in z])
    d = np.linalg.norm(r[None, :] - xyz, axis=1)
    cell = np.ones(n)
    for i in range(n):
        for j in range(n):
            if i == j:
                continue
            mu = (d[i] - d[j]) / np.linalg.norm(xyz[i] - xyz[j])
            chi = rad[i] / rad[j]
            u = (chi - 1.0) / (chi + 1.0)
            a = min(0.5, max(-0.5, u / (u * u - 1.0)))
            f = mu + a * (1.0 - mu * mu)
            for _ in range(3):
                f = 1.5 * f - 0.5 * f ** 3
            cell[i] *= 0.5 * (1.0 - f)
    return cell / cell.sum()


def test_bragg_slater_radii_of_the_workload_elements():
    for zz, ang in SLATER_1964_ANGSTROM.items():
        assert O.lib().or_bragg_slater_bohr(zz) == pytest.approx(ang * ANG, rel=1e-15), zz


@pytest.mark.parametrize("which", ["water", "c8h18"])
def test_becke_weights_match_unscreened_becke_1988(which):
    from calchem_amd import configs
    mol = distorted_water_grid() if which == "water" else configs.alkane_c8h18()
    z = mol.z_array()
    xyz = mol.coords_bohr()
    rng = np.random.default_rng(3)
    g = O.build_grid(mol, "explicit", 30, 110)
    # points of the grid itself (where weights matter) plus random points around the atoms
    pts = np.vstack([g[rng.choice(g.shape[0], 300, replace=False), :3],
                     xyz[rng.integers(0, len(z), 300)] + rng.normal(scale=1.5, size=(300, 3))])
    worst = 0.0
    for r in pts:
        ref = becke1988_weights(z, xyz, r)
        got = np.array([O.becke_weight(mol, h, r) for h in range(len(z))])
        worst = max(worst, float(np.abs(got - ref).max()))
    # the screened form drops cell products below 1e-20 and cells with 1 - s < 1e-17
    assert worst < 1e-14, worst


@pytest.mark.parametrize("heavy,frac", [(8, 0.35 / (0.35 + 0.60)), (6, 0.35 / (0.35 + 0.70)), (1, 0.5)])
def test_becke_cell_boundary_sits_at_the_radius_ratio(heavy, frac):
    """On the axis of a diatomic H-X, w_H = 1/2 where r_H / r_X = R_H / R_X
    (nu = 0 at mu = u_HX, Becke 1988 eq. A2): the H cell is the smaller one
    beside O or C, which fixes the sign of a_ij and the radii's ratio."""
    L = 1.9
    mol = Molecule([Atom.from_bohr("H", (0.0, 0.0, 0.0)), Atom.from_bohr(
        {8: "O", 6: "C", 1: "H"}[heavy], (0.0, 0.0, L))])
    lo, hi = 1e-6, L - 1e-6
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if O.becke_weight(mol, 0, np.array([0.0, 0.0, mid])) > 0.5:
            lo = mid
        else:
            hi = mid
    assert abs(0.5 * (lo + hi) / L - frac) < 1e-12, (heavy, 0.5 * (lo + hi) / L, frac)
    if heavy != 1:
        assert frac < 0.5
