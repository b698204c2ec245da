"""GPU parity of the HIP Vxc path against the CPU FP64 oracle (through the C ABI).

Tolerances (the FP64 bar): |dE_xc| < 1e-10 Eh, max|dV| < 1e-9, |dN| < 1e-9.
"""
import numpy as np
import pytest

import oracle_lib as O
from calchem_amd import configs
from calchem_amd.basis import BasisSet
from calchem_amd.grid import MolecularGrid
from calchem_amd.molecule import Atom, Molecule
from calchem_amd.vxc import GpuVxcGrid, RksVxcEngine

pytestmark = pytest.mark.gpu

E_TOL, V_TOL, N_TOL = 1e-10, 1e-9, 1e-9


def _case(name):
    w = configs.WORKLOADS[name]
    mol = w["build"]()
    basis = BasisSet.from_name(w["basis"], mol)
    P = configs.random_density(basis, w["n_occ"], w["seed"])
    return mol, basis, P, w


def distorted_water():
    # crates/calchem-dft/src/vxc.rs:2345-2362
    return Molecule([Atom.from_angstrom("O", (0.02, -0.04, 0.11)), Atom.from_angstrom("H", (0.81, 0.59, -0.37)),
                     Atom.from_angstrom("H", (-0.74, 0.63, -0.42))])


@pytest.mark.parametrize("level", ["coarse", "medium", "fine"])
def test_grid_levels_match_oracle(level):
    mol = distorted_water()
    g = MolecularGrid.build(mol, level)
    ref = O.build_grid(mol, level)
    assert g.points.shape == ref.shape
    np.testing.assert_array_equal(g.points[:, 5], ref[:, 5])
    np.testing.assert_allclose(g.points[:, :3], ref[:, :3], rtol=0, atol=1e-12)
    np.testing.assert_allclose(g.points[:, 3:5], ref[:, 3:5], rtol=1e-12, atol=1e-300)


def test_grid_explicit_75x302_c8h18_matches_oracle():
    mol = configs.alkane_c8h18()
    g = MolecularGrid.explicit(mol, 75, 302)
    ref = O.build_grid(mol, "explicit", 75, 302)
    print("c8h18 grid points", g.n_points, "raw", 26 * 75 * 302)
    assert g.points.shape == ref.shape
    np.testing.assert_allclose(g.points[:, :3], ref[:, :3], rtol=0, atol=1e-12)
    np.testing.assert_allclose(g.points[:, 3:5], ref[:, 3:5], rtol=1e-11, atol=1e-300)


def test_eval_rho_matches_oracle():
    mol, basis, P, _ = _case("h2o")
    grid = MolecularGrid.build(mol, "coarse")
    G = GpuVxcGrid(basis, grid)
    rg = G.eval_rho_grad(P)
    ref = O.rho_points(O.OracleBasis(basis), grid.points[G.keep], P)
    err = np.abs(rg - ref).max(axis=0)
    scale = np.abs(ref).max(axis=0)
    print("eval_rho max abs err", err, "scale", scale)
    assert np.all(err <= 1e-11 * np.maximum(scale, 1.0))


def test_accum_matches_oracle_via_host_functional():
    mol, basis, P, _ = _case("h2o")
    grid = MolecularGrid.build(mol, "coarse")
    G = GpuVxcGrid(basis, grid)
    rg = G.eval_rho_grad(P)
    w = G.weights
    wv = np.zeros_like(rg)
    e = 0.0
    for g in range(rg.shape[0]):
        rho = rg[g, 0]
        if rho < 1e-18:
            continue
        gx, gy, gz = rg[g, 1:]
        gss = 0.25 * (gx * gx + gy * gy + gz * gz)
        f, va, _, gaa, gab, _ = O.eval_xc("b3lyp", 0.5 * rho, 0.5 * rho, gss, gss, gss)
        e += w[g] * f
        wv[g, 0] = 0.5 * w[g] * va
        us = w[g] * (gaa + 0.5 * gab)
        wv[g, 1:] = us * rg[g, 1:]
    V = G.accumulate_vxc(wv)
    e_ref, V_ref, _ = O.vxc_rks(O.OracleBasis(basis), grid.points, "b3lyp", P)
    print("accum |dE|", abs(e - e_ref), "max|dV|", np.abs(V - V_ref).max())
    assert abs(e - e_ref) < E_TOL
    assert np.abs(V - V_ref).max() < V_TOL


@pytest.mark.parametrize("func", ["b3lyp", "pbe", "svwn5", "blyp", "b3lyp5", "pbe0"])
def test_fused_fp64_h2o_75x302(func):
    mol, basis, P, w = _case("h2o")
    grid = MolecularGrid.explicit(mol, w["n_rad"], w["n_ang"])
    eng = RksVxcEngine(basis, grid)
    e, V, ne = eng.evaluate_full(func, P)
    e_ref, V_ref, ne_ref = O.vxc_rks(O.OracleBasis(basis), grid.points, func, P)
    dv = np.abs(V - V_ref).max()
    print(f"h2o {func}: E={e:.12f} ref={e_ref:.12f} dE={e - e_ref:.2e} dV={dv:.2e} N={ne:.10f} dN={ne - ne_ref:.2e}")
    assert abs(e - e_ref) < E_TOL and dv < V_TOL and abs(ne - ne_ref) < N_TOL
    assert np.abs(V - V.T).max() == 0.0


@pytest.mark.parametrize("func", ["pbe", "b3lyp"])
def test_fused_fp64_c8h18_75x302(func):
    mol, basis, P, w = _case("c8h18")
    grid = MolecularGrid.explicit(mol, w["n_rad"], w["n_ang"])
    eng = RksVxcEngine(basis, grid)
    e, V, ne = eng.evaluate_full(func, P)
    e_ref, V_ref, ne_ref = O.vxc_rks(O.OracleBasis(basis), grid.points, func, P)
    dv = np.abs(V - V_ref).max()
    print(f"c8h18 {func}: E={e:.12f} ref={e_ref:.12f} dE={e - e_ref:.2e} dV={dv:.2e} N={ne:.10f} "
          f"dN={ne - ne_ref:.2e}")
    print("stats", eng.stats())
    assert abs(e - e_ref) < E_TOL and dv < V_TOL and abs(ne - ne_ref) < N_TOL


@pytest.mark.parametrize("func", ["b3lyp", "pbe"])
def test_fused_fp32_c8h18_error_report(func):
    mol, basis, P, w = _case("c8h18")
    grid = MolecularGrid.explicit(mol, w["n_rad"], w["n_ang"])
    eng = RksVxcEngine(basis, grid)
    e64, V64, n64 = eng.evaluate_full(func, P, "fp64")
    e32, V32, n32 = eng.evaluate_full(func, P, "fp32")
    print(f"c8h18 {func} fp32 vs fp64: dE={e32 - e64:.3e} max|dV|={np.abs(V32 - V64).max():.3e} dN={n32 - n64:.3e}")
    assert abs(e32 - e64) < 1e-4
