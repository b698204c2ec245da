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
    dv = np.abs(V32 - V64).max()
    print(f"c8h18 {func} fp32 vs fp64: dE={e32 - e64:.3e} max|dV|={dv:.3e} dN={n32 - n64:.3e}")
    # The mixed build's accuracy gate (|dE_xc| < 1e-7 Eh against FP64).
    assert abs(e32 - e64) < 1e-7 and dv < 1e-4 and abs(n32 - n64) < 1e-7
    assert np.abs(V32 - V32.T).max() == 0.0


def test_small_group_v_operand_forms_agree():
    """The mixed build's small-group V kernel in its two operand forms (FP32,
    and three BF16 terms per FP32 operand): E_xc and N come from the rho stage
    and the functional, so they are bitwise equal; V differs at FP32 rounding,
    and both stay within the mixed build's bar of the FP64 build."""
    mol, basis, P, w = _case("c8h18")
    grid = MolecularGrid.explicit(mol, w["n_rad"], w["n_ang"])
    e64, V64, _ = RksVxcEngine(basis, grid).evaluate_full("b3lyp", P, "fp64")
    out = {}
    for form in (1, 2):
        eng = RksVxcEngine(basis, grid, v_small_bf16=form)
        out[form] = eng.evaluate_full("b3lyp", P, "fp32")
        eng.close()
    (e1, V1, n1), (e2, V2, n2) = out[1], out[2]
    d12, d1, d2 = np.abs(V1 - V2).max(), np.abs(V1 - V64).max(), np.abs(V2 - V64).max()
    print(f"small-group V forms: max|V_fp32 - V_bf16x3| {d12:.2e}; against FP64 {d1:.2e} / {d2:.2e}")
    assert e1 == e2 and n1 == n2
    assert d12 < 2e-6 and d1 < 1e-4 and d2 < 1e-4
    assert np.abs(V2 - V2.T).max() == 0.0
    with pytest.raises(RuntimeError):
        RksVxcEngine(basis, grid, v_small_bf16=3)


@pytest.mark.parametrize("block_points,group_points", [(64, 256), (96, 512), (512, 4096)])
def test_fused_fp64_c8h18_block_shapes(block_points, group_points):
    """Other block / group sizes give the same result."""
    mol, basis, P, w = _case("c8h18")
    grid = MolecularGrid.explicit(mol, w["n_rad"], w["n_ang"])
    eng = RksVxcEngine(basis, grid, block_points=block_points, group_points=group_points)
    st = eng.stats()
    e, V, ne = eng.evaluate_full("b3lyp", P)
    e_ref, V_ref, ne_ref = O.vxc_rks(O.OracleBasis(basis), grid.points, "b3lyp", P)
    dv = np.abs(V - V_ref).max()
    print(f"blocks {block_points}/{group_points}: n_blocks={st['n_blocks']} n_groups={st['n_groups']} "
          f"pairs={st['n_v_pairs']} max_m={st['max_m_rho']}/{st['max_m_v']} dE={e - e_ref:.2e} dV={dv:.2e}")
    assert abs(e - e_ref) < E_TOL and dv < V_TOL and abs(ne - ne_ref) < N_TOL
    assert np.abs(V - V.T).max() == 0.0


@pytest.mark.parametrize("block_points,group_points,group_max_rows",
                         [(256, 1024, 192), (32, 32, 0), (16, 48, 0)])
def test_group_shapes_both_precisions(block_points, group_points, group_max_rows):
    """192-row register-resident groups (the 12-tile kernels) and groups of one
    to three 16-point steps (the pipelined mixed V kernel's prologue cases):
    FP64 against the oracle, mixed against FP64 within its gate."""
    mol, basis, P, w = _case("c8h18")
    grid = MolecularGrid.explicit(mol, 40, 110)
    eng = RksVxcEngine(basis, grid, block_points=block_points, group_points=group_points,
                       group_max_rows=group_max_rows)
    st = eng.stats()
    e, V, ne = eng.evaluate_full("b3lyp", P)
    e_ref, V_ref, ne_ref = O.vxc_rks(O.OracleBasis(basis), grid.points, "b3lyp", P)
    dv = np.abs(V - V_ref).max()
    e32, V32, n32 = eng.evaluate_full("b3lyp", P, "fp32")
    dv32 = np.abs(V32 - V).max()
    print(f"{block_points}/{group_points}/{group_max_rows}: groups={st['n_groups']} max_m={st['max_m_v']} "
          f"dE={e - e_ref:.2e} dV={dv:.2e} mixed dE={e32 - e:.2e} dV={dv32:.2e}")
    assert abs(e - e_ref) < E_TOL and dv < V_TOL and abs(ne - ne_ref) < N_TOL
    # the gate is on E_xc; N (66 electrons) is held to 1e-6, a relative 1.5e-8
    assert abs(e32 - e) < 1e-7 and dv32 < 1e-4 and abs(n32 - ne) < 1e-6
    assert np.abs(V32 - V32.T).max() == 0.0


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_wide_blocks_use_tile_pair_kernel(prec):
    """The 20-water cluster in 4096-point blocks has blocks with more than 272
    significant AOs: their V goes through the 64x64 tile-pair kernel and the
    rho kernel's 16-point tiles; the result still matches the oracle."""
    mol = configs.water_cluster(20, 2)
    basis = BasisSet.from_name("6-31g(d,p)", mol)
    P = configs.random_density(basis, 100, 2)
    grid = MolecularGrid.explicit(mol, 50, 194)
    eng = RksVxcEngine(basis, grid, block_points=4096, group_points=4096)
    st = eng.stats()
    e, V, ne = eng.evaluate_full("b3lyp", P, prec)
    e_ref, V_ref, ne_ref = O.vxc_rks(O.OracleBasis(basis), grid.points, "b3lyp", P)
    dv = np.abs(V - V_ref).max()
    print(f"20 waters {prec}: n_blocks={st['n_blocks']} pairs={st['n_v_pairs']} max_m={st['max_m_rho']}/"
          f"{st['max_m_v']} rho pts {st['rho_points_fp64']}/{st['rho_points_fp32']} dE={e - e_ref:.2e} dV={dv:.2e}")
    assert st["n_v_pairs"] > 0
    tol = (E_TOL, V_TOL) if prec == "fp64" else (1e-7, 1e-4)
    assert abs(e - e_ref) < tol[0] and dv < tol[1] and abs(ne - ne_ref) < tol[0]
    assert np.abs(V - V.T).max() == 0.0


def test_points_without_significant_aos():
    """Blocks far from every atom have no significant AO: their points add
    nothing to E_xc, N or V (the oracle agrees), in both precisions."""
    mol, basis, P, _ = _case("h2o")
    g = MolecularGrid.build(mol, "coarse")
    rng = np.random.default_rng(5)
    far = np.zeros((700, 6))
    far[:, :3] = 200.0 + rng.uniform(-5.0, 5.0, (700, 3))
    far[:, 3] = far[:, 4] = 0.5
    pts = np.vstack([g.points, far])
    grid = MolecularGrid(pts, g.level)
    eng = RksVxcEngine(basis, grid)
    e_ref, V_ref, ne_ref = O.vxc_rks(O.OracleBasis(basis), pts, "pbe", P)
    for prec in ("fp64", "fp32", "fp64"):
        e, V, ne = eng.evaluate_full("pbe", P, prec)
        tol = (E_TOL, V_TOL) if prec == "fp64" else (1e-7, 1e-4)
        assert abs(e - e_ref) < tol[0] and np.abs(V - V_ref).max() < tol[1] and abs(ne - ne_ref) < tol[0]
