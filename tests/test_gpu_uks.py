"""Unrestricted (UKS) Vxc on the GPU against the CPU FP64 oracle, through the C ABI.

Reference semantics: UksVxcEngine::evaluate / accumulate_uks_points
(crates/calchem-dft/src/vxc.rs:1141-1348, 1662-1766).  The oracle's per-point
UKS path is pinned by UKS(P/2, P/2) == RKS(P) and per-spin finite differences
(tests/test_oracle_vxc.py).  FP64 bar: |dE_xc| < 1e-10 Eh, max|dV| < 1e-9,
|dN| < 1e-9.
"""
import numpy as np
import pytest

import oracle_lib as O
from calchem_amd import configs
from calchem_amd.basis import BasisSet
from calchem_amd.grid import MolecularGrid
from calchem_amd.vxc import RksVxcEngine, UksVxcEngine

pytestmark = pytest.mark.gpu

E_TOL, V_TOL, N_TOL = 1e-10, 1e-9, 1e-9


def _spin_densities(basis, n_a, n_b, seed):
    """Spin densities C_s C_s^T with seeded S-orthonormal occupied orbitals."""
    Pa = 0.5 * configs.random_density(basis, n_a, seed)
    Pb = 0.5 * configs.random_density(basis, n_b, seed + 17) if n_b > 0 else np.zeros_like(Pa)
    return Pa, Pb


@pytest.mark.parametrize("func", ["b3lyp", "pbe", "svwn5", "blyp"])
def test_uks_fp64_h2o_cation_like(func):
    mol = configs.water()
    basis = BasisSet.from_name("6-31g(d,p)", mol)
    grid = MolecularGrid.explicit(mol, 75, 302)
    Pa, Pb = _spin_densities(basis, 5, 4, 3)
    eng = UksVxcEngine(basis, grid)
    e, Va, Vb, ne = eng.evaluate_full(func, Pa, Pb)
    e_ref, Va_ref, Vb_ref, ne_ref = O.vxc_uks(O.OracleBasis(basis), grid.points, func, Pa, Pb)
    dva, dvb = np.abs(Va - Va_ref).max(), np.abs(Vb - Vb_ref).max()
    print(f"uks h2o {func}: E={e:.12f} dE={e - e_ref:.2e} dVa={dva:.2e} dVb={dvb:.2e} N={ne:.10f} "
          f"dN={ne - ne_ref:.2e}")
    assert abs(e - e_ref) < E_TOL and dva < V_TOL and dvb < V_TOL and abs(ne - ne_ref) < N_TOL
    assert np.abs(Va - Va.T).max() == 0.0 and np.abs(Vb - Vb.T).max() == 0.0


def test_uks_fp64_empty_beta_spin():
    """rho_b = 0 everywhere (a one-spin density) matches the oracle."""
    mol = configs.water()
    basis = BasisSet.from_name("6-31g(d,p)", mol)
    grid = MolecularGrid.build(mol, "medium")
    Pa, Pb = _spin_densities(basis, 5, 0, 4)
    eng = UksVxcEngine(basis, grid)
    e, Va, Vb, ne = eng.evaluate_full("b3lyp", Pa, Pb)
    e_ref, Va_ref, Vb_ref, ne_ref = O.vxc_uks(O.OracleBasis(basis), grid.points, "b3lyp", Pa, Pb)
    print(f"uks empty beta: dE={e - e_ref:.2e} dVa={np.abs(Va - Va_ref).max():.2e} "
          f"dVb={np.abs(Vb - Vb_ref).max():.2e}")
    assert abs(e - e_ref) < E_TOL and abs(ne - ne_ref) < N_TOL
    assert np.abs(Va - Va_ref).max() < V_TOL and np.abs(Vb - Vb_ref).max() < V_TOL


@pytest.mark.parametrize("func", ["b3lyp", "pbe"])
def test_uks_half_densities_equal_rks_c8h18(func):
    """UKS(P/2, P/2) == RKS(P) on the GPU (vxc.rs:2436-2448), FP64."""
    w = configs.WORKLOADS["c8h18"]
    mol = w["build"]()
    basis = BasisSet.from_name(w["basis"], mol)
    grid = MolecularGrid.explicit(mol, w["n_rad"], w["n_ang"])
    P = configs.random_density(basis, w["n_occ"], w["seed"])
    eng = UksVxcEngine(basis, grid)
    e_u, Va, Vb, n_u = eng.evaluate_full(func, 0.5 * P, 0.5 * P)
    e_r, V_r, n_r = RksVxcEngine.evaluate_full(eng, func, P)
    print(f"uks/rks {func}: dE={e_u - e_r:.2e} dVa={np.abs(Va - V_r).max():.2e} dN={n_u - n_r:.2e}")
    assert abs(e_u - e_r) < 1e-11 * max(1.0, abs(e_r))
    assert np.abs(Va - V_r).max() < 1e-11 and np.abs(Vb - V_r).max() < 1e-11
    assert np.abs(Va - Vb).max() < 1e-13
    assert abs(n_u - n_r) < 1e-10


def test_uks_mixed_gate_c8h18():
    """The mixed UKS build stays within the |dE_xc| < 1e-7 Eh gate of FP64."""
    w = configs.WORKLOADS["c8h18"]
    mol = w["build"]()
    basis = BasisSet.from_name(w["basis"], mol)
    grid = MolecularGrid.explicit(mol, w["n_rad"], w["n_ang"])
    Pa, Pb = _spin_densities(basis, 33, 32, 5)
    eng = UksVxcEngine(basis, grid)
    e64, Va64, Vb64, n64 = eng.evaluate_full("b3lyp", Pa, Pb, "fp64")
    e32, Va32, Vb32, n32 = eng.evaluate_full("b3lyp", Pa, Pb, "fp32")
    dv = max(np.abs(Va32 - Va64).max(), np.abs(Vb32 - Vb64).max())
    print(f"uks mixed: dE={e32 - e64:.3e} max|dV|={dv:.3e} dN={n32 - n64:.3e}")
    assert abs(e32 - e64) < 1e-7 and dv < 1e-4 and abs(n32 - n64) < 1e-7


def test_uks_mixed_half_densities_equal_rks_mixed():
    """Mixed UKS(P/2, P/2): Va and Vb bitwise equal (the two spins run the same
    kernels on the same inputs) and within the mixed build's own V error of
    the mixed RKS(P) build (max|dV| < 1e-6)."""
    w = configs.WORKLOADS["c8h18"]
    mol = w["build"]()
    basis = BasisSet.from_name(w["basis"], mol)
    grid = MolecularGrid.explicit(mol, w["n_rad"], w["n_ang"])
    P = configs.random_density(basis, w["n_occ"], w["seed"])
    eng = UksVxcEngine(basis, grid)
    e_u, Va, Vb, n_u = eng.evaluate_full("b3lyp", 0.5 * P, 0.5 * P, "fp32")
    e_r, V_r, n_r = RksVxcEngine.evaluate_full(eng, "b3lyp", P, "fp32")
    print(f"uks/rks mixed: dE={e_u - e_r:.2e} dVa={np.abs(Va - V_r).max():.2e} dN={n_u - n_r:.2e}")
    assert np.array_equal(Va, Vb)
    assert abs(e_u - e_r) < 1e-9 and np.abs(Va - V_r).max() < 1e-6 and abs(n_u - n_r) < 1e-9


def test_uks_device_entry_matches_host_entry():
    """calchem_hip_vxc_build_uks_device on device buffers gives the host-buffer result."""
    from calchem_amd import hip_runtime as hr

    mol = configs.water()
    basis = BasisSet.from_name("6-31g(d,p)", mol)
    grid = MolecularGrid.build(mol, "coarse")
    Pa, Pb = _spin_densities(basis, 5, 4, 6)
    eng = UksVxcEngine(basis, grid)
    e, Va, Vb, ne = eng.evaluate_full("pbe", Pa, Pb)
    n = basis.n_ao
    dPa, dPb = hr.DeviceArray.from_numpy(Pa), hr.DeviceArray.from_numpy(Pb)
    dVa, dVb, den = hr.DeviceArray(n * n * 8), hr.DeviceArray(n * n * 8), hr.DeviceArray(16)
    st = hr.Stream()
    eng.build_device_uks("pbe", dPa.address, dPb.address, dVa.address, dVb.address, den.address, st.address)
    st.synchronize()
    en = den.to_numpy((2,))
    assert en[0] == e and en[1] == ne
    assert np.array_equal(dVa.to_numpy((n, n)), Va) and np.array_equal(dVb.to_numpy((n, n)), Vb)
