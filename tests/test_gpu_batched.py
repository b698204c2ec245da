"""Many molecules in one plan, and captured (HIP graph) builds, through the C ABI.

A batched build must give every molecule exactly what its own single-molecule
plan gives (same blocks, same reduction order), and the single-molecule path is
pinned against the oracle in test_gpu_parity.py; two molecules are also
checked against the oracle directly.  Graph replays are checked against the
eager build after each replay, with eager work and synchronisation between
replays.
"""
import numpy as np
import pytest

import oracle_lib as O
from calchem_amd import configs
from calchem_amd import hip_runtime as hr
from calchem_amd.basis import BasisSet
from calchem_amd.grid import MolecularGrid
from calchem_amd.vxc import BatchedVxcEngine, RksVxcEngine

pytestmark = pytest.mark.gpu

E_TOL, V_TOL, N_TOL = 1e-10, 1e-9, 1e-9


def _systems(sizes, seed=11, n_rad=30, n_ang=110):
    out = []
    for k, n in enumerate(sizes):
        mol = configs.water_cluster(n, 1000 * seed + k)
        basis = BasisSet.from_name("6-31g(d)", mol)
        grid = MolecularGrid.explicit(mol, n_rad, n_ang)
        P = configs.random_density(basis, 5 * n, 100 + k)
        out.append((basis, grid, P))
    return out


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_batched_equals_single_molecule_plans(prec):
    systems = _systems([2, 5, 3, 4, 2, 6])
    eng = BatchedVxcEngine([(b, g) for b, g, _ in systems])
    res = eng.evaluate("pbe", [P for _, _, P in systems], prec)
    for k, (b, g, P) in enumerate(systems):
        one = RksVxcEngine(b, g)
        e1, V1, n1 = one.evaluate_full("pbe", P, prec)
        e, V, ne = res[k]
        print(f"{prec} molecule {k} ({b.n_ao} AO): dE={e - e1:.1e} dV={np.abs(V - V1).max():.1e} dN={ne - n1:.1e}")
        assert e == e1 and ne == n1 and np.array_equal(V, V1)


def test_batched_matches_oracle():
    systems = _systems([3, 2], seed=12, n_rad=40, n_ang=194)
    eng = BatchedVxcEngine([(b, g) for b, g, _ in systems])
    res = eng.evaluate("b3lyp", [P for _, _, P in systems])
    for k, (b, g, P) in enumerate(systems):
        e_ref, V_ref, n_ref = O.vxc_rks(O.OracleBasis(b), g.points, "b3lyp", P)
        e, V, ne = res[k]
        dv = np.abs(V - V_ref).max()
        print(f"molecule {k}: dE={e - e_ref:.2e} dV={dv:.2e} dN={ne - n_ref:.2e}")
        assert abs(e - e_ref) < E_TOL and dv < V_TOL and abs(ne - n_ref) < N_TOL


def test_batched_device_entry_and_graph():
    systems = _systems([4, 2, 3], seed=13)
    eng = BatchedVxcEngine([(b, g) for b, g, _ in systems])
    dens = [P for _, _, P in systems]
    ref = eng.evaluate("pbe", dens)
    Pp = eng.pack(dens)
    dP = hr.DeviceArray.from_numpy(Pp)
    dV = hr.DeviceArray(Pp.nbytes)
    den = hr.DeviceArray(16 * eng.n_mol)
    st = hr.Stream()
    eng.build_device("pbe", dP.address, dV.address, den.address, st.address)
    st.synchronize()
    V = eng.unpack(dV.to_numpy((Pp.size,)))
    en = den.to_numpy((2 * eng.n_mol,))
    for k in range(eng.n_mol):
        assert en[2 * k] == ref[k][0] and en[2 * k + 1] == ref[k][2] and np.array_equal(V[k], ref[k][1])
    g = eng.graph_create("pbe", dP.address, dV.address, den.address)
    for rep in range(3):
        # new densities between replays (eager copies + synchronisation), each replay checked
        scale = 1.0 + 0.1 * rep
        dens_r = [scale * P for P in dens]
        ref_r = eng.evaluate("pbe", dens_r)
        dP.copy_from_numpy(eng.pack(dens_r))
        hr.device_synchronize()
        eng.graph_launch(g, st.address)
        st.synchronize()
        V = eng.unpack(dV.to_numpy((Pp.size,)))
        en = den.to_numpy((2 * eng.n_mol,))
        for k in range(eng.n_mol):
            assert en[2 * k] == ref_r[k][0] and en[2 * k + 1] == ref_r[k][2]
            assert np.array_equal(V[k], ref_r[k][1])
    eng.graph_destroy(g)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_single_molecule_graph_replays(prec):
    w = configs.WORKLOADS["c8h18"]
    mol = w["build"]()
    basis = BasisSet.from_name(w["basis"], mol)
    grid = MolecularGrid.explicit(mol, 50, 194)
    P0 = configs.random_density(basis, w["n_occ"], w["seed"])
    eng = RksVxcEngine(basis, grid)
    n = basis.n_ao
    dP = hr.DeviceArray.from_numpy(P0)
    dV = hr.DeviceArray(n * n * 8)
    den = hr.DeviceArray(16)
    st = hr.Stream()
    g = eng.graph_create("b3lyp", dP.address, dV.address, den.address, prec)
    for rep in range(4):
        P = P0 * (1.0 - 0.05 * rep)
        e_ref, V_ref, n_ref = eng.evaluate_full("b3lyp", P, prec)
        dP.copy_from_numpy(P)
        hr.device_synchronize()
        eng.graph_launch(g, st.address)
        st.synchronize()
        en = den.to_numpy((2,))
        V = dV.to_numpy((n, n))
        print(f"replay {rep}: dE={en[0] - e_ref:.1e} dV={np.abs(V - V_ref).max():.1e}")
        assert en[0] == e_ref and en[1] == n_ref and np.array_equal(V, V_ref)
    eng.graph_destroy(g)


def test_batched_with_an_800_ao_cluster_matches_oracle():
    """The many-molecule workload's largest member: 32 waters, 6-31G(d,p)
    (800 AO), beside small clusters of 2 and 5 waters, in one batched plan on a
    reduced grid (20 x 50 per atom); each molecule against the oracle (FP64)
    and the mixed build against the 1e-7 Eh gate, or for molecules larger than
    the 20 waters it is stated for its relative size, 1.2e-9 |E_xc| (1e-7 Eh on
    about -87 Eh; the error grows with the molecule)."""
    out = []
    for k, n in enumerate([32, 2, 5]):
        mol = configs.water_cluster(n, 3000 + k)
        basis = BasisSet.from_name("6-31g(d,p)", mol)
        grid = MolecularGrid.explicit(mol, 20, 50)
        out.append((basis, grid, configs.random_density(basis, 5 * n, 300 + k)))
    assert out[0][0].n_ao == 800
    eng = BatchedVxcEngine([(b, g) for b, g, _ in out])
    res = eng.evaluate("pbe", [P for _, _, P in out])
    res32 = eng.evaluate("pbe", [P for _, _, P in out], "fp32")
    for k, (b, g, P) in enumerate(out):
        e_ref, V_ref, n_ref = O.vxc_rks(O.OracleBasis(b), g.points, "pbe", P)
        e, V, ne = res[k]
        dv = np.abs(V - V_ref).max()
        print(f"molecule {k} ({b.n_ao} AO, {g.n_points} points): dE={e - e_ref:.2e} dV={dv:.2e} dN={ne - n_ref:.2e}, "
              f"mixed dE={res32[k][0] - e:.2e}")
        assert abs(e - e_ref) < E_TOL and dv < V_TOL and abs(ne - n_ref) < N_TOL
        assert abs(res32[k][0] - e) < max(1e-7, 1.2e-9 * abs(e))
