"""The mixed build's per-geometry AO cache (option ao_cache) through the C ABI.

With the cache the mixed rho stage reads (phi, grad phi) per (AO, point) from
device memory, FP64 for the block's leading rows and FP32 for the others,
instead of evaluating them every build.  Bars: the FP64 build is bitwise the
same with and without a cache (it never reads it); the mixed build with the
cache stays within the mixed gate of the FP64 build (|dE_xc| < 1e-7 Eh) and
within the same bar of the mixed build without it, with N and V within three
times the mixed build's own deviation from FP64; batched plans with the cache
give every molecule exactly what its own cached plan gives.
"""
import numpy as np
import pytest

from calchem_amd import _native, configs
from calchem_amd.basis import BasisSet
from calchem_amd.grid import MolecularGrid
from calchem_amd.vxc import BatchedVxcEngine, RksVxcEngine, UksVxcEngine

pytestmark = pytest.mark.gpu

GATE = 1e-7


def _case(name):
    if name == "c8h18":
        mol, n_occ, grid_shape = configs.alkane_c8h18(), 33, (50, 194)
    else:
        mol, n_occ, grid_shape = configs.water_cluster(20, seed=5), 100, (30, 110)
    basis = BasisSet.from_name("6-31g(d,p)", mol)
    grid = MolecularGrid.explicit(mol, *grid_shape)
    return mol, basis, grid, configs.random_density(basis, n_occ, 11)


@pytest.mark.parametrize("name,block_points", [("c8h18", 0), ("water20", 1024)])
def test_cache_on_and_off(name, block_points):
    # 1024-point blocks give blocks of more than 128 rows (row tiles past the
    # two a wave keeps in registers are read from the cache twice)
    _, basis, grid, P = _case(name)
    off = RksVxcEngine(basis, grid, block_points=block_points, ao_cache=1)
    on = RksVxcEngine(basis, grid, block_points=block_points, ao_cache=2)
    st_on, st_off = on.stats(), off.stats()
    assert st_off["ao_cache_bytes"] == 0 and st_on["ao_cache_bytes"] > 0
    print(f"{name}: cache {st_on['ao_cache_bytes'] / 2**20:.1f} MiB, max_m_rho {st_on['max_m_rho']}", flush=True)
    if block_points:
        assert st_on["max_m_rho"] > 128
    e64, V64, n64 = off.evaluate_full("b3lyp", P, "fp64")
    f64 = on.evaluate_full("b3lyp", P, "fp64")
    assert f64[0] == e64 and np.array_equal(f64[1], V64) and f64[2] == n64
    e0, V0, n0 = off.evaluate_full("b3lyp", P, "fp32")
    e1, V1, n1 = on.evaluate_full("b3lyp", P, "fp32")
    print(f"mixed vs fp64: no cache dE={e0 - e64:.2e} max|dV|={np.abs(V0 - V64).max():.2e} dN={n0 - n64:.2e}; "
          f"cache dE={e1 - e64:.2e} max|dV|={np.abs(V1 - V64).max():.2e} dN={n1 - n64:.2e}")
    # the gate is on E_xc; N and V are held to a few times what the mixed
    # build without the cache gives on the same case
    assert abs(e1 - e64) < GATE and abs(e1 - e0) < GATE
    assert abs(n1 - n64) < max(1e-7, 3.0 * abs(n0 - n64))
    assert np.abs(V1 - V64).max() < max(1e-6, 3.0 * np.abs(V0 - V64).max())
    assert np.abs(V1 - V1.T).max() == 0.0
    # repeated builds read the same cache: bitwise reproducible
    e2, V2, n2 = on.evaluate_full("b3lyp", P, "fp32")
    assert e2 == e1 and np.array_equal(V2, V1) and n2 == n1
    on.close()
    off.close()


def test_uks_mixed_with_cache():
    _, basis, grid, P = _case("c8h18")
    Pa, Pb = 0.5 * P, 0.5 * configs.random_density(basis, 32, 19)
    off = UksVxcEngine(basis, grid, ao_cache=1)
    on = UksVxcEngine(basis, grid, ao_cache=2)
    e64, Va64, Vb64, n64 = off.evaluate_full("b3lyp", Pa, Pb)
    e1, Va1, Vb1, n1 = on.evaluate_full("b3lyp", Pa, Pb, precision="fp32")
    print(f"uks mixed with cache vs fp64: dE={e1 - e64:.2e} dVa={np.abs(Va1 - Va64).max():.2e} "
          f"dVb={np.abs(Vb1 - Vb64).max():.2e}")
    e0, Va0, Vb0, n0 = off.evaluate_full("b3lyp", Pa, Pb, precision="fp32")
    assert abs(e1 - e64) < GATE and abs(e1 - e0) < GATE
    assert abs(n1 - n64) < max(1e-7, 3.0 * abs(n0 - n64))
    for V1, V0, V64 in ((Va1, Va0, Va64), (Vb1, Vb0, Vb64)):
        assert np.abs(V1 - V64).max() < max(1e-6, 3.0 * np.abs(V0 - V64).max())
    on.close()
    off.close()


def test_batched_plan_with_cache_equals_single_plans():
    systems, dens = [], []
    for k, n_w in enumerate((2, 4)):
        mol = configs.water_cluster(n_w, seed=40 + k)
        basis = BasisSet.from_name("6-31g(d,p)", mol)
        systems.append((basis, MolecularGrid.explicit(mol, 40, 110)))
        dens.append(configs.random_density(basis, 5 * n_w, 7 + k))
    bat = BatchedVxcEngine(systems, ao_cache=2)
    assert bat.stats()["ao_cache_bytes"] > 0
    res = bat.evaluate("pbe", dens, "fp32")
    for (basis, grid), P, (e, V, ne) in zip(systems, dens, res):
        single = RksVxcEngine(basis, grid, ao_cache=2)
        e1, V1, n1 = single.evaluate_full("pbe", P, "fp32")
        assert e == e1 and np.array_equal(V, V1) and ne == n1
        single.close()
    bat.close()


def test_bad_option_is_rejected():
    _, basis, grid, _ = _case("c8h18")
    with pytest.raises(_native.CalchemHipError) as ei:
        RksVxcEngine(basis, grid, ao_cache=3)
    assert ei.value.code == -1
