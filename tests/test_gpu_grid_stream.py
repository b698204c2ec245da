"""Grids generated on the device, materialised and streaming
(MolecularGrid::build / streaming_tasks / generate_task_points,
crates/calchem-dft/src/grid.rs:614-782).

The concatenated task points equal the materialised grid bit for bit (the
materialised grid itself equals the oracle's, tests/test_gpu_parity.py);
tasks are 8 radial shells of one atom; the Vxc build over the streaming grid
(chunked plans) equals the build over the materialised grid (the reference's
streaming-vs-materialised check, vxc.rs tests, holds it to 1e-13).
"""
import numpy as np
import pytest

import oracle_lib as O
from calchem_amd import configs
from calchem_amd.basis import BasisSet
from calchem_amd.grid import MolecularGrid
from calchem_amd.vxc import RksVxcEngine, StreamingVxcEngine

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n_w,level", [(1, "medium"), (3, "fine"), (2, "coarse")])
def test_streaming_tasks_concatenate_to_the_grid(n_w, level):
    mol = configs.water_cluster(n_w, 5)
    basis = BasisSet.from_name("6-31g(d,p)", mol)
    full = MolecularGrid.build_for_basis(mol, basis, level)
    st = MolecularGrid.build_streaming_for_basis(mol, basis, level)
    assert st.is_streaming() and st.n_points == 0
    tasks = st.streaming_tasks()
    assert all(e - s + 1 <= 8 for _, s, e in tasks)
    per_atom = {}
    for a, s, e in tasks:
        per_atom.setdefault(a, []).append((s, e))
    for a, rng in per_atom.items():
        assert rng[0][0] == 1 and all(rng[k + 1][0] == rng[k][1] + 1 for k in range(len(rng) - 1))
    pts = np.vstack([st.task_points(t) for t in tasks])
    assert np.array_equal(pts, full.points)
    ref = O.build_grid(mol, level, radial_min=0)
    assert pts.shape == ref.shape
    print(f"{n_w} waters {level}: {len(tasks)} tasks, {pts.shape[0]} points")


def test_streaming_vxc_equals_materialised():
    mol = configs.water_cluster(3, 6)
    basis = BasisSet.from_name("6-31g(d,p)", mol)
    P = configs.random_density(basis, 15, 6)
    full = MolecularGrid.build_for_basis(mol, basis, "medium")
    st = MolecularGrid.build_streaming_for_basis(mol, basis, "medium")
    se = StreamingVxcEngine(basis, st, max_points=8_000)
    fe = RksVxcEngine(basis, full)
    assert len(se.engines) > 2 and se.n_points == full.n_points
    for prec, tol in (("fp64", 1e-12), ("fp32", 1e-7)):
        e1, V1, n1 = se.evaluate_full("b3lyp", P, prec)
        e2, V2, n2 = fe.evaluate_full("b3lyp", P, prec)
        print(f"{prec}: dE={e1 - e2:.2e} max|dV|={np.abs(V1 - V2).max():.2e} dN={n1 - n2:.2e}")
        assert abs(e1 - e2) < tol and np.abs(V1 - V2).max() < 10 * tol and abs(n1 - n2) < tol
    se.close()
    fe.close()
