"""The trained learned guess (calchem_amd/data/learned_guess_b3lyp_631gd.npz,
made by `python -m calchem_amd.learned_guess train`) on held-out perturbed
16-water clusters (B3LYP/6-31G(d), Medium grid): the SCF from the learned
density reaches the SAD run's energy (1e-8 Eh), in fewer iterations, from a
starting density much closer to the converged one; molecules the model does
not cover fall back to SAD."""
import os

import numpy as np
import pytest

from calchem_amd import configs
from calchem_amd import learned_guess as LG
from calchem_amd.basis import BasisSet

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def model():
    if not os.path.exists(LG.DEFAULT_MODEL):
        pytest.skip("no trained model in calchem_amd/data (run python -m calchem_amd.learned_guess train)")
    return LG.DEFAULT_MODEL


def test_learned_start_beats_sad_on_held_out_clusters(model):
    rows = LG._eval_chunk((model, [(k, 16, 100000 * 1004 + k) for k in range(4)]))
    for r in rows:
        print(f"cluster {r['cluster']}: iterations SAD {r['iters_sad']} learned {r['iters_learned']} "
              f"(purified {r['iters_learned_purified']}), dE {r['dE']:.1e}, max|dP| of the start "
              f"SAD {r['guess_err_sad']:.3f} learned {r['guess_err_learned']:.3f}, rms|dP| SAD "
              f"{r['guess_rms_sad']:.2e} learned {r['guess_rms_learned']:.2e}, inference {r['inference_s']:.3f}s")
        assert abs(r["dE"]) < 1e-8 and abs(r["dE_purified"]) < 1e-8
        assert r["iters_learned"] < r["iters_sad"] and r["iters_learned_purified"] < r["iters_sad"]
        # the start is closer to the converged density element by element
        # (RMS), and in its worst element
        assert r["guess_rms_learned"] < 0.25 * r["guess_rms_sad"]
        assert r["guess_err_learned"] < 0.5 * r["guess_err_sad"]


def test_model_declines_what_it_does_not_cover(model):
    m = LG.LearnedGuess.load(model)
    mol = configs.alkane_c8h18(2)
    basis = BasisSet.from_name("6-31g(d)", mol)
    assert m.density(mol, basis, np.eye(basis.n_ao), 9) is None
    w = LG.perturbed_cluster(3, 5)
    assert m.density(w, BasisSet.from_name("6-31g(d,p)", w), np.eye(75), 15) is None


def test_diagnosis_orders_the_starts(model):
    # exact blocks everywhere start next to the converged density; the
    # learned start sits between that and SAD
    rep = LG.diagnose(model, 1, 1004, n_waters=8)
    it = {k: v[0] for k, v in rep["iterations"].items()}
    print("iterations:", it)
    assert it["exact_all"] < it["learned"] < it["sad"]
    assert it["exact_intra"] <= it["learned"] + 1
