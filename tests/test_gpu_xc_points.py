"""The device functional (xc_device.hpp, the code every build runs) at the
golden points, through calchem_hip_xc_eval.

tests/golden/xc_points.json holds 42 records (the five points of
crates/calchem-dft/src/xc.rs:1228-1236 and a core / tail sweep that reaches
the density cut-off, the gamma clamp, the LYP underflow cut and the PBE
large-w and small-gamma forms, xc.rs:13,167-173,505-516,744-781) written
from the literature formulas in 50-digit arithmetic.  The bar is the one
tests/test_oracle_xc.py holds the oracle to: 1e-12 relative.
"""
import ctypes as C
import json
import os

import numpy as np
import pytest

from calchem_amd import _native
from calchem_amd.hip_runtime import DeviceArray, Stream
from calchem_amd.xc import XC_IDS

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "xc_points.json")


def _eval(func, pts, w=None, want_sums=True):
    L = _native.require_device()
    pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 5)
    n = pts.shape[0]
    out = np.zeros((max(n, 1), 6))
    sums = np.zeros(2)
    wa = None if w is None else np.ascontiguousarray(w, dtype=np.float64)
    rc = L.calchem_hip_xc_eval(XC_IDS[func], n, pts.ctypes.data, None if wa is None else wa.ctypes.data,
                               out.ctypes.data, sums.ctypes.data if want_sums else None)
    _native.check(rc, "calchem_hip_xc_eval")
    return out[:n], sums


def test_golden_points_on_the_device():
    recs = json.load(open(GOLDEN))
    assert len(recs) == 42
    worst = 0.0
    for func in XC_IDS:
        rs = [r for r in recs if r["func"] == func]
        if not rs:
            continue
        out, _ = _eval(func, [r["point"] for r in rs])
        for r, o in zip(rs, out):
            ref = [float(r["f"])] + [None if d is None else float(d) for d in r["d"]]
            for k in range(6):
                if ref[k] is None:
                    continue
                err = abs(o[k] - ref[k])
                worst = max(worst, err / max(abs(ref[k]), 1e-300))
                assert err <= 1e-12 * max(abs(ref[k]), 1e-300) + 1e-300, (func, r["point"], k, o[k], ref[k])
    print(f"42 golden records: worst relative error {worst:.2e}")


def test_weighted_sums_and_cutoff():
    rng = np.random.default_rng(11)
    n = 100_003  # not a multiple of the 256-point workgroup
    ra = 10.0 ** rng.uniform(-20, 2, n)
    rb = ra * rng.uniform(0.0, 1.5, n)
    g = 10.0 ** rng.uniform(-40, 3, (n, 3))
    pts = np.column_stack([ra, rb, g[:, 0], np.sqrt(g[:, 0] * g[:, 2]) * rng.uniform(-1, 1, n), g[:, 2]])
    w = rng.uniform(0.0, 2.0, n)
    for func in ("b3lyp", "pbe"):
        out, sums = _eval(func, pts, w)
        assert np.all(np.isfinite(out))
        dead = ra + rb < 1e-18  # xc.rs:13,167-171
        assert np.all(out[dead] == 0.0)
        e_ref = float(np.sum(w * out[:, 0]))
        n_ref = float(np.sum(w * (ra + rb)))
        assert abs(sums[0] - e_ref) <= 1e-12 * abs(e_ref) and abs(sums[1] - n_ref) <= 1e-12 * abs(n_ref)
        # deterministic: the same call gives the same bits
        out2, sums2 = _eval(func, pts, w)
        assert np.array_equal(out, out2) and np.array_equal(sums, sums2)
    _, s1 = _eval("svwn5", pts[:1000], None)
    out1, _ = _eval("svwn5", pts[:1000], np.ones(1000))
    assert abs(s1[0] - np.sum(out1[:, 0])) <= 1e-12 * abs(s1[0])


def test_device_entry_on_a_stream_and_edges():
    L = _native.require_device()
    pts = np.array([[0.3, 0.3, 0.2, 0.2, 0.2], [0.1, 0.05, 0.01, 0.0, 0.02]])
    ref, ref_sums = _eval("blyp", pts)
    d_in, d_out, d_s = DeviceArray.from_numpy(pts), DeviceArray(6 * 2 * 8), DeviceArray(16)
    st = Stream()
    rc = L.calchem_hip_xc_eval_device(XC_IDS["blyp"], 2, d_in.address, None, d_out.address, d_s.address,
                                      st.address)
    _native.check(rc, "calchem_hip_xc_eval_device")
    st.synchronize()
    assert np.array_equal(d_out.to_numpy((2, 6)), ref)
    assert np.array_equal(d_s.to_numpy(2), ref_sums)
    # no points: zero sums
    rc = L.calchem_hip_xc_eval_device(XC_IDS["blyp"], 0, None, None, None, d_s.address, st.address)
    _native.check(rc, "calchem_hip_xc_eval_device n=0")
    st.synchronize()
    assert np.array_equal(d_s.to_numpy(2), np.zeros(2))
    out = np.zeros(12)
    assert L.calchem_hip_xc_eval(99, 2, pts.ctypes.data, None, out.ctypes.data, None) == -7
    assert L.calchem_hip_xc_eval(0, -1, pts.ctypes.data, None, out.ctypes.data, None) == -1
