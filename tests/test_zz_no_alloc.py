"""calchem_hip_xc_eval_device_noalloc (the entry point that allocates nothing)
against calchem_hip_xc_eval_device, through the C ABI: two calls in a row,
the second smaller, both bit-identical to the allocating entry point (same
kernels, same summation order); sizes above its 16,777,216-point limit are
refused with -6 before anything runs."""
import numpy as np
import pytest

from calchem_amd import _native
from calchem_amd.hip_runtime import DeviceArray, Stream
from calchem_amd.xc import XC_IDS

pytestmark = pytest.mark.gpu


def _points(n, seed):
    rng = np.random.default_rng(seed)
    ra = 10.0 ** rng.uniform(-6, 1.5, n)
    rb = ra * rng.uniform(0.2, 1.2, n)
    gaa = 10.0 ** rng.uniform(-8, 2, n)
    gbb = 10.0 ** rng.uniform(-8, 2, n)
    gab = np.sqrt(gaa * gbb) * rng.uniform(-1, 1, n)
    return np.column_stack([ra, rb, gaa, gab, gbb]), rng.uniform(0.0, 1.0, n)


def _run(entry, xc, pts, w, st):
    n = pts.shape[0]
    d_in, d_w = DeviceArray.from_numpy(pts), DeviceArray.from_numpy(w)
    d_out, d_s = DeviceArray(n * 6 * 8), DeviceArray(16)
    rc = entry(XC_IDS[xc], n, d_in.address, d_w.address, d_out.address, d_s.address, st.address)
    _native.check(rc, "xc eval")
    st.synchronize()
    return d_out.to_numpy((n, 6)), d_s.to_numpy(2)


def test_noalloc_matches_allocating_entry_on_two_calls():
    L = _native.require_device()
    st = Stream()
    for xc, n, seed in (("b3lyp", 1_000_003, 1), ("pbe", 70_001, 2)):
        pts, w = _points(n, seed)
        ref_out, ref_s = _run(L.calchem_hip_xc_eval_device, xc, pts, w, st)
        out, s = _run(L.calchem_hip_xc_eval_device_noalloc, xc, pts, w, st)
        print(f"{xc} n={n}: sums {s} vs {ref_s}", flush=True)
        assert np.array_equal(out, ref_out)
        assert np.array_equal(s, ref_s)


def test_noalloc_limit_and_empty():
    L = _native.require_device()
    st = Stream()
    d = DeviceArray(64)
    d_s = DeviceArray(16)
    assert L.calchem_hip_xc_eval_device_noalloc(XC_IDS["pbe"], 16_777_217, d.address, None, d.address,
                                                d_s.address, st.address) == -6
    rc = L.calchem_hip_xc_eval_device_noalloc(XC_IDS["pbe"], 0, None, None, None, d_s.address, st.address)
    _native.check(rc, "noalloc n=0")
    st.synchronize()
    assert np.array_equal(d_s.to_numpy(2), np.zeros(2))
