"""ctypes binding of the CPU FP64 oracle (oracle/liboracle.so).

Test infrastructure: the oracle is the parity checker, never the product.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")

FUNC_IDS = {"svwn5": 0, "blyp": 1, "pbe": 2, "b3lyp": 3, "b3lyp5": 4, "pbe0": 5}
LEVELS = {"coarse": 0, "medium": 1, "fine": 2, "ultrafine": 3, "explicit": 4}

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


class _Basis(C.Structure):
    _fields_ = [("n_shells", C.c_int), ("atom", C.c_void_p), ("center", C.c_void_p),
                ("l", C.c_void_p), ("n_prim", C.c_void_p), ("prim_start", C.c_void_p),
                ("exps", C.c_void_p), ("coefs", C.c_void_p)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
        L = C.CDLL(ORACLE_SO)
        L.or_primitive_norm.restype = C.c_double
        L.or_primitive_norm.argtypes = [C.c_double, C.c_int, C.c_int, C.c_int]
        L.or_normalized_contraction.argtypes = [C.c_int, _dp, _dp, C.c_int, _dp]
        L.or_lebedev.argtypes = [C.c_int, _dp]
        L.or_build_grid.restype = C.c_long
        L.or_build_grid.argtypes = [C.c_int, _ip, _dp, C.c_int, C.c_int, C.c_int, C.c_int,
                                    C.c_void_p, C.c_long]
        L.or_becke_weight_fast.restype = C.c_double
        L.or_becke_weight_fast.argtypes = [C.c_int, _ip, _dp, C.c_int, _dp]
        L.or_bragg_slater_bohr.restype = C.c_double
        L.or_eval_aos.argtypes = [C.POINTER(_Basis), _dp, _dp, _dp]
        L.or_eval_xc.argtypes = [C.c_int] + [C.c_double] * 5 + [_dp]
        for nm in ("or_vxc_rks_perpoint", "or_vxc_rks_cached"):
            f = getattr(L, nm)
            f.argtypes = [C.POINTER(_Basis), C.c_long, _dp, C.c_int, _dp, C.c_int, _dp,
                          C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.or_vxc_rks_cached_range.restype = C.c_long
        L.or_vxc_rks_cached_range.argtypes = [C.POINTER(_Basis), C.c_long, _dp, C.c_int, _dp,
                                              C.c_int, C.c_long, C.c_long, _dp,
                                              C.POINTER(C.c_double), C.POINTER(C.c_double),
                                              C.POINTER(C.c_double)]
        L.or_vxc_uks_perpoint.argtypes = [C.POINTER(_Basis), C.c_long, _dp, C.c_int, _dp, _dp,
                                          C.c_int, _dp, _dp, C.POINTER(C.c_double),
                                          C.POINTER(C.c_double)]
        L.or_overlap.argtypes = [C.POINTER(_Basis), _dp]
        L.or_xc_gradient.argtypes = [C.POINTER(_Basis), C.c_int, _ip, _dp, C.c_long, _dp, C.c_int, _dp, _dp,
                                     C.c_int, _dp]
        L.or_rho_points.argtypes = [C.POINTER(_Basis), C.c_long, _dp, _dp, C.c_int, _dp]
        L.or_ao_dense.argtypes = [C.POINTER(_Basis), C.c_long, _dp, _dp, _dp]
        L.or_boys.argtypes = [C.c_int, C.c_double, _dp]
        L.or_one_electron.argtypes = [C.POINTER(_Basis), C.c_int, _ip, _dp, _dp, _dp, _dp]
        L.or_eri_full.argtypes = [C.POINTER(_Basis), C.c_int, _dp]
        _lib = L
    return _lib


class OracleBasis:
    """Keeps the raw arrays alive for the lifetime of the C struct."""

    def __init__(self, basis):
        raw = basis.raw_arrays()
        self.arrays = {k: np.ascontiguousarray(v) for k, v in raw.items()}
        a = self.arrays
        self.n_ao = basis.n_ao
        self.struct = _Basis(len(a["l"]), a["atom"].ctypes.data, a["center"].ctypes.data,
                             a["l"].ctypes.data, a["n_prim"].ctypes.data,
                             a["prim_start"].ctypes.data, a["exps"].ctypes.data,
                             a["coefs"].ctypes.data)


def primitive_norm(alpha, lmn):
    return lib().or_primitive_norm(alpha, *lmn)


def normalized_contraction(exps, coefs, L):
    e = np.ascontiguousarray(exps, dtype=np.float64)
    c = np.ascontiguousarray(coefs, dtype=np.float64)
    out = np.zeros_like(e)
    rc = lib().or_normalized_contraction(e.size, e, c, L, out)
    if rc != 0:
        raise ValueError(rc)
    return out


def lebedev(n_ang):
    out = np.zeros(n_ang * 4)
    n = lib().or_lebedev(n_ang, out)
    if n != n_ang:
        raise ValueError(f"no Lebedev rule with {n_ang} points")
    return out.reshape(n_ang, 4)


def build_grid(mol, level="explicit", n_rad=0, n_ang=0, radial_min=0):
    z = np.ascontiguousarray(mol.z_array(), dtype=np.int32)
    xyz = np.ascontiguousarray(mol.coords_bohr().ravel())
    lv = LEVELS[level]
    need = lib().or_build_grid(len(z), z, xyz, lv, n_rad, n_ang, radial_min, None, 0)
    if need == -1:
        raise ValueError("bad grid arguments")
    cnt = -need
    out = np.zeros((max(cnt, 1), 6))
    got = lib().or_build_grid(len(z), z, xyz, lv, n_rad, n_ang, radial_min, out.ctypes.data, cnt)
    assert got == cnt
    return out[:cnt]


def becke_weight(mol, host, r):
    z = np.ascontiguousarray(mol.z_array(), dtype=np.int32)
    xyz = np.ascontiguousarray(mol.coords_bohr().ravel())
    return lib().or_becke_weight_fast(len(z), z, xyz, host, np.ascontiguousarray(r, dtype=np.float64))


def eval_aos(ob: OracleBasis, r):
    val = np.zeros(ob.n_ao)
    grad = np.zeros(ob.n_ao * 3)
    lib().or_eval_aos(C.byref(ob.struct), np.ascontiguousarray(r, dtype=np.float64), val, grad)
    return val, grad.reshape(-1, 3)


def eval_xc(func, ra, rb, gaa, gab, gbb):
    out = np.zeros(6)
    lib().or_eval_xc(FUNC_IDS[func], ra, rb, gaa, gab, gbb, out)
    return out


def _pts4(grid_rows):
    g = np.asarray(grid_rows)
    return np.ascontiguousarray(g[:, :4])


def vxc_rks(ob: OracleBasis, grid_rows, func, P, path="cached", threads=None):
    pts = _pts4(grid_rows)
    n = ob.n_ao
    V = np.zeros(n * n)
    e = C.c_double()
    ne = C.c_double()
    nt = threads or min(16, os.cpu_count() or 1)
    fn = lib().or_vxc_rks_cached if path == "cached" else lib().or_vxc_rks_perpoint
    rc = fn(C.byref(ob.struct), pts.shape[0], pts.ravel(), FUNC_IDS[func],
            np.ascontiguousarray(P, dtype=np.float64).ravel(), nt, V, C.byref(e), C.byref(ne))
    if rc != 0:
        raise RuntimeError(rc)
    return e.value, V.reshape(n, n), ne.value


def vxc_rks_cached_range(ob: OracleBasis, grid_rows, func, P, batch_begin, batch_end, threads):
    pts = _pts4(grid_rows)
    n = ob.n_ao
    V = np.zeros(n * n)
    e = C.c_double()
    ne = C.c_double()
    fl = C.c_double()
    nb = lib().or_vxc_rks_cached_range(C.byref(ob.struct), pts.shape[0], pts.ravel(),
                                       FUNC_IDS[func], np.ascontiguousarray(P).ravel(), threads,
                                       batch_begin, batch_end, V, C.byref(e), C.byref(ne),
                                       C.byref(fl))
    return nb, e.value, V.reshape(n, n), ne.value, fl.value


def vxc_uks(ob: OracleBasis, grid_rows, func, Pa, Pb, threads=None):
    pts = _pts4(grid_rows)
    n = ob.n_ao
    Va = np.zeros(n * n)
    Vb = np.zeros(n * n)
    e = C.c_double()
    ne = C.c_double()
    nt = threads or min(16, os.cpu_count() or 1)
    rc = lib().or_vxc_uks_perpoint(C.byref(ob.struct), pts.shape[0], pts.ravel(), FUNC_IDS[func],
                                   np.ascontiguousarray(Pa).ravel(), np.ascontiguousarray(Pb).ravel(),
                                   nt, Va, Vb, C.byref(e), C.byref(ne))
    if rc != 0:
        raise RuntimeError(rc)
    return e.value, Va.reshape(n, n), Vb.reshape(n, n), ne.value


def xc_gradient(ob: OracleBasis, mol, grid_rows, func, Pa, Pb, threads=None):
    """dE_xc/dR_A, [n_atoms, 3], for spin densities Pa, Pb (RKS: Pa = Pb = P/2)."""
    z = np.ascontiguousarray(mol.z_array(), dtype=np.int32)
    xyz = np.ascontiguousarray(mol.coords_bohr().ravel())
    pts = np.ascontiguousarray(np.asarray(grid_rows)[:, :6])
    g = np.zeros(3 * len(z))
    nt = threads or min(16, os.cpu_count() or 1)
    rc = lib().or_xc_gradient(C.byref(ob.struct), len(z), z, xyz, pts.shape[0], pts.ravel(), FUNC_IDS[func],
                              np.ascontiguousarray(Pa, dtype=np.float64).ravel(),
                              np.ascontiguousarray(Pb, dtype=np.float64).ravel(), nt, g)
    if rc != 0:
        raise RuntimeError(rc)
    return g.reshape(-1, 3)


def overlap(ob: OracleBasis):
    S = np.zeros(ob.n_ao * ob.n_ao)
    lib().or_overlap(C.byref(ob.struct), S)
    return S.reshape(ob.n_ao, ob.n_ao)


def rho_points(ob: OracleBasis, grid_rows, P, threads=None):
    pts = _pts4(grid_rows)
    out = np.zeros(pts.shape[0] * 4)
    nt = threads or min(16, os.cpu_count() or 1)
    rc = lib().or_rho_points(C.byref(ob.struct), pts.shape[0], pts.ravel(),
                             np.ascontiguousarray(P, dtype=np.float64).ravel(), nt, out)
    if rc != 0:
        raise RuntimeError(rc)
    return out.reshape(-1, 4)


def ao_dense(ob: OracleBasis, grid_rows):
    pts = _pts4(grid_rows)
    n = pts.shape[0]
    val = np.zeros(n * ob.n_ao)
    grad = np.zeros(n * ob.n_ao * 3)
    rc = lib().or_ao_dense(C.byref(ob.struct), n, pts.ravel(), val, grad)
    if rc != 0:
        raise RuntimeError(rc)
    return val.reshape(n, ob.n_ao), grad.reshape(n, ob.n_ao, 3)


def boys(mmax, T):
    out = np.zeros(mmax + 1)
    lib().or_boys(mmax, float(T), out)
    return out


def one_electron(ob: OracleBasis, mol):
    """(S, T, V) of the basis on molecule mol (V over its nuclei)."""
    z = np.ascontiguousarray(mol.z_array(), dtype=np.int32)
    xyz = np.ascontiguousarray(mol.coords_bohr().ravel())
    n = ob.n_ao
    S, T, V = np.zeros(n * n), np.zeros(n * n), np.zeros(n * n)
    rc = lib().or_one_electron(C.byref(ob.struct), len(z), z, xyz, S, T, V)
    if rc != 0:
        raise RuntimeError(rc)
    return S.reshape(n, n), T.reshape(n, n), V.reshape(n, n)


def eri_full(ob: OracleBasis, threads=None):
    n = ob.n_ao
    out = np.zeros(n ** 4)
    nt = threads or min(16, os.cpu_count() or 1)
    rc = lib().or_eri_full(C.byref(ob.struct), nt, out)
    if rc != 0:
        raise RuntimeError(rc)
    return out.reshape(n, n, n, n)
