"""RKS exchange-correlation engine on the GPU, mirroring ``calchem_dft::vxc``.

RksVxcEngine(basis, grid)            RksVxcEngine::new_with_gpu (vxc.rs:429-716)
engine.evaluate(basis, grid, xc, P)  RksVxcEngine::evaluate (vxc.rs:720-985) -> (E_xc, V)
engine.evaluate_full(...)            same, plus N = sum_g w_g rho_g
engine.telemetry()                   RksVxcEngine::telemetry (vxc.rs:420-425)
UksVxcEngine                         UksVxcEngine (vxc.rs:1473-1766): (E_xc, Va, Vb)
GpuVxcGrid                           calchem_cuda::GpuVxcGrid (lib.rs:1178-1306): the
                                     reference GPU backend contract (rho on the device,
                                     functional on the host, V on the device).

Differences from the reference, by design: there is no CPU fallback (a GPU
failure raises CalchemHipError), the functional runs on the device in the
fused build, and V is returned full-symmetric.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Tuple

import numpy as np

from . import _native
from .basis import AO_COMP_DTYPE, AO_PRIM_DTYPE, AO_SHELL_DTYPE, BasisSet
from .grid import MolecularGrid
from .xc import xc_id

PRECISIONS = {"fp64": 0, "fp32": 1}

GRID_PT_DTYPE = np.dtype([("x", "<f8"), ("y", "<f8"), ("z", "<f8"), ("batch_id", "<i4"),
                          ("host_atom", "<i4")], align=True)
assert GRID_PT_DTYPE.itemsize == 32


def _records(basis: BasisSet):
    shells, prims, comps = basis.ao_records()
    assert shells.dtype == AO_SHELL_DTYPE and prims.dtype == AO_PRIM_DTYPE and comps.dtype == AO_COMP_DTYPE
    return np.ascontiguousarray(shells), np.ascontiguousarray(prims), np.ascontiguousarray(comps)


class RksVxcEngine:
    def __init__(self, basis: BasisSet, grid: MolecularGrid, block_points: int = 0,
                 group_points: int = 0, memory_limit_bytes: int = 0, mixed_exact_threshold: float = 0.0,
                 rho_tile_points: int = 0, group_max_rows: int = 0, ao_cache: int = 0, v_small_bf16: int = 0):
        self._lib = _native.require_device()
        self.n_ao = basis.n_ao
        self.n_points = grid.n_points
        self.handle = 0
        self._kernel_invocations = 0
        shells, prims, comps = _records(basis)
        xyzw = grid.xyzw()
        opts = _native.VxcOptions(block_points, group_points, memory_limit_bytes, mixed_exact_threshold,
                                  rho_tile_points, group_max_rows, ao_cache, v_small_bf16)
        h = _native.new_handle()
        rc = self._lib.calchem_hip_vxc_create(
            _native.ptr(shells), len(shells), _native.ptr(prims), len(prims), _native.ptr(comps),
            len(comps), _native.ptr(xyzw), xyzw.shape[0], self.n_ao, C.byref(opts), h)
        _native.check(rc, "calchem_hip_vxc_create")
        self.handle = h

    @classmethod
    def new_with_gpu(cls, basis: BasisSet, grid: MolecularGrid, use_gpu: bool = True) -> "RksVxcEngine":
        if not use_gpu:
            raise ValueError("calchem_amd has no CPU Vxc engine; use_gpu must be True")
        return cls(basis, grid)

    def close(self):
        if self.handle:
            self._lib.calchem_hip_vxc_release(self.handle)
            self.handle = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def evaluate_full(self, func: str, P: np.ndarray, precision: str = "fp64") -> Tuple[float, np.ndarray, float]:
        P = np.ascontiguousarray(P, dtype=np.float64)
        n = self.n_ao
        if P.shape != (n, n):
            raise ValueError(f"density has shape {P.shape}, expected {(n, n)}")
        V = np.empty((n, n))
        e = C.c_double()
        ne = C.c_double()
        rc = self._lib.calchem_hip_vxc_build(self.handle, xc_id(func), PRECISIONS[precision],
                                             _native.ptr(P), _native.ptr(V), C.byref(e), C.byref(ne))
        _native.check(rc, "calchem_hip_vxc_build")
        self._kernel_invocations += 2
        return e.value, V, ne.value

    def evaluate(self, basis: Optional[BasisSet], grid: Optional[MolecularGrid], func: str,
                 density_total: np.ndarray, precision: str = "fp64") -> Tuple[float, np.ndarray]:
        e, V, _ = self.evaluate_full(func, density_total, precision)
        return e, V

    def build_device(self, func: str, d_P: int, d_V: int, d_exc_nelec: int, stream: int = 0,
                     precision: str = "fp64") -> None:
        """Enqueue a build on device pointers (no synchronisation)."""
        rc = self._lib.calchem_hip_vxc_build_device(self.handle, xc_id(func), PRECISIONS[precision],
                                                    d_P, d_V, d_exc_nelec, stream or None)
        _native.check(rc, "calchem_hip_vxc_build_device")
        self._kernel_invocations += 2

    def graph_create(self, func: str, d_P: int, d_V: int, d_exc_nelec: int, precision: str = "fp64") -> int:
        """Capture the build on these device buffers into a HIP graph; returns its id."""
        g = C.c_uint64()
        rc = self._lib.calchem_hip_vxc_graph_create(self.handle, xc_id(func), PRECISIONS[precision], d_P, d_V,
                                                    d_exc_nelec, C.byref(g))
        _native.check(rc, "calchem_hip_vxc_graph_create")
        return g.value

    def graph_launch(self, graph: int, stream: int = 0) -> None:
        _native.check(self._lib.calchem_hip_vxc_graph_launch(self.handle, graph, stream or None),
                      "calchem_hip_vxc_graph_launch")
        self._kernel_invocations += 2

    def graph_destroy(self, graph: int) -> None:
        _native.check(self._lib.calchem_hip_vxc_graph_destroy(self.handle, graph), "calchem_hip_vxc_graph_destroy")

    def set_atoms(self, mol, grid: MolecularGrid) -> None:
        """Atoms (z, bohr) and the grid's raw weights / host atoms, for xc_gradient."""
        z = np.ascontiguousarray(mol.z_array(), dtype=np.int32)
        xyz = np.ascontiguousarray(mol.coords_bohr().ravel(), dtype=np.float64)
        raw = np.ascontiguousarray(grid.points[:, 4], dtype=np.float64)
        host = np.ascontiguousarray(grid.points[:, 5].astype(np.int32))
        if raw.shape[0] != self.n_points:
            raise ValueError("grid does not match the engine's points")
        _native.check(self._lib.calchem_hip_vxc_set_atoms(self.handle, len(z), _native.ptr(z), _native.ptr(xyz),
                                                          _native.ptr(raw), _native.ptr(host)),
                      "calchem_hip_vxc_set_atoms")
        self._n_atoms = len(z)

    def xc_gradient(self, func: str, P: np.ndarray, Pb: Optional[np.ndarray] = None) -> np.ndarray:
        """dE_xc/dR_A [n_atoms, 3] at fixed density (xc_nuclear_gradient_with_gpu,
        vxc.rs:2030-2334): restricted for the total density P, or unrestricted for
        spin densities (P, Pb)."""
        n = self.n_ao
        Pa = np.ascontiguousarray(P, dtype=np.float64)
        if Pa.shape != (n, n) or (Pb is not None and np.shape(Pb) != (n, n)):
            raise ValueError("density shape does not match the basis")
        Pb_c = None if Pb is None else np.ascontiguousarray(Pb, dtype=np.float64)
        na = getattr(self, "_n_atoms", 0)
        if na <= 0:
            raise _native.CalchemHipError("calchem_hip_vxc_xc_gradient (set_atoms first)", -2)
        g = np.zeros(3 * na)
        rc = self._lib.calchem_hip_vxc_xc_gradient(self.handle, xc_id(func), _native.ptr(Pa),
                                                   None if Pb_c is None else _native.ptr(Pb_c), _native.ptr(g))
        _native.check(rc, "calchem_hip_vxc_xc_gradient")
        return g.reshape(na, 3)

    def set_profiling(self, on: bool) -> None:
        _native.check(self._lib.calchem_hip_vxc_set_profiling(self.handle, 1 if on else 0),
                      "calchem_hip_vxc_set_profiling")

    def stats(self) -> dict:
        s = _native.VxcStats()
        _native.check(self._lib.calchem_hip_vxc_stats(self.handle, C.byref(s)), "calchem_hip_vxc_stats")
        return s.as_dict()

    def telemetry(self) -> Tuple[int, int]:
        """(GPU kernel invocations, CPU fallback invocations); the latter is always 0."""
        return self._kernel_invocations, 0


class UksVxcEngine(RksVxcEngine):
    """Unrestricted engine, mirroring ``calchem_dft::vxc::UksVxcEngine``.

    UksVxcEngine(basis, grid)                  UksVxcEngine::new_with_gpu (vxc.rs:1473-1660)
    engine.evaluate(basis, grid, xc, Pa, Pb)   UksVxcEngine::evaluate (vxc.rs:1662-1766) -> (E_xc, Va, Vb)
    engine.evaluate_full(xc, Pa, Pb)           same, plus N = sum_g w_g (rho_a + rho_b)

    It shares the plan machinery of the restricted engine (one plan per
    geometry); the restricted entry points stay usable on the same object.
    """

    @classmethod
    def new_with_gpu(cls, basis: BasisSet, grid: MolecularGrid, use_gpu: bool = True) -> "UksVxcEngine":
        if not use_gpu:
            raise ValueError("calchem_amd has no CPU Vxc engine; use_gpu must be True")
        return cls(basis, grid)

    def evaluate_full(self, func: str, Pa: np.ndarray, Pb: np.ndarray = None,
                      precision: str = "fp64") -> Tuple[float, np.ndarray, np.ndarray, float]:
        if Pb is None:
            raise TypeError("UksVxcEngine.evaluate_full needs both spin densities")
        n = self.n_ao
        Pa = np.ascontiguousarray(Pa, dtype=np.float64)
        Pb = np.ascontiguousarray(Pb, dtype=np.float64)
        for P in (Pa, Pb):
            if P.shape != (n, n):
                raise ValueError(f"density has shape {P.shape}, expected {(n, n)}")
        Va, Vb = np.empty((n, n)), np.empty((n, n))
        e, ne = C.c_double(), C.c_double()
        rc = self._lib.calchem_hip_vxc_build_uks(self.handle, xc_id(func), PRECISIONS[precision], _native.ptr(Pa),
                                                 _native.ptr(Pb), _native.ptr(Va), _native.ptr(Vb), C.byref(e),
                                                 C.byref(ne))
        _native.check(rc, "calchem_hip_vxc_build_uks")
        self._kernel_invocations += 2
        return e.value, Va, Vb, ne.value

    def evaluate(self, basis: Optional[BasisSet], grid: Optional[MolecularGrid], func: str, Pa: np.ndarray,
                 Pb: np.ndarray = None, precision: str = "fp64") -> Tuple[float, np.ndarray, np.ndarray]:
        e, Va, Vb, _ = self.evaluate_full(func, Pa, Pb, precision)
        return e, Va, Vb

    def build_device_uks(self, func: str, d_Pa: int, d_Pb: int, d_Va: int, d_Vb: int, d_exc_nelec: int,
                         stream: int = 0, precision: str = "fp64") -> None:
        """Enqueue an unrestricted build on device pointers (no synchronisation)."""
        rc = self._lib.calchem_hip_vxc_build_uks_device(self.handle, xc_id(func), PRECISIONS[precision], d_Pa,
                                                        d_Pb, d_Va, d_Vb, d_exc_nelec, stream or None)
        _native.check(rc, "calchem_hip_vxc_build_uks_device")
        self._kernel_invocations += 2


class BatchedVxcEngine(RksVxcEngine):
    """Many independent molecules in one plan: each build stage is one launch
    over all their blocks.  No single reference counterpart (the reference
    builds one molecule per RksVxcEngine, vxc.rs:429-985).

    engine = BatchedVxcEngine([(basis_0, grid_0), (basis_1, grid_1), ...])
    engine.evaluate(func, [P_0, P_1, ...]) -> [(E_xc, V, N), ...]
    Densities / V on the device are packed: molecule k's n_k x n_k block at
    element offset engine.p_offsets[k].
    """

    def __init__(self, systems, block_points: int = 0, group_points: int = 0, memory_limit_bytes: int = 0,
                 mixed_exact_threshold: float = 0.0, rho_tile_points: int = 0, group_max_rows: int = 0,
                 ao_cache: int = 0, v_small_bf16: int = 0):
        self._lib = _native.require_device()
        self.handle = 0
        self._kernel_invocations = 0
        shells_l, prims_l, comps_l, pts_l = [], [], [], []
        shell_begin, point_begin, n_ao = [0], [0], []
        np_, nc_ = 0, 0
        for basis, grid in systems:
            sh, pr, cp = _records(basis)
            sh = sh.copy()
            sh["prim_start"] += np_
            sh["comp_start"] += nc_
            np_ += len(pr)
            nc_ += len(cp)
            shells_l.append(sh)
            prims_l.append(pr)
            comps_l.append(cp)
            pts_l.append(grid.xyzw())
            shell_begin.append(shell_begin[-1] + len(sh))
            point_begin.append(point_begin[-1] + grid.n_points)
            n_ao.append(basis.n_ao)
        shells = np.ascontiguousarray(np.concatenate(shells_l))
        prims = np.ascontiguousarray(np.concatenate(prims_l))
        comps = np.ascontiguousarray(np.concatenate(comps_l))
        xyzw = np.ascontiguousarray(np.concatenate(pts_l))
        self.n_mol = len(n_ao)
        self.n_aos = n_ao
        self.n_ao = int(sum(n_ao))
        self.n_points = int(point_begin[-1])
        self.p_offsets = np.concatenate([[0], np.cumsum([n * n for n in n_ao])]).astype(np.int64)
        sb = np.asarray(shell_begin, dtype=np.int32)
        pb = np.asarray(point_begin, dtype=np.int32)
        na = np.asarray(n_ao, dtype=np.int32)
        opts = _native.VxcOptions(block_points, group_points, memory_limit_bytes, mixed_exact_threshold,
                                  rho_tile_points, group_max_rows, ao_cache, v_small_bf16)
        h = _native.new_handle()
        rc = self._lib.calchem_hip_vxc_create_batched(self.n_mol, _native.ptr(sb), _native.ptr(pb), _native.ptr(na),
                                                      _native.ptr(shells), _native.ptr(prims), len(prims),
                                                      _native.ptr(comps), len(comps), _native.ptr(xyzw), C.byref(opts),
                                                      h)
        _native.check(rc, "calchem_hip_vxc_create_batched")
        self.handle = h

    def pack(self, mats) -> np.ndarray:
        """Concatenate per-molecule n_k x n_k matrices into the packed layout."""
        return np.ascontiguousarray(np.concatenate([np.ascontiguousarray(m, dtype=np.float64).ravel()
                                                    for m in mats]))

    def unpack(self, packed: np.ndarray):
        return [packed[self.p_offsets[k]:self.p_offsets[k + 1]].reshape(n, n) for k, n in enumerate(self.n_aos)]

    def evaluate(self, func: str, densities, precision: str = "fp64"):
        if len(densities) != self.n_mol:
            raise ValueError(f"{len(densities)} densities for {self.n_mol} molecules")
        for P, n in zip(densities, self.n_aos):
            if np.shape(P) != (n, n):
                raise ValueError(f"density has shape {np.shape(P)}, expected {(n, n)}")
        Pp = self.pack(densities)
        Vp = np.empty_like(Pp)
        e = np.empty(self.n_mol)
        ne = np.empty(self.n_mol)
        rc = self._lib.calchem_hip_vxc_build_batched(self.handle, xc_id(func), PRECISIONS[precision],
                                                     _native.ptr(Pp), _native.ptr(Vp), _native.ptr(e),
                                                     _native.ptr(ne))
        _native.check(rc, "calchem_hip_vxc_build_batched")
        self._kernel_invocations += 2
        return [(float(e[k]), V, float(ne[k])) for k, V in enumerate(self.unpack(Vp))]

    def evaluate_full(self, func: str, P, precision: str = "fp64"):
        raise TypeError("BatchedVxcEngine: use evaluate(func, [P_0, P_1, ...])")

    def build_device(self, func: str, d_P: int, d_V: int, d_exc_nelec: int, stream: int = 0,
                     precision: str = "fp64") -> None:
        """Enqueue a batched build on packed device buffers (d_exc_nelec holds 2 n_mol doubles)."""
        rc = self._lib.calchem_hip_vxc_build_batched_device(self.handle, xc_id(func), PRECISIONS[precision], d_P,
                                                            d_V, d_exc_nelec, stream or None)
        _native.check(rc, "calchem_hip_vxc_build_batched_device")
        self._kernel_invocations += 2


class GpuVxcGrid:
    """The reference GPU backend contract: upload, eval_rho_grad, accumulate_vxc."""

    def __init__(self, basis: BasisSet, grid: MolecularGrid, batch_size: int = 128,
                 memory_limit_bytes: int = 0):
        self._lib = _native.require_device()
        self.n_ao = basis.n_ao
        shells, prims, comps = _records(basis)
        w = grid.weights
        keep = np.nonzero(np.abs(w) >= 1e-24)[0]
        self.keep = keep
        n = len(keep)
        pts = np.zeros(n, dtype=GRID_PT_DTYPE)
        pts["x"], pts["y"], pts["z"] = (grid.coords[keep, 0], grid.coords[keep, 1], grid.coords[keep, 2])
        pts["batch_id"] = np.arange(n) // batch_size
        pts["host_atom"] = grid.points[keep, 5].astype(np.int32)
        n_b = (n + batch_size - 1) // batch_size
        # Every shell is a candidate of every batch: a valid (if loose) CSR.
        cand_shells = np.tile(np.arange(len(shells), dtype=np.int32), n_b)
        cand_offsets = (np.arange(n_b + 1, dtype=np.int32) * len(shells)).astype(np.int32)
        self.n_pts = n
        self.weights = np.ascontiguousarray(w[keep])
        h = _native.new_handle()
        rc = self._lib.calchem_hip_vxc_upload_grid(
            _native.ptr(shells), len(shells), _native.ptr(prims), len(prims), _native.ptr(comps),
            len(comps), _native.ptr(pts), n, _native.ptr(cand_offsets), len(cand_offsets),
            _native.ptr(cand_shells), len(cand_shells), self.n_ao, h, memory_limit_bytes)
        _native.check(rc, "calchem_hip_vxc_upload_grid")
        self.grid_id = h

    def eval_rho_grad(self, P: np.ndarray) -> np.ndarray:
        P = np.ascontiguousarray(P, dtype=np.float64)
        out = np.zeros((self.n_pts, 4))
        rc = self._lib.calchem_hip_vxc_eval_rho(self.grid_id, _native.ptr(P), _native.ptr(out), self.n_pts,
                                                self.n_ao)
        _native.check(rc, "calchem_hip_vxc_eval_rho")
        return out

    def accumulate_vxc(self, wv: np.ndarray) -> np.ndarray:
        wv = np.ascontiguousarray(wv, dtype=np.float64)
        V = np.zeros((self.n_ao, self.n_ao))
        rc = self._lib.calchem_hip_vxc_accum(self.grid_id, _native.ptr(wv), _native.ptr(V), self.n_pts,
                                             self.n_ao)
        _native.check(rc, "calchem_hip_vxc_accum")
        return V

    def close(self):
        if self.grid_id:
            self._lib.calchem_hip_vxc_release(self.grid_id)
            self.grid_id = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class StreamingVxcEngine:
    """The Vxc build over a streaming grid (MolecularGrid.build_streaming,
    grid.rs:636-729; the reference evaluates such grids task by task in
    evaluate_vxc_rks, vxc.rs:280-384): the grid's (atom, radial range) tasks
    are generated on the device and grouped in task order into chunks of at
    most `max_points` points; each chunk has its own resident plan and a build
    sums the chunks' E_xc, V and N in chunk order.  No full point list is kept
    on the host."""

    def __init__(self, basis: BasisSet, grid: MolecularGrid, max_points: int = 2_000_000, **plan_opts):
        if not grid.is_streaming():
            raise ValueError("StreamingVxcEngine needs a streaming grid (MolecularGrid.build_streaming)")
        self.n_ao = basis.n_ao
        self.engines = []
        self.n_points = 0
        chunk = []
        rows = 0
        for task in grid.streaming_tasks() + [None]:
            pts = grid.task_points(task) if task is not None else None
            if chunk and (pts is None or rows + pts.shape[0] > max_points):
                sub = MolecularGrid(np.vstack(chunk), grid.level)
                self.engines.append(RksVxcEngine(basis, sub, **plan_opts))
                self.n_points += sub.n_points
                chunk, rows = [], 0
            if pts is not None and pts.shape[0]:
                chunk.append(pts)
                rows += pts.shape[0]

    def close(self):
        for e in self.engines:
            e.close()
        self.engines = []

    def evaluate_full(self, func: str, P: np.ndarray, precision: str = "fp64") -> Tuple[float, np.ndarray, float]:
        e_tot, V_tot, n_tot = 0.0, np.zeros((self.n_ao, self.n_ao)), 0.0
        for eng in self.engines:
            e, V, ne = eng.evaluate_full(func, P, precision)
            e_tot += e
            V_tot += V
            n_tot += ne
        return e_tot, V_tot, n_tot

    def evaluate(self, basis, grid, func: str, density_total: np.ndarray, precision: str = "fp64"):
        e, V, _ = self.evaluate_full(func, density_total, precision)
        return e, V
