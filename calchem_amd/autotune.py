"""Pick block / group shapes for a geometry by timing builds on the device.

The rho stage works on k-d blocks of at most `block_points` points and the V
stage on groups of neighbouring blocks of at most `group_points` points and
`group_max_rows` AO rows (register-resident up to that many rows).  Which
shape is fastest depends on the basis density and the grid, so for a
geometry that will be built many times (every SCF iteration) it pays to try
a few plans once and keep the fastest: each candidate plan is built, warmed
up, and timed over `steps` builds with HIP events on one stream; the fastest
candidate whose E_xc agrees with the first candidate's to `e_tol` wins.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Iterable, List, Optional, Tuple

import numpy as np

from . import hip_runtime as hr
from .basis import BasisSet
from .grid import MolecularGrid
from .vxc import RksVxcEngine

# (block_points, group_points, group_max_rows[, v_small_bf16]); 0 = library
# default (320 / 384 / 272 / FP32 small-group V operands).
DEFAULT_CANDIDATES: Tuple[Tuple[int, ...], ...] = (
    (0, 0, 0), (0, 0, 0, 2), (256, 384, 0), (384, 384, 0), (512, 512, 0), (320, 512, 0), (320, 256, 0),
    (192, 384, 0), (320, 1024, 192),
)


@dataclass
class TuneResult:
    block_points: int
    group_points: int
    group_max_rows: int
    ms_per_build: float
    table: List[Tuple[Tuple[int, ...], float, float]]  # (candidate, ms, E_xc)
    v_small_bf16: int = 0

    @property
    def shape(self) -> Tuple[int, int, int, int]:
        return (self.block_points, self.group_points, self.group_max_rows, self.v_small_bf16)

    def engine(self, basis: BasisSet, grid: MolecularGrid, **kw) -> RksVxcEngine:
        return RksVxcEngine(basis, grid, block_points=self.block_points, group_points=self.group_points,
                            group_max_rows=self.group_max_rows, v_small_bf16=self.v_small_bf16, **kw)


def autotune(basis: BasisSet, grid: MolecularGrid, func: str, P: np.ndarray, precision: str = "fp64",
             candidates: Optional[Iterable[Tuple[int, ...]]] = None, steps: int = 5, warmup: int = 1,
             e_tol: Optional[float] = None, keep_first_within: float = 0.0) -> TuneResult:
    """e_tol: largest accepted |E_xc - E_xc(first candidate)|; default 1e-9 Eh for
    FP64 builds, 1e-7 Eh for mixed ones (their rounding depends on the blocking).
    keep_first_within: the first candidate is kept unless another is faster by
    more than this fraction of its time (so that run-to-run spread between
    near-equal shapes does not change the plan)."""
    if e_tol is None:
        e_tol = 1e-9 if precision == "fp64" else 1e-7
    cands = list(candidates or DEFAULT_CANDIDATES)
    if not cands:
        raise ValueError("no candidates")
    n = basis.n_ao
    stream = hr.Stream()
    dP = hr.DeviceArray.from_numpy(np.ascontiguousarray(P, dtype=np.float64))
    dV = hr.DeviceArray(n * n * 8)
    den = hr.DeviceArray(16)
    table = []
    e_first = None
    for cand in cands:
        bp, gp, gr = cand[:3]
        vs = cand[3] if len(cand) > 3 else 0
        eng = RksVxcEngine(basis, grid, block_points=bp, group_points=gp, group_max_rows=gr, v_small_bf16=vs)
        for _ in range(max(warmup, 1)):
            eng.build_device(func, dP.address, dV.address, den.address, stream.address, precision)
        stream.synchronize()
        e = float(den.to_numpy((2,))[0])
        ev0, ev1 = hr.Event(), hr.Event()
        ev0.record(stream)
        for _ in range(steps):
            eng.build_device(func, dP.address, dV.address, den.address, stream.address, precision)
        ev1.record(stream)
        ev1.synchronize()
        ms = ev0.elapsed_ms(ev1) / steps
        eng.close()
        if e_first is None:
            e_first = e
        table.append(((bp, gp, gr, vs), ms, e))
    best, ms = pick(table, e_tol, keep_first_within)
    return TuneResult(best[0], best[1], best[2], ms, table, best[3])


def pick(table, e_tol: float, keep_first_within: float = 0.0):
    """(candidate, ms) of the fastest entry of table [(candidate, ms, E_xc)] whose E_xc is within
    e_tol of the first entry's; the first entry itself unless that one is faster by more than
    keep_first_within of the first entry's time."""
    e_first = table[0][2]
    ok = [t for t in table if abs(t[2] - e_first) <= e_tol]
    best, ms, _ = min(ok, key=lambda t: t[1])
    if ms >= table[0][1] * (1.0 - keep_first_within):
        best, ms, _ = table[0]
    return best, ms
