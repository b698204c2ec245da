"""The C ABI: the HIP library loads and exports every entry point declared in
include/calchem_hip.h, record layouts match the header byte for byte, and the
host side refuses to run without a device (no CPU fallback).  No GPU work.
"""
import ctypes as C
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

from calchem_amd import _native
from calchem_amd.basis import AO_COMP_DTYPE, AO_PRIM_DTYPE, AO_SHELL_DTYPE
from calchem_amd.vxc import GRID_PT_DTYPE

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "calchem_hip.h")


def declared_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(calchem_hip_\w+)\s*\(", txt)))


def test_header_declares_the_python_export_list():
    assert declared_functions() == sorted(_native.EXPORTS)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_native.LIB_PATH):
        pytest.skip("libcalchem_hip.so not built (run __graft_entry__.build())")
    out = subprocess.run(["nm", "-D", "--defined-only", _native.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\sT\s(calchem_hip_\w+)", out))
    missing = set(declared_functions()) - exported
    assert not missing, missing
    lib = C.CDLL(_native.LIB_PATH)
    for name in declared_functions():
        assert hasattr(lib, name)


def test_probe_library_exports_its_entry_point():
    # the MFMA throughput probe that bench.py reports beside the datasheet peak
    path = os.path.join(ROOT, "calchem_amd", "libcalchem_probe.so")
    if not os.path.exists(path):
        pytest.skip("libcalchem_probe.so not built (run __graft_entry__.build())")
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True, check=True).stdout
    assert re.search(r"\sT\scalchem_probe_mfma_tflops\b", out)
    lib = C.CDLL(path)
    lib.calchem_probe_mfma_tflops.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double)]
    out_v = C.c_double(0.0)
    # argument checks come before any device work
    assert lib.calchem_probe_mfma_tflops(2, 1, 1, C.byref(out_v)) == -1
    assert lib.calchem_probe_mfma_tflops(0, 0, 1, C.byref(out_v)) == -1
    assert lib.calchem_probe_mfma_tflops(0, 1, 0, C.byref(out_v)) == -1


def _c_layout():
    src = r"""
#include <stddef.h>
#include <stdio.h>
#include "calchem_hip.h"
#define P(T, f) printf(#T " " #f " %zu\n", offsetof(T, f))
int main(void) {
  printf("CalchemAoShell size %zu\n", sizeof(CalchemAoShell));
  P(CalchemAoShell, r2_cutoff); P(CalchemAoShell, ao_offset); P(CalchemAoShell, atom_index);
  printf("CalchemAoPrim size %zu\n", sizeof(CalchemAoPrim));
  printf("CalchemAoComp size %zu\n", sizeof(CalchemAoComp));
  P(CalchemAoComp, lx); P(CalchemAoComp, lz);
  printf("CalchemGridPt size %zu\n", sizeof(CalchemGridPt));
  P(CalchemGridPt, batch_id); P(CalchemGridPt, host_atom);
  printf("CalchemVxcOptions size %zu\n", sizeof(CalchemVxcOptions));
  P(CalchemVxcOptions, memory_limit_bytes);
  printf("CalchemVxcStats size %zu\n", sizeof(CalchemVxcStats));
  P(CalchemVxcStats, flops_rho); P(CalchemVxcStats, device_bytes); P(CalchemVxcStats, lds_bytes_per_block);
  P(CalchemVxcStats, ms_reduce_en);
  return 0;
}
"""
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "layout.c")
        exe = os.path.join(d, "layout")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        lines = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split("\n")
    return {" ".join(l.split()[:2]): int(l.split()[2]) for l in lines if l.strip()}


def test_record_layouts_match_header():
    L = _c_layout()
    assert L["CalchemAoShell size"] == AO_SHELL_DTYPE.itemsize == 56
    assert L["CalchemAoShell r2_cutoff"] == AO_SHELL_DTYPE.fields["r2_cutoff"][1]
    assert L["CalchemAoShell ao_offset"] == AO_SHELL_DTYPE.fields["ao_offset"][1]
    assert L["CalchemAoShell atom_index"] == AO_SHELL_DTYPE.fields["atom_index"][1]
    assert L["CalchemAoPrim size"] == AO_PRIM_DTYPE.itemsize == 16
    assert L["CalchemAoComp size"] == AO_COMP_DTYPE.itemsize == 24
    assert L["CalchemAoComp lx"] == AO_COMP_DTYPE.fields["lx"][1]
    assert L["CalchemGridPt size"] == GRID_PT_DTYPE.itemsize == 32
    assert L["CalchemGridPt host_atom"] == GRID_PT_DTYPE.fields["host_atom"][1]
    assert L["CalchemVxcOptions size"] == C.sizeof(_native.VxcOptions)
    assert L["CalchemVxcOptions memory_limit_bytes"] == _native.VxcOptions.memory_limit_bytes.offset
    assert L["CalchemVxcStats size"] == C.sizeof(_native.VxcStats)
    for f in ("flops_rho", "device_bytes", "lds_bytes_per_block", "ms_reduce_en"):
        assert L[f"CalchemVxcStats {f}"] == getattr(_native.VxcStats, f).offset


def test_disabled_device_refuses_without_fallback(monkeypatch):
    if not os.path.exists(_native.LIB_PATH):
        pytest.skip("libcalchem_hip.so not built")
    monkeypatch.setenv("CALCHEM_HIP_DISABLE", "1")
    assert not _native.is_available()
    from calchem_amd import configs
    from calchem_amd.basis import BasisSet
    from calchem_amd.grid import MolecularGrid
    from calchem_amd.vxc import RksVxcEngine

    mol = configs.water()
    b = BasisSet.from_name("6-31g(d,p)", mol)
    with pytest.raises(_native.CalchemHipError) as ei:
        MolecularGrid.explicit(mol, 10, 26)
    assert ei.value.code == -3
    g = MolecularGrid(np.array([[0.0, 0.0, 0.0, 1.0, 1.0, 0.0]]), "x")
    with pytest.raises(_native.CalchemHipError):
        RksVxcEngine(b, g)


def test_entry_points_reject_bad_arguments_before_device_work(monkeypatch):
    if not os.path.exists(_native.LIB_PATH):
        pytest.skip("libcalchem_hip.so not built")
    L = _native.load()
    # Null pointers and zero sizes are rejected with -1 (checked before any device query).
    assert L.calchem_hip_vxc_eval_rho(1, None, None, 10, 5) == -1
    assert L.calchem_hip_vxc_accum(1, None, None, 10, 5) == -1
    assert L.calchem_hip_vxc_build(0, 0, 0, None, None, None, None) == -1
    assert L.calchem_hip_vxc_release(0) == 0
