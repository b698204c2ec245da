"""Build the gfx950 HIP library calchem_amd/libcalchem_hip.so with hipcc, and
the MFMA throughput probe calchem_amd/libcalchem_probe.so used by bench.py.

Every call compiles from source (no incremental reuse), so the library always
matches the tree.  Also used by __graft_entry__.build().
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libcalchem_hip.so")
PROBE_LIB = os.path.join(HERE, "libcalchem_probe.so")
PROBE_SOURCE = "probe_mfma.hip"
SOURCES = ["api.hip", "vxc_kernels.hip", "grid.hip", "grad_kernels.hip", "xc_points.hip"]
ARCH = os.environ.get("CALCHEM_HIP_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    # keep the compiler from emitting packed FP32 VALU ops
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # keep MFMA accumulators in ordinary VGPRs (no AGPR copies around loops)
    "-mllvm", "-amdgpu-mfma-vgpr-form",
    "-Wno-unused-value", "-Wno-unused-result", "-Wno-pass-failed",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the calchem_amd HIP library cannot be built")


def build(verbose: bool = False) -> str:
    cc = hipcc()
    tmp = tempfile.mkdtemp(prefix="calchem_hip_build_")
    try:
        objs = []
        for src in SOURCES:
            obj = os.path.join(tmp, src.replace(".hip", ".o"))
            cmd = [cc, *HIPCC_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
            r = subprocess.run(cmd, capture_output=True, text=True)
            if r.returncode != 0:
                sys.stderr.write(r.stdout + r.stderr)
                raise RuntimeError(f"hipcc failed on {src}")
            if verbose and r.stderr.strip():
                sys.stderr.write(r.stderr)
            objs.append(obj)
        out_tmp = LIB + ".tmp"
        cmd = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", out_tmp, *objs]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError("hipcc link of libcalchem_hip.so failed")
        os.replace(out_tmp, LIB)
        # the measurement probe: a library of its own, outside the Vxc interface
        probe_tmp = PROBE_LIB + ".tmp"
        cmd = [cc, *HIPCC_FLAGS, "-shared", "-o", probe_tmp, os.path.join(CSRC, PROBE_SOURCE)]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError("hipcc build of libcalchem_probe.so failed")
        os.replace(probe_tmp, PROBE_LIB)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
