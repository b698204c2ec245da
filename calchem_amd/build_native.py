"""Build the gfx950 HIP library calchem_amd/libcalchem_hip.so with hipcc, and
the MFMA throughput probe calchem_amd/libcalchem_probe.so used by bench.py.

Every call compiles from source (no incremental reuse), so the library always
matches the tree; the objects compile side by side (MAX_JOBS caps the workers).  Also used by __graft_entry__.build().
"""
from __future__ import annotations

import concurrent.futures
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
SOURCES = ["api.hip", "vxc_kernels.hip", "grid.hip", "grad_kernels.hip", "xc_points.hip", "integrals.hip", "mlp.hip"]
ERI_SOURCE = "eri_classes.hip"
ERI_BRA_CLASSES = 6
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


def _compile_jobs(tmp: str):
    """(source, extra flags, object) for every object of the library."""
    jobs = [(src, [], os.path.join(tmp, src.replace(".hip", ".o"))) for src in SOURCES]
    # the repulsion kernels: one object per bra angular-momentum class
    for k in range(ERI_BRA_CLASSES):
        jobs.append((ERI_SOURCE, [f"-DERI_BRA_CLASS={k}"], os.path.join(tmp, f"eri_classes_{k}.o")))
    return jobs


def build(verbose: bool = False) -> str:
    cc = hipcc()
    tmp = tempfile.mkdtemp(prefix="calchem_hip_build_")
    try:
        jobs = _compile_jobs(tmp)

        def run(job):
            src, extra, obj = job
            cmd = [cc, *HIPCC_FLAGS, *extra, "-c", os.path.join(CSRC, src), "-o", obj]
            return job, subprocess.run(cmd, capture_output=True, text=True)

        workers = max(1, min(len(jobs), int(os.environ.get("MAX_JOBS", "0")) or (os.cpu_count() or 4)))
        with concurrent.futures.ThreadPoolExecutor(max_workers=workers) as pool:
            results = list(pool.map(run, jobs))
        for (src, extra, obj), r in results:
            if r.returncode != 0:
                sys.stderr.write(r.stdout + r.stderr)
                raise RuntimeError(f"hipcc failed on {src} {' '.join(extra)}")
            if verbose and r.stderr.strip():
                sys.stderr.write(r.stderr)
        objs = [obj for (_, _, obj), _ in results]
        out_tmp = LIB + ".tmp"
        cmd = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", out_tmp, *objs, "-lrocblas"]
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
