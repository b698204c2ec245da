"""bench.py's digests of rocprofv3 counter runs (the summaries its roofline
and VALU fractions read): a synthetic counter file with known totals."""
import csv
import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402

FIELDS = ["Correlation_Id", "Dispatch_Id", "Agent_Id", "Queue_Id", "Process_Id", "Thread_Id", "Grid_Size",
          "Kernel_Id", "Kernel_Name", "Workgroup_Size", "LDS_Block_Size", "Scratch_Size", "VGPR_Count",
          "Accum_VGPR_Count", "SGPR_Count", "Counter_Name", "Counter_Value", "Start_Timestamp", "End_Timestamp"]


def _write(path, rows):
    os.makedirs(path, exist_ok=True)
    with open(os.path.join(path, "pass_counter_collection.csv"), "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=FIELDS)
        w.writeheader()
        for i, (name, grid, counter, value) in enumerate(rows):
            w.writerow({"Correlation_Id": i, "Dispatch_Id": i, "Agent_Id": "Agent 1", "Queue_Id": 1,
                        "Process_Id": 1, "Thread_Id": 1, "Grid_Size": grid, "Kernel_Id": 1, "Kernel_Name": name,
                        "Workgroup_Size": 256, "LDS_Block_Size": 0, "Scratch_Size": 0, "VGPR_Count": 64,
                        "Accum_VGPR_Count": 0, "SGPR_Count": 32, "Counter_Name": counter,
                        "Counter_Value": value, "Start_Timestamp": 0, "End_Timestamp": 1})


def test_stage_totals_and_copy_calibration(tmp_path, monkeypatch):
    digest = bench.sources_digest()
    monkeypatch.setattr(bench, "ROOT", str(tmp_path))
    monkeypatch.setattr(bench, "sources_digest", lambda: digest)
    os.makedirs(tmp_path / "profiles")
    # two rho stages of two classes each (grid sizes 512 and 256), one
    # calibration copy reading half its bytes through FETCH_SIZE
    fetch = [("__amd_rocclr_copyBuffer", 64, "FETCH_SIZE", 524288.0)]
    write = [("__amd_rocclr_copyBuffer", 64, "WRITE_SIZE", 1048576.0)]
    for _ in range(2):
        fetch += [("void chx::k_rho_aoc2<1>(chx::RhoArgs)", 512, "FETCH_SIZE", 3000.0),
                  ("void chx::k_rho_aoc2<1>(chx::RhoArgs)", 256, "FETCH_SIZE", 1000.0),
                  ("void chx::k_rho<double, false, 1, 0>(chx::RhoArgs)", 512, "FETCH_SIZE", 50.0)]
        write += [("void chx::k_rho_aoc2<1>(chx::RhoArgs)", 512, "WRITE_SIZE", 30.0),
                  ("void chx::k_rho_aoc2<1>(chx::RhoArgs)", 256, "WRITE_SIZE", 10.0),
                  ("void chx::k_rho<double, false, 1, 0>(chx::RhoArgs)", 512, "WRITE_SIZE", 5.0)]
        # V stage: two group kernels launched on the same grid size, one launch each per build
        fetch += [("void chx::k_vxc_group_pipe<17, true>(chx::VArgs)", 1024, "FETCH_SIZE", 700.0),
                  ("void chx::k_vxc_group_pipe<8, false>(chx::VArgs)", 1024, "FETCH_SIZE", 300.0)]
        write += [("void chx::k_vxc_group_pipe<17, true>(chx::VArgs)", 1024, "WRITE_SIZE", 70.0),
                  ("void chx::k_vxc_group_pipe<8, false>(chx::VArgs)", 1024, "WRITE_SIZE", 30.0)]
    _write(str(tmp_path / "mf"), fetch)
    _write(str(tmp_path / "mw"), write)
    bench.digest_traffic(str(tmp_path / "mf"), str(tmp_path / "mw"), "water20", "test")
    t = json.load(open(tmp_path / "profiles" / "k_rho_water20_fp32_traffic.json"))
    assert t["stages"] == {"fetch": 2, "write": 2}
    assert t["fetch_kib_per_build_raw"] == 4000.0 and t["write_kib_per_build_raw"] == 40.0
    # FETCH_SIZE scaled by 2 (the copy's 1 GiB read as 512 MiB), WRITE_SIZE by 1
    assert t["hbm_bytes_per_build"] == pytest.approx((4000.0 * 2 + 40.0) * 1024)
    assert t["sources_sha256"] == digest
    f64 = json.load(open(tmp_path / "profiles" / "k_rho_water20_fp64_traffic.json"))
    assert f64["hbm_bytes_per_build"] == pytest.approx((50.0 * 2 + 5.0) * 1024)
    v = json.load(open(tmp_path / "profiles" / "k_vxc_group_water20_fp32_traffic.json"))
    assert v["kernel"] == "k_vxc_group" and v["stages"] == {"fetch": 2, "write": 2}
    assert v["hbm_bytes_per_build"] == pytest.approx((1000.0 * 2 + 100.0) * 1024)
    assert not os.path.exists(tmp_path / "profiles" / "k_vxc_group_water20_fp64_traffic.json")


def test_uneven_class_dispatches_are_refused(tmp_path):
    rows = [("void chx::k_rho_aoc2<1>(chx::RhoArgs)", 512, "FETCH_SIZE", 1.0),
            ("void chx::k_rho_aoc2<1>(chx::RhoArgs)", 256, "FETCH_SIZE", 1.0),
            ("void chx::k_rho_aoc2<1>(chx::RhoArgs)", 512, "FETCH_SIZE", 1.0)]
    _write(str(tmp_path / "x"), rows)
    with pytest.raises(ValueError):
        bench._per_stage(bench._pmc_rows(str(tmp_path / "x")), "FETCH_SIZE", ("k_rho_aoc2<1>",))
