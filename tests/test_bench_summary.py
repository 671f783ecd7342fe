"""bench.py's rocprofv3 summarizer on a synthetic kernel trace (CPU only).

The committed profile summaries and the bench line's `profiled_pass_us` come
from `bench.py --summarize`; this pins how it cuts the trace: the last
batched step runs from its tensor product up to the first roofline launch,
without the roofline's calibration copy; the roofline passes are the first
launches of the 960-limb shape; the 32-limb passes are every launch of the
small-transform kernels with 32 limbs.
"""
import csv
import os

import bench

FIELDS = ["Kind", "Agent_Id", "Queue_Id", "Stream_Id", "Thread_Id", "Dispatch_Id", "Kernel_Id", "Kernel_Name",
          "Correlation_Id", "Start_Timestamp", "End_Timestamp", "LDS_Block_Size", "Scratch_Size", "VGPR_Count",
          "Accum_VGPR_Count", "SGPR_Count", "Workgroup_Size_X", "Workgroup_Size_Y", "Workgroup_Size_Z",
          "Grid_Size_X", "Grid_Size_Y", "Grid_Size_Z"]


def write_trace(path, launches):
    """launches: (name, grid_x, grid_y, duration_ns), back to back with 1 us gaps."""
    t = 1_000_000
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=FIELDS)
        w.writeheader()
        for i, (name, gx, gy, d) in enumerate(launches):
            w.writerow({"Kind": "KERNEL_DISPATCH", "Agent_Id": 1, "Queue_Id": 1, "Stream_Id": 0, "Thread_Id": 1,
                        "Dispatch_Id": i + 1, "Kernel_Id": i, "Kernel_Name": name + "(args)", "Correlation_Id": i,
                        "Start_Timestamp": t, "End_Timestamp": t + d, "LDS_Block_Size": 0, "Scratch_Size": 0,
                        "VGPR_Count": 64, "Accum_VGPR_Count": 0, "SGPR_Count": 32, "Workgroup_Size_X": 256,
                        "Workgroup_Size_Y": 1, "Workgroup_Size_Z": 1, "Grid_Size_X": gx, "Grid_Size_Y": gy,
                        "Grid_Size_Z": 1})
            t += d + 1000


def test_summary_cuts_the_step_roofline_and_small_transforms(tmp_path):
    col, row = bench.ROOF_KERNELS
    scol, srow = bench.CFG2_KERNELS
    step = [("k_tensor_batch", 4096, 1177, 700_000), ("void k_intt_row<16, 0>", 4096, 1177, 400_000),
            ("void k_ntt_row_dot_mixed<16, 3, 1>", 4096, 1785, 2_000_000)]
    launches = [("k_tensor_batch", 4096, 10, 50_000)]           # an earlier (autotuner) step
    launches += [(scol, 4096, 32, 12_000), (srow, 4096, 32, 15_000)]
    launches += step + step                                       # the last step is the second
    launches += [("__amd_rocclr_copyBuffer", 65536, 1, 98_000)]   # calibration copy
    launches += [(col, 4096, 960, 230_000), (row, 4096, 960, 300_000)] * 12
    launches += [(col, 4096, 960, 999_000)]                       # beyond the roofline's 12 launches
    launches += [(scol, 4096, 32, 13_000), (srow, 4096, 32, 14_000)]
    d = tmp_path / "trace"
    os.makedirs(d)
    write_trace(str(d / "run_kernel_trace.csv"), launches)
    text, res = bench.summarize(str(d))
    lines = text.split("\n")
    head = next(l for l in lines if l.startswith("last batched step:"))
    assert head.startswith("last batched step: 3 launches, 3.100 ms")
    assert not any("copyBuffer" in l for l in lines[lines.index(head):lines.index(head) + 8])
    roof = next(l for l in lines if l.startswith("roofline launch"))
    assert "%s avg 230.00 us over 12" % col in roof and "%s avg 300.00 us over 12" % row in roof
    small = res["cfg2_32_limbs"]
    assert small["pass_durations_us"] == {scol: 12.5, srow: 14.5}
    assert small["launches"] == {scol: 2, srow: 2}
