"""The rocprofv3 CSV summariser that turns --pmc passes into HBM bytes per
dispatch (scaled by a copy kernel of known size)."""
import json

from axiomr_amd import rocprof_summary as RS

HDR = ("Correlation_Id,Dispatch_Id,Agent_Id,Queue_Id,Process_Id,Thread_Id,Grid_Size,Kernel_Id,Kernel_Name,"
       "Workgroup_Size,LDS_Block_Size,Scratch_Size,VGPR_Count,Accum_VGPR_Count,SGPR_Count,Counter_Name,"
       "Counter_Value,Start_Timestamp,End_Timestamp\n")


def _pmc(dirpath, rows):
    dirpath.mkdir(parents=True)
    lines = [HDR]
    for i, (kernel, counter, value) in enumerate(rows):
        lines.append(f'{i},{i},1,1,1,1,256,3,"{kernel}",256,0,0,32,0,16,{counter},{value},0,1\n')
    (dirpath / "run_counter_collection.csv").write_text("".join(lines))


def test_short_names():
    assert RS.short_name("void axr::k_raster<16, 256, false>(axr::KArgs, axr::Arena)") == "k_raster<16,256,false>"
    assert RS.short_name("k_scatter(axr::Arena, unsigned int)") == "k_scatter"


def test_traffic_scaling(tmp_path):
    k = "void axr::k_setup<16, false>(axr::KArgs, axr::Arena)"
    _pmc(tmp_path / "f", [(k, "FETCH_SIZE", 100.0), (k, "FETCH_SIZE", 120.0), (k, "FETCH_SIZE", 110.0)])
    _pmc(tmp_path / "w", [(k, "WRITE_SIZE", 50.0)])
    # copy of 1 MiB: fetch counter reads half, write counter exact
    _pmc(tmp_path / "cf", [("copy_kernel", "FETCH_SIZE", 512.0), ("fill", "FETCH_SIZE", 0.0)])
    _pmc(tmp_path / "cw", [("copy_kernel", "WRITE_SIZE", 1024.0)])
    res = RS.traffic(tmp_path / "f", tmp_path / "w", tmp_path / "cf", tmp_path / "cw", 1 << 20)
    assert res["copy_fetch"]["factor"] == 2048.0
    assert res["copy_write"]["factor"] == 1024.0
    rec = res["kernels"]["k_setup<16,false>"]
    assert rec["fetch_kib"] == 110.0          # median over dispatches
    assert rec["traffic_bytes"] == 110.0 * 2048 + 50.0 * 1024
    assert rec["dispatches"] == 3
    json.dumps(res)


def test_stats_markdown(tmp_path):
    d = tmp_path / "kt"
    d.mkdir()
    (d / "run_kernel_stats.csv").write_text(
        '"Name","Calls","TotalDurationNs","AverageNs","Percentage","MinNs","MaxNs","StdDev"\n'
        '"void axr::k_raster<16, 256, false>(axr::KArgs, axr::Arena)",10,300000,30000,75.0,20000,40000,1.0\n'
        '"k_clear(unsigned int*, float*, unsigned int*, unsigned long, unsigned int, float)",10,100000,10000,25.0,'
        '9000,11000,1.0\n')
    md = RS.stats_markdown(d, "t", "cmd")
    assert "| `k_raster<16,256,false>` | 10 | 30.00 |" in md
    assert md.index("k_raster") < md.index("k_clear")
