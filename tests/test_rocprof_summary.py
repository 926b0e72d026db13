"""The rocprof summary's joins on synthetic rocprofv3 CSV output (CPU)."""
import json

from fastmulticolsort_amd import rocprof_summary as rs


def _write(path, header, rows):
    path.write_text(",".join(header) + "\n" + "".join(",".join(str(x) for x in r) + "\n" for r in rows))


def test_roofline_joins_trace_counters_and_algorithmic_bytes(tmp_path):
    t, f, w = tmp_path / "t", tmp_path / "f", tmp_path / "w"
    for d in (t, f, w):
        d.mkdir()
    _write(t / "run_kernel_stats.csv", ["Name", "Calls", "TotalDurationNs", "AverageNs", "Percentage"],
           [["\"void fmcs::k_stitch<unsigned long>(int)\"", 3, 300000, 100000, 60.0],
            ["\"void fmcs::k_stitch<unsigned int>(int)\"", 3, 150000, 50000, 30.0],
            ["__amd_rocclr_copyBuffer", 2, 50000, 25000, 10.0]])
    hdr = ["Dispatch_Id", "Kernel_Name", "Grid_Size", "Counter_Value"]
    # the first copy is the 256 MiB probe (FETCH_SIZE counts half of it here);
    # a later, larger copy must not be taken for it
    _write(f / "run_counter_collection.csv", hdr,
           [[1, "__amd_rocclr_copyBuffer", 64, 131072], [2, "\"fmcs::k_stitch<unsigned long>(int)\"", 64, 1000],
            [3, "\"fmcs::k_stitch<unsigned int>(int)\"", 64, 3000], [4, "__amd_rocclr_copyBuffer", 64, 262144]])
    _write(w / "run_counter_collection.csv", hdr,
           [[1, "__amd_rocclr_copyBuffer", 64, 262144], [2, "\"fmcs::k_stitch<unsigned long>(int)\"", 64, 500],
            [3, "\"fmcs::k_stitch<unsigned int>(int)\"", 64, 1500], [4, "__amd_rocclr_copyBuffer", 64, 524288]])
    j = tmp_path / "b.jsonl"
    j.write_text(json.dumps({"config": 2, "kernels": {"k_stitch": {"calls_in_process": 6,
                                                                   "bytes_in_process": 6 * 1_000_000}}}) + "\n"
                 + json.dumps({"config": 4, "plans_run": [{"kernels": {}}]}) + "\n")
    md = rs.roofline(str(t), str(f), str(w), str(j), "T", "F", "W")
    row = [l for l in md.splitlines() if l.startswith("| k_stitch |")]
    assert len(row) == 1
    cells = [c.strip() for c in row[0].strip("|").split("|")]
    # 6 calls, 450 us total -> 75 us average; 1 MB a launch -> 13.3 GB/s
    assert cells[1] == "6" and cells[2] == "6" and cells[3] == "75.0" and cells[4] == "1.0"
    assert cells[5] == "13" and cells[6] == "0.002"
    # fetch: mean 2000 KiB x 2 (scale), write: mean 1000 KiB x 1 -> 5000 KiB
    assert cells[7] == f"{5000 * 1024 / 1e6:.1f}"
    assert "scale 2.000" in md and "scale 1.000" in md
