"""The benchmark's rocprofv3 summary helpers (bench.py), on small synthetic
CSV files shaped like rocprofv3's: kernel names as the tracer prints them,
per-kernel statistics, counters summed per dispatch, and the copy
calibration that scales FETCH_SIZE / WRITE_SIZE."""
import bench


def test_short_kernel_names():
    sk = bench._short_kernel
    assert sk('"void rt::k_fused_o8<1>(rt::UnitDesc const*, int)"') == "k_fused_o8<1>"
    assert sk("rt::(anonymous namespace)::k_bi_edge<64, 16, 1024>(unsigned char const*, int)") \
        == "k_bi_edge<64, 16, 1024>"
    assert sk("void rt::(anonymous namespace)::k_bi_init(rt::StreamTable, int, unsigned int)") \
        == "k_bi_init"
    assert sk("rt::k_pair_lozenge(rt::UnitDesc const*, rt::WorkItem const*, int)") \
        == "k_pair_lozenge"
    assert sk("__amd_rocclr_copyBuffer") == "__amd_rocclr_copyBuffer"


def _write(path, header, rows):
    with open(path, "w") as f:
        f.write(",".join(f'"{h}"' for h in header) + "\n")
        for r in rows:
            f.write(",".join(f'"{v}"' if isinstance(v, str) else str(v) for v in r) + "\n")


def test_kernel_stats_and_counters(tmp_path):
    d = tmp_path / "run"
    d.mkdir()
    _write(d / "run_kernel_stats.csv",
           ["Name", "Calls", "TotalDurationNs", "AverageNs", "Percentage", "MinNs", "MaxNs",
            "StdDev"],
           [["rt::(anonymous namespace)::k_bi_edge<64, 16, 1024>(int)", 4, 80000, 20000.0,
             99.0, 19000, 21000, 1.0]])
    st = bench._kernel_stats(str(tmp_path))
    assert st["k_bi_edge<64, 16, 1024>"] == {"calls": 4, "total_ms": 0.08, "avg_us": 20.0,
                                             "min_us": 19.0, "max_us": 21.0, "pct": 99.0}
    # counters: two rows (two agents / dimensions) of one dispatch are summed
    _write(d / "run_counter_collection.csv",
           ["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"],
           [[1, "k_a(int)", "FETCH_SIZE", 10.0], [1, "k_a(int)", "FETCH_SIZE", 5.0],
            [2, "k_a(int)", "FETCH_SIZE", 20.0], [3, "__amd_rocclr_copyBuffer", "FETCH_SIZE",
                                                  1.0],
            [4, "__amd_rocclr_copyBuffer", "FETCH_SIZE", 524288.0],
            [5, "k_a(int)", "WRITE_SIZE", 7.0]])
    assert sorted(bench._counter_dispatches(str(tmp_path), "FETCH_SIZE")["k_a"]) == [15.0, 20.0]
    avg, n = bench._counter_per_dispatch(str(tmp_path), "FETCH_SIZE")["k_a"]
    assert (avg, n) == (17.5, 2)
    # the copy calibration takes the largest copy dispatches, not the small one
    assert bench._copy_kib(str(tmp_path), "FETCH_SIZE", copies=1) == 524288.0
