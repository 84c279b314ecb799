"""The rocprofv3 CSV summariser (host logic, no GPU)."""
import csv

from ecsprt_amd import profile_summary as ps


def write(path, header, rows):
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(header)
        w.writerows(rows)


def test_stats_and_step_breakdown(tmp_path):
    d = tmp_path / "run"
    d.mkdir()
    write(d / "run_kernel_stats.csv", ["Name", "Calls", "TotalDurationNs", "AverageNs", "Percentage", "MinNs", "MaxNs"],
          [["void ecsprt::k_tile<31, 8, 8, 4, 2>(ecsprt::TileArgs)", 3, 30e6, 10e6, 90.0, 9.5e6, 10.5e6],
           ["k_sweep_z<double, 32, 16>", 10, 3.3e5, 3.3e4, 10.0, 3.0e4, 3.5e4]])
    rows, t = [], 0
    for _ in range(3):
        for name, dur in (("void ecsprt::k_tile<31, 8, 8, 4, 2>(ecsprt::TileArgs)", 10_000_000),
                          ("k_sweep_z<double, 32, 16>", 33_000), ("k_sweep_z<double, 32, 16>", 33_000)):
            rows.append([name, t, t + dur])
            t += dur + 1000
    write(d / "run_kernel_trace.csv", ["Kernel_Name", "Start_Timestamp", "End_Timestamp"], rows)
    out = tmp_path / "s.txt"
    ps.stats(str(tmp_path), str(out), "cmd", 10, "k_tile<", drop=True)
    text = out.read_text()
    assert "command: cmd" in text
    assert "k_tile<31, 8, 8, 4, 2>" in text and "10000.00" in text
    assert "one step (from the last-but-one launch of k_tile<): 3 launches" in text
    assert not list(tmp_path.rglob("*.csv"))


def test_pmc_means(tmp_path):
    write(tmp_path / "run_counter_collection.csv", ["Kernel_Name", "Counter_Name", "Counter_Value"],
          [["k_tile_mx<31, 8, 8, 8, 2>", "FETCH_SIZE", 10.0], ["k_tile_mx<31, 8, 8, 8, 2>", "FETCH_SIZE", 30.0]])
    out = tmp_path / "p.txt"
    ps.pmc(str(tmp_path), str(out), "cmd", drop=False)
    line = [l for l in out.read_text().splitlines() if l.startswith("k_tile_mx")][0]
    assert line.split()[-2:] == ["2", "20.0"]


def test_short_names_drop_anonymous_namespaces():
    assert ps.short("void ecsprt::(anonymous namespace)::k_push2d(long, double const*)") == "k_push2d"
    assert ps.short("void ecsprt::k_tile<31, 8, 8, 4, 2>(ecsprt::TileArgs)") == "k_tile<31, 8, 8, 4, 2>"
