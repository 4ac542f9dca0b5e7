"""The counter reduction behind profiles/traffic_*.json: calibration on the
copy probe, kernel kinds recognised from the step-flag template argument,
empty graph nodes left out, and per-mode merging under a key."""
import csv
import json

from torchwi_amd import profiling as P


def _write(path, counter, rows):
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Kernel_Name", "Counter_Name", "Counter_Value", "Grid_Size"])
        w.writeheader()
        for name, v in rows:
            w.writerow({"Kernel_Name": name, "Counter_Name": counter, "Counter_Value": v, "Grid_Size": 256})


def test_summarize_calibrates_on_the_copy_and_skips_empty_nodes(tmp_path):
    cells = 1000
    copy_mib = 1
    # FETCH_SIZE reads half the bytes, WRITE_SIZE all of them (as measured on MI355X)
    fetch = [("__amd_rocclr_copyBuffer", copy_mib * 1024 / 2),
             ("void tw::k_vstep<4, 24u, 1, 16>(...)", 10.0), ("void tw::k_vstep<4, 24u, 1, 16>(...)", 10.0),
             ("void tw::k_vstep<4, 24u, 1, 16>(...)", 0.1),          # a graph node past the loop's end
             ("void tw::k_pipe<4, 3u, 3>(...)", 4.0), ("at::native::something", 99.0)]
    write = [("__amd_rocclr_copyBuffer", copy_mib * 1024),
             ("void tw::k_vstep<4, 24u, 1, 16>(...)", 5.0), ("void tw::k_vstep<4, 24u, 1, 16>(...)", 5.0),
             ("void tw::k_pipe<4, 3u, 3>(...)", 3.0)]
    fc, wc = str(tmp_path / "f.csv"), str(tmp_path / "w.csv")
    _write(fc, "FETCH_SIZE", fetch)
    _write(wc, "WRITE_SIZE", write)
    out = P.summarize(fc, wc, copy_mib, cells, "synthetic")
    assert out["calibration"]["fetch_factor"] == 2.0 and out["calibration"]["write_factor"] == 1.0
    adj = out["kernels"]["adj"]
    assert adj["read_bytes"] == 2 * 10 * 1024 and adj["write_bytes"] == 5 * 1024
    assert adj["algorithmic_bytes_per_cell"] == 28 and adj["steps_per_launch"] == 1
    fwd = out["kernels"]["fwd"]
    assert fwd["algorithmic_bytes_per_cell"] == 20 and fwd["steps_per_launch"] == 3
    assert fwd["bytes_per_cell"] == round((2 * 4 + 3) * 1024 / cells, 2)


def test_main_merges_modes_under_keys(tmp_path):
    fc, wc = str(tmp_path / "f.csv"), str(tmp_path / "w.csv")
    _write(fc, "FETCH_SIZE", [("copyBuffer", 512), ("tw::k_vstep<4, 7u, 1, 16>", 8.0)])
    _write(wc, "WRITE_SIZE", [("copyBuffer", 1024), ("tw::k_vstep<4, 7u, 1, 16>", 4.0)])
    out = str(tmp_path / "t.json")
    assert P.main([fc, wc, "1", "100", out, "w", "full_storage"]) == 0
    assert P.main([fc, wc, "1", "100", out, "w", "checkpointed"]) == 0
    d = json.load(open(out))
    assert set(d["full_storage"]) == {"fwd_virt"} and d["checkpointed"] == d["full_storage"]
    assert d["cells_per_launch"] == 100 and "kernels" not in d
