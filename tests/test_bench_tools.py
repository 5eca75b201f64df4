"""bench.py's counter bookkeeping, on synthetic rocprofv3 counter files (CPU only): the copy
and gather calibrations give the byte scales, the traversal's FETCH_SIZE / WRITE_SIZE per
dispatch (several rows per dispatch are summed, other kernels ignored) give bytes per ray,
and the roofline's `traffic` is only reported for the traversal sources it was measured on."""
import csv
import json
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
import bench  # noqa: E402


def _counters(d: Path, rows):
    d.mkdir(parents=True)
    with open(d / "p_counter_collection.csv", "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"])
        w.writeheader()
        for r in rows:
            w.writerow(dict(zip(["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"], r)))


def _json(p: Path, obj):
    p.write_text("# a log line first\n" + json.dumps(obj) + "\n")


@pytest.fixture
def counter_dir(tmp_path, monkeypatch):
    monkeypatch.setattr(bench, "PMC_FILE", tmp_path / "pmc.json")
    d = tmp_path / "runs"
    K = bench.TRACE_KERNEL
    # copy: 5 launches moving 1 MiB each way; FETCH_SIZE counts half of it, WRITE_SIZE all
    d.mkdir()
    _json(d / "copy.json", {"copy_bytes_read": 1 << 20, "copy_bytes_written": 1 << 20, "launches": 5})
    _counters(d / "copy_fetch", [(i, "elementwise_kernel", "FETCH_SIZE", 512.0) for i in range(5)] +
              [(9, "tiny_kernel", "FETCH_SIZE", 1.0)])
    _counters(d / "copy_write", [(i, "elementwise_kernel", "WRITE_SIZE", 1024.0) for i in range(5)])
    # gather: 1000 distinct lines of 128 B plus 8 KiB of index per launch, counted as 100 KiB
    _json(d / "gather.json", {"gather_rows": 1024, "unique_rows": 1000, "row_bytes": 64, "unique_lines": 1000,
                              "line_bytes": 128, "index_bytes": 8192, "launches": 5})
    _counters(d / "gather_fetch", [(i, "vectorized_gather_kernel", "FETCH_SIZE", 100.0) for i in range(5)] +
              [(7, "index_copy", "FETCH_SIZE", 5000.0)])
    # traversal: two dispatches, each counter split over two rows; another kernel's rows ignored
    _json(d / "fetch.json", {"warmup": 0, "rays": {"extension": 1000}, "config": {"workload": "synthetic"}})
    _counters(d / "fetch", [(1, K, "FETCH_SIZE", 10.0), (1, K, "FETCH_SIZE", 30.0), (2, K, "FETCH_SIZE", 60.0),
                            (3, "k_shade<false>", "FETCH_SIZE", 999.0)])
    _counters(d / "write", [(1, K, "WRITE_SIZE", 5.0), (1, K, "WRITE_SIZE", 5.0), (2, K, "WRITE_SIZE", 10.0)])
    return d


def test_pmc_summary_scales_and_bytes_per_ray(counter_dir, capsys):
    assert bench.pmc_summary(counter_dir) == 0
    s = json.loads(bench.PMC_FILE.read_text())
    assert s["fetch_scale"] == 2.0 and s["write_scale"] == 1.0
    assert s["gather_fetch_scale"] == pytest.approx((1000 * 128 + 8192) / (100 * 1024), rel=1e-15)
    assert s["dispatches"] == 2 and s["extension_rays"] == 1000
    # (40 + 60) KiB fetched x 2 + (10 + 10) KiB written x 1, over 1000 rays
    assert s["hbm_bytes_per_ray"] == pytest.approx(1024 * (100 * 2.0 + 20 * 1.0) / 1000, rel=1e-15)
    assert s["hbm_bytes_per_ray_gather_scaled"] == pytest.approx(
        1024 * (100 * s["gather_fetch_scale"] + 20) / 1000, rel=1e-15)
    assert s["source_hash"] == bench.traversal_source_hash()
    assert "| 1 | 40 | 10 |" in capsys.readouterr().out


def test_traffic_only_for_the_sources_it_was_measured_on(counter_dir):
    bench.pmc_summary(counter_dir)
    per_ray = json.loads(bench.PMC_FILE.read_text())["hbm_bytes_per_ray"]
    traffic, why = bench._pmc_traffic(5e6)
    assert traffic == pytest.approx(per_ray * 5e6, rel=1e-15) and "source" in why
    s = json.loads(bench.PMC_FILE.read_text())
    s["source_hash"] = "0" * 16
    bench.PMC_FILE.write_text(json.dumps(s))
    traffic, why = bench._pmc_traffic(5e6)
    assert traffic is None and why.startswith("stale")


def test_counter_run_must_count_every_launch(counter_dir):
    _json(counter_dir / "fetch.json", {"warmup": 3, "rays": {"extension": 1000}, "config": {"workload": "x"}})
    with pytest.raises(SystemExit, match="warmup 0"):
        bench.pmc_summary(counter_dir)
