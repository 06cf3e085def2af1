"""bench.py's command-line contract on the CPU: one GPU only, and the counted
HBM traffic applied only to a run with the profiled arguments."""
import argparse
import importlib.util
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)


def test_counted_traffic_matches_profiled_arguments(tmp_path, monkeypatch):
    b = _bench()
    (tmp_path / "profiles").mkdir()
    (tmp_path / b.TRAFFIC_SUMMARY).write_text(
        "text\n    traffic_bytes_per_launch precond 1234.5 config=cfg1 inner_tol=0.0003 spmv_variant=None\n"
        "traffic_bytes_per_launch spmv 99 config=cfg1\n")
    monkeypatch.setattr(b, "ROOT", str(tmp_path))
    a = argparse.Namespace(config="cfg1", inner_tol=3e-4, spmv_variant=None)
    assert b.counted_traffic("precond", a) == 1234.5
    assert b.counted_traffic("spmv", a) == 99.0
    assert b.counted_traffic("other", a) is None
    a.inner_tol = 1e-4                      # a different launch sequence
    assert b.counted_traffic("precond", a) is None


def test_committed_traffic_summary_parses():
    """The committed counter summary carries a figure for the default run."""
    b = _bench()
    path = os.path.join(ROOT, b.TRAFFIC_SUMMARY)
    lines = [l.split() for l in open(path) if l.split()[:1] == ["traffic_bytes_per_launch"]]
    assert {w[1] for w in lines} >= {"precond", "spmv"}
    for w in lines:
        assert float(w[2]) > 0
        assert all("=" in kv for kv in w[3:])
