"""bench.py's command line on the CPU: it measures one GPU only, so any other
--gpus value exits with status 2 and says so, before touching a device; the
argument defaults are the ones its workload description states."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_defaults():
    sys.path.insert(0, ROOT)
    import bench

    old = sys.argv
    try:
        sys.argv = ["bench.py"]
        a = bench.parse()
    finally:
        sys.argv = old
    assert (a.gpus, a.batch, a.n_theta) == (1, 128, 256)
    assert a.launch == "graph" and a.graph_steps == 32 and a.groups == 3
    assert a.uneven_steps == 200 and a.large_steps == 500 and a.serve_requests == 10000
