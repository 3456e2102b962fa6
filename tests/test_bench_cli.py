"""bench.py's command-line contract on the host: more than one GPU is refused with a message
(this build measures one GPU), and the option list names the side legs."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args],
                          capture_output=True, text=True, timeout=120)


def test_more_than_one_gpu_is_refused():
    r = run("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_help_lists_the_side_legs():
    r = run("--help")
    assert r.returncode == 0
    for leg in ("mpc_3d", "realfield_3d", "cavity_1d", "batched_1d", "adjoint_3d"):
        assert leg in r.stdout


def test_measured_traffic_is_read_from_the_profile():
    """The bench line's roofline.traffic comes from the committed counter summary."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("bench", os.path.join(ROOT, "bench.py"))
    b = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(b)
    t = b.measured_traffic(128, 1024)
    assert t is not None and 2.6e8 < t < 2.8e8   # ~2 complex64 fields of 128 x 128 x 1024
