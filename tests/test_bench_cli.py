"""bench.py's command-line contract on the CPU: more than one GPU is refused
with a message and exit status 2, before anything touches a device, and
--horlimit without tracers is a usage error."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                          timeout=300, cwd=ROOT)


def test_more_than_one_gpu_is_refused():
    r = _run("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout.strip() == ""


def test_horlimit_needs_tracers():
    r = _run("--horlimit")
    assert r.returncode == 2
    assert "--horlimit needs tracers" in r.stderr
