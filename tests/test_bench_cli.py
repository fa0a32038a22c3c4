"""bench.py's command line without a GPU: more than one GPU is refused with a
message (every run here uses one MI355X), and the options parse."""
import pathlib
import subprocess
import sys

ROOT = pathlib.Path(__file__).resolve().parent.parent


def run(*args):
    return subprocess.run([sys.executable, str(ROOT / "bench.py"), *args], capture_output=True, text=True,
                          timeout=120, cwd=str(ROOT))


def test_more_than_one_gpu_is_refused():
    r = run("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU" in r.stderr and "--gpus 2" in r.stderr
    assert r.stdout == ""  # no JSON line


def test_help_lists_the_legs():
    r = run("--help")
    assert r.returncode == 0
    for opt in ("--steps", "--warmup", "--only", "--no-cpu-baseline", "--copy-calibration", "--composite-steps"):
        assert opt in r.stdout
