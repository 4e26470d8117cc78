"""bench.py's command line on the CPU: it runs on one GPU only, so --gpus above 1 exits
with status 2 and says why, before touching a device."""
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2", "--steps", "1"],
                       capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr
    assert r.stdout.strip() == ""


def test_help_lists_the_timing_options():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--help"],
                       capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 0
    for opt in ("--steps", "--warmup", "--replays", "--formats", "--headline", "--batch"):
        assert opt in r.stdout
