"""bench.py's one-GPU contract (no GPU needed: it exits before any device
work): more than one GPU, asked by flag or by a launcher's WORLD_SIZE, is
refused with a message and exit status 2."""
import os
import subprocess
import sys
from pathlib import Path

import pytest

BENCH = Path(__file__).resolve().parents[1] / "bench.py"


@pytest.mark.parametrize("argv,world", [(["--gpus", "2"], "1"), (["--gpus", "8"], "1"), ([], "2")])
def test_more_than_one_gpu_refused(argv, world):
    env = dict(os.environ, WORLD_SIZE=world)
    r = subprocess.run([sys.executable, str(BENCH), *argv], capture_output=True, text=True, timeout=120, env=env)
    assert r.returncode == 2
    assert "measures one GPU" in r.stderr
    assert r.stdout == ""
