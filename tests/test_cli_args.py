"""The CLI's argument handling (no GPU: every case exits before a context is
created)."""
import subprocess
from pathlib import Path

import pytest

EXE = Path(__file__).resolve().parents[1] / "poissonrecon_gpu_amd" / "poisson_recon"

pytestmark = pytest.mark.skipif(not EXE.exists(), reason="poisson_recon not built")


@pytest.mark.parametrize("args", [[], ["--in", "x.ply"], ["--out", "m.ply"], ["--bogus"], ["--in"],
                                  ["--in", "x.ply", "--out", "m.ply", "--weldd"], ["--in", "x.ply", "--out", "m.ply", "--nosubdivide"],
                                  ["--in", "x.ply", "--out", "m.ply", "--edgeCollapse", "0"],
                                  ["--in", "x.ply", "--out", "m.ply", "--triangleCollapse", "-1"],
                                  ["--in", "x.ply", "--out", "m.ply", "--edgeCollapse"]])
def test_usage_errors(args):
    r = subprocess.run([str(EXE), *args], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2
    assert "usage: poisson_recon" in r.stderr and "--no-subdivide" in r.stderr and "--weld" in r.stderr
    assert "--edgeCollapse" in r.stderr


def test_missing_input_file(tmp_path):
    r = subprocess.run([str(EXE), "--in", str(tmp_path / "none.ply"), "--out", str(tmp_path / "m.ply"),
                        "--subdivide", "--weld"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "cannot read" in r.stderr
    assert not (tmp_path / "m.ply").exists()
