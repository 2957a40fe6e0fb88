"""The C-ABI library loads and exports every entry point declared in include/*.h;
host-only functions (generators, argument checks) work without a GPU."""
import ctypes as C
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]


def _declared():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s+\**(prg_[a-z0-9_]+)\s*\(", text, re.M):
            names.add(m.group(1))
    return names


def test_every_declared_symbol_is_exported(prg):
    declared = _declared()
    assert len(declared) >= 25
    out = subprocess.run(["nm", "-D", "--defined-only", str(prg.LIB_PATH)], capture_output=True, text=True, check=True)
    exported = {l.split()[-1] for l in out.stdout.splitlines() if " T " in l}
    missing = declared - exported
    assert not missing, f"declared but not exported: {sorted(missing)}"
    lib = prg.load_library()
    for name in declared:
        assert getattr(lib, name) is not None
    # the Python mirror binds every declared function
    assert declared <= set(prg.SIGNATURES)


def test_config_default(prg):
    cfg = prg.PrgConfig()
    assert prg.lib().prg_config_default(C.byref(cfg)) == prg.OK
    assert cfg.max_depth == 8 and abs(cfg.cg_tol - 1e-6) < 1e-12 and cfg.cg_refresh == 20
    assert prg.lib().prg_config_default(None) == prg.E_ARG


def test_create_rejects_bad_depth(prg):
    cfg = prg.PrgConfig()
    prg.lib().prg_config_default(C.byref(cfg))
    h = C.c_void_p()
    for bad in (1, 11, -3):
        cfg.max_depth = bad
        assert prg.lib().prg_create(C.byref(h), C.byref(cfg)) == prg.E_ARG


def test_null_context_is_an_argument_error(prg):
    L = prg.lib()
    assert L.prg_build_octree(None) == prg.E_ARG
    assert L.prg_solve(None, None) == prg.E_ARG
    assert L.prg_last_error(None) == b"null context"


def test_synth_generators_are_deterministic(prg):
    a = prg.synth_sphere(1000, 1)
    b = prg.synth_sphere(1000, 1)
    c = prg.synth_sphere(1000, 2)
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
    assert not np.array_equal(a[0], c[0])
    r = np.linalg.norm(a[0].astype(np.float64) - 0.5, axis=1)
    assert np.allclose(r, 0.4, atol=1e-6)
    assert np.allclose(np.linalg.norm(a[1], axis=1), 1.0, atol=1e-6)
    t = prg.synth_torus(2000, 2)
    assert np.isfinite(t[0]).all() and np.allclose(np.linalg.norm(t[1], axis=1), 1.0, atol=1e-5)
    s = prg.synth_superquadrics(3000, 3, 0.001, 3)
    assert np.isfinite(s[0]).all() and np.allclose(np.linalg.norm(s[1], axis=1), 1.0, atol=1e-5)


def test_synth_rejects_bad_arguments(prg):
    p = np.empty((1, 3), np.float32)
    assert prg.lib().prg_synth_sphere(0, 1, p.ctypes.data, p.ctypes.data) == prg.E_ARG
    assert prg.lib().prg_synth_superquadrics(10, 0, 0.0, 1, p.ctypes.data, p.ctypes.data) == prg.E_ARG
