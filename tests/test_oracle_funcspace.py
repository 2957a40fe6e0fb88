"""Pins the oracle's function-space restatement (A9; FunctionData.inl:111-300).

The reference ships no tables and its host code is not buildable here, so the
oracle's float polynomial arithmetic is checked against:
  1. known answers obtained by running the reference's own host code
     (recorded with their source in tests/golden/funcspace_known_answers.json);
  2. exact rational inner products of the B-spline basis (the float tables
     carry ~1e-5 relative error, far below the size of a dropped term, a wrong
     sign or a transposed index);
  3. translation invariance, symmetry of dot/d2Dot and antisymmetry of dDot
     off the diagonal.
"""
import json
from fractions import Fraction as Fr
from pathlib import Path

import numpy as np
import pytest

import oracle as orc

GOLDEN = Path(__file__).resolve().parent / "golden"

# exact F = (4/3) * quadratic B-spline on [-3/2, 3/2] (F(0) = 1): pieces (lo, hi, coefficients in x)
F_EXACT = [(Fr(-3, 2), Fr(-1, 2), [Fr(3, 2), Fr(2), Fr(2, 3)]),
           (Fr(-1, 2), Fr(1, 2), [Fr(1), Fr(0), Fr(-4, 3)]),
           (Fr(1, 2), Fr(3, 2), [Fr(3, 2), Fr(-2), Fr(2, 3)])]


def _pmul(a, b):
    r = [Fr(0)] * (len(a) + len(b) - 1)
    for i, x in enumerate(a):
        for j, y in enumerate(b):
            r[i + j] += x * y
    return r


def _affine(p, a, b):  # p(a x + b)
    out, powk = [Fr(0)], [Fr(1)]
    for c in p:
        n = max(len(out), len(powk))
        out = [(out[i] if i < len(out) else 0) + (c * powk[i] if i < len(powk) else 0) for i in range(n)]
        powk = _pmul(powk, [b, a])
    return out


def _basis(depth, off, deriv=False):
    w = Fr(1, 2 ** depth)
    c = (Fr(1, 2) + off) * w
    out = []
    for lo, hi, p in F_EXACT:
        q = [p[i] * i for i in range(1, len(p))] if deriv else p
        q = _affine(q, 1 / w, -c / w)
        if deriv:
            q = [v / w for v in q]
        out.append((c + lo * w, c + hi * w, q))
    return out


def _inner(A, B):
    s = Fr(0)
    for a0, a1, p in A:
        for b0, b1, q in B:
            lo, hi = max(a0, b0), min(a1, b1)
            if lo < hi:
                pq = _pmul(p, q)
                s += sum(cf * (hi ** (i + 1) - lo ** (i + 1)) / (i + 1) for i, cf in enumerate(pq))
    return s


def exact_entry(which, da, oa, db, ob):
    """<B_a, B_b>, <B_a', B_b>, <B_a', B_b'> - the quantities the reference tabulates."""
    if which == 0:
        return float(_inner(_basis(da, oa), _basis(db, ob)))
    if which == 1:
        return float(_inner(_basis(da, oa, True), _basis(db, ob)))
    return float(_inner(_basis(da, oa, True), _basis(db, ob, True)))


def test_known_answers_from_reference_host_code():
    ka = json.loads((GOLDEN / "funcspace_known_answers.json").read_text())
    D = ka["max_depth"]
    for name, which in (("dot_0_0", 0), ("d2dot_0_0", 2)):
        v = orc.fs_entry(D, which, 0, 0, 0, 0)
        assert v == pytest.approx(ka[name], rel=ka["rel_tol"]), name
    assert orc.fs_entry(D, 1, 0, 0, 0, 0) == pytest.approx(ka["ddot_0_0"], rel=ka["ddot_rel_tol"])
    # the sign of dDot[0,0] is rounding residue (exact value 0) and matches the reference
    assert np.sign(orc.fs_entry(D, 1, 0, 0, 0, 0)) == np.sign(ka["ddot_0_0"])
    for depth_key, vals in ka["stencils"].items():
        st = orc.fs_stencil(D, int(depth_key))
        for name, j in (("centre", 13), ("face", 4), ("edge", 1), ("corner", 0)):
            assert st[j] == pytest.approx(vals[name], rel=ka["rel_tol"]), (depth_key, name)


@pytest.mark.parametrize("which", [0, 1, 2])
@pytest.mark.parametrize("da,oa,db,ob", [
    (8, 5, 8, 5), (8, 5, 8, 6), (8, 6, 8, 5), (8, 7, 8, 9), (4, 5, 4, 6), (5, 3, 8, 20), (5, 3, 8, 30),
    (3, 2, 8, 70), (0, 0, 8, 100), (8, 100, 5, 10), (2, 1, 8, 150), (7, 40, 8, 81), (6, 20, 8, 83)])
def test_entries_match_exact_inner_products(which, da, oa, db, ob):
    D = 8
    v = orc.fs_entry(D, which, da, oa, db, ob)
    e = exact_entry(which, da, oa, db, ob)
    # Float error scales with the largest entries of the same kind and depth pair
    # (and grows with the depth gap: coefficients are scaled by 2^(k*gap)); a
    # dropped term, a sign flip or an index shifted by one is a change of the
    # order of the family's largest entry.
    step = 2 ** abs(db - da)
    scale = max(abs(exact_entry(which, da, oa, db, ob + k)) for k in range(-2 * step, 2 * step + 1, max(1, step // 4)))
    tol = 1e-4 if da == db else 1e-3
    assert abs(v - e) <= tol * max(scale, 1e-12), (v, e)


def test_translation_invariance_and_symmetry():
    D = 9
    rng = np.random.default_rng(0)
    for which in (0, 1, 2):
        for _ in range(60):
            d = int(rng.integers(2, D + 1))
            o = int(rng.integers(1, 2 ** d - 1))
            s = int(rng.integers(-1, 2))
            base = orc.fs_entry(D, which, d, o, d, o + s)
            other = int(rng.integers(1, 2 ** d - 1)) if 2 ** d > 3 else o
            if 0 <= other + s < 2 ** d:
                assert orc.fs_entry(D, which, d, other, d, other + s) == base
            sw = orc.fs_entry(D, which, d, o + s, d, o)
            if which == 1 and s != 0:
                assert sw == -base
            elif which != 1:
                assert sw == base
    # coarse-node / leaf pairs depend on the relative offset only
    for d in range(0, D):
        sc = 2 ** (D - d)
        for u in range(0, 3 * sc, max(1, sc // 4)):
            a = orc.fs_entry(D, 1, d, 0, D, u - sc)
            if 2 ** d > 2:
                assert orc.fs_entry(D, 1, d, 1, D, u) == a


def test_base_function_values_track_the_exact_spline():
    # float basis evaluation is ill-conditioned but must follow the spline
    D = 6
    for idx in (0, 1, 5, 30, 64, 100):
        depth = int(np.floor(np.log2(idx + 1)))
        off = idx + 1 - 2 ** depth
        w = 1.0 / 2 ** depth
        c = (0.5 + off) * w
        for t in np.linspace(-1.4, 1.4, 15):
            x = np.float32(c + t * w)
            exact = float(sum(sum(cf * Fr(float(x)) ** i for i, cf in enumerate(p))
                              for lo, hi, p in _basis(depth, off) if lo < Fr(float(x)) <= hi))
            assert abs(orc.fs_value(D, idx, x) - exact) <= 2e-3 * 4 ** depth * 1e-2 + 1e-3


def test_stencil_rows_sum_positive_and_symmetric():
    for D in (8, 10):
        for d in range(D + 1):
            st = orc.fs_stencil(D, d).reshape(3, 3, 3)
            assert np.array_equal(st, st[::-1, :, :]) and np.array_equal(st, st[:, ::-1, :])
            assert np.array_equal(st, st.transpose(1, 0, 2)) and np.array_equal(st, st.transpose(2, 1, 0))
            assert st[1, 1, 1] > 0
            # scales as 2^-d (truncated Galerkin Laplacian of the B-spline basis)
            if d > 0:
                prev = orc.fs_stencil(D, d - 1)[13]
                assert st[1, 1, 1] == pytest.approx(prev / 2, rel=1e-5)
