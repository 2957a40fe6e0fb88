"""The numpy warm-start U-Net (the oracle of the GPU network): its analytic
gradient against central finite differences of the exact (unrounded)
function, and its building blocks against direct definitions."""
import numpy as np
import pytest

from oracle import unet_ref as U


@pytest.fixture
def exact():
    U.ROUND = False
    yield
    U.ROUND = True


def _sample(G, seed):
    rng = np.random.default_rng(seed)
    x = np.zeros((G, G, G, U.CIN0), np.float32)
    x[..., 0] = rng.standard_normal((G, G, G))
    x[..., 1] = rng.random((G, G, G)) < 0.5
    t = rng.standard_normal((G, G, G)).astype(np.float32)
    m = (rng.random((G, G, G)) < 0.6).astype(np.float32)
    return x, t, m


def test_conv3_matches_direct_sum(exact):
    rng = np.random.default_rng(1)
    a = rng.standard_normal((4, 4, 4, 3))
    w = rng.standard_normal((27, 2, 3))
    out = U.conv3(a, w)
    x, y, z = 1, 3, 0
    ref = np.zeros(2)
    for t in range(27):
        dx, dy, dz = t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1
        u, v, s = x + dx, y + dy, z + dz
        if 0 <= u < 4 and 0 <= v < 4 and 0 <= s < 4:
            ref += w[t] @ a[u, v, s]
    assert np.allclose(out[x, y, z], ref)


def test_gradients_match_finite_differences(exact):
    G = 8
    params = U.init_params(3)
    for k in params:
        if k.endswith(".b"):
            params[k] = (np.random.default_rng(hash(k) % 1000).standard_normal(params[k].shape) * 0.1).astype(np.float32)
    params = {k: v.astype(np.float64) for k, v in params.items()}
    x, t, m = _sample(G, 5)
    loss, grads, _ = U.loss_and_grad(params, x, t, m)
    rng = np.random.default_rng(9)
    checked = 0
    for name in ["e0a.w", "e1b.w", "ba.w", "d2a.w", "d0b.w", "head.w", "d1a.b", "e2b.b", "head.b"]:
        shape = params[name].shape
        for _ in range(3):
            idx = tuple(rng.integers(0, s) for s in shape)
            h = 1e-6 * max(1.0, abs(params[name][idx]))
            old = params[name][idx]
            params[name][idx] = old + h
            lp = U.loss_and_grad(params, x, t, m)[0]
            params[name][idx] = old - h
            lm = U.loss_and_grad(params, x, t, m)[0]
            params[name][idx] = old
            fd = (lp - lm) / (2 * h)
            an = grads[name][idx]
            assert abs(fd - an) <= 1e-5 * max(1.0, abs(fd)) + 1e-7, (name, idx, fd, an)
            checked += 1
    assert checked == 27


def test_bf16_rounding_is_nearest_even():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.0e38], np.float32)
    r = U.bf16(v)
    assert r[0] == 1.0 and r[1] == 1.0  # tie to even
    assert r[2] == np.float32(1.0 + 2 ** -7)
    assert r[3] == -2.5
    assert np.isfinite(r[4]) or r[4] == np.inf
