"""The learned warm start's U-Net on the GPU (BF16 matrix-core convolutions,
FP32 accumulation) against its numpy oracle (oracle/unet_ref.py, which
mirrors the GPU's BF16 rounding points and accumulates in float64), and
the warm start inside the solver."""
import numpy as np
import pytest

from oracle import unet_ref as U

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)


def _sample(G, seed):
    rng = np.random.default_rng(seed)
    b = rng.standard_normal((G, G, G)).astype(np.float32)
    m = (rng.random((G, G, G)) < 0.4).astype(np.float32)
    t = rng.standard_normal((G, G, G)).astype(np.float32)
    return b, m, t


@pytest.fixture(scope="module")
def net16(prg):
    net = prg.UNet(grid=16, seed=7)
    p = net.get_params()
    assert p.size == U.num_params()
    d = U.unflatten(p)
    rng = np.random.default_rng(3)
    for k in d:
        if k.endswith(".b"):
            d[k] = (rng.standard_normal(d[k].shape) * 0.05).astype(np.float32)
    net.set_params(U.flatten(d))
    yield net, d
    net.close()


def test_forward_matches_oracle(net16):
    net, d = net16
    print("\n[unet forward G=16]", flush=True)
    b, m, _ = _sample(16, 1)
    x = np.zeros((16, 16, 16, U.CIN0), np.float32)
    x[..., 0] = b
    x[..., 1] = m
    ref, _ = U.forward(d, x)
    out = net.predict(b, m, 1.0)
    err = _rel(out, ref)
    print(f"prediction rel L2 {err:.2e}")
    assert err <= 2e-3


def test_gradients_match_oracle(net16):
    net, d = net16
    print("\n[unet gradients G=16]", flush=True)
    b, m, t = _sample(16, 2)
    x = np.zeros((16, 16, 16, U.CIN0), np.float32)
    x[..., 0] = b
    x[..., 1] = m
    loss_ref, grads_ref, _ = U.loss_and_grad(d, x, t, m)
    loss, g = net.gradients(b, m, t, 1.0)
    assert abs(loss - loss_ref) <= 1e-3 * abs(loss_ref)
    gd = U.unflatten(g)
    worst = 0.0
    for name, _ in U.param_layout():
        err = _rel(gd[name], grads_ref[name])
        worst = max(worst, err)
        assert err <= 3e-2, (name, err)
    print(f"loss {loss:.6f} (oracle {loss_ref:.6f}); worst gradient rel L2 {worst:.2e}")


def test_training_reduces_loss(prg):
    print("\n[unet training G=16]", flush=True)
    net = prg.UNet(grid=16, seed=11)
    try:
        rng = np.random.default_rng(5)
        G = 16
        B = 4
        # a learnable map: the target is a smoothed copy of the input on the mask
        bs = rng.standard_normal((B, G, G, G)).astype(np.float32)
        ms = np.ones((B, G, G, G), np.float32)
        ts = 0.5 * (bs + np.roll(bs, 1, axis=1) + np.roll(bs, -1, axis=2)) / 3.0
        first = net.train_step(bs, ms, ts, np.ones(B, np.float32), 2e-3)
        for _ in range(80):
            last = net.train_step(bs, ms, ts, np.ones(B, np.float32), 2e-3)
        print(f"loss {first:.4f} -> {last:.4f}")
        assert last < 0.6 * first
    finally:
        net.close()


def test_learned_warm_start_solves_to_tolerance(prg):
    print("\n[warm start: sphere N=35947 D=7, network on depth 5]", flush=True)
    p, n = prg.synth_sphere(35947, 1)
    net = prg.UNet(grid=32, seed=3)
    plain = prg.Reconstructor(max_depth=7)
    warm = prg.Reconstructor(max_depth=7, warm_start=prg.WARM_LEARNED)
    try:
        warm.set_warm_start_net(net)
        for g in (plain, warm):
            g.set_points(p, n)
            g.build_octree()
            g.compute_divergence()
        sp, sw = plain.solve(), warm.solve()
        for s in sp + sw:
            assert s["true_rel_residual"] <= 2e-6
        # the untrained network's start changes the depth-5 trajectory only
        assert [s["iters"] for s in sp][:5] == [s["iters"] for s in sw][:5]
        assert [s["iters"] for s in sp][6:] == [s["iters"] for s in sw][6:]
        xw = warm.read_buffer(prg.BUF_SOLUTION)
        xp = plain.read_buffer(prg.BUF_SOLUTION)
        counts, bases = plain.level_info()
        sl = slice(int(bases[5]), int(bases[5] + counts[5]))
        assert _rel(xw[sl], xp[sl]) <= 1e-4
    finally:
        warm.close()
        plain.close()
        net.close()


@pytest.mark.parametrize("grid", [16, 64])
def test_worse_than_zero_warm_start_falls_back(prg, monkeypatch, grid):
    """A warm start with ||b - A x0|| > ||b|| is replaced by x0 = 0: with the
    prediction scaled by 1e6 (test setting PRG_WARM_START_SCALE), the solve
    equals the plain solve bit for bit, and prg_warm_start_info reports the
    rejection.  grid 16 warm-starts depth 4 (the one-workgroup small-depth
    CG), grid 64 depth 6 (the multi-kernel CG of the large depths)."""
    print(f"\n[warm start fallback: sphere N=35947 D=8, network grid {grid}]", flush=True)
    p, n = prg.synth_sphere(35947, 1)
    net = prg.UNet(grid=grid, seed=3)
    plain = prg.Reconstructor(max_depth=8)
    monkeypatch.setenv("PRG_WARM_START_SCALE", "1e6")
    warm = prg.Reconstructor(max_depth=8, warm_start=prg.WARM_LEARNED)
    try:
        warm.set_warm_start_net(net)
        for g in (plain, warm):
            g.set_points(p, n)
            g.build_octree()
            g.compute_divergence()
        counts, _ = plain.level_info()
        wd = {16: 4, 64: 6}[grid]
        assert (counts[wd] > 2048) == (grid == 64), counts.tolist()
        sp, sw = plain.solve(), warm.solve()
        assert warm.warm_start_info() == (wd, True)
        assert [s["iters"] for s in sp] == [s["iters"] for s in sw]
        assert np.array_equal(plain.read_buffer(prg.BUF_SOLUTION).view(np.uint32),
                              warm.read_buffer(prg.BUF_SOLUTION).view(np.uint32))
        assert plain.warm_start_info() == (-1, False)
    finally:
        warm.close()
        plain.close()
        net.close()
