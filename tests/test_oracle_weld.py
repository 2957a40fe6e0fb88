"""The welding oracle (oracle/weld_ref.py) pinned by a hand-worked case, by
an independent construction (np.unique) and by invariants that a wrong
index or order would break."""
import numpy as np

from oracle import weld_ref


def _mesh_with_duplicates(seed, n=3000, dup=0.3, tris=5000):
    rng = np.random.default_rng(seed)
    base = rng.random((n, 3), dtype=np.float32)
    k = int(n * dup)
    src = rng.integers(0, n, k)
    dst = rng.choice(n, k, replace=False)
    base[dst] = base[src]
    base[5] = [0.0, 0.5, 0.5]
    base[6] = [-0.0, 0.5, 0.5]  # -0 and +0 differ in their bits: not merged
    t = rng.integers(0, n, (tris, 3)).astype(np.int32)
    return base, t


def test_hand_worked_case():
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 0, 0], [0, 1, 0], [1, 0, 0]], np.float32)
    t = np.array([[0, 1, 3], [2, 4, 3]], np.int32)
    wv, wt, nleaf, removed = weld_ref.weld(v, t, 2)
    assert removed == 2 and nleaf == 2
    assert np.array_equal(wv, np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32))
    assert np.array_equal(wt, np.array([[0, 1, 2], [0, 1, 2]], np.int32))
    # n_leaf counts the kept vertices among the first n_leaf
    assert weld_ref.weld(v, t, 3)[2] == 2 and weld_ref.weld(v, t, 4)[2] == 3 and weld_ref.weld(v, t, 0)[2] == 0


def test_two_constructions_agree():
    for seed in range(4):
        v, t = _mesh_with_duplicates(seed)
        a = weld_ref.weld(v, t, 1000)
        b = weld_ref.weld_unique(v, t, 1000)
        assert np.array_equal(a[0].view(np.uint32), b[0].view(np.uint32))
        assert np.array_equal(a[1], b[1]) and a[2] == b[2] and a[3] == b[3]


def test_invariants():
    v, t = _mesh_with_duplicates(9)
    wv, wt, _, removed = weld_ref.weld(v, t, len(v))
    # every triangle corner keeps its position, bit for bit
    assert np.array_equal(wv[wt].view(np.uint32), v[t].view(np.uint32))
    # positions are distinct after welding, and -0 / +0 stay apart
    assert len(np.unique(wv.view(np.uint32), axis=0)) == len(wv) == len(v) - removed
    assert removed > 0
    # the kept vertices are the first occurrences, in order (idempotent)
    again = weld_ref.weld(wv, wt, len(wv))
    assert again[3] == 0 and np.array_equal(again[1], wt)


def test_empty_and_single():
    wv, wt, nleaf, removed = weld_ref.weld(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.int32), 0)
    assert len(wv) == 0 and len(wt) == 0 and removed == 0
    wv, wt, nleaf, removed = weld_ref.weld(np.ones((1, 3), np.float32), np.zeros((2, 3), np.int32), 1)
    assert len(wv) == 1 and nleaf == 1 and removed == 0 and np.array_equal(wt, np.zeros((2, 3), np.int32))
