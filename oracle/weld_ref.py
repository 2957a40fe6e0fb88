"""ORACLE (test infrastructure only): numpy restatement of the optional
cross-batch vertex welding that ``poissonrecon_gpu_amd/csrc/weld.hip`` runs
on the GPU (``prg_weld_mesh``).

Only tests/ use this module; the product never imports it.  The welding has
no counterpart in the reference (it concatenates its mesh batches without
merging the vertices they share, main.cu:3793,4183,4543); this file is its
specification:

  - two vertices are the same when their three float32 coordinates have the
    same bit patterns;
  - each position keeps its first occurrence, and the kept vertices stay in
    their original order;
  - every triangle entry becomes the output number of its vertex's first
    occurrence; triangles are neither removed nor reordered;
  - n_leaf_vertices becomes the number of kept vertices among the first
    n_leaf_vertices.
"""
from __future__ import annotations

import numpy as np


def weld(vertices: np.ndarray, triangles: np.ndarray, n_leaf: int):
    """(vertices (V', 3), triangles (T, 3), n_leaf', removed) by a dictionary
    pass in vertex order."""
    v = np.ascontiguousarray(vertices, np.float32).reshape(-1, 3)
    bits = v.view(np.uint32)
    first: dict[tuple[int, int, int], int] = {}
    remap = np.empty(len(v), np.int64)
    kept = []
    for j in range(len(v)):
        k = (int(bits[j, 0]), int(bits[j, 1]), int(bits[j, 2]))
        if k not in first:
            first[k] = len(kept)
            kept.append(j)
        remap[j] = first[k]
    kept_idx = np.array(kept, np.int64)
    out_v = v[kept_idx] if len(kept) else np.zeros((0, 3), np.float32)
    t = np.asarray(triangles, np.int64).reshape(-1, 3)
    out_t = remap[t].astype(np.int32) if len(t) else np.zeros((0, 3), np.int32)
    n_leaf_out = int(np.count_nonzero(kept_idx < min(n_leaf, len(v)))) if n_leaf > 0 else 0
    return out_v, out_t, n_leaf_out, len(v) - len(kept)


def weld_unique(vertices: np.ndarray, triangles: np.ndarray, n_leaf: int):
    """The same result through np.unique on the coordinate bit patterns (an
    independent construction of the first-occurrence numbering)."""
    v = np.ascontiguousarray(vertices, np.float32).reshape(-1, 3)
    if len(v) == 0:
        return v.copy(), np.asarray(triangles, np.int32).reshape(-1, 3), 0, 0
    bits = v.view(np.uint32)
    _, first_idx, inverse = np.unique(bits, axis=0, return_index=True, return_inverse=True)
    inverse = inverse.reshape(-1)
    order = np.argsort(first_idx, kind="stable")  # groups in order of first occurrence
    rank = np.empty_like(order)
    rank[order] = np.arange(len(order))
    remap = rank[inverse]
    kept_idx = np.sort(first_idx)
    t = np.asarray(triangles, np.int64).reshape(-1, 3)
    n_leaf_out = int(np.count_nonzero(kept_idx < min(n_leaf, len(v)))) if n_leaf > 0 else 0
    return v[kept_idx], remap[t].astype(np.int32), n_leaf_out, len(v) - len(kept_idx)
