"""ORACLE (test infrastructure): numpy restatement of query_pairs semantics.

All (i, j) with i < j whose distance in lon/lat degree space is at most r
(closed ball), positions being the float32 record widened to double, as
cKDTree(points).query_pairs(r, p) returns them at
interaction_simulator.py:93,98.  Sorted lexicographically.  O(n^2) in blocks:
small cases only; tests/ pin it against scipy.spatial.cKDTree.
"""
import numpy as np


def pairs_numpy(x, y, r, p=2.0, block=2048):
    x = np.asarray(x, dtype=np.float32).astype(np.float64)
    y = np.asarray(y, dtype=np.float32).astype(np.float64)
    n = x.size
    out_i, out_j = [], []
    for a in range(0, n, block):
        xa, ya = x[a:a + block, None], y[a:a + block, None]
        dx = x[None, :] - xa
        dy = y[None, :] - ya
        if p == 2:
            hit = dx * dx + dy * dy <= r * r
        elif p == 1:
            hit = np.abs(dx) + np.abs(dy) <= r
        elif np.isinf(p):
            hit = np.maximum(np.abs(dx), np.abs(dy)) <= r
        else:
            hit = np.abs(dx) ** p + np.abs(dy) ** p <= r ** p
        ii, jj = np.nonzero(hit)
        ii = ii + a
        keep = jj > ii
        out_i.append(ii[keep])
        out_j.append(jj[keep])
    pi = np.concatenate(out_i) if out_i else np.zeros(0, np.int64)
    pj = np.concatenate(out_j) if out_j else np.zeros(0, np.int64)
    order = np.lexsort((pj, pi))
    return pi[order].astype(np.int32), pj[order].astype(np.int32)


def sorted_kdtree_pairs(x, y, r, p=2.0):
    """scipy cKDTree.query_pairs (the reference's call) as a sorted int32 array pair."""
    from scipy.spatial import cKDTree
    pts = np.stack([np.asarray(x, np.float32), np.asarray(y, np.float32)], axis=-1).astype(np.float64)
    arr = cKDTree(pts).query_pairs(r=r, p=p, output_type="ndarray")
    if arr.size == 0:
        return np.zeros(0, np.int32), np.zeros(0, np.int32)
    arr = np.sort(arr, axis=1)
    order = np.lexsort((arr[:, 1], arr[:, 0]))
    arr = arr[order]
    return arr[:, 0].astype(np.int32), arr[:, 1].astype(np.int32)
