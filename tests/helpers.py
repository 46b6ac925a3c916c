"""Shared comparison helpers for the parity tests."""
from __future__ import annotations

import numpy as np


def normwise(a, b) -> float:
    """max_i |a_i - b_i| / max_i |b_i| with |.| the Euclidean norm of a row."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    if a.ndim == 1:
        a, b = a[:, None], b[:, None]
    den = np.max(np.linalg.norm(b, axis=1))
    num = np.max(np.linalg.norm(a - b, axis=1)) if a.size else 0.0
    if den == 0.0:
        return float(num)
    return float(num / den)


def rel_l2(a, b) -> float:
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.linalg.norm(b)
    num = np.linalg.norm(a - b)
    return float(num / den) if den else float(num)


def by_id(field_cell_order: np.ndarray, sorted_ids: np.ndarray) -> np.ndarray:
    """Reorder a cell-order field to particle-id order."""
    out = np.empty_like(field_cell_order)
    out[sorted_ids] = field_cell_order
    return out


def flat_table_census(params, pos, order, start):
    """For each particle in cell order (pos[order], cell table `start` of the
    oracle), the number of non-empty 32-slot pseudo-rows of its nine stencil
    rows' exact neighbour masks (the flattened walk's table), whether all its
    rows hold at most 64 slots, and whether an empty pseudo-row follows the
    16th non-empty one.  The FP32 pair predicate is emulated (FMA products
    exact in FP64)."""
    P = np.ascontiguousarray(pos[order], np.float32)
    n = P.shape[0]
    gx, gy, gz = (int(g) for g in params.grid)
    inv = np.array([np.float32(params.inv_cell[a]) for a in range(3)], np.float32)
    h2 = np.float32(params.h) * np.float32(params.h)
    c = np.floor(P * inv).astype(np.int64)
    inside = np.all((c >= 0) & (c < np.array([gx, gy, gz])), axis=1)
    c = np.where(inside[:, None], c, 0)
    x0 = np.maximum(c[:, 0] - 1, 0)
    x1 = np.minimum(c[:, 0] + 1, gx - 1)
    ent = np.zeros(n, np.int64)
    fits = inside.copy()
    trailing = np.zeros(n, bool)
    idx = np.arange(n)
    st = start.astype(np.int64)
    for r in range(9):
        z = c[:, 2] + r // 3 - 1
        y = c[:, 1] + r % 3 - 1
        ok = (z >= 0) & (z < gz) & (y >= 0) & (y < gy) & inside
        row = (np.where(ok, z, 0) * gy + np.where(ok, y, 0)) * gx
        b = st[row + x0]
        e = np.where(ok, st[row + x1 + 1], b)
        length = e - b
        fits &= length <= 64
        words = np.zeros((n, 2), bool)
        for k in range(64):
            j = b + k
            live = (k < length) & (j != idx)
            jj = np.where(live, j, 0)
            d = (P[jj] - P).astype(np.float32)
            sq = d.astype(np.float64) ** 2
            rr = np.float32(sq[:, 2] + np.float64(np.float32(sq[:, 1] + np.float64(np.float32(sq[:, 0])))))
            words[:, k // 32] |= live & (rr <= h2)
        for w in range(2):
            empty = ~words[:, w]
            trailing |= empty & (ent == 16)
            ent += words[:, w]
    return ent, fits, trailing


def flat_table_hazards(params, pos, order, start, wave=64):
    """Lanes whose flattened-walk table has exactly 16 entries followed by an
    empty pseudo-row, in waves that take the flattened walk (every lane's
    rows fit 64-slot masks and no lane has more than 16 entries)."""
    ent, fits, trailing = flat_table_census(params, pos, order, start)
    lanes = 0
    for w0 in range(0, ent.size, wave):
        sl = slice(w0, w0 + wave)
        if fits[sl].all() and not (ent[sl] > 16).any():
            lanes += int((trailing[sl] & (ent[sl] == 16)).sum())
    return lanes
