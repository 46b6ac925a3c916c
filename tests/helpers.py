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
