"""The reference's formula design expansion in numpy -- TEST INFRASTRUCTURE ONLY.

Restates src/src_main.f90 for checking the device expansion (fehip_design_levels /
fehip_design_expand): collect_levels_for_column (:1389-1432; MAX_CATEGORIES = 40, CATEGORY_TOL = 1e-6
at :15-16), append_main_term_columns (:1178-1230) and append_interaction_columns with
build_factor_block (:1276-1387).  Terms are (name, column, categorical, base) tuples.
"""
from __future__ import annotations

import numpy as np

MAX_CATEGORIES = 40
CATEGORY_TOL = 1e-6


def nint(x):
    """Fortran nint: nearest integer, halves away from zero."""
    return np.where(x >= 0, np.floor(x + 0.5), np.ceil(x - 0.5)).astype(np.int64)


def collect_levels(column):
    """First MAX_CATEGORIES distinct near-integer values in row order, sorted; and whether the list filled."""
    seen = []
    for x in np.asarray(column, dtype=np.float64):
        v = int(nint(np.array(x)))
        if abs(x - v) > CATEGORY_TOL:
            continue
        if v not in seen:
            seen.append(v)
            if len(seen) == MAX_CATEGORIES:
                return sorted(seen), True
    return sorted(seen), False


def _block(raw, term, levels):
    name, col, cat, base = term
    x = raw[:, col]
    if not cat:
        return x[:, None].copy(), [name]
    lv = levels[col]
    if len(lv) <= 1:
        return None
    b = lv.index(base) if base in lv else 0
    cols, names = [], []
    for i, v in enumerate(lv):
        if i == b:
            continue
        cols.append((nint(x) == v).astype(np.float64))
        names.append(f"{name}:{v}")
    return np.column_stack(cols), names


def expand(raw, terms, interactions):
    """(design matrix, names) for the main terms then the interactions."""
    raw = np.asarray(raw, dtype=np.float64)
    levels = {}
    for t in list(terms) + [f for it in interactions for f in it]:
        if t[2] and t[1] not in levels:
            levels[t[1]] = collect_levels(raw[:, t[1]])[0]
    blocks, names = [], []
    for t in terms:
        b = _block(raw, t, levels)
        if b is not None:
            blocks.append(b[0])
            names += b[1]
    for it in interactions:
        if len(it) <= 1:
            continue
        cur = _block(raw, it[0], levels)
        for f in it[1:]:
            nxt = _block(raw, f, levels)
            if cur is None or nxt is None:
                cur = None
                break
            m = np.column_stack([cur[0][:, a] * nxt[0][:, b] for a in range(cur[0].shape[1])
                                 for b in range(nxt[0].shape[1])])
            cur = (m, [f"{x}&{y}" for x in cur[1] for y in nxt[1]])
        if cur is not None:
            blocks.append(cur[0])
            names += cur[1]
    n = raw.shape[0]
    return (np.column_stack(blocks) if blocks else np.zeros((n, 0))), names, levels
