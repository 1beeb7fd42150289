"""Formula design expansion on the device (fehip_design_levels / fehip_design_expand).

The column list follows build_formula_design_matrix (src/src_main.f90:595-736): main terms in formula
order (a continuous term is its column; a categorical term one dummy per level but the base, levels
ascending, append_main_term_columns :1178-1230), then interactions (each factor's block -- its column,
or its dummies -- multiplied out with the first factor outermost, append_interaction_columns
:1276-1335).  A categorical variable with at most one level gives no columns; the base is the level
named by the term, else the smallest (find_level_index / count_columns_for_term :808-826).  Names are
'x', 'x:3' and 'a:1&b:2' as the reference writes them.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

from . import _coarse_sigs as S
from ._lib import check, load
from .device import DeviceBuffer


@dataclass
class Factor:
    name: str
    col: int                    # raw column index
    categorical: bool = False
    base: int | None = None     # base level of a categorical factor (None: the smallest level)


def column_levels(col_ptr: int, n: int):
    """The levels of a raw device column (first 40 distinct near-integer values in row order, sorted)
    and whether the list filled up."""
    lib = load()
    lv = (S.I32 * 40)()
    nl, tr = S.I32(), S.I32()
    check(lib.fehip_design_levels(col_ptr, n, lv, C.byref(nl), C.byref(tr)), "fehip_design_levels")
    return [int(v) for v in lv[:nl.value]], bool(tr.value)


def _block(f: Factor, levels):
    """(factors of each column, names) of one factor's block; None when it gives no columns."""
    if not f.categorical:
        return [[(f.col, 0, 0)]], [f.name]
    lv = levels.get(f.col, [])
    if len(lv) <= 1:
        return None
    base = lv.index(f.base) if f.base in lv else 0
    keep = [v for i, v in enumerate(lv) if i != base]
    return [[(f.col, v, 1)] for v in keep], [f"{f.name}:{v}" for v in keep]


def plan(terms, interactions, levels):
    """Column list and names for main terms and interactions; levels maps raw column -> its levels."""
    cols, names = [], []
    for t in terms:
        b = _block(t, levels)
        if b:
            cols += b[0]
            names += b[1]
    for inter in interactions:
        if len(inter) <= 1:
            continue
        cur = _block(inter[0], levels)
        for f in inter[1:]:
            nxt = _block(f, levels)
            if cur is None or nxt is None:
                cur = None
                break
            cur = ([a + b for a in cur[0] for b in nxt[0]], [f"{x}&{y}" for x in cur[1] for y in nxt[1]])
        if cur:
            cols += cur[0]
            names += cur[1]
    return cols, names


def expand(raw_ptr: int, ld_raw: int, n_raw: int, n: int, cols) -> DeviceBuffer:
    """The expanded design (n x len(cols), column-major FP64) in a new device buffer."""
    if any(len(c) > 4 for c in cols):
        raise ValueError("at most 4 factors per column")
    arr = (S.DesignColumn * max(1, len(cols)))()
    for j, c in enumerate(cols):
        arr[j].n_factors = len(c)
        for k, (col, level, cat) in enumerate(c):
            arr[j].col[k], arr[j].level[k], arr[j].categorical[k] = col, level, cat
    out = DeviceBuffer(8 * n * max(1, len(cols)))
    check(load().fehip_design_expand(raw_ptr, ld_raw, n_raw, n, arr, len(cols), out.ptr, n), "fehip_design_expand")
    return out
