"""Pins of the oracle's design expansion (src/src_main.f90 restated in oracle/design.py) by hand-worked
examples, and the product's column planner (fe_gpu_amd.design.plan, host code) against it."""
import numpy as np

from oracle import design as od


def test_levels_hand_example():
    col = [3, 1.0000001, 2.4, 2, 3, -1, 2.0000019]
    # 2.4 and 2.0000019 are not within 1e-6 of an integer; first seen 3, 1, 2, -1
    assert od.collect_levels(col) == ([-1, 1, 2, 3], False)


def test_levels_truncate_in_row_order():
    order = [50 - i for i in range(50)]          # 50, 49, ..., 1
    lv, tr = od.collect_levels(order)
    assert tr and lv == list(range(11, 51))      # the first 40 seen, sorted
    lv, tr = od.collect_levels(list(range(40)) * 2)
    assert tr and lv == list(range(40))          # the list filling up counts as truncation


def test_expand_hand_example():
    x = np.array([1.5, -2.0, 0.5, 3.0, 2.0])
    c = np.array([1, 2, 3, 2.4, 3])               # 2.4 rounds to level 2 in the dummies
    d = np.array([5, 7, 5, 7, 7])
    raw = np.column_stack([x, c, d])
    terms = [("x", 0, False, None), ("c", 1, True, 2)]
    inter = [[("c", 1, True, None), ("x", 0, False, None)], [("c", 1, True, 2), ("d", 2, True, None)]]
    M, names, levels = od.expand(raw, terms, inter)
    assert levels == {1: [1, 2, 3], 2: [5, 7]}
    assert names == ["x", "c:1", "c:3", "c:2&x", "c:3&x", "c:1&d:7", "c:3&d:7"]
    c1, c2, c3 = (c == 1) * 1.0, (np.round(c) == 2) * 1.0, (c == 3) * 1.0
    d7 = (d == 7) * 1.0
    expect = np.column_stack([x, c1, c3, c2 * x, c3 * x, c1 * d7, c3 * d7])
    assert np.array_equal(M, expect)


def test_planner_matches_oracle_order_and_names():
    from fe_gpu_amd import design
    rng = np.random.default_rng(4)
    raw = np.column_stack([rng.normal(size=300), rng.integers(1, 6, 300), rng.integers(0, 3, 300),
                           rng.normal(size=300)])
    terms_o = [("a", 0, False, None), ("g", 1, True, 3), ("h", 2, True, None)]
    inter_o = [[("g", 1, True, None), ("h", 2, True, 2)], [("a", 0, False, None), ("b", 3, False, None)],
               [("g", 1, True, None), ("a", 0, False, None), ("h", 2, True, None)]]
    M, names, levels = od.expand(raw, terms_o, inter_o)
    F = lambda t: design.Factor(t[0], t[1], t[2], t[3])  # noqa: E731
    cols, pnames = design.plan([F(t) for t in terms_o], [[F(t) for t in it] for it in inter_o], levels)
    assert pnames == names and len(cols) == M.shape[1]
