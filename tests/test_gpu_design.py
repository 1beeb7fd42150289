"""GPU parity of the design expansion: fehip_design_levels equals the reference's level collection
(oracle/design.py) on columns whose first occurrences spread over many blocks, truncation included;
fehip_design_expand equals the oracle's expanded matrix bit for bit."""
import sys

import numpy as np
import pytest

from oracle import design as od

pytestmark = pytest.mark.gpu


def note(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def device():
    from fe_gpu_amd import device as d
    d.init(0)


def test_levels_match_reference_collection():
    from fe_gpu_amd import design
    from fe_gpu_amd.device import DeviceBuffer
    note("design levels")
    rng = np.random.default_rng(6)
    n = 200_000
    cols = [rng.integers(1, 9, n).astype(np.float64),                     # 8 levels
            rng.integers(-30, 30, n).astype(np.float64),                  # 60 distinct: truncation
            np.where(rng.random(n) < 0.5, rng.integers(0, 5, n) + 0.3, rng.integers(0, 5, n)).astype(np.float64)]
    late = np.zeros(n)
    late[150_000:] = np.arange(n - 150_000) % 45 + 100                     # first occurrences late in the column
    cols.append(late)
    for c in cols:
        b = DeviceBuffer.from_array(c)
        assert design.column_levels(b.ptr, n) == od.collect_levels(c)
        b.free()


def test_expand_matches_reference_expansion():
    from fe_gpu_amd import design
    from fe_gpu_amd.device import DeviceBuffer
    note("design expand")
    rng = np.random.default_rng(9)
    n = 50_001
    raw = np.column_stack([rng.normal(size=n), rng.integers(1, 9, n), rng.integers(1, 31, n) + 0.2 * (rng.random(n) < 0.1),
                           rng.integers(1, 8, n), rng.normal(size=n)])
    terms = [("x1", 0, False, None), ("cat8", 1, True, None), ("x2", 4, False, None)]
    inter = [[("ind", 2, True, None), ("msp", 3, True, 4)], [("x1", 0, False, None), ("x2", 4, False, None)],
             [("cat8", 1, True, 2), ("x1", 0, False, None)]]
    M, names, levels = od.expand(raw, terms, inter)
    F = lambda t: design.Factor(t[0], t[1], t[2], t[3])  # noqa: E731
    draw = DeviceBuffer.from_array(np.asfortranarray(raw))
    dev_levels = {c: design.column_levels(draw.ptr + 8 * n * c, n)[0] for c in levels}
    assert dev_levels == levels
    cols, pnames = design.plan([F(t) for t in terms], [[F(t) for t in it] for it in inter], dev_levels)
    assert pnames == names
    out = design.expand(draw.ptr, n, raw.shape[1], n, cols)
    got = out.to_array(np.float64, (n, len(cols)), order="F")
    assert np.array_equal(got, M)
    out.free(), draw.free()
