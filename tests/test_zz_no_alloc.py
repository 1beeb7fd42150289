"""GPU parity of the allocation-free SYRK (fe_gpu_syrk_noalloc) against fe_gpu_syrk.

Both are called through the C-ABI with the same inputs.  fe_gpu_syrk_noalloc keeps its split-K
partials in a fixed device array that every call on the GPU reuses, so each case makes two calls in
a row, the second with a smaller input, and checks both against fe_gpu_syrk: a stale partial of the
first call would show in the second.  FP64 products of the same kernels: the upper triangles agree
to 1e-12 relative (bit-equal whenever both pick the same split count, which these shapes do).
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def note(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def lib():
    from fe_gpu_amd import device
    from fe_gpu_amd._lib import load
    device.init(0)
    return load()


def _upper(Q):
    return np.triu(Q)


def _syrk_both(lib, W, beta, Q0):
    from fe_gpu_amd.device import DeviceBuffer
    n, K = W.shape
    dW = DeviceBuffer.from_array(np.asfortranarray(W))
    out = []
    for fn in (lib.fe_gpu_syrk, lib.fe_gpu_syrk_noalloc):
        dQ = DeviceBuffer.from_array(np.asfortranarray(Q0))
        assert fn(n, K, 1.25, dW.ptr, n, beta, dQ.ptr) == 0
        out.append(dQ.to_array(np.float64, (K, K), order="F"))
    return out


@pytest.mark.parametrize("shapes", [
    [(300_000, 40), (60_000, 20)],      # matrix-core path (>= 16 columns, >= 4096 rows), split-K
    [(400_000, 9), (5_000, 6)],         # tiled path with split-K partials
    [(20_000, 160), (4_096, 17)],       # wide tiles, then the smallest matrix-core shape
])
def test_noalloc_syrk_matches_syrk_twice(lib, shapes):
    rng = np.random.default_rng(len(shapes[0]) + shapes[0][1])
    for call, (n, K) in enumerate(shapes):
        note(f"syrk_noalloc call {call}: n={n} K={K}")
        W = rng.normal(size=(n, K))
        Q0 = rng.normal(size=(K, K))
        ref, got = _syrk_both(lib, W, 0.5 if call else 0.0, Q0)
        expect = 1.25 * (W.T @ W) + (0.5 * Q0 if call else 0.0)
        scale = np.max(np.abs(_upper(ref)))
        assert np.max(np.abs(_upper(got) - _upper(ref))) <= 1e-12 * scale
        assert np.max(np.abs(_upper(ref) - _upper(expect))) <= 1e-11 * scale
        # the lower triangle is not written by either
        low = np.tril_indices(K, -1)
        assert np.array_equal(got[low], Q0[low]) and np.array_equal(ref[low], Q0[low])


def test_noalloc_syrk_rejects_sizes_beyond_its_scratch(lib):
    from fe_gpu_amd._lib import last_error
    note("syrk_noalloc size check")
    # 2049 columns need 2049^2 * 8 bytes of partials: above the 32 MiB maximum; nothing is launched
    assert lib.fe_gpu_syrk_noalloc(10, 2049, 1.0, 0, 10, 0.0, 0) != 0
    assert "maximum" in last_error()
    assert lib.fe_gpu_runtime_clear_error() == 0
    assert lib.fe_gpu_syrk_noalloc(10, 0, 1.0, 0, 10, 0.0, 0) == 0
    assert lib.fe_gpu_syrk_noalloc(10, -1, 1.0, 0, 10, 0.0, 0) != 0
    assert lib.fe_gpu_runtime_clear_error() == 0
