"""oxi_dlaswp_noalloc (no reference counterpart: the laswp for callers that
cannot allocate) through the C ABI against oxi_dlaswp, to the bar of the
laswp tests: bit-identical rows.  Two calls in a row, the second with a
smaller input, then an invalid pivot rejected before any write."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run(ox, fn, A, k1, k2, piv, inverse):
    import torch
    h = ox.handle()
    Ad = ox.DMatrix.from_numpy(A)
    pv = torch.from_numpy(np.ascontiguousarray(piv, dtype=np.int64)).to("cuda")
    rc = getattr(h._L, fn)(h.ptr, Ad.n, Ad.ptr, Ad.ld, int(k1), int(k2), pv.data_ptr(), 1 if inverse else 0)
    return rc, Ad.numpy()


def test_noalloc_laswp_matches_laswp(gpu):
    rng = np.random.default_rng(0x0B1A5)
    # first call: more than one block of 512 interchanges, both directions
    m, ncols = 1500, 33
    A = rng.standard_normal((m, ncols))
    piv = np.array([rng.integers(j, m) + 1 for j in range(m)], dtype=np.int64)
    for inverse in (False, True):
        rc1, got = _run(gpu, "oxi_dlaswp_noalloc", A, 1, m, piv, inverse)
        rc2, ref = _run(gpu, "oxi_dlaswp", A, 1, m, piv, inverse)
        assert rc1 == 0 and rc2 == 0
        assert np.array_equal(got, ref)
        assert not np.array_equal(got, A)
    # second call, smaller: arbitrary pivots (the general replay), a sub-range
    m2, n2 = 300, 5
    A2 = rng.standard_normal((m2, n2))
    piv2 = rng.integers(1, m2 + 1, m2).astype(np.int64)
    rc1, got = _run(gpu, "oxi_dlaswp_noalloc", A2, 20, 250, piv2, False)
    rc2, ref = _run(gpu, "oxi_dlaswp", A2, 20, 250, piv2, False)
    assert rc1 == 0 and rc2 == 0 and np.array_equal(got, ref)
    # an out-of-range pivot: rejected (-7), nothing written
    bad = piv2.copy(); bad[30] = m2 + 10
    rc, after = _run(gpu, "oxi_dlaswp_noalloc", A2, 1, 100, bad, False)
    assert rc == -7 and np.array_equal(after, A2)
    # and a valid call afterwards still works (the status word is cleared)
    rc1, got = _run(gpu, "oxi_dlaswp_noalloc", A2, 1, m2, piv2, True)
    rc2, ref = _run(gpu, "oxi_dlaswp", A2, 1, m2, piv2, True)
    assert rc1 == 0 and rc2 == 0 and np.array_equal(got, ref)
