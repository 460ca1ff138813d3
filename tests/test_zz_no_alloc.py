"""mgm_spmv_rows_static, the entry point that allocates no device memory,
against mgm_spmv_rows (the same SpMV over caller rows, staged in buffers the
library allocates per call) and the oracle's SpMV (hf/mat_lib.h:70-89).

Two calls in a row, the second smaller, so that stale contents of the shared
scratch array from the first call would show; sizes beyond the array are
rejected before any device work."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _call(fn, D, cl, f):
    n, m = D.shape
    out = np.full(n, np.nan)
    rc = fn(n, m, D.ctypes.data, cl.ctypes.data, f.size, f.ctypes.data, out.ctypes.data)
    return rc, out


def test_static_staging_matches_allocating_entry_point(gpu, oracle):
    from multigrid_meshless_amd._lib import lib
    L = lib()
    rng = np.random.default_rng(3)
    for n, m, ncol in ((60000, 70, 150000), (3000, 20, 5000)):
        print(f"[gpu no-alloc] spmv rows {n} x {m}, vector {ncol}", flush=True)
        D = rng.standard_normal((n, m))
        cl = rng.integers(0, ncol, (n, m), dtype=np.int32)
        f = rng.standard_normal(ncol)
        rc, want = _call(L.mgm_spmv_rows, D, cl, f)
        assert rc == 0, L.mgm_last_error()
        rc, got = _call(L.mgm_spmv_rows_static, D, cl, f)
        assert rc == 0, L.mgm_last_error()
        np.testing.assert_array_equal(got, want)
        np.testing.assert_array_equal(got, oracle.spmv(D, cl, f))


def test_static_staging_rejects_what_does_not_fit(gpu):
    from multigrid_meshless_amd._lib import lib
    L = lib()
    n, m, ncol = 100000, 70, 1000            # 84 MB of rows and columns: beyond 64 MiB
    D = np.zeros((n, m))
    cl = np.zeros((n, m), np.int32)
    f = np.zeros(ncol)
    rc, out = _call(L.mgm_spmv_rows_static, D, cl, f)
    assert rc != 0 and b"scratch" in L.mgm_last_error()
    cl = np.full((10, 4), 7, np.int32)       # a column past the vector
    rc, out = _call(L.mgm_spmv_rows_static, np.ones((10, 4)), cl, np.ones(5))
    assert rc != 0 and b"column index" in L.mgm_last_error()
