"""sjcs_index_count_noalloc (the index count that allocates nothing: its
working memory is the library's fixed per-device scratch array) against
sjcs_index_count, the entry point it is a variant of, on two calls in a row
-- the second smaller, with other shells -- plus the size limit.

Runs on an MI355X only (marker `gpu`), through the C-ABI."""
import numpy as np
import pytest

import sjcs_testlib as t

pytestmark = pytest.mark.gpu


def test_noalloc_equals_index_count(built):
    rng = np.random.default_rng(77)
    p = rng.random((300_000, 3), dtype=np.float32) * np.float32(1000.0)
    eng = built.Engine(0)
    ix = built.ArraySTR(eng, p.shape[0])
    try:
        ix.insert(p)
        ix.make_tree()
        calls = [(20_000, (0.001, 5.0, 40)), (3_000, (0.5, 30.0, 13)), (1, (0.01, 2.0, 8))]
        for n, (mn, mx, k) in calls:
            h = rng.random((n, 3), dtype=np.float32) * np.float32(1000.0)
            h[: min(n, 2)] = [[0.0, 999.5, 3.0], [1000.0, 0.5, 500.0]][: min(n, 2)]
            _, r2 = t.oracle_thresholds(mn, mx, k)
            want = ix.multi_range_count(h, r2, np.full((n, k), 7, np.uint64))
            d_h = eng.device_array(h)
            d_c = eng.device_array(np.full((n, k), 7, np.uint64))
            try:
                ix.multi_range_count_noalloc(d_h, n, r2, d_c)
                got = np.zeros((n, k), np.uint64)
                eng._check(eng._lib.sjcs_copy_to_host(eng._h, got.ctypes.data, d_c, got.nbytes))
            finally:
                eng.device_free(d_h)
                eng.device_free(d_c)
            print(f"noalloc: {n} halos x {k} shells, {int(got.sum()) - 7 * n * k} pairs",
                  flush=True)
            assert np.array_equal(got, want)
        # beyond the scratch: 2e6 halos x 8 shells need 104 MB
        _, r2 = t.oracle_thresholds(0.001, 5.0, 8)
        d = eng.device_array(np.zeros((4, 3), np.float32))
        try:
            with pytest.raises(built.SJCSError, match="scratch"):
                ix.multi_range_count_noalloc(d, 2_000_000, r2, d)
        finally:
            eng.device_free(d)
    finally:
        ix.close()
        eng.close()
