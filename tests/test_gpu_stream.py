"""Device StreamingSparseTensor (streaming_sptensor.cpp:69-264, tt_sort
sort.cpp:477-519, next_dynamic_batch at timeslice limit 1) against the
oracle's stream_relabel: relabelled coordinates and values bit-exact, slice
boundaries and running-max dims exact.  Covers duplicate coordinates (ties
keep input order), a stream mode that is not first, and a 5-mode stream
whose lexicographic key needs more than 128 bits (two sort passes)."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _expected(cidx, vals, s):
    c2, v2 = O.stream_relabel(cidx, vals, s)
    t = c2[s].astype(np.int64)
    T = int(t.max()) + 1
    counts = np.bincount(t, minlength=T)
    ptr = np.concatenate([[0], np.cumsum(counts)]).astype(np.uint64)
    dims = np.zeros((T, len(cidx)), dtype=np.uint64)
    for m in range(len(cidx)):
        if m == s:
            dims[:, m] = 1
            continue
        run = 0
        for k in range(T):
            seg = c2[m][int(ptr[k]): int(ptr[k + 1])]
            run = max(run, int(seg.max()) + 1)
            dims[k, m] = run
    return c2, v2, ptr, dims


def _check(gpu, dims, cidx, vals, s):
    got_c, got_v, got_ptr, got_dims = gpu.stream_prepare(dims, cidx, vals, s)
    c2, v2, ptr, edims = _expected(cidx, vals, s)
    for a, b in zip(got_c, c2):
        np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(got_v, v2)
    np.testing.assert_array_equal(got_ptr, ptr)
    np.testing.assert_array_equal(got_dims, edims)


@pytest.mark.parametrize("s", [0, 2])
def test_stream_prepare_with_duplicates(gpu, s):
    rng = np.random.default_rng(130)
    n, T = 20000, 7
    dims = [T, 3000, 500] if s == 0 else [3000, 500, T]
    cidx = [rng.integers(0, d, size=n).astype(np.uint64) for d in dims]
    cidx[s][:T] = np.arange(T)
    dup = rng.integers(0, n, size=4000)               # repeated coordinates
    for c in cidx:
        c[n - 4000:] = c[dup]
    vals = rng.random(n)
    _check(gpu, dims, cidx, vals, s)


def test_stream_prepare_key_wider_than_128_bits(gpu):
    rng = np.random.default_rng(131)
    n, T = 5000, 4
    dims = [T, 1 << 40, 1 << 40, 1 << 30, 1 << 20]   # 2 + 40 + 40 + 30 + 20 = 132 bits
    cidx = [rng.integers(0, d, size=n, dtype=np.uint64) for d in dims]
    cidx[0] = rng.integers(0, T, size=n).astype(np.uint64)
    cidx[0][:T] = np.arange(T)
    hot = rng.integers(0, 200, size=n)                # a few shared indices per mode
    for m in range(1, 5):
        cidx[m] = np.where(rng.random(n) < 0.5, cidx[m][hot], cidx[m])
    vals = rng.random(n)
    _check(gpu, dims, cidx, vals, 0)


def test_stream_prepare_errors(gpu):
    rng = np.random.default_rng(132)
    dims = [5, 30, 20]
    cidx = [rng.integers(0, d, size=300).astype(np.uint64) for d in dims]
    cidx[0] = np.where(cidx[0] == 3, 4, cidx[0]).astype(np.uint64)
    cidx[0][:3] = [0, 1, 2]
    cidx[0][3] = 4
    vals = rng.random(300)
    with pytest.raises(gpu.AltoError, match="empty time slice"):
        gpu.stream_prepare(dims, cidx, vals, 0)
    bad = [c.copy() for c in cidx]
    bad[1][7] = 30
    with pytest.raises(gpu.AltoError):
        gpu.stream_prepare(dims, bad, vals, 0)
