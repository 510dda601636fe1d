"""alto_hip_cpd_error_noalloc against alto_hip_cpd_error (cpd_error,
src/streaming_cpd.cpp:368-464) through the C ABI: two calls in a row on the
shared device scratch, the second with a smaller input (fewer modes, rows,
nonzeros and a lower rank), each equal to the allocating entry point's
result within 1e-10 relative (FP64 sums in another order); a coordinate
beyond its mode gives NaN; a rank beyond the scratch limit is refused."""
import numpy as np
import pytest

from conftest import rand_factors, rand_tensor

pytestmark = pytest.mark.gpu


def _run_both(gpu, dims, nnz, R, seed):
    dims, cidx, vals = rand_tensor(dims, nnz, seed=seed, dup_frac=0.1)
    F = rand_factors(dims, R, seed=seed + 1, lo=-1, hi=1)
    err, parts = gpu.cpd_error(dims, cidx, vals, F)
    dc = [gpu.DeviceArray.from_numpy(c.astype(np.uint64)) for c in cidx]
    dv = gpu.DeviceArray.from_numpy(vals)
    dF = [gpu.DeviceArray.from_numpy(f) for f in F]
    out = gpu.DeviceArray.zeros(4, np.float64)
    gpu.cpd_error_noalloc(dims, dc, dv, dF, out)
    gpu.synchronize()
    got = out.numpy()
    want = np.array([err] + list(parts))
    print("dims", dims, "nnz", nnz, "R", R, "noalloc", got, "cpd_error", want, flush=True)
    np.testing.assert_allclose(got, want, rtol=1e-10, atol=0)
    return dims, dc, dv, dF


def test_noalloc_matches_cpd_error_twice_then_smaller(gpu):
    _run_both(gpu, [400, 300, 200], 50000, 32, seed=301)
    _run_both(gpu, [50, 40, 30, 20], 3000, 8, seed=303)


def test_noalloc_bad_coordinate_is_nan_and_rank_limit(gpu):
    dims, dc, dv, dF = _run_both(gpu, [60, 50, 40], 2000, 16, seed=305)
    bad = dc[1].numpy()
    bad[7] = 50  # not below its mode's size
    db = gpu.DeviceArray.from_numpy(bad)
    out = gpu.DeviceArray.zeros(4, np.float64)
    gpu.cpd_error_noalloc(dims, [dc[0], db, dc[2]], dv, dF, out)
    gpu.synchronize()
    assert np.all(np.isnan(out.numpy()))
    F65 = [gpu.DeviceArray.zeros((d, 65), np.float64) for d in dims]
    with pytest.raises(gpu.AltoError, match="not built|scratch"):
        gpu.cpd_error_noalloc(dims, dc, dv, F65, out)
