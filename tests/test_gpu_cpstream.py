"""GPU CP-stream (cpstream + cpstream_alto_iter + admm_cf_, src/streaming_cpd.cpp,
src/admm.cpp) against the oracle on the same input and the same initial rows.

Bars (FP64 MTTKRP path): slice count, slice dims, outer and ADMM iteration
counts equal; time rows, final factors and Grams within 1e-8 relative.
FP32 MTTKRP path: factors within 1e-4 relative (the ADMM, solves and Grams
stay FP64)."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _stream_tensor(T, dims, nnz, seed, zipf=False):
    rng = np.random.default_rng(seed)
    t = np.sort(rng.integers(0, T, size=nnz)).astype(np.uint64)
    t[:T] = np.arange(T)
    rng.shuffle(t)
    cidx = [t]
    for d in dims:
        if zipf:
            w = np.arange(1, d + 1, dtype=np.float64) ** -1.05
            p = w / w.sum()
            cidx.append(rng.permutation(d)[rng.choice(d, size=nnz, p=p)].astype(np.uint64))
        else:
            cidx.append(rng.integers(0, d, size=nnz).astype(np.uint64))
    vals = rng.poisson(3, size=nnz) + 1.0
    return [T] + list(dims), cidx, vals


def _init(dims, R, seed):
    rng = np.random.default_rng(seed)
    return [rng.random((int(d), R)) for d in dims]


@pytest.fixture(params=["cached", "streamed"])
def admm_state(request, monkeypatch):
    """The non-negative ADMM loop keeps its row state in registers when each
    workgroup owns few row tiles (every slice here); ALTO_ADMM_CACHE=0 makes it
    stream the state from memory every iteration, the path of large slices."""
    if request.param == "streamed":
        monkeypatch.setenv("ALTO_ADMM_CACHE", "0")
    return request.param


@pytest.mark.parametrize("nonneg", [True, False])
@pytest.mark.parametrize("R", [4, 16, 32])
def test_cpstream_matches_oracle(gpu, nonneg, R, admm_state):
    dims, cidx, vals = _stream_tensor(6, [300, 200], 6000, seed=90)
    init = _init(dims, R, 91)
    O.set_threads(1)
    ref = O.cpstream(dims, cidx, vals, 0, R, 6, 1e-5, nonneg, init)
    got = gpu.cpstream(dims, cidx, vals, 0, R, 6, 1e-5, nonneg, init)
    assert got["slices"] == ref["slices"] == 6
    np.testing.assert_array_equal(got["dims"], ref["dims"])
    np.testing.assert_array_equal(got["outer_iters"], ref["outer_iters"])
    if nonneg:
        np.testing.assert_array_equal(got["admm_iters"], ref["admm_iters"])
    np.testing.assert_allclose(got["time_rows"], ref["time_rows"], rtol=1e-8, atol=1e-10)
    for a, b in zip(got["factors"], ref["factors"]):
        np.testing.assert_allclose(a, b, rtol=1e-8, atol=1e-10)
    # the reference writes only the upper triangle of the factor Grams
    iu = np.triu_indices(R)
    np.testing.assert_allclose(got["grams"][:, iu[0], iu[1]], ref["grams"][:, iu[0], iu[1]],
                               rtol=1e-8, atol=1e-9)


def test_cpstream_stream_mode_not_first_and_4_modes(gpu, admm_state):
    rng = np.random.default_rng(92)
    T, nnz = 5, 5000
    dims = [120, T, 90, 40]
    t = rng.integers(0, T, size=nnz).astype(np.uint64)
    t[:T] = np.arange(T)
    cidx = [rng.integers(0, 120, size=nnz).astype(np.uint64), t,
            rng.integers(0, 90, size=nnz).astype(np.uint64), rng.integers(0, 40, size=nnz).astype(np.uint64)]
    vals = rng.random(nnz) + 0.5
    init = _init(dims, 8, 93)
    ref = O.cpstream(dims, cidx, vals, 1, 8, 4, 1e-6, True, init)
    got = gpu.cpstream(dims, cidx, vals, 1, 8, 4, 1e-6, True, init)
    np.testing.assert_array_equal(got["outer_iters"], ref["outer_iters"])
    for a, b in zip(got["factors"], ref["factors"]):
        np.testing.assert_allclose(a, b, rtol=1e-8, atol=1e-10)


def test_cpstream_f32_mttkrp_zipf(gpu):
    dims, cidx, vals = _stream_tensor(5, [2000, 1000], 40000, seed=94, zipf=True)
    init = _init(dims, 16, 95)
    ref = O.cpstream(dims, cidx, vals, 0, 16, 3, 1e-30, True, init)
    got = gpu.cpstream(dims, cidx, vals, 0, 16, 3, 1e-30, True, init, prec=gpu.PREC_F32)
    assert np.all(got["factors"][1] >= 0) and np.all(got["factors"][2] >= 0)
    for a, b in zip(got["factors"][1:], ref["factors"][1:]):
        scale = np.abs(b).max()
        assert np.abs(a - b).max() <= 1e-4 * scale


def test_cpstream_empty_slice_is_an_error(gpu):
    dims, cidx, vals = _stream_tensor(4, [30, 20], 500, seed=96)
    cidx[0] = np.where(cidx[0] == 2, 3, cidx[0]).astype(np.uint64)  # slice 2 empty
    with pytest.raises(gpu.AltoError, match="empty time slice"):
        gpu.cpstream(dims, cidx, vals, 0, 4, 2, 1e-5, True, _init(dims, 4, 97))


@pytest.mark.parametrize("thr,lim", [(0, 2), (2500, 10), (0, 7)])
@pytest.mark.parametrize("nonneg", [True, False])
def test_cpstream_dynamic_batches_match_oracle(gpu, thr, lim, nonneg):
    """Batches of several time slices (next_dynamic_batch with -g lim and
    -k thr): the same batches, dims and outer / ADMM iteration counts as the
    oracle, time rows (one per time index) and factors within 1e-8."""
    dims, cidx, vals = _stream_tensor(7, [300, 200], 7000, seed=98)
    R = 8
    init = _init(dims, R, 99)
    print("thr", thr, "lim", lim, "nonneg", nonneg, flush=True)
    ref = O.cpstream(dims, cidx, vals, 0, R, 5, 1e-6, nonneg, init, nnz_threshold=thr, timeslice_limit=lim)
    got = gpu.cpstream(dims, cidx, vals, 0, R, 5, 1e-6, nonneg, init, nnz_threshold=thr, timeslice_limit=lim)
    assert got["slices"] == ref["slices"] and got["time_slices"] == ref["time_slices"] == 7
    assert ref["slices"] < 7
    np.testing.assert_array_equal(got["dims"][1:], ref["dims"][1:])
    np.testing.assert_array_equal(got["outer_iters"], ref["outer_iters"])
    if nonneg:
        np.testing.assert_array_equal(got["admm_iters"], ref["admm_iters"])
    np.testing.assert_allclose(got["time_rows"], ref["time_rows"], rtol=1e-8, atol=1e-10)
    for a, b in zip(got["factors"][1:], ref["factors"][1:]):
        np.testing.assert_allclose(a, b, rtol=1e-8, atol=1e-10)


def test_exact_closing_error_with_merged_duplicates(gpu):
    """The bench's exact closing error of a streamed model (duplicate
    coordinates merged into one value, which the reference's sum of squared
    stored values does not do) equals ||X - M||_F / ||X||_F computed densely
    in numpy from the relabelled stream, the time rows and the factors."""
    import os
    import sys
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    keep = {k: os.environ.get(k) for k in ("OMP_PROC_BIND", "OMP_PLACES", "OMP_WAIT_POLICY")}
    import bench  # sets OpenMP defaults for its CPU legs: undone here
    for k, v in keep.items():
        if v is None:
            os.environ.pop(k, None)
        else:
            os.environ[k] = v
    dims, cidx, vals = _stream_tensor(6, [40, 30], 3000, seed=95, zipf=True)
    init = _init(dims, 4, 96)
    res = gpu.cpstream(dims, cidx, vals, 0, 4, 5, 1e-5, True, init)
    got = bench.stream_exact_errors(gpu, dims, cidx, vals, {"cpstream_nonneg": res})
    sc, sv, _, _ = gpu.stream_prepare(dims, cidx, vals, 0)
    T, I, J = len(res["time_rows"]), int(res["dims"][1]), int(res["dims"][2])
    X = np.zeros((T, I, J))
    np.add.at(X, (sc[0].astype(np.int64), sc[1].astype(np.int64), sc[2].astype(np.int64)), sv)
    assert np.count_nonzero(X) < len(sv)  # the stream does repeat coordinates
    M = np.einsum("tr,ir,jr->tij", res["time_rows"], res["factors"][1], res["factors"][2])
    want = np.linalg.norm(X - M) / np.linalg.norm(X)
    assert abs(got["cpstream_nonneg"] - want) <= 1e-9 * want
