"""GPU spCP-stream (cpstream with use_spcpstream + spcpstream_alto_iter,
src/streaming_cpd.cpp:26-366, 466-1085) against the oracle on the same input
and initial rows, and against the GPU CP-stream without constraints, which it
reformulates.

Bars (FP64 MTTKRP path): slice count, slice dims and outer-iteration counts
equal; time rows, final factors and the Grams' upper triangles within 1e-8
relative.  FP32 MTTKRP path: factors within 1e-3 of the largest entry (the
solves and R x R bookkeeping stay FP64; the unconstrained solves amplify the
MTTKRP's FP32 rounding)."""
import numpy as np
import pytest

import oracle as O
from test_gpu_cpstream import _init, _stream_tensor

pytestmark = pytest.mark.gpu


def _compare(got, ref, R, rtol=1e-8):
    assert got["slices"] == ref["slices"]
    np.testing.assert_array_equal(got["dims"], ref["dims"])
    np.testing.assert_array_equal(got["outer_iters"], ref["outer_iters"])
    np.testing.assert_allclose(got["time_rows"], ref["time_rows"], rtol=rtol, atol=1e-10)
    for a, b in zip(got["factors"], ref["factors"]):
        np.testing.assert_allclose(a, b, rtol=rtol, atol=1e-10)
    iu = np.triu_indices(R)
    np.testing.assert_allclose(got["grams"][:, iu[0], iu[1]], ref["grams"][:, iu[0], iu[1]],
                               rtol=rtol, atol=1e-9)


@pytest.mark.parametrize("R", [4, 16, 32])
@pytest.mark.parametrize("eps", [1e-5, 0.0])
def test_spcpstream_matches_oracle(gpu, R, eps):
    dims, cidx, vals = _stream_tensor(6, [300, 200], 6000, seed=110)
    init = _init(dims, R, 111)
    print("R", R, "eps", eps, flush=True)
    ref = O.spcpstream(dims, cidx, vals, 0, R, 6, eps, init)
    got = gpu.spcpstream(dims, cidx, vals, 0, R, 6, eps, init)
    assert got["slices"] == 6
    _compare(got, ref, R)


def test_spcpstream_equals_gpu_cpstream_unconstrained(gpu):
    """The reformulation on the device: every outer iteration runs
    (epsilon 0), so both paths give the same model to rounding."""
    dims, cidx, vals = _stream_tensor(5, [400, 150], 3000, seed=112)
    init = _init(dims, 8, 113)
    sp = gpu.spcpstream(dims, cidx, vals, 0, 8, 4, 0.0, init)
    cs = gpu.cpstream(dims, cidx, vals, 0, 8, 4, 0.0, False, init)
    np.testing.assert_allclose(sp["time_rows"], cs["time_rows"], rtol=1e-8, atol=1e-10)
    for a, b in zip(sp["factors"], cs["factors"]):
        np.testing.assert_allclose(a, b, rtol=1e-8, atol=1e-10)


def test_spcpstream_stream_mode_not_first_and_4_modes(gpu):
    rng = np.random.default_rng(114)
    T, nnz = 5, 4000
    dims = [120, T, 90, 40]
    t = rng.integers(0, T, size=nnz).astype(np.uint64)
    t[:T] = np.arange(T)
    cidx = [rng.integers(0, 120, size=nnz).astype(np.uint64), t,
            rng.integers(0, 90, size=nnz).astype(np.uint64), rng.integers(0, 40, size=nnz).astype(np.uint64)]
    vals = rng.random(nnz) + 0.5
    init = _init(dims, 8, 115)
    ref = O.spcpstream(dims, cidx, vals, 1, 8, 4, 1e-6, init)
    got = gpu.spcpstream(dims, cidx, vals, 1, 8, 4, 1e-6, init)
    _compare(got, ref, 8)


def test_spcpstream_rank_not_power_of_two(gpu):
    dims, cidx, vals = _stream_tensor(4, [250, 120], 2500, seed=116)
    init = _init(dims, 10, 117)
    ref = O.spcpstream(dims, cidx, vals, 0, 10, 3, 0.0, init)
    got = gpu.spcpstream(dims, cidx, vals, 0, 10, 3, 0.0, init)
    _compare(got, ref, 10)


def test_spcpstream_f32_mttkrp_zipf(gpu):
    """FP32 MTTKRP: within 1e-3 of the largest entry.  The unconstrained
    solves (Phi + 1e-12 I) amplify the FP32 rounding of the MTTKRP by Phi's
    conditioning (CP-stream's non-negative ADMM damps it); the FP64 path on
    the same input stays within 1e-8, so the difference is the FP32 rounding."""
    dims, cidx, vals = _stream_tensor(5, [2000, 1000], 40000, seed=118, zipf=True)
    init = _init(dims, 16, 119)
    ref = O.spcpstream(dims, cidx, vals, 0, 16, 3, 0.0, init)
    got = gpu.spcpstream(dims, cidx, vals, 0, 16, 3, 0.0, init, prec=gpu.PREC_F32)
    for a, b in zip(got["factors"][1:], ref["factors"][1:]):
        scale = np.abs(b).max()
        assert np.abs(a - b).max() <= 1e-3 * scale
    got64 = gpu.spcpstream(dims, cidx, vals, 0, 16, 3, 0.0, init)
    for a, b in zip(got64["factors"][1:], ref["factors"][1:]):
        np.testing.assert_allclose(a, b, rtol=1e-8, atol=1e-10)


def test_spcpstream_final_error_is_cpd_error(gpu):
    """The closing error is cpd_error over the processed stream in relabelled
    coordinates with the time rows as the time factor (oracle cpd_error)."""
    dims, cidx, vals = _stream_tensor(4, [200, 100], 3000, seed=120)
    init = _init(dims, 8, 121)
    got = gpu.spcpstream(dims, cidx, vals, 0, 8, 3, 0.0, init)
    c2, v2 = O.stream_relabel(cidx, vals, 0)
    err, _ = O.cpd_error(c2, v2, [got["time_rows"]] + list(got["factors"][1:]))
    assert abs(got["final_error"] - err) <= 1e-10 * max(1.0, abs(err))


def test_spcpstream_argument_errors(gpu):
    dims, cidx, vals = _stream_tensor(4, [30, 20], 500, seed=122)
    init = _init(dims, 4, 123)
    with pytest.raises(gpu.AltoError):
        gpu.spcpstream(dims, cidx, vals, 0, 4, 0, 0.0, init)  # no outer iteration
    with pytest.raises(gpu.AltoError):
        gpu.spcpstream(dims, cidx, vals, 0, 65, 2, 0.0, _init(dims, 65, 124))
    bad = [c.copy() for c in cidx]
    bad[0] = np.where(bad[0] == 2, 3, bad[0]).astype(np.uint64)  # slice 2 empty
    with pytest.raises(gpu.AltoError, match="empty time slice"):
        gpu.spcpstream(dims, bad, vals, 0, 4, 2, 0.0, init)


@pytest.mark.parametrize("thr,lim", [(0, 2), (2500, 10)])
def test_spcpstream_dynamic_batches_match_oracle(gpu, thr, lim):
    """spCP-stream over batches of several time slices equals the oracle
    (same batches and outer iterations; time rows and factors within 1e-8)."""
    rng = np.random.default_rng(126)
    T, nnz = 7, 7000
    t = np.sort(rng.integers(0, T, size=nnz)).astype(np.uint64)
    t[:T] = np.arange(T)
    cidx = [t, rng.integers(0, 300, nnz).astype(np.uint64), rng.integers(0, 200, nnz).astype(np.uint64)]
    vals = rng.poisson(3, size=nnz) + 1.0
    dims = [T, 300, 200]
    init = _init(dims, 8, 127)
    print("thr", thr, "lim", lim, flush=True)
    ref = O.spcpstream(dims, cidx, vals, 0, 8, 4, 1e-6, init, nnz_threshold=thr, timeslice_limit=lim)
    got = gpu.spcpstream(dims, cidx, vals, 0, 8, 4, 1e-6, init, nnz_threshold=thr, timeslice_limit=lim)
    assert got["slices"] == ref["slices"] < T and got["time_slices"] == ref["time_slices"] == T
    np.testing.assert_array_equal(got["outer_iters"], ref["outer_iters"])
    np.testing.assert_allclose(got["time_rows"], ref["time_rows"], rtol=1e-8, atol=1e-10)
    for a, b in zip(got["factors"][1:], ref["factors"][1:]):
        np.testing.assert_allclose(a, b, rtol=1e-8, atol=1e-10)
