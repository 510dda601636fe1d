"""GPU cpd_error (src/streaming_cpd.cpp:368-464, kruskal_norm
kruskal_model.cpp:400-442) against the oracle, and CP-stream's closing
error (streaming_cpd.cpp:307-345).

Bars: err and its parts within 1e-10 relative (the COO MTTKRP adds rows
with atomics, so the last bits vary)."""
import numpy as np
import pytest

import oracle as O
from conftest import rand_factors, rand_tensor

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dims,nnz,R", [([40, 30, 20], 3000, 8), ([60, 5, 30, 20], 5000, 16), ([300, 200, 100], 20000, 32)])
def test_cpd_error_matches_oracle(gpu, dims, nnz, R):
    d, c, v = rand_tensor(dims, nnz, seed=400 + R, dup_frac=0.05)
    F = rand_factors(d, R, seed=401 + R, lo=-1, hi=1)
    err, parts = gpu.cpd_error(d, c, v, F)
    rerr, rparts = O.cpd_error(c, v, F)
    assert abs(err - rerr) <= 1e-10 * abs(rerr)
    for a, b in zip(parts, rparts):
        assert abs(a - b) <= 1e-10 * max(abs(b), 1.0)


def test_cpd_error_rejects_out_of_range_coordinates(gpu):
    d, c, v = rand_tensor([10, 10, 10], 100, seed=402)
    c[2][0] = 9
    F = rand_factors([10, 10, 9], 4, seed=403)     # mode 2 declared with 9 rows; a coordinate is 9
    with pytest.raises(gpu.AltoError, match="invalid argument"):
        gpu.cpd_error([10, 10, 9], c, v, F)
    with pytest.raises(gpu.AltoError, match="factor"):
        gpu.cpd_error([10, 10, 10], c, v, F)       # factor shape does not match the dims


@pytest.mark.parametrize("nonneg", [True, False])
def test_cpstream_final_error_matches_oracle(gpu, nonneg):
    from test_gpu_cpstream import _init, _stream_tensor
    dims, cidx, vals = _stream_tensor(6, [300, 200], 6000, seed=404)
    R = 8
    init = _init(dims, R, 405)
    got = gpu.cpstream(dims, cidx, vals, 0, R, 5, 1e-5, nonneg, init)
    ref = O.cpstream(dims, cidx, vals, 0, R, 5, 1e-5, nonneg, init)
    rc, rv = O.stream_relabel(cidx, vals, 0)
    Fr = [ref["time_rows"]] + [ref["factors"][1], ref["factors"][2]]
    rerr, _ = O.cpd_error(rc, rv, Fr)
    assert np.isfinite(got["final_error"])
    assert abs(got["final_error"] - rerr) <= 1e-7 * abs(rerr)


def test_device_coo_takes_device_factors(gpu):
    """With the COO on the device the C ABI reads the factors there too: the
    binding refuses host factors before any call (a kernel reading a host
    address would fault the GPU), and device factors give the host result."""
    d, c, v = rand_tensor([30, 20, 10], 500, seed=71)
    F = rand_factors(d, 8, seed=72)
    dc = [gpu.DeviceArray.from_numpy(x) for x in c]
    dv = gpu.DeviceArray.from_numpy(v)
    with pytest.raises(gpu.AltoError, match="DeviceArray"):
        gpu.cpd_error(d, dc, dv, F, on_device=True)
    dF = [gpu.DeviceArray.from_numpy(f) for f in F]
    err_d, parts_d = gpu.cpd_error(d, dc, dv, dF, on_device=True)
    err_h, parts_h = gpu.cpd_error(d, c, v, F)
    # the <X, M> MTTKRP adds with atomics: equal up to the order of the sums
    np.testing.assert_allclose(err_d, err_h, rtol=1e-12)
    np.testing.assert_allclose(parts_d, parts_h, rtol=1e-12)
    for x in dc + dF + [dv]:
        x.close()
