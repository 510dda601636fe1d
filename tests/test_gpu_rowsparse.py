"""GPU row-sparse MTTKRP (rowsparse_mttkrp, src/rowsparse_mttkrp.cpp:97-187)
against the oracle's restatement, through the C ABI.

Bars: the output row indices bit-exact; rows within 1e-12 of the
|X|-MTTKRP bound (the same computation on |values| and |rows|)."""
import numpy as np
import pytest

import oracle as O
from test_oracle import _rowsparse_case

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dims,stream_mode,absent", [([6, 9, 7, 5], -1, 0.0), ([6, 9, 7, 5], 0, 0.2),
                                                      ([1, 300, 200], 0, 0.0), ([40, 30, 20], -1, 0.3)])
@pytest.mark.parametrize("R", [4, 16, 32])
def test_rowsparse_mttkrp_matches_oracle(gpu, dims, stream_mode, absent, R):
    print("rowsparse dims", dims, "stream", stream_mode, "R", R, flush=True)
    cidx, vals, rsp, ridx = _rowsparse_case(dims, 2000, R, 78 + R, stream_mode, absent)
    for mode in range(len(dims)):
        if mode == stream_mode:
            continue
        exp_ind, exp = O.rowsparse_mttkrp(cidx, vals, mode, stream_mode, rsp, ridx, R)
        got_ind, got = gpu.rowsparse_mttkrp(dims, cidx, vals, mode, stream_mode, R, rsp, ridx)
        np.testing.assert_array_equal(got_ind, exp_ind)
        _, bnd = O.rowsparse_mttkrp(cidx, np.abs(vals), mode, stream_mode, [None if r is None else np.abs(r)
                                                                         for r in rsp], ridx, R)
        assert np.all(np.abs(got - exp) <= 1e-12 * bnd + 1e-300), (dims, mode)


def test_rowsparse_mttkrp_rejects_bad_modes(gpu):
    cidx, vals, rsp, ridx = _rowsparse_case([5, 6, 7], 50, 4, 79)
    with pytest.raises(gpu.AltoError, match="invalid argument"):
        gpu.rowsparse_mttkrp([5, 6, 7], cidx, vals, 1, 1, 4, rsp, ridx)
