"""Batched CP-ALS (alto_hip_cpd_als_batch): many independent tensors on
several streams give each tensor the result of cpd_alto on it alone
(cpd.hpp:47-191), which the single-run tests pin to the oracle.

Bars: FP64 fits within 1e-10 and factors within 1e-9 relative of the
single runs (boundary rows are summed with atomics, so the last bits may
differ); FP32 and BF16 fits within 1e-6; one tensor also against the oracle."""
import numpy as np
import pytest

import oracle as O
from conftest import rand_factors, rand_tensor

pytestmark = pytest.mark.gpu


def _instances(n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        nm = 3 if k % 4 else 4
        dims = [int(x) for x in rng.integers(20, 400, size=nm)]
        nnz = int(10 ** rng.uniform(3, 4.5))
        R = (8, 16, 32)[k % 3]
        d, c, v = rand_tensor(dims, nnz, seed=seed + k, dup_frac=0.05)
        out.append((d, c, v, R, rand_factors(d, R, seed=seed + 100 + k)))
    return out


@pytest.mark.parametrize("prec,ftol", [(0, 1e-10), (1, 1e-6), (2, 1e-6)])
@pytest.mark.parametrize("nstreams", [1, 4])
@pytest.mark.parametrize("eps", [1e-30, 0.0])
def test_batch_equals_single_runs(gpu, prec, ftol, nstreams, eps):
    inst = _instances(9, 300)
    ats = [gpu.create_alto(d, c, v, nprtn=4) for d, c, v, _, _ in inst]
    iters = 6
    bf, bl, bfits, wall = gpu.cpd_alto_batch(ats, [x[4] for x in inst], [x[3] for x in inst], iters,
                                             eps=eps, prec=prec, nstreams=nstreams)
    assert wall > 0
    for k, (d, c, v, R, F) in enumerate(inst):
        sf, sl, sfits, _, _ = gpu.cpd_alto(ats[k], F, R, iters, prec=prec)
        assert len(bfits[k]) == len(sfits) == iters
        np.testing.assert_allclose(bfits[k], sfits, rtol=0, atol=ftol)
        if prec == 0:
            np.testing.assert_allclose(bl[k], sl, rtol=1e-9)
            for a, b in zip(bf[k], sf):
                np.testing.assert_allclose(a, b, rtol=1e-9, atol=1e-12)
    for a in ats:
        a.close()


def test_batch_instance_matches_oracle(gpu):
    d, c, v, R, F = _instances(1, 310)[0]
    ref = O.AltoTensor(d, c, v, nprtn=4)
    _, _, rfits, _ = ref.cpd_als(F, R, 5)
    at = gpu.create_alto(d, c, v, nprtn=4)
    _, _, bfits, _ = gpu.cpd_alto_batch([at], [F], [R], 5, prec=0, nstreams=2)
    np.testing.assert_allclose(bfits[0], rfits, rtol=0, atol=1e-8)
    at.close()


def test_batch_rejects_verbose_and_bad_stream_count(gpu):
    d, c, v, R, F = _instances(1, 320)[0]
    at = gpu.create_alto(d, c, v)
    with pytest.raises(gpu.AltoError, match="invalid argument"):
        gpu.cpd_alto_batch([at], [F], [R], 2, nstreams=0)
    at.close()


@pytest.mark.parametrize("prec,ftol", [(0, 1e-10), (1, 1e-6), (2, 1e-6)])
def test_batch_device_models_equal_host_models(gpu, prec, ftol):
    """alto_hip_cpd_als_batch_dev (models in device memory, converted into
    and out of the working precision on the device) gives the host-model
    batch's results: the same kernels run on the same converted values
    (FP64 -> FP32 / BF16 rounds to nearest on the host and on the device).
    The mix holds 3- and 4-mode tensors at ranks 8, 16, 32 (the grouped
    classes) and a rank-24 tensor, which the grouped path does not take (a
    host round trip of its device model).  Workgroup-boundary rows of the
    FP64 MTTKRP outputs are added with atomics, so bars as for the batch
    against single runs: fits within ftol, FP64 lambda and factors within
    1e-9 relative."""
    inst = _instances(7, 700)
    d, c, v = rand_tensor([150, 90, 60], 4000, seed=777, dup_frac=0.05)
    inst.append((d, c, v, 24, rand_factors(d, 24, seed=778)))
    ats = [gpu.create_alto(d, c, v, nprtn=4) for d, c, v, _, _ in inst]
    ranks = [x[3] for x in inst]
    iters = 5
    hf, hl, hfits, _ = gpu.cpd_alto_batch(ats, [x[4] for x in inst], ranks, iters, prec=prec, nstreams=2)
    dF = [[gpu.DeviceArray.from_numpy(np.ascontiguousarray(f)) for f in x[4]] for x in inst]
    dl, dfits, wall = gpu.cpd_alto_batch_dev(ats, dF, ranks, iters, prec=prec, nstreams=2)
    assert wall > 0
    for k in range(len(inst)):
        assert len(hfits[k]) == len(dfits[k]) == iters
        np.testing.assert_allclose(dfits[k], hfits[k], rtol=0, atol=ftol)
        got = [f.numpy() for f in dF[k]]
        if prec == 0:
            np.testing.assert_allclose(dl[k], hl[k], rtol=1e-9)
            for a, b in zip(got, hf[k]):
                np.testing.assert_allclose(a, b, rtol=1e-9, atol=1e-12)
        else:  # the final model is the stored FP32 / BF16 factor, widened
            for a, b in zip(got, hf[k]):
                np.testing.assert_allclose(a, b, rtol=1e-4, atol=1e-6)
    for a in ats:
        a.close()
