"""GPU parity of the MTTKRP launch knobs and the autotuned default.

Tensors with at least 2^20 schedule entries take the autotuned launch shape
(segment length and kernel chosen by timing, cached per shape bucket).  Every
choice the tuner can make, and the explicit knobs, must give the oracle's
MTTKRP to the same bars as the default path: FP64 within 1e-12, FP32 within
2e-6 of the |X|-MTTKRP bound per element, and BF16 factors (FP32 arithmetic)
within 2e-6 against the oracle fed the same rounded factors (2e-5 with an
FP32 output).
"""
import numpy as np
import pytest

import oracle as O
from conftest import rand_factors, rand_tensor

pytestmark = pytest.mark.gpu

DIMS = [20000, 5000, 300]
NNZ = 1_300_000  # above the tuner's 2^20-entry floor after merging duplicates


@pytest.fixture(scope="module")
def big():
    dims, cidx, vals = rand_tensor(DIMS, NNZ, seed=91, dup_frac=0.02)
    vals = vals - 0.5
    return dims, cidx, vals


def _bf16_round(a):
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def _expect(cidx, vals, F, mode):
    exp = O.mttkrp_coo(cidx, vals, F, mode)
    bnd = O.mttkrp_coo(cidx, np.abs(vals), [np.abs(f) for f in F], mode)
    return exp, bnd


@pytest.mark.parametrize("prec,tol", [("f64", 1e-12), ("f32", 2e-6), ("bf16", 2e-6)])
def test_autotuned_default_matches_oracle(gpu, big, prec, tol):
    dims, cidx, vals = big
    p = {"f64": gpu.PREC_F64, "f32": gpu.PREC_F32, "bf16": gpu.PREC_BF16}[prec]
    R = 32
    F = rand_factors(dims, R, seed=92, lo=-1, hi=1)
    if prec == "f32":
        F = [f.astype(np.float32).astype(np.float64) for f in F]
    elif prec == "bf16":  # BF16 storage, FP32 arithmetic: the oracle takes the rounded factors
        F = [_bf16_round(f) for f in F]
    at = gpu.create_alto(dims, cidx, vals, nprtn=8)
    devF = [gpu.to_prec(f, p) for f in F]
    for mode in range(3):
        print("prec", prec, "mode", mode, flush=True)
        exp, bnd = _expect(cidx, vals, F, mode)
        for call in range(2):  # the first call tunes, the second reads the cache
            out = gpu.DeviceArray.zeros((dims[mode], R), np.float64)
            gpu.mttkrp_alto_par(mode, devF, R, at, out, prec=p, out_f64=True)
            got = out.numpy()
            assert np.all(np.abs(got - exp) <= tol * bnd + 1e-300), (prec, mode, call)
    at.close()


@pytest.mark.parametrize("segment", [1, 4, 16, 64, 128])
@pytest.mark.parametrize("flags", [-1, 3 | 16, 3 | 8, 0, 3 | 32, 3 | 16 | 32])
def test_explicit_knobs_match_oracle(gpu, big, segment, flags):
    dims, cidx, vals = big
    R = 16
    F = rand_factors(dims, R, seed=93, lo=-1, hi=1)
    F32 = [f.astype(np.float32).astype(np.float64) for f in F]
    at = gpu.create_alto(dims, cidx, vals, nprtn=8)
    for p, FF, tol in ((gpu.PREC_F64, F, 1e-12), (gpu.PREC_F32, F32, 2e-6)):
        devF = [gpu.to_prec(f, p) for f in FF]
        for mode in (0, 2):
            print("segment", segment, "flags", flags, "prec", p, "mode", mode, flush=True)
            exp, bnd = _expect(cidx, vals, FF, mode)
            out = gpu.DeviceArray.zeros((dims[mode], R), np.float64)
            gpu.mttkrp_config(mode, devF, R, at, out, prec=p, out_f64=True, algo=2, segment=segment,
                              flags=flags)
            got = out.numpy()
            assert np.all(np.abs(got - exp) <= tol * bnd + 1e-300), (segment, flags, p, mode)
    at.close()


@pytest.mark.parametrize("segment", [4, 64])
@pytest.mark.parametrize("flags", [3, 3 | 64, 3 | 64 | 32, 2 | 64, 3 | 16 | 64])
def test_bf16_fiber_and_decode_match_oracle(gpu, big, segment, flags):
    """BF16 factors through the decode-once kernel (flags 3) and the fiber
    kernel (bit 6), FP32 arithmetic: within 2e-6 of the bound against the
    oracle on the BF16-rounded factors."""
    dims, cidx, vals = big
    R = 32
    F = [_bf16_round(f) for f in rand_factors(dims, R, seed=94, lo=-1, hi=1)]
    at = gpu.create_alto(dims, cidx, vals, nprtn=8)
    devF = [gpu.to_prec(f, gpu.PREC_BF16) for f in F]
    for mode in range(3):
        print("segment", segment, "flags", flags, "mode", mode, flush=True)
        exp, bnd = _expect(cidx, vals, F, mode)
        for out_f64 in (True, False):
            out = gpu.DeviceArray.zeros((dims[mode], R), np.float64 if out_f64 else np.float32)
            gpu.mttkrp_config(mode, devF, R, at, out, prec=gpu.PREC_BF16, out_f64=out_f64, algo=2,
                              segment=segment, flags=flags)
            got = out.numpy().astype(np.float64)
            tol = 2e-6 if out_f64 else 2e-5
            assert np.all(np.abs(got - exp) <= tol * bnd + 1e-300), (segment, flags, mode, out_f64)
    at.close()
