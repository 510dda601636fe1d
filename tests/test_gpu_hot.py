"""GPU parity of the hot-row MTTKRP (one workgroup per CU, the most popular
factor rows of the gathered modes in LDS, re-encoded schedule keys, FP64
boundary combining) against the oracle's COO MTTKRP (src/mttkrp.cpp:122-156).

Every workgroup shape (1024 / 768 / 512 threads), ranks 16, 32 and 64, FP64,
FP32 and BF16 factors, FP32 and FP64 outputs, 3-, 4- and 5-mode tensors with
64- and 128-bit keys, skewed (Zipf) and uniform coordinates, tensors with
fewer entries than lane groups, and the accumulate / zeroed-output modes.
Bars as the other MTTKRP tests: FP64 within 1e-12 of the |X|-MTTKRP bound per
element, FP32 within 2e-6 (FP64 output) or 2e-5 (FP32 output), BF16 against
the oracle fed the same rounded factors at the same bars.
"""
import numpy as np
import pytest

import oracle as O
from conftest import rand_factors, rand_tensor

pytestmark = pytest.mark.gpu

HOT = 128          # flags bit 7: hot-row kernel
ZEROED = 32        # flags bit 5: the output holds zeros on entry


def _bf16_round(a):
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def _zipf(gpu, dims, nnz, seed):
    dc, dv = gpu.generate(dims, nnz, seed, [1] * len(dims), [1.1] * len(dims), 1, 0.0, 1.0)
    gpu.synchronize()
    cidx = [c.numpy() for c in dc]
    vals = dv.numpy() - 1.0  # signed values: the bound is not the MTTKRP itself
    return dims, cidx, vals


def _cases(gpu):
    return {
        "zipf3": _zipf(gpu, [40000, 9000, 700], 400_000, 11),
        "zipf4": _zipf(gpu, [50000, 20000, 5000, 100], 300_000, 12),
        "u128_5mode": rand_tensor([1 << 20, 1 << 20, 1 << 18, 1 << 16, 1 << 12], 60000, seed=13, dup_frac=0.1),
        "tiny": rand_tensor([30, 20, 10], 500, seed=14, dup_frac=0.2),
        "one_row_out": rand_tensor([1, 3000, 200], 20000, seed=15),
    }


@pytest.fixture(scope="module")
def cases(gpu):
    out = {}
    for name, (dims, cidx, vals) in _cases(gpu).items():
        at = gpu.create_alto(dims, cidx, vals, nprtn=8)
        out[name] = (dims, cidx, vals, at)
    return out


def _factors(dims, R, prec, seed):
    F = rand_factors(dims, R, seed=seed, lo=-1, hi=1)
    if prec == "f32":
        F = [f.astype(np.float32).astype(np.float64) for f in F]
    elif prec == "bf16":
        F = [_bf16_round(f) for f in F]
    return F


PREC = {"f64": 0, "f32": 1, "bf16": 2}


@pytest.mark.parametrize("name", ["zipf3", "zipf4", "u128_5mode", "tiny", "one_row_out"])
@pytest.mark.parametrize("prec,out64", [("f64", True), ("f32", True), ("f32", False), ("bf16", False)])
@pytest.mark.parametrize("R", [16, 32, 64])
def test_hot_kernel_matches_oracle(gpu, cases, name, prec, out64, R):
    dims, cidx, vals, at = cases[name]
    F = _factors(dims, R, prec, seed=R + 7)
    p = PREC[prec]
    devF = [gpu.to_prec(f, p) for f in F]
    tol = 1e-12 if prec == "f64" else (2e-6 if out64 else 2e-5)
    odt = np.float64 if out64 else np.float32
    for mode in range(len(dims)):
        exp = O.mttkrp_coo(cidx, vals, F, mode)
        bnd = O.mttkrp_coo(cidx, np.abs(vals), [np.abs(f) for f in F], mode)
        base = np.random.default_rng(mode).uniform(-1, 1, exp.shape).astype(odt)
        for shape in range(3):
            print("case", name, prec, "out64", out64, "R", R, "mode", mode, "shape", shape, flush=True)
            # accumulate into a non-zero output
            out = gpu.DeviceArray.from_numpy(base)
            gpu.mttkrp_config(mode, devF, R, at, out, prec=p, out_f64=out64, algo=2,
                              flags=3 | HOT | (shape << 9))
            got = out.numpy().astype(np.float64) - base.astype(np.float64)
            err = np.abs(got - exp) / np.maximum(bnd + np.abs(base), 1e-300)
            assert err.max() <= tol, ("accumulate", shape, err.max())
            # zeroed output: owned rows are stored
            out = gpu.DeviceArray.zeros(exp.shape, odt)
            gpu.mttkrp_config(mode, devF, R, at, out, prec=p, out_f64=out64, algo=2,
                              flags=3 | HOT | ZEROED | (shape << 9))
            got = out.numpy().astype(np.float64)
            err = np.abs(got - exp) / np.maximum(bnd, 1e-300)
            err[bnd == 0] = np.abs(got - exp)[bnd == 0]
            assert err.max() <= tol, ("zeroed", shape, err.max())


def test_hot_kernel_is_deterministic(gpu, cases):
    """No atomics: repeated calls give identical bits."""
    dims, cidx, vals, at = cases["zipf4"]
    F = _factors(dims, 32, "f32", seed=3)
    devF = [gpu.to_prec(f, 1) for f in F]
    for mode in range(4):
        outs = []
        for _ in range(3):
            out = gpu.DeviceArray.zeros((dims[mode], 32), np.float32)
            gpu.mttkrp_config(mode, devF, 32, at, out, prec=1, out_f64=False, algo=2, flags=3 | HOT | ZEROED)
            outs.append(out.numpy())
        assert all(np.array_equal(outs[0], o) for o in outs[1:])


def test_hot_kernel_rejects_unsupported_rank(gpu, cases):
    dims, cidx, vals, at = cases["zipf3"]
    F = _factors(dims, 8, "f64", seed=4)
    devF = [gpu.to_prec(f, 0) for f in F]
    out = gpu.DeviceArray.zeros((dims[0], 8), np.float64)
    with pytest.raises(gpu.AltoError):
        gpu.mttkrp_config(0, devF, 8, at, out, prec=0, out_f64=True, algo=2, flags=3 | HOT)
