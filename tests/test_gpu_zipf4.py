"""GPU parity on the headline workload's shape: 4-mode 500000 x 200000 x
50000 x 1000, every index Zipf(1.1) over a seeded permutation of its mode,
lognormal(0, 1) values, rank 32, made by the same device generator and seed
as the benchmark, at 3,000,000 nonzeros (the oracle finishes in seconds and
the hottest rows still collect hundreds of thousands of merged entries).

The MTTKRP runs the default path -- output-ordered schedule, autotuned launch
shape, every kernel the tuner may pick -- for every mode, in FP64 (FP64
output), FP32 and BF16 factor storage (FP32 output, the benchmark's launch),
and is compared with the oracle's COO MTTKRP (src/mttkrp.cpp:122-156, the
reference's -c check, main.cpp:532-649).  Values and factors are positive,
so the |X|-MTTKRP bound of the other MTTKRP tests is the MTTKRP itself:
FP64 within 1e-12 relative per element, FP32 within 2e-5 (FP32 output:
runs accumulate in FP32; the partial sums of rows that cross workgroup
boundaries are added in FP64), BF16 within 2e-5 against the oracle fed the
same rounded factors.  The FP32 output takes no atomics, so repeated calls
give identical bits.

CP-ALS (cpd_alto, include/cpd.hpp:47-191), 3 iterations from the same
initial factors: the plain FP64 path equals the oracle (fits within 1e-8,
lambda and factors within 1e-6 relative); FP32 and BF16 storage are held to
stated bounds on the relative Frobenius error of every factor and of lambda
against the oracle's FP64 model.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

DIMS = [500000, 200000, 50000, 1000]
NNZ = 3_000_000
SEED = 0xA1700001
R = 32


def _bf16_round(a):
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


@pytest.fixture(scope="module")
def zipf4(gpu):
    dcidx, dvals = gpu.generate(DIMS, NNZ, SEED, [1, 1, 1, 1], [1.1] * 4, 1, 0.0, 1.0)
    gpu.synchronize()
    cidx = [c.numpy() for c in dcidx]
    vals = dvals.numpy()
    at = gpu.create_alto(DIMS, dcidx, dvals, nprtn=64, on_device=True)
    for c in dcidx:
        c.close()
    dvals.close()
    # the workload's skew: the hottest output rows hold a large share of the entries
    top = max(int(np.bincount(c.astype(np.int64)).max()) for c in cidx)
    print("zipf4: %d nnz, hottest row %d entries" % (NNZ, top), flush=True)
    assert top > 100_000
    return cidx, vals, at


@pytest.mark.parametrize("prec,tol", [("f64", 1e-12), ("f32", 2e-5), ("bf16", 2e-5)])
def test_zipf4_default_mttkrp_matches_oracle(gpu, zipf4, prec, tol):
    cidx, vals, at = zipf4
    p = {"f64": gpu.PREC_F64, "f32": gpu.PREC_F32, "bf16": gpu.PREC_BF16}[prec]
    rng = np.random.default_rng(SEED)
    F = [rng.random((d, R)) for d in DIMS]
    if prec == "f32":
        F = [f.astype(np.float32).astype(np.float64) for f in F]
    elif prec == "bf16":
        F = [_bf16_round(f) for f in F]
    devF = [gpu.to_prec(f, p) for f in F]
    out_f64 = prec == "f64"
    for mode in range(4):
        print("prec", prec, "mode", mode, flush=True)
        exp = O.mttkrp_coo(cidx, vals, F, mode)
        for call in range(2):  # the first call builds the schedule and tunes
            out = gpu.DeviceArray.zeros((DIMS[mode], R), np.float64 if out_f64 else np.float32)
            gpu.mttkrp_alto_par(mode, devF, R, at, out, prec=p, out_f64=out_f64)
            got = out.numpy().astype(np.float64)
            err = np.abs(got - exp) / np.maximum(exp, 1e-300)
            print("  call %d max rel err %.3g" % (call, err.max()), flush=True)
            assert err.max() <= tol, (prec, mode, call, err.max())
            out.close()


def test_zipf4_fp32_output_is_deterministic(gpu, zipf4):
    cidx, vals, at = zipf4
    rng = np.random.default_rng(SEED + 1)
    devF = [gpu.to_prec(rng.random((d, R)), gpu.PREC_F32) for d in DIMS]
    for mode in range(4):
        outs = []
        for _ in range(3):
            out = gpu.DeviceArray.zeros((DIMS[mode], R), np.float32)
            gpu.mttkrp_alto_par(mode, devF, R, at, out, prec=gpu.PREC_F32, out_f64=False)
            outs.append(out.numpy())
            out.close()
        assert all(np.array_equal(outs[0], o) for o in outs[1:]), mode


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def zipf4_als_ref(zipf4):
    cidx, vals, _ = zipf4
    rng = np.random.default_rng(SEED + 7)
    F = [rng.random((d, R)) for d in DIMS]
    ref = O.AltoTensor(DIMS, cidx, vals, nprtn=64)
    rf, rl, rfits, rfe = ref.cpd_als(F, R, 3, fit_mode=O.FIT_EXACT)
    return F, rf, rl, rfits, rfe


# FP32: MTTKRP rows within ~1e-6 relative (measured above), amplified by the
# solve: the Hadamard of U(0,1) Grams is ill-conditioned, and FP32 ends ~40x
# above its MTTKRP error (4e-5).  BF16 storage rounds every factor entry by up
# to 2^-9 relative before each MTTKRP, so its MTTKRP rows of few entries err by
# ~1e-3 and the same amplification gives a few percent.
@pytest.mark.parametrize("prec,ftol", [(0, 1e-6), (1, 1e-4), (2, 1e-1)])
def test_zipf4_cp_als_matches_oracle(gpu, zipf4, zipf4_als_ref, prec, ftol):
    _, _, at = zipf4
    F, rf, rl, rfits, rfe = zipf4_als_ref
    gf, gl, gfits, gfe, _ = gpu.cpd_alto(at, F, R, 3, prec=prec, fit_mode=gpu.FIT_EXACT)
    lam_err = _rel(gl, rl)
    f_err = [_rel(a, b) for a, b in zip(gf, rf)]
    print("prec %d: lambda rel err %.3g, factor rel errs %s, fits %s vs %s"
          % (prec, lam_err, ["%.3g" % e for e in f_err], gfe, rfe), flush=True)
    if prec == 0:
        np.testing.assert_allclose(gfits, rfits, rtol=0, atol=1e-8)
        np.testing.assert_allclose(gfe, rfe, rtol=0, atol=1e-8)
        np.testing.assert_allclose(gl, rl, rtol=1e-6)
        for a, b in zip(gf, rf):
            np.testing.assert_allclose(a, b, rtol=1e-6, atol=1e-9)
    assert lam_err <= ftol
    assert max(f_err) <= ftol
