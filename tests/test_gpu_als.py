"""GPU CP-ALS (cpd_alto, include/cpd.hpp:47-191) against the oracle at the
same initial factors.  Bars: FP64 path fit within 1e-8 and factors within
1e-6 relative; FP32 path |delta fit| <= 1e-5; BF16 factor storage
|delta fit| <= 2e-3."""
import numpy as np
import pytest

import oracle as O
from conftest import golden, rand_factors, rand_tensor

pytestmark = pytest.mark.gpu


def _setup(dims, nnz, seed, rank, dup=0.0, distinct=False):
    dims, cidx, vals = rand_tensor(dims, nnz, seed=seed, dup_frac=dup, distinct=distinct)
    F = rand_factors(dims, rank, seed=seed + 1)
    return dims, cidx, vals, F


@pytest.mark.parametrize("fit_mode", [O.FIT_REFERENCE_DIAG, O.FIT_EXACT])
def test_als_f64_matches_oracle(gpu, fit_mode):
    dims, cidx, vals, F = _setup([120, 90, 70], 30000, 80, 16, distinct=True)
    ref = O.AltoTensor(dims, cidx, vals, nprtn=8)
    rf, rl, rfits, rfe = ref.cpd_als(F, 16, 6, fit_mode=fit_mode)
    at = gpu.create_alto(dims, cidx, vals, nprtn=8)
    gf, gl, gfits, gfe, tim = gpu.cpd_alto(at, F, 16, 6, prec=gpu.PREC_F64, fit_mode=fit_mode)
    assert len(gfits) == 6
    np.testing.assert_allclose(gfits, rfits, rtol=0, atol=1e-8)
    np.testing.assert_allclose(gfe, rfe, rtol=0, atol=1e-8)
    np.testing.assert_allclose(gl, rl, rtol=1e-6)
    for a, b in zip(gf, rf):
        np.testing.assert_allclose(a, b, rtol=1e-6, atol=1e-9)


@pytest.mark.parametrize("rank,rows", [(32, [300, 200, 100]), (64, [500, 300, 130]), (24, [300, 200, 100])])
def test_als_f64_ranks_match_oracle(gpu, rank, rows):
    """Ranks 32 and 64 take the matrix-core Gram, 24 the LDS one; the row
    counts are not multiples of four (partial row groups)."""
    dims, cidx, vals, F = _setup(rows, 40000, 87, rank, dup=0.05)
    print("rank", rank, "dims", dims, flush=True)
    ref = O.AltoTensor(dims, cidx, vals, nprtn=8)
    rf, rl, rfits, rfe = ref.cpd_als(F, rank, 4, fit_mode=O.FIT_EXACT)
    at = gpu.create_alto(dims, cidx, vals, nprtn=8)
    gf, gl, gfits, gfe, _ = gpu.cpd_alto(at, F, rank, 4, prec=gpu.PREC_F64, fit_mode=gpu.FIT_EXACT)
    np.testing.assert_allclose(gfits, rfits, rtol=0, atol=1e-8)
    np.testing.assert_allclose(gfe, rfe, rtol=0, atol=1e-8)
    np.testing.assert_allclose(gl, rl, rtol=1e-6)
    for a, b in zip(gf, rf):
        np.testing.assert_allclose(a, b, rtol=1e-6, atol=1e-9)


def test_als_f64_fixture_med4(gpu):
    dims, cidx, vals = O.read_tns(golden("med4.tns"))
    F = rand_factors(dims, 8, seed=81)
    ref = O.AltoTensor(dims, cidx, vals, nprtn=16)
    rf, rl, rfits, _ = ref.cpd_als(F, 8, 5)
    at = gpu.create_alto(dims, cidx, vals, nprtn=16)
    gf, gl, gfits, _, _ = gpu.cpd_alto(at, F, 8, 5, prec=gpu.PREC_F64)
    np.testing.assert_allclose(gfits, rfits, rtol=0, atol=1e-8)


@pytest.mark.parametrize("rank", [16, 32, 64])
@pytest.mark.parametrize("prec,tol", [(1, 1e-5), (2, 2e-3)])
def test_als_reduced_precision_fit(gpu, prec, tol, rank):
    dims, cidx, vals, F = _setup([300, 200, 150, 40], 60000, 82, rank, dup=0.1)
    ref = O.AltoTensor(dims, cidx, vals, nprtn=8)
    _, _, rfits, rfe = ref.cpd_als(F, rank, 5)
    at = gpu.create_alto(dims, cidx, vals, nprtn=8)
    _, _, gfits, gfe, _ = gpu.cpd_alto(at, F, rank, 5, prec=prec)
    print("prec", prec, "rank", rank, "max |dfit|", np.abs(gfits - rfits).max(), flush=True)
    assert np.abs(gfits - rfits).max() <= tol
    assert np.abs(gfe - rfe).max() <= tol


def test_als_128bit_keys(gpu):
    dims, cidx, vals, F = _setup([1 << 20, 1 << 20, 1 << 18, 1 << 16, 1 << 12], 20000, 83, 8)
    ref = O.AltoTensor(dims, cidx, vals, nprtn=4)
    _, _, rfits, _ = ref.cpd_als(F, 8, 3)
    at = gpu.create_alto(dims, cidx, vals, nprtn=4)
    assert at.kw == 2
    _, _, gfits, _, _ = gpu.cpd_alto(at, F, 8, 3, prec=gpu.PREC_F64)
    np.testing.assert_allclose(gfits, rfits, rtol=0, atol=1e-8)


@pytest.mark.parametrize("prec,tol", [(0, 1e-8), (2, 2e-3)])
def test_als_5mode_128bit_rank64(gpu, prec, tol):
    """Five modes of 2^14 (70 key bits: 128-bit keys, two packed words),
    rank 64: the decode-once kernel with four gathered modes, the
    matrix-core Gram and solve and the four-wave inverse."""
    dims, cidx, vals, F = _setup([1 << 14] * 5, 30000, 88, 64)
    ref = O.AltoTensor(dims, cidx, vals, nprtn=4)
    rf, rl, rfits, rfe = ref.cpd_als(F, 64, 3, fit_mode=O.FIT_EXACT)
    at = gpu.create_alto(dims, cidx, vals, nprtn=4)
    assert at.kw == 2
    gf, gl, gfits, gfe, _ = gpu.cpd_alto(at, F, 64, 3, prec=prec, fit_mode=gpu.FIT_EXACT)
    print("prec", prec, "max |dfit|", np.abs(gfits - rfits).max(), flush=True)
    assert np.abs(gfits - rfits).max() <= tol
    assert np.abs(gfe - rfe).max() <= tol
    if prec == 0:  # the model itself, not only its fit
        np.testing.assert_allclose(gl, rl, rtol=1e-6)
        for a, b in zip(gf, rf):
            np.testing.assert_allclose(a, b, rtol=1e-6, atol=1e-9)


def test_als_epsilon_stop_and_exact_fit_is_residual(gpu):
    dims, cidx, vals, F = _setup([8, 7, 6], 100, 84, 3, distinct=True)
    at = gpu.create_alto(dims, cidx, vals)
    _, _, fits, _, _ = gpu.cpd_alto(at, F, 3, 20, eps=1e30)
    assert len(fits) == 2
    gf, gl, fits, fe, _ = gpu.cpd_alto(at, F, 3, 4, fit_mode=gpu.FIT_EXACT)
    X = np.zeros(dims)
    np.add.at(X, tuple(c.astype(np.int64) for c in cidx), vals)
    M = np.einsum("ir,jr,kr,r->ijk", gf[0], gf[1], gf[2], gl)
    assert abs(fe[-1] - np.linalg.norm(X - M) / np.linalg.norm(X)) < 1e-9


def test_als_f64_signed_init_matches_oracle(gpu):
    # the reference initialises with +-3 U(0,1) (util.cpp:28-35): signed
    # factors exercise the signed max-norm (kruskal_model.cpp:277-317)
    dims, cidx, vals = rand_tensor([150, 100, 80], 20000, seed=85)
    F = rand_factors(dims, 8, seed=86, lo=-3, hi=3)
    ref = O.AltoTensor(dims, cidx, vals, nprtn=4)
    rf, rl, rfits, rfe = ref.cpd_als(F, 8, 5, fit_mode=O.FIT_EXACT)
    at = gpu.create_alto(dims, cidx, vals, nprtn=4)
    gf, gl, gfits, gfe, _ = gpu.cpd_alto(at, F, 8, 5, prec=gpu.PREC_F64, fit_mode=gpu.FIT_EXACT)
    np.testing.assert_allclose(gfe, rfe, rtol=0, atol=1e-8)
    np.testing.assert_allclose(gl, rl, rtol=1e-6)


@pytest.mark.parametrize("prec,tol", [(0, 1e-10), (1, 1e-6), (2, 1e-6)])
def test_als_graph_replay_matches_checked_loop(gpu, prec, tol):
    """epsilon <= 0 runs every iteration without the per-iteration check:
    iteration 0 directly, the rest as replays of one captured graph.  Every
    replay's fit and the final model equal the checked loop's."""
    dims, cidx, vals, F = _setup([200, 150, 90, 30], 40000, 86, 16, dup=0.1)
    at = gpu.create_alto(dims, cidx, vals, nprtn=8)
    sf, sl, sfits, sfe, _ = gpu.cpd_alto(at, F, 16, 7, eps=1e-30, prec=prec)
    gf, gl, gfits, gfe, tim = gpu.cpd_alto(at, F, 16, 7, eps=0.0, prec=prec)
    assert len(gfits) == len(sfits) == 7
    np.testing.assert_allclose(gfits, sfits, rtol=0, atol=tol)
    np.testing.assert_allclose(gfe, sfe, rtol=0, atol=tol)
    np.testing.assert_allclose(gl, sl, rtol=10 * tol)
    for a, b in zip(gf, sf):
        np.testing.assert_allclose(a, b, rtol=10 * tol, atol=10 * tol)
    assert tim[0] > 0 and tim[7] > 0


@pytest.mark.parametrize("R", [4, 8, 32])
def test_cpd_fit_entry_matches_closed_forms(gpu, R):
    """alto_hip_cpd_fit (cpd_fit_alto, cpd.hpp:193-275): the reference's
    diagonal convention equals the sum of the dense rank-1 terms' squared
    norms, the exact one ||X - M|| / ||X||, and both equal the oracle's."""
    dims, cidx, vals = rand_tensor([9, 7, 6], 150, seed=97, distinct=True)
    F = rand_factors(dims, R, seed=98)
    lam = np.random.default_rng(99).uniform(0.5, 2.0, R)
    X = np.zeros(dims)
    np.add.at(X, tuple(c.astype(np.int64) for c in cidx), vals)
    terms = [lam[r] * np.einsum("i,j,k->ijk", F[0][:, r], F[1][:, r], F[2][:, r]) for r in range(R)]
    M = sum(terms)
    nx = float((X ** 2).sum())
    mt = np.einsum("ijk,ir,jr->kr", X, F[0], F[1])
    grams = [O.gram_upper(f) for f in F]
    inner = float((X * M).sum())
    r = nx + sum(float((t ** 2).sum()) for t in terms) - 2 * inner
    want_diag = (np.sqrt(r) if r > 0 else r) / np.sqrt(nx)
    want_exact = np.linalg.norm(X - M) / np.sqrt(nx)
    got_diag = gpu.cpd_fit(F[2], mt, lam, grams, nx, gpu.FIT_REFERENCE_DIAG)
    got_exact = gpu.cpd_fit(F[2], mt, lam, grams, nx, gpu.FIT_EXACT)
    assert abs(got_diag - want_diag) < 1e-10 * max(1.0, abs(want_diag))
    assert abs(got_exact - want_exact) < 1e-10
    assert abs(got_diag - O.cpd_fit(dims, F, lam, grams, mt, nx, O.FIT_REFERENCE_DIAG)) < 1e-12 * max(1.0, abs(want_diag))
