"""The R x R inverse of the dense step (pseudo_inverse, cpd.hpp:377-480)
against the oracle's restatement (Cholesky + per-row solves; pinned in the
oracle tests), at every rank with its own kernel: 8/16/32/64 (Gauss-Jordan),
4 and 24 (one-wave Cholesky), 72 (256-thread Cholesky); and the fallback
(V not positive definite) at ranks 4 and 8, which must take the diagonal
minimum-norm solve.  Bar: 1e-9 relative for the well-conditioned cases,
exact flags."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("R", [4, 8, 16, 24, 32, 64, 72])
def test_pseudo_inverse_matches_oracle(gpu, R):
    rng = np.random.default_rng(150 + R)
    facs = [rng.random((3 * R, R)) for _ in range(3)]
    grams = [f.T @ f for f in facs]
    B = rng.normal(size=(40, R))
    for mode in (0, 2):
        ref, _, fb = O.pseudo_inverse_als(grams, mode, B)
        got, gfb = gpu.pseudo_inverse(grams, mode, B)
        assert fb == gfb == 0
        np.testing.assert_allclose(got, ref, rtol=1e-9, atol=1e-12 * np.abs(ref).max())


@pytest.mark.parametrize("R", [4, 8])
def test_pseudo_inverse_fallback(gpu, R):
    """V + 1e-3 I indefinite (a negative diagonal entry); Grams as the
    reference holds them (upper triangle written, lower zero)."""
    rng = np.random.default_rng(160 + R)
    g1 = np.eye(R) * 2.0 + np.triu(np.full((R, R), 0.5), 1)
    d = rng.uniform(1.0, 3.0, R)
    d[1] = -1.0
    g2 = np.diag(d) + np.triu(np.full((R, R), 0.25), 1)
    grams = [np.full((R, R), 7.0), g1, g2]
    B = rng.normal(size=(5, R))
    ref, _, fb = O.pseudo_inverse_als(grams, 0, B)
    got, gfb = gpu.pseudo_inverse(grams, 0, B)
    assert fb == gfb == 1
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=0)
