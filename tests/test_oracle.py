"""Pins the CPU oracle (oracle/alto_oracle.c) against known answers from the
reference's fixtures, closed forms, and independent mathematics (dense
tensors, numpy/scipy linear algebra), so that a dropped term, a wrong index or
a transposed operand in the oracle fails here."""
import numpy as np
import pytest

import oracle as O
from conftest import golden, rand_factors, rand_tensor


@pytest.fixture(autouse=True, scope="module")
def _single_thread():
    O.set_threads(1)


# ---------- masks (setup_packed_alto, alto.hpp:954-1057) ----------

def test_masks_small_kat():
    dims, cidx, vals = O.read_tns(golden("small.tns"))
    assert dims == [2, 3, 2]
    m, total = O.alto_masks(dims)
    assert total == 4
    assert [int(x) for x in m[:, 0]] == [0x1, 0xC, 0x2]


@pytest.mark.parametrize("dims,expect", [
    ([1000, 800, 600], [0x9249249, 0x12492492, 0x24924924]),
    ([500000, 200000, 50000, 1000], [0x6a49248888888888, 0x1524924444444444, 0x92492222222222, 0x1111111111]),
])
def test_masks_config_kat(dims, expect):
    m, total = O.alto_masks(dims)
    assert [int(x) for x in m[:, 0]] == expect


def test_masks_128_partition_the_key():
    dims = [1 << 20, 1 << 20, 1 << 18, 1 << 16, 1 << 12]
    m, total = O.alto_masks(dims)
    assert total == 86 and m.shape == (5, 2)
    whole = 0
    for n, d in enumerate(dims):
        v = int(m[n, 0]) | (int(m[n, 1]) << 64)
        assert bin(v).count("1") == O.num_bits(d)
        assert whole & v == 0
        whole |= v
    assert whole == (1 << 86) - 1


def test_num_bits_size_one_mode():
    assert O.num_bits(1) == 0 and O.num_bits(2) == 1 and O.num_bits(1024) == 10 and O.num_bits(1025) == 11


def test_key_too_wide_is_rejected():
    with pytest.raises(ValueError):
        O.alto_masks([1 << 40, 1 << 40], kw=1)


# ---------- pdep / pext (bitops.hpp) ----------

def _pdep_brute(x, mask):
    out, b = 0, 0
    for pos in range(128):
        if (mask >> pos) & 1:
            if (x >> b) & 1:
                out |= 1 << pos
            b += 1
    return out


@pytest.mark.parametrize("dims", [[1000, 800, 600], [500000, 200000, 50000, 1000],
                                  [1 << 20, 1 << 20, 1 << 18, 1 << 16, 1 << 12]])
def test_pdep_pext_round_trip(dims):
    m, _ = O.alto_masks(dims)
    rng = np.random.default_rng(3)
    for n, d in enumerate(dims):
        mask = int(m[n, 0]) | ((int(m[n, 1]) << 64) if m.shape[1] == 2 else 0)
        for x in list(rng.integers(0, d, size=20)) + [0, d - 1]:
            dep = O.pdep(int(x), m[n])
            v = int(dep[0]) | ((int(dep[1]) << 64) if len(dep) == 2 else 0)
            assert v == _pdep_brute(int(x), mask)
            assert O.pext(dep, m[n]) == int(x)


# ---------- sort + partitions ----------

def test_small_sorted_pairs_kat():
    dims, cidx, vals = O.read_tns(golden("small.tns"))
    at = O.AltoTensor(dims, cidx, vals, nprtn=8)
    assert [int(k) for k in at.keys[:, 0]] == [1, 2, 4, 7, 8, 9]
    assert list(at.vals) == [4.0, 3.0, 1.0, 6.0, 2.0, 5.0]


def test_alt_pext_small_kat():
    """ALT_PEXT re-encoding of small.tns (dims 2x3x2, bits 1/2/1, masks
    0x1/0xC/0x2): by hand, key 1 -> (1,0,0) -> 1; 2 -> (0,0,1) -> 8;
    4 -> (0,1,0) -> 2; 7 -> (1,1,1) -> 11; 8 -> (0,2,0) -> 4; 9 -> (1,2,0) -> 5
    with fields at bit positions 0, 1, 3."""
    dims, cidx, vals = O.read_tns(golden("small.tns"))
    at = O.AltoTensor(dims, cidx, vals, nprtn=8)
    packed, pos = O.alt_pext(dims, at.keys, at.masks)
    assert list(pos) == [0, 1, 3]
    assert [int(k) for k in packed[:, 0]] == [1, 8, 2, 11, 4, 5]


def test_alt_pext_fields_are_the_sorted_coordinates():
    """Each packed field is the coordinate of the sorted nonzero (128-bit
    keys included)."""
    rng = np.random.default_rng(42)
    dims = [1 << 20, 1 << 20, 1 << 18, 1 << 16, 1 << 12]
    cidx = [rng.integers(0, d, size=300).astype(np.uint64) for d in dims]
    at = O.AltoTensor(dims, cidx, rng.random(300), nprtn=4)
    packed, pos = O.alt_pext(dims, at.keys, at.masks)
    for i in range(300):
        v = int(packed[i, 0]) | (int(packed[i, 1]) << 64)
        src = int(at.perm[i])
        for m, d in enumerate(dims):
            w = O.num_bits(d)
            assert (v >> int(pos[m])) & ((1 << w) - 1) == int(cidx[m][src])


def test_partition_closed_form_kat():
    dims, cidx, vals = O.read_tns(golden("small.tns"))
    at = O.AltoTensor(dims, cidx, vals, nprtn=8)
    assert list(at.prtn_ptr) == [0, 1, 2, 3, 4, 5, 6, 6, 6]
    assert at.intervals[6].tolist() == [[0, 0]] * 3 and at.intervals[7].tolist() == [[0, 0]] * 3


def test_sort_is_ascending_and_stable_on_ties():
    dims, cidx, vals = rand_tensor([7, 5, 3], 300, seed=1, dup_frac=0.5)
    at = O.AltoTensor(dims, cidx, vals, nprtn=3)
    k = at.keys[:, 0]
    assert np.all(k[1:] >= k[:-1])
    perm = at.perm
    for i in range(1, len(k)):
        if k[i] == k[i - 1]:
            assert perm[i] > perm[i - 1]
    # keys decode back to the input coordinates of the permuted nonzero
    for i in range(0, len(k), 17):
        for n in range(3):
            assert O.pext(at.keys[i], at.masks[n]) == int(cidx[n][perm[i]])


def test_partition_intervals_brute_force():
    dims, cidx, vals = rand_tensor([50, 40, 30], 1000, seed=2)
    at = O.AltoTensor(dims, cidx, vals, nprtn=7)
    per = -(-1000 // 7)
    for p in range(7):
        s, e = min(p * per, 1000), min((p + 1) * per, 1000)
        assert at.prtn_ptr[p + 1] == e
        for n in range(3):
            idx = [O.pext(at.keys[i], at.masks[n]) for i in range(s, e)]
            assert at.intervals[p, n].tolist() == [min(idx), max(idx)]


# ---------- conflict-resolution bits (GENERATE_CR_BITS, alto.hpp:1250-1346) ----------

def _cr_case(dims, nnz, seed, hot_rows=None):
    """Uniform coordinates; with hot_rows, mode 0 draws from that many rows
    only, so that rows recur across partitions."""
    dims, cidx, vals = rand_tensor(dims, nnz, seed=seed)
    if hot_rows:
        rng = np.random.default_rng(seed + 1)
        cidx[0] = rng.integers(0, hot_rows, size=nnz).astype(np.uint64)
    return dims, cidx, vals


def _cr_brute(dims, cidx, perm, ptr, cr_masks):
    """Independent statement: a nonzero gets mode n's bit iff the rows of
    mode n it lies in is held by nonzeros of two or more partitions."""
    nnz = len(perm)
    part = np.searchsorted(np.asarray(ptr[1:], dtype=np.int64), np.arange(nnz), side="right")
    extra = np.zeros((nnz, cr_masks.shape[1]), dtype=np.uint64)
    for n in range(len(dims)):
        if not cr_masks[n].any():
            continue
        rows = np.asarray(cidx[n], dtype=np.int64)[np.asarray(perm, dtype=np.int64)]
        owners = {}
        for r, p in zip(rows.tolist(), part.tolist()):
            owners.setdefault(r, set()).add(p)
        hit = np.array([len(owners[r]) > 1 for r in rows.tolist()])
        extra[hit] |= cr_masks[n]
    return extra


@pytest.mark.parametrize("nprtn", [1, 2, 7])
def test_cr_bits_mode_choice_and_brute_force(nprtn):
    # bits: 11 + 10 + 10 + 3 = 34; nnz 3000: reuse 1, 4, 5, 600; mode 3 has
    # 5 <= 512 rows.  So mode 0 takes bit 34 and mode 1 bit 35 (reuse 4 is
    # still <= MIN_FIBER_REUSE); mode 2 (reuse 5) and mode 3 take none.
    dims, cidx, vals = _cr_case([2000, 700, 600, 5], 3000, seed=40, hot_rows=120)
    at = O.AltoTensor(dims, cidx, vals, nprtn=nprtn)
    keys, cm = O.cr_bits(dims, at.keys, at.masks, nprtn, at.prtn_ptr)
    assert cm[:, 0].tolist() == [1 << 34, 1 << 35, 0, 0]
    extra = _cr_brute(dims, cidx, at.perm, at.prtn_ptr, cm)
    assert np.array_equal(keys, at.keys | extra)
    assert np.array_equal(keys & ~np.uint64((1 << 34) | (1 << 35)), at.keys)
    if nprtn == 1:
        assert not extra.any()
    else:  # conflicts occur, and mode 1 (uniform rows) also has rows held by one partition
        col0 = (extra[:, 0] >> np.uint64(34)) & np.uint64(1)
        col1 = (extra[:, 0] >> np.uint64(35)) & np.uint64(1)
        assert int(col0.sum()) > 0
        assert 0 < int(col1.sum()) < len(col1)


def test_cr_bits_thresholds_are_strict_and_inclusive():
    # dims 512 is not > MIN_NUM_FIBERS; 513 is.  nnz 2565: reuse 5 for 513
    # rows (no bit), so use nnz 2052: reuse 4 (bit).
    dims, cidx, vals = _cr_case([512, 513, 4], 2052, seed=41)
    at = O.AltoTensor(dims, cidx, vals, nprtn=3)
    _, cm = O.cr_bits(dims, at.keys, at.masks, 3, at.prtn_ptr)
    assert cm[:, 0].tolist() == [0, 1 << 21, 0]
    dims, cidx, vals = _cr_case([512, 513, 4], 2565, seed=41)
    at = O.AltoTensor(dims, cidx, vals, nprtn=3)
    _, cm = O.cr_bits(dims, at.keys, at.masks, 3, at.prtn_ptr)
    assert not cm.any()


def test_cr_bits_run_out_of_free_bits():
    # 21 + 21 + 21 = 63 bits: one free bit, which mode 0 takes (bit 63)
    dims, cidx, vals = _cr_case([1 << 21, 1 << 21, 1 << 21], 400, seed=42, hot_rows=30)
    at = O.AltoTensor(dims, cidx, vals, nprtn=4)
    keys, cm = O.cr_bits(dims, at.keys, at.masks, 4, at.prtn_ptr)
    assert cm[:, 0].tolist() == [1 << 63, 0, 0]
    assert np.array_equal(keys, at.keys | _cr_brute(dims, cidx, at.perm, at.prtn_ptr, cm))


def test_cr_bits_128_bit_keys():
    # 30 + 30 + 10 = 70 bits: two-word keys; all three modes qualify and take
    # bits 70, 71, 72 (word 1, bits 6, 7, 8)
    dims, cidx, vals = _cr_case([1 << 30, 1 << 30, 1000], 500, seed=43, hot_rows=40)
    at = O.AltoTensor(dims, cidx, vals, nprtn=5)
    assert at.kw == 2
    keys, cm = O.cr_bits(dims, at.keys, at.masks, 5, at.prtn_ptr)
    assert cm.tolist() == [[0, 1 << 6], [0, 1 << 7], [0, 1 << 8]]
    extra = _cr_brute(dims, cidx, at.perm, at.prtn_ptr, cm)
    assert extra[:, 1].any()
    assert np.array_equal(keys, at.keys | extra)


def test_get_num_ptrn():
    # util.hpp:30-48
    assert O.get_num_ptrn(5, 8) == 1
    assert O.get_num_ptrn(1000, 8) == 8
    assert O.get_num_ptrn(40, 8) == 4


# ---------- MTTKRP ----------

def _dense(dims, cidx, vals):
    X = np.zeros(dims)
    np.add.at(X, tuple(c.astype(np.int64) for c in cidx), vals)
    return X


def _mttkrp_dense(X, factors, mode):
    N = X.ndim
    letters = "abcdefgh"[:N]
    ins = [letters] + [letters[m] + "r" for m in range(N) if m != mode]
    ops = [X] + [factors[m] for m in range(N) if m != mode]
    return np.einsum(",".join(ins) + "->" + letters[mode] + "r", *ops)


@pytest.mark.parametrize("nprtn", [1, 3, 16])
def test_mttkrp_variants_match_dense_definition(nprtn):
    dims, cidx, vals = rand_tensor([9, 700, 6, 5], 2000, seed=4, dup_frac=0.2)
    F = rand_factors(dims, 8, seed=5, lo=-1, hi=1)
    X = _dense(dims, cidx, vals)
    at = O.AltoTensor(dims, cidx, vals, nprtn=nprtn)
    seen = set()
    for mode in range(4):
        ref = _mttkrp_dense(X, F, mode)
        out, var = at.mttkrp(F, mode)
        seen.add(var)
        np.testing.assert_allclose(out, ref, rtol=1e-12, atol=1e-12)
    assert O.VAR_ATOMIC in seen and O.VAR_DA_MEM_PULL in seen  # mode 1: I > 512, reuse <= 4


def test_mttkrp_stream_variant_size_one_mode():
    dims, cidx, vals = rand_tensor([1, 40, 30], 500, seed=6)
    F = rand_factors(dims, 4, seed=7)
    at = O.AltoTensor(dims, cidx, vals, nprtn=4)
    assert at.variant(0, stream_mode=0) == O.VAR_STREAM
    out, var = at.mttkrp(F, 0, stream_mode=0)
    np.testing.assert_allclose(out, _mttkrp_dense(_dense(dims, cidx, vals), F, 0), rtol=1e-12)


def test_mttkrp_accumulates():
    dims, cidx, vals = rand_tensor([10, 12, 14], 400, seed=8)
    F = rand_factors(dims, 4, seed=9)
    at = O.AltoTensor(dims, cidx, vals, nprtn=2)
    base = np.full((12, 4), 2.5)
    out, _ = at.mttkrp(F, 1, out=base.copy())
    np.testing.assert_allclose(out - 2.5, _mttkrp_dense(_dense(dims, cidx, vals), F, 1), rtol=1e-12)


def test_mttkrp_coo_med_fixture():
    dims, cidx, vals = O.read_tns(golden("med.tns"))
    F = rand_factors(dims, 8, seed=10)
    at = O.AltoTensor(dims, cidx, vals, nprtn=8)
    for mode in range(3):
        ref = O.mttkrp_coo(cidx, vals, F, mode)
        out, _ = at.mttkrp(F, mode)
        np.testing.assert_allclose(out, ref, rtol=1e-12, atol=1e-9)


def test_zero_based_fixture_reads_like_reference():
    d1, c1, v1 = O.read_tns(golden("small4.tns"))
    d0, c0, v0 = O.read_tns(golden("small4_zeroidx.tns"), zero_based=True)
    assert d1 == [2, 2, 3, 3] and d0 == [3, 3, 4, 3]
    np.testing.assert_array_equal(v0, v1)


# ---------- dense step ----------

def test_gram_upper_triangle_only():
    U = np.random.default_rng(11).random((50, 6))
    G = O.gram_upper(U)
    full = U.T @ U
    iu = np.triu_indices(6)
    np.testing.assert_allclose(G[iu], full[iu], rtol=1e-13)
    assert np.all(G[np.tril_indices(6, -1)] == 0)


def test_kruskal_norms():
    U = np.random.default_rng(12).random((40, 5)) - 0.3
    U2, lam2 = O.kruskal_norm(U, 0)
    np.testing.assert_allclose(lam2, np.linalg.norm(U, axis=0), rtol=1e-13)
    np.testing.assert_allclose(np.linalg.norm(U2, axis=0), 1.0, rtol=1e-13)
    Um, lamm = O.kruskal_norm(U * 4, 1)
    np.testing.assert_allclose(lamm, np.maximum(1.0, (U * 4).max(axis=0)), rtol=0)
    Uneg = -np.abs(U)
    _, lneg = O.kruskal_norm(Uneg, 1)
    assert np.all(lneg == 1.0)  # signed maximum, floored at 1


def test_pseudo_inverse_solves_regularised_hadamard():
    rng = np.random.default_rng(13)
    R = 6
    facs = [rng.random((30, R)) for _ in range(3)]
    grams = [np.triu(f.T @ f) for f in facs]
    B = rng.random((20, R))
    X, g2, fb = O.pseudo_inverse_als(grams, 1, B)
    assert fb == 0
    V = (facs[0].T @ facs[0]) * (facs[2].T @ facs[2]) + 1e-3 * np.eye(R)
    np.testing.assert_allclose(X, np.linalg.solve(V, B.T).T, rtol=1e-9)


def test_pseudo_inverse_fallback_uses_diagonal_min_norm():
    R = 4
    # Grams as the reference holds them: upper triangle written, lower zero.
    g1 = np.eye(R) * 2.0 + np.triu(np.full((R, R), 0.5), 1)
    g2 = np.diag([3.0, -1.0, 1.0, 2.0]) + np.triu(np.full((R, R), 0.25), 1)
    grams = [np.full((R, R), np.nan), g1, g2]
    B = np.ones((3, R))
    X, _, fb = O.pseudo_inverse_als(grams, 0, B)
    assert fb == 1  # V + 1e-3 I is indefinite (pivot 1)
    # fallback symmetrises V from its (zero) lower triangle: a diagonal solve
    np.testing.assert_allclose(X, np.tile([1 / 6.0, -0.5, 0.5, 0.25], (3, 1)))


def test_cpd_fit_exact_is_relative_residual():
    dims, cidx, vals = rand_tensor([6, 5, 4], 60, seed=14, distinct=True)
    R = 3
    at = O.AltoTensor(dims, cidx, vals, nprtn=2)
    F = rand_factors(dims, R, seed=15)
    facs, lam, fits, fits_exact = at.cpd_als(F, R, 4, fit_mode=O.FIT_EXACT)
    X = _dense(dims, cidx, vals)
    M = np.einsum("ir,jr,kr,r->ijk", facs[0], facs[1], facs[2], lam)
    assert abs(fits_exact[-1] - np.linalg.norm(X - M) / np.linalg.norm(X)) < 1e-10
    # rank 1: the reference's diagonal-only norm equals the exact one
    F1 = rand_factors(dims, 1, seed=16)
    f1, l1, fd, fe = at.cpd_als(F1, 1, 3, fit_mode=O.FIT_REFERENCE_DIAG)
    np.testing.assert_allclose(fd, fe, rtol=1e-12)
    M1 = np.einsum("ir,jr,kr,r->ijk", f1[0], f1[1], f1[2], l1)
    assert abs(fd[-1] - np.linalg.norm(X - M1) / np.linalg.norm(X)) < 1e-10


@pytest.mark.parametrize("R", [4, 8])
def test_cpd_fit_reference_diag_is_sum_of_rank_one_norms(R):
    """At rank > 1 the reference's fit (cpd.hpp:228-268 reading Gram lower
    triangles that update_gram never writes, gram.cpp:35-41) keeps only the
    diagonal of lambda lambda^T (*) Had_m G_m, which is the sum over r of the
    squared norms of the dense rank-1 terms lambda_r u_1r o u_2r o u_3r; the
    exact convention is ||X - M|| / ||X||.  Non-orthogonal factors, so the
    two differ.  <X, M> and the norms come from dense numpy tensors."""
    dims, cidx, vals = rand_tensor([7, 6, 5], 90, seed=19, distinct=True)
    F = rand_factors(dims, R, seed=20)
    lam = np.random.default_rng(21).uniform(0.5, 2.0, R)
    X = _dense(dims, cidx, vals)
    terms = [lam[r] * np.einsum("i,j,k->ijk", F[0][:, r], F[1][:, r], F[2][:, r]) for r in range(R)]
    M = sum(terms)
    diag_norm = sum(float((t ** 2).sum()) for t in terms)
    inner = float((X * M).sum())
    nx = float((X ** 2).sum())
    mt = np.einsum("ijk,ir,jr->kr", X, F[0], F[1])  # MTTKRP of the last mode
    grams = [O.gram_upper(f) for f in F]            # upper triangles, zero below
    r = nx + diag_norm - 2 * inner
    want_diag = (np.sqrt(r) if r > 0 else r) / np.sqrt(nx)
    got_diag = O.cpd_fit(dims, F, lam, grams, mt, nx, O.FIT_REFERENCE_DIAG)
    assert abs(got_diag - want_diag) < 1e-10 * max(1.0, abs(want_diag))
    got_exact = O.cpd_fit(dims, F, lam, grams, mt, nx, O.FIT_EXACT)
    assert abs(got_exact - np.linalg.norm(X - M) / np.sqrt(nx)) < 1e-10
    assert abs(got_diag - got_exact) > 1e-3  # the conventions do differ here


def test_cpd_als_decreases_residual_and_stops_on_epsilon():
    dims, cidx, vals = rand_tensor([20, 15, 10], 800, seed=17)
    at = O.AltoTensor(dims, cidx, vals, nprtn=4)
    F = rand_factors(dims, 4, seed=18)
    _, _, fits, fe = at.cpd_als(F, 4, 30, eps=1e-30, fit_mode=O.FIT_EXACT)
    assert len(fits) == 30 and fe[-1] < fe[0]
    _, _, fits2, _ = at.cpd_als(F, 4, 30, eps=1e30)
    assert len(fits2) == 2  # stops at it = 1 (cpd.hpp:163)


# ---------- ADMM (admm.cpp:216-429) ----------

def test_admm_unconstrained_is_least_squares():
    rng = np.random.default_rng(19)
    R = 5
    facs = [rng.random((25, R)) for _ in range(3)]
    grams = [f.T @ f for f in facs]
    mt = rng.random((12, R))
    x, d, it = O.admm_cf(grams, 2, np.zeros((12, R)), np.zeros((12, R)), mt, nonneg=False)
    Phi = grams[0] * grams[1] + 1e-12 * np.eye(R)
    np.testing.assert_allclose(x, np.linalg.solve(Phi, mt.T).T, rtol=1e-9)


def test_admm_nonneg_approaches_nnls_solution():
    from scipy.optimize import nnls
    rng = np.random.default_rng(20)
    R = 4
    facs = [rng.random((30, R)) for _ in range(3)]
    grams = [f.T @ f for f in facs]
    Phi = grams[1] * grams[2]
    mt = rng.normal(size=(10, R)) * 5
    x = np.zeros((10, R))
    d = np.zeros((10, R))
    for _ in range(40):  # repeated calls keep the dual, as CP-stream's outer loop does
        x, d, it = O.admm_cf(grams, 0, x, d, mt, nonneg=True)
    assert np.all(x >= 0)
    Lc = np.linalg.cholesky(Phi)
    for i in range(10):
        # min 1/2 x^T Phi x - b^T x, x >= 0  <=>  NNLS on ||L^T x - L^-1 b||
        ref, _ = nnls(Lc.T, np.linalg.solve(Lc, mt[i]))
        np.testing.assert_allclose(x[i], ref, atol=2e-3 * (1 + np.abs(ref).max()))


# ---------- CP-stream (streaming_cpd.cpp) ----------

def test_cpstream_batching_and_invariants():
    rng = np.random.default_rng(21)
    T, I, J, nnz = 6, 30, 25, 1200
    t = np.sort(rng.integers(0, T, size=nnz)).astype(np.uint64)
    t[:T] = np.arange(T)  # every slice non-empty
    i = rng.integers(0, I, size=nnz).astype(np.uint64)
    j = rng.integers(0, J, size=nnz).astype(np.uint64)
    v = rng.poisson(3, size=nnz) + 1.0
    R = 4
    init = [rng.random((T, R)), rng.random((I, R)), rng.random((J, R))]
    res = O.cpstream([T, I, J], [t, i, j], v, 0, R, 5, 1e-5, True, init)
    assert res["slices"] == T
    # running max of first-appearance labels: final dims = number of distinct indices
    assert int(res["dims"][1]) == len(np.unique(i)) and int(res["dims"][2]) == len(np.unique(j))
    assert all(np.all(f >= 0) for f in res["factors"][1:])
    assert np.all((res["outer_iters"] >= 1) & (res["outer_iters"] <= 5))
    G = res["grams"]
    np.testing.assert_allclose(G[0], G[0].T, rtol=1e-12)


def _stream_case(seed, T=5, I=14, J=11, nnz=260, R=3, distinct=True):
    """A 3-mode stream (mode 0 = time), every slice non-empty; distinct
    coordinates unless asked otherwise."""
    rng = np.random.default_rng(seed)
    if distinct:
        lin = rng.choice(T * I * J, size=nnz, replace=False)
        t, i, j = np.unravel_index(lin, (T, I, J))
    else:
        t, i, j = rng.integers(0, T, nnz), rng.integers(0, I, nnz), rng.integers(0, J, nnz)
    t = t.copy()
    t[:T] = np.arange(T)
    cidx = [t.astype(np.uint64), i.astype(np.uint64), j.astype(np.uint64)]
    vals = rng.uniform(0.5, 2.0, nnz)
    init = [rng.random((T, R)), rng.random((I, R)), rng.random((J, R))]
    return [T, I, J], cidx, vals, init, R


def _np_cpstream_unconstrained(dims, cidx, vals, R, iters, init, batches=None):
    """Dense numpy restatement of CP-stream without constraints, stream mode 0
    (cpstream, streaming_cpd.cpp:26-366; cpstream_alto_iter 1198-1519): per
    batch of time slices (default: one slice each) the time rows
    S = MTTKRP_t (Had_{m!=t} G_m + 1e-12 I)^-1, G_t += S^T S, then per mode
    U_j = (MTTKRP_j + P_j (G_t,old Had_{m!=j,t} P_m^T U_m))
    (Had_{m!=j} G_m + 1e-12 I)^-1; G_t *= 0.99 once per batch.
    batches: list of (first time index, number of time indices)."""
    c2, v2 = O.stream_relabel(cidx, vals, 0)
    T = int(c2[0].max()) + 1
    if batches is None:
        batches = [(t, 1) for t in range(T)]
    U = [None] * 3
    P = [None] * 3
    G = [None] * 3
    prev = [1, 0, 0]
    trows = []
    eye = np.eye(R)
    for b, (t0, tb) in enumerate(batches):
        sel = (c2[0] >= t0) & (c2[0] < t0 + tb)
        bt = (c2[0][sel] - t0).astype(np.int64)
        bi, bj, bv = c2[1][sel].astype(np.int64), c2[2][sel].astype(np.int64), v2[sel]
        bd = [tb, max(int(bi.max()) + 1, prev[1]), max(int(bj.max()) + 1, prev[2])]
        prev = bd
        X = np.zeros((tb, bd[1], bd[2]))
        np.add.at(X, (bt, bi, bj), bv)
        if b == 0:
            U = [np.zeros((tb, R)), init[1][: bd[1]].copy(), init[2][: bd[2]].copy()]
            P = [np.zeros((tb, R)), np.zeros((bd[1], R)), np.zeros((bd[2], R))]
            for m in (1, 2):
                lam = np.sqrt(np.maximum((U[m] ** 2).sum(0), 1e-12))
                U[m] /= lam
            G = [np.zeros((R, R)), U[1].T @ U[1], U[2].T @ U[2]]
        else:
            for m in (1, 2):
                old = U[m].shape[0]
                if bd[m] > old:
                    U[m] = np.vstack([U[m], init[m][old: bd[m]]])
                    P[m] = np.vstack([P[m], np.zeros((bd[m] - old, R))])
                G[m] = U[m].T @ U[m]
        for _ in range(iters):
            S = np.einsum("tij,ir,jr->tr", X, U[1], U[2])
            S = np.linalg.solve(G[1] * G[2] + 1e-12 * eye, S.T).T
            U[0] = S
            Gold = G[0].copy()
            G[0] = G[0] + S.T @ S
            for j in (1, 2):
                k = 3 - j
                if j == 1:
                    mt = np.einsum("tij,tr,jr->ir", X, S, U[2])
                else:
                    mt = np.einsum("tij,tr,ir->jr", X, S, U[1])
                ata = Gold * (P[k].T @ U[k])
                rhs = mt + P[j] @ ata
                U[j] = np.linalg.solve((G[0] * G[k] + 1e-12 * eye).T, rhs.T).T
                G[j] = U[j].T @ U[j]
        G[0] *= 0.99
        P = [u.copy() for u in U]
        trows.extend(S.copy())
    return U, np.array(trows)


def _stream_counts_case(counts, seed, I=12, J=9, R=3):
    """A stream whose time slice t has exactly counts[t] nonzeros."""
    rng = np.random.default_rng(seed)
    t = np.repeat(np.arange(len(counts)), counts).astype(np.uint64)
    n = len(t)
    i = rng.integers(0, I, n).astype(np.uint64)
    j = rng.integers(0, J, n).astype(np.uint64)
    vals = rng.uniform(0.5, 2.0, n)
    init = [rng.random((len(counts), R)), rng.random((I, R)), rng.random((J, R))]
    return [len(counts), I, J], [t, i, j], vals, init, R


@pytest.mark.parametrize("thr,lim,expect", [
    (0, 1, [1, 1, 1, 1, 1]),      # the reference defaults: one time index per batch
    (10, 3, [3, 2]),              # 5+3 < 10 -> take the third (3 slices); 2+7 < 10, stream ends
    (6, 10, [2, 1, 2]),           # 5 < 6 -> +3; 8 >= 6 alone; 2 < 6 -> +7
    (0, 2, [2, 2, 1]),            # pairs
    (1, 5, [1, 1, 1, 1, 1]),      # any slice reaches a threshold of 1
])
def test_cpstream_dynamic_batches_known_answers(thr, lim, expect):
    """next_dynamic_batch (streaming_sptensor.cpp:196-264) on time slices of
    5, 3, 8, 2, 7 nonzeros: hand-derived batch sizes."""
    dims, cidx, vals, init, R = _stream_counts_case([5, 3, 8, 2, 7], 41)
    res = O.cpstream(dims, cidx, vals, 0, R, 2, 0.0, False, init, nnz_threshold=thr, timeslice_limit=lim)
    assert res["slices"] == len(expect)
    assert res["time_slices"] == 5 and res["time_rows"].shape == (5, R)
    sp = O.spcpstream(dims, cidx, vals, 0, R, 2, 0.0, init, nnz_threshold=thr, timeslice_limit=lim)
    assert sp["slices"] == len(expect) and sp["time_slices"] == 5


@pytest.mark.parametrize("thr,lim,batches", [(0, 2, [(0, 2), (2, 2), (4, 1)]), (6, 10, [(0, 2), (2, 1), (3, 2)])])
def test_cpstream_batched_oracle_matches_dense_restatement(thr, lim, batches):
    """Batches of several time indices: the oracle's unconstrained CP-stream
    equals the dense restatement over the same batches (a stream mode with
    one row per time index, one Gram update and one forgetting step per
    batch)."""
    dims, cidx, vals, init, R = _stream_counts_case([5, 3, 8, 2, 7], 42)
    iters = 3
    res = O.cpstream(dims, cidx, vals, 0, R, iters, 0.0, False, init, nnz_threshold=thr, timeslice_limit=lim)
    U, trows = _np_cpstream_unconstrained(dims, cidx, vals, R, iters, init, batches=batches)
    assert res["slices"] == len(batches)
    np.testing.assert_allclose(res["time_rows"], trows, rtol=1e-9, atol=1e-12)
    for m in (1, 2):
        np.testing.assert_allclose(res["factors"][m], U[m], rtol=1e-9, atol=1e-11)
    # spCP-stream reformulates the same update over batches as well
    sp = O.spcpstream(dims, cidx, vals, 0, R, iters, 0.0, init, nnz_threshold=thr, timeslice_limit=lim)
    np.testing.assert_allclose(sp["time_rows"], trows, rtol=1e-8, atol=1e-11)
    for m in (1, 2):
        np.testing.assert_allclose(sp["factors"][m], U[m], rtol=1e-8, atol=1e-10)


def test_cpstream_unconstrained_oracle_matches_dense_restatement():
    """The oracle's CP-stream without constraints (ALTO MTTKRP, Cholesky
    solves, upper-triangle Grams) equals a dense numpy restatement of the
    same updates for a fixed number of outer iterations (epsilon 0)."""
    dims, cidx, vals, init, R = _stream_case(31)
    iters = 3
    res = O.cpstream(dims, cidx, vals, 0, R, iters, 0.0, False, init)
    U, trows = _np_cpstream_unconstrained(dims, cidx, vals, R, iters, init)
    assert res["slices"] == dims[0] and np.all(res["outer_iters"] == iters)
    np.testing.assert_allclose(res["time_rows"], trows, rtol=1e-9, atol=1e-12)
    for m in (1, 2):
        np.testing.assert_allclose(res["factors"][m], U[m], rtol=1e-9, atol=1e-11)


@pytest.mark.parametrize("seed,distinct", [(32, True), (33, False)])
def test_spcpstream_oracle_equals_unconstrained_cpstream(seed, distinct):
    """spCP-stream is CP-stream without constraints with the rows absent from
    a slice updated implicitly: c = c_nz + c_z and h = h_nz + h_z track
    U^T U and prevU^T U exactly, and the z rows end as prevU Q Phi^-1, the
    value the full update gives them.  With epsilon 0 both run every outer
    iteration, so factors, time rows and the temporal Gram agree to rounding."""
    dims, cidx, vals, init, R = _stream_case(seed, T=6, I=20, J=15, nnz=300, R=4, distinct=distinct)
    iters = 4
    sp = O.spcpstream(dims, cidx, vals, 0, R, iters, 0.0, init)
    cs = O.cpstream(dims, cidx, vals, 0, R, iters, 0.0, False, init)
    assert sp["slices"] == cs["slices"] == dims[0]
    assert np.all(sp["outer_iters"] == iters)
    np.testing.assert_array_equal(sp["dims"], cs["dims"])
    np.testing.assert_allclose(sp["time_rows"], cs["time_rows"], rtol=1e-8, atol=1e-11)
    for m in (1, 2):
        np.testing.assert_allclose(sp["factors"][m], cs["factors"][m], rtol=1e-8, atol=1e-10)
    np.testing.assert_allclose(sp["grams"][0], cs["grams"][0], rtol=1e-8, atol=1e-10)


def test_spcpstream_oracle_rejects_zero_iterations():
    dims, cidx, vals, init, R = _stream_case(34)
    with pytest.raises(ValueError):
        O.spcpstream(dims, cidx, vals, 0, R, 0, 0.0, init)


def _rowsparse_case(dims, nnz, R, seed, stream_mode=-1, absent=0.0):
    rng = np.random.default_rng(seed)
    cidx = [rng.integers(0, d, size=nnz).astype(np.uint64) for d in dims]
    vals = rng.uniform(-1, 1, nnz)
    rsp, ridx = [], []
    for m, d in enumerate(dims):
        perm = rng.permutation(d).astype(np.int32)          # stored-row order differs from index order
        keep = rng.random(d) >= absent
        r = np.where(keep, perm, -1).astype(np.int32)
        ridx.append(r)
        rsp.append(rng.uniform(-1, 1, (d if m != stream_mode else 1, R)))
    return cidx, vals, rsp, ridx


@pytest.mark.parametrize("stream_mode", [-1, 0])
def test_rowsparse_mttkrp_oracle_is_dense_mttkrp_restricted(stream_mode):
    """rowsparse_mttkrp == the dense COO MTTKRP (mttkrp.cpp:122-156, pinned
    above) over the expanded factors, restricted to the occurring rows."""
    dims, R = [6, 9, 7, 5], 4
    cidx, vals, rsp, ridx = _rowsparse_case(dims, 300, R, 77, stream_mode, absent=0.2)
    for mode in (1, 3):
        rowind, rows = O.rowsparse_mttkrp(cidx, vals, mode, stream_mode, rsp, ridx, R)
        dense = []
        for m, d in enumerate(dims):
            if m == stream_mode:
                dense.append(np.repeat(rsp[m][:1], d, axis=0))
            else:
                D = np.zeros((d, R))
                ok = ridx[m] >= 0
                D[ok] = rsp[m][ridx[m][ok]]
                dense.append(D)
        full = O.mttkrp_coo(cidx, vals, dense, mode)
        np.testing.assert_array_equal(rowind, np.unique(cidx[mode]))
        np.testing.assert_allclose(rows, full[rowind.astype(np.int64)], rtol=1e-13, atol=1e-14)


def test_rowsparse_mttkrp_oracle_brute_force_tiny():
    """One output row by hand: sum over its nonzeros of val * product of the
    stored rows the map selects."""
    cidx = [np.array([0, 2, 0], np.uint64), np.array([1, 0, 1], np.uint64), np.array([0, 0, 1], np.uint64)]
    vals = np.array([2.0, 3.0, 5.0])
    rsp = [None, np.array([[1.0, 2.0], [3.0, 4.0]]), np.array([[0.5, 1.0], [2.0, -1.0]])]
    ridx = [None, np.array([1, 0], np.int32), np.array([0, 1], np.int32)]
    rowind, rows = O.rowsparse_mttkrp(cidx, vals, 0, -1, rsp, ridx, 2)
    np.testing.assert_array_equal(rowind, [0, 2])
    # row 0: nonzeros 0 and 2; mode-1 index 1 -> stored row 0 = [1, 2]
    exp0 = 2.0 * np.array([1.0, 2.0]) * np.array([0.5, 1.0]) + 5.0 * np.array([1.0, 2.0]) * np.array([2.0, -1.0])
    exp2 = 3.0 * np.array([3.0, 4.0]) * np.array([0.5, 1.0])
    np.testing.assert_allclose(rows, [exp0, exp2], rtol=0, atol=0)


def test_cpd_error_oracle_is_dense_relative_residual():
    """cpd_error's closed form equals ||X - M|| / ||X|| computed densely
    (weights 1), for a tensor with distinct coordinates."""
    rng = np.random.default_rng(81)
    dims, R = [5, 4, 6], 3
    lin = rng.choice(np.prod(dims), size=40, replace=False)
    cidx = list(np.unravel_index(lin, dims))
    cidx = [c.astype(np.uint64) for c in cidx]
    vals = rng.uniform(-1, 2, 40)
    F = [rng.uniform(-1, 1, (d, R)) for d in dims]
    err, (xx, zz, inner) = O.cpd_error(cidx, vals, F)
    X = np.zeros(dims)
    X[tuple(c.astype(np.int64) for c in cidx)] = vals
    M = np.einsum("ir,jr,kr->ijk", *F)
    assert abs(err - np.linalg.norm(X - M) / np.linalg.norm(X)) < 1e-12
    assert abs(zz - np.linalg.norm(M) ** 2) < 1e-10 * zz
    assert abs(inner - np.sum(X * M)) < 1e-12 * max(1.0, abs(inner))


def test_stream_relabel_is_sorted_first_appearance_bijection():
    rng = np.random.default_rng(82)
    dims = [7, 50, 40]
    cidx = [rng.integers(0, d, 400).astype(np.uint64) for d in dims]
    vals = rng.random(400)
    rc, rv = O.stream_relabel(cidx, vals, 0)
    assert np.all(np.diff(rc[0].astype(np.int64)) >= 0)
    for m in (1, 2):
        seen = []
        for x in rc[m]:
            if int(x) not in seen:
                seen.append(int(x))
        assert seen == list(range(len(seen)))          # labels in order of first appearance
    # the relabelling is a bijection per mode: the multiset of (coords, value) is kept
    order = np.lexsort([cidx[2], cidx[1], cidx[0]])
    np.testing.assert_array_equal(rv, vals[order])
    for m in (1, 2):
        pairs = set(zip(cidx[m][order].tolist(), rc[m].tolist()))
        assert len({a for a, _ in pairs}) == len(pairs) == len({b for _, b in pairs})
