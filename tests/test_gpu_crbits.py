"""GPU conflict-resolution bits (GENERATE_CR_BITS, alto.hpp:1250-1346;
alto_hip_tensor_cr_bits) against the oracle, through the C ABI.

Bars: masks and keys bit-exact in both layouts (the ALTO keys against the
oracle's, the ALT_PEXT keys against the oracle's re-encoding plus the same
bits); the MTTKRP reads only the mode fields, so its FP32-output result is
identical bit for bit with and without the bits and the FP64 one stays
within 1e-12 of the |X|-MTTKRP bound; a rebuild (update_salto) clears them."""
import numpy as np
import pytest

import oracle as O
from conftest import rand_factors, rand_tensor

pytestmark = pytest.mark.gpu


def _case(dims, nnz, seed, hot_rows=None):
    dims, cidx, vals = rand_tensor(dims, nnz, seed=seed)
    if hot_rows:
        rng = np.random.default_rng(seed + 1)
        cidx[0] = rng.integers(0, hot_rows, size=nnz).astype(np.uint64)
    return dims, cidx, vals


CASES = [
    ("4mode_34bits", [2000, 700, 600, 5], 3000, 40, 120),
    ("63bits_one_free", [1 << 21, 1 << 21, 1 << 21], 400, 42, 30),
    ("70bits_two_words", [1 << 30, 1 << 30, 1000], 500, 43, 40),
    ("200k_two_modes", [100000, 80000, 600], 200000, 44, None),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("nprtn", [1, 2, 7, 64])
def test_cr_bits_match_oracle(gpu, case, nprtn):
    name, dims, nnz, seed, hot = case
    print("case", name, "nprtn", nprtn, flush=True)
    dims, cidx, vals = _case(dims, nnz, seed, hot)
    ref = O.AltoTensor(dims, cidx, vals, nprtn=nprtn)
    want_keys, want_cm = O.cr_bits(dims, ref.keys, ref.masks, nprtn, ref.prtn_ptr)
    at = gpu.create_alto(dims, cidx, vals, nprtn=nprtn)
    cm = at.cr_bits()
    assert np.array_equal(cm, want_cm)
    assert np.array_equal(at.keys(), want_keys)
    packed, _ = O.alt_pext(dims, ref.keys, ref.masks)
    got_packed, _ = at.packed_keys()
    assert np.array_equal(got_packed, packed | (want_keys ^ ref.keys))
    assert np.array_equal(at.cr_bits(), cm)  # a second call changes nothing
    assert np.array_equal(at.keys(), want_keys)
    at.close()


def test_cr_bits_beyond_packed_keys_are_refused(gpu):
    # 63 bits in forced two-word keys: free bits 65, mode 1's bit (64) lies
    # beyond the one-word packed keys
    dims, cidx, vals = _case([1 << 21, 1 << 21, 1 << 21], 300, 45)
    at = gpu.create_alto(dims, cidx, vals, nprtn=3, key_words=2)
    before = at.keys()
    with pytest.raises(gpu.AltoError, match="unsupported"):
        at.cr_bits()
    assert np.array_equal(at.keys(), before)
    at.close()


def test_cr_bits_mode_too_long_is_refused(gpu):
    # a qualifying mode of 2^33 rows: its owner array would need an entry per row
    dims, cidx, vals = _case([600, 700, 4], 200, 47)
    dims = [1 << 33, 700, 4]
    at = gpu.create_alto(dims, cidx, vals, nprtn=2)
    with pytest.raises(gpu.AltoError, match="unsupported"):
        at.cr_bits()
    at.close()


@pytest.mark.parametrize("rank", [16, 32])
def test_mttkrp_ignores_cr_bits(gpu, rank):
    dims, cidx, vals = _case([100000, 80000, 600], 200000, 44)
    F = rand_factors(dims, rank, seed=46, lo=-1, hi=1)
    a0 = gpu.create_alto(dims, cidx, vals, nprtn=16)
    a1 = gpu.create_alto(dims, cidx, vals, nprtn=16)
    cm = a1.cr_bits()
    assert cm[0].any() and cm[1].any() and not cm[2].any()
    ref = O.AltoTensor(dims, cidx, vals, nprtn=16)
    bnd_t = O.AltoTensor(dims, cidx, np.abs(vals), nprtn=16)
    d32 = [gpu.to_prec(f, gpu.PREC_F32) for f in F]
    d64 = [gpu.to_prec(f, gpu.PREC_F64) for f in F]
    for mode in range(3):
        I = dims[mode]
        outs = []
        for at in (a0, a1):
            o = gpu.DeviceArray.zeros((I, rank), np.float32)
            gpu.mttkrp_alto_par(mode, d32, rank, at, o, prec=gpu.PREC_F32, out_f64=False)
            outs.append(o.numpy())
        assert np.array_equal(outs[0], outs[1]), mode
        for algo in (1, 2):
            o = gpu.DeviceArray.zeros((I, rank), np.float64)
            gpu.mttkrp_alto_par(mode, d64, rank, a1, o, prec=gpu.PREC_F64, algo=algo)
            exp = ref.mttkrp(F, mode)[0]
            bnd = bnd_t.mttkrp([np.abs(f) for f in F], mode)[0]
            assert np.all(np.abs(o.numpy() - exp) <= 1e-12 * bnd + 1e-300), (mode, algo)
    a0.close()
    a1.close()


def test_update_clears_cr_bits(gpu):
    dims, cidx, vals = _case([2000, 700, 600, 5], 3000, 40, 120)
    at = gpu.create_alto(dims, cidx, vals, nprtn=7)
    assert at.cr_bits().any()
    ref = O.AltoTensor(dims, cidx, vals, nprtn=7)
    at.update(dims, cidx, vals, nprtn=7)
    assert np.array_equal(at.keys(), ref.keys)
    at.close()
