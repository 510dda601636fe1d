"""GPU parity of the ALTO encoding, sort, partitions and MTTKRP against the
oracle, through the C ABI (libalto_amd.so).

Bars: keys, order (permutation), values and partitions bit-exact; FP64
MTTKRP within 1e-12 of the |X|-MTTKRP bound (MTTKRP of |X| with |factors|,
per element); FP32 within 2e-6 of that bound; BF16-stored factors compared
with the oracle fed the same BF16-rounded factors, within 2e-6 of the bound.
"""
import numpy as np
import pytest

import oracle as O
from conftest import golden, rand_factors, rand_tensor

pytestmark = pytest.mark.gpu

F64_TOL = 1e-12
F32_TOL = 2e-6


def bf16_round(a):
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def _cases():
    out = []
    for name in ["small.tns", "small4.tns", "med.tns", "med4.tns", "med5.tns"]:
        out.append(("fixture:" + name,) + tuple(O.read_tns(golden(name))))
    out.append(("zero_based:small4",) + tuple(O.read_tns(golden("small4_zeroidx.tns"), zero_based=True)))
    out.append(("dups3",) + rand_tensor([37, 1000, 5], 20000, seed=31, dup_frac=0.4))
    out.append(("size1",) + rand_tensor([1, 3000, 700], 5000, seed=32))
    out.append(("one_nnz",) + rand_tensor([4, 5, 6], 1, seed=33))
    out.append(("u128",) + rand_tensor([1 << 20, 1 << 20, 1 << 18, 1 << 16, 1 << 12], 30000, seed=34, dup_frac=0.1))
    out.append(("big_tiles",) + rand_tensor([3000, 2000, 1000], 300000, seed=35, dup_frac=0.05))
    out.append(("hot_dups",) + _hot_dups())
    return out


def _hot_dups():
    """Runs of equal coordinates around the 64-entry short-run limit of the
    schedule merge (63, 64, 65) and one of 5000, mixed with random entries."""
    dims, cidx, vals = rand_tensor([50, 40, 30], 3000, seed=38)
    rng = np.random.default_rng(39)
    extra = [((1, 2, 3), 5000), ((4, 5, 6), 65), ((7, 8, 9), 64), ((10, 11, 12), 63)]
    add_c = [np.concatenate([np.full(k, c[m], np.uint64) for c, k in extra]) for m in range(3)]
    add_v = rng.uniform(-1, 1, sum(k for _, k in extra))
    perm = rng.permutation(len(vals) + len(add_v))
    cidx = [np.concatenate([cidx[m], add_c[m]])[perm] for m in range(3)]
    vals = np.concatenate([vals, add_v])[perm]
    return dims, cidx, vals


CASES = _cases()


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("nprtn", [1, 7, 64])
def test_encode_sort_partition_bit_exact(gpu, case, nprtn):
    name, dims, cidx, vals = case
    print("case", name, "nprtn", nprtn, flush=True)
    ref = O.AltoTensor(dims, cidx, vals, nprtn=nprtn)
    at = gpu.create_alto(dims, cidx, vals, nprtn=nprtn)
    assert at.kw == ref.kw and at.total_bits == ref.bits
    np.testing.assert_array_equal(at.masks(), ref.masks)
    np.testing.assert_array_equal(at.keys(), ref.keys)
    np.testing.assert_array_equal(at.perm().astype(np.uint64), ref.perm)
    np.testing.assert_array_equal(at.vals(), ref.vals)
    ptr, iv = at.partitions()
    np.testing.assert_array_equal(ptr, ref.prtn_ptr)
    np.testing.assert_array_equal(iv, ref.intervals)


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_alt_pext_layout_bit_exact(gpu, case):
    """The packed keys the MTTKRP reads are the reference's ALT_PEXT
    re-encoding (alto.hpp:260-292) of the sorted keys."""
    name, dims, cidx, vals = case
    ref = O.AltoTensor(dims, cidx, vals, nprtn=4)
    at = gpu.create_alto(dims, cidx, vals, nprtn=4)
    packed, pos = at.packed_keys()
    exp, epos = O.alt_pext(dims, ref.keys, ref.masks)
    np.testing.assert_array_equal(pos, epos)
    np.testing.assert_array_equal(packed, exp)


def test_forced_128bit_key_matches_oracle(gpu):
    dims, cidx, vals = rand_tensor([1000, 800, 600], 5000, seed=36, dup_frac=0.1)
    ref = O.AltoTensor(dims, cidx, vals, nprtn=4, kw=2)
    at = gpu.create_alto(dims, cidx, vals, nprtn=4, key_words=2)
    assert at.kw == 2
    np.testing.assert_array_equal(at.keys(), ref.keys)
    np.testing.assert_array_equal(at.vals(), ref.vals)


def test_empty_tensor(gpu):
    """No nonzeros: masks as for any tensor of these dims, no keys, every
    partition empty ([0, 0] intervals, alto.hpp:1220-1225), the MTTKRP adds
    nothing (out += 0) in every mode and precision, and the conflict bits'
    masks follow the mode rule with no key to mark."""
    dims = [600, 700, 8]
    cidx = [np.zeros(0, np.uint64) for _ in dims]
    vals = np.zeros(0)
    ref = O.AltoTensor(dims, cidx, vals, nprtn=4)
    at = gpu.create_alto(dims, cidx, vals, nprtn=4)
    assert at.nnz == 0
    np.testing.assert_array_equal(at.masks(), ref.masks)
    assert at.keys().shape == (0, at.kw)
    ptr, iv = at.partitions()
    np.testing.assert_array_equal(ptr, ref.prtn_ptr)
    np.testing.assert_array_equal(iv, ref.intervals)
    assert not ptr.any() and not iv.any()
    R = 16
    F = rand_factors(dims, R, seed=7)
    for prec in (gpu.PREC_F64, gpu.PREC_F32):
        devF = [gpu.to_prec(f, prec) for f in F]
        for algo in (1, 2):
            for mode in range(3):
                dt = np.float64 if prec == gpu.PREC_F64 else np.float32
                out = gpu.DeviceArray.from_numpy(np.ones((dims[mode], R), dt))
                gpu.mttkrp_alto_par(mode, devF, R, at, out, prec=prec, out_f64=prec == gpu.PREC_F64, algo=algo)
                assert np.array_equal(out.numpy(), np.ones((dims[mode], R))), (prec, algo, mode)
    _, want = O.cr_bits(dims, ref.keys, ref.masks, 4, ref.prtn_ptr)
    np.testing.assert_array_equal(at.cr_bits(), want)
    at.close()


def test_errors_are_status_codes(gpu):
    dims, cidx, vals = rand_tensor([10, 10], 50, seed=37)
    bad = [c.copy() for c in cidx]
    bad[1][3] = 10  # not below its mode size
    with pytest.raises(gpu.AltoError, match="invalid argument"):
        gpu.create_alto(dims, bad, vals)
    with pytest.raises(gpu.AltoError, match="more bits"):
        gpu.create_alto([1 << 40, 1 << 40], [np.zeros(1, np.uint64)] * 2, np.ones(1), key_words=1)


def test_c_abi_refuses_a_short_output_buffer(gpu):
    """alto_hip_mttkrp takes the output's size in bytes and refuses a buffer
    smaller than the rows the kernel writes (FP64 rows for out_f64, FP32
    rows otherwise), before anything is launched: the output keeps its bytes."""
    import ctypes
    dims, cidx, vals = rand_tensor([40, 30, 20], 800, seed=38)
    at = gpu.create_alto(dims, cidx, vals)
    R = 8
    F = rand_factors(dims, R, seed=39)
    devF = [gpu.to_prec(f, gpu.PREC_F32) for f in F]
    out = gpu.DeviceArray.from_numpy(np.full((dims[0], R), 7.0, np.float64))
    L = gpu.lib()
    ptrs = (ctypes.c_void_p * 3)(*[f.ptr for f in devF])
    call = lambda out_f64, nbytes: L.alto_hip_mttkrp(ctypes.c_void_p(at.handle), 0, gpu.PREC_F32, R, ptrs,
                                                     ctypes.c_void_p(out.ptr), out_f64, ctypes.c_size_t(nbytes),
                                                     None)
    assert call(1, dims[0] * R * 8 - 1) == -1  # ALTO_ERR_ARG: FP64 rows need 8 bytes each
    assert call(0, dims[0] * R * 4 - 4) == -1  # FP32 rows, one element short
    gpu.synchronize()
    assert np.array_equal(out.numpy(), np.full((dims[0], R), 7.0))
    assert call(0, dims[0] * R * 4) == 0  # exactly the FP32 rows: accepted
    gpu.synchronize()
    at.close()


class _Ref:
    def __init__(self, dims, cidx, vals):
        self.t = O.AltoTensor(dims, cidx, vals, nprtn=8)
        self.tabs = O.AltoTensor(dims, cidx, np.abs(vals), nprtn=8)

    def mttkrp(self, F, mode):
        return self.t.mttkrp(F, mode)[0]

    def bound(self, F, mode):
        return self.tabs.mttkrp([np.abs(f) for f in F], mode)[0]


def _run_gpu_mttkrp(gpu, at, F, mode, prec, rank, algo, out_f64=True):
    devF = [gpu.to_prec(f, prec) for f in F]
    I = int(at.dims[mode])
    out = gpu.DeviceArray.zeros((I, rank), np.float64 if (out_f64 or prec == gpu.PREC_F64) else np.float32)
    gpu.mttkrp_alto_par(mode, devF, rank, at, out, prec=prec, out_f64=out_f64, algo=algo)
    return out.numpy().astype(np.float64)


MT_CASES = [c for c in CASES if c[0] in ("fixture:med.tns", "fixture:med4.tns", "fixture:med5.tns", "dups3",
                                         "size1", "u128", "hot_dups")]


@pytest.mark.parametrize("case", MT_CASES, ids=[c[0] for c in MT_CASES])
@pytest.mark.parametrize("rank", [8, 16, 32, 64, 10])
@pytest.mark.parametrize("algo", [1, 2])
def test_mttkrp_f64_parity(gpu, case, rank, algo):
    name, dims, cidx, vals = case
    print("case", name, "rank", rank, "algo", algo, flush=True)
    ref = _Ref(dims, cidx, vals)
    F = rand_factors(dims, rank, seed=40 + rank, lo=-1, hi=1)
    at = gpu.create_alto(dims, cidx, vals, nprtn=8)
    for mode in range(len(dims)):
        got = _run_gpu_mttkrp(gpu, at, F, mode, gpu.PREC_F64, rank, algo)
        exp = ref.mttkrp(F, mode)
        bnd = ref.bound(F, mode)
        assert np.all(np.abs(got - exp) <= F64_TOL * bnd + 1e-300), (name, mode)


@pytest.mark.parametrize("case", MT_CASES, ids=[c[0] for c in MT_CASES])
@pytest.mark.parametrize("rank", [16, 32, 64])
@pytest.mark.parametrize("out_f64", [True, False])
def test_mttkrp_f32_parity(gpu, case, rank, out_f64):
    name, dims, cidx, vals = case
    print("case", name, "rank", rank, "out_f64", out_f64, flush=True)
    ref = _Ref(dims, cidx, vals)
    F = rand_factors(dims, rank, seed=50 + rank, lo=-1, hi=1)
    F32 = [f.astype(np.float32).astype(np.float64) for f in F]
    at = gpu.create_alto(dims, cidx, vals, nprtn=8)
    tol = F32_TOL if out_f64 else 2e-5
    for algo in (1, 2):
        for mode in range(len(dims)):
            got = _run_gpu_mttkrp(gpu, at, F32, mode, gpu.PREC_F32, rank, algo, out_f64)
            exp = ref.mttkrp(F32, mode)
            bnd = ref.bound(F32, mode)
            assert np.all(np.abs(got - exp) <= tol * bnd + 1e-30), (name, mode, algo)


@pytest.mark.parametrize("case", MT_CASES, ids=[c[0] for c in MT_CASES])
@pytest.mark.parametrize("rank", [8, 32, 64])
def test_mttkrp_bf16_storage_parity(gpu, case, rank):
    name, dims, cidx, vals = case
    ref = _Ref(dims, cidx, vals)
    F = rand_factors(dims, rank, seed=60 + rank, lo=-1, hi=1)
    Fb = [bf16_round(f) for f in F]
    at = gpu.create_alto(dims, cidx, vals, nprtn=8)
    for mode in range(len(dims)):
        got = _run_gpu_mttkrp(gpu, at, F, mode, gpu.PREC_BF16, rank, 2, True)
        exp = ref.mttkrp(Fb, mode)
        bnd = ref.bound(Fb, mode)
        assert np.all(np.abs(got - exp) <= F32_TOL * bnd + 1e-30), (name, mode)
        # storage error against unrounded FP64 factors: bounded by BF16 rounding
        exact = ref.mttkrp(F, mode)
        assert np.all(np.abs(got - exact) <= (len(dims) * 2.0 ** -8) * ref.bound(F, mode) + 1e-30)


def test_mttkrp_accumulates_into_output(gpu):
    dims, cidx, vals = rand_tensor([40, 30, 20], 3000, seed=70)
    ref = _Ref(dims, cidx, vals)
    F = rand_factors(dims, 16, seed=71)
    at = gpu.create_alto(dims, cidx, vals)
    devF = [gpu.to_prec(f, gpu.PREC_F64) for f in F]
    out = gpu.DeviceArray.from_numpy(np.full((30, 16), 3.0))
    gpu.mttkrp_alto_par(1, devF, 16, at, out)
    got = out.numpy()
    np.testing.assert_allclose(got - 3.0, ref.mttkrp(F, 1), rtol=1e-11, atol=1e-12)


def test_coo_mttkrp_on_device(gpu):
    dims, cidx, vals = O.read_tns(golden("med4.tns"))
    F = rand_factors(dims, 16, seed=72, lo=-1, hi=1)
    dc = [gpu.DeviceArray.from_numpy(np.ascontiguousarray(c, dtype=np.uint64)) for c in cidx]
    dv = gpu.DeviceArray.from_numpy(vals)
    devF = [gpu.to_prec(f, gpu.PREC_F64) for f in F]
    ref = _Ref(dims, cidx, vals)
    for mode in range(4):
        out = gpu.DeviceArray.zeros((dims[mode], 16), np.float64)
        gpu.mttkrp_coo(dims, dc, dv, mode, 16, devF, out)
        got = out.numpy()
        exp = O.mttkrp_coo(cidx, vals, F, mode)
        assert np.all(np.abs(got - exp) <= F64_TOL * ref.bound(F, mode) + 1e-300)


def test_generator_is_deterministic_and_in_range(gpu):
    dims = [5000, 2000, 300]
    c1, v1 = gpu.generate(dims, 100000, 7, [1, 0, 1], [1.1, 1.0, 1.05], value_dist=1, p1=0.0, p2=1.0)
    c2, v2 = gpu.generate(dims, 100000, 7, [1, 0, 1], [1.1, 1.0, 1.05], value_dist=1, p1=0.0, p2=1.0)
    for a, b, d in zip(c1, c2, dims):
        x = a.numpy()
        np.testing.assert_array_equal(x, b.numpy())
        assert x.max() < d
    v = v1.numpy()
    np.testing.assert_array_equal(v, v2.numpy())
    assert np.all(v > 0) and abs(np.median(np.log(v))) < 0.02
    # Zipf(1.1) over 5000: the most popular index holds about 1/H of the draws
    x0 = c1[0].numpy()
    top = np.bincount(x0.astype(np.int64)).max() / len(x0)
    H = np.sum(np.arange(1, 5001, dtype=np.float64) ** -1.1)
    assert abs(top - 1 / H) < 0.01
