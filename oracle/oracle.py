"""ctypes binding of the CPU oracle (oracle/alto_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker or the timed CPU port,
never as the product path.

Arrays are numpy; keys are uint64 arrays of shape (nnz, kw).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")

VAR_ATOMIC, VAR_STREAM, VAR_DA_MEM_PULL = 0, 1, 2
FIT_EXACT, FIT_REFERENCE_DIAG = 0, 1

_u64p = ctypes.POINTER(ctypes.c_uint64)
_f64p = ctypes.POINTER(ctypes.c_double)
_lib = None


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        L = _lib
        L.orc_num_bits.restype = ctypes.c_int
        L.orc_num_bits.argtypes = [ctypes.c_uint64]
        L.orc_alto_masks.restype = ctypes.c_int
        L.orc_pext.restype = ctypes.c_uint64
        L.orc_get_num_ptrn.restype = ctypes.c_int
        L.orc_get_num_ptrn.argtypes = [ctypes.c_uint64, ctypes.c_int]
        L.orc_mttkrp_variant.restype = ctypes.c_int
        L.orc_mttkrp_alto.restype = ctypes.c_int
        L.orc_pseudo_inverse_als.restype = ctypes.c_int
        L.orc_cpd_fit.restype = ctypes.c_double
        L.orc_cpd_als.restype = ctypes.c_int
        L.orc_admm_cf.restype = ctypes.c_double
        L.orc_cpstream.restype = ctypes.c_int
        L.orc_get_threads.restype = ctypes.c_int
        L.orc_cr_bits.restype = ctypes.c_int
    return _lib


def _p64(a):
    return a.ctypes.data_as(_u64p)


def _pf(a):
    return a.ctypes.data_as(_f64p)


def _ptr_array(arrs, kind):
    T = _u64p if kind == "u" else _f64p
    arr = (T * len(arrs))()
    for i, a in enumerate(arrs):
        arr[i] = a.ctypes.data_as(T) if a is not None else None
    return arr


def set_threads(n):
    lib().orc_set_threads(ctypes.c_int(int(n)))


def num_bits(dim):
    return lib().orc_num_bits(ctypes.c_uint64(int(dim)))


def key_words(dims):
    total = sum(num_bits(d) for d in dims)
    return 1 if total <= 64 else 2


def alto_masks(dims, kw=None):
    dims = np.ascontiguousarray(dims, dtype=np.uint64)
    if kw is None:
        kw = key_words(dims)
    masks = np.zeros((len(dims), kw), dtype=np.uint64)
    total = lib().orc_alto_masks(ctypes.c_int(len(dims)), _p64(dims), ctypes.c_int(kw), _p64(masks))
    if total < 0:
        raise ValueError("ALTO key needs more than %d bits" % (64 * kw))
    return masks, total


def pdep(x, mask):
    mask = np.ascontiguousarray(mask, dtype=np.uint64)
    out = np.zeros(len(mask), dtype=np.uint64)
    lib().orc_pdep(ctypes.c_uint64(int(x)), _p64(mask), ctypes.c_int(len(mask)), _p64(out))
    return out


def pext(key, mask):
    key = np.ascontiguousarray(key, dtype=np.uint64)
    mask = np.ascontiguousarray(mask, dtype=np.uint64)
    return int(lib().orc_pext(_p64(key), _p64(mask), ctypes.c_int(len(mask))))


def linearize(cidx, masks):
    cidx = [np.ascontiguousarray(c, dtype=np.uint64) for c in cidx]
    nnz = len(cidx[0])
    kw = masks.shape[1]
    keys = np.zeros((nnz, kw), dtype=np.uint64)
    lib().orc_linearize(ctypes.c_int(len(cidx)), ctypes.c_uint64(nnz), _ptr_array(cidx, "u"),
                        _p64(np.ascontiguousarray(masks)), ctypes.c_int(kw), _p64(keys))
    return keys


def sort_alto(keys, vals):
    keys = np.ascontiguousarray(keys, dtype=np.uint64).copy()
    vals = np.ascontiguousarray(vals, dtype=np.float64).copy()
    nnz, kw = keys.shape
    perm = np.zeros(nnz, dtype=np.uint64)
    lib().orc_sort_alto(ctypes.c_uint64(nnz), ctypes.c_int(kw), _p64(keys), _pf(vals), _p64(perm))
    return keys, vals, perm


def alt_pext(dims, keys, masks):
    """ALT_PEXT re-encoding (alto.hpp:260-292, mode_pos alto.hpp:982-988):
    each sorted key's mode fields pext'ed out and packed at pos[m] = prefix
    sum of num_bits over the modes in mode order.  keys: (nnz, kw) words.
    Returns (packed keys (nnz, kw), pos)."""
    keys = np.asarray(keys, dtype=np.uint64).reshape(len(keys), -1)
    masks = np.asarray(masks, dtype=np.uint64).reshape(len(dims), -1)
    kw = keys.shape[1]
    bits = [num_bits(d) for d in dims]
    pos = np.concatenate([[0], np.cumsum(bits)[:-1]]).astype(int)
    out = np.zeros_like(keys)
    one = np.uint64(1)
    for m in range(len(dims)):
        j = 0  # next field bit: mask bits in ascending order (pext)
        for w in range(kw):
            mw = int(masks[m, w])
            for b in range(64):
                if (mw >> b) & 1:
                    bit = (keys[:, w] >> np.uint64(b)) & one
                    dst = int(pos[m]) + j
                    out[:, dst // 64] |= bit << np.uint64(dst % 64)
                    j += 1
    return out, pos


def prtn_alto(dims, keys, masks, nprtn):
    dims = np.ascontiguousarray(dims, dtype=np.uint64)
    keys = np.ascontiguousarray(keys, dtype=np.uint64)
    nnz, kw = keys.shape
    ptr = np.zeros(nprtn + 1, dtype=np.uint64)
    iv = np.zeros((nprtn, len(dims), 2), dtype=np.uint64)
    lib().orc_prtn_alto(ctypes.c_int(len(dims)), _p64(dims), ctypes.c_uint64(nnz), ctypes.c_int(kw),
                        _p64(keys), _p64(np.ascontiguousarray(masks)), ctypes.c_int(nprtn),
                        _p64(ptr), _p64(iv))
    return ptr, iv


def cr_bits(dims, keys, masks, nprtn, prtn_ptr):
    """orc_cr_bits: GENERATE_CR_BITS (alto.hpp:1250-1346) on sorted keys
    (nnz, kw) and their partitions.  Returns (keys with the conflict bits,
    cr_masks (nmode, kw))."""
    dims = np.ascontiguousarray(dims, dtype=np.uint64)
    keys = np.ascontiguousarray(keys, dtype=np.uint64).copy()
    nnz, kw = keys.shape
    cm = np.zeros((len(dims), kw), dtype=np.uint64)
    lib().orc_cr_bits(ctypes.c_int(len(dims)), _p64(dims), ctypes.c_uint64(nnz), ctypes.c_int(kw), _p64(keys),
                      _p64(np.ascontiguousarray(masks, dtype=np.uint64)), ctypes.c_int(nprtn),
                      _p64(np.ascontiguousarray(prtn_ptr, dtype=np.uint64)), _p64(cm))
    return keys, cm


def get_num_ptrn(nnz, max_threads):
    return lib().orc_get_num_ptrn(ctypes.c_uint64(int(nnz)), ctypes.c_int(int(max_threads)))


def mttkrp_coo(cidx, vals, factors, mode, out=None):
    cidx = [np.ascontiguousarray(c, dtype=np.uint64) for c in cidx]
    vals = np.ascontiguousarray(vals, dtype=np.float64)
    factors = [np.ascontiguousarray(f, dtype=np.float64) for f in factors]
    rank = factors[0].shape[1]
    if out is None:
        out = np.zeros_like(factors[mode])
    lib().orc_mttkrp_coo(ctypes.c_int(len(cidx)), ctypes.c_uint64(len(vals)), _ptr_array(cidx, "u"),
                         _pf(vals), _ptr_array(factors, "f"), ctypes.c_int(rank), ctypes.c_int(mode),
                         _pf(out))
    return out


def rowsparse_mttkrp(cidx, vals, mode, stream_mode, rsp_vals, ridx, rank):
    """Row-sparse MTTKRP as src/rowsparse_mttkrp.cpp:97-187 computes it after
    idxsort_hist (8-95): nonzeros bucketed by their `mode` index (ascending),
    per nonzero accum = val * prod_{m != mode} rsp_vals[m][m_idx] in mode
    order (m_idx = 0 for stream_mode, else ridx[m][i_m]), summed per bucket.
    ridx -1 contributes a zero row (the reference would read out of bounds).
    Pure Python loops: small inputs only.  Returns (rowind, rows)."""
    cidx = [np.asarray(c, dtype=np.int64) for c in cidx]
    vals = np.asarray(vals, dtype=np.float64)
    order = np.argsort(cidx[mode], kind="stable")
    keys = cidx[mode][order]
    starts = [0] + [i for i in range(1, len(keys)) if keys[i] != keys[i - 1]] if len(keys) else []
    bounds = list(zip(starts, starts[1:] + [len(keys)]))
    rowind = np.array([keys[a] for a, _ in bounds], dtype=np.uint64)
    out = np.zeros((len(bounds), rank))
    N = len(cidx)
    for b, (a, e) in enumerate(bounds):
        row_accum = np.zeros(rank)
        for i in order[a:e]:
            accum = np.full(rank, vals[i])
            for m in range(N):
                if m == mode:
                    continue
                r = 0 if m == stream_mode else int(ridx[m][cidx[m][i]])
                accum = accum * (rsp_vals[m][r] if r >= 0 else 0.0)
            row_accum = row_accum + accum
        out[b] += row_accum
    return rowind, out


class AltoTensor:
    """Oracle-side ALTO tensor: masks, sorted keys/values, partitions."""

    def __init__(self, dims, cidx, vals, nprtn, kw=None):
        self.dims = np.ascontiguousarray(dims, dtype=np.uint64)
        self.nmode = len(dims)
        self.masks, self.bits = alto_masks(self.dims, kw)
        self.kw = self.masks.shape[1]
        keys = linearize(cidx, self.masks)
        self.keys, self.vals, self.perm = sort_alto(keys, vals)
        self.nnz = len(self.vals)
        self.nprtn = nprtn
        self.prtn_ptr, self.intervals = prtn_alto(self.dims, self.keys, self.masks, nprtn)

    def variant(self, target, stream_mode=-1):
        return lib().orc_mttkrp_variant(ctypes.c_int(self.nmode), _p64(self.dims),
                                        ctypes.c_uint64(self.nnz), ctypes.c_int(stream_mode),
                                        ctypes.c_int(target))

    def mttkrp(self, factors, target, stream_mode=-1, out=None):
        factors = [np.ascontiguousarray(f, dtype=np.float64) for f in factors]
        rank = factors[0].shape[1]
        if out is None:
            out = np.zeros((int(self.dims[target]), rank), dtype=np.float64)
        var = lib().orc_mttkrp_alto(
            ctypes.c_int(self.nmode), _p64(self.dims), ctypes.c_uint64(self.nnz), ctypes.c_int(self.kw),
            _p64(self.keys), _pf(self.vals), _p64(self.masks), ctypes.c_int(self.nprtn),
            _p64(self.prtn_ptr), _p64(self.intervals), ctypes.c_int(stream_mode),
            _ptr_array(factors, "f"), ctypes.c_int(rank), ctypes.c_int(target), _pf(out))
        return out, var

    def cpd_als(self, factors, rank, max_iters, eps=1e-30, fit_mode=FIT_REFERENCE_DIAG):
        factors = [np.ascontiguousarray(f, dtype=np.float64).copy() for f in factors]
        lam = np.ones(rank, dtype=np.float64)
        fits = np.zeros(max_iters, dtype=np.float64)
        fits_exact = np.zeros(max_iters, dtype=np.float64)
        n = lib().orc_cpd_als(
            ctypes.c_int(self.nmode), _p64(self.dims), ctypes.c_uint64(self.nnz), ctypes.c_int(self.kw),
            _p64(self.keys), _pf(self.vals), _p64(self.masks), ctypes.c_int(self.nprtn),
            _p64(self.prtn_ptr), _p64(self.intervals), ctypes.c_int(rank), ctypes.c_int(max_iters),
            ctypes.c_double(eps), ctypes.c_int(fit_mode), _ptr_array(factors, "f"), _pf(lam),
            _pf(fits), _pf(fits_exact))
        return factors, lam, fits[:n], fits_exact[:n]


def cpd_fit(dims, factors, lam, grams, last_mttkrp, normX, fit_mode=FIT_REFERENCE_DIAG):
    """orc_cpd_fit: cpd_fit_alto (cpd.hpp:193-275) from the factors, lambda,
    the Gram matrices as given (the reference's are upper-triangle only) and
    the last mode's MTTKRP."""
    dims = np.ascontiguousarray(dims, dtype=np.uint64)
    factors = [np.ascontiguousarray(f, dtype=np.float64) for f in factors]
    grams = [np.ascontiguousarray(g, dtype=np.float64) for g in grams]
    lam = np.ascontiguousarray(lam, dtype=np.float64)
    mt = np.ascontiguousarray(last_mttkrp, dtype=np.float64)
    return lib().orc_cpd_fit(ctypes.c_int(len(dims)), _p64(dims), ctypes.c_int(factors[0].shape[1]),
                             _ptr_array(factors, "f"), _pf(lam), _ptr_array(grams, "f"), _pf(mt),
                             ctypes.c_double(normX), ctypes.c_int(fit_mode))


def gram_upper(U):
    U = np.ascontiguousarray(U, dtype=np.float64)
    R = U.shape[1]
    G = np.zeros((R, R), dtype=np.float64)
    lib().orc_gram_upper(ctypes.c_uint64(U.shape[0]), ctypes.c_int(R), _pf(U), _pf(G))
    return G


def kruskal_norm(U, which):
    U = np.ascontiguousarray(U, dtype=np.float64).copy()
    lam = np.zeros(U.shape[1], dtype=np.float64)
    lib().orc_kruskal_norm(ctypes.c_uint64(U.shape[0]), ctypes.c_int(U.shape[1]), _pf(U), _pf(lam),
                           ctypes.c_int(which))
    return U, lam


def pseudo_inverse_als(grams, mode, U):
    grams = [np.ascontiguousarray(g, dtype=np.float64).copy() for g in grams]
    U = np.ascontiguousarray(U, dtype=np.float64).copy()
    R = U.shape[1]
    fb = lib().orc_pseudo_inverse_als(ctypes.c_int(len(grams)), ctypes.c_int(R), _ptr_array(grams, "f"),
                                      ctypes.c_int(mode), ctypes.c_uint64(U.shape[0]), _pf(U))
    return U, grams, fb


def admm_cf(grams, mode, primal, dual, mttkrp, nonneg=True):
    grams = [np.ascontiguousarray(g, dtype=np.float64) for g in grams]
    primal = np.ascontiguousarray(primal, dtype=np.float64).copy()
    dual = np.ascontiguousarray(dual, dtype=np.float64).copy()
    mttkrp = np.ascontiguousarray(mttkrp, dtype=np.float64).copy()
    R = primal.shape[1]
    it = lib().orc_admm_cf(ctypes.c_uint64(primal.shape[0]), ctypes.c_int(R), ctypes.c_int(len(grams)),
                           ctypes.c_int(mode), _ptr_array(grams, "f"), ctypes.c_int(1 if nonneg else 0),
                           _pf(primal), _pf(dual), _pf(mttkrp))
    return primal, dual, it


def cpstream(dims, cidx, vals, stream_mode, rank, max_iters, eps, nonneg, init_factors,
             max_slices=0, nnz_threshold=0, timeslice_limit=1):
    """orc_cpstream: cpstream (src/streaming_cpd.cpp:26-366) with batches as
    next_dynamic_batch (streaming_sptensor.cpp:196-264).  outer_iters and
    admm_iters have one entry per batch, time_rows one row per time index."""
    dims = np.ascontiguousarray(dims, dtype=np.uint64)
    N = len(dims)
    cidx = [np.ascontiguousarray(c, dtype=np.uint64) for c in cidx]
    vals = np.ascontiguousarray(vals, dtype=np.float64)
    init = [np.ascontiguousarray(f, dtype=np.float64) for f in init_factors]
    nslices_max = int(dims[stream_mode])
    outs = [np.zeros((int(dims[m]), rank)) for m in range(N)]
    out_dims = np.zeros(N, dtype=np.uint64)
    trows = np.zeros((nslices_max, rank))
    outer = np.zeros(nslices_max, dtype=np.int32)
    admm = np.zeros(nslices_max)
    grams = np.zeros((N, rank, rank))
    nt = ctypes.c_uint64()
    n = lib().orc_cpstream(
        ctypes.c_int(N), _p64(dims), ctypes.c_uint64(len(vals)), _ptr_array(cidx, "u"), _pf(vals),
        ctypes.c_int(stream_mode), ctypes.c_int(rank), ctypes.c_int(max_iters), ctypes.c_double(eps),
        ctypes.c_int(1 if nonneg else 0), ctypes.c_int(max_slices), ctypes.c_uint64(nnz_threshold),
        ctypes.c_uint64(timeslice_limit), _ptr_array(init, "f"),
        _ptr_array(outs, "f"), _p64(out_dims), _pf(trows),
        outer.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _pf(admm), _pf(grams), ctypes.byref(nt))
    outs = [o[: int(out_dims[m])] for m, o in enumerate(outs)]
    return dict(factors=outs, dims=out_dims, time_rows=trows[: nt.value], outer_iters=outer[:n],
                admm_iters=admm[:n], grams=grams, slices=n, time_slices=int(nt.value))


def spcpstream(dims, cidx, vals, stream_mode, rank, max_iters, eps, init_factors, max_slices=0,
               nnz_threshold=0, timeslice_limit=1):
    """orc_spcpstream: cpstream with use_spcpstream (src/streaming_cpd.cpp:26-366)
    driving spcpstream_alto_iter (streaming_cpd.cpp:466-1085).  Same outputs as
    cpstream, without the ADMM counts."""
    dims = np.ascontiguousarray(dims, dtype=np.uint64)
    N = len(dims)
    cidx = [np.ascontiguousarray(c, dtype=np.uint64) for c in cidx]
    vals = np.ascontiguousarray(vals, dtype=np.float64)
    init = [np.ascontiguousarray(f, dtype=np.float64) for f in init_factors]
    nslices_max = int(dims[stream_mode])
    outs = [np.zeros((int(dims[m]), rank)) for m in range(N)]
    out_dims = np.zeros(N, dtype=np.uint64)
    trows = np.zeros((nslices_max, rank))
    outer = np.zeros(nslices_max, dtype=np.int32)
    grams = np.zeros((N, rank, rank))
    nt = ctypes.c_uint64()
    n = lib().orc_spcpstream(
        ctypes.c_int(N), _p64(dims), ctypes.c_uint64(len(vals)), _ptr_array(cidx, "u"), _pf(vals),
        ctypes.c_int(stream_mode), ctypes.c_int(rank), ctypes.c_int(max_iters), ctypes.c_double(eps),
        ctypes.c_int(max_slices), ctypes.c_uint64(nnz_threshold), ctypes.c_uint64(timeslice_limit),
        _ptr_array(init, "f"), _ptr_array(outs, "f"), _p64(out_dims), _pf(trows),
        outer.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _pf(grams), ctypes.byref(nt))
    if n < 0:
        raise ValueError("orc_spcpstream: bad arguments")
    outs = [o[: int(out_dims[m])] for m, o in enumerate(outs)]
    return dict(factors=outs, dims=out_dims, time_rows=trows[: nt.value], outer_iters=outer[:n], grams=grams,
                slices=n, time_slices=int(nt.value))


def cpd_error(cidx, vals, factors):
    """cpd_error (src/streaming_cpd.cpp:368-464) with kruskal_norm
    (src/kruskal_model.cpp:400-442), weights 1: the mode with the fewest
    factor rows (first such) gives <X, M> = sum_r sum_i U[i, r] MTTKRP[i, r]
    (per-column sums first, then over columns, as the reference's accumF);
    ||X||^2 = sum of squared values; ||M||^2 = sum_{i<=j} (2 - [i==j]) *
    prod_m (U_m^T U_m)[i, j] (absolute value).  Returns (err, parts)."""
    factors = [np.ascontiguousarray(f, dtype=np.float64) for f in factors]
    rows = [f.shape[0] for f in factors]
    mn = int(np.argmin(rows))
    mt = mttkrp_coo(cidx, vals, factors, mn, out=np.zeros_like(factors[mn]))
    acc = np.zeros(factors[mn].shape[1])
    for i in range(rows[mn]):
        acc += factors[mn][i] * mt[i]
    inner = 0.0
    for r in range(len(acc)):
        inner += acc[r]
    xx = float(np.dot(vals, vals))
    R = factors[0].shape[1]
    scratch = np.ones((R, R))
    for f in factors:
        scratch *= f.T @ f
    zz = 0.0
    for i in range(R):
        zz += scratch[i, i]
        for j in range(i + 1, R):
            zz += scratch[i, j] * 2
    zz = abs(zz)
    resid = np.sqrt(xx + zz - 2 * inner)
    return resid / np.sqrt(xx), (xx, zz, inner)


def stream_relabel(cidx, vals, stream_mode):
    """The tensor as StreamingSparseTensor leaves it (src/streaming_sptensor.cpp:69-124):
    sorted by (stream_mode, stream_mode+1, ...) mod N (tt_sort, sort.cpp:477-519;
    ties keep input order here) and every other mode relabelled by first
    appearance in that order.  Returns (cidx, vals)."""
    N = len(cidx)
    keys = [np.asarray(cidx[(stream_mode + t) % N], dtype=np.uint64) for t in range(N)]
    order = np.lexsort(keys[::-1])
    out = [np.asarray(c, dtype=np.uint64)[order].copy() for c in cidx]
    for m in range(N):
        if m == stream_mode:
            continue
        _, first = np.unique(out[m], return_index=True)
        lab = np.empty(len(first), dtype=np.uint64)
        lab[np.argsort(first, kind="stable")] = np.arange(len(first), dtype=np.uint64)
        uniq = np.unique(out[m])
        out[m] = lab[np.searchsorted(uniq, out[m])]
    return out, np.asarray(vals, dtype=np.float64)[order]


def read_tns(path, zero_based=False):
    """Reads a FROSTT .tns file like ImportSparseTensor (sptensor.cpp:74-158):
    '#' lines skipped; dims = max index (+1 if zero based)."""
    rows = []
    with open(path) as f:
        for line in f:
            if len(line) > 1 and line[0] != "#":
                rows.append(line.split())
    nmode = len(rows[0]) - 1
    idx = np.array([[int(t) for t in r[:nmode]] for r in rows], dtype=np.uint64)
    vals = np.array([float(r[nmode]) for r in rows], dtype=np.float64)
    off = 0 if zero_based else 1
    cidx = [idx[:, m] - np.uint64(off) for m in range(nmode)]
    dims = [int(idx[:, m].max()) + (1 if zero_based else 0) for m in range(nmode)]
    return dims, cidx, vals
