"""Python binding of libalto_amd.so (the C ABI in include/alto_hip.h).

Mirrors the reference's host interface names so that callers and tests read
like the reference's own code:
  create_alto        include/alto.hpp:436-582   -> AltoTensor
  mttkrp_alto_par    include/alto.hpp:902-908   -> mttkrp_alto_par()
  mttkrp (COO)       src/mttkrp.cpp:122-156     -> mttkrp_coo()
  cpd_alto           include/cpd.hpp:47-191     -> cpd_alto()
  cpstream           src/streaming_cpd.cpp:26-366 -> cpstream()

There is no CPU fallback: if the shared library is missing or no GPU is
usable, every call raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libalto_amd.so")

OK = 0
PREC_F64, PREC_F32, PREC_BF16 = 0, 1, 2
FIT_EXACT, FIT_REFERENCE_DIAG = 0, 1
_PREC_DTYPE = {PREC_F64: np.float64, PREC_F32: np.float32, PREC_BF16: np.uint16}

_u64p = ctypes.POINTER(ctypes.c_uint64)
_f64p = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p
_lib = None


class AltoError(RuntimeError):
    pass


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE, "-j8"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise AltoError("libalto_amd.so is not built (%s); run __graft_entry__.build()" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        L.alto_hip_error_string.restype = ctypes.c_char_p
        L.alto_hip_last_error.restype = ctypes.c_char_p
        _lib = L
    return _lib


def check(status, what=""):
    if status != OK:
        L = lib()
        raise AltoError("%s failed: %s (%d): %s" % (
            what, L.alto_hip_error_string(status).decode(), status, L.alto_hip_last_error().decode()))
    return status


def init(device=0):
    check(lib().alto_hip_init(ctypes.c_int(device)), "alto_hip_init")


def device_name():
    buf = ctypes.create_string_buffer(256)
    check(lib().alto_hip_device_name(buf, ctypes.c_size_t(256)), "device_name")
    return buf.value.decode()


def synchronize(stream=None):
    check(lib().alto_hip_stream_sync(_vp(stream)), "stream_sync")


class DeviceArray:
    """A device allocation with numpy-like shape/dtype, freed on close/GC."""

    def __init__(self, shape, dtype):
        self.shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = _vp()
        check(lib().alto_hip_malloc(ctypes.byref(p), ctypes.c_size_t(max(self.nbytes, 16))), "malloc")
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, a, stream=None):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        if d.nbytes:
            check(lib().alto_hip_memcpy_h2d(_vp(d.ptr), a.ctypes.data_as(_vp), ctypes.c_size_t(d.nbytes),
                                            _vp(stream)), "h2d")
        return d

    @classmethod
    def zeros(cls, shape, dtype, stream=None):
        d = cls(shape, dtype)
        check(lib().alto_hip_memset(_vp(d.ptr), 0, ctypes.c_size_t(max(d.nbytes, 16)), _vp(stream)), "memset")
        return d

    def zero_(self, stream=None):
        check(lib().alto_hip_memset(_vp(self.ptr), 0, ctypes.c_size_t(max(self.nbytes, 16)), _vp(stream)),
              "memset")
        return self

    def numpy(self, stream=None):
        a = np.empty(self.shape, dtype=self.dtype)
        if self.nbytes:
            check(lib().alto_hip_memcpy_d2h(a.ctypes.data_as(_vp), _vp(self.ptr), ctypes.c_size_t(self.nbytes),
                                            _vp(stream)), "d2h")
        return a

    def close(self):
        if getattr(self, "ptr", None):
            lib().alto_hip_free(_vp(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def to_prec(a, prec, stream=None):
    """Device copy of a host FP64 array in the given precision."""
    src = DeviceArray.from_numpy(np.ascontiguousarray(a, dtype=np.float64), stream)
    if prec == PREC_F64:
        return src
    dst = DeviceArray(src.shape, _PREC_DTYPE[prec])
    check(lib().alto_hip_convert(_vp(dst.ptr), ctypes.c_int(prec), _vp(src.ptr), ctypes.c_int(PREC_F64),
                                 ctypes.c_size_t(int(np.prod(src.shape))), _vp(stream)), "convert")
    src.close()
    return dst


def convert(dst, dst_prec, src, src_prec, n, stream=None):
    check(lib().alto_hip_convert(_vp(dst.ptr), ctypes.c_int(dst_prec), _vp(src.ptr), ctypes.c_int(src_prec),
                                 ctypes.c_size_t(n), _vp(stream)), "convert")


class _Coo(ctypes.Structure):
    _fields_ = [("nmode", ctypes.c_int), ("dims", _u64p), ("nnz", ctypes.c_uint64),
                ("cidx", ctypes.POINTER(_u64p)), ("vals", _f64p), ("on_device", ctypes.c_int)]


def _coo(dims, cidx, vals, on_device):
    dims_a = np.ascontiguousarray(dims, dtype=np.uint64)
    N = len(dims_a)
    arr = (_u64p * N)()
    keep = [dims_a]
    if on_device:
        for m in range(N):
            arr[m] = ctypes.cast(_vp(cidx[m].ptr), _u64p)
        vp = ctypes.cast(_vp(vals.ptr), _f64p)
        nnz = vals.shape[0]
    else:
        cs = [np.ascontiguousarray(c, dtype=np.uint64) for c in cidx]
        vv = np.ascontiguousarray(vals, dtype=np.float64)
        keep += cs + [vv]
        for m in range(N):
            arr[m] = cs[m].ctypes.data_as(_u64p)
        vp = vv.ctypes.data_as(_f64p)
        nnz = len(vv)
    c = _Coo(N, dims_a.ctypes.data_as(_u64p), nnz, arr, vp, 1 if on_device else 0)
    return c, keep + [arr]


class AltoTensor:
    """Device ALTO tensor: create_alto (include/alto.hpp:436-582)."""

    def __init__(self, dims, cidx, vals, nprtn=1, key_words=0, on_device=False, stream=None):
        c, keep = _coo(dims, cidx, vals, on_device)
        h = _vp()
        check(lib().alto_hip_tensor_create(ctypes.byref(c), ctypes.c_int(nprtn), ctypes.c_int(key_words),
                                           _vp(stream), ctypes.byref(h)), "create_alto")
        del keep
        self.handle = h.value
        self._info(dims)

    def _info(self, dims):
        nm, nnz, kw, npr, tb = ctypes.c_int(), ctypes.c_uint64(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        check(lib().alto_hip_tensor_info(_vp(self.handle), ctypes.byref(nm), ctypes.byref(nnz), ctypes.byref(kw),
                                         ctypes.byref(npr), ctypes.byref(tb)), "tensor_info")
        self.nmode, self.nnz, self.kw, self.nprtn, self.total_bits = nm.value, nnz.value, kw.value, npr.value, tb.value
        self.dims = np.ascontiguousarray(dims, dtype=np.uint64).copy()

    def update(self, dims, cidx, vals, nprtn=1, on_device=False, stream=None):
        """update_salto (alto.hpp:195-299): rebuild in place from a new COO tensor."""
        c, keep = _coo(dims, cidx, vals, on_device)
        check(lib().alto_hip_tensor_update(_vp(self.handle), ctypes.byref(c), ctypes.c_int(nprtn), _vp(stream)),
              "update_salto")
        del keep
        self._info(dims)

    def masks(self):
        m = np.zeros((self.nmode, self.kw), dtype=np.uint64)
        check(lib().alto_hip_tensor_masks(_vp(self.handle), m.ctypes.data_as(_u64p)), "masks")
        return m

    def keys(self):
        k = np.zeros((self.nnz, self.kw), dtype=np.uint64)
        check(lib().alto_hip_tensor_keys(_vp(self.handle), k.ctypes.data_as(_u64p)), "keys")
        return k

    def packed_keys(self):
        """Sorted keys in the ALT_PEXT layout and the per-mode bit positions."""
        k = np.zeros((self.nnz, self.kw), dtype=np.uint64)
        pos = np.zeros(len(self.dims), dtype=np.int32)
        check(lib().alto_hip_tensor_packed_keys(_vp(self.handle), k.ctypes.data_as(_u64p),
                                                pos.ctypes.data_as(ctypes.POINTER(ctypes.c_int))), "packed_keys")
        return k, pos

    def cr_bits(self):
        """GENERATE_CR_BITS (alto.hpp:1250-1346): sets the conflict-resolution
        bits in the keys; returns the per-mode masks, shape (nmode, kw)."""
        m = np.zeros((len(self.dims), self.kw), dtype=np.uint64)
        check(lib().alto_hip_tensor_cr_bits(_vp(self.handle), m.ctypes.data_as(_u64p), None), "cr_bits")
        return m

    def vals(self):
        v = np.zeros(self.nnz, dtype=np.float64)
        check(lib().alto_hip_tensor_vals(_vp(self.handle), v.ctypes.data_as(_f64p)), "vals")
        return v

    def perm(self):
        p = np.zeros(self.nnz, dtype=np.uint32)
        check(lib().alto_hip_tensor_perm(_vp(self.handle), p.ctypes.data_as(_vp)), "perm")
        return p

    def partitions(self):
        ptr = np.zeros(self.nprtn + 1, dtype=np.uint64)
        iv = np.zeros((self.nprtn, self.nmode, 2), dtype=np.uint64)
        check(lib().alto_hip_tensor_partitions(_vp(self.handle), ptr.ctypes.data_as(_u64p),
                                               iv.ctypes.data_as(_u64p)), "partitions")
        return ptr, iv

    def close(self):
        if getattr(self, "handle", None):
            lib().alto_hip_tensor_destroy(_vp(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def create_alto(dims, cidx, vals, nprtn=1, key_words=0, on_device=False, stream=None):
    return AltoTensor(dims, cidx, vals, nprtn, key_words, on_device, stream)


def _ptrs(devs):
    arr = (_vp * len(devs))()
    for i, d in enumerate(devs):
        arr[i] = d.ptr if d is not None else None
    return arr


def _check_out(out, rows, rank, f64):
    """The kernels write rows x rank elements of FP64 (f64) or FP32: refuse an
    output array of another element type or a smaller shape."""
    want = np.dtype(np.float64 if f64 else np.float32)
    if out.dtype != want:
        raise AltoError("output array is %s, the MTTKRP writes %s" % (out.dtype, want))
    if out.nbytes < rows * rank * want.itemsize:
        raise AltoError("output array holds %d bytes, the MTTKRP writes %d x %d" % (out.nbytes, rows, rank))


def mttkrp_alto_par(target, factors, rank, at, out, prec=PREC_F64, out_f64=True, algo=0, stream=None):
    """out += MTTKRP(target) on device arrays (mttkrp_alto_par, alto.hpp:902-908).
    out: FP64 when prec is PREC_F64 or out_f64, else FP32."""
    _check_out(out, int(at.dims[target]), rank, out_f64 or prec == PREC_F64)
    check(lib().alto_hip_mttkrp_algo(_vp(at.handle), ctypes.c_int(target), ctypes.c_int(prec), ctypes.c_int(rank),
                                     _ptrs(factors), _vp(out.ptr), ctypes.c_int(1 if out_f64 else 0),
                                     ctypes.c_size_t(out.nbytes), ctypes.c_int(algo), _vp(stream)),
          "mttkrp_alto_par")


def mttkrp_config(target, factors, rank, at, out, prec=PREC_F64, out_f64=True, algo=2, segment=0, flags=-1,
                  stream=None):
    """mttkrp_alto_par with explicit launch knobs (segment: entries per lane
    group, 0 = from the entry count; flags -1 = defaults)."""
    _check_out(out, int(at.dims[target]), rank, out_f64 or prec == PREC_F64)
    check(lib().alto_hip_mttkrp_config(_vp(at.handle), ctypes.c_int(target), ctypes.c_int(prec),
                                       ctypes.c_int(rank), _ptrs(factors), _vp(out.ptr),
                                       ctypes.c_int(1 if out_f64 else 0), ctypes.c_size_t(out.nbytes),
                                       ctypes.c_int(algo), ctypes.c_uint64(segment), ctypes.c_int(flags),
                                       _vp(stream)),
          "mttkrp_config")


def mttkrp_coo(dims, cidx_dev, vals_dev, target, rank, factors, out, stream=None):
    c, keep = _coo(dims, cidx_dev, vals_dev, True)
    check(lib().alto_hip_mttkrp_coo(ctypes.byref(c), ctypes.c_int(target), ctypes.c_int(rank), _ptrs(factors),
                                    _vp(out.ptr), _vp(stream)), "mttkrp_coo")


class _AlsOpts(ctypes.Structure):
    _fields_ = [("rank", ctypes.c_int), ("max_iters", ctypes.c_int), ("epsilon", ctypes.c_double),
                ("prec", ctypes.c_int), ("fit_mode", ctypes.c_int), ("verbose", ctypes.c_int)]


def cpd_alto(at, factors, rank, max_iters, eps=1e-30, prec=PREC_F64, fit_mode=FIT_REFERENCE_DIAG,
             verbose=False, stream=None):
    """CP-ALS (cpd_alto, include/cpd.hpp:47-191).  factors: host FP64 initial
    model.  Returns (factors, lambda, fits, fits_exact, timings)."""
    fs = [np.ascontiguousarray(f, dtype=np.float64).copy() for f in factors]
    arr = (_f64p * len(fs))()
    for i, f in enumerate(fs):
        arr[i] = f.ctypes.data_as(_f64p)
    lam = np.zeros(rank)
    fits = np.zeros(max(max_iters, 1))
    fits_exact = np.zeros(max(max_iters, 1))
    tim = np.zeros(8)
    iters = ctypes.c_int()
    o = _AlsOpts(rank, max_iters, eps, prec, fit_mode, 1 if verbose else 0)
    check(lib().alto_hip_cpd_als(_vp(at.handle), ctypes.byref(o), arr, lam.ctypes.data_as(_f64p),
                                 fits.ctypes.data_as(_f64p), fits_exact.ctypes.data_as(_f64p),
                                 ctypes.byref(iters), tim.ctypes.data_as(_f64p), _vp(stream)), "cpd_alto")
    n = iters.value
    return fs, lam, fits[:n], fits_exact[:n], tim


def cpd_fit(last_factor, last_mttkrp, lam, grams, normX, fit_mode=FIT_REFERENCE_DIAG, stream=None):
    """cpd_fit_alto (cpd.hpp:193-275) on the GPU from host arrays: the last
    mode's factor and MTTKRP (rows x R), lambda, the nmode Gram matrices
    (their upper triangles are read) and ||X||^2."""
    U = np.ascontiguousarray(last_factor, dtype=np.float64)
    MT = np.ascontiguousarray(last_mttkrp, dtype=np.float64)
    lam = np.ascontiguousarray(lam, dtype=np.float64)
    g = np.ascontiguousarray(np.stack([np.asarray(x, dtype=np.float64) for x in grams]))
    fit = ctypes.c_double()
    check(lib().alto_hip_cpd_fit(ctypes.c_int(g.shape[0]), ctypes.c_int(U.shape[1]), ctypes.c_uint64(U.shape[0]),
                                 U.ctypes.data_as(_f64p), MT.ctypes.data_as(_f64p), lam.ctypes.data_as(_f64p),
                                 g.ctypes.data_as(_f64p), ctypes.c_double(normX), ctypes.c_int(fit_mode),
                                 ctypes.byref(fit), _vp(stream)), "cpd_fit")
    return fit.value


def pseudo_inverse(grams, mode, U, stream=None):
    """pseudo_inverse (cpd.hpp:377-480) on the GPU: returns (U V^-1, fallback)."""
    g = np.ascontiguousarray(np.stack([np.asarray(x, dtype=np.float64) for x in grams]))
    U = np.ascontiguousarray(U, dtype=np.float64).copy()
    fb = ctypes.c_int()
    check(lib().alto_hip_pseudo_inverse(ctypes.c_int(g.shape[0]), ctypes.c_int(U.shape[1]), g.ctypes.data_as(_f64p),
                                        ctypes.c_int(mode), ctypes.c_uint64(U.shape[0]), U.ctypes.data_as(_f64p),
                                        ctypes.byref(fb), _vp(stream)), "pseudo_inverse")
    return U, fb.value


def cpd_alto_batch(tensors, factors, ranks, max_iters, eps=1e-30, prec=PREC_F64, fit_mode=FIT_REFERENCE_DIAG,
                   nstreams=8):
    """CP-ALS on many independent ALTO tensors concurrently
    (alto_hip_cpd_als_batch).  factors[k]: host FP64 initial model of tensor k;
    ranks[k]: its rank.  Returns (factors, lambdas, fits, wall_seconds)."""
    n = len(tensors)
    fs = [[np.ascontiguousarray(f, dtype=np.float64).copy() for f in fk] for fk in factors]
    lams = [np.zeros(int(r)) for r in ranks]
    fits = [np.zeros(max(int(max_iters), 1)) for _ in range(n)]
    iters = np.zeros(n, dtype=np.int32)
    handles = (_vp * n)(*[t.handle for t in tensors])
    opts = (_AlsOpts * n)(*[_AlsOpts(int(r), int(max_iters), eps, prec, fit_mode, 0) for r in ranks])
    farr = (ctypes.POINTER(_f64p) * n)()
    keep = []
    for k in range(n):
        a = (_f64p * len(fs[k]))(*[f.ctypes.data_as(_f64p) for f in fs[k]])
        keep.append(a)
        farr[k] = ctypes.cast(a, ctypes.POINTER(_f64p))
    larr = (_f64p * n)(*[l.ctypes.data_as(_f64p) for l in lams])
    fitarr = (_f64p * n)(*[f.ctypes.data_as(_f64p) for f in fits])
    tim = np.zeros(2)
    check(lib().alto_hip_cpd_als_batch(ctypes.c_int(n), handles, opts, farr, larr, fitarr,
                                       iters.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), ctypes.c_int(nstreams),
                                       tim.ctypes.data_as(_f64p)), "cpd_als_batch")
    return fs, lams, [f[: iters[k]] for k, f in enumerate(fits)], float(tim[0])


def cpd_alto_batch_dev(tensors, factors, ranks, max_iters, eps=1e-30, prec=PREC_F64, fit_mode=FIT_REFERENCE_DIAG,
                       nstreams=8):
    """alto_hip_cpd_als_batch_dev: as cpd_alto_batch with device-resident
    models.  factors[k]: DeviceArrays (FP64, dims x rank) holding tensor k's
    initial model, overwritten with the final one.  Returns (lambdas, fits,
    wall_seconds)."""
    n = len(tensors)
    for k in range(n):
        for m, f in enumerate(factors[k]):
            if f.dtype != np.float64 or f.shape != (int(tensors[k].dims[m]), int(ranks[k])):
                raise ValueError("factor %d of tensor %d: need FP64 %d x %d" %
                                 (m, k, int(tensors[k].dims[m]), int(ranks[k])))
    lams = [np.zeros(int(r)) for r in ranks]
    fits = [np.zeros(max(int(max_iters), 1)) for _ in range(n)]
    iters = np.zeros(n, dtype=np.int32)
    handles = (_vp * n)(*[t.handle for t in tensors])
    opts = (_AlsOpts * n)(*[_AlsOpts(int(r), int(max_iters), eps, prec, fit_mode, 0) for r in ranks])
    farr = (ctypes.POINTER(_f64p) * n)()
    keep = []
    for k in range(n):
        a = (_f64p * len(factors[k]))(*[ctypes.cast(_vp(f.ptr), _f64p) for f in factors[k]])
        keep.append(a)
        farr[k] = ctypes.cast(a, ctypes.POINTER(_f64p))
    larr = (_f64p * n)(*[l.ctypes.data_as(_f64p) for l in lams])
    fitarr = (_f64p * n)(*[f.ctypes.data_as(_f64p) for f in fits])
    tim = np.zeros(2)
    check(lib().alto_hip_cpd_als_batch_dev(ctypes.c_int(n), handles, opts, farr, larr, fitarr,
                                           iters.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                           ctypes.c_int(nstreams), tim.ctypes.data_as(_f64p)), "cpd_als_batch_dev")
    return lams, [f[: iters[k]] for k, f in enumerate(fits)], float(tim[0])


def cpstream(dims, cidx, vals, stream_mode, rank, max_iters, eps, nonneg, init_factors, max_slices=0,
             prec=PREC_F64, nnz_threshold=0, timeslice_limit=1, stream=None):
    """CP-stream (cpstream, src/streaming_cpd.cpp:26-366), host COO input.
    Time slices are batched as next_dynamic_batch does (timeslice_limit,
    nnz_threshold; 0 = no threshold); per-batch outputs (outer_iters,
    admm_iters) have one entry per batch, time_rows one row per time index."""
    c, keep = _coo(dims, cidx, vals, False)
    dims = np.ascontiguousarray(dims, dtype=np.uint64)
    N = len(dims)
    init = [np.ascontiguousarray(f, dtype=np.float64) for f in init_factors]
    iarr = (_f64p * N)()
    for i, f in enumerate(init):
        iarr[i] = f.ctypes.data_as(_f64p)
    nsl = int(dims[stream_mode])
    outs = [np.zeros((int(dims[m]), rank)) for m in range(N)]
    oarr = (_f64p * N)()
    for i, f in enumerate(outs):
        oarr[i] = f.ctypes.data_as(_f64p)
    out_dims = np.zeros(N, dtype=np.uint64)
    trows = np.zeros((nsl, rank))
    outer = np.zeros(nsl, dtype=np.int32)
    admm = np.zeros(nsl)
    grams = np.zeros((N, rank, rank))
    tim = np.zeros(4)
    ferr = np.zeros(1)
    n = lib().alto_hip_cpstream(ctypes.byref(c), ctypes.c_int(stream_mode), ctypes.c_int(rank),
                                ctypes.c_int(max_iters), ctypes.c_double(eps), ctypes.c_int(1 if nonneg else 0),
                                ctypes.c_int(max_slices), ctypes.c_uint64(nnz_threshold),
                                ctypes.c_uint64(timeslice_limit), ctypes.c_int(prec), iarr, oarr,
                                out_dims.ctypes.data_as(_u64p),
                                trows.ctypes.data_as(_f64p), outer.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                admm.ctypes.data_as(_f64p), grams.ctypes.data_as(_f64p),
                                ferr.ctypes.data_as(_f64p), tim.ctypes.data_as(_f64p), _vp(stream))
    if n < 0:
        check(n, "cpstream")
    outs = [o[: int(out_dims[m])] for m, o in enumerate(outs)]
    nt = int(tim[3])
    return dict(factors=outs, dims=out_dims, time_rows=trows[:nt], outer_iters=outer[:n],
                admm_iters=admm[:n], grams=grams, slices=n, time_slices=nt, timings=tim,
                final_error=float(ferr[0]))


def stream_prepare(dims, cidx, vals, stream_mode, stream=None):
    """StreamingSparseTensor on the device (streaming_sptensor.cpp:69-264):
    host COO in; returns (cidx, vals, slice_ptr, slice_dims) of the sorted,
    relabelled stream."""
    c, keep = _coo(dims, cidx, vals, False)
    N = len(dims)
    n = len(vals)
    outs = [np.zeros(n, dtype=np.uint64) for _ in range(N)]
    oarr = (_u64p * N)()
    for i, o in enumerate(outs):
        oarr[i] = o.ctypes.data_as(_u64p)
    ov = np.zeros(n)
    ns = ctypes.c_uint64()
    cap = max(n, 1)
    ptr = np.zeros(cap + 1, dtype=np.uint64)
    sd = np.zeros((cap, N), dtype=np.uint64)
    check(lib().alto_hip_stream_prepare(ctypes.byref(c), ctypes.c_int(stream_mode), oarr, ov.ctypes.data_as(_f64p),
                                        ctypes.c_uint64(cap), ctypes.byref(ns), ptr.ctypes.data_as(_u64p),
                                        sd.ctypes.data_as(_u64p), _vp(stream)), "stream_prepare")
    T = int(ns.value)
    return outs, ov, ptr[: T + 1], sd[:T]


def spcpstream(dims, cidx, vals, stream_mode, rank, max_iters, eps, init_factors, max_slices=0,
               prec=PREC_F64, nnz_threshold=0, timeslice_limit=1, stream=None):
    """spCP-stream (cpstream with use_spcpstream, src/streaming_cpd.cpp:26-366,
    spcpstream_alto_iter 466-1085), host COO input.  Outputs as cpstream,
    without ADMM counts."""
    c, keep = _coo(dims, cidx, vals, False)
    dims = np.ascontiguousarray(dims, dtype=np.uint64)
    N = len(dims)
    init = [np.ascontiguousarray(f, dtype=np.float64) for f in init_factors]
    iarr = (_f64p * N)()
    for i, f in enumerate(init):
        iarr[i] = f.ctypes.data_as(_f64p)
    nsl = int(dims[stream_mode])
    outs = [np.zeros((int(dims[m]), rank)) for m in range(N)]
    oarr = (_f64p * N)()
    for i, f in enumerate(outs):
        oarr[i] = f.ctypes.data_as(_f64p)
    out_dims = np.zeros(N, dtype=np.uint64)
    trows = np.zeros((nsl, rank))
    outer = np.zeros(nsl, dtype=np.int32)
    grams = np.zeros((N, rank, rank))
    tim = np.zeros(4)
    ferr = np.zeros(1)
    n = lib().alto_hip_spcpstream(ctypes.byref(c), ctypes.c_int(stream_mode), ctypes.c_int(rank),
                                  ctypes.c_int(max_iters), ctypes.c_double(eps), ctypes.c_int(max_slices),
                                  ctypes.c_uint64(nnz_threshold), ctypes.c_uint64(timeslice_limit),
                                  ctypes.c_int(prec), iarr, oarr, out_dims.ctypes.data_as(_u64p),
                                  trows.ctypes.data_as(_f64p), outer.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                  grams.ctypes.data_as(_f64p), ferr.ctypes.data_as(_f64p),
                                  tim.ctypes.data_as(_f64p), _vp(stream))
    if n < 0:
        check(n, "spcpstream")
    outs = [o[: int(out_dims[m])] for m, o in enumerate(outs)]
    nt = int(tim[3])
    return dict(factors=outs, dims=out_dims, time_rows=trows[:nt], outer_iters=outer[:n], grams=grams,
                slices=n, time_slices=nt, timings=tim, final_error=float(ferr[0]))


def rowsparse_mttkrp(dims, cidx, vals, mode, stream_mode, rank, rsp_vals, ridx, stream=None):
    """Row-sparse MTTKRP (rowsparse_mttkrp, src/rowsparse_mttkrp.cpp:97-187) on
    host arrays.  rsp_vals[m]: stored rows (k x rank) of mode m; ridx[m]:
    int32 map index -> stored row (-1 absent); the stream mode uses its row 0.
    Returns (rowind, rows)."""
    c, keep = _coo(dims, cidx, vals, False)
    N = len(dims)
    I = int(dims[mode])
    rv = [np.ascontiguousarray(rsp_vals[m] if m != mode and rsp_vals[m] is not None else np.zeros((1, rank)),
                               dtype=np.float64) for m in range(N)]
    rx = [np.ascontiguousarray(ridx[m] if m not in (mode, stream_mode) else np.zeros(1, np.int32), dtype=np.int32)
          for m in range(N)]
    varr = (_f64p * N)()
    xarr = (ctypes.POINTER(ctypes.c_int32) * N)()
    for m in range(N):
        varr[m] = rv[m].ctypes.data_as(_f64p)
        xarr[m] = rx[m].ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
    rows = np.ascontiguousarray([r.shape[0] for r in rv], dtype=np.uint64)
    rowind = np.zeros(I, dtype=np.uint64)
    out = np.zeros((I, rank))
    nnzr = ctypes.c_uint64()
    check(lib().alto_hip_rowsparse_mttkrp(ctypes.byref(c), ctypes.c_int(mode), ctypes.c_int(stream_mode),
                                          ctypes.c_int(rank), varr, rows.ctypes.data_as(_u64p), xarr,
                                          rowind.ctypes.data_as(_u64p), out.ctypes.data_as(_f64p),
                                          ctypes.byref(nnzr), _vp(stream)), "rowsparse_mttkrp")
    n = nnzr.value
    return rowind[:n], out[:n]


def cpd_error(dims, cidx, vals, factors, stream=None, on_device=False):
    """cpd_error (src/streaming_cpd.cpp:368-464): relative residual of the
    Kruskal model (weights 1).  Host COO and host FP64 factors, or, with
    on_device, DeviceArrays for both (the C ABI reads every array where
    coo->on_device says).  Returns (err, (||X||^2, ||M||^2, <X, M>))."""
    N = len(dims)
    if on_device:
        if not all(isinstance(f, DeviceArray) for f in factors):
            raise AltoError("cpd_error(on_device=True) takes DeviceArray factors")
        R = int(factors[0].shape[1])
        for m, f in enumerate(factors):
            if f.shape != (int(dims[m]), R) or f.dtype != np.float64:
                raise AltoError("factor %d is %s %s, expected (%d, %d) float64" % (m, f.shape, f.dtype,
                                                                                  int(dims[m]), R))
        c, keep = _coo(dims, cidx, vals, True)
        arr = _ptrs(factors)
        err = ctypes.c_double()
        parts = np.zeros(3)
        check(lib().alto_hip_cpd_error(ctypes.byref(c), ctypes.c_int(R), arr, ctypes.byref(err),
                                       parts.ctypes.data_as(_f64p), _vp(stream)), "cpd_error")
        return err.value, tuple(parts)
    c, keep = _coo(dims, cidx, vals, False)
    R = int(np.asarray(factors[0]).shape[1])
    fs = [np.ascontiguousarray(f, dtype=np.float64) for f in factors]
    for m, f in enumerate(fs):
        if f.shape != (int(dims[m]), R):
            raise AltoError("factor %d has shape %s, expected (%d, %d)" % (m, f.shape, int(dims[m]), R))
    arr = (_f64p * N)(*[f.ctypes.data_as(_f64p) for f in fs])
    err = ctypes.c_double()
    parts = np.zeros(3)
    check(lib().alto_hip_cpd_error(ctypes.byref(c), ctypes.c_int(R), arr, ctypes.byref(err),
                                   parts.ctypes.data_as(_f64p), _vp(stream)), "cpd_error")
    return err.value, tuple(parts)


def cpd_error_noalloc(dims, cidx_dev, vals_dev, factors_dev, out_dev, stream=None):
    """alto_hip_cpd_error_noalloc: device COO and factors in; out_dev (device,
    4 doubles) receives err, ||X||^2, ||M||^2, <X, M> once the stream reaches
    the call.  Allocates nothing on the device."""
    c, keep = _coo(dims, cidx_dev, vals_dev, True)
    R = int(factors_dev[0].shape[1])
    check(lib().alto_hip_cpd_error_noalloc(ctypes.byref(c), ctypes.c_int(R), _ptrs(factors_dev),
                                           _vp(out_dev.ptr), _vp(stream)), "cpd_error_noalloc")


def gather_probe(rows, gathers=1 << 25, dist=0, zipf_s=1.1, layout=0, seed=1, stream=None):
    """ms per launch of the 128-byte row gather probe over a table of `rows`
    rows: dist 0 uniform, 1 Zipf(zipf_s) over a permutation, 2 Zipf(zipf_s)
    with the popular rows first; layout 0 = 8 lanes x 16 B, 1 = 4
    lanes x 2 x 16 B per row."""
    ms = ctypes.c_float()
    check(lib().alto_hip_gather_probe(ctypes.c_uint64(rows), ctypes.c_uint64(gathers), ctypes.c_int(dist),
                                      ctypes.c_double(zipf_s), ctypes.c_int(layout), ctypes.c_uint64(seed),
                                      ctypes.byref(ms), _vp(stream)), "gather_probe")
    return ms.value


def generate(dims, nnz, seed, mode_dist, zipf_s=None, value_dist=0, p1=0.0, p2=1.0, block_frac=0.0,
             block=0, stream=None):
    """Synthetic tensor in device memory -> (cidx device arrays, vals device array)."""
    dims = np.ascontiguousarray(dims, dtype=np.uint64)
    N = len(dims)
    cidx = [DeviceArray((nnz,), np.uint64) for _ in range(N)]
    vals = DeviceArray((nnz,), np.float64)
    md = np.ascontiguousarray(mode_dist, dtype=np.int32)
    zs = np.ascontiguousarray(zipf_s if zipf_s is not None else [1.1] * N, dtype=np.float64)
    carr = (_vp * N)()
    for m in range(N):
        carr[m] = cidx[m].ptr
    check(lib().alto_hip_generate(ctypes.c_int(N), dims.ctypes.data_as(_u64p), ctypes.c_uint64(nnz),
                                  ctypes.c_uint64(seed), md.ctypes.data_as(_vp), zs.ctypes.data_as(_f64p),
                                  ctypes.c_int(value_dist), ctypes.c_double(p1), ctypes.c_double(p2),
                                  ctypes.c_double(block_frac), ctypes.c_uint64(block), carr, _vp(vals.ptr),
                                  _vp(stream)), "generate")
    return cidx, vals


class Timer:
    """HIP-event timer on a stream (milliseconds)."""

    def __init__(self, stream=None):
        self.stream = stream

    def __enter__(self):
        t = _vp()
        check(lib().alto_hip_timer_start(ctypes.byref(t), _vp(self.stream)), "timer_start")
        self.t = t.value
        return self

    def __exit__(self, *exc):
        ms = ctypes.c_float()
        check(lib().alto_hip_timer_stop(_vp(self.t), _vp(self.stream), ctypes.byref(ms)), "timer_stop")
        self.ms = ms.value
        return False
