#!/usr/bin/env python3
"""Benchmark: all-mode ALTO MTTKRP on one MI355X.

Workload (default): 4-mode tensor 500000 x 200000 x 50000 x 1000, 1e8
nonzeros, every mode's index Zipf(s=1.1) over a seeded permutation of the
mode, values lognormal(0, 1); rank 32, factors U(0,1); FP32 values, factors
and output.  A step is one MTTKRP per mode (4 launches) over the tensor
already resident in HBM, accumulating into the outputs as the reference's
BenchmarkAlto does (src/main.cpp:498-512).

Prints one JSON line.  `value` = algorithmic HBM bytes of the step
(sum over modes of nnz*(k+v) + sum_{m!=n} I_m*R*f + I_n*R*o) / step time.
Also reported: CP-ALS s/iter on the same tensor in FP32, BF16 storage and
FP64 with the exact fit of each final model (FP64, duplicates merged) and
their lambda/factor errors against FP64; the other workloads; and the
oracle (the CPU restatement of the reference's OpenMP algorithm) timed over
the whole tensor on the process's CPU share, pinned one thread per core.
"""
import argparse
import json
import os
import sys
import time

# The CPU legs (the oracle, the only OpenMP code here) run pinned, one thread
# per core, threads spinning between regions; libgomp reads these when it
# loads, and the CPU legs run after every GPU leg so that the binding of the
# main thread touches no GPU work.
os.environ.setdefault("OMP_PROC_BIND", "close")
os.environ.setdefault("OMP_PLACES", "cores")
os.environ.setdefault("OMP_WAIT_POLICY", "active")

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "alto-stream_amd"))

HBM_PEAK_GBS = 8000.0
# |fit_exact(prec) - fit_exact(FP64)| within which a reduced-precision CP-ALS
# counts as reaching the FP64 fit
EQUAL_FIT_TOL = 1e-5


def _json_scalar(x):
    """numpy scalars (np.bool_, np.float64, np.int64) as plain JSON values."""
    if isinstance(x, np.generic):
        return x.item()
    raise TypeError("not JSON serializable: %r" % type(x))


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload(name):
    if name == "zipf4":
        return dict(dims=[500000, 200000, 50000, 1000], nnz=100_000_000, seed=0xA1700001, rank=32,
                    mode_dist=[1, 1, 1, 1], zipf=[1.1] * 4, vdist=1, p1=0.0, p2=1.0,
                    desc="4-mode 500000x200000x50000x1000, 1e8 nnz, Zipf(1.1) indices, lognormal(0,1) values, rank 32")
    if name == "bf16_5mode":  # the 5-mode 128-bit-key workload (for --counters)
        return dict(dims=[1 << 20, 1 << 20, 1 << 18, 1 << 16, 1 << 12], nnz=50_000_000, seed=0xA1700003, rank=64,
                    mode_dist=[0] * 5, zipf=None, vdist=2, p1=0.0, p2=1.0, block_frac=0.3, block=256,
                    desc="5-mode 2^20x2^20x2^18x2^16x2^12, 5e7 nnz (30% in a 256^5 block), N(0,1) values, rank 64")
    if name == "zipf4_small":  # reduced shape for quick checks (not a bench line)
        return dict(dims=[50000, 20000, 5000, 100], nnz=2_000_000, seed=0xA1700001, rank=32,
                    mode_dist=[1, 1, 1, 1], zipf=[1.1] * 4, vdist=1, p1=0.0, p2=1.0,
                    desc="reduced shape of the 4-mode Zipf workload")
    raise SystemExit("unknown workload " + name)


def alg_bytes(dims, nnz, rank, k, v, f, o):
    tot = 0
    per = []
    for n in range(len(dims)):
        b = nnz * (k + v) + sum(dims[m] * rank * f for m in range(len(dims)) if m != n) + dims[n] * rank * o
        per.append(b)
        tot += b
    return tot, per


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="zipf4")
    ap.add_argument("--prec", default="f32", choices=["f32", "bf16", "f64"])
    ap.add_argument("--algo", type=int, default=0)
    ap.add_argument("--als-iters", type=int, default=10)
    ap.add_argument("--no-als", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=0,
                    help="CPU baseline over the first N nonzeros only (0: the whole tensor)")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the other workloads (plain CP-ALS, 5-mode BF16, CP-stream sample)")
    ap.add_argument("--stream-slices", type=int, default=1000, help="CP-stream: number of time slices (of 1000)")
    ap.add_argument("--batch-tensors", type=int, default=512, help="batched CP-ALS: number of tensors")
    ap.add_argument("--batch-ranks", default="8,16,32",
                    help="batched CP-ALS, with --leg: run only the tensors of these ranks (same tensors)")
    ap.add_argument("--probe", action="store_true",
                    help="random 128-byte row gather rate vs table size (the MTTKRP's miss-service bound)")
    ap.add_argument("--leg", choices=["plain_als", "bf16_5mode", "cp_stream", "batched_als"],
                    help="run one other-workload leg alone and print its JSON (for profiling)")
    ap.add_argument("--counters", action="store_true",
                    help="untimed pass for rocprofv3 --pmc: warm-up, a calibration copy, then each mode once in FP32, BF16 and FP64")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    cpu_threads()

    import alto_amd as A
    A.init(0)
    dev = A.device_name()
    if args.leg:
        legs = {"plain_als": lambda: leg_plain_als(A), "bf16_5mode": lambda: leg_bf16_5mode(A),
                "cp_stream": lambda: leg_cp_stream(A, args.stream_slices),
                "batched_als": lambda: leg_batched_als(A, args.batch_tensors,
                                                       [int(r) for r in args.batch_ranks.split(",")])}
        print(json.dumps(legs[args.leg](), default=_json_scalar), flush=True)
        return
    if args.probe:
        # 128-byte row gathers: uniform rows over growing tables (L2, MALL, HBM),
        # and Zipf(1.1) rows as the workload's modes draw them, in both lane layouts
        g = 1 << 25
        for dist in (0, 1):
            for rows in (16384, 65536, 200000, 500000, 1 << 20, 1 << 24):
                for layout in (0, 1):
                    ms = A.gather_probe(rows, g, dist=dist, zipf_s=1.1, layout=layout)
                    print(json.dumps({"rows": rows, "table_MiB": round(rows * 128 / 2 ** 20, 1),
                                      "dist": "zipf1.1" if dist else "uniform",
                                      "layout": "8x16B" if layout == 0 else "4x2x16B", "ms": round(ms, 4),
                                      "G_rows_per_s": round(g / ms / 1e6, 2)}), flush=True)
        return
    W = workload(args.workload)
    dims, nnz, R = W["dims"], W["nnz"], W["rank"]
    N = len(dims)
    prec = {"f32": A.PREC_F32, "bf16": A.PREC_BF16, "f64": A.PREC_F64}[args.prec]
    log("device:", dev)
    t0 = time.time()
    cidx, vals = A.generate(dims, nnz, W["seed"], W["mode_dist"], W["zipf"], W["vdist"], W["p1"], W["p2"],
                            block_frac=W.get("block_frac", 0.0), block=W.get("block", 0))
    A.synchronize()
    log("generated %d nnz in %.1f s" % (nnz, time.time() - t0))
    t0 = time.time()
    at = A.create_alto(dims, cidx, vals, nprtn=256, on_device=True)
    A.synchronize()
    t_create = time.time() - t0
    log("ALTO creation time: %.3f s (%d bits, %d-bit key)" % (t_create, at.total_bits, 64 * at.kw))
    for c in cidx:
        c.close()

    rng = np.random.default_rng(W["seed"])
    F64 = [rng.random((d, R)) for d in dims]
    if args.counters:
        # Warm-up call of every (precision, mode): schedules and launch tuning.
        # Then the calibration copy, k_f64_to_f32 over 16,000,000 elements
        # (128,000,000 B read, 64,000,000 B written), which also marks where the
        # measured calls start: one per precision (FP32, BF16, FP64) and mode.
        sets = []
        for p in (A.PREC_F32, A.PREC_BF16, A.PREC_F64):
            dF = [A.to_prec(f, p) for f in F64]
            os_ = [A.DeviceArray.zeros((d, R), np.float64 if p == A.PREC_F64 else np.float32) for d in dims]
            for n in range(N):
                A.mttkrp_alto_par(n, dF, R, at, os_[n], prec=p, out_f64=False, algo=args.algo)
            sets.append((p, dF, os_))
        A.synchronize()
        src = A.DeviceArray.zeros(16_000_000, np.float64)
        dst = A.DeviceArray.zeros(16_000_000, np.float32)
        A.convert(dst, A.PREC_F32, src, A.PREC_F64, 16_000_000)
        A.synchronize()
        for p, dF, os_ in sets:
            for n in range(N):
                A.mttkrp_alto_par(n, dF, R, at, os_[n], prec=p, out_f64=False, algo=args.algo)
        A.synchronize()
        print(json.dumps({"counters_pass": True}), flush=True)
        return
    devF = [A.to_prec(f, prec) for f in F64]
    odt = np.float64 if prec == A.PREC_F64 else np.float32
    outs = [A.DeviceArray.zeros((d, R), odt) for d in dims]

    def step(timers=None):
        for n in range(N):
            if timers is not None:
                with A.Timer() as tm:
                    A.mttkrp_alto_par(n, devF, R, at, outs[n], prec=prec, out_f64=False, algo=args.algo)
                timers[n].append(tm.ms)
            else:
                A.mttkrp_alto_par(n, devF, R, at, outs[n], prec=prec, out_f64=False, algo=args.algo)

    t0 = time.time()
    step()  # builds the per-mode schedules (and hot-row encodings), tunes the launches
    A.synchronize()
    t_prep = time.time() - t0
    for _ in range(args.warmup - 1):
        step()
    A.synchronize()
    log("warmup (incl. per-mode schedules) %.2f s, of which the first step %.2f s" % (time.time() - t0, t_prep))
    with A.Timer() as tm:
        for _ in range(args.steps):
            step()
    ms_step = tm.ms / args.steps
    per_mode = [[] for _ in range(N)]
    for _ in range(3):
        step(per_mode)
    mode_ms = [float(np.median(x)) for x in per_mode]

    k = 8 * at.kw
    v = 8 if prec == A.PREC_F64 else 4
    f = {A.PREC_F64: 8, A.PREC_F32: 4, A.PREC_BF16: 2}[prec]
    o = 8 if prec == A.PREC_F64 else 4
    tot, per = alg_bytes(dims, nnz, R, k, v, f, o)
    achieved = tot / (ms_step * 1e-3) / 1e9
    log("all-mode MTTKRP: %.3f ms/step (%s ms per mode), %.1f GB/s algorithmic" %
        (ms_step, ["%.3f" % x for x in mode_ms], achieved))

    result = {
        "metric": "all-mode MTTKRP ms & HBM GB/s; CP-ALS s/iter, FP32/BF16 vs FP64 at equal fit",
        "value": round(achieved, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"f32": "f32", "bf16": "bf16-storage/f32", "f64": "f64"}[args.prec],
        "data": "synthetic (device generator, seed %#x)" % W["seed"],
        "config": {"workload": W["desc"], "nnz": nnz, "dims": dims, "rank": R, "key_bits": 64 * at.kw,
                   "prec": args.prec, "mode_ms": [round(x, 4) for x in mode_ms],
                   "alto_create_s": round(t_create, 3),
                   "format_prep_s": round(t_prep, 3),
                   "format_prep_note": "the first all-mode step, untimed: builds every mode's output-ordered "
                                       "schedule (a sorted, duplicate-merged copy of the nonzeros, ~1 GB per "
                                       "mode here), the hot-row encodings, and tunes the launches",
                   "device": dev},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic(args.workload, args.prec),
                     "alg_bytes_per_launch": int(tot / N)},
    }
    for d in outs:
        d.close()

    # host copy of the ALTO-ordered nonzeros: the merged-duplicate norm and the
    # CPU legs read it (the tensor itself is released before the other legs)
    t0 = time.time()
    hkeys, hvals = host_nonzeros(A, at, at.nnz if args.cpu_sample <= 0 else min(args.cpu_sample, at.nnz))
    normX2_merged = merged_norm2(hkeys, hvals) if hkeys.shape[0] == at.nnz else None
    log("host copy of %d nonzeros and merged norm in %.1f s" % (hkeys.shape[0], time.time() - t0))

    if not args.no_als:
        als = {}
        models = {}
        for name, p in (("f32", A.PREC_F32), ("bf16", A.PREC_BF16), ("f64", A.PREC_F64)):
            A.cpd_alto(at, F64, R, 2, eps=0.0, prec=p)  # warm-up (one-time runtime set-up)
            t0 = time.time()
            fs, lam, fits, fe, tim = A.cpd_alto(at, F64, R, args.als_iters, eps=1e-30, prec=p)
            models[name] = (fs, lam)
            als[name] = dict(s_per_iter=round(float(tim[7]), 5), fit_reference=float(fits[-1]),
                             fit_exact_unmerged_norm=float(fe[-1]), iters=len(fits))
            if normX2_merged is not None:
                als[name]["fit_exact"] = exact_fit(A, at, fs, lam, normX2_merged)
            log("CP-ALS %s: %.4f s/iter, reference fit %.8f, exact fit %s (wall %.1f s)" %
                (name, tim[7], fits[-1], als[name].get("fit_exact"), time.time() - t0))
        als["fit_note"] = ("fit_exact: 1 - ||X - M|| / ||X|| in FP64 over the tensor with duplicate coordinates "
                           "merged (%.1f%% of the stored nonzeros repeat a coordinate); fit_reference: the "
                           "reference's printed fit, whose ||X||^2 is the sum of squared stored values "
                           "(cpd.hpp:80-83) and so is not a fit on this tensor" %
                           (100.0 * (1.0 - distinct_count(hkeys) / hkeys.shape[0]) if normX2_merged is not None
                            else float("nan")))
        for name in ("f32", "bf16"):
            fs, lam = models[name]
            fs64, lam64 = models["f64"]
            als["speedup_%s_vs_f64" % name] = round(als["f64"]["s_per_iter"] / als[name]["s_per_iter"], 3)
            if normX2_merged is not None:
                als["dfit_exact_%s" % name] = abs(als[name]["fit_exact"] - als["f64"]["fit_exact"])
                als["%s_at_equal_fit" % name] = bool(als["dfit_exact_%s" % name] <= EQUAL_FIT_TOL)
            als["lambda_rel_err_%s" % name] = float(np.linalg.norm(lam - lam64) / np.linalg.norm(lam64))
            als["factor_rel_err_%s" % name] = float(max(np.linalg.norm(a - b) / np.linalg.norm(b)
                                                        for a, b in zip(fs, fs64)))
            als["dfit_reference_%s" % name] = abs(als[name]["fit_reference"] - als["f64"]["fit_reference"])
        als["equal_fit_tolerance"] = EQUAL_FIT_TOL
        result["cp_als"] = als
    kw = at.kw
    at.close()
    if not args.no_extra:
        result["other_workloads"] = {"plain_cp_als_1000x800x600": leg_plain_als(A),
                                     "bf16_5mode_5e7": leg_bf16_5mode(A),
                                     "cp_stream_200000x100000": leg_cp_stream(A, args.stream_slices),
                                     "batched_cp_als_512": leg_batched_als(A, args.batch_tensors)}
    # CPU legs last: the pinned OpenMP threads must not share cores with GPU work
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(hkeys, hvals, dims, R, F64, kw, nnz)
        if not args.no_extra:
            result["other_workloads"]["plain_cp_als_1000x800x600"].update(plain_als_cpu())
    print(json.dumps(result, default=_json_scalar), flush=True)


def host_nonzeros(A, at, count):
    """The first `count` nonzeros of the ALTO order (keys, FP64 values) on the host."""
    import ctypes
    keys_d, vals_d = ctypes.c_void_p(), ctypes.c_void_p()
    A.check(A.lib().alto_hip_tensor_device_arrays(ctypes.c_void_p(at.handle), ctypes.byref(keys_d),
                                                   ctypes.byref(vals_d)), "device_arrays")
    keys = np.empty((count, at.kw), dtype=np.uint64)
    vals = np.empty(count, dtype=np.float64)
    A.check(A.lib().alto_hip_memcpy_d2h(keys.ctypes.data_as(ctypes.c_void_p), keys_d,
                                        ctypes.c_size_t(keys.nbytes), None), "d2h")
    A.check(A.lib().alto_hip_memcpy_d2h(vals.ctypes.data_as(ctypes.c_void_p), vals_d,
                                        ctypes.c_size_t(vals.nbytes), None), "d2h")
    return keys, vals


def _run_starts(keys):
    """Start of each run of equal keys (the ALTO order is sorted by key, so a
    coordinate's duplicates are adjacent)."""
    diff = np.any(keys[1:] != keys[:-1], axis=1)
    return np.concatenate(([0], np.flatnonzero(diff) + 1))


def distinct_count(keys):
    return int(_run_starts(keys).shape[0])


def merged_norm2(keys, vals):
    """||X||^2 of the tensor whose duplicate coordinates are summed."""
    sums = np.add.reduceat(vals, _run_starts(keys))
    return float(np.dot(sums, sums))


def exact_fit(A, at, fs, lam, normX2):
    """1 - ||X - M|| / ||X|| in FP64 for the model M = sum_r lam_r u_1r o ... o u_Nr:
    ||M||^2 from the Gram matrices, <X, M> from an FP64 MTTKRP of the last
    mode with the model's factors (the MTTKRP sums every stored nonzero, so
    duplicates count as their merged sum)."""
    N = len(fs)
    R = fs[0].shape[1]
    G = np.ones((R, R))
    for f in fs:
        G *= f.T @ f
    nM2 = float(lam @ G @ lam)
    dF = [A.to_prec(f, A.PREC_F64) for f in fs]
    o = A.DeviceArray.zeros((fs[-1].shape[0], R), np.float64)
    A.mttkrp_alto_par(N - 1, dF, R, at, o, prec=A.PREC_F64)
    K = o.numpy()
    for d in dF:
        d.close()
    o.close()
    ip = float(np.sum((K * fs[-1]) @ lam))
    res2 = max(normX2 - 2.0 * ip + nM2, 0.0)
    return float(1.0 - np.sqrt(res2) / np.sqrt(normX2))


def measured_traffic(workload, prec):
    """HBM bytes per MTTKRP launch (FETCH_SIZE + WRITE_SIZE, scaled by a copy
    kernel of known size) from the committed rocprofv3 summary of this
    workload, or None when there is none."""
    if workload != "zipf4":
        return None
    path = os.path.join(ROOT, "profiles", "mttkrp_zipf4mode_rank32.md")
    key = "traffic_bytes_per_launch_%s:" % {"f32": "fp32", "bf16": "bf16", "f64": "fp64"}[prec]
    try:
        for line in open(path):
            if line.startswith(key):
                return int(line.split(":")[1])
    except OSError:
        pass
    return None


def cpu_info():
    """CPU model, the host's physical cores and logical CPUs, and this
    process's CPU share: its affinity set, bounded by a cgroup quota."""
    model, cores = None, set()
    phys = core = None
    try:
        for line in open("/proc/cpuinfo"):
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k == "model name" and model is None:
                model = v
            elif k == "physical id":
                phys = v
            elif k == "core id":
                core = v
            elif not k:
                if core is not None:
                    cores.add((phys, core))
                phys = core = None
        if core is not None:
            cores.add((phys, core))
    except OSError:
        pass
    logical = os.cpu_count() or 1
    affinity = len(os.sched_getaffinity(0))
    quota = None
    try:
        q, per = open("/sys/fs/cgroup/cpu.max").read().split()
        if q != "max":
            quota = int(q) / int(per)
    except (OSError, ValueError):
        pass
    share = affinity if quota is None else max(1, min(affinity, int(quota)))
    return dict(cpu_model=model, host_physical_cores=len(cores) or logical, host_logical_cpus=logical,
                process_affinity_cpus=affinity, cgroup_cpu_quota=quota, process_cpu_share=share)


# at most 16 threads: the CPU share allotted to one GPU's job on the measuring host
CPU_THREADS_CAP = 16


_CPU = None


def cpu_threads():
    """Threads for the CPU legs and the CPU description, read once: after the
    first pinned OpenMP region the main thread's affinity is one core."""
    global _CPU
    if _CPU is None:
        info = cpu_info()
        _CPU = (min(info["process_cpu_share"], CPU_THREADS_CAP), info)
    return _CPU


def cpu_baseline(keys, vals, dims, R, F64, kw, nnz_total):
    """The oracle's OpenMP restatement of the reference MTTKRP (da_mem_pull /
    atomic dispatch, FP64) over the ALTO-ordered nonzeros copied from the GPU
    tensor: one untimed all-mode pass, then one timed."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    threads, info = cpu_threads()
    S = keys.shape[0]
    O.set_threads(threads)

    class T:  # oracle-side ALTO tensor over the host copy
        pass
    t = T()
    t.dims = np.array(dims, dtype=np.uint64)
    t.nmode = len(dims)
    t.masks, _ = O.alto_masks(dims, kw)
    t.kw, t.keys, t.vals, t.nnz, t.nprtn = kw, keys, vals, S, threads
    t0 = time.time()
    t.prtn_ptr, t.intervals = O.prtn_alto(t.dims, keys, t.masks, threads)
    t_prtn = time.time() - t0
    outs = [np.zeros((d, R)) for d in dims]
    for n in range(len(dims)):
        O.AltoTensor.mttkrp(t, F64, n, out=outs[n])
    t0 = time.time()
    for n in range(len(dims)):
        O.AltoTensor.mttkrp(t, F64, n, out=outs[n])
    dt = time.time() - t0
    tot, _ = alg_bytes(dims, S, R, 8 * kw, 8, 8, 8)
    log("cpu baseline: %.2f s for one all-mode FP64 MTTKRP over %d nnz on %d threads (%s)" %
        (dt, S, threads, info["cpu_model"]))
    full = S == nnz_total
    out = {"value": round(tot / dt / 1e9, 3), "unit": "GB/s", "cores": threads, "kind": "port",
           "sample": "one all-mode FP64 MTTKRP (oracle OpenMP restatement of the reference dispatch) over %s "
                     "(%d nonzeros in ALTO order), after one untimed pass: %.3f s per all-mode step" %
                     ("the whole tensor" if full else "a prefix of the tensor", S, dt),
           "ms_per_step": round(dt * 1e3, 1), "partition_s": round(t_prtn, 3),
           "omp": {k: os.environ.get(k) for k in ("OMP_PROC_BIND", "OMP_PLACES", "OMP_WAIT_POLICY")}}
    out.update(info)
    return out


def plain_als_data():
    """3-mode 1000x800x600, 1e6 distinct uniform coordinates, U(0,1) values,
    rank 16: coordinates, values and the initial model."""
    dims, nnz, R = [1000, 800, 600], 1_000_000, 16
    rng = np.random.default_rng(0xA1700000)
    total = dims[0] * dims[1] * dims[2]
    lin = np.unique(rng.integers(0, total, size=int(nnz * 1.05)))
    lin = rng.permutation(lin)[:nnz]
    cidx = [(lin // (dims[1] * dims[2])).astype(np.uint64), ((lin // dims[2]) % dims[1]).astype(np.uint64),
            (lin % dims[2]).astype(np.uint64)]
    vals = rng.random(nnz)
    F = [rng.random((d, R)) for d in dims]
    return dims, cidx, vals, F, R


def leg_plain_als(A):
    """Plain CP-ALS as the reference runs it: 3-mode 1000x800x600, 1e6
    distinct uniform coordinates, U(0,1) values, rank 16, 20 iterations;
    plus 100 mode-0 MTTKRPs (the reference's -p benchmark)."""
    log("plain CP-ALS, 1000x800x600, 1e6 nnz, rank 16")
    dims, cidx, vals, F, R = plain_als_data()
    at = A.create_alto(dims, cidx, vals, nprtn=256)
    out = {"workload": "3-mode 1000x800x600, 1e6 distinct nnz, U(0,1), rank 16"}
    for name, p in (("f64", A.PREC_F64), ("f32", A.PREC_F32)):
        # warm-up: the first graph instantiation in a process pays one-time
        # runtime set-up, which the per-iteration figure should not carry
        A.cpd_alto(at, F, R, 2, eps=0.0, prec=p)
        _, _, fits, _, tim = A.cpd_alto(at, F, R, 20, eps=1e-30, prec=p)
        out["cp_als_%s_s_per_iter" % name] = round(float(tim[7]), 6)
        out["cp_als_%s_fit" % name] = float(fits[-1])
    dF = [A.to_prec(f, A.PREC_F64) for f in F]
    o = A.DeviceArray.zeros((dims[0], R), np.float64)
    A.mttkrp_alto_par(0, dF, R, at, o, prec=A.PREC_F64)
    A.synchronize()
    with A.Timer() as tm:
        for _ in range(100):
            A.mttkrp_alto_par(0, dF, R, at, o, prec=A.PREC_F64)
    out["mttkrp_mode0_f64_ms"] = round(tm.ms / 100, 5)
    for d in dF:
        d.close()
    o.close()
    at.close()
    log("plain CP-ALS: f64 %.5f s/iter, f32 %.5f s/iter, MTTKRP mode 0 %.4f ms" %
        (out["cp_als_f64_s_per_iter"], out["cp_als_f32_s_per_iter"], out["mttkrp_mode0_f64_ms"]))
    return out


def plain_als_cpu():
    """The oracle's CP-ALS on the plain workload, 5 iterations, on the CPU share."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    dims, cidx, vals, F, R = plain_als_data()
    threads, info = cpu_threads()
    O.set_threads(threads)
    ref = O.AltoTensor(dims, cidx, vals, nprtn=threads)
    t0 = time.time()
    _, _, rfits, _ = ref.cpd_als(F, R, 5, eps=1e-30)
    dt = (time.time() - t0) / 5
    log("plain CP-ALS on the CPU: %.4f s/iter on %d threads" % (dt, threads))
    return {"cpu_port_cp_als_s_per_iter": round(dt, 5), "cpu_cores": threads, "cpu_fit_after_5": float(rfits[-1])}


def leg_bf16_5mode(A):
    """5-mode 2^20 x 2^20 x 2^18 x 2^16 x 2^12, 5e7 nnz (70% uniform, 30% in a
    256^5 sub-block), N(0,1) values, rank 64, 128-bit keys: all-mode MTTKRP
    with BF16-stored factors (FP32 accumulate) and FP64, and 10 CP-ALS
    iterations BF16 vs FP64 with the fit difference."""
    dims, nnz, R = [1 << 20, 1 << 20, 1 << 18, 1 << 16, 1 << 12], 50_000_000, 64
    seed = 0xA1700003
    log("5-mode 2^20x2^20x2^18x2^16x2^12, 5e7 nnz, rank 64, BF16 / FP64")
    t0 = time.time()
    cidx, vals = A.generate(dims, nnz, seed, [0] * 5, None, 2, 0.0, 1.0, block_frac=0.3, block=256)
    at = A.create_alto(dims, cidx, vals, nprtn=256, on_device=True)
    A.synchronize()
    t_create = time.time() - t0
    for c in cidx:
        c.close()
    vals.close()
    rng = np.random.default_rng(seed)
    F = [rng.random((d, R)) for d in dims]
    out = {"workload": "5-mode 2^20x2^20x2^18x2^16x2^12, 5e7 nnz (30%% in a 256^5 block), N(0,1), rank 64, "
                       "%d-bit keys" % (64 * at.kw), "alto_create_s": round(t_create, 3)}
    single = {}  # one MTTKRP per mode from zero, for the error against FP64
    for name, p in (("bf16", A.PREC_BF16), ("f64", A.PREC_F64)):
        dF = [A.to_prec(f, p) for f in F]
        odt = np.float64 if p == A.PREC_F64 else np.float32
        outs = [A.DeviceArray.zeros((d, R), odt) for d in dims]
        for n in range(5):
            A.mttkrp_alto_par(n, dF, R, at, outs[n], prec=p, out_f64=False)
        A.synchronize()
        with A.Timer() as tm:
            for _ in range(3):
                for n in range(5):
                    A.mttkrp_alto_par(n, dF, R, at, outs[n], prec=p, out_f64=False)
        ms = tm.ms / 3
        k, v = 8 * at.kw, 8 if p == A.PREC_F64 else 4
        f, o = (8, 8) if p == A.PREC_F64 else (2, 4)
        tot, _ = alg_bytes(dims, nnz, R, k, v, f, o)
        out["mttkrp_%s_ms_all_mode" % name] = round(ms, 4)
        out["mttkrp_%s_GBs_algorithmic" % name] = round(tot / (ms * 1e-3) / 1e9, 1)
        single[name] = []
        for n in range(5):
            outs[n].zero_()
            A.mttkrp_alto_par(n, dF, R, at, outs[n], prec=p, out_f64=False)
            single[name].append(outs[n].numpy().astype(np.float64))
        for d in dF:
            d.close()
        for d in outs:
            d.close()
    # error of the BF16-stored MTTKRP against the FP64 one on the same FP64
    # model (the BF16 rounding of the factors included), per mode
    errs = [float(np.linalg.norm(b - f) / np.linalg.norm(f)) for b, f in zip(single["bf16"], single["f64"])]
    out["mttkrp_bf16_rel_err_vs_f64"] = max(errs)
    out["mttkrp_bf16_rel_err_vs_f64_by_mode"] = errs
    del single
    fits, models = {}, {}
    for name, p in (("bf16", A.PREC_BF16), ("f64", A.PREC_F64)):
        A.cpd_alto(at, F, R, 2, eps=0.0, prec=p)  # warm-up (one-time runtime set-up)
        fs, lam, ft, fe, tim = A.cpd_alto(at, F, R, 10, eps=1e-30, prec=p, fit_mode=A.FIT_EXACT)
        out["cp_als_%s_s_per_iter" % name] = round(float(tim[7]), 5)
        fits[name] = float(ft[-1])
        models[name] = (fs, lam)
        out["cp_als_%s_fit" % name] = fits[name]
    out["cp_als_speedup_bf16_vs_f64"] = round(out["cp_als_f64_s_per_iter"] / out["cp_als_bf16_s_per_iter"], 3)
    out["fit_abs_diff_bf16_vs_f64"] = abs(fits["bf16"] - fits["f64"])
    (fb, lb), (f64m, l64) = models["bf16"], models["f64"]
    out["cp_als_lambda_rel_err_bf16"] = float(np.linalg.norm(lb - l64) / np.linalg.norm(l64))
    out["cp_als_factor_rel_err_bf16"] = float(max(np.linalg.norm(a - b) / np.linalg.norm(b)
                                                  for a, b in zip(fb, f64m)))
    out["fit_note"] = ("cp_als_*_fit: the relative residual ||X - M|| / ||X|| (1: the model explains nothing); "
                       "on N(0,1) values a rank-64 model explains almost nothing, so the fit difference does "
                       "not measure accuracy here: mttkrp_bf16_rel_err_vs_f64 and the model errors do")
    del models
    at.close()
    log("5-mode: MTTKRP bf16 %.3f ms, f64 %.3f ms all-mode; CP-ALS bf16 %.4f s/iter, f64 %.4f s/iter, |dfit| %.2e" %
        (out["mttkrp_bf16_ms_all_mode"], out["mttkrp_f64_ms_all_mode"], out["cp_als_bf16_s_per_iter"],
         out["cp_als_f64_s_per_iter"], out["fit_abs_diff_bf16_vs_f64"]))
    return out


def leg_cp_stream(A, slices):
    """CP-stream with non-negative ADMM on the first `slices` time slices of
    the 1000-slice workload: slices of 200000 x 100000, per-slice nnz
    lognormal(ln 2e5, 1) clipped to [2e4, 2e6], Zipf(1.05) indices,
    Poisson(3)+1 values, rank 16, at most 10 outer iterations per slice;
    FP32 MTTKRP, FP64 ADMM/solve."""
    seed = 0xA1700004
    log("CP-stream 200000x100000, %d slices" % slices)
    rng = np.random.default_rng(seed)
    sizes = np.clip(np.round(rng.lognormal(np.log(2e5), 1.0, 1000)), 2e4, 2e6).astype(np.int64)[:slices]
    nnz = int(sizes.sum())
    dims = [slices, 200000, 100000]
    cidx, vals = A.generate([1, 200000, 100000], nnz, seed, [0, 1, 1], [1.0, 1.05, 1.05], 3, 3.0, 0.0)
    hc = [c.numpy() for c in cidx]
    hv = vals.numpy()
    for c in cidx:
        c.close()
    vals.close()
    hc[0] = np.repeat(np.arange(slices, dtype=np.uint64), sizes)
    R = 16
    init = [rng.random((d, R)) for d in dims]
    # warm-up on the first two slices: one-time runtime set-up (code objects,
    # the first graph instantiation) stays out of the timed slice loops
    A.cpstream(dims, hc, hv, 0, R, 10, 1e-6, True, init, max_slices=2, prec=A.PREC_F32)
    A.cpstream(dims, hc, hv, 0, R, 10, 1e-6, False, init, max_slices=2, prec=A.PREC_F32)
    A.spcpstream(dims, hc, hv, 0, R, 10, 1e-6, init, max_slices=2, prec=A.PREC_F32)
    t0 = time.time()
    res = A.cpstream(dims, hc, hv, 0, R, 10, 1e-6, True, init, prec=A.PREC_F32)
    wall = time.time() - t0
    tim = res["timings"]
    out = {"workload": "CP-stream, first %d of 1000 slices of 200000x100000 (%d nnz), Zipf(1.05), Poisson(3)+1, "
                       "rank 16, nonneg ADMM, <= 10 outer iterations" % (slices, nnz),
           "slices_per_s": round(slices / float(tim[0]), 2), "slice_loop_s": round(float(tim[0]), 3),
           "wall_s_incl_upload_and_stream_prep": round(wall, 3),
           "outer_iters_mean": round(float(np.mean(res["outer_iters"])), 2),
           "admm_iters_mean": round(float(np.mean(res["admm_iters"])), 2)}
    log("CP-stream: %d slices in %.3f s (%.2f slices/s), wall %.1f s" % (slices, tim[0], slices / tim[0], wall))
    # spCP-stream on the same stream (the reference's SPCPSTREAM_ALTO: row-sparse
    # updates, no constraint), beside CP-stream without the constraint
    sp = A.spcpstream(dims, hc, hv, 0, R, 10, 1e-6, init, prec=A.PREC_F32)
    un = A.cpstream(dims, hc, hv, 0, R, 10, 1e-6, False, init, prec=A.PREC_F32)
    ts, tu = float(sp["timings"][0]), float(un["timings"][0])
    # The closing cpd_error takes ||X||^2 as the sum of squared stored values
    # (streaming_cpd.cpp:368-464); this stream repeats coordinates (Zipf
    # indices), so that sum falls short of the tensor's norm and the
    # residual's square can go negative: NaN, as in the reference.  Reported
    # as null then.
    def fin(x):
        return float(x) if np.isfinite(x) else None
    out["final_error"] = fin(res["final_error"])
    out["spcpstream"] = {"slices_per_s": round(slices / ts, 2), "slice_loop_s": round(ts, 3),
                         "outer_iters_mean": round(float(np.mean(sp["outer_iters"])), 2),
                         "final_error": fin(sp["final_error"])}
    out["cpstream_unconstrained"] = {"slices_per_s": round(slices / tu, 2), "slice_loop_s": round(tu, 3),
                                     "outer_iters_mean": round(float(np.mean(un["outer_iters"])), 2),
                                     "final_error": fin(un["final_error"])}
    log("CP-stream: spCP-stream %.2f slices/s (%.2f outer iterations), CP-stream unconstrained %.2f slices/s (%.2f)" %
        (slices / ts, np.mean(sp["outer_iters"]), slices / tu, np.mean(un["outer_iters"])))
    try:
        exact = stream_exact_errors(A, dims, hc, hv, {"cpstream_nonneg": res, "spcpstream": sp,
                                                      "cpstream_unconstrained": un})
        out["final_error_exact"] = exact["cpstream_nonneg"]
        out["spcpstream"]["final_error_exact"] = exact["spcpstream"]
        out["cpstream_unconstrained"]["final_error_exact"] = exact["cpstream_unconstrained"]
        out["final_error_exact_note"] = exact["note"]
    except Exception as e:  # reported, not fatal to the bench line
        out["final_error_exact"] = None
        out["final_error_exact_note"] = "not computed: %s" % e
    return out


def stream_exact_errors(A, dims, hc, hv, models):
    """||X - M|| / ||X|| of each streamed model over the processed stream with
    duplicate coordinates merged (||X||^2 = sum over distinct coordinates of
    the squared summed values), in the relabelled index space the closing
    cpd_error uses (the time factor = the time rows).  <X, M> and ||M||^2 come
    from alto_hip_cpd_error, which is linear in the values; only ||X||^2 needs
    the merge: equal coordinates have equal ALTO keys, which sort adjacent."""
    sc, sv, _, _ = A.stream_prepare(dims, hc, hv, 0)
    res0 = models["cpstream_nonneg"]
    rel_dims = [len(res0["time_rows"])] + [int(d) for d in res0["dims"][1:]]
    for m in range(1, len(dims)):
        rel_dims[m] = max(rel_dims[m], int(sc[m].max()) + 1)
    dc = [A.DeviceArray.from_numpy(c) for c in sc]
    dv = A.DeviceArray.from_numpy(sv)
    at = A.create_alto(rel_dims, dc, dv, nprtn=1, on_device=True)
    keys, vals = at.keys(), at.vals()
    at.close()
    diff = keys[1:, 0] != keys[:-1, 0] if keys.shape[1] == 1 else np.any(keys[1:] != keys[:-1], axis=1)
    heads = np.concatenate([[0], np.flatnonzero(diff) + 1])
    merged = np.add.reduceat(vals, heads)
    nx = float(np.dot(merged, merged))
    del keys, vals, diff, heads
    out = {"note": ("relative residual over the processed stream with duplicate coordinates merged (%d stored, "
                    "%d distinct); final_error keeps the reference's ||X||^2 = sum of squared stored values, "
                    "NaN here" % (len(sv), len(merged)))}
    for name, r in models.items():
        fs = [np.asarray(r["time_rows"])] + [np.asarray(f) for f in r["factors"][1:]]
        fs = [np.vstack([f, np.zeros((rel_dims[m] - f.shape[0], f.shape[1]))]) if f.shape[0] < rel_dims[m] else f
              for m, f in enumerate(fs)]
        df = [A.DeviceArray.from_numpy(np.ascontiguousarray(f, dtype=np.float64)) for f in fs]
        _, (_, nm, inner) = A.cpd_error(rel_dims, dc, dv, df, on_device=True)
        for f in df:
            f.close()
        r2 = nx + nm - 2.0 * inner
        out[name] = float(np.sqrt(r2) / np.sqrt(nx)) if r2 >= 0 else None
    for c in dc:
        c.close()
    dv.close()
    return out


def leg_batched_als(A, ntensors, only_ranks=(8, 16, 32)):
    """Batched CP-ALS of uneven instances: 3-mode tensors, each mode size
    uniform in [500, 50000], nnz log-uniform in [1e4, 5e6], uniform
    coordinates, U(0,1) values; ranks 8/16/32 round robin; 25 iterations
    each (epsilon 0: every iteration runs, as in the reference with any
    epsilon <= 0); FP32 MTTKRP, FP64 solve.  Timed with the initial models
    resident in HBM (alto_hip_cpd_als_batch_dev, after a warm-up call), and
    with host models (alto_hip_cpd_als_batch: every model crosses PCIe in and
    out, converted on host threads) on 8 streams."""
    seed = 0xA1700005
    log("batched CP-ALS, %d tensors" % ntensors)
    rng = np.random.default_rng(seed)
    ats, Fs, ranks, total = [], [], [], 0
    for k in range(ntensors):
        dims = [int(x) for x in rng.integers(500, 50001, size=3)]
        nnz = int(round(10 ** rng.uniform(4, np.log10(5e6))))
        R = (8, 16, 32)[k % 3]
        F = [rng.random((d, R)) for d in dims]  # drawn for every tensor: the same models whatever is run
        if R not in only_ranks:
            continue
        cidx, vals = A.generate(dims, nnz, seed + k, [0, 0, 0], None, 0)
        ats.append(A.create_alto(dims, cidx, vals, nprtn=1, on_device=True))
        A.synchronize()
        for c in cidx:
            c.close()
        vals.close()
        Fs.append(F)
        ranks.append(R)
        total += nnz
    out = {"workload": "%d 3-mode tensors, modes U[500, 50000], nnz log-uniform [1e4, 5e6] (%d total), "
                       "ranks %s, 25 iterations, FP32 MTTKRP + FP64 solve"
                       % (len(ats), total, "/".join(str(r) for r in sorted(set(ranks))))}
    def dev_models():
        return [[A.DeviceArray.from_numpy(f) for f in fk] for fk in Fs]

    dF = dev_models()
    A.cpd_alto_batch_dev(ats, dF, ranks, 25, eps=0.0, prec=A.PREC_F32, nstreams=8)  # warm-up
    for fk in dF:
        for f in fk:
            f.close()
    dF = dev_models()
    A.synchronize()
    _, fits_d, wall = A.cpd_alto_batch_dev(ats, dF, ranks, 25, eps=0.0, prec=A.PREC_F32, nstreams=8)
    for fk in dF:
        for f in fk:
            f.close()
    out["wall_s"] = round(wall, 3)
    out["tensors_per_s"] = round(ntensors / wall, 2)
    out["nnz_iterations_per_s"] = round(total * 25 / wall / 1e9, 3)
    _, _, fits_h, wall_h = A.cpd_alto_batch(ats, Fs, ranks, 25, eps=0.0, prec=A.PREC_F32, nstreams=8)
    out["wall_s_host_models"] = round(wall_h, 3)
    out["tensors_per_s_host_models"] = round(ntensors / wall_h, 2)
    out["same_fits_host_and_device_models"] = bool(all(np.array_equal(a, b) for a, b in zip(fits_d, fits_h)))
    out["unit_nnz_iterations"] = "1e9 nonzero-iterations/s (nnz x 25 ALS iterations / wall)"
    out["note"] = ("wall_s: models resident in HBM before and after (the batch's inputs and outputs); "
                   "wall_s_host_models: host FP64 models in and out over PCIe (%.1f GB each way)"
                   % (sum(f.nbytes for fk in Fs for f in fk) / 1e9))
    for a in ats:
        a.close()
    log("batched CP-ALS: %d tensors (%d nnz): %.3f s with models in HBM, %.3f s with host models" %
        (len(ats), total, wall, wall_h))
    return out

if __name__ == "__main__":
    main()
