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
Also reported: CP-ALS s/iter on the same tensor for the FP32 and the plain
FP64 GPU paths with their fits, and the oracle (the CPU restatement of the
reference's OpenMP algorithm) timed on the host cores on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "alto-stream_amd"))

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload(name):
    if name == "c2":
        return dict(dims=[500000, 200000, 50000, 1000], nnz=100_000_000, seed=0xA1700001, rank=32,
                    mode_dist=[1, 1, 1, 1], zipf=[1.1] * 4, vdist=1, p1=0.0, p2=1.0,
                    desc="4-mode 500000x200000x50000x1000, 1e8 nnz, Zipf(1.1) indices, lognormal(0,1) values, rank 32")
    if name == "c2s":  # reduced shape for quick checks (not a bench line)
        return dict(dims=[50000, 20000, 5000, 100], nnz=2_000_000, seed=0xA1700001, rank=32,
                    mode_dist=[1, 1, 1, 1], zipf=[1.1] * 4, vdist=1, p1=0.0, p2=1.0,
                    desc="reduced c2 shape")
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
    ap.add_argument("--workload", default="c2")
    ap.add_argument("--prec", default="f32", choices=["f32", "bf16", "f64"])
    ap.add_argument("--algo", type=int, default=0)
    ap.add_argument("--als-iters", type=int, default=5)
    ap.add_argument("--no-als", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=25_000_000)
    ap.add_argument("--probe", action="store_true",
                    help="random 128-byte row gather rate vs table size (the MTTKRP's miss-service bound)")
    ap.add_argument("--counters", action="store_true",
                    help="untimed pass for rocprofv3 --pmc: each mode twice in FP32, BF16 and FP64")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import alto_amd as A
    A.init(0)
    dev = A.device_name()
    if args.probe:
        for lg in range(21, 32, 2):
            g = 1 << 26
            ms = A.gather_probe(1 << lg, g)
            print(json.dumps({"table_MiB": (1 << lg) >> 20, "ms": round(ms, 4), "G_rows_per_s": round(g / ms / 1e6, 2),
                              "GB_per_s": round(g * 128 / ms / 1e6, 1)}), flush=True)
        return
    W = workload(args.workload)
    dims, nnz, R = W["dims"], W["nnz"], W["rank"]
    N = len(dims)
    prec = {"f32": A.PREC_F32, "bf16": A.PREC_BF16, "f64": A.PREC_F64}[args.prec]
    log("device:", dev)
    t0 = time.time()
    cidx, vals = A.generate(dims, nnz, W["seed"], W["mode_dist"], W["zipf"], W["vdist"], W["p1"], W["p2"])
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
        for p in (A.PREC_F32, A.PREC_BF16, A.PREC_F64):
            dF = [A.to_prec(f, p) for f in F64]
            os_ = [A.DeviceArray.zeros((d, R), np.float64 if p == A.PREC_F64 else np.float32) for d in dims]
            for _ in range(2):
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
    for _ in range(args.warmup):
        step()
    A.synchronize()
    log("warmup (incl. per-mode schedules) %.2f s" % (time.time() - t0))
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
                   "alto_create_s": round(t_create, 3), "device": dev},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic(args.workload, args.prec),
                     "alg_bytes_per_launch": int(tot / N)},
    }
    for d in outs:
        d.close()

    if not args.no_als:
        als = {}
        for name, p in (("f32", A.PREC_F32), ("f64", A.PREC_F64)):
            t0 = time.time()
            _, lam, fits, fe, tim = A.cpd_alto(at, F64, R, args.als_iters, eps=1e-30, prec=p)
            als[name] = dict(s_per_iter=round(float(tim[7]), 5), mttkrp_s_per_iter=round(float(tim[1] / len(fits)), 5),
                             fit=float(fits[-1]), fit_exact=float(fe[-1]), iters=len(fits))
            log("CP-ALS %s: %.4f s/iter, fit %.8f (wall %.1f s)" % (name, tim[7], fits[-1], time.time() - t0))
        als["speedup_f32_vs_f64"] = round(als["f64"]["s_per_iter"] / als["f32"]["s_per_iter"], 3)
        als["dfit"] = abs(als["f32"]["fit"] - als["f64"]["fit"])
        result["cp_als"] = als

    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(A, at, dims, R, F64, args.cpu_sample)
    at.close()
    print(json.dumps(result), flush=True)


def measured_traffic(workload, prec):
    """HBM bytes per MTTKRP launch (FETCH_SIZE + WRITE_SIZE, scaled by a copy
    kernel of known size) from the committed rocprofv3 summary of this
    workload, or None when there is none."""
    if workload != "c2":
        return None
    path = os.path.join(ROOT, "profiles", "mttkrp_c2_rank32.md")
    key = "traffic_bytes_per_launch_%s:" % {"f32": "fp32", "bf16": "bf16", "f64": "fp64"}[prec]
    try:
        for line in open(path):
            if line.startswith(key):
                return int(line.split(":")[1])
    except OSError:
        pass
    return None


def cpu_baseline(A, at, dims, R, F64, sample):
    """The oracle's OpenMP restatement of the reference MTTKRP (da_mem_pull /
    atomic dispatch, FP64) on the first `sample` nonzeros of the ALTO order."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    import ctypes
    threads = min(16, os.cpu_count() or 1)
    S = min(sample, at.nnz)
    keys_d, vals_d = ctypes.c_void_p(), ctypes.c_void_p()
    A.check(A.lib().alto_hip_tensor_device_arrays(ctypes.c_void_p(at.handle), ctypes.byref(keys_d),
                                                   ctypes.byref(vals_d)), "device_arrays")
    keys = np.empty((S, at.kw), dtype=np.uint64)
    vals = np.empty(S, dtype=np.float64)
    A.check(A.lib().alto_hip_memcpy_d2h(keys.ctypes.data_as(ctypes.c_void_p), keys_d,
                                        ctypes.c_size_t(keys.nbytes), None), "d2h")
    A.check(A.lib().alto_hip_memcpy_d2h(vals.ctypes.data_as(ctypes.c_void_p), vals_d,
                                        ctypes.c_size_t(vals.nbytes), None), "d2h")
    O.set_threads(threads)

    class T:  # oracle-side ALTO tensor over the host copy
        pass
    t = T()
    t.dims = np.array(dims, dtype=np.uint64)
    t.nmode = len(dims)
    t.masks, _ = O.alto_masks(dims, at.kw)
    t.kw, t.keys, t.vals, t.nnz, t.nprtn = at.kw, keys, vals, S, threads
    t.prtn_ptr, t.intervals = O.prtn_alto(t.dims, keys, t.masks, threads)
    outs = [np.zeros((d, R)) for d in dims]
    t0 = time.time()
    for n in range(len(dims)):
        O.AltoTensor.mttkrp(t, F64, n, out=outs[n])
    dt = time.time() - t0
    tot, _ = alg_bytes(dims, S, R, 8 * at.kw, 8, 8, 8)
    log("cpu baseline: %.2f s for one all-mode FP64 MTTKRP over %d nnz on %d threads" % (dt, S, threads))
    return {"value": round(tot / dt / 1e9, 3), "unit": "GB/s", "cores": threads, "kind": "port",
            "sample": "one all-mode FP64 MTTKRP (oracle OpenMP restatement of the reference dispatch) over "
                      "the first %d nonzeros in ALTO order: %.3f s; extrapolated %.1f ms per all-mode step "
                      "at 1e8 nnz" % (S, dt, dt * 1e3 * 1e8 / S)}


if __name__ == "__main__":
    main()
