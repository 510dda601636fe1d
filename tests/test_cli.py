"""cpd64 / cpd128: the reference's command line (src/main.cpp) on the GPU
library.  CPU tests check the binary; GPU tests run its modes end to end."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import oracle as O
from conftest import ROOT, golden

CPD64 = os.path.join(ROOT, "alto-stream_amd", "_build", "cpd64")
CPD128 = os.path.join(ROOT, "alto-stream_amd", "_build", "cpd128")


def run(args, timeout=120, exe=CPD64):
    p = subprocess.run([exe] + args, capture_output=True, text=True, timeout=timeout)
    return p.returncode, p.stdout, p.stderr


def test_cli_built_and_usage():
    assert os.path.exists(CPD64) and os.path.exists(CPD128)
    rc, out, err = run(["-h"])
    assert rc == 0 and "--bench" in err and "cpstream_alto" in err


def test_cli_without_device_fails_loudly():
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    rc, out, err = run(["-i", golden("small.tns"), "-l", "als"])
    assert rc != 0 and "no usable GPU" in err


@pytest.mark.gpu
def test_cli_check_all_modes(gpu):
    """-c compares the GPU ALTO MTTKRP with the sequential COO MTTKRP on the
    host (mttkrp.cpp:122-156), element by element at 1e-6 (main.cpp:632)."""
    rc, out, err = run(["-c", "-t", "-1", "-r", "16", "-x", "3", "-i", golden("med4.tns")])
    assert rc == 0, err
    assert out.count("Results of mttkrp_alto is correct") == 4, out


@pytest.mark.gpu
def test_cli_bench_prints_runtime(gpu):
    # med5 needs 66 key bits: the 64-bit binary refuses it, the 128-bit one runs it
    rc, out, err = run(["-p", "-t", "-1", "-m", "5", "-r", "32", "--prec", "f32", "-i", golden("med5.tns")])
    assert rc != 0 and "more bits" in err
    rc, out, err = run(["-p", "-t", "-1", "-m", "5", "-r", "32", "--prec", "f32", "-i", golden("med5.tns")],
                       exe=CPD128)
    assert rc == 0, err
    assert re.search(r"ALTO runtime:\s+[0-9.]+", out) and "ALTO creation time" in out


def _glibc_init(dims, R, seed):
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(ctypes.c_uint(seed))
    RAND_MAX = 2147483647
    out = []
    for d in dims:
        a = np.empty(d * R)
        for i in range(d * R):
            v = 3.0 * (libc.rand() / RAND_MAX)
            if libc.rand() % 2 == 0:
                v = -v
            a[i] = v
        out.append(a.reshape(d, R))
    return out


@pytest.mark.gpu
def test_cli_als_matches_oracle_with_reference_init(gpu):
    dims, cidx, vals = O.read_tns(golden("med.tns"))
    R, seed, iters = 8, 7, 4
    rc, out, err = run(["-l", "als", "-r", str(R), "-m", str(iters), "-e", "1e-30", "-x", str(seed),
                        "-i", golden("med.tns")])
    assert rc == 0, err
    fits = [float(x) for x in re.findall(r"it: \d+\t fit: ([-0-9.e+]+)", out)]
    assert len(fits) == iters, out
    init = _glibc_init(dims, R, seed)
    ref = O.AltoTensor(dims, cidx, vals, nprtn=1)
    _, _, rfits, _ = ref.cpd_als(init, R, iters)
    np.testing.assert_allclose(fits, rfits, rtol=1e-5)


@pytest.mark.gpu
def test_cli_cpstream(gpu, tmp_path):
    rng = np.random.default_rng(5)
    T, nnz = 4, 3000
    t = rng.integers(0, T, size=nnz)
    t[:T] = np.arange(T)
    i = rng.integers(0, 80, size=nnz)
    j = rng.integers(0, 60, size=nnz)
    v = rng.poisson(3, size=nnz) + 1
    p = tmp_path / "s.tns"
    with open(p, "w") as f:
        for a, b, c, d in zip(t, i, j, v):
            f.write("%d %d %d %d\n" % (a + 1, b + 1, c + 1, d))
    rc, out, err = run(["-l", "cpstream_alto", "-a", "0", "-z", "nonneg", "-r", "4", "-m", "3", "-e", "1e-6",
                        "-i", str(p)])
    assert rc == 0, err
    assert "Total CPSTREAM time" in out and "(4 time slices in 4 batches" in out


def _write_stream(p, seed, T=4, nnz=3000):
    rng = np.random.default_rng(seed)
    t = rng.integers(0, T, size=nnz)
    t[:T] = np.arange(T)
    i = rng.integers(0, 80, size=nnz)
    j = rng.integers(0, 60, size=nnz)
    v = rng.poisson(3, size=nnz) + 1
    with open(p, "w") as f:
        for a, b, c, d in zip(t, i, j, v):
            f.write("%d %d %d %d\n" % (a + 1, b + 1, c + 1, d))


@pytest.mark.gpu
def test_cli_cpstream_timeslice_limit(gpu, tmp_path):
    """-g 2 groups the 4 time slices into 2 batches (next_dynamic_batch)."""
    p = tmp_path / "s.tns"
    _write_stream(p, 7)
    rc, out, err = run(["-l", "cpstream_alto", "-a", "0", "-z", "nonneg", "-r", "4", "-m", "3", "-e", "1e-6",
                        "-g", "2", "-i", str(p)])
    assert rc == 0, err
    assert "timeslice_limit: 2" in out and "(4 time slices in 2 batches" in out


@pytest.mark.parametrize("flag", [["-d", "10,10,10"], ["-s", "0.1"], ["-f", "1"]])
def test_cli_refuses_generator_flags(flag):
    """The synthetic generator is not built: its flags are errors, not ignored."""
    rc, out, err = run(flag + ["-i", golden("small.tns"), "-l", "als"])
    assert rc != 0 and "generator" in err


def test_cli_refuses_unknown_precision():
    rc, out, err = run(["--prec", "f16", "-i", golden("small.tns"), "-l", "als"])
    assert rc != 0 and "Invalid --prec" in err


@pytest.mark.gpu
def test_cli_refuses_bf16_stream(gpu, tmp_path):
    p = tmp_path / "s.tns"
    _write_stream(p, 8)
    rc, out, err = run(["-l", "cpstream_alto", "-a", "0", "--prec", "bf16", "-i", str(p)])
    assert rc != 0 and "bf16" in err


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["f32", "bf16"])
def test_cli_bench_reduced_precision(gpu, prec):
    """-p with FP32 or BF16 factors writes the MTTKRP in the factors' own
    storage (BF16 through an FP32 accumulation copy): no output overrun."""
    rc, out, err = run(["-p", "-t", "-1", "-m", "3", "-r", "16", "--prec", prec, "-i", golden("med4.tns")])
    assert rc == 0, err
    assert re.search(r"ALTO runtime:\s+[0-9.]+", out)


@pytest.mark.gpu
def test_cli_spcpstream(gpu, tmp_path):
    """-l spcpstream_alto runs spCP-stream (main.cpp:163, 397); the constraint
    flag is accepted and unused, as in the reference."""
    rng = np.random.default_rng(6)
    T, nnz = 4, 3000
    t = rng.integers(0, T, size=nnz)
    t[:T] = np.arange(T)
    i = rng.integers(0, 80, size=nnz)
    j = rng.integers(0, 60, size=nnz)
    v = rng.poisson(3, size=nnz) + 1
    p = tmp_path / "s.tns"
    with open(p, "w") as f:
        for a, b, c, d in zip(t, i, j, v):
            f.write("%d %d %d %d\n" % (a + 1, b + 1, c + 1, d))
    rc, out, err = run(["-l", "spcpstream_alto", "-a", "0", "-z", "nonneg", "-r", "4", "-m", "3", "-e", "1e-6",
                        "-i", str(p)])
    assert rc == 0, err
    assert "spCPSTREAM" in err and "Total CPSTREAM time" in out and "(4 time slices in 4 batches" in out
    assert "final fit error" in out
