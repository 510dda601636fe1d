"""Host tensor import (ImportSparseTensor, sptensor.cpp:74-242) of the C++
mirror, on the CPU: a small driver built against host/alto_amd.cpp reads
.tns files (the reference's fixtures and a large synthetic file whose
parallel chunks split it unevenly, with comments, blank lines, tabs and no
final newline) and a .bin file; everything must equal the oracle's read_tns
bit for bit (indices, values via %.17g, dims, mode count)."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from conftest import golden

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "alto-stream_amd")
LIBDIR = os.path.join(PKG, "_build")

DRIVER = r'''
#include "alto_amd.hpp"
#include <stdio.h>
#include <stdlib.h>
int main(int argc, char **argv) {
    SparseTensor *X;
    ImportSparseTensor(argv[1], argv[2][0] == 'b' ? BINARY_FORMAT : TEXT_FORMAT, &X, atoi(argv[3]));
    printf("%d %llu\n", X->nmodes, (unsigned long long)X->nnz);
    for (int m = 0; m < X->nmodes; ++m) printf("%llu ", (unsigned long long)X->dims[m]);
    printf("\n");
    for (IType i = 0; i < X->nnz; ++i) {
        for (int m = 0; m < X->nmodes; ++m) printf("%llu ", (unsigned long long)X->cidx[m][i]);
        printf("%.17g\n", X->vals[i]);
    }
    DestroySparseTensor(X);
    return 0;
}
'''


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    if not os.path.exists(os.path.join(LIBDIR, "libalto_amd.so")):
        pytest.skip("libalto_amd.so is not built")
    d = tmp_path_factory.mktemp("io")
    src = d / "drv.cpp"
    src.write_text(DRIVER)
    exe = d / "drv"
    subprocess.run(["g++", "-std=c++17", "-O2", "-I", os.path.join(PKG, "host"), "-I", os.path.join(ROOT, "include"),
                    str(src), os.path.join(PKG, "host", "alto_amd.cpp"), "-L", LIBDIR, "-lalto_amd",
                    "-Wl,-rpath," + LIBDIR, "-o", str(exe)], check=True)
    return str(exe)


def _read(driver, path, fmt="t", zero_based=0, threads=None):
    env = dict(os.environ)
    if threads:
        env["OMP_NUM_THREADS"] = str(threads)
    out = subprocess.run([driver, str(path), fmt, str(zero_based)], check=True, capture_output=True, text=True,
                         env=env).stdout.split("\n")
    nm, nnz = (int(x) for x in out[0].split())
    dims = [int(x) for x in out[1].split()]
    rows = [r.split() for r in out[2: 2 + nnz]]
    idx = np.array([[int(t) for t in r[:nm]] for r in rows], dtype=np.uint64).reshape(nnz, nm)
    vals = np.array([float(r[nm]) for r in rows])
    return dims, [idx[:, m] for m in range(nm)], vals


def _same(got, ref):
    gd, gc, gv = got
    rd, rc, rv = ref
    assert gd == rd
    for a, b in zip(gc, rc):
        np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(gv, rv)


@pytest.mark.parametrize("name", ["small.tns", "small4.tns", "med.tns", "med4.tns", "med5.tns"])
def test_import_fixtures(driver, name):
    _same(_read(driver, golden(name), threads=3), O.read_tns(golden(name)))


def test_import_zero_based_fixture(driver):
    p = golden("small4_zeroidx.tns")
    _same(_read(driver, p, zero_based=1), O.read_tns(p, zero_based=True))


@pytest.mark.parametrize("threads", [1, 7])
def test_import_large_synthetic(driver, tmp_path, threads):
    rng = np.random.default_rng(140)
    n = 120000
    idx = rng.integers(1, [900, 70000, 13], size=(n, 3))
    vals = rng.normal(size=n) * 10.0 ** rng.integers(-12, 12, size=n)
    p = tmp_path / "big.tns"
    with open(p, "w") as f:
        f.write("# synthetic\n# more comments\n")
        for k in range(n):
            sep = "\t" if k % 7 == 0 else " "
            f.write("%d%s%d%s%d%s%.17g" % (idx[k, 0], sep, idx[k, 1], sep, idx[k, 2], sep, vals[k]))
            if k + 1 < n:
                f.write("\n")
                if k % 5000 == 0:
                    f.write("\n# interleaved comment\n")
    _same(_read(driver, p, threads=threads), O.read_tns(p))


def test_import_bin(driver, tmp_path):
    dims, cidx, vals = O.read_tns(golden("med4.tns"))
    p = tmp_path / "x.bin"
    with open(p, "wb") as f:
        np.array([len(dims)] + list(dims) + [len(vals)], dtype=np.uint64).tofile(f)
        for c in cidx:
            np.asarray(c, dtype=np.uint64).tofile(f)
        np.asarray(vals, dtype=np.float64).tofile(f)
    _same(_read(driver, p, fmt="b"), (dims, cidx, vals))
