import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "alto-stream_amd"))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def gpu():
    """Initialises the HIP library; the GPU tests fail loudly if it cannot."""
    import alto_amd
    alto_amd.init(0)
    print("device:", alto_amd.device_name(), flush=True)
    return alto_amd


def golden(name):
    return os.path.join(GOLDEN, name)


def rand_tensor(dims, nnz, seed, dup_frac=0.0, distinct=False):
    """Seeded COO tensor with uniform coordinates (numpy); dup_frac of the
    nonzeros repeat earlier coordinates."""
    rng = np.random.default_rng(seed)
    dims = [int(d) for d in dims]
    if distinct:
        total = int(np.prod(dims, dtype=np.float64))
        assert nnz <= total
        lin = set()
        out = []
        while len(out) < nnz:
            cand = rng.integers(0, total, size=2 * (nnz - len(out)) + 8)
            for c in cand:
                if int(c) not in lin:
                    lin.add(int(c))
                    out.append(int(c))
                    if len(out) == nnz:
                        break
        lin = np.array(out, dtype=np.uint64)
        cidx = []
        for d in reversed(dims):
            cidx.append(lin % np.uint64(d))
            lin = lin // np.uint64(d)
        cidx = cidx[::-1]
    else:
        cidx = [rng.integers(0, d, size=nnz, dtype=np.uint64) for d in dims]
        ndup = int(dup_frac * nnz)
        if ndup:
            src = rng.integers(0, nnz - ndup, size=ndup)
            for m in range(len(dims)):
                cidx[m][nnz - ndup:] = cidx[m][src]
    vals = rng.random(nnz) + 0.1
    return dims, cidx, vals


def rand_factors(dims, rank, seed, lo=0.0, hi=1.0):
    rng = np.random.default_rng(seed)
    return [lo + (hi - lo) * rng.random((int(d), rank)) for d in dims]
