"""The C-ABI library loads and exports every symbol include/*.h declares
(no compute calls: runs without a GPU)."""
import ctypes
import glob
import os
import re

import numpy as np
import pytest

import alto_amd
from conftest import ROOT


def _declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(alto_hip_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_library_is_built_in_tree():
    assert os.path.exists(alto_amd.LIB_PATH)
    assert alto_amd.LIB_PATH.startswith(ROOT)


def test_every_declared_symbol_is_exported():
    lib = alto_amd.lib()
    names = _declared()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert missing == []


def test_error_strings():
    lib = alto_amd.lib()
    assert lib.alto_hip_error_string(0) == b"ok"
    assert lib.alto_hip_error_string(-2).startswith(b"coordinates need more bits")


def test_no_device_is_reported_not_faked():
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    lib = alto_amd.lib()
    assert lib.alto_hip_init(0) == -5  # ALTO_ERR_NO_DEVICE
    with pytest.raises(alto_amd.AltoError):
        alto_amd.init(0)


def test_mttkrp_output_type_is_checked_before_launch():
    """The binding refuses an output array whose element type or size does
    not match what the kernel writes (FP64 rows for out_f64 / PREC_F64)."""
    import alto_amd as A

    class Fake:
        def __init__(self, dtype, n):
            self.dtype = np.dtype(dtype)
            self.nbytes = n * self.dtype.itemsize

    A._check_out(Fake(np.float64, 40 * 16), 40, 16, True)
    A._check_out(Fake(np.float32, 40 * 16), 40, 16, False)
    with pytest.raises(A.AltoError, match="writes float64"):
        A._check_out(Fake(np.float32, 40 * 16), 40, 16, True)
    with pytest.raises(A.AltoError, match="holds"):
        A._check_out(Fake(np.float64, 39 * 16), 40, 16, True)
