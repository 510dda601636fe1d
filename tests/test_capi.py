"""The C-ABI library loads and exports every symbol include/*.h declares
(no compute calls: runs without a GPU)."""
import ctypes
import glob
import os
import re

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
