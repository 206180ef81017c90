"""The C ABI library loads, exports every entry point of include/pyiid_hip.h,
and fails with a clear error (not a crash) where no device is present."""
import ctypes
import os
import re

import pytest

from pyiid_amd import _lib

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                      "include", "pyiid_hip.h")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(pyiid_\w+)\s*\(", text)))


def test_header_declares_the_binding():
    assert declared() == sorted(_lib.EXPORTED)


def test_library_exports_every_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for name in declared():
        assert hasattr(lib, name), name


def test_abi_version_and_error_channel():
    lib = _lib.load()
    assert lib.pyiid_abi_version() == 1
    # invalid call: null output pointer -> error code and message, no crash
    assert lib.pyiid_device_count(None) == -1
    assert b"null" in lib.pyiid_last_error()


def test_null_context_is_rejected():
    lib = _lib.load()
    assert lib.pyiid_set_positions(None, None, 0, 0) == -1
    assert lib.pyiid_fq(None, 0, 0.1, None, 0) == -1


@pytest.mark.skipif(_lib.device_count() > 0, reason="a HIP device is present")
def test_create_without_device_fails_loudly():
    with pytest.raises(RuntimeError, match="no HIP device"):
        _lib.Context(0)
