"""The C-ABI library, CPU only: it loads, exports every symbol that
include/sjcs.h declares, and refuses to compute without a HIP device (no
CPU fallback)."""
import ctypes
import os
import re

import pytest

import sjcs_testlib as t
from conftest import have_gpu

HEADER = os.path.join(t.ROOT, "include", "sjcs.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sjcs_\w+)\s*\(", text)))


def test_every_declared_symbol_is_exported(built):
    lib = ctypes.CDLL(built.LIB_PATH)
    syms = declared_symbols()
    assert len(syms) >= 20
    missing = [s for s in syms if not hasattr(lib, s)]
    assert missing == []
    assert sorted(built.EXPORTED) == syms


def test_host_library_loads(built):
    lib = ctypes.CDLL(built.HOST_LIB_PATH)
    for s in ("sjcs_read_halos", "sjcs_write_result", "sjcs_cli_parse",
              "sjcs_host_shell_thresholds"):
        assert hasattr(lib, s)


@pytest.mark.skipif(have_gpu(), reason="checks the no-device path")
def test_no_device_fails_loudly(built):
    with pytest.raises(built.SJCSError) as ei:
        built.Engine(0)
    assert ei.value.code == built.SJCS_ERR_DEVICE
    lib = built.load_library()
    h = ctypes.c_void_p()
    assert lib.sjcs_engine_create(0, ctypes.byref(h)) == built.SJCS_ERR_DEVICE
    assert b"device" in lib.sjcs_last_error(None)


def test_null_arguments_are_rejected(built):
    lib = built.load_library()
    assert lib.sjcs_engine_create(0, None) == built.SJCS_ERR_INVALID
    assert lib.sjcs_set_shells(None, None, 0) == built.SJCS_ERR_INVALID
    assert lib.sjcs_add_particles(None, None, 0) == built.SJCS_ERR_INVALID
    assert lib.sjcs_get_counts(None, None) == built.SJCS_ERR_INVALID
    assert lib.sjcs_index_build(None) == built.SJCS_ERR_INVALID
    lib.sjcs_engine_destroy(None)
    lib.sjcs_index_destroy(None)
