"""CPU checks of the C ABI library: it loads, exports every symbol that
include/mosaic_hip.h declares, and reports its version and error text without
touching a GPU."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "mosaic_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mh_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_binding_exports(mh):
    assert _declared() == sorted(mh.EXPORTS)


def test_library_exports_every_declared_symbol(mh):
    lib = ctypes.CDLL(mh.LIB_PATH)
    missing = [s for s in _declared() if not hasattr(lib, s)]
    assert not missing, missing


def test_version_and_error_string(mh):
    lib = mh.load()
    assert lib.mh_abi_version() == 2
    assert isinstance(lib.mh_last_error(), bytes)


def test_null_arguments_are_rejected_without_a_device(mh):
    lib = mh.load()
    assert lib.mh_context_create(0, None) == mh.MH_ERR_INVALID_ARGUMENT
    assert b"NULL" in lib.mh_last_error()
