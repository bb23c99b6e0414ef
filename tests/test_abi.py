"""The C ABI without a GPU: the engine library loads, and it exports every function that
include/nnp_capi.h declares (no compute call is made)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "nnp_capi.h")
LIB = os.path.join(ROOT, "atomnn_amd", "lib", "libatomnn_amd.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)  # drop comments
    names = set(re.findall(r"\b(nnp_[A-Za-z0-9_]+)\s*\(", text))
    return sorted(n for n in names if not n.startswith("nnp_structure") and not n.startswith("nnp_result"))


def test_header_declares_the_expected_surface():
    names = declared_functions()
    for must in ("nnp_load", "nnp_free", "nnp_compute", "nnp_compute_batch", "nnp_last_error", "nnp_type_index",
                 "nnp_server_create", "nnp_server_submit", "nnp_server_wait", "nnp_md_create", "nnp_md_step",
                 "nnp_train_create", "nnp_train_error"):
        assert must in names, must
    assert len(names) >= 30


def test_library_exports_every_declared_function():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
