"""The C ABI without a GPU: the engine library loads, and it exports every function that
include/nnp_capi.h declares (no compute call is made)."""
import ctypes
import os
import re

import pytest

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


def test_cpp_facade_compiles_and_links(tmp_path):
    """include/nnp_pot.hpp (the ptnl::Pot-style facade) compiles against the header and links
    against the library; the program makes no device call unless given a potential file."""
    import shutil
    import subprocess

    if shutil.which("g++") is None:
        pytest.skip("no g++")
    src = tmp_path / "facade.cpp"
    src.write_text(
        '#include "nnp_pot.hpp"\n'
        "#include <cstdio>\n"
        "int main(int argc, char** argv) {\n"
        "    if (argc < 2) return 0;\n"
        "    atomnn::GpuPot pot(argv[1]);\n"
        "    const int32_t type[2] = {0, 0};\n"
        "    const double pos[6] = {0, 0, 0, 1.5, 0, 0};\n"
        "    double f[6] = {0, 0, 0, 0, 0, 0};\n"
        "    const double e = pot.compute(2, type, pos, nullptr, f);\n"
        "    std::printf(\"%.17g %.17g %.17g\\n\", e, pot.energy(2, type, pos, nullptr), f[0]);\n"
        "    return 0;\n"
        "}\n")
    exe = tmp_path / "facade"
    lib = os.path.dirname(LIB)
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe),
                        "-L", lib, "-latomnn_amd", "-Wl,-rpath," + lib, "-L/opt/rocm/lib", "-lamdhip64",
                        "-Wl,-rpath,/opt/rocm/lib"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert subprocess.run([str(exe)], capture_output=True).returncode == 0
