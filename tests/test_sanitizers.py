"""AddressSanitizer and UndefinedBehaviorSanitizer builds of the CPU-side code (host compiler only):

* the host library's potential-file parser and device-plan builder (atomnn_amd/csrc/model.cpp) on
  the workload potentials and on several hundred corrupted copies (truncations, garbage or
  out-of-range numbers, absurd counts, dropped and repeated lines): every file either parses or is
  rejected with the parser's exception (the reference throws std::invalid_argument /
  std::runtime_error from NNP::read, src/nnp/nnp.cpp:254, 265), with no memory error or undefined
  behaviour;
* the oracle driven through every entry point of its interface (neighbours, descriptors, energies
  and forces under every flag combination, energy and force training) on periodic, sheared, open
  and one-atom structures, for the workload potentials and a generic-basis variant.
"""
import os
import re
import subprocess

import numpy as np
import pytest

from atomnn_amd import synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")
REPORT = re.compile(r"AddressSanitizer|LeakSanitizer|runtime error:|UndefinedBehaviorSanitizer")


@pytest.fixture(scope="module")
def tools():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"], check=True)
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "atomnn_amd", "csrc"), "sanitize"], check=True)
    return (os.path.join(ROOT, "oracle", "lib", "sanitize_oracle"),
            os.path.join(ROOT, "atomnn_amd", "bin", "sanitize_model"))


def potentials(d):
    paths = []
    for name in ("plain64", "plain4096", "mixed64x4096"):
        p = os.path.join(d, f"{name}.ann")
        synth.workload_potential(name).write(p)
        paths.append(p)
    pot = synth.make_potential(["A", "B"], 6.0, 4, [0.005, 0.05], [1.5, 4.0], [1, -1], [8], np.random.default_rng(3))
    pot.phi_r, pot.phi_a, pot.neuron, pot.cut, pot.norm = "SWISH", "SECHS", "SOFTPLUS", "STEP", "VOL"
    p = os.path.join(d, "generic.ann")
    pot.write(p)
    paths.append(p)
    return paths


def run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True, env=ENV, timeout=600)
    out = r.stdout + r.stderr
    assert not REPORT.search(out), out[-4000:]
    assert r.returncode == 0, out[-4000:]
    return r.stdout


def test_oracle_under_sanitizers(tools, tmp_path):
    out = run([tools[0]] + potentials(str(tmp_path)))
    assert "0 failed calls" in out


def corrupted(src, d):
    text = open(src).read()
    lines = text.split("\n")
    rng = np.random.default_rng(17)
    out = []

    def put(t):
        p = os.path.join(d, f"bad{len(out):04d}.ann")
        with open(p, "w") as f:
            f.write(t)
        out.append(p)

    for cut in np.linspace(0, len(text), 60).astype(int):  # truncations
        put(text[:cut])
    for k in range(0, len(lines), max(1, len(lines) // 40)):  # one line dropped, one repeated
        put("\n".join(lines[:k] + lines[k + 1:]))
        put("\n".join(lines[:k + 1] + lines[k:]))
    numbers = [m.span() for m in re.finditer(r"-?\d+(\.\d+)?(e-?\d+)?", text)]
    for q in rng.choice(len(numbers), size=min(150, len(numbers)), replace=False):  # one number replaced
        a, b = numbers[q]
        for bad in ("nan", "-1", "1e400", "abc", "99999999", "0", ""):
            put(text[:a] + bad + text[b:])
    ntrunc = 60
    nstruct_from = None
    for pat, rep in ((r"ntypes \d+", "ntypes 100000"), (r"ntypes \d+", "ntypes -3"), (r"ntypes \d+", "ntypes 0"),
                     (r"nn 80 32 32 1", "nn 80 -32 32 1"), (r"nn 80 32 32 1", "nn 0"),
                     (r"nn 80 32 32 1", "nn 80 32 32 2"), (r"GAUSSIAN 16", "GAUSSIAN 100000"),
                     (r"GAUSS 16 2", "GAUSS -16 2"), (r"GAUSS 16 2", "GAUSS 16 0"), (r"rc \S+", "rc -6"),
                     (r"COS", "BOGUS"), (r"name B", "name A"), (r"TANH", "NOPE")):
        if nstruct_from is None:
            nstruct_from = len(out)
        put(re.sub(pat, rep, text, count=1))
    return out, ntrunc, nstruct_from


def test_parser_under_sanitizers_on_corrupted_files(tools, tmp_path):
    src = potentials(str(tmp_path))[1]
    bad, ntrunc, nstruct = corrupted(src, str(tmp_path))
    out = run([tools[1], src] + bad)
    lines = out.strip().split("\n")
    assert len(lines) == len(bad) + 2 and lines[0].startswith(f"parsed {src}")
    outcome = [l.split(" ", 1)[0] for l in lines[1:-1]]
    # every truncation short of the whole file is rejected
    assert all(v == "rejected" for v in outcome[:ntrunc - 1]), lines[1:ntrunc]
    # every structural corruption (bad counts, cutoff, names, kinds) is rejected; a replaced weight
    # or parameter may parse, as the reference's own number reading accepts it
    assert all(v == "rejected" for v in outcome[nstruct:]), lines[nstruct + 1:]
