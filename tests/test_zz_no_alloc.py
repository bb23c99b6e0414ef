"""nnp_compute_noalloc: the energy+force pass with no device or pinned allocation by the call (its
workspace is the library's fixed per-device scratch array), against nnp_compute (the same kernels
on allocated workspaces): results bit for bit, on two calls in a row (the second smaller), in the
FP64, mixed and all-images forms; device free memory unchanged across calls; oversize
structures rejected with a clear error, after which both entry points keep working."""
import ctypes

import numpy as np
import pytest

from atomnn_amd import synth
from atomnn_amd.nnp import NNP, NNPError, NNP_ALL_IMAGES, NNP_ATOM_ENERGY, NNP_FORCE, NNP_PREC_MIXED, NNP_VACUUM_ENERGY

pytestmark = pytest.mark.gpu

FLAGS = NNP_FORCE | NNP_ATOM_ENERGY | NNP_VACUUM_ENERGY


def device_free_bytes():
    hip = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so")
    free, total = ctypes.c_size_t(0), ctypes.c_size_t(0)
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


def small_cell(n, seed):
    rng = np.random.default_rng(seed)
    cell = synth.cubic_cell(n, 0.05)
    return synth.Structure(synth.species_exact(n, [0.5, 0.5], rng), synth.random_positions(n, cell, 2.0, seed + 1),
                           cell)


def same(a, b):
    (E, Ea, F), (E2, Ea2, F2) = a, b
    assert E == E2
    assert np.array_equal(Ea, Ea2)
    assert (F is None and F2 is None) or np.array_equal(F, F2)


@pytest.mark.parametrize("extra", [0, NNP_PREC_MIXED, NNP_ALL_IMAGES], ids=["fp64", "mixed", "all_images"])
def test_noalloc_equals_compute_on_two_calls(potentials, extra):
    m = NNP.read(potentials["plain4096"])
    try:
        flags = FLAGS | extra
        big, small = synth.plain4096()[0], small_cell(300, 7)
        for s in (big, small):  # the second call is smaller: the scratch keeps the first call's contents
            same(m.compute_noalloc(s, flags), m.compute(s, flags))
        # and a structure with many images per pair (all-pairs neighbour search)
        tri = synth.batch512(3)[2]
        same(m.compute_noalloc(tri, flags), m.compute(tri, flags))
    finally:
        m.close()


def test_noalloc_allocates_no_device_memory(potentials):
    m = NNP.read(potentials["mixed64x4096"])
    try:
        s1, s2 = synth.mixed64x4096(1)[0], small_cell(500, 11)
        m.compute_noalloc(s1, FLAGS)  # first use: the runtime's own staging for pageable copies
        free0 = device_free_bytes()
        for s in (s1, s2, s1):
            m.compute_noalloc(s, FLAGS | NNP_PREC_MIXED)
            m.compute_noalloc(s, FLAGS)
        assert device_free_bytes() == free0
    finally:
        m.close()


def test_noalloc_rejects_oversize_and_keeps_working(potentials):
    m = NNP.read(potentials["plain4096"])
    try:
        big = small_cell(20000, 3)
        with pytest.raises(NNPError) as ei:
            m.compute_noalloc(big, FLAGS)
        assert "natoms" in str(ei.value) or "scratch" in str(ei.value)
        mid = small_cell(14000, 4)  # within the atom limit, beyond what the scratch array holds
        with pytest.raises(NNPError) as ei:
            m.compute_noalloc(mid, FLAGS)
        assert "scratch" in str(ei.value)
        s = small_cell(400, 5)
        same(m.compute_noalloc(s, FLAGS), m.compute(s, FLAGS))
        E, Ea, F = m.compute(big, FLAGS)
        assert np.isfinite(E) and F.shape == (20000, 3)
    finally:
        m.close()
