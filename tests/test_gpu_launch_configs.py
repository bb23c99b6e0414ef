"""Launch configurations that must not change any result, bit for bit: the FP64 network kernel's
weight staging and persistent grid (option net_launch, fixed and tuned), on a batch of more than
two 16-atom tiles per CU, through nnp_compute and through the resident graph-replayed pass."""
import numpy as np
import pytest

from atomnn_amd import synth
from atomnn_amd.nnp import NNP, NNP_ATOM_ENERGY, NNP_FORCE, NNP_VACUUM_ENERGY

pytestmark = pytest.mark.gpu
FLAGS = NNP_FORCE | NNP_ATOM_ENERGY | NNP_VACUUM_ENERGY


def flat(res):
    return np.concatenate([np.r_[E, Ea, F.ravel()] for E, Ea, F in res])


def test_net_launch_configurations_bitwise(tmp_path):
    path = str(tmp_path / "batch.ann")
    synth.workload_potential("batch512").write(path)
    m = NNP.read(path)
    try:
        strucs = synth.batch512(72)  # about 19k atoms: over 512 network tiles
        assert sum(s.natoms for s in strucs) > 2 * 256 * 16
        ref = None
        for opt in (0, 1, 2, -1):
            m.set_option("net_launch", opt)
            got = flat(m.compute_batch(strucs))
            if ref is None:
                ref = got
            assert np.array_equal(got, ref), opt
        # the resident pass sums the structure energies on the device (in another order than
        # nnp_compute's host sum): per-atom energies and forces bit for bit, totals to rounding
        per_atom = lambda res: np.concatenate([np.r_[Ea, F.ravel()] for _, Ea, F in res])
        want = m.compute_batch(strucs)
        r = m.resident(strucs, FLAGS)
        for _ in range(3):  # eager first pass (tuned), then graph replays
            r.run()
            got = r.fetch()
            assert np.array_equal(per_atom(got), per_atom(want))
            assert np.allclose([g[0] for g in got], [w[0] for w in want], rtol=1e-13, atol=0)
        small = synth.batch512(4)
        m.set_option("net_launch", 0)
        a = flat(m.compute_batch(small))
        m.set_option("net_launch", -1)
        assert np.array_equal(flat(m.compute_batch(small)), a)
    finally:
        m.close()


def test_synchronous_calls_with_speculated_rows(tmp_path):
    """nnp_compute stages the rows at the previous call's neighbour maximum and redoes a call
    whose maximum exceeds them: a sparse cell, then denser ones (the redo), then the sparse one
    again, each bit for bit equal to a fresh model's first call (no speculation)."""
    import numpy as np  # noqa: F811

    path = str(tmp_path / "plain.ann")
    synth.workload_potential("plain4096").write(path)
    rng = np.random.default_rng(5)
    strucs = []
    for rho in (0.03, 0.05, 0.07, 0.03):
        cell = synth.cubic_cell(1200, rho)
        strucs.append(synth.Structure(synth.species_exact(1200, [0.5, 0.5], rng),
                                      synth.random_positions(1200, cell, 1.8, 11), cell))
    m = NNP.read(path)
    try:
        got = [m.compute(s, FLAGS) for s in strucs]
    finally:
        m.close()
    for s, (E, Ea, F) in zip(strucs, got):
        f = NNP.read(path)
        try:
            E0, Ea0, F0 = f.compute(s, FLAGS)
        finally:
            f.close()
        assert E == E0 and np.array_equal(Ea, Ea0) and np.array_equal(F, F0)


def test_compute_into_fixed_buffers(tmp_path):
    """The fixed-buffer caller (the benchmark's nnpce-style loop) returns what compute returns."""
    import numpy as np  # noqa: F811

    path = str(tmp_path / "plain.ann")
    synth.workload_potential("plain4096").write(path)
    m = NNP.read(path)
    try:
        s = synth.batch512(3)[2]
        E0, Ea0, F0 = m.compute(s, FLAGS)
        call = m.compute_into(s, FLAGS)
        for _ in range(3):
            E, Ea, F = call()
            assert E == E0 and np.array_equal(Ea, Ea0) and np.array_equal(F, F0)
    finally:
        m.close()


def test_upload_tables_reused_across_calls(tmp_path):
    """A call with the previous call's atom counts, species and search modes uploads only the
    positions and cells (the per-atom tables stay on the device); a change of any of them uploads
    everything.  Each call of one model, in a sequence that alternates between the two, equals
    a fresh model's call bit for bit."""
    import numpy as np  # noqa: F811

    path = str(tmp_path / "plain.ann")
    synth.workload_potential("plain4096").write(path)
    rng = np.random.default_rng(8)
    n = 300
    cell = synth.cubic_cell(n, 0.05)
    a = synth.Structure(synth.species_exact(n, [0.5, 0.5], rng), synth.random_positions(n, cell, 1.8, 3), cell)
    moved = synth.Structure(a.types, a.pos + rng.normal(0, 0.05, a.pos.shape), cell)
    swapped = synth.Structure((1 - a.types).astype(np.int32), a.pos, cell)  # same counts, other species
    open_a = synth.Structure(a.types, a.pos, None)  # same species, all-pairs search instead of cell list
    small = synth.batch512(6)[1:3]
    seq = [[a], [moved], [moved], [swapped], [a], [open_a], [a], small,
           [small[0], synth.Structure(small[1].types, small[1].pos * 1.01, small[1].cell * 1.01)],
           [small[0], synth.Structure((1 - small[1].types).astype(np.int32), small[1].pos, small[1].cell)]]
    m = NNP.read(path)
    try:
        got = [m.compute_batch(ss, FLAGS) for ss in seq]
    finally:
        m.close()
    for ss, res in zip(seq, got):
        f = NNP.read(path)
        try:
            want = f.compute_batch(ss, FLAGS)
        finally:
            f.close()
        assert np.array_equal(flat(res), flat(want))
