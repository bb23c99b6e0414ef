"""Device-resident batches and hipGraph replay.

Every replay of the captured pass must reproduce the eager pass bit for bit, also with
unrelated eager GPU work (another model's evaluation) between the replays.  The resident
energies use a fixed-order device reduction; they match the per-call results to 1e-12.
"""
import numpy as np
import pytest

from atomnn_amd import synth
from atomnn_amd.nnp import NNP, NNP_FORCE, NNP_PREC_MIXED, NNP_VACUUM_ENERGY

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mixed", [False, True])
def test_graph_replays_match_eager(potentials, mixed):
    flags = NNP_FORCE | NNP_VACUUM_ENERGY | (NNP_PREC_MIXED if mixed else 0)
    m = NNP.read(potentials["plain4096"])
    other = NNP.read(potentials["plain64"])
    strucs = synth.plain4096() + synth.batch512(6)
    R = m.resident(strucs, flags)
    R.run()
    eager = R.fetch()
    for rep in range(4):
        R.run()  # rep 0 captures the graph, later reps replay it
        got = R.fetch()
        for (E0, Ea0, F0), (E1, Ea1, F1) in zip(eager, got):
            assert E0 == E1
            np.testing.assert_array_equal(Ea0, Ea1)
            np.testing.assert_array_equal(F0, F1)
        other.compute(synth.plain64()[0])  # unrelated eager work between replays
    single = m.compute_batch(strucs, flags)
    for (E0, Ea0, F0), (E1, Ea1, F1) in zip(eager, single):
        assert abs(E0 - E1) <= 1e-12 * max(1.0, np.abs(Ea1).sum())
        np.testing.assert_array_equal(Ea0, Ea1)
        np.testing.assert_array_equal(F0, F1)
    m.close()
    other.close()


def test_profiled_passes_match(potentials):
    m = NNP.read(potentials["plain64"])
    strucs = synth.plain64()
    R = m.resident(strucs, NNP_FORCE)
    R.run()
    E0 = R.fetch()[0]
    R.profile(True)
    for _ in range(3):
        R.run()
        R.sync()
    ms, calls = R.stage_ms()
    R.profile(False)
    assert (calls[:5] == 3).all() and (ms[:5] > 0).all()
    E1 = R.fetch()[0]
    assert E0[0] == E1[0]
    m.close()
