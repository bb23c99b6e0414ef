"""Serving form: continuous batching with overlapping lanes.

* every result is the same alone (nnp_compute) and inside the server's mixed batches
  (bit for bit: kernels are per atom, reductions per structure in a fixed order);
* a request larger than a lane's HBM budget is rejected with NNP_EBUDGET and the server
  keeps serving;
* under an open-loop stream, several batches are in flight at once.
"""
import numpy as np
import pytest

from atomnn_amd import serve, synth
from atomnn_amd.nnp import NNP, NNP_FORCE, NNP_PREC_MIXED, NNP_VACUUM_ENERGY, NNPError

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mixed", [False, True])
def test_results_same_alone_and_in_mix(potentials, mixed):
    flags = NNP_FORCE | NNP_VACUUM_ENERGY | (NNP_PREC_MIXED if mixed else 0)
    m = NNP.read(potentials["serve20k"])
    pool = [s for s in synth.serve_pool(40) if s.natoms <= 3000][:14]
    alone = [m.compute(s, flags) for s in pool]
    srv = serve.Server(m, flags, lanes=3, max_batch_atoms=6000)
    arrivals = synth.serve_arrivals(len(pool) * 3, rate=400.0)
    structs = [pool[k % len(pool)] for k in range(len(arrivals))]
    res = serve.run_stream(srv, structs, arrivals, keep_results=True)
    for k, (E, Ea, F) in enumerate(res.results):
        E0, Ea0, F0 = alone[k % len(pool)]
        assert E == E0
        np.testing.assert_array_equal(Ea, Ea0)
        np.testing.assert_array_equal(F, F0)
    st = srv.stats()
    print(st, f"p50 {res.p50_ms:.2f} ms p99 {res.p99_ms:.2f} ms")
    assert st["completed"] == len(structs) and st["batches"] >= 2
    srv.close()
    m.close()


def test_budget_rejection_keeps_serving(potentials):
    m = NNP.read(potentials["serve20k"])
    srv = serve.Server(m, NNP_FORCE, lanes=2, hbm_budget=64 << 20)
    big = synth.serve_pool(60)
    big = max(big, key=lambda s: s.natoms)
    with pytest.raises(NNPError) as ei:
        srv.submit(big)
    assert ei.value.code == -101
    small = [s for s in synth.serve_pool(30) if s.natoms < 200][:3]
    for s in small:
        E, Ea, F = srv.submit(s).wait()
        E0, Ea0, F0 = m.compute(s, NNP_FORCE)
        assert E == E0
    srv.close()
    m.close()
