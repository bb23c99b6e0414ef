"""Parity against the FP64 oracle on the benchmark configurations themselves.

* serving requests (log-uniform 16-20000 atoms, three species in Dirichlet proportions, strained
  cubes at 0.055 /A^3): the plain FP64 form through nnp_compute within 1e-10, the mixed form
  through the server's continuous batches within 1 meV/atom and 20 meV/A RMS, per request;
* one 4096-atom cell of the 64 x 4096 precision batch (three species 40/30/30, 0.06 /A^3,
  rc 6.5, 84-64-64-1): plain FP64 energy, per-atom energies and forces within 1e-10, mixed within
  the gates;
* the 4096-atom two-species cell: per-atom energies within 1e-10;
* the batch of uneven triclinic cells in FP32 ("FP64 then FP32"): within the gates, in the
  separate-kernel and the fused pipelines.

Conventions (NNP::energy / NNP::force, src/nnp/nnp.cpp:390-482): energy |dE| / max(|E|, sum|E_i|),
per-atom max|dE_i| / max|E_i|, forces max_i |dF_i| / max_i |F_i|; mixed: |dE| / N in meV and the
RMS over atoms of |dF| in meV/A.
"""
import numpy as np
import pytest

from atomnn_amd import serve, synth
from atomnn_amd.nnp import NNP, NNP_ATOM_ENERGY, NNP_FORCE, NNP_PREC_MIXED, NNP_VACUUM_ENERGY
import oracle_binding as ob

pytestmark = pytest.mark.gpu

TOL = 1e-10
FLAGS = NNP_FORCE | NNP_ATOM_ENERGY | NNP_VACUUM_ENERGY
OFLAGS = ob.F_FORCE | ob.F_ATOM_ENERGY | ob.F_VACUUM


def check_fp64(res, ref, tag):
    (E, Ea, F), (Eo, Eao, Fo) = res, ref
    assert abs(E - Eo) <= TOL * max(abs(Eo), np.abs(Eao).sum()), (tag, E, Eo)
    assert np.abs(Ea - Eao).max() <= TOL * np.abs(Eao).max(), tag
    assert np.linalg.norm(F - Fo, axis=1).max() <= TOL * np.linalg.norm(Fo, axis=1).max(), tag


def mixed_errors(res, ref, n):
    (E, Ea, F), (Eo, Eao, Fo) = res, ref
    return abs(E - Eo) / n * 1e3, float(np.sqrt(np.mean(np.sum((F - Fo) ** 2, axis=1)))) * 1e3


def serving_requests():
    """Eight requests of the serving distribution: the three smallest of a 400-request pool
    (strained cells of 16-20 atoms, hundreds of images per pair), two of a few hundred atoms, two
    of a few thousand and the smallest above 10,000."""
    pool = synth.serve_pool(400)
    by = sorted(range(len(pool)), key=lambda k: pool[k].natoms)
    pick = by[:3]
    pick += [k for k in by if 200 <= pool[k].natoms <= 800][:2]
    pick += [k for k in by if 2000 <= pool[k].natoms <= 5000][:2]
    pick += [k for k in by if pool[k].natoms >= 10000][:1]
    return [pool[k] for k in pick]


@pytest.fixture(scope="module")
def serving(potentials):
    reqs = serving_requests()
    o = ob.OracleModel(potentials["serve20k"])
    ref = []
    for s in reqs:
        print(f"oracle: request of {s.natoms} atoms", flush=True)
        ref.append(o.compute(s.types, s.pos, s.cell, OFLAGS))
    m = NNP.read(potentials["serve20k"])
    yield m, reqs, ref
    m.close()


def test_serving_requests_plain_fp64(serving):
    m, reqs, ref = serving
    assert max(s.natoms for s in reqs) >= 10000 and min(s.natoms for s in reqs) < 100
    for s, r in zip(reqs, ref):
        check_fp64(m.compute(s, FLAGS), r, s.natoms)


def test_serving_requests_mixed_through_the_server(serving):
    """All eight submitted at once to a server of three lanes: they are packed into shared
    batches; each result is checked against the oracle on its own."""
    m, reqs, ref = serving
    srv = serve.Server(m, FLAGS | NNP_PREC_MIXED, lanes=3, max_batch_atoms=16384)
    try:
        pend = [srv.submit(s) for s in reqs]
        for s, p, r in zip(reqs, pend, ref):
            de, df = mixed_errors(p.wait(), r, s.natoms)
            print(f"{s.natoms:6d} atoms: {de:.5f} meV/atom, {df:.4f} meV/A RMS")
            assert de <= 1.0 and df <= 20.0, (s.natoms, de, df)
        assert srv.stats()["completed"] == len(reqs)
    finally:
        srv.close()


@pytest.fixture(scope="module")
def precision_cell(potentials):
    s = synth.mixed64x4096(1)[0]
    o = ob.OracleModel(potentials["mixed64x4096"])
    return s, o.compute(s.types, s.pos, s.cell, OFLAGS)


def test_precision_cell_plain_fp64(potentials, precision_cell):
    s, ref = precision_cell
    m = NNP.read(potentials["mixed64x4096"])
    try:
        check_fp64(m.compute(s, FLAGS), ref, "C3 fp64")
        # the resident, graph-replayed pass
        r = m.resident([s], FLAGS)
        for _ in range(3):
            r.run()
        check_fp64(r.fetch()[0], ref, "C3 fp64 resident")
    finally:
        m.close()


def test_precision_cell_mixed(potentials, precision_cell):
    s, ref = precision_cell
    m = NNP.read(potentials["mixed64x4096"])
    try:
        de, df = mixed_errors(m.compute(s, FLAGS | NNP_PREC_MIXED), ref, s.natoms)
        print(f"C3 mixed: {de:.5f} meV/atom, {df:.4f} meV/A RMS")
        assert de <= 1.0 and df <= 20.0
    finally:
        m.close()


def test_plain_cell_per_atom_energies(potentials):
    s = synth.plain4096()[0]
    ref = ob.OracleModel(potentials["plain4096"]).compute(s.types, s.pos, s.cell, OFLAGS)
    m = NNP.read(potentials["plain4096"])
    try:
        check_fp64(m.compute(s, FLAGS), ref, "C1")
    finally:
        m.close()


@pytest.mark.parametrize("pipeline", [0, 1], ids=["separate", "fused"])
def test_batch_fp32_leg(potentials, pipeline):
    """24 cells of the uneven triclinic batch in the FP32 form (FP32 descriptors and force terms,
    BF16 network products in the separate pipeline, FP32 in-wave networks in the fused one)."""
    strucs = synth.batch512(24)
    o = ob.OracleModel(potentials["batch512"])
    m = NNP.read(potentials["batch512"])
    try:
        m.set_option("pipeline", pipeline)
        out = m.compute_batch(strucs, FLAGS | NNP_PREC_MIXED)
        worst = (0.0, 0.0)
        for s, res in zip(strucs, out):
            de, df = mixed_errors(res, o.compute(s.types, s.pos, s.cell, OFLAGS), s.natoms)
            worst = (max(worst[0], de), max(worst[1], df))
            assert de <= 1.0 and df <= 20.0, (s.natoms, de, df)
        print(f"FP32 leg ({'fused' if pipeline else 'separate'}): worst {worst[0]:.5f} meV/atom, {worst[1]:.4f} meV/A")
    finally:
        m.close()


@pytest.mark.parametrize("waves", [1, 2])
def test_precision_cell_force_waves(potentials, precision_cell, waves):
    """The moment force kernel with one or two wavefronts per centre (option force_waves), which
    reads the moment tables the descriptor kernel stored, against the oracle: FP64 within 1e-10,
    mixed within the gates; a model that sums its moments itself (the option moments without the
    descriptor bit: triplet descriptors, moment forces) gives the same forces to rounding."""
    s, ref = precision_cell
    m = NNP.read(potentials["mixed64x4096"])
    try:
        m.set_option("force_waves", waves)
        got = m.compute(s, FLAGS)
        check_fp64(got, ref, f"C3 fp64 force_waves {waves}")
        de, df = mixed_errors(m.compute(s, FLAGS | NNP_PREC_MIXED), ref, s.natoms)
        assert de <= 1.0 and df <= 20.0, (waves, de, df)
        m.set_option("moments", 2)
        own = m.compute(s, FLAGS)
        check_fp64(own, ref, f"C3 fp64 force_waves {waves}, moments summed in the force kernel")
        assert np.linalg.norm(own[2] - got[2], axis=1).max() <= 1e-12 * np.linalg.norm(got[2], axis=1).max()
    finally:
        m.close()
