"""The batched form's persistent fused descriptor + network + force kernel (pipeline 1).

Same parity bar as the separate kernels: FP64 within 1e-10 of the oracle, mixed precision
within 1 meV/atom and 20 meV/A RMS; results independent of the batch and of graph replay;
the autotuned launch geometry does not change results.
"""
import numpy as np
import pytest

from atomnn_amd import synth
from atomnn_amd.nnp import NNP, NNP_FORCE, NNP_PREC_MIXED, NNP_VACUUM_ENERGY
import oracle_binding as ob

pytestmark = pytest.mark.gpu
TOL = 1e-10


def rel_e(E, Eo, Eao):
    return abs(E - Eo) / max(abs(Eo), np.abs(Eao).sum())


def rel_f(F, Fo):
    return np.linalg.norm(F - Fo, axis=1).max() / np.linalg.norm(Fo, axis=1).max()


@pytest.fixture(scope="module", params=[0, 3], ids=["triplets", "moments"])
def fused_models(potentials, request):
    out = {}
    for k in ("plain64", "plain4096", "mixed64x4096"):
        m = NNP.read(potentials[k])
        m.set_option("pipeline", 1)
        m.set_option("moments", request.param)
        out[k] = (m, ob.OracleModel(potentials[k]))
    yield out
    for m, _ in out.values():
        m.close()


@pytest.mark.parametrize("name", ["plain64", "plain4096"])
def test_fused_parity_fp64(fused_models, name):
    m, o = fused_models[name]
    strucs = synth.plain64() if name == "plain64" else synth.plain4096() + synth.batch512(6)
    res = m.compute_batch(strucs)
    for s, (E, Ea, F) in zip(strucs, res):
        Eo, Eao, Fo = o.compute(s.types, s.pos, s.cell)
        assert rel_e(E, Eo, Eao) <= TOL
        assert rel_f(F, Fo) <= TOL


def test_fused_mixed_accuracy(fused_models):
    m, o = fused_models["mixed64x4096"]
    s = synth.mixed64x4096(1)[0]
    E, Ea, F = m.compute(s, NNP_FORCE | NNP_VACUUM_ENERGY | NNP_PREC_MIXED)
    Eo, Eao, Fo = o.compute(s.types, s.pos, s.cell)
    assert abs(E - Eo) / s.natoms <= 1e-3
    assert np.sqrt(np.mean(np.sum((F - Fo) ** 2, axis=1))) <= 20e-3


def test_fused_batch_single_and_geometry_independent(fused_models):
    m, _ = fused_models["plain4096"]
    strucs = synth.batch512(5)
    res = m.compute_batch(strucs)
    for s, (E, Ea, F) in zip(strucs, res):
        E1, Ea1, F1 = m.compute(s)
        np.testing.assert_array_equal(Ea1, Ea)
        np.testing.assert_array_equal(F1, F)
    m.set_option("fused_blocks_per_cu", 1)
    res1 = m.compute_batch(strucs)
    m.set_option("fused_blocks_per_cu", 0)
    for (E, Ea, F), (E1, Ea1, F1) in zip(res, res1):
        np.testing.assert_array_equal(Ea1, Ea)
        np.testing.assert_array_equal(F1, F)


@pytest.mark.parametrize("mixed", [False, True])
def test_fused_graph_replays(fused_models, mixed):
    m, _ = fused_models["mixed64x4096"]
    flags = NNP_FORCE | NNP_VACUUM_ENERGY | (NNP_PREC_MIXED if mixed else 0)
    R = m.resident(synth.mixed64x4096(2), flags)
    R.run()
    eager = R.fetch()
    assert m.info()["fused_blocks_per_cu"] > 0
    for _ in range(3):
        R.run()
        got = R.fetch()
        for (E0, Ea0, F0), (E1, Ea1, F1) in zip(eager, got):
            assert E0 == E1
            np.testing.assert_array_equal(F0, F1)
