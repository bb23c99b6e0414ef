"""The energy-training inner loop on the GPU (nnp_train_*: NNPTE::error's training part,
src/nnp/nnpte.cpp:869-927, with Cost::grad, src/ml/nn.cpp:1554-1589) against the CPU oracle's
restatement (itself pinned by finite differences and closed forms in test_oracle_train.py):
losses within 1e-10 relative, gradients within 1e-10 of their largest component, for every
loss function; plus the properties the loop must have at sizes the oracle does not reach."""
import numpy as np
import pytest

from atomnn_amd import synth
from atomnn_amd.nnp import NNP, NNPError, Train, NNP_ATOM_ENERGY, LOSS_MSE, LOSS_MAE, LOSS_HUBER, LOSS_ASINH
import oracle_binding as ob

pytestmark = pytest.mark.gpu

TOL = 1e-10


def training_set(seed=3):
    rng = np.random.default_rng(seed)
    out = []
    for k, n in enumerate((40, 17, 64, 9, 33, 120, 25, 58)):
        cell = synth.strained_cell(n, 0.05, 0.08, rng) if k != 3 else None
        box = cell if cell is not None else synth.cubic_cell(n, 0.05)
        pos = synth.random_positions(n, box, 2.0, 100 + k)
        out.append(synth.Structure(synth.species_exact(n, [0.5, 0.5], rng), pos, cell))
    eref = np.array([rng.normal(-4.0, 0.5) * s.natoms for s in out])
    return out, eref


@pytest.fixture(scope="module")
def pair(potentials):
    strucs, eref = training_set()
    m = NNP.read(potentials["plain4096"])
    t = Train(m, strucs, eref)
    o = ob.OracleTrain(ob.OracleModel(potentials["plain4096"]), strucs, eref)
    yield m, t, o, strucs, eref
    t.close()
    m.close()


def test_parameters_match_the_oracle(pair):
    m, t, o, strucs, eref = pair
    assert t.nparams() == o.params().size
    assert t.nparams(0) == o.nparams(0) and t.nparams(1) == o.nparams(1)
    assert np.array_equal(t.params(), o.params())


@pytest.mark.parametrize("loss,delta", [(LOSS_MSE, 1.0), (LOSS_MAE, 1.0), (LOSS_HUBER, 0.02), (LOSS_ASINH, 0.02)])
def test_error_and_gradient_match_the_oracle(pair, loss, delta):
    m, t, o, strucs, eref = pair
    rng = np.random.default_rng(loss + 11)
    x = t.params() + rng.normal(0, 0.02, t.nparams())
    for batch in ([0, 1, 2, 3, 4, 5, 6, 7], [5, 2, 7], [3]):
        e, g = t.error(x, batch, loss, delta)
        eo, go = o.error(x, batch, loss, delta)
        assert abs(e[0] - eo[0]) <= TOL * abs(eo[0]), (batch, e, eo)
        assert abs(e[1] - eo[1]) <= TOL * abs(eo[1])
        assert np.abs(g - go).max() <= TOL * np.abs(go).max(), batch
        # without the gradient: the same errors
        e2, _ = t.error(x, batch, loss, delta, grad=False)
        assert np.array_equal(e2, e)


def test_gradient_is_additive_over_structures(pair):
    m, t, o, strucs, eref = pair
    x = t.params()
    _, g = t.error(x, list(range(8)))
    parts = sum(t.error(x, [b])[1] for b in range(8))
    assert np.abs(g - parts).max() <= 1e-12 * np.abs(g).max()


def test_training_leaves_the_model_unchanged_and_descends(pair):
    m, t, o, strucs, eref = pair
    s = strucs[0]
    E0, Ea0, _ = m.compute(s, NNP_ATOM_ENERGY)
    x = t.params()
    batch = list(range(8))
    l0, g = t.error(x, batch)
    for _ in range(5):
        x = x - 0.5 * g / np.linalg.norm(g) * 1e-2
        l1, g = t.error(x, batch)
    assert l1[0] < l0[0]
    E1, Ea1, _ = m.compute(s, NNP_ATOM_ENERGY)
    assert E1 == E0 and np.array_equal(Ea0, Ea1)


@pytest.mark.parametrize("one_pass", [1, 0])
def test_large_batch_gradient_in_chunks(potentials, one_pass):
    """32768 atoms of a three-species potential (84-64-64-1 networks): one pass over
    single-structure tiles, or two passes with the rows summed in chunks of tiles; either equals
    the sum of the per-structure gradients, and the two agree."""
    rng = np.random.default_rng(19)
    strucs = [s for s in synth.mixed64x4096(8)]
    eref = np.array([rng.normal(-3.0, 0.2) * s.natoms for s in strucs])
    m = NNP.read(potentials["mixed64x4096"])
    try:
        m.set_option("train_one_pass", one_pass)
        if not one_pass:
            m.set_option("train_chunk_elems", 1 << 22)  # the rows of 2048 tiles in five chunks
        t = Train(m, strucs, eref)
        x = t.params()
        err, g = t.error(x, list(range(8)), LOSS_HUBER, 0.05)
        err2, g2 = t.error(x, list(range(8)), LOSS_HUBER, 0.05)  # tiles reused
        assert np.array_equal(err, err2) and np.array_equal(g, g2)
        parts_e = np.zeros(2)
        parts_g = np.zeros_like(g)
        for b in range(8):
            e, gb = t.error(x, [b], LOSS_HUBER, 0.05)
            parts_e += e
            parts_g += gb
        assert np.allclose(err, parts_e, rtol=1e-12)
        assert np.abs(g - parts_g).max() <= 1e-11 * np.abs(g).max()
        m.set_option("train_one_pass", 1 - one_pass)
        e3, g3 = t.error(x, list(range(8)), LOSS_HUBER, 0.05)
        assert np.allclose(err, e3, rtol=1e-12)
        assert np.abs(g - g3).max() <= 1e-11 * np.abs(g).max()
        t.close()
    finally:
        m.close()


def test_many_small_structures_take_two_passes(potentials):
    """40 structures of 3-6 atoms: single-structure tiles would mostly be padding, so the
    species tiles across structures are used; the oracle bar holds."""
    rng = np.random.default_rng(29)
    strucs = []
    for k in range(40):
        n = int(rng.integers(3, 7))
        box = synth.cubic_cell(n, 0.05)
        strucs.append(synth.Structure(rng.integers(0, 2, n).astype(np.int32), synth.random_positions(n, box, 1.8, 300 + k),
                                      box if k % 3 else None))
    eref = np.array([rng.normal(-4.0, 0.5) * s.natoms for s in strucs])
    m = NNP.read(potentials["plain4096"])
    try:
        t = Train(m, strucs, eref)
        o = ob.OracleTrain(ob.OracleModel(potentials["plain4096"]), strucs, eref)
        x = t.params() + rng.normal(0, 0.02, t.nparams())
        e, g = t.error(x, list(range(40)), LOSS_MSE)
        eo, go = o.error(x, list(range(40)), LOSS_MSE)
        assert abs(e[0] - eo[0]) <= TOL * abs(eo[0])
        assert np.abs(g - go).max() <= TOL * np.abs(go).max()
        t.close()
    finally:
        m.close()


def test_bad_arguments(pair):
    m, t, o, strucs, eref = pair
    x = t.params()
    with pytest.raises(NNPError, match="out of range"):
        t.error(x, [0, 8])
    with pytest.raises(NNPError, match="unknown loss"):
        t.error(x, [0], loss=7)
    with pytest.raises(NNPError, match="delta"):
        t.error(x, [0], LOSS_HUBER, 0.0)
    with pytest.raises(ValueError):
        t.error(x[:-1], [0])


# ---------------------------------------------------------------- force training
def reference_forces(om, strucs, seed):
    """Oracle forces at the model's parameters, perturbed (so the force loss is not zero)."""
    rng = np.random.default_rng(seed)
    out = []
    for s in strucs:
        _, _, f = om.compute(s.types, s.pos, s.cell, ob.F_FORCE)
        out.append(f + rng.normal(0, 0.05, f.shape))
    return out


@pytest.fixture(scope="module")
def fpair(potentials):
    strucs, eref = training_set(5)
    m = NNP.read(potentials["plain4096"])
    om = ob.OracleModel(potentials["plain4096"])
    fref = reference_forces(om, strucs, 8)
    t = Train(m, strucs, eref)
    t.set_forces(fref)
    o = ob.OracleTrain(om, strucs, eref)
    o.set_forces(fref)
    yield m, t, o, strucs, eref
    t.close()
    m.close()


def check_force_parity(t, o, x, batch, loss, delta, beta, ai):
    e, g = t.error_force(x, batch, loss, delta, beta, all_images=ai)
    eo, go = o.error_force(x, batch, loss, delta, beta, all_images=ai)
    for k in range(3):
        assert abs(e[k] - eo[k]) <= TOL * abs(eo[k]), (k, e, eo)
    assert np.abs(g - go).max() <= TOL * np.abs(go).max()
    e2, _ = t.error_force(x, batch, loss, delta, beta, all_images=ai, grad=False)
    assert np.array_equal(e2, e)


@pytest.mark.parametrize("loss,ai", [(LOSS_MSE, False), (LOSS_HUBER, False), (LOSS_MSE, True)])
def test_force_training_matches_the_oracle(fpair, loss, ai):
    m, t, o, strucs, eref = fpair
    x = t.params() + np.random.default_rng(21).normal(0, 0.02, t.nparams())
    for batch in ([0, 1, 2, 3, 4, 5, 6, 7], [6, 3]):
        check_force_parity(t, o, x, batch, loss, 0.05, 2.5, ai)


def test_force_training_triplet_path(potentials, fpair):
    """With the moment expansion switched off the descriptor tangents take the pair loops; the
    same parity bar."""
    _, _, o, strucs, eref = fpair
    m2 = NNP.read(potentials["plain4096"])
    try:
        m2.set_option("moments", 0)
        t2 = Train(m2, strucs, eref)
        t2.set_forces(reference_forces(ob.OracleModel(potentials["plain4096"]), strucs, 8))
        x = t2.params() + np.random.default_rng(22).normal(0, 0.02, t2.nparams())
        check_force_parity(t2, o, x, [0, 3, 5, 7], LOSS_MSE, 1.0, 2.5, False)
        t2.close()
    finally:
        m2.close()


def test_force_training_cache_gives_the_same_results(potentials, fpair):
    """Chunks whose neighbour rows and descriptors are kept between calls, and chunks rebuilt
    every call (cache size 0), give identical losses and gradients."""
    _, t, _, strucs, eref = fpair
    x = t.params() + np.random.default_rng(31).normal(0, 0.02, t.nparams())
    e1, g1 = t.error_force(x, [1, 4, 6], LOSS_MSE, 1.0, 2.0)
    e1b, g1b = t.error_force(x, [1, 4, 6], LOSS_MSE, 1.0, 2.0)  # the cached chunk
    m2 = NNP.read(potentials["plain4096"])
    try:
        m2.set_option("train_cache_atoms", 0)
        t2 = Train(m2, strucs, eref)
        t2.set_forces(reference_forces(ob.OracleModel(potentials["plain4096"]), strucs, 8))
        e2, g2 = t2.error_force(x, [1, 4, 6], LOSS_MSE, 1.0, 2.0)
        e2b, g2b = t2.error_force(x, [1, 4, 6], LOSS_MSE, 1.0, 2.0)
        for a, b in ((e1, e1b), (e1, e2), (e1, e2b), (g1, g1b), (g1, g2), (g1, g2b)):
            assert np.array_equal(a, b)
        t2.close()
    finally:
        m2.close()


def test_force_training_without_force_weight_is_energy_training(fpair):
    m, t, o, strucs, eref = fpair
    x = t.params()
    e, g = t.error_force(x, list(range(8)), LOSS_ASINH, 0.05, 0.0)
    ee, ge = t.error(x, list(range(8)), LOSS_ASINH, 0.05)
    assert abs(e[0] - ee[0]) <= 1e-12 * abs(ee[0]) and abs(e[1] - ee[1]) <= 1e-12 * abs(ee[1])
    assert np.abs(g - ge).max() <= 1e-12 * np.abs(ge).max()


def test_force_training_variants(tmp_path):
    """Generic kernels: another activation (its second derivative), radial and non-separable
    angular basis functions."""
    pot = synth.make_potential(["A", "B"], 5.0, 4, [0.02, 0.2], [1.0, 2.0], [1, -1], [12, 9],
                               np.random.default_rng(13))
    pot.neuron, pot.phi_r, pot.phi_a = "SOFTPLUS", "TANH", "IPOWS"
    path = str(tmp_path / "gen.ann")
    pot.write(path)
    strucs, eref = training_set(9)
    om = ob.OracleModel(path)
    fref = reference_forces(om, strucs, 10)
    m = NNP.read(path)
    try:
        t = Train(m, strucs, eref)
        t.set_forces(fref)
        o = ob.OracleTrain(om, strucs, eref)
        o.set_forces(fref)
        x = t.params() + np.random.default_rng(2).normal(0, 0.02, t.nparams())
        check_force_parity(t, o, x, [0, 2, 4, 6, 7], LOSS_MSE, 1.0, 1.5, False)
        t.close()
    finally:
        m.close()


def test_force_training_wide_networks(potentials):
    """84-64-64-1 networks of three species (the tile kernel's largest LDS): a 64-atom cell
    against the oracle, and 2 x 4096 atoms equal to the sum of the per-structure calls."""
    rng = np.random.default_rng(23)
    big = synth.mixed64x4096(2)
    cell = synth.cubic_cell(64, 0.06)
    small = synth.Structure(synth.species_exact(64, [0.4, 0.3, 0.3], rng), synth.random_positions(64, cell, 1.8, 5),
                            cell)
    strucs = [small] + big
    eref = np.array([rng.normal(-3.0, 0.2) * s.natoms for s in strucs])
    om = ob.OracleModel(potentials["mixed64x4096"])
    fref = reference_forces(om, [small], 3) + [rng.normal(0, 0.3, (s.natoms, 3)) for s in big]
    m = NNP.read(potentials["mixed64x4096"])
    try:
        t = Train(m, strucs, eref)
        t.set_forces(fref)
        x = t.params()
        o = ob.OracleTrain(om, [small], eref[:1])
        o.set_forces(fref[:1])
        check_force_parity(t, o, x, [0], LOSS_MSE, 1.0, 2.0, False)
        err, g = t.error_force(x, [1, 2], LOSS_HUBER, 0.05, 2.0)
        e1, g1 = t.error_force(x, [1], LOSS_HUBER, 0.05, 2.0)
        e2, g2 = t.error_force(x, [2], LOSS_HUBER, 0.05, 2.0)
        assert np.allclose(err, e1 + e2, rtol=1e-12)
        assert np.abs(g - (g1 + g2)).max() <= 1e-11 * np.abs(g).max()
        t.close()
    finally:
        m.close()


def test_force_training_bad_arguments(pair):
    m, t, o, strucs, eref = pair
    x = t.params()
    with pytest.raises(NNPError, match="reference forces"):
        t.error_force(x, [0])
    t.set_forces([np.zeros((s.natoms, 3)) for s in strucs])
    with pytest.raises(NNPError, match="beta"):
        t.error_force(x, [0], beta=-1.0)
    with pytest.raises(NNPError, match="out of range"):
        t.error_force(x, [9])
