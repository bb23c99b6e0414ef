"""The oracle's restatement of the energy-training inner loop (NNPTE::error, training part,
src/nnp/nnpte.cpp:869-927, and Cost::grad, src/ml/nn.cpp:1554-1589), pinned without the GPU:

* the parameter vector follows ANN operator>> (nn.cpp:797-831): every layer's biases, then
  every weight matrix column-major, checked against the matrices the potential file holds;
* the structure energy is the sum of the atomic network outputs without the type energies,
  checked against the oracle's per-atom energies of NNP::compute (no vacuum energy);
* the loss values follow the four loss functions' closed forms in dE = (E - E_ref) / N;
* the gradient is the derivative of the loss: central finite differences on parameters of
  every layer (biases and weights) of every species, for every loss.
"""
import numpy as np
import pytest

from atomnn_amd import synth
import oracle_binding as ob


@pytest.fixture(scope="module")
def setup(tmp_path_factory):
    pot = synth.make_potential(["A", "B"], 5.0, 4, [0.02, 0.2], [1.0, 2.0], [1, -1], [8, 6],
                               np.random.default_rng(31))
    path = str(tmp_path_factory.mktemp("train") / "train.ann")
    pot.write(path)
    q = pot.quantized()
    m = ob.OracleModel(path)
    rng = np.random.default_rng(5)
    strucs = []
    for k, n in enumerate((9, 14, 6, 11)):
        cell = synth.cubic_cell(n, 0.04) if k != 2 else None
        box = cell if cell is not None else synth.cubic_cell(n, 0.04)
        pos = synth.random_positions(n, box, 1.9, 40 + k)
        strucs.append(synth.Structure(rng.integers(0, 2, n).astype(np.int32), pos, cell))
    eref = np.array([rng.normal(-0.5, 0.3) * s.natoms for s in strucs])
    return m, q, strucs, eref


def test_parameter_order_follows_the_network_matrices(setup):
    m, q, strucs, eref = setup
    t = ob.OracleTrain(m, strucs, eref)
    x = t.params()
    off = 0
    for s in range(2):
        Ws, bs = q.weights[s], q.biases[s]
        assert t.nparams(s) == sum(W.size + b.size for W, b in zip(Ws, bs))
        c = off
        for b in bs:
            assert np.array_equal(x[c:c + b.size], b)
            c += b.size
        for W in Ws:
            n_l, n_prev = W.shape
            assert np.array_equal(x[c:c + W.size], W.T.reshape(-1))  # column-major
            assert x[c + 1 * n_l + 0] == W[0, 1]
            c += W.size
        off = c
    assert off == x.size


def test_energy_and_mse_loss_from_atomic_outputs(setup):
    m, q, strucs, eref = setup
    t = ob.OracleTrain(m, strucs, eref)
    x = t.params()
    batch = [2, 0, 3]
    err, _ = t.error(x, batch, ob.LOSS_MSE, grad=False)
    loss = rmse = 0.0
    for b in batch:
        s = strucs[b]
        _, ea, _ = m.compute(s.types, s.pos, s.cell, ob.F_ATOM_ENERGY)  # NNP::compute: outputs only
        dE = (ea.sum() - eref[b]) / s.natoms
        loss += 0.5 * dE * dE
        rmse += dE * dE
    assert err[0] == pytest.approx(loss, rel=1e-12)
    assert err[1] == pytest.approx(rmse, rel=1e-12)


@pytest.mark.parametrize("loss,delta", [(ob.LOSS_MAE, 1.0), (ob.LOSS_HUBER, 0.05), (ob.LOSS_ASINH, 0.05)])
def test_loss_closed_forms(setup, loss, delta):
    m, q, strucs, eref = setup
    t = ob.OracleTrain(m, strucs, eref)
    x = t.params()
    err, _ = t.error(x, [0, 1, 2, 3], loss, delta, grad=False)
    want = 0.0
    for b, s in enumerate(strucs):
        _, ea, _ = m.compute(s.types, s.pos, s.cell, ob.F_ATOM_ENERGY)
        dE = (ea.sum() - eref[b]) / s.natoms
        a = dE / delta
        if loss == ob.LOSS_MAE:
            want += abs(dE)
        elif loss == ob.LOSS_HUBER:
            want += delta * delta * (np.sqrt(1 + a * a) - 1)
        else:
            want += delta * delta * (1 - np.sqrt(1 + a * a) + a * np.arcsinh(a))
    assert err[0] == pytest.approx(want, rel=1e-12)


def _probe_indices(t):
    """A few parameters of every block: each layer's biases and weights of each species."""
    idx, off = [], 0
    rng = np.random.default_rng(2)
    for s in range(2):
        n = t.nparams(s)
        idx += [off, off + 1, off + n // 3, off + n // 2, off + n - 1]
        idx += list(off + rng.integers(0, n, 4))
        off += n
    return sorted(set(int(i) for i in idx))


@pytest.mark.parametrize("loss,delta", [(ob.LOSS_MSE, 1.0), (ob.LOSS_MAE, 1.0), (ob.LOSS_HUBER, 0.05),
                                        (ob.LOSS_ASINH, 0.05)])
def test_gradient_matches_finite_differences(setup, loss, delta):
    m, q, strucs, eref = setup
    t = ob.OracleTrain(m, strucs, eref)
    rng = np.random.default_rng(9)
    x = t.params() + rng.normal(0, 0.05, t.params().size)  # away from zero biases
    batch = [3, 1, 0, 2]
    _, g = t.error(x, batch, loss, delta)
    h = 1e-6
    for k in _probe_indices(t):
        xp, xm = x.copy(), x.copy()
        xp[k] += h
        xm[k] -= h
        fd = (t.error(xp, batch, loss, delta, grad=False)[0][0] - t.error(xm, batch, loss, delta, grad=False)[0][0]) / (2 * h)
        assert g[k] == pytest.approx(fd, rel=2e-5, abs=1e-9), k


def test_gradient_sums_over_structures(setup):
    """The batch gradient is the sum of the single-structure gradients (per-atom accumulation)."""
    m, q, strucs, eref = setup
    t = ob.OracleTrain(m, strucs, eref)
    x = t.params()
    _, g = t.error(x, [0, 1, 2, 3])
    parts = sum(t.error(x, [b])[1] for b in range(4))
    assert np.allclose(g, parts, rtol=1e-12, atol=1e-15)


# ---------------------------------------------------------------- force training
def test_activation_second_derivatives():
    """a'' of every activation against central differences of its a' (the reference's AFFPBP2
    where it is the derivative, the true second derivative where it is not)."""
    h = 1e-5
    for name in ob.NEURONS:
        for z in (-3.1, -1.7, -0.6, -0.23, 0.31, 0.8, 1.9, 2.7):
            d2 = ob.activation(name, z)[2]
            fd = (ob.activation(name, z + h)[1] - ob.activation(name, z - h)[1]) / (2 * h)
            assert d2 == pytest.approx(fd, rel=1e-7, abs=1e-8), (name, z)


def _force_setup(tmp_path_factory, neuron="TANH", phi_r="GAUSSIAN", phi_a="GAUSS", seed=31):
    pot = synth.make_potential(["A", "B"], 5.0, 4, [0.02, 0.2], [1.0, 2.0], [1, -1], [7, 5],
                               np.random.default_rng(seed))
    pot.neuron, pot.phi_r, pot.phi_a = neuron, phi_r, phi_a
    path = str(tmp_path_factory.mktemp("trainf") / "trainf.ann")
    pot.write(path)
    m = ob.OracleModel(path)
    rng = np.random.default_rng(seed + 1)
    strucs = []
    # a periodic cell with L < 2 rc (several images, min-image reactions matter), a larger
    # cell and an open cluster
    for k, (n, dens) in enumerate(((6, 0.05), (13, 0.04), (7, 0.05))):
        cell = synth.cubic_cell(n, dens) if k != 2 else None
        box = cell if cell is not None else synth.cubic_cell(n, dens)
        pos = synth.random_positions(n, box, 1.9, 60 + k + seed)
        strucs.append(synth.Structure(rng.integers(0, 2, n).astype(np.int32), pos, cell))
    eref = np.array([rng.normal(-0.5, 0.3) * s.natoms for s in strucs])
    fref = []
    for s in strucs:
        _, _, f = m.compute(s.types, s.pos, s.cell, ob.F_FORCE)
        fref.append(f + rng.normal(0, 0.05, f.shape))
    return m, strucs, eref, fref


@pytest.fixture(scope="module")
def fsetup(tmp_path_factory):
    return _force_setup(tmp_path_factory)


def test_force_loss_values(fsetup):
    """err = {energy loss + beta/2 sum |fdiff|^2, sum dE^2, sum |fdiff|^2} with
    fdiff = (F - F_ref) / N and F the oracle's NNP::compute forces."""
    m, strucs, eref, fref = fsetup
    t = ob.OracleTrain(m, strucs, eref)
    t.set_forces(fref)
    beta = 3.0
    for ai in (False, True):
        err, _ = t.error_force(t.params(), [2, 0, 1], ob.LOSS_MSE, 1.0, beta, all_images=ai, grad=False)
        le = f2 = e2 = 0.0
        for b in (2, 0, 1):
            s = strucs[b]
            flags = ob.F_FORCE | ob.F_ATOM_ENERGY | (ob.F_ALL_IMAGES if ai else 0)
            _, ea, f = m.compute(s.types, s.pos, s.cell, flags)
            dE = (ea.sum() - eref[b]) / s.natoms
            le += 0.5 * dE * dE
            e2 += dE * dE
            f2 += float(np.sum(((f - fref[b]) / s.natoms) ** 2))
        assert err[0] == pytest.approx(le + 0.5 * beta * f2, rel=1e-12)
        assert err[1] == pytest.approx(e2, rel=1e-12)
        assert err[2] == pytest.approx(f2, rel=1e-12)


def test_force_gradient_without_force_weight_is_the_energy_gradient(fsetup):
    m, strucs, eref, fref = fsetup
    t = ob.OracleTrain(m, strucs, eref)
    t.set_forces(fref)
    x = t.params()
    g0 = t.error_force(x, [0, 1, 2], ob.LOSS_HUBER, 0.1, 0.0)[1]
    ge = t.error(x, [0, 1, 2], ob.LOSS_HUBER, 0.1)[1]
    assert np.abs(g0 - ge).max() <= 1e-13 * np.abs(ge).max()


def _fd_check(t, x, batch, loss, delta, beta, ai, rel):
    _, g = t.error_force(x, batch, loss, delta, beta, all_images=ai)
    h = 1e-6
    for k in _probe_indices(t):
        xp, xm = x.copy(), x.copy()
        xp[k] += h
        xm[k] -= h
        lp = t.error_force(xp, batch, loss, delta, beta, all_images=ai, grad=False)[0][0]
        lm = t.error_force(xm, batch, loss, delta, beta, all_images=ai, grad=False)[0][0]
        assert g[k] == pytest.approx((lp - lm) / (2 * h), rel=rel, abs=1e-9), k


@pytest.mark.parametrize("loss,ai", [(ob.LOSS_MSE, False), (ob.LOSS_ASINH, False), (ob.LOSS_MSE, True)])
def test_force_gradient_matches_finite_differences(fsetup, loss, ai):
    """The gradient of energy + force loss against central differences of the loss, with the
    reactions on minimum images only (the reference's rule) and on every image."""
    m, strucs, eref, fref = fsetup
    t = ob.OracleTrain(m, strucs, eref)
    t.set_forces(fref)
    x = t.params() + np.random.default_rng(4).normal(0, 0.05, t.params().size)
    _fd_check(t, x, [1, 2, 0], loss, 0.05, 2.5, ai, 2e-5)


@pytest.mark.parametrize("neuron,phi_r,phi_a", [("SOFTPLUS", "TANH", "IPOWS"), ("SINSIG", "MISH", "SECHP"),
                                                ("SOFTSIGN", "LOGCOSH", "SECHS"), ("IQUAD", "SWISH", "IPOWP")])
def test_force_gradient_variants(tmp_path_factory, neuron, phi_r, phi_a):
    """Other activations (second derivatives included) and basis functions (per-function force
    shares of non-separable angular factors)."""
    m, strucs, eref, fref = _force_setup(tmp_path_factory, neuron, phi_r, phi_a, seed=47)
    t = ob.OracleTrain(m, strucs, eref)
    t.set_forces(fref)
    x = t.params() + np.random.default_rng(6).normal(0, 0.05, t.params().size)
    _fd_check(t, x, [0, 1, 2], ob.LOSS_MSE, 1.0, 1.5, False, 5e-5)
