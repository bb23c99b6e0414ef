"""Pins the oracle's remaining basis and activation variants (radial TANH, SOFTPLUS, LOGCOSH,
SWISH, MISH; angular IPOWP, IPOWS, SECHP, SECHS; the 17 neurons; output scaling).

Each variant is checked against an independent closed form of the function it stands for
(numpy/scipy expressions written from the mathematics, not from the oracle's code) and its
derivative path against finite differences:
  * radial: a dimer, G2_m = phi_m(-eta (r - rs)) fc(r);
  * angular: an asymmetric trimer, G4 = |(1 + lambda cos)/2|^zeta fc(r1) fc(r2) D(x1, x2);
  * neurons: a one-unit network, out = outb + outw act(x); dO/dI by finite differences;
  * forces = -dE/dr on an L > 2 rc cell (the reference's min-image rule is inert there).
SOFTPLUS and LOGCOSH use the reference's truncated series for log(1 + x)
(math/special.cpp:131-135), within 2e-6 of log1p on [0, 1]: those closed forms carry that
tolerance, which still fails a dropped term or a wrong sign.
"""
import math

import numpy as np
import pytest
from scipy.special import expit

from atomnn_amd import synth
import oracle_binding as ob
from test_oracle_pinning import fc_cos, tiny_potential, write, fd_forces

LN2 = math.log(2.0)
RC = 6.0

RADIAL = {
    "GAUSSIAN": (lambda x, eta, dr: math.exp(-eta * dr * dr), 1e-14),
    "TANH": (lambda x, eta, dr: 0.5 * (math.tanh(x) + 1.0), 1e-14),
    "SOFTPLUS": (lambda x, eta, dr: math.log1p(math.exp(x)), 3e-6),
    "LOGCOSH": (lambda x, eta, dr: 0.5 * math.log1p(math.exp(2.0 * x)), 3e-6),
    "SWISH": (lambda x, eta, dr: x * expit(x), 1e-14),
    "MISH": (lambda x, eta, dr: x * math.tanh(math.log1p(math.exp(x))), 1e-13),
}

ANGULAR = {
    "GAUSS": lambda x1, x2: math.exp(-(x1 + x2)),
    "IPOWP": lambda x1, x2: 1.0 / ((1.0 + x1) * (1.0 + x2)),
    "IPOWS": lambda x1, x2: 1.0 / (1.0 + x1 + x2),
    "SECHP": lambda x1, x2: 1.0 / (math.cosh(x1) * math.cosh(x2)),
    "SECHS": lambda x1, x2: 1.0 / math.cosh(x1 + x2),
}

NEURONS = {
    "LINEAR": (lambda z: z, 1e-15),
    "SIGMOID": (lambda z: expit(z), 1e-15),
    "TANH": (lambda z: math.tanh(z), 1e-15),
    "ISRU": (lambda z: z / math.sqrt(1.0 + z * z), 1e-15),
    "ARCTAN": (lambda z: 2.0 / math.pi * math.atan(z), 1e-15),
    "SOFTSIGN": (lambda z: z / (1.0 + abs(z)), 1e-15),
    "RELU": (lambda z: max(z, 0.0), 1e-15),
    "SOFTPLUS": (lambda z: math.log1p(math.exp(z)) - LN2, 3e-6),
    "ELU": (lambda z: z if z >= 0 else math.expm1(z), 1e-15),
    "GELU": (lambda z: z * expit(1.702 * z), 1e-14),
    "SWISH": (lambda z: z * expit(z), 1e-14),
    "MISH": (lambda z: z * math.tanh(math.log1p(math.exp(z))), 1e-13),
    "TANHRE": (lambda z: z if z > 0 else math.tanh(z), 1e-15),
    # the reference's LOGCOSH neuron is log((1 + e^2z)/2)/2, not log cosh z (nn.cpp:629-637)
    "LOGCOSH": (lambda z: 0.5 * (math.log1p(math.exp(2.0 * z)) - LN2), 3e-6),
    "IQUAD": (lambda z: 0.5 * (z + z * z / math.sqrt(1.0 + z * z)), 1e-14),
    "IFABS": (lambda z: 0.5 * (z + z * z / (1.0 + abs(z))), 1e-14),
    "SINSIG": (lambda z: z * math.sin(0.5 * math.pi * expit(z)), 1e-14),
}


def one_type(phi_r="GAUSSIAN", phi_a="GAUSS", radial=None, angular=None, hidden=(4,), neuron="TANH", seed=0):
    pot = tiny_potential(["A"], rc=RC, radial=radial, angular=angular, hidden=hidden, seed=seed)
    pot.phi_r, pot.phi_a, pot.neuron = phi_r, phi_a, neuron
    return pot


@pytest.mark.parametrize("phi", list(RADIAL))
def test_radial_variants_dimer_closed_form(tmp_path, phi):
    radial = [(1.0, 0.5), (2.5, 1.2), (4.0, 0.8), (5.5, 2.0), (3.0, 0.05)]
    m = ob.OracleModel(write(one_type(phi_r=phi, radial=radial), tmp_path))
    f, tol = RADIAL[phi]
    for r in (1.5, 3.0, 4.7, 5.9):
        pos = np.array([[0.0, 0.0, 0.0], [r, 0.0, 0.0]])
        G = m.symm(np.zeros(2, dtype=np.int32), pos, None)
        for k, (rs, eta) in enumerate(radial):
            want = f(-eta * (r - rs), eta, r - rs) * fc_cos(r, RC)
            assert abs(G[0, k] - want) <= tol * max(1.0, abs(want)), (phi, r, k, G[0, k], want)


@pytest.mark.parametrize("phi", list(ANGULAR))
def test_angular_variants_trimer_closed_form(tmp_path, phi):
    angular = [(2.0, 1.0, 1), (3.5, 4.0, -1), (1.5, 2.0, 1), (5.0, 1.0, -1)]
    m = ob.OracleModel(write(one_type(phi_a=phi, angular=angular), tmp_path))
    r1, r2, th = 2.1, 3.4, 1.9
    pos = np.array([[0, 0, 0], [r1, 0, 0], [r2 * math.cos(th), r2 * math.sin(th), 0.0]])
    G = m.symm(np.zeros(3, dtype=np.int32), pos, None)
    nR = 1
    cs = math.cos(th)
    for k, (eta, zeta, lam) in enumerate(angular):
        ietap = (1.0 / eta) ** 2
        want = abs(0.5 * (1 + lam * cs)) ** zeta * fc_cos(r1, RC) * fc_cos(r2, RC) * \
            ANGULAR[phi](ietap * r1 * r1, ietap * r2 * r2)
        assert abs(G[0, nR + k] - want) <= 1e-13 * max(1.0, abs(want)), (phi, k)


def _cell_structure(n=150, seed=5):
    cell = synth.cubic_cell(n, 0.05)
    assert cell[0, 0] > 2 * RC
    pos = synth.random_positions(n, cell, 1.6, seed)
    return synth.Structure(np.zeros(n, dtype=np.int32), pos, cell)


@pytest.mark.parametrize("phi_r,phi_a", [(p, "GAUSS") for p in RADIAL if p != "GAUSSIAN"] +
                         [("GAUSSIAN", p) for p in ANGULAR if p != "GAUSS"])
def test_variant_forces_are_minus_gradient(tmp_path, phi_r, phi_a):
    radial = [(1.5, 0.6), (3.0, 1.1), (4.5, 0.4)]
    angular = [(2.0, 1.0, 1), (3.5, 4.0, -1), (2.5, 2.0, 1)]
    pot = one_type(phi_r=phi_r, phi_a=phi_a, radial=radial, angular=angular, hidden=(5,), seed=2)
    m = ob.OracleModel(write(pot, tmp_path))
    s = _cell_structure()
    E, Ea, F = m.compute(s.types, s.pos, s.cell, ob.F_FORCE)
    atoms = [0, 77, 149]
    fd = fd_forces(m, s, atoms, 0)
    # SOFTPLUS / LOGCOSH: the reference differentiates the exact function while its values use
    # the truncated log(1 + x) series, so F and -dE/dr differ at the series' 1e-6 level
    rtol = 3e-5 if phi_r in ("SOFTPLUS", "LOGCOSH") else 1e-6
    np.testing.assert_allclose(F[atoms], fd, rtol=rtol, atol=rtol / 10 * np.abs(F).max())


@pytest.mark.parametrize("neuron", list(NEURONS))
def test_neuron_closed_forms_and_input_gradient(tmp_path, neuron):
    pot = one_type(neuron=neuron, hidden=(1,))
    pot.weights[0] = [np.array([[1.0, 0.0]]), np.array([[1.0]])]
    pot.biases[0] = [np.zeros(1), np.zeros(1)]
    pot.outw[0], pot.outb[0] = 1.75, -0.25
    m = ob.OracleModel(write(pot, tmp_path))
    f, tol = NEURONS[neuron]
    for x in (-2.5, -0.3, 0.4, 3.1):
        out, g = m.network(0, np.array([x, 0.7]))
        assert abs(out - (-0.25 + 1.75 * f(x))) <= tol * 2 * max(1.0, abs(f(x))), (neuron, x, out)
        h = 1e-6
        fd = 1.75 * (f(x + h) - f(x - h)) / (2 * h)
        assert abs(g[0] - fd) <= 1e-6 * max(1.0, abs(fd)) + (2 * tol / h if tol > 1e-10 else 0.0), (neuron, x)
        assert g[1] == 0.0


@pytest.mark.parametrize("neuron", ["SIGMOID", "SOFTPLUS", "MISH", "SINSIG", "IFABS"])
def test_neuron_deep_network_forces(tmp_path, neuron):
    """Forces through a two-hidden-layer network of each family = -dE/dr."""
    radial = [(1.5, 0.6), (3.0, 1.1)]
    angular = [(2.0, 1.0, 1), (3.5, 4.0, -1)]
    pot = one_type(radial=radial, angular=angular, hidden=(6, 4), neuron=neuron, seed=9)
    pot.outw[0], pot.outb[0] = 0.8, 0.3
    m = ob.OracleModel(write(pot, tmp_path))
    s = _cell_structure(seed=6)
    E, Ea, F = m.compute(s.types, s.pos, s.cell, ob.F_FORCE)
    atoms = [3, 99]
    fd = fd_forces(m, s, atoms, 0)
    rtol = 3e-5 if neuron == "SOFTPLUS" else 1e-6  # see test_variant_forces_are_minus_gradient
    np.testing.assert_allclose(F[atoms], fd, rtol=rtol, atol=rtol / 10 * np.abs(F).max())


def test_variant_names_rejected_like_the_reference(tmp_path):
    for kw, msg in ((dict(phi_r="GAUSS"), "invalid radial function type"),
                    (dict(phi_a="GAUSSIAN"), "invalid angular function type"),
                    (dict(neuron="RELU6"), "Invalid neuron")):
        pot = one_type(**kw)
        with pytest.raises(ValueError, match=msg):
            ob.OracleModel(write(pot, tmp_path, "bad.ann"))
