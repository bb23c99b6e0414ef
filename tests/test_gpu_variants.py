"""The engine's basis and activation variants against the oracle (FP64, <= 1e-10 relative):
integer zeta beyond the fast kernels' polynomial range (8, 16) and zeta = 0,
radial TANH / SOFTPLUS / LOGCOSH / SWISH / MISH, angular IPOWP / SECHP (separable) and
IPOWS / SECHS (per-pair factor), and the 17 activations, through the separate kernels and the
fused kernel.  One mixed-precision case checks the FP32 forms against the FP64 oracle.
"""
import numpy as np
import pytest

from atomnn_amd import synth
from atomnn_amd.nnp import NNP, NNP_ATOM_ENERGY, NNP_FORCE, NNP_PREC_MIXED, NNP_VACUUM_ENERGY
import oracle_binding as ob

pytestmark = pytest.mark.gpu
TOL = 1e-10

NEURONS = ["LINEAR", "SIGMOID", "TANH", "ISRU", "ARCTAN", "SOFTSIGN", "RELU", "SOFTPLUS", "ELU", "GELU", "SWISH",
           "MISH", "TANHRE", "LOGCOSH", "IQUAD", "IFABS", "SINSIG"]


def rel_e(E, Eo, Eao):
    return abs(E - Eo) / max(abs(Eo), np.abs(Eao).sum())


def rel_f(F, Fo):
    return np.linalg.norm(F - Fo, axis=1).max() / np.linalg.norm(Fo, axis=1).max()


def variant_potential(tmp_path, phi_r="GAUSSIAN", phi_a="GAUSS", neuron="TANH", seed=0, zetas=(1.0, 4.0), cut="COS",
                      norm="UNIT", hidden=(16, 16)):
    """Writes the potential (preconditioned on the oracle's descriptors of a calibration cell)
    and returns its path."""
    rng = np.random.default_rng(seed)
    pot = synth.make_potential(["A", "B"], 6.0, 6, [0.005, 0.05], list(zetas), [1, -1], list(hidden), rng)
    # radial functions centred across the cutoff so each variant sees both signs of its argument
    for key in pot.radial:
        pot.radial[key] = [(rs, eta) for rs, eta in zip(np.linspace(0.5, 5.5, 6), np.linspace(0.3, 2.0, 6))]
    pot.phi_r, pot.phi_a, pot.neuron = phi_r, phi_a, neuron
    pot.cut, pot.norm = cut, norm
    pot.outw[0], pot.outb[1] = 0.9, 0.05
    path = str(tmp_path / (f"v_{phi_r}_{phi_a}_{neuron}_{cut}_{norm}_{'_'.join(str(z) for z in zetas)}"
                           f"_h{'x'.join(str(h) for h in hidden)}.ann"))
    pot.write(path)
    cell = synth.cubic_cell(128, 0.05)
    s = synth.Structure(synth.species_exact(128, [0.5, 0.5], rng), synth.random_positions(128, cell, 2.0, seed + 1),
                        cell)
    G = ob.OracleModel(path).symm(s.types, s.pos, s.cell)
    for t in range(2):
        rows = G[s.types == t][:, :pot.ninput(t)]
        pot.inpb[t] = -rows.mean(axis=0)
        pot.inpw[t] = 1.0 / np.maximum(rows.std(axis=0), 1e-3)
    pot.write(path)
    return path


def structures():
    rng = np.random.default_rng(41)
    cell = synth.cubic_cell(300, 0.05)
    big = synth.Structure(synth.species_exact(300, [0.5, 0.5], rng), synth.random_positions(300, cell, 2.0, 42), cell)
    return synth.batch512(3) + [big]


# cutoffs and normalisations of src/nnp/basis.cpp:80-120, 160-168 (the common fast kernels and
# the moment expansion take them like COS / UNIT)
CUTS = [dict(cut=c, norm=n) for c in ("STEP", "TANH") for n in ("UNIT", "VOL")] + [dict(norm="VOL")]

CASES = ([dict(zetas=(2.0, 8.0)), dict(zetas=(3.0, 16.0)), dict(zetas=(0.0, 2.0))] +
         [dict(phi_r=p) for p in ("TANH", "SOFTPLUS", "LOGCOSH", "SWISH", "MISH")] +
         [dict(phi_a=p) for p in ("IPOWP", "SECHP", "IPOWS", "SECHS")] +
         [dict(neuron=n) for n in NEURONS if n != "TANH"] + CUTS +
         [dict(phi_a="IPOWS", cut="TANH", norm="VOL"), dict(phi_r="SWISH", cut="STEP", norm="VOL")])


def case_id(c):
    return "-".join(f"{k}={v}" for k, v in c.items()).replace(" ", "")


@pytest.mark.parametrize("case", CASES, ids=[case_id(c) for c in CASES])
@pytest.mark.parametrize("pipeline", [0, 1])
def test_variant_parity_fp64(tmp_path, case, pipeline):
    path = variant_potential(tmp_path, **case)
    o = ob.OracleModel(path)
    m = NNP.read(path)
    try:
        m.set_option("pipeline", pipeline)
        strucs = structures()
        res = m.compute_batch(strucs)
        for s, (E, Ea, F) in zip(strucs, res):
            Eo, Eao, Fo = o.compute(s.types, s.pos, s.cell)
            assert rel_e(E, Eo, Eao) <= TOL, (case, s.natoms)
            assert np.abs(Ea - Eao).max() <= TOL * np.abs(Eao).sum()
            assert rel_f(F, Fo) <= TOL, (case, s.natoms)
    finally:
        m.close()


DEPTHS = [(), (16,), (24, 12, 40)]


@pytest.mark.parametrize("hidden", DEPTHS, ids=["h" + "x".join(map(str, h)) for h in DEPTHS])
def test_network_depths(tmp_path, hidden):
    """Networks with no hidden layer, one, and three of widths that are not multiples of 16: the
    output layer's energy dot product and elementwise first backward step, FP64 at 1e-10 (weights
    staged in LDS for these few tiles, read from L2 on a batch of more tiles than two per CU) and
    the mixed form within 1 meV/atom."""
    path = variant_potential(tmp_path, hidden=hidden, neuron="SWISH" if len(hidden) == 1 else "TANH")
    o = ob.OracleModel(path)
    m = NNP.read(path)
    try:
        strucs = structures()
        if len(hidden) == 3:
            strucs = strucs + synth.batch512(48)[8:]  # ~13k atoms: over 512 tiles per species pass
        for s, (E, Ea, F) in zip(strucs, m.compute_batch(strucs)):
            Eo, Eao, Fo = o.compute(s.types, s.pos, s.cell)
            assert rel_e(E, Eo, Eao) <= TOL, (hidden, s.natoms)
            assert np.abs(Ea - Eao).max() <= TOL * np.abs(Eao).sum()
            assert rel_f(F, Fo) <= TOL, (hidden, s.natoms)
        s = structures()[-1]
        E, Ea, F = m.compute(s, NNP_FORCE | NNP_ATOM_ENERGY | NNP_VACUUM_ENERGY | NNP_PREC_MIXED)
        Eo, Eao, Fo = o.compute(s.types, s.pos, s.cell)
        assert abs(E - Eo) / s.natoms <= 1e-3, hidden
        assert np.sqrt(np.mean(np.sum((F - Fo) ** 2, axis=1))) <= 20e-3, hidden
    finally:
        m.close()


def test_variant_descriptors_fp64(tmp_path):
    """Descriptor rows of every angular kind, including the per-pair factors, at 1e-10."""
    s = structures()[-1]
    for phi_a in ("IPOWP", "SECHP", "IPOWS", "SECHS"):
        path = variant_potential(tmp_path, phi_r="SWISH", phi_a=phi_a)
        m = NNP.read(path)
        try:
            G = m.symm(s)
            Go = ob.OracleModel(path).symm(s.types, s.pos, s.cell)
            assert np.abs(G - Go).max() / np.abs(Go).max() <= TOL, phi_a
        finally:
            m.close()


def test_variant_mixed_precision(tmp_path):
    path = variant_potential(tmp_path, phi_r="MISH", phi_a="SECHS", neuron="SWISH")
    o = ob.OracleModel(path)
    m = NNP.read(path)
    try:
        s = structures()[-1]
        E, Ea, F = m.compute(s, NNP_FORCE | NNP_VACUUM_ENERGY | NNP_PREC_MIXED)
        Eo, Eao, Fo = o.compute(s.types, s.pos, s.cell)
        assert abs(E - Eo) / s.natoms <= 1e-3
        assert np.sqrt(np.mean(np.sum((F - Fo) ** 2, axis=1))) <= 20e-3
    finally:
        m.close()


@pytest.mark.parametrize("case", CUTS, ids=[case_id(c) for c in CUTS])
def test_cutoffs_and_norms_triplet_path_and_resident(tmp_path, case):
    """The same cutoff / norm cases with the G4 sums and force terms by the triplet loops
    (moments 0), and through the resident graph-replayed pass with the moment expansion."""
    path = variant_potential(tmp_path, **case)
    o = ob.OracleModel(path)
    m = NNP.read(path)
    try:
        strucs = structures()
        ref = [o.compute(s.types, s.pos, s.cell) for s in strucs]
        m.set_option("moments", 0)
        for s, (Eo, Eao, Fo) in zip(strucs, ref):
            E, Ea, F = m.compute(s)
            assert rel_e(E, Eo, Eao) <= TOL and rel_f(F, Fo) <= TOL, (case, s.natoms)
        m.set_option("moments", 3)
        r = m.resident(strucs, NNP_FORCE | NNP_ATOM_ENERGY | NNP_VACUUM_ENERGY)
        for _ in range(3):
            r.run()
        for (E, Ea, F), (Eo, Eao, Fo) in zip(r.fetch(), ref):
            assert rel_e(E, Eo, Eao) <= TOL and rel_f(F, Fo) <= TOL, case
    finally:
        m.close()


@pytest.mark.parametrize("case", CUTS, ids=[case_id(c) for c in CUTS])
def test_cutoffs_and_norms_mixed_precision(tmp_path, case):
    path = variant_potential(tmp_path, **case)
    o = ob.OracleModel(path)
    m = NNP.read(path)
    try:
        for s in structures():
            E, Ea, F = m.compute(s, NNP_FORCE | NNP_VACUUM_ENERGY | NNP_PREC_MIXED)
            Eo, Eao, Fo = o.compute(s.types, s.pos, s.cell)
            assert abs(E - Eo) / s.natoms <= 1e-3, case
            assert np.sqrt(np.mean(np.sum((F - Fo) ** 2, axis=1))) <= 20e-3, case
    finally:
        m.close()
