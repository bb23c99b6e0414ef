"""Edge cases of the evaluation path on the GPU against the oracle: atoms with no neighbours,
one-atom structures (open, and periodic with only self images), a pair at exactly the cutoff
(inclusive, neighbor.cpp:113), positions outside the cell, a ragged batch mixing all of these
with the 4096-atom cell, invalid input; and, at the 20,000-atom size where the oracle is not
run, the properties that do not depend on size (Newton's third law with one image per pair,
translation invariance, batch equal to single)."""
import numpy as np
import pytest

from atomnn_amd import synth
from atomnn_amd.nnp import NNP, NNPError, NNP_FORCE, NNP_VACUUM_ENERGY
import oracle_binding as ob
from test_gpu_parity import TOL, check_neighbors

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=[0, 3], ids=["triplets", "moments"])
def pair(potentials, request):
    m = NNP.read(potentials["plain4096"])
    m.set_option("moments", request.param)
    o = ob.OracleModel(potentials["plain4096"])
    yield m, o
    m.close()


def edge_structures():
    rng = np.random.default_rng(41)
    out = []
    # one atom, open: no neighbours at all
    out.append(synth.Structure(np.array([1], np.int32), np.array([[0.3, -0.2, 0.1]]), None))
    # one atom in a small triclinic cell: self images only (no min flags, no reactions)
    cell = np.array([[3.1, 0.0, 0.0], [0.4, 2.9, 0.0], [-0.3, 0.5, 3.3]])
    out.append(synth.Structure(np.array([0], np.int32), np.array([[0.5, 0.7, 1.1]]), cell))
    # a pair at exactly rc = 6 (kept, fc = 0) and a third atom within reach of one of them
    out.append(synth.Structure(np.array([0, 1, 1], np.int32),
                               np.array([[0.0, 0.0, 0.0], [6.0, 0.0, 0.0], [1.5, 1.0, -0.5]]), None))
    # a cluster with isolated atoms (far from everything) among bonded ones
    pos = synth.random_positions(12, synth.cubic_cell(12, 0.05), 2.0, 9)
    pos = np.vstack([pos, [[60.0, 0.0, 0.0], [0.0, -55.0, 3.0]]])
    out.append(synth.Structure(rng.integers(0, 2, 14).astype(np.int32), pos, None))
    # positions far outside a periodic cell (several lattice vectors away)
    s = synth.batch512(3)[2]
    shift = np.array([3, -2, 5]) @ s.cell
    out.append(synth.Structure(s.types, s.pos + shift, s.cell))
    return out


def check_against_oracle(got, o, s):
    E, Ea, F = got
    Eo, Eao, Fo = o.compute(s.types, s.pos, s.cell)
    assert abs(E - Eo) <= TOL * max(abs(Eo), np.abs(Eao).sum(), 1e-300)
    # forces that cancel to rounding level (one atom among its own images) get an absolute floor
    scale = max(np.linalg.norm(Fo, axis=1).max(), 1e-2)
    assert np.linalg.norm(F - Fo, axis=1).max() <= TOL * scale


def test_edge_structures_match_the_oracle(pair):
    m, o = pair
    for k, s in enumerate(edge_structures()):
        check_neighbors(m, s, o.rc)
        check_against_oracle(m.compute(s), o, s)


def test_isolated_atom_has_its_bias_energy_and_no_force(pair):
    """With no neighbours every descriptor is 0: E is the network at G = 0 plus the vacuum
    energy, the force is exactly zero."""
    m, o = pair
    s = edge_structures()[0]
    E, Ea, F = m.compute(s)
    Eo, Eao, Fo = o.compute(s.types, s.pos, s.cell)
    assert np.all(F == 0.0)
    assert abs(E - Eo) <= TOL * abs(Eo)
    assert m.neighbors(s)["counts"][0] == 0


def test_pair_at_the_cutoff_is_a_neighbour(pair):
    m, o = pair
    s = edge_structures()[2]
    nb = m.neighbors(s)
    assert list(nb["counts"]) == [2, 2, 2]  # 0-1 at exactly rc, 0-2 and 1-2 closer
    d = nb["d"][nb["offsets"][0]:nb["offsets"][1]]
    assert 6.0 in d


def test_ragged_batch(pair):
    """Every edge structure batched with the 4096-atom cell and a 64-atom cell: each result
    equals its single-structure result bit for bit and the oracle within the bar."""
    m, o = pair
    rng = np.random.default_rng(5)
    cell = synth.cubic_cell(64, 0.05)
    small = synth.Structure(rng.integers(0, 2, 64).astype(np.int32), synth.random_positions(64, cell, 2.0, 3), cell)
    edges = edge_structures()
    batch = [edges[0], synth.plain4096()[0], edges[1], edges[2], small, edges[3], edges[4]]
    got = m.compute_batch(batch)
    for s, g in zip(batch, got):
        E, Ea, F = m.compute(s)
        assert g[0] == E and np.array_equal(g[1], Ea) and np.array_equal(g[2], F)
    for s, g in zip(edges, [got[0], got[2], got[3], got[5], got[6]]):
        check_against_oracle(g, o, s)


def test_invalid_input(pair):
    m, o = pair
    with pytest.raises(NNPError):
        m.compute(synth.Structure(np.zeros(0, np.int32), np.zeros((0, 3)), None))
    with pytest.raises(NNPError):
        m.compute(synth.Structure(np.array([0, 5], np.int32), np.zeros((2, 3)) + [[0, 0, 0], [1, 1, 1]], None))
    with pytest.raises(NNPError):
        m.compute_batch([])


def test_size_independent_properties_at_20000_atoms(potentials):
    """A 20,000-atom cubic cell (L = 73.7 A > 2 rc: one image per pair, every entry carries the
    reaction): the forces sum to zero, a rigid translation changes nothing, and the batch of the
    cell with a copy of itself equals the single result."""
    m = NNP.read(potentials["plain4096"])
    try:
        rng = np.random.default_rng(77)
        cell = synth.cubic_cell(20000, 0.05)
        s = synth.Structure(synth.species_exact(20000, [0.5, 0.5], rng),
                            synth.random_positions(20000, cell, 1.9, 123), cell)
        E, Ea, F = m.compute(s)
        fmax = np.linalg.norm(F, axis=1).max()
        assert np.linalg.norm(F.sum(axis=0)) <= 1e-11 * fmax * np.sqrt(s.natoms)
        t = synth.Structure(s.types, s.pos + np.array([0.37, -1.21, 2.93]), cell)
        E2, Ea2, F2 = m.compute(t)
        assert abs(E2 - E) <= TOL * np.abs(Ea).sum()
        assert np.linalg.norm(F2 - F, axis=1).max() <= TOL * fmax
        (Eb, Eab, Fb), _ = m.compute_batch([s, t])
        assert Eb == E and np.array_equal(Fb, F)
    finally:
        m.close()
