"""The premises of the GPU reaction tests, checked on the oracle (CPU): the crystals carry exact
distance ties that make the reference's min-image rule differ from the all-images rule, and the
one-way pairs are one-way under the reference's own neighbour arithmetic."""
import numpy as np
import pytest

from atomnn_amd import synth
import oracle_binding as ob
from reaction_cases import ASYM, CRYSTALS, lists, pair_at_rc


@pytest.fixture(scope="module")
def oracle(tmp_path_factory):
    p = str(tmp_path_factory.mktemp("pot") / "plain4096.ann")
    synth.workload_potential("plain4096").write(p)
    return ob.OracleModel(p)


@pytest.mark.parametrize("name", ["bcc_1x1x1", "cscl_1x1x1", "fcc_1x1x1", "fcc_2x2x2"])
def test_min_image_rule_differs_from_all_images_on_crystals(oracle, name):
    """Perfect crystals: the all-images forces vanish by inversion symmetry, while the reference
    applies a reaction to one of several equidistant images only (the first in its table order),
    so its forces do not."""
    s = CRYSTALS[name]
    _, _, Fmin = oracle.compute(s.types, s.pos, s.cell, ob.F_FORCE)
    _, _, Fall = oracle.compute(s.types, s.pos, s.cell, ob.F_FORCE | ob.F_ALL_IMAGES)
    assert np.abs(Fmin).max() > 1e-3
    assert np.abs(Fall).max() < 1e-12 * max(1.0, np.abs(Fmin).max()) * 1e3


@pytest.mark.parametrize("name", sorted(CRYSTALS))
def test_crystals_have_exact_ties(name):
    """Some pair of distinct atoms in each crystal sits exactly half a cell edge apart along an
    axis (both fractional differences wrap to -1/2)."""
    s = CRYSTALS[name]
    L = s.cell[0, 0]
    frac = s.pos / L
    d = np.abs(frac[:, None, :] - frac[None, :, :])
    off = ~np.eye(s.natoms, dtype=bool)
    assert np.any((d == 0.5).any(axis=2) & off)


@pytest.mark.parametrize("name", sorted(ASYM))
def test_one_way_pairs(name):
    n, L, seed = ASYM[name]
    s = pair_at_rc(n, L, seed)
    nb = ob.neighbors(s.types, s.pos, s.cell, 6.0)
    assert lists(nb, 0, 1) != lists(nb, 1, 0)
    assert abs(np.linalg.norm(s.pos[1] - s.pos[0]) - 6.0) < 1e-12
    if n >= 256:  # the cell-list search applies (every image shell 0)
        assert s.cell[0, 0] >= 12.0
