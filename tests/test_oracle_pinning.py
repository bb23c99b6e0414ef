"""Pins the CPU oracle against what the mathematics and the reference's own tests fix.

The reference ships no golden vectors (its tests print FD errors and timings only), so
each oracle function is checked against something independent of its code:
  * closed forms: dimer G2, symmetric-trimer G4, integral of the cosine cutoff, one-hot networks;
  * finite differences of every derivative path (forces = -dE/dr where the reference's
    min-image reaction rule is inert, i.e. L > 2 rc, or in the all-images mode);
  * invariants the reference tests use: supercell consistency and NaCl symmetry
    equivalence (src/nnp/test_nnp_symm.cpp), neighbour no-duplicates and supercell match
    (src/struc/test_neigh_unit.cpp), FD of dO/dI (src/ml/test_nn_grad.cpp);
  * an independent numpy enumeration of neighbours and descriptors (atomnn_amd.synth).
Each check is chosen so that a dropped term, a wrong sign or index, or a transposed
operand fails it.
"""
import math
import os

import numpy as np
import pytest

from atomnn_amd import synth
import oracle_binding as ob


# ------------------------------------------------------------------ helpers
def tiny_potential(species, rc=6.0, radial=None, angular=None, hidden=(4,), seed=0, norm="UNIT", cut="COS",
                   evac=None):
    """A potential with explicit basis lists; radial: [(rs, eta)], angular: [(eta_len, zeta, lambda)]."""
    T = len(species)
    radial = radial if radial is not None else [(0.0, 0.5)]
    angular = angular if angular is not None else [(2.0, 1.0, 1)]
    rng = np.random.default_rng(seed)
    pot = synth.Potential(
        species=list(species), rc=rc,
        radial={(i, j): list(radial) for i in range(T) for j in range(T)},
        angular={(i, j, k): (2, list(angular)) for i in range(T) for j in range(T) for k in range(j, T)},
        layers={}, weights={}, biases={}, inpw={}, inpb={}, norm=norm, cut=cut)
    for s in range(T):
        nin = pot.ninput(s)
        lay = [nin] + list(hidden) + [1]
        pot.layers[s] = lay
        pot.weights[s] = [rng.normal(0, 0.5, size=(lay[l + 1], lay[l])) for l in range(len(lay) - 1)]
        pot.biases[s] = [rng.normal(0, 0.1, size=lay[l + 1]) for l in range(len(lay) - 1)]
        pot.inpw[s] = np.ones(nin)
        pot.inpb[s] = np.zeros(nin)
    if evac:
        pot.evac = dict(evac)
    return pot


def write(pot, tmp_path, name="p.ann"):
    p = str(tmp_path / name)
    pot.write(p)
    return p


def fc_cos(r, rc):
    return 0.5 * (math.cos(math.pi * r / rc) + 1.0) if r <= rc else 0.0


def fc_closed(cut, r, rc):
    """The reference's cutoffs (src/nnp/basis.cpp:80-99): COS, STEP (1 inside, 0 outside),
    TANH = tanh^3(1 - r/rc)."""
    if r > rc:
        return 0.0
    if cut == "COS":
        return fc_cos(r, rc)
    if cut == "STEP":
        return 1.0
    return math.tanh(1.0 - r / rc) ** 3


def norm_closed(norm, rc):
    """UNIT = 1; VOL = 1 / (1/2 4/3 (pi^2 - 6)/pi rc^3) (src/nnp/basis.cpp:160-168)."""
    return 1.0 if norm == "UNIT" else 1.0 / (0.5 * 4.0 / 3.0 * (math.pi ** 2 - 6.0) / math.pi * rc ** 3)


# ------------------------------------------------------------------ cell / neighbours
def test_cell_diff_cubic_closed_form():
    cell = np.eye(3) * 10.0
    np.testing.assert_allclose(ob.cell_diff(cell, [1, 1, 1], [9, 9, 9]), [2, 2, 2], atol=1e-12)
    np.testing.assert_allclose(ob.cell_diff(cell, [9, 1, 5], [1, 9, 5.5]), [-2, 2, -0.5], atol=1e-12)


def test_cell_diff_minimum_image_and_translation_invariance():
    rng = np.random.default_rng(1)
    cell = synth.strained_cell(64, 0.05, 0.10, rng)
    for _ in range(200):
        v1, v2 = rng.uniform(-20, 20, 3), rng.uniform(-20, 20, 3)
        d = ob.cell_diff(cell, v1, v2)
        imgs = np.array([(a, b, c) for a in range(-4, 5) for b in range(-4, 5) for c in range(-4, 5)]) @ cell
        cand = (v1 - v2)[None, :] - imgs
        best = cand[np.argmin(np.einsum("ij,ij->i", cand, cand))]
        np.testing.assert_allclose(d, best, atol=1e-9)
        n = rng.integers(-3, 4, 3)
        np.testing.assert_allclose(ob.cell_diff(cell, v1 + n @ cell, v2), d, atol=1e-9)


def _nbr_sets(nl, n):
    out = []
    for i in range(n):
        a, b = nl["offsets"][i], nl["offsets"][i + 1]
        out.append(sorted((int(j), tuple(np.round(r, 7))) for j, r in zip(nl["index"][a:b], nl["r"][a:b])))
    return out


@pytest.mark.parametrize("which", ["plain64", "tri0", "tri1", "tri2", "open"])
def test_neighbor_sets_match_independent_enumeration(which):
    if which == "plain64":
        s = synth.plain64()[0]
    elif which == "open":
        cell = synth.cubic_cell(80, 0.05)
        s = synth.Structure(np.zeros(80, dtype=np.int32), synth.random_positions(80, cell, 2.0, 4), None)
    else:
        s = synth.batch512(int(which[-1]) + 1)[-1]
        s = synth.Structure(s.types, s.pos, s.cell)
    nl = ob.neighbors(s.types, s.pos, s.cell, 6.0)
    ref = synth.numpy_neighbors(s, 6.0)
    got = _nbr_sets(nl, s.natoms)
    for i in range(s.natoms):
        r, jj = ref[i]
        exp = sorted((int(j), tuple(np.round(v, 7))) for j, v in zip(jj, r))
        assert got[i] == exp, f"atom {i}"


def test_neighbor_order_min_flags_and_no_duplicates():
    for s in [synth.plain64()[0]] + synth.batch512(4):
        nl = ob.neighbors(s.types, s.pos, s.cell, 6.0)
        nimg = int(np.prod(2 * nl["shell"] + 1))
        for i in range(s.natoms):
            a, b = nl["offsets"][i], nl["offsets"][i + 1]
            key = nl["index"][a:b].astype(np.int64) * nimg + nl["image"][a:b]
            assert np.all(np.diff(key) > 0)  # (j, image) strictly ascending: the reference order
            r = nl["r"][a:b]
            assert len({tuple(np.round(v, 9)) for v in r}) == len(r)  # no duplicate vectors
            for j in np.unique(nl["index"][a:b]):
                sel = nl["index"][a:b] == j
                mins = nl["min"][a:b][sel]
                if j == i:
                    assert mins.sum() == 0  # self images never carry the min flag
                else:
                    assert mins.sum() == 1
                    d = nl["d"][a:b][sel]
                    assert d[mins == 1][0] == d.min()


def test_neighbor_supercell_consistency():
    s = synth.plain64()[0]
    nl = ob.neighbors(s.types, s.pos, s.cell, 6.0)
    sup = 2
    shifts = np.array([(a, b, c) for a in range(sup) for b in range(sup) for c in range(sup)]) @ s.cell
    pos2 = np.concatenate([s.pos + t for t in shifts])
    nl2 = ob.neighbors(np.zeros(len(pos2), dtype=np.int32), pos2, s.cell * sup, 6.0)
    for i in range(s.natoms):
        d1 = np.sort(nl["d"][nl["offsets"][i]:nl["offsets"][i + 1]])
        for k in range(sup ** 3):
            ii = k * s.natoms + i
            d2 = np.sort(nl2["d"][nl2["offsets"][ii]:nl2["offsets"][ii + 1]])
            np.testing.assert_allclose(d2, d1, atol=1e-9)


# ------------------------------------------------------------------ basis closed forms
def test_g2_dimer_closed_form_and_cutoff_integral(tmp_path):
    rc = 6.0
    etas = [0.0, 0.3, 1.7]
    pot = tiny_potential(["A"], rc=rc, radial=[(0.0, e) for e in etas] + [(1.5, 0.8)])
    m = ob.OracleModel(write(pot, tmp_path))
    rs_list = np.linspace(0.5, 6.5, 61)
    vals = []
    for r in rs_list:
        pos = np.array([[0.0, 0, 0], [r, 0, 0]])
        G = m.symm([0, 0], pos, None)
        exp_vals = [math.exp(-e * r * r) * fc_cos(r, rc) for e in etas] + [math.exp(-0.8 * (r - 1.5) ** 2) * fc_cos(r, rc)]
        np.testing.assert_allclose(G[0, :4], exp_vals, rtol=1e-13, atol=1e-15)
        vals.append(G[0, 0])
    # integral of the cosine cutoff over [0, rc] is rc/2 (the reference's test_cutoff.cpp check)
    x = np.linspace(0, rc, 20001)
    fcs = []
    for r in x[::400]:
        G = m.symm([0, 0], np.array([[0.0, 0, 0], [max(r, 1e-2), 0, 0]]), None)
        fcs.append(G[0, 0])
    integral = np.trapezoid(fcs, x[::400])
    assert abs(integral - rc / 2) < 2e-3


@pytest.mark.parametrize("norm", ["UNIT", "VOL"])
@pytest.mark.parametrize("cut", ["STEP", "TANH", "COS"])
def test_g2_g4_closed_forms_every_cutoff_and_norm(tmp_path, cut, norm):
    """Dimer G2 = norm exp(-eta (r - rs)^2) fc(r) and symmetric-trimer G4 = norm |(1 + lambda
    cos)/2|^zeta fc(r)^2 exp(-2 r^2 / eta^2), for each cutoff and normalisation, on both sides of
    rc (a STEP cutoff that leaked past rc, a TANH cutoff missing its cube, or a dropped VOL factor
    fails here)."""
    rc = 6.0
    rad = [(0.0, 0.05), (0.0, 0.7), (2.0, 0.4)]
    ang = [(3.0, 1.0, 1), (2.5, 4.0, -1)]
    pot = tiny_potential(["A"], rc=rc, radial=rad, angular=ang, norm=norm, cut=cut)
    m = ob.OracleModel(write(pot, tmp_path))
    nv = norm_closed(norm, rc)
    # TANH near rc: 1 - r/rc cancels, so its relative rounding grows by the condition number
    # r / (rc - r) (the reference forms it as 1 - r * (1/rc))
    cond = lambda r: 1.0 if cut != "TANH" or r > rc else max(1.0, 3.0 * r / (rc - r))
    for r in [0.7, 1.9, 3.3, 4.8, 5.9, 5.999, 6.001, 6.6]:
        G = m.symm([0, 0], np.array([[0.0, 0, 0], [r, 0, 0]]), None)
        want = [nv * math.exp(-eta * (r - rs) ** 2) * fc_closed(cut, r, rc) for rs, eta in rad]
        np.testing.assert_allclose(G[0, :3], want, rtol=1e-13 * cond(r), atol=1e-300)
    for r, theta in [(1.5, 1.0), (2.6, 2.2), (3.9, 0.5), (5.95, 1.7), (6.2, 1.0)]:
        pos = np.array([[0.0, 0, 0], [r, 0, 0], [r * math.cos(theta), r * math.sin(theta), 0]])
        G = m.symm([0, 0, 0], pos, None)
        for f, (eta, zeta, lam) in enumerate(ang):
            want = (nv * abs(0.5 * (1 + lam * math.cos(theta))) ** zeta * fc_closed(cut, r, rc) ** 2
                    * math.exp(-2 * r * r / eta ** 2))
            assert abs(G[0, 3 + f] - want) <= 1e-13 * cond(r) * max(abs(want), 1e-300), (cut, norm, r, theta, f)


@pytest.mark.parametrize("cut", ["STEP", "TANH"])
def test_forces_are_minus_gradient_step_and_tanh_cutoffs(tmp_path, cut):
    """F = -dE/dr on an L > 2 rc cell with the STEP and TANH cutoffs (VOL norm), so cut_grad is
    pinned: STEP has derivative 0 inside rc (no pair lies within the difference step of rc), TANH
    -3 tanh^2 (1 - tanh^2) / rc."""
    pot = tiny_potential(["A", "B"], radial=[(0.0, 0.3), (0.0, 1.1), (1.0, 0.5)],
                         angular=[(2.5, 1.0, 1), (3.0, 4.0, -1)], hidden=(6,), seed=4, norm="VOL", cut=cut)
    m = ob.OracleModel(write(pot, tmp_path))
    n = 120
    cell = synth.cubic_cell(n, 0.05) * (14.0 / synth.cubic_cell(n, 0.05)[0, 0])
    rng = np.random.default_rng(12)
    s = synth.Structure(synth.species_exact(n, [0.5, 0.5], rng), synth.random_positions(n, cell, 2.0, 13), cell)
    nl = ob.neighbors(s.types, s.pos, s.cell, 6.0)
    assert np.abs(nl["d"] - 6.0).min() > 1e-4  # no pair crosses rc within the difference step
    E, Ea, F = m.compute(s.types, s.pos, s.cell, ob.F_FORCE)
    atoms = [0, 41, 119]
    fd = fd_forces(m, s, atoms, 0)
    np.testing.assert_allclose(F[atoms], fd, rtol=1e-6, atol=1e-7 * np.abs(F).max())


def test_g4_symmetric_trimer_closed_form(tmp_path):
    rc = 6.0
    ang = [(2.0, 1.0, 1), (2.0, 1.0, -1), (3.5, 4.0, 1), (3.5, 4.0, -1), (1.7, 2.5, 1)]
    pot = tiny_potential(["A"], rc=rc, radial=[(0.0, 0.1)], angular=ang)
    m = ob.OracleModel(write(pot, tmp_path))
    for r, theta in [(1.5, 1.0), (2.5, 2.0), (3.2, 0.4), (4.0, 2.9)]:
        pos = np.array([[0.0, 0, 0], [r, 0, 0], [r * math.cos(theta), r * math.sin(theta), 0]])
        G = m.symm([0, 0, 0], pos, None)
        for f, (eta, zeta, lam) in enumerate(ang):
            ietap = (1.0 / eta) ** 2
            exp_v = abs(0.5 * (1 + lam * math.cos(theta))) ** zeta * fc_cos(r, rc) ** 2 * math.exp(-ietap * 2 * r * r)
            assert abs(G[0, 1 + f] - exp_v) <= 1e-13 * max(1.0, abs(exp_v)), (r, theta, f)


def test_vol_norm_scales_every_function(tmp_path):
    s = synth.plain64()[0]
    p_unit = tiny_potential(["A"], radial=[(0.0, 0.2), (0.0, 1.0)], angular=[(3.0, 1.0, 1), (2.0, 4.0, -1)])
    p_vol = tiny_potential(["A"], radial=[(0.0, 0.2), (0.0, 1.0)], angular=[(3.0, 1.0, 1), (2.0, 4.0, -1)], norm="VOL")
    Gu = ob.OracleModel(write(p_unit, tmp_path, "u.ann")).symm(s.types, s.pos, s.cell)
    Gv = ob.OracleModel(write(p_vol, tmp_path, "v.ann")).symm(s.types, s.pos, s.cell)
    rc = 6.0
    norm = 1.0 / (0.5 * 4.0 / 3.0 * (math.pi ** 2 - 6.0) / math.pi * rc ** 3)
    np.testing.assert_allclose(Gv, Gu * norm, rtol=1e-13)


def test_descriptor_layout_three_species(tmp_path):
    """Radial blocks by neighbour species, angular blocks in LMat order (0,0),(1,0),(1,1),(2,0),..."""
    pot = tiny_potential(["A", "B", "C"], radial=[(0.0, 0.3)], angular=[(2.5, 1.0, 1)])
    m = ob.OracleModel(write(pot, tmp_path))
    # centre A at origin; one B and one C neighbour: only the radial B and C slots and the (B,C) angular block
    pos = np.array([[0.0, 0, 0], [2.0, 0, 0], [0, 3.0, 0]])
    G = m.symm([0, 1, 2], pos, None)
    g = G[0]
    assert g[0] == 0.0 and g[1] > 0 and g[2] > 0  # radial A, B, C
    ang = g[3:]
    # LMat linear order: (0,0)=AA, (1,0)=AB, (1,1)=BB, (2,0)=AC, (2,1)=BC, (2,2)=CC
    assert np.count_nonzero(ang) == 1 and ang[4] > 0
    np.testing.assert_allclose(g[1], math.exp(-0.3 * 4) * fc_cos(2.0, 6.0), rtol=1e-14)
    np.testing.assert_allclose(g[2], math.exp(-0.3 * 9) * fc_cos(3.0, 6.0), rtol=1e-14)
    # against the independent numpy evaluator on a random 3-species cell
    cell = synth.cubic_cell(60, 0.06)
    s = synth.Structure(np.array([i % 3 for i in range(60)], dtype=np.int32), synth.random_positions(60, cell, 1.8, 2),
                        cell)
    Go = m.symm(s.types, s.pos, s.cell)
    Gn = np.array(synth.numpy_descriptors(pot.quantized(), s))
    np.testing.assert_allclose(Go, Gn, rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("name", ["plain64", "plain4096"])
def test_descriptors_match_numpy_evaluator(potentials, name):
    m = ob.OracleModel(potentials[name])
    pot = synth.workload_potential(name).quantized()
    strucs = synth.plain64() if name == "plain64" else synth.batch512(3)
    for s in strucs:
        Go = m.symm(s.types, s.pos, s.cell)
        Gn = np.array(synth.numpy_descriptors(pot, s))
        assert np.abs(Go - Gn).max() / np.abs(Gn).max() < 1e-13


def test_nacl_symmetry_equivalence(tmp_path):
    """Rock salt: every Na (Cl) descriptor is equal in the cubic and the primitive fcc cell
    (AtomNN src/nnp/test_nnp_symm.cpp:25-290, a0 = 5.6199998856)."""
    a0 = 5.6199998856
    pot = tiny_potential(["Na", "Cl"], radial=[(0.0, 0.1), (0.0, 0.5)], angular=[(3.0, 1.0, 1), (2.0, 2.0, -1)])
    m = ob.OracleModel(write(pot, tmp_path))
    fcc = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0]])
    frac = np.concatenate([fcc, fcc + [0.5, 0, 0]])
    cubic = np.eye(3) * a0
    types = np.array([0] * 4 + [1] * 4, dtype=np.int32)
    Gc = m.symm(types, (frac % 1.0) @ cubic, cubic)
    prim = np.array([[0, .5, .5], [.5, 0, .5], [.5, .5, 0]]) * a0
    Gp = m.symm([0, 1], np.array([[0, 0, 0], [0.5 * a0, 0, 0]]), prim)
    for i in range(8):
        np.testing.assert_allclose(Gc[i], Gp[0 if i < 4 else 1], rtol=1e-10, atol=1e-12)


# ------------------------------------------------------------------ network
def test_network_closed_form_and_fd(tmp_path):
    pot = tiny_potential(["A"], radial=[(0.0, 0.3)], angular=[(2.0, 1.0, 1)], hidden=(2,))
    W0 = np.array([[0.7, 0.0], [0.0, -1.3]])
    W1 = np.array([[0.9, 0.4]])
    pot.weights[0] = [W0, W1]
    pot.biases[0] = [np.array([0.1, -0.2]), np.array([0.05])]
    pot.inpw[0] = np.array([2.0, 0.5])
    pot.inpb[0] = np.array([-0.25, 0.75])
    m = ob.OracleModel(write(pot, tmp_path))
    x = np.array([0.8, -0.3])
    ins = pot.inpw[0] * (x + pot.inpb[0])
    h = np.tanh(W0 @ ins + pot.biases[0][0])
    E = float(W1 @ h + pot.biases[0][1])
    dE = (W1[0] * (1 - h ** 2)) @ W0 * pot.inpw[0]
    out, dEdG = m.network(0, x)
    assert abs(out - E) < 1e-14
    np.testing.assert_allclose(dEdG, dE, rtol=1e-13)
    # deeper random network: central finite differences of dO/dI (test_nn_grad.cpp:12-93)
    pot2 = tiny_potential(["A"], radial=[(0.0, 0.3), (0.0, 1.1)], angular=[(2.0, 1.0, 1), (3.0, 4.0, -1)],
                          hidden=(7, 5), seed=3)
    m2 = ob.OracleModel(write(pot2, tmp_path, "deep.ann"))
    x = np.random.default_rng(4).normal(size=4)
    _, g = m2.network(0, x)
    for k in range(4):
        e = np.zeros(4)
        e[k] = 1e-6
        fd = (m2.network(0, x + e)[0] - m2.network(0, x - e)[0]) / 2e-6
        assert abs(fd - g[k]) < 1e-8 * max(1, abs(g[k]))


# ------------------------------------------------------------------ energy / forces
def fd_forces(m, s, atoms, flags, h=1e-5):
    out = []
    for i in atoms:
        row = []
        for c in range(3):
            p = s.pos.copy()
            p[i, c] += h
            ep = m.compute(s.types, p, s.cell, flags)[0]
            p[i, c] -= 2 * h
            em = m.compute(s.types, p, s.cell, flags)[0]
            row.append(-(ep - em) / (2 * h))
        out.append(row)
    return np.array(out)


def test_forces_are_minus_gradient_large_cell(potentials):
    """L > 2 rc: the reference's min-image reaction rule is inert, so F = -dE/dr exactly."""
    m = ob.OracleModel(potentials["plain4096"])
    n = 300
    cell = synth.cubic_cell(n, 0.05)
    assert cell[0, 0] > 12.0
    rng = np.random.default_rng(8)
    s = synth.Structure(synth.species_exact(n, [0.5, 0.5], rng), synth.random_positions(n, cell, 2.0, 9), cell)
    E, Ea, F = m.compute(s.types, s.pos, s.cell, ob.F_FORCE)
    atoms = [0, 17, 123, 299]
    fd = fd_forces(m, s, atoms, 0)
    np.testing.assert_allclose(F[atoms], fd, rtol=1e-6, atol=1e-7 * np.abs(F).max())
    np.testing.assert_allclose(F.sum(axis=0), 0, atol=1e-10 * np.abs(F).max() * n)


def test_forces_all_images_small_cell_and_min_image_rule(potentials):
    """L < 2 rc (the 64-atom cell): all-images forces are -dE/dr; the reference rule differs
    from it exactly by the reactions of the non-minimum images."""
    m = ob.OracleModel(potentials["plain64"])
    s = synth.plain64()[0]
    E1, _, Fall = m.compute(s.types, s.pos, s.cell, ob.F_FORCE | ob.F_ALL_IMAGES)
    fd = fd_forces(m, s, [0, 31, 63], 0)
    np.testing.assert_allclose(Fall[[0, 31, 63]], fd, rtol=1e-6, atol=1e-7 * np.abs(Fall).max())
    np.testing.assert_allclose(Fall.sum(axis=0), 0, atol=1e-9)
    E2, _, Fmin = m.compute(s.types, s.pos, s.cell, ob.F_FORCE)
    assert E1 == E2
    assert np.abs(Fmin.sum(axis=0)).max() > 1e-6  # reactions dropped: momentum not conserved
    assert np.abs(Fmin - Fall).max() > 1e-6


def test_min_image_rule_hand_built_two_atom_cell(tmp_path):
    """Two atoms in a 5 A cube (rc = 6): radial-only linear potential E_i = sum_m w_m G_m(i).
    Reference rule: F_i = sum_entries X - sum_{min entries of j pointing at i} X."""
    pot = tiny_potential(["A"], radial=[(0.0, 0.2), (0.0, 0.9)], angular=[(2.0, 1.0, 1)], hidden=(1,))
    pot.neuron = "LINEAR"
    w = np.array([0.7, -0.4, 0.0])
    pot.weights[0] = [w[None, :], np.array([[1.0]])]
    pot.biases[0] = [np.zeros(1), np.zeros(1)]
    m = ob.OracleModel(write(pot, tmp_path))
    L, rc = 5.0, 6.0
    cell = np.eye(3) * L
    pos = np.array([[0.3, 0.2, 0.1], [2.9, 1.1, 4.2]])
    nl = ob.neighbors([0, 0], pos, cell, rc)

    def dGdd(d):  # derivative of each radial function
        f, g = fc_cos(d, rc), (-0.5 * math.pi / rc * math.sin(math.pi * d / rc) if d <= rc else 0.0)
        return [math.exp(-e * d * d) * (-2 * e * d * f + g) for e in (0.2, 0.9)]

    X = []
    for q in range(len(nl["d"])):
        d = nl["d"][q]
        dE = float(np.dot(w[:2], dGdd(d)))
        X.append(-dE * nl["r"][q] / d)
    X = np.array(X)
    F = np.zeros((2, 3))
    for i in range(2):
        a, b = nl["offsets"][i], nl["offsets"][i + 1]
        F[i] += X[a:b].sum(axis=0)
        for q in range(a, b):
            if nl["min"][q]:
                F[nl["index"][q]] -= X[q]
    _, _, Fo = m.compute([0, 0], pos, cell, ob.F_FORCE)
    np.testing.assert_allclose(Fo, F, rtol=1e-12, atol=1e-14)


def test_rotation_translation_invariance(potentials):
    m = ob.OracleModel(potentials["plain64"])
    s = synth.plain64()[0]
    E, Ea, F = m.compute(s.types, s.pos, s.cell, ob.F_FORCE)
    th = 0.7
    Rz = np.array([[math.cos(th), -math.sin(th), 0], [math.sin(th), math.cos(th), 0], [0, 0, 1]])
    Rx = np.array([[1, 0, 0], [0, math.cos(0.3), -math.sin(0.3)], [0, math.sin(0.3), math.cos(0.3)]])
    Q = Rz @ Rx
    E2, Ea2, F2 = m.compute(s.types, s.pos @ Q.T + [1.3, -2.2, 0.7], s.cell @ Q.T, ob.F_FORCE)
    assert abs(E2 - E) < 1e-11 * abs(E) + 1e-12
    np.testing.assert_allclose(F2, F @ Q.T, atol=1e-10)


def test_supercell_extensivity(potentials):
    m = ob.OracleModel(potentials["plain64"])
    s = synth.plain64()[0]
    E, Ea, F = m.compute(s.types, s.pos, s.cell, ob.F_FORCE | ob.F_ALL_IMAGES)
    shifts = np.array([(a, b, c) for a in range(2) for b in range(2) for c in range(2)]) @ s.cell
    pos2 = np.concatenate([s.pos + t for t in shifts])
    E2, Ea2, F2 = m.compute(np.zeros(len(pos2), dtype=np.int32), pos2, 2 * s.cell, ob.F_FORCE | ob.F_ALL_IMAGES)
    assert abs(E2 - 8 * E) < 1e-10 * abs(E2)
    np.testing.assert_allclose(F2, np.tile(F, (8, 1)), atol=1e-10)


def test_vacuum_energy_flag(tmp_path):
    pot = tiny_potential(["A", "B"], evac={0: -1.25, 1: 0.5})
    m = ob.OracleModel(write(pot, tmp_path))
    s = synth.batch512(1)[0]
    Ev, Eav, _ = m.compute(s.types, s.pos, s.cell, ob.F_VACUUM)
    E0, Ea0, _ = m.compute(s.types, s.pos, s.cell, 0)
    np.testing.assert_allclose(Eav - Ea0, np.where(s.types == 0, -1.25, 0.5), atol=1e-14)
    assert abs((Ev - E0) - float(np.where(s.types == 0, -1.25, 0.5).sum())) < 1e-10


# ------------------------------------------------------------------ parser
def test_parser_round_trip_and_errors(potentials, tmp_path):
    m = ob.OracleModel(potentials["mixed64x4096"])
    assert m.ntypes == 3 and m.rc == 6.5
    assert [m.ninput(t) for t in range(3)] == [84, 84, 84]
    assert m.type_index("B") == 1 and m.type_index("Zz") == -1
    bad = tmp_path / "bad.ann"
    text = open(potentials["plain64"]).read()
    bad.write_text(text.replace("ntypes 1", "ntypes 0"))
    with pytest.raises(ValueError, match="invalid number of types"):
        ob.OracleModel(str(bad))
    bad.write_text(text[: len(text) // 2])
    with pytest.raises(ValueError):
        ob.OracleModel(str(bad))
    bad.write_text(text.replace("rc 6.000000", "rc -1"))
    with pytest.raises(ValueError, match="invalid cutoff"):
        ob.OracleModel(str(bad))
    with pytest.raises(ValueError, match="Could not open"):
        ob.OracleModel(str(tmp_path / "missing.ann"))
