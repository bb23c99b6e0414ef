"""Molecular dynamics on the GPU (velocity Verlet, the reference's Verlet::compute,
torch/integrator.cpp:170-188) against the same integrator driven by the CPU oracle's forces,
plus the physics a correct integrator must show: energy conservation in NVE and time
reversibility.

METAL units: a = F / m / 103.642696 A/fs^2, KE = 0.5 m v^2 * 103.642696 eV.
"""
import numpy as np
import pytest

from atomnn_amd import synth
from atomnn_amd.nnp import MD, NNP, NNP_ALL_IMAGES, NNP_FORCE
import oracle_binding as ob

pytestmark = pytest.mark.gpu

CONV = 103.642696
MASSES = [28.0855, 15.999]


def cell_300(seed=8):
    rng = np.random.default_rng(seed)
    cell = synth.cubic_cell(300, 0.05)  # L = 18.2 A > 2 rc: the min-image reaction rule is inert
    s = synth.Structure(synth.species_exact(300, [0.5, 0.5], rng), synth.random_positions(300, cell, 2.0, seed + 1),
                        cell)
    return s, rng


def velocities(s, rng, temp, masses=MASSES):
    m = np.array([masses[t] for t in s.types])[:, None]
    v = rng.normal(size=(s.natoms, 3)) * np.sqrt(8.617333262e-5 * temp / (m * CONV))
    v -= (m * v).sum(axis=0) / m.sum()
    return v


def wrap(x, cell):
    f = x @ np.linalg.inv(cell)
    f = np.fmod(np.fmod(f, 1.0) + 1.0, 1.0)
    return f @ cell


def test_md_matches_cpu_verlet_with_oracle_forces(potentials):
    s, rng = cell_300()
    v0 = velocities(s, rng, 600.0)
    dt, nsteps = 1.0, 4
    m = NNP.read(potentials["plain4096"])
    o = ob.OracleModel(potentials["plain4096"])
    md = MD(m, s, MASSES, v0, dt, NNP_FORCE)
    try:
        md.step(nsteps)
        st = md.state()
    finally:
        md.close()
        m.close()
    mass = np.array([MASSES[t] for t in s.types])[:, None]
    x, v = s.pos.copy(), v0.copy()
    E, _, F = o.compute(s.types, x, s.cell, ob.F_FORCE)
    for _ in range(nsteps):
        v += 0.5 * dt * F / (mass * CONV)
        x = wrap(x + v * dt, s.cell)
        E, _, F = o.compute(s.types, x, s.cell, ob.F_FORCE)
        v += 0.5 * dt * F / (mass * CONV)
    ke = 0.5 * float(np.sum(mass * v * v)) * CONV
    assert st["steps"] == nsteps
    assert np.abs(st["pos"] - x).max() < 1e-9
    assert np.abs(st["vel"] - v).max() < 1e-11
    assert abs(st["epot"] - E) <= 1e-9 * max(abs(E), 1.0)
    assert abs(st["ekin"] - ke) <= 1e-10 * ke
    assert abs(st["temperature"] - ke * 2 / 3 / (s.natoms * 8.617333262e-5)) < 1e-9


HEAVY = [120.0, 80.0]  # heavier masses keep the random starting structure's dynamics gentle


def test_md_energy_conservation_and_reversibility(potentials):
    s, rng = cell_300(seed=21)
    m = NNP.read(potentials["plain4096"])
    md = MD(m, s, HEAVY, None, 0.25, NNP_FORCE | NNP_ALL_IMAGES)
    try:
        st0 = md.state()
        e0 = st0["epot"] + st0["ekin"]
        etot = []
        for _ in range(10):
            md.step(20)
            st = md.state()
            etot.append(st["epot"] + st["ekin"])
        drift = np.abs(np.array(etot) - e0).max() / s.natoms
        print(f"NVE: max |dE_tot| = {drift * 1e3:.4f} meV/atom over 200 steps of 0.25 fs, "
              f"T = {st['temperature']:.1f} K")
        assert st["ekin"] > 0.0
        assert drift < 1e-3
        assert st["steps"] == 200
    finally:
        md.close()
    # time reversal: 20 steps forward, flip the velocities, 20 steps back
    v0 = velocities(s, rng, 300.0, HEAVY)
    md = MD(m, s, HEAVY, v0, 0.25, NNP_FORCE | NNP_ALL_IMAGES)
    try:
        md.step(20)
        mid = md.state()
    finally:
        md.close()
    back = synth.Structure(s.types, mid["pos"], s.cell)
    md = MD(m, back, HEAVY, -mid["vel"], 0.25, NNP_FORCE | NNP_ALL_IMAGES)
    try:
        md.step(20)
        end = md.state()
    finally:
        md.close()
        m.close()
    d = end["pos"] - s.pos
    d -= np.round(d @ np.linalg.inv(s.cell)) @ s.cell
    assert np.abs(d).max() < 1e-6
    assert np.abs(end["vel"] + v0).max() < 1e-8


def test_md_rejects_bad_arguments(potentials):
    s, _ = cell_300()
    m = NNP.read(potentials["plain4096"])
    try:
        with pytest.raises(Exception, match="masses must be positive"):
            MD(m, s, [28.0, 0.0], None, 1.0)
        with pytest.raises(Exception, match="invalid argument"):
            MD(m, s, MASSES, None, -1.0)
    finally:
        m.close()


def test_md_steps_in_one_call_equal_single_steps(potentials):
    """20 steps in one call (graphs of several steps) leave exactly the state of 20 one-step
    calls (one step per graph)."""
    m = NNP.read(potentials["plain4096"])
    try:
        s, rng = cell_300()
        v0 = velocities(s, rng, 300.0)
        a = MD(m, s, MASSES, v0, 0.5, NNP_FORCE)
        b = MD(m, s, MASSES, v0, 0.5, NNP_FORCE)
        a.step(20)
        for _ in range(20):
            b.step(1)
        sa, sb = a.state(), b.state()
        assert sa["steps"] == sb["steps"] == 20
        for k in ("pos", "vel", "force"):
            assert np.array_equal(sa[k], sb[k]), k
        assert sa["epot"] == sb["epot"] and sa["ekin"] == sb["ekin"]
        a.close()
        b.close()
    finally:
        m.close()


def test_md_with_receiver_rows_on_a_small_crystal(potentials):
    """A rattled 128-atom BCC cell (L = 11.5 A < 2 rc, so the min-image reaction rule is active)
    with the force assembly on receiver rows (option reaction_rows = 1): the trajectory matches
    the CPU Verlet driven by the oracle's forces."""
    from reaction_cases import BCC, crystal

    rng = np.random.default_rng(4)
    c = crystal(2.87, BCC, 4, [0, 1])
    s = synth.Structure(c.types, c.pos + rng.normal(0, 0.05, c.pos.shape), c.cell)
    v0 = velocities(s, rng, 300.0)
    dt, nsteps = 1.0, 4
    m = NNP.read(potentials["plain4096"])
    m.set_option("reaction_rows", 1)
    o = ob.OracleModel(potentials["plain4096"])
    md = MD(m, s, MASSES, v0, dt, NNP_FORCE)
    try:
        md.step(nsteps)
        st = md.state()
    finally:
        md.close()
        m.close()
    mass = np.array([MASSES[t] for t in s.types])[:, None]
    x, v = s.pos.copy(), v0.copy()
    E, _, F = o.compute(s.types, x, s.cell, ob.F_FORCE)
    for _ in range(nsteps):
        v += 0.5 * dt * F / (mass * CONV)
        x = wrap(x + v * dt, s.cell)
        E, _, F = o.compute(s.types, x, s.cell, ob.F_FORCE)
        v += 0.5 * dt * F / (mass * CONV)
    assert np.abs(st["pos"] - x).max() < 1e-9
    assert np.abs(st["vel"] - v).max() < 1e-11
    assert abs(st["epot"] - E) <= 1e-9 * max(abs(E), 1.0)


def run_md(path, s, v0, dt, calls, flags=NNP_FORCE, vlist=None, masses=MASSES):
    m = NNP.read(path)
    md = MD(m, s, masses, v0, dt, flags)
    try:
        if vlist:
            md.vlist(*vlist)
        for k in calls:
            md.step(k)
        st = md.state()
    finally:
        md.close()
        m.close()
    return st


@pytest.mark.parametrize("case", ["cell300", "plain64"])
def test_md_verlet_list_matches_builds_every_step(potentials, case):
    """A Verlet list of radius rc + 1 A rebuilt every 5 steps (the reference's MD driver,
    torch/torch.cpp:269) gives the trajectory of a fresh list every step: the same entries in the
    same order with the min-image flags recomputed, displacements equal up to rounding.  The
    64-atom cell (L = 10.9 A < 2 rc, 27 images per pair) exercises the min-image rule's flags."""
    if case == "cell300":
        s, rng = cell_300()
        path = potentials["plain4096"]
        v0, masses = velocities(s, rng, 600.0), MASSES
    else:
        s = synth.plain64()[0]
        rng = np.random.default_rng(3)
        path = potentials["plain64"]
        masses = [28.0855]
        v0 = velocities(s, rng, 600.0, masses)
    dt = 1.0
    ref = run_md(path, s, v0, dt, [12], masses=masses)
    for calls in ([12], [4, 4, 4], [7, 5]):
        st = run_md(path, s, v0, dt, calls, vlist=(7.0, 5), masses=masses)
        assert st["steps"] == 12
        np.testing.assert_allclose(wrap(st["pos"], s.cell), wrap(ref["pos"], s.cell), rtol=0, atol=1e-9)
        np.testing.assert_allclose(st["vel"], ref["vel"], rtol=0, atol=1e-10)
        np.testing.assert_allclose(st["force"], ref["force"], rtol=0, atol=1e-8 * np.abs(ref["force"]).max())
        assert abs(st["epot"] - ref["epot"]) <= 1e-10 * abs(ref["epot"])


def test_md_verlet_list_against_oracle_verlet(potentials):
    s, rng = cell_300(12)
    v0 = velocities(s, rng, 600.0)
    dt, nsteps = 1.0, 6
    st = run_md(potentials["plain4096"], s, v0, dt, [nsteps], vlist=(7.5, 3))
    o = ob.OracleModel(potentials["plain4096"])
    mass = np.array([MASSES[t] for t in s.types])[:, None]
    x, v = s.pos.copy(), v0.copy()
    E, _, F = o.compute(s.types, x, s.cell, ob.F_FORCE)
    for _ in range(nsteps):
        v += 0.5 * dt * F / (mass * CONV)
        x = wrap(x + v * dt, s.cell)
        E, _, F = o.compute(s.types, x, s.cell, ob.F_FORCE)
        v += 0.5 * dt * F / (mass * CONV)
    np.testing.assert_allclose(wrap(st["pos"], s.cell), x, rtol=0, atol=1e-9)
    assert np.abs(st["vel"] - v).max() < 1e-10
    assert abs(st["epot"] - E) <= 1e-9 * max(abs(E), 1.0)


def test_md_verlet_list_thin_skin_falls_back(potentials):
    """A skin of 1e-4 A is outrun at once: each call is redone with a fresh list every step, which
    is the stride-1 trajectory bit for bit."""
    s, rng = cell_300(5)
    v0 = velocities(s, rng, 900.0)
    ref = run_md(potentials["plain4096"], s, v0, 1.0, [6, 3])
    st = run_md(potentials["plain4096"], s, v0, 1.0, [6, 3], vlist=(6.0001, 4))
    for k in ("pos", "vel", "force"):
        assert np.array_equal(st[k], ref[k]), k
    assert st["epot"] == ref["epot"] and st["steps"] == ref["steps"]


def test_md_verlet_list_all_images(potentials):
    s = synth.plain64()[0]
    rng = np.random.default_rng(9)
    v0 = velocities(s, rng, 400.0, [28.0855])
    ref = run_md(potentials["plain64"], s, v0, 0.5, [10], flags=NNP_FORCE | NNP_ALL_IMAGES, masses=[28.0855])
    st = run_md(potentials["plain64"], s, v0, 0.5, [10], flags=NNP_FORCE | NNP_ALL_IMAGES, vlist=(7.2, 10),
                masses=[28.0855])
    np.testing.assert_allclose(wrap(st["pos"], s.cell), wrap(ref["pos"], s.cell), rtol=0, atol=1e-9)
    np.testing.assert_allclose(st["vel"], ref["vel"], rtol=0, atol=1e-10)
