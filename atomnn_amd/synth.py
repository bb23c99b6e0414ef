"""Synthetic workloads for the Behler-Parrinello energy/force engine.

Five workloads, each seeded with ``20261202 + k``:

* ``plain64``   - 1 species, 64 atoms, cubic L = 10.86 A, rc = 6.0, 8 G2 + 8 G4, 16-16-16-1.
* ``plain4096`` - 2 species 50/50, 4096 atoms at 0.05 /A^3, rc = 6.0, 32 G2 + 48 G4, 80-32-32-1.
* ``batch512``  - 512 triclinic structures of U{8..512} atoms (cube at 0.05 /A^3 plus a random
  symmetric strain of +-0.10), basis and networks of ``plain4096``.
* ``mixed64x4096`` - 64 structures of 4096 atoms, 3 species 40/30/30, 0.06 /A^3, rc = 6.5,
  36 G2 + 48 G4, 84-64-64-1.
* ``serve20k``  - a Poisson stream of structures with log-uniform atom counts in [16, 20000],
  Dirichlet(1,1,1) species proportions, 0.055 /A^3 with +-5% strain, basis and networks of
  ``mixed64x4096``.

Potentials are written in AtomNN's ``ann`` text format (the grammar of src/nnp/nnp.cpp:198-237,
BasisR::write basis_radial.cpp:107-116, BasisA::write basis_angular.cpp:97-106,
ANN::write ml/nn.cpp:1284-1324), so the engine and any AtomNN reader load the same numbers.
Angular "eta" in that format is a length (ietap = eta^-alpha, basis_angular.cpp:168-173); a
Behler exponent eta_B maps to eta = eta_B^-1/2 with alpha = 2.

Network inputs are preconditioned as AtomNN's ``pre_cond`` option does (inpb = -mean(G),
inpw = 1/std(G) per species, nn.cpp:1194), with the statistics taken from a calibration
structure of the same workload by :func:`numpy_descriptors`, a small vectorised numpy
evaluator that is independent of both the engine and the CPU oracle.
"""
from __future__ import annotations

import ctypes
import math
import os
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

SEED0 = 20261202

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBDIR = os.path.join(_HERE, "lib")


# ------------------------------------------------------------------ structures
@dataclass
class Structure:
    """One atomic structure: ``cell`` rows are the lattice vectors a, b, c (None = open)."""

    types: np.ndarray  # int32 [N]
    pos: np.ndarray  # float64 [N,3], Angstrom
    cell: Optional[np.ndarray]  # float64 [3,3], row q = lattice vector q

    @property
    def natoms(self) -> int:
        return int(self.types.shape[0])


_rsa_lib = None


def _rsa():
    global _rsa_lib
    if _rsa_lib is None:
        path = os.path.join(_LIBDIR, "libatomnn_synth.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing: run __graft_entry__.build() first")
        lib = ctypes.CDLL(path)
        lib.synth_rsa.restype = ctypes.c_int
        lib.synth_rsa.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_double, ctypes.c_uint64,
                                  ctypes.c_longlong, ctypes.c_void_p]
        _rsa_lib = lib
    return _rsa_lib


def random_positions(n: int, cell: np.ndarray, dmin: float, seed: int) -> np.ndarray:
    """Random sequential insertion with a periodic minimum separation ``dmin``."""
    lv = np.ascontiguousarray(cell, dtype=np.float64).reshape(9)
    out = np.empty((n, 3), dtype=np.float64)
    rc = _rsa().synth_rsa(n, lv.ctypes.data, float(dmin), int(seed) & (2**64 - 1), 10**7, out.ctypes.data)
    if rc != 0:
        raise RuntimeError(f"random sequential insertion failed (code {rc}) for n={n}")
    return out


def cubic_cell(n: int, rho: float) -> np.ndarray:
    return np.eye(3) * (n / rho) ** (1.0 / 3.0)


def strained_cell(n: int, rho: float, amp: float, rng: np.random.Generator) -> np.ndarray:
    """Cube at density rho, then (I + eps) with eps symmetric, entries U(-amp, amp) for i <= j."""
    L = (n / rho) ** (1.0 / 3.0)
    eps = np.zeros((3, 3))
    for i in range(3):
        for j in range(i, 3):
            eps[i, j] = eps[j, i] = rng.uniform(-amp, amp)
    R = (np.eye(3) + eps) @ (np.eye(3) * L)  # columns are lattice vectors
    return np.ascontiguousarray(R.T)  # rows = lattice vectors


def species_exact(n: int, fractions: Sequence[float], rng: np.random.Generator) -> np.ndarray:
    counts = [int(math.floor(f * n + 1e-9)) for f in fractions[:-1]]
    counts.append(n - sum(counts))
    t = np.concatenate([np.full(c, s, dtype=np.int32) for s, c in enumerate(counts)])
    rng.shuffle(t)
    return t.astype(np.int32)


# ------------------------------------------------------------------ potentials
@dataclass
class Potential:
    species: List[str]
    rc: float
    radial: Dict[Tuple[int, int], List[Tuple[float, float]]]  # (centre, nbr) -> [(rs, eta)]
    angular: Dict[Tuple[int, int, int], Tuple[int, List[Tuple[float, float, int]]]]  # (centre,j,k) j<=k
    layers: Dict[int, List[int]]  # species -> [nin, h..., 1]
    weights: Dict[int, List[np.ndarray]]  # species -> W_l (n_l x n_{l-1})
    biases: Dict[int, List[np.ndarray]]
    inpw: Dict[int, np.ndarray]
    inpb: Dict[int, np.ndarray]
    evac: Dict[int, float] = field(default_factory=dict)
    neuron: str = "TANH"
    norm: str = "UNIT"
    cut: str = "COS"
    phi_r: str = "GAUSSIAN"  # radial function (basis_radial.cpp:24-32)
    phi_a: str = "GAUSS"     # angular distance function (basis_angular.cpp:24-30)
    outw: Dict[int, float] = field(default_factory=dict)  # output scaling per species (default 1)
    outb: Dict[int, float] = field(default_factory=dict)  # output shift per species (default 0)

    @property
    def ntypes(self) -> int:
        return len(self.species)

    def write(self, path: str) -> None:
        """Write the potential in the ``ann`` text format (nnp.cpp:198-237)."""
        T = self.ntypes
        with open(path, "w") as f:
            f.write("ann\n")
            f.write(f"ntypes {T}\n")
            for s in range(T):
                line = f"name {self.species[s]} "
                if s in self.evac:
                    line += f"energy {self.evac[s]:f} "
                f.write(line + "\n")
            f.write(f"rc {self.rc:f}\n")
            for i in range(T):
                f.write(f"basis {self.species[i]}\n")
                for j in range(T):
                    f.write(f"basis_radial {self.species[j]}\n")
                    fr = self.radial[(i, j)]
                    f.write(f"BasisR {self.norm} {self.cut} {self.rc:f} {self.phi_r} {len(fr)}\n")
                    for rs, eta in fr:
                        f.write(f"\t{rs:f} {eta:f}\n")
                for j in range(T):
                    for k in range(j, T):
                        f.write(f"basis_angular {self.species[j]} {self.species[k]}\n")
                        alpha, fa = self.angular[(i, j, k)]
                        f.write(f"BasisA {self.norm} {self.cut} {self.rc:f} {self.phi_a} {len(fa)} {alpha}\n")
                        for eta, zeta, lam in fa:
                            f.write(f"\t{eta:f} {zeta:f} {lam}\n")
            for s in range(T):
                f.write(f"nn {self.species[s]}\n")
                lay = self.layers[s]
                f.write("nn " + " ".join(str(v) for v in lay) + " \n")
                f.write(f"neuron {self.neuron}\n")
                f.write("inpw " + "".join(f"{v:.15f} " for v in self.inpw[s]) + "\n")
                f.write("outw " + f"{self.outw.get(s, 1.0):.15f} " + "\n")
                f.write("inpb " + "".join(f"{v:.15f} " for v in self.inpb[s]) + "\n")
                f.write("outb " + f"{self.outb.get(s, 0.0):.15f} " + "\n")
                for l, b in enumerate(self.biases[s]):
                    f.write(f"bias[{l + 1}] " + "".join(f"{v:.15f} " for v in b) + "\n")
                for l, W in enumerate(self.weights[s]):
                    f.write(f"weight[{l},{l + 1}] " + "".join(f"{v:.15f} " for v in W.reshape(-1)) + "\n")

    def quantized(self) -> "Potential":
        """The potential as a reader sees it after the ``%f`` / ``%.15f`` round trip."""
        q6 = lambda x: float(f"{x:f}")
        q15 = lambda x: float(f"{x:.15f}")
        return Potential(
            species=list(self.species), rc=q6(self.rc),
            radial={k: [(q6(a), q6(b)) for a, b in v] for k, v in self.radial.items()},
            angular={k: (a, [(q6(e), q6(z), l) for e, z, l in v]) for k, (a, v) in self.angular.items()},
            layers=dict(self.layers),
            weights={s: [np.vectorize(q15)(W) for W in Ws] for s, Ws in self.weights.items()},
            biases={s: [np.vectorize(q15)(b) for b in bs] for s, bs in self.biases.items()},
            inpw={s: np.vectorize(q15)(v) for s, v in self.inpw.items()},
            inpb={s: np.vectorize(q15)(v) for s, v in self.inpb.items()},
            evac={s: q6(v) for s, v in self.evac.items()},
            neuron=self.neuron, norm=self.norm, cut=self.cut, phi_r=self.phi_r, phi_a=self.phi_a,
            outw={s: q15(v) for s, v in self.outw.items()}, outb={s: q15(v) for s, v in self.outb.items()})

    def ninput(self, s: int) -> int:
        T = self.ntypes
        n = sum(len(self.radial[(s, j)]) for j in range(T))
        n += sum(len(self.angular[(s, j, k)][1]) for j in range(T) for k in range(j, T))
        return n


def _xavier(rng: np.random.Generator, nout: int, nin: int) -> np.ndarray:
    return rng.normal(0.0, math.sqrt(2.0 / (nin + nout)), size=(nout, nin))


def make_potential(species: Sequence[str], rc: float, n_radial: int, eta_b_ang: Sequence[float],
                   zetas: Sequence[float], lambdas: Sequence[int], hidden: Sequence[int],
                   rng: np.random.Generator) -> Potential:
    """Radial: rs=0, eta log-spaced 0.01..2.0 (A^-2).  Angular: eta_B x zeta x lambda grid."""
    T = len(species)
    etas = [0.01 * (200.0 ** (m / (n_radial - 1))) for m in range(n_radial)] if n_radial > 1 else [0.01]
    radial = {(i, j): [(0.0, e) for e in etas] for i in range(T) for j in range(T)}
    afun = [(1.0 / math.sqrt(eb), float(z), int(l)) for eb in eta_b_ang for z in zetas for l in lambdas]
    angular = {(i, j, k): (2, list(afun)) for i in range(T) for j in range(T) for k in range(j, T)}
    nin = T * n_radial + (T * (T + 1) // 2) * len(afun)
    layers, weights, biases, inpw, inpb = {}, {}, {}, {}, {}
    for s in range(T):
        lay = [nin] + list(hidden) + [1]
        layers[s] = lay
        weights[s] = [_xavier(rng, lay[l + 1], lay[l]) for l in range(len(lay) - 1)]
        biases[s] = [np.zeros(lay[l + 1]) for l in range(len(lay) - 1)]
        inpw[s] = np.ones(nin)
        inpb[s] = np.zeros(nin)
    return Potential(list(species), rc, radial, angular, layers, weights, biases, inpw, inpb)


# ------------------------------------------------------------------ numpy descriptors
def _images(cell: np.ndarray, rc: float) -> np.ndarray:
    R = cell.T
    vol = abs(np.linalg.det(R))
    h = [vol / np.linalg.norm(np.cross(cell[(d + 1) % 3], cell[(d + 2) % 3])) for d in range(3)]
    m = [int(math.ceil(rc / h[d])) + 1 for d in range(3)]
    g = np.array([(a, b, c) for a in range(-m[0], m[0] + 1) for b in range(-m[1], m[1] + 1)
                  for c in range(-m[2], m[2] + 1)], dtype=np.float64)
    return g @ cell  # translations


def numpy_neighbors(s: Structure, rc: float):
    """All (j, translation) with 1e-6 < |r_i - r_j - T|^2 <= rc^2; returns per-atom r (P,3), j (P,)."""
    out = []
    if s.cell is None:
        T = np.zeros((1, 3))
    else:
        T = _images(s.cell, rc)
        inv = np.linalg.inv(s.cell.T)
    for i in range(s.natoms):
        d0 = s.pos[i] - s.pos  # (N,3)
        if s.cell is not None:
            f = d0 @ inv.T
            f -= np.round(f)
            d0 = f @ s.cell
        r = d0[:, None, :] - T[None, :, :]  # (N, nT, 3)
        d2 = np.einsum("ijk,ijk->ij", r, r)
        jj, tt = np.nonzero((d2 <= rc * rc) & (d2 > 1e-6))
        out.append((r[jj, tt], jj.astype(np.int64)))
    return out


def _fc(d, rc, cut="COS"):
    if cut == "COS":
        return np.where(d > rc, 0.0, 0.5 * (np.cos(d * math.pi / rc) + 1.0))
    raise NotImplementedError(cut)


def numpy_descriptors(pot: Potential, s: Structure) -> List[np.ndarray]:
    """Descriptor vectors by direct vectorised summation (GAUSSIAN radial, GAUSS angular).

    Layout: radial blocks by neighbour species, then angular blocks in lower-triangular
    order (0,0),(1,0),(1,1),(2,0),... (AtomNN nnh.cpp:95-121, lmat.hpp:43-46).
    """
    T = pot.ntypes
    norm = 1.0 if pot.norm == "UNIT" else 1.0 / (0.5 * 4.0 / 3.0 * (math.pi ** 2 - 6.0) / math.pi * pot.rc ** 3)
    nbrs = numpy_neighbors(s, pot.rc)
    res = []
    for i in range(s.natoms):
        ti = int(s.types[i])
        r, jj = nbrs[i]
        tj = s.types[jj]
        d = np.sqrt(np.einsum("ij,ij->i", r, r))
        G = []
        for b in range(T):
            m = tj == b
            for rs, eta in pot.radial[(ti, b)]:
                G.append(float(np.sum(np.exp(-eta * (d[m] - rs) ** 2) * _fc(d[m], pot.rc, pot.cut)) * norm))
        P = len(d)
        cosm = (r @ r.T) / np.outer(d, d) if P else np.zeros((0, 0))
        iu = np.triu_indices(P, 1)
        blocks = {}
        for j in range(T):
            for k in range(j, T):
                alpha, fa = pot.angular[(ti, j, k)]
                vals = []
                if P > 1:
                    c = cosm[iu]
                    a, b2 = iu
                    sel = ((tj[a] == j) & (tj[b2] == k)) | ((tj[a] == k) & (tj[b2] == j))
                    c = c[sel]
                    da, db = d[a][sel], d[b2][sel]
                    fcp = _fc(da, pot.rc, pot.cut) * _fc(db, pot.rc, pot.cut) * norm
                    for eta, zeta, lam in fa:
                        ietap = (1.0 / eta) ** alpha
                        vals.append(float(np.sum(np.abs(0.5 * (1.0 + lam * c)) ** zeta * fcp *
                                                 np.exp(-ietap * (da ** alpha + db ** alpha)))))
                else:
                    vals = [0.0] * len(fa)
                blocks[(j, k)] = vals
        for row in range(T):  # LMat linear order: (0,0),(1,0),(1,1),(2,0),...
            for col in range(row + 1):
                G.extend(blocks[(col, row)])
        res.append(np.array(G))
    return res


def precondition(pot: Potential, s: Structure) -> Potential:
    """Set inpb = -mean(G), inpw = 1/std(G) per species from a calibration structure."""
    Gs = numpy_descriptors(pot, s)
    for sp in range(pot.ntypes):
        rows = [Gs[i] for i in range(s.natoms) if int(s.types[i]) == sp]
        if not rows:
            continue
        A = np.vstack(rows)
        mu = A.mean(axis=0)
        sd = A.std(axis=0)
        sd = np.where(sd > 1e-8 * np.maximum(1.0, np.abs(mu)), sd, 1.0)
        pot.inpb[sp] = -mu
        pot.inpw[sp] = 1.0 / sd
    return pot


# ------------------------------------------------------------------ workloads
@dataclass
class Workload:
    name: str
    potential: Potential
    structures: List[Structure]
    description: str


def _pot_plain64(rng):
    return make_potential(["A"], 6.0, 8, [0.005, 0.05], [1.0, 4.0], [1, -1], [16, 16], rng)


def _pot_two_species(rng):
    return make_potential(["A", "B"], 6.0, 16, [0.001, 0.005, 0.02, 0.05], [1.0, 4.0], [1, -1], [32, 32], rng)


def _pot_three_species(rng):
    return make_potential(["A", "B", "C"], 6.5, 12, [0.005, 0.05], [1.0, 4.0], [1, -1], [64, 64], rng)


def _calibrate(pot: Potential, n: int, fractions, rho: float, dmin: float, seed: int) -> Potential:
    rng = np.random.default_rng(seed)
    cell = cubic_cell(n, rho)
    s = Structure(species_exact(n, fractions, rng), random_positions(n, cell, dmin, seed + 7), cell)
    return precondition(pot, s)


def workload_potential(name: str) -> Potential:
    k = {"plain64": 0, "plain4096": 1, "batch512": 1, "mixed64x4096": 3, "serve20k": 3}[name]
    rng = np.random.default_rng(SEED0 + k)
    if k == 0:
        pot = _pot_plain64(rng)
        return _calibrate(pot, 64, [1.0], 0.05, 2.0, SEED0 + 100)
    if k == 1:
        pot = _pot_two_species(rng)
        return _calibrate(pot, 128, [0.5, 0.5], 0.05, 2.0, SEED0 + 101)
    pot = _pot_three_species(rng)
    return _calibrate(pot, 128, [0.4, 0.3, 0.3], 0.06, 1.8, SEED0 + 103)


def plain64() -> List[Structure]:
    rng = np.random.default_rng(SEED0 + 0 + 1000)
    cell = np.eye(3) * 10.86
    return [Structure(np.zeros(64, dtype=np.int32), random_positions(64, cell, 2.0, SEED0 + 0), cell)]


def plain4096() -> List[Structure]:
    rng = np.random.default_rng(SEED0 + 1 + 1000)
    cell = cubic_cell(4096, 0.05)
    return [Structure(species_exact(4096, [0.5, 0.5], rng), random_positions(4096, cell, 2.0, SEED0 + 1), cell)]


def batch512(count: int = 512) -> List[Structure]:
    rng = np.random.default_rng(SEED0 + 2 + 1000)
    out = []
    for s in range(512):
        n = int(rng.integers(8, 513))
        cell = strained_cell(n, 0.05, 0.10, rng)
        types = species_exact(n, [0.5, 0.5], rng)
        seed = int(rng.integers(0, 2**62))
        if s < count:
            out.append(Structure(types, random_positions(n, cell, 1.8, seed), cell))
    return out


def mixed64x4096(count: int = 64) -> List[Structure]:
    rng = np.random.default_rng(SEED0 + 3 + 1000)
    out = []
    cell = cubic_cell(4096, 0.06)
    for s in range(64):
        types = species_exact(4096, [0.4, 0.3, 0.3], rng)
        seed = int(rng.integers(0, 2**62))
        if s < count:
            out.append(Structure(types, random_positions(4096, cell, 1.8, seed), cell))
    return out


@dataclass
class Request:
    arrival: float  # seconds after stream start
    structure: Structure


def serve_structure(rng: np.random.Generator) -> Structure:
    """One serving request: N = round(exp(U(ln 16, ln 20000))), Dirichlet(1,1,1) species
    proportions, cube at 0.055 /A^3 with a +-5% symmetric strain, 1.8 A minimum separation."""
    n = int(round(math.exp(rng.uniform(math.log(16), math.log(20000)))))
    p = rng.dirichlet([1.0, 1.0, 1.0])
    cell = strained_cell(n, 0.055, 0.05, rng)
    types = species_exact(n, list(p), rng)
    seed = int(rng.integers(0, 2**62))
    return Structure(types, random_positions(n, cell, 1.8, seed), cell)


def serve_pool(size: int, seed_offset: int = 0) -> List[Structure]:
    """A pool of distinct serving structures drawn from the request distribution."""
    rng = np.random.default_rng(SEED0 + 4 + 1000 + seed_offset)
    return [serve_structure(rng) for _ in range(size)]


def serve_arrivals(count: int, rate: float, seed_offset: int = 0) -> np.ndarray:
    """Open-loop Poisson arrival times (s) at `rate` requests/s."""
    rng = np.random.default_rng(SEED0 + 4 + 2000 + seed_offset)
    return np.cumsum(rng.exponential(1.0 / rate, size=count))


def serve_stream(count: int = 20000, rate: float = 1000.0, pool: Optional[List[Structure]] = None,
                 seed_offset: int = 0) -> List[Request]:
    """Poisson request stream; with a pool, request k carries pool[k % len(pool)] (every request is
    still evaluated in full), otherwise every request is a fresh draw."""
    t = serve_arrivals(count, rate, seed_offset)
    if pool is None:
        rng = np.random.default_rng(SEED0 + 4 + 1000 + seed_offset)
        return [Request(float(t[k]), serve_structure(rng)) for k in range(count)]
    return [Request(float(t[k]), pool[k % len(pool)]) for k in range(count)]


def write_workload_potential(name: str, path: str) -> Potential:
    pot = workload_potential(name)
    pot.write(path)
    return pot
