"""Python mirror of AtomNN's NNP evaluation interface over the engine's C ABI.

The engine is libatomnn_amd.so (HIP kernels for gfx950 behind include/nnp_capi.h).  This
module only marshals arrays; every number comes from the GPU.  There is no CPU fallback:
if the library or a GPU is missing, the calls raise.

Reference interface mirrored (AtomNN src/):
  NNP.read(path)          NNP::read                      nnp/nnp.cpp:182-191, 244-308
  NNP.index(name)         NNP::index                     nnp/nnp.hpp:83-84
  NNP.symm(struc)         NNP::init + NNP::symm          nnp/nnp.cpp:317-381
  NNP.energy(struc)       NNP::energy (with E_vac)       nnp/nnp.cpp:390-402
  NNP.force(struc)        NNP::force                     nnp/nnp.cpp:416-482
  NNP.compute(struc)      NNP::compute (no E_vac)        nnp/nnp.cpp:496-566
  NNP.neighbors(struc)    NeighborList::build            struc/neighbor.cpp:59-171
Errors raise NNPError carrying the engine's code and message (the reference throws
std::invalid_argument / std::runtime_error with the same texts).
"""
from __future__ import annotations

import ctypes
import os
from typing import List, Optional, Sequence, Tuple

import numpy as np

from .synth import Structure

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libatomnn_amd.so")

NNP_FORCE = 0x1
NNP_ATOM_ENERGY = 0x2
NNP_VACUUM_ENERGY = 0x4
NNP_ALL_IMAGES = 0x8
NNP_PREC_MIXED = 0x10

ERRORS = {-22: "EINVAL", -5: "EIO", -12: "ENOMEM", -100: "EDEVICE", -101: "EBUDGET", -102: "EUNSUPPORTED",
          -11: "EAGAIN"}


class NNPError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class _Struc(ctypes.Structure):
    _fields_ = [("natoms", ctypes.c_int32), ("type", ctypes.c_void_p), ("pos", ctypes.c_void_p),
                ("cell", ctypes.c_void_p)]


class _Result(ctypes.Structure):
    _fields_ = [("energy", ctypes.c_double), ("energy_atom", ctypes.c_void_p), ("force", ctypes.c_void_p)]


_lib = None


def lib():
    """Load the engine library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: the HIP engine must be built (__graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        vp, i32, i64, dp = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_double
        L.nnp_last_error.restype = ctypes.c_char_p
        L.nnp_load.argtypes = [ctypes.c_char_p, ctypes.POINTER(vp)]
        L.nnp_load.restype = ctypes.c_int
        L.nnp_free.argtypes = [vp]
        L.nnp_ntypes.argtypes = [vp]
        L.nnp_rc.argtypes = [vp]
        L.nnp_rc.restype = dp
        L.nnp_ninput.argtypes = [vp, ctypes.c_int]
        L.nnp_type_index.argtypes = [vp, ctypes.c_char_p]
        L.nnp_neighbors.argtypes = [vp, ctypes.POINTER(_Struc), vp, i64, vp, vp, vp, vp, vp]
        L.nnp_neighbors.restype = i64
        L.nnp_symm.argtypes = [vp, ctypes.POINTER(_Struc), vp, ctypes.c_int]
        L.nnp_compute.argtypes = [vp, ctypes.POINTER(_Struc), ctypes.c_uint32, ctypes.POINTER(_Result)]
        L.nnp_compute_batch.argtypes = [vp, ctypes.c_int, ctypes.POINTER(_Struc), ctypes.c_uint32,
                                        ctypes.POINTER(_Result)]
        L.nnp_compute_noalloc.argtypes = [vp, ctypes.POINTER(_Struc), ctypes.c_uint32, ctypes.POINTER(_Result)]
        L.nnp_resident_load.argtypes = [vp, ctypes.c_int, ctypes.POINTER(_Struc), ctypes.c_uint32]
        L.nnp_resident_run.argtypes = [vp, ctypes.c_uint32]
        L.nnp_resident_fetch.argtypes = [vp, ctypes.c_uint32, ctypes.POINTER(_Result)]
        L.nnp_sync.argtypes = [vp]
        L.nnp_profile.argtypes = [vp, ctypes.c_int]
        L.nnp_profile_read.argtypes = [vp, vp, vp, ctypes.c_int]
        L.nnp_info.argtypes = [vp, vp, ctypes.c_int]
        L.nnp_set_option.argtypes = [vp, ctypes.c_char_p, i64]
        L.nnp_md_create.argtypes = [vp, ctypes.POINTER(_Struc), vp, vp, dp, ctypes.c_uint32, ctypes.POINTER(vp)]
        L.nnp_md_step.argtypes = [vp, ctypes.c_int]
        L.nnp_md_vlist.argtypes = [vp, ctypes.c_double, ctypes.c_int]
        L.nnp_md_state.argtypes = [vp, vp, vp, vp, vp]
        L.nnp_md_destroy.argtypes = [vp]
        L.nnp_train_create.argtypes = [vp, ctypes.c_int, ctypes.POINTER(_Struc), vp, ctypes.POINTER(vp)]
        L.nnp_train_nparams.restype = i64
        L.nnp_train_nparams.argtypes = [vp, ctypes.c_int]
        L.nnp_train_get_params.argtypes = [vp, vp]
        L.nnp_train_error.argtypes = [vp, vp, ctypes.c_int, vp, ctypes.c_int, dp, vp, vp]
        L.nnp_train_set_forces.argtypes = [vp, vp]
        L.nnp_train_error_force.argtypes = [vp, vp, ctypes.c_int, vp, ctypes.c_int, dp, dp, ctypes.c_uint32, vp, vp]
        L.nnp_train_destroy.argtypes = [vp]
        L.nnp_md_destroy.restype = None
        _lib = L
    return _lib


def _check(code: int) -> int:
    if code < 0:
        raise NNPError(code, lib().nnp_last_error().decode())
    return code


class _Pinned:
    """Keeps the numpy arrays of a marshalled structure alive for the duration of a call."""

    def __init__(self, s: Structure):
        self.t = np.ascontiguousarray(s.types, dtype=np.int32)
        self.p = np.ascontiguousarray(s.pos, dtype=np.float64)
        self.c = None if s.cell is None else np.ascontiguousarray(s.cell, dtype=np.float64).reshape(9)
        self.c_struct = _Struc(int(self.t.shape[0]), self.t.ctypes.data, self.p.ctypes.data,
                               None if self.c is None else self.c.ctypes.data)


class NNP:
    """A potential resident on the GPU."""

    def __init__(self, handle, path: str):
        self._h = handle
        self.path = path

    @staticmethod
    def read(path: str) -> "NNP":
        h = ctypes.c_void_p()
        _check(lib().nnp_load(path.encode(), ctypes.byref(h)))
        return NNP(h, path)

    def close(self):
        if getattr(self, "_h", None):
            lib().nnp_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def ntypes(self) -> int:
        return lib().nnp_ntypes(self._h)

    @property
    def rc(self) -> float:
        return lib().nnp_rc(self._h)

    def ninput(self, t: int) -> int:
        return _check(lib().nnp_ninput(self._h, t))

    def index(self, name: str) -> int:
        return lib().nnp_type_index(self._h, name.encode())

    def neighbors(self, s: Structure) -> dict:
        ps = _Pinned(s)
        n = s.natoms
        counts = np.zeros(n, dtype=np.int32)
        tot = _check(lib().nnp_neighbors(self._h, ctypes.byref(ps.c_struct), counts.ctypes.data, 0, None, None,
                                         None, None, None))
        idx = np.zeros(tot, dtype=np.int32)
        img = np.zeros(tot, dtype=np.int32)
        mn = np.zeros(tot, dtype=np.int32)
        r = np.zeros((tot, 3))
        d = np.zeros(tot)
        _check(lib().nnp_neighbors(self._h, ctypes.byref(ps.c_struct), counts.ctypes.data, tot, idx.ctypes.data,
                                   img.ctypes.data, mn.ctypes.data, r.ctypes.data, d.ctypes.data))
        offsets = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
        return dict(counts=counts, offsets=offsets, index=idx, image=img, min=mn, r=r, d=d)

    def symm(self, s: Structure) -> np.ndarray:
        ps = _Pinned(s)
        ldg = max(self.ninput(t) for t in range(self.ntypes))
        G = np.zeros((s.natoms, ldg))
        _check(lib().nnp_symm(self._h, ctypes.byref(ps.c_struct), G.ctypes.data, ldg))
        return G

    def compute(self, s: Structure, flags: int = NNP_FORCE | NNP_ATOM_ENERGY | NNP_VACUUM_ENERGY
                ) -> Tuple[float, np.ndarray, Optional[np.ndarray]]:
        (out,) = self.compute_batch([s], flags)
        return out

    def compute_noalloc(self, s: Structure, flags: int = NNP_FORCE | NNP_ATOM_ENERGY | NNP_VACUUM_ENERGY
                        ) -> Tuple[float, np.ndarray, Optional[np.ndarray]]:
        """nnp_compute_noalloc: the same results as compute(), with no device or pinned memory
        allocated by the call (the library's fixed scratch array is the workspace)."""
        pin = _Pinned(s)
        ea = np.zeros(s.natoms)
        f = np.zeros((s.natoms, 3))
        res = _Result(0.0, ea.ctypes.data, f.ctypes.data)
        _check(lib().nnp_compute_noalloc(self._h, ctypes.byref(pin.c_struct), flags, ctypes.byref(res)))
        return res.energy, ea, (f if flags & NNP_FORCE else None)

    def compute_batch(self, strucs: Sequence[Structure], flags: int = NNP_FORCE | NNP_ATOM_ENERGY | NNP_VACUUM_ENERGY
                      ) -> List[Tuple[float, np.ndarray, Optional[np.ndarray]]]:
        pins = [_Pinned(s) for s in strucs]
        arr = (_Struc * len(strucs))(*[p.c_struct for p in pins])
        eas = [np.zeros(s.natoms) for s in strucs]
        fs = [np.zeros((s.natoms, 3)) for s in strucs]
        res = (_Result * len(strucs))(*[_Result(0.0, ea.ctypes.data, f.ctypes.data) for ea, f in zip(eas, fs)])
        _check(lib().nnp_compute_batch(self._h, len(strucs), arr, flags | NNP_ATOM_ENERGY, res))
        out = []
        for k in range(len(strucs)):
            out.append((res[k].energy, eas[k], fs[k] if flags & NNP_FORCE else None))
        return out

    def compute_into(self, s: Structure, flags: int = NNP_FORCE | NNP_ATOM_ENERGY | NNP_VACUUM_ENERGY):
        """A zero-argument callable that runs nnp_compute on s into fixed result arrays, as a C
        driver's loop calls it (the same host buffers every call, no per-call marshalling), and
        returns (energy, per-atom energies, forces) views of those arrays."""
        pin = _Pinned(s)
        ea = np.zeros(s.natoms)
        f = np.zeros((s.natoms, 3))
        res = _Result(0.0, ea.ctypes.data, f.ctypes.data)
        fn, h, st, rs = lib().nnp_compute, self._h, ctypes.byref(pin.c_struct), ctypes.byref(res)

        def call():
            _check(fn(h, st, flags, rs))
            return res.energy, ea, f

        call.keep = (pin, ea, f, res)  # the arrays live as long as the callable
        return call

    def set_option(self, key: str, value: int) -> None:
        _check(lib().nnp_set_option(self._h, key.encode(), int(value)))

    def info(self) -> dict:
        out = np.zeros(14, dtype=np.int64)
        _check(lib().nnp_info(self._h, out.ctypes.data, 14))
        keys = ["maxp", "max_entries", "atoms", "structures", "shape", "ne", "nc", "ldg", "pipeline",
                "fused_blocks_per_cu", "compute_units", "moments", "assemble_mode", "reactions_missed"]
        return dict(zip(keys, (int(v) for v in out)))

    def resident(self, strucs: Sequence[Structure], flags: int) -> "Resident":
        return Resident(self, strucs, flags)

    # reference-named conveniences
    def energy(self, s: Structure) -> float:
        return self.compute(s, NNP_VACUUM_ENERGY)[0]

    def force(self, s: Structure) -> np.ndarray:
        return self.compute(s, NNP_FORCE | NNP_VACUUM_ENERGY)[2]


STAGES = ["neighbors", "descriptors", "networks", "force_terms", "force_assembly", "energy"]


class Resident:
    """A batch kept in HBM: load once, enqueue passes (replayed from a captured hipGraph), fetch."""

    def __init__(self, model: NNP, strucs: Sequence[Structure], flags: int):
        self.m = model
        self.flags = flags
        self.pins = [_Pinned(s) for s in strucs]
        self.arr = (_Struc * len(strucs))(*[p.c_struct for p in self.pins])
        self.natoms = [s.natoms for s in strucs]
        _check(lib().nnp_resident_load(model._h, len(strucs), self.arr, flags))

    def run(self) -> None:
        _check(lib().nnp_resident_run(self.m._h, self.flags))

    def sync(self) -> None:
        _check(lib().nnp_sync(self.m._h))

    def fetch(self):
        eas = [np.zeros(n) for n in self.natoms]
        fs = [np.zeros((n, 3)) for n in self.natoms]
        res = (_Result * len(self.natoms))(*[_Result(0.0, ea.ctypes.data, f.ctypes.data) for ea, f in zip(eas, fs)])
        _check(lib().nnp_resident_fetch(self.m._h, self.flags, res))
        return [(res[k].energy, eas[k], fs[k] if self.flags & NNP_FORCE else None) for k in range(len(self.natoms))]

    def profile(self, on: bool, force_repeat: int = 1) -> None:
        """Stage timing on or off; force_repeat > 1 runs the force-term launch that many times
        back to back inside its events (the launch is idempotent) and counts each."""
        _check(lib().nnp_profile(self.m._h, max(1, int(force_repeat)) if on else 0))

    def stage_ms(self):
        ms = np.zeros(len(STAGES))
        calls = np.zeros(len(STAGES), dtype=np.int64)
        _check(lib().nnp_profile_read(self.m._h, ms.ctypes.data, calls.ctypes.data, len(STAGES)))
        return ms, calls


class MD:
    """Velocity-Verlet molecular dynamics on the GPU (METAL units: A, eV, amu, fs), the step of
    the reference's Verlet::compute (torch/integrator.cpp:170-188).  masses: one per species."""

    def __init__(self, model: NNP, s: Structure, masses: Sequence[float], vel=None, dt: float = 1.0,
                 flags: int = NNP_FORCE):
        self.n = s.natoms
        pin = _Pinned(s)
        m = np.ascontiguousarray(masses, dtype=np.float64)
        v = None if vel is None else np.ascontiguousarray(vel, dtype=np.float64).reshape(self.n, 3)
        h = ctypes.c_void_p()
        _check(lib().nnp_md_create(model._h, ctypes.byref(pin.c_struct), m.ctypes.data,
                                   None if v is None else v.ctypes.data, float(dt), flags, ctypes.byref(h)))
        self._h = h

    def vlist(self, rc_list: float, stride: int) -> None:
        """Verlet list of radius rc_list (>= rc) rebuilt every `stride` steps (1: every step)."""
        _check(lib().nnp_md_vlist(self._h, float(rc_list), int(stride)))

    def step(self, nsteps: int = 1) -> None:
        _check(lib().nnp_md_step(self._h, int(nsteps)))

    def state(self) -> dict:
        pos = np.zeros((self.n, 3))
        vel = np.zeros((self.n, 3))
        frc = np.zeros((self.n, 3))
        e = np.zeros(4)
        _check(lib().nnp_md_state(self._h, pos.ctypes.data, vel.ctypes.data, frc.ctypes.data, e.ctypes.data))
        return {"pos": pos, "vel": vel, "force": frc, "epot": float(e[0]), "ekin": float(e[1]),
                "temperature": float(e[2]), "steps": int(e[3])}

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().nnp_md_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


LOSS_MSE, LOSS_MAE, LOSS_HUBER, LOSS_ASINH = 0, 1, 2, 3


class Train:
    """The energy-training inner loop on the GPU: NNPTE::error's training part
    (src/nnp/nnpte.cpp:869-927) with Cost::grad (src/ml/nn.cpp:1554-1589).  Descriptors of the
    structures are computed once; error(x, batch) returns ([sum of losses, sum of dE^2], gradient)
    for the parameter vector x (species in order, each in ANN operator>> order)."""

    def __init__(self, model: NNP, structures: Sequence[Structure], energies):
        pins = [_Pinned(s) for s in structures]
        arr = (_Struc * len(structures))(*[p.c_struct for p in pins])
        e = np.ascontiguousarray(energies, dtype=np.float64)
        if e.shape[0] != len(structures):
            raise ValueError("one reference energy per structure")
        h = ctypes.c_void_p()
        _check(lib().nnp_train_create(model._h, len(structures), arr, e.ctypes.data, ctypes.byref(h)))
        self._h = h
        self.ntypes = model.ntypes

    def nparams(self, species: int = -1) -> int:
        n = lib().nnp_train_nparams(self._h, int(species))
        if n < 0:
            _check(int(n))
        return int(n)

    def params(self) -> np.ndarray:
        x = np.zeros(self.nparams())
        _check(lib().nnp_train_get_params(self._h, x.ctypes.data))
        return x

    def error(self, x, batch, loss: int = LOSS_MSE, delta: float = 1.0, grad: bool = True):
        xx = np.ascontiguousarray(x, dtype=np.float64)
        if xx.shape[0] != self.nparams():
            raise ValueError("parameter vector of the wrong length")
        b = np.ascontiguousarray(batch, dtype=np.int32)
        err = np.zeros(2)
        g = np.zeros(xx.shape[0]) if grad else None
        _check(lib().nnp_train_error(self._h, xx.ctypes.data, int(b.shape[0]), b.ctypes.data, int(loss), float(delta),
                                     err.ctypes.data, g.ctypes.data if grad else None))
        return err, g

    def set_forces(self, forces) -> None:
        """Reference forces of every structure (natoms x 3 each, in creation order)."""
        f = np.ascontiguousarray(np.concatenate([np.asarray(x, dtype=np.float64).reshape(-1, 3) for x in forces]))
        _check(lib().nnp_train_set_forces(self._h, f.ctypes.data))

    def error_force(self, x, batch, loss: int = LOSS_MSE, delta: float = 1.0, beta: float = 1.0,
                    all_images: bool = False, grad: bool = True):
        """Energy + force loss (NNPTEF::error's training part, src/nnp/nnptef.cpp:1040-1293):
        ([sum of losses, sum of dE^2, sum of |fdiff|^2], exact gradient)."""
        xx = np.ascontiguousarray(x, dtype=np.float64)
        if xx.shape[0] != self.nparams():
            raise ValueError("parameter vector of the wrong length")
        b = np.ascontiguousarray(batch, dtype=np.int32)
        err = np.zeros(3)
        g = np.zeros(xx.shape[0]) if grad else None
        _check(lib().nnp_train_error_force(self._h, xx.ctypes.data, int(b.shape[0]), b.ctypes.data, int(loss),
                                           float(delta), float(beta), NNP_ALL_IMAGES if all_images else 0,
                                           err.ctypes.data, g.ctypes.data if grad else None))
        return err, g

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().nnp_train_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
