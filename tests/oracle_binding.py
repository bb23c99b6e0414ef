"""ctypes binding to the CPU oracle (oracle/lib/libnnp_oracle.so).

Test infrastructure: the oracle is the checker for the HIP engine, never the thing
measured or shipped.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg load it.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional

import numpy as np

_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_ROOT, "oracle", "lib", "libnnp_oracle.so")
LIB_PATH_V4 = os.path.join(_ROOT, "oracle", "lib", "libnnp_oracle_v4.so")
_V4_FLAGS = {"avx512f", "avx512bw", "avx512cd", "avx512dq", "avx512vl"}


def cpu_has_v4() -> bool:
    """Whether this host's CPU runs x86-64-v4 code (AVX-512 F/BW/CD/DQ/VL)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    return _V4_FLAGS <= set(line.split(":", 1)[1].split())
    except OSError:
        pass
    return False


def use_host_isa_build() -> str:
    """Load the oracle built for the widest ISA this host supports (the CPU-baseline timing; the
    reference builds with -march=native).  Must be called before the first oracle call of the
    process; returns the -march the loaded library was built with."""
    global LIB_PATH
    if _lib is not None:
        raise RuntimeError("the oracle library is already loaded")
    if cpu_has_v4() and os.path.exists(LIB_PATH_V4):
        LIB_PATH = LIB_PATH_V4
        return "x86-64-v4"
    return "x86-64-v3"

F_FORCE = 1
F_ATOM_ENERGY = 2
F_VACUUM = 4
F_ALL_IMAGES = 8

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: build the oracle (make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        vp, ip, dp = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        L.orc_last_error.restype = ctypes.c_char_p
        L.orc_load.restype = vp
        L.orc_load.argtypes = [ctypes.c_char_p]
        L.orc_free.argtypes = [vp]
        L.orc_ntypes.argtypes = [vp]
        L.orc_rc.restype = dp
        L.orc_rc.argtypes = [vp]
        L.orc_ninput.argtypes = [vp, ip]
        L.orc_type_index.argtypes = [vp, ctypes.c_char_p]
        L.orc_cell_diff.argtypes = [vp, vp, vp, vp]
        L.orc_neighbors.restype = ctypes.c_longlong
        L.orc_neighbors.argtypes = [ip, vp, vp, vp, dp, vp, ctypes.c_longlong, vp, vp, vp, vp, vp, vp]
        L.orc_symm.argtypes = [vp, ip, vp, vp, vp, vp, ip]
        L.orc_network.argtypes = [vp, ip, vp, vp, vp]
        L.orc_compute.argtypes = [vp, ip, vp, vp, vp, ip, vp, vp, vp]
        L.orc_bench.restype = dp
        L.orc_bench.argtypes = [vp, ip, vp, vp, vp, vp, ip, ip, vp]
        L.orc_train_create.restype = vp
        L.orc_train_create.argtypes = [vp, ip, vp, vp, vp, vp, vp]
        L.orc_train_free.argtypes = [vp]
        L.orc_train_nparams.argtypes = [vp, ip]
        L.orc_train_get_params.argtypes = [vp, vp]
        L.orc_train_error.argtypes = [vp, vp, ip, vp, ip, dp, vp, vp]
        L.orc_train_set_forces.argtypes = [vp, vp]
        L.orc_train_error_force.argtypes = [vp, vp, ip, vp, ip, dp, dp, ip, vp, vp]
        L.orc_activation_d2.restype = dp
        L.orc_activation_d2.argtypes = [ip, dp]
        L.orc_activation.argtypes = [ip, dp, vp]
        _lib = L
    return _lib


def _err() -> str:
    return lib().orc_last_error().decode()


def _arr(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def _cell_ptr(cell):
    if cell is None:
        return None, None
    c = _arr(cell, np.float64).reshape(9)
    return c, c.ctypes.data


class OracleModel:
    def __init__(self, path: str):
        self.h = lib().orc_load(path.encode())
        if not self.h:
            raise ValueError(_err())

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_free(self.h)
            self.h = None

    @property
    def ntypes(self) -> int:
        return lib().orc_ntypes(self.h)

    @property
    def rc(self) -> float:
        return lib().orc_rc(self.h)

    def ninput(self, t: int) -> int:
        return lib().orc_ninput(self.h, t)

    def type_index(self, name: str) -> int:
        return lib().orc_type_index(self.h, name.encode())

    def symm(self, types, pos, cell) -> np.ndarray:
        t = _arr(types, np.int32)
        p = _arr(pos, np.float64)
        n = t.shape[0]
        ldg = max(self.ninput(k) for k in range(self.ntypes))
        G = np.zeros((n, ldg))
        c, cp = _cell_ptr(cell)
        if lib().orc_symm(self.h, n, cp, p.ctypes.data, t.ctypes.data, G.ctypes.data, ldg) != 0:
            raise ValueError(_err())
        return G

    def network(self, t: int, G) -> tuple:
        g = _arr(G, np.float64)
        out = np.zeros(1)
        dEdG = np.zeros(self.ninput(t))
        if lib().orc_network(self.h, t, g.ctypes.data, out.ctypes.data, dEdG.ctypes.data) != 0:
            raise ValueError(_err())
        return float(out[0]), dEdG

    def compute(self, types, pos, cell, flags: int = F_FORCE | F_ATOM_ENERGY | F_VACUUM):
        t = _arr(types, np.int32)
        p = _arr(pos, np.float64)
        n = t.shape[0]
        e = np.zeros(1)
        ea = np.zeros(n)
        f = np.zeros((n, 3))
        c, cp = _cell_ptr(cell)
        if lib().orc_compute(self.h, n, cp, p.ctypes.data, t.ctypes.data, flags, e.ctypes.data, ea.ctypes.data,
                             f.ctypes.data) != 0:
            raise ValueError(_err())
        return float(e[0]), ea, f

    def bench(self, structures, flags: int, nthreads: int = 1):
        """Wall seconds of nnpce's per-structure loop over ``structures`` (oracle CPU baseline)."""
        nat = _arr([s.natoms for s in structures], np.int32)
        cells = _arr(np.concatenate([np.asarray(s.cell, dtype=np.float64).reshape(9) for s in structures]),
                     np.float64)
        pos = _arr(np.concatenate([s.pos for s in structures]), np.float64)
        typ = _arr(np.concatenate([s.types for s in structures]), np.int32)
        en = np.zeros(len(structures))
        sec = lib().orc_bench(self.h, len(structures), nat.ctypes.data, cells.ctypes.data, pos.ctypes.data,
                              typ.ctypes.data, flags, nthreads, en.ctypes.data)
        if sec < 0:
            raise ValueError(_err())
        return sec, en


LOSS_MSE, LOSS_MAE, LOSS_HUBER, LOSS_ASINH = 0, 1, 2, 3


class OracleTrain:
    """NNPTE::error's training part on the CPU (descriptors computed once at creation)."""

    def __init__(self, model: OracleModel, structures, energies):
        nat = _arr([s.natoms for s in structures], np.int32)
        cells = _arr(np.concatenate([np.zeros(9) if s.cell is None else np.asarray(s.cell, dtype=np.float64).reshape(9)
                                     for s in structures]), np.float64)
        pos = _arr(np.concatenate([s.pos for s in structures]), np.float64)
        typ = _arr(np.concatenate([s.types for s in structures]), np.int32)
        e = _arr(energies, np.float64)
        self.ntypes = model.ntypes
        self.h = lib().orc_train_create(model.h, len(structures), nat.ctypes.data, cells.ctypes.data, pos.ctypes.data,
                                        typ.ctypes.data, e.ctypes.data)
        if not self.h:
            raise ValueError(_err())

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_train_free(self.h)
            self.h = None

    def nparams(self, t: int) -> int:
        return lib().orc_train_nparams(self.h, t)

    def params(self) -> np.ndarray:
        x = np.zeros(sum(self.nparams(t) for t in range(self.ntypes)))
        lib().orc_train_get_params(self.h, x.ctypes.data)
        return x

    def error(self, x, batch, loss: int = LOSS_MSE, delta: float = 1.0, grad: bool = True):
        xx = _arr(x, np.float64)
        b = _arr(batch, np.int32)
        err = np.zeros(2)
        g = np.zeros(xx.shape[0]) if grad else None
        if lib().orc_train_error(self.h, xx.ctypes.data, b.shape[0], b.ctypes.data, loss, delta, err.ctypes.data,
                                 g.ctypes.data if grad else None) != 0:
            raise ValueError(_err())
        return err, g

    def set_forces(self, forces):
        """Reference forces of every atom (structures in creation order, natoms x 3 each)."""
        f = _arr(np.concatenate([np.asarray(x, dtype=np.float64).reshape(-1, 3) for x in forces]), np.float64)
        lib().orc_train_set_forces(self.h, f.ctypes.data)

    def error_force(self, x, batch, loss: int = LOSS_MSE, delta: float = 1.0, beta: float = 1.0,
                    all_images: bool = False, grad: bool = True):
        """Energy + force loss ({loss, sum dE^2, sum |fdiff|^2}) and its gradient."""
        xx = _arr(x, np.float64)
        b = _arr(batch, np.int32)
        err = np.zeros(3)
        g = np.zeros(xx.shape[0]) if grad else None
        if lib().orc_train_error_force(self.h, xx.ctypes.data, b.shape[0], b.ctypes.data, loss, delta, beta,
                                       1 if all_images else 0, err.ctypes.data, g.ctypes.data if grad else None) != 0:
            raise ValueError(_err())
        return err, g


NEURONS = ["LINEAR", "SIGMOID", "TANH", "ISRU", "ARCTAN", "SOFTSIGN", "RELU", "SOFTPLUS", "ELU", "GELU", "SWISH",
           "MISH", "TANHRE", "LOGCOSH", "IQUAD", "IFABS", "SINSIG"]


def activation(name: str, z: float):
    """(a, a', a'') of a hidden-layer activation."""
    n = NEURONS.index(name) + 1
    ad = np.zeros(2)
    if lib().orc_activation(n, float(z), ad.ctypes.data) != 0:
        raise ValueError(_err())
    return ad[0], ad[1], lib().orc_activation_d2(n, float(z))


def neighbors(types, pos, cell, rc: float) -> dict:
    """The reference-ordered neighbour list (NeighborList::build) as flat arrays."""
    t = _arr(types, np.int32)
    p = _arr(pos, np.float64)
    n = t.shape[0]
    counts = np.zeros(n, dtype=np.int32)
    shell = np.zeros(3, dtype=np.int32)
    c, cp = _cell_ptr(cell)
    tot = lib().orc_neighbors(n, cp, p.ctypes.data, t.ctypes.data, rc, counts.ctypes.data, 0, None, None, None,
                              None, None, shell.ctypes.data)
    if tot < 0:
        raise ValueError(_err())
    idx = np.zeros(tot, dtype=np.int32)
    img = np.zeros(tot, dtype=np.int32)
    mn = np.zeros(tot, dtype=np.int32)
    r = np.zeros((tot, 3))
    d = np.zeros(tot)
    lib().orc_neighbors(n, cp, p.ctypes.data, t.ctypes.data, rc, counts.ctypes.data, tot, idx.ctypes.data,
                        img.ctypes.data, mn.ctypes.data, r.ctypes.data, d.ctypes.data, shell.ctypes.data)
    offsets = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    return dict(counts=counts, offsets=offsets, index=idx, image=img, min=mn, r=r, d=d, shell=shell)


def cell_diff(cell, v1, v2) -> np.ndarray:
    c = _arr(cell, np.float64).reshape(9)
    a = _arr(v1, np.float64)
    b = _arr(v2, np.float64)
    out = np.zeros(3)
    lib().orc_cell_diff(c.ctypes.data, a.ctypes.data, b.ctypes.data, out.ctypes.data)
    return out
