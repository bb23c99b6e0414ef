"""Benchmark: energy+force atoms/s of the MI355X engine on one GPU.

Default workload (the plain FP64 GPU configuration): one periodic cubic cell of 4096 atoms,
two species 50/50 at 0.05 /A^3 (L = 43.4 A), rc = 6.0, 32 G2 + 48 G4 descriptors, one
80-32-32-1 tanh network per species, energy and forces in FP64.

A step is one full pass of the hot path over the resident batch: neighbour list (cell
list), descriptors, per-species networks (MFMA), per-entry force terms, force assembly and
the per-structure energy, with the structure already in HBM.  K steps are timed between two
stream synchronisations; a second, profiled pass of K steps brackets every stage with HIP
events to obtain the force kernel's launch duration for the roofline line.

Also reported:
  * "mixed": the precision form (FP32 descriptors, BF16 MFMA networks, FP64 reductions) on
    the 3-species 64 x 4096-atom batch beside the plain FP64 form on the same batch, with the
    energy / force error of the mixed form against the FP64 result;
  * "cpu_baseline": the CPU oracle (a restatement of the reference evaluation loop, the
    reference itself cannot be built) timed on this host's cores on one 4096-atom structure.

Prints one JSON line.
"""
import argparse
import ctypes
import json
import math
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from atomnn_amd import synth  # noqa: E402
from atomnn_amd import nnp as nnpmod  # noqa: E402
from atomnn_amd.nnp import NNP, NNP_FORCE, NNP_PREC_MIXED, NNP_VACUUM_ENERGY  # noqa: E402

METRIC = "E+F atoms/s, mixed vs plain FP64 at ≤1 meV/atom; p50/p99 request latency"
FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0


def resident_api():
    L = nnpmod.lib()
    vp = ctypes.c_void_p
    L.nnp_resident_load.argtypes = [vp, ctypes.c_int, ctypes.POINTER(nnpmod._Struc), ctypes.c_uint32]
    L.nnp_resident_run.argtypes = [vp, ctypes.c_uint32]
    L.nnp_resident_fetch.argtypes = [vp, ctypes.c_uint32, ctypes.POINTER(nnpmod._Result)]
    L.nnp_sync.argtypes = [vp]
    L.nnp_profile.argtypes = [vp, ctypes.c_int]
    L.nnp_profile_read.argtypes = [vp, vp, vp, ctypes.c_int]
    L.nnp_info.argtypes = [vp, vp, ctypes.c_int]
    return L


class Resident:
    def __init__(self, model: NNP, strucs, flags):
        self.L = resident_api()
        self.m = model
        self.flags = flags
        self.pins = [nnpmod._Pinned(s) for s in strucs]
        self.arr = (nnpmod._Struc * len(strucs))(*[p.c_struct for p in self.pins])
        self.strucs = strucs
        nnpmod._check(self.L.nnp_resident_load(model._h, len(strucs), self.arr, flags))

    def run(self):
        nnpmod._check(self.L.nnp_resident_run(self.m._h, self.flags))

    def sync(self):
        nnpmod._check(self.L.nnp_sync(self.m._h))

    def fetch(self):
        eas = [np.zeros(s.natoms) for s in self.strucs]
        fs = [np.zeros((s.natoms, 3)) for s in self.strucs]
        res = (nnpmod._Result * len(self.strucs))(
            *[nnpmod._Result(0.0, ea.ctypes.data, f.ctypes.data) for ea, f in zip(eas, fs)])
        nnpmod._check(self.L.nnp_resident_fetch(self.m._h, self.flags, res))
        return [(res[k].energy, eas[k], fs[k]) for k in range(len(self.strucs))]

    def profile(self, on: bool):
        nnpmod._check(self.L.nnp_profile(self.m._h, 1 if on else 0))

    def stage_ms(self):
        ms = np.zeros(6)
        calls = np.zeros(6, dtype=np.int64)
        nnpmod._check(self.L.nnp_profile_read(self.m._h, ms.ctypes.data, calls.ctypes.data, 6))
        return ms, calls


def timed_steps(R: Resident, steps: int, warmup: int):
    for _ in range(warmup):
        R.run()
    R.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        R.run()
    R.sync()
    return time.perf_counter() - t0


def profiled_steps(R: Resident, steps: int):
    R.profile(True)
    for _ in range(steps):
        R.run()
        R.sync()
    ms, calls = R.stage_ms()
    R.profile(False)
    return ms / np.maximum(calls, 1)


def work_counts(model: NNP, strucs):
    """Realised neighbour entries P and triplets T summed over atoms."""
    P = 0
    T = 0
    for s in strucs:
        c = model.neighbors(s)["counts"].astype(np.int64)
        P += int(c.sum())
        T += int((c * (c - 1) // 2).sum())
    return P, T


def force_ops(P, T, nR, nA):
    """Algorithmic ops of the force pass: c_p + 13 per (pair, radial fn), c_t + 28 per (triplet, angular fn),
    with c_p = 30 and c_t = 75 (the per-pair and per-triplet geometry and assembly counts)."""
    return P * (30 + 13 * nR) + T * (75 + 28 * nA)


def traffic_from_profile(path):
    """HBM bytes per launch from a committed rocprofv3 --pmc summary (FETCH_SIZE / WRITE_SIZE in KiB,
    already divided by the copy-kernel calibration factor); None when absent."""
    if not os.path.exists(path):
        return None
    vals = {}
    with open(path) as f:
        for line in f:
            if ":" in line:
                k, v = line.split(":", 1)
                vals[k.strip()] = v.strip()
    try:
        return float(vals["traffic_bytes_per_launch"])
    except (KeyError, ValueError):
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--mixed-structures", type=int, default=64)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); structures and requests shard "
              "across GPUs with no exchange, so the whole-job figure is N times the one-GPU figure.",
              file=sys.stderr)
        sys.exit(2)
    tmp = tempfile.mkdtemp()
    pot_path = os.path.join(tmp, "plain4096.ann")
    synth.workload_potential("plain4096").write(pot_path)
    strucs = synth.plain4096()
    natoms = sum(s.natoms for s in strucs)
    model = NNP.read(pot_path)
    flags = NNP_FORCE | NNP_VACUUM_ENERGY
    R = Resident(model, strucs, flags)
    sec = timed_steps(R, args.steps, args.warmup)
    value = natoms * args.steps / sec
    stage = profiled_steps(R, args.steps)
    P, T = work_counts(model, strucs)
    nR, nA = 16, 16
    ops = force_ops(P, T, nR, nA)
    t_force = stage[3] * 1e-3
    achieved = ops / t_force / 1e12
    traffic = traffic_from_profile(os.path.join(ROOT, "profiles", "k_force_plain4096_traffic.txt"))
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "atoms/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sec / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "plain4096: 1 cubic cell, 4096 atoms, 2 species, 0.05/A^3, rc 6.0, 32 G2 + 48 G4, "
                        "80-32-32-1 tanh, FP64 energy + forces",
            "atoms_per_step": natoms,
            "neighbor_entries": P,
            "triplets": T,
        },
        "stage_ms": {k: round(float(v), 5) for k, v in
                     zip(["neighbors", "descriptors", "networks", "force_terms", "force_assembly", "energy"], stage)},
        "roofline": {
            "kernel": "k_force<double,4,4> (per-entry force terms)",
            "bound": "valu",
            "achieved": achieved,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP64_PEAK_TFLOPS,
            "traffic": traffic,
            "ops_per_launch": ops,
        },
    }
    model_mixed = None
    if not args.no_mixed:
        mpath = os.path.join(tmp, "mixed64x4096.ann")
        synth.workload_potential("mixed64x4096").write(mpath)
        ms = synth.mixed64x4096(args.mixed_structures)
        mat = sum(s.natoms for s in ms)
        model_mixed = NNP.read(mpath)
        res = {}
        for label, fl in (("plain_fp64", NNP_FORCE | NNP_VACUUM_ENERGY),
                          ("mixed", NNP_FORCE | NNP_VACUUM_ENERGY | NNP_PREC_MIXED)):
            Rm = Resident(model_mixed, ms, fl)
            steps = max(3, args.steps // 4)
            secm = timed_steps(Rm, steps, 1)
            Rm.run()
            res[label] = (mat * steps / secm, Rm.fetch())
            del Rm
        ref = res["plain_fp64"][1]
        mix = res["mixed"][1]
        e_err = max(abs(a[0] - b[0]) / s.natoms for a, b, s in zip(mix, ref, ms))
        f_rms = math.sqrt(sum(float(np.sum((a[2] - b[2]) ** 2)) for a, b in zip(mix, ref)) / mat)
        out["mixed"] = {
            "workload": f"mixed64x4096: {len(ms)} cubic cells x 4096 atoms, 3 species 40/30/30, 0.06/A^3, rc 6.5, "
                        "36 G2 + 48 G4, 84-64-64-1; FP32 descriptors, BF16 MFMA networks, FP64 reductions",
            "atoms_per_s": res["mixed"][0],
            "plain_fp64_atoms_per_s": res["plain_fp64"][0],
            "speedup_vs_plain_fp64": res["mixed"][0] / res["plain_fp64"][0],
            "max_energy_err_meV_per_atom": e_err * 1e3,
            "force_rms_err_meV_per_A": f_rms * 1e3,
            "error_reference": "plain FP64 GPU form (parity-tested against the FP64 CPU oracle at 1e-10)",
        }
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_binding as ob

        om = ob.OracleModel(pot_path)
        sec_cpu, _ = om.bench(strucs, ob.F_FORCE | ob.F_VACUUM, nthreads=1)
        cpu_rate = natoms / sec_cpu
        out["cpu_baseline"] = {
            "value": cpu_rate,
            "unit": "atoms/s",
            "cores": 1,
            "kind": "port",
            "sample": "one plain4096 structure (4096 atoms): neighbour list + descriptors + both network passes + "
                      f"forces, single thread, {sec_cpu:.2f} s",
        }
        out["speedup_vs_cpu"] = value / cpu_rate
    print(json.dumps(out))
    model.close()
    if model_mixed is not None:
        model_mixed.close()


if __name__ == "__main__":
    main()
