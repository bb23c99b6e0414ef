"""Benchmark: energy+force atoms/s of the MI355X engine on one GPU.

`value` (the plain FP64 GPU configuration): one periodic cubic cell of 4096 atoms, two species
50/50 at 0.05 /A^3 (L = 43.4 A), rc = 6.0, 32 G2 + 48 G4 descriptors, one 80-32-32-1 tanh
network per species, energy and forces in FP64, with the structure resident in HBM.  A step is
one full pass of the hot path: neighbour list (cell list), descriptors, per-species networks
(MFMA), per-entry force terms, force assembly and the per-structure energy.  K steps are timed
between two stream synchronisations; a second, profiled pass of K steps brackets every stage with
HIP events on the engine's stream to obtain the force kernel's launch duration for the roofline
(the force-term launch, which is idempotent, runs 8 times back to back inside its events).

Also reported:
  * "roofline": the force-term kernel's throughput in algorithmic operations (the fixed per-unit
    convention: per neighbour entry 30 + 13 per radial function, per triplet 75 + 28 per angular
    function) against the FP64 peak, the operations it actually executes (moment expansion), its
    HBM traffic and VALU-issue floor (from the committed rocprofv3 summaries in profiles/);
  * "end_to_end": the synchronous nnp_compute call from host positions to energies and forces in
    host memory (PCIe included), the metric as the reference's driver loop sees it;
  * "md": velocity-Verlet steps per second on the resident cell;
  * "mixed": the precision form (FP32 descriptors, split-BF16 MFMA networks, FP64 reductions)
    on the 3-species 64 x 4096-atom batch beside the plain FP64 form on the same batch, with stage
    times, its force kernel's roofline, and errors against the GPU FP64 form;
  * "batching": 512 uneven triclinic cells in one pass (separate kernels, and the persistent
    fused kernel);
  * "training": error + gradient calls of the energy-training loop, and of force training;
  * "serving": open-loop Poisson requests at fractions of the burst capacity, p50 / p99;
  * "ratios": the mixed serving form over the plain FP64 form (end to end and resident);
  * "cpu_baseline": the CPU oracle (a restatement of the reference evaluation loop; the reference
    itself cannot be built) timed on this host's cores: `value` on the 4096-atom cell with one
    thread, plus bounded subsets of every configuration with up to 16 threads over structures,
    the CPU model and the -march of the build.  The oracle's energies and forces of the first two
    cells of the precision batch and of eight serving requests are the FP64 reference of the
    "oracle_errors" lists (per structure) in "mixed" and "serving".

Prints one JSON line.
"""
import argparse
import json
import math
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from atomnn_amd import serve, synth  # noqa: E402
from atomnn_amd import nnp as nnpmod  # noqa: E402
from atomnn_amd.nnp import LOSS_MSE, MD, NNP, NNP_FORCE, NNP_PREC_MIXED, NNP_VACUUM_ENERGY, Resident, Train  # noqa: E402

METRIC = "E+F atoms/s, mixed vs plain FP64 at ≤1 meV/atom; p50/p99 request latency"
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0


def log(msg):
    """Progress line on stderr (the JSON result goes to stdout)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def timed_steps(R: Resident, steps: int, warmup: int):
    for _ in range(warmup):
        R.run()
    R.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        R.run()
    R.sync()
    return time.perf_counter() - t0


def profiled_steps(R: Resident, steps: int, force_repeat: int = 8):
    # the force-term launch runs force_repeat times back to back inside its stage events, so its
    # per-launch time is the kernel's duration with the events' own cost spread over the launches
    R.profile(True, force_repeat)
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


def force_ops_triplets(P, T, nR, nA):
    """Algorithmic ops of the triplet force pass: c_p + 13 per (pair, radial fn), c_t + 28 per (triplet,
    angular fn), with c_p = 30 and c_t = 75 (the per-pair and per-triplet geometry and assembly counts)."""
    return P * (30 + 13 * nR) + T * (75 + 28 * nA)


def force_ops_moments(P, N, S, nR, NE, stored_moments=True):
    """Executed ops of the moment force pass (35 monomials of degree <= 4, NE radial exponents, S
    species).  Per entry: 30 + 13 nR (cutoff, radial functions), 13 NE (angular radial factors),
    64 + 216 NE (monomials, the adjoint polynomial's value and gradient per exponent: 35 + 60 FMAs,
    and the force assembly), and 35 (2 + 2 NE) (moments: monomial and one FMA per exponent) unless
    the force kernel reads the moment tables the descriptor kernel left; per centre 35 NE S (3 S + 1)
    (adjoint moments)."""
    per_entry = 30 + 13 * nR + 13 * NE + 64 + 216 * NE + (0 if stored_moments else 35 * (2 + 2 * NE))
    return P * per_entry + N * 35 * NE * S * (3 * S + 1)


def traffic_from_profile(path, key="traffic_bytes_per_launch"):
    """A figure of a committed rocprofv3 --pmc summary (HBM bytes from FETCH_SIZE / WRITE_SIZE, scaled
    by the copy-kernel calibration factors); None when absent."""
    if not os.path.exists(path):
        return None
    vals = {}
    with open(path) as f:
        for line in f:
            if ":" in line:
                k, v = line.split(":", 1)
                vals[k.strip()] = v.strip()
    try:
        return float(vals[key])
    except (KeyError, ValueError):
        return None


def train_line(tmp, args):
    """Error + gradient of the energy loss over a resident training set (NNPTE::error's training
    part with Cost::grad): the mixed64x4096 structures and potential, FP64 throughout.  Each call
    uploads the parameter vector and the batch, runs the per-species networks forward and
    backward, sums the gradient rows and downloads the loss and the gradient."""
    tpath = os.path.join(tmp, "mixed64x4096.ann")
    pot = synth.workload_potential("mixed64x4096")
    pot.write(tpath)
    ts = synth.mixed64x4096(args.train_structures)
    tat = sum(s.natoms for s in ts)
    rng = np.random.default_rng(7)
    eref = np.array([rng.normal(-3.0, 0.2) * s.natoms for s in ts])
    model_t = NNP.read(tpath)
    tr = Train(model_t, ts, eref)
    x = tr.params()
    batch = list(range(len(ts)))
    tr.error(x, batch, LOSS_MSE)
    calls = max(3, args.steps // 2)
    t0 = time.perf_counter()
    for _ in range(calls):
        tr.error(x, batch, LOSS_MSE)
    sec = (time.perf_counter() - t0) / calls
    # per atom of species s: forward 2 n_l n_(l-1), back-propagation 2 n_l n_(l-1) (l >= 1), gradient rows
    # 2 n_l n_(l-1) multiply-adds over every weight matrix
    counts = np.bincount(np.concatenate([s.types for s in ts]), minlength=pot.ntypes)
    flops = 0.0
    for sp in range(pot.ntypes):
        w = [W.size for W in pot.weights[sp]]
        flops += counts[sp] * (4 * sum(w) + 2 * sum(w[1:]))
    line = {
        "workload": f"{len(ts)} mixed64x4096 structures ({tat} atoms, 3 species, 84-64-64-1 networks), MSE loss, "
                    "error + gradient over the whole set per call, FP64",
        "atoms_per_s": tat / sec,
        "ms_per_call": sec * 1e3,
        "nparams": tr.nparams(),
        "network_tflops": flops / sec / 1e12,
    }
    # force training on a quarter of the set: each call rebuilds neighbour lists and descriptors,
    # evaluates the forces at x, and runs the force pass backwards for the gradient
    fb = list(range(max(1, len(ts) // 4)))
    fat = sum(ts[b].natoms for b in fb)
    frng = np.random.default_rng(8)
    tr.set_forces([frng.normal(0, 0.5, (s.natoms, 3)) for s in ts])
    tr.error_force(x, fb, LOSS_MSE, 1.0, 1.0)
    t0 = time.perf_counter()
    for _ in range(calls):
        tr.error_force(x, fb, LOSS_MSE, 1.0, 1.0)
    fsec = (time.perf_counter() - t0) / calls
    line["force_training"] = {
        "workload": f"{len(fb)} mixed64x4096 structures ({fat} atoms), MSE energy + force loss (beta 1), "
                    "error + exact gradient per call, FP64",
        "atoms_per_s": fat / fsec,
        "ms_per_call": fsec * 1e3,
    }
    tr.close()
    model_t.close()
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_binding as ob

        sample = ts[:2]
        ot = ob.OracleTrain(ob.OracleModel(tpath), sample, eref[:2])
        xo = ot.params()
        reps = 3
        t0 = time.perf_counter()
        for _ in range(reps):
            ot.error(xo, [0, 1], ob.LOSS_MSE)
        sec_cpu = (time.perf_counter() - t0) / reps
        nat_cpu = sum(s.natoms for s in sample)
        line["cpu_baseline"] = {"value": nat_cpu / sec_cpu, "unit": "atoms/s", "cores": 1, "kind": "port",
                                "sample": f"2 structures ({nat_cpu} atoms), error + gradient, single thread, "
                                          f"{sec_cpu:.2f} s per call"}
    return line


def md_line(model, s, args):
    """Velocity-Verlet MD on the resident 4096-atom cell (half kick, drift and wrap, the full
    pass, half kick, kinetic energy; one captured graph per step): steps/s from a 300 K
    Maxwell-Boltzmann start, dt 0.5 fs.  (The synthetic cell is far from equilibrium, so it heats
    quickly; energy conservation is a test, not a benchmark figure.)"""
    rng = np.random.default_rng(11)
    masses = [30.0, 60.0]
    kT = 8.617333262e-5 * 300.0
    mv2e = 103.642696  # eV per amu A^2 / fs^2
    m_at = np.array(masses)[s.types]
    vel = rng.normal(0.0, 1.0, (s.natoms, 3)) * np.sqrt(kT / (m_at * mv2e))[:, None]
    vel -= vel.mean(axis=0)
    nsteps = max(50, args.steps * 5)

    def rate(vlist):
        md = MD(model, s, masses, vel, dt=0.5)
        if vlist:
            md.vlist(*vlist)
        md.step(10)
        e0 = md.state()
        t0 = time.perf_counter()
        md.step(nsteps)
        st = md.state()  # waits for the steps
        sec = time.perf_counter() - t0
        md.close()
        assert st["steps"] == e0["steps"] + nsteps
        return sec

    sec = rate(None)
    # a Verlet list of radius rc + 1 A rebuilt every 10 steps (the reference's MD driver rebuilds
    # its list every `stride` steps); 100 steps = 10 calls' worth of builds at the call's start
    secv = rate((model.rc + 1.0, 10))
    return {
        "workload": f"plain4096 cell, velocity Verlet, dt 0.5 fs, {nsteps} steps after 10 warm-up steps, FP64",
        "steps_per_s": nsteps / sec,
        "atom_steps_per_s": nsteps * s.natoms / sec,
        "us_per_step": sec / nsteps * 1e6,
        "verlet_list": {
            "list": "radius rc + 1.0 A, rebuilt every 10 steps; other steps refresh the entries' displacements",
            "steps_per_s": nsteps / secv,
            "atom_steps_per_s": nsteps * s.natoms / secv,
            "us_per_step": secv / nsteps * 1e6,
        },
    }


def valu_floor_from_profile(path, kernel):
    """The kernel's VALU-issue floor (us per launch: its SQ_INSTS_VALU x 4 cycles over the 1024
    SIMDs at 2.4 GHz) from the committed SQ-counter summary, or None."""
    try:
        with open(path) as f:
            for line in f:
                cells = [c.strip() for c in line.split("|")]
                if len(cells) > 8 and cells[1].replace("atomnn::", "") == kernel:
                    return float(cells[7])
    except (OSError, ValueError):
        pass
    return None


def host_cpu():
    """CPU model name and the CPUs this process may run on."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        cpus = len(os.sched_getaffinity(0))
    except AttributeError:
        cpus = os.cpu_count() or 1
    return model, cpus


def oracle_errors(got, ref, strucs):
    """Per-structure errors of GPU results against oracle results: |dE| / N (meV/atom) and the RMS
    over atoms of |dF| (meV/A)."""
    out = []
    for (E, _, F), (Eo, _, Fo), s in zip(got, ref, strucs):
        out.append({"atoms": s.natoms, "energy_meV_per_atom": abs(E - Eo) / s.natoms * 1e3,
                    "force_rms_meV_per_A": float(np.sqrt(np.mean(np.sum((F - Fo) ** 2, axis=1)))) * 1e3})
    return out


def cpu_baseline_leg(pot_paths, subsets, gpu_rates, march):
    """The CPU oracle (a restatement of the reference's per-structure loop, nnpce.cpp:275-279:
    NeighborList::build + NNP::symm + NNP::energy + NNP::force with both network passes) on this
    host's cores, on bounded subsets of each configuration, with up to 16
    threads over structures as nnpce spreads structures over MPI ranks (nnpce.cpp:222).  Also
    returns the oracle's energies and forces of the accuracy samples (the FP64 reference the
    mixed and serving results are checked against)."""
    import oracle_binding as ob

    model_name, cpus = host_cpu()
    threads = min(16, cpus)
    flags = ob.F_FORCE | ob.F_VACUUM
    oracles = {k: ob.OracleModel(v) for k, v in pot_paths.items()}

    def rate(key, strucs, nthreads, reps=1):
        sec = 0.0
        for _ in range(reps):
            sec += oracles[key].bench(strucs, flags, nthreads=nthreads)[0]
        return sum(s.natoms for s in strucs) * reps / sec, sec

    leg = {"cpu_model": model_name, "host_cpus": cpus, "threads_max": threads, "march": march,
           "compiler": "g++ -O3 -std=gnu++17 (oracle/Makefile)", "configs": {}}
    c1_rate, c1_sec = rate("plain4096", subsets["plain4096"], 1)
    leg["value"] = c1_rate
    leg["sample"] = (f"one plain4096 structure (4096 atoms): neighbour list + descriptors + both network passes + "
                     f"forces, 1 thread, {c1_sec:.2f} s")
    cfg = leg["configs"]
    r0, _ = rate("plain64", subsets["plain64"], 1, reps=10)
    cfg["plain64"] = {"sample": "the 64-atom cell, 10 repetitions", "threads_1": r0}
    cfg["plain4096"] = {"sample": "the 4096-atom cell", "threads_1": c1_rate}
    b = subsets["batch512"]
    r1, _ = rate("plain4096", b[:8], 1)
    rn, _ = rate("plain4096", b[:64], threads)
    cfg["batch512"] = {"sample": f"first 8 cells (1 thread) and first 64 cells ({sum(s.natoms for s in b[:64])} atoms, "
                                 f"{threads} threads)", "threads_1": r1, f"threads_{threads}": rn}
    # precision batch: the first structures one at a time (their forces are the accuracy reference)
    m3 = subsets["mixed64x4096"]
    t0 = time.perf_counter()
    ref3 = [oracles["mixed64x4096"].compute(s.types, s.pos, s.cell, flags) for s in m3]
    sec3 = time.perf_counter() - t0
    cfg["mixed64x4096"] = {"sample": f"first {len(m3)} of the 64 cells ({sum(s.natoms for s in m3)} atoms), 1 thread",
                           "threads_1": sum(s.natoms for s in m3) / sec3}
    sv = subsets["serve20k"]
    rn4, sec4 = rate("serve20k", sv, threads)
    cfg["serve20k"] = {"sample": f"first {len(sv)} requests of the stream ({sum(s.natoms for s in sv)} atoms), "
                                 f"{threads} threads", f"threads_{threads}": rn4}
    acc = subsets["serve_accuracy"]
    ref4 = [oracles["serve20k"].compute(s.types, s.pos, s.cell, flags) for s in acc]
    for key, g in gpu_rates.items():
        if key in cfg and g:
            best = max(v for k, v in cfg[key].items() if k.startswith("threads_"))
            cfg[key]["gpu_atoms_per_s"] = g
            cfg[key]["gpu_speedup_vs_best_cpu"] = g / best
    return leg, ref3, ref4


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--mixed-structures", type=int, default=64)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-train", action="store_true")
    ap.add_argument("--no-md", action="store_true")
    ap.add_argument("--train-structures", type=int, default=64)
    ap.add_argument("--serve-requests", type=int, default=20000)
    ap.add_argument("--serve-pool", type=int, default=1024)
    ap.add_argument("--serve-load", type=float, default=0.5,
                    help="arrival rate of the reported line, as a fraction of the measured burst capacity")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run a plain device copy of BYTES bytes 5 times (HBM counter calibration) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch

        n = args.copy_calibration // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": n * 4, "launches": 5}))
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); structures and requests shard "
              "across GPUs with no exchange, so the whole-job figure is N times the one-GPU figure.",
              file=sys.stderr)
        sys.exit(2)
    march = None
    if not args.no_cpu_baseline:
        # the oracle build for this host's widest ISA (the reference builds with -march=native);
        # chosen before any oracle call of the process
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_binding as ob

        march = ob.use_host_isa_build()
    tmp = tempfile.mkdtemp()
    pot_paths = {}
    for name in ("plain64", "plain4096", "mixed64x4096", "serve20k"):
        pot_paths[name] = os.path.join(tmp, f"{name}.ann")
        synth.workload_potential(name).write(pot_paths[name])
    pot_path = pot_paths["plain4096"]
    strucs = synth.plain4096()
    natoms = sum(s.natoms for s in strucs)
    model = NNP.read(pot_path)
    flags = NNP_FORCE | NNP_VACUUM_ENERGY
    log("plain4096 resident passes")
    R = Resident(model, strucs, flags)
    sec = timed_steps(R, args.steps, args.warmup)
    value = natoms * args.steps / sec
    stage = profiled_steps(R, args.steps)
    P, T = work_counts(model, strucs)
    nR, nA, NE, S = 16, 16, 4, 2
    mom_forces = bool(model.info()["moments"] & 2)
    # roofline: the force pass's algorithmic operations by the fixed per-unit convention (per
    # entry 30 + 13 per radial function, per triplet 75 + 28 per angular function, from the
    # reference expressions), over the kernel's launch time measured with HIP events on the
    # engine's stream; the operations the moment kernel actually executes are reported beside it
    ops = force_ops_triplets(P, T, nR, nA)
    ops_exec = force_ops_moments(P, natoms, S, nR, NE) if mom_forces else ops
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
                        "80-32-32-1 tanh, FP64 energy + forces, structure resident in HBM",
            "atoms_per_step": natoms,
            "neighbor_entries": P,
            "triplets": T,
        },
        # with forces, the per-structure energy sums run as extra blocks of the assembly launch
        "stage_ms": {k: round(float(v), 5) for k, v in
                     zip(["neighbors", "descriptors", "networks", "force_terms", "force_assembly_and_energy"], stage)},
        "roofline": {
            "kernel": ("k_force_mom<double,4,1> (per-entry force terms, moment expansion)" if mom_forces
                       else "k_force<double,4,4> (per-entry force terms, triplet loops)"),
            "bound": "valu",
            "achieved": achieved,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP64_PEAK_TFLOPS,
            "traffic": traffic,
            "ops_per_launch": ops,
            "ops_convention": "P (30 + 13 nR) + T (75 + 28 nA): the reference's pair and triplet force expressions",
            # the moment expansion reaches the same forces with far fewer operations (O(P) per
            # centre instead of O(P^2)), so the rate by this convention can pass the ALU peak
            # (frac > 1); executed_frac is the fraction of the peak the kernel's own operations reach
            "ops_note": "algorithmic rate by the reference's operation count; the kernel executes "
                        "executed_ops_per_launch, so frac can exceed 1 while executed_frac stays below it",
            # what the moment kernel executes (35 monomials per entry instead of the triplet loops)
            "executed_ops_per_launch": ops_exec,
            "executed_frac": ops_exec / t_force / 1e12 / FP64_PEAK_TFLOPS,
            # counter view: the VALU instructions the kernel issues need this long at one per
            # 4 cycles on every SIMD; the launch time over it is the issue efficiency
            "valu_issue_floor_us": valu_floor_from_profile(
                os.path.join(ROOT, "profiles", "plain4096_fp64_sq_counters.md"),
                "k_force_mom<double, 4, 1>" if mom_forces else "k_force<double, 4, 4, false>"),
            "launch_us": t_force * 1e6,
        },
    }
    # the metric as the reference's driver sees it: the synchronous C ABI call from host positions
    # to energies and forces in host memory (upload, the pass, download; PCIe included), called in
    # a loop on the same host arrays as nnpce's loop does
    call = model.compute_into(strucs[0], flags)
    for _ in range(max(args.warmup, 8)):  # the runtime loads copy and kernel code on first uses
        call()
    ncall = max(args.steps, 100)
    t0 = time.perf_counter()
    for _ in range(ncall):
        call()
    hsec = (time.perf_counter() - t0) / ncall
    e2e = natoms / hsec
    out["end_to_end"] = {
        "call": "nnp_compute on the plain4096 cell (host positions in; host energy, per-atom energies and forces "
                "out), as nnpce's loop calls it",
        "atoms_per_s": e2e,
        "ms_per_call": hsec * 1e3,
    }
    if not args.no_md:
        log("md")
        out["md"] = md_line(model, strucs[0], args)
    model_mixed = None
    gpu_rates = {"plain4096": value}
    mixed_results = None
    ms = synth.mixed64x4096(args.mixed_structures)
    if not args.no_mixed:
        log("mixed64x4096")
        mat = sum(s.natoms for s in ms)
        model_mixed = NNP.read(pot_paths["mixed64x4096"])
        res = {}
        for label, fl in (("plain_fp64", NNP_FORCE | NNP_VACUUM_ENERGY),
                          ("mixed", NNP_FORCE | NNP_VACUUM_ENERGY | NNP_PREC_MIXED)):
            Rm = Resident(model_mixed, ms, fl)
            steps = max(3, args.steps // 4)
            secm = timed_steps(Rm, steps, 1)
            Rm.run()
            got = Rm.fetch()
            st = profiled_steps(Rm, 2, 1)
            res[label] = (mat * steps / secm, got, st, secm / steps)
            del Rm
            # the same batch through the synchronous call (host positions in, host results out)
            for _ in range(2):
                model_mixed.compute_batch(ms, fl)
            t0 = time.perf_counter()
            for _ in range(3):
                model_mixed.compute_batch(ms, fl)
            res[label + "_e2e"] = mat * 3 / (time.perf_counter() - t0)
        ref = res["plain_fp64"][1]
        mixed_results = res["mixed"][1]
        e_err = max(abs(a[0] - b[0]) / s.natoms for a, b, s in zip(mixed_results, ref, ms))
        f_rms = math.sqrt(sum(float(np.sum((a[2] - b[2]) ** 2)) for a, b in zip(mixed_results, ref)) / mat)
        Pm, Tm = work_counts(model_mixed, ms)
        stm = res["mixed"][2]
        ops_m = force_ops_triplets(Pm, Tm, 12, 8)
        out["mixed"] = {
            "workload": f"mixed64x4096: {len(ms)} cubic cells x 4096 atoms, 3 species 40/30/30, 0.06/A^3, rc 6.5, "
                        "36 G2 + 48 G4, 84-64-64-1; FP32 descriptors, BF16 MFMA networks, FP64 reductions",
            "atoms_per_s": res["mixed"][0],
            "ms_per_pass": res["mixed"][3] * 1e3,
            "plain_fp64_atoms_per_s": res["plain_fp64"][0],
            "plain_fp64_ms_per_pass": res["plain_fp64"][3] * 1e3,
            "speedup_vs_plain_fp64": res["mixed"][0] / res["plain_fp64"][0],
            # the same ratio through nnp_compute_batch (upload, pass, download; PCIe included)
            "end_to_end": {"mixed_atoms_per_s": res["mixed_e2e"], "plain_fp64_atoms_per_s": res["plain_fp64_e2e"],
                           "speedup_vs_plain_fp64": res["mixed_e2e"] / res["plain_fp64_e2e"]},
            "neighbor_entries": Pm,
            "triplets": Tm,
            "stage_ms": {k: round(float(v), 4) for k, v in
                         zip(["neighbors", "descriptors", "networks", "force_terms", "force_assembly_and_energy"], stm)},
            "plain_fp64_stage_ms": {k: round(float(v), 4) for k, v in
                                    zip(["neighbors", "descriptors", "networks", "force_terms",
                                         "force_assembly_and_energy"], res["plain_fp64"][2])},
            "roofline": {
                "kernel": "k_force_mom<float,2,1> (per-entry force terms, FP32)",
                "bound": "valu",
                "achieved": ops_m / (stm[3] * 1e-3) / 1e12,
                "peak": FP32_PEAK_TFLOPS,
                "unit": "TFLOP/s",
                "frac": ops_m / (stm[3] * 1e-3) / 1e12 / FP32_PEAK_TFLOPS,
                "ops_per_launch": ops_m,
                "launch_us": stm[3] * 1e3,
                # per-atom HBM bytes of the kernel (profiled on 16 of the cells) times the batch's atoms
                "traffic": (lambda b: b * mat if b else None)(traffic_from_profile(
                    os.path.join(ROOT, "profiles", "k_force_mixed64x4096_traffic.txt"), "traffic_bytes_per_atom")),
            },
            "max_energy_err_meV_per_atom_vs_gpu_fp64": e_err * 1e3,
            "force_rms_err_meV_per_A_vs_gpu_fp64": f_rms * 1e3,
        }
        gpu_rates["mixed64x4096"] = res["mixed"][0]
    if not args.no_batch:
        log("batch512")
        bstrucs = synth.batch512()
        bat = sum(s.natoms for s in bstrucs)
        model_b = NNP.read(pot_path)
        res = {}
        for pipe in (0, 1):
            model_b.set_option("pipeline", pipe)
            for label, fl in (("fp64", NNP_FORCE | NNP_VACUUM_ENERGY),
                              ("mixed", NNP_FORCE | NNP_VACUUM_ENERGY | NNP_PREC_MIXED)):
                Rb = Resident(model_b, bstrucs, fl)
                steps = max(3, args.steps // 2)
                secb = timed_steps(Rb, steps, 1)
                res[(pipe, label)] = (bat * steps / secb, model_b.info()["fused_blocks_per_cu"], secb / steps * 1e3)
                del Rb
        out["batching"] = {
            "workload": "batch512: 512 triclinic cells of U{8..512} atoms (cube at 0.05/A^3 + symmetric strain "
                        "+-0.10), 2 species, rc 6.0, 80-32-32-1; packed into one pass, hipGraph replay",
            "atoms": bat,
            "pipeline": "separate kernels (neighbours, descriptors, MFMA networks, force terms, assembly)",
            "fp64_atoms_per_s": res[(0, "fp64")][0],
            "fp64_ms_per_pass": res[(0, "fp64")][2],
            "mixed_atoms_per_s": res[(0, "mixed")][0],
            "mixed_ms_per_pass": res[(0, "mixed")][2],
            "fused_kernel": {
                "pipeline": "one persistent fused descriptor+network+force kernel, autotuned blocks per CU",
                "fp64_atoms_per_s": res[(1, "fp64")][0],
                "fp64_blocks_per_cu": res[(1, "fp64")][1],
                "mixed_atoms_per_s": res[(1, "mixed")][0],
                "mixed_blocks_per_cu": res[(1, "mixed")][1],
            },
        }
        gpu_rates["batch512"] = res[(0, "fp64")][0]
        model_b.close()
    if not args.no_train:
        log("training")
        out["training"] = train_line(tmp, args)
    pool = synth.serve_pool(args.serve_pool)
    serve_acc = [p for p in pool if p.natoms <= 3000][:8]
    serve_acc_results = None
    if not args.no_serve:
        log("serving")
        mean_atoms = float(np.mean([p.natoms for p in pool]))
        model_s = NNP.read(pot_paths["serve20k"])
        sflags = NNP_FORCE | NNP_VACUUM_ENERGY | NNP_PREC_MIXED
        srv = serve.Server(model_s, sflags, lanes=3)
        # capacity probe: everything at t = 0, after one untimed burst that sizes the lanes'
        # workspaces (first-use allocations are not part of steady-state serving)
        serve.run_stream(srv, pool[:512], np.zeros(512))
        bursts = sorted((serve.run_stream(srv, pool[:512], np.zeros(512)) for _ in range(3)),
                        key=lambda r: r.atoms_per_s)
        burst = bursts[1]  # median of three
        structs = [pool[k % len(pool)] for k in range(args.serve_requests)]
        # latency against load: open-loop streams at fractions of the burst capacity (a burst
        # forms large batches, so the sustainable rate of a Poisson stream is lower)
        sweep = []
        for load in sorted({0.3, args.serve_load, 0.7}):
            r = serve.run_stream(srv, structs, synth.serve_arrivals(args.serve_requests, load * burst.atoms_per_s / mean_atoms))
            sweep.append((load, r))
        rate = args.serve_load * burst.atoms_per_s / mean_atoms
        res = [r for load, r in sweep if load == args.serve_load][0]
        # served (mixed) results of an accuracy sample, against the GPU's FP64 form here and
        # against the oracle in the CPU-baseline leg
        serve_acc_results = serve.run_stream(srv, serve_acc, np.zeros(len(serve_acc)), keep_results=True).results
        ref = model_s.compute_batch(serve_acc, NNP_FORCE | NNP_VACUUM_ENERGY)
        e_err = max(abs(a[0] - b[0]) / s.natoms for a, b, s in zip(serve_acc_results, ref, serve_acc))
        f_rms = math.sqrt(sum(float(np.sum((a[2] - b[2]) ** 2)) for a, b in zip(serve_acc_results, ref)) /
                          sum(s.natoms for s in serve_acc))
        srv.close()
        out["serving"] = {
            "workload": f"serve20k: {args.serve_requests} open-loop Poisson requests at {args.serve_load:g} of the burst "
                        "capacity, atoms log-uniform 16-20000 "
                        f"(pool of {args.serve_pool} distinct structures cycled), 3 species Dirichlet(1,1,1), "
                        "0.055/A^3 +-5% strain, mixed precision, separate kernels, 3 lanes",
            "atoms_per_s": res.atoms_per_s,
            "requests_per_s": res.requests / res.wall_s,
            "arrival_rate_per_s": rate,
            "p50_ms": res.p50_ms,
            "p99_ms": res.p99_ms,
            "mean_ms": res.mean_ms,
            "capacity_atoms_per_s": burst.atoms_per_s,
            "batches": res.stats["batches"],
            "max_concurrent_batches": res.stats["max_concurrent_batches"],
            "peak_workspace_bytes": res.stats["peak_bytes"],
            "max_energy_err_meV_per_atom_vs_gpu_fp64": e_err * 1e3,
            "force_rms_err_meV_per_A_vs_gpu_fp64": f_rms * 1e3,
            "load": args.serve_load,
            "load_sweep": [{"load": load, "arrival_rate_per_s": load * burst.atoms_per_s / mean_atoms,
                            "atoms_per_s": r.atoms_per_s, "p50_ms": r.p50_ms, "p99_ms": r.p99_ms}
                           for load, r in sweep],
        }
        gpu_rates["serve20k"] = burst.atoms_per_s
        # the target ratio, mixed batched serving over the plain FP64 GPU form, both ways the plain
        # form can be timed
        out["ratios"] = {
            "serving_capacity_vs_plain_fp64_end_to_end": burst.atoms_per_s / e2e,
            "serving_capacity_vs_plain_fp64_resident": burst.atoms_per_s / value,
            "serving_at_load_vs_plain_fp64_end_to_end": res.atoms_per_s / e2e,
        }
        if "mixed" in out:
            out["ratios"]["mixed_vs_plain_fp64_same_batch_resident"] = out["mixed"]["speedup_vs_plain_fp64"]
        model_s.close()
    if not args.no_cpu_baseline:
        log("cpu baseline")
        subsets = {
            "plain64": synth.plain64(),
            "plain4096": strucs,
            "batch512": synth.batch512(64),
            "mixed64x4096": ms[:2],
            "serve20k": pool[:24],
            "serve_accuracy": serve_acc,
        }
        leg, ref3, ref4 = cpu_baseline_leg(pot_paths, subsets, gpu_rates, march)
        out["cpu_baseline"] = {
            "value": leg.pop("value"),
            "unit": "atoms/s",
            "cores": 1,
            "kind": "port",
            "sample": leg.pop("sample"),
            **leg,
        }
        out["speedup_vs_cpu"] = value / out["cpu_baseline"]["value"]
        if mixed_results is not None:
            out["mixed"]["oracle_errors"] = oracle_errors(mixed_results[:len(ref3)], ref3, ms[:len(ref3)])
        if serve_acc_results is not None:
            out["serving"]["oracle_errors"] = oracle_errors(serve_acc_results, ref4, serve_acc)
    print(json.dumps(out))
    model.close()
    if model_mixed is not None:
        model_mixed.close()


if __name__ == "__main__":
    main()
