// Molecular dynamics on the device-resident structure: velocity Verlet with the potential's
// energy and forces, the step the reference's MD driver takes (Verlet::compute,
// torch/integrator.cpp:170-188: half kick, drift, return to cell, Engine::compute
// (torch/engine.cpp:107-119), half kick, kinetic energy and temperature).
//
// Units: METAL (Angstrom, eV, amu, fs).  a = F / m in A/fs^2 is F[eV/A] / m[amu] times
// 1 / 103.642696 (1 amu A^2 / fs^2 = 103.642696 eV); kinetic energy in eV; temperature from
// T = (2/3) KE / (N kB), kB = 8.617333262e-5 eV/K.
//
// Every step is one graph of launches on the engine's stream (half kick, drift and wrap, the
// full pass of neighbour list / descriptors / networks / forces / energies, half kick, kinetic
// energy), captured once and replayed; the host only waits when the state is read.  The
// neighbour list is rebuilt every step (it costs less than the skin bookkeeping of a reused
// list).  Rows are sized with head room; a step whose neighbour count outgrows them is detected
// at the end of nnp_md_step, which then restores the state saved at its start and redoes the
// steps with larger rows.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "../../include/nnp_capi.h"
#include "engine.hpp"

namespace atomnn {

constexpr double kAmuA2PerFs2 = 103.642696;  // eV per amu A^2 / fs^2
constexpr double kBoltzmann = 8.617333262e-5;  // eV / K

// math/special.hpp:46-47, mod(x, 0, 1)
__device__ __forceinline__ double mod01(double x) { return fmod(fmod(x, 1.0) + 1.0, 1.0); }

// The first half kick and the drift of a step in one launch: v += c_i F with c_i = dt / (2 m_i)
// (A/fs per eV/A), then x += v dt and back into the cell (Cell::returnToCell,
// struc/cell.cpp:84-94).
__global__ __launch_bounds__(256) void k_kick_drift(double* __restrict__ x, double* __restrict__ v,
                                                   const double* __restrict__ F, const double* __restrict__ c, double dt,
                                                   int n, const DevStruc* __restrict__ S) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const double ci = c[i];
    const double vx = v[3 * i] + ci * F[3 * i], vy = v[3 * i + 1] + ci * F[3 * i + 1], vz = v[3 * i + 2] + ci * F[3 * i + 2];
    v[3 * i] = vx;
    v[3 * i + 1] = vy;
    v[3 * i + 2] = vz;
    double px = x[3 * i] + vx * dt, py = x[3 * i + 1] + vy * dt, pz = x[3 * i + 2] + vz * dt;
    const DevStruc& s = S[0];
    if (s.periodic) {
        const double f0 = mod01(s.RInv[0] * px + s.RInv[1] * py + s.RInv[2] * pz);
        const double f1 = mod01(s.RInv[3] * px + s.RInv[4] * py + s.RInv[5] * pz);
        const double f2 = mod01(s.RInv[6] * px + s.RInv[7] * py + s.RInv[8] * pz);
        px = s.R[0] * f0 + s.R[1] * f1 + s.R[2] * f2;
        py = s.R[3] * f0 + s.R[4] * f1 + s.R[5] * f2;
        pz = s.R[6] * f0 + s.R[7] * f1 + s.R[8] * f2;
    }
    x[3 * i] = px;
    x[3 * i + 1] = py;
    x[3 * i + 2] = pz;
}

// The second half kick and the kinetic energy in one launch: each block kicks its 256 atoms and
// sums m v^2 over them (fixed tree), the last block to finish adds the block sums in block order
// (ke = 0.5 of that), folds the pass's neighbour maximum into the sticky maximum and resets the
// block counter for the next step (no memset node in the captured step).
__global__ __launch_bounds__(256) void k_kick_kinetic(double* __restrict__ v, const double* __restrict__ F,
                                                     const double* __restrict__ c, const double* __restrict__ m, int n,
                                                     double* __restrict__ bsum, unsigned* __restrict__ done,
                                                     double* __restrict__ ke, const int* __restrict__ maxcnt,
                                                     int* __restrict__ maxseen) {
    __shared__ double part[256];
    __shared__ bool last;
    const int i = blockIdx.x * 256 + threadIdx.x;
    double s = 0.0;
    if (i < n) {
        const double ci = c[i];
        const double vx = v[3 * i] + ci * F[3 * i], vy = v[3 * i + 1] + ci * F[3 * i + 1],
                     vz = v[3 * i + 2] + ci * F[3 * i + 2];
        v[3 * i] = vx;
        v[3 * i + 1] = vy;
        v[3 * i + 2] = vz;
        s = m[i] * (vx * vx + vy * vy + vz * vz);
    }
    part[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) part[threadIdx.x] += part[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        bsum[blockIdx.x] = part[0];
        __threadfence();
        last = atomicAdd(done, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    double t = 0.0;
    for (int b = threadIdx.x; b < (int)gridDim.x; b += 256) t += __hip_atomic_load(&bsum[b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    part[threadIdx.x] = t;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) part[threadIdx.x] += part[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        *ke = 0.5 * part[0];
        // neighbour maximum, missed reactions, longest receiver row: maxima since the last reset
        for (int q = 0; q < 3; ++q)
            if (maxcnt[q] > maxseen[q]) maxseen[q] = maxcnt[q];
        *done = 0u;
    }
}

// ke = 0.5 sum_i m_i v_i^2 (amu A^2/fs^2), one block, fixed-order tree sum; also tracks the
// largest neighbour count seen since the last reset
__global__ __launch_bounds__(256) void k_kinetic(const double* __restrict__ v, const double* __restrict__ m, int n,
                                                double* __restrict__ ke, const int* __restrict__ maxcnt,
                                                int* __restrict__ maxseen) {
    __shared__ double part[256];
    double s = 0;
    for (int i = threadIdx.x; i < n; i += 256)
        s += m[i] * (v[3 * i] * v[3 * i] + v[3 * i + 1] * v[3 * i + 1] + v[3 * i + 2] * v[3 * i + 2]);
    part[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) part[threadIdx.x] += part[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        *ke = 0.5 * part[0];
        for (int q = 0; q < 3; ++q)
            if (maxcnt[q] > maxseen[q]) maxseen[q] = maxcnt[q];
    }
}

struct Md {
    std::unique_ptr<Engine> eng;
    uint32_t flags = 0;
    double dt = 1.0;
    int n = 0;
    std::vector<double> mass;  // per atom, amu
    DevMem d_vel, d_half, d_mass, d_ke, d_maxseen, d_save, d_bsum, d_done;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // kStepsPerGraph consecutive steps in one graph: one launch instead of eight
    static constexpr int kStepsPerGraph = 8;
    hipGraph_t graph8 = nullptr;
    hipGraphExec_t exec8 = nullptr;
    int graph_lmaxp = 0, graph_maxp = 0;
    bool eager_done = false;
    int64_t steps = 0;

    ~Md() { drop(); }
    void drop() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (exec8) (void)hipGraphExecDestroy(exec8);
        if (graph8) (void)hipGraphDestroy(graph8);
        exec = nullptr;
        graph = nullptr;
        exec8 = nullptr;
        graph8 = nullptr;
    }
    double* pos() { return (double*)eng->d_pos.p; }
    double* force() { return (double*)eng->d_F.p; }

    void enqueue_step() {
        Engine& E = *eng;
        const int nb = (n + 255) / 256;
        hipLaunchKernelGGL(k_kick_drift, dim3(nb), dim3(256), 0, E.stream, pos(), (double*)d_vel.p, force(),
                           (const double*)d_half.p, dt, n, (const DevStruc*)E.d_strucs.p);
        HIPCHK(hipGetLastError());
        E.enqueue_pass(flags);
        hipLaunchKernelGGL(k_kick_kinetic, dim3(nb), dim3(256), 0, E.stream, (double*)d_vel.p, force(),
                           (const double*)d_half.p, (const double*)d_mass.p, n, (double*)d_bsum.p,
                           (unsigned*)d_done.p, (double*)d_ke.p, (const int*)E.d_maxcnt.p, (int*)d_maxseen.p);
        HIPCHK(hipGetLastError());
    }

    // one step: the first runs eagerly (it sizes every buffer), the second is captured, later
    // steps replay the graph
    void step_once() {
        Engine& E = *eng;
        if (exec && graph_lmaxp == E.lmaxp && graph_maxp == E.maxp) {
            HIPCHK(hipGraphLaunch(exec, E.stream));
            return;
        }
        if (!eager_done || graph_lmaxp != E.lmaxp || graph_maxp != E.maxp) {
            drop();
            graph_lmaxp = E.lmaxp;
            graph_maxp = E.maxp;
            enqueue_step();
            eager_done = true;
            return;
        }
        HIPCHK(hipStreamBeginCapture(E.stream, hipStreamCaptureModeThreadLocal));
        try {
            enqueue_step();
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(E.stream, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        HIPCHK(hipStreamEndCapture(E.stream, &graph));
        HIPCHK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        HIPCHK(hipGraphLaunch(exec, E.stream));
    }

    // kStepsPerGraph steps (after the eager step has sized the buffers for these rows)
    void step_block() {
        Engine& E = *eng;
        if (!exec8) {
            HIPCHK(hipStreamBeginCapture(E.stream, hipStreamCaptureModeThreadLocal));
            try {
                for (int q = 0; q < kStepsPerGraph; ++q) enqueue_step();
            } catch (...) {
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture(E.stream, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            HIPCHK(hipStreamEndCapture(E.stream, &graph8));
            HIPCHK(hipGraphInstantiate(&exec8, graph8, nullptr, nullptr, 0));
        }
        HIPCHK(hipGraphLaunch(exec8, E.stream));
    }

    void run(int nsteps) {
        Engine& E = *eng;
        const size_t vb = (size_t)n * 24;
        // save x, v, F so an overflowing run can be redone
        d_save.reserve(3 * vb);
        HIPCHK(hipMemcpyAsync(d_save.p, pos(), vb, hipMemcpyDeviceToDevice, E.stream));
        HIPCHK(hipMemcpyAsync((char*)d_save.p + vb, d_vel.p, vb, hipMemcpyDeviceToDevice, E.stream));
        HIPCHK(hipMemcpyAsync((char*)d_save.p + 2 * vb, force(), vb, hipMemcpyDeviceToDevice, E.stream));
        HIPCHK(hipMemsetAsync(d_maxseen.p, 0, 16, E.stream));
        for (int attempt = 0; attempt < 4; ++attempt) {
            for (int k = 0; k < nsteps;) {
                const bool ready = eager_done && graph_lmaxp == E.lmaxp && graph_maxp == E.maxp;
                if (ready && nsteps - k >= kStepsPerGraph) {
                    step_block();
                    k += kStepsPerGraph;
                    steps += kStepsPerGraph;
                } else {
                    step_once();
                    ++k;
                    ++steps;
                }
            }
            int c[4] = {0, 0, 0, 0};
            HIPCHK(hipMemcpyAsync(c, d_maxseen.p, 16, hipMemcpyDeviceToHost, E.stream));
            HIPCHK(hipStreamSynchronize(E.stream));
            const int seen = c[0];
            const bool rows = seen > E.lmaxp;
            // reactions the assembly's gather missed (or receiver rows too short): the steps
            // are redone with receiver rows
            const bool react = !rows && E.reactions_redo(c);
            if (!rows && !react) return;
            // restore and redo
            steps -= nsteps;
            HIPCHK(hipMemcpyAsync(pos(), d_save.p, vb, hipMemcpyDeviceToDevice, E.stream));
            HIPCHK(hipMemcpyAsync(d_vel.p, (char*)d_save.p + vb, vb, hipMemcpyDeviceToDevice, E.stream));
            HIPCHK(hipMemcpyAsync(force(), (char*)d_save.p + 2 * vb, vb, hipMemcpyDeviceToDevice, E.stream));
            HIPCHK(hipMemsetAsync(d_maxseen.p, 0, 16, E.stream));
            if (rows) {
                if (seen > E.maxp) {
                    E.maxp = (seen + seen / 4 + 16 + 31) / 32 * 32;
                    E.build_neighbors(flags & NNP_ALL_IMAGES);
                }
                E.lmaxp = std::min(E.maxp, (seen + seen / 4 + 8 + 7) / 8 * 8);
            }
            drop();
            eager_done = false;
        }
        throw std::runtime_error("nnp_md_step: neighbour rows or reaction rows did not converge");
    }
};

}  // namespace atomnn

using namespace atomnn;

struct nnp_md {
    Md* md;
};

namespace atomnn {
Engine& engine_of(nnp_model* m);
}

template <typename F>
static int md_guard(F&& f) {
    try {
        set_error("");
        return f();
    } catch (const std::bad_alloc&) {
        set_error("out of memory");
        return NNP_ENOMEM;
    } catch (const hip_failure& e) {
        set_error(e.what());
        return NNP_EDEVICE;
    } catch (const std::exception& e) {
        set_error(e.what());
        return NNP_EINVAL;
    }
}

extern "C" {

int nnp_md_create(nnp_model* m, const nnp_structure* s, const double* masses, const double* vel, double dt,
                  uint32_t flags, nnp_md** out) {
    return md_guard([&]() {
        if (!m || !s || !masses || !out || s->natoms <= 0 || !(dt > 0))
            throw std::invalid_argument("nnp_md_create: invalid argument");
        Engine& owner = engine_of(m);
        for (int t = 0; t < owner.host.T; ++t)
            if (!(masses[t] > 0)) throw std::invalid_argument("nnp_md_create: masses must be positive");
        auto md = std::make_unique<Md>();
        md->eng.reset(new Engine(owner, 0));
        md->flags = (flags | NNP_FORCE) & ~(uint32_t)NNP_ATOM_ENERGY;
        md->dt = dt;
        md->n = s->natoms;
        Engine& E = *md->eng;
        E.resident_load(std::vector<nnp_structure>(1, *s), md->flags);
        // MD moves atoms: rows and LDS staging with extra head room
        const int mc = E.max_entries;
        const int want = (mc + mc / 4 + 16 + 31) / 32 * 32;
        if (want > E.maxp) {
            E.maxp = want;
            E.build_neighbors(md->flags & NNP_ALL_IMAGES);
        }
        E.lmaxp = std::min(E.maxp, (mc + mc / 4 + 8 + 7) / 8 * 8);
        const int n = md->n;
        md->mass.resize(n);
        std::vector<double> half(n);
        for (int i = 0; i < n; ++i) {
            const int t = s->type[i];
            md->mass[i] = masses[t];
            half[i] = 0.5 * dt / (masses[t] * kAmuA2PerFs2);
        }
        md->d_vel.reserve((size_t)n * 24);
        md->d_half.reserve((size_t)n * 8);
        md->d_mass.reserve((size_t)n * 8);
        md->d_ke.reserve(8);
        md->d_maxseen.reserve(16);
        md->d_bsum.reserve((size_t)((n + 255) / 256) * 8);
        md->d_done.reserve(4);
        HIPCHK(hipMemsetAsync(md->d_done.p, 0, 4, E.stream));  // block counter; each step leaves it at 0
        HIPCHK(hipMemcpyAsync(md->d_half.p, half.data(), (size_t)n * 8, hipMemcpyHostToDevice, E.stream));
        HIPCHK(hipMemcpyAsync(md->d_mass.p, md->mass.data(), (size_t)n * 8, hipMemcpyHostToDevice, E.stream));
        if (vel) HIPCHK(hipMemcpyAsync(md->d_vel.p, vel, (size_t)n * 24, hipMemcpyHostToDevice, E.stream));
        else HIPCHK(hipMemsetAsync(md->d_vel.p, 0, (size_t)n * 24, E.stream));
        HIPCHK(hipMemsetAsync(md->d_maxseen.p, 0, 16, E.stream));
        // initial forces and energies at the starting positions
        E.enqueue_pass(md->flags);
        hipLaunchKernelGGL(k_kinetic, dim3(1), dim3(256), 0, E.stream, (const double*)md->d_vel.p,
                           (const double*)md->d_mass.p, n, (double*)md->d_ke.p, (const int*)E.d_maxcnt.p,
                           (int*)md->d_maxseen.p);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(E.stream));
        *out = new nnp_md{md.release()};
        return NNP_OK;
    });
}

int nnp_md_step(nnp_md* md, int nsteps) {
    return md_guard([&]() {
        if (!md || nsteps < 0) throw std::invalid_argument("nnp_md_step: invalid argument");
        if (nsteps > 0) md->md->run(nsteps);
        return NNP_OK;
    });
}

int nnp_md_state(nnp_md* h, double* pos, double* vel, double* force, double* energies) {
    return md_guard([&]() {
        if (!h) throw std::invalid_argument("nnp_md_state: invalid argument");
        Md& md = *h->md;
        Engine& E = *md.eng;
        const size_t vb = (size_t)md.n * 24;
        double e[2] = {0, 0};
        if (pos) HIPCHK(hipMemcpyAsync(pos, md.pos(), vb, hipMemcpyDeviceToHost, E.stream));
        if (vel) HIPCHK(hipMemcpyAsync(vel, md.d_vel.p, vb, hipMemcpyDeviceToHost, E.stream));
        if (force) HIPCHK(hipMemcpyAsync(force, md.force(), vb, hipMemcpyDeviceToHost, E.stream));
        HIPCHK(hipMemcpyAsync(&e[0], E.d_E.p, 8, hipMemcpyDeviceToHost, E.stream));
        HIPCHK(hipMemcpyAsync(&e[1], md.d_ke.p, 8, hipMemcpyDeviceToHost, E.stream));
        HIPCHK(hipStreamSynchronize(E.stream));
        if (energies) {
            energies[0] = e[0];                                       // potential energy (eV)
            energies[1] = e[1] * kAmuA2PerFs2;                        // kinetic energy (eV)
            energies[2] = energies[1] * (2.0 / 3.0) / (md.n * kBoltzmann);  // temperature (K)
            energies[3] = (double)md.steps;                           // steps taken
        }
        return NNP_OK;
    });
}

void nnp_md_destroy(nnp_md* h) {
    if (!h) return;
    delete h->md;
    delete h;
}

}  // extern "C"
