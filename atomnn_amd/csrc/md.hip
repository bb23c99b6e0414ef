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
// energy), captured once and replayed; the host only waits when the state is read.  By default
// the neighbour list is rebuilt every step.  With a Verlet list (nnp_md_vlist: list radius
// rc_list > rc, rebuilt every `stride` steps, as the reference's MD driver rebuilds its
// verlet::List, torch/torch.cpp:269 and struc/verlet.cpp:51-106) a step that does not rebuild
// recomputes the displacement of every list entry from the current positions and the entry's
// lattice translation stored at the last build, and keeps the entries within rc, in list order,
// with the min-image flags of NeighborList::build recomputed (k_md_refresh); the rest of the pass
// is unchanged, so every step evaluates the potential on the list a fresh build would give.
// Positions are returned to the cell only at list builds.  Rows are sized with head room; a step
// whose neighbour count outgrows them, or whose atoms moved more than half the skin since the
// last build, is detected at the end of nnp_md_step, which then restores the state saved at its
// start and redoes the steps (with larger rows, or with a build every step).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "../../include/nnp_capi.h"
#include "engine.hpp"

namespace atomnn {

hipError_t launch_neighbors(const NeighborArgs& a, hipStream_t st);  // neigh.hip

constexpr double kAmuA2PerFs2 = 103.642696;  // eV per amu A^2 / fs^2
constexpr double kBoltzmann = 8.617333262e-5;  // eV / K

// math/special.hpp:46-47, mod(x, 0, 1)
__device__ __forceinline__ double mod01(double x) { return fmod(fmod(x, 1.0) + 1.0, 1.0); }

// The first half kick and the drift of a step in one launch: v += c_i F with c_i = dt / (2 m_i)
// (A/fs per eV/A), then x += v dt and, with wrap, back into the cell (Cell::returnToCell,
// struc/cell.cpp:84-94).
__global__ __launch_bounds__(256) void k_kick_drift(double* __restrict__ x, double* __restrict__ v,
                                                   const double* __restrict__ F, const double* __restrict__ c, double dt,
                                                   int n, const DevStruc* __restrict__ S, int wrap) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const double ci = c[i];
    const double vx = v[3 * i] + ci * F[3 * i], vy = v[3 * i + 1] + ci * F[3 * i + 1], vz = v[3 * i + 2] + ci * F[3 * i + 2];
    v[3 * i] = vx;
    v[3 * i + 1] = vy;
    v[3 * i + 2] = vz;
    double px = x[3 * i] + vx * dt, py = x[3 * i + 1] + vy * dt, pz = x[3 * i + 2] + vz * dt;
    const DevStruc& s = S[0];
    if (wrap && s.periodic) {
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

// Verlet-list step (one wavefront per centre i): for each list entry (i -> j, translation T),
// r = (x_i - x_j) - T from the current positions, kept when |r|^2 <= rc^2 (self images only
// above 1e-6, neighbor.cpp:113-121), written in list order into the pass's rows; then the
// min-image flag of NeighborList::build (neighbor.cpp:131-141): the first strictly closest kept
// entry of each neighbour, when closer than ||R||_F.  At a list build (built = 1) the entries'
// displacements are the build's own (so the rows equal a fresh build's bit for bit) and T and the
// build positions are stored.  Also: clears the pass counters the neighbour launch would clear;
// flags an atom that moved more than half the skin since the build (maxseen[3]); carries the
// build's longest list row into maxseen[4].
__global__ __launch_bounds__(kWave) void k_md_refresh(const DevStruc* __restrict__ S, const double* __restrict__ x,
                                                     const int* __restrict__ lcnt, const int* __restrict__ lj,
                                                     const int* __restrict__ lmeta, const double4* __restrict__ lr,
                                                     int lmaxp, double4* __restrict__ T, double* __restrict__ xb,
                                                     int built, double rc2, double half_skin2,
                                                     int* __restrict__ cnt, int* __restrict__ nbr_j,
                                                     int* __restrict__ nbr_meta, double4* __restrict__ nbr_r, int maxp,
                                                     int* __restrict__ maxcnt, const int* __restrict__ list_maxcnt,
                                                     int* __restrict__ maxseen) {
    extern __shared__ __align__(8) unsigned char md_lds[];
    double* dk = (double*)md_lds;          // [maxp] kept distances
    int* jk = (int*)(dk + maxp);           // [maxp] kept neighbour indices
    int* mk = jk + maxp;                   // [maxp] kept entries' metadata without the flag
    const int i = blockIdx.x, lane = threadIdx.x;
    const DevStruc& s = S[0];
    if (i == 0 && lane == 0) {
        maxcnt[0] = 0;
        maxcnt[1] = 0;
        if (built) atomicMax(&maxseen[4], list_maxcnt[0]);
    }
    const double xi = x[3 * i], yi = x[3 * i + 1], zi = x[3 * i + 2];
    if (lane == 0) {
        if (built) {
            xb[3 * i] = xi;
            xb[3 * i + 1] = yi;
            xb[3 * i + 2] = zi;
        } else {
            const double ux = xi - xb[3 * i], uy = yi - xb[3 * i + 1], uz = zi - xb[3 * i + 2];
            if (ux * ux + uy * uy + uz * uz >= half_skin2) atomicMax(&maxseen[3], 1);
        }
    }
    const int P = min(lcnt[i], lmaxp);
    const int64_t lrow = (int64_t)i * lmaxp, row = (int64_t)i * maxp;
    int kept = 0;
    for (int s0 = 0; s0 < P; s0 += kWave) {
        const int sidx = s0 + lane;
        bool keep = false;
        int j = 0, meta = 0;
        double rx = 0, ry = 0, rz = 0, d2 = 0;
        if (sidx < P) {
            const int64_t e = lrow + sidx;
            j = lj[e];
            meta = lmeta[e];
            if (built) {
                const double4 r = lr[e];
                rx = r.x;
                ry = r.y;
                rz = r.z;
                T[e] = make_double4((xi - x[3 * j]) - rx, (yi - x[3 * j + 1]) - ry, (zi - x[3 * j + 2]) - rz, 0.0);
            } else {
                const double4 t = T[e];
                rx = (xi - x[3 * j]) - t.x;
                ry = (yi - x[3 * j + 1]) - t.y;
                rz = (zi - x[3 * j + 2]) - t.z;
            }
            d2 = rx * rx + ry * ry + rz * rz;
            keep = (j != i) ? (d2 <= rc2) : (1e-6 < d2 && d2 <= rc2);
        }
        const unsigned long long m = __ballot(keep);
        const int at = kept + __popcll(m & ((1ull << lane) - 1ull));
        if (keep && at < maxp) {
            const double d = sqrt(d2);
            nbr_j[row + at] = j;
            nbr_r[row + at] = make_double4(rx, ry, rz, d);
            dk[at] = d;
            jk[at] = j;
            mk[at] = meta & ~1;
        }
        kept += __popcll(m);
    }
    __syncthreads();
    // min flags over the kept entries (runs of one neighbour are contiguous)
    const int K = min(kept, maxp);
    for (int k = lane; k < K; k += kWave) {
        const int j = jk[k];
        const double d = dk[k];
        bool mn = j != i && d < s.fro;
        for (int q = k - 1; mn && q >= 0 && jk[q] == j; --q) mn = dk[q] > d;
        for (int q = k + 1; mn && q < K && jk[q] == j; ++q) mn = dk[q] >= d;
        nbr_meta[row + k] = mk[k] | (mn ? 1 : 0);
    }
    if (lane == 0) cnt[i] = kept;
}

struct Md {
    std::unique_ptr<Engine> eng;
    uint32_t flags = 0;
    double dt = 1.0;
    int n = 0;
    std::vector<double> mass;  // per atom, amu
    DevMem d_vel, d_half, d_mass, d_ke, d_maxseen, d_save, d_bsum, d_done;
    // Verlet list (nnp_md_vlist): a workspace whose cutoff is the list radius builds the list
    // every `stride` steps; d_T holds each entry's lattice translation, d_xb the build positions
    int stride = 1;
    double rc_list = 0.0;
    std::unique_ptr<Engine> list;
    DevMem d_T, d_xb;
    std::vector<int32_t> h_type;
    std::vector<double> h_cell;
    // step kinds: 0 = a full pass with its own neighbour build (no list), 1 = list build,
    // 2 = list reuse.  One captured graph per kind (after one eager run that sizes every
    // buffer), and one graph of a block of steps: kStepsPerGraph kind-0 steps, or one list
    // period (a build and stride - 1 reuses) when the period is at most kMaxBlock steps
    static constexpr int kStepsPerGraph = 8;
    static constexpr int kMaxBlock = 64;
    hipGraph_t graph[3] = {nullptr, nullptr, nullptr};
    hipGraphExec_t exec[3] = {nullptr, nullptr, nullptr};
    bool eager_done[3] = {false, false, false};
    hipGraph_t graphb = nullptr;
    hipGraphExec_t execb = nullptr;
    int64_t steps = 0;

    ~Md() { drop(); }
    void drop() {
        for (int k = 0; k < 3; ++k) {
            if (exec[k]) (void)hipGraphExecDestroy(exec[k]);
            if (graph[k]) (void)hipGraphDestroy(graph[k]);
            exec[k] = nullptr;
            graph[k] = nullptr;
            eager_done[k] = false;
        }
        if (execb) (void)hipGraphExecDestroy(execb);
        if (graphb) (void)hipGraphDestroy(graphb);
        execb = nullptr;
        graphb = nullptr;
    }
    double* pos() { return (double*)eng->d_pos.p; }
    double* force() { return (double*)eng->d_F.p; }

    // the list's rows sized for the current structure with head room (atoms move between builds)
    void size_list() {
        Engine& L = *list;
        L.build_neighbors(flags & NNP_ALL_IMAGES);
        const int mc = L.max_entries;
        const int want = (mc + mc / 4 + 16 + 31) / 32 * 32;
        if (want > L.maxp) {
            L.maxp = want;
            L.build_neighbors(flags & NNP_ALL_IMAGES);
        }
        d_T.reserve((size_t)n * L.maxp * sizeof(double4));
        d_xb.reserve((size_t)n * 24);
    }

    void refresh(int built) {
        Engine& E = *eng;
        Engine& L = *list;
        const size_t lds = (size_t)E.maxp * 16;
        const double half = 0.5 * (rc_list - E.host.rc);
        hipLaunchKernelGGL(k_md_refresh, dim3(n), dim3(kWave), lds, E.stream, (const DevStruc*)E.d_strucs.p,
                           (const double*)E.d_pos.p, (const int*)L.d_cnt.p, (const int*)L.d_nbrj.p,
                           (const int*)L.d_meta.p, (const double4*)L.d_nbrr.p, L.maxp, (double4*)d_T.p,
                           (double*)d_xb.p, built, E.host.rc * E.host.rc, half * half * (1.0 - 1e-9),
                           (int*)E.d_cnt.p, (int*)E.d_nbrj.p, (int*)E.d_meta.p, (double4*)E.d_nbrr.p, E.maxp,
                           (int*)E.d_maxcnt.p, (const int*)L.d_maxcnt.p, (int*)d_maxseen.p);
        HIPCHK(hipGetLastError());
    }

    void enqueue_step(int kind) {
        Engine& E = *eng;
        const int nb = (n + 255) / 256;
        hipLaunchKernelGGL(k_kick_drift, dim3(nb), dim3(256), 0, E.stream, pos(), (double*)d_vel.p, force(),
                           (const double*)d_half.p, dt, n, (const DevStruc*)E.d_strucs.p, kind == 2 ? 0 : 1);
        HIPCHK(hipGetLastError());
        if (kind == 0) {
            E.enqueue_pass(flags);
        } else {
            Engine& L = *list;
            if (kind == 1) {
                // the list at the (returned-to-cell) positions: the list workspace's neighbour
                // launch at rc_list, on this stream
                HIPCHK(hipMemcpyAsync(L.d_pos.p, E.d_pos.p, (size_t)n * 24, hipMemcpyDeviceToDevice, E.stream));
                HIPCHK(launch_neighbors(L.neighbor_args(flags & NNP_ALL_IMAGES), E.stream));
            }
            refresh(kind == 1 ? 1 : 0);
            E.enqueue_pass(flags, true);
        }
        hipLaunchKernelGGL(k_kick_kinetic, dim3(nb), dim3(256), 0, E.stream, (double*)d_vel.p, force(),
                           (const double*)d_half.p, (const double*)d_mass.p, n, (double*)d_bsum.p,
                           (unsigned*)d_done.p, (double*)d_ke.p, (const int*)E.d_maxcnt.p, (int*)d_maxseen.p);
        HIPCHK(hipGetLastError());
    }

    template <typename F>
    void capture(hipGraph_t& g, hipGraphExec_t& ex, F&& body) {
        Engine& E = *eng;
        HIPCHK(hipStreamBeginCapture(E.stream, hipStreamCaptureModeThreadLocal));
        try {
            body();
        } catch (...) {
            hipGraph_t gg = nullptr;
            (void)hipStreamEndCapture(E.stream, &gg);
            if (gg) (void)hipGraphDestroy(gg);
            throw;
        }
        HIPCHK(hipStreamEndCapture(E.stream, &g));
        HIPCHK(hipGraphInstantiate(&ex, g, nullptr, nullptr, 0));
    }

    // one step of a kind: the first runs eagerly (it sizes every buffer), the second is captured,
    // later steps replay the graph
    void step_once(int kind) {
        Engine& E = *eng;
        if (!eager_done[kind]) {
            enqueue_step(kind);
            eager_done[kind] = true;
            return;
        }
        if (!exec[kind]) capture(graph[kind], exec[kind], [&]() { enqueue_step(kind); });
        HIPCHK(hipGraphLaunch(exec[kind], E.stream));
    }

    // a block of steps in one graph (each kind it holds has run eagerly)
    void step_block(int len, bool use_list) {
        Engine& E = *eng;
        if (!execb)
            capture(graphb, execb, [&]() {
                for (int q = 0; q < len; ++q) enqueue_step(use_list ? (q == 0 ? 1 : 2) : 0);
            });
        HIPCHK(hipGraphLaunch(execb, E.stream));
    }

    void run(int nsteps) {
        Engine& E = *eng;
        const size_t vb = (size_t)n * 24;
        // save x, v, F so an overflowing run can be redone
        d_save.reserve(3 * vb);
        HIPCHK(hipMemcpyAsync(d_save.p, pos(), vb, hipMemcpyDeviceToDevice, E.stream));
        HIPCHK(hipMemcpyAsync((char*)d_save.p + vb, d_vel.p, vb, hipMemcpyDeviceToDevice, E.stream));
        HIPCHK(hipMemcpyAsync((char*)d_save.p + 2 * vb, force(), vb, hipMemcpyDeviceToDevice, E.stream));
        HIPCHK(hipMemsetAsync(d_maxseen.p, 0, 32, E.stream));
        bool use_list = list && stride > 1;
        for (int attempt = 0; attempt < 5; ++attempt) {
            // list steps: a build at the first step of the call and every stride steps after
            const int blen = use_list ? stride : kStepsPerGraph;
            const bool blocks = !use_list || stride <= kMaxBlock;
            for (int k = 0; k < nsteps;) {
                const int kind = use_list ? (k % stride == 0 ? 1 : 2) : 0;
                const bool ready = use_list ? (eager_done[1] && eager_done[2]) : eager_done[0];
                if (blocks && ready && k % blen == 0 && nsteps - k >= blen) {
                    step_block(blen, use_list);
                    k += blen;
                    steps += blen;
                } else {
                    step_once(kind);
                    ++k;
                    ++steps;
                }
            }
            int c[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            HIPCHK(hipMemcpyAsync(c, d_maxseen.p, 32, hipMemcpyDeviceToHost, E.stream));
            HIPCHK(hipStreamSynchronize(E.stream));
            const int seen = c[0];
            const bool rows = seen > E.lmaxp;
            const bool list_rows = use_list && c[4] > list->maxp;
            const bool skin = use_list && !list_rows && c[3] != 0;
            // reactions the assembly's gather missed (or receiver rows too short): the steps
            // are redone with receiver rows
            const bool react = !rows && !list_rows && !skin && E.reactions_redo(c);
            if (!rows && !react && !skin && !list_rows) return;
            // restore and redo
            steps -= nsteps;
            HIPCHK(hipMemcpyAsync(pos(), d_save.p, vb, hipMemcpyDeviceToDevice, E.stream));
            HIPCHK(hipMemcpyAsync(d_vel.p, (char*)d_save.p + vb, vb, hipMemcpyDeviceToDevice, E.stream));
            HIPCHK(hipMemcpyAsync(force(), (char*)d_save.p + 2 * vb, vb, hipMemcpyDeviceToDevice, E.stream));
            HIPCHK(hipMemsetAsync(d_maxseen.p, 0, 32, E.stream));
            if (rows) {
                if (seen > E.maxp) {
                    E.maxp = (seen + seen / 4 + 16 + 31) / 32 * 32;
                    E.build_neighbors(flags & NNP_ALL_IMAGES);
                }
                E.lmaxp = std::min(E.maxp, (seen + seen / 4 + 8 + 7) / 8 * 8);
            }
            if (list_rows) {
                list->maxp = (c[4] + c[4] / 4 + 16 + 31) / 32 * 32;
                HIPCHK(hipMemcpyAsync(list->d_pos.p, pos(), vb, hipMemcpyDeviceToDevice, E.stream));
                size_list();
            }
            if (skin) use_list = false;  // an atom outran half the skin: this call builds every step
            drop();
        }
        throw std::runtime_error("nnp_md_step: neighbour rows or reaction rows did not converge");
    }

    // Verlet list: radius rc_list (>= rc), rebuilt every `stride` steps (1: a fresh list every step)
    void set_list(double rcl, int str) {
        Engine& E = *eng;
        HIPCHK(hipStreamSynchronize(E.stream));
        drop();
        stride = str;
        rc_list = rcl;
        if (stride <= 1) {
            list.reset();
            return;
        }
        list.reset(new Engine(E, 0));
        list->host.rc = rc_list;
        nnp_structure s;
        s.natoms = n;
        s.type = h_type.data();
        std::vector<double> x((size_t)n * 3);
        HIPCHK(hipMemcpy(x.data(), pos(), x.size() * 8, hipMemcpyDeviceToHost));
        s.pos = x.data();
        s.cell = h_cell.empty() ? nullptr : h_cell.data();
        list->upload(std::vector<nnp_structure>(1, s));
        size_list();
        HIPCHK(hipStreamSynchronize(list->stream));
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
        md->d_maxseen.reserve(32);
        md->h_type.assign(s->type, s->type + n);
        if (s->cell) md->h_cell.assign(s->cell, s->cell + 9);
        md->d_bsum.reserve((size_t)((n + 255) / 256) * 8);
        md->d_done.reserve(4);
        HIPCHK(hipMemsetAsync(md->d_done.p, 0, 4, E.stream));  // block counter; each step leaves it at 0
        HIPCHK(hipMemcpyAsync(md->d_half.p, half.data(), (size_t)n * 8, hipMemcpyHostToDevice, E.stream));
        HIPCHK(hipMemcpyAsync(md->d_mass.p, md->mass.data(), (size_t)n * 8, hipMemcpyHostToDevice, E.stream));
        if (vel) HIPCHK(hipMemcpyAsync(md->d_vel.p, vel, (size_t)n * 24, hipMemcpyHostToDevice, E.stream));
        else HIPCHK(hipMemsetAsync(md->d_vel.p, 0, (size_t)n * 24, E.stream));
        HIPCHK(hipMemsetAsync(md->d_maxseen.p, 0, 32, E.stream));
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

int nnp_md_vlist(nnp_md* md, double rc_list, int stride) {
    return md_guard([&]() {
        if (!md || stride < 1) throw std::invalid_argument("nnp_md_vlist: invalid argument");
        if (stride > 1 && !(rc_list >= md->md->eng->host.rc))
            throw std::invalid_argument("nnp_md_vlist: the list radius must be at least the potential's cutoff");
        md->md->set_list(rc_list, stride);
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
