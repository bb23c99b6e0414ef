// Host orchestration of the energy/force pipeline and the C ABI (include/nnp_capi.h).
//
// Pipeline for a batch of structures (all kernels cover every atom of every structure):
//   1. neighbour list      k_zero, k_bin_atoms / k_scan_excl / k_fill_bins / k_neigh_cells
//                          (large cells), k_neigh_allpairs (small cells)
//   2. descriptors         k_symm<T, NE, NC>                 -> G
//   3. networks            k_net_f64 | k_net_mixed, all species' 16-atom tiles -> E_i, dE/dG
//   4. force terms         k_force<T, NE, NC>                 -> X (per neighbour entry)
//   5. force assembly      k_assemble (gathers each atom's reactions from its neighbours' rows) -> F (FP64)
//   6. energy              per structure, sum of E_i in atom order (FP64), as
//                          NNP::energy accumulates (nnp.cpp:394-400)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/nnp_capi.h"
#include "device.hpp"
#include "engine.hpp"
#include "model.hpp"

namespace atomnn {

// launchers (symm.hip, net.hip, neigh.hip kernels)
int grid_shape(int ne, int nc);
int shape_ne(int shape);  // exponents per centre of a grid shape
constexpr int kMomK = 35;  // monomials of degree <= 4: the moment table's row length
constexpr int kMomZ = 4;   // their highest degree
size_t symm_lds(int shape, bool fp64, bool ns, int lmaxp, int ldg);
size_t force_lds(int shape, bool fp64, bool ns, int lmaxp, int ldg);
size_t net_lds_bytes(const DevNet& n, bool fp64);
template <typename T>
hipError_t launch_symm(int shape, int ns, const DevPlan* plan, const double* rfe, const double* rfr, const int* atom_type,
                       const int* centres, int ncentres, const int* cnt, const int* meta, const double4* r, int maxp,
                       int lmaxp, T* G, int ldg, int use_mom, int ntypes, hipStream_t st, T* mom_out);
template <typename T>
hipError_t launch_force(int shape, int ns, const DevPlan* plan, const double* rfe, const double* rfr, const int* atom_type,
                        const int* centres, int ncentres, const int* cnt, const int* meta, const double4* r, int maxp,
                        int lmaxp, const T* dEdG, int ldg, T* X, int use_mom, int ntypes, hipStream_t st);
size_t force_mom_lds(int shape, bool fp64, int waves, int lmaxp, int ldg, int ntypes);
template <typename T>
hipError_t launch_force_mom(int shape, int waves, const DevPlan* plan, const double* rfe, const double* rfr,
                            const int* atom_type, const int* centres, int ncentres, const int* cnt, const int* meta,
                            const double4* r, int maxp, int lmaxp, const T* dEdG, int ldg, T* X, int ntypes,
                            hipStream_t st, const T* mom_in);
template <typename T>
hipError_t launch_assemble(const AssembleArgs& a, const T* X, const EnergyTail* tail, hipStream_t st);
hipError_t launch_net(bool fp64, const DevPlan* plan, const DevPlan& hplan, const void* W, const int* sp_off,
                      const int* spatoms, const void* G, int ldg, int vac, double* eatom, void* dEdG, int bf16_terms,
                      NetLaunch cfg, hipStream_t st);
int net_tiles(const int* sp_off, int ntypes);
bool net_wl_fits(const DevPlan& hplan);
hipError_t launch_neighbors(const NeighborArgs& a, hipStream_t st);
hipError_t launch_count_max(const int* cnt, int n, int* maxcnt, hipStream_t st);
size_t fused_lds(int shape, bool fp64, bool ns, int lmaxp, int ldg, int maxw, int dersum);
size_t mom_lds(int shape, bool fp64, int lmaxp, int ntypes, int bits);
template <typename T>
hipError_t launch_fused(int shape, int ns, int nblocks, const DevPlan* plan, const double* rfe, const double* rfr, const T* Wt,
                        const int* atom_type, const int* order, int ntot, int* work, const int* cnt, const int* meta,
                        const double4* r, int maxp, int lmaxp, int ldg, int maxw, int dersum, int vac, int do_force,
                        double* eatom, T* X, int use_mom, int ntypes, hipStream_t st);

// The end-of-pass reductions on their own (passes without forces): one wavefront per block.
__global__ __launch_bounds__(kWave) void k_energy(EnergyTail t) { energy_tail_block(t, blockIdx.x, threadIdx.x); }

// Results of a pass packed into one buffer for a single device-to-host copy: the four counters
// (16 bytes), then E (per structure), E_i and F at the given 8-byte-word offsets.
__global__ __launch_bounds__(256) void k_pack_out(const int64_t* __restrict__ cnt, const int64_t* __restrict__ E,
                                                  int ne, const int64_t* __restrict__ ea, int nea,
                                                  const int64_t* __restrict__ F, int nf, int64_t* __restrict__ out,
                                                  int o_e, int o_ea, int o_f) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t < 2) out[t] = cnt[t];
    if (t < ne) out[o_e + t] = E[t];
    if (t < nea) out[o_ea + t] = ea[t];
    if (t < nf) out[o_f + t] = F[t];
}

// ------------------------------------------------------------------ errors
static thread_local std::string g_err;

void set_error(const std::string& s) { g_err = s; }
const char* last_error() { return g_err.c_str(); }


// ------------------------------------------------------------------ scratch
// The workspace of nnp_compute_noalloc: one fixed array per device, held until the process ends.
// Host code never uses the array's name as an address (in host code it is a stand-in, not device
// memory): its device address is looked up once per device, with that device current, and kept.
constexpr size_t kScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices] = {};

// The scratch array's address on the current device (looked up on the first call per device).
static void* scratch_on_current_device(int& dev) {
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("device index out of range for the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) throw hip_failure("hipGetSymbolAddress returned no address for the scratch array");
        g_scratch_ptr[dev] = ptr;
    }
    return g_scratch_ptr[dev];
}

// ------------------------------------------------------------------ device buffers
// Failure injection for tests (option inject_alloc_failure, process-wide): when set to k > 0,
// the k-th device allocation after it fails as an exhausted device would (std::bad_alloc).
static std::atomic<int64_t> g_alloc_fail_in{0};

void DevMem::reserve(size_t bytes) {
    if (bytes <= cap) return;
    if (bound)
        throw unsupported_error("nnp_compute_noalloc: the structure needs more workspace than the library's " +
                                std::to_string(kScratchBytes >> 20) + " MiB scratch array holds");
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    size_t want = std::max<size_t>(bytes, 256);
    want = (want + 4095) & ~size_t(4095);
    if (g_alloc_fail_in.load() > 0 && g_alloc_fail_in.fetch_sub(1) == 1) throw std::bad_alloc();
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) {
        p = nullptr;
        throw std::bad_alloc();
    }
    cap = want;
}
void HostPinned::reserve(size_t bytes) {
    if (bytes <= cap) return;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
    size_t want = std::max<size_t>(bytes, 4096);
    want = (want + 65535) & ~size_t(65535);
    if (hipHostMalloc(&p, want, hipHostMallocDefault) != hipSuccess) {
        p = nullptr;
        throw std::bad_alloc();
    }
    cap = want;
}
void HostPinned::release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
}

void DevMem::release() {
    if (p && !bound) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    bound = false;
}

// ------------------------------------------------------------------ structures
// Cell::init (cell.cpp:45-70) and the image table of NeighborList::build (neighbor.cpp:67-95)
DevStruc make_struc(const nnp_structure& s, double rc, int atom0) {
    DevStruc d;
    std::memset(&d, 0, sizeof(d));
    d.atom0 = atom0;
    d.natoms = s.natoms;
    double R[3][3] = {{0}};
    if (s.cell)
        for (int c = 0; c < 3; ++c)
            for (int r = 0; r < 3; ++r) R[r][c] = s.cell[3 * c + r];
    double fro = 0;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) fro += R[r][c] * R[r][c];
    fro = std::sqrt(fro);
    d.fro = fro;
    d.periodic = fro > 0;
    d.nimg = 1;
    if (!d.periodic) return d;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) d.R[3 * r + c] = R[r][c];
    const double c00 = R[1][1] * R[2][2] - R[1][2] * R[2][1];
    const double c01 = R[1][2] * R[2][0] - R[1][0] * R[2][2];
    const double c02 = R[1][0] * R[2][1] - R[1][1] * R[2][0];
    const double det = R[0][0] * c00 + R[0][1] * c01 + R[0][2] * c02;
    if (det == 0) throw std::invalid_argument("Cell::init: singular lattice");
    const double id = 1.0 / det;
    double Ri[3][3];
    Ri[0][0] = c00 * id;
    Ri[1][0] = c01 * id;
    Ri[2][0] = c02 * id;
    Ri[0][1] = (R[0][2] * R[2][1] - R[0][1] * R[2][2]) * id;
    Ri[1][1] = (R[0][0] * R[2][2] - R[0][2] * R[2][0]) * id;
    Ri[2][1] = (R[0][1] * R[2][0] - R[0][0] * R[2][1]) * id;
    Ri[0][2] = (R[0][1] * R[1][2] - R[0][2] * R[1][1]) * id;
    Ri[1][2] = (R[0][2] * R[1][0] - R[0][0] * R[1][2]) * id;
    Ri[2][2] = (R[0][0] * R[1][1] - R[0][1] * R[1][0]) * id;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) d.RInv[3 * r + c] = Ri[r][c];
    auto nrm = [](const double* v) { return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); };
    double col[3][3];
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) col[c][r] = R[r][c];
    double dmax = nrm(col[0]);
    for (int c = 1; c < 3; ++c) dmax = std::min(dmax, nrm(col[c]));
    dmax *= 0.5;
    d.dMax2 = dmax * dmax;
    // lattice vectors sorted by descending norm, stable (neighbor.cpp:71-74)
    int perm[3] = {0, 1, 2};
    for (int i = 1; i < 3; ++i) {
        const int p = perm[i];
        int j = i;
        while (j > 0 && nrm(col[p]) > nrm(col[perm[j - 1]])) {
            perm[j] = perm[j - 1];
            --j;
        }
        perm[j] = p;
    }
    double Rv[3][3];
    for (int q = 0; q < 3; ++q)
        for (int r = 0; r < 3; ++r) Rv[q][r] = col[perm[q]][r];
    for (int q = 0; q < 3; ++q)
        for (int r = 0; r < 3; ++r) d.Rv[3 * q + r] = Rv[q][r];
    auto cross = [](const double* a, const double* b, double* o) {
        o[0] = a[1] * b[2] - a[2] * b[1];
        o[1] = a[2] * b[0] - a[0] * b[2];
        o[2] = a[0] * b[1] - a[1] * b[0];
    };
    double x01[3];
    cross(Rv[0], Rv[1], x01);
    d.shell[0] = (int)std::floor(2.0 * rc / nrm(Rv[0]));
    d.shell[1] = (int)std::floor(2.0 * rc / (nrm(x01) / nrm(Rv[0])));
    d.shell[2] = (int)std::floor(2.0 * rc / std::fabs((Rv[2][0] * x01[0] + Rv[2][1] * x01[1] + Rv[2][2] * x01[2]) /
                                                      (nrm(Rv[0]) * nrm(Rv[1]))));
    d.nimg = (2 * d.shell[0] + 1) * (2 * d.shell[1] + 1) * (2 * d.shell[2] + 1);
    {
        double x12[3], x20[3];
        cross(Rv[1], Rv[2], x12);
        cross(Rv[2], Rv[0], x20);
        const double V = Rv[0][0] * x12[0] + Rv[0][1] * x12[1] + Rv[0][2] * x12[2];
        const double* g[3] = {x12, x20, x01};
        for (int q = 0; q < 3; ++q) {
            for (int r = 0; r < 3; ++r) d.Gv[3 * q + r] = g[q][r] / V;
            d.reach[q] = rc * nrm(&d.Gv[3 * q]) * (1.0 + 1e-12) + 1e-9;
        }
    }
    // cell list when every shell is 0 and the perpendicular heights hold at least one rc bin
    if (d.shell[0] == 0 && d.shell[1] == 0 && d.shell[2] == 0 && s.natoms >= 256) {
        const double vol = std::fabs(det);
        bool ok = true;
        for (int a = 0; a < 3; ++a) {
            double xc[3];
            cross(col[(a + 1) % 3], col[(a + 2) % 3], xc);
            const double h = vol / nrm(xc);
            d.nbin[a] = (int)std::floor(h / rc);
            if (d.nbin[a] < 1) ok = false;
        }
        if (ok) d.mode = 1;
    }
    return d;
}

// ------------------------------------------------------------------ engine
Engine::Engine(HostModel&& hm) : host(std::move(hm)) {
    pb = build_plan(host);
    shape = grid_shape(pb.plan.ne_max, pb.plan.nc_max);
    HIPCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipDeviceGetAttribute(&num_cus, hipDeviceAttributeMultiprocessorCount, dev));
    build_adjoint_table();
    d_plan.reserve(sizeof(DevPlan));
    HIPCHK(hipMemcpy(d_plan.p, &pb.plan, sizeof(DevPlan), hipMemcpyHostToDevice));
    d_rfe.reserve(pb.rfun_eta.size() * 8);
    d_rfr.reserve(pb.rfun_rs.size() * 8);
    HIPCHK(hipMemcpy(d_rfe.p, pb.rfun_eta.data(), pb.rfun_eta.size() * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d_rfr.p, pb.rfun_rs.data(), pb.rfun_rs.size() * 8, hipMemcpyHostToDevice));
    d_w64.reserve(pb.weights.size() * 8);
    HIPCHK(hipMemcpy(d_w64.p, pb.weights.data(), pb.weights.size() * 8, hipMemcpyHostToDevice));
    std::vector<float> wf(pb.weights.begin(), pb.weights.end());
    d_w32.reserve(wf.size() * 4);
    HIPCHK(hipMemcpy(d_w32.p, wf.data(), wf.size() * 4, hipMemcpyHostToDevice));
}

// The w table of the moment path's adjoint (moments.hpp adjoint_moments) as a linear map of the
// dE/dG row: w[blk][e][p] = norm (1/2 for a same-species block) sum_c dE/dG[f(e, c)] C(zeta_c, p)
// lambda_c^p 2^-zeta_c over the block's angular terms c with f(e, c) defined and p <= zeta_c.
// Stored per item as kMaxC (index, coefficient) pairs, so the kernel reads one row of constants
// instead of walking the block's description.
void Engine::build_adjoint_table() {
    const DevPlan& P = pb.plan;
    const int ne = shape_ne(shape), T = P.T, nblk = T * (T + 1) / 2;
    const size_t items = (size_t)T * nblk * ne * (kMomZ + 1);
    std::vector<double> coef(items * kMaxC, 0.0);
    std::vector<int> index(items * kMaxC, 0);
    auto binom = [](int z, int p) { int r = 1; for (int k = 0; k < p; ++k) r = r * (z - k) / (k + 1); return r; };
    for (int t = 0; t < T; ++t) {
        if (!P.mom_ok[t]) continue;
        for (int blk = 0; blk < nblk; ++blk) {
            int bb = 0;
            while ((bb + 1) * (bb + 2) / 2 <= blk) ++bb;
            const int a = blk - bb * (bb + 1) / 2;
            const DevAngular& A = P.ang[t][blk];
            for (int e = 0; e < ne && e < A.ne; ++e)
                for (int p = 0; p <= kMomZ; ++p) {
                    const size_t base = ((((size_t)t * nblk + blk) * ne + e) * (kMomZ + 1) + p) * kMaxC;
                    int k = 0;
                    for (int c = 0; c < A.nc; ++c) {
                        const int f = A.fmap[e * kMaxC + c], z = A.zint[c];
                        if (f < 0 || p > z) continue;
                        double lp = 1.0;
                        for (int q = 0; q < p; ++q) lp *= A.lambda[c];
                        coef[base + k] = binom(z, p) * lp * (1.0 / (double)(1 << z)) * A.norm * (a == bb ? 0.5 : 1.0);
                        index[base + k] = A.goff + f;
                        ++k;
                    }
                }
        }
    }
    d_adjc.reserve(coef.size() * 8);
    d_adji.reserve(index.size() * 4);
    HIPCHK(hipMemcpy(d_adjc.p, coef.data(), coef.size() * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d_adji.p, index.data(), index.size() * 4, hipMemcpyHostToDevice));
    pb.plan.adj_coef = (const double*)d_adjc.p;
    pb.plan.adj_index = (const int*)d_adji.p;
    pb.plan.adj_ne = ne;
    // the angular descriptors from the moment products: G_f = norm 2^-zeta sum_p C(zeta, p)
    // lambda^p X_p for the function f of each (block, exponent, term)
    const size_t gitems = (size_t)T * nblk * ne * kMaxC;
    std::vector<double> gcoef(gitems * (kMomZ + 1), 0.0);
    std::vector<int> gindex(gitems, -1);
    for (int t = 0; t < T; ++t) {
        if (!P.mom_ok[t]) continue;
        for (int blk = 0; blk < nblk; ++blk) {
            const DevAngular& A = P.ang[t][blk];
            for (int e = 0; e < ne && e < A.ne; ++e)
                for (int c = 0; c < A.nc; ++c) {
                    const int f = A.fmap[e * kMaxC + c], z = A.zint[c];
                    if (f < 0) continue;
                    const size_t it = (((size_t)t * nblk + blk) * ne + e) * kMaxC + c;
                    gindex[it] = A.goff + f;
                    double lp = 1.0;
                    for (int p = 0; p <= z && p <= kMomZ; ++p) {
                        gcoef[it * (kMomZ + 1) + p] = binom(z, p) * lp * A.norm * (1.0 / (double)(1 << z));
                        lp *= A.lambda[c];
                    }
                }
        }
    }
    d_angc.reserve(gcoef.size() * 8);
    d_angi.reserve(gindex.size() * 4);
    HIPCHK(hipMemcpy(d_angc.p, gcoef.data(), gcoef.size() * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d_angi.p, gindex.data(), gindex.size() * 4, hipMemcpyHostToDevice));
    pb.plan.ang_coef = (const double*)d_angc.p;
    pb.plan.ang_index = (const int*)d_angi.p;
}

Engine::Engine(const Engine& o, int) : host(o.host), pb(o.pb), shape(o.shape) {
    owns_model = false;
    pipeline = o.pipeline;
    fused_blocks_per_cu = o.fused_blocks_per_cu;
    moments = o.moments;
    force_waves = o.force_waves;
    net_bf16_terms = o.net_bf16_terms;
    react_rows = o.react_rows;
    assemble_mode = o.react_rows;
    num_cus = o.num_cus;
    tune_cache = o.tune_cache;
    d_plan = o.d_plan;
    d_rfe = o.d_rfe;
    d_rfr = o.d_rfr;
    d_w64 = o.d_w64;
    d_w32 = o.d_w32;
    HIPCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
}

size_t Engine::workspace_bytes() const {
    size_t b = 0;
    for (const DevMem* m : {&d_up, &d_cnt, &d_nbrj, &d_meta, &d_nbrr, &d_maxcnt, &d_abin, &d_bcount, &d_bstart,
                            &d_bfill, &d_G, &d_dEdG, &d_X, &d_eatom, &d_F, &d_E, &d_sstart, &d_rcount,
                            &d_rlist, &d_out, &d_mom})
        b += m->cap;
    return b;
}

Engine::~Engine() {
    if (stream) (void)hipStreamSynchronize(stream);
    drop_graph();
    for (int st = 0; st < ST_COUNT; ++st)
        for (int q = 0; q < 2; ++q)
            if (ev[st][q]) (void)hipEventDestroy(ev[st][q]);
    d_E.release();
    d_sstart.release();
    d_work.release();
    if (owns_model)
        for (DevMem* m : {&d_plan, &d_rfe, &d_rfr, &d_w64, &d_w32, &d_adjc, &d_adji, &d_angc, &d_angi}) m->release();
    for (DevMem* m : {&d_up, &d_cnt, &d_nbrj, &d_meta, &d_nbrr, &d_maxcnt, &d_abin, &d_bcount, &d_bstart,
                      &d_bfill, &d_G, &d_dEdG, &d_X, &d_eatom, &d_F, &d_rcount, &d_rlist, &d_mom})
        m->release();
    h_up.release();
    h_down.release();
    if (stream) (void)hipStreamDestroy(stream);
}

// Estimated entries per atom from the density, plus head room; refined after the first build.
static int estimate_maxp(Strucs ss, double rc) {
    double worst = 0;
    for (const auto& s : ss) {
        double rho = 0;
        if (s.cell) {
            const double* a = s.cell;
            const double vol = std::fabs(a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
                                         a[2] * (a[3] * a[7] - a[4] * a[6]));
            rho = vol > 0 ? s.natoms / vol : 0;
        }
        if (!(rho > 0) && s.natoms > 0) {
            // open structure: the density of its bounding box (each side padded by 2 A) bounds
            // the mean count; atoms in a denser region are caught by the count check after the
            // build
            double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
            for (int i = 0; i < s.natoms; ++i)
                for (int c = 0; c < 3; ++c) {
                    lo[c] = std::min(lo[c], s.pos[3 * i + c]);
                    hi[c] = std::max(hi[c], s.pos[3 * i + c]);
                }
            const double vol = (hi[0] - lo[0] + 2.0) * (hi[1] - lo[1] + 2.0) * (hi[2] - lo[2] + 2.0);
            rho = s.natoms / vol;
        }
        const double p = std::min((double)s.natoms, rho * 4.0 / 3.0 * M_PI * rc * rc * rc);
        worst = std::max(worst, p);
    }
    int m = (int)std::ceil(worst * 1.5 + 16);
    m = (m + 31) / 32 * 32;
    return std::max(64, std::min(m, 4096));
}

void Engine::upload(Strucs ss) {
    HIPCHK(hipStreamSynchronize(stream));  // the staging buffer may still feed an earlier copy
    resident_valid = false;
    drop_graph();
    eager_runs = 0;
    mom_fresh = 0;
    assemble_mode = react_rows;
    react_missed = 0;
    nstruc = (int)ss.size();
    cell_natoms_max = 0;
    h_natoms.clear();
    for (const auto& s : ss) h_natoms.push_back(s.natoms);
    ntot = 0;
    for (const auto& s : ss) {
        if (s.natoms <= 0) throw std::invalid_argument("natoms must be positive");
        if (!s.pos || !s.type) throw std::invalid_argument("structure without positions or types");
        ntot += s.natoms;
    }
    // per structure: cell, images, cell-list bins
    h_strucs.clear();
    up_modes.clear();
    int bins = 0;
    for (int k = 0, a0 = 0; k < nstruc; a0 += ss[k].natoms, ++k) {
        DevStruc d = make_struc(ss[k], host.rc, a0);
        if (d.mode == 1) {
            d.bin0 = bins;
            bins += d.nbin[0] * d.nbin[1] * d.nbin[2];
            cell_natoms_max = std::max(cell_natoms_max, ss[k].natoms);
        }
        h_strucs.push_back(d);
        up_modes.push_back((char)d.mode);
    }
    nbins = bins;
    // per atom (species, structure of each atom, centre lists, species-sorted atoms): kept from
    // the previous upload when the atom counts, species and search modes are the same, as in a
    // driver loop over one structure
    bool same = up_tables && !bound && up_natoms == h_natoms && up_modes_prev == up_modes;
    for (int k = 0, a0 = 0; same && k < nstruc; a0 += ss[k].natoms, ++k)
        same = std::memcmp(ss[k].type, &h_type[a0], sizeof(int32_t) * ss[k].natoms) == 0;
    up_tables = false;
    if (!same) {
        h_type.resize(ntot);
        h_astruc.resize(ntot);
        h_cell_centres.clear();
        h_pair_centres.clear();
        for (int k = 0, a0 = 0; k < nstruc; a0 += ss[k].natoms, ++k) {
            const auto& st = ss[k];
            for (int i = 0; i < st.natoms; ++i) {
                const int t = st.type[i];
                if (t < 0 || t >= host.T) throw std::invalid_argument("atom type out of range");
                h_type[a0 + i] = t;
                h_astruc[a0 + i] = k;
                (h_strucs[k].mode == 1 ? h_cell_centres : h_pair_centres).push_back(a0 + i);
            }
        }
        // per-species atom lists for the network GEMMs
        sp_off.assign(host.T + 1, 0);
        for (int i = 0; i < ntot; ++i) sp_off[h_type[i] + 1]++;
        for (int t = 0; t < host.T; ++t) sp_off[t + 1] += sp_off[t];
        h_spatoms.resize(ntot);
        sp_fill.assign(sp_off.begin(), sp_off.end() - 1);
        for (int i = 0; i < ntot; ++i) h_spatoms[sp_fill[h_type[i]]++] = i;
        h_centres.assign(h_cell_centres.begin(), h_cell_centres.end());
        h_centres.insert(h_centres.end(), h_pair_centres.begin(), h_pair_centres.end());
    }
    // one device buffer: positions | structures | species | structure of each atom | centres
    // (cell-list first) | species-sorted atoms, each 256-byte aligned; the first two change with
    // every call
    auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
    const size_t o_pos = 0, o_st = o_pos + al((size_t)ntot * 24), o_type = o_st + al(h_strucs.size() * sizeof(DevStruc)),
                 o_ast = o_type + al((size_t)ntot * 4), o_cen = o_ast + al((size_t)ntot * 4),
                 o_spa = o_cen + al((size_t)ntot * 4), total = o_spa + al((size_t)ntot * 4);
    const size_t up_cap = d_up.cap;
    d_up.reserve(total);
    same = same && d_up.cap == up_cap;  // a new device buffer (it only grows) takes every table
    char* ds = (char*)d_up.p;
    d_pos.p = ds + o_pos;
    d_type.p = ds + o_type;
    d_astruc.p = ds + o_ast;
    d_strucs.p = ds + o_st;
    d_centres.p = ds + o_cen;
    d_spatoms.p = ds + o_spa;
    if (bound) {
        // no staging buffer: each piece goes straight from host memory into its slice
        h_pos.resize(3 * (size_t)ntot);
        for (int k = 0, a0 = 0; k < nstruc; a0 += ss[k].natoms, ++k)
            std::memcpy(&h_pos[3 * (size_t)a0], ss[k].pos, 3 * sizeof(double) * ss[k].natoms);
        HIPCHK(hipMemcpyAsync(d_pos.p, h_pos.data(), h_pos.size() * 8, hipMemcpyHostToDevice, stream));
        HIPCHK(hipMemcpyAsync(d_type.p, h_type.data(), (size_t)ntot * 4, hipMemcpyHostToDevice, stream));
        HIPCHK(hipMemcpyAsync(d_astruc.p, h_astruc.data(), (size_t)ntot * 4, hipMemcpyHostToDevice, stream));
        HIPCHK(hipMemcpyAsync(d_strucs.p, h_strucs.data(), h_strucs.size() * sizeof(DevStruc), hipMemcpyHostToDevice,
                              stream));
        HIPCHK(hipMemcpyAsync(d_centres.p, h_centres.data(), (size_t)ntot * 4, hipMemcpyHostToDevice, stream));
        HIPCHK(hipMemcpyAsync(d_spatoms.p, h_spatoms.data(), (size_t)ntot * 4, hipMemcpyHostToDevice, stream));
        maxp = std::min(estimate_maxp(ss, host.rc), maxp_cap);
        return;
    }
    h_up.reserve(total);
    char* hs = (char*)h_up.p;
    for (int k = 0, a0 = 0; k < nstruc; a0 += ss[k].natoms, ++k)
        std::memcpy(hs + o_pos + 24 * (size_t)a0, ss[k].pos, 3 * sizeof(double) * ss[k].natoms);
    std::memcpy(hs + o_st, h_strucs.data(), h_strucs.size() * sizeof(DevStruc));
    if (!same) {
        std::memcpy(hs + o_type, h_type.data(), (size_t)ntot * 4);
        std::memcpy(hs + o_ast, h_astruc.data(), (size_t)ntot * 4);
        std::memcpy(hs + o_cen, h_centres.data(), (size_t)ntot * 4);
        std::memcpy(hs + o_spa, h_spatoms.data(), (size_t)ntot * 4);
    }
    HIPCHK(hipMemcpyAsync(d_up.p, h_up.p, same ? o_type : total, hipMemcpyHostToDevice, stream));
    up_natoms = h_natoms;
    up_modes_prev = up_modes;
    up_tables = true;
    // serving lanes size each batch's rows afresh (a batch of large neighbour counts must not
    // leave every later batch on wide rows); other workspaces keep the stride they reached
    if (maxp == 0 || defer_count) maxp = estimate_maxp(ss, host.rc);
}

// Whether the per-centre kernels' LDS holds rows of lm staged entries (FP64 sizes, which bound
// the FP32 ones).
bool Engine::lds_fits(int lm) const {
    const bool gen = pb.plan.generic;
    if (symm_lds(shape, true, gen, lm, pb.plan.ldg) > 65536) return false;
    bool all = true;
    for (int t = 0; t < pb.plan.T; ++t) all = all && pb.plan.mom_ok[t];
    if (all && !gen && (moments & 2) && force_mom_lds(shape, true, force_waves, lm, pb.plan.ldg, pb.plan.T) <= 65536)
        return true;
    return force_lds(shape, true, gen, lm, pb.plan.ldg) <= 65536;
}

// A workspace for calls that allocate nothing: the host bookkeeping reserved for max_atoms
// atoms up front (so the calls' resizes stay within capacity), every batch buffer bound to a slice
// of the scratch array per call.
void Engine::make_bound(int max_atoms) {
    bound = true;
    pipeline = 0;
    const size_t n = (size_t)max_atoms;
    h_pos.reserve(3 * n);
    h_type.reserve(n);
    h_astruc.reserve(n);
    h_cell_centres.reserve(n);
    h_pair_centres.reserve(n);
    h_centres.reserve(n);
    h_spatoms.reserve(n);
    h_eatom.reserve(n);
    h_F.reserve(3 * n);
    h_E.reserve(1);
    h_natoms.reserve(1);
    h_strucs.reserve(1);
    sp_off.reserve(kMaxSpecies + 1);
    sp_fill.reserve(kMaxSpecies + 1);
}

// Slices of [base, base + bytes) for one structure: the per-atom buffers, the cell-list bins of
// its cell, the descriptor rows (FP64 widths, which hold the FP32 forms too), and what remains for
// the neighbour rows, force terms and receiver rows at the largest stride that fits (maxp_cap).
void Engine::bind_scratch(void* base, size_t bytes, const nnp_structure& s) {
    const size_t N = (size_t)s.natoms;
    const DevStruc d = make_struc(s, host.rc, 0);
    const size_t nb = d.mode == 1 ? (size_t)d.nbin[0] * d.nbin[1] * d.nbin[2] : 1;
    auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
    const size_t up = al(N * 24) + 4 * al(N * 4) + al(sizeof(DevStruc));
    const size_t ldg = (size_t)pb.plan.ldg;
    size_t off = 0;
    char* b = (char*)base;
    auto take = [&](DevMem& m, size_t sz) {
        m.bind(b + off, al(sz));
        off += al(sz);
    };
    take(d_up, up);
    take(d_maxcnt, 64);
    take(d_work, 64);
    take(d_E, 8);
    take(d_sstart, 8);
    for (DevMem* m : {&d_cnt, &d_abin, &d_rcount}) take(*m, N * 4);
    take(d_eatom, N * 8);
    take(d_F, N * 24);
    for (DevMem* m : {&d_bcount, &d_bstart, &d_bfill}) take(*m, nb * 4);
    take(d_G, N * ldg * 8);
    take(d_dEdG, N * ldg * 8);
    // per row slot: neighbour index 4, metadata 4, displacement 32, FP64 force term 32, receiver
    // row entry 4 bytes
    const size_t slot = 76;
    const size_t rest = bytes > off + 5 * 256 ? bytes - off - 5 * 256 : 0;
    const int fit = (int)std::min<size_t>(rest / (N * slot) / 32 * 32, 8192);
    if (fit < 32)
        throw std::invalid_argument("nnp_compute_noalloc: " + std::to_string(N) + " atoms need more than the " +
                                    std::to_string(kScratchBytes >> 20) + " MiB scratch array holds");
    maxp_cap = fit;
    take(d_nbrj, N * fit * 4);
    take(d_meta, N * fit * 4);
    take(d_nbrr, N * fit * 32);
    take(d_X, N * fit * 32);
    take(d_rlist, N * fit * 4);
    HIPCHK(hipMemsetAsync(d_maxcnt.p, 0, 64, stream));  // counters the mode does not write read as zero
    maxp = 0;
    rcap = 0;
}

void Engine::build_neighbors(bool all_images) {
    all_images_ = all_images;
    speculated = false;
    count_pending = false;
    mom_fresh = 0;  // new rows
    for (int attempt = 0; attempt < 4; ++attempt) {
        const size_t ent = (size_t)ntot * maxp;
        if (ent >= (size_t)1 << 31) throw std::invalid_argument("batch too large for 32-bit entry indices");
        d_cnt.reserve((size_t)ntot * 4);
        d_nbrj.reserve(ent * 4);
        d_meta.reserve(ent * 4);
        d_nbrr.reserve(ent * sizeof(double4));
        if (!d_maxcnt.p) {  // counters the current mode does not write are read back as zero
            d_maxcnt.reserve(64);
            HIPCHK(hipMemsetAsync(d_maxcnt.p, 0, 64, stream));
        }
        d_work.reserve(64);
        if (nbins > 0) {
            d_abin.reserve((size_t)ntot * 4);
            d_bcount.reserve((size_t)nbins * 4);
            d_bstart.reserve((size_t)nbins * 4);
            d_bfill.reserve((size_t)nbins * 4);
        }
        // serving lanes check the count at the download instead (one host sync per batch), staging
        // full rows in LDS; when full rows do not fit, the count is read here and the rows staged
        // at the realised maximum
        const bool defer = defer_count && lds_fits(maxp);
        // synchronous calls stage the rows at the previous call's maximum with the resident path's
        // head room (LDS rows decide the per-centre kernels' occupancy)
        int lm = 0;
        const bool spec = !defer && speculate_rows && spec_entries > 0 && spec_entries <= maxp &&
                          lds_fits(lm = std::min(maxp, (spec_entries + spec_entries / 16 + 4 + 7) / 8 * 8));
        NeighborArgs a = neighbor_args(all_images);
        // without a read here, the count maximum is folded into the force assembly's launch (or
        // taken just before the download)
        a.count_max = !(defer || spec);
        HIPCHK(launch_neighbors(a, stream));
        if (defer || spec) {
            count_pending = true;
            lmaxp = defer ? maxp : lm;
            speculated = spec;
            return;
        }
        int mc = 0;
        HIPCHK(hipMemcpyAsync(&mc, d_maxcnt.p, 4, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        max_entries = mc;
        if (mc <= maxp) {
            lmaxp = std::min(maxp, std::max(8, (mc + 7) / 8 * 8));
            return;
        }
        maxp = (mc + 31) / 32 * 32;
    }
    throw std::runtime_error("neighbour list did not converge");
}

// The moment bits (of `bits`: 1 descriptors, 2 forces) a kernel whose own LDS is base_lds bytes
// runs with: the option's bits, when some species qualifies and the moment tables fit in 64 KiB.
int Engine::mom_flag(size_t base_lds, bool fp64, int bits) {
    int any = 0;
    for (int t = 0; t < pb.plan.T; ++t) any |= pb.plan.mom_ok[t];
    const int use = moments & bits;
    const int got = (use && any && base_lds + mom_lds(shape, fp64, lmaxp, pb.plan.T, use) <= 65536) ? use : 0;
    last_mom |= got;
    return got;
}

template <typename T>
void Engine::descriptors() {
    const size_t lds = symm_lds(shape, sizeof(T) == 8, pb.plan.generic, lmaxp, pb.plan.ldg);
    if (lds > 65536) throw unsupported_error("too many neighbours per atom for the descriptor kernel's LDS");
    // mom = 2 when every species takes the moments: no triplet blocks, so the reduction buffer
    // only holds power-table chunks and the kernel's LDS drops to about 9 KB (FP64)
    bool all = true;
    for (int t = 0; t < pb.plan.T; ++t) all = all && pb.plan.mom_ok[t];
    const int mom = mom_flag(lds, sizeof(T) == 8, 1) ? (all ? 2 : 1) : 0;
    d_G.reserve((size_t)ntot * pb.plan.ldg * sizeof(T));
    // the moment tables go to the force-term kernel when it takes the moment-only form (not in the
    // workspace of calls that allocate nothing)
    const bool keep = mom == 2 && !bound && !pb.plan.generic && (moments & 2);
    mom_fresh = 0;
    if (keep) d_mom.reserve((size_t)ntot * pb.plan.T * shape_ne(shape) * kMomK * sizeof(T));
    HIPCHK(launch_symm<T>(shape, pb.plan.generic, (const DevPlan*)d_plan.p, (const double*)d_rfe.p, (const double*)d_rfr.p,
                          (const int*)d_type.p, (const int*)d_centres.p, ntot, (const int*)d_cnt.p,
                          (const int*)d_meta.p, (const double4*)d_nbrr.p, maxp, lmaxp, (T*)d_G.p, pb.plan.ldg, mom,
                          pb.plan.T, stream, keep ? (T*)d_mom.p : nullptr));
    if (keep) mom_fresh = (int)sizeof(T);
}

void Engine::networks(bool mixed, bool vac) {
    d_eatom.reserve((size_t)ntot * 8);
    d_dEdG.reserve((size_t)ntot * pb.plan.ldg * (mixed ? 4 : 8));
    const NetLaunch cfg = net_launch(mixed, vac);
    HIPCHK(launch_net(!mixed, (const DevPlan*)d_plan.p, pb.plan, mixed ? d_w32.p : d_w64.p, sp_off.data(),
                      (const int*)d_spatoms.p, d_G.p, pb.plan.ldg, vac ? 1 : 0, (double*)d_eatom.p, d_dEdG.p,
                      net_bf16_terms, cfg, stream));
}

// Launch configuration of the network kernel.  FP64 batches of at most two 16-atom tiles per CU
// copy the weights into LDS, one block per tile (the latency-bound small-batch form).  Larger
// FP64 batches time, the first time their log2 tile count is seen, the weights read through L2
// (one block per tile) against the weights in LDS on a persistent grid of 1 or 2 blocks per CU,
// on the batch at hand, and keep the fastest.  Every candidate computes the same results.
// Serving lanes and captured passes read the cache but do not time kernels.
NetLaunch Engine::net_launch(bool mixed, bool vac) {
    NetLaunch cfg;
    if (mixed || !net_wl_fits(pb.plan)) return cfg;
    if (net_launch_opt >= 0) {  // fixed by the option
        cfg.wl = net_launch_opt > 0;
        cfg.grid = net_launch_opt * num_cus;
        return cfg;
    }
    const int ntiles = net_tiles(sp_off.data(), pb.plan.T);
    if (ntiles <= 2 * num_cus) {
        cfg.wl = 1;
        return cfg;
    }
    int lg = 0;
    while ((1 << (lg + 1)) <= ntiles) ++lg;
    for (const auto& kv : net_cache)
        if (kv.first == lg) return kv.second;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    HIPCHK(hipStreamIsCapturing(stream, &cs));
    if (!may_tune || cs != hipStreamCaptureStatusNone) return cfg;
    const NetLaunch cand[] = {{0, 0}, {1, num_cus}, {1, 2 * num_cus}};
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    last_net_tune.clear();
    double best_ms = 1e300;
    for (const NetLaunch& c : cand) {
        double tot = 0;
        for (int rep = 0; rep < 3; ++rep) {
            HIPCHK(hipEventRecord(e0, stream));
            HIPCHK(launch_net(true, (const DevPlan*)d_plan.p, pb.plan, d_w64.p, sp_off.data(), (const int*)d_spatoms.p,
                              d_G.p, pb.plan.ldg, vac ? 1 : 0, (double*)d_eatom.p, d_dEdG.p, net_bf16_terms, c, stream));
            HIPCHK(hipEventRecord(e1, stream));
            HIPCHK(hipEventSynchronize(e1));
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, e0, e1));
            if (rep > 0) tot += ms;  // the first repetition warms the caches
        }
        last_net_tune.push_back({c, tot / 2});
        if (tot / 2 < best_ms) {
            best_ms = tot / 2;
            cfg = c;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    net_cache.push_back({lg, cfg});
    return cfg;
}

template <typename T>
void Engine::force_terms() {
    const bool fp64 = sizeof(T) == 8;
    d_X.reserve((size_t)ntot * maxp * 4 * sizeof(T));
    bool all = true;
    for (int t = 0; t < pb.plan.T; ++t) all = all && pb.plan.mom_ok[t];
    if (all && !pb.plan.generic && (moments & 2) &&
        force_mom_lds(shape, fp64, force_waves, lmaxp, pb.plan.ldg, pb.plan.T) <= 65536) {
        last_mom |= 2;
        HIPCHK(launch_force_mom<T>(shape, force_waves, (const DevPlan*)d_plan.p, (const double*)d_rfe.p, (const double*)d_rfr.p,
                                   (const int*)d_type.p, (const int*)d_centres.p, ntot, (const int*)d_cnt.p,
                                   (const int*)d_meta.p, (const double4*)d_nbrr.p, maxp, lmaxp, (const T*)d_dEdG.p,
                                   pb.plan.ldg, (T*)d_X.p, pb.plan.T, stream,
                                   mom_fresh == (int)sizeof(T) ? (const T*)d_mom.p : nullptr));
        return;
    }
    const size_t lds = force_lds(shape, fp64, pb.plan.generic, lmaxp, pb.plan.ldg);
    if (lds > 65536) throw unsupported_error("too many neighbours per atom for the force kernel's LDS");
    const int mom = mom_flag(lds, fp64, 2) ? 1 : 0;
    HIPCHK(launch_force<T>(shape, pb.plan.generic, (const DevPlan*)d_plan.p, (const double*)d_rfe.p, (const double*)d_rfr.p,
                           (const int*)d_type.p, (const int*)d_centres.p, ntot, (const int*)d_cnt.p,
                           (const int*)d_meta.p, (const double4*)d_nbrr.p, maxp, lmaxp, (const T*)d_dEdG.p, pb.plan.ldg,
                           (T*)d_X.p, mom, pb.plan.T, stream));
}

template <typename T>
void Engine::assemble(const EnergyTail* tail) {
    d_F.reserve((size_t)ntot * 3 * 8);
    AssembleArgs a;
    a.S = (const DevStruc*)d_strucs.p;
    a.atom_struc = (const int*)d_astruc.p;
    a.cnt = (const int*)d_cnt.p;
    a.nbr_j = (const int*)d_nbrj.p;
    a.meta = (const int*)d_meta.p;
    a.maxp = maxp;
    a.ntot = ntot;
    a.all_images = all_images_ ? 1 : 0;
    a.F = (double*)d_F.p;
    a.react = (int*)d_maxcnt.p + 1;
    a.mode = assemble_mode;
    a.rcount = nullptr;
    a.rlist = nullptr;
    a.rcap = 0;
    // atoms in bin order (cell-list structures) then the rest, for the XCD-local block order
    a.ncell = nbins > 0 ? (int)h_cell_centres.size() : 0;
    a.bin_atoms = a.ncell > 0 ? (const int*)d_centres.p : nullptr;
    a.pair_centres = (const int*)d_centres.p + a.ncell;
    if (assemble_mode == 1) {
        // a receiver row holds about as many entries as a centre's row
        rcap = std::max(rcap, std::min(maxp, 8192));
        if ((size_t)rcap * 8 > 65536) throw unsupported_error("too many reactions per atom for the receiver rows");
        d_rcount.reserve((size_t)ntot * 4);
        d_rlist.reserve((size_t)ntot * rcap * 4);
        a.rcount = (int*)d_rcount.p;
        a.rlist = (int*)d_rlist.p;
        a.rcap = rcap;
    }
    EnergyTail counts;
    if (!tail && count_pending) {  // count blocks only: the energies of a synchronous call are summed on the host
        counts = energy_tail(true);
        counts.nstruc = 0;
        tail = &counts;
    }
    if (tail && tail->ncount > 0) count_pending = false;
    HIPCHK(launch_assemble<T>(a, (const T*)d_X.p, tail, stream));
}

void Engine::assemble_prec(bool mixed, const EnergyTail* tail) {
    if (mixed) assemble<float>(tail);
    else assemble<double>(tail);
}

bool Engine::reactions_redo(const int* c) {
    if (assemble_mode == 0) {
        react_missed = c[1];
        if (c[1] == 0) return false;
        assemble_mode = 1;
        return true;
    }
    if (c[2] <= rcap) return false;
    rcap = (c[2] + c[2] / 8 + 8 + 31) / 32 * 32;
    return true;
}

template <typename T>
void Engine::forces() {
    force_terms<T>();
    assemble<T>();
}

void Engine::run(Strucs ss, uint32_t flags) {
    const bool mixed = flags & NNP_PREC_MIXED;
    last_mom = 0;
    upload(ss);
    build_neighbors(flags & NNP_ALL_IMAGES);
    if (pipeline == 1) {
        fused(mixed, flags, fused_blocks(mixed, flags));
        if (flags & NNP_FORCE) assemble_prec(mixed);
        return;
    }
    if (mixed) descriptors<float>();
    else descriptors<double>();
    networks(mixed, flags & NNP_VACUUM_ENERGY);
    if (flags & NNP_FORCE) {
        if (mixed) forces<float>();
        else forces<double>();
    }
}

void Engine::net_dims(int& maxw, int& dersum) const {
    maxw = 0;
    dersum = 0;
    for (int t = 0; t < pb.plan.T; ++t) {
        const DevNet& n = pb.plan.net[t];
        int ds = 0;
        maxw = std::max(maxw, n.nin_pad);
        for (int l = 0; l < n.L; ++l) {
            maxw = std::max(maxw, n.wpad[l]);
            ds += n.wpad[l];
        }
        dersum = std::max(dersum, ds);
    }
}

// Enqueue the fused persistent kernel over the current batch (work counter must be zero:
// k_zero in the neighbour launch clears it).
void Engine::fused(bool mixed, uint32_t flags, int nblocks) {
    int maxw, dersum;
    net_dims(maxw, dersum);
    const size_t lds = fused_lds(shape, !mixed, pb.plan.generic, lmaxp, pb.plan.ldg, maxw, dersum);
    if (lds > 65536) throw unsupported_error("too many neighbours per atom for the fused kernel's LDS");
    const int mom = mom_flag(lds, !mixed, (flags & NNP_FORCE) ? 3 : 1);
    d_eatom.reserve((size_t)ntot * 8);
    d_X.reserve((size_t)ntot * maxp * 4 * (mixed ? 4 : 8));
    d_F.reserve((size_t)ntot * 3 * 8);
    const int vac = (flags & NNP_VACUUM_ENERGY) ? 1 : 0, frc = (flags & NNP_FORCE) ? 1 : 0;
    if (mixed)
        HIPCHK(launch_fused<float>(shape, pb.plan.generic, nblocks, (const DevPlan*)d_plan.p, (const double*)d_rfe.p,
                                   (const double*)d_rfr.p, (const float*)d_w32.p, (const int*)d_type.p,
                                   (const int*)d_spatoms.p, ntot, (int*)d_work.p, (const int*)d_cnt.p,
                                   (const int*)d_meta.p, (const double4*)d_nbrr.p, maxp, lmaxp, pb.plan.ldg, maxw, dersum,
                                   vac, frc, (double*)d_eatom.p, (float*)d_X.p, mom, pb.plan.T, stream));
    else
        HIPCHK(launch_fused<double>(shape, pb.plan.generic, nblocks, (const DevPlan*)d_plan.p, (const double*)d_rfe.p,
                                    (const double*)d_rfr.p, (const double*)d_w64.p, (const int*)d_type.p,
                                    (const int*)d_spatoms.p, ntot, (int*)d_work.p, (const int*)d_cnt.p,
                                    (const int*)d_meta.p, (const double4*)d_nbrr.p, maxp, lmaxp, pb.plan.ldg, maxw, dersum,
                                    vac, frc, (double*)d_eatom.p, (double*)d_X.p, mom, pb.plan.T, stream));
}

// Persistent grid of the fused kernel.  The autotuner times the kernel on the current batch
// for several blocks-per-CU counts the first time a (shape, precision, maxp, batch size)
// combination is seen and keeps the fastest; the launch geometry does not change any result.
int Engine::fused_blocks(bool mixed, uint32_t flags) {
    if (fused_blocks_per_cu > 0) return std::min(ntot, num_cus * fused_blocks_per_cu);
    int lg = 0;
    while ((1 << (lg + 1)) <= ntot) ++lg;
    const TuneKey key{shape, mixed ? 0 : 1, lg};
    for (const auto& kv : tune_cache)
        if (kv.first == key) return std::min(ntot, num_cus * kv.second);
    // small batches, and serving lanes: the nearest tuned batch size of this shape, else 4
    if (!may_tune || ntot < 4 * num_cus) {
        int best = 4, dist = 1 << 30;
        for (const auto& kv : tune_cache)
            if (kv.first.shape == key.shape && kv.first.fp64 == key.fp64 && std::abs(kv.first.ntot_log2 - lg) < dist) {
                dist = std::abs(kv.first.ntot_log2 - lg);
                best = kv.second;
            }
        return std::min(ntot, num_cus * best);
    }
    const int cand[] = {1, 2, 3, 4, 6, 8, 12, 16};
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    last_tune.clear();
    int best = 4;
    double best_ms = 1e300;
    for (int m : cand) {
        const int nb = std::min(ntot, num_cus * m);
        double tot = 0;
        for (int rep = 0; rep < 3; ++rep) {
            HIPCHK(hipMemsetAsync(d_work.p, 0, 4, stream));
            HIPCHK(hipEventRecord(e0, stream));
            fused(mixed, flags, nb);
            HIPCHK(hipEventRecord(e1, stream));
            HIPCHK(hipEventSynchronize(e1));
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, e0, e1));
            if (rep > 0) tot += ms;  // first repetition warms the caches
        }
        last_tune.push_back({m, tot / 2});
        if (tot / 2 < best_ms) {
            best_ms = tot / 2;
            best = m;
        }
        if (nb == ntot) break;  // more blocks than atoms changes nothing
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    HIPCHK(hipMemsetAsync(d_work.p, 0, 4, stream));
    tune_cache.push_back({key, best});
    return std::min(ntot, num_cus * best);
}

void Engine::stage_begin(int s) {
    if (!profiling) return;
    if (!ev[s][0]) {
        HIPCHK(hipEventCreate(&ev[s][0]));
        HIPCHK(hipEventCreate(&ev[s][1]));
    }
    HIPCHK(hipEventRecord(ev[s][0], stream));
}
void Engine::stage_end(int s, int launches) {
    if (!profiling) return;
    HIPCHK(hipEventRecord(ev[s][1], stream));
    pending[s] = launches;
}
void Engine::profile_collect() {
    for (int s = 0; s < ST_COUNT; ++s) {
        if (!pending[s]) continue;
        HIPCHK(hipEventSynchronize(ev[s][1]));
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, ev[s][0], ev[s][1]));
        stage_ms[s] += ms;
        stage_calls[s] += pending[s];
        pending[s] = 0;
    }
}

EnergyTail Engine::energy_tail(bool count_max) {
    d_E.reserve((size_t)nstruc * 8);
    EnergyTail t;
    t.eatom = (const double*)d_eatom.p;
    t.sstart = (const int*)d_sstart.p;
    t.E = (double*)d_E.p;
    t.nstruc = nstruc;
    t.ncount = count_max ? std::min(1024, (ntot + 1023) / 1024) : 0;
    t.cnt = (const int*)d_cnt.p;
    t.ntot = ntot;
    t.maxcnt = (int*)d_maxcnt.p;
    return t;
}

void Engine::energies(bool count_max) {
    const EnergyTail t = energy_tail(count_max);
    hipLaunchKernelGGL(k_energy, dim3(t.nstruc + t.ncount), dim3(kWave), 0, stream, t);
    HIPCHK(hipGetLastError());
}

// Upload a batch and size the neighbour rows once (with head room), so later passes run
// without a host round trip.
void Engine::resident_load(Strucs ss, uint32_t flags) {
    upload(ss);
    build_neighbors(flags & NNP_ALL_IMAGES);
    const int want = (max_entries + max_entries / 8 + 8 + 31) / 32 * 32;
    if (want > maxp) {
        maxp = want;
        build_neighbors(flags & NNP_ALL_IMAGES);
    }
    // LDS rows with a little head room over the realised maximum (positions may move)
    lmaxp = std::min(maxp, (max_entries + max_entries / 16 + 4 + 7) / 8 * 8);
    std::vector<int> st(nstruc + 1, 0);
    for (int k = 0; k < nstruc; ++k) st[k + 1] = st[k] + ss[k].natoms;
    d_sstart.reserve(st.size() * 4);
    HIPCHK(hipMemcpyAsync(d_sstart.p, st.data(), st.size() * 4, hipMemcpyHostToDevice, stream));
    if (pipeline == 1) (void)fused_blocks(flags & NNP_PREC_MIXED, flags);  // tune before any capture
    HIPCHK(hipStreamSynchronize(stream));
    resident_valid = true;
}

void Engine::drop_graph() {
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    graph_exec = nullptr;
    graph = nullptr;
}

// One full pass on the resident batch.  The first pass with given flags runs eagerly (it
// allocates every buffer); later passes replay a captured graph of the same launches.  With
// profiling on, passes run eagerly so the stage events can be recorded.
void Engine::resident_run(uint32_t flags) {
    if (!use_graph || profiling) {
        enqueue_pass(flags);
        return;
    }
    if (graph_exec && graph_flags == flags && graph_maxp == maxp && graph_lmaxp == lmaxp) {
        HIPCHK(hipGraphLaunch(graph_exec, stream));
        return;
    }
    if (eager_runs == 0 || graph_flags != flags || graph_maxp != maxp || graph_lmaxp != lmaxp) {
        drop_graph();
        graph_flags = flags;
        graph_maxp = maxp;
        graph_lmaxp = lmaxp;
        eager_runs = 1;
        enqueue_pass(flags);
        return;
    }
    HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    try {
        enqueue_pass(flags);
    } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(stream, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    HIPCHK(hipStreamEndCapture(stream, &graph));
    HIPCHK(hipGraphInstantiate(&graph_exec, graph, nullptr, nullptr, 0));
    HIPCHK(hipGraphLaunch(graph_exec, stream));
}

// The neighbour launch's arguments over the current batch and rows (buffers already sized).
NeighborArgs Engine::neighbor_args(bool all_images) const {
    NeighborArgs a;
    a.S = (const DevStruc*)d_strucs.p;
    a.atom_struc = (const int*)d_astruc.p;
    a.pos = (const double*)d_pos.p;
    a.type = (const int*)d_type.p;
    a.ntot = ntot;
    a.nstruc = nstruc;
    a.rc2 = host.rc * host.rc;
    a.maxp = maxp;
    a.cnt = (int*)d_cnt.p;
    a.nbr_j = (int*)d_nbrj.p;
    a.nbr_meta = (int*)d_meta.p;
    a.nbr_r = (double4*)d_nbrr.p;
    a.maxcnt = (int*)d_maxcnt.p;
    a.work = (int*)d_work.p;
    a.cell_centres = (const int*)d_centres.p;
    a.ncell = (int)h_cell_centres.size();
    a.pair_centres = (const int*)d_centres.p + h_cell_centres.size();
    a.npair = (int)h_pair_centres.size();
    a.nbins = nbins;
    a.cell_natoms_max = cell_natoms_max;
    a.all_images = all_images ? 1 : 0;
    a.atom_bin = (int*)d_abin.p;
    a.bin_count = (int*)d_bcount.p;
    a.bin_start = (int*)d_bstart.p;
    a.bin_fill = (int*)d_bfill.p;
    // the binning writes the cell-list atoms in bin order over the first ncell centres, so every
    // per-centre kernel walking the centres in XCD-local order takes spatially compact runs
    a.bin_atoms = (int*)d_centres.p;
    return a;
}

// Enqueue one full pass on the engine's stream (no host sync).  With rows_ready the neighbour
// rows (and the cleared per-pass counters) were provided by the caller (the MD step's list
// refresh), so the pass starts with the descriptors.
void Engine::enqueue_pass(uint32_t flags, bool rows_ready) {
    last_mom = 0;
    const int fused_nb = pipeline == 1 ? fused_blocks(flags & NNP_PREC_MIXED, flags) : 0;
    const bool mixed = flags & NNP_PREC_MIXED;
    if (!rows_ready) {
        NeighborArgs a = neighbor_args(flags & NNP_ALL_IMAGES);
        a.count_max = false;  // folded into the energy launch at the end of the pass
        stage_begin(ST_NEIGH);
        HIPCHK(launch_neighbors(a, stream));
        stage_end(ST_NEIGH);
    }
    if (pipeline == 1) {
        stage_begin(ST_SYMM);  // descriptors + networks + force terms in one kernel
        fused(mixed, flags, fused_nb);
        stage_end(ST_SYMM);
        if (flags & NNP_FORCE) {
            stage_begin(ST_ASSEMBLE);  // forces, energies and the count check in one launch
            const EnergyTail tail = energy_tail(true);
            assemble_prec(mixed, &tail);
            stage_end(ST_ASSEMBLE);
            return;
        }
        stage_begin(ST_ENERGY);
        energies(true);
        stage_end(ST_ENERGY);
        return;
    }
    stage_begin(ST_SYMM);
    if (mixed) descriptors<float>();
    else descriptors<double>();
    stage_end(ST_SYMM);
    stage_begin(ST_NET);
    networks(mixed, flags & NNP_VACUUM_ENERGY);
    stage_end(ST_NET);
    if (flags & NNP_FORCE) {
        // profiling may repeat the (idempotent) force-term launch back to back inside its events,
        // so the per-launch time carries a fraction of the events' own cost
        const int reps = profiling ? profile_repeat : 1;
        stage_begin(ST_FORCE);
        for (int r = 0; r < reps; ++r) {
            if (mixed) force_terms<float>();
            else force_terms<double>();
        }
        stage_end(ST_FORCE, reps);
        stage_begin(ST_ASSEMBLE);  // forces, energies and the count check in one launch
        const EnergyTail tail = energy_tail(true);
        assemble_prec(mixed, &tail);
        stage_end(ST_ASSEMBLE);
        return;
    }
    stage_begin(ST_ENERGY);
    energies(true);
    stage_end(ST_ENERGY);
}

// Staged download: [counters | E per structure | E_i | F] into pinned memory, one sync.
int Engine::download(bool energies, bool forces, bool keep_host) {
    if (count_pending) {  // a pass without the force assembly: the count maximum on its own
        HIPCHK(launch_count_max((const int*)d_cnt.p, ntot, (int*)d_maxcnt.p, stream));
        count_pending = false;
    }
    if (bound) {  // straight into the pre-reserved host vectors, no staging buffer
        h_eatom.resize(ntot);
        if (energies) h_E.resize(nstruc);
        if (forces) h_F.resize(3 * (size_t)ntot);
        HIPCHK(hipMemcpyAsync(counters, d_maxcnt.p, 16, hipMemcpyDeviceToHost, stream));
        if (energies) HIPCHK(hipMemcpyAsync(h_E.data(), d_E.p, (size_t)nstruc * 8, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipMemcpyAsync(h_eatom.data(), d_eatom.p, (size_t)ntot * 8, hipMemcpyDeviceToHost, stream));
        if (forces) HIPCHK(hipMemcpyAsync(h_F.data(), d_F.p, (size_t)ntot * 24, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        down_ea = h_eatom.data();
        down_f = forces ? h_F.data() : nullptr;
        return counters[0];
    }
    auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
    const size_t o_e = 256, o_ea = o_e + (energies ? al((size_t)nstruc * 8) : 0), o_f = o_ea + al((size_t)ntot * 8),
                 total = o_f + (forces ? al((size_t)ntot * 24) : 0);
    h_down.reserve(total);
    d_out.reserve(total);
    char* hd = (char*)h_down.p;
    const int ne = energies ? nstruc : 0, nf = forces ? 3 * ntot : 0;
    const int nthr = std::max(std::max(ntot, nf), std::max(ne, 2));
    hipLaunchKernelGGL(k_pack_out, dim3((nthr + 255) / 256), dim3(256), 0, stream, (const int64_t*)d_maxcnt.p,
                       (const int64_t*)d_E.p, ne, (const int64_t*)d_eatom.p, ntot, (const int64_t*)d_F.p, nf,
                       (int64_t*)d_out.p, (int)(o_e / 8), (int)(o_ea / 8), (int)(o_f / 8));
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(hd, d_out.p, total, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    std::memcpy(counters, hd, 16);
    const int mc = counters[0];
    if (energies) {
        h_E.resize(nstruc);
        std::memcpy(h_E.data(), hd + o_e, (size_t)nstruc * 8);
    }
    down_ea = (const double*)(hd + o_ea);
    down_f = forces ? (const double*)(hd + o_f) : nullptr;
    if (keep_host) {
        h_eatom.resize(ntot);
        std::memcpy(h_eatom.data(), hd + o_ea, (size_t)ntot * 8);
        if (forces) {
            h_F.resize(3 * (size_t)ntot);
            std::memcpy(h_F.data(), hd + o_f, (size_t)ntot * 24);
        }
        down_ea = h_eatom.data();
        down_f = forces ? h_F.data() : nullptr;
    }
    return mc;
}

void Engine::resident_fetch(uint32_t flags, nnp_result* res) {
    // the counters come back with the results (one round trip); a pass they show to be invalid
    // is redone and fetched again
    const bool frc = flags & NNP_FORCE;
    int mc = download(true, frc, false);
    profile_collect();
    if (mc > lmaxp) {  // the rows (or their LDS staging) overflowed: enlarge and redo the pass
        if (mc > maxp) {
            maxp = (mc + mc / 8 + 8 + 31) / 32 * 32;
            build_neighbors(flags & NNP_ALL_IMAGES);
        }
        lmaxp = std::min(maxp, (mc + mc / 16 + 4 + 7) / 8 * 8);
        drop_graph();
        enqueue_pass(flags);
        mc = download(true, frc, false);
        profile_collect();
    }
    // reactions the gather missed (or receiver rows too short): the batch takes receiver rows
    for (int attempt = 0; frc && reactions_redo(counters); ++attempt) {
        if (attempt == 3) throw std::runtime_error("force assembly: receiver rows did not converge");
        drop_graph();
        eager_runs = 0;
        enqueue_pass(flags);
        (void)download(true, frc, false);
        profile_collect();
    }
    int a0 = 0;
    for (int k = 0; k < nstruc; ++k) {
        const int n = h_natoms[k];
        res[k].energy = h_E[k];
        if (res[k].energy_atom) std::memcpy(res[k].energy_atom, down_ea + a0, (size_t)n * 8);
        if (frc && res[k].force) std::memcpy(res[k].force, down_f + 3 * (size_t)a0, (size_t)n * 24);
        a0 += n;
    }
}

void Engine::fetch(Strucs ss, uint32_t flags, nnp_result* res) {
    int mc = download(false, flags & NNP_FORCE, false);
    if ((defer_count && mc > maxp) || (speculated && mc > lmaxp)) {
        // rows overflowed, or the rows staged at the previous call's maximum were too short: redo
        // the batch with the count read back
        const bool dc = defer_count, sp = speculate_rows;
        defer_count = false;
        speculate_rows = false;
        try {
            run(ss, flags);
        } catch (...) {
            defer_count = dc;
            speculate_rows = sp;
            throw;
        }
        defer_count = dc;
        speculate_rows = sp;
        mc = download(false, flags & NNP_FORCE, false);
    }
    spec_entries = mc;
    // reactions the gather missed (or receiver rows too short): assemble again from receiver rows
    for (int attempt = 0; (flags & NNP_FORCE) && reactions_redo(counters); ++attempt) {
        if (attempt == 3) throw std::runtime_error("force assembly: receiver rows did not converge");
        assemble_prec(flags & NNP_PREC_MIXED);
        (void)download(false, true, false);
    }
    int a0 = 0;
    for (size_t k = 0; k < ss.size(); ++k) {
        const int n = ss[k].natoms;
        double e = 0;
        for (int i = 0; i < n; ++i) e += down_ea[a0 + i];  // nnp.cpp:394-399, atom order
        res[k].energy = e;
        if (res[k].energy_atom) std::memcpy(res[k].energy_atom, down_ea + a0, (size_t)n * 8);
        if ((flags & NNP_FORCE) && res[k].force) std::memcpy(res[k].force, down_f + 3 * (size_t)a0, (size_t)n * 24);
        a0 += n;
    }
}

}  // namespace atomnn

using namespace atomnn;

struct nnp_model {
    Engine* eng;
    // nnp_compute_noalloc: a workspace bound to the scratch array of the device the model was
    // loaded on (set up here, at load, so the calls themselves allocate nothing)
    Engine* noalloc = nullptr;
    void* scratch = nullptr;
    int dev = 0;
};

namespace atomnn {
// atoms a call of nnp_compute_noalloc may hold (host bookkeeping reserved at load); the scratch
// array bounds it further by the structure's neighbour rows
constexpr int kNoallocMaxAtoms = 16384;
}  // namespace atomnn

namespace atomnn {
Engine& engine_of(nnp_model* m) { return *m->eng; }
}  // namespace atomnn

template <typename F>
static int guarded(F&& f) {
    try {
        set_error("");
        return f();
    } catch (const io_error& e) {
        set_error(e.what());
        return NNP_EIO;
    } catch (const unsupported_error& e) {
        set_error(e.what());
        return NNP_EUNSUPPORTED;
    } catch (const hip_failure& e) {
        set_error(e.what());
        return NNP_EDEVICE;
    } catch (const std::bad_alloc&) {
        set_error("out of memory");
        return NNP_ENOMEM;
    } catch (const std::exception& e) {
        set_error(e.what());
        return NNP_EINVAL;
    }
}

extern "C" {

const char* nnp_last_error(void) { return last_error(); }

int nnp_load(const char* path, nnp_model** out) {
    return guarded([&]() {
        if (!path || !out) throw std::invalid_argument("nnp_load: null argument");
        *out = nullptr;
        HostModel hm = parse_ann(path);
        auto* m = new nnp_model{nullptr};
        try {
            m->eng = new Engine(std::move(hm));
            m->scratch = scratch_on_current_device(m->dev);
            m->noalloc = new Engine(*m->eng, 0);
            m->noalloc->make_bound(kNoallocMaxAtoms);
        } catch (...) {
            delete m->noalloc;
            delete m->eng;
            delete m;
            throw;
        }
        *out = m;
        return NNP_OK;
    });
}

void nnp_free(nnp_model* m) {
    if (!m) return;
    delete m->noalloc;
    delete m->eng;
    delete m;
}

int nnp_ntypes(const nnp_model* m) { return m ? m->eng->host.T : NNP_EINVAL; }
double nnp_rc(const nnp_model* m) { return m ? m->eng->host.rc : 0.0; }
int nnp_ninput(const nnp_model* m, int t) {
    if (!m || t < 0 || t >= m->eng->host.T) return NNP_EINVAL;
    return m->eng->host.sp[t].nG;
}
int nnp_type_index(const nnp_model* m, const char* name) {
    if (!m || !name) return -1;
    return m->eng->host.index(name);
}

int64_t nnp_neighbors(nnp_model* m, const nnp_structure* s, int32_t* counts, int64_t cap, int32_t* index,
                      int32_t* image, int32_t* is_min, double* r, double* d) {
    int64_t total = 0;
    const int rc = guarded([&]() {
        if (!m || !s || !counts) throw std::invalid_argument("nnp_neighbors: null argument");
        Engine& E = *m->eng;
        std::vector<nnp_structure> ss(1, *s);
        E.upload(ss);
        E.build_neighbors(false);
        const int n = s->natoms, mp = E.maxp;
        std::vector<int> cnt(n), nj((size_t)n * mp), meta((size_t)n * mp);
        std::vector<double4> rr((size_t)n * mp);
        HIPCHK(hipMemcpy(cnt.data(), E.d_cnt.p, (size_t)n * 4, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(nj.data(), E.d_nbrj.p, (size_t)n * mp * 4, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(meta.data(), E.d_meta.p, (size_t)n * mp * 4, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(rr.data(), E.d_nbrr.p, (size_t)n * mp * sizeof(double4), hipMemcpyDeviceToHost));
        const bool periodic = E.h_strucs[0].periodic;
        for (int i = 0; i < n; ++i) {
            counts[i] = cnt[i];
            total += cnt[i];
        }
        if (total <= cap) {
            int64_t q = 0;
            for (int i = 0; i < n; ++i)
                for (int k = 0; k < cnt[i]; ++k, ++q) {
                    const size_t e = (size_t)i * mp + k;
                    if (index) index[q] = nj[e];
                    if (image) image[q] = periodic ? (meta[e] >> 8) : -1;
                    if (is_min) is_min[q] = meta[e] & 1;
                    if (r) {
                        r[3 * q] = rr[e].x;
                        r[3 * q + 1] = rr[e].y;
                        r[3 * q + 2] = rr[e].z;
                    }
                    if (d) d[q] = rr[e].w;
                }
        }
        return NNP_OK;
    });
    return rc < 0 ? rc : total;
}

int nnp_symm(nnp_model* m, const nnp_structure* s, double* G, int ldg) {
    return guarded([&]() {
        if (!m || !s || !G) throw std::invalid_argument("nnp_symm: null argument");
        Engine& E = *m->eng;
        std::vector<nnp_structure> ss(1, *s);
        E.upload(ss);
        E.build_neighbors(false);
        E.descriptors<double>();
        const int n = s->natoms, lg = E.pb.plan.ldg;
        std::vector<double> g((size_t)n * lg);
        HIPCHK(hipMemcpyAsync(g.data(), E.d_G.p, g.size() * 8, hipMemcpyDeviceToHost, E.stream));
        HIPCHK(hipStreamSynchronize(E.stream));
        for (int i = 0; i < n; ++i)
            for (int k = 0; k < ldg; ++k) G[(size_t)i * ldg + k] = k < lg ? g[(size_t)i * lg + k] : 0.0;
        return NNP_OK;
    });
}

int nnp_compute(nnp_model* m, const nnp_structure* s, uint32_t flags, nnp_result* r) {
    return nnp_compute_batch(m, 1, s, flags, r);
}

int nnp_compute_batch(nnp_model* m, int n, const nnp_structure* s, uint32_t flags, nnp_result* r) {
    return guarded([&]() {
        if (!m || !s || !r || n <= 0) throw std::invalid_argument("nnp_compute: invalid argument");
        Engine& E = *m->eng;
        std::vector<nnp_structure> ss(s, s + n);
        struct Spec {  // row speculation for this call only
            Engine& e;
            explicit Spec(Engine& x) : e(x) { e.speculate_rows = true; }
            ~Spec() { e.speculate_rows = false; }
        } spec(E);
        E.run(ss, flags);
        E.fetch(ss, flags, r);
        return NNP_OK;
    });
}

int nnp_compute_noalloc(nnp_model* m, const nnp_structure* s, uint32_t flags, nnp_result* r) {
    return guarded([&]() {
        if (!m || !s || !r) throw std::invalid_argument("nnp_compute_noalloc: invalid argument");
        if (s->natoms <= 0 || s->natoms > kNoallocMaxAtoms)
            throw std::invalid_argument("nnp_compute_noalloc: natoms must be in 1.." + std::to_string(kNoallocMaxAtoms));
        int dev = 0;
        HIPCHK(hipGetDevice(&dev));
        if (dev != m->dev) throw std::invalid_argument("nnp_compute_noalloc: call it on the device the model was loaded on");
        Engine& B = *m->noalloc;
        const Engine& E = *m->eng;
        B.moments = E.moments;
        B.force_waves = E.force_waves;
        B.net_bf16_terms = E.net_bf16_terms;
        B.react_rows = E.react_rows;
        HIPCHK(hipStreamSynchronize(B.stream));
        B.bind_scratch(m->scratch, kScratchBytes, *s);
        const Strucs ss(s, 1);
        B.run(ss, flags);
        B.fetch(ss, flags, r);
        return NNP_OK;
    });
}

int nnp_resident_load(nnp_model* m, int n, const nnp_structure* s, uint32_t flags) {
    return guarded([&]() {
        if (!m || !s || n <= 0) throw std::invalid_argument("nnp_resident_load: invalid argument");
        m->eng->resident_load(std::vector<nnp_structure>(s, s + n), flags);
        return NNP_OK;
    });
}

int nnp_resident_run(nnp_model* m, uint32_t flags) {
    return guarded([&]() {
        if (!m || !m->eng->resident_valid) throw std::invalid_argument("nnp_resident_run: no resident batch");
        m->eng->resident_run(flags);
        return NNP_OK;
    });
}

int nnp_resident_fetch(nnp_model* m, uint32_t flags, nnp_result* r) {
    return guarded([&]() {
        if (!m || !r || !m->eng->resident_valid) throw std::invalid_argument("nnp_resident_fetch: no resident batch");
        m->eng->resident_fetch(flags, r);
        return NNP_OK;
    });
}

int nnp_sync(nnp_model* m) {
    return guarded([&]() {
        if (!m) throw std::invalid_argument("nnp_sync: null model");
        HIPCHK(hipStreamSynchronize(m->eng->stream));
        m->eng->profile_collect();
        return NNP_OK;
    });
}

int nnp_profile(nnp_model* m, int enable) {
    return guarded([&]() {
        if (!m) throw std::invalid_argument("nnp_profile: null model");
        Engine& E = *m->eng;
        HIPCHK(hipStreamSynchronize(E.stream));
        E.profile_collect();
        E.profiling = enable != 0;
        E.profile_repeat = enable > 1 ? enable : 1;
        for (int s = 0; s < Engine::ST_COUNT; ++s) {
            E.stage_ms[s] = 0;
            E.stage_calls[s] = 0;
        }
        return NNP_OK;
    });
}

int nnp_profile_read(nnp_model* m, double* ms, int64_t* calls, int nstage) {
    return guarded([&]() {
        if (!m || !ms) throw std::invalid_argument("nnp_profile_read: invalid argument");
        Engine& E = *m->eng;
        for (int s = 0; s < nstage && s < Engine::ST_COUNT; ++s) {
            ms[s] = E.stage_ms[s];
            if (calls) calls[s] = E.stage_calls[s];
        }
        return Engine::ST_COUNT;
    });
}

int nnp_set_option(nnp_model* m, const char* key, int64_t value) {
    return guarded([&]() {
        if (!m || !key) throw std::invalid_argument("nnp_set_option: invalid argument");
        Engine& E = *m->eng;
        const std::string k(key);
        if (k == "graph") {
            HIPCHK(hipStreamSynchronize(E.stream));
            E.use_graph = value != 0;
            E.drop_graph();
            E.eager_runs = 0;
        } else if (k == "net_launch") {
            if (value < -1 || value > 2) throw std::invalid_argument("nnp_set_option: net_launch is -1..2");
            HIPCHK(hipStreamSynchronize(E.stream));
            E.net_launch_opt = (int)value;
            E.drop_graph();
            E.eager_runs = 0;
        } else if (k == "pipeline") {
            if (value != 0 && value != 1) throw std::invalid_argument("nnp_set_option: pipeline is 0 or 1");
            HIPCHK(hipStreamSynchronize(E.stream));
            E.pipeline = (int)value;
            E.drop_graph();
            E.eager_runs = 0;
        } else if (k == "moments") {
            if (value < 0 || value > 3) throw std::invalid_argument("nnp_set_option: moments is 0..3");
            HIPCHK(hipStreamSynchronize(E.stream));
            E.moments = (int)value;
            E.tune_cache.clear();
            E.drop_graph();
            E.eager_runs = 0;
        } else if (k == "net_bf16_terms") {
            if (value != 1 && value != 3) throw std::invalid_argument("nnp_set_option: net_bf16_terms is 1 or 3");
            HIPCHK(hipStreamSynchronize(E.stream));
            E.net_bf16_terms = (int)value;
            E.drop_graph();
            E.eager_runs = 0;
        } else if (k == "force_waves") {
            if (value != 1 && value != 2) throw std::invalid_argument("nnp_set_option: force_waves is 1 or 2");
            HIPCHK(hipStreamSynchronize(E.stream));
            E.force_waves = (int)value;
            E.drop_graph();
            E.eager_runs = 0;
        } else if (k == "inject_alloc_failure") {
            if (value < 0) throw std::invalid_argument("nnp_set_option: inject_alloc_failure must be >= 0");
            g_alloc_fail_in.store(value);
        } else if (k == "reaction_rows") {
            if (value != 0 && value != 1) throw std::invalid_argument("nnp_set_option: reaction_rows is 0 or 1");
            HIPCHK(hipStreamSynchronize(E.stream));
            E.react_rows = (int)value;
            E.assemble_mode = (int)value;
            E.drop_graph();
            E.eager_runs = 0;
        } else if (k == "train_chunk_elems") {
            if (value < (1 << 16)) throw std::invalid_argument("nnp_set_option: train_chunk_elems must be >= 65536");
            E.train_chunk_elems = value;
        } else if (k == "train_cache_atoms") {
            if (value < 0) throw std::invalid_argument("nnp_set_option: train_cache_atoms must be >= 0");
            E.train_cache_atoms = value;
        } else if (k == "train_one_pass") {
            if (value != 0 && value != 1) throw std::invalid_argument("nnp_set_option: train_one_pass is 0 or 1");
            E.train_one_pass = (int)value;
        } else if (k == "fused_blocks_per_cu") {
            if (value < 0 || value > 64) throw std::invalid_argument("nnp_set_option: fused_blocks_per_cu in 0..64");
            E.fused_blocks_per_cu = (int)value;
            E.drop_graph();
            E.eager_runs = 0;
        } else {
            throw std::invalid_argument("nnp_set_option: unknown option " + k);
        }
        return NNP_OK;
    });
}

int nnp_info(nnp_model* m, int64_t* out, int nout) {
    return guarded([&]() {
        if (!m || !out) throw std::invalid_argument("nnp_info: invalid argument");
        const Engine& E = *m->eng;
        int64_t tuned = E.fused_blocks_per_cu;
        if (!E.tune_cache.empty()) tuned = E.tune_cache.back().second;
        const int64_t v[] = {E.maxp, E.max_entries, E.ntot, E.nstruc, E.shape, E.pb.plan.ne_max, E.pb.plan.nc_max,
                             E.pb.plan.ldg, E.pipeline, tuned, E.num_cus, E.last_mom, E.assemble_mode,
                             E.react_missed};
        for (int k = 0; k < nout && k < (int)(sizeof(v) / sizeof(v[0])); ++k) out[k] = v[k];
        return (int)(sizeof(v) / sizeof(v[0]));
    });
}

}  // extern "C"
