// Engine state: the device-resident model and the reusable batch workspace.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/nnp_capi.h"
#include "device.hpp"
#include "model.hpp"

namespace atomnn {

void set_error(const std::string& s);
const char* last_error();

struct hip_failure : std::runtime_error {
    using std::runtime_error::runtime_error;
};

#define HIPCHK(expr)                                                                                          \
    do {                                                                                                      \
        hipError_t e_ = (expr);                                                                               \
        if (e_ != hipSuccess)                                                                                 \
            throw hip_failure(std::string("HIP error in ") + #expr + ": " + hipGetErrorString(e_));            \
    } while (0)

// A device allocation that only grows; or, bound, a fixed slice of memory owned elsewhere, which
// refuses (throws) instead of allocating when a larger size is asked for.
struct DevMem {
    void* p = nullptr;
    size_t cap = 0;
    bool bound = false;
    void reserve(size_t bytes);
    void release();
    void bind(void* ptr, size_t bytes) {
        p = ptr;
        cap = bytes;
        bound = true;
    }
};

// The structures of a call (a view: no copy of the caller's array).
struct Strucs {
    const nnp_structure* p = nullptr;
    size_t n = 0;
    Strucs(const std::vector<nnp_structure>& v) : p(v.data()), n(v.size()) {}
    Strucs(const nnp_structure* q, size_t m) : p(q), n(m) {}
    size_t size() const { return n; }
    const nnp_structure& operator[](size_t k) const { return p[k]; }
    const nnp_structure* begin() const { return p; }
    const nnp_structure* end() const { return p + n; }
};

// A non-owning device pointer into a larger allocation.
struct DevView {
    void* p = nullptr;
};

// Grow-only page-locked host buffer (hipHostMalloc) for single-copy staging of uploads and
// downloads.
struct HostPinned {
    void* p = nullptr;
    size_t cap = 0;
    void reserve(size_t bytes);
    void release();
};

struct NeighborArgs {
    const DevStruc* S;
    const int* atom_struc;
    const double* pos;
    const int* type;
    int ntot;
    int nstruc = 0;  // structures in the batch (1: every atom in S[0]; 0: not stated)
    double rc2;
    int maxp;
    int* cnt;
    int* nbr_j;
    int* nbr_meta;
    double4* nbr_r;
    int* maxcnt;
    int* work = nullptr;  // work counter of the fused persistent kernel (cleared with maxcnt)
    const int* cell_centres;
    int ncell;
    const int* pair_centres;
    int npair;
    int nbins;
    int* atom_bin = nullptr;
    int* bin_count = nullptr;
    int* bin_start = nullptr;
    int* bin_fill = nullptr;
    int* bin_atoms = nullptr;
    int all_images;
    int cell_natoms_max = 0;  // atoms of the largest cell-list structure (LDS bitmap size)
    bool count_max = true;    // fold the largest neighbour count into maxcnt after the search
};

DevStruc make_struc(const nnp_structure& s, double rc, int atom0);

class Engine {
  public:
    explicit Engine(HostModel&& hm);
    // A second workspace (own stream and batch buffers) on the same device-resident model.
    explicit Engine(const Engine& model_owner, int tag);
    ~Engine();
    Engine(const Engine&) = delete;
    Engine& operator=(const Engine&) = delete;

    void upload(Strucs ss);
    bool lds_fits(int lm) const;
    void build_neighbors(bool all_images);
    void build_adjoint_table();
    template <typename T>
    void descriptors();
    void networks(bool mixed, bool vac);
    template <typename T>
    void forces();
    template <typename T>
    void force_terms();
    // forces of the pass from the per-entry terms; with tail, the end-of-pass reductions run in
    // the same launch
    template <typename T>
    void assemble(const EnergyTail* tail = nullptr);
    void assemble_prec(bool mixed, const EnergyTail* tail = nullptr);
    // Reactions of the force assembly: the gather from the neighbours' rows (mode 0) reports in
    // maxcnt[1] how many reactions it missed; receiver rows (mode 1) report their longest row in
    // maxcnt[2].  Given the counters read back after a pass with forces, switch to receiver rows
    // or lengthen them; true when the pass's forces must be assembled again.
    bool reactions_redo(const int* counters);
    // option: 0 = gather, receiver rows once a miss is seen in the batch; 1 = receiver rows always
    int react_rows = 0;
    int assemble_mode = 0;  // the current batch's mode (reset to react_rows by upload)
    int rcap = 0;           // receiver-row length (mode 1)
    int64_t react_missed = 0;  // reactions the last mode-0 pass missed (0 in the usual case)
    DevMem d_rcount, d_rlist;
    int counters[4] = {0, 0, 0, 0};  // maxcnt[0..3] as of the last download
    // Bound workspace (nnp_compute_noalloc): every batch buffer is a slice of a fixed array
    // (bind_scratch), uploads and downloads copy straight between the caller's / pre-reserved host
    // memory and those slices, and no call allocates device or pinned memory.
    bool bound = false;
    int maxp_cap = 0;  // largest row stride the bound slices hold
    void make_bound(int max_atoms);
    void bind_scratch(void* base, size_t bytes, const nnp_structure& s);
    std::vector<int> sp_fill;
    void run(Strucs ss, uint32_t flags);
    void fetch(Strucs ss, uint32_t flags, nnp_result* res);

    // device-resident batches: upload once, enqueue passes, fetch results
    void resident_load(Strucs ss, uint32_t flags);
    void resident_run(uint32_t flags);
    void resident_fetch(uint32_t flags, nnp_result* res);
    EnergyTail energy_tail(bool count_max);
    void energies(bool count_max = false);

    // stage timing with HIP events on the engine's stream
    enum Stage { ST_NEIGH = 0, ST_SYMM, ST_NET, ST_FORCE, ST_ASSEMBLE, ST_ENERGY, ST_COUNT };
    bool profiling = false;
    hipEvent_t ev[ST_COUNT][2] = {};
    double stage_ms[ST_COUNT] = {};
    int64_t stage_calls[ST_COUNT] = {};
    void stage_begin(int s);
    void stage_end(int s, int launches = 1);
    void profile_collect();
    int pending[ST_COUNT] = {};   // launches the stage's pending events bracket
    int profile_repeat = 1;       // force-term launches per profiled pass
    // pipeline: 0 = separate kernels (descriptors, MFMA networks, force terms),
    //           1 = one persistent fused descriptor + network + force kernel
    int pipeline = 0;
    int fused_blocks_per_cu = 0;  // 0: chosen by the autotuner
    // moments: bit 0 = descriptor angular sums, bit 1 = angular force terms by the moment
    // expansion, where the model qualifies (integer zeta <= 4, |lambda| <= 1, one shared
    // exponent table and separable distance factor) and the LDS fits; clear bits use the
    // triplet loops.  Default: both (O(P) per centre against the triplet loops' O(P^2)).
    int moments = 3;
    // wavefronts per centre of the moment-only force kernel (every species takes the moments on
    // a fast-path model): 1, or 2 = exponents, radial functions and moment entries split between
    // two wavefronts.  Other models take the general force kernel.
    int force_waves = 1;
    // mixed precision networks: BF16 MFMA products with operands split into hi + lo BF16 parts
    // (3 products, about 16 mantissa bits per operand) or single BF16 products (1)
    int net_bf16_terms = 3;
    // energy training: one forward + backward pass over single-structure tiles when it fits (1),
    // or always a forward pass and a second pass with the loss seeds (0)
    int train_one_pass = 1;
    // force training: atoms whose neighbour rows and descriptors are kept between calls
    int64_t train_cache_atoms = (int64_t)1 << 20;
    // training: gradient-row elements (doubles) held at a time
    int64_t train_chunk_elems = (int64_t)1 << 28;
    int last_mom = 0;  // the moment bits the last pass used
    int mom_flag(size_t base_lds, bool fp64, int bits);
    int num_cus = 0;
    DevMem d_work;
    // Autotuned persistent grid of the fused kernel per (shape, precision, log2 batch atoms).
    // Serving lanes read the cache but do not tune (concurrent lanes would time each other).
    struct TuneKey {
        int shape, fp64, ntot_log2;
        bool operator==(const TuneKey& o) const {
            return shape == o.shape && fp64 == o.fp64 && ntot_log2 == o.ntot_log2;
        }
    };
    bool may_tune = true;
    // defer the neighbour-count check of run() to the download in fetch() (one host sync per
    // batch instead of two; LDS rows then span the full row stride)
    bool defer_count = false;
    std::vector<std::pair<TuneKey, int>> tune_cache;  // key -> blocks per CU
    std::vector<std::pair<int, double>> last_tune;    // candidates and their times (ms) of the last tuning
    int fused_blocks(bool mixed, uint32_t flags);
    // Network launch configuration per log2 tile count of FP64 batches (tuned once each), and the
    // candidates and times (ms) of the last tuning
    std::vector<std::pair<int, NetLaunch>> net_cache;
    int net_launch_opt = -1;  // option net_launch: -1 tuned, 0 weights through L2, k weights in LDS on k blocks per CU
    std::vector<std::pair<NetLaunch, double>> last_net_tune;
    NetLaunch net_launch(bool mixed, bool vac);
    void fused(bool mixed, uint32_t flags, int nblocks);
    void net_dims(int& maxw, int& dersum) const;
    // captured resident pass (replayed with hipGraphLaunch)
    bool use_graph = true;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    uint32_t graph_flags = 0;
    int graph_maxp = 0;
    int graph_lmaxp = 0;
    int eager_runs = 0;
    void drop_graph();
    void enqueue_pass(uint32_t flags, bool rows_ready = false);
    NeighborArgs neighbor_args(bool all_images) const;
    std::vector<int> h_natoms;  // atoms per structure of the current batch
    bool resident_valid = false;
    std::vector<double> h_E;
    DevMem d_E, d_sstart;

    HostModel host;
    PlanBuffers pb;
    int shape = 0;
    hipStream_t stream = nullptr;
    int maxp = 0;   // neighbour row stride (entries per centre, multiple of 32)
    int lmaxp = 0;  // entries the per-centre kernels stage in LDS (multiple of 8, <= maxp)
    int max_entries = 0;
    int ntot = 0, nstruc = 0, nbins = 0;
    int cell_natoms_max = 0;
    bool all_images_ = false;

    // model (owned by the first engine; workspaces alias it)
    DevMem d_plan, d_rfe, d_rfr, d_w64, d_w32, d_adjc, d_adji, d_angc, d_angi;
    bool owns_model = true;
    // bytes of device memory held by the batch buffers
    size_t workspace_bytes() const;
    // batch
    // the per-batch inputs are one allocation filled by one copy from pinned staging
    DevMem d_up;
    HostPinned h_up, h_down;
    DevView d_pos, d_type, d_astruc, d_strucs, d_centres, d_spatoms;
    DevMem d_cnt, d_nbrj, d_meta, d_nbrr, d_maxcnt;
    DevMem d_abin, d_bcount, d_bstart, d_bfill;
    DevMem d_G, d_dEdG, d_X, d_eatom, d_F;
    // per-centre moment tables [S][NE][35] the moment-only descriptor kernel leaves for the
    // force-term kernel of the same rows (which then skips its own moment sums); mom_fresh is the
    // element size of the tables that match the current rows, 0 when there are none
    DevMem d_mom;
    int mom_fresh = 0;
    // copy E (per structure, optional), E_i, F (optional) and the neighbour maximum to the host
    // in one staged download (packed on the device into d_out, one copy); returns the neighbour
    // maximum.  keep_host = false leaves E_i and F in the staging buffer (down_ea, down_f)
    // instead of copying them into h_eatom / h_F.
    int download(bool energies, bool forces, bool keep_host = true);
    DevMem d_out;
    const double* down_ea = nullptr;
    const double* down_f = nullptr;
    // Synchronous calls (nnp_compute): stage the rows at the previous call's realised neighbour
    // maximum (with head room) instead of reading this call's maximum back before the
    // descriptors; the maximum comes back with the results and a pass that exceeded the
    // staged rows is run again with the count read first.
    bool speculate_rows = false;
    bool speculated = false;
    int spec_entries = 0;
    // the neighbour maximum of the pass is still to be taken (deferred or speculated rows): the
    // force assembly's launch carries the count blocks, else the download launches them
    bool count_pending = false;
    // the per-atom tables of the last upload (species, structure of each atom, centre lists,
    // species-sorted atoms) are on the device for these atom counts and search modes
    bool up_tables = false;
    std::vector<int> up_natoms;
    std::vector<char> up_modes, up_modes_prev;

    std::vector<double> h_pos, h_eatom, h_F;
    std::vector<int> h_type, h_astruc, h_cell_centres, h_pair_centres, h_centres, h_spatoms, sp_off;
    std::vector<DevStruc> h_strucs;
};

}  // namespace atomnn
