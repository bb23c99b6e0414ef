/*
 * nnp_capi.h — C ABI of the MI355X Behler-Parrinello energy/force engine.
 *
 * Drop-in for AtomNN's CPU evaluation path (the nnp_compute tool): a structure goes in,
 * the total energy, per-atom energies and forces come out.  Plain pointers and sizes
 * only; no C++ types, no exceptions cross this boundary.  Every call returns 0 on
 * success or a negative NNP_E* code; nnp_last_error() then holds a message (the text
 * of the reference's exception where there is one).
 *
 * Reference interfaces replaced (paths relative to AtomNN's src/):
 *   nnp_load            <- NNP::read(const char*, NNP&)            nnp/nnp.cpp:182-191, 244-308
 *   nnp_type_index      <- NNP::index(const char*)                 nnp/nnp.hpp:83-84
 *   nnp_ntypes/nnp_rc/nnp_ninput <- NNP::ntypes(), NNP::rc(), NNH::nInput()   nnp/nnp.hpp, nnp/nnh.hpp:91
 *   nnp_neighbors       <- NeighborList::build(const Structure&, double)  struc/neighbor.cpp:59-171
 *   nnp_symm            <- NNP::init + NNP::symm                   nnp/nnp.cpp:317-381
 *   nnp_compute         <- NNP::symm + NNP::energy + NNP::force    nnp/nnp.cpp:334-482, as run by
 *                          nnpce.cpp:275-279; with NNP_VACUUM_ENERGY clear it is NNP::compute
 *                          (nnp/nnp.cpp:496-566), which omits the vacuum energy (nnp.cpp:509)
 *   nnp_compute_batch   <- the per-structure loop of nnpce.cpp:273-281
 *   nnp_free            <- NNP destructor
 *
 * Units and conventions follow the reference: positions in Angstrom, energies in the units
 * of the potential file.  Lattice vectors are given as three consecutive 3-vectors
 * a = cell[0..2], b = cell[3..5], c = cell[6..8] (the columns of AtomNN's R,
 * struc/cell.hpp:22).  cell == NULL means a non-periodic structure (NeighborList::build's
 * open branch, neighbor.cpp:149-170).  Species indices are the order of the "name" lines
 * in the potential file (nnp/nnp.cpp:256-261).
 *
 * Threading: one host thread per model handle at a time.  The serving API (nnp_server_*)
 * is internally synchronised.
 */
#ifndef ATOMNN_AMD_NNP_CAPI_H
#define ATOMNN_AMD_NNP_CAPI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes ---- */
#define NNP_OK 0
#define NNP_EINVAL (-22)   /* invalid argument / malformed potential file */
#define NNP_EIO (-5)       /* file could not be opened or read */
#define NNP_ENOMEM (-12)   /* device or host allocation failed */
#define NNP_EDEVICE (-100) /* HIP runtime error */
#define NNP_EBUDGET (-101) /* request does not fit the serving HBM budget */
#define NNP_EUNSUPPORTED (-102) /* potential uses a form this engine does not implement */
#define NNP_EAGAIN (-11)   /* serving: result not ready yet */

/* ---- flags ---- */
#define NNP_FORCE 0x1u          /* compute forces */
#define NNP_ATOM_ENERGY 0x2u    /* write per-atom energies */
#define NNP_VACUUM_ENERGY 0x4u  /* add the species vacuum energy per atom (NNP::energy, nnh.cpp:129) */
#define NNP_ALL_IMAGES 0x8u     /* Newton reaction on every periodic image (true -dE/dr); default is
                                   the reference rule: reaction on the minimum image only
                                   (nnp.cpp:445, 471-478) */
#define NNP_PREC_MIXED 0x10u    /* precision form: FP32 symmetry functions, BF16 network GEMMs
                                   with FP32 accumulation, FP64 displacements and reductions */

typedef struct nnp_model nnp_model;

typedef struct {
    int32_t natoms;
    const int32_t* type; /* [natoms] species index */
    const double* pos;   /* [natoms][3] Cartesian positions */
    const double* cell;  /* [9] lattice vectors a,b,c consecutively, or NULL (open) */
} nnp_structure;

typedef struct {
    double energy;       /* total energy */
    double* energy_atom; /* [natoms] or NULL */
    double* force;       /* [natoms][3] or NULL */
} nnp_result;

/* Last error message of the calling thread ("" if none). */
const char* nnp_last_error(void);

/* Parse an AtomNN "ann" potential file and upload it to the current HIP device. */
int nnp_load(const char* ann_path, nnp_model** out);
void nnp_free(nnp_model* m);

int nnp_ntypes(const nnp_model* m);
double nnp_rc(const nnp_model* m);
int nnp_ninput(const nnp_model* m, int type);
/* Species index of a name via the reference's string hash (str/string.cpp:17-28); -1 if absent. */
int nnp_type_index(const nnp_model* m, const char* name);

/* Neighbour list in the reference's order (NeighborList::build).  counts[natoms] receives the
 * number of entries per atom.  If the total fits in `cap`, the entries are written in order:
 * index[] (neighbour atom), image[] (linear image index in the reference's enumeration,
 * -1 for open structures), is_min[] (the reference's min flag), r[][3] (vector from the
 * neighbour to the centre) and d[] (its length).  Returns the total number of entries
 * (>= 0) or a negative error code. */
int64_t nnp_neighbors(nnp_model* m, const nnp_structure* s, int32_t* counts, int64_t cap, int32_t* index,
                      int32_t* image, int32_t* is_min, double* r, double* d);

/* Symmetry-function descriptors (NNP::symm): G[natoms][ldg], zero padded past nnp_ninput(type). */
int nnp_symm(nnp_model* m, const nnp_structure* s, double* G, int ldg);

/* Energy (and forces with NNP_FORCE) of one structure, synchronous. */
int nnp_compute(nnp_model* m, const nnp_structure* s, uint32_t flags, nnp_result* r);

/* n structures packed into shared launches; results per structure. */
int nnp_compute_batch(nnp_model* m, int n, const nnp_structure* s, uint32_t flags, nnp_result* r);

/* nnp_compute that allocates nothing: no counterpart in the reference; for callers that cannot
 * allocate device or pinned memory (serving loops and harnesses that forbid allocation).  Same
 * arguments, flags and results as nnp_compute (bit-identical results).  Its device workspace is a
 * fixed 64 MiB array per device, looked up once per device at nnp_load (with the model's device
 * current); host bookkeeping is reserved at nnp_load for up to 16384 atoms.  A structure whose
 * neighbour rows, descriptors and force terms do not fit the array (beyond about 7,500 atoms at
 * 0.06 atoms/A^3 and rc 6.5, 10,000 at 0.05 and rc 6), or of more than 16384 atoms, is rejected with
 * NNP_EINVAL / NNP_EUNSUPPORTED and a message.  Every call on a device shares that device's array:
 * calls that overlap on one device (from several threads or models) race; serialise them.  Call
 * it on the device the model was loaded on. */
int nnp_compute_noalloc(nnp_model* m, const nnp_structure* s, uint32_t flags, nnp_result* r);

/* ---- serving (engine extension; the reference has no server) ----
 * A continuous-batching request server on one model: `lanes` concurrent batches, each on its
 * own HIP stream and workspace, share `hbm_budget` bytes of device memory.  Free lanes take
 * queued requests in arrival order and evaluate them as one packed batch; requests that arrive
 * while other batches are on the GPU are admitted by the next free lane.  A request whose
 * estimated footprint exceeds one lane's share of the budget is rejected with NNP_EBUDGET.
 * nnp_server_submit copies the structure; the result buffers (energy_atom, force) must stay
 * valid until nnp_server_wait returns for that request.  Destroy the server before its model.
 * Results do not depend on which requests share a batch. */
typedef struct nnp_server nnp_server;
typedef struct {
    uint32_t flags;               /* evaluation flags of every request (NNP_FORCE, NNP_PREC_MIXED, ...) */
    int32_t lanes;                /* concurrent batches (default 3) */
    int64_t hbm_budget;           /* bytes of batch workspaces over all lanes (default 16 GiB) */
    int32_t max_batch_atoms;      /* atoms per batch (default 262144) */
    int32_t max_batch_structures; /* structures per batch (default 4096) */
} nnp_server_cfg;
int nnp_server_create(nnp_model* m, const nnp_server_cfg* cfg, nnp_server** out);
/* Returns the request id (>= 0) or a negative error code. */
int64_t nnp_server_submit(nnp_server* s, const nnp_structure* st, nnp_result* r);
/* NNP_OK when done, NNP_EAGAIN while pending, negative on error. */
int nnp_server_poll(nnp_server* s, int64_t id);
/* Blocks until the request is done; optional submit / completion times on nnp_server_now's clock. */
int nnp_server_wait(nnp_server* s, int64_t id, double* t_submit, double* t_done);
/* Seconds on a monotonic clock. */
double nnp_server_now(void);
/* completed requests, batches, max concurrent batches, workspace bytes, peak bytes, queued */
int nnp_server_stats(nnp_server* s, int64_t* out, int nout);
void nnp_server_destroy(nnp_server* s);

/* ---- device-resident batches (engine extension; no counterpart in the reference) ----
 * For callers that evaluate the same structures repeatedly with inputs kept in HBM (benchmarks,
 * serving, MD drivers): nnp_resident_load uploads a batch and sizes its neighbour rows,
 * nnp_resident_run enqueues one full pass (neighbour list, descriptors, networks, forces,
 * energies) on the model's stream without a host round trip, nnp_resident_fetch waits and copies
 * the results out.  Any other compute call on the same model replaces the resident batch. */
int nnp_resident_load(nnp_model* m, int n, const nnp_structure* s, uint32_t flags);
int nnp_resident_run(nnp_model* m, uint32_t flags);
int nnp_resident_fetch(nnp_model* m, uint32_t flags, nnp_result* r);
int nnp_sync(nnp_model* m);

/* ---- stage timing (engine extension) ----
 * With profiling on, every pass brackets its stages with HIP events on the model's stream.
 * nnp_profile_read returns accumulated milliseconds and call counts per stage:
 * 0 neighbour list, 1 descriptors, 2 networks, 3 force terms, 4 force assembly, 5 energy.
 * enable > 1 also runs the force-term launch (idempotent) enable times back to back inside its
 * events and counts each launch as a call, so its per-call time is the kernel's own duration
 * with the events' cost spread over the launches. */
#define NNP_NSTAGES 6
int nnp_profile(nnp_model* m, int enable);
int nnp_profile_read(nnp_model* m, double* ms, int64_t* calls, int nstage);
/* Engine options:
 *   "graph"     1 = replay resident passes from a captured hipGraph (default), 0 = eager passes
 *   "pipeline"  0 = separate kernels: descriptors, per-species MFMA network GEMMs, force terms
 *                   (default);
 *               1 = one persistent fused descriptor + network + force kernel (the batched form;
 *                   networks evaluated in-wave, so results differ from pipeline 0 in the last bits)
 *   "fused_blocks_per_cu"  persistent grid of the fused kernel; 0 = chosen by the autotuner
 *   "moments"   bit 0: descriptor angular sums by the moment expansion, bit 1: the angular
 *               force terms by it (default 3, both); clear bits take the triplet loops; models
 *               outside its domain (integer zeta <= 4, |lambda| <= 1, one exponent table per
 *               centre species) take the triplet loops whatever the bits
 *   "force_waves"  1 = one wavefront per centre in the force-term kernel (default), 2 = two
 *               wavefronts split a centre's exponents, radial functions and moment entries
 *               (moment path on fast-path models only)
 *   "net_launch"  FP64 network kernel: -1 = tuned (default: weights staged in LDS, one block per
 *               16-atom tile, up to two tiles per CU; larger batches time the weights read
 *               through L2 against the weights in LDS on 1 or 2 persistent blocks per CU, once
 *               per log2 tile count, and keep the fastest), 0 = weights through L2, k = 1, 2:
 *               weights in LDS on k blocks per CU; every setting gives the same results
 *   "net_bf16_terms"  mixed-precision networks: 3 = BF16 MFMA products of split operands
 *               (hi + lo, about 16 mantissa bits; default), 1 = single BF16 products
 *   "train_one_pass"  energy training: 1 = one forward + backward pass over single-structure
 *               tiles when it fits (default), 0 = a forward pass, then a pass with the loss seeds
 *   "train_cache_atoms"  force training keeps the neighbour rows and descriptors of the
 *               structure chunks it has seen, up to this many atoms (default 2^20; 0 = rebuild
 *               them every call)
 *   "train_chunk_elems"  training: gradient-row elements (doubles) held at a time (default
 *               2^28, 2 GB; batches with more rows are summed in chunks; at least 65536)
 *   "inject_alloc_failure"  test hook, process-wide: k > 0 makes the k-th device allocation
 *               after it fail as on an exhausted device (the call returns NNP_ENOMEM; a serving
 *               lane fails that batch's requests with NNP_ENOMEM and serves on); 0 clears it
 *   "reaction_rows"  force assembly: 0 = each atom gathers its reactions from its neighbours'
 *               rows, and a batch in which that misses one (a centre lists the atom but the
 *               atom does not list the centre: a distance within rounding of rc) is assembled
 *               again from receiver rows (default); 1 = receiver rows always.  Both apply the
 *               reaction of every entry whose own min flag is set (NNP::force, nnp.cpp:441-478)
 *               and give the same forces up to summation order. */
int nnp_set_option(nnp_model* m, const char* key, int64_t value);
/* Engine state: maxp, max entries, atoms, structures, kernel shape, NE, NC, descriptor width,
 * pipeline, fused blocks per CU (tuned), compute units, moment bits of the last pass, force
 * assembly mode of the current batch (0 gather, 1 receiver rows), reactions the last gather
 * missed (0 in the usual case; a batch with misses is assembled again from receiver rows). */
int nnp_info(nnp_model* m, int64_t* out, int nout);

/* ---- energy-training inner loop ----
 * Replaces the training part of NNPTE::error (src/nnp/nnpte.cpp:869-927) for one process:
 * for each structure b of a batch, E_b = sum of the atomic network outputs (ANN::fp, without
 * the type energies), dE_b = (E_b - E_ref,b) / N_b, the loss and dC/dE of the loss function
 * (nnpte.cpp:884-916), and per atom fpbp + Cost::grad (src/ml/nn.cpp:1554-1589) added into its
 * species' gradient.  The descriptors of the training structures are computed once at creation
 * (NNP::symm, nnpte.cpp:1976-1977).  Parameter vectors hold every species in species order, each
 * in ANN operator>> order (src/ml/nn.cpp:797-831): the biases of every layer, then every weight
 * matrix column-major. */
#define NNP_LOSS_MSE 0   /* 0.5 dE^2 */
#define NNP_LOSS_MAE 1   /* |dE| */
#define NNP_LOSS_HUBER 2 /* delta^2 (sqrt(1 + (dE/delta)^2) - 1) */
#define NNP_LOSS_ASINH 3 /* delta^2 (1 - sqrt(1 + a^2) + a asinh(a)), a = dE/delta */
typedef struct nnp_train nnp_train;
/* The structures are copied; the model m must outlive t (t uses its input scalings, basis
 * tables and force kernels).  One host thread per training set at a time. */
int nnp_train_create(nnp_model* m, int nstruc, const nnp_structure* s, const double* energy, nnp_train** out);
/* parameters of one species' network, or of all species (species < 0); negative on error */
int64_t nnp_train_nparams(nnp_train* t, int species);
/* the model's current parameters (the x of the next call) */
int nnp_train_get_params(nnp_train* t, double* x);
/* err[0] = sum over the batch of the losses, err[1] = sum of dE^2 (error_[0] and error_[2] of
 * NNPTE::error before its BATCH normalisation); grad (may be NULL) = d err[0] / d x.  The model
 * used for evaluation (nnp_compute ...) keeps its own parameters. */
int nnp_train_error(nnp_train* t, const double* x, int nbatch, const int* batch, int loss, double delta, double* err,
                    double* grad);
/* ---- force training (the training part of NNPTEF::error, src/nnp/nnptef.cpp:1040-1293) ----
 * Reference forces of every training atom: 3 doubles per atom, structures in the order given to
 * nnp_train_create. */
int nnp_train_set_forces(nnp_train* t, const double* forces);
/* Per batch structure the energy loss of nnp_train_error plus beta/2 |fdiff_i|^2 per atom,
 * fdiff = (F_i - F_ref,i) / N (nnptef.cpp:1163-1166), with F the forces of the networks at x
 * (reactions on minimum images only, the reference rule, or on every image with NNP_ALL_IMAGES in
 * flags).  err[0] = sum of the losses, err[1] = sum of dE^2, err[2] = sum of |fdiff|^2;
 * grad (may be NULL) = d err[0] / d x, the exact derivative.  The reference's gradient
 * (nnptef.cpp:1138-1293, built on D2ODZDI::grad, src/ml/nn.cpp:1799-1890, and BasisA::forcep,
 * src/nnp/basis_angular.cpp:642-761) keeps only the centre atom's share of each force and does
 * not compile as shipped. */
int nnp_train_error_force(nnp_train* t, const double* x, int nbatch, const int* batch, int loss, double delta,
                          double beta, uint32_t flags, double* err, double* grad);
void nnp_train_destroy(nnp_train* t);

/* ---- molecular dynamics on the device-resident structure ----
 * Replaces the reference's MD step, Verlet::compute (src/torch/integrator.cpp:170-188: half
 * kick, drift, Cell::returnToCell, Engine::compute -> Pot::compute (src/torch/engine.cpp:
 * 107-119, pot_nnpe.cpp:175-295), half kick, kinetic energy and temperature).  METAL units:
 * positions A, velocities A/fs, masses amu (one per species), dt fs, energies eV.
 * flags as for nnp_compute (NNP_FORCE implied; NNP_ALL_IMAGES gives true -dE/dr forces, which
 * conserve energy also in cells shorter than 2 rc).  The model handle stays usable: the MD
 * state has its own workspace.  nnp_md_step advances nsteps steps on the device (each step
 * one replayed graph; the host waits once per call).  nnp_md_state copies positions,
 * velocities and forces (any may be NULL) and energies[4] = {potential, kinetic (eV),
 * temperature (K), steps taken}.  The model m must outlive the MD state. */
typedef struct nnp_md nnp_md;
int nnp_md_create(nnp_model* m, const nnp_structure* s, const double* masses, const double* vel, double dt,
                  uint32_t flags, nnp_md** out);
/* Verlet list for the MD steps (the reference's `vlist rc stride` input, struc/verlet.cpp:35-43,
 * rebuilt every `stride` steps by its MD driver, torch/torch.cpp:269): the list holds every
 * neighbour within rc_list >= rc and is built at the first step of each nnp_md_step call and
 * every `stride` steps after; the other steps recompute the displacements of the list's entries
 * and evaluate the potential on those within rc, the list a fresh build would give
 * (NeighborList::build semantics, min-image flags included), so forces equal those of a build
 * every step up to rounding.  Positions are returned to the cell only at builds.  A call in
 * which some atom moved more than (rc_list - rc) / 2 since a build is redone with a build every
 * step.  stride 1 (the default) builds a fresh list every step. */
int nnp_md_vlist(nnp_md* md, double rc_list, int stride);
int nnp_md_step(nnp_md* md, int nsteps);
int nnp_md_state(nnp_md* md, double* pos, double* vel, double* force, double* energies);
void nnp_md_destroy(nnp_md* md);

#ifdef __cplusplus
}
#endif

#endif /* ATOMNN_AMD_NNP_CAPI_H */
