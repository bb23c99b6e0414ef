// Host-side model of an AtomNN Behler-Parrinello potential and its device plan.
//
// HostModel is what the "ann" text file holds (grammar of AtomNN nnp/nnp.cpp:244-308);
// DevPlan is the flattened, kernel-ready form uploaded once per model: per (centre
// species, neighbour species) radial blocks, per (centre species, pair block) angular
// blocks factorised into a grid of distinct radial exponents x distinct angular terms,
// and zero-padded network weights for the MFMA GEMM kernels.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace atomnn {

enum CutKind : int { CUT_STEP = 1, CUT_COS = 2, CUT_TANH = 3 };

// Radial functions phi(x), x = -eta (r - rs) (basis_radial.cpp:24-32, 302-359).
enum RadialKind : int { RAD_GAUSSIAN = 0, RAD_TANH, RAD_SOFTPLUS, RAD_LOGCOSH, RAD_SWISH, RAD_MISH };
// Angular distance factors of x = ietap r^alpha (basis_angular.cpp:24-30, 415-441).  The first
// three factorise over the two legs of a triplet (separable); IPOWS and SECHS do not.
enum AngularKind : int { ANG_GAUSS = 0, ANG_IPOWP, ANG_SECHP, ANG_IPOWS, ANG_SECHS };
inline bool angular_separable(int k) { return k <= ANG_SECHP; }
// Activations, numbered as the reference's Neuron::Type (ml/nn.hpp:84-103).
enum NeuronKind : int {
    NEU_LINEAR = 1, NEU_SIGMOID, NEU_TANH, NEU_ISRU, NEU_ARCTAN, NEU_SOFTSIGN, NEU_RELU, NEU_SOFTPLUS, NEU_ELU,
    NEU_GELU, NEU_SWISH, NEU_MISH, NEU_TANHRE, NEU_LOGCOSH, NEU_IQUAD, NEU_IFABS, NEU_SINSIG
};

// Limits of the device plan.  The parser rejects potentials beyond them with NNP_EUNSUPPORTED.
constexpr int kMaxSpecies = 8;
constexpr int kMaxE = 8;       // distinct radial exponents per angular block
constexpr int kMaxC = 8;       // distinct (zeta, lambda) terms per angular block
constexpr int kMaxLayers = 8;  // network layers (hidden + output)
constexpr int kMaxWidth = 256; // network layer width

struct unsupported_error : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct io_error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

struct CutoffSpec {
    int kind = 0;
    double rc = 0, rci = 0, pirci = 0, norm = 0;
};

struct RadialSpec {
    CutoffSpec cut;
    int phi = RAD_GAUSSIAN;
    std::vector<double> rs, eta;
};

struct AngularSpec {
    CutoffSpec cut;
    int phi = ANG_GAUSS;
    int alpha = 0;
    std::vector<double> eta, ietap, zeta;
    std::vector<int> lambda;
};

struct NetSpec {
    int nin = 0;
    std::vector<int> nodes;  // hidden..., output (=1)
    int neuron = NEU_TANH;  // hidden layers; the output layer is linear
    std::vector<double> inpw, inpb, outw, outb;
    std::vector<std::vector<double>> W, b;  // W[l]: nodes[l] x in(l), row-major
};

struct SpeciesSpec {
    std::string name;
    int hash = 0;
    double evac = 0;
    std::vector<RadialSpec> radial;    // [neighbour species]
    std::vector<AngularSpec> angular;  // [lower-triangular pair block]
    std::vector<int> offR, offA;
    int nR = 0, nA = 0, nG = 0;
    NetSpec net;
};

struct HostModel {
    int T = 0;
    double rc = 0;
    std::vector<SpeciesSpec> sp;
    int index(const std::string& name) const;  // -1 if absent
};

inline int lmat_index(int i, int j) { return (i >= j) ? i * (i + 1) / 2 + j : j * (j + 1) / 2 + i; }
int name_hash(const char* s);

// Parse an ann file; throws std::invalid_argument / io_error / unsupported_error.
HostModel parse_ann(const char* path);

// ---------------------------------------------------------------- device plan (POD)
struct DevRadial {
    int n;       // functions in this block
    int goff;    // descriptor offset
    int fbase;   // first function in the plan's radial function arrays
    int cut;
    int phi;     // RadialKind
    int pad_;
    double rc, rci, pirci, norm;
};

struct DevAngular {
    int goff;    // descriptor offset (nInputR + offsetA)
    int alpha;
    int ne, nc;  // distinct exponents / angular terms
    int cut;
    int phi;     // AngularKind
    double rc, rci, pirci, norm;
    double ietap[kMaxE];
    double zeta[kMaxC];
    double lambda[kMaxC];
    int zint[kMaxC];          // integer zeta (>= 0) or -1
    int fmap[kMaxE * kMaxC];  // grid cell (e, c) -> function index within the block, or -1
};

struct DevNet {
    int L;                     // number of weight layers
    int nin, nin_pad;
    int width[kMaxLayers];     // real output widths
    int wpad[kMaxLayers];      // output widths padded to 16
    int64_t woff[kMaxLayers];  // offsets into the weight buffer (row-major wpad x in_pad)
    int64_t wtoff[kMaxLayers]; // the same weights transposed (in_pad x wpad)
    int64_t boff[kMaxLayers];
    int64_t inpw_off, inpb_off;
    int64_t pack_off;          // contiguous copy [W_0 | W_1 | ... | b_0 | b_1 | ...] (row-major, padded;
                               // element (o, k) of W_l at o in_pad + (k ^ (o & 15)); of the
                               // output layer only row 0)
    int pack_len;              // its length in doubles (even)
    int pw_off[kMaxLayers];    // offsets of W_l and b_l inside the copy
    int pb_off[kMaxLayers];
    int act;                   // NeuronKind of the hidden layers
    int pad_;
    double outw, outb, evac;
};

struct DevPlan {
    int T;
    int ldg;  // max descriptor length over species
    int ne_max, nc_max;
    // moment path per centre species: every angular block has integer zeta in [0, 4], |lambda| <= 1,
    // and all blocks share one exponent table, alpha and cutoff (see symm.hip)
    int mom_ok[kMaxSpecies];
    // every angular block of the centre species shares one exponent table, alpha and cutoff and
    // has a separable distance factor: the entry tables R_e(p), R'_e(p) are computed once
    int ang_shared[kMaxSpecies];
    // the model uses a radial function other than GAUSSIAN, an angular factor other than GAUSS,
    // a non-integer zeta or an activation other than tanh / linear: the kernels' generic
    // instantiations run it
    int generic;
    int nG[kMaxSpecies];
    int nR[kMaxSpecies];
    DevRadial rad[kMaxSpecies][kMaxSpecies];
    DevAngular ang[kMaxSpecies][kMaxSpecies * (kMaxSpecies + 1) / 2];
    DevNet net[kMaxSpecies];
    // adjoint coefficients of the moment path (device arrays, set by the engine): for centre
    // species t and item (block, exponent e < adj_ne, power p) of the w table, kMaxC pairs
    // (dE/dG index, coefficient) with w = sum_c dE/dG[index] coefficient (unused pairs: 0, 0.0)
    const double* adj_coef;
    const int* adj_index;
    int adj_ne;
    // angular descriptors from the moment sums: for centre species t and item (block, exponent
    // e < adj_ne, term c < kMaxC), the descriptor index (-1: none) and five coefficients with
    // G = sum_p coefficient_p X_p over the powers p of the block's moment products
    const double* ang_coef;
    const int* ang_index;
};

struct PlanBuffers {
    DevPlan plan;
    std::vector<double> rfun_eta, rfun_rs;  // radial functions, all blocks
    std::vector<double> weights;            // network weights/biases/scalings (padded)
};

PlanBuffers build_plan(const HostModel& m);

}  // namespace atomnn
