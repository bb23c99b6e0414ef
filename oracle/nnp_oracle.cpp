// =====================================================================================
// ORACLE — TEST INFRASTRUCTURE ONLY.
//
// A CPU, FP64, dependency-free C++17 restatement of AtomNN's energy/force evaluation
// path (nnp_compute).  It is the checker for the HIP engine in atomnn_amd/ and the
// CPU baseline timed beside it; it is never linked into, called by, or shipped with
// the product library.  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg load it.
//
// Every function cites the reference file:line whose arithmetic it restates
// (paths relative to the AtomNN source tree, src/...).  The reference itself cannot
// be compiled here: it needs the un-vendored Eigen 3.4 headers (src/Makefile:4).
//
// Parity status of each part (how it is pinned; see tests/test_oracle_*.py):
//   cell_diff ............ pinned: closed forms on cubic/sheared cells, lattice-translation
//                          invariance, agreement with an exhaustive image search.
//   neighbor_build ....... pinned: exact set equality against an independent numpy
//                          enumeration; ordering rule (j asc, image a,b,c over
//                          norm-sorted lattice vectors) pinned by code reading;
//                          no-duplicate and supercell invariants.
//   cutoff / G2 / G4 ..... pinned: closed forms (dimer, symmetric trimer), integral
//                          of fc, finite differences of every derivative.
//   basis variants ....... pinned: radial TANH/SOFTPLUS/LOGCOSH/SWISH/MISH against dimer closed
//                          forms, angular IPOWP/IPOWS/SECHP/SECHS against an asymmetric trimer,
//                          forces of each by finite differences (tests/test_oracle_variants.py).
//   neurons .............. pinned: all 17 activations against closed forms on a one-unit
//                          network, dO/dI and forces by finite differences.
//   descriptor layout .... pinned: hand-built 3-species cases (LMat block order).
//   ANN fp / DODZ ........ pinned: closed-form one-hot networks, finite differences.
//   energy / force ....... pinned: F = -dE/dr by finite differences on L > 2rc cells,
//                          sum F = 0, rotation/translation invariance, supercell
//                          extensivity, NaCl symmetry equivalence, a hand-built
//                          small cell for the min-image-only reaction rule.
//   ann parser ........... pinned: write->read round trip.
//   Summation order inside Eigen GEMVs and glibc last-ulp behaviour: parity unpinned
//   at bit level (only a running reference could pin it); the 1e-10 relative bar
//   absorbs it.
// =====================================================================================
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

namespace orc {

// math/const.hpp:8-9
static const double ZERO = 1e-6;
static const double PI = 3.141592653589793;

// ---------------------------------------------------------------- small vector algebra
struct V3 {
    double x, y, z;
};
static inline V3 v3(double x, double y, double z) { return V3{x, y, z}; }
static inline V3 add(const V3& a, const V3& b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline V3 sub(const V3& a, const V3& b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline V3 scl(double s, const V3& a) { return v3(s * a.x, s * a.y, s * a.z); }
static inline double dot(const V3& a, const V3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline double sqn(const V3& a) { return dot(a, a); }
static inline double nrm(const V3& a) { return std::sqrt(sqn(a)); }
static inline V3 cross(const V3& a, const V3& b) {
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

// math/special.hpp:46-47 — mod(n, lo, hi) = fmod(fmod(n-lo, w)+w, w)+lo
static inline double mod_range(double n, double lo, double hi) {
    const double w = hi - lo;
    return std::fmod(std::fmod(n - lo, w) + w, w) + lo;
}

// math/special.cpp:111-120 — integer power by repeated squaring
static double powint(double x, int n) {
    if (n == 0) return 1.0;
    if (x == 0.0) return 0.0;
    int nn = (n > 0) ? n : -n;
    double ww = x, yy = 1.0;
    for (; nn != 0; nn >>= 1, ww *= ww)
        if (nn & 1) yy *= ww;
    return (n > 0) ? yy : 1.0 / yy;
}

// math/special.cpp:131-135 — the reference's log(1+x): 2 atanh(y), y = x/(x+2), by its odd
// series truncated after the y^9 term (not the library log1p; error ~2e-7 at x = 1)
static double logp1(double x) {
    const double y = x / (x + 2.0);
    const double y2 = y * y;
    return 2.0 * y * (1.0 + y2 * (1.0 / 3.0 + y2 * (1.0 / 5.0 + y2 * (1.0 / 7.0 + y2 * 1.0 / 9.0))));
}

// math/special.cpp:71-79 — sech from one exponential of -|x|
static double sech(double x) {
    const double e = std::exp(-std::fabs(x));
    return 2.0 * e / (1.0 + e * e);
}

// math/const.hpp:15 (the reference's truncated ln 2)
static const double LOG2 = 0.693147180559945;

// str/string.cpp:17-28 — polynomial string hash (int arithmetic, mod 2^31-1)
static int str_hash(const char* s) {
    const int p = 31, m = 2147483647;
    int val = 0, pw = 1;
    while (*s) {
        // same int arithmetic as the reference; names are short so nothing overflows
        val = (int)(((long long)val + (long long)(*s - 'a' + 1) * pw) % m);
        pw = (int)(((long long)pw * p) % m);
        ++s;
    }
    return val;
}

// ---------------------------------------------------------------- cell
// struc/cell.cpp:45-70 (init) and 103-128 (diff).  R holds lattice vectors as columns.
struct Cell {
    bool periodic = false;
    double R[3][3] = {{0}};     // R[row][col]
    double RInv[3][3] = {{0}};
    double dMax2 = 0;
    V3 shifts[6];
    double fro = 0;             // Frobenius norm ||R|| (neighbor.cpp:67, 132)
    V3 col(int c) const { return v3(R[0][c], R[1][c], R[2][c]); }

    // lattice vectors given as a[0..2], b[3..5], c[6..8] (columns of R)
    void init(const double* lv) {
        for (int c = 0; c < 3; ++c)
            for (int r = 0; r < 3; ++r) R[r][c] = lv[3 * c + r];
        fro = 0;
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) fro += R[r][c] * R[r][c];
        fro = std::sqrt(fro);
        periodic = fro > 0;
        if (!periodic) return;
        // 3x3 inverse by cofactors
        const double (&m)[3][3] = R;
        const double c00 = m[1][1] * m[2][2] - m[1][2] * m[2][1];
        const double c01 = m[1][2] * m[2][0] - m[1][0] * m[2][2];
        const double c02 = m[1][0] * m[2][1] - m[1][1] * m[2][0];
        const double det = m[0][0] * c00 + m[0][1] * c01 + m[0][2] * c02;
        if (det == 0) throw std::invalid_argument("Cell::init: singular lattice");
        const double id = 1.0 / det;
        RInv[0][0] = c00 * id;
        RInv[1][0] = c01 * id;
        RInv[2][0] = c02 * id;
        RInv[0][1] = (m[0][2] * m[2][1] - m[0][1] * m[2][2]) * id;
        RInv[1][1] = (m[0][0] * m[2][2] - m[0][2] * m[2][0]) * id;
        RInv[2][1] = (m[0][1] * m[2][0] - m[0][0] * m[2][1]) * id;
        RInv[0][2] = (m[0][1] * m[1][2] - m[0][2] * m[1][1]) * id;
        RInv[1][2] = (m[0][2] * m[1][0] - m[0][0] * m[1][2]) * id;
        RInv[2][2] = (m[0][0] * m[1][1] - m[0][1] * m[1][0]) * id;
        // cell.cpp:56-62: half of the shortest lattice vector
        double dmax = nrm(col(0));
        for (int i = 1; i < 3; ++i) dmax = std::min(dmax, nrm(col(i)));
        dmax *= 0.5;
        dMax2 = dmax * dmax;
        // cell.cpp:64-69
        shifts[0] = col(0);
        shifts[1] = scl(-1.0, col(0));
        shifts[2] = col(1);
        shifts[3] = scl(-1.0, col(1));
        shifts[4] = col(2);
        shifts[5] = scl(-1.0, col(2));
    }

    // cell.cpp:103-128 — minimum-image difference v1 - v2
    V3 diff(const V3& v1, const V3& v2) const {
        const V3 d = sub(v1, v2);
        double f[3];
        for (int r = 0; r < 3; ++r) f[r] = RInv[r][0] * d.x + RInv[r][1] * d.y + RInv[r][2] * d.z;
        for (int r = 0; r < 3; ++r) f[r] = mod_range(f[r], -0.5, 0.5);
        V3 out = v3(R[0][0] * f[0] + R[0][1] * f[1] + R[0][2] * f[2],
                    R[1][0] * f[0] + R[1][1] * f[1] + R[1][2] * f[2],
                    R[2][0] * f[0] + R[2][1] * f[1] + R[2][2] * f[2]);
        if (sqn(out) > dMax2) {
            V3 rmin = out;
            double dmin2 = sqn(rmin);
            for (int i = 0; i < 6; ++i) {
                const V3 t = add(out, shifts[i]);
                const double d2 = sqn(t);
                if (d2 < dmin2) {
                    dmin2 = d2;
                    rmin = t;
                }
            }
            out = rmin;
        }
        return out;
    }
};

// ---------------------------------------------------------------- neighbour list
// struc/neighbor.hpp:28-54 — r points from the neighbour to the centre atom
struct Neighbor {
    V3 r;
    double dr;
    int type;
    int index;
    int image;  // linear index n into the image table (order of neighbor.cpp:89-95); -1 non-periodic
    bool min;
};

struct NeighborList {
    std::vector<std::vector<Neighbor>> neigh;
    int shell[3] = {0, 0, 0};
    int perm[3] = {0, 1, 2};  // sorted lattice vector q = column perm[q]
};

// struc/neighbor.cpp:59-171
static void neighbor_build(const Cell& cell, int n, const V3* pos, const int* type, double rc, NeighborList& nl) {
    nl.neigh.assign(n, {});
    const double rc2 = rc * rc;
    if (cell.periodic) {
        // neighbor.cpp:71-74: lattice vectors sorted by descending norm (stable for 3 items)
        V3 Rv[3] = {cell.col(0), cell.col(1), cell.col(2)};
        int perm[3] = {0, 1, 2};
        for (int i = 1; i < 3; ++i) {  // insertion sort, comparator |a| > |b|
            V3 v = Rv[i];
            int p = perm[i];
            int j = i;
            while (j > 0 && nrm(v) > nrm(Rv[j - 1])) {
                Rv[j] = Rv[j - 1];
                perm[j] = perm[j - 1];
                --j;
            }
            Rv[j] = v;
            perm[j] = p;
        }
        for (int q = 0; q < 3; ++q) nl.perm[q] = perm[q];
        // neighbor.cpp:77-81
        int shell[3];
        shell[0] = (int)std::floor(2.0 * rc / nrm(Rv[0]));
        shell[1] = (int)std::floor(2.0 * rc / (nrm(cross(Rv[0], Rv[1])) / nrm(Rv[0])));
        shell[2] = (int)std::floor(2.0 * rc / std::fabs(dot(Rv[2], cross(Rv[0], Rv[1])) / (nrm(Rv[0]) * nrm(Rv[1]))));
        for (int q = 0; q < 3; ++q) nl.shell[q] = shell[q];
        // neighbor.cpp:87-95: a-major, then b, then c
        std::vector<V3> img;
        for (int a = -shell[0]; a <= shell[0]; ++a)
            for (int b = -shell[1]; b <= shell[1]; ++b)
                for (int c = -shell[2]; c <= shell[2]; ++c)
                    img.push_back(add(add(scl((double)a, Rv[0]), scl((double)b, Rv[1])), scl((double)c, Rv[2])));
        const int nimg = (int)img.size();
        for (int i = 0; i < n; ++i) {
            auto& li = nl.neigh[i];
            for (int j = 0; j < n; ++j) {
                std::vector<Neighbor> nj;
                const V3 base = cell.diff(pos[i], pos[j]);
                for (int m = 0; m < nimg; ++m) {
                    const V3 r = sub(base, img[m]);
                    const double d2 = sqn(r);
                    // neighbor.cpp:112-128: inclusive cutoff; self images need d2 > ZERO
                    const bool keep = (j != i) ? (d2 <= rc2) : (ZERO < d2 && d2 <= rc2);
                    if (keep) nj.push_back(Neighbor{r, std::sqrt(d2), type[j], j, m, false});
                }
                // neighbor.cpp:131-141: strict minimum, first occurrence, below ||R||
                if (j != i && !nj.empty()) {
                    double mn = cell.fro;
                    int loc = -1;
                    for (size_t q = 0; q < nj.size(); ++q)
                        if (nj[q].dr < mn) {
                            mn = nj[q].dr;
                            loc = (int)q;
                        }
                    if (loc >= 0) nj[loc].min = true;
                }
                li.insert(li.end(), nj.begin(), nj.end());
            }
        }
    } else {
        // neighbor.cpp:149-170: no images, min never set
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                const V3 r = sub(pos[i], pos[j]);
                const double d2 = sqn(r);
                if (ZERO < d2 && d2 <= rc2) nl.neigh[i].push_back(Neighbor{r, std::sqrt(d2), type[j], j, -1, false});
            }
    }
}

// ---------------------------------------------------------------- basis
enum CutName { CUT_UNKNOWN = 0, CUT_STEP, CUT_COS, CUT_TANH };  // nnp/cutoff.hpp:56-61
enum CutNorm { NORM_UNKNOWN = 0, NORM_UNIT, NORM_VOL };          // nnp/cutoff.hpp:29-33
// basis_radial.cpp:24-32 and basis_angular.cpp:24-30: function names
enum PhiR { PHIR_UNKNOWN = 0, PHIR_GAUSSIAN, PHIR_TANH, PHIR_SOFTPLUS, PHIR_LOGCOSH, PHIR_SWISH, PHIR_MISH };
enum PhiA { PHIA_UNKNOWN = 0, PHIA_GAUSS, PHIA_IPOWP, PHIA_IPOWS, PHIA_SECHP, PHIA_SECHS };

static PhiR phir_read(const std::string& s) {
    if (s == "GAUSSIAN") return PHIR_GAUSSIAN;
    if (s == "TANH") return PHIR_TANH;
    if (s == "SOFTPLUS") return PHIR_SOFTPLUS;
    if (s == "LOGCOSH") return PHIR_LOGCOSH;
    if (s == "SWISH") return PHIR_SWISH;
    if (s == "MISH") return PHIR_MISH;
    return PHIR_UNKNOWN;
}
static PhiA phia_read(const std::string& s) {
    if (s == "GAUSS") return PHIA_GAUSS;
    if (s == "IPOWP") return PHIA_IPOWP;
    if (s == "IPOWS") return PHIA_IPOWS;
    if (s == "SECHP") return PHIA_SECHP;
    if (s == "SECHS") return PHIA_SECHS;
    return PHIA_UNKNOWN;
}

static CutName cut_read(const std::string& s) {  // cutoff.cpp:42-47
    if (s == "STEP") return CUT_STEP;
    if (s == "COS") return CUT_COS;
    if (s == "TANH") return CUT_TANH;
    return CUT_UNKNOWN;
}
static CutNorm norm_read(const std::string& s) {  // cutoff.cpp:15-19
    if (s == "UNIT") return NORM_UNIT;
    if (s == "VOL") return NORM_VOL;
    return NORM_UNKNOWN;
}

struct Basis {
    CutName cut = CUT_UNKNOWN;
    CutNorm cnorm = NORM_UNKNOWN;
    double rc = 0, rci = 0, pirci = 0, norm = 0;

    // basis.cpp:33-46
    void init(double rc_, CutName c, CutNorm nm) {
        if (rc_ <= 0) throw std::invalid_argument("Basis: invalid cutoff radius");
        cut = c;
        cnorm = nm;
        rc = rc_;
        rci = 1.0 / rc;
        pirci = PI / rc;
        // basis.cpp:160-168
        if (nm == NORM_UNIT) norm = 1.0;
        else if (nm == NORM_VOL) norm = 1.0 / (0.5 * 4.0 / 3.0 * (PI * PI - 6.0) / PI * rc * rc * rc);
        else throw std::invalid_argument("Basis::norm: invalid normalization scheme");
        if (cut == CUT_UNKNOWN) throw std::invalid_argument("Basis: invalid cutoff type");
    }
    // basis.cpp:80-98
    double cut_func(double dr) const {
        switch (cut) {
            case CUT_STEP: return (dr > rc) ? 0.0 : 1.0;
            case CUT_COS: return (dr > rc) ? 0 : 0.5 * (std::cos(dr * pirci) + 1.0);
            case CUT_TANH: {
                const double f = (dr > rc) ? 0 : std::tanh(1.0 - dr * rci);
                return f * f * f;
            }
            default: throw std::invalid_argument("Basis::cut_func: invalid cutoff type");
        }
    }
    // basis.cpp:100-120
    double cut_grad(double dr) const {
        switch (cut) {
            case CUT_STEP: return 0.0;
            case CUT_COS: return (dr > rc) ? 0 : -0.5 * pirci * std::sin(dr * pirci);
            case CUT_TANH: {
                if (dr < rc) {
                    const double f = std::tanh(1.0 - dr * rci);
                    return -3.0 * f * f * (1.0 - f * f) * rci;
                }
                return 0;
            }
            default: throw std::invalid_argument("Basis::cut_grad: invalid cutoff type");
        }
    }
};

struct BasisR : Basis {
    PhiR phi = PHIR_UNKNOWN;
    std::vector<double> rs, eta;
    int size() const { return (int)eta.size(); }

    // Radial function of x = -eta (r - rs) and its derivative with respect to x.
    // basis_radial.cpp:302-359 (values) and 371-438 (derivatives, there multiplied by -eta).
    // SWISH and MISH at x = 0 exactly: the reference leaves the previous value in place
    // (basis_radial.cpp:341-357 assign only for x > 0 and x < 0); here the function's value 0.
    static void phi_eval(PhiR f, double x, double& v, double& dv) {
        switch (f) {
            case PHIR_TANH: {  // 0.5 (tanh(x) + 1)
                const double t = std::tanh(x);
                v = 0.5 * (t + 1.0);
                dv = 0.5 * (1.0 - t * t);
            } break;
            case PHIR_SOFTPLUS: {  // log(1 + e^x), stable split at 0
                if (x > 0.0) {
                    const double e = std::exp(-x);
                    v = x + logp1(e);
                    dv = 1.0 / (1.0 + e);
                } else {
                    const double e = std::exp(x);
                    v = logp1(e);
                    dv = e / (1.0 + e);
                }
            } break;
            case PHIR_LOGCOSH: {  // x + log(1 + e^-2x)/2 (x > 0), log(1 + e^2x)/2 (x <= 0)
                if (x > 0.0) {
                    const double e = std::exp(-2.0 * x);
                    v = x + 0.5 * logp1(e);
                    dv = 1.0 / (1.0 + e);
                } else {
                    const double e = std::exp(2.0 * x);
                    v = 0.5 * logp1(e);
                    dv = e / (1.0 + e);
                }
            } break;
            case PHIR_SWISH: {  // x sigmoid(x)
                if (x > 0.0) {
                    const double e = std::exp(-x), den = 1.0 / (1.0 + e);
                    v = x * den;
                    dv = (1.0 + e * (1.0 + x)) * den * den;
                } else if (x < 0.0) {
                    const double e = std::exp(x), den = 1.0 / (1.0 + e);
                    v = x * e * den;
                    dv = (1.0 + x + e) * e * den * den;
                } else {
                    v = 0.0;
                    dv = 0.5;
                }
            } break;
            case PHIR_MISH: {  // x tanh(softplus(x))
                if (x > 0.0) {
                    const double e = std::exp(-x), den = 1.0 / (2.0 * e * (e + 1.0) + 1.0);
                    v = x * (2.0 * e + 1.0) * den;
                    dv = (1.0 + e * (4.0 + e * (4.0 * x + 6.0 + e * 4.0 * (x + 1.0)))) * den * den;
                } else if (x < 0.0) {
                    const double e = std::exp(x), den = 1.0 / (e * (e + 2.0) + 2.0);
                    v = x * e * (e + 2.0) * den;
                    dv = (e * (e * (e + 4.0) + 4.0 * x + 6.0) + 4.0 * (x + 1.0)) * den * den * e;
                } else {
                    v = 0.0;
                    dv = 0.6;
                }
            } break;
            default: throw std::invalid_argument("BasisR::symm(double): Invalid symmetry function.");
        }
    }

    // basis_radial.cpp:302-359
    void symm(double dr, double* out) const {
        const double cutv = cut_func(dr) * norm;
        if (phi == PHIR_GAUSSIAN) {
            for (int i = 0; i < size(); ++i) out[i] = std::exp(-eta[i] * (dr - rs[i]) * (dr - rs[i])) * cutv;
            return;
        }
        for (int i = 0; i < size(); ++i) {
            double v, dv;
            phi_eval(phi, -eta[i] * (dr - rs[i]), v, dv);
            out[i] = v * cutv;
        }
    }
    // basis_radial.cpp:371-443: amp = -norm sum_i dEdG_i d/dr[phi_i(r) fc(r)]
    double force(double dr, const double* dEdG) const {
        double amp = 0;
        const double cutv = cut_func(dr), gcut = cut_grad(dr);
        if (phi == PHIR_GAUSSIAN) {
            for (int i = 0; i < size(); ++i)
                amp -= dEdG[i] * std::exp(-eta[i] * (dr - rs[i]) * (dr - rs[i])) *
                       (-2.0 * eta[i] * (dr - rs[i]) * cutv + gcut);
            return amp * norm;
        }
        for (int i = 0; i < size(); ++i) {
            double v, dv;
            phi_eval(phi, -eta[i] * (dr - rs[i]), v, dv);
            amp -= dEdG[i] * (-eta[i] * dv * cutv + v * gcut);
        }
        return amp * norm;
    }
};

struct BasisA : Basis {
    PhiA phi = PHIA_UNKNOWN;
    int alpha = 0;
    std::vector<double> eta, ietap, zeta;
    std::vector<int> lambda;
    int size() const { return (int)eta.size(); }
    // basis_angular.cpp:168-173 — the stored eta is a length; ietap = (1/eta)^alpha
    void init_ietap() {
        ietap.resize(eta.size());
        for (int i = 0; i < size(); ++i) ietap[i] = powint(1.0 / eta[i], alpha);
    }
    // Distance factor of function i for xj = ietap r_ij^alpha, xk = ietap r_ik^alpha, and the
    // factor tj with d(factor)/d(xj) = -tj factor (tk likewise):
    //   GAUSS exp(-(xj+xk)): t = 1;  IPOWP 1/((1+xj)(1+xk)): tj = 1/(1+xj);
    //   IPOWS 1/(1+xj+xk): t = the factor;  SECHP sech(xj) sech(xk): tj = tanh(xj);
    //   SECHS sech(xj+xk): t = tanh(xj+xk).  tanh is taken as sqrt(1 - sech^2).
    // basis_angular.cpp:415-441 (values), 478-630 (the derivative factors).
    void dist_factor(double xj, double xk, double& f, double& tj, double& tk) const {
        switch (phi) {
            case PHIA_GAUSS:
                f = std::exp(-(xj + xk));
                tj = tk = 1.0;
                break;
            case PHIA_IPOWP: {
                const double dj = 1.0 / (1.0 + xj), dk = 1.0 / (1.0 + xk);
                f = dj * dk;
                tj = dj;
                tk = dk;
            } break;
            case PHIA_IPOWS:
                f = 1.0 / (1.0 + (xj + xk));
                tj = tk = f;
                break;
            case PHIA_SECHP: {
                const double sj = sech(xj), sk = sech(xk);
                f = sj * sk;
                tj = std::sqrt(1.0 - sj * sj);
                tk = std::sqrt(1.0 - sk * sk);
            } break;
            case PHIA_SECHS: {
                const double sv = sech(xj + xk);
                f = sv;
                tj = tk = std::sqrt(1.0 - sv * sv);
            } break;
            default: throw std::invalid_argument("BasisA::symm(double): Invalid symmetry function.");
        }
    }
    // basis_angular.cpp:406-441; no fc(r_jk), factor 2^-zeta
    void symm(double cs, const double dr[3], double* out) const {
        const double c0 = cut_func(dr[0]), c1 = cut_func(dr[1]);
        const double cprod = c0 * c1 * norm;
        const double rijp = powint(dr[0], alpha), rikp = powint(dr[1], alpha);
        for (int i = 0; i < size(); ++i) {
            double f, tj, tk;
            dist_factor(ietap[i] * rijp, ietap[i] * rikp, f, tj, tk);
            out[i] = std::pow(std::fabs(0.5 * (1.0 + lambda[i] * cs)), zeta[i]) * cprod * f;
        }
    }
    // basis_angular.cpp:456-640
    void force(double& phiv, double* etav, double cs, const double dr[3], const double* dEdG) const {
        const double c[2] = {cut_func(dr[0]), cut_func(dr[1])};
        const double g[2] = {cut_grad(dr[0]), cut_grad(dr[1])};
        phiv = 0;
        etav[0] = etav[1] = etav[2] = 0;
        const double cprod = c[0] * c[1];
        const double rijp = powint(dr[0], alpha), rikp = powint(dr[1], alpha);
        const double dij = alpha * rijp / dr[0] * c[0];
        const double dik = alpha * rikp / dr[1] * c[1];
        for (int i = 0; i < size(); ++i) {
            double f, tj, tk;
            dist_factor(ietap[i] * rijp, ietap[i] * rikp, f, tj, tk);
            const double cw = std::fabs(0.5 * (1.0 + lambda[i] * cs));
            const double gangle = std::pow(cw, zeta[i] - 1.0) * dEdG[i] * f;
            const double fangle = cw * gangle;
            phiv -= 0.5 * zeta[i] * lambda[i] * cprod * gangle;
            etav[0] -= fangle * (-dij * ietap[i] * tj + g[0]) * c[1];
            etav[1] -= fangle * (-dik * ietap[i] * tk + g[1]) * c[0];
        }
        phiv *= norm;
        etav[0] *= norm;
        etav[1] *= norm;
        etav[2] *= norm;
    }
};

// ---------------------------------------------------------------- network
// ml/nn.hpp:84-103 and nn.cpp:115-134: activation names
enum Neuron {
    NEURON_UNKNOWN = 0, NEURON_LINEAR, NEURON_SIGMOID, NEURON_TANH, NEURON_ISRU, NEURON_ARCTAN, NEURON_SOFTSIGN,
    NEURON_RELU, NEURON_SOFTPLUS, NEURON_ELU, NEURON_GELU, NEURON_SWISH, NEURON_MISH, NEURON_TANHRE, NEURON_LOGCOSH,
    NEURON_IQUAD, NEURON_IFABS, NEURON_SINSIG
};

static Neuron neuron_read(const std::string& s) {
    static const char* names[] = {"LINEAR", "SIGMOID", "TANH", "ISRU", "ARCTAN", "SOFTSIGN", "RELU", "SOFTPLUS", "ELU",
                                  "GELU", "SWISH", "MISH", "TANHRE", "LOGCOSH", "IQUAD", "IFABS", "SINSIG"};
    for (int k = 0; k < 17; ++k)
        if (s == names[k]) return (Neuron)(k + 1);
    return NEURON_UNKNOWN;
}

// Activation a(z) and derivative d(z) of the hidden layers (AFFPBP, ml/nn.cpp:144-766).
static void activate(Neuron n, double z, double& a, double& d) {
    switch (n) {
        case NEURON_LINEAR: a = z; d = 1.0; return;                       // 144-148
        case NEURON_SIGMOID: {                                            // 168-183
            const double e = std::exp(-std::fabs(z)), fr = 1.0 / (1.0 + e);
            a = z >= 0 ? fr : e * fr;
            d = e * fr * fr;
        } return;
        case NEURON_TANH: a = std::tanh(z); d = 1.0 - a * a; return;      // 209-216
        case NEURON_ISRU: {                                               // 233-240
            const double r = 1.0 / std::sqrt(1.0 + z * z);
            a = z * r;
            d = r * r * r;
        } return;
        case NEURON_ARCTAN:                                               // 258-264
            a = (2.0 / PI) * std::atan(z);
            d = (2.0 / PI) / (1.0 + z * z);
            return;
        case NEURON_SOFTSIGN: {                                           // 281-288
            const double r = 1.0 / (1.0 + std::fabs(z));
            a = z * r;
            d = r * r;
        } return;
        case NEURON_RELU: a = z > 0.0 ? z : 0.0; d = z > 0.0 ? 1.0 : 0.0; return;  // 309-321
        case NEURON_SOFTPLUS:                                             // 350-366, shifted by -ln 2
            if (z >= 0.0) {
                const double e = std::exp(-z);
                a = z + logp1(e) - LOG2;
                d = 1.0 / (1.0 + e);
            } else {
                const double e = std::exp(z);
                a = logp1(e) - LOG2;
                d = e / (1.0 + e);
            }
            return;
        case NEURON_ELU:                                                  // 397-409
            if (z >= 0.0) {
                a = z;
                d = 1.0;
            } else {
                const double e = std::exp(z);
                a = e - 1.0;
                d = e;
            }
            return;
        case NEURON_GELU:                                                 // 446-469: z sigmoid(1.702 z)
        case NEURON_SWISH: {                                              // 510-525: z sigmoid(z)
            const double X = (n == NEURON_GELU) ? 1.702 : 1.0;
            if (z > 0.0) {
                const double e = std::exp(-X * z), den = 1.0 / (1.0 + e);
                a = z * den;
                d = (1.0 + e * (1.0 + X * z)) * den * den;
            } else {
                const double e = std::exp(X * z), den = 1.0 / (1.0 + e);
                a = z * e * den;
                d = e * (X * z + 1.0 + e) * den * den;
            }
        } return;
        case NEURON_MISH:                                                 // 557-572
            if (z > 0.0) {
                const double e = std::exp(-z), den = 1.0 / (2.0 * e * (e + 1.0) + 1.0);
                a = z * (2.0 * e + 1.0) * den;
                d = (1.0 + e * (4.0 + e * (4.0 * z + 6.0 + e * 4.0 * (z + 1.0)))) * den * den;
            } else {
                const double e = std::exp(z), den = 1.0 / (e * (e + 2.0) + 2.0);
                a = z * e * (e + 2.0) * den;
                d = e * (e * (e * (e + 4.0) + 4.0 * z + 6.0) + 4.0 * (z + 1.0)) * den * den;
            }
            return;
        case NEURON_TANHRE:                                               // 602-614
            if (z > 0) {
                a = z;
                d = 1.0;
            } else {
                a = std::tanh(z);
                d = 1.0 - a * a;
            }
            return;
        case NEURON_LOGCOSH:                                              // 639-652, shifted by -ln 2 / 2
            if (z >= 0.0) {
                const double e = std::exp(-2.0 * z);
                a = z + 0.5 * (logp1(e) - LOG2);
                d = 1.0 / (e + 1.0);
            } else {
                const double e = std::exp(2.0 * z);
                a = 0.5 * (logp1(e) - LOG2);
                d = e / (e + 1.0);
            }
            return;
        case NEURON_IQUAD: {                                              // 678-685
            const double r = 1.0 / std::sqrt(1.0 + z * z);
            a = 0.5 * z * (1.0 + z * r);
            d = 0.5 * (1.0 + z * (z * z + 2.0) * r * r * r);
        } return;
        case NEURON_IFABS:                                                // 706-720
            if (z > 0) {
                const double r = 1.0 / (1.0 + z);
                a = 0.5 * z * (1.0 + z * r);
                d = 1.0 - 0.5 * r * r;
            } else {
                const double r = 1.0 / (1.0 - z);
                a = 0.5 * z * (1.0 + z * r);
                d = 0.5 * r * r;
            }
            return;
        case NEURON_SINSIG: {                                             // 743-754
            const double e = std::exp(-z), den = 1.0 / (1.0 + e);
            const double sv = std::sin(0.5 * PI * den), cv = std::cos(0.5 * PI * den);
            a = z * sv;
            d = sv + 0.5 * PI * z * e * cv * den * den;
        } return;
        default: throw std::invalid_argument("ANN: invalid neuron");
    }
}

// Second derivative of the hidden-layer activations (AFFPBP2, ml/nn.cpp:149-766), as the
// reference writes them where they are the derivative of its AFFPBP d(z).  Where its AFFPBP2
// leaves d2 unset (SOFTSIGN, 289-297) or sets it to 0 (IQUAD 686-694, IFABS 720-735, SINSIG
// 754-765) the true second derivative of its a(z) is used instead.
static double activate_d2(Neuron n, double z) {
    switch (n) {
        case NEURON_LINEAR: case NEURON_RELU: return 0.0;                 // 149-154, 321-334
        case NEURON_SIGMOID: {                                            // 184-201
            const double e = std::exp(-std::fabs(z)), fr = 1.0 / (1.0 + e);
            return (z >= 0 ? -e * (1.0 - e) : -e * (e - 1.0)) * fr * fr * fr;
        }
        case NEURON_TANH: {                                               // 217-225
            const double t = std::tanh(z);
            return -2.0 * t * (1.0 - t * t);
        }
        case NEURON_ISRU: {                                               // 241-250
            const double r = 1.0 / std::sqrt(1.0 + z * z);
            return -3.0 * z * r * r * r * r * r;
        }
        case NEURON_ARCTAN: {                                             // 265-273
            const double fr = 1.0 / (1.0 + z * z);
            return -2.0 * z * (2.0 / PI) * fr * fr;
        }
        case NEURON_SOFTSIGN: {  // d/dz (1 + |z|)^-2
            const double r = 1.0 / (1.0 + std::fabs(z));
            return (z > 0 ? -2.0 : (z < 0 ? 2.0 : 0.0)) * r * r * r;
        }
        case NEURON_SOFTPLUS: {                                           // 366-385
            const double e = std::exp(-std::fabs(z)), fr = 1.0 / (1.0 + e);
            return e * fr * fr;
        }
        case NEURON_ELU: return z >= 0.0 ? 0.0 : std::exp(z);             // 410-424
        case NEURON_GELU:                                                 // 471-497
        case NEURON_SWISH: {                                              // 526-543
            const double X = (n == NEURON_GELU) ? 1.702 : 1.0;
            if (z > 0.0) {
                const double e = std::exp(-X * z), den = 1.0 / (1.0 + e);
                return X * e * (2.0 - X * z + e * (2.0 + X * z)) * den * den * den;
            }
            const double e = std::exp(X * z), den = 1.0 / (1.0 + e);
            return X * e * (2.0 + X * z + e * (2.0 - X * z)) * den * den * den;
        }
        case NEURON_MISH:                                                 // 573-590
            if (z > 0.0) {
                const double e = std::exp(-z), den = 1.0 / (2.0 * e * (e + 1.0) + 1.0);
                return 4.0 * e * e * (e * (2.0 * e * (e * (z + 2.0) + z + 4.0) + 3.0 * (-z + 2.0)) - 2.0 * (z - 1.0)) * den *
                       den * den;
            } else {
                const double e = std::exp(z), den = 1.0 / (e * (e + 2.0) + 2.0);
                return 4.0 * e * (2.0 * (z + 2.0) + e * (e * (-2.0 * e * (z - 1.0) - 3.0 * (z - 2.0)) + 2.0 * (z + 4.0))) *
                       den * den * den;
            }
        case NEURON_TANHRE:                                               // 614-627
            if (z > 0) return 0.0;
            else {
                const double t = std::tanh(z);
                return -2.0 * t * (1.0 - t * t);
            }
        case NEURON_LOGCOSH: {                                            // 653-670
            const double e = std::exp(-2.0 * std::fabs(z)), den = 1.0 / (e + 1.0);
            return 2.0 * e * den * den;
        }
        case NEURON_IQUAD: {  // d/dz of 0.5 (1 + z (z^2 + 2) (1 + z^2)^-3/2) = (2 - z^2) (1 + z^2)^-5/2 / 2
            const double r = 1.0 / std::sqrt(1.0 + z * z);
            return 0.5 * (2.0 - z * z) * r * r * r * r * r;
        }
        case NEURON_IFABS: {  // both branches: (1 + |z|)^-3
            const double r = 1.0 / (1.0 + std::fabs(z));
            return r * r * r;
        }
        case NEURON_SINSIG: {  // a = z sin(pi s / 2), s = 1 / (1 + e^-z), s' = s (1 - s), s'' = s' (1 - 2 s)
            const double e = std::exp(-z), s = 1.0 / (1.0 + e), s1 = e * s * s, s2 = s1 * (1.0 - 2.0 * s);
            const double sv = std::sin(0.5 * PI * s), cv = std::cos(0.5 * PI * s);
            return PI * s1 * cv + z * 0.5 * PI * (cv * s2 - 0.5 * PI * s1 * s1 * sv);
        }
        default: throw std::invalid_argument("ANN: invalid neuron");
    }
}

// ml/nn.cpp:969-1138 (resize), 1191-1233 (fp/fpbp), 1620-1628 (DODZ::grad)
struct ANN {
    int nin = 0;
    std::vector<int> nodes;  // hidden..., output
    Neuron neuron = NEURON_UNKNOWN;
    std::vector<double> inpw, inpb, outw, outb;
    std::vector<std::vector<double>> b;  // b[l][n_l]
    std::vector<std::vector<double>> w;  // w[l] row-major n_l x n_{l-1}
    int nlayer() const { return (int)nodes.size(); }
    int in_size(int l) const { return l == 0 ? nin : nodes[l - 1]; }

    // activation of layer l: hidden layers use `neuron`, the last layer is linear (nn.cpp:1134-1137)
    void act(int l, double z, double& a, double& d) const {
        if (l == nlayer() - 1) {
            a = z;
            d = 1.0;
        } else {
            activate(neuron, z, a, d);
        }
    }
    // forward (fpbp: activations and derivatives); returns out[0]
    double fpbp(const double* inp, std::vector<std::vector<double>>& a, std::vector<std::vector<double>>& d) const {
        std::vector<double> ins(nin);
        for (int i = 0; i < nin; ++i) ins[i] = inpw[i] * (inp[i] + inpb[i]);  // nn.cpp:1218
        a.assign(nlayer(), {});
        d.assign(nlayer(), {});
        const double* prev = ins.data();
        for (int l = 0; l < nlayer(); ++l) {
            const int no = nodes[l], ni = in_size(l);
            a[l].resize(no);
            d[l].resize(no);
            for (int o = 0; o < no; ++o) {
                double z = b[l][o];
                for (int k = 0; k < ni; ++k) z += w[l][(size_t)o * ni + k] * prev[k];
                act(l, z, a[l][o], d[l][o]);
            }
            prev = a[l].data();
        }
        return outb[0] + a.back()[0] * outw[0];  // nn.cpp:1229-1230
    }
    // Parameter vector in ANN operator>> / operator<< order (nn.cpp:797-831): the biases of
    // every layer, then every weight matrix in Eigen's column-major order (w(l)(n,m) at
    // m n_l + n).
    int nparams() const {
        int p = 0;
        for (int l = 0; l < nlayer(); ++l) p += nodes[l] + nodes[l] * in_size(l);
        return p;
    }
    void get_params(double* x) const {
        int c = 0;
        for (int l = 0; l < nlayer(); ++l)
            for (int n = 0; n < nodes[l]; ++n) x[c++] = b[l][n];
        for (int l = 0; l < nlayer(); ++l)
            for (int m = 0; m < in_size(l); ++m)
                for (int n = 0; n < nodes[l]; ++n) x[c++] = w[l][(size_t)n * in_size(l) + m];
    }
    void set_params(const double* x) {
        int c = 0;
        for (int l = 0; l < nlayer(); ++l)
            for (int n = 0; n < nodes[l]; ++n) b[l][n] = x[c++];
        for (int l = 0; l < nlayer(); ++l)
            for (int m = 0; m < in_size(l); ++m)
                for (int n = 0; n < nodes[l]; ++n) w[l][(size_t)n * in_size(l) + m] = x[c++];
    }
    // Cost::grad (nn.cpp:1554-1589) after fpbp on input `inp`: dcdz_L = dcdo outw .* dadz_L,
    // dcdz_{l-1} = (w_l^T dcdz_l) .* dadz_{l-1}; gradient = the dcdz (biases), then
    // dcdz_l[n] a_{l-1}[m] (weights, column-major; a_{-1} = ins).  Adds into g.
    void cost_grad(const double* inp, const std::vector<std::vector<double>>& a,
                   const std::vector<std::vector<double>>& d, double dcdo, double* g) const {
        const int L = nlayer();
        std::vector<std::vector<double>> dz(L);
        dz[L - 1].assign(nodes[L - 1], 0.0);
        for (int i = 0; i < nodes[L - 1]; ++i) dz[L - 1][i] = dcdo * outw[i] * d[L - 1][i];
        for (int l = L - 1; l > 0; --l) {
            const int no = nodes[l], ni = nodes[l - 1];
            dz[l - 1].assign(ni, 0.0);
            for (int k = 0; k < ni; ++k) {
                double s = 0;
                for (int o = 0; o < no; ++o) s += w[l][(size_t)o * ni + k] * dz[l][o];
                dz[l - 1][k] = s * d[l - 1][k];
            }
        }
        std::vector<double> ins(nin);
        for (int i = 0; i < nin; ++i) ins[i] = inpw[i] * (inp[i] + inpb[i]);
        int c = 0;
        for (int l = 0; l < L; ++l)
            for (int n = 0; n < nodes[l]; ++n) g[c++] += dz[l][n];
        for (int l = 0; l < L; ++l) {
            const double* prev = l == 0 ? ins.data() : a[l - 1].data();
            for (int m = 0; m < in_size(l); ++m)
                for (int n = 0; n < nodes[l]; ++n) g[c++] += dz[l][n] * prev[m];
        }
    }
    // fpbp with second derivatives (fpbp2 / AFFPBP2, nn.cpp:149-766): a, d = a', e = a'' per layer
    double fpbp2(const double* inp, std::vector<std::vector<double>>& a, std::vector<std::vector<double>>& d,
                 std::vector<std::vector<double>>& e) const {
        const double out = fpbp(inp, a, d);
        std::vector<double> ins(nin);
        for (int i = 0; i < nin; ++i) ins[i] = inpw[i] * (inp[i] + inpb[i]);
        e.assign(nlayer(), {});
        const double* prev = ins.data();
        for (int l = 0; l < nlayer(); ++l) {
            const int no = nodes[l], ni = in_size(l);
            e[l].assign(no, 0.0);
            if (l < nlayer() - 1)
                for (int o = 0; o < no; ++o) {
                    double z = b[l][o];
                    for (int k = 0; k < ni; ++k) z += w[l][(size_t)o * ni + k] * prev[k];
                    e[l][o] = activate_d2(neuron, z);
                }
            prev = a[l].data();
        }
        return out;
    }
    // Gradient, with respect to the parameters, of dcdo * out + (dout/dinp . v) for one atom
    // (the energy term of the loss and the term through which the forces depend on the network:
    // the forces are linear in dout/dinp, and v is the loss's derivative with respect to it).
    // dout/dinp . v is the network's tangent output for the input tangent inpw .* v:
    //   t_-1 = inpw .* v,  zt_l = w_l t_(l-1),  t_l = a'_l .* zt_l,  tangent out = outw t_L;
    // reverse pass over the primal and tangent graphs (p: adjoint of t, q: adjoint of a):
    //   pz_l = p_l .* a'_l,  qz_l = q_l .* a'_l + p_l .* a''_l .* zt_l,
    //   grad b_l += qz_l,  grad w_l(n, m) += pz_l[n] t_(l-1)[m] + qz_l[n] a_(l-1)[m],
    //   p_(l-1) = w_l^T pz_l,  q_(l-1) = w_l^T qz_l,  with p_L = outw, q_L = dcdo outw.
    // This is the product of the reference's D2ODZDI (nn.cpp:1799-1890, d2out/dparam dinp) with
    // v, plus Cost::grad (nn.cpp:1554-1589).  Adds into g (parameter order of get_params).
    void force_grad(const double* inp, const double* v, const std::vector<std::vector<double>>& a,
                    const std::vector<std::vector<double>>& d, const std::vector<std::vector<double>>& e, double dcdo,
                    double* g) const {
        const int L = nlayer();
        std::vector<double> ins(nin), tin(nin);
        for (int i = 0; i < nin; ++i) {
            ins[i] = inpw[i] * (inp[i] + inpb[i]);
            tin[i] = inpw[i] * v[i];
        }
        std::vector<std::vector<double>> zt(L), t(L);
        const double* tp = tin.data();
        for (int l = 0; l < L; ++l) {
            const int no = nodes[l], ni = in_size(l);
            zt[l].assign(no, 0.0);
            t[l].assign(no, 0.0);
            for (int o = 0; o < no; ++o) {
                double s = 0;
                for (int k = 0; k < ni; ++k) s += w[l][(size_t)o * ni + k] * tp[k];
                zt[l][o] = s;
                t[l][o] = d[l][o] * s;
            }
            tp = t[l].data();
        }
        std::vector<double> p(nodes[L - 1], 0.0), q(nodes[L - 1], 0.0);
        p[0] = outw[0];
        q[0] = dcdo * outw[0];
        std::vector<int> rb(L), rw(L);
        int c = 0;
        for (int l = 0; l < L; ++l) rb[l] = c, c += nodes[l];
        for (int l = 0; l < L; ++l) rw[l] = c, c += nodes[l] * in_size(l);
        for (int l = L - 1; l >= 0; --l) {
            const int no = nodes[l], ni = in_size(l);
            std::vector<double> pz(no), qz(no);
            for (int o = 0; o < no; ++o) {
                pz[o] = p[o] * d[l][o];
                qz[o] = q[o] * d[l][o] + p[o] * e[l][o] * zt[l][o];
                g[rb[l] + o] += qz[o];
            }
            const double* tprev = l == 0 ? tin.data() : t[l - 1].data();
            const double* aprev = l == 0 ? ins.data() : a[l - 1].data();
            for (int m = 0; m < ni; ++m)
                for (int o = 0; o < no; ++o) g[rw[l] + m * no + o] += pz[o] * tprev[m] + qz[o] * aprev[m];
            if (l > 0) {
                p.assign(ni, 0.0);
                q.assign(ni, 0.0);
                for (int m = 0; m < ni; ++m)
                    for (int o = 0; o < no; ++o) {
                        p[m] += w[l][(size_t)o * ni + m] * pz[o];
                        q[m] += w[l][(size_t)o * ni + m] * qz[o];
                    }
            }
        }
    }
    // DODZ::grad (nn.cpp:1620-1628), output row 0: dodi = d out / d inp
    void dodi(const std::vector<std::vector<double>>& d, double* g) const {
        const int L = nlayer();
        std::vector<double> cur(nodes[L - 1], 0.0);
        // dodz_L = diag(outw .* d_L) row 0 (single output)
        cur.assign(nodes[L - 1], 0.0);
        cur[0] = outw[0] * d[L - 1][0];
        for (int l = L - 1; l > 0; --l) {
            const int no = nodes[l], ni = nodes[l - 1];
            std::vector<double> nxt(ni, 0.0);
            for (int k = 0; k < ni; ++k) {
                double s = 0;
                for (int o = 0; o < no; ++o) s += cur[o] * w[l][(size_t)o * ni + k];
                nxt[k] = s * d[l - 1][k];
            }
            cur.swap(nxt);
        }
        for (int k = 0; k < nin; ++k) {
            double s = 0;
            for (int o = 0; o < nodes[0]; ++o) s += cur[o] * w[0][(size_t)o * nin + k];
            g[k] = s * inpw[k];
        }
    }
};

// ---------------------------------------------------------------- model (NNP / NNH)
struct NNH {
    std::string name;
    double evac = 0;              // Type energy (type.cpp:45-47), 0 when absent
    std::vector<BasisR> basisR;   // by neighbour species
    std::vector<BasisA> basisA;   // LMat linear order (lmat.hpp:43-46)
    std::vector<int> offsetR;     // nnh.cpp:103-106
    std::vector<int> offsetA;     // nnh.cpp:115-118, LMat linear
    int nInputR = 0, nInputA = 0, nInput = 0;
    ANN nn;
};

static inline int lmat(int i, int j) { return (i >= j) ? i * (i + 1) / 2 + j : j * (j + 1) / 2 + i; }

struct Model {
    int ntypes = 0;
    double rc = 0;
    std::vector<int> hashes;
    std::vector<NNH> nnh;
    int index(const std::string& name) const {  // nnp.hpp:83; mem/map.hpp:116-123
        const int h = str_hash(name.c_str());
        for (int i = 0; i < ntypes; ++i)
            if (hashes[i] == h) return i;
        throw std::runtime_error("No match for key in map: " + name);
    }
};

// ---------------------------------------------------------------- parser
// Tokenised line reader (str/token.cpp:9-38, delimiters " \r\t\n").
struct LineReader {
    FILE* f;
    std::vector<std::string> tok;
    explicit LineReader(FILE* f_) : f(f_) {}
    void next_line(const char* what) {
        std::string line;
        int ch;
        while ((ch = fgetc(f)) != EOF) {
            line.push_back((char)ch);
            if (ch == '\n') break;
        }
        if (line.empty()) throw std::runtime_error(std::string("unexpected end of file reading ") + what);
        tok.clear();
        size_t p = line.find_first_not_of(" \r\t\n");
        while (p != std::string::npos) {
            size_t e = line.find_first_of(" \r\t\n", p);
            tok.push_back(line.substr(p, e == std::string::npos ? std::string::npos : e - p));
            p = (e == std::string::npos) ? e : line.find_first_not_of(" \r\t\n", e);
        }
    }
    const std::string& at(size_t i, const char* what) const {
        if (i >= tok.size()) throw std::runtime_error(std::string("missing token reading ") + what);
        return tok[i];
    }
};

// type.cpp:32-70 (only name and energy matter on this path)
static void read_type(LineReader& lr, NNH& h) {
    lr.next_line("type");
    for (size_t i = 0; i < lr.tok.size(); ++i) {
        std::string tag = lr.tok[i];
        if (!tag.empty() && tag[0] == '#') break;  // WSC includes '#'
        for (auto& c : tag) c = (char)std::toupper((unsigned char)c);
        if (tag == "NAME") h.name = lr.at(++i, "type name");
        else if (tag == "ENERGY") h.evac = std::atof(lr.at(++i, "type energy").c_str());
        else if (tag == "MASS" || tag == "CHARGE" || tag == "CHI" || tag == "ETA" || tag == "JS" || tag == "C6" ||
                 tag == "RVDW" || tag == "RCOV")
            ++i;
    }
}

// basis_radial.cpp:123-144
static void read_basisR(LineReader& lr, BasisR& b) {
    lr.next_line("BasisR");
    const CutNorm nm = norm_read(lr.at(1, "BasisR norm"));
    const CutName cn = cut_read(lr.at(2, "BasisR cutoff"));
    const double rc = std::atof(lr.at(3, "BasisR rc").c_str());
    const PhiR phi = phir_read(lr.at(4, "BasisR function"));
    const int n = std::atoi(lr.at(5, "BasisR size").c_str());
    b.init(rc, cn, nm);
    if (phi == PHIR_UNKNOWN)  // basis_radial.cpp:70
        throw std::invalid_argument("BasisR(rc,cutoff::Name,cutoff::Norm,int,PhiRN): invalid radial function type");
    if (n < 0) throw std::invalid_argument("Basis::resize(int): invalid number of functions.");
    b.phi = phi;
    b.rs.resize(n);
    b.eta.resize(n);
    for (int i = 0; i < n; ++i) {
        lr.next_line("BasisR parameters");
        b.rs[i] = std::atof(lr.at(0, "rs").c_str());
        b.eta[i] = std::atof(lr.at(1, "eta").c_str());
    }
}

// basis_angular.cpp:113-138
static void read_basisA(LineReader& lr, BasisA& b) {
    lr.next_line("BasisA");
    const CutNorm nm = norm_read(lr.at(1, "BasisA norm"));
    const CutName cn = cut_read(lr.at(2, "BasisA cutoff"));
    const double rc = std::atof(lr.at(3, "BasisA rc").c_str());
    const PhiA phi = phia_read(lr.at(4, "BasisA function"));
    const int n = std::atoi(lr.at(5, "BasisA size").c_str());
    const int alpha = std::atoi(lr.at(6, "BasisA alpha").c_str());
    b.init(rc, cn, nm);
    if (phi == PHIA_UNKNOWN)  // basis_angular.cpp:62
        throw std::invalid_argument("BasisA(rc,cutoff::Name,cutoff::Norm,int,PhiAN): invalid angular function type");
    if (n < 0) throw std::invalid_argument("Basis::resize(int): invalid number of functions.");
    b.phi = phi;
    b.eta.resize(n);
    b.zeta.resize(n);
    b.lambda.resize(n);
    for (int i = 0; i < n; ++i) {
        lr.next_line("BasisA parameters");
        b.eta[i] = std::atof(lr.at(0, "eta").c_str());
        b.zeta[i] = std::atof(lr.at(1, "zeta").c_str());
        b.lambda[i] = std::atoi(lr.at(2, "lambda").c_str());
    }
    b.alpha = alpha;
    b.init_ietap();
}

// ml/nn.cpp:1349-1425
static void read_ann(LineReader& lr, ANN& nn) {
    lr.next_line("nn layout");
    std::vector<int> nodes;
    for (size_t i = 1; i < lr.tok.size(); ++i) nodes.push_back(std::atoi(lr.tok[i].c_str()));
    if (nodes.size() < 2) throw std::invalid_argument("ANN::resize: Invalid number of layers.");
    for (int v : nodes)
        if (v <= 0) throw std::invalid_argument("ANN::resize: Invalid layer size.");
    lr.next_line("neuron");
    nn.neuron = neuron_read(lr.at(1, "neuron"));
    if (nn.neuron == NEURON_UNKNOWN) throw std::invalid_argument("ANN::read(FILE*,ANN&): Invalid neuron.");  // nn.cpp:1376
    nn.nin = nodes[0];
    nn.nodes.assign(nodes.begin() + 1, nodes.end());
    if (nn.nodes.back() != 1) throw std::invalid_argument("ANN: the potential needs exactly one output");
    auto read_vec = [&](const char* what, std::vector<double>& v, int n) {
        lr.next_line(what);
        v.resize(n);
        for (int i = 0; i < n; ++i) v[i] = std::atof(lr.at(1 + i, what).c_str());
    };
    read_vec("inpw", nn.inpw, nn.nin);
    read_vec("outw", nn.outw, 1);
    read_vec("inpb", nn.inpb, nn.nin);
    read_vec("outb", nn.outb, 1);
    const int L = nn.nlayer();
    nn.b.resize(L);
    nn.w.resize(L);
    for (int l = 0; l < L; ++l) read_vec("bias", nn.b[l], nn.nodes[l]);
    for (int l = 0; l < L; ++l) read_vec("weight", nn.w[l], nn.nodes[l] * nn.in_size(l));
}

// nnp.cpp:244-308 with nnh.cpp:95-121
static void read_model(const char* path, Model& m) {
    FILE* f = std::fopen(path, "r");
    if (!f) throw std::runtime_error(std::string("NNP::read(const char*,NNP&): Could not open nnp file: \"") + path + "\"");
    std::unique_ptr<FILE, int (*)(FILE*)> guard(f, &std::fclose);
    LineReader lr(f);
    lr.next_line("header");
    lr.next_line("ntypes");
    const int ntypes = std::atoi(lr.at(1, "ntypes").c_str());
    if (ntypes <= 0) throw std::invalid_argument("NNP::read(FILE*,NNP&): invalid number of types.");
    m.ntypes = ntypes;
    m.nnh.assign(ntypes, NNH());
    m.hashes.resize(ntypes);
    for (int n = 0; n < ntypes; ++n) {
        read_type(lr, m.nnh[n]);
        m.hashes[n] = str_hash(m.nnh[n].name.c_str());
    }
    lr.next_line("rc");
    m.rc = std::atof(lr.at(1, "rc").c_str());
    if (m.rc <= 0) throw std::invalid_argument("NNP::read(FILE*,NNP&): invalid cutoff.");
    for (int n = 0; n < ntypes; ++n) {
        m.nnh[n].basisR.assign(ntypes, BasisR());
        m.nnh[n].basisA.assign(ntypes * (ntypes + 1) / 2, BasisA());
    }
    for (int i = 0; i < ntypes; ++i) {
        lr.next_line("basis");
        const int II = m.index(lr.at(1, "basis species"));
        for (int j = 0; j < ntypes; ++j) {
            lr.next_line("basis_radial");
            const int JJ = m.index(lr.at(1, "basis_radial species"));
            read_basisR(lr, m.nnh[II].basisR[JJ]);
        }
        for (int j = 0; j < ntypes; ++j)
            for (int k = j; k < ntypes; ++k) {
                lr.next_line("basis_angular");
                const int JJ = m.index(lr.at(1, "basis_angular species"));
                const int KK = m.index(lr.at(2, "basis_angular species"));
                read_basisA(lr, m.nnh[II].basisA[lmat(JJ, KK)]);
            }
    }
    for (int i = 0; i < ntypes; ++i) {  // nnh.cpp:95-121
        NNH& h = m.nnh[i];
        h.nInputR = 0;
        for (int t = 0; t < ntypes; ++t) h.nInputR += h.basisR[t].size();
        h.offsetR.assign(ntypes, 0);
        for (int t = 1; t < ntypes; ++t) h.offsetR[t] = h.offsetR[t - 1] + h.basisR[t - 1].size();
        h.nInputA = 0;
        for (auto& b : h.basisA) h.nInputA += b.size();
        h.offsetA.assign(h.basisA.size(), 0);
        for (size_t t = 1; t < h.basisA.size(); ++t) h.offsetA[t] = h.offsetA[t - 1] + h.basisA[t - 1].size();
        h.nInput = h.nInputR + h.nInputA;
    }
    for (int n = 0; n < ntypes; ++n) {
        lr.next_line("nn");
        const int II = m.index(lr.at(1, "nn species"));
        read_ann(lr, m.nnh[II].nn);
        if (m.nnh[II].nn.nin != m.nnh[II].nInput)
            throw std::invalid_argument("NNP::read: network input size does not match the basis of " + m.nnh[II].name);
    }
}

// ---------------------------------------------------------------- evaluation
struct Structure {
    int n = 0;
    Cell cell;
    std::vector<V3> pos;
    std::vector<int> type;
    std::vector<std::vector<double>> symm;
    std::vector<V3> force;
    std::vector<double> eatom;
    double energy = 0;
};

// nnp.cpp:334-381
static void nnp_symm(const Model& m, Structure& s, const NeighborList& nl) {
    s.symm.resize(s.n);
    std::vector<double> tmp(256);
    for (int i = 0; i < s.n; ++i) {
        const NNH& h = m.nnh[s.type[i]];
        auto& G = s.symm[i];
        G.assign(h.nInput, 0.0);
        const auto& L = nl.neigh[i];
        for (size_t j = 0; j < L.size(); ++j) {
            const int JJ = L[j].type;
            const V3& rIJ = L[j].r;
            const double dIJ = L[j].dr;
            {
                const BasisR& br = h.basisR[JJ];
                if ((int)tmp.size() < br.size()) tmp.resize(br.size());
                br.symm(dIJ, tmp.data());
                for (int q = 0; q < br.size(); ++q) G[h.offsetR[JJ] + q] += tmp[q];
            }
            for (size_t k = j + 1; k < L.size(); ++k) {
                const int KK = L[k].type;
                const V3& rIK = L[k].r;
                const double dIK = L[k].dr;
                const double dJK = nrm(sub(rIK, rIJ));
                const double cs = dot(rIJ, rIK) / (dIJ * dIK);
                const double d[3] = {dIJ, dIK, dJK};
                const int blk = lmat(JJ, KK);
                const int off = h.nInputR + h.offsetA[blk];
                const BasisA& ba = h.basisA[blk];
                if ((int)tmp.size() < ba.size()) tmp.resize(ba.size());
                ba.symm(cs, d, tmp.data());
                for (int q = 0; q < ba.size(); ++q) G[off + q] += tmp[q];
            }
        }
    }
}

enum {
    F_FORCE = 1,        // compute forces
    F_ATOM_ENERGY = 2,  // per-atom energies
    F_VACUUM = 4,       // add the species vacuum energy (NNP::energy semantics, nnh.cpp:129)
    F_ALL_IMAGES = 8,   // reaction on every image (true -dE/dr); default is the reference's min-image rule
};

// nnp.cpp:390-402 (energy) and 416-482 (force), run as nnpce does (nnpce.cpp:275-279):
// a forward pass for the energy, then fpbp + DODZ for the forces.
static void nnp_energy_force(const Model& m, Structure& s, const NeighborList& nl, int flags) {
    const bool vac = flags & F_VACUUM;
    s.eatom.assign(s.n, 0.0);
    double E = 0;
    std::vector<std::vector<double>> a, d;
    for (int i = 0; i < s.n; ++i) {  // nnp.cpp:394-399
        const NNH& h = m.nnh[s.type[i]];
        const double out = h.nn.fpbp(s.symm[i].data(), a, d);
        s.eatom[i] = out + (vac ? h.evac : 0.0);
        E += s.eatom[i];
    }
    s.energy = E;
    s.force.assign(s.n, v3(0, 0, 0));
    if (!(flags & F_FORCE)) return;
    const bool all_images = flags & F_ALL_IMAGES;
    std::vector<double> dEdG;
    for (int i = 0; i < s.n; ++i) {
        const NNH& h = m.nnh[s.type[i]];
        h.nn.fpbp(s.symm[i].data(), a, d);  // nnp.cpp:425 (second network pass, as the reference)
        dEdG.assign(h.nInput, 0.0);
        h.nn.dodi(d, dEdG.data());  // nnp.cpp:427-429
        const auto& L = nl.neigh[i];
        for (size_t j = 0; j < L.size(); ++j) {
            const int JJ = L[j].type, jj = L[j].index;
            const bool jmin = all_images ? true : L[j].min;
            const V3& rIJ = L[j].r;
            const double dIJ = L[j].dr, dIJi = 1.0 / dIJ;
            {  // nnp.cpp:441-446
                const double amp = h.basisR[JJ].force(dIJ, dEdG.data() + h.offsetR[JJ]) * dIJi;
                s.force[i] = add(s.force[i], scl(amp, rIJ));
                if (jmin) s.force[jj] = sub(s.force[jj], scl(amp, rIJ));
            }
            for (size_t k = j + 1; k < L.size(); ++k) {  // nnp.cpp:448-479
                const int KK = L[k].type, kk = L[k].index;
                const bool kmin = all_images ? true : L[k].min;
                const V3& rIK = L[k].r;
                const double dIK = L[k].dr, dIKi = 1.0 / dIK;
                const V3 rJK = sub(rIK, rIJ);
                const double dJK = nrm(rJK), dJKi = 1.0 / dJK;
                const int blk = lmat(JJ, KK);
                const int off = h.nInputR + h.offsetA[blk];
                double phi = 0, eta[3] = {0, 0, 0};
                const double cs = dot(rIJ, rIK) * dIJi * dIKi;
                const double dd[3] = {dIJ, dIK, dJK};
                h.basisA[blk].force(phi, eta, cs, dd, dEdG.data() + off);
                s.force[i] = add(s.force[i], scl((phi * (dIKi - cs * dIJi) + eta[0]) * dIJi, rIJ));
                s.force[i] = add(s.force[i], scl((phi * (dIJi - cs * dIKi) + eta[1]) * dIKi, rIK));
                if (jmin) {
                    s.force[jj] = sub(s.force[jj], add(scl((-phi * cs * dIJi + eta[0]) * dIJi, rIJ), scl(phi * dIJi * dIKi, rIK)));
                    s.force[jj] = sub(s.force[jj], scl(eta[2] * dJKi, rJK));
                }
                if (kmin) {
                    s.force[kk] = sub(s.force[kk], add(scl((-phi * cs * dIKi + eta[1]) * dIKi, rIK), scl(phi * dIKi * dIJi, rIJ)));
                    s.force[kk] = add(s.force[kk], scl(eta[2] * dJKi, rJK));
                }
            }
        }
    }
}

// Adjoint of the force loop above with respect to dE/dG: given gF[k] = dL/dF_k for every atom,
// v[i][g] = dL/d(dE_i/dG_g).  The forces are linear in dE/dG, so each basis function's share of a
// pair or triplet force is the force() of a one-hot dE/dG (the reference's per-function force
// amplitudes, BasisA::forcep, basis_angular.cpp:642-761); v collects gF . (that share) over the
// same updates nnp.cpp:441-478 makes, reactions included.
static void force_adjoint(const Model& m, const Structure& s, const NeighborList& nl, bool all_images,
                          const std::vector<V3>& gF, std::vector<std::vector<double>>& v) {
    v.assign(s.n, {});
    std::vector<double> oh;
    for (int i = 0; i < s.n; ++i) {
        const NNH& h = m.nnh[s.type[i]];
        v[i].assign(h.nInput, 0.0);
        const auto& L = nl.neigh[i];
        for (size_t j = 0; j < L.size(); ++j) {
            const int JJ = L[j].type, jj = L[j].index;
            const bool jmin = all_images ? true : L[j].min;
            const V3& rIJ = L[j].r;
            const double dIJ = L[j].dr, dIJi = 1.0 / dIJ;
            {
                const BasisR& br = h.basisR[JJ];
                oh.assign(br.size(), 0.0);
                const double gr = dot(gF[i], rIJ) - (jmin ? dot(gF[jj], rIJ) : 0.0);
                for (int q = 0; q < br.size(); ++q) {
                    oh[q] = 1.0;
                    v[i][h.offsetR[JJ] + q] += br.force(dIJ, oh.data()) * dIJi * gr;
                    oh[q] = 0.0;
                }
            }
            for (size_t k = j + 1; k < L.size(); ++k) {
                const int KK = L[k].type, kk = L[k].index;
                const bool kmin = all_images ? true : L[k].min;
                const V3& rIK = L[k].r;
                const double dIK = L[k].dr, dIKi = 1.0 / dIK;
                const V3 rJK = sub(rIK, rIJ);
                const double dJK = nrm(rJK), dJKi = 1.0 / dJK;
                const int blk = lmat(JJ, KK);
                const int off = h.nInputR + h.offsetA[blk];
                const double cs = dot(rIJ, rIK) * dIJi * dIKi;
                const double dd[3] = {dIJ, dIK, dJK};
                const BasisA& ba = h.basisA[blk];
                oh.assign(ba.size(), 0.0);
                for (int q = 0; q < ba.size(); ++q) {
                    oh[q] = 1.0;
                    double phi = 0, eta[3] = {0, 0, 0};
                    ba.force(phi, eta, cs, dd, oh.data());
                    oh[q] = 0.0;
                    const V3 fi = add(scl((phi * (dIKi - cs * dIJi) + eta[0]) * dIJi, rIJ),
                                      scl((phi * (dIJi - cs * dIKi) + eta[1]) * dIKi, rIK));
                    double acc = dot(gF[i], fi);
                    if (jmin) {
                        const V3 fj = add(add(scl((-phi * cs * dIJi + eta[0]) * dIJi, rIJ), scl(phi * dIJi * dIKi, rIK)),
                                          scl(eta[2] * dJKi, rJK));
                        acc -= dot(gF[jj], fj);
                    }
                    if (kmin) {
                        const V3 fk = sub(add(scl((-phi * cs * dIKi + eta[1]) * dIKi, rIK), scl(phi * dIKi * dIJi, rIJ)),
                                          scl(eta[2] * dJKi, rJK));
                        acc -= dot(gF[kk], fk);
                    }
                    v[i][off + q] += acc;
                }
            }
        }
    }
}

static void load_structure(Structure& s, int n, const double* cell, const double* pos, const int* type, const Model* m) {
    if (n <= 0) throw std::invalid_argument("natoms must be positive");
    s.n = n;
    double zero[9] = {0};
    s.cell.init(cell ? cell : zero);
    s.pos.resize(n);
    s.type.assign(type, type + n);
    for (int i = 0; i < n; ++i) {
        s.pos[i] = v3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
        if (m && (type[i] < 0 || type[i] >= m->ntypes)) throw std::invalid_argument("atom type out of range");
    }
}

static thread_local std::string g_err;

}  // namespace orc

using namespace orc;

// Loss of one structure's per-atom energy error dE (nnpte.cpp:876-916): adds the loss to L and
// returns dC/d(dE).
static double energy_loss(int loss, double dE, double delta, double& L) {
    const double deltai = 1.0 / delta, delta2 = delta * delta;
    switch (loss) {
        case 0: L += 0.5 * dE * dE; return dE;                                         // MSE
        case 1: L += std::fabs(dE); return (double)((0.0 < dE) - (dE < 0.0));          // MAE
        case 2: {                                                                      // HUBER
            const double arg = dE * deltai, sq = std::sqrt(1.0 + arg * arg);
            L += delta2 * (sq - 1.0);
            return dE / sq;
        }
        case 3: {                                                                      // ASINH
            const double arg = dE * deltai, sq = std::sqrt(1.0 + arg * arg), lg = std::log(arg + sq);
            L += delta2 * (1.0 - sq + arg * lg);
            return lg * delta;
        }
        default: throw std::invalid_argument("orc_train_error: unknown loss");
    }
}

// ---------------------------------------------------------------- C entry points (ctypes)
extern "C" {

const char* orc_last_error() { return g_err.c_str(); }

void* orc_load(const char* path) {
    try {
        auto* m = new Model();
        read_model(path, *m);
        return m;
    } catch (std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}
void orc_free(void* m) { delete (Model*)m; }
int orc_ntypes(void* m) { return ((Model*)m)->ntypes; }
double orc_rc(void* m) { return ((Model*)m)->rc; }
int orc_ninput(void* m, int t) { return ((Model*)m)->nnh[t].nInput; }
int orc_type_index(void* m, const char* name) {
    try {
        return ((Model*)m)->index(name);
    } catch (std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

// minimum-image difference (cell.cpp:103-128); lv = 3 lattice vectors
void orc_cell_diff(const double* lv, const double* v1, const double* v2, double* out) {
    Cell c;
    c.init(lv);
    const V3 d = c.diff(v3(v1[0], v1[1], v1[2]), v3(v2[0], v2[1], v2[2]));
    out[0] = d.x;
    out[1] = d.y;
    out[2] = d.z;
}

// Neighbour list (neighbor.cpp:59-171).  Writes per-atom counts; if the total fits in cap,
// also the entries in order.  Returns the total number of entries, or -1 on error.
long long orc_neighbors(int n, const double* cell, const double* pos, const int* type, double rc, int* counts, long long cap,
                        int* idx, int* img, int* mn, double* r, double* d, int* shell_out) {
    try {
        Structure s;
        load_structure(s, n, cell, pos, type, nullptr);
        NeighborList nl;
        neighbor_build(s.cell, n, s.pos.data(), s.type.data(), rc, nl);
        long long tot = 0;
        for (int i = 0; i < n; ++i) {
            counts[i] = (int)nl.neigh[i].size();
            tot += counts[i];
        }
        if (shell_out)
            for (int q = 0; q < 3; ++q) shell_out[q] = nl.shell[q];
        if (tot <= cap) {
            long long p = 0;
            for (int i = 0; i < n; ++i)
                for (auto& e : nl.neigh[i]) {
                    if (idx) idx[p] = e.index;
                    if (img) img[p] = e.image;
                    if (mn) mn[p] = e.min ? 1 : 0;
                    if (r) {
                        r[3 * p] = e.r.x;
                        r[3 * p + 1] = e.r.y;
                        r[3 * p + 2] = e.r.z;
                    }
                    if (d) d[p] = e.dr;
                    ++p;
                }
        }
        return tot;
    } catch (std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

// Descriptors (NNP::symm, nnp.cpp:334-381).  G is natoms x ldg, zero padded.
int orc_symm(void* mp, int n, const double* cell, const double* pos, const int* type, double* G, int ldg) {
    try {
        const Model& m = *(Model*)mp;
        Structure s;
        load_structure(s, n, cell, pos, type, &m);
        NeighborList nl;
        neighbor_build(s.cell, n, s.pos.data(), s.type.data(), m.rc, nl);
        nnp_symm(m, s, nl);
        for (int i = 0; i < n; ++i) {
            for (int q = 0; q < ldg; ++q) G[(size_t)i * ldg + q] = q < (int)s.symm[i].size() ? s.symm[i][q] : 0.0;
        }
        return 0;
    } catch (std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

// Network only: energy and dE/dG for one descriptor (ANN::fpbp + DODZ::grad).
int orc_network(void* mp, int type, const double* G, double* out, double* dEdG) {
    try {
        const Model& m = *(Model*)mp;
        const NNH& h = m.nnh[type];
        std::vector<std::vector<double>> a, d;
        *out = h.nn.fpbp(G, a, d);
        if (dEdG) h.nn.dodi(d, dEdG);
        return 0;
    } catch (std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

// Full evaluation as nnpce runs it (nnpce.cpp:275-279).
int orc_compute(void* mp, int n, const double* cell, const double* pos, const int* type, int flags, double* energy,
                double* eatom, double* force) {
    try {
        const Model& m = *(Model*)mp;
        Structure s;
        load_structure(s, n, cell, pos, type, &m);
        NeighborList nl;
        neighbor_build(s.cell, n, s.pos.data(), s.type.data(), m.rc, nl);
        nnp_symm(m, s, nl);
        nnp_energy_force(m, s, nl, flags);
        *energy = s.energy;
        if (eatom)
            for (int i = 0; i < n; ++i) eatom[i] = s.eatom[i];
        if (force && (flags & F_FORCE))
            for (int i = 0; i < n; ++i) {
                force[3 * i] = s.force[i].x;
                force[3 * i + 1] = s.force[i].y;
                force[3 * i + 2] = s.force[i].z;
            }
        return 0;
    } catch (std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

// CPU baseline: nstruc structures packed back to back (natoms[s] atoms each, cells 9 doubles
// each), evaluated with energy + forces by `nthreads` threads that take whole structures
// (the reference spreads structures over MPI ranks, nnpce.cpp:222).  Returns wall seconds
// (steady clock) of the evaluation loop, or a negative value on error.  energies[s] receives
// each total energy.
double orc_bench(void* mp, int nstruc, const int* natoms, const double* cells, const double* pos, const int* type, int flags,
                 int nthreads, double* energies) {
    try {
        const Model& m = *(Model*)mp;
        std::vector<long long> off(nstruc + 1, 0);
        for (int s = 0; s < nstruc; ++s) off[s + 1] = off[s] + natoms[s];
        std::atomic<int> next(0);
        std::atomic<bool> failed(false);
        auto work = [&]() {
            try {
                for (;;) {
                    const int s = next.fetch_add(1);
                    if (s >= nstruc) break;
                    Structure st;
                    load_structure(st, natoms[s], cells + 9 * (size_t)s, pos + 3 * off[s], type + off[s], &m);
                    NeighborList nl;
                    neighbor_build(st.cell, st.n, st.pos.data(), st.type.data(), m.rc, nl);
                    nnp_symm(m, st, nl);
                    nnp_energy_force(m, st, nl, flags);
                    energies[s] = st.energy;
                }
            } catch (...) {
                failed = true;
            }
        };
        const auto t0 = std::chrono::steady_clock::now();
        if (nthreads <= 1) {
            work();
        } else {
            std::vector<std::thread> th;
            for (int t = 0; t < nthreads; ++t) th.emplace_back(work);
            for (auto& t : th) t.join();
        }
        const auto t1 = std::chrono::steady_clock::now();
        if (failed) throw std::runtime_error("orc_bench: a structure failed");
        return std::chrono::duration<double>(t1 - t0).count();
    } catch (std::exception& e) {
        g_err = e.what();
        return -1.0;
    }
}

// ---------------------------------------------------------------- training inner loop
// NNPTE::error, training part (nnpte.cpp:869-927): for each structure of the batch,
// E = sum_i out_i (ANN::fp, no type energy), dE = (E - E_ref) / N, the loss and dcdo = dC/dE by
// the loss function (MSE, MAE, HUBER, ASINH with delta, deltai = 1/delta, delta2 = delta^2;
// nnpte.cpp:1195-1197), then per atom fpbp + Cost::grad(nn, dcdo) added to its species'
// gradient.  err[0] = sum of losses, err[1] = sum of dE^2 (error_[0], error_[2] before the
// BATCH normalisation, which is 1 on one process).  Descriptors are computed once at creation
// (NNP::symm, nnpte.cpp:1976-1977).
struct TrainSet {
    Model m;  // a copy: the parameters change per call
    std::vector<Structure> st;
    std::vector<NeighborList> nl;
    std::vector<double> eref;
    std::vector<std::vector<V3>> fref;  // reference forces (force training)
};

int orc_train_nparams(void* tp, int type) { return ((TrainSet*)tp)->m.nnh[type].nn.nparams(); }

void* orc_train_create(void* mp, int nstruc, const int* natoms, const double* cells, const double* pos, const int* type,
                       const double* eref) {
    try {
        auto* t = new TrainSet();
        t->m = *(Model*)mp;
        long long off = 0;
        for (int s = 0; s < nstruc; ++s) {
            Structure st;
            const double* cell = cells ? cells + 9 * (size_t)s : nullptr;
            bool open = true;
            if (cell)
                for (int q = 0; q < 9; ++q) open = open && cell[q] == 0.0;
            load_structure(st, natoms[s], open ? nullptr : cell, pos + 3 * off, type + off, &t->m);
            NeighborList nl;
            neighbor_build(st.cell, st.n, st.pos.data(), st.type.data(), t->m.rc, nl);
            nnp_symm(t->m, st, nl);
            t->st.push_back(std::move(st));
            t->nl.push_back(std::move(nl));
            t->eref.push_back(eref[s]);
            off += natoms[s];
        }
        return t;
    } catch (std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}

void orc_train_free(void* tp) { delete (TrainSet*)tp; }

int orc_train_get_params(void* tp, double* x) {
    TrainSet& t = *(TrainSet*)tp;
    int c = 0;
    for (int n = 0; n < t.m.ntypes; ++n) {
        t.m.nnh[n].nn.get_params(x + c);
        c += t.m.nnh[n].nn.nparams();
    }
    return c;
}

int orc_train_error(void* tp, const double* x, int nbatch, const int* batch, int loss, double delta, double* err,
                    double* grad) {
    try {
        TrainSet& t = *(TrainSet*)tp;
        const int T = t.m.ntypes;
        std::vector<int> off(T + 1, 0);
        for (int n = 0; n < T; ++n) off[n + 1] = off[n] + t.m.nnh[n].nn.nparams();
        for (int n = 0; n < T; ++n) t.m.nnh[n].nn.set_params(x + off[n]);
        if (grad)
            for (int q = 0; q < off[T]; ++q) grad[q] = 0.0;
        err[0] = err[1] = 0.0;
        std::vector<std::vector<double>> a, d;
        for (int b = 0; b < nbatch; ++b) {
            const Structure& s = t.st.at(batch[b]);
            double energy = 0;
            for (int i = 0; i < s.n; ++i) energy += t.m.nnh[s.type[i]].nn.fpbp(s.symm[i].data(), a, d);
            const double norm = 1.0 / s.n;
            const double dE = (energy - t.eref[batch[b]]) * norm;
            const double dcdo = norm * energy_loss(loss, dE, delta, err[0]);
            err[1] += dE * dE;
            if (grad)
                for (int i = 0; i < s.n; ++i) {
                    const ANN& nn = t.m.nnh[s.type[i]].nn;
                    nn.fpbp(s.symm[i].data(), a, d);
                    nn.cost_grad(s.symm[i].data(), a, d, dcdo, grad + off[s.type[i]]);
                }
        }
        return 0;
    } catch (std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

// Reference forces of every training atom (3 per atom, structures in creation order).
int orc_train_set_forces(void* tp, const double* f) {
    TrainSet& t = *(TrainSet*)tp;
    t.fref.assign(t.st.size(), {});
    size_t off = 0;
    for (size_t b = 0; b < t.st.size(); ++b) {
        t.fref[b].resize(t.st[b].n);
        for (int i = 0; i < t.st[b].n; ++i, ++off) t.fref[b][i] = v3(f[3 * off], f[3 * off + 1], f[3 * off + 2]);
    }
    return 0;
}

// Energy + force training loss and its gradient (the training part of NNPTEF::error,
// nnptef.cpp:1040-1293): per structure the energy loss of dE = (E - E_ref) / N (as
// orc_train_error) plus beta/2 |fdiff_i|^2 per atom with fdiff = (F_i - F_ref,i) / N
// (nnptef.cpp:1163-1166), the forces being NNP::compute's (reactions on minimum images only,
// or on every image with all_images).  err = {loss, sum dE^2, sum |fdiff|^2}.  The gradient is
// the exact derivative of that loss: the reference's own (nnptef.cpp:1138-1293, which does not
// compile) keeps only the centre atom's share of each force.
int orc_train_error_force(void* tp, const double* x, int nbatch, const int* batch, int loss, double delta, double beta,
                          int all_images, double* err, double* grad) {
    try {
        TrainSet& t = *(TrainSet*)tp;
        if (t.fref.size() != t.st.size()) throw std::invalid_argument("orc_train_error_force: no reference forces");
        const int T = t.m.ntypes;
        std::vector<int> off(T + 1, 0);
        for (int n = 0; n < T; ++n) off[n + 1] = off[n] + t.m.nnh[n].nn.nparams();
        for (int n = 0; n < T; ++n) t.m.nnh[n].nn.set_params(x + off[n]);
        if (grad)
            for (int q = 0; q < off[T]; ++q) grad[q] = 0.0;
        err[0] = err[1] = err[2] = 0.0;
        std::vector<std::vector<double>> a, d, e, v;
        std::vector<V3> gF;
        for (int b = 0; b < nbatch; ++b) {
            const int sb = batch[b];
            Structure s = t.st.at(sb);
            const NeighborList& nl = t.nl[sb];
            nnp_energy_force(t.m, s, nl, F_FORCE | (all_images ? F_ALL_IMAGES : 0));
            const double norm = 1.0 / s.n;
            const double dE = (s.energy - t.eref[sb]) * norm;
            const double dcdo = norm * energy_loss(loss, dE, delta, err[0]);
            err[1] += dE * dE;
            gF.assign(s.n, v3(0, 0, 0));
            for (int i = 0; i < s.n; ++i) {
                const V3 fd = scl(norm, sub(s.force[i], t.fref[sb][i]));
                err[0] += 0.5 * beta * sqn(fd);
                err[2] += sqn(fd);
                gF[i] = scl(beta * norm, fd);
            }
            if (!grad) continue;
            force_adjoint(t.m, s, nl, all_images != 0, gF, v);
            for (int i = 0; i < s.n; ++i) {
                const ANN& nn = t.m.nnh[s.type[i]].nn;
                nn.fpbp2(s.symm[i].data(), a, d, e);
                nn.force_grad(s.symm[i].data(), v[i].data(), a, d, e, dcdo, grad + off[s.type[i]]);
            }
        }
        return 0;
    } catch (std::exception& ex) {
        g_err = ex.what();
        return -1;
    }
}

// Second derivative of an activation (tests).
double orc_activation_d2(int neuron, double z) { return activate_d2((Neuron)neuron, z); }
int orc_activation(int neuron, double z, double* ad) {
    try {
        activate((Neuron)neuron, z, ad[0], ad[1]);
        return 0;
    } catch (std::exception& ex) {
        g_err = ex.what();
        return -1;
    }
}

}  // extern "C"
