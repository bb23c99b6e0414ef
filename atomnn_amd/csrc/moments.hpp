// Moment expansion of the angular (G4, BasisA GAUSS) sums, O(P) per centre instead of O(P^2).
//
// For integer zeta <= 4 and |lambda| <= 1 (then 1 + lambda cos >= 0 and the reference's fabs is
// inert, basis_angular.cpp:419):
//   ((1 + lambda cos)/2)^zeta = 2^-zeta sum_p C(zeta,p) lambda^p cos^p,
//   cos_jk^p = (u_j . u_k)^p = sum_{|a| = p} (p! / a!) u_j^a u_k^a      (multinomial theorem)
// so with per-species moments  M[s][e][a] = sum_{j in s} R_e(j) u_j^a  (R_e(j) = fc(d_j)
// exp(-ietap_e d_j^alpha)) the ordered-pair sums are  S_ab(e,p) = sum_{|a|=p} (p!/a!) M_a M_b.
// The reference sums unordered pairs j < k (nnp.cpp:359); for a != b that is S_ab, for a == b
// it is (S_aa - D_a)/2 with the diagonal D_a(e) = sum_{j in a} R_e(j)^2 (cos_jj = 1).
//
// Forces: with w_{ab,e,p} = sum_c dE/dG_f norm 2^-zeta_c C(zeta_c,p) lambda_c^p (x 1/2 for a == b)
// the adjoint moments are  L[s][a][e] = (p!/a!) (sum_{b != s} w_{sb,e,p} M_b + 2 w_{ss,e,p} M_s)
// and, for entry j of species s, with the polynomial  Q_e(u) = sum_a L[s][a][e] u^a  and its
// gradient h_e = grad_u Q_e (u taken as three free variables),
//   dE/dr_j = sum_e [ (Q_e(u_j) - 2 Delta_s(e) R_e) R'_e u_j + R_e (h_e - (u_j . h_e) u_j) / d_j ],
//   Delta_s(e) = sum_p w_{ss,e,p},
// using grad_r u^a = (grad_u u^a - |a| u^a u) / d and Euler's theorem sum_a |a| L u^a = u . h.
//
// GPU mapping (one wavefront per centre):
//  * moments: the 32 monomials q < 32 are summed directly, lane 32 h + q taking monomial q over
//    the entries of parity h, reading a per-entry power table (1, u, u^2, u^3, u^4 per axis, 15
//    values) from LDS: three broadcast-row reads and two products per entry instead of a stored
//    35-column monomial table; the halves meet by a shuffle, and the last three monomials of
//    degree 4 follow from the others because |u| = 1;
//  * forces: lane p owns entry p, keeps the 20 monomials of degree <= 3 in registers and walks
//    the 35 monomials unrolled at compile time, so each adjoint row L[s][q][0..NE) is read once
//    and feeds the value and all three gradient components.
#pragma once

#include <utility>

#include "device.hpp"
#include "model.hpp"

namespace atomnn {

constexpr int kMomZ = 4;    // highest monomial degree
constexpr int kMomK = 35;   // monomials of degree <= 4 in three variables
constexpr int kMomK3 = 20;  // monomials of degree <= 3
constexpr int kMomPw = 15;  // per-entry power table: (1, x, x^2, x^3, x^4) per axis
constexpr int kMomCh = 32;  // entries per power-table chunk in the force path

// Monomials in degree order (x exponent descending, then y): the first 20 have degree <= 3.
struct MomTab {
    signed char a[kMomK][3];    // exponents
    signed char low[kMomK][3];  // index of the monomial with a_c lowered by one, or -1
    unsigned char mult[kMomK];  // multinomial coefficient |a|! / (ax! ay! az!)
    unsigned char deg[kMomK];
};

constexpr MomTab make_mom_tab() {
    MomTab t{};
    int q = 0;
    for (int d = 0; d <= kMomZ; ++d)
        for (int ax = d; ax >= 0; --ax)
            for (int ay = d - ax; ay >= 0; --ay) {
                t.a[q][0] = (signed char)ax;
                t.a[q][1] = (signed char)ay;
                t.a[q][2] = (signed char)(d - ax - ay);
                t.deg[q] = (unsigned char)d;
                ++q;
            }
    const int f[5] = {1, 1, 2, 6, 24};
    for (int r = 0; r < kMomK; ++r) {
        t.mult[r] = (unsigned char)(f[t.deg[r]] / (f[t.a[r][0]] * f[t.a[r][1]] * f[t.a[r][2]]));
        for (int c = 0; c < 3; ++c) {
            t.low[r][c] = -1;
            if (t.a[r][c] == 0) continue;
            for (int s = 0; s < kMomK; ++s) {
                bool same = true;
                for (int k = 0; k < 3; ++k)
                    if (t.a[s][k] != t.a[r][k] - (k == c ? 1 : 0)) same = false;
                if (same) t.low[r][c] = (signed char)s;
            }
        }
    }
    return t;
}

// compile-time table (folded in unrolled code); lane-varying look-ups use mom_exps / mom_deg /
// mom_mult below, which compute the same entries without a memory access
constexpr MomTab kMomTabC = make_mom_tab();

constexpr int mom_first_dir(int q) {
    return kMomTabC.a[q][0] > 0 ? 0 : (kMomTabC.a[q][1] > 0 ? 1 : 2);
}

// LDS of the moment path, in elements of T.
//   descriptors: Mom [S][NE][K] | D [S][NE] | S table [blocks][NE][5]  (the power table goes
//                through the kernel's reduction buffer in chunks)
//   forces:      Mom | D (Delta) | w table [blocks][NE][5] | U
//                U = power table [kMomCh][15], then the adjoint table L [S][K][NE]
template <int NE>
__host__ __device__ constexpr int mom_symm_elems(int ntypes_max) {
    return ntypes_max * NE * kMomK + ntypes_max * NE + (ntypes_max * (ntypes_max + 1) / 2) * NE * 5;
}
template <int NE>
__host__ __device__ constexpr int mom_u_elems(int ntypes_max) {
    return kMomCh * kMomPw > ntypes_max * kMomK * NE ? kMomCh * kMomPw : ntypes_max * kMomK * NE;
}
template <int NE>
__host__ __device__ constexpr int mom_lds_elems(int ntypes_max) {
    return mom_symm_elems<NE>(ntypes_max) + mom_u_elems<NE>(ntypes_max);
}

// Power table rows for the entries [c0, c1): pw[(p - c0) * 15 + 5 c + k] = u_c^k.
template <typename T>
__device__ __forceinline__ void power_rows(const T* su, int c0, int c1, T* pw, int lane) {
    for (int p = c0 + lane; p < c1; p += kWave) {
        T* row = pw + (p - c0) * kMomPw;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const T x = su[4 * p + c], x2 = x * x;
            row[5 * c + 0] = T(1);
            row[5 * c + 1] = x;
            row[5 * c + 2] = x2;
            row[5 * c + 3] = x2 * x;
            row[5 * c + 4] = x2 * x2;
        }
    }
}

// Exponents, degree and multinomial coefficient of monomial q < 35 by integer arithmetic (the
// order of the table above): lane-varying look-ups in the constant table are global-memory loads,
// a round trip each time a loop or barrier separates them from their use.
__host__ __device__ constexpr void mom_exps(int q, int& ax, int& ay, int& az) {
    const int d = (q >= 1) + (q >= 4) + (q >= 10) + (q >= 20);
    const int r = q - d * (d + 1) * (d + 2) / 6;  // position inside the degree-d block
    const int t = (r >= 1) + (r >= 3) + (r >= 6) + (r >= 10);  // d - ax
    const int j = r - t * (t + 1) / 2;
    ax = d - t;
    ay = t - j;
    az = j;
}
__host__ __device__ constexpr int mom_deg(int q) {
    return (q >= 1) + (q >= 4) + (q >= 10) + (q >= 20);
}
__host__ __device__ constexpr int small_fact(int n) {
    return n < 2 ? 1 : (n == 2 ? 2 : (n == 3 ? 6 : 24));
}
// the coefficient as four bits of one of three 64-bit constants (no division, no memory access)
__host__ __device__ constexpr unsigned long long mom_mult_word(int w) {
    unsigned long long v = 0;
    for (int k = 0; k < 16 && 16 * w + k < kMomK; ++k) {
        const int q = 16 * w + k;
        int ax = 0, ay = 0, az = 0;
        mom_exps(q, ax, ay, az);
        v |= (unsigned long long)(small_fact(ax + ay + az) / (small_fact(ax) * small_fact(ay) * small_fact(az)))
             << (4 * k);
    }
    return v;
}
__host__ __device__ constexpr int mom_mult(int q) {
    constexpr unsigned long long w0 = mom_mult_word(0), w1 = mom_mult_word(1), w2 = mom_mult_word(2);
    const unsigned long long w = q < 16 ? w0 : (q < 32 ? w1 : w2);
    return (int)((w >> (4 * (q & 15))) & 15);
}
constexpr bool mom_arith_matches_table() {
    for (int q = 0; q < kMomK; ++q) {
        int ax = 0, ay = 0, az = 0;
        mom_exps(q, ax, ay, az);
        if (ax != kMomTabC.a[q][0] || ay != kMomTabC.a[q][1] || az != kMomTabC.a[q][2]) return false;
        if (mom_deg(q) != kMomTabC.deg[q] || mom_mult(q) != kMomTabC.mult[q]) return false;
    }
    return true;
}
static_assert(mom_arith_matches_table(), "monomial arithmetic and table agree");

// Monomials summed directly: q < 32.  The last three of degree 4 follow from |u| = 1
// (x^2 + y^2 + z^2 = 1 for every entry, so each moment identity below holds term by term):
//   y^2 z^2 = y^2 - x^2 y^2 - y^4,  y z^3 = y z - x^2 y z - y^3 z,  z^4 = z^2 - x^2 z^2 - y^2 z^2.
constexpr int kMomDirect = 32;
constexpr int mom_index(int ax, int ay, int az) {
    for (int q = 0; q < kMomK; ++q)
        if (kMomTabC.a[q][0] == ax && kMomTabC.a[q][1] == ay && kMomTabC.a[q][2] == az) return q;
    return -1;
}
static_assert(mom_index(0, 2, 2) == 32 && mom_index(0, 1, 3) == 33 && mom_index(0, 0, 4) == 34,
              "the derived monomials are the last three");

// Moment rows M[q] (stride 1) of one (species, exponent): fill q = 32, 33, 34 from the others.
template <typename T>
__device__ __forceinline__ void mom_derive(T* M) {
    constexpr int y2 = mom_index(0, 2, 0), x2y2 = mom_index(2, 2, 0), y4 = mom_index(0, 4, 0);
    constexpr int yz = mom_index(0, 1, 1), x2yz = mom_index(2, 1, 1), y3z = mom_index(0, 3, 1);
    constexpr int z2 = mom_index(0, 0, 2), x2z2 = mom_index(2, 0, 2);
    static_assert(y2 >= 0 && x2y2 >= 0 && y4 >= 0 && yz >= 0 && x2yz >= 0 && y3z >= 0 && z2 >= 0 && x2z2 >= 0 &&
                      x2y2 < kMomDirect && y4 < kMomDirect && x2yz < kMomDirect && y3z < kMomDirect &&
                      x2z2 < kMomDirect,
                  "the identities use directly summed monomials");
    M[32] = M[y2] - M[x2y2] - M[y4];
    M[33] = M[yz] - M[x2yz] - M[y3z];
    M[34] = M[z2] - M[x2z2] - M[32];
}

// One chunk's contribution to the moments of the entries [c0, c1) whose power rows are in pw:
// lane 32 h + q sums monomial q over the entries of parity h of each species' part of the chunk,
// the two halves are added by a shuffle and lane q (h = 0) adds the sum to Mom[s][e][q].  Every
// moment is a fixed-order sum.
template <typename T, int NE>
__device__ __forceinline__ void moments_chunk(const T* Rf, int c0, int c1, int ntypes, const int* seg, const T* pw,
                                              T* Mom, int lane) {
    const int q = lane & 31, hf = lane >> 5;
    int ax = 0, ay = 0, az = 0;
    mom_exps(q, ax, ay, az);
    const int ox = ax, oy = 5 + ay, oz = 10 + az;
    for (int s = 0; s < ntypes; ++s) {
        const int lo = max(seg[s], c0), hi = min(seg[s + 1], c1);
        if (lo >= hi) continue;
        T a[NE];
#pragma unroll
        for (int e = 0; e < NE; ++e) a[e] = T(0);
#pragma unroll 2
        for (int p = lo + hf; p < hi; p += 2) {
            const T* row = pw + (p - c0) * kMomPw;
            const T m = row[ox] * row[oy] * row[oz];
#pragma unroll
            for (int e = 0; e < NE; ++e) a[e] += Rf[p * NE + e] * m;
        }
#pragma unroll
        for (int e = 0; e < NE; ++e) a[e] += __shfl_xor(a[e], kWave / 2);
        if (hf == 0)
#pragma unroll
            for (int e = 0; e < NE; ++e) Mom[(s * NE + e) * kMomK + q] += a[e];
    }
}

// Per-species moments Mom[s][e][q] of the staged entries (Rf holds R_e(p) for every entry) and,
// when D is given, the diagonals D[s][e] = sum_{p in s} R_e(p)^2.  The power table pw holds
// `ch` entries at a time.
template <typename T, int NE>
__device__ __forceinline__ void build_moments(const T* su, const T* Rf, int P, int ch, int ntypes, const int* seg,
                                              T* pw, T* Mom, T* D, int lane) {
    ch = max(ch, 1);
    for (int it = lane; it < ntypes * NE * kMomK; it += kWave) Mom[it] = T(0);
    for (int c0 = 0; c0 < P || c0 == 0; c0 += ch) {
        const int c1 = min(P, c0 + ch);
        __syncthreads();  // the previous chunk's readers are done (and the moments cleared)
        power_rows<T>(su, c0, c1, pw, lane);
        __syncthreads();
        moments_chunk<T, NE>(Rf, c0, c1, ntypes, seg, pw, Mom, lane);
        if (c1 >= P) break;
    }
    __syncthreads();
    for (int it = lane; it < ntypes * NE; it += kWave) mom_derive<T>(Mom + it * kMomK);
    if (D) {
        for (int it = lane; it < ntypes * NE; it += kWave) {
            const int e = it % NE, s = it / NE;
            T a = T(0);
            for (int p = seg[s]; p < seg[s + 1]; ++p) a += Rf[p * NE + e] * Rf[p * NE + e];
            D[it] = a;  // [s][e]
        }
    }
    __syncthreads();
}

__device__ __forceinline__ int binom4(int z, int p) {
    return z == 4 ? (p == 0 || p == 4 ? 1 : (p == 2 ? 6 : 4)) : (z == 3 ? (p == 0 || p == 3 ? 1 : 3) : (z == 2 && p == 1 ? 2 : (p <= z ? 1 : 0)));
}

__device__ __forceinline__ void blk_species(int blk, int& a, int& b) {
    int bb = 0;
    while ((bb + 1) * (bb + 2) / 2 <= blk) ++bb;
    b = bb;
    a = blk - bb * (bb + 1) / 2;
}

// Angular descriptors of centre species ti from the moments.
template <typename T, int NE>
__device__ __forceinline__ void symm_angular_moments(const DevPlan* __restrict__ plan, int ti, int ntypes,
                                                     const T* Mom, const T* D, T* S, T* Grow, int lane) {
    const int nblk = ntypes * (ntypes + 1) / 2;
    // S[blk][e][p] = sum_{|a| = p} (p!/a!) M_a[e][a] M_b[e][a]
    for (int it = lane; it < nblk * NE * (kMomZ + 1); it += kWave) {
        const int pd = it % (kMomZ + 1), e = (it / (kMomZ + 1)) % NE, blk = it / ((kMomZ + 1) * NE);
        int a, bb;
        blk_species(blk, a, bb);
        const T* Ma = Mom + (a * NE + e) * kMomK;
        const T* Mb = Mom + (bb * NE + e) * kMomK;
        T acc = T(0);
        const int q0 = pd * (pd + 1) * (pd + 2) / 6, q1 = (pd + 1) * (pd + 2) * (pd + 3) / 6;
        for (int q = q0; q < q1; ++q) acc += T(mom_mult(q)) * Ma[q] * Mb[q];
        S[it] = acc;
    }
    __syncthreads();
    // G_f = norm 2^-zeta sum_p C(zeta,p) lambda^p X_p, X_p = S_ab or (S_aa - D_a)/2
    if (plan->ang_coef && plan->adj_ne == NE) {  // the engine's rows: index and five coefficients
        const size_t r0 = (size_t)ti * nblk * NE * kMaxC;
        for (int it = lane; it < nblk * NE * kMaxC; it += kWave) {
            const int gi = plan->ang_index[r0 + it];
            if (gi < 0) continue;
            const int e = (it / kMaxC) % NE, blk = it / (kMaxC * NE);
            int a, bb;
            blk_species(blk, a, bb);
            const double* cf = plan->ang_coef + (r0 + it) * (kMomZ + 1);
            T acc = T(0);
#pragma unroll
            for (int p = 0; p <= kMomZ; ++p) {
                T x = S[(blk * NE + e) * (kMomZ + 1) + p];
                if (a == bb) x = T(0.5) * (x - D[a * NE + e]);
                acc += T(cf[p]) * x;
            }
            Grow[gi] = acc;
        }
        return;
    }
    for (int it = lane; it < nblk * NE * kMaxC; it += kWave) {
        const int c = it % kMaxC, e = (it / kMaxC) % NE, blk = it / (kMaxC * NE);
        const DevAngular& A = plan->ang[ti][blk];
        if (e >= A.ne || c >= A.nc) continue;
        const int f = A.fmap[e * kMaxC + c];
        if (f < 0) continue;
        int a, bb;
        blk_species(blk, a, bb);
        const int z = A.zint[c];
        const T lam = T(A.lambda[c]);
        T acc = T(0), lp = T(1);
        for (int p = 0; p <= z; ++p) {
            T x = S[(blk * NE + e) * (kMomZ + 1) + p];
            if (a == bb) x = T(0.5) * (x - D[a * NE + e]);
            acc += T(binom4(z, p)) * lp * x;
            lp *= lam;
        }
        Grow[A.goff + f] = acc * T(A.norm) * T(1.0 / (double)(1 << z));
    }
}

// Tangents of the moments along entry displacements (force training).  tg[4p..4p+3] holds
// (u_x', u_y', u_z', d') of entry p, Rf / Rd hold R_e(p) and R_e'(p).  Per species and exponent:
//   M[s][e][q]  = sum_p R_e u^q,
//   dM[s][e][q] = sum_p (R_e' d' u^q + R_e grad(u^q) . u'),  grad(u^q) . u' = sum_c q_c u^(q - e_c) u_c',
//   dD[s][e]    = sum_p 2 R_e R_e' d'   (the tangent of the diagonal D = sum_p R_e^2).
// Lane q < 35 owns monomial q; entries in index order; the power table goes through pw in chunks.
template <int NE>
__device__ __forceinline__ void build_moments_jvp(const double* su, const double* tg, const double* Rf,
                                                  const double* Rd, int P, int ch, int ntypes, const int* seg,
                                                  double* pw, double* Mom, double* dMom, double* D, double* dD,
                                                  int lane) {
    ch = max(ch, 1);
    const int q = lane;
    int ax = 0, ay = 0, az = 0;
    if (q < kMomK) {
        mom_exps(q, ax, ay, az);
    }
    // power-table offsets of u_c^(a_c) and u_c^(a_c - 1) (the latter only read when a_c > 0)
    const int ox = ax, oy = 5 + ay, oz = 10 + az;
    const int lx = ax > 0 ? ox - 1 : ox, ly = ay > 0 ? oy - 1 : oy, lz = az > 0 ? oz - 1 : oz;
    const double fx = ax, fy = ay, fz = az;
    for (int c0 = 0; c0 < P || c0 == 0; c0 += ch) {
        const int c1 = min(P, c0 + ch);
        __syncthreads();
        power_rows<double>(su, c0, c1, pw, lane);
        __syncthreads();
        if (q < kMomK) {
            for (int s = 0; s < ntypes; ++s) {
                const int lo = max(seg[s], c0), hi = min(seg[s + 1], c1);
                double a[NE], da[NE];
#pragma unroll
                for (int e = 0; e < NE; ++e) {
                    a[e] = c0 == 0 ? 0.0 : Mom[(s * NE + e) * kMomK + q];
                    da[e] = c0 == 0 ? 0.0 : dMom[(s * NE + e) * kMomK + q];
                }
                for (int p = lo; p < hi; ++p) {
                    const double* row = pw + (p - c0) * kMomPw;
                    const double X = row[ox], Y = row[oy], Z = row[oz];
                    const double m = X * Y * Z;
                    const double dm = fx * row[lx] * Y * Z * tg[4 * p] + fy * X * row[ly] * Z * tg[4 * p + 1] +
                                      fz * X * Y * row[lz] * tg[4 * p + 2];
                    const double dd = tg[4 * p + 3];
#pragma unroll
                    for (int e = 0; e < NE; ++e) {
                        const double r = Rf[p * NE + e];
                        a[e] += r * m;
                        da[e] += Rd[p * NE + e] * dd * m + r * dm;
                    }
                }
#pragma unroll
                for (int e = 0; e < NE; ++e) {
                    Mom[(s * NE + e) * kMomK + q] = a[e];
                    dMom[(s * NE + e) * kMomK + q] = da[e];
                }
            }
        }
        if (c1 >= P) break;
    }
    for (int it = lane; it < ntypes * NE; it += kWave) {
        const int e = it % NE, s = it / NE;
        double a = 0.0, da = 0.0;
        for (int p = seg[s]; p < seg[s + 1]; ++p) {
            const double r = Rf[p * NE + e];
            a += r * r;
            da += 2.0 * r * Rd[p * NE + e] * tg[4 * p + 3];
        }
        if (D) D[it] = a;
        dD[it] = da;
    }
    __syncthreads();
}

// Tangents of the angular descriptors from the moments and their tangents: per block (a, b),
// exponent e and power p, dS = sum_{|q| = p} mult(q) (dM_a M_b + M_a dM_b), dX_p = dS for a != b
// and (dS - dD_a) / 2 for a == b; G_f' = norm 2^-zeta sum_p C(zeta, p) lambda^p dX_p.  Writes
// -G_f' into Vrow (the sign of the force-training adjoint).  S: nblk NE 5 doubles of scratch.
template <int NE>
__device__ __forceinline__ void symm_angular_moments_jvp(const DevPlan* __restrict__ plan, int ti, int ntypes,
                                                         const double* Mom, const double* dMom, const double* dD,
                                                         double* S, double* Vrow, int lane) {
    const int nblk = ntypes * (ntypes + 1) / 2;
    for (int it = lane; it < nblk * NE * (kMomZ + 1); it += kWave) {
        const int pd = it % (kMomZ + 1), e = (it / (kMomZ + 1)) % NE, blk = it / ((kMomZ + 1) * NE);
        int a, bb;
        blk_species(blk, a, bb);
        const double* Ma = Mom + (a * NE + e) * kMomK;
        const double* Mb = Mom + (bb * NE + e) * kMomK;
        const double* dMa = dMom + (a * NE + e) * kMomK;
        const double* dMb = dMom + (bb * NE + e) * kMomK;
        double acc = 0.0;
        const int q0 = pd * (pd + 1) * (pd + 2) / 6, q1 = (pd + 1) * (pd + 2) * (pd + 3) / 6;
        for (int q = q0; q < q1; ++q) acc += double(mom_mult(q)) * (dMa[q] * Mb[q] + Ma[q] * dMb[q]);
        S[it] = acc;
    }
    __syncthreads();
    for (int it = lane; it < nblk * NE * kMaxC; it += kWave) {
        const int c = it % kMaxC, e = (it / kMaxC) % NE, blk = it / (kMaxC * NE);
        const DevAngular& A = plan->ang[ti][blk];
        if (e >= A.ne || c >= A.nc) continue;
        const int f = A.fmap[e * kMaxC + c];
        if (f < 0) continue;
        int a, bb;
        blk_species(blk, a, bb);
        const int z = A.zint[c];
        const double lam = A.lambda[c];
        double acc = 0.0, lp = 1.0;
        for (int p = 0; p <= z; ++p) {
            double x = S[(blk * NE + e) * (kMomZ + 1) + p];
            if (a == bb) x = 0.5 * (x - dD[a * NE + e]);
            acc += double(binom4(z, p)) * lp * x;
            lp *= lam;
        }
        Vrow[A.goff + f] = -acc * A.norm * (1.0 / (double)(1 << z));
    }
}

// Adjoint tables from dE/dG (row W): w table, Delta[s][e] (into Delta) and L[s][q][e] (into L,
// which held the power table: the caller's barrier after the moments separates the two uses).
// The moments are the sum of `nparts` partial tables `pstride` elements apart, added in order.
// Threads tid of nthreads share the items; the barriers are block-wide.  TAB: the w table from the
// engine's coefficient rows (plan->adj_coef) when present.
template <typename T, int NE, bool TAB = false>
__device__ __forceinline__ void adjoint_moments(const DevPlan* __restrict__ plan, int ti, int ntypes, const T* W,
                                                const T* Mom, int nparts, int pstride, T* Delta, T* wtab, T* L,
                                                int tid, int nthreads) {
    const int nblk = ntypes * (ntypes + 1) / 2;
    if (TAB && plan->adj_coef && plan->adj_ne == NE) {
        // the w table from the engine's coefficient rows: one row of constants per item
        const size_t b0 = (size_t)ti * nblk * NE * (kMomZ + 1);
        for (int it = tid; it < nblk * NE * (kMomZ + 1); it += nthreads) {
            const double* cf = plan->adj_coef + (b0 + it) * kMaxC;
            const int* ix = plan->adj_index + (b0 + it) * kMaxC;
            T acc = T(0);
#pragma unroll
            for (int c = 0; c < kMaxC; ++c) acc += W[ix[c]] * T(cf[c]);
            wtab[it] = acc;
        }
    } else
    for (int it = tid; it < nblk * NE * (kMomZ + 1); it += nthreads) {
        const int pd = it % (kMomZ + 1), e = (it / (kMomZ + 1)) % NE, blk = it / ((kMomZ + 1) * NE);
        const DevAngular& A = plan->ang[ti][blk];
        int a, bb;
        blk_species(blk, a, bb);
        T acc = T(0);
        if (e < A.ne) {
            for (int c = 0; c < A.nc; ++c) {
                const int f = A.fmap[e * kMaxC + c];
                const int z = A.zint[c];
                if (f < 0 || pd > z) continue;
                T lp = T(1);
                for (int k = 0; k < pd; ++k) lp *= T(A.lambda[c]);
                acc += W[A.goff + f] * T(binom4(z, pd)) * lp * T(1.0 / (double)(1 << z));
            }
            acc *= T(A.norm);
            if (a == bb) acc *= T(0.5);
        }
        wtab[it] = acc;
    }
    __syncthreads();
    for (int it = tid; it < ntypes * kMomK * NE; it += nthreads) {
        const int e = it % NE, q = (it / NE) % kMomK, s = it / (NE * kMomK);
        const int pd = mom_deg(q);
        T acc = T(0);
        for (int b = 0; b < ntypes; ++b) {
            const int blk = b >= s ? b * (b + 1) / 2 + s : s * (s + 1) / 2 + b;
            const T wv = wtab[(blk * NE + e) * (kMomZ + 1) + pd];
            const int mi = (b * NE + e) * kMomK + q;
            T mb = Mom[mi];
            for (int k = 1; k < nparts; ++k) mb += Mom[k * pstride + mi];
            acc += (b == s ? T(2) : T(1)) * wv * mb;
        }
        L[it] = T(mom_mult(q)) * acc;  // [s][q][e]
    }
    for (int it = tid; it < ntypes * NE; it += nthreads) {
        const int e = it % NE, s = it / NE;
        const int blk = s * (s + 1) / 2 + s;
        T acc = T(0);
#pragma unroll
        for (int pd = 0; pd <= kMomZ; ++pd) acc += wtab[(blk * NE + e) * (kMomZ + 1) + pd];
        Delta[it] = acc;
    }
    __syncthreads();
}

// Q_e(u) and the gradient h_e accumulated over monomial Q (compile-time index) for NK exponents
// e = e0 + ES k: Lb points at L[s][0][e0], RS = NE is the row stride of the L table.
template <typename T, int NK, int Q, int C>
__device__ __forceinline__ void mom_grad(const T* m, const T (&Lq)[NK], T (&h)[3][NK]) {
    constexpr int ac = kMomTabC.a[Q][C];
    if constexpr (ac > 0) {
        constexpr int lo = kMomTabC.low[Q][C];
        const T t = ac == 1 ? m[lo] : T(ac) * m[lo];
#pragma unroll
        for (int k = 0; k < NK; ++k) h[C][k] += Lq[k] * t;
    }
}

template <typename T, int NK, int RS, int ES, int Q>
__device__ __forceinline__ void mom_term(const T* Lb, const T* m, const T (&u)[3], T (&g0)[NK], T (&h)[3][NK]) {
    T Lq[NK];
#pragma unroll
    for (int k = 0; k < NK; ++k) Lq[k] = Lb[Q * RS + ES * k];
    T mq;
    if constexpr (Q < kMomK3) {
        mq = m[Q];
    } else {
        constexpr int c = mom_first_dir(Q);
        constexpr int lo = kMomTabC.low[Q][c];
        mq = m[lo] * u[c];
    }
#pragma unroll
    for (int k = 0; k < NK; ++k) g0[k] += Lq[k] * mq;
    mom_grad<T, NK, Q, 0>(m, Lq, h);
    mom_grad<T, NK, Q, 1>(m, Lq, h);
    mom_grad<T, NK, Q, 2>(m, Lq, h);
}

template <typename T, int NK, int RS, int ES, int... Qs>
__device__ __forceinline__ void mom_terms(const T* Lb, const T* m, const T (&u)[3], T (&g0)[NK], T (&h)[3][NK],
                                          std::integer_sequence<int, Qs...>) {
    (mom_term<T, NK, RS, ES, Qs>(Lb, m, u, g0, h), ...);
}

// monomial Q (1 <= Q < 20) as the product of a lower one and one coordinate
template <typename T, int Q>
__device__ __forceinline__ void mom_mono(T* m, const T (&u)[3]) {
    if constexpr (Q > 0) {
        constexpr int c = mom_first_dir(Q);
        constexpr int lo = kMomTabC.low[Q][c];
        m[Q] = m[lo] * u[c];
    } else {
        m[0] = T(1);
    }
}

template <typename T, int... Qs>
__device__ __forceinline__ void mom_monos(T* m, const T (&u)[3], std::integer_sequence<int, Qs...>) {
    (mom_mono<T, Qs>(m, u), ...);
}

// Angular force term of every entry from the adjoint tables, subtracted from Xs.
template <typename T, int NE>
__device__ __forceinline__ void force_from_moments(int ntypes, int P, const T* su, const T* Rf, const T* Rd,
                                                   const int* seg, const T* L, const T* Delta, T* Xs, int lane) {
    for (int p = lane; p < P; p += kWave) {
        int s = 0;
        while (s + 1 < ntypes && p >= seg[s + 1]) ++s;
        const T u[3] = {su[4 * p], su[4 * p + 1], su[4 * p + 2]};
        const T di = T(1) / su[4 * p + 3];
        // monomials of degree <= 3, each the product of a lower one and one coordinate
        T m[kMomK3];
        mom_monos<T>(m, u, std::make_integer_sequence<int, kMomK3>{});
        T g0[NE], h[3][NE];
#pragma unroll
        for (int e = 0; e < NE; ++e) g0[e] = h[0][e] = h[1][e] = h[2][e] = T(0);
        mom_terms<T, NE, NE, 1>(L + s * kMomK * NE, m, u, g0, h, std::make_integer_sequence<int, kMomK>{});
        T gx = T(0), gy = T(0), gz = T(0);
#pragma unroll
        for (int e = 0; e < NE; ++e) {
            const T r = Rf[p * NE + e], rd = Rd[p * NE + e];
            const T gu = u[0] * h[0][e] + u[1] * h[1][e] + u[2] * h[2][e];  // sum_q |q| L_q u^q
            const T dEdR = g0[e] - T(2) * Delta[s * NE + e] * r;
            const T a1 = dEdR * rd - r * gu * di;
            const T a2 = r * di;
            gx += a1 * u[0] + a2 * h[0][e];
            gy += a1 * u[1] + a2 * h[1][e];
            gz += a1 * u[2] + a2 * h[2][e];
        }
        Xs[3 * p] -= gx;
        Xs[3 * p + 1] -= gy;
        Xs[3 * p + 2] -= gz;
    }
}

}  // namespace atomnn
