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
// the adjoint moments are  L[s][e][a] = (p!/a!) (sum_{b != s} w_{sb,e,p} M_b + 2 w_{ss,e,p} M_s)
// and, for entry j of species s,
//   dE/dr_j = sum_e [ (sum_a L u^a - 2 Delta_s(e) R_e) R'_e u_j + R_e (sum_a L grad_u u^a
//                     - (sum_a |a| L u^a) u_j) / d_j ],   Delta_s(e) = sum_p w_{ss,e,p},
// using grad_r u^a = (grad_u u^a - |a| u^a u) / d (Euler's theorem for degree-|a| monomials).
#pragma once

#include "device.hpp"
#include "model.hpp"

namespace atomnn {

constexpr int kMomZ = 4;   // highest monomial degree
constexpr int kMomK = 35;  // monomials of degree <= 4 in three variables
constexpr int kMomK3 = 20; // monomials of degree <= 3

// Exponents (ax, ay, az) of the monomials in degree order; the first 20 have degree <= 3.
__constant__ static const signed char kMono[kMomK][3] = {
    {0, 0, 0},                                                                                // degree 0
    {1, 0, 0}, {0, 1, 0}, {0, 0, 1},                                                          // 1
    {2, 0, 0}, {1, 1, 0}, {1, 0, 1}, {0, 2, 0}, {0, 1, 1}, {0, 0, 2},                         // 2
    {3, 0, 0}, {2, 1, 0}, {2, 0, 1}, {1, 2, 0}, {1, 1, 1}, {1, 0, 2}, {0, 3, 0}, {0, 2, 1},   // 3
    {0, 1, 2}, {0, 0, 3},
    {4, 0, 0}, {3, 1, 0}, {3, 0, 1}, {2, 2, 0}, {2, 1, 1}, {2, 0, 2}, {1, 3, 0}, {1, 2, 1},   // 4
    {1, 1, 2}, {1, 0, 3}, {0, 4, 0}, {0, 3, 1}, {0, 2, 2}, {0, 1, 3}, {0, 0, 4}};

__device__ __forceinline__ int mono_deg(int q) { return kMono[q][0] + kMono[q][1] + kMono[q][2]; }

// multinomial coefficient p! / (ax! ay! az!)
__device__ __forceinline__ int mono_mult(int q) {
    const int f[5] = {1, 1, 2, 6, 24};
    return f[mono_deg(q)] / (f[kMono[q][0]] * f[kMono[q][1]] * f[kMono[q][2]]);
}

// index of the monomial with exponents (ax, ay, az), degree <= 4
__device__ __forceinline__ int mono_index(int ax, int ay, int az) {
    for (int q = 0; q < kMomK; ++q)
        if (kMono[q][0] == ax && kMono[q][1] == ay && kMono[q][2] == az) return q;
    return -1;
}

__device__ __forceinline__ int binom4(int z, int p) {
    const int c[5][5] = {{1, 0, 0, 0, 0}, {1, 1, 0, 0, 0}, {1, 2, 1, 0, 0}, {1, 3, 3, 1, 0}, {1, 4, 6, 4, 1}};
    return c[z][p];
}

// u^a for all 35 monomials of one unit vector
template <typename T>
__device__ __forceinline__ void monomials(T ux, T uy, T uz, T* m) {
    T px[5], py[5], pz[5];
    px[0] = py[0] = pz[0] = T(1);
#pragma unroll
    for (int k = 1; k < 5; ++k) {
        px[k] = px[k - 1] * ux;
        py[k] = py[k - 1] * uy;
        pz[k] = pz[k - 1] * uz;
    }
#pragma unroll
    for (int q = 0; q < kMomK; ++q) m[q] = px[kMono[q][0]] * py[kMono[q][1]] * pz[kMono[q][2]];
}

// LDS of the moment path, sizes in elements of T.  Descriptors: Mom [S][NE][K] | D [S][NE] |
// S table [blocks][NE][5] (the monomial table lives in chunks in the kernel's reduction buffer).
// Forces: mono [K][lmaxp+1] | Mom | D | w table [blocks][NE][5] | L, Lp [S][NE][K] | Lx, Ly, Lz
// [S][NE][20].
template <int NE>
__host__ __device__ constexpr int mom_symm_elems(int ntypes_max) {
    return ntypes_max * NE * kMomK + ntypes_max * NE + (ntypes_max * (ntypes_max + 1) / 2) * NE * 5;
}
template <int NE>
__host__ __device__ constexpr int mom_lds_elems(int lmaxp, int ntypes_max) {
    return kMomK * (lmaxp + 1) + mom_symm_elems<NE>(ntypes_max) + 2 * ntypes_max * NE * kMomK +
           3 * ntypes_max * NE * kMomK3;
}

// Per-species moments M and diagonals D of the staged entries (Rf holds R_e(p) for every entry,
// radial_tables over [0, P)).  The monomial table mono [K][ch + 1] holds `ch` entries at a time.
template <typename T, int NE>
__device__ __forceinline__ void build_moments(const T* su, const T* Rf, int P, int ch, int ntypes, const int* seg,
                                              T* mono, T* Mom, T* D, int lane) {
    const int ms = ch + 1;
    const int items = ntypes * NE * kMomK;
    for (int it = lane; it < items; it += kWave) Mom[it] = T(0);
    for (int c0 = 0; c0 < P; c0 += ch) {
        const int c1 = min(P, c0 + ch);
        __syncthreads();
        for (int p = c0 + lane; p < c1; p += kWave) {
            T m[kMomK];
            monomials<T>(su[4 * p], su[4 * p + 1], su[4 * p + 2], m);
#pragma unroll
            for (int q = 0; q < kMomK; ++q) mono[q * ms + (p - c0)] = m[q];
        }
        __syncthreads();
        for (int it = lane; it < items; it += kWave) {
            const int q = it % kMomK, e = (it / kMomK) % NE, s = it / (kMomK * NE);
            const int lo = max(seg[s], c0), hi = min(seg[s + 1], c1);
            T acc = Mom[it];
            for (int p = lo; p < hi; ++p) acc += mono[q * ms + (p - c0)] * Rf[p * NE + e];
            Mom[it] = acc;  // [s][e][q]
        }
    }
    for (int it = lane; it < ntypes * NE; it += kWave) {
        const int e = it % NE, s = it / NE;
        T acc = T(0);
        for (int p = seg[s]; p < seg[s + 1]; ++p) acc += Rf[p * NE + e] * Rf[p * NE + e];
        D[it] = acc;  // [s][e]
    }
    __syncthreads();
}

// Angular descriptors of centre species ti from the moments.
template <typename T, int NE>
__device__ __forceinline__ void symm_angular_moments(const DevPlan* __restrict__ plan, int ti, int ntypes,
                                                     const T* Mom, const T* D, T* S, T* Grow, int lane) {
    const int nblk = ntypes * (ntypes + 1) / 2;
    // S[blk][e][p] = sum_{|a| = p} (p!/a!) M_a[e][a] M_b[e][a]
    for (int it = lane; it < nblk * NE * (kMomZ + 1); it += kWave) {
        const int pd = it % (kMomZ + 1), e = (it / (kMomZ + 1)) % NE, blk = it / ((kMomZ + 1) * NE);
        int bb = 0;
        while ((bb + 1) * (bb + 2) / 2 <= blk) ++bb;
        const int a = blk - bb * (bb + 1) / 2;
        const T* Ma = Mom + (a * NE + e) * kMomK;
        const T* Mb = Mom + (bb * NE + e) * kMomK;
        T acc = T(0);
        const int q0 = pd == 0 ? 0 : (pd == 1 ? 1 : (pd == 2 ? 4 : (pd == 3 ? 10 : 20)));
        const int q1 = pd == 0 ? 1 : (pd == 1 ? 4 : (pd == 2 ? 10 : (pd == 3 ? 20 : 35)));
        for (int q = q0; q < q1; ++q) acc += T(mono_mult(q)) * Ma[q] * Mb[q];
        S[it] = acc;
    }
    __syncthreads();
    // G_f = norm 2^-zeta sum_p C(zeta,p) lambda^p X_p, X_p = S_ab or (S_aa - D_a)/2
    for (int it = lane; it < nblk * NE * kMaxC; it += kWave) {
        const int c = it % kMaxC, e = (it / kMaxC) % NE, blk = it / (kMaxC * NE);
        const DevAngular& A = plan->ang[ti][blk];
        if (e >= A.ne || c >= A.nc) continue;
        const int f = A.fmap[e * kMaxC + c];
        if (f < 0) continue;
        int bb = 0;
        while ((bb + 1) * (bb + 2) / 2 <= blk) ++bb;
        const int a = blk - bb * (bb + 1) / 2;
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

// Adjoint moments from dE/dG (row W), then the angular force term of every entry into Xs.
template <typename T, int NE>
__device__ __forceinline__ void force_angular_moments(const DevPlan* __restrict__ plan, int ti, int ntypes, int P,
                                                      int maxp, const T* su, const T* Rf, const T* Rd, const T* W,
                                                      const int* seg, const T* mono, const T* Mom, T* wtab, T* Lam,
                                                      T* Lamp, T* Lx, T* Ly, T* Lz, T* Xs, int lane) {
    const int nblk = ntypes * (ntypes + 1) / 2;
    // w_{blk,e,p}
    for (int it = lane; it < nblk * NE * (kMomZ + 1); it += kWave) {
        const int pd = it % (kMomZ + 1), e = (it / (kMomZ + 1)) % NE, blk = it / ((kMomZ + 1) * NE);
        const DevAngular& A = plan->ang[ti][blk];
        int bb = 0;
        while ((bb + 1) * (bb + 2) / 2 <= blk) ++bb;
        const int a = blk - bb * (bb + 1) / 2;
        T acc = T(0);
        if (e < A.ne) {
            for (int c = 0; c < A.nc; ++c) {
                const int f = A.fmap[e * kMaxC + c];
                const int z = A.zint[c];
                if (f < 0 || pd > z) continue;
                T lp = T(1);
                for (int q = 0; q < pd; ++q) lp *= T(A.lambda[c]);
                acc += W[A.goff + f] * T(binom4(z, pd)) * lp * T(1.0 / (double)(1 << z));
            }
            acc *= T(A.norm);
            if (a == bb) acc *= T(0.5);
        }
        wtab[it] = acc;
    }
    __syncthreads();
    // L[s][e][q]
    for (int it = lane; it < ntypes * NE * kMomK; it += kWave) {
        const int q = it % kMomK, e = (it / kMomK) % NE, s = it / (kMomK * NE);
        const int pd = mono_deg(q);
        T acc = T(0);
        for (int b = 0; b < ntypes; ++b) {
            const int blk = b >= s ? b * (b + 1) / 2 + s : s * (s + 1) / 2 + b;
            const T wv = wtab[(blk * NE + e) * (kMomZ + 1) + pd];
            acc += (b == s ? T(2) : T(1)) * wv * Mom[(b * NE + e) * kMomK + q];
        }
        acc *= T(mono_mult(q));
        Lam[it] = acc;
        Lamp[it] = T(pd) * acc;
    }
    __syncthreads();
    // gradient coefficients: L_x[b] = (b_x + 1) L[b + e_x] for the 20 monomials b of degree <= 3
    for (int it = lane; it < ntypes * NE * kMomK3; it += kWave) {
        const int q = it % kMomK3, se = it / kMomK3;
        const int ax = kMono[q][0], ay = kMono[q][1], az = kMono[q][2];
        const T* L = Lam + se * kMomK;
        Lx[it] = T(ax + 1) * L[mono_index(ax + 1, ay, az)];
        Ly[it] = T(ay + 1) * L[mono_index(ax, ay + 1, az)];
        Lz[it] = T(az + 1) * L[mono_index(ax, ay, az + 1)];
    }
    __syncthreads();
    // entries: X_j -= dE/dr_j (angular part)
    const int ms = maxp + 1;  // the monomial table holds every entry
    for (int p = lane; p < P; p += kWave) {
        int s = 0;
        while (s + 1 < ntypes && p >= seg[s + 1]) ++s;
        const T ux = su[4 * p], uy = su[4 * p + 1], uz = su[4 * p + 2], di = T(1) / su[4 * p + 3];
        T gx = T(0), gy = T(0), gz = T(0);
#pragma unroll
        for (int e = 0; e < NE; ++e) {
            const int se = s * NE + e;
            const T* L = Lam + se * kMomK;
            const T* Lp = Lamp + se * kMomK;
            T g0 = T(0), g1 = T(0), hx = T(0), hy = T(0), hz = T(0);
            for (int q = 0; q < kMomK; ++q) {
                const T m = mono[q * ms + p];
                g0 += L[q] * m;
                g1 += Lp[q] * m;
                if (q < kMomK3) {
                    hx += Lx[se * kMomK3 + q] * m;
                    hy += Ly[se * kMomK3 + q] * m;
                    hz += Lz[se * kMomK3 + q] * m;
                }
            }
            // Delta_s(e) = sum_p w_{ss,e,p}
            const int blk = s * (s + 1) / 2 + s;
            T delta = T(0);
            for (int pd = 0; pd <= kMomZ; ++pd) delta += wtab[(blk * NE + e) * (kMomZ + 1) + pd];
            const T r = Rf[p * NE + e], rd = Rd[p * NE + e];
            const T dEdR = g0 - T(2) * delta * r;
            const T a1 = dEdR * rd - r * g1 * di;
            const T a2 = r * di;
            gx += a1 * ux + a2 * hx;
            gy += a1 * uy + a2 * hy;
            gz += a1 * uz + a2 * hz;
        }
        Xs[3 * p] -= gx;
        Xs[3 * p + 1] -= gy;
        Xs[3 * p + 2] -= gz;
    }
}

}  // namespace atomnn
