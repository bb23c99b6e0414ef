// Symmetry-function (descriptor) and force-term kernels.
//
// One wavefront owns one centre atom i and its P neighbour entries.  The entries are
// staged in LDS and partitioned by neighbour species (a stable partition), so that every
// angular block (a, b) of AtomNN's descriptor layout (nnh.cpp:95-121, LMat order
// lmat.hpp:43-46) is a contiguous triangle (a == b) or rectangle (a != b) of pairs.  Each
// block's pair space is cut into 64 contiguous, equal chunks, one per lane, walked row by
// row so a lane keeps its row atom's data in registers.
//
// Radial G2 (BasisR GAUSSIAN, basis_radial.cpp:302-310, 371-380, 443):
//   G[offR(b)+m] = norm * sum_j exp(-eta_m (d_j - rs_m)^2) fc(d_j)
// Angular G4 (BasisA GAUSS, basis_angular.cpp:406-421, 456-507, 635-640):
//   G[goff+m] = norm * sum_{j<k} |(1 + lambda_m cos_jk)/2|^zeta_m R_m(j) R_m(k),
//   R_m(j) = fc(d_j) exp(-ietap_m d_j^alpha)
// The block's functions are factorised into a grid of distinct radial exponents e and
// distinct angular terms c (ietap_e x (zeta_c, lambda_c)), so the exponential is evaluated
// once per neighbour and exponent instead of once per pair and function.  This changes the
// rounding of the sum, not its terms.
//
// Forces: with dE/dG from the network, the kernel computes for every entry j of centre i
//   X_ij = -dE_i/d r_ij   (r_ij = r_i - r_j - T, the entry's vector)
// summing the radial term and every pair (j, k) that contains j.  The reference's force
// loop (nnp.cpp:416-482) adds X_ij to F_i and subtracts it from F_j when the entry is the
// minimum image of j; the assembly kernel does exactly that from X, without atomics.
// Per pair (j, k) the kernel needs dE/dcos and dE/dd_j:
//   dE/dcos = sum_c A'_c(cos) sum_e W_ec R_e(j) R_e(k)
//   dE/dd_j = sum_c A_c(cos)  sum_e W_ec R'_e(j) R_e(k)
//   X_j    -= dE/dd_j u_j + dE/dcos (u_k - cos u_j) / d_j
// with W_ec = norm dE/dG[goff + fmap(e,c)].  A'_c = zeta lambda |cw|^(zeta-1) / 2 as in
// basis_angular.cpp:498-501 (no sign of cw, which is >= 0 up to rounding).
#include "device.hpp"
#include "model.hpp"
#include "basis.hpp"
#include "moments.hpp"

namespace atomnn {

// sin(pi x), cos(pi x): the reduction by pi is exact, so no large-argument path is needed
template <typename T>
__device__ __forceinline__ void sincos_pi(T x, T& s, T& c);
template <>
__device__ __forceinline__ void sincos_pi<double>(double x, double& s, double& c) {
    sincospi(x, &s, &c);
}
template <>
__device__ __forceinline__ void sincos_pi<float>(float x, float& s, float& c) {
    sincospif(x, &s, &c);
}

template <typename T>
__device__ __forceinline__ void cutoff_fg(int kind, T d, T rc, T rci, T pirci, T& f, T& g) {
    if (kind == CUT_COS) {  // basis.cpp:86-88, 106-108: cos(pi d / rc)
        if (d > rc) {
            f = T(0);
            g = T(0);
        } else {
            T sn, cs;
            sincos_pi<T>(d * rci, sn, cs);
            f = T(0.5) * (cs + T(1));
            g = T(-0.5) * pirci * sn;
        }
    } else if (kind == CUT_TANH) {  // basis.cpp:89-92, 109-114
        const T t = tanh(T(1) - d * rci);
        f = (d > rc) ? T(0) : t * t * t;
        g = (d < rc) ? T(-3) * t * t * (T(1) - t * t) * rci : T(0);
    } else {  // STEP, basis.cpp:83-85, 103-105
        f = (d > rc) ? T(0) : T(1);
        g = T(0);
    }
}

// e^x in the per-entry loops.  FP64: the library exp.  FP32 (the mixed form): the hardware 2^x
// (v_exp_f32) of x log2(e), two instructions instead of about ten: the rounding of x log2(e)
// adds a relative error of |x| 2^-24 at most (3e-6 at x = -50, where e^x is 2e-22), inside what
// the mixed form's 1 meV/atom gate allows by orders of magnitude.
template <typename T>
__device__ __forceinline__ T kexp(T x);
template <>
__device__ __forceinline__ double kexp<double>(double x) {
    return exp(x);
}
template <>
__device__ __forceinline__ float kexp<float>(float x) {
    return __expf(x);
}

// math/special.cpp:111-120 (n >= 0)
template <typename T>
__device__ __forceinline__ T ipow(T x, int n) {
    if (n == 0) return T(1);
    if (x == T(0)) return T(0);
    int nn = n > 0 ? n : -n;
    T yy = T(1), ww = x;
    for (; nn != 0; nn >>= 1, ww *= ww)
        if (nn & 1) yy *= ww;
    return n > 0 ? yy : T(1) / yy;
}

// x^n for a wave-uniform n >= 0: straight-line products for the usual small exponents (the
// same products repeated squaring forms), the loop beyond.
template <typename T>
__device__ __forceinline__ T ipow_u(T x, int n) {
    switch (n) {
        case 0: return T(1);
        case 1: return x;
        case 2: return x * x;
        case 3: return (x * x) * x;
        case 4: {
            const T y = x * x;
            return y * y;
        }
        case 5: {
            const T y = x * x;
            return x * (y * y);
        }
        case 7: {
            const T y = x * x;
            return (x * y) * (y * y);
        }
        case 8: {
            const T y = x * x, z = y * y;
            return z * z;
        }
        default: return ipow(x, n);
    }
}

template <typename T>
__device__ __forceinline__ T rpow(T x, T e) {
    return pow(x, e);
}

// |(1 + lambda cos)/2|^zeta (and the derivative factor zeta lambda |.|^(zeta-1) / 2).  The
// non-generic kernels only see integer zeta (zint >= 0): the general pow stays out of them.
// C(z, k) for 0 <= k <= z <= 4
__device__ __forceinline__ int c_binom4(int z, int k) {
    return z == 4 ? (k == 0 || k == 4 ? 1 : (k == 2 ? 6 : 4)) : (z == 3 ? (k == 0 || k == 3 ? 1 : 3) : (z == 2 && k == 1 ? 2 : 1));
}

// x^n for a wave-uniform 0 <= n <= 15 without branches: selects on the bits of n over the
// repeated-squaring powers.
template <typename T>
__device__ __forceinline__ T ipow_sel(T x, int n) {
    const T x2 = x * x, x4 = x2 * x2;
    T p = (n & 1) ? x : T(1);
    p = (n & 2) ? p * x2 : p;
    p = (n & 4) ? p * x4 : p;
    p = (n & 8) ? p * (x4 * x4) : p;
    return p;
}

// |(1 + lambda cos)/2|^zeta and the derivative factor zeta lambda |.|^(zeta-1) / 2.  The
// non-generic kernels see integer 0 <= zeta <= 16 only (larger or fractional zeta takes the
// generic instantiation), so they evaluate the terms without branches.
template <typename T, bool GEN>
__device__ __forceinline__ void angular_term(T cs, T lam, T zeta, int zint, T& A, T& dA) {
    const T acw = fabs(T(0.5) * (T(1) + lam * cs));
    T p1;
    if (!GEN) {
        p1 = zint > 0 ? ipow_sel(acw, zint - 1) : T(0);  // zeta = 0: A = 1, dA = 0
    } else {
        if (zint >= 1) p1 = ipow_u(acw, zint - 1);
        else if (zint == 0) p1 = T(0);
        else p1 = rpow(acw, zeta - T(1));
    }
    A = (zint == 0) ? T(1) : p1 * acw;
    dA = T(0.5) * zeta * lam * p1;
}

template <typename T, bool GEN>
__device__ __forceinline__ T angular_value(T cs, T lam, T zeta, int zint) {
    const T acw = fabs(T(0.5) * (T(1) + lam * cs));
    if (!GEN) return ipow_sel(acw, zint);
    return zint >= 0 ? ipow_u(acw, zint) : rpow(acw, zeta);
}

// Row index of flat pair q in the strict upper triangle of an n x n matrix (row-major).
__device__ __forceinline__ int tri_row(long long q, int n) {
    const double b = 2.0 * n - 1.0;
    int j = (int)floor((b - sqrt(b * b - 8.0 * (double)q)) * 0.5);
    if (j < 0) j = 0;
    auto start = [n](int r) { return (long long)r * (2LL * n - r - 1) / 2; };
    while (j + 1 < n && start(j + 1) <= q) ++j;
    while (j > 0 && start(j) > q) --j;
    return j;
}

struct Segs {
    int* start;  // [kMaxSpecies + 1] in LDS
};

// Stage the P entries of centre i into LDS, partitioned by neighbour species (stable).
// su[4p..4p+3] = (ux, uy, uz, d); orig[p] = entry index in the reference order.
template <typename T>
__device__ __forceinline__ void stage_entries(int i, int P, int maxp, int ntypes, const int* __restrict__ nbr_meta,
                                              const double4* __restrict__ nbr_r, T* su, int* orig, Segs& sg, int lane) {
    int count[kMaxSpecies];
#pragma unroll
    for (int t = 0; t < kMaxSpecies; ++t) count[t] = 0;
    const int64_t row = (int64_t)i * maxp;
    for (int p0 = 0; p0 < P; p0 += kWave) {
        const int p = p0 + lane;
        const int t = p < P ? ((nbr_meta[row + p] >> 1) & 127) : -1;
#pragma unroll
        for (int s = 0; s < kMaxSpecies; ++s)
            if (s < ntypes) count[s] += __popcll(__ballot(t == s));
    }
    int fill[kMaxSpecies];
    int run = 0;
#pragma unroll
    for (int s = 0; s < kMaxSpecies; ++s) {
        fill[s] = run;
        if (lane == 0) sg.start[s] = run;
        run += (s < ntypes ? count[s] : 0);
    }
    if (lane == 0) sg.start[kMaxSpecies] = run;
    const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int p0 = 0; p0 < P; p0 += kWave) {
        const int p = p0 + lane;
        int t = -1;
        double4 r = make_double4(0, 0, 0, 1);
        if (p < P) {
            t = (nbr_meta[row + p] >> 1) & 127;
            r = nbr_r[row + p];
        }
        int dst = -1;
#pragma unroll
        for (int s = 0; s < kMaxSpecies; ++s) {
            if (s < ntypes) {
                const unsigned long long m = __ballot(t == s);
                if (t == s) dst = fill[s] + __popcll(m & lt);
                fill[s] += __popcll(m);
            }
        }
        if (p < P) {
            const double inv = 1.0 / r.w;
            su[4 * dst + 0] = T(r.x * inv);
            su[4 * dst + 1] = T(r.y * inv);
            su[4 * dst + 2] = T(r.z * inv);
            su[4 * dst + 3] = T(r.w);
            if (orig) orig[dst] = p;
        }
    }
}

// Radial exponent factors of one entry at distance d for an angular block with a separable
// distance factor, given the block's cutoff value fc and derivative gc: R_e = fc h_e,
// R'_e = dR_e/dd (when Rd is given).
template <typename T, int NE, bool GEN>
__device__ __forceinline__ void ang_sep_entry(const DevAngular& A, T d, T fc, T gc, T* Rf, T* Rd) {
    const T dp = ipow(d, A.alpha);
    const T ddp = T(A.alpha) * dp / d;  // d(d^alpha)/dd
    if (!GEN || A.phi == ANG_GAUSS) {
#pragma unroll
        for (int e = 0; e < NE; ++e) {
            const T ie = T(A.ietap[e]);
            const T ex = kexp(-ie * dp);
            Rf[e] = fc * ex;
            if (Rd) Rd[e] = (gc - ie * ddp * fc) * ex;
        }
    } else {  // IPOWP, SECHP: leg factor h(ietap d^alpha)
#pragma unroll
        for (int e = 0; e < NE; ++e) {
            const T ie = T(A.ietap[e]);
            T h, dh;
            ang_leg<T>(A.phi, ie * dp, h, dh);
            Rf[e] = fc * h;
            if (Rd) Rd[e] = gc * h + fc * dh * ie * ddp;
        }
    }
}

// Radial exponent tables of one angular block for the slots [lo, hi): Rf (and Rd).
template <typename T, int NE, bool GEN>
__device__ __forceinline__ void radial_tables(const DevAngular& A, const T* su, int lo, int hi, T* Rf, T* Rd, int lane) {
    const T rc = T(A.rc), rci = T(A.rci), pirci = T(A.pirci);
    for (int p = lo + lane; p < hi; p += kWave) {
        const T d = su[4 * p + 3];
        T fc, gc;
        cutoff_fg<T>(A.cut, d, rc, rci, pirci, fc, gc);
        ang_sep_entry<T, NE, GEN>(A, d, fc, gc, Rf + p * NE, Rd ? Rd + p * NE : nullptr);
    }
}

// Entry tables of a non-separable angular block (IPOWS, SECHS) for the slots [lo, hi):
// Rf[p][e] = a(ietap_e d^alpha), Rd[p][e] = ietap_e d(d^alpha)/dd, nsx[2p] = fc, nsx[2p+1] = fc'.
template <typename T, int NE>
__device__ __forceinline__ void ns_tables(const DevAngular& A, const T* su, int lo, int hi, T* Rf, T* Rd, T* nsx,
                                          int lane) {
    const T rc = T(A.rc), rci = T(A.rci), pirci = T(A.pirci);
    for (int p = lo + lane; p < hi; p += kWave) {
        const T d = su[4 * p + 3];
        T fc, gc;
        cutoff_fg<T>(A.cut, d, rc, rci, pirci, fc, gc);
        const T dp = ipow(d, A.alpha);
        const T ddp = T(A.alpha) * dp / d;
#pragma unroll
        for (int e = 0; e < NE; ++e) {
            const T ie = T(A.ietap[e]);
            Rf[p * NE + e] = ang_entry<T>(A.phi, ie * dp);
            if (Rd) Rd[p * NE + e] = ie * ddp;
        }
        nsx[2 * p] = fc;
        nsx[2 * p + 1] = gc;
    }
}

// One entry's force term (x, y, z, 0) as a single 32-byte (FP64) or 16-byte (FP32) store.
template <typename T>
__device__ __forceinline__ void store_x4(T* dst, T x, T y, T z);
template <>
__device__ __forceinline__ void store_x4<double>(double* dst, double x, double y, double z) {
    *reinterpret_cast<double4*>(dst) = make_double4(x, y, z, 0.0);
}
template <>
__device__ __forceinline__ void store_x4<float>(float* dst, float x, float y, float z) {
    *reinterpret_cast<float4*>(dst) = make_float4(x, y, z, 0.f);
}

template <typename T>
__device__ __forceinline__ double4 load_x4(const T* src);
template <>
__device__ __forceinline__ double4 load_x4<double>(const double* src) {
    return *reinterpret_cast<const double4*>(src);
}
template <>
__device__ __forceinline__ double4 load_x4<float>(const float* src) {
    const float4 v = *reinterpret_cast<const float4*>(src);
    return make_double4(v.x, v.y, v.z, 0.0);
}

// ------------------------------------------------------------------ per-atom bodies
// Descriptors of one staged centre into Grow (LDS, zeroed by the caller).
template <typename T, int NE, int NC, bool GEN>
__device__ __forceinline__ void symm_body(const DevPlan* __restrict__ plan, const double* __restrict__ rfun_eta,
                                          const double* __restrict__ rfun_rs, int ti, int ntypes, int P, int maxp,
                                          const T* su, T* Rf, T* red, int red_elems, T* Grow, const Segs& sg,
                                          T* nsx, T* mom, int lane, T* mom_out = nullptr) {
    // radial: cutoff once per neighbour, then (function, part) items over 64 lanes; the parts of
    // a function are summed in a fixed order
    {
        T* fcr = Rf;    // [maxp] cutoff values (Rf is free until the angular blocks)
        T* part = red;  // [64] partial sums
        for (int b = 0; b < ntypes; ++b) {
            const DevRadial& R = plan->rad[ti][b];
            for (int p = sg.start[b] + lane; p < sg.start[b + 1]; p += kWave) {
                T fc, gc;
                cutoff_fg<T>(R.cut, su[4 * p + 3], T(R.rc), T(R.rci), T(R.pirci), fc, gc);
                fcr[p] = fc;
            }
        }
        __syncthreads();
        int nrt = 0;
        for (int b = 0; b < ntypes; ++b) nrt += plan->rad[ti][b].n;
        const int nparts = nrt > 0 ? max(1, kWave / nrt) : 1;
        for (int f0 = 0; f0 < nrt; f0 += kWave / nparts) {
            const int item = lane;
            const int f = f0 + item / nparts, prt = item % nparts;
            T acc = T(0);
            int goff = -1;
            T norm = T(0);
            if (item < (kWave / nparts) * nparts && f < nrt) {
                int b = 0, fl = f;
                while (fl >= plan->rad[ti][b].n) fl -= plan->rad[ti][b++].n;
                const DevRadial& R = plan->rad[ti][b];
                const T eta = T(rfun_eta[R.fbase + fl]), rs = T(rfun_rs[R.fbase + fl]);
                if (!GEN || R.phi == RAD_GAUSSIAN) {
                    for (int p = sg.start[b] + prt; p < sg.start[b + 1]; p += nparts) {
                        const T x = su[4 * p + 3] - rs;
                        acc += kexp(-eta * x * x) * fcr[p];
                    }
                } else {
                    for (int p = sg.start[b] + prt; p < sg.start[b + 1]; p += nparts) {
                        T v, dv;
                        rad_phi<T>(R.phi, -eta * (su[4 * p + 3] - rs), v, dv);
                        acc += v * fcr[p];
                    }
                }
                goff = R.goff + fl;
                norm = T(R.norm);
            }
            part[lane] = acc;
            __syncthreads();
            if (goff >= 0 && prt == 0) {
                T s = T(0);
                for (int q = 0; q < nparts; ++q) s += part[lane + q];
                Grow[goff] = s * norm;
            }
            __syncthreads();
        }
    }

    // angular blocks: moment expansion when the centre species allows it
    if (mom && plan->mom_ok[ti]) {
        // the power table goes through the reduction buffer `red` in chunks of `ch` entries
        const int ch = red_elems / kMomPw;
        T* Mom = mom;
        T* D = Mom + ntypes * NE * kMomK;
        T* S = D + ntypes * NE;
        radial_tables<T, NE, GEN>(plan->ang[ti][0], su, 0, P, Rf, nullptr, lane);
        __syncthreads();
        build_moments<T, NE>(su, Rf, P, ch, ntypes, sg.start, red, Mom, D, lane);
        // the centre's moment table for the force-term kernel of the same pass, which then skips
        // its own moment sums
        if (mom_out)
            for (int it = lane; it < ntypes * NE * kMomK; it += kWave) mom_out[it] = Mom[it];
        symm_angular_moments<T, NE>(plan, ti, ntypes, Mom, D, S, Grow, lane);
        __syncthreads();
        return;
    }
    // angular blocks in LMat order, pair by pair
    const bool shared = plan->ang_shared[ti];
    if (shared) radial_tables<T, NE, GEN>(plan->ang[ti][0], su, 0, P, Rf, nullptr, lane);
    for (int bb = 0; bb < ntypes; ++bb) {
        for (int a = 0; a <= bb; ++a) {
            const DevAngular& A = plan->ang[ti][bb * (bb + 1) / 2 + a];
            const int sa = sg.start[a], na = sg.start[a + 1] - sa;
            const int sb = sg.start[bb], nb = sg.start[bb + 1] - sb;
            const bool tri = (a == bb);
            const long long Q = tri ? (long long)na * (na - 1) / 2 : (long long)na * nb;
            if (Q == 0) continue;  // block stays zero
            const bool nsb = GEN && !ang_separable(A.phi);
            if (nsb) {
                ns_tables<T, NE>(A, su, sa, sa + na, Rf, nullptr, nsx, lane);
                if (!tri) ns_tables<T, NE>(A, su, sb, sb + nb, Rf, nullptr, nsx, lane);
            } else if (!shared) {
                radial_tables<T, NE, GEN>(A, su, sa, sa + na, Rf, nullptr, lane);
                if (!tri) radial_tables<T, NE, GEN>(A, su, sb, sb + nb, Rf, nullptr, lane);
            }
            __syncthreads();
            T lam[NC], zet[NC];
            int zin[NC];
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                lam[c] = T(A.lambda[c]);
                zet[c] = T(A.zeta[c]);
                zin[c] = A.zint[c];
            }
            T acc[NE][NC];
#pragma unroll
            for (int e = 0; e < NE; ++e)
#pragma unroll
                for (int c = 0; c < NC; ++c) acc[e][c] = T(0);
            const long long L = (Q + kWave - 1) / kWave;
            const long long q0 = (long long)lane * L, q1 = min(q0 + L, Q);
            if (q0 < q1) {
                int j, k;
                if (tri) {
                    j = tri_row(q0, na);
                    k = j + 1 + (int)(q0 - (long long)j * (2LL * na - j - 1) / 2);
                } else {
                    j = (int)(q0 / nb);
                    k = (int)(q0 % nb);
                }
                long long q = q0;
                while (q < q1) {
                    const int kend = tri ? na : nb;
                    const int n = (int)min((long long)(kend - k), q1 - q);
                    const int pj = sa + j;
                    const T ujx = su[4 * pj], ujy = su[4 * pj + 1], ujz = su[4 * pj + 2];
                    T rj[NE];
#pragma unroll
                    for (int e = 0; e < NE; ++e) rj[e] = Rf[pj * NE + e];
                    const T fcj = nsb ? nsx[2 * pj] : T(0);
                    for (int t = 0; t < n; ++t) {
                        const int pk = sb + k + t;
                        const T cs = ujx * su[4 * pk] + ujy * su[4 * pk + 1] + ujz * su[4 * pk + 2];
                        T Ac[NC];
#pragma unroll
                        for (int c = 0; c < NC; ++c) Ac[c] = angular_value<T, GEN>(cs, lam[c], zet[c], zin[c]);
                        const T fcjk = nsb ? fcj * nsx[2 * pk] : T(0);
#pragma unroll
                        for (int e = 0; e < NE; ++e) {
                            T pe;
                            if (nsb) {
                                T F, tt;
                                ang_pair<T>(A.phi, rj[e], Rf[pk * NE + e], F, tt);
                                pe = fcjk * F;
                            } else {
                                pe = rj[e] * Rf[pk * NE + e];
                            }
#pragma unroll
                            for (int c = 0; c < NC; ++c) acc[e][c] += Ac[c] * pe;
                        }
                    }
                    q += n;
                    ++j;
                    k = tri ? j + 1 : 0;
                }
            }
#pragma unroll
            for (int e = 0; e < NE; ++e)
#pragma unroll
                for (int c = 0; c < NC; ++c) red[(e * NC + c) * kWave + lane] = acc[e][c];
            __syncthreads();
            for (int t = lane; t < NE * NC; t += kWave) {
                const int e = t / NC, c = t % NC;
                const int f = A.fmap[e * kMaxC + c];
                if (f >= 0) {
                    T s = T(0);
                    for (int l = 0; l < kWave; ++l) s += red[t * kWave + l];
                    Grow[A.goff + f] = s * T(A.norm);
                }
            }
            __syncthreads();
        }
    }
}

// Force terms of one staged centre into Xs (LDS).  Wrad: dE/dG row in LDS (radial weights);
// Wang: the same row in global memory (uniform, scalar loads) or LDS.
template <typename T, int NE, int NC, bool GEN>
__device__ __forceinline__ void force_body(const DevPlan* __restrict__ plan, const double* __restrict__ rfun_eta,
                                           const double* __restrict__ rfun_rs, int ti, int ntypes, int P, int maxp,
                                           const T* su, T* Rf, T* Rd, T* Xs, T* slots, int* slot_row, const T* W,
                                           const T* Wang, const Segs& sg, T* nsx, T* mom, int lane) {
    // radial term, one lane per entry: X_j = -dE/dd_j u_j; with one exponent table for every
    // angular block, the entry's angular factors R_e, R'_e too (the cutoff is shared when the
    // angular basis uses the radial one's)
    const bool shared = plan->ang_shared[ti];
    const DevAngular& A0 = plan->ang[ti][0];
    for (int p = lane; p < P; p += kWave) {
        int b = 0;
        while (b + 1 < ntypes && p >= sg.start[b + 1]) ++b;
        const DevRadial& R = plan->rad[ti][b];
        const T d = su[4 * p + 3];
        T fc, gc;
        cutoff_fg<T>(R.cut, d, T(R.rc), T(R.rci), T(R.pirci), fc, gc);
        if (shared) {
            T fa = fc, ga = gc;
            if (A0.cut != R.cut || A0.rc != R.rc) cutoff_fg<T>(A0.cut, d, T(A0.rc), T(A0.rci), T(A0.pirci), fa, ga);
            ang_sep_entry<T, NE, GEN>(A0, d, fa, ga, Rf + p * NE, Rd + p * NE);
        }
        T s = T(0);
        if (!GEN || R.phi == RAD_GAUSSIAN) {
            for (int f = 0; f < R.n; ++f) {
                const T eta = T(rfun_eta[R.fbase + f]), rs = T(rfun_rs[R.fbase + f]);
                const T x = d - rs;
                s += W[R.goff + f] * kexp(-eta * x * x) * (T(-2) * eta * x * fc + gc);
            }
        } else {
            // d/dd [phi(-eta (d - rs)) fc(d)] = -eta phi' fc + phi fc'
            for (int f = 0; f < R.n; ++f) {
                const T eta = T(rfun_eta[R.fbase + f]), rs = T(rfun_rs[R.fbase + f]);
                T v, dv;
                rad_phi<T>(R.phi, -eta * (d - rs), v, dv);
                s += W[R.goff + f] * (-eta * dv * fc + v * gc);
            }
        }
        s *= T(R.norm);
        Xs[3 * p + 0] = -s * su[4 * p];
        Xs[3 * p + 1] = -s * su[4 * p + 1];
        Xs[3 * p + 2] = -s * su[4 * p + 2];
    }
    __syncthreads();

    if (mom && plan->mom_ok[ti]) {
        // mom_ok implies shared tables: Rf, Rd are complete
        T* Mom = mom;
        T* Delta = Mom + ntypes * NE * kMomK;
        T* wtab = Delta + ntypes * NE;
        T* U = wtab + (ntypes * (ntypes + 1) / 2) * NE * (kMomZ + 1);
        build_moments<T, NE>(su, Rf, P, kMomCh, ntypes, sg.start, U, Mom, nullptr, lane);
        adjoint_moments<T, NE>(plan, ti, ntypes, W, Mom, 1, 0, Delta, wtab, U, lane, kWave);
        force_from_moments<T, NE>(ntypes, P, su, Rf, Rd, sg.start, U, Delta, Xs, lane);
        __syncthreads();
        return;
    }
    for (int bb = 0; bb < ntypes; ++bb) {
        for (int a = 0; a <= bb; ++a) {
            const DevAngular& A = plan->ang[ti][bb * (bb + 1) / 2 + a];
            const int sa = sg.start[a], na = sg.start[a + 1] - sa;
            const int sb = sg.start[bb], nb = sg.start[bb + 1] - sb;
            const bool same = (a == bb);
            if (same ? (na < 2) : (na == 0 || nb == 0)) continue;
            const bool nsb = GEN && !ang_separable(A.phi);
            if (nsb) {
                ns_tables<T, NE>(A, su, sa, sa + na, Rf, Rd, nsx, lane);
                if (!same) ns_tables<T, NE>(A, su, sb, sb + nb, Rf, Rd, nsx, lane);
            } else if (!shared) {
                radial_tables<T, NE, GEN>(A, su, sa, sa + na, Rf, Rd, lane);
                if (!same) radial_tables<T, NE, GEN>(A, su, sb, sb + nb, Rf, Rd, lane);
            }
            // block weights W_ec = dE/dG[goff + fmap(e,c)] (uniform; the block norm scales each
            // row's sum).  Fast kernels (integer 0 <= zeta <= 4, |lambda| <= 1 in every block):
            // for each exponent e the angular sum  S_e(cos) = sum_c W_ec ((1 + lambda_c cos)/2)^zeta_c
            // is a polynomial of degree <= 4, P[e][p] = sum_c W_ec 2^-zeta C(zeta,p) lambda^p, and
            // a pair needs S_e and S_e' by Horner's rule.  Generic kernels: the W_ec and the terms.
            const T* dEdGi = Wang + A.goff;
            const T bnorm = T(A.norm);
            T w[GEN ? NE : 1][GEN ? NC : 1];
            T lam[GEN ? NC : 1], zet[GEN ? NC : 1];
            int zin[GEN ? NC : 1];
            T P[GEN ? 1 : NE][5];
            if constexpr (GEN) {
#pragma unroll
                for (int e = 0; e < NE; ++e)
#pragma unroll
                    for (int c = 0; c < NC; ++c) {
                        const int f = A.fmap[e * kMaxC + c];
                        w[e][c] = f >= 0 ? dEdGi[f] : T(0);
                    }
#pragma unroll
                for (int c = 0; c < NC; ++c) {
                    lam[c] = T(A.lambda[c]);
                    zet[c] = T(A.zeta[c]);
                    zin[c] = A.zint[c];
                }
            } else {
#pragma unroll
                for (int e = 0; e < NE; ++e) {
#pragma unroll
                    for (int k = 0; k < 5; ++k) P[e][k] = T(0);
#pragma unroll
                    for (int c = 0; c < NC; ++c) {
                        const int f = A.fmap[e * kMaxC + c];
                        if (f < 0) continue;
                        const int z = A.zint[c];
                        const T lc = T(A.lambda[c]);
                        T coef = dEdGi[f] * T(1.0 / (double)(1 << z));
#pragma unroll
                        for (int k = 0; k < 5; ++k) {
                            if (k <= z) P[e][k] += coef * T(c_binom4(z, k));
                            coef *= lc;
                        }
                    }
                }
            }
            __syncthreads();
            // ordered pairs: rows in one segment, columns in the other (or the same, minus the diagonal)
            for (int pass = 0; pass < (same ? 1 : 2); ++pass) {
                const int rs0 = pass == 0 ? sa : sb, nrow = pass == 0 ? na : nb;
                const int cs0 = pass == 0 ? sb : sa, ncol = pass == 0 ? nb : na;
                const int ncol_eff = same ? ncol - 1 : ncol;
                const long long Q = (long long)nrow * ncol_eff;
                const long long L = (Q + kWave - 1) / kWave;
                const long long q0 = (long long)lane * L, q1 = min(q0 + L, Q);
                slot_row[lane] = -1;
                slot_row[kWave + lane] = -1;
                if (q0 < q1) {
                    int r = (int)(q0 / ncol_eff), c = (int)(q0 % ncol_eff);
                    long long q = q0;
                    while (q < q1) {
                        const int n = (int)min((long long)(ncol_eff - c), q1 - q);
                        const bool whole = (c == 0 && n == ncol_eff);
                        const int pj = rs0 + r;
                        const T ujx = su[4 * pj], ujy = su[4 * pj + 1], ujz = su[4 * pj + 2];
                        const T dji = T(1) / su[4 * pj + 3];
                        T rfj[NE], rdj[NE];
#pragma unroll
                        for (int e = 0; e < NE; ++e) {
                            rfj[e] = Rf[pj * NE + e];
                            rdj[e] = Rd[pj * NE + e];
                        }
                        T xx = T(0), xy = T(0), xz = T(0);
                        const T fcj = nsb ? nsx[2 * pj] : T(0), gcj = nsb ? nsx[2 * pj + 1] : T(0);
                        for (int t = 0; t < n; ++t) {
                            const int cc0 = c + t;
                            const int k = (same && cc0 >= r) ? cc0 + 1 : cc0;
                            const int pk = cs0 + k;
                            const T ukx = su[4 * pk], uky = su[4 * pk + 1], ukz = su[4 * pk + 2];
                            const T cs = ujx * ukx + ujy * uky + ujz * ukz;
                            T pf[NE], pd[NE];
                            if (nsb) {
                                // F(x_j, x_k) fc_j fc_k and its d/dd_j: fc_k F (fc_j' - fc_j t dx_j/dd_j)
                                const T fck = nsx[2 * pk];
#pragma unroll
                                for (int e = 0; e < NE; ++e) {
                                    T F, tt;
                                    ang_pair<T>(A.phi, rfj[e], Rf[pk * NE + e], F, tt);
                                    pf[e] = fcj * fck * F;
                                    pd[e] = fck * F * (gcj - fcj * tt * rdj[e]);
                                }
                            } else {
#pragma unroll
                                for (int e = 0; e < NE; ++e) {
                                    const T rfk = Rf[pk * NE + e];
                                    pf[e] = rfj[e] * rfk;
                                    pd[e] = rdj[e] * rfk;
                                }
                            }
                            T dEdcos = T(0), dEdd = T(0);
                            if constexpr (GEN) {
#pragma unroll
                                for (int cc = 0; cc < NC; ++cc) {
                                    T tp = T(0), tq = T(0);
#pragma unroll
                                    for (int e = 0; e < NE; ++e) {
                                        tp += w[e][cc] * pf[e];
                                        tq += w[e][cc] * pd[e];
                                    }
                                    T Av, dAv;
                                    angular_term<T, GEN>(cs, lam[cc], zet[cc], zin[cc], Av, dAv);
                                    dEdcos += dAv * tp;
                                    dEdd += Av * tq;
                                }
                            } else {
#pragma unroll
                                for (int e = 0; e < NE; ++e) {
                                    T b = P[e][4], d = T(0);
#pragma unroll
                                    for (int k = 3; k >= 0; --k) {
                                        d = d * cs + b;
                                        b = b * cs + P[e][k];
                                    }
                                    dEdd += b * pd[e];
                                    dEdcos += d * pf[e];
                                }
                            }
                            const T g = dEdcos * dji;
                            xx -= dEdd * ujx + g * (ukx - cs * ujx);
                            xy -= dEdd * ujy + g * (uky - cs * ujy);
                            xz -= dEdd * ujz + g * (ukz - cs * ujz);
                        }
                        xx *= bnorm;
                        xy *= bnorm;
                        xz *= bnorm;
                        if (whole) {
                            Xs[3 * pj] += xx;
                            Xs[3 * pj + 1] += xy;
                            Xs[3 * pj + 2] += xz;
                        } else {
                            const int sl = (q == q0) ? 0 : 1;
                            slot_row[sl * kWave + lane] = pj;
                            slots[(sl * kWave + lane) * 3 + 0] = xx;
                            slots[(sl * kWave + lane) * 3 + 1] = xy;
                            slots[(sl * kWave + lane) * 3 + 2] = xz;
                        }
                        q += n;
                        ++r;
                        c = 0;
                    }
                }
                __syncthreads();
                // rows split between lanes: add the partial sums in lane order
                for (int r = lane; r < nrow; r += kWave) {
                    const int pj = rs0 + r;
                    const long long lo = (long long)r * ncol_eff, hi = lo + ncol_eff - 1;
                    const int lf = (int)(lo / L), ll = (int)min((long long)(kWave - 1), hi / L);
                    for (int l = lf; l <= ll; ++l) {
#pragma unroll
                        for (int sl = 0; sl < 2; ++sl) {
                            if (slot_row[sl * kWave + l] == pj) {
                                Xs[3 * pj] += slots[(sl * kWave + l) * 3 + 0];
                                Xs[3 * pj + 1] += slots[(sl * kWave + l) * 3 + 1];
                                Xs[3 * pj + 2] += slots[(sl * kWave + l) * 3 + 2];
                            }
                        }
                    }
                }
                __syncthreads();
            }
        }
    }
}

// LDS sizes (bytes) of the kernels for a given maxp.
// Each kernel's LDS: its own arrays, then (moment path) the moment region of mom_lds_elems.
// GEN: the kernels for models beyond GAUSSIAN / GAUSS / tanh with integer zeta (other radial
// functions, angular distance factors, activations, non-integer zeta); they add the per-entry
// cutoff table nsx [maxp][2] that the non-separable angular factors use.  GEN = false keeps the
// common case's register budget.
template <typename T, int NE, int NC, bool GEN>
__host__ __device__ constexpr size_t symm_lds_bytes(int maxp, int ldg, int red_elems = kWave * NE * NC) {
    return sizeof(T) * (size_t)(4 * maxp + NE * maxp + red_elems) + sizeof(T) * (size_t)((ldg + 1) & ~1) +
           (GEN ? sizeof(T) * 2 * (size_t)maxp : 0);
}
template <typename T, int NE, int NC, bool GEN>
__host__ __device__ constexpr size_t force_lds_bytes(int maxp, int ldg) {
    return sizeof(T) * (size_t)(4 * maxp + 2 * NE * maxp + 3 * maxp + 6 * kWave + ((ldg + 1) & ~1)) +
           sizeof(int) * (size_t)(maxp + ((2 * kWave + 3) & ~3)) + (GEN ? sizeof(T) * 2 * (size_t)maxp : 0);
}
// bits: 1 descriptors, 2 forces (the fused kernel's region serves both phases)
template <typename T, int NE>
__host__ __device__ constexpr size_t mom_lds_bytes(int lmaxp, int ntypes, int bits = 3) {
    (void)lmaxp;
    return sizeof(T) * (size_t)((bits & 2) ? mom_lds_elems<NE>(ntypes) : mom_symm_elems<NE>(ntypes));
}

// ------------------------------------------------------------------ descriptors
template <typename T, int NE, int NC, bool GEN>
__global__ __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(NE * NC <= 16 ? (GEN ? 2 : 3) : 1, 8))) void k_symm(
    const DevPlan* __restrict__ plan, const double* __restrict__ rfun_eta, const double* __restrict__ rfun_rs,
    const int* __restrict__ atom_type, const int* __restrict__ centres, int ncentres, const int* __restrict__ cnt,
    const int* __restrict__ nbr_meta, const double4* __restrict__ nbr_r, int maxp, int lmaxp, T* __restrict__ G,
    int ldg, int use_mom, int red_elems, T* __restrict__ Mout) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    T* su = (T*)lds_raw;
    T* Rf = su + 4 * lmaxp;
    T* red = Rf + NE * lmaxp;  // red_elems: kWave NE NC, or less when every species takes the moments
    T* Grow = red + red_elems;
    T* nsx = Grow + ((ldg + 1) & ~1);
    T* mom = use_mom ? nsx + (GEN ? 2 * lmaxp : 0) : nullptr;
    // centres in the XCD-local order of the neighbour kernel, which wrote this centre's row from a
    // block on the same XCD (its lines can still be in that L2)
    const int slot = xcd_slot(ncentres);
    if (slot < 0) return;
    const int i = centres[slot];
    const int lane = threadIdx.x;
    const int ti = atom_type[i];
    const int P = min(cnt[i], lmaxp);
    const int ntypes = plan->T;
    __shared__ int seg_lds[kMaxSpecies + 1];
    Segs sg{seg_lds};
    stage_entries<T>(i, P, maxp, ntypes, nbr_meta, nbr_r, su, nullptr, sg, lane);
    for (int k = lane; k < ldg; k += kWave) Grow[k] = T(0);
    __syncthreads();
    symm_body<T, NE, NC, GEN>(plan, rfun_eta, rfun_rs, ti, ntypes, P, lmaxp, su, Rf, red, red_elems, Grow, sg, nsx,
                              mom, lane, Mout ? Mout + (int64_t)i * (ntypes * NE * kMomK) : nullptr);
    __syncthreads();
    for (int k = lane; k < ldg; k += kWave) G[(int64_t)i * ldg + k] = Grow[k];
}

// ------------------------------------------------------------------ force terms
template <typename T, int NE, int NC, bool GEN>
__global__ __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(NE * NC <= 16 ? (GEN ? 2 : 3) : 1, 8))) void k_force(
    const DevPlan* __restrict__ plan, const double* __restrict__ rfun_eta, const double* __restrict__ rfun_rs,
    const int* __restrict__ atom_type, const int* __restrict__ centres, int ncentres, const int* __restrict__ cnt,
    const int* __restrict__ nbr_meta, const double4* __restrict__ nbr_r, int maxp, int lmaxp,
    const T* __restrict__ dEdG, int ldg, T* __restrict__ Xout, int use_mom) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    T* su = (T*)lds_raw;
    T* Rf = su + 4 * lmaxp;
    T* Rd = Rf + NE * lmaxp;
    T* Xs = Rd + NE * lmaxp;          // [lmaxp][3]
    T* slots = Xs + 3 * lmaxp;        // [2][kWave][3]
    T* W = slots + 6 * kWave;         // dE/dG row
    int* orig = (int*)(W + ((ldg + 1) & ~1));
    int* slot_row = orig + lmaxp;     // [2][kWave]
    T* nsx = (T*)(slot_row + ((2 * kWave + 3) & ~3));
    T* mom = use_mom ? nsx + (GEN ? 2 * lmaxp : 0) : nullptr;
    const int slot = xcd_slot(ncentres);  // the XCD-local order of the neighbour kernel
    if (slot < 0) return;
    const int i = centres[slot];
    const int lane = threadIdx.x;
    const int ti = atom_type[i];
    const int P = min(cnt[i], lmaxp);
    const int ntypes = plan->T;
    __shared__ int seg_lds[kMaxSpecies + 1];
    Segs sg{seg_lds};
    stage_entries<T>(i, P, maxp, ntypes, nbr_meta, nbr_r, su, orig, sg, lane);
    for (int k = lane; k < ldg; k += kWave) W[k] = dEdG[(int64_t)i * ldg + k];
    __syncthreads();
    force_body<T, NE, NC, GEN>(plan, rfun_eta, rfun_rs, ti, ntypes, P, lmaxp, su, Rf, Rd, Xs, slots, slot_row, W,
                              dEdG + (int64_t)i * ldg, sg, nsx, mom, lane);
    const int64_t row = (int64_t)i * maxp;
    for (int p = lane; p < P; p += kWave) {
        const int o = orig[p];
        store_x4<T>(Xout + 4 * (row + o), Xs[3 * p], Xs[3 * p + 1], Xs[3 * p + 2]);
    }
}

// ------------------------------------------------------------------ descriptor tangents
// Adjoint of the force pass with respect to dE/dG (force training).  Given gF[k] = dL/dF_k, the
// loss depends on entry j of centre i through X_ij (F_i += X_ij, and F_j -= X_ij when the entry
// carries the reaction), so phi_ij = dL/dX_ij = gF_i - [reaction] gF_j, and since
// X_ij = -sum_g dE/dG_g dG_g/dr_ij:
//   V_i = dL/d(dE_i/dG) = -sum_j phi_ij . dG_i/dr_ij,
// the directional derivative of the centre's descriptors along the entry displacements phi_ij,
// negated.  With d_j = |r_j|, u_j = r_j / d_j: d_j' = u_j . phi_j, u_j' = (phi_j - d_j' u_j) / d_j,
// cos_jk' = u_k . w_j + u_j . w_k - cos_jk (s_j + s_k) with w = phi / d, s = d' / d.
// Radial: G_f' = norm sum_j d_j' (phi_f fc)'(d_j).  Angular, per pair and (exponent e, term c):
// (A_c R_e(j) R_e(k))' = A_c' cos' R_e(j) R_e(k) + A_c (R_e'(j) d_j' R_e(k) + R_e(j) R_e'(k) d_k').
// One wavefront per centre, the same staging, tables and fixed-order lane reductions as k_symm.
// Moment form (centre species with mom_ok): the tangents of the moments instead of the pair
// loops, O(35 NE) per entry; LDS adds Mom, dMom [S][NE][35], dD [S][NE] and S [blocks][NE][5].
template <int NE>
__host__ __device__ constexpr int jvp_mom_elems(int ntypes) {
    return 2 * ntypes * NE * kMomK + ntypes * NE + (ntypes * (ntypes + 1) / 2) * NE * (kMomZ + 1);
}
template <int NE, int NC, bool GEN>
__host__ __device__ constexpr size_t jvp_lds_bytes(int maxp, int ldg, int mom_elems = 0) {
    return 8 * (size_t)(8 * maxp + 2 * NE * maxp + kWave * NE * NC + ((ldg + 1) & ~1) + (GEN ? 2 * maxp : 0) +
                        mom_elems) +
           4 * (size_t)((maxp + 3) & ~3);
}

template <int NE, int NC, bool GEN>
__global__ __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(1, 8))) void k_symm_jvp(
    const DevPlan* __restrict__ plan, const double* __restrict__ rfun_eta, const double* __restrict__ rfun_rs,
    const int* __restrict__ atom_type, const DevStruc* __restrict__ S, const int* __restrict__ atom_struc,
    const int* __restrict__ cnt, const int* __restrict__ nbr_j, const int* __restrict__ nbr_meta,
    const double4* __restrict__ nbr_r, int maxp, int lmaxp, const double* __restrict__ gF, int all_images,
    double* __restrict__ V, int ldg, int use_mom) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    double* su = (double*)lds_raw;
    double* tg = su + 4 * lmaxp;  // (w_x, w_y, w_z, d')
    double* Rf = tg + 4 * lmaxp;
    double* Rd = Rf + NE * lmaxp;
    double* red = Rd + NE * lmaxp;
    double* Vrow = red + kWave * NE * NC;
    double* nsx = Vrow + ((ldg + 1) & ~1);
    double* momr = nsx + (GEN ? 2 * lmaxp : 0);
    int* orig = (int*)(momr + (use_mom ? jvp_mom_elems<NE>(plan->T) : 0));
    const int i = blockIdx.x;
    const int lane = threadIdx.x;
    const int ti = atom_type[i];
    const int P = min(cnt[i], lmaxp);
    const int ntypes = plan->T;
    __shared__ int seg_lds[kMaxSpecies + 1];
    Segs sg{seg_lds};
    stage_entries<double>(i, P, maxp, ntypes, nbr_meta, nbr_r, su, orig, sg, lane);
    for (int k = lane; k < ldg; k += kWave) Vrow[k] = 0.0;
    __syncthreads();
    {
        const DevStruc& st = S[atom_struc[i]];
        const double gx = gF[3 * (int64_t)i], gy = gF[3 * (int64_t)i + 1], gz = gF[3 * (int64_t)i + 2];
        const int64_t row = (int64_t)i * maxp;
        for (int p = lane; p < P; p += kWave) {
            const int64_t e = row + orig[p];
            double px = gx, py = gy, pz = gz;
            if ((nbr_meta[e] & 1) || all_images) {
                const int64_t gj = st.atom0 + nbr_j[e];
                px -= gF[3 * gj];
                py -= gF[3 * gj + 1];
                pz -= gF[3 * gj + 2];
            }
            const double d = su[4 * p + 3], di = 1.0 / d;
            tg[4 * p + 0] = px * di;
            tg[4 * p + 1] = py * di;
            tg[4 * p + 2] = pz * di;
            tg[4 * p + 3] = su[4 * p] * px + su[4 * p + 1] * py + su[4 * p + 2] * pz;
        }
    }
    __syncthreads();
    // radial: cutoff value and slope once per entry, then (function, part) items over the lanes
    {
        double* fcr = Rf;
        double* gcr = Rd;
        double* part = red;
        for (int b = 0; b < ntypes; ++b) {
            const DevRadial& R = plan->rad[ti][b];
            for (int p = sg.start[b] + lane; p < sg.start[b + 1]; p += kWave) {
                double fc, gc;
                cutoff_fg<double>(R.cut, su[4 * p + 3], R.rc, R.rci, R.pirci, fc, gc);
                fcr[p] = fc;
                gcr[p] = gc;
            }
        }
        __syncthreads();
        int nrt = 0;
        for (int b = 0; b < ntypes; ++b) nrt += plan->rad[ti][b].n;
        const int nparts = nrt > 0 ? max(1, kWave / nrt) : 1;
        for (int f0 = 0; f0 < nrt; f0 += kWave / nparts) {
            const int f = f0 + lane / nparts, prt = lane % nparts;
            double acc = 0.0, norm = 0.0;
            int goff = -1;
            if (lane < (kWave / nparts) * nparts && f < nrt) {
                int b = 0, fl = f;
                while (fl >= plan->rad[ti][b].n) fl -= plan->rad[ti][b++].n;
                const DevRadial& R = plan->rad[ti][b];
                const double eta = rfun_eta[R.fbase + fl], rs = rfun_rs[R.fbase + fl];
                for (int p = sg.start[b] + prt; p < sg.start[b + 1]; p += nparts) {
                    const double d = su[4 * p + 3];
                    double dv;
                    if (!GEN || R.phi == RAD_GAUSSIAN) {
                        const double x = d - rs;
                        dv = exp(-eta * x * x) * (-2.0 * eta * x * fcr[p] + gcr[p]);
                    } else {
                        double v, dphi;
                        rad_phi<double>(R.phi, -eta * (d - rs), v, dphi);
                        dv = -eta * dphi * fcr[p] + v * gcr[p];
                    }
                    acc += tg[4 * p + 3] * dv;
                }
                goff = R.goff + fl;
                norm = R.norm;
            }
            part[lane] = acc;
            __syncthreads();
            if (goff >= 0 && prt == 0) {
                double s = 0.0;
                for (int q = 0; q < nparts; ++q) s += part[lane + q];
                Vrow[goff] = -s * norm;
            }
            __syncthreads();
        }
    }
    // angular blocks: tangents of the moments when the centre species takes them
    if (use_mom && plan->mom_ok[ti]) {
        // tg: (w, d') -> (u' = w - (d'/d) u, d')
        for (int p = lane; p < P; p += kWave) {
            const double sp = tg[4 * p + 3] / su[4 * p + 3];
#pragma unroll
            for (int c = 0; c < 3; ++c) tg[4 * p + c] -= sp * su[4 * p + c];
        }
        radial_tables<double, NE, GEN>(plan->ang[ti][0], su, 0, P, Rf, Rd, lane);
        double* Mom = momr;
        double* dMom = Mom + ntypes * NE * kMomK;
        double* dD = dMom + ntypes * NE * kMomK;
        double* S = dD + ntypes * NE;
        build_moments_jvp<NE>(su, tg, Rf, Rd, P, (kWave * NE * NC) / kMomPw, ntypes, sg.start, red, Mom, dMom, nullptr,
                              dD, lane);
        symm_angular_moments_jvp<NE>(plan, ti, ntypes, Mom, dMom, dD, S, Vrow, lane);
        __syncthreads();
        for (int k = lane; k < ldg; k += kWave) V[(int64_t)i * ldg + k] = Vrow[k];
        return;
    }
    // angular blocks in LMat order
    const bool shared = plan->ang_shared[ti];
    if (shared) radial_tables<double, NE, GEN>(plan->ang[ti][0], su, 0, P, Rf, Rd, lane);
    for (int bb = 0; bb < ntypes; ++bb) {
        for (int a = 0; a <= bb; ++a) {
            const DevAngular& A = plan->ang[ti][bb * (bb + 1) / 2 + a];
            const int sa = sg.start[a], na = sg.start[a + 1] - sa;
            const int sb = sg.start[bb], nb = sg.start[bb + 1] - sb;
            const bool tri = (a == bb);
            const long long Q = tri ? (long long)na * (na - 1) / 2 : (long long)na * nb;
            if (Q == 0) continue;
            const bool nsb = GEN && !ang_separable(A.phi);
            if (nsb) {
                ns_tables<double, NE>(A, su, sa, sa + na, Rf, Rd, nsx, lane);
                if (!tri) ns_tables<double, NE>(A, su, sb, sb + nb, Rf, Rd, nsx, lane);
            } else if (!shared) {
                radial_tables<double, NE, GEN>(A, su, sa, sa + na, Rf, Rd, lane);
                if (!tri) radial_tables<double, NE, GEN>(A, su, sb, sb + nb, Rf, Rd, lane);
            }
            __syncthreads();
            double lam[NC], zet[NC];
            int zin[NC];
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                lam[c] = A.lambda[c];
                zet[c] = A.zeta[c];
                zin[c] = A.zint[c];
            }
            double acc[NE][NC];
#pragma unroll
            for (int e = 0; e < NE; ++e)
#pragma unroll
                for (int c = 0; c < NC; ++c) acc[e][c] = 0.0;
            const long long L = (Q + kWave - 1) / kWave;
            const long long q0 = (long long)lane * L, q1 = min(q0 + L, Q);
            if (q0 < q1) {
                int j, k;
                if (tri) {
                    j = tri_row(q0, na);
                    k = j + 1 + (int)(q0 - (long long)j * (2LL * na - j - 1) / 2);
                } else {
                    j = (int)(q0 / nb);
                    k = (int)(q0 % nb);
                }
                long long q = q0;
                while (q < q1) {
                    const int kend = tri ? na : nb;
                    const int n = (int)min((long long)(kend - k), q1 - q);
                    const int pj = sa + j;
                    const double ujx = su[4 * pj], ujy = su[4 * pj + 1], ujz = su[4 * pj + 2];
                    const double wjx = tg[4 * pj], wjy = tg[4 * pj + 1], wjz = tg[4 * pj + 2], ddj = tg[4 * pj + 3];
                    const double sj = ddj / su[4 * pj + 3];
                    double rfj[NE], rdj[NE];
#pragma unroll
                    for (int e = 0; e < NE; ++e) {
                        rfj[e] = Rf[pj * NE + e];
                        rdj[e] = Rd[pj * NE + e];
                    }
                    const double fcj = nsb ? nsx[2 * pj] : 0.0, gcj = nsb ? nsx[2 * pj + 1] : 0.0;
                    for (int t = 0; t < n; ++t) {
                        const int pk = sb + k + t;
                        const double ukx = su[4 * pk], uky = su[4 * pk + 1], ukz = su[4 * pk + 2];
                        const double ddk = tg[4 * pk + 3];
                        const double cs = ujx * ukx + ujy * uky + ujz * ukz;
                        const double cdot = ukx * wjx + uky * wjy + ukz * wjz + ujx * tg[4 * pk] + ujy * tg[4 * pk + 1] +
                                            ujz * tg[4 * pk + 2] - cs * (sj + ddk / su[4 * pk + 3]);
                        double pf[NE], pdt[NE];
                        if (nsb) {
                            const double fck = nsx[2 * pk], gck = nsx[2 * pk + 1];
#pragma unroll
                            for (int e = 0; e < NE; ++e) {
                                double F, tt;
                                ang_pair<double>(A.phi, rfj[e], Rf[pk * NE + e], F, tt);
                                pf[e] = fcj * fck * F;
                                pdt[e] = fck * F * (gcj - fcj * tt * rdj[e]) * ddj +
                                         fcj * F * (gck - fck * tt * Rd[pk * NE + e]) * ddk;
                            }
                        } else {
#pragma unroll
                            for (int e = 0; e < NE; ++e) {
                                const double rfk = Rf[pk * NE + e];
                                pf[e] = rfj[e] * rfk;
                                pdt[e] = rdj[e] * rfk * ddj + rfj[e] * Rd[pk * NE + e] * ddk;
                            }
                        }
#pragma unroll
                        for (int c = 0; c < NC; ++c) {
                            double Av, dAv;
                            angular_term<double, GEN>(cs, lam[c], zet[c], zin[c], Av, dAv);
                            const double dc = dAv * cdot;
#pragma unroll
                            for (int e = 0; e < NE; ++e) acc[e][c] += dc * pf[e] + Av * pdt[e];
                        }
                    }
                    q += n;
                    ++j;
                    k = tri ? j + 1 : 0;
                }
            }
#pragma unroll
            for (int e = 0; e < NE; ++e)
#pragma unroll
                for (int c = 0; c < NC; ++c) red[(e * NC + c) * kWave + lane] = acc[e][c];
            __syncthreads();
            for (int t = lane; t < NE * NC; t += kWave) {
                const int e = t / NC, c = t % NC;
                const int f = A.fmap[e * kMaxC + c];
                if (f >= 0) {
                    double s = 0.0;
                    for (int l = 0; l < kWave; ++l) s += red[t * kWave + l];
                    Vrow[A.goff + f] = -s * A.norm;
                }
            }
            __syncthreads();
        }
    }
    for (int k = lane; k < ldg; k += kWave) V[(int64_t)i * ldg + k] = Vrow[k];
}

template <int NE, int NC, bool GEN>
static hipError_t launch_jvp_t(const DevPlan* plan, const double* rfe, const double* rfr, const int* atom_type,
                               const DevStruc* S, const int* atom_struc, int ntot, const int* cnt, const int* nbr_j,
                               const int* meta, const double4* r, int maxp, int lmaxp, const double* gF, int all_images,
                               double* V, int ldg, int use_mom, int ntypes, hipStream_t st) {
    size_t lds = jvp_lds_bytes<NE, NC, GEN>(lmaxp, ldg);
    if (lds > 65536) return hipErrorInvalidConfiguration;
    if (use_mom) {
        const size_t ml = jvp_lds_bytes<NE, NC, GEN>(lmaxp, ldg, jvp_mom_elems<NE>(ntypes));
        if (ml <= 65536) lds = ml;
        else use_mom = 0;
    }
    hipLaunchKernelGGL((k_symm_jvp<NE, NC, GEN>), dim3(ntot), dim3(kWave), lds, st, plan, rfe, rfr, atom_type, S,
                       atom_struc, cnt, nbr_j, meta, r, maxp, lmaxp, gF, all_images, V, ldg, use_mom);
    return hipGetLastError();
}

size_t symm_jvp_lds(int shape, bool ns, int lmaxp, int ldg) {
    if (shape == 0) return ns ? jvp_lds_bytes<2, 4, true>(lmaxp, ldg) : jvp_lds_bytes<2, 4, false>(lmaxp, ldg);
    if (shape == 1) return ns ? jvp_lds_bytes<4, 4, true>(lmaxp, ldg) : jvp_lds_bytes<4, 4, false>(lmaxp, ldg);
    return ns ? jvp_lds_bytes<8, 8, true>(lmaxp, ldg) : jvp_lds_bytes<8, 8, false>(lmaxp, ldg);
}

hipError_t launch_symm_jvp(int shape, int ns, const DevPlan* plan, const double* rfe, const double* rfr,
                           const int* atom_type, const DevStruc* S, const int* atom_struc, int ntot, const int* cnt,
                           const int* nbr_j, const int* meta, const double4* r, int maxp, int lmaxp, const double* gF,
                           int all_images, double* V, int ldg, int use_mom, int ntypes, hipStream_t st) {
#define JVP_ARGS \
    plan, rfe, rfr, atom_type, S, atom_struc, ntot, cnt, nbr_j, meta, r, maxp, lmaxp, gF, all_images, V, ldg, use_mom, ntypes, st
    if (shape == 0) return ns ? launch_jvp_t<2, 4, true>(JVP_ARGS) : launch_jvp_t<2, 4, false>(JVP_ARGS);
    if (shape == 1) return ns ? launch_jvp_t<4, 4, true>(JVP_ARGS) : launch_jvp_t<4, 4, false>(JVP_ARGS);
    return ns ? launch_jvp_t<8, 8, true>(JVP_ARGS) : launch_jvp_t<8, 8, false>(JVP_ARGS);
#undef JVP_ARGS
}

// ------------------------------------------------------------------ fused persistent kernel
// Network of one atom inside the wavefront (the batched form's fused launch): lanes own output
// units; forward layers read the transposed weights (coalesced over lanes), backward layers the
// row-major weights.  Same reference arithmetic as the MFMA kernels (net.hip): fp/fpbp and
// DODZ::grad, ml/nn.cpp:1191-1233, 1620-1628.
template <typename T, bool GEN>
__device__ __forceinline__ double net_body(const DevNet& N, const T* __restrict__ Wt, const T* G, T* act0, T* act1,
                                           T* der, T* dErow, int maxw, int vac, int lane) {
    for (int k = lane; k < N.nin_pad; k += kWave)
        act0[k] = k < N.nin ? Wt[N.inpw_off + k] * (G[k] + Wt[N.inpb_off + k]) : T(0);
    __syncthreads();
    T* cur = act0;
    T* nxt = act1;
    int in_pad = N.nin_pad, doff = 0;
    for (int l = 0; l < N.L; ++l) {
        const int out_pad = N.wpad[l];
        const T* WT = Wt + N.wtoff[l];
        const T* b = Wt + N.boff[l];
        const int act = (l == N.L - 1) ? NEU_LINEAR : N.act;
        for (int o = lane; o < out_pad; o += kWave) {
            T z = b[o];
            for (int k = 0; k < in_pad; ++k) z += WT[(int64_t)k * out_pad + o] * cur[k];
            T a, d;
            if (GEN) {
                act_eval<T>(act, z, a, d);
            } else if (act == NEU_LINEAR) {
                a = z;
                d = T(1);
            } else {
                a = tanh_m1(z);
                d = T(1) - a * a;
            }
            nxt[o] = a;
            der[doff + o] = d;
        }
        __syncthreads();
        T* t = cur;
        cur = nxt;
        nxt = t;
        in_pad = out_pad;
        doff += out_pad;
    }
    const double E = (double)N.outb + (double)cur[0] * N.outw + (vac ? N.evac : 0.0);
    const int last = N.wpad[N.L - 1];
    doff -= last;
    for (int o = lane; o < last; o += kWave) nxt[o] = o == 0 ? T(N.outw) * der[doff] : T(0);
    __syncthreads();
    {
        T* t = cur;
        cur = nxt;
        nxt = t;
    }
    for (int l = N.L - 1; l >= 0; --l) {
        const int K = N.wpad[l], J = l == 0 ? N.nin_pad : N.wpad[l - 1];
        const T* W = Wt + N.woff[l];  // K x J row-major
        if (l > 0) doff -= J;
        for (int j = lane; j < J; j += kWave) {
            T sacc = T(0);
            for (int k = 0; k < K; ++k) sacc += W[(int64_t)k * J + j] * cur[k];
            if (l > 0) nxt[j] = sacc * der[doff + j];
            else dErow[j] = j < N.nin ? sacc * Wt[N.inpw_off + j] : T(0);
        }
        __syncthreads();
        T* t = cur;
        cur = nxt;
        nxt = t;
    }
    return E;
}

// ------------------------------------------------------------------ force terms, NW waves per centre
// Moment-path force terms when every species takes the moments (fast-path models).  NW
// wavefronts share one centre: wave w owns the exponents e = w NE/NW ... (w + 1) NE/NW - 1 and the radial
// functions f = w, w + NW, ... of every entry, and a contiguous 1/NW of the entries in the moment
// sums; partial moments and partial force terms meet in LDS and are added in wave order.  Per
// centre this divides the latency chain by about NW at the same LDS, so twice the wavefronts
// are resident.
// power-table rows per wave and chunk (the L table, ntypes x 35 x NE, sets the region's size)
template <int NW>
__host__ __device__ constexpr int mom_chw() {
    return NW == 1 ? 16 : 8;
}

// partial moments [NW][S][NE][K]; after the adjoint tables the force exchange [NW-1][kWave][3]
// followed by the cutoff values of every entry [lmaxp][4]
template <int NE, int NW>
__host__ __device__ constexpr int force_mom_mp_elems(int ntypes, int lmaxp) {
    return NW * ntypes * NE * kMomK > (NW - 1) * kWave * 3 + 4 * lmaxp ? NW * ntypes * NE * kMomK
                                                                         : (NW - 1) * kWave * 3 + 4 * lmaxp;
}
template <int NE, int NW>
__host__ __device__ constexpr int force_mom_u_elems(int ntypes) {
    return NW * mom_chw<NW>() * kMomPw > ntypes * kMomK * NE ? NW * mom_chw<NW>() * kMomPw : ntypes * kMomK * NE;
}
template <typename T, int NE, int NW>
__host__ __device__ constexpr size_t force_mom_lds_bytes(int lmaxp, int ldg, int ntypes) {
    return sizeof(T) * (size_t)(4 * lmaxp + NE * lmaxp + ((ldg + 1) & ~1) + force_mom_mp_elems<NE, NW>(ntypes, lmaxp) +
                                ntypes * NE + (ntypes * (ntypes + 1) / 2) * NE * (kMomZ + 1) +
                                force_mom_u_elems<NE, NW>(ntypes)) +
           sizeof(int) * (size_t)lmaxp;
}

// dE/dr of entry p (the negated force term) over the radial functions f0, f0 + fstep, ... and
// the NK exponents e0 ... e0 + NK - 1 of the angular term (the caller adds the parts of the
// other functions and exponents)
template <typename T, int NE, int NK>
__device__ __forceinline__ void force_mom_entry(const DevPlan* __restrict__ plan, const double* __restrict__ rfun_eta,
                                                const double* __restrict__ rfun_rs, int ti, int ntypes, const Segs& sg,
                                                const T* su, const T* Cut, const T* W, const T* U, const T* Delta,
                                                int p, int f0, int fstep, int e0, T& gx, T& gy, T& gz) {
    const DevAngular& A0 = plan->ang[ti][0];
    int s = 0;
    while (s + 1 < ntypes && p >= sg.start[s + 1]) ++s;
    const T u[3] = {su[4 * p], su[4 * p + 1], su[4 * p + 2]};
    const T d = su[4 * p + 3], di = T(1) / d;
    // radial term dE/dd u
    const DevRadial& R = plan->rad[ti][s];
    const T fc = Cut[4 * p], gc = Cut[4 * p + 1];
    T sr = T(0);
    for (int f = f0; f < R.n; f += fstep) {
        const T eta = T(rfun_eta[R.fbase + f]), rs = T(rfun_rs[R.fbase + f]);
        const T x = d - rs;
        sr += W[R.goff + f] * kexp(-eta * x * x) * (T(-2) * eta * x * fc + gc);
    }
    sr *= T(R.norm);
    gx = sr * u[0];
    gy = sr * u[1];
    gz = sr * u[2];
    // angular term of the exponents e0 + k
    const T fa = Cut[4 * p + 2], ga = Cut[4 * p + 3];
    const T dp = ipow(d, A0.alpha);
    const T ddp = T(A0.alpha) * dp * di;
    T m[kMomK3];
    mom_monos<T>(m, u, std::make_integer_sequence<int, kMomK3>{});
    T g0[NK], h[3][NK];
#pragma unroll
    for (int k = 0; k < NK; ++k) g0[k] = h[0][k] = h[1][k] = h[2][k] = T(0);
    mom_terms<T, NK, NE, 1>(U + s * kMomK * NE + e0, m, u, g0, h, std::make_integer_sequence<int, kMomK>{});
#pragma unroll
    for (int k = 0; k < NK; ++k) {
        const int e = e0 + k;
        const T ie = T(A0.ietap[e]);
        const T ex = kexp(-ie * dp);
        const T r = fa * ex, rd = (ga - ie * ddp * fa) * ex;
        const T gu = u[0] * h[0][k] + u[1] * h[1][k] + u[2] * h[2][k];
        const T dEdR = g0[k] - T(2) * Delta[s * NE + e] * r;
        const T a1 = dEdR * rd - r * gu * di;
        const T a2 = r * di;
        gx += a1 * u[0] + a2 * h[0][k];
        gy += a1 * u[1] + a2 * h[1][k];
        gz += a1 * u[2] + a2 * h[2][k];
    }
}

template <typename T, int NE, int NW>
__global__ __launch_bounds__(kWave * NW) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 8 ? 4 : 6, 8))) void k_force_mom(
    const DevPlan* __restrict__ plan, const double* __restrict__ rfun_eta, const double* __restrict__ rfun_rs,
    const int* __restrict__ atom_type, const int* __restrict__ centres, int ncentres, const int* __restrict__ cnt,
    const int* __restrict__ nbr_meta, const double4* __restrict__ nbr_r, int maxp, int lmaxp,
    const T* __restrict__ dEdG, int ldg, T* __restrict__ Xout, const T* __restrict__ Min) {
    constexpr int NK = NE / NW;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const int ntypes = plan->T;
    const int nblk = ntypes * (ntypes + 1) / 2;
    const int tid = threadIdx.x, lane = tid & (kWave - 1), w = tid / kWave;
    T* su = (T*)lds_raw;
    T* Rf = su + 4 * lmaxp;
    T* W = Rf + NE * lmaxp;
    T* Mp = W + ((ldg + 1) & ~1);  // partial moments; then the force exchange and the cutoff values
    T* Delta = Mp + force_mom_mp_elems<NE, NW>(ntypes, lmaxp);
    T* wtab = Delta + ntypes * NE;
    T* U = wtab + nblk * NE * (kMomZ + 1);  // power rows [NW][kMomChW][15]; then L [S][K][NE]
    int* orig = (int*)(U + force_mom_u_elems<NE, NW>(ntypes));
    __shared__ int seg_lds[kMaxSpecies + 1];
    Segs sg{seg_lds};
    const int slot = xcd_slot(ncentres);  // the XCD-local order of the neighbour kernel
    if (slot < 0) return;
    const int i = centres[slot];
    const int ti = atom_type[i];
    const int P = min(cnt[i], lmaxp);
    if (w == 0) stage_entries<T>(i, P, maxp, ntypes, nbr_meta, nbr_r, su, orig, sg, lane);
    if (NW == 1 || w > 0) {
        const int t0 = NW == 1 ? tid : tid - kWave, nt = NW == 1 ? kWave : kWave * (NW - 1);
        for (int k = t0; k < ldg; k += nt) W[k] = dEdG[(int64_t)i * ldg + k];
    }
    // the descriptor kernel of the same pass left this centre's moments: they are read with the
    // entries, and the exponent factors and moment sums below are skipped
    if (Min) {
        const T* src = Min + (int64_t)i * (ntypes * NE * kMomK);
        for (int it = tid; it < ntypes * NE * kMomK; it += kWave * NW) Mp[it] = src[it];
    }
    __syncthreads();

    const DevAngular& A0 = plan->ang[ti][0];
    if (!Min) {
    // angular radial factors R_e(p) of this wave's exponents (one table for every block)
    for (int p = lane; p < P; p += kWave) {
        const T d = su[4 * p + 3];
        T fc, gc;
        cutoff_fg<T>(A0.cut, d, T(A0.rc), T(A0.rci), T(A0.pirci), fc, gc);
        const T dp = ipow(d, A0.alpha);
#pragma unroll
        for (int k = 0; k < NK; ++k) {
            const int e = w * NK + k;
            Rf[p * NE + e] = fc * kexp(-T(A0.ietap[e]) * dp);
        }
    }
    __syncthreads();

    // partial moments of this wave's entries [lo, hi) (32 monomials summed over the two half-waves,
    // the last three derived from |u| = 1)
    {
        T* Mw = Mp + w * ntypes * NE * kMomK;
        constexpr int kMomChW = mom_chw<NW>();
        T* pw = U + w * kMomChW * kMomPw;
        for (int it = lane; it < ntypes * NE * kMomK; it += kWave) Mw[it] = T(0);
        const int per = (P + NW - 1) / NW;
        const int lo = min(P, w * per), hi = min(P, lo + per);
        const int nch = (per + kMomChW - 1) / kMomChW;  // the same for every wave
        for (int c = 0; c < nch; ++c) {
            const int c0 = min(hi, lo + c * kMomChW), c1 = min(hi, c0 + kMomChW);
            __syncthreads();  // the previous chunk's readers are done (and the moments cleared)
            power_rows<T>(su, c0, c1, pw, lane);
            __syncthreads();
            moments_chunk<T, NE>(Rf, c0, c1, ntypes, sg.start, pw, Mw, lane);
        }
        __syncthreads();
        for (int it = lane; it < ntypes * NE; it += kWave) mom_derive<T>(Mw + it * kMomK);
        __syncthreads();
    }
    }
    adjoint_moments<T, NE, true>(plan, ti, ntypes, W, Mp, Min ? 1 : NW, ntypes * NE * kMomK, Delta, wtab, U, tid,
                                 kWave * NW);

    // cutoff value and derivative of every entry, radial then angular, in the partial moments'
    // place.  Computed ahead of the entry loop: inside it, the polynomial constants of the cosine
    // are loop invariants that the compiler keeps in registers across the loop, beyond its budget.
    T* Xx = Mp;
    T* Cut = Mp + (NW - 1) * kWave * 3;
    for (int p = tid; p < P; p += kWave * NW) {
        int s = 0;
        while (s + 1 < ntypes && p >= sg.start[s + 1]) ++s;
        const DevRadial& R = plan->rad[ti][s];
        const T d = su[4 * p + 3];
        T fc, gc;
        cutoff_fg<T>(R.cut, d, T(R.rc), T(R.rci), T(R.pirci), fc, gc);
        T fa = fc, ga = gc;
        if (A0.cut != R.cut || A0.rc != R.rc) cutoff_fg<T>(A0.cut, d, T(A0.rc), T(A0.rci), T(A0.pirci), fa, ga);
        Cut[4 * p] = fc;
        Cut[4 * p + 1] = gc;
        Cut[4 * p + 2] = fa;
        Cut[4 * p + 3] = ga;
    }
    __syncthreads();

    // per entry: this wave's radial functions and exponents; X = -(sum of the waves' dE/dr)
    const int64_t row = (int64_t)i * maxp;
    for (int j0 = 0; j0 < P; j0 += kWave) {
        if constexpr (NW == 1 && NK % 2 == 0 && sizeof(T) == 4 && NE <= 4) {
            if (P - j0 <= kWave / 2) {
                // at most 32 entries left: two lanes per entry, the even radial functions and the
                // first half of the exponents in lanes 0-31, the rest in lanes 32-63, added by a
                // shuffle across the halves (the last pass of a centre with 64 < P <= 96 costs
                // half a pass instead of a whole one).  FP32 only: in FP64 the second form of the
                // body pushes the kernel past its register budget.
                const int hf = lane >> 5, p = j0 + (lane & 31);
                T gx = T(0), gy = T(0), gz = T(0);
                if (p < P)
                    force_mom_entry<T, NE, NK / 2>(plan, rfun_eta, rfun_rs, ti, ntypes, sg, su, Cut, W, U, Delta, p, hf,
                                                   2, hf * (NK / 2), gx, gy, gz);
                gx += __shfl_xor(gx, kWave / 2);
                gy += __shfl_xor(gy, kWave / 2);
                gz += __shfl_xor(gz, kWave / 2);
                if (hf == 0 && p < P) store_x4<T>(Xout + 4 * (row + orig[p]), -gx, -gy, -gz);
                continue;
            }
        }
        const int p = j0 + lane;
        T gx = T(0), gy = T(0), gz = T(0);
        if (p < P)
            force_mom_entry<T, NE, NK>(plan, rfun_eta, rfun_rs, ti, ntypes, sg, su, Cut, W, U, Delta, p, w, NW, w * NK,
                                       gx, gy, gz);
        if (w > 0) {
            T* x = Xx + ((w - 1) * kWave + lane) * 3;
            x[0] = gx;
            x[1] = gy;
            x[2] = gz;
        }
        if (NW > 1) __syncthreads();
        if (w == 0 && p < P) {
            for (int v = 1; v < NW; ++v) {
                const T* x = Xx + ((v - 1) * kWave + lane) * 3;
                gx += x[0];
                gy += x[1];
                gz += x[2];
            }
            store_x4<T>(Xout + 4 * (row + orig[p]), -gx, -gy, -gz);
        }
        if (NW > 1) __syncthreads();
    }
}

template <typename T, int NE, int NW>
static hipError_t launch_force_mom_t(const DevPlan* plan, const double* rfe, const double* rfr, const int* atom_type,
                                     const int* centres, int ncentres, const int* cnt, const int* meta,
                                     const double4* r, int maxp, int lmaxp, const T* dEdG, int ldg, T* X, int ntypes,
                                     hipStream_t st, const T* Min) {
    const size_t lds = force_mom_lds_bytes<T, NE, NW>(lmaxp, ldg, ntypes);
    if (lds > 65536) return hipErrorInvalidConfiguration;
    hipLaunchKernelGGL((k_force_mom<T, NE, NW>), dim3(xcd_grid(ncentres)), dim3(kWave * NW), lds, st, plan, rfe,
                       rfr, atom_type, centres, ncentres, cnt, meta, r, maxp, lmaxp, dEdG, ldg, X, Min);
    return hipGetLastError();
}

template <typename T, int NE, int NC, bool GEN>
__host__ __device__ constexpr size_t fused_lds_bytes(int maxp, int ldg, int maxw, int dersum) {
    return sizeof(T) * (size_t)(4 * maxp + 2 * NE * maxp + 3 * maxp + 6 * kWave + kWave * NE * NC +
                                2 * ((ldg + 1) & ~1) + 2 * maxw + dersum) +
           sizeof(int) * (size_t)(maxp + ((2 * kWave + 3) & ~3)) + (GEN ? sizeof(T) * 2 * (size_t)maxp : 0);
}

// Persistent fused pass: each wavefront takes chunks of atoms from a work counter until the
// batch is done (every wave reaches the exit), and for each atom stages its entries once, sums the
// descriptors, runs its network in-wave and computes its per-entry force terms.  The order of
// atoms (species-sorted) keeps a species' weights hot in the caches.
template <typename T, int NE, int NC, bool GEN>
__global__ __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(NE * NC <= 16 ? (sizeof(T) == 8 ? (GEN ? 1 : 2) : 3) : 1, 8))) void k_fused(
    const DevPlan* __restrict__ plan, const double* __restrict__ rfun_eta, const double* __restrict__ rfun_rs,
    const T* __restrict__ Wt, const int* __restrict__ atom_type, const int* __restrict__ order, int ntot,
    int* __restrict__ work, const int* __restrict__ cnt, const int* __restrict__ nbr_meta,
    const double4* __restrict__ nbr_r, int maxp, int lmaxp, int ldg, int maxw, int dersum, int vac, int do_force,
    double* __restrict__ eatom, T* __restrict__ Xout, int use_mom, int chunk) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    T* su = (T*)lds_raw;
    T* Rf = su + 4 * lmaxp;
    T* Rd = Rf + NE * lmaxp;
    T* Xs = Rd + NE * lmaxp;
    T* slots = Xs + 3 * lmaxp;
    T* red = slots + 6 * kWave;
    T* Grow = red + kWave * NE * NC;
    T* dErow = Grow + ((ldg + 1) & ~1);
    T* act0 = dErow + ((ldg + 1) & ~1);
    T* act1 = act0 + maxw;
    T* der = act1 + maxw;
    int* orig = (int*)(der + dersum);
    int* slot_row = orig + lmaxp;
    T* nsx = (T*)(slot_row + ((2 * kWave + 3) & ~3));
    T* mom = use_mom ? nsx + (GEN ? 2 * lmaxp : 0) : nullptr;
    __shared__ int seg_lds[kMaxSpecies + 1];
    __shared__ int next_atom;
    Segs sg{seg_lds};
    const int lane = threadIdx.x;
    const int ntypes = plan->T;
    // each block starts on its own chunk of `chunk` atoms, then takes chunks from the counter
    int base = blockIdx.x * chunk, k = 0;
    for (;;) {
        if (k == chunk) {
            if (lane == 0) next_atom = (int)gridDim.x * chunk + atomicAdd(work, chunk);
            __syncthreads();
            base = next_atom;
            __syncthreads();
            k = 0;
        }
        const int idx = base + k++;
        if (idx >= ntot) break;
        const int i = order[idx];
        const int ti = atom_type[i];
        const int P = min(cnt[i], lmaxp);
        stage_entries<T>(i, P, maxp, ntypes, nbr_meta, nbr_r, su, orig, sg, lane);
        for (int k = lane; k < ldg; k += kWave) Grow[k] = T(0);
        __syncthreads();
        symm_body<T, NE, NC, GEN>(plan, rfun_eta, rfun_rs, ti, ntypes, P, lmaxp, su, Rf, red, kWave * NE * NC, Grow,
                                  sg, nsx,
                                  (use_mom & 1) ? mom : nullptr, lane);
        __syncthreads();
        const double E = net_body<T, GEN>(plan->net[ti], Wt, Grow, act0, act1, der, dErow, maxw, vac, lane);
        if (lane == 0) eatom[i] = E;
        if (do_force) {
            force_body<T, NE, NC, GEN>(plan, rfun_eta, rfun_rs, ti, ntypes, P, lmaxp, su, Rf, Rd, Xs, slots,
                                       slot_row, dErow, dErow, sg, nsx, (use_mom & 2) ? mom : nullptr, lane);
            const int64_t row = (int64_t)i * maxp;
            for (int p = lane; p < P; p += kWave) {
                const int o = orig[p];
                store_x4<T>(Xout + 4 * (row + o), Xs[3 * p], Xs[3 * p + 1], Xs[3 * p + 2]);
            }
        }
        __syncthreads();
    }
}

template <typename T, int NE, int NC, bool GEN>
static hipError_t launch_fused_t(int nblocks, const DevPlan* plan, const double* rfe, const double* rfr, const T* Wt,
                                 const int* atom_type, const int* order, int ntot, int* work, const int* cnt,
                                 const int* meta, const double4* r, int maxp, int lmaxp, int ldg, int maxw, int dersum, int vac,
                                 int do_force, double* eatom, T* X, int use_mom, int ntypes, hipStream_t st) {
    const size_t lds = fused_lds_bytes<T, NE, NC, GEN>(lmaxp, ldg, maxw, dersum) +
                       (use_mom ? mom_lds_bytes<T, NE>(lmaxp, ntypes, use_mom) : 0);
    if (lds > 65536) return hipErrorInvalidConfiguration;
    hipLaunchKernelGGL((k_fused<T, NE, NC, GEN>), dim3(nblocks), dim3(kWave), lds, st, plan, rfe, rfr, Wt, atom_type,
                       order, ntot, work, cnt, meta, r, maxp, lmaxp, ldg, maxw, dersum, vac, do_force, eatom, X, use_mom,
                       std::max(1, std::min(8, ntot / (8 * std::max(nblocks, 1)))));
    return hipGetLastError();
}

size_t mom_lds(int shape, bool fp64, int lmaxp, int ntypes, int bits) {
    if (shape == 0) return fp64 ? mom_lds_bytes<double, 2>(lmaxp, ntypes, bits) : mom_lds_bytes<float, 2>(lmaxp, ntypes, bits);
    if (shape == 1) return fp64 ? mom_lds_bytes<double, 4>(lmaxp, ntypes, bits) : mom_lds_bytes<float, 4>(lmaxp, ntypes, bits);
    return fp64 ? mom_lds_bytes<double, 8>(lmaxp, ntypes, bits) : mom_lds_bytes<float, 8>(lmaxp, ntypes, bits);
}

template <typename T, bool GEN>
static size_t fused_lds_t(int shape, int maxp, int ldg, int maxw, int dersum) {
    if (shape == 0) return fused_lds_bytes<T, 2, 4, GEN>(maxp, ldg, maxw, dersum);
    if (shape == 1) return fused_lds_bytes<T, 4, 4, GEN>(maxp, ldg, maxw, dersum);
    return fused_lds_bytes<T, 8, 8, GEN>(maxp, ldg, maxw, dersum);
}

size_t fused_lds(int shape, bool fp64, bool ns, int maxp, int ldg, int maxw, int dersum) {
    if (fp64) return ns ? fused_lds_t<double, true>(shape, maxp, ldg, maxw, dersum)
                        : fused_lds_t<double, false>(shape, maxp, ldg, maxw, dersum);
    return ns ? fused_lds_t<float, true>(shape, maxp, ldg, maxw, dersum)
              : fused_lds_t<float, false>(shape, maxp, ldg, maxw, dersum);
}

template <typename T, bool GEN>
static hipError_t launch_fused_s(int shape, int nblocks, const DevPlan* plan, const double* rfe, const double* rfr,
                                 const T* Wt, const int* atom_type, const int* order, int ntot, int* work,
                                 const int* cnt, const int* meta, const double4* r, int maxp, int lmaxp, int ldg,
                                 int maxw, int dersum, int vac, int do_force, double* eatom, T* X, int use_mom,
                                 int ntypes, hipStream_t st) {
    if (shape == 0)
        return launch_fused_t<T, 2, 4, GEN>(nblocks, plan, rfe, rfr, Wt, atom_type, order, ntot, work, cnt, meta, r,
                                           maxp, lmaxp, ldg, maxw, dersum, vac, do_force, eatom, X, use_mom, ntypes, st);
    if (shape == 1)
        return launch_fused_t<T, 4, 4, GEN>(nblocks, plan, rfe, rfr, Wt, atom_type, order, ntot, work, cnt, meta, r,
                                           maxp, lmaxp, ldg, maxw, dersum, vac, do_force, eatom, X, use_mom, ntypes, st);
    return launch_fused_t<T, 8, 8, GEN>(nblocks, plan, rfe, rfr, Wt, atom_type, order, ntot, work, cnt, meta, r, maxp,
                                       lmaxp, ldg, maxw, dersum, vac, do_force, eatom, X, use_mom, ntypes, st);
}

template <typename T>
hipError_t launch_fused(int shape, int ns, int nblocks, const DevPlan* plan, const double* rfe, const double* rfr,
                        const T* Wt, const int* atom_type, const int* order, int ntot, int* work, const int* cnt,
                        const int* meta, const double4* r, int maxp, int lmaxp, int ldg, int maxw, int dersum, int vac,
                        int do_force, double* eatom, T* X, int use_mom, int ntypes, hipStream_t st) {
    if (ns)
        return launch_fused_s<T, true>(shape, nblocks, plan, rfe, rfr, Wt, atom_type, order, ntot, work, cnt, meta, r,
                                       maxp, lmaxp, ldg, maxw, dersum, vac, do_force, eatom, X, use_mom, ntypes, st);
    return launch_fused_s<T, false>(shape, nblocks, plan, rfe, rfr, Wt, atom_type, order, ntot, work, cnt, meta, r,
                                    maxp, lmaxp, ldg, maxw, dersum, vac, do_force, eatom, X, use_mom, ntypes, st);
}

template hipError_t launch_fused<double>(int, int, int, const DevPlan*, const double*, const double*, const double*,
                                         const int*, const int*, int, int*, const int*, const int*, const double4*, int,
                                         int, int, int, int, int, int, double*, double*, int, int, hipStream_t);
template hipError_t launch_fused<float>(int, int, int, const DevPlan*, const double*, const double*, const float*,
                                        const int*, const int*, int, int*, const int*, const int*, const double4*, int,
                                        int, int, int, int, int, int, double*, float*, int, int, hipStream_t);

// F_a = sum_s X[a,s] - sum_{entries (i -> a) that carry a reaction} X[i -> a], in FP64.
//
// NNP::force (nnp.cpp:441-446, 469-478) subtracts the term of entry (i -> a, image n) from F_a
// when that entry, in centre i's own row, carries the min flag (every entry with all_images).
// One wavefront per atom a, lanes over its own entries.  The reactions are gathered from the
// rows of a's neighbours: for the first entry of each neighbour j in a's row (rows are sorted by
// neighbour index, then image), the run of entries with index a in j's row is located (lower
// bound of a) and every entry of that run that carries a reaction is subtracted.  No image
// correspondence is assumed, so images that tie in distance (half-lattice offsets, where both
// differences wrap to -1/2) take exactly the entries the reference takes.
//
// That gathers everything unless some centre i lists a while a's own row does not list i (a
// distance within rounding of rc, or an image table that is not symmetric after the 6-shift
// fallback of Cell::diff).  Each atom therefore also counts its own reaction-carrying entries and
// the entries it consumed; summed over atoms the difference is exactly the number of reactions
// missed.  Atoms whose two counts differ add the difference to react[0] (zero atomics in the
// usual case); the host reads it with the pass's results and, when it is not zero, redoes the
// assembly from receiver rows (MODE 1 below), which does not depend on that symmetry.
//
// MODE 1: the receiver rows built by k_react_fill (the entries carrying a reaction onto a, in
// an order set by atomics) are ranked by entry index in LDS and subtracted in that order, so the
// result does not depend on the order of the atomics.  A row longer than rcap is reported in
// react[1] (largest length seen) and the host redoes the pass with longer rows.
//
// Lane partials are summed in a fixed tree order (deterministic, no atomics on F).
//
// Block order: workgroups are dealt to the 8 XCDs round-robin, each XCD with its own L2.  Block b
// (after the tail blocks, padded to a multiple of 8) takes position (b mod 8) * chunk + b / 8 of the
// processing order (cell-list atoms by bin, then the rest), so the blocks an XCD runs at any time
// are neighbours in space and the rows and terms they look up in each other's rows are shared in
// that XCD's L2, instead of every XCD touching the rows of the whole batch.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

__device__ __forceinline__ int wave_sum_i(int v) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

// First index in [0, P) of row `rowj` whose neighbour index is >= want (P if none).  The row's
// indices ascend, so the entry sits near want / natoms of the way along it: the 16 indices around
// that guess come in four aligned 16-byte loads and, when they bracket want, their count places
// it; otherwise an 8-ary search (7 independent probes per level, a short dependent chain) takes
// the side that holds it and the last <= 16 indices are counted.
__device__ __forceinline__ int row_lower_bound(const int* __restrict__ nbr_j, int64_t rowj, int P, int want,
                                               int natoms) {
    int lo = 0, hi = P;
    if (P >= 16) {
        const int g = (int)(((int64_t)want * P) / max(natoms, 1));
        const int w0 = min(max(0, g - 8), P - 16) & ~3;  // 4-aligned, and w0 + 16 <= P
        const int4* kv = reinterpret_cast<const int4*>(nbr_j + rowj + w0);
        int c = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int4 k4 = kv[q];
            c += (k4.x < want) + (k4.y < want) + (k4.z < want) + (k4.w < want);
        }
        const int w1 = w0 + 16;
        if (c > 0 && (c < 16 || w1 == P)) {
            lo = hi = w0 + c;
        } else if (c == 0) {
            hi = w0;
        } else {
            lo = w1;
        }
    }
    while (hi - lo > 16) {
        const int len = hi - lo;
        int c = 0;
#pragma unroll
        for (int m = 1; m < 8; ++m) c += nbr_j[rowj + lo + (len * m) / 8] < want ? 1 : 0;
        const int nlo = c == 0 ? lo : lo + (len * c) / 8 + 1;
        const int nhi = c == 7 ? hi : lo + (len * (c + 1)) / 8;
        lo = nlo;
        hi = nhi;
    }
    int c = 0;
#pragma unroll
    for (int m = 0; m < 16; ++m) c += (lo + m < hi && nbr_j[rowj + lo + m] < want) ? 1 : 0;
    return lo + c;
}


template <typename T, int MODE>
__global__ __launch_bounds__(kWave) void k_assemble(const DevStruc* __restrict__ S, const int* __restrict__ atom_struc,
                                                    const int* __restrict__ cnt, const int* __restrict__ nbr_j,
                                                    const int* __restrict__ nbr_meta, const T* __restrict__ X, int maxp,
                                                    int all_images, double* __restrict__ F, EnergyTail tail,
                                                    int* __restrict__ react, const int* __restrict__ rcount,
                                                    const int* __restrict__ rlist, int rcap,
                                                    const int* __restrict__ bin_atoms,
                                                    const int* __restrict__ pair_centres, int ncell, int ntot,
                                                    int chunk) {
    // the end-of-pass reduction blocks come first, so they run beside the atoms' blocks
    const int ntail = tail.nstruc + tail.ncount;
    const int ntail_pad = (ntail + kXcds - 1) / kXcds * kXcds;
    const int lane = threadIdx.x;
    if ((int)blockIdx.x < ntail) {
        energy_tail_block(tail, blockIdx.x, lane);
        return;
    }
    if ((int)blockIdx.x < ntail_pad) return;
    const int b = blockIdx.x - ntail_pad;
    const int pos = (b % kXcds) * chunk + b / kXcds;
    if (pos >= ntot) return;
    const int a = pos < ncell ? bin_atoms[pos] : pair_centres[pos - ncell];
    const int P = min(cnt[a], maxp);
    const int64_t row = (int64_t)a * maxp;
    double fx = 0, fy = 0, fz = 0;
    if (MODE == 0) {
        const DevStruc& st = S[atom_struc[a]];
        const int la = a - st.atom0;
        int own = 0, used = 0;
        for (int s = lane; s < P; s += kWave) {
            const int64_t e = row + s;
            {
                const double4 v = load_x4<T>(X + 4 * e);
                fx += v.x;
                fy += v.y;
                fz += v.z;
            }
            const int j = nbr_j[e];
            own += (all_images || (nbr_meta[e] & 1)) ? 1 : 0;
            // one lane per neighbour: the first of its entries.  Self entries carry no min flag
            // (neighbor.cpp:131), so under the min-image rule the atom's own row is skipped.
            const bool first = s == 0 || nbr_j[e - 1] != j;
            if (first && (all_images || j != la)) {
                const int gj = st.atom0 + j;
                const int64_t rowj = (int64_t)gj * maxp;
                const int Pj = min(cnt[gj], maxp);
                for (int q = row_lower_bound(nbr_j, rowj, Pj, la, st.natoms); q < Pj && nbr_j[rowj + q] == la; ++q) {
                    if (all_images || (nbr_meta[rowj + q] & 1)) {
                        const double4 v = load_x4<T>(X + 4 * (rowj + q));
                        fx -= v.x;
                        fy -= v.y;
                        fz -= v.z;
                        ++used;
                    }
                }
            }
        }
        const int miss = wave_sum_i(own - used);
        if (lane == 0 && miss != 0) atomicAdd(react, miss);
    } else {
        extern __shared__ int rl_lds[];
        int* ids = rl_lds;           // [rcap] receiver row as filled
        int* sorted = rl_lds + rcap;  // [rcap] ranked by entry index
        for (int s = lane; s < P; s += kWave) {
            const double4 v = load_x4<T>(X + 4 * (row + s));
            fx += v.x;
            fy += v.y;
            fz += v.z;
        }
        int n = rcount[a];
        if (n > rcap) {
            if (lane == 0) atomicMax(react + 1, n);
            n = rcap;
        }
        for (int q = lane; q < n; q += kWave) ids[q] = rlist[(int64_t)a * rcap + q];
        __syncthreads();
        for (int q = lane; q < n; q += kWave) {
            const int id = ids[q];
            int r = 0;
            for (int k = 0; k < n; ++k) r += ids[k] < id ? 1 : 0;  // entry indices are distinct
            sorted[r] = id;
        }
        __syncthreads();
        for (int q = lane; q < n; q += kWave) {
            const double4 v = load_x4<T>(X + 4 * (int64_t)sorted[q]);
            fx -= v.x;
            fy -= v.y;
            fz -= v.z;
        }
    }
    fx = wave_sum(fx);
    fy = wave_sum(fy);
    fz = wave_sum(fz);
    if (lane == 0) {
        F[3 * (int64_t)a] = fx;
        F[3 * (int64_t)a + 1] = fy;
        F[3 * (int64_t)a + 2] = fz;
    }
}

// Receiver rows for MODE 1 of k_assemble: rcount cleared, react[1] (longest row seen) cleared.
__global__ void k_react_zero(int* __restrict__ rcount, int n, int* __restrict__ react) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) rcount[t] = 0;
    if (t == 0) react[1] = 0;
}

// One wavefront per centre i: every entry (i -> j) that carries a reaction is appended to the
// receiver row of j (slot order set by the atomics; k_assemble ranks the row before summing).
__global__ __launch_bounds__(kWave) void k_react_fill(const DevStruc* __restrict__ S, const int* __restrict__ atom_struc,
                                                      const int* __restrict__ cnt, const int* __restrict__ nbr_j,
                                                      const int* __restrict__ nbr_meta, int maxp, int all_images,
                                                      int* __restrict__ rcount, int* __restrict__ rlist, int rcap) {
    const int i = blockIdx.x;
    const int atom0 = S[atom_struc[i]].atom0;
    const int P = min(cnt[i], maxp);
    const int64_t row = (int64_t)i * maxp;
    for (int s = threadIdx.x; s < P; s += kWave) {
        const int64_t e = row + s;
        if (!(all_images || (nbr_meta[e] & 1))) continue;
        const int recv = atom0 + nbr_j[e];
        const int slot = atomicAdd(&rcount[recv], 1);
        if (slot < rcap) rlist[(int64_t)recv * rcap + slot] = (int)e;
    }
}

// ------------------------------------------------------------------ launchers
template <typename T, int NE, int NC, bool GEN>
static hipError_t launch_symm_t(const DevPlan* plan, const double* rfe, const double* rfr, const int* atom_type,
                                const int* centres, int ncentres, const int* cnt, const int* meta, const double4* r,
                                int maxp, int lmaxp, T* G, int ldg, int use_mom, int ntypes, hipStream_t st, T* Mout) {
    // use_mom 2: every species takes the moments, the reduction buffer only holds power-table
    // chunks of 16 entries (and the radial partial sums, kWave values)
    const int red = use_mom == 2 ? 16 * kMomPw : kWave * NE * NC;
    const size_t lds =
        symm_lds_bytes<T, NE, NC, GEN>(lmaxp, ldg, red) + (use_mom ? mom_lds_bytes<T, NE>(lmaxp, ntypes, 1) : 0);
    if (lds > 65536) return hipErrorInvalidConfiguration;
    hipLaunchKernelGGL((k_symm<T, NE, NC, GEN>), dim3(xcd_grid(ncentres)), dim3(kWave), lds, st, plan, rfe, rfr,
                       atom_type, centres, ncentres, cnt, meta, r, maxp, lmaxp, G, ldg, use_mom, red,
                       use_mom == 2 ? Mout : nullptr);
    return hipGetLastError();
}

template <typename T, int NE, int NC, bool GEN>
static hipError_t launch_force_t(const DevPlan* plan, const double* rfe, const double* rfr, const int* atom_type,
                                 const int* centres, int ncentres, const int* cnt, const int* meta, const double4* r,
                                 int maxp, int lmaxp, const T* dEdG, int ldg, T* X, int use_mom, int ntypes,
                                 hipStream_t st) {
    const size_t lds = force_lds_bytes<T, NE, NC, GEN>(lmaxp, ldg) + (use_mom ? mom_lds_bytes<T, NE>(lmaxp, ntypes, 2) : 0);
    if (lds > 65536) return hipErrorInvalidConfiguration;
    hipLaunchKernelGGL((k_force<T, NE, NC, GEN>), dim3(xcd_grid(ncentres)), dim3(kWave), lds, st, plan, rfe, rfr,
                       atom_type, centres, ncentres, cnt, meta, r, maxp, lmaxp, dEdG, ldg, X, use_mom);
    return hipGetLastError();
}

// Grid shapes instantiated: (NE, NC) in {(2,4), (4,4), (8,8)}.
int grid_shape(int ne, int nc) {
    if (ne <= 2 && nc <= 4) return 0;
    if (ne <= 4 && nc <= 4) return 1;
    return 2;
}
int shape_ne(int shape) {
    return shape == 0 ? 2 : (shape == 1 ? 4 : 8);
}

template <typename T, bool GEN>
static size_t symm_lds_t(int shape, int maxp, int ldg) {
    if (shape == 0) return symm_lds_bytes<T, 2, 4, GEN>(maxp, ldg);
    if (shape == 1) return symm_lds_bytes<T, 4, 4, GEN>(maxp, ldg);
    return symm_lds_bytes<T, 8, 8, GEN>(maxp, ldg);
}
template <typename T, bool GEN>
static size_t force_lds_t(int shape, int maxp, int ldg) {
    if (shape == 0) return force_lds_bytes<T, 2, 4, GEN>(maxp, ldg);
    if (shape == 1) return force_lds_bytes<T, 4, 4, GEN>(maxp, ldg);
    return force_lds_bytes<T, 8, 8, GEN>(maxp, ldg);
}

size_t symm_lds(int shape, bool fp64, bool ns, int maxp, int ldg) {
    if (fp64) return ns ? symm_lds_t<double, true>(shape, maxp, ldg) : symm_lds_t<double, false>(shape, maxp, ldg);
    return ns ? symm_lds_t<float, true>(shape, maxp, ldg) : symm_lds_t<float, false>(shape, maxp, ldg);
}

size_t force_lds(int shape, bool fp64, bool ns, int maxp, int ldg) {
    if (fp64) return ns ? force_lds_t<double, true>(shape, maxp, ldg) : force_lds_t<double, false>(shape, maxp, ldg);
    return ns ? force_lds_t<float, true>(shape, maxp, ldg) : force_lds_t<float, false>(shape, maxp, ldg);
}

template <typename T, bool GEN>
static hipError_t launch_symm_s(int shape, const DevPlan* plan, const double* rfe, const double* rfr,
                                const int* atom_type, const int* centres, int ncentres, const int* cnt, const int* meta,
                                const double4* r, int maxp, int lmaxp, T* G, int ldg, int use_mom, int ntypes,
                                hipStream_t st, T* Mout) {
    if (shape == 0)
        return launch_symm_t<T, 2, 4, GEN>(plan, rfe, rfr, atom_type, centres, ncentres, cnt, meta, r, maxp, lmaxp, G, ldg,
                                          use_mom, ntypes, st, Mout);
    if (shape == 1)
        return launch_symm_t<T, 4, 4, GEN>(plan, rfe, rfr, atom_type, centres, ncentres, cnt, meta, r, maxp, lmaxp, G, ldg,
                                          use_mom, ntypes, st, Mout);
    return launch_symm_t<T, 8, 8, GEN>(plan, rfe, rfr, atom_type, centres, ncentres, cnt, meta, r, maxp, lmaxp, G, ldg, use_mom,
                                      ntypes, st, Mout);
}

template <typename T>
hipError_t launch_symm(int shape, int ns, const DevPlan* plan, const double* rfe, const double* rfr,
                       const int* atom_type, const int* centres, int ncentres, const int* cnt, const int* meta,
                       const double4* r, int maxp, int lmaxp, T* G, int ldg, int use_mom, int ntypes,
                       hipStream_t st, T* Mout) {
    if (ns)
        return launch_symm_s<T, true>(shape, plan, rfe, rfr, atom_type, centres, ncentres, cnt, meta, r, maxp, lmaxp, G, ldg,
                                      use_mom, ntypes, st, Mout);
    return launch_symm_s<T, false>(shape, plan, rfe, rfr, atom_type, centres, ncentres, cnt, meta, r, maxp, lmaxp, G, ldg,
                                   use_mom, ntypes, st, Mout);
}

template <typename T, bool GEN>
static hipError_t launch_force_s(int shape, const DevPlan* plan, const double* rfe, const double* rfr,
                                 const int* atom_type, const int* centres, int ncentres, const int* cnt,
                                 const int* meta, const double4* r, int maxp, int lmaxp, const T* dEdG, int ldg, T* X,
                                 int use_mom, int ntypes, hipStream_t st) {
    if (shape == 0)
        return launch_force_t<T, 2, 4, GEN>(plan, rfe, rfr, atom_type, centres, ncentres, cnt, meta, r, maxp, lmaxp, dEdG, ldg,
                                           X, use_mom, ntypes, st);
    if (shape == 1)
        return launch_force_t<T, 4, 4, GEN>(plan, rfe, rfr, atom_type, centres, ncentres, cnt, meta, r, maxp, lmaxp, dEdG, ldg,
                                           X, use_mom, ntypes, st);
    return launch_force_t<T, 8, 8, GEN>(plan, rfe, rfr, atom_type, centres, ncentres, cnt, meta, r, maxp, lmaxp, dEdG, ldg, X,
                                       use_mom, ntypes, st);
}

template <typename T>
hipError_t launch_force(int shape, int ns, const DevPlan* plan, const double* rfe, const double* rfr,
                        const int* atom_type, const int* centres, int ncentres, const int* cnt, const int* meta,
                        const double4* r, int maxp, int lmaxp, const T* dEdG, int ldg, T* X, int use_mom, int ntypes,
                        hipStream_t st) {
    if (ns)
        return launch_force_s<T, true>(shape, plan, rfe, rfr, atom_type, centres, ncentres, cnt, meta, r, maxp, lmaxp, dEdG,
                                       ldg, X, use_mom, ntypes, st);
    return launch_force_s<T, false>(shape, plan, rfe, rfr, atom_type, centres, ncentres, cnt, meta, r, maxp, lmaxp, dEdG, ldg,
                                    X, use_mom, ntypes, st);
}

template <typename T, int NW>
static size_t force_mom_lds_w(int shape, int lmaxp, int ldg, int ntypes) {
    if (shape == 0) return force_mom_lds_bytes<T, 2, NW>(lmaxp, ldg, ntypes);
    if (shape == 1) return force_mom_lds_bytes<T, 4, NW>(lmaxp, ldg, ntypes);
    return force_mom_lds_bytes<T, 8, NW>(lmaxp, ldg, ntypes);
}
size_t force_mom_lds(int shape, bool fp64, int waves, int lmaxp, int ldg, int ntypes) {
    if (waves == 1)
        return fp64 ? force_mom_lds_w<double, 1>(shape, lmaxp, ldg, ntypes) : force_mom_lds_w<float, 1>(shape, lmaxp, ldg, ntypes);
    return fp64 ? force_mom_lds_w<double, 2>(shape, lmaxp, ldg, ntypes) : force_mom_lds_w<float, 2>(shape, lmaxp, ldg, ntypes);
}

template <typename T, int NW>
static hipError_t launch_force_mom_w(int shape, const DevPlan* plan, const double* rfe, const double* rfr,
                                     const int* atom_type, const int* centres, int ncentres, const int* cnt,
                                     const int* meta, const double4* r, int maxp, int lmaxp, const T* dEdG, int ldg,
                                     T* X, int ntypes, hipStream_t st, const T* Min) {
    if (shape == 0)
        return launch_force_mom_t<T, 2, NW>(plan, rfe, rfr, atom_type, centres, ncentres, cnt, meta, r, maxp, lmaxp,
                                            dEdG, ldg, X, ntypes, st, Min);
    if (shape == 1)
        return launch_force_mom_t<T, 4, NW>(plan, rfe, rfr, atom_type, centres, ncentres, cnt, meta, r, maxp, lmaxp,
                                            dEdG, ldg, X, ntypes, st, Min);
    return launch_force_mom_t<T, 8, NW>(plan, rfe, rfr, atom_type, centres, ncentres, cnt, meta, r, maxp, lmaxp, dEdG,
                                        ldg, X, ntypes, st, Min);
}

template <typename T>
hipError_t launch_force_mom(int shape, int waves, const DevPlan* plan, const double* rfe, const double* rfr,
                            const int* atom_type, const int* centres, int ncentres, const int* cnt, const int* meta,
                            const double4* r, int maxp, int lmaxp, const T* dEdG, int ldg, T* X, int ntypes,
                            hipStream_t st, const T* Min) {
    if (waves == 1)
        return launch_force_mom_w<T, 1>(shape, plan, rfe, rfr, atom_type, centres, ncentres, cnt, meta, r, maxp, lmaxp,
                                        dEdG, ldg, X, ntypes, st, Min);
    return launch_force_mom_w<T, 2>(shape, plan, rfe, rfr, atom_type, centres, ncentres, cnt, meta, r, maxp, lmaxp, dEdG,
                                    ldg, X, ntypes, st, Min);
}
template hipError_t launch_force_mom<double>(int, int, const DevPlan*, const double*, const double*, const int*,
                                             const int*, int, const int*, const int*, const double4*, int, int,
                                             const double*, int, double*, int, hipStream_t, const double*);
template hipError_t launch_force_mom<float>(int, int, const DevPlan*, const double*, const double*, const int*,
                                            const int*, int, const int*, const int*, const double4*, int, int,
                                            const float*, int, float*, int, hipStream_t, const float*);

template <typename T>
hipError_t launch_assemble(const AssembleArgs& a, const T* X, const EnergyTail* tail, hipStream_t st) {
    EnergyTail t{};
    if (tail) t = *tail;  // without a tail, nstruc = ncount = 0: no extra blocks
    const int ntail_pad = (t.nstruc + t.ncount + kXcds - 1) / kXcds * kXcds;
    const int chunk = (a.ntot + kXcds - 1) / kXcds;
    const dim3 grid(ntail_pad + kXcds * chunk);
    const int ncell = a.bin_atoms ? a.ncell : 0;
    if (a.mode == 0) {
        hipLaunchKernelGGL((k_assemble<T, 0>), grid, dim3(kWave), 0, st, a.S, a.atom_struc, a.cnt, a.nbr_j, a.meta, X,
                           a.maxp, a.all_images, a.F, t, a.react, nullptr, nullptr, 0, a.bin_atoms, a.pair_centres,
                           ncell, a.ntot, chunk);
        return hipGetLastError();
    }
    const size_t lds = (size_t)a.rcap * 8;
    if (lds > 65536) return hipErrorInvalidConfiguration;
    hipLaunchKernelGGL(k_react_zero, dim3((a.ntot + 255) / 256), dim3(256), 0, st, a.rcount, a.ntot, a.react);
    hipLaunchKernelGGL(k_react_fill, dim3(a.ntot), dim3(kWave), 0, st, a.S, a.atom_struc, a.cnt, a.nbr_j, a.meta,
                       a.maxp, a.all_images, a.rcount, a.rlist, a.rcap);
    hipLaunchKernelGGL((k_assemble<T, 1>), grid, dim3(kWave), lds, st, a.S, a.atom_struc, a.cnt, a.nbr_j, a.meta, X,
                       a.maxp, a.all_images, a.F, t, a.react, a.rcount, a.rlist, a.rcap, a.bin_atoms, a.pair_centres,
                       ncell, a.ntot, chunk);
    return hipGetLastError();
}

template hipError_t launch_symm<double>(int, int, const DevPlan*, const double*, const double*, const int*, const int*,
                                        int, const int*, const int*, const double4*, int, int, double*, int, int, int,
                                        hipStream_t, double*);
template hipError_t launch_symm<float>(int, int, const DevPlan*, const double*, const double*, const int*, const int*,
                                       int, const int*, const int*, const double4*, int, int, float*, int, int, int,
                                       hipStream_t, float*);
template hipError_t launch_force<double>(int, int, const DevPlan*, const double*, const double*, const int*,
                                         const int*, int, const int*, const int*, const double4*, int, int,
                                         const double*, int, double*, int, int, hipStream_t);
template hipError_t launch_force<float>(int, int, const DevPlan*, const double*, const double*, const int*, const int*,
                                        int, const int*, const int*, const double4*, int, int, const float*, int,
                                        float*, int, int, hipStream_t);
template hipError_t launch_assemble<double>(const AssembleArgs&, const double*, const EnergyTail*, hipStream_t);
template hipError_t launch_assemble<float>(const AssembleArgs&, const float*, const EnergyTail*, hipStream_t);

}  // namespace atomnn
