// Energy-training inner loop on the device: the training part of NNPTE::error
// (src/nnp/nnpte.cpp:869-927) with Cost::grad (src/ml/nn.cpp:1554-1589), for a training set
// whose descriptors are computed once on the GPU (NNP::symm, nnpte.cpp:1976-1977).
//
// Per call (parameter vector x, batch of structures):
//   1. forward pass of every batch atom (per-species 16-atom tiles, FP64 MFMA as the network
//      kernels): out_i = outb + outw a_L (no type energy, as the reference's training energy);
//   2. per structure b: E_b = sum_i out_i (fixed-order tree), dE_b = (E_b - E_ref) / N_b, the loss
//      and dcdo_b = dC/dE_b / N_b of the loss function (MSE, MAE, HUBER, ASINH);
//   3. forward again, then backward per tile: dcdz_L = dcdo outw .* dadz_L, dcdz_{l-1} =
//      (W_l^T dcdz_l) .* dadz_{l-1}; the tile's gradient (bias: sum over its 16 atoms of dcdz_l;
//      weights: dcdz_l^T a_{l-1} as 16 x 16 MFMA tiles with K = 16 atoms) goes to its own row of
//      a partial buffer in the reference's parameter order (operator>>, nn.cpp:797-831: every
//      layer's biases, then every weight matrix column-major);
//   4. the partial rows of a species are summed in a fixed order: groups of consecutive tiles in
//      tile order, then the groups in order.
// Force training (error_force) adds beta/2 |(F_i - F_ref,i) / N|^2 per atom, the forces being the
// model's force kernels applied to the networks' dout/dG at x; its gradient runs the force pass
// backwards: force residuals, descriptor tangents (k_symm_jvp) and a tangent forward/backward
// network pass per tile (k_train_force_tile).
// Every sum has a fixed order, so results do not depend on timing; they match the CPU
// restatement to rounding (different summation order).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <vector>

#include "../../include/nnp_capi.h"
#include "basis.hpp"
#include "engine.hpp"
#include "netmm.hpp"

namespace atomnn {

hipError_t launch_symm_jvp(int shape, int ns, const DevPlan* plan, const double* rfe, const double* rfr,
                           const int* atom_type, const DevStruc* S, const int* atom_struc, int ntot, const int* cnt,
                           const int* nbr_j, const int* meta, const double4* r, int maxp, int lmaxp, const double* gF,
                           int all_images, double* V, int ldg, int use_mom, int ntypes, hipStream_t st);

// Per-species network for training.  W_l (padded, row-major wpad_l x in_pad_l) and b_l (padded)
// live in the training parameter buffer; inpw / inpb stay in the model's weight buffer.
struct TrainNet {
    int L, nin, nin_pad, act;
    int nodes[kMaxLayers], wpad[kMaxLayers];
    int64_t woff[kMaxLayers], boff[kMaxLayers];  // into the training parameter buffer
    int rb[kMaxLayers], rw[kMaxLayers];          // offsets of b_l and w_l in the species' parameter vector
    int nparams;
    int64_t inpw_off, inpb_off;  // into the model's weight buffer
    double outw, outb;
};
struct TrainPlan {
    int T;
    TrainNet net[kMaxSpecies];
};
// Species ranges of a batch: tiles, species-sorted batch atoms, partial-buffer rows.
struct TrainTiles {
    int tile_start[kMaxSpecies + 1];
    int atom_start[kMaxSpecies + 1];
    int64_t pbase[kMaxSpecies + 1];  // first partial element of the species' first tile
};

constexpr int kTrainWaves = 4;
constexpr int kTrainThreads = kTrainWaves * kWave;

__device__ __forceinline__ int train_species(const TrainTiles& tt, int ntypes, int tile) {
    int s = 0;
    while (s + 1 < ntypes && tile >= tt.tile_start[s + 1]) ++s;
    return s;
}

// Hidden-layer activation and its derivatives.  GEN = false: tanh hidden layers only (the common
// case; the general switch over every activation costs the kernels about 50 VGPRs).
template <bool GEN>
__device__ __forceinline__ void train_act(int act, double z, double& a, double& d) {
    if constexpr (GEN) {
        act_eval<double>(act, z, a, d);
    } else if (act == NEU_LINEAR) {
        a = z;
        d = 1.0;
    } else {
        a = tanh_m1(z);
        d = 1.0 - a * a;
    }
}
template <bool GEN>
__device__ __forceinline__ double train_act_d2(int act, double z, double a) {
    if constexpr (GEN) return act_d2(act, z);
    return act == NEU_LINEAR ? 0.0 : -2.0 * a * (1.0 - a * a);
}

// LDS: ins [16][nin_pad], then per layer A_l [16][wpad_l] and D_l [16][wpad_l] (dadz; dcdz in the
// backward pass, in place).
static size_t train_lds_bytes(const TrainNet& n) {
    size_t e = 16 * (size_t)n.nin_pad;
    for (int l = 0; l < n.L; ++l) e += 2 * 16 * (size_t)n.wpad[l];
    return 8 * e;
}

// MODE 0: forward, out[slot] = outb + outw a_L.  MODE 1: forward and backward, the tile's
// gradient row into `partial`.  MODE 2: forward and the input gradient (fpbp + DODZ::grad,
// nn.cpp:1191-1233, 1620-1628): out[slot] and Y[slot] = dout/dG = inpw .* (w_0^T dodz_0).
// MODE 3: out[slot] and the gradient row of the outputs (seed outw instead of dcdo outw), for
// tiles whose atoms belong to one structure: the reduction then weights the row by dcdo.
// Rows beyond the species' atoms or with batom < 0 (padding) carry nothing.
template <int MODE, bool GEN>
__global__ __launch_bounds__(kTrainThreads) void k_train_tile(
    const TrainPlan* __restrict__ tp, const double* __restrict__ Wm, const double* __restrict__ par, TrainTiles tt,
    int tile0, const int* __restrict__ batom, const int* __restrict__ bslot, const int* __restrict__ bstruct,
    const double* __restrict__ G, int ldg, double* __restrict__ out, const double* __restrict__ dcdo,
    double* __restrict__ partial, double* __restrict__ Y) {
    extern __shared__ __align__(16) double tsh[];
    const int tile = tile0 + blockIdx.x;
    const int s = train_species(tt, tp->T, tile);
    const TrainNet& N = tp->net[s];
    const int a0 = tt.atom_start[s], M = tt.atom_start[s + 1] - a0;
    const int r0 = (tile - tt.tile_start[s]) * 16;
    const int tid = threadIdx.x, wv = tid / kWave, lane = tid % kWave, li = lane & 15, lk = lane >> 4;
    double* ins = tsh;
    double* A[kMaxLayers];
    double* D[kMaxLayers];
    {
        double* p = ins + 16 * N.nin_pad;
        for (int l = 0; l < N.L; ++l) {
            A[l] = p;
            D[l] = p + 16 * N.wpad[l];
            p += 32 * N.wpad[l];
        }
    }
    for (int idx = tid; idx < 16 * N.nin_pad; idx += kTrainThreads) {
        const int r = idx / N.nin_pad, k = idx % N.nin_pad;
        double v = 0.0;
        const int src = r0 + r < M ? batom[a0 + r0 + r] : -1;  // -1: padding row
        if (src >= 0 && k < N.nin) v = Wm[N.inpw_off + k] * (G[(int64_t)src * ldg + k] + Wm[N.inpb_off + k]);
        ins[r * N.nin_pad + k] = v;
    }
    __syncthreads();
    for (int l = 0; l < N.L; ++l) {
        const double* prev = l == 0 ? ins : A[l - 1];
        const int in_pad = l == 0 ? N.nin_pad : N.wpad[l - 1], wp = N.wpad[l];
        const int act = l == N.L - 1 ? NEU_LINEAR : N.act;
        for (int jt = wv; jt < wp / 16; jt += kTrainWaves) {
            const d4_t acc = tile_mm_f64<true>(prev, in_pad, par + N.woff[l], in_pad, in_pad, jt, li, lk);
            const int col = jt * 16 + li;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = 4 * v + lk;
                double av, dv;
                train_act<GEN>(act, acc[v] + par[N.boff[l] + col], av, dv);
                A[l][row * wp + col] = av;
                D[l][row * wp + col] = dv;
            }
        }
        __syncthreads();
    }
    const int L = N.L;
    if constexpr (MODE != 1) {
        if (tid < 16 && r0 + tid < M && batom[a0 + r0 + tid] >= 0)
            out[bslot[a0 + r0 + tid]] = N.outb + A[L - 1][tid * N.wpad[L - 1]] * N.outw;
    }
    if constexpr (MODE == 0) {
        return;
    } else if constexpr (MODE == 2) {
        // dodz_L = outw .* dadz_L (unit 0), dodz_{l-1} = (w_l^T dodz_l) .* dadz_{l-1}
        {
            const int wp = N.wpad[L - 1];
            for (int idx = tid; idx < 16 * wp; idx += kTrainThreads) D[L - 1][idx] *= (idx % wp == 0) ? N.outw : 0.0;
        }
        __syncthreads();
        for (int l = L - 1; l > 0; --l) {
            const int wp = N.wpad[l], wq = N.wpad[l - 1];
            for (int jt = wv; jt < wq / 16; jt += kTrainWaves) {
                const d4_t acc = tile_mm_f64<false>(D[l], wp, par + N.woff[l], wq, wp, jt, li, lk);
#pragma unroll
                for (int v = 0; v < 4; ++v) D[l - 1][(4 * v + lk) * wq + jt * 16 + li] *= acc[v];
            }
            __syncthreads();
        }
        const int wp = N.wpad[0];
        for (int jt = wv; jt < N.nin_pad / 16; jt += kTrainWaves) {
            const d4_t acc = tile_mm_f64<false>(D[0], wp, par + N.woff[0], N.nin_pad, wp, jt, li, lk);
            const int col = jt * 16 + li;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = 4 * v + lk;
                if (r0 + row < M && col < N.nin && batom[a0 + r0 + row] >= 0)
                    Y[(int64_t)bslot[a0 + r0 + row] * ldg + col] = acc[v] * Wm[N.inpw_off + col];
            }
        }
    } else {
        // dcdz_L = dcdo outw .* dadz_L (single output unit; padded units and rows carry zero)
        {
            const int wp = N.wpad[L - 1];
            for (int idx = tid; idx < 16 * wp; idx += kTrainThreads) {
                const int r = idx / wp, n = idx % wp;
                const bool real = r0 + r < M && n < N.nodes[L - 1] && batom[a0 + r0 + r] >= 0;
                const double c = real ? (MODE == 3 ? N.outw : dcdo[bstruct[a0 + r0 + r]] * N.outw) : 0.0;
                D[L - 1][idx] *= c;
            }
        }
        __syncthreads();
        double* prow = partial + tt.pbase[s] + (int64_t)(tile - tt.tile_start[s]) * N.nparams;
        for (int l = L - 1; l >= 0; --l) {
            const double* prev = l == 0 ? ins : A[l - 1];
            const int wp = N.wpad[l], in_pad = l == 0 ? N.nin_pad : N.wpad[l - 1];
            const int in_real = l == 0 ? N.nin : N.nodes[l - 1], no = N.nodes[l];
            const double* dz = D[l];  // dcdz_l [16][wp]
            // bias gradient: sum over the tile's 16 atoms
            for (int n = tid; n < no; n += kTrainThreads) {
                double g = 0.0;
#pragma unroll
                for (int r = 0; r < 16; ++r) g += dz[r * wp + n];
                prow[N.rb[l] + n] = g;
            }
            // weight gradient dcdz_l^T a_{l-1}: tiles (nt, mt) of 16 x 16, K = the 16 atoms
            const int ntn = wp / 16, ntm = in_pad / 16;
            for (int t = wv; t < ntn * ntm; t += kTrainWaves) {
                const int nt = t / ntm, mt = t % ntm;
                d4_t acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int r = 4 * q + lk;
                    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(dz[r * wp + nt * 16 + li], prev[r * in_pad + mt * 16 + li],
                                                               acc, 0, 0, 0);
                }
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int n = nt * 16 + 4 * v + lk, m = mt * 16 + li;
                    if (n < no && m < in_real) prow[N.rw[l] + m * no + n] = acc[v];
                }
            }
            // dcdz_{l-1} = (W_l^T dcdz_l) .* dadz_{l-1}, in place over dadz_{l-1}
            if (l > 0) {
                const int wq = N.wpad[l - 1];
                for (int jt = wv; jt < wq / 16; jt += kTrainWaves) {
                    const d4_t acc = tile_mm_f64<false>(dz, wp, par + N.woff[l], wq, wp, jt, li, lk);
                    const int col = jt * 16 + li;
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int row = 4 * v + lk;
                        D[l - 1][row * wq + col] *= acc[v];
                    }
                }
            }
            __syncthreads();
        }
    }
}

// Force training, one 16-atom tile: the gradient of dcdo out + (dout/dG . V) with respect to the
// parameters (the energy term, and the term through which the forces depend on the network:
// V = dL/d(dout/dG) from the descriptor tangents).  dout/dG . V is the network's tangent output
// for the input tangent inpw .* V, so:
//   forward:  z_l = w_l a_(l-1) + b_l,  a_l = act(z_l),  zt_l = w_l t_(l-1),  t_l = act'(z_l) zt_l
//   reverse:  pz_l = p_l .* act'_l,  qz_l = q_l .* act'_l + p_l .* act''_l .* zt_l,
//             grad b_l = sum qz_l,  grad w_l = pz_l^T t_(l-1) + qz_l^T a_(l-1)  (K = the 16 atoms),
//             p_(l-1) = w_l^T pz_l,  q_(l-1) = w_l^T qz_l,  p_L = outw,  q_L = dcdo outw
// (D2ODZDI::grad, nn.cpp:1799-1890, contracted with V; Cost::grad, nn.cpp:1554-1589).
// LDS: ins and tin [16][nin_pad], then per layer A (a), D (act', then pz), T (t) and H (act'' zt,
// then qz), each [16][wpad_l].
static size_t train_force_lds_bytes(const TrainNet& n) {
    size_t e = 2 * 16 * (size_t)n.nin_pad;
    for (int l = 0; l < n.L; ++l) e += 4 * 16 * (size_t)n.wpad[l];
    return 8 * e;
}

template <bool GEN>
__global__ __launch_bounds__(kTrainThreads) void k_train_force_tile(
    const TrainPlan* __restrict__ tp, const double* __restrict__ Wm, const double* __restrict__ par, TrainTiles tt,
    int tile0, const int* __restrict__ batom, const int* __restrict__ bstruct, const double* __restrict__ G,
    const double* __restrict__ V, int ldg, const double* __restrict__ dcdo, double* __restrict__ partial) {
    extern __shared__ __align__(16) double tsh[];
    const int tile = tile0 + blockIdx.x;
    const int s = train_species(tt, tp->T, tile);
    const TrainNet& N = tp->net[s];
    const int a0 = tt.atom_start[s], M = tt.atom_start[s + 1] - a0;
    const int r0 = (tile - tt.tile_start[s]) * 16;
    const int tid = threadIdx.x, wv = tid / kWave, lane = tid % kWave, li = lane & 15, lk = lane >> 4;
    const int L = N.L;
    double* ins = tsh;
    double* tin = ins + 16 * N.nin_pad;
    double *A[kMaxLayers], *D[kMaxLayers], *T[kMaxLayers], *H[kMaxLayers];
    {
        double* p = tin + 16 * N.nin_pad;
        for (int l = 0; l < L; ++l) {
            const int sz = 16 * N.wpad[l];
            A[l] = p;
            D[l] = p + sz;
            T[l] = p + 2 * sz;
            H[l] = p + 3 * sz;
            p += 4 * sz;
        }
    }
    for (int idx = tid; idx < 16 * N.nin_pad; idx += kTrainThreads) {
        const int r = idx / N.nin_pad, k = idx % N.nin_pad;
        double v = 0.0, t = 0.0;
        if (r0 + r < M && k < N.nin) {
            const int64_t src = (int64_t)batom[a0 + r0 + r] * ldg + k;
            const double w = Wm[N.inpw_off + k];
            v = w * (G[src] + Wm[N.inpb_off + k]);
            t = w * V[src];
        }
        ins[idx] = v;
        tin[idx] = t;
    }
    __syncthreads();
    for (int l = 0; l < L; ++l) {
        const double* prev = l == 0 ? ins : A[l - 1];
        const double* tprev = l == 0 ? tin : T[l - 1];
        const int in_pad = l == 0 ? N.nin_pad : N.wpad[l - 1], wp = N.wpad[l];
        const bool last = l == L - 1;
        const int act = last ? NEU_LINEAR : N.act;
        for (int jt = wv; jt < wp / 16; jt += kTrainWaves) {
            const d4_t az = tile_mm_f64<true>(prev, in_pad, par + N.woff[l], in_pad, in_pad, jt, li, lk);
            const d4_t at = tile_mm_f64<true>(tprev, in_pad, par + N.woff[l], in_pad, in_pad, jt, li, lk);
            const int col = jt * 16 + li;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int idx = (4 * v + lk) * wp + col;
                const double z = az[v] + par[N.boff[l] + col];
                double av, dv;
                train_act<GEN>(act, z, av, dv);
                const double d2 = last ? 0.0 : train_act_d2<GEN>(act, z, av);
                A[l][idx] = av;
                D[l][idx] = dv;
                T[l][idx] = dv * at[v];
                H[l][idx] = d2 * at[v];
            }
        }
        __syncthreads();
    }
    // seeds on the output unit of real rows: p = outw, q = dcdo outw
    {
        const int wp = N.wpad[L - 1];
        for (int idx = tid; idx < 16 * wp; idx += kTrainThreads) {
            const int r = idx / wp, n = idx % wp;
            const double p = (r0 + r < M && n == 0) ? N.outw : 0.0;
            const double q = p != 0.0 ? p * dcdo[bstruct[a0 + r0 + r]] : 0.0;
            const double dv = D[L - 1][idx], hv = H[L - 1][idx];
            D[L - 1][idx] = p * dv;
            H[L - 1][idx] = q * dv + p * hv;
        }
    }
    __syncthreads();
    double* prow = partial + tt.pbase[s] + (int64_t)(tile - tt.tile_start[s]) * N.nparams;
    for (int l = L - 1; l >= 0; --l) {
        const double* aprev = l == 0 ? ins : A[l - 1];
        const double* tprev = l == 0 ? tin : T[l - 1];
        const int wp = N.wpad[l], in_pad = l == 0 ? N.nin_pad : N.wpad[l - 1];
        const int in_real = l == 0 ? N.nin : N.nodes[l - 1], no = N.nodes[l];
        const double* pz = D[l];
        const double* qz = H[l];
        for (int n = tid; n < no; n += kTrainThreads) {
            double g = 0.0;
#pragma unroll
            for (int r = 0; r < 16; ++r) g += qz[r * wp + n];
            prow[N.rb[l] + n] = g;
        }
        const int ntn = wp / 16, ntm = in_pad / 16;
        for (int t = wv; t < ntn * ntm; t += kTrainWaves) {
            const int nt = t / ntm, mt = t % ntm;
            d4_t acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int r = 4 * q + lk;
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(pz[r * wp + nt * 16 + li], tprev[r * in_pad + mt * 16 + li], acc,
                                                           0, 0, 0);
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(qz[r * wp + nt * 16 + li], aprev[r * in_pad + mt * 16 + li], acc,
                                                           0, 0, 0);
            }
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int n = nt * 16 + 4 * v + lk, m = mt * 16 + li;
                if (n < no && m < in_real) prow[N.rw[l] + m * no + n] = acc[v];
            }
        }
        if (l > 0) {
            const int wq = N.wpad[l - 1];
            for (int jt = wv; jt < wq / 16; jt += kTrainWaves) {
                const d4_t pm = tile_mm_f64<false>(pz, wp, par + N.woff[l], wq, wp, jt, li, lk);
                const d4_t qm = tile_mm_f64<false>(qz, wp, par + N.woff[l], wq, wp, jt, li, lk);
                const int col = jt * 16 + li;
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int idx = (4 * v + lk) * wq + col;
                    const double dv = D[l - 1][idx], hv = H[l - 1][idx];
                    D[l - 1][idx] = pm[v] * dv;
                    H[l - 1][idx] = qm[v] * dv + pm[v] * hv;
                }
            }
        }
        __syncthreads();
    }
}

// Force residuals of one batch structure b (atoms [bstart[b], bstart[b+1]) of the workspace,
// reference forces from training atom fstart[b]): fdiff = (F - F_ref) / N, gF = beta fdiff / N
// (dL/dF of beta/2 |fdiff|^2), f2[b] = sum |fdiff|^2 (fixed-order tree; nnptef.cpp:1163-1166).
__global__ __launch_bounds__(256) void k_train_fresid(const double* __restrict__ F, const double* __restrict__ fref,
                                                      const int* __restrict__ bstart, const int64_t* __restrict__ fstart,
                                                      double beta, double* __restrict__ gF, double* __restrict__ f2) {
    __shared__ double part[256];
    const int b = blockIdx.x, t = threadIdx.x;
    const int o0 = bstart[b], n = bstart[b + 1] - o0;
    const double norm = 1.0 / n;
    double acc = 0.0;
    for (int i = t; i < n; i += 256) {
        const int64_t a = o0 + i, g = fstart[b] + i;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double fd = (F[3 * a + c] - fref[3 * g + c]) * norm;
            gF[3 * a + c] = beta * norm * fd;
            acc += fd * fd;
        }
    }
    part[t] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (t < o) part[t] += part[t + o];
        __syncthreads();
    }
    if (t == 0) f2[b] = part[0];
}

// Per batch structure: E = sum of its atoms' outputs (fixed-order tree), dE = (E - E_ref) / N,
// the loss and dcdo (nnpte.cpp:876-916).
__global__ __launch_bounds__(256) void k_train_struct(const double* __restrict__ out, const int* __restrict__ bstart,
                                                      const double* __restrict__ eref_b, int loss, double delta,
                                                      double* __restrict__ dcdo, double* __restrict__ lossv,
                                                      double* __restrict__ dE2) {
    __shared__ double part[256];
    const int b = blockIdx.x, t = threadIdx.x;
    const int o0 = bstart[b], n = bstart[b + 1] - o0, chunk = (n + 255) / 256;
    double acc = 0.0;
    for (int i = o0 + t * chunk; i < o0 + min(n, (t + 1) * chunk); ++i) acc += out[i];
    part[t] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (t < o) part[t] += part[t + o];
        __syncthreads();
    }
    if (t == 0) {
        const double norm = 1.0 / n;
        const double dE = (part[0] - eref_b[b]) * norm;
        double c = norm, L = 0.0;
        const double deltai = 1.0 / delta, delta2 = delta * delta;
        if (loss == NNP_LOSS_MSE) {
            L = 0.5 * dE * dE;
            c *= dE;
        } else if (loss == NNP_LOSS_MAE) {
            L = fabs(dE);
            c *= (double)((0.0 < dE) - (dE < 0.0));
        } else if (loss == NNP_LOSS_HUBER) {
            const double arg = dE * deltai, sq = sqrt(1.0 + arg * arg);
            L = delta2 * (sq - 1.0);
            c *= dE / sq;
        } else {
            const double arg = dE * deltai, sq = sqrt(1.0 + arg * arg), lg = log(arg + sq);
            L = delta2 * (1.0 - sq + arg * lg);
            c *= lg * delta;
        }
        dcdo[b] = c;
        lossv[b] = L;
        dE2[b] = dE * dE;
    }
}

// Gradient rows summed in two fixed-order levels (all rows of a level in parallel):
// level 1: groups of kReduceGroup consecutive tiles of a species, each summed in tile order into
// grp[s][g][p]; level 2: grad[goff_s + p] (+)= the groups in order.
constexpr int kReduceGroup = 32;
__global__ __launch_bounds__(256) void k_train_reduce1(const TrainPlan* __restrict__ tp, TrainTiles tt, int tile0,
                                                       int tile1, const double* __restrict__ partial, int gmax,
                                                       int64_t gstride, double* __restrict__ grp,
                                                       const int* __restrict__ tstruct, const double* __restrict__ dcdo) {
    const int s = blockIdx.z, g = blockIdx.y;
    const TrainNet& N = tp->net[s];
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= N.nparams || g >= gmax) return;
    const int t0 = max(tile0, tt.tile_start[s]), t1 = min(tile1, tt.tile_start[s + 1]);
    const int a = t0 + g * kReduceGroup, b = min(t1, a + kReduceGroup);
    double acc = 0.0;
    if (tstruct) {  // rows of single-structure tiles, weighted by their structure's dcdo
        for (int t = a; t < b; ++t)
            acc += dcdo[tstruct[t]] * partial[tt.pbase[s] + (int64_t)(t - tt.tile_start[s]) * N.nparams + p];
    } else {
        for (int t = a; t < b; ++t) acc += partial[tt.pbase[s] + (int64_t)(t - tt.tile_start[s]) * N.nparams + p];
    }
    grp[(int64_t)s * gstride + (int64_t)g * N.nparams + p] = acc;
}
__global__ __launch_bounds__(256) void k_train_reduce2(const TrainPlan* __restrict__ tp, TrainTiles tt, int tile0,
                                                       int tile1, const double* __restrict__ grp, int64_t gstride,
                                                       const int64_t* __restrict__ goff, double* __restrict__ grad,
                                                       int accumulate) {
    const int s = blockIdx.y;
    const TrainNet& N = tp->net[s];
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= N.nparams) return;
    const int t0 = max(tile0, tt.tile_start[s]), t1 = min(tile1, tt.tile_start[s + 1]);
    const int ng = t1 > t0 ? (t1 - t0 + kReduceGroup - 1) / kReduceGroup : 0;
    double acc = accumulate ? grad[goff[s] + p] : 0.0;
    for (int g = 0; g < ng; ++g) acc += grp[(int64_t)s * gstride + (int64_t)g * N.nparams + p];
    grad[goff[s] + p] = acc;
}

// ------------------------------------------------------------------ host side
template <int MODE>
static void launch_tile(bool gen, dim3 grid, size_t lds, hipStream_t st, const TrainPlan* tp, const double* Wm,
                        const double* par, const TrainTiles& tt, int tile0, const int* batom, const int* bslot,
                        const int* bstruct, const double* G, int ldg, double* out, const double* dcdo, double* partial,
                        double* Y) {
    if (gen)
        hipLaunchKernelGGL((k_train_tile<MODE, true>), grid, dim3(kTrainThreads), lds, st, tp, Wm, par, tt, tile0, batom,
                           bslot, bstruct, G, ldg, out, dcdo, partial, Y);
    else
        hipLaunchKernelGGL((k_train_tile<MODE, false>), grid, dim3(kTrainThreads), lds, st, tp, Wm, par, tt, tile0, batom,
                           bslot, bstruct, G, ldg, out, dcdo, partial, Y);
}

class TrainSet {
  public:
    TrainSet(Engine& owner, int n, const nnp_structure* ss, const double* energy) : owner_(owner) {
        if (n <= 0 || !ss || !energy) throw std::invalid_argument("nnp_train_create: invalid argument");
        const HostModel& hm = owner.host;
        const DevPlan& P = owner.pb.plan;
        T_ = hm.T;
        ldg_ = P.ldg;
        // network layout
        std::memset(&plan_, 0, sizeof(plan_));
        plan_.T = T_;
        int64_t pofs = 0;
        goff_.assign(T_ + 1, 0);
        for (int t = 0; t < T_; ++t) {
            const NetSpec& ns = hm.sp[t].net;
            const DevNet& dn = P.net[t];
            TrainNet& N = plan_.net[t];
            N.L = dn.L;
            N.nin = ns.nin;
            N.nin_pad = dn.nin_pad;
            N.act = dn.act;
            N.inpw_off = dn.inpw_off;
            N.inpb_off = dn.inpb_off;
            N.outw = dn.outw;
            N.outb = dn.outb;
            gen_ = gen_ || (N.L > 1 && N.act != NEU_TANH && N.act != NEU_LINEAR);
            if (ns.nodes.back() != 1) throw unsupported_error("nnp_train: networks with more than one output");
            int rb = 0, nb = 0;
            for (int l = 0; l < N.L; ++l) nb += ns.nodes[l];
            int rw = nb;
            for (int l = 0; l < N.L; ++l) {
                const int in_pad = l == 0 ? N.nin_pad : dn.wpad[l - 1];
                const int in_real = l == 0 ? ns.nin : ns.nodes[l - 1];
                N.nodes[l] = ns.nodes[l];
                N.wpad[l] = dn.wpad[l];
                N.woff[l] = pofs;
                pofs += (int64_t)dn.wpad[l] * in_pad;
                N.boff[l] = pofs;
                pofs += dn.wpad[l];
                pofs = (pofs + 1) & ~int64_t(1);
                N.rb[l] = rb;
                rb += ns.nodes[l];
                N.rw[l] = rw;
                rw += ns.nodes[l] * in_real;
            }
            N.nparams = rw;
            goff_[t + 1] = goff_[t] + N.nparams;
        }
        npar_dev_ = pofs;
        hpar_.assign(npar_dev_, 0.0);
        d_plan_.reserve(sizeof(TrainPlan));
        HIPCHK(hipMemcpy(d_plan_.p, &plan_, sizeof(TrainPlan), hipMemcpyHostToDevice));
        d_goff_.reserve(sizeof(int64_t) * (T_ + 1));
        HIPCHK(hipMemcpy(d_goff_.p, goff_.data(), sizeof(int64_t) * (T_ + 1), hipMemcpyHostToDevice));
        // descriptors of the training set, computed once in chunks by a workspace of the model
        natoms_.resize(n);
        aoff_.assign(n + 1, 0);
        for (int k = 0; k < n; ++k) {
            if (ss[k].natoms <= 0) throw std::invalid_argument("nnp_train_create: empty structure");
            natoms_[k] = ss[k].natoms;
            aoff_[k + 1] = aoff_[k] + ss[k].natoms;
            for (int i = 0; i < ss[k].natoms; ++i) types_.push_back(ss[k].type[i]);
            if (!ss[k].pos || !ss[k].type) throw std::invalid_argument("nnp_train_create: null structure arrays");
            hpos_.insert(hpos_.end(), ss[k].pos, ss[k].pos + 3 * (size_t)ss[k].natoms);
            hopen_.push_back(ss[k].cell ? 0 : 1);
            for (int q = 0; q < 9; ++q) hcell_.push_back(ss[k].cell ? ss[k].cell[q] : 0.0);
        }
        eref_.assign(energy, energy + n);
        ntot_ = aoff_[n];
        d_G_.reserve((size_t)ntot_ * ldg_ * 8);
        Engine E(owner, 1000);
        int k0 = 0;
        while (k0 < n) {
            int k1 = k0;
            int64_t atoms = 0;
            while (k1 < n && (k1 == k0 || atoms + natoms_[k1] <= (1 << 18))) atoms += natoms_[k1++];
            std::vector<nnp_structure> chunk(ss + k0, ss + k1);
            E.upload(chunk);
            E.build_neighbors(false);
            E.descriptors<double>();
            HIPCHK(hipMemcpyAsync((double*)d_G_.p + (size_t)aoff_[k0] * ldg_, E.d_G.p, (size_t)atoms * ldg_ * 8,
                                  hipMemcpyDeviceToDevice, E.stream));
            HIPCHK(hipStreamSynchronize(E.stream));
            k0 = k1;
        }
        HIPCHK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    }
    ~TrainSet() {
        if (stream_) (void)hipStreamDestroy(stream_);
    }

    int64_t nparams(int t) const {
        if (t < 0 || t >= T_) throw std::invalid_argument("nnp_train_nparams: species out of range");
        return plan_.net[t].nparams;
    }
    int64_t total() const { return goff_[T_]; }

    // the model's current parameters in the reference order (operator>>)
    void get_params(double* x) const {
        for (int t = 0; t < T_; ++t) {
            const NetSpec& ns = owner_.host.sp[t].net;
            const TrainNet& N = plan_.net[t];
            double* y = x + goff_[t];
            for (int l = 0; l < N.L; ++l)
                for (int n = 0; n < N.nodes[l]; ++n) y[N.rb[l] + n] = ns.b[l][n];
            for (int l = 0; l < N.L; ++l) {
                const int in_real = l == 0 ? ns.nin : ns.nodes[l - 1];
                for (int m = 0; m < in_real; ++m)
                    for (int n = 0; n < N.nodes[l]; ++n) y[N.rw[l] + m * N.nodes[l] + n] = ns.W[l][(size_t)n * in_real + m];
            }
        }
    }

    struct Tiles {
        TrainTiles tt;
        int tiles = 0;
        int64_t pb = 0;  // partial elements of all tiles
    };

    void check_args(const char* fn, const double* x, int nbatch, const int* batch, int loss, double delta,
                    const double* err) const {
        if (!x || !err || nbatch <= 0 || !batch) throw std::invalid_argument(std::string(fn) + ": invalid argument");
        if (loss < NNP_LOSS_MSE || loss > NNP_LOSS_ASINH) throw std::invalid_argument(std::string(fn) + ": unknown loss");
        if ((loss == NNP_LOSS_HUBER || loss == NNP_LOSS_ASINH) && !(delta > 0))
            throw std::invalid_argument(std::string(fn) + ": delta must be positive");
        const int nst = (int)natoms_.size();
        for (int b = 0; b < nbatch; ++b)
            if (batch[b] < 0 || batch[b] >= nst) throw std::invalid_argument(std::string(fn) + ": structure out of range");
    }

    // parameters: reference order -> padded row-major layers, uploaded on stream st
    void pack_params(const double* x, hipStream_t st) {
        for (int t = 0; t < T_; ++t) {
            const TrainNet& N = plan_.net[t];
            const double* y = x + goff_[t];
            for (int l = 0; l < N.L; ++l) {
                const int in_pad = l == 0 ? N.nin_pad : N.wpad[l - 1];
                const int in_real = l == 0 ? N.nin : N.nodes[l - 1];
                for (int n = 0; n < N.nodes[l]; ++n) {
                    hpar_[N.boff[l] + n] = y[N.rb[l] + n];
                    for (int m = 0; m < in_real; ++m) hpar_[N.woff[l] + (int64_t)n * in_pad + m] = y[N.rw[l] + m * N.nodes[l] + n];
                }
            }
        }
        d_par_.reserve(8 * (size_t)npar_dev_);
        HIPCHK(hipMemcpyAsync(d_par_.p, hpar_.data(), 8 * (size_t)npar_dev_, hipMemcpyHostToDevice, st));
        HIPCHK(hipStreamSynchronize(st));  // hpar_ is reused by the next call
    }

    // Species tiles of 16 atoms over the given (source row, slot, structure) lists per species,
    // uploaded into d_batom_, d_bslot_, d_bstruct_ on stream st.
    Tiles make_tiles(const std::vector<std::vector<int>>& atom_s, const std::vector<std::vector<int>>& slot_s,
                     const std::vector<std::vector<int>>& struct_s, hipStream_t st) {
        Tiles r;
        std::memset(&r.tt, 0, sizeof(r.tt));
        hbatom_.clear();
        hbslot_.clear();
        hbstruct_.clear();
        for (int t = 0; t <= kMaxSpecies; ++t) {
            r.tt.tile_start[t] = r.tiles;
            r.tt.atom_start[t] = (int)hbatom_.size();
            r.tt.pbase[t] = r.pb;
            if (t < T_) {
                hbatom_.insert(hbatom_.end(), atom_s[t].begin(), atom_s[t].end());
                hbslot_.insert(hbslot_.end(), slot_s[t].begin(), slot_s[t].end());
                hbstruct_.insert(hbstruct_.end(), struct_s[t].begin(), struct_s[t].end());
                const int nt = ((int)atom_s[t].size() + 15) / 16;
                r.tiles += nt;
                r.pb += (int64_t)nt * plan_.net[t].nparams;
            }
        }
        const size_t nba = std::max<size_t>(hbatom_.size(), 1);
        d_batom_.reserve(4 * nba);
        d_bslot_.reserve(4 * nba);
        d_bstruct_.reserve(4 * nba);
        HIPCHK(hipMemcpyAsync(d_batom_.p, hbatom_.data(), 4 * hbatom_.size(), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(d_bslot_.p, hbslot_.data(), 4 * hbslot_.size(), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(d_bstruct_.p, hbstruct_.data(), 4 * hbstruct_.size(), hipMemcpyHostToDevice, st));
        return r;
    }

    // Gradient rows of every tile, in chunks of tiles (at most the model's train_chunk_elems row
    // elements at a time, 2^28 = 2 GB by default):
    // launch(tc, c0, c1) writes the rows of tiles [c0, c1) at the shifted bases of tc; the rows are
    // then added into d_grad_ in a fixed order (groups of consecutive tiles, then the groups).
    template <typename F>
    void sum_tiles(const Tiles& tl, hipStream_t st, F&& launch) {
        if (tl.tiles == 0) return;
        int64_t maxnp = 1;
        for (int t = 0; t < T_; ++t) maxnp = std::max<int64_t>(maxnp, plan_.net[t].nparams);
        const int chunk = (int)std::max<int64_t>(
            1, std::min<int64_t>({(int64_t)tl.tiles, owner_.train_chunk_elems / maxnp, (int64_t)kReduceGroup * 65535}));
        d_partial_.reserve(8 * (size_t)std::min<int64_t>(tl.pb, (int64_t)chunk * maxnp + maxnp));
        for (int c0 = 0; c0 < tl.tiles; c0 += chunk) {
            const int c1 = std::min(tl.tiles, c0 + chunk);
            // partial rows of this chunk start at element 0: shift each species' base
            TrainTiles tc = tl.tt;
            int64_t base = 0;
            for (int t = 0; t < T_; ++t) {
                const int t0 = std::max(c0, tl.tt.tile_start[t]), t1 = std::min(c1, tl.tt.tile_start[t + 1]);
                tc.pbase[t] = base - (int64_t)(t0 - tl.tt.tile_start[t]) * plan_.net[t].nparams;
                if (t1 > t0) base += (int64_t)(t1 - t0) * plan_.net[t].nparams;
            }
            launch(tc, c0, c1);
            HIPCHK(hipGetLastError());
            reduce_rows(tl, tc, c0, c1, st, nullptr, nullptr);
        }
    }

    // d_grad_ += the rows of tiles [c0, c1) (bases tc), optionally weighted by dcdo[tstruct[tile]].
    void reduce_rows(const Tiles& tl, const TrainTiles& tc, int c0, int c1, hipStream_t st, const int* tstruct,
                     const double* dcdo) {
        (void)tl;
        if (c1 <= c0) return;
        int64_t maxnp = 1;
        for (int t = 0; t < T_; ++t) maxnp = std::max<int64_t>(maxnp, plan_.net[t].nparams);
        const int gmax = (c1 - c0 + kReduceGroup - 1) / kReduceGroup;
        if (gmax > 65535) throw std::invalid_argument("nnp_train: too many tiles in one reduction");
        const int64_t gstride = (int64_t)gmax * maxnp;
        d_grp_.reserve(8 * (size_t)(T_ * gstride));
        const unsigned pblocks = (unsigned)((maxnp + 255) / 256);
        hipLaunchKernelGGL(k_train_reduce1, dim3(pblocks, gmax, T_), dim3(256), 0, st, (const TrainPlan*)d_plan_.p, tc, c0,
                           c1, (const double*)d_partial_.p, gmax, gstride, (double*)d_grp_.p, tstruct, dcdo);
        hipLaunchKernelGGL(k_train_reduce2, dim3(pblocks, T_), dim3(256), 0, st, (const TrainPlan*)d_plan_.p, tc, c0, c1,
                           (const double*)d_grp_.p, gstride, (const int64_t*)d_goff_.p, (double*)d_grad_.p, 1);
        HIPCHK(hipGetLastError());
    }

    // A tile kernel's dynamic LDS within the device's per-workgroup limit (above 64 KiB the kernel
    // is allowed the larger allocation first).
    void check_lds(size_t lds, const void* fn) {
        if (lds_limit_ == 0) {
            int dev = 0, v = 0;
            HIPCHK(hipGetDevice(&dev));
            HIPCHK(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
            lds_limit_ = v > 0 ? (size_t)v : 65536;
        }
        if (lds > lds_limit_) throw unsupported_error("nnp_train: network too wide for the training kernels' LDS");
        if (lds > 65536) HIPCHK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    }

    // Reference forces of every training atom (3 per atom, structures in creation order).
    void set_forces(const double* f) {
        if (!f) throw std::invalid_argument("nnp_train_set_forces: null argument");
        d_fref_.reserve(8 * 3 * (size_t)ntot_);
        HIPCHK(hipMemcpy(d_fref_.p, f, 8 * 3 * (size_t)ntot_, hipMemcpyHostToDevice));
        has_forces_ = true;
    }

    void error(const double* x, int nbatch, const int* batch, int loss, double delta, double* err, double* grad) {
        check_args("nnp_train_error", x, nbatch, batch, loss, delta, err);
        pack_params(x, stream_);
        int64_t maxnp = 1;
        for (int t = 0; t < T_; ++t) maxnp = std::max<int64_t>(maxnp, plan_.net[t].nparams);
        // Tiles: with the gradient, single-structure tiles (each structure's atoms of a species
        // padded to whole tiles) when that adds few tiles and every row fits in one buffer, so one
        // forward + backward pass per tile gives the outputs and the gradient rows of the outputs,
        // which the reduction weights by dcdo; otherwise species tiles across structures, a forward
        // pass for the energies and a second pass with the loss seeds.  The tiles of the previous
        // call are reused when the batch and mode repeat.
        const int mode = grad ? (owner_.train_one_pass ? 2 : 1) : 0;
        std::vector<int> key(batch, batch + nbatch);
        key.push_back(mode);
        if (key != tiles_key_) {
            tiles_key_.clear();
            std::vector<int> bstart(nbatch + 1, 0);
            for (int b = 0; b < nbatch; ++b) bstart[b + 1] = bstart[b] + natoms_[batch[b]];
            std::vector<std::vector<int>> atom_s(T_), slot_s(T_), struct_s(T_);
            for (int b = 0; b < nbatch; ++b) {
                const int st = batch[b];
                for (int i = 0; i < natoms_[st]; ++i) {
                    const int g = (int)aoff_[st] + i, t = types_[g];
                    atom_s[t].push_back(g);
                    slot_s[t].push_back(bstart[b] + i);
                    struct_s[t].push_back(b);
                }
            }
            one_pass_ = false;
            if (grad) {
                int64_t tm = 0, th = 0;
                for (int t = 0; t < T_; ++t) {
                    tm += ((int64_t)atom_s[t].size() + 15) / 16;
                    for (size_t k = 0; k < atom_s[t].size();) {  // runs of one structure
                        size_t e = k;
                        while (e < atom_s[t].size() && struct_s[t][e] == struct_s[t][k]) ++e;
                        th += ((int64_t)(e - k) + 15) / 16;
                        k = e;
                    }
                }
                one_pass_ = mode == 2 && th <= tm + tm / 8 + 8 && th * maxnp <= owner_.train_chunk_elems;
            }
            if (one_pass_) {
                // pad each (structure, species) run to whole tiles with rows marked -1
                for (int t = 0; t < T_; ++t) {
                    std::vector<int> pa, ps, pst;
                    for (size_t k = 0; k < atom_s[t].size();) {
                        size_t e = k;
                        while (e < atom_s[t].size() && struct_s[t][e] == struct_s[t][k]) ++e;
                        for (size_t q = k; q < e; ++q) {
                            pa.push_back(atom_s[t][q]);
                            ps.push_back(slot_s[t][q]);
                            pst.push_back(struct_s[t][q]);
                        }
                        while (pa.size() % 16) {
                            pa.push_back(-1);
                            ps.push_back(0);
                            pst.push_back(struct_s[t][k]);
                        }
                        k = e;
                    }
                    atom_s[t].swap(pa);
                    slot_s[t].swap(ps);
                    struct_s[t].swap(pst);
                }
            }
            tiles_ = make_tiles(atom_s, slot_s, struct_s, stream_);
            if (one_pass_) {
                std::vector<int> ts;
                for (int t = 0; t < T_; ++t)
                    for (size_t k = 0; k < struct_s[t].size(); k += 16) ts.push_back(struct_s[t][k]);
                d_tstruct_.reserve(4 * std::max<size_t>(ts.size(), 1));
                HIPCHK(hipMemcpyAsync(d_tstruct_.p, ts.data(), 4 * ts.size(), hipMemcpyHostToDevice, stream_));
            }
            d_bstart_.reserve(4 * (size_t)(nbatch + 1));
            d_out_.reserve(8 * (size_t)std::max(bstart[nbatch], 1));
            HIPCHK(hipMemcpyAsync(d_bstart_.p, bstart.data(), 4 * (size_t)(nbatch + 1), hipMemcpyHostToDevice, stream_));
            HIPCHK(hipStreamSynchronize(stream_));  // the host vectors above go out of scope
            tiles_key_ = key;
        }
        const Tiles& tl = tiles_;
        std::vector<double> eb(nbatch);
        for (int b = 0; b < nbatch; ++b) eb[b] = eref_[batch[b]];
        d_eref_.reserve(8 * (size_t)nbatch);
        d_dcdo_.reserve(8 * (size_t)nbatch);
        d_loss_.reserve(8 * (size_t)nbatch);
        d_dE2_.reserve(8 * (size_t)nbatch);
        d_grad_.reserve(8 * (size_t)total());
        HIPCHK(hipMemcpyAsync(d_eref_.p, eb.data(), 8 * (size_t)nbatch, hipMemcpyHostToDevice, stream_));
        size_t lds = 0;
        for (int t = 0; t < T_; ++t) lds = std::max(lds, train_lds_bytes(plan_.net[t]));
        for (int md = 0; md < 4; md += (md == 1 ? 2 : 1)) {  // modes 0, 1, 3
            const void* fn = md == 0 ? (gen_ ? (const void*)k_train_tile<0, true> : (const void*)k_train_tile<0, false>)
                             : md == 1 ? (gen_ ? (const void*)k_train_tile<1, true> : (const void*)k_train_tile<1, false>)
                                       : (gen_ ? (const void*)k_train_tile<3, true> : (const void*)k_train_tile<3, false>);
            check_lds(lds, fn);
        }
        const double* Wm = (const double*)owner_.d_w64.p;
        const TrainPlan* tpp = (const TrainPlan*)d_plan_.p;
        const double* par = (const double*)d_par_.p;
        const int *ba = (const int*)d_batom_.p, *bsl = (const int*)d_bslot_.p, *bst = (const int*)d_bstruct_.p;
        const double* Gp = (const double*)d_G_.p;
        if (one_pass_) {
            d_partial_.reserve(8 * (size_t)std::max<int64_t>(tl.pb, 1));
            if (tl.tiles > 0)
                launch_tile<3>(gen_, dim3(tl.tiles), lds, stream_, tpp, Wm, par, tl.tt, 0, ba, bsl, bst, Gp, ldg_,
                               (double*)d_out_.p, nullptr, (double*)d_partial_.p, nullptr);
        } else if (tl.tiles > 0) {
            launch_tile<0>(gen_, dim3(tl.tiles), lds, stream_, tpp, Wm, par, tl.tt, 0, ba, bsl, bst, Gp, ldg_,
                           (double*)d_out_.p, nullptr, nullptr, nullptr);
        }
        hipLaunchKernelGGL(k_train_struct, dim3(nbatch), dim3(256), 0, stream_, (const double*)d_out_.p,
                           (const int*)d_bstart_.p, (const double*)d_eref_.p, loss, delta, (double*)d_dcdo_.p,
                           (double*)d_loss_.p, (double*)d_dE2_.p);
        HIPCHK(hipGetLastError());
        if (grad) {
            HIPCHK(hipMemsetAsync(d_grad_.p, 0, 8 * (size_t)total(), stream_));
            if (one_pass_) {
                reduce_rows(tl, tl.tt, 0, tl.tiles, stream_, (const int*)d_tstruct_.p, (const double*)d_dcdo_.p);
            } else {
                sum_tiles(tl, stream_, [&](const TrainTiles& tc, int c0, int c1) {
                    launch_tile<1>(gen_, dim3(c1 - c0), lds, stream_, tpp, Wm, par, tc, c0, ba, bsl, bst, Gp, ldg_,
                                   nullptr, (const double*)d_dcdo_.p, (double*)d_partial_.p, nullptr);
                });
            }
        }
        std::vector<double> lv(nbatch), e2(nbatch);
        HIPCHK(hipMemcpyAsync(lv.data(), d_loss_.p, 8 * (size_t)nbatch, hipMemcpyDeviceToHost, stream_));
        HIPCHK(hipMemcpyAsync(e2.data(), d_dE2_.p, 8 * (size_t)nbatch, hipMemcpyDeviceToHost, stream_));
        if (grad) HIPCHK(hipMemcpyAsync(grad, d_grad_.p, 8 * (size_t)total(), hipMemcpyDeviceToHost, stream_));
        HIPCHK(hipStreamSynchronize(stream_));
        err[0] = err[1] = 0.0;
        for (int b = 0; b < nbatch; ++b) {
            err[0] += lv[b];
            err[1] += e2[b];
        }
    }

    // Energy + force loss and its gradient (the training part of NNPTEF::error, nnptef.cpp:1040-1293,
    // with the exact derivative of the force term).  Per chunk of batch structures a workspace of the
    // model builds the neighbour list and descriptors; the networks at x give the atomic outputs and
    // dout/dG; the model's force kernels turn those into forces; the force residuals give gF = dL/dF,
    // the descriptor tangents V = dL/d(dout/dG), and the tile kernel the gradient.
    void error_force(const double* x, int nbatch, const int* batch, int loss, double delta, double beta, uint32_t flags,
                     double* err, double* grad) {
        check_args("nnp_train_error_force", x, nbatch, batch, loss, delta, err);
        if (!has_forces_) throw std::invalid_argument("nnp_train_error_force: reference forces not set");
        if (!(beta >= 0)) throw std::invalid_argument("nnp_train_error_force: beta must be non-negative");
        if (!ws_) ws_ = std::make_unique<Engine>(owner_, 1001);
        const hipStream_t st = ws_->stream;  // every chunk's work is ordered on this stream
        const bool all_images = flags & NNP_ALL_IMAGES;
        pack_params(x, st);
        size_t lds = 0, flds = 0;
        for (int t = 0; t < T_; ++t) {
            lds = std::max(lds, train_lds_bytes(plan_.net[t]));
            flds = std::max(flds, train_force_lds_bytes(plan_.net[t]));
        }
        check_lds(lds, gen_ ? (const void*)k_train_tile<2, true> : (const void*)k_train_tile<2, false>);
        if (grad) check_lds(flds, gen_ ? (const void*)k_train_force_tile<true> : (const void*)k_train_force_tile<false>);
        const double* Wm = (const double*)owner_.d_w64.p;
        if (grad) {
            d_grad_.reserve(8 * (size_t)total());
            HIPCHK(hipMemsetAsync(d_grad_.p, 0, 8 * (size_t)total(), st));
        }
        err[0] = err[1] = err[2] = 0.0;
        int k0 = 0;
        while (k0 < nbatch) {
            // a chunk of whole structures, at most ~128k atoms (one structure if larger)
            int k1 = k0;
            int64_t atoms = 0;
            while (k1 < nbatch && (k1 == k0 || atoms + natoms_[batch[k1]] <= (1 << 17))) atoms += natoms_[batch[k1++]];
            const int nb = k1 - k0;
            std::vector<nnp_structure> cs(nb);
            std::vector<int> bstart(nb + 1, 0);
            std::vector<int64_t> fstart(nb);
            std::vector<double> eb(nb);
            for (int b = 0; b < nb; ++b) {
                const int sb = batch[k0 + b];
                cs[b].natoms = natoms_[sb];
                cs[b].type = types_.data() + aoff_[sb];
                cs[b].pos = hpos_.data() + 3 * aoff_[sb];
                cs[b].cell = hopen_[sb] ? nullptr : hcell_.data() + 9 * (size_t)sb;
                bstart[b + 1] = bstart[b] + natoms_[sb];
                fstart[b] = aoff_[sb];
                eb[b] = eref_[sb];
            }
            // The structures do not move between calls: a chunk seen before keeps its neighbour
            // rows and descriptors in a workspace of its own (up to the model's train_cache_atoms
            // option in all, 2^20 by default);
            // otherwise the scratch workspace builds them.
            std::vector<int> key(batch + k0, batch + k1);
            key.push_back(all_images ? 1 : 0);
            Engine* Ep = nullptr;
            for (auto& c : fcache_)
                if (c.first == key) Ep = c.second.get();
            if (!Ep) {
                const bool keep = fcache_atoms_ + atoms <= owner_.train_cache_atoms;
                if (keep) {
                    fcache_.emplace_back(key, std::make_unique<Engine>(owner_, 1002));
                    fcache_atoms_ += atoms;
                    Ep = fcache_.back().second.get();
                } else {
                    Ep = ws_.get();
                }
                Engine& B = *Ep;
                B.upload(cs);
                B.build_neighbors(all_images);
                B.descriptors<double>();
                if (keep) HIPCHK(hipStreamSynchronize(B.stream));
            }
            Engine& E = *Ep;
            // the chunk's kernels, the workspace's force kernels included, run in order on st
            struct OnStream {
                Engine& e;
                hipStream_t own;
                OnStream(Engine& w, hipStream_t s) : e(w), own(w.stream) { e.stream = s; }
                ~OnStream() { e.stream = own; }
            } on_stream(E, st);
            const int n = bstart[nb];
            // species tiles over the workspace atoms (source row = slot = workspace index)
            std::vector<std::vector<int>> atom_s(T_), slot_s(T_), struct_s(T_);
            for (int b = 0; b < nb; ++b)
                for (int i = bstart[b]; i < bstart[b + 1]; ++i) {
                    const int t = types_[aoff_[batch[k0 + b]] + (i - bstart[b])];
                    atom_s[t].push_back(i);
                    slot_s[t].push_back(i);
                    struct_s[t].push_back(b);
                }
            tiles_key_.clear();  // the energy path's cached tiles are overwritten
            const Tiles tl = make_tiles(atom_s, slot_s, struct_s, st);
            d_bstart_.reserve(4 * (size_t)(nb + 1));
            d_fstart_.reserve(8 * (size_t)nb);
            d_eref_.reserve(8 * (size_t)nb);
            d_out_.reserve(8 * (size_t)n);
            d_dcdo_.reserve(8 * (size_t)nb);
            d_loss_.reserve(8 * (size_t)nb);
            d_dE2_.reserve(8 * (size_t)nb);
            d_f2_.reserve(8 * (size_t)nb);
            d_gF_.reserve(8 * 3 * (size_t)n);
            d_V_.reserve(8 * (size_t)n * ldg_);
            E.d_dEdG.reserve(8 * (size_t)n * ldg_);
            HIPCHK(hipMemcpyAsync(d_bstart_.p, bstart.data(), 4 * (size_t)(nb + 1), hipMemcpyHostToDevice, st));
            HIPCHK(hipMemcpyAsync(d_fstart_.p, fstart.data(), 8 * (size_t)nb, hipMemcpyHostToDevice, st));
            HIPCHK(hipMemcpyAsync(d_eref_.p, eb.data(), 8 * (size_t)nb, hipMemcpyHostToDevice, st));
            // outputs and dout/dG at x, then the forces by the model's force kernels
            if (tl.tiles > 0)
                launch_tile<2>(gen_, dim3(tl.tiles), lds, st,
                                   (const TrainPlan*)d_plan_.p, Wm, (const double*)d_par_.p, tl.tt, 0,
                                   (const int*)d_batom_.p, (const int*)d_bslot_.p, (const int*)d_bstruct_.p,
                                   (const double*)E.d_G.p, ldg_, (double*)d_out_.p, nullptr, nullptr,
                                   (double*)E.d_dEdG.p);
            HIPCHK(hipGetLastError());
            E.force_terms<double>();
            // reactions: the missed-reaction counter is cleared here (a cached chunk has no
            // neighbour launch in this call); a miss switches the chunk's workspace to receiver rows
            HIPCHK(hipMemsetAsync((int*)E.d_maxcnt.p + 1, 0, 4, st));
            E.assemble<double>();
            for (int attempt = 0;; ++attempt) {
                int c[4] = {0, 0, 0, 0};
                HIPCHK(hipMemcpyAsync(c, E.d_maxcnt.p, 16, hipMemcpyDeviceToHost, st));
                HIPCHK(hipStreamSynchronize(st));
                if (!E.reactions_redo(c)) break;
                if (attempt == 3) throw std::runtime_error("force assembly: receiver rows did not converge");
                E.assemble<double>();
            }
            hipLaunchKernelGGL(k_train_struct, dim3(nb), dim3(256), 0, st, (const double*)d_out_.p,
                               (const int*)d_bstart_.p, (const double*)d_eref_.p, loss, delta, (double*)d_dcdo_.p,
                               (double*)d_loss_.p, (double*)d_dE2_.p);
            hipLaunchKernelGGL(k_train_fresid, dim3(nb), dim3(256), 0, st, (const double*)E.d_F.p,
                               (const double*)d_fref_.p, (const int*)d_bstart_.p, (const int64_t*)d_fstart_.p, beta,
                               (double*)d_gF_.p, (double*)d_f2_.p);
            HIPCHK(hipGetLastError());
            if (grad) {
                HIPCHK(launch_symm_jvp(E.shape, owner_.pb.plan.generic, (const DevPlan*)E.d_plan.p,
                                       (const double*)E.d_rfe.p, (const double*)E.d_rfr.p, (const int*)E.d_type.p,
                                       (const DevStruc*)E.d_strucs.p, (const int*)E.d_astruc.p, n,
                                       (const int*)E.d_cnt.p, (const int*)E.d_nbrj.p, (const int*)E.d_meta.p,
                                       (const double4*)E.d_nbrr.p, E.maxp, E.lmaxp, (const double*)d_gF_.p,
                                       all_images ? 1 : 0, (double*)d_V_.p, ldg_, (E.moments & 1) ? 1 : 0, T_, st));
                sum_tiles(tl, st, [&](const TrainTiles& tc, int c0, int c1) {
                    const TrainPlan* tpp = (const TrainPlan*)d_plan_.p;
                    const double* par = (const double*)d_par_.p;
                    const int *ba = (const int*)d_batom_.p, *bs = (const int*)d_bstruct_.p;
                    const double *Gp = (const double*)E.d_G.p, *Vp = (const double*)d_V_.p, *dc = (const double*)d_dcdo_.p;
                    double* pp = (double*)d_partial_.p;
                    if (gen_)
                        hipLaunchKernelGGL(k_train_force_tile<true>, dim3(c1 - c0), dim3(kTrainThreads), flds, st, tpp, Wm,
                                           par, tc, c0, ba, bs, Gp, Vp, ldg_, dc, pp);
                    else
                        hipLaunchKernelGGL(k_train_force_tile<false>, dim3(c1 - c0), dim3(kTrainThreads), flds, st, tpp, Wm,
                                           par, tc, c0, ba, bs, Gp, Vp, ldg_, dc, pp);
                });
            }
            std::vector<double> lv(nb), e2(nb), f2(nb);
            HIPCHK(hipMemcpyAsync(lv.data(), d_loss_.p, 8 * (size_t)nb, hipMemcpyDeviceToHost, st));
            HIPCHK(hipMemcpyAsync(e2.data(), d_dE2_.p, 8 * (size_t)nb, hipMemcpyDeviceToHost, st));
            HIPCHK(hipMemcpyAsync(f2.data(), d_f2_.p, 8 * (size_t)nb, hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            for (int b = 0; b < nb; ++b) {
                err[0] += lv[b] + 0.5 * beta * f2[b];
                err[1] += e2[b];
                err[2] += f2[b];
            }
            k0 = k1;
        }
        if (grad) {
            HIPCHK(hipMemcpyAsync(grad, d_grad_.p, 8 * (size_t)total(), hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
        }
    }

  private:
    Engine& owner_;
    int T_ = 0, ldg_ = 0;
    int64_t ntot_ = 0, npar_dev_ = 0;
    TrainPlan plan_;
    std::vector<int64_t> goff_;
    std::vector<int> natoms_, types_;
    std::vector<int64_t> aoff_;
    std::vector<double> eref_, hpar_;
    hipStream_t stream_ = nullptr;
    DevMem d_plan_, d_goff_, d_G_, d_par_, d_batom_, d_bslot_, d_bstruct_, d_bstart_, d_eref_, d_out_, d_dcdo_, d_loss_,
        d_dE2_, d_grad_, d_partial_, d_grp_, d_fref_, d_fstart_, d_f2_, d_gF_, d_V_;
    std::vector<double> hpos_, hcell_;
    std::vector<char> hopen_;
    std::vector<int> hbatom_, hbslot_, hbstruct_;
    bool has_forces_ = false;
    size_t lds_limit_ = 0;
    bool gen_ = false;  // some hidden layer is not tanh
    // tiles of the last energy-training call (batch + mode) and whether they are single-structure
    std::vector<int> tiles_key_;
    Tiles tiles_;
    bool one_pass_ = false;
    DevMem d_tstruct_;
    std::unique_ptr<Engine> ws_;  // workspace of the force passes
    // workspaces that keep the neighbour rows and descriptors of force-training chunks
    std::vector<std::pair<std::vector<int>, std::unique_ptr<Engine>>> fcache_;
    int64_t fcache_atoms_ = 0;
};

Engine& engine_of(nnp_model* m);

}  // namespace atomnn

using namespace atomnn;

struct nnp_train {
    std::unique_ptr<TrainSet> t;
};

template <typename F>
static int train_guard(F&& f) {
    try {
        set_error("");
        return f();
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

int nnp_train_create(nnp_model* m, int nstruc, const nnp_structure* s, const double* energy, nnp_train** out) {
    return train_guard([&]() {
        if (!m || !out) throw std::invalid_argument("nnp_train_create: null argument");
        *out = new nnp_train{std::make_unique<TrainSet>(engine_of(m), nstruc, s, energy)};
        return NNP_OK;
    });
}

int64_t nnp_train_nparams(nnp_train* t, int species) {
    int64_t n = -1;
    const int rc = train_guard([&]() {
        if (!t) throw std::invalid_argument("nnp_train_nparams: null argument");
        n = species < 0 ? t->t->total() : t->t->nparams(species);
        return NNP_OK;
    });
    return rc < 0 ? rc : n;
}

int nnp_train_get_params(nnp_train* t, double* x) {
    return train_guard([&]() {
        if (!t || !x) throw std::invalid_argument("nnp_train_get_params: null argument");
        t->t->get_params(x);
        return NNP_OK;
    });
}

int nnp_train_error(nnp_train* t, const double* x, int nbatch, const int* batch, int loss, double delta, double* err,
                    double* grad) {
    return train_guard([&]() {
        if (!t) throw std::invalid_argument("nnp_train_error: null argument");
        t->t->error(x, nbatch, batch, loss, delta, err, grad);
        return NNP_OK;
    });
}

int nnp_train_set_forces(nnp_train* t, const double* forces) {
    return train_guard([&]() {
        if (!t) throw std::invalid_argument("nnp_train_set_forces: null argument");
        t->t->set_forces(forces);
        return NNP_OK;
    });
}

int nnp_train_error_force(nnp_train* t, const double* x, int nbatch, const int* batch, int loss, double delta,
                          double beta, uint32_t flags, double* err, double* grad) {
    return train_guard([&]() {
        if (!t) throw std::invalid_argument("nnp_train_error_force: null argument");
        t->t->error_force(x, nbatch, batch, loss, delta, beta, flags, err, grad);
        return NNP_OK;
    });
}

void nnp_train_destroy(nnp_train* t) { delete t; }

}  // extern "C"
