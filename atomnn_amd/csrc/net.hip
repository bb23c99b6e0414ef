// Per-species network forward and input gradient, batched over all atoms of a species so
// each layer is a GEMM on the matrix cores.
//
// Reference arithmetic (ml/nn.cpp):
//   ins = inpw .* (G + inpb)                         fp/fpbp, nn.cpp:1194, 1218
//   z_l = b_l + W_l a_{l-1};  a_l = act(z_l), d_l = act'(z_l)   (hidden layers: the potential's
//                                                    neuron, AFFPBP nn.cpp:144-766)
//   last layer linear: a_L = z_L, d_L = 1            (nn.cpp:144-148, 1134-1137)
//   out = outb + outw .* a_L                         (nn.cpp:1229-1230)
//   dodz_L = outw .* d_L;  dodz_{l-1} = (dodz_l W_l) .* d_{l-1};  dE/dG = (dodz_0 W_0) .* inpw
//                                                    DODZ::grad, nn.cpp:1620-1628
// E_i = out_i (+ E_vac of the species with NNP_VACUUM_ENERGY, nnh.cpp:129).
//
// A block of kNetWaves wavefronts handles a tile of 16 atoms.  FP64 form: v_mfma_f64_16x16x4f64, 16x16
// output tiles, K in steps of 4; A lane l holds A[l%16][l/16], B lane l holds B[l/16][l%16],
// and result register v of lane l is D[4v + l/16][l%16].  Mixed form: v_mfma_f32_16x16x32_bf16
// with the activations and weights rounded to BF16 and FP32 accumulation; tanh and the
// derivative products stay in FP32.  Weights are zero padded to multiples of 16: a padded unit
// may carry act(0) != 0 (SIGMOID), but its outgoing weights are zero, so it adds nothing in
// either direction.
#include <algorithm>

#include "device.hpp"
#include "basis.hpp"
#include "model.hpp"
#include "netmm.hpp"

namespace atomnn {

typedef float f4_t __attribute__((ext_vector_type(4)));
typedef __bf16 bf8_t __attribute__((ext_vector_type(8)));

// LDS: two activation buffers [16][lda], the derivatives of every hidden layer [16][wpad[l]] (the
// output layer is linear: derivative 1), then
// (FP64 form) the input scale inpw [nin_pad] and the tile's 16 atom indices (16 slots).  The
// activation rows are padded (lda = widest layer + 1 doubles, + 4 floats) so the 16 rows an MFMA
// operand load reads start in different LDS banks: with the unpadded stride, a multiple of 16
// elements, all 16 rows of a K-step fall in one bank.
static __host__ __device__ int net_lda(int maxw, bool fp64) { return maxw + (fp64 ? 1 : 4); }
size_t net_lds_bytes(const DevNet& n, bool fp64) {
    int maxw = n.nin_pad, dsum = 0;
    for (int l = 0; l < n.L; ++l) {
        if (n.wpad[l] > maxw) maxw = n.wpad[l];
        if (l < n.L - 1) dsum += n.wpad[l];
    }
    return (size_t)(fp64 ? 8 : 4) * (size_t)(2 * 16 * net_lda(maxw, fp64) + 16 * dsum + n.nin_pad + 16);
}

// LDS copy of a species' network description, staged with the tile's first loads: the layer loops
// then read each layer's widths and offsets from LDS instead of waiting on a chain of scalar
// loads from the plan at the start of every layer.
static_assert(sizeof(DevNet) % 8 == 0, "DevNet is copied in 8-byte words");
__device__ __forceinline__ void stage_net(const DevNet& src, DevNet& dst, int tid) {
    if (tid < (int)(sizeof(DevNet) / 8)) reinterpret_cast<int2*>(&dst)[tid] = reinterpret_cast<const int2*>(&src)[tid];
}

// Elements of the LDS copy of the weights-in-LDS form: every layer's matrix and bias.
static __host__ __device__ int net_weight_elems(const DevNet& n) { return n.pack_len; }

static __device__ __forceinline__ int net_maxw(const DevNet& n) {
    int maxw = n.nin_pad;
    for (int l = 0; l < n.L; ++l) maxw = n.wpad[l] > maxw ? n.wpad[l] : maxw;
    return maxw;
}

// Activation of the 4 result values of a lane.  GEN = false: tanh hidden layers (and the linear
// output layer) only, straight-line code whose four dependent FP64 / FP32 chains interleave;
// GEN = true: any activation kind, one switch per value.
template <bool GEN, typename T>
__device__ __forceinline__ void net_act4(int act, const T (&z)[4], T (&a)[4], T (&d)[4]) {
    if constexpr (GEN) {
#pragma unroll
        for (int v = 0; v < 4; ++v) act_eval<T>(act, z[v], a[v], d[v]);
    } else if (act == NEU_LINEAR) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            a[v] = z[v];
            d[v] = T(1);
        }
    } else {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            a[v] = tanh_m1(z[v]);
            d[v] = T(1) - a[v] * a[v];
        }
    }
}

// ------------------------------------------------------------------ FP64 (tile_mm_f64: netmm.hpp)
// Waves per 16-atom tile: they share the tile's activations in LDS and split each layer's
// 16-column output tiles between them (the layer widths, 32-256, give 2-16 tiles).
constexpr int kNetWaves = 4;
constexpr int kNetThreads = kNetWaves * kWave;

struct NetTiles {
    int tile_start[kMaxSpecies + 1];  // first 16-atom tile of each species
    int atom_start[kMaxSpecies + 1];  // first atom of each species in the species-sorted list
};

__device__ __forceinline__ int tile_species(const NetTiles& nt, int ntypes, int tile) {
    int s = 0;
    while (s + 1 < ntypes && tile >= nt.tile_start[s + 1]) ++s;
    return s;
}

// WL: the species' weight matrices are copied into LDS (all loads in flight together), so the
// layers read weights from LDS instead of waiting on global loads layer by layer.  A block takes
// tiles blockIdx.x, blockIdx.x + gridDim.x, ... (species-major, LOOP), and copies the weights
// again only when the species changes, so with a grid of a few blocks per CU the copy is spread
// over many tiles; without LOOP a block takes the one tile blockIdx.x (the small-batch form, in
// straight-line code: the weight loads stay in flight together with the descriptor gather).
template <bool WL, bool GEN, bool LOOP>
__global__ __launch_bounds__(kNetWaves * kWave) void k_net_f64(const DevPlan* __restrict__ plan, const double* __restrict__ Wb,
                                                   NetTiles nt, int ntiles, const int* __restrict__ spatoms,
                                                   const double* __restrict__ G, int ldg, int vac,
                                                   double* __restrict__ eatom, double* __restrict__ dEdG) {
    extern __shared__ __align__(16) double shd[];
    __shared__ DevNet snet;
    const int tid = threadIdx.x, wv = tid / kWave, lane = tid % kWave;
    int s_prev = -1;
    for (int tile = blockIdx.x; tile < ntiles; tile += LOOP ? gridDim.x : ntiles) {
        const int s = tile_species(nt, plan->T, tile);
        const bool reload = !LOOP || s != s_prev;
        s_prev = s;
        const DevNet& N = plan->net[s];
        if (reload) stage_net(N, snet, tid);
        const int* atoms = spatoms + nt.atom_start[s];
        const int M = nt.atom_start[s + 1] - nt.atom_start[s];
        const int r0 = (tile - nt.tile_start[s]) * 16;
        const int lda = net_lda(net_maxw(N), true);
        double* cur = shd;
        double* nxt = shd + 16 * lda;
        double* der = shd + 32 * lda;
        const int li = lane & 15, lk = lane >> 4;
        int dsum = 0;
        for (int l = 0; l < N.L - 1; ++l) dsum += N.wpad[l];
        double* sinw = der + 16 * dsum;                          // inpw, for dE/dG
        int* sat = reinterpret_cast<int*>(sinw + N.nin_pad);     // atom index of each tile row, -1 past M
        // input staging: thread (gr, gk) fills row gr, columns gk + 16 j; its atom index is loaded
        // first, so that load is in flight together with the weight copy
        static_assert(kNetThreads == 16 * 16, "input staging maps 16 threads to each of the 16 rows");
        const int gr = tid >> 4, gk = tid & 15;
        const int gat = r0 + gr < M ? atoms[r0 + gr] : -1;
        double* wl = nullptr;
        // Weights and the first 8 column groups of the descriptor row are loaded into registers
        // before any LDS store, so the gather's loads (which wait only for the atom index) overlap
        // the weight copy's instead of following them.
        constexpr int kWReg = 16;  // 16-byte weight loads per thread: n2 <= 16 x 256 under the 64 KB LDS bound
        constexpr int kGReg = 8;   // column groups of 16 gathered through registers (nin_pad <= 128)
        double2 wr[kWReg];
        const double2* src = nullptr;
        double2* dst = nullptr;
        int n2 = 0;
        if constexpr (WL) {
            wl = sinw + N.nin_pad + 16;
            // one flat 16-byte copy of [W_0 | W_1 | ... | b_0 | ...]: all loads in flight together
            src = reinterpret_cast<const double2*>(Wb + N.pack_off);
            dst = reinterpret_cast<double2*>(wl);
            n2 = N.pack_len / 2;
            // unconditional loads of clamped (in-bounds) indices: no branch, so no wait, between them
            // (the persistent form copies directly, on a species change only)
            if constexpr (!LOOP) {
#pragma unroll
                for (int j = 0; j < kWReg; ++j) wr[j] = src[min(tid + j * kNetThreads, n2 - 1)];
            }
        }
        double gq[kGReg], wq[kGReg], bq[kGReg];
        const int64_t grow = (int64_t)(gat >= 0 ? gat : 0) * ldg;  // row past M: any valid row, unused
#pragma unroll
        for (int j = 0; j < kGReg; ++j) {
            const int kc = min(gk + 16 * j, N.nin - 1);
            wq[j] = Wb[N.inpw_off + kc];
            bq[j] = Wb[N.inpb_off + kc];
            gq[j] = G[grow + kc];
        }
        if constexpr (WL) {
            // unconditional too (threads past n2 rewrite the last element with its own value), so no
            // load is sunk into a branch of its own
            if constexpr (!LOOP) {
#pragma unroll
                for (int j = 0; j < kWReg; ++j) dst[min(tid + j * kNetThreads, n2 - 1)] = wr[j];
                for (int idx = tid + kWReg * kNetThreads; idx < n2; idx += kNetThreads) dst[idx] = src[idx];
            } else if (reload) {
                for (int idx = tid; idx < n2; idx += kNetThreads) dst[idx] = src[idx];
            }
        }
        // layer l's weights and biases: LDS copy (bank-swizzled rows, as in the packed copy) or global
        auto wmat = [&](int l) -> const double* {
            if constexpr (WL) return wl + snet.pw_off[l];
            else return Wb + snet.woff[l];
        };
        auto bvec = [&](int l) -> const double* {
            if constexpr (WL) return wl + snet.pb_off[l];
            else return Wb + snet.boff[l];
        };
        // scaled input ins = inpw .* (G + inpb), zero past nin and on rows past M
#pragma unroll
        for (int j = 0; j < kGReg; ++j) {
            const int k = gk + 16 * j;
            if (k < N.nin_pad) {
                const bool in = k < N.nin;
                cur[gr * lda + k] = in && gat >= 0 ? wq[j] * (gq[j] + bq[j]) : 0.0;
                if (gr == 0) sinw[k] = in ? wq[j] : 0.0;
            }
        }
        for (int k = gk + 16 * kGReg; k < N.nin_pad; k += 16) {
            double v = 0.0, w = 0.0;
            if (k < N.nin) {
                w = Wb[N.inpw_off + k];
                if (gat >= 0) v = w * (G[(int64_t)gat * ldg + k] + Wb[N.inpb_off + k]);
            }
            cur[gr * lda + k] = v;
            if (gr == 0) sinw[k] = w;
        }
        if (gk == 0) sat[gr] = gat;
        __syncthreads();
        const DevNet& NL = snet;
        const int L1 = NL.L - 1;  // the output layer (linear, one real unit)
        int in_pad = NL.nin_pad;
        int doff = 0;
        // hidden layers
        for (int l = 0; l < L1; ++l) {
            const int out_pad = NL.wpad[l];
            const double* W = wmat(l);
            const double* b = bvec(l);
            for (int jt = wv; jt < out_pad / 16; jt += kNetWaves) {
                const d4_t acc = tile_mm_f64<true, WL>(cur, lda, W, in_pad, in_pad, jt, li, lk);
                const int col = jt * 16 + li;
                const double bc = b[col];
                double z[4], av[4], dv[4];
#pragma unroll
                for (int v = 0; v < 4; ++v) z[v] = acc[v] + bc;
                net_act4<GEN>(NL.act, z, av, dv);
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int row = 4 * v + lk;
                    nxt[row * lda + col] = av[v];
                    der[doff + row * out_pad + col] = dv[v];
                }
            }
            __syncthreads();
            double* t = cur;
            cur = nxt;
            nxt = t;
            in_pad = out_pad;
            doff += 16 * out_pad;
        }
        // The output layer has one real unit and is linear (derivative 1), so neither of its passes
        // needs the matrix cores: wave 0 forms the energies as row dot products with W_L's first row
        // (4 lanes per atom), and the other waves the first backward step, which is elementwise:
        // delta_{L-1} = (outw W_L[0][.]) .* d_{L-1}.  With no hidden layer that step is dE/dG itself.
        {
            const double* wo = wmat(L1);  // row 0: unswizzled in the LDS copy as well
            if (L1 > 0) doff -= 16 * in_pad;  // d_{L-1}
            if (wv == 0) {
                const int r = lane >> 2, q = lane & 3;
                double acc = 0.0;
                for (int k = q; k < in_pad; k += 4) acc += wo[k] * cur[r * lda + k];
                acc += __shfl_xor(acc, 1, 4);
                acc += __shfl_xor(acc, 2, 4);
                const int at = sat[r];
                if (q == 0 && at >= 0) eatom[at] = NL.outb + (bvec(L1)[0] + acc) * NL.outw + (vac ? NL.evac : 0.0);
            } else if (L1 > 0) {
                for (int r = wv - 1; r < 16; r += kNetWaves - 1)
                    for (int j = lane; j < in_pad; j += kWave)
                        nxt[r * lda + j] = NL.outw * wo[j] * der[doff + r * in_pad + j];
            } else {
                for (int r = wv - 1; r < 16; r += kNetWaves - 1) {
                    const int at = sat[r];
                    if (at >= 0)
                        for (int j = lane; j < NL.nin; j += kWave)
                            dEdG[(int64_t)at * ldg + j] = NL.outw * wo[j] * sinw[j];
                }
            }
        }
        if (L1 > 0) {
            __syncthreads();
            {
                double* t = cur;
                cur = nxt;
                nxt = t;
            }
            // backward through the hidden layers: delta_{l-1} = (delta_l W_l) .* d_{l-1}
            for (int l = L1 - 1; l >= 1; --l) {
                const int K = NL.wpad[l], J = NL.wpad[l - 1];
                const double* W = wmat(l);  // wpad[l] x wpad[l-1], row-major
                doff -= 16 * J;
                for (int jt = wv; jt < J / 16; jt += kNetWaves) {
                    const d4_t acc = tile_mm_f64<false, WL>(cur, lda, W, J, K, jt, li, lk);
                    const int col = jt * 16 + li;
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int row = 4 * v + lk;
                        nxt[row * lda + col] = acc[v] * der[doff + row * J + col];
                    }
                }
                __syncthreads();
                double* t = cur;
                cur = nxt;
                nxt = t;
            }
            // dE/dG = (delta_0 W_0) .* inpw
            {
                const int K = NL.wpad[0], J = NL.nin_pad;
                const double* W = wmat(0);
                for (int jt = wv; jt < J / 16; jt += kNetWaves) {
                    const d4_t acc = tile_mm_f64<false, WL>(cur, lda, W, J, K, jt, li, lk);
                    const int col = jt * 16 + li;
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int at = sat[4 * v + lk];
                        if (at >= 0 && col < NL.nin) dEdG[(int64_t)at * ldg + col] = acc[v] * sinw[col];
                    }
                }
            }
        }  // L1 > 0
        if constexpr (LOOP) __syncthreads();  // the next tile's staging overwrites this tile's LDS
    }
}

// ------------------------------------------------------------------ mixed (BF16 MFMA, FP32 accumulate)
// v_mfma_f32_16x16x32_bf16 (gfx950): A lane l holds A[l%16][8(l/16) .. 8(l/16)+7], B lane l
// holds B[8(l/16) .. +7][l%16], result register v of lane l is D[4(l/16) + v][l%16].
__device__ __forceinline__ __bf16 to_bf16(float x) { return (__bf16)x; }

// Split-BF16 products: x = hi + lo with hi = bf16(x), lo = bf16(x - hi) (the subtraction is exact),
// and A B ~ Ah Bh + Ah Bl + Al Bh, three BF16 MFMAs with FP32 accumulation: about 16 mantissa
// bits per operand instead of 8 (the dropped Al Bl term is 2^-16 relative).  TERMS = 1 is the
// single BF16 product.
template <int TERMS>
__device__ __forceinline__ f4_t mfma_split(const float (&a)[8], const float (&b)[8], f4_t acc) {
    bf8_t ah, bh;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        ah[q] = to_bf16(a[q]);
        bh[q] = to_bf16(b[q]);
    }
    if constexpr (TERMS == 3) {
        bf8_t al, bl;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            al[q] = to_bf16(a[q] - (float)ah[q]);
            bl[q] = to_bf16(b[q] - (float)bh[q]);
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, acc, 0, 0, 0);
    }
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, acc, 0, 0, 0);
}

template <int TERMS, bool GEN>
__global__ __launch_bounds__(kNetWaves * kWave) void k_net_mixed(const DevPlan* __restrict__ plan, const float* __restrict__ Wf,
                                                     NetTiles nt, const int* __restrict__ spatoms,
                                                     const float* __restrict__ G, int ldg, int vac,
                                                     double* __restrict__ eatom, float* __restrict__ dEdG) {
    extern __shared__ __align__(16) float shf[];
    const int s = tile_species(nt, plan->T, blockIdx.x);
    const DevNet& N = plan->net[s];
    __shared__ DevNet snet;
    stage_net(N, snet, threadIdx.x);
    const int* atoms = spatoms + nt.atom_start[s];
    const int M = nt.atom_start[s + 1] - nt.atom_start[s];
    const int tid = threadIdx.x, wv = tid / kWave, lane = tid % kWave;
    const int r0 = (blockIdx.x - nt.tile_start[s]) * 16;
    const int lda = net_lda(net_maxw(N), false);
    float* cur = shf;
    float* nxt = shf + 16 * lda;
    float* der = shf + 32 * lda;
    const int li = lane & 15, lk = lane >> 4;
    for (int idx = tid; idx < 16 * N.nin_pad; idx += kNetThreads) {
        const int r = idx / N.nin_pad, k = idx % N.nin_pad;
        const int row = r0 + r;
        float v = 0.f;
        if (row < M && k < N.nin) {
            const int at = atoms[row];
            v = Wf[N.inpw_off + k] * (G[(int64_t)at * ldg + k] + Wf[N.inpb_off + k]);
        }
        cur[r * lda + k] = v;
    }
    __syncthreads();
    const DevNet& NL = snet;
    const int L1 = NL.L - 1;  // the output layer (linear, one real unit)
    int in_pad = NL.nin_pad;
    int doff = 0;
    // hidden layers
    for (int l = 0; l < L1; ++l) {
        const int out_pad = NL.wpad[l];
        const float* W = Wf + NL.woff[l];
        const float* b = Wf + NL.boff[l];
        for (int jt = wv; jt < out_pad / 16; jt += kNetWaves) {
            f4_t acc = {0.f, 0.f, 0.f, 0.f};
            for (int k0 = 0; k0 < in_pad; k0 += 32) {
                float a[8], bv[8];
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    const int k = k0 + 8 * lk + q;
                    const bool ok = k < in_pad;
                    a[q] = ok ? cur[li * lda + k] : 0.f;
                    bv[q] = ok ? W[(int64_t)(jt * 16 + li) * in_pad + k] : 0.f;
                }
                acc = mfma_split<TERMS>(a, bv, acc);
            }
            const int col = jt * 16 + li;
            const float bc = b[col];
            float z[4], av[4], dv[4];
#pragma unroll
            for (int v = 0; v < 4; ++v) z[v] = acc[v] + bc;
            net_act4<GEN>(NL.act, z, av, dv);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = 4 * lk + v;
                nxt[row * lda + col] = av[v];
                der[doff + row * out_pad + col] = dv[v];
            }
        }
        __syncthreads();
        float* t = cur;
        cur = nxt;
        nxt = t;
        in_pad = out_pad;
        doff += 16 * out_pad;
    }
    // output layer without the matrix cores (as in the FP64 kernel): wave 0 forms the energies as
    // FP32 row dot products with W_L's first row, the other waves the elementwise first backward
    // step delta_{L-1} = (outw W_L[0][.]) .* d_{L-1} (with no hidden layer: dE/dG itself)
    {
        const float* wo = Wf + NL.woff[L1];
        if (L1 > 0) doff -= 16 * in_pad;  // d_{L-1}
        if (wv == 0) {
            const int r = lane >> 2, q = lane & 3;
            float acc = 0.f;
            for (int k = q; k < in_pad; k += 4) acc += wo[k] * cur[r * lda + k];
            acc += __shfl_xor(acc, 1, 4);
            acc += __shfl_xor(acc, 2, 4);
            if (q == 0 && r0 + r < M)
                eatom[atoms[r0 + r]] =
                    (double)NL.outb + (double)(Wf[NL.boff[L1]] + acc) * NL.outw + (vac ? NL.evac : 0.0);
        } else if (L1 > 0) {
            for (int r = wv - 1; r < 16; r += kNetWaves - 1)
                for (int j = lane; j < in_pad; j += kWave)
                    nxt[r * lda + j] = (float)NL.outw * wo[j] * der[doff + r * in_pad + j];
        } else {
            for (int r = wv - 1; r < 16; r += kNetWaves - 1) {
                if (r0 + r < M) {
                    const int64_t row = (int64_t)atoms[r0 + r] * ldg;
                    for (int j = lane; j < NL.nin; j += kWave)
                        dEdG[row + j] = (float)NL.outw * wo[j] * Wf[NL.inpw_off + j];
                }
            }
        }
    }
    if (L1 == 0) return;
    __syncthreads();
    {
        float* t = cur;
        cur = nxt;
        nxt = t;
    }
    // backward through the hidden layers, then dE/dG = (delta_0 W_0) .* inpw
    for (int l = L1 - 1; l >= 0; --l) {
        const int K = NL.wpad[l], J = (l == 0) ? NL.nin_pad : NL.wpad[l - 1];
        const float* W = Wf + NL.woff[l];
        if (l > 0) doff -= 16 * J;
        for (int jt = wv; jt < J / 16; jt += kNetWaves) {
            f4_t acc = {0.f, 0.f, 0.f, 0.f};
            for (int k0 = 0; k0 < K; k0 += 32) {
                float a[8], bv[8];
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    const int k = k0 + 8 * lk + q;
                    const bool ok = k < K;
                    a[q] = ok ? cur[li * lda + k] : 0.f;
                    bv[q] = ok ? W[(int64_t)k * J + jt * 16 + li] : 0.f;
                }
                acc = mfma_split<TERMS>(a, bv, acc);
            }
            const int col = jt * 16 + li;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int rr = 4 * lk + v;
                if (l > 0) {
                    nxt[rr * lda + col] = acc[v] * der[doff + rr * J + col];
                } else {
                    const int row = r0 + rr;
                    if (row < M && col < NL.nin) dEdG[(int64_t)atoms[row] * ldg + col] = acc[v] * Wf[NL.inpw_off + col];
                }
            }
        }
        __syncthreads();
        float* t = cur;
        cur = nxt;
        nxt = t;
    }
}

// One launch covers the 16-atom tiles of every species; the LDS size is the largest species'.
static NetTiles make_tiles(const int* sp_off, int ntypes, int& ntiles) {
    NetTiles nt;
    ntiles = 0;
    for (int s = 0; s <= kMaxSpecies; ++s) {
        nt.tile_start[s] = 0;
        nt.atom_start[s] = 0;
    }
    for (int s = 0; s < ntypes; ++s) {
        nt.tile_start[s] = ntiles;
        nt.atom_start[s] = sp_off[s];
        ntiles += (sp_off[s + 1] - sp_off[s] + 15) / 16;
    }
    for (int s = ntypes; s <= kMaxSpecies; ++s) {
        nt.tile_start[s] = ntiles;
        nt.atom_start[s] = sp_off[ntypes];
    }
    return nt;
}

int net_tiles(const int* sp_off, int ntypes) {
    int ntiles = 0;
    (void)make_tiles(sp_off, ntypes, ntiles);
    return ntiles;
}

// Whether the FP64 weights-in-LDS form fits: the weights beside the activations and the kernels'
// static copy of the network description, under 64 KiB.
bool net_wl_fits(const DevPlan& hplan) {
    size_t ldsw = 0;
    for (int s = 0; s < hplan.T; ++s)
        ldsw = std::max(ldsw, net_lds_bytes(hplan.net[s], true) + 8 * (size_t)net_weight_elems(hplan.net[s]));
    return ldsw + sizeof(DevNet) <= 65536;
}

hipError_t launch_net(bool fp64, const DevPlan* plan, const DevPlan& hplan, const void* W, const int* sp_off,
                      const int* spatoms, const void* G, int ldg, int vac, double* eatom, void* dEdG, int bf16_terms,
                      NetLaunch cfg, hipStream_t st) {
    int ntiles = 0;
    const NetTiles nt = make_tiles(sp_off, hplan.T, ntiles);
    if (ntiles == 0) return hipSuccess;
    size_t lds = 0, ldsw = 0;
    for (int s = 0; s < hplan.T; ++s) {
        lds = std::max(lds, net_lds_bytes(hplan.net[s], fp64));
        ldsw = std::max(ldsw, net_lds_bytes(hplan.net[s], fp64) + 8 * (size_t)net_weight_elems(hplan.net[s]));
    }
    // the tanh-only kernels when every species' hidden layers are tanh
    bool gen = false;
    for (int s = 0; s < hplan.T; ++s) gen = gen || hplan.net[s].act != NEU_TANH;
    const bool wl = fp64 && cfg.wl && net_wl_fits(hplan);
    const size_t bytes = wl ? ldsw : lds;
    auto kern = [&](auto k) { return (const void*)k; };
    if (fp64) {
        const int grid = wl && cfg.grid > 0 && cfg.grid < ntiles ? cfg.grid : ntiles;
        const void* f = !wl ? (gen ? kern(k_net_f64<false, true, false>) : kern(k_net_f64<false, false, false>))
                        : grid < ntiles ? (gen ? kern(k_net_f64<true, true, true>) : kern(k_net_f64<true, false, true>))
                                        : (gen ? kern(k_net_f64<true, true, false>) : kern(k_net_f64<true, false, false>));
        void* args[] = {(void*)&plan,  (void*)&W, (void*)&nt,  (void*)&ntiles, (void*)&spatoms,
                        (void*)&G,     (void*)&ldg, (void*)&vac, (void*)&eatom,  (void*)&dEdG};
        const hipError_t e = hipLaunchKernel(f, dim3(grid), dim3(kNetThreads), args, bytes, st);
        if (e != hipSuccess) return e;
        return hipGetLastError();
    }
    const void* f = bf16_terms == 1 ? (gen ? kern(k_net_mixed<1, true>) : kern(k_net_mixed<1, false>))
                                    : (gen ? kern(k_net_mixed<3, true>) : kern(k_net_mixed<3, false>));
    void* args[] = {(void*)&plan, (void*)&W, (void*)&nt, (void*)&spatoms, (void*)&G, (void*)&ldg, (void*)&vac,
                    (void*)&eatom, (void*)&dEdG};
    const hipError_t e = hipLaunchKernel(f, dim3(ntiles), dim3(kNetThreads), args, bytes, st);
    if (e != hipSuccess) return e;
    return hipGetLastError();
}

}  // namespace atomnn
