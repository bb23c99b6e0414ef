// FP64 MFMA tile product shared by the network kernels (net.hip) and the training kernels
// (train.hip).
#pragma once

#include "device.hpp"

namespace atomnn {

typedef double d4_t __attribute__((ext_vector_type(4)));

// Tile product D(16 x 16, column tile jt) = A(16 x K, LDS rows of stride lda) * B(K x 16) with
// v_mfma_f64_16x16x4f64: A lane l holds A[l%16][l/16], B lane l holds B[l/16][l%16], result
// register v of lane l is D[4v + l/16][l%16].  Forward layers use B[k][j] = W[(16 jt + j) ldw + k]
// (W row-major out x in); backward layers use B[k][j] = W[k ldw + 16 jt + j].  Eight K-steps of
// operands are loaded before their eight MFMAs so the weight loads overlap.  K is a multiple of
// 16: a remainder of 16 takes four K-steps rather than eight half-empty ones.  SWZ: W holds
// element (o, k) at o ldw + (k ^ (o & 15)) (the network kernel's LDS copy of the weights).  One
// accumulator: a dependent f64 MFMA issues every 64 cycles, as fast as independent ones.
template <bool FWD, bool SWZ = false>
__device__ __forceinline__ d4_t tile_mm_f64(const double* A, int lda, const double* W, int ldw, int K, int jt, int li,
                                            int lk) {
    auto bval = [&](int k) -> double {
        if constexpr (FWD) return W[(int64_t)(jt * 16 + li) * ldw + (SWZ ? (k ^ li) : k)];
        else return W[(int64_t)k * ldw + jt * 16 + (SWZ ? (li ^ (k & 15)) : li)];
    };
    d4_t acc = {0.0, 0.0, 0.0, 0.0};
    int k0 = 0;
    for (; k0 + 32 <= K; k0 += 32) {
        double a[8], b[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int k = k0 + 4 * q + lk;
            a[q] = A[li * lda + k];
            b[q] = bval(k);
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[q], b[q], acc, 0, 0, 0);
    }
    if (k0 < K) {
        double a[4], b[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int k = k0 + 4 * q + lk;
            a[q] = A[li * lda + k];
            b[q] = bval(k);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[q], b[q], acc, 0, 0, 0);
    }
    return acc;
}

}  // namespace atomnn
