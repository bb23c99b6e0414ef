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
// operands are loaded before their eight MFMAs so the weight loads overlap.
template <bool FWD>
__device__ __forceinline__ d4_t tile_mm_f64(const double* A, int lda, const double* W, int ldw, int K, int jt, int li,
                                            int lk) {
    d4_t acc = {0.0, 0.0, 0.0, 0.0};
    for (int k0 = 0; k0 < K; k0 += 32) {
        double a[8], b[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int k = k0 + 4 * q + lk;
            const bool ok = k < K;
            a[q] = ok ? A[li * lda + k] : 0.0;
            b[q] = ok ? (FWD ? W[(int64_t)(jt * 16 + li) * ldw + k] : W[(int64_t)k * ldw + jt * 16 + li]) : 0.0;
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[q], b[q], acc, 0, 0, 0);
    }
    return acc;
}

}  // namespace atomnn
