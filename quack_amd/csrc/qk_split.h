// Split-precision operands: two or three BF16 slices of an FP64 matrix,
// row-major with k contiguous and k padded to a multiple of 32
// (qk_gemm_bf16x3.hip).
#pragma once
#include <algorithm>

#include "qk_common.h"

namespace qk {

struct Split3 {
    __bf16* s[3] = {nullptr, nullptr, nullptr};
    int rows = 0;
    int kpad = 0;
    int nsl = 3;  // slices held: 3 (six slice products kept) or 2 (three kept)
};

inline int pad32(int k) { return (k + 31) / 32 * 32; }

// out.s[t][r * kpad + k] = slice t of src(r, k), t < out.nsl;
// src(r, k) = src[k + ld*r] if k_contiguous else src[r + ld*k]; k >= K gives 0.
void split3(hipStream_t s, const double* src, long long ld, bool k_contiguous, int rows, int K, Split3& out);

// Row blocks of a strided array: out row r = r' + nr*b holds src(r', k) =
// src[r' + ld*k + bs*b] (r' < nr, b < nb), the slices as split3's.
void split3_blocks(hipStream_t s, const double* src, long long ld, long long bs, int nr, int nb, int K,
                   Split3& out);

// C[i + ldc*j] = sum_k A(i, k) B(j, k), FP64 result from BF16 slice products:
// with three slices per operand the six products of weight >= 2^-16, with two
// the three of weight >= 2^-8 (A.nsl == B.nsl).
// Returns the number of kernel launches it issued (more than one while tuning).
int gemm_bf16x3(hipStream_t s, int M, int N, const Split3& A, const Split3& B, double* C, long long ldc);
// The same product rounded to FP32 on the store (for consumers that need only
// the FP32-level accuracy of the slicing).
int gemm_bf16x3(hipStream_t s, int M, int N, const Split3& A, const Split3& B, float* C, long long ldc);

}  // namespace qk
