// Split-precision GEMM for the excitation densities: BF16 slices on CDNA4
// matrix cores (v_mfma_f32_16x16x32_bf16), FP32 accumulation over K-chunks,
// FP64 recombination.
//
// Every FP64 operand element x is split once into three BF16 slices
//   x = s0 + s1 + s2 + O(2^-24 |x|),   s0 = bf16(x), s1 = bf16(x - s0), s2 = bf16(x - s0 - s1)
// and the product keeps the six slice products whose weights are >= 2^-16
// relative: s0t0 + s0t1 + s1t0 + s0t2 + s1t1 + s2t0 (Ozaki-style slicing; the
// dropped terms are O(2^-24)).  Each wave accumulates KC = 256 k-values in
// FP32 registers, then adds them to FP64 accumulators, so the FP32
// accumulation error is bounded per chunk instead of over all of K.
//
// Slices are stored row-major with k contiguous and k padded to a multiple of
// 32 (zeros), so one 16-byte load gives a lane the 8 k-values that
// v_mfma_f32_16x16x32_bf16 takes per operand lane: lane l holds row l%16,
// k = 8*(l/16) .. +7.  The f32 accumulator holds D(i, j) in lane j + 16*(i/4),
// register i%4.  srcA is fed from B and srcB from A, so lane l of D holds
// row l%16 of C and stores are 16 consecutive doubles per column.
#include "qk_split.h"

namespace qk {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int KS = 32;    // k per MFMA
constexpr int KC = 256;   // k per FP32 accumulation chunk

// out_s[r * kpad + k] = slice s of src(r, k); src(r, k) = src[k + ld*r] (k contiguous)
__global__ void k_split3_kc(const double* __restrict__ src, long long ld, int rows, int K, int kpad,
                            __bf16* __restrict__ s0, __bf16* __restrict__ s1, __bf16* __restrict__ s2) {
    long long total = (long long)rows * kpad;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total; t += (long long)gridDim.x * blockDim.x) {
        long long r = t / kpad;
        int k = (int)(t % kpad);
        double x = k < K ? src[k + ld * r] : 0.0;
        __bf16 a = (__bf16)(float)x;
        double y = x - (double)(float)a;
        __bf16 b = (__bf16)(float)y;
        double z = y - (double)(float)b;
        s0[t] = a;
        s1[t] = b;
        s2[t] = (__bf16)(float)z;
    }
}

// Same with src(r, k) = src[r + ld*k] (r contiguous): 32x32 tiles through LDS.
__global__ void k_split3_rc(const double* __restrict__ src, long long ld, int rows, int K, int kpad,
                            __bf16* __restrict__ s0, __bf16* __restrict__ s1, __bf16* __restrict__ s2) {
    __shared__ double t[32][33];
    int r0 = blockIdx.x * 32, k0 = blockIdx.y * 32;
    for (int kk = threadIdx.y; kk < 32; kk += blockDim.y) {
        int r = r0 + threadIdx.x, k = k0 + kk;
        t[kk][threadIdx.x] = (r < rows && k < K) ? src[r + ld * k] : 0.0;
    }
    __syncthreads();
    for (int rr = threadIdx.y; rr < 32; rr += blockDim.y) {
        int r = r0 + rr, k = k0 + threadIdx.x;
        if (r < rows && k < kpad) {
            double x = t[threadIdx.x][rr];
            __bf16 a = (__bf16)(float)x;
            double y = x - (double)(float)a;
            __bf16 b = (__bf16)(float)y;
            double z = y - (double)(float)b;
            long long o = (long long)r * kpad + k;
            s0[o] = a;
            s1[o] = b;
            s2[o] = (__bf16)(float)z;
        }
    }
}

template <int WAVES_M, int WAVES_N, int WM, int WN>
__global__ __launch_bounds__(64 * WAVES_M * WAVES_N) void k_gemm_bf16x3(
    int M, int N, int kpad, const __bf16* __restrict__ A0, const __bf16* __restrict__ A1,
    const __bf16* __restrict__ A2, const __bf16* __restrict__ B0, const __bf16* __restrict__ B1,
    const __bf16* __restrict__ B2, double* __restrict__ C, long long ldc, int tiles_m) {
    constexpr int BM = WAVES_M * WM * 16, BN = WAVES_N * WN * 16;
    const int tile = blockIdx.x;
    const int m0 = (tile % tiles_m) * BM, n0 = (tile / tiles_m) * BN;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm0 = m0 + (wave % WAVES_M) * WM * 16;
    const int wn0 = n0 + (wave / WAVES_M) * WN * 16;
    const int l16 = lane & 15, lk = (lane >> 4) * 8;

    // row pointers (clamped to row 0 when outside; their results are never stored)
    const __bf16 *pa[WM][3], *pb[WN][3];
#pragma unroll
    for (int a = 0; a < WM; ++a) {
        int i = wm0 + a * 16 + l16;
        long long off = (long long)(i < M ? i : 0) * kpad + lk;
        pa[a][0] = A0 + off; pa[a][1] = A1 + off; pa[a][2] = A2 + off;
    }
#pragma unroll
    for (int b = 0; b < WN; ++b) {
        int j = wn0 + b * 16 + l16;
        long long off = (long long)(j < N ? j : 0) * kpad + lk;
        pb[b][0] = B0 + off; pb[b][1] = B1 + off; pb[b][2] = B2 + off;
    }

    d4 acc64[WM][WN];
#pragma unroll
    for (int a = 0; a < WM; ++a)
#pragma unroll
        for (int b = 0; b < WN; ++b) acc64[a][b] = d4{0.0, 0.0, 0.0, 0.0};

    for (int kc = 0; kc < kpad; kc += KC) {
        f4 acc[WM][WN];
#pragma unroll
        for (int a = 0; a < WM; ++a)
#pragma unroll
            for (int b = 0; b < WN; ++b) acc[a][b] = f4{0.f, 0.f, 0.f, 0.f};
        const int kend = min(kpad, kc + KC);
        for (int k = kc; k < kend; k += KS) {
            bf16x8 fa[WM][3], fb[WN][3];
#pragma unroll
            for (int a = 0; a < WM; ++a)
#pragma unroll
                for (int s = 0; s < 3; ++s) fa[a][s] = *reinterpret_cast<const bf16x8*>(pa[a][s] + k);
#pragma unroll
            for (int b = 0; b < WN; ++b)
#pragma unroll
                for (int s = 0; s < 3; ++s) fb[b][s] = *reinterpret_cast<const bf16x8*>(pb[b][s] + k);
#pragma unroll
            for (int a = 0; a < WM; ++a)
#pragma unroll
                for (int b = 0; b < WN; ++b) {
                    // smallest terms first
                    f4 c = acc[a][b];
                    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fb[b][2], fa[a][0], c, 0, 0, 0);
                    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fb[b][1], fa[a][1], c, 0, 0, 0);
                    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fb[b][0], fa[a][2], c, 0, 0, 0);
                    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fb[b][1], fa[a][0], c, 0, 0, 0);
                    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fb[b][0], fa[a][1], c, 0, 0, 0);
                    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fb[b][0], fa[a][0], c, 0, 0, 0);
                    acc[a][b] = c;
                }
        }
#pragma unroll
        for (int a = 0; a < WM; ++a)
#pragma unroll
            for (int b = 0; b < WN; ++b)
#pragma unroll
                for (int r = 0; r < 4; ++r) acc64[a][b][r] += (double)acc[a][b][r];
    }

    // acc64[a][b][r] = C(row wm0 + a*16 + l16, col wn0 + b*16 + 4*(lane/16) + r)
#pragma unroll
    for (int a = 0; a < WM; ++a) {
        int i = wm0 + a * 16 + l16;
        if (i >= M) continue;
#pragma unroll
        for (int b = 0; b < WN; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                int j = wn0 + b * 16 + 4 * (lane >> 4) + r;
                if (j < N) C[i + ldc * (long long)j] = acc64[a][b][r];
            }
    }
}

inline int grid_for(long long total) { return (int)std::min<long long>((total + 255) / 256, 65536); }

}  // namespace

void split3(hipStream_t s, const double* src, long long ld, bool k_contiguous, int rows, int K, Split3& out) {
    QK_CHECK(out.s[0] && out.rows == rows && out.kpad >= K && out.kpad % KS == 0, QK_E_ARG, "split3: bad target");
    if (k_contiguous) {
        hipLaunchKernelGGL(k_split3_kc, dim3(grid_for((long long)rows * out.kpad)), dim3(256), 0, s, src, ld, rows, K,
                           out.kpad, out.s[0], out.s[1], out.s[2]);
    } else {
        dim3 g((rows + 31) / 32, (out.kpad + 31) / 32);
        QK_CHECK(g.y <= 65535, QK_E_ARG, "split3: K too large");
        hipLaunchKernelGGL(k_split3_rc, g, dim3(32, 8), 0, s, src, ld, rows, K, out.kpad, out.s[0], out.s[1],
                           out.s[2]);
    }
    QK_HIP(hipGetLastError());
}

void gemm_bf16x3(hipStream_t s, int M, int N, const Split3& A, const Split3& B, double* C, long long ldc) {
    QK_CHECK(A.kpad == B.kpad && A.rows == M && B.rows == N, QK_E_ARG, "gemm_bf16x3: operand shapes");
    constexpr int WAVES_M = 2, WAVES_N = 2, WM = 2, WN = 4;
    constexpr int BM = WAVES_M * WM * 16, BN = WAVES_N * WN * 16;
    int tm = (M + BM - 1) / BM, tn = (N + BN - 1) / BN;
    QK_CHECK((long long)tm * tn < (1LL << 31), QK_E_ARG, "gemm_bf16x3: grid too large");
    hipLaunchKernelGGL((k_gemm_bf16x3<WAVES_M, WAVES_N, WM, WN>), dim3(tm * tn), dim3(64 * WAVES_M * WAVES_N), 0, s,
                       M, N, A.kpad, A.s[0], A.s[1], A.s[2], B.s[0], B.s[1], B.s[2], C, ldc, tm);
    QK_HIP(hipGetLastError());
}

}  // namespace qk
