// FP64 GEMM on CDNA4 matrix cores: v_mfma_f64_16x16x4_f64, LDS-staged tiles.
//
// Used for every dense FP64 contraction of the restricted GW path: the
// AO->MO quarter transforms (AOtoMO/AOtoMO_ERI_RHF.f90:31-51) and the
// excitation densities rho = M * W (GW/RGW_excitation_density.f90:48-50).
//
// Tile: workgroup of WAVES_M x WAVES_N wavefronts (64 lanes each); every wave
// owns BLK_M x BLK_N 16x16 accumulators.  K advances in steps of KT = 16:
// the next K-slab is fetched into registers while the current slab in LDS is
// consumed, so global latency overlaps the MFMA chain.
//
// MFMA operand roles are swapped (srcA = B^T fragment, srcB = A^T fragment) so
// that lane l of an accumulator holds row l%16 of C (in 4 columns): a
// wavefront's store then writes 16 consecutive doubles of each column, i.e.
// 128-byte segments of the column-major output.
#include "qk_common.h"

namespace qk {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int KT = 16;

template <int WAVES_M, int WAVES_N, int BLK_M, int BLK_N, bool TA, bool TB>
__device__ __forceinline__ void dgemm_tile(int tile, int M, int N, int K, double alpha, const double* __restrict__ A,
                                           long long lda, const double* __restrict__ B, long long ldb, double beta,
                                           double* __restrict__ C, long long ldc, int tiles_m) {
    constexpr int NT = 64 * WAVES_M * WAVES_N;
    constexpr int BM = WAVES_M * BLK_M * 16;
    constexpr int BN = WAVES_N * BLK_N * 16;
    constexpr int PADM = BM + 2;  // +2 doubles: breaks the power-of-two LDS stride for transposed stores
    constexpr int PADN = BN + 2;
    constexpr int LA = BM * KT / NT;  // A elements loaded per thread per slab
    constexpr int LB = BN * KT / NT;
    static_assert(BM * KT % NT == 0 && BN * KT % NT == 0, "tile/thread mismatch");

    __shared__ double As[KT * PADM];
    __shared__ double Bs[KT * PADN];

    const int m0 = (tile % tiles_m) * BM;
    const int n0 = (tile / tiles_m) * BN;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm = (wave % WAVES_M) * BLK_M * 16;
    const int wn = (wave / WAVES_M) * BLK_N * 16;
    const int l16 = lane & 15;
    const int lk = lane >> 4;

    double ra[LA], rb[LB];

    auto fetch = [&](int k0) {
#pragma unroll
        for (int s = 0; s < LA; ++s) {
            int idx = tid + s * NT;
            int i, k;
            if (!TA) { i = idx % BM; k = idx / BM; } else { k = idx % KT; i = idx / KT; }
            int gi = m0 + i, gk = k0 + k;
            double v = 0.0;
            if (gi < M && gk < K) v = TA ? A[gk + lda * gi] : A[gi + lda * gk];
            ra[s] = v;
        }
#pragma unroll
        for (int s = 0; s < LB; ++s) {
            int idx = tid + s * NT;
            int j, k;
            if (!TB) { k = idx % KT; j = idx / KT; } else { j = idx % BN; k = idx / BN; }
            int gj = n0 + j, gk = k0 + k;
            double v = 0.0;
            if (gj < N && gk < K) v = TB ? B[gj + ldb * gk] : B[gk + ldb * gj];
            rb[s] = v;
        }
    };
    auto stash = [&]() {
#pragma unroll
        for (int s = 0; s < LA; ++s) {
            int idx = tid + s * NT;
            int i, k;
            if (!TA) { i = idx % BM; k = idx / BM; } else { k = idx % KT; i = idx / KT; }
            As[k * PADM + i] = ra[s];
        }
#pragma unroll
        for (int s = 0; s < LB; ++s) {
            int idx = tid + s * NT;
            int j, k;
            if (!TB) { k = idx % KT; j = idx / KT; } else { j = idx % BN; k = idx / BN; }
            Bs[k * PADN + j] = rb[s];
        }
    };

    d4 acc[BLK_M][BLK_N];
#pragma unroll
    for (int a = 0; a < BLK_M; ++a)
#pragma unroll
        for (int b = 0; b < BLK_N; ++b) acc[a][b] = d4{0.0, 0.0, 0.0, 0.0};

    fetch(0);
    for (int k0 = 0; k0 < K; k0 += KT) {
        __syncthreads();
        stash();
        __syncthreads();
        if (k0 + KT < K) fetch(k0 + KT);
#pragma unroll
        for (int kk = 0; kk < KT; kk += 4) {
            double fa[BLK_M], fb[BLK_N];
#pragma unroll
            for (int a = 0; a < BLK_M; ++a) fa[a] = As[(kk + lk) * PADM + wm + a * 16 + l16];
#pragma unroll
            for (int b = 0; b < BLK_N; ++b) fb[b] = Bs[(kk + lk) * PADN + wn + b * 16 + l16];
#pragma unroll
            for (int a = 0; a < BLK_M; ++a)
#pragma unroll
                for (int b = 0; b < BLK_N; ++b)
                    acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(fb[b], fa[a], acc[a][b], 0, 0, 0);
        }
    }

    // v_mfma_f64_16x16x4f64 leaves D(i', j') in lane j' + 16 (i' % 4), register i' / 4.
    // With srcA = B^T and srcB = A^T, D(i', j') = C(row j', col i'): acc[a][b][r]
    // at this lane is C(m0+wm+a*16+l16, n0+wn+b*16+lk+4r).
#pragma unroll
    for (int a = 0; a < BLK_M; ++a) {
        int gi = m0 + wm + a * 16 + l16;
        if (gi >= M) continue;
#pragma unroll
        for (int b = 0; b < BLK_N; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                int gj = n0 + wn + b * 16 + lk + 4 * r;
                if (gj < N) {
                    double* c = C + gi + ldc * (long long)gj;
                    double v = alpha * acc[a][b][r];
                    if (beta != 0.0) v += beta * *c;
                    *c = v;
                }
            }
    }
}

template <int WAVES_M, int WAVES_N, int BLK_M, int BLK_N, bool TA, bool TB>
__global__ __launch_bounds__(64 * WAVES_M * WAVES_N, 2) void dgemm_mfma(
    int M, int N, int K, double alpha, const double* __restrict__ A, long long lda, const double* __restrict__ B,
    long long ldb, double beta, double* __restrict__ C, long long ldc, Batch bt, int tiles_m) {
    if (bt.upper) {
        constexpr int BM = WAVES_M * BLK_M * 16, BN = WAVES_N * BLK_N * 16;
        if ((int)(blockIdx.x % tiles_m) * BM >= (int)(blockIdx.x / tiles_m) * BN + BN) return;
    }
    const int z = blockIdx.z;
    const long long z0 = z % bt.zdiv, z1 = z / bt.zdiv;
    dgemm_tile<WAVES_M, WAVES_N, BLK_M, BLK_N, TA, TB>(blockIdx.x, M, N, K, alpha, A + z0 * bt.sa0 + z1 * bt.sa1, lda,
                                                       B + z0 * bt.sb0 + z1 * bt.sb1, ldb, beta,
                                                       C + z0 * bt.sc0 + z1 * bt.sc1, ldc, tiles_m);
}

// Grouped form: problem g owns tiles [tile0[g], tile0[g+1]); every workgroup
// finds its problem by binary search, so problems of any shape share one launch.
template <int WAVES_M, int WAVES_N, int BLK_M, int BLK_N, bool TA, bool TB>
__global__ __launch_bounds__(64 * WAVES_M * WAVES_N, 2) void dgemm_mfma_grouped(const GemmProblem* __restrict__ pr,
                                                                                 const int* __restrict__ tile0,
                                                                                 int ngroup) {
    int t = blockIdx.x;
    int lo = 0, hi = ngroup - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (tile0[mid] <= t) lo = mid; else hi = mid - 1;
    }
    const GemmProblem& q = pr[lo];
    constexpr int BM = WAVES_M * BLK_M * 16;
    int tiles_m = (q.M + BM - 1) / BM;
    dgemm_tile<WAVES_M, WAVES_N, BLK_M, BLK_N, TA, TB>(t - tile0[lo], q.M, q.N, q.K, q.alpha, q.A, q.lda, q.B, q.ldb,
                                                       q.beta, q.C, q.ldc, tiles_m);
}

template <int WM, int WN, int BMK, int BNK, bool TA, bool TB>
void launch(hipStream_t s, int M, int N, int K, double alpha, const double* A, long long lda, const double* B,
            long long ldb, double beta, double* C, long long ldc, const Batch& bt) {
    constexpr int BM = WM * BMK * 16, BN = WN * BNK * 16;
    int tm = (M + BM - 1) / BM, tn = (N + BN - 1) / BN;
    QK_CHECK((long long)tm * tn < (1LL << 31), QK_E_ARG, "dgemm: grid too large");
    QK_CHECK(bt.count <= 65535, QK_E_ARG, "dgemm: too many batches");
    dim3 grid(tm * tn, 1, bt.count);
    hipLaunchKernelGGL((dgemm_mfma<WM, WN, BMK, BNK, TA, TB>), grid, dim3(64 * WM * WN), 0, s, M, N, K, alpha, A, lda,
                       B, ldb, beta, C, ldc, bt, tm);
    QK_HIP(hipGetLastError());
}

template <bool TA, bool TB>
void dispatch(hipStream_t s, int M, int N, int K, double alpha, const double* A, long long lda, const double* B,
              long long ldb, double beta, double* C, long long ldc, const Batch& bt) {
    if (N <= 32)
        launch<4, 1, 4, 2, TA, TB>(s, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, bt);  // 256 x 32
    else if (N <= 64 || M <= 64)
        launch<2, 2, 2, 2, TA, TB>(s, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, bt);  // 64 x 64
    else
        launch<2, 2, 4, 4, TA, TB>(s, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, bt);  // 128 x 128
}

template <int WM, int WN, int BMK, int BNK, bool TA, bool TB>
void launch_grouped(hipStream_t s, const GemmProblem* dprob, const int* dtile0, int ngroup, int ntiles) {
    hipLaunchKernelGGL((dgemm_mfma_grouped<WM, WN, BMK, BNK, TA, TB>), dim3(ntiles), dim3(64 * WM * WN), 0, s, dprob,
                       dtile0, ngroup);
    QK_HIP(hipGetLastError());
}

}  // namespace

int dgemm_grouped_tile_rows(int cfg) { return cfg == 0 ? 64 : 128; }

void dgemm_grouped_tt(hipStream_t s, int cfg, const GemmProblem* dprob, const int* dtile0, int ngroup, int ntiles) {
    if (ntiles <= 0) return;
    if (cfg == 0)
        launch_grouped<2, 2, 2, 2, true, true>(s, dprob, dtile0, ngroup, ntiles);
    else
        launch_grouped<2, 2, 4, 4, true, true>(s, dprob, dtile0, ngroup, ntiles);
}

void dgemm(hipStream_t s, bool transA, bool transB, int M, int N, int K, double alpha, const double* A, long long lda,
           const double* B, long long ldb, double beta, double* C, long long ldc, const Batch& bt) {
    if (M <= 0 || N <= 0 || bt.count <= 0) return;
    if (transA) {
        if (transB) dispatch<true, true>(s, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, bt);
        else dispatch<true, false>(s, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, bt);
    } else {
        if (transB) dispatch<false, true>(s, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, bt);
        else dispatch<false, false>(s, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, bt);
    }
}

}  // namespace qk
