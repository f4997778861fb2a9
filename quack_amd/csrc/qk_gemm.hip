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
#include <rocblas/rocblas.h>

#include <atomic>
#include <map>
#include <mutex>
#include <tuple>

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

// Tile configurations: 0: 256 x 32, 1: 64 x 64, 2: 128 x 64, 3: 128 x 128;
// 4: the library's rocblas_dgemm (only where lib_ok allows it).
constexpr int NOWN = 4, CFG_LIB = 4, NCFG = 5;
const char* const CFG_NAME[NCFG] = {"256x32", "64x64", "128x64", "128x128", "rocblas"};

// Library FP64 GEMM as a tuning candidate (qk_set_library_gemm, off by
// default): plain products only (one matrix, all tiles), 32-bit indexable,
// at least 512 in every dimension, and never inside a graph capture.  It sums
// in its own order, so with it on, results can differ from the own kernels'
// in the last bits.
std::atomic<int> g_libgemm{0};
struct ThreadBlas {
    rocblas_handle h = nullptr;
    ~ThreadBlas() {
        if (h) (void)rocblas_destroy_handle(h);
    }
};
thread_local ThreadBlas t_blas;

bool lib_ok(hipStream_t s, int M, int N, int K, long long lda, long long ldb, long long ldc, const Batch& bt) {
    const long long lim = (1LL << 31) - 1;
    return g_libgemm.load() && bt.count == 1 && !bt.upper && M >= 512 && N >= 512 && K >= 512 &&
           (long long)M * N < lim && (long long)M * K < lim && (long long)K * N < lim && lda < lim && ldb < lim &&
           ldc < lim && !stream_capturing(s);
}

template <bool TA, bool TB>
void launch_lib(hipStream_t s, int M, int N, int K, double alpha, const double* A, long long lda, const double* B,
                long long ldb, double beta, double* C, long long ldc) {
    if (!t_blas.h) {
        rocblas_status st = rocblas_create_handle(&t_blas.h);
        QK_CHECK(st == rocblas_status_success, QK_E_HIP, "rocblas_create_handle: status " + std::to_string((int)st));
    }
    rocblas_status st = rocblas_set_stream(t_blas.h, s);
    QK_CHECK(st == rocblas_status_success, QK_E_HIP, "rocblas_set_stream: status " + std::to_string((int)st));
    st = rocblas_dgemm(t_blas.h, TA ? rocblas_operation_transpose : rocblas_operation_none,
                       TB ? rocblas_operation_transpose : rocblas_operation_none, M, N, K, &alpha, A, (int)lda, B,
                       (int)ldb, &beta, C, (int)ldc);
    QK_CHECK(st == rocblas_status_success, QK_E_HIP, "rocblas_dgemm: status " + std::to_string((int)st));
}

template <bool TA, bool TB>
void launch_cfg(int cfg, hipStream_t s, int M, int N, int K, double alpha, const double* A, long long lda,
                const double* B, long long ldb, double beta, double* C, long long ldc, const Batch& bt) {
    if (cfg == CFG_LIB) {
        if (lib_ok(s, M, N, K, lda, ldb, ldc, bt)) {
            launch_lib<TA, TB>(s, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
            return;
        }
        cfg = 3;  // not allowed here (switched off meanwhile, or capturing): the largest own tile
    }
    switch (cfg) {
        case 0: launch<4, 1, 4, 2, TA, TB>(s, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, bt); break;
        case 1: launch<2, 2, 2, 2, TA, TB>(s, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, bt); break;
        case 2: launch<2, 2, 4, 2, TA, TB>(s, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, bt); break;
        default: launch<2, 2, 4, 4, TA, TB>(s, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, bt); break;
    }
}

// fixed shape rule, used when tuning is off or not worth it
int default_cfg(int M, int N) {
    if (N <= 32) return 0;
    if (N <= 64 || M <= 64) return 1;
    return 3;
}

// Autotuner: the first call of a (TA, TB, M, N, K, batch, upper) shape with
// beta = 0 and at least 1e9 flop times every configuration once with HIP
// events on the caller's stream and keeps the fastest for the process.  All
// configurations add the k-terms of an element in the same order, so the
// tuning runs leave the same C and later calls give identical results.
struct TuneKey {
    int ta, tb, M, N, K, count, upper, lib;
    bool operator<(const TuneKey& o) const {
        return std::tie(ta, tb, M, N, K, count, upper, lib) <
               std::tie(o.ta, o.tb, o.M, o.N, o.K, o.count, o.upper, o.lib);
    }
};
struct TuneEntry {
    int cfg;
    float ms[NCFG];
};
std::mutex g_tune_mu;
std::map<TuneKey, TuneEntry> g_tune;
std::atomic<int> g_autotune{1};

template <bool TA, bool TB>
int dispatch(hipStream_t s, int M, int N, int K, double alpha, const double* A, long long lda, const double* B,
              long long ldb, double beta, double* C, long long ldc, const Batch& bt) {
    const double flop = 2.0 * M * N * (double)K * bt.count;
    if (!g_autotune.load() || beta != 0.0 || flop < 1e9) {
        launch_cfg<TA, TB>(default_cfg(M, N), s, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, bt);
        return 1;
    }
    const bool lib = lib_ok(s, M, N, K, lda, ldb, ldc, bt);
    TuneKey key{TA, TB, M, N, K, bt.count, bt.upper, lib ? 1 : 0};
    std::lock_guard<std::mutex> lock(g_tune_mu);
    auto it = g_tune.find(key);
    if (it != g_tune.end()) {
        launch_cfg<TA, TB>(it->second.cfg, s, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, bt);
        return 1;
    }
    if (stream_capturing(s)) {  // no timing inside a graph capture: an untuned shape keeps its default
        launch_cfg<TA, TB>(default_cfg(M, N), s, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, bt);
        return 1;
    }
    TuneEntry e{default_cfg(M, N), {0.f, 0.f, 0.f, 0.f, 0.f}};
    hipEvent_t ev[2] = {nullptr, nullptr};
    QK_HIP(hipEventCreate(&ev[0]));
    QK_HIP(hipEventCreate(&ev[1]));
    float best = 1e30f;
    const int ncand = lib ? NCFG : NOWN;
    for (int c = 0; c < NCFG; ++c) e.ms[c] = 0.f;
    for (int c = 0; c < ncand; ++c) {
        (void)hipEventRecord(ev[0], s);
        launch_cfg<TA, TB>(c, s, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, bt);
        (void)hipEventRecord(ev[1], s);
        float ms = 1e30f;
        if (hipEventSynchronize(ev[1]) != hipSuccess || hipEventElapsedTime(&ms, ev[0], ev[1]) != hipSuccess)
            ms = 1e30f;
        e.ms[c] = ms;
        if (ms < best) {
            best = ms;
            e.cfg = c;
        }
    }
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    g_tune.emplace(key, e);
    return ncand;
}

}  // namespace

void set_autotune(bool on) { g_autotune.store(on ? 1 : 0); }
void autotune_reset() {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    g_tune.clear();
}
void set_library_gemm(bool on) { g_libgemm.store(on ? 1 : 0); }
bool autotune_on() { return g_autotune.load() != 0; }

std::string autotune_report() {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    std::string r;
    for (const auto& kv : g_tune) {
        const TuneKey& k = kv.first;
        char buf[256];
        std::snprintf(buf, sizeof buf, "dgemm %c%c M=%d N=%d K=%d batch=%d%s -> %s (ms:", k.ta ? 'T' : 'N',
                      k.tb ? 'T' : 'N', k.M, k.N, k.K, k.count, k.upper ? " upper" : "", CFG_NAME[kv.second.cfg]);
        r += buf;
        for (int c = 0; c < (k.lib ? NCFG : NOWN); ++c) {
            std::snprintf(buf, sizeof buf, " %s %.3f", CFG_NAME[c], kv.second.ms[c]);
            r += buf;
        }
        r += ")\n";
    }
    return r;
}

int dgemm(hipStream_t s, bool transA, bool transB, int M, int N, int K, double alpha, const double* A, long long lda,
          const double* B, long long ldb, double beta, double* C, long long ldc, const Batch& bt) {
    if (M <= 0 || N <= 0 || bt.count <= 0) return 0;
    if (transA) {
        if (transB) return dispatch<true, true>(s, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, bt);
        return dispatch<true, false>(s, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, bt);
    }
    if (transB) return dispatch<false, true>(s, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, bt);
    return dispatch<false, false>(s, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, bt);
}

}  // namespace qk
