// Split-precision GEMM for the excitation densities: BF16 slices on CDNA4
// matrix cores (v_mfma_f32_16x16x32_bf16), FP32 accumulation over K-chunks,
// FP64 recombination.
//
// Every FP64 operand element x is split once into three BF16 slices
//   x = s0 + s1 + s2 + O(2^-24 |x|),   s0 = bf16(x), s1 = bf16(x - s0), s2 = bf16(x - s0 - s1)
// and the product keeps the six slice products whose weights are >= 2^-16
// relative: s0t0 + s0t1 + s1t0 + s0t2 + s1t1 + s2t0 (Ozaki-style slicing; the
// dropped terms are O(2^-24)).  The two-slice form (x = s0 + s1 + O(2^-16 |x|))
// keeps the three products of weight >= 2^-8, s0t0 + s0t1 + s1t0 (dropped
// terms O(2^-16)): half the matrix-core work and two thirds of the operand
// bytes.  Each wave accumulates KC = 256 k-values in
// FP32 registers, then adds them to FP64 accumulators, so the FP32
// accumulation error is bounded per chunk instead of over all of K.
//
// Slices are stored row-major with k contiguous and k padded to a multiple of
// 32 (zeros), so one 16-byte load gives a lane the 8 k-values that
// v_mfma_f32_16x16x32_bf16 takes per operand lane: lane l holds row l%16,
// k = 8*(l/16) .. +7.  The f32 accumulator holds D(i, j) in lane j + 16*(i/4),
// register i%4.  srcA is fed from B and srcB from A, so lane l of D holds
// row l%16 of C and stores are 16 consecutive doubles per column.
#include <array>
#include <map>
#include <mutex>
#include <tuple>

#include "qk_split.h"

namespace qk {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int KS = 32;    // k per MFMA
constexpr int KC = 256;   // k per FP32 accumulation chunk

// out_s[r * kpad + k] = slice s of src(r, k); src(r, k) = src[k + ld*r] (k contiguous);
// s2 == nullptr: two slices only
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
        if (s2) s2[t] = (__bf16)(float)z;
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
            if (s2) s2[o] = (__bf16)(float)z;
        }
    }
}

// Row blocks: src(r, k) = src[r' + ld*k + bs*b] with r = r' + nr*b (r' < nr
// contiguous within block b), 32x32 tiles through LDS, grid z = block.
__global__ void k_split3_blk(const double* __restrict__ src, long long ld, long long bs, int nr, int K, int kpad,
                             __bf16* __restrict__ s0, __bf16* __restrict__ s1, __bf16* __restrict__ s2) {
    __shared__ double t[32][33];
    const long long b = blockIdx.z;
    src += b * bs;
    int r0 = blockIdx.x * 32, k0 = blockIdx.y * 32;
    for (int kk = threadIdx.y; kk < 32; kk += blockDim.y) {
        int r = r0 + threadIdx.x, k = k0 + kk;
        t[kk][threadIdx.x] = (r < nr && k < K) ? src[r + ld * k] : 0.0;
    }
    __syncthreads();
    for (int rr = threadIdx.y; rr < 32; rr += blockDim.y) {
        int r = r0 + rr, k = k0 + threadIdx.x;
        if (r < nr && k < kpad) {
            double x = t[threadIdx.x][rr];
            __bf16 a = (__bf16)(float)x;
            double y = x - (double)(float)a;
            __bf16 c = (__bf16)(float)y;
            double z = y - (double)(float)c;
            long long o = (r + nr * b) * (long long)kpad + k;
            s0[o] = a;
            s1[o] = c;
            if (s2) s2[o] = (__bf16)(float)z;
        }
    }
}

// Workgroup tile BM x BN = 128 x 128 in one of two wave layouts (tuned):
// 4 wavefronts as 2 x 2, each 64 x 64 = 4 x 4 MFMA tiles (WR = 4), or 8
// wavefronts as 4 x 2, each 32 x 64 = 2 x 4 MFMA tiles (WR = 2: half the
// accumulators per wave, so two waves fit on a SIMD).  Every output element
// sees the same sequence of matrix-core operations and FP64 chunk additions in
// both layouts, so they give identical results.  A 32-wide k-slab of the three
// slices of the A rows and B rows is 48 KB of LDS.
//
// Tile order.  Workgroups are dispatched round-robin over the 8 XCDs, each
// with its own L2; workgroup p runs logical tile L = x q + min(x, r) + p / 8
// (x = p % 8, T = 8 q + r tiles), so every XCD works through a contiguous
// range of logical tiles.  Logical tiles run in groups of gm m-tiles (tuned) by all
// n-tiles, m fastest, so the ~32 workgroups an XCD holds at once share about
// 8 A row-blocks and 4 B row-blocks: each slab fetched from HBM feeds several
// workgroups from L2.
constexpr int BM = 128, BN = 128, NXCD = 8;
// threads of a workgroup in wave layout WR (wave rows = 16 WR)
__host__ __device__ constexpr int threads_of(int WR) { return 64 * (BM / (16 * WR)) * (BN / 64); }
static_assert(BM == 128 && BN == 128, "slab moves assume a 128 x 128 tile");
inline int grid_for(long long total) { return (int)std::min<long long>((total + 255) / 256, 65536); }

__device__ __forceinline__ void tile_of(int pid, int tm, int tn, int gm, int& mb, int& nb) {
    const int T = tm * tn, q = T / NXCD, r = T % NXCD;
    const int x = pid % NXCD, slot = pid / NXCD;
    const int L = x * q + min(x, r) + slot;
    const int per = gm * tn, g = L / per, first = g * gm;
    const int gsm = min(tm - first, gm), w = L % per;
    mb = first + w % gsm;
    nb = w / gsm;
}

// The slabs move straight from HBM into LDS (global_load_lds_dwordx4: each
// lane's 16 bytes land at the wave's LDS base + 16 * lane), with three slabs in
// flight in a ring of three LDS stages of 48 KB.
// LDS rows are 64 bytes (32 k of one slice); the 16-byte k-chunk c of row r is
// stored in slot c ^ ((r >> 1) & 3), so the matrix-core operand reads (16 rows
// x one chunk per lane group) hit distinct banks.  Rows past M or N load row
// 0: their products land only in outputs that are never stored.
// bytes of one slab stage: A and B, NSL slices each, 128 rows of 64 bytes
template <int NSL>
constexpr int slab_bytes() { return 2 * NSL * 128 * 64; }
// s_waitcnt immediate for "vector-memory count <= n" (gfx9 layout), other counters left at their maximum
#define QK_VMCNT(n) __builtin_amdgcn_s_waitcnt(((n) & 15) | (((n) >> 4) << 14) | (7 << 4) | (15 << 8))

// OutT: double, or float where the consumer needs only the FP32-level accuracy
// the slicing gives (the FP64 recombination is then rounded once on the store).
template <typename OutT, int NSL, int WR>
__global__ __launch_bounds__(threads_of(WR)) void k_gemm_bf16x3(int M, int N, int kpad, const __bf16* __restrict__ A0,
                                                         const __bf16* __restrict__ A1, const __bf16* __restrict__ A2,
                                                         const __bf16* __restrict__ B0, const __bf16* __restrict__ B1,
                                                         const __bf16* __restrict__ B2, OutT* __restrict__ C,
                                                         long long ldc, int tiles_m, int tiles_n, int gm) {
    static_assert(NSL == 2 || NSL == 3, "two or three slices");
    static_assert(WR == 4 || WR == 2, "64 or 32 wave rows");
    constexpr int NW = threads_of(WR) / 64;  // waves
    constexpr int WM = BM / (16 * WR);       // waves along m
    constexpr int SLAB_B = slab_bytes<NSL>();
    constexpr int LPW = 16 * NSL / NW;  // 1 KB slab loads per wave per slab
    static_assert(LPW * NW == 16 * NSL, "slab blocks split evenly over the waves");
    __shared__ __attribute__((aligned(1024))) char st0[SLAB_B];
    __shared__ __attribute__((aligned(1024))) char st1[SLAB_B];
    __shared__ __attribute__((aligned(1024))) char st2[SLAB_B];
    int mb, nb;
    tile_of(blockIdx.x, tiles_m, tiles_n, gm, mb, nb);
    const int m0 = mb * BM, n0 = nb * BN;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm = (wave % WM) * 16 * WR, wn = (wave / WM) * 64;
    const int l16 = lane & 15;
    // operand reads: lane reads row (.. + l16), k-chunk lane/16, stored in slot
    const int rslot = ((lane >> 4) ^ ((l16 >> 1) & 3)) * 16;
    // slab writes: wave instruction j of LPW fills the 1 KB block idx = LPW wave + j:
    // operand idx / (8 NSL) (A, B), slice (idx % (8 NSL)) / 8, rows 16 (idx % 8) .. +15;
    // lane writes row 16 (idx % 8) + lane / 4, slot lane % 4 = chunk (lane % 4) ^ ((lane >> 3) & 3)
    const int wchunk = ((lane & 3) ^ ((lane >> 3) & 3)) * 8;
    const __bf16* src[LPW];
    int ldo[LPW];
#pragma unroll
    for (int j = 0; j < LPW; ++j) {
        const int idx = wave * LPW + j;
        const int op = idx / (8 * NSL), sl = (idx % (8 * NSL)) / 8, g = idx % 8;
        const int row = g * 16 + (lane >> 2);
        int gr = (op ? n0 : m0) + row;
        if (gr >= (op ? N : M)) gr = 0;
        const __bf16* base = op ? (sl == 0 ? B0 : sl == 1 ? B1 : B2) : (sl == 0 ? A0 : sl == 1 ? A1 : A2);
        src[j] = base + (long long)gr * kpad + wchunk;
        ldo[j] = ((op * NSL + sl) * 128 + g * 16) * 64;
    }
#define QK_ISSUE(stage, k0)                                                                                  \
    do {                                                                                                     \
        _Pragma("unroll") for (int j = 0; j < LPW; ++j)                                                      \
            __builtin_amdgcn_global_load_lds((const void*)(src[j] + (k0)),                                   \
                                             (__attribute__((address_space(3))) void*)((stage) + ldo[j]), 16, 0, 0); \
    } while (0)

    d4 acc64[WR][4];
    f4 acc[WR][4];
#pragma unroll
    for (int a = 0; a < WR; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            acc64[a][b] = d4{0.0, 0.0, 0.0, 0.0};
            acc[a][b] = f4{0.f, 0.f, 0.f, 0.f};
        }

#define QK_FRAG(stage, opsl, row) (*reinterpret_cast<const bf16x8*>((stage) + ((opsl) * 128 + (row)) * 64 + rslot))
#define QK_SLAB(stage, k0)                                                                                   \
    do {                                                                                                     \
        bf16x8 fb[4][NSL];                                                                                   \
        _Pragma("unroll") for (int b = 0; b < 4; ++b)                                                        \
            _Pragma("unroll") for (int sl = 0; sl < NSL; ++sl)                                               \
                fb[b][sl] = QK_FRAG(stage, NSL + sl, wn + b * 16 + l16);                                     \
        _Pragma("unroll") for (int a = 0; a < WR; ++a) {                                                     \
            bf16x8 fa[NSL];                                                                                  \
            _Pragma("unroll") for (int sl = 0; sl < NSL; ++sl) fa[sl] = QK_FRAG(stage, sl, wm + a * 16 + l16); \
            if constexpr (NSL == 3) {                                                                        \
                _Pragma("unroll") for (int b = 0; b < 4; ++b)                                                \
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fb[b][2], fa[0], acc[a][b], 0, 0, 0); \
                _Pragma("unroll") for (int b = 0; b < 4; ++b)                                                \
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fb[b][1], fa[1], acc[a][b], 0, 0, 0); \
                _Pragma("unroll") for (int b = 0; b < 4; ++b)                                                \
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fb[b][0], fa[2], acc[a][b], 0, 0, 0); \
            }                                                                                                \
            _Pragma("unroll") for (int b = 0; b < 4; ++b)                                                    \
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fb[b][1], fa[0], acc[a][b], 0, 0, 0);   \
            _Pragma("unroll") for (int b = 0; b < 4; ++b)                                                    \
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fb[b][0], fa[1], acc[a][b], 0, 0, 0);   \
            _Pragma("unroll") for (int b = 0; b < 4; ++b)                                                    \
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fb[b][0], fa[0], acc[a][b], 0, 0, 0);   \
        }                                                                                                    \
        if (((k0) + KS) % KC == 0 || (k0) + KS >= kpad) {                                                    \
            _Pragma("unroll") for (int a = 0; a < WR; ++a)                                                   \
                _Pragma("unroll") for (int b = 0; b < 4; ++b) {                                              \
                    _Pragma("unroll") for (int r = 0; r < 4; ++r) acc64[a][b][r] += (double)acc[a][b][r];    \
                    acc[a][b] = f4{0.f, 0.f, 0.f, 0.f};                                                      \
                }                                                                                            \
        }                                                                                                    \
    } while (0)
// slab s from stage cur; stage nxt2 (read two slabs ago) receives slab s + 2.
// The issue is unconditional (past the last slab it reloads the last slab into
// a stage nobody reads again), so every path keeps exactly three slabs in flight
// and the same wait applies everywhere.
// (plain s_barrier: a __syncthreads() fence would also wait for the slabs still
// in flight; the explicit vector-memory wait covers exactly the slab needed)
#define QK_STEP(cur, nxt2, s)                                              \
    do {                                                                   \
        __builtin_amdgcn_s_waitcnt(0xC07F); /* own LDS reads of nxt2 done */ \
        __builtin_amdgcn_s_barrier();                                      \
        QK_ISSUE(nxt2, min((s) + 2, nslab - 1) * KS);                      \
        QK_VMCNT(2 * LPW);                                                 \
        __builtin_amdgcn_s_barrier();                                      \
        if ((s) < nslab) QK_SLAB(cur, (s) * KS);                           \
    } while (0)

    const int nslab = kpad / KS;
    QK_ISSUE(st0, 0);
    QK_ISSUE(st1, min(1, nslab - 1) * KS);
    for (int s = 0; s < nslab; s += 3) {
        QK_STEP(st0, st2, s);
        QK_STEP(st1, st0, s + 1);
        QK_STEP(st2, st1, s + 2);
    }
    QK_VMCNT(0);  // no slab may land in LDS after the workgroup ends
#undef QK_STEP
#undef QK_SLAB
#undef QK_FRAG
#undef QK_ISSUE

    // acc64[a][b][r] = C(row m0+wm+a*16+l16, col n0+wn+b*16+4*(lane/16)+r)
#pragma unroll
    for (int a = 0; a < WR; ++a) {
        int i = m0 + wm + a * 16 + l16;
        if (i >= M) continue;
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                int j = n0 + wn + b * 16 + 4 * (lane >> 4) + r;
                if (j < N) C[i + ldc * (long long)j] = (OutT)acc64[a][b][r];
            }
    }
}
#undef QK_VMCNT

}  // namespace

void split3(hipStream_t s, const double* src, long long ld, bool k_contiguous, int rows, int K, Split3& out) {
    QK_CHECK(out.s[0] && out.s[1] && (out.nsl == 2 || (out.nsl == 3 && out.s[2])) && out.rows == rows &&
                 out.kpad >= K && out.kpad % KS == 0,
             QK_E_ARG, "split3: bad target");
    __bf16* s2 = out.nsl == 3 ? out.s[2] : nullptr;
    if (k_contiguous) {
        hipLaunchKernelGGL(k_split3_kc, dim3(grid_for((long long)rows * out.kpad)), dim3(256), 0, s, src, ld, rows, K,
                           out.kpad, out.s[0], out.s[1], s2);
    } else {
        dim3 g((rows + 31) / 32, (out.kpad + 31) / 32);
        QK_CHECK(g.y <= 65535, QK_E_ARG, "split3: K too large");
        hipLaunchKernelGGL(k_split3_rc, g, dim3(32, 8), 0, s, src, ld, rows, K, out.kpad, out.s[0], out.s[1], s2);
    }
    QK_HIP(hipGetLastError());
}

void split3_blocks(hipStream_t s, const double* src, long long ld, long long bs, int nr, int nb, int K,
                   Split3& out) {
    QK_CHECK(out.s[0] && out.s[1] && (out.nsl == 2 || (out.nsl == 3 && out.s[2])) &&
                 (long long)out.rows == (long long)nr * nb && out.kpad >= K && out.kpad % KS == 0,
             QK_E_ARG, "split3_blocks: bad target");
    dim3 g((nr + 31) / 32, (out.kpad + 31) / 32, nb);
    QK_CHECK(g.y <= 65535 && nb <= 65535, QK_E_ARG, "split3_blocks: K or block count too large");
    hipLaunchKernelGGL(k_split3_blk, g, dim3(32, 8), 0, s, src, ld, bs, nr, K, out.kpad, out.s[0], out.s[1],
                       out.nsl == 3 ? out.s[2] : nullptr);
    QK_HIP(hipGetLastError());
}

namespace {
// Autotuned tile order and wave layout, per (M, N, kpad, output type, slices):
// candidate c has group size GM_CAND[c % NGM] and wave layout WR 4 (c < NGM)
// or 2.  The first call of a shape with at least 1e10 flop times each
// candidate once (HIP events on the caller's stream); every candidate writes
// the same C.
constexpr int NGM = 4, NCAND = 2 * NGM, CAND_DEFAULT = 2;  // group 8, 64-row waves
const int GM_CAND[NGM] = {2, 4, 8, 16};
std::mutex g_mu;
std::map<std::tuple<int, int, int, int, int>, std::pair<int, std::array<float, NCAND>>> g_gm;

template <typename OutT, int NSL, int WR>
void launch_k(hipStream_t s, int gm, int M, int N, const Split3& A, const Split3& B, OutT* C, long long ldc, int tm,
              int tn) {
    hipLaunchKernelGGL((k_gemm_bf16x3<OutT, NSL, WR>), dim3(tm * tn), dim3(threads_of(WR)), 0, s, M, N, A.kpad,
                       A.s[0], A.s[1], NSL == 3 ? A.s[2] : (const __bf16*)nullptr, B.s[0], B.s[1],
                       NSL == 3 ? B.s[2] : (const __bf16*)nullptr, C, ldc, tm, tn, gm);
}

template <typename OutT>
void launch_gm(hipStream_t s, int cand, int M, int N, const Split3& A, const Split3& B, OutT* C, long long ldc, int tm,
               int tn) {
    const int gm = GM_CAND[cand % NGM];
    const bool wide = cand < NGM;
    if (A.nsl == 3) {
        if (wide)
            launch_k<OutT, 3, 4>(s, gm, M, N, A, B, C, ldc, tm, tn);
        else
            launch_k<OutT, 3, 2>(s, gm, M, N, A, B, C, ldc, tm, tn);
    } else {
        if (wide)
            launch_k<OutT, 2, 4>(s, gm, M, N, A, B, C, ldc, tm, tn);
        else
            launch_k<OutT, 2, 2>(s, gm, M, N, A, B, C, ldc, tm, tn);
    }
    QK_HIP(hipGetLastError());
}

template <typename OutT>
int gemm_bf16x3_t(hipStream_t s, int M, int N, const Split3& A, const Split3& B, OutT* C, long long ldc) {
    QK_CHECK(A.kpad == B.kpad && A.rows == M && B.rows == N, QK_E_ARG, "gemm_bf16x3: operand shapes");
    QK_CHECK(A.nsl == B.nsl && (A.nsl == 2 || A.nsl == 3), QK_E_ARG, "gemm_bf16x3: slice counts");
    int tm = (M + BM - 1) / BM, tn = (N + BN - 1) / BN;
    QK_CHECK((long long)tm * tn < (1LL << 31), QK_E_ARG, "gemm_bf16x3: grid too large");
    const double flop = 2.0 * (A.nsl == 3 ? 6 : 3) * M * (double)N * A.kpad;  // slice products
    if (!autotune_on() || flop < 1e10) {
        launch_gm(s, CAND_DEFAULT, M, N, A, B, C, ldc, tm, tn);
        return 1;
    }
    std::lock_guard<std::mutex> lock(g_mu);
    auto key = std::make_tuple(M, N, A.kpad, (int)sizeof(OutT), A.nsl);
    auto it = g_gm.find(key);
    if (it != g_gm.end()) {
        launch_gm(s, it->second.first, M, N, A, B, C, ldc, tm, tn);
        return 1;
    }
    if (stream_capturing(s)) {  // no timing inside a graph capture
        launch_gm(s, CAND_DEFAULT, M, N, A, B, C, ldc, tm, tn);
        return 1;
    }
    hipEvent_t ev[2] = {nullptr, nullptr};
    QK_HIP(hipEventCreate(&ev[0]));
    QK_HIP(hipEventCreate(&ev[1]));
    std::array<float, NCAND> ms{};
    int best = CAND_DEFAULT;
    float tbest = 1e30f;
    for (int c = 0; c < NCAND; ++c) {
        (void)hipEventRecord(ev[0], s);
        launch_gm(s, c, M, N, A, B, C, ldc, tm, tn);
        (void)hipEventRecord(ev[1], s);
        float t = 1e30f;
        if (hipEventSynchronize(ev[1]) != hipSuccess || hipEventElapsedTime(&t, ev[0], ev[1]) != hipSuccess) t = 1e30f;
        ms[c] = t;
        if (t < tbest) {
            tbest = t;
            best = c;
        }
    }
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    g_gm.emplace(key, std::make_pair(best, ms));
    return NCAND;
}
}  // namespace

int gemm_bf16x3(hipStream_t s, int M, int N, const Split3& A, const Split3& B, double* C, long long ldc) {
    return gemm_bf16x3_t(s, M, N, A, B, C, ldc);
}

int gemm_bf16x3(hipStream_t s, int M, int N, const Split3& A, const Split3& B, float* C, long long ldc) {
    return gemm_bf16x3_t(s, M, N, A, B, C, ldc);
}

void autotune_reset_bf16() {
    std::lock_guard<std::mutex> lock(g_mu);
    g_gm.clear();
}

std::string autotune_report_bf16() {
    std::lock_guard<std::mutex> lock(g_mu);
    std::string r;
    for (const auto& kv : g_gm) {
        char buf[256];
        const int best = kv.second.first;
        std::snprintf(buf, sizeof buf,
                      "gemm_bf16x3 M=%d N=%d kpad=%d out=f%d slices=%d -> group %d, %d-row waves (ms:",
                      std::get<0>(kv.first), std::get<1>(kv.first), std::get<2>(kv.first), 8 * std::get<3>(kv.first),
                      std::get<4>(kv.first), GM_CAND[best % NGM], best < NGM ? 64 : 32);
        r += buf;
        for (int c = 0; c < NCAND; ++c) {
            std::snprintf(buf, sizeof buf, " g%d/w%d %.3f", GM_CAND[c % NGM], c < NGM ? 64 : 32, kv.second.second[c]);
            r += buf;
        }
        r += ")\n";
    }
    return r;
}

}  // namespace qk
