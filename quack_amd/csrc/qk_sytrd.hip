// Householder tridiagonalisation of a symmetric matrix (upper triangle),
// LAPACK dsytrd conventions (UPLO = 'U'): T = Q^T A Q with Q = H(n-1)...H(1),
// H(i) = I - tau v v^T, v(i+1:n) = 0, v(i) = 1, v(1:i-1) stored in A(1:i-1, i+1);
// D = diag(T), E = superdiagonal, TAU = tau.  The reflectors are therefore the
// ones rocSOLVER's dormtr applies for the back-transformation.
//
// Blocked as LAPACK's dsytrd/dlatrd (Golub & Van Loan, sec. 8.3; the panel
// recurrence of dlatrd): panels of nb columns from the right; per column
//   K_a  finalise the previous W column, update column c with the panel,
//        partial sums of squares of the vector to annihilate;
//   K_b1 Householder scalars (every workgroup re-reduces the norm), symmetric
//        matrix-vector product on the upper triangle in 64 x 64 tiles (row and
//        column sums of each tile to a work array), panel dots t1 = W^T v,
//        t2 = V^T v;
//   K_b2 sums the tile partials, panel corrections, scaling by tau, partial
//        dots w^T v (the alpha of the next W update is applied lazily by the
//        next K_a), stores v and the scalars;
// and after each panel the trailing update A -= V W^T + W V^T runs as two
// matrix-core GEMMs on the upper tiles.  No host synchronisation inside.
#include <algorithm>

#include "qk_kernels.h"

namespace qk {
namespace {

constexpr int TB = 256;  // threads per block
constexpr int TS = 64;   // symv tile

__device__ inline double wg_sum(double v, double* sh) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double r = 0.0;
    for (int k = 0; k < (int)(blockDim.x + 63) / 64; ++k) r += sh[k];
    __syncthreads();
    return r;  // valid in every thread
}

// scalars of a column: [0] beta, [1] tau, [2] 1/(alpha - beta) (the scale of x)
// Householder scalars of column c from the norm partials (dlarfg)
__device__ inline void householder(const double* A, int n, int c, const double* pnorm, int npart, double* sh,
                                   double& beta, double& tau, double& scale) {
    double s = 0.0;
    for (int k = threadIdx.x; k < npart; k += blockDim.x) s += pnorm[k];
    const double xx = wg_sum(s, sh);
    const double alpha = A[(c - 1) + (long long)n * c];
    if (xx == 0.0) {
        beta = alpha;
        tau = 0.0;
        scale = 0.0;
    } else {
        const double nrm = sqrt(alpha * alpha + xx);
        beta = alpha >= 0.0 ? -nrm : nrm;
        tau = (beta - alpha) / beta;
        scale = 1.0 / (alpha - beta);
    }
}

// K_a: rows r <= c.  The previous W column piw (= iw + 1) still lacks
// palpha v_prev (v_prev = A(:, c+1), its unit element A(c, c+1) stored as 1):
// rows r < c are finalised in memory; row c, read by every thread, is only
// ever used here and is finalised on the fly.  Then
// A(r, c) -= sum_k A(r, c+1+k) W(c, iw+1+k) + W(r, iw+1+k) A(c, c+1+k), k < j;
// pnorm[block] = sum over r < c-1 of the updated A(r, c)^2;  D[c] = A(c, c).
__global__ __launch_bounds__(TB) void k_col_update(double* __restrict__ A, int n, double* __restrict__ W, int ldw,
                                                   int c, int iw, int j, int piw, const double* __restrict__ pdot,
                                                   int npdot, const double* __restrict__ scal_prev,
                                                   double* __restrict__ pnorm, double* __restrict__ D) {
    __shared__ double sh[TB / 64];
    __shared__ double red[4][64];
    double palpha = 0.0;
    if (piw >= 0) {
        double s = 0.0;
        for (int k = threadIdx.x; k < npdot; k += blockDim.x) s += pdot[k];
        palpha = -0.5 * scal_prev[1] * wg_sum(s, sh);
    }
    // 64 rows per block; the four waves split the panel columns k = q, q+4, ...
    const int rr = threadIdx.x & 63, q = threadIdx.x >> 6;
    const int r = blockIdx.x * 64 + rr;
    const bool on = r <= c;
    double wfin = 0.0;  // finalised W(r, piw)
    if (on && piw >= 0) wfin = W[r + (long long)ldw * piw] + palpha * A[r + (long long)n * (c + 1)];
    double acc = 0.0;
    if (on) {
        for (int k = q; k < j; k += 4) {
            const long long col = (long long)n * (c + 1 + k);
            double wck = W[c + (long long)ldw * (iw + 1 + k)];
            double wrk = W[r + (long long)ldw * (iw + 1 + k)];
            if (k == 0 && piw >= 0) {
                wck += palpha;  // v_prev(c) = 1
                wrk = wfin;
            }
            acc += A[r + col] * wck + wrk * A[c + col];
        }
    }
    red[q][rr] = acc;
    __syncthreads();
    double sq = 0.0;
    if (q == 0 && on) {
        if (piw >= 0 && r < c) W[r + (long long)ldw * piw] = wfin;
        const double a = A[r + (long long)n * c] - (red[0][rr] + red[1][rr] + red[2][rr] + red[3][rr]);
        A[r + (long long)n * c] = a;
        if (r < c - 1) sq = a * a;
        if (r == c) D[c] = a;
    }
    const double t = wg_sum(sq, sh);
    if (threadIdx.x == 0) pnorm[blockIdx.x] = t;
}

// K_b1: blocks [0, ntiles) are upper tiles (I <= J) of A(0:L, 0:L), L = c; the
// next 2j blocks are the panel dots.  v(r) = A(r, c) scale (r < c-1), v(c-1) = 1.
// part[tile][0..63] = row sums for block I, part[tile][64..127] = column sums
// for block J (strictly upper part; zero on the diagonal tile's lower half).
__global__ __launch_bounds__(TB) void k_symv_tiles(const double* __restrict__ A, int n, const double* __restrict__ W,
                                                   int ldw, int c, int iw, int j, const double* __restrict__ pnorm,
                                                   int npart, int nt, int ntiles, double* __restrict__ part,
                                                   double* __restrict__ t12, double* __restrict__ scal) {
    __shared__ double sh[TB / 64];
    __shared__ double vI[TS], vJ[TS];
    __shared__ double rs[4][TS], cs[4][TS];
    double beta, tau, scale;
    householder(A, n, c, pnorm, npart, sh, beta, tau, scale);
    const int L = c;
    auto vof = [&](int r) -> double {
        if (r >= L) return 0.0;
        if (r == L - 1) return 1.0;
        return A[r + (long long)n * c] * scale;
    };
    const int b = blockIdx.x;
    if (b < ntiles) {
        if (b == 0 && threadIdx.x == 0) {
            scal[0] = beta;
            scal[1] = tau;
            scal[2] = scale;
        }
        // tile index -> (I, J), I <= J, column-major over the upper triangle
        int J = (int)((sqrt(8.0 * b + 1.0) - 1.0) * 0.5);
        while ((J + 1) * (J + 2) / 2 <= b) ++J;
        while (J * (J + 1) / 2 > b) --J;
        const int I = b - J * (J + 1) / 2;
        const int r0 = I * TS, c0 = J * TS;
        if (threadIdx.x < TS) vI[threadIdx.x] = vof(r0 + threadIdx.x);
        else if (threadIdx.x < 2 * TS) vJ[threadIdx.x - TS] = vof(c0 + threadIdx.x - TS);
        __syncthreads();
        // thread: row rr = tid % 64 of the tile, column quarter q = tid / 64 (16 columns)
        const int rr = threadIdx.x & 63, q = threadIdx.x >> 6;
        const int r = r0 + rr;
        double rsum = 0.0;
        for (int cc = q * 16; cc < q * 16 + 16; ++cc) {
            const int col = c0 + cc;
            double a = 0.0;
            bool use = r < L && col < L && (I < J || rr <= cc);
            if (use) a = A[r + (long long)n * col];
            rsum += a * vJ[cc];
            // column sums: strictly upper only (the diagonal entry is counted once, in the row sum)
            double cpart = (use && (I < J || rr < cc)) ? a * vI[rr] : 0.0;
            // reduce cpart over the 64 rows of this wave
            for (int off = 32; off > 0; off >>= 1) cpart += __shfl_down(cpart, off, 64);
            if (rr == 0) cs[q][cc - q * 16] = cpart;
        }
        rs[q][rr] = rsum;
        __syncthreads();
        double* o = part + (long long)b * 2 * TS;
        if (threadIdx.x < TS) {
            o[threadIdx.x] = rs[0][threadIdx.x] + rs[1][threadIdx.x] + rs[2][threadIdx.x] + rs[3][threadIdx.x];
        } else if (threadIdx.x < 2 * TS) {
            const int cc = threadIdx.x - TS;
            o[TS + cc] = cs[cc / 16][cc % 16];
        }
        return;
    }
    // panel dots: k < j -> t1[k] = W(0:L, iw+1+k) . v ; k >= j -> t2[k-j] = A(0:L, c+1+k-j) . v
    const int k = b - ntiles;
    const double* colp = k < j ? W + (long long)ldw * (iw + 1 + k) : A + (long long)n * (c + 1 + (k - j));
    double s = 0.0;
    for (int r = threadIdx.x; r < L; r += blockDim.x) s += colp[r] * vof(r);
    const double t = wg_sum(s, sh);
    if (threadIdx.x == 0) t12[k] = t;
}

// K_b2: rows r < L.  y(r) = sum of tile partials; w = tau (y - A_p t1 - W_p t2);
// W(r, iw) = w; pdot[block] = sum w v; stores v into A(r, c) and the unit, E, TAU.
__global__ __launch_bounds__(TB) void k_symv_finish(double* __restrict__ A, int n, double* __restrict__ W, int ldw,
                                                    int c, int iw, int j, int nt, const double* __restrict__ part,
                                                    const double* __restrict__ t12, const double* __restrict__ scal,
                                                    double* __restrict__ pdot, double* __restrict__ E,
                                                    double* __restrict__ TAU) {
    __shared__ double sh[TB / 64];
    __shared__ double red[4][64];
    const double tau = scal[1], scale = scal[2];
    const int L = c;
    // block Ib = 64 rows; the four waves split the tile partials and the panel columns
    const int Ib = blockIdx.x, rr = threadIdx.x & 63, q = threadIdx.x >> 6;
    const int r = Ib * TS + rr;
    double y = 0.0;
    if (r < L) {
        // row sums of tiles (Ib, J >= Ib) and column sums of tiles (I <= Ib, Ib)
        for (int J = Ib + q; J < nt; J += 4) y += part[(long long)(J * (J + 1) / 2 + Ib) * 2 * TS + rr];
        for (int I = q; I <= Ib; I += 4) y += part[(long long)(Ib * (Ib + 1) / 2 + I) * 2 * TS + TS + rr];
        for (int k = q; k < j; k += 4) {
            y -= A[r + (long long)n * (c + 1 + k)] * t12[k];
            y -= W[r + (long long)ldw * (iw + 1 + k)] * t12[j + k];
        }
    }
    red[q][rr] = y;
    __syncthreads();
    double d = 0.0;
    if (q == 0 && r < L) {
        const double w = tau * (red[0][rr] + red[1][rr] + red[2][rr] + red[3][rr]);
        W[r + (long long)ldw * iw] = w;
        double v;
        if (r == L - 1) {
            v = 1.0;
        } else {
            v = A[r + (long long)n * c] * scale;
            A[r + (long long)n * c] = v;
        }
        d = w * v;
    }
    const double t = wg_sum(d, sh);
    if (threadIdx.x == 0) pdot[blockIdx.x] = t;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        A[(c - 1) + (long long)n * c] = 1.0;
        E[c - 1] = scal[0];
        TAU[c - 1] = tau;
    }
}

// W(r, piw) += palpha A(r, c+1) for r <= c (the last column of a panel)
__global__ void k_w_finish(const double* __restrict__ A, int n, double* __restrict__ W, int ldw, int c, int piw,
                           const double* __restrict__ pdot, int npdot, const double* __restrict__ scal_prev) {
    __shared__ double sh[TB / 64];
    double s = 0.0;
    for (int k = threadIdx.x; k < npdot; k += blockDim.x) s += pdot[k];
    const double palpha = -0.5 * scal_prev[1] * wg_sum(s, sh);
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r <= c) W[r + (long long)ldw * piw] += palpha * A[r + (long long)n * (c + 1)];
}

__global__ void k_d0(const double* __restrict__ A, double* __restrict__ D) { D[0] = A[0]; }

}  // namespace

size_t sytrd_work_doubles(int n, int nb) {
    const int nt = (n + TS - 1) / TS;
    const size_t ntiles = (size_t)nt * (nt + 1) / 2;
    const size_t nblk = (n + 63) / 64;
    return (size_t)n * nb + ntiles * 2 * TS + 2 * nb + 2 * nblk + 8;
}

void sytrd_upper(hipStream_t s, int n, double* A, double* D, double* E, double* TAU, double* work, int nb) {
    QK_CHECK(n >= 1 && nb >= 1 && nb <= 64, QK_E_ARG, "sytrd: bad size");
    const int nt_all = (n + TS - 1) / TS;
    const size_t ntiles_all = (size_t)nt_all * (nt_all + 1) / 2;
    const int nblk_all = (n + 63) / 64;
    double* W = work;
    double* part = W + (size_t)n * nb;
    double* t12 = part + ntiles_all * 2 * TS;
    double* pnorm = t12 + 2 * nb;
    double* pdot = pnorm + nblk_all;
    double* scal = pdot + nblk_all;  // two slots of (beta, tau, scale): [0..2] and [4..6]
    const int ldw = n;
    int c_hi = n - 1;  // 0-based columns c_hi down to 1 carry reflectors
    while (c_hi >= 1) {
        const int nbp = std::min(nb, c_hi);
        const int c_lo = c_hi - nbp + 1;
        int piw = -1, npdot_prev = 0;
        for (int c = c_hi; c >= c_lo; --c) {
            const int iw = c - c_lo;
            const int j = c_hi - c;  // finished panel columns to the right
            const int nbu = (c + 1 + 63) / 64;  // 64 rows per block
            // scalars ping-pong between scal[0..2] and scal[4..6] by column parity
            double* sc_cur = scal + 4 * (c & 1);
            double* sc_prev = scal + 4 * ((c + 1) & 1);
            hipLaunchKernelGGL(k_col_update, dim3(nbu), dim3(TB), 0, s, A, n, W, ldw, c, iw, j, piw, pdot,
                               npdot_prev, sc_prev, pnorm, D);
            const int L = c;
            const int nt = (L + TS - 1) / TS;
            const int ntiles = nt * (nt + 1) / 2;
            const int npart = nbu;
            hipLaunchKernelGGL(k_symv_tiles, dim3(ntiles + 2 * j), dim3(TB), 0, s, A, n, W, ldw, c, iw, j, pnorm,
                               npart, nt, ntiles, part, t12, sc_cur);
            const int nbl = nt;  // one block per 64-row block
            hipLaunchKernelGGL(k_symv_finish, dim3(nbl), dim3(TB), 0, s, A, n, W, ldw, c, iw, j, nt, part, t12,
                               sc_cur, pdot, E, TAU);
            QK_HIP(hipGetLastError());
            piw = iw;
            npdot_prev = nbl;
        }
        // finalise the panel's last W column (that of c_lo), then the trailing update
        const int m = c_lo;  // rows/cols 0..c_lo-1 remain
        hipLaunchKernelGGL(k_w_finish, dim3((c_lo + TB - 1) / TB), dim3(TB), 0, s, A, n, W, ldw, c_lo - 1, piw, pdot,
                           npdot_prev, scal + 4 * (c_lo & 1));
        QK_HIP(hipGetLastError());
        if (m > 0) {
            Batch up;
            up.upper = 1;
            const double* V = A + (long long)n * c_lo;
            dgemm(s, false, true, m, m, nbp, -1.0, V, n, W, ldw, 1.0, A, n, up);
            dgemm(s, false, true, m, m, nbp, -1.0, W, ldw, V, n, 1.0, A, n, up);
        }
        c_hi = c_lo - 1;
    }
    hipLaunchKernelGGL(k_d0, dim3(1), dim3(1), 0, s, A, D);
    QK_HIP(hipGetLastError());
}

}  // namespace qk
