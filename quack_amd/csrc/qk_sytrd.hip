// Householder tridiagonalisation of a symmetric matrix (upper triangle),
// LAPACK dsytrd conventions (UPLO = 'U'): T = Q^T A Q with Q = H(n-1)...H(1),
// H(i) = I - tau v v^T, v(i+1:n) = 0, v(i) = 1, v(1:i-1) stored in A(1:i-1, i+1);
// D = diag(T), E = superdiagonal, TAU = tau.  The reflectors are therefore the
// ones rocSOLVER's dormtr applies for the back-transformation.
//
// Blocked as LAPACK's dsytrd/dlatrd (Golub & Van Loan, sec. 8.3; the panel
// recurrence of dlatrd): panels of nb columns from the right.  Two kernels per
// column c:
//   K_b1 Householder scalars (every workgroup re-reduces the norm partials),
//        symmetric matrix-vector product y = A v on the upper triangle in
//        64 x 64 tiles, grouped by column strip (row sums per tile, column
//        sums per group), partial sums of the quadratic form v^T A v;
//   K_b2 finishes column c (panel dots, y, the W column with its alpha term,
//        which v^T A v gives every workgroup without a further pass) and
//        updates column c-1 with the panel (its norm and panel-dot partials);
// a small kernel prepares the first column of each panel, and after each
// panel the trailing update A -= V W^T + W V^T runs as two matrix-core GEMMs
// on the upper tiles.  No host synchronisation inside.
#include <algorithm>

#include "qk_kernels.h"

namespace qk {
namespace {

constexpr int TB = 256;    // threads per block (tile products)
constexpr int TBW = 1024;  // threads per block (row-wise kernels: 64 rows x 16 waves)
constexpr int TS = 64;     // symv tile
constexpr int GT = 4;      // tiles per symv workgroup (one column strip)

// Workgroup barrier that orders LDS only: no kernel here exchanges global
// memory between its waves, so outstanding global loads and stores (the tile
// loads in flight, the column stores) need not drain at every barrier, as a
// __syncthreads() fence would make them.
__device__ inline void lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

__device__ inline double wg_sum(double v, double* sh) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    lds_sync();
    if (lane == 0) sh[w] = v;
    lds_sync();
    double r = 0.0;
    for (int k = 0; k < (int)(blockDim.x + 63) / 64; ++k) r += sh[k];
    lds_sync();
    return r;  // valid in every thread
}

// scalars of a column: [0] beta, [1] tau, [2] 1/(alpha - beta) (the scale of x)
// Householder scalars (dlarfg) from this thread's share of the norm partials
// and alpha = A(c-1, c); every thread of the block gets them
__device__ inline void householder(double pn, double alpha, double* sh, double& beta, double& tau, double& scale) {
    const double xx = wg_sum(pn, sh);
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

// strips of GT row tiles: strip J (columns 64J..64J+63) holds tiles I = 0..J
// in J/GT + 1 groups; groups_before(J) = sum over J' < J of (J'/GT + 1)
__host__ __device__ inline int groups_before(int J) {
    const int a = J / GT, b = J % GT;
    return GT * a * (a - 1) / 2 + a * b + J;
}

// K_a (first column of a panel, j = 0, piw = -1; general form kept): rows
// r <= c, 64 rows per block, the sixteen waves split the panel columns.  The previous W column piw (= iw + 1) still lacks palpha v_prev
// (v_prev = A(:, c+1), its unit element A(c, c+1) stored as 1): rows r < c are
// finalised in memory; row c, read by every block, is finalised on the fly.
// Then A(r, c) -= sum_k A(r, c+1+k) W(c, iw+1+k) + W(r, iw+1+k) A(c, c+1+k),
// k < j; pnorm[block] = sum over r < c-1 of the updated A(r, c)^2; D[c] = A(c, c).
__global__ __launch_bounds__(TBW) void k_col_update(double* __restrict__ A, int n, double* __restrict__ W, int ldw,
                                                    int c, int iw, int j, int piw, const double* __restrict__ pdot,
                                                    int npdot, const double* __restrict__ scal_prev,
                                                    double* __restrict__ pnorm, double* __restrict__ D,
                                                    double* __restrict__ tpart, int ldt) {
    __shared__ double sh[TBW / 64];
    __shared__ double red[TBW / 64][64];
    __shared__ double xs[64];
    constexpr int NW = TBW / 64;
    const int rr = threadIdx.x & 63, q = threadIdx.x >> 6;
    const int r = blockIdx.x * 64 + rr;
    const bool on = r <= c;
    // panel loads first (they do not depend on palpha)
    constexpr int KPW = 64 / NW;  // panel columns per wave (nb <= 64)
    double ar[KPW], wr[KPW], wc[KPW], ac[KPW];
#pragma unroll
    for (int t = 0; t < KPW; ++t) {
        const int k = q + NW * t;
        const bool use = on && k < j;
        const long long col = (long long)n * (c + 1 + k), wcol = (long long)ldw * (iw + 1 + k);
        ar[t] = use ? A[r + col] : 0.0;
        wr[t] = use ? W[r + wcol] : 0.0;
        wc[t] = use ? W[c + wcol] : 0.0;
        ac[t] = use ? A[c + col] : 0.0;
    }
    const double aold = on ? A[r + (long long)n * c] : 0.0;
    double palpha = 0.0;
    if (piw >= 0) {
        double s = 0.0;
        for (int k = threadIdx.x; k < npdot; k += blockDim.x) s += pdot[k];
        palpha = -0.5 * scal_prev[1] * wg_sum(s, sh);
    }
    double acc = 0.0;
#pragma unroll
    for (int t = 0; t < KPW; ++t) {
        double wck = wc[t], wrk = wr[t];
        if (q + NW * t == 0 && piw >= 0) {
            wck += palpha;               // v_prev(c) = 1
            wrk += palpha * ar[t];       // finalised W(r, piw)
        }
        acc += ar[t] * wck + wrk * ac[t];
    }
    red[q][rr] = acc;
    lds_sync();
    double sq = 0.0;
    if (q == 0) xs[rr] = 0.0;
    if (q == 0 && on) {
        if (piw >= 0 && r < c && j > 0) W[r + (long long)ldw * piw] = wr[0] + palpha * ar[0];
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < NW; ++k) s += red[k][rr];
        const double a = aold - s;
        A[r + (long long)n * c] = a;
        if (r < c - 1) {
            sq = a * a;
            xs[rr] = a;
        }
        if (r == c) D[c] = a;
    }
    const double t = wg_sum(sq, sh);  // its barriers also publish xs
    if (threadIdx.x == 0) pnorm[blockIdx.x] = t;
    // panel-dot partials over this block's rows of x (= A(r, c), r < c-1):
    // tpart[k][block] = W(:, iw+1+k) . x, tpart[j+k][block] = A(:, c+1+k) . x
    const double x = xs[rr];
#pragma unroll
    for (int t2 = 0; t2 < KPW; ++t2) {
        const int k = q + NW * t2;
        if (k >= j) break;
        double pw = (k == 0 && piw >= 0 ? wr[t2] + palpha * ar[t2] : wr[t2]) * x;
        double pa = ar[t2] * x;
        for (int off = 32; off > 0; off >>= 1) {
            pw += __shfl_down(pw, off, 64);
            pa += __shfl_down(pa, off, 64);
        }
        if (rr == 0) {
            tpart[(long long)k * ldt + blockIdx.x] = pw;
            tpart[(long long)(j + k) * ldt + blockIdx.x] = pa;
        }
    }
}

// K_b1: one block per group of up to GT upper tiles (I, J), I = GT g .., of
// one column strip J of A(0:L, 0:L), L = c.  v(r) = A(r, c) scale (r < c-1),
// v(c-1) = 1.  Every block loads its norm partial and vector entries, then its
// tiles, and recomputes the Householder scalars while the tiles arrive.  Each
// tile's row sums (its product with v_J) go to prow[tile], the strip's column
// sums (A_IJ^T v_I summed over the group's tiles, transposed through LDS) to
// pcol[group]; on a diagonal tile the row sums take the upper part with the
// diagonal, the column sums the strict part.
__global__ __launch_bounds__(TB) void k_symv_tiles(const double* __restrict__ A, int n, int c,
                                                   const double* __restrict__ pnorm, int npart,
                                                   double* __restrict__ prow, double* __restrict__ pcol,
                                                   double* __restrict__ qpart, double* __restrict__ scal) {
    __shared__ double sh[TB / 64];
    __shared__ double vJ[TS], vI[GT][TS];
    __shared__ double rs[GT][4][TS];
    __shared__ double cq[4][TS];
    __shared__ double ct[TS][TS + 1];
    const int L = c;
    const int gb = blockIdx.x;
    int J = (int)sqrt(8.0 * gb);
    while (J > 0 && groups_before(J) > gb) --J;
    while (groups_before(J + 1) <= gb) ++J;
    const int I0 = GT * (gb - groups_before(J));
    const int nI = min(GT, J - I0 + 1);
    const int c0 = J * TS;
    const int rr = threadIdx.x & 63, q = threadIdx.x >> 6;
    // small loads first: norm partials, alpha, raw vector entries (x(r) = A(r, c), r < L-1)
    double pn = 0.0;
    for (int k = threadIdx.x; k < npart; k += TB) pn += pnorm[k];
    const double alpha = A[(c - 1) + (long long)n * c];
    constexpr int NV = ((GT + 1) * TS + TB - 1) / TB;
    double xr[NV];
#pragma unroll
    for (int u = 0; u < NV; ++u) {
        const int e = threadIdx.x + u * TB;
        const int t = e / TS, x = e % TS;
        const int r = t == 0 ? c0 + x : (I0 + t - 1) * TS + x;
        xr[u] = (e < (GT + 1) * TS && r < L - 1) ? A[r + (long long)n * c] : 0.0;
    }
    // thread (rr, q): row rr of each tile, columns q*16 .. q*16+15 of the strip
    double a[GT][16];
#pragma unroll
    for (int t = 0; t < GT; ++t) {
        const int I = I0 + t;
        const int r = I * TS + rr;
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const int cc = q * 16 + k;
            const int col = c0 + cc;
            const bool use = t < nI && r < L && col < L && (I < J || rr <= cc);
            a[t][k] = use ? A[r + (long long)n * col] : 0.0;
        }
    }
    double beta, tau, scale;
    householder(pn, alpha, sh, beta, tau, scale);
    if (gb == 0 && threadIdx.x == 0) {
        scal[0] = beta;
        scal[1] = tau;
        scal[2] = scale;
    }
#pragma unroll
    for (int u = 0; u < NV; ++u) {
        const int e = threadIdx.x + u * TB;
        if (e < (GT + 1) * TS) {
            const int t = e / TS, x = e % TS;
            const int r = t == 0 ? c0 + x : (I0 + t - 1) * TS + x;
            const double v = r == L - 1 ? 1.0 : xr[u] * scale;
            if (t == 0) vJ[x] = v;
            else vI[t - 1][x] = v;
        }
    }
    lds_sync();
    double colacc[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) colacc[k] = 0.0;
#pragma unroll
    for (int t = 0; t < GT; ++t) {
        const double vi = vI[t][rr];
        const bool diag = I0 + t == J;
        double rsum = 0.0;
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const int cc = q * 16 + k;
            rsum += a[t][k] * vJ[cc];
            colacc[k] += (!diag || rr < cc) ? a[t][k] * vi : 0.0;
        }
        rs[t][q][rr] = rsum;
    }
#pragma unroll
    for (int k = 0; k < 16; ++k) ct[q * 16 + k][rr] = colacc[k];
    lds_sync();
    if (q < nI) prow[(long long)(J * (J + 1) / 2 + I0 + q) * TS + rr] = rs[q][0][rr] + rs[q][1][rr] + rs[q][2][rr] + rs[q][3][rr];
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < 16; ++k) s += ct[rr][q * 16 + k];
    cq[q][rr] = s;
    lds_sync();
    // v^T A v over the group: sum_t v_It . (A_IJ v_J) + v_J . (column sums); off-diagonal
    // tiles are so counted twice (A_IJ and A_JI), the diagonal tile once
    double qv = 0.0;
    if (q < nI) qv = vI[q][rr] * (rs[q][0][rr] + rs[q][1][rr] + rs[q][2][rr] + rs[q][3][rr]);
    if (q == 0) {
        const double cs = cq[0][rr] + cq[1][rr] + cq[2][rr] + cq[3][rr];
        pcol[(long long)gb * TS + rr] = cs;
        qv += vJ[rr] * cs;
    }
    const double qs = wg_sum(qv, sh);
    if (threadIdx.x == 0) qpart[gb] = qs;
}

// K_b2: rows r < L, 64 rows per block, sixteen waves; finishes column c and,
// unless c is the last column of its panel, updates column c-1.
//  1. panel dots t1 = W_p^T v, t2 = V_p^T v from the partials of x = A(:, c)
//     (16 lanes per dot) plus the unit row L-1 of v = x scale + e_{L-1};
//  2. w^T v = tau (v^T A v - 2 t1 . t2) from the tile kernel's quadratic-form
//     partials (w = tau (A v - V_p t1 - W_p t2)), so every block has
//     palpha = -tau w^T v / 2 without another pass;
//  3. y(r) from the tile partials, w = tau (y - V_p t1 - W_p t2), the finished
//     W(r, iw) = w + palpha v, A(r, c) = v (the unit at L-1); row L-1 of w is
//     computed by every block in the same order, for the update below;
//  4. column c-1: A(r, c-1) -= sum over panel columns k = c .. c_hi of
//     A(r, k) W(c-1, k) + W(r, k) A(c-1, k) (dlatrd's update of the next
//     column), D(c-1), the norm partial of the next vector x' (rows < c-2)
//     and its panel-dot partials W(:, k) . x', A(:, k) . x'.
__global__ __launch_bounds__(TBW) void k_finish_update(
    double* __restrict__ A, int n, double* __restrict__ W, int ldw, int c, int iw, int j, int nt, int last,
    const double* __restrict__ prow, const double* __restrict__ pcol, const double* __restrict__ qpart, int ngroups,
    const double* __restrict__ tcur, double* __restrict__ tnext, int ldt, int ntp, const double* __restrict__ scal,
    double* __restrict__ pnorm, double* __restrict__ D, double* __restrict__ E, double* __restrict__ TAU) {
    __shared__ double sh[TBW / 64];
    __shared__ double red[TBW / 64][64];
    __shared__ double red2[TBW / 64][64];
    __shared__ double t12[128];
    __shared__ double ylp[TBW / 64];  // per-wave panel corrections of row L-1
    __shared__ double wsh[64], vsh[64], xsh[64];
    constexpr int NW = TBW / 64;
    constexpr int KPW = 64 / NW;
    const double beta = scal[0], tau = scal[1], scale = scal[2];
    const int L = c;
    const int Ib = blockIdx.x, rr = threadIdx.x & 63, q = threadIdx.x >> 6;
    const int r = Ib * TS + rr;
    const bool on = r < L;
    // every load first.  Panel columns of this block's rows and of row L-1:
    double ar[KPW], wr[KPW], aL[KPW], wL[KPW];
#pragma unroll
    for (int t = 0; t < KPW; ++t) {
        const int k = q + NW * t;
        const bool use = k < j;
        const long long col = (long long)n * (c + 1 + k), wcol = (long long)ldw * (iw + 1 + k);
        ar[t] = use && on ? A[r + col] : 0.0;
        wr[t] = use && on ? W[r + wcol] : 0.0;
        aL[t] = use ? A[(L - 1) + col] : 0.0;
        wL[t] = use ? W[(L - 1) + wcol] : 0.0;
    }
    // A(r, c) and A(r, c-1) are rewritten below by wave 0 only, which alone reads them
    const double xr = q == 0 && on && r < L - 1 ? A[r + (long long)n * c] : 0.0;
    const double anext = q == 0 && !last && on ? A[r + (long long)n * (c - 1)] : 0.0;
    // panel-dot partials of x = A(:, c) (thread (kk, s16): 16 lanes per dot) and their unit-row terms
    const int s16 = threadIdx.x & 15, kk0 = threadIdx.x >> 4;
    double tp[2] = {0.0, 0.0}, tu[2] = {0.0, 0.0};
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int kk = kk0 + (TBW / 16) * h;
        if (kk < 2 * j) {
            for (int b = s16; b < ntp; b += 16) tp[h] += tcur[(long long)kk * ldt + b];
            if (s16 == 0)
                tu[h] = kk < j ? W[(L - 1) + (long long)ldw * (iw + 1 + kk)]
                               : A[(L - 1) + (long long)n * (c + 1 + (kk - j))];
        }
    }
    // quadratic-form partials (this thread's share)
    double qf = 0.0;
    for (int g = threadIdx.x; g < ngroups; g += TBW) qf += qpart[g];
    // tile partials of y for this row
    double y = 0.0;
    if (on) {
        const int nrow = nt - Ib, ng = Ib / GT + 1, g0 = groups_before(Ib);
        for (int e = q; e < nrow + ng; e += NW) {
            if (e < nrow) {
                const int J = Ib + e;
                y += prow[(long long)(J * (J + 1) / 2 + Ib) * TS + rr];
            } else {
                y += pcol[(long long)(g0 + e - nrow) * TS + rr];
            }
        }
    }
    // tile partials of y for row L-1 (wave 0, one per lane; the same order in every block)
    double yl = 0.0;
    if (q == 0) {
        const int IbL = (L - 1) / TS, rL = (L - 1) % TS;
        const int nrow = nt - IbL, ng = IbL / GT + 1, g0 = groups_before(IbL);
        for (int e = rr; e < nrow + ng; e += 64) {
            if (e < nrow) {
                const int J = IbL + e;
                yl += prow[(long long)(J * (J + 1) / 2 + IbL) * TS + rL];
            } else {
                yl += pcol[(long long)(g0 + e - nrow) * TS + rL];
            }
        }
        for (int off = 32; off > 0; off >>= 1) yl += __shfl_xor(yl, off, 64);
    }
    // panel dots t1 = W_p^T v, t2 = V_p^T v (v = x scale + e_{L-1})
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        double t = tp[h];
        for (int off = 8; off > 0; off >>= 1) t += __shfl_xor(t, off, 64);
        const int kk = kk0 + (TBW / 16) * h;
        if (s16 == 0 && kk < 2 * j) t12[kk] = scale * t + tu[h];
    }
    const double vAv = wg_sum(qf, sh);  // its barriers also publish t12
    // panel corrections of y; update partials of column c-1 (panel columns k >= 1);
    // the corrections of row L-1, the same in every lane of a wave
    double u = 0.0, ylc = 0.0;
#pragma unroll
    for (int t = 0; t < KPW; ++t) {
        const int k = q + NW * t;
        if (k < j) {
            y -= ar[t] * t12[k] + wr[t] * t12[j + k];
            u += ar[t] * wL[t] + wr[t] * aL[t];
            ylc += aL[t] * t12[k] + wL[t] * t12[j + k];
        }
    }
    red[q][rr] = y;
    red2[q][rr] = u;
    if (rr == 0) ylp[q] = ylc;
    lds_sync();
    double sq = 0.0;
    if (q == 0) {
        // row L-1 of w and alpha = -tau w^T v / 2, w^T v = tau (v^T A v - 2 t1 . t2)
        double corr = 0.0;
#pragma unroll
        for (int k = 0; k < NW; ++k) corr += ylp[k];
        double tt = 0.0;
        for (int k = 0; k < j; ++k) tt += t12[k] * t12[j + k];
        const double palpha = -0.5 * tau * (tau * (vAv - 2.0 * tt));
        const double wLfin = tau * (yl - corr) + palpha;  // v(L-1) = 1
        wsh[rr] = 0.0;
        vsh[rr] = 0.0;
        xsh[rr] = 0.0;
        if (on) {
            double ys = 0.0, us = 0.0;
#pragma unroll
            for (int k = 0; k < NW; ++k) {
                ys += red[k][rr];
                us += red2[k][rr];
            }
            const double v = r == L - 1 ? 1.0 : xr * scale;
            const double wf = r == L - 1 ? wLfin : tau * ys + palpha * v;
            W[r + (long long)ldw * iw] = wf;
            A[r + (long long)n * c] = v;
            wsh[rr] = wf;
            vsh[rr] = v;
            if (!last) {
                // column c-1, rows r <= c-1: the k = 0 panel column is c itself (A(c-1, c) = 1)
                const double a = anext - (us + v * wLfin + wf);
                A[r + (long long)n * (c - 1)] = a;
                if (r == L - 1) D[c - 1] = a;
                if (r < L - 2) {
                    sq = a * a;
                    xsh[rr] = a;
                }
            }
        }
    }
    if (!last) {
        const double t = wg_sum(sq, sh);  // its barriers also publish wsh, vsh, xsh
        if (threadIdx.x == 0) pnorm[blockIdx.x] = t;
        // panel-dot partials of x' for column c-1: panel columns k = 0 (c itself) .. j
        const double x = xsh[rr];
        const int jn = j + 1;
#pragma unroll
        for (int t = 0; t <= KPW; ++t) {
            // wave q handles k = 0 (wave 0 only, t = 0) and k = 1 + q + NW (t - 1)
            int k;
            double wv, av;
            if (t == 0) {
                if (q != 0) continue;
                k = 0;
                wv = wsh[rr];
                av = vsh[rr];
            } else {
                k = 1 + q + NW * (t - 1);
                if (k > j) break;
                wv = wr[t - 1];
                av = ar[t - 1];
            }
            double pw = wv * x, pa = av * x;
            for (int off = 32; off > 0; off >>= 1) {
                pw += __shfl_down(pw, off, 64);
                pa += __shfl_down(pa, off, 64);
            }
            if (rr == 0) {
                tnext[(long long)k * ldt + blockIdx.x] = pw;
                tnext[(long long)(jn + k) * ldt + blockIdx.x] = pa;
            }
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        E[c - 1] = beta;
        TAU[c - 1] = tau;
    }
}

__global__ void k_d0(const double* __restrict__ A, double* __restrict__ D) { D[0] = A[0]; }

// The panel's reflector columns V = A(0:m, c_lo .. c_lo+nbp-1) copied on both
// sides of W in the work array, so that [V W] and [W V] are column blocks of
// one array: A -= [V W] [W V]^T = V W^T + W V^T is a single rank-2nbp update.
__global__ void k_copy_v(const double* __restrict__ V, int n, int m, int nbp, double* __restrict__ left,
                         double* __restrict__ right) {
    const long long total = (long long)m * nbp;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total;
         t += (long long)gridDim.x * blockDim.x) {
        const int r = (int)(t % m), k = (int)(t / m);
        const double v = V[r + (long long)n * k];
        left[r + (long long)n * k] = v;
        right[r + (long long)n * k] = v;
    }
}


// ---- back-transformation Z <- Q Z (dormtr, SIDE = 'L', UPLO = 'U', TRANS = 'N') ----
// Q = H(n-1) ... H(1); reflector j (0-based, column c = j + 1) is v_j with
// v_j(r) = A(r, j+1) for r < j, v_j(j) = 1, zero below; tau_j = TAU[j].  Blocks of
// OK reflectors in increasing j are applied in turn, each in compact WY form
// H_{j0+OK-1} ... H_{j0} = I - V T V^T with T lower triangular (LAPACK dlarft,
// DIRECT = 'B'): Z(0:m, :) -= V (T (V^T Z(0:m, :))), m = j0 + OK.
constexpr int OK = 128;   // reflectors per block
constexpr int OKC = 512;  // rows per split-K chunk of V^T Z

// V (n x nbk, zero-padded columns past n-1) and the padded taus
__global__ void k_build_v(const double* __restrict__ A, int n, const double* __restrict__ TAU, int nbk,
                          double* __restrict__ V, double* __restrict__ taup) {
    const long long total = (long long)n * nbk;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total;
         t += (long long)gridDim.x * blockDim.x) {
        const int r = (int)(t % n), j = (int)(t / n);
        double v = 0.0;
        if (j < n - 1) v = r < j ? A[r + (long long)n * (j + 1)] : (r == j ? 1.0 : 0.0);
        V[t] = v;
        if (r == 0) taup[j] = j < n - 1 ? TAU[j] : 0.0;
    }
}

// T of each block (one workgroup per block) from G = V^T V:
// T(j, j) = tau_j; T(j+1:, j) = -tau_j T(j+1:, j+1:) G(j+1:, j), j descending.
// Written OKS times side by side, [T T ... T] (OK x OKS OK), so that one GEMM
// applies T to the sum of the split-K partials.
__global__ __launch_bounds__(256) void k_build_t(const double* __restrict__ G, const double* __restrict__ taup,
                                                 double* __restrict__ Trep, int oks) {
    __shared__ double T[OK][OK + 1];
    __shared__ double gc[OK];
    const int b = blockIdx.x;
    const double* Gb = G + (long long)b * OK * OK;
    const double* tb = taup + (long long)b * OK;
    for (int e = threadIdx.x; e < OK * OK; e += blockDim.x) T[e % OK][e / OK] = 0.0;
    lds_sync();
    if (threadIdx.x == 0) T[OK - 1][OK - 1] = tb[OK - 1];
    for (int j = OK - 2; j >= 0; --j) {
        if (threadIdx.x < OK) gc[threadIdx.x] = Gb[threadIdx.x + (long long)OK * j];
        lds_sync();
        const int i = threadIdx.x;
        if (i > j && i < OK) {
            double t = 0.0;
            for (int l = j + 1; l <= i; ++l) t += T[i][l] * gc[l];
            T[i][j] = -tb[j] * t;
        }
        if (i == j) T[j][j] = tb[j];
        lds_sync();
    }
    double* o = Trep + (long long)b * OK * OK * oks;
    for (long long e = threadIdx.x; e < (long long)OK * OK * oks; e += blockDim.x) {
        const int i = (int)(e % OK), col = (int)((e / OK) % OK);
        o[e] = T[i][col];
    }
}

}  // namespace

size_t sytrd_work_doubles(int n, int nb) {
    const int nt = (n + TS - 1) / TS;
    const size_t ntiles = (size_t)nt * (nt + 1) / 2;
    const size_t ng = groups_before(nt);
    const size_t nblk = (n + 63) / 64;
    return 3 * (size_t)n * nb + (ntiles + ng) * TS + ng + 2 * (2 * (size_t)nb * nblk) + nblk + 8;
}

void sytrd_upper(hipStream_t s, int n, double* A, double* D, double* E, double* TAU, double* work, int nb) {
    QK_CHECK(n >= 1 && nb >= 1 && nb <= 64, QK_E_ARG, "sytrd: bad size");
    const int nt_all = (n + TS - 1) / TS;
    const size_t ntiles_all = (size_t)nt_all * (nt_all + 1) / 2;
    const size_t ng_all = groups_before(nt_all);
    const int nblk_all = (n + 63) / 64;
    // work: P = [V copy | W | V copy] (n x 3nb), then the partial sums
    double* P = work;
    double* W = P + (size_t)n * nb;
    double* prow = P + 3 * (size_t)n * nb;
    double* pcol = prow + ntiles_all * TS;
    double* qpart = pcol + ng_all * TS;
    double* tpart0 = qpart + ng_all;  // two buffers of [2 nb][nblk_all], by column parity
    double* tpart1 = tpart0 + 2 * (size_t)nb * nblk_all;
    double* pnorm = tpart1 + 2 * (size_t)nb * nblk_all;
    double* scal = pnorm + nblk_all;  // two slots of (beta, tau, scale): [0..2] and [4..6]
    const int ldw = n, ldt = nblk_all;
    int c_hi = n - 1;  // 0-based columns c_hi down to 1 carry reflectors
    while (c_hi >= 1) {
        const int nbp = std::min(nb, c_hi);
        const int c_lo = c_hi - nbp + 1;
        // first column of the panel: no panel terms; its norm partials and diagonal
        hipLaunchKernelGGL(k_col_update, dim3((c_hi + 1 + 63) / 64), dim3(TBW), 0, s, A, n, W, ldw, c_hi, 0, 0, -1,
                           (const double*)nullptr, 0, (const double*)nullptr, pnorm, D, tpart0, ldt);
        for (int c = c_hi; c >= c_lo; --c) {
            const int iw = c - c_lo;
            const int j = c_hi - c;  // finished panel columns to the right
            const int npart = (c + 1 + 63) / 64;  // partial blocks over rows 0..c
            double* sc_cur = scal + 4 * (c & 1);
            double* tcur = (c & 1) ? tpart1 : tpart0;
            double* tnext = (c & 1) ? tpart0 : tpart1;
            const int L = c;
            const int nt = (L + TS - 1) / TS;
            const int ng = groups_before(nt);
            hipLaunchKernelGGL(k_symv_tiles, dim3(ng), dim3(TB), 0, s, A, n, c, pnorm, npart, prow, pcol, qpart,
                               sc_cur);
            hipLaunchKernelGGL(k_finish_update, dim3(nt), dim3(TBW), 0, s, A, n, W, ldw, c, iw, j, nt,
                               c == c_lo ? 1 : 0, prow, pcol, qpart, ng, tcur, tnext, ldt, npart, sc_cur, pnorm, D, E,
                               TAU);
            QK_HIP(hipGetLastError());
        }
        // trailing update A -= V W^T + W V^T on the upper tiles, as one
        // product of K = 2 nbp: [V W] = P(:, nb-nbp .. nb+nbp), [W V] = P(:, nb .. nb+2nbp)
        const int m = c_lo;  // rows/cols 0..c_lo-1 remain
        if (m > 0) {
            Batch up;
            up.upper = 1;
            const double* V = A + (long long)n * c_lo;
            double* left = P + (long long)n * (nb - nbp);
            double* right = P + (long long)n * (nb + nbp);
            const long long tot = (long long)m * nbp;
            hipLaunchKernelGGL(k_copy_v, dim3((unsigned)std::min<long long>((tot + 255) / 256, 4096)), dim3(256), 0,
                               s, V, n, m, nbp, left, right);
            QK_HIP(hipGetLastError());
            dgemm(s, false, true, m, m, 2 * nbp, -1.0, left, n, W, n, 1.0, A, n, up);
        }
        c_hi = c_lo - 1;
    }
    hipLaunchKernelGGL(k_d0, dim3(1), dim3(1), 0, s, A, D);
    QK_HIP(hipGetLastError());
}


size_t ormtr_work_doubles(int n) {
    const int nblk = (n - 1 + OK - 1) / OK;
    const size_t nbk = (size_t)nblk * OK;
    const int oks = (n + OKC - 1) / OKC;
    return (size_t)n * nbk + nbk + (size_t)nblk * OK * OK + (size_t)nblk * OK * OK * oks + (size_t)oks * OK * n +
           (size_t)OK * n;
}

void ormtr_upper(hipStream_t s, int n, const double* A, const double* TAU, double* Z, double* work) {
    QK_CHECK(n >= 1, QK_E_ARG, "ormtr: bad size");
    if (n == 1) return;
    const int nblk = (n - 1 + OK - 1) / OK;
    const int nbk = nblk * OK;
    const int oks = (n + OKC - 1) / OKC;  // largest number of split-K chunks
    double* V = work;
    double* taup = V + (size_t)n * nbk;
    double* G = taup + nbk;
    double* Trep = G + (size_t)nblk * OK * OK;
    double* W1p = Trep + (size_t)nblk * OK * OK * oks;  // (oks OK) x n, partial V^T Z stacked by chunk
    double* W2 = W1p + (size_t)oks * OK * n;            // OK x n
    const long long tot = (long long)n * nbk;
    hipLaunchKernelGGL(k_build_v, dim3((unsigned)std::min<long long>((tot + 255) / 256, 65536)), dim3(256), 0, s, A,
                       n, TAU, nbk, V, taup);
    QK_HIP(hipGetLastError());
    {
        Batch bg;
        bg.sa0 = (long long)OK * n;
        bg.sb0 = (long long)OK * n;
        bg.sc0 = (long long)OK * OK;
        bg.count = nblk;
        bg.zdiv = nblk;  // one-level batch: z0 = z runs over the sa0/sb0/sc0 strides
        dgemm(s, true, false, OK, OK, n, 1.0, V, n, V, n, 0.0, G, OK, bg);
    }
    hipLaunchKernelGGL(k_build_t, dim3(nblk), dim3(256), 0, s, G, taup, Trep, oks);
    QK_HIP(hipGetLastError());
    const long long ldp = (long long)oks * OK;
    for (int b = 0; b < nblk; ++b) {
        const int j0 = b * OK;
        const int m = std::min(j0 + OK, n);
        const double* Vb = V + (long long)n * j0;
        // W1p = V(0:m, :)^T Z(0:m, :) in chunks of OKC rows, stacked
        const int s1 = m / OKC, rem = m % OKC;
        if (s1 > 0) {
            Batch bp;
            bp.sa0 = OKC;
            bp.sb0 = OKC;
            bp.sc0 = OK;
            bp.count = s1;
            bp.zdiv = s1;
            dgemm(s, true, false, OK, n, OKC, 1.0, Vb, n, Z, n, 0.0, W1p, ldp, bp);
        }
        if (rem > 0)
            dgemm(s, true, false, OK, n, rem, 1.0, Vb + (long long)s1 * OKC, n, Z + (long long)s1 * OKC, n, 0.0,
                  W1p + (long long)s1 * OK, ldp);
        const int sc = s1 + (rem > 0 ? 1 : 0);
        // W2 = T (sum of the partials) = [T ... T] W1p
        dgemm(s, false, false, OK, n, sc * OK, 1.0, Trep + (long long)b * OK * OK * oks, OK, W1p, ldp, 0.0, W2, OK);
        // Z(0:m, :) -= V(0:m, :) W2
        dgemm(s, false, false, m, n, OK, -1.0, Vb, n, W2, OK, 1.0, Z, n);
    }
}

}  // namespace qk
