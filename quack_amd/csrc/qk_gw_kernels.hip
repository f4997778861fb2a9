// HBM-streaming FP64 kernels of the restricted GW path (gfx950).
//
// Index conventions on the device (all column-major, 0-based):
//   M[p + n q + n^2 jb]        (pq|jb), the only ERI block the path needs
//   K[ia + nS jb]              (ia|jb), ia = a' + v*i', jb = b' + v*j'
//   Cm[ia + nS jb]             RPA matrix, then its eigenvectors (syevd)
//   rhoT[m + nS (p + n q)]     excitation densities rho(p,q,m): contiguous in m
//   where n = nOrb, o = nO - nC, v = nOrb - nR - nO, nS = o*v, i = nC + i',
//   a = nO + a'.
#include "qk_kernels.h"

namespace qk {
namespace {

constexpr int RB = 256;  // reduction block

__device__ inline double block_sum(double v, double* sh) {
    // wavefront reduction (64 lanes) then across the waves of the block
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0)
        for (int k = 0; k < (int)(blockDim.x + 63) / 64; ++k) r += sh[k];
    return r;  // valid in thread 0
}

// ---- ERI block extraction -------------------------------------------------

// X[p + n j' + n o q + n o n b'] = ERI(p, nC+j', q, nO+b')  ->  M[p + n q + n^2 (b' + v j')]
__global__ void k_permute_block(const double* __restrict__ X, double* __restrict__ M, int n, int o, int v) {
    long long total = (long long)n * n * o * v;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total; t += (long long)gridDim.x * blockDim.x) {
        long long p = t % n, r = t / n;
        long long q = r % n; r /= n;
        long long jb = r;  // b' + v j'
        long long b = jb % v, j = jb / v;
        M[t] = X[p + n * (j + o * (q + (long long)n * b))];
    }
}

// K[ia + nS jb] = M[(nC+i') + n (nO+a') + n^2 jb]
__global__ void k_extract_k(const double* __restrict__ M, double* __restrict__ K, int n, int nC, int nO, int v, int nS) {
    long long total = (long long)nS * nS;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total; t += (long long)gridDim.x * blockDim.x) {
        long long ia = t % nS, jb = t / nS;
        long long a = ia % v, i = ia / v;
        K[t] = M[(nC + i) + (long long)n * (nO + a) + (long long)n * n * jb];
    }
}

// Unrestricted pairs (no frozen orbitals): pair ia of spin s is a' + v_s i'
// within its spin block, the beta block after the alpha one.  The restricted
// RPA kernels then build the unrestricted matrices unchanged from
// K[ia + nSt jb] = (ia|jb) / 2: A = D + K_u, B = K_u give A + B = D + 2 K_u and
// Tr A = sum D + K_u(ia, ia), the restricted forms with 2K in place of K_u.
__device__ inline void upair(int ia, const UPairs& u, int& s, int& i, int& a) {
    s = ia >= u.nSa;
    const int l = ia - (s ? u.nSa : 0);
    const int v = u.n - u.nR[s] - u.nO[s];
    a = u.nO[s] + l % v;
    i = u.nC[s] + l / v;
}

// K[ia + nSt jb] = (ia|jb) / 2 from M_s[p + n q + n^2 jb] = (pq|jb), s = spin of ia
__global__ void k_extract_ku(const double* __restrict__ M0, const double* __restrict__ M1, double* __restrict__ K,
                             UPairs u) {
    const int n = u.n, nSt = u.nSt;
    const long long total = (long long)nSt * nSt;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total;
         t += (long long)gridDim.x * blockDim.x) {
        const int ia = (int)(t % nSt);
        const long long jb = t / nSt;
        int s, i, a;
        upair(ia, u, s, i, a);
        const double* M = s ? M1 : M0;
        K[t] = 0.5 * M[i + (long long)n * a + (long long)n * n * jb];
    }
}

// dE[ia] = e_s[a] - e_s[i], sd[ia] = sqrt(dE), e holding e_alpha then e_beta
__global__ void k_delta_u(const double* __restrict__ e, double* __restrict__ dE, double* __restrict__ sd, UPairs u) {
    const int ia = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = u.n;
    if (ia < u.nSt) {
        int s, i, a;
        upair(ia, u, s, i, a);
        const double d = e[n * s + a] - e[n * s + i];
        dE[ia] = d;
        sd[ia] = sqrt(d);
    }
}

// ---- RPA matrix ---------------------------------------------------------

// dE[ia] = e_a - e_i ; sd[ia] = sqrt(dE)
__global__ void k_delta(const double* __restrict__ e, double* __restrict__ dE, double* __restrict__ sd, int nC, int nO,
                        int v, int nS, Zs zs) {
    const long long z = blockIdx.z;
    e += z * zs.n;
    dE += z * zs.S;
    sd += z * zs.S;
    int ia = blockIdx.x * blockDim.x + threadIdx.x;
    if (ia < nS) {
        int a = ia % v, i = ia / v;
        double d = e[nO + a] - e[nC + i];
        dE[ia] = d;
        sd[ia] = sqrt(d);
    }
}

// TDA: Cm = A = diag(dE) + 2K ; RPA: Cm = (A-B)^1/2 (A+B) (A-B)^1/2 = diag(dE^2) + 4 sd K sd
__global__ void k_build_c(const double* __restrict__ K, const double* __restrict__ dE, const double* __restrict__ sd,
                          double* __restrict__ Cm, int nS, int tda, Zs zs) {
    const long long z = blockIdx.z;
    K += z * zs.S2;
    dE += z * zs.S;
    sd += z * zs.S;
    Cm += z * zs.S2;
    long long total = (long long)nS * nS;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total; t += (long long)gridDim.x * blockDim.x) {
        int ia = (int)(t % nS), jb = (int)(t / nS);
        double k = K[t];
        double c;
        if (tda) {
            c = 2.0 * k;
            if (ia == jb) c += dE[ia];
        } else {
            c = 4.0 * sd[ia] * k * sd[jb];
            if (ia == jb) c += dE[ia] * dE[ia];
        }
        Cm[t] = c;
    }
}

// out[0] = Tr(A) = sum_ia dE[ia] + 2 K[ia,ia]
__global__ void k_trace_a(const double* __restrict__ K, const double* __restrict__ dE, int nS, double* out, Zs zs) {
    const long long z = blockIdx.z;
    K += z * zs.S2;
    dE += z * zs.S;
    out += z * zs.c;
    __shared__ double sh[RB / 64];
    double s = 0.0;
    for (int ia = threadIdx.x; ia < nS; ia += blockDim.x) s += dE[ia] + 2.0 * K[(long long)ia * nS + ia];
    double r = block_sum(s, sh);
    if (threadIdx.x == 0) out[0] = r;
}

// Om = sqrt|lambda| (RPA) or lambda (TDA); out[1] = EcRPA = (sum Om - TrA)/2 with TrA = out[0]
__global__ void k_finalize_om(double* __restrict__ Om, int nS, int tda, double* out, Zs zs) {
    const long long z = blockIdx.z;
    Om += z * zs.S;
    out += z * zs.c;
    __shared__ double sh[RB / 64];
    double s = 0.0;
    for (int m = threadIdx.x; m < nS; m += blockDim.x) {
        double w = Om[m];
        if (!tda) w = sqrt(fabs(w));
        Om[m] = w;
        s += w;
    }
    double r = block_sum(s, sh);
    if (threadIdx.x == 0) out[1] = 0.5 * (r - out[0]);
}

// W[jb + nS m] = Z[jb + nS m] * sd[jb] / sqrt(Om[m])   (so that rho = M W)
__global__ void k_scale_w(double* __restrict__ W, const double* __restrict__ sd, const double* __restrict__ Om, int nS,
                          Zs zs) {
    const long long z = blockIdx.z;
    W += z * zs.S2;
    sd += z * zs.S;
    Om += z * zs.S;
    long long total = (long long)nS * nS;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total; t += (long long)gridDim.x * blockDim.x) {
        int jb = (int)(t % nS), m = (int)(t / nS);
        W[t] = W[t] * sd[jb] / sqrt(Om[m]);
    }
}

// ---- self-energy --------------------------------------------------------

// Per (p, chunk of m): partial sums over q and m of
//   occupied q:  eps = w_p - e_q + Om_m ;  virtual q: eps = w_p - e_q - Om_m
//   sig += 2 rho^2 eps/(eps^2+eta^2) ; dsig -= 2 rho^2 (eps^2-eta^2)/(eps^2+eta^2)^2
// (SRG: sig += 2 rho^2 (1-exp(-2 s eps^2))/eps, dsig -= 2 rho^2 (1-exp(-2 s eps^2))/eps^2).
// part[3*(p*nchunk + c) + {0,1,2}] = {sig_occ, sig_vir, dsig}.
// p runs over the active orbitals p = nC + blockIdx.y; skip[p] != 0 exits.
// Each thread owns SIG_MPT modes m (chunk = RB * SIG_MPT), keeps their Om in
// registers and streams rho(p, q, m) for all q: the loads of row q+1 are issued
// before the arithmetic of row q, and every load is unconditional (padding
// lanes read the last mode and discard it), so no load waits inside a branch.
// For 2 s eps^2 > 38, exp(-2 s eps^2) < 2^-54 and 1 - exp(...) rounds to 1.0,
// so the exponential is skipped there with bit-identical results.
// RhoT: double, or float where rho comes from the split-precision contraction
// (whose FP32-level accuracy the FP32 copy keeps); every sum is FP64.
constexpr int SIG_MPT = 4;
// 1/x from the hardware reciprocal and two Newton steps (r <- r + r (1 - x r)),
// each squaring the relative error: within an ulp or two of the rounded
// quotient, in 5 instructions instead of the 10 of an IEEE division.  x is
// an energy difference here (no overflow or denormal scaling is needed).
__device__ __forceinline__ double recip(double x) {
    double r = __builtin_amdgcn_rcp(x);
    r = fma(r, fma(-x, r, 1.0), r);
    return fma(r, fma(-x, r, 1.0), r);
}
template <typename RhoT>
__global__ __launch_bounds__(RB) void k_sigma_partial(const RhoT* __restrict__ rhoT, const double* __restrict__ w,
                                                      const double* __restrict__ e, const double* __restrict__ Om,
                                                      const int* __restrict__ skip, double eta, int srg, double flow,
                                                      int n, int nC, int nO, int nR, int nS,
                                                      double* __restrict__ part, Zs zs) {
    {
        // a float copy of rho occupies the first half of each molecule's double slot
        const long long z = blockIdx.z;
        rhoT += z * zs.R * (long long)(sizeof(double) / sizeof(RhoT));
        w += z * zs.n;
        e += z * zs.n;
        Om += z * zs.S;
        if (skip) skip += z * zs.n;
        part += z * zs.P;
    }
    const int p = nC + blockIdx.y;
    const int c = blockIdx.x;
    const int nchunk = gridDim.x;
    if (skip && skip[p]) return;
    const double wp = w[p];
    const double eta2 = eta * eta;
    const int mb = c * (RB * SIG_MPT) + threadIdx.x;
    double om[SIG_MPT];
    int mk[SIG_MPT];
    bool vk[SIG_MPT];
#pragma unroll
    for (int k = 0; k < SIG_MPT; ++k) {
        const int m = mb + k * RB;
        vk[k] = m < nS;
        mk[k] = vk[k] ? m : nS - 1;
        om[k] = vk[k] ? Om[m] : 0.0;
    }
    const int q0 = nC, q1 = n - nR;
    RhoT cur[SIG_MPT], nxt[SIG_MPT];
    if (q0 < q1) {
        const RhoT* rq = rhoT + (long long)nS * (p + (long long)n * q0);
#pragma unroll
        for (int k = 0; k < SIG_MPT; ++k) cur[k] = rq[mk[k]];
    }
    double so = 0.0, sv = 0.0, ds = 0.0;
    for (int q = q0; q < q1; ++q) {
        const int qn = q + 1 < q1 ? q + 1 : q;  // the last row reloads itself
        const RhoT* rn = rhoT + (long long)nS * (p + (long long)n * qn);  // rho(p, q+1, :)
#pragma unroll
        for (int k = 0; k < SIG_MPT; ++k) nxt[k] = rn[mk[k]];
        const double sgn = q < nO ? 1.0 : -1.0;
        const double wq = wp - e[q];
        double s_acc = 0.0, d_acc = 0.0;
#pragma unroll
        for (int k = 0; k < SIG_MPT; ++k) {
            if (!vk[k]) continue;  // padding modes contribute nothing (eps may be 0 there)
            const double r = (double)cur[k];
            double num = 2.0 * r * r;
            double eps = wq + sgn * om[k];
            double e2 = eps * eps;
            if (!srg) {
                double den = recip(e2 + eta2);
                s_acc += num * eps * den;
                d_acc -= num * (e2 - eta2) * den * den;
            } else {
                double x = 2.0 * flow * e2;
                double g = x > 38.0 ? 1.0 : 1.0 - exp(-x);
                double inv = recip(eps);
                double reg = num * g * inv;
                s_acc += reg;
                d_acc -= reg * inv;
            }
        }
        if (q < nO) so += s_acc; else sv += s_acc;
        ds += d_acc;
#pragma unroll
        for (int k = 0; k < SIG_MPT; ++k) cur[k] = nxt[k];
    }
    __shared__ double sh[RB / 64];
    double r0 = block_sum(so, sh);
    double r1 = block_sum(sv, sh);
    double r2 = block_sum(ds, sh);
    if (threadIdx.x == 0) {
        double* o = part + 3LL * ((long long)p * nchunk + c);
        o[0] = r0; o[1] = r1; o[2] = r2;
    }
}

// Static self-energy (RGW_self_energy_diag): Sig, Z = 1/(1-dSig) for active p;
// zero outside (GW/RGW_self_energy_diag.f90:115-116,191); EcGM = -2 sum_{p virtual} sig_occ(p)
// (the same eps = e_a - e_i + Om_m as GW/RGW_self_energy_diag.f90:179-181).
__global__ void k_sigma_finalize(const double* __restrict__ part, int nchunk, int n, int nC, int nO, int nR,
                                 double* __restrict__ Sig, double* __restrict__ Z, double* out, Zs zs) {
    const long long z = blockIdx.z;
    part += z * zs.P;
    Sig += z * zs.n;
    Z += z * zs.n;
    out += z * zs.c;
    __shared__ double sh[RB / 64];
    double gm = 0.0;
    for (int p = threadIdx.x; p < n; p += blockDim.x) {
        double s = 0.0, d = 0.0, so = 0.0;
        if (p >= nC && p < n - nR) {
            for (int c = 0; c < nchunk; ++c) {
                const double* o = part + 3LL * ((long long)p * nchunk + c);
                so += o[0];
                s += o[0] + o[1];
                d += o[2];
            }
            if (p >= nO) gm += so;
        }
        Sig[p] = s;
        Z[p] = 1.0 / (1.0 - d);
    }
    double r = block_sum(gm, sh);
    if (threadIdx.x == 0) out[2] = -2.0 * r;
}

// Sig(p), dSig(p) for every active p from the partial sums (zero elsewhere)
__global__ void k_sigma_out(const double* __restrict__ part, int nchunk, int n, int nC, int nR, double* __restrict__ Sig,
                            double* __restrict__ dSig) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    double s = 0.0, d = 0.0;
    if (p >= nC && p < n - nR)
        for (int c = 0; c < nchunk; ++c) {
            const double* o = part + 3LL * ((long long)p * nchunk + c);
            s += o[0] + o[1];
            d += o[2];
        }
    Sig[p] = s;
    dSig[p] = d;
}

// eGWlin = e0 + Z*Sig (GW/RG0W0.f90:188) ; or e0 + Sig (evRGW.f90:147, zflag = 0)
__global__ void k_lin(const double* e0, const double* Z, const double* Sig, double* out, int n, int zflag, Zs zs) {
    const long long z = blockIdx.z;
    e0 += z * zs.n;
    Z += z * zs.n;
    Sig += z * zs.n;
    out += z * zs.n;
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) out[p] = e0[p] + (zflag ? Z[p] * Sig[p] : Sig[p]);
}

// ---- QP Newton (GW/RGW_QP_graph.f90:50-92) --------------------------------

__global__ void k_qp_init(const double* __restrict__ eGWlin, double* __restrict__ w, double* __restrict__ f,
                          int* __restrict__ nit, int* __restrict__ done, int n, int nC, int nR, Zs zs) {
    const long long z = blockIdx.z;
    eGWlin += z * zs.n;
    w += z * zs.n;
    f += z * zs.n;
    nit += z * zs.n;
    done += z * zs.n;
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) {
        w[p] = eGWlin[p];
        f[p] = 1.0;
        nit[p] = 0;
        done[p] = (p < nC || p >= n - nR) ? 1 : 0;
    }
}

// One Newton step for every active orbital; active[0] = number still active.
__global__ void k_qp_update(const double* __restrict__ part, int nchunk, const double* __restrict__ eHF,
                            double* __restrict__ w, double* __restrict__ f, double* __restrict__ df,
                            int* __restrict__ nit, int* __restrict__ done, int n, double thresh, int maxit,
                            int* __restrict__ active, Zs zs) {
    {
        const long long z = blockIdx.z;
        part += z * zs.P;
        eHF += z * zs.n;
        w += z * zs.n;
        f += z * zs.n;
        df += z * zs.n;
        nit += z * zs.n;
        done += z * zs.n;
        active += z * zs.i;
    }
    __shared__ double sh[RB / 64];
    double cnt = 0.0;
    for (int p = threadIdx.x; p < n; p += blockDim.x) {
        if (done[p]) continue;
        double s = 0.0, d = 0.0;
        for (int c = 0; c < nchunk; ++c) {
            const double* o = part + 3LL * ((long long)p * nchunk + c);
            s += o[0] + o[1];
            d += o[2];
        }
        int it = nit[p] + 1;
        double wp = w[p];
        double fp = wp - eHF[p] - s;
        double dfp = 1.0 / (1.0 - d);
        w[p] = wp - dfp * fp;
        f[p] = fp;
        df[p] = dfp;
        nit[p] = it;
        if (!(fabs(fp) > thresh && it < maxit)) done[p] = 1;
        else cnt += 1.0;
    }
    double r = block_sum(cnt, sh);
    if (threadIdx.x == 0) active[0] = (int)r;
}

// eGW(p) = w, Z(p) = df unless nIt == maxIt (then eGW = eGWlin, Z kept)
__global__ void k_qp_finish(const double* __restrict__ w, const double* __restrict__ df, const int* __restrict__ nit,
                            const double* __restrict__ eGWlin, const double* __restrict__ eG,
                            double* __restrict__ eGW, double* __restrict__ Z, int n, int nC, int nR, int maxit, Zs zs) {
    {
        const long long z = blockIdx.z;
        w += z * zs.n;
        df += z * zs.n;
        nit += z * zs.n;
        eGWlin += z * zs.n;
        eG += z * zs.n;
        eGW += z * zs.n;
        Z += z * zs.n;
    }
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    if (p < nC || p >= n - nR) {
        // GW/RGW_QP_graph.f90:50 never assigns frozen entries of eGW, so they
        // keep what the caller's array held: in GW/evRGW.f90:154 that is
        // eOld, the energies in G.  Z keeps its input value there.
        eGW[p] = eG[p];
        return;
    }
    if (nit[p] == maxit) {
        eGW[p] = eGWlin[p];
    } else {
        eGW[p] = w[p];
        Z[p] = df[p];
    }
}

// ---- evGW bookkeeping on the device (batched evGW cycles) ------------------

constexpr int DIIS_MAX = 32;  // history length the device DIIS supports

// Bordered DIIS system of order N <= DIIS_MAX + 1 (column-major in A), solved
// by LU with partial pivoting; returns the exact 1-norm reciprocal condition
// number from the explicit inverse (what LAPACK's dsysvx estimates) and the
// solution in x, 0 when the matrix is singular.  One thread; LU, piv and col
// are that thread's work arrays.
__device__ double diis_solve(int N, const double* A, const double* b, double* x, double* LU, int* piv, double* col) {
    double anorm = 0.0;
    for (int j = 0; j < N; ++j) {
        double sj = 0.0;
        for (int i = 0; i < N; ++i) sj += fabs(A[i + N * j]);
        anorm = fmax(anorm, sj);
    }
    for (int t = 0; t < N * N; ++t) LU[t] = A[t];
    for (int k = 0; k < N; ++k) {
        int pk = k;
        for (int i = k + 1; i < N; ++i)
            if (fabs(LU[i + N * k]) > fabs(LU[pk + N * k])) pk = i;
        piv[k] = pk;
        if (LU[pk + N * k] == 0.0) {
            for (int i = 0; i < N; ++i) x[i] = 0.0;
            return 0.0;
        }
        if (pk != k)
            for (int j = 0; j < N; ++j) {
                const double t = LU[k + N * j];
                LU[k + N * j] = LU[pk + N * j];
                LU[pk + N * j] = t;
            }
        for (int i = k + 1; i < N; ++i) {
            LU[i + N * k] /= LU[k + N * k];
            for (int j = k + 1; j < N; ++j) LU[i + N * j] -= LU[i + N * k] * LU[k + N * j];
        }
    }
    auto lusolve = [&](double* r) {
        for (int k = 0; k < N; ++k) {
            const double t = r[k];
            r[k] = r[piv[k]];
            r[piv[k]] = t;
        }
        for (int i = 0; i < N; ++i)
            for (int j = 0; j < i; ++j) r[i] -= LU[i + N * j] * r[j];
        for (int i = N - 1; i >= 0; --i) {
            for (int j = i + 1; j < N; ++j) r[i] -= LU[i + N * j] * r[j];
            r[i] /= LU[i + N * i];
        }
    };
    double inorm = 0.0;
    for (int j = 0; j < N; ++j) {
        for (int i = 0; i < N; ++i) col[i] = i == j ? 1.0 : 0.0;
        lusolve(col);
        double sj = 0.0;
        for (int i = 0; i < N; ++i) sj += fabs(col[i]);
        inorm = fmax(inorm, sj);
    }
    for (int i = 0; i < N; ++i) x[i] = b[i];
    lusolve(x);
    if (!(anorm > 0.0) || !(inorm > 0.0) || !isfinite(inorm)) return 0.0;
    return 1.0 / (anorm * inorm);
}

// One evGW cycle's bookkeeping for molecule blockIdx.x (GW/evRGW.f90:156-183),
// so that a whole cycle can run without the host: for a live molecule the
// cycle's outputs are kept (Sig, Z, Om; EcRPA and EcGM from scal[1], scal[2]),
// Conv = max |eNew - e|, DIIS when maxd > 1 (utils/DIIS_extrapolation.f90:
// history prepended, bordered system, extrapolation only when rcond > 1e-13),
// e <- eNew (extrapolated), nscf += 1, live <- Conv > thresh and nscf <= maxSCF.
// A finished molecule is left as it is.  hist holds per molecule the error and
// energy histories (n x maxd each).
__global__ __launch_bounds__(RB) void k_evgw_update(const double* __restrict__ eNew, double* __restrict__ e,
                                                    const double* __restrict__ Sig, const double* __restrict__ Z,
                                                    const double* __restrict__ Om, const double* __restrict__ scal,
                                                    double* __restrict__ hist, int* __restrict__ state,
                                                    double* __restrict__ conv, double* __restrict__ outSig,
                                                    double* __restrict__ outZ, double* __restrict__ outOm,
                                                    double* __restrict__ outEc, int n, int nS, int maxd,
                                                    double thresh, int maxSCF, Zs zs) {
    const long long m = blockIdx.x;
    int* st = state + 4 * m;  // [0] live, [1] nscf, [2] history length
    if (!st[0]) return;
    eNew += m * zs.n;
    e += m * zs.n;
    Sig += m * zs.n;
    Z += m * zs.n;
    Om += m * zs.S;
    scal += m * zs.c;
    outSig += m * zs.n;
    outZ += m * zs.n;
    outOm += m * zs.S;
    double* errH = hist + m * 2LL * n * maxd;
    double* eH = errH + (long long)n * maxd;
    __shared__ double sh[RB / 64];
    __shared__ double A[(DIIS_MAX + 1) * (DIIS_MAX + 1)], LU[(DIIS_MAX + 1) * (DIIS_MAX + 1)];
    __shared__ double rhs[DIIS_MAX + 1], x[DIIS_MAX + 1], col[DIIS_MAX + 1];
    __shared__ int piv[DIIS_MAX + 1];
    __shared__ int use_x;
    for (int p = threadIdx.x; p < n; p += blockDim.x) {
        outSig[p] = Sig[p];
        outZ[p] = Z[p];
    }
    for (int k = threadIdx.x; k < nS; k += blockDim.x) outOm[k] = Om[k];
    if (threadIdx.x == 0) {
        outEc[2 * m] = scal[1];
        outEc[2 * m + 1] = scal[2];
    }
    double cmax = 0.0;
    for (int p = threadIdx.x; p < n; p += blockDim.x) cmax = fmax(cmax, fabs(eNew[p] - e[p]));
    // block maximum through the sum reduction's layout: |x| >= 0, so reduce max by shuffles
    for (int off = 32; off > 0; off >>= 1) cmax = fmax(cmax, __shfl_down(cmax, off, 64));
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = cmax;
    __syncthreads();
    double cv = 0.0;
    for (int k = 0; k < (int)(blockDim.x + 63) / 64; ++k) cv = fmax(cv, sh[k]);
    __syncthreads();
    if (maxd > 1) {
        const int md = min(st[2] + 1, maxd);
        for (int p = threadIdx.x; p < n; p += blockDim.x) {
            for (int j = md - 2; j >= 0; --j) {
                errH[p + (long long)n * (j + 1)] = errH[p + (long long)n * j];
                eH[p + (long long)n * (j + 1)] = eH[p + (long long)n * j];
            }
            errH[p] = eNew[p] - e[p];
            eH[p] = eNew[p];
        }
        const int N = md + 1;
        for (int i = 0; i < md; ++i)
            for (int j = i; j < md; ++j) {
                double d = 0.0;
                for (int p = threadIdx.x; p < n; p += blockDim.x)
                    d += errH[p + (long long)n * i] * errH[p + (long long)n * j];
                const double r = block_sum(d, sh);
                if (threadIdx.x == 0) {
                    A[i + N * j] = r;
                    A[j + N * i] = r;
                }
            }
        if (threadIdx.x == 0) {
            for (int i = 0; i < md; ++i) {
                A[i + N * md] = -1.0;
                A[md + N * i] = -1.0;
                rhs[i] = 0.0;
            }
            A[md + N * md] = 0.0;
            rhs[md] = -1.0;
            const double rc = diis_solve(N, A, rhs, x, LU, piv, col);
            use_x = rc > 1e-13;
            st[2] = md;
        }
        __syncthreads();
        for (int p = threadIdx.x; p < n; p += blockDim.x) {
            double v = eNew[p];
            if (use_x) {
                v = 0.0;
                for (int j = 0; j < md; ++j) v += x[j] * eH[p + (long long)n * j];
            }
            e[p] = v;
        }
    } else {
        for (int p = threadIdx.x; p < n; p += blockDim.x) e[p] = eNew[p];
    }
    if (threadIdx.x == 0) {
        conv[m] = cv;
        st[1] += 1;
        st[0] = (cv > thresh && st[1] <= maxSCF) ? 1 : 0;
    }
}

// ---- layout helpers -------------------------------------------------------

// out[j + ld_out*i] = in[i + ld_in*j], tiled through LDS (rows x cols of `in`)
__global__ void k_transpose(const double* __restrict__ in, double* __restrict__ out, int rows, int cols,
                            long long ld_in, long long ld_out, long long bs_in, long long bs_out) {
    __shared__ double t[32][33];
    in += blockIdx.z * bs_in;
    out += blockIdx.z * bs_out;
    int i0 = blockIdx.x * 32, j0 = blockIdx.y * 32;
    for (int k = threadIdx.y; k < 32; k += blockDim.y) {
        int i = i0 + threadIdx.x, j = j0 + k;
        if (i < rows && j < cols) t[k][threadIdx.x] = in[i + ld_in * j];
    }
    __syncthreads();
    for (int k = threadIdx.y; k < 32; k += blockDim.y) {
        int j = j0 + threadIdx.x, i = i0 + k;
        if (i < rows && j < cols) out[j + ld_out * i] = t[threadIdx.x][k];
    }
}

// general phRLR helpers
__global__ void k_axpby2(const double* A, const double* B, double* ApB, double* AmB, long long total) {
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total; t += (long long)gridDim.x * blockDim.x) {
        ApB[t] = A[t] + B[t];
        AmB[t] = A[t] - B[t];
    }
}
// U[:, k] *= d[k]^pw  (pw = +1/2 or -1/2), written to out
__global__ void k_colscale_pow(const double* U, const double* d, double* out, int n, double pw) {
    long long total = (long long)n * n;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total; t += (long long)gridDim.x * blockDim.x) {
        int k = (int)(t / n);
        out[t] = U[t] * (pw > 0 ? sqrt(d[k]) : 1.0 / sqrt(d[k]));
    }
}
// X[m + n ia] *= s[m]^pw with s = Om
__global__ void k_rowscale_pow(double* X, const double* s, int n, double pw) {
    long long total = (long long)n * n;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total; t += (long long)gridDim.x * blockDim.x) {
        int m = (int)(t % n);
        X[t] *= (pw > 0 ? sqrt(s[m]) : 1.0 / sqrt(s[m]));
    }
}
__global__ void k_trace(const double* A, int n, double* out) {
    __shared__ double sh[RB / 64];
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) s += A[(long long)i * n + i];
    double r = block_sum(s, sh);
    if (threadIdx.x == 0) out[0] = r;
}

// phRLR_A / phRLR_B (singlet) from a full ERI(n^4):
// A(ia,jb) = 2 lam ERI(b,i,j,a) - (1-dRPA) lam ERI(b,j,a,i) + delta (e_a - e_i)
// B(ia,jb) = 2 lam ERI(b,i,j,a) - (1-dRPA) lam ERI(b,j,i,a)
__global__ void k_phrlr_ab(const double* __restrict__ G, const double* __restrict__ e, double* __restrict__ out, int n,
                           int nC, int nO, int v, int nS, double lam, int dRPA, int isB) {
    long long total = (long long)nS * nS;
    long long n1 = n, n2 = n1 * n, n3 = n2 * n;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total; t += (long long)gridDim.x * blockDim.x) {
        long long ia = t % nS, jb = t / nS;
        long long a = nO + ia % v, i = nC + ia / v, b = nO + jb % v, j = nC + jb / v;
        double x = 2.0 * lam * G[b + n1 * i + n2 * j + n3 * a];
        if (!dRPA) x -= lam * (isB ? G[b + n1 * j + n2 * i + n3 * a] : G[b + n1 * j + n2 * a + n3 * i]);
        if (!isB && ia == jb) x += e[a] - e[i];
        out[t] = x;
    }
}

// ---- eigenvector refinement ----------------------------------------------
// For symmetric C and near-eigenvectors X (columns): S = X^T C X, G = X^T X.

// Rayleigh quotients lam_i = S_ii / G_ii
__global__ void k_eig_lambda(const double* __restrict__ S, const double* __restrict__ G, double* __restrict__ lam,
                             int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) lam[i] = S[(long long)i * n + i] / G[(long long)i * n + i];
}

// One Newton step of the eigenvector equations (Ogita & Aishima, Japan J.
// Indust. Appl. Math. 35 (2018) 1007-1035), with R = I - G and clusters of
// (nearly) equal eigenvalues (lab[i] = cluster label; inside a cluster the
// vectors have been rotated so that the cluster block of S is diagonal):
//   E_ii = R_ii / 2;  E_ij = R_ij / 2 (same cluster);
//   E_ij = (S_ij + lam_j R_ij) / (lam_j - lam_i) otherwise.
// S is overwritten with F = I + E (so that X <- X F).
__global__ void k_eig_update(double* __restrict__ S, const double* __restrict__ G, const double* __restrict__ lam,
                             const int* __restrict__ lab, int n) {
    long long total = (long long)n * n;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total; t += (long long)gridDim.x * blockDim.x) {
        int i = (int)(t % n), j = (int)(t / n);
        double g = G[t], f;
        if (i == j) {
            f = 1.0 + 0.5 * (1.0 - g);
        } else if (lab[i] == lab[j]) {
            f = -0.5 * g;
        } else {
            double lj = lam[j];
            f = (S[t] - lj * g) / (lj - lam[i]);
            if (!(fabs(f) <= 1e300)) f = 0.0;  // excluded by the check pass before any update
        }
        S[t] = f;
    }
}

// Check pass, one block per column j: colmax[j] = max |E_ij| over i outside
// j's cluster (non-finite -> 1e300), colarg[j] = that i.
__global__ __launch_bounds__(RB) void k_eig_check(const double* __restrict__ S, const double* __restrict__ G,
                                                  const double* __restrict__ lam, const int* __restrict__ lab, int n,
                                                  double* __restrict__ colmax, int* __restrict__ colarg) {
    __shared__ double shv[RB];
    __shared__ int shi[RB];
    const int j = blockIdx.x;
    const double lj = lam[j];
    const int labj = lab[j];
    const double* Sj = S + (long long)j * n;
    const double* Gj = G + (long long)j * n;
    double best = 0.0;
    int arg = j;
    for (int i = threadIdx.x; i < n; i += RB) {
        if (i == j || lab[i] == labj) continue;
        double a = fabs((Sj[i] - lj * Gj[i]) / (lj - lam[i]));
        if (!(a <= 1e300)) a = 1e300;
        if (a > best) {
            best = a;
            arg = i;
        }
    }
    shv[threadIdx.x] = best;
    shi[threadIdx.x] = arg;
    __syncthreads();
    for (int off = RB / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off && shv[threadIdx.x + off] > shv[threadIdx.x]) {
            shv[threadIdx.x] = shv[threadIdx.x + off];
            shi[threadIdx.x] = shi[threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        colmax[j] = shv[0];
        colarg[j] = shi[0];
    }
}

// Diagonal blocks of clusters: for cluster c (members idx[ptr[c] .. ptr[c+1]),
// size k), outA[boff[c] + r + k q] = A[idx[r] + lda idx[q]], same for B.
__global__ void k_gather_blocks(const double* __restrict__ A, const double* __restrict__ B, long long lda,
                                const int* __restrict__ ptr, const int* __restrict__ idx,
                                const long long* __restrict__ boff, double* __restrict__ outA,
                                double* __restrict__ outB) {
    const int c = blockIdx.x;
    const int p0 = ptr[c], k = ptr[c + 1] - p0;
    for (int t = threadIdx.x; t < k * k; t += blockDim.x) {
        int r = t % k, q = t / k;
        long long src = idx[p0 + r] + lda * idx[p0 + q];
        outA[boff[c] + t] = A[src];
        outB[boff[c] + t] = B[src];
    }
}

// Column mixing inside clusters: member slot s (cluster c = scl[s], local
// column r = s - ptr[c], size k) gets
//   tmp[i + nrows s] = sum_q A[i + lda cidx[ptr[c] + q]] Q[boff[c] + q + k r].
__global__ void k_mix_cols(const double* __restrict__ A, long long lda, int nrows, const int* __restrict__ cidx,
                           const int* __restrict__ ptr, const int* __restrict__ scl, const double* __restrict__ Q,
                           const long long* __restrict__ boff, int m, double* __restrict__ tmp) {
    long long total = (long long)nrows * m;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total; t += (long long)gridDim.x * blockDim.x) {
        int i = (int)(t % nrows), sl = (int)(t / nrows);
        int c = scl[sl], p0 = ptr[c], k = ptr[c + 1] - p0, r = sl - p0;
        const double* q = Q + boff[c] + (long long)k * r;
        double acc = 0.0;
        for (int u = 0; u < k; ++u) acc += A[i + lda * cidx[p0 + u]] * q[u];
        tmp[t] = acc;
    }
}
// A[i + lda idx[s]] = tmp[i + nrows s]
__global__ void k_scatter_cols(double* __restrict__ A, long long lda, int nrows, const int* __restrict__ idx, int m,
                               const double* __restrict__ tmp) {
    long long total = (long long)nrows * m;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total; t += (long long)gridDim.x * blockDim.x) {
        int i = (int)(t % nrows), sl = (int)(t / nrows);
        A[i + lda * idx[sl]] = tmp[t];
    }
}
// Row mixing: tmp[s + m j] = sum_q Q[boff[c] + q + k r] A[idx[ptr[c] + q] + n j]
__global__ void k_mix_rows(const double* __restrict__ A, int n, const int* __restrict__ idx,
                           const int* __restrict__ ptr, const int* __restrict__ scl, const double* __restrict__ Q,
                           const long long* __restrict__ boff, int m, double* __restrict__ tmp) {
    long long total = (long long)n * m;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total; t += (long long)gridDim.x * blockDim.x) {
        int sl = (int)(t % m), j = (int)(t / m);
        int c = scl[sl], p0 = ptr[c], k = ptr[c + 1] - p0, r = sl - p0;
        const double* q = Q + boff[c] + (long long)k * r;
        const double* Aj = A + (long long)n * j;
        double acc = 0.0;
        for (int u = 0; u < k; ++u) acc += q[u] * Aj[idx[p0 + u]];
        tmp[t] = acc;
    }
}
// A[idx[s] + n j] = tmp[s + m j]
__global__ void k_scatter_rows(double* __restrict__ A, int n, const int* __restrict__ idx, int m,
                               const double* __restrict__ tmp) {
    long long total = (long long)n * m;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total; t += (long long)gridDim.x * blockDim.x) {
        int sl = (int)(t % m), j = (int)(t / m);
        A[idx[sl] + (long long)n * j] = tmp[t];
    }
}
// out[i + n s] = X[i + n idx[s]]
__global__ void k_gather_cols(const double* __restrict__ X, int n, const int* __restrict__ idx, int m,
                              double* __restrict__ out) {
    long long total = (long long)n * m;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total; t += (long long)gridDim.x * blockDim.x) {
        int i = (int)(t % n), sl = (int)(t / n);
        out[t] = X[i + (long long)n * idx[sl]];
    }
}

// A[i + n j] = A[j + n i] for i > j (lower triangle from the upper), 32 x 32
// tiles through LDS; grid (tiles, tiles), only tiles with bi >= bj work.
__global__ void k_sym_lower(double* __restrict__ A, int n) {
    __shared__ double t[32][33];
    const int bi = blockIdx.x, bj = blockIdx.y;
    if (bi < bj) return;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8 threads
    // source tile: rows bj*32.., columns bi*32.. (upper part)
    for (int r = ty; r < 32; r += 8) {
        int row = bj * 32 + tx, col = bi * 32 + r;
        t[r][tx] = (row < n && col < n) ? A[row + (long long)n * col] : 0.0;
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        int row = bi * 32 + tx, col = bj * 32 + r;  // destination (row, col) = source (col, row)
        if (row < n && col < n && row > col) A[row + (long long)n * col] = t[tx][r];
    }
}

// Bov[P + naux ia] = B[P + naux (i + n a)], i = nC + ia / v, a = nO + ia % v
__global__ void k_gather_bov(const double* __restrict__ B, double* __restrict__ Bov, int naux, int n, int nC, int nO,
                             int v, int nS) {
    long long total = (long long)naux * nS;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total; t += (long long)gridDim.x * blockDim.x) {
        long long P = t % naux, ia = t / naux;
        long long i = nC + ia / v, a = nO + ia % v;
        Bov[t] = B[P + naux * (i + (long long)n * a)];
    }
}

// out[:, k] = X[:, perm[k]]
__global__ void k_permute_cols(const double* __restrict__ X, const int* __restrict__ perm, double* __restrict__ out,
                               int n) {
    long long total = (long long)n * n;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total; t += (long long)gridDim.x * blockDim.x) {
        int r = (int)(t % n), k = (int)(t / n);
        out[t] = X[r + (long long)n * perm[k]];
    }
}

inline int grid_for(long long total, int bs = 256) {
    long long g = (total + bs - 1) / bs;
    return (int)std::min<long long>(g, 65536);
}

}  // namespace

// ---- host wrappers --------------------------------------------------------

// Full AO array in QuAcK's ERI.bin order, f[mu + N la + N^2 nu + N^3 si] =
// (mu nu|la si), from the 8-fold packed chemists' array of int/ERI_chem.bin
// (utils/read_2e_integrals.f90:85-107): element (ab|cd) sits at
// pair(pair(a, b), pair(c, d)), pair(x, y) = x (x + 1) / 2 + y for x >= y
// (0-based; the 1-based Yoshimine index of utils/utils.f90:1186-1217 minus one).
__global__ void k_unpack_s8(const double* __restrict__ packed, double* __restrict__ f, int N) {
    const long long total = (long long)N * N * N * N;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total;
         t += (long long)gridDim.x * blockDim.x) {
        const long long mu = t % N, la = (t / N) % N, nu = (t / ((long long)N * N)) % N,
                        si = t / ((long long)N * N * N);
        const long long ij = mu >= nu ? mu * (mu + 1) / 2 + nu : nu * (nu + 1) / 2 + mu;
        const long long kl = la >= si ? la * (la + 1) / 2 + si : si * (si + 1) / 2 + la;
        f[t] = packed[ij >= kl ? ij * (ij + 1) / 2 + kl : kl * (kl + 1) / 2 + ij];
    }
}

void unpack_eri_s8(hipStream_t s, const double* packed, double* f, int N) {
    hipLaunchKernelGGL(k_unpack_s8, dim3(grid_for((long long)N * N * N * N)), dim3(256), 0, s, packed, f, N);
    QK_HIP(hipGetLastError());
}

void permute_block(hipStream_t s, const double* X, double* M, int n, int o, int v) {
    hipLaunchKernelGGL(k_permute_block, dim3(grid_for((long long)n * n * o * v)), dim3(256), 0, s, X, M, n, o, v);
    QK_HIP(hipGetLastError());
}
void extract_k(hipStream_t s, const double* M, double* K, int n, int nC, int nO, int v, int nS) {
    hipLaunchKernelGGL(k_extract_k, dim3(grid_for((long long)nS * nS)), dim3(256), 0, s, M, K, n, nC, nO, v, nS);
    QK_HIP(hipGetLastError());
}
void rpa_build(hipStream_t s, const double* e, const double* K, double* dE, double* sd, double* Cm, double* scal,
               int nC, int nO, int v, int nS, int tda, const Zs& zs, int nz) {
    hipLaunchKernelGGL(k_delta, dim3((nS + 255) / 256, 1, nz), dim3(256), 0, s, e, dE, sd, nC, nO, v, nS, zs);
    rpa_build_c(s, K, dE, sd, Cm, scal, nS, tda, zs, nz);
}
void rpa_build_c(hipStream_t s, const double* K, const double* dE, const double* sd, double* Cm, double* scal, int nS,
                 int tda, const Zs& zs, int nz) {
    hipLaunchKernelGGL(k_build_c, dim3(grid_for((long long)nS * nS), 1, nz), dim3(256), 0, s, K, dE, sd, Cm, nS, tda,
                       zs);
    hipLaunchKernelGGL(k_trace_a, dim3(1, 1, nz), dim3(RB), 0, s, K, dE, nS, scal, zs);
    QK_HIP(hipGetLastError());
}
void extract_ku(hipStream_t s, const double* M0, const double* M1, double* K, const UPairs& up) {
    hipLaunchKernelGGL(k_extract_ku, dim3(grid_for((long long)up.nSt * up.nSt)), dim3(256), 0, s, M0, M1, K, up);
    QK_HIP(hipGetLastError());
}
void delta_u(hipStream_t s, const double* e, double* dE, double* sd, const UPairs& up) {
    hipLaunchKernelGGL(k_delta_u, dim3((up.nSt + 255) / 256), dim3(256), 0, s, e, dE, sd, up);
    QK_HIP(hipGetLastError());
}
void rpa_finalize(hipStream_t s, double* Om, double* W, const double* sd, double* scal, int nS, int tda, bool scale_w,
                  const Zs& zs, int nz) {
    hipLaunchKernelGGL(k_finalize_om, dim3(1, 1, nz), dim3(RB), 0, s, Om, nS, tda, scal, zs);
    if (scale_w && !tda)
        hipLaunchKernelGGL(k_scale_w, dim3(grid_for((long long)nS * nS), 1, nz), dim3(256), 0, s, W, sd, Om, nS, zs);
    QK_HIP(hipGetLastError());
}
int sigma_chunks(int nS, int nact) {
    (void)nact;
    return (nS + RB * SIG_MPT - 1) / (RB * SIG_MPT);
}
void sigma_partial(hipStream_t s, const void* rhoT, bool rho_f32, const double* w, const double* e, const double* Om,
                   const int* skip, double eta, int srg, double flow, int n, int nC, int nO, int nR, int nS,
                   int nchunk, double* part, const Zs& zs, int nz) {
    int nact = n - nR - nC;
    QK_CHECK((long long)nchunk * RB * SIG_MPT >= nS, QK_E_ARG, "sigma_partial: too few chunks");
    if (rho_f32)
        hipLaunchKernelGGL(k_sigma_partial<float>, dim3(nchunk, nact, nz), dim3(RB), 0, s,
                           static_cast<const float*>(rhoT), w, e, Om, skip, eta, srg, flow, n, nC, nO, nR, nS, part,
                           zs);
    else
        hipLaunchKernelGGL(k_sigma_partial<double>, dim3(nchunk, nact, nz), dim3(RB), 0, s,
                           static_cast<const double*>(rhoT), w, e, Om, skip, eta, srg, flow, n, nC, nO, nR, nS, part,
                           zs);
    QK_HIP(hipGetLastError());
}
void sigma_finalize(hipStream_t s, const double* part, int nchunk, int n, int nC, int nO, int nR, double* Sig,
                    double* Z, double* scal, const Zs& zs, int nz) {
    hipLaunchKernelGGL(k_sigma_finalize, dim3(1, 1, nz), dim3(RB), 0, s, part, nchunk, n, nC, nO, nR, Sig, Z, scal, zs);
    QK_HIP(hipGetLastError());
}
void sigma_out(hipStream_t s, const double* part, int nchunk, int n, int nC, int nR, double* Sig, double* dSig) {
    hipLaunchKernelGGL(k_sigma_out, dim3((n + 255) / 256), dim3(256), 0, s, part, nchunk, n, nC, nR, Sig, dSig);
    QK_HIP(hipGetLastError());
}
// out[m + nS (p' + P q')] = w_m rho(p0 + p', q0 + q', m), w_m = Om_m / (Om_m^2 + eta^2)
// when weighted, else 1; rho(p, q, m) = rhoT[m + nS (p + n q)] (FP64).  The
// (occ, occ), (virt, virt), (occ, virt) and (virt, occ) blocks that the static
// BSE kernels contract over m (GW/RGW_phBSE_static_kernel_A.f90, _B.f90).
__global__ void k_rho_block(const double* __restrict__ rhoT, const double* __restrict__ Om, double eta, int weighted,
                            int n, int nS, int p0, int P, int q0, int Q, double* __restrict__ out) {
    const long long total = (long long)nS * P * Q;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total;
         t += (long long)gridDim.x * blockDim.x) {
        const int m = (int)(t % nS);
        const long long pq = t / nS;
        const int pp = (int)(pq % P), qq = (int)(pq / P);
        double v = rhoT[m + (long long)nS * ((p0 + pp) + (long long)n * (q0 + qq))];
        if (weighted) {
            const double om = Om[m];
            v *= om / (om * om + eta * eta);
        }
        out[t] = v;
    }
}

void rho_block(hipStream_t s, const double* rhoT, const double* Om, double eta, int weighted, int n, int nS, int p0,
               int P, int q0, int Q, double* out) {
    const long long total = (long long)nS * P * Q;
    hipLaunchKernelGGL(k_rho_block, dim3((unsigned)std::min<long long>((total + 255) / 256, 65536)), dim3(256), 0, s,
                       rhoT, Om, eta, weighted, n, nS, p0, P, q0, Q, out);
    QK_HIP(hipGetLastError());
}

// The full correlation self-energy of qsGW (GW/RGW_self_energy.f90,
// GW/RGW_SRG_self_energy.f90): Sig(p,q) = 2 sum_{r,m} rho(p,r,m) rho(q,r,m) w,
// D_x = e_x - e_r + Om_m (r occupied) or e_x - e_r - Om_m (r virtual),
//   w = D_p / (D_p^2 + eta^2)                                      (plain)
//   w = (1 - exp(-s D_p^2) exp(-s D_q^2)) (D_p + D_q) / (D_p^2 + D_q^2)  (SRG).
// One thread per (p, q) of a 16 x 16 tile of the active orbitals; the tile's
// rho(p, r, m) and rho(q, r, m) rows come through LDS in chunks of 64 modes
// (m contiguous in rhoT, so every row load is coalesced).
template <bool SRG>
__global__ __launch_bounds__(256) void k_sigma_full(const double* __restrict__ rhoT, const double* __restrict__ e,
                                                    const double* __restrict__ Om, double eta, double flow, int n,
                                                    int nC, int nO, int nR, int nS, double* __restrict__ Sig) {
    constexpr int TT = 16, MC = 64;
    __shared__ double sp[TT][MC + 1], sq[TT][MC + 1], om[MC];
    const int hi = n - nR;  // active orbitals nC .. hi-1
    const int tx = threadIdx.x % TT, ty = threadIdx.x / TT;
    const int p0 = nC + blockIdx.x * TT, q0 = nC + blockIdx.y * TT;
    const int p = p0 + tx, q = q0 + ty;
    const bool on = p < hi && q < hi;
    const double ep = p < hi ? e[p] : 0.0, eq = q < hi ? e[q] : 0.0;
    double acc = 0.0;
    for (int r = nC; r < hi; ++r) {
        const double er = e[r];
        const double sgn = r < nO ? 1.0 : -1.0;
        for (int m0 = 0; m0 < nS; m0 += MC) {
            for (int t = threadIdx.x; t < TT * MC; t += blockDim.x) {
                const int row = t / MC, mm = t % MC, m = m0 + mm;
                const int pr = p0 + row, qr = q0 + row;
                sp[row][mm] = (m < nS && pr < hi) ? rhoT[m + (long long)nS * (pr + (long long)n * r)] : 0.0;
                sq[row][mm] = (m < nS && qr < hi) ? rhoT[m + (long long)nS * (qr + (long long)n * r)] : 0.0;
            }
            if (threadIdx.x < MC) om[threadIdx.x] = m0 + (int)threadIdx.x < nS ? Om[m0 + threadIdx.x] : 0.0;
            __syncthreads();
            const int mmax = min(MC, nS - m0);
            for (int mm = 0; mm < mmax; ++mm) {
                const double o = sgn * om[mm];
                const double dp = ep - er + o;
                double w;
                if constexpr (SRG) {
                    const double dq = eq - er + o;
                    w = (1.0 - exp(-flow * dp * dp) * exp(-flow * dq * dq)) * (dp + dq) / (dp * dp + dq * dq);
                } else {
                    w = dp / (dp * dp + eta * eta);
                }
                acc += sp[tx][mm] * sq[ty][mm] * w;
            }
            __syncthreads();
        }
    }
    if (on) Sig[p + (long long)n * q] = 2.0 * acc;
}

// Plain weight, which depends on p, r and m only: X[m + nS (p + n r)] =
// 2 rho(p,r,m) D_p / (D_p^2 + eta^2) for active p and r, else 0; then
// Sig = sum_r X_r^T R_r (R_r the nS x n slab of rhoT at r) is a sum of
// matrix-core products.
__global__ void k_sigma_weight(const double* __restrict__ rhoT, const double* __restrict__ e,
                               const double* __restrict__ Om, double eta, int n, int nC, int nO, int nR, int nS,
                               double* __restrict__ X) {
    const long long total = (long long)nS * n * n;
    const int hi = n - nR;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total;
         t += (long long)gridDim.x * blockDim.x) {
        const int m = (int)(t % nS);
        const long long pr = t / nS;
        const int p = (int)(pr % n), r = (int)(pr / n);
        double v = 0.0;
        if (p >= nC && p < hi && r >= nC && r < hi) {
            const double d = e[p] - e[r] + (r < nO ? Om[m] : -Om[m]);
            v = 2.0 * rhoT[t] * d / (d * d + eta * eta);
        }
        X[t] = v;
    }
}

// Sig(nC + i, nC + j) = sum over the nact products of parts[k nn + i + nact j]
__global__ void k_sum_parts(const double* __restrict__ parts, int nact, long long nn, int n, int nC,
                            double* __restrict__ Sig) {
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < nn; t += (long long)gridDim.x * blockDim.x) {
        double v = 0.0;
        for (int k = 0; k < nact; ++k) v += parts[k * nn + t];
        const int i = (int)(t % nact), j = (int)(t / nact);
        Sig[(nC + i) + (long long)n * (nC + j)] = v;
    }
}

void sigma_full(hipStream_t s, const double* rhoT, const double* e, const double* Om, double eta, int srg, double flow,
                int n, int nC, int nO, int nR, int nS, double* Sig, double* work) {
    QK_HIP(hipMemsetAsync(Sig, 0, sizeof(double) * (size_t)n * n, s));
    const int nact = n - nR - nC;
    if (nact <= 0) return;
    if (!srg && work) {
        const long long total = (long long)nS * n * n;
        hipLaunchKernelGGL(k_sigma_weight, dim3((unsigned)std::min<long long>((total + 255) / 256, 65536)), dim3(256),
                           0, s, rhoT, e, Om, eta, n, nC, nO, nR, nS, work);
        QK_HIP(hipGetLastError());
        // the active block only (inactive rows and columns stay zero): one
        // product per active r, batched, then their sum in a fixed order
        const long long slab = (long long)nS * n, off = (long long)nS * nC;
        double* parts = work + slab * n;  // nact products of nact x nact
        Batch bt;
        bt.count = nact;
        bt.zdiv = nact;
        bt.sa0 = slab;
        bt.sb0 = slab;
        bt.sc0 = (long long)nact * nact;
        dgemm(s, true, false, nact, nact, nS, 1.0, work + slab * nC + off, nS, rhoT + slab * nC + off, nS, 0.0, parts,
              nact, bt);
        const long long nn = (long long)nact * nact;
        hipLaunchKernelGGL(k_sum_parts, dim3((unsigned)std::min<long long>((nn + 255) / 256, 4096)), dim3(256), 0, s,
                           parts, nact, nn, n, nC, Sig);
        QK_HIP(hipGetLastError());
        return;
    }
    dim3 g((nact + 15) / 16, (nact + 15) / 16);
    if (srg)
        hipLaunchKernelGGL(k_sigma_full<true>, g, dim3(256), 0, s, rhoT, e, Om, eta, flow, n, nC, nO, nR, nS, Sig);
    else
        hipLaunchKernelGGL(k_sigma_full<false>, g, dim3(256), 0, s, rhoT, e, Om, eta, flow, n, nC, nO, nR, nS, Sig);
    QK_HIP(hipGetLastError());
}

void lin_qp(hipStream_t s, const double* e0, const double* Z, const double* Sig, double* out, int n, int zflag,
            const Zs& zs, int nz) {
    hipLaunchKernelGGL(k_lin, dim3((n + 255) / 256, 1, nz), dim3(256), 0, s, e0, Z, Sig, out, n, zflag, zs);
    QK_HIP(hipGetLastError());
}
void qp_init(hipStream_t s, const double* eGWlin, double* w, double* f, int* nit, int* done, int n, int nC, int nR,
             const Zs& zs, int nz) {
    hipLaunchKernelGGL(k_qp_init, dim3((n + 255) / 256, 1, nz), dim3(256), 0, s, eGWlin, w, f, nit, done, n, nC, nR,
                       zs);
    QK_HIP(hipGetLastError());
}
void qp_update(hipStream_t s, const double* part, int nchunk, const double* eHF, double* w, double* f, double* df,
               int* nit, int* done, int n, double thresh, int maxit, int* active, const Zs& zs, int nz) {
    hipLaunchKernelGGL(k_qp_update, dim3(1, 1, nz), dim3(RB), 0, s, part, nchunk, eHF, w, f, df, nit, done, n, thresh,
                       maxit, active, zs);
    QK_HIP(hipGetLastError());
}
void qp_finish(hipStream_t s, const double* w, const double* df, const int* nit, const double* eGWlin,
               const double* eG, double* eGW, double* Z, int n, int nC, int nR, int maxit, const Zs& zs, int nz) {
    hipLaunchKernelGGL(k_qp_finish, dim3((n + 255) / 256, 1, nz), dim3(256), 0, s, w, df, nit, eGWlin, eG, eGW, Z, n,
                       nC, nR, maxit, zs);
    QK_HIP(hipGetLastError());
}
void evgw_update(hipStream_t s, const double* eNew, double* e, const double* Sig, const double* Z, const double* Om,
                 const double* scal, double* hist, int* state, double* conv, double* outSig, double* outZ,
                 double* outOm, double* outEc, int n, int nS, int maxd, double thresh, int maxSCF, const Zs& zs,
                 int nmol) {
    QK_CHECK(maxd >= 1 && maxd <= DIIS_MAX, QK_E_ARG, "evgw_update: DIIS history longer than 32");
    hipLaunchKernelGGL(k_evgw_update, dim3(nmol), dim3(RB), 0, s, eNew, e, Sig, Z, Om, scal, hist, state, conv, outSig,
                       outZ, outOm, outEc, n, nS, maxd, thresh, maxSCF, zs);
    QK_HIP(hipGetLastError());
}
void transpose(hipStream_t s, const double* in, double* out, int rows, int cols, long long ld_in, long long ld_out,
               int batch, long long bs_in, long long bs_out) {
    dim3 g((rows + 31) / 32, (cols + 31) / 32, batch);
    hipLaunchKernelGGL(k_transpose, g, dim3(32, 8), 0, s, in, out, rows, cols, ld_in, ld_out, bs_in, bs_out);
    QK_HIP(hipGetLastError());
}
void apb_amb(hipStream_t s, const double* A, const double* B, double* ApB, double* AmB, long long total) {
    hipLaunchKernelGGL(k_axpby2, dim3(grid_for(total)), dim3(256), 0, s, A, B, ApB, AmB, total);
    QK_HIP(hipGetLastError());
}
void colscale_pow(hipStream_t s, const double* U, const double* d, double* out, int n, double pw) {
    hipLaunchKernelGGL(k_colscale_pow, dim3(grid_for((long long)n * n)), dim3(256), 0, s, U, d, out, n, pw);
    QK_HIP(hipGetLastError());
}
void rowscale_pow(hipStream_t s, double* X, const double* sv, int n, double pw) {
    hipLaunchKernelGGL(k_rowscale_pow, dim3(grid_for((long long)n * n)), dim3(256), 0, s, X, sv, n, pw);
    QK_HIP(hipGetLastError());
}
void trace(hipStream_t s, const double* A, int n, double* out) {
    hipLaunchKernelGGL(k_trace, dim3(1), dim3(RB), 0, s, A, n, out);
    QK_HIP(hipGetLastError());
}
void phrlr_ab(hipStream_t s, const double* G, const double* e, double* out, int n, int nC, int nO, int v, int nS,
              double lam, int dRPA, int isB) {
    hipLaunchKernelGGL(k_phrlr_ab, dim3(grid_for((long long)nS * nS)), dim3(256), 0, s, G, e, out, n, nC, nO, v, nS,
                       lam, dRPA, isB);
    QK_HIP(hipGetLastError());
}

void eig_lambda(hipStream_t s, const double* S, const double* G, double* lam, int n) {
    hipLaunchKernelGGL(k_eig_lambda, dim3((n + 255) / 256), dim3(256), 0, s, S, G, lam, n);
    QK_HIP(hipGetLastError());
}
void eig_update(hipStream_t s, double* S, const double* G, const double* lam, const int* lab, int n) {
    hipLaunchKernelGGL(k_eig_update, dim3(grid_for((long long)n * n)), dim3(256), 0, s, S, G, lam, lab, n);
    QK_HIP(hipGetLastError());
}
void eig_check(hipStream_t s, const double* S, const double* G, const double* lam, const int* lab, int n,
               double* colmax, int* colarg) {
    hipLaunchKernelGGL(k_eig_check, dim3(n), dim3(RB), 0, s, S, G, lam, lab, n, colmax, colarg);
    QK_HIP(hipGetLastError());
}
void gather_blocks(hipStream_t s, const double* A, const double* B, long long lda, const int* ptr, const int* idx,
                   const long long* boff, int ncl, double* outA, double* outB) {
    if (ncl <= 0) return;
    hipLaunchKernelGGL(k_gather_blocks, dim3(ncl), dim3(256), 0, s, A, B, lda, ptr, idx, boff, outA, outB);
    QK_HIP(hipGetLastError());
}
void mix_cols(hipStream_t s, const double* A, long long lda, int nrows, const int* cidx, double* D, long long ldd,
              const int* idx, const int* ptr, const int* scl, const double* Q, const long long* boff, int m,
              double* tmp) {
    if (m <= 0) return;
    int g = grid_for((long long)nrows * m);
    hipLaunchKernelGGL(k_mix_cols, dim3(g), dim3(256), 0, s, A, lda, nrows, cidx, ptr, scl, Q, boff, m, tmp);
    hipLaunchKernelGGL(k_scatter_cols, dim3(g), dim3(256), 0, s, D, ldd, nrows, idx, m, tmp);
    QK_HIP(hipGetLastError());
}
void mix_rows(hipStream_t s, double* A, int n, const int* idx, const int* ptr, const int* scl, const double* Q,
              const long long* boff, int m, double* tmp) {
    if (m <= 0) return;
    int g = grid_for((long long)n * m);
    hipLaunchKernelGGL(k_mix_rows, dim3(g), dim3(256), 0, s, A, n, idx, ptr, scl, Q, boff, m, tmp);
    hipLaunchKernelGGL(k_scatter_rows, dim3(g), dim3(256), 0, s, A, n, idx, m, tmp);
    QK_HIP(hipGetLastError());
}
void gather_cols(hipStream_t s, const double* X, int n, const int* idx, int m, double* out) {
    if (m <= 0) return;
    hipLaunchKernelGGL(k_gather_cols, dim3(grid_for((long long)n * m)), dim3(256), 0, s, X, n, idx, m, out);
    QK_HIP(hipGetLastError());
}
void gather_bov(hipStream_t s, const double* B, double* Bov, int naux, int n, int nC, int nO, int v, int nS) {
    hipLaunchKernelGGL(k_gather_bov, dim3(grid_for((long long)naux * nS)), dim3(256), 0, s, B, Bov, naux, n, nC, nO, v,
                       nS);
    QK_HIP(hipGetLastError());
}
void sym_lower(hipStream_t s, double* A, int n) {
    int nt = (n + 31) / 32;
    hipLaunchKernelGGL(k_sym_lower, dim3(nt, nt), dim3(256), 0, s, A, n);
    QK_HIP(hipGetLastError());
}
void permute_cols(hipStream_t s, const double* X, const int* perm, double* out, int n) {
    hipLaunchKernelGGL(k_permute_cols, dim3(grid_for((long long)n * n)), dim3(256), 0, s, X, perm, out, n);
    QK_HIP(hipGetLastError());
}

}  // namespace qk
