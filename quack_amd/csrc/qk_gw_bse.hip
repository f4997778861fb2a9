// Kernels of the second-order BSE kernel and of the first-order dynamical
// correction of phBSE@GW (GW/RGW_phBSE.f90 with dophBSE2 / dBSE).  Pairs ia
// and jb run over i, j occupied and a, b virtual (nC = nR = 0: QuAcK's pair
// index (a-nO) + (i-1)(nBas-nO) of these routines), ia = (a - nO) + v i,
// column-major nS x nS matrices, one thread per element.
#include "qk_kernels.h"

namespace qk {
namespace {

constexpr int TB = 256;

__device__ __forceinline__ double fw(double x, double eta) { return x / (x * x + eta * eta); }
__device__ __forceinline__ double gw(double x, double eta) {
    const double d = x * x + eta * eta;
    return (x * x - eta * eta) / (d * d);
}

// W(p,q,r,s) of QuAcK's 4-index array (column-major, 0-based)
struct W4 {
    const double* w;
    long long n;
    __device__ __forceinline__ double operator()(int p, int q, int r, int s) const {
        return w[p + n * (q + n * (r + n * (long long)s))];
    }
};

// GW/RGW_phBSE2_static_kernel_A.f90:36-90 and _B.f90:36-90: the second-order
// static kernels (KB2 skipped when null), energies e = eW
__global__ void k_bse2_static(const double* __restrict__ Wp, const double* __restrict__ e, double eta, int n, int nO,
                              int nS, double* __restrict__ KA2, double* __restrict__ KB2) {
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= (long long)nS * nS) return;
    const int v = n - nO;
    const int ia = (int)(t % nS), jb = (int)(t / nS);
    const int i = ia / v, a = nO + ia % v, j = jb / v, b = nO + jb % v;
    const W4 W{Wp, n};
    double ka = 0.0, kb = 0.0;
    for (int k = 0; k < nO; ++k)
        for (int c = nO; c < n; ++c) {
            const double d1 = -(e[c] - e[k]), d2 = e[c] - e[k];
            ka -= 2.0 * W(j, k, i, c) * W(a, c, b, k) * fw(d1, eta);
            ka += 2.0 * W(j, c, i, k) * W(a, k, b, c) * fw(d2, eta);
            if (KB2) {
                kb -= 2.0 * W(b, k, i, c) * W(a, c, j, k) * fw(d1, eta);
                kb += 2.0 * W(b, c, i, k) * W(a, k, j, c) * fw(d2, eta);
            }
        }
    for (int c = nO; c < n; ++c)
        for (int d = nO; d < n; ++d) {
            const double dd = -(e[c] + e[d]);
            ka += 2.0 * W(a, j, c, d) * W(c, d, i, b) * fw(dd, eta);
            if (KB2) kb += 2.0 * W(a, b, c, d) * W(c, d, i, j) * fw(dd, eta);
        }
    for (int k = 0; k < nO; ++k)
        for (int l = 0; l < nO; ++l) {
            const double dd = -(e[k] + e[l]);
            ka -= 2.0 * W(a, j, k, l) * W(k, l, i, b) * fw(dd, eta);
            if (KB2) kb -= 2.0 * W(a, b, k, l) * W(k, l, i, j) * fw(dd, eta);
        }
    KA2[t] = ka;
    if (KB2) KB2[t] = kb;
}

// The six quadratic forms of GW/RGW_phBSE_dynamic_perturbation.f90:86-112 for
// one BSE excitation Om with vectors X, Y: per element (ia, jb) the dynamic
// kernels KA(+Om), ZA(+Om) (GW/RGW_phBSE_dynamic_kernel_A.f90:38-82, plus the
// second-order GW/RGW_phBSE2_dynamic_kernel_A.f90 when W is given) and, unless
// dTDA, KA(-Om), ZA(-Om) and KB (GW/RGW_phBSE_dynamic_kernel_B.f90:34-62, plus
// GW/RGW_phBSE2_dynamic_kernel_B.f90), contracted with X and Y; each block
// writes its partial sums of X KA X, X ZA X, Y KA(-) Y, Y ZA(-) Y, X KB Y and
// Y KB X to part[6 block + k].  rhoT holds rho(p, q, m) at m + nS (p + n q).
__global__ __launch_bounds__(TB) void k_bse_dyn(const double* __restrict__ rhoT, const double* __restrict__ OmR,
                                                 const double* __restrict__ e, const double* __restrict__ Wp,
                                                 const double* __restrict__ X, const double* __restrict__ Y,
                                                 double Om, double eta, int n, int nO, int nS, int dtda,
                                                 double* __restrict__ part) {
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    double q[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (t < (long long)nS * nS) {
        const int v = n - nO;
        const int ia = (int)(t % nS), jb = (int)(t / nS);
        const int i = ia / v, a = nO + ia % v, j = jb / v, b = nO + jb % v;
        const double* rij = rhoT + (long long)nS * (i + (long long)n * j);
        const double* rab = rhoT + (long long)nS * (a + (long long)n * b);
        const double daj = e[a] - e[j], dbi = e[b] - e[i];
        double kap = 0.0, zap = 0.0, kam = 0.0, zam = 0.0, kb = 0.0;
        for (int m = 0; m < nS; ++m) {
            const double r = rij[m] * rab[m];
            const double p1 = Om - OmR[m] - daj, p2 = Om - OmR[m] - dbi;
            kap += r * (fw(p1, eta) + fw(p2, eta));
            zap += r * (gw(p1, eta) + gw(p2, eta));
            if (!dtda) {
                const double m1 = -Om - OmR[m] - daj, m2 = -Om - OmR[m] - dbi;
                kam += r * (fw(m1, eta) + fw(m2, eta));
                zam += r * (gw(m1, eta) + gw(m2, eta));
            }
        }
        kap *= -2.0;
        zap *= 2.0;
        kam *= -2.0;
        zam *= 2.0;
        if (!dtda) {
            const double* rib = rhoT + (long long)nS * (i + (long long)n * b);
            const double* rja = rhoT + (long long)nS * (j + (long long)n * a);
            for (int m = 0; m < nS; ++m) {
                const double r = rib[m] * rja[m];
                kb += r * (fw(-OmR[m] - dbi, eta) + fw(-OmR[m] - daj, eta));
            }
            kb *= -2.0;
        }
        if (Wp) {
            const W4 W{Wp, n};
            for (int k = 0; k < nO; ++k)
                for (int c = nO; c < n; ++c) {
                    const double d1 = Om - e[a] - e[c] + e[k] + e[j];
                    const double n1 = 2.0 * W(j, k, i, c) * W(a, c, b, k);
                    kap -= n1 * fw(d1, eta);
                    zap += n1 * gw(d1, eta);
                    const double d2 = Om + e[i] - e[c] + e[k] - e[b];
                    const double n2 = 2.0 * W(j, c, i, k) * W(a, k, b, c);
                    kap -= n2 * fw(d2, eta);
                    zap += n2 * gw(d2, eta);
                    if (!dtda) {
                        kb -= 2.0 * W(b, k, i, c) * W(a, c, j, k) * fw(-e[a] - e[c] + e[k] + e[j], eta);
                        kb -= 2.0 * W(b, c, i, k) * W(a, k, j, c) * fw(e[i] - e[c] + e[k] - e[b], eta);
                    }
                }
            for (int c = nO; c < n; ++c)
                for (int d = nO; d < n; ++d) {
                    const double d3 = Om + e[i] + e[j] - e[c] - e[d];
                    const double n3 = 2.0 * W(a, j, c, d) * W(c, d, i, b);
                    kap += n3 * fw(d3, eta);
                    zap -= n3 * gw(d3, eta);
                    if (!dtda) kb += 2.0 * W(a, b, c, d) * W(c, d, i, j) * fw(e[i] + e[j] - e[c] - e[d], eta);
                }
            for (int k = 0; k < nO; ++k)
                for (int l = 0; l < nO; ++l) {
                    const double d4 = Om - e[a] - e[b] + e[k] + e[l];
                    const double n4 = 2.0 * W(a, j, k, l) * W(k, l, i, b);
                    kap += n4 * fw(d4, eta);
                    zap -= n4 * gw(d4, eta);
                    if (!dtda) kb += 2.0 * W(a, b, k, l) * W(k, l, i, j) * fw(-e[a] - e[b] + e[k] + e[l], eta);
                }
        }
        const double xx = X[ia] * X[jb], yy = Y[ia] * Y[jb];
        q[0] = xx * kap;
        q[1] = xx * zap;
        q[2] = yy * kam;
        q[3] = yy * zam;
        q[4] = X[ia] * Y[jb] * kb;
        q[5] = Y[ia] * X[jb] * kb;
    }
    // fixed-order block reduction
    __shared__ double red[6][TB];
    for (int k = 0; k < 6; ++k) red[k][threadIdx.x] = q[k];
    __syncthreads();
    for (int w = TB / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int k = 0; k < 6; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 6; ++k) part[6 * (long long)blockIdx.x + k] = red[k][0];
}

}  // namespace

void bse2_static(hipStream_t s, const double* W, const double* e, double eta, int n, int nO, int nS, double* KA2,
                 double* KB2) {
    const long long tot = (long long)nS * nS;
    hipLaunchKernelGGL(k_bse2_static, dim3((unsigned)((tot + TB - 1) / TB)), dim3(TB), 0, s, W, e, eta, n, nO, nS,
                       KA2, KB2);
    QK_HIP(hipGetLastError());
}

int bse_dyn_blocks(int nS) { return (int)(((long long)nS * nS + TB - 1) / TB); }

void bse_dyn(hipStream_t s, const double* rhoT, const double* OmR, const double* e, const double* W, const double* X,
             const double* Y, double Om, double eta, int n, int nO, int nS, int dtda, double* part) {
    hipLaunchKernelGGL(k_bse_dyn, dim3((unsigned)bse_dyn_blocks(nS)), dim3(TB), 0, s, rhoT, OmR, e, W, X, Y, Om, eta,
                       n, nO, nS, dtda, part);
    QK_HIP(hipGetLastError());
}

}  // namespace qk
