// Two-electron integrals over contracted s and p Gaussian shells on the
// device (McMurchie-Davidson: Hermite expansion coefficients E^{ij}_t and
// Hermite Coulomb integrals R_{tuv}; McMurchie & Davidson, J. Comput. Phys. 26,
// 218 (1978); Helgaker, Jorgensen, Olsen, "Molecular Electronic-Structure
// Theory", ch. 9).  This is the device counterpart of the host producer
// (quack_amd/intgen/intgen.c) for basis sets without d shells, such as
// cc-pVDZ hydrogen: it stands in for PySCF's mol.intor('int2e') of
// PyDuck.py:329-345 and writes the array QuAcK streams from int/ERI.bin
// (utils/read_2e_integrals.f90:55-63):
//   f[a + N b + N^2 c + N^3 d] = <ab|cd> = (ac|bd),
// p functions ordered x, y, z, every contracted function normalised to one
// (the host producer's conventions, so that both give the same array).
//
// Work layout: shell pairs (a >= b) are grouped by their angular momenta
// (la, lb); one launch per (bra class, ket class) runs one thread per
// (bra pair, ket pair), so every thread of a launch has the same loop and
// array shapes (compile-time) and no divergence.  A thread writes its
// quartet's integrals at the four positions that permute within the bra and
// within the ket; the bra-ket images are the (ket, bra) thread's.
#include <cmath>
#include <vector>

#include "qk_common.h"

namespace qk {
namespace {

constexpr int NPRIM = 8;  // primitives per shell (cc-pVDZ hydrogen: 4)

struct DShell {
    double A[3];
    double ex[NPRIM];
    double co[NPRIM];  // primitive- and contraction-normalised coefficients
    int l, np, off, pad;
};

// Boys function F_m(T), m = 0..M: power series of F_M below the switch point
// with downward recursion, erf plus upward recursion above it.
template <int M>
__device__ __forceinline__ void boys(double T, double* F) {
    if (T < 1e-14) {
#pragma unroll
        for (int m = 0; m <= M; ++m) F[m] = 1.0 / (2 * m + 1) - T / (2 * m + 3);
        return;
    }
    const double e = exp(-T);
    if (T > 2.0 * M + 25.0) {
        const double st = sqrt(T);
        F[0] = 0.8862269254527580 * erf(st) / st;
#pragma unroll
        for (int m = 0; m < M; ++m) F[m + 1] = ((2 * m + 1) * F[m] - e) / (2.0 * T);
        return;
    }
    double term = 1.0 / (2 * M + 1), sum = term;
    for (int k = 1; k < 400; ++k) {
        term *= 2.0 * T / (2 * M + 2 * k + 1);
        sum += term;
        if (term < 1e-17 * sum) break;
    }
    F[M] = e * sum;
#pragma unroll
    for (int m = M - 1; m >= 0; --m) F[m] = (2.0 * T * F[m + 1] + e) / (2 * m + 1);
}

// 1D Hermite coefficients E[i][j][t] (i <= LI, j <= LJ, t <= i + j), without
// the exp(-mu X^2) factor: E^{i+1,j}_t = E^{ij}_{t-1}/(2p) + X_PA E^{ij}_t + (t+1) E^{ij}_{t+1},
// and the same in j with X_PB.
template <int LI, int LJ>
__device__ __forceinline__ void hermite_1d(double p, double xpa, double xpb, double (&E)[LI + 1][LJ + 1][LI + LJ + 1]) {
    const double h = 0.5 / p;
#pragma unroll
    for (int i = 0; i <= LI; ++i)
#pragma unroll
        for (int j = 0; j <= LJ; ++j)
#pragma unroll
            for (int t = 0; t <= LI + LJ; ++t) E[i][j][t] = 0.0;
    E[0][0][0] = 1.0;
#pragma unroll
    for (int i = 0; i < LI; ++i)
#pragma unroll
        for (int t = 0; t <= i + 1; ++t) {
            double v = t <= i ? xpa * E[i][0][t] : 0.0;
            if (t > 0) v += h * E[i][0][t - 1];
            if (t + 1 <= i) v += (t + 1) * E[i][0][t + 1];
            E[i + 1][0][t] = v;
        }
#pragma unroll
    for (int i = 0; i <= LI; ++i)
#pragma unroll
        for (int j = 0; j < LJ; ++j)
#pragma unroll
            for (int t = 0; t <= i + j + 1; ++t) {
                double v = t <= i + j ? xpb * E[i][j][t] : 0.0;
                if (t > 0) v += h * E[i][j][t - 1];
                if (t + 1 <= i + j) v += (t + 1) * E[i][j][t + 1];
                E[i][j + 1][t] = v;
            }
}

// Hermite Coulomb integrals R_{tuv} = R^0_{tuv}, t+u+v <= L, from
// R^n_{000} = (-2 alpha)^n F_n(alpha |PQ|^2) and
// R^n_{t+1,u,v} = t R^{n+1}_{t-1,u,v} + X R^{n+1}_{t,u,v} (same in u with Y, v with Z).
template <int L>
__device__ __forceinline__ void hermite_R(double alpha, double X, double Y, double Z,
                                          double (&R0)[L + 1][L + 1][L + 1]) {
    double F[L + 1];
    boys<L>(alpha * (X * X + Y * Y + Z * Z), F);
    double R[L + 1][L + 1][L + 1][L + 1];  // [n][t][u][v]; entries with n+t+u+v <= L are used
    double f = 1.0;
#pragma unroll
    for (int n = 0; n <= L; ++n) {
        R[n][0][0][0] = f * F[n];
        f *= -2.0 * alpha;
    }
#pragma unroll
    for (int n = L - 1; n >= 0; --n)
#pragma unroll
        for (int s = 1; s <= L - n; ++s)
#pragma unroll
            for (int t = 0; t <= s; ++t)
#pragma unroll
                for (int u = 0; u <= s - t; ++u) {
                    const int v = s - t - u;
                    double val;
                    if (t > 0) {
                        val = X * R[n + 1][t - 1][u][v];
                        if (t > 1) val += (t - 1) * R[n + 1][t - 2][u][v];
                    } else if (u > 0) {
                        val = Y * R[n + 1][t][u - 1][v];
                        if (u > 1) val += (u - 1) * R[n + 1][t][u - 2][v];
                    } else {
                        val = Z * R[n + 1][t][u][v - 1];
                        if (v > 1) val += (v - 1) * R[n + 1][t][u][v - 2];
                    }
                    R[n][t][u][v] = val;
                }
#pragma unroll
    for (int t = 0; t <= L; ++t)
#pragma unroll
        for (int u = 0; u <= L - t; ++u)
#pragma unroll
            for (int v = 0; v <= L - t - u; ++v) R0[t][u][v] = R[0][t][u][v];
}

// cartesian exponent of component c of an l <= 1 shell along axis x
__device__ __forceinline__ constexpr int cexp(int l, int c, int x) { return l == 0 ? 0 : (c == x ? 1 : 0); }

template <int LA, int LB, int LC, int LD>
__global__ __launch_bounds__(256) void k_eri_class(const DShell* __restrict__ sh, const int2* __restrict__ bra,
                                                   int nbra, const int2* __restrict__ ket, int nket,
                                                   double* __restrict__ f, int N) {
    constexpr int NA = 2 * LA + 1, NB = 2 * LB + 1, NC = 2 * LC + 1, ND = 2 * LD + 1;
    constexpr int LAB = LA + LB, LCD = LC + LD, L = LAB + LCD;
    const int kq = blockIdx.x * blockDim.x + threadIdx.x;
    const int bq = blockIdx.y;
    if (kq >= nket || bq >= nbra) return;
    const DShell& A = sh[bra[bq].x];
    const DShell& B = sh[bra[bq].y];
    const DShell& C = sh[ket[kq].x];
    const DShell& D = sh[ket[kq].y];
    double I[NA][NB][NC][ND];
#pragma unroll
    for (int a = 0; a < NA; ++a)
#pragma unroll
        for (int b = 0; b < NB; ++b)
#pragma unroll
            for (int c = 0; c < NC; ++c)
#pragma unroll
                for (int d = 0; d < ND; ++d) I[a][b][c][d] = 0.0;
    double AB[3], CD[3];
#pragma unroll
    for (int x = 0; x < 3; ++x) {
        AB[x] = A.A[x] - B.A[x];
        CD[x] = C.A[x] - D.A[x];
    }
    const double ab2 = AB[0] * AB[0] + AB[1] * AB[1] + AB[2] * AB[2];
    const double cd2 = CD[0] * CD[0] + CD[1] * CD[1] + CD[2] * CD[2];
    for (int i = 0; i < A.np; ++i)
        for (int j = 0; j < B.np; ++j) {
            const double a = A.ex[i], b = B.ex[j], p = a + b;
            const double Kab = exp(-a * b / p * ab2) * A.co[i] * B.co[j];
            if (fabs(Kab) < 1e-20) continue;
            double P[3];
            double Eb[3][LA + 1][LB + 1][LAB + 1];
#pragma unroll
            for (int x = 0; x < 3; ++x) {
                P[x] = (a * A.A[x] + b * B.A[x]) / p;
                hermite_1d<LA, LB>(p, P[x] - A.A[x], P[x] - B.A[x], Eb[x]);
            }
            for (int k = 0; k < C.np; ++k)
                for (int l = 0; l < D.np; ++l) {
                    const double c = C.ex[k], d = D.ex[l], q = c + d;
                    const double Kcd = exp(-c * d / q * cd2) * C.co[k] * D.co[l];
                    double pref = Kab * Kcd;
                    if (fabs(pref) < 1e-22) continue;
                    double Q[3];
                    double Ek[3][LC + 1][LD + 1][LCD + 1];
#pragma unroll
                    for (int x = 0; x < 3; ++x) {
                        Q[x] = (c * C.A[x] + d * D.A[x]) / q;
                        hermite_1d<LC, LD>(q, Q[x] - C.A[x], Q[x] - D.A[x], Ek[x]);
                    }
                    const double alpha = p * q / (p + q);
                    pref *= 2.0 * 17.493418327624862 / (p * q * sqrt(p + q));  // 2 pi^(5/2)
                    double R[L + 1][L + 1][L + 1];
                    hermite_R<L>(alpha, P[0] - Q[0], P[1] - Q[1], P[2] - Q[2], R);
#pragma unroll
                    for (int ca = 0; ca < NA; ++ca)
#pragma unroll
                        for (int cb = 0; cb < NB; ++cb)
#pragma unroll
                            for (int cc = 0; cc < NC; ++cc)
#pragma unroll
                                for (int cd = 0; cd < ND; ++cd) {
                                    const int ix = cexp(LA, ca, 0), iy = cexp(LA, ca, 1), iz = cexp(LA, ca, 2);
                                    const int jx = cexp(LB, cb, 0), jy = cexp(LB, cb, 1), jz = cexp(LB, cb, 2);
                                    const int kx = cexp(LC, cc, 0), ky = cexp(LC, cc, 1), kz = cexp(LC, cc, 2);
                                    const int lx = cexp(LD, cd, 0), ly = cexp(LD, cd, 1), lz = cexp(LD, cd, 2);
                                    double s = 0.0;
#pragma unroll
                                    for (int t = 0; t <= ix + jx; ++t)
#pragma unroll
                                        for (int u = 0; u <= iy + jy; ++u)
#pragma unroll
                                            for (int v = 0; v <= iz + jz; ++v) {
                                                const double eb = Eb[0][ix][jx][t] * Eb[1][iy][jy][u] * Eb[2][iz][jz][v];
                                                double g = 0.0;
#pragma unroll
                                                for (int tt = 0; tt <= kx + lx; ++tt)
#pragma unroll
                                                    for (int uu = 0; uu <= ky + ly; ++uu)
#pragma unroll
                                                        for (int vv = 0; vv <= kz + lz; ++vv) {
                                                            const double ek = Ek[0][kx][lx][tt] * Ek[1][ky][ly][uu] *
                                                                              Ek[2][kz][lz][vv];
                                                            const double r = R[t + tt][u + uu][v + vv];
                                                            g += ((tt + uu + vv) & 1) ? -ek * r : ek * r;
                                                        }
                                                s += eb * g;
                                            }
                                    I[ca][cb][cc][cd] += pref * s;
                                }
                }
        }
    const long long n1 = N, n2 = n1 * N, n3 = n2 * N;
#pragma unroll
    for (int ca = 0; ca < NA; ++ca)
#pragma unroll
        for (int cb = 0; cb < NB; ++cb)
#pragma unroll
            for (int cc = 0; cc < NC; ++cc)
#pragma unroll
                for (int cd = 0; cd < ND; ++cd) {
                    const double v = I[ca][cb][cc][cd];
                    const long long m = A.off + ca, n = B.off + cb, s = C.off + cc, t = D.off + cd;
                    // (mn|st), (nm|st), (mn|ts), (nm|ts); (ac|bd) sits at a + N b + N^2 c + N^3 d
                    f[m + n1 * s + n2 * n + n3 * t] = v;
                    f[n + n1 * s + n2 * m + n3 * t] = v;
                    f[m + n1 * t + n2 * n + n3 * s] = v;
                    f[n + n1 * t + n2 * m + n3 * s] = v;
                }
}

double dfact(int n) {  // n!! with (-1)!! = 1
    double r = 1.0;
    for (int k = n; k > 1; k -= 2) r *= k;
    return r;
}

template <int LA, int LB, int LC, int LD>
void launch_class(hipStream_t s, const DShell* sh, const int2* bra, int nbra, const int2* ket, int nket, double* f,
                  int N) {
    dim3 grid((nket + 255) / 256, nbra);
    hipLaunchKernelGGL((k_eri_class<LA, LB, LC, LD>), grid, dim3(256), 0, s, sh, bra, nbra, ket, nket, f, N);
    QK_HIP(hipGetLastError());
}

using LaunchFn = void (*)(hipStream_t, const DShell*, const int2*, int, const int2*, int, double*, int);
const LaunchFn LAUNCH[16] = {
    launch_class<0, 0, 0, 0>, launch_class<0, 0, 0, 1>, launch_class<0, 0, 1, 0>, launch_class<0, 0, 1, 1>,
    launch_class<0, 1, 0, 0>, launch_class<0, 1, 0, 1>, launch_class<0, 1, 1, 0>, launch_class<0, 1, 1, 1>,
    launch_class<1, 0, 0, 0>, launch_class<1, 0, 0, 1>, launch_class<1, 0, 1, 0>, launch_class<1, 0, 1, 1>,
    launch_class<1, 1, 0, 0>, launch_class<1, 1, 0, 1>, launch_class<1, 1, 1, 0>, launch_class<1, 1, 1, 1>,
};

}  // namespace
}  // namespace qk

int qk_int_eri_device(int nshell, const int* sl, const int* snp, const int* satom, const double* exps,
                      const double* coefs, const double* xyz, double* eri_device, long long n_eri) {
    using namespace qk;
    if (nshell < 1 || !sl || !snp || !satom || !exps || !coefs || !xyz || !eri_device) return QK_E_ARG;
    {
        long long nbf = 0;
        for (int i = 0; i < nshell; ++i) nbf += 2LL * sl[i] + 1;
        if (n_eri != nbf * nbf * nbf * nbf) return QK_E_ARG;  // the buffer must hold nBas^4 doubles
    }
    DShell* dsh = nullptr;
    int2* dpairs = nullptr;
    hipStream_t s = nullptr;
    try {
        std::vector<DShell> shells(nshell);
        int off = 0, poff = 0;
        for (int i = 0; i < nshell; ++i) {
            DShell& S = shells[i];
            if (sl[i] < 0 || snp[i] < 1 || snp[i] > NPRIM) return QK_E_ARG;
            if (sl[i] > 1) return QK_E_UNSUPPORTED;  // d shells: the host producer
            S.l = sl[i];
            S.np = snp[i];
            S.off = off;
            S.pad = 0;
            off += 2 * S.l + 1;
            for (int x = 0; x < 3; ++x) S.A[x] = xyz[3 * satom[i] + x];
            const int l = S.l;
            // primitive normalisation of x^l exp(-a r^2), then the contraction's
            for (int k = 0; k < NPRIM; ++k) {
                S.ex[k] = k < S.np ? exps[poff + k] : 1.0;
                S.co[k] = 0.0;
            }
            for (int k = 0; k < S.np; ++k) {
                const double a = S.ex[k];
                const double n2 = std::pow(2.0 * a / M_PI, 1.5) * std::pow(4.0 * a, l) / dfact(2 * l - 1);
                S.co[k] = coefs[poff + k] * std::sqrt(n2);
            }
            double ss = 0.0;
            for (int k = 0; k < S.np; ++k)
                for (int m = 0; m < S.np; ++m) {
                    const double p = S.ex[k] + S.ex[m];
                    ss += S.co[k] * S.co[m] * std::pow(M_PI / p, 1.5) * dfact(2 * l - 1) / std::pow(2.0 * p, l);
                }
            for (int k = 0; k < S.np; ++k) S.co[k] /= std::sqrt(ss);
            poff += S.np;
        }
        const int N = off;
        // shell pairs a >= b grouped by (la, lb)
        std::vector<int2> cls[4];
        for (int a = 0; a < nshell; ++a)
            for (int b = 0; b <= a; ++b) cls[2 * shells[a].l + shells[b].l].push_back(make_int2(a, b));
        size_t npair = 0;
        for (auto& v : cls) npair += v.size();
        QK_HIP(hipMalloc(&dsh, sizeof(DShell) * nshell));
        QK_HIP(hipMalloc(&dpairs, sizeof(int2) * npair));
        QK_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        QK_HIP(hipMemcpyAsync(dsh, shells.data(), sizeof(DShell) * nshell, hipMemcpyHostToDevice, s));
        const int2* base[4];
        size_t o = 0;
        for (int c = 0; c < 4; ++c) {
            base[c] = dpairs + o;
            if (!cls[c].empty())
                QK_HIP(hipMemcpyAsync(dpairs + o, cls[c].data(), sizeof(int2) * cls[c].size(), hipMemcpyHostToDevice,
                                      s));
            o += cls[c].size();
        }
        for (int x = 0; x < 4; ++x)
            for (int y = 0; y < 4; ++y) {
                if (cls[x].empty() || cls[y].empty()) continue;
                QK_CHECK(cls[x].size() <= 65535, QK_E_ARG, "too many shell pairs of one class");
                LAUNCH[4 * x + y](s, dsh, base[x], (int)cls[x].size(), base[y], (int)cls[y].size(), eri_device, N);
            }
        QK_HIP(hipStreamSynchronize(s));
        (void)hipStreamDestroy(s);
        (void)hipFree(dsh);
        (void)hipFree(dpairs);
        return QK_OK;
    } catch (const Error& e) {
        if (s) {
            (void)hipStreamSynchronize(s);
            (void)hipStreamDestroy(s);
        }
        (void)hipFree(dsh);
        (void)hipFree(dpairs);
        return e.code;
    } catch (...) {
        if (s) (void)hipStreamDestroy(s);
        (void)hipFree(dsh);
        (void)hipFree(dpairs);
        return QK_E_HIP;
    }
}
