/*
 * Gaussian integral producer for contracted s/p/d shells (McMurchie-Davidson).
 *
 * Role: this is the input producer in front of the restricted G0W0/evGW path.
 * It replaces the PySCF calls of PyDuck.py:156-159 (one-electron integrals)
 * and PyDuck.py:329-345 (int/ERI.bin), which cannot run without PySCF.  It
 * writes exactly the arrays that QuAcK reads:
 *   S, T, V            nbf x nbf, symmetric (PyDuck.py:156-158)
 *   eri[a + N b + N^2 c + N^3 d] = <ab|cd> = (ac|bd)
 *                      the physicist-order column-major array that
 *                      utils/read_2e_integrals.f90:55-63 streams into G(n,n,n,n)
 * Real spherical harmonics for l = 2 (5 functions, PySCF's default
 * mol.cart = False, PyDuck.py:126); p functions are ordered x, y, z.
 * Each contracted function is normalised to one.
 *
 * Algorithm: Hermite expansion coefficients E^{ij}_t and Hermite Coulomb
 * integrals R^n_{tuv} (McMurchie & Davidson, J. Comput. Phys. 26, 218 (1978);
 * Helgaker, Jorgensen, Olsen, "Molecular Electronic-Structure Theory", ch. 9).
 * Boys function: power series below the switch point, erf plus upward
 * recursion above it.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdint.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define LMAX 2                 /* highest shell angular momentum */
#define LTOT (4 * LMAX)        /* highest total L of a quartet */
#define DH (LTOT + 1)
#define NCMAX 6                /* cartesian components of a d shell */

typedef struct {
    int l, nprim, ncart, nsph, off;   /* off = first spherical function */
    double A[3];
    const double *exps;
    double coef[32];                  /* primitive-normalised coefficients */
} shell_t;

static const int cart_l[3][NCMAX][3] = {
    {{0, 0, 0}},
    {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}},
    {{2, 0, 0}, {1, 1, 0}, {1, 0, 1}, {0, 2, 0}, {0, 1, 1}, {0, 0, 2}},
};

/* rows: spherical function, cols: cartesian xx,xy,xz,yy,yz,zz (all with the
 * radial normalisation of the xx component).  Order xy, yz, z2, xz, x2-y2. */
static const double d_c2s[5][6] = {
    {0.0, 1.7320508075688772, 0.0, 0.0, 0.0, 0.0},
    {0.0, 0.0, 0.0, 0.0, 1.7320508075688772, 0.0},
    {-0.5, 0.0, 0.0, -0.5, 0.0, 1.0},
    {0.0, 0.0, 1.7320508075688772, 0.0, 0.0, 0.0},
    {0.8660254037844386, 0.0, 0.0, -0.8660254037844386, 0.0, 0.0},
};

static double dfact(int n) { /* (n)!! with (-1)!! = 1 */
    double r = 1.0;
    for (int k = n; k > 1; k -= 2) r *= k;
    return r;
}

void qk_boys(int mmax, double T, double *F) {
    if (T < 1e-14) {
        for (int m = 0; m <= mmax; ++m) F[m] = 1.0 / (2 * m + 1) - T / (2 * m + 3);
        return;
    }
    double e = exp(-T);
    if (T > 2.0 * mmax + 25.0) {
        double st = sqrt(T);
        F[0] = 0.8862269254527580 * erf(st) / st;
        for (int m = 0; m < mmax; ++m) F[m + 1] = ((2 * m + 1) * F[m] - e) / (2.0 * T);
        return;
    }
    double term = 1.0 / (2 * mmax + 1), sum = term;
    for (int k = 1; k < 400; ++k) {
        term *= 2.0 * T / (2 * mmax + 2 * k + 1);
        sum += term;
        if (term < 1e-17 * sum) break;
    }
    F[mmax] = e * sum;
    for (int m = mmax - 1; m >= 0; --m) F[m] = (2.0 * T * F[m + 1] + e) / (2 * m + 1);
}

/* 1D Hermite coefficients E[i][j][t], i<=imax, j<=jmax, without exp(-mu X^2). */
#define EI(i, j, t) (((i) * (JD) + (j)) * (TD) + (t))
static void hermite_1d(int imax, int jmax, double p, double xpa, double xpb, double *E, int JD, int TD) {
    double h = 0.5 / p;
    for (int k = 0; k < (imax + 1) * JD * TD; ++k) E[k] = 0.0;
    E[EI(0, 0, 0)] = 1.0;
    for (int i = 0; i < imax; ++i)
        for (int t = 0; t <= i + 1; ++t) {
            double v = xpa * (t <= i ? E[EI(i, 0, t)] : 0.0);
            if (t > 0) v += h * E[EI(i, 0, t - 1)];
            if (t + 1 <= i) v += (t + 1) * E[EI(i, 0, t + 1)];
            E[EI(i + 1, 0, t)] = v;
        }
    for (int i = 0; i <= imax; ++i)
        for (int j = 0; j < jmax; ++j)
            for (int t = 0; t <= i + j + 1; ++t) {
                double v = xpb * (t <= i + j ? E[EI(i, j, t)] : 0.0);
                if (t > 0) v += h * E[EI(i, j, t - 1)];
                if (t + 1 <= i + j) v += (t + 1) * E[EI(i, j, t + 1)];
                E[EI(i, j + 1, t)] = v;
            }
}

/* Hermite Coulomb integrals R^0_{tuv}, t+u+v <= L, stored R[(t*DH+u)*DH+v]. */
static void hermite_R(int L, double alpha, double X, double Y, double Z, double *Rout) {
    static __thread double R[DH][DH * DH * DH];
    double F[DH];
    qk_boys(L, alpha * (X * X + Y * Y + Z * Z), F);
    double f = 1.0;
    for (int n = 0; n <= L; ++n) {
        R[n][0] = f * F[n];
        f *= -2.0 * alpha;
    }
    for (int n = L - 1; n >= 0; --n) {
        double *Rn = R[n], *R1 = R[n + 1];
        for (int s = 1; s <= L - n; ++s)
            for (int t = 0; t <= s; ++t)
                for (int u = 0; u <= s - t; ++u) {
                    int v = s - t - u;
                    double val;
                    if (t > 0) {
                        val = X * R1[((t - 1) * DH + u) * DH + v];
                        if (t > 1) val += (t - 1) * R1[((t - 2) * DH + u) * DH + v];
                    } else if (u > 0) {
                        val = Y * R1[(t * DH + u - 1) * DH + v];
                        if (u > 1) val += (u - 1) * R1[(t * DH + u - 2) * DH + v];
                    } else {
                        val = Z * R1[(t * DH + u) * DH + v - 1];
                        if (v > 1) val += (v - 1) * R1[(t * DH + u) * DH + v - 2];
                    }
                    Rn[(t * DH + u) * DH + v] = val;
                }
    }
    for (int t = 0; t <= L; ++t)
        for (int u = 0; u <= L - t; ++u)
            for (int v = 0; v <= L - t - u; ++v) Rout[(t * DH + u) * DH + v] = R[0][(t * DH + u) * DH + v];
}

static int build_shells(int nshell, const int *sl, const int *snp, const int *satom, const double *exps,
                        const double *coefs, const double *xyz, shell_t *sh) {
    int off = 0, poff = 0;
    for (int s = 0; s < nshell; ++s) {
        shell_t *S = &sh[s];
        S->l = sl[s];
        S->nprim = snp[s];
        if (S->l < 0 || S->l > LMAX || S->nprim < 1 || S->nprim > 32) return -1;
        S->ncart = (S->l + 1) * (S->l + 2) / 2;
        S->nsph = 2 * S->l + 1;
        S->off = off;
        off += S->nsph;
        for (int k = 0; k < 3; ++k) S->A[k] = xyz[3 * satom[s] + k];
        S->exps = exps + poff;
        int l = S->l;
        /* primitive normalisation of x^l exp(-a r^2) */
        for (int k = 0; k < S->nprim; ++k) {
            double a = S->exps[k];
            double n2 = pow(2.0 * a / M_PI, 1.5) * pow(4.0 * a, l) / dfact(2 * l - 1);
            S->coef[k] = coefs[poff + k] * sqrt(n2);
        }
        /* contracted normalisation of the x^l component */
        double ss = 0.0;
        for (int k = 0; k < S->nprim; ++k)
            for (int m = 0; m < S->nprim; ++m) {
                double p = S->exps[k] + S->exps[m];
                ss += S->coef[k] * S->coef[m] * pow(M_PI / p, 1.5) * dfact(2 * l - 1) / pow(2.0 * p, l);
            }
        for (int k = 0; k < S->nprim; ++k) S->coef[k] /= sqrt(ss);
        poff += S->nprim;
    }
    return off;
}

int qk_int_nbf(int nshell, const int *sl) {
    int n = 0;
    for (int s = 0; s < nshell; ++s) n += 2 * sl[s] + 1;
    return n;
}

/* cartesian block (na x nb) -> spherical block written through a callback-free
 * transform: out[i*nsb + j] = sum_ab C_a[i,a] C_b[j,b] in[a*ncb + b] */
static void c2s_vec(int l, const double *in, int stride, int n, double *out, int ostride) {
    /* transforms the leading index of in (ncart rows of length n) */
    if (l < 2) {
        for (int i = 0; i < 2 * l + 1; ++i)
            for (int k = 0; k < n; ++k) out[i * ostride + k] = in[i * stride + k];
        return;
    }
    for (int i = 0; i < 5; ++i)
        for (int k = 0; k < n; ++k) {
            double v = 0.0;
            for (int a = 0; a < 6; ++a) v += d_c2s[i][a] * in[a * stride + k];
            out[i * ostride + k] = v;
        }
}

void qk_int_1e(int nshell, const int *sl, const int *snp, const int *satom, const double *exps, const double *coefs,
               int natom, const double *xyz, const double *charge, double *S, double *T, double *V) {
    shell_t *sh = (shell_t *)malloc(sizeof(shell_t) * nshell);
    int N = build_shells(nshell, sl, snp, satom, exps, coefs, xyz, sh);
    if (N < 0) { free(sh); return; }
#pragma omp parallel for schedule(dynamic)
    for (int ab = 0; ab < nshell * nshell; ++ab) {
        int sa = ab / nshell, sb = ab % nshell;
        const shell_t *A = &sh[sa], *B = &sh[sb];
        int la = A->l, lb = B->l;
        double Sc[NCMAX * NCMAX], Tc[NCMAX * NCMAX], Vc[NCMAX * NCMAX];
        memset(Sc, 0, sizeof Sc); memset(Tc, 0, sizeof Tc); memset(Vc, 0, sizeof Vc);
        double AB[3] = {A->A[0] - B->A[0], A->A[1] - B->A[1], A->A[2] - B->A[2]};
        double ab2 = AB[0] * AB[0] + AB[1] * AB[1] + AB[2] * AB[2];
        const int JD = LMAX + 3, TD = 2 * LMAX + 3;
        double E[3][(LMAX + 1) * (LMAX + 3) * (2 * LMAX + 3)];
        double R[DH * DH * DH];
        for (int i = 0; i < A->nprim; ++i)
            for (int j = 0; j < B->nprim; ++j) {
                double a = A->exps[i], b = B->exps[j], p = a + b, mu = a * b / p;
                double K = exp(-mu * ab2) * A->coef[i] * B->coef[j];
                double P[3];
                for (int k = 0; k < 3; ++k) {
                    P[k] = (a * A->A[k] + b * B->A[k]) / p;
                    hermite_1d(la, lb + 2, p, P[k] - A->A[k], P[k] - B->A[k], E[k], JD, TD);
                }
                double s0 = sqrt(M_PI / p);
                for (int ca = 0; ca < A->ncart; ++ca)
                    for (int cb = 0; cb < B->ncart; ++cb) {
                        const int *ia = cart_l[la][ca], *ib = cart_l[lb][cb];
                        double s1[3], t1[3];
                        for (int k = 0; k < 3; ++k) {
                            int ii = ia[k], jj = ib[k];
                            double *Ek = E[k];
                            s1[k] = Ek[EI(ii, jj, 0)] * s0;
                            double tv = -2.0 * b * b * Ek[EI(ii, jj + 2, 0)] * s0 + b * (2 * jj + 1) * s1[k];
                            if (jj >= 2) tv -= 0.5 * jj * (jj - 1) * Ek[EI(ii, jj - 2, 0)] * s0;
                            t1[k] = tv;
                        }
                        Sc[ca * NCMAX + cb] += K * s1[0] * s1[1] * s1[2];
                        Tc[ca * NCMAX + cb] += K * (t1[0] * s1[1] * s1[2] + s1[0] * t1[1] * s1[2] + s1[0] * s1[1] * t1[2]);
                    }
                int L = la + lb;
                for (int c = 0; c < natom; ++c) {
                    hermite_R(L, p, P[0] - xyz[3 * c], P[1] - xyz[3 * c + 1], P[2] - xyz[3 * c + 2], R);
                    double pref = -charge[c] * 2.0 * M_PI / p * K;
                    for (int ca = 0; ca < A->ncart; ++ca)
                        for (int cb = 0; cb < B->ncart; ++cb) {
                            const int *ia = cart_l[la][ca], *ib = cart_l[lb][cb];
                            double v = 0.0;
                            for (int t = 0; t <= ia[0] + ib[0]; ++t)
                                for (int u = 0; u <= ia[1] + ib[1]; ++u)
                                    for (int w = 0; w <= ia[2] + ib[2]; ++w)
                                        v += E[0][EI(ia[0], ib[0], t)] * E[1][EI(ia[1], ib[1], u)] *
                                             E[2][EI(ia[2], ib[2], w)] * R[(t * DH + u) * DH + w];
                            Vc[ca * NCMAX + cb] += pref * v;
                        }
                }
            }
        double tmp[NCMAX * NCMAX], outb[NCMAX * NCMAX];
        double *src[3] = {Sc, Tc, Vc}, *dst[3] = {S, T, V};
        for (int m = 0; m < 3; ++m) {
            /* rows (a) then columns (b) */
            c2s_vec(la, src[m], NCMAX, B->ncart, tmp, NCMAX);
            double tt[NCMAX * NCMAX];
            for (int i = 0; i < A->nsph; ++i)
                for (int k = 0; k < B->ncart; ++k) tt[k * NCMAX + i] = tmp[i * NCMAX + k];
            c2s_vec(lb, tt, NCMAX, A->nsph, outb, NCMAX);
            for (int j = 0; j < B->nsph; ++j)
                for (int i = 0; i < A->nsph; ++i) dst[m][(size_t)(A->off + i) * N + B->off + j] = outb[j * NCMAX + i];
        }
    }
    free(sh);
}

/* Electron repulsion integrals over all unique shell quartets, expanded into
 * the full N^4 array eri[a + N b + N^2 c + N^3 d] = (ac|bd). */
int qk_int_eri(int nshell, const int *sl, const int *snp, const int *satom, const double *exps, const double *coefs,
               const double *xyz, double *eri) {
    shell_t *sh = (shell_t *)malloc(sizeof(shell_t) * nshell);
    int N = build_shells(nshell, sl, snp, satom, exps, coefs, xyz, sh);
    if (N < 0) { free(sh); return -1; }
    const size_t n1 = (size_t)N, n2 = n1 * n1, n3 = n2 * n1;
    int npair = nshell * (nshell + 1) / 2;
    int *pa = (int *)malloc(sizeof(int) * npair), *pb = (int *)malloc(sizeof(int) * npair);
    for (int a = 0, k = 0; a < nshell; ++a)
        for (int b = 0; b <= a; ++b, ++k) { pa[k] = a; pb[k] = b; }
    long long nquart = (long long)npair * (npair + 1) / 2;
    int maxnp = 1;
    for (int s = 0; s < nshell; ++s) maxnp = sh[s].nprim > maxnp ? sh[s].nprim : maxnp;
#pragma omp parallel
    {
        const int JD = LMAX + 1, TD = 2 * LMAX + 1;
        double *Eab = (double *)malloc(sizeof(double) * maxnp * maxnp * NCMAX * NCMAX * 125);
        double Ex[3][(LMAX + 1) * (LMAX + 1) * (2 * LMAX + 1)];
        double R[DH * DH * DH];
        double G[NCMAX * NCMAX][125];
        double Ic[NCMAX * NCMAX * NCMAX * NCMAX];
        double t1[NCMAX * NCMAX * NCMAX * NCMAX], t2[NCMAX * NCMAX * NCMAX * NCMAX];
#pragma omp for schedule(dynamic, 16)
        for (long long q = 0; q < nquart; ++q) {
            /* decode q -> (pair1 >= pair2) */
            long long r = (long long)((sqrt(8.0 * (double)q + 1.0) - 1.0) / 2.0);
            while (r * (r + 1) / 2 > q) --r;
            while ((r + 1) * (r + 2) / 2 <= q) ++r;
            int p1 = (int)r, p2 = (int)(q - r * (r + 1) / 2);
            const shell_t *A = &sh[pa[p1]], *B = &sh[pb[p1]], *C = &sh[pa[p2]], *D = &sh[pb[p2]];
            int la = A->l, lb = B->l, lc = C->l, ld = D->l;
            int nab = A->ncart * B->ncart, ncd = C->ncart * D->ncart;
            int Lab = la + lb, Lcd = lc + ld, L = Lab + Lcd;
            memset(Ic, 0, sizeof(double) * nab * ncd);
            double AB[3], CD[3];
            for (int k = 0; k < 3; ++k) { AB[k] = A->A[k] - B->A[k]; CD[k] = C->A[k] - D->A[k]; }
            double ab2 = AB[0] * AB[0] + AB[1] * AB[1] + AB[2] * AB[2];
            double cd2 = CD[0] * CD[0] + CD[1] * CD[1] + CD[2] * CD[2];
            /* ket primitive pairs: Hermite tables E^{cd}_{tuv} with sign folded in */
            int nket = C->nprim * D->nprim;
            double kp[1024], kK[1024], kP[1024][3];
            const int HT = 125; /* (Lcd+1)^3 <= 125 */
            for (int i = 0; i < C->nprim; ++i)
                for (int j = 0; j < D->nprim; ++j) {
                    int k = i * D->nprim + j;
                    double c = C->exps[i], d = D->exps[j], p = c + d;
                    kp[k] = p;
                    kK[k] = exp(-c * d / p * cd2) * C->coef[i] * D->coef[j];
                    for (int x = 0; x < 3; ++x) {
                        kP[k][x] = (c * C->A[x] + d * D->A[x]) / p;
                        hermite_1d(lc, ld, p, kP[k][x] - C->A[x], kP[k][x] - D->A[x], Ex[x], JD, TD);
                    }
                    double *Ek = Eab + (size_t)k * NCMAX * NCMAX * HT;
                    for (int cc = 0; cc < C->ncart; ++cc)
                        for (int cd = 0; cd < D->ncart; ++cd) {
                            const int *ic = cart_l[lc][cc], *id = cart_l[ld][cd];
                            double *e = Ek + (cc * D->ncart + cd) * HT;
                            for (int t = 0; t <= Lcd; ++t)
                                for (int u = 0; u <= Lcd; ++u)
                                    for (int v = 0; v <= Lcd; ++v) {
                                        double val = 0.0;
                                        if (t <= ic[0] + id[0] && u <= ic[1] + id[1] && v <= ic[2] + id[2]) {
                                            val = Ex[0][(ic[0] * JD + id[0]) * TD + t] * Ex[1][(ic[1] * JD + id[1]) * TD + u] *
                                                  Ex[2][(ic[2] * JD + id[2]) * TD + v];
                                            if ((t + u + v) & 1) val = -val;
                                        }
                                        e[(t * (Lcd + 1) + u) * (Lcd + 1) + v] = val;
                                    }
                        }
                }
            for (int i = 0; i < A->nprim; ++i)
                for (int j = 0; j < B->nprim; ++j) {
                    double a = A->exps[i], b = B->exps[j], p = a + b;
                    double Kab = exp(-a * b / p * ab2) * A->coef[i] * B->coef[j];
                    if (fabs(Kab) < 1e-20) continue;
                    double P[3];
                    for (int x = 0; x < 3; ++x) {
                        P[x] = (a * A->A[x] + b * B->A[x]) / p;
                        hermite_1d(la, lb, p, P[x] - A->A[x], P[x] - B->A[x], Ex[x], JD, TD);
                    }
                    for (int k = 0; k < nket; ++k) {
                        double q2 = kp[k];
                        double pref = Kab * kK[k];
                        if (fabs(pref) < 1e-22) continue;
                        double alpha = p * q2 / (p + q2);
                        pref *= 2.0 * pow(M_PI, 2.5) / (p * q2 * sqrt(p + q2));
                        hermite_R(L, alpha, P[0] - kP[k][0], P[1] - kP[k][1], P[2] - kP[k][2], R);
                        const double *Ek = Eab + (size_t)k * NCMAX * NCMAX * HT;
                        /* G[cd][tuv] = sum_{tau nu phi} Ecd * R_{t+tau,...} for t+u+v <= Lab */
                        int D1 = Lcd + 1;
                        for (int c = 0; c < ncd; ++c) {
                            const double *e = Ek + c * HT;
                            for (int t = 0; t <= Lab; ++t)
                                for (int u = 0; u <= Lab - t; ++u)
                                    for (int v = 0; v <= Lab - t - u; ++v) {
                                        double s = 0.0;
                                        for (int tt = 0; tt <= Lcd; ++tt)
                                            for (int uu = 0; uu <= Lcd - tt; ++uu)
                                                for (int vv = 0; vv <= Lcd - tt - uu; ++vv)
                                                    s += e[(tt * D1 + uu) * D1 + vv] *
                                                         R[((t + tt) * DH + u + uu) * DH + v + vv];
                                        G[c][(t * 5 + u) * 5 + v] = s;
                                    }
                        }
                        for (int ca = 0; ca < A->ncart; ++ca)
                            for (int cb = 0; cb < B->ncart; ++cb) {
                                const int *ia = cart_l[la][ca], *ib = cart_l[lb][cb];
                                double *row = Ic + (size_t)(ca * B->ncart + cb) * ncd;
                                for (int t = 0; t <= ia[0] + ib[0]; ++t) {
                                    double et = Ex[0][(ia[0] * JD + ib[0]) * TD + t];
                                    for (int u = 0; u <= ia[1] + ib[1]; ++u) {
                                        double eu = et * Ex[1][(ia[1] * JD + ib[1]) * TD + u];
                                        for (int v = 0; v <= ia[2] + ib[2]; ++v) {
                                            double ev = pref * eu * Ex[2][(ia[2] * JD + ib[2]) * TD + v];
                                            int h = (t * 5 + u) * 5 + v;
                                            for (int c = 0; c < ncd; ++c) row[c] += ev * G[c][h];
                                        }
                                    }
                                }
                            }
                    }
                }
            /* spherical transform of the four indices: Ic[a][b][c][d] */
            int nb = B->ncart, nc = C->ncart, nd = D->ncart;
            int sa = A->nsph, sb = B->nsph, sc = C->nsph, sd = D->nsph;
            double *cur = Ic;
            /* index a */
            c2s_vec(la, cur, nb * nc * nd, nb * nc * nd, t1, nb * nc * nd);
            /* index b: view t1 as [sa][nb][nc*nd] */
            for (int i = 0; i < sa; ++i) c2s_vec(lb, t1 + i * nb * nc * nd, nc * nd, nc * nd, t2 + i * sb * nc * nd, nc * nd);
            /* index c: view t2 as [sa*sb][nc][nd] */
            for (int i = 0; i < sa * sb; ++i) c2s_vec(lc, t2 + i * nc * nd, nd, nd, t1 + i * sc * nd, nd);
            /* index d: view t1 as [sa*sb*sc][nd] -> transpose trick */
            for (int i = 0; i < sa * sb * sc; ++i) {
                double colv[NCMAX], outv[5];
                for (int k = 0; k < nd; ++k) colv[k] = t1[i * nd + k];
                c2s_vec(ld, colv, 1, 1, outv, 1);
                for (int k = 0; k < sd; ++k) t2[i * sd + k] = outv[k];
            }
            for (int i = 0; i < sa; ++i)
                for (int j = 0; j < sb; ++j)
                    for (int k = 0; k < sc; ++k)
                        for (int l = 0; l < sd; ++l) {
                            double v = t2[((i * sb + j) * sc + k) * sd + l];
                            size_t m = A->off + i, n = B->off + j, s = C->off + k, t = D->off + l;
                            /* (mn|st) and its 8 images: (ac|bd) sits at a + N b + N^2 c + N^3 d */
                            eri[m + n1 * s + n2 * n + n3 * t] = v;
                            eri[n + n1 * s + n2 * m + n3 * t] = v;
                            eri[m + n1 * t + n2 * n + n3 * s] = v;
                            eri[n + n1 * t + n2 * m + n3 * s] = v;
                            eri[s + n1 * m + n2 * t + n3 * n] = v;
                            eri[t + n1 * m + n2 * s + n3 * n] = v;
                            eri[s + n1 * n + n2 * t + n3 * m] = v;
                            eri[t + n1 * n + n2 * s + n3 * m] = v;
                        }
        }
        free(Eab);
    }
    free(pa);
    free(pb);
    free(sh);
    return N;
}
