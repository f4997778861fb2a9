// Coulomb and exchange matrices of the restricted Hartree-Fock SCF from the
// AO integrals resident on the device, in one pass over them.
//
// QuAcK (AOtoMO/Hartree_matrix_AO_basis.f90:22-32,
// AOtoMO/exchange_matrix_AO_basis.f90:22-32), with ERI(mu,la,nu,si) =
// <mu la|nu si> = (mu nu|la si) stored at mu + N la + N^2 nu + N^3 si:
//   J(mu, nu) =  sum_{la,si} P(la, si) ERI(mu, la, nu, si)
//   K(mu, nu) = -sum_{la,si} P(la, si) ERI(mu, la, si, nu)
// Element E = ERI(mu, la, a, b) therefore feeds J(mu, a) with weight P(la, b)
// and K(mu, b) with weight -P(la, a): for every slab (a, b), the N x N plane
// M_ab(mu, la), two matrix-vector products with the same matrix,
//   J(:, a) += M_ab P(:, b),   K(:, b) -= M_ab P(:, a).
// Each workgroup reads the slabs of a TA x TB block of (a, b) once (mu along
// the threads, so every wave load is 64 consecutive doubles) and keeps the
// partial J columns of its a's and K columns of its b's in registers; a second
// kernel sums the partials of the b-blocks (J) and a-blocks (K) in a fixed
// order, so the result does not depend on scheduling.  HBM-bound: N^4 x 8
// bytes per Fock build.
#include <algorithm>
#include <string>

#include "qk_common.h"

namespace qk {
namespace {

constexpr int TA = 4, TB = 4;  // slabs per workgroup: a-block x b-block
constexpr int NT = 256;        // threads; thread t owns mu = t, t + NT
constexpr int MU = 2;          // so N <= 512 (the N^4 array of N = 512 is 550 GB)
constexpr int NMAX = NT * MU;

// Jp[(bb * N + a) * N + mu] = sum over b in b-block bb of (M_ab P(:, b))(mu)
// Kp[(ab * N + b) * N + mu] = sum over a in a-block ab of (M_ab P(:, a))(mu)
__global__ __launch_bounds__(NT) void k_jk_partial(const double* __restrict__ E, const double* __restrict__ P,
                                                   int N, double* __restrict__ Jp, double* __restrict__ Kp) {
    __shared__ double pa[TA][NMAX], pb[TB][NMAX];
    const int a0 = blockIdx.x * TA, b0 = blockIdx.y * TB;
    const int na = min(TA, N - a0), nb = min(TB, N - b0);
    for (int e = threadIdx.x; e < TA * N; e += NT) {
        const int i = e / N, l = e % N;
        pa[i][l] = i < na ? P[l + (long long)N * (a0 + i)] : 0.0;
    }
    for (int e = threadIdx.x; e < TB * N; e += NT) {
        const int j = e / N, l = e % N;
        pb[j][l] = j < nb ? P[l + (long long)N * (b0 + j)] : 0.0;
    }
    __syncthreads();
    const long long N2 = (long long)N * N, N3 = N2 * N;
    double jacc[TA][MU], kacc[TB][MU];
#pragma unroll
    for (int i = 0; i < TA; ++i)
#pragma unroll
        for (int u = 0; u < MU; ++u) jacc[i][u] = 0.0;
#pragma unroll
    for (int j = 0; j < TB; ++j)
#pragma unroll
        for (int u = 0; u < MU; ++u) kacc[j][u] = 0.0;
    int mu[MU];
    bool on[MU];
#pragma unroll
    for (int u = 0; u < MU; ++u) {
        mu[u] = threadIdx.x + NT * u;
        on[u] = mu[u] < N;
    }
#pragma unroll
    for (int i = 0; i < TA; ++i) {
        if (i >= na) break;
#pragma unroll
        for (int j = 0; j < TB; ++j) {
            if (j >= nb) break;
            const double* slab = E + N2 * (a0 + i) + N3 * (b0 + j);
            double sj[MU] = {}, sk[MU] = {};
#pragma unroll 8
            for (int l = 0; l < N; ++l) {
                const double wb = pb[j][l], wa = pa[i][l];
#pragma unroll
                for (int u = 0; u < MU; ++u) {
                    const double e = on[u] ? slab[mu[u] + (long long)N * l] : 0.0;
                    sj[u] += e * wb;
                    sk[u] += e * wa;
                }
            }
#pragma unroll
            for (int u = 0; u < MU; ++u) {
                jacc[i][u] += sj[u];
                kacc[j][u] += sk[u];
            }
        }
    }
#pragma unroll
    for (int u = 0; u < MU; ++u) {
        if (!on[u]) continue;
        for (int i = 0; i < na; ++i) Jp[((long long)blockIdx.y * N + a0 + i) * N + mu[u]] = jacc[i][u];
        for (int j = 0; j < nb; ++j) Kp[((long long)blockIdx.x * N + b0 + j) * N + mu[u]] = kacc[j][u];
    }
}

// J(mu, a) = sum_bb Jp[bb][a][mu], K(mu, b) = -sum_ab Kp[ab][b][mu] (fixed order)
__global__ void k_jk_reduce(const double* __restrict__ Jp, const double* __restrict__ Kp, int N, int nblk_a,
                            int nblk_b, double* __restrict__ J, double* __restrict__ K) {
    const long long N2 = (long long)N * N;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < N2; t += (long long)gridDim.x * blockDim.x) {
        double j = 0.0, k = 0.0;
        for (int bb = 0; bb < nblk_b; ++bb) j += Jp[bb * N2 + t];
        for (int ab = 0; ab < nblk_a; ++ab) k += Kp[ab * N2 + t];
        J[t] = j;
        K[t] = -k;
    }
}

}  // namespace

void rhf_jk(hipStream_t s, int N, const double* E, const double* P, double* J, double* K, double* work) {
    QK_CHECK(N >= 1 && N <= NMAX, QK_E_ARG, "rhf_jk: nBas out of range");
    const int nba = (N + TA - 1) / TA, nbb = (N + TB - 1) / TB;
    double* Jp = work;
    double* Kp = work + (size_t)nbb * N * N;
    hipLaunchKernelGGL(k_jk_partial, dim3(nba, nbb), dim3(NT), 0, s, E, P, N, Jp, Kp);
    QK_HIP(hipGetLastError());
    const long long N2 = (long long)N * N;
    hipLaunchKernelGGL(k_jk_reduce, dim3((unsigned)std::min<long long>((N2 + 255) / 256, 4096)), dim3(256), 0, s, Jp,
                       Kp, N, nba, nbb, J, K);
    QK_HIP(hipGetLastError());
}

size_t rhf_jk_work_doubles(int N) {
    const size_t nba = (N + TA - 1) / TA, nbb = (N + TB - 1) / TB;
    return (nba + nbb) * (size_t)N * N;
}

}  // namespace qk

struct qk_scf_ctx {
    int nBas = 0;
    hipStream_t s = nullptr;
    double *P = nullptr, *J = nullptr, *K = nullptr, *work = nullptr;
    double* eri = nullptr;  // the handle's own copy of the AO integrals (qk_rhf_set_eri_ao)
    hipEvent_t ev[2] = {nullptr, nullptr};
    float last_ms = 0.f;
    std::string err;
};

namespace {
void scf_free(qk_scf_ctx* h) {
    if (h->s) (void)hipStreamSynchronize(h->s);
    (void)hipFree(h->P);
    (void)hipFree(h->J);
    (void)hipFree(h->K);
    (void)hipFree(h->work);
    (void)hipFree(h->eri);
    for (auto& e : h->ev)
        if (e) (void)hipEventDestroy(e);
    if (h->s) (void)hipStreamDestroy(h->s);
}
}  // namespace

int qk_rhf_create(qk_scf_handle* out, int nBas) {
    if (!out || nBas < 1 || nBas > qk::NMAX) return QK_E_ARG;
    *out = nullptr;
    qk_scf_ctx* h = new qk_scf_ctx;
    h->nBas = nBas;
    try {
        const size_t n2 = (size_t)nBas * nBas;
        QK_HIP(hipStreamCreateWithFlags(&h->s, hipStreamNonBlocking));
        QK_HIP(hipMalloc(&h->P, sizeof(double) * n2));
        QK_HIP(hipMalloc(&h->J, sizeof(double) * n2));
        QK_HIP(hipMalloc(&h->K, sizeof(double) * n2));
        QK_HIP(hipMalloc(&h->work, sizeof(double) * qk::rhf_jk_work_doubles(nBas)));
        QK_HIP(hipEventCreate(&h->ev[0]));
        QK_HIP(hipEventCreate(&h->ev[1]));
    } catch (const qk::Error& e) {
        scf_free(h);
        delete h;
        return e.code;
    }
    *out = h;
    return QK_OK;
}

int qk_rhf_destroy(qk_scf_handle h) {
    if (!h) return QK_E_ARG;
    scf_free(h);
    delete h;
    return QK_OK;
}

const char* qk_rhf_last_error(qk_scf_handle h) { return h ? h->err.c_str() : ""; }

int qk_rhf_set_eri_ao(qk_scf_handle h, const double* ERI_AO) {
    if (!h || !ERI_AO) return QK_E_ARG;
    const long long N = h->nBas;
    try {
        if (!h->eri) QK_HIP(hipMalloc(&h->eri, sizeof(double) * N * N * N * N));
        QK_HIP(hipMemcpyAsync(h->eri, ERI_AO, sizeof(double) * N * N * N * N, hipMemcpyHostToDevice, h->s));
        QK_HIP(hipStreamSynchronize(h->s));
    } catch (const qk::Error& e) {
        h->err = e.msg;
        return e.code;
    }
    return QK_OK;
}

int qk_rhf_jk(qk_scf_handle h, const double* ERI_device, long long n_eri, const double* P, double* J, double* K) {
    if (!h || !P || !J || !K) return QK_E_ARG;
    const long long N = h->nBas;
    if (n_eri != N * N * N * N) return QK_E_ARG;
    if (!ERI_device) ERI_device = h->eri;  // the copy made by qk_rhf_set_eri_ao
    if (!ERI_device) return QK_E_ARG;
    try {
        const size_t bytes = sizeof(double) * N * N;
        QK_HIP(hipMemcpyAsync(h->P, P, bytes, hipMemcpyHostToDevice, h->s));
        QK_HIP(hipEventRecord(h->ev[0], h->s));
        qk::rhf_jk(h->s, h->nBas, ERI_device, h->P, h->J, h->K, h->work);
        QK_HIP(hipEventRecord(h->ev[1], h->s));
        QK_HIP(hipMemcpyAsync(J, h->J, bytes, hipMemcpyDeviceToHost, h->s));
        QK_HIP(hipMemcpyAsync(K, h->K, bytes, hipMemcpyDeviceToHost, h->s));
        QK_HIP(hipStreamSynchronize(h->s));
        QK_HIP(hipEventElapsedTime(&h->last_ms, h->ev[0], h->ev[1]));
    } catch (const qk::Error& e) {
        h->err = e.msg;
        (void)hipStreamSynchronize(h->s);
        return e.code;
    }
    return QK_OK;
}

int qk_rhf_jk_ms(qk_scf_handle h, double* ms) {
    if (!h || !ms) return QK_E_ARG;
    *ms = h->last_ms;
    return QK_OK;
}
