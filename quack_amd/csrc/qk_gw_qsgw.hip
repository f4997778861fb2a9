// Quasiparticle self-consistent GW (qk_qsgw_*; GW/qsRGW.f90:193-341 without
// linDM).  One handle holds a qsGW calculation: the AO integrals in HBM (read
// by the Hartree/exchange builds and by the engine's AO->MO block every
// cycle), the engine handle that screens and builds the full self-energy, and
// the N^3 matrices of the cycle on the device.  qk_qsgw_cycle runs one pass
// of QuAcK's loop body, so a Fortran caller keeps QuAcK's loop, convergence
// test and printing around it.
#include "qk_gw_internal.h"

using namespace qk;

struct qk_qsgw_ctx {
    int nBas = 0, nOrb = 0, nC = 0, nO = 0, nV = 0, nR = 0, nS = 0;
    std::string err;
    qk_gw_handle g = nullptr;
    qk_scf_handle jk = nullptr;
    hipStream_t s = nullptr;
    rocblas_handle rb = nullptr;
    double* eri = nullptr;  // nBas^4 AO integrals (QuAcK's ERI.bin order)
    // nBas x nBas: S, P, F and three products; nBas x nOrb: X, c, c2, SC, T;
    // nOrb x nOrb: Sig, Fp; w, work: nOrb
    double *S = nullptr, *P = nullptr, *F = nullptr, *A1 = nullptr, *A2 = nullptr, *A3 = nullptr;
    double *X = nullptr, *c = nullptr, *c2 = nullptr, *SC = nullptr, *Tm = nullptr;
    double *Sig = nullptr, *Fp = nullptr, *w = nullptr, *work = nullptr;
    int* info = nullptr;
    std::vector<double> hT, hV, hHc, hP, hJ, hK, hF, hErr, hSig, hSigp, eGW;
    Diis* diis = nullptr;
    int max_diis = 1;
};

namespace {

void qsgw_release(qk_qsgw_ctx* h) {
    if (h->s) (void)hipStreamSynchronize(h->s);
    for (double** p : {&h->eri, &h->S, &h->P, &h->F, &h->A1, &h->A2, &h->A3, &h->X, &h->c, &h->c2, &h->SC, &h->Tm,
                       &h->Sig, &h->Fp, &h->w, &h->work})
        dfree(*p);
    dfree(h->info);
    if (h->g) (void)qk_gw_destroy(h->g);
    if (h->jk) (void)qk_rhf_destroy(h->jk);
    if (h->rb) (void)rocblas_destroy_handle(h->rb);
    if (h->s) (void)hipStreamDestroy(h->s);
    delete h->diis;
    h->g = nullptr;
    h->jk = nullptr;
    h->rb = nullptr;
    h->s = nullptr;
    h->diis = nullptr;
}

void up(hipStream_t s, double* d, const double* h, size_t n) {
    QK_HIP(hipMemcpyAsync(d, h, sizeof(double) * n, hipMemcpyHostToDevice, s));
}
void down(hipStream_t s, double* h, const double* d, size_t n) {
    QK_HIP(hipMemcpyAsync(h, d, sizeof(double) * n, hipMemcpyDeviceToHost, s));
}

// sum_ij A(i,j) B(j,i) of two n x n column-major matrices (utils/utils.f90
// trace_matrix of matmul(A,B))
double trace_prod(int n, const std::vector<double>& A, const std::vector<double>& B) {
    double t = 0.0;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) t += A[i + (size_t)n * j] * B[j + (size_t)n * i];
    return t;
}

}  // namespace

extern "C" {

int qk_qsgw_create(qk_qsgw_handle* out, int nBas, int nOrb, int nC, int nO, int nV, int nR, int nS, int max_diis,
                   const double* S, const double* T, const double* V, const double* Hc, const double* X,
                   const double* ERI_AO, const double* PHF, const double* cHF, const double* eHF) {
    if (!out || !S || !T || !V || !Hc || !X || !ERI_AO || !PHF || !cHF || !eHF) return QK_E_ARG;
    *out = nullptr;
    if (nBas <= 0 || nOrb <= 0 || nOrb > nBas || max_diis < 0) return QK_E_ARG;
    qk_qsgw_ctx* h = new qk_qsgw_ctx;
    int rc = guarded(&h->err, [&] {
        check_sizes(nBas, nOrb, nC, nO, nV, nR, nS);
        h->nBas = nBas;
        h->nOrb = nOrb;
        h->nC = nC;
        h->nO = nO;
        h->nV = nV;
        h->nR = nR;
        h->nS = nS;
        h->max_diis = std::max(max_diis, 1);
        const size_t N = nBas, n = nOrb, N2 = N * N;
        QK_HIP(hipStreamCreateWithFlags(&h->s, hipStreamNonBlocking));
        rb_check(rocblas_create_handle(&h->rb), "rocblas_create_handle");
        rb_check(rocblas_set_stream(h->rb, h->s), "rocblas_set_stream");
        int r = qk_rhf_create(&h->jk, nBas);
        QK_CHECK(r == QK_OK, r, "qsGW: Hartree/exchange builder for nBas = " + std::to_string(nBas));
        r = qk_gw_create(&h->g, nBas, nOrb, nC, nO, nV, nR, nS);
        QK_CHECK(r == QK_OK, r, "qsGW: GW engine: " + std::string(qk_last_error()));
        h->eri = dalloc<double>(N2 * N2);
        for (double** p : {&h->S, &h->P, &h->F, &h->A1, &h->A2, &h->A3}) *p = dalloc<double>(N2);
        for (double** p : {&h->X, &h->c, &h->c2, &h->SC, &h->Tm}) *p = dalloc<double>(N * n);
        for (double** p : {&h->Sig, &h->Fp}) *p = dalloc<double>(n * n);
        h->w = dalloc<double>(n);
        h->work = dalloc<double>(n);
        h->info = dalloc<int>(1);
        up(h->s, h->eri, ERI_AO, N2 * N2);
        up(h->s, h->S, S, N2);
        up(h->s, h->X, X, N * n);
        up(h->s, h->c, cHF, N * n);
        h->hT.assign(T, T + N2);
        h->hV.assign(V, V + N2);
        h->hHc.assign(Hc, Hc + N2);
        h->hP.assign(PHF, PHF + N2);
        for (auto* v : {&h->hJ, &h->hK, &h->hF, &h->hErr, &h->hSigp}) v->assign(N2, 0.0);
        h->hSig.assign(n * n, 0.0);
        h->eGW.assign(eHF, eHF + n);
        h->diis = new Diis((int)N2, h->max_diis);
        QK_HIP(hipStreamSynchronize(h->s));
    });
    if (rc != QK_OK) {
        g_last_error = h->err;
        qsgw_release(h);
        delete h;
        return rc;
    }
    *out = h;
    return QK_OK;
}

int qk_qsgw_destroy(qk_qsgw_handle h) {
    if (!h) return QK_E_ARG;
    qsgw_release(h);
    delete h;
    return QK_OK;
}

const char* qk_qsgw_last_error(qk_qsgw_handle h) { return h ? h->err.c_str() : ""; }

int qk_qsgw_cycle(qk_qsgw_handle h, double eta, int TDA_W, int doSRG, double* eGW, double* c, double* SigC, double* Z,
                  double* E, double* EcRPA, double* EcGM, double* err_max) {
    if (!h || !eGW || !c || !SigC || !Z || !E || !err_max) return QK_E_ARG;
    return guarded(&h->err, [&] {
        const int N = h->nBas, n = h->nOrb, nO = h->nO;
        const size_t N2 = (size_t)N * N;
        hipStream_t s = h->s;
        // Hartree and exchange matrices of the current density (GW/qsRGW.f90:201-202)
        int r = qk_rhf_jk(h->jk, h->eri, (long long)N2 * N2, h->hP.data(), h->hJ.data(), h->hK.data());
        QK_CHECK(r == QK_OK, r, "qsGW: J/K build: " + std::string(qk_rhf_last_error(h->jk)));
        // MO integrals of the current orbitals (:210), dRPA screening at the
        // current energies and excitation densities (:214-220), full
        // self-energy (:256-264)
        QK_HIP(hipStreamSynchronize(s));
        r = qk_gw_set_eri_ao_device(h->g, h->eri, h->c);
        QK_CHECK(r == QK_OK, r, "qsGW: AO->MO block: " + std::string(qk_gw_last_error(h->g)));
        r = qk_gw_screen(h->g, h->eGW.data(), TDA_W);
        QK_CHECK(r == QK_OK, r, "qsGW: screening: " + std::string(qk_gw_last_error(h->g)));
        double ecgm = 0.0, ecrpa = 0.0;
        r = qk_gw_sigma_full(h->g, h->eGW.data(), eta, doSRG, h->hSig.data(), Z, &ecgm, &ecrpa);
        QK_CHECK(r == QK_OK, r, "qsGW: self-energy: " + std::string(qk_gw_last_error(h->g)));
        // Hermitian part (:269), then MOtoAO: SigCp = S c SigC c^T S (:276)
        for (int q = 0; q < n; ++q)
            for (int p = q + 1; p < n; ++p) {
                double& a = h->hSig[p + (size_t)n * q];
                double& b = h->hSig[q + (size_t)n * p];
                const double m = 0.5 * (a + b);
                a = m;
                b = m;
            }
        up(s, h->Sig, h->hSig.data(), (size_t)n * n);
        dgemm(s, false, false, N, n, N, 1.0, h->S, N, h->c, N, 0.0, h->SC, N);
        dgemm(s, false, true, n, N, n, 1.0, h->Sig, n, h->SC, N, 0.0, h->Tm, n);  // Sig (S c)^T, n x N
        dgemm(s, false, false, N, N, n, 1.0, h->SC, N, h->Tm, n, 0.0, h->A1, N);
        down(s, h->hSigp.data(), h->A1, N2);
        QK_HIP(hipStreamSynchronize(s));
        // F = Hc + J + K/2 + SigCp (:280)
        for (size_t k = 0; k < N2; ++k) h->hF[k] = h->hHc[k] + h->hJ[k] + 0.5 * h->hK[k] + h->hSigp[k];
        up(s, h->F, h->hF.data(), N2);
        if (N != n) {
            // project on the kept orbitals: AOtoMO then MOtoAO (:281-284)
            dgemm(s, false, false, N, n, N, 1.0, h->F, N, h->c, N, 0.0, h->Tm, N);
            dgemm(s, true, false, n, n, N, 1.0, h->c, N, h->Tm, N, 0.0, h->Fp, n);
            dgemm(s, false, true, n, N, n, 1.0, h->Fp, n, h->SC, N, 0.0, h->Tm, n);
            dgemm(s, false, false, N, N, n, 1.0, h->SC, N, h->Tm, n, 0.0, h->F, N);
            down(s, h->hF.data(), h->F, N2);
        }
        // commutator F (P S) - (S P) F (:288) and its largest element
        up(s, h->P, h->hP.data(), N2);
        dgemm(s, false, false, N, N, N, 1.0, h->P, N, h->S, N, 0.0, h->A1, N);
        dgemm(s, false, false, N, N, N, 1.0, h->F, N, h->A1, N, 0.0, h->A2, N);
        dgemm(s, false, false, N, N, N, 1.0, h->S, N, h->P, N, 0.0, h->A1, N);
        dgemm(s, false, false, N, N, N, 1.0, h->A1, N, h->F, N, 0.0, h->A3, N);
        std::vector<double> fps(N2), spf(N2);
        down(s, fps.data(), h->A2, N2);
        down(s, spf.data(), h->A3, N2);
        QK_HIP(hipStreamSynchronize(s));
        double emax = 0.0;
        for (size_t k = 0; k < N2; ++k) {
            h->hErr[k] = fps[k] - spf[k];
            emax = std::max(emax, std::fabs(h->hErr[k]));
        }
        // energies (:292-310)
        const double ET = trace_prod(N, h->hP, h->hT), EV = trace_prod(N, h->hP, h->hV);
        const double EJ = 0.5 * trace_prod(N, h->hP, h->hJ), EK = 0.25 * trace_prod(N, h->hP, h->hK);
        // DIIS on F (:314-318)
        if (h->max_diis > 1) h->diis->step(h->hErr.data(), h->hF.data());
        up(s, h->F, h->hF.data(), N2);
        // new orbitals: X^T F X (nBas = nOrb) or c^T F c (:317-333)
        const double* B = N == n ? h->X : h->c;
        dgemm(s, false, false, N, n, N, 1.0, h->F, N, B, N, 0.0, h->Tm, N);
        dgemm(s, true, false, n, n, N, 1.0, B, N, h->Tm, N, 0.0, h->Fp, n);
        syevd(s, h->rb, true, n, h->Fp, h->w, h->work, h->info);
        dgemm(s, false, false, N, n, n, 1.0, B, N, h->Fp, n, 0.0, h->c2, N);
        std::swap(h->c, h->c2);
        // SigC in the new orbitals (:335) and the density (:339)
        up(s, h->A1, h->hSigp.data(), N2);
        dgemm(s, false, false, N, n, N, 1.0, h->A1, N, h->c, N, 0.0, h->Tm, N);
        dgemm(s, true, false, n, n, N, 1.0, h->c, N, h->Tm, N, 0.0, h->Sig, n);
        dgemm(s, false, true, N, N, nO, 2.0, h->c, N, h->c, N, 0.0, h->P, N);
        down(s, h->hP.data(), h->P, N2);
        down(s, h->eGW.data(), h->w, n);
        down(s, SigC, h->Sig, (size_t)n * n);
        down(s, c, h->c, (size_t)N * n);
        QK_HIP(hipStreamSynchronize(s));
        std::memcpy(eGW, h->eGW.data(), sizeof(double) * n);
        E[0] = ET;
        E[1] = EV;
        E[2] = EJ;
        E[3] = EK;
        E[4] = ET + EV + EJ + EK;
        if (EcRPA) *EcRPA = ecrpa;
        if (EcGM) *EcGM = ecgm;
        *err_max = emax;
    });
}

}  // extern "C"
