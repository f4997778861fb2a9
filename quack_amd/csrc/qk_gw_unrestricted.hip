// Unrestricted G0W0 and evGW (qk_ug0w0, qk_evugw; GW/UG0W0.f90:92-150,
// GW/evUGW.f90:103-215, frozen orbitals per spin), and the host-array self-energy and
// graphical QP entries (qk_rgw_self_energy_diag, qk_rgw_qp_graph) that share
// their device work buffers.
#include "qk_gw_internal.h"

namespace qk {
namespace {
// Device work of the diagonal self-energy and the graphical QP solve for n
// orbitals and nch chunks of modes: partial sums, Newton state and the
// orbital-energy vectors, allocated once by their owner.
struct QpWork {
    int n = 0, nch = 0;
    double *part = nullptr, *w = nullptr, *f = nullptr, *df = nullptr, *sig = nullptr, *z = nullptr, *sc = nullptr;
    double *e0 = nullptr, *eG = nullptr, *eL = nullptr, *eo = nullptr;
    int *nit = nullptr, *done = nullptr, *act = nullptr;
    QpWork(int n_, int nch_) : n(n_), nch(nch_) {
        try {
            part = dalloc<double>((size_t)3 * n * nch);
            for (double** p : {&w, &f, &df, &sig, &z, &e0, &eG, &eL, &eo}) *p = dalloc<double>(n);
            sc = dalloc<double>(4);
            nit = dalloc<int>(n);
            done = dalloc<int>(n);
            act = dalloc<int>(1);
        } catch (...) {
            release();
            throw;
        }
    }
    QpWork(const QpWork&) = delete;
    QpWork& operator=(const QpWork&) = delete;
    ~QpWork() { release(); }
    void release() {
        for (double** p : {&part, &w, &f, &df, &sig, &z, &sc, &e0, &eG, &eL, &eo}) dfree(*p);
        dfree(nit);
        dfree(done);
        dfree(act);
    }
};

// Diagonal self-energy (GW/RGW_self_energy_diag.f90 or the SRG form) at the
// host energies e, from device rT (rho(p, q, m) at m + nS (p + n q)) and Om;
// SigC, Z and EcGM to the host.
void sigma_dev(hipStream_t s, QpWork& q, const double* rT, const double* dO, int doSRG, double eta, int nC, int nO,
               int nR, int nS, const double* e, double* SigC, double* Z, double* EcGM) {
    const int n = q.n;
    QK_HIP(hipMemcpyAsync(q.e0, e, sizeof(double) * n, hipMemcpyHostToDevice, s));
    sigma_partial(s, rT, false, q.e0, q.e0, dO, nullptr, eta, doSRG, SRG_FLOW, n, nC, nO, nR, nS, q.nch, q.part);
    sigma_finalize(s, q.part, q.nch, n, nC, nO, nR, q.sig, q.z, q.sc);
    QK_HIP(hipMemcpyAsync(SigC, q.sig, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    QK_HIP(hipMemcpyAsync(Z, q.z, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    double gm = 0.0;
    QK_HIP(hipMemcpyAsync(&gm, q.sc + 2, sizeof(double), hipMemcpyDeviceToHost, s));
    QK_HIP(hipStreamSynchronize(s));
    if (EcGM) *EcGM = gm;
}

// Graphical QP solve (GW/RGW_QP_graph.f90:50-92) from host energies, device rT
// and Om; the Newton sweeps stop once no orbital is left active (read every
// QP_SYNC_EVERY sweeps).  Z is
// read (frozen orbitals keep it) and written on the host.
void qp_dev(hipStream_t s, QpWork& q, const double* rT, const double* dO, int doSRG, double eta, double flow, int nC,
            int nO, int nR, int nS, const double* eHF, const double* eGWlin, const double* eOld, double* eGW,
            double* Z) {
    const int n = q.n;
    QK_HIP(hipMemcpyAsync(q.e0, eHF, sizeof(double) * n, hipMemcpyHostToDevice, s));
    QK_HIP(hipMemcpyAsync(q.eG, eOld, sizeof(double) * n, hipMemcpyHostToDevice, s));
    QK_HIP(hipMemcpyAsync(q.eL, eGWlin, sizeof(double) * n, hipMemcpyHostToDevice, s));
    QK_HIP(hipMemcpyAsync(q.z, Z, sizeof(double) * n, hipMemcpyHostToDevice, s));
    qp_init(s, q.eL, q.w, q.f, q.nit, q.done, n, nC, nR);
    int h_active = 1;
    for (int it = 0; it < QP_MAXIT && h_active > 0;) {
        for (int k = 0; k < QP_SYNC_EVERY && it < QP_MAXIT; ++k, ++it) {
            sigma_partial(s, rT, false, q.w, q.eG, dO, q.done, eta, doSRG, flow, n, nC, nO, nR, nS, q.nch, q.part);
            qp_update(s, q.part, q.nch, q.e0, q.w, q.f, q.df, q.nit, q.done, n, QP_THRESH, QP_MAXIT, q.act);
        }
        QK_HIP(hipMemcpyAsync(&h_active, q.act, sizeof(int), hipMemcpyDeviceToHost, s));
        QK_HIP(hipStreamSynchronize(s));
    }
    qp_finish(s, q.w, q.df, q.nit, q.eL, q.eG, q.eo, q.z, n, nC, nR, QP_MAXIT);
    QK_HIP(hipMemcpyAsync(eGW, q.eo, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    QK_HIP(hipMemcpyAsync(Z, q.z, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    QK_HIP(hipStreamSynchronize(s));
}

// Host rho(p,q,m) (n x n x nS, column-major) and Om uploaded for one
// self-energy or QP call of the C-ABI (qk_rgw_self_energy_diag, qk_rgw_qp_graph).
struct HostRho {
    hipStream_t s;
    double *rT = nullptr, *dO = nullptr;
    HostRho(hipStream_t s_, const double* rho, const double* Om, int n, int nS) : s(s_) {
        try {
            rT = upload_rho(s, rho, n, nS);
            dO = dalloc<double>(nS);
            QK_HIP(hipMemcpyAsync(dO, Om, sizeof(double) * nS, hipMemcpyHostToDevice, s));
        } catch (...) {
            release();
            throw;
        }
    }
    ~HostRho() { release(); }
    void release() {
        (void)hipStreamSynchronize(s);
        dfree(rT);
        dfree(dO);
    }
};

// Orbital counts of qk_ug0w0 / qk_evugw: per spin 0 <= nC < nO < nBas - nR
// (at least one active occupied and one active virtual orbital).
// per spin 0 <= nC <= nO <= nBas - nR with at least one active orbital
// (nC + nR < nBas); a spin may have no active pair (no active occupied or no
// active virtual orbital, e.g. the empty beta spin of a one-electron system),
// as QuAcK's loops allow, but both spins together need one
bool upairs_ok(int nBas, const int* nC, const int* nO, const int* nR) {
    if (nBas <= 0 || !nC || !nO || !nR) return false;
    long long pairs = 0;
    for (int s = 0; s < 2; ++s) {
        if (nC[s] < 0 || nR[s] < 0 || nO[s] < nC[s] || nO[s] > nBas - nR[s] || nC[s] + nR[s] >= nBas) return false;
        pairs += (long long)(nO[s] - nC[s]) * (nBas - nR[s] - nO[s]);
    }
    return pairs > 0;
}

// Eigensolver of the unrestricted screening (qk_set_unrestricted_eigensolver),
// process-wide, and the eigensolver counts of the calling thread's last
// qk_ug0w0 / qk_evugw (qk_unrestricted_eig_stats).
std::atomic<int> g_u_eig_mode{0};
std::atomic<double> g_u_eig_demax{5e-3};
thread_local long long g_u_eig_count[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};

// Unrestricted screening and excitation densities, shared by UG0W0 and evUGW.
// Frozen core (nC) and frozen virtuals (nR) are counted per spin.  Pair index
// ia = a' + v_s i' over the active i, a of spin block s, the beta block after
// the alpha one (LR/phULR_A.f90, LR/phULR_B.f90; UPairs).  The integral
// slabs M_s(pq, k) = (pq|k) of each spin's densities
// (GW/UGW_excitation_density.f90:24-62) and the coupling K = (ia|jb) / 2
// gathered from them stay on the device for the whole calculation.  The dRPA
// A = D + K_u, B = K_u (K_u = 2K) has A - B = D, so each screening is the
// restricted engine's (screening(): C = D^2 + 2 D^1/2 K_u D^1/2, or A under
// TDA, built on the device; the engine's eigensolver, with Newton refinement
// of the previous cycle's vectors when it is switched on), held by a
// screening-only engine handle whose pairs are the unrestricted ones.  It
// leaves Om and XpY^T = W on the device, then rho_s = M_s W / sqrt(2) in the
// restricted kernels' rhoT layout (m contiguous): the unrestricted sums weigh
// rho^2 where the restricted kernels weigh 2 rho^2.
struct UScreen {
    int n, nC[2], nO[2], nR[2], v[2], nSa, nSb, nSt, TDA_W;
    UPairs up;
    size_t S2, n2;
    const double* G[3];
    Ctx* c = nullptr;
    hipStream_t st = nullptr;
    double* de = nullptr;
    double *dM[2] = {nullptr, nullptr}, *dR[2] = {nullptr, nullptr};
    QpWork* q = nullptr;

    UScreen(int n_, const int* nC_, const int* nO_, const int* nR_, const double* Gaaaa, const double* Gaabb,
            const double* Gbbbb, int tda)
        : n(n_), nC{nC_[0], nC_[1]}, nO{nO_[0], nO_[1]}, nR{nR_[0], nR_[1]}, TDA_W(tda), up(n_, nC_, nO_, nR_),
          G{Gaaaa, Gaabb, Gbbbb} {
        for (int sp = 0; sp < 2; ++sp) v[sp] = up.vir(sp);
        nSa = up.nSa;
        nSb = up.nSt - up.nSa;
        nSt = up.nSt;
        QK_CHECK((long long)nSt * nSt < (1LL << 40), QK_E_ARG, "nSa + nSb too large");
        S2 = (size_t)nSt * nSt;
        n2 = (size_t)n * n;
        QK_CHECK((long long)n2 < (1LL << 31), QK_E_ARG, "nBas^2 too large");
        try {
            c = new Ctx();
            c->nBas = c->nOrb = n;
            c->nC = c->nR = 0;  // the restricted pair counts are unused: screening() takes up
            c->nO = nO[0];
            c->nOb = nO[1];
            c->up = up;
            c->nV = n - nO[0];
            c->o = nO[0];
            c->v = n - nO[0];
            c->nS = nSt;
            QK_HIP(hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking));
            rb_check(rocblas_create_handle(&c->rb), "rocblas_create_handle");
            rb_check(rocblas_set_stream(c->rb, c->s), "rocblas_set_stream");
            for (auto* e : {&c->ev[0], &c->ev[1], &c->ev2[0], &c->ev2[1], &c->ev3[0], &c->ev3[1], &c->ev3[2],
                            &c->ev3[3]})
                QK_HIP(hipEventCreate(e));
            st = c->s;
            c->K = dalloc<double>(S2);
            c->Cm = dalloc<double>(S2);
            for (double** p : {&c->dE, &c->sd, &c->Om, &c->work}) *p = dalloc<double>(nSt);
            c->scal = dalloc<double>(8);
            c->info = dalloc<int>(1);
            de = dalloc<double>(2 * (size_t)n);
            std::vector<double> Ms(n2 * nSt);
            for (int sp = 0; sp < 2; ++sp) {
                dM[sp] = dalloc<double>(n2 * nSt);
                dR[sp] = dalloc<double>(n2 * nSt);
                for (int sj = 0; sj < 2; ++sj)
                    for (int j = nC[sj]; j < nO[sj]; ++j)
                        for (int b = 0; b < v[sj]; ++b) {
                            const size_t k = (sj ? nSa : 0) + b + (size_t)v[sj] * (j - nC[sj]);
                            const int Bb = nO[sj] + b;
                            for (int qq = 0; qq < n; ++qq)
                                for (int p = 0; p < n; ++p)
                                    Ms[(p + (size_t)n * qq) + n2 * k] =
                                        sp == 0 ? at(sj == 0 ? 0 : 1, p, j, qq, Bb)
                                                : (sj == 0 ? at(1, j, p, Bb, qq) : at(2, p, j, qq, Bb));
                        }
                QK_HIP(hipMemcpyAsync(dM[sp], Ms.data(), sizeof(double) * n2 * nSt, hipMemcpyHostToDevice, st));
                QK_HIP(hipStreamSynchronize(st));  // Ms is refilled for the next spin
            }
            extract_ku(st, dM[0], dM[1], c->K, up);
            const int rc = qk_gw_set_eigensolver(c, g_u_eig_mode.load(), g_u_eig_demax.load());
            QK_CHECK(rc == QK_OK, rc, "unrestricted screening: " + c->err);
            q = new QpWork(n, sigma_chunks(nSt, n));
        } catch (...) {
            release();
            throw;
        }
    }
    UScreen(const UScreen&) = delete;
    UScreen& operator=(const UScreen&) = delete;
    ~UScreen() { release(); }
    void release() {
        if (st) (void)hipStreamSynchronize(st);
        delete q;
        q = nullptr;
        for (double** p : {&de, &dM[0], &dM[1], &dR[0], &dR[1]}) dfree(*p);
        if (c) {
            std::copy(c->eig_count, c->eig_count + 9, g_u_eig_count);
            qk::release(c);
            delete c;
            c = nullptr;
        }
        st = nullptr;
    }
    double at(int k, int p, int qq, int r, int t) const {  // ERI_k(p, q, r, t), 0-based
        return G[k][p + (size_t)n * (qq + (size_t)n * (r + (size_t)n * t))];
    }
    // LR/phULR.f90 at energies en (n x 2, host); returns EcRPA, leaves Om and
    // W = XpY^T on the device
    double rpa(const double* en, const char* who) {
        if (!TDA_W) {
            // LR/phRLR.f90:57-58 (the unrestricted solver shares it): A - B
            // must be positive definite; here it is diag(e_a - e_i)
            double dmin = HUGE_VAL;
            for (int sp = 0; sp < 2; ++sp)
                for (int i = nC[sp]; i < nO[sp]; ++i)
                    for (int a = nO[sp]; a < n - nR[sp]; ++a)
                        dmin = std::min(dmin, en[(size_t)n * sp + a] - en[(size_t)n * sp + i]);
            QK_CHECK(dmin > 0.0, QK_E_UNSTABLE,
                     std::string(who) + ": A-B is not positive definite (smallest eigenvalue " +
                         std::to_string(dmin) + "): linear-response instability");
        }
        QK_HIP(hipMemcpyAsync(de, en, sizeof(double) * 2 * n, hipMemcpyHostToDevice, st));
        screening(c, de, en, TDA_W, true);
        double ec = 0.0;
        QK_HIP(hipMemcpyAsync(&ec, c->scal + 1, sizeof(double), hipMemcpyDeviceToHost, st));
        QK_HIP(hipStreamSynchronize(st));
        return ec;
    }
    // GW/UGW_excitation_density.f90 for both spins from the last rpa():
    // rho_s(m, pq) = sum_k W(k, m) M_s(pq, k) / sqrt(2)
    void densities() {
        for (int sp = 0; sp < 2; ++sp)
            dgemm(st, true, true, nSt, (int)n2, nSt, 1.0 / std::sqrt(2.0), c->Cm, nSt, dM[sp], n2, 0.0, dR[sp], nSt);
    }
    void om_host(double* out) {
        QK_HIP(hipMemcpyAsync(out, c->Om, sizeof(double) * nSt, hipMemcpyDeviceToHost, st));
        QK_HIP(hipStreamSynchronize(st));
    }
    // per spin: diagonal self-energy at e_s (GW/UGW_self_energy_diag.f90 or the SRG form);
    // EcGM per spin, halved (the restricted sum carries 4 rho^2, the unrestricted rho^2, x2 here)
    void sigma(int sp, const double* es, double eta, int doSRG, double* sig, double* z, double* ecgm) {
        double gm = 0.0;
        sigma_dev(st, *q, dR[sp], c->Om, doSRG, eta, nC[sp], nO[sp], nR[sp], nSt, es, sig, z, &gm);
        if (ecgm) *ecgm = 0.5 * gm;
    }
    // GW/UGW_QP_graph.f90 for spin sp
    void qp(int sp, int doSRG, double eta, const double* eHF_s, const double* eLin, const double* eOld, double* eGW_s,
            double* z) {
        qp_dev(st, *q, dR[sp], c->Om, doSRG, eta, SRG_FLOW, nC[sp], nO[sp], nR[sp], nSt, eHF_s, eLin, eOld, eGW_s, z);
    }
    // One block of the static BSE kernels (GW/UGW_phBSE_static_kernel_A.f90,
    // _B.f90) from the last densities():
    //   G[(x + X1 x2) + X1 X2 (y + Y1 y2)] = 4 sum_m w_m rhoX(px0 + x, qx0 + x2, m) rhoY(py0 + y, qy0 + y2, m)
    // with w_m = Om_m/(Om_m^2 + eta^2) and rho_s = sqrt(2) dR[s] (the unrestricted
    // kernels carry 2 rho rho), as one FP64 matrix-core product over the modes.
    std::vector<double> kernel_block(double eta, int sx, int px0, int X1, int qx0, int X2, int sy, int py0, int Y1,
                                     int qy0, int Y2) {
        const size_t nx = (size_t)X1 * X2, ny = (size_t)Y1 * Y2;
        std::vector<double> h(nx * ny);
        if (nx == 0 || ny == 0) return h;  // a spin block without pairs
        double *X = nullptr, *Y = nullptr, *G = nullptr;
        try {
            X = dalloc<double>((size_t)nSt * nx);
            Y = dalloc<double>((size_t)nSt * ny);
            G = dalloc<double>(nx * ny);
            rho_block(st, dR[sx], c->Om, eta, 1, n, nSt, px0, X1, qx0, X2, X);
            rho_block(st, dR[sy], c->Om, eta, 0, n, nSt, py0, Y1, qy0, Y2, Y);
            dgemm(st, true, false, (int)nx, (int)ny, nSt, 4.0, X, nSt, Y, nSt, 0.0, G, (long long)nx);
            QK_HIP(hipMemcpyAsync(h.data(), G, sizeof(double) * h.size(), hipMemcpyDeviceToHost, st));
            QK_HIP(hipStreamSynchronize(st));
        } catch (...) {
            (void)hipStreamSynchronize(st);
            dfree(X); dfree(Y); dfree(G);
            throw;
        }
        dfree(X); dfree(Y); dfree(G);
        return h;
    }
};
}  // namespace
}  // namespace qk

using namespace qk;

extern "C" {

int qk_rgw_self_energy_diag(int doSRG, double eta, int nOrb, int nC, int nO, int nV, int nR, int nS, const double* e,
                            const double* Om, const double* rho, double* EcGM, double* SigC, double* Z) {
    if (!e || !Om || !rho || !SigC || !Z) return QK_E_ARG;
    return guarded(nullptr, [&] {
        check_sizes(nOrb, nOrb, nC, nO, nV, nR, nS);
        hipStream_t s = scratch().s;
        HostRho r(s, rho, Om, nOrb, nS);
        QpWork q(nOrb, sigma_chunks(nS, nOrb - nR - nC));
        sigma_dev(s, q, r.rT, r.dO, doSRG, eta, nC, nO, nR, nS, e, SigC, Z, EcGM);
    });
}

int qk_rgw_qp_graph(int doSRG, double eta, double flow, int nOrb, int nC, int nO, int nV, int nR, int nS,
                    const double* eHF, const double* Om, const double* rho, const double* eGWlin, const double* eOld,
                    double* eGW, double* Z) {
    if (!eHF || !Om || !rho || !eGWlin || !eOld || !eGW || !Z) return QK_E_ARG;
    return guarded(nullptr, [&] {
        check_sizes(nOrb, nOrb, nC, nO, nV, nR, nS);
        hipStream_t s = scratch().s;
        HostRho r(s, rho, Om, nOrb, nS);
        QpWork q(nOrb, sigma_chunks(nS, nOrb - nR - nC));
        qp_dev(s, q, r.rT, r.dO, doSRG, eta, flow, nC, nO, nR, nS, eHF, eGWlin, eOld, eGW, Z);
    });
}


int qk_ug0w0(int nBas, const int* nC, const int* nO, const int* nR, const double* ERI_aaaa, const double* ERI_aabb,
             const double* ERI_bbbb, const double* eHF, double eta, int linearize, int TDA_W, int doSRG, double* eGW,
             double* SigC, double* Z, double* Om, double* EcRPA, double* EcGM) {
    if (!upairs_ok(nBas, nC, nO, nR) || !ERI_aaaa || !ERI_aabb || !ERI_bbbb || !eHF || !eGW || !SigC || !Z)
        return QK_E_ARG;
    return guarded(nullptr, [&] {
        const int n = nBas;
        UScreen u(n, nC, nO, nR, ERI_aaaa, ERI_aabb, ERI_bbbb, TDA_W);
        u.rpa(eHF, "qk_ug0w0");
        u.densities();
        if (Om) u.om_host(Om);
        // GW/UG0W0.f90:104-134 per spin: self-energy at eHF, linearised energies,
        // graphical QP with the linearised guess and G from eHF
        for (int sp = 0; sp < 2; ++sp) {
            const double* es = eHF + (size_t)n * sp;
            double *sig = SigC + (size_t)n * sp, *z = Z + (size_t)n * sp, *eg = eGW + (size_t)n * sp;
            u.sigma(sp, es, eta, doSRG, sig, z, EcGM ? EcGM + sp : nullptr);
            std::vector<double> lin(n);
            for (int p = 0; p < n; ++p) lin[p] = es[p] + z[p] * sig[p];
            if (linearize) std::memcpy(eg, lin.data(), sizeof(double) * n);
            else u.qp(sp, doSRG, eta, es, lin.data(), es, eg, z);
        }
        // RPA correlation energy at eGW (GW/UG0W0.f90:140-142)
        const double ec = u.rpa(eGW, "qk_ug0w0 (RPA at eGW)");
        if (EcRPA) *EcRPA = ec;
    });
}

int qk_evugw(int nBas, const int* nC, const int* nO, const int* nR, const double* ERI_aaaa, const double* ERI_aabb,
             const double* ERI_bbbb, const double* eHF, int maxSCF, double thresh, int max_diis, double eta,
             int linearize, int TDA_W, int doSRG, double* eGW, double* SigC, double* Z, double* Om, double* EcRPA,
             double* EcGM, int* nSCF_out) {
    if (!upairs_ok(nBas, nC, nO, nR) || !ERI_aaaa || !ERI_aabb || !ERI_bbbb || !eHF || !eGW || !SigC || !Z ||
        maxSCF < 0)
        return QK_E_ARG;
    int conv_status = QK_OK;
    int rc = guarded(nullptr, [&] {
        const int n = nBas;
        const size_t n2s = (size_t)n * 2;
        UScreen u(n, nC, nO, nR, ERI_aaaa, ERI_aabb, ERI_bbbb, TDA_W);
        // GW/evUGW.f90:103-215: eGW = eOld = eHF, Z = 1; per cycle screening at eGW,
        // self-energy at eGW, QP energies (linearised: eHF + SigC; graphical: guess and
        // G from eOld), Conv = max |eGW - eOld| over both spins, DIIS per spin with
        // one shared history length
        std::vector<double> e(eHF, eHF + n2s), eOld = e, err(n2s);
        for (size_t p = 0; p < n2s; ++p) Z[p] = 1.0;
        Diis diis[2] = {Diis(n, std::max(max_diis, 1)), Diis(n, std::max(max_diis, 1))};
        int nSCF = 0;
        double Conv = 1.0, ecrpa = 0.0, ecgm[2] = {0.0, 0.0};
        while (Conv > thresh && nSCF <= maxSCF) {
            ecrpa = u.rpa(e.data(), "qk_evugw");
            u.densities();
            for (int sp = 0; sp < 2; ++sp) {
                const size_t o = (size_t)n * sp;
                u.sigma(sp, e.data() + o, eta, doSRG, SigC + o, Z + o, ecgm + sp);
                if (linearize) {
                    for (int p = 0; p < n; ++p) e[o + p] = eHF[o + p] + SigC[o + p];
                } else {
                    std::vector<double> eg(n);
                    u.qp(sp, doSRG, eta, eHF + o, eOld.data() + o, eOld.data() + o, eg.data(), Z + o);
                    std::memcpy(e.data() + o, eg.data(), sizeof(double) * n);
                }
            }
            Conv = 0.0;
            for (size_t p = 0; p < n2s; ++p) {
                err[p] = e[p] - eOld[p];
                Conv = std::max(Conv, std::fabs(err[p]));
            }
            if (max_diis > 1)
                for (int sp = 0; sp < 2; ++sp) diis[sp].step(err.data() + (size_t)n * sp, e.data() + (size_t)n * sp);
            eOld = e;
            ++nSCF;
        }
        std::memcpy(eGW, e.data(), sizeof(double) * n2s);
        if (Om) u.om_host(Om);
        if (EcRPA) *EcRPA = ecrpa;
        if (EcGM) {
            EcGM[0] = ecgm[0];
            EcGM[1] = ecgm[1];
        }
        if (nSCF_out) *nSCF_out = nSCF;
        if (nSCF == maxSCF + 1) {
            conv_status = QK_E_NOCONV;
            g_last_error = "evUGW: convergence failed after " + std::to_string(nSCF) + " cycles";
        }
    });
    return rc != QK_OK ? rc : conv_status;
}

int qk_ugw_phbse(int nBas, const int* nC, const int* nO, const int* nR, const double* ERI_aaaa, const double* ERI_aabb,
                 const double* ERI_bbbb, const double* eW, const double* eGW, double eta, int TDA_W, int TDA,
                 int spin_conserved, int spin_flip, int exchange_kernel, double* EcBSE, double* OmSC, double* OmSF) {
    if (!upairs_ok(nBas, nC, nO, nR) || !ERI_aaaa || !ERI_aabb || !ERI_bbbb || !eW || !eGW || !EcBSE ||
        (spin_conserved && !OmSC) || (spin_flip && !OmSF))
        return QK_E_ARG;
    return guarded(nullptr, [&] {
        const int n = nBas;
        UScreen u(n, nC, nO, nR, ERI_aaaa, ERI_aabb, ERI_bbbb, TDA_W);
        // GW/UGW_phBSE.f90:90-100: spin-conserved dRPA screening at eW and its densities
        u.rpa(eW, "qk_ugw_phbse");
        u.densities();
        const double* Gk[3] = {ERI_aaaa, ERI_aabb, ERI_bbbb};
        auto G = [&](int k, int p, int q, int r, int t) {  // ERI_k(p, q, r, s), 0-based
            return Gk[k][p + (size_t)n * (q + (size_t)n * (r + (size_t)n * t))];
        };
        const double* eg[2] = {eGW, eGW + n};
        const int o[2] = {nO[0] - nC[0], nO[1] - nC[1]};
        const int v[2] = {n - nR[0] - nO[0], n - nR[1] - nO[1]};
        EcBSE[0] = EcBSE[1] = 0.0;
        // GW/UGW_phBSE.f90:108-140 (ispin 1) and :150-182 (ispin 2): the BSE at eGW
        // (phULR_A/B with dRPA = false, LR/phULR_A.f90, LR/phULR_B.f90) plus the
        // static kernels, each solved by phULR (the general solver of qk_phrlr)
        for (int ispin = 1; ispin <= 2; ++ispin) {
            if ((ispin == 1 && !spin_conserved) || (ispin == 2 && !spin_flip)) continue;
            // pair blocks: (occupied spin, virtual spin) of the first and second block
            const int so[2] = {0, 1};
            const int sv[2] = {ispin == 1 ? 0 : 1, ispin == 1 ? 1 : 0};
            const int nb0 = o[so[0]] * v[sv[0]], nb1 = o[so[1]] * v[sv[1]], N = nb0 + nb1;
            const int off[2] = {0, nb0};
            if (N == 0) continue;  // no pair in this manifold: EcBSE stays 0
            std::vector<double> A((size_t)N * N, 0.0), B((size_t)N * N, 0.0), XpY((size_t)N * N), XmY((size_t)N * N);
            for (int bI = 0; bI < 2; ++bI) {
                const int si = so[bI], sa = sv[bI], oi = o[si], va = v[sa];
                // KA on the diagonal blocks: rho_si(i, j) rho_sa(a, b)
                std::vector<double> ka = u.kernel_block(eta, si, nC[si], oi, nC[si], oi, sa, nO[sa], va, nO[sa], va);
                for (int i = 0; i < oi; ++i)
                    for (int a = 0; a < va; ++a)
                        for (int j = 0; j < oi; ++j)
                            for (int b = 0; b < va; ++b) {
                                const size_t ia = off[bI] + a + (size_t)va * i, jb = off[bI] + b + (size_t)va * j;
                                const int I = nC[si] + i, J = nC[si] + j, Aa = nO[sa] + a, Bb = nO[sa] + b;
                                double x = ka[(i + (size_t)oi * j) + (size_t)oi * oi * (a + (size_t)va * b)];
                                if (ispin == 1) {
                                    const int k = si == 0 ? 0 : 2;  // aaaa or bbbb
                                    x += G(k, I, Bb, Aa, J) - G(k, I, Bb, J, Aa);
                                } else {
                                    x -= si == 0 ? G(1, I, Bb, J, Aa) : G(1, Bb, I, Aa, J);
                                }
                                if (i == j && a == b) x += eg[sa][Aa] - eg[si][I];
                                A[ia + (size_t)N * jb] = x;
                            }
            }
            if (ispin == 1) {
                // the coupling blocks of A: lam ERI_aabb(i,b,a,j) and ERI_aabb(b,i,j,a)
                for (int i = 0; i < o[0]; ++i)
                    for (int a = 0; a < v[0]; ++a)
                        for (int j = 0; j < o[1]; ++j)
                            for (int b = 0; b < v[1]; ++b) {
                                const size_t ia = a + (size_t)v[0] * i, jb = nb0 + b + (size_t)v[1] * j;
                                const int I = nC[0] + i, Aa = nO[0] + a, J = nC[1] + j, Bb = nO[1] + b;
                                A[ia + (size_t)N * jb] = G(1, I, Bb, Aa, J);
                                A[jb + (size_t)N * ia] = G(1, Aa, J, I, Bb);
                            }
            }
            if (!TDA) {
                if (ispin == 1) {
                    for (int bI = 0; bI < 2; ++bI) {
                        const int sp = bI, oi = o[sp], va = v[sp];
                        std::vector<double> kb =
                            u.kernel_block(eta, sp, nC[sp], oi, nO[sp], va, sp, nO[sp], va, nC[sp], oi);
                        const int k = sp == 0 ? 0 : 2;
                        for (int i = 0; i < oi; ++i)
                            for (int a = 0; a < va; ++a)
                                for (int j = 0; j < oi; ++j)
                                    for (int b = 0; b < va; ++b) {
                                        const size_t ia = off[bI] + a + (size_t)va * i, jb = off[bI] + b + (size_t)va * j;
                                        const int I = nC[sp] + i, J = nC[sp] + j, Aa = nO[sp] + a, Bb = nO[sp] + b;
                                        B[ia + (size_t)N * jb] =
                                            kb[(i + (size_t)oi * b) + (size_t)oi * va * (a + (size_t)va * j)] +
                                            G(k, I, J, Aa, Bb) - G(k, I, J, Bb, Aa);
                                    }
                    }
                    for (int i = 0; i < o[0]; ++i)
                        for (int a = 0; a < v[0]; ++a)
                            for (int j = 0; j < o[1]; ++j)
                                for (int b = 0; b < v[1]; ++b) {
                                    const size_t ia = a + (size_t)v[0] * i, jb = nb0 + b + (size_t)v[1] * j;
                                    const int I = nC[0] + i, Aa = nO[0] + a, J = nC[1] + j, Bb = nO[1] + b;
                                    B[ia + (size_t)N * jb] = G(1, I, J, Aa, Bb);
                                    B[jb + (size_t)N * ia] = G(1, I, J, Aa, Bb);
                                }
                } else {
                    // (alpha i -> beta a, beta j -> alpha b) and (beta i -> alpha a, alpha j -> beta b)
                    std::vector<double> k01 =
                        u.kernel_block(eta, 0, nC[0], o[0], nO[0], v[0], 1, nO[1], v[1], nC[1], o[1]);
                    std::vector<double> k10 =
                        u.kernel_block(eta, 1, nC[1], o[1], nO[1], v[1], 0, nO[0], v[0], nC[0], o[0]);
                    for (int i = 0; i < o[0]; ++i)
                        for (int a = 0; a < v[1]; ++a)
                            for (int j = 0; j < o[1]; ++j)
                                for (int b = 0; b < v[0]; ++b) {
                                    const size_t ia = a + (size_t)v[1] * i, jb = nb0 + b + (size_t)v[0] * j;
                                    const int I = nC[0] + i, Aa = nO[1] + a, J = nC[1] + j, Bb = nO[0] + b;
                                    B[ia + (size_t)N * jb] =
                                        k01[(i + (size_t)o[0] * b) + (size_t)o[0] * v[0] * (a + (size_t)v[1] * j)] -
                                        G(1, I, J, Bb, Aa);
                                }
                    for (int i = 0; i < o[1]; ++i)
                        for (int a = 0; a < v[0]; ++a)
                            for (int j = 0; j < o[0]; ++j)
                                for (int b = 0; b < v[1]; ++b) {
                                    const size_t ia = nb0 + a + (size_t)v[0] * i, jb = b + (size_t)v[1] * j;
                                    const int I = nC[1] + i, Aa = nO[0] + a, J = nC[0] + j, Bb = nO[1] + b;
                                    B[ia + (size_t)N * jb] =
                                        k10[(i + (size_t)o[1] * b) + (size_t)o[1] * v[1] * (a + (size_t)v[0] * j)] -
                                        G(1, J, I, Aa, Bb);
                                }
                }
            }
            double ec = 0.0;
            const int r = qk_phrlr(TDA, N, A.data(), B.data(), &ec, ispin == 1 ? OmSC : OmSF, XpY.data(), XmY.data());
            QK_CHECK(r == QK_OK, r, "qk_ugw_phbse: phULR failed: " + g_last_error);
            EcBSE[ispin - 1] = ec;
        }
        if (exchange_kernel) {  // GW/UGW_phBSE.f90:189-193
            EcBSE[0] *= 0.5;
            EcBSE[1] *= 0.5;
        }
    });
}

int qk_set_unrestricted_eigensolver(int mode, double de_max) {
    if (mode < 0 || mode > 3 || !(de_max >= 0.0)) return QK_E_ARG;
    qk::g_u_eig_mode.store(mode);
    qk::g_u_eig_demax.store(de_max);
    return QK_OK;
}

int qk_unrestricted_eig_stats(long long* counts, int n) {
    if (!counts || n < 0 || n > 9) return QK_E_ARG;
    for (int i = 0; i < n; ++i) counts[i] = qk::g_u_eig_count[i];
    return QK_OK;
}

}  // extern "C"
