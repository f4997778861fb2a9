// Batched G0W0 and evGW (qk_gwb_*): many molecules of one size on one
// stream, one launch per kernel for the whole group, the evGW bookkeeping on
// the device and hipGraph replay of the evGW cycle.
#include "qk_gw_internal.h"

using namespace qk;

// ---- batched G0W0: many molecules of one size, stage by stage -----------------

struct qk_gwb_ctx {
    int nmol = 0;
    hipStream_t s = nullptr;
    rocblas_handle rb = nullptr;
    std::vector<qk_gw_ctx*> mol;
    // contiguous per-molecule arrays with uniform strides (zs), so that one
    // launch serves every molecule (grid z = molecule); the molecules'
    // handles point into them
    double *M = nullptr, *Cm = nullptr, *rhoT = nullptr, *Om = nullptr, *E = nullptr;
    double *K = nullptr, *dE = nullptr, *sd = nullptr, *Om0 = nullptr, *part = nullptr, *scal = nullptr;
    double* vec[10] = {};  // e, e0, eG, Sig, Z, eLin, eGW, w, f, df: nOrb per molecule
    int *info = nullptr, *active = nullptr, *nit = nullptr, *done = nullptr;
    qk::Zs zs;
    int nchunk = 1;
    // evGW state on the device (qk_gwb_evgw): DIIS histories, per molecule
    // [live, cycles, history length, -], Conv, and the live molecules' outputs
    double *hist = nullptr, *conv = nullptr, *outSig = nullptr, *outZ = nullptr, *outOm = nullptr, *outEc = nullptr;
    int* state = nullptr;
    int hist_maxd = 0;
    // hipGraph replay of the evGW cycle (qk_gwb_set_graphs): the launches
    // before the batched eigensolve and those after it, captured on the
    // second cycle of a call and replayed while the parameters stay the same
    int graphs = 0;
    hipGraphExec_t gpre = nullptr, gpost = nullptr;
    std::vector<double> gkey;
    long long graph_cycles = 0, eager_cycles = 0;
    int precision = 0;
    bool have_eri = false;
    double t_stage[4] = {0, 0, 0, 0};  // AO->MO, eigensolves, excitation densities, self-energy + QP
    std::string err;
};

namespace qk {
namespace {

using BCtx = qk_gwb_ctx;

enum { V_E, V_E0, V_EG, V_SIG, V_Z, V_ELIN, V_EGW, V_W, V_F, V_DF };

void drop_graphs(BCtx* b) {
    for (hipGraphExec_t* g : {&b->gpre, &b->gpost}) {
        if (*g) (void)hipGraphExecDestroy(*g);
        *g = nullptr;
    }
    b->gkey.clear();
}

void release_batch(BCtx* b) {
    if (b->s) (void)hipStreamSynchronize(b->s);
    drop_graphs(b);
    for (Ctx* c : b->mol) {
        release(c);
        delete c;
    }
    b->mol.clear();
    for (double* p : {b->M, b->Cm, b->rhoT, b->Om, b->E, b->K, b->dE, b->sd, b->Om0, b->part, b->scal, b->hist, b->conv,
                      b->outSig, b->outZ, b->outOm, b->outEc})
        if (p) (void)hipFree(p);
    for (double* p : b->vec)
        if (p) (void)hipFree(p);
    for (int* p : {b->info, b->active, b->nit, b->done, b->state})
        if (p) (void)hipFree(p);
    if (b->rb) (void)rocblas_destroy_handle(b->rb);
    if (b->s) (void)hipStreamDestroy(b->s);
}

// One strided-batched rocSOLVER dsyevd for every molecule; with check, the
// info flags are read back (the graph-replayed evGW reads them after the cycle)
void syevd_batched(BCtx* b, bool vectors, bool check) {
    const int nS = b->mol[0]->nS;
    const long long S2 = (long long)nS * nS;
    rb_check(rocsolver_dsyevd_strided_batched(b->rb, vectors ? rocblas_evect_original : rocblas_evect_none,
                                              rocblas_fill_upper, nS, b->Cm, nS, S2, b->Om, nS, b->E, nS, b->info,
                                              b->nmol),
             "rocsolver_dsyevd_strided_batched");
    if (check) {
        std::vector<int> info(b->nmol);
        QK_HIP(hipMemcpyAsync(info.data(), b->info, sizeof(int) * b->nmol, hipMemcpyDeviceToHost, b->s));
        QK_HIP(hipStreamSynchronize(b->s));
        for (int m = 0; m < b->nmol; ++m)
            QK_CHECK(info[m] == 0, QK_E_EIG,
                     "rocsolver_dsyevd_strided_batched failed for molecule " + std::to_string(m));
    }
}

// Event-timed region on the batch stream (nothing is timed inside a graph capture)
struct BStage {
    BCtx* b;
    int k;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    BStage(BCtx* b_, int k_) : b(b_), k(k_) {
        if (stream_capturing(b->s)) return;
        QK_HIP(hipEventCreate(&e0));
        QK_HIP(hipEventCreate(&e1));
        QK_HIP(hipEventRecord(e0, b->s));
    }
    ~BStage() {
        if (!e0) return;
        float ms = 0.f;
        if (hipEventRecord(e1, b->s) == hipSuccess && hipEventSynchronize(e1) == hipSuccess &&
            hipEventElapsedTime(&ms, e0, e1) == hipSuccess)
            b->t_stage[k] += ms;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
};

// RPA matrices of every molecule at its energies e (device, nOrb per molecule):
// one launch per kernel for the batch
void rpa_build_batch(BCtx* b, const double* e, int tda) {
    Ctx* c0 = b->mol[0];
    rpa_build(b->s, e, b->K, b->dE, b->sd, b->Cm, b->scal, c0->nC, c0->nO, c0->v, c0->nS, tda, b->zs, b->nmol);
}
void rpa_finalize_batch(BCtx* b, int tda, bool vectors) {
    rpa_finalize(b->s, b->Om, b->Cm, b->sd, b->scal, b->mol[0]->nS, tda, vectors, b->zs, b->nmol);
}

// screening of every molecule at its energies e: RPA build, one batched
// eigensolve, finalisation (W scaling, EcRPA)
void screening_batch(BCtx* b, const double* e, int tda, bool vectors) {
    rpa_build_batch(b, e, tda);
    {
        BStage st(b, 1);
        syevd_batched(b, vectors, true);
    }
    rpa_finalize_batch(b, tda, vectors);
}

void excitation_batch(BCtx* b) {
    BStage st(b, 2);
    Ctx* c0 = b->mol[0];
    const long long n2 = (long long)c0->nOrb * c0->nOrb, nS = c0->nS;
    for (Ctx* c : b->mol) c->rho_f32 = b->precision != 0;
    if (b->precision == 0) {
        Batch bt;
        bt.count = b->nmol;
        bt.zdiv = b->nmol;
        bt.sa0 = nS * nS;
        bt.sb0 = n2 * nS;
        bt.sc0 = n2 * nS;
        dgemm(b->s, true, true, (int)nS, (int)n2, (int)nS, 1.0, b->Cm, nS, b->M, n2, 0.0, b->rhoT, nS, bt);
        return;
    }
    for (Ctx* c : b->mol) {
        if (!c->Msplit.s[0] || c->Msplit.nsl != slices_of(b->precision)) {
            QK_CHECK(!stream_capturing(b->s), QK_E_ARG, "split operands allocated inside a graph capture");
            free_split(c->Wsplit);
            free_split(c->Msplit);
            alloc_split(c->Wsplit, c->nS, c->nS, slices_of(b->precision));
            alloc_split(c->Msplit, (int)n2, c->nS, slices_of(b->precision));
            c->msplit_valid = false;
        }
        if (!c->msplit_valid) {
            split3(c->s, c->M, n2, false, (int)n2, c->nS, c->Msplit);
            c->msplit_valid = true;
        }
        split3(c->s, c->Cm, c->nS, true, c->nS, c->nS, c->Wsplit);
        gemm_bf16x3(c->s, c->nS, (int)n2, c->Wsplit, c->Msplit, reinterpret_cast<float*>(c->rhoT), c->nS);
    }
}

// Diagonal self-energy of every molecule with w = e (the energies in G): Sig, Z, EcGM in scal[2]
void sigma_batch(BCtx* b, const double* e, double eta, int srg) {
    Ctx* c0 = b->mol[0];
    const int n = c0->nOrb;
    sigma_partial(b->s, b->rhoT, b->precision != 0, e, e, b->Om, nullptr, eta, srg, SRG_FLOW, n, c0->nC, c0->nO, c0->nR,
                  c0->nS, b->nchunk, b->part, b->zs, b->nmol);
    sigma_finalize(b->s, b->part, b->nchunk, n, c0->nC, c0->nO, c0->nR, b->vec[V_SIG], b->vec[V_Z], b->scal, b->zs,
                   b->nmol);
}

// GW/RGW_QP_graph.f90:50-92 for every active orbital of every molecule, each
// Newton sweep one launch pair for the batch.  eG: energies in G (eHF in
// G0W0, eOld in evGW); start: eGWlin (G0W0) or eOld (evGW), also the fallback
// of a failed solve.  Eagerly the sweeps stop once no orbital of any molecule
// is active (one host read per sweep); with fixed_sweeps all QP_MAXIT sweeps
// are issued, finished orbitals masked on the device, with the same results
// and no host read (the form a graph captures).
void qp_graph_batch(BCtx* b, double eta, int srg, const double* eG, const double* start, bool fixed_sweeps) {
    Ctx* c0 = b->mol[0];
    const int nm = b->nmol, n = c0->nOrb;
    qp_init(b->s, start, b->vec[V_W], b->vec[V_F], b->nit, b->done, n, c0->nC, c0->nR, b->zs, nm);
    std::vector<int> act(nm, 1);
    for (int it = 0; it < QP_MAXIT; ++it) {
        sigma_partial(b->s, b->rhoT, b->precision != 0, b->vec[V_W], eG, b->Om, b->done, eta, srg, SRG_FLOW, n, c0->nC,
                      c0->nO, c0->nR, c0->nS, b->nchunk, b->part, b->zs, nm);
        qp_update(b->s, b->part, b->nchunk, b->vec[V_E0], b->vec[V_W], b->vec[V_F], b->vec[V_DF], b->nit, b->done, n,
                  QP_THRESH, QP_MAXIT, b->active, b->zs, nm);
        if (fixed_sweeps) continue;
        QK_HIP(hipMemcpyAsync(act.data(), b->active, sizeof(int) * nm, hipMemcpyDeviceToHost, b->s));
        QK_HIP(hipStreamSynchronize(b->s));
        if (std::all_of(act.begin(), act.end(), [](int a) { return a == 0; })) break;
    }
    qp_finish(b->s, b->vec[V_W], b->vec[V_DF], b->nit, start, eG, b->vec[V_EGW], b->vec[V_Z], n, c0->nC, c0->nR,
              QP_MAXIT, b->zs, nm);
}

// Capture what body enqueues on the batch stream into an executable graph
// (thread-local capture mode: other host threads are unaffected).
template <class F>
hipGraphExec_t capture_graph(hipStream_t s, F&& body) {
    QK_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    try {
        body();
    } catch (...) {
        hipGraph_t junk = nullptr;
        (void)hipStreamEndCapture(s, &junk);
        if (junk) (void)hipGraphDestroy(junk);
        throw;
    }
    hipGraph_t g = nullptr;
    QK_HIP(hipStreamEndCapture(s, &g));
    hipGraphExec_t ex = nullptr;
    const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    QK_HIP(e);
    return ex;
}

}  // namespace
}  // namespace qk

extern "C" {

int qk_gwb_create(qk_gwb_handle* h, int nmol, int nBas, int nOrb, int nC, int nO, int nV, int nR, int nS) {
    if (!h || nmol <= 0) return QK_E_ARG;
    *h = nullptr;
    BCtx* b = new (std::nothrow) BCtx();
    if (!b) return QK_E_ALLOC;
    int rc = guarded(&b->err, [&] {
        check_sizes(nBas, nOrb, nC, nO, nV, nR, nS);
        QK_CHECK(nmol <= 65535, QK_E_ARG, "at most 65535 molecules per batch");
        b->nmol = nmol;
        QK_HIP(hipStreamCreateWithFlags(&b->s, hipStreamNonBlocking));
        rb_check(rocblas_create_handle(&b->rb), "rocblas_create_handle");
        rb_check(rocblas_set_stream(b->rb, b->s), "rocblas_set_stream");
        const size_t n = nOrb, S = nS, nm = nmol;
        b->nchunk = sigma_chunks(nS, nOrb - nR - nC);
        b->zs.n = (long long)n;
        b->zs.S = (long long)S;
        b->zs.S2 = (long long)(S * S);
        b->zs.R = (long long)(n * n * S);
        b->zs.P = 3LL * (long long)n * b->nchunk;
        b->zs.c = 8;
        b->zs.i = 1;
        b->M = dalloc<double>(nm * n * n * S);
        b->Cm = dalloc<double>(nm * S * S);
        b->rhoT = dalloc<double>(nm * n * n * S);
        b->Om = dalloc<double>(nm * S);
        b->E = dalloc<double>(nm * S);
        b->K = dalloc<double>(nm * S * S);
        b->dE = dalloc<double>(nm * S);
        b->sd = dalloc<double>(nm * S);
        b->Om0 = dalloc<double>(nm * S);
        b->part = dalloc<double>(nm * (size_t)b->zs.P);
        b->scal = dalloc<double>(nm * 8);
        for (double*& v : b->vec) v = dalloc<double>(nm * n);
        b->info = dalloc<int>(nm);
        b->active = dalloc<int>(nm);
        b->nit = dalloc<int>(nm * n);
        b->done = dalloc<int>(nm * n);
        for (int m = 0; m < nmol; ++m) {
            Ctx* c = new Ctx();
            b->mol.push_back(c);
            c->shared = true;
            c->s = b->s;
            c->rb = b->rb;
            c->nBas = nBas; c->nOrb = nOrb; c->nC = nC; c->nO = nO; c->nV = nV; c->nR = nR; c->nS = nS;
            c->o = nO - nC; c->v = nOrb - nR - nO;
            c->M = b->M + m * n * n * S;
            c->Cm = b->Cm + m * S * S;
            c->rhoT = b->rhoT + m * n * n * S;
            c->Om = b->Om + m * S;
            c->work = b->E + m * S;
            c->info = b->info + m;
            c->active = b->active + m;
            c->K = b->K + m * S * S;
            c->dE = b->dE + m * S;
            c->sd = b->sd + m * S;
            c->Om0 = b->Om0 + m * S;
            c->part = b->part + m * (size_t)b->zs.P;
            c->scal = b->scal + m * 8;
            double** vp[10] = {&c->e, &c->e0, &c->eG, &c->Sig, &c->Z, &c->eLin, &c->eGW, &c->w, &c->f, &c->df};
            for (int k = 0; k < 10; ++k) *vp[k] = b->vec[k] + m * n;
            c->nit = b->nit + m * n;
            c->done = b->done + m * n;
            c->nchunk = b->nchunk;
        }
    });
    if (rc != QK_OK) {
        g_last_error = b->err;
        release_batch(b);
        delete b;
        return rc;
    }
    *h = b;
    return QK_OK;
}

int qk_gwb_destroy(qk_gwb_handle h) {
    if (!h) return QK_E_ARG;
    release_batch(h);
    delete h;
    return QK_OK;
}

const char* qk_gwb_last_error(qk_gwb_handle h) { return h ? h->err.c_str() : g_last_error.c_str(); }

int qk_gwb_set_precision(qk_gwb_handle h, int mode) {
    if (!h || mode < 0 || mode > 2) return QK_E_ARG;
    h->precision = mode;
    for (Ctx* c : h->mol) c->precision = mode;
    return QK_OK;
}

int qk_gwb_set_graphs(qk_gwb_handle h, int on) {
    if (!h) return QK_E_ARG;
    return guarded(&h->err, [&] {
        QK_HIP(hipStreamSynchronize(h->s));
        drop_graphs(h);
        h->graphs = on ? 1 : 0;
    });
}

int qk_gwb_graph_stats(qk_gwb_handle h, long long* counts, int n) {
    if (!h || !counts || n < 0 || n > 2) return QK_E_ARG;
    const long long v[2] = {h->eager_cycles, h->graph_cycles};
    for (int i = 0; i < n; ++i) counts[i] = v[i];
    return QK_OK;
}

int qk_gwb_set_eri_ao_device(qk_gwb_handle h, const double* ERI_AO, long long stride_eri, const double* c,
                             long long stride_c) {
    if (!h || !ERI_AO || !c) return QK_E_ARG;
    return guarded(&h->err, [&] {
        BStage st(h, 0);
        const bool sp = split_aotomo();
        for (int m = 0; m < h->nmol; ++m) {
            Ctx* x = h->mol[m];
            aotomo_block(h->s, ERI_AO + m * stride_eri, c + m * stride_c, x->M, x->nBas, x->nOrb, x->nC, x->nO,
                         x->nR, sp, x->K);
            if (!sp) extract_k(h->s, x->M, x->K, x->nOrb, x->nC, x->nO, x->v, x->nS);
            x->msplit_valid = false;
            x->have_eri = true;
        }
        QK_HIP(hipStreamSynchronize(h->s));
        h->have_eri = true;
    });
}

int qk_gwb_set_eri_mo(qk_gwb_handle h, int mol, const double* ERI_MO) {
    if (!h || !ERI_MO || mol < 0 || mol >= h->nmol) return QK_E_ARG;
    return guarded(&h->err, [&] {
        Ctx* x = h->mol[mol];
        upload_mo_block(h->s, ERI_MO, x->M, x->nOrb, x->nC, x->nO, x->nR);
        extract_k(h->s, x->M, x->K, x->nOrb, x->nC, x->nO, x->v, x->nS);
        QK_HIP(hipStreamSynchronize(h->s));
        x->msplit_valid = false;
        x->have_eri = true;
        h->have_eri = std::all_of(h->mol.begin(), h->mol.end(), [](Ctx* c) { return c->have_eri; });
    });
}

int qk_gwb_g0w0(qk_gwb_handle h, const double* eHF, double eta, int linearize, int TDA_W, int doSRG, double* eGW,
                double* SigC, double* Z, double* Om, double* EcRPA, double* EcGM) {
    if (!h || !eHF || !eGW) return QK_E_ARG;
    return guarded(&h->err, [&] {
        QK_CHECK(h->have_eri, QK_E_ARG, "no integrals for some molecule: call qk_gwb_set_eri_* first");
        for (double& t : h->t_stage) t = 0.0;
        const int n = h->mol[0]->nOrb, nS = h->mol[0]->nS, nm = h->nmol;
        const size_t nn = (size_t)nm * n, nnS = (size_t)nm * nS;
        hipStream_t s = h->s;
        QK_HIP(hipMemcpyAsync(h->vec[V_E0], eHF, sizeof(double) * nn, hipMemcpyHostToDevice, s));
        // GW/RG0W0.f90:116-127: screening at eHF, one batched eigensolve
        screening_batch(h, h->vec[V_E0], TDA_W, true);
        QK_HIP(hipMemcpyAsync(h->Om0, h->Om, sizeof(double) * nnS, hipMemcpyDeviceToDevice, s));
        excitation_batch(h);
        {
            BStage st(h, 3);
            sigma_batch(h, h->vec[V_E0], eta, doSRG);
            lin_qp(s, h->vec[V_E0], h->vec[V_Z], h->vec[V_SIG], h->vec[V_ELIN], n, 1, h->zs, nm);
            if (linearize)
                QK_HIP(hipMemcpyAsync(h->vec[V_EGW], h->vec[V_ELIN], sizeof(double) * nn, hipMemcpyDeviceToDevice, s));
            else
                qp_graph_batch(h, eta, doSRG, h->vec[V_E0], h->vec[V_ELIN], false);
        }
        std::vector<double> sig(nn), sc1((size_t)nm * 8), sc2((size_t)nm * 8);
        QK_HIP(hipMemcpyAsync(sig.data(), h->vec[V_SIG], sizeof(double) * nn, hipMemcpyDeviceToHost, s));
        QK_HIP(hipMemcpyAsync(sc1.data(), h->scal, sizeof(double) * nm * 8, hipMemcpyDeviceToHost, s));  // [2] = EcGM
        // RPA correlation energy at eGW (GW/RG0W0.f90:222-225): eigenvalues only
        screening_batch(h, h->vec[V_EGW], TDA_W, false);
        QK_HIP(hipMemcpyAsync(sc2.data(), h->scal, sizeof(double) * nm * 8, hipMemcpyDeviceToHost, s));
        QK_HIP(hipMemcpyAsync(eGW, h->vec[V_EGW], sizeof(double) * nn, hipMemcpyDeviceToHost, s));
        if (Z) QK_HIP(hipMemcpyAsync(Z, h->vec[V_Z], sizeof(double) * nn, hipMemcpyDeviceToHost, s));
        if (Om) QK_HIP(hipMemcpyAsync(Om, h->Om0, sizeof(double) * nnS, hipMemcpyDeviceToHost, s));
        QK_HIP(hipStreamSynchronize(s));
        if (SigC) std::memcpy(SigC, sig.data(), sizeof(double) * nn);
        for (int m = 0; m < nm; ++m) {
            if (EcRPA) EcRPA[m] = sc2[(size_t)m * 8 + 1];
            if (EcGM) EcGM[m] = sc1[(size_t)m * 8 + 2];
        }
    });
}

int qk_gwb_evgw(qk_gwb_handle h, const double* eHF, const double* eGW_start, int maxSCF, double thresh, int max_diis,
                double eta, int linearize, int TDA_W, int doSRG, double* eGW, double* SigC, double* Z, double* Om,
                double* EcRPA, double* EcGM, int* nSCF_out) {
    if (!h || !eHF || !eGW) return QK_E_ARG;
    int conv_status = QK_OK;
    int rc = guarded(&h->err, [&] {
        QK_CHECK(h->have_eri, QK_E_ARG, "no integrals for some molecule: call qk_gwb_set_eri_* first");
        QK_CHECK(maxSCF >= 0, QK_E_ARG, "maxSCF < 0");
        QK_CHECK(max_diis <= 32, QK_E_ARG, "batched evGW: max_diis > 32");
        for (double& t : h->t_stage) t = 0.0;
        const int n = h->mol[0]->nOrb, nS = h->mol[0]->nS, nm = h->nmol;
        const size_t nn = (size_t)nm * n, nnS = (size_t)nm * nS;
        const int maxd = std::max(max_diis, 1);
        hipStream_t s = h->s;
        if (h->hist_maxd < maxd) {
            QK_HIP(hipStreamSynchronize(s));
            drop_graphs(h);  // the graphs hold the old history buffer
            dfree(h->hist);
            h->hist = dalloc<double>((size_t)nm * 2 * n * maxd);
            h->hist_maxd = maxd;
        }
        if (!h->state) {
            h->conv = dalloc<double>(nm);
            h->outSig = dalloc<double>(nn);
            h->outZ = dalloc<double>(nn);
            h->outOm = dalloc<double>(nnS);
            h->outEc = dalloc<double>((size_t)nm * 2);
            h->state = dalloc<int>((size_t)nm * 4);
        }
        // per molecule: GW/evRGW.f90:104-183 (current energies, DIIS history,
        // cycle count, Conv) on the device; a molecule that has converged
        // keeps its last cycle's outputs and rides along with frozen energies
        std::vector<int> st((size_t)nm * 4, 0), info(nm);
        for (int m = 0; m < nm; ++m) st[(size_t)m * 4] = 1.0 > thresh ? 1 : 0;  // Conv = 1 before the first cycle
        std::vector<double> zeros(std::max(nn, nnS), 0.0);
        QK_HIP(hipMemcpyAsync(h->state, st.data(), sizeof(int) * st.size(), hipMemcpyHostToDevice, s));
        QK_HIP(hipMemcpyAsync(h->vec[V_E], eGW_start ? eGW_start : eHF, sizeof(double) * nn, hipMemcpyHostToDevice, s));
        QK_HIP(hipMemcpyAsync(h->vec[V_E0], eHF, sizeof(double) * nn, hipMemcpyHostToDevice, s));
        QK_HIP(hipMemcpyAsync(h->outSig, zeros.data(), sizeof(double) * nn, hipMemcpyHostToDevice, s));
        QK_HIP(hipMemcpyAsync(h->outZ, zeros.data(), sizeof(double) * nn, hipMemcpyHostToDevice, s));
        QK_HIP(hipMemcpyAsync(h->outOm, zeros.data(), sizeof(double) * nnS, hipMemcpyHostToDevice, s));
        QK_HIP(hipMemcpyAsync(h->outEc, zeros.data(), sizeof(double) * nm * 2, hipMemcpyHostToDevice, s));
        QK_HIP(hipStreamSynchronize(s));
        // the cycle: launches before the batched eigensolve, the eigensolve,
        // launches after it (GW/evRGW.f90:119-183)
        auto pre = [&] { rpa_build_batch(h, h->vec[V_E], TDA_W); };
        auto post = [&](bool fixed_sweeps) {
            rpa_finalize_batch(h, TDA_W, true);
            excitation_batch(h);
            BStage stq(h, 3);
            sigma_batch(h, h->vec[V_E], eta, doSRG);
            if (linearize)
                lin_qp(s, h->vec[V_E0], h->vec[V_Z], h->vec[V_SIG], h->vec[V_EGW], n, 0, h->zs, nm);
            else
                qp_graph_batch(h, eta, doSRG, h->vec[V_E], h->vec[V_E], fixed_sweeps);
            evgw_update(s, h->vec[V_EGW], h->vec[V_E], h->vec[V_SIG], h->vec[V_Z], h->Om, h->scal, h->hist, h->state,
                        h->conv, h->outSig, h->outZ, h->outOm, h->outEc, n, nS, maxd, thresh, maxSCF, h->zs, nm);
        };
        const std::vector<double> key = {eta, (double)linearize, (double)TDA_W, (double)doSRG, (double)maxd, thresh,
                                         (double)maxSCF, (double)h->precision};
        for (int cyc = 0;; ++cyc) {
            // the first cycle of a call runs eagerly: it tunes the GEMM shapes
            // and allocates the split operands, which a capture cannot do
            if (h->graphs && cyc > 0) {
                if (!h->gpre || h->gkey != key) {
                    drop_graphs(h);
                    h->gpre = capture_graph(s, pre);
                    h->gpost = capture_graph(s, [&] { post(true); });
                    h->gkey = key;
                }
                QK_HIP(hipGraphLaunch(h->gpre, s));
                {
                    BStage se(h, 1);
                    syevd_batched(h, true, false);
                }
                QK_HIP(hipGraphLaunch(h->gpost, s));
                ++h->graph_cycles;
            } else {
                pre();
                {
                    BStage se(h, 1);
                    syevd_batched(h, true, false);
                }
                post(false);
                ++h->eager_cycles;
            }
            QK_HIP(hipMemcpyAsync(st.data(), h->state, sizeof(int) * st.size(), hipMemcpyDeviceToHost, s));
            QK_HIP(hipMemcpyAsync(info.data(), h->info, sizeof(int) * nm, hipMemcpyDeviceToHost, s));
            QK_HIP(hipStreamSynchronize(s));
            for (int m = 0; m < nm; ++m)
                QK_CHECK(info[m] == 0, QK_E_EIG,
                         "rocsolver_dsyevd_strided_batched failed for molecule " + std::to_string(m));
            bool any = false;
            for (int m = 0; m < nm; ++m) any = any || st[(size_t)m * 4] != 0;
            if (!any) break;
        }
        std::vector<double> ec((size_t)nm * 2);
        QK_HIP(hipMemcpyAsync(eGW, h->vec[V_E], sizeof(double) * nn, hipMemcpyDeviceToHost, s));
        if (SigC) QK_HIP(hipMemcpyAsync(SigC, h->outSig, sizeof(double) * nn, hipMemcpyDeviceToHost, s));
        if (Z) QK_HIP(hipMemcpyAsync(Z, h->outZ, sizeof(double) * nn, hipMemcpyDeviceToHost, s));
        if (Om) QK_HIP(hipMemcpyAsync(Om, h->outOm, sizeof(double) * nnS, hipMemcpyDeviceToHost, s));
        QK_HIP(hipMemcpyAsync(ec.data(), h->outEc, sizeof(double) * nm * 2, hipMemcpyDeviceToHost, s));
        QK_HIP(hipStreamSynchronize(s));
        bool failed = false;
        for (int m = 0; m < nm; ++m) {
            if (EcRPA) EcRPA[m] = ec[(size_t)m * 2];
            if (EcGM) EcGM[m] = ec[(size_t)m * 2 + 1];
            if (nSCF_out) nSCF_out[m] = st[(size_t)m * 4 + 1];
            failed = failed || st[(size_t)m * 4 + 1] == maxSCF + 1;
        }
        if (failed) {
            conv_status = QK_E_NOCONV;
            h->err = "evGW: convergence failed for some molecule (nSCF = maxSCF + 1)";
        }
    });
    return rc != QK_OK ? rc : conv_status;
}

int qk_gwb_stage_times(qk_gwb_handle h, double* ms, int n) {
    if (!h || !ms || n < 0 || n > 4) return QK_E_ARG;
    for (int i = 0; i < n; ++i) ms[i] = h->t_stage[i];
    return QK_OK;
}

}  // extern "C"
