// Private declarations shared by the translation units of the GW C-ABI:
// qk_gw_api.hip (the single-molecule engine and the per-routine entries),
// qk_gw_unrestricted.hip (UG0W0, evUGW) and qk_gw_batch.hip (the batched
// form).  Not part of the library's interface.
#pragma once
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "qk_kernels.h"
#include "qk_split.h"

namespace qk {

constexpr int QP_MAXIT = 64;       // GW/RGW_QP_graph.f90:31
constexpr double QP_THRESH = 1e-6; // GW/RGW_QP_graph.f90:32
constexpr int QP_SYNC_EVERY = 4;   // Newton sweeps issued between host reads of the active count
constexpr double SRG_FLOW = 500.0; // GW/RG0W0.f90:98

// message of the calling thread's last failed call (qk_last_error)
extern thread_local std::string g_last_error;

template <class T>
T* dalloc(size_t count) {
    if (count == 0) count = 1;
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, count * sizeof(T));
    if (e != hipSuccess) {
        (void)hipGetLastError();
        throw Error{QK_E_ALLOC, "hipMalloc of " + std::to_string(count * sizeof(T)) + " bytes failed: " +
                                    hipGetErrorString(e)};
    }
    return static_cast<T*>(p);
}
template <class T>
void dfree(T*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

inline void rb_check(rocblas_status st, const char* what) {
    if (st != rocblas_status_success) throw Error{QK_E_HIP, std::string(what) + ": rocblas status " + std::to_string((int)st)};
}

// rocSOLVER dsyevd on the stream bound to rb; eigenvalues ascending in w.
void syevd(hipStream_t s, rocblas_handle rb, bool vectors, int n, double* A, double* w, double* work, int* dinfo);

struct GraphSlot {
    const void* key = nullptr;
    hipGraphExec_t exec = nullptr;
};
struct SolveGraphs {
    GraphSlot trd[2], orm[2];
    void clear() {
        for (GraphSlot* g : {&trd[0], &trd[1], &orm[0], &orm[1]}) {
            if (g->exec) (void)hipGraphExecDestroy(g->exec);
            *g = GraphSlot();
        }
    }
};

// Tiny symmetric solve for DIIS (LU with partial pivoting); returns the exact
// 1-norm reciprocal condition number.
double small_solve(int n, std::vector<double> A, std::vector<double> b, std::vector<double>& x);

// utils/DIIS_extrapolation.f90:1-61 (history prepend utils/utils.f90:506-536).
struct Diis {
    int n, max_diis, n_diis = 0;
    std::vector<double> err, e;  // column-major (n, max_diis)
    Diis(int n_, int m_) : n(n_), max_diis(m_), err((size_t)n_ * m_, 0.0), e((size_t)n_ * m_, 0.0) {}
    void step(const double* err_in, double* e_inout) {
        n_diis = std::min(n_diis + 1, max_diis);
        int m = n_diis;
        for (int j = m - 2; j >= 0; --j) {
            std::memcpy(&err[(size_t)n * (j + 1)], &err[(size_t)n * j], sizeof(double) * n);
            std::memcpy(&e[(size_t)n * (j + 1)], &e[(size_t)n * j], sizeof(double) * n);
        }
        std::memcpy(&err[0], err_in, sizeof(double) * n);
        std::memcpy(&e[0], e_inout, sizeof(double) * n);
        int N = m + 1;
        std::vector<double> A((size_t)N * N, 0.0), b(N, 0.0), w;
        for (int i = 0; i < m; ++i)
            for (int j = 0; j < m; ++j) {
                double s = 0.0;
                for (int k = 0; k < n; ++k) s += err[(size_t)n * i + k] * err[(size_t)n * j + k];
                A[i + N * j] = s;
            }
        for (int i = 0; i < m; ++i) { A[i + N * m] = -1.0; A[m + N * i] = -1.0; }
        b[m] = -1.0;
        double rcond = small_solve(N, A, b, w);
        if (rcond > 1e-13) {
            for (int k = 0; k < n; ++k) {
                double s = 0.0;
                for (int j = 0; j < m; ++j) s += w[j] * e[(size_t)n * j + k];
                e_inout[k] = s;
            }
        }
    }
};

}  // namespace qk

struct qk_gw_ctx {
    int nBas = 0, nOrb = 0, nC = 0, nO = 0, nV = 0, nR = 0, nS = 0, o = 0, v = 0;
    hipStream_t s = nullptr;
    rocblas_handle rb = nullptr;
    double *M = nullptr, *K = nullptr, *Cm = nullptr, *rhoT = nullptr;
    // rhoT holds FP32 values (first half of the allocation) after a
    // split-precision contraction, whose accuracy is FP32-level; FP64 otherwise
    bool rho_f32 = false;
    double *e = nullptr, *e0 = nullptr, *eG = nullptr, *dE = nullptr, *sd = nullptr, *Om = nullptr, *Om0 = nullptr;
    double *Sig = nullptr, *Z = nullptr, *eLin = nullptr, *eGW = nullptr, *w = nullptr, *f = nullptr, *df = nullptr;
    double *part = nullptr, *scal = nullptr, *work = nullptr;
    int *nit = nullptr, *done = nullptr, *info = nullptr, *active = nullptr;
    int nchunk = 1;
    bool have_eri = false;
    bool have_screen = false;
    // 0: FP64 MFMA; split precision with FP64 recombination: 1 three BF16 slices
    // (six products), 2 two slices (three products) in the excitation densities
    int precision = 0;
    qk::Split3 Msplit, Wsplit;  // M^T and W slices (mixed precision)
    bool msplit_valid = false;
    // density-fitted integrals (pq|rs) = sum_P B(P,pq) B(P,rs): B[P + naux (p + nOrb q)],
    // Bov[P + naux ia] its (i, a) block, T = Bov W (naux x nS)
    bool use_df = false;
    int naux = 0;
    double *Bdf = nullptr, *Bov = nullptr, *Tdf = nullptr;
    // work of the full self-energy's matrix-core path (qk_gw_sigma_full), kept across calls
    double* sigw = nullptr;
    qk::Split3 Bsplit, BovTsplit, Tsplit;  // B (rows pq), Bov^T (rows P, k = jb), T (rows m)
    bool bsplit_valid = false;
    // eigensolver: 0 = rocSOLVER dsyevd for every RPA (QuAcK's scheme); 1 = refine
    // the previous eigenvectors by Newton steps when the energies moved by at
    // most eig_demax since they were computed, dsyevd otherwise or on failure
    int eig_mode = 0;  // bit 0: refine; bit 1: own tridiagonalisation instead of dsyevd
    double eig_demax = 5e-3;
    double *Zev = nullptr, *tau = nullptr, *trd_work = nullptr, *orm_work = nullptr;  // own full-solve path
    qk::SolveGraphs graphs;  // captured launch sequences of the own full solve
    double *Xv = nullptr, *Ym = nullptr, *Sm = nullptr, *Gm = nullptr, *lam = nullptr, *eig_dbuf = nullptr;
    int *perm = nullptr, *eig_ibuf = nullptr;
    long long* eig_lbuf = nullptr;
    bool have_vec = false;
    int tda_vec = 0;
    // unrestricted pairs (the screening of qk_ug0w0 and qk_evugw): nOb = beta
    // occupied count, -1 for the restricted handle; up holds the pairs of both
    // spins (nS = up.nSt), the energies are (e_alpha, e_beta)
    int nOb = -1;
    qk::UPairs up;
    std::vector<double> e_vec;
    // dsyevd calls, refined solves, Newton steps, fallbacks; fallbacks because a
    // cluster outgrew EIG_KMAX, because merging did not settle, because the
    // steps did not converge; largest cluster met
    long long eig_count[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};  // [8]: full solves by the own tridiagonalisation
    double t_ao = 0.0;
    // [0..4] RPA build, eigensolve, excitation densities, self-energy + QP, total;
    // [5] excitation-density GEMM launches alone, [6] their number
    double t_stage[7] = {0, 0, 0, 0, 0, 0, 0};
    hipEvent_t ev[2] = {nullptr, nullptr};
    hipEvent_t ev2[2] = {nullptr, nullptr};
    double t_gemm_total = 0.0, n_gemm_total = 0.0;  // excitation GEMM ms and launches since creation
    hipEvent_t ev3[4] = {nullptr, nullptr, nullptr, nullptr};
    double t_trd_total = 0.0, n_trd_total = 0.0;    // own tridiagonalisations: ms and count since creation
    double t_stedc_total = 0.0, t_ormtr_total = 0.0; // their dstedc and back-transformation (vector solves)
    bool shared = false;  // stream, rocBLAS handle and the batch slots belong to a qk_gwb handle
    std::string err;
};

namespace qk {

using Ctx = qk_gw_ctx;

void check_sizes(int nBas, int nOrb, int nC, int nO, int nV, int nR, int nS);
void h2d(Ctx* c, double* dst, const double* src, size_t count);
void d2h(Ctx* c, double* dst, const double* src, size_t count);
void alloc_split(Split3& sp, int rows, int K, int nsl = 3);
void free_split(Split3& sp);
// slices per operand of a split-precision mode (1: three, 2: two)
int slices_of(int precision);
// RPA screening at energies e (device; eh the same energies on the host):
// eigenvectors scaled to W in Cm when vectors, Om
void screening(Ctx* c, const double* e, const double* eh, int tda, bool vectors);
// split-precision AO->MO block switch (qk_set_split_aotomo)
bool split_aotomo();
void aotomo_block(hipStream_t s, const double* f, const double* cc, double* M, int N, int n, int nC, int nO, int nR,
                  bool split = false, double* K = nullptr);
void upload_mo_block(hipStream_t s, const double* ERI_MO, double* M, int n, int nC, int nO, int nR);
// rho(p,q,m) (QuAcK layout, host) into rhoT[m + nS (p + n q)] on the device (caller frees)
double* upload_rho(hipStream_t s, const double* rho, int n, int nS);

// Shared stream and rocBLAS handle of the handle-less entry points.
struct Scratch {
    hipStream_t s = nullptr;
    rocblas_handle rb = nullptr;
    Scratch() {
        QK_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        rb_check(rocblas_create_handle(&rb), "rocblas_create_handle");
        rb_check(rocblas_set_stream(rb, s), "rocblas_set_stream");
    }
};
Scratch& scratch();

template <class F>
int guarded(std::string* slot, F&& fn) {
    try {
        fn();
        return QK_OK;
    } catch (const Error& e) {
        if (slot) *slot = e.msg;
        g_last_error = e.msg;
        return e.code;
    } catch (const std::bad_alloc&) {
        if (slot) *slot = "host allocation failed";
        g_last_error = "host allocation failed";
        return QK_E_ALLOC;
    } catch (...) {
        if (slot) *slot = "unknown error";
        g_last_error = "unknown error";
        return QK_E_HIP;
    }
}

void release(Ctx* c);

}  // namespace qk
