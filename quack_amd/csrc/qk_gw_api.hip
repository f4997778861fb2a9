// C-ABI of the restricted G0W0/evGW engine: device-resident handle plus the
// per-routine entry points declared in include/quack_amd.h.
//
// Control flow follows GW/RG0W0.f90:88-233 and GW/evRGW.f90:104-200.  The
// screening uses that, with dRPA coupling (isp_W = 1, dRPA_W = .true.,
// GW/RG0W0.f90:88-89), A and B share the element 2(ia|jb)
// (LR/phRLR_A.f90:160, LR/phRLR_B.f90:290), so A-B = diag(e_a-e_i) and
// (A-B)^1/2 (A+B) (A-B)^1/2 = D^2 + 4 D^1/2 K D^1/2: one symmetric eigensolve
// instead of the two of LR/phRLR.f90:54,67, and XpY/XmY follow from its
// eigenvectors by diagonal scalings (LR/phRLR.f90:77-81).
#include <chrono>
#include <cstdio>
#include <cstdlib>

#include "qk_gw_internal.h"

namespace qk {

thread_local std::string g_last_error;

// rocSOLVER dsyevd on the stream bound to rb; eigenvalues ascending in w.
void syevd(hipStream_t s, rocblas_handle rb, bool vectors, int n, double* A, double* w, double* work, int* dinfo) {
    rb_check(rocsolver_dsyevd(rb, vectors ? rocblas_evect_original : rocblas_evect_none, rocblas_fill_upper, n, A, n,
                              w, work, dinfo),
             "rocsolver_dsyevd");
    int info = 0;
    QK_HIP(hipMemcpyAsync(&info, dinfo, sizeof(int), hipMemcpyDeviceToHost, s));
    QK_HIP(hipStreamSynchronize(s));
    QK_CHECK(info == 0, QK_E_EIG, "rocsolver_dsyevd failed, info = " + std::to_string(info));
}

// Tiny symmetric solve for DIIS: LU with partial pivoting; rcond is the exact
// 1-norm reciprocal condition number (what dsysvx estimates,
// utils/wrap_lapack.f90:560-598) from the explicit inverse of the (<= 33)x(<= 33) matrix.
double small_solve(int n, std::vector<double> A, std::vector<double> b, std::vector<double>& x) {
    std::vector<double> Ainv(n * n, 0.0), LU = A;
    std::vector<int> piv(n);
    double anorm = 0.0;
    for (int j = 0; j < n; ++j) {
        double s = 0.0;
        for (int i = 0; i < n; ++i) s += std::fabs(A[i + n * j]);
        anorm = std::max(anorm, s);
    }
    for (int k = 0; k < n; ++k) {
        int p = k;
        for (int i = k + 1; i < n; ++i)
            if (std::fabs(LU[i + n * k]) > std::fabs(LU[p + n * k])) p = i;
        piv[k] = p;
        if (LU[p + n * k] == 0.0) { x.assign(n, 0.0); return 0.0; }
        if (p != k)
            for (int j = 0; j < n; ++j) std::swap(LU[k + n * j], LU[p + n * j]);
        for (int i = k + 1; i < n; ++i) {
            LU[i + n * k] /= LU[k + n * k];
            for (int j = k + 1; j < n; ++j) LU[i + n * j] -= LU[i + n * k] * LU[k + n * j];
        }
    }
    auto lusolve = [&](std::vector<double>& r) {
        for (int k = 0; k < n; ++k) std::swap(r[k], r[piv[k]]);
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < i; ++j) r[i] -= LU[i + n * j] * r[j];
        for (int i = n - 1; i >= 0; --i) {
            for (int j = i + 1; j < n; ++j) r[i] -= LU[i + n * j] * r[j];
            r[i] /= LU[i + n * i];
        }
    };
    double inorm = 0.0;
    for (int j = 0; j < n; ++j) {
        std::vector<double> ej(n, 0.0);
        ej[j] = 1.0;
        lusolve(ej);
        double s = 0.0;
        for (int i = 0; i < n; ++i) s += std::fabs(ej[i]);
        inorm = std::max(inorm, s);
    }
    x = b;
    lusolve(x);
    if (!(anorm > 0.0) || !(inorm > 0.0) || !std::isfinite(inorm)) return 0.0;
    return 1.0 / (anorm * inorm);
}

namespace {

// hipGraph replay of the own full solve's launch sequences (process-wide
// switch, qk_set_graphs; off by default).  The tridiagonalisation (two
// launches per column) and the back-transformation contain only the
// library's own kernels, with fixed arguments for a given matrix buffer, so
// each is captured once per buffer and then replayed by one graph launch.  The
// solve swaps the matrix and eigenvector buffers, so a handle meets two keys.
// Measured at n = 5600 on MI355X (ROCm 7.2): 142 ms per tridiagonalisation
// replayed against 137 ms launched on the stream, hence off by default.
std::atomic<int> g_graphs{0};

// Run body on stream s: replayed from the slot keyed by key, captured into a
// free slot on first use (thread-local capture mode, so other host threads
// working on other streams are unaffected), or run eagerly when the slots are
// taken or graphs are off.
template <class F>
void run_graph(hipStream_t s, GraphSlot* slots, const void* key, F&& body) {
    if (!g_graphs.load()) {
        body();
        return;
    }
    for (int i = 0; i < 2; ++i)
        if (slots[i].exec && slots[i].key == key) {
            QK_HIP(hipGraphLaunch(slots[i].exec, s));
            return;
        }
    GraphSlot* slot = !slots[0].exec ? &slots[0] : !slots[1].exec ? &slots[1] : nullptr;
    if (!slot) {
        body();
        return;
    }
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
    slot->key = key;
    slot->exec = ex;
    QK_HIP(hipGraphLaunch(ex, s));
}

// Full eigen-decomposition without dsyevd: own Householder tridiagonalisation
// (qk_sytrd.hip), rocSOLVER's divide-and-conquer dstedc on the tridiagonal
// (also for eigenvalues only), the own blocked back-transformation with the
// reflectors.  Same outputs as syevd(): w ascending, A = eigenvectors.
void syevd_own(hipStream_t s, rocblas_handle rb, bool vectors, int n, double*& A, double* w, double* E, double* tau,
               double*& Z, double* work, double* owork, int* dinfo, hipEvent_t* tev = nullptr,
               SolveGraphs* gr = nullptr) {
    if (tev) QK_HIP(hipEventRecord(tev[0], s));
    double* Ain = A;
    auto trd = [&] { sytrd_upper(s, n, Ain, w, E, tau, work, 32); };
    if (gr) run_graph(s, gr->trd, Ain, trd);
    else trd();
    if (tev) QK_HIP(hipEventRecord(tev[1], s));
    if (vectors) {
        rb_check(rocsolver_dstedc(rb, rocblas_evect_tridiagonal, n, w, E, Z, n, dinfo), "rocsolver_dstedc");
        if (tev) QK_HIP(hipEventRecord(tev[2], s));
        double* Zin = Z;
        auto orm = [&] { ormtr_upper(s, n, Ain, tau, Zin, owork); };
        if (gr) run_graph(s, gr->orm, Ain, orm);
        else orm();
        if (tev) QK_HIP(hipEventRecord(tev[3], s));
        std::swap(A, Z);
    } else {
        // eigenvalues only: divide-and-conquer into the spare matrix (rocSOLVER's
        // dsterf, a sequential QL sweep, takes seconds at n ~ 5000)
        rb_check(rocsolver_dstedc(rb, rocblas_evect_tridiagonal, n, w, E, Z, n, dinfo), "rocsolver_dstedc");
    }
    int info = 0;
    QK_HIP(hipMemcpyAsync(&info, dinfo, sizeof(int), hipMemcpyDeviceToHost, s));
    QK_HIP(hipStreamSynchronize(s));
    QK_CHECK(info == 0, QK_E_EIG, "tridiagonal eigensolver failed, info = " + std::to_string(info));
}

}  // namespace

// Event-timed stage on the handle's stream (accumulates into t_stage[k]).
struct Stage {
    Ctx* c;
    int k;
    hipEvent_t* ev;
    Stage(Ctx* c_, int k_, bool inner = false) : c(c_), k(k_), ev(inner ? c_->ev2 : c_->ev) {
        QK_HIP(hipEventRecord(ev[0], c->s));
    }
    ~Stage() {
        // never throws: a failed timing read leaves the stage time unchanged
        float ms = 0.f;
        if (hipEventRecord(ev[1], c->s) == hipSuccess && hipEventSynchronize(ev[1]) == hipSuccess &&
            hipEventElapsedTime(&ms, ev[0], ev[1]) == hipSuccess)
        {
            c->t_stage[k] += ms;
            if (k == 5) c->t_gemm_total += ms;
        }
    }
};

void check_sizes(int nBas, int nOrb, int nC, int nO, int nV, int nR, int nS) {
    QK_CHECK(nBas > 0 && nOrb > 0 && nOrb <= nBas, QK_E_ARG, "need 0 < nOrb <= nBas");
    QK_CHECK(nC >= 0 && nR >= 0 && nO > nC && nO < nOrb - nR, QK_E_ARG, "need nC < nO < nOrb - nR");
    QK_CHECK(nV == nOrb - nO, QK_E_ARG, "need nV = nOrb - nO");
    QK_CHECK((long long)nS == (long long)(nO - nC) * (nV - nR), QK_E_ARG, "need nS = (nO-nC)*(nV-nR)");
    QK_CHECK((long long)nS * nS < (1LL << 40), QK_E_ARG, "nS too large");
}

void h2d(Ctx* c, double* dst, const double* src, size_t count) {
    QK_HIP(hipMemcpyAsync(dst, src, count * sizeof(double), hipMemcpyHostToDevice, c->s));
}
void d2h(Ctx* c, double* dst, const double* src, size_t count) {
    if (dst) QK_HIP(hipMemcpyAsync(dst, src, count * sizeof(double), hipMemcpyDeviceToHost, c->s));
}
double d2h_scalar(Ctx* c, const double* src) {
    double v = 0.0;
    QK_HIP(hipMemcpyAsync(&v, src, sizeof(double), hipMemcpyDeviceToHost, c->s));
    QK_HIP(hipStreamSynchronize(c->s));
    return v;
}

void alloc_split(Split3& sp, int rows, int K, int nsl) {
    sp.rows = rows;
    sp.kpad = pad32(K);
    sp.nsl = nsl;
    for (int t = 0; t < nsl; ++t) sp.s[t] = dalloc<__bf16>((size_t)rows * sp.kpad);
}
void free_split(Split3& sp) {
    for (auto& p : sp.s) dfree(p);
}
// slices per operand of a split-precision mode (1: three, 2: two)
int slices_of(int precision) { return precision == 2 ? 2 : 3; }

constexpr int EIG_MAXIT = 8;       // Newton steps before falling back to dsyevd
// |E_ij| above this: the pair is outside the Newton step's convergence region
// and joins a cluster.  Measured at nS = 5600: 0.25 let one evGW cycle wander
// through 8 steps with corrections near 0.25; at 0.1 that cycle's pairs form a
// cluster too large for EIG_KMAX and it takes a full solve instead, and the
// later cycles converge in 3-5 steps.
constexpr double EIG_BIG = 0.1;
constexpr double EIG_TOL = 1e-8;   // last step: the next correction would be O(1e-16)
// Split-precision forms, whose vectors enter only an FP32-level contraction:
// the last step is the one with max|E| below this; its update X (I + E) leaves
// the vectors within O(1e-10), and its Rayleigh quotients S_ii/G_ii are
// already exact to O(1e-10) relative (second order in E), one Newton step
// (three of its four n^3 products) earlier than EIG_TOL.
constexpr double EIG_TOL_SPLIT = 1e-5;
constexpr double EIG_DEG = 1e-9;   // eigenvalues closer than this (relative) form a cluster
constexpr int EIG_KMAX = 32;       // largest cluster handled
// QK_EIG_TRACE=1 prints the largest correction of every refinement step to stderr
bool eig_trace() {
    static const bool on = [] {
        const char* v = std::getenv("QK_EIG_TRACE");
        return v && v[0] == '1';
    }();
    return on;
}

// Cyclic Jacobi eigen-decomposition of a small symmetric matrix A (k x k,
// column-major, destroyed): V = eigenvectors (columns), w = eigenvalues,
// ascending.  Rotation angles as in Golub & Van Loan, Matrix Computations,
// sec. 8.5 (symmetric Schur decomposition).
void jacobi_eig(int k, std::vector<double>& A, std::vector<double>& V, std::vector<double>& w) {
    V.assign((size_t)k * k, 0.0);
    for (int i = 0; i < k; ++i) V[i + (size_t)k * i] = 1.0;
    for (int sweep = 0; sweep < 64; ++sweep) {
        double off = 0.0, tot = 0.0;
        for (int q = 0; q < k; ++q)
            for (int p = 0; p < k; ++p) {
                double a = A[p + (size_t)k * q] * A[p + (size_t)k * q];
                tot += a;
                if (p != q) off += a;
            }
        if (off <= 1e-34 * tot) break;
        for (int p = 0; p < k - 1; ++p)
            for (int q = p + 1; q < k; ++q) {
                double apq = A[p + (size_t)k * q];
                if (apq == 0.0) continue;
                double tau = (A[q + (size_t)k * q] - A[p + (size_t)k * p]) / (2.0 * apq);
                double t = (tau >= 0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(1.0 + tau * tau));
                double cs = 1.0 / std::sqrt(1.0 + t * t), sn = t * cs;
                for (int r = 0; r < k; ++r) {  // A <- A J
                    double x = A[r + (size_t)k * p], y = A[r + (size_t)k * q];
                    A[r + (size_t)k * p] = cs * x - sn * y;
                    A[r + (size_t)k * q] = sn * x + cs * y;
                }
                for (int r = 0; r < k; ++r) {  // A <- J^T A
                    double x = A[p + (size_t)k * r], y = A[q + (size_t)k * r];
                    A[p + (size_t)k * r] = cs * x - sn * y;
                    A[q + (size_t)k * r] = sn * x + cs * y;
                }
                for (int r = 0; r < k; ++r) {  // V <- V J
                    double x = V[r + (size_t)k * p], y = V[r + (size_t)k * q];
                    V[r + (size_t)k * p] = cs * x - sn * y;
                    V[r + (size_t)k * q] = sn * x + cs * y;
                }
            }
    }
    std::vector<int> o(k);
    for (int i = 0; i < k; ++i) o[i] = i;
    std::sort(o.begin(), o.end(), [&](int a, int b) { return A[a + (size_t)k * a] < A[b + (size_t)k * b]; });
    std::vector<double> V2((size_t)k * k);
    w.resize(k);
    for (int j = 0; j < k; ++j) {
        w[j] = A[o[j] + (size_t)k * o[j]];
        for (int r = 0; r < k; ++r) V2[r + (size_t)k * j] = V[r + (size_t)k * o[j]];
    }
    V.swap(V2);
}

// Clusters of the refinement: labels (merge only) and their member lists.
struct EigClusters {
    std::vector<int> lab, ptr, idx, scl, loc;
    std::vector<long long> boff;
    int ncl = 0, m = 0, biggest = 1;
    long long nblk = 0;
    explicit EigClusters(int n) : lab(n) {
        for (int i = 0; i < n; ++i) lab[i] = i;
    }
    void merge(int a, int b) {
        int la = lab[a], lb = lab[b];
        if (la == lb) return;
        int keep = std::min(la, lb), drop = std::max(la, lb);
        for (int& l : lab)
            if (l == drop) l = keep;
    }
    // members of clusters of size >= 2; false if one exceeds EIG_KMAX
    bool build() {
        int n = (int)lab.size();
        std::vector<int> cnt(n, 0);
        for (int l : lab) ++cnt[l];
        biggest = *std::max_element(cnt.begin(), cnt.end());
        ptr.assign(1, 0);
        idx.clear(); scl.clear(); boff.clear();
        std::vector<int> first(n, -1);
        ncl = 0;
        for (int l = 0; l < n; ++l)
            if (cnt[l] >= 2) {
                if (cnt[l] > EIG_KMAX) return false;
                first[l] = ncl++;
            }
        std::vector<std::vector<int>> mem(ncl);
        for (int i = 0; i < n; ++i)
            if (first[lab[i]] >= 0) mem[first[lab[i]]].push_back(i);
        nblk = 0;
        for (int c = 0; c < ncl; ++c) {
            boff.push_back(nblk);
            nblk += (long long)mem[c].size() * mem[c].size();
            for (int i : mem[c]) {
                idx.push_back(i);
                scl.push_back(c);
            }
            ptr.push_back((int)idx.size());
        }
        m = (int)idx.size();
        loc.resize(m);
        for (int s = 0; s < m; ++s) loc[s] = s;
        return true;
    }
};

// Upload the member lists; ib: ptr | idx | scl | loc | lab, lb: boff.
void upload_clusters(Ctx* c, const EigClusters& cl) {
    int n = c->nS;
    int* ib = c->eig_ibuf;
    QK_HIP(hipMemcpyAsync(ib, cl.ptr.data(), sizeof(int) * cl.ptr.size(), hipMemcpyHostToDevice, c->s));
    if (cl.m > 0) {
        QK_HIP(hipMemcpyAsync(ib + (n + 1), cl.idx.data(), sizeof(int) * cl.m, hipMemcpyHostToDevice, c->s));
        QK_HIP(hipMemcpyAsync(ib + (2 * n + 1), cl.scl.data(), sizeof(int) * cl.m, hipMemcpyHostToDevice, c->s));
        QK_HIP(hipMemcpyAsync(ib + (3 * n + 1), cl.loc.data(), sizeof(int) * cl.m, hipMemcpyHostToDevice, c->s));
        QK_HIP(hipMemcpyAsync(c->eig_lbuf, cl.boff.data(), sizeof(long long) * cl.ncl, hipMemcpyHostToDevice, c->s));
    }
    QK_HIP(hipMemcpyAsync(ib + (4 * n + 1), cl.lab.data(), sizeof(int) * n, hipMemcpyHostToDevice, c->s));
}

// Rotation inside each cluster that diagonalises the cluster block of S in the
// metric G: T = S + (R S + S R)/2 (R = I - G, first order), T = Q diag(w) Q^T;
// with orth, W = (I + R/2) Q (orthonormalising too).  Blocks from the device
// arrays bS, bG (packed, cl.boff); W packed the same way into Wh.
void cluster_rotations(const EigClusters& cl, const std::vector<double>& bS, const std::vector<double>& bG, bool orth,
                       std::vector<double>& Wh, std::vector<double>* wout) {
    Wh.assign(cl.nblk, 0.0);
    std::vector<double> T, Q, w;
    for (int c = 0; c < cl.ncl; ++c) {
        int k = cl.ptr[c + 1] - cl.ptr[c];
        const double* S = bS.data() + cl.boff[c];
        const double* G = bG.data() + cl.boff[c];
        T.assign((size_t)k * k, 0.0);
        for (int q = 0; q < k; ++q)
            for (int p = 0; p < k; ++p) {
                double t = S[p + k * q];
                for (int u = 0; u < k; ++u) {
                    double rpu = (p == u ? 1.0 : 0.0) - G[p + k * u];
                    double ruq = (u == q ? 1.0 : 0.0) - G[u + k * q];
                    t += 0.5 * (rpu * S[u + k * q] + S[p + k * u] * ruq);
                }
                T[p + (size_t)k * q] = t;
            }
        for (int q = 0; q < k; ++q)  // symmetrise against rounding
            for (int p = 0; p < q; ++p) {
                double a = 0.5 * (T[p + (size_t)k * q] + T[q + (size_t)k * p]);
                T[p + (size_t)k * q] = T[q + (size_t)k * p] = a;
            }
        jacobi_eig(k, T, Q, w);
        double* W = Wh.data() + cl.boff[c];
        for (int q = 0; q < k; ++q)
            for (int p = 0; p < k; ++p) {
                double v = Q[p + (size_t)k * q];
                if (orth)
                    for (int u = 0; u < k; ++u) v += 0.5 * ((p == u ? 1.0 : 0.0) - G[p + k * u]) * Q[u + (size_t)k * q];
                W[p + k * q] = v;
            }
        if (wout)
            for (int r = 0; r < k; ++r) (*wout)[cl.idx[cl.ptr[c] + r]] = w[r];
    }
}

// Eigen-decomposition of C (in Cm, intact on failure) from the eigenvectors Xv
// of a nearby matrix.  Quadratically convergent Newton iteration for
// X^T X = I, X^T C X diagonal: Y = C X, S = X^T Y, G = X^T X, X <- X (I + E).
// Exactly or nearly degenerate eigenvalues (relative gap <= 1e-9, or a
// correction above 0.25 between them) form clusters whose vectors are rotated
// to diagonalise their block before each step and once more at the end.
// On success Cm = eigenvectors and Om = eigenvalues, ascending (as dsyevd).
bool refine_eig(Ctx* c) {
    const int n = c->nS;
    hipStream_t s = c->s;
    double *X = c->Xv, *Y = c->Ym;
    std::vector<double> lam(n), colmax(n), bS, bG, Wh;
    std::vector<int> colarg(n), order(n), pos(n);
    EigClusters cl(n);
    const long long before_fail = c->eig_count[4] + c->eig_count[5];
    int* d_ptr = c->eig_ibuf;
    int* d_idx = d_ptr + (n + 1);
    int* d_scl = d_ptr + (2 * n + 1);
    int* d_loc = d_ptr + (3 * n + 1);
    int* d_lab = d_ptr + (4 * n + 1);
    int* d_arg = d_ptr + (5 * n + 1);
    double* d_Q = c->eig_dbuf;
    double* d_bS = d_Q + (size_t)EIG_KMAX * n;
    double* d_bG = d_bS + (size_t)EIG_KMAX * n;
    double* d_cmax = d_bG + (size_t)EIG_KMAX * n;

    auto fetch_lam = [&] {
        eig_lambda(s, c->Sm, c->Gm, c->lam, n);
        QK_HIP(hipMemcpyAsync(lam.data(), c->lam, sizeof(double) * n, hipMemcpyDeviceToHost, s));
        QK_HIP(hipStreamSynchronize(s));
        for (int i = 0; i < n; ++i) order[i] = i;
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return lam[a] < lam[b]; });
        for (int k = 0; k < n; ++k) pos[order[k]] = k;
    };
    auto fetch_blocks = [&](const double* A, const double* B, long long lda, const int* dix) {
        gather_blocks(s, A, B, lda, d_ptr, dix, c->eig_lbuf, cl.ncl, d_bS, d_bG);
        bS.resize(cl.nblk);
        bG.resize(cl.nblk);
        QK_HIP(hipMemcpyAsync(bS.data(), d_bS, sizeof(double) * cl.nblk, hipMemcpyDeviceToHost, s));
        QK_HIP(hipMemcpyAsync(bG.data(), d_bG, sizeof(double) * cl.nblk, hipMemcpyDeviceToHost, s));
        QK_HIP(hipStreamSynchronize(s));
    };

    bool ok = false;
    for (int it = 0; it < EIG_MAXIT && !ok; ++it) {
        Batch up;
        up.upper = 1;  // S and G are symmetric: upper tiles, then mirror
        dgemm(s, false, false, n, n, n, 1.0, c->Cm, n, X, n, 0.0, Y, n);
        dgemm(s, true, false, n, n, n, 1.0, X, n, Y, n, 0.0, c->Sm, n, up);
        dgemm(s, true, false, n, n, n, 1.0, X, n, X, n, 0.0, c->Gm, n, up);
        sym_lower(s, c->Sm, n);
        sym_lower(s, c->Gm, n);
        ++c->eig_count[2];
        fetch_lam();
        double lmax = 0.0;
        for (double l : lam) lmax = std::max(lmax, std::fabs(l));
        if (!(lmax < 1e300)) break;
        for (int k = 0; k + 1 < n; ++k)
            if (lam[order[k + 1]] - lam[order[k]] <= EIG_DEG * lmax) cl.merge(order[k], order[k + 1]);
        bool settled = false;
        for (int rep = 0; rep < 12 && !settled; ++rep) {
            if (!cl.build()) {
                ++c->eig_count[4];
                c->eig_count[7] = std::max<long long>(c->eig_count[7], cl.biggest);
                c->Xv = X;
                c->Ym = Y;
                return false;
            }
            c->eig_count[7] = std::max<long long>(c->eig_count[7], cl.biggest);
            upload_clusters(c, cl);
            if (cl.ncl > 0) {
                fetch_blocks(c->Sm, c->Gm, n, d_idx);
                cluster_rotations(cl, bS, bG, false, Wh, nullptr);
                QK_HIP(hipMemcpyAsync(d_Q, Wh.data(), sizeof(double) * cl.nblk, hipMemcpyHostToDevice, s));
                for (double* A : {c->Sm, c->Gm}) {
                    mix_cols(s, A, n, n, d_idx, A, n, d_idx, d_ptr, d_scl, d_Q, c->eig_lbuf, cl.m, Y);
                    mix_rows(s, A, n, d_idx, d_ptr, d_scl, d_Q, c->eig_lbuf, cl.m, Y);
                }
                mix_cols(s, X, n, n, d_idx, X, n, d_idx, d_ptr, d_scl, d_Q, c->eig_lbuf, cl.m, Y);
                fetch_lam();
            }
            eig_check(s, c->Sm, c->Gm, c->lam, d_lab, n, d_cmax, d_arg);
            QK_HIP(hipMemcpyAsync(colmax.data(), d_cmax, sizeof(double) * n, hipMemcpyDeviceToHost, s));
            QK_HIP(hipMemcpyAsync(colarg.data(), d_arg, sizeof(int) * n, hipMemcpyDeviceToHost, s));
            QK_HIP(hipStreamSynchronize(s));
            settled = true;
            for (int j = 0; j < n; ++j)
                if (!(colmax[j] <= EIG_BIG)) {
                    settled = false;
                    int a = std::min(pos[j], pos[colarg[j]]), b = std::max(pos[j], pos[colarg[j]]);
                    for (int k = a; k < b; ++k) cl.merge(order[k], order[k + 1]);
                }
        }
        if (!settled) {
            ++c->eig_count[5];
            break;
        }
        double emax = 0.0;
        for (double v : colmax) emax = std::max(emax, v);
        if (eig_trace())
            std::fprintf(stderr, "refine: n %d step %d max|E| %.3e clusters %d largest %d\n", n, it, emax, cl.ncl,
                         cl.biggest);
        // FP64 form: the last correction is below 1e-8, the eigenvalues S_ii/G_ii
        // of this step are exact to O(1e-16), and the final X (I + E) is applied.
        // Split-precision forms: the last correction is below EIG_TOL_SPLIT and
        // the final X (I + E) is applied too.
        ok = emax < (c->precision != 0 ? EIG_TOL_SPLIT : EIG_TOL);
        eig_update(s, c->Sm, c->Gm, c->lam, d_lab, n);
        dgemm(s, false, false, n, n, n, 1.0, X, n, c->Sm, n, 0.0, Y, n);
        std::swap(X, Y);
    }
    c->Xv = X;
    c->Ym = Y;
    if (!ok) {
        if (c->eig_count[4] + c->eig_count[5] == before_fail) ++c->eig_count[6];
        return false;
    }
    // final Rayleigh-Ritz inside the clusters with the converged vectors
    if (!cl.build()) return false;
    if (cl.ncl > 0) {
        upload_clusters(c, cl);
        const int m = cl.m;
        gather_cols(s, X, n, d_idx, m, c->Sm);                                  // Xc (n x m)
        dgemm(s, false, false, n, m, n, 1.0, c->Cm, n, c->Sm, n, 0.0, c->Gm, n);  // C Xc
        dgemm(s, true, false, m, m, n, 1.0, c->Sm, n, c->Gm, n, 0.0, Y, m);      // Xc^T C Xc
        dgemm(s, true, false, m, m, n, 1.0, c->Sm, n, c->Sm, n, 0.0, c->Gm, m);  // Xc^T Xc
        fetch_blocks(Y, c->Gm, m, d_loc);
        cluster_rotations(cl, bS, bG, true, Wh, &lam);
        QK_HIP(hipMemcpyAsync(d_Q, Wh.data(), sizeof(double) * cl.nblk, hipMemcpyHostToDevice, s));
        mix_cols(s, c->Sm, n, n, d_loc, X, n, d_idx, d_ptr, d_scl, d_Q, c->eig_lbuf, m, Y);
    }
    std::vector<int> perm(n);
    for (int i = 0; i < n; ++i) perm[i] = i;
    std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return lam[a] < lam[b]; });
    std::vector<double> sorted(n);
    for (int i = 0; i < n; ++i) sorted[i] = lam[perm[i]];
    QK_HIP(hipMemcpyAsync(c->perm, perm.data(), sizeof(int) * n, hipMemcpyHostToDevice, s));
    QK_HIP(hipMemcpyAsync(c->Om, sorted.data(), sizeof(double) * n, hipMemcpyHostToDevice, s));
    permute_cols(s, X, c->perm, c->Cm, n);
    QK_HIP(hipStreamSynchronize(s));  // host vectors leave scope
    return true;
}

// RPA screening at energies e (device; eh = the same energies on the host).
// Cm = eigenvectors (scaled to W when vectors), Om = excitation energies.
int n_energies(const Ctx* c) { return c->nOb >= 0 ? 2 * c->nOrb : c->nOrb; }

void screening(Ctx* c, const double* e, const double* eh, int tda, bool vectors) {
    {
        Stage st(c, 0);
        if (c->nOb >= 0) {
            delta_u(c->s, e, c->dE, c->sd, c->up);
            rpa_build_c(c->s, c->K, c->dE, c->sd, c->Cm, c->scal, c->nS, tda);
        } else {
            rpa_build(c->s, e, c->K, c->dE, c->sd, c->Cm, c->scal, c->nC, c->nO, c->v, c->nS, tda);
        }
    }
    {
        Stage st(c, 1);
        bool use_refine = (c->eig_mode & 1) && vectors && c->have_vec && c->tda_vec == tda;
        if (use_refine) {
            double d = 0.0;
            for (int p = 0; p < n_energies(c); ++p) d = std::max(d, std::fabs(eh[p] - c->e_vec[p]));
            use_refine = d <= c->eig_demax;
        }
        bool done = false;
        if (use_refine) {
            done = refine_eig(c);
            ++c->eig_count[done ? 1 : 3];
        }
        if (!done) {
            if (c->eig_mode & 2) {
                syevd_own(c->s, c->rb, vectors, c->nS, c->Cm, c->Om, c->work, c->tau, c->Zev, c->trd_work,
                          c->orm_work, c->info, c->ev3, &c->graphs);
                float ms = 0.f;
                QK_HIP(hipEventSynchronize(c->ev3[1]));
                QK_HIP(hipEventElapsedTime(&ms, c->ev3[0], c->ev3[1]));
                c->t_trd_total += ms;
                c->n_trd_total += 1.0;
                if (vectors) {
                    QK_HIP(hipEventSynchronize(c->ev3[3]));
                    QK_HIP(hipEventElapsedTime(&ms, c->ev3[1], c->ev3[2]));
                    c->t_stedc_total += ms;
                    QK_HIP(hipEventElapsedTime(&ms, c->ev3[2], c->ev3[3]));
                    c->t_ormtr_total += ms;
                }
                ++c->eig_count[8];
            } else {
                syevd(c->s, c->rb, vectors, c->nS, c->Cm, c->Om, c->work, c->info);
            }
            ++c->eig_count[0];
            if ((c->eig_mode & 1) && vectors)
                QK_HIP(hipMemcpyAsync(c->Xv, c->Cm, sizeof(double) * c->nS * c->nS, hipMemcpyDeviceToDevice, c->s));
        }
        if ((c->eig_mode & 1) && vectors) {
            c->have_vec = true;
            c->tda_vec = tda;
            c->e_vec.assign(eh, eh + n_energies(c));
        }
    }
    {
        Stage st(c, 0);
        rpa_finalize(c->s, c->Om, c->Cm, c->sd, c->scal, c->nS, tda, vectors);
    }
}

// rhoT[m + nS (p + n q)] = sum_jb W[jb, m] M[pq, jb]  (rho^T = W^T M^T)
void excitation_df(Ctx* c);

// excitation GEMM launches (several on a shape's first, tuning call)
void count_gemm(Ctx* c, int launches) {
    c->t_stage[6] += launches;
    c->n_gemm_total += launches;
}

void excitation(Ctx* c) {
    Stage st(c, 2);
    long long n2 = (long long)c->nOrb * c->nOrb;
    QK_CHECK(n2 < (1LL << 31), QK_E_ARG, "nOrb^2 too large");
    if (c->use_df) {
        excitation_df(c);
        return;
    }
    if (c->precision != 0 && (!c->Msplit.s[0] || c->Msplit.nsl != slices_of(c->precision))) {
        free_split(c->Wsplit);
        free_split(c->Msplit);
        alloc_split(c->Wsplit, c->nS, c->nS, slices_of(c->precision));
        alloc_split(c->Msplit, (int)n2, c->nS, slices_of(c->precision));
        c->msplit_valid = false;
    }
    c->rho_f32 = c->precision != 0;
    if (c->precision == 0) {
        Stage g(c, 5, true);
        count_gemm(c, dgemm(c->s, true, true, c->nS, (int)n2, c->nS, 1.0, c->Cm, c->nS, c->M, n2, 0.0, c->rhoT,
                            c->nS));
        return;
    }
    // rho^T(m, pq) = sum_jb W(jb, m) M(pq, jb): A rows m (k = jb contiguous in W),
    // B rows pq (M^T, split once per molecule)
    if (!c->msplit_valid) {
        split3(c->s, c->M, n2, false, (int)n2, c->nS, c->Msplit);
        c->msplit_valid = true;
    }
    split3(c->s, c->Cm, c->nS, true, c->nS, c->nS, c->Wsplit);
    Stage g(c, 5, true);
    count_gemm(c, gemm_bf16x3(c->s, c->nS, (int)n2, c->Wsplit, c->Msplit, reinterpret_cast<float*>(c->rhoT), c->nS));
}

// Density-fitted excitation densities: T(P, m) = sum_jb Bov(P, jb) W(jb, m),
// rho^T(m, pq) = sum_P T(P, m) B(P, pq).  The second product (2 nS nOrb^2 naux
// flop) is the one timed as the excitation GEMM.
void excitation_df(Ctx* c) {
    const int nS = c->nS, naux = c->naux;
    const long long n2 = (long long)c->nOrb * c->nOrb;
    c->rho_f32 = c->precision != 0;
    if (c->precision == 0) {
        dgemm(c->s, false, false, naux, nS, nS, 1.0, c->Bov, naux, c->Cm, nS, 0.0, c->Tdf, naux);
        Stage g(c, 5, true);
        count_gemm(c, dgemm(c->s, true, false, nS, (int)n2, naux, 1.0, c->Tdf, naux, c->Bdf, naux, 0.0, c->rhoT, nS));
        return;
    }
    const int nsl = slices_of(c->precision);
    if (!c->Wsplit.s[0] || c->Wsplit.nsl != nsl) {
        free_split(c->Wsplit);
        alloc_split(c->Wsplit, nS, nS, nsl);
    }
    if (!c->Tsplit.s[0] || c->Tsplit.nsl != nsl) {
        free_split(c->Tsplit);
        free_split(c->BovTsplit);
        free_split(c->Bsplit);
        alloc_split(c->Tsplit, nS, naux, nsl);
        alloc_split(c->BovTsplit, naux, nS, nsl);
        alloc_split(c->Bsplit, (int)n2, naux, nsl);
        c->bsplit_valid = false;
    }
    if (!c->bsplit_valid) {
        split3(c->s, c->Bov, naux, false, naux, nS, c->BovTsplit);
        split3(c->s, c->Bdf, naux, true, (int)n2, naux, c->Bsplit);
        c->bsplit_valid = true;
    }
    split3(c->s, c->Cm, nS, true, nS, nS, c->Wsplit);
    gemm_bf16x3(c->s, naux, nS, c->BovTsplit, c->Wsplit, c->Tdf, naux);
    split3(c->s, c->Tdf, naux, true, nS, naux, c->Tsplit);
    Stage g(c, 5, true);
    count_gemm(c, gemm_bf16x3(c->s, nS, (int)n2, c->Tsplit, c->Bsplit, reinterpret_cast<float*>(c->rhoT), nS));
}

void sigma_static(Ctx* c, const double* e, double eta, int srg) {
    Stage st(c, 3);
    sigma_partial(c->s, c->rhoT, c->rho_f32, e, e, c->Om, nullptr, eta, srg, SRG_FLOW, c->nOrb, c->nC, c->nO, c->nR, c->nS,
                  c->nchunk, c->part);
    sigma_finalize(c->s, c->part, c->nchunk, c->nOrb, c->nC, c->nO, c->nR, c->Sig, c->Z, c->scal);
}

// GW/RGW_QP_graph.f90:50-92 for all active orbitals at once: e0 = eHF, eG =
// energies in G, start = eGWlin; Z in/out.  Result in c->eGW.
void qp_graph(Ctx* c, const double* e0, const double* eG, const double* start, double eta, int srg, double flow) {
    Stage st(c, 3);
    int n = c->nOrb;
    qp_init(c->s, start, c->w, c->f, c->nit, c->done, n, c->nC, c->nR);
    // the host reads the active count after every QP_SYNC_EVERY sweeps: the
    // sweeps issued after the last orbital finished are masked no-ops, so the
    // results are those of a read after every sweep
    int h_active = 1;
    for (int it = 0; it < QP_MAXIT && h_active > 0;) {
        for (int k = 0; k < QP_SYNC_EVERY && it < QP_MAXIT; ++k, ++it) {
            sigma_partial(c->s, c->rhoT, c->rho_f32, c->w, eG, c->Om, c->done, eta, srg, flow, n, c->nC, c->nO, c->nR,
                          c->nS, c->nchunk, c->part);
            qp_update(c->s, c->part, c->nchunk, e0, c->w, c->f, c->df, c->nit, c->done, n, QP_THRESH, QP_MAXIT,
                      c->active);
        }
        QK_HIP(hipMemcpyAsync(&h_active, c->active, sizeof(int), hipMemcpyDeviceToHost, c->s));
        QK_HIP(hipStreamSynchronize(c->s));
    }
    qp_finish(c->s, c->w, c->df, c->nit, start, eG, c->eGW, c->Z, n, c->nC, c->nR, QP_MAXIT);
}


// Split-precision form of the AO->MO block (qk_set_split_aotomo), process-wide, off by default.
std::atomic<int> g_split_aotomo{0};
bool split_aotomo() { return g_split_aotomo.load() != 0; }

// (pq|jb) from AO integrals on the device: four quarter transforms (occupied
// index first: 2N^4 o + 6 N^3 o v flops instead of 8 N^5 for the full ERI_MO).
// With split, the last two quarter transforms of M
// (2 n N^2 v o flops each) run as three-slice BF16 products with FP32
// accumulation and FP64 recombination (relative error O(2^-24), far below the
// excitation densities' own split error), and K = (ia|jb), which sets the RPA
// energies, comes from a separate FP64 transform of the occupied-virtual block
// (2 o N^2 v o + 2 o v N v o flops); without split, K is taken from M later.
void aotomo_block(hipStream_t s, const double* f, const double* cc, double* M, int N, int n, int nC, int nO, int nR,
                  bool split, double* K) {
    int o = nO - nC, v = n - nR - nO;
    long long N2 = (long long)N * N, N3 = N2 * N;
    QK_CHECK(N3 < (1LL << 31), QK_E_ARG, "nBas^3 too large");
    QK_CHECK(!split || K, QK_E_ARG, "aotomo_block: the split form needs K");
    const double* Cocc = cc + (long long)N * nC;
    const double* Cvir = cc + (long long)N * nO;
    double* T1 = dalloc<double>((size_t)N3 * o);
    double* T2 = nullptr;
    Split3 As, Bs;
    try {
        // T1[mu, lam, nu, j] = sum_sig (mu nu|lam sig) C[sig, j]
        dgemm(s, false, false, (int)N3, o, N, 1.0, f, N3, Cocc, N, 0.0, T1, N3);
        // T2[mu, nu, b, j] = sum_lam T1[mu, lam, nu, j] C[lam, b]; batch (nu, j)
        T2 = dalloc<double>((size_t)N2 * v * o);
        Batch b2;
        b2.zdiv = N; b2.count = N * o;
        b2.sa0 = N2; b2.sa1 = N3;
        b2.sc0 = N; b2.sc1 = N2 * v;
        dgemm(s, false, false, N, v, N, 1.0, T1, N, Cvir, N, 0.0, T2, N2, b2);
        QK_HIP(hipStreamSynchronize(s));
        dfree(T1);
        long long cols = (long long)N * v * o;
        QK_CHECK(cols < (1LL << 31) && (long long)n * v * o < (1LL << 31), QK_E_ARG, "block too large");
        if (split) {
            const long long nS = (long long)o * v;
            // K[a' + v i' + nS jb] = sum_nu Cvir[nu, a'] sum_mu Cocc[mu, i'] T2[mu, nu, jb] in FP64
            T1 = dalloc<double>((size_t)o * cols);
            dgemm(s, true, false, o, (int)cols, N, 1.0, Cocc, N, T2, N, 0.0, T1, o);  // T3o[i' + o (nu + N jb)]
            Batch bk;
            bk.zdiv = v * o; bk.count = v * o;
            bk.sb0 = (long long)o * N;
            bk.sc0 = nS;
            dgemm(s, true, true, v, o, N, 1.0, Cvir, N, T1, o, 0.0, K, v, bk);
            QK_HIP(hipStreamSynchronize(s));
            dfree(T1);
            // T3[p + n (nu + N jb)] = sum_mu C[mu, p] T2[mu + N (nu + N jb)]: A(p, mu) = C[mu + N p],
            // B(r, mu) = T2[mu + N r], both k contiguous
            alloc_split(As, n, N, 3);
            alloc_split(Bs, (int)cols, N, 3);
            split3(s, cc, N, true, n, N, As);
            split3(s, T2, N, true, (int)cols, N, Bs);
            T1 = dalloc<double>((size_t)n * cols);
            gemm_bf16x3(s, n, (int)cols, As, Bs, T1, n);
            QK_HIP(hipStreamSynchronize(s));
            free_split(Bs);
            dfree(T2);
            // M[q + n (p + n jb)] = sum_nu C[nu + N q] T3[p + n nu + n N jb]: B rows r = p + n jb,
            // k = nu, split from the row blocks jb of T3 (M is symmetric in p and q)
            alloc_split(Bs, n * v * o, N, 3);
            split3_blocks(s, T1, n, (long long)n * N, n, v * o, N, Bs);
            QK_HIP(hipStreamSynchronize(s));
            dfree(T1);
            gemm_bf16x3(s, n, n * v * o, As, Bs, M, n);
            QK_HIP(hipStreamSynchronize(s));
            free_split(As);
            free_split(Bs);
            return;
        }
        // T3[p, nu, b, j] = sum_mu C[mu, p] T2[mu, nu, b, j]
        T1 = dalloc<double>((size_t)n * N * v * o);
        dgemm(s, true, false, n, (int)cols, N, 1.0, cc, N, T2, N, 0.0, T1, n);
        QK_HIP(hipStreamSynchronize(s));
        dfree(T2);
        // M[p, q, b, j] = sum_nu T3[p, nu, b, j] C[nu, q]; batch over (b + v j)
        Batch b4;
        b4.zdiv = v * o; b4.count = v * o;
        b4.sa0 = (long long)n * N;
        b4.sc0 = (long long)n * n;
        dgemm(s, false, false, n, n, N, 1.0, T1, n, cc, N, 0.0, M, n, b4);
        QK_HIP(hipStreamSynchronize(s));
        dfree(T1);
    } catch (...) {
        (void)hipStreamSynchronize(s);
        dfree(T1);
        dfree(T2);
        free_split(As);
        free_split(Bs);
        throw;
    }
}


// The four-index block M is allocated on first use (a density-fitted handle never needs it).
void ensure_m(Ctx* c) {
    if (!c->M) c->M = dalloc<double>((size_t)c->nOrb * c->nOrb * c->nS);
    c->use_df = false;
}

// k_done: K already formed (by the split form of aotomo_block)
void after_eri(Ctx* c, bool k_done = false) {
    c->msplit_valid = false;
    c->have_screen = false;
    c->have_vec = false;
    if (!k_done) extract_k(c->s, c->M, c->K, c->nOrb, c->nC, c->nO, c->v, c->nS);
    QK_HIP(hipStreamSynchronize(c->s));
    c->have_eri = true;
}

// gather X[p + n j' + n o q + n o n b'] = ERI_MO(p, nC+j', q, nO+b') on the host
void gather_mo_block(const double* ERI, double* X, int n, int nC, int nO, int nR) {
    int o = nO - nC, v = n - nR - nO;
    size_t n1 = n, n2 = n1 * n, n3 = n2 * n;
    for (int b = 0; b < v; ++b)
        for (int q = 0; q < n; ++q)
            std::memcpy(X + n1 * o * (q + n1 * b), ERI + n1 * nC + n2 * q + n3 * (nO + b), sizeof(double) * n1 * o);
}

void upload_mo_block(hipStream_t s, const double* ERI_MO, double* M, int n, int nC, int nO, int nR) {
    int o = nO - nC, v = n - nR - nO;
    size_t cnt = (size_t)n * n * o * v;
    double* hX = nullptr;
    QK_HIP(hipHostMalloc((void**)&hX, cnt * sizeof(double), hipHostMallocDefault));
    double* dX = nullptr;
    try {
        gather_mo_block(ERI_MO, hX, n, nC, nO, nR);
        dX = dalloc<double>(cnt);
        QK_HIP(hipMemcpyAsync(dX, hX, cnt * sizeof(double), hipMemcpyHostToDevice, s));
        permute_block(s, dX, M, n, o, v);
        QK_HIP(hipStreamSynchronize(s));
    } catch (...) {
        (void)hipStreamSynchronize(s);
        dfree(dX);
        (void)hipHostFree(hX);
        throw;
    }
    dfree(dX);
    (void)hipHostFree(hX);
}

// upload rho(p,q,m) (QuAcK layout) into rhoT[m + nS (p + n q)]
double* upload_rho(hipStream_t s, const double* rho, int n, int nS) {
    size_t n2 = (size_t)n * n;
    double* tmp = dalloc<double>(n2 * nS);
    double* out = nullptr;
    try {
        out = dalloc<double>(n2 * nS);
        QK_HIP(hipMemcpyAsync(tmp, rho, sizeof(double) * n2 * nS, hipMemcpyHostToDevice, s));
        transpose(s, tmp, out, (int)n2, nS, n2, nS);
        QK_HIP(hipStreamSynchronize(s));
    } catch (...) {
        (void)hipStreamSynchronize(s);
        dfree(tmp); dfree(out);
        throw;
    }
    dfree(tmp);
    return out;
}

Scratch& scratch() {
    static Scratch* sc = new Scratch();  // lives until process exit
    return *sc;
}

void release(Ctx* c) {
    if (c->s) (void)hipStreamSynchronize(c->s);
    c->graphs.clear();
    if (c->shared) {  // slots of the batch's contiguous arrays
        c->M = c->Cm = c->rhoT = c->Om = c->work = nullptr;
        c->K = c->dE = c->sd = c->Om0 = c->part = c->scal = nullptr;
        c->e = c->e0 = c->eG = c->Sig = c->Z = c->eLin = c->eGW = c->w = c->f = c->df = nullptr;
        c->info = c->active = c->nit = c->done = nullptr;
    }
    double* ds[] = {c->M, c->K, c->Cm, c->rhoT, c->e, c->e0, c->eG, c->dE, c->sd, c->Om, c->Om0, c->Sig, c->Z,
                    c->eLin, c->eGW, c->w, c->f, c->df, c->part, c->scal, c->work, c->sigw};
    for (double* p : ds)
        if (p) (void)hipFree(p);
    int* is[] = {c->nit, c->done, c->info, c->active};
    for (int* p : is)
        if (p) (void)hipFree(p);
    for (double* p : {c->Xv, c->Ym, c->Sm, c->Gm, c->lam, c->eig_dbuf, c->Zev, c->tau, c->trd_work, c->orm_work})
        if (p) (void)hipFree(p);
    for (int* p : {c->perm, c->eig_ibuf})
        if (p) (void)hipFree(p);
    if (c->eig_lbuf) (void)hipFree(c->eig_lbuf);
    free_split(c->Msplit);
    free_split(c->Wsplit);
    free_split(c->Bsplit);
    free_split(c->BovTsplit);
    free_split(c->Tsplit);
    for (double* p : {c->Bdf, c->Bov, c->Tdf})
        if (p) (void)hipFree(p);
    for (auto& e : c->ev)
        if (e) (void)hipEventDestroy(e);
    for (auto& e : c->ev2)
        if (e) (void)hipEventDestroy(e);
    for (auto& e : c->ev3)
        if (e) (void)hipEventDestroy(e);
    if (!c->shared) {
        if (c->rb) (void)rocblas_destroy_handle(c->rb);
        if (c->s) (void)hipStreamDestroy(c->s);
    }
}

}  // namespace qk

using namespace qk;

extern "C" {

const char* qk_last_error(void) { return g_last_error.c_str(); }

const char* qk_gw_last_error(qk_gw_handle h) { return h ? h->err.c_str() : g_last_error.c_str(); }

int qk_gw_create(qk_gw_handle* h, int nBas, int nOrb, int nC, int nO, int nV, int nR, int nS) {
    if (!h) return QK_E_ARG;
    *h = nullptr;
    Ctx* c = new (std::nothrow) Ctx();
    if (!c) return QK_E_ALLOC;
    int rc = guarded(nullptr, [&] {
        check_sizes(nBas, nOrb, nC, nO, nV, nR, nS);
        c->nBas = nBas; c->nOrb = nOrb; c->nC = nC; c->nO = nO; c->nV = nV; c->nR = nR; c->nS = nS;
        c->o = nO - nC; c->v = nOrb - nR - nO;
        QK_HIP(hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking));
        rb_check(rocblas_create_handle(&c->rb), "rocblas_create_handle");
        rb_check(rocblas_set_stream(c->rb, c->s), "rocblas_set_stream");
        for (auto* e : {&c->ev[0], &c->ev[1], &c->ev2[0], &c->ev2[1], &c->ev3[0], &c->ev3[1], &c->ev3[2], &c->ev3[3]})
            QK_HIP(hipEventCreate(e));
        size_t n = nOrb, S = nS;
        c->K = dalloc<double>(S * S);
        c->Cm = dalloc<double>(S * S);
        c->rhoT = dalloc<double>(n * n * S);
        for (double** p : {&c->e, &c->e0, &c->eG, &c->Sig, &c->Z, &c->eLin, &c->eGW, &c->w, &c->f, &c->df})
            *p = dalloc<double>(n);
        for (double** p : {&c->dE, &c->sd, &c->Om, &c->Om0, &c->work}) *p = dalloc<double>(S);
        c->nchunk = sigma_chunks(nS, nOrb - nR - nC);
        c->part = dalloc<double>(3 * n * c->nchunk);
        c->scal = dalloc<double>(8);
        c->nit = dalloc<int>(n);
        c->done = dalloc<int>(n);
        c->info = dalloc<int>(1);
        c->active = dalloc<int>(1);
    });
    if (rc != QK_OK) {
        release(c);
        delete c;
        return rc;
    }
    *h = c;
    return QK_OK;
}

int qk_gw_destroy(qk_gw_handle h) {
    if (!h) return QK_E_ARG;
    release(h);
    delete h;
    return QK_OK;
}

int qk_gw_set_eri_ao_device(qk_gw_handle c, const double* f, const double* cc) {
    if (!c || !f || !cc) return QK_E_ARG;
    return guarded(&c->err, [&] {
        ensure_m(c);
        const bool sp = split_aotomo();
        QK_HIP(hipEventRecord(c->ev[0], c->s));
        aotomo_block(c->s, f, cc, c->M, c->nBas, c->nOrb, c->nC, c->nO, c->nR, sp, c->K);
        QK_HIP(hipEventRecord(c->ev[1], c->s));
        QK_HIP(hipEventSynchronize(c->ev[1]));
        float ms = 0.f;
        QK_HIP(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
        c->t_ao = ms;
        after_eri(c, sp);
    });
}

int qk_gw_set_eri_ao(qk_gw_handle c, const double* ERI_AO, const double* cc) {
    if (!c || !ERI_AO || !cc) return QK_E_ARG;
    return guarded(&c->err, [&] {
        size_t N = c->nBas;
        double *df = nullptr, *dc = nullptr;
        try {
            ensure_m(c);
            df = dalloc<double>(N * N * N * N);
            dc = dalloc<double>(N * c->nOrb);
            h2d(c, df, ERI_AO, N * N * N * N);
            h2d(c, dc, cc, N * c->nOrb);
            const bool sp = split_aotomo();
            QK_HIP(hipEventRecord(c->ev[0], c->s));
            aotomo_block(c->s, df, dc, c->M, c->nBas, c->nOrb, c->nC, c->nO, c->nR, sp, c->K);
            QK_HIP(hipEventRecord(c->ev[1], c->s));
            QK_HIP(hipEventSynchronize(c->ev[1]));
            float ms = 0.f;
            QK_HIP(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
            c->t_ao = ms;
            after_eri(c, sp);
        } catch (...) {
            (void)hipStreamSynchronize(c->s);
            dfree(df);
            dfree(dc);
            throw;
        }
        dfree(df);
        dfree(dc);
    });
}

int qk_gw_set_eri_ao_packed(qk_gw_handle c, const double* ERI_chem, long long n_packed, const double* cc) {
    if (!c || !ERI_chem || !cc) return QK_E_ARG;
    return guarded(&c->err, [&] {
        const long long N = c->nBas, P = N * (N + 1) / 2;
        QK_CHECK(n_packed == P * (P + 1) / 2, QK_E_ARG,
                 "packed ERI: expected nBas(nBas+1)/2 pairs packed to " + std::to_string(P * (P + 1) / 2) +
                     " elements, got " + std::to_string(n_packed));
        double *dp = nullptr, *df = nullptr, *dc = nullptr;
        try {
            ensure_m(c);
            dp = dalloc<double>((size_t)n_packed);
            h2d(c, dp, ERI_chem, (size_t)n_packed);
            df = dalloc<double>((size_t)(N * N * N * N));
            unpack_eri_s8(c->s, dp, df, (int)N);
            QK_HIP(hipStreamSynchronize(c->s));
            dfree(dp);
            dp = nullptr;
            dc = dalloc<double>((size_t)N * c->nOrb);
            h2d(c, dc, cc, (size_t)N * c->nOrb);
            const bool sp = split_aotomo();
            QK_HIP(hipEventRecord(c->ev[0], c->s));
            aotomo_block(c->s, df, dc, c->M, c->nBas, c->nOrb, c->nC, c->nO, c->nR, sp, c->K);
            QK_HIP(hipEventRecord(c->ev[1], c->s));
            QK_HIP(hipEventSynchronize(c->ev[1]));
            float ms = 0.f;
            QK_HIP(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
            c->t_ao = ms;
            after_eri(c, sp);
        } catch (...) {
            (void)hipStreamSynchronize(c->s);
            dfree(dp);
            dfree(df);
            dfree(dc);
            throw;
        }
        dfree(df);
        dfree(dc);
    });
}

int qk_gw_set_eri_mo(qk_gw_handle c, const double* ERI_MO) {
    if (!c || !ERI_MO) return QK_E_ARG;
    return guarded(&c->err, [&] {
        ensure_m(c);
        upload_mo_block(c->s, ERI_MO, c->M, c->nOrb, c->nC, c->nO, c->nR);
        c->t_ao = 0.0;
        after_eri(c);
    });
}

int qk_gw_set_eri_df_device(qk_gw_handle c, const double* B, int naux) {
    if (!c || !B || naux <= 0) return QK_E_ARG;
    return guarded(&c->err, [&] {
        QK_CHECK(c->nBas == c->nOrb, QK_E_ARG, "density-fitted integrals are in the MO basis: need nBas = nOrb");
        const long long n2 = (long long)c->nOrb * c->nOrb;
        QK_CHECK(n2 < (1LL << 31) && (long long)naux * c->nS < (1LL << 40), QK_E_ARG, "sizes too large");
        if (c->naux != naux) {
            dfree(c->Bdf); dfree(c->Bov); dfree(c->Tdf);
            c->Bdf = c->Bov = c->Tdf = nullptr;
            free_split(c->Bsplit); free_split(c->BovTsplit); free_split(c->Tsplit);
            c->naux = naux;
        }
        if (!c->Bdf) {
            c->Bdf = dalloc<double>((size_t)naux * n2);
            c->Bov = dalloc<double>((size_t)naux * c->nS);
            c->Tdf = dalloc<double>((size_t)naux * c->nS);
        }
        QK_HIP(hipMemcpyAsync(c->Bdf, B, sizeof(double) * naux * n2, hipMemcpyDeviceToDevice, c->s));
        QK_HIP(hipEventRecord(c->ev[0], c->s));
        gather_bov(c->s, c->Bdf, c->Bov, naux, c->nOrb, c->nC, c->nO, c->v, c->nS);
        // K(ia, jb) = (ia|jb) = sum_P Bov(P, ia) Bov(P, jb), in the handle's precision
        if (c->precision == 0) {
            Batch up;
            up.upper = 1;
            dgemm(c->s, true, false, c->nS, c->nS, naux, 1.0, c->Bov, naux, c->Bov, naux, 0.0, c->K, c->nS, up);
            sym_lower(c->s, c->K, c->nS);
        } else {
            Split3 sp;
            alloc_split(sp, c->nS, naux);
            split3(c->s, c->Bov, naux, true, c->nS, naux, sp);
            gemm_bf16x3(c->s, c->nS, c->nS, sp, sp, c->K, c->nS);
            QK_HIP(hipStreamSynchronize(c->s));
            free_split(sp);
        }
        QK_HIP(hipEventRecord(c->ev[1], c->s));
        QK_HIP(hipEventSynchronize(c->ev[1]));
        float ms = 0.f;
        QK_HIP(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
        c->t_ao = ms;
        c->use_df = true;
        c->bsplit_valid = false;
        c->have_screen = false;
        c->have_vec = false;
        c->have_eri = true;
    });
}

int qk_gw_set_eri_df(qk_gw_handle c, const double* B, int naux) {
    if (!c || !B || naux <= 0) return QK_E_ARG;
    double* d = nullptr;
    int rc = guarded(&c->err, [&] {
        size_t cnt = (size_t)naux * c->nOrb * c->nOrb;
        d = dalloc<double>(cnt);
        QK_HIP(hipMemcpyAsync(d, B, sizeof(double) * cnt, hipMemcpyHostToDevice, c->s));
        QK_HIP(hipStreamSynchronize(c->s));
    });
    if (rc == QK_OK) rc = qk_gw_set_eri_df_device(c, d, naux);
    if (c) (void)hipStreamSynchronize(c->s);
    dfree(d);
    return rc;
}

int qk_gw_screen(qk_gw_handle c, const double* e, int TDA_W) {
    if (!c || !e) return QK_E_ARG;
    return guarded(&c->err, [&] {
        QK_CHECK(c->have_eri, QK_E_ARG, "no integrals: call qk_gw_set_eri_* first");
        for (double& t : c->t_stage) t = 0.0;
        h2d(c, c->e, e, c->nOrb);
        screening(c, c->e, e, TDA_W, true);
        excitation(c);
        QK_HIP(hipStreamSynchronize(c->s));
        c->have_screen = true;
    });
}

int qk_gw_sigma_full(qk_gw_handle c, const double* e, double eta, int doSRG, double* SigC, double* Z, double* EcGM,
                     double* EcRPA) {
    if (!c || !e || !SigC || !Z) return QK_E_ARG;
    return guarded(&c->err, [&] {
        QK_CHECK(c->have_screen, QK_E_ARG, "no screening: call qk_gw_screen first");
        QK_CHECK(!c->rho_f32, QK_E_UNSUPPORTED, "the full self-energy reads FP64 excitation densities (precision 0)");
        const int n = c->nOrb;
        const double ecrpa = d2h_scalar(c, c->scal + 1);  // the screening's, before the sums below overwrite scal
        h2d(c, c->eG, e, n);
        // Z and EcGM: the diagonal routine's (GW/RGW_self_energy.f90 computes them with the same formulas)
        sigma_partial(c->s, c->rhoT, false, c->eG, c->eG, c->Om, nullptr, eta, doSRG, SRG_FLOW, n, c->nC, c->nO, c->nR,
                      c->nS, c->nchunk, c->part);
        sigma_finalize(c->s, c->part, c->nchunk, n, c->nC, c->nO, c->nR, c->Sig, c->Z, c->scal);
        double* full = dalloc<double>((size_t)n * n);
        try {
            // weighted densities (n^2 nS) and one n x n product per orbital r (n^3)
            if (!doSRG && !c->sigw) c->sigw = dalloc<double>((size_t)n * n * c->nS + (size_t)n * n * n);
            sigma_full(c->s, c->rhoT, c->eG, c->Om, eta, doSRG, SRG_FLOW, n, c->nC, c->nO, c->nR, c->nS, full,
                       doSRG ? nullptr : c->sigw);
            QK_HIP(hipMemcpyAsync(SigC, full, sizeof(double) * (size_t)n * n, hipMemcpyDeviceToHost, c->s));
            d2h(c, Z, c->Z, n);
            const double ecgm = d2h_scalar(c, c->scal + 2);
            QK_HIP(hipStreamSynchronize(c->s));
            if (EcGM) *EcGM = ecgm;
            if (EcRPA) *EcRPA = ecrpa;
        } catch (...) {
            (void)hipStreamSynchronize(c->s);
            dfree(full);
            throw;
        }
        dfree(full);
    });
}

int qk_gw_sigma_at(qk_gw_handle c, const double* eG, const double* w, double eta, int doSRG, double* Sig,
                   double* dSig) {
    if (!c || !eG || !w || !Sig || !dSig) return QK_E_ARG;
    return guarded(&c->err, [&] {
        QK_CHECK(c->have_screen, QK_E_ARG, "no screening: call qk_gw_screen first");
        int n = c->nOrb;
        h2d(c, c->eG, eG, n);
        h2d(c, c->w, w, n);
        sigma_partial(c->s, c->rhoT, c->rho_f32, c->w, c->eG, c->Om, nullptr, eta, doSRG, SRG_FLOW, n, c->nC, c->nO, c->nR,
                      c->nS, c->nchunk, c->part);
        sigma_out(c->s, c->part, c->nchunk, n, c->nC, c->nR, c->f, c->df);
        d2h(c, Sig, c->f, n);
        d2h(c, dSig, c->df, n);
        QK_HIP(hipStreamSynchronize(c->s));
    });
}

int qk_gw_set_precision(qk_gw_handle c, int mode) {
    if (!c || mode < 0 || mode > 2) return QK_E_ARG;
    return guarded(&c->err, [&] {
        c->precision = mode;  // slices are allocated on first use
    });
}

int qk_gw_set_eigensolver(qk_gw_handle c, int mode, double de_max) {
    if (!c || mode < 0 || mode > 3 || !(de_max >= 0.0)) return QK_E_ARG;
    return guarded(&c->err, [&] {
        if ((mode & 2) && !c->Zev) {
            size_t S = c->nS;
            c->Zev = dalloc<double>(S * S);
            c->tau = dalloc<double>(S);
            c->trd_work = dalloc<double>(sytrd_work_doubles(c->nS, 32));
            c->orm_work = dalloc<double>(ormtr_work_doubles(c->nS));
        }
        if ((mode & 1) && !c->Xv) {
            size_t S = c->nS;
            c->Xv = dalloc<double>(S * S);
            c->Ym = dalloc<double>(S * S);
            c->Sm = dalloc<double>(S * S);
            c->Gm = dalloc<double>(S * S);
            c->lam = dalloc<double>(S);
            c->eig_dbuf = dalloc<double>((3 * EIG_KMAX + 1) * S);
            c->eig_ibuf = dalloc<int>(6 * S + 1);
            c->eig_lbuf = dalloc<long long>(S);
            c->perm = dalloc<int>(S);
        }
        c->eig_mode = mode;
        c->eig_demax = de_max;
        c->have_vec = false;
    });
}

int qk_gw_eig_stats(qk_gw_handle c, long long* counts, int n) {
    if (!c || !counts || n < 0 || n > 9) return QK_E_ARG;
    for (int i = 0; i < n; ++i) counts[i] = c->eig_count[i];
    return QK_OK;
}

int qk_gw_qp_iters(qk_gw_handle c, int* nit) {
    if (!c || !nit) return QK_E_ARG;
    return guarded(&c->err, [&] {
        QK_HIP(hipMemcpyAsync(nit, c->nit, sizeof(int) * c->nOrb, hipMemcpyDeviceToHost, c->s));
        QK_HIP(hipStreamSynchronize(c->s));
    });
}

int qk_gw_stage_times(qk_gw_handle c, double* ms, int n) {
    if (!c || !ms || n < 0 || n > 14) return QK_E_ARG;
    double all[14] = {c->t_ao, c->t_stage[0], c->t_stage[1], c->t_stage[2], c->t_stage[3], c->t_stage[4],
                      c->t_stage[5], c->t_stage[6], c->t_gemm_total, c->n_gemm_total, c->t_trd_total, c->n_trd_total,
                      c->t_stedc_total, c->t_ormtr_total};
    for (int i = 0; i < n; ++i) ms[i] = all[i];
    return QK_OK;
}

int qk_rg0w0(qk_gw_handle c, const double* eHF, double eta, int linearize, int TDA_W, int doSRG, double* eGW,
             double* SigC, double* Z, double* Om, double* EcRPA, double* EcGM) {
    if (!c || !eHF || !eGW) return QK_E_ARG;
    return guarded(&c->err, [&] {
        QK_CHECK(c->have_eri, QK_E_ARG, "no integrals: call qk_gw_set_eri_* first");
        for (double& t : c->t_stage) t = 0.0;
        auto t0 = std::chrono::steady_clock::now();
        int n = c->nOrb;
        h2d(c, c->e0, eHF, n);
        // screening at eHF (GW/RG0W0.f90:116-119) and spectral weights (:127)
        screening(c, c->e0, eHF, TDA_W, true);
        QK_HIP(hipMemcpyAsync(c->Om0, c->Om, sizeof(double) * c->nS, hipMemcpyDeviceToDevice, c->s));
        excitation(c);
        c->have_screen = true;
        // self-energy (:169-177) and QP equation (:188-208)
        sigma_static(c, c->e0, eta, doSRG);
        double ecgm = d2h_scalar(c, c->scal + 2);
        lin_qp(c->s, c->e0, c->Z, c->Sig, c->eLin, n, 1);
        d2h(c, SigC, c->Sig, n);
        if (linearize) {
            QK_HIP(hipMemcpyAsync(c->eGW, c->eLin, sizeof(double) * n, hipMemcpyDeviceToDevice, c->s));
        } else {
            qp_graph(c, c->e0, c->e0, c->eLin, eta, doSRG, SRG_FLOW);
        }
        // RPA correlation energy at eGW (:222-225): eigenvalues only
        std::vector<double> hgw(n);
        d2h(c, hgw.data(), c->eGW, n);
        QK_HIP(hipStreamSynchronize(c->s));
        screening(c, c->eGW, hgw.data(), TDA_W, false);
        c->have_screen = false;  // the eigenvectors are gone
        double ecrpa = d2h_scalar(c, c->scal + 1);
        d2h(c, eGW, c->eGW, n);
        d2h(c, Z, c->Z, n);
        d2h(c, Om, c->Om0, c->nS);
        QK_HIP(hipStreamSynchronize(c->s));
        if (EcRPA) *EcRPA = ecrpa;
        if (EcGM) *EcGM = ecgm;
        c->t_stage[4] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    });
}

int qk_evrgw(qk_gw_handle c, const double* eHF, const double* eGW_start, int maxSCF, double thresh, int max_diis,
             double eta, int linearize, int TDA_W, int doSRG, double* eGW, double* SigC, double* Z, double* Om,
             double* EcRPA, double* EcGM, int* nSCF_out) {
    if (!c || !eHF || !eGW) return QK_E_ARG;
    int conv_status = QK_OK;
    int rc = guarded(&c->err, [&] {
        QK_CHECK(c->have_eri, QK_E_ARG, "no integrals: call qk_gw_set_eri_* first");
        QK_CHECK(maxSCF >= 0, QK_E_ARG, "maxSCF < 0");
        for (double& t : c->t_stage) t = 0.0;
        auto t0 = std::chrono::steady_clock::now();
        int n = c->nOrb;
        std::vector<double> hGW(eGW_start ? eGW_start : eHF, (eGW_start ? eGW_start : eHF) + n), hOld = hGW,
            herr(n);
        Diis diis(n, std::max(max_diis, 1));
        h2d(c, c->e0, eHF, n);
        int nSCF = 0;
        double Conv = 1.0, ecrpa = 0.0, ecgm = 0.0;
        // GW/evRGW.f90:119-183
        while (Conv > thresh && nSCF <= maxSCF) {
            h2d(c, c->e, hGW.data(), n);  // eGW == eOld here
            screening(c, c->e, hGW.data(), TDA_W, true);
            ecrpa = d2h_scalar(c, c->scal + 1);
            excitation(c);
            c->have_screen = true;
            sigma_static(c, c->e, eta, doSRG);
            ecgm = d2h_scalar(c, c->scal + 2);
            if (linearize) {
                lin_qp(c->s, c->e0, c->Z, c->Sig, c->eGW, n, 0);
            } else {
                qp_graph(c, c->e0, c->e, c->e, eta, doSRG, SRG_FLOW);
            }
            d2h(c, hGW.data(), c->eGW, n);
            QK_HIP(hipStreamSynchronize(c->s));
            Conv = 0.0;
            for (int p = 0; p < n; ++p) {
                herr[p] = hGW[p] - hOld[p];
                Conv = std::max(Conv, std::fabs(herr[p]));
            }
            if (max_diis > 1) diis.step(herr.data(), hGW.data());
            hOld = hGW;
            ++nSCF;
        }
        std::memcpy(eGW, hGW.data(), sizeof(double) * n);
        d2h(c, SigC, c->Sig, n);
        d2h(c, Z, c->Z, n);
        d2h(c, Om, c->Om, c->nS);
        QK_HIP(hipStreamSynchronize(c->s));
        if (EcRPA) *EcRPA = ecrpa;
        if (EcGM) *EcGM = ecgm;
        if (nSCF_out) *nSCF_out = nSCF;
        if (nSCF == maxSCF + 1) {
            conv_status = QK_E_NOCONV;
            c->err = "evGW: convergence failed after " + std::to_string(nSCF) + " cycles";
        }
        c->t_stage[4] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    });
    return rc != QK_OK ? rc : conv_status;
}

// ---- per-routine entry points ----------------------------------------------

int qk_aotomo_eri_rhf(int nBas, int nOrb, const double* c, const double* ERI_AO, double* ERI_MO) {
    if (nBas <= 0 || nOrb <= 0 || nOrb > nBas || !c || !ERI_AO || !ERI_MO) return QK_E_ARG;
    return guarded(nullptr, [&] {
        hipStream_t s = scratch().s;
        size_t N = nBas, n = nOrb;
        QK_CHECK(N * N * N < (1ULL << 31), QK_E_ARG, "nBas too large for the full transform");
        double *a = nullptr, *b = nullptr, *dc = nullptr;
        try {
            a = dalloc<double>(N * N * N * N);
            b = dalloc<double>(N * N * N * n);
            dc = dalloc<double>(N * n);
            QK_HIP(hipMemcpyAsync(a, ERI_AO, sizeof(double) * N * N * N * N, hipMemcpyHostToDevice, s));
            QK_HIP(hipMemcpyAsync(dc, c, sizeof(double) * N * n, hipMemcpyHostToDevice, s));
            // AOtoMO/AOtoMO_ERI_RHF.f90:232-252: dgemm('T','N', rest, nOrb, nBas) four times
            long long rest = N * N * N;
            dgemm(s, true, false, (int)rest, (int)n, (int)N, 1.0, a, N, dc, N, 0.0, b, rest);
            rest = N * N * n;
            dgemm(s, true, false, (int)rest, (int)n, (int)N, 1.0, b, N, dc, N, 0.0, a, rest);
            rest = N * n * n;
            dgemm(s, true, false, (int)rest, (int)n, (int)N, 1.0, a, N, dc, N, 0.0, b, rest);
            rest = n * n * n;
            dgemm(s, true, false, (int)rest, (int)n, (int)N, 1.0, b, N, dc, N, 0.0, a, rest);
            QK_HIP(hipMemcpyAsync(ERI_MO, a, sizeof(double) * n * n * n * n, hipMemcpyDeviceToHost, s));
            QK_HIP(hipStreamSynchronize(s));
        } catch (...) {
            (void)hipStreamSynchronize(s);
            dfree(a); dfree(b); dfree(dc);
            throw;
        }
        dfree(a); dfree(b); dfree(dc);
    });
}

int qk_phrlr_a(int dRPA, int nBas, int nC, int nO, int nV, int nR, int nS, double lambda, const double* e,
               const double* ERI, double* Aph) {
    if (!e || !ERI || !Aph) return QK_E_ARG;
    return guarded(nullptr, [&] {
        check_sizes(nBas, nBas, nC, nO, nV, nR, nS);
        hipStream_t s = scratch().s;
        size_t n = nBas;
        double *G = nullptr, *de = nullptr, *A = nullptr;
        try {
            G = dalloc<double>(n * n * n * n);
            de = dalloc<double>(n);
            A = dalloc<double>((size_t)nS * nS);
            QK_HIP(hipMemcpyAsync(G, ERI, sizeof(double) * n * n * n * n, hipMemcpyHostToDevice, s));
            QK_HIP(hipMemcpyAsync(de, e, sizeof(double) * n, hipMemcpyHostToDevice, s));
            phrlr_ab(s, G, de, A, (int)n, nC, nO, nBas - nR - nO, nS, lambda, dRPA, 0);
            QK_HIP(hipMemcpyAsync(Aph, A, sizeof(double) * nS * nS, hipMemcpyDeviceToHost, s));
            QK_HIP(hipStreamSynchronize(s));
        } catch (...) {
            (void)hipStreamSynchronize(s);
            dfree(G); dfree(de); dfree(A);
            throw;
        }
        dfree(G); dfree(de); dfree(A);
    });
}

int qk_phrlr_b(int dRPA, int nBas, int nC, int nO, int nV, int nR, int nS, double lambda, const double* ERI,
               double* Bph) {
    if (!ERI || !Bph) return QK_E_ARG;
    return guarded(nullptr, [&] {
        check_sizes(nBas, nBas, nC, nO, nV, nR, nS);
        hipStream_t s = scratch().s;
        size_t n = nBas;
        double *G = nullptr, *B = nullptr;
        try {
            G = dalloc<double>(n * n * n * n);
            B = dalloc<double>((size_t)nS * nS);
            QK_HIP(hipMemcpyAsync(G, ERI, sizeof(double) * n * n * n * n, hipMemcpyHostToDevice, s));
            phrlr_ab(s, G, nullptr, B, (int)n, nC, nO, nBas - nR - nO, nS, lambda, dRPA, 1);
            QK_HIP(hipMemcpyAsync(Bph, B, sizeof(double) * nS * nS, hipMemcpyDeviceToHost, s));
            QK_HIP(hipStreamSynchronize(s));
        } catch (...) {
            (void)hipStreamSynchronize(s);
            dfree(G); dfree(B);
            throw;
        }
        dfree(G); dfree(B);
    });
}

int qk_phrlr(int TDA, int nS, const double* Aph, const double* Bph, double* EcRPA, double* Om, double* XpY,
             double* XmY) {
    if (nS <= 0 || !Aph || (!TDA && !Bph) || !Om || !XpY || !XmY) return QK_E_ARG;
    return guarded(nullptr, [&] {
        hipStream_t s = scratch().s;
        rocblas_handle rb = scratch().rb;
        size_t S2 = (size_t)nS * nS;
        std::vector<double*> bufs;
        auto get = [&](size_t cnt) { bufs.push_back(dalloc<double>(cnt)); return bufs.back(); };
        try {
            double* A = get(S2);
            double* U = get(S2);
            double* w = get(nS);
            double* work = get(nS);
            double* sc = get(4);
            int* info = dalloc<int>(1);
            bufs.push_back(reinterpret_cast<double*>(info));
            QK_HIP(hipMemcpyAsync(A, Aph, sizeof(double) * S2, hipMemcpyHostToDevice, s));
            trace(s, A, nS, sc);
            if (TDA) {
                // LR/phRLR.f90:40-43
                QK_HIP(hipMemcpyAsync(U, A, sizeof(double) * S2, hipMemcpyDeviceToDevice, s));
                syevd(s, rb, true, nS, U, w, work, info);
                double* X = get(S2);
                transpose(s, U, X, nS, nS, nS, nS);
                rpa_finalize(s, w, nullptr, nullptr, sc, nS, 1, false);
                QK_HIP(hipMemcpyAsync(XpY, X, sizeof(double) * S2, hipMemcpyDeviceToHost, s));
                QK_HIP(hipMemcpyAsync(XmY, X, sizeof(double) * S2, hipMemcpyDeviceToHost, s));
            } else {
                // LR/phRLR.f90:47-81
                double* B = get(S2);
                double* ApB = get(S2);
                double* Sq = get(S2);
                double* Iv = get(S2);
                double* T = get(S2);
                QK_HIP(hipMemcpyAsync(B, Bph, sizeof(double) * S2, hipMemcpyHostToDevice, s));
                apb_amb(s, A, B, ApB, U, (long long)S2);
                syevd(s, rb, true, nS, U, w, work, info);
                // LR/phRLR.f90:57-58: QuAcK warns when A - B has a negative
                // eigenvalue and then takes its square root; stop before the
                // NaN (or the infinite inverse root of a zero mode) spreads
                double wmin = 0.0;
                QK_HIP(hipMemcpyAsync(&wmin, w, sizeof(double), hipMemcpyDeviceToHost, s));
                QK_HIP(hipStreamSynchronize(s));
                QK_CHECK(wmin > 0.0, QK_E_UNSTABLE,
                         "phRLR: A-B is not positive definite (smallest eigenvalue " + std::to_string(wmin) +
                             "): linear-response instability");
                colscale_pow(s, U, w, T, nS, 0.5);
                dgemm(s, false, true, nS, nS, nS, 1.0, T, nS, U, nS, 0.0, Sq, nS);   // ADAt(U, sqrt(w))
                colscale_pow(s, U, w, T, nS, -0.5);
                dgemm(s, false, true, nS, nS, nS, 1.0, T, nS, U, nS, 0.0, Iv, nS);   // ADAt(U, 1/sqrt(w))
                dgemm(s, false, false, nS, nS, nS, 1.0, ApB, nS, Sq, nS, 0.0, T, nS);  // tmp = ApB AmBSq
                dgemm(s, false, false, nS, nS, nS, 1.0, Sq, nS, T, nS, 0.0, U, nS);    // Z = AmBSq tmp
                syevd(s, rb, true, nS, U, w, work, info);
                rpa_finalize(s, w, nullptr, nullptr, sc, nS, 0, false);                 // Om = sqrt|w|
                dgemm(s, true, false, nS, nS, nS, 1.0, U, nS, Sq, nS, 0.0, T, nS);     // XpY = Z^T AmBSq
                rowscale_pow(s, T, w, nS, -0.5);
                QK_HIP(hipMemcpyAsync(XpY, T, sizeof(double) * S2, hipMemcpyDeviceToHost, s));
                dgemm(s, true, false, nS, nS, nS, 1.0, U, nS, Iv, nS, 0.0, ApB, nS);   // XmY = Z^T AmBIv
                rowscale_pow(s, ApB, w, nS, 0.5);
                QK_HIP(hipMemcpyAsync(XmY, ApB, sizeof(double) * S2, hipMemcpyDeviceToHost, s));
            }
            QK_HIP(hipMemcpyAsync(Om, w, sizeof(double) * nS, hipMemcpyDeviceToHost, s));
            double ec = 0.0;
            QK_HIP(hipMemcpyAsync(&ec, sc + 1, sizeof(double), hipMemcpyDeviceToHost, s));
            QK_HIP(hipStreamSynchronize(s));
            if (EcRPA) *EcRPA = ec;
        } catch (...) {
            (void)hipStreamSynchronize(s);
            for (double* p : bufs) (void)hipFree(p);
            throw;
        }
        for (double* p : bufs) (void)hipFree(p);
    });
}

int qk_rgw_excitation_density(int nOrb, int nC, int nO, int nR, int nS, const double* ERI, const double* XpY,
                              double* rho) {
    if (!ERI || !XpY || !rho) return QK_E_ARG;
    return guarded(nullptr, [&] {
        check_sizes(nOrb, nOrb, nC, nO, nOrb - nO, nR, nS);
        hipStream_t s = scratch().s;
        size_t n = nOrb, n2 = n * n;
        double *M = nullptr, *X = nullptr, *R = nullptr;
        try {
            M = dalloc<double>(n2 * nS);
            X = dalloc<double>((size_t)nS * nS);
            R = dalloc<double>(n2 * nS);
            upload_mo_block(s, ERI, M, nOrb, nC, nO, nR);
            QK_HIP(hipMemcpyAsync(X, XpY, sizeof(double) * nS * nS, hipMemcpyHostToDevice, s));
            // GW/RGW_excitation_density.f90:48-50: dgemm('N','T', nOrb^2, nS, nS)
            dgemm(s, false, true, (int)n2, nS, nS, 1.0, M, n2, X, nS, 0.0, R, n2);
            QK_HIP(hipMemcpyAsync(rho, R, sizeof(double) * n2 * nS, hipMemcpyDeviceToHost, s));
            QK_HIP(hipStreamSynchronize(s));
        } catch (...) {
            (void)hipStreamSynchronize(s);
            dfree(M); dfree(X); dfree(R);
            throw;
        }
        dfree(M); dfree(X); dfree(R);
    });
}

int qk_rgw_phbse(int dophBSE2, int exchange_kernel, int TDA_W, int TDA, int dBSE, int dTDA, int singlet, int triplet,
                 double eta, int nOrb, int nC, int nO, int nV, int nR, int nS, const double* ERI, const double* eW,
                 const double* eGW, double* EcBSE, double* OmBSE) {
    if (dBSE) {
        g_last_error = "qk_rgw_phbse: the dynamical correction has its outputs in qk_rgw_phbse_dyn";
        return QK_E_ARG;
    }
    return qk_rgw_phbse_dyn(dophBSE2, exchange_kernel, TDA_W, TDA, 0, dTDA, singlet, triplet, eta, nOrb, nC, nO, nV,
                            nR, nS, ERI, eW, eGW, EcBSE, OmBSE, nullptr, nullptr);
}

int qk_rgw_phbse_dyn(int dophBSE2, int exchange_kernel, int TDA_W, int TDA, int dBSE, int dTDA, int singlet,
                     int triplet, double eta, int nOrb, int nC, int nO, int nV, int nR, int nS, const double* ERI,
                     const double* eW, const double* eGW, double* EcBSE, double* OmBSE, double* OmDyn, double* ZDyn) {
    if (nOrb <= 0 || nC < 0 || nO <= nC || nR < 0 || nV != nOrb - nO || nS != (nO - nC) * (nV - nR) || nS <= 0 ||
        !ERI || !eW || !eGW || !EcBSE || !OmBSE || (dBSE && (!OmDyn || !ZDyn)))
        return QK_E_ARG;
    if ((dophBSE2 || dBSE) && (nC || nR)) {
        // QuAcK's BSE2 and dynamical kernels index pairs as (a-nO) + (i-1)(nBas-nO)
        g_last_error = "qk_rgw_phbse_dyn: the second-order kernel and the dynamical correction need nC = nR = 0";
        return QK_E_UNSUPPORTED;
    }
    qk_gw_handle c = nullptr;
    int rc = qk_gw_create(&c, nOrb, nOrb, nC, nO, nV, nR, nS);
    if (rc != QK_OK) return rc;
    rc = qk_gw_set_eri_mo(c, ERI);
    if (rc == QK_OK) rc = guarded(&c->err, [&] {
        const int n = nOrb, o = nO - nC, v = nV - nR;
        const size_t S2 = (size_t)nS * nS;
        // GW/RGW_phBSE.f90:113-118: dRPA screening at eW and its excitation densities (FP64)
        c->precision = 0;
        c->eig_mode = 0;
        h2d(c, c->e, eW, n);
        screening(c, c->e, eW, TDA_W, true);
        excitation(c);
        // static kernels (GW/RGW_phBSE_static_kernel_A.f90, _B.f90), as two
        // matrix-core products over the RPA modes m:
        //   KA(ia,jb) = 4 sum_m [w_m rho(i,j,m)] rho(a,b,m) -> G1[(i + o j), (a + v b)]
        //   KB(ia,jb) = 4 sum_m [w_m rho(i,b,m)] rho(a,j,m) -> G2[(i + o b), (a + v j)]
        std::vector<double> KA(S2, 0.0), KB(S2, 0.0);
        {
            double* X = dalloc<double>((size_t)nS * o * std::max(o, v));
            double* Y = nullptr;
            double* G = nullptr;
            try {
                Y = dalloc<double>((size_t)nS * v * std::max(o, v));
                G = dalloc<double>(S2 > (size_t)o * o * v * v ? S2 : (size_t)o * o * v * v);
                std::vector<double> hG((size_t)o * o * v * v);
                rho_block(c->s, c->rhoT, c->Om, eta, 1, n, nS, nC, o, nC, o, X);
                rho_block(c->s, c->rhoT, c->Om, eta, 0, n, nS, nO, v, nO, v, Y);
                dgemm(c->s, true, false, o * o, v * v, nS, 1.0, X, nS, Y, nS, 0.0, G, o * o);
                QK_HIP(hipMemcpyAsync(hG.data(), G, sizeof(double) * hG.size(), hipMemcpyDeviceToHost, c->s));
                QK_HIP(hipStreamSynchronize(c->s));
                for (int i = 0; i < o; ++i)
                    for (int a = 0; a < v; ++a)
                        for (int j = 0; j < o; ++j)
                            for (int b = 0; b < v; ++b)
                                KA[(a + (size_t)v * i) + nS * (b + (size_t)v * j)] =
                                    4.0 * hG[(i + (size_t)o * j) + (size_t)o * o * (a + (size_t)v * b)];
                if (!TDA) {
                    rho_block(c->s, c->rhoT, c->Om, eta, 1, n, nS, nC, o, nO, v, X);
                    rho_block(c->s, c->rhoT, c->Om, eta, 0, n, nS, nO, v, nC, o, Y);
                    dgemm(c->s, true, false, o * v, v * o, nS, 1.0, X, nS, Y, nS, 0.0, G, o * v);
                    hG.resize(S2);
                    QK_HIP(hipMemcpyAsync(hG.data(), G, sizeof(double) * S2, hipMemcpyDeviceToHost, c->s));
                    QK_HIP(hipStreamSynchronize(c->s));
                    for (int i = 0; i < o; ++i)
                        for (int a = 0; a < v; ++a)
                            for (int j = 0; j < o; ++j)
                                for (int b = 0; b < v; ++b)
                                    KB[(a + (size_t)v * i) + nS * (b + (size_t)v * j)] =
                                        4.0 * hG[(i + (size_t)o * b) + (size_t)o * v * (a + (size_t)v * j)];
                }
            } catch (...) {
                (void)hipStreamSynchronize(c->s);
                dfree(X); dfree(Y); dfree(G);
                throw;
            }
            dfree(X); dfree(Y); dfree(G);
        }
        auto G = [&](int p, int q, int r, int t) {  // ERI(p,q,r,s), 0-based
            return ERI[p + (size_t)n * (q + (size_t)n * (r + (size_t)n * t))];
        };
        // device buffers of the second-order kernel and the dynamical correction
        double *dW = nullptr, *dE = nullptr, *dX = nullptr, *dY = nullptr, *dPart = nullptr, *dK = nullptr;
        auto free_extra = [&] {
            (void)hipStreamSynchronize(c->s);
            dfree(dW); dfree(dE); dfree(dX); dfree(dY); dfree(dPart); dfree(dK);
        };
        try {
        if (dophBSE2 || dBSE) {
            dE = dalloc<double>(2 * (size_t)n);  // eW, then eGW
            QK_HIP(hipMemcpyAsync(dE, eW, sizeof(double) * n, hipMemcpyHostToDevice, c->s));
            QK_HIP(hipMemcpyAsync(dE + n, eGW, sizeof(double) * n, hipMemcpyHostToDevice, c->s));
        }
        if (dophBSE2) {
            // GW/RGW_phBSE_static_kernel.f90:40-60: W(p,s,q,r) = -ERI(p,s,q,r) + 4 chi(pq, rs),
            // chi = sum_m rho(p,q,m) w_m rho(r,s,m) as one FP64 matrix-core product
            const size_t n2 = (size_t)n * n, n4 = n2 * n2;
            double* Xw = dalloc<double>(n2 * nS);
            double* chi = nullptr;
            try {
                chi = dalloc<double>(n4);
                rho_block(c->s, c->rhoT, c->Om, eta, 1, n, nS, 0, n, 0, n, Xw);
                dgemm(c->s, true, false, (int)n2, (int)n2, nS, 1.0, Xw, nS, c->rhoT, nS, 0.0, chi, n2);
                std::vector<double> hchi(n4), hW(n4);
                QK_HIP(hipMemcpyAsync(hchi.data(), chi, sizeof(double) * n4, hipMemcpyDeviceToHost, c->s));
                QK_HIP(hipStreamSynchronize(c->s));
                for (int r = 0; r < n; ++r)
                    for (int q = 0; q < n; ++q)
                        for (int sI = 0; sI < n; ++sI)
                            for (int p = 0; p < n; ++p)
                                hW[p + n * (sI + n * (q + (size_t)n * r))] =
                                    -G(p, sI, q, r) + 4.0 * hchi[(p + (size_t)n * q) + n2 * (r + (size_t)n * sI)];
                dW = dalloc<double>(n4);
                QK_HIP(hipMemcpyAsync(dW, hW.data(), sizeof(double) * n4, hipMemcpyHostToDevice, c->s));
                QK_HIP(hipStreamSynchronize(c->s));
            } catch (...) {
                (void)hipStreamSynchronize(c->s);
                dfree(Xw); dfree(chi);
                throw;
            }
            dfree(Xw); dfree(chi);
        }
        if (dBSE) {
            dX = dalloc<double>(nS);
            dY = dalloc<double>(nS);
            dPart = dalloc<double>(6 * (size_t)bse_dyn_blocks(nS));
            for (int k = 0; k < 20; ++k) OmDyn[k] = ZDyn[k] = 0.0;
        }
        // singlet (ispin 1) and triplet (ispin 2): phRLR_A/B with dRPA = false at
        // eGW (LR/phRLR_A.f90:43-115, LR/phRLR_B.f90:34-95) plus the kernels,
        // then phRLR (GW/RGW_phBSE.f90:133-190)
        std::vector<double> A(S2), B(S2), XpY(S2), XmY(S2);
        for (int k = 0; k < 2 * nS; ++k) OmBSE[k] = 0.0;
        EcBSE[0] = EcBSE[1] = 0.0;
        for (int ispin = 1; ispin <= 2; ++ispin) {
            if ((ispin == 1 && !singlet) || (ispin == 2 && !triplet)) continue;
            if (ispin == 1 && dophBSE2) {
                // GW/RGW_phBSE.f90:123-138: the second-order kernels are added onto
                // KA/KB (the BSE2 routines accumulate into the array RGW_phBSE
                // passes, which holds the first-order kernel); the triplet
                // block after this one uses the kernels so changed
                dK = dalloc<double>(2 * S2);
                bse2_static(c->s, dW, dE, eta, n, nO, nS, dK, TDA ? nullptr : dK + S2);
                std::vector<double> h(2 * S2);
                QK_HIP(hipMemcpyAsync(h.data(), dK, sizeof(double) * (TDA ? S2 : 2 * S2), hipMemcpyDeviceToHost,
                                      c->s));
                QK_HIP(hipStreamSynchronize(c->s));
                for (size_t k = 0; k < S2; ++k) KA[k] += h[k];
                if (!TDA)
                    for (size_t k = 0; k < S2; ++k) KB[k] += h[S2 + k];
                dfree(dK);
            }
            const double ct1 = ispin == 1 ? 2.0 : 0.0, ct2 = -1.0;
            for (int i = 0; i < o; ++i)
                for (int a = 0; a < v; ++a) {
                    const size_t ia = a + (size_t)v * i;
                    const int I = nC + i, Aa = nO + a;
                    for (int j = 0; j < o; ++j)
                        for (int b = 0; b < v; ++b) {
                            const size_t jb = b + (size_t)v * j;
                            const int J = nC + j, Bb = nO + b;
                            const double d = G(Bb, I, J, Aa);
                            double aij = ct1 * d + ct2 * G(Bb, J, Aa, I) + KA[ia + nS * jb];
                            if (i == j && a == b) aij += eGW[Aa] - eGW[I];
                            A[ia + nS * jb] = aij;
                            B[ia + nS * jb] = TDA ? 0.0 : ct1 * d + ct2 * G(Bb, J, I, Aa) + KB[ia + nS * jb];
                        }
                }
            double ec = 0.0;
            double* Om = OmBSE + (size_t)(ispin - 1) * nS;
            const int r = qk_phrlr(TDA, nS, A.data(), B.data(), &ec, Om, XpY.data(), XmY.data());
            QK_CHECK(r == QK_OK, r, "qk_rgw_phbse: phRLR failed: " + g_last_error);
            EcBSE[ispin - 1] = ec;
            if (!dBSE) continue;
            // GW/RGW_phBSE_dynamic_perturbation.f90:73-122 for the first
            // min(nS, 10) excitations of this manifold: the dynamic kernels'
            // quadratic forms on the device, the static ones here
            const int nb = bse_dyn_blocks(nS);
            std::vector<double> X(nS), Y(nS), part(6 * (size_t)nb);
            for (int ia = 0; ia < std::min(nS, 10); ++ia) {
                for (int k = 0; k < nS; ++k) {
                    X[k] = 0.5 * (XpY[ia + (size_t)nS * k] + XmY[ia + (size_t)nS * k]);
                    Y[k] = 0.5 * (XpY[ia + (size_t)nS * k] - XmY[ia + (size_t)nS * k]);
                }
                QK_HIP(hipMemcpyAsync(dX, X.data(), sizeof(double) * nS, hipMemcpyHostToDevice, c->s));
                QK_HIP(hipMemcpyAsync(dY, Y.data(), sizeof(double) * nS, hipMemcpyHostToDevice, c->s));
                bse_dyn(c->s, c->rhoT, c->Om, dE + n, dophBSE2 ? dW : nullptr, dX, dY, Om[ia], eta, n, nO, nS,
                        dTDA ? 1 : 0, dPart);
                QK_HIP(hipMemcpyAsync(part.data(), dPart, sizeof(double) * part.size(), hipMemcpyDeviceToHost, c->s));
                QK_HIP(hipStreamSynchronize(c->s));
                double qd[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
                for (int bI = 0; bI < nb; ++bI)
                    for (int k = 0; k < 6; ++k) qd[k] += part[6 * (size_t)bI + k];
                auto quad = [&](const std::vector<double>& u, const std::vector<double>& M, const std::vector<double>& w) {
                    double t = 0.0;
                    for (int jb = 0; jb < nS; ++jb) {
                        double row = 0.0;
                        for (int kk = 0; kk < nS; ++kk) row += M[kk + (size_t)nS * jb] * u[kk];
                        t += row * w[jb];
                    }
                    return t;
                };
                double z, om;
                if (dTDA) {
                    z = qd[1];
                    om = qd[0] - quad(X, KA, X);
                } else {
                    z = qd[1] + qd[3];
                    om = (qd[0] - quad(X, KA, X)) - (qd[2] - quad(Y, KA, Y)) + (qd[4] - quad(X, KB, Y)) -
                         (qd[5] - quad(Y, KB, X));
                }
                z = 1.0 / (1.0 - z);
                ZDyn[10 * (ispin - 1) + ia] = z;
                OmDyn[10 * (ispin - 1) + ia] = z * om;
            }
        }
        EcBSE[1] *= 3.0;  // GW/RGW_phBSE.f90:193
        if (exchange_kernel) {
            EcBSE[0] *= 0.5;
            EcBSE[1] *= 0.5;
        }
        } catch (...) {
            free_extra();
            throw;
        }
        free_extra();
    });
    if (rc != QK_OK && c) g_last_error = c->err;
    qk_gw_destroy(c);
    return rc;
}

int qk_set_split_aotomo(int on) {
    qk::g_split_aotomo.store(on ? 1 : 0);
    return QK_OK;
}

int qk_set_graphs(int on) {
    qk::g_graphs.store(on ? 1 : 0);
    return QK_OK;
}

int qk_set_autotune(int on) {
    qk::set_autotune(on != 0);
    return QK_OK;
}

int qk_autotune_reset(void) {
    qk::autotune_reset();
    qk::autotune_reset_bf16();
    return QK_OK;
}

int qk_set_library_gemm(int on) {
    qk::set_library_gemm(on != 0);
    return QK_OK;
}

int qk_autotune_report(char* buf, int len) {
    if (!buf || len <= 0) return QK_E_ARG;
    std::string r = qk::autotune_report() + qk::autotune_report_bf16();
    std::snprintf(buf, (size_t)len, "%s", r.c_str());
    return (int)r.size() < len ? QK_OK : QK_E_ARG;
}

}  // extern "C"
