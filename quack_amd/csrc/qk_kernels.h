// Host-side launchers of the GW-path kernels (qk_gw_kernels.hip).
#pragma once
#include <algorithm>

#include "qk_common.h"

namespace qk {

// Strides (in elements) between the molecules of a molecule-batched launch,
// whose grid z index is the molecule: n-vectors (orbital energies, Sig, Z,
// the Newton state, done/nit flags), nS-vectors (Om, dE, sd), nS x nS
// matrices (K, the RPA matrix and its eigenvectors), excitation densities
// (n^2 nS doubles), self-energy partial sums, the scalar block (8 doubles)
// and the one-int counters.  Zero, with one molecule, for a single engine.
struct Zs {
    long long n = 0, S = 0, S2 = 0, R = 0, P = 0, c = 0, i = 0;
};

// Unrestricted particle-hole pairs (LR/phULR_A.f90:55-118): spin s has active
// occupied i = nC[s] .. nO[s]-1 and virtuals a = nO[s] .. n-nR[s]-1 (0-based);
// pair ia = (a - nO[s]) + v_s (i - nC[s]), v_s = n - nR[s] - nO[s], the beta
// block after the nSa alpha pairs.
struct UPairs {
    int n = 0, nC[2] = {0, 0}, nO[2] = {0, 0}, nR[2] = {0, 0}, nSa = 0, nSt = 0;
    UPairs() = default;
    UPairs(int n_, const int* nC_, const int* nO_, const int* nR_) : n(n_) {
        for (int s = 0; s < 2; ++s) {
            nC[s] = nC_[s];
            nO[s] = nO_[s];
            nR[s] = nR_[s];
        }
        nSa = occ(0) * vir(0);
        nSt = nSa + occ(1) * vir(1);
    }
    int occ(int s) const { return nO[s] - nC[s]; }
    int vir(int s) const { return n - nR[s] - nO[s]; }
};

void permute_block(hipStream_t s, const double* X, double* M, int n, int o, int v);
// f (N^4, QuAcK's ERI.bin order) from the 8-fold packed chemists' array
// (N(N+1)/2 pairs, pair-of-pairs packed)
void unpack_eri_s8(hipStream_t s, const double* packed, double* f, int N);
void extract_k(hipStream_t s, const double* M, double* K, int n, int nC, int nO, int v, int nS);
// scal[0] = Tr(A), scal[1] = EcRPA, scal[2] = EcGM
void rpa_build(hipStream_t s, const double* e, const double* K, double* dE, double* sd, double* Cm, double* scal,
               int nC, int nO, int v, int nS, int tda, const Zs& zs = Zs(), int nz = 1);
// rpa_build without the pair energies: dE, sd given
void rpa_build_c(hipStream_t s, const double* K, const double* dE, const double* sd, double* Cm, double* scal, int nS,
                 int tda, const Zs& zs = Zs(), int nz = 1);
// Unrestricted pairs (alpha block, then beta): K = (ia|jb) / 2 from the spin
// slabs M_s[p + n q + n^2 jb] = (pq|jb), and the pair energies from
// e = (e_alpha, e_beta)
void extract_ku(hipStream_t s, const double* M0, const double* M1, double* K, const UPairs& up);
void delta_u(hipStream_t s, const double* e, double* dE, double* sd, const UPairs& up);
void rpa_finalize(hipStream_t s, double* Om, double* W, const double* sd, double* scal, int nS, int tda, bool scale_w,
                  const Zs& zs = Zs(), int nz = 1);
int sigma_chunks(int nS, int nact);
// rhoT holds doubles, or floats when rho_f32
void sigma_partial(hipStream_t s, const void* rhoT, bool rho_f32, const double* w, const double* e, const double* Om,
                   const int* skip, double eta, int srg, double flow, int n, int nC, int nO, int nR, int nS,
                   int nchunk, double* part, const Zs& zs = Zs(), int nz = 1);
void sigma_finalize(hipStream_t s, const double* part, int nchunk, int n, int nC, int nO, int nR, double* Sig,
                    double* Z, double* scal, const Zs& zs = Zs(), int nz = 1);
void sigma_out(hipStream_t s, const double* part, int nchunk, int n, int nC, int nR, double* Sig, double* dSig);
void lin_qp(hipStream_t s, const double* e0, const double* Z, const double* Sig, double* out, int n, int zflag,
            const Zs& zs = Zs(), int nz = 1);
// work: nS n^2 + n^3 doubles for the plain weight's matrix-core path (NULL: the direct kernel)
void sigma_full(hipStream_t s, const double* rhoT, const double* e, const double* Om, double eta, int srg, double flow,
                int n, int nC, int nO, int nR, int nS, double* Sig, double* work);
void rho_block(hipStream_t s, const double* rhoT, const double* Om, double eta, int weighted, int n, int nS, int p0,
               int P, int q0, int Q, double* out);
// Second-order static BSE kernels (GW/RGW_phBSE2_static_kernel_A.f90, _B.f90)
// from the 4-index W of GW/RGW_phBSE_static_kernel.f90 (KB2 skipped when null).
void bse2_static(hipStream_t s, const double* W, const double* e, double eta, int n, int nO, int nS, double* KA2,
                 double* KB2);
// The quadratic forms of the first-order dynamical BSE correction for one
// excitation (GW/RGW_phBSE_dynamic_perturbation.f90): 6 partial sums per
// block of bse_dyn_blocks(nS) blocks in part.
int bse_dyn_blocks(int nS);
void bse_dyn(hipStream_t s, const double* rhoT, const double* OmR, const double* e, const double* W, const double* X,
             const double* Y, double Om, double eta, int n, int nO, int nS, int dtda, double* part);
void qp_init(hipStream_t s, const double* eGWlin, double* w, double* f, int* nit, int* done, int n, int nC, int nR,
             const Zs& zs = Zs(), int nz = 1);
void qp_update(hipStream_t s, const double* part, int nchunk, const double* eHF, double* w, double* f, double* df,
               int* nit, int* done, int n, double thresh, int maxit, int* active, const Zs& zs = Zs(), int nz = 1);
// eG: energies in G; frozen orbitals of eGW are set to them (eOld in evGW).
void qp_finish(hipStream_t s, const double* w, const double* df, const int* nit, const double* eGWlin,
               const double* eG, double* eGW, double* Z, int n, int nC, int nR, int maxit, const Zs& zs = Zs(),
               int nz = 1);
// One evGW cycle's bookkeeping for every molecule of a batch (GW/evRGW.f90:156-183,
// DIIS of utils/DIIS_extrapolation.f90), on the device: state[4 m + {0, 1, 2}] =
// live flag, cycles run, DIIS history length; hist: 2 n maxd doubles per
// molecule; conv[m] = Conv of the cycle; outSig/outZ/outOm/outEc (EcRPA, EcGM)
// keep the live molecules' outputs of the cycle; e <- the next cycle's
// energies.  maxd <= 32.
void evgw_update(hipStream_t s, const double* eNew, double* e, const double* Sig, const double* Z, const double* Om,
                 const double* scal, double* hist, int* state, double* conv, double* outSig, double* outZ,
                 double* outOm, double* outEc, int n, int nS, int maxd, double thresh, int maxSCF, const Zs& zs,
                 int nmol);
void transpose(hipStream_t s, const double* in, double* out, int rows, int cols, long long ld_in, long long ld_out,
               int batch = 1, long long bs_in = 0, long long bs_out = 0);
void apb_amb(hipStream_t s, const double* A, const double* B, double* ApB, double* AmB, long long total);
void colscale_pow(hipStream_t s, const double* U, const double* d, double* out, int n, double pw);
void rowscale_pow(hipStream_t s, double* X, const double* sv, int n, double pw);
void trace(hipStream_t s, const double* A, int n, double* out);
void phrlr_ab(hipStream_t s, const double* G, const double* e, double* out, int n, int nC, int nO, int v, int nS,
              double lam, int dRPA, int isB);

// Eigenvector refinement (clusters: lab[i] = cluster label of vector i;
// members of cluster c are idx[ptr[c] .. ptr[c+1]), scl[slot] = cluster of a
// member slot, boff[c] = offset of its k x k block in packed Q / block arrays).
void eig_lambda(hipStream_t s, const double* S, const double* G, double* lam, int n);
void eig_update(hipStream_t s, double* S, const double* G, const double* lam, const int* lab, int n);
void eig_check(hipStream_t s, const double* S, const double* G, const double* lam, const int* lab, int n,
               double* colmax, int* colarg);
void gather_blocks(hipStream_t s, const double* A, const double* B, long long lda, const int* ptr, const int* idx,
                   const long long* boff, int ncl, double* outA, double* outB);
// D[:, idx[slot]] <- sum_q A[:, cidx[ptr[c]+q]] Q_c[q, r] (A read in full before
// D is written, so D = A is allowed); tmp: nrows*m
void mix_cols(hipStream_t s, const double* A, long long lda, int nrows, const int* cidx, double* D, long long ldd,
              const int* idx, const int* ptr, const int* scl, const double* Q, const long long* boff, int m,
              double* tmp);
// A[idx[slot], :] <- sum_q Q_c[q, r] A'[idx[ptr[c]+q], :]; tmp: n*m
void mix_rows(hipStream_t s, double* A, int n, const int* idx, const int* ptr, const int* scl, const double* Q,
              const long long* boff, int m, double* tmp);
void gather_cols(hipStream_t s, const double* X, int n, const int* idx, int m, double* out);
// (i, a) block of density-fitting factors: Bov[P + naux ia] = B[P + naux (i + n a)]
void gather_bov(hipStream_t s, const double* B, double* Bov, int naux, int n, int nC, int nO, int v, int nS);
// Householder tridiagonalisation, upper triangle, LAPACK dsytrd layout (qk_sytrd.hip);
// work holds sytrd_work_doubles(n, nb) doubles, nb <= 64.
size_t sytrd_work_doubles(int n, int nb);
void sytrd_upper(hipStream_t s, int n, double* A, double* D, double* E, double* TAU, double* work, int nb);
// Z <- Q Z with the reflectors sytrd_upper left in A, TAU (LAPACK dormtr,
// SIDE = 'L', UPLO = 'U', TRANS = 'N'); work holds ormtr_work_doubles(n).
size_t ormtr_work_doubles(int n);
void ormtr_upper(hipStream_t s, int n, const double* A, const double* TAU, double* Z, double* work);
// A symmetric from its upper triangle: A[i, j] = A[j, i] for i > j
void sym_lower(hipStream_t s, double* A, int n);
void permute_cols(hipStream_t s, const double* X, const int* perm, double* out, int n);

}  // namespace qk
