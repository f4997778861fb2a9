/*
 * quack_amd: C-ABI of the MI355X (gfx950) G0W0/evGW engine: QuAcK's restricted
 * path, and beside it phBSE@GW, the qsGW cycle, the unrestricted G0W0/evGW
 * and static phBSE, the RHF Fock builds and the device integral producer.
 *
 * Conventions (identical to QuAcK's Fortran, so a bind(C) interface passes
 * its arrays straight through):
 *   - all arrays are FP64, Fortran column-major, caller-owned host memory
 *     unless the name ends in _device;
 *   - ERI(p,q,r,s) = <pq|rs> = (pr|qs), physicist order
 *     (utils/read_2e_integrals.f90:55-63);
 *   - particle-hole index ia = (a-nO) + (i-nC-1)*(nBas-nR-nO) with 1-based
 *     i, a (LR/phRLR_A.f90:151), a fastest;
 *   - XpY(m, ia): row = RPA mode m, column = ph pair (LR/phRLR.f90:77-78);
 *   - rho(p, q, m) (GW/RGW_excitation_density.f90:24);
 *   - unrestricted entries take per-spin counts nC[2], nO[2], nR[2] (alpha,
 *     beta) and nBas x 2 energies; their pairs are the alpha pairs, then the
 *     beta pairs, each indexed as above (LR/phULR_A.f90).
 * Every entry point returns QK_OK (0) or one of the QK_E_* codes below; a
 * message for the last failure is kept per handle (qk_gw_last_error) and per
 * thread for the handle-less entry points (qk_last_error).
 * Calls on one handle are not re-entrant; every call is synchronous at
 * return.  Call from outside OpenMP parallel regions.
 */
#ifndef QUACK_AMD_H
#define QUACK_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
    QK_OK = 0,
    QK_E_ARG = 1,          /* invalid size, flag or pointer */
    QK_E_HIP = 2,          /* HIP runtime or device error */
    QK_E_EIG = 3,          /* eigensolver failure (rocSOLVER info != 0) */
    QK_E_NOCONV = 4,       /* evGW reached maxSCF+1 cycles (GW/evRGW.f90:190-199) */
    QK_E_UNSUPPORTED = 5,  /* option not implemented on the GPU path */
    QK_E_ALLOC = 6,        /* device memory exhausted */
    QK_E_UNSTABLE = 7      /* A-B not positive definite: the linear-response instability
                              QuAcK warns about (LR/phRLR.f90:57-58) before it takes
                              square roots of the negative modes */
};

typedef struct qk_gw_ctx *qk_gw_handle;

/* ---- handle-based path: device-resident state across evGW cycles ------- */

/* Sizes as in GW/RGW.f90:44-50.  nS must equal (nO-nC)*(nOrb-nR-nO).
 * Replaces the allocation block of GW/RG0W0.f90:109-110 / evRGW.f90:99-100. */
int qk_gw_create(qk_gw_handle *h, int nBas, int nOrb, int nC, int nO, int nV, int nR, int nS);
int qk_gw_destroy(qk_gw_handle h);
const char *qk_gw_last_error(qk_gw_handle h);

/* AO->MO transformation of the block (pq|jb) that the GW path needs
 * (j active occupied, b active virtual), on the device.  Replaces the call
 * AOtoMO_ERI_RHF(nBas,nOrb,c,ERI_AO,ERI_MO) of QuAcK/RQuAcK.f90:459 for this
 * path (AOtoMO/AOtoMO_ERI_RHF.f90:1-55).  ERI_AO: nBas^4, c: nBas x nOrb. */
int qk_gw_set_eri_ao(qk_gw_handle h, const double *ERI_AO, const double *c);
/* Same with ERI_AO and c already in device memory (the handle keeps no
 * reference to them after return). */
int qk_gw_set_eri_ao_device(qk_gw_handle h, const double *ERI_AO_device, const double *c_device);
/* The same from the 8-fold packed AO integrals that PyDuck writes to
 * int/ERI_chem.bin (chemists' notation, pyscf's s8 order) and QuAcK's HPC
 * branch reads (utils/read_2e_integrals.f90:85-107; element (ab|cd) at the
 * Yoshimine index of utils/utils.f90:1186-1217).  n_packed must be
 * p(p+1)/2 with p = nBas(nBas+1)/2.  Only the packed array crosses to the
 * device (an eighth of nBas^4); it is unpacked there. */
int qk_gw_set_eri_ao_packed(qk_gw_handle h, const double *ERI_chem, long long n_packed, const double *c);
/* Two-electron integrals of contracted s and p shells (e.g. cc-pVDZ hydrogen)
 * generated on the device into eri_device (n_eri = nBas^4 doubles, nBas =
 * sum of 2l+1), in the order QuAcK streams from int/ERI.bin
 * (utils/read_2e_integrals.f90:55-63): eri[a + N b + N^2 c + N^3 d] = <ab|cd>.
 * No counterpart in QuAcK: it stands in for PySCF's mol.intor('int2e') of
 * PyDuck.py:329-345, with the conventions of the project's host producer
 * (p functions x, y, z; contracted functions normalised to one).  Shell i has
 * angular momentum sl[i] (0 or 1; QK_E_UNSUPPORTED otherwise), snp[i] <= 8
 * primitives with exponents and raw contraction coefficients taken in order
 * from exps/coefs, and sits on atom satom[i] at xyz[3*satom[i] .. +2] (bohr). */
int qk_int_eri_device(int nshell, const int *sl, const int *snp, const int *satom, const double *exps,
                      const double *coefs, const double *xyz, double *eri_device, long long n_eri);
/* MO integrals as QuAcK passes them to RG0W0/evRGW (GW/RGW.f90:85-86,102-103):
 * ERI_MO is nOrb^4; only the (pq|jb) block is copied to the device. */
int qk_gw_set_eri_mo(qk_gw_handle h, const double *ERI_MO);

/* Density-fitted MO integrals (pq|rs) = sum_P B(P,p,q) B(P,r,s) (chemists'
 * notation; ERI(p,q,r,s) = <pq|rs> = sum_P B(P,p,r) B(P,q,s)), B[P + naux*(p +
 * nOrb*q)], symmetric in p, q; requires nBas = nOrb.  The (ia|jb) block is
 * formed at once in the handle's current precision; the excitation densities
 * then run as rho = B^T (Bov W) and the four-index block is never formed.
 * The path of QuAcK's density-fitted callers; QuAcK's RGW itself takes ERI_MO. */
int qk_gw_set_eri_df(qk_gw_handle h, const double *B, int naux);
int qk_gw_set_eri_df_device(qk_gw_handle h, const double *B_device, int naux);

/* G0W0@RHF: GW/RG0W0.f90:88-233 (screening at eHF, excitation densities,
 * diagonal self-energy, linearised or graphical QP solve, second RPA at eGW
 * for EcRPA).  Om receives the screening energies at eHF (nS); any output
 * pointer except eGW may be NULL.  linDM/BSE/ACFDT/ppBSE are not on this
 * path: callers keep QuAcK's CPU routines for them. */
int qk_rg0w0(qk_gw_handle h, const double *eHF, double eta, int linearize, int TDA_W, int doSRG, double *eGW,
             double *SigC, double *Z, double *Om, double *EcRPA, double *EcGM);

/* evGW@RHF: GW/evRGW.f90:104-200.  eGW_start (may be NULL) replaces the
 * eGW = eHF start of evRGW.f90:110; NULL reproduces QuAcK.  Returns
 * QK_E_NOCONV (outputs filled with the last cycle) where QuAcK stops with
 * "Convergence failed". */
int qk_evrgw(qk_gw_handle h, const double *eHF, const double *eGW_start, int maxSCF, double thresh, int max_diis,
             double eta, int linearize, int TDA_W, int doSRG, double *eGW, double *SigC, double *Z, double *Om,
             double *EcRPA, double *EcGM, int *nSCF);

/* Screening at energies e (phRLR at e + excitation densities, as one cycle of
 * GW/evRGW.f90:123-130), kept in the handle for qk_gw_sigma_at.  Used to build
 * the learned warm start's features; no separate counterpart in QuAcK. */
int qk_gw_screen(qk_gw_handle h, const double *e, int TDA_W);
/* With the last screening (qk_gw_screen, or the last cycle of qk_evrgw):
 * Sig(p) = Sigma_c(w_p) and dSig(p) = dSigma_c/dw(w_p)
 * for every active orbital, G built from eG (GW/RGW_Re_SigC.f90,
 * GW/RGW_Re_dSigC.f90 and the SRG variants; zero for frozen orbitals). */
/* The full correlation self-energy matrix of qsGW at energies e
 * (GW/RGW_self_energy.f90, or GW/RGW_SRG_self_energy.f90 with doSRG), from the
 * screening of the last qk_gw_screen (which must run in precision 0: FP64
 * excitation densities, else QK_E_UNSUPPORTED): SigC nOrb x nOrb
 * column-major (not yet symmetrised), Z, EcGM, and the screening's EcRPA. */
int qk_gw_sigma_full(qk_gw_handle h, const double *e, double eta, int doSRG, double *SigC, double *Z, double *EcGM,
                     double *EcRPA);
int qk_gw_sigma_at(qk_gw_handle h, const double *eG, const double *w, double eta, int doSRG, double *Sig,
                   double *dSig);

/* Precision of the excitation-density contraction rho = sum_jb (pq|jb) XpY(m,jb)
 * (GW/RGW_excitation_density.f90:48-50): 0 = FP64 MFMA (default, QuAcK's
 * arithmetic); 1 = split precision: each FP64 operand as three BF16 slices,
 * six slice products on BF16 matrix cores, FP32 accumulation over 256-term
 * chunks, FP64 recombination (dropped terms O(2^-24) relative); 2 = the same
 * with two BF16 slices and three slice products (dropped terms O(2^-16)
 * relative, half the matrix-core work).  Modes 1 and 2 store rho in FP32.  The
 * RPA eigensolve and the QP Newton solve stay FP64 in every mode; with
 * density-fitted input the (ia|jb) product that builds the RPA matrix uses
 * three slices in both split modes.  No counterpart in QuAcK (which is FP64
 * throughout). */
int qk_gw_set_precision(qk_gw_handle h, int mode);

/* Eigensolver of the RPA problem inside qk_evrgw, qk_gw_screen and the first
 * RPA of qk_rg0w0 (LR/phRLR.f90:77, diagonalize_matrix = dsyev):
 *   mode 0 (default): rocSOLVER dsyevd for every RPA, as QuAcK re-diagonalises
 *           every cycle;
 *   mode 2: full solves by the library's own Householder tridiagonalisation
 *           (LAPACK dsytrd layout), rocSOLVER dstedc on the tridiagonal and the
 *           library's own blocked back-transformation (in place of dormtr);
 *   mode 1 (or 3 = 1 with the full solves of mode 2):
 *           when no orbital energy moved by more than de_max (Ha) since the
 *           handle's last eigenvectors were computed (same TDA_W), refine those
 *           eigenvectors by Newton steps for X^T X = I, X^T C X diagonal (FP64
 *           matrix-core GEMMs, quadratic convergence, stopped when the largest
 *           correction is below 1e-8, or below 1e-5 in the split-precision
 *           modes of qk_gw_set_precision, with that step's update applied;
 *           close eigenvalues handled as clusters);
 *           a full solve otherwise, or when a cluster outgrows 32 vectors,
 *           merging clusters does not settle, or eight steps do not converge.
 * Both modes return FP64 eigenpairs of the same matrix.  No counterpart in
 * QuAcK.  Mode 1 keeps four nS x nS FP64 work arrays on the device. */
int qk_gw_set_eigensolver(qk_gw_handle h, int mode, double de_max);
/* counts[0..n-1], n <= 9, since the handle was created: full solves, refined
 * solves, Newton steps, refinements abandoned for a full solve; of those,
 * abandoned because a cluster of close eigenvalues outgrew 32, because merging
 * clusters did not settle, because the steps did not converge; the largest
 * cluster met; full solves by the own tridiagonalisation. */
int qk_gw_eig_stats(qk_gw_handle h, long long *counts, int n);

/* Newton iterations of every orbital in the handle's last graphical QP solve
 * (nit[nOrb]; 64 = QuAcK's maxIt, where the orbital keeps its starting energy,
 * GW/RGW_QP_graph.f90:78-81; 0 for frozen orbitals).  No counterpart in QuAcK,
 * which prints them. */
int qk_gw_qp_iters(qk_gw_handle h, int *nit);

/* Device time (ms, HIP events on the handle's stream) of the stages of the
 * last qk_rg0w0 / qk_evrgw call: [0] AO->MO block (last set_eri_*),
 * [1] RPA matrix build, [2] eigensolve, [3] excitation densities,
 * [4] self-energy + QP solve, [5] total of the GW call (host clock),
 * [6] the excitation-density GEMM launches alone (the product that writes rho),
 * [7] their number (tuning trials of the autotuner included);
 * [8], [9] the same two summed over every call since the handle was created;
 * [10] device time of the own Householder tridiagonalisations (eigensolver
 * modes 2 and 3) since the handle was created, [11] their number;
 * [12], [13] device time of the tridiagonal divide-and-conquer (rocSOLVER
 * dstedc) and of the own back-transformation in those solves with vectors.
 * n <= 14. */
int qk_gw_stage_times(qk_gw_handle h, double *ms, int n);

/* ---- per-routine entry points (host in, host out) ---------------------- */

/* AOtoMO/AOtoMO_ERI_RHF.f90:1-55: full ERI_MO(nOrb^4). */
int qk_aotomo_eri_rhf(int nBas, int nOrb, const double *c, const double *ERI_AO, double *ERI_MO);

/* LR/phRLR.f90:1-97 for general Aph, Bph (nS x nS): EcRPA, Om (ascending),
 * XpY, XmY.  TDA != 0 uses Aph only (phRLR.f90:38-43).  Returns QK_E_UNSTABLE,
 * outputs untouched, when the smallest eigenvalue of A - B is not positive
 * (QuAcK prints its warning there and goes on with NaN square roots). */
int qk_phrlr(int TDA, int nS, const double *Aph, const double *Bph, double *EcRPA, double *Om, double *XpY,
             double *XmY);

/* LR/phRLR_A.f90:43-89 and LR/phRLR_B.f90:34-75, singlet (ispin = 1). */
int qk_phrlr_a(int dRPA, int nBas, int nC, int nO, int nV, int nR, int nS, double lambda, const double *e,
               const double *ERI, double *Aph);
int qk_phrlr_b(int dRPA, int nBas, int nC, int nO, int nV, int nR, int nS, double lambda, const double *ERI,
               double *Bph);

/* GW/RGW_excitation_density.f90:1-80: rho(nOrb,nOrb,nS) from ERI_MO(nOrb^4). */
int qk_rgw_excitation_density(int nOrb, int nC, int nO, int nR, int nS, const double *ERI, const double *XpY,
                              double *rho);

/* GW/RGW_self_energy_diag.f90:1-113 (doSRG = 0) or
 * GW/RGW_SRG_self_energy_diag.f90:1-160 with flow 500 (doSRG != 0). */
int qk_rgw_self_energy_diag(int doSRG, double eta, int nOrb, int nC, int nO, int nV, int nR, int nS, const double *e,
                            const double *Om, const double *rho, double *EcGM, double *SigC, double *Z);

/* GW/RGW_QP_graph.f90:1-97: Newton solve of the QP equation; Z is in/out as
 * in QuAcK (unchanged for an orbital whose solve hits maxIt = 64, and for
 * frozen orbitals).  Frozen entries of eGW, which QuAcK leaves unassigned,
 * are set to eOld: the value they hold at QuAcK's call (GW/evRGW.f90:154). */
int qk_rgw_qp_graph(int doSRG, double eta, double flow, int nOrb, int nC, int nO, int nV, int nR, int nS,
                    const double *eHF, const double *Om, const double *rho, const double *eGWlin, const double *eOld,
                    double *eGW, double *Z);
/* Allocation-free form of qk_rgw_qp_graph.  No counterpart in QuAcK: it is for
 * callers that cannot allocate device memory (code under hipGraph capture,
 * serving loops, harnesses that forbid hipMalloc).  Every array is a device
 * pointer; rhoT holds rho(p, q, m) at rhoT[m + nS (p + nOrb q)] (the engine's
 * layout, modes contiguous).  The Newton state and partial sums live in a
 * library-owned static device array of 4 MiB per GPU (held until the process
 * ends): 8 (3 nOrb nchunk + 3 nOrb) + 4 (2 nOrb + 1) bytes with
 * nchunk = ceil(nS / 1024) must fit, else QK_E_ARG and nothing runs.  The
 * call enqueues a fixed 64 Newton sweeps (finished orbitals masked on the
 * device, results identical to qk_rgw_qp_graph) on `stream` (a hipStream_t;
 * NULL = the null stream) and returns without synchronising; it calls no
 * hipMalloc/hipFree.  Calls that overlap on one GPU share the array and race.
 * Z is in/out as in qk_rgw_qp_graph.  Messages: qk_noalloc_last_error. */
int qk_rgw_qp_graph_noalloc(int doSRG, double eta, double flow, int nOrb, int nC, int nO, int nV, int nR, int nS,
                            const double *eHF_device, const double *Om_device, const double *rhoT_device,
                            const double *eGWlin_device, const double *eOld_device, double *eGW_device,
                            double *Z_device, void *stream);
const char *qk_noalloc_last_error(void);

/* Unrestricted G0W0 (GW/UG0W0.f90:92-150): spin-conserving
 * dRPA screening at eHF over both spin blocks (LR/phULR_A.f90, LR/phULR_B.f90,
 * LR/phULR.f90; nSa + nSb modes), excitation densities
 * (GW/UGW_excitation_density.f90) as FP64 matrix-core products, diagonal
 * self-energies (GW/UGW_self_energy_diag.f90 or the SRG form), linearised or
 * graphical (GW/UGW_QP_graph.f90) quasiparticle energies per spin, and the RPA
 * again at eGW for EcRPA.  ERI_aaaa, ERI_aabb, ERI_bbbb: nBas^4 MO integrals
 * (physicist, column-major; in ERI_aabb the first and third indices are
 * alpha).  nC[2], nO[2], nR[2]: frozen core, occupied and frozen virtual
 * counts per spin (alpha, beta), as QuAcK's nC(nspin), nO(nspin), nR(nspin);
 * per spin 0 <= nC <= nO <= nBas - nR and nC + nR < nBas, nS_s = (nO - nC)(nBas -
 * nR - nO); one spin may have no pair (an empty beta spin, say), but
 * nS_a + nS_b >= 1.
 * eHF, eGW, SigC, Z: nBas x 2 (alpha column, then beta); frozen entries of
 * eGW are eHF (QuAcK leaves them undefined), SigC 0 and Z 1 there.  Om: nSa +
 * nSb.  EcGM[2]: per spin (QuAcK's SRG routine returns their sum). */
int qk_ug0w0(int nBas, const int *nC, const int *nO, const int *nR, const double *ERI_aaaa, const double *ERI_aabb,
             const double *ERI_bbbb, const double *eHF, double eta, int linearize, int TDA_W, int doSRG, double *eGW,
             double *SigC, double *Z, double *Om, double *EcRPA, double *EcGM);

/* Unrestricted evGW (GW/evUGW.f90:103-215), orbital counts as qk_ug0w0
 * (frozen entries of eGW stay at eHF, as in QuAcK): per cycle the
 * screening of qk_ug0w0 at the current energies (both spins), the diagonal
 * self-energies at those energies, then eHF + SigC (linearize) or the graphical
 * solve started from the previous energies; Conv = max |eGW - eOld| over both
 * spins; DIIS per spin (utils/DIIS_extrapolation.f90) when max_diis > 1.
 * Screening always uses TDA_W (QuAcK builds B unless its separate TDA flag is
 * set).  Outputs as qk_ug0w0 from the last cycle (EcRPA is the last cycle's
 * screening at its input energies); nSCF_out = cycles run.  Returns
 * QK_E_NOCONV, outputs filled, where QuAcK stops after maxSCF + 1 cycles. */
int qk_evugw(int nBas, const int *nC, const int *nO, const int *nR, const double *ERI_aaaa, const double *ERI_aabb,
             const double *ERI_bbbb, const double *eHF, int maxSCF, double thresh, int max_diis, double eta,
             int linearize, int TDA_W, int doSRG, double *eGW, double *SigC, double *Z, double *Om, double *EcRPA,
             double *EcGM, int *nSCF_out);

/* Unrestricted static phBSE@GW (GW/UGW_phBSE.f90:60-196 with the static
 * kernels GW/UGW_phBSE_static_kernel_A.f90 and _B.f90): spin-conserved dRPA
 * screening at eW, its excitation densities, then the spin-conserved
 * (spin_conserved != 0; LR/phULR_A.f90, _B.f90 ispin 1 with dRPA = false)
 * and spin-flip (spin_flip != 0; ispin 2) BSE matrices at eGW plus the
 * kernels (FP64 matrix-core products over the RPA modes), each solved by
 * phULR (TDA for the BSE as asked).  Orbital counts and integrals as
 * qk_ug0w0; eW, eGW: nBas x 2.  Outputs: EcBSE[2] (spin-conserved,
 * spin-flip; halved with exchange_kernel, :189-193), OmSC[nS_aa + nS_bb] and
 * OmSF[nS_ab + nS_ba] (nS_ab = (nO_a - nC_a)(nBas - nR_b - nO_b), nS_ba the
 * same with the spins exchanged), each ignored for a manifold not asked.  The
 * dynamical correction (dBSE) and oscillator strengths are not computed. */
int qk_ugw_phbse(int nBas, const int *nC, const int *nO, const int *nR, const double *ERI_aaaa, const double *ERI_aabb,
                 const double *ERI_bbbb, const double *eW, const double *eGW, double eta, int TDA_W, int TDA,
                 int spin_conserved, int spin_flip, int exchange_kernel, double *EcBSE, double *OmSC, double *OmSF);

/* Eigensolver of the unrestricted screening of qk_ug0w0 and qk_evugw,
 * process-wide: mode and de_max as qk_gw_set_eigensolver (0: rocSOLVER dsyevd
 * for every RPA, QuAcK's sequence and the default; 1: Newton refinement of the
 * previous cycle's eigenvectors when no energy of either spin moved by more
 * than de_max Ha, dsyevd otherwise; 2: the own tridiagonalisation for every
 * full solve; 3: both).  No counterpart in QuAcK (LR/phULR.f90 always calls
 * diagonalize_matrix). */
int qk_set_unrestricted_eigensolver(int mode, double de_max);
/* Eigensolver counts (as qk_gw_eig_stats) of the calling thread's last
 * qk_ug0w0 or qk_evugw. */
int qk_unrestricted_eig_stats(long long *counts, int n);

/* Static phBSE@GW (GW/RGW_phBSE.f90:1-201, with GW/RGW_phBSE_static_kernel_A.f90
 * and _B.f90): dRPA screening at eW (TDA_W), its excitation densities, the
 * static kernels KA/KB (two FP64 matrix-core products over the RPA modes), then
 * the singlet (singlet != 0) and triplet (triplet != 0) BSE matrices at eGW
 * (phRLR_A/B with dRPA = false, LR/phRLR_A.f90:43-115, LR/phRLR_B.f90:34-95)
 * plus the kernels, each solved by phRLR (TDA for the BSE).  ERI: nOrb^4 MO
 * integrals as RG0W0 receives them.  Outputs: EcBSE[2] (triplet x 3 as at
 * :193, both halved with exchange_kernel), OmBSE[2 nS] (singlet, then
 * triplet; zeros for a manifold not asked).  dophBSE2: the second-order
 * static kernels (GW/RGW_phBSE_static_kernel.f90 for W, FP64 matrix-core
 * product; GW/RGW_phBSE2_static_kernel_A.f90, _B.f90) are added to the
 * first-order ones before the singlet BSE, and the triplet BSE after it uses
 * the kernels so changed (QuAcK's BSE2 routines accumulate into the array
 * they receive); W costs 2 nOrb^4 doubles on the device and the host.  dBSE
 * needs qk_rgw_phbse_dyn, which returns the corrections (QK_E_ARG here).
 * dophBSE2 needs nC = nR = 0 (QuAcK's BSE2 and dynamical kernels index pairs
 * as (a-nO) + (i-1)(nBas-nO)): QK_E_UNSUPPORTED otherwise.  Oscillator
 * strengths (phLR_transition_vectors) are not computed. */
int qk_rgw_phbse(int dophBSE2, int exchange_kernel, int TDA_W, int TDA, int dBSE, int dTDA, int singlet, int triplet,
                 double eta, int nOrb, int nC, int nO, int nV, int nR, int nS, const double *ERI, const double *eW,
                 const double *eGW, double *EcBSE, double *OmBSE);
/* qk_rgw_phbse with the first-order dynamical correction when dBSE != 0
 * (GW/RGW_phBSE_dynamic_perturbation.f90:73-122, nC = nR = 0): for the first
 * min(nS, 10) excitations of each manifold asked, the dynamic kernels at the
 * excitation energy (GW/RGW_phBSE_dynamic_kernel_A.f90, _B.f90; with
 * dophBSE2 also GW/RGW_phBSE2_dynamic_kernel_A.f90, _B.f90) contracted with
 * the BSE vectors on the device; dTDA keeps the resonant part only.
 * OmDyn[20], ZDyn[20]: the renormalised corrections Z (X^T (K(Om) - K_sta) X
 * ...) and the factors Z = 1/(1 - X^T ZA X ...), singlet at 0..9, triplet at
 * 10..19 (zeros beyond nS and for a manifold not asked); QuAcK prints
 * OmBSE + OmDyn and Z. */
int qk_rgw_phbse_dyn(int dophBSE2, int exchange_kernel, int TDA_W, int TDA, int dBSE, int dTDA, int singlet,
                     int triplet, double eta, int nOrb, int nC, int nO, int nV, int nR, int nS, const double *ERI,
                     const double *eW, const double *eGW, double *EcBSE, double *OmBSE, double *OmDyn, double *ZDyn);

const char *qk_last_error(void);

/* Tile autotuner of the library's GEMMs (process-wide; on by default): the first
 * call of each large GEMM shape times every tile configuration (FP64) or tile
 * grouping and wave layout (split precision) once on the caller's stream and
 * keeps the fastest; the tuning call's output is the last candidate's.  All
 * candidates give identical results.  No counterpart in QuAcK.
 * qk_autotune_report writes one line per tuned shape (QK_E_ARG if len is too
 * small); qk_autotune_reset forgets every tuned shape, so the next large call
 * of each shape tunes again (calls that overlap it on other threads race). */
int qk_set_autotune(int on);
int qk_autotune_report(char *buf, int len);
int qk_autotune_reset(void);
/* rocBLAS's dgemm as one more autotuner candidate (process-wide; off by
 * default) for plain FP64 products: unbatched, all tiles, every dimension at
 * least 512, 32-bit indexable, outside graph capture.  It sums in its own
 * order, so with it on, results can differ from the default's in the last
 * bits; off, every candidate gives identical results.  No counterpart in
 * QuAcK (its DGEMMs are the host BLAS's). */
int qk_set_library_gemm(int on);

/* hipGraph replay of the own full eigensolve's launch sequences (process-wide;
 * off by default, since the replay measured slower than stream launches): the
 * tridiagonalisation (two launches per column) and the back-transformation are
 * captured once per matrix buffer of a handle and then replayed by one graph
 * launch each.  Results are identical either way.  No counterpart in QuAcK. */
int qk_set_graphs(int on);

/* Split-precision AO->MO block (process-wide, off by default; applies to the
 * qk_gw_set_eri_ao* and qk_gwb_set_eri_ao_device calls made while it is on):
 * the last two quarter transforms of the (pq|jb) block (AOtoMO_ERI_RHF.f90:31-51
 * restricted to the block) run as three-slice BF16 products with FP32
 * accumulation over 256-term k-chunks and FP64 recombination, and (ia|jb),
 * which sets the RPA energies, comes from a separate FP64 transform of the
 * occupied-virtual block.  No counterpart in QuAcK. */
int qk_set_split_aotomo(int on);

/* ---- batched G0W0: many molecules of one size in one handle ------------- */

/* nmol molecules with the same (nBas, nOrb, nC, nO, nV, nR, nS) on one stream:
 * each stage runs for all of them before the next, one launch per kernel for
 * the whole batch (grid z = molecule): one strided-batched rocSOLVER
 * eigensolve per RPA, one batched excitation-density GEMM in the FP64 form,
 * one launch pair per QP Newton sweep.  GW/RG0W0.f90:88-233 per molecule; no
 * counterpart in QuAcK, which runs one molecule per process.  Callers with
 * uneven sizes group molecules by size. */
typedef struct qk_gwb_ctx *qk_gwb_handle;
int qk_gwb_create(qk_gwb_handle *h, int nmol, int nBas, int nOrb, int nC, int nO, int nV, int nR, int nS);
int qk_gwb_destroy(qk_gwb_handle h);
const char *qk_gwb_last_error(qk_gwb_handle h);
int qk_gwb_set_precision(qk_gwb_handle h, int mode); /* as qk_gw_set_precision */
/* molecule m: ERI_AO + m*stride_eri (nBas^4), c + m*stride_c (nBas x nOrb), device memory */
int qk_gwb_set_eri_ao_device(qk_gwb_handle h, const double *ERI_AO_device, long long stride_eri,
                             const double *c_device, long long stride_c);
int qk_gwb_set_eri_mo(qk_gwb_handle h, int mol, const double *ERI_MO); /* host nOrb^4, one molecule */
/* eHF, eGW, SigC, Z: nOrb x nmol; Om: nS x nmol; EcRPA, EcGM: nmol (outputs may be NULL except eGW) */
int qk_gwb_g0w0(qk_gwb_handle h, const double *eHF, double eta, int linearize, int TDA_W, int doSRG, double *eGW,
                double *SigC, double *Z, double *Om, double *EcRPA, double *EcGM);
/* [0] AO->MO, [1] eigensolves, [2] excitation densities, [3] self-energy + QP (ms), n <= 4 */
/* evGW (GW/evRGW.f90:104-200) for every molecule of the batch, cycle by cycle
 * in step: one batched eigensolve, one excitation stage and one QP stage per
 * cycle for the whole batch.  Each molecule keeps its own DIIS history
 * (max_diis <= 32), convergence test and cycle count, all on the device; one
 * that has converged keeps its last cycle's outputs while the others go on.  Arrays as qk_gwb_g0w0 (nOrb x nmol,
 * nS x nmol), eGW_start NULL or nOrb x nmol (as qk_evrgw), nSCF: nmol.
 * QK_E_NOCONV when some molecule reached maxSCF + 1 cycles (its outputs are
 * those of its last cycle). */
int qk_gwb_evgw(qk_gwb_handle h, const double *eHF, const double *eGW_start, int maxSCF, double thresh, int max_diis,
                double eta, int linearize, int TDA_W, int doSRG, double *eGW, double *SigC, double *Z, double *Om,
                double *EcRPA, double *EcGM, int *nSCF);
int qk_gwb_stage_times(qk_gwb_handle h, double *ms, int n);
/* hipGraph replay of qk_gwb_evgw's cycles (off by default): from the second
 * cycle of a call, the launches before the batched eigensolve and those after
 * it (RPA finalisation, excitation densities, self-energy, all 64 Newton
 * sweeps with finished orbitals masked on the device, the DIIS and
 * convergence bookkeeping) are replayed from two graphs captured once per set
 * of cycle parameters; the host reads only the molecules' live flags and the
 * eigensolver's info after each cycle.  Results are identical with it on or
 * off.  No counterpart in QuAcK (GW/evRGW.f90:119-183 is the cycle). */
int qk_gwb_set_graphs(qk_gwb_handle h, int on);
/* counts[0] = evGW cycles run eagerly, counts[1] = cycles replayed from graphs
 * (since the handle was created), n <= 2 */
int qk_gwb_graph_stats(qk_gwb_handle h, long long *counts, int n);

/* ---- learned warm start: per-orbital MLP (no counterpart in QuAcK) ------ */

/* nin -> width (SiLU) x depth hidden layers -> 1; FP32 weights, He-uniform
 * initialisation from `seed`, zero biases.  width <= 256, multiple of 16. */
typedef struct qk_mlp_ctx *qk_mlp_handle;
int qk_mlp_create(qk_mlp_handle *h, int nin, int width, int depth, unsigned long long seed);
int qk_mlp_destroy(qk_mlp_handle h);
int qk_mlp_nparams(qk_mlp_handle h, int *n);
/* flat FP32 parameters, layer by layer: W_l (out x in, row-major), b_l */
int qk_mlp_get_params(qk_mlp_handle h, float *out);
int qk_mlp_set_params(qk_mlp_handle h, const float *in);
/* Y[n] from X[n x nin] (row-major); bf16 != 0 runs the fused BF16 matrix-core
 * kernel (FP32 accumulation), else FP32 kernels. */
int qk_mlp_forward(qk_mlp_handle h, const float *X, int n, float *Y, int bf16);
/* MSE loss mean((Y - T)^2) and its gradient with respect to the parameters. */
int qk_mlp_loss_grad(qk_mlp_handle h, const float *X, const float *T, int n, float *loss, float *grad);
/* Adam (beta 0.9/0.999, eps 1e-8) on shuffled mini-batches, on the device. */
int qk_mlp_train(qk_mlp_handle h, const float *X, const float *T, int n, int epochs, int batch, float lr,
                 unsigned long long seed, float *loss_per_epoch);
const char *qk_mlp_last_error(qk_mlp_handle h);

/* ---- restricted Hartree-Fock Fock builds (the SCF before the GW path) ----
 * The Coulomb and exchange matrices of QuAcK's RHF loop (HF/RHF.f90:178-180)
 * from the nBas^4 AO integrals resident on the device, in QuAcK's ERI.bin
 * order (ERI(mu,la,nu,si) = <mu la|nu si> at mu + N la + N^2 nu + N^3 si).
 * One handle per basis size; it owns the device copies of P, J, K and the
 * partial sums.  Not re-entrant per handle; synchronous. */
/* ---- quasiparticle self-consistent GW (GW/qsRGW.f90) -------------------- */

/* A qsGW calculation (GW/qsRGW.f90:1-431 without linDM): sizes as qk_gw_create
 * (nOrb <= nBas <= 512), DIIS on the Fock matrix with max_diis vectors, and
 * QuAcK's inputs S, T, V, Hc (nBas x nBas), X (nBas x nOrb), ERI_AO (nBas^4,
 * QuAcK's ERI.bin order), PHF (nBas x nBas), cHF (nBas x nOrb), eHF (nOrb).
 * The handle keeps the AO integrals in device memory. */
typedef struct qk_qsgw_ctx *qk_qsgw_handle;
int qk_qsgw_create(qk_qsgw_handle *h, int nBas, int nOrb, int nC, int nO, int nV, int nR, int nS, int max_diis,
                   const double *S, const double *T, const double *V, const double *Hc, const double *X,
                   const double *ERI_AO, const double *PHF, const double *cHF, const double *eHF);
/* One pass of the body of qsRGW's loop (GW/qsRGW.f90:199-341): J and K of the
 * current density, the MO integrals of the current orbitals, dRPA screening
 * at the current quasiparticle energies, the full self-energy
 * (GW/RGW_self_energy.f90, or GW/RGW_SRG_self_energy.f90 with doSRG),
 * symmetrised and taken to the AO basis, F = Hc + J + K/2 + SigC_AO
 * (projected on the kept orbitals when nOrb < nBas), the commutator
 * FPS - SPF, the energies, DIIS on F, the new orbitals and energies and the
 * density.  Outputs: eGW (nOrb), c (nBas x nOrb), SigC (nOrb x nOrb, in the
 * new orbitals), Z (nOrb), E[5] = ET, EV, EJ, EK, EqsGW, EcRPA, EcGM, and
 * err_max = max |FPS - SPF|, which QuAcK takes as Conv once its cycle counter
 * (starting at -1) exceeds 1.  The caller keeps the loop and its stopping rule
 * (Conv > thresh and nSCF < maxSCF). */
int qk_qsgw_cycle(qk_qsgw_handle h, double eta, int TDA_W, int doSRG, double *eGW, double *c, double *SigC, double *Z,
                  double *E, double *EcRPA, double *EcGM, double *err_max);
int qk_qsgw_destroy(qk_qsgw_handle h);
const char *qk_qsgw_last_error(qk_qsgw_handle h);

typedef struct qk_scf_ctx *qk_scf_handle;
int qk_rhf_create(qk_scf_handle *h, int nBas); /* nBas <= 512 */
int qk_rhf_destroy(qk_scf_handle h);
const char *qk_rhf_last_error(qk_scf_handle h);
/* Copy the host AO integrals (nBas^4, the ERI array QuAcK's RHF receives,
 * HF/RHF.f90:1-2) to device memory owned by the handle, for callers without
 * device memory of their own (Fortran). */
int qk_rhf_set_eri_ao(qk_scf_handle h, const double *ERI_AO);
/* J(mu,nu) = sum P(la,si) ERI(mu,la,nu,si)   (AOtoMO/Hartree_matrix_AO_basis.f90:22-32)
 * K(mu,nu) = -sum P(la,si) ERI(mu,la,si,nu)  (AOtoMO/exchange_matrix_AO_basis.f90:22-32)
 * P, J, K: host nBas x nBas, column-major; n_eri must be nBas^4.  ERI_device
 * NULL uses the handle's copy from qk_rhf_set_eri_ao.  One pass over the
 * integrals; the sums run in a fixed order (same result every call). */
int qk_rhf_jk(qk_scf_handle h, const double *ERI_device, long long n_eri, const double *P, double *J, double *K);
/* device time (ms, HIP events) of the last qk_rhf_jk's kernels */
int qk_rhf_jk_ms(qk_scf_handle h, double *ms);

#ifdef __cplusplus
}
#endif
#endif
