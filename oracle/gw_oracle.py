"""TEST INFRASTRUCTURE ONLY: CPU restatement of QuAcK's G0W0/evGW (restricted
and unrestricted), phBSE@GW and qsGW.

This module is the checker for the HIP path.  Only tests/, the smoke check
and the CPU-baseline leg of bench.py import it; the product library never
does.  Every function follows one QuAcK routine call-for-call (file:line in
each docstring; paths relative to QuAcK's src/), in FP64, with QuAcK's index
conventions:

* ERI arrays are the flat physicist column-major arrays QuAcK holds:
  ``G(p,q,r,s) = <pq|rs> = (pr|qs)`` at ``p + n q + n^2 r + n^3 s``;
  here they are viewed as ``G = eri.reshape((n,)*4, order='F')``.
* particle-hole index ``ia = (a-nO) + (i-nC)*nn`` (0-based), a fastest
  (LR/phRLR_A.f90:151).
* ``XpY[m, ia]``: row = RPA mode, column = ph pair (LR/phRLR.f90:77-78).

dsyev('V','U') of utils/wrap_lapack.f90:99-136 is called through SciPy's
LAPACK binding with the same driver and triangle; dsysvx of
wrap_lapack.f90:560-598 likewise.

Unrestricted functions take (alpha, beta) pairs of counts and n x 2 energies
and index spin-blocked pairs as LR/phULR_A.f90 (alpha pairs first).

Pinning: tests/test_oracle.py checks each function against the compiled
QuAcK routines (oracle/_ref), the FeatherBench / Rtest_ref.dat goldens
(He/6-31G, H2O/cc-pVDZ) and closed forms (nS = 1 He, RPA supermatrix
eigenvalues, the (X+Y)^T(X-Y) = 1 normalisation, brute-force einsums).
"""
import numpy as np
from scipy.linalg import lapack

FLOW = 500.0          # SRG flow parameter, GW/RG0W0.f90:98, GW/evRGW.f90:88
QP_MAXIT = 64         # GW/RGW_QP_graph.f90:31
QP_THRESH = 1e-6      # GW/RGW_QP_graph.f90:32
HA_TO_EV = 27.211386245988  # include/parameters.h:16


def G4(eri, n):
    return np.asarray(eri).reshape((n, n, n, n), order="F")


def diagonalize_matrix(A):
    """utils/wrap_lapack.f90:99-136: dsyev('V','U'), ascending eigenvalues."""
    w, v, info = lapack.dsyev(np.asfortranarray(A), compute_v=1, lower=0)
    if info != 0:
        raise RuntimeError("dsyev info=%d" % info)
    return w, v


def aotomo_eri_rhf(c, eri_ao):
    """AOtoMO/AOtoMO_ERI_RHF.f90:31-51: four dgemm('T','N') quarter transforms,
    each contracting the leading AO index and appending the MO index."""
    nBas, nOrb = c.shape
    a = np.asarray(eri_ao).reshape(nBas, -1, order="F")         # (i1, i2 i3 i4)
    a = (a.T @ c)                                               # (i2 i3 i4, p)
    a = a.reshape(nBas, -1, order="F").T @ c                    # (i3 i4 p, q)
    a = a.reshape(nBas, -1, order="F").T @ c                    # (i4 p q, r)
    a = a.reshape(nBas, -1, order="F").T @ c                    # (p q r, s)
    return a.reshape(-1, order="F")


def ph_dims(nOrb, nO, nC=0, nR=0):
    nV = nOrb - nO
    return nC, nO, nV, nR, (nO - nC) * (nV - nR)


def phRLR_A(e, eri, nO, lam=1.0, dRPA=True, nC=0, nR=0, ispin=1):
    """LR/phRLR_A.f90:43-115: A(ia,jb) = 2 lam <bi|ja> (singlet only)
    - (1-dRPA) lam <bj|ai> + delta_ij delta_ab (e_a - e_i); ispin 1 singlet,
    2 triplet."""
    n = len(e)
    G = G4(eri, n)
    occ, vir = slice(nC, nO), slice(nO, n - nR)
    nS = (nO - nC) * (n - nR - nO)
    A = np.zeros((nS, nS))
    if ispin == 1:
        A += 2.0 * lam * G[vir, occ, occ, vir].transpose(1, 3, 2, 0).reshape(nS, nS)
    if not dRPA:
        A -= lam * G[vir, occ, vir, occ].transpose(3, 2, 1, 0).reshape(nS, nS)
    dE = (e[None, vir] - e[occ, None]).reshape(-1)
    A[np.diag_indices(nS)] += dE
    return A


def phRLR_B(eri, n, nO, lam=1.0, dRPA=True, nC=0, nR=0, ispin=1):
    """LR/phRLR_B.f90:34-95: B(ia,jb) = 2 lam <bi|ja> (singlet only)
    - (1-dRPA) lam <bj|ia>; ispin 1 singlet, 2 triplet."""
    G = G4(eri, n)
    occ, vir = slice(nC, nO), slice(nO, n - nR)
    nS = (nO - nC) * (n - nR - nO)
    B = np.zeros((nS, nS))
    if ispin == 1:
        B += 2.0 * lam * G[vir, occ, occ, vir].transpose(1, 3, 2, 0).reshape(nS, nS)
    if not dRPA:
        B -= lam * G[vir, occ, occ, vir].transpose(2, 3, 1, 0).reshape(nS, nS)
    return B


def ADAt(A, D):
    """utils/utils.f90:634-678: B = A diag(D) A^T."""
    return A @ (D[:, None] * A.T)


def phRLR(TDA, A, B):
    """LR/phRLR.f90:36-95: returns EcRPA, Om, XpY, XmY."""
    nS = A.shape[0]
    if TDA:
        Om, V = diagonalize_matrix(A)
        XpY = V.T.copy()
        XmY = XpY.copy()
    else:
        ApB = A + B
        AmB = A - B
        Om, U = diagonalize_matrix(AmB)
        AmBSq = ADAt(U, np.sqrt(Om))
        AmBIv = ADAt(U, 1.0 / np.sqrt(Om))
        tmp = ApB @ AmBSq
        Z = AmBSq @ tmp
        Om, Z = diagonalize_matrix(Z)
        Om = np.sqrt(np.abs(Om))
        XpY = (Z.T @ AmBSq) * (1.0 / np.sqrt(Om))[:, None]
        XmY = (Z.T @ AmBIv) * np.sqrt(Om)[:, None]
    EcRPA = 0.5 * (np.sum(Om) - np.trace(A))
    return EcRPA, Om, XpY, XmY


def excitation_density(eri, XpY, n, nO, nC=0, nR=0):
    """GW/RGW_excitation_density.f90:26-50: tmp(p,q,jb) = <pj|qb>;
    rho(p,q,m) = sum_jb tmp(p,q,jb) XpY(m,jb).  Returned as rho[p, q, m]."""
    G = G4(eri, n)
    tmp = G[:, nC:nO, :, nO:n - nR].transpose(0, 2, 1, 3)           # [p, q, j, b]
    nS = XpY.shape[0]
    tmp = tmp.reshape(n, n, nS)                                      # jb = j*nn + b
    return np.einsum("pqk,mk->pqm", tmp, XpY, optimize=True)


def self_energy_diag(e, Om, rho, nO, eta, nC=0, nR=0):
    """GW/RGW_self_energy_diag.f90:115-191: SigC(p), Z(p) = 1/(1-dSig), EcGM."""
    n = len(e)
    occ, vir = slice(nC, nO), slice(nO, n - nR)
    act = slice(nC, n - nR)
    Sig = np.zeros(n)
    dS = np.zeros(n)
    r2 = rho ** 2
    eps = e[act, None, None] - e[None, occ, None] + Om[None, None, :]
    num = 2.0 * r2[act, occ, :]
    Sig[act] += np.sum(num * eps / (eps ** 2 + eta ** 2), axis=(1, 2))
    dS[act] -= np.sum(num * (eps ** 2 - eta ** 2) / (eps ** 2 + eta ** 2) ** 2, axis=(1, 2))
    eps = e[act, None, None] - e[None, vir, None] - Om[None, None, :]
    num = 2.0 * r2[act, vir, :]
    Sig[act] += np.sum(num * eps / (eps ** 2 + eta ** 2), axis=(1, 2))
    dS[act] -= np.sum(num * (eps ** 2 - eta ** 2) / (eps ** 2 + eta ** 2) ** 2, axis=(1, 2))
    eps = e[vir, None, None] - e[None, occ, None] + Om[None, None, :]
    num = 4.0 * r2[vir, occ, :]
    EcGM = -np.sum(num * eps / (eps ** 2 + eta ** 2))
    Z = 1.0 / (1.0 - dS)
    return Sig, Z, EcGM


def srg_self_energy_diag(e, Om, rho, nO, flow=FLOW, nC=0, nR=0):
    """GW/RGW_SRG_self_energy_diag.f90:64-160."""
    n = len(e)
    occ, vir = slice(nC, nO), slice(nO, n - nR)
    act = slice(nC, n - nR)
    s = flow
    Sig = np.zeros(n)
    dS = np.zeros(n)
    r2 = rho ** 2
    for blk, sgn in ((occ, 1.0), (vir, -1.0)):
        D = e[act, None, None] - e[None, blk, None] + sgn * Om[None, None, :]
        reg = 1.0 - np.exp(-2.0 * s * D * D)
        Sig[act] += np.sum(2.0 * r2[act, blk, :] * reg / D, axis=(1, 2))
        dS[act] -= np.sum(2.0 * r2[act, blk, :] * reg / D ** 2, axis=(1, 2))
    D = e[vir, None, None] - e[None, occ, None] + Om[None, None, :]
    EcGM = -np.sum(4.0 * r2[vir, occ, :] * (1.0 - np.exp(-2.0 * s * D * D)) / D)
    return Sig, 1.0 / (1.0 - dS), EcGM


def re_sigc_dsigc(p, w, e, Om, rho, nO, eta, srg=False, flow=FLOW, nC=0, nR=0):
    """GW/RGW_Re_SigC.f90:127-147 and GW/RGW_Re_dSigC.f90:179-199 (SRG:
    GW/RGW_SRG_Re_SigC.f90, GW/RGW_SRG_Re_dSigC.f90) for one orbital p."""
    n = len(e)
    occ, vir = slice(nC, nO), slice(nO, n - nR)
    sig = 0.0
    dsig = 0.0
    for blk, sgn in ((occ, 1.0), (vir, -1.0)):
        eps = w - e[blk, None] + sgn * Om[None, :]
        num = 2.0 * rho[p, blk, :] ** 2
        if srg:
            reg = 1.0 - np.exp(-2.0 * flow * eps * eps)
            sig += np.sum(num * reg / eps)
            dsig -= np.sum(num * reg / eps ** 2)
        else:
            sig += np.sum(num * eps / (eps ** 2 + eta ** 2))
            dsig -= np.sum(num * (eps ** 2 - eta ** 2) / (eps ** 2 + eta ** 2) ** 2)
    return sig, dsig


def qp_graph(eHF, Om, rho, eGWlin, eOld, Z, nO, eta, srg=False, nC=0, nR=0):
    """GW/RGW_QP_graph.f90:50-92: per-orbital Newton on w - eHF(p) - Sig(w) = 0.
    Start at eGWlin(p); Sig built from eOld; |f| > 1e-6 tested before each
    update; at nIt == 64 eGW(p) = eGWlin(p) and Z(p) keeps its input value,
    otherwise Z(p) = df of the last step.  Frozen orbitals (p < nC, p >= n-nR)
    are never assigned there, so eGW keeps the caller's values: eOld at the
    call in GW/evRGW.f90:154.  Returns eGW, Z, iterations."""
    n = len(eHF)
    eGW = np.array(eOld, dtype=np.float64)
    Z = np.array(Z, dtype=np.float64)
    its = np.zeros(n, dtype=np.int64)
    for p in range(nC, n - nR):
        w = eGWlin[p]
        nIt = 0
        f = 1.0
        df = 0.0
        while abs(f) > QP_THRESH and nIt < QP_MAXIT:
            nIt += 1
            sig, dsig = re_sigc_dsigc(p, w, eOld, Om, rho, nO, eta, srg, nC=nC, nR=nR)
            f = w - eHF[p] - sig
            df = 1.0 / (1.0 - dsig)
            w = w - df * f
        its[p] = nIt
        if nIt == QP_MAXIT:
            eGW[p] = eGWlin[p]
        else:
            eGW[p] = w
            Z[p] = df
    return eGW, Z, its


def linear_solve(A, b):
    """utils/wrap_lapack.f90:560-598: dsysvx('N','U'); returns x, rcond."""
    res = lapack.dsysvx(np.asfortranarray(A), b.reshape(-1, 1), lower=0)
    # SciPy returns (a_s, udut, ipiv, b_s, x, rcond, ferr, berr, info)
    x, rcond = res[4], res[5]
    return x.reshape(-1), float(rcond)


class DIIS:
    """utils/DIIS_extrapolation.f90:1-61 with utils/utils.f90:506-536 (prepend)."""

    def __init__(self, n, max_diis):
        self.err = np.zeros((n, max_diis))
        self.e = np.zeros((n, max_diis))
        self.n_diis = 0
        self.max_diis = max_diis

    def __call__(self, err_in, e_inout):
        self.n_diis = min(self.n_diis + 1, self.max_diis)
        m = self.n_diis
        self.err[:, 1:m] = self.err[:, 0:m - 1].copy()
        self.err[:, 0] = err_in
        self.e[:, 1:m] = self.e[:, 0:m - 1].copy()
        self.e[:, 0] = e_inout
        A = np.zeros((m + 1, m + 1))
        A[:m, :m] = self.err[:, :m].T @ self.err[:, :m]
        A[:m, m] = -1.0
        A[m, :m] = -1.0
        b = np.zeros(m + 1)
        b[m] = -1.0
        w, rcond = linear_solve(A, b)
        if rcond > 1e-13:
            e_inout = self.e[:, :m] @ w[:m]
        return e_inout, rcond


def screening(e, eri, nO, TDA_W=False, nC=0, nR=0):
    n = len(e)
    A = phRLR_A(e, eri, nO, nC=nC, nR=nR)
    B = np.zeros_like(A) if TDA_W else phRLR_B(eri, n, nO, nC=nC, nR=nR)
    return phRLR(TDA_W, A, B)


def rg0w0(eri_mo, eHF, nO, eta=0.0, linearize=False, TDA_W=False, doSRG=False, nC=0, nR=0):
    """GW/RG0W0.f90:88-233 (linDM, BSE, ACFDT and ppBSE branches off).
    Frozen orbitals of eGW, undefined in QuAcK's G0W0 (eGW is freshly
    allocated and GW/RGW_QP_graph.f90 does not assign them), are eHF here."""
    n = len(eHF)
    EcRPA, Om, XpY, XmY = screening(eHF, eri_mo, nO, TDA_W, nC, nR)
    rho = excitation_density(eri_mo, XpY, n, nO, nC, nR)
    if doSRG:
        SigC, Z, EcGM = srg_self_energy_diag(eHF, Om, rho, nO, nC=nC, nR=nR)
    else:
        SigC, Z, EcGM = self_energy_diag(eHF, Om, rho, nO, eta, nC=nC, nR=nR)
    eGWlin = eHF + Z * SigC
    its = None
    if linearize:
        eGW = eGWlin.copy()
    else:
        eGW, Z, its = qp_graph(eHF, Om, rho, eGWlin, eHF, Z, nO, eta, doSRG, nC=nC, nR=nR)
    EcRPA2, _, _, _ = screening(eGW, eri_mo, nO, TDA_W, nC, nR)
    return {"eGW": eGW, "SigC": SigC, "Z": Z, "Om": Om, "EcRPA": EcRPA2, "EcGM": EcGM,
            "eGWlin": eGWlin, "rho": rho, "XpY": XpY, "XmY": XmY, "EcRPA0": EcRPA, "qp_iters": its}


def evrgw(eri_mo, eHF, nO, maxSCF=256, thresh=1e-5, max_diis=5, eta=0.0, linearize=False, TDA_W=False,
          doSRG=False, nC=0, nR=0):
    """GW/evRGW.f90:104-200.  status 1 where QuAcK prints 'Convergence failed'.
    Frozen orbitals keep eHF throughout (eGW = eHF at :110, never assigned by
    GW/RGW_QP_graph.f90, zero DIIS error there)."""
    n = len(eHF)
    eGW = eHF.copy()
    eOld = eGW.copy()
    Z = np.ones(n)
    diis = DIIS(n, max(max_diis, 1))
    nSCF = 0
    Conv = 1.0
    hist = []
    while Conv > thresh and nSCF <= maxSCF:
        EcRPA, Om, XpY, XmY = screening(eGW, eri_mo, nO, TDA_W, nC, nR)
        rho = excitation_density(eri_mo, XpY, n, nO, nC, nR)
        if doSRG:
            SigC, Z, EcGM = srg_self_energy_diag(eGW, Om, rho, nO, nC=nC, nR=nR)
        else:
            SigC, Z, EcGM = self_energy_diag(eGW, Om, rho, nO, eta, nC=nC, nR=nR)
        if linearize:
            eGW = eHF + SigC
        else:
            eGW, Z, _ = qp_graph(eHF, Om, rho, eOld, eOld, Z, nO, eta, doSRG, nC=nC, nR=nR)
        Conv = np.max(np.abs(eGW - eOld))
        if max_diis > 1:
            eGW, _ = diis(eGW - eOld, eGW)
        eOld = eGW.copy()
        hist.append(eGW.copy())
        nSCF += 1
    status = 1 if nSCF == maxSCF + 1 else 0
    return {"eGW": eGW, "SigC": SigC, "Z": Z, "Om": Om, "EcRPA": EcRPA, "EcGM": EcGM, "nSCF": nSCF,
            "status": status, "Conv": Conv, "history": hist}


def phbse_static_kernel_A(eta, Om, rho, nO, lam=1.0, nC=0, nR=0):
    """GW/RGW_phBSE_static_kernel_A.f90:20-40:
    KA(ia,jb) = 4 lam^2 sum_m rho(i,j,m) rho(a,b,m) Om_m / (Om_m^2 + eta^2)."""
    n = rho.shape[0]
    occ, vir = slice(nC, nO), slice(nO, n - nR)
    w = Om / (Om ** 2 + eta ** 2)
    K = np.einsum("ijm,abm,m->iajb", rho[occ, occ], rho[vir, vir], w, optimize=True)
    nS = (nO - nC) * (n - nR - nO)
    return 4.0 * lam ** 2 * K.reshape(nS, nS)


def phbse_static_kernel_B(eta, Om, rho, nO, lam=1.0, nC=0, nR=0):
    """GW/RGW_phBSE_static_kernel_B.f90:20-40:
    KB(ia,jb) = 4 lam^2 sum_m rho(i,b,m) rho(a,j,m) Om_m / (Om_m^2 + eta^2)."""
    n = rho.shape[0]
    occ, vir = slice(nC, nO), slice(nO, n - nR)
    w = Om / (Om ** 2 + eta ** 2)
    K = np.einsum("ibm,ajm,m->iajb", rho[occ, vir], rho[vir, occ], w, optimize=True)
    nS = (nO - nC) * (n - nR - nO)
    return 4.0 * lam ** 2 * K.reshape(nS, nS)


def _fw(x, eta):
    """x/(x^2 + eta^2), the regularised 1/x of the BSE kernels."""
    return x / (x ** 2 + eta ** 2)


def _gw(x, eta):
    """(x^2 - eta^2)/(x^2 + eta^2)^2, minus its derivative."""
    return (x ** 2 - eta ** 2) / (x ** 2 + eta ** 2) ** 2


def phbse_static_screening(eta, eri, Om, rho, lam=1.0):
    """GW/RGW_phBSE_static_kernel.f90:40-60: the static screened interaction
    W(p,s,q,r) = -lam ERI(p,s,q,r) + 4 lam sum_m rho(p,q,m) rho(r,s,m) Om_m/(Om_m^2+eta^2),
    returned as W[p, s, q, r] (the 4-index array QuAcK holds)."""
    n = rho.shape[0]
    chi = np.einsum("pqm,rsm,m->pqrs", rho, rho, Om / (Om ** 2 + eta ** 2), optimize=True)
    return -lam * G4(eri, n) + 4.0 * lam * chi.transpose(0, 3, 1, 2)


def _bse2_blocks(W, nO):
    n = W.shape[0]
    o, v = slice(0, nO), slice(nO, n)
    return o, v, n - nO


def phbse2_static_kernel_A(eta, eW, W, nO):
    """GW/RGW_phBSE2_static_kernel_A.f90:36-90 (nC = nR = 0): the second-order
    static kernel of the resonant block, from W and the energies eW."""
    o, v, nv = _bse2_blocks(W, nO)
    eo, ev = eW[o], eW[v]
    F1 = _fw(-(ev[None, :] - eo[:, None]), eta)            # [k, c]
    F2 = _fw(+(ev[None, :] - eo[:, None]), eta)
    Fvv = _fw(-(ev[:, None] + ev[None, :]), eta)            # [c, d]
    Foo = _fw(-(eo[:, None] + eo[None, :]), eta)            # [k, l]
    K = -2.0 * np.einsum("jkic,acbk,kc->iajb", W[o, o, o, v], W[v, v, v, o], F1, optimize=True)
    K += 2.0 * np.einsum("jcik,akbc,kc->iajb", W[o, v, o, o], W[v, o, v, v], F2, optimize=True)
    K += 2.0 * np.einsum("ajcd,cdib,cd->iajb", W[v, o, v, v], W[v, v, o, v], Fvv, optimize=True)
    K -= 2.0 * np.einsum("ajkl,klib,kl->iajb", W[v, o, o, o], W[o, o, o, v], Foo, optimize=True)
    nS = nO * nv
    return K.reshape(nS, nS)


def phbse2_static_kernel_B(eta, eW, W, nO):
    """GW/RGW_phBSE2_static_kernel_B.f90:36-90 (nC = nR = 0): the
    second-order static kernel of the coupling block."""
    o, v, nv = _bse2_blocks(W, nO)
    eo, ev = eW[o], eW[v]
    F1 = _fw(-(ev[None, :] - eo[:, None]), eta)
    F2 = _fw(+(ev[None, :] - eo[:, None]), eta)
    Fvv = _fw(-(ev[:, None] + ev[None, :]), eta)
    Foo = _fw(-(eo[:, None] + eo[None, :]), eta)
    K = -2.0 * np.einsum("bkic,acjk,kc->iajb", W[v, o, o, v], W[v, v, o, o], F1, optimize=True)
    K += 2.0 * np.einsum("bcik,akjc,kc->iajb", W[v, v, o, o], W[v, o, o, v], F2, optimize=True)
    K += 2.0 * np.einsum("abcd,cdij,cd->iajb", W[v, v, v, v], W[v, v, o, o], Fvv, optimize=True)
    K -= 2.0 * np.einsum("abkl,klij,kl->iajb", W[v, v, o, o], W[o, o, o, o], Foo, optimize=True)
    nS = nO * nv
    return K.reshape(nS, nS)


def phbse_dynamic_kernel_A(eta, eGW, OmRPA, rho, nO, Om):
    """GW/RGW_phBSE_dynamic_kernel_A.f90:38-82 (nC = nR = 0, lambda 1) at the
    excitation energy Om: KA(ia,jb) = -2 sum_m rho_ij rho_ab [f(Om - Om_m -
    (e_a - e_j)) + f(Om - Om_m - (e_b - e_i))], ZA with g in place of f and +2."""
    n = rho.shape[0]
    o, v = slice(0, nO), slice(nO, n)
    roo, rvv = rho[o, o, :], rho[v, v, :]
    D1 = Om - OmRPA[None, None, :] - (eGW[v][:, None, None] - eGW[o][None, :, None])   # [a, j, m]
    D2 = Om - OmRPA[None, None, :] - (eGW[v][:, None, None] - eGW[o][None, :, None])   # [b, i, m]
    KA = -2.0 * (np.einsum("ijm,abm,ajm->iajb", roo, rvv, _fw(D1, eta), optimize=True)
                 + np.einsum("ijm,abm,bim->iajb", roo, rvv, _fw(D2, eta), optimize=True))
    ZA = 2.0 * (np.einsum("ijm,abm,ajm->iajb", roo, rvv, _gw(D1, eta), optimize=True)
                + np.einsum("ijm,abm,bim->iajb", roo, rvv, _gw(D2, eta), optimize=True))
    nS = nO * (n - nO)
    return KA.reshape(nS, nS), ZA.reshape(nS, nS)


def phbse_dynamic_kernel_B(eta, eGW, OmRPA, rho, nO):
    """GW/RGW_phBSE_dynamic_kernel_B.f90:34-62 (nC = nR = 0, lambda 1):
    KB(ia,jb) = -2 sum_m rho_ib rho_ja [f(-Om_m - (e_b - e_i)) + f(-Om_m - (e_a - e_j))]."""
    n = rho.shape[0]
    o, v = slice(0, nO), slice(nO, n)
    rov = rho[o, v, :]
    E = -OmRPA[None, None, :] - (eGW[v][:, None, None] - eGW[o][None, :, None])   # [b, i, m] / [a, j, m]
    KB = -2.0 * (np.einsum("ibm,jam,bim->iajb", rov, rov, _fw(E, eta), optimize=True)
                 + np.einsum("ibm,jam,ajm->iajb", rov, rov, _fw(E, eta), optimize=True))
    nS = nO * (n - nO)
    return KB.reshape(nS, nS)


def phbse2_dynamic_kernel_A(eta, eGW, W, nO, Om):
    """GW/RGW_phBSE2_dynamic_kernel_A.f90:36-90 (nC = nR = 0) at Om: the
    second-order dynamic kernel and its derivative term."""
    o, v, nv = _bse2_blocks(W, nO)
    eo, ev = eGW[o], eGW[v]
    # [a, c, k, j]: Om - e_a - e_c + e_k + e_j
    d1 = Om - ev[:, None, None, None] - ev[None, :, None, None] + eo[None, None, :, None] + eo[None, None, None, :]
    # [i, c, k, b]: Om + e_i - e_c + e_k - e_b
    d2 = Om + eo[:, None, None, None] - ev[None, :, None, None] + eo[None, None, :, None] - ev[None, None, None, :]
    # [i, j, c, d]: Om + e_i + e_j - e_c - e_d
    d3 = Om + eo[:, None, None, None] + eo[None, :, None, None] - ev[None, None, :, None] - ev[None, None, None, :]
    # [a, b, k, l]: Om - e_a - e_b + e_k + e_l
    d4 = Om - ev[:, None, None, None] - ev[None, :, None, None] + eo[None, None, :, None] + eo[None, None, None, :]
    KA = np.zeros((nO, nv, nO, nv))
    ZA = np.zeros((nO, nv, nO, nv))
    t1 = ("jkic,acbk,ackj->iajb", W[o, o, o, v], W[v, v, v, o], d1)
    t2 = ("jcik,akbc,ickb->iajb", W[o, v, o, o], W[v, o, v, v], d2)
    t3 = ("ajcd,cdib,ijcd->iajb", W[v, o, v, v], W[v, v, o, v], d3)
    t4 = ("ajkl,klib,abkl->iajb", W[v, o, o, o], W[o, o, o, v], d4)
    for (eq, A1, A2, d), sg in ((t1, -1.0), (t2, -1.0), (t3, 1.0), (t4, 1.0)):
        KA += sg * 2.0 * np.einsum(eq, A1, A2, _fw(d, eta), optimize=True)
        ZA -= sg * 2.0 * np.einsum(eq, A1, A2, _gw(d, eta), optimize=True)
    nS = nO * nv
    return KA.reshape(nS, nS), ZA.reshape(nS, nS)


def phbse2_dynamic_kernel_B(eta, eGW, W, nO):
    """GW/RGW_phBSE2_dynamic_kernel_B.f90:36-84 (nC = nR = 0)."""
    o, v, nv = _bse2_blocks(W, nO)
    eo, ev = eGW[o], eGW[v]
    d1 = -ev[:, None, None, None] - ev[None, :, None, None] + eo[None, None, :, None] + eo[None, None, None, :]  # [a,c,k,j]
    d2 = eo[:, None, None, None] - ev[None, :, None, None] + eo[None, None, :, None] - ev[None, None, None, :]   # [i,c,k,b]
    d3 = eo[:, None, None, None] + eo[None, :, None, None] - ev[None, None, :, None] - ev[None, None, None, :]   # [i,j,c,d]
    d4 = -ev[:, None, None, None] - ev[None, :, None, None] + eo[None, None, :, None] + eo[None, None, None, :]  # [a,b,k,l]
    KB = -2.0 * np.einsum("bkic,acjk,ackj->iajb", W[v, o, o, v], W[v, v, o, o], _fw(d1, eta), optimize=True)
    KB -= 2.0 * np.einsum("bcik,akjc,ickb->iajb", W[v, v, o, o], W[v, o, o, v], _fw(d2, eta), optimize=True)
    KB += 2.0 * np.einsum("abcd,cdij,ijcd->iajb", W[v, v, v, v], W[v, v, o, o], _fw(d3, eta), optimize=True)
    KB += 2.0 * np.einsum("abkl,klij,abkl->iajb", W[v, v, o, o], W[o, o, o, o], _fw(d4, eta), optimize=True)
    nS = nO * nv
    return KB.reshape(nS, nS)


DYN_MAXS = 10  # GW/RGW_phBSE_dynamic_perturbation.f90:36


def phbse_dynamic_perturbation(dophBSE2, dTDA, eta, eGW, OmRPA, rho, nO, OmBSE, XpY, XmY, KA_sta, KB_sta, W=None):
    """GW/RGW_phBSE_dynamic_perturbation.f90:73-122: first-order dynamical
    correction of the first min(nS, 10) BSE excitations.  Returns (Om_dyn,
    Z_dyn): the renormalised corrections Z (X^T (K(Om) - K_sta) X ...) and Z."""
    nS = len(OmBSE)
    m = min(nS, DYN_MAXS)
    Om_dyn, Z_dyn = np.zeros(m), np.zeros(m)
    for ia in range(m):
        X = 0.5 * (XpY[ia, :] + XmY[ia, :])
        Y = 0.5 * (XpY[ia, :] - XmY[ia, :])
        KAp, ZAp = phbse_dynamic_kernel_A(eta, eGW, OmRPA, rho, nO, +OmBSE[ia])
        if dophBSE2:
            K2, Z2 = phbse2_dynamic_kernel_A(eta, eGW, W, nO, OmBSE[ia])
            KAp, ZAp = KAp + K2, ZAp + Z2
        if dTDA:
            z = X @ ZAp @ X
            om = X @ (KAp - KA_sta) @ X
        else:
            KAm, ZAm = phbse_dynamic_kernel_A(eta, eGW, OmRPA, rho, nO, -OmBSE[ia])
            KB = phbse_dynamic_kernel_B(eta, eGW, OmRPA, rho, nO)
            if dophBSE2:
                KB = KB + phbse2_dynamic_kernel_B(eta, eGW, W, nO)
            z = X @ ZAp @ X + Y @ ZAm @ Y
            om = (X @ (KAp - KA_sta) @ X - Y @ (KAm - KA_sta) @ Y + X @ (KB - KB_sta) @ Y
                  - Y @ (KB - KB_sta) @ X)
        Z_dyn[ia] = 1.0 / (1.0 - z)
        Om_dyn[ia] = Z_dyn[ia] * om
    return Om_dyn, Z_dyn


def rgw_phbse(eri, eW, eGW, nO, eta=0.0, TDA_W=False, TDA=False, singlet=True, triplet=True,
              exchange_kernel=False, nC=0, nR=0, dophBSE2=False, dBSE=False, dTDA=True, dyn=None):
    """GW/RGW_phBSE.f90:100-195: dRPA screening at eW, rho, the static kernels
    KA/KB, then the singlet and triplet BSE matrices at eGW (phRLR_A/B with
    dRPA = false) plus the kernels, each solved by phRLR.  dophBSE2 (singlet
    block, :123-138): the second-order static kernels from W are added to
    KA/KB (QuAcK's BSE2 routines accumulate into their output array, which
    holds the first-order kernel), and the triplet block after it uses the
    kernels so changed.  dBSE (:150-152, :181-183): the first-order dynamical
    correction of each manifold (dTDA: its resonant part only), returned in
    dyn["singlet"/"triplet"] = (Om_dyn, Z_dyn) when a dict is passed.
    BSE2 and dBSE follow QuAcK's pair index (a-nO) + (i-1)(nBas-nO), valid for
    nC = nR = 0 only.  Returns EcBSE(2) (triplet x 3, halved with
    exchange_kernel) and the singlet and triplet excitation energies."""
    n = len(eW)
    nS = ph_dims(n, nO, nC, nR)[4]
    if (dophBSE2 or dBSE) and (nC or nR):
        raise NotImplementedError("phBSE2 and dBSE index pairs as nC = nR = 0")
    A = phRLR_A(eW, eri, nO, dRPA=True, nC=nC, nR=nR)
    B = np.zeros((nS, nS)) if TDA_W else phRLR_B(eri, n, nO, dRPA=True, nC=nC, nR=nR)
    _, OmRPA, XpY, _ = phRLR(TDA_W, A, B)
    rho = excitation_density(eri, XpY, n, nO, nC, nR)
    KA = phbse_static_kernel_A(eta, OmRPA, rho, nO, nC=nC, nR=nR)
    KB = np.zeros((nS, nS)) if TDA else phbse_static_kernel_B(eta, OmRPA, rho, nO, nC=nC, nR=nR)
    W = phbse_static_screening(eta, eri, OmRPA, rho) if dophBSE2 else None
    EcBSE = np.zeros(2)
    Om = {1: np.zeros(nS), 2: np.zeros(nS)}
    for ispin, on in ((1, singlet), (2, triplet)):
        if not on:
            continue
        A = phRLR_A(eGW, eri, nO, dRPA=False, ispin=ispin, nC=nC, nR=nR)
        B = (np.zeros((nS, nS)) if TDA
             else phRLR_B(eri, n, nO, dRPA=False, ispin=ispin, nC=nC, nR=nR))
        if ispin == 1 and dophBSE2:
            KA = KA + phbse2_static_kernel_A(eta, eW, W, nO)
            if not TDA:
                KB = KB + phbse2_static_kernel_B(eta, eW, W, nO)
        A = A + KA
        if not TDA:
            B = B + KB
        EcBSE[ispin - 1], Om[ispin], Xp, Xm = phRLR(TDA, A, B)
        if dBSE:
            r = phbse_dynamic_perturbation(dophBSE2, dTDA, eta, eGW, OmRPA, rho, nO, Om[ispin], Xp, Xm, KA, KB, W)
            if dyn is not None:
                dyn["singlet" if ispin == 1 else "triplet"] = r
    EcBSE[1] *= 3.0
    if exchange_kernel:
        EcBSE *= 0.5
    return EcBSE, Om[1], Om[2]


def rgw_self_energy(eta, e, Om, rho, nO, nC=0, nR=0):
    """GW/RGW_self_energy.f90: the full correlation self-energy
    SigC(p,q) = sum_{r,m} 2 rho(p,r,m) rho(q,r,m) eps/(eps^2 + eta^2) with
    eps = e_p - e_r + Om_m (r occupied) or e_p - e_r - Om_m (r virtual), the
    renormalisation Z(p) and EcGM.  Returns EcGM, SigC, Z."""
    n = len(e)
    act, occ, vir = slice(nC, n - nR), slice(nC, nO), slice(nO, n - nR)
    Sig = np.zeros((n, n))
    Zs = np.zeros(n)
    for rs, sgn in ((occ, 1.0), (vir, -1.0)):
        eps = e[act, None, None] - e[None, rs, None] + sgn * Om[None, None, :]
        r_ = rho[act, rs, :]
        w = eps / (eps ** 2 + eta ** 2)
        Sig[act, act] += 2.0 * np.einsum("prm,prm,qrm->pq", r_, w, r_, optimize=True)
        Zs[act] -= np.einsum("prm,prm->p", 2.0 * r_ ** 2, (eps ** 2 - eta ** 2) / (eps ** 2 + eta ** 2) ** 2)
    Z = 1.0 / (1.0 - Zs)
    eps = e[vir, None, None] - e[None, occ, None] + Om[None, None, :]
    EcGM = -np.sum(4.0 * rho[vir, occ, :] ** 2 * eps / (eps ** 2 + eta ** 2))
    return EcGM, Sig, Z


def rgw_srg_self_energy(flow, e, Om, rho, nO, nC=0, nR=0):
    """GW/RGW_SRG_self_energy.f90: SigC(p,q) = sum_{r,m} 2 rho(p,r,m) rho(q,r,m)
    (1 - exp(-s D_p^2) exp(-s D_q^2)) (D_p + D_q)/(D_p^2 + D_q^2), D_x = e_x -
    e_r +- Om_m; Z(p) and EcGM with the SRG factors.  Returns EcGM, SigC, Z."""
    n = len(e)
    s = flow
    act, occ, vir = slice(nC, n - nR), slice(nC, nO), slice(nO, n - nR)
    Sig = np.zeros((n, n))
    Zs = np.zeros(n)
    for rs, sgn in ((occ, 1.0), (vir, -1.0)):
        D = e[act, None, None] - e[None, rs, None] + sgn * Om[None, None, :]       # [p, r, m]
        r_ = rho[act, rs, :]
        Dp, Dq = D[:, None], D[None, :]                                            # [p, q, r, m]
        w = (1.0 - np.exp(-s * Dp * Dp) * np.exp(-s * Dq * Dq)) * (Dp + Dq) / (Dp * Dp + Dq * Dq)
        Sig[act, act] += 2.0 * np.einsum("prm,qrm,pqrm->pq", r_, r_, w, optimize=True)
        Zs[act] -= np.sum(2.0 * r_ ** 2 * (1.0 - np.exp(-2.0 * s * D * D)) / D ** 2, axis=(1, 2))
    Z = 1.0 / (1.0 - Zs)
    D = e[vir, None, None] - e[None, occ, None] + Om[None, None, :]
    EcGM = -np.sum(4.0 * rho[vir, occ, :] ** 2 * (1.0 - np.exp(-2.0 * s * D * D)) / D)
    return EcGM, Sig, Z


def qsrgw(S, T, V, eri_ao, X, cHF, PHF, eHF, nO, maxSCF=64, thresh=1e-5, max_diis=5, eta=0.0, doSRG=False,
          TDA_W=False, nC=0, nR=0):
    """GW/qsRGW.f90:193-352 (no linDM, no BSE): per cycle J and
    K from the AO integrals, the MO integrals of the current orbitals, dRPA
    screening at eGW, rho, the full self-energy (symmetrised, to the AO basis),
    F = Hc + J + K/2 + SigC_AO, the commutator (Conv from nSCF > 1, nSCF
    starting at -1), DIIS on F, diagonalisation of X^T F X (nBas == nOrb) or
    of C^T F C (nBas > nOrb, where F is first projected on the kept orbitals,
    :281-284, :323-333).  X: nBas x nOrb.  Returns a dict with eGW, c, EqsGW
    (electronic), EcRPA, EcGM, Z, nSCF, Conv, converged, and c_last_cycle: the
    orbitals whose MO integrals the last cycle used, which qsRGW leaves in
    ERI_MO for the BSE after it (:210, :375)."""
    from .scf_oracle import exchange_matrix, hartree_matrix
    nB, n = X.shape
    Hc = T + V
    nSCF, Conv = -1, 1.0
    P, eGW, c = PHF.copy(), eHF.copy(), cHF.copy()
    diis = DIIS(nB * nB, max(1, max_diis))
    EqsGW = EcRPA = EcGM = 0.0
    Z = np.ones(n)
    c_prev = c
    while Conv > thresh and nSCF < maxSCF:
        nSCF += 1
        c_prev = c
        J = hartree_matrix(P, eri_ao, nB)
        K = exchange_matrix(P, eri_ao, nB)
        mo = aotomo_eri_rhf(c, eri_ao)
        A = phRLR_A(eGW, mo, nO, dRPA=True, nC=nC, nR=nR)
        B = np.zeros_like(A) if TDA_W else phRLR_B(mo, n, nO, dRPA=True, nC=nC, nR=nR)
        EcRPA, Om, XpY, _ = phRLR(TDA_W, A, B)
        rho = excitation_density(mo, XpY, n, nO, nC, nR)
        if doSRG:
            EcGM, SigC, Z = rgw_srg_self_energy(FLOW, eGW, Om, rho, nO, nC, nR)
        else:
            EcGM, SigC, Z = rgw_self_energy(eta, eGW, Om, rho, nO, nC, nR)
        SigC = 0.5 * (SigC + SigC.T)
        SigCp = S @ c @ SigC @ c.T @ S                       # AOtoMO/MOtoAO.f90
        F = Hc + J + 0.5 * K + SigCp
        if nB != n:
            F = S @ c @ (c.T @ F @ c) @ c.T @ S
        err = F @ P @ S - S @ P @ F
        if nSCF > 1:
            Conv = np.abs(err).max()
        EqsGW = np.trace(P @ T) + np.trace(P @ V) + 0.5 * np.trace(P @ J) + 0.25 * np.trace(P @ K)
        if max_diis > 1:
            f, _ = diis(err.ravel(order="F"), F.ravel(order="F").copy())
            F = f.reshape(nB, nB, order="F")
        if nB == n:
            eGW, cp = diagonalize_matrix(X.T @ F @ X)
            c = X @ cp
        else:
            eGW, cp = diagonalize_matrix(c.T @ F @ c)
            c = c @ cp
        P = 2.0 * c[:, :nO] @ c[:, :nO].T
    return {"eGW": eGW, "c": c, "c_last_cycle": c_prev, "EqsGW": EqsGW, "EcRPA": EcRPA, "EcGM": EcGM, "Z": Z,
            "nSCF": nSCF, "Conv": Conv, "converged": nSCF < maxSCF}


# ---- unrestricted G0W0 (GW/UG0W0.f90) ----------------------------------------
# Spin-resolved MO integrals: G_aa, G_ab, G_bb = G4(ERI_aaaa), G4(ERI_aabb),
# G4(ERI_bbbb) (in ERI_aabb the first and third indices are alpha, the second
# and fourth beta); e[:, 0] alpha, e[:, 1] beta; nO = (nO_alpha, nO_beta).


def _ublocks(n, nO, nC=(0, 0), nR=(0, 0)):
    """Active occupied and virtual slices of each spin (nC(s)+1..nO(s) and
    nO(s)+1..nBas-nR(s) of LR/phULR_A.f90:55-118) and the pair counts."""
    oa, ob = slice(nC[0], nO[0]), slice(nC[1], nO[1])
    va, vb = slice(nO[0], n - nR[0]), slice(nO[1], n - nR[1])
    return (oa, va, ob, vb, (nO[0] - nC[0]) * (n - nR[0] - nO[0]), (nO[1] - nC[1]) * (n - nR[1] - nO[1]))


def _sf_blocks(n, nO, nC=(0, 0), nR=(0, 0)):
    """Spin-flip pairs (LR/phULR_A.f90:120-168): alpha-occupied to
    beta-virtual (nS_ab of them), then beta-occupied to alpha-virtual."""
    oa, ob = slice(nC[0], nO[0]), slice(nC[1], nO[1])
    va, vb = slice(nO[0], n - nR[0]), slice(nO[1], n - nR[1])
    return (oa, vb, ob, va, (nO[0] - nC[0]) * (n - nR[1] - nO[1]), (nO[1] - nC[1]) * (n - nR[0] - nO[0]))


def phULR_A(e, eri_aa, eri_ab, eri_bb, nO, lam=1.0, dRPA=True, nC=(0, 0), nR=(0, 0), ispin=1):
    """LR/phULR_A.f90:26-170: ispin 1 spin-conserving (alpha pairs, then
    beta), ispin 2 spin-flip (alpha-to-beta pairs, then beta-to-alpha); pairs
    ia with i, a active in their spin, a fastest."""
    n = e.shape[0]
    Ga, Gm, Gb = G4(eri_aa, n), G4(eri_ab, n), G4(eri_bb, n)
    dr = 1.0 if dRPA else 0.0
    if ispin == 2:
        oa, vb, ob, va, nab, nba = _sf_blocks(n, nO, nC, nR)
        A = np.zeros((nab + nba, nab + nba))
        # alpha i -> beta a: delta (e_a(b) - e_i(a)) - (1-dRPA) lam <ib|ja>_aabb
        A[:nab, :nab] = -(1 - dr) * lam * np.einsum("ibja->iajb", Gm[oa, vb, oa, vb]).reshape(nab, nab)
        # beta i -> alpha a: - (1-dRPA) lam ERI_aabb(b,i,a,j)
        A[nab:, nab:] = -(1 - dr) * lam * np.einsum("biaj->iajb", Gm[va, ob, va, ob]).reshape(nba, nba)
        d = np.concatenate([(e[vb, 1][None, :] - e[oa, 0][:, None]).reshape(-1),
                            (e[va, 0][None, :] - e[ob, 1][:, None]).reshape(-1)])
        A[np.diag_indices(nab + nba)] += d
        return A
    oa, va, ob, vb, nSa, nSb = _ublocks(n, nO, nC, nR)
    A = np.zeros((nSa + nSb, nSa + nSb))
    A[:nSa, :nSa] = lam * Ga[oa, va, va, oa].transpose(0, 2, 3, 1).reshape(nSa, nSa) \
        - (1 - dr) * lam * Ga[oa, va, oa, va].transpose(0, 3, 2, 1).reshape(nSa, nSa)
    A[:nSa, nSa:] = lam * Gm[oa, vb, va, ob].transpose(0, 2, 3, 1).reshape(nSa, nSb)
    A[nSa:, :nSa] = lam * Gm[va, ob, oa, vb].transpose(1, 3, 2, 0).reshape(nSb, nSa)
    A[nSa:, nSa:] = lam * Gb[ob, vb, vb, ob].transpose(0, 2, 3, 1).reshape(nSb, nSb) \
        - (1 - dr) * lam * Gb[ob, vb, ob, vb].transpose(0, 3, 2, 1).reshape(nSb, nSb)
    d = np.concatenate([(e[va, 0][None, :] - e[oa, 0][:, None]).reshape(-1),
                        (e[vb, 1][None, :] - e[ob, 1][:, None]).reshape(-1)])
    A[np.diag_indices(nSa + nSb)] += d
    return A


def phULR_B(n, eri_aa, eri_ab, eri_bb, nO, lam=1.0, dRPA=True, nC=(0, 0), nR=(0, 0), ispin=1):
    """LR/phULR_B.f90:26-168: ispin 1 spin-conserving, ispin 2 spin-flip
    (only the alpha-to-beta / beta-to-alpha coupling blocks)."""
    Ga, Gm, Gb = G4(eri_aa, n), G4(eri_ab, n), G4(eri_bb, n)
    dr = 1.0 if dRPA else 0.0
    if ispin == 2:
        oa, vb, ob, va, nab, nba = _sf_blocks(n, nO, nC, nR)
        B = np.zeros((nab + nba, nab + nba))
        # (alpha i -> beta a, beta j -> alpha b): - (1-dRPA) lam ERI_aabb(i,j,b,a)
        B[:nab, nab:] = -(1 - dr) * lam * np.einsum("ijba->iajb", Gm[oa, ob, va, vb]).reshape(nab, nba)
        # (beta i -> alpha a, alpha j -> beta b): - (1-dRPA) lam ERI_aabb(j,i,a,b)
        B[nab:, :nab] = -(1 - dr) * lam * np.einsum("jiab->iajb", Gm[oa, ob, va, vb]).reshape(nba, nab)
        return B
    oa, va, ob, vb, nSa, nSb = _ublocks(n, nO, nC, nR)
    B = np.zeros((nSa + nSb, nSa + nSb))
    B[:nSa, :nSa] = lam * Ga[oa, oa, va, va].transpose(0, 2, 1, 3).reshape(nSa, nSa) \
        - (1 - dr) * lam * Ga[oa, oa, va, va].transpose(0, 3, 1, 2).reshape(nSa, nSa)
    B[:nSa, nSa:] = lam * Gm[oa, ob, va, vb].transpose(0, 2, 1, 3).reshape(nSa, nSb)
    B[nSa:, :nSa] = lam * Gm[oa, ob, va, vb].transpose(1, 3, 0, 2).reshape(nSb, nSa)
    B[nSa:, nSa:] = lam * Gb[ob, ob, vb, vb].transpose(0, 2, 1, 3).reshape(nSb, nSb) \
        - (1 - dr) * lam * Gb[ob, ob, vb, vb].transpose(0, 3, 1, 2).reshape(nSb, nSb)
    return B


def ugw_excitation_density(n, eri_aa, eri_ab, eri_bb, XpY, nO, nC=(0, 0), nR=(0, 0)):
    """GW/UGW_excitation_density.f90: rho[p, q, m, spin], summed over the
    active pairs jb; QuAcK fills only active p, q of each spin (:43-44,
    :86-87), the others are never read."""
    Ga, Gm, Gb = G4(eri_aa, n), G4(eri_ab, n), G4(eri_bb, n)
    oa, va, ob, vb, nSa, nSb = _ublocks(n, nO, nC, nR)
    nm = XpY.shape[0]  # explicit: a spin may have no active pairs
    Xa = XpY[:, :nSa].reshape(nm, oa.stop - oa.start, va.stop - va.start)
    Xb = XpY[:, nSa:].reshape(nm, ob.stop - ob.start, vb.stop - vb.start)
    rho = np.zeros((n, n, nSa + nSb, 2))
    # alpha: sum_jb(a) <pj|qb>_aaaa X(m, jb) + sum_jb(b) <pj|qb>_aabb X(m, nSa + jb)
    rho[:, :, :, 0] = np.einsum("pjqb,mjb->pqm", Ga[:, oa, :, va], Xa, optimize=True) \
        + np.einsum("pjqb,mjb->pqm", Gm[:, ob, :, vb], Xb, optimize=True)
    # beta: sum_jb(a) <jp|bq>_aabb X(m, jb) + sum_jb(b) <pj|qb>_bbbb X(m, nSa + jb)
    rho[:, :, :, 1] = np.einsum("jpbq,mjb->pqm", Gm[oa, :, va, :], Xa, optimize=True) \
        + np.einsum("pjqb,mjb->pqm", Gb[:, ob, :, vb], Xb, optimize=True)
    return rho


def ugw_phbse_static_kernel_A(ispin, eta, Om, rho, nO, nC=(0, 0), nR=(0, 0), lam=1.0):
    """GW/UGW_phBSE_static_kernel_A.f90:36-151: 2 lam^2 sum_m rho_s(i,j,m)
    rho_s'(a,b,m) Om_m/(Om_m^2+eta^2), on the aaaa and bbbb blocks (ispin 1;
    no coupling block) or the abab and baba blocks (ispin 2)."""
    n = rho.shape[0]
    w = Om / (Om ** 2 + eta ** 2)
    if ispin == 1:
        oa, va, ob, vb, na, nb = _ublocks(n, nO, nC, nR)
        blocks = ((oa, va, 0, 0), (ob, vb, 1, 1))
    else:
        oa, vb, ob, va, na, nb = _sf_blocks(n, nO, nC, nR)
        blocks = ((oa, vb, 0, 1), (ob, va, 1, 0))
    K = np.zeros((na + nb, na + nb))
    for (o, v, so, sv), off, m in zip(blocks, (0, na), (na, nb)):
        k = np.einsum("ijm,abm,m->iajb", rho[o, o, :, so], rho[v, v, :, sv], w, optimize=True).reshape(m, m)
        K[off:off + m, off:off + m] = 2.0 * lam ** 2 * k
    return K


def ugw_phbse_static_kernel_B(ispin, eta, Om, rho, nO, nC=(0, 0), nR=(0, 0), lam=1.0):
    """GW/UGW_phBSE_static_kernel_B.f90:36-153: 2 lam^2 sum_m rho(i,b,m)
    rho(a,j,m) Om_m/(Om_m^2+eta^2) on the aaaa and bbbb blocks (ispin 1) or
    the two coupling blocks of the spin-flip pairs (ispin 2)."""
    n = rho.shape[0]
    w = Om / (Om ** 2 + eta ** 2)
    if ispin == 1:
        oa, va, ob, vb, na, nb = _ublocks(n, nO, nC, nR)
        K = np.zeros((na + nb, na + nb))
        for o, v, sp, off, m in ((oa, va, 0, 0, na), (ob, vb, 1, na, nb)):
            k = np.einsum("ibm,ajm,m->iajb", rho[o, v, :, sp], rho[v, o, :, sp], w, optimize=True).reshape(m, m)
            K[off:off + m, off:off + m] = 2.0 * lam ** 2 * k
        return K
    oa, vb, ob, va, na, nb = _sf_blocks(n, nO, nC, nR)
    K = np.zeros((na + nb, na + nb))
    # (alpha i -> beta a, beta j -> alpha b): rho_a(i,b) rho_b(a,j)
    K[:na, na:] = 2.0 * lam ** 2 * np.einsum("ibm,ajm,m->iajb", rho[oa, va, :, 0], rho[vb, ob, :, 1], w,
                                              optimize=True).reshape(na, nb)
    # (beta i -> alpha a, alpha j -> beta b): rho_b(i,b) rho_a(a,j)
    K[na:, :na] = 2.0 * lam ** 2 * np.einsum("ibm,ajm,m->iajb", rho[ob, vb, :, 1], rho[va, oa, :, 0], w,
                                              optimize=True).reshape(nb, na)
    return K


def ugw_phbse(eri_aa, eri_ab, eri_bb, eW, eGW, nO, eta=0.0, TDA_W=False, TDA=False, spin_conserved=True,
              spin_flip=True, exchange_kernel=False, nC=(0, 0), nR=(0, 0)):
    """GW/UGW_phBSE.f90:60-196 with the static kernel (no dynamical
    correction): spin-conserved dRPA screening at eW, its excitation
    densities, then the spin-conserved (ispin 1) and spin-flip (ispin 2) BSE
    at eGW (phULR_A/B with dRPA = false) plus the static kernels, each solved
    by phULR.  eW, eGW: n x 2.  Returns EcBSE(2) (halved with
    exchange_kernel), the spin-conserved and the spin-flip excitation energies."""
    n = eW.shape[0]
    A = phULR_A(eW, eri_aa, eri_ab, eri_bb, nO, nC=nC, nR=nR)
    B = np.zeros_like(A) if TDA_W else phULR_B(n, eri_aa, eri_ab, eri_bb, nO, nC=nC, nR=nR)
    _, OmRPA, XpY, _ = phRLR(TDA_W, A, B)
    rho = ugw_excitation_density(n, eri_aa, eri_ab, eri_bb, XpY, nO, nC, nR)
    EcBSE = np.zeros(2)
    Om = {1: None, 2: None}
    for ispin, on in ((1, spin_conserved), (2, spin_flip)):
        if not on:
            continue
        A = phULR_A(eGW, eri_aa, eri_ab, eri_bb, nO, dRPA=False, nC=nC, nR=nR, ispin=ispin)
        A = A + ugw_phbse_static_kernel_A(ispin, eta, OmRPA, rho, nO, nC, nR)
        if TDA:
            B = np.zeros_like(A)
        else:
            B = phULR_B(n, eri_aa, eri_ab, eri_bb, nO, dRPA=False, nC=nC, nR=nR, ispin=ispin)
            B = B + ugw_phbse_static_kernel_B(ispin, eta, OmRPA, rho, nO, nC, nR)
        EcBSE[ispin - 1], Om[ispin], _, _ = phRLR(TDA, A, B)
    if exchange_kernel:
        EcBSE *= 0.5
    return EcBSE, Om[1], Om[2]


def ug0w0(eri_aa, eri_ab, eri_bb, eHF, nO, eta=0.0, linearize=False, TDA_W=False, doSRG=False, nC=(0, 0),
          nR=(0, 0)):
    """GW/UG0W0.f90:92-150 (nC, nR per spin): spin-conserving dRPA screening at
    eHF, excitation densities, diagonal self-energies (GW/UGW_self_energy_diag.f90
    or the SRG form), linearised energies and the graphical QP solve of each
    spin (GW/UGW_QP_graph.f90), then the RPA again at eGW for EcRPA.  The
    unrestricted sums carry rho^2 where the restricted ones carry 2 rho^2, so
    the restricted restatements are used with rho / sqrt(2) (and EcGM, 4 rho^2
    there, halved).  Frozen entries of eGW (undefined in QuAcK's G0W0) are
    eHF.  Returns dict(eGW, SigC, Z (n x 2), Om, EcRPA, EcGM (2))."""
    n = eHF.shape[0]
    A = phULR_A(eHF, eri_aa, eri_ab, eri_bb, nO, nC=nC, nR=nR)
    B = np.zeros_like(A) if TDA_W else phULR_B(n, eri_aa, eri_ab, eri_bb, nO, nC=nC, nR=nR)
    _, Om, XpY, _ = phRLR(TDA_W, A, B)
    rho = ugw_excitation_density(n, eri_aa, eri_ab, eri_bb, XpY, nO, nC, nR) / np.sqrt(2.0)
    Sig, Z, EcGM = np.zeros((n, 2)), np.zeros((n, 2)), np.zeros(2)
    eGW = np.zeros((n, 2))
    for s in range(2):
        if doSRG:
            Sig[:, s], Z[:, s], ec = srg_self_energy_diag(eHF[:, s], Om, rho[:, :, :, s], nO[s], nC=nC[s],
                                                          nR=nR[s])
        else:
            Sig[:, s], Z[:, s], ec = self_energy_diag(eHF[:, s], Om, rho[:, :, :, s], nO[s], eta, nC=nC[s],
                                                      nR=nR[s])
        EcGM[s] = 0.5 * ec
        lin = eHF[:, s] + Z[:, s] * Sig[:, s]
        if linearize:
            eGW[:, s] = lin
        else:
            eGW[:, s], Z[:, s], _ = qp_graph(eHF[:, s], Om, rho[:, :, :, s], lin, eHF[:, s], Z[:, s], nO[s], eta,
                                             srg=doSRG, nC=nC[s], nR=nR[s])
    A = phULR_A(eGW, eri_aa, eri_ab, eri_bb, nO, nC=nC, nR=nR)
    B = np.zeros_like(A) if TDA_W else phULR_B(n, eri_aa, eri_ab, eri_bb, nO, nC=nC, nR=nR)
    EcRPA, _, _, _ = phRLR(TDA_W, A, B)
    return {"eGW": eGW, "SigC": Sig, "Z": Z, "Om": Om, "EcRPA": EcRPA, "EcGM": EcGM}


def evugw(eri_aa, eri_ab, eri_bb, eHF, nO, maxSCF=64, thresh=1e-5, max_diis=5, eta=0.0, linearize=False,
          TDA_W=False, doSRG=False, nC=(0, 0), nR=(0, 0)):
    """GW/evUGW.f90:103-215 (nC, nR per spin): eGW = eOld = eHF; per cycle the
    screening at eGW, excitation densities, self-energies at eGW, then eHF + SigC
    (linearize) or the graphical solve with guess and G from eOld; Conv =
    max |eGW - eOld| over both spins; DIIS per spin with one shared history
    length.  B is built unless TDA_W (QuAcK tests its separate TDA flag there).
    Returns dict(eGW, SigC, Z, Om, EcRPA, EcGM, nSCF, converged)."""
    n = eHF.shape[0]
    eHF = np.asarray(eHF, dtype=np.float64)
    eGW, eOld = eHF.copy(), eHF.copy()
    Z = np.ones((n, 2))
    Sig, EcGM = np.zeros((n, 2)), np.zeros(2)
    diis = [DIIS(n, max(max_diis, 1)) for _ in range(2)]
    nSCF, Conv, EcRPA, Om = 0, 1.0, 0.0, None
    while Conv > thresh and nSCF <= maxSCF:
        A = phULR_A(eGW, eri_aa, eri_ab, eri_bb, nO, nC=nC, nR=nR)
        B = np.zeros_like(A) if TDA_W else phULR_B(n, eri_aa, eri_ab, eri_bb, nO, nC=nC, nR=nR)
        EcRPA, Om, XpY, _ = phRLR(TDA_W, A, B)
        rho = ugw_excitation_density(n, eri_aa, eri_ab, eri_bb, XpY, nO, nC, nR) / np.sqrt(2.0)
        new = np.zeros((n, 2))
        for s in range(2):
            if doSRG:
                Sig[:, s], Z[:, s], ec = srg_self_energy_diag(eGW[:, s], Om, rho[:, :, :, s], nO[s], nC=nC[s],
                                                              nR=nR[s])
            else:
                Sig[:, s], Z[:, s], ec = self_energy_diag(eGW[:, s], Om, rho[:, :, :, s], nO[s], eta, nC=nC[s],
                                                          nR=nR[s])
            EcGM[s] = 0.5 * ec
            if linearize:
                new[:, s] = eHF[:, s] + Sig[:, s]
            else:
                new[:, s], Z[:, s], _ = qp_graph(eHF[:, s], Om, rho[:, :, :, s], eOld[:, s], eOld[:, s], Z[:, s],
                                                 nO[s], eta, srg=doSRG, nC=nC[s], nR=nR[s])
        eGW = new
        Conv = np.abs(eGW - eOld).max()
        if max_diis > 1:
            for s in range(2):
                eGW[:, s], _ = diis[s](eGW[:, s] - eOld[:, s], eGW[:, s].copy())
        eOld = eGW.copy()
        nSCF += 1
    return {"eGW": eGW, "SigC": Sig, "Z": Z, "Om": Om, "EcRPA": EcRPA, "EcGM": EcGM, "nSCF": nSCF,
            "converged": nSCF <= maxSCF}
