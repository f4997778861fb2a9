"""Restricted Hartree-Fock SCF with its Fock builds on the device.

QuAcK's RHF loop (HF/RHF.f90:100-249): core guess (HF/mo_guess.f90:31-34,
HF/core_guess.f90), then per cycle J and K (the N^4 step, on the GPU through
qk_rhf_jk: one pass over the AO integrals resident in HBM), F = Hc + J + K/2,
the commutator FPS - SPF and its largest element as the convergence measure
(from the third cycle on), the energy terms, DIIS on F
(utils/DIIS_extrapolation.f90), diagonalisation of X^T F X and the new density
P = 2 C_occ C_occ^T.  The O(N^3) steps stay on the host (N <= 512 here).
QuAcK stops when maxSCF cycles pass without convergence; here that raises
RHFNotConverged carrying the last cycle's results.
"""
import ctypes

import numpy as np

from . import gw


class RHFNotConverged(RuntimeError):
    def __init__(self, msg, result):
        super().__init__(msg)
        self.result = result


class JK:
    """Coulomb and exchange matrices of QuAcK's Hartree_matrix_AO_basis and
    exchange_matrix_AO_basis (AOtoMO/*.f90:22-32) from AO integrals resident
    on the device (a pointer to nBas^4 doubles in QuAcK's ERI.bin order)."""

    def __init__(self, nBas):
        lib = gw.load()
        self.nBas = nBas
        self.h = gw._H()
        rc = lib.qk_rhf_create(ctypes.byref(self.h), nBas)
        if rc != gw.QK_OK:
            raise gw.QKError(rc, "qk_rhf_create(nBas=%d)" % nBas)

    def __call__(self, eri_ptr, P):
        n = self.nBas
        P = np.asfortranarray(P, dtype=np.float64)
        if P.shape != (n, n):
            raise ValueError("P must be %d x %d" % (n, n))
        J = np.zeros((n, n), order="F")
        K = np.zeros((n, n), order="F")
        lib = gw.load()
        rc = lib.qk_rhf_jk(self.h, ctypes.c_void_p(eri_ptr), n ** 4, gw._p(P), gw._p(J), gw._p(K))
        if rc != gw.QK_OK:
            raise gw.QKError(rc, lib.qk_rhf_last_error(self.h).decode())
        return J, K

    def last_ms(self):
        ms = ctypes.c_double(0.0)
        gw.load().qk_rhf_jk_ms(self.h, ctypes.byref(ms))
        return ms.value

    def close(self):
        if self.h:
            gw.load().qk_rhf_destroy(self.h)
            self.h = gw._H()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def orthogonalization_matrix(S, thresh=1e-6):
    """Canonical orthogonalisation X = U s^-1/2 over the eigenvalues of S above
    thresh (utils/orthogonalization_matrix.f90, QuAcK's default type)."""
    s, U = np.linalg.eigh(S)
    keep = s > thresh
    return U[:, keep] / np.sqrt(s[keep])


def _prepend(hist, v, n_diis):
    """utils/utils.f90:506-536: shift columns 1..n_diis-1 right, v in column 1."""
    hist[:, 1:n_diis] = hist[:, 0:n_diis - 1].copy()
    hist[:, 0] = v


def diis_extrapolation(n_diis, err_hist, e_hist, err_in, e_inout):
    """utils/DIIS_extrapolation.f90: prepend the new error and value, solve the
    bordered system [E^T E, -1; -1, 0] w = (0, -1) and replace e_inout by the
    extrapolation when rcond (1-norm) exceeds 1e-13.  Returns rcond."""
    _prepend(err_hist, err_in, n_diis)
    _prepend(e_hist, e_inout, n_diis)
    E = err_hist[:, :n_diis]
    A = np.empty((n_diis + 1, n_diis + 1))
    A[:n_diis, :n_diis] = E.T @ E
    A[:n_diis, n_diis] = -1.0
    A[n_diis, :n_diis] = -1.0
    A[n_diis, n_diis] = 0.0
    b = np.zeros(n_diis + 1)
    b[n_diis] = -1.0
    try:
        Ainv = np.linalg.inv(A)
        rcond = 1.0 / (np.abs(A).sum(axis=0).max() * np.abs(Ainv).sum(axis=0).max())
        w = Ainv @ b
    except np.linalg.LinAlgError:
        rcond, w = 0.0, None
    if rcond > 1e-13:
        e_inout[:] = e_hist[:, :n_diis] @ w[:n_diis]
    return rcond


def RHF(eri_ptr, nO, S, T, V, X=None, maxSCF=256, thresh=1e-7, max_diis=5, guess_type=1, level_shift=0.0,
        mix=0.0, log=None):
    """QuAcK's RHF (HF/RHF.f90:100-249) for the AO integrals at device address
    eri_ptr (nBas^4 doubles, QuAcK's ERI.bin order).  Defaults are those of
    input/options.default.  Returns dict(eHF, c, P, ERHF (electronic), nSCF,
    Conv, ET, EV, EJ, EK, jk_ms)."""
    if guess_type != 1:
        raise NotImplementedError("only the core guess (guess_type = 1) is supported")
    if level_shift > 0.0 or mix > 0.0:
        raise NotImplementedError("level shifting and guess mixing are not supported")
    nBas = S.shape[0]
    if X is None:
        X = orthogonalization_matrix(S)
    Hc = T + V
    jk = JK(nBas)
    try:
        # core guess: diagonalise X^T Hc X
        eHF, cp = np.linalg.eigh(X.T @ Hc @ X)
        c = X @ cp
        P = 2.0 * c[:, :nO] @ c[:, :nO].T
        n2 = nBas * nBas
        err_diis = np.zeros((n2, max(1, max_diis)))
        F_diis = np.zeros((n2, max(1, max_diis)))
        n_diis, Conv, nSCF = 0, 1.0, 0
        jk_ms = []
        ERHF = ET = EV = EJ = EK = 0.0
        while Conv > thresh and nSCF < maxSCF:
            nSCF += 1
            J, K = jk(eri_ptr, P)
            jk_ms.append(jk.last_ms())
            F = Hc + J + 0.5 * K
            err = F @ P @ S - S @ P @ F
            if nSCF > 2:
                Conv = float(np.abs(err).max())
            ET = float(np.sum(P * T.T))          # trace(P T)
            EV = float(np.sum(P * V.T))
            EJ = 0.5 * float(np.sum(P * J.T))
            EK = 0.25 * float(np.sum(P * K.T))
            ERHF = ET + EV + EJ + EK
            if max_diis > 1:
                n_diis = min(n_diis + 1, max_diis)
                f = np.asfortranarray(F).ravel(order="F").copy()
                diis_extrapolation(n_diis, err_diis, F_diis, np.asfortranarray(err).ravel(order="F"), f)
                F = f.reshape(nBas, nBas, order="F")
            eHF, cp = np.linalg.eigh(X.T @ F @ X)
            c = X @ cp
            P = 2.0 * c[:, :nO] @ c[:, :nO].T
            if log:
                log("RHF %3d  E %.10f  Conv %.2e" % (nSCF, ERHF, Conv))
        res = {"eHF": eHF, "c": np.asfortranarray(c), "P": P, "ERHF": ERHF, "nSCF": nSCF, "Conv": Conv, "ET": ET,
               "EV": EV, "EJ": EJ, "EK": EK, "nOrb": X.shape[1], "jk_ms": jk_ms}
    finally:
        jk.close()
    if nSCF == maxSCF:  # HF/RHF.f90:253: reaching maxSCF counts as failure
        raise RHFNotConverged("RHF: convergence failed after %d cycles (Conv %.2e)" % (nSCF, Conv), res)
    return res


def jk_closure(nBas, eri_ptr):
    """Fock-build closure with the calling convention of the input producer's
    rhf() (quack_amd.intgen.rhf: jk(eri, D) -> (J, K) with F = H + 2J - K for
    the half density D = C_occ C_occ^T): J = J_QuAcK(D), K = -K_QuAcK(D)."""
    eng = JK(nBas)

    def jk(eri_unused, D):
        J, K = eng(eri_ptr, D)
        return J, -K

    jk.engine = eng
    return jk
