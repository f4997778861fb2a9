"""Quasiparticle self-consistent GW (qsGW@RHF) on the device.

QuAcK's qsRGW (GW/qsRGW.f90:1-431, without linDM): the loop and its stopping
rule stay here, as in QuAcK; every pass of the loop body is one call of the
library (qk_qsgw_cycle):
  * J and K from the AO integrals resident in HBM,
  * the MO integral block of the current orbitals and the dRPA screening at
    the current quasiparticle energies with its excitation densities,
  * the full correlation self-energy at those energies (GW/RGW_self_energy.f90
    or, with SRG, GW/RGW_SRG_self_energy.f90), symmetrised and taken to the AO
    basis (S C Sigma C^T S, AOtoMO/MOtoAO.f90),
  * F = Hc + J + K/2 + Sigma_AO, the commutator FPS - SPF, the energy terms,
    DIIS on F (utils/DIIS_extrapolation.f90),
  * diagonalisation of X^T F X (or C^T F C when functions were dropped) for
    the new quasiparticle energies and orbitals, and the new density.
QuAcK stops when maxSCF cycles pass; here that raises QsGWNotConverged with
the results.
"""
import ctypes

import numpy as np

from . import gw


class QsGWNotConverged(RuntimeError):
    def __init__(self, msg, result):
        super().__init__(msg)
        self.result = result


class QsGW:
    """One qsGW calculation (a qk_qsgw_handle): inputs as GW/qsRGW.f90:1-4."""

    def __init__(self, S, T, V, X, eri_ao, cHF, eHF, nO, max_diis=5, nC=0, nR=0, Hc=None, PHF=None):
        lib = gw.load()
        X = gw._f64(X)
        self.nBas, self.nOrb = X.shape
        self.nC, self.nO, self.nV, self.nR, self.nS = gw.dims(self.nOrb, nO, nC, nR)
        cHF = gw._f64(cHF)
        if Hc is None:
            Hc = np.asarray(T) + np.asarray(V)
        if PHF is None:
            PHF = 2.0 * cHF[:, :nO] @ cHF[:, :nO].T
        self._keep = [gw._f64(a) for a in (S, T, V, Hc, X, PHF, cHF)] + [gw._f64(eHF)]
        eri = np.ascontiguousarray(eri_ao, dtype=np.float64).reshape(-1)
        assert eri.size == self.nBas ** 4, "eri_ao: nBas^4 values in QuAcK's ERI.bin order"
        self.h = gw._H()
        S_, T_, V_, Hc_, X_, P_, c_, e_ = self._keep
        rc = lib.qk_qsgw_create(ctypes.byref(self.h), self.nBas, self.nOrb, self.nC, self.nO, self.nV, self.nR,
                                self.nS, int(max_diis), gw._p(S_), gw._p(T_), gw._p(V_), gw._p(Hc_), gw._p(X_),
                                gw._p(eri), gw._p(P_), gw._p(c_), gw._p(e_))
        if rc != gw.QK_OK:
            raise gw.QKError(rc, lib.qk_last_error().decode())

    def cycle(self, eta=0.0, TDA_W=False, doSRG=False):
        """One pass of the loop body (qk_qsgw_cycle); returns its outputs."""
        n, N = self.nOrb, self.nBas
        eGW, Z = np.zeros(n), np.zeros(n)
        c = np.zeros((N, n), order="F")
        SigC = np.zeros((n, n), order="F")
        E = np.zeros(5)
        Ec, Egm, err = (np.zeros(1) for _ in range(3))
        rc = gw.load().qk_qsgw_cycle(self.h, eta, int(TDA_W), int(doSRG), gw._p(eGW), gw._p(c), gw._p(SigC),
                                     gw._p(Z), gw._p(E), gw._p(Ec), gw._p(Egm), gw._p(err))
        if rc != gw.QK_OK:
            raise gw.QKError(rc, gw.load().qk_qsgw_last_error(self.h).decode())
        return {"eGW": eGW, "c": c, "SigC": SigC, "Z": Z, "ET": E[0], "EV": E[1], "EJ": E[2], "EK": E[3],
                "EqsGW": E[4], "EcRPA": Ec[0], "EcGM": Egm[0], "err_max": err[0]}

    def close(self):
        if self.h:
            gw.load().qk_qsgw_destroy(self.h)
            self.h = gw._H()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def qsRGW(eri_ao, S, T, V, X, cHF, eHF, nO, maxSCF=64, thresh=1e-5, max_diis=5, eta=0.0, doSRG=False,
          TDA_W=False, nC=0, nR=0, log=None):
    """qsGW from the RHF orbitals cHF (nBas x nOrb) and energies eHF, with
    QuAcK's loop (GW/qsRGW.f90:193-352): the cycle counter starts at -1 and
    the commutator becomes the convergence measure once it exceeds 1.
    eri_ao: the nBas^4 AO integrals (QuAcK's ERI.bin order).  X: the
    orthogonalisation matrix (nBas x nOrb).  Returns dict(eGW, c, EqsGW
    (electronic), ET, EV, EJ, EK, EcRPA, EcGM, Z, SigC (MO basis), nSCF,
    Conv)."""
    nSCF, Conv = -1, 1.0
    res = {}
    with QsGW(S, T, V, X, eri_ao, cHF, eHF, nO, max_diis=max_diis, nC=nC, nR=nR) as q:
        while Conv > thresh and nSCF < maxSCF:
            nSCF += 1
            res = q.cycle(eta=eta, TDA_W=TDA_W, doSRG=doSRG)
            if nSCF > 1:
                Conv = res["err_max"]
            res.update(nSCF=nSCF, Conv=Conv, nS=q.nS)
            if log:
                log("qsGW %3d  E %.10f  Conv %.2e" % (nSCF, res["EqsGW"], Conv))
    if nSCF == maxSCF:  # GW/qsRGW.f90:354-362
        raise QsGWNotConverged("qsGW: convergence failed after %d cycles (Conv %.2e)" % (nSCF, Conv), res)
    return res
