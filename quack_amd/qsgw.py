"""Quasiparticle self-consistent GW (qsGW@RHF) on the device.

QuAcK's qsRGW loop (GW/qsRGW.f90:193-352, without linDM):
per cycle
  * J and K from the AO integrals resident in HBM (qk_rhf_jk),
  * the MO integral block of the current orbitals (the engine's AO->MO block)
    and the dRPA screening at the current quasiparticle energies with its
    excitation densities (qk_gw_screen),
  * the full correlation self-energy at those energies (qk_gw_sigma_full:
    GW/RGW_self_energy.f90 or, with SRG, GW/RGW_SRG_self_energy.f90),
    symmetrised and taken to the AO basis (S C Sigma C^T S, AOtoMO/MOtoAO.f90),
  * F = Hc + J + K/2 + Sigma_AO, the commutator FPS - SPF (its largest element
    is the convergence measure once the cycle counter, which starts at -1,
    exceeds 1), the energy terms, DIIS on F (utils/DIIS_extrapolation.f90),
  * diagonalisation of X^T F X for the new quasiparticle energies and
    orbitals, and the new density P = 2 C_occ C_occ^T.
The N^4 and N^5 steps run on the GPU; the N^3 steps on the host.  QuAcK stops
when maxSCF cycles pass; here that raises QsGWNotConverged with the results.
"""
import numpy as np

from . import gw, scf


class QsGWNotConverged(RuntimeError):
    def __init__(self, msg, result):
        super().__init__(msg)
        self.result = result


def qsRGW(eri_ptr, S, T, V, X, cHF, eHF, nO, maxSCF=64, thresh=1e-5, max_diis=5, eta=0.0, doSRG=False,
          TDA_W=False, log=None):
    """qsGW from the RHF orbitals cHF (nBas x nOrb) and energies eHF.  eri_ptr:
    device address of the nBas^4 AO integrals (QuAcK's ERI.bin order).  X: the
    orthogonalisation matrix (nBas x nOrb).  Returns dict(eGW, c, EqsGW
    (electronic), EcRPA, EcGM, Z, SigC (MO basis, symmetrised), nSCF, Conv)."""
    import torch  # device memory for the MO coefficients handed to the AO->MO block

    nBas, n = X.shape  # n = nOrb
    Hc = T + V
    c = np.array(cHF, dtype=np.float64, order="F")
    eGW = np.array(eHF, dtype=np.float64)
    P = 2.0 * c[:, :nO] @ c[:, :nO].T
    nS = nO * (n - nO)
    err_diis = np.zeros((nBas * nBas, max(1, max_diis)))
    F_diis = np.zeros((nBas * nBas, max(1, max_diis)))
    n_diis, nSCF, Conv = 0, -1, 1.0
    res = {}
    c_t = torch.empty(nBas * n, dtype=torch.float64, device="cuda")
    with scf.JK(nBas) as jk, gw.GWEngine(nBas, n, nO) as eng:
        while Conv > thresh and nSCF < maxSCF:
            nSCF += 1
            J, K = jk(eri_ptr, P)
            c_t.copy_(torch.from_numpy(np.asfortranarray(c).ravel(order="F").copy()))
            eng.set_eri_ao_device(eri_ptr, c_t.data_ptr())
            eng.screen(eGW, TDA_W=TDA_W)
            r = eng.sigma_full(eGW, eta=eta, doSRG=doSRG)
            SigC = 0.5 * (r["SigC"] + r["SigC"].T)
            SigCp = S @ c @ SigC @ c.T @ S
            F = Hc + J + 0.5 * K + SigCp
            if nBas != n:  # project F onto the kept orbitals (qsRGW.f90:281-284)
                F = S @ c @ (c.T @ F @ c) @ c.T @ S
            err = F @ P @ S - S @ P @ F
            if nSCF > 1:
                Conv = float(np.abs(err).max())
            EqsGW = float(np.sum(P * T.T) + np.sum(P * V.T) + 0.5 * np.sum(P * J.T) + 0.25 * np.sum(P * K.T))
            if max_diis > 1:
                n_diis = min(n_diis + 1, max_diis)
                f = np.asfortranarray(F).ravel(order="F").copy()
                scf.diis_extrapolation(n_diis, err_diis, F_diis, np.asfortranarray(err).ravel(order="F"), f)
                F = f.reshape(nBas, nBas, order="F")
            if nBas == n:
                eGW, cp = np.linalg.eigh(X.T @ F @ X)
                c = np.asfortranarray(X @ cp)
            else:  # qsRGW.f90:323-333: rotate the current orbitals
                eGW, cp = np.linalg.eigh(c.T @ F @ c)
                c = np.asfortranarray(c @ cp)
            P = 2.0 * c[:, :nO] @ c[:, :nO].T
            res = {"eGW": eGW, "c": c, "EqsGW": EqsGW, "EcRPA": r["EcRPA"], "EcGM": r["EcGM"], "Z": r["Z"],
                   "SigC": c.T @ SigCp @ c, "nSCF": nSCF, "Conv": Conv, "nS": nS}
            if log:
                log("qsGW %3d  E %.10f  Conv %.2e" % (nSCF, EqsGW, Conv))
    if nSCF == maxSCF:  # GW/qsRGW.f90:354-362
        raise QsGWNotConverged("qsGW: convergence failed after %d cycles (Conv %.2e)" % (nSCF, Conv), res)
    return res
