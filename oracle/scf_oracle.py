"""TEST INFRASTRUCTURE ONLY: CPU restatement of QuAcK's restricted
Hartree-Fock loop and its Fock builds.

The checker for qk_rhf_jk and quack_amd.scf.RHF.  Only tests/ import it; the
product library never does.  Conventions as in oracle/gw_oracle.py: the AO
integrals are QuAcK's flat physicist array, ERI(mu,la,nu,si) = <mu la|nu si>
= (mu nu|la si) at mu + n la + n^2 nu + n^3 si, viewed as
G = eri.reshape((n,)*4, order='F').

Pinning (tests/test_scf_oracle.py):
* hartree_matrix / exchange_matrix against literal transcriptions of the
  Fortran loops (hartree_matrix_loops / exchange_matrix_loops) on random
  integrals with the 8-fold symmetry, and against the input producer's
  independently written Fock build (quack_amd.intgen.jk_numpy, another index
  route: J = J_QuAcK(D), K = -K_QuAcK(D) for the half density D);
* the single-orbital identity: for P = 2 phi phi^T,
  phi^T J phi = 2 (phi phi|phi phi) = -phi^T K phi;
* rhf against QuAcK's RHF energies and HOMO/LUMO of He/6-31G and
  H2O/cc-pVDZ (FeatherBench, tests/golden/quack_goldens.json).
"""
import numpy as np

from .gw_oracle import DIIS, G4, diagonalize_matrix


def hartree_matrix_loops(P, eri, n):
    """AOtoMO/Hartree_matrix_AO_basis.f90:22-32, the four loops as written."""
    G = G4(eri, n)
    H = np.zeros((n, n))
    for si in range(n):
        for nu in range(n):
            for la in range(n):
                for mu in range(n):
                    H[mu, nu] += P[la, si] * G[mu, la, nu, si]
    return H


def exchange_matrix_loops(P, eri, n):
    """AOtoMO/exchange_matrix_AO_basis.f90:22-32, the four loops as written."""
    G = G4(eri, n)
    K = np.zeros((n, n))
    for nu in range(n):
        for si in range(n):
            for la in range(n):
                for mu in range(n):
                    K[mu, nu] -= P[la, si] * G[mu, la, si, nu]
    return K


def hartree_matrix(P, eri, n):
    """J(mu,nu) = sum_{la,si} P(la,si) ERI(mu,la,nu,si)."""
    return np.einsum("ls,mlns->mn", P, G4(eri, n), optimize=True)


def exchange_matrix(P, eri, n):
    """K(mu,nu) = -sum_{la,si} P(la,si) ERI(mu,la,si,nu)."""
    return -np.einsum("ls,mlsn->mn", P, G4(eri, n), optimize=True)


def rhf(S, T, V, eri, nO, X, maxSCF=256, thresh=1e-7, max_diis=5):
    """HF/RHF.f90:100-249 with the core guess (HF/core_guess.f90), no level
    shift, no mixing.  Returns dict(eHF, c, P, ERHF, nSCF, Conv)."""
    n = S.shape[0]
    Hc = T + V
    eHF, cp = diagonalize_matrix(X.T @ Hc @ X)
    c = X @ cp
    P = 2.0 * c[:, :nO] @ c[:, :nO].T
    diis = DIIS(n * n, max_diis)
    Conv, nSCF, ERHF = 1.0, 0, 0.0
    while Conv > thresh and nSCF < maxSCF:
        nSCF += 1
        J = hartree_matrix(P, eri, n)
        K = exchange_matrix(P, eri, n)
        F = Hc + J + 0.5 * K
        err = F @ P @ S - S @ P @ F
        if nSCF > 2:
            Conv = np.abs(err).max()
        ERHF = np.trace(P @ T) + np.trace(P @ V) + 0.5 * np.trace(P @ J) + 0.25 * np.trace(P @ K)
        if max_diis > 1:
            f, _ = diis(err.ravel(order="F"), F.ravel(order="F").copy())
            F = f.reshape(n, n, order="F")
        eHF, cp = diagonalize_matrix(X.T @ F @ X)
        c = X @ cp
        P = 2.0 * c[:, :nO] @ c[:, :nO].T
    return {"eHF": eHF, "c": c, "P": P, "ERHF": ERHF, "nSCF": nSCF, "Conv": Conv, "converged": nSCF < maxSCF}
