"""CPU restatement of QuAcK's RHF loop and Fock builds (oracle/scf_oracle.py)
pinned by literal transcriptions of the Fortran loops, the input producer's
independently written Fock build, a closed-form identity and QuAcK's RHF
energies of He/6-31G and H2O/cc-pVDZ."""
import json
import os

import numpy as np
import pytest

from oracle import scf_oracle as so
from quack_amd import intgen as ig

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "quack_goldens.json")))


def random_eri(n, seed):
    """Flat physicist array of random integrals with the 8-fold symmetry of
    real orbitals: (ab|cd) = (ba|cd) = (ab|dc) = (cd|ab)."""
    rng = np.random.default_rng(seed)
    c = rng.standard_normal((n, n, n, n))
    c = c + c.transpose(1, 0, 2, 3)
    c = c + c.transpose(0, 1, 3, 2)
    c = c + c.transpose(2, 3, 0, 1)        # chem[a, b, c, d] = (ab|cd)
    G = c.transpose(0, 2, 1, 3)            # G[a, c, b, d] = <ac|bd> = (ab|cd)
    return np.asfortranarray(G).ravel(order="F")


@pytest.mark.parametrize("n", [1, 3, 5])
def test_fock_builds_match_fortran_loops(n):
    eri = random_eri(n, seed=n)
    rng = np.random.default_rng(10 + n)
    P = rng.standard_normal((n, n))
    P = P + P.T
    assert np.abs(so.hartree_matrix(P, eri, n) - so.hartree_matrix_loops(P, eri, n)).max() < 1e-12
    assert np.abs(so.exchange_matrix(P, eri, n) - so.exchange_matrix_loops(P, eri, n)).max() < 1e-12
    if n > 1:  # a plausible slip, the exchange indices in the Coulomb build, is caught
        assert np.abs(so.exchange_matrix(P, eri, n) + so.hartree_matrix(P, eri, n)).max() > 1e-3


def test_fock_builds_match_producer():
    """J(D) and K(D) against quack_amd.intgen.jk_numpy, which forms the same
    matrices from a C-order view of the array (J = J(D), K = -K(D))."""
    mol = ig.h2o_feather()
    eri = mol.eri()
    n = mol.nBas
    rng = np.random.default_rng(3)
    C = rng.standard_normal((n, 5))
    D = C @ C.T
    Ji, Ki = ig.jk_numpy(eri, D)
    assert np.abs(so.hartree_matrix(D, eri, n) - Ji).max() < 1e-12
    assert np.abs(so.exchange_matrix(D, eri, n) + Ki).max() < 1e-12


def test_single_orbital_identity():
    """P = 2 phi phi^T: phi^T J phi = 2 (phi phi|phi phi) = -phi^T K phi."""
    mol = ig.h2o_feather()
    eri = mol.eri()
    n = mol.nBas
    phi = np.random.default_rng(5).standard_normal(n)
    P = 2.0 * np.outer(phi, phi)
    G = eri.reshape((n,) * 4, order="F")           # <ab|cd> = (ac|bd)
    pppp = np.einsum("a,b,c,d,abcd->", phi, phi, phi, phi, G)
    assert abs(phi @ so.hartree_matrix(P, eri, n) @ phi - 2 * pppp) < 1e-10 * abs(pppp)
    assert abs(phi @ so.exchange_matrix(P, eri, n) @ phi + 2 * pppp) < 1e-10 * abs(pppp)


@pytest.mark.parametrize("name", ["He", "H2O"])
def test_rhf_anchors(name):
    """QuAcK's defaults (input/options.default: maxSCF 256, thresh 1e-7,
    DIIS 5, core guess) reproduce QuAcK's RHF energy and frontier orbitals."""
    mol = ig.he_631g() if name == "He" else ig.h2o_feather()
    S, T, V = mol.one_electron()
    X = np.linalg.cholesky(np.linalg.inv(S))  # S^-1 = L L^T: X = L gives X^T S X = 1
    assert np.abs(X.T @ S @ X - np.eye(mol.nBas)).max() < 1e-10
    r = so.rhf(S, T, V, mol.eri(), mol.nO, X)
    g = GOLD[name]["feather"]
    nO = mol.nO
    assert r["converged"]
    # same guess, DIIS and stopping rule as QuAcK: the energy agrees to ~1e-13,
    # the orbital energies to the 1e-7 commutator threshold's ~1e-8
    assert abs(r["ERHF"] - g["RHF energy"]) < 1e-11
    assert abs(r["eHF"][nO - 1] - g["RHF HOMO energy"]) < 5e-8
    assert abs(r["eHF"][nO] - g["RHF LUMO energy"]) < 5e-8
