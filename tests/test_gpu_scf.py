"""HIP Fock builds (qk_rhf_jk) and the device RHF loop (quack_amd.scf.RHF)
against the CPU restatement of QuAcK's Hartree_matrix_AO_basis,
exchange_matrix_AO_basis and RHF (oracle/scf_oracle.py) and QuAcK's RHF
energies (He/6-31G, H2O/cc-pVDZ)."""
import ctypes
import json
import os
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "quack_goldens.json")))


def say(*a):
    print(*a, flush=True)


def device_eri(eri):
    import torch
    return torch.from_numpy(np.ascontiguousarray(eri)).to("cuda")


@pytest.mark.parametrize("name", ["He", "H2O", "H10"])
def test_jk_matches_oracle(systems, name):
    from oracle import scf_oracle as so
    from quack_amd import scf
    say("jk", name)
    s = systems[name]
    n = s["nBas"]
    eri_t = device_eri(s["eri"])
    rng = np.random.default_rng(7)
    A = rng.standard_normal((n, n))
    Ps = [2.0 * s["c"][:, :s["nO"]] @ s["c"][:, :s["nO"]].T, A + A.T]
    with scf.JK(n) as jk:
        for P in Ps:
            J, K = jk(eri_t.data_ptr(), P)
            Jr, Kr = so.hartree_matrix(P, s["eri"], n), so.exchange_matrix(P, s["eri"], n)
            scale = max(1.0, np.abs(Jr).max())
            assert np.abs(J - Jr).max() < 1e-13 * scale * n, name
            assert np.abs(K - Kr).max() < 1e-13 * scale * n, name
            # fixed summation order: the same bits every call
            J2, K2 = jk(eri_t.data_ptr(), P)
            assert np.array_equal(J, J2) and np.array_equal(K, K2)
        # host integrals copied into the handle (the Fortran caller's route)
        from quack_amd import gw
        lib = gw.load()
        assert lib.qk_rhf_set_eri_ao(jk.h, gw._p(np.ascontiguousarray(s["eri"]))) == gw.QK_OK
        P = Ps[0]
        J3, K3 = np.zeros((n, n), order="F"), np.zeros((n, n), order="F")
        assert lib.qk_rhf_jk(jk.h, None, n ** 4, gw._p(np.asfortranarray(P)), gw._p(J3), gw._p(K3)) == gw.QK_OK
        J, K = jk(eri_t.data_ptr(), P)
        assert np.array_equal(J, J3) and np.array_equal(K, K3)


def test_jk_rejects_wrong_size(systems):
    from quack_amd import gw, scf
    s = systems["He"]
    n = s["nBas"]
    eri_t = device_eri(s["eri"])
    with scf.JK(n) as jk:
        J = np.zeros((n, n), order="F")
        rc = gw.load().qk_rhf_jk(jk.h, ctypes.c_void_p(eri_t.data_ptr()), n ** 4 - 1, gw._p(np.eye(n)),
                                 gw._p(J), gw._p(J.copy()))
        assert rc == gw.QK_E_ARG


@pytest.mark.parametrize("name", ["He", "H2O"])
def test_rhf_anchors_on_device(systems, name):
    """QuAcK's defaults reproduce QuAcK's RHF energy and frontier orbitals, and
    follow the CPU restatement cycle for cycle."""
    from oracle import scf_oracle as so
    from quack_amd import intgen as ig
    from quack_amd import scf
    say("rhf", name)
    mol = ig.he_631g() if name == "He" else ig.h2o_feather()
    S, T, V = mol.one_electron()
    eri = mol.eri()
    eri_t = device_eri(eri)
    X = np.linalg.cholesky(np.linalg.inv(S))
    r = scf.RHF(eri_t.data_ptr(), mol.nO, S, T, V, X)
    ro = so.rhf(S, T, V, eri, mol.nO, X)
    g = GOLD[name]["feather"]
    nO = mol.nO
    assert r["nSCF"] == ro["nSCF"]
    assert abs(r["ERHF"] - ro["ERHF"]) < 1e-12
    assert np.abs(r["eHF"] - ro["eHF"]).max() < 1e-10
    assert abs(r["ERHF"] - g["RHF energy"]) < 1e-11
    assert abs(r["eHF"][nO - 1] - g["RHF HOMO energy"]) < 5e-8
    assert abs(r["eHF"][nO] - g["RHF LUMO energy"]) < 5e-8


def test_rhf_h10_and_lattice_timing():
    """H10 chain (N = 50): the device loop against the restatement; then one
    Fock build of the H50 lattice (N = 250, 31 GB of AO integrals) timed
    against the HBM roofline."""
    import torch

    from oracle import scf_oracle as so
    from quack_amd import intgen as ig
    from quack_amd import scf
    say("rhf H10")
    mol = ig.h_chain(10, seed=0)
    S, T, V = mol.one_electron()
    eri = mol.eri()
    eri_t = device_eri(eri)
    X = scf.orthogonalization_matrix(S)
    r = scf.RHF(eri_t.data_ptr(), mol.nO, S, T, V, X)
    ro = so.rhf(S, T, V, eri, mol.nO, X)
    assert r["nSCF"] == ro["nSCF"]
    assert abs(r["ERHF"] - ro["ERHF"]) < 1e-11
    assert np.abs(r["eHF"] - ro["eHF"]).max() < 1e-9
    del eri_t
    say("jk H50 lattice")
    lat = ig.h_lattice(5, 10, seed=0)
    n = lat.nBas
    big = torch.empty(n ** 4, dtype=torch.float64, device="cuda")
    lat.eri_device(big.data_ptr(), big.numel())
    torch.cuda.synchronize()
    P = np.eye(n)
    with scf.JK(n) as jk:
        jk(big.data_ptr(), P)
        t = time.time()
        for _ in range(5):
            J, K = jk(big.data_ptr(), P)
        wall = (time.time() - t) / 5
        ms = jk.last_ms()
    # P = 1: J(mu,nu) = sum_la ERI(mu,la,nu,la), K(mu,nu) = -sum_la ERI(mu,la,la,nu);
    # the N^4 array stays on the device (C-order view [si, nu, la, mu]): one column of each
    col, ar = 7, torch.arange(n, device="cuda")
    V4 = big.view(n, n, n, n)
    Jref = V4[ar, col, ar, :].sum(dim=0).cpu().numpy()   # sum over la of ERI(mu, la, col, la)
    Kref = -V4[col, ar, ar, :].sum(dim=0).cpu().numpy()  # -sum over la of ERI(mu, la, la, col)
    assert np.abs(J[:, col] - Jref).max() < 1e-11 * n
    assert np.abs(K[:, col] - Kref).max() < 1e-11 * n
    gbs = n ** 4 * 8 / (ms * 1e-3) / 1e9
    say("H50 lattice Fock build: %.2f ms device (%.0f GB/s of the 8000 GB/s HBM peak), %.2f ms per call" %
        (ms, gbs, wall * 1e3))
    assert ms > 0
