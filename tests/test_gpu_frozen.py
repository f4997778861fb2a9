"""GPU: frozen core (nC) and frozen virtuals (nR) through the C-ABI against
the oracle (oracle/gw_oracle.py, pinned to QuAcK's compiled routines with
the same nC, nR).  Frozen orbitals drop out of every sum; the frozen entries
of eGW keep the energies in G (eHF in G0W0, the previous cycle in evGW, as
GW/RGW_QP_graph.f90 never assigns them), so evGW's DIIS error is zero there
and its cycle count is QuAcK's.  Bar: 1e-8 Ha on eGW, Om and EcRPA."""
import sys

import numpy as np
import pytest

from oracle import gw_oracle as go

pytestmark = pytest.mark.gpu
CASES = [(1, 0), (1, 2), (0, 3)]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def gw():
    from quack_amd import gw as m
    m.load()
    return m


@pytest.mark.parametrize("nC,nR", CASES)
def test_frozen_g0w0_evgw(gw, mo_systems, nC, nR):
    s = mo_systems["H2O"]
    nO, e, mo = s["nO"], s["eHF"], s["eri_mo"]
    for kw in ({"eta": 0.0}, {"eta": 0.1}, {"doSRG": True}, {"linearize": True}):
        say("frozen g0w0", nC, nR, kw)
        r = gw.RG0W0(mo, e, nO, nC=nC, nR=nR, **kw)
        o = go.rg0w0(mo, e, nO, nC=nC, nR=nR, **kw)
        assert np.abs(r["eGW"] - o["eGW"]).max() < 1e-8, kw
        assert np.abs(r["Om"] - o["Om"]).max() < 1e-8 and abs(r["EcRPA"] - o["EcRPA"]) < 1e-8
    for kw in ({"eta": 0.1, "max_diis": 5}, {"doSRG": True, "max_diis": 5}, {"eta": 0.1, "max_diis": 1}):
        say("frozen evgw", nC, nR, kw)
        r = gw.evRGW(mo, e, nO, thresh=1e-6, nC=nC, nR=nR, **kw)
        o = go.evrgw(mo, e, nO, thresh=1e-6, nC=nC, nR=nR, **kw)
        assert r["nSCF"] == o["nSCF"], kw
        assert np.abs(r["eGW"] - o["eGW"]).max() < 1e-8, kw
        frozen = np.r_[np.arange(nC), np.arange(len(e) - nR, len(e))]
        assert np.abs(r["eGW"][frozen] - e[frozen]).max(initial=0) < 1e-12


def test_frozen_qp_graph_keeps_eold(gw, mo_systems):
    """qk_rgw_qp_graph with nC = 1, nR = 2: active orbitals as the oracle,
    frozen entries of eGW equal to eOld (their value at QuAcK's evGW call)."""
    say("frozen qp_graph")
    s = mo_systems["H2O"]
    nO, e, mo = s["nO"], s["eHF"], s["eri_mo"]
    nC, nR, n = 1, 2, len(e)
    Ec, Om, XpY, _ = go.screening(e, mo, nO, nC=nC, nR=nR)
    rho = go.excitation_density(mo, XpY, n, nO, nC, nR)
    Sig, Z, _ = go.self_energy_diag(e, Om, rho, nO, 0.1, nC=nC, nR=nR)
    eOld = e + 0.01 * np.arange(n)
    eg, zz = gw.RGW_QP_graph(e, Om, rho, eOld, eOld, Z, nO, eta=0.1, nC=nC, nR=nR)
    og, oz, _ = go.qp_graph(e, Om, rho, eOld, eOld, Z, nO, 0.1, nC=nC, nR=nR)
    assert np.abs(eg - og).max() < 1e-10 and np.abs(zz - oz).max() < 1e-8
    assert np.array_equal(eg[:nC], eOld[:nC]) and np.array_equal(eg[n - nR:], eOld[n - nR:])


@pytest.mark.parametrize("nC,nR", CASES)
def test_frozen_phbse(gw, mo_systems, nC, nR):
    say("frozen phbse", nC, nR)
    s = mo_systems["H2O"]
    nO, e, mo = s["nO"], s["eHF"], s["eri_mo"]
    eGW = go.rg0w0(mo, e, nO, nC=nC, nR=nR)["eGW"]
    for tda_w, tda in ((False, False), (True, True)):
        r = gw.RGW_phBSE(mo, e, eGW, nO, TDA_W=tda_w, TDA=tda, nC=nC, nR=nR)
        Ec, OmS, OmT = go.rgw_phbse(mo, e, eGW, nO, TDA_W=tda_w, TDA=tda, nC=nC, nR=nR)
        assert np.abs(r["OmS"] - OmS).max() < 1e-8 and np.abs(r["OmT"] - OmT).max() < 1e-8
        assert np.abs(r["EcBSE"] - Ec).max() < 1e-8


def test_frozen_batch_evgw(gw):
    """Batched evGW (one launch per kernel for all molecules, device DIIS)
    with nC = 1, nR = 1 on three H8 chains against the oracle per molecule."""
    from quack_amd import intgen as ig
    chains = [ig.system(ig.h_chain(8, seed=k)) for k in (0, 1, 2)]
    nC, nR = 1, 1
    s0 = chains[0]
    nBas, nOrb, nO = s0["nBas"], s0["nOrb"], s0["nO"]
    mos = [go.aotomo_eri_rhf(s["c"], s["eri"]) for s in chains]
    eHF = np.stack([s["eHF"] for s in chains])
    with gw.GWBatch(3, nBas, nOrb, nO, nC, nR) as b:
        for m, mo in enumerate(mos):
            b.set_eri_mo(m, mo)
        say("frozen batch g0w0")
        rg = b.g0w0(eHF, eta=0.1)
        say("frozen batch evgw")
        re = b.evgw(eHF, thresh=1e-6, max_diis=5, eta=0.1)
    for m, (s, mo) in enumerate(zip(chains, mos)):
        og = go.rg0w0(mo, s["eHF"], nO, eta=0.1, nC=nC, nR=nR)
        assert np.abs(rg["eGW"][m] - og["eGW"]).max() < 1e-8, m
        oe = go.evrgw(mo, s["eHF"], nO, thresh=1e-6, max_diis=5, eta=0.1, nC=nC, nR=nR)
        assert re["nSCF"][m] == oe["nSCF"], m
        assert np.abs(re["eGW"][m] - oe["eGW"]).max() < 1e-8, m
