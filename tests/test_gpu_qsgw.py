"""GPU: the full self-energy (qk_gw_sigma_full) and qsGW (quack_amd.qsgw)
against the CPU restatements of GW/RGW_self_energy.f90,
GW/RGW_SRG_self_energy.f90 and GW/qsRGW.f90 (oracle/gw_oracle.py, whose
self-energy and first qsGW cycle are checked against QuAcK's compiled
routines in tests/test_oracle.py)."""
import numpy as np
import pytest

from oracle import gw_oracle as go

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["H2O", "H10"])
def test_sigma_full_matches_oracle(mo_systems, name):
    from quack_amd import gw
    print("sigma full", name, flush=True)
    s = mo_systems[name]
    n, nO, mo = len(s["eHF"]), s["nO"], s["eri_mo"]
    A, B = go.phRLR_A(s["eHF"], mo, nO), go.phRLR_B(mo, n, nO)
    EcRPA, Om, XpY, _ = go.phRLR(False, A, B)
    rho = go.excitation_density(mo, XpY, n, nO)
    with gw.GWEngine(n, n, nO) as g:
        g.set_eri_mo(mo)
        g.screen(s["eHF"])
        for srg, eta in ((False, 0.0), (False, 0.05), (True, 0.0)):
            r = g.sigma_full(s["eHF"], eta=eta, doSRG=srg)
            if srg:
                Ec, Sig, Z = go.rgw_srg_self_energy(go.FLOW, s["eHF"], Om, rho, nO)
            else:
                Ec, Sig, Z = go.rgw_self_energy(eta, s["eHF"], Om, rho, nO)
            # eta = 0: denominators e_p - e_r +- Om_m close to zero amplify the
            # rounding of the two eigensolvers' vectors (4.8e-10 measured at H10)
            scale = max(1.0, np.abs(Sig).max())
            assert np.abs(r["SigC"] - Sig).max() < (1e-8 if eta == 0.0 and not srg else 1e-10) * scale, \
                (name, srg, eta)
            # Z = 1/(1 - dSigma) amplifies the rounding by Z^2: at H10 with eta =
            # 0.05 one orbital has Z = 10.35 and the two paths differ by 7e-9
            assert np.abs(r["Z"] - Z).max() < 1e-8 * np.abs(Z).max() ** 2
            assert abs(r["EcGM"] - Ec) < 1e-10
            assert abs(r["EcRPA"] - EcRPA) < 1e-10


@pytest.mark.parametrize("name", ["He", "H2O"])
def test_qsgw_matches_oracle(systems, name):
    """qsGW with QuAcK's defaults (maxSCF 64, thresh 1e-5, DIIS 5), SRG and
    eta = 0.05, through qk_qsgw_create / qk_qsgw_cycle: same cycle count,
    quasiparticle energies within 1e-8 Ha, and on H2O the same with one
    frozen core and two frozen virtual orbitals."""
    from quack_amd import qsgw, scf
    print("qsgw", name, flush=True)
    s = systems[name]
    mol = s["mol"]
    S, T, V = mol.one_electron()
    X = scf.orthogonalization_matrix(S)
    nO = s["nO"]
    c = np.asarray(s["c"])
    P = 2.0 * c[:, :nO] @ c[:, :nO].T
    cases = [{"doSRG": True}, {"eta": 0.05}]
    if name == "H2O":
        cases.append({"doSRG": True, "nC": 1, "nR": 2})
    for kw in cases:
        print("qsgw", name, kw, flush=True)
        r = qsgw.qsRGW(s["eri"], S, T, V, X, c, s["eHF"], nO, **kw)
        ro = go.qsrgw(S, T, V, s["eri"], X, c, P, s["eHF"], nO, **kw)
        print(kw, "cycles", r["nSCF"], ro["nSCF"], "EqsGW", r["EqsGW"], flush=True)
        assert ro["converged"] and r["nSCF"] == ro["nSCF"], kw
        assert np.abs(r["eGW"] - ro["eGW"]).max() < 1e-8, kw
        assert abs(r["EqsGW"] - ro["EqsGW"]) < 1e-9 and abs(r["EcGM"] - ro["EcGM"]) < 1e-9


def test_qsgw_dropped_functions():
    """nBas > nOrb (a compressed H4 chain: canonical orthogonalisation drops one
    of 20 functions): QuAcK's second branch (F projected on the kept
    orbitals, the current orbitals rotated) against the restatement."""
    from quack_amd import intgen as ig
    from quack_amd import qsgw, scf
    print("qsgw H4 nOrb < nBas", flush=True)
    mol = ig.h_chain(4, spacing=0.3, jitter=0.0, seed=0)
    s = ig.system(mol)
    S, T, V = mol.one_electron()
    X = scf.orthogonalization_matrix(S)
    assert X.shape == (20, 19)
    nO, c = s["nO"], np.asarray(s["c"])
    P = 2.0 * c[:, :nO] @ c[:, :nO].T
    r = qsgw.qsRGW(s["eri"], S, T, V, X, c, s["eHF"], nO, doSRG=True)
    ro = go.qsrgw(S, T, V, s["eri"], X, c, P, s["eHF"], nO, doSRG=True)
    assert ro["converged"] and r["nSCF"] == ro["nSCF"]
    assert np.abs(r["eGW"] - ro["eGW"]).max() < 1e-8
    # eta = 0.05 does not converge here in QuAcK's 64 cycles (nor in the
    # restatement): the loop stops with the error and the last cycle's results
    with pytest.raises(qsgw.QsGWNotConverged) as ex:
        qsgw.qsRGW(s["eri"], S, T, V, X, c, s["eHF"], nO, eta=0.05)
    assert ex.value.result["nSCF"] == 64 and np.all(np.isfinite(ex.value.result["eGW"]))
