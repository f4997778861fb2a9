"""GPU: phBSE@G0W0 (qk_rgw_phbse, qk_rgw_phbse_dyn, GW/RGW_phBSE.f90) against the CPU
restatement (oracle/gw_oracle.py rgw_phbse, itself checked against QuAcK's
compiled routines in tests/test_oracle.py): singlet and triplet excitation
energies within 1e-8 Ha, EcBSE within 1e-8 Ha, for full and TDA BSE, full and
TDA screening, eta = 0 and 0.05."""
import numpy as np
import pytest

from oracle import gw_oracle as go

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["He", "H2O", "H10"])
def test_phbse_matches_oracle(mo_systems, name):
    from quack_amd import gw
    print("phbse", name, flush=True)
    s = mo_systems[name]
    nO, mo = s["nO"], s["eri_mo"]
    eGW = gw.RG0W0(mo, s["eHF"], nO)["eGW"]
    for tda_w, tda, eta in ((False, False, 0.0), (False, True, 0.0), (True, False, 0.05)):
        r = gw.RGW_phBSE(mo, s["eHF"], eGW, nO, eta=eta, TDA_W=tda_w, TDA=tda)
        Ec, OmS, OmT = go.rgw_phbse(mo, s["eHF"], eGW, nO, eta=eta, TDA_W=tda_w, TDA=tda)
        assert np.abs(r["OmS"] - OmS).max() < 1e-8, (name, tda_w, tda)
        assert np.abs(r["OmT"] - OmT).max() < 1e-8, (name, tda_w, tda)
        assert np.abs(r["EcBSE"] - Ec).max() < 1e-8, (name, tda_w, tda)


def test_phbse_options(mo_systems):
    """Singlet or triplet alone, exchange_kernel halving, and the argument
    checks of the second-order kernel and the dynamical correction: dBSE
    through qk_rgw_phbse (which has no outputs for it) is QK_E_ARG, and both
    with frozen orbitals are QK_E_UNSUPPORTED."""
    from quack_amd import gw
    s = mo_systems["H2O"]
    nO, mo = s["nO"], s["eri_mo"]
    full = gw.RGW_phBSE(mo, s["eHF"], s["eHF"], nO)
    only_s = gw.RGW_phBSE(mo, s["eHF"], s["eHF"], nO, triplet=False)
    only_t = gw.RGW_phBSE(mo, s["eHF"], s["eHF"], nO, singlet=False)
    half = gw.RGW_phBSE(mo, s["eHF"], s["eHF"], nO, exchange_kernel=True)
    assert np.array_equal(only_s["OmS"], full["OmS"]) and not only_s["OmT"].any()
    assert np.array_equal(only_t["OmT"], full["OmT"]) and not only_t["OmS"].any()
    assert np.allclose(half["EcBSE"], 0.5 * full["EcBSE"], rtol=0, atol=1e-14)
    n = len(s["eHF"])
    nS = nO * (n - nO)
    Ec, Om = np.zeros(2), np.zeros(2 * nS)
    lib = gw.load()
    rc = lib.qk_rgw_phbse(0, 0, 0, 0, 1, 0, 1, 1, 0.0, n, 0, nO, n - nO, 0, nS,
                          gw._p(gw._f64(mo)), gw._p(s["eHF"]), gw._p(s["eHF"]), gw._p(Ec), gw._p(Om))
    assert rc == gw.QK_E_ARG
    nS1 = (nO - 1) * (n - nO)
    Om1, Od, Zd = np.zeros(2 * nS1), np.zeros(20), np.zeros(20)
    for flags in ((1, 0), (0, 1)):
        rc = lib.qk_rgw_phbse_dyn(flags[0], 0, 0, 0, flags[1], 0, 1, 1, 0.0, n, 1, nO, n - nO, 0, nS1,
                                  gw._p(gw._f64(mo)), gw._p(s["eHF"]), gw._p(s["eHF"]), gw._p(Ec), gw._p(Om1),
                                  gw._p(Od), gw._p(Zd))
        assert rc == gw.QK_E_UNSUPPORTED


@pytest.mark.parametrize("name", ["He", "H2O"])
def test_phbse2_dynamic_matches_oracle(mo_systems, name):
    """The second-order static kernel (W of RGW_phBSE_static_kernel as one
    FP64 matrix-core product, k_bse2_static) and the first-order dynamical
    correction (k_bse_dyn's quadratic forms), full and dynamical TDA, full
    and TDA BSE, with and without the second-order dynamic kernels, against
    the oracle (pinned to QuAcK's compiled routines): excitation energies,
    EcBSE and the corrections and their factors within 1e-8."""
    from quack_amd import gw
    s = mo_systems[name]
    nO, mo, e = s["nO"], s["eri_mo"], s["eHF"]
    eGW = gw.RG0W0(mo, e, nO, eta=0.1)["eGW"]
    cases = [{"dophBSE2": True}, {"dBSE": True, "dTDA": False}, {"dBSE": True, "dTDA": True},
             {"dophBSE2": True, "dBSE": True, "dTDA": False}, {"dophBSE2": True, "dBSE": True, "TDA": True},
             {"dophBSE2": True, "singlet": False}]
    for kw in cases:
        print("phbse2/dyn", name, kw, flush=True)
        r = gw.RGW_phBSE(mo, e, eGW, nO, eta=0.1, **kw)
        dyn = {}
        Ec, OmS, OmT = go.rgw_phbse(mo, e, eGW, nO, eta=0.1, dyn=dyn, **kw)
        assert np.abs(r["OmS"] - OmS).max() < 1e-8 and np.abs(r["OmT"] - OmT).max() < 1e-8, kw
        assert np.abs(r["EcBSE"] - Ec).max() < 1e-8, kw
        assert set(r["dyn"]) == set(dyn), kw
        for man, (od, zd) in dyn.items():
            assert np.abs(r["dyn"][man][0] - od).max() < 1e-8 and np.abs(r["dyn"][man][1] - zd).max() < 1e-8, \
                (kw, man)
