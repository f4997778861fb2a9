"""GPU: planar H lattices of the headline recipe (1.6 bohr spacing, +-0.15 bohr
in-plane jitter, numpy default_rng(0), cc-pVDZ), the GW path from AO
integrals resident in HBM (GW/RG0W0.f90:88-233, GW/evRGW.f90:104-200).

* 2 x 5 (N = 50) and 5 x 6 (N = 150, nS = 2025): against QuAcK's own routines
  compiled from source (tests/golden/lattice_*_ref.npz, written by
  tests/golden/make_golden.py).  Both keep every basis function, so QuAcK's
  results are valid there.
* 5 x 10 (N = 250, the headline lattice): canonical orthogonalisation keeps
  nOrb = 249 < nBas, where QuAcK's RGW_self_energy_diag (which sizes e and rho
  with nBas, GW/RGW_self_energy_diag.f90:18-20, while RG0W0.f90:175 passes
  nOrb arrays) has no valid answer; the engine uses nOrb throughout.  It is
  held by properties instead: the FP64 form with the library's own
  eigensolver against rocSOLVER's dsyevd; the RPA energies and EcRPA against
  an independent torch route from the same AO integrals; the headline
  split-precision form (two BF16 slices, refined and own full solves) within
  1 meV of FP64 for all 249 orbitals with the same cycle count.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
MEV = 1e-3 / 27.211386245988


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def h50():
    import torch

    from quack_amd import gw, intgen, workloads
    gw.load()
    say("h50 lattice inputs")
    mol = intgen.h_lattice(5, 10, seed=0)
    scf, eri_t, c_t = workloads.device_inputs(mol, log=say)
    eng = gw.GWEngine(mol.nBas, scf["nOrb"], mol.nO)
    eng.set_eri_ao_device(eri_t.data_ptr(), c_t.data_ptr())
    yield {"mol": mol, "scf": scf, "eri_t": eri_t, "eng": eng, "nO": mol.nO}
    eng.close()
    del eri_t, c_t
    torch.cuda.empty_cache()


def _evgw(eng, eHF, precision, eig):
    eng.set_precision(precision)
    eng.set_eigensolver(eig, 0.003)
    return eng.evgw(eHF, maxSCF=256, thresh=1e-6, max_diis=1, doSRG=True)


def test_h50_shape(h50):
    assert h50["mol"].nBas == 250 and h50["scf"]["nOrb"] == 249 and h50["eng"].nS == 5600


def test_h50_own_eigensolver_equals_dsyevd(h50):
    """SRG-evGW in FP64: every full RPA solve by rocSOLVER's dsyevd (QuAcK's
    dsyev, LR/phRLR.f90:67) against the library's own tridiagonalisation +
    dstedc + back-transformation: the same cycle count, QP and RPA energies
    equal to rounding."""
    say("h50 fp64 dsyevd vs own eigensolver")
    eng, eHF = h50["eng"], h50["scf"]["eHF"]
    a = _evgw(eng, eHF, "fp64", "syevd")
    b = _evgw(eng, eHF, "fp64", "sytrd")
    h50["fp64"] = a
    dE, dO = np.abs(a["eGW"] - b["eGW"]).max(), np.abs(a["Om"] - b["Om"]).max()
    say("  nSCF %d / %d, max|dE| %.2e Ha, max|dOm| %.2e Ha" % (a["nSCF"], b["nSCF"], dE, dO))
    assert a["nSCF"] == b["nSCF"]
    assert dE < 1e-10 and dO < 1e-10 and abs(a["EcRPA"] - b["EcRPA"]) < 1e-9


def test_h50_rpa_vs_torch(h50):
    """G0W0 screening at eHF: the RPA energies Om (the engine's AO->MO block,
    RPA build and eigensolve) against sqrt(eigvalsh(D^2 + 4 D^1/2 K D^1/2))
    with K = (ia|jb) built by torch's FP64 matrix products from the same
    resident AO integrals, and EcRPA = (sum Om - Tr A)/2 (LR/phRLR.f90:95) with
    Tr A = sum_ia (e_a - e_i + 2 (ia|ia)) from that K."""
    import torch

    from quack_amd import gw, workloads
    say("h50 rpa vs torch")
    eng, scf, nO = h50["eng"], h50["scf"], h50["nO"]
    eng.set_precision("fp64")
    eng.set_eigensolver("syevd")
    r = eng.g0w0(scf["eHF"], doSRG=True)
    K = workloads.ia_block_torch(h50["eri_t"], scf["c"], nO)
    e = torch.from_numpy(scf["eHF"]).to(K.device)
    D = (e[None, nO:] - e[:nO, None]).reshape(-1)
    sd = torch.sqrt(D)
    C = 4.0 * sd[:, None] * K * sd[None, :]
    C.diagonal().add_(D * D)
    lam = torch.linalg.eigvalsh(C)
    Om = torch.sqrt(lam).cpu().numpy()
    trA = float((D + 2.0 * K.diagonal()).sum())
    del K, C
    torch.cuda.empty_cache()
    dO = np.abs(r["Om"] - Om).max()
    ec = 0.5 * (Om.sum() - trA)
    say("  max|dOm| %.2e Ha, EcRPA %.12f vs %.12f" % (dO, r["EcRPA"], ec))
    assert dO < 1e-8
    # G0W0's EcRPA comes from its second RPA at eGW (RG0W0.f90:222-225); the
    # one at eHF is evGW's first cycle (maxSCF = 0 runs exactly one cycle and
    # reports non-convergence with that cycle's outputs)
    try:
        r1 = eng.evgw(scf["eHF"], maxSCF=0, thresh=0.0, max_diis=1, doSRG=True)
    except gw.NotConverged as ex:
        r1 = ex.result
    say("  EcRPA(eHF) engine %.12f" % r1["EcRPA"])
    assert abs(r1["EcRPA"] - ec) < 1e-8 and np.abs(r1["Om"] - Om).max() < 1e-8


def test_h50_headline_form_within_1meV(h50):
    """The split-precision form (two-slice excitation densities, refined and
    own full solves) against the FP64 dsyevd form: all 249 QP energies within
    1 meV with the same number of evGW cycles, SRG-regularised and at
    eta = 0.1 (DIIS 5).  The RPA energies of the last cycles are taken at
    energies that differ by the split-precision error, so the RPA bar is
    checked on G0W0's screening at eHF, the same input in both forms."""
    say("h50 split-precision form vs fp64")
    eng, eHF = h50["eng"], h50["scf"]["eHF"]
    ref = h50.get("fp64") or _evgw(eng, eHF, "fp64", "syevd")
    r = _evgw(eng, eHF, "mixed2", "refine+sytrd")
    dE = np.abs(r["eGW"] - ref["eGW"])
    say("  SRG: nSCF %d / %d, max|dE| %.3e meV (orbital %d)" % (r["nSCF"], ref["nSCF"], dE.max() / MEV,
                                                               int(np.argmax(dE))))
    assert r["nSCF"] == ref["nSCF"]
    assert dE.max() < MEV
    kw = {"maxSCF": 256, "thresh": 1e-6, "max_diis": 5, "eta": 0.1}
    eng.set_precision("fp64")
    eng.set_eigensolver("syevd")
    ref = eng.evgw(eHF, **kw)
    g_ref = eng.g0w0(eHF, eta=0.1)
    eng.set_precision("mixed2")
    eng.set_eigensolver("refine+sytrd", 0.003)
    r = eng.evgw(eHF, **kw)
    g = eng.g0w0(eHF, eta=0.1)
    dE = np.abs(r["eGW"] - ref["eGW"])
    say("  eta 0.1: nSCF %d / %d, max|dE| %.3e meV; G0W0 max|dOm| %.2e Ha max|dE| %.3e meV" %
        (r["nSCF"], ref["nSCF"], dE.max() / MEV, np.abs(g["Om"] - g_ref["Om"]).max(),
         np.abs(g["eGW"] - g_ref["eGW"]).max() / MEV))
    assert r["nSCF"] == ref["nSCF"] and dE.max() < MEV
    assert np.abs(g["Om"] - g_ref["Om"]).max() < 1e-8 and np.abs(g["eGW"] - g_ref["eGW"]).max() < MEV


def test_h50_learned_form(h50):
    """The learned warm start on the headline lattice (a molecule the
    training set did not hold): every QP energy within 1 meV of plain evGW of
    the same method (eta = 0.1, DIIS 5, 1e-6 Ha) in FP64, in FP64 with fewer
    full RPA diagonalisations, and in the headline form (two-slice split
    precision) within 1 meV of the plain FP64 dsyevd form."""
    from quack_amd import learned
    say("h50 learned form")
    eng, scf = h50["eng"], h50["scf"]
    sysd = {"eHF": scf["eHF"], "Sx": scf["Sx"], "nO": h50["nO"]}
    eng.set_precision("fp64")
    eng.set_eigensolver("syevd")
    ref = learned.plain(eng, sysd)
    eng.set_eigensolver("refine+sytrd", 0.003)
    s0 = eng.eig_stats()
    plain = learned.plain(eng, sysd)
    s1 = eng.eig_stats()
    ws = learned.WarmStart()
    try:
        r = ws.evgw(eng, sysd)
        s2 = eng.eig_stats()
        eng.set_precision("mixed2")
        rm = ws.evgw(eng, sysd)
    finally:
        ws.close()
    full_plain, full_learned = s1["syevd"] - s0["syevd"], s2["syevd"] - s1["syevd"]
    d = np.abs(r["eGW"] - ref["eGW"])
    dm = np.abs(rm["eGW"] - ref["eGW"])
    say("  plain %d cycles %d full solves; learned %d cycles %d full solves; max|dE| fp64 %.3e meV, split "
        "precision %.3e meV; start error %.2f meV" % (plain["nSCF"], full_plain, r["nSCF_total"], full_learned,
                                                       d.max() / MEV, dm.max() / MEV,
                                                       np.abs(r["start"] - ref["eGW"]).max() / MEV))
    assert np.abs(plain["eGW"] - ref["eGW"]).max() < 1e-8
    assert d.max() < MEV and dm.max() < MEV
    assert full_learned < full_plain


@pytest.mark.parametrize("name", ["lattice_2x5", "lattice_5x6"])
def test_lattice_vs_quack(name):
    """G0W0 (linearised at eta = 0, graphical at eta = 0.05 and 0.1, SRG) and
    evGW (SRG with DIIS off; eta = 0.05 and 0.1 with DIIS 5) from device-resident AO
    integrals against QuAcK's compiled routines on the same RHF solution: QP
    and RPA energies within 1e-8 Ha, the same evGW cycle counts; the
    split-precision form within 1 meV of QuAcK's evGW.  (The eta = 0
    graphical G0W0 is not compared: near a pole one orbital's root switched
    between two FP64 builds whose RPA energies agree to 1e-13.)"""
    import torch

    from quack_amd import gw, intgen, workloads
    say("lattice vs quack", name)
    ref = np.load(os.path.join(HERE, "golden", name + "_ref.npz"))
    mol = intgen.h_lattice(int(ref["nx"]), int(ref["ny"]), seed=0)
    scf, eri_t, c_t = workloads.device_inputs(mol, c=ref["c"], eHF=ref["eHF"])
    nO = int(ref["nO"])
    with gw.GWEngine(mol.nBas, scf["nOrb"], nO) as eng:
        eng.set_eri_ao_device(eri_t.data_ptr(), c_t.data_ptr())
        for tag, kw in (("g0w0_lin", {"linearize": True}), ("g0w0_eta", {"eta": 0.05}), ("g0w0_eta01", {"eta": 0.1}),
                        ("g0w0_srg", {"doSRG": True})):
            r = eng.g0w0(ref["eHF"], **kw)
            say("  %s max|dE| %.2e max|dOm| %.2e" % (tag, np.abs(r["eGW"] - ref[tag + "_eGW"]).max(),
                                                     np.abs(r["Om"] - ref[tag + "_Om"]).max()))
            assert np.abs(r["Om"] - ref[tag + "_Om"]).max() < 1e-8, tag
            assert np.abs(r["eGW"] - ref[tag + "_eGW"]).max() < 1e-8, tag
            assert abs(r["EcRPA"] - ref[tag + "_EcRPA"]) < 1e-8 and abs(r["EcGM"] - ref[tag + "_EcGM"]) < 1e-8
        ev = (("evgw_srg", {"max_diis": 1, "doSRG": True}), ("evgw_eta", {"max_diis": 5, "eta": 0.05}),
              ("evgw_eta01", {"max_diis": 5, "eta": 0.1}))
        for tag, kw in ev:
            r = eng.evgw(ref["eHF"], maxSCF=256, thresh=1e-6, **kw)
            say("  %s nSCF %d / %d max|dE| %.2e" % (tag, r["nSCF"], int(ref[tag + "_nSCF"]),
                                                   np.abs(r["eGW"] - ref[tag + "_eGW"]).max()))
            assert r["nSCF"] == int(ref[tag + "_nSCF"]), tag
            assert np.abs(r["eGW"] - ref[tag + "_eGW"]).max() < 1e-8, tag
            assert abs(r["EcRPA"] - ref[tag + "_EcRPA"]) < 1e-8, tag
        eng.set_precision("mixed2")
        eng.set_eigensolver("refine+sytrd", 0.003)
        for tag, kw in ev:
            r = eng.evgw(ref["eHF"], maxSCF=256, thresh=1e-6, **kw)
            assert r["nSCF"] == int(ref[tag + "_nSCF"]), tag
            assert np.abs(r["eGW"] - ref[tag + "_eGW"]).max() < MEV, tag
    del eri_t, c_t
    torch.cuda.empty_cache()
