"""Writes the committed fixtures of tests/golden/ (run in a tree with the QuAcK sources).

* quack_goldens.json: the scalar results QuAcK's own test files hold for this
  path (He/6-31G and H2O/cc-pVDZ; tests/feather_bench.py:5-54,58-101 and
  test/Rtest_ref.dat:51-62), plus the options they were produced with.
* h10_ref.npz: outputs of QuAcK's own routines (oracle/_ref/libqkref.so,
  compiled from the QuAcK sources by oracle/Makefile) for the H10 chain
  workload (1.40 bohr + U(-0.10, 0.10), numpy default_rng(0), cc-pVDZ,
  N = 50): the RHF inputs the routines were given (eHF and the MO
  coefficients c, so that tests rebuild identical MO integrals from the
  deterministic AO integrals) and the sign- and rotation-invariant outputs
  (Om, eGW, Z, SigC, EcRPA, EcGM, nSCF).
* lattice_2x5_ref.npz, lattice_5x6_ref.npz: the same for planar H lattices
  of the headline recipe (1.6 bohr spacing, +-0.15 bohr in-plane jitter,
  default_rng(0), cc-pVDZ): 2 x 5 (N = 50, nS = 225) and 5 x 6 (N = 150,
  nS = 2025).  Both keep every basis function (nOrb = nBas), so QuAcK's
  RGW_self_energy_diag, which sizes its arrays with nBas, is valid there.
  G0W0 (linearised at eta = 0, graphical at eta = 0.05 and 0.1, SRG),
  SRG-evGW to 1e-6 Ha with DIIS off and eta = 0.05 and 0.1 evGW with DIIS 5.

    python tests/golden/make_golden.py [h10] [lattice]
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import qkref  # noqa: E402
from quack_amd import intgen as ig  # noqa: E402

GOLDENS = {
    "source": {
        "feather_bench": "tests/feather_bench.py:13-39 (He), :68-94 (H2O)",
        "Rtest_ref": "test/Rtest_ref.dat:1-6,51-62 (He)",
    },
    "gw_options": {"maxSCF": 256, "thresh": 1e-5, "lin": False, "eta": 0.0, "TDA_W": False, "reg": False},
    "He": {
        "basis": "6-31g",
        "feather": {"RHF energy": -2.855160426154444, "RHF HOMO energy": -0.914126628640145,
                    "RHF LUMO energy": 1.399859335255765,
                    "G0W0 correlation energy": -0.019314094399372, "G0W0 HOMO energy": -0.87053388021722,
                    "G0W0 LUMO energy": 1.377171287041735, "evGW correlation energy": -0.019335511771337,
                    "evGW HOMO energy": -0.868460640984803, "evGW LUMO energy": 1.376287581502582},
        "Rtest_ref": {"RHF energy": -2.855160426884076, "RHF HOMO energy": -0.914126628614305,
                      "RHF LUMO energy": 1.399859335225087,
                      "G0W0 correlation energy": -0.019314094399756, "G0W0 HOMO energy": -0.870533880190454,
                      "G0W0 LUMO energy": 1.377171287010956, "evGW correlation energy": -0.019335511771724,
                      "evGW HOMO energy": -0.868460640957913, "evGW LUMO energy": 1.376287581471769},
    },
    "H2O": {
        "basis": "cc-pvdz",
        "geometry_angstrom": [["O", 0.0, 0.0, 0.0], ["H", 0.7571, 0.0, 0.5861], ["H", -0.7571, 0.0, 0.5861]],
        "feather": {"RHF energy": -85.21935817501823, "RHF HOMO energy": -0.493132793449897,
                    "RHF LUMO energy": 0.185534869842355,
                    "G0W0 correlation energy": -0.23853664665404, "G0W0 HOMO energy": -0.446828623007469,
                    "G0W0 LUMO energy": 0.173026609033024, "evGW correlation energy": -0.239414217281308,
                    "evGW HOMO energy": -0.443076613314424, "evGW LUMO energy": 0.172691758111392},
    },
}


def lattice(nx, ny, name):
    """Reference outputs for one lattice.  The RHF solution (c, eHF) of an
    existing fixture file is kept, so that regenerating the outputs leaves the
    inputs bit-identical; the AO integrals are deterministic."""
    mol = ig.h_lattice(nx, ny, seed=0)
    path = os.path.join(HERE, name)
    if os.path.exists(path):
        old = np.load(path)
        s = {"eri": mol.eri(), "c": old["c"], "eHF": old["eHF"], "ERHF": old["ERHF"], "nO": int(old["nO"]),
             "nOrb": old["c"].shape[1], "nBas": mol.nBas}
    else:
        s = ig.system(mol)
    assert s["nOrb"] == s["nBas"], "a lattice whose RHF drops functions has no valid QuAcK result"
    nO = s["nO"]
    mo = qkref.aotomo(s["c"], s["eri"])
    out = {"eHF": s["eHF"], "ERHF": s["ERHF"], "nO": nO, "c": s["c"], "nx": nx, "ny": ny}
    # eta = 0 only linearised: its graphical solve is sensitive at the 1e-13 level
    # near poles (a root can switch for one orbital between two FP64 builds)
    for tag, kw in (("g0w0_lin", {"linearize": True}), ("g0w0_eta", {"eta": 0.05}), ("g0w0_eta01", {"eta": 0.1}),
                    ("g0w0_srg", {"srg": True})):
        r = qkref.rg0w0(mo, s["eHF"], nO, **kw)
        for k in ("eGW", "SigC", "Z", "Om", "EcRPA", "EcGM"):
            out[tag + "_" + k] = r[k]
    for tag, kw in (("evgw_srg", {"max_diis": 1, "srg": True}), ("evgw_eta", {"max_diis": 5, "eta": 0.05}),
                    ("evgw_eta01", {"max_diis": 5, "eta": 0.1})):
        r = qkref.evrgw(mo, s["eHF"], nO, maxSCF=256, thresh=1e-6, **kw)
        for k in ("eGW", "SigC", "Z", "EcRPA", "EcGM", "nSCF", "status"):
            out[tag + "_" + k] = r[k]
    np.savez(path, **{k: np.asarray(v) for k, v in out.items()})


def main():
    what = sys.argv[1:] or ["h10", "lattice"]
    if "lattice" in what:
        lattice(2, 5, "lattice_2x5_ref.npz")
        lattice(5, 6, "lattice_5x6_ref.npz")
    if "h10" not in what:
        return
    with open(os.path.join(HERE, "quack_goldens.json"), "w") as f:
        json.dump(GOLDENS, f, indent=1)
    s = ig.system(ig.h_chain(10, seed=0))
    nO = s["nO"]
    mo = qkref.aotomo(s["c"], s["eri"])
    out = {"eHF": s["eHF"], "ERHF": s["ERHF"], "nO": nO, "c": s["c"]}
    for tag, kw in (("g0w0", {}), ("g0w0_lin", {"linearize": True}), ("g0w0_srg", {"srg": True}),
                    ("g0w0_tda", {"tda_w": True})):
        r = qkref.rg0w0(mo, s["eHF"], nO, **kw)
        for k in ("eGW", "SigC", "Z", "Om", "EcRPA", "EcGM"):
            out[tag + "_" + k] = r[k]
    for tag, kw in (("evgw", {"max_diis": 5, "thresh": 1e-6}), ("evgw_nodiis", {"max_diis": 1, "thresh": 1e-6}),
                    ("evgw_lin", {"linearize": True, "max_diis": 5, "thresh": 1e-6, "maxSCF": 1})):
        r = qkref.evrgw(mo, s["eHF"], nO, **kw)
        for k in ("eGW", "SigC", "Z", "EcRPA", "EcGM", "nSCF", "status"):
            out[tag + "_" + k] = r[k]
    np.savez(os.path.join(HERE, "h10_ref.npz"), **{k: np.asarray(v) for k, v in out.items()})


if __name__ == "__main__":
    main()
