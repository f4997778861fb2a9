"""Learned warm start (quack_amd/learned.py + models/mlp_evgw_srg.npz, trained
on the GPU by the project's MLP kernels) against plain evGW, which judges it:
frontier orbitals within 1 meV, no more diagonalisations than plain evGW."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
MEV = 1e-3 / 27.211386245988


def test_warm_start_h10(systems):
    from quack_amd import gw, learned
    print("learned warm start H10", flush=True)
    s = systems["H10"]
    ws = learned.WarmStart(learned.DEFAULT_MODEL)
    with gw.GWEngine(s["nBas"], s["nOrb"], s["nO"]) as eng:
        eng.set_eri_ao(s["eri"], s["c"])
        ref = eng.evgw(s["eHF"], maxSCF=128, thresh=1e-6, max_diis=1, doSRG=True)
        r = ws.evgw(eng, s, doSRG=True, thresh=1e-5)
    nO = s["nO"]
    assert r["nSCF"] + 1 <= ref["nSCF"]
    assert abs(r["eGW"][nO - 1] - ref["eGW"][nO - 1]) < MEV
    assert abs(r["eGW"][nO] - ref["eGW"][nO]) < MEV
    e = s["eHF"]
    win = np.abs(e - 0.5 * (e[nO - 1] + e[nO])) < 1.0
    assert np.abs(r["eGW"] - ref["eGW"])[win].max() < MEV


def test_model_file_consistent():
    from quack_amd import learned
    d = np.load(learned.DEFAULT_MODEL)
    assert d["params"].dtype == np.float32 and d["mean"].shape == (learned.NFEAT,)
    width, depth = int(d["width"]), int(d["depth"])
    n = learned.NFEAT * width + width + (depth - 1) * (width * width + width) + width + 1
    assert d["params"].size == n
