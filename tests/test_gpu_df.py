"""GPU: density-fitted input path (qk_gw_set_eri_df) and the precision form on
the synthetic DF workload recipe (quack_amd/dfgen.py).

FP64 form against the oracle on the same integrals (ERI_MO built from B):
eGW within 1e-8 Ha, Om within 1e-8 Ha; DF against the four-index path of the
same engine within 1e-10 Ha; split-precision form against FP64 within 1 meV
(3.675e-5 Ha) for every QP energy.
"""
import sys

import numpy as np
import pytest

from oracle import gw_oracle as go

pytestmark = pytest.mark.gpu
MEV = 1e-3 / 27.211386245988


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def small():
    from quack_amd import dfgen
    n, nO, naux = 40, 8, 100
    e, B, _ = dfgen.synthetic_df(n, nO, naux, seed=0)
    return {"n": n, "nO": nO, "naux": naux, "e": e, "B": B, "eri_mo": dfgen.eri_mo_from_df(B)}


@pytest.mark.parametrize("kw", [{"doSRG": True}, {"linearize": True}, {"doSRG": True, "TDA_W": True}])
def test_df_g0w0_vs_oracle_and_four_index(small, kw):
    from quack_amd import gw
    say("df g0w0", kw)
    n, nO, e = small["n"], small["nO"], small["e"]
    ref = go.rg0w0(small["eri_mo"], e, nO, **kw)
    with gw.GWEngine(n, n, nO) as g:
        g.set_eri_df(small["B"])
        r = g.g0w0(e, **kw)
        g.set_eri_mo(small["eri_mo"])
        r4 = g.g0w0(e, **kw)
    assert np.abs(r["eGW"] - ref["eGW"]).max() < 1e-8
    assert np.abs(r["Om"] - ref["Om"]).max() < 1e-8
    assert abs(r["EcRPA"] - ref["EcRPA"]) < 1e-9 and abs(r["EcGM"] - ref["EcGM"]) < 1e-9
    assert np.abs(r["eGW"] - r4["eGW"]).max() < 1e-10


def test_df_evgw_vs_oracle(small):
    from quack_amd import gw
    say("df evgw")
    n, nO, e = small["n"], small["nO"], small["e"]
    ref = go.evrgw(small["eri_mo"], e, nO, max_diis=1, thresh=1e-6, doSRG=True)
    with gw.GWEngine(n, n, nO) as g:
        g.set_eri_df(small["B"])
        r = g.evgw(e, max_diis=1, thresh=1e-6, doSRG=True)
    assert r["nSCF"] == ref["nSCF"]
    assert np.abs(r["eGW"] - ref["eGW"]).max() < 1e-8


def test_df_mixed_within_1meV():
    """Scaled-down precision-form workload (N = 240, o = 24, naux = 720,
    nS = 5184; the full recipe is N = 600, naux = 1800), generated on the
    device: split-precision (ia|jb) and excitation densities against FP64."""
    import torch

    from quack_amd import dfgen, gw
    say("df mixed N=240")
    n, nO, naux = 240, 24, 720
    e, Bt, scale = dfgen.synthetic_df(n, nO, naux, seed=3, device="cuda")
    out = {}
    with gw.GWEngine(n, n, nO) as g:
        for form in ("fp64", "mixed", "mixed2"):
            g.set_precision(form)
            g.set_eri_df_device(Bt.data_ptr(), naux)  # K is formed in the handle's precision
            out[form] = (g.g0w0(e, doSRG=True), g.evgw(e, max_diis=1, thresh=1e-6, doSRG=True))
    torch.cuda.synchronize()
    for form in ("mixed", "mixed2"):
        for k in range(2):
            d = np.abs(out[form][k]["eGW"] - out["fp64"][k]["eGW"]).max()
            say(form, k, "max |dE| meV", d / MEV)
            assert d < MEV
        # (ia|jb) uses three slices in both split forms: the same screening
        assert np.abs(out[form][0]["Om"] - out["fp64"][0]["Om"]).max() < 1e-6
    assert np.array_equal(out["mixed"][0]["Om"], out["mixed2"][0]["Om"])


def test_df600_precision_form_within_1meV():
    """The precision form's full configuration (N = 600, o = 60, nS = 32400,
    naux = 1800, seed 0, factors generated on the device): SRG-G0W0 in the
    two-slice split-precision form within 1 meV of FP64 on every orbital, and
    its RPA energies within 1e-7 Ha ((ia|jb) keeps three BF16 slices).  About
    70 s: each nS = 32400 eigensolve takes ~20 s."""
    import torch

    from quack_amd import dfgen, gw
    say("df600 precision form vs fp64")
    n, nO, naux = 600, 60, 1800
    e, Bt, _ = dfgen.synthetic_df(n, nO, naux, seed=0, device="cuda")
    out = {}
    with gw.GWEngine(n, n, nO) as g:
        for form in ("fp64", "mixed2"):
            g.set_precision(form)
            g.set_eri_df_device(Bt.data_ptr(), naux)
            out[form] = g.g0w0(e, doSRG=True)
            say(" ", form, "done")
    torch.cuda.synchronize()
    d = np.abs(out["mixed2"]["eGW"] - out["fp64"]["eGW"])
    dom = np.abs(out["mixed2"]["Om"] - out["fp64"]["Om"]).max()
    say("  max |dE| %.2e meV, max |dOm| %.2e Ha" % (d.max() / MEV, dom))
    assert d.max() < MEV
    assert dom < 1e-7


def test_df_device_layout_matches_host(small):
    """The device-resident layout [q, p, P] (P fastest) is the host array's."""
    import torch

    from quack_amd import gw
    say("df device layout")
    n, nO, naux, e = small["n"], small["nO"], small["naux"], small["e"]
    Bt = torch.from_numpy(np.ascontiguousarray(np.transpose(small["B"], (2, 1, 0)))).to("cuda")
    with gw.GWEngine(n, n, nO) as g:
        g.set_eri_df(small["B"])
        a = g.g0w0(e, doSRG=True)
        g.set_eri_df_device(Bt.data_ptr(), naux)
        b = g.g0w0(e, doSRG=True)
    assert np.abs(a["eGW"] - b["eGW"]).max() < 1e-12


def test_autotuner_identical_results():
    """Tuned and fixed tile choices give bit-identical results (every candidate
    sums the k-terms of an element in the same order), and the tuner reports
    what it chose."""
    import torch

    from quack_amd import dfgen, gw
    say("autotuner")
    n, nO, naux = 240, 24, 720
    e, Bt, _ = dfgen.synthetic_df(n, nO, naux, seed=4, device="cuda")
    out = {}
    try:
        with gw.GWEngine(n, n, nO) as g:
            for form in ("fp64", "mixed"):
                g.set_precision(form)
                for tuned in (False, True):
                    gw.set_autotune(tuned)
                    g.set_eri_df_device(Bt.data_ptr(), naux)
                    out[(form, tuned)] = g.g0w0(e, doSRG=True)
    finally:
        gw.set_autotune(True)
    torch.cuda.synchronize()
    for form in ("fp64", "mixed"):
        for k in ("eGW", "Om", "SigC"):
            assert np.array_equal(out[(form, False)][k], out[(form, True)][k]), (form, k)
    rep = gw.autotune_report()
    say(*rep)
    assert any(l.startswith("dgemm") for l in rep) and any(l.startswith("gemm_bf16x3") for l in rep)


def test_split_gemm_wave_layouts_identical():
    """The split-precision GEMM's two wave layouts (64- and 32-row waves) give
    bit-identical results.  After qk_autotune_reset the first call of each
    large shape tunes, and its output is the last candidate's (32-row waves,
    group 16); the run with the tuner off uses the default (64-row waves,
    group 8) throughout.  Both split forms (three and two slices), FP32 and FP64
    stores (the excitation densities and the DF (ia|jb) product)."""
    import torch

    from quack_amd import dfgen, gw
    say("split gemm wave layouts")
    n, nO, naux = 240, 24, 720
    e, Bt, _ = dfgen.synthetic_df(n, nO, naux, seed=6, device="cuda")
    out = {}
    try:
        with gw.GWEngine(n, n, nO) as g:
            for form in ("mixed", "mixed2"):
                g.set_precision(form)
                for tuned in (True, False):
                    gw.set_autotune(tuned)
                    if tuned:
                        gw.autotune_reset()
                    g.set_eri_df_device(Bt.data_ptr(), naux)
                    out[(form, tuned)] = g.g0w0(e, doSRG=True)
    finally:
        gw.set_autotune(True)
    torch.cuda.synchronize()
    for form in ("mixed", "mixed2"):
        for k in ("eGW", "Om", "SigC"):
            assert np.array_equal(out[(form, False)][k], out[(form, True)][k]), (form, k)
    rep = [l for l in gw.autotune_report() if l.startswith("gemm_bf16x3")]
    say(*rep)
    assert rep and all("w32" in l for l in rep)


def test_library_gemm_candidate():
    """rocBLAS's dgemm as an autotuner candidate (qk_set_library_gemm): the
    large plain FP64 products of the DF path (nS = 5184) time it beside the
    own tiles, and the results stay within rounding of the default (the
    library sums in its own order)."""
    import torch

    from quack_amd import dfgen, gw
    say("library gemm candidate")
    n, nO, naux = 240, 24, 720
    e, Bt, _ = dfgen.synthetic_df(n, nO, naux, seed=5, device="cuda")
    out = {}
    try:
        with gw.GWEngine(n, n, nO) as g:
            g.set_precision("fp64")
            for lib in (False, True):
                gw.set_library_gemm(lib)
                g.set_eri_df_device(Bt.data_ptr(), naux)
                out[lib] = g.g0w0(e, doSRG=True)
    finally:
        gw.set_library_gemm(False)
    torch.cuda.synchronize()
    assert np.abs(out[True]["eGW"] - out[False]["eGW"]).max() < 1e-10
    assert np.abs(out[True]["Om"] - out[False]["Om"]).max() < 1e-10
    rep = gw.autotune_report()
    assert any("rocblas" in l for l in rep if l.startswith("dgemm")), rep
