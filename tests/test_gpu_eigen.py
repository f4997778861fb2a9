"""GPU: refined RPA eigensolver (qk_gw_set_eigensolver mode 1) against dsyevd.

Mode 1 replaces all but the first few dsyevd calls of an evGW run by Newton
refinement of the previous cycle's eigenvectors.  It must reproduce QuAcK's
results to the FP64-form bars (eGW, EcRPA, EcGM within 1e-8 Ha), and its
fallback to dsyevd must give the same results when the refinement is tried
outside its convergence region.
"""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
H10 = np.load(os.path.join(HERE, "golden", "h10_ref.npz"))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_refined_evgw_h10_vs_quack(systems):
    from quack_amd import gw
    say("refined evgw H10")
    s = systems["H10"]
    with gw.GWEngine(s["nBas"], s["nOrb"], s["nO"]) as g:
        g.set_eri_ao(s["eri"], s["c"])
        g.set_eigensolver("refine")
        for tag, kw in (("evgw", {"max_diis": 5, "thresh": 1e-6}), ("evgw_nodiis", {"max_diis": 1, "thresh": 1e-6})):
            before = g.eig_stats()
            r = g.evgw(s["eHF"], **kw)
            st = {k: g.eig_stats()[k] - before[k] for k in before}
            say(tag, st)
            assert st["refined"] >= 3, (tag, st)
            assert st["syevd"] + st["refined"] == r["nSCF"], (tag, st)
            assert r["nSCF"] == int(H10[tag + "_nSCF"]), tag
            assert np.abs(r["eGW"] - H10[tag + "_eGW"]).max() < 1e-8, tag
            assert abs(r["EcRPA"] - H10[tag + "_EcRPA"]) < 1e-8, tag
            assert abs(r["EcGM"] - H10[tag + "_EcGM"]) < 1e-8, tag


def test_refined_screening_matches_syevd(systems):
    """Om and the sign-invariant self-energy from a refined solve equal dsyevd's."""
    from quack_amd import gw
    say("refined screening H10")
    s = systems["H10"]
    e = s["eHF"]
    rng = np.random.default_rng(3)
    e2 = e + 1e-3 * rng.standard_normal(e.size)
    with gw.GWEngine(s["nBas"], s["nOrb"], s["nO"]) as g:
        g.set_eri_ao(s["eri"], s["c"])
        g.screen(e2)
        ref = g.sigma_at(e2, e2 + 0.3, eta=0.05)
        g.set_eigensolver("refine", de_max=0.01)
        before = g.eig_stats()
        g.screen(e)  # dsyevd: no eigenvectors yet
        g.screen(e2)  # refined from the eigenvectors at e
        st = {k: g.eig_stats()[k] - before[k] for k in before}
        assert st["refined"] == 1 and st["syevd"] == 1, st
        got = g.sigma_at(e2, e2 + 0.3, eta=0.05)
    for a, b in zip(got, ref):
        assert np.abs(a - b).max() < 1e-10 * max(1.0, np.abs(b).max())


def test_refinement_fallback(systems):
    """Refinement tried from the eigenvectors of a very different matrix
    (virtual energies tripled) is abandoned for dsyevd; the result is dsyevd's."""
    from quack_amd import gw
    say("refinement fallback H10")
    s = systems["H10"]
    e = s["eHF"]
    nO = s["nO"]
    far = e.copy()
    far[nO:] *= 3.0
    with gw.GWEngine(s["nBas"], s["nOrb"], nO) as g:
        g.set_eri_ao(s["eri"], s["c"])
        g.screen(e)
        ref = g.sigma_at(e, e + 0.3, eta=0.05)
        g.set_eigensolver("refine", de_max=100.0)
        g.screen(far)
        before = g.eig_stats()
        g.screen(e)
        st = {k: g.eig_stats()[k] - before[k] for k in before}
        say(st)
        assert st["fallbacks"] == 1 and st["syevd"] == 1, st
        got = g.sigma_at(e, e + 0.3, eta=0.05)
    for a, b in zip(got, ref):
        assert np.abs(a - b).max() < 1e-10 * max(1.0, np.abs(b).max())


def test_refined_evgw_any_start_distance(systems):
    """With refinement allowed from any distance (de_max 100), every cycle of
    evGW is refined or falls back, and the result is QuAcK's."""
    from quack_amd import gw
    say("refined evgw de_max=100 H10")
    s = systems["H10"]
    with gw.GWEngine(s["nBas"], s["nOrb"], s["nO"]) as g:
        g.set_eri_ao(s["eri"], s["c"])
        g.set_eigensolver("refine", de_max=100.0)
        r = g.evgw(s["eHF"], max_diis=5, thresh=1e-6)
        st = g.eig_stats()
        say(st)
        assert st["syevd"] == 1 + st["fallbacks"] and st["refined"] >= r["nSCF"] - st["syevd"]
        assert np.abs(r["eGW"] - H10["evgw_eGW"]).max() < 1e-8
        assert r["nSCF"] == int(H10["evgw_nSCF"])


@pytest.mark.parametrize("mode", ["sytrd", "refine+sytrd"])
def test_own_tridiagonalisation_evgw_h10_vs_quack(systems, mode):
    """Full solves through the library's own Householder tridiagonalisation
    reproduce QuAcK's evGW (values, cycle counts) like dsyevd does."""
    from quack_amd import gw
    say("own tridiagonalisation evgw H10", mode)
    s = systems["H10"]
    with gw.GWEngine(s["nBas"], s["nOrb"], s["nO"]) as g:
        g.set_eri_ao(s["eri"], s["c"])
        g.set_eigensolver(mode)
        for tag, kw in (("evgw", {"max_diis": 5, "thresh": 1e-6}), ("evgw_nodiis", {"max_diis": 1, "thresh": 1e-6})):
            r = g.evgw(s["eHF"], **kw)
            assert r["nSCF"] == int(H10[tag + "_nSCF"]), tag
            assert np.abs(r["eGW"] - H10[tag + "_eGW"]).max() < 1e-8, tag
            assert abs(r["EcRPA"] - H10[tag + "_EcRPA"]) < 1e-8, tag
        assert g.eig_stats()["own_tridiagonal"] >= 1


def test_own_tridiagonalisation_matches_dsyevd():
    """A larger problem (synthetic DF, nS = 1296: many 32-column panels and
    64 x 64 tiles): G0W0 with the own full solver, including the values-only
    second RPA, against dsyevd."""
    from quack_amd import dfgen, gw
    say("own tridiagonalisation vs dsyevd, nS = 1296")
    n, nO, naux = 120, 12, 360
    e, B, _ = dfgen.synthetic_df(n, nO, naux, seed=7)
    out = {}
    with gw.GWEngine(n, n, nO) as g:
        g.set_eri_df(B)
        for mode in ("syevd", "sytrd"):
            g.set_eigensolver(mode)
            out[mode] = g.g0w0(e, doSRG=True)
        assert g.eig_stats()["own_tridiagonal"] == 2
    a, b = out["sytrd"], out["syevd"]
    assert np.abs(a["Om"] - b["Om"]).max() < 1e-10
    assert np.abs(a["eGW"] - b["eGW"]).max() < 1e-9
    assert abs(a["EcRPA"] - b["EcRPA"]) < 1e-10 and abs(a["EcGM"] - b["EcGM"]) < 1e-10


def test_graph_replay_matches_eager(systems):
    """The own full solve replayed from captured graphs equals eager launches bit
    for bit, replay after replay: four screenings alternate the handle's two
    matrix buffers, with eager self-energy sweeps and a G0W0 in between."""
    from quack_amd import gw
    say("graph replay vs eager, H10")
    s = systems["H10"]
    rng = np.random.default_rng(11)
    es = [s["eHF"] + 1e-3 * rng.standard_normal(s["eHF"].size) for _ in range(4)]

    def run(graphs):
        gw.set_graphs(graphs)
        out = []
        with gw.GWEngine(s["nBas"], s["nOrb"], s["nO"]) as g:
            g.set_eri_ao(s["eri"], s["c"])
            g.set_eigensolver("sytrd")
            for ek in es:
                g.screen(ek)
                out.append(g.sigma_at(ek, ek + 0.3, eta=0.05))
            r = g.g0w0(s["eHF"])
            out.append((r["eGW"], r["Om"]))
            assert g.eig_stats()["own_tridiagonal"] == 6
        return out

    try:
        eager = run(False)
        graph = run(True)
    finally:
        gw.set_graphs(False)
    for k, (a, b) in enumerate(zip(eager, graph)):
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), k
    assert np.abs(graph[-1][0] - H10["g0w0_eGW"]).max() < 1e-8
