import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and libquack_amd.so")


def pytest_collection_finish(session):
    """torch ships its own HIP runtime and math libraries; in a process that
    uses both torch device memory and libquack_amd.so, torch's must be loaded
    first, so that the library binds to the same runtime (loading the library
    first leaves torch without a device).  Done only when GPU tests were
    selected: the CPU tests load the compiled reference, which must not share
    a process with torch."""
    if any(item.get_closest_marker("gpu") for item in session.items):
        try:
            import torch
            torch.cuda.is_available()
        except Exception:
            pass


@pytest.fixture(scope="session")
def systems():
    """Small closed-shell systems with integrals and RHF from the producer."""
    from quack_amd import intgen as ig
    out = {}
    for name, mol in (("He", ig.he_631g()), ("H2O", ig.h2o_feather()), ("H10", ig.h_chain(10, seed=0))):
        out[name] = ig.system(mol)
    # H10: the RHF solution the QuAcK fixture was computed from (the RHF is
    # converged to 1e-10, so a fresh run differs in the 11th digit, which the
    # eta = 0 graphical and linearised solves amplify)
    import numpy as np
    ref = np.load(os.path.join(ROOT, "tests", "golden", "h10_ref.npz"))
    s = out["H10"]
    assert np.abs(s["eHF"] - ref["eHF"]).max() < 1e-8
    s["eHF"], s["c"] = ref["eHF"].copy(), np.asfortranarray(ref["c"])
    return out


@pytest.fixture(scope="session")
def mo_systems(systems):
    """Same systems with MO integrals (oracle restatement of AOtoMO_ERI_RHF)."""
    from oracle import gw_oracle as go
    out = {}
    for k, s in systems.items():
        d = dict(s)
        d["eri_mo"] = go.aotomo_eri_rhf(s["c"], s["eri"])
        out[k] = d
    return out


def _mo4(G, c1, c2, c3, c4):
    """<pq|rs> = sum c1[mu,p] c2[nu,q] c3[la,r] c4[si,s] <mu nu|la si>, indexed [p,q,r,s]."""
    import numpy as np
    t = np.tensordot(G, c4, axes=([3], [0]))   # [mu, nu, la, s]
    t = np.tensordot(t, c3, axes=([2], [0]))   # [mu, nu, s, r]
    t = np.tensordot(t, c2, axes=([1], [0]))   # [mu, s, r, q]
    t = np.tensordot(t, c1, axes=([0], [0]))   # [s, r, q, p]
    return t.transpose(3, 2, 1, 0)


@pytest.fixture(scope="session")
def uhf_h2o(systems):
    """H2O/cc-pVDZ with genuinely different alpha and beta orbitals: beta's
    coefficients are the RHF ones rotated by exp(X), X antisymmetric with a
    seeded occupied-virtual block of scale 0.05, so the three MO blocks
    (aa|aa), (aa|bb), (bb|bb) all differ and (aa|bb) loses the symmetry that
    would hide a swapped alpha/beta index slot.  Returned flat, QuAcK's
    column-major order; ERI_aabb(p,q,r,s) = <p_a q_b|r_a s_b>."""
    import numpy as np
    from scipy.linalg import expm
    s = systems["H2O"]
    c, nO = s["c"], s["nO"]
    N, n = c.shape
    rng = np.random.default_rng(11)
    X = np.zeros((n, n))
    K = 0.05 * rng.standard_normal((nO, n - nO))
    X[:nO, nO:] = K
    X[nO:, :nO] = -K.T
    cb = c @ expm(X)
    G = np.asarray(s["eri"]).reshape((N,) * 4, order="F")
    flat = lambda t: np.asfortranarray(t).reshape(-1, order="F")
    e = s["eHF"]
    return {"aa": flat(_mo4(G, c, c, c, c)), "ab": flat(_mo4(G, c, cb, c, cb)), "bb": flat(_mo4(G, cb, cb, cb, cb)),
            "e2": np.stack([e, e + 0.01 * np.sin(np.arange(n))], axis=1), "nO": nO, "n": n}
