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
