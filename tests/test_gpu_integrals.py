"""GPU: two-electron integrals of s and p shells generated on the device
(qk_int_eri_device) against the host producer (quack_amd/intgen/intgen.c),
which is pinned by the Boys-function, closed-form, symmetry and RHF-anchor
tests of tests/test_intgen.py.  Both follow the same conventions, so the two
arrays agree element by element to rounding; a G0W0 run from the device array
matches the one from the host array."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def device_eri(mol):
    import torch
    n = mol.nBas
    buf = torch.empty(n ** 4, dtype=torch.float64, device="cuda")
    mol.eri_device(buf.data_ptr(), buf.numel())
    torch.cuda.synchronize()
    return buf


@pytest.mark.parametrize("name", ["He", "H10", "H2x3"])
def test_device_eri_matches_host(name):
    from quack_amd import intgen as ig
    mol = {"He": ig.he_631g(), "H10": ig.h_chain(10, seed=0), "H2x3": ig.h_lattice(2, 3, seed=5)}[name]
    say("device ERIs", name, mol.nBas)
    host = mol.eri()
    dev = device_eri(mol).cpu().numpy()
    err = float(np.abs(dev - host).max())
    say("max |device - host| = %.3e (max |eri| %.3f)" % (err, float(np.abs(host).max())))
    assert err < 1e-13


def test_device_eri_g0w0_h10(systems):
    from quack_amd import gw
    from quack_amd import intgen as ig
    s = systems["H10"]
    say("G0W0 from device ERIs, H10")
    buf = device_eri(ig.h_chain(10, seed=0))
    import torch
    c = torch.from_numpy(np.asfortranarray(s["c"]).ravel(order="F").copy()).to("cuda")
    out = []
    with gw.GWEngine(s["nBas"], s["nOrb"], s["nO"]) as g:
        g.set_eri_ao_device(buf.data_ptr(), c.data_ptr())
        out.append(g.g0w0(s["eHF"], doSRG=True))
    with gw.GWEngine(s["nBas"], s["nOrb"], s["nO"]) as g:
        g.set_eri_ao(s["eri"], s["c"])
        out.append(g.g0w0(s["eHF"], doSRG=True))
    assert np.abs(out[0]["eGW"] - out[1]["eGW"]).max() < 1e-10
    assert np.abs(out[0]["Om"] - out[1]["Om"]).max() < 1e-10


def test_device_eri_rejects_d_shells():
    from quack_amd import gw
    from quack_amd import intgen as ig
    mol = ig.h2o_feather()  # oxygen cc-pVDZ has a d shell
    with pytest.raises(gw.QKError):
        mol.eri_device(1, mol.nBas ** 4)  # rejected before any device access
    with pytest.raises(gw.QKError):
        ig.h_chain(4).eri_device(1, 7)  # wrong buffer size
