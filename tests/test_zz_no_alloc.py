"""GPU: the allocation-free QP solve (qk_rgw_qp_graph_noalloc) against the
existing entry point qk_rgw_qp_graph on the same inputs: two calls in a row on
one GPU (H10 chain, then the smaller H2O) through the C-ABI with device
pointers on torch's stream, each held to the existing entry point's result
(the same kernels in the same order: equal to rounding, bar 1e-12 Ha), and an
input larger than the static scratch buffer rejected with QK_E_ARG."""
import ctypes

import numpy as np
import pytest

from oracle import gw_oracle as go

pytestmark = pytest.mark.gpu


def _case(gw, s, doSRG, eta):
    import torch
    e, nO, n = s["eHF"], s["nO"], len(s["eHF"])
    Ec, Om, XpY, XmY = go.screening(e, s["eri_mo"], nO)
    rho = go.excitation_density(s["eri_mo"], XpY, n, nO)
    Sig, Z0, _ = go.srg_self_energy_diag(e, Om, rho, nO) if doSRG else go.self_energy_diag(e, Om, rho, nO, eta)
    lin = e + Z0 * Sig
    ref_e, ref_z = gw.RGW_QP_graph(e, Om, rho, lin, e, Z0, nO, eta=eta, doSRG=doSRG)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()
    nS = len(Om)
    # rhoT[m + nS (p + n q)] = rho[p, q, m]: C order (q, p, m)
    rhoT = dev(np.transpose(rho, (1, 0, 2)))
    d_e, d_om, d_lin = dev(e), dev(Om), dev(lin)
    d_out = torch.zeros(n, dtype=torch.float64, device="cuda")
    d_z = dev(Z0)
    stream = torch.cuda.current_stream().cuda_stream
    rc = gw.load().qk_rgw_qp_graph_noalloc(int(doSRG), eta, 500.0, n, 0, nO, n - nO, 0, nS, d_e.data_ptr(),
                                           d_om.data_ptr(), rhoT.data_ptr(), d_lin.data_ptr(), d_e.data_ptr(),
                                           d_out.data_ptr(), d_z.data_ptr(), ctypes.c_void_p(stream))
    assert rc == 0, gw.load().qk_noalloc_last_error()
    torch.cuda.synchronize()
    out_e, out_z = d_out.cpu().numpy(), d_z.cpu().numpy()
    assert np.abs(out_e - ref_e).max() < 1e-12 and np.abs(out_z - ref_z).max() < 1e-12
    return out_e


def test_noalloc_qp_graph_matches_existing_entry(mo_systems):
    from quack_amd import gw
    gw.load()
    print("noalloc qp graph", flush=True)
    for name, doSRG, eta in (("H10", False, 0.05), ("H2O", True, 0.0), ("H10", True, 0.0), ("H2O", False, 0.05)):
        _case(gw, mo_systems[name], doSRG, eta)


def test_noalloc_rejects_oversize():
    import torch

    from quack_amd import gw
    gw.load()
    x = torch.zeros(1, dtype=torch.float64, device="cuda")
    p = x.data_ptr()
    # nOrb = 2000, nS = 1000 * 1000: 3 nOrb ceil(nS / 1024) doubles alone exceed 4 MiB
    rc = gw.load().qk_rgw_qp_graph_noalloc(0, 0.05, 500.0, 2000, 0, 1000, 1000, 0, 1000 * 1000, p, p, p, p, p, p, p,
                                           None)
    assert rc == gw.QK_E_ARG and b"scratch" in gw.load().qk_noalloc_last_error()
