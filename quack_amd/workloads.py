"""Device-resident inputs of the GW path for seeded synthetic systems.

The AO integrals go straight into device memory (the project's s/p Gaussian
producer, qk_int_eri_device); the RHF that gives eHF and the MO coefficients
runs QuAcK's SCF loop (quack_amd.intgen.rhf) with its Fock builds done by the
library's HIP kernels over those resident integrals (qk_rhf_jk).  torch is
plumbing here: it owns the device buffers.
"""
import time

import numpy as np

from . import intgen as ig


def device_inputs(mol, log=None, c=None, eHF=None):
    """(scf, eri_t, c_t): scf is intgen.rhf's dict (eHF, c, nOrb, ERHF, ...);
    eri_t the nBas^4 AO integrals on the device in QuAcK's ERI.bin order;
    c_t the MO coefficients (nBas x nOrb, column-major) on the device.  With c
    and eHF given (a stored RHF solution), the SCF is skipped and those are
    used, so that outputs can be compared with a fixture computed from them."""
    import torch

    t = time.time()
    eri_t = torch.empty(mol.nBas ** 4, dtype=torch.float64, device="cuda")
    mol.eri_device(eri_t.data_ptr(), eri_t.numel())
    torch.cuda.synchronize()
    if log:
        log("inputs: AO integrals nBas=%d generated on the device in %.2f s" % (mol.nBas, time.time() - t))
    if c is None:
        from . import scf as qscf
        S, T, V = mol.one_electron()
        t = time.time()
        jk = qscf.jk_closure(mol.nBas, eri_t.data_ptr())
        try:
            scf = ig.rhf(S, T, V, None, mol.nO, jk=jk)
        finally:
            jk.engine.close()
        if log:
            log("inputs: RHF %d iterations, nOrb=%d, ERHF=%.10f, gap=%.4f Ha in %.2f s" %
                (scf["niter"], scf["nOrb"], scf["ERHF"], scf["eHF"][mol.nO] - scf["eHF"][mol.nO - 1],
                 time.time() - t))
    else:
        c = np.asfortranarray(c, dtype=np.float64)
        scf = {"c": c, "eHF": np.asarray(eHF, dtype=np.float64), "nOrb": c.shape[1]}
    c_t = torch.from_numpy(np.asfortranarray(scf["c"]).ravel(order="F").copy()).to("cuda")
    torch.cuda.synchronize()
    return scf, eri_t, c_t


def ia_block_torch(eri_t, c, nO):
    """K[ia, jb] = (ia|jb) (rows ia = a + v i, QuAcK's ph order) from the
    device AO integrals by four quarter transforms in torch FP64 matrix
    products: an independent route to the RPA coupling that the engine builds
    with its own kernels (used by tests as a property check)."""
    import torch

    N = c.shape[0]
    n = c.shape[1]
    o, v = nO, n - nO
    C = torch.from_numpy(np.ascontiguousarray(c)).to(eri_t.device)
    Co, Cv = C[:, :o].contiguous(), C[:, o:].contiguous()
    # eri_t[a + N b + N^2 c + N^3 d] = <ab|cd>, viewed C-order as T[d, c, b, a];
    # (ia|jb) = <ij|ab> = sum C[mu,i] C[nu,j] C[la,a] C[si,b] T[si, la, nu, mu]
    X1 = eri_t.view(N * N * N, N) @ Co                                    # [si la nu, i]
    X2 = torch.matmul(Co.T.contiguous(), X1.view(N * N, N, o))            # [si la, j, i]
    del X1
    X3 = torch.matmul(Cv.T.contiguous(), X2.reshape(N, N, o * o))         # [si, a, j i]
    del X2
    X4 = (Cv.T.contiguous() @ X3.reshape(N, v * o * o)).view(v, v, o, o)  # [b, a, j, i]
    return X4.permute(3, 1, 2, 0).reshape(o * v, o * v).contiguous()     # [i a, j b]
