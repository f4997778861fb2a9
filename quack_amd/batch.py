"""Batched G0W0 over molecules of uneven size.

Molecules are grouped by (nBas, nOrb, nO); each group runs through one
qk_gwb handle (quack_amd/gw.py GWBatch): the AO->MO block, the RPA build,
one strided-batched eigensolve, the excitation densities, the self-energy and
the QP Newton sweeps for all molecules of the group before the next stage.

The seeded workload family: H_2k chains (1.40 bohr + U(-0.10, 0.10)) and 2 x k
lattices (1.6 bohr + U(-0.15, 0.15)), cc-pVDZ, k drawn uniformly from
[kmin, kmax] (N = 10k basis functions, nS = 9k^2 before any dropped function).
"""
import time

import numpy as np

from . import gw
from . import intgen as ig


def workload(nmol, seed=0, kmin=2, kmax=16):
    """Seeded list of molecules of the batching workload."""
    rng = np.random.default_rng(seed)
    mols = []
    for _ in range(nmol):
        k = int(rng.integers(kmin, kmax + 1))
        s = int(rng.integers(1 << 30))
        if rng.random() < 0.5 or k < 2:
            mols.append(ig.h_chain(2 * k, seed=s))
        else:
            mols.append(ig.h_lattice(2, k, seed=s))
    return mols


def prepare(mols, log=None):
    """AO integrals (generated on the device for s/p bases such as the hydrogen
    systems here, the host producer otherwise) and RHF (Fock builds by the HIP
    J/K kernel over the resident integrals); the AO integrals and MO
    coefficients stay on the device."""
    import torch

    from . import gw, scf as qscf
    out = []
    engines = {}  # one J/K handle per basis size
    for i, mol in enumerate(mols):
        S, T, V = mol.one_electron()
        eri_t = torch.empty(mol.nBas ** 4, dtype=torch.float64, device="cuda")
        try:
            mol.eri_device(eri_t.data_ptr(), eri_t.numel())
        except gw.QKError as ex:
            if ex.code != gw.QK_E_UNSUPPORTED:
                raise
            eri_t.copy_(torch.from_numpy(mol.eri()))
        if mol.nBas not in engines:
            engines[mol.nBas] = qscf.JK(mol.nBas)
        eng, ptr = engines[mol.nBas], eri_t.data_ptr()

        def jk(_eri, D):  # the producer's convention: F = H + 2 J - K for D = C_occ C_occ^T
            J, K = eng(ptr, D)
            return J, -K
        try:
            scf = ig.rhf(S, T, V, None, mol.nO, maxit=600, jk=jk)
        except gw.QKError:
            raise
        except RuntimeError:
            if log:
                log("molecule %d: RHF did not converge in 600 iterations; left out" % i)
            del eri_t
            continue
        c_t = torch.from_numpy(np.asfortranarray(scf["c"]).ravel(order="F").copy()).to("cuda")
        out.append({"nBas": mol.nBas, "nOrb": scf["nOrb"], "nO": mol.nO, "eHF": scf["eHF"], "eri_t": eri_t,
                    "c_t": c_t})
        if log and (i + 1) % 16 == 0:
            log("prepared %d/%d molecules" % (i + 1, len(mols)))
    for eng in engines.values():
        eng.close()
    return out


def groups(systems):
    """{(nBas, nOrb, nO): [index, ...]}"""
    g = {}
    for i, s in enumerate(systems):
        g.setdefault((s["nBas"], s["nOrb"], s["nO"]), []).append(i)
    return g


class Runner:
    """One GWBatch per size group, with the group's AO integrals and MO
    coefficients stacked contiguously on the device once."""

    def __init__(self, systems):
        import torch
        self.systems = systems
        self.plan = []
        for key, idx in sorted(groups(systems).items()):
            nBas, nOrb, nO = key
            eri = torch.stack([systems[i]["eri_t"].reshape(-1) for i in idx])
            cc = torch.stack([systems[i]["c_t"] for i in idx])
            eHF = np.stack([systems[i]["eHF"] for i in idx])
            self.plan.append((idx, gw.GWBatch(len(idx), nBas, nOrb, nO), eri, cc, eHF))
        torch.cuda.synchronize()

    def run(self, precision="mixed", doSRG=True, method="g0w0", **kw):
        """G0W0 (or evGW, method="evgw") for every molecule (AO->MO
        included); returns per-molecule results in input order and the wall
        time."""
        res = [None] * len(self.systems)
        t = time.time()
        for idx, b, eri, cc, eHF in self.plan:
            b.set_precision(precision)
            b.set_eri_ao_device(eri.data_ptr(), eri.shape[1], cc.data_ptr(), cc.shape[1])
            r = b.evgw(eHF, doSRG=doSRG, **kw) if method == "evgw" else b.g0w0(eHF, doSRG=doSRG, **kw)
            for j, i in enumerate(idx):
                res[i] = {k: (v[j] if np.ndim(v) else v) for k, v in r.items()}
        return res, time.time() - t

    def set_graphs(self, on):
        """hipGraph replay of the evGW cycles of every group's handle."""
        for _, b, _, _, _ in self.plan:
            b.set_graphs(on)

    def close(self):
        for _, b, _, _, _ in self.plan:
            b.close()
        self.plan = []
