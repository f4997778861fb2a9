"""Seeded synthetic density-fitted systems (the "precision form" workload).

Factors B[P, p, q] for p <= q are drawn from N(0, 1) * exp(-|p - q| / decay)
and mirrored (B_pq = B_qp), so (pq|rs) = sum_P B_pq^P B_rs^P is a Gram matrix
and the RPA matrices are positive definite.  One overall scale makes the
largest (ia|jb) equal kmax (for a Gram matrix that is the largest diagonal
(ia|ia)).  Occupied energies are sorted U[-2, -0.3], virtual energies sorted
U[0.05, 6] Hartree.
"""
import numpy as np


def energies(n, nO, seed):
    rng = np.random.default_rng(seed)
    eo = np.sort(rng.uniform(-2.0, -0.3, nO))
    ev = np.sort(rng.uniform(0.05, 6.0, n - nO))
    return np.concatenate([eo, ev])


def synthetic_df(n, nO, naux, seed=0, decay=40.0, kmax=0.1, device=None):
    """Returns (e, B, scale).  device None: B is a numpy array (naux, n, n) from
    numpy default_rng(seed + 1); otherwise a torch tensor on `device` of shape
    (n, n, naux), element [q, p, P] = B(P, p, q) (P fastest, the engine's
    layout), from torch.Generator(device).manual_seed(seed + 1)."""
    e = energies(n, nO, seed)
    dist = np.abs(np.arange(n)[:, None] - np.arange(n)[None, :])
    if device is None:
        rng = np.random.default_rng(seed + 1)
        G = rng.standard_normal((naux, n, n))
        B = np.triu(G) + np.transpose(np.triu(G, 1), (0, 2, 1))
        B *= np.exp(-dist / decay)[None]
        diag = np.einsum("Pia,Pia->ia", B[:, :nO, nO:], B[:, :nO, nO:])
        scale = np.sqrt(kmax / diag.max())
        return e, B * scale, float(scale)
    import torch
    g = torch.Generator(device=device).manual_seed(seed + 1)
    G = torch.randn((n, n, naux), generator=g, device=device, dtype=torch.float64)
    iu = torch.triu(torch.ones(n, n, device=device, dtype=torch.bool))
    B = torch.where(iu[:, :, None], G, G.transpose(0, 1))  # [q, p, P]: symmetric in p, q
    del G
    B *= torch.exp(-torch.from_numpy(dist).to(device=device, dtype=torch.float64) / decay)[:, :, None]
    ov = B[nO:, :nO, :]  # [a, i, P]
    diag = (ov * ov).sum(dim=2)
    scale = float(torch.sqrt(kmax / diag.max()))
    B *= scale
    return e, B.contiguous(), scale


def eri_mo_from_df(B):
    """Physicist ERI_MO(p,q,r,s) = <pq|rs> = sum_P B[P,p,r] B[P,q,s] (Fortran
    order), from numpy B (naux, n, n)."""
    n = B.shape[1]
    chem = np.einsum("Ppq,Prs->pqrs", B, B)  # (pq|rs)
    return np.asfortranarray(np.transpose(chem, (0, 2, 1, 3)))  # <pr|qs> -> index (p, r, q, s)
