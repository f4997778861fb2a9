"""Learned warm start (quack_amd/learned.py + models/mlp_evgw_eta01.npz,
trained on the GPU by the project's MLP kernels on QP shifts from the batched
plain evGW) against plain evGW, which judges it: every QP energy within 1 meV
of plain evGW (the same method: eta = 0.1, DIIS 5, 1e-6 Ha), and fewer full
RPA diagonalisations than plain evGW needs.  The held-out molecules are drawn
with a seed the training data never used."""
import json

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
MEV = 1e-3 / 27.211386245988


def test_model_file_consistent():
    from quack_amd import learned
    d = np.load(learned.DEFAULT_MODEL)
    assert d["params"].dtype == np.float32 and d["mean"].shape == (learned.NFEAT,)
    width, depth = int(d["width"]), int(d["depth"])
    n = learned.NFEAT * width + width + (depth - 1) * (width * width + width) + width + 1
    assert d["params"].size == n
    assert json.loads(str(d["method"])) == learned.METHOD


def test_heldout_all_orbitals_within_1meV():
    """Eight held-out seeded H chains and planar lattices (N = 20..250 drawn
    log-uniformly): the learned form lands within 1 meV of plain evGW on every
    orbital of every molecule, and needs fewer full RPA diagonalisations in
    total than plain evGW (both with refinement of the previous eigenvectors
    where the energies moved little)."""
    import torch

    from quack_amd import gw, learned
    torch.cuda.init()
    gw.load()
    print("learned held-out", flush=True)
    rows = learned.evaluate_device(learned.DEFAULT_MODEL, nsys=8, seed=20241203,
                                   log=lambda *a: print(*a, flush=True))
    assert len(rows) >= 6
    worst = max(r["max_err_meV"] for r in rows)
    assert worst < 1.0, worst
    assert sum(r["learned_full_solves"] for r in rows) < sum(r["plain_full_solves"] for r in rows)
