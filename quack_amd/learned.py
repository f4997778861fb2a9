"""Learned warm start of evGW: a per-orbital MLP trained on QP shifts that the
GPU engine's own plain evGW produces, run before the evGW cycles.

The learned form runs plain evGW's first cycle unchanged (screening at eHF,
self-energy, graphical QP: GW/evRGW.f90:123-154), giving eGW1.  Features of
orbital p (14), from the RHF reference and that cycle's screening:
  eHF_p, diagonal exchange Sigma_x(p), occupied flag,
  Sigma_c and dSigma_c/dw at w = eHF_p + {-0.25, 0, +0.25} Ha (6),
  HOMO-LUMO gap, nO/nOrb, the linearised shift Z(eHF_p) Sigma_c(eHF_p),
  eHF_p - Fermi level, the first-cycle QP shift eGW1_p - eHF_p.
Target: the converged evGW energy minus eGW1_p.  evGW then restarts from
eGW1 + prediction, so the learned form's diagonalisations are 1 + its cycles.
The network (quack_amd/csrc/qk_mlp.hip: 14 -> 128 -> 128 -> 128 -> 1, SiLU,
FP32 weights) is trained on the device by the project's own forward/backward
and Adam kernels and evaluated with the BF16 matrix-core kernel.  evGW then
starts from the predicted energies (qk_evrgw's eGW_start) instead of eHF;
plain evGW (eGW_start = NULL) stays the reference that judges its error.

    python -m quack_amd.learned make-data --out data.npz --nsys 160
    python -m quack_amd.learned train --data data.npz --out quack_amd/models/mlp_evgw_srg.npz
    python -m quack_amd.learned eval --model quack_amd/models/mlp_evgw_srg.npz
"""
import argparse
import json
import os
import sys
import time

import numpy as np

from . import gw
from . import intgen as ig

OFFSETS = (-0.25, 0.0, 0.25)
NFEAT = 14
HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_MODEL = os.path.join(HERE, "models", "mlp_evgw_srg.npz")


def first_cycle(eng, sysd, doSRG=True, eta=0.0):
    """Plain evGW's first cycle (maxSCF = 0 runs exactly one): eGW1."""
    try:
        r = eng.evgw(sysd["eHF"], maxSCF=0, thresh=0.0, max_diis=1, eta=eta, doSRG=doSRG)
    except gw.NotConverged as ex:
        r = ex.result
    return r["eGW"]


def features(eng, sysd, eGW1, doSRG=True, eta=0.0):
    """(n x 14) features from the screening left by the first cycle (at eHF)."""
    e = np.asarray(sysd["eHF"], np.float64)
    n, nO = len(e), sysd["nO"]
    sig, dsig = [], []
    for off in OFFSETS:
        s, d = eng.sigma_at(e, e + off, eta=eta, doSRG=doSRG)
        sig.append(s)
        dsig.append(d)
    Z0 = 1.0 / (1.0 - dsig[1])
    lin = Z0 * sig[1]
    gap = e[nO] - e[nO - 1]
    mu = 0.5 * (e[nO] + e[nO - 1])
    X = np.stack([e, sysd["Sx"], (np.arange(n) < nO).astype(float), sig[0], sig[1], sig[2],
                  dsig[0], dsig[1], dsig[2], np.full(n, gap), np.full(n, nO / n), lin, e - mu, eGW1 - e], axis=1)
    return X


def random_system(rng):
    """Seeded H chain or planar lattice, cc-pVDZ (N = 20..100)."""
    if rng.random() < 0.5:
        nat = int(rng.choice([4, 6, 8, 10, 12, 14, 16, 18, 20]))
        return ig.h_chain(nat, spacing=rng.uniform(1.3, 1.9), jitter=rng.uniform(0.0, 0.15),
                          seed=int(rng.integers(1 << 30)))
    nx, ny = [(2, 2), (2, 3), (2, 4), (3, 4), (2, 5), (2, 6), (4, 4), (3, 6), (4, 5)][int(rng.integers(9))]
    return ig.h_lattice(nx, ny, spacing=rng.uniform(1.5, 1.9), jitter=rng.uniform(0.0, 0.15),
                        seed=int(rng.integers(1 << 30)))


def plain_and_features(mol, doSRG=True, thresh=1e-6):
    s = ig.system(mol)
    with gw.GWEngine(s["nBas"], s["nOrb"], s["nO"]) as eng:
        eng.set_eri_ao(s["eri"], s["c"])
        e1 = first_cycle(eng, s, doSRG=doSRG)
        X = features(eng, s, e1, doSRG=doSRG)
        r = eng.evgw(s["eHF"], maxSCF=128, thresh=thresh, max_diis=1, doSRG=doSRG)
    y = r["eGW"] - e1
    return s, X, y, r


def make_data(out, nsys, seed=0, doSRG=True):
    rng = np.random.default_rng(seed)
    Xs, ys, meta = [], [], []
    t0 = time.time()
    while len(meta) < nsys:
        mol = random_system(rng)
        try:
            s, X, y, r = plain_and_features(mol, doSRG=doSRG)
        except (gw.NotConverged, gw.QKError, RuntimeError) as ex:
            print("skip system:", ex, flush=True)
            continue
        Xs.append(X)
        ys.append(y)
        meta.append((s["nOrb"], r["nSCF"]))
        if len(meta) % 10 == 0:
            print("data: %d systems, %d orbitals, %.0f s" % (len(meta), sum(m[0] for m in meta), time.time() - t0),
                  flush=True)
    np.savez(out, X=np.concatenate(Xs), y=np.concatenate(ys), meta=np.array(meta))
    return out


def train(data, out, epochs=400, width=128, depth=3, lr=1e-3, batch=512, seed=0):
    d = np.load(data)
    X, y = d["X"].astype(np.float64), d["y"].astype(np.float64)
    mean, std = X.mean(0), X.std(0) + 1e-12
    ys = y.std() + 1e-12
    Xn = ((X - mean) / std).astype(np.float32)
    yn = (y / ys).astype(np.float32)
    rng = np.random.default_rng(seed)
    perm = rng.permutation(len(y))
    nval = max(1, len(y) // 10)
    va, tr = perm[:nval], perm[nval:]
    net = gw.MLP(nin=NFEAT, width=width, depth=depth, seed=seed)
    t0 = time.time()
    hist = net.train(Xn[tr], yn[tr], epochs=epochs, batch=batch, lr=lr, seed=seed + 1)
    t_train = time.time() - t0
    pv = net.forward(Xn[va], bf16=True) * ys
    stats = {"train_loss_first": float(hist[0]), "train_loss_last": float(hist[-1]), "epochs": epochs,
             "n_train": int(len(tr)), "n_val": int(nval), "val_mae_Ha": float(np.mean(np.abs(pv - y[va]))),
             "val_max_Ha": float(np.max(np.abs(pv - y[va]))), "val_mae_lin_only_Ha": float(np.mean(np.abs(y[va]))),
             "train_seconds": t_train}
    os.makedirs(os.path.dirname(os.path.abspath(out)), exist_ok=True)
    np.savez(out, params=net.get_params(), mean=mean, std=std, yscale=ys, width=width, depth=depth,
             stats=json.dumps(stats))
    return stats


class WarmStart:
    """Loaded model: predicts evGW starting energies for a molecule."""

    def __init__(self, path=DEFAULT_MODEL):
        d = np.load(path)
        self.mean, self.std, self.ys = d["mean"], d["std"], float(d["yscale"])
        self.net = gw.MLP(nin=NFEAT, width=int(d["width"]), depth=int(d["depth"]), seed=0)
        self.net.set_params(d["params"])

    def predict(self, eng, sysd, doSRG=True, zmin=0.5, cap=0.1):
        """eGW1 + learned correction.  The correction is applied only where the
        first-cycle renormalisation Z(eHF) exceeds zmin (a quasiparticle-like
        orbital, a single well-separated root) and is capped at `cap` Ha;
        elsewhere the start is eGW1, where plain evGW's second cycle starts."""
        e1 = first_cycle(eng, sysd, doSRG=doSRG)
        X = features(eng, sysd, e1, doSRG=doSRG)
        r = self.net.forward(((X - self.mean) / self.std).astype(np.float32), bf16=True).astype(np.float64)
        r *= self.ys
        Z0 = 1.0 / (1.0 - X[:, 7])
        ok = (Z0 > zmin) & (np.abs(r) < cap)
        return e1 + np.where(ok, r, 0.0)

    def evgw(self, eng, sysd, doSRG=True, thresh=1e-6, maxSCF=128):
        start = self.predict(eng, sysd, doSRG=doSRG)
        r = eng.evgw(sysd["eHF"], maxSCF=maxSCF, thresh=thresh, max_diis=1, doSRG=doSRG, eGW_start=start)
        r["start"] = start
        return r


def evaluate(model, nsys=8, seed=12345, doSRG=True):
    ws = WarmStart(model)
    rng = np.random.default_rng(seed)
    rows = []
    for k in range(nsys):
        mol = random_system(rng)
        s = ig.system(mol)
        with gw.GWEngine(s["nBas"], s["nOrb"], s["nO"]) as eng:
            eng.set_eri_ao(s["eri"], s["c"])
            t = time.time()
            ref = eng.evgw(s["eHF"], maxSCF=128, thresh=1e-6, max_diis=1, doSRG=doSRG)
            t_plain = time.time() - t
            t = time.time()
            r = ws.evgw(eng, s, doSRG=doSRG)
            t_warm = time.time() - t
        d = np.abs(r["eGW"] - ref["eGW"]) * 27211.386
        e, nO = s["eHF"], s["nO"]
        win = np.abs(e - 0.5 * (e[nO - 1] + e[nO])) < 1.0
        rows.append({"nOrb": s["nOrb"], "plain_cycles": ref["nSCF"], "warm_cycles": r["nSCF"] + 1,
                     "max_err_meV": float(d.max()), "max_err_meV_1Ha_window": float(d[win].max()),
                     "homo_lumo_err_meV": float(max(d[nO - 1], d[nO])), "n_over_1meV": int((d > 1.0).sum()),
                     "start_err_meV": float(np.abs(r["start"] - ref["eGW"]).max() * 27211.386),
                     "t_plain": t_plain, "t_warm": t_warm})
        print(json.dumps(rows[-1]), flush=True)
    return rows


def evaluate_device(model, nsys=32, seed=12345, doSRG=True, log=None):
    """Held-out evaluation with the inputs produced on the device (integrals,
    RHF Fock builds): for each seeded system, plain evGW (from eHF) and the
    learned form (first cycle, prediction, evGW from the prediction), both on
    the same engine.  Returns per-system rows (cycles, wall times, errors
    against plain evGW in meV)."""
    import torch

    from . import scf
    ws = WarmStart(model)
    rng = np.random.default_rng(seed)
    rows = []
    while len(rows) < nsys:
        mol = random_system(rng)
        S, T, V = mol.one_electron()
        eri_t = torch.empty(mol.nBas ** 4, dtype=torch.float64, device="cuda")
        mol.eri_device(eri_t.data_ptr(), eri_t.numel())
        jk = scf.jk_closure(mol.nBas, eri_t.data_ptr())
        try:
            s = ig.rhf(S, T, V, None, mol.nO, maxit=600, jk=jk)
        except gw.QKError:
            raise
        except RuntimeError:
            continue  # RHF not converged: drawn again
        finally:
            jk.engine.close()
        s["nO"] = mol.nO
        c_t = torch.from_numpy(np.asfortranarray(s["c"]).ravel(order="F").copy()).to("cuda")
        with gw.GWEngine(mol.nBas, s["nOrb"], mol.nO) as eng:
            eng.set_eri_ao_device(eri_t.data_ptr(), c_t.data_ptr())
            try:
                t = time.time()
                ref = eng.evgw(s["eHF"], maxSCF=128, thresh=1e-6, max_diis=1, doSRG=doSRG)
                t_plain = time.time() - t
                t = time.time()
                r = ws.evgw(eng, s, doSRG=doSRG)
                t_warm = time.time() - t
            except gw.NotConverged:
                continue
        d = np.abs(r["eGW"] - ref["eGW"]) * 27211.386
        e, nO = s["eHF"], mol.nO
        win = np.abs(e - 0.5 * (e[nO - 1] + e[nO])) < 1.0
        rows.append({"nBas": mol.nBas, "nOrb": s["nOrb"], "plain_cycles": ref["nSCF"], "warm_cycles": r["nSCF"] + 1,
                     "max_err_meV": float(d.max()), "max_err_meV_1Ha_window": float(d[win].max()),
                     "homo_lumo_err_meV": float(max(d[nO - 1], d[nO])), "n_over_1meV": int((d > 1.0).sum()),
                     "n_orbitals": int(len(d)), "t_plain": t_plain, "t_warm": t_warm})
        if log:
            log(json.dumps(rows[-1]))
        del eri_t, c_t
    return rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("cmd", choices=["make-data", "train", "eval"])
    ap.add_argument("--out", default="data.npz")
    ap.add_argument("--data", default="data.npz")
    ap.add_argument("--model", default=DEFAULT_MODEL)
    ap.add_argument("--nsys", type=int, default=160)
    ap.add_argument("--epochs", type=int, default=400)
    ap.add_argument("--seed", type=int, default=0)
    a = ap.parse_args()
    if a.cmd == "make-data":
        make_data(a.out, a.nsys, a.seed)
    elif a.cmd == "train":
        print(json.dumps(train(a.data, a.out, epochs=a.epochs, seed=a.seed)), flush=True)
    else:
        evaluate(a.model, a.nsys, seed=12345 + a.seed)


if __name__ == "__main__":
    sys.exit(main())
