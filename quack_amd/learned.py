"""Learned warm start of evGW: a per-orbital MLP trained on the QP shifts that
the batched plain evGW produces, run before the evGW cycles.

Method.  evGW with QuAcK's broadened self-energy (GW/RGW_self_energy_diag.f90
with eta > 0, its exact derivative in the graphical Newton solve of
GW/RGW_QP_graph.f90), eta = 0.1 Ha, DIIS 5 (QuAcK's default,
GW/evRGW.f90:168-173), converged to 1e-6 Ha.  Its self-consistent solution is
unique in the sense that matters here: evGW restarted from its own
second-cycle input, or 5 mHa away from its solution, returns to it for every
orbital (measured on 40 seeded H chains and lattices, N = 30..250).  At
eta = 0.05 with DIIS 5 six of the same 40 are path-dependent (a restart from
the second-cycle input alone, which only drops the DIIS history, lands on
another solution), so no warm start can be held to plain evGW there.  Neither
QuAcK's unbroadened eta = 0 form (chaotic: 1e-12 Ha moves QP energies by eV)
nor its SRG form has a unique solution: the SRG Newton solve uses a
pseudo-derivative (GW/RGW_SRG_Re_dSigC.f90) and orbitals whose solve fails keep
their previous energy (GW/RGW_QP_graph.f90:78-81), so their final values depend
on the path and any warm start moves them.

Flow of the learned form for one molecule:
  1. plain evGW's first cycle at eHF (screening, excitation densities,
     self-energy, graphical QP: GW/evRGW.f90:123-154) gives eGW1;
  2. features of orbital p (14) from the RHF reference and that cycle's
     screening: eHF_p, diagonal exchange Sigma_x(p), occupied flag, Sigma_c
     and dSigma_c/dw at w = eHF_p + {-0.25, 0, +0.25} Ha (6), the HOMO-LUMO
     gap, nO/nOrb, the linearised shift Z(eHF_p) Sigma_c(eHF_p), eHF_p - Fermi
     level, the first-cycle shift eGW1_p - eHF_p;
  3. the MLP (quack_amd/csrc/qk_mlp.hip: 14 -> 128 -> 128 -> 128 -> 1, SiLU,
     FP32 weights, BF16 matrix-core inference) predicts e*_p - eGW1_p;
  4. evGW restarts from eGW1 + prediction (qk_evrgw's eGW_start, which
     replaces the eGW = eHF start of GW/evRGW.f90:110).  Its screenings
     after the first are refined from the previous eigenvectors when the
     energies moved little, so the learned form needs two full RPA
     diagonalisations where plain evGW from eHF needs several.
Plain evGW (eGW_start = NULL) stays the fallback and the judge of its error.

Training data: seeded H chains and planar lattices of 4 to 50 atoms in
cc-pVDZ (N = 20..250), AO integrals and RHF produced on the device, the
targets from the batched plain evGW (qk_gwb_evgw, FP64), the features from a
single engine's first cycle.  Training runs on the device with the library's
own forward, backward and Adam kernels.

    python -m quack_amd.learned make-data --out data.npz --nsys 1024
    python -m quack_amd.learned train --data data.npz --out quack_amd/models/mlp_evgw_eta01.npz
    python -m quack_amd.learned eval --nsys 64
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
DEFAULT_MODEL = os.path.join(HERE, "models", "mlp_evgw_eta01.npz")
# the evGW the model is trained for and judged against
METHOD = {"eta": 0.1, "doSRG": False, "max_diis": 5, "thresh": 1e-6, "maxSCF": 256}
# predictions larger than this (Ha) are not trusted: the start is eGW1 there
CAP = 0.25


def first_cycle(eng, sysd, doSRG=False, eta=0.1):
    """Plain evGW's first cycle (maxSCF = 0 runs exactly one): eGW1."""
    try:
        r = eng.evgw(sysd["eHF"], maxSCF=0, thresh=0.0, max_diis=1, eta=eta, doSRG=doSRG)
    except gw.NotConverged as ex:
        r = ex.result
    return r["eGW"]


def features(eng, sysd, eGW1, doSRG=False, eta=0.1):
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


def random_molecule(rng):
    """Seeded H chain or planar lattice, cc-pVDZ, 4..50 atoms (N = 20..250),
    the atom count drawn log-uniformly."""
    nat = int(round(np.exp(rng.uniform(np.log(4), np.log(50)))))
    seed = int(rng.integers(1 << 30))
    if rng.random() < 0.5:
        nat += nat % 2
        return ig.h_chain(min(nat, 50), spacing=rng.uniform(1.35, 1.9), jitter=rng.uniform(0.0, 0.15), seed=seed)
    shapes = [(nx, ny) for nx in range(2, 6) for ny in range(nx, 11) if 4 <= nx * ny <= 50 and nx * ny % 2 == 0]
    nx, ny = min(shapes, key=lambda s: (abs(s[0] * s[1] - nat), rng.random()))
    return ig.h_lattice(nx, ny, spacing=rng.uniform(1.5, 1.9), jitter=rng.uniform(0.0, 0.15), seed=seed)


def device_systems(mols, log=None, chunk_gb=64.0, rng=None):
    """Yields lists of prepared systems of one basis size whose AO integrals
    sit in one contiguous device tensor (at most chunk_gb GB): each system has
    eHF, c (host), Sx, nOrb, nO, nBas, the tensor and its index there.  RHFs
    that do not converge in 600 iterations are left out.  Chunks come in
    order of size, or in an order shuffled by rng (so that a run cut short
    still covers every size)."""
    import torch

    from . import scf as qscf
    by_size = {}
    for m in mols:
        by_size.setdefault(m.nBas, []).append(m)
    chunks = []
    for nBas, ms in sorted(by_size.items()):
        per = max(1, int(chunk_gb * 1e9 // (8 * nBas ** 4)))
        chunks += [(nBas, ms[k0:k0 + per]) for k0 in range(0, len(ms), per)]
    if rng is not None:
        chunks = [chunks[i] for i in rng.permutation(len(chunks))]
    for nBas, part in chunks:
        eri = torch.empty((len(part), nBas ** 4), dtype=torch.float64, device="cuda")
        out = []
        with qscf.JK(nBas) as jkeng:
            for j, mol in enumerate(part):
                mol.eri_device(eri[j].data_ptr(), nBas ** 4)
                S, T, V = mol.one_electron()
                ptr = eri[j].data_ptr()

                def jk(_eri, D, ptr=ptr):
                    J, K = jkeng(ptr, D)
                    return J, -K
                try:
                    s = ig.rhf(S, T, V, None, mol.nO, maxit=600, jk=jk)
                except gw.QKError:
                    raise
                except RuntimeError:
                    if log:
                        log("RHF not converged: molecule left out (nBas %d)" % nBas)
                    continue
                s.update({"nO": mol.nO, "nBas": nBas, "eri": eri, "slot": j})
                out.append(s)
        yield out
        # this chunk's integrals go back to the device before the next chunk's
        # are allocated (the batched handles allocate outside torch's cache):
        # the systems handed out drop their reference to the tensor
        for s in out:
            s["eri"] = None
        del eri
        torch.cuda.empty_cache()


def plain_targets_batched(systems, **method):
    """Plain evGW (FP64) of every system through batched handles: systems of
    one (nBas, nOrb, nO) whose slots are equally strided share one handle.
    Returns {index in systems: result}."""
    import torch
    res = {}
    groups = {}
    for i, s in enumerate(systems):
        groups.setdefault((s["nBas"], s["nOrb"], s["nO"]), []).append(i)
    for (nBas, nOrb, nO), idx in groups.items():
        slots = [systems[i]["slot"] for i in idx]
        runs = [idx] if all(b - a == slots[1] - slots[0] for a, b in zip(slots, slots[1:])) and len(idx) > 1 \
            else [[i] for i in idx]
        for run in runs:
            s0 = systems[run[0]]
            step = (systems[run[1]]["slot"] - s0["slot"]) if len(run) > 1 else 1
            c = torch.from_numpy(np.stack([np.asfortranarray(systems[i]["c"]).ravel(order="F") for i in run])).cuda()
            with gw.GWBatch(len(run), nBas, nOrb, nO) as b:
                b.set_precision("fp64")
                b.set_eri_ao_device(s0["eri"][s0["slot"]].data_ptr(), step * nBas ** 4, c.data_ptr(), c.shape[1])
                try:
                    r = b.evgw(np.stack([systems[i]["eHF"] for i in run]), **method)
                except gw.NotConverged as ex:
                    r = ex.result
                    r["failed"] = True
            for j, i in enumerate(run):
                res[i] = {"eGW": r["eGW"][j], "nSCF": int(r["nSCF"][j]),
                          "converged": not r.get("failed") or int(r["nSCF"][j]) <= method.get("maxSCF", 256)}
    return res


def make_data_device(out, nsys, seed=0, log=None, t_max=1e9):
    """Training data: nsys seeded molecules (fewer if t_max seconds pass),
    targets from the batched plain evGW, features from each molecule's first
    cycle.  Saves X (n_orbitals x 14), y (e* - eGW1) and per-molecule meta;
    rewritten after every chunk, so a cut-off run keeps what it finished."""
    rng = np.random.default_rng(seed)
    mols = [random_molecule(rng) for _ in range(nsys)]
    Xs, ys, meta = [], [], []
    t0 = time.time()
    kw = {k: METHOD[k] for k in ("eta", "doSRG", "max_diis", "thresh", "maxSCF")}
    for systems in device_systems(mols, log=log, rng=rng):
        if not systems:
            continue
        tgt = plain_targets_batched(systems, **kw)
        engines = {}
        for i, s in enumerate(systems):
            if not tgt[i]["converged"]:
                continue
            key = (s["nBas"], s["nOrb"], s["nO"])
            if key not in engines:
                engines[key] = gw.GWEngine(*key)
            eng = engines[key]
            import torch
            c_t = torch.from_numpy(np.asfortranarray(s["c"]).ravel(order="F").copy()).cuda()
            eng.set_eri_ao_device(s["eri"][s["slot"]].data_ptr(), c_t.data_ptr())
            e1 = first_cycle(eng, s, doSRG=kw["doSRG"], eta=kw["eta"])
            X = features(eng, s, e1, doSRG=kw["doSRG"], eta=kw["eta"])
            Xs.append(X)
            ys.append(tgt[i]["eGW"] - e1)
            meta.append((s["nBas"], s["nOrb"], s["nO"], tgt[i]["nSCF"]))
        for e in engines.values():
            e.close()
        np.savez(out, X=np.concatenate(Xs), y=np.concatenate(ys), meta=np.array(meta), method=json.dumps(METHOD))
        if log:
            log("data: %d molecules, %d orbitals, largest nBas %d, %.0f s" %
                (len(meta), sum(m[1] for m in meta), max(m[0] for m in meta), time.time() - t0))
        if time.time() - t0 > t_max:
            break
    return out


def load_data(data):
    """X, y and per-molecule meta of one make-data file, or of several
    (comma-separated paths, each made with its own seed) concatenated."""
    parts = [np.load(p) for p in data.split(",")]
    for d in parts:
        assert json.loads(str(d["method"])) == METHOD, "training data of another method"
    return (np.concatenate([d["X"] for d in parts]), np.concatenate([d["y"] for d in parts]),
            np.concatenate([d["meta"] for d in parts]))


def train(data, out, epochs=600, width=128, depth=3, lr=1e-3, batch=512, seed=0, log=None, every=25):
    X, y, meta = load_data(data)
    X, y = X.astype(np.float64), y.astype(np.float64)
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
    # epochs in segments of `every` (the Adam state stays with the network; each
    # segment shuffles with its own seed), reporting the loss after each
    hist = []
    for k0 in range(0, epochs, every):
        ne = min(every, epochs - k0)
        hist += list(net.train(Xn[tr], yn[tr], epochs=ne, batch=batch, lr=lr, seed=seed + 1 + k0))
        if log:
            log("train: epoch %d/%d, loss %.4g, %.0f s" % (k0 + ne, epochs, hist[-1], time.time() - t0))
    t_train = time.time() - t0
    pv = net.forward(Xn[va], bf16=True) * ys
    err = np.abs(pv - y[va])
    stats = {"train_loss_first": float(hist[0]), "train_loss_last": float(hist[-1]), "epochs": epochs,
             "n_train": int(len(tr)), "n_val": int(nval), "n_molecules": int(len(meta)),
             "largest_nBas": int(meta[:, 0].max()), "val_mae_Ha": float(err.mean()),
             "val_p99_Ha": float(np.quantile(err, 0.99)), "val_max_Ha": float(err.max()),
             "val_mae_first_cycle_Ha": float(np.mean(np.abs(y[va]))), "train_seconds": t_train}
    os.makedirs(os.path.dirname(os.path.abspath(out)), exist_ok=True)
    np.savez(out, params=net.get_params(), mean=mean, std=std, yscale=ys, width=width, depth=depth,
             stats=json.dumps(stats), method=json.dumps(METHOD))
    net.close()
    return stats


class WarmStart:
    """Loaded model: predicts evGW starting energies for a molecule."""

    def __init__(self, path=DEFAULT_MODEL):
        d = np.load(path)
        self.mean, self.std, self.ys = d["mean"], d["std"], float(d["yscale"])
        self.method = json.loads(str(d["method"]))
        self.net = gw.MLP(nin=NFEAT, width=int(d["width"]), depth=int(d["depth"]), seed=0)
        self.net.set_params(d["params"])

    def close(self):
        self.net.close()

    def predict(self, eng, sysd, window=None, cap=CAP, diag=None, zmax=None):
        """eGW1 + the learned correction, applied where |correction| < cap Ha
        and, with `window`, only to orbitals whose eHF lies within `window` Ha
        of the Fermi level, and, with `zmax`, only where the renormalisation
        factor at eHF, Z0 = 1 / (1 - dSigma_c/dw), lies in (0, zmax] (Z0 > 1
        means dSigma_c/dw > 0: the orbital sits on the near side of a
        broadened pole of Sigma_c, where its QP equation can have several
        roots); eGW1 elsewhere.  Also returns eGW1.  diag (a dict), when
        given, receives per-molecule quantities known after the first cycle."""
        m = self.method
        e1 = first_cycle(eng, sysd, doSRG=m["doSRG"], eta=m["eta"])
        nit1 = eng.qp_iters()
        X = features(eng, sysd, e1, doSRG=m["doSRG"], eta=m["eta"])
        r = self.net.forward(((X - self.mean) / self.std).astype(np.float32), bf16=True).astype(np.float64)
        r *= self.ys
        ok = np.abs(r) < cap
        if window is not None:
            ok &= np.abs(X[:, 12]) < window
        Z0 = 1.0 / (1.0 - X[:, 7])
        if zmax is not None:
            ok &= (Z0 > 0.0) & (Z0 <= zmax)
        if diag is not None:
            diag.update({"first_cycle_newton_failures": int((nit1 >= 64).sum()), "first_cycle_max_newton": int(nit1.max()),
                         "max_abs_pred": float(np.abs(r).max()), "min_Z0": float(Z0.min()),
                         "n_Z0_below_half": int((Z0 < 0.5).sum()), "max_first_shift": float(np.abs(e1 - X[:, 0]).max()),
                         "e1_minus_start_max": float(np.abs(np.where(ok, r, 0.0)).max())})
        return e1 + np.where(ok, r, 0.0), e1

    def evgw(self, eng, sysd, window=None, cap=CAP, diag=None, zmax=None):
        """The learned form: first cycle, prediction, evGW from the prediction.
        Returns evGW's result dict with "start", and "nSCF_total" counting the
        first cycle."""
        m = self.method
        start, e1 = self.predict(eng, sysd, window=window, cap=cap, diag=diag, zmax=zmax)
        r = eng.evgw(sysd["eHF"], maxSCF=m["maxSCF"], thresh=m["thresh"], max_diis=m["max_diis"], eta=m["eta"],
                     doSRG=m["doSRG"], eGW_start=start)
        r["start"] = start
        r["nSCF_total"] = r["nSCF"] + 1
        return r


def plain(eng, sysd, method=METHOD):
    return eng.evgw(sysd["eHF"], maxSCF=method["maxSCF"], thresh=method["thresh"], max_diis=method["max_diis"],
                    eta=method["eta"], doSRG=method["doSRG"])


def evaluate_device(model, nsys=64, seed=12345, log=None, precision="fp64", eig="refine+sytrd", de_max=0.03,
                    variants=None):
    """Held-out evaluation (molecules drawn with a seed the training never
    used; inputs produced on the device): per molecule plain evGW from eHF and
    the learned form on the same engine.  Returns per-molecule rows (cycles,
    full diagonalisations, wall times, errors against plain evGW in meV).
    variants: {name: WarmStart.evgw keyword arguments} run in addition to the
    default form on the same molecules, reported as name_cycles,
    name_full_solves, name_max_err_meV, name_n_over_1meV."""
    import torch
    ws = WarmStart(model)
    rng = np.random.default_rng(seed)
    mols = [random_molecule(rng) for _ in range(nsys)]
    rows = []
    for systems in device_systems(mols, log=log):
        for s in systems:
            c_t = torch.from_numpy(np.asfortranarray(s["c"]).ravel(order="F").copy()).cuda()
            with gw.GWEngine(s["nBas"], s["nOrb"], s["nO"]) as eng:
                eng.set_precision(precision)
                eng.set_eigensolver(eig, de_max)
                eng.set_eri_ao_device(s["eri"][s["slot"]].data_ptr(), c_t.data_ptr())
                try:
                    plain(eng, s)  # warm-up of this size's GEMM shapes
                    st0 = eng.eig_stats()
                    t = time.time()
                    ref = plain(eng, s)
                    t_plain = time.time() - t
                    st1 = eng.eig_stats()
                    nitp = eng.qp_iters()
                    t = time.time()
                    dg = {}
                    r = ws.evgw(eng, s, diag=dg)
                    t_warm = time.time() - t
                    st2 = eng.eig_stats()
                    extra = {}
                    for name, kw in (variants or {}).items():
                        sa = eng.eig_stats()
                        rv = ws.evgw(eng, s, **kw)
                        sb = eng.eig_stats()
                        dv = np.abs(rv["eGW"] - ref["eGW"]) * 27211.386
                        extra.update({name + "_cycles": rv["nSCF_total"], name + "_full_solves": sb["syevd"] - sa["syevd"],
                                      name + "_max_err_meV": float(dv.max()), name + "_n_over_1meV": int((dv > 1.0).sum())})
                except gw.NotConverged:
                    continue
            d = np.abs(r["eGW"] - ref["eGW"]) * 27211.386
            e, nO = s["eHF"], s["nO"]
            rows.append({"nBas": s["nBas"], "nOrb": s["nOrb"], "plain_cycles": ref["nSCF"],
                         "learned_cycles": r["nSCF_total"], "plain_full_solves": st1["syevd"] - st0["syevd"],
                         "learned_full_solves": st2["syevd"] - st1["syevd"], "max_err_meV": float(d.max()),
                         "homo_lumo_err_meV": float(max(d[nO - 1], d[nO])), "n_over_1meV": int((d > 1.0).sum()),
                         "n_orbitals": int(len(d)), "start_err_meV": float(np.abs(r["start"] - ref["eGW"]).max() * 27211.386),
                         "t_plain": t_plain, "t_learned": t_warm,
                         "plain_last_cycle_newton_failures": int((nitp >= 64).sum()), **dg, **extra})
            if log:
                log(json.dumps(rows[-1]))
    ws.close()
    return rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("cmd", choices=["make-data", "train", "eval"])
    ap.add_argument("--out", default="data.npz")
    ap.add_argument("--data", default="data.npz", help="train: one make-data file or several, comma-separated")
    ap.add_argument("--model", default=DEFAULT_MODEL)
    ap.add_argument("--nsys", type=int, default=1024)
    ap.add_argument("--epochs", type=int, default=600)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--t-max", type=float, default=1e9, help="make-data: stop drawing chunks after this many seconds")
    ap.add_argument("--zmax", type=float, default=None, help="eval: also run the form gated to 0 < Z0 <= zmax")
    a = ap.parse_args()

    def log(*x):
        print(*x, flush=True)
    if a.cmd in ("make-data", "eval"):
        import torch
        torch.cuda.init()  # torch's HIP runtime first, then the library binds to it
    gw.load()
    if a.cmd == "make-data":
        make_data_device(a.out, a.nsys, a.seed, log=log, t_max=a.t_max)
    elif a.cmd == "train":
        print(json.dumps(train(a.data, a.out, epochs=a.epochs, seed=a.seed, log=log)), flush=True)
    else:
        variants = {"zgate": {"zmax": a.zmax}} if a.zmax is not None else None
        rows = evaluate_device(a.model, a.nsys, seed=12345 + a.seed, log=log, variants=variants)
        summary = {"molecules": len(rows), "max_err_meV": max(r["max_err_meV"] for r in rows),
                   "molecules_over_1meV": sum(r["n_over_1meV"] > 0 for r in rows),
                   "plain_full": sum(r["plain_full_solves"] for r in rows),
                   "learned_full": sum(r["learned_full_solves"] for r in rows),
                   "plain_cycles": sum(r["plain_cycles"] for r in rows),
                   "learned_cycles": sum(r["learned_cycles"] for r in rows)}
        if variants:
            summary.update({"zgate_max_err_meV": max(r["zgate_max_err_meV"] for r in rows),
                            "zgate_molecules_over_1meV": sum(r["zgate_n_over_1meV"] > 0 for r in rows),
                            "zgate_full": sum(r["zgate_full_solves"] for r in rows),
                            "zgate_cycles": sum(r["zgate_cycles"] for r in rows)})
        print(json.dumps(summary), flush=True)


if __name__ == "__main__":
    sys.exit(main())
