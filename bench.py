"""Benchmark of the MI355X restricted G0W0/evGW engine.

Workload (default): seeded 2D H50 lattices (5 x 10 grid, 1.6 bohr spacing,
+-0.15 bohr in-plane jitter, numpy default_rng(seed + k) for the k-th
molecule in flight), cc-pVDZ: nBas = 250 basis functions; QuAcK's canonical
orthogonalisation (threshold 1e-6) keeps nOrb = 249, so o = 25, v = 224,
nS = 5600.  Every molecule is the AO->MO transformation of the (pq|jb) block
from its AO integrals resident in HBM, then evGW@RHF with QuAcK's broadened
self-energy (eta = 0.1 Ha, the graphical QP solve with its exact derivative)
and DIIS 5 (QuAcK's default) to 1e-6 Ha.  eta = 0 is QuAcK's default but its
graphical solution of these lattices is chaotic in FP64 itself (a 1e-12 Ha
perturbation of eHF moves QP energies by up to 1.5 eV), and QuAcK's SRG form
has no unique self-consistent solution (orbitals whose Newton solve fails keep
their previous energy, so restarts 2 mHa away land up to 0.8 eV apart).  With
eta = 0.1 evGW restarted from its own second-cycle input or 5 mHa from its
solution returns to that solution within 0.01 meV on every one of 40 seeded
held-out H systems (eta = 0.05 with DIIS 5 does not on 6 of them), which is
what makes the 1 meV comparisons meaningful.

Forms: the plain FP64 form (QuAcK's arithmetic, dsyevd every cycle, one
molecule at a time; median of --baseline-reps runs) is measured first as the
accuracy reference and the baseline rate, then the same FP64 arithmetic with
the project's own tridiagonalisation in every full solve; the timed form
(default) starts evGW from the learned warm start (quack_amd/learned.py: plain
evGW's first cycle, the MLP's prediction of the converged energies, evGW from
there), runs the excitation densities in split precision (two BF16 slices,
three slice products: `--form mixed2`; `--form mixed` keeps three slices and
six products), refines the previous cycle's eigenvectors where it can, does
the remaining full solves through the own tridiagonalisation, and keeps four
molecules in flight (one engine, stream and host thread each).  A step is one
such batch of molecules.  Inputs (AO ERIs, MO coefficients, eHF) are produced
before the timed region by the integral producer and an RHF whose Fock builds
run through the library's HIP kernels on the device.

    python bench.py [--steps K] [--warmup W] [--form fp64|mixed|mixed2] [--eig syevd|refine|sytrd|refine+sytrd]
                    [--concurrent C] [--method evgw|g0w0] [--reg eta|srg|none] [--eta ETA] [--max-diis D]
                    [--learned on|off] [--cpu-baseline auto|off] [--lib-gemm on|off]
                    [--workload h50|df600|batch|learned|ugw|cpu250|bse|qsgw|ubse]

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64 = 78.6       # TFLOP/s, MI355X dense FP64 matrix
PEAK_HBM = 8000.0      # GB/s
PEAK_BF16 = 2500.0     # TFLOP/s, MI355X dense BF16 matrix
# BF16 slice products per element product of the split-precision forms
PRODUCTS = {"mixed": 6, "mixed2": 3}
DTYPE = {"fp64": "f64", "mixed": "bf16x3-slices/f32-acc/f64 (contractions), f64 (eigensolve, QP)",
         "mixed2": "bf16x2-slices/f32-acc/f64 (contractions), f64 (eigensolve, QP)"}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_inputs(nx, ny, seed):
    import numpy as np
    import torch

    from quack_amd import intgen as ig
    mol = ig.h_lattice(nx, ny, seed=seed)
    t = time.time()
    S, T, V = mol.one_electron()
    eri_t = torch.empty(mol.nBas ** 4, dtype=torch.float64, device="cuda")
    mol.eri_device(eri_t.data_ptr(), eri_t.numel())
    torch.cuda.synchronize()
    eri = None  # the AO integrals live only on the device
    log("inputs: AO integrals nBas=%d generated on the device in %.2f s" % (mol.nBas, time.time() - t))
    from quack_amd import scf as qscf
    t = time.time()
    jk = qscf.jk_closure(mol.nBas, eri_t.data_ptr())  # HIP Fock builds over the resident AO integrals
    scf = ig.rhf(S, T, V, eri, mol.nO, jk=jk)
    log("inputs: RHF %d iterations, nOrb=%d, ERHF=%.10f, gap=%.4f Ha in %.2f s (last Fock build %.2f ms)" %
        (scf["niter"], scf["nOrb"], scf["ERHF"], scf["eHF"][mol.nO] - scf["eHF"][mol.nO - 1], time.time() - t,
         jk.engine.last_ms()))
    jk.engine.close()
    c_t = torch.from_numpy(np.asfortranarray(scf["c"]).ravel(order="F").copy()).to("cuda")
    torch.cuda.synchronize()
    return mol, scf, eri, eri_t, c_t


CPU_SAMPLE = r"""
import json, os, sys, time
sys.path.insert(0, sys.argv[1])
method, max_diis, srg, eta = sys.argv[2], int(sys.argv[3]), bool(int(sys.argv[4])), float(sys.argv[5])
from oracle import qkref
from quack_amd import intgen as ig
s = ig.system(ig.h_chain(10, seed=0))
nO = s["nO"]
def one():
    mo = qkref.aotomo(s["c"], s["eri"])
    if method == "evgw":
        return qkref.evrgw(mo, s["eHF"], nO, maxSCF=256, thresh=1e-6, max_diis=max_diis, srg=srg, eta=eta)
    return qkref.rg0w0(mo, s["eHF"], nO, srg=srg, eta=eta)
sys.stdout.flush()
saved = os.dup(1); null = os.open(os.devnull, os.O_WRONLY); os.dup2(null, 1)
r = one()
ts = []
t0 = time.time()
while time.time() - t0 < 10.0 or len(ts) < 5:
    t = time.time(); r = one(); ts.append(time.time() - t)
ts.sort()
os.dup2(saved, 1)
print(json.dumps({"t": ts[len(ts) // 2], "reps": len(ts), "eGW": list(r["eGW"]), "EcRPA": float(r["EcRPA"])}), flush=True)
"""

# QuAcK's compiled path at N = 250 on the host cores: AOtoMO_ERI_RHF and one
# evGW cycle (evRGW.f90:119-183 with maxSCF = 0: phRLR_A/B, phRLR with its two
# dsyev, RGW_excitation_density, RGW_self_energy_diag, RGW_QP_graph), each
# timed once, on the RHF solution the GPU computed (passed in a file).
CPU250_SAMPLE = r"""
import json, os, sys, time
import numpy as np
sys.path.insert(0, sys.argv[1])
from oracle import qkref
from quack_amd import intgen as ig
d = np.load(sys.argv[2])
mol = ig.h_lattice(int(d["nx"]), int(d["ny"]), seed=int(d["seed"]))
t = time.time(); eri = mol.eri(); t_int = time.time() - t
print("cpu250: AO integrals on the host in %.1f s" % t_int, file=sys.stderr, flush=True)
saved = os.dup(1); null = os.open(os.devnull, os.O_WRONLY); os.dup2(null, 1)
t = time.time(); mo = qkref.aotomo(d["c"], eri); t_ao = time.time() - t
del eri
print("cpu250: AOtoMO_ERI_RHF %.1f s" % t_ao, file=sys.stderr, flush=True)
t = time.time()
r = qkref.evrgw(mo, d["eHF"], int(d["nO"]), maxSCF=0, thresh=1e-6, max_diis=int(d["max_diis"]),
                eta=float(d["eta"]), srg=bool(d["srg"]))
t_cyc = time.time() - t
os.dup2(saved, 1)
print(json.dumps({"t_aotomo": t_ao, "t_cycle": t_cyc, "eGW1": list(r["eGW"]), "t_integrals": t_int}), flush=True)
"""


def host_cpu():
    """Physical cores, logical CPUs and model of the host (lscpu)."""
    import subprocess
    info = {"model": None, "physical_cores": None, "logical_cpus": os.cpu_count()}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        kv = dict((ln.split(":", 1)[0].strip(), ln.split(":", 1)[1].strip()) for ln in out.splitlines() if ":" in ln)
        info["model"] = kv.get("Model name")
        if kv.get("Core(s) per socket") and kv.get("Socket(s)"):
            info["physical_cores"] = int(kv["Core(s) per socket"]) * int(kv["Socket(s)"])
    except Exception:
        pass
    return info


def sample_threads():
    """Host threads the reference's MKL/OpenMP use (OMP_NUM_THREADS, as set on the host)."""
    return int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count())


def cpu_baseline(method, max_diis, mk):
    """QuAcK's own routines (oracle/_ref: GW/, LR/, AOtoMO/ compiled from
    source with flang + MKL) on the host cores, on a bounded sample: one
    complete molecule of the smallest member of the workload family, the H10
    chain (cc-pVDZ, N = 50, nS = 225), AO->MO + the same method, median of at
    least 5 runs over ~10 s.  It runs in a child process (the reference's MKL
    must not share a process with torch).  The GPU runs the same molecule
    right after, so the ratio at equal work is reported beside it.  The N = 250
    figure (QuAcK's AOtoMO_ERI_RHF plus one evGW cycle, timed by
    `bench.py --workload cpu250`, extrapolated over the cycles) is read from
    profiles/cpu_reference_h50_lattice.md when present."""
    import subprocess

    import numpy as np

    from oracle import qkref
    from quack_amd import gw
    from quack_amd import intgen as ig
    if not qkref.available():
        return None
    p = subprocess.run([sys.executable, "-c", CPU_SAMPLE, ROOT, method, str(max_diis), str(int(mk["doSRG"])),
                        str(mk["eta"])], capture_output=True, text=True, timeout=300)
    if p.returncode != 0:
        raise RuntimeError("cpu sample failed: " + p.stderr[-2000:])
    ref = json.loads(p.stdout.strip().splitlines()[-1])
    s = ig.system(ig.h_chain(10, seed=0))
    nO = s["nO"]
    with gw.GWEngine(s["nBas"], s["nOrb"], nO) as g:
        def gpu_one():
            g.set_eri_ao(s["eri"], s["c"])
            if method == "evgw":
                return g.evgw(s["eHF"], maxSCF=256, thresh=1e-6, max_diis=max_diis, **mk)
            return g.g0w0(s["eHF"], **mk)
        r = gpu_one()
        ts, t0 = [], time.time()
        while time.time() - t0 < 2.0 or len(ts) < 5:
            t = time.time()
            gpu_one()
            ts.append(time.time() - t)
        t_gpu = float(np.median(ts))
    err = float(np.abs(np.asarray(ref["eGW"]) - r["eGW"]).max())
    log("cpu baseline: H10 %s reference %.3f s/molecule (median), GPU %.4f s/molecule, max|eGW diff| %.2e Ha" %
        (method, ref["t"], t_gpu, err))
    hc = host_cpu()
    out = {"value": 1.0 / ref["t"], "unit": "molecules/s", "cores": sample_threads(), "kind": "reference",
           "host": hc,
           "sample": ("median of %d runs of one H10-chain molecule (cc-pVDZ, N=50, nS=225), same options: QuAcK's "
                      "AOtoMO_ERI_RHF + %s (phRLR_A/B, phRLR, RGW_excitation_density, RGW_self_energy_diag, "
                      "RGW_QP_graph%s), host ERIs" %
                      (ref["reps"], "evRGW loop" if method == "evgw" else "RG0W0 sequence",
                       ", DIIS" if (method == "evgw" and max_diis > 1) else "")),
           "gpu_same_sample": 1.0 / t_gpu, "gpu_over_cpu_same_sample": ref["t"] / t_gpu,
           "max_abs_eGW_diff_same_sample": err}
    n250 = read_cpu250()
    if n250:
        out["n250"] = n250
    return out


def read_cpu250():
    """The recorded N = 250 reference figure (profiles/cpu_reference_h50_lattice.md)."""
    f = os.path.join(ROOT, "profiles", "cpu_reference_h50_lattice.md")
    if not os.path.exists(f):
        return None
    for line in open(f):
        if line.startswith("json: "):
            try:
                return json.loads(line[6:])
            except ValueError:
                return None
    return None


def run_cpu250(args):
    """QuAcK's compiled path on the host cores at N = 250 (the first H50
    lattice): AOtoMO_ERI_RHF and one evGW cycle, each timed once on the RHF
    solution and AO integrals of the GPU inputs, against the plain FP64 GPU
    form of the same molecule; the reference's whole-molecule time is
    extrapolated as AO->MO + cycles x one cycle (labelled so)."""
    import subprocess
    import tempfile

    import numpy as np
    import torch

    from quack_amd import gw, intgen, workloads
    torch.cuda.init()
    gw.load()
    mk = method_kw(args)
    mol = intgen.h_lattice(args.nx, args.ny, seed=args.seed)
    scf, eri_t, c_t = workloads.device_inputs(mol, log=log)
    with gw.GWEngine(mol.nBas, scf["nOrb"], mol.nO) as g:
        g.set_precision("fp64")
        g.set_eigensolver("syevd")
        g.set_eri_ao_device(eri_t.data_ptr(), c_t.data_ptr())
        g.evgw(scf["eHF"], maxSCF=256, thresh=1e-6, max_diis=args.max_diis, **mk)
        ts = []
        for _ in range(3):
            torch.cuda.synchronize()
            t = time.time()
            g.set_eri_ao_device(eri_t.data_ptr(), c_t.data_ptr())
            ref = g.evgw(scf["eHF"], maxSCF=256, thresh=1e-6, max_diis=args.max_diis, **mk)
            ts.append(time.time() - t)
        try:
            r1 = g.evgw(scf["eHF"], maxSCF=0, thresh=0.0, max_diis=args.max_diis, **mk)
        except gw.NotConverged as ex:
            r1 = ex.result
    del eri_t, c_t
    torch.cuda.empty_cache()
    t_gpu = float(np.median(ts))
    path = os.path.join(tempfile.mkdtemp(), "in.npz")
    np.savez(path, c=scf["c"], eHF=scf["eHF"], nO=mol.nO, nx=args.nx, ny=args.ny, seed=args.seed,
             max_diis=args.max_diis, eta=mk["eta"], srg=int(mk["doSRG"]))
    log("cpu250: GPU plain FP64 %.3f s/molecule (%d cycles); reference on the host cores next" % (t_gpu, ref["nSCF"]))
    p = subprocess.Popen([sys.executable, "-u", "-c", CPU250_SAMPLE, ROOT, path], stdout=subprocess.PIPE, text=True)
    outp, _ = p.communicate(timeout=3000)
    if p.returncode != 0:
        raise RuntimeError("reference run failed")
    d = json.loads(outp.strip().splitlines()[-1])
    err1 = float(np.abs(np.asarray(d["eGW1"]) - r1["eGW"]).max())
    cycles = int(ref["nSCF"])
    t_ref = d["t_aotomo"] + cycles * d["t_cycle"]
    out = {"metric": "QuAcK's compiled restricted evGW path at N=250 on the host cores (extrapolated)",
           "value": 1.0 / t_ref, "unit": "molecules/s", "n_gpus": 1, "higher_is_better": True,
           "config": {"workload": "H50 5x10 lattice seed %d cc-pVDZ, nBas %d, nOrb %d, %s evGW DIIS %d" %
                                  (args.seed, mol.nBas, scf["nOrb"],
                                   "SRG" if mk["doSRG"] else "eta=%g" % mk["eta"], args.max_diis)},
           "reference_aotomo_s": d["t_aotomo"], "reference_one_cycle_s": d["t_cycle"], "cycles": cycles,
           "reference_s_per_molecule_extrapolated": t_ref,
           "extrapolation": "AOtoMO_ERI_RHF + cycles x (one evGW cycle), each timed once",
           "first_cycle_max_abs_dE_vs_gpu_Ha": err1, "gpu_plain_fp64_s_per_molecule_median_of_3": t_gpu,
           "gpu_plain_fp64_over_reference": t_ref / t_gpu, "cores": sample_threads(), "host": host_cpu()}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--method", choices=["evgw", "g0w0"], default="evgw")
    ap.add_argument("--nx", type=int, default=5)
    ap.add_argument("--ny", type=int, default=10)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--reg", choices=["eta", "srg", "none"], default="eta",
                    help="eta: QuAcK's broadened self-energy (eta = --eta); srg: QuAcK's SRG regularisation; none: "
                         "eta = 0 as in QuAcK's default options")
    ap.add_argument("--eta", type=float, default=0.1)
    ap.add_argument("--max-diis", type=int, default=5)
    ap.add_argument("--form", choices=["fp64", "mixed", "mixed2"], default="mixed2",
                    help="fp64: QuAcK's arithmetic; mixed: excitation densities from three BF16 slices "
                         "(six products); mixed2: two slices (three products)")
    ap.add_argument("--eig", choices=["syevd", "refine", "sytrd", "refine+sytrd"], default="refine+sytrd",
                    help="syevd: dsyevd every cycle (QuAcK); refine: Newton refinement of the previous eigenvectors, "
                         "dsyevd for the rest; sytrd: full solves through the project's own tridiagonalisation; "
                         "refine+sytrd: refinement, own-tridiagonalisation full solves for the rest")
    ap.add_argument("--de-max", type=float, default=0.03,
                    help="refinement is tried when no orbital energy moved by more than this (Ha) since the stored "
                         "eigenvectors")
    ap.add_argument("--concurrent", type=int, default=4,
                    help="molecules in flight at once (one engine, stream and host thread each, each its own seeded "
                         "lattice); a step is one such batch")
    ap.add_argument("--learned", choices=["on", "off"], default="on",
                    help="on: the timed form starts evGW from the MLP's prediction (quack_amd/learned.py)")
    ap.add_argument("--baseline-reps", type=int, default=5, help="runs of the plain FP64 form (median reported)")
    ap.add_argument("--lib-gemm", choices=["on", "off"], default="off",
                    help="on: rocBLAS's dgemm joins the autotuner's candidates for large plain FP64 products, in "
                         "every form (the plain FP64 baseline included); off keeps every product on the own "
                         "kernels, whose tile choices give bit-identical results")
    ap.add_argument("--cpu-baseline", choices=["auto", "off"], default="auto")
    ap.add_argument("--ugw-atoms", type=int, default=10, help="ugw workload: H chain length (one more alpha electron)")
    ap.add_argument("--wide-atoms", type=int, default=20, help="bse and qsgw workloads: H chain length")
    ap.add_argument("--workload", choices=["h50", "df600", "batch", "learned", "ugw", "cpu250", "bse", "qsgw", "ubse"],
                    default="h50",
                    help="h50: the headline N=250 lattices; df600: the precision-form measurement on the synthetic "
                         "density-fitted system (N=600, o=60, naux=1800), FP64 against split precision; batch: "
                         "batched G0W0 over seeded H chains/lattices of uneven size; learned: the MLP warm start "
                         "against plain evGW on held-out systems; cpu250: QuAcK's compiled path on the host cores "
                         "at N=250 (AO->MO + one evGW cycle); bse: static phBSE@G0W0, ubse: its unrestricted form "
                         "(spin-conserved and spin-flip, closed-shell occupations) and qsgw: one qsGW cycle, each "
                         "against QuAcK's compiled routines on an H chain (--wide-atoms)")
    ap.add_argument("--nmol", type=int, default=64, help="batch and learned workloads: number of molecules")
    ap.add_argument("--kmax", type=int, default=12, help="batch workload: largest H_2k (N = 10k)")
    ap.add_argument("--chunk", type=int, default=256, help="batch workload: most molecules resident at once")
    ap.add_argument("--batch-method", choices=["g0w0", "evgw"], default="g0w0",
                    help="batch workload: SRG-G0W0, or SRG-evGW to 1e-6 Ha with DIIS off, for every molecule")
    ap.add_argument("--chunk-gb", type=float, default=48.0,
                    help="batch workload: most AO-integral bytes (GB) resident at once")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU: the GW engine is single-GPU in this build (molecules shard across GPUs "
              "with no exchange; see the sharding notes).  Use --gpus 1.", file=sys.stderr)
        sys.exit(2)
    if args.workload == "df600":
        return run_df600(args)
    if args.workload == "batch":
        return run_batch(args)
    if args.workload == "ugw":
        return run_ugw(args)
    if args.workload == "learned":
        return run_learned(args)
    if args.workload == "cpu250":
        return run_cpu250(args)
    if args.workload in ("bse", "qsgw", "ubse"):
        return run_widened(args)
    return run_h50(args)


def method_kw(args):
    if args.reg == "srg":
        return {"doSRG": True, "eta": 0.0}
    return {"doSRG": False, "eta": args.eta if args.reg == "eta" else 0.0}


def run_h50(args):
    """The headline: --concurrent seeded H50 lattices (seeds seed .. seed+K-1),
    one engine each, evGW at eta = 0.1 with DIIS 5 by default."""
    K = max(1, args.concurrent)
    # one hardware queue per engine stream: with four (HIP's default), a fourth
    # molecule's stream shares a queue with another (plus torch's) and waits
    # behind it (measured: 4 in flight 0.807 molecules/s with 4 queues, 0.934
    # with 8).  Raised, never lowered, for this process only; the HIP runtime
    # reads it when the GPU is first used below.
    want = min(32, 2 * K)
    if int(os.environ.get("GPU_MAX_HW_QUEUES", "4")) < want:
        os.environ["GPU_MAX_HW_QUEUES"] = str(want)

    from concurrent.futures import ThreadPoolExecutor

    import numpy as np
    import torch

    from quack_amd import gw, intgen, learned, workloads
    torch.cuda.init()  # torch's HIP runtime first, then the library binds to it
    gw.load()
    gw.set_library_gemm(args.lib_gemm == "on")
    mk = method_kw(args)
    evkw = {"maxSCF": 256, "thresh": 1e-6, "max_diis": args.max_diis, **mk}
    mols, inputs = [], []
    for k in range(K):
        mol = intgen.h_lattice(args.nx, args.ny, seed=args.seed + k)
        mols.append(mol)
        inputs.append(workloads.device_inputs(mol, log=log))
    nBas, nO = mols[0].nBas, mols[0].nO
    engs = [gw.GWEngine(nBas, inp[0]["nOrb"], nO) for inp in inputs]
    eng = engs[0]
    nOrb, nS = eng.nOrb, eng.nS
    use_learned = args.learned == "on" and args.method == "evgw" and os.path.exists(learned.DEFAULT_MODEL)
    if use_learned:
        wsm = learned.WarmStart(learned.DEFAULT_MODEL)
        if wsm.method["eta"] != mk["eta"] or wsm.method["doSRG"] != mk["doSRG"] or \
                wsm.method["max_diis"] != args.max_diis:
            log("learned model trained for %s, not this method: learned form off" % wsm.method)
            use_learned = False
        wsm.close()
    wss = [learned.WarmStart(learned.DEFAULT_MODEL) for _ in range(K)] if use_learned else [None] * K

    def step_on(i, learned_on=use_learned):
        g, (scf, eri_t, c_t) = engs[i], inputs[i]
        g.set_eri_ao_device(eri_t.data_ptr(), c_t.data_ptr())
        if args.method == "g0w0":
            return g.g0w0(scf["eHF"], **mk)
        if learned_on:
            return wss[i].evgw(g, {"eHF": scf["eHF"], "Sx": scf["Sx"], "nO": nO})
        return g.evgw(scf["eHF"], **evkw)

    # plain FP64 form (dsyevd every cycle, one molecule at a time): the accuracy
    # reference of every molecule and, as the median of --baseline-reps runs on
    # the first molecule, the baseline rate
    for g in engs:
        g.set_precision("fp64")
        g.set_eigensolver("syevd")
    step_on(0, False)  # warm-up of the FP64 shapes
    refs, t_fp64s = [], []
    for rep in range(max(1, args.baseline_reps)):
        torch.cuda.synchronize()
        t = time.time()
        refs_0 = step_on(0, False)
        t_fp64s.append(time.time() - t)
    refs.append(refs_0)
    for i in range(1, K):
        refs.append(step_on(i, False))
    t_fp64 = float(np.median(t_fp64s))
    log("plain fp64 form: %s s/molecule (median %.3f)%s" % ([round(x, 3) for x in t_fp64s], t_fp64,
                                                            (" nSCF=%s" % [r.get("nSCF") for r in refs])))
    # the same FP64 arithmetic with the project's own tridiagonalisation in every full solve
    eng.set_eigensolver("sytrd")
    t = time.time()
    ref_own = step_on(0, False)
    t_fp64_own = time.time() - t
    log("plain fp64 form, own tridiagonalisation: %.3f s/molecule, max |dE| vs dsyevd %.2e Ha" %
        (t_fp64_own, float(np.abs(ref_own["eGW"] - refs[0]["eGW"]).max())))
    st0 = eng.stage_times()  # the plain-form launches, part of the fp64 kernel's population only

    for g in engs:
        g.set_precision(args.form)
        g.set_eigensolver(args.eig, args.de_max)
    # first call of every GEMM shape in this form: the tile autotuner's trials
    step_on(0, False)
    # one molecule alone, without and with the learned start: per-stage device
    # times, full-solve counts and the dominant kernels' solo launch times
    st_a = eng.stage_times()
    e_a = eng.eig_stats()
    t = time.time()
    r_nolearn = step_on(0, False)
    t_nolearn = time.time() - t
    e_b = eng.eig_stats()
    st_b = eng.stage_times()
    t = time.time()
    r1 = step_on(0)
    t_single = time.time() - t
    st1 = eng.stage_times()
    e_c = eng.eig_stats()
    # per-stage device times of the run without the learned start (stage_times
    # reports the last GW call); for the learned run, the cumulative kernel
    # counters over its two calls (first cycle, then evGW from the prediction)
    solo_stages = {k: st_b[k] for k in ("aotomo", "rpa_build", "eigensolve", "excitation", "sigma_qp", "total")}
    cum = ("excitation_gemm_all_calls", "excitation_gemm_launches_all_calls", "tridiagonalisation_all_calls",
           "tridiagonalisations_all_calls", "stedc_all_calls", "back_transformation_all_calls")
    learned_cum = {k: st1[k] - st_b[k] for k in cum}
    log("%s/%s form, one molecule alone: without the learned start %.3f s (%d cycles, %d full solves); with it "
        "%.3f s (%s cycles, %d full solves); stages %s" %
        (args.form, args.eig, t_nolearn, r_nolearn.get("nSCF", 1), e_b["syevd"] - e_a["syevd"], t_single,
         r1.get("nSCF_total", r1.get("nSCF")), e_c["syevd"] - e_b["syevd"],
         {k: round(v, 1) for k, v in solo_stages.items()}))
    pool = ThreadPoolExecutor(K)
    for w in range(max(0, args.warmup - 1)):
        list(pool.map(step_on, range(K)))
        log("warmup %d done" % (w + 1))
    torch.cuda.synchronize()
    c0 = [g.stage_times() for g in engs]
    ec0 = [g.eig_stats() for g in engs]
    log("timed region starts, monotonic ns %d" % time.monotonic_ns())
    t0 = time.time()
    for k in range(args.steps):
        rs = list(pool.map(step_on, range(K)))
        log("step %d done (%d molecules)" % (k, K))
    torch.cuda.synchronize()
    t1 = time.time()
    log("timed region ends, monotonic ns %d" % time.monotonic_ns())
    c1 = [g.stage_times() for g in engs]
    ec1 = [g.eig_stats() for g in engs]
    pool.shutdown()
    ms_step = (t1 - t0) * 1e3 / args.steps
    value = K * args.steps / (t1 - t0)
    mev = 1e-3 / 27.211386245988
    dEs = [np.abs(x["eGW"] - ref["eGW"]) for x, ref in zip(rs, refs)]
    dE = np.max(dEs, axis=0)
    eHF0 = inputs[0][0]["eHF"]
    window = np.abs(eHF0 - 0.5 * (eHF0[nO - 1] + eHF0[nO])) < 1.0
    full_timed = sum(b["syevd"] - a["syevd"] for a, b in zip(ec0, ec1))
    refined_timed = sum(b["refined"] - a["refined"] for a, b in zip(ec0, ec1))

    # the excitation-density GEMM rho^T = W^T M^T (the largest matrix-core
    # kernel), one launch per screening cycle: HIP events on the engine's own
    # stream.  "solo" = one molecule alone (the kernel's own device time, the
    # roofline figure); "concurrent" = averaged over the timed region, where the
    # launches of the molecules in flight overlap and each takes longer.
    # FP64 form: 2 nOrb^2 nS^2 flop on FP64 MFMA; split forms: six (mixed) or
    # three (mixed2) BF16 slice products, 6 or 3 x 2 nOrb^2 nS^2 flop.
    gemm_t = sum(b["excitation_gemm_all_calls"] - a["excitation_gemm_all_calls"] for a, b in zip(c0, c1))
    gemm_n = sum(b["excitation_gemm_launches_all_calls"] - a["excitation_gemm_launches_all_calls"] for a, b in zip(c0, c1))
    gemm_ms = gemm_t / max(1.0, gemm_n)
    solo_ms = (st1["excitation_gemm_all_calls"] - st_a["excitation_gemm_all_calls"]) / \
        max(1.0, st1["excitation_gemm_launches_all_calls"] - st_a["excitation_gemm_launches_all_calls"])
    if args.form == "fp64":
        flops, peak, kname = 2.0 * nOrb * nOrb * nS * nS, PEAK_FP64, "dgemm_mfma (excitation densities rho = M W)"
    else:
        npr = PRODUCTS[args.form]
        flops, peak, kname = 2.0 * npr * nOrb * nOrb * nS * nS, PEAK_BF16, \
            "k_gemm_bf16x3 (excitation densities, %d BF16 slice products)" % npr
    achieved = flops / (solo_ms * 1e-3) / 1e12
    gemm_roof = {"kernel": kname, "bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                 "frac": achieved / peak, "traffic": read_traffic(args.form), "flops_per_launch": flops,
                 "avg_launch_ms": solo_ms, "timing": "solo: one molecule alone, HIP events on the engine's stream",
                 "concurrent": {"avg_launch_ms": gemm_ms, "launches": int(gemm_n), "molecules_in_flight": K,
                                "achieved": flops / (gemm_ms * 1e-3) / 1e12}}
    # the own Householder tridiagonalisation (two kernels per column plus the
    # panel GEMMs): one "launch" is one tridiagonalisation of the nS x nS RPA
    # matrix.  Algorithmic bytes: the symmetric product of every column reads
    # the trailing upper triangle once, 8 sum_{L<nS} L (L+1) / 2 bytes.
    trd_bytes = 8.0 * sum(L * (L + 1) / 2.0 for L in range(1, nS))
    trd_solo_n = st1["tridiagonalisations_all_calls"] - st_a["tridiagonalisations_all_calls"]
    trd_t = sum(b["tridiagonalisation_all_calls"] - a["tridiagonalisation_all_calls"] for a, b in zip(c0, c1))
    trd_n = sum(b["tridiagonalisations_all_calls"] - a["tridiagonalisations_all_calls"] for a, b in zip(c0, c1))
    trd_roof = None
    if trd_solo_n > 0:
        solo_trd = (st1["tridiagonalisation_all_calls"] - st_a["tridiagonalisation_all_calls"]) / trd_solo_n
        gbs = trd_bytes / (solo_trd * 1e-3) / 1e9
        trd_roof = {"kernel": "own tridiagonalisation (k_symv_tiles + k_finish_update per column, panel GEMMs)",
                    "bound": "hbm", "achieved": gbs, "peak": PEAK_HBM, "unit": "GB/s", "frac": gbs / PEAK_HBM,
                    "traffic": read_trd_traffic(), "bytes_per_launch": trd_bytes, "avg_launch_ms": solo_trd,
                    "timing": "solo: one molecule alone, HIP events on the engine's stream"}
        if trd_n > 0:
            trd_roof["concurrent"] = {"avg_launch_ms": trd_t / trd_n, "launches": int(trd_n),
                                      "molecules_in_flight": K}
    # the dominant kernel by device time of one molecule alone
    dom = trd_roof if (trd_roof is not None and trd_solo_n * trd_roof["avg_launch_ms"] >
                       (st1["excitation_gemm_launches_all_calls"] - st_a["excitation_gemm_launches_all_calls"]) *
                       solo_ms) else gemm_roof

    cpu = None
    if args.cpu_baseline == "auto":
        try:
            cpu = cpu_baseline(args.method, args.max_diis, mk)
        except Exception as e:  # the baseline is reported, never required
            log("cpu baseline failed: %r" % (e,))
    if cpu and cpu.get("n250") and cpu["n250"].get("value"):
        n250 = cpu["n250"]
        same = "eta=%g" % mk["eta"] in n250.get("config", {}).get("workload", "") and not mk["doSRG"]
        # the recorded N = 250 reference counts the same method's cycles only when the methods match
        cpu["n250_headline_over_reference"] = value / n250["value"] if same else None
    tag = "SRG-" if mk["doSRG"] else ("eta=%g " % mk["eta"])
    out = {
        "metric": "G0W0/evGW molecules/sec at N=250, mixed+learned vs FP64 GPU at 1 meV QP error",
        "value": value,
        "unit": "molecules/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPE[args.form],
        "data": "synthetic: %d seeded H50 lattices, integrals from the project's own Gaussian-integral producer" % K,
        "config": {"workload": "H50 5x10 lattices (seeds %d..%d) cc-pVDZ %s%s DIIS %d, %s form, %s eigensolver%s, "
                               "%d molecules in flight" %
                               (args.seed, args.seed + K - 1, tag, args.method, args.max_diis, args.form, args.eig,
                                ", learned warm start" if use_learned else "", K),
                   "nBas": nBas, "nOrb": [e.nOrb for e in engs], "nO": nO, "nS": [e.nS for e in engs],
                   "method": args.method, "thresh": 1e-6 if args.method == "evgw" else None,
                   "max_diis": args.max_diis, "eta": mk["eta"], "srg_flow": 500.0 if mk["doSRG"] else None,
                   "cycles_per_molecule": [x.get("nSCF_total", x.get("nSCF", 1)) for x in rs],
                   "plain_cycles_per_molecule": [x.get("nSCF", 1) for x in refs], "form": args.form,
                   "eigensolver": args.eig, "refine_de_max": args.de_max, "learned": use_learned,
                   "molecules_in_flight": K, "hip_hw_queues": int(os.environ.get("GPU_MAX_HW_QUEUES", "4")),
                   "library_gemm": args.lib_gemm == "on",
                   "full_rpa_solves_per_molecule_timed": full_timed / (K * args.steps),
                   "refined_rpa_solves_per_molecule_timed": refined_timed / (K * args.steps),
                   "step": "%d molecules" % K},
        "plain_fp64_form": {"molecules_per_s": 1.0 / t_fp64, "s_per_molecule_runs": t_fp64s,
                            "statistic": "median of %d runs, one molecule at a time" % len(t_fp64s),
                            "cycles": refs[0].get("nSCF"), "eigensolver": "dsyevd every cycle",
                            "speedup_of_this_form": value * t_fp64},
        "plain_fp64_form_own_tridiagonal": {"molecules_per_s": 1.0 / t_fp64_own,
                                            "eigensolver": "own tridiagonalisation every cycle",
                                            "max_abs_dE_vs_dsyevd_Ha": float(np.abs(ref_own["eGW"] - refs[0]["eGW"]).max()),
                                            "speedup_of_this_form": value * t_fp64_own},
        "this_form_one_molecule_alone": {"s": t_single, "speedup_vs_plain_fp64": t_fp64 / t_single,
                                         "without_learned_start_s": t_nolearn,
                                         "cycles": r1.get("nSCF_total", r1.get("nSCF")),
                                         "cycles_without_learned_start": r_nolearn.get("nSCF"),
                                         "full_rpa_solves": e_c["syevd"] - e_b["syevd"],
                                         "full_rpa_solves_without_learned_start": e_b["syevd"] - e_a["syevd"],
                                         "stages_ms_without_learned_start": solo_stages,
                                         "learned_run_kernel_ms": learned_cum},
        "accuracy_vs_fp64": {"max_abs_dE_meV_all": float(dE.max() / mev),
                             "orbitals_over_1meV": int(sum((d > mev).sum() for d in dEs)),
                             "orbitals": int(sum(len(d) for d in dEs)),
                             "max_abs_dE_meV_within_1Ha_of_gap": float(dE[window].max() / mev),
                             "HOMO_dE_meV": float(dE[nO - 1] / mev), "LUMO_dE_meV": float(dE[nO] / mev),
                             "learned_start_max_err_meV": (float(np.abs(r1["start"] - refs[0]["eGW"]).max() / mev)
                                                           if "start" in r1 else None)},
        "roofline": dom,
        "roofline_excitation_gemm": gemm_roof,
        # the split-precision GEMM shapes the autotuner timed, with the tile group
        # and wave layout it kept and every candidate's time
        "split_gemm_autotune": [l for l in gw.autotune_report() if l.startswith("gemm_bf16x3")],
        "roofline_tridiagonalisation": trd_roof,
        "cpu_baseline": cpu,
        "result": {"HOMO": float(rs[0]["eGW"][nO - 1]), "LUMO": float(rs[0]["eGW"][nO]), "EcRPA": rs[0]["EcRPA"],
                   "nSCF": rs[0].get("nSCF")},
    }
    print(json.dumps(out), flush=True)
    for g in engs:
        g.close()
    for w in wss:
        if w is not None:
            w.close()


def run_df600(args):
    """Precision form on the synthetic density-fitted system (quack_amd/dfgen.py:
    N = 600, o = 60, v = 540, nS = 32400, naux = 1800, decay 40, max (ia|jb) =
    0.1 Ha, seed 0, factors generated on the device): SRG-G0W0 in the FP64 form
    and in the split-precision form from the same factors.  The split form
    forms (ia|jb) = Bov^T Bov, T = Bov W and rho = T^T B from BF16 slices; the
    eigensolve and the QP solve are FP64 in both."""
    import numpy as np
    import torch

    from quack_amd import dfgen, gw
    n, nO, naux = 600, 60, 1800
    t = time.time()
    e, Bt, scale = dfgen.synthetic_df(n, nO, naux, seed=args.seed, device="cuda")
    torch.cuda.synchronize()
    log("df600 inputs: B (%d x %d x %d) on the device in %.1f s, scale %.3e" % (n, n, naux, time.time() - t, scale))
    g = gw.GWEngine(n, n, nO)
    nS = g.nS
    res = {}
    for form in ("fp64", "mixed", "mixed2"):
        g.set_precision(form)
        t = time.time()
        g.set_eri_df_device(Bt.data_ptr(), naux)
        t_k = time.time() - t
        t = time.time()
        r = g.g0w0(e, doSRG=True)
        dt = time.time() - t
        st = g.stage_times()
        res[form] = (r, dt, st, t_k)
        log("df600 %s: (ia|jb) %.2f s, G0W0 %.1f s, stages %s" % (form, t_k, dt, {k: round(v, 1) for k, v in st.items()}))
    mev = 1e-3 / 27.211386245988
    flops_rho = 2.0 * nS * n * n * naux
    forms = {}
    acc = {}
    for form in ("mixed", "mixed2"):
        d = np.abs(res[form][0]["eGW"] - res["fp64"][0]["eGW"])
        dom = np.abs(res[form][0]["Om"] - res["fp64"][0]["Om"])
        acc[form] = {"max_abs_dE_meV_all": float(d.max() / mev), "orbitals_over_1meV": int((d > mev).sum()),
                     "max_abs_dOm_Ha": float(dom.max()), "HOMO_dE_meV": float(d[nO - 1] / mev),
                     "LUMO_dE_meV": float(d[nO] / mev)}
    report = gw.autotune_report()
    for form, peak, mult in (("fp64", PEAK_FP64, 1.0), ("mixed", PEAK_BF16, 6.0), ("mixed2", PEAK_BF16, 3.0)):
        r, dt, st, t_k = res[form]
        # the rho product is the largest GEMM of the DF path; its first call tuned
        # it, so the kept candidate's time is the smallest of the tuner's HIP-event
        # timings of that shape (the call's stage time also holds every trial)
        ms = tuned_gemm_ms(report, form)
        if ms is None:
            ms = st["excitation_gemm"] / max(1.0, st["excitation_gemm_launches"])
        ach = mult * flops_rho / (ms * 1e-3) / 1e12
        forms[form] = {"g0w0_s": dt, "ia_jb_s": t_k, "stages_ms": st, "rho_gemm_ms": ms,
                       "rho_gemm_roofline": {"achieved": ach, "peak": peak, "unit": "TFLOP/s", "frac": ach / peak,
                                             "flops_per_launch": mult * flops_rho}}
    out = {"metric": "precision form: split-precision vs FP64 contractions on the synthetic DF system",
           "value": forms["fp64"]["rho_gemm_ms"] / forms["mixed"]["rho_gemm_ms"],
           "unit": "x (rho GEMM speed-up, mixed over fp64)", "n_gpus": 1, "higher_is_better": True,
           "dtype": "f64 vs bf16x3-slices/f32-acc/f64", "data": "synthetic: seeded DF factors (dfgen)",
           "mixed2_rho_gemm_speedup_vs_fp64": forms["fp64"]["rho_gemm_ms"] / forms["mixed2"]["rho_gemm_ms"],
           "rho_gemm_timing": "the kept autotuner candidate's HIP-event time for the largest GEMM shape of each form",
           "config": {"workload": "synthetic DF N=600 o=60 naux=1800 SRG-G0W0", "nOrb": n, "nO": nO, "nS": nS,
                      "naux": naux, "decay": 40.0, "kmax_Ha": 0.1, "scale": scale, "seed": args.seed},
           "forms": forms,
           "accuracy_mixed_vs_fp64": acc["mixed"], "accuracy_mixed2_vs_fp64": acc["mixed2"]}
    print(json.dumps(out), flush=True)
    g.close()


def tuned_gemm_ms(report, form):
    """Time (ms) of the kept candidate for the largest GEMM shape of a form in
    the autotuner report: split forms ("mixed": three slices, "mixed2": two)
    from the gemm_bf16x3 lines, "fp64" from the dgemm lines (M N K x batch)."""
    import re
    best = None
    for line in report:
        if form == "fp64" and line.startswith("dgemm"):
            m = re.search(r"M=(\d+) N=(\d+) K=(\d+) batch=(\d+)", line)
            size = int(m.group(1)) * int(m.group(2)) * int(m.group(3)) * int(m.group(4)) if m else 0
        elif form != "fp64" and line.startswith("gemm_bf16x3"):
            m = re.search(r"M=(\d+) N=(\d+) kpad=(\d+) out=f\d+ slices=(\d)", line)
            if not m or int(m.group(4)) != (3 if form == "mixed" else 2):
                continue
            size = int(m.group(1)) * int(m.group(2)) * int(m.group(3))
        else:
            continue
        times = [float(x) for x in re.findall(r" ([0-9.]+)(?= |\)$)", line.split("(ms:", 1)[1])]
        if times and (best is None or size > best[0]):
            best = (size, min(times))
    return best[1] if best else None


def run_batch(args):
    """Batched SRG-G0W0 (quack_amd/batch.py) over args.nmol seeded H_2k chains
    and 2 x k lattices, k uniform in [2, args.kmax], cc-pVDZ: the split-precision
    batch against the FP64 batch (accuracy per molecule) and against the plain
    FP64 engine one molecule at a time (rate).  AO integrals, MO coefficients
    and eHF are resident in HBM before each timed region; AO->MO is timed.
    The molecules go through sorted by size, in chunks of at most args.chunk
    molecules and args.chunk_gb GB of AO integrals (those of all 512 molecules
    of the full recipe, ~0.6 TB, do not fit in HBM at once); the rates are
    totals over the chunks' timed regions."""
    import numpy as np
    import torch

    from quack_amd import batch, gw
    gw.load()
    mols = batch.workload(args.nmol, seed=args.seed, kmin=2, kmax=args.kmax)
    # molecules of one size share a batched handle: go through them sorted by
    # size, in chunks whose AO integrals fit the HBM budget
    mols = sorted(mols, key=lambda m: m.nBas)
    chunks, cur, cur_b = [], [], 0.0
    for m in mols:
        b = 8.0 * m.nBas ** 4
        if cur and (cur_b + b > args.chunk_gb * 1e9 or len(cur) >= args.chunk):
            chunks.append(cur)
            cur, cur_b = [], 0.0
        cur.append(m)
        cur_b += b
    if cur:
        chunks.append(cur)
    reps = max(1, args.steps)
    # every molecule: SRG-G0W0, or evGW of the bench's method (eta = 0.1, DIIS 5
    # by default) to 1e-6 Ha, its cycles replayed from hipGraphs (and eagerly,
    # for the comparison)
    evgw = args.batch_method == "evgw"
    if evgw:
        mk = method_kw(args)
        bkw = {"method": "evgw", "thresh": 1e-6, "max_diis": args.max_diis, **mk}
        meth = ("SRG-" if mk["doSRG"] else "eta=%g " % mk["eta"]) + "evGW"
    else:
        bkw = {"doSRG": True, "method": "g0w0"}
        meth = "SRG-G0W0"
    t_prep = t64 = tb = t_single = t_eager = 0.0
    n_graph_identical = 0
    n_done = 0
    d_all, nS_all, nbas_all, groups_all = [], [], [], set()
    n_seen = 0
    for chunk in chunks:
        t = time.time()
        systems = batch.prepare(chunk, log=log)
        n_seen += len(chunk)
        t_prep += time.time() - t
        n_done += len(systems)
        if not systems:
            continue
        run = batch.Runner(systems)
        groups_all |= set(batch.groups(systems))
        run.run("fp64", **bkw)  # warm-up of every group's handle
        ref, t = run.run("fp64", **bkw)
        t64 += t
        if evgw:
            # eager cycles first (the same work without graphs), then the graph-replayed form
            run.set_graphs(False)
            run.run(args.form, **bkw)
            torch.cuda.synchronize()
            t0 = time.time()
            for _ in range(reps):
                res_eager, _ = run.run(args.form, **bkw)
            t_eager += (time.time() - t0) / reps
            run.set_graphs(True)
        run.run(args.form, **bkw)
        torch.cuda.synchronize()
        t0 = time.time()
        for _ in range(reps):
            res, _ = run.run(args.form, **bkw)
        tb += (time.time() - t0) / reps
        if evgw:
            n_graph_identical += sum(all(np.array_equal(np.asarray(a[k]), np.asarray(b_[k])) for k in ("eGW", "nSCF"))
                                     for a, b_ in zip(res, res_eager))
        run.close()
        # plain FP64 engine, one molecule at a time (same work)
        engines = {}
        t0 = time.time()
        for s_ in systems:
            key = (s_["nBas"], s_["nOrb"], s_["nO"])
            if key not in engines:
                engines[key] = gw.GWEngine(*key)
            g = engines[key]
            g.set_eri_ao_device(s_["eri_t"].data_ptr(), s_["c_t"].data_ptr())
            if evgw:
                g.evgw(s_["eHF"], thresh=1e-6, max_diis=args.max_diis, **mk)
            else:
                g.g0w0(s_["eHF"], doSRG=True)
        t_single += time.time() - t0
        for g in engines.values():
            g.close()
        d_all += [np.abs(a["eGW"] - b["eGW"]).max() for a, b in zip(res, ref)]
        nS_all += [s_["nO"] * (s_["nOrb"] - s_["nO"]) for s_ in systems]
        nbas_all += [s_["nBas"] for s_ in systems]
        del systems, run, res, ref
        torch.cuda.empty_cache()
        log("batch: %d/%d molecules done" % (n_seen, len(mols)))
    log("batch inputs: %d of %d molecules prepared in %.1f s (device integrals, RHF)" % (n_done, len(mols), t_prep))
    mev = 1e-3 / 27.211386245988
    d = np.array(d_all)
    out = {"metric": "batched %s molecules/sec (uneven H chains/lattices)" % ("evGW" if evgw else "G0W0"), "value": n_done / tb,
           "unit": "molecules/s", "n_gpus": 1, "steps": reps, "higher_is_better": True,
           "dtype": DTYPE[args.form],
           "data": "synthetic: seeded H chains and 2 x k lattices, integrals from the project's own producer",
           "config": {"workload": "%d seeded H_2k chains/lattices, k in [2, %d], cc-pVDZ, %s, %s batch" %
                                  (n_done, args.kmax, meth, args.form), "nBas_range": [min(nbas_all), max(nbas_all)],
                      "nS_range": [min(nS_all), max(nS_all)], "groups": len(groups_all), "seed": args.seed,
                      "chunks": len(chunks), "drawn": len(mols), "left_out_rhf_unconverged": len(mols) - n_done},
           "fp64_batch_molecules_per_s": n_done / t64,
           "plain_fp64_one_at_a_time_molecules_per_s": n_done / t_single,
           "speedup_vs_plain_fp64_one_at_a_time": t_single / tb,
           "accuracy_vs_fp64_per_molecule": {"max_abs_dE_meV": float(d.max() / mev),
                                             "molecules_over_1meV": int((d > mev).sum())}}
    if evgw:
        out["config"]["graphs"] = "evGW cycles after the first of each call replayed from hipGraphs"
        out["eager_cycles_molecules_per_s"] = n_done / t_eager
        out["graph_speedup_vs_eager"] = t_eager / tb
        out["molecules_graph_bit_identical_to_eager"] = n_graph_identical
    print(json.dumps(out), flush=True)


def run_learned(args):
    """Learned form on held-out systems (seeded H chains and planar lattices,
    N = 20-250, drawn with a seed the training never used): plain evGW and the
    learned warm start on each (the model's method: eta = 0.1, DIIS 5, FP64
    with refined and own full RPA solves), inputs produced on the device.
    Reports full RPA diagonalisations and evGW cycles saved, wall time, and
    the QP errors against plain evGW."""
    import numpy as np
    import torch

    from quack_amd import gw, learned
    torch.cuda.init()  # torch's HIP runtime first, then the library binds to it
    gw.load()
    rows = learned.evaluate_device(learned.DEFAULT_MODEL, nsys=args.nmol, seed=12345 + args.seed, log=log,
                                   precision=args.form, eig=args.eig)
    t_plain = sum(r["t_plain"] for r in rows)
    t_warm = sum(r["t_learned"] for r in rows)
    err = np.array([r["max_err_meV"] for r in rows])
    hl = np.array([r["homo_lumo_err_meV"] for r in rows])
    out = {"metric": "learned warm start: held-out evGW molecules/sec (plain evGW vs MLP warm start)",
           "value": len(rows) / t_warm, "unit": "molecules/s", "n_gpus": 1, "steps": 1, "higher_is_better": True,
           "dtype": "f64 (evGW), f32/bf16 (MLP)", "data": "synthetic: seeded held-out H chains and planar lattices",
           "config": {"workload": "%d held-out seeded H chains/lattices, cc-pVDZ, evGW eta=0.1 DIIS 5 to 1e-6 Ha, "
                                  "%s form, %s eigensolver" % (len(rows), args.form, args.eig),
                      "nBas_range": [min(r["nBas"] for r in rows), max(r["nBas"] for r in rows)],
                      "model": os.path.relpath(learned.DEFAULT_MODEL, ROOT), "seed": 12345 + args.seed},
           "plain_molecules_per_s": len(rows) / t_plain, "speedup_vs_plain": t_plain / t_warm,
           "full_rpa_solves": {"plain": sum(r["plain_full_solves"] for r in rows),
                               "learned": sum(r["learned_full_solves"] for r in rows)},
           "evgw_cycles": {"plain": sum(r["plain_cycles"] for r in rows),
                           "learned_incl_first_cycle": sum(r["learned_cycles"] for r in rows)},
           "accuracy_vs_plain_evgw": {"max_abs_dE_meV_all": float(err.max()),
                                      "molecules_all_orbitals_within_1meV": int((err <= 1.0).sum()),
                                      "max_HOMO_LUMO_dE_meV": float(hl.max()),
                                      "orbitals_over_1meV": int(sum(r["n_over_1meV"] for r in rows)),
                                      "orbitals": int(sum(r["n_orbitals"] for r in rows))},
           "molecules": len(rows)}
    print(json.dumps(out), flush=True)


def read_trd_traffic():
    """HBM bytes per tridiagonalisation (nS = 5600) from the committed rocprofv3
    PMC summary of its kernels (FETCH_SIZE + WRITE_SIZE, copy-kernel scaled)."""
    f = os.path.join(ROOT, "profiles", "pmc_tridiagonalisation_h50_lattice.md")
    if not os.path.exists(f):
        return None
    for line in open(f):
        if line.startswith("| total per tridiagonalisation |"):
            try:
                return float(line.split("|")[-2])
            except ValueError:
                return None
    return None


def read_traffic(form):
    """HBM bytes per launch of the excitation GEMM from the committed rocprofv3
    PMC summary (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel factor):
    the form's last row, which is the kernel layout the bench now runs."""
    f = os.path.join(ROOT, "profiles", "pmc_excitation_gemm_h50_lattice.md")
    if not os.path.exists(f):
        return None
    val = None
    for line in open(f):
        if line.startswith("| %s |" % form):
            try:
                val = float(line.split("|")[-2])
            except ValueError:
                pass
    return val


UGW_SAMPLE = r"""
import sys, time
import numpy as np
sys.path.insert(0, sys.argv[1])
from quack_amd import intgen as ig
from oracle import qkref
s = ig.system(ig.h_chain(int(sys.argv[2]), seed=0))
mo = qkref.aotomo(s["c"], s["eri"])
e2 = np.stack([s["eHF"], s["eHF"]], axis=1)
occ = (s["nO"] + 1, s["nO"])
t = time.time()
r = qkref.evugw(mo, mo, mo, e2, occ, srg=True)
t = time.time() - t
np.savez(sys.argv[3], mo=mo, e2=e2, occ=np.array(occ), eGW=r["eGW"], nSCF=r["nSCF"], t=t)
"""


def run_ugw(args):
    """Unrestricted SRG-evGW (GW/evUGW.f90) on an open-shell H chain (cc-pVDZ,
    RHF orbitals for both spins, one more alpha electron): one complete
    molecule on the GPU (qk_evugw) against QuAcK's own unrestricted routines
    compiled from source (oracle/_ref, flang + MKL on the host cores, in a
    child process without torch) on the same inputs."""
    import subprocess
    import tempfile

    import numpy as np
    import torch

    from quack_amd import gw
    path = os.path.join(tempfile.mkdtemp(), "ugw.npz")
    p = subprocess.run([sys.executable, "-c", UGW_SAMPLE, ROOT, str(args.ugw_atoms), path], capture_output=True,
                       text=True, timeout=900)
    if p.returncode != 0:
        raise RuntimeError("reference run failed: " + p.stderr[-2000:])
    d = np.load(path)
    mo, e2, occ = d["mo"], d["e2"], tuple(int(x) for x in d["occ"])
    n = e2.shape[0]
    torch.cuda.init()

    def timed(eig):
        gw.set_unrestricted_eigensolver(eig, args.de_max)
        for _ in range(max(args.warmup, 1)):
            gw.evUGW(mo, mo, mo, e2, occ, doSRG=True, maxSCF=1)
        ts = []
        for _ in range(max(3, args.steps)):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            g = gw.evUGW(mo, mo, mo, e2, occ, doSRG=True)
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        st = gw.unrestricted_eig_stats()
        gw.set_unrestricted_eigensolver("syevd")
        return float(np.median(ts)), g, st

    t0, g0, st0 = timed("syevd")
    t, g, st = timed(args.eig)
    log("ugw: dsyevd sequence %.3f s, %s %.3f s (full solves %d, refined %d), reference %.3f s" % (
        t0, args.eig, t, st["syevd"], st["refined"], float(d["t"])))
    print(json.dumps({
        "metric": "unrestricted SRG-evGW molecules per second", "value": 1.0 / t, "unit": "mol/s", "n_gpus": 1,
        "steps": max(3, args.steps), "warmup": args.warmup, "ms_per_step": 1e3 * t,
        "statistic": "median of %d runs" % max(3, args.steps), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "ugw: H%d chain cc-pVDZ, nBas %d, nO %s, nSa + nSb %d" % (
            args.ugw_atoms, n, list(occ), occ[0] * (n - occ[0]) + occ[1] * (n - occ[1])),
            "eigensolver": args.eig, "refine_de_max": args.de_max},
        "cycles": int(g["nSCF"]), "cycles_reference": int(d["nSCF"]),
        "max_abs_dE_vs_reference_Ha": float(np.abs(g["eGW"] - d["eGW"]).max()),
        "full_rpa_solves": int(st["syevd"]), "refined_rpa_solves": int(st["refined"]),
        "dsyevd_sequence": {"molecules_per_s": 1.0 / t0, "cycles": int(g0["nSCF"]),
                            "max_abs_dE_vs_reference_Ha": float(np.abs(g0["eGW"] - d["eGW"]).max())},
        "cpu_baseline": {"value": 1.0 / float(d["t"]), "unit": "mol/s", "cores": sample_threads(), "host": host_cpu(),
                         "kind": "reference", "sample": "the same molecule, QuAcK's compiled unrestricted routines, "
                                                        "one run"},
    }), flush=True)


# QuAcK's compiled routines (oracle/_ref) on an H chain, for the bse and qsgw
# workloads: the static phBSE@G0W0 of GW/RGW_phBSE.f90:113-190 at eta = 0.05
# (W and G from G0W0 at eta = 0.05), or the first qsGW cycle of
# GW/qsRGW.f90:199-341 with SRG from the RHF orbitals; each timed as the median
# of the runs within 10 s (at least 3); the inputs and results go to a file.
WIDE_SAMPLE = r"""
import sys, time
import numpy as np
sys.path.insert(0, sys.argv[1])
from quack_amd import intgen as ig
from oracle import qkref
work, natoms, out = sys.argv[2], int(sys.argv[3]), sys.argv[4]
# 1.8 bohr spacing: no near-dependent functions (nOrb = nBas), the case of
# qsRGW's first branch that the composed reference cycle covers
s = ig.system(ig.h_chain(natoms, spacing=1.8, seed=0))
nO = s["nO"]
mol = s["mol"]
def timed(f):
    r = f()
    ts = []
    t0 = time.time()
    while time.time() - t0 < 10.0 or len(ts) < 3:
        t = time.time(); r = f(); ts.append(time.time() - t)
    ts.sort()
    return r, ts[len(ts) // 2], len(ts)
if work == "bse":
    mo = qkref.aotomo(s["c"], s["eri"])
    eGW = qkref.rg0w0(mo, s["eHF"], nO, eta=0.05)["eGW"]
    r, t, reps = timed(lambda: qkref.phbse(mo, s["eHF"], eGW, nO, eta=0.05))
    np.savez(out, mo=mo, eHF=s["eHF"], eGW=eGW, nO=nO, Ec=r["EcBSE"], OmS=r["OmS"], OmT=r["OmT"], t=t, reps=reps)
elif work == "ubse":
    mo = qkref.aotomo(s["c"], s["eri"])
    e2 = np.stack([s["eHF"], s["eHF"]], axis=1)
    eGW = qkref.ug0w0(mo, mo, mo, e2, (nO, nO), eta=0.05)["eGW"]
    r, t, reps = timed(lambda: qkref.uphbse(mo, mo, mo, e2, eGW, (nO, nO), eta=0.05))
    np.savez(out, mo=mo, e2=e2, eGW=eGW, nO=nO, Ec=r["EcBSE"], OmS=r["OmSC"], OmT=r["OmSF"], t=t, reps=reps)
else:
    S, T, V = mol.one_electron()
    w, U = np.linalg.eigh(S)
    X = U / np.sqrt(w)
    c = np.asarray(s["c"])
    P = 2.0 * c[:, :nO] @ c[:, :nO].T
    r, t, reps = timed(lambda: qkref.qsgw_cycle(S, T, V, X, c, P, s["eHF"], s["eri"], nO, srg=True))
    np.savez(out, S=S, T=T, V=V, X=X, c=c, eHF=s["eHF"], eri=s["eri"], nO=nO, eGW=r["eGW"], t=t, reps=reps)
"""


def run_widened(args):
    """The rows beside the path, each on an H chain of --wide-atoms atoms
    (1.8 bohr spacing, cc-pVDZ), against QuAcK's own routines compiled from source (oracle/_ref,
    flang + MKL on the host cores, in a child process without torch) on the
    same inputs:
      bse:  static phBSE@G0W0 (qk_rgw_phbse: screening, kernels, singlet and
            triplet BSE) from the MO integrals in host memory, as the Fortran
            drop-in passes them;
      qsgw: the first qsGW cycle (qk_qsgw_cycle with SRG) from the RHF
            orbitals, the AO integrals uploaded beforehand (qk_qsgw_create,
            outside the timed call)."""
    import subprocess
    import tempfile

    import numpy as np
    import torch

    from quack_amd import gw, qsgw
    path = os.path.join(tempfile.mkdtemp(), "wide.npz")
    p = subprocess.run([sys.executable, "-c", WIDE_SAMPLE, ROOT, args.workload, str(args.wide_atoms), path],
                       capture_output=True, text=True, timeout=900)
    if p.returncode != 0:
        raise RuntimeError("reference run failed: " + p.stderr[-2000:])
    d = np.load(path)
    nO = int(d["nO"])
    torch.cuda.init()
    reps = max(5, args.steps)
    ts = []
    if args.workload == "bse":
        mo, eHF, eGW = d["mo"], d["eHF"], d["eGW"]
        n = len(eHF)
        for k in range(max(args.warmup, 1) + reps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            r = gw.RGW_phBSE(mo, eHF, eGW, nO, eta=0.05)
            torch.cuda.synchronize()
            if k >= max(args.warmup, 1):
                ts.append(time.perf_counter() - t0)
        err = max(float(np.abs(r["OmS"] - d["OmS"]).max()), float(np.abs(r["OmT"] - d["OmT"]).max()),
                  float(np.abs(r["EcBSE"] - d["Ec"]).max()))
        metric, what = "static phBSE@G0W0 calculations per second", "phBSE"
        extra = {"max_abs_dOm_dEc_vs_reference_Ha": err}
    elif args.workload == "ubse":
        mo, e2, eGW = d["mo"], d["e2"], d["eGW"]
        n = e2.shape[0]
        for k in range(max(args.warmup, 1) + reps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            r = gw.UGW_phBSE(mo, mo, mo, e2, eGW, (nO, nO), eta=0.05)
            torch.cuda.synchronize()
            if k >= max(args.warmup, 1):
                ts.append(time.perf_counter() - t0)
        err = max(float(np.abs(r["OmSC"] - d["OmS"]).max()), float(np.abs(r["OmSF"] - d["OmT"]).max()),
                  float(np.abs(r["EcBSE"] - d["Ec"]).max()))
        metric, what = "unrestricted static phBSE@G0W0 calculations per second", "UGW phBSE"
        extra = {"max_abs_dOm_dEc_vs_reference_Ha": err}
    else:
        S, T, V, X, c, eHF, eri = (d[k] for k in ("S", "T", "V", "X", "c", "eHF", "eri"))
        n = len(eHF)
        for k in range(max(args.warmup, 1) + reps):
            with qsgw.QsGW(S, T, V, X, eri, c, eHF, nO) as q:
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                r = q.cycle(doSRG=True)
                torch.cuda.synchronize()
                if k >= max(args.warmup, 1):
                    ts.append(time.perf_counter() - t0)
        err = float(np.abs(r["eGW"] - d["eGW"]).max())
        metric, what = "qsGW cycles per second", "one qsGW cycle"
        extra = {"max_abs_dE_vs_reference_Ha": err}
    t = float(np.median(ts))
    tr = float(d["t"])
    log("%s: GPU %.4f s, reference %.3f s (%.1fx), max |d| %.1e Ha" % (what, t, tr, tr / t, err))
    print(json.dumps({
        "metric": metric, "value": 1.0 / t, "unit": "1/s", "n_gpus": 1, "steps": reps, "warmup": args.warmup,
        "ms_per_step": 1e3 * t, "statistic": "median of %d runs" % reps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "%s: H%d chain (1.8 bohr spacing, seed 0) cc-pVDZ, nBas %d, nO %d, nS %d%s" % (
            args.workload, args.wide_atoms, n, nO, nO * (n - nO),
            ", SRG, from RHF" if args.workload == "qsgw" else ", eta 0.05, W and G from G0W0 at eta 0.05" +
            (", both spins from the RHF orbitals" if args.workload == "ubse" else ""))},
        **extra,
        "cpu_baseline": {"value": 1.0 / tr, "unit": "1/s", "cores": sample_threads(), "host": host_cpu(),
                         "kind": "reference", "sample": "the same calculation, QuAcK's compiled routines, median of "
                                                        "%d runs" % int(d["reps"])},
        "gpu_over_reference": tr / t,
    }), flush=True)


if __name__ == "__main__":
    main()

