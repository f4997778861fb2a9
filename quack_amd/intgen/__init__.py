"""Input producer for the restricted G0W0/evGW path: molecules, integrals, RHF.

QuAcK obtains its integrals from PySCF inside PyDuck.py (PyDuck.py:115-159,
329-345) and its reference state from its own RHF (HF/RHF.f90:100-282).
PySCF is not available here, so this module produces the same quantities with
the McMurchie-Davidson code in ``intgen.c`` and a closed-shell RHF that
follows QuAcK's conventions:

* canonical orthogonalisation with eigenvalue threshold 1e-6
  (utils/orthogonalization_matrix.f90:16-59), nOrb <= nBas;
* orbital energies ascending, MO coefficients ``c(nBas, nOrb)``;
* ERHF is the electronic energy, without ENuc (HF/RHF.f90:305).

Two-electron integrals are returned as the flat array that QuAcK streams
from int/ERI.bin: ``eri[a + N b + N^2 c + N^3 d] = <ab|cd> = (ac|bd)``.
"""
import ctypes
import json
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libqkint.so")
ANGSTROM_TO_BOHR = 1.0 / 0.52917721092  # PySCF's conversion (mol.unit = 'Angstrom')
Z_OF = {"H": 1, "He": 2, "O": 8}

_lib = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libqkint.so is not built: run __graft_entry__.build()")
        lib = ctypes.CDLL(LIB_PATH)
        i_p = ctypes.POINTER(ctypes.c_int)
        d_p = ctypes.POINTER(ctypes.c_double)
        lib.qk_int_eri.restype = ctypes.c_int
        lib.qk_int_eri.argtypes = [ctypes.c_int, i_p, i_p, i_p, d_p, d_p, d_p, d_p]
        lib.qk_int_1e.restype = None
        lib.qk_int_1e.argtypes = [ctypes.c_int, i_p, i_p, i_p, d_p, d_p, ctypes.c_int, d_p, d_p, d_p, d_p, d_p]
        lib.qk_boys.restype = None
        lib.qk_boys.argtypes = [ctypes.c_int, ctypes.c_double, d_p]
        _lib = lib
    return _lib


def _ptr(a, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))


def boys(mmax, T):
    F = np.zeros(mmax + 1)
    _load().qk_boys(mmax, float(T), _ptr(F, ctypes.c_double))
    return F


class Molecule:
    """Nuclei (bohr) plus a basis name; closed shell."""

    def __init__(self, symbols, xyz_bohr, basis="cc-pvdz", charge=0):
        self.symbols = list(symbols)
        self.xyz = np.ascontiguousarray(np.asarray(xyz_bohr, dtype=np.float64).reshape(-1, 3))
        self.Z = np.array([Z_OF[s] for s in self.symbols], dtype=np.float64)
        self.basis = basis
        nel = int(self.Z.sum()) - charge
        if nel % 2:
            raise ValueError("closed-shell molecules only")
        self.nO = nel // 2
        with open(os.path.join(_HERE, "basis_data.json")) as f:
            data = json.load(f)[basis]
        sl, snp, sat, ex, co = [], [], [], [], []
        for ia, s in enumerate(self.symbols):
            for shell in data[s]:
                sl.append(shell["l"])
                snp.append(len(shell["exps"]))
                sat.append(ia)
                ex += shell["exps"]
                co += shell["coefs"]
        self.shell_l = np.array(sl, dtype=np.int32)
        self.shell_np = np.array(snp, dtype=np.int32)
        self.shell_atom = np.array(sat, dtype=np.int32)
        self.exps = np.array(ex, dtype=np.float64)
        self.coefs = np.array(co, dtype=np.float64)
        self.nBas = int(np.sum(2 * self.shell_l + 1))

    def enuc(self):
        e = 0.0
        for i in range(len(self.Z)):
            for j in range(i):
                e += self.Z[i] * self.Z[j] / np.linalg.norm(self.xyz[i] - self.xyz[j])
        return e

    def _args(self):
        i = ctypes.c_int
        return (len(self.shell_l), _ptr(self.shell_l, i), _ptr(self.shell_np, i), _ptr(self.shell_atom, i),
                _ptr(self.exps, ctypes.c_double), _ptr(self.coefs, ctypes.c_double))

    def one_electron(self):
        n = self.nBas
        S, T, V = (np.zeros((n, n)) for _ in range(3))
        d = ctypes.c_double
        _load().qk_int_1e(*self._args(), len(self.Z), _ptr(self.xyz, d), _ptr(self.Z, d),
                          _ptr(S, d), _ptr(T, d), _ptr(V, d))
        return S, T, V

    def eri(self, out=None):
        n = self.nBas
        if out is None:
            out = np.empty(n ** 4, dtype=np.float64)
        if out.size != n ** 4 or out.dtype != np.float64 or not out.flags.c_contiguous:
            raise ValueError("eri buffer must be a contiguous float64 array of nBas^4")
        d = ctypes.c_double
        rc = _load().qk_int_eri(*self._args(), _ptr(self.xyz, d), _ptr(out, d))
        if rc != n:
            raise RuntimeError("qk_int_eri failed")
        return out


    def eri_device(self, ptr, count):
        """The same array generated on the GPU (s and p shells only) into device
        memory at ptr (count = nBas^4 doubles), by libquack_amd's
        qk_int_eri_device; the host producer above stays the reference."""
        from quack_amd import gw
        lib = gw.load()
        d, i = ctypes.c_double, ctypes.c_int
        rc = lib.qk_int_eri_device(len(self.shell_l), _ptr(self.shell_l, i), _ptr(self.shell_np, i),
                                   _ptr(self.shell_atom, i), _ptr(self.exps, d), _ptr(self.coefs, d),
                                   _ptr(self.xyz, d), ctypes.c_void_p(ptr), ctypes.c_longlong(count))
        if rc != 0:
            raise gw.QKError(rc, lib.qk_last_error().decode() if rc != 5 else "s and p shells only")
        return ptr


def h_chain(natom, spacing=1.40, jitter=0.10, seed=0, basis="cc-pvdz"):
    """Linear H chain along z: atom k at spacing*k + U(-jitter, +jitter) bohr."""
    rng = np.random.default_rng(seed)
    z = spacing * np.arange(natom) + rng.uniform(-jitter, jitter, natom)
    xyz = np.zeros((natom, 3))
    xyz[:, 2] = z
    return Molecule(["H"] * natom, xyz, basis)


def h_lattice(nx, ny, spacing=1.60, jitter=0.15, seed=0, basis="cc-pvdz"):
    """Planar nx-by-ny H lattice (z = 0) with in-plane U(-jitter, +jitter) bohr."""
    rng = np.random.default_rng(seed)
    gx, gy = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    xyz = np.zeros((nx * ny, 3))
    xyz[:, 0] = spacing * gx.ravel() + rng.uniform(-jitter, jitter, nx * ny)
    xyz[:, 1] = spacing * gy.ravel() + rng.uniform(-jitter, jitter, nx * ny)
    return Molecule(["H"] * (nx * ny), xyz, basis)


def h2o_feather():
    """H2O of tests/feather_bench.py:62-64 (angstrom), cc-pVDZ."""
    xyz = np.array([[0.0, 0.0, 0.0], [0.7571, 0.0, 0.5861], [-0.7571, 0.0, 0.5861]]) * ANGSTROM_TO_BOHR
    return Molecule(["O", "H", "H"], xyz, "cc-pvdz")


def he_631g():
    return Molecule(["He"], np.zeros((1, 3)), "6-31g")


# --------------------------------------------------------------------------
# Fock builds over the flat physicist array eri[a + N b + N^2 c + N^3 d] = (ac|bd)
# viewed in C order as X[d, c, b, a].
# --------------------------------------------------------------------------

def jk_numpy(eri, D, chunk=16):
    n = D.shape[0]
    X = eri.reshape(n, n, n, n)
    J = np.zeros((n, n))
    K = np.zeros(n * n)
    for d0 in range(0, n, chunk):
        d1 = min(n, d0 + chunk)
        Xc = X[d0:d1]
        # K_ab = sum_cd (ac|bd) D_cd ; rows (d, c), cols (b, a)
        K += D[:, d0:d1].T.reshape(-1) @ Xc.reshape((d1 - d0) * n, n * n)
        for d in range(d0, d1):
            # J_ac = sum_bd (ac|bd) D_bd ; X[d] is (c, b, a)
            J += np.matmul(D[:, d], X[d]).T
    return J, K.reshape(n, n).T


def rhf(S, T, V, eri, nO, maxit=200, conv=1e-10, jk=jk_numpy, ndiis=8):
    """Closed-shell RHF.  Returns dict(eHF, c, ERHF, nOrb, X)."""
    n = S.shape[0]
    s, U = np.linalg.eigh(S)
    keep = s > 1e-6
    X = U[:, keep] / np.sqrt(s[keep])
    H = T + V
    e, Cp = np.linalg.eigh(X.T @ H @ X)
    c = X @ Cp
    fs, errs = [], []
    for it in range(maxit):
        Co = c[:, :nO]
        D = Co @ Co.T
        J, K = jk(eri, D)
        F = H + 2.0 * J - K
        E = float(np.sum(D * (H + F)))
        # commutator in the orthogonalised basis (zero at convergence even
        # when canonical orthogonalisation dropped near-dependent functions)
        err = X.T @ (F @ D @ S - S @ D @ F) @ X
        errmax = np.abs(err).max()
        if errmax < conv:
            break
        fs.append(F)
        errs.append(err)
        fs, errs = fs[-ndiis:], errs[-ndiis:]
        m = len(fs)
        if m > 1:
            B = -np.ones((m + 1, m + 1))
            B[m, m] = 0.0
            for i in range(m):
                for j in range(m):
                    B[i, j] = np.sum(errs[i] * errs[j])
            rhs = np.zeros(m + 1)
            rhs[m] = -1.0
            try:
                w = np.linalg.solve(B, rhs)[:m]
                F = sum(wi * fi for wi, fi in zip(w, fs))
            except np.linalg.LinAlgError:
                pass
        e, Cp = np.linalg.eigh(X.T @ F @ X)
        c = X @ Cp
    else:
        raise RuntimeError("RHF did not converge")
    # final orbitals from the converged Fock matrix
    e, Cp = np.linalg.eigh(X.T @ F @ X)
    c = X @ Cp
    # diagonal exchange self-energy in the MO basis, -(c^T K c)_pp
    Co = c[:, :nO]
    J, K = jk(eri, Co @ Co.T)
    Sx = -np.einsum("mp,mn,np->p", c, K, c)
    return {"eHF": e, "c": np.asfortranarray(c), "ERHF": E, "nOrb": X.shape[1], "niter": it, "Sx": Sx}


def system(mol, jk=jk_numpy):
    """Integrals + RHF for a molecule: everything the GW path takes as input."""
    S, T, V = mol.one_electron()
    eri = mol.eri()
    scf = rhf(S, T, V, eri, mol.nO, jk=jk)
    return {"mol": mol, "S": S, "T": T, "V": V, "eri": eri, "ENuc": mol.enuc(), "nBas": mol.nBas,
            "nO": mol.nO, **scf}


def jk_torch(eri_t, chunk=32):
    """Fock-build closure over a device tensor holding the flat eri array.

    Used to prepare RHF references for large workloads (N = 250: 31 GB) on the
    GPU; torch is only plumbing here, the GW path never calls it."""
    import torch

    def jk(eri_unused, D):
        n = D.shape[0]
        X = eri_t.view(n, n, n, n)                      # X[d, c, b, a] = (ac|bd)
        Dt = torch.from_numpy(np.ascontiguousarray(D)).to(eri_t.device)
        K = (Dt.T.reshape(1, -1) @ X.view(n * n, n * n)).view(n, n).T
        J = torch.zeros(n, n, dtype=torch.float64, device=eri_t.device)
        for d0 in range(0, n, chunk):
            d1 = min(n, d0 + chunk)
            # sum_b D[b, d] X[d, c, b, a] for d in the chunk -> (c, a)
            v = Dt[:, d0:d1].T.reshape(d1 - d0, 1, 1, n)
            J += torch.matmul(v, X[d0:d1]).sum(dim=0).squeeze(1)
        return J.T.cpu().numpy(), K.cpu().numpy()

    return jk
