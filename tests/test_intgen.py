"""Input producer (integrals + RHF) against closed forms and QuAcK's RHF anchors."""
import json
import math
import os

import numpy as np
import pytest
from scipy import special

from quack_amd import intgen as ig

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "quack_goldens.json")))


def boys_ref(m, T):
    if T < 1e-12:
        return 1.0 / (2 * m + 1)
    a = m + 0.5
    return special.gamma(a) * special.gammainc(a, T) / (2.0 * T ** a)


@pytest.mark.parametrize("T", [0.0, 1e-9, 0.3, 2.5, 11.0, 24.9, 25.1, 33.0, 60.0, 300.0])
def test_boys_function(T):
    F = ig.boys(8, T)
    for m in range(9):
        assert abs(F[m] - boys_ref(m, T)) <= 1e-14 * max(1.0, abs(boys_ref(m, T))) + 1e-16
    # downward recursion identity
    for m in range(8):
        assert abs(F[m] - (2 * T * F[m + 1] + math.exp(-T)) / (2 * m + 1)) < 1e-14


def _two_s(a, b, R):
    """Two normalised single-primitive s functions a, b at distance R (bohr)."""
    m = ig.Molecule(["H", "H"], np.array([[0, 0, 0], [0, 0, R]], float), "cc-pvdz")
    m.shell_l = np.array([0, 0], np.int32)
    m.shell_np = np.array([1, 1], np.int32)
    m.shell_atom = np.array([0, 1], np.int32)
    m.exps = np.array([a, b])
    m.coefs = np.array([1.0, 1.0])
    m.nBas = 2
    return m


def test_s_closed_forms():
    a, b, R = 0.8, 0.3, 1.3
    m = _two_s(a, b, R)
    S, T, V = m.one_electron()
    p, mu = a + b, a * b / (a + b)
    S_ab = (4 * a * b / p ** 2) ** 0.75 * math.exp(-mu * R * R)
    assert abs(S[0, 1] - S_ab) < 1e-14
    T_ab = mu * (3 - 2 * mu * R * R) * S_ab
    assert abs(T[0, 1] - T_ab) < 1e-14
    # nuclear attraction of the product at P to nucleus C: -2 sqrt(p/pi) F0(p PC^2) S_ab
    P = b * R / p
    Vref = 0.0
    for C in (0.0, R):
        Vref -= 2.0 * math.sqrt(p / math.pi) * boys_ref(0, p * (P - C) ** 2) * S_ab
    assert abs(V[0, 1] - Vref) < 1e-13
    eri = m.eri().reshape(2, 2, 2, 2, order="F")   # eri[a,b,c,d] = (ac|bd)
    # (aa|bb) for normalised s: 2 sqrt(alpha/pi) F0(alpha R^2), alpha = (2a)(2b)/(2a+2b)
    al = 4 * a * b / (2 * a + 2 * b)
    assert abs(eri[0, 1, 0, 1] - 2 * math.sqrt(al / math.pi) * boys_ref(0, al * R * R)) < 1e-14
    # (aa|aa) = 2 sqrt(a/pi)
    assert abs(eri[0, 0, 0, 0] - 2 * math.sqrt(a / math.pi)) < 1e-14


def test_eri_symmetry_and_normalisation():
    mol = ig.h2o_feather()
    S, T, V = mol.one_electron()
    n = mol.nBas
    assert n == 24
    assert np.abs(np.diag(S) - 1).max() < 1e-13
    for M in (S, T, V):
        assert np.abs(M - M.T).max() < 1e-13
    G = mol.eri().reshape(n, n, n, n, order="F")  # G[a,b,c,d] = <ab|cd> = (ac|bd)
    chem = G.transpose(0, 2, 1, 3)                # chem[a,c,b,d] = (ac|bd)
    for perm in [(1, 0, 2, 3), (0, 1, 3, 2), (2, 3, 0, 1), (3, 2, 1, 0)]:
        assert np.abs(chem - chem.transpose(perm)).max() < 1e-12
    # (aa|bb) Coulomb matrix is positive definite
    J = np.einsum("aabb->ab", chem)
    assert np.linalg.eigvalsh(J).min() > 0


@pytest.mark.parametrize("name", ["He", "H2O"])
def test_rhf_anchors(name):
    mol = ig.he_631g() if name == "He" else ig.h2o_feather()
    s = ig.system(mol)
    g = GOLD[name]["feather"]
    nO = mol.nO
    # electronic energy (HF/RHF.f90:305); QuAcK's SCF threshold 1e-7 limits
    # the orbital-energy agreement to ~1e-8
    assert abs(s["ERHF"] - g["RHF energy"]) < 1e-9
    assert abs(s["eHF"][nO - 1] - g["RHF HOMO energy"]) < 1e-8
    assert abs(s["eHF"][nO] - g["RHF LUMO energy"]) < 1e-8


def test_workload_geometries():
    m = ig.h_chain(10, seed=0)
    z = m.xyz[:, 2]
    assert m.nBas == 50 and m.nO == 5
    assert np.all(np.abs(z - 1.4 * np.arange(10)) <= 0.1)
    lat = ig.h_lattice(5, 10, seed=0)
    assert lat.nBas == 250 and lat.nO == 25
    assert np.all(lat.xyz[:, 2] == 0)
