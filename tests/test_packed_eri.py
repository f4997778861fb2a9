"""CPU: the 8-fold packed AO-integral format of int/ERI_chem.bin.

PyDuck writes pyscf's s8 array (PyDuck.py:318-326) and QuAcK reads it
(utils/read_2e_integrals.f90:85-107), addressing (ab|cd) by the 1-based
Yoshimine index (utils/utils.f90:1186-1217).  The packer must put every
element where that index looks for it, and unpacking must give back the full
array exactly (the producer's integrals are exactly 8-fold symmetric).
"""
import numpy as np
import pytest

from quack_amd import intgen as ig
from quack_amd import quackio


def yoshimine_2(a, b):
    """utils/utils.f90:1205-1217 (1-based)."""
    return a * (a - 1) // 2 + b if a > b else b * (b - 1) // 2 + a


@pytest.fixture(scope="module")
def h2o():
    return ig.system(ig.h2o_feather())


def test_packed_size_and_yoshimine_addressing(h2o):
    N = h2o["nBas"]
    f = np.asarray(h2o["eri"]).reshape(N, N, N, N)  # f[si, nu, la, mu] = (mu nu|la si)
    pk = quackio.pack_eri_s8(h2o["eri"], N)
    P = N * (N + 1) // 2
    assert pk.size == P * (P + 1) // 2
    rng = np.random.default_rng(1)
    for mu, nu, la, si in rng.integers(0, N, size=(4000, 4)):
        idx = yoshimine_2(yoshimine_2(mu + 1, nu + 1), yoshimine_2(la + 1, si + 1)) - 1
        assert pk[idx] == f[si, nu, la, mu]


def test_unpack_round_trip(h2o, tmp_path):
    N = h2o["nBas"]
    quackio.write_eri_chem(str(tmp_path), h2o["eri"], N)
    pk = quackio.read_eri_chem(str(tmp_path), N)
    full = quackio.unpack_eri_s8(pk, N)
    assert np.array_equal(full, np.asarray(h2o["eri"]).reshape(-1))
    with pytest.raises(ValueError):
        quackio.read_eri_chem(str(tmp_path), N + 1)
