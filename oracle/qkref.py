"""TEST INFRASTRUCTURE ONLY: ctypes view of oracle/_ref/libqkref.so.

libqkref.so is QuAcK's own Fortran for the restricted and unrestricted
G0W0/evGW paths, phBSE and the qsGW cycle (GW/, LR/, utils/, AOtoMO/ sources
compiled unchanged by oracle/Makefile) plus the bind(C) entry points of
oracle/ref_driver.f90, which compose them in QuAcK's order.  Only tests, the smoke check and the
CPU-baseline leg of bench.py use it, and only as the checker.
"""
import ctypes
import os

import numpy as np

LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_ref", "libqkref.so")
_lib = None

_I = ctypes.c_int
_D = ctypes.c_double
_P = ctypes.POINTER(ctypes.c_double)
_IP = ctypes.POINTER(ctypes.c_int)


def available():
    return os.path.exists(LIB)


def load():
    global _lib
    if _lib is None:
        if not available():
            raise RuntimeError("oracle/_ref/libqkref.so is not built (make -C oracle)")
        # MKL behind this library gives wrong dsyev/dsysvx results in a process
        # that has imported torch (torch carries its own MKL); callers that use
        # torch run the reference in a child process (bench.py does).
        if "torch" in __import__("sys").modules:
            raise RuntimeError("load oracle/_ref in a process without torch (MKL clash)")
        lib = ctypes.CDLL(LIB, mode=os.RTLD_NOW | os.RTLD_LOCAL | os.RTLD_DEEPBIND)
        lib.qkref_aotomo.argtypes = [_I, _I, _P, _P, _P]
        lib.qkref_phrlr.argtypes = [_I] * 7 + [_P, _P, _P, _P, _P, _P]
        lib.qkref_excitation_density.argtypes = [_I] * 5 + [_P, _P, _P]
        lib.qkref_self_energy_diag.argtypes = [_I, _D] + [_I] * 6 + [_P] * 6
        lib.qkref_qp_graph.argtypes = [_I, _D] + [_I] * 6 + [_P] * 7
        lib.qkref_rg0w0.argtypes = [_I, _I, _I, _D] + [_I] * 6 + [_D, _D] + [_P] * 8
        lib.qkref_evrgw.argtypes = [_I, _I, _I, _I, _D, _I, _D] + [_I] * 6 + [_D, _D] + [_P] * 7 + [_IP]
        lib.qkref_evrgw.restype = _I
        lib.qkref_phbse.argtypes = [_I] * 4 + [_D] + [_I] * 6 + [_P] * 6
        lib.qkref_uphbse.argtypes = [_I] * 4 + [_D] + [_I] * 7 + [_P] * 8
        lib.qkref_phbse_full.argtypes = [_I] * 7 + [_D] + [_I] * 6 + [_P] * 8
        lib.qkref_self_energy_full.argtypes = [_I, _D] + [_I] * 6 + [_P] * 6
        lib.qkref_qsgw_cycle.argtypes = [_I, _I, _D] + [_I] * 6 + [_P] * 10
        lib.qkref_ug0w0.argtypes = [_I, _I, _I, _D, _I, _I, _I] + [_I] * 4 + [_P] * 9
        lib.qkref_evugw.argtypes = [_I, _D, _I, _I, _I, _I, _D, _I, _I, _I] + [_I] * 4 + [_P] * 9 + \
            [ctypes.POINTER(_I)]
        _lib = lib
    return _lib


def _p(a):
    assert a.dtype == np.float64 and (a.flags.f_contiguous or a.flags.c_contiguous)
    return a.ctypes.data_as(_P)


def _f(a):
    return np.asfortranarray(a, dtype=np.float64)


def _eri(a):
    """A four-index array for the Fortran side: a flat array is taken as
    QuAcK's column-major data as it is; an (n, n, n, n) array indexed
    [p, q, r, s] is laid out column-major, so Fortran reads ERI(p, q, r, s)
    (a C-order buffer would reach it as ERI(s, r, q, p), which differs for the
    mixed-spin block ERI_aabb)."""
    return np.asfortranarray(a, dtype=np.float64)


def _dims(nOrb, nO, nC=0, nR=0):
    nV = nOrb - nO
    nS = (nO - nC) * (nV - nR)
    return nC, nO, nV, nR, nS


def aotomo(c, eri_ao):
    nBas, nOrb = c.shape
    c = _f(c)
    out = np.zeros(nOrb ** 4)
    load().qkref_aotomo(nBas, nOrb, _p(c), _p(_eri(eri_ao)), _p(out))
    return out


def phrlr(e, eri_mo, nO, tda=False, nC=0, nR=0):
    nOrb = len(e)
    nC, nO, nV, nR, nS = _dims(nOrb, nO, nC, nR)
    Ec = np.zeros(1)
    Om = np.zeros(nS)
    XpY = np.zeros((nS, nS), order="F")
    XmY = np.zeros((nS, nS), order="F")
    load().qkref_phrlr(int(tda), nOrb, nC, nO, nV, nR, nS, _p(_f(e)), _p(_eri(eri_mo)),
                       _p(Ec), _p(Om), _p(XpY), _p(XmY))
    return {"EcRPA": Ec[0], "Om": Om, "XpY": XpY, "XmY": XmY}


def excitation_density(eri_mo, XpY, nOrb, nO, nC=0, nR=0):
    nC, nO, nV, nR, nS = _dims(nOrb, nO, nC, nR)
    rho = np.zeros((nOrb, nOrb, nS), order="F")
    load().qkref_excitation_density(nOrb, nC, nO, nR, nS, _p(_eri(eri_mo)), _p(_f(XpY)), _p(rho))
    return rho


def self_energy_diag(e, Om, rho, nO, eta=0.0, srg=False, nC=0, nR=0):
    nOrb = len(e)
    nC, nO, nV, nR, nS = _dims(nOrb, nO, nC, nR)
    Ec = np.zeros(1)
    Sig = np.zeros(nOrb)
    Z = np.zeros(nOrb)
    load().qkref_self_energy_diag(int(srg), eta, nOrb, nC, nO, nV, nR, nS, _p(_f(e)), _p(_f(Om)), _p(_f(rho)),
                                  _p(Ec), _p(Sig), _p(Z))
    return {"EcGM": Ec[0], "SigC": Sig, "Z": Z}


def qp_graph(eHF, Om, rho, eGWlin, eOld, Z, nO, eta=0.0, srg=False, nC=0, nR=0):
    nOrb = len(eHF)
    nC, nO, nV, nR, nS = _dims(nOrb, nO, nC, nR)
    eGW = np.zeros(nOrb)
    Zio = np.array(Z, dtype=np.float64)
    load().qkref_qp_graph(int(srg), eta, nOrb, nC, nO, nV, nR, nS, _p(_f(eHF)), _p(_f(Om)), _p(_f(rho)),
                          _p(_f(eGWlin)), _p(_f(eOld)), _p(eGW), _p(Zio))
    return {"eGW": eGW, "Z": Zio}


def rg0w0(eri_mo, eHF, nO, eta=0.0, linearize=False, tda_w=False, srg=False, ENuc=0.0, ERHF=0.0, nC=0, nR=0):
    nOrb = len(eHF)
    nC, nO, nV, nR, nS = _dims(nOrb, nO, nC, nR)
    eGW, Sig, Z = np.zeros(nOrb), np.zeros(nOrb), np.zeros(nOrb)
    Om, Ec, Egm = np.zeros(nS), np.zeros(1), np.zeros(1)
    load().qkref_rg0w0(int(tda_w), int(linearize), int(srg), eta, nOrb, nC, nO, nV, nR, nS, ENuc, ERHF,
                       _p(_eri(eri_mo)), _p(_f(eHF)), _p(eGW), _p(Sig), _p(Z), _p(Om),
                       _p(Ec), _p(Egm))
    return {"eGW": eGW, "SigC": Sig, "Z": Z, "Om": Om, "EcRPA": Ec[0], "EcGM": Egm[0]}


def evrgw(eri_mo, eHF, nO, maxSCF=256, thresh=1e-5, max_diis=5, eta=0.0, linearize=False, tda_w=False,
          srg=False, ENuc=0.0, ERHF=0.0, nC=0, nR=0):
    nOrb = len(eHF)
    nC, nO, nV, nR, nS = _dims(nOrb, nO, nC, nR)
    eGW, Sig, Z = np.zeros(nOrb), np.zeros(nOrb), np.zeros(nOrb)
    Ec, Egm = np.zeros(1), np.zeros(1)
    nscf = ctypes.c_int(0)
    st = load().qkref_evrgw(int(tda_w), int(linearize), int(srg), maxSCF, thresh, max_diis, eta, nOrb, nC, nO, nV,
                            nR, nS, ENuc, ERHF, _p(_eri(eri_mo)), _p(_f(eHF)), _p(eGW), _p(Sig),
                            _p(Z), _p(Ec), _p(Egm), ctypes.byref(nscf))
    return {"eGW": eGW, "SigC": Sig, "Z": Z, "EcRPA": Ec[0], "EcGM": Egm[0], "nSCF": nscf.value, "status": st}


def phbse(eri_mo, eW, eGW, nO, eta=0.0, tda_w=False, tda=False, singlet=True, triplet=True, nC=0, nR=0):
    """Static phBSE@GW through QuAcK's own routines (ref_driver.f90
    qkref_phbse, the order of GW/RGW_phBSE.f90:113-190)."""
    nOrb = len(eW)
    nC, nO, nV, nR, nS = _dims(nOrb, nO, nC, nR)
    Ec = np.zeros(2)
    OmS, OmT = np.zeros(nS), np.zeros(nS)
    load().qkref_phbse(int(tda_w), int(tda), int(singlet), int(triplet), eta, nOrb, nC, nO, nV, nR, nS,
                       _p(_eri(eri_mo)), _p(_f(eW)), _p(_f(eGW)), _p(Ec), _p(OmS), _p(OmT))
    return {"EcBSE": Ec, "OmS": OmS, "OmT": OmT}


def phbse_full(eri_mo, eW, eGW, nO, eta=0.0, tda_w=False, tda=False, singlet=True, triplet=True, bse2=False,
               dbse=False, dtda=True):
    """phBSE@GW with QuAcK's second-order kernel and first-order dynamical
    correction (ref_driver.f90 qkref_phbse_full).  Returns EcBSE, OmS, OmT
    and the corrections OmDyn/ZDyn of the first min(nS, 10) states of each
    manifold (singlet, triplet)."""
    nOrb = len(eW)
    nC, nO, nV, nR, nS = _dims(nOrb, nO)
    Ec = np.zeros(2)
    OmS, OmT = np.zeros(nS), np.zeros(nS)
    Od, Zd = np.zeros(20), np.zeros(20)
    load().qkref_phbse_full(int(tda_w), int(tda), int(singlet), int(triplet), int(bse2), int(dbse), int(dtda), eta,
                            nOrb, nC, nO, nV, nR, nS, _p(_eri(eri_mo)), _p(_f(eW)), _p(_f(eGW)), _p(Ec), _p(OmS),
                            _p(OmT), _p(Od), _p(Zd))
    m = min(nS, 10)
    return {"EcBSE": Ec, "OmS": OmS, "OmT": OmT, "dyn": {"singlet": (Od[:m], Zd[:m]), "triplet": (Od[10:10 + m],
                                                                                                 Zd[10:10 + m])}}


def uphbse(eri_aa, eri_ab, eri_bb, eW, eGW, nO, eta=0.0, tda_w=False, tda=False, spin_conserved=True,
           spin_flip=True, nC=(0, 0), nR=(0, 0)):
    """Unrestricted static phBSE@GW through QuAcK's own routines (ref_driver.f90
    qkref_uphbse, the order of GW/UGW_phBSE.f90:60-196); eW, eGW n x 2."""
    n = eW.shape[0]
    nSsc = sum((nO[s] - nC[s]) * (n - nR[s] - nO[s]) for s in range(2))
    nSsf = (nO[0] - nC[0]) * (n - nR[1] - nO[1]) + (nO[1] - nC[1]) * (n - nR[0] - nO[0])
    Ec = np.zeros(2)
    OmSC, OmSF = np.zeros(nSsc), np.zeros(nSsf)
    load().qkref_uphbse(int(tda_w), int(tda), int(spin_conserved), int(spin_flip), eta, n, nO[0], nO[1], nC[0], nC[1],
                        nR[0], nR[1], _p(_eri(eri_aa)), _p(_eri(eri_ab)), _p(_eri(eri_bb)), _p(_f(eW)),
                        _p(_f(eGW)), _p(Ec), _p(OmSC), _p(OmSF))
    return {"EcBSE": Ec, "OmSC": OmSC, "OmSF": OmSF}


def self_energy_full(e, Om, rho, nO, eta=0.0, srg=False, nC=0, nR=0):
    """GW/RGW_self_energy.f90 (or RGW_SRG_self_energy.f90): full SigC, Z, EcGM."""
    nOrb = len(e)
    nC, nO, nV, nR, nS = _dims(nOrb, nO, nC, nR)
    Ec = np.zeros(1)
    Sig = np.zeros((nOrb, nOrb), order="F")
    Z = np.zeros(nOrb)
    load().qkref_self_energy_full(int(srg), eta, nOrb, nC, nO, nV, nR, nS, _p(_f(e)), _p(_f(Om)), _p(_f(rho)),
                                  _p(Ec), _p(Sig), _p(Z))
    return {"EcGM": Ec[0], "SigC": Sig, "Z": Z}


def qsgw_cycle(S, T, V, X, c, P, eGW, eri_ao, nO, eta=0.0, srg=False, tda_w=False, nC=0, nR=0):
    """One cycle of GW/qsRGW.f90 through QuAcK's own routines: F and the new
    quasiparticle energies."""
    n = S.shape[0]
    nC, nO, nV, nR, nS = _dims(n, nO, nC, nR)
    F = np.zeros((n, n), order="F")
    e = np.zeros(n)
    load().qkref_qsgw_cycle(int(srg), int(tda_w), eta, n, nC, nO, nV, nR, nS, _p(_f(S)), _p(_f(T)), _p(_f(V)),
                            _p(_f(X)), _p(_f(c)), _p(_f(P)), _p(_f(eGW)), _p(_eri(eri_ao)), _p(F),
                            _p(e))
    return {"F": F, "eGW": e}


def ug0w0(eri_aa, eri_ab, eri_bb, eHF, nO, eta=0.0, linearize=False, tda_w=False, srg=False, nC=(0, 0),
          nR=(0, 0)):
    """GW/UG0W0.f90 through QuAcK's own unrestricted routines
    (ref_driver.f90 qkref_ug0w0); eHF n x 2, nO, nC, nR = (alpha, beta).
    Frozen entries of eGW are left as QuAcK leaves them (zero here)."""
    n = eHF.shape[0]
    eGW, Sig, Z = (np.zeros((n, 2), order="F") for _ in range(3))
    Ec = np.zeros(1)
    Egm = np.zeros(2)
    load().qkref_ug0w0(int(linearize), int(tda_w), int(srg), eta, n, nO[0], nO[1], nC[0], nC[1], nR[0], nR[1],
                       _p(_eri(eri_aa)), _p(_eri(eri_ab)),
                       _p(_eri(eri_bb)), _p(_f(eHF)), _p(eGW), _p(Sig), _p(Z), _p(Ec), _p(Egm))
    return {"eGW": eGW, "SigC": Sig, "Z": Z, "EcRPA": Ec[0], "EcGM": Egm}


def evugw(eri_aa, eri_ab, eri_bb, eHF, nO, maxSCF=64, thresh=1e-5, max_diis=5, eta=0.0, linearize=False,
          tda_w=False, srg=False, nC=(0, 0), nR=(0, 0)):
    """GW/evUGW.f90's cycle through QuAcK's own routines (ref_driver.f90
    qkref_evugw); eHF n x 2, nO, nC, nR = (alpha, beta)."""
    n = eHF.shape[0]
    eGW, Sig, Z = (np.zeros((n, 2), order="F") for _ in range(3))
    Ec = np.zeros(1)
    Egm = np.zeros(2)
    nscf = ctypes.c_int(0)
    load().qkref_evugw(int(maxSCF), thresh, int(max_diis), int(linearize), int(tda_w), int(srg), eta, n, nO[0], nO[1],
                       nC[0], nC[1], nR[0], nR[1], _p(_eri(eri_aa)),
                       _p(_eri(eri_ab)), _p(_eri(eri_bb)), _p(_f(eHF)), _p(eGW),
                       _p(Sig), _p(Z), _p(Ec), _p(Egm), ctypes.byref(nscf))
    return {"eGW": eGW, "SigC": Sig, "Z": Z, "EcRPA": Ec[0], "EcGM": Egm, "nSCF": nscf.value}

