"""Python host mirror of QuAcK's restricted GW entry points over the HIP C-ABI.

The functions keep the names, argument meaning and output of QuAcK's Fortran
routines (GW/RG0W0.f90, GW/evRGW.f90, LR/phRLR*.f90, GW/RGW_*.f90,
AOtoMO/AOtoMO_ERI_RHF.f90) and call libquack_amd.so, the gfx950 build of
quack_amd/csrc.  There is no CPU fallback: a missing library or a device
error raises.

Arrays are numpy FP64.  4-index integrals are the flat column-major arrays
QuAcK holds (``G(p,q,r,s) = <pq|rs>`` at ``p + n q + n^2 r + n^3 s``); 2-D
arrays are returned in Fortran order with QuAcK's index meaning
(``XpY[m, ia]``, ``rho[p, q, m]``).
"""
import ctypes
import os

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libquack_amd.so")

QK_OK, QK_E_ARG, QK_E_HIP, QK_E_EIG, QK_E_NOCONV, QK_E_UNSUPPORTED, QK_E_ALLOC, QK_E_UNSTABLE = range(8)
_NAMES = {1: "QK_E_ARG", 2: "QK_E_HIP", 3: "QK_E_EIG", 4: "QK_E_NOCONV", 5: "QK_E_UNSUPPORTED", 6: "QK_E_ALLOC",
          7: "QK_E_UNSTABLE"}

EXPORTS = ["qk_gw_create", "qk_gw_destroy", "qk_gw_last_error", "qk_gw_set_eri_ao", "qk_gw_set_eri_ao_device",
           "qk_gw_set_eri_mo", "qk_rg0w0", "qk_evrgw", "qk_gw_stage_times", "qk_aotomo_eri_rhf", "qk_phrlr",
           "qk_phrlr_a", "qk_phrlr_b", "qk_rgw_excitation_density", "qk_rgw_self_energy_diag", "qk_rgw_qp_graph",
           "qk_last_error", "qk_gw_set_precision", "qk_mlp_create", "qk_mlp_destroy", "qk_mlp_nparams",
           "qk_mlp_get_params", "qk_mlp_set_params", "qk_mlp_forward", "qk_mlp_loss_grad", "qk_mlp_train",
           "qk_mlp_last_error", "qk_gw_screen", "qk_gw_sigma_at", "qk_gw_set_eigensolver", "qk_gw_eig_stats",
           "qk_gw_set_eri_df", "qk_gw_set_eri_df_device", "qk_gwb_create", "qk_gwb_destroy", "qk_gwb_last_error",
           "qk_gwb_set_precision", "qk_gwb_set_eri_ao_device", "qk_gwb_set_eri_mo", "qk_gwb_g0w0",
           "qk_gwb_stage_times", "qk_gwb_evgw", "qk_set_autotune", "qk_autotune_report", "qk_autotune_reset", "qk_set_graphs", "qk_gw_set_eri_ao_packed", "qk_int_eri_device",
           "qk_rgw_phbse", "qk_gw_sigma_full", "qk_ug0w0", "qk_evugw", "qk_rhf_create", "qk_rhf_destroy", "qk_rhf_last_error", "qk_rhf_set_eri_ao", "qk_rhf_jk", "qk_rhf_jk_ms",
           "qk_rgw_qp_graph_noalloc", "qk_noalloc_last_error", "qk_gw_qp_iters",
           "qk_set_unrestricted_eigensolver", "qk_unrestricted_eig_stats", "qk_gwb_set_graphs", "qk_gwb_graph_stats", "qk_set_split_aotomo",
           "qk_qsgw_create", "qk_qsgw_cycle", "qk_qsgw_destroy", "qk_qsgw_last_error", "qk_rgw_phbse_dyn", "qk_set_library_gemm", "qk_ugw_phbse"]


EIG_MODES = {"syevd": 0, "refine": 1, "sytrd": 2, "refine+sytrd": 3}


def _eig_counts(c):
    return {"syevd": c[0], "refined": c[1], "newton_steps": c[2], "fallbacks": c[3], "fail_cluster_size": c[4],
            "fail_unsettled": c[5], "fail_noconv": c[6], "largest_cluster": c[7], "own_tridiagonal": c[8]}


class QKError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s: %s" % (_NAMES.get(code, str(code)), msg))
        self.code = code


class NotConverged(QKError):
    """evGW reached maxSCF+1 cycles (QuAcK prints 'Convergence failed' and stops)."""

    def __init__(self, code, msg, result):
        super().__init__(code, msg)
        self.result = result


_I, _D = ctypes.c_int, ctypes.c_double
_P = ctypes.POINTER(ctypes.c_double)
_IP = ctypes.POINTER(ctypes.c_int)
_H = ctypes.c_void_p
_lib = None


def library_path():
    return LIB_PATH


def load():
    """Load libquack_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libquack_amd.so is not built; run __graft_entry__.build() (no CPU fallback exists)")
    lib = ctypes.CDLL(LIB_PATH)
    lib.qk_gw_create.argtypes = [ctypes.POINTER(_H)] + [_I] * 7
    lib.qk_gw_destroy.argtypes = [_H]
    lib.qk_gw_last_error.argtypes = [_H]
    lib.qk_gw_last_error.restype = ctypes.c_char_p
    lib.qk_last_error.restype = ctypes.c_char_p
    lib.qk_gw_set_eri_ao.argtypes = [_H, _P, _P]
    lib.qk_gw_set_eri_ao_device.argtypes = [_H, _H, _H]
    lib.qk_gw_set_eri_ao_packed.argtypes = [_H, _P, ctypes.c_longlong, _P]
    _IPs = ctypes.POINTER(ctypes.c_int)
    lib.qk_int_eri_device.argtypes = [_I, _IPs, _IPs, _IPs, _P, _P, _P, ctypes.c_void_p, ctypes.c_longlong]
    lib.qk_gw_set_eri_mo.argtypes = [_H, _P]
    lib.qk_rg0w0.argtypes = [_H, _P, _D, _I, _I, _I, _P, _P, _P, _P, _P, _P]
    lib.qk_evrgw.argtypes = [_H, _P, _P, _I, _D, _I, _D, _I, _I, _I, _P, _P, _P, _P, _P, _P, _IP]
    lib.qk_gw_stage_times.argtypes = [_H, _P, _I]
    lib.qk_gw_set_precision.argtypes = [_H, _I]
    lib.qk_gw_screen.argtypes = [_H, _P, _I]
    lib.qk_gw_sigma_at.argtypes = [_H, _P, _P, _D, _I, _P, _P]
    lib.qk_gw_set_eigensolver.argtypes = [_H, _I, _D]
    lib.qk_gw_set_eri_df.argtypes = [_H, _P, _I]
    _LL = ctypes.c_longlong
    lib.qk_gwb_create.argtypes = [ctypes.POINTER(_H)] + [_I] * 8
    lib.qk_gwb_destroy.argtypes = [_H]
    lib.qk_gwb_last_error.argtypes = [_H]
    lib.qk_gwb_last_error.restype = ctypes.c_char_p
    lib.qk_gwb_set_precision.argtypes = [_H, _I]
    lib.qk_gwb_set_eri_ao_device.argtypes = [_H, _H, _LL, _H, _LL]
    lib.qk_gwb_set_eri_mo.argtypes = [_H, _I, _P]
    lib.qk_gwb_g0w0.argtypes = [_H, _P, _D, _I, _I, _I, _P, _P, _P, _P, _P, _P]
    lib.qk_gwb_stage_times.argtypes = [_H, _P, _I]
    lib.qk_gwb_evgw.argtypes = [_H, _P, _P, _I, _D, _I, _D, _I, _I, _I, _P, _P, _P, _P, _P, _P, _IP]
    lib.qk_set_autotune.argtypes = [_I]
    lib.qk_autotune_reset.argtypes = []
    lib.qk_set_graphs.argtypes = [_I]
    lib.qk_autotune_report.argtypes = [ctypes.c_char_p, _I]
    lib.qk_gw_set_eri_df_device.argtypes = [_H, _H, _I]
    lib.qk_gw_eig_stats.argtypes = [_H, ctypes.POINTER(ctypes.c_longlong), _I]
    _F = ctypes.POINTER(ctypes.c_float)
    lib.qk_mlp_create.argtypes = [ctypes.POINTER(_H), _I, _I, _I, ctypes.c_ulonglong]
    lib.qk_mlp_destroy.argtypes = [_H]
    lib.qk_mlp_nparams.argtypes = [_H, _IP]
    lib.qk_mlp_get_params.argtypes = [_H, _F]
    lib.qk_mlp_set_params.argtypes = [_H, _F]
    lib.qk_mlp_forward.argtypes = [_H, _F, _I, _F, _I]
    lib.qk_mlp_loss_grad.argtypes = [_H, _F, _F, _I, _F, _F]
    lib.qk_mlp_train.argtypes = [_H, _F, _F, _I, _I, _I, ctypes.c_float, ctypes.c_ulonglong, _F]
    lib.qk_mlp_last_error.argtypes = [_H]
    lib.qk_mlp_last_error.restype = ctypes.c_char_p
    lib.qk_aotomo_eri_rhf.argtypes = [_I, _I, _P, _P, _P]
    lib.qk_phrlr.argtypes = [_I, _I, _P, _P, _P, _P, _P, _P]
    lib.qk_phrlr_a.argtypes = [_I] * 7 + [_D, _P, _P, _P]
    lib.qk_phrlr_b.argtypes = [_I] * 7 + [_D, _P, _P]
    lib.qk_rgw_excitation_density.argtypes = [_I] * 5 + [_P, _P, _P]
    lib.qk_rgw_self_energy_diag.argtypes = [_I, _D] + [_I] * 6 + [_P] * 6
    lib.qk_rgw_qp_graph.argtypes = [_I, _D, _D] + [_I] * 6 + [_P] * 7
    lib.qk_gw_qp_iters.argtypes = [_H, _IP]
    lib.qk_rgw_qp_graph_noalloc.argtypes = [_I, _D, _D] + [_I] * 6 + [_H] * 8
    lib.qk_noalloc_last_error.restype = ctypes.c_char_p
    lib.qk_rgw_phbse.argtypes = [_I] * 8 + [_D] + [_I] * 6 + [_P] * 5
    lib.qk_set_library_gemm.argtypes = [_I]
    lib.qk_ugw_phbse.argtypes = [_I, _IP, _IP, _IP] + [_P] * 5 + [_D] + [_I] * 5 + [_P] * 3
    lib.qk_rgw_phbse_dyn.argtypes = [_I] * 8 + [_D] + [_I] * 6 + [_P] * 7
    lib.qk_gw_sigma_full.argtypes = [_H, _P, _D, _I, _P, _P, _P, _P]
    lib.qk_ug0w0.argtypes = [_I, _IP, _IP, _IP, _P, _P, _P, _P, _D, _I, _I, _I, _P, _P, _P, _P, _P, _P]
    lib.qk_evugw.argtypes = [_I, _IP, _IP, _IP, _P, _P, _P, _P, _I, _D, _I, _D, _I, _I, _I, _P, _P, _P, _P, _P, _P,
                             _IP]
    lib.qk_set_unrestricted_eigensolver.argtypes = [_I, _D]
    lib.qk_gwb_set_graphs.argtypes = [_H, _I]
    lib.qk_set_split_aotomo.argtypes = [_I]
    lib.qk_gwb_graph_stats.argtypes = [_H, ctypes.POINTER(ctypes.c_longlong), _I]
    lib.qk_unrestricted_eig_stats.argtypes = [ctypes.POINTER(ctypes.c_longlong), _I]
    lib.qk_rhf_create.argtypes = [ctypes.POINTER(_H), _I]
    lib.qk_rhf_destroy.argtypes = [_H]
    lib.qk_rhf_last_error.argtypes = [_H]
    lib.qk_rhf_last_error.restype = ctypes.c_char_p
    lib.qk_rhf_set_eri_ao.argtypes = [_H, _P]
    lib.qk_rhf_jk.argtypes = [_H, _H, _LL, _P, _P, _P]
    lib.qk_rhf_jk_ms.argtypes = [_H, _P]
    lib.qk_qsgw_create.argtypes = [ctypes.POINTER(_H)] + [_I] * 8 + [_P] * 9
    lib.qk_qsgw_cycle.argtypes = [_H, _D, _I, _I] + [_P] * 8
    lib.qk_qsgw_destroy.argtypes = [_H]
    lib.qk_qsgw_last_error.argtypes = [_H]
    lib.qk_qsgw_last_error.restype = ctypes.c_char_p
    _lib = lib
    return lib


def _p(a):
    if a is None:
        return None
    return a.ctypes.data_as(_P)


def _f64(a, order="F"):
    return np.require(np.asarray(a, dtype=np.float64), requirements=[order, "A"])


def _check(rc, handle=None):
    if rc != QK_OK:
        lib = load()
        msg = (lib.qk_gw_last_error(handle) if handle else lib.qk_last_error()) or b""
        raise QKError(rc, msg.decode(errors="replace"))


# names of the qk_gw_set_precision modes: FP64, three BF16 slices (six slice
# products), two BF16 slices (three slice products) in the excitation densities
PRECISION = {"fp64": 0, "mixed": 1, "mixed2": 2}


def dims(nOrb, nO, nC=0, nR=0):
    nV = nOrb - nO
    return nC, nO, nV, nR, (nO - nC) * (nV - nR)


class GWEngine:
    """Device-resident G0W0/evGW state for one molecule (one qk_gw_handle)."""

    def __init__(self, nBas, nOrb, nO, nC=0, nR=0):
        lib = load()
        self.nBas, self.nOrb = nBas, nOrb
        self.nC, self.nO, self.nV, self.nR, self.nS = dims(nOrb, nO, nC, nR)
        h = _H()
        _check(lib.qk_gw_create(ctypes.byref(h), nBas, nOrb, nC, nO, self.nV, nR, self.nS))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            load().qk_gw_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_precision(self, mode):
        """0 / "fp64": QuAcK's FP64 arithmetic; 1 / "mixed": excitation densities
        from three BF16 slices (six products); 2 / "mixed2": two slices (three
        products)."""
        _check(load().qk_gw_set_precision(self.h, PRECISION.get(mode, mode)), self.h)

    def set_eigensolver(self, mode, de_max=5e-3):
        """"syevd" (0): dsyevd for every RPA, QuAcK's scheme; "refine" (1):
        Newton refinement of the previous eigenvectors when no orbital energy
        moved by more than de_max Ha, dsyevd otherwise; "sytrd" (2): full solves
        through the library's own tridiagonalisation; "refine+sytrd" (3)."""
        _check(load().qk_gw_set_eigensolver(self.h, EIG_MODES.get(mode, mode), de_max), self.h)

    def eig_stats(self):
        c = (ctypes.c_longlong * 9)()
        _check(load().qk_gw_eig_stats(self.h, c, 9), self.h)
        return _eig_counts(c)

    def set_eri_df(self, B):
        """Density-fitted MO integrals: B of shape (naux, nOrb, nOrb) (any float64
        array; stored with P fastest), (pq|rs) = sum_P B[P,p,q] B[P,r,s]."""
        B = np.asarray(B, dtype=np.float64)
        naux = B.shape[0]
        assert B.shape == (naux, self.nOrb, self.nOrb) and self.nBas == self.nOrb
        flat = np.ascontiguousarray(np.transpose(B, (2, 1, 0))).reshape(-1)  # P fastest, then p, then q
        _check(load().qk_gw_set_eri_df(self.h, flat.ctypes.data_as(_P), naux), self.h)

    def set_eri_df_device(self, ptr, naux):
        """B already on the device as B[P + naux*(p + nOrb*q)] (a data_ptr())."""
        _check(load().qk_gw_set_eri_df_device(self.h, ptr, naux), self.h)

    def set_eri_ao(self, eri_ao, c):
        eri_ao = np.ascontiguousarray(eri_ao, dtype=np.float64).reshape(-1)
        assert eri_ao.size == self.nBas ** 4
        c = _f64(c)
        assert c.shape == (self.nBas, self.nOrb)
        _check(load().qk_gw_set_eri_ao(self.h, _p(eri_ao), _p(c)), self.h)

    def set_eri_ao_packed(self, eri_s8, c):
        """AO integrals in the 8-fold packed chemists' form of int/ERI_chem.bin
        (quackio.pack_eri_s8 writes one); unpacked on the device."""
        eri_s8 = np.ascontiguousarray(eri_s8, dtype=np.float64).reshape(-1)
        c = _f64(c)
        assert c.shape == (self.nBas, self.nOrb)
        _check(load().qk_gw_set_eri_ao_packed(self.h, _p(eri_s8), eri_s8.size, _p(c)), self.h)

    def set_eri_ao_device(self, eri_ptr, c_ptr):
        """Device pointers (e.g. torch tensor.data_ptr()) of ERI_AO and c (column-major)."""
        _check(load().qk_gw_set_eri_ao_device(self.h, _H(eri_ptr), _H(c_ptr)), self.h)

    def set_eri_mo(self, eri_mo):
        eri_mo = np.ascontiguousarray(eri_mo, dtype=np.float64).reshape(-1)
        assert eri_mo.size == self.nOrb ** 4
        _check(load().qk_gw_set_eri_mo(self.h, _p(eri_mo)), self.h)

    def g0w0(self, eHF, eta=0.0, linearize=False, TDA_W=False, doSRG=False):
        n, nS = self.nOrb, self.nS
        eHF = _f64(eHF)
        eGW, Sig, Z, Om = np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(nS)
        Ec, Egm = _D(0.0), _D(0.0)
        rc = load().qk_rg0w0(self.h, _p(eHF), eta, int(linearize), int(TDA_W), int(doSRG), _p(eGW), _p(Sig), _p(Z),
                             _p(Om), ctypes.byref(Ec), ctypes.byref(Egm))
        _check(rc, self.h)
        return {"eGW": eGW, "SigC": Sig, "Z": Z, "Om": Om, "EcRPA": Ec.value, "EcGM": Egm.value}

    def evgw(self, eHF, maxSCF=256, thresh=1e-5, max_diis=5, eta=0.0, linearize=False, TDA_W=False, doSRG=False,
             eGW_start=None):
        n, nS = self.nOrb, self.nS
        eHF = _f64(eHF)
        start = None if eGW_start is None else _f64(eGW_start)
        eGW, Sig, Z, Om = np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(nS)
        Ec, Egm, ns = _D(0.0), _D(0.0), _I(0)
        rc = load().qk_evrgw(self.h, _p(eHF), _p(start), maxSCF, thresh, max_diis, eta, int(linearize), int(TDA_W),
                             int(doSRG), _p(eGW), _p(Sig), _p(Z), _p(Om), ctypes.byref(Ec), ctypes.byref(Egm),
                             ctypes.byref(ns))
        res = {"eGW": eGW, "SigC": Sig, "Z": Z, "Om": Om, "EcRPA": Ec.value, "EcGM": Egm.value, "nSCF": ns.value}
        if rc == QK_E_NOCONV:
            raise NotConverged(rc, load().qk_gw_last_error(self.h).decode(), res)
        _check(rc, self.h)
        return res

    def screen(self, e, TDA_W=False):
        _check(load().qk_gw_screen(self.h, _p(_f64(e)), int(TDA_W)), self.h)

    def sigma_full(self, e, eta=0.0, doSRG=False):
        """Full self-energy matrix of qsGW at e (qk_gw_sigma_full) from the last
        screen(); returns dict(SigC (nOrb x nOrb), Z, EcGM, EcRPA)."""
        n = self.nOrb
        Sig = np.zeros((n, n), order="F")
        Z = np.zeros(n)
        Ec, Er = _D(0.0), _D(0.0)
        _check(load().qk_gw_sigma_full(self.h, _p(_f64(e)), eta, int(doSRG), _p(Sig), _p(Z), ctypes.byref(Ec),
                                       ctypes.byref(Er)), self.h)
        return {"SigC": Sig, "Z": Z, "EcGM": Ec.value, "EcRPA": Er.value}

    def sigma_at(self, eG, w, eta=0.0, doSRG=False):
        n = self.nOrb
        Sig, dSig = np.zeros(n), np.zeros(n)
        _check(load().qk_gw_sigma_at(self.h, _p(_f64(eG)), _p(_f64(w)), eta, int(doSRG), _p(Sig), _p(dSig)), self.h)
        return Sig, dSig

    def qp_iters(self):
        """Newton iterations per orbital of the last graphical QP solve (64: the
        solve failed and the orbital kept its starting energy)."""
        nit = np.zeros(self.nOrb, dtype=np.int32)
        _check(load().qk_gw_qp_iters(self.h, nit.ctypes.data_as(_IP)), self.h)
        return nit

    def stage_times(self):
        ms = np.zeros(14)
        _check(load().qk_gw_stage_times(self.h, _p(ms), 14), self.h)
        return dict(zip(["aotomo", "rpa_build", "eigensolve", "excitation", "sigma_qp", "total", "excitation_gemm",
                         "excitation_gemm_launches", "excitation_gemm_all_calls", "excitation_gemm_launches_all_calls",
                         "tridiagonalisation_all_calls", "tridiagonalisations_all_calls", "stedc_all_calls",
                         "back_transformation_all_calls"], ms))


def set_autotune(on):
    """Process-wide GEMM tile autotuner (on by default; identical results either way)."""
    load().qk_set_autotune(int(bool(on)))


def set_graphs(on):
    """Process-wide hipGraph replay of the own full eigensolve (off by default;
    identical results either way)."""
    load().qk_set_graphs(int(bool(on)))


def autotune_reset():
    """Forget every tuned GEMM shape (qk_autotune_reset): the next large call
    of each shape tunes again."""
    _check(load().qk_autotune_reset())


def autotune_report():
    buf = ctypes.create_string_buffer(1 << 16)
    rc = load().qk_autotune_report(buf, len(buf))
    if rc != 0:
        raise QKError(rc, "autotune report larger than 64 KiB")
    return [l for l in buf.value.decode().splitlines() if l]


class GWBatch:
    """nmol molecules of one size (qk_gwb_*): G0W0 or evGW for all of them
    stage by stage on one stream.  Uneven batches: group by size
    (quack_amd/batch.py)."""

    def __init__(self, nmol, nBas, nOrb, nO, nC=0, nR=0):
        self.nmol, self.nBas, self.nOrb, self.nO = nmol, nBas, nOrb, nO
        nV = nOrb - nO
        self.nS = (nO - nC) * (nV - nR)
        h = _H()
        rc = load().qk_gwb_create(ctypes.byref(h), nmol, nBas, nOrb, nC, nO, nV, nR, self.nS)
        if rc != 0:
            raise QKError(rc, load().qk_gwb_last_error(None).decode())
        self.h = h

    def _check(self, rc):
        if rc != 0:
            raise QKError(rc, load().qk_gwb_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            load().qk_gwb_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_precision(self, mode):
        self._check(load().qk_gwb_set_precision(self.h, PRECISION.get(mode, mode)))

    def set_graphs(self, on):
        """hipGraph replay of evgw()'s cycles from the second cycle of a call."""
        self._check(load().qk_gwb_set_graphs(self.h, int(bool(on))))

    def graph_stats(self):
        c = (ctypes.c_longlong * 2)()
        self._check(load().qk_gwb_graph_stats(self.h, c, 2))
        return {"eager_cycles": c[0], "graph_cycles": c[1]}

    def set_eri_ao_device(self, eri_ptr, stride_eri, c_ptr, stride_c):
        """Device pointers: molecule m's AO ERIs at eri_ptr + 8*m*stride_eri bytes
        (nBas^4, physicist order), its MO coefficients at c_ptr + 8*m*stride_c
        (nBas x nOrb, column-major)."""
        self._check(load().qk_gwb_set_eri_ao_device(self.h, eri_ptr, stride_eri, c_ptr, stride_c))

    def set_eri_mo(self, m, eri_mo):
        eri_mo = np.asfortranarray(eri_mo, dtype=np.float64)
        self._check(load().qk_gwb_set_eri_mo(self.h, m, eri_mo.ctypes.data_as(_P)))

    def g0w0(self, eHF, eta=0.0, linearize=False, TDA_W=False, doSRG=False):
        n, nS, nm = self.nOrb, self.nS, self.nmol
        eHF = np.asfortranarray(np.asarray(eHF, dtype=np.float64).reshape(nm, n).T)
        eGW, Sig, Z = (np.zeros((n, nm), order="F") for _ in range(3))
        Om = np.zeros((nS, nm), order="F")
        Ec, Egm = np.zeros(nm), np.zeros(nm)
        self._check(load().qk_gwb_g0w0(self.h, _p(eHF), eta, int(linearize), int(TDA_W), int(doSRG), _p(eGW),
                                       _p(Sig), _p(Z), _p(Om), _p(Ec), _p(Egm)))
        return {"eGW": eGW.T.copy(), "SigC": Sig.T.copy(), "Z": Z.T.copy(), "Om": Om.T.copy(), "EcRPA": Ec,
                "EcGM": Egm}

    def evgw(self, eHF, maxSCF=256, thresh=1e-5, max_diis=5, eta=0.0, linearize=False, TDA_W=False, doSRG=False,
             eGW_start=None):
        """evGW of every molecule (qk_gwb_evgw); eHF (and eGW_start) nmol x nOrb.
        Raises NotConverged (with the results) when some molecule reached
        maxSCF + 1 cycles."""
        n, nS, nm = self.nOrb, self.nS, self.nmol
        eHF = np.asfortranarray(np.asarray(eHF, dtype=np.float64).reshape(nm, n).T)
        start = None if eGW_start is None else \
            np.asfortranarray(np.asarray(eGW_start, dtype=np.float64).reshape(nm, n).T)
        eGW, Sig, Z = (np.zeros((n, nm), order="F") for _ in range(3))
        Om = np.zeros((nS, nm), order="F")
        Ec, Egm = np.zeros(nm), np.zeros(nm)
        ns = np.zeros(nm, dtype=np.int32)
        rc = load().qk_gwb_evgw(self.h, _p(eHF), _p(start), maxSCF, thresh, max_diis, eta, int(linearize),
                                int(TDA_W), int(doSRG), _p(eGW), _p(Sig), _p(Z), _p(Om), _p(Ec), _p(Egm),
                                ns.ctypes.data_as(_IP))
        res = {"eGW": eGW.T.copy(), "SigC": Sig.T.copy(), "Z": Z.T.copy(), "Om": Om.T.copy(), "EcRPA": Ec,
               "EcGM": Egm, "nSCF": ns.astype(int)}
        if rc == QK_E_NOCONV:
            raise NotConverged(rc, load().qk_gwb_last_error(self.h).decode(), res)
        self._check(rc)
        return res

    def stage_times(self):
        ms = np.zeros(4)
        self._check(load().qk_gwb_stage_times(self.h, _p(ms), 4))
        return dict(zip(["aotomo", "eigensolve", "excitation", "sigma_qp"], ms))


# ---- QuAcK-named entry points -------------------------------------------------

def RG0W0(ERI, eHF, nO, eta=0.0, linearize=False, TDA_W=False, doSRG=False, nC=0, nR=0):
    """GW/RG0W0.f90 on the GPU from MO integrals ERI(nOrb^4) and eHF."""
    n = len(eHF)
    with GWEngine(n, n, nO, nC, nR) as g:
        g.set_eri_mo(ERI)
        return g.g0w0(eHF, eta, linearize, TDA_W, doSRG)


def evRGW(ERI, eHF, nO, maxSCF=256, thresh=1e-5, max_diis=5, eta=0.0, linearize=False, TDA_W=False, doSRG=False,
          nC=0, nR=0):
    """GW/evRGW.f90 on the GPU from MO integrals ERI(nOrb^4) and eHF."""
    n = len(eHF)
    with GWEngine(n, n, nO, nC, nR) as g:
        g.set_eri_mo(ERI)
        return g.evgw(eHF, maxSCF, thresh, max_diis, eta, linearize, TDA_W, doSRG)


def AOtoMO_ERI_RHF(c, ERI_AO):
    c = _f64(c)
    nBas, nOrb = c.shape
    eri_ao = np.ascontiguousarray(ERI_AO, dtype=np.float64).reshape(-1)
    out = np.zeros(nOrb ** 4)
    _check(load().qk_aotomo_eri_rhf(nBas, nOrb, _p(c), _p(eri_ao), _p(out)))
    return out


def phRLR_A(e, ERI, nO, lam=1.0, dRPA=True, nC=0, nR=0):
    n = len(e)
    nC, nO, nV, nR, nS = dims(n, nO, nC, nR)
    A = np.zeros((nS, nS), order="F")
    eri = np.ascontiguousarray(ERI, dtype=np.float64).reshape(-1)
    _check(load().qk_phrlr_a(int(dRPA), n, nC, nO, nV, nR, nS, lam, _p(_f64(e)), _p(eri), _p(A)))
    return A


def phRLR_B(ERI, n, nO, lam=1.0, dRPA=True, nC=0, nR=0):
    nC, nO, nV, nR, nS = dims(n, nO, nC, nR)
    B = np.zeros((nS, nS), order="F")
    eri = np.ascontiguousarray(ERI, dtype=np.float64).reshape(-1)
    _check(load().qk_phrlr_b(int(dRPA), n, nC, nO, nV, nR, nS, lam, _p(eri), _p(B)))
    return B


def phRLR(TDA, A, B):
    """LR/phRLR.f90: returns EcRPA, Om, XpY, XmY."""
    A = _f64(A)
    nS = A.shape[0]
    B = A if (TDA or B is None) else _f64(B)
    Om = np.zeros(nS)
    XpY = np.zeros((nS, nS), order="F")
    XmY = np.zeros((nS, nS), order="F")
    Ec = _D(0.0)
    _check(load().qk_phrlr(int(TDA), nS, _p(A), _p(B), ctypes.byref(Ec), _p(Om), _p(XpY), _p(XmY)))
    return Ec.value, Om, XpY, XmY


def RGW_excitation_density(ERI, XpY, nOrb, nO, nC=0, nR=0):
    nC, nO, nV, nR, nS = dims(nOrb, nO, nC, nR)
    rho = np.zeros((nOrb, nOrb, nS), order="F")
    eri = np.ascontiguousarray(ERI, dtype=np.float64).reshape(-1)
    _check(load().qk_rgw_excitation_density(nOrb, nC, nO, nR, nS, _p(eri), _p(_f64(XpY)), _p(rho)))
    return rho


def RGW_self_energy_diag(e, Om, rho, nO, eta=0.0, doSRG=False, nC=0, nR=0):
    n = len(e)
    nC, nO, nV, nR, nS = dims(n, nO, nC, nR)
    Sig, Z = np.zeros(n), np.zeros(n)
    Ec = _D(0.0)
    _check(load().qk_rgw_self_energy_diag(int(doSRG), eta, n, nC, nO, nV, nR, nS, _p(_f64(e)), _p(_f64(Om)),
                                          _p(_f64(rho)), ctypes.byref(Ec), _p(Sig), _p(Z)))
    return Sig, Z, Ec.value


def RGW_QP_graph(eHF, Om, rho, eGWlin, eOld, Z, nO, eta=0.0, doSRG=False, flow=500.0, nC=0, nR=0):
    n = len(eHF)
    nC, nO, nV, nR, nS = dims(n, nO, nC, nR)
    eGW = np.zeros(n)
    Zio = np.array(Z, dtype=np.float64)
    _check(load().qk_rgw_qp_graph(int(doSRG), eta, flow, n, nC, nO, nV, nR, nS, _p(_f64(eHF)), _p(_f64(Om)),
                                  _p(_f64(rho)), _p(_f64(eGWlin)), _p(_f64(eOld)), _p(eGW), _p(Zio)))
    return eGW, Zio


# ---- per-orbital MLP (learned warm start) ---------------------------------

class MLP:
    """Device MLP over the qk_mlp_* C-ABI (FP32 weights; FP32 or BF16 inference)."""

    def __init__(self, nin=14, width=128, depth=3, seed=0):
        lib = load()
        h = _H()
        _check(lib.qk_mlp_create(ctypes.byref(h), nin, width, depth, seed))
        self.h, self.nin, self.width, self.depth = h, nin, width, depth
        n = _I(0)
        lib.qk_mlp_nparams(h, ctypes.byref(n))
        self.nparams = n.value

    def _chk(self, rc):
        if rc != QK_OK:
            raise QKError(rc, load().qk_mlp_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            load().qk_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _f(a):
        a = np.ascontiguousarray(a, dtype=np.float32)
        return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))

    def get_params(self):
        out = np.zeros(self.nparams, np.float32)
        self._chk(load().qk_mlp_get_params(self.h, self._f(out)[1]))
        return out

    def set_params(self, p):
        a, ptr = self._f(p)
        assert a.size == self.nparams
        self._chk(load().qk_mlp_set_params(self.h, ptr))

    def forward(self, X, bf16=False):
        X, px = self._f(X)
        n = X.shape[0]
        Y = np.zeros(n, np.float32)
        self._chk(load().qk_mlp_forward(self.h, px, n, self._f(Y)[1], int(bf16)))
        return Y

    def loss_grad(self, X, T):
        X, px = self._f(X)
        T, pt = self._f(T)
        loss = np.zeros(1, np.float32)
        g = np.zeros(self.nparams, np.float32)
        self._chk(load().qk_mlp_loss_grad(self.h, px, pt, X.shape[0], self._f(loss)[1], self._f(g)[1]))
        return float(loss[0]), g

    def train(self, X, T, epochs=100, batch=1024, lr=1e-3, seed=0):
        X, px = self._f(X)
        T, pt = self._f(T)
        hist = np.zeros(max(epochs, 1), np.float32)
        self._chk(load().qk_mlp_train(self.h, px, pt, X.shape[0], epochs, batch, lr, seed, self._f(hist)[1]))
        return hist[:epochs]


def RGW_phBSE(ERI, eW, eGW, nO, eta=0.0, TDA_W=False, TDA=False, singlet=True, triplet=True, exchange_kernel=False,
              nC=0, nR=0, dophBSE2=False, dBSE=False, dTDA=True):
    """GW/RGW_phBSE.f90 (qk_rgw_phbse_dyn): the static kernel, with
    dophBSE2 the second-order one added, with dBSE the first-order dynamical
    correction of the first min(nS, 10) states of each manifold.  Returns
    dict(EcBSE (2), OmS, OmT, dyn: {"singlet"/"triplet": (Om_dyn, Z_dyn)})."""
    eW, eGW = _f64(eW), _f64(eGW)
    n = len(eW)
    nV = n - nO
    nS = (nO - nC) * (nV - nR)
    ERI = _f64(ERI)  # QuAcK's column-major nOrb^4 array (a flat array as it is)
    Ec = np.zeros(2)
    Om = np.zeros(2 * nS)
    Od, Zd = np.zeros(20), np.zeros(20)
    _check(load().qk_rgw_phbse_dyn(int(dophBSE2), int(exchange_kernel), int(TDA_W), int(TDA), int(dBSE), int(dTDA),
                                   int(singlet), int(triplet), eta, n, nC, nO, nV, nR, nS, _p(ERI), _p(eW), _p(eGW),
                                   _p(Ec), _p(Om), _p(Od), _p(Zd)))
    m = min(nS, 10)
    dyn = {}
    if dBSE:
        if singlet:
            dyn["singlet"] = (Od[:m].copy(), Zd[:m].copy())
        if triplet:
            dyn["triplet"] = (Od[10:10 + m].copy(), Zd[10:10 + m].copy())
    return {"EcBSE": Ec, "OmS": Om[:nS].copy(), "OmT": Om[nS:].copy(), "dyn": dyn}


def set_library_gemm(on):
    """rocBLAS's dgemm as an autotuner candidate for large plain FP64 products
    (process-wide, off by default; results can then differ in the last bits)."""
    _check(load().qk_set_library_gemm(int(bool(on))))


def set_split_aotomo(on):
    """Split-precision AO->MO block for the integrals set while it is on
    (process-wide, off by default; (ia|jb) stays FP64)."""
    _check(load().qk_set_split_aotomo(int(bool(on))))


def set_unrestricted_eigensolver(mode, de_max=5e-3):
    """Eigensolver of UG0W0/evUGW's screening (process-wide), the modes of
    GWEngine.set_eigensolver; "syevd" (QuAcK's sequence) by default."""
    _check(load().qk_set_unrestricted_eigensolver(EIG_MODES.get(mode, mode), de_max))


def unrestricted_eig_stats():
    """Eigensolver counts of this thread's last UG0W0/evUGW (as GWEngine.eig_stats)."""
    c = (ctypes.c_longlong * 9)()
    _check(load().qk_unrestricted_eig_stats(c, 9))
    return _eig_counts(c)


def _ucounts(n, nC, nO, nR):
    """Per-spin int32 arrays nC, nO, nR and the pair count nSa + nSb."""
    c, o, r = (np.ascontiguousarray(x, dtype=np.int32).reshape(2) for x in (nC, nO, nR))
    nSt = int(sum((o[s] - c[s]) * (n - r[s] - o[s]) for s in range(2)))
    return c, o, r, nSt


def _ip(a):
    return a.ctypes.data_as(_IP)


def UG0W0(ERI_aaaa, ERI_aabb, ERI_bbbb, eHF, nO, eta=0.0, linearize=False, TDA_W=False, doSRG=False, nC=(0, 0),
          nR=(0, 0)):
    """GW/UG0W0.f90 on the GPU (qk_ug0w0).  eHF: nBas x 2 (alpha, beta); nO,
    nC, nR = (alpha, beta) occupied, frozen core and frozen virtual counts.
    Returns dict(eGW, SigC, Z (nBas x 2), Om, EcRPA, EcGM (2))."""
    eHF = np.asfortranarray(eHF, dtype=np.float64)
    n = eHF.shape[0]
    c, o, r, nSt = _ucounts(n, nC, nO, nR)
    eGW, Sig, Z = (np.zeros((n, 2), order="F") for _ in range(3))
    Om = np.zeros(nSt)
    Ec = _D(0.0)
    Egm = np.zeros(2)
    _check(load().qk_ug0w0(n, _ip(c), _ip(o), _ip(r), _p(_f64(ERI_aaaa)), _p(_f64(ERI_aabb)), _p(_f64(ERI_bbbb)),
                           _p(eHF), eta, int(linearize), int(TDA_W), int(doSRG), _p(eGW), _p(Sig), _p(Z), _p(Om),
                           ctypes.byref(Ec), _p(Egm)))
    return {"eGW": eGW, "SigC": Sig, "Z": Z, "Om": Om, "EcRPA": Ec.value, "EcGM": Egm}


def UGW_phBSE(ERI_aaaa, ERI_aabb, ERI_bbbb, eW, eGW, nO, eta=0.0, TDA_W=False, TDA=False, spin_conserved=True,
              spin_flip=True, exchange_kernel=False, nC=(0, 0), nR=(0, 0)):
    """GW/UGW_phBSE.f90 with the static kernel (qk_ugw_phbse).  eW, eGW:
    nBas x 2.  Returns dict(EcBSE (2), OmSC, OmSF)."""
    eW = np.asfortranarray(eW, dtype=np.float64)
    eGW = np.asfortranarray(eGW, dtype=np.float64)
    n = eW.shape[0]
    c, o, r, nSc = _ucounts(n, nC, nO, nR)
    nSf = int((o[0] - c[0]) * (n - r[1] - o[1]) + (o[1] - c[1]) * (n - r[0] - o[0]))
    Ec, OmSC, OmSF = np.zeros(2), np.zeros(nSc), np.zeros(nSf)
    _check(load().qk_ugw_phbse(n, _ip(c), _ip(o), _ip(r), _p(_f64(ERI_aaaa)), _p(_f64(ERI_aabb)),
                               _p(_f64(ERI_bbbb)), _p(eW), _p(eGW), eta, int(TDA_W), int(TDA), int(spin_conserved),
                               int(spin_flip), int(exchange_kernel), _p(Ec), _p(OmSC), _p(OmSF)))
    return {"EcBSE": Ec, "OmSC": OmSC, "OmSF": OmSF}


def evUGW(ERI_aaaa, ERI_aabb, ERI_bbbb, eHF, nO, maxSCF=64, thresh=1e-5, max_diis=5, eta=0.0, linearize=False,
          TDA_W=False, doSRG=False, nC=(0, 0), nR=(0, 0)):
    """GW/evUGW.f90 on the GPU (qk_evugw), QuAcK's defaults for maxSCF, thresh
    and max_diis; orbital counts as UG0W0.  Returns dict(eGW, SigC, Z, Om,
    EcRPA, EcGM, nSCF, converged); non-convergence is reported in "converged"
    with the last cycle's outputs (QuAcK stops there)."""
    eHF = np.asfortranarray(eHF, dtype=np.float64)
    n = eHF.shape[0]
    c, o, r, nSt = _ucounts(n, nC, nO, nR)
    eGW, Sig, Z = (np.zeros((n, 2), order="F") for _ in range(3))
    Om = np.zeros(nSt)
    Ec = _D(0.0)
    Egm = np.zeros(2)
    nscf = _I(0)
    rc = load().qk_evugw(n, _ip(c), _ip(o), _ip(r), _p(_f64(ERI_aaaa)), _p(_f64(ERI_aabb)), _p(_f64(ERI_bbbb)),
                         _p(eHF), int(maxSCF), thresh, int(max_diis), eta, int(linearize), int(TDA_W), int(doSRG),
                         _p(eGW), _p(Sig), _p(Z), _p(Om), ctypes.byref(Ec), _p(Egm), ctypes.byref(nscf))
    if rc != QK_E_NOCONV:
        _check(rc)
    return {"eGW": eGW, "SigC": Sig, "Z": Z, "Om": Om, "EcRPA": Ec.value, "EcGM": Egm, "nSCF": nscf.value,
            "converged": rc == QK_OK}

