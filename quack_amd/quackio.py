"""QuAcK-format input files for the Fortran drop-in driver (quack_amd/fortran/qk_gw_driver).

Writes the files QuAcK reads (input/methods, input/options, input/molecule,
input/hpc_flags, int/ENuc.dat, int/nBas.dat, int/ERI.bin with the layouts of
PyDuck.py:137-153,329-345 and input/*.default) plus the two arrays QuAcK
computes in-process before RGW (int/MO.dat: nOrb, ERHF, eHF; int/ERI_MO.bin).
For qsGW also the one-electron integrals int/Ov.dat, Kin.dat, Nuc.dat in
PyDuck's "mu nu value" lines (PyDuck.py:277-284, read by
utils/read_1e_integrals.f90), and two more in-process arrays: int/X.bin (the
orthogonalisation matrix, nBas x nOrb) and int/cHF.bin (the RHF orbitals,
nBas x nOrb), raw FP64 column-major.
"""
import os

import numpy as np

METHODS = """# RHF UHF GHF ROHF HFB cRHF cUHF eRHF scGHF
  {rhf}   {uhf}   F   F    F   F    F    F    F
# MP2 MP3
  F   F
# CCD pCCD DCD CCSD CCSD(T)
  F   F    F   F    F
# drCCD rCCD crCCD lCCD
  F     F    F     F
# CIS CIS(D) CID CISD FCI
  F   F      F   F    F
# phRPA phRPAx crRPA ppRPA BRPA
  F     F      F     F     F
# OORPA
  F
# G0F2 evGF2 qsGF2 G0F3 evGF3 psdG0F3 scGF2
  F    F     F     F    F     F     F
# G0W0 evGW qsGW scGW
  {g0w0}    {evgw}    {qsgw}    F
# G0T0pp evGTpp qsGTpp ufG0T0pp
  F      F        F       F
# G0T0eh evGTeh qsGTeh
  F      F      F
# evParquet qsParquet
  F         F
# IPEA-ADC2 IPEA-ADC3 SOSEX 2SOSEX G3W2 psdG3W2
  F         F         F     F      F    F
# ADC-GW ADC-2SOSEX ADC(3)-G3W2 ADC(3x)-G3W2 ADC-G3W2
  F      F          F           F            F
# Rtest Utest Gtest
  {test}     F     F
"""

OPTIONS = """# HF: maxSCF thresh    DIIS guess mix shift stab search aordm readFCIDUMP MOM WriteMOs
      256    0.0000001   5    1   0.0  0.0   F     F      F       F        F    F
# MP: reg
      F
# CC: maxSCF thresh  DIIS
      64     0.00001 5
# spin: TDA singlet triplet
         F     T      T
# RPA/MP2: CVS-Alpha CVS-Beta  FC-Alpha FC-Beta
               0         0          0      0
# OORPA:  maxIter thresh  dRPA state diagHess
          256     0.00001 T    0     T
# GF: maxSCF thresh  DIIS lin eta renorm reg linDM restart_scGF2 verbose_scGF2
      256    0.00001 5    F   0.0 0      F   F     F         F
# GW: maxSCF thresh  DIIS lin eta TDA_W reg nfreqs shift linDM restart_scGW verbose_scGW adjust_mu_scGW change_sign_XoB
      {maxscf}    {thresh} {diis}    {lin}   {eta} {tda}     {reg}   1      0.0   F     F            F            F              F
# GT: maxSCF thresh  DIIS lin eta TDA_T reg linDM
      256    0.00001 5    F   0.0 F     F   F
# ACFDT: AC Kx  XBS
         F  F   F
# BSE: phBSE phBSE2 ppBSE dBSE dTDA
       {bse}     {bse2}      F     {dbse}    {dtda}
"""


def tf(b):
    return "T" if b else "F"


def write_inputs(wd, sysd, g0w0=True, evgw=False, test=True, maxSCF=256, thresh=1e-5, max_diis=5, lin=False,
                 eta=0.0, TDA_W=False, reg=False, eri_mo=None, bse=False, uhf=False, nel=None, use_gpu=True,
                 ncore=0, nryd=0, qsgw=False, bse2=False, dbse=False, dtda=False):
    """sysd: a dict from quack_amd.intgen.system(); eri_mo: MO integrals (nOrb^4, flat);
    bse, bse2, dbse, dtda: the phBSE, phBSE2, dBSE and dTDA flags of the
    "# BSE:" options line (read_options.f90:300-315); uhf: the
    methods file selects UHF instead of RHF (the unrestricted GW branch);
    nel: (nEla, nElb) for input/molecule (default: the closed shell); use_gpu:
    the GPU switch of input/hpc_flags (its fourth line, read_hpc_flags.f90:31-33),
    None to leave the file out (QuAcK then takes both switches as F); ncore,
    nryd: core and Rydberg electrons of input/molecule (even; QuAcK freezes
    ncore/2 and nryd/2 orbitals per spin, utils/read_molecule.f90:43-53)."""
    for d in ("input", "int", "test"):
        os.makedirs(os.path.join(wd, d), exist_ok=True)
    mol = sysd["mol"]
    with open(os.path.join(wd, "input", "methods"), "w") as f:
        f.write(METHODS.format(g0w0=tf(g0w0), evgw=tf(evgw), qsgw=tf(qsgw), test=tf(test), rhf=tf(not uhf),
                               uhf=tf(uhf)))
    with open(os.path.join(wd, "input", "options"), "w") as f:
        f.write(OPTIONS.format(maxscf=maxSCF, thresh="%.10g" % thresh, diis=max_diis, lin=tf(lin), eta="%.10g" % eta,
                               tda=tf(TDA_W), reg=tf(reg), bse=tf(bse), bse2=tf(bse2), dbse=tf(dbse),
                               dtda=tf(dtda)))
    hpc = os.path.join(wd, "input", "hpc_flags")
    if use_gpu is None:
        if os.path.exists(hpc):
            os.remove(hpc)
    else:
        with open(hpc, "w") as f:
            f.write("# if True (T), switch to HPC mode\n  F\n# if True (T), use GPU\n  %s\n" % tf(use_gpu))
    with open(os.path.join(wd, "input", "molecule"), "w") as f:
        nea, neb = nel if nel is not None else (mol.nO, mol.nO)
        f.write("# nAt nEla nElb nCore nRyd\n%d %d %d %d %d\n# Znuc x  y  z\n"
                % (len(mol.symbols), nea, neb, ncore, nryd))
        for s, r in zip(mol.symbols, mol.xyz):
            f.write("%s %.17g %.17g %.17g\n" % (s, r[0], r[1], r[2]))
    with open(os.path.join(wd, "int", "ENuc.dat"), "w") as f:
        f.write("%.17g \n" % float(sysd["ENuc"]))
    with open(os.path.join(wd, "int", "nBas.dat"), "w") as f:
        f.write(" %d " % sysd["nBas"])
    np.ascontiguousarray(sysd["eri"], dtype=np.float64).tofile(os.path.join(wd, "int", "ERI.bin"))
    with open(os.path.join(wd, "int", "MO.dat"), "w") as f:
        f.write("%d\n%.17g\n" % (sysd["nOrb"], sysd["ERHF"]))
        for e in sysd["eHF"]:
            f.write("%.17g\n" % e)
    if eri_mo is not None:
        np.ascontiguousarray(eri_mo, dtype=np.float64).tofile(os.path.join(wd, "int", "ERI_MO.bin"))
    if qsgw:
        from .scf import orthogonalization_matrix
        S, T, V = mol.one_electron()
        for name, M in (("Ov", S), ("Kin", T), ("Nuc", V)):
            write_matrix(os.path.join(wd, "int", name + ".dat"), M)
        np.asfortranarray(orthogonalization_matrix(S)).ravel(order="F").tofile(os.path.join(wd, "int", "X.bin"))
        np.asfortranarray(sysd["c"], dtype=np.float64).ravel(order="F").tofile(os.path.join(wd, "int", "cHF.bin"))


def write_matrix(path, M):
    """A symmetric matrix as PyDuck's "mu nu value" lines, 1-based, mu <= nu
    (utils/read_1e_integrals.f90:43-48 fills both triangles)."""
    n = M.shape[0]
    with open(path, "w") as f:
        for nu in range(n):
            for mu in range(nu + 1):
                f.write("%d %d %.17g\n" % (mu + 1, nu + 1, M[mu, nu]))


def s8_pairs(nBas):
    """0-based (a, b), a >= b, in packed pair order: pair(a, b) = a (a + 1) / 2 + b."""
    a, b = np.tril_indices(nBas)
    return a, b


def pack_eri_s8(eri, nBas):
    """The 8-fold packed chemists' array of int/ERI_chem.bin (PyDuck.py:318-326,
    pyscf's s8 order) from the full AO array in ERI.bin order
    (flat index mu + N la + N^2 nu + N^3 si holds (mu nu|la si)): element
    (ab|cd), a >= b, c >= d, pair(a, b) >= pair(c, d), sits at
    pair(pair(a, b), pair(c, d)) -- the Yoshimine index of
    utils/utils.f90:1186-1217 minus one."""
    N = nBas
    f = np.asarray(eri, dtype=np.float64).reshape(N, N, N, N)  # f[si, nu, la, mu] = (mu nu|la si)
    a, b = s8_pairs(N)
    P = a.size
    i, j = np.tril_indices(P)  # pair indices ij >= kl
    return np.ascontiguousarray(f[b[j], b[i], a[j], a[i]])


def unpack_eri_s8(packed, nBas):
    """The full AO array in ERI.bin order from the packed one (what the device
    kernel k_unpack_s8 does), for host-side readers of int/ERI_chem.bin."""
    N = nBas
    mu, la, nu, si = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), np.arange(N), indexing="ij")

    def pair(x, y):
        hi, lo = np.maximum(x, y), np.minimum(x, y)
        return hi * (hi + 1) // 2 + lo

    idx = pair(pair(mu, nu), pair(la, si))
    # axes (mu, la, nu, si) -> flat order mu + N la + N^2 nu + N^3 si
    return np.ascontiguousarray(np.asarray(packed)[idx].transpose(3, 2, 1, 0)).reshape(-1)


def write_eri_chem(wd, eri, nBas):
    """int/ERI_chem.bin as PyDuck writes it (raw float64 stream)."""
    os.makedirs(os.path.join(wd, "int"), exist_ok=True)
    path = os.path.join(wd, "int", "ERI_chem.bin")
    pack_eri_s8(eri, nBas).tofile(path)
    return path


def read_eri_chem(wd, nBas):
    """The packed array of int/ERI_chem.bin (utils/read_2e_integrals.f90:85-107)."""
    P = nBas * (nBas + 1) // 2
    x = np.fromfile(os.path.join(wd, "int", "ERI_chem.bin"), dtype=np.float64)
    if x.size != P * (P + 1) // 2:
        raise ValueError("int/ERI_chem.bin holds %d values, expected %d for nBas = %d" % (x.size, P * (P + 1) // 2,
                                                                                          nBas))
    return x


def read_test_dump_pairs(path):
    """The same dump as (description, value) pairs in file order (the
    unrestricted drivers write some descriptions twice, once per spin)."""
    lines = [ln.strip() for ln in open(path) if ln.strip()]
    return [(lines[k], float(lines[k + 1])) for k in range(0, len(lines) - 1, 2)]


def read_test_dump(path):
    """Parse QuAcK's test dump (description line, value line), src/test/dump_test_value.f90:19-20."""
    out = {}
    lines = [ln.strip() for ln in open(path) if ln.strip()]
    for k in range(0, len(lines) - 1, 2):
        out[lines[k]] = float(lines[k + 1])
    return out
