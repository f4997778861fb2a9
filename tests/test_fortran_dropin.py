"""The Fortran drop-in (RG0W0_hip / evRGW_hip behind QuAcK's signatures) from
QuAcK-format input files, checked against QuAcK's own test values."""
import json
import os
import subprocess

import pytest

from quack_amd import quackio

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "quack_amd", "fortran", "qk_gw_driver")
GOLD = json.load(open(os.path.join(ROOT, "tests", "golden", "quack_goldens.json")))


def test_driver_built_and_linked():
    assert os.path.exists(DRIVER), "run __graft_entry__.build()"
    out = subprocess.run(["ldd", DRIVER], capture_output=True, text=True).stdout
    assert "libquack_amd.so" in out and "not found" not in out


def test_input_files_roundtrip(tmp_path, systems):
    s = systems["He"]
    quackio.write_inputs(str(tmp_path), s, g0w0=True, evgw=True, max_diis=1)
    txt = open(tmp_path / "input" / "options").read().splitlines()
    k = [i for i, ln in enumerate(txt) if ln.startswith("# GW:")][0]
    fields = txt[k + 1].split()
    assert len(fields) == 14 and fields[2] == "1" and fields[3] == "F"
    m = open(tmp_path / "input" / "methods").read().splitlines()
    k = [i for i, ln in enumerate(m) if ln.startswith("# G0W0")][0]
    assert m[k + 1].split()[:2] == ["T", "T"]


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["He", "H2O"])
def test_driver_reproduces_quack_goldens(tmp_path, mo_systems, name):
    print("fortran driver", name, flush=True)
    s = mo_systems[name]
    quackio.write_inputs(str(tmp_path), s, g0w0=True, evgw=True, maxSCF=256, thresh=1e-5, max_diis=1,
                         eri_mo=s["eri_mo"])
    r = subprocess.run([DRIVER, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "G0W0@RHF HOMO      energy" in r.stdout and "evGW@RHF HOMO      energy" in r.stdout
    dump = quackio.read_test_dump(str(tmp_path / "test" / "Rtest.dat"))
    g = GOLD[name]["feather"]
    for key in ("G0W0 correlation energy", "G0W0 HOMO energy", "G0W0 LUMO energy", "evGW correlation energy",
                "evGW HOMO energy", "evGW LUMO energy"):
        assert abs(dump[key] - g[key]) < 1e-7 * abs(g[key]), key


@pytest.mark.gpu
def test_driver_phbse_matches_oracle(tmp_path, mo_systems):
    """phBSE switched on in input/options: the drop-ins run the static BSE
    after G0W0 and evGW and print QuAcK's Tr@BSE lines, which match the CPU
    restatement of RGW_phBSE at the printed precision."""
    import numpy as np

    from oracle import gw_oracle as go
    print("fortran driver phBSE H2O", flush=True)
    s = mo_systems["H2O"]
    quackio.write_inputs(str(tmp_path), s, g0w0=True, evgw=True, maxSCF=256, thresh=1e-5, max_diis=1,
                         eri_mo=s["eri_mo"], bse=True)
    r = subprocess.run([DRIVER, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    vals = {}
    for ln in r.stdout.splitlines():
        for tag in ("G0W0", "evGW"):
            for man in ("singlet", "triplet"):
                key = "Tr@BSE@%s@RHF correlation energy (%s) =" % (tag, man)
                if key in ln:
                    vals[(tag, man)] = float(ln.split("=")[1].split()[0])
    assert len(vals) == 4, r.stdout[-3000:]
    nO, mo = s["nO"], s["eri_mo"]
    eG0W0 = go.rg0w0(mo, s["eHF"], nO)["eGW"]
    Ec, _, _ = go.rgw_phbse(mo, s["eHF"], eG0W0, nO)
    assert abs(vals[("G0W0", "singlet")] - Ec[0]) < 2e-10 and abs(vals[("G0W0", "triplet")] - Ec[1]) < 2e-10
    # the excitation tables of LR/print_excitation_energies.f90 (method and manifold in A15 fields)
    assert "phBSE@GW@RHF" in r.stdout and "singlet manifold" in r.stdout and "triplet manifold" in r.stdout
    assert np.isfinite(vals[("evGW", "singlet")])


@pytest.mark.gpu
def test_driver_phbse2_dynamic(tmp_path, mo_systems):
    """phBSE with the second-order kernel and the dynamical correction
    switched on in input/options (phBSE2, dBSE, dTDA F): the drop-in prints
    the Tr@BSE energies and the dynamical-correction tables, which match the
    CPU restatement at the printed precision."""
    import numpy as np

    from oracle import gw_oracle as go
    print("fortran driver phBSE2 + dBSE H2O", flush=True)
    s = mo_systems["H2O"]
    nO, mo, e = s["nO"], s["eri_mo"], s["eHF"]
    quackio.write_inputs(str(tmp_path), s, g0w0=True, evgw=False, eta=0.1, eri_mo=mo, bse=True, bse2=True, dbse=True,
                         dtda=False)
    r = subprocess.run([DRIVER, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "Second-order BSE static kernel activated" in r.stdout
    eGW = go.rg0w0(mo, e, nO, eta=0.1)["eGW"]
    dyn = {}
    Ec, OmS, OmT = go.rgw_phbse(mo, e, eGW, nO, eta=0.1, dophBSE2=True, dBSE=True, dTDA=False, dyn=dyn)
    vals = {}
    for ln in r.stdout.splitlines():
        for man in ("singlet", "triplet"):
            key = "Tr@BSE@G0W0@RHF correlation energy (%s) =" % man
            if key in ln:
                vals[man] = float(ln.split("=")[1].split()[0])
    assert abs(vals["singlet"] - Ec[0]) < 2e-10 and abs(vals["triplet"] - Ec[1]) < 2e-10
    # the two dynamical-correction tables: state, static, dynamic, correction (eV), Z
    lines = r.stdout.splitlines()
    tables = [k for k, ln in enumerate(lines) if "First-order dynamical correction" in ln]
    assert len(tables) == 2
    HA = 27.211386245988
    for t0, (man, Om) in zip(tables, (("singlet", OmS), ("triplet", OmT))):
        # the rows: the first ten lines after the title that start with a state number
        # (the runtime may wrap the 99-character rules of list-directed output)
        rows = [ln.split() for ln in lines[t0 + 1:] if ln.split() and ln.split()[0].isdigit()][:10]
        od, zd = dyn[man]
        for k, row in enumerate(rows):
            assert int(row[0]) == k + 1
            assert abs(float(row[1]) - Om[k] * HA) < 2e-6 and abs(float(row[3]) - od[k] * HA) < 2e-6
            assert abs(float(row[4]) - zd[k]) < 2e-6


@pytest.mark.gpu
def test_driver_unrestricted_phbse(tmp_path, mo_systems):
    """phBSE switched on in the UHF branch: UG0W0_hip and evUGW_hip run the
    static unrestricted BSE (spin-conserved and spin-flip, the "# spin:" line's
    answers) and print QuAcK's Tr@BSE lines, which match the CPU restatement
    at the printed precision."""
    import numpy as np

    from oracle import gw_oracle as go
    print("fortran driver UHF phBSE H2O", flush=True)
    s = mo_systems["H2O"]
    nO, mo, e = s["nO"], s["eri_mo"], s["eHF"]
    quackio.write_inputs(str(tmp_path), s, g0w0=True, evgw=True, maxSCF=64, thresh=1e-5, max_diis=5, eri_mo=mo,
                         uhf=True, bse=True)
    r = subprocess.run([DRIVER, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "phBSE@GW@UHF" in r.stdout and "spin-flip" in r.stdout
    e2 = np.stack([e, e], axis=1)
    occ = (nO, nO)
    g0 = go.ug0w0(mo, mo, mo, e2, occ)["eGW"]
    ev = go.evugw(mo, mo, mo, e2, occ, maxSCF=64)["eGW"]
    for tag, eW, eG in (("G0W0", e2, g0), ("evGW", ev, ev)):
        Ec, _, _ = go.ugw_phbse(mo, mo, mo, eW, eG, occ)
        vals = {}
        for ln in r.stdout.splitlines():
            for man in ("spin-conserved", "spin-flip"):
                # evUGW's A50 label field cuts the label before ") =" (GW/evUGW.f90:250-251)
                key = "Tr@BSE@%s@UHF correlation energy (%s" % (tag, man)
                if key in ln:
                    vals[man] = float(ln.split()[-2])
        assert abs(vals["spin-conserved"] - Ec[0]) < 2e-9 and abs(vals["spin-flip"] - Ec[1]) < 2e-9, (tag, vals, Ec)


def test_input_files_uhf_flag(tmp_path, systems):
    """The unrestricted branch is selected by the HF line of input/methods and
    the electron counts of input/molecule, as QuAcK reads them."""
    s = systems["H2O"]
    quackio.write_inputs(str(tmp_path), s, uhf=True, nel=(s["nO"] + 1, s["nO"]))
    m = open(tmp_path / "input" / "methods").read().splitlines()
    assert m[0].startswith("# RHF UHF") and m[1].split()[:2] == ["F", "T"]
    mol = open(tmp_path / "input" / "molecule").read().splitlines()
    assert mol[1].split()[1:3] == [str(s["nO"] + 1), str(s["nO"])]


@pytest.mark.gpu
def test_driver_unrestricted(tmp_path, mo_systems):
    """UG0W0_hip and evUGW_hip through the driver's UHF branch: a closed shell
    reproduces QuAcK's restricted goldens (UHF = RHF there), and one more
    alpha electron matches the CPU restatement of UG0W0/evUGW on the same
    orbitals; the values come from test/Utest.dat in QuAcK's dump format."""
    import numpy as np

    from oracle import gw_oracle as go
    s = mo_systems["H2O"]
    nO, mo = s["nO"], s["eri_mo"]
    print("fortran driver UHF H2O closed shell", flush=True)
    wd = tmp_path / "closed"
    quackio.write_inputs(str(wd), s, g0w0=True, evgw=True, maxSCF=256, thresh=1e-5, max_diis=1, eri_mo=mo, uhf=True)
    r = subprocess.run([DRIVER, str(wd)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "G0W0@UHF HOMO      energy" in r.stdout and "evGW@UHF HOMO      energy" in r.stdout
    d = quackio.read_test_dump_pairs(str(wd / "test" / "Utest.dat"))
    g = GOLD["H2O"]["feather"]
    for tag, k0 in (("G0W0", 0), ("evGW", 5)):
        assert d[k0][0] == tag + " correlation energy"
        assert abs(d[k0][1] - g[tag + " correlation energy"]) < 1e-7 * abs(g[tag + " correlation energy"])
        for k, key in ((1, "HOMO"), (2, "LUMO"), (3, "HOMO"), (4, "LUMO")):
            ref = g["%s %s energy" % (tag, key)]
            assert abs(d[k0 + k][1] - ref) < 1e-7 * abs(ref), (tag, k)
    print("fortran driver UHF H2O open shell", flush=True)
    wd = tmp_path / "open"
    occ = (nO + 1, nO)
    quackio.write_inputs(str(wd), s, g0w0=True, evgw=True, maxSCF=64, thresh=1e-5, max_diis=5, eri_mo=mo, uhf=True,
                         nel=occ)
    r = subprocess.run([DRIVER, str(wd)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    d = quackio.read_test_dump_pairs(str(wd / "test" / "Utest.dat"))
    e2 = np.stack([s["eHF"], s["eHF"]], axis=1)
    for k0, o in ((0, go.ug0w0(mo, mo, mo, e2, occ)), (5, go.evugw(mo, mo, mo, e2, occ, maxSCF=64, max_diis=5))):
        ref = [o["EcRPA"], o["eGW"][occ[0] - 1, 0], o["eGW"][occ[0], 0], o["eGW"][occ[1] - 1, 1], o["eGW"][occ[1], 1]]
        for k in range(5):
            assert abs(d[k0 + k][1] - ref[k]) < 1e-8, (k0, k, d[k0 + k], ref[k])


@pytest.mark.gpu
def test_driver_frozen_core(tmp_path, mo_systems):
    """Core and Rydberg electrons in input/molecule (2 and 4: one frozen core
    and two frozen virtual orbitals per spin, utils/read_molecule.f90:43-53)
    reach RG0W0_hip / evRGW_hip and UG0W0_hip / evUGW_hip as QuAcK's nC, nR;
    the dumped values match the CPU restatement with the same counts."""
    import numpy as np

    from oracle import gw_oracle as go
    s = mo_systems["H2O"]
    nO, mo, e = s["nO"], s["eri_mo"], s["eHF"]
    print("fortran driver frozen core RHF H2O", flush=True)
    wd = tmp_path / "rhf"
    quackio.write_inputs(str(wd), s, g0w0=True, evgw=True, maxSCF=64, thresh=1e-6, max_diis=5, eta=0.1, eri_mo=mo,
                         ncore=2, nryd=4)
    r = subprocess.run([DRIVER, str(wd)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    dump = quackio.read_test_dump(str(wd / "test" / "Rtest.dat"))
    for tag, o in (("G0W0", go.rg0w0(mo, e, nO, eta=0.1, nC=1, nR=2)),
                   ("evGW", go.evrgw(mo, e, nO, maxSCF=64, thresh=1e-6, max_diis=5, eta=0.1, nC=1, nR=2))):
        assert abs(dump[tag + " correlation energy"] - o["EcRPA"]) < 1e-8, tag
        assert abs(dump[tag + " HOMO energy"] - o["eGW"][nO - 1]) < 1e-8, tag
        assert abs(dump[tag + " LUMO energy"] - o["eGW"][nO]) < 1e-8, tag
    print("fortran driver frozen core UHF H2O open shell", flush=True)
    wd = tmp_path / "uhf"
    occ = (nO + 1, nO)
    quackio.write_inputs(str(wd), s, g0w0=True, evgw=True, maxSCF=64, thresh=1e-5, max_diis=5, eri_mo=mo, uhf=True,
                         nel=occ, ncore=2, nryd=4)
    r = subprocess.run([DRIVER, str(wd)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    d = quackio.read_test_dump_pairs(str(wd / "test" / "Utest.dat"))
    e2 = np.stack([e, e], axis=1)
    fz = {"nC": (1, 1), "nR": (2, 2)}
    for k0, o in ((0, go.ug0w0(mo, mo, mo, e2, occ, **fz)), (5, go.evugw(mo, mo, mo, e2, occ, maxSCF=64, **fz))):
        ref = [o["EcRPA"], o["eGW"][occ[0] - 1, 0], o["eGW"][occ[0], 0], o["eGW"][occ[1] - 1, 1], o["eGW"][occ[1], 1]]
        for k in range(5):
            assert abs(d[k0 + k][1] - ref[k]) < 1e-8, (k0, k, d[k0 + k], ref[k])


@pytest.mark.gpu
def test_driver_qsgw(tmp_path, mo_systems):
    """qsRGW_hip through the driver (qsGW switched on in input/methods, SRG,
    with static phBSE after it): the dumped qsGW values and the Tr@BSE@qsGW
    lines match the CPU restatement of qsRGW and RGW_phBSE (the BSE on the MO
    integrals of the last cycle's orbitals, as qsRGW leaves them)."""
    import numpy as np

    from oracle import gw_oracle as go
    from quack_amd import scf
    print("fortran driver qsGW H2O", flush=True)
    s = mo_systems["H2O"]
    nO, mo = s["nO"], s["eri_mo"]
    quackio.write_inputs(str(tmp_path), s, g0w0=False, evgw=False, qsgw=True, maxSCF=64, thresh=1e-5, max_diis=5,
                         reg=True, eri_mo=mo, bse=True)
    r = subprocess.run([DRIVER, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "qsGW@RHF HOMO      energy" in r.stdout and "qsRGW        energy" in r.stdout
    S, T, V = s["mol"].one_electron()
    X = scf.orthogonalization_matrix(S)
    c = np.asarray(s["c"])
    o = go.qsrgw(S, T, V, s["eri"], X, c, 2.0 * c[:, :nO] @ c[:, :nO].T, s["eHF"], nO, doSRG=True)
    assert o["converged"]
    dump = quackio.read_test_dump(str(tmp_path / "test" / "Rtest.dat"))
    assert abs(dump["qsGW correlation energy"] - o["EcRPA"]) < 1e-8
    assert abs(dump["qsGW HOMO energy"] - o["eGW"][nO - 1]) < 1e-8
    assert abs(dump["qsGW LUMO energy"] - o["eGW"][nO]) < 1e-8
    vals = {}
    for ln in r.stdout.splitlines():
        for man in ("singlet", "triplet"):
            key = "Tr@BSE@qsGW@RHF correlation energy (%s) =" % man
            if key in ln:
                vals[man] = float(ln.split("=")[1].split()[0])
    mo_last = go.aotomo_eri_rhf(o["c_last_cycle"], s["eri"])
    Ec, _, _ = go.rgw_phbse(mo_last, o["eGW"], o["eGW"], nO)
    assert abs(vals["singlet"] - Ec[0]) < 1e-8 and abs(vals["triplet"] - Ec[1]) < 1e-8


def test_input_files_qsgw(tmp_path, systems):
    """The qsGW inputs: the methods switch, PyDuck's "mu nu value" lines of the
    one-electron integrals (read back as QuAcK's read_1e_integrals does) and
    the X and cHF arrays."""
    import numpy as np
    s = systems["H2O"]
    quackio.write_inputs(str(tmp_path), s, g0w0=False, qsgw=True)
    m = open(tmp_path / "input" / "methods").read().splitlines()
    k = [i for i, ln in enumerate(m) if ln.startswith("# G0W0")][0]
    assert m[k + 1].split()[:3] == ["F", "F", "T"]
    S, T, V = s["mol"].one_electron()
    n = S.shape[0]
    for name, M in (("Ov", S), ("Kin", T), ("Nuc", V)):
        R = np.zeros((n, n))
        for ln in open(tmp_path / "int" / (name + ".dat")):
            mu, nu, v = ln.split()
            R[int(mu) - 1, int(nu) - 1] = R[int(nu) - 1, int(mu) - 1] = float(v)
        U = np.triu(M)
        assert np.array_equal(R, U + np.triu(M, 1).T), name  # the upper triangle, exactly
        assert np.abs(R - M).max() < 1e-14, name
    c = np.fromfile(tmp_path / "int" / "cHF.bin").reshape(s["c"].shape, order="F")
    assert np.array_equal(c, np.asarray(s["c"]))


def test_driver_odd_core_electrons_rejected(tmp_path, systems):
    """An odd number of core electrons stops the driver before any device
    call, as QuAcK's read_molecule does."""
    s = systems["He"]
    quackio.write_inputs(str(tmp_path), s, ncore=1)
    r = subprocess.run([DRIVER, str(tmp_path)], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "must be even" in r.stdout


@pytest.mark.parametrize("use_gpu", [False, None])
def test_driver_gpu_switch_off(tmp_path, systems, use_gpu):
    """input/hpc_flags with the GPU switch F (or no file, which QuAcK's
    read_hpc_flags.f90 reads as F): the run belongs to QuAcK's CPU RG0W0 /
    evRGW, so the driver prints why and stops with code 3 before any device
    call (the patched dispatch of src/GW/RGW.f90:85,102).  The T case is every
    GPU test of this file."""
    s = systems["He"]
    quackio.write_inputs(str(tmp_path), s, g0w0=True, evgw=True, max_diis=1, use_gpu=use_gpu)
    r = subprocess.run([DRIVER, str(tmp_path)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 3, (r.returncode, r.stdout[-2000:], r.stderr[-2000:])
    assert "use_gpu is F" in r.stdout
    assert not os.path.exists(tmp_path / "test" / "Rtest.dat")
