"""The reference's own program run with the device drop-ins: oracle/_ref/
static_tbk (its Fortran objects linked against bande_/fermi_ of
lib/libtbk_fortran.a) on the ex1 inputs writes the published SKENG, and on
ex2 (hcp Ti, mesh line aligned) the SKENG of the unmodified program;
oracle/_ref/static_diag (the reference's own bande and k-loop, with diag_ and
perturb_ from the archive: one device solve per k) writes the published
ex1 SKENG too; tbk_perturb against the oracle's perturb."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

import ex2_case
import oracle
from conftest import ROOT
from tbk.api import perturb

pytestmark = pytest.mark.gpu

EXE = os.path.join(ROOT, "oracle", "_ref", "static_tbk")
INPUTS = os.path.join(ROOT, "oracle", "_ref", "ex1")
GOLD = os.path.join(ROOT, "tests", "golden", "ex1.json")
INPUTS2 = os.path.join(ROOT, "oracle", "_ref", "ex2")
GOLD2 = os.path.join(ROOT, "tests", "golden", "ex2.json")


@pytest.mark.skipif(not os.path.exists(EXE), reason="drop-in program not built (oracle/Makefile dropin)")
def test_reference_program_with_device_bande(tmp_path):
    work = tmp_path / "ex1"
    shutil.copytree(INPUTS, work)
    print("case: static_tbk on ex1 inputs", flush=True)
    r = subprocess.run("ulimit -s unlimited; " + EXE, shell=True, cwd=work, capture_output=True,
                       text=True, timeout=300, env=dict(os.environ, TBK_DEVICE="0"))
    assert r.returncode == 0, (r.stdout[-1000:], r.stderr[-2000:])
    rows = [line.split() for line in open(work / "SKENG") if line.strip()]
    gold = json.load(open(GOLD))["structures"]
    assert len(rows) == len(gold)
    worst = 0.0
    for row, g in zip(rows, gold):
        vol, ef, et, pt = (float(x) for x in row[2:6])
        assert row[0] == g["lattice"]
        for got, ref in ((vol, g["volume"]), (ef, g["efermi"]), (et, g["et"]), (pt, g["pt"])):
            worst = max(worst, abs(got - ref))
    print(f"  worst |SKENG - published| = {worst:.1e} (9 printed decimals)", flush=True)
    assert worst <= 1e-9 * 1.5


@pytest.mark.skipif(not os.path.exists(EXE) or not os.path.isdir(INPUTS2),
                    reason="drop-in program not built (oracle/Makefile dropin)")
def test_reference_program_with_device_bande_ex2_hcp(tmp_path):
    """Two atoms of nine orbitals (N = 18), 133 irreducible k of a 12x12x12
    mesh, 13 c/a values, mode 3 (E_F and E(T), no pressure): the SKENG rows
    agree with the reference program's to their 9 printed decimals."""
    work = tmp_path / "ex2"
    shutil.copytree(INPUTS2, work)
    ex2_case.write_skin(work)
    print("case: static_tbk on ex2 inputs (hcp Ti, mesh line aligned)", flush=True)
    r = subprocess.run("ulimit -s unlimited; " + EXE, shell=True, cwd=work, capture_output=True,
                       text=True, timeout=300, env=dict(os.environ, TBK_DEVICE="0"))
    assert r.returncode == 0, (r.stdout[-1000:], r.stderr[-2000:])
    rows = ex2_case.read_skeng(work / "SKENG")
    gold = json.load(open(GOLD2))["structures"]
    assert len(rows) == len(gold)
    worst = 0.0
    for (label, vol, ef, et), g in zip(rows, gold):
        assert label == g["label"] and vol == g["volume"]
        for got, ref in ((ef, g["efermi"]), (et, g["et"])):
            worst = max(worst, abs(got - ref))
    print(f"  worst |SKENG - reference program| = {worst:.1e} (9 printed decimals)", flush=True)
    assert worst <= 1e-9 * 1.5


EXE_DIAG = os.path.join(ROOT, "oracle", "_ref", "static_diag")


@pytest.mark.skipif(not os.path.exists(EXE_DIAG), reason="drop-in program not built (oracle/Makefile dropin)")
def test_reference_program_with_device_diag_and_perturb(tmp_path):
    work = tmp_path / "ex1"
    shutil.copytree(INPUTS, work)
    print("case: static_diag (device diag_/perturb_ per k) on ex1 inputs", flush=True)
    r = subprocess.run("ulimit -s unlimited; " + EXE_DIAG, shell=True, cwd=work, capture_output=True,
                       text=True, timeout=300, env=dict(os.environ, TBK_DEVICE="0"))
    assert r.returncode == 0, (r.stdout[-1000:], r.stderr[-2000:])
    rows = [line.split() for line in open(work / "SKENG") if line.strip()]
    gold = json.load(open(GOLD))["structures"]
    assert len(rows) == len(gold)
    worst = 0.0
    for row, g in zip(rows, gold):
        vol, ef, et, pt = (float(x) for x in row[2:6])
        assert row[0] == g["lattice"]
        for got, ref in ((vol, g["volume"]), (ef, g["efermi"]), (et, g["et"]), (pt, g["pt"])):
            worst = max(worst, abs(got - ref))
    print(f"  worst |SKENG - published| = {worst:.1e} (9 printed decimals)", flush=True)
    assert worst <= 1e-9 * 1.5


def _packed_upper(a):
    n = a.shape[0]
    return np.array([a[i, j] for j in range(n) for i in range(j + 1)], np.complex128)


@pytest.mark.parametrize("n,degen", [(1, False), (7, False), (40, False), (40, True)])
def test_device_perturb_matches_oracle(n, degen):
    """tbk_perturb (perturb.f on the device) against the oracle's perturb on
    random pencils; with degen, every eigenvalue three times (groups)."""
    print(f"case: tbk_perturb n={n} degenerate={degen}", flush=True)
    rng = np.random.default_rng(7 + n)

    def herm(m):
        a = rng.standard_normal((m, m)) + 1j * rng.standard_normal((m, m))
        return a + a.conj().T

    if degen:
        b = herm(n // 3 + 1)
        h = np.kron(np.eye(3), b)[:n, :n]
        s = np.eye(n)
    else:
        h = herm(n)
        g = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
        s = g @ g.conj().T / n + np.eye(n)
    info, e, z = oracle.zhegv(h, s)
    assert info == 0
    dh, ds = herm(n), herm(n) * 0.1
    ref = oracle.perturb(_packed_upper(dh), _packed_upper(ds), z, e, eigcut=1e3)
    got = perturb(dh, ds, z, e, eigcut=1e3)
    scale = max(1.0, float(np.abs(ref).max()))
    print(f"  ntot {len(got)} vs {len(ref)}, max|d deig| = {np.abs(got - ref).max():.2e}", flush=True)
    assert len(got) == len(ref)
    assert np.abs(got - ref).max() < 1e-11 * scale


@pytest.mark.skipif(not os.path.exists(EXE) or not os.path.isdir(INPUTS) or not os.path.isdir(INPUTS2),
                    reason="drop-in program not built (oracle/Makefile dropin)")
def test_reference_program_with_device_bande_twenty_spins(tmp_path):
    """A 20-atom non-collinear cell whose atoms all point different ways
    (tests/skin_cases.py noncol20): the reference program through the device
    bande_ (every atom's spinat through the drop-in) writes the SKENG line
    of the unmodified program (tests/golden/static.json)."""
    import skin_cases as SC
    want = json.load(open(os.path.join(ROOT, "tests", "golden", "static.json")))["noncol20"]["skeng"]
    SC.build_static_case("noncol20", tmp_path, INPUTS, INPUTS2)
    print("case: static_tbk on a 20-atom non-collinear cell", flush=True)
    r = subprocess.run("ulimit -s unlimited; " + EXE, shell=True, cwd=tmp_path, capture_output=True, text=True,
                       timeout=300, env=dict(os.environ, TBK_DEVICE="0"))
    assert r.returncode == 0, (r.stdout[-1000:], r.stderr[-2000:])
    got = [l.rstrip("\n") for l in open(tmp_path / "SKENG") if l.strip()]
    assert len(got) == len(want) == 1
    a, b = got[0].split(), want[0].split()
    assert a[:2] == b[:2]
    worst = max(abs(float(x) - float(y)) for x, y in zip(a[2:], b[2:]))
    print(f"  {got[0]}\n  {want[0]}\n  worst |SKENG - reference program| = {worst:.1e}", flush=True)
    assert worst <= 1.01e-9
