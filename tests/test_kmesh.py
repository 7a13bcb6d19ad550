"""k-point meshes (csrc/tbk_kmesh.cpp: kptin's reciprocal lattice and
weights, rdspgp's operations, kptgen's irreducible mesh; the Fortran
drop-in kptgen_).

Pinned by
* the reference's own kptin/rdspgp/kptgen (oracle/_ref, compiled from its
  sources where the tree is present): lattice, operations, k-points and
  weights bit-identical, and the reference program linked against the
  drop-in kptgen_ writing the same SKOUT and SKENG;
* the fixture those routines wrote (tests/golden/kmesh.json), everywhere;
* Monkhorst & Pack, Phys. Rev. B 13, 5188 (1976): 2, 10, 28, 60 special
  points for the fcc lattice at q = 2, 4, 6, 8;
* star weights by brute force: every irreducible point's weight is the
  share of the half-open mesh cell in its orbit under the operations and
  time reversal;
* a_i . b_j = 2 pi delta_ij, the volume, integer operations closed under
  products;
* sums over the irreducible mesh equal those over the full mesh (E_F,
  E(T), P(T) of the ex1 Pd structure through the oracle).
"""
import json
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

import ex2_case
import kmesh_cases as KC
import oracle
from conftest import ROOT
from tbk import _lib, kmesh
from tbk import structures as S

GOLD = os.path.join(ROOT, "tests", "golden", "kmesh.json")
EX1_JSON = os.path.join(ROOT, "tests", "golden", "ex1.json")
EX2_JSON = os.path.join(ROOT, "tests", "golden", "ex2.json")
REF = "/root/reference/NRLtb/ex1"
ORACLE = os.path.join(ROOT, "oracle")

needs_ref = pytest.mark.skipif(not oracle.ref_available(), reason="reference build (oracle/_ref) not present")
CASES = KC.cases()
IDS = [c[0] for c in CASES]


def ours(case, tmp_path):
    name, lat, lattic, ilat, rot, tnp, shift, kdiv, lspec, kpst, kpend, info, sfk = case
    gpath = str(tmp_path / f"{name}.grp")
    KC.write_group(gpath, name, rot, tnp, ilat, shift or (0.0, 0.0, 0.0))
    L = kmesh.reciprocal_lattice(lat, lattic)
    g = kmesh.read_space_group(gpath, L, ilat)
    q, w = kmesh.kptgen(L, g, kdiv, lspec, info=info, sfk=sfk, kpst=kpst, kpend=kpend)
    return L, g, q, w, gpath


@needs_ref
@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_bit_exact_vs_reference_kptin(case, tmp_path):
    name, lat, lattic, ilat, rot, tnp, shift, kdiv, lspec, kpst, kpend, info, sfk = case
    L, g, q, w, gpath = ours(case, tmp_path)
    r = oracle.ref_kptin(lat, KC.newsym_block(gpath, lattic, ilat, kdiv, lspec, kpst, kpend, info, sfk),
                         str(tmp_path))
    assert r["lattic"] == lattic
    for key, mine in (("avec", L.avec), ("bvec", L.bvec), ("bij", L.bij), ("rspc", g.rspc),
                      ("rspct", g.rspct), ("tnpspc", g.tnpspc)):
        assert np.array_equal(mine, r[key]), key
    assert L.wsvol == r["wsvol"]
    assert q.shape == r["q"].shape, (len(q), len(r["q"]))
    assert np.array_equal(q, r["q"]) and np.array_equal(w, r["w"])


@pytest.mark.skipif(not os.path.exists(GOLD), reason="fixture not written")
@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_matches_reference_fixture(case, tmp_path):
    gold = json.load(open(GOLD))["cases"][case[0]]
    L, g, q, w, _ = ours(case, tmp_path)
    assert np.array_equal(q, np.array(gold["q"]).reshape(-1, 3))
    assert np.array_equal(w, np.array(gold["w"]))
    assert np.array_equal(g.rspct, np.array(gold["rspct"]).reshape(-1, 3, 3))
    assert np.array_equal(g.tnpspc, np.array(gold["tnpspc"]).reshape(-1, 3))
    assert np.array_equal(L.bvec, np.array(gold["bvec"]).reshape(3, 3)) and L.wsvol == gold["wsvol"]


def test_monkhorst_pack_special_point_counts_fcc(tmp_path):
    counts = {}
    for case in CASES:
        m = re.fullmatch(r"fcc_mp(\d)", case[0])
        if m:
            _, _, q, w, _ = ours(case, tmp_path)
            counts[int(m.group(1))] = len(q)
            assert abs(w.sum() - 1.0) < 1e-14
    assert counts == {2: 2, 4: 10, 6: 28, 8: 60}


def full_mesh(lattic, kdiv, lspec):
    """The points kptgen.f:99-139 visits (lattices other than bcc and bct)
    and whether each lies in the half-open cell."""
    d1, d2, d3 = kdiv
    if lspec:
        m1 = d1 + d1 % 2 if lattic == 4 else d1 + 1
        m2 = m1 if lattic == 4 else d2 + 1
        m3 = d3 + 1
    else:
        m1, m2, m3 = d1, d2, d3
    pts, inside = [], []
    for ii in range(-m1, m1 + 1, 2):
        for jj in range(-m2, m2 + 1, 2):
            for kk in range(-m3, m3 + 1, 2):
                pts.append((ii / (2 * d1), jj / (2 * d2), kk / (2 * d3)))
                inside.append(-d1 < ii <= d1 and -d2 < jj <= d2 and -d3 < kk <= d3)
    return np.array(pts), np.array(inside)


def key(u):
    f = np.round(np.mod(u, 1.0) * 1e6).astype(np.int64) % 1000000
    return tuple(int(x) for x in f)


@pytest.mark.parametrize("case", [c for c in CASES if c[2] not in (2, 7, 8, 9) and c[9] == 0
                                  and c[11] is None], ids=lambda c: c[0])
def test_star_weights_by_brute_force(case, tmp_path):
    """Weight of each irreducible point = (points of the half-open cell in
    its orbit under {+-R}) / (points of the cell); orbits of distinct
    irreducible points are disjoint and together hold the whole cell."""
    name, lat, lattic, ilat, rot, tnp, shift, kdiv, lspec, *_ = case
    L, g, q, w, _ = ours(case, tmp_path)
    ops = np.rint(g.rspct)
    pts, inside = full_mesh(lattic, kdiv, lspec)
    kin = inside.sum()
    owner = {}
    for n, qn in enumerate(q):
        for R in ops:
            for sgn in (1.0, -1.0):
                k = key(sgn * (R @ qn))
                assert owner.get(k, n) == n, "two irreducible points in one orbit"
                owner[k] = n
    count = np.zeros(len(q))
    for p, ins in zip(pts, inside):
        n = owner.get(key(p))
        assert n is not None, "a mesh point outside every orbit"
        if ins:
            count[n] += 1
    assert np.allclose(w, count / kin, rtol=0, atol=1e-15)


def test_reciprocal_lattice_and_groups(tmp_path):
    for case in CASES:
        L, g, _, _, _ = ours(case, tmp_path)
        assert np.allclose(L.avec @ L.bvec.T, 2 * np.pi * np.eye(3), atol=1e-13)
        assert np.isclose(L.wsvol, abs(np.linalg.det(case[1])), rtol=1e-14)
        assert np.allclose(L.bij, L.bvec @ L.bvec.T, rtol=1e-15, atol=1e-15)
        ops = np.rint(g.rspct)
        assert np.abs(g.rspct - ops).max() < 1e-12
        keys = {tuple(m.ravel()) for m in ops}
        for a in ops:
            for b in ops:
                assert tuple((a @ b).ravel()) in keys
        for m in g.rspc:                                # Cartesian operations orthogonal
            assert np.allclose(m @ m.T, np.eye(3), atol=1e-9)


def test_hexagonal_entries_snapped(tmp_path):
    case = [c for c in CASES if c[0] == "hex_cartesian_special"][0]
    _, g, _, _, _ = ours(case, tmp_path)
    vals = np.abs(g.rspc.ravel())
    s = np.sqrt(3.0) / 2.0
    near = vals[np.abs(vals - s) < 1e-3]
    assert len(near) > 0 and np.all(near == s)       # 10-decimal input snapped to sqrt(3)/2


@needs_ref
def test_explicit_list_and_path_vs_reference(tmp_path):
    ex1 = json.load(open(EX1_JSON))
    raw = np.array(ex1["kpoints"]["fcc.08"])
    gpath = str(tmp_path / "oh.grp")
    KC.write_group(gpath, "Oh", KC.oh_cartesian(), np.zeros((48, 3)), False)
    kfile = tmp_path / "list.k"
    kfile.write_text(f"{len(raw):5d}{1.5:10.6f}\n" + "\n".join(" ".join(repr(float(x)) for x in r) for r in raw) + "\n")
    lines = ["NEWSYM=T", "LATTIC=1", gpath, "ILAT=F", "-1313", str(kfile)]
    r = oracle.ref_kptin(KC.fcc(), lines, str(tmp_path))
    q, w = kmesh.klist(raw, sfk=1.5)
    assert np.array_equal(q, r["q"]) and np.array_equal(w, r["w"])
    segs = [(11, (0.0, 0.0, 0.0), (0.5, -0.5, 0.5)), (7, (0.5, -0.5, 0.5), (0.25, 0.25, 0.25)),
            (1, (0.1, 0.2, 0.3), (0.3, 0.2, 0.1)), (5, (0.25, 0.25, 0.25), (0.0, 0.0, 0.5))]
    lines = ["NEWSYM=T", "LATTIC=3", gpath, "ILAT=F", "  999", f"{len(segs):5d}"]
    lines += [f"{kd} " + " ".join(repr(x) for x in a + b) for kd, a, b in segs]
    r = oracle.ref_kptin(KC.sc(), lines, str(tmp_path))
    q, w = kmesh.kpath(segs)
    assert np.array_equal(q, r["q"]) and np.array_equal(w, r["w"])


def test_errors_where_the_reference_stops(tmp_path):
    hx = KC.hexagonal()
    L = kmesh.reciprocal_lattice(hx, 4)
    g = kmesh.space_group(L, KC.d6h_cartesian(), np.zeros((24, 3)), ilat=False)
    with pytest.raises(_lib.TbkError):                 # kptgen.f:104
        kmesh.kptgen(L, g, (4, 6, 3), lspec=True)
    with pytest.raises(_lib.TbkError):                 # kptgen.f:141 (nkbzd)
        kmesh.kptgen(L, g, (8, 8, 8), lspec=False, maxfull=100)
    with pytest.raises(_lib.TbkError):                 # more irreducible points than room
        kmesh.kptgen(L, g, (8, 8, 8), lspec=False, maxk=3)
    gpath = str(tmp_path / "oh.grp")
    KC.write_group(gpath, "Oh", KC.oh_cartesian(), np.zeros((48, 3)), False)
    with pytest.raises(_lib.TbkError):                 # rdspgp.f:121 (nop > nopd)
        kmesh.read_space_group(gpath, L, False, maxop=24)
    with pytest.raises(FileNotFoundError):
        kmesh.read_space_group(str(tmp_path / "none.grp"), L, False)
    with pytest.raises(_lib.TbkError):
        kmesh.kptgen(L, g, (0, 4, 4), lspec=False)


@needs_ref
def test_reference_program_with_dropin_kptgen(tmp_path):
    """static.f and the rest of the reference's objects, kptgen.o replaced by
    kptgen_ of lib/libtbk_fortran.a (oracle/_ref/static_kpt), on ex1 with
    its k-point files replaced by generated 6x6x6 special-point meshes:
    SKOUT (which lists the k-points and weights) and SKENG identical to the
    unmodified program's."""
    r = subprocess.run(["make", "-C", ORACLE, "dropin"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    outs = {}
    for exe in ("static_ref", "static_kpt"):
        work = tmp_path / exe
        shutil.copytree(os.path.join(ORACLE, "_ref", "ex1"), work)
        skin = (work / "SKIN").read_text()
        gen = re.sub(r"-1313[^\n]*\n(fcc|bcc)\.08[^\n]*\n", "    0\nLSPEC=T LINFOK=F\n    6    6    6    0    0\n",
                     skin)
        assert gen.count("LSPEC=T") == 2
        (work / "SKIN").unlink()                        # the copy may be read-only
        (work / "SKIN").write_text(gen)
        p = subprocess.run("ulimit -s unlimited; " + os.path.join(ORACLE, "_ref", exe), shell=True, cwd=work,
                           capture_output=True, text=True, timeout=300)
        assert p.returncode == 0, p.stderr[-2000:]
        outs[exe] = ((work / "SKOUT").read_text(), (work / "SKENG").read_text())
    assert "K-POINT GENERATOR IS INVOKED" in outs["static_kpt"][0]
    assert outs["static_kpt"] == outs["static_ref"]


def _run_ex2(tmp_path, exe, aligned):
    work = tmp_path / f"{exe}_{int(aligned)}"
    shutil.copytree(os.path.join(ORACLE, "_ref", "ex2"), work)
    ex2_case.write_skin(work, aligned)
    p = subprocess.run("ulimit -s unlimited; " + os.path.join(ORACLE, "_ref", exe), shell=True, cwd=work,
                       capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr[-2000:]
    skeng = (work / "SKENG").read_text() if (work / "SKENG").exists() else ""
    return p.stdout + p.stderr, (work / "SKOUT").read_text(), skeng


@needs_ref
def test_dropin_kptgen_stops_where_the_reference_does(tmp_path):
    """ex2 as published asks kptgen for 12000^3 divisions (tests/ex2_case.py):
    the reference stops at its full-mesh bound (kptgen.f:141-145) before any
    band energy; the drop-in kptgen_ prints the same lines and stops the
    program the same way, with the same SKOUT and no SKENG row (not rows of
    zeros from an empty mesh)."""
    subprocess.run(["make", "-C", ORACLE, "dropin"], capture_output=True, check=True)
    out_ref, skout_ref, skeng_ref = _run_ex2(tmp_path, "static_ref", aligned=False)
    out_kpt, skout_kpt, skeng_kpt = _run_ex2(tmp_path, "static_kpt", aligned=False)
    assert "KDIV1,KDIV2,KDIV3=120001200012000" in skout_ref
    for out in (out_ref, out_kpt):
        assert "KPTGEN: NKBZD DIM. EXCEED., NK=  100001" in out
        assert "KPTGEN: NKBZD DIMENSION EXCEEDED" in out
    assert skout_kpt == skout_ref
    assert skeng_ref.strip() == "" and skeng_kpt.strip() == ""


@needs_ref
def test_dropin_kptgen_on_ex2_hcp(tmp_path):
    """ex2 with its mesh line in 5I5 fields: hcp with the operations given
    in direct-lattice coordinates (ILAT=T), a 12x12x12 uniform mesh, 13 c/a
    values; the program with the drop-in kptgen_ writes the SKOUT and SKENG
    of the unmodified one, and both match tests/golden/ex2.json."""
    subprocess.run(["make", "-C", ORACLE, "dropin"], capture_output=True, check=True)
    _, skout_ref, skeng_ref = _run_ex2(tmp_path, "static_ref", aligned=True)
    _, skout_kpt, skeng_kpt = _run_ex2(tmp_path, "static_kpt", aligned=True)
    assert (skout_kpt, skeng_kpt) == (skout_ref, skeng_ref)
    gold = json.load(open(EX2_JSON))
    assert [int(x) for x in re.findall(r"NUMBER OF K-POINTS=\s+(\d+)", skout_ref)] == gold["nk"]
    ef = [float(x) for x in re.findall(r"EFermi =\s+([0-9.Ee+-]+)", skout_ref)]
    et = [float(x) for x in re.findall(r"E\(T\) =\s+([0-9.Ee+-]+)", skout_ref)]
    assert ef == [g["efermi"] for g in gold["structures"]]
    assert et == [g["et"] for g in gold["structures"]]


@pytest.mark.parametrize("lattice", ["fcc", "bcc"])
def test_irreducible_mesh_reproduces_full_mesh_sums(lattice):
    """Folding is exact: the ex1 Pd structure (O_h) summed over the
    irreducible special points with kptgen's weights gives the E_F, E(T) and
    P(T) of the full 6x6x6 Monkhorst-Pack mesh with equal weights (oracle,
    CPU).  A wrong weight or a missing star member changes the sums."""
    gold = json.load(open(EX1_JSON))
    p = S.NRLParams(1, [9], [gold["valence"]], [gold["rcut"]], [gold["screenl"]],
                    np.array(gold["param"]), 0, True)
    a = 6.9
    lat = S.fcc_primitive(a) if lattice == "fcc" else S.bcc_primitive(a)
    st = S.Structure(lat, np.zeros((1, 3)), [1], p, gold["nd"])
    L = kmesh.reciprocal_lattice(lat, 1)                 # generic code: the primitive cell's mesh
    g = kmesh.space_group(L, KC.oh_cartesian(), np.zeros((48, 3)), ilat=False)
    q, w = kmesh.kptgen(L, g, (6, 6, 6), lspec=True)
    full = S.monkhorst_pack((6, 6, 6))
    assert len(q) < len(full) // 4
    setup = oracle.Setup(st)
    r_ibz = setup.bande(q, w, mode=4, tkb=gold["tkb"], eigcut=10.0, nthreads=4)
    r_full = setup.bande(full, np.full(len(full), 1.0 / len(full)), mode=4, tkb=gold["tkb"], eigcut=10.0,
                         nthreads=4)
    assert abs(r_ibz["efermi"] - r_full["efermi"]) < 1e-11
    assert abs(r_ibz["earg1"] - r_full["earg1"]) < 1e-11
    assert abs(r_ibz["earg2"] - r_full["earg2"]) * S.RYBOHR3_KBAR < 1e-6


def test_space_group_reader_fortran_fields(tmp_path):
    """rdspgp's fixed fields as a Fortran runtime reads them: implied
    decimals in fields without a point (F8.5, F4.0), D exponents, blank
    fields as zero, records shorter than the format padded with blanks,
    text past the last field ignored."""
    L = kmesh.reciprocal_lattice(KC.sc(), 3)
    ops = KC.oh_cartesian()[:2]                          # identity and one more operation
    lines = ["Test      " + "    2" + "     " + "   0.00000" + "      0.25" + "          " + " trailing text"]
    for m in ops:
        for row in m:
            lines.append("".join(f"{v:20.10E}".replace("E", "D") for v in row))
    # translations: op 1 blank (zero), op 2 (0.5, 0.25, -0.125) with implied decimals
    lines.append(" " * 24 + "   50000" + "   25000" + " -12500")
    path = tmp_path / "g.grp"
    path.write_text("\n".join(lines) + "\n")
    g = kmesh.read_space_group(str(path), L, ilat=False)
    assert g.nop == 2 and g.name == "Test"
    assert np.array_equal(g.rspc, ops)
    # the origin shift (0, 0.25, 0) moves the translations: t' = R s + t - s
    s = np.array([0.0, 0.25, 0.0])
    t = np.array([[0.0, 0.0, 0.0], [0.5, 0.25, -0.125]])
    for k in range(2):
        rdir = np.rint(np.linalg.solve(KC.sc().T, ops[k] @ KC.sc().T))
        assert np.allclose(g.tnpspc[k], rdir @ s + t[k] - s, atol=1e-15)
    # lattice coordinates in 18F4.0 records, a field without a point
    lines = ["Test2         2     " + "   0.00000" * 3]
    vals = [int(v) for v in np.concatenate([m.ravel() for m in ops])]
    rec = "".join(f"{v:4d}" for v in vals)
    lines.append(rec)
    lines.append("")                                   # blank translation record: zeros
    path.write_text("\n".join(lines) + "\n")
    g2 = kmesh.read_space_group(str(path), L, ilat=True)
    assert np.array_equal(g2.rspc, ops) and np.array_equal(g2.tnpspc, np.zeros((2, 3)))


@needs_ref
def test_space_group_reader_fields_vs_reference(tmp_path):
    """The same odd-format file read by the reference's rdspgp (through its
    kptin) and by tbk_space_group_read: identical operations and
    translations."""
    L = kmesh.reciprocal_lattice(KC.sc(), 3)
    ops = KC.oh_cartesian()
    lines = ["Oh-odd    " + "   48" + "     " + "   0.00000" + "      0.25" + "          "]
    for m in ops:
        for row in m:
            lines.append("".join(f"{v:20.10E}".replace("E", "D") for v in row))
    tv = [f"{int(x):8d}" for x in np.arange(48 * 3) * 1000 % 100000]     # implied decimals
    for i in range(0, len(tv), 9):
        lines.append("".join(tv[i:i + 9]))
    path = tmp_path / "odd.grp"
    path.write_text("\n".join(lines) + "\n")
    g = kmesh.read_space_group(str(path), L, ilat=False)
    r = oracle.ref_kptin(KC.sc(), KC.newsym_block(str(path), 3, False, (2, 2, 2), False), str(tmp_path))
    for key, mine in (("rspc", g.rspc), ("rspct", g.rspct), ("tnpspc", g.tnpspc)):
        assert np.array_equal(mine, r[key]), key
