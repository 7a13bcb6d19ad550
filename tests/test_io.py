"""The native readers and writers of the reference program's files
(tbk_read_params, tbk_skin_*, tbk_skeng_line) against the reference's own
readers (input.f, setlat.f, input1.f, kptin.f, built in oracle/_ref and
called through oracle/ref/ref_driver.f on the same files), bit for bit, and
the SKENG writer against the published NRLtb/ex1/SKENG lines.
"""
import json
import os
import shutil

import numpy as np
import pytest

import oracle
import skin_cases as SC
from tbk import _lib
from tbk import io as tio
from tbk import structures as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EX1 = os.path.join(ROOT, "oracle", "_ref", "ex1")
EX2 = os.path.join(ROOT, "oracle", "_ref", "ex2")
C_PAR = "/root/reference/NRLtb/atom_params/C_par"
needs_ref = pytest.mark.skipif(not oracle.ref_available() or not os.path.isdir(EX1),
                               reason="reference build (oracle/_ref) not present")


def ref_params(tmp_path, path):
    f = tmp_path / "skin_line"
    f.write_text(path + "\n")
    return oracle.ref_input(str(f))


def assert_params_equal(mine, ref):
    p, chans, ps = mine
    assert ps.nparam == ref["nparam"] and ps.jspins == ref["jspins"]
    assert int(ps.jsover) == ref["jsover"] and ps.nltype == ref["nltype"] and ps.kinds == ref["kinds"]
    assert list(p.kbas) == list(ref["kbas"])
    assert np.array_equal(p.valence, ref["valence"])
    assert np.array_equal(p.dnn, ref["dnn"]) and np.array_equal(p.screenl, ref["screenl"])
    assert np.array_equal(chans, ref["param"])


@needs_ref
@pytest.mark.parametrize("name", ["pd_par", "ti_par", "C_par"])
def test_parameter_files_of_the_reference(tmp_path, name):
    path = {"pd_par": os.path.join(EX1, "pd_par"), "ti_par": os.path.join(EX2, "ti_par"), "C_par": C_PAR}[name]
    if not os.path.exists(path):
        pytest.skip(f"{path} not present")
    assert_params_equal(tio.read_params(path), ref_params(tmp_path, path))


@needs_ref
@pytest.mark.parametrize("variant", ["orthogonal", "magnetic", "two_kinds", "short", "dexp", "nl90000",
                                     "nl_unknown", "nl_blank", "blank_line"])
def test_parameter_file_variants(tmp_path, variant):
    """Header letters O and M, the style field (90000, out of range, blank),
    a file short of values (zeroed), D exponents, a blank record among the
    values, and two kinds."""
    p = S.synthetic_params(11)
    kw, chans = {}, None
    if variant == "orthogonal":
        kw["head"] = "ON"
    elif variant == "magnetic":
        kw["head"] = "NM"
        chans = [p.param, p.param * 1.01]
    elif variant == "two_kinds":
        import branch_cases as B
        d = B._two_kind_params([9, 4])
        p = S.NRLParams(2, d["kbas"], d["valence"], d["dnn"], d["screenl"], np.array(d["param"]), 0, True)
    elif variant == "short":
        kw["drop"] = 7
    elif variant == "dexp":
        kw["dexp"] = True
    elif variant == "nl90000":
        kw["nltype_field"] = "90000"
    elif variant == "nl_unknown":
        kw["nltype_field"] = "    7"
    elif variant == "nl_blank":
        kw["nltype_field"] = "     "
    elif variant == "blank_line":
        kw["blank_lines"] = True
    path = str(tmp_path / "par")
    SC.write_params(path, p, channels=chans, **kw)
    mine = tio.read_params(path)
    assert_params_equal(mine, ref_params(tmp_path, path))
    if variant == "short":
        assert mine[2].nparam == len(p.param) - 7 and np.all(mine[1][0, -7:] == 0.0)


def skin_dir(tmp_path, structures, mode=4):
    shutil.copy(os.path.join(EX1, "pd_par"), tmp_path / "pd_par")
    SC.write_trivial_group(str(tmp_path / "trivial.grp"))
    SC.write_skin(str(tmp_path / "SKIN"), "pd_par", structures, mode=mode)
    return tmp_path


@needs_ref
def test_structures_of_every_lattice_menu_and_strain(tmp_path):
    """setlat's 33 lattice inputs with the strain options, atoms in lattice
    and Cartesian coordinates with phase origins and extra electrons, and
    frozen-phonon shifts: lattice, volume, Cartesian positions, phase
    origins, search widths and electron count identical to the reference's
    setlat + input1 (and its main program's conversion, static.f:1088-1103)."""
    cases = SC.lattice_structures()
    d = skin_dir(tmp_path, cases)
    skin = tio.Skin(str(d / "SKIN"))
    assert (skin.mode, skin.tkb, skin.eigcut) == (4, 0.005, 10.0)
    got = list(skin)
    skin.close()
    assert len(got) == len(cases)
    for i, ((title, label, z, block, _), g) in enumerate(zip(cases, got)):
        bf = d / f"block{i}"
        bf.write_text("\n".join(block) + "\n")
        r = oracle.ref_input1(str(bf), [9], [10.0], z)
        assert r["inpstat"] == 0
        st = g.structure
        assert g.label == f"{label:<20s}"[:20]
        assert np.array_equal(st.lat, r["lat"]), (label, st.lat, r["lat"])
        assert g.thisvol == r["volume"], label
        assert np.array_equal(st.posn, r["posn"]), label
        assert np.array_equal(st.posq, r["posq"]), label
        assert list(st.nd) == list(r["nd"]) and list(st.kkind) == list(r["kkind"])
        assert g.zfill == r["zfill"]
        assert len(g.weights) == 8


@needs_ref
def test_spin_options_of_the_atom_lines(tmp_path):
    """Spin flips (magdir = 2) and non-collinear angles (degrees) on the
    atom lines (input1.f:735-812)."""
    atoms_flip = [" 1  0.0 0.0 0.0  0 0 0  2", " 1  0.5 0.5 0.5  0 0 0  1"]
    atoms_nc = [" 1  0.0 0.0 0.0  0 0 0  30.0 45.0", " 1  0.5 0.5 0.5  0 0 0  120.0 -60.0"]
    cases = [("flip", "flip", 0.0, SC.structure_block((4, "4.9"), (0, None), atoms_flip, flags="FT"),
              SC.kblock_explicit(SC.K8, SC.W8)),
             ("noncol", "noncol", 0.0, SC.structure_block((4, "4.9"), (0, None), atoms_nc, flags="FN"),
              ["NEWSYM=F"])]
    d = skin_dir(tmp_path, cases)
    got = list(tio.Skin(str(d / "SKIN")))
    for i, ((_, _, z, block, _), g) in enumerate(zip(cases, got)):
        bf = d / f"block{i}"
        bf.write_text("\n".join(block) + "\n")
        r = oracle.ref_input1(str(bf), [9], [10.0], z)
        assert list(g.flipos) == list(r["flipos"])
        assert g.flags["lnoncol"] == r["lnoncol"]
        if r["lnoncol"]:
            assert np.array_equal(g.spinat, r["spinat"])
    assert list(got[0].flipos) == [1, 0] and got[1].flags["lnoncol"] == 1


@needs_ref
@pytest.mark.parametrize("which", ["explicit", "explicit_sfk", "file", "mesh", "special", "info_range", "line"])
def test_kpoint_blocks_vs_reference_kptin(tmp_path, which):
    """The k-point block forms of kptin.f (explicit lists in SKIN and in a
    -1313 file, generated uniform and special-point meshes, information
    points with a kpst..kpend range, line mode) read natively after a
    one-atom fcc structure: k-points and weights identical to the
    reference's kptin on the same block."""
    for f in ("spcgrp.fcc", "fcc.08"):
        shutil.copy(os.path.join(EX1, f), tmp_path / f)
    blocks = dict(SC.kpoint_blocks("spcgrp.fcc", "fcc.08"))
    kb = blocks[which]
    block = SC.structure_block((1, "6.9"), (0, None), [" 1  0.0 0.0 0.0  0 0 0"])
    d = skin_dir(tmp_path, [("k", "k", 0.0, block, kb)])
    g = list(tio.Skin(str(d / "SKIN")))[0]
    kf = d / "kblock"
    kf.write_text("\n".join(kb) + "\n")
    cwd = os.getcwd()
    os.chdir(d)                  # the block names its files relative to the run directory
    try:
        L = oracle.ref()
        import ctypes as C
        maxk = 50000
        q, w = np.zeros(3 * maxk), np.zeros(maxk)
        av, bv, bij = np.zeros(9), np.zeros(9), np.zeros(9)
        rsp, rspt, tnp = np.zeros(9 * 48), np.zeros(9 * 48), np.zeros(3 * 48)
        nk, lattic, nop, ierr = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        vol = C.c_double()
        fb, sb = os.fsencode(str(kf)), os.fsencode(str(d / "scratch15"))
        L.ref_kptin(np.ascontiguousarray(g.structure.lat.ravel()), fb, C.byref(C.c_int(len(fb))), sb,
                    C.byref(C.c_int(len(sb))), C.byref(C.c_int(maxk)), C.byref(nk), q, w, av, bv, bij,
                    C.byref(vol), C.byref(lattic), C.byref(nop), rsp, rspt, tnp, C.byref(ierr))
    finally:
        os.chdir(cwd)
    assert ierr.value == 0
    n = nk.value
    assert len(g.weights) == n, (which, len(g.weights), n)
    assert np.array_equal(g.kpts, q[:3 * n].reshape(n, 3))
    assert np.array_equal(g.weights, w[:n])


def test_skeng_lines_format():
    """static.f:1238-1254's formats: label (a20), f12.6, f15.9 ..., f10.6."""
    line = tio.skeng_line(" fcc  6.80", 78.608, 0.25679727656536383, 0.038518610966493151,
                          4.6249402027521728e-3, mode=4)
    assert line == " fcc  6.80             78.608000    0.256797277    0.038518611    0.004624940"
    assert tio.skeng_line("x", 1.0, -0.5, 2.0, 9.0, mode=3) == "x" + " " * 19 + "    1.000000   -0.500000000    2.000000000"
    assert tio.skeng_line("m", 1.0, 0.1, 0.2, 0.3, smag=0.25, mode=4, with_smag=True).endswith("  0.250000")
    assert tio.skeng_line("big", 1e9, 0.0, 0.0, 0.0, mode=4)[20:32] == "*" * 12


@needs_ref
def test_ex1_skin_read_natively():
    """ex1's SKIN: 18 structures whose volumes and k-point counts are those
    of the published run (ex1/SKENG volumes, 85 fcc and 55 bcc k-points),
    and whose SKENG lines, written from the published full-precision values
    (ex1/SKOUT), are the published ex1/SKENG lines."""
    gold = json.load(open(os.path.join(ROOT, "tests", "golden", "ex1.json")))
    skeng = open("/root/reference/NRLtb/ex1/SKENG").read().splitlines() \
        if os.path.exists("/root/reference/NRLtb/ex1/SKENG") else None
    got = list(tio.Skin(os.path.join(EX1, "SKIN")))
    assert len(got) == len(gold["structures"]) == 18
    for i, (g, s) in enumerate(zip(got, gold["structures"])):
        assert g.thisvol == pytest.approx(s["volume"], rel=1e-15)
        assert len(g.weights) == (85 if s["lattice"] == "fcc" else 55)
        assert abs(g.weights.sum() - 1.0) < 1e-12
        if skeng:
            line = tio.skeng_line(g.label, g.thisvol, s["efermi"], s["et"], s["pt"], mode=4)
            assert line == skeng[i], (line, skeng[i])


def test_skin_errors_are_codes(tmp_path):
    """An unreadable lattice line, an unknown lattice type and a missing
    parameter file are error codes, not stops."""
    p = tmp_path / "pd"
    SC.write_params(str(p), S.pd_like_params())
    SC.write_trivial_group(str(tmp_path / "trivial.grp"))
    SC.write_skin(str(tmp_path / "SKIN"), "pd", [("t", "l", 0.0, [" 99", "5.0", " 0", " 1", "4 4 4", "F",
                                                                   " 1 0 0 0 0 0 0"],
                                                   SC.kblock_explicit(SC.K8, SC.W8))])
    skin = tio.Skin(str(tmp_path / "SKIN"))
    with pytest.raises(_lib.TbkError):
        next(skin)
    skin.close()
    SC.write_skin(str(tmp_path / "SKIN2"), "nofile", [])
    with pytest.raises(_lib.TbkError):
        tio.Skin(str(tmp_path / "SKIN2"))
