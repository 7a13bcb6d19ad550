"""Input files of the reference program written for the tests of the
native readers (tbk_read_params, tbk_skin_*, tbk_static_run): NRL parameter
files in input.f's layout, SKIN structure blocks for every lattice menu of
setlat.f and the strain options, atoms in lattice and Cartesian coordinates
with frozen-phonon shifts and spin options (input1.f), and k-point blocks
(kptin.f), plus the SKIN files of the end-to-end cases.
"""
import os

import numpy as np

import kmesh_cases as K
from tbk import structures as S

# the lattice menus of setlat.f:167-204 with parameters that give cells of
# about Pd's nearest-neighbour distance (bohr, radians)
LATTICES = [
    (1, "6.9"), (-1, "82.1"), (2, "5.5"), (-2, "83.2"), (3, "5.2 8.5"), (-3, "238 1.63"),
    (4, "4.9"), (-4, "117.6"), (5, "5.0 5.6"), (-5, "140.0 1.1"), (6, "5.4 6.0"), (-6, "87.5 1.1"),
    (7, "5.0 5.3 5.6"), (-7, "148.0 1.05 1.1"), (8, "5.0 5.4 5.8"), (-8, "78.5 1.08 1.16"),
    (9, "6.9 4.9"), (-9, "116.6 0.71"), (10, "6.9 6.7"), (-10, "80.0 0.97"), (11, "6.9 7.1"),
    (-11, "84.5 1.03"), (12, "0.3 3.45"), (-12, "81.0 1.05"), (13, "5.5 5.8"), (-13, "87.5 1.05"),
    (14, "4.9 5.1 1.3"), (-14, "118.0 1.04 1.25"), (15, "5.4 0.05 0.3"), (-15, "78.7 0.1 -0.4"),
    (16, "6.9 0.2310490601866 0.0"), (-16, "82.0 0.12 1.0"),
]
ARBITRARY = (0, "4.9 0.1 0.0\n0.2 4.8 0.0\n0.3 0.1 5.1\n1.0 1.02 0")
# setlat.f:811-976
STRAINS = [(0, None), (1, "0.01 -0.02 0.015 0.004 -0.003 0.002"), (2, "0.0004"), (-2, "0.0004"),
           (21, "0.02"), (-21, "-0.02"), (3, "0.0009"), (-3, "0.0009"), (4, "0.03"), (5, "0.02"),
           (6, "0.0016"), (6, "-0.0016")]


def write_params(path, p: S.NRLParams, head="NN", channels=None, nltype_field=None, drop=0, dexp=False,
                 blank_lines=False):
    """An NRL parameter file for p in input.f:181-359's layout: header
    (2a1,i5), title, kinds, cutoffs, per kind orbitals / weight / s p d
    occupancies, then one value per line for each channel (channels: more
    than p.param for a magnetic file).  drop: leave out the last values;
    dexp: D exponents; blank_lines: a blank line before the values."""
    chans = [np.asarray(p.param, float)] if channels is None else [np.asarray(c, float) for c in channels]
    nlt = f"{p.nltype:5d}" if nltype_field is None else nltype_field
    lines = [f"{head}{nlt}                           (parameter style)", "test parameters",
             f"{p.kinds}                                 (kinds)"]
    for i in range(p.kinds * (p.kinds + 1) // 2):
        lines.append(f"{p.dnn[i]}   {p.screenl[i]}                     (RCUT and SCREENL)")
    for k in range(p.kinds):
        v = p.valence[k]
        lines += [f"{p.kbas[k]}", "100.0", f" {v - 2.0 if v >= 2 else 0.0:.1f}  {min(v, 2.0):.1f} 0.0"]
    if blank_lines:
        lines.append("")
    for c in chans:
        vals = c[:len(c) - drop] if drop else c
        for j, x in enumerate(vals):
            s = f"{x:.15E}"
            if dexp:
                s = s.replace("E", "D")
            lines.append(f"  {s}   0  {j + 1}     value")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def structure_block(lattice, strain, atoms, nd="4 4 4", flags="F", phonon=None):
    """A SKIN structure block from its lattice line on (setlat.f, input1.f):
    lattice (latnum, parameter text), strain (option, values or None),
    atoms: lines 'kind x y z qx qy qz [...]'."""
    latnum, par = lattice
    out = [f" {latnum}", par, f" {strain[0]}"]
    if strain[0] != 0:
        out.append(strain[1])
    out += [f" {len(atoms)}", nd, flags]
    if phonon:
        out.append(f" {len(phonon)}")
        out += phonon
    out += atoms
    return out


def kblock_explicit(kpts, weights, sfk=1.0, group="trivial.grp", lattic=0):
    """NEWSYM block with an explicit list in SKIN itself (kptin.f:253-273)."""
    out = ["NEWSYM=T", f"LATTIC={lattic}", group, "ILAT=T", f"{len(kpts):5d}{sfk:10.6f}"]
    out += [f"{k[0]!r} {k[1]!r} {k[2]!r} {w!r}" for k, w in zip(kpts, weights)]
    return out


def write_trivial_group(path):
    """The identity alone, in rdspgp's layout (lattice coordinates)."""
    K.write_group(path, "P1", [np.eye(3)], [np.zeros(3)], True)


def write_skin(path, param_file, structures, mode=4, tkb=0.005, eigcut=10.0):
    """A SKIN file: structures = [(title, label, zcharge, block lines,
    k-block lines)]."""
    lines = [f"Mode={mode}               (mode)", f"{tkb} {eigcut}        (kT, eigcut)", param_file]
    for title, label, z, block, kb in structures:
        lines += [title, f"{label:<20s}(label)", f" {z}"] + block + kb
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def pd_atoms_fcc_conv():
    return [" 1  0.0 0.0 0.0  0 0 0", " 1  0.0 0.5 0.5  0 0 0", " 1  0.5 0.0 0.5  0 0 0",
            " 1  0.5 0.5 0.0  0 0 0"]


K8 = [[0.125, 0.125, 0.125], [0.375, 0.125, 0.125], [0.375, 0.375, 0.125], [0.375, 0.375, 0.375],
      [-0.125, 0.375, 0.125], [0.125, -0.375, 0.375], [0.375, -0.125, -0.375], [-0.375, -0.375, 0.125]]
W8 = [1.0, 3.0, 3.0, 1.0, 2.0, 2.0, 1.0, 3.0]


def lattice_structures():
    """(title, label, zcharge, block, k-block) for a SKIN that walks the
    lattice menus and the input1 options on one or two Pd atoms: explicit
    k lists of 8 points (the first structure's list is re-used with
    NEWSYM=F by every second structure)."""
    out = []
    atoms1 = [" 1  0.0 0.0 0.0  0 0 0"]
    for i, lat in enumerate(LATTICES + [ARBITRARY]):
        strain = STRAINS[i % len(STRAINS)]
        kb = kblock_explicit(K8, W8) if i % 2 == 0 else ["NEWSYM=F"]
        out.append((f"lattice {lat[0]}", f"lat{lat[0]:+d} s{strain[0]:+d}", 0.0,
                    structure_block(lat, strain, atoms1), kb))
    # two atoms: a Cartesian position (negative kind), a phase origin, extra
    # electrons
    out.append(("two atoms, Cartesian", "cart", 0.5,
                structure_block((4, "9.6"), (0, None), [" 1  0.0 0.0 0.0  0 0 0",
                                                          "-1  4.8 4.8 0.0  0.1 0.0 0.0"]),
                kblock_explicit(K8, W8)))
    # frozen phonon: one shift in lattice, one in Cartesian coordinates
    out.append(("phonon", "phonon", 0.0,
                structure_block((4, "4.9"), (0, None), [" 1  0.0 0.0 0.0  0 0 0", " 1  0.5 0.5 0.5  0 0 0"],
                                flags="T", phonon=[" 1  F  0.01 0.0 -0.02", " 2  T  0.05 0.03 0.0"]),
                kblock_explicit(K8, W8)))
    return out


def kpoint_blocks(group_fcc, kfile):
    """k-point blocks for a one-atom fcc cell (kptin.f): (name, lines)."""
    return [
        ("explicit", kblock_explicit(K8, W8, sfk=1.0, group=group_fcc, lattic=1)),
        ("explicit_sfk", kblock_explicit([[1.0, 1.0, 1.0], [3.0, 1.0, 1.0]], [1.0, 3.0], sfk=8.0,
                                         group=group_fcc, lattic=1)),
        ("file", ["NEWSYM=T", "LATTIC=1", group_fcc, "ILAT=F", "-1313", kfile]),
        ("mesh", ["NEWSYM=T", "LATTIC=1", group_fcc, "ILAT=F", "    0  1.000000", "lspec=f,linfok=f",
                  "    6    6    6    0    0"]),
        ("special", ["NEWSYM=T", "LATTIC=1", group_fcc, "ILAT=F", "    0  1.000000", "lspec=t,linfok=f",
                     "    4    4    4    0    0"]),
        ("info_range", ["NEWSYM=T", "LATTIC=1", group_fcc, "ILAT=F", "   -2  2.000000", "lspec=f,linfok=t",
                        "    4    4    4    2    7", "0.5 0.0 0.0", "1.0 1.0 0.0"]),
        ("line", ["NEWSYM=T", "LATTIC=1", group_fcc, "ILAT=F", "  999  1.000000", "    2",
                  "5 0.0 0.0 0.0 0.5 0.0 0.5", "4 0.5 0.0 0.5 0.5 0.25 0.75"]),
    ]


# ---------------------------------------------------------------------------
# end-to-end cases: a directory with SKIN and the files it names, run by the
# reference program (oracle/_ref/static_ref) for the fixture and by the
# native program (tbk_static) on the GPU
# ---------------------------------------------------------------------------
STATIC_CASES = ("ex1", "ex1_mode1", "ex1_mode2", "ex1_mode3", "ex2_aligned", "ex2_published", "lattices",
                "spin_flip", "noncol", "ex1_qlmt6", "ex1_qlmt10", "ex1_qlmt9", "ex1_qapw8", "ex2_qlmt6",
                "ex2_qlmt10", "spin_qlmt6", "noncol_qlmt6", "noncol_spdir", "noncol20")
# cases run by the reference program built for the spin-direction output
# (oracle/Makefile spdir: msdir = 1), and the files they add
SPDIR_CASES = {"noncol_spdir": "SPNDIR"}
# the QLMT / QAPW file each case writes (modes 6, 9, 10 / 8, static.f:757-778)
QFILE = {"ex1_qlmt6": "QLMT", "ex1_qlmt10": "QLMT", "ex1_qlmt9": "QLMT", "ex1_qapw8": "QAPW",
         "ex2_qlmt6": "QLMT", "ex2_qlmt10": "QLMT", "spin_qlmt6": "QLMT", "noncol_qlmt6": "QLMT"}


def _first_structures(text, n, title="Palladium"):
    """A SKIN cut after its n-th structure (each starts with its title line,
    'Palladium ...' in ex1, 'Titanium ...' in ex2)."""
    lines = text.splitlines()
    starts = [i for i, l in enumerate(lines) if l.startswith(title)]
    return "\n".join(lines[:starts[n]] if n < len(starts) else lines) + "\n"


def _mesh4(text):
    """ex1's first k-point block with the k-point file (-1313, fcc.08)
    replaced by a generated uniform 4x4x4 mesh (kptin.f:276-351)."""
    lines = text.splitlines()
    i = next(j for j, l in enumerate(lines) if l.startswith("-1313"))
    lines[i:i + 2] = ["    0  1.000000     (generate)", "lspec=f,linfok=f", "    4    4    4    0    0"]
    return "\n".join(lines) + "\n"


def _mode(text, mode):
    """The SKIN's mode line set to `mode` (static.f:706, read as (5x,a2))."""
    lines = text.splitlines()
    lines[0] = f"Mode={mode:<2d}" + lines[0][7:]
    return "\n".join(lines) + "\n"


def spdir_structure(flag="S"):
    """A 20-atom Pd cell for the spin-direction output (input1.f: spin flag
    'S', non-collinear with SPNDIR; 'N' plain non-collinear): five conventional fcc cells stacked
    along z (simple tetragonal, a = 7.35, c = 5a), every atom with its own
    spin direction (theta, phi in degrees), on an explicit list of 8
    k-points with the trivial group.  Twenty atoms fill the reference
    build's matom, so its orbital loop kbas(kkind(matom)) (diag.f:488)
    covers every orbital of an atom."""
    a = 7.35
    base = [(0.0, 0.0, 0.0), (0.0, 0.5, 0.5), (0.5, 0.0, 0.5), (0.5, 0.5, 0.0)]
    atoms = []
    for n in range(5):
        for i, (x, y, z) in enumerate(base):
            t = 4 * n + i
            theta, phi = 15.0 + 7.0 * t, 20.0 + 13.0 * t
            atoms.append(f" 1  {x:.3f}  {y:.3f}  {(z + n) / 5:.3f}  0 0 0  {theta:.1f} {phi:.1f}")
    block = structure_block((5, f"{a} {5 * a}"), (0, None), atoms, flags="F" + flag)
    return ("Palladium spin directions", "pd20 spdir" if flag == "S" else "pd20 noncol", 0.0, block,
            kblock_explicit(K8, W8))


def build_static_case(name, work, ex1_dir, ex2_dir):
    """Write case `name` into the directory `work` from the examples' input
    files in ex1_dir / ex2_dir (oracle/_ref/ex1, ex2)."""
    import shutil
    import ex2_case
    import spin_case
    work = str(work)
    os.makedirs(work, exist_ok=True)
    if name.startswith("ex2"):
        for f in ex2_case.INPUTS:
            shutil.copy(os.path.join(ex2_dir, f), os.path.join(work, f))
        text = open(os.path.join(ex2_dir, "SKIN")).read()
        os.remove(os.path.join(work, "SKIN"))
        text = ex2_case.skin_text(text, aligned=name != "ex2_published")
        if "qlmt" in name:           # the first structure on a 4x4x4 mesh, mode 6 or 10
            text = _mode(_first_structures(text.replace(ex2_case.MESH_ALIGNED, "    4    4    4"), 1, "Titanium"),
                         int(name.split("qlmt")[1]))
        open(os.path.join(work, "SKIN"), "w").write(text)
        return
    for f in ("pd_par", "fcc.08", "bcc.08", "spcgrp.fcc", "spcgrp.bcc"):
        shutil.copy(os.path.join(ex1_dir, f), os.path.join(work, f))
    skin = open(os.path.join(ex1_dir, "SKIN")).read()
    if name in ("noncol_spdir", "noncol20"):   # two-channel pd_par, one 20-atom structure
        par = open(os.path.join(ex1_dir, "pd_par")).read()
        pdn = spin_case.spin_down_values(spin_case.read_channel(par))
        os.remove(os.path.join(work, "pd_par"))
        open(os.path.join(work, "pd_par"), "w").write(spin_case.nm_text(par, pdn))
        write_trivial_group(os.path.join(work, "trivial.grp"))
        write_skin(os.path.join(work, "SKIN"), "pd_par", [spdir_structure("S" if name == "noncol_spdir" else "N")])
        return
    if name == "ex1":
        out = skin
    elif name.startswith("ex1_mode"):
        out = _first_structures(skin, 3).replace("Mode=4", "Mode=" + name[-1], 1)
    elif name.startswith("ex1_q"):
        out = _mode(_mesh4(_first_structures(skin, 1)), int(name[len(name.rstrip("0123456789")):]))
    elif name == "lattices":
        write_trivial_group(os.path.join(work, "trivial.grp"))
        write_skin(os.path.join(work, "SKIN"), "pd_par", lattice_structures())
        return
    else:                                  # spin_flip, noncol: ex1 with a two-channel pd_par
        par = open(os.path.join(ex1_dir, "pd_par")).read()
        pdn = spin_case.spin_down_values(spin_case.read_channel(par))
        os.remove(os.path.join(work, "pd_par"))
        open(os.path.join(work, "pd_par"), "w").write(spin_case.nm_text(par, pdn))
        kind = "flip" if name.startswith("spin") else "noncol"
        if name.endswith("qlmt6"):
            out = _mode(_mesh4(_first_structures(spin_case.skin_spin(skin, kind), 1)), 6)
        else:
            out = _first_structures(spin_case.skin_spin(skin, kind), 3)
    open(os.path.join(work, "SKIN"), "w").write(out)
