"""Cases for the k-mesh tests: lattices, point groups generated from their
definitions (not read from the reference's files), the space-group file
writer in rdspgp's fixed format, and the NEWSYM input block kptin reads.
Shared by tests/test_kmesh.py and tests/golden/make_kmesh_fixture.py."""
from __future__ import annotations

import itertools

import numpy as np

A_CUBIC = 6.8          # bohr, as ex1's first structure


def fcc(a=A_CUBIC):
    return 0.5 * a * np.array([[0, 1, 1], [1, 0, 1], [1, 1, 0]], float)


def bcc(a=A_CUBIC):
    return 0.5 * a * np.array([[-1, 1, 1], [1, -1, 1], [1, 1, -1]], float)


def sc(a=A_CUBIC):
    return a * np.eye(3)


def hexagonal(a=5.957092, c_over_a=1.3):
    """a1 = a (1/2, -sqrt3/2, 0), a2 = a (1/2, sqrt3/2, 0), a3 = (0, 0, c)."""
    s = np.sqrt(3.0) / 2.0
    return np.array([[0.5 * a, -s * a, 0.0], [0.5 * a, s * a, 0.0], [0.0, 0.0, c_over_a * a]])


def bct(a=6.0, c=8.4):
    return 0.5 * np.array([[-a, a, c], [a, -a, c], [a, a, -c]], float)


def triclinic():
    return np.array([[6.1, 0.2, -0.1], [0.4, 5.7, 0.3], [-0.2, 0.5, 7.3]])


def oh_cartesian():
    """The 48 signed permutation matrices (O_h)."""
    ops = []
    for perm in itertools.permutations(range(3)):
        for signs in itertools.product((1, -1), repeat=3):
            m = np.zeros((3, 3))
            for i in range(3):
                m[i, perm[i]] = signs[i]
            ops.append(m)
    ident = [i for i, m in enumerate(ops) if np.array_equal(m, np.eye(3))][0]
    ops.insert(0, ops.pop(ident))
    return np.array(ops)


def d4h_cartesian():
    """The 16 operations of O_h that keep the z axis."""
    return np.array([m for m in oh_cartesian() if abs(m[2, 2]) == 1.0])


def d6h_cartesian():
    """Rotations by n 60 degrees about z, the six in-plane two-fold axes,
    and their products with the inversion (24)."""
    ops = []
    for n in range(6):
        t = n * np.pi / 3.0
        c, s = np.cos(t), np.sin(t)
        ops.append(np.array([[c, -s, 0], [s, c, 0], [0, 0, 1.0]]))
    for n in range(6):
        t = n * np.pi / 3.0                      # reflection line at angle t/2 ... two-fold axis
        c, s = np.cos(t), np.sin(t)
        ops.append(np.array([[c, s, 0], [s, -c, 0], [0, 0, -1.0]]))
    ops = ops + [-m for m in ops]
    return np.array([np.where(np.abs(m) < 1e-15, 0.0, m) for m in ops])


def ci():
    return np.array([np.eye(3), -np.eye(3)])


def to_direct(lat, ops):
    """Cartesian operations in direct-lattice coordinates (integer)."""
    A = np.asarray(lat).T                        # columns = lattice vectors
    out = np.array([np.linalg.solve(A, m @ A) for m in ops])
    r = np.rint(out)
    assert np.abs(out - r).max() < 1e-9, "operation not in the lattice's group"
    return r


def write_group(path, name, rot, tnp, ilat, shift=(0.0, 0.0, 0.0)):
    """A space-group file in rdspgp's format (rdspgp.f:83-90): header
    A10,I5,5X,3F10.5; rotations 18F4.0 (ilat) or 3D20.10 (Cartesian, one
    row per record); translations 9F8.5."""
    rot = np.asarray(rot, float).reshape(-1, 3, 3)
    tnp = np.asarray(tnp, float).reshape(-1, 3)
    lines = [f"{name[:10]:<10s}{len(rot):5d}     " + "".join(f"{x:10.5f}" for x in shift)]
    if ilat:
        vals = rot.reshape(-1)
        for i in range(0, len(vals), 18):
            lines.append("".join(f"{v:3.0f}." for v in vals[i:i + 18]))
    else:
        for m in rot:
            for row in m:
                lines.append("".join(f"{v:20.10f}" for v in row))
    tv = tnp.reshape(-1)
    for i in range(0, len(tv), 9):
        lines.append("".join(f"{v:8.5f}" for v in tv[i:i + 9]))
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def newsym_block(group_path, lattic, ilat, kdiv, lspec, kpst=0, kpend=0, info=None, sfk=0.0):
    """kptin's input for a generated mesh (kptin.f:84-96, 184-186, 278-280)."""
    ninfo = 0 if info is None else len(info)
    lines = ["NEWSYM=T", f"LATTIC={lattic:d}", group_path, f"ILAT={'T' if ilat else 'F'}",
             f"{-ninfo:5d}{sfk:10.6f}",
             f"LSPEC={'T' if lspec else 'F'} LINFOK={'T' if ninfo else 'F'}",
             "".join(f"{x:5d}" for x in (*kdiv, kpst, kpend))]
    for p in (info or []):
        lines.append(" ".join(repr(float(x)) for x in p))
    return lines


def cases():
    """(name, lattice rows, lattic, ilat, rotations as written, tnp, shift,
    kdiv, lspec, kpst, kpend, info, sfk)."""
    oh = oh_cartesian()
    z48 = np.zeros((48, 3))
    hx = hexagonal()
    d6 = d6h_cartesian()
    tn6 = np.zeros((24, 3))
    tn6[6:12, 2] = 0.5                           # screw / glide part of P6_3/mmc
    tn6[18:24, 2] = 0.5
    d4 = d4h_cartesian()
    out = []
    for q in (2, 4, 6, 8):
        out.append((f"fcc_mp{q}", fcc(), 1, False, oh, z48, None, (q, q, q), True, 0, 0, None, 0.0))
    for q in (4, 5):
        out.append((f"fcc_uniform{q}", fcc(), 1, False, oh, z48, None, (q, q, q), False, 0, 0, None, 0.0))
    out.append(("bcc_mp4", bcc(), 2, False, oh, z48, None, (4, 4, 4), True, 0, 0, None, 0.0))
    out.append(("bcc_uniform6", bcc(), 2, False, oh, z48, None, (6, 6, 6), False, 0, 0, None, 0.0))
    out.append(("sc_mp4", sc(), 3, False, oh, z48, None, (4, 4, 4), True, 0, 0, None, 0.0))
    out.append(("hex_direct_uniform", hx, 4, True, to_direct(hx, d6), tn6, None, (6, 6, 4), False, 0, 0,
                None, 0.0))
    out.append(("hex_direct_special", hx, 4, True, to_direct(hx, d6), tn6, None, (4, 4, 3), True, 0, 0,
                None, 0.0))
    out.append(("hex_cartesian_special", hx, 4, False, d6, tn6, None, (5, 5, 2), True, 0, 0, None, 0.0))
    for lt in (7, 8, 9):
        out.append((f"bct{lt}", bct(), lt, False, d4, np.zeros((16, 3)), None, (3, 3, 4), lt != 8, 0, 0,
                    None, 0.0))
    tn = np.zeros((48, 3))
    tn[1::2] = 0.25
    out.append(("fcc_shifted_origin", fcc(), 1, False, oh, tn, (0.25, 0.25, 0.25), (4, 4, 4), True, 0, 0,
                None, 0.0))
    out.append(("fcc_range", fcc(), 1, False, oh, z48, None, (6, 6, 6), True, 3, 12, None, 0.0))
    out.append(("fcc_info_points", fcc(), 1, False, oh, z48, None, (4, 4, 4), True, 0, 0,
                [(0.5, 0.5, 0.5), (1.0, 0.0, 0.25)], 2.0))
    out.append(("triclinic_ci", triclinic(), 5, False, ci(), np.zeros((2, 3)), None, (3, 4, 5), False, 0, 0,
                None, 0.0))
    return out
