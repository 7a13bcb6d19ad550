"""The reference's second example, NRLtb/ex2 (hcp Ti, two atoms per cell,
c/a from 1.30 to 1.90 at V = 238 bohr^3, mode 3: E_F and E(T)), as input to
the reference program.

As published, its SKIN gives the k-mesh divisions left-aligned
("12   12   12") while kptin reads them in 5I5 fields (kptin.f:279, 410)
from a unit opened with BLANK='ZERO' (static.f:682), so they read as
12000 12000 12000 and kptgen stops at its full-mesh bound
(kptgen.f:141-145); the published SKOUT ends there ("KDIV1,KDIV2,KDIV3=
120001200012000").  skin_text(aligned=True) right-aligns the line in its
fields, as the line's own comment ("must be (5I5) format") asks; nothing
else changes.
"""
import re

MESH_AS_PUBLISHED = "12   12   12 "
MESH_ALIGNED = "   12   12   12"
INPUTS = ("SKIN", "ti_par", "spcgp1.hcp")


def skin_text(published: str, aligned: bool = True) -> str:
    if not aligned:
        return published
    out, n = re.subn(r"(?m)^12   12   12 ", MESH_ALIGNED, published)
    assert n == 1, "the ex2 mesh line was not found"
    return out


def write_skin(work, aligned: bool = True):
    """Rewrite work/SKIN in place (as a new file: the copied one may be
    read-only)."""
    path = work / "SKIN"
    text = skin_text(path.read_text(), aligned)
    path.unlink()
    path.write_text(text)


def read_skeng(path):
    """SKENG rows: (label, volume, E_F, E(T)) in mode 3 (static.f's writer)."""
    rows = []
    for line in open(path):
        if line.strip():
            t = line.split()
            rows.append((" ".join(t[:2]), float(t[2]), float(t[3]), float(t[4])))
    return rows
