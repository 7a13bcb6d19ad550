"""A spin-polarised ("NM") NRL parameter file for the reference program:
the spin-up channel is the file's own block, the spin-down channel follows
it (input.f:181-182 reads the second header letter, input.f:327-359 the
channels in turn).  Shared by tests/golden/make_spin_fixture.py and the
device drop-in test."""
import numpy as np


def spin_down_values(param, seed=7, n_changed=57):
    """The first n_changed values (decay, on-site and bond integrals for one
    kind) scaled by 1 + 0.02 N(0,1); the overlap block is kept."""
    rng = np.random.default_rng(seed)
    pdn = np.array(param, float)
    pdn[:n_changed] = pdn[:n_changed] * (1.0 + 0.02 * rng.standard_normal(n_changed))
    return pdn


def read_channel(text, nvalues=97, first_line=7):
    lines = text.splitlines()
    return np.array([float(l.split()[0]) for l in lines[first_line:first_line + nvalues]])


def nm_text(text, pdn, nvalues=97, first_line=7):
    """The parameter file `text` (one channel) as a two-channel file."""
    lines = text.splitlines()
    assert lines[0][:2] == "NN"
    head = "NM" + lines[0][2:]
    body = lines[1:first_line + nvalues]
    extra = [f"{v: .17E}" for v in pdn]
    return "\n".join([head] + body + extra) + "\n"


def read_skeng(path):
    """SKENG rows of a collinear magnetic run in modes 1, 2, 4: label,
    volume, E_F, earg1, earg2, smag (static.f:1246-1248)."""
    rows = []
    for line in open(path):
        if line.strip():
            t = line.split()
            rows.append((" ".join(t[:2]),) + tuple(float(x) for x in t[2:7]))
    return rows


def skin_spin(text, kind):
    """ex1's SKIN with the spin input of input1.f:518-545, 736-786 on every
    structure: kind "noncol" turns on non-collinear magnetisation ("FN")
    with the atom's direction theta = 30, phi = 45 degrees; kind "flip"
    allows spin flips ("FT") and flips the atom (magdir = 2)."""
    import re
    flag, extra = {"noncol": ("N", "  30.0 45.0"), "flip": ("T", "  2")}[kind]
    out, n1 = re.subn(r"(?m)^F (\s+\(Logical variable)", "F" + flag + r"\1", text)
    out, n2 = re.subn(r"(?m)^( 1  0\.000  0\.000  0\.000  0 0 0)", r"\1" + extra, out)
    assert n1 == n2 == 18
    return out


def read_skeng_rows(path):
    """SKENG rows as label and numbers (any column count)."""
    rows = []
    for line in open(path):
        if line.strip():
            t = line.split()
            rows.append([" ".join(t[:2])] + [float(x) for x in t[2:]])
    return rows
