"""Structures, NRL parameter sets, k-meshes and the seeded synthetic workloads.

A :class:`Structure` carries exactly what the reference's per-structure
COMMON blocks carry into ``bande`` (NRLtb/src/static.f:1031-1075):
lattice vectors, Cartesian atom positions, phase origins, kinds, the
neighbour-search widths, and the NRL parameter set of ``input.f:139-359``.

Units follow the reference: bohr, Ry, k in reciprocal-lattice units.
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Optional, Sequence

import numpy as np

BOHR_ANGSTROM = 0.529177210903          # 1 bohr in Angstrom
RYBOHR3_KBAR = 147105.164               # Ry/bohr^3 -> kbar (ex1/pdfit.gnu:121)
NPK_ONSITE, NPK_LIKE, NPK_UNLIKE = 4, 10, 14   # npkind (static.f:610)


def npard(kinds: int) -> int:
    """Number of parameters of one spin channel (input.f:275)."""
    return kinds * (29 + 68 * kinds)


@dataclasses.dataclass
class NRLParams:
    """One NRL parameter file (input.f:181-359), one spin channel."""

    kinds: int
    kbas: List[int]
    valence: List[float]
    dnn: List[float]          # cutoff R_c per interaction type
    screenl: List[float]      # screening length l per interaction type
    param: np.ndarray         # float64[npard(kinds)]
    nltype: int = 0           # 0 standard, 1 new overlap, 2 Harrison, 3 both
    jsover: bool = True       # non-orthogonal basis
    title: str = ""

    @property
    def realov(self) -> bool:
        return self.nltype in (1, 3)

    @property
    def harrison(self) -> bool:
        return self.nltype in (2, 3)


def read_nrl_params(path: str) -> NRLParams:
    """Read an NRL parameter file (format of input.f:181-359)."""
    with open(path) as fh:
        lines = fh.read().splitlines()
    head = lines[0]
    ovtype, nltype = head[0], int(head[2:7] or 0)
    jsover = ovtype.upper() != "O"
    title = lines[1].rstrip()
    kinds = int(lines[2].split()[0])
    pos = 3
    dnn, scr = [], []
    for _ in range(kinds * (kinds + 1) // 2):
        a, b = lines[pos].split()[:2]
        dnn.append(float(a)); scr.append(float(b)); pos += 1
    kbas, val = [], []
    for _ in range(kinds):
        kbas.append(int(lines[pos].split()[0])); pos += 1
        pos += 1                                   # atomic weight
        val.append(sum(float(x) for x in lines[pos].split()[:3])); pos += 1
    npar = npard(kinds)
    param = np.zeros(npar)
    for i in range(npar):
        if pos + i >= len(lines):
            break
        param[i] = float(lines[pos + i].split()[0])
    return NRLParams(kinds, kbas, val, dnn, scr, param, nltype, jsover, title)


@dataclasses.dataclass
class Structure:
    """One crystal structure as ``bande`` sees it."""

    lat: np.ndarray           # (3,3): lat[v] = lattice vector v (bohr)
    posn: np.ndarray          # (natoms,3) Cartesian (bohr)
    kkind: np.ndarray         # (natoms,) 1-based kinds
    params: NRLParams
    nd: Sequence[int] = (4, 4, 4)   # neighbour search half-widths
    posq: Optional[np.ndarray] = None   # (natoms,3) phase origins
    zcharge: float = 0.0      # electrons above the nominal valence
    label: str = ""

    def __post_init__(self):
        self.lat = np.ascontiguousarray(self.lat, dtype=np.float64)
        self.posn = np.ascontiguousarray(self.posn, dtype=np.float64)
        self.kkind = np.ascontiguousarray(self.kkind, dtype=np.int32)
        if self.posq is None:
            self.posq = np.zeros_like(self.posn)
        self.posq = np.ascontiguousarray(self.posq, dtype=np.float64)

    @property
    def natoms(self) -> int:
        return len(self.kkind)

    @property
    def ksk(self) -> np.ndarray:
        """First orbital of each atom (input1.f:835-836)."""
        kb = np.array([self.params.kbas[k - 1] for k in self.kkind], np.int32)
        return np.concatenate([[0], np.cumsum(kb)[:-1]]).astype(np.int32)

    @property
    def nv(self) -> int:
        return int(sum(self.params.kbas[k - 1] for k in self.kkind))

    @property
    def zfill(self) -> float:
        """Electron count (input1.f:733, 834)."""
        z = self.zcharge
        for k in self.kkind:
            z = z + self.params.valence[k - 1]
        return z

    @property
    def volume(self) -> float:
        """Cell volume with setlat.f:997-1000's expansion (a(i,j) = comp i of vec j)."""
        a = self.lat.T
        return abs(a[0, 0] * (a[1, 1] * a[2, 2] - a[2, 1] * a[1, 2]) +
                   a[1, 0] * (a[2, 1] * a[0, 2] - a[0, 1] * a[2, 2]) +
                   a[2, 0] * (a[0, 1] * a[1, 2] - a[1, 1] * a[0, 2]))


# ---------------------------------------------------------------------------
# lattices and k-meshes
# ---------------------------------------------------------------------------

def fcc_primitive(a: float) -> np.ndarray:
    """fcc primitive vectors as setlat.f builds them (ex1/SKOUT:121-123)."""
    h = 0.5 * a
    return np.array([[0.0, h, h], [h, 0.0, h], [h, h, 0.0]])


def bcc_primitive(a: float) -> np.ndarray:
    """bcc primitive vectors (ex1/SKOUT:576-578)."""
    h = 0.5 * a
    return np.array([[-h, h, h], [h, -h, h], [h, h, -h]])


def conventional_cell(kind: str, a: float, c_over_a: float = math.sqrt(8.0 / 3.0)):
    """Lattice vectors and fractional basis of a conventional cell."""
    if kind == "fcc":
        lat = np.eye(3) * a
        frac = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0]])
    elif kind == "bcc":
        lat = np.eye(3) * a
        frac = np.array([[0, 0, 0], [.5, .5, .5]])
    elif kind == "hcp":
        c = c_over_a * a
        lat = np.array([[a, 0, 0], [-0.5 * a, 0.5 * math.sqrt(3) * a, 0], [0, 0, c]])
        frac = np.array([[1 / 3, 2 / 3, 0.25], [2 / 3, 1 / 3, 0.75]])
    else:
        raise ValueError(kind)
    return lat, frac


def supercell(lat, frac, reps):
    """Replicate a cell reps=(n1,n2,n3) times; returns (lat, Cartesian posn)."""
    reps = tuple(int(r) for r in reps)
    pos = []
    for i in range(reps[0]):
        for j in range(reps[1]):
            for k in range(reps[2]):
                for f in frac:
                    pos.append((np.asarray(f) + [i, j, k]) @ lat)
    return lat * np.array(reps)[:, None], np.array(pos)


def search_widths(lat: np.ndarray, rcut: float) -> List[int]:
    """Smallest nd per direction so that every image within rcut is visited."""
    rec = np.linalg.inv(lat).T          # rows: reciprocal vectors / 2pi
    # distance between lattice planes of family i = 1/|b_i|
    return [int(math.ceil(rcut * np.linalg.norm(rec[i]) + 1e-9)) + 1 for i in range(3)]


def monkhorst_pack(q: Sequence[int]) -> np.ndarray:
    """Full Monkhorst-Pack grid u_r = (2r-q-1)/(2q) (kptgen.f:99-108 LSPEC=T)."""
    axes = [[(2.0 * r - qq - 1.0) / (2.0 * qq) for r in range(1, qq + 1)] for qq in q]
    k = np.array([[a, b, c] for a in axes[0] for b in axes[1] for c in axes[2]])
    return k


# ---------------------------------------------------------------------------
# synthetic NRL parameters
# ---------------------------------------------------------------------------

def pd_like_params(rcut: float = 16.5, screenl: float = 0.5) -> NRLParams:
    """Load the full 97-value Pd set from the committed fixture."""
    import json, os
    here = os.path.dirname(os.path.abspath(__file__))
    fix = os.path.join(here, "data", "pd_par.json")
    with open(fix) as fh:
        d = json.load(fh)
    return NRLParams(1, [9], [11.0], [rcut], [screenl], np.array(d["param"]),
                     0, True, "Pd-centred synthetic")


def synthetic_params(seed: int, rcut: float = 16.5, screenl: float = 0.5,
                     r0: float = 4.824, valence: float = 11.0,
                     noise: float = 0.03, overlap_scale: float = 0.65) -> NRLParams:
    """Seeded NRL-form parameters for one sp3d5 kind.

    Draws centre on the Pd set.  Each bond integral (e + fR + fbar R^2)
    exp(-d^2 R) (setpar.f:592-598) gets d ~ U[0.8, 1.2]; its polynomial is
    rescaled so that its value at the nearest-neighbour distance r0 is the
    Pd value times (1 + noise*N(0,1)).  Overlap integrals keep the Pd decay
    constants (times 1 + noise*N(0,1), clipped to [0.8, 1.2]) and are scaled
    by overlap_scale, which keeps S(k) positive definite at every k for
    close-packed cells at a = 3.61 A with a 16.5 bohr cutoff.  On-site
    coefficients get the same relative noise.
    """
    rng = np.random.default_rng(seed)
    base = pd_like_params(rcut, screenl)
    p = base.param.copy()
    p[0] = p[0] * (1.0 + noise * rng.standard_normal())          # lambda
    p[1:17] = p[1:17] * (1.0 + noise * rng.standard_normal(16))  # on-site
    for j in range(10):                                           # H
        i = 17 + 4 * j
        e, f, fb, g = p[i:i + 4]
        dnew = rng.uniform(0.8, 1.2)
        scale = math.exp(-(g * g - dnew * dnew) * r0) * (1.0 + noise * rng.standard_normal())
        p[i:i + 3] = np.array([e, f, fb]) * scale
        p[i + 3] = dnew
    for j in range(10):                                           # S
        i = 57 + 4 * j
        g = p[i + 3]
        dnew = min(1.2, max(0.8, g * (1.0 + noise * rng.standard_normal())))
        scale = math.exp(-(g * g - dnew * dnew) * r0) * overlap_scale
        p[i:i + 3] = p[i:i + 3] * scale
        p[i + 3] = dnew
    return NRLParams(1, [9], [valence], [rcut], [screenl], p, 0, True,
                     f"synthetic seed {seed}")


def synthetic_structure(kind: str, reps, a_bohr: float, params: NRLParams,
                        sigma_bohr: float = 0.0, strain: float = 0.0,
                        seed: int = 0, zcharge: float = 0.0) -> Structure:
    """A displaced (and optionally strained) supercell of one species."""
    rng = np.random.default_rng(seed)
    lat, frac = conventional_cell(kind, a_bohr)
    lat, pos = supercell(lat, frac, reps)
    if strain:
        e = rng.uniform(-strain, strain, (3, 3))
        e = 0.5 * (e + e.T)
        m = np.eye(3) + e
        lat = lat @ m
        pos = pos @ m
    if sigma_bohr:
        pos = pos + sigma_bohr * rng.standard_normal(pos.shape)
    # phase origins: fractional coordinates of the atoms (input1.f:795)
    posq = pos @ np.linalg.inv(lat)
    nd = search_widths(lat, params.dnn[0])
    return Structure(lat, pos, np.ones(len(pos), np.int32), params, nd, posq,
                     zcharge, f"{kind}{tuple(reps)}")


ANGSTROM = 1.0 / BOHR_ANGSTROM


def config_c1(seed: int = 20261203) -> Structure:
    """4-atom conventional fcc, a = 3.61 A, N = 36."""
    return synthetic_structure("fcc", (1, 1, 1), 3.61 * ANGSTROM, synthetic_params(seed))


def config_c2(seed: int = 20261203) -> Structure:
    """32-atom fcc 2x2x2 supercell, sigma = 0.05 A, N = 288."""
    return synthetic_structure("fcc", (2, 2, 2), 3.61 * ANGSTROM, synthetic_params(seed),
                               sigma_bohr=0.05 * ANGSTROM, seed=seed + 1)


def config_c3(seed: int = 20261203) -> Structure:
    """108-atom fcc 3x3x3 supercell, sigma = 0.03 A, <=1% strain, N = 972."""
    return synthetic_structure("fcc", (3, 3, 3), 3.61 * ANGSTROM, synthetic_params(seed),
                               sigma_bohr=0.03 * ANGSTROM, strain=0.01, seed=seed + 2)


def primitive_cell(kind: str, a: float, c_over_a: float = math.sqrt(8.0 / 3.0)):
    """Lattice vectors and fractional basis of the primitive cell (1 atom for
    fcc and bcc, 2 for hcp)."""
    if kind == "fcc":
        return fcc_primitive(a), np.zeros((1, 3))
    if kind == "bcc":
        return bcc_primitive(a), np.zeros((1, 3))
    return conventional_cell("hcp", a, c_over_a)


def uneven_structure(kind: str, reps, a_bohr: float, params: NRLParams, sigma_bohr: float,
                     seed: int) -> Structure:
    """Displaced supercell of a primitive cell."""
    rng = np.random.default_rng(seed)
    lat, frac = primitive_cell(kind, a_bohr)
    lat, pos = supercell(lat, frac, reps)
    if sigma_bohr:
        pos = pos + sigma_bohr * rng.standard_normal(pos.shape)
    posq = pos @ np.linalg.inv(lat)
    nd = search_widths(lat, params.dnn[0])
    return Structure(lat, pos, np.ones(len(pos), np.int32), params, nd, posq, 0.0,
                     f"{kind}{tuple(reps)}")


# nearest-neighbour distance of fcc at a = 3.61 A, kept for bcc and hcp
_NN = 3.61 / math.sqrt(2.0)
_A_OF = {"fcc": _NN * math.sqrt(2.0), "bcc": _NN * 2.0 / math.sqrt(3.0), "hcp": _NN}


def config_c4(seed: int = 20261203, count: int = 256, max_atoms: int = 64, max_q: int = 8,
              nk_budget: float = 9 * 8 ** 3):
    """Uneven batch: `count` seeded structures, each fcc, bcc or hcp
    (uniform), 1..max_atoms atoms (log-uniform; supercells of the primitive
    cell with n_i <= 4), lattice constant +-3% around the same
    nearest-neighbour distance, sigma = 0.04 A displacements, Monkhorst-Pack
    q^3 grids (2 <= q <= max_q) with N q^3 close to nk_budget.  Returns
    [(structure, q)]."""
    rng = np.random.default_rng(seed)
    params = synthetic_params(seed)
    triples = [(a, b, c) for a in range(1, 5) for b in range(1, 5) for c in range(1, 5)]
    out = []
    for i in range(count):
        kind = ("fcc", "bcc", "hcp")[int(rng.integers(3))]
        target = math.exp(rng.uniform(0.0, math.log(max_atoms)))
        per = 2 if kind == "hcp" else 1
        best = min(triples, key=lambda t: (abs(math.log(per * t[0] * t[1] * t[2] / target)),
                                           max(t) - min(t)))
        a = _A_OF[kind] * (1.0 + rng.uniform(-0.03, 0.03)) * ANGSTROM
        st = uneven_structure(kind, best, a, params, 0.04 * ANGSTROM, seed + 100 + i)
        q = int(min(max_q, max(2, round((nk_budget / st.nv) ** (1.0 / 3.0)))))
        out.append((st, q))
    return out
