"""k-point meshes for the plans: reciprocal lattice, space-group operations,
the irreducible mesh and kptin's weights, through the C ABI of
lib/libtbk.so (tbk_reciprocal_lattice, tbk_space_group, tbk_space_group_read,
tbk_kptgen, tbk_kmesh_weights, tbk_klist, tbk_kpath).

The k-points come back in reciprocal-lattice coordinates with weights
summing to one: the arguments :class:`tbk.api.Plan` takes.  Each function
follows the reference routine it names (kptin.f, rdspgp.f, kptgen.f) bit for
bit, including its tolerances; the reference's STOPs raise
:class:`tbk._lib.TbkError`.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

from . import _lib

NKD = 50000          # the reference's nkd (NRLtb/ex1/P3): most k-points
NOPD = 48            # nopd: most operations
NKBZD = 2 * NKD      # kptgen.f:33: largest full mesh

# lattice codes of kptin.f:84 that kptgen.f:72-74 treats specially
LATTIC_BCC, LATTIC_HEX = 2, 4
LATTIC_BCT = (7, 8, 9)


def _check(rc: int, what: str):
    if rc != 0:
        raise _lib.TbkError(rc, what)


@dataclass
class Lattice:
    """kptin's /LATTYP/: rows of avec are the lattice vectors, rows of bvec
    the reciprocal ones (with 2 pi), bij the reciprocal metric."""
    avec: np.ndarray
    bvec: np.ndarray
    bij: np.ndarray
    wsvol: float
    lattic: int


@dataclass
class SpaceGroup:
    """rdspgp's /SPACE/: rspc[k] in Cartesian coordinates, rspct[k] in
    reciprocal-lattice coordinates, tnpspc[k] the non-primitive translation
    in direct-lattice coordinates."""
    name: str
    rspc: np.ndarray
    rspct: np.ndarray
    tnpspc: np.ndarray

    @property
    def nop(self) -> int:
        return self.rspc.shape[0]


def reciprocal_lattice(lat, lattic: int = 0) -> Lattice:
    """Lattice of plv rows ``lat`` (bohr), kptin.f:94-155."""
    L = _lib.lib()
    lat = np.ascontiguousarray(lat, dtype=np.float64).reshape(9)
    a, b, g = np.zeros(9), np.zeros(9), np.zeros(9)
    vol = C.c_double()
    _check(L.tbk_reciprocal_lattice(lat, a, b, g, C.byref(vol)), "tbk_reciprocal_lattice")
    return Lattice(a.reshape(3, 3), b.reshape(3, 3), g.reshape(3, 3), vol.value, int(lattic))


def space_group(lattice: Lattice, rot, tnp, ilat: bool, shift=None, name: str = "") -> SpaceGroup:
    """Operations from rotations as read (direct-lattice coordinates if
    ``ilat`` else Cartesian) and translations, rdspgp.f:106-289."""
    L = _lib.lib()
    rot = np.ascontiguousarray(rot, dtype=np.float64).reshape(-1, 9)
    nop = rot.shape[0]
    tnp = np.ascontiguousarray(tnp, dtype=np.float64).reshape(nop * 3)
    sh = None
    if shift is not None:
        sh_arr = np.ascontiguousarray(shift, dtype=np.float64).reshape(3)
        sh = sh_arr.ctypes.data_as(C.c_void_p)
    rspc, rspct, tn = np.zeros(9 * nop), np.zeros(9 * nop), np.zeros(3 * nop)
    _check(L.tbk_space_group(int(bool(ilat)), lattice.lattic, _flat(lattice.avec), _flat(lattice.bvec),
                             nop, rot.reshape(-1), tnp, sh, rspc, rspct, tn), "tbk_space_group")
    return SpaceGroup(name, rspc.reshape(nop, 3, 3), rspct.reshape(nop, 3, 3), tn.reshape(nop, 3))


def read_space_group(path: str, lattice: Lattice, ilat: bool, maxop: int = NOPD) -> SpaceGroup:
    """A space-group file in rdspgp's fixed format (rdspgp.f:117-147)."""
    L = _lib.lib()
    if not os.path.exists(path):
        raise FileNotFoundError(path)
    rspc, rspct, tn = np.zeros(9 * maxop), np.zeros(9 * maxop), np.zeros(3 * maxop)
    nop = C.c_int()
    name = C.create_string_buffer(11)
    _check(L.tbk_space_group_read(os.fsencode(path), int(bool(ilat)), lattice.lattic,
                                  _flat(lattice.avec), _flat(lattice.bvec), maxop, C.byref(nop),
                                  rspc, rspct, tn, name), f"tbk_space_group_read({path})")
    n = nop.value
    return SpaceGroup(name.value.decode(errors="replace").strip(), rspc[:9 * n].reshape(n, 3, 3),
                      rspct[:9 * n].reshape(n, 3, 3), tn[:3 * n].reshape(n, 3))


def kptgen(lattice: Lattice, group: SpaceGroup, kdiv, lspec: bool, info=None, sfk: float = 1.0,
           kpst: int = 0, kpend: int = 0, maxfull: int = NKBZD, maxk: int = NKD):
    """Irreducible mesh and weights as kptin hands them to bande (its
    generated-mesh branch, kptin.f:276-351): kptgen.f, then the
    information-only points ``info`` (k in units of 1/sfk, weight zero),
    then the weight normalisation and the kpst..kpend range.  Returns
    (q [n, 3], w [n])."""
    L = _lib.lib()
    kd = np.ascontiguousarray(kdiv, dtype=np.int32).reshape(3)
    extra = 0 if info is None else len(info)
    cap = maxk + extra
    q, w = np.zeros(3 * cap), np.zeros(cap)
    nk = C.c_int()
    _check(L.tbk_kptgen(lattice.lattic, _flat(lattice.avec), _flat(lattice.bvec), _flat(lattice.bij),
                        lattice.wsvol, group.nop, _flat(group.rspct), kd, int(bool(lspec)), maxfull, maxk,
                        C.byref(nk), q, w), "tbk_kptgen")
    n = nk.value
    if extra:
        s = sfk if sfk > 1e-6 else 1.0
        pts = np.asarray(info, dtype=np.float64).reshape(extra, 3)
        for i in range(extra):
            q[3 * (n + i):3 * (n + i) + 3] = pts[i] / s
            w[n + i] = 0.0
        n += extra
        if n > maxk:
            raise _lib.TbkError(_lib.TBK_ERR_ARG, "information points exceed nkd")
    out = C.c_int()
    _check(L.tbk_kmesh_weights(n, q, w, kpst, kpend, C.byref(out)), "tbk_kmesh_weights")
    m = out.value
    return q[:3 * m].reshape(m, 3).copy(), w[:m].copy()


def klist(raw, sfk: float = 1.0):
    """An explicit list (rows k1 k2 k3 weight, k in units of 1/sfk), as
    kptin reads the -1313 / nkv > 0 files (kptin.f:266-273)."""
    L = _lib.lib()
    raw = np.ascontiguousarray(raw, dtype=np.float64).reshape(-1, 4)
    n = raw.shape[0]
    q, w = np.zeros(3 * n), np.zeros(n)
    _check(L.tbk_klist(n, raw.reshape(-1), float(sfk), q, w), "tbk_klist")
    return q.reshape(n, 3), w


def kpath(segments, maxk: int = NKD):
    """Line mode (nkv = 999, kptin.f:223-251): ``segments`` of (kd, start,
    end).  Every weight is the point count, as the reference leaves it."""
    L = _lib.lib()
    kd = np.ascontiguousarray([s[0] for s in segments], dtype=np.int32)
    ends = np.ascontiguousarray([list(s[1]) + list(s[2]) for s in segments], dtype=np.float64)
    q, w = np.zeros(3 * maxk), np.zeros(maxk)
    nk = C.c_int()
    _check(L.tbk_kpath(len(segments), kd, ends.reshape(-1), maxk, C.byref(nk), q, w), "tbk_kpath")
    n = nk.value
    return q[:3 * n].reshape(n, 3).copy(), w[:n].copy()


def _flat(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
