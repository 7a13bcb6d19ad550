"""The reference program's files through the library's native readers and
writers: NRL parameter files (input.f), SKIN control files (static.f,
setlat.f, input1.f, kptin.f), SKENG lines (static.f:1238-1254), and the
program's structure loop on the device (``static_run``).

Everything is done in lib/libtbk.so (tbk_read_params, tbk_skin_*,
tbk_skeng_line, tbk_static_run); this module only marshals arrays.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import _lib
from .structures import NRLParams, Structure


def read_params(path: str):
    """(NRLParams of channel 1, all channels [jspins, npard], the header
    as a ParamSet) of an NRL parameter file."""
    L = _lib.lib()
    ps = _lib.ParamSet()
    cap = 2 * _lib.TBK_MAXKINDS * (29 + 68 * _lib.TBK_MAXKINDS)
    buf = np.zeros(cap)
    _lib.check(L.tbk_read_params(os.fsencode(path), C.byref(ps), cap, buf), f"tbk_read_params({path})")
    k = ps.kinds
    npard = k * (29 + 68 * k)
    chans = buf[:ps.jspins * npard].reshape(ps.jspins, npard).copy()
    nint = k * (k + 1) // 2
    p = NRLParams(k, list(ps.kbas[:k]), list(ps.valence[:k]), list(ps.dnn[:nint]), list(ps.screenl[:nint]),
                  chans[0].copy(), int(ps.nltype), bool(ps.jsover), ps.title.decode(errors="replace").rstrip())
    return p, chans, ps


@dataclass
class SkinStructure:
    """One structure of a SKIN file as the reference program reads it."""
    title: str
    label: str
    structure: Structure
    thisvol: float
    zfill: float
    kpts: np.ndarray
    weights: np.ndarray
    latnum: int
    strain: int
    flipos: np.ndarray
    spinat: np.ndarray
    flags: dict = field(default_factory=dict)


class Skin:
    """A SKIN file opened with tbk_skin_open; iterate for its structures."""

    def __init__(self, path: str, maxatoms: int = 4096, maxk: int = 50000):
        L = _lib.lib()
        self.h = C.c_void_p()
        mode, tkb, eig = C.c_int(), C.c_double(), C.c_double()
        pp = C.create_string_buffer(1024)
        _lib.check(L.tbk_skin_open(os.fsencode(path), C.byref(self.h), C.byref(mode), C.byref(tkb),
                                   C.byref(eig), pp, 1024), f"tbk_skin_open({path})")
        self.mode, self.tkb, self.eigcut = mode.value, tkb.value, eig.value
        self.param_path = pp.value.decode()
        self.params, self.channels, self._ps = read_params(self.param_path)
        self.maxatoms, self.maxk = maxatoms, maxk

    def __iter__(self):
        return self

    def __next__(self) -> SkinStructure:
        L = _lib.lib()
        n, m = self.maxatoms, self.maxk
        st = _lib.SkinStructure()
        kk = np.zeros(n, np.int32); fl = np.zeros(n, np.int32)
        posn, posq, sp = np.zeros(3 * n), np.zeros(3 * n), np.zeros(3 * n)
        q, w = np.zeros(3 * m), np.zeros(m)
        rc = L.tbk_skin_next(self.h, C.byref(self._ps), n, m, C.byref(st), kk, posn, posq, fl, sp, q, w)
        if rc == _lib.TBK_END:
            raise StopIteration
        _lib.check(rc, f"tbk_skin_next (inpstat {st.inpstat})")
        na, nk = st.natoms, st.nk
        lat = np.array(st.lat[:], float).reshape(3, 3)
        s = Structure(lat, posn[:3 * na].reshape(na, 3).copy(), kk[:na].copy(), self.params, list(st.nd[:]),
                      posq[:3 * na].reshape(na, 3).copy(), st.zcharge, st.label.decode(errors="replace"))
        flags = dict(lphonon=st.lphonon, spinflip=st.spinflip, lnoncol=st.lnoncol, ltorque=st.ltorque,
                     lspdir=st.lspdir, inpstat=st.inpstat)
        return SkinStructure(st.title.decode(errors="replace"), st.label.decode(errors="replace"), s, st.thisvol,
                             st.zfill, q[:3 * nk].reshape(nk, 3).copy(), w[:nk].copy(), st.latnum, st.strain,
                             fl[:na].copy(), sp[:3 * na].reshape(na, 3).copy(), flags)

    def close(self):
        if getattr(self, "h", None):
            _lib.lib().tbk_skin_close(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def skeng_line(label: str, thisvol: float, efermi: float, earg1: float, earg2: float = 0.0,
               smag: float = 0.0, mode: int = 4, with_smag: bool = False) -> str:
    """One SKENG line (static.f:1238-1254)."""
    buf = C.create_string_buffer(160)
    _lib.check(_lib.lib().tbk_skeng_line(label.encode(), thisvol, efermi, earg1, earg2, smag, int(mode),
                                         int(with_smag), buf, 160), "tbk_skeng_line")
    return buf.value.decode()


def static_run(directory: str, device: int = 0, precision: int = _lib.TBK_FP64) -> int:
    """Run directory/SKIN through the device plans and write
    directory/SKENG, as the reference program does; returns the number of
    structures done (raises TbkError where the program would stop)."""
    n = C.c_int()
    _lib.check(_lib.lib().tbk_static_run(os.fsencode(directory), int(device), int(precision), C.byref(n)),
               "tbk_static_run")
    return n.value
