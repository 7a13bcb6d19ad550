"""ctypes front end of the oracle (test infrastructure only).

Two libraries live here:

* ``libnrl_oracle.so`` -- the C restatement (nrl_oracle.c) of the reference's
  per-k path; the checker for the HIP product and the CPU baseline.
* ``_ref/libnrlref.so`` -- the reference's own Fortran (NRLtb/ex1 sources,
  compiled where they lie) behind the bind(C) driver ref/ref_driver.f.  It
  exists only where the reference tree was present at build time and is
  used to pin the restatement.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg
import this package.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "libnrl_oracle.so")
REF_SO = os.path.join(HERE, "_ref", "libnrlref.so")
# the same routines built with matom = 128, nkd = 1024 (make -C oracle refbig)
REF_SO_BIG = os.path.join(HERE, "_ref", "big", "libnrlref_big.so")

_d = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_i = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_c = np.ctypeslib.ndpointer(dtype=np.complex128, flags="C_CONTIGUOUS")

_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"oracle not built: {ORACLE_SO} (run make -C oracle)")
        L = C.CDLL(ORACLE_SO)
        L.orc_search2.argtypes = [C.c_int, _d, _d, _i, C.c_int, _i, _d, _d, C.c_int,
                                  C.POINTER(C.c_int), _d, _d, _d, _d, _d, _i, _i,
                                  C.POINTER(C.c_double)]
        L.orc_setpar.argtypes = [C.c_int, _i, C.c_int, _d, C.c_int, C.c_int, C.c_int,
                                 C.c_int, C.c_int, _d, _d, _i, _i, _d, _d, _d, _d]
        L.orc_setparv.argtypes = [C.c_int, _i, C.c_int, _d, C.c_int, C.c_int, C.c_int,
                                  C.c_int, C.c_double, C.c_int, _d, _d, _d, _i, _i, _d,
                                  _d, _d, _d]
        L.orc_setham.argtypes = [C.c_int, _d, C.c_int, _i, _i, _i, _d, C.c_int, _d, _d,
                                 _i, _i, _d, _d, _d, C.c_int, C.c_int, _c, _c]
        L.orc_setham.restype = None
        L.orc_unpack.argtypes = [C.c_int, _c, _c]
        L.orc_unpack.restype = None
        L.orc_zhegv.argtypes = [C.c_int, _c, _c, _d, C.c_void_p]
        L.orc_perturb.argtypes = [C.c_int, _c, _c, _c, _d, C.c_double,
                                  C.POINTER(C.c_int), _d]
        L.orc_fermi.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, _d, _d,
                                C.c_double, C.c_double]
        L.orc_fermi.restype = C.c_double
        L.orc_bande_sums.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, _d,
                                     _d, _d, _i, C.c_double, C.c_double,
                                     C.POINTER(C.c_double), C.POINTER(C.c_double),
                                     C.POINTER(C.c_double)]
        L.orc_bande_sums.restype = None
        L.orc_bande.argtypes = [C.c_int, C.c_int, C.c_int, _i, _i, _i, _d, C.c_int, _d,
                                _d, _i, _i, _d, _d, _d, _d, _d, _d, C.c_int, _d, _d,
                                C.c_double, C.c_double, C.c_double, C.c_int,
                                C.POINTER(C.c_double), C.POINTER(C.c_double),
                                C.POINTER(C.c_double), C.c_void_p, C.c_void_p]
        _lib = L
    return _lib


_ref_path = REF_SO


def ref_available(big: bool = False) -> bool:
    return os.path.exists(REF_SO_BIG if big else REF_SO)


def use_big_reference():
    """Make ref() load the build with matom = 128 (structures of up to 128
    atoms, up to 1024 k-points per call).  Only before the first ref()."""
    global _ref_path
    if _ref is not None and _ref_path != REF_SO_BIG:
        raise RuntimeError("the ex1-dimensioned reference is already loaded in this process")
    _ref_path = REF_SO_BIG


def ref():
    """The reference's own routines (only where it was built)."""
    global _ref
    if _ref is None:
        if not os.path.exists(_ref_path):
            raise RuntimeError(f"reference build not present: {_ref_path} (make -C oracle ref / refbig)")
        L = C.CDLL(_ref_path)
        ip = C.POINTER(C.c_int)
        dp = C.POINTER(C.c_double)
        L.ref_setup.argtypes = [ip, _d, _d, _d, _i, _i, ip, _i, _d, _d, _d, _d, ip, ip,
                                ip, ip, dp, ip]
        L.ref_get_pairs.argtypes = [_d, _d, _d, _d, _d, _i, _i]
        L.ref_setpar.argtypes = [dp, _d, _d, _d, _d, _d, _d, _d, ip]
        L.ref_setham.argtypes = [_d, ip, _c, _c]
        L.ref_diag.argtypes = [ip, _c, _c, _d, _c, ip]
        L.ref_perturb.argtypes = [_c, _c, _d, dp, ip, _d]
        L.ref_fermi.argtypes = [ip, dp, _d, _d, dp, dp]
        L.ref_fermi.restype = C.c_double
        L.ref_bande.argtypes = [ip, ip, _d, _d, dp, dp, dp, dp, dp, dp, dp, ip]
        L.ref_setpar_spin.argtypes = [ip, dp, _d, ip, _i, _d, _d, _d, _d, _d, _d, _d, ip]
        L.ref_setpar_nc.argtypes = [dp, _d, ip, _d, _d, _d, _d, _d, _d, _d, _d, _d, ip]
        L.ref_bande_nc.argtypes = [ip, ip, _d, _d, dp, dp, dp, dp, _d, ip, _d, dp, dp, dp, ip]
        L.ref_setham_nc.argtypes = [_d, ip, _d, _d, _c, _c, ip]
        L.ref_bande_spin.argtypes = [ip, ip, _d, _d, dp, dp, dp, dp, _d, ip, _i, dp, dp, dp, dp, ip]
        L.ref_kptin.argtypes = [_d, C.c_char_p, ip, C.c_char_p, ip, ip, ip, _d, _d, _d, _d, _d,
                                dp, ip, ip, _d, _d, _d, ip]
        L.ref_input.argtypes = [C.c_char_p, ip, ip, ip, ip, ip, ip, _i, _d, _d, _d, _d, ip]
        L.ref_input1.argtypes = [C.c_char_p, ip, ip, _i, _d, dp, _d, dp, ip, _i, _i, _d, _d, dp, _i, ip, _d,
                                 ip]
        _ref = L
    return _ref


def ref_kptin(lat, lines, workdir, maxk: int = 50000) -> dict:
    """The reference's kptin (with its rdspgp and kptgen) on the NEWSYM block
    ``lines`` of a SKIN (written to workdir/kin), for the lattice rows
    ``lat``: the k-points and weights bande receives, /LATTYP/ and /SPACE/."""
    L = ref()
    fin = os.path.join(workdir, "kin")
    with open(fin, "w") as f:
        f.write("\n".join(lines) + "\n")
    fscr = os.path.join(workdir, "kin.15")
    a = np.ascontiguousarray(np.asarray(lat, float))       # rows = vectors: lat(c, v) in Fortran
    q, w = np.zeros(3 * maxk), np.zeros(maxk)
    av, bv, bij = np.zeros(9), np.zeros(9), np.zeros(9)
    rsp, rspt, tnp = np.zeros(9 * 48), np.zeros(9 * 48), np.zeros(3 * 48)
    nk, lattic, nop, ierr = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    vol = C.c_double()
    fb, sb = os.fsencode(fin), os.fsencode(fscr)
    L.ref_kptin(a.ravel(), fb, C.byref(C.c_int(len(fb))), sb, C.byref(C.c_int(len(sb))),
                C.byref(C.c_int(maxk)), C.byref(nk), q, w, av, bv, bij, C.byref(vol),
                C.byref(lattic), C.byref(nop), rsp, rspt, tnp, C.byref(ierr))
    if ierr.value:
        raise RuntimeError(f"ref_kptin: ierr {ierr.value}")
    n, m = nk.value, nop.value
    # Fortran column-major (i, j) -> rows i, columns j
    f3 = lambda x: x.reshape(3, 3).T.copy()
    return dict(q=q[:3 * n].reshape(n, 3).copy(), w=w[:n].copy(), avec=f3(av).T.copy(),
                bvec=f3(bv).T.copy(), bij=f3(bij), wsvol=vol.value, lattic=lattic.value,
                rspc=rsp[:9 * m].reshape(m, 3, 3).transpose(0, 2, 1).copy(),
                rspct=rspt[:9 * m].reshape(m, 3, 3).transpose(0, 2, 1).copy(),
                tnpspc=tnp[:3 * m].reshape(m, 3).copy())


class Setup:
    """Pair list + SK parameters of one structure, from the C restatement."""

    def __init__(self, st, maxpairs: Optional[int] = None):
        L = lib()
        p = st.params
        nat = st.natoms
        maxpairs = maxpairs or max(1000, 400 * nat)
        self.st = st
        self.tt = np.zeros(3 * maxpairs); self.dlv = np.zeros(3 * maxpairs)
        self.dist = np.zeros(maxpairs); self.screen = np.zeros(maxpairs)
        self.dscrn = np.zeros(maxpairs)
        self.jkind = np.zeros(2 * maxpairs, np.int32); self.jatm = np.zeros(2 * maxpairs, np.int32)
        npairs = C.c_int(0); dmin = C.c_double(0)
        ierr = L.orc_search2(nat, np.ascontiguousarray(st.lat.ravel()),
                             np.ascontiguousarray(st.posn.ravel()), st.kkind, p.kinds,
                             np.array(st.nd, np.int32), np.array(p.dnn, float),
                             np.array(p.screenl, float), maxpairs, C.byref(npairs),
                             self.tt, self.dlv, self.dist, self.screen, self.dscrn,
                             self.jkind, self.jatm, C.byref(dmin))
        if ierr:
            raise RuntimeError("pair overflow")
        self.npairs = npairs.value
        self.dismin = dmin.value
        n = self.npairs
        self.den = np.zeros(nat * p.kinds); self.paros = np.zeros(4 * nat)
        self.hop = np.zeros(14 * max(n, 1)); self.over = np.zeros(14 * max(n, 1))
        self.dparos = np.zeros(4 * nat)
        self.dhop = np.zeros(14 * max(n, 1)); self.dover = np.zeros(14 * max(n, 1))
        args = (nat, st.kkind, p.kinds, np.ascontiguousarray(p.param, float), p.nltype,
                int(p.realov), int(p.harrison), int(p.jsover))
        L.orc_setpar(*args, n, self.dist, self.screen, self.jkind, self.jatm,
                     self.den, self.paros, self.hop, self.over)
        L.orc_setparv(*args, st.volume, n, self.dist, self.screen, self.dscrn,
                      self.jkind, self.jatm, self.den, self.dparos, self.dhop, self.dover)
        self.kbas = np.array(p.kbas, np.int32)
        self.ksk = st.ksk
        self.nv = st.nv
        self.posq = np.ascontiguousarray(st.posq.ravel())

    def hs(self, ak, deriv: bool = False):
        """Full Hermitian H(k), S(k) (column-major as Fortran, returned as
        numpy arrays indexed [row, col])."""
        L = lib()
        nv = self.nv
        np_ = nv * (nv + 1) // 2
        hp = np.zeros(np_, np.complex128); sp = np.zeros(np_, np.complex128)
        if deriv:
            L.orc_setham(nv, np.asarray(ak, float), self.st.natoms, self.st.kkind, self.ksk,
                         self.kbas, self.posq, self.npairs, self.tt, self.dlv, self.jkind,
                         self.jatm, self.dhop, self.dover, self.dparos, 1, 0, hp, sp)
        else:
            L.orc_setham(nv, np.asarray(ak, float), self.st.natoms, self.st.kkind, self.ksk,
                         self.kbas, self.posq, self.npairs, self.tt, self.dlv, self.jkind,
                         self.jatm, self.hop, self.over, self.paros, 1, 1, hp, sp)
        return hp, sp

    @staticmethod
    def unpack(n, ap):
        a = np.zeros(n * n, np.complex128)
        lib().orc_unpack(n, np.ascontiguousarray(ap), a)
        return a.reshape(n, n).T.copy()   # column-major -> [row, col]

    def bande(self, kpts, weights, mode=4, tkb=0.005, eigcut=10.0, nthreads=1,
              want_eig=False):
        L = lib()
        nk = len(kpts)
        ef, e1, e2 = C.c_double(0), C.c_double(0), C.c_double(0)
        eig = np.zeros(nk * self.nv) if want_eig else None
        desk = np.zeros(nk * self.nv) if want_eig else None
        err = L.orc_bande(mode, self.nv, self.st.natoms, self.st.kkind, self.ksk, self.kbas,
                          self.posq, self.npairs, self.tt, self.dlv, self.jkind, self.jatm,
                          self.hop, self.over, self.paros, self.dhop, self.dover,
                          self.dparos, nk, np.ascontiguousarray(np.asarray(kpts, float).ravel()),
                          np.ascontiguousarray(np.asarray(weights, float)), self.st.zfill,
                          tkb, eigcut, nthreads, C.byref(ef), C.byref(e1), C.byref(e2),
                          eig.ctypes.data if want_eig else None,
                          desk.ctypes.data if want_eig else None)
        out = dict(err=err, efermi=ef.value, earg1=e1.value, earg2=e2.value)
        if want_eig:
            out["eig"] = eig.reshape(nk, self.nv)
            out["desk"] = desk.reshape(nk, self.nv)
        return out


def zhegv(h, s, vectors=True):
    """Oracle generalized eigensolve of full Hermitian h, s ([row, col])."""
    n = h.shape[0]
    hc = np.ascontiguousarray(h.T, np.complex128).ravel()   # column-major
    sc = np.ascontiguousarray(s.T, np.complex128).ravel()
    w = np.zeros(n)
    z = np.zeros(n * n, np.complex128) if vectors else None
    info = lib().orc_zhegv(n, hc, sc, w, z.ctypes.data if vectors else None)
    return info, w, (z.reshape(n, n).T.copy() if vectors else None)


def perturb(dhp, dsp, z, eig, eigcut=10.0):
    """Oracle perturb (perturb.f:87-221): dhp, dsp packed upper (LAPACK 'U'),
    z full ([row, col]) eigenvectors, eig ascending; returns deig[:ntot]."""
    n = len(eig)
    zc = np.ascontiguousarray(np.asarray(z).T, np.complex128).ravel()
    e = np.ascontiguousarray(eig, float)
    d = np.zeros(n)
    nt = C.c_int(0)
    rc = lib().orc_perturb(n, np.ascontiguousarray(dhp, np.complex128), np.ascontiguousarray(dsp, np.complex128),
                           zc, e, float(eigcut), C.byref(nt), d)
    if rc:
        raise RuntimeError("oracle perturb: degenerate group larger than 9")
    return d[:nt.value]


class RefSetup:
    """The same structure pushed through the reference's own routines."""

    def __init__(self, st):
        L = ref()
        p = st.params
        ci = lambda v: C.byref(C.c_int(int(v)))
        lat = np.ascontiguousarray(st.lat.ravel())       # lat(c,v) column-major
        npo = C.c_int(0); dmin = C.c_double(0); ierr = C.c_int(0)
        L.ref_setup(ci(st.natoms), lat, np.ascontiguousarray(st.posn.ravel()),
                    np.ascontiguousarray(st.posq.ravel()), st.kkind,
                    np.array(st.nd, np.int32), ci(p.kinds), np.array(p.kbas, np.int32),
                    np.array(p.valence, float), np.array(p.dnn, float),
                    np.array(p.screenl, float), np.ascontiguousarray(p.param, float),
                    ci(len(p.param)), ci(p.nltype), ci(int(p.jsover)), C.byref(npo),
                    C.byref(dmin), C.byref(ierr))
        if ierr.value:
            raise RuntimeError(f"ref_setup ierr={ierr.value}")
        self.st = st
        self.npairs = npo.value
        self.dismin = dmin.value
        n = self.npairs
        self.tt = np.zeros(3 * n); self.dlv = np.zeros(3 * n); self.dist = np.zeros(n)
        self.screen = np.zeros(n); self.dscrn = np.zeros(n)
        self.jkind = np.zeros(2 * n, np.int32); self.jatm = np.zeros(2 * n, np.int32)
        L.ref_get_pairs(self.tt, self.dlv, self.dist, self.screen, self.dscrn,
                        self.jkind, self.jatm)
        nat = st.natoms
        self.paros = np.zeros(4 * nat); self.dparos = np.zeros(4 * nat)
        self.hop = np.zeros(14 * n); self.over = np.zeros(14 * n)
        self.dhop = np.zeros(14 * n); self.dover = np.zeros(14 * n)
        self.den = np.zeros(2 * nat)
        L.ref_setpar(C.byref(C.c_double(st.volume)), self.paros, self.hop, self.over,
                     self.dparos, self.dhop, self.dover, self.den, C.byref(ierr))
        self.nv = st.nv

    def setpar_spin(self, jspin, param_dn, flipos=None):
        """The reference's setpar/setparv for spin channel jspin of a
        spin-polarised set (param(:,2) = param_dn; atoms flipped where
        flipos is non-zero); replaces this object's per-structure arrays,
        which the later setham calls use."""
        st = self.st
        ierr = C.c_int(0)
        pdn = np.ascontiguousarray(np.asarray(param_dn, float))
        fl = np.zeros(st.natoms, np.int32) if flipos is None else np.ascontiguousarray(flipos, np.int32)
        ref().ref_setpar_spin(C.byref(C.c_int(int(jspin))), C.byref(C.c_double(st.volume)), pdn,
                              C.byref(C.c_int(len(pdn))), fl, self.paros, self.hop, self.over,
                              self.dparos, self.dhop, self.dover, self.den, C.byref(ierr))
        if ierr.value:
            raise RuntimeError(f"ref_setpar_spin ierr={ierr.value}")

    def setpar_noncol(self, param_dn):
        """The reference's setpar/setparv with lnoncol = .true.: paros,
        spinos, dparos, dspnos, den [2][natoms][mkind] and channel 1's
        integrals."""
        st = self.st
        n, nat = self.npairs, st.natoms
        out = dict(paros=np.zeros(4 * nat), spinos=np.zeros(4 * nat), dparos=np.zeros(4 * nat),
                   dspnos=np.zeros(4 * nat), hop=np.zeros(14 * max(n, 1)), over=np.zeros(14 * max(n, 1)),
                   dhop=np.zeros(14 * max(n, 1)), dover=np.zeros(14 * max(n, 1)), den=np.zeros(2 * 20 * 2))
        ierr = C.c_int(0)
        pdn = np.ascontiguousarray(np.asarray(param_dn, float))
        ref().ref_setpar_nc(C.byref(C.c_double(st.volume)), pdn, C.byref(C.c_int(len(pdn))), out["paros"],
                            out["spinos"], out["dparos"], out["dspnos"], out["hop"], out["over"], out["dhop"],
                            out["dover"], out["den"], C.byref(ierr))
        if ierr.value:
            raise RuntimeError(f"ref_setpar_nc ierr={ierr.value}")
        return out

    def hs_noncol(self, ak, spinat, spin_onsite, deriv=False):
        """The reference's spinor H(k), S(k) (ideriv 0) or their volume
        derivatives (ideriv 1): spinham + setham/sethamv + addspins, as
        bande.f:897-962 and 1024-1034 build them. spin_onsite is
        setpar_noncol's spinos (dspnos for deriv), which must run first;
        returns the packed upper triangles of order 2 nv."""
        n2 = 2 * self.nv
        hp = np.zeros(n2 * (n2 + 1) // 2, np.complex128); sp = np.zeros_like(hp)
        ierr = C.c_int(0)
        ref().ref_setham_nc(np.asarray(ak, float), C.byref(C.c_int(int(deriv))),
                            np.ascontiguousarray(np.asarray(spinat, float).ravel()),
                            np.ascontiguousarray(np.asarray(spin_onsite, float)), hp, sp, C.byref(ierr))
        if ierr.value:
            raise RuntimeError(f"ref_setham_nc ierr={ierr.value}")
        return hp, sp

    def bande_noncol(self, kpts, weights, param_dn, spinat, mode=4, tkb=0.005, eigcut=10.0):
        """The reference's bande with lnoncol (jspins = 2): param(:,2) =
        param_dn, spinat [natoms, 3] unit spin directions."""
        st = self.st
        ef, e1, e2 = C.c_double(0), C.c_double(0), C.c_double(0)
        ib = C.c_int(0)
        kp = np.ascontiguousarray(np.asarray(kpts, float).ravel())
        pdn = np.ascontiguousarray(np.asarray(param_dn, float))
        sp = np.ascontiguousarray(np.asarray(spinat, float).ravel())
        ref().ref_bande_nc(C.byref(C.c_int(mode)), C.byref(C.c_int(len(kpts))), kp,
                           np.ascontiguousarray(np.asarray(weights, float)),
                           C.byref(C.c_double(st.volume)), C.byref(C.c_double(st.zfill)),
                           C.byref(C.c_double(tkb)), C.byref(C.c_double(eigcut)), pdn,
                           C.byref(C.c_int(len(pdn))), sp, C.byref(ef), C.byref(e1), C.byref(e2), C.byref(ib))
        return dict(err=ib.value, efermi=ef.value, earg1=e1.value, earg2=e2.value)

    def bande_spin(self, kpts, weights, param_dn, flipos=None, mode=4, tkb=0.005, eigcut=10.0):
        """The reference's bande with jspins = 2 (param(:,1) = the
        structure's parameters, param(:,2) = param_dn)."""
        st = self.st
        ef, e1, e2, sm = C.c_double(0), C.c_double(0), C.c_double(0), C.c_double(0)
        ib = C.c_int(0)
        kp = np.ascontiguousarray(np.asarray(kpts, float).ravel())
        pdn = np.ascontiguousarray(np.asarray(param_dn, float))
        fl = np.zeros(st.natoms, np.int32) if flipos is None else np.ascontiguousarray(flipos, np.int32)
        ref().ref_bande_spin(C.byref(C.c_int(mode)), C.byref(C.c_int(len(kpts))), kp,
                             np.ascontiguousarray(np.asarray(weights, float)),
                             C.byref(C.c_double(st.volume)), C.byref(C.c_double(st.zfill)),
                             C.byref(C.c_double(tkb)), C.byref(C.c_double(eigcut)), pdn,
                             C.byref(C.c_int(len(pdn))), fl, C.byref(ef), C.byref(e1), C.byref(e2),
                             C.byref(sm), C.byref(ib))
        return dict(err=ib.value, efermi=ef.value, earg1=e1.value, earg2=e2.value, smag=sm.value)

    def hs(self, ak, deriv=False):
        np_ = self.nv * (self.nv + 1) // 2
        hp = np.zeros(np_, np.complex128); sp = np.zeros(np_, np.complex128)
        ref().ref_setham(np.asarray(ak, float), C.byref(C.c_int(int(deriv))), hp, sp)
        return hp, sp

    def diag(self, hp, sp, mode=4):
        n = self.nv
        e = np.zeros(n); z = np.zeros(n * n, np.complex128); io = C.c_int(0)
        ref().ref_diag(C.byref(C.c_int(mode)), np.ascontiguousarray(hp),
                       np.ascontiguousarray(sp), e, z, C.byref(io))
        return io.value, e, z.reshape(n, n).T.copy()

    def perturb(self, dhp, dsp, eig, eigcut=10.0):
        d = np.zeros(self.nv); nt = C.c_int(0)
        ref().ref_perturb(np.ascontiguousarray(dhp), np.ascontiguousarray(dsp),
                          np.ascontiguousarray(eig), C.byref(C.c_double(eigcut)),
                          C.byref(nt), d)
        return nt.value, d

    def bande(self, kpts, weights, mode=4, tkb=0.005, eigcut=10.0):
        ef, e1, e2 = C.c_double(0), C.c_double(0), C.c_double(0)
        ib = C.c_int(0)
        kp = np.ascontiguousarray(np.asarray(kpts, float).ravel())
        ref().ref_bande(C.byref(C.c_int(mode)), C.byref(C.c_int(len(kpts))), kp,
                        np.ascontiguousarray(np.asarray(weights, float)),
                        C.byref(C.c_double(self.st.volume)),
                        C.byref(C.c_double(self.st.zfill)), C.byref(C.c_double(tkb)),
                        C.byref(C.c_double(eigcut)), C.byref(ef), C.byref(e1),
                        C.byref(e2), C.byref(ib))
        return dict(err=ib.value, efermi=ef.value, earg1=e1.value, earg2=e2.value)


# dimensions of the ex1 reference build (NRLtb/ex1/P1)
REF_MATOM, REF_MKIND = 20, 2
REF_NPARD = REF_MKIND * (29 + 68 * REF_MKIND)


def ref_input(skin_line_file: str) -> dict:
    """The reference's own parameter-file reader, input(nparam, jspins),
    on a unit-10 file whose line is the parameter file's name."""
    L = ref()
    fn = os.fsencode(skin_line_file)
    ci = lambda: C.c_int(0)
    npo, jsp, jso, nlt, kin, ierr = ci(), ci(), ci(), ci(), ci(), ci()
    kb = np.zeros(REF_MKIND, np.int32)
    val = np.zeros(REF_MKIND)
    mint = REF_MKIND * (REF_MKIND + 1) // 2
    dnn, scr = np.zeros(mint), np.zeros(mint)
    par = np.zeros(2 * REF_NPARD)
    L.ref_input(fn, C.byref(C.c_int(len(fn))), C.byref(npo), C.byref(jsp), C.byref(jso), C.byref(nlt),
                C.byref(kin), kb, val, dnn, scr, par, C.byref(ierr))
    k = kin.value
    npard = k * (29 + 68 * k)
    # param(npard_dim, mspin) column-major: channel j at [j * REF_NPARD, ...]
    chans = np.stack([par[j * REF_NPARD:j * REF_NPARD + npard] for j in range(jsp.value)])
    return dict(nparam=npo.value, jspins=jsp.value, jsover=jso.value, nltype=nlt.value, kinds=k,
                kbas=kb[:k].copy(), valence=val[:k].copy(), dnn=dnn[:k * (k + 1) // 2].copy(),
                screenl=scr[:k * (k + 1) // 2].copy(), param=chans)


def ref_input1(block_file: str, kbas, valence, zcharge: float = 0.0) -> dict:
    """The reference's own structure reader (setlat + input1, then the main
    program's conversion to Cartesian positions) on a unit-10 file holding a
    SKIN structure block from its lattice line on."""
    L = ref()
    fn = os.fsencode(block_file)
    nk = len(kbas)
    plv, vol, nat = np.zeros(9), C.c_double(0), C.c_int(0)
    nd = np.zeros(3, np.int32)
    kk = np.zeros(REF_MATOM, np.int32); fl = np.zeros(REF_MATOM, np.int32)
    posn, posq, sp = np.zeros(3 * REF_MATOM), np.zeros(3 * REF_MATOM), np.zeros(3 * REF_MATOM)
    zf, lnc, inp = C.c_double(0), C.c_int(0), C.c_int(0)
    L.ref_input1(fn, C.byref(C.c_int(len(fn))), C.byref(C.c_int(nk)), np.asarray(kbas, np.int32),
                 np.asarray(valence, float), C.byref(C.c_double(zcharge)), plv, C.byref(vol), C.byref(nat), nd,
                 kk, posn, posq, C.byref(zf), fl, C.byref(lnc), sp, C.byref(inp))
    n = nat.value
    return dict(inpstat=inp.value, lat=plv.reshape(3, 3).copy(), volume=vol.value, natoms=n, nd=nd.copy(),
                kkind=kk[:n].copy(), posn=posn[:3 * n].reshape(n, 3).copy(),
                posq=posq[:3 * n].reshape(n, 3).copy(), zfill=zf.value, flipos=fl[:n].copy(),
                lnoncol=lnc.value, spinat=sp[:3 * n].reshape(n, 3).copy())
