"""Host-side mirror of the reference's per-structure interface.

``bande(structure, kpts, weights, mode, tkb, eigcut)`` plays the role of
``bande(mode,...,efermi,earg1,earg2,...,iberr)`` (NRLtb/src/bande.f:1-4):
it returns E_F, earg1 and earg2 with the meaning of bande.f:99-123 and
raises :class:`TbkError` where the reference sets iberr / ioverr or stops.
Everything below runs in lib/libtbk.so (HIP kernels on the GPU); this
module only marshals arrays.
"""
from __future__ import annotations

import ctypes as C
from typing import Dict, Optional

import numpy as np

from . import _lib
from ._lib import TBK_FP64, TBK_MIXED, TbkError, check, lib
from .structures import Structure


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


class HostSetup:
    """Neighbour list and SK parameters from the library's host set-up
    (search2.f, setpar.f, setparv.f restated bit-exactly).  With param_dn,
    spin channel jspin of a spin-polarised set (st.params the spin-up
    channel; flipos the flipped atoms)."""

    def __init__(self, st: Structure, maxpairs: Optional[int] = None, param_dn=None, jspin: int = 1,
                 flipos=None):
        L = lib()
        p = st.params
        nat = st.natoms
        maxpairs = maxpairs or max(4096, 4000 * nat)
        self.tt = np.zeros(3 * maxpairs); self.dlv = np.zeros(3 * maxpairs)
        self.dist = np.zeros(maxpairs); self.screen = np.zeros(maxpairs)
        self.dscrn = np.zeros(maxpairs)
        self.jkind = np.zeros(2 * maxpairs, np.int32); self.jatm = np.zeros(2 * maxpairs, np.int32)
        npairs = C.c_int(0); dmin = C.c_double(0.0)
        rc = L.tbk_search2(nat, np.ascontiguousarray(st.lat.ravel()),
                           np.ascontiguousarray(st.posn.ravel()), st.kkind, p.kinds,
                           np.array(st.nd, np.int32), np.array(p.dnn, float),
                           np.array(p.screenl, float), maxpairs, C.byref(npairs), self.tt,
                           self.dlv, self.dist, self.screen, self.dscrn, self.jkind, self.jatm,
                           C.byref(dmin))
        check(rc, "tbk_search2")
        n = self.npairs = npairs.value
        self.dismin = dmin.value
        for nm in ("tt", "dlv"):
            setattr(self, nm, getattr(self, nm)[:3 * n].copy())
        for nm in ("dist", "screen", "dscrn"):
            setattr(self, nm, getattr(self, nm)[:n].copy())
        self.jkind = self.jkind[:2 * n].copy(); self.jatm = self.jatm[:2 * n].copy()
        self.den = np.zeros(nat * p.kinds); self.paros = np.zeros(4 * nat)
        self.hop = np.zeros(14 * max(n, 1)); self.over = np.zeros(14 * max(n, 1))
        param = np.ascontiguousarray(p.param, float)
        self.dparos = np.zeros(4 * nat)
        self.dhop = np.zeros(14 * max(n, 1)); self.dover = np.zeros(14 * max(n, 1))
        if param_dn is None:
            check(L.tbk_setpar(nat, st.kkind, p.kinds, param, p.nltype, int(p.jsover), n, self.dist,
                               self.screen, self.jkind, self.jatm, self.den, self.paros, self.hop,
                               self.over), "tbk_setpar")
            check(L.tbk_setparv(nat, st.kkind, p.kinds, param, p.nltype, int(p.jsover), st.volume, n,
                                self.dist, self.screen, self.dscrn, self.jkind, self.jatm, self.den,
                                self.dparos, self.dhop, self.dover), "tbk_setparv")
        else:
            p2 = np.ascontiguousarray(np.concatenate([param, np.asarray(param_dn, float)]))
            if len(p2) != 2 * len(param):
                raise ValueError("param_dn must have as many values as the spin-up channel")
            fl = None
            if flipos is not None:
                self._flip = np.ascontiguousarray(np.asarray(flipos, np.int32))
                fl = self._flip.ctypes.data
            check(L.tbk_setpar_spin(nat, st.kkind, p.kinds, p2, len(param), int(jspin), fl, p.nltype,
                                    int(p.jsover), n, self.dist, self.screen, self.jkind, self.jatm,
                                    self.den, self.paros, self.hop, self.over), "tbk_setpar_spin")
            check(L.tbk_setparv_spin(nat, st.kkind, p.kinds, p2, len(param), int(jspin), fl, p.nltype,
                                     int(p.jsover), st.volume, n, self.dist, self.screen, self.dscrn,
                                     self.jkind, self.jatm, self.den, self.dparos, self.dhop, self.dover),
                  "tbk_setparv_spin")


def _structure_desc(st: Structure, keep: list) -> "_lib.StructureDesc":
    """Fill the C descriptor of a structure; the arrays it points to are
    appended to `keep` (they must outlive the create call)."""
    p = st.params
    k = dict(
        lat=np.ascontiguousarray(st.lat.ravel(), float),
        posn=np.ascontiguousarray(st.posn.ravel(), float),
        posq=np.ascontiguousarray(st.posq.ravel(), float),
        kkind=np.ascontiguousarray(st.kkind, np.int32),
        kbas=np.array(p.kbas, np.int32), valence=np.array(p.valence, float),
        dnn=np.array(p.dnn, float), screenl=np.array(p.screenl, float),
        param=np.ascontiguousarray(p.param, float))
    keep.append(k)
    desc = _lib.StructureDesc()
    desc.natoms = st.natoms
    desc.lat, desc.posn, desc.posq = _dp(k["lat"]), _dp(k["posn"]), _dp(k["posq"])
    desc.kkind = _ip(k["kkind"])
    desc.nd = (C.c_int * 3)(*[int(x) for x in st.nd])
    desc.kinds = p.kinds
    desc.kbas, desc.valence = _ip(k["kbas"]), _dp(k["valence"])
    desc.dnn, desc.screenl = _dp(k["dnn"]), _dp(k["screenl"])
    desc.param, desc.nparam = _dp(k["param"]), len(k["param"])
    desc.nltype, desc.jsover = p.nltype, int(p.jsover)
    desc.thisvol, desc.zfill = st.volume, st.zfill
    return desc


class Plan:
    """A structure, its k-mesh and device-resident H/S/Z state."""

    def __init__(self, st: Structure, kpts, weights, batch: int = 0,
                 precision: int = TBK_FP64, device: int = 0, timing: bool = False,
                 graph: bool = False):
        L = lib()
        self.st = st
        self._keep = []
        desc = _structure_desc(st, self._keep)
        self.kpts = np.ascontiguousarray(np.asarray(kpts, float).reshape(-1, 3))
        self.weights = np.ascontiguousarray(np.asarray(weights, float))
        self.nk = len(self.weights)
        h = C.c_void_p()
        check(L.tbk_plan_create(device, C.byref(desc), self.nk, self.kpts.ravel(), self.weights,
                                int(batch), int(precision), C.byref(h)), "tbk_plan_create")
        self.h = h
        self.nv = L.tbk_plan_nv(h)
        self.npairs = L.tbk_plan_npairs(h)
        if timing:
            L.tbk_plan_set_timing(h, 1)
        if graph:
            L.tbk_plan_set_graph(h, 1)

    def set_graph(self, on: bool = True):
        check(lib().tbk_plan_set_graph(self.h, int(on)), "tbk_plan_set_graph")

    def set_tridiag(self, stages: int):
        """1: one-stage Householder panels; 2: two-stage (band + bulge
        chasing); 0: the default (two-stage from order 512)."""
        check(lib().tbk_plan_set_tridiag(self.h, int(stages)), "tbk_plan_set_tridiag")

    def set_window(self, on: bool = True):
        """Occupation window (default on): pressure runs with every k in one
        batch form eigenvectors and derivatives only below E_F + 20 kT."""
        check(lib().tbk_plan_set_window(self.h, int(on)), "tbk_plan_set_window")

    def set_fused_build(self, on: bool = True):
        """Pressure modes: H, S, dH/dV, dS/dV in one build pass (on), or
        dH/dV, dS/dV after the eigensolve with no extra memory (default)."""
        check(lib().tbk_plan_set_fused_build(self.h, int(on)), "tbk_plan_set_fused_build")

    def device_bytes(self) -> int:
        return int(lib().tbk_plan_device_bytes(self.h))

    def run(self, mode: int = 4, tkb: float = 0.005, eigcut: float = 10.0) -> Dict[str, float]:
        ef, e1, e2 = C.c_double(), C.c_double(), C.c_double()
        check(lib().tbk_plan_run(self.h, mode, tkb, eigcut, C.byref(ef), C.byref(e1),
                                 C.byref(e2)), "tbk_plan_run")
        return dict(efermi=ef.value, earg1=e1.value, earg2=e2.value)

    def launch(self, mode: int = 4, tkb: float = 0.005, eigcut: float = 10.0):
        check(lib().tbk_plan_launch(self.h, mode, tkb, eigcut), "tbk_plan_launch")

    def wait(self) -> Dict[str, float]:
        ef, e1, e2 = C.c_double(), C.c_double(), C.c_double()
        check(lib().tbk_plan_wait(self.h, C.byref(ef), C.byref(e1), C.byref(e2)), "tbk_plan_wait")
        return dict(efermi=ef.value, earg1=e1.value, earg2=e2.value)

    def eigenvalues(self):
        e = np.zeros((self.nk, self.nv)); d = np.zeros((self.nk, self.nv))
        check(lib().tbk_plan_eigenvalues(self.h, e.ctypes.data, d.ctypes.data))
        return e, d

    def eigenvectors(self, k: int) -> np.ndarray:
        z = np.zeros(self.nv * self.nv, np.complex128)
        check(lib().tbk_plan_eigenvectors(self.h, int(k), z.ctypes.data), "tbk_plan_eigenvectors")
        return z.reshape(self.nv, self.nv).T.copy()          # [row, col]

    def projections(self, k: int, eigcut: float = 10.0):
        """Lowdin projections of item k after a run in mode 6 or 10
        (diag.f:526-758): (levels e <= eigcut, weights [state, orbital] =
        |(S^{1/2} z)_i|^2)."""
        n = self.nv
        e = np.zeros(n); w = np.zeros(n * n); nt = C.c_int()
        check(lib().tbk_plan_projections(self.h, int(k), float(eigcut), C.byref(nt), e.ctypes.data,
                                         w.ctypes.data), "tbk_plan_projections")
        m = nt.value
        return e[:m].copy(), w[:m * n].reshape(m, n).copy()

    def hs(self, k: int, deriv: bool = False):
        n = self.nv
        h = np.zeros(n * n, np.complex128); s = np.zeros(n * n, np.complex128)
        check(lib().tbk_plan_hs(self.h, int(k), int(deriv), h.ctypes.data, s.ctypes.data))
        return h.reshape(n, n).T.copy(), s.reshape(n, n).T.copy()

    def autotune(self, reps: int = 3) -> int:
        """Time every GEMM tile for this plan's shapes and keep the fastest
        (per process).  Results are undefined until the next run."""
        n = C.c_int(0)
        check(lib().tbk_plan_autotune(self.h, int(reps), C.byref(n)), "tbk_plan_autotune")
        return n.value

    def set_timing(self, on: bool = True):
        """Enable per-stage HIP event timing (clears accumulated times)."""
        check(lib().tbk_plan_set_timing(self.h, int(on)))

    def stage_times(self) -> Dict[str, float]:
        names = (C.c_char_p * 32)(); ms = np.zeros(32); n = C.c_int(0)
        check(lib().tbk_plan_stage_times(self.h, 32, names, _dp(ms), C.byref(n)))
        return {names[i].decode(): float(ms[i]) for i in range(n.value)}

    def close(self):
        if getattr(self, "h", None):
            lib().tbk_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class GroupPlan(Plan):
    """Several structures of uneven orbital count solved in the same batched
    launches (padded to the largest order).  run() returns one result dict
    per structure; a structure whose solve failed gets {"error": code}."""

    def __init__(self, structures, kpts_list, weights_list, batch: int = 0,
                 precision: int = TBK_FP64, device: int = 0, graph: bool = False):
        L = lib()
        self.structures = list(structures)
        self._keep = []
        ns = len(self.structures)
        descs = (_lib.StructureDesc * ns)()
        for i, st in enumerate(self.structures):
            descs[i] = _structure_desc(st, self._keep)
        ks = [np.ascontiguousarray(np.asarray(k, float).reshape(-1, 3)) for k in kpts_list]
        ws = [np.ascontiguousarray(np.asarray(w, float)) for w in weights_list]
        self.nks = np.array([len(w) for w in ws], np.int32)
        self.kpts = np.ascontiguousarray(np.concatenate(ks))
        self.weights = np.ascontiguousarray(np.concatenate(ws))
        self.nk = len(self.weights)
        h = C.c_void_p()
        check(L.tbk_plan_create_group(device, ns, descs, self.nks, self.kpts.ravel(), self.weights,
                                      int(batch), int(precision), C.byref(h)), "tbk_plan_create_group")
        self.h = h
        self.nstruct = ns
        self.order = L.tbk_plan_order(h)
        self.nvs = [st.nv for st in self.structures]
        if graph:
            L.tbk_plan_set_graph(h, 1)

    def run(self, mode: int = 4, tkb: float = 0.005, eigcut: float = 10.0):
        self.launch(mode, tkb, eigcut)
        return self.wait()

    def wait(self):
        L = lib()
        ef, e1, e2 = C.c_double(), C.c_double(), C.c_double()
        L.tbk_plan_wait(self.h, C.byref(ef), C.byref(e1), C.byref(e2))   # per-structure below
        out = []
        for s in range(self.nstruct):
            rc = L.tbk_plan_result(self.h, s, C.byref(ef), C.byref(e1), C.byref(e2))
            out.append({"error": rc} if rc else dict(efermi=ef.value, earg1=e1.value, earg2=e2.value))
        return out

    def eigenvalues(self, s: int = 0):
        nk, nv = int(self.nks[s]), self.nvs[s]
        e = np.zeros((nk, nv)); d = np.zeros((nk, nv))
        check(lib().tbk_plan_structure_eigenvalues(self.h, int(s), e.ctypes.data, d.ctypes.data))
        return e, d


class SpinPlan(Plan):
    """A collinear spin-polarised structure (bande with jspins = 2,
    bande.f:816-889, 1069-1081): st.params holds the spin-up channel,
    param_dn the spin-down parameters in the same order (an "NM"/"OM"
    parameter file's second block, input.f:327-359), flipos the atoms whose
    densities and on-site terms come from the other channel
    (setpar.f:500-536).  One E_F over both channels with one electron per
    state; run() and wait() add "smag", spin-up minus spin-down electrons
    (bande.f:1388-1402)."""

    def __init__(self, st: Structure, param_dn, kpts, weights, flipos=None, batch: int = 0,
                 precision: int = TBK_FP64, device: int = 0, graph: bool = False):
        L = lib()
        self.st = st
        self._keep = []
        desc = _structure_desc(st, self._keep)
        self.param_dn = np.ascontiguousarray(np.asarray(param_dn, float))
        if len(self.param_dn) != len(st.params.param):
            raise ValueError("param_dn must have as many values as the spin-up channel")
        self.flipos = None
        if flipos is not None:
            self.flipos = np.ascontiguousarray(np.asarray(flipos, np.int32))
            if len(self.flipos) != st.natoms:
                raise ValueError("flipos needs one entry per atom")
        self.kpts = np.ascontiguousarray(np.asarray(kpts, float).reshape(-1, 3))
        self.weights = np.ascontiguousarray(np.asarray(weights, float))
        self.nk = len(self.weights)
        h = C.c_void_p()
        fl = None if self.flipos is None else self.flipos.ctypes.data
        check(L.tbk_plan_create_spin(device, C.byref(desc), self.param_dn, fl, self.nk, self.kpts.ravel(),
                                     self.weights, int(batch), int(precision), C.byref(h)),
              "tbk_plan_create_spin")
        self.h = h
        self.nv = L.tbk_plan_nv(h)
        self.npairs = L.tbk_plan_npairs(h)
        if graph:
            L.tbk_plan_set_graph(h, 1)

    def magnetization(self) -> float:
        m = C.c_double()
        check(lib().tbk_plan_magnetization(self.h, C.byref(m)), "tbk_plan_magnetization")
        return m.value

    def run(self, mode: int = 4, tkb: float = 0.005, eigcut: float = 10.0) -> Dict[str, float]:
        r = super().run(mode, tkb, eigcut)
        r["smag"] = self.magnetization()
        return r

    def wait(self) -> Dict[str, float]:
        r = super().wait()
        r["smag"] = self.magnetization()
        return r

    def eigenvalues(self, spin: int = 0):
        """Eigenvalues and derivatives of channel spin (0 up, 1 down)."""
        e = np.zeros((self.nk, self.nv)); d = np.zeros((self.nk, self.nv))
        check(lib().tbk_plan_structure_eigenvalues(self.h, int(spin), e.ctypes.data, d.ctypes.data))
        return e, d


class NonCollinearPlan(Plan):
    """A non-collinear structure (bande with jspins = 2, lnoncol true):
    st.params is channel 1, param_dn channel 2, spinat [natoms, 3] each
    atom's unit spin direction.  Items are the 2 nv spinor matrices
    (addspins.f), one electron per state; eigenvalues() has 2 nv columns."""

    def __init__(self, st: Structure, param_dn, spinat, kpts, weights, batch: int = 0,
                 precision: int = TBK_FP64, device: int = 0, graph: bool = False):
        L = lib()
        self.st = st
        self._keep = []
        desc = _structure_desc(st, self._keep)
        self.param_dn = np.ascontiguousarray(np.asarray(param_dn, float))
        if len(self.param_dn) != len(st.params.param):
            raise ValueError("param_dn must have as many values as channel 1")
        self.spinat = np.ascontiguousarray(np.asarray(spinat, float).reshape(-1))
        if len(self.spinat) != 3 * st.natoms:
            raise ValueError("spinat needs three components per atom")
        self.kpts = np.ascontiguousarray(np.asarray(kpts, float).reshape(-1, 3))
        self.weights = np.ascontiguousarray(np.asarray(weights, float))
        self.nk = len(self.weights)
        h = C.c_void_p()
        check(L.tbk_plan_create_noncol(device, C.byref(desc), self.param_dn, self.spinat, self.nk,
                                       self.kpts.ravel(), self.weights, int(batch), int(precision),
                                       C.byref(h)), "tbk_plan_create_noncol")
        self.h = h
        self.nv = L.tbk_plan_nv(h)
        self.npairs = L.tbk_plan_npairs(h)
        if graph:
            L.tbk_plan_set_graph(h, 1)

    def set_spin_directions(self, on: bool = True):
        """Form each atom's spin vector in the following runs (the lspdir
        output of bande.f:1484-1561)."""
        check(lib().tbk_plan_set_spin_directions(self.h, 1 if on else 0), "tbk_plan_set_spin_directions")

    def spin_directions(self) -> np.ndarray:
        """[natoms, 3]: each atom's spin vector after a run with
        set_spin_directions(True)."""
        out = np.zeros(3 * self.st.natoms)
        check(lib().tbk_plan_spin_directions(self.h, 0, out.ctypes.data_as(C.POINTER(C.c_double))),
              "tbk_plan_spin_directions")
        return out.reshape(-1, 3)


class MultiPlan:
    """Many structures of uneven orbital count, grouped by padded order
    (tbk_multi_*): each group is one batched plan, groups run concurrently."""

    def __init__(self, structures, kpts_list, weights_list, precision: int = TBK_FP64,
                 granularity: int = 32, max_batch_bytes: int = 16 << 30, device: int = 0,
                 graph: bool = False):
        L = lib()
        self.structures = list(structures)
        self._keep = []
        ns = len(self.structures)
        descs = (_lib.StructureDesc * ns)()
        for i, st in enumerate(self.structures):
            descs[i] = _structure_desc(st, self._keep)
        ks = [np.ascontiguousarray(np.asarray(k, float).reshape(-1, 3)) for k in kpts_list]
        ws = [np.ascontiguousarray(np.asarray(w, float)) for w in weights_list]
        self.nks = np.array([len(w) for w in ws], np.int32)
        self.kpts = np.ascontiguousarray(np.concatenate(ks))
        self.weights = np.ascontiguousarray(np.concatenate(ws))
        h = C.c_void_p()
        check(L.tbk_multi_create(device, ns, descs, self.nks, self.kpts.ravel(), self.weights,
                                 int(precision), int(granularity), int(max_batch_bytes), C.byref(h)),
              "tbk_multi_create")
        self.h = h
        self.nstruct = ns
        self.ngroups = L.tbk_multi_ngroups(h)
        if graph:
            L.tbk_multi_set_graph(h, 1)

    def autotune(self, reps: int = 2):
        check(lib().tbk_multi_autotune(self.h, int(reps)), "tbk_multi_autotune")

    def set_graph(self, on: bool = True):
        check(lib().tbk_multi_set_graph(self.h, int(on)), "tbk_multi_set_graph")

    def launch(self, mode: int = 4, tkb: float = 0.005, eigcut: float = 10.0):
        check(lib().tbk_multi_launch(self.h, mode, tkb, eigcut), "tbk_multi_launch")

    def wait(self):
        L = lib()
        rc = L.tbk_multi_wait(self.h)
        if rc in (_lib.TBK_ERR_HIP, _lib.TBK_ERR_ARG):
            check(rc, "tbk_multi_wait")
        out = []
        ef, e1, e2 = C.c_double(), C.c_double(), C.c_double()
        for s in range(self.nstruct):
            r = L.tbk_multi_result(self.h, s, C.byref(ef), C.byref(e1), C.byref(e2))
            out.append({"error": r} if r else dict(efermi=ef.value, earg1=e1.value, earg2=e2.value))
        return out

    def run(self, mode: int = 4, tkb: float = 0.005, eigcut: float = 10.0):
        self.launch(mode, tkb, eigcut)
        return self.wait()

    def eigenvalues(self, s: int):
        nk, nv = int(self.nks[s]), self.structures[s].nv
        e = np.zeros((nk, nv)); d = np.zeros((nk, nv))
        check(lib().tbk_multi_eigenvalues(self.h, int(s), e.ctypes.data, d.ctypes.data))
        return e, d

    def device_bytes(self) -> int:
        return int(lib().tbk_multi_device_bytes(self.h))

    def close(self):
        if getattr(self, "h", None):
            lib().tbk_multi_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Request loop with continuous batching (tbk_server_*): submit() queues
    a structure and returns at once; result() waits for it."""

    def __init__(self, precision: int = TBK_MIXED, budget_bytes: int = 64 << 30,
                 max_inflight: int = 8, granularity: int = 32, device: int = 0):
        h = C.c_void_p()
        check(lib().tbk_server_create(device, int(precision), int(budget_bytes), int(max_inflight),
                                      int(granularity), C.byref(h)), "tbk_server_create")
        self.h = h

    def submit(self, st: Structure, kpts, weights, mode: int = 4, tkb: float = 0.005,
               eigcut: float = 10.0) -> int:
        keep = []
        desc = _structure_desc(st, keep)
        k = np.ascontiguousarray(np.asarray(kpts, float).reshape(-1, 3))
        w = np.ascontiguousarray(np.asarray(weights, float))
        rid = C.c_longlong()
        check(lib().tbk_server_submit(self.h, C.byref(desc), len(w), k.ravel(), w, int(mode), tkb,
                                      eigcut, C.byref(rid)), "tbk_server_submit")
        return rid.value

    def result(self, rid: int, timeout_ms: int = -1):
        st = C.c_int(); ef, e1, e2, lat, q = (C.c_double() for _ in range(5))
        rc = lib().tbk_server_result(self.h, int(rid), int(timeout_ms), C.byref(st), C.byref(ef),
                                     C.byref(e1), C.byref(e2), C.byref(lat), C.byref(q))
        if rc == 1:
            return None
        check(rc, "tbk_server_result")
        return dict(status=st.value, efermi=ef.value, earg1=e1.value, earg2=e2.value,
                    latency_ms=lat.value, queue_ms=q.value)

    def release(self, rid: int):
        check(lib().tbk_server_release(self.h, int(rid)), "tbk_server_release")

    def stats(self):
        a = (C.c_longlong * 8)()
        check(lib().tbk_server_stats(self.h, a), "tbk_server_stats")
        keys = ("completed", "kpoints", "groups", "max_group", "pending", "running", "inflight_bytes",
                "last_error")
        return dict(zip(keys, [int(x) for x in a]))

    def timing(self):
        """Device busy ms (union of the groups' device intervals), work done
        (sum of nv^3 nk), ms from the first submission to the last
        completion, summed host set-up and plan creation + launch ms, and
        the set-up threads."""
        a = (C.c_double * 6)()
        check(lib().tbk_server_timing(self.h, a), "tbk_server_timing")
        keys = ("busy_ms", "work_n3k", "span_ms", "setup_ms", "launch_ms", "setup_threads")
        return dict(zip(keys, [float(x) for x in a]))

    def close(self):
        if getattr(self, "h", None):
            lib().tbk_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def bande(st: Structure, kpts, weights, mode: int = 4, tkb: float = 0.005,
          eigcut: float = 10.0, batch: int = 0, precision: int = TBK_FP64,
          device: int = 0) -> Dict[str, float]:
    """One structure through the whole device pipeline (bande.f:1)."""
    plan = Plan(st, kpts, weights, batch=batch, precision=precision, device=device)
    try:
        return plan.run(mode, tkb, eigcut)
    finally:
        plan.close()


def zhegv(h: np.ndarray, s: np.ndarray, vectors: bool = True, device: int = 0):
    """Device generalized eigensolve of full Hermitian h, s ([row, col])."""
    n = h.shape[0]
    hc = np.ascontiguousarray(h.T, np.complex128)
    sc = np.ascontiguousarray(s.T, np.complex128)
    w = np.zeros(n)
    z = np.zeros(n * n, np.complex128) if vectors else None
    check(lib().tbk_zhegv(device, n, hc.ctypes.data, sc.ctypes.data, w.ctypes.data,
                          z.ctypes.data if vectors else None), "tbk_zhegv")
    return w, (z.reshape(n, n).T.copy() if vectors else None)


def perturb(dh: np.ndarray, ds: np.ndarray, z: np.ndarray, eig, eigcut: float = 10.0, device: int = 0):
    """Device first-order derivatives of one k's eigenvalues (perturb.f:1):
    dh, ds, z full ([row, col]), eig ascending; returns deig[:ntot]."""
    n = dh.shape[0]
    a = [np.ascontiguousarray(m.T, np.complex128) for m in (dh, ds, z)]
    e = np.ascontiguousarray(eig, float)
    d = np.zeros(n)
    nt = C.c_int()
    check(lib().tbk_perturb(device, n, a[0].ctypes.data, a[1].ctypes.data, a[2].ctypes.data, e.ctypes.data,
                            float(eigcut), C.byref(nt), d.ctypes.data), "tbk_perturb")
    return d[:nt.value]


def fermi(eig: np.ndarray, weights, zfill: float, tkb: float = 0.005, el: float = 2.0,
          device: int = 0) -> float:
    """Device Fermi level (fermi.f:1) of eigenvalues eig[k, i]."""
    eig = np.ascontiguousarray(eig, float)
    nk, nv = eig.shape
    ef = C.c_double()
    check(lib().tbk_fermi(device, nv, nk, nv, el, np.ascontiguousarray(weights, float),
                          eig.ravel(), tkb, zfill, C.byref(ef)), "tbk_fermi")
    return ef.value
