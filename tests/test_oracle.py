"""Pins the CPU oracle (oracle/nrl_oracle.c) before it is trusted.

* the reference's own published ex1 results (NRLtb/ex1/SKOUT, 18 structures);
* the reference's own routines built from its sources (oracle/_ref), where
  the reference tree was present at build time;
* LAPACK through scipy for the eigensolver;
* closed forms and invariants the mathematics fixes (s-band dispersion,
  H(k+G) = H(k), rotation invariance, N(E_F) = Z, P = -dE/dV).
"""
import json
import math
import os

import numpy as np
import pytest
import scipy.linalg as sl

import oracle
from tbk import structures as S

GOLD = os.path.join(os.path.dirname(__file__), "golden", "ex1.json")


def ex1_structure(gold, s):
    p = S.NRLParams(1, [9], [gold["valence"]], [gold["rcut"]], [gold["screenl"]],
                    np.array(gold["param"]), 0, True)
    lat = S.fcc_primitive(s["a"]) if s["lattice"] == "fcc" else S.bcc_primitive(s["a"])
    return S.Structure(lat, np.zeros((1, 3)), [1], p, gold["nd"])


def ex1_kpoints(gold, s):
    kl = np.array(gold["kpoints"]["fcc.08" if s["lattice"] == "fcc" else "bcc.08"])
    return kl[:, :3], kl[:, 3] / kl[:, 3].sum()


def test_ex1_published_results():
    """E_F, E(T), P(T), pair counts and minimum distances of ex1/SKOUT."""
    gold = json.load(open(GOLD))
    for s in gold["structures"]:
        st = ex1_structure(gold, s)
        o = oracle.Setup(st)
        assert o.npairs == s["pairs"]
        assert o.dismin == pytest.approx(s["dismin"], abs=1e-13)
        assert st.volume == pytest.approx(s["volume"], rel=1e-15)
        k, w = ex1_kpoints(gold, s)
        r = o.bande(k, w, mode=gold["mode"], tkb=gold["tkb"], eigcut=gold["eigcut"])
        assert r["err"] == 0
        assert abs(r["efermi"] - s["efermi"]) < 1e-12
        assert abs(r["earg1"] - s["et"]) < 1e-12
        assert abs(r["earg2"] - s["pt"]) * S.RYBOHR3_KBAR < 1e-6


needs_ref = pytest.mark.skipif(not oracle.ref_available(),
                               reason="reference build (oracle/_ref) not present")


@needs_ref
def test_setup_bit_exact_vs_reference():
    st = S.config_c1()
    o, R = oracle.Setup(st), oracle.RefSetup(st)
    n = R.npairs
    assert o.npairs == n and o.dismin == R.dismin
    for nm in ("tt", "dlv", "dist", "screen", "dscrn"):
        a, b = getattr(o, nm), getattr(R, nm)
        assert np.array_equal(a[:len(b)], b), nm
    assert np.array_equal(o.jatm[:2 * n], R.jatm) and np.array_equal(o.jkind[:2 * n], R.jkind)
    for nm in ("paros", "hop", "over", "dparos", "dhop", "dover"):
        a, b = getattr(o, nm), getattr(R, nm)
        assert np.array_equal(a[:len(b)], b), nm


@needs_ref
def test_two_kind_setup_vs_reference():
    """Unlike-kind pairs exercise the parity rearrangement (setham.f:326-376)."""
    base = S.synthetic_params(3)
    kinds = 2
    npar = S.npard(kinds)
    rng = np.random.default_rng(5)
    par = np.zeros(npar)
    # kind-1 and kind-2 on-site and like-pair blocks from the one-kind set,
    # unlike pairs from a scaled copy with 14 integrals
    par[0:17] = base.param[0:17]
    par[17:34] = base.param[0:17] * (1 + 0.02 * rng.standard_normal(17))
    head = kinds * (1 + 4 * (3 * kinds + 1))
    par[head:head + 40] = base.param[17:57]
    par[head + 40:head + 80] = base.param[17:57]
    par[head + 80:head + 80 + 56] = np.concatenate([base.param[17:57], base.param[17:33]]) * 0.9
    o0 = head + 80 + 56
    par[o0:o0 + 40] = base.param[57:97]
    par[o0 + 40:o0 + 80] = base.param[57:97]
    par[o0 + 80:o0 + 136] = np.concatenate([base.param[57:97], base.param[57:73]]) * 0.9
    p = S.NRLParams(2, [9, 9], [11.0, 11.0], [12.0, 12.0, 12.0], [0.5, 0.5, 0.5], par, 0, True)
    lat, frac = S.conventional_cell("bcc", 5.4)
    st = S.Structure(lat, frac @ lat, [1, 2], p, (3, 3, 3), frac)
    o, R = oracle.Setup(st), oracle.RefSetup(st)
    assert o.npairs == R.npairs
    for nm in ("hop", "over", "dhop", "dover", "paros", "dparos"):
        a, b = getattr(o, nm), getattr(R, nm)
        assert np.array_equal(a[:len(b)], b), nm
    for ak in ([0.1, 0.2, 0.3], [0.5, 0.0, 0.25]):
        h1, s1 = o.hs(ak)
        h2, s2 = R.hs(ak)
        assert np.array_equal(h1, h2) and np.array_equal(s1, s2)


@needs_ref
def test_hs_eig_perturb_vs_reference():
    st = S.config_c1()
    o, R = oracle.Setup(st), oracle.RefSetup(st)
    n = st.nv
    for ak in ([0.05, 0.15, 0.25], [0.45, 0.05, 0.35], [0.5, 0.5, 0.5]):
        h1, s1 = o.hs(ak)
        h2, s2 = R.hs(ak)
        assert np.array_equal(h1, h2) and np.array_equal(s1, s2)
        info, w, z = oracle.zhegv(o.unpack(n, h1), o.unpack(n, s1))
        io, wr, zr = R.diag(h2, s2)
        assert info == 0 and io == 0
        assert np.abs(w - wr).max() < 1e-12
        dh, ds = o.hs(ak, deriv=True)
        dh2, ds2 = R.hs(ak, deriv=True)
        assert np.array_equal(dh, dh2) and np.array_equal(ds, ds2)
        nt, dr = R.perturb(dh2, ds2, wr)
        import ctypes as C
        d = np.zeros(n)
        ntc = C.c_int(0)
        oracle.lib().orc_perturb(n, dh, ds, np.ascontiguousarray(z.T).ravel(), w, 10.0,
                                 C.byref(ntc), d)
        assert nt == ntc.value == n
        assert np.abs(d - dr).max() < 1e-10


@needs_ref
def test_bande_vs_reference_c1_slice():
    st = S.config_c1()
    k = S.monkhorst_pack((4, 4, 4))
    w = np.full(len(k), 1.0 / len(k))
    for mode in (1, 2, 3, 4):
        r = oracle.Setup(st).bande(k, w, mode=mode, nthreads=4)
        rr = oracle.RefSetup(st).bande(k, w, mode=mode)
        assert abs(r["efermi"] - rr["efermi"]) < 1e-12
        assert abs(r["earg1"] - rr["earg1"]) / st.natoms < 1e-12
        scale = S.RYBOHR3_KBAR if mode in (2, 4) else 1.0
        assert abs(r["earg2"] - rr["earg2"]) * scale < 1e-7


@needs_ref
def test_fermi_vs_reference():
    rng = np.random.default_rng(2)
    eig = np.sort(rng.uniform(-0.5, 1.5, (50, 36)), axis=1)
    w = rng.uniform(0.5, 1.5, 50)
    w /= w.sum()
    for z in (44.0, 43.5, 20.0):
        a = oracle.lib().orc_fermi(36, 50, 36, 2.0, w, eig.ravel(), 0.005, z)
        import ctypes as C
        b = oracle.ref().ref_fermi(C.byref(C.c_int(50)), C.byref(C.c_double(2.0)), w,
                                   eig.ravel(), C.byref(C.c_double(0.005)),
                                   C.byref(C.c_double(z)))
        assert a == b


def test_zhegv_vs_lapack():
    rng = np.random.default_rng(0)
    for n in (2, 9, 36, 90):
        a = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
        h = a + a.conj().T
        b = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
        s = b @ b.conj().T / n + np.eye(n)
        info, w, z = oracle.zhegv(h, s)
        assert info == 0
        ws = sl.eigh(h, s, eigvals_only=True)
        scale = max(1.0, np.abs(ws).max())
        assert np.abs(w - ws).max() < 1e-12 * scale
        assert np.abs(z.conj().T @ s @ z - np.eye(n)).max() < 1e-12
        assert np.abs(h @ z - s @ z * w).max() < 1e-11 * scale


def test_zhegv_reports_indefinite_overlap():
    h = np.eye(3, dtype=complex)
    s = np.diag([1.0, -1.0, 1.0]).astype(complex)
    info, _, _ = oracle.zhegv(h, s)
    assert info == 2     # zpptrf: leading minor of order 2


def test_s_band_closed_form():
    """s-only, orthogonal-like check: one atom, only ss-sigma, the band is
    e(k) = eps_s + sum_R h(R) cos(2 pi k.R) divided by the overlap sum."""
    base = S.synthetic_params(1)
    p = S.NRLParams(1, [1], [1.0], [9.0], [0.5], base.param.copy(), 0, True)
    st = S.Structure(S.fcc_primitive(6.8), np.zeros((1, 3)), [1], p, (3, 3, 3))
    o = oracle.Setup(st)
    for ak in ([0.1, 0.2, 0.3], [0.0, 0.0, 0.0], [0.5, 0.5, 0.0]):
        hp, sp = o.hs(ak)
        n = o.npairs
        dlv = o.dlv[:3 * n].reshape(n, 3)
        ph = np.cos(2 * math.pi * dlv @ np.array(ak))
        hk = o.paros[0] + np.sum(o.hop[:14 * n].reshape(n, 14)[:, 0] * ph)
        sk = 1.0 + np.sum(o.over[:14 * n].reshape(n, 14)[:, 0] * ph)
        assert hp[0].real == pytest.approx(hk, abs=1e-13)
        assert sp[0].real == pytest.approx(sk, abs=1e-13)


def test_periodicity_and_rotation_invariance():
    st = S.config_c1()
    # NRL keeps separate t2g and e_g on-site terms (setham.f:430-446), a cubic
    # split fixed in the lab frame; rotation invariance needs them equal.
    par = st.params.param.copy()
    par[13:17] = par[9:13]
    st.params = S.NRLParams(1, [9], [11.0], st.params.dnn, st.params.screenl, par, 0, True)
    o = oracle.Setup(st)
    n = st.nv
    ak = np.array([0.13, 0.27, 0.41])
    e1 = sl.eigh(*[o.unpack(n, m) for m in o.hs(ak)], eigvals_only=True)
    e2 = sl.eigh(*[o.unpack(n, m) for m in o.hs(ak + [1, -2, 3])], eigvals_only=True)
    assert np.abs(e1 - e2).max() < 1e-11
    # rigid rotation of the crystal: the SK table must rotate with it
    th = 0.7
    rot = np.array([[math.cos(th), -math.sin(th), 0], [math.sin(th), math.cos(th), 0], [0, 0, 1]])
    ax = np.array([1.0, 2.0, 0.5]) / math.sqrt(5.25)
    k = np.array([[0, -ax[2], ax[1]], [ax[2], 0, -ax[0]], [-ax[1], ax[0], 0]])
    rot2 = sl.expm(0.9 * k) @ rot
    st2 = S.Structure(st.lat @ rot2.T, st.posn @ rot2.T, st.kkind, st.params, st.nd, st.posq)
    o2 = oracle.Setup(st2)
    e3 = sl.eigh(*[o2.unpack(n, m) for m in o2.hs(ak)], eigvals_only=True)
    assert np.abs(e1 - e3).max() < 1e-10


def test_fermi_counts_electrons_and_insulator():
    rng = np.random.default_rng(3)
    eig = np.sort(rng.uniform(-0.5, 1.5, (20, 36)), axis=1)
    w = np.full(20, 0.05)
    z = 30.0
    ef = oracle.lib().orc_fermi(36, 20, 36, 2.0, w, eig.ravel(), 0.005, z)
    f = 1.0 / (1.0 + np.exp((eig - ef) / 0.005))
    f[eig < ef - 0.1] = 1.0
    f[eig >= ef + 0.1] = 0.0
    assert abs(2.0 * np.sum(w[:, None] * f) - z) < 1e-7
    # gapped spectrum: 15 bands at -1, 21 at +1 -> insulator shortcut, mid gap
    eg = np.concatenate([np.full((20, 15), -1.0), np.full((20, 21), 1.0)], axis=1)
    assert oracle.lib().orc_fermi(36, 20, 36, 2.0, w, eg.ravel(), 0.005, 30.0) == 0.0


def test_pressure_is_minus_dE_dV():
    """First-order pressure (perturb.f + bande.f sums) against a central
    finite difference of E(T) under isotropic scaling."""
    base = S.config_c1()
    k = S.monkhorst_pack((3, 3, 3))
    w = np.full(len(k), 1.0 / len(k))

    def scaled(f):
        return S.Structure(base.lat * f, base.posn * f, base.kkind, base.params, base.nd,
                           base.posq)

    r0 = oracle.Setup(base).bande(k, w, mode=4, nthreads=4)
    h = 1e-4
    ep = oracle.Setup(scaled(1 + h)).bande(k, w, mode=4, nthreads=4)["earg1"]
    em = oracle.Setup(scaled(1 - h)).bande(k, w, mode=4, nthreads=4)["earg1"]
    dv = base.volume * ((1 + h) ** 3 - (1 - h) ** 3)
    p_fd = -(ep - em) / dv
    assert abs(r0["earg2"] - p_fd) < 2e-6 * max(1.0, abs(p_fd)) + 1e-7


def test_oracle_vs_reference_at_supercell_size():
    """The restatement against the reference's own runs at the sizes of the
    benchmark workloads (tests/golden/ref_large.npz, the reference built with
    matom = 128): eigenvalues of the 32-atom (N = 288) and 108-atom (N = 972)
    cells at a few k-points, and E_F, E(T), P(T) of three structures of the
    uneven batch."""
    gold = np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_large.npz"))
    seed = int(gold["seed"][0])
    for name, st, k, idx in (("c2", S.config_c2(seed), S.monkhorst_pack((8, 8, 8)), (0, 137, 511)),
                             ("c3", S.config_c3(seed), S.monkhorst_pack((4, 4, 4)), (21,))):
        o = oracle.Setup(st)
        for i in idx:
            hp, sp = o.hs(k[i])
            info, e, _ = oracle.zhegv(o.unpack(st.nv, hp), o.unpack(st.nv, sp), vectors=False)
            assert info == 0
            assert np.abs(e - gold[name + "_eig"][i]).max() < 1e-12, (name, i)
    pool = S.config_c4(seed)
    for i in (0, 1, 2):
        st, q = pool[i]
        k = S.monkhorst_pack((q, q, q))
        r = oracle.Setup(st).bande(k, np.full(len(k), 1.0 / len(k)), mode=4, nthreads=8)
        ef, e1, e2 = gold["c4_out"][i]
        assert abs(r["efermi"] - ef) < 1e-12
        assert abs(r["earg1"] - e1) / st.natoms < 1e-12
        assert abs(r["earg2"] - e2) * S.RYBOHR3_KBAR < 1e-7
