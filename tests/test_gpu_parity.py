"""Parity of the HIP path (lib/libtbk.so on the GPU) with the oracle.

Tolerances of the plain FP64 method: eigenvalues 1e-10 Ry, E_F and E_tot
1e-9 Ry (per atom), pressure 1e-6 kbar; H/S elements 1e-14 (the build is
the reference's arithmetic except for the ulp of sin/cos).
"""
import json
import os

import numpy as np
import pytest

import oracle
from tbk import _lib
from tbk import structures as S
from tbk.api import Plan, fermi, zhegv

pytestmark = pytest.mark.gpu

EIG_TOL, EF_TOL, E_TOL, P_TOL_KBAR = 1e-10, 1e-9, 1e-9, 1e-6
GOLD = os.path.join(os.path.dirname(__file__), "golden", "ex1.json")


def full(o, packed, n):
    return o.unpack(n, packed)


def compare_bande(st, k, w, mode=4, tkb=0.005, eigcut=10.0, batch=0, window=True):
    print(f"case: natoms={st.natoms} N={st.nv} nk={len(k)} mode={mode} batch={batch} window={window}",
          flush=True)
    plan = Plan(st, k, w, batch=batch)
    plan.set_window(window)
    r = plan.run(mode, tkb, eigcut)
    e, d = plan.eigenvalues()
    plan.close()
    ref = oracle.Setup(st).bande(k, w, mode=mode, tkb=tkb, eigcut=eigcut, nthreads=16,
                                 want_eig=True)
    assert ref["err"] == 0
    de = float(np.abs(e - ref["eig"]).max())
    def_ = abs(r["efermi"] - ref["efermi"])
    dE = abs(r["earg1"] - ref["earg1"]) / st.natoms
    print(f"  max|de|={de:.2e} dEF={def_:.2e} dE/atom={dE:.2e}", flush=True)
    assert de < EIG_TOL
    assert def_ < EF_TOL
    assert dE < E_TOL
    if mode in (2, 4):
        dP = abs(r["earg2"] - ref["earg2"]) * S.RYBOHR3_KBAR
        # with the occupation window (one batch) derivatives are formed only
        # below E_F + 20 kT, the states the sums use; zero above
        used = ref["eig"] < r["efermi"] + 20 * tkb if (window and batch == 0) else np.ones(d.shape, bool)
        dd = float(np.abs(d - ref["desk"])[used].max())
        print(f"  dP={dP:.2e} kbar max|d desk|={dd:.2e}", flush=True)
        assert dP < P_TOL_KBAR
        assert dd < 1e-8
    else:
        assert abs(r["earg2"] - ref["earg2"]) / st.natoms < E_TOL
    return r, e


@pytest.mark.parametrize("cfg", ["c1", "c2"])
def test_hs_build_matches_oracle(cfg):
    st = S.config_c1() if cfg == "c1" else S.config_c2()
    k = np.array([[0.0, 0.0, 0.0], [0.125, 0.375, -0.25], [0.5, 0.5, 0.5], [0.31, -0.17, 0.44]])
    plan = Plan(st, k, np.full(len(k), 0.25), batch=1)
    o = oracle.Setup(st)
    n = st.nv
    for i, ak in enumerate(k):
        for deriv in (False, True):
            h, s = plan.hs(i, deriv)
            hp, sp = o.hs(ak, deriv)
            dh = float(np.abs(h - full(o, hp, n)).max())
            ds = float(np.abs(s - full(o, sp, n)).max())
            assert dh < 1e-14 and ds < 1e-14, (i, deriv, dh, ds)
    plan.close()


@pytest.mark.parametrize("n", [2, 9, 36, 97, 288])
def test_zhegv_random_matches_oracle(n):
    rng = np.random.default_rng(n)
    a = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    h = a + a.conj().T
    b = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    s = b @ b.conj().T / n + np.eye(n)
    w, z = zhegv(h, s)
    info, wo, _ = oracle.zhegv(h, s, vectors=False)
    scale = max(1.0, np.abs(wo).max())
    assert info == 0
    de = float(np.abs(w - wo).max())
    orth = float(np.abs(z.conj().T @ s @ z - np.eye(n)).max())
    res = float(np.abs(h @ z - s @ z * w).max())
    print(f"n={n}: |dw|={de:.2e} |Z^H S Z - I|={orth:.2e} |HZ-SZW|={res:.2e}", flush=True)
    assert de < 1e-12 * scale
    assert orth < 1e-11
    assert res < 1e-11 * scale


def test_zhegv_degenerate_spectrum():
    """Exactly degenerate eigenvalues (inverse iteration with
    reorthogonalisation inside the group)."""
    n = 40
    rng = np.random.default_rng(7)
    q, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    lam = np.repeat(np.linspace(-1, 1, 8), 5)
    h = (q * lam) @ q.conj().T
    h = 0.5 * (h + h.conj().T)
    s = np.eye(n, dtype=complex)
    w, z = zhegv(h, s)
    de = float(np.abs(w - np.sort(lam)).max())
    orth = float(np.abs(z.conj().T @ z - np.eye(n)).max())
    res = float(np.abs(h @ z - z * w).max())
    print(f"degenerate: |dw|={de:.2e} orth={orth:.2e} res={res:.2e}", flush=True)
    assert de < 1e-12 and orth < 1e-11 and res < 1e-11


def test_zhegv_indefinite_overlap_is_an_error():
    h = np.eye(4, dtype=complex)
    s = np.diag([1.0, 1.0, -0.5, 1.0]).astype(complex)
    with pytest.raises(_lib.TbkError) as ei:
        zhegv(h, s)
    assert ei.value.code == _lib.TBK_ERR_OVERLAP


def test_c1_full_grid_all_modes():
    st = S.config_c1()
    k = S.monkhorst_pack((10, 10, 10))
    w = np.full(len(k), 1.0 / len(k))
    for mode in (4, 1, 2, 3):
        compare_bande(st, k, w, mode=mode)
    compare_bande(st, k, w, mode=4, window=False)      # every derivative against the oracle


def test_c1_odd_grid_degenerate_gamma():
    """An odd grid contains Gamma, where the perfect fcc cell has exactly
    degenerate levels (perturb.f degenerate groups)."""
    st = S.config_c1()
    k = S.monkhorst_pack((5, 5, 5))
    compare_bande(st, k, np.full(len(k), 1.0 / len(k)))


def test_c2_slice():
    st = S.config_c2()
    k = S.monkhorst_pack((8, 8, 8))[::16]
    compare_bande(st, k, np.full(len(k), 1.0 / len(k)))


def test_ex1_published_results_on_gpu():
    """The reference's own example: 18 Pd structures, mode 4, eigcut 0.5."""
    gold = json.load(open(GOLD))
    for s in gold["structures"]:
        p = S.NRLParams(1, [9], [gold["valence"]], [gold["rcut"]], [gold["screenl"]],
                        np.array(gold["param"]), 0, True)
        lat = S.fcc_primitive(s["a"]) if s["lattice"] == "fcc" else S.bcc_primitive(s["a"])
        st = S.Structure(lat, np.zeros((1, 3)), [1], p, gold["nd"])
        kl = np.array(gold["kpoints"]["fcc.08" if s["lattice"] == "fcc" else "bcc.08"])
        plan = Plan(st, kl[:, :3], kl[:, 3] / kl[:, 3].sum())
        r = plan.run(gold["mode"], gold["tkb"], gold["eigcut"])
        plan.close()
        print(f"ex1 {s['lattice']} a={s['a']}: dEF={r['efermi'] - s['efermi']:.2e}", flush=True)
        assert abs(r["efermi"] - s["efermi"]) < EF_TOL
        assert abs(r["earg1"] - s["et"]) < E_TOL
        assert abs(r["earg2"] - s["pt"]) * S.RYBOHR3_KBAR < P_TOL_KBAR


def test_fermi_on_device_matches_oracle():
    rng = np.random.default_rng(11)
    eig = np.sort(rng.uniform(-0.5, 1.5, (64, 72)), axis=1)
    w = rng.uniform(0.5, 1.5, 64)
    w /= w.sum()
    for z in (88.0, 87.25, 30.0):
        ef = fermi(eig, w, z)
        ef_o = oracle.lib().orc_fermi(72, 64, 72, 2.0, w, eig.ravel(), 0.005, z)
        assert abs(ef - ef_o) < 1e-12


def test_batch_invariance():
    """Results do not depend on how k-points are batched."""
    st = S.config_c1()
    k = S.monkhorst_pack((4, 4, 4))
    w = np.full(len(k), 1.0 / len(k))
    outs = []
    for b in (0, 1, 7):
        plan = Plan(st, k, w, batch=b)
        outs.append((plan.run(4), plan.eigenvalues()))
        plan.close()
    # one batch uses the occupation window: derivatives agree where formed
    used = outs[0][1][0] < outs[0][0]["efermi"] + 20 * 0.005
    for r, (e, d) in outs[1:]:
        assert r == outs[0][0]
        assert np.array_equal(e, outs[0][1][0]) and np.array_equal(d[used], outs[0][1][1][used])
    plan = Plan(st, k, w, batch=0)
    plan.set_window(False)
    assert plan.run(4) == outs[0][0]
    assert np.array_equal(plan.eigenvalues()[1], outs[1][1][1])
    plan.close()


def test_indefinite_overlap_structure_is_an_error():
    p = S.synthetic_params(1, overlap_scale=3.0)
    st = S.synthetic_structure("fcc", (1, 1, 1), 3.3 * S.ANGSTROM, p)
    with pytest.raises(_lib.TbkError) as ei:
        Plan(st, [[0.0, 0.0, 0.0]], [1.0]).run(4)
    assert ei.value.code == _lib.TBK_ERR_OVERLAP


# Mixed precision (FP32 tridiagonalisation and back-transform, one FP64
# refinement step with clusters rotated exactly): bounds about ten times the
# largest error measured on these cases (C1 6^3 and C2: eigenvalues 8.2e-13 Ry,
# E_F 1.4e-14 Ry, E/atom 7.9e-15 Ry, P 1.0e-8 kbar, |Z^H S Z - I| 7.2e-9);
# residual |HZ - SZ diag(e)| 1e-9.  One refinement step leaves
# |Z^H S Z - I| of order |E|^2 (|E| <= 1e-2 between states of different
# clusters) and of the FP32 orthogonality defect squared.
MIX_EF, MIX_E, MIX_EIG, MIX_P_KBAR, MIX_ORTH = 2e-13, 1e-13, 1e-11, 2e-7, 1e-7


@pytest.mark.parametrize("cfg", ["c1", "c2"])
def test_mixed_precision_error_vs_oracle(cfg):
    if cfg == "c1":
        st, k = S.config_c1(), S.monkhorst_pack((6, 6, 6))
    else:
        st, k = S.config_c2(), S.monkhorst_pack((8, 8, 8))[::16]
    w = np.full(len(k), 1.0 / len(k))
    print(f"mixed case {cfg}: N={st.nv} nk={len(k)}", flush=True)
    plan = Plan(st, k, w, precision=_lib.TBK_MIXED)
    plan.set_window(False)            # every eigenvector, for the orthogonality check
    r = plan.run(4, 0.005, 10.0)
    e, d = plan.eigenvalues()
    z0 = plan.eigenvectors(0)
    h0, s0 = plan.hs(0)
    plan.close()
    ref = oracle.Setup(st).bande(k, w, mode=4, nthreads=16, want_eig=True)
    de = float(np.abs(e - ref["eig"]).max())
    def_ = abs(r["efermi"] - ref["efermi"])
    dE = abs(r["earg1"] - ref["earg1"]) / st.natoms
    dP = abs(r["earg2"] - ref["earg2"]) * S.RYBOHR3_KBAR
    orth = float(np.abs(z0.conj().T @ s0 @ z0 - np.eye(st.nv)).max())
    res = float(np.abs(h0 @ z0 - s0 @ z0 * e[0]).max())
    print(f"  max|de|={de:.2e} dEF={def_:.2e} dE/atom={dE:.2e} dP={dP:.2e} kbar "
          f"orth={orth:.2e} res={res:.2e}", flush=True)
    assert def_ < MIX_EF and dE < MIX_E
    assert de < MIX_EIG
    assert dP < MIX_P_KBAR
    assert orth < MIX_ORTH and res < 1e-9


@pytest.mark.parametrize("prec", [_lib.TBK_FP64, _lib.TBK_MIXED])
def test_tile_and_panel_tridiagonalisation_agree(prec):
    """Batches below 160 items of order >= 128 tridiagonalise over 64x64
    tiles (several workgroups per item); larger batches with one workgroup
    per item.  Both give the same results."""
    st = S.config_c2()
    k = S.monkhorst_pack((8, 8, 8))[:192]
    w = np.full(len(k), 1.0 / len(k))
    out = {}
    for b in (192, 16):
        p = Plan(st, k, w, batch=b, precision=prec)
        out[b] = (p.run(4), p.eigenvalues()[0])
        p.close()
    de = float(np.abs(out[192][1] - out[16][1]).max())
    dr = max(abs(out[192][0][x] - out[16][0][x]) for x in ("efermi", "earg1", "earg2"))
    print(f"tile vs panel (precision {prec}): max|de|={de:.2e} max|dresult|={dr:.2e}", flush=True)
    tol = 1e-11 if prec == _lib.TBK_FP64 else 1e-9
    assert de < tol and dr < tol


@pytest.fixture(scope="module")
def c2_full_ref():
    st = S.config_c2()
    k = S.monkhorst_pack((8, 8, 8))
    w = np.full(len(k), 1.0 / len(k))
    ref = oracle.Setup(st).bande(k, w, mode=4, nthreads=16, want_eig=True)
    assert ref["err"] == 0
    return st, k, w, ref


@pytest.mark.parametrize("prec", [_lib.TBK_FP64, _lib.TBK_MIXED])
def test_c2_full_bench_workload_vs_oracle(c2_full_ref, prec):
    """The bench's own workload -- all 512 k of C2 in one batch, occupation
    window on -- against the oracle: the plain bounds in FP64, the mixed
    bounds in mixed precision."""
    st, k, w, ref = c2_full_ref
    print(f"case: C2 full, 512 k, precision {prec}", flush=True)
    plan = Plan(st, k, w, precision=prec)
    r = plan.run(4, 0.005, 10.0)
    e, d = plan.eigenvalues()
    plan.close()
    de = float(np.abs(e - ref["eig"]).max())
    def_ = abs(r["efermi"] - ref["efermi"])
    dE = abs(r["earg1"] - ref["earg1"]) / st.natoms
    dP = abs(r["earg2"] - ref["earg2"]) * S.RYBOHR3_KBAR
    used = ref["eig"] < r["efermi"] + 20 * 0.005
    dd = float(np.abs(d - ref["desk"])[used].max())
    print(f"  max|de|={de:.2e} dEF={def_:.2e} dE/atom={dE:.2e} dP={dP:.2e} kbar max|d desk|={dd:.2e}",
          flush=True)
    if prec == _lib.TBK_FP64:
        assert de < EIG_TOL and def_ < EF_TOL and dE < E_TOL and dP < P_TOL_KBAR and dd < 1e-8
    else:
        assert def_ < MIX_EF and dE < MIX_E and de < MIX_EIG and dP < MIX_P_KBAR
