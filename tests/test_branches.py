"""Branches of the path the main workloads do not reach (tests/branch_cases.py:
two species, sp and s-only bases, nltype 1/2/3/90000, an orthogonal basis,
a too-large degenerate group, zfill = 0), on the CPU.

* the oracle against the reference's own results (tests/golden/branches.json,
  written by tests/golden/make_branch_fixture.py from the reference's
  sources compiled where they lie): E_F, earg1, earg2 in modes 1-4, the
  packed H(k), S(k) and the eigenvalues at a general k;
* the oracle's set-up bit-identical to the reference's own setpar/setparv
  where the reference is built;
* the library's host set-up bit-identical to the oracle's;
* the refusals: the oracle reports the degenerate group the reference stops
  on, and the library's Fermi entry point returns TBK_ERR_ZFILL for
  zfill <= 0 before touching a device.
"""
import json
import os

import numpy as np
import pytest

import branch_cases as B
import oracle
from tbk import _lib
from tbk import structures as S
from tbk.api import HostSetup

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "branches.json")))
CASES = {c["name"]: c for c in B.cases(GOLD["c_par"])}
OK = [n for n, c in CASES.items() if c["expect"] == "ok"]
needs_ref = pytest.mark.skipif(not oracle.ref_available(), reason="reference build (oracle/_ref) not present")


def packed(gold, key):
    a = np.array(gold[key], float)
    return a[:, 0] + 1j * a[:, 1]


def test_cases_reach_their_branches():
    c = {n: B.structure(CASES[n]) for n in CASES}
    assert c["two_kind"].params.kinds == 2 and list(c["two_kind_sp"].params.kbas) == [9, 4]
    assert c["diamond_C"].params.kbas == [4] and c["diamond_C"].params.nltype == 1
    assert c["diamond_C_nl3"].params.realov and c["diamond_C_nl3"].params.harrison
    assert c["pd_harrison"].params.harrison and not c["pd_harrison"].params.realov
    assert c["h_atom"].params.nltype == 90000 and c["h_atom"].params.kbas == [1]
    assert not c["orthogonal"].params.jsover
    assert c["no_electrons"].zfill == 0.0
    # the two-kind cell has like and unlike pairs
    o = oracle.Setup(c["two_kind"])
    kinds = {(int(a), int(b)) for a, b in zip(o.jkind[0:2 * o.npairs:2], o.jkind[1:2 * o.npairs:2])}
    assert len(kinds) >= 2, kinds


@pytest.mark.parametrize("name", OK)
def test_oracle_vs_reference_fixture(name):
    case, gold = CASES[name], GOLD["cases"][name]
    st = B.structure(case)
    o = oracle.Setup(st)
    assert o.npairs == gold["npairs"] and st.nv == gold["nv"]
    hp, sp = o.hs(B.GENERAL_K)
    assert np.array_equal(hp, packed(gold, "hp"))
    if st.params.jsover:
        assert np.array_equal(sp, packed(gold, "sp"))
    else:
        # setham leaves smat alone for an orthogonal basis (setham.f:279,
        # 443) and diag solves with H alone (diag.f:385-414): S = I
        assert np.array_equal(o.unpack(st.nv, sp), np.eye(st.nv))
    info, e, _ = oracle.zhegv(o.unpack(st.nv, hp), o.unpack(st.nv, sp), vectors=False)
    assert info == 0
    de = float(np.abs(e - np.array(gold["eig"])).max())
    assert de < 1e-12, de
    k, w = B.mesh(case)
    for mode in (1, 2, 3, 4):
        r = o.bande(k, w, mode=mode, tkb=GOLD["tkb"], eigcut=GOLD["eigcut"])
        ef, e1, e2 = gold["modes"][str(mode)]
        assert r["err"] == 0
        assert abs(r["efermi"] - ef) < 1e-12, (mode, r["efermi"], ef)
        assert abs(r["earg1"] - e1) / st.natoms < 1e-12, (mode, r["earg1"], e1)
        scale = S.RYBOHR3_KBAR if mode in (2, 4) else 1.0 / st.natoms
        assert abs(r["earg2"] - e2) * scale < 1e-7, (mode, r["earg2"], e2)


@needs_ref
@pytest.mark.parametrize("name", OK)
def test_oracle_setup_bit_exact_vs_reference(name):
    st = B.structure(CASES[name])
    o, R = oracle.Setup(st), oracle.RefSetup(st)
    n = R.npairs
    assert o.npairs == n and o.dismin == R.dismin
    for nm in ("tt", "dlv", "dist", "screen", "dscrn"):
        assert np.array_equal(getattr(o, nm)[:len(getattr(R, nm))], getattr(R, nm)), nm
    for nm in ("paros", "hop", "over", "dparos", "dhop", "dover"):
        a, b = getattr(o, nm), getattr(R, nm)
        assert np.array_equal(a[:len(b)], b), nm
    dh, ds = o.hs(B.GENERAL_K, deriv=True)
    dh2, ds2 = R.hs(B.GENERAL_K, deriv=True)
    assert np.array_equal(dh, dh2) and np.array_equal(ds, ds2)


@pytest.mark.parametrize("name", list(CASES))
def test_library_setup_bit_exact_vs_oracle(name):
    st = B.structure(CASES[name])
    o, h = oracle.Setup(st), HostSetup(st)
    n = o.npairs
    assert h.npairs == n and h.dismin == o.dismin
    for nm, m in (("tt", 3 * n), ("dlv", 3 * n), ("dist", n), ("screen", n), ("dscrn", n),
                  ("jatm", 2 * n), ("jkind", 2 * n), ("paros", None), ("dparos", None),
                  ("hop", 14 * n), ("over", 14 * n), ("dhop", 14 * n), ("dover", 14 * n), ("den", None)):
        a, b = getattr(o, nm), getattr(h, nm)
        m = m or len(b)
        assert np.array_equal(a[:m], b[:m]), nm


def test_oracle_refuses_the_degenerate_group():
    """The reference stops in perturb (tests/golden/branches.json holds the
    line it prints); the oracle reports the same case as an error in the
    pressure modes and runs the energy-only ones."""
    assert GOLD["cases"]["degenerate"]["stop"] == B.DEGEN_STOP
    case = CASES["degenerate"]
    st = B.structure(case)
    o = oracle.Setup(st)
    k, w = B.mesh(case)
    assert o.bande(k, w, mode=4)["err"] == 5
    r = o.bande(k, w, mode=3)
    assert r["err"] == 0 and np.isfinite(r["efermi"])


def test_fermi_refuses_zfill_zero():
    assert GOLD["cases"]["no_electrons"]["stop"] == B.ZFILL_STOP
    L = _lib.lib()
    e = np.linspace(-1.0, 1.0, 8)
    w = np.ones(1)
    ef = np.zeros(1)
    for z in (0.0, -3.0):
        rc = L.tbk_fermi(0, 8, 1, 8, 2.0, w, e, 0.005, z, ef.ctypes.data_as(_lib.C.POINTER(_lib.C.c_double)))
        assert rc == _lib.TBK_ERR_ZFILL
