"""The device path on the branch cases of tests/branch_cases.py against the
reference's own results (tests/golden/branches.json): two species, sp and
s-only bases, nltype 1/2/3/90000, an orthogonal basis; and the two
refusals, a degenerate group above perturb's limit of 9 (TBK_ERR_DEGEN where
the reference stops, perturb.f:106) and zfill = 0 (TBK_ERR_ZFILL,
fermi.f:165).

Bounds: FP64 as the plain method's (eigenvalues 1e-10 Ry, E_F and E/atom
1e-9 Ry, P 1e-6 kbar; H(k), S(k) 1e-14); mixed precision within a few times
the largest error measured on these cases (eigenvalues 5e-10 Ry, E_F 1e-10
Ry, E/atom 1e-10 Ry, P 3e-5 kbar).
"""
import json
import os

import numpy as np
import pytest

import branch_cases as B
import oracle
from tbk import _lib
from tbk import structures as S
from tbk.api import Plan

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "branches.json")))
CASES = {c["name"]: c for c in B.cases(GOLD["c_par"])}
OK = [n for n, c in CASES.items() if c["expect"] == "ok"]
TOL = {_lib.TBK_FP64: dict(eig=1e-10, ef=1e-9, e=1e-9, p=1e-6),
       _lib.TBK_MIXED: dict(eig=5e-14, ef=2e-13, e=2e-13, p=2e-8)}


def packed_full(gold, key, n):
    a = np.array(gold[key], float)
    ap = a[:, 0] + 1j * a[:, 1]
    m = np.zeros((n, n), complex)
    # packed upper, column-major: AP(i + j(j+1)/2) = A(i, j), i <= j
    idx = [(i, j) for j in range(n) for i in range(j + 1)]
    for v, (i, j) in zip(ap, idx):
        m[j, i] = np.conj(v)
        m[i, j] = v
    return m


@pytest.mark.parametrize("name", OK)
def test_hs_build_matches_reference(name):
    case, gold = CASES[name], GOLD["cases"][name]
    st = B.structure(case)
    print(f"case: branch {name} H/S at a general k (N={st.nv})", flush=True)
    plan = Plan(st, [B.GENERAL_K], [1.0], batch=1)
    try:
        h, s = plan.hs(0)
        dh = float(np.abs(h - packed_full(gold, "hp", st.nv)).max())
        if st.params.jsover:
            ds = float(np.abs(s - packed_full(gold, "sp", st.nv)).max())
        else:
            ds = float(np.abs(s - np.eye(st.nv)).max())
        o = oracle.Setup(st)
        dhv, dsv = plan.hs(0, deriv=True)
        hp, sp = o.hs(B.GENERAL_K, deriv=True)
        ddh = float(np.abs(dhv - o.unpack(st.nv, hp)).max())
        dds = float(np.abs(dsv - o.unpack(st.nv, sp)).max())
        print(f"  |dH|={dh:.1e} |dS|={ds:.1e} |d dH/dV|={ddh:.1e} |d dS/dV|={dds:.1e}", flush=True)
        assert dh < 1e-14 and ds < 1e-14 and ddh < 1e-14 and dds < 1e-14
    finally:
        plan.close()


@pytest.mark.parametrize("prec", [_lib.TBK_FP64, _lib.TBK_MIXED])
@pytest.mark.parametrize("name", OK)
def test_bande_matches_reference(name, prec):
    case, gold = CASES[name], GOLD["cases"][name]
    st = B.structure(case)
    tol = TOL[prec]
    print(f"case: branch {name} N={st.nv} precision {prec}", flush=True)
    plan = Plan(st, [B.GENERAL_K], [1.0], batch=1, precision=prec)
    try:
        plan.run(mode=3, tkb=GOLD["tkb"], eigcut=GOLD["eigcut"])
        e, _ = plan.eigenvalues()
        de = float(np.abs(e[0] - np.array(gold["eig"])).max())
    finally:
        plan.close()
    k, w = B.mesh(case)
    plan = Plan(st, k, w, batch=0, precision=prec)
    worst = [de, 0.0, 0.0, 0.0]
    try:
        for mode in (1, 2, 3, 4):
            r = plan.run(mode=mode, tkb=GOLD["tkb"], eigcut=GOLD["eigcut"])
            ef, e1, e2 = gold["modes"][str(mode)]
            worst[1] = max(worst[1], abs(r["efermi"] - ef))
            worst[2] = max(worst[2], abs(r["earg1"] - e1) / st.natoms)
            if mode in (2, 4):
                worst[3] = max(worst[3], abs(r["earg2"] - e2) * S.RYBOHR3_KBAR)
            else:
                worst[2] = max(worst[2], abs(r["earg2"] - e2) / st.natoms)
    finally:
        plan.close()
    print(f"  max|de|={worst[0]:.1e} dEF={worst[1]:.1e} dE/atom={worst[2]:.1e} dP={worst[3]:.1e} kbar",
          flush=True)
    assert worst[0] < tol["eig"] and worst[1] < tol["ef"] and worst[2] < tol["e"] and worst[3] < tol["p"]


@pytest.mark.parametrize("prec", [_lib.TBK_FP64, _lib.TBK_MIXED])
def test_degenerate_group_is_refused(prec):
    """More than 9 states within 1e-8 Ry: the reference stops in perturb
    (the fixture holds the line it prints); the device returns
    TBK_ERR_DEGEN in the pressure modes and runs the energy-only ones."""
    assert GOLD["cases"]["degenerate"]["stop"] == B.DEGEN_STOP
    case = CASES["degenerate"]
    st = B.structure(case)
    k, w = B.mesh(case)
    print(f"case: degenerate group N={st.nv} precision {prec}", flush=True)
    plan = Plan(st, k, w, precision=prec)
    try:
        for mode in (2, 4):
            with pytest.raises(_lib.TbkError) as ei:
                plan.run(mode=mode)
            assert ei.value.code == _lib.TBK_ERR_DEGEN
        r = plan.run(mode=3)
    finally:
        plan.close()
    ref = oracle.Setup(st).bande(k, w, mode=3)
    assert abs(r["efermi"] - ref["efermi"]) < 1e-9
    assert abs(r["earg1"] - ref["earg1"]) / st.natoms < 1e-9


def test_zfill_zero_is_refused():
    assert GOLD["cases"]["no_electrons"]["stop"] == B.ZFILL_STOP
    case = CASES["no_electrons"]
    st = B.structure(case)
    k, w = B.mesh(case)
    print("case: zfill = 0", flush=True)
    plan = Plan(st, k, w)
    try:
        with pytest.raises(_lib.TbkError) as ei:
            plan.run(mode=4)
        assert ei.value.code == _lib.TBK_ERR_ZFILL
    finally:
        plan.close()
