"""Lowdin projections on the device (tbk_plan_projections, the QLMT modes 6
and 10 of diag.f:526-758) against the same projections formed on the CPU
from the oracle's H(k), S(k) with LAPACK (scipy): S^{1/2} from S's own
eigendecomposition, y = S^{1/2} z, weights |y_i|^2, compared summed over
each group of degenerate levels (within one the individual weights depend on
the basis chosen inside the subspace); every state's weights sum to one
(to the mixed path's orthogonality error in mixed precision).
Cases: two kinds with 9 and 4 orbitals (tests/branch_cases.py) and the
32-atom cell at two k-points, both precisions.
"""
import json
import os

import numpy as np
import pytest
import scipy.linalg as sl

import branch_cases as B
import oracle
from tbk import _lib
from tbk import structures as S
from tbk.api import Plan

pytestmark = pytest.mark.gpu
GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "branches.json")))


def cpu_projections(st, ak, eigcut):
    o = oracle.Setup(st)
    hp, sp = o.hs(ak)
    h, s = o.unpack(st.nv, hp), o.unpack(st.nv, sp)
    e, z = sl.eigh(h, s)
    lam, v = np.linalg.eigh(s)
    u = (v * np.sqrt(lam)) @ v.conj().T
    y = u @ z
    m = int(np.sum(e <= eigcut))
    return e[:m], (np.abs(y[:, :m]) ** 2).T


@pytest.mark.parametrize("prec", [_lib.TBK_FP64, _lib.TBK_MIXED])
@pytest.mark.parametrize("case", ["two_kind_sp", "c2"])
def test_projections_match_cpu(case, prec):
    if case == "c2":
        st = S.config_c2()
        ks = np.array([[0.0, 0.0, 0.0], [0.31, -0.17, 0.44]])
    else:
        st = B.structure({c["name"]: c for c in B.cases(GOLD["c_par"])}[case])
        ks = np.array([B.GENERAL_K, [0.0, 0.0, 0.0]])
    eigcut = 0.7
    print(f"case: projections {case} N={st.nv} precision {prec}", flush=True)
    plan = Plan(st, ks, np.full(len(ks), 1.0 / len(ks)), precision=prec)
    try:
        plan.run(6, 0.005, eigcut)
        worst = 0.0
        for i, ak in enumerate(ks):
            e, w = plan.projections(i, eigcut)
            ec, wc = cpu_projections(st, ak, eigcut)
            assert len(e) == len(ec)
            assert np.abs(e - ec).max() < (1e-10 if prec == _lib.TBK_FP64 else 1e-11)
            # sum_i |y_i|^2 = z^H S z: 1 to rounding in FP64, to the mixed
            # path's orthogonality (|Z^H S Z - I| <= 1e-7 after one refinement)
            assert np.abs(w.sum(axis=1) - 1.0).max() < (1e-10 if prec == _lib.TBK_FP64 else 1e-7)
            # groups of levels within perturb's 1e-8, in which any rotation of
            # the vectors is an eigenbasis
            gtol = 1e-8
            g0 = 0
            while g0 < len(ec):
                g1 = g0 + 1
                while g1 < len(ec) and ec[g1] - ec[g0] < gtol:
                    g1 += 1
                d = np.abs(w[g0:g1].sum(axis=0) - wc[g0:g1].sum(axis=0)).max()
                worst = max(worst, d)
                g0 = g1
    finally:
        plan.close()
    print(f"  max |group weights - CPU| = {worst:.1e}", flush=True)
    assert worst < (1e-11 if prec == _lib.TBK_FP64 else 1e-7)
