"""The request loop: requests that arrive while others run are admitted and
batched with them, and every result is the same alone and in the mix."""
import time

import numpy as np
import pytest

import oracle
from tbk import _lib
from tbk import structures as S
from tbk.api import Plan, Server

pytestmark = pytest.mark.gpu


def test_server_same_alone_and_in_mix():
    c4 = S.config_c4(count=64, max_atoms=16, max_q=4)
    reqs = [(st, S.monkhorst_pack((q, q, q))) for st, q in c4[:16]]
    srv = Server(precision=_lib.TBK_MIXED, budget_bytes=8 << 30, max_inflight=4)
    ids = []
    for i, (st, k) in enumerate(reqs):
        ids.append(srv.submit(st, k, np.full(len(k), 1.0 / len(k))))
        if i == 7:
            time.sleep(0.05)          # the second half arrives while the first runs
    res = [srv.result(r) for r in ids]
    stats = srv.stats()
    srv.close()
    print(f"stats {stats}; latencies ms {[round(r['latency_ms'], 1) for r in res]}", flush=True)
    assert stats["completed"] == len(reqs)
    assert stats["max_group"] >= 2               # requests were batched together
    for (st, k), r in zip(reqs, res):
        assert r["status"] == 0
        w = np.full(len(k), 1.0 / len(k))
        p = Plan(st, k, w, precision=_lib.TBK_MIXED)
        a = p.run(4)
        p.close()
        # the mixed path's own rounding: alone the structure is solved at its
        # order, in the mix at the group's padded order, and the FP32 stages
        # round differently before the FP64 refinement (about 1e-13 relative
        # on the cell's energy)
        d = max(abs(r[x] - a[x]) for x in ("efermi", "earg1", "earg2"))
        assert d < 1e-11, (st.nv, d)
        ref = oracle.Setup(st).bande(k, w, mode=4, nthreads=16)
        assert abs(r["efermi"] - ref["efermi"]) < 1e-10
        assert abs(r["earg1"] - ref["earg1"]) / st.natoms < 1e-10


def test_server_failing_request_does_not_affect_others():
    p_bad = S.synthetic_params(1, overlap_scale=3.0)
    bad = S.synthetic_structure("fcc", (1, 1, 1), 3.3 * S.ANGSTROM, p_bad)
    good = S.config_c1()
    k = S.monkhorst_pack((2, 2, 2))
    w = np.full(len(k), 1.0 / len(k))
    srv = Server(precision=_lib.TBK_FP64, budget_bytes=4 << 30)
    ib = srv.submit(bad, k, w)
    ig = srv.submit(good, k, w)
    rb, rg = srv.result(ib), srv.result(ig)
    srv.close()
    assert rb["status"] == _lib.TBK_ERR_OVERLAP
    assert rg["status"] == 0
    p = Plan(good, k, w)
    a = p.run(4)
    p.close()
    assert abs(rg["efermi"] - a["efermi"]) < 1e-12


def test_server_fp64_matches_oracle_and_reuses_nothing_stale():
    """FP64 requests of repeated and different structures, submitted in two
    waves through one server: every result equals the oracle's, and a
    structure's result is the same bits in both waves although the groups it
    joins differ (each group is solved at the padded order of its key)."""
    c4 = S.config_c4(count=32, max_atoms=8, max_q=3)
    reqs = [(st, S.monkhorst_pack((2, 2, 2))) for st, _ in c4[:6]] * 2
    srv = Server(precision=_lib.TBK_FP64, budget_bytes=2 << 30, max_inflight=2)
    ids = [srv.submit(st, k, np.full(len(k), 1.0 / len(k))) for st, k in reqs[:6]]
    res = [srv.result(r) for r in ids]
    ids2 = [srv.submit(st, k, np.full(len(k), 1.0 / len(k))) for st, k in reqs[6:]]
    res += [srv.result(r) for r in ids2]
    srv.close()
    for (st, k), r in zip(reqs, res):
        assert r["status"] == 0
        ref = oracle.Setup(st).bande(k, np.full(len(k), 1.0 / len(k)), mode=4, nthreads=16)
        assert abs(r["efermi"] - ref["efermi"]) < 1e-9
        assert abs(r["earg1"] - ref["earg1"]) / st.natoms < 1e-9
        assert abs(r["earg2"] - ref["earg2"]) * S.RYBOHR3_KBAR < 1e-6
    for a, b in zip(res[:6], res[6:]):
        assert a["efermi"] == b["efermi"] and a["earg2"] == b["earg2"]


def test_server_timing_and_admission_limits():
    """One slot and a budget no request fits: groups run one at a time
    (each admitted only when nothing runs), results agree with a plan's (the
    server solves at the padded order, a plan at nv: the rounding differs),
    and the timing counters add up: work is the sum of nv^3 nk, the device
    intervals lie inside the submission-to-completion span."""
    c4 = S.config_c4(count=16, max_atoms=6, max_q=3)
    reqs = [(st, S.monkhorst_pack((2, 2, 2))) for st, _ in c4[:5]]
    srv = Server(precision=_lib.TBK_FP64, budget_bytes=1, max_inflight=1)
    ids = [srv.submit(st, k, np.full(len(k), 1.0 / len(k))) for st, k in reqs]
    res = [srv.result(r) for r in ids]
    stats, tm = srv.stats(), srv.timing()
    srv.close()
    print(f"stats {stats} timing {tm}", flush=True)
    assert stats["completed"] == len(reqs) and stats["groups"] == len(reqs) and stats["max_group"] == 1
    assert tm["work_n3k"] == sum(float(st.nv) ** 3 * len(k) for st, k in reqs)
    assert 0.0 < tm["busy_ms"] <= tm["span_ms"]
    assert tm["setup_threads"] >= 2 and tm["setup_ms"] > 0.0
    for (st, k), r in zip(reqs, res):
        w = np.full(len(k), 1.0 / len(k))
        p = Plan(st, k, w)
        a = p.run(4)
        p.close()
        assert r["status"] == 0
        assert abs(r["efermi"] - a["efermi"]) < 1e-12
        assert abs(r["earg1"] - a["earg1"]) / st.natoms < 1e-12
        assert abs(r["earg2"] - a["earg2"]) * S.RYBOHR3_KBAR < 1e-8
