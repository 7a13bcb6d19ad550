"""The two-stage tridiagonalisation (panel QR + GEMMs to a band of width 16,
bulge chasing to tridiagonal, back-transform through both stages) against
the one-stage Householder panels and against the oracle's eigensolver, at
orders covering every kernel configuration: the panel QR's register layouts
(panel rows up to 128, 256, 512, 1024 and beyond), the band held in LDS and
in global memory, and orders that are not multiples of 16."""
import numpy as np
import pytest
import scipy.linalg

import oracle
from tbk import _lib
from tbk import structures as S
from tbk.api import Plan, zhegv

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [577, 1100])
def test_two_stage_zhegv_random(n):
    """Random Hermitian pencils through the one-shot eigensolver (two-stage
    from order 512): eigenvalues against scipy's, residual and
    S-orthogonality of the vectors."""
    print(f"case: zhegv two-stage n={n}", flush=True)
    rng = np.random.default_rng(1000 + n)
    a = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    h = a + a.conj().T
    b = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    s = b @ b.conj().T / n + np.eye(n)
    w, z = zhegv(h, s)
    wo = scipy.linalg.eigh(h, s, eigvals_only=True)
    scale = max(1.0, np.abs(wo).max())
    de = float(np.abs(w - wo).max())
    orth = float(np.abs(z.conj().T @ s @ z - np.eye(n)).max())
    res = float(np.abs(h @ z - s @ z * w).max())
    print(f"  |dw|={de:.2e} (scale {scale:.1f}) |Z^H S Z - I|={orth:.2e} |HZ-SZW|={res:.2e}", flush=True)
    assert de < 1e-12 * scale
    assert orth < 1e-11
    assert res < 1e-11 * scale


@pytest.mark.parametrize("prec", [_lib.TBK_FP64, _lib.TBK_MIXED])
def test_two_stage_same_as_one_stage_c2(prec):
    """C2 (N = 288) on 64 of its k-points, mode 4: E_F, E, P and every
    eigenvalue and derivative of the two-stage solve agree with the
    one-stage panels' to rounding."""
    print(f"case: C2 two-stage vs one-stage, precision {prec}", flush=True)
    st = S.config_c2()
    k = S.monkhorst_pack((8, 8, 8))[::8]
    w = np.full(len(k), 1.0 / len(k))
    out = {}
    for stages in (1, 2):
        p = Plan(st, k, w, precision=prec)
        p.set_tridiag(stages)
        r = p.run(4)
        e, d = p.eigenvalues()
        p.close()
        out[stages] = (r, e, d)
    (r1, e1, d1), (r2, e2, d2) = out[1], out[2]
    de = float(np.abs(e1 - e2).max())
    dd = float(np.abs(d1 - d2).max())
    print(f"  dE_F={abs(r1['efermi'] - r2['efermi']):.2e} dE={abs(r1['earg1'] - r2['earg1']):.2e} "
          f"dP={abs(r1['earg2'] - r2['earg2']):.2e} max|de|={de:.2e} max|ddesk|={dd:.2e}", flush=True)
    tol = 1e-11 if prec == _lib.TBK_FP64 else 2e-12
    assert de < tol
    assert abs(r1["efermi"] - r2["efermi"]) < tol
    assert abs(r1["earg1"] - r2["earg1"]) / st.natoms < tol
    assert abs(r1["earg2"] - r2["earg2"]) * S.RYBOHR3_KBAR < (1e-8 if prec == _lib.TBK_FP64 else 1e-7)


def test_two_stage_odd_orders_same_as_one_stage():
    """Structures of the uneven pool whose orders are not multiples of 16
    (the last panel and the last sweeps partial) and span the panel QR's
    register layouts and the band in LDS or global memory: two-stage forced
    against one-stage, FP64, mode 4 on a 3^3 mesh."""
    pool = S.config_c4(count=256, max_atoms=64, max_q=3)
    want = [81, 153, 261, 342, 567]
    picks = []
    for nv in want:
        best = min(pool, key=lambda x: abs(x[0].nv - nv))[0]
        if best.nv not in [p.nv for p in picks]:
            picks.append(best)
    k = S.monkhorst_pack((3, 3, 3))
    w = np.full(len(k), 1.0 / len(k))
    for st in picks:
        print(f"case: two-stage vs one-stage, N={st.nv}", flush=True)
        out = []
        for stages in (1, 2):
            p = Plan(st, k, w)
            p.set_tridiag(stages)
            r = p.run(4)
            e, d = p.eigenvalues()
            p.close()
            out.append((r, e, d))
        (r1, e1, d1), (r2, e2, d2) = out
        de, dd = float(np.abs(e1 - e2).max()), float(np.abs(d1 - d2).max())
        print(f"  max|de|={de:.2e} max|ddesk|={dd:.2e} dE={abs(r1['earg1'] - r2['earg1']):.2e} "
              f"dP={abs(r1['earg2'] - r2['earg2']) * S.RYBOHR3_KBAR:.2e} kbar", flush=True)
        assert de < 1e-11 and dd < 1e-9
        assert abs(r1["efermi"] - r2["efermi"]) < 1e-11
        assert abs(r1["earg1"] - r2["earg1"]) / st.natoms < 1e-11
        assert abs(r1["earg2"] - r2["earg2"]) * S.RYBOHR3_KBAR < 1e-7


def test_two_stage_small_cell_against_oracle():
    """An 8-atom fcc cell (N = 72, two-stage forced) against the oracle's
    bande over a 6^3 mesh, mode 4; a plan below order 40 refuses the
    two-stage reduction."""
    print("case: 8-atom fcc (N=72) two-stage vs oracle", flush=True)
    st = S.synthetic_structure("fcc", (2, 1, 1), 3.61 * S.ANGSTROM, S.synthetic_params(20261203))
    k = S.monkhorst_pack((6, 6, 6))
    w = np.full(len(k), 1.0 / len(k))
    p = Plan(st, k, w)
    p.set_tridiag(2)
    r = p.run(4)
    p.close()
    ref = oracle.Setup(st).bande(k, w, mode=4, nthreads=16)
    print(f"  {r} vs {ref}", flush=True)
    assert abs(r["efermi"] - ref["efermi"]) < 1e-12
    assert abs(r["earg1"] - ref["earg1"]) / st.natoms < 1e-12
    assert abs(r["earg2"] - ref["earg2"]) * S.RYBOHR3_KBAR < 1e-8
    c1 = S.config_c1()
    p = Plan(c1, k[:2], w[:2])
    with pytest.raises(_lib.TbkError):
        p.set_tridiag(2)
    p.close()
