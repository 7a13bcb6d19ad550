"""Occupation window: with every k-point in one batch, pressure runs form
eigenvectors and first-order derivatives only for the states below
E_F + 20 kT that the Fermi sums use (bande.f:1170-1321 skips all others).
E_F, E, P, the eigenvalues and every derivative below the window must be
bit-identical to the run that forms all eigenvectors."""
import numpy as np
import pytest

from tbk import _lib
from tbk import structures as S
from tbk.api import GroupPlan, Plan

pytestmark = pytest.mark.gpu


def run_both(make, mode):
    out = []
    for on in (False, True):
        p = make()
        p.set_window(on)
        r = p.run(mode, 0.005, 10.0)
        out.append((p, r))
    return out


@pytest.mark.parametrize("prec", [_lib.TBK_FP64, _lib.TBK_MIXED])
@pytest.mark.parametrize("mode", [2, 4])
def test_window_bit_identical_c2(prec, mode):
    st = S.config_c2()
    k = S.monkhorst_pack((8, 8, 8))[::8]
    w = np.full(len(k), 1.0 / len(k))
    print(f"window: C2 nk={len(k)} precision {prec} mode {mode}", flush=True)
    (p0, r0), (p1, r1) = run_both(lambda: Plan(st, k, w, precision=prec), mode)
    e0, d0 = p0.eigenvalues()
    e1, d1 = p1.eigenvalues()
    assert r0 == r1, (r0, r1)
    assert np.array_equal(e0, e1)
    ef = r1["efermi"]
    below = e0 < ef + 20 * 0.005
    assert np.array_equal(d0[below], d1[below])
    # the window keeps the occupied states and their thermal tail only
    kept = int(below.sum(axis=1).max())
    print(f"  states below E_F + 20 kT: <= {kept} of {st.nv}", flush=True)
    assert kept < st.nv
    # vectors: identical below the window's cut, zero above it
    z0, z1 = p0.eigenvectors(len(k) - 1), p1.eigenvectors(len(k) - 1)
    nz = np.flatnonzero(np.abs(z1).max(axis=0) > 0)
    c = int(nz.max()) + 1
    assert np.sum(below[-1]) <= c < st.nv
    assert np.array_equal(z0[:, :c], z1[:, :c]) and not np.any(z1[:, c:])
    p0.close(); p1.close()


@pytest.mark.parametrize("prec", [_lib.TBK_FP64, _lib.TBK_MIXED])
def test_window_group_plan_and_batches(prec):
    c4 = S.config_c4(count=256)
    sel = []
    for n in (18, 36, 54):
        for st, q in c4:
            if st.nv == n:
                sel.append(st)
                break
    ks = [S.monkhorst_pack((3, 3, 3)) for _ in sel]
    ws = [np.full(len(k), 1.0 / len(k)) for k in ks]
    (g0, r0), (g1, r1) = run_both(lambda: GroupPlan(sel, ks, ws, precision=prec), 4)
    for i in range(len(sel)):
        assert r0[i] == r1[i], (i, r0[i], r1[i])
        assert np.array_equal(g0.eigenvalues(i)[0], g1.eigenvalues(i)[0])
    g0.close(); g1.close()
    # several batches: the window stays off and results match the one-batch run
    st, k = sel[1], ks[1]
    w = np.full(len(k), 1.0 / len(k))
    pb = Plan(st, k, w, batch=5, precision=prec)
    rb = pb.run(4, 0.005, 10.0)
    pb.close()
    tol = 1e-12 if prec == _lib.TBK_FP64 else 1e-10
    assert max(abs(rb[q] - r1[1][q]) for q in rb) < tol


@pytest.mark.parametrize("prec", [_lib.TBK_FP64, _lib.TBK_MIXED])
@pytest.mark.parametrize("mode", [1, 3])
def test_energy_modes_form_no_vectors(prec, mode):
    st = S.config_c1()
    k = S.monkhorst_pack((4, 4, 4))
    w = np.full(len(k), 1.0 / len(k))
    (p0, r0), (p1, r1) = run_both(lambda: Plan(st, k, w, precision=prec), mode)
    assert r0 == r1
    assert np.array_equal(p0.eigenvalues()[0], p1.eigenvalues()[0])
    assert not np.any(p1.eigenvectors(len(k) - 1))
    z0 = p0.eigenvectors(len(k) - 1)
    assert np.abs(z0).max() > 0
    p0.close(); p1.close()
