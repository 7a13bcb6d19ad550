"""The resident tridiagonalisation (unblocked Householder steps with the
item's matrix held in registers, one workgroup per k-point) against the
one-stage Householder panels and against LAPACK: the mixed method's FP32
form up to order 288 (every tile configuration: tile dimensions 8, 12, 16
and 18), the FP64 form up to order 160 (8 and 10), orders that are not multiples of
16, and the tridiagonal itself (its spectrum is the reduced matrix's)."""
import ctypes as C

import numpy as np
import pytest
import scipy.linalg

from tbk import _lib
from tbk import structures as S
from tbk.api import Plan

pytestmark = pytest.mark.gpu


def _solve(st, k, w, prec, stages):
    p = Plan(st, k, w, precision=prec)
    p.set_tridiag(stages)
    r = p.run(4)
    e, d = p.eigenvalues()
    p.close()
    return r, e, d


@pytest.mark.parametrize("prec", [_lib.TBK_FP64, _lib.TBK_MIXED])
def test_resident_same_as_panels(prec):
    """C2 (N = 288, the mixed form's largest tile layout) in the mixed
    method and a 16-atom cell (N = 144) in both, on a slice of their
    meshes, mode 4: E_F, E, P, every eigenvalue and derivative of the
    resident reduction agree with the one-stage panels' to rounding (FP64)
    or within the refinement's rounding (mixed)."""
    cells = [S.config_c2()] if prec == _lib.TBK_MIXED else []
    cells.append(S.synthetic_structure("fcc", (2, 2, 1), 3.61 * S.ANGSTROM, S.synthetic_params(7),
                                       sigma_bohr=0.05 * S.ANGSTROM,
                                       seed=11))
    for st in cells:
        print(f"case: resident vs panels, N={st.nv}, precision {prec}", flush=True)
        k = S.monkhorst_pack((4, 4, 4))[::2]
        w = np.full(len(k), 1.0 / len(k))
        (r1, e1, d1), (r3, e3, d3) = _solve(st, k, w, prec, 1), _solve(st, k, w, prec, 3)
        de, dd = float(np.abs(e1 - e3).max()), float(np.abs(d1 - d3).max())
        dp = abs(r1["earg2"] - r3["earg2"]) * S.RYBOHR3_KBAR
        print(f"  dE_F={abs(r1['efermi'] - r3['efermi']):.2e} dE/atom={abs(r1['earg1'] - r3['earg1']) / st.natoms:.2e} "
              f"dP={dp:.2e} kbar max|de|={de:.2e} max|ddesk|={dd:.2e}", flush=True)
        tol = 1e-11 if prec == _lib.TBK_FP64 else 2e-12
        assert de < tol
        assert abs(r1["efermi"] - r3["efermi"]) < tol
        assert abs(r1["earg1"] - r3["earg1"]) / st.natoms < tol
        assert dp < (1e-8 if prec == _lib.TBK_FP64 else 1e-6)


@pytest.mark.parametrize("prec", [_lib.TBK_FP64, _lib.TBK_MIXED])
def test_resident_odd_orders(prec):
    """Structures of the uneven pool whose orders are not multiples of 16
    (partial last tiles) and that span the tile layouts of each precision:
    resident against one-stage, mode 4 on a 3^3 mesh."""
    pool = S.config_c4(count=256, max_atoms=32, max_q=3)
    want = [9, 45, 99, 153, 207, 261] if prec == _lib.TBK_MIXED else [9, 45, 99, 135, 153]
    top = 288 if prec == _lib.TBK_MIXED else 160
    picks = []
    for nv in want:
        best = min((x[0] for x in pool if x[0].nv <= top), key=lambda s: abs(s.nv - nv))
        if best.nv not in [p.nv for p in picks]:
            picks.append(best)
    k = S.monkhorst_pack((3, 3, 3))
    w = np.full(len(k), 1.0 / len(k))
    for st in picks:
        print(f"case: resident vs panels, N={st.nv}, precision {prec}", flush=True)
        (r1, e1, d1), (r3, e3, d3) = _solve(st, k, w, prec, 1), _solve(st, k, w, prec, 3)
        de, dd = float(np.abs(e1 - e3).max()), float(np.abs(d1 - d3).max())
        dp = abs(r1["earg2"] - r3["earg2"]) * S.RYBOHR3_KBAR
        print(f"  max|de|={de:.2e} max|ddesk|={dd:.2e} dE={abs(r1['earg1'] - r3['earg1']):.2e} dP={dp:.2e} kbar",
              flush=True)
        if prec == _lib.TBK_FP64:
            assert de < 1e-11 and dd < 1e-9
            assert abs(r1["efermi"] - r3["efermi"]) < 1e-11
        else:
            assert de < 1e-9
            assert abs(r1["efermi"] - r3["efermi"]) < 1e-11
            assert abs(r1["earg1"] - r3["earg1"]) / st.natoms < 1e-10


def test_resident_refused_beyond_its_orders():
    """The resident form is refused above order 160 in FP64 (order 288 in the
    mixed method)."""
    st = S.config_c2()
    k = S.monkhorst_pack((2, 2, 2))
    p = Plan(st, k, np.full(len(k), 1.0 / len(k)), precision=_lib.TBK_FP64)
    with pytest.raises(_lib.TbkError) as ei:
        p.set_tridiag(3)
    assert ei.value.code == _lib.TBK_ERR_ARG
    p.close()


@pytest.mark.parametrize("n", [2, 17, 64, 130, 160])
def test_resident_tridiagonal_has_the_reduced_spectrum(n):
    """FP64, through the stage-by-stage diagnostic: the tridiagonal (d, e)
    the resident reduction leaves has the eigenvalues of the reduced matrix
    C = U^-H H U^-1 it started from (LAPACK on the CPU), and the reflectors
    it stores are unit below the diagonal (v_{j+1} = 1, zeros above)."""
    print(f"case: resident tridiagonal spectrum n={n}", flush=True)
    rng = np.random.default_rng(500 + n)
    a = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    h = np.ascontiguousarray((a + a.conj().T).T)
    b = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    s = np.ascontiguousarray((b @ b.conj().T / n + np.eye(n)).T)
    bufs = {k: np.zeros((n, n), dtype=complex) for k in ("uinv", "c", "refl", "z")}
    d, e, wv = np.zeros(n), np.zeros(n), np.zeros(n)
    vt = np.zeros((n, n))
    ptr = lambda x: x.ctypes.data_as(C.c_void_p)
    rc = _lib.lib().tbk_debug_stages(0, n, ptr(h), ptr(s), ptr(bufs["uinv"]), ptr(bufs["c"]), ptr(bufs["refl"]),
                                     ptr(d), ptr(e), ptr(wv), ptr(vt), ptr(bufs["z"]))
    assert rc == 0
    cm = bufs["c"].T                         # column-major on the device
    cm = np.tril(cm) + np.tril(cm, -1).conj().T
    lam = scipy.linalg.eigvalsh(cm)
    lt = scipy.linalg.eigvalsh_tridiagonal(d, e[:-1]) if n > 1 else d
    scale = max(1.0, np.abs(lam).max())
    dl = float(np.abs(np.sort(lt) - lam).max())
    print(f"  max|eig(T) - eig(C)|={dl:.2e} (scale {scale:.1f})", flush=True)
    assert dl < 1e-12 * scale
    refl = bufs["refl"].T
    for j in range(n - 1):
        assert np.all(refl[: j + 1, j] == 0)
        assert refl[j + 1, j] == 1
