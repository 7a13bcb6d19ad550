"""Large order: the 108-atom strained supercell (N = 972), FP64 and mixed
precision against the oracle on a slice of its 4x4x4 mesh."""
import numpy as np
import pytest

import oracle
from tbk import _lib
from tbk import structures as S
from tbk.api import Plan

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c3_ref():
    st = S.config_c3()
    k = S.monkhorst_pack((4, 4, 4))[::16]                     # 4 of the 64 k-points
    w = np.full(len(k), 1.0 / len(k))
    ref = oracle.Setup(st).bande(k, w, mode=4, nthreads=16, want_eig=True)
    return st, k, w, ref


@pytest.mark.parametrize("prec", [_lib.TBK_FP64, _lib.TBK_MIXED])
def test_c3_slice_vs_oracle(c3_ref, prec):
    st, k, w, ref = c3_ref
    print(f"C3 slice: N={st.nv} nk={len(k)} precision={prec}", flush=True)
    p = Plan(st, k, w, precision=prec)
    r = p.run(4)
    e, _ = p.eigenvalues()
    p.close()
    de = float(np.abs(e - ref["eig"]).max())
    def_ = abs(r["efermi"] - ref["efermi"])
    dE = abs(r["earg1"] - ref["earg1"]) / st.natoms
    dP = abs(r["earg2"] - ref["earg2"]) * S.RYBOHR3_KBAR
    print(f"  max|de|={de:.2e} dEF={def_:.2e} dE/atom={dE:.2e} dP={dP:.2e} kbar", flush=True)
    if prec == _lib.TBK_FP64:
        assert de < 1e-10 and def_ < 1e-9 and dE < 1e-9 and dP < 1e-6
    else:
        # about ten times the errors measured (4.6e-11, 3.4e-13, 1.6e-11 Ry, 1.6e-7 kbar)
        assert de < 5e-10 and def_ < 4e-12 and dE < 2e-10 and dP < 2e-6
