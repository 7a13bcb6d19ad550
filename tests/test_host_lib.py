"""CPU checks of lib/libtbk.so: it loads, exports every symbol of
include/tbk.h, and its host set-up (neighbour list, SK parameters) is
bit-identical to the oracle's (itself bit-identical to the reference)."""
import os
import re

import numpy as np
import pytest

import oracle
from tbk import _lib
from tbk import structures as S
from tbk.api import HostSetup

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "tbk.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tbk_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = declared_functions()
    assert len(names) >= 15
    for nm in names:
        assert hasattr(L, nm), nm
    assert set(names) == set(_lib.EXPORTS)


def test_error_strings():
    L = _lib.lib()
    for code in (0, 1, 2, 3, 4, 5, 6, 10, 11, 12):
        assert L.tbk_error_string(code)


@pytest.mark.parametrize("cfg", ["c1", "c2", "ex1", "hcp", "large"])
def test_host_setup_bit_exact(cfg):
    """Neighbour list and parameters bit-identical to the oracle's serial
    restatement; "large" (128 atoms, 16,712 pairs) runs the set-up's
    per-atom and per-pair-chunk work on several host threads."""
    if cfg == "large":
        pool = S.config_c4(seed=20261203, count=256, max_atoms=128, max_q=6, nk_budget=9 * 6 ** 3)
        st = max((p[0] for p in pool), key=lambda s: s.nv)
    elif cfg == "c1":
        st = S.config_c1()
    elif cfg == "c2":
        st = S.config_c2()
    elif cfg == "ex1":
        p = S.pd_like_params(10.5, 0.5)
        st = S.Structure(S.fcc_primitive(6.9), np.zeros((1, 3)), [1], p, (4, 4, 4))
    else:
        st = S.synthetic_structure("hcp", (2, 2, 1), 3.61 * S.ANGSTROM * 0.71,
                                   S.synthetic_params(4), sigma_bohr=0.05, seed=9)
    o = oracle.Setup(st)
    h = HostSetup(st)
    n = o.npairs
    assert h.npairs == n and h.dismin == o.dismin
    for nm, m in (("tt", 3 * n), ("dlv", 3 * n), ("dist", n), ("screen", n), ("dscrn", n),
                  ("jatm", 2 * n), ("jkind", 2 * n), ("paros", None), ("dparos", None),
                  ("hop", 14 * n), ("over", 14 * n), ("dhop", 14 * n), ("dover", 14 * n),
                  ("den", None)):
        a, b = getattr(o, nm), getattr(h, nm)
        m = m or len(b)
        assert np.array_equal(a[:m], b[:m]), nm


def test_pair_overflow_is_an_error_code():
    st = S.config_c1()
    with pytest.raises(_lib.TbkError) as ei:
        HostSetup(st, maxpairs=100)
    assert ei.value.code == _lib.TBK_ERR_PAIRS


def test_no_device_is_an_error_code_not_a_crash():
    """On a machine without a GPU plan creation must fail with a code."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from tbk.api import Plan
    with pytest.raises(_lib.TbkError) as ei:
        Plan(S.config_c1(), [[0, 0, 0]], [1.0])
    assert ei.value.code in (_lib.TBK_ERR_NODEVICE, _lib.TBK_ERR_HIP)
