"""The irreducible mesh of tbk.kmesh feeding a device plan: the ex1 Pd
structure over kptgen's special points and weights gives, on the GPU, the
sums of the full Monkhorst-Pack mesh (also on the GPU) and of the oracle over
the same irreducible list."""
import json
import os

import numpy as np
import pytest

import kmesh_cases as KC
import oracle
from tbk import kmesh
from tbk import structures as S
from tbk.api import Plan

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden", "ex1.json")


@pytest.mark.parametrize("q", [4, 6])
def test_irreducible_mesh_on_device(q):
    gold = json.load(open(GOLD))
    p = S.NRLParams(1, [9], [gold["valence"]], [gold["rcut"]], [gold["screenl"]],
                    np.array(gold["param"]), 0, True)
    lat = S.fcc_primitive(6.9)
    st = S.Structure(lat, np.zeros((1, 3)), [1], p, gold["nd"])
    L = kmesh.reciprocal_lattice(lat, 1)
    g = kmesh.space_group(L, KC.oh_cartesian(), np.zeros((48, 3)), ilat=False)
    kq, kw = kmesh.kptgen(L, g, (q, q, q), lspec=True)
    full = S.monkhorst_pack((q, q, q))
    print(f"case: fcc Pd a=6.9, MP {q}^3: {len(kq)} irreducible of {len(full)}", flush=True)
    res = {}
    for name, kp, w in (("ibz", kq, kw), ("full", full, np.full(len(full), 1.0 / len(full)))):
        plan = Plan(st, kp, w)
        res[name] = plan.run(4, gold["tkb"], 10.0)
        plan.close()
    ref = oracle.Setup(st).bande(kq, kw, mode=4, tkb=gold["tkb"], eigcut=10.0, nthreads=4)
    for other in (res["full"], ref):
        assert abs(res["ibz"]["efermi"] - other["efermi"]) < 1e-11
        assert abs(res["ibz"]["earg1"] - other["earg1"]) < 1e-11
        assert abs(res["ibz"]["earg2"] - other["earg2"]) * S.RYBOHR3_KBAR < 1e-6
