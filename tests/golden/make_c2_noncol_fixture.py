"""Writes tests/golden/c2_noncol.json: the reference's own lnoncol bande
(mode 4) on the C2 structure (32-atom fcc 2x2x2 supercell, seeded
displacements, spinor order 2 x 288 = 576) over the 2x2x2 Monkhorst-Pack
mesh, with a seeded spin-down channel and a spin direction per atom that
has x, y and z components.

Data only: inputs (seeds, spin directions, k-points, weights) and the
reference's E_F, E(T), P(T), from the reference's sources compiled with
matom = 128 (oracle/_ref/big/libnrlref_big.so), on the CPU.  Run in a tree
where that build is present:
    make -C oracle refbig && python tests/golden/make_c2_noncol_fixture.py
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tight-binding_amd"))
sys.path.insert(0, os.path.dirname(HERE))
import oracle  # noqa: E402
import spin_case  # noqa: E402
from tbk import structures as S  # noqa: E402

SEED = 20261203          # the C2 structure's displacements
DN_SEED, DN_CHANGED = 11, 57
MODE, TKB, EIGCUT = 4, 0.005, 10.0


def spin_directions(natoms):
    """Atom t's spin at polar angle 15 + 7t and azimuth 20 + 13t degrees."""
    t = np.arange(natoms)
    th, ph = np.radians(15.0 + 7.0 * t), np.radians(20.0 + 13.0 * t)
    return np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], axis=1)


def case():
    st = S.config_c2(SEED)
    k = S.monkhorst_pack((2, 2, 2))
    w = np.full(len(k), 1.0 / len(k))
    pdn = spin_case.spin_down_values(st.params.param, seed=DN_SEED, n_changed=DN_CHANGED)
    return st, k, w, pdn, spin_directions(st.natoms)


def main():
    oracle.use_big_reference()
    st, k, w, pdn, spins = case()
    t0 = time.time()
    r = oracle.RefSetup(st).bande_noncol(k, w, pdn, spins, mode=MODE, tkb=TKB, eigcut=EIGCUT)
    assert r.pop("err") == 0
    print(f"C2 lnoncol, {len(k)} k at order {2 * st.nv}: {r} ({time.time() - t0:.1f} s)")
    out = dict(source="reference bande (lnoncol) through oracle/_ref/big/libnrlref_big.so",
               seed=SEED, dn_seed=DN_SEED, dn_changed=DN_CHANGED, mode=MODE, tkb=TKB, eigcut=EIGCUT,
               spinat=spins.tolist(), kpts=np.asarray(k).tolist(), weights=w.tolist(), result=r)
    path = os.path.join(HERE, "c2_noncol.json")
    json.dump(out, open(path, "w"), indent=1)
    print(f"wrote {path}")


if __name__ == "__main__":
    main()
