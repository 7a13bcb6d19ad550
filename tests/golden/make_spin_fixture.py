"""Writes tests/golden/spin.json: the reference's own bande with jspins = 2
(collinear spin polarisation, and non-collinear with lnoncol) on small
structures, for the device path's parity tests where the reference is not
present.

Data only: inputs (structure, the spin-down parameter channel, flipped
atoms, k-points) and the reference's E_F, earg1, earg2 and smag per mode,
from oracle/_ref/libnrlref.so (the reference's sources compiled where they
lie, NRLtb/ex1's P1 with mspin = 2) through its bande, on the CPU.  The
spin-up channel is the Pd set of tight-binding_amd/tbk/data/pd_par.json.
Also writes tests/golden/spin_hs.npz: the reference's spinor H(k), S(k) and
their volume derivatives (spinham + setham/sethamv + addspins) of the 4-atom
non-collinear case at one general k-point, packed upper triangles.
Run in a tree where the reference is built:
    make -C oracle ref && python tests/golden/make_spin_fixture.py
"""
import json
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tight-binding_amd"))
sys.path.insert(0, os.path.dirname(HERE))
import oracle  # noqa: E402
import spin_case  # noqa: E402
from tbk import structures as S  # noqa: E402


def spin_down_channel(param, seed, n_changed):
    """The spin-up values with the first n_changed (decay, on-site and, for
    57, the bond integrals) scaled by 1 + 0.02 N(0,1); the overlap block is
    kept, so S(k) stays positive definite."""
    rng = np.random.default_rng(seed)
    pdn = np.array(param, float)
    pdn[:n_changed] = pdn[:n_changed] * (1.0 + 0.02 * rng.standard_normal(n_changed))
    return pdn


def cases():
    p = S.pd_like_params()
    out = []
    for name, cell, a, seed, nch, flip, q in (
            ("bcc2_ferro", "bcc", 6.0, 3, 57, [0, 0], 4),
            ("bcc2_flip", "bcc", 6.0, 3, 57, [1, 0], 4),
            ("fcc4_flip", "fcc", 7.2, 5, 57, [0, 1, 1, 0], 3)):
        lat, frac = S.conventional_cell(cell, a)
        st = S.Structure(lat, frac @ lat, np.ones(len(frac), np.int32), p, nd=(4, 4, 4))
        k = S.monkhorst_pack((q, q, q))
        w = np.full(len(k), 1.0 / len(k))
        out.append(dict(name=name, cell=cell, a=a, lat=lat.tolist(), posn=(frac @ lat).tolist(),
                        param_dn=spin_down_channel(p.param, seed, nch).tolist(), flipos=flip,
                        kpts=k.tolist(), weights=w.tolist(), st=st))
    return out


def ex1_magnetic():
    """The reference program (oracle/_ref/static_ref) on ex1 with pd_par
    made a two-channel ("NM") file: SKENG rows with the moment column."""
    ref = os.path.join(ROOT, "oracle", "_ref")
    with tempfile.TemporaryDirectory() as tmp:
        work = os.path.join(tmp, "ex1")
        shutil.copytree(os.path.join(ref, "ex1"), work)
        src = open(os.path.join(work, "pd_par")).read()
        pdn = spin_case.spin_down_values(spin_case.read_channel(src))
        os.unlink(os.path.join(work, "pd_par"))
        open(os.path.join(work, "pd_par"), "w").write(spin_case.nm_text(src, pdn))
        subprocess.run("ulimit -s unlimited; " + os.path.join(ref, "static_ref"), shell=True, cwd=work,
                       check=True, capture_output=True)
        rows = spin_case.read_skeng(os.path.join(work, "SKENG"))
    assert len(rows) == 18 and all(len(r) == 6 for r in rows)
    out = dict(param_dn=pdn.tolist(), skeng=[list(r) for r in rows])
    # the same with spin input on every structure (input1.f:518-545)
    for kind in ("noncol", "flip"):
        with tempfile.TemporaryDirectory() as tmp:
            work = os.path.join(tmp, "ex1")
            shutil.copytree(os.path.join(ref, "ex1"), work)
            for f in ("pd_par", "SKIN"):
                os.unlink(os.path.join(work, f))
            open(os.path.join(work, "pd_par"), "w").write(spin_case.nm_text(src, pdn))
            open(os.path.join(work, "SKIN"), "w").write(
                spin_case.skin_spin(open(os.path.join(ref, "ex1", "SKIN")).read(), kind))
            subprocess.run("ulimit -s unlimited; " + os.path.join(ref, "static_ref"), shell=True, cwd=work,
                           check=True, capture_output=True)
            out["skeng_" + kind] = spin_case.read_skeng_rows(os.path.join(work, "SKENG"))
        assert len(out["skeng_" + kind]) == 18
    return out


def noncol_cases():
    """Non-collinear structures (lnoncol): spin directions per atom."""
    out = []
    for c in cases():
        st = c.pop("st")
        nat = len(c["posn"])
        if nat == 2:
            t = 0.7
            spn = [[0.0, 0.0, 1.0], [np.sin(t), 0.0, np.cos(t)]]
        else:
            r3 = 1.0 / np.sqrt(3.0)
            spn = [[0.0, 0.0, 1.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [r3, r3, r3]][:nat]
        c["spinat"] = spn
        c.pop("flipos")
        out.append((c, st))
    return out[0::2]


# a k-point off every symmetry element (the units of the cases' k-points)
GENERAL_K = [0.137, -0.291, 0.413]


def spinor_matrices():
    """The reference's spinor H, S, dH/dV, dS/dV (packed upper triangles of
    order 2 nv) of the 4-atom non-collinear case, whose spins point along
    z, x, y and (1,1,1), at GENERAL_K."""
    c, st = [x for x in noncol_cases() if len(x[0]["posn"]) == 4][0]
    R = oracle.RefSetup(st)
    par = R.setpar_noncol(np.array(c["param_dn"]))
    h, s = R.hs_noncol(GENERAL_K, c["spinat"], par["spinos"], deriv=False)
    dh, ds = R.hs_noncol(GENERAL_K, c["spinat"], par["dspnos"], deriv=True)
    return dict(name=np.array(c["name"]), k=np.array(GENERAL_K), nv2=np.array(2 * st.nv),
                h=h, s=s, dh=dh, ds=ds)


def main():
    res = []
    for c in cases():
        st = c.pop("st")
        R = oracle.RefSetup(st)
        c["modes"] = {}
        for mode in (1, 2, 3, 4):
            r = R.bande_spin(np.array(c["kpts"]), np.array(c["weights"]), np.array(c["param_dn"]),
                             np.array(c["flipos"], np.int32), mode=mode, tkb=0.005, eigcut=10.0)
            assert r.pop("err") == 0
            c["modes"][str(mode)] = r
        res.append(c)
        print(c["name"], c["modes"]["4"])
    nc = []
    for c, st in noncol_cases():
        R = oracle.RefSetup(st)
        c["modes"] = {}
        for mode in (1, 2, 3, 4):
            r = R.bande_noncol(np.array(c["kpts"]), np.array(c["weights"]), np.array(c["param_dn"]),
                               np.array(c["spinat"]), mode=mode, tkb=0.005, eigcut=10.0)
            assert r.pop("err") == 0
            c["modes"][str(mode)] = r
        nc.append(c)
        print("noncol", c["name"], c["modes"]["4"])
    ex1 = ex1_magnetic()
    print("ex1 NM:", ex1["skeng"][0])
    out = dict(source="reference bande (jspins = 2) through oracle/_ref/libnrlref.so; Pd spin-up channel; "
                      "ex1 with a two-channel pd_par through oracle/_ref/static_ref",
               tkb=0.005, eigcut=10.0, nd=[4, 4, 4], cases=res, ex1_magnetic=ex1, noncol=nc)
    path = os.path.join(HERE, "spin.json")
    json.dump(out, open(path, "w"), indent=1)
    print(f"wrote {path}")
    mats = spinor_matrices()
    path = os.path.join(HERE, "spin_hs.npz")
    np.savez_compressed(path, **mats)
    print(f"wrote {path}")


if __name__ == "__main__":
    sys.exit(main())
