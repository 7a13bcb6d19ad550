"""Writes tests/golden/branches.json: the reference's own results for the
branch cases of tests/branch_cases.py (two species, sp and s-only bases,
nltype 1/2/3/90000, an orthogonal basis, a too-large degenerate group,
zfill = 0), for the parity tests where the reference is not present.

Data only, from the reference's sources compiled where they lie
(oracle/_ref/libnrlref.so; oracle/_ref/big/libnrlref_big.so for the
32-atom case), on the CPU:

* c_par: NRLtb/atom_params/C_par read in input.f:181-359's order;
* per case with expect "ok": E_F, earg1, earg2 of the reference's bande in
  modes 1-4 on the case's Monkhorst-Pack mesh (kT = 0.005 Ry, eigcut =
  10 Ry), and at branch_cases.GENERAL_K the packed H(k) (and S(k) for a
  non-orthogonal basis) of its setham and the eigenvalues of its diag (the
  one-orbital case runs first, as in a fresh program);
* per case with expect "degen" / "zfill": the line the reference's program
  prints when it stops (perturb.f:106, fermi.f:165), from a child process.

Run where the reference is built:
    make -C oracle ref refbig && python tests/golden/make_branch_fixture.py
"""
import json
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tight-binding_amd"))
sys.path.insert(0, os.path.dirname(HERE))
import oracle  # noqa: E402
import branch_cases as B  # noqa: E402
from tbk import structures as S  # noqa: E402

C_PAR = "/root/reference/NRLtb/atom_params/C_par"
TKB, EIGCUT = 0.005, 10.0


def c_par_dict():
    p = S.read_nrl_params(C_PAR)
    return dict(kinds=p.kinds, kbas=list(p.kbas), valence=list(p.valence), dnn=list(p.dnn),
                screenl=list(p.screenl), param=np.asarray(p.param, float).tolist(), nltype=p.nltype,
                jsover=bool(p.jsover))


def run_stop_case(name):
    """Child process: the reference's bande on a case it refuses; prints
    what the reference prints and exits the way it exits."""
    case = next(c for c in B.cases(c_par_dict()) if c["name"] == name)
    st = B.structure(case)
    if st.natoms > 20:
        oracle.use_big_reference()
    k, w = B.mesh(case)
    r = oracle.RefSetup(st).bande(k, w, mode=4, tkb=TKB, eigcut=EIGCUT)
    print("returned", r, flush=True)


def main():
    cp = c_par_dict()
    out = dict(c_par=cp, general_k=B.GENERAL_K, tkb=TKB, eigcut=EIGCUT, cases={})
    # A one-orbital cell first: the reference's diag forms no eigenvector for
    # a 1 x 1 pencil (diag.f:320-349), so its perturb reads the eigenvector
    # storage as the previous structure left it -- zero only in a fresh run
    cases = sorted(B.cases(cp), key=lambda c: c["name"] != "h_atom_single")
    for case in cases:
        name = case["name"]
        if case["expect"] != "ok":
            p = subprocess.run([sys.executable, __file__, "--stop", name], capture_output=True, text=True,
                               timeout=600)
            text = (p.stdout + p.stderr).strip()
            want = B.DEGEN_STOP if case["expect"] == "degen" else B.ZFILL_STOP
            if want not in text:
                raise SystemExit(f"{name}: the reference did not stop as expected:\n{text}")
            out["cases"][name] = dict(stop=want, returncode=p.returncode)
            print(f"{name}: stops with '{want}' (exit {p.returncode})", flush=True)
            continue
        st = B.structure(case)
        R = oracle.RefSetup(st)
        k, w = B.mesh(case)
        modes = {}
        for mode in (1, 2, 3, 4):
            r = R.bande(k, w, mode=mode, tkb=TKB, eigcut=EIGCUT)
            if r["err"]:
                raise SystemExit(f"{name} mode {mode}: iberr {r['err']}")
            modes[str(mode)] = [r["efermi"], r["earg1"], r["earg2"]]
        hp, sp = R.hs(B.GENERAL_K)
        info, e, _ = R.diag(hp, sp, mode=3)
        if info:
            raise SystemExit(f"{name}: ioverr {info}")
        out["cases"][name] = dict(npairs=R.npairs, nv=st.nv, modes=modes, eig=e.tolist(),
                                  hp=np.stack([hp.real, hp.imag], 1).tolist())
        if st.params.jsover:       # setham writes no smat otherwise (setham.f:279)
            out["cases"][name]["sp"] = np.stack([sp.real, sp.imag], 1).tolist()
        print(f"{name}: N={st.nv} pairs={R.npairs} mode 4 {modes['4']}", flush=True)
    path = os.path.join(HERE, "branches.json")
    with open(path, "w") as fh:
        json.dump(out, fh, indent=0)
    print(f"wrote {path} ({os.path.getsize(path)} bytes)")


if __name__ == "__main__":
    if len(sys.argv) == 3 and sys.argv[1] == "--stop":
        run_stop_case(sys.argv[2])
    else:
        main()
