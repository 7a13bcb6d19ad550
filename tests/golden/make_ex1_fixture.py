"""Writes tests/golden/ex1.json from the reference's own ex1 example data.

Data only: the Pd parameter values (ex1/pd_par), the two k-point lists
(ex1/fcc.08, ex1/bcc.08), the 18 structures of ex1/SKIN, and the published
outputs E_F, E(T), P(T), pair counts and minimum distances of ex1/SKOUT.
Run in a tree where the reference is mounted (python tests/golden/make_ex1_fixture.py).
"""
import json
import os
import re
import sys

REF = os.environ.get("NRL_REF", "/root/reference/NRLtb/ex1")


def main():
    lines = open(os.path.join(REF, "pd_par")).read().splitlines()
    par = [float(l.split()[0]) for l in lines[7:7 + 97]]
    kl = {}
    for nm in ("fcc.08", "bcc.08"):
        rows = open(os.path.join(REF, nm)).read().splitlines()
        n = int(rows[0].split()[0])
        kl[nm] = [[float(x) for x in r.split()[:4]] for r in rows[1:1 + n]]
    skout = open(os.path.join(REF, "SKOUT")).read()
    blocks = skout.split("structure ")[1:]
    structs = []
    for b in blocks:
        lat = "fcc" if "fcc lattice" in b else "bcc"
        a = float(re.search(r"a =\s+([0-9.]+)", b).group(1))
        pairs = int(re.search(r"Number of pairs =\s+(\d+)", b).group(1))
        dmin = float(re.search(r"Minimum distance this configuration =\s+([0-9.Ee+-]+)", b).group(1))
        ef = float(re.search(r"EFermi =\s+([0-9.Ee+-]+)", b).group(1))
        et = float(re.search(r"E\(T\) =\s+([0-9.Ee+-]+)", b).group(1))
        pt = float(re.search(r"P\(T\) =\s+([0-9.Ee+-]+)", b).group(1))
        vol = float(re.search(r"Volume=\s+([0-9.Ee+-]+)", b).group(1))
        structs.append(dict(lattice=lat, a=a, pairs=pairs, dismin=dmin, efermi=ef, et=et, pt=pt,
                            volume=vol))
    out = dict(source="NRLtb/ex1 (pd_par, fcc.08, bcc.08, SKIN, SKOUT)", mode=4, tkb=0.005,
               eigcut=0.5, rcut=10.5, screenl=0.5, valence=10.0, nd=[4, 4, 4], param=par,
               kpoints=kl, structures=structs)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ex1.json")
    json.dump(out, open(path, "w"), indent=1)
    print(f"wrote {path}: {len(structs)} structures")


if __name__ == "__main__":
    sys.exit(main())
