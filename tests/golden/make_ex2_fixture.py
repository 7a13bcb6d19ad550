"""Writes tests/golden/ex2.json: the reference program's own results on its
ex2 example (hcp Ti) with the k-mesh line aligned in its 5I5 fields
(tests/ex2_case.py says why the published input stops in kptgen).

Data only: E_F and E(T) of the 13 structures, as printed to SKOUT with 17
digits, the SKENG rows, and the number of irreducible k-points, from
oracle/_ref/static_ref (the reference's sources compiled where they lie by
oracle/Makefile's dropin target) run on the CPU.
Run in a tree where the reference is mounted:
    make -C oracle dropin && python tests/golden/make_ex2_fixture.py
"""
import json
import os
import pathlib
import re
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(HERE))
import ex2_case  # noqa: E402


def main():
    ref = os.path.join(ROOT, "oracle", "_ref")
    with tempfile.TemporaryDirectory() as tmp:
        for f in ex2_case.INPUTS:
            shutil.copy(os.path.join(ref, "ex2", f), tmp)
        ex2_case.write_skin(pathlib.Path(tmp))
        subprocess.run("ulimit -s unlimited; " + os.path.join(ref, "static_ref"), shell=True, cwd=tmp,
                       check=True, capture_output=True)
        skout = open(os.path.join(tmp, "SKOUT")).read()
        rows = ex2_case.read_skeng(os.path.join(tmp, "SKENG"))
    nk = [int(x) for x in re.findall(r"NUMBER OF K-POINTS=\s+(\d+)", skout)]
    ef = [float(x) for x in re.findall(r"EFermi =\s+([0-9.Ee+-]+)", skout)]
    et = [float(x) for x in re.findall(r"E\(T\) =\s+([0-9.Ee+-]+)", skout)]
    pairs = [int(x) for x in re.findall(r"Number of pairs =\s+(\d+)", skout)]
    assert len(ef) == len(et) == len(rows) == len(pairs) == 13 and nk
    out = dict(source="NRLtb/ex2 (SKIN with the mesh line in 5I5 fields, ti_par, spcgp1.hcp) "
                      "through oracle/_ref/static_ref",
               mode=3, nk=nk, structures=[dict(label=r[0], volume=r[1], efermi=e, et=t, pairs=p,
                                               skeng=[r[2], r[3]])
                                          for r, e, t, p in zip(rows, ef, et, pairs)])
    path = os.path.join(HERE, "ex2.json")
    json.dump(out, open(path, "w"), indent=1)
    print(f"wrote {path}: {len(rows)} structures, {nk} irreducible k-points")


if __name__ == "__main__":
    sys.exit(main())
