"""Writes tests/golden/static.json: the SKENG lines the reference program
itself writes on the end-to-end cases of tests/skin_cases.py (ex1 as
published and in modes 1-3, ex2 with its mesh line aligned and as
published, a SKIN walking every lattice menu and strain option, ex1 with a
two-channel parameter file, spin-flipped and non-collinear, and the QLMT /
QAPW modes 6, 8, 9, 10 on ex1, ex2 and the spin inputs, and the SPNDIR file
of a 20-atom non-collinear cell from the build of oracle/Makefile's spdir
target), for the test of the native program tbk_static on the GPU.

Data only: the SKENG lines, the QLMT or QAPW file text and the program's
exit status from
oracle/_ref/static_ref (the reference's sources compiled where they lie by
oracle/Makefile's dropin target), run on the CPU in a scratch directory.
Run where the reference is built:
    make -C oracle dropin spdir && python tests/golden/make_static_fixture.py
"""
import json
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tight-binding_amd"))
import skin_cases as SC  # noqa: E402


def main():
    ref = os.path.join(ROOT, "oracle", "_ref")
    out = {}
    for name in SC.STATIC_CASES:
        with tempfile.TemporaryDirectory() as tmp:
            SC.build_static_case(name, tmp, os.path.join(ref, "ex1"), os.path.join(ref, "ex2"))
            prog = os.path.join(ref, "spdir", "static_ref_spdir") if name in SC.SPDIR_CASES else \
                os.path.join(ref, "static_ref")
            p = subprocess.run("ulimit -s unlimited; " + prog, shell=True, cwd=tmp,
                               capture_output=True, text=True, timeout=1800)
            sk = os.path.join(tmp, "SKENG")
            lines = [l.rstrip("\n") for l in open(sk)] if os.path.exists(sk) else []
            stops = [l.strip() for l in (p.stdout + p.stderr).splitlines() if "STOP" in l or "stop" in l.lower()]
            qname = SC.QFILE.get(name) or SC.SPDIR_CASES.get(name)
            qtext = open(os.path.join(tmp, qname)).read() if qname else None
        out[name] = dict(skeng=lines, returncode=p.returncode, stop=stops[-1] if stops else "")
        if qtext is not None:
            out[name]["qfile"] = qname
            out[name]["qtext"] = qtext
        print(f"{name}: {len(lines)} SKENG lines, exit {p.returncode} {out[name]['stop']}", flush=True)
    path = os.path.join(HERE, "static.json")
    json.dump(out, open(path, "w"), indent=1)
    print(f"wrote {path}")


if __name__ == "__main__":
    main()
