"""Writes tests/golden/kmesh.json: k-points, weights, reciprocal lattice and
space-group operations as the reference's own kptin/rdspgp/kptgen produce
them (oracle/_ref/libnrlref.so, built by `make -C oracle ref` where the
reference tree is present) for the cases of tests/kmesh_cases.py.

Run from the repository root: python tests/golden/make_kmesh_fixture.py
"""
import json
import os
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import kmesh_cases as KC  # noqa: E402
import oracle  # noqa: E402


def main():
    out = {"source": "reference kptin/rdspgp/kptgen (oracle/_ref) on tests/kmesh_cases.py", "cases": {}}
    for name, lat, lattic, ilat, rot, tnp, shift, kdiv, lspec, kpst, kpend, info, sfk in KC.cases():
        with tempfile.TemporaryDirectory() as d:
            g = os.path.join(d, "g.grp")
            KC.write_group(g, name, rot, tnp, ilat, shift or (0.0, 0.0, 0.0))
            r = oracle.ref_kptin(lat, KC.newsym_block(g, lattic, ilat, kdiv, lspec, kpst, kpend, info, sfk), d)
        out["cases"][name] = {"q": [float(x) for x in r["q"].ravel()], "w": [float(x) for x in r["w"]],
                              "rspct": [float(x) for x in r["rspct"].ravel()],
                              "tnpspc": [float(x) for x in r["tnpspc"].ravel()],
                              "bvec": [float(x) for x in r["bvec"].ravel()], "wsvol": float(r["wsvol"])}
    path = os.path.join(HERE, "kmesh.json")
    with open(path, "w") as f:
        json.dump(out, f)
    print(f"wrote {path}: {len(out['cases'])} cases")


if __name__ == "__main__":
    sys.exit(main())
