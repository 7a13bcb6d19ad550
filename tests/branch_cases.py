"""Small structures for the branches of the path that the 4-, 32- and
108-atom workloads do not reach, each named for the reference lines it
drives.

=============  ===============================================================
two_kind       two species: unlike pairs with 14 integrals and the parity
               rearrangement (setham.f:326-376, rotate.f:226-258), cross-kind
               densities (setpar.f:510-536)
two_kind_sp    the same with an sp atom beside an spd atom (kbas 9 and 4):
               unlike blocks of 9 x 4 and 4 x 9 (rotate.f:260-326)
diamond_C      NRLtb/atom_params/C_par as published: sp basis (kbas = 4) and
               new-style overlaps, nltype 1 (setpar.f:643-719,
               setparv.f:640-684)
diamond_C_nl3  the same parameters with nltype 3: new-style overlaps and
               Harrison signs (setpar.f:604-614, 680-692)
pd_harrison    Pd-like parameters with nltype 2 (Harrison signs only)
h_atom         s-only atoms (kbas = 1) with nltype 90000, the H-atom form of
               both integrals (setpar.f:615-628, 693-705), whose volume
               derivative takes the standard form (setparv.f:640 tests 9000)
h_atom_single  one such atom in the cell: order N = 1, the 1 x 1 pencil
               (zhpgv with n = 1 as diag.f:396-412 calls it)
orthogonal     an orthogonal basis, jsover false: S = I, zhpev in place of
               zhpgv (diag.f:385-414, setham.f:399-453)
degenerate     the perfect 32-atom fcc supercell at Gamma: groups of 12 equal
               eigenvalues, which perturb refuses (perturb.f:106)
no_electrons   zfill = 0, which fermi refuses (fermi.f:165)
=============  ===============================================================

Each case is a plain dictionary of inputs (JSON-able: parameters, lattice,
positions, kinds, the Monkhorst-Pack divisions q) from which
:func:`structure` builds the :class:`tbk.structures.Structure`.
tests/golden/make_branch_fixture.py writes the reference's own results for
them to tests/golden/branches.json.
"""
import numpy as np

from tbk import structures as S

SEED = 20261203
DEGEN_STOP = "dgen:  Degeneracy is too large"
ZFILL_STOP = "FERMI:  No solution for zfill <= 0"


def _two_kind_params(kbas):
    """Two kinds from the seeded one-kind set: kind 2's on-site and like-pair
    values perturbed by 2 %, unlike pairs (14 integrals) at 0.9 times the
    like ones (input.f:327-359 order)."""
    base = S.synthetic_params(3)
    kinds = 2
    rng = np.random.default_rng(5)
    par = np.zeros(S.npard(kinds))
    par[0:17] = base.param[0:17]
    par[17:34] = base.param[0:17] * (1 + 0.02 * rng.standard_normal(17))
    head = kinds * (1 + 4 * (3 * kinds + 1))
    par[head:head + 40] = base.param[17:57]
    par[head + 40:head + 80] = base.param[17:57] * (1 + 0.02 * rng.standard_normal(40))
    par[head + 80:head + 136] = np.concatenate([base.param[17:57], base.param[17:33]]) * 0.9
    o0 = head + 136
    par[o0:o0 + 40] = base.param[57:97]
    par[o0 + 40:o0 + 80] = base.param[57:97]
    par[o0 + 80:o0 + 136] = np.concatenate([base.param[57:97], base.param[57:73]]) * 0.9
    val = [11.0 if b == 9 else 4.0 for b in kbas]
    return dict(kinds=2, kbas=list(kbas), valence=val, dnn=[12.0, 12.0, 12.0],
                screenl=[0.5, 0.5, 0.5], param=par.tolist(), nltype=0, jsover=True)


def _h_atom_params():
    """One s-only kind in the H-atom form: H_ss = (-0.8 + 0.02 R) e^{-0.75^2 R} F,
    S_ss = 0.4 e^{-0.75^2 R} F, on-site -0.3 + 0.1 rho^{2/3}."""
    p = np.zeros(S.npard(1))
    p[0] = 1.2
    p[1:5] = [-0.3, 0.1, 0.0, 0.0]
    p[17:23] = [-0.8, 0.02, 0.0, 0.75, 0.0, 0.0]
    p[57:63] = [0.4, 0.0, 0.0, 0.75, 0.0, 0.0]
    return dict(kinds=1, kbas=[1], valence=[1.0], dnn=[12.0], screenl=[0.5], param=p.tolist(),
                nltype=90000, jsover=True)


def _params_dict(p: S.NRLParams, **over):
    d = dict(kinds=p.kinds, kbas=list(p.kbas), valence=list(p.valence), dnn=list(p.dnn),
             screenl=list(p.screenl), param=np.asarray(p.param, float).tolist(), nltype=p.nltype,
             jsover=bool(p.jsover))
    d.update(over)
    return d


def _cell(kind, a, reps=(1, 1, 1)):
    lat, frac = S.conventional_cell(kind, a)
    return S.supercell(lat, frac, reps)


def cases(c_par: dict):
    """The cases; c_par: NRLtb/atom_params/C_par as a parameter dictionary
    (tests/golden/branches.json holds it)."""
    out = []
    lat, pos = _cell("bcc", 6.8)       # S(k) positive definite (min eigenvalue 0.17)
    out.append(dict(name="two_kind", params=_two_kind_params([9, 9]), lat=lat, posn=pos, kkind=[1, 2],
                    q=4, expect="ok"))
    out.append(dict(name="two_kind_sp", params=_two_kind_params([9, 4]), lat=lat, posn=pos, kkind=[1, 2],
                    q=4, expect="ok"))
    a = 3.567 * S.ANGSTROM
    lat = S.fcc_primitive(a)
    pos = np.array([[0.0, 0.0, 0.0], [0.25, 0.25, 0.25]]) * a
    out.append(dict(name="diamond_C", params=dict(c_par), lat=lat, posn=pos, kkind=[1, 1], q=5,
                    expect="ok"))
    out.append(dict(name="diamond_C_nl3", params=dict(c_par, nltype=3), lat=lat, posn=pos, kkind=[1, 1],
                    q=5, expect="ok"))
    pd = S.synthetic_params(7)
    lat, pos = _cell("fcc", 3.61 * S.ANGSTROM)
    out.append(dict(name="pd_harrison", params=_params_dict(pd, nltype=2), lat=lat, posn=pos,
                    kkind=[1, 1, 1, 1], q=3, expect="ok"))
    out.append(dict(name="orthogonal", params=_params_dict(pd, jsover=False), lat=lat, posn=pos,
                    kkind=[1, 1, 1, 1], q=3, expect="ok"))
    lat, pos = _cell("fcc", 5.0)
    out.append(dict(name="h_atom", params=_h_atom_params(), lat=lat, posn=pos, kkind=[1, 1, 1, 1], q=4,
                    expect="ok"))
    out.append(dict(name="h_atom_single", params=_h_atom_params(), lat=S.fcc_primitive(5.0).tolist(),
                    posn=[[0.0, 0.0, 0.0]], kkind=[1], q=6, expect="ok"))
    lat, pos = _cell("fcc", 3.61 * S.ANGSTROM, (2, 2, 2))
    out.append(dict(name="degenerate", params=_params_dict(S.synthetic_params(SEED)), lat=lat, posn=pos,
                    kkind=[1] * 32, q=1, expect="degen"))
    lat, pos = _cell("fcc", 3.61 * S.ANGSTROM)
    out.append(dict(name="no_electrons", params=_params_dict(pd), lat=lat, posn=pos, kkind=[1, 1, 1, 1],
                    q=2, zcharge=-44.0, expect="zfill"))
    for c in out:
        c["lat"] = np.asarray(c["lat"], float).tolist()
        c["posn"] = np.asarray(c["posn"], float).tolist()
    return out


def structure(case) -> S.Structure:
    p = case["params"]
    prm = S.NRLParams(p["kinds"], list(p["kbas"]), list(p["valence"]), list(p["dnn"]), list(p["screenl"]),
                      np.array(p["param"], float), int(p["nltype"]), bool(p["jsover"]))
    lat = np.array(case["lat"], float)
    pos = np.array(case["posn"], float)
    posq = pos @ np.linalg.inv(lat)
    rc = max(prm.dnn)
    return S.Structure(lat, pos, np.array(case["kkind"], np.int32), prm, S.search_widths(lat, rc), posq,
                       float(case.get("zcharge", 0.0)), case["name"])


def mesh(case):
    q = int(case["q"])
    k = S.monkhorst_pack((q, q, q)) if q > 1 else np.zeros((1, 3))
    return k, np.full(len(k), 1.0 / len(k))


# a k-point off every symmetry element, for eigenvalue and H/S comparisons
GENERAL_K = [0.137, -0.291, 0.413]
