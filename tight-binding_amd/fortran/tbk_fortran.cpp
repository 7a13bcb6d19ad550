// tbk_fortran.cpp -- Fortran-ABI drop-ins for the reference's bande and
// fermi (NRLtb/src/bande.f:1-4, NRLtb/src/fermi.f:1-2).
//
// bande_ reads the structure, parameters and k-points from the COMMON
// blocks the reference's main program fills (input, input1, kptin,
// search2: static.f:1218-1222) and runs the whole k-loop, Fermi search and
// sums through the C ABI (tbk_plan_*).  Link this file's object (built into
// lib/libtbk_fortran.a) with the Fortran program objects in place of
// bande.o / fermi.o, and lib/libtbk.so.
//
// The COMMON layouts follow the reference's P1/P2/P3 dimensions, given here
// as TBK_MATOM, TBK_MKIND, TBK_MSPIN, TBK_NKD (defaults: NRLtb/ex1/P1, P3).
// Rebuild with -D for other dimensions.
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "tbk.h"

#ifndef TBK_MATOM
#define TBK_MATOM 20
#endif
#ifndef TBK_MKIND
#define TBK_MKIND 2
#endif
#ifndef TBK_MSPIN
#define TBK_MSPIN 2
#endif
#ifndef TBK_NKD
#define TBK_NKD 50000
#endif
#define TBK_NPARD (TBK_MKIND * (29 + 68 * TBK_MKIND))     // P1: npard
#define TBK_MINT (TBK_MKIND * (TBK_MKIND + 1) / 2)         // P1: mint

typedef int f_logical;                  // LOGICAL*4: 0 / nonzero

// COMMON blocks (column-major Fortran arrays, hence reversed C indices)
extern "C" {
extern struct { double posn[3][TBK_MATOM]; int kkind[TBK_MATOM]; int natoms; } codes_;
extern struct { double plv[3][3]; } latt1_;                 // plv(v, c): component c of vector v
extern struct { double posq[3][TBK_MATOM]; } fazon_;
extern struct { double valence[TBK_MKIND]; int kinds; int kbas[TBK_MKIND]; } parinfo_;
extern struct { double dnnsq[TBK_MINT], dnn[TBK_MINT], screenl[TBK_MINT], scrninv[TBK_MINT]; } cutcom_;
extern struct { int nkdarry[6]; } disply_;
extern struct { int npts; f_logical jsover; } lia_;
extern struct { double x[TBK_NKD], y[TBK_NKD], z[TBK_NKD]; } cd1_;
extern struct { double param[TBK_MSPIN][TBK_NPARD]; } params_;
extern struct { int nltype; f_logical realov, harrison; } parcom_;
}

static int device_ordinal()
{
    const char *e = std::getenv("TBK_DEVICE");
    return e ? std::atoi(e) : 0;
}

// iberr beyond the reference's 1..4: 5 = feature outside this path
// (torques, spin directions out, modes other than 1..4), 6 = device or library error,
// 7 = S(k) not positive definite (the reference's ioverr), 8 = more than 9
// degenerate states (perturb.f STOP).
static int iberr_of(int rc)
{
    switch (rc) {
    case TBK_OK: return 0;
    case TBK_ERR_OVERLAP: return 7;
    case TBK_ERR_DEGEN: return 8;
    case TBK_ERR_ARG: return 5;
    default: return 6;
    }
}

extern "C" void bande_(int *mode, int *jspins, f_logical *flipos, f_logical *lnoncol, double *spinat,
                       f_logical *ltorque, double *dspth, double *dspphi, f_logical *lspdir, double *thisvol,
                       double *zfill, double *tkb, double *eigcut, double *weight, double *efermi,
                       double *earg1, double *earg2, double *smag, f_logical *lp_par, f_logical *lp_parv,
                       f_logical *lp_klst, int *kptpr, char *label, int *iberr, size_t label_len)
{
    (void)dspth; (void)dspphi; (void)lp_par; (void)lp_parv; (void)lp_klst;
    (void)kptpr; (void)label; (void)label_len;
    *iberr = 0;
    if (*lnoncol && *jspins == 1) { *iberr = 1; return; }          // bande.f:695-703
    if (*jspins < 1 || *jspins > 2 || *ltorque || *lspdir || *mode < 1 || *mode > 4) {
        *iberr = 5;
        return;
    }
    if (*jspins == 2 && TBK_MSPIN < 2) { *iberr = 5; return; }
    const int nat = codes_.natoms, kinds = parinfo_.kinds, nk = lia_.npts;
    if (nat < 1 || nat > TBK_MATOM || kinds < 1 || kinds > TBK_MKIND || nk < 1 || nk > TBK_NKD) {
        *iberr = 5;
        return;
    }
    std::vector<double> lat(9), posn(3 * nat), posq(3 * nat), kp(3 * (size_t)nk), w(weight, weight + nk);
    for (int v = 0; v < 3; ++v)
        for (int c = 0; c < 3; ++c) lat[3 * v + c] = latt1_.plv[c][v];
    for (int a = 0; a < nat; ++a)
        for (int c = 0; c < 3; ++c) {
            posn[3 * a + c] = codes_.posn[c][a];
            posq[3 * a + c] = fazon_.posq[c][a];
        }
    for (int k = 0; k < nk; ++k) {
        kp[3 * k] = cd1_.x[k]; kp[3 * k + 1] = cd1_.y[k]; kp[3 * k + 2] = cd1_.z[k];
    }
    tbk_structure_desc d;
    d.natoms = nat;
    d.lat = lat.data(); d.posn = posn.data(); d.posq = posq.data(); d.kkind = codes_.kkind;
    for (int c = 0; c < 3; ++c) d.nd[c] = disply_.nkdarry[c];
    d.kinds = kinds; d.kbas = parinfo_.kbas; d.valence = parinfo_.valence;
    d.dnn = cutcom_.dnn; d.screenl = cutcom_.screenl;
    d.param = params_.param[0]; d.nparam = TBK_NPARD;
    d.nltype = parcom_.nltype; d.jsover = lia_.jsover ? 1 : 0;
    d.thisvol = *thisvol; d.zfill = *zfill;
    tbk_plan *p = nullptr;
    int rc;
    *smag = 0.0;
    if (*jspins == 1) {
        rc = tbk_plan_create(device_ordinal(), &d, nk, kp.data(), w.data(), 0, TBK_FP64, &p);
    } else if (*lnoncol) {
        // spinat(3, matom): the atoms' spin directions (input1.f); smag
        // stays 0 as in the reference's non-collinear branch
        rc = tbk_plan_create_noncol(device_ordinal(), &d, params_.param[1], spinat, nk, kp.data(), w.data(), 0,
                                    TBK_FP64, &p);
    } else {
        // both channels of /params/ (param(npard, mspin)); flipos from input1
        std::vector<int> flip(nat);
        for (int a = 0; a < nat; ++a) flip[a] = flipos[a] ? 1 : 0;
        rc = tbk_plan_create_spin(device_ordinal(), &d, params_.param[1], flip.data(), nk, kp.data(), w.data(),
                                  0, TBK_FP64, &p);
    }
    if (!rc) rc = tbk_plan_run(p, *mode, *tkb, *eigcut, efermi, earg1, earg2);
    if (!rc && *jspins == 2 && !*lnoncol) rc = tbk_plan_magnetization(p, smag);
    if (p) tbk_plan_destroy(p);
    *iberr = iberr_of(rc);
    if (rc) std::fprintf(stderr, "bande (device): %s\n", tbk_error_string(rc));
}

// fermi(ntbd, nkptd, nkloc, istart, ifin, nv, jspins, el, weight, eigen,
// tkb, zfill): eigen(ntbd, nkloc, jspins) for k-points istart..ifin of the
// global list (one process holds them all here), weight(nkptd).
extern "C" double fermi_(int *ntbd, int *nkptd, int *nkloc, int *istart, int *ifin, int *nv, int *jspins,
                         double *el, double *weight, double *eigen, double *tkb, double *zfill)
{
    (void)nkptd;
    const int nk = *ifin - *istart + 1, js = *jspins;
    if (nk < 1 || nk > *nkloc || js < 1 || js > 2) return 0.0;
    // spin channels as extra k-points with the same weights
    std::vector<double> e((size_t)(*ntbd) * nk * js), w((size_t)nk * js);
    for (int s = 0; s < js; ++s)
        for (int k = 0; k < nk; ++k) {
            w[(size_t)s * nk + k] = weight[*istart - 1 + k];
            for (int i = 0; i < *ntbd; ++i)
                e[((size_t)s * nk + k) * (*ntbd) + i] = eigen[((size_t)s * (*nkloc) + k) * (*ntbd) + i];
        }
    double ef = 0.0;
    const int rc = tbk_fermi(device_ordinal(), *ntbd, nk * js, *nv, *el, w.data(), e.data(), *tkb, *zfill, &ef);
    if (rc) std::fprintf(stderr, "fermi (device): %s\n", tbk_error_string(rc));
    return ef;
}
