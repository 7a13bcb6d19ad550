/*
 * nrl_oracle.c -- CPU restatement of the NRL tight-binding per-k-point path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Only tests/, the smoke entry and
 * the CPU-baseline leg of bench.py may load this library; the product path
 * never links or calls it.
 *
 * Built with -O2 -ffp-contract=off and no -march flags so that the
 * k-independent set-up (neighbour list, SK parameters, SK tables, H/S) uses
 * the same IEEE operation sequence as the reference Fortran; each routine
 * cites the reference lines it follows.  The eigensolver follows the
 * algorithm of LAPACK 1.1 zhpgv (lapack.f:2274-2473): Cholesky of S,
 * reduction to standard form, Householder tridiagonalisation, implicit QL
 * with vector accumulation and back-transformation; it is written on full
 * storage and sorts eigenpairs ascending as zsteqr does.
 */
#include "nrl_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NPK1 4   /* npkind(1): on-site terms (static.f:610)        */
#define NPK2 10  /* npkind(2): like-pair integrals (static.f:610)   */
#define NPK3 14  /* npkind(3): unlike-pair integrals (static.f:610) */

/* ------------------------------------------------------------------ */
/* search2.f:121-223                                                   */
/* ------------------------------------------------------------------ */
int orc_search2(int natoms, const double *lat, const double *posn,
                const int *kkind, int kinds, const int *nd,
                const double *dnn, const double *screenl, int maxpairs,
                int *npairs_out, double *tt, double *dlv, double *dist,
                double *screen, double *dscrn, int *jkind, int *jatm,
                double *dismin)
{
    const double eps = 1e-3, five = 5.0, one = 1.0;
    int kd[3] = {2 * nd[0] + 1, 2 * nd[1] + 1, 2 * nd[2] + 1};
    int npairs = 0;
    int kneigh = 0; /* COMMON /relat/ kneigh: zero in a fresh process */
    double dismin2 = 1e10;
    for (int iat = 0; iat < natoms; ++iat) {
        for (int jat = iat; jat < natoms; ++jat) {
            int k1 = kkind[iat], k2 = kkind[jat], intknd;
            if (k1 == k2) {
                intknd = k1;
            } else {
                int jk1 = k1 < k2 ? k1 : k2, jk2 = k1 < k2 ? k2 : k1;
                intknd = kinds + (jk2 - 1) * (jk2 - 2) / 2 + jk1;
            }
            const double dnnv = dnn[intknd - 1];
            const double dnnsq = dnnv * dnnv;            /* input.f:294 */
            const double scrninv = one / screenl[intknd - 1]; /* input.f:295 */
            const double screenr = dnnv - five * screenl[intknd - 1];
            for (int ix = 0; ix < kd[0]; ++ix)
            for (int iy = 0; iy < kd[1]; ++iy)
            for (int iz = 0; iz < kd[2]; ++iz) {
                int imat[3] = {ix - nd[0], iy - nd[1], iz - nd[2]};
                double t[3], dsq = 0.0;
                for (int i = 0; i < 3; ++i) {
                    double rlv = 0.0;
                    for (int j = 0; j < 3; ++j)
                        rlv = rlv + lat[3 * j + i] * (double)imat[j];
                    t[i] = posn[3 * jat + i] + rlv - posn[3 * iat + i];
                    dsq = dsq + t[i] * t[i];
                }
                double dd = dsq - eps;
                if (dd > dnnsq) continue;
                if (dd < dnnsq) kneigh = 1;
                if (dd < eps) kneigh = 0;
                if (dd > eps && dsq < dismin2) dismin2 = dsq;
                if (kneigh == 0) continue;
                if (npairs >= maxpairs) { *npairs_out = npairs; return 1; }
                int p = npairs++;
                double d = sqrt(dsq);
                double tinv = one / d;
                for (int j = 0; j < 3; ++j) {
                    tt[3 * p + j] = t[j] * tinv;
                    dlv[3 * p + j] = (double)imat[j];
                }
                double expdl = exp((d - screenr) * scrninv);
                screen[p] = one / (expdl + one);
                dscrn[p] = -expdl * screen[p] * scrninv;
                jkind[2 * p] = k1; jkind[2 * p + 1] = k2;
                jatm[2 * p] = iat + 1; jatm[2 * p + 1] = jat + 1;
                dist[p] = d;
            }
        }
    }
    *npairs_out = npairs;
    *dismin = sqrt(dismin2);
    return 0;
}

/* Offsets into param (0-based) -- setpar.f:461, 558-572, 655-666. */
static int onsite_ip(int kind) { return (kind - 1) * (1 + NPK1 * 4); }
static int hop_ip(int kinds, int jk1, int jk2)
{
    if (jk1 == jk2)
        return kinds * (1 + NPK1 * (3 * kinds + 1)) + 4 * NPK2 * (jk1 - 1);
    return kinds * (1 + NPK1 * (3 * kinds + 1)) + 4 * NPK2 * kinds
         + 4 * NPK3 * ((jk2 - 1) * (jk2 - 2) / 2 + jk1 - 1);
}
static int over_ip(int kinds, int jk1, int jk2)
{
    if (jk1 == jk2)
        return kinds * (1 + NPK1 * (3 * kinds + 1)) + 4 * NPK2 * kinds
             + 4 * NPK3 * (kinds * (kinds - 1) / 2) + 4 * NPK2 * (jk1 - 1);
    return kinds * (1 + NPK1 * (3 * kinds + 1)) + 8 * NPK2 * kinds
         + 4 * NPK3 * (kinds * (kinds - 1) / 2)
         + 4 * NPK3 * ((jk2 - 1) * (jk2 - 2) / 2 + jk1 - 1);
}
/* param(ip+x) in 1-based Fortran == P(ip + x - 1) here. */
#define P(i) param[(i) - 1]

static const double ovlnull[10] = {1, 0, 1, 1, 0, 0, 0, 1, 1, 1};
static const double harsign[9] = {-1, 1, 1, -1, -1, -1, 1, -1, 1};

/* ------------------------------------------------------------------ */
/* setpar.f:398-884                                                    */
/* ------------------------------------------------------------------ */
int orc_setpar(int natoms, const int *kkind, int kinds, const double *param,
               int nltype, int realov, int harrison, int jsover,
               int npairs, const double *dist, const double *screen,
               const int *jkind, const int *jatm,
               double *den, double *paros, double *hop, double *over)
{
    const double zero = 0.0, two = 2.0, three = 3.0, twtrd = two / three;
    for (int i = 0; i < natoms * kinds; ++i) den[i] = zero;
    for (int p = 0; p < npairs; ++p) {
        int iat = jatm[2 * p] - 1, jat = jatm[2 * p + 1] - 1;
        int k1 = jkind[2 * p], k2 = jkind[2 * p + 1];
        double d = dist[p];
        /* density (setpar.f:508-536) */
        int ip = onsite_ip(k2) + 1;
        double b = P(ip) * P(ip);
        double dent = exp(-b * d) * screen[p];
        den[iat * kinds + (k2 - 1)] = den[iat * kinds + (k2 - 1)] + dent;
        if (jat != iat) {
            int ipt = onsite_ip(k1) + 1;
            b = P(ipt) * P(ipt);
            dent = exp(-b * d) * screen[p];
            den[jat * kinds + (k1 - 1)] = den[jat * kinds + (k1 - 1)] + dent;
        }
        int jk1 = k1 < k2 ? k1 : k2, jk2 = k1 < k2 ? k2 : k1;
        int npk = (jk1 == jk2) ? NPK2 : NPK3;
        ip = hop_ip(kinds, jk1, jk2);
        double *hp = hop + ORC_NPK * p;
        for (int j = 0; j < ORC_NPK; ++j) hp[j] = 0.0;
        if (nltype != 90000) {
            for (int j = 1; j <= npk; ++j) {
                double abr = P(ip + 4 * j - 3) + d * (P(ip + 4 * j - 2) +
                             d * P(ip + 4 * j - 1));
                double p32m = -(P(ip + 4 * j) * P(ip + 4 * j));
                hp[j - 1] = abr * exp(p32m * d) * screen[p];
            }
            if (harrison && jk1 == jk2)
                for (int j = 0; j < 9; ++j)
                    hp[j] = harsign[j] * fabs(hp[j]);   /* setpar.f:604-614 */
        } else {
            int j = 1;
            double abr = P(ip + 4 * j - 3) + d * (P(ip + 4 * j - 2) +
                         d * (P(ip + 4 * j - 1) + d * (P(ip + 4 * j + 1) +
                         d * P(ip + 4 * j + 2))));
            double p32m = -(P(ip + 4 * j) * P(ip + 4 * j));
            hp[0] = abr * exp(p32m * d) * screen[p];
        }
        double *op = over + ORC_NPK * p;
        for (int j = 0; j < ORC_NPK; ++j) op[j] = 0.0;
        if (jsover) {                                  /* setpar.f:643-719 */
            int usenew = realov && (jk1 == jk2);
            int iop = over_ip(kinds, jk1, jk2);
            if (nltype != 90000) {
                for (int j = 1; j <= npk; ++j) {
                    double abr = P(iop + 4 * j - 3) + d * (P(iop + 4 * j - 2) +
                                 d * P(iop + 4 * j - 1));
                    double p32m = -(P(iop + 4 * j) * P(iop + 4 * j));
                    if (usenew) abr = ovlnull[j - 1] + d * abr;
                    op[j - 1] = abr * exp(p32m * d) * screen[p];
                }
                if (harrison && jk1 == jk2)
                    for (int j = 0; j < 9; ++j)
                        op[j] = -harsign[j] * fabs(op[j]);
            } else {
                int j = 1;
                double abr = P(iop + 4 * j - 3) + d * (P(iop + 4 * j - 2) +
                             d * (P(iop + 4 * j - 1) + d * (P(iop + 4 * j + 1) +
                             d * P(iop + 4 * j + 2))));
                double p32m = -(P(iop + 4 * j) * P(iop + 4 * j));
                op[0] = abr * exp(p32m * d) * screen[p];
            }
        }
    }
    /* on-site terms (setpar.f:795-867) */
    for (int iat = 0; iat < natoms; ++iat) {
        int jk1 = kkind[iat];
        int ip = onsite_ip(jk1) + 1;
        double vm23 = pow(den[iat * kinds + jk1 - 1], twtrd);
        for (int k = 0; k < 4; ++k) {
            paros[4 * iat + k] = P(ip + 1) + vm23 * (P(ip + 2) + vm23 *
                                 (P(ip + 3) + vm23 * P(ip + 4)));
            ip += 4;
        }
        for (int jk2 = 1; jk2 <= kinds; ++jk2) {
            if (jk2 == jk1 || den[iat * kinds + jk2 - 1] == zero) continue;
            vm23 = pow(den[iat * kinds + jk2 - 1], twtrd);
            int jkx = jk1 > jk2 ? jk1 : jk2, jkm = jk1 < jk2 ? jk1 : jk2;
            ip = kinds * (1 + NPK1 * 4) +
                 6 * NPK1 * ((jkx - 1) * (jkx - 2) / 2 + jkm - 1);
            if (jk2 < jkx) ip += 3 * NPK1;
            for (int k = 0; k < 4; ++k) {
                paros[4 * iat + k] = paros[4 * iat + k] + vm23 * (P(ip + 1) +
                                     vm23 * (P(ip + 2) + vm23 * P(ip + 3)));
                ip += 3;
            }
        }
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* setparv.f:426-886                                                   */
/* ------------------------------------------------------------------ */
int orc_setparv(int natoms, const int *kkind, int kinds, const double *param,
                int nltype, int realov, int harrison, int jsover,
                double thisvol, int npairs, const double *dist,
                const double *screen, const double *dscrn,
                const int *jkind, const int *jatm, const double *den,
                double *dparos, double *dhop, double *dover)
{
    const double zero = 0.0, one = 1.0, two = 2.0, three = 3.0, four = 4.0;
    const double third = one / three, twtrd = two / three;
    const double threeh = three / two;
    double *vden = (double *)calloc((size_t)natoms * kinds, sizeof(double));
    for (int p = 0; p < npairs; ++p) {
        int iat = jatm[2 * p] - 1, jat = jatm[2 * p + 1] - 1;
        int k1 = jkind[2 * p], k2 = jkind[2 * p + 1];
        double d = dist[p];
        double vdist = third * d / thisvol;            /* setparv.f:459 */
        int ip = onsite_ip(k2) + 1;
        double b = P(ip) * P(ip);
        double dent = exp(-b * d) * screen[p];
        vden[iat * kinds + k2 - 1] = vden[iat * kinds + k2 - 1] +
                                     (dscrn[p] - b) * dent * vdist;
        if (jat != iat) {
            int ipt = onsite_ip(k1) + 1;
            b = P(ipt) * P(ipt);
            dent = exp(-b * d) * screen[p];
            vden[jat * kinds + k1 - 1] = vden[jat * kinds + k1 - 1] +
                                         (dscrn[p] - b) * dent * vdist;
        }
        int jk1 = k1 < k2 ? k1 : k2, jk2 = k1 < k2 ? k2 : k1;
        int npk = (jk1 == jk2) ? NPK2 : NPK3;
        int usehar = harrison && (jk1 == jk2);
        ip = hop_ip(kinds, jk1, jk2);
        double *hp = dhop + ORC_NPK * p;
        for (int j = 0; j < ORC_NPK; ++j) hp[j] = 0.0;
        if (nltype != 90000) {
            for (int j = 1; j <= npk; ++j) {
                double abr = P(ip + 4 * j - 3) + d * (P(ip + 4 * j - 2) +
                             d * P(ip + 4 * j - 1));
                double p32m = -(P(ip + 4 * j) * P(ip + 4 * j));
                double par21 = abr * exp(p32m * d) * screen[p];
                hp[j - 1] = vdist * (((P(ip + 4 * j - 2) + two * d *
                            P(ip + 4 * j - 1)) + p32m * abr) * exp(p32m * d) *
                            screen[p] + dscrn[p] * par21);
                if (usehar && j < npk)
                    hp[j - 1] = harsign[j - 1] * hp[j - 1] * copysign(one, abr);
            }
        } else {
            int j = 1;
            double abr = P(ip + 4 * j - 3) + d * (P(ip + 4 * j - 2) +
                         d * (P(ip + 4 * j - 1) + d * (P(ip + 4 * j + 1) +
                         d * P(ip + 4 * j + 2))));
            double p32m = -(P(ip + 4 * j) * P(ip + 4 * j));
            double par21 = abr * exp(p32m * d) * screen[p];
            hp[0] = vdist * (((P(ip + 4 * j - 2) + d * (two * P(ip + 4 * j - 1) +
                    d * (three * P(ip + 4 * j + 1) + d * (four *
                    P(ip + 4 * j + 2))))) + p32m * abr) * exp(p32m * d) *
                    screen[p] + dscrn[p] * par21);
        }
        double *op = dover + ORC_NPK * p;
        for (int j = 0; j < ORC_NPK; ++j) op[j] = 0.0;
        if (jsover) {
            int usenew = realov && (jk1 == jk2);
            int iop = over_ip(kinds, jk1, jk2);
            /* setparv.f:640 tests nltype against 9000 (not 90000). */
            if (nltype != 9000) {
                for (int j = 1; j <= npk; ++j) {
                    double abr, ovl21, p32m = -(P(iop + 4 * j) * P(iop + 4 * j));
                    if (usenew) {
                        abr = ovlnull[j - 1] + d * (P(iop + 4 * j - 3) +
                              d * (P(iop + 4 * j - 2) + d * P(iop + 4 * j - 1)));
                        ovl21 = abr * exp(p32m * d) * screen[p];
                        op[j - 1] = vdist * (((P(iop + 4 * j - 3) + two * d *
                                    P(iop + 4 * j - 2) + three * d * d *
                                    P(iop + 4 * j - 1)) + p32m * abr) *
                                    exp(p32m * d) * screen[p] + dscrn[p] * ovl21);
                    } else {
                        abr = P(iop + 4 * j - 3) + d * (P(iop + 4 * j - 2) +
                              d * P(iop + 4 * j - 1));
                        ovl21 = abr * exp(p32m * d) * screen[p];
                        op[j - 1] = vdist * (((P(iop + 4 * j - 2) + two * d *
                                    P(iop + 4 * j - 1)) + p32m * abr) *
                                    exp(p32m * d) * screen[p] + dscrn[p] * ovl21);
                    }
                    if (usehar && j < npk)
                        op[j - 1] = -harsign[j - 1] * op[j - 1] * copysign(one, abr);
                }
            } else {
                int j = 1;
                double abr = P(iop + 4 * j - 3) + d * (P(iop + 4 * j - 2) +
                             d * (P(iop + 4 * j - 1) + d * (P(iop + 4 * j + 1) +
                             d * P(iop + 4 * j + 2))));
                double p32m = -(P(iop + 4 * j) * P(iop + 4 * j));
                double par21 = abr * exp(p32m * d) * screen[p];
                op[0] = vdist * (((P(iop + 4 * j - 2) + d * (two *
                        P(iop + 4 * j - 1) + d * (three * P(iop + 4 * j + 1) +
                        d * (four * P(iop + 4 * j + 2))))) + p32m * abr) *
                        exp(p32m * d) * screen[p] + dscrn[p] * par21);
            }
        }
    }
    for (int iat = 0; iat < natoms; ++iat) {       /* setparv.f:820-877 */
        int jk1 = kkind[iat];
        int ip = onsite_ip(jk1) + 1;
        double vm13 = pow(den[iat * kinds + jk1 - 1], third);
        double vm23 = vm13 * vm13;
        double vm2t = two * vm23, vm3t = threeh * vm23;
        double drho = twtrd * vden[iat * kinds + jk1 - 1] / vm13;
        for (int k = 0; k < 4; ++k) {
            dparos[4 * iat + k] = drho * (P(ip + 2) + vm2t * (P(ip + 3) +
                                  vm3t * P(ip + 4)));
            ip += 4;
        }
        for (int jk2 = 1; jk2 <= kinds; ++jk2) {
            if (jk2 == jk1 || den[iat * kinds + jk2 - 1] == zero) continue;
            vm13 = pow(den[iat * kinds + jk2 - 1], third);
            vm23 = vm13 * vm13;
            vm2t = two * vm23; vm3t = threeh * vm23;
            drho = twtrd * vden[iat * kinds + jk2 - 1] / vm13;
            int jkx = jk1 > jk2 ? jk1 : jk2, jkm = jk1 < jk2 ? jk1 : jk2;
            ip = kinds * (1 + NPK1 * 4) +
                 6 * NPK1 * ((jkx - 1) * (jkx - 2) / 2 + jkm - 1);
            if (jk2 < jkx) ip += 3 * NPK1;
            for (int k = 0; k < 4; ++k) {
                dparos[4 * iat + k] = dparos[4 * iat + k] + drho * (P(ip + 1) +
                                      vm2t * (P(ip + 2) + vm3t * P(ip + 3)));
                ip += 3;
            }
        }
    }
    free(vden);
    return 0;
}
#undef P

/* ------------------------------------------------------------------ */
/* rotate.f:112-258: the 9x9 Slater-Koster table e(i,j) (row i, col j)  */
/* ------------------------------------------------------------------ */
static void sk_table(double l, double m, double n, const double *p,
                     int unlike, double e[9][9])
{
    const double rt3 = 1.732050807568877, half = 0.5;
    const double rt3h = half * rt3, rt3q = half * rt3h, rt32 = 2.0 * rt3;
    static const double parity[9] = {1, -1, -1, -1, 1, 1, 1, 1, 1};
    double ll = l * l, mm = m * m, nn = n * n, lm = l * m, mn = m * n,
           nl = n * l, lmn = lm * n;
#define E(i, j) e[(i) - 1][(j) - 1]
#define PP(i) p[(i) - 1]
    E(1, 1) = PP(1);
    E(1, 2) = l * PP(2); E(1, 3) = m * PP(2); E(1, 4) = n * PP(2);
    double rp5 = rt3 * PP(5);
    E(1, 5) = mn * rp5; E(1, 6) = nl * rp5; E(1, 7) = lm * rp5;
    E(1, 8) = half * rp5 * (ll - mm);
    E(1, 9) = (nn - half * (ll + mm)) * PP(5);
    double p34 = PP(3) - PP(4);
    E(2, 2) = PP(4) + ll * p34; E(3, 3) = PP(4) + mm * p34;
    E(4, 4) = PP(4) + nn * p34;
    E(2, 3) = lm * p34; E(2, 4) = nl * p34; E(3, 4) = mn * p34;
    double p67 = rt3 * PP(6) - 2.0 * PP(7);
    E(2, 5) = lmn * p67;
    E(2, 6) = n * (ll * p67 + PP(7));
    E(2, 7) = m * (ll * p67 + PP(7));
    E(3, 5) = n * (mm * p67 + PP(7));
    E(3, 6) = E(2, 5);
    E(3, 7) = l * (mm * p67 + PP(7));
    E(4, 5) = m * (nn * p67 + PP(7));
    E(4, 6) = l * (nn * p67 + PP(7));
    E(4, 7) = E(2, 5);
    E(2, 8) = l * (half * (ll - mm) * p67 + PP(7));
    double p67a = PP(6) - rt3 * PP(7);
    E(2, 9) = l * (nn * p67a - (half * (ll + mm)) * PP(6));
    E(3, 8) = m * (half * (ll - mm) * p67 - PP(7));
    E(3, 9) = m * (nn * p67a - (half * (ll + mm)) * PP(6));
    E(4, 8) = n * (half * (ll - mm) * p67);
    E(4, 9) = n * (nn * PP(6) - half * (ll + mm) * (PP(6) - rt32 * PP(7)));
    double p8910 = (3.0 * PP(8) + PP(10)) - 4.0 * PP(9);
    E(5, 5) = ll * PP(10) + mm * (nn * p8910) + (mm + nn) * PP(9);
    double p910m = PP(9) - PP(10);
    E(5, 6) = lm * (p910m + nn * p8910);
    E(5, 7) = nl * (p910m + mm * p8910);
    E(6, 6) = mm * PP(10) + ll * (nn * p8910) + (ll + nn) * PP(9);
    E(6, 7) = mn * (p910m + ll * p8910);
    E(7, 7) = nn * PP(10) + ll * (mm * p8910) + (ll + mm) * PP(9);
    E(5, 8) = mn * ((half * (ll - mm)) * p8910 - p910m);
    E(6, 8) = nl * ((half * (ll - mm)) * p8910 + p910m);
    E(7, 8) = lm * (half * (ll - mm)) * p8910;
    double p89 = PP(8) - 2.0 * PP(9);
    double p8910a = p89 + PP(10);
    double p8910b = p8910a + p89;
    double p89m = PP(8) - PP(9);
    E(5, 9) = rt3 * mn * (nn * p89m - (half * (ll + mm)) * p8910a);
    E(6, 9) = rt3 * nl * (nn * p89m - (half * (ll + mm)) * p8910a);
    E(7, 9) = rt3h * lm * (PP(10) - (ll + mm) * PP(8) + nn * p8910b);
    E(8, 8) = (ll + mm) * PP(9) + nn * PP(10) + ((half * (ll - mm)) *
              (half * (ll - mm))) * p8910;
    E(8, 9) = rt3q * (ll - mm) * (PP(10) - (ll + mm) * PP(8) + nn * p8910b);
    E(9, 9) = nn * nn * PP(8) + (ll + mm) * (2.5e-1 * (ll + mm) *
              (3.0 * PP(10) + PP(8)) - nn * (PP(8) - 3.0 * PP(9)));
    for (int i = 2; i <= 9; ++i)
        for (int j = 1; j <= i - 1; ++j)
            E(i, j) = parity[i - 1] * (E(j, i) * parity[j - 1]);
    if (unlike) {
        E(2, 1) = l * PP(11); E(3, 1) = m * PP(11); E(4, 1) = n * PP(11);
        double rp12 = rt3 * PP(12);
        E(5, 1) = mn * rp12; E(6, 1) = nl * rp12; E(7, 1) = lm * rp12;
        E(8, 1) = half * rp12 * (ll - mm);
        E(9, 1) = (nn - half * (ll + mm)) * PP(12);
        double p134 = rt3 * PP(13) - 2.0 * PP(14);
        E(5, 2) = lmn * p134;
        E(6, 2) = n * (ll * p134 + PP(14));
        E(7, 2) = m * (ll * p134 + PP(14));
        E(5, 3) = n * (mm * p134 + PP(14));
        E(6, 3) = E(5, 2);
        E(7, 3) = l * (mm * p134 + PP(14));
        E(5, 4) = m * (nn * p134 + PP(14));
        E(6, 4) = l * (nn * p134 + PP(14));
        E(7, 4) = E(5, 2);
        E(8, 2) = l * (half * (ll - mm) * p134 + PP(14));
        double p134a = PP(13) - rt3 * PP(14);
        E(9, 2) = l * (nn * p134a - (half * (ll + mm)) * PP(13));
        E(8, 3) = m * (half * (ll - mm) * p134 - PP(14));
        E(9, 3) = m * (nn * p134a - (half * (ll + mm)) * PP(13));
        E(8, 4) = n * (half * (ll - mm) * p134);
        E(9, 4) = n * (nn * PP(13) - half * (ll + mm) * (PP(13) - rt32 * PP(14)));
    }
#undef E
#undef PP
}

/* setham.f:326-376: the par(.,jk2,jk1)/par(.,jk1,jk2) arrangement; returns
 * the vector rotate.f:155 reads, p(j) = par(j, jkind(2), jkind(1)). */
static void arrange_pair(const double *hp, int k1, int k2, double out[14])
{
    int jk1 = k1 < k2 ? k1 : k2, jk2 = k1 < k2 ? k2 : k1;
    int npk = (jk1 == jk2) ? NPK2 : NPK3;
    double a[14] = {0}, b[14] = {0};   /* a = par(.,jk2,jk1), b = par(.,jk1,jk2) */
    for (int j = 0; j < npk; ++j) a[j] = hp[j];
    if (jk1 != jk2) {
        for (int j = 0; j < NPK2; ++j) b[j] = a[j];
        b[1] = -a[10]; b[4] = a[11]; b[5] = -a[12]; b[6] = -a[13];
    }
    b[10] = -a[1]; b[11] = a[4]; b[12] = -a[5]; b[13] = -a[6];
    if (jk1 == jk2) {
        /* par(.,jk1,jk1) is one array: the like case writes 11..14 into it */
        for (int j = 0; j < NPK2; ++j) out[j] = a[j];
        for (int j = 10; j < 14; ++j) out[j] = b[j];
        return;
    }
    /* p(j) = par(j, jkind(2), jkind(1)) */
    if (k2 == jk2) memcpy(out, a, sizeof(a)); else memcpy(out, b, sizeof(b));
}

void orc_setham(int nv, const double *ak, int natoms, const int *kkind,
                const int *ksk, const int *kbas, const double *posq,
                int npairs, const double *tt, const double *dlv,
                const int *jkind, const int *jatm, const double *hop,
                const double *over, const double *paros, int jsover,
                int onsite_s, double *hp, double *sp)
{
    const double tpi = 6.283185307179586;
    size_t np = (size_t)nv * (nv + 1) / 2;
    memset(hp, 0, 2 * np * sizeof(double));
    if (sp) memset(sp, 0, 2 * np * sizeof(double));
    double e[9][9];
    for (int p = 0; p < npairs; ++p) {
        int i1 = jatm[2 * p] - 1, i2 = jatm[2 * p + 1] - 1;
        int k1 = jkind[2 * p], k2 = jkind[2 * p + 1];
        int kk1 = ksk[i1], kk2 = ksk[i2];
        int kb1 = kbas[k1 - 1], kb2 = kbas[k2 - 1];
        double dot = ak[0] * (dlv[3 * p + 0] - posq[3 * i2 + 0] + posq[3 * i1 + 0])
                   + ak[1] * (dlv[3 * p + 1] - posq[3 * i2 + 1] + posq[3 * i1 + 1])
                   + ak[2] * (dlv[3 * p + 2] - posq[3 * i2 + 2] + posq[3 * i1 + 2]);
        double cs = cos(tpi * dot), sn = sin(tpi * dot);
        for (int pass = 0; pass < (jsover && sp ? 2 : 1); ++pass) {
            double pv[14];
            arrange_pair((pass == 0 ? hop : over) + ORC_NPK * p, k1, k2, pv);
            sk_table(tt[3 * p], tt[3 * p + 1], tt[3 * p + 2], pv, k1 != k2, e);
            double *m = pass == 0 ? hp : sp;
            for (int j = 0; j < kb2; ++j) {
                int col = kk2 + j + 1;                 /* 1-based column */
                size_t base = (size_t)(col - 1) * col / 2;
                int imax = (kk2 > kk1) ? kb1 : j + 1;
                for (int i = 0; i < imax; ++i) {
                    size_t mij = base + kk1 + i;       /* 0-based packed */
                    m[2 * mij] = m[2 * mij] + e[i][j] * cs;
                    m[2 * mij + 1] = m[2 * mij + 1] + e[i][j] * sn;
                }
            }
        }
    }
    /* on-site (setham.f:399-453) */
    for (int iat = 0; iat < natoms; ++iat) {
        int kk = ksk[iat], kb = kbas[kkind[iat] - 1];
        for (int j = 1; j <= kb; ++j) {
            int kkj = kk + j;
            size_t d = (size_t)kkj * (kkj + 1) / 2 - 1;
            int cls = (j == 1) ? 0 : (j <= 4) ? 1 : (j <= 7) ? 2 : 3;
            hp[2 * d] = hp[2 * d] + paros[4 * iat + cls];
            if (jsover && sp && onsite_s) sp[2 * d] = sp[2 * d] + 1.0;
        }
    }
}

void orc_unpack(int n, const double *ap, double *a)
{
    for (int j = 0; j < n; ++j)
        for (int i = 0; i <= j; ++i) {
            size_t k = (size_t)j * (j + 1) / 2 + i;
            a[2 * ((size_t)j * n + i)] = ap[2 * k];
            a[2 * ((size_t)j * n + i) + 1] = ap[2 * k + 1];
            a[2 * ((size_t)i * n + j)] = ap[2 * k];
            a[2 * ((size_t)i * n + j) + 1] = -ap[2 * k + 1];
        }
    /* zhpgv reads only the real part of diagonal entries */
    for (int j = 0; j < n; ++j) a[2 * ((size_t)j * n + j) + 1] = 0.0;
}

/* ------------------------------------------------------------------ */
/* Generalized Hermitian eigensolver (zhpgv algorithm, full storage)   */
/* ------------------------------------------------------------------ */
#define RE(a, i, j, n) (a)[2 * ((size_t)(j) * (n) + (i))]
#define IM(a, i, j, n) (a)[2 * ((size_t)(j) * (n) + (i)) + 1]

/* Implicit QL with Wilkinson shifts on a real symmetric tridiagonal
 * (d[0..n-1], e[0..n-2]), accumulating rotations into v (n x n real,
 * column-major); the algorithm zsteqr (lapack.f:4673) uses.  Returns 0 or
 * the index of a non-converged eigenvalue. */
static int tql2(int n, double *d, double *e, double *v)
{
    double *ee = (double *)calloc((size_t)n + 1, sizeof(double));
    for (int i = 0; i < n - 1; ++i) ee[i] = e[i];
    ee[n - 1] = 0.0;
    const double epsm = 2.220446049250313e-16;
    for (int l = 0; l < n; ++l) {
        int iter = 0;
        for (;;) {
            int m;
            for (m = l; m < n - 1; ++m) {
                double dd = fabs(d[m]) + fabs(d[m + 1]);
                if (fabs(ee[m]) <= epsm * dd) break;
            }
            if (m == l) break;
            if (++iter > 60) { free(ee); return l + 1; }
            double g = (d[l + 1] - d[l]) / (2.0 * ee[l]);
            double r = hypot(g, 1.0);
            g = d[m] - d[l] + ee[l] / (g + (g >= 0 ? fabs(r) : -fabs(r)));
            double s = 1.0, c = 1.0, p = 0.0;
            int i;
            for (i = m - 1; i >= l; --i) {
                double f = s * ee[i], b = c * ee[i];
                r = hypot(f, g);
                ee[i + 1] = r;
                if (r == 0.0) { d[i + 1] -= p; ee[m] = 0.0; break; }
                s = f / r; c = g / r;
                g = d[i + 1] - p;
                r = (d[i] - g) * s + 2.0 * c * b;
                p = s * r;
                d[i + 1] = g + p;
                g = c * r - b;
                for (int k = 0; k < n; ++k) {
                    double *vk1 = v + (size_t)(i + 1) * n + k, *vk = v + (size_t)i * n + k;
                    f = *vk1;
                    *vk1 = s * *vk + c * f;
                    *vk = c * *vk - s * f;
                }
            }
            if (r == 0.0 && i >= l) continue;
            d[l] -= p; ee[l] = g; ee[m] = 0.0;
        }
    }
    /* selection sort ascending, as zsteqr ends */
    for (int i = 0; i < n - 1; ++i) {
        int k = i; double pmin = d[i];
        for (int j = i + 1; j < n; ++j) if (d[j] < pmin) { k = j; pmin = d[j]; }
        if (k != i) {
            d[k] = d[i]; d[i] = pmin;
            for (int r = 0; r < n; ++r) {
                double t = v[(size_t)i * n + r];
                v[(size_t)i * n + r] = v[(size_t)k * n + r];
                v[(size_t)k * n + r] = t;
            }
        }
    }
    free(ee);
    return 0;
}

int orc_zhegv(int n, double *h, double *s, double *w, double *z)
{
    /* 1. S = U^H U, U upper (zpptrf, lapack.f:4466-4645) */
    for (int j = 0; j < n; ++j) {
        for (int i = 0; i < j; ++i) {
            double xr = RE(s, i, j, n), xi = IM(s, i, j, n);
            for (int k = 0; k < i; ++k) {
                double ar = RE(s, k, i, n), ai = -IM(s, k, i, n);
                double br = RE(s, k, j, n), bi = IM(s, k, j, n);
                xr -= ar * br - ai * bi; xi -= ar * bi + ai * br;
            }
            double dii = RE(s, i, i, n);
            RE(s, i, j, n) = xr / dii; IM(s, i, j, n) = xi / dii;
        }
        double ajj = RE(s, j, j, n);
        for (int k = 0; k < j; ++k)
            ajj -= RE(s, k, j, n) * RE(s, k, j, n) + IM(s, k, j, n) * IM(s, k, j, n);
        if (ajj <= 0.0) return j + 1;
        RE(s, j, j, n) = sqrt(ajj); IM(s, j, j, n) = 0.0;
    }
    /* 2. C = U^-H H U^-1 (zhpgst itype 1, lapack.f:2058) via two
     *    triangular solves on full storage. X = H U^-1: X U = H. */
    for (int j = 0; j < n; ++j) {
        for (int k = 0; k < j; ++k) {
            double ur = RE(s, k, j, n), ui = IM(s, k, j, n);
            for (int i = 0; i < n; ++i) {
                double xr = RE(h, i, k, n), xi = IM(h, i, k, n);
                RE(h, i, j, n) -= xr * ur - xi * ui;
                IM(h, i, j, n) -= xr * ui + xi * ur;
            }
        }
        double ujj = RE(s, j, j, n);
        for (int i = 0; i < n; ++i) { RE(h, i, j, n) /= ujj; IM(h, i, j, n) /= ujj; }
    }
    /* U^H C = X, row by row */
    for (int i = 0; i < n; ++i) {
        for (int k = 0; k < i; ++k) {
            double ur = RE(s, k, i, n), ui = -IM(s, k, i, n); /* conj(U[k,i]) */
            for (int j = 0; j < n; ++j) {
                double cr = RE(h, k, j, n), ci = IM(h, k, j, n);
                RE(h, i, j, n) -= ur * cr - ui * ci;
                IM(h, i, j, n) -= ur * ci + ui * cr;
            }
        }
        double uii = RE(s, i, i, n);
        for (int j = 0; j < n; ++j) { RE(h, i, j, n) /= uii; IM(h, i, j, n) /= uii; }
    }
    /* Hermitian part from the lower triangle */
    for (int j = 0; j < n; ++j) {
        IM(h, j, j, n) = 0.0;
        for (int i = j + 1; i < n; ++i) {
            RE(h, j, i, n) = RE(h, i, j, n); IM(h, j, i, n) = -IM(h, i, j, n);
        }
    }
    /* 3. Householder tridiagonalisation, lower (zhetd2 algorithm) */
    double *d = (double *)malloc(sizeof(double) * n);
    double *e = (double *)malloc(sizeof(double) * (n > 1 ? n : 1));
    double *tau = (double *)calloc(2 * (size_t)n, sizeof(double));
    double *wv = (double *)malloc(2 * sizeof(double) * n);
    for (int j = 0; j < n - 1; ++j) {
        int m = n - j - 1;                    /* length of v */
        double alr = RE(h, j + 1, j, n), ali = IM(h, j + 1, j, n);
        double xn = 0.0;
        for (int i = j + 2; i < n; ++i)
            xn += RE(h, i, j, n) * RE(h, i, j, n) + IM(h, i, j, n) * IM(h, i, j, n);
        xn = sqrt(xn);
        double tr = 0.0, ti = 0.0, beta;
        if (xn == 0.0 && ali == 0.0) {
            beta = alr;
        } else {
            beta = -copysign(sqrt(alr * alr + ali * ali + xn * xn), alr);
            tr = (beta - alr) / beta; ti = -ali / beta;
            /* v = x / (alpha - beta) */
            double dr = alr - beta, di = ali, den = dr * dr + di * di;
            double sr = dr / den, si = -di / den;   /* 1/(alpha-beta) */
            for (int i = j + 2; i < n; ++i) {
                double xr = RE(h, i, j, n), xi = IM(h, i, j, n);
                RE(h, i, j, n) = xr * sr - xi * si;
                IM(h, i, j, n) = xr * si + xi * sr;
            }
        }
        e[j] = beta;
        tau[2 * j] = tr; tau[2 * j + 1] = ti;
        RE(h, j + 1, j, n) = 1.0; IM(h, j + 1, j, n) = 0.0;
        if (tr != 0.0 || ti != 0.0) {
            /* w = tau * A v (A = trailing block rows/cols j+1..n-1) */
            for (int r = 0; r < m; ++r) {
                double ar = 0, ai = 0;
                for (int c = 0; c < m; ++c) {
                    double mr = RE(h, j + 1 + r, j + 1 + c, n), mi = IM(h, j + 1 + r, j + 1 + c, n);
                    double vr = RE(h, j + 1 + c, j, n), vi = IM(h, j + 1 + c, j, n);
                    ar += mr * vr - mi * vi; ai += mr * vi + mi * vr;
                }
                wv[2 * r] = tr * ar - ti * ai; wv[2 * r + 1] = tr * ai + ti * ar;
            }
            /* alpha = -1/2 tau (w^H v) */
            double pr = 0, pi = 0;
            for (int r = 0; r < m; ++r) {
                double a_r = wv[2 * r], a_i = -wv[2 * r + 1];
                double vr = RE(h, j + 1 + r, j, n), vi = IM(h, j + 1 + r, j, n);
                pr += a_r * vr - a_i * vi; pi += a_r * vi + a_i * vr;
            }
            double alphr = -0.5 * (tr * pr - ti * pi), alphi = -0.5 * (tr * pi + ti * pr);
            for (int r = 0; r < m; ++r) {
                double vr = RE(h, j + 1 + r, j, n), vi = IM(h, j + 1 + r, j, n);
                wv[2 * r] += alphr * vr - alphi * vi;
                wv[2 * r + 1] += alphr * vi + alphi * vr;
            }
            /* A -= v w^H + w v^H */
            for (int c = 0; c < m; ++c)
                for (int r = 0; r < m; ++r) {
                    double vr = RE(h, j + 1 + r, j, n), vi = IM(h, j + 1 + r, j, n);
                    double wcr = wv[2 * c], wci = -wv[2 * c + 1];
                    double wr = wv[2 * r], wi = wv[2 * r + 1];
                    double vcr = RE(h, j + 1 + c, j, n), vci = -IM(h, j + 1 + c, j, n);
                    RE(h, j + 1 + r, j + 1 + c, n) -= (vr * wcr - vi * wci) + (wr * vcr - wi * vci);
                    IM(h, j + 1 + r, j + 1 + c, n) -= (vr * wci + vi * wcr) + (wr * vci + wi * vcr);
                }
        }
        d[j] = RE(h, j, j, n);
    }
    d[n - 1] = RE(h, n - 1, n - 1, n);
    /* 4. eigen-decomposition of T */
    double *v = (double *)calloc((size_t)n * n, sizeof(double));
    for (int i = 0; i < n; ++i) v[(size_t)i * n + i] = 1.0;
    int info = tql2(n, d, e, v);
    for (int i = 0; i < n; ++i) w[i] = d[i];
    if (z && !info) {
        /* 5. Y = Q V, Q = H(0) H(1) ... H(n-2), H(j) = I - tau v v^H */
        for (int c = 0; c < n; ++c)
            for (int r = 0; r < n; ++r) { RE(z, r, c, n) = v[(size_t)c * n + r]; IM(z, r, c, n) = 0.0; }
        for (int j = n - 2; j >= 0; --j) {
            double tr = tau[2 * j], ti = tau[2 * j + 1];
            if (tr == 0.0 && ti == 0.0) continue;
            for (int c = 0; c < n; ++c) {
                double sr = 0, si = 0;            /* v^H y */
                for (int r = j + 1; r < n; ++r) {
                    double vr = RE(h, r, j, n), vi = -IM(h, r, j, n);
                    double yr = RE(z, r, c, n), yi = IM(z, r, c, n);
                    sr += vr * yr - vi * yi; si += vr * yi + vi * yr;
                }
                double fr = tr * sr - ti * si, fi = tr * si + ti * sr;
                for (int r = j + 1; r < n; ++r) {
                    double vr = RE(h, r, j, n), vi = IM(h, r, j, n);
                    RE(z, r, c, n) -= vr * fr - vi * fi;
                    IM(z, r, c, n) -= vr * fi + vi * fr;
                }
            }
        }
        /* 6. Z = U^-1 Y (ztpsv, lapack.f:2445) */
        for (int c = 0; c < n; ++c)
            for (int i = n - 1; i >= 0; --i) {
                double xr = RE(z, i, c, n), xi = IM(z, i, c, n);
                for (int k = i + 1; k < n; ++k) {
                    double ur = RE(s, i, k, n), ui = IM(s, i, k, n);
                    double yr = RE(z, k, c, n), yi = IM(z, k, c, n);
                    xr -= ur * yr - ui * yi; xi -= ur * yi + ui * yr;
                }
                double uii = RE(s, i, i, n);
                RE(z, i, c, n) = xr / uii; IM(z, i, c, n) = xi / uii;
            }
    }
    free(v); free(d); free(e); free(tau); free(wv);
    return info ? n + info : 0;
}

/* Eigenvalues (ascending) of a small Hermitian matrix via cyclic Jacobi on
 * the real symmetric embedding [[Re,-Im],[Im,Re]] (each eigenvalue twice). */
static void small_heev(int n, const double *a /*n*n complex col-major*/, double *w)
{
    int m = 2 * n;
    double b[18 * 18];
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            double ar = a[2 * (j * n + i)], ai = a[2 * (j * n + i) + 1];
            b[j * m + i] = ar;           b[(j + n) * m + i + n] = ar;
            b[j * m + i + n] = ai;       b[(j + n) * m + i] = -ai;
        }
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0;
        for (int j = 0; j < m; ++j) for (int i = 0; i < m; ++i) if (i != j) off += b[j * m + i] * b[j * m + i];
        if (off < 1e-300) break;
        for (int p = 0; p < m - 1; ++p)
            for (int q = p + 1; q < m; ++q) {
                double apq = b[q * m + p];
                if (apq == 0.0) continue;
                double theta = (b[q * m + q] - b[p * m + p]) / (2.0 * apq);
                double t = copysign(1.0, theta) / (fabs(theta) + sqrt(theta * theta + 1.0));
                double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < m; ++k) {
                    double bkp = b[p * m + k], bkq = b[q * m + k];
                    b[p * m + k] = c * bkp - s * bkq; b[q * m + k] = s * bkp + c * bkq;
                }
                for (int k = 0; k < m; ++k) {
                    double bpk = b[k * m + p], bqk = b[k * m + q];
                    b[k * m + p] = c * bpk - s * bqk; b[k * m + q] = s * bpk + c * bqk;
                }
            }
    }
    double ev[18];
    for (int i = 0; i < m; ++i) ev[i] = b[i * m + i];
    for (int i = 1; i < m; ++i) { double x = ev[i]; int j = i - 1; while (j >= 0 && ev[j] > x) { ev[j + 1] = ev[j]; --j; } ev[j + 1] = x; }
    for (int i = 0; i < n; ++i) w[i] = 0.5 * (ev[2 * i] + ev[2 * i + 1]);
}

/* perturb.f:87-221 */
int orc_perturb(int n, const double *dhp, const double *dsp, const double *z,
                const double *eig, double eigcut, int *ntot, double *deig)
{
    const double tol = 1e-8;
    double *work = (double *)malloc(2 * sizeof(double) * n);
    int start = 0;
    while (start < n) {
        if (eig[start] > eigcut) { *ntot = start; free(work); return 0; }
        int num = n - start;
        for (int j = start + 1; j < n; ++j)
            if (fabs(eig[j] - eig[start]) > tol) { num = j - start; break; }
        if (num > 9) { free(work); return 1; }
        double es = eig[start];
        double hd[2 * 81];
        for (int i = 0; i < num; ++i) {
            int is = start + i;
            for (int j = 0; j < n; ++j) {            /* work = (dH - e dS) z */
                double cr = 0, ci = 0;
                for (int k = 0; k < n; ++k) {
                    size_t mjk; double ar, ai;
                    if (k < j) { mjk = (size_t)j * (j + 1) / 2 + k;
                        ar = dhp[2 * mjk] - es * dsp[2 * mjk];
                        ai = -(dhp[2 * mjk + 1] - es * dsp[2 * mjk + 1]); }
                    else { mjk = (size_t)k * (k + 1) / 2 + j;
                        ar = dhp[2 * mjk] - es * dsp[2 * mjk];
                        ai = dhp[2 * mjk + 1] - es * dsp[2 * mjk + 1]; }
                    double zr = RE(z, k, is, n), zi = IM(z, k, is, n);
                    cr += ar * zr - ai * zi; ci += ar * zi + ai * zr;
                }
                work[2 * j] = cr; work[2 * j + 1] = ci;
            }
            for (int j = 0; j <= i; ++j) {
                int js = start + j;
                double cr = 0, ci = 0;
                for (int k = 0; k < n; ++k) {
                    double zr = RE(z, k, js, n), zi = -IM(z, k, js, n);
                    cr += zr * work[2 * k] - zi * work[2 * k + 1];
                    ci += zr * work[2 * k + 1] + zi * work[2 * k];
                }
                /* hd(j,i) of the upper-packed num x num matrix, stored full */
                hd[2 * (i * num + j)] = cr; hd[2 * (i * num + j) + 1] = ci;
                hd[2 * (j * num + i)] = cr; hd[2 * (j * num + i) + 1] = -ci;
            }
        }
        if (num == 1) {
            deig[start] = hd[0];
        } else if (num == 2) {
            double h1 = hd[0], h3 = hd[2 * 3], h2r = hd[2 * 2], h2i = hd[2 * 2 + 1];
            double sq = sqrt((h1 - h3) * (h1 - h3) + 4.0 * (h2r * h2r + h2i * h2i));
            deig[start] = 0.5 * ((h1 + h3) - sq);
            deig[start + 1] = 0.5 * ((h1 + h3) + sq);
        } else {
            small_heev(num, hd, deig + start);
        }
        start += num;
    }
    *ntot = n;
    free(work);
    return 0;
}

/* ------------------------------------------------------------------ */
/* fermi.f:163-433                                                     */
/* ------------------------------------------------------------------ */
double orc_fermi(int ldn, int nk, int nv, double el, const double *weight,
                 const double *eigen, double tkb, double zfill)
{
    const double nearz = 1e-5, tol = 1e-8, twenty = 20.0, half = 0.5;
    double beta = (tkb < nearz) ? 1.0 / nearz : 1.0 / tkb;
    int nfill = (int)(zfill / el + nearz);
    int linst = fabs(zfill / el - (double)nfill) < tol;
#define EIG(i, k) eigen[(size_t)(k) * ldn + (i) - 1]
/* levels above nv: rows no k-point wrote, zero in the reference's static
   eigenvalue storage (full bands, or one electron per state) */
#define EIGZ(i, k) ((i) <= nv ? EIG(i, k) : 0.0)
    if (nfill < 1) nfill = 1;          /* guards for reading eigen(nfill+1) */
    double muhigh = EIG(nv, 0), mulow = EIG(1, 0);
    double stlow = EIGZ(nfill, 0), sthigh = EIGZ(nfill + 1, 0);
    for (int k = 0; k < nk; ++k) {
        if (EIGZ(nfill, k) > stlow) stlow = EIGZ(nfill, k);
        if (EIGZ(nfill + 1, k) < sthigh) sthigh = EIGZ(nfill + 1, k);
        for (int i = 1; i <= nv; ++i) {
            if (EIG(i, k) > muhigh) muhigh = EIG(i, k);
            if (EIG(i, k) < mulow) mulow = EIG(i, k);
        }
    }
    if (linst && (sthigh - stlow > twenty * tkb)) return half * (stlow + sthigh);
    double mu = half * (stlow + sthigh);
    for (;;) {
        double nsum = 0.0, dnsum = 0.0;
        double mum20 = mu - tkb * twenty, mup20 = mu + tkb * twenty;
        double btmu = -beta * mu;
        for (int k = 0; k < nk; ++k) {
            double nks = 0.0, dnks = 0.0;
            for (int i = 1; i <= nv; ++i) {
                double ev = EIG(i, k);
                if (ev < mum20) nks = nks + 1.0;
                else if (ev < mup20) {
                    double ex = exp(beta * ev + btmu);
                    double f = 1.0 / (1.0 + ex);
                    nks = nks + f;
                    dnks = dnks + ex * f * f;
                } else break;
            }
            nsum = nsum + weight[k] * nks;
            dnsum = dnsum + weight[k] * dnks;
        }
        nsum = el * nsum;
        dnsum = beta * el * dnsum;
        if (fabs(nsum - zfill) < tol) return mu;
        if (nsum < zfill) mulow = mu; else muhigh = mu;
        double mu1;
        if (dnsum < tol) mu1 = half * (mulow + muhigh);
        else {
            mu1 = mu + (zfill - nsum) / dnsum;
            if (mu1 > muhigh) mu1 = half * (mu + muhigh);
            else if (mu1 < mulow) mu1 = half * (mu + mulow);
        }
        if (muhigh - mulow > tol) { mu = mu1; continue; }
        return mu;
    }
#undef EIGZ
#undef EIG
}

/* ------------------------------------------------------------------ */
/* bande.f:1057-1477 (jspins = 1, collinear)                           */
/* ------------------------------------------------------------------ */
void orc_bande_sums(int mode, int ldn, int nk, int nv, double el,
                    const double *weight, const double *esk,
                    const double *desk, const int *ncut, double tkb,
                    double zfill, double *efermi_o, double *earg1,
                    double *earg2)
{
    const double nearz = 1e-5, small = 1e-10, one = 1.0, half = 0.5;
    int lpress = (mode == 2) || (mode == 4);
    int lentro = mode < 3;
    double efermi = orc_fermi(ldn, nk, nv, el, weight, esk, tkb, zfill);
    double tkbi = (tkb < nearz) ? 1.0 / nearz : one / tkb;
#define FB(x) (one / (one + exp(tkbi * (x))))
#define DFB(x) (-tkbi * FB(x) * FB(-(x)))
    double etol = 20.0 * tkb, efp = efermi + etol, efm = efermi - etol;
    double esum = 0, desum1 = 0, desum2 = 0, desum3 = 0, dfbsum = 0, dferm = 0;
    double entropy = 0, dent1 = 0, dent2 = 0;
    for (int k = 0; k < nk; ++k) {
        double wt = el * weight[k];
        int ni = lpress ? ncut[k] : nv;
        for (int i = 0; i < ni; ++i) {
            double e = esk[(size_t)k * ldn + i];
            if (!(e < efp)) continue;
            double ewt = wt * e;
            if (lpress) {
                double de = desk[(size_t)k * ldn + i];
                if (e < efm) { esum = esum + ewt; desum1 = desum1 + wt * de; }
                else {
                    double fbe = FB(e - efermi);
                    esum = esum + ewt * fbe;
                    double dfbe = DFB(e - efermi);
                    dfbsum = dfbsum + dfbe * wt;
                    desum3 = desum3 + dfbe * ewt;
                    double dewt = wt * de;
                    desum1 = desum1 + dewt * fbe;
                    dferm = dferm + dfbe * dewt;
                    desum2 = desum2 + dfbe * ewt * de;
                    if (lentro) {
                        entropy = entropy - wt * (fbe * log(fbe) + (one - fbe) * log(one - fbe));
                        double con1 = dfbe * wt * (e - efermi);
                        dent1 = dent1 + con1 * de;
                        dent2 = dent2 + con1;
                    }
                }
            } else {
                if (e < efm) esum = esum + ewt;
                else {
                    double fbe = FB(e - efermi);
                    if (lentro) entropy = entropy - wt * (fbe * log(fbe) + (one - fbe) * log(one - fbe));
                    esum = esum + ewt * fbe;
                }
            }
        }
    }
    *efermi_o = efermi;
    double e1 = 0, e2 = 0;
    if (lentro) {
        double fr = esum - tkb * entropy;
        e1 = half * (esum + fr);
        e2 = esum;
    } else {
        e1 = esum;
    }
    if (lpress) {
        double dfermi = (fabs(dfbsum) > small) ? dferm / dfbsum : 0.0;
        double press = -(desum1 + (desum2 - dfermi * desum3));
        if (lentro) {
            double pent = dfermi * dent2 - dent1;
            double fpress = press - pent;
            e2 = half * (press + fpress);
        } else {
            e2 = press;
        }
    }
    *earg1 = e1; *earg2 = e2;
#undef FB
#undef DFB
}

int orc_bande(int mode, int nv, int natoms, const int *kkind, const int *ksk,
              const int *kbas, const double *posq, int npairs,
              const double *tt, const double *dlv, const int *jkind,
              const int *jatm, const double *hop, const double *over,
              const double *paros, const double *dhop, const double *dover,
              const double *dparos, int nk, const double *kpts,
              const double *weight, double zfill, double tkb, double eigcut,
              int nthreads, double *efermi, double *earg1, double *earg2,
              double *eig_out, double *desk_out)
{
    int lpress = (mode == 2) || (mode == 4);
    double *esk = (double *)calloc((size_t)nk * nv, sizeof(double));
    double *desk = (double *)calloc((size_t)nk * nv, sizeof(double));
    int *ncut = (int *)calloc((size_t)nk, sizeof(int));
    int err = 0;
    size_t np = (size_t)nv * (nv + 1) / 2;
#ifdef _OPENMP
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
#endif
    for (int k = 0; k < nk; ++k) {
        double *hp = (double *)malloc(2 * np * sizeof(double));
        double *sp = (double *)malloc(2 * np * sizeof(double));
        double *hf = (double *)malloc(2 * sizeof(double) * nv * nv);
        double *sf = (double *)malloc(2 * sizeof(double) * nv * nv);
        double *zf = lpress ? (double *)malloc(2 * sizeof(double) * nv * nv) : NULL;
        orc_setham(nv, kpts + 3 * k, natoms, kkind, ksk, kbas, posq, npairs,
                   tt, dlv, jkind, jatm, hop, over, paros, 1, 1, hp, sp);
        orc_unpack(nv, hp, hf);
        orc_unpack(nv, sp, sf);
        int info;
        if (nv == 1) {
            /* diag.f:320-349: the 1 x 1 pencil in closed form, e = Re h / Re s
             * (s <= 0 flags ioverr); no eigenvector is formed, so perturb
             * sees zeig = 0 and the derivative is zero (ncut counts the
             * state when it lies below eigcut) */
            const double h1 = hp[0], s1 = sp[0];
            info = s1 > 0.0 ? 0 : 1;
            esk[(size_t)k * nv] = h1 / s1;
            if (!info && lpress) {
                desk[(size_t)k * nv] = 0.0;
                ncut[k] = esk[(size_t)k * nv] > eigcut ? 0 : 1;
            }
        } else {
            info = orc_zhegv(nv, hf, sf, esk + (size_t)k * nv, zf);
        }
        if (info) {
#ifdef _OPENMP
#pragma omp atomic write
#endif
            err = 3;
        } else if (lpress && nv > 1) {
            orc_setham(nv, kpts + 3 * k, natoms, kkind, ksk, kbas, posq, npairs,
                       tt, dlv, jkind, jatm, dhop, dover, dparos, 1, 0, hp, sp);
            if (orc_perturb(nv, hp, sp, zf, esk + (size_t)k * nv, eigcut,
                            ncut + k, desk + (size_t)k * nv)) {
#ifdef _OPENMP
#pragma omp atomic write
#endif
                err = 5;
            }
        }
        free(hp); free(sp); free(hf); free(sf); free(zf);
    }
    if (!err)
        orc_bande_sums(mode, nv, nk, nv, 2.0, weight, esk, desk, ncut, tkb,
                       zfill, efermi, earg1, earg2);
    if (eig_out) memcpy(eig_out, esk, sizeof(double) * nk * nv);
    if (desk_out) memcpy(desk_out, desk, sizeof(double) * nk * nv);
    free(esk); free(desk); free(ncut);
    return err;
}
