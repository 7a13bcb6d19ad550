// tbk_setup.cpp -- k-independent host set-up for the k-point engine.
//
// Neighbour list (search2.f), density-dependent on-site terms and bond /
// overlap integrals (setpar.f), their volume derivatives (setparv.f) and the
// per-pair Slater-Koster tables of rotate.f.  Compiled with
// -ffp-contract=off and without -march so that every value carries the
// same IEEE operation sequence as the reference: the neighbour list and the
// parameters are bit-identical to those of static v1.33.
#include <cmath>
#include <cstring>
#include <vector>

#include "tbk.h"
#include "tbk_internal.h"

namespace tbk {

namespace {
constexpr int kOnsite = 4, kLike = 10, kUnlike = 14;   // npkind (static.f:610)

// 0-based parameter offsets of setpar.f:461, 558-572, 655-666.
int onsite_base(int kind) { return (kind - 1) * (1 + kOnsite * 4); }
int hop_base(int kinds, int lo, int hi)
{
    int head = kinds * (1 + kOnsite * (3 * kinds + 1));
    if (lo == hi) return head + 4 * kLike * (lo - 1);
    return head + 4 * kLike * kinds + 4 * kUnlike * ((hi - 1) * (hi - 2) / 2 + lo - 1);
}
int over_base(int kinds, int lo, int hi)
{
    int head = kinds * (1 + kOnsite * (3 * kinds + 1)) + 4 * kUnlike * (kinds * (kinds - 1) / 2);
    if (lo == hi) return head + 4 * kLike * kinds + 4 * kLike * (lo - 1);
    return head + 8 * kLike * kinds + 4 * kUnlike * ((hi - 1) * (hi - 2) / 2 + lo - 1);
}
const double kOvlNull[10] = {1, 0, 1, 1, 0, 0, 0, 1, 1, 1};   // setpar.f:33
const double kHarSign[9] = {-1, 1, 1, -1, -1, -1, 1, -1, 1};  // setparv.f:412

// Fortran param(i) with 1-based i.
struct ParamView {
    const double *p;
    double operator()(int i) const { return p[i - 1]; }
};
}  // namespace

int search2(int natoms, const double *lat, const double *posn, const int *kkind,
            int kinds, const int *nd, const double *dnn, const double *screenl,
            int maxpairs, PairList &out)
{
    out.clear();
    const double eps = 1e-3;
    const int kd0 = 2 * nd[0] + 1, kd1 = 2 * nd[1] + 1, kd2 = 2 * nd[2] + 1;
    // Pass 1, one atom ia per task (host threads): every candidate within
    // the cutoff, in the reference's order, with the pair values and its
    // effect on the kneigh flag of COMMON /relat/ (search2.f:168-175):
    // 1 sets it, 0 clears it, -1 (distance exactly at the cutoff) keeps it.
    struct Cand { Pair p; signed char eff; };
    std::vector<std::vector<Cand>> per(natoms);
    std::vector<double> dmin(natoms, 1e10);
    parallel_for(natoms, [&](int ia) {
        std::vector<Cand> &list = per[ia];
        double dm = 1e10;
        for (int ja = ia; ja < natoms; ++ja) {
            const int ki = kkind[ia], kj = kkind[ja];
            int itype;
            if (ki == kj) {
                itype = ki;
            } else {
                const int lo = ki < kj ? ki : kj, hi = ki < kj ? kj : ki;
                itype = kinds + (hi - 1) * (hi - 2) / 2 + lo;
            }
            const double rc = dnn[itype - 1], ell = screenl[itype - 1];
            const double rcsq = rc * rc;
            const double ellinv = 1.0 / ell;
            const double rscreen = rc - 5.0 * ell;
            for (int ix = 0; ix < kd0; ++ix)
            for (int iy = 0; iy < kd1; ++iy)
            for (int iz = 0; iz < kd2; ++iz) {
                const int m[3] = {ix - nd[0], iy - nd[1], iz - nd[2]};
                double t[3];
                double dsq = 0.0;
                for (int c = 0; c < 3; ++c) {
                    double r = 0.0;
                    r = r + lat[0 * 3 + c] * (double)m[0];
                    r = r + lat[1 * 3 + c] * (double)m[1];
                    r = r + lat[2 * 3 + c] * (double)m[2];
                    t[c] = posn[3 * ja + c] + r - posn[3 * ia + c];
                    dsq = dsq + t[c] * t[c];
                }
                const double dd = dsq - eps;
                if (dd > rcsq) continue;
                Cand cd{};
                cd.eff = dd < eps ? 0 : (dd < rcsq ? 1 : -1);
                if (dd > eps && dsq < dm) dm = dsq;
                if (cd.eff != 0) {
                    const double r = std::sqrt(dsq);
                    const double rinv = 1.0 / r;
                    const double ex = std::exp((r - rscreen) * ellinv);
                    const double f = 1.0 / (ex + 1.0);
                    for (int c = 0; c < 3; ++c) {
                        cd.p.cosines[c] = t[c] * rinv;
                        cd.p.dlv[c] = (double)m[c];
                    }
                    cd.p.dist = r;
                    cd.p.screen = f;
                    cd.p.dscrn = -ex * f * ellinv;
                    cd.p.kind[0] = ki; cd.p.kind[1] = kj;
                    cd.p.atom[0] = ia; cd.p.atom[1] = ja;
                }
                list.push_back(cd);
            }
        }
        dmin[ia] = dm;
    });
    // Pass 2, in the reference's order: the flag (that of a fresh process
    // at the start) decides which candidates become pairs.
    int kneigh = 0;
    double dismin2 = 1e10;
    for (int ia = 0; ia < natoms; ++ia) {
        dismin2 = std::min(dismin2, dmin[ia]);
        for (const Cand &cd : per[ia]) {
            if (cd.eff >= 0) kneigh = cd.eff;
            if (!kneigh) continue;
            if (out.n >= maxpairs) return TBK_ERR_PAIRS;
            out.push(cd.p);
        }
    }
    out.dismin = std::sqrt(dismin2);
    return TBK_OK;
}

// setpar.f:398-884, collinear.  param: the channel of jspin (bond and
// overlap integrals); an atom with flip[a] != 0 takes its density decay and
// its on-site terms from param_flip, the other channel (setpar.f:500-536,
// 795-867: jsfrm = 3 - jspin).  flip may be null (no flips).
int setpar(int natoms, const int *kkind, int kinds, const double *param, int nltype,
           int jsover, const PairList &pl, double *den, double *paros, double *hop,
           double *over, const double *param_flip, const int *flip)
{
    const ParamView P{param};
    auto site = [&](int a) { return ParamView{flip && flip[a] ? param_flip : param}; };
    const bool realov = nltype == 1 || nltype == 3;
    const bool harrison = nltype == 2 || nltype == 3;
    const double twtrd = 2.0 / 3.0;
    for (int i = 0; i < natoms * kinds; ++i) den[i] = 0.0;
    // densities, both directions, summed in pair order (setpar.f:508-536)
    for (int ip = 0; ip < pl.n; ++ip) {
        const Pair &pr = pl.pairs[ip];
        const int ia = pr.atom[0], ja = pr.atom[1];
        const int ki = pr.kind[0], kj = pr.kind[1];
        const double r = pr.dist, f = pr.screen;
        const int i0 = onsite_base(kj) + 1;
        const ParamView Pi = site(ia);
        const double b = Pi(i0) * Pi(i0);
        den[ia * kinds + kj - 1] = den[ia * kinds + kj - 1] + std::exp(-b * r) * f;
        if (ja != ia) {
            const int i1 = onsite_base(ki) + 1;
            const ParamView Pj = site(ja);
            const double b1 = Pj(i1) * Pj(i1);
            den[ja * kinds + ki - 1] = den[ja * kinds + ki - 1] + std::exp(-b1 * r) * f;
        }
    }
    // bond and overlap integrals, each pair on its own (chunks of pairs on
    // host threads)
    constexpr int kChunk = 512;
    parallel_for((pl.n + kChunk - 1) / kChunk, [&](int ch) {
    const int ipe = std::min(pl.n, (ch + 1) * kChunk);
    for (int ip = ch * kChunk; ip < ipe; ++ip) {
        const Pair &pr = pl.pairs[ip];
        const int ki = pr.kind[0], kj = pr.kind[1];
        const double r = pr.dist, f = pr.screen;
        const int lo = ki < kj ? ki : kj, hi = ki < kj ? kj : ki;
        const int npk = lo == hi ? kLike : kUnlike;
        double *h = hop + 14 * ip;
        double *s = over + 14 * ip;
        for (int j = 0; j < 14; ++j) { h[j] = 0.0; s[j] = 0.0; }
        int b = hop_base(kinds, lo, hi);
        if (nltype != 90000) {
            for (int j = 1; j <= npk; ++j) {
                const double poly = P(b + 4 * j - 3) + r * (P(b + 4 * j - 2) + r * P(b + 4 * j - 1));
                const double g2 = -(P(b + 4 * j) * P(b + 4 * j));
                h[j - 1] = poly * std::exp(g2 * r) * f;
            }
            if (harrison && lo == hi)
                for (int j = 0; j < 9; ++j) h[j] = kHarSign[j] * std::fabs(h[j]);
        } else {
            const double poly = P(b + 1) + r * (P(b + 2) + r * (P(b + 3) + r * (P(b + 5) + r * P(b + 6))));
            const double g2 = -(P(b + 4) * P(b + 4));
            h[0] = poly * std::exp(g2 * r) * f;
        }
        if (jsover) {
            const bool usenew = realov && lo == hi;
            b = over_base(kinds, lo, hi);
            if (nltype != 90000) {
                for (int j = 1; j <= npk; ++j) {
                    double poly = P(b + 4 * j - 3) + r * (P(b + 4 * j - 2) + r * P(b + 4 * j - 1));
                    const double g2 = -(P(b + 4 * j) * P(b + 4 * j));
                    if (usenew) poly = kOvlNull[j - 1] + r * poly;
                    s[j - 1] = poly * std::exp(g2 * r) * f;
                }
                if (harrison && lo == hi)
                    for (int j = 0; j < 9; ++j) s[j] = -kHarSign[j] * std::fabs(s[j]);
            } else {
                const double poly = P(b + 1) + r * (P(b + 2) + r * (P(b + 3) + r * (P(b + 5) + r * P(b + 6))));
                const double g2 = -(P(b + 4) * P(b + 4));
                s[0] = poly * std::exp(g2 * r) * f;
            }
        }
    }
    });
    for (int ia = 0; ia < natoms; ++ia) {      // setpar.f:795-867
        const ParamView Ps = site(ia);
        const int k1 = kkind[ia];
        int b = onsite_base(k1) + 1;
        double v = std::pow(den[ia * kinds + k1 - 1], twtrd);
        for (int c = 0; c < 4; ++c, b += 4)
            paros[4 * ia + c] = Ps(b + 1) + v * (Ps(b + 2) + v * (Ps(b + 3) + v * Ps(b + 4)));
        for (int k2 = 1; k2 <= kinds; ++k2) {
            if (k2 == k1 || den[ia * kinds + k2 - 1] == 0.0) continue;
            v = std::pow(den[ia * kinds + k2 - 1], twtrd);
            const int hi = k1 > k2 ? k1 : k2, lo = k1 < k2 ? k1 : k2;
            b = kinds * (1 + kOnsite * 4) + 6 * kOnsite * ((hi - 1) * (hi - 2) / 2 + lo - 1);
            if (k2 < hi) b += 3 * kOnsite;
            for (int c = 0; c < 4; ++c, b += 3)
                paros[4 * ia + c] = paros[4 * ia + c] + v * (Ps(b + 1) + v * (Ps(b + 2) + v * Ps(b + 3)));
        }
    }
    return TBK_OK;
}

// setparv.f:426-886, collinear; param, param_flip and flip as in setpar
// (setparv.f:507-539, 800-884).
int setparv(int natoms, const int *kkind, int kinds, const double *param, int nltype,
            int jsover, double thisvol, const PairList &pl, const double *den,
            double *dparos, double *dhop, double *dover, const double *param_flip, const int *flip)
{
    const ParamView P{param};
    auto site = [&](int a) { return ParamView{flip && flip[a] ? param_flip : param}; };
    const bool realov = nltype == 1 || nltype == 3;
    const bool harrison = nltype == 2 || nltype == 3;
    const double third = 1.0 / 3.0, twtrd = 2.0 / 3.0, threeh = 3.0 / 2.0;
    std::vector<double> vden((size_t)natoms * kinds, 0.0);
    // density derivatives, summed in pair order (setparv.f:464-544)
    for (int ip = 0; ip < pl.n; ++ip) {
        const Pair &pr = pl.pairs[ip];
        const int ia = pr.atom[0], ja = pr.atom[1];
        const int ki = pr.kind[0], kj = pr.kind[1];
        const double r = pr.dist, f = pr.screen, df = pr.dscrn;
        const double vdist = third * r / thisvol;          // setparv.f:459
        const int i0 = onsite_base(kj) + 1;
        const ParamView Pi = site(ia);
        const double b = Pi(i0) * Pi(i0);
        const double dent = std::exp(-b * r) * f;
        vden[ia * kinds + kj - 1] = vden[ia * kinds + kj - 1] + (df - b) * dent * vdist;
        if (ja != ia) {
            const int i1 = onsite_base(ki) + 1;
            const ParamView Pj = site(ja);
            const double b1 = Pj(i1) * Pj(i1);
            const double dent1 = std::exp(-b1 * r) * f;
            vden[ja * kinds + ki - 1] = vden[ja * kinds + ki - 1] + (df - b1) * dent1 * vdist;
        }
    }
    // derivatives of the bond and overlap integrals, each pair on its own
    constexpr int kChunk = 512;
    parallel_for((pl.n + kChunk - 1) / kChunk, [&](int ch) {
    const int ipe = std::min(pl.n, (ch + 1) * kChunk);
    for (int ip = ch * kChunk; ip < ipe; ++ip) {
        const Pair &pr = pl.pairs[ip];
        const int ki = pr.kind[0], kj = pr.kind[1];
        const double r = pr.dist, f = pr.screen, df = pr.dscrn;
        const double vdist = third * r / thisvol;          // setparv.f:459
        const int lo = ki < kj ? ki : kj, hi = ki < kj ? kj : ki;
        const int npk = lo == hi ? kLike : kUnlike;
        const bool usehar = harrison && lo == hi;
        double *h = dhop + 14 * ip;
        double *s = dover + 14 * ip;
        for (int j = 0; j < 14; ++j) { h[j] = 0.0; s[j] = 0.0; }
        int b = hop_base(kinds, lo, hi);
        if (nltype != 90000) {
            for (int j = 1; j <= npk; ++j) {
                const double poly = P(b + 4 * j - 3) + r * (P(b + 4 * j - 2) + r * P(b + 4 * j - 1));
                const double g2 = -(P(b + 4 * j) * P(b + 4 * j));
                const double val = poly * std::exp(g2 * r) * f;
                h[j - 1] = vdist * (((P(b + 4 * j - 2) + 2.0 * r * P(b + 4 * j - 1)) + g2 * poly) *
                           std::exp(g2 * r) * f + df * val);
                if (usehar && j < npk) h[j - 1] = kHarSign[j - 1] * h[j - 1] * std::copysign(1.0, poly);
            }
        } else {
            const double poly = P(b + 1) + r * (P(b + 2) + r * (P(b + 3) + r * (P(b + 5) + r * P(b + 6))));
            const double g2 = -(P(b + 4) * P(b + 4));
            const double val = poly * std::exp(g2 * r) * f;
            h[0] = vdist * (((P(b + 2) + r * (2.0 * P(b + 3) + r * (3.0 * P(b + 5) +
                   r * (4.0 * P(b + 6))))) + g2 * poly) * std::exp(g2 * r) * f + df * val);
        }
        if (jsover) {
            const bool usenew = realov && lo == hi;
            b = over_base(kinds, lo, hi);
            if (nltype != 9000) {       // setparv.f:640 compares with 9000
                for (int j = 1; j <= npk; ++j) {
                    const double g2 = -(P(b + 4 * j) * P(b + 4 * j));
                    double poly;
                    if (usenew) {
                        poly = kOvlNull[j - 1] + r * (P(b + 4 * j - 3) + r * (P(b + 4 * j - 2) +
                               r * P(b + 4 * j - 1)));
                        const double val = poly * std::exp(g2 * r) * f;
                        s[j - 1] = vdist * (((P(b + 4 * j - 3) + 2.0 * r * P(b + 4 * j - 2) +
                                   3.0 * r * r * P(b + 4 * j - 1)) + g2 * poly) *
                                   std::exp(g2 * r) * f + df * val);
                    } else {
                        poly = P(b + 4 * j - 3) + r * (P(b + 4 * j - 2) + r * P(b + 4 * j - 1));
                        const double val = poly * std::exp(g2 * r) * f;
                        s[j - 1] = vdist * (((P(b + 4 * j - 2) + 2.0 * r * P(b + 4 * j - 1)) +
                                   g2 * poly) * std::exp(g2 * r) * f + df * val);
                    }
                    if (usehar && j < npk) s[j - 1] = -kHarSign[j - 1] * s[j - 1] * std::copysign(1.0, poly);
                }
            } else {
                const double poly = P(b + 1) + r * (P(b + 2) + r * (P(b + 3) + r * (P(b + 5) + r * P(b + 6))));
                const double g2 = -(P(b + 4) * P(b + 4));
                const double val = poly * std::exp(g2 * r) * f;
                s[0] = vdist * (((P(b + 2) + r * (2.0 * P(b + 3) + r * (3.0 * P(b + 5) +
                       r * (4.0 * P(b + 6))))) + g2 * poly) * std::exp(g2 * r) * f + df * val);
            }
        }
    }
    });
    for (int ia = 0; ia < natoms; ++ia) {
        const ParamView Ps = site(ia);
        const int k1 = kkind[ia];
        int b = onsite_base(k1) + 1;
        double v13 = std::pow(den[ia * kinds + k1 - 1], third);
        double v23 = v13 * v13;
        double v2t = 2.0 * v23, v3t = threeh * v23;
        double drho = twtrd * vden[ia * kinds + k1 - 1] / v13;
        for (int c = 0; c < 4; ++c, b += 4)
            dparos[4 * ia + c] = drho * (Ps(b + 2) + v2t * (Ps(b + 3) + v3t * Ps(b + 4)));
        for (int k2 = 1; k2 <= kinds; ++k2) {
            if (k2 == k1 || den[ia * kinds + k2 - 1] == 0.0) continue;
            v13 = std::pow(den[ia * kinds + k2 - 1], third);
            v23 = v13 * v13;
            v2t = 2.0 * v23; v3t = threeh * v23;
            drho = twtrd * vden[ia * kinds + k2 - 1] / v13;
            const int hi = k1 > k2 ? k1 : k2, lo = k1 < k2 ? k1 : k2;
            b = kinds * (1 + kOnsite * 4) + 6 * kOnsite * ((hi - 1) * (hi - 2) / 2 + lo - 1);
            if (k2 < hi) b += 3 * kOnsite;
            for (int c = 0; c < 4; ++c, b += 3)
                dparos[4 * ia + c] = dparos[4 * ia + c] + drho * (Ps(b + 1) + v2t * (Ps(b + 2) + v3t * Ps(b + 3)));
        }
    }
    return TBK_OK;
}

// setpar.f:466-485, 744-795 (lnoncol): densities of both spins from each
// channel's decay, on-site terms as the average (paros) and half difference
// (spinos) of the two channels.  The terms from other kinds replace, not add
// to, the own-kind values and take four coefficients per orbital type with
// the full difference, as setpar.f:772-794 does.  den2[s*natoms*kinds + ...].
void setpar_noncol(int natoms, const int *kkind, int kinds, const double *param_up, const double *param_dn,
                   const PairList &pl, double *den2, double *paros, double *spinos)
{
    const ParamView U{param_up}, D{param_dn};
    const double twtrd = 2.0 / 3.0, half = 0.5;
    const size_t na = (size_t)natoms * kinds;
    for (size_t i = 0; i < 2 * na; ++i) den2[i] = 0.0;
    double *du = den2, *dd = den2 + na;
    for (int ip = 0; ip < pl.n; ++ip) {
        const Pair &pr = pl.pairs[ip];
        const int ia = pr.atom[0], ja = pr.atom[1];
        const int ki = pr.kind[0], kj = pr.kind[1];
        const double r = pr.dist, f = pr.screen;
        const int i0 = onsite_base(kj) + 1;
        const double bup = U(i0) * U(i0), bdn = D(i0) * D(i0);
        du[ia * kinds + kj - 1] = du[ia * kinds + kj - 1] + std::exp(-bup * r) * f;
        dd[ia * kinds + kj - 1] = dd[ia * kinds + kj - 1] + std::exp(-bdn * r) * f;
        if (ja != ia) {
            const int i1 = onsite_base(ki) + 1;
            const double bu1 = U(i1) * U(i1), bd1 = D(i1) * D(i1);
            du[ja * kinds + ki - 1] = du[ja * kinds + ki - 1] + std::exp(-bu1 * r) * f;
            dd[ja * kinds + ki - 1] = dd[ja * kinds + ki - 1] + std::exp(-bd1 * r) * f;
        }
    }
    for (int ia = 0; ia < natoms; ++ia) {
        const int k1 = kkind[ia];
        int b = onsite_base(k1) + 1;
        double vu = std::pow(du[ia * kinds + k1 - 1], twtrd), vd = std::pow(dd[ia * kinds + k1 - 1], twtrd);
        for (int c = 0; c < 4; ++c, b += 4) {
            const double pup = U(b + 1) + vu * (U(b + 2) + vu * (U(b + 3) + vu * U(b + 4)));
            const double pdn = D(b + 1) + vd * (D(b + 2) + vd * (D(b + 3) + vd * D(b + 4)));
            paros[4 * ia + c] = half * (pup + pdn);
            spinos[4 * ia + c] = half * (pup - pdn);
        }
        for (int k2 = 1; k2 <= kinds; ++k2) {
            if (k2 == k1 || !(std::max(du[ia * kinds + k2 - 1], dd[ia * kinds + k2 - 1]) > 0.0)) continue;
            vu = std::pow(du[ia * kinds + k2 - 1], twtrd);
            vd = std::pow(dd[ia * kinds + k2 - 1], twtrd);
            const int hi = k1 > k2 ? k1 : k2, lo = k1 < k2 ? k1 : k2;
            b = kinds * (1 + kOnsite * 4) + 6 * kOnsite * ((hi - 1) * (hi - 2) / 2 + lo - 1);
            if (k2 < hi) b += 3 * kOnsite;
            for (int c = 0; c < 4; ++c, b += 4) {
                const double pup = U(b + 1) + vu * (U(b + 2) + vu * (U(b + 3) + vu * U(b + 4)));
                const double pdn = D(b + 1) + vd * (D(b + 2) + vd * (D(b + 3) + vd * D(b + 4)));
                paros[4 * ia + c] = half * (pup + pdn);
                spinos[4 * ia + c] = pup - pdn;
            }
        }
    }
}

// setparv.f:436-490, 725-798 (lnoncol): volume derivatives of both spins'
// densities and of the average / half-difference on-site terms.  For other
// kinds the reference takes the spin-down density derivative from the
// spin-up one and steps three coefficients per orbital type
// (setparv.f:760-797); both are kept.
void setparv_noncol(int natoms, const int *kkind, int kinds, const double *param_up, const double *param_dn,
                    double thisvol, const PairList &pl, const double *den2, double *dparos, double *dspnos)
{
    const ParamView U{param_up}, D{param_dn};
    const double third = 1.0 / 3.0, twtrd = 2.0 / 3.0, threeh = 3.0 / 2.0, half = 0.5;
    const size_t na = (size_t)natoms * kinds;
    std::vector<double> vden(2 * na, 0.0);
    double *vu = vden.data(), *vd = vden.data() + na;
    const double *du = den2, *dd = den2 + na;
    for (int ip = 0; ip < pl.n; ++ip) {
        const Pair &pr = pl.pairs[ip];
        const int ia = pr.atom[0], ja = pr.atom[1];
        const int ki = pr.kind[0], kj = pr.kind[1];
        const double r = pr.dist, f = pr.screen, df = pr.dscrn;
        const double vdist = third * r / thisvol;
        const int i0 = onsite_base(kj) + 1;
        double bup = U(i0) * U(i0), bdn = D(i0) * D(i0);
        vu[ia * kinds + kj - 1] = vu[ia * kinds + kj - 1] + vdist * (df - bup) * std::exp(-bup * r) * f;
        vd[ia * kinds + kj - 1] = vd[ia * kinds + kj - 1] + vdist * (df - bdn) * std::exp(-bdn * r) * f;
        if (ja != ia) {
            const int i1 = onsite_base(ki) + 1;
            bup = U(i1) * U(i1); bdn = D(i1) * D(i1);
            vu[ja * kinds + ki - 1] = vu[ja * kinds + ki - 1] + vdist * (df - bup) * std::exp(-bup * r) * f;
            vd[ja * kinds + ki - 1] = vd[ja * kinds + ki - 1] + vdist * (df - bdn) * std::exp(-bdn * r) * f;
        }
    }
    for (int ia = 0; ia < natoms; ++ia) {
        const int k1 = kkind[ia];
        int b = onsite_base(k1) + 1;
        double u13 = std::pow(du[ia * kinds + k1 - 1], third), d13 = std::pow(dd[ia * kinds + k1 - 1], third);
        double u23 = u13 * u13, d23 = d13 * d13;
        double u2t = 2.0 * u23, u3t = threeh * u23, d2t = 2.0 * d23, d3t = threeh * d23;
        double dru = twtrd * vu[ia * kinds + k1 - 1] / u13, drd = twtrd * vd[ia * kinds + k1 - 1] / d13;
        for (int c = 0; c < 4; ++c, b += 4) {
            const double dpup = dru * (U(b + 2) + u2t * (U(b + 3) + u3t * U(b + 4)));
            const double dpdn = drd * (D(b + 2) + d2t * (D(b + 3) + d3t * D(b + 4)));
            dparos[4 * ia + c] = half * (dpup + dpdn);
            dspnos[4 * ia + c] = half * (dpup - dpdn);
        }
        for (int k2 = 1; k2 <= kinds; ++k2) {
            if (k2 == k1 || du[ia * kinds + k2 - 1] == 0.0) continue;
            u13 = std::pow(du[ia * kinds + k2 - 1], third); u23 = u13 * u13;
            u2t = 2.0 * u23; u3t = threeh * u23;
            dru = twtrd * vu[ia * kinds + k2 - 1] / u13;
            d13 = std::pow(dd[ia * kinds + k2 - 1], third); d23 = d13 * d13;
            d2t = 2.0 * d23; d3t = threeh * d23;
            drd = twtrd * vu[ia * kinds + k2 - 1] / d13;
            const int hi = k1 > k2 ? k1 : k2, lo = k1 < k2 ? k1 : k2;
            b = kinds * (1 + kOnsite * 4) + 6 * kOnsite * ((hi - 1) * (hi - 2) / 2 + lo - 1);
            if (k2 < hi) b += 3 * kOnsite;
            for (int c = 0; c < 4; ++c, b += 3) {
                const double dpup = dru * (U(b + 2) + u2t * (U(b + 3) + u3t * U(b + 4)));
                const double dpdn = drd * (D(b + 2) + d2t * (D(b + 3) + d3t * D(b + 4)));
                dparos[4 * ia + c] = half * (dpup + dpdn);
                dspnos[4 * ia + c] = half * (dpup - dpdn);
            }
        }
    }
}

// spinham.f:110-150: per orbital (rows in atom order), the diagonal spin
// term t(q, atom) * s_z and the mixing term t(q, atom) * (s_x + i s_y) of
// an atom of spin direction s, q the orbital type (s; p; t2g; eg).
void spinor_terms(int natoms, const int *nbas, const double *spinos, const double *spinat, double *d,
                  double *mre, double *mim)
{
    int i = 0;
    for (int a = 0; a < natoms; ++a)
        for (int k = 1; k <= nbas[a]; ++k, ++i) {
            const int q = k == 1 ? 0 : (k < 5 ? 1 : (k < 8 ? 2 : 3));
            const double t = spinos[4 * a + q];
            d[i] = t * spinat[3 * a + 2];
            mre[i] = t * spinat[3 * a];
            mim[i] = t * spinat[3 * a + 1];
        }
}

// rotate.f:112-258: two-centre table e(i,j) for direction cosines (l,m,n)
// and integrals p(1..14) (p(11..14) used only for unlike kinds).
void sk_block(double l, double m, double n, const double *pp, bool unlike, double *e /*9x9 row-major*/)
{
    const double rt3 = 1.732050807568877, half = 0.5;
    const double rt3h = half * rt3, rt3q = half * rt3h, rt32 = 2.0 * rt3;
    static const double parity[9] = {1, -1, -1, -1, 1, 1, 1, 1, 1};
    auto E = [e](int i, int j) -> double & { return e[(i - 1) * 9 + (j - 1)]; };
    auto p = [pp](int i) { return pp[i - 1]; };
    const double ll = l * l, mm = m * m, nn = n * n, lm = l * m, mn = m * n, nl = n * l;
    const double lmn = lm * n;
    E(1, 1) = p(1);
    E(1, 2) = l * p(2); E(1, 3) = m * p(2); E(1, 4) = n * p(2);
    const double rp5 = rt3 * p(5);
    E(1, 5) = mn * rp5; E(1, 6) = nl * rp5; E(1, 7) = lm * rp5;
    E(1, 8) = half * rp5 * (ll - mm);
    E(1, 9) = (nn - half * (ll + mm)) * p(5);
    const double p34 = p(3) - p(4);
    E(2, 2) = p(4) + ll * p34; E(3, 3) = p(4) + mm * p34; E(4, 4) = p(4) + nn * p34;
    E(2, 3) = lm * p34; E(2, 4) = nl * p34; E(3, 4) = mn * p34;
    const double p67 = rt3 * p(6) - 2.0 * p(7);
    E(2, 5) = lmn * p67;
    E(2, 6) = n * (ll * p67 + p(7));
    E(2, 7) = m * (ll * p67 + p(7));
    E(3, 5) = n * (mm * p67 + p(7));
    E(3, 6) = E(2, 5);
    E(3, 7) = l * (mm * p67 + p(7));
    E(4, 5) = m * (nn * p67 + p(7));
    E(4, 6) = l * (nn * p67 + p(7));
    E(4, 7) = E(2, 5);
    E(2, 8) = l * (half * (ll - mm) * p67 + p(7));
    const double p67a = p(6) - rt3 * p(7);
    E(2, 9) = l * (nn * p67a - (half * (ll + mm)) * p(6));
    E(3, 8) = m * (half * (ll - mm) * p67 - p(7));
    E(3, 9) = m * (nn * p67a - (half * (ll + mm)) * p(6));
    E(4, 8) = n * (half * (ll - mm) * p67);
    E(4, 9) = n * (nn * p(6) - half * (ll + mm) * (p(6) - rt32 * p(7)));
    const double p8910 = (3.0 * p(8) + p(10)) - 4.0 * p(9);
    E(5, 5) = ll * p(10) + mm * (nn * p8910) + (mm + nn) * p(9);
    const double p910m = p(9) - p(10);
    E(5, 6) = lm * (p910m + nn * p8910);
    E(5, 7) = nl * (p910m + mm * p8910);
    E(6, 6) = mm * p(10) + ll * (nn * p8910) + (ll + nn) * p(9);
    E(6, 7) = mn * (p910m + ll * p8910);
    E(7, 7) = nn * p(10) + ll * (mm * p8910) + (ll + mm) * p(9);
    E(5, 8) = mn * ((half * (ll - mm)) * p8910 - p910m);
    E(6, 8) = nl * ((half * (ll - mm)) * p8910 + p910m);
    E(7, 8) = lm * (half * (ll - mm)) * p8910;
    const double p89 = p(8) - 2.0 * p(9);
    const double p8910a = p89 + p(10);
    const double p8910b = p8910a + p89;
    const double p89m = p(8) - p(9);
    E(5, 9) = rt3 * mn * (nn * p89m - (half * (ll + mm)) * p8910a);
    E(6, 9) = rt3 * nl * (nn * p89m - (half * (ll + mm)) * p8910a);
    E(7, 9) = rt3h * lm * (p(10) - (ll + mm) * p(8) + nn * p8910b);
    const double hlm = half * (ll - mm);
    E(8, 8) = (ll + mm) * p(9) + nn * p(10) + (hlm * hlm) * p8910;
    E(8, 9) = rt3q * (ll - mm) * (p(10) - (ll + mm) * p(8) + nn * p8910b);
    E(9, 9) = nn * nn * p(8) + (ll + mm) * (2.5e-1 * (ll + mm) * (3.0 * p(10) + p(8)) -
              nn * (p(8) - 3.0 * p(9)));
    for (int i = 2; i <= 9; ++i)
        for (int j = 1; j < i; ++j) E(i, j) = parity[i - 1] * (E(j, i) * parity[j - 1]);
    if (unlike) {
        E(2, 1) = l * p(11); E(3, 1) = m * p(11); E(4, 1) = n * p(11);
        const double rp12 = rt3 * p(12);
        E(5, 1) = mn * rp12; E(6, 1) = nl * rp12; E(7, 1) = lm * rp12;
        E(8, 1) = half * rp12 * (ll - mm);
        E(9, 1) = (nn - half * (ll + mm)) * p(12);
        const double p134 = rt3 * p(13) - 2.0 * p(14);
        E(5, 2) = lmn * p134;
        E(6, 2) = n * (ll * p134 + p(14));
        E(7, 2) = m * (ll * p134 + p(14));
        E(5, 3) = n * (mm * p134 + p(14));
        E(6, 3) = E(5, 2);
        E(7, 3) = l * (mm * p134 + p(14));
        E(5, 4) = m * (nn * p134 + p(14));
        E(6, 4) = l * (nn * p134 + p(14));
        E(7, 4) = E(5, 2);
        E(8, 2) = l * (half * (ll - mm) * p134 + p(14));
        const double p134a = p(13) - rt3 * p(14);
        E(9, 2) = l * (nn * p134a - (half * (ll + mm)) * p(13));
        E(8, 3) = m * (half * (ll - mm) * p134 - p(14));
        E(9, 3) = m * (nn * p134a - (half * (ll + mm)) * p(13));
        E(8, 4) = n * (half * (ll - mm) * p134);
        E(9, 4) = n * (nn * p(13) - half * (ll + mm) * (p(13) - rt32 * p(14)));
    }
}

// setham.f:326-376: the integral vector rotate.f:155 reads for a pair of
// kinds (k1, k2) = (jkind(1), jkind(2)), from the stored hop/over row.
void arrange_integrals(const double *stored, int k1, int k2, double *out /*14*/)
{
    const int lo = k1 < k2 ? k1 : k2, hi = k1 < k2 ? k2 : k1;
    double fwd[14] = {0}, rev[14] = {0};    // par(.,hi,lo) and par(.,lo,hi)
    const int npk = lo == hi ? kLike : kUnlike;
    for (int j = 0; j < npk; ++j) fwd[j] = stored[j];
    if (lo != hi) {
        for (int j = 0; j < kLike; ++j) rev[j] = fwd[j];
        rev[1] = -fwd[10]; rev[4] = fwd[11]; rev[5] = -fwd[12]; rev[6] = -fwd[13];
    }
    rev[10] = -fwd[1]; rev[11] = fwd[4]; rev[12] = -fwd[5]; rev[13] = -fwd[6];
    if (lo == hi) {
        for (int j = 0; j < 10; ++j) out[j] = fwd[j];
        for (int j = 10; j < 14; ++j) out[j] = rev[j];
        return;
    }
    const double *src = (k2 == hi) ? fwd : rev;
    for (int j = 0; j < 14; ++j) out[j] = src[j];
}

}  // namespace tbk

// ---------------------------------------------------------------------------
// C ABI of the host set-up
// ---------------------------------------------------------------------------
extern "C" int tbk_search2(int natoms, const double *lat, const double *posn,
                           const int *kkind, int kinds, const int *nd,
                           const double *dnn, const double *screenl, int maxpairs,
                           int *npairs, double *tt, double *dlv, double *dist,
                           double *screen, double *dscrn, int *jkind, int *jatm,
                           double *dismin)
{
    if (natoms <= 0 || !lat || !posn || !kkind || !nd || !npairs) return TBK_ERR_ARG;
    tbk::PairList pl;
    int rc = tbk::search2(natoms, lat, posn, kkind, kinds, nd, dnn, screenl, maxpairs, pl);
    *npairs = pl.n;
    for (int p = 0; p < pl.n; ++p) {
        const tbk::Pair &q = pl.pairs[p];
        for (int c = 0; c < 3; ++c) { tt[3 * p + c] = q.cosines[c]; dlv[3 * p + c] = q.dlv[c]; }
        dist[p] = q.dist; screen[p] = q.screen; dscrn[p] = q.dscrn;
        jkind[2 * p] = q.kind[0]; jkind[2 * p + 1] = q.kind[1];
        jatm[2 * p] = q.atom[0] + 1; jatm[2 * p + 1] = q.atom[1] + 1;
    }
    if (dismin) *dismin = pl.dismin;
    return rc;
}

static tbk::PairList pairs_from_arrays(int npairs, const double *dist, const double *screen,
                                       const double *dscrn, const int *jkind, const int *jatm)
{
    tbk::PairList pl;
    for (int p = 0; p < npairs; ++p) {
        tbk::Pair q{};
        q.dist = dist[p]; q.screen = screen[p]; q.dscrn = dscrn ? dscrn[p] : 0.0;
        q.kind[0] = jkind[2 * p]; q.kind[1] = jkind[2 * p + 1];
        q.atom[0] = jatm[2 * p] - 1; q.atom[1] = jatm[2 * p + 1] - 1;
        pl.push(q);
    }
    return pl;
}

extern "C" int tbk_setpar(int natoms, const int *kkind, int kinds, const double *param,
                          int nltype, int jsover, int npairs, const double *dist,
                          const double *screen, const int *jkind, const int *jatm,
                          double *den, double *paros, double *hop, double *over)
{
    if (natoms <= 0 || npairs < 0 || !param) return TBK_ERR_ARG;
    tbk::PairList pl = pairs_from_arrays(npairs, dist, screen, nullptr, jkind, jatm);
    return tbk::setpar(natoms, kkind, kinds, param, nltype, jsover, pl, den, paros, hop, over);
}

extern "C" int tbk_setpar_spin(int natoms, const int *kkind, int kinds, const double *param2, int nparam,
                               int jspin, const int *flipos, int nltype, int jsover, int npairs,
                               const double *dist, const double *screen, const int *jkind, const int *jatm,
                               double *den, double *paros, double *hop, double *over)
{
    if (natoms <= 0 || npairs < 0 || !param2 || nparam < 1 || (jspin != 1 && jspin != 2)) return TBK_ERR_ARG;
    tbk::PairList pl = pairs_from_arrays(npairs, dist, screen, nullptr, jkind, jatm);
    const double *own = param2 + (size_t)(jspin - 1) * nparam, *other = param2 + (size_t)(2 - jspin) * nparam;
    return tbk::setpar(natoms, kkind, kinds, own, nltype, jsover, pl, den, paros, hop, over, other, flipos);
}

extern "C" int tbk_setparv_spin(int natoms, const int *kkind, int kinds, const double *param2, int nparam,
                                int jspin, const int *flipos, int nltype, int jsover, double thisvol,
                                int npairs, const double *dist, const double *screen, const double *dscrn,
                                const int *jkind, const int *jatm, const double *den, double *dparos,
                                double *dhop, double *dover)
{
    if (natoms <= 0 || npairs < 0 || !param2 || nparam < 1 || (jspin != 1 && jspin != 2) || thisvol <= 0)
        return TBK_ERR_ARG;
    tbk::PairList pl = pairs_from_arrays(npairs, dist, screen, dscrn, jkind, jatm);
    const double *own = param2 + (size_t)(jspin - 1) * nparam, *other = param2 + (size_t)(2 - jspin) * nparam;
    return tbk::setparv(natoms, kkind, kinds, own, nltype, jsover, thisvol, pl, den, dparos, dhop, dover,
                        other, flipos);
}

extern "C" int tbk_setpar_noncol(int natoms, const int *kkind, int kinds, const double *param2, int nparam,
                                 int nltype, int jsover, double thisvol, int npairs, const double *dist,
                                 const double *screen, const double *dscrn, const int *jkind, const int *jatm,
                                 double *den2, double *paros, double *spinos, double *dparos, double *dspnos,
                                 double *hop, double *over, double *dhop, double *dover)
{
    if (natoms <= 0 || npairs < 0 || !param2 || nparam < 1 || thisvol <= 0) return TBK_ERR_ARG;
    tbk::PairList pl = pairs_from_arrays(npairs, dist, screen, dscrn, jkind, jatm);
    const double *up = param2, *dn = param2 + nparam;
    // bond and overlap integrals from channel 1 (setpar/setparv with jspin = 1)
    std::vector<double> den1((size_t)natoms * kinds), scratch(4 * (size_t)natoms);
    int rc = tbk::setpar(natoms, kkind, kinds, up, nltype, jsover, pl, den1.data(), scratch.data(), hop, over);
    if (!rc) tbk::setpar_noncol(natoms, kkind, kinds, up, dn, pl, den2, paros, spinos);
    if (!rc) rc = tbk::setparv(natoms, kkind, kinds, up, nltype, jsover, thisvol, pl, den1.data(),
                               scratch.data(), dhop, dover);
    if (!rc) tbk::setparv_noncol(natoms, kkind, kinds, up, dn, thisvol, pl, den2, dparos, dspnos);
    return rc;
}

extern "C" int tbk_setparv(int natoms, const int *kkind, int kinds, const double *param,
                           int nltype, int jsover, double thisvol, int npairs,
                           const double *dist, const double *screen, const double *dscrn,
                           const int *jkind, const int *jatm, const double *den,
                           double *dparos, double *dhop, double *dover)
{
    if (natoms <= 0 || npairs < 0 || !param || thisvol <= 0) return TBK_ERR_ARG;
    tbk::PairList pl = pairs_from_arrays(npairs, dist, screen, dscrn, jkind, jatm);
    return tbk::setparv(natoms, kkind, kinds, param, nltype, jsover, thisvol, pl, den,
                        dparos, dhop, dover);
}
