// tbk_io.cpp -- the input and output files of the reference's program for
// this path, read and written natively: the NRL parameter file (input.f),
// the SKIN control file with its lattice menus, structure blocks and k-point
// blocks (static.f, setlat.f, input1.f, kptin.f), and SKENG (static.f); and
// the program's loop over the structures of a SKIN file with the device
// plans in place of bande (tbk_static_run).
//
// The reads follow the reference's Fortran I/O statement by statement
// (tbk_fortran_io.h gives the edit descriptors and list-directed semantics,
// SKIN with BLANK='ZERO' as static.f:682 opens it), and the lattice and
// position arithmetic follows setlat.f / input1.f / static.f in their
// operation order (compiled with -ffp-contract=off), so the structures are
// bit-identical to those the reference program passes to bande.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "tbk.h"
#include "tbk_fortran_io.h"

namespace {

constexpr double kPi = 3.14159265358979323846;
// kptin.f's dimensions (NRLtb/ex1/P3): k-points, full meshes, operations
constexpr int kNkd = 50000, kNopd = 48;

std::string trim(const std::string &s)
{
    size_t b = 0, e = s.size();
    while (b < e && (s[b] == ' ' || s[b] == '\t')) ++b;
    while (e > b && (s[e - 1] == ' ' || s[e - 1] == '\t')) --e;
    return s.substr(b, e - b);
}

std::string resolve(const std::string &dir, const std::string &name)
{
    if (name.empty() || name[0] == '/' || dir.empty()) return name;
    return dir + "/" + name;
}

std::string dirname_of(const std::string &path)
{
    const size_t p = path.rfind('/');
    return p == std::string::npos ? std::string(".") : path.substr(0, p);
}

// setlat.f:207-1025.  alat[3*i + j] = component j of primitive vector i
// (plv(i, j)).  Returns setlat's ierr: 0, 1 bad lattice (unknown type,
// zero volume, lattice 12 angle), 2 latnum input, 3 lattice parameters,
// 4 strain input; *latnum and *strain as read.
int setlat(tbk::FortranReader &rd, double *alat, double *vol_out, int *latnum_out, int *strain_out)
{
    const double zero = 0.0, one = 1.0, two = 2.0, three = 3.0, four = 4.0;
    const double third = one / three, thirdm = -third, half = 5e-1, halfm = -half, quarter = one / four;
    const double two3rd = two / three;
    const double fac1 = 0.866025403784438647, fac1m = -fac1, fac2 = one / fac1;
    const double tpi3 = two * kPi / three;
    auto A = [&](int i, int j) -> double & { return alat[3 * (i - 1) + (j - 1)]; };
    for (int i = 0; i < 9; ++i) alat[i] = zero;
    int latnum = 0;
    if (!rd.next_record() || !rd.list_int(&latnum)) return 2;
    *latnum_out = latnum;
    double a = 0, b = 0, c = 0, vol = 0, coa = 0, boa = 0, theta = 0, scale[3] = {1, 1, 1};
    auto read_list = [&](std::initializer_list<double *> xs) {
        if (!rd.next_record()) return false;
        for (double *x : xs)
            if (!rd.list_real(x)) return false;
        return true;
    };
    const int al = std::abs(latnum);
    if (latnum == 0) {
        if (!rd.next_record()) return 3;
        for (int i = 1; i <= 3; ++i)
            for (int j = 1; j <= 3; ++j)
                if (!rd.list_real(&A(i, j))) return 3;
        if (!read_list({scale, scale + 1, scale + 2})) return 3;
        for (int i = 0; i < 3; ++i)
            if (scale[i] == zero) scale[i] = one;
    } else if (al == 1) {
        if (latnum == 1) {
            if (!read_list({&a})) return 3;
        } else {
            if (!read_list({&vol})) return 3;
            a = std::pow(four * std::fabs(vol), third);
        }
        for (int i = 1; i <= 3; ++i)
            for (int j = 1; j <= 3; ++j)
                if (j != i) A(j, i) = half * a;
    } else if (al == 2) {
        if (latnum == 2) {
            if (!read_list({&a})) return 3;
        } else {
            if (!read_list({&vol})) return 3;
            a = std::pow(two * std::fabs(vol), third);
        }
        for (int i = 1; i <= 3; ++i)
            for (int j = 1; j <= 3; ++j) A(j, i) = (j == i) ? -half * a : half * a;
    } else if (al == 3) {
        if (latnum == 3) {
            if (!read_list({&a, &c})) return 3;
            coa = c / a;
        } else {
            if (!read_list({&vol, &coa})) return 3;
            a = std::pow(fac2 * std::fabs(vol) / std::fabs(coa), third);
            c = std::fabs(coa) * a;
        }
        A(1, 1) = half * a;
        A(2, 1) = A(1, 1);
        A(1, 2) = fac1m * a;
        A(2, 2) = -A(1, 2);
        A(3, 3) = c;
    } else if (al == 4) {
        if (latnum == 4) {
            if (!read_list({&a})) return 3;
        } else {
            if (!read_list({&vol})) return 3;
            a = std::pow(std::fabs(vol), third);
        }
        for (int i = 1; i <= 3; ++i) A(i, i) = a;
    } else if (al == 5) {
        if (latnum == 5) {
            if (!read_list({&a, &c})) return 3;
        } else {
            if (!read_list({&vol, &coa})) return 3;
            a = std::pow(std::fabs(vol) / std::fabs(coa), third);
            c = coa * a;
        }
        A(1, 1) = a; A(2, 2) = a; A(3, 3) = c;
    } else if (al == 6) {
        if (latnum == 6) {
            if (!read_list({&a, &c})) return 3;
        } else {
            if (!read_list({&vol, &coa})) return 3;
            a = std::pow(two * std::fabs(vol) / std::fabs(coa), third);
            c = coa * a;
        }
        A(1, 1) = a; A(2, 2) = a; A(3, 1) = half * a; A(3, 2) = half * a; A(3, 3) = half * c;
    } else if (al == 7) {
        if (latnum == 7) {
            if (!read_list({&a, &b, &c})) return 3;
        } else {
            if (!read_list({&vol, &boa, &coa})) return 3;
            a = std::pow(std::fabs(vol) / (std::fabs(boa * coa)), third);
            b = boa * a;
            c = coa * a;
        }
        A(1, 1) = a; A(2, 2) = b; A(3, 3) = c;
    } else if (al == 8) {
        if (latnum == 8) {
            if (!read_list({&a, &b, &c})) return 3;
        } else {
            if (!read_list({&vol, &boa, &coa})) return 3;
            a = std::pow(std::fabs(two * vol) / (std::fabs(boa * coa)), third);
            b = boa * a;
            c = coa * a;
        }
        A(1, 1) = half * a; A(2, 1) = A(1, 1); A(2, 2) = half * b; A(1, 2) = -A(2, 2); A(3, 3) = c;
    } else if (al == 9) {
        if (latnum == 9) {
            if (!read_list({&a, &c})) return 3;
        } else {
            if (!read_list({&vol, &coa})) return 3;
            a = std::pow(std::fabs(two * vol / coa), third);
            c = coa * a;
        }
        A(1, 1) = half * a; A(1, 2) = -A(1, 1); A(2, 1) = A(1, 1); A(2, 2) = A(1, 1); A(3, 3) = c;
    } else if (al == 10) {
        if (latnum == 10) {
            if (!read_list({&a, &c})) return 3;
        } else {
            if (!read_list({&vol, &coa})) return 3;
            a = std::pow(std::fabs(four * vol / coa), third);
            c = coa * a;
        }
        A(1, 1) = half * a; A(1, 2) = -A(1, 1); A(2, 1) = A(1, 1); A(2, 2) = A(1, 1); A(3, 1) = A(1, 1);
        A(3, 3) = half * c;
    } else if (al == 11) {
        if (latnum == 11) {
            if (!read_list({&a, &c})) return 3;
        } else {
            if (!read_list({&vol, &coa})) return 3;
            a = std::pow(four * std::fabs(vol) / std::fabs(coa), third);
            c = coa * a;
        }
        A(2, 1) = half * a; A(3, 1) = half * a; A(1, 2) = half * a; A(3, 2) = half * a;
        A(1, 3) = half * c; A(2, 3) = half * c;
    } else if (al == 12) {
        if (latnum == 12) {
            if (!read_list({&a, &b})) return 3;
        } else {
            if (!read_list({&vol, &theta})) return 3;
            coa = std::cos(theta);
            if (coa < halfm) return 1;                         // the reference stops here
            boa = third * (std::sqrt((one + two * coa) / (one - coa)) - one);
            c = std::pow(std::fabs(vol) / (one + three * boa), third);
            a = c * (one + boa);
            b = c * boa;
        }
        for (int i = 1; i <= 3; ++i)
            for (int j = 1; j <= 3; ++j) A(j, i) = (j == i) ? a : b;
    } else if (al == 13) {
        if (latnum == 13) {
            if (!read_list({&a, &c})) return 3;
        } else {
            if (!read_list({&vol, &coa})) return 3;
            a = std::pow(two * std::fabs(vol) / std::fabs(coa), third);
            c = coa * a;
        }
        A(1, 1) = -half * a; A(1, 2) = half * a; A(1, 3) = half * c;
        A(2, 1) = half * a; A(2, 2) = -half * a; A(2, 3) = half * c;
        A(3, 1) = half * a; A(3, 2) = half * a; A(3, 3) = -half * c;
    } else if (al == 14) {
        if (latnum == 14) {
            if (!read_list({&a, &c, &theta})) return 3;
            vol = a * a * c * std::sin(theta);
        } else {
            if (!read_list({&vol, &coa, &theta})) return 3;
            a = std::pow(vol / (coa * std::sin(theta)), third);
            c = a * coa;
        }
        theta = half * theta;
        A(1, 1) = a * std::cos(theta);
        A(1, 2) = -a * std::sin(theta);
        A(2, 1) = A(1, 1);
        A(2, 2) = -A(1, 2);
        A(3, 3) = c;
    } else if (al == 15 || al == 16) {
        double aref = 0, rmagic = 0;
        if (latnum > 0) {
            if (!read_list({&aref, &rmagic, &theta})) return 3;
            vol = (al == 15 ? half : quarter) * (aref * aref * aref);
        } else {
            if (!read_list({&vol, &rmagic, &theta})) return 3;
        }
        const double w = std::pow((al == 15 ? half : quarter) * vol, third);
        const double sgn = al == 15 ? -1.0 : 1.0;
        const double x = w * std::exp(sgn * rmagic * std::cos(theta + tpi3));
        const double y = w * std::exp(sgn * rmagic * std::cos(theta - tpi3));
        const double z = w * std::exp(sgn * rmagic * std::cos(theta));
        if (al == 15) {
            A(1, 2) = y; A(1, 3) = z; A(2, 1) = x; A(2, 3) = z; A(3, 1) = x; A(3, 2) = y;
        } else {
            A(1, 1) = -x; A(1, 2) = y; A(1, 3) = z;
            A(2, 1) = x; A(2, 2) = -y; A(2, 3) = z;
            A(3, 1) = x; A(3, 2) = y; A(3, 3) = -z;
        }
    } else {
        return 1;                                              // unknown lattice type
    }
    // strains (setlat.f:811-976)
    int strain = 0;
    if (!rd.next_record() || !rd.list_int(&strain)) return 4;
    *strain_out = strain;
    if (strain != 0) {
        double e[7] = {0, 0, 0, 0, 0, 0, 0};      // e[1..6]
        double x = 0, x2 = 0;
        if (strain == 1) {
            if (!rd.next_record()) return 4;
            for (int i = 1; i <= 6; ++i)
                if (!rd.list_real(e + i)) return 4;
        } else if (std::abs(strain) == 2) {
            if (!read_list({&x2})) return 4;
            x = std::sqrt(x2);
            if (strain == 2) {
                e[1] = std::sqrt((one + x) / (one - x)) - one;
                e[2] = std::sqrt((one - x) / (one + x)) - one;
            } else {
                e[1] = x; e[2] = -x; e[3] = x2 / (one - x2);
            }
        } else if (std::abs(strain) == 21) {
            if (!read_list({&x})) return 4;
            x2 = x * x;
            if (strain == 21) {
                e[1] = std::sqrt((one + x) / (one - x)) - one;
                e[2] = std::sqrt((one - x) / (one + x)) - one;
            } else {
                e[1] = x; e[2] = -x; e[3] = x2 / (one - x2);
            }
        } else if (std::abs(strain) == 3) {
            if (!read_list({&x2})) return 4;
            e[6] = std::sqrt(x2);
            if (strain == 3) {
                e[1] = std::sqrt(one + quarter * x2) - one;
                e[2] = e[1];
            } else {
                e[3] = x2 / (four - x2);
            }
        } else if (strain == 4) {
            if (!read_list({&x})) return 4;
            e[1] = std::pow(one + x, thirdm) - one;
            e[2] = e[1];
            e[3] = std::pow(one + x, two3rd) - one;
        } else if (strain == 5) {
            if (!read_list({&x})) return 4;
            e[1] = (one + x) * (std::pow(one + three * x, thirdm)) - one;
            e[2] = e[1]; e[3] = e[1];
            e[4] = two * x * (std::pow(one + three * x, thirdm));
            e[5] = e[4]; e[6] = e[5];
        } else if (strain == 6) {
            if (!read_list({&x2})) return 4;
            if (x2 < zero) {
                x2 = std::fabs(x2);
                x = -std::sqrt(x2);
            } else {
                x = std::sqrt(x2);
            }
            e[4] = x;
            e[3] = quarter * x * x;
        } else {
            return 4;                                          // unknown strain type
        }
        double est[3][3];
        for (int i = 0; i < 3; ++i) est[i][i] = one + e[i + 1];
        est[1][2] = half * e[4]; est[2][1] = est[1][2];
        est[0][2] = half * e[5]; est[2][0] = est[0][2];
        est[0][1] = half * e[6]; est[1][0] = est[0][1];
        double alats[9];
        for (int i = 1; i <= 3; ++i)
            for (int j = 1; j <= 3; ++j) {
                double sum = zero;
                for (int k = 1; k <= 3; ++k) sum = sum + A(j, k) * est[k - 1][i - 1];
                alats[3 * (j - 1) + (i - 1)] = sum;
            }
        for (int i = 0; i < 9; ++i) alat[i] = alats[i];
    }
    if (latnum == 0)
        for (int j = 1; j <= 3; ++j)
            for (int i = 1; i <= 3; ++i) A(i, j) = A(i, j) * scale[j - 1];
    vol = std::fabs(A(1, 1) * (A(2, 2) * A(3, 3) - A(3, 2) * A(2, 3)) +
                    A(2, 1) * (A(3, 2) * A(1, 3) - A(1, 2) * A(3, 3)) +
                    A(3, 1) * (A(1, 2) * A(2, 3) - A(2, 2) * A(1, 3)));
    if (vol == zero) return 1;
    *vol_out = vol;
    return 0;
}

}  // namespace

// ---------------------------------------------------------------------------
// NRL parameter file (input.f:139-359)
// ---------------------------------------------------------------------------
extern "C" int tbk_read_params(const char *path, tbk_param_set *ps, int cap, double *param)
{
    if (!path || !ps || !param || cap < 1) return TBK_ERR_ARG;
    FILE *f = std::fopen(path, "r");
    if (!f) return TBK_ERR_ARG;
    tbk::FortranReader rd(f);                 // unit 32: blanks ignored
    int rc = TBK_OK;
    std::memset(ps, 0, sizeof(*ps));
    do {
        // (2a1,i5): basis type, magnetism, parameter style
        if (!rd.next_record()) { rc = TBK_ERR_ARG; break; }
        const std::string head = rd.field(2);
        int nltype = 0;
        if (!rd.integer(5, &nltype)) { rc = TBK_ERR_ARG; break; }
        const char ov = head[0], mg = head[1];
        ps->jsover = !(ov == 'O' || ov == 'o');          // anything but O: non-orthogonal
        ps->jspins = (mg == 'M' || mg == 'm') ? 2 : 1;   // anything but M: one spin
        if (nltype != 90000 && (nltype > 3 || nltype < 0)) nltype = 0;
        ps->nltype = nltype;
        if (!rd.next_record()) { rc = TBK_ERR_ARG; break; }
        const std::string title = rd.chars(79);
        std::snprintf(ps->title, sizeof(ps->title), "%s", title.c_str());
        int kinds = 0;
        if (!rd.next_record() || !rd.list_int(&kinds) || kinds < 1 || kinds > TBK_MAXKINDS) {
            rc = TBK_ERR_ARG;
            break;
        }
        ps->kinds = kinds;
        for (int i = 0; i < kinds * (kinds + 1) / 2 && !rc; ++i)
            if (!rd.next_record() || !rd.list_real(ps->dnn + i) || !rd.list_real(ps->screenl + i))
                rc = TBK_ERR_ARG;
        for (int i = 0; i < kinds && !rc; ++i) {
            double wt = 0, s = 0, p = 0, d = 0;
            if (!rd.next_record() || !rd.list_int(ps->kbas + i) || !rd.next_record() || !rd.list_real(&wt) ||
                !rd.next_record() || !rd.list_real(&s) || !rd.list_real(&p) || !rd.list_real(&d))
                rc = TBK_ERR_ARG;
            ps->valence[i] = s + p + d;
        }
        if (rc) break;
        const int npards = kinds * (29 + 68 * kinds);
        if (cap < ps->jspins * npards) { rc = TBK_ERR_ARG; break; }
        // one value per record; the first unreadable record ends a channel
        // and zeroes the rest of it (input.f:327-357)
        for (int js = 0; js < ps->jspins; ++js) {
            double *pr = param + (size_t)js * npards;
            int n = npards;
            for (int i = 0; i < npards; ++i) {
                double v = 0.0;
                if (!rd.next_record() || !rd.list_real(&v)) {
                    n = i;
                    for (int j = i; j < npards; ++j) pr[j] = 0.0;
                    break;
                }
                pr[i] = v;
            }
            if (js == 0) ps->nparam = n;
        }
    } while (false);
    std::fclose(f);
    return rc;
}

// ---------------------------------------------------------------------------
// SKIN (static.f:682-1255, input.f:139, setlat.f, input1.f, kptin.f)
// ---------------------------------------------------------------------------
struct tbk_skin {
    FILE *f = nullptr;
    tbk::FortranReader *rd = nullptr;
    std::string dir;
    int nstruct = 0;
    std::vector<double> q, w;          // the k-points of the last NEWSYM=T block
    bool have_k = false;
};

extern "C" int tbk_skin_open(const char *path, tbk_skin **out, int *mode, double *tkb, double *eigcut,
                             char *param_path, int param_path_cap)
{
    if (!path || !out || !mode || !tkb || !eigcut || !param_path || param_path_cap < 2) return TBK_ERR_ARG;
    *out = nullptr;
    FILE *f = std::fopen(path, "r");
    if (!f) return TBK_ERR_ARG;
    tbk_skin *s = new tbk_skin;
    s->f = f;
    s->rd = new tbk::FortranReader(f, true);   // static.f:682 opens SKIN with blank='zero'
    s->dir = dirname_of(path);
    tbk::FortranReader &rd = *s->rd;
    int rc = TBK_OK;
    // (5x,a2) and modeparse.f: two digits, a non-digit first one taken as 0
    if (!rd.next_record()) rc = TBK_ERR_ARG;
    if (!rc) {
        rd.skip(5);
        const std::string m = rd.field(2);
        int d0 = std::isdigit((unsigned char)m[0]) ? m[0] - '0' : -1;
        const int d1 = std::isdigit((unsigned char)m[1]) ? m[1] - '0' : -1;
        if (d0 < 0) d0 = 0;
        *mode = d1 < 0 ? d0 : 10 * d0 + d1;
        if (!rd.next_record() || !rd.list_real(tkb) || !rd.list_real(eigcut)) rc = TBK_ERR_ARG;
    }
    if (!rc) {
        if (!rd.next_record()) rc = TBK_ERR_ARG;
        else {
            const std::string name = resolve(s->dir, trim(rd.chars(79)));
            if ((int)name.size() + 1 > param_path_cap) rc = TBK_ERR_ARG;
            else std::memcpy(param_path, name.c_str(), name.size() + 1);
        }
    }
    if (rc) { tbk_skin_close(s); return rc; }
    *out = s;
    return TBK_OK;
}

extern "C" int tbk_skin_close(tbk_skin *s)
{
    if (!s) return TBK_ERR_ARG;
    delete s->rd;
    if (s->f) std::fclose(s->f);
    delete s;
    return TBK_OK;
}

namespace {

// kptin.f:60-366: the k-point block of a structure.  Returns TBK_OK or
// TBK_ERR_ARG where the reference stops or its reads fail.
int read_kpoints(tbk_skin *s, const double *lat, int maxk)
{
    tbk::FortranReader &rd = *s->rd;
    int newsym = 0;
    if (!rd.next_record()) return TBK_ERR_ARG;
    rd.skip(7);
    if (!rd.logical(1, &newsym)) return TBK_ERR_ARG;
    if (!newsym) return s->have_k ? TBK_OK : TBK_ERR_ARG;     // the previous structure's k-points
    int lattic = 0, ilat = 0;
    if (!rd.next_record()) return TBK_ERR_ARG;
    rd.skip(7);
    if (!rd.integer(1, &lattic)) return TBK_ERR_ARG;
    double avec[9], bvec[9], bij[9], wsvol = 0;
    if (tbk_reciprocal_lattice(lat, avec, bvec, bij, &wsvol)) return TBK_ERR_ARG;
    if (!rd.next_record()) return TBK_ERR_ARG;
    const std::string sg = resolve(s->dir, trim(rd.record()));
    if (!rd.next_record()) return TBK_ERR_ARG;
    rd.skip(5);
    if (!rd.logical(1, &ilat)) return TBK_ERR_ARG;
    std::vector<double> rspc(9 * kNopd), rspct(9 * kNopd), tnp(3 * kNopd);
    int nop = 0;
    if (tbk_space_group_read(sg.c_str(), ilat, lattic, avec, bvec, kNopd, &nop, rspc.data(), rspct.data(),
                             tnp.data(), nullptr))
        return TBK_ERR_ARG;
    // READ(10,1000) NKV, SFK with FORMAT(I5,2F10.6); -1313: the rest from a file
    int nkv = 0;
    double sfk = 0;
    if (!rd.next_record() || !rd.integer(5, &nkv) || !rd.real(10, 6, &sfk)) return TBK_ERR_ARG;
    FILE *kf = nullptr;
    tbk::FortranReader *krd = &rd;
    tbk::FortranReader file_rd(nullptr);
    if (nkv == -1313) {
        if (!rd.next_record()) return TBK_ERR_ARG;
        const std::string kfile = resolve(s->dir, trim(rd.field(255)));
        kf = std::fopen(kfile.c_str(), "r");
        if (!kf) return TBK_ERR_ARG;                            // kptin.f:13000 stops
        file_rd = tbk::FortranReader(kf);                       // unit 33: blanks ignored
        krd = &file_rd;
        if (!krd->next_record() || !krd->integer(5, &nkv) || !krd->real(10, 6, &sfk)) {
            std::fclose(kf);
            return TBK_ERR_ARG;
        }
    }
    const double eps = 1e-6;
    if (sfk <= eps) sfk = 1.0;
    int rc = TBK_OK;
    std::vector<double> q, w;
    if (nkv > 0) {
        if (nkv == 999) {                                     // line mode
            int kpdir = 0;
            if (!krd->next_record() || !krd->integer(5, &kpdir) || kpdir < 1) rc = TBK_ERR_ARG;
            std::vector<int> kd(kpdir > 0 ? kpdir : 0);
            std::vector<double> ends(6 * kd.size());
            for (int p = 0; p < kpdir && !rc; ++p) {
                if (!krd->next_record() || !krd->list_int(&kd[p])) { rc = TBK_ERR_ARG; break; }
                for (int j = 0; j < 6 && !rc; ++j)
                    if (!krd->list_real(&ends[6 * p + j])) rc = TBK_ERR_ARG;
            }
            if (!rc) {
                q.assign(3 * (size_t)maxk, 0.0);
                w.assign(maxk, 0.0);
                int nk = 0;
                rc = tbk_kpath(kpdir, kd.data(), ends.data(), maxk, &nk, q.data(), w.data());
                q.resize(3 * (size_t)nk);
                w.resize(nk);
            }
        } else {                                              // explicit list
            if (nkv > kNkd || nkv > maxk) rc = TBK_ERR_ARG;
            std::vector<double> raw(4 * (size_t)nkv);
            for (int i = 0; i < nkv && !rc; ++i) {
                if (!krd->next_record()) { rc = TBK_ERR_ARG; break; }
                for (int j = 0; j < 4 && !rc; ++j)
                    if (!krd->list_real(&raw[4 * (size_t)i + j])) rc = TBK_ERR_ARG;
            }
            if (!rc) {
                q.assign(3 * (size_t)nkv, 0.0);
                w.assign(nkv, 0.0);
                rc = tbk_klist(nkv, raw.data(), sfk, q.data(), w.data());
            }
        }
    } else {                                                  // generated mesh
        int lspec = 0, linfok = 0, kdiv[3] = {0, 0, 0}, kpst = 0, kpend = 0;
        if (!krd->next_record()) rc = TBK_ERR_ARG;
        if (!rc) {
            krd->skip(6);
            if (!krd->logical(1, &lspec)) rc = TBK_ERR_ARG;
            krd->skip(8);
            if (!rc && !krd->logical(1, &linfok)) rc = TBK_ERR_ARG;
        }
        if (!rc && (!krd->next_record() || !krd->integer(5, kdiv) || !krd->integer(5, kdiv + 1) ||
                    !krd->integer(5, kdiv + 2) || !krd->integer(5, &kpst) || !krd->integer(5, &kpend)))
            rc = TBK_ERR_ARG;
        const int nkinfo = linfok ? -nkv : 0;
        if (!rc) {
            q.assign(3 * (size_t)(kNkd + nkinfo), 0.0);
            w.assign(kNkd + nkinfo, 0.0);
            int nk = 0;
            rc = tbk_kptgen(lattic, avec, bvec, bij, wsvol, nop, rspct.data(), kdiv, lspec, 2 * kNkd, kNkd, &nk,
                            q.data(), w.data());
            for (int i = 0; i < nkinfo && !rc; ++i, ++nk) {
                if (nk >= kNkd || !krd->next_record()) { rc = TBK_ERR_ARG; break; }
                for (int j = 0; j < 3 && !rc; ++j)
                    if (!krd->list_real(&q[3 * (size_t)nk + j])) rc = TBK_ERR_ARG;
                for (int j = 0; j < 3; ++j) q[3 * (size_t)nk + j] = q[3 * (size_t)nk + j] / sfk;
                w[nk] = 0.0;
            }
            int nout = 0;
            if (!rc) rc = tbk_kmesh_weights(nk, q.data(), w.data(), kpst, kpend, &nout);
            if (!rc) {
                q.resize(3 * (size_t)nout);
                w.resize(nout);
            }
        }
    }
    if (kf) std::fclose(kf);
    if (rc) return TBK_ERR_ARG;
    if ((int)w.size() > maxk || w.empty()) return TBK_ERR_ARG;
    s->q = q;
    s->w = w;
    s->have_k = true;
    return TBK_OK;
}

}  // namespace

extern "C" int tbk_skin_next(tbk_skin *s, const tbk_param_set *ps, int maxatoms, int maxk, tbk_skin_structure *st,
                             int *kkind, double *posn, double *posq, int *flipos, double *spinat, double *kpts,
                             double *weights)
{
    if (!s || !ps || !st || !kkind || !posn || !posq || maxatoms < 1 || maxk < 1 || !kpts || !weights)
        return TBK_ERR_ARG;
    tbk::FortranReader &rd = *s->rd;
    std::memset(st, 0, sizeof(*st));
    // static.f:938-950: the title, the label and zcharge; the end of the file
    // (or an unreadable title) ends the program normally
    if (!rd.next_record()) return TBK_END;
    std::snprintf(st->title, sizeof(st->title), "%s", rd.field(70).c_str());
    if (!rd.next_record()) return TBK_END;
    std::snprintf(st->label, sizeof(st->label), "%s", rd.field(20).c_str());
    if (!rd.next_record() || !rd.list_real(&st->zcharge)) return TBK_END;
    // input1.f:469: the lattice
    int ierr = setlat(rd, st->lat, &st->thisvol, &st->latnum, &st->strain);
    if (ierr) { st->inpstat = 1000 + ierr; return TBK_ERR_ARG; }
    const double *plv = st->lat;                 // plv[3*v + c]
    int natoms = 0;
    if (!rd.next_record() || !rd.list_int(&natoms)) return TBK_ERR_ARG;
    if (!rd.next_record() || !rd.list_int(st->nd) || !rd.list_int(st->nd + 1) || !rd.list_int(st->nd + 2))
        return TBK_ERR_ARG;
    // (2a1): phonon and spin flags (input1.f:522-548)
    if (!rd.next_record()) { st->inpstat = 102; return TBK_ERR_ARG; }
    const std::string fl = rd.field(2);
    auto is = [](char c, char u) { return c == u || c == (char)(u - 'A' + 'a'); };
    st->lphonon = is(fl[0], 'T');
    st->spinflip = is(fl[1], 'T');
    st->ltorque = is(fl[1], 'D');
    st->lspdir = is(fl[1], 'S');
    st->lnoncol = is(fl[1], 'N') || st->ltorque || st->lspdir;
    // bvec(i, j) = (plv(i1,j1) plv(i2,j2) - plv(i1,j2) plv(i2,j1)) / vol
    double bv[3][3];
    for (int i = 0; i < 3; ++i) {
        const int i1 = (i + 1) % 3, i2 = (i1 + 1) % 3;
        for (int j = 0; j < 3; ++j) {
            const int j1 = (j + 1) % 3, j2 = (j1 + 1) % 3;
            bv[i][j] = (plv[3 * i1 + j1] * plv[3 * i2 + j2] - plv[3 * i1 + j2] * plv[3 * i2 + j1]) / st->thisvol;
        }
    }
    // frozen-phonon shifts (input1.f:569-687), kept in lattice coordinates
    std::vector<int> lshift;
    std::vector<double> epsl;
    if (st->lphonon) {
        int nshift = 0;
        if (!rd.next_record() || !rd.list_int(&nshift)) return TBK_ERR_ARG;
        if (nshift > natoms) { st->inpstat = 103; return TBK_ERR_ARG; }
        lshift.assign(natoms > 0 ? natoms : 0, 0);
        epsl.assign(3 * lshift.size(), 0.0);
        for (int k = 0; k < nshift; ++k) {
            int iatom = 0, lc = 0;
            double et[3] = {0, 0, 0};
            if (!rd.next_record() || !rd.list_int(&iatom) || !rd.list_logical(&lc) || !rd.list_real(et) ||
                !rd.list_real(et + 1) || !rd.list_real(et + 2))
                return TBK_ERR_ARG;
            if (iatom > natoms || iatom < 1) { st->inpstat = 104; return TBK_ERR_ARG; }
            if (lshift[iatom - 1]) st->inpstat = 105;              // reported, the run goes on
            lshift[iatom - 1] = 1;
            for (int i = 0; i < 3; ++i) {
                if (lc) {
                    double sum = 0.0;
                    for (int j = 0; j < 3; ++j) sum = sum + et[j] * bv[i][j];
                    epsl[3 * (iatom - 1) + i] = sum;
                } else {
                    epsl[3 * (iatom - 1) + i] = et[i];
                }
            }
        }
    }
    if (natoms > maxatoms || natoms < 1) { st->inpstat = 101; return TBK_ERR_ARG; }
    st->natoms = natoms;
    // atoms (input1.f:735-861): kind, position, phase origin [, angles | spin]
    const double deg2rad = kPi / 180.0;
    double zfill = st->zcharge;
    for (int i = 0; i < natoms; ++i) {
        double p[3], pq[3];
        if (!rd.next_record() || !rd.list_int(kkind + i)) return TBK_ERR_ARG;
        for (int j = 0; j < 3; ++j)
            if (!rd.list_real(p + j)) return TBK_ERR_ARG;
        for (int j = 0; j < 3; ++j)
            if (!rd.list_real(pq + j)) return TBK_ERR_ARG;
        if (st->lnoncol) {
            double th = 0, ph = 0;
            if (!rd.list_real(&th) || !rd.list_real(&ph)) return TBK_ERR_ARG;
            if (spinat) {
                spinat[3 * i] = std::sin(deg2rad * th) * std::cos(deg2rad * ph);
                spinat[3 * i + 1] = std::sin(deg2rad * th) * std::sin(deg2rad * ph);
                spinat[3 * i + 2] = std::cos(deg2rad * th);
            }
        } else if (st->spinflip) {
            int magdir = 0;
            if (!rd.list_int(&magdir)) return TBK_ERR_ARG;
            if (flipos) flipos[i] = magdir == 2;
        } else if (flipos) {
            flipos[i] = 0;
        }
        if (kkind[i] < 0) {                      // Cartesian position: to lattice coordinates
            kkind[i] = -kkind[i];
            double v[3];
            for (int a = 0; a < 3; ++a) {
                double sum = 0.0;
                for (int b = 0; b < 3; ++b) sum = sum + bv[a][b] * p[b];
                v[a] = sum;
            }
            for (int a = 0; a < 3; ++a) p[a] = v[a];
        }
        if (kkind[i] < 1 || kkind[i] > ps->kinds) return TBK_ERR_ARG;
        zfill = zfill + ps->valence[kkind[i] - 1];
        if (st->lphonon && lshift[i])
            for (int j = 0; j < 3; ++j) p[j] = p[j] + epsl[3 * i + j];
        // static.f:1088-1103: lattice to Cartesian coordinates
        for (int c = 0; c < 3; ++c) {
            double t = 0.0;
            for (int v = 0; v < 3; ++v) t = t + plv[3 * v + c] * p[v];
            posn[3 * i + c] = t;
        }
        for (int j = 0; j < 3; ++j) posq[3 * i + j] = pq[j];
    }
    st->zfill = zfill;
    // kptin.f: the k-point block
    if (read_kpoints(s, plv, maxk)) return TBK_ERR_ARG;
    st->nk = (int)s->w.size();
    std::memcpy(kpts, s->q.data(), sizeof(double) * s->q.size());
    std::memcpy(weights, s->w.data(), sizeof(double) * s->w.size());
    ++s->nstruct;
    return TBK_OK;
}

// ---------------------------------------------------------------------------
// SKENG (static.f:1238-1254)
// ---------------------------------------------------------------------------
namespace {
// Fortran Fw.d output: right-justified, w asterisks when it does not fit
std::string fmt_f(double x, int w, int d)
{
    char buf[128];
    if (std::isnan(x)) {                      // Fortran writes NaN right-justified
        std::snprintf(buf, sizeof(buf), "%*s", w, "NaN");
        return std::string(buf);
    }
    std::snprintf(buf, sizeof(buf), "%*.*f", w, d, x);
    std::string s(buf);
    if ((int)s.size() > w) s.assign(w, '*');
    return s;
}
// Fortran Ew.d / Dw.d output: 0.d1d2...dd followed by E+xx (or D+xx)
std::string fmt_e(double x, int w, int d, char letter = 'E')
{
    std::string m;
    int ex = 0;
    if (x == 0.0) {
        m = "0." + std::string(d, '0');
    } else {
        char buf[64];
        std::snprintf(buf, sizeof(buf), "%.*e", d - 1, std::fabs(x));   // d.ddde+xx: the same d digits
        std::string t(buf);
        const size_t epos = t.find('e');
        std::string digits;
        for (size_t i = 0; i < epos; ++i)
            if (t[i] != '.') digits.push_back(t[i]);
        ex = std::atoi(t.c_str() + epos + 1) + 1;
        m = (x < 0 ? "-0." : "0.") + digits;
    }
    char eb[16];
    std::snprintf(eb, sizeof(eb), "%c%c%02d", letter, ex < 0 ? '-' : '+', std::abs(ex));
    std::string s = m + eb;
    if ((int)s.size() > w) return std::string(w, '*');
    return std::string(w - s.size(), ' ') + s;
}
std::string fmt_i(long v, int w)
{
    char buf[64];
    std::snprintf(buf, sizeof(buf), "%*ld", w, v);
    std::string s(buf);
    if ((int)s.size() > w) s.assign(w, '*');
    return s;
}

// One k-point's record header of a QLMT file, (3f10.6,2X,f10.5,i10,d20.10,' K')
// (bande.f:1628-1629)
std::string qlmt_k_line(const double *ak, double wt, int ntot)
{
    return fmt_f(ak[0], 10, 6) + fmt_f(ak[1], 10, 6) + fmt_f(ak[2], 10, 6) + "  " + fmt_f(wt, 10, 5) +
           fmt_i(ntot, 10) + fmt_e(wt, 20, 10, 'D') + " K";
}

// A QAPW record, (1x,3f9.6,10x,4f9.5/(38x,4f9.5)) (bande.f:1690-1691): with
// exactly four levels the slash still ends the record and an empty one
// follows, as the Fortran runtime writes it.
std::string qapw_record(const double *ak, const double *e, int n)
{
    std::string out = " " + fmt_f(ak[0], 9, 6) + fmt_f(ak[1], 9, 6) + fmt_f(ak[2], 9, 6) + std::string(10, ' ');
    for (int i = 0; i < n && i < 4; ++i) out += fmt_f(e[i], 9, 5);
    out += "\n";
    if (n == 4) out += "\n";
    for (int i = 4; i < n; i += 4) {
        out += std::string(38, ' ');
        for (int j = i; j < n && j < i + 4; ++j) out += fmt_f(e[j], 9, 5);
        out += "\n";
    }
    return out;
}
}  // namespace

extern "C" int tbk_skeng_line(const char *label, double thisvol, double efermi, double earg1, double earg2,
                              double smag, int mode, int with_smag, char *buf, int cap)
{
    if (!label || !buf || cap < 96) return TBK_ERR_ARG;
    std::string lab(label);
    if (lab.size() > 20) lab.resize(20);
    lab.append(20 - lab.size(), ' ');
    std::string s = lab + fmt_f(thisvol, 12, 6) + fmt_f(efermi, 15, 9) + fmt_f(earg1, 15, 9);
    if (mode < 3 || mode == 4) s += fmt_f(earg2, 15, 9);
    if (with_smag) s += fmt_f(smag, 10, 6);
    if ((int)s.size() + 1 > cap) return TBK_ERR_ARG;
    std::memcpy(buf, s.c_str(), s.size() + 1);
    return TBK_OK;
}

// ---------------------------------------------------------------------------
// QLMT / QAPW records of one structure after its run (bande.f:1568-1703, as
// diag.f:635-758 forms them): for each spin channel and k-point, modes 6 and
// 10 the states up to eigcut with the Lowdin weights per atom (s, p, d sums;
// mode 10 every orbital), mode 9 every eigenvalue, mode 8 (QAPW) every
// eigenvalue in its 4-per-line layout.
// ---------------------------------------------------------------------------
namespace {
int write_qlmt(FILE *qf, tbk_plan *p, int mode, double eigcut, const tbk_skin_structure &st, const int *kkind,
               const tbk_param_set &ps, bool noncol, int natx, const double *kp, const double *wt)
{
    const int nv = tbk_plan_nv(p), nk = st.nk, nstr = tbk_plan_nstruct(p);
    std::vector<double> esk((size_t)nk * nv), desk((size_t)nk * nv), e(nv), proj((size_t)nv * nv);
    for (int s = 0; s < nstr; ++s) {               // spin channels (jsx, bande.f:711-721)
        int rc = tbk_plan_structure_eigenvalues(p, s, esk.data(), desk.data());
        if (rc) return rc;
        for (int k = 0; k < nk; ++k) {
            const double *ak = kp + 3 * (size_t)k;
            const double *ek = esk.data() + (size_t)k * nv;
            if (mode == 8) {
                std::fputs(qapw_record(ak, ek, nv).c_str(), qf);
                continue;
            }
            if (mode == 9) {
                std::fprintf(qf, "%s\n", qlmt_k_line(ak, wt[k], nv).c_str());
                for (int i = 0; i < nv; ++i)
                    std::fprintf(qf, "%s%s\n", fmt_f(ek[i], 12, 7).c_str(), fmt_e(0.0, 15, 6).c_str());
                continue;
            }
            int ntot = 0;
            rc = tbk_plan_projections(p, s * nk + k, eigcut, &ntot, e.data(), proj.data());
            if (rc) return rc;
            std::fprintf(qf, "%s\n", qlmt_k_line(ak, wt[k], ntot).c_str());
            for (int j = 0; j < ntot; ++j) {
                const double en = e[j];
                const bool wide = mode == 6 && (en <= -10.0 || en >= 100.0);
                std::fprintf(qf, "%s%s\n", fmt_f(en, 10, wide ? 6 : 7).c_str(), fmt_e(0.0, 15, 6).c_str());
                int i = 0;
                for (int a = 0; a < natx; ++a) {
                    const int atom = a < st.natoms ? a : a - st.natoms;
                    const int nbas = ps.kbas[kkind[atom] - 1];
                    const double *occ = proj.data() + (size_t)j * nv + i;
                    i += nbas;
                    std::string line;
                    if (mode == 10) {
                        for (int b = 0; b < nbas; ++b) line += fmt_f(occ[b], 15, 9);
                    } else {
                        double sv = occ[0], pv = 0.0, dv = 0.0;
                        for (int b = 1; b < std::min(4, nbas); ++b) pv = pv + occ[b];
                        for (int b = 4; b < nbas; ++b) dv = dv + occ[b];
                        line = fmt_f(sv, 15, 9) + fmt_f(pv, 15, 9) + fmt_f(dv, 15, 9) + fmt_f(0.0, 15, 9);
                    }
                    std::fprintf(qf, "%s\n", line.c_str());
                }
            }
        }
    }
    std::fflush(qf);
    return TBK_OK;
}
}  // namespace

// ---------------------------------------------------------------------------
// The program's loop over the structures of a SKIN file (static.f:925-1255)
// ---------------------------------------------------------------------------
extern "C" int tbk_static_run(const char *dir, int device, int precision, int *nstruct)
{
    if (!dir || !nstruct) return TBK_ERR_ARG;
    *nstruct = 0;
    const std::string d(dir);
    tbk_skin *skin = nullptr;
    int mode = 0;
    double tkb = 0, eigcut = 0;
    char ppath[1024];
    int rc = tbk_skin_open((d + "/SKIN").c_str(), &skin, &mode, &tkb, &eigcut, ppath, (int)sizeof(ppath));
    if (rc) return rc;
    tbk_param_set ps;
    std::vector<double> param(2 * TBK_MAXKINDS * (29 + 68 * TBK_MAXKINDS));
    rc = tbk_read_params(ppath, &ps, (int)param.size(), param.data());
    const int npard = rc ? 0 : ps.kinds * (29 + 68 * ps.kinds);
    if (!rc && ps.nparam != npard) rc = TBK_ERR_ARG;            // static.f:864-874
    // modes 1-4 and the QLMT (6, 9, 10) / QAPW (8) modes; 5 aborts in the
    // reference's own program, 7 it does not implement
    const bool qmode = mode == 6 || mode == 8 || mode == 9 || mode == 10;
    if (!rc && !((mode >= 1 && mode <= 4) || qmode)) rc = TBK_ERR_ARG;
    FILE *out = rc ? nullptr : std::fopen((d + "/SKENG").c_str(), "w");
    if (!rc && !out) rc = TBK_ERR_ARG;
    FILE *qf = nullptr;                                          // static.f:757-778
    FILE *spf = nullptr;                                         // SPNDIR (unit 42), static.f:1010-1017
    if (!rc && qmode) {
        qf = std::fopen((d + (mode == 8 ? "/QAPW" : "/QLMT")).c_str(), "w");
        if (!qf) rc = TBK_ERR_ARG;
    }
    const int maxatoms = 4096, maxk = kNkd;
    std::vector<int> kkind(maxatoms), flipos(maxatoms);
    std::vector<double> posn(3 * maxatoms), posq(3 * maxatoms), spinat(3 * maxatoms), kp(3 * (size_t)maxk),
        wt(maxk);
    while (!rc) {
        tbk_skin_structure st;
        const int r = tbk_skin_next(skin, &ps, maxatoms, maxk, &st, kkind.data(), posn.data(), posq.data(),
                                    flipos.data(), spinat.data(), kp.data(), wt.data());
        if (r == TBK_END) break;                 // end of the SKIN file
        if (r) { rc = r; break; }
        bool noncol = st.lnoncol != 0;
        if (noncol && ps.jspins == 1) noncol = false;              // static.f:982-996
        // static.f:983-1004 turns ltorque off with the non-collinear run but
        // leaves lspdir on, and diag.f:268-283 then stops; refused here
        if (!noncol && st.lspdir) { rc = TBK_ERR_ARG; break; }
        const bool spdir = noncol && st.lspdir != 0;
        if (noncol && st.ltorque) {
            // static.f:1012 opens SKDERIV; bande computes the spin-angle
            // derivatives of spinham (bande.f:897-901) and writes nothing there
            FILE *fd = std::fopen((d + "/SKDERIV").c_str(), "a");
            if (!fd) { rc = TBK_ERR_ARG; break; }
            std::fclose(fd);
        }
        if (spdir && !spf) {
            spf = std::fopen((d + "/SPNDIR").c_str(), "w");
            if (!spf) { rc = TBK_ERR_ARG; break; }
        }
        const int natx = noncol ? 2 * st.natoms : st.natoms;
        if (qf && mode != 8) {        // static.f:1127-1137, 1152-1157: the structure's header lines
            std::fprintf(qf, "%s%s%s%s     SPINS, WIND., SC. WIND.\n", fmt_i(1, 5).c_str(), fmt_i(1, 5).c_str(),
                         fmt_i(0, 5).c_str(), fmt_i(mode == 9 ? 0 : st.natoms, 5).c_str());
            std::fprintf(qf, "%s K-POINTS%s%s%s\n", fmt_i(st.nk, 5).c_str(), fmt_i(st.natoms, 6).c_str(),
                         fmt_f(st.zfill, 10, 5).c_str(), fmt_f(tkb, 10, 5).c_str());
        }
        tbk_structure_desc sd;
        sd.natoms = st.natoms;
        sd.lat = st.lat; sd.posn = posn.data(); sd.posq = posq.data(); sd.kkind = kkind.data();
        for (int c = 0; c < 3; ++c) sd.nd[c] = st.nd[c];
        sd.kinds = ps.kinds; sd.kbas = ps.kbas; sd.valence = ps.valence; sd.dnn = ps.dnn;
        sd.screenl = ps.screenl; sd.param = param.data(); sd.nparam = npard; sd.nltype = ps.nltype;
        sd.jsover = ps.jsover; sd.thisvol = st.thisvol; sd.zfill = st.zfill;
        tbk_plan *p = nullptr;
        double ef = 0, e1 = 0, e2 = 0, smag = 0;
        if (ps.jspins == 1)
            rc = tbk_plan_create(device, &sd, st.nk, kp.data(), wt.data(), 0, precision, &p);
        else if (noncol)
            rc = tbk_plan_create_noncol(device, &sd, param.data() + npard, spinat.data(), st.nk, kp.data(),
                                        wt.data(), 0, precision, &p);
        else
            rc = tbk_plan_create_spin(device, &sd, param.data() + npard, flipos.data(), st.nk, kp.data(),
                                      wt.data(), 0, precision, &p);
        if (!rc && (mode == 6 || mode == 10)) rc = tbk_plan_set_window(p, 0);
        if (!rc && spdir) rc = tbk_plan_set_spin_directions(p, 1);
        if (!rc) rc = tbk_plan_run(p, mode, tkb, eigcut, &ef, &e1, &e2);
        if (!rc && spdir) {         // bande.f:1541-1560: (a20,i5,4f10.6,2f12.6) per atom
            std::vector<double> sv(3 * (size_t)st.natoms);
            rc = tbk_plan_spin_directions(p, 0, sv.data());
            const double rad2deg = 180.0 / 3.141592653589793116;
            std::string lab(st.label);
            lab.resize(20, ' ');
            for (int i = 0; !rc && i < st.natoms; ++i) {
                const double x = sv[3 * i], y = sv[3 * i + 1], z = sv[3 * i + 2];
                const double smagl = std::sqrt(x * x + y * y + z * z);
                const double theta = rad2deg * std::acos(z / smagl), phi = rad2deg * std::atan2(y, x);
                std::fprintf(spf, "%s%s%s%s%s%s%s%s\n", lab.c_str(), fmt_i(i + 1, 5).c_str(), fmt_f(x, 10, 6).c_str(),
                             fmt_f(y, 10, 6).c_str(), fmt_f(z, 10, 6).c_str(), fmt_f(smagl, 10, 6).c_str(),
                             fmt_f(theta, 12, 6).c_str(), fmt_f(phi, 12, 6).c_str());
            }
            std::fflush(spf);
        }
        const bool with_smag = ps.jspins == 2 && !noncol;
        if (!rc && with_smag) rc = tbk_plan_magnetization(p, &smag);
        if (!rc && qf) rc = write_qlmt(qf, p, mode, eigcut, st, kkind.data(), ps, noncol, natx, kp.data(), wt.data());
        if (p) tbk_plan_destroy(p);
        if (rc) break;
        char line[160];
        rc = tbk_skeng_line(st.label, st.thisvol, ef, e1, e2, smag, mode, with_smag ? 1 : 0, line, (int)sizeof(line));
        if (rc) break;
        std::fprintf(out, "%s\n", line);
        std::fflush(out);
        ++*nstruct;
    }
    if (out) std::fclose(out);
    if (qf) std::fclose(qf);
    if (spf) std::fclose(spf);
    tbk_skin_close(skin);
    return rc;
}
