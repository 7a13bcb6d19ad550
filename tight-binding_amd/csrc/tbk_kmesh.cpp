// tbk_kmesh.cpp -- k-point meshes: the reciprocal lattice, space-group
// operations, the irreducible Monkhorst-Pack / special-point mesh and the
// weight handling that feed the k-point engine.
//
// Restates kptin.f (reciprocal lattice and weight normalisation), rdspgp.f
// (space-group file and the similarity transforms of its operations),
// kptgen.f (mesh, Wigner-Seitz shift, shell sort and folding by the
// operations and time reversal), mapk.f, gdot.f, matmlt.f and hsort.f of
// NRLtb/src.  Compiled with -ffp-contract=off and without -march, and every
// sum is formed in the reference's order, so the k-points and weights are
// bit-identical to those the reference program passes to bande.
//
// Conventions of the C interface (all row-major C arrays):
//   lat[3*v + c]   component c of lattice vector v (bohr), as plv(v, c)
//   avec[3*v + c]  the same vectors (AVEC(c, v))
//   bvec[3*v + c]  reciprocal vectors with 2 pi (BVEC(c, v))
//   bij[3*i + j]   metric b_i . b_j (BIJ(i, j))
//   rot[9*k + 3*i + j]  operation k, row i, column j (R(i, j, k))
//   q[3*n + c]     k-point n in reciprocal-lattice coordinates
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "tbk.h"
#include "tbk_fortran_io.h"

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kTwoPi = 2.0 * kPi;
constexpr double kFourPi = 4.0 * kPi;
constexpr double kInvTwoPi = 1.0 / kTwoPi;   // tpiinv of mapk.f / rdspgp.f
constexpr double kShellTol = 1e-4;           // toler, kptgen.f:47
constexpr double kFoldTol = 1e-4;            // kptgen.f:243-247
constexpr double kWeightEps = 1e-6;          // EPS, kptin.f:36
// Single-precision literals of rdspgp.f:155-157 and :266 (0.001 without a
// kind suffix), promoted to double as the reference's compiler does.
const double kHexSnap = (double)0.001f;
const double kShiftMin = (double)0.001f;

// Fortran-indexed views of the row-major arrays: F(i, j) is element (i, j),
// 1-based, of the column-major Fortran array the reference holds.
struct Mat3 {
    double m[3][3];   // m[i][j] = A(i + 1, j + 1)
};

// matmlt.f: C = A B, each element accumulated over the inner index from 0.
Mat3 matmul(const Mat3 &a, const Mat3 &b)
{
    Mat3 c{};
    for (int k = 0; k < 3; ++k)
        for (int j = 0; j < 3; ++j) {
            const double s = b.m[j][k];
            for (int i = 0; i < 3; ++i) c.m[i][k] = a.m[i][j] * s + c.m[i][k];
        }
    return c;
}

struct Lattice {
    Mat3 a, b, g;     // AVEC, BVEC, BIJ (Fortran index order)
    double vol;
};

Lattice lattice_from_arrays(const double *avec, const double *bvec, const double *bij, double wsvol)
{
    Lattice L{};
    for (int v = 0; v < 3; ++v)
        for (int c = 0; c < 3; ++c) {
            L.a.m[c][v] = avec[3 * v + c];
            L.b.m[c][v] = bvec[3 * v + c];
        }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) L.g.m[i][j] = bij[3 * i + j];
    L.vol = wsvol;
    return L;
}

// gdot.f: a^T G b, inner sums over the first index.
double gdot(const double *a, const double *b, const Mat3 &g)
{
    double r = 0.0;
    for (int j = 0; j < 3; ++j) {
        double s = 0.0;
        for (int i = 0; i < 3; ++i) s = s + a[i] * g.m[i][j];
        r = r + s * b[j];
    }
    return r;
}

// mapk.f with inv = 1: Cartesian k -> reciprocal-lattice coordinates.
void cart_to_lattice(const Lattice &L, const double *rk, double *uk)
{
    for (int j = 0; j < 3; ++j)
        uk[j] = (rk[0] * L.a.m[0][j] + rk[1] * L.a.m[1][j] + rk[2] * L.a.m[2][j]) * kInvTwoPi;
}

// hsort.f: heap sort of an index vector by key (ascending).  The order of
// equal keys is the one this particular heap sort leaves, which fixes the
// representative that kptgen keeps of each star, so the sift steps follow
// hsort.f:1-48 exactly.  Indices are 1-based, as in the reference.
void heap_index_sort(int n, const std::vector<double> &key, std::vector<int> &ind)
{
    ind.assign(n + 1, 0);
    for (int j = 1; j <= n; ++j) ind[j] = j;
    if (n <= 1) return;
    auto a = [&](int i) { return key[i - 1]; };
    int l = n / 2 + 1, ir = n;
    for (;;) {
        int held;
        double q;
        if (l > 1) {
            --l;
            held = ind[l];
            q = a(held);
        } else {
            held = ind[ir];
            q = a(held);
            ind[ir] = ind[1];
            if (--ir == 1) {
                ind[1] = held;
                return;
            }
        }
        int i = l, j = l + l;
        while (j <= ir) {
            if (j < ir && a(ind[j]) < a(ind[j + 1])) ++j;
            if (q < a(ind[j])) {
                ind[i] = ind[j];
                i = j;
                j += j;
            } else {
                j = ir + 1;
            }
        }
        ind[i] = held;
    }
}


}  // namespace

extern "C" {

// kptin.f:94-155: AVEC from plv, BVEC = 2 pi (a2 x a3, a3 x a1, a1 x a2)/V
// with the signed volume V = a1 . (a2 x a3), BIJ = BVEC^T BVEC, |V|.
int tbk_reciprocal_lattice(const double *lat, double *avec, double *bvec, double *bij, double *wsvol)
{
    if (!lat || !avec || !bvec || !bij || !wsvol) return TBK_ERR_ARG;
    double A[9], B[9];
    for (int v = 0; v < 3; ++v)
        for (int c = 0; c < 3; ++c) A[3 * v + c] = lat[3 * v + c];   // A(L), L = c + 3 v
    // cross products in the reference's operand order (kptin.f:102-110)
    B[0] = A[4] * A[8] - A[5] * A[7];
    B[1] = A[5] * A[6] - A[3] * A[8];
    B[2] = A[3] * A[7] - A[4] * A[6];
    B[3] = A[7] * A[2] - A[8] * A[1];
    B[4] = A[8] * A[0] - A[6] * A[2];
    B[5] = A[6] * A[1] - A[7] * A[0];
    B[6] = A[1] * A[5] - A[2] * A[4];
    B[7] = A[2] * A[3] - A[0] * A[5];
    B[8] = A[0] * A[4] - A[1] * A[3];
    const double v = A[0] * B[0] + A[1] * B[1] + A[2] * B[2];
    const double pref = kTwoPi / v;
    for (int i = 0; i < 9; ++i) {
        avec[i] = A[i];
        bvec[i] = pref * B[i];
    }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k) s = bvec[3 * i + k] * bvec[3 * j + k] + s;
            bij[3 * i + j] = s;
        }
    *wsvol = std::fabs(v);
    return TBK_OK;
}

// rdspgp.f:106-289 after the read: rotations given in direct-lattice (ilat)
// or Cartesian coordinates become RSPC (Cartesian) and RSPCT (reciprocal-
// lattice coordinates); hexagonal Cartesian entries within 0.001 of
// +-sqrt(3)/2 are snapped to it; a nonzero origin shift moves the
// non-primitive translations.
int tbk_space_group(int ilat, int lattic, const double *avec, const double *bvec, int nop, const double *rot,
                    const double *tnp, const double *shift, double *rspc, double *rspct, double *tnpspc)
{
    if (!avec || !bvec || !rot || !tnp || !rspc || !rspct || !tnpspc || nop < 1) return TBK_ERR_ARG;
    Mat3 a{}, b{};
    for (int v = 0; v < 3; ++v)
        for (int c = 0; c < 3; ++c) {
            a.m[c][v] = avec[3 * v + c];
            b.m[c][v] = bvec[3 * v + c];
        }
    // direct <-> Cartesian similarity: T = BVEC^T / 2 pi, TINV = AVEC
    Mat3 t{}, tinv = a;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) t.m[i][j] = b.m[j][i] * kInvTwoPi;
    std::vector<Mat3> dir(nop);
    for (int k = 0; k < nop; ++k) {
        Mat3 r{};
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) r.m[i][j] = rot[9 * k + 3 * i + j];
        Mat3 cart;
        if (ilat) {
            dir[k] = r;
            cart = matmul(tinv, matmul(r, t));
        } else {
            cart = r;
            if (lattic == 4) {
                const double c = std::sqrt(3.0) / 2.0;
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j) {
                        if (std::fabs(cart.m[i][j] - c) < kHexSnap) cart.m[i][j] = c;
                        if (std::fabs(cart.m[i][j] + c) < kHexSnap) cart.m[i][j] = -c;
                    }
            }
            dir[k] = matmul(t, matmul(cart, tinv));
        }
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) rspc[9 * k + 3 * i + j] = cart.m[i][j];
    }
    // Cartesian -> reciprocal-lattice similarity: T = AVEC^T / 2 pi, TINV = BVEC
    Mat3 t2{}, tinv2 = b;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) t2.m[i][j] = a.m[j][i] / kTwoPi;
    for (int k = 0; k < nop; ++k) {
        Mat3 cart{};
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) cart.m[i][j] = rspc[9 * k + 3 * i + j];
        const Mat3 rc = matmul(t2, matmul(cart, tinv2));
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) rspct[9 * k + 3 * i + j] = rc.m[i][j];
    }
    double s = 0.0;
    if (shift)
        for (int j = 0; j < 3; ++j) s = s + std::fabs(shift[j]);
    for (int k = 0; k < nop; ++k) {
        double r[3] = {0.0, 0.0, 0.0};
        if (s > kShiftMin)
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) r[i] = r[i] + dir[k].m[i][j] * shift[j];
        for (int j = 0; j < 3; ++j)
            tnpspc[3 * k + j] = s > kShiftMin ? r[j] + tnp[3 * k + j] - shift[j] : tnp[3 * k + j];
    }
    return TBK_OK;
}

// rdspgp.f:117-147: the space-group file (name, count and origin shift;
// rotations as 18F4.0 records in direct-lattice coordinates or 3D20.10
// records in Cartesian ones; translations as 9F8.5 records), then the
// transforms of tbk_space_group.
int tbk_space_group_read(const char *path, int ilat, int lattic, const double *avec, const double *bvec,
                         int maxop, int *nop, double *rspc, double *rspct, double *tnpspc, char *name)
{
    if (!path || !avec || !bvec || !nop || !rspc || !rspct || !tnpspc || maxop < 1) return TBK_ERR_ARG;
    FILE *f = std::fopen(path, "r");
    if (!f) return TBK_ERR_ARG;
    tbk::FortranReader rd(f);        // rdspgp.f:83-90 (unit 25, blanks ignored)
    int rc = TBK_OK, n = 0;
    double shift[3] = {0.0, 0.0, 0.0};
    std::vector<double> rot, tnp;
    do {
        if (!rd.next_record()) { rc = TBK_ERR_ARG; break; }
        const std::string nm = rd.field(10);
        if (name) {
            std::memcpy(name, nm.data(), 10);
            name[10] = '\0';
        }
        if (!rd.integer(5, &n) || n < 1 || n > maxop) { rc = TBK_ERR_ARG; break; }
        rd.skip(5);
        for (int j = 0; j < 3; ++j)
            if (!rd.real(10, 5, shift + j)) rc = TBK_ERR_ARG;
        if (rc) break;
        rot.assign(9 * (size_t)n, 0.0);
        tnp.assign(3 * (size_t)n, 0.0);
        bool ok = true;
        if (ilat) {
            ok = rd.real_list(9 * n, 18, 4, 0, rot.data());
        } else {
            for (int k = 0; k < n && ok; ++k) ok = rd.real_list(9, 3, 20, 10, rot.data() + 9 * k);
        }
        ok = ok && rd.real_list(3 * n, 9, 8, 5, tnp.data());
        if (!ok) { rc = TBK_ERR_ARG; break; }
    } while (false);
    std::fclose(f);
    if (rc) return rc;
    *nop = n;
    return tbk_space_group(ilat, lattic, avec, bvec, n, rot.data(), tnp.data(), shift, rspc, rspct, tnpspc);
}

// kptgen.f:72-268: the uniform (lspec = 0) or special-point (lspec = 1)
// mesh of kdiv divisions in reciprocal-lattice coordinates (the bcc lattice
// through its simple-cubic cell, the body-centred tetragonal ones 7-9 with
// doubled in-plane divisions), each point moved into the Wigner-Seitz cell
// by the shortest of its images under -3..3 reciprocal translations, sorted
// by length, grouped into shells 1e-4 wide, and folded: a point equal
// within 1e-4 (sum of component differences) to an operation image of an
// earlier point of its shell, or to its negative, adds its weight 1/KIN to
// that point (KIN = the points of the half-open cell).  maxfull bounds the
// full mesh (kptgen's nkbzd; <= 0: no bound), maxk the irreducible points.
int tbk_kptgen(int lattic, const double *avec, const double *bvec, const double *bij, double wsvol, int nop,
               const double *rspct, const int *kdiv, int lspec, int maxfull, int maxk, int *nk, double *q,
               double *w)
{
    if (!avec || !bvec || !bij || !rspct || !kdiv || !nk || !q || !w || nop < 1 || maxk < 1) return TBK_ERR_ARG;
    if (kdiv[0] < 1 || kdiv[1] < 1 || kdiv[2] < 1) return TBK_ERR_ARG;
    const Lattice L = lattice_from_arrays(avec, bvec, bij, wsvol);
    const bool bcc = lattic == 2, hex = lattic == 4, tet = lattic == 7 || lattic == 8 || lattic == 9;
    double alat = 0.0;
    if (bcc) {
        alat = std::pow(2.0 * L.vol, 0.33333333333333333);
        alat = kFourPi / alat;
    }
    int d1 = kdiv[0], d2 = kdiv[1];
    const int d3 = kdiv[2];
    if (tet) { d1 *= 2; d2 *= 2; }
    int m1 = d1, m2 = d2, m3 = d3;
    if (lspec) {
        if (hex) {
            if (d1 != d2) return TBK_ERR_ARG;     // kptgen.f:104 (STOP)
            m1 = d1 + d1 % 2;
            m2 = m1;
        } else {
            m1 = d1 + 1;
            m2 = d2 + 1;
        }
        m3 = d3 + 1;
    }
    const double f1 = 0.5 / (double)d1, f2 = 0.5 / (double)d2, f3 = 0.5 / (double)d3;
    std::vector<double> rk, rad;
    std::vector<char> inside;
    int kin = 0;
    for (int ii = -m1; ii <= m1; ii += 2)
        for (int jj = -m2; jj <= m2; jj += 2)
            for (int kk = -m3; kk <= m3; kk += 2) {
                double uk[3] = {f1 * (double)ii, f2 * (double)jj, f3 * (double)kk};
                if (lattic == 7 || lattic == 8 || lattic == 9) {
                    uk[0] = (double)ii / (double)d1;
                    uk[1] = (double)jj / (double)d2;
                    if (lattic == 7) uk[2] = uk[2] + 0.5 * (uk[0] - uk[1]);
                    if (lattic == 8) uk[2] = uk[2] + 0.5 * uk[0];
                    if (lattic == 9) uk[2] = uk[2] + 0.5 * (uk[0] + uk[1]);
                }
                if (maxfull > 0 && (int)rad.size() + 1 > maxfull) return TBK_ERR_ARG;   // kptgen.f:141 (STOP)
                const bool in = ii <= d1 && ii > -d1 && jj <= d2 && jj > -d2 && kk <= d3 && kk > -d3;
                inside.push_back(in);
                if (in) ++kin;
                double u[3] = {uk[0], uk[1], uk[2]};
                if (bcc) {
                    const double uu[3] = {uk[0] * alat, uk[1] * alat, uk[2] * alat};
                    cart_to_lattice(L, uu, u);
                }
                double best = 1e50, bt[3] = {0.0, 0.0, 0.0};
                for (int i1 = -3; i1 <= 3; ++i1)
                    for (int i2 = -3; i2 <= 3; ++i2)
                        for (int i3 = -3; i3 <= 3; ++i3) {
                            const double uu[3] = {u[0] + (double)i1, u[1] + (double)i2, u[2] + (double)i3};
                            const double r = gdot(uu, uu, L.g);
                            if (r <= best) {
                                best = r;
                                bt[0] = uu[0]; bt[1] = uu[1]; bt[2] = uu[2];
                            }
                        }
                rk.insert(rk.end(), bt, bt + 3);
                rad.push_back(best);
            }
    const int nfull = (int)rad.size();
    if (kin == 0) return TBK_ERR_ARG;
    std::vector<int> idx;
    heap_index_sort(nfull, rad, idx);
    const double fw = 1.0 / (double)kin;
    std::vector<double> srad(nfull), srk(3 * (size_t)nfull);
    for (int j = 0; j < nfull; ++j) {
        const int o = idx[j + 1] - 1;
        srad[j] = rad[o];
        for (int c = 0; c < 3; ++c) srk[3 * j + c] = rk[3 * o + c];
    }
    // shells: a new shell starts where the length exceeds the previous
    // shell's first length by more than the tolerance
    std::vector<int> start{0};
    double lim = srad[0] + kShellTol;
    for (int j = 1; j < nfull; ++j)
        if (srad[j] > lim) {
            lim = srad[j] + kShellTol;
            start.push_back(j);
        }
    start.push_back(nfull);
    int nir = 0;
    for (size_t s = 0; s + 1 < start.size(); ++s) {
        const int k1 = nir;           // first irreducible point of this shell
        for (int j = start[s]; j < start[s + 1]; ++j) {
            const double xx = srk[3 * j], yy = srk[3 * j + 1], zz = srk[3 * j + 2];
            int ieq = 0;              // 1-based match, the last one found
            for (int k = k1; k < nir; ++k) {
                const double *u = q + 3 * (size_t)k;
                for (int op = 0; op < nop; ++op) {
                    const double *r = rspct + 9 * (size_t)op;
                    for (int i1 = -1; i1 <= 1; ++i1) {
                        const double x = r[0] * u[0] + r[1] * u[1] + r[2] * u[2] + (double)i1;
                        for (int i2 = -1; i2 <= 1; ++i2) {
                            const double y = r[3] * u[0] + r[4] * u[1] + r[5] * u[2] + (double)i2;
                            for (int i3 = -1; i3 <= 1; ++i3) {
                                const double z = r[6] * u[0] + r[7] * u[1] + r[8] * u[2] + (double)i3;
                                if (std::fabs(x - xx) + std::fabs(y - yy) + std::fabs(z - zz) <= kFoldTol)
                                    ieq = k + 1;
                                if (ieq == 0 &&
                                    std::fabs(x + xx) + std::fabs(y + yy) + std::fabs(z + zz) <= kFoldTol)
                                    ieq = k + 1;
                            }
                        }
                    }
                }
            }
            const bool in = inside[idx[j + 1] - 1] != 0;
            if (ieq == 0) {
                if (nir >= maxk) return TBK_ERR_ARG;
                q[3 * (size_t)nir] = xx;
                q[3 * (size_t)nir + 1] = yy;
                q[3 * (size_t)nir + 2] = zz;
                w[nir] = in ? fw : 0.0;
                ++nir;
            } else if (in) {
                w[ieq - 1] = w[ieq - 1] + fw;
            }
        }
    }
    *nk = nir;
    return TBK_OK;
}

// kptin.f:312-351, the weights of a generated mesh (with any zero-weight
// information points already appended): the sum S of the weights; if
// |S - 1| > 1e-6 every weight divided by S; then either the range
// kpst..kpend (1-based, kpst <= 0 meaning 1, kpend = 0 meaning nk) kept and
// scaled by the inverse of its weight sum, or, for the whole list, every
// weight multiplied by S.  On return *nk_out points (moved to the front).
int tbk_kmesh_weights(int nk, double *q, double *w, int kpst, int kpend, int *nk_out)
{
    if (nk < 1 || !q || !w || !nk_out) return TBK_ERR_ARG;
    double sum = 0.0;
    for (int i = 0; i < nk; ++i) sum = sum + w[i];
    if (std::fabs(sum - 1.0) > kWeightEps) {
        const double inv = 1.0 / sum;
        for (int i = 0; i < nk; ++i) w[i] = w[i] * inv;
    }
    if (kpst <= 0) kpst = 1;
    if (kpend == 0) kpend = nk;
    if (kpst > kpend || kpend > nk || kpend < 0) return TBK_ERR_ARG;
    int n = nk;
    if (kpst > 1 || kpend < nk) {
        double s = 0.0;
        n = 0;
        for (int i = kpst - 1; i < kpend; ++i, ++n) {
            w[n] = w[i];
            s = s + w[n];
            for (int c = 0; c < 3; ++c) q[3 * n + c] = q[3 * i + c];
        }
        sum = 1.0 / s;
    }
    for (int i = 0; i < n; ++i) w[i] = w[i] * sum;
    *nk_out = n;
    return TBK_OK;
}

// kptin.f:266-273 and :348-351, an explicit list (k in units of 1/sfk,
// sfk <= 1e-6 meaning 1; raw[4 n] = k, weight): k divided by sfk, the
// weights scaled by the inverse of their sum.
int tbk_klist(int nk, const double *raw, double sfk, double *q, double *w)
{
    if (nk < 1 || !raw || !q || !w) return TBK_ERR_ARG;
    if (sfk <= kWeightEps) sfk = 1.0;
    double sum = 0.0;
    for (int i = 0; i < nk; ++i) {
        for (int c = 0; c < 3; ++c) q[3 * i + c] = raw[4 * i + c] / sfk;
        w[i] = raw[4 * i + 3];
        sum = sum + w[i];
    }
    sum = 1.0 / sum;
    for (int i = 0; i < nk; ++i) w[i] = w[i] * sum;
    return TBK_OK;
}

// kptin.f:223-251 (nkv = 999), straight paths: segment p runs from
// ends[6p..6p+2] in kd[p] points, each point the previous one plus
// (end - start)/(kd - 1); every weight is the total point count, as the
// reference leaves it.  maxk bounds the total.
int tbk_kpath(int npath, const int *kd, const double *ends, int maxk, int *nk, double *q, double *w)
{
    if (npath < 1 || !kd || !ends || !nk || !q || !w) return TBK_ERR_ARG;
    int n = 0;
    for (int p = 0; p < npath; ++p) {
        if (kd[p] < 1 || n + kd[p] > maxk) return TBK_ERR_ARG;
        const double *a = ends + 6 * p, *b = a + 3;
        double step[3] = {0.0, 0.0, 0.0};
        if (kd[p] > 1)
            for (int c = 0; c < 3; ++c) step[c] = (b[c] - a[c]) / (double)(kd[p] - 1);
        // the start and end share one slot when kd = 1 (read into the same
        // QKPT column, the end last)
        for (int c = 0; c < 3; ++c) q[3 * n + c] = kd[p] == 1 ? b[c] : a[c];
        w[n++] = 1.0;
        for (int s = 2; s <= kd[p]; ++s, ++n) {
            for (int c = 0; c < 3; ++c) q[3 * n + c] = q[3 * (n - 1) + c] + step[c];
            w[n] = 1.0;
        }
    }
    const double total = (double)n;
    for (int i = 0; i < n; ++i) w[i] = w[i] * total;
    *nk = n;
    return TBK_OK;
}

}  // extern "C"
