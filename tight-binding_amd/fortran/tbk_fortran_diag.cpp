// tbk_fortran_diag.cpp -- Fortran-ABI drop-ins for the reference's per-k
// eigensolve and derivatives: diag (NRLtb/src/diag.f:1-3) and perturb
// (NRLtb/src/perturb.f:1-2), for a caller that keeps its own k-loop (the
// reference's bande, or a program of the user's) and swaps in only the
// device solver.  Each call moves one k's matrices to the device and back:
// the batched plans (bande_ in tbk_fortran.cpp, the C ABI) are the fast
// path; these are the finer-grained boundary.
//
// Link this file's object (its own member of lib/libtbk_fortran.a) in place
// of diag.o / perturb.o, with lib/libtbk.so.  zeig's leading dimension is
// the reference's mhe (P2: mcol (meigen mbas (matom - 1) + mbas)), given
// here by TBK_MCOL, TBK_MEIGEN, TBK_MBAS, TBK_MATOM (defaults: NRLtb/ex1/P1,
// P2); rebuild with -D for other dimensions.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "tbk.h"

#ifndef TBK_MATOM
#define TBK_MATOM 20
#endif
#ifndef TBK_MBAS
#define TBK_MBAS 9
#endif
#ifndef TBK_MCOL
#define TBK_MCOL 2
#endif
#ifndef TBK_MEIGEN
#define TBK_MEIGEN 1
#endif
#define TBK_MHE (TBK_MCOL * (TBK_MEIGEN * TBK_MBAS * (TBK_MATOM - 1) + TBK_MBAS))   // P2: mhe

typedef int f_logical;

static int device_ordinal()
{
    const char *e = std::getenv("TBK_DEVICE");
    return e ? std::atoi(e) : 0;
}

// Packed upper (LAPACK 'U', column-major: AP(i + j(j-1)/2) = A(i, j), i <= j,
// 1-based) to a full Hermitian n x n matrix, column-major, interleaved.
static void unpack_upper(int n, const double *ap, std::vector<double> &a)
{
    a.assign(2 * (size_t)n * n, 0.0);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i <= j; ++i) {
            const size_t p = (size_t)i + (size_t)j * (j + 1) / 2;
            const double re = ap[2 * p], im = ap[2 * p + 1];
            a[2 * ((size_t)j * n + i)] = re;
            a[2 * ((size_t)j * n + i) + 1] = im;
            a[2 * ((size_t)i * n + j)] = re;
            a[2 * ((size_t)i * n + j) + 1] = i == j ? im : -im;
        }
}

static int zhegv_device(int n, const double *h, const double *s, double *w, double *z)
{
    std::vector<double> zz;
    if (!z) { zz.resize(2 * (size_t)n * n); z = zz.data(); }
    return n <= 512 ? tbk_zhegv_noalloc(device_ordinal(), n, h, s, w, z)
                    : tbk_zhegv(device_ordinal(), n, h, s, w, z);
}

// diag.f:1-455 for modes 1-4: zhpgv (or zhpev for an orthogonal basis) on
// the device, the 1 x 1 pencil in closed form (diag.f:320-349), and on an
// S(k) that is not positive definite ioverr = 1 with en = the eigenvalues
// of S when meigen = 1 (diag.f:429-455).  The QLMT/QAPW writers (modes 5,
// 6, 8, 9, 10) and spin directions (lspdir) are not provided here: ioverr =
// -1, on which bande stops (bande.f:987-995).
extern "C" void diag_(int *mode, f_logical *lnoncol, double *hmat, double *smat, double *en, f_logical *lspdir,
                      double *spev, int *n, double *zeig, f_logical *jsover, double *ak, double *eigcut,
                      double *weight, int *ioverr, int *natoms, int *kkind, int *kbas, int *ksk)
{
    (void)lnoncol; (void)spev; (void)ak; (void)eigcut; (void)weight; (void)natoms; (void)kkind; (void)kbas;
    (void)ksk;
    const int m = *mode, nn = *n;
    *ioverr = 0;
    if (m < 1 || m > 4 || *lspdir) {
        std::printf(" diag (device drop-in): mode %d%s not provided\n", m, *lspdir ? " with spin directions" : "");
        *ioverr = -1;
        return;
    }
    const bool leigen = m == 2 || m == 4;
    if (nn == 1) {
        const double h1 = hmat[0];
        if (*jsover) {
            const double s1 = smat[0];
            if (s1 > 0.0) {
                en[0] = h1 / s1;
            } else {
                en[0] = std::fabs(h1 / s1);
                *ioverr = 1;
                std::printf(" Found s(1) = %g use %g\n", s1, en[0]);
                std::fprintf(stderr, " Found s(1) = %g use %g\n", s1, en[0]);
            }
        } else {
            en[0] = h1;
        }
        return;
    }
    std::vector<double> H, S, W(nn), Z(leigen ? 2 * (size_t)nn * nn : 0);
    unpack_upper(nn, hmat, H);
    if (*jsover) {
        unpack_upper(nn, smat, S);
    } else {
        S.assign(2 * (size_t)nn * nn, 0.0);
        for (int i = 0; i < nn; ++i) S[2 * ((size_t)i * nn + i)] = 1.0;
    }
    const int rc = zhegv_device(nn, H.data(), S.data(), W.data(), leigen ? Z.data() : nullptr);
    if (rc == TBK_ERR_OVERLAP) {
        std::printf(" Error return from zhpgv:  S(k) is not positive definite\n");
        std::printf(" Probably a negative eigenvalue in S\n");
        if (TBK_MEIGEN == 1) {
            std::vector<double> I(2 * (size_t)nn * nn, 0.0);
            for (int i = 0; i < nn; ++i) I[2 * ((size_t)i * nn + i)] = 1.0;
            if (zhegv_device(nn, S.data(), I.data(), en, nullptr) == TBK_OK) {
                std::printf(" The eigenvalues of S are:  \n");
                for (int j = 0; j < nn; ++j) std::printf("%5d%15.8f\n", j + 1, en[j]);
            }
        }
        *ioverr = 1;
        return;
    }
    if (rc) {
        std::fprintf(stderr, " diag (device drop-in): %s\n", tbk_error_string(rc));
        *ioverr = -1;
        return;
    }
    for (int j = 0; j < nn; ++j) en[j] = W[j];
    if (leigen)
        for (int j = 0; j < nn; ++j)
            for (int i = 0; i < nn; ++i) {
                zeig[2 * ((size_t)j * TBK_MHE + i)] = Z[2 * ((size_t)j * nn + i)];
                zeig[2 * ((size_t)j * TBK_MHE + i) + 1] = Z[2 * ((size_t)j * nn + i) + 1];
            }
}

// perturb.f:1-221 on the device (tbk_perturb): dhmat, dsmat packed upper,
// z(mxd, mxd) the eigenvectors of the last diag.  A degenerate group of more
// than 9 states stops the program as the reference's does (perturb.f:106).
extern "C" void perturb_(int *mxd, int *n, double *dhmat, double *dsmat, double *z, double *eig, double *eigcut,
                         int *ntot, double *deig, double *work)
{
    (void)work;
    const int nn = *n, ld = *mxd;
    std::vector<double> DH, DS, Z(2 * (size_t)nn * nn), D(nn);
    unpack_upper(nn, dhmat, DH);
    unpack_upper(nn, dsmat, DS);
    for (int j = 0; j < nn; ++j)
        for (int i = 0; i < nn; ++i) {
            Z[2 * ((size_t)j * nn + i)] = z[2 * ((size_t)j * ld + i)];
            Z[2 * ((size_t)j * nn + i) + 1] = z[2 * ((size_t)j * ld + i) + 1];
        }
    int nt = 0;
    const int rc = tbk_perturb(device_ordinal(), nn, DH.data(), DS.data(), Z.data(), eig, *eigcut, &nt, D.data());
    if (rc == TBK_ERR_DEGEN) {
        std::printf(" dgen:  Degeneracy is too large\n");
        std::fflush(stdout);
        std::exit(0);
    }
    if (rc) {
        std::fprintf(stderr, " perturb (device drop-in): %s\n", tbk_error_string(rc));
        std::exit(1);
    }
    *ntot = nt;
    for (int j = 0; j < nt; ++j) deig[j] = D[j];
}
