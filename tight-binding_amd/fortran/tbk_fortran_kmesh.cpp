// tbk_fortran_kmesh.cpp -- Fortran-ABI drop-in for the reference's kptgen
// (NRLtb/src/kptgen.f:1, called by kptin.f:293).
//
// kptgen_ reads the lattice (/LATTYP/, filled by kptin), the operations
// (/SPACE/, filled by rdspgp) and the mesh divisions (/KMESH2/), and writes
// the irreducible k-points and weights into /KMESH1/ through tbk_kptgen,
// bit-identical to the reference routine.  It sits in its own archive
// member of lib/libtbk_fortran.a, so a program can take it alone (with the
// reference's bande.o) or together with bande_.
//
// COMMON dimensions follow NRLtb/ex1/P3 (TBK_NKD, TBK_NOPD, TBK_NWDD);
// rebuild with -D for others.  Where the reference stops the program
// (hexagonal special points with kdiv1 != kdiv2, kptgen.f:103; a full mesh
// above nkbzd = 2 nkd points, kptgen.f:141-145) the drop-in prints the same
// lines and ends the program with status 0, as Fortran's STOP does; more
// than nkd irreducible points (past the end of /KMESH1/ in the reference)
// stops it the same way.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "tbk.h"

#ifndef TBK_NKD
#define TBK_NKD 50000
#endif
#ifndef TBK_NOPD
#define TBK_NOPD 48
#endif
#ifndef TBK_NWDD
#define TBK_NWDD 1
#endif

typedef int f_logical;

extern "C" {
// COMMON /LATTYP/ AVEC(3,3),BVEC(3,3),BIJ(3,3),WSVOL,LATTIC (kptgen.f:36)
extern struct { double avec[3][3], bvec[3][3], bij[3][3], wsvol; int lattic; } lattyp_;
// COMMON /SPACE/ RSPC(3,3,NOPD),RSPCT(3,3,NOPD),TNPSPC(3,NOPD),NOP (kptgen.f:37)
extern struct {
    double rspc[TBK_NOPD][3][3], rspct[TBK_NOPD][3][3], tnpspc[TBK_NOPD][3];
    int nop;
} space_;
// common /kmesh1/ qkpt(3,nkd,nwdd),wghtk(nkd,nwdd),nkv(nwdd) (kptgen.f:38)
extern struct {
    double qkpt[TBK_NWDD][TBK_NKD][3], wghtk[TBK_NWDD][TBK_NKD];
    int nkv[TBK_NWDD];
} kmesh1_;
// COMMON /KMESH2/ KPDIV1(NWDD),KPDIV2(NWDD),KPDIV3(NWDD),LSPEC(NWDD) (kptgen.f:39-40)
extern struct { int kpdiv1[TBK_NWDD], kpdiv2[TBK_NWDD], kpdiv3[TBK_NWDD]; f_logical lspec[TBK_NWDD]; } kmesh2_;

// Fortran STOP 'text': the message on standard output, then a normal end
// of the program (exit handlers flush and close the Fortran units).
[[noreturn]] static void fortran_stop(const char *text)
{
    std::printf(" %s\n", text);
    std::fflush(stdout);
    std::exit(0);
}

void kptgen_(int *nw)
{
    const int m = *nw - 1;
    if (m < 0 || m >= TBK_NWDD) {
        std::printf(" KPTGEN (drop-in): window %d outside 1..%d\n", *nw, TBK_NWDD);
        fortran_stop("KPTGEN: NO K-POINT SET");
    }
    // Fortran (i, j) of a 3x3 COMMON array sits at C [j][i]: the row-major
    // vectors of the C interface are the Fortran columns
    double avec[9], bvec[9], bij[9];
    for (int v = 0; v < 3; ++v)
        for (int c = 0; c < 3; ++c) {
            avec[3 * v + c] = lattyp_.avec[v][c];
            bvec[3 * v + c] = lattyp_.bvec[v][c];
        }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) bij[3 * i + j] = lattyp_.bij[j][i];
    const int nop = space_.nop;
    if (nop < 1 || nop > TBK_NOPD) {
        std::printf(" KPTGEN (drop-in): %d operations, outside 1..%d\n", nop, TBK_NOPD);
        fortran_stop("KPTGEN: NO K-POINT SET");
    }
    std::vector<double> rspct(9 * (size_t)nop);
    for (int k = 0; k < nop; ++k)
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) rspct[9 * k + 3 * i + j] = space_.rspct[k][j][i];
    const int kdiv[3] = {kmesh2_.kpdiv1[m], kmesh2_.kpdiv2[m], kmesh2_.kpdiv3[m]};
    {
        // the full mesh's size, as kptgen.f:78-112 sets its loop bounds
        const bool tet = lattyp_.lattic == 7 || lattyp_.lattic == 8 || lattyp_.lattic == 9;
        long d1 = kdiv[0], d2 = kdiv[1];
        const long d3 = kdiv[2];
        if (tet) { d1 *= 2; d2 *= 2; }
        long m1 = d1, m2 = d2, m3 = d3;
        if (kmesh2_.lspec[m]) {
            if (lattyp_.lattic == 4) {
                if (d1 != d2) fortran_stop("KPTGEN: KDIV ERROR FOR HEX LAT.");
                m1 = d1 + d1 % 2;
                m2 = m1;
            } else {
                m1 = d1 + 1;
                m2 = d2 + 1;
            }
            m3 = d3 + 1;
        }
        const long nkbzd = 2L * TBK_NKD;
        if (d1 >= 1 && d2 >= 1 && d3 >= 1 && (m1 + 1) * (m2 + 1) * (m3 + 1) > nkbzd) {
            std::printf(" KPTGEN: NKBZD DIM. EXCEED., NK= %7ld\n", nkbzd + 1);
            std::printf(" KPTGEN: NKBZD DIM. EXCEED., NK= %7ld\n", nkbzd + 1);
            fortran_stop("KPTGEN: NKBZD DIMENSION EXCEEDED");
        }
    }
    std::vector<double> q(3 * (size_t)TBK_NKD), w(TBK_NKD);
    int nk = 0;
    const int rc = tbk_kptgen(lattyp_.lattic, avec, bvec, bij, lattyp_.wsvol, nop, rspct.data(), kdiv,
                              kmesh2_.lspec[m] ? 1 : 0, 2 * TBK_NKD, TBK_NKD, &nk, q.data(), w.data());
    if (rc) {
        std::printf(" KPTGEN (drop-in): %s (divisions %d %d %d; at most %d irreducible points)\n",
                    tbk_error_string(rc), kdiv[0], kdiv[1], kdiv[2], TBK_NKD);
        fortran_stop("KPTGEN: NO K-POINT SET");
    }
    for (int k = 0; k < nk; ++k) {
        for (int c = 0; c < 3; ++c) kmesh1_.qkpt[m][k][c] = q[3 * (size_t)k + c];
        kmesh1_.wghtk[m][k] = w[k];
    }
    kmesh1_.nkv[m] = nk;
}
}
