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
// rebuild with -D for others.  Where the reference stops (full mesh above
// 2 nkd, more than nkd irreducible points, hexagonal special points with
// kdiv1 != kdiv2) the drop-in prints the reason and leaves nkv(nw) = 0.
#include <cstdio>
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

void kptgen_(int *nw)
{
    const int m = *nw - 1;
    if (m < 0 || m >= TBK_NWDD) {
        std::fprintf(stderr, "kptgen (drop-in): window %d outside 1..%d\n", *nw, TBK_NWDD);
        return;
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
        std::fprintf(stderr, "kptgen (drop-in): %d operations\n", nop);
        kmesh1_.nkv[m] = 0;
        return;
    }
    std::vector<double> rspct(9 * (size_t)nop);
    for (int k = 0; k < nop; ++k)
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) rspct[9 * k + 3 * i + j] = space_.rspct[k][j][i];
    const int kdiv[3] = {kmesh2_.kpdiv1[m], kmesh2_.kpdiv2[m], kmesh2_.kpdiv3[m]};
    std::vector<double> q(3 * (size_t)TBK_NKD), w(TBK_NKD);
    int nk = 0;
    const int rc = tbk_kptgen(lattyp_.lattic, avec, bvec, bij, lattyp_.wsvol, nop, rspct.data(), kdiv,
                              kmesh2_.lspec[m] ? 1 : 0, 2 * TBK_NKD, TBK_NKD, &nk, q.data(), w.data());
    if (rc) {
        std::fprintf(stderr, "kptgen (drop-in): %s (divisions %d %d %d)\n", tbk_error_string(rc), kdiv[0],
                     kdiv[1], kdiv[2]);
        kmesh1_.nkv[m] = 0;
        return;
    }
    for (int k = 0; k < nk; ++k) {
        for (int c = 0; c < 3; ++c) kmesh1_.qkpt[m][k][c] = q[3 * (size_t)k + c];
        kmesh1_.wghtk[m][k] = w[k];
    }
    kmesh1_.nkv[m] = nk;
}
}
