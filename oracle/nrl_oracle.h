/*
 * nrl_oracle.h -- CPU restatement of the NRL tight-binding per-k-point path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * product path: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  Nothing in the product links or calls it.
 *
 * Every routine restates one routine of the reference program `static`
 * (NRLtb/src, v1.33) and cites the file:line it follows.  Array layouts are
 * C layouts and are stated per argument.  Indices of atoms and kinds are
 * 1-based in the stored lists (as in the reference) so fixtures can be
 * compared element by element.
 */
#ifndef NRL_ORACLE_H
#define NRL_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Number of Slater-Koster integrals per pair (mpkind, NRLtb/src/P1:109). */
#define ORC_NPK 14

/*
 * search2.f:1-240 -- neighbour list.
 *   lat[3*v+c]    component c of primitive vector v (bohr)  (= plv(v,c))
 *   posn[3*a+c]   Cartesian position of atom a (bohr)        (= posn(a,c))
 *   kkind[a]      kind of atom a (1-based)
 *   nd[3]         search half-widths nd1..nd3 (input1.f:488-492)
 *   dnn[int], screenl[int]  cutoff and screening length per interaction type
 * Outputs (size maxpairs): tt[3p+c], dlv[3p+c], dist[p], screen[p], dscrn[p],
 *   jkind[2p+{0,1}], jatm[2p+{0,1}]; *npairs; *dismin.
 * Returns 0, or 1 on pair overflow (search2.f:180-188).
 */
int orc_search2(int natoms, const double *lat, const double *posn,
                const int *kkind, int kinds, const int *nd,
                const double *dnn, const double *screenl, int maxpairs,
                int *npairs, double *tt, double *dlv, double *dist,
                double *screen, double *dscrn, int *jkind, int *jatm,
                double *dismin);

/*
 * setpar.f:1-891 (collinear, jspin = 1, no spin flips).
 *   param[npard]  parameters of one spin channel (input.f:327-359 order)
 * Outputs: den[kinds*natoms] (den[(a)*kinds + k]), paros[4*natoms],
 *   hop[14*npairs], over[14*npairs].
 */
int orc_setpar(int natoms, const int *kkind, int kinds, const double *param,
               int nltype, int realov, int harrison, int jsover,
               int npairs, const double *dist, const double *screen,
               const int *jkind, const int *jatm,
               double *den, double *paros, double *hop, double *over);

/* setparv.f:1-888: volume derivatives dparos[4*natoms], dhop, dover[14*np]. */
int orc_setparv(int natoms, const int *kkind, int kinds, const double *param,
                int nltype, int realov, int harrison, int jsover,
                double thisvol, int npairs, const double *dist,
                const double *screen, const double *dscrn,
                const int *jkind, const int *jatm, const double *den,
                double *dparos, double *dhop, double *dover);

/*
 * setham.f:1-499 + rotate.f:1-337: packed upper-triangle H(k), S(k).
 *   ak[3]           k in reciprocal-lattice units
 *   ksk[a]          first orbital of atom a (0-based offset, input1.f:835)
 *   kbas[kind-1]    orbitals per kind
 *   posq[3a+c]      phase origin of atom a (lattice units)
 *   hp, sp          outputs: complex128 packed upper, column-major,
 *                   interleaved (re,im), length nv(nv+1)/2 complex.
 * sethamv.f is the same routine fed with dhop/dover/dparos and onsite_s = 0.
 */
void orc_setham(int nv, const double *ak, int natoms, const int *kkind,
                const int *ksk, const int *kbas, const double *posq,
                int npairs, const double *tt, const double *dlv,
                const int *jkind, const int *jatm, const double *hop,
                const double *over, const double *paros, int jsover,
                int onsite_s, double *hp, double *sp);

/* Unpack a packed-upper complex matrix into full Hermitian column-major. */
void orc_unpack(int n, const double *ap, double *a);

/*
 * diag.f:300-455 -> zhpgv(1,'V','U') (lapack.f:2274-2473): generalized
 * Hermitian eigensolve on full column-major complex H, S (destroyed).
 * Outputs w[n] ascending, z[n*n] complex column-major with Z^H S Z = I
 * (z may be NULL for eigenvalues only).  Returns 0, or j>0 when S is not
 * positive definite at column j (zpptrf semantics, lapack.f:4466).
 */
int orc_zhegv(int n, double *h, double *s, double *w, double *z);

/*
 * perturb.f:1-222: first-order eigenvalue derivatives.  dh, ds are packed
 * upper (as from orc_setham); z full column-major (leading dim n).
 * Returns 0, or 1 when a degenerate group exceeds 9 (perturb.f:106).
 */
int orc_perturb(int n, const double *dhp, const double *dsp, const double *z,
                const double *eig, double eigcut, int *ntot, double *deig);

/*
 * fermi.f:1-434 (one process, jspins = 1).  eigen[k*ldn + i] eigenvalue i of
 * k-point k (ascending per k), weight[k].
 */
double orc_fermi(int ldn, int nk, int nv, double el, const double *weight,
                 const double *eigen, double tkb, double zfill);

/*
 * bande.f:1057-1477 sums.  mode as in bande.f:99-123 (1..4).
 * ncut[k] states of k enter the pressure sums (pressure modes).
 * Outputs efermi, earg1, earg2.
 */
void orc_bande_sums(int mode, int ldn, int nk, int nv, double el,
                    const double *weight, const double *esk,
                    const double *desk, const int *ncut, double tkb,
                    double zfill, double *efermi, double *earg1,
                    double *earg2);

/*
 * Whole bande.f k-loop (bande.f:816-1055) for one structure:
 * per k setham -> zhegv -> [sethamv -> perturb], then fermi and sums.
 * eig_out[k*nv+i] (may be NULL), desk_out likewise.  nthreads > 1 runs
 * k-points on OpenMP threads.  Returns 0, 3 (S not positive definite at some
 * k) or 5 (degeneracy > 9).
 */
int orc_bande(int mode, int nv, int natoms, const int *kkind, const int *ksk,
              const int *kbas, const double *posq, int npairs,
              const double *tt, const double *dlv, const int *jkind,
              const int *jatm, const double *hop, const double *over,
              const double *paros, const double *dhop, const double *dover,
              const double *dparos, int nk, const double *kpts,
              const double *weight, double zfill, double tkb, double eigcut,
              int nthreads, double *efermi, double *earg1, double *earg2,
              double *eig_out, double *desk_out);

#ifdef __cplusplus
}
#endif
#endif
