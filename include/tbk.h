/*
 * tbk.h -- C ABI of the MI355X-native NRL tight-binding k-point engine.
 *
 * The engine replaces the per-k-point hot path of the NRL program `static`
 * (boiled-greggs/Tight-Binding, NRLtb/src): the neighbour list and
 * k-independent Slater-Koster parameters feeding it, the per-k H(k)/S(k)
 * build, the generalized Hermitian eigensolve, the first-order pressure
 * derivatives, the Fermi search and the band-energy/pressure sums.
 *
 * Conventions (all entry points):
 *   - plain pointers and sizes; no allocation is visible to the caller;
 *   - lengths in bohr, energies in Ry, pressures in Ry/bohr^3, k-points in
 *     reciprocal-lattice units (the x,y,z of COMMON /cd1/, kptin.f:349-366);
 *   - lat[3*v+c]  = component c of primitive vector v   (plv(v,c))
 *     posn[3*a+c] = Cartesian coordinate c of atom a     (posn(a,c))
 *     posq[3*a+c] = phase origin of atom a, lattice units (posq(a,c))
 *     kkind[a]    = 1-based kind of atom a;
 *   - complex matrices are complex128 interleaved (re, im), column-major;
 *   - every function returns TBK_OK or an error code; none aborts.
 */
#ifndef TBK_H
#define TBK_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Error codes.  The reference reports the same conditions through iberr,
 * ioverr, ierr or STOP (bande.f:987-995, diag.f:429-455, search2.f:180-188,
 * perturb.f:106, fermi.f:165); here each is a return value. */
#define TBK_OK            0
#define TBK_ERR_ARG       1   /* invalid argument or size                 */
#define TBK_ERR_PAIRS     2   /* pair list exceeds maxpairs (search2)      */
#define TBK_ERR_OVERLAP   3   /* S(k) not positive definite at some k      */
#define TBK_ERR_CONVERGE  4   /* tridiagonal eigensolver did not converge  */
#define TBK_ERR_ZFILL     5   /* zfill <= 0 (fermi.f:165)                  */
#define TBK_ERR_DEGEN     6   /* degenerate group larger than 9 (perturb.f:106) */
#define TBK_ERR_HIP       10  /* HIP runtime error                         */
#define TBK_ERR_NOMEM     11  /* device allocation failed                  */
#define TBK_ERR_NODEVICE  12  /* no usable GPU                             */
#define TBK_END           (-1) /* tbk_skin_next: no more structures        */

/* Precision of a plan: the plain all-FP64 method or the mixed method. */
#define TBK_FP64          0   /* all-FP64 plain method                     */
#define TBK_MIXED         1   /* FP32 tridiagonalisation + back-transform,
                                 FP64 Cholesky, Fermi sums and refinement  */

/* bande modes supported (bande.f:99-123). */
#define TBK_MODE_E0_ET    1   /* earg1 = E(0) (Gillan), earg2 = E(T)       */
#define TBK_MODE_E0_P0    2   /* earg1 = E(0), earg2 = P(0)                */
#define TBK_MODE_ET       3   /* earg1 = E(T)                              */
#define TBK_MODE_ET_PT    4   /* earg1 = E(T), earg2 = P(T)                */

/* ---------------------------------------------------------------------
 * Host set-up (k-independent).  Bit-exact restatements.
 * --------------------------------------------------------------------- */

/* Neighbour list.  Replaces search2(npfound,dismin,ierr), search2.f:1
 * (pair order iat, jat >= iat, ix, iy, iz; cutoff test search2.f:168-175).
 * nd[3]: search half-widths (input1.f:488).  dnn, screenl: per interaction
 * type.  Outputs sized maxpairs: tt[3p+c] direction cosines, dlv[3p+c]
 * lattice translation, dist[p], screen[p], dscrn[p], jkind[2p+i],
 * jatm[2p+i] (1-based).  Returns TBK_ERR_PAIRS on overflow. */
int tbk_search2(int natoms, const double *lat, const double *posn,
                const int *kkind, int kinds, const int *nd,
                const double *dnn, const double *screenl, int maxpairs,
                int *npairs, double *tt, double *dlv, double *dist,
                double *screen, double *dscrn, int *jkind, int *jatm,
                double *dismin);

/* Density-dependent on-site terms and bond/overlap integrals.  Replaces
 * setpar(jspin,param,jsover,den,flipos,lnoncol,paros,spinos,hop,over,
 * lprint,ierr), setpar.f:1 (collinear, jspin = 1).  param: one spin channel
 * (input.f:327-359 order).  Outputs den[a*kinds+k], paros[4a+j],
 * hop[14p+j], over[14p+j]. */
int tbk_setpar(int natoms, const int *kkind, int kinds, const double *param,
               int nltype, int jsover, int npairs, const double *dist,
               const double *screen, const int *jkind, const int *jatm,
               double *den, double *paros, double *hop, double *over);

/* Volume derivatives.  Replaces setparv(...), setparv.f:1. */
int tbk_setparv(int natoms, const int *kkind, int kinds, const double *param,
                int nltype, int jsover, double thisvol, int npairs,
                const double *dist, const double *screen, const double *dscrn,
                const int *jkind, const int *jatm, const double *den,
                double *dparos, double *dhop, double *dover);

/* Spin channel jspin (1 or 2) of a collinear spin-polarised set.  Replaces
 * setpar(jspin,param,jsover,den,flipos,.false.,...), setpar.f:1, and
 * setparv(jspin,...), setparv.f:1, with jspins = 2: param2 [2*nparam] holds
 * param(:,1) then param(:,2); bond and overlap integrals come from channel
 * jspin, the densities and on-site terms of an atom with flipos[a] != 0
 * from the other channel (setpar.f:500-536, 795-867); flipos may be NULL.
 * Outputs as tbk_setpar / tbk_setparv. */
int tbk_setpar_spin(int natoms, const int *kkind, int kinds,
                    const double *param2, int nparam, int jspin,
                    const int *flipos, int nltype, int jsover, int npairs,
                    const double *dist, const double *screen, const int *jkind,
                    const int *jatm, double *den, double *paros, double *hop,
                    double *over);
int tbk_setparv_spin(int natoms, const int *kkind, int kinds,
                     const double *param2, int nparam, int jspin,
                     const int *flipos, int nltype, int jsover, double thisvol,
                     int npairs, const double *dist, const double *screen,
                     const double *dscrn, const int *jkind, const int *jatm,
                     const double *den, double *dparos, double *dhop,
                     double *dover);

/* Non-collinear set-up (lnoncol): setpar(1,param,jsover,den,flipos,.true.,
 * paros,spinos,hop,over,...) and setparv(1,...,.true.,dparos,dspnos,dhop,
 * dover,...), setpar.f:1, setparv.f:1, for a two-channel param2
 * [2*nparam].  den2 [2*natoms*kinds] (spin up, then down), paros/spinos and
 * dparos/dspnos [4*natoms] (average and half difference of the channels'
 * on-site terms), hop/over/dhop/dover [14*npairs] from channel 1. */
int tbk_setpar_noncol(int natoms, const int *kkind, int kinds,
                      const double *param2, int nparam, int nltype, int jsover,
                      double thisvol, int npairs, const double *dist,
                      const double *screen, const double *dscrn,
                      const int *jkind, const int *jatm, double *den2,
                      double *paros, double *spinos, double *dparos,
                      double *dspnos, double *hop, double *over, double *dhop,
                      double *dover);

/* ---------------------------------------------------------------------
 * k-point meshes (host).  Bit-exact restatements of the reference's
 * kptin/rdspgp/kptgen, whose k-points and weights the plans take.
 * Arrays are row-major: lat/avec/bvec[3*v + c] = component c of vector v
 * (bvec with the factor 2 pi), bij[3*i + j] = b_i . b_j, operations
 * rot[9*k + 3*i + j] = R_k(i, j), k-points q[3*n + c] in reciprocal-lattice
 * coordinates (the units bande and tbk_plan_create take).
 * --------------------------------------------------------------------- */

/* Reciprocal lattice, metric and cell volume.  Replaces the NEWSYM block of
 * kptin (kptin.f:94-155, AVEC, BVEC, BIJ, WSVOL of COMMON /LATTYP/). */
int tbk_reciprocal_lattice(const double *lat, double *avec, double *bvec,
                           double *bij, double *wsvol);

/* Space-group operations from rotations as read (direct-lattice
 * coordinates if ilat, else Cartesian) and non-primitive translations tnp
 * [3*nop] (direct-lattice coordinates), with an origin shift (may be NULL).
 * Outputs rspc (Cartesian), rspct (reciprocal-lattice coordinates), tnpspc
 * [3*nop].  Replaces the transforms of rdspgp (rdspgp.f:106-289; lattic 4
 * snaps Cartesian entries to +-sqrt(3)/2). */
int tbk_space_group(int ilat, int lattic, const double *avec,
                    const double *bvec, int nop, const double *rot,
                    const double *tnp, const double *shift, double *rspc,
                    double *rspct, double *tnpspc);

/* The same from a space-group file in rdspgp's format (A10,I5,5X,3F10.5
 * header; 18F4.0 or 3D20.10 rotations; 9F8.5 translations).  Replaces
 * rdspgp(iprt,ilat,io), rdspgp.f:1 (TBK_ERR_ARG for an unreadable file or
 * nop > maxop, where the reference stops).  name: 11 chars or NULL. */
int tbk_space_group_read(const char *path, int ilat, int lattic,
                         const double *avec, const double *bvec, int maxop,
                         int *nop, double *rspc, double *rspct,
                         double *tnpspc, char *name);

/* Irreducible k-points and weights of the kdiv mesh (lspec: special-point
 * mesh), folded by the nop operations rspct and time reversal.  Replaces
 * kptgen(nw), kptgen.f:1 (lattic 2 = bcc through the simple-cubic cell,
 * 4 = hexagonal, 7-9 = body-centred tetragonal).  maxfull bounds the full
 * mesh (the reference's nkbzd = 2 nkd; <= 0: unbounded), maxk the
 * irreducible points; TBK_ERR_ARG where the reference stops. */
int tbk_kptgen(int lattic, const double *avec, const double *bvec,
               const double *bij, double wsvol, int nop, const double *rspct,
               const int *kdiv, int lspec, int maxfull, int maxk, int *nk,
               double *q, double *w);

/* kptin's weights of a generated mesh (zero-weight information points
 * appended by the caller): renormalisation, the kpst..kpend range and the
 * final scaling, in place (kptin.f:312-351).  *nk_out points remain. */
int tbk_kmesh_weights(int nk, double *q, double *w, int kpst, int kpend,
                      int *nk_out);

/* kptin's explicit list: raw[4*n] = k (units of 1/sfk), weight; q = k/sfk,
 * w normalised to sum 1 (kptin.f:214, 266-273, 348-351). */
int tbk_klist(int nk, const double *raw, double sfk, double *q, double *w);

/* kptin's line mode (nkv = 999): npath segments of kd[p] points from
 * ends[6p .. 6p+2] to ends[6p+3 .. 6p+5] (kptin.f:223-251). */
int tbk_kpath(int npath, const int *kd, const double *ends, int maxk,
              int *nk, double *q, double *w);

/* ---------------------------------------------------------------------
 * The reference program's input and output files, read and written
 * natively (host).  Reads follow its Fortran I/O statement by statement;
 * lattices and positions are bit-identical to those it passes to bande.
 * --------------------------------------------------------------------- */
#define TBK_MAXKINDS 8

/* One NRL parameter file (input.f:181-359). */
typedef struct tbk_param_set {
    int jsover;           /* first header character: 'O' orthogonal, else 1  */
    int jspins;           /* second: 'M' two spin channels, else 1           */
    int nltype;           /* I5 style field: 0..3 or 90000 (others read as 0)  */
    int kinds;
    int nparam;           /* values the file holds per channel (the rest are
                             zero); the program needs kinds*(29+68*kinds)    */
    int kbas[TBK_MAXKINDS];
    double valence[TBK_MAXKINDS];                        /* s + p + d      */
    double dnn[TBK_MAXKINDS * (TBK_MAXKINDS + 1) / 2];    /* cutoff R_c     */
    double screenl[TBK_MAXKINDS * (TBK_MAXKINDS + 1) / 2];
    char title[80];
} tbk_param_set;

/* Read an NRL parameter file.  Replaces input(nparam,jspins) (input.f:139,
 * its reads of unit 32, input.f:181-359).  param [cap]: jspins channels of
 * kinds*(29+68*kinds) values, param(:,1) then param(:,2).  TBK_ERR_ARG for
 * an unreadable file, kinds > TBK_MAXKINDS or cap too small. */
int tbk_read_params(const char *path, tbk_param_set *ps, int cap,
                    double *param);

/* A SKIN control file, read structure by structure. */
typedef struct tbk_skin tbk_skin;

/* What the reference program reads of one structure (static.f:938-950,
 * setlat.f, input1.f, kptin.f). */
typedef struct tbk_skin_structure {
    char title[72];       /* (a70)                                        */
    char label[24];       /* (a20) SKENG label, 20 characters             */
    double zcharge;       /* electrons above the kinds' valence           */
    double zfill;         /* electrons (input1.f:735, 820-829)            */
    double thisvol;       /* cell volume (setlat.f:997-1000)              */
    double lat[9];        /* lat[3*v+c] = plv(v,c), bohr                  */
    int latnum, strain;   /* setlat's lattice and strain options          */
    int natoms, nd[3];
    int lphonon, spinflip, lnoncol, ltorque, lspdir;
    int nk;               /* k-points (the previous ones for NEWSYM=F)    */
    int inpstat;          /* input1's error code, 1000 + setlat's         */
} tbk_skin_structure;

/* Open a SKIN file and read its header: the mode (static.f:706,
 * modeparse.f), kT and eigcut (static.f:818) and the parameter file name
 * (input.f:139), resolved against the SKIN file's directory. */
int tbk_skin_open(const char *path, tbk_skin **out, int *mode, double *tkb,
                  double *eigcut, char *param_path, int param_path_cap);

/* The next structure: TBK_OK, TBK_END at the end of the file (where the
 * program ends normally), or an error where the program stops.  ps gives the
 * kinds' valence.  Arrays of capacity maxatoms / maxk: kkind, posn
 * (Cartesian, static.f:1088-1103), posq (phase origins, lattice units),
 * flipos and spinat [3*natoms] (may be NULL), kpts [3*maxk], weights
 * [maxk] (kptin.f, bit-identical). */
int tbk_skin_next(tbk_skin *s, const tbk_param_set *ps, int maxatoms,
                  int maxk, tbk_skin_structure *st, int *kkind, double *posn,
                  double *posq, int *flipos, double *spinat, double *kpts,
                  double *weights);
int tbk_skin_close(tbk_skin *s);

/* One SKENG line (static.f:1238-1254): (a20,f12.6,3f15.9) in modes 1, 2 and
 * 4, (a20,f12.6,2f15.9) in mode 3, with f10.6 of the moment appended when
 * with_smag (collinear jspins = 2).  buf: at least 96 characters. */
int tbk_skeng_line(const char *label, double thisvol, double efermi,
                   double earg1, double earg2, double smag, int mode,
                   int with_smag, char *buf, int cap);

/* The reference program's loop for this path (static.f:925-1255): every
 * structure of dir/SKIN through the device plans (modes 1-4; jspins 1 and 2,
 * collinear and non-collinear), its line written to dir/SKENG as each one
 * finishes.  Stops at the first structure the reference would stop on, or
 * the device cannot run (torques and spin directions out), with the
 * lines of the structures before it written.  *nstruct: structures done. */
int tbk_static_run(const char *dir, int device, int precision, int *nstruct);

/* ---------------------------------------------------------------------
 * Device plans: one structure, its k-mesh, device-resident H/S/Z state.
 * --------------------------------------------------------------------- */
typedef struct tbk_plan tbk_plan;
typedef struct tbk_multi tbk_multi;
typedef struct tbk_server tbk_server;

/* Structure description passed to tbk_plan_create. */
typedef struct tbk_structure_desc {
    int natoms;
    const double *lat;        /* [9]            */
    const double *posn;       /* [3*natoms]     */
    const double *posq;       /* [3*natoms]     */
    const int *kkind;         /* [natoms]       */
    int nd[3];
    int kinds;
    const int *kbas;          /* [kinds]        */
    const double *valence;    /* [kinds]        */
    const double *dnn;        /* [kinds*(kinds+1)/2] */
    const double *screenl;    /* [kinds*(kinds+1)/2] */
    const double *param;      /* [nparam]       */
    int nparam;
    int nltype;
    int jsover;
    double thisvol;           /* cell volume (setlat.f:997)            */
    double zfill;             /* electrons (input1.f:733, 834)         */
} tbk_structure_desc;

/* Build the pair list and SK tables on the host, upload them, and size the
 * device state for nk k-points solved `batch` at a time (batch <= 0: all).
 * precision: TBK_FP64 or TBK_MIXED.  device: HIP device ordinal. */
int tbk_plan_create(int device, const tbk_structure_desc *st, int nk,
                    const double *kpts, const double *weights, int batch,
                    int precision, tbk_plan **out);

/* A plan over several structures of uneven orbital count, solved together
 * in the same batched launches (no reference counterpart: the reference
 * runs one structure per call of bande, static.f:1218-1222).  Structure i
 * has nk[i] k-points; kpts [3*sum(nk)] and weights [sum(nk)] hold them
 * structure after structure.  Items are padded to the largest orbital
 * count N with decoupled rows (H = 1e4 Ry * I, S = I), which leaves each
 * structure's eigenpairs, E_F and sums those of the structure alone.
 * tbk_plan_create is the case nstruct = 1 (no padding). */
int tbk_plan_create_group(int device, int nstruct,
                          const tbk_structure_desc *st, const int *nk,
                          const double *kpts, const double *weights,
                          int batch, int precision, tbk_plan **out);

/* A collinear spin-polarised structure: bande with jspins = 2 and lnoncol
 * false (bande.f:816-889, 1069-1081, 1342-1402).  st->param is the spin-up
 * channel (param(:,1), input.f:327-359), param_dn [st->nparam] the
 * spin-down channel (param(:,2)); flipos [natoms] (may be NULL: none)
 * marks atoms whose densities and on-site terms come from the other
 * channel (setpar.f:500-536, 795-867; input1.f:189-195).  The plan holds two
 * structures, 0 = spin up and 1 = spin down, each with the nk k-points and
 * weights given, solved in the same launches; tbk_plan_run returns one E_F
 * over both channels with one electron per state and the sums over both,
 * tbk_plan_result gives the same for s = 0 and 1, and
 * tbk_plan_structure_eigenvalues the channels' eigenvalues. */
int tbk_plan_create_spin(int device, const tbk_structure_desc *st,
                         const double *param_dn, const int *flipos, int nk,
                         const double *kpts, const double *weights, int batch,
                         int precision, tbk_plan **out);

/* A non-collinear structure: bande with jspins = 2 and lnoncol true, no
 * torques or spin directions out (bande.f:713-720, 859-905, 945-1036).
 * st->param is channel 1 (its bond and overlap integrals are used),
 * param_dn [st->nparam] channel 2; the on-site terms are the channels'
 * average and the spin terms half their difference (setpar.f:466-485,
 * 744-795), turned along spinat [3*natoms], each atom's unit spin direction
 * (spinham.f:110-150).  The items are the 2 nv spinor matrices of
 * addspins.f:68-96 (spin up rows first), one electron per state; results as
 * tbk_plan_run, tbk_plan_nv gives 2 nv. */
int tbk_plan_create_noncol(int device, const tbk_structure_desc *st,
                           const double *param_dn, const double *spinat,
                           int nk, const double *kpts, const double *weights,
                           int batch, int precision, tbk_plan **out);

/* Magnetic moment of the last run of a spin-polarised plan, spin-up minus
 * spin-down electrons per cell (the smag of bande.f:1388-1402); 0 for other
 * plans. */
int tbk_plan_magnetization(const tbk_plan *plan, double *smag);

/* Spin directions of a non-collinear plan (replaces the lspdir output of
 * bande.f:1484-1561 with diag.f:458-523): on = 1 makes the following runs
 * form every state's spin on each atom, (-2 Re, 2 Im, |z_dn|^2 - |z_up|^2)
 * of conj(z_up) z_dn summed over the atom's orbitals, and its
 * occupation-weighted sum over the k-points (states below E_F + 20 kT,
 * Fermi-Dirac weights within 20 kT); eigenvectors are formed in every mode.
 * TBK_ERR_ARG for a plan that is not non-collinear.  Memory: 24 N maxnat
 * bytes per item. */
int tbk_plan_set_spin_directions(tbk_plan *plan, int on);
/* After such a run: spdir[3*i + l], component l of atom i's spin vector,
 * of structure s (natoms of it). */
int tbk_plan_spin_directions(tbk_plan *plan, int s, double *spdir);

/* Run the whole bande k-loop + Fermi + sums on the device.  Replaces
 * bande(mode,...,efermi,earg1,earg2,...,iberr), bande.f:1, for the
 * collinear case (spin-polarised through tbk_plan_create_spin).  Modes 1-4,
 * and the energy-only modes whose sums are mode 3's and which write QLMT
 * (6, 9, 10: with 6 and 10 the run also forms the Lowdin projections,
 * tbk_plan_projections) or QAPW (8) files; mode 5 is refused (the
 * reference's program aborts on it).  Blocks until results are on the host. */
int tbk_plan_run(tbk_plan *plan, int mode, double tkb, double eigcut,
                 double *efermi, double *earg1, double *earg2);

/* Same, enqueued on the plan's stream without the final synchronisation;
 * tbk_plan_wait collects the results. */
int tbk_plan_launch(tbk_plan *plan, int mode, double tkb, double eigcut);
int tbk_plan_wait(tbk_plan *plan, double *efermi, double *earg1,
                  double *earg2);

/* TBK_OK when a launched run has finished (the plan's stream is idle), 1
 * while it still runs. */
int tbk_plan_query(const tbk_plan *plan);

/* Results of structure s after tbk_plan_wait: E_F, earg1, earg2 as in
 * tbk_plan_run; returns the structure's own status (TBK_ERR_OVERLAP,
 * TBK_ERR_DEGEN) so one failing structure does not hide the others.
 * tbk_plan_wait returns the first failing status and structure 0's
 * results. */
int tbk_plan_result(const tbk_plan *plan, int s, double *efermi,
                    double *earg1, double *earg2);

/* Results of the last run: eigenvalues esk[k*nv+i] (ascending per k) and
 * first-order derivatives desk[k*nv+i] (pressure modes, else stale; zero
 * above the occupation window when it was used), host copies, of structure
 * s (tbk_plan_eigenvalues: s = 0). */
int tbk_plan_eigenvalues(tbk_plan *plan, double *esk, double *desk);
int tbk_plan_structure_eigenvalues(tbk_plan *plan, int s, double *esk,
                                   double *desk);

/* Eigenvectors of item k (k-points numbered across structures) of the
 * last run when the plan holds it (last batch), complex128 column-major
 * nv x nv of the item's structure, Z^H S Z = I.  After a run that used the
 * occupation window (tbk_plan_set_window) the columns of states at or above
 * the window are zero (all columns after an energy-only run). */
int tbk_plan_eigenvectors(tbk_plan *plan, int k, double *z);

/* Lowdin projections of item k's states (diag.f:526-758) after a run in
 * mode 6 or 10 (QLMT): the states with e <= eigcut (*ntot of them, the
 * reference's ntot, diag.f:635-641), their eigenvalues e[j] and weights
 * proj[j*nv + i] = |(S^{1/2} z_j)_i|^2 on the item's nv orbitals (atoms in
 * order, kbas orbitals each; for a non-collinear item the spin-down block
 * follows the spin-up one).  S^{1/2} is formed on the device from S's own
 * eigendecomposition; sum_i proj = 1.  e [nv], proj [nv*nv]. */
int tbk_plan_projections(tbk_plan *plan, int k, double eigcut, int *ntot,
                         double *e, double *proj);

/* Occupation window (no reference counterpart; on by default).  The Fermi
 * sums skip every state at or above E_F + 20 kT (bande.f:1170-1321); when a
 * pressure run (modes 2, 4) holds all k-points in one batch, E_F is found as
 * soon as the eigenvalues are final and eigenvectors and derivatives are
 * formed only below that level (moved up to the next spectral gap), and
 * energy-only runs (modes 1, 3) form no eigenvector.  E_F, energies,
 * pressures and eigenvalues are the same as without it; on = 0 forms every
 * eigenvector. */
int tbk_plan_set_window(tbk_plan *plan, int on);
/* Build of the pressure modes (no reference counterpart; the derivative
 * matrices of sethamv.f): on = 0 (default) builds dH/dV, dS/dV after the
 * eigensolve into its workspace; on = 1 builds H, S, dH/dV and dS/dV in one
 * pass over the phases and Slater-Koster blocks, into two more matrices per
 * k-point allocated on the first pressure run (32 N^2 bytes per item).
 * Results are bit-identical; at order 288 the one pass is the slower. */
int tbk_plan_set_fused_build(tbk_plan *plan, int on);
/* Tridiagonalisation of the plan's eigensolves (no reference counterpart:
 * both replace ZHPTRD, lapack.f:3218, as diag.f:396-412 calls it):
 * 1 one-stage blocked Householder panels; 2 two-stage (reduction to a band
 * of width 16 by panel QR and GEMMs, then bulge chasing to tridiagonal);
 * 3 resident (unblocked Householder steps with the whole matrix held on
 * chip, one workgroup per k-point: orders up to 288 in the mixed method,
 * 160 in FP64); 0 the default (resident where the order allows it, else
 * two-stage from order 512, else one-stage).  Results agree to rounding
 * (FP64) or to the mixed method's stated errors.  Two-stage needs order
 * >= 40, resident the orders above (TBK_ERR_ARG otherwise). */
int tbk_plan_set_tridiag(tbk_plan *plan, int stages);

/* H(k), S(k) of k-point k as built on the device (full Hermitian,
 * column-major complex128), for parity checks of the build.  deriv = 1
 * returns dH/dV, dS/dV (sethamv.f).  The build is the one the plan's
 * pressure runs use (the four-matrix pass unless tbk_plan_set_fused_build
 * turned it off). */
int tbk_plan_hs(tbk_plan *plan, int k, int deriv, double *h, double *s);

int tbk_plan_nv(const tbk_plan *plan);
int tbk_plan_npairs(const tbk_plan *plan);
int tbk_plan_destroy(tbk_plan *plan);

/* Stage timing (ms per stage, HIP events on the plan's stream) summed over
 * the runs since timing was last enabled: names[i] and ms[i] for i < *n
 * (at most cap entries). */
int tbk_plan_stage_times(const tbk_plan *plan, int cap, const char **names,
                         double *ms, int *n);

/* Enable (on = 1) or disable per-stage HIP event timing of later runs;
 * either call clears the accumulated times. */
int tbk_plan_set_timing(tbk_plan *plan, int on);

/* Capture the launch sequence of a run into a hipGraph on first use and
 * replay it while (mode, tkb, eigcut) stay the same (on = 1); stage timing
 * and the autotuner run uncaptured.  Off by default. */
int tbk_plan_set_graph(tbk_plan *plan, int on);

/* Padded order N, number of structures, device bytes held by the plan. */
int tbk_plan_order(const tbk_plan *plan);
int tbk_plan_nstruct(const tbk_plan *plan);
long long tbk_plan_device_bytes(const tbk_plan *plan);

/* GEMM tile autotuner (no reference counterpart: the reference calls
 * unblocked LAPACK).  Runs one mode-4 pass of the plan to collect its GEMM
 * shapes, times every tile for each shape not tuned yet (reps launches
 * each) and keeps the fastest for the rest of the process.  The plan's
 * results are undefined until its next run.  *ntuned (may be NULL) receives
 * the number of shapes tuned. */
int tbk_plan_autotune(tbk_plan *plan, int reps, int *ntuned);
/* Tile used for a GEMM shape: 0 = 64x64, 1 = 96x96, 2 = 128x64, 3 = 32x32. */
int tbk_gemm_tile(int precision, int opa, int opb, int m, int n, int k, int tri, int nb);

/* ---------------------------------------------------------------------
 * Many structures of uneven orbital count (no reference counterpart: the
 * reference solves one structure per bande call).  Structures are grouped
 * by padded order (orbital count rounded up to `granularity`), each group
 * is one group plan, and the groups run concurrently on their own streams.
 * max_batch_bytes bounds the workspace of each group (items per launch);
 * <= 0: all items of a group in one launch.
 * --------------------------------------------------------------------- */
int tbk_multi_create(int device, int nstruct, const tbk_structure_desc *st,
                     const int *nk, const double *kpts,
                     const double *weights, int precision, int granularity,
                     long long max_batch_bytes, tbk_multi **out);
int tbk_multi_launch(tbk_multi *m, int mode, double tkb, double eigcut);
/* Waits for every group; returns the first failing structure status. */
int tbk_multi_wait(tbk_multi *m);
int tbk_multi_run(tbk_multi *m, int mode, double tkb, double eigcut);
int tbk_multi_result(const tbk_multi *m, int s, double *efermi,
                     double *earg1, double *earg2);
int tbk_multi_eigenvalues(tbk_multi *m, int s, double *esk, double *desk);
int tbk_multi_ngroups(const tbk_multi *m);
int tbk_multi_group_order(const tbk_multi *m, int g);
int tbk_multi_set_graph(tbk_multi *m, int on);
int tbk_multi_autotune(tbk_multi *m, int reps);
long long tbk_multi_device_bytes(const tbk_multi *m);
int tbk_multi_destroy(tbk_multi *m);

/* ---------------------------------------------------------------------
 * Request loop with continuous batching (no reference counterpart).  Host
 * threads build each queued request's set-up (neighbour list, SK blocks) as
 * it arrives; a worker thread admits prepared requests while their device
 * footprint fits in budget_bytes and one of max_inflight streams is free;
 * requests admitted together are grouped by padded order (granularity) and
 * (mode, tkb, eigcut) into group plans, one per stream, overlapping with
 * the groups already on the device.  Results are each structure's own (the
 * padding is decoupled).
 * --------------------------------------------------------------------- */
int tbk_server_create(int device, int precision, long long budget_bytes,
                      int max_inflight, int granularity, tbk_server **out);
/* Queue one structure (copied) with its k-mesh; *id names the request. */
int tbk_server_submit(tbk_server *s, const tbk_structure_desc *st, int nk,
                      const double *kpts, const double *weights, int mode,
                      double tkb, double eigcut, long long *id);
/* Wait up to timeout_ms (< 0: no limit) for request id: returns TBK_OK with
 * its status and results, or 1 if it is still queued or running.  Latency
 * is submit -> result, queue time submit -> admission, in ms. */
int tbk_server_result(tbk_server *s, long long id, int timeout_ms,
                      int *status, double *efermi, double *earg1,
                      double *earg2, double *latency_ms, double *queue_ms);
int tbk_server_release(tbk_server *s, long long id);
/* completed, k-points done, groups launched, largest group, pending,
 * running groups, in-flight bytes, last error. */
int tbk_server_stats(tbk_server *s, long long *stats8);
/* t6: device busy ms (union over finished groups of launch -> completion on
 * the device), work done (sum of nv^3 nk over completed requests), ms from
 * the first submission to the last completion, host set-up ms and plan
 * creation + launch ms (summed), set-up threads. */
int tbk_server_timing(tbk_server *s, double *t6);
/* Finishes every queued request, then stops the worker. */
int tbk_server_destroy(tbk_server *s);

/* ---------------------------------------------------------------------
 * Single-call drop-ins on host arrays.
 * --------------------------------------------------------------------- */

/* Generalized Hermitian eigensolve of one pair of full column-major
 * complex matrices on the device.  Replaces zhpgv(1,'V','U',...) as called
 * from diag.f:396-412 (diag_ for the collinear case).  z may be NULL.
 * Returns TBK_ERR_OVERLAP when S is not positive definite (diag.f:429). */
int tbk_zhegv(int device, int n, const double *h, const double *s,
              double *w, double *z);

/* tbk_zhegv without any allocation.  No counterpart in the reference: for
 * callers that cannot allocate device memory (tight serving loops, code
 * that forbids hipMalloc).  The same kernels and results as tbk_zhegv; the
 * device workspace is a fixed array the library reserves on each GPU when
 * it is loaded there (about 39 MB, held until the process ends), so
 * n <= 512 (TBK_ERR_ARG beyond).  Runs on the null stream and returns when
 * the results are on the host.  Calls on one GPU share the array: this
 * process's threads are serialised per GPU. */
int tbk_zhegv_noalloc(int device, int n, const double *h, const double *s,
                      double *w, double *z);

/* Fermi level.  Replaces fermi(ntbd,nkptd,nkloc,istart,ifin,nv,jspins,el,
 * weight,eigen,tkb,zfill), fermi.f:1 (one process, jspins = 1).
 * eigen[k*ldn+i].  Evaluated on the device. */
/* First-order derivatives of one k's eigenvalues on the device.  Replaces
 * perturb(mxd,n,dhmat,dsmat,z,eig,eigcut,ntot,deig,work), perturb.f:1-2:
 * dh, ds full n x n (complex, column-major) dH/dV, dS/dV; z the S-orthonormal
 * eigenvectors (column j = state j); eig ascending.  Groups of eigenvalues
 * within 1e-8 are diagonalised together (perturb.f:97-221); *ntot = states
 * at or below eigcut, deig[0..ntot).  TBK_ERR_DEGEN where a group exceeds 9
 * (the reference stops, perturb.f:106). */
int tbk_perturb(int device, int n, const double *dh, const double *ds,
                const double *z, const double *eig, double eigcut, int *ntot,
                double *deig);
int tbk_fermi(int device, int ldn, int nk, int nv, double el,
              const double *weight, const double *eigen, double tkb,
              double zfill, double *efermi);

const char *tbk_error_string(int code);

/* ---------------------------------------------------------------------
 * Diagnostics (no counterpart in the reference): stage-by-stage outputs of
 * the device eigensolver for one matrix pair, and one device GEMM.  Any
 * output pointer may be NULL.  uinv, c, refl, z: n x n complex128; vt:
 * n x n real (row-major, vt[i*n+j] = component i of tridiagonal
 * eigenvector j); d, e, w: n reals.
 * --------------------------------------------------------------------- */
int tbk_debug_stages(int device, int n, const double *h, const double *s,
                     double *uinv, double *c, double *refl, double *d,
                     double *e, double *w, double *vt, double *z);
/* One launch of a plain copy kernel over `bytes` bytes (a multiple of 16):
 * the calibration point for hardware traffic counters. */
int tbk_debug_copy(int device, long long bytes);

/* Average time (ms) of reps launches of the batched GEMM over nb items of
 * shape m x n x k (zero operands): the microbenchmark of the tile
 * autotuner.  tile -1: the tuned / modelled choice. */
int tbk_debug_gemm_time(int device, int precision, int tile, int m, int n,
                        int k, int opa, int opb, int tri, int nb, int reps,
                        double *ms);

/* C = op(A) op(B) of column-major complex matrices (complex128, or
 * complex64 when precision is TBK_MIXED) with the library's batched GEMM;
 * tile -1 = the tuned / modelled choice, else 0..3 (see tbk_gemm_tile).
 * tri: 0 full; 1 op(A) upper triangular; 2 op(A) lower triangular; 3 op(B)
 * upper triangular; 4 only the tiles touching the lower triangle of C;
 * 5 (m == n) the lower triangle of the product and its conjugate transpose
 * above the diagonal; 6 = 5 with op(A) lower triangular; 7 only the tiles
 * touching the upper triangle of C; 9 (opa = 0, k == m) A Hermitian, read
 * as tril(A, -1) + diag(A) / 2 (its upper triangle is not read). */
int tbk_debug_zgemm(int device, int precision, int tile, int m, int n, int k,
                    int opa, int opb, int tri, const void *a, int lda,
                    const void *b, int ldb, void *c, int ldc);

#ifdef __cplusplus
}
#endif
#endif
