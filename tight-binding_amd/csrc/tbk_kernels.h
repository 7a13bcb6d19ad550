// tbk_kernels.h -- argument blocks and launchers of the device kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <type_traits>

#include <cstddef>
#include <vector>

namespace tbk {

// One structure of a plan (device pointers into the plan's concatenated
// set-up arrays).
struct StructDev {
    int natoms, nv, npairs;
    const int *row_atom, *row_orb;   // [nv] atom and orbital of each row
    const int *ksk, *nbas;           // [natoms] first row, orbital count
    const int *blk_ptr;              // [natoms*natoms+1] CSR of pairs by (i, j)
    const double *dvec;              // [3*npairs] dlv - posq_j + posq_i
    const double *eh, *es;           // [81*npairs] SK blocks e(a,b) of H, S
    const double *edh, *eds;         // [81*npairs] volume derivatives
    const double *paros, *dparos;    // [4*natoms] on-site terms and derivatives
    // non-collinear structures: [6*nv] per orbital the spin terms d, Re m,
    // Im m and their volume derivatives; the items have order 2 nv.
    // nullptr otherwise.
    const double *spin = nullptr;
};

// Build of H(k), S(k) (or dH, dS) for a batch of items (structure, k-point)
// at the plan order n >= nv: rows/columns nv..n-1 are padding, decoupled
// (H = pad * I, S = I; zero in the derivative).
struct BuildArgs {
    int n, deriv, maxpairs;
    size_t stride;                 // elements between item slabs
    double pad;
    const StructDev *sd;
    const int *item_s;             // [nb] structure of each item
    const double2 *ph;             // [nb][maxpairs] Bloch phases
    double2 *h, *s;
    double2 *dh = nullptr, *ds = nullptr;   // non-null: H, S, dH/dV, dS/dV in one pass (deriv unused)
    // 1: the strictly upper triangle is not written (for consumers that
    // read only the lower triangle and the diagonal, GEMM tri 9)
    int lower_only = 0;
    // 1: the strictly lower triangle of S is not written (the Cholesky
    // reads the upper triangle only)
    int s_upper_only = 0;
};

template <class R> struct cplx_t;
template <> struct cplx_t<double> { typedef double2 type; };
template <> struct cplx_t<float> { typedef float2 type; };

template <class R>
struct GemmArgsT {
    typedef typename cplx_t<R>::type C;
    int m, n, k;
    C alpha, beta;
    const C *a; int lda; size_t sa;
    const C *b; int ldb; size_t sb;
    C *c; int ldc; size_t sc;
    int tri;                       // 0 full; 1 A upper triangular; 2 A lower triangular;
                                   // 3 B upper triangular; 4 lower triangle of C only;
                                   // 5 lower triangle of C mirrored into the upper
                                   //   (Hermitian C; alpha real, beta 0); 6 = 5 with
                                   //   A lower triangular as in 2; 7 upper triangle of C only;
                                   //   8 = 5 with beta (C = alpha A B + beta C on the
                                   //   lower triangle, mirrored); 9 A (not transposed)
                                   //   Hermitian, read as tril(A, -1) + diag(A) / 2
    const int *ncol = nullptr;     // [nb] per item: column tiles at or beyond ncol[k]
                                   // are skipped (C left as it was); null: all
    // optional copy of every element written to C, in the other precision
    typedef typename cplx_t<typename std::conditional<sizeof(R) == 8, float, double>::type>::type CX;
    CX *cx = nullptr;
    int ldcx = 0;
    size_t scx = 0;
};
typedef GemmArgsT<double> GemmArgs;

template <class R>
struct HetrdArgsT {
    typedef typename cplx_t<R>::type C;
    int n;
    size_t stride;
    C *a, *vp, *wp, *t;            // vp: [nb][n*96] panel [V|W|V]; t: [nb][tstride]
    C *vt;                         // [nb][n*n]: V T^H of every panel (columns i0..)
    double *d, *e;                 // [nb][n]
    size_t tstride;                // ceil(n/32) panels of 32x32 WY factors
};
typedef HetrdArgsT<double> HetrdArgs;

// Two-stage tridiagonalisation: band width TB; the band is stored by
// columns, TBLD entries each (diagonals 0 .. 2 TB - 1, room for the fill of
// the bulge chasing).
constexpr int TB = 16;
constexpr int TBLD = 2 * TB;
constexpr int kMaxBandBlocks = 1280 / TB + 2;     // sweep blocks of the largest order

template <class R>
struct BandArgsT {                 // stage 1 (full -> band)
    typedef typename cplx_t<R>::type C;
    int n;
    size_t ss;
    C *a, *vt;                     // [nb][n*n]: matrix (both triangles); V T^H of the panels
    C *p;                          // [nb][n*96]: columns 0..TB-1 V, TB..2TB-1 Y/Z, 2TB..3TB-1 V, 3TB..4TB-1 V T
    size_t sp;
    C *t;                          // [nb][tstride]: TB x TB T factor of each panel
    size_t st;
    C *band;                       // [nb][n*TBLD]
    size_t sb;
};
template <class R>
struct BulgeArgsT {                // stage 2 (band -> tridiagonal)
    typedef typename cplx_t<R>::type C;
    int n;
    C *band;                       // [nb][n*TBLD] (overwritten)
    size_t sb;
    C *v2;                         // [nb][diamonds*TB*TB] reflectors; entry 0 of each is its tau
    size_t sv;
    double *d, *e;                 // [nb][n] real tridiagonal
    size_t sde;
    double2 *ph;                   // [nb][n] phases: eigenvector rows of the real tridiagonal times ph[r]
    size_t sph;
};
template <class R>
struct Q2ArgsT {                   // back-transform through the stage-2 reflectors
    typedef typename cplx_t<R>::type C;
    int n;
    C *y;                          // [nb][n*n] vectors, Y <- Q2 Y
    size_t sy;
    C *v2;                         // as BulgeArgsT
    size_t sv;
    C *t2;                         // [nb][diamonds*TB*TB] T factors
    size_t st2;
    const int *ncol = nullptr;     // [nb] columns to transform (null: all)
};

// Scratch of the multi-workgroup tridiagonalisation (per item): v, the
// column / w vector, the next column's delayed update (carry), tile
// partial sums ydot, yax [tmax][n], the panel dots (2*32) and tau.
template <class R>
struct WideArgsT {
    typedef typename cplx_t<R>::type C;
    int tmax;                      // ceil((n - 1) / 64)
    C *v, *col, *carry, *ydot, *yax, *udot, *scal;
};

struct SteinArgs {
    int n;
    double ortol_rel;              // group threshold relative to ||T||_1
    const int *nact;               // [nb] active order of each item (null: n)
    const double *d, *e;
    double *w, *v, *scratch;       // w [nb][n]; v [nb][n*n]; scratch [nb][5*n*stein_threads(n)]
    int *info;
    double abstol_rel = 0.0;       // > 0: bisection stops at abstol_rel * ||T||_1 (else dstebz's)
    const int *cut = nullptr;      // [nb] vectors j >= cut[k] are set to zero, not formed; null: all
};

struct PerturbArgs {
    int n, nb, k0;
    size_t stride;
    double eigcut;
    const int *nact;               // [nb] active order of each item (null: n)
    const double *w;
    // x = L_dH z, y = L_dS z with L = tril(M, -1) + diag(M) / 2 (GEMM tri 9):
    // z_a^H M z_b = z_a^H (L z_b) + conj(z_b^H (L z_a)); px, py their column dots
    const double2 *px, *py, *z, *x, *y;
    double *desk;
    int *ncut, *err;               // ncut [k0+nb], err [k0+nb] (2: > 9 degenerate)
    const int *cut = nullptr;      // [nb] derivatives only below cut[k] (0 above); null: all
    // an item of one active orbital takes a zero derivative, as the
    // reference's perturb computes it on the eigenvector its closed-form
    // 1 x 1 diag never formed (diag.f:320-349)
    bool one_orbital_zero = false;
};

struct FermiArgs {
    int nk, nv, ld, mode;
    double el, tkb, zfill;
    const double *esk, *desk, *w;
    const int *ncut;
    double *out;                   // efermi, earg1, earg2, esum, entropy, iters, smag
    int *cut = nullptr;            // [nk] occupation window per k (k_fermi_window)
    int k0 = 0;                    // first item of the group in the plan
    // > 0: items [0, khalf) are spin up and [khalf, nk) spin down of one
    // structure (collinear, el = 1); out[6] = up - down electrons
    int khalf = 0;
};

// GEMM tiles and their choice (tbk_tune.cpp)
constexpr int kGemmTiles = 4;      // 0 = 64x64, 1 = 96x96, 2 = 128x64, 3 = 32x32
struct GemmShape {
    int prec, opa, opb, m, n, k, tri, nb;   // prec 0 = complex FP64, 1 = complex FP32
    GemmArgsT<double> g64;                  // arguments of the first launch seen
    GemmArgsT<float> g32;
};
int gemm_tile(int prec, int opa, int opb, int m, int n, int k, int tri, int nb);
int gemm_tile_model(int m, int n, int nb);
void gemm_force_tile(int tile);            // -1: table / model
void gemm_store_tile(const GemmShape &s, int tile);
bool gemm_tuned(const GemmShape &s);
void gemm_record(bool on);                  // collect the shapes of launches
void gemm_note(const GemmShape &s);
std::vector<GemmShape> gemm_recorded();

// Non-collinear items: the spatial H, S (or dH, dS) just built in the
// leading nv block become the 2 nv spinor matrices of addspins.f:68-96.
void launch_spinor_expand(const BuildArgs &a, int maxnv, int nb, hipStream_t st);
void launch_phases(const StructDev *sd, const int *item_s, const double *kpts, int maxpairs,
                   double2 *ph, int nb, hipStream_t st);
void launch_build_hs(const BuildArgs &a, int maxnat, int nb, hipStream_t st);
void launch_potrf_panel(int n, double2 *s, size_t stride, int j0, int *info, int nb,
                        hipStream_t st);
void launch_zgemm(const GemmArgs &g, int opa, int opb, int nb, hipStream_t st);
void launch_cgemm(const GemmArgsT<float> &g, int opa, int opb, int nb, hipStream_t st);
void launch_gemm_tile(const GemmShape &s, int tile, hipStream_t st);   // a recorded shape, given tile
void launch_hetrd_panel(const HetrdArgs &h, int i0, int nb, hipStream_t st);
void launch_hetrd_finish(const HetrdArgs &h, int nb, hipStream_t st);
void launch_hetrd_panel(const HetrdArgsT<float> &h, int i0, int nb, hipStream_t st);
// multi-workgroup form: one column (three launches), and V T^H after a panel
void launch_hetrd_wide_column(const HetrdArgs &h, const WideArgsT<double> &w, int i0, int i, int nb,
                              hipStream_t st);
void launch_hetrd_wide_column(const HetrdArgsT<float> &h, const WideArgsT<float> &w, int i0, int i, int nb,
                              hipStream_t st);
void launch_hetrd_wide_vt(const HetrdArgs &h, int i0, int nb, hipStream_t st);
void launch_hetrd_wide_vt(const HetrdArgsT<float> &h, int i0, int nb, hipStream_t st);
void launch_hetrd_finish(const HetrdArgsT<float> &h, int nb, hipStream_t st);
// resident form (the whole reduction of an item in one workgroup, matrix in
// registers; reflectors, WY factors and V T^H as the panels leave them) for
// orders up to hetrd_resident_max_n
int hetrd_resident_max_n(bool fp32);
void launch_hetrd_resident(const HetrdArgsT<float> &h, int nb, hipStream_t st);
void launch_hetrd_resident(const HetrdArgs &h, int nb, hipStream_t st);
int stein_threads(int n);        // eigenvector threads per k-point
void launch_stein(const SteinArgs &s, int nb, hipStream_t st);          // eigenvalues and vectors
void launch_stebz(const SteinArgs &s, int nb, hipStream_t st);          // eigenvalues only
void launch_stein_vectors(const SteinArgs &s, int nb, hipStream_t st);  // vectors of known eigenvalues
// ph (may be null): row i of Y times ph[k][i] (phases of the two-stage reduction)
void launch_real_to_complex_t(int n, const double *vt, double2 *y, size_t stride, int nb,
                              hipStream_t st, const double2 *ph = nullptr);
// y64_row0 (may be null): also writes row 0 of Y in FP64 there
void launch_real_to_c32_t(int n, const double *vt, float2 *y, size_t stride, int nb, hipStream_t st,
                          double2 *y64_row0 = nullptr, const double2 *ph = nullptr);
// two-stage tridiagonalisation (panel c0 = panel * TB of stage 1)
int band_diamonds(int n);        // reflector groups of stage 2 per item
void launch_band_qr(const BandArgsT<double> &g, int panel, int nb, hipStream_t st);
void launch_band_qr(const BandArgsT<float> &g, int panel, int nb, hipStream_t st);
void launch_band_tail(const BandArgsT<double> &g, int c_end, int nb, hipStream_t st);
void launch_band_tail(const BandArgsT<float> &g, int c_end, int nb, hipStream_t st);
void launch_bulge(const BulgeArgsT<double> &g, int nb, hipStream_t st);
void launch_bulge(const BulgeArgsT<float> &g, int nb, hipStream_t st);
void launch_q2(const Q2ArgsT<double> &g, int nb, hipStream_t st);      // T factors, then Y <- Q2 Y
void launch_q2(const Q2ArgsT<float> &g, int nb, hipStream_t st);
void launch_rows_c32_to_c64(int n, int r0, int r1, const float2 *y32, double2 *y64, size_t stride, int nb,
                            hipStream_t st);
int oa_max_n();                  // largest N of the refinement kernels
// The refinement step of the mixed method on C (FP64) and the approximate
// eigenvectors Y (FP64 copy Y, FP32 copy Y32 of the FP32 back-transform):
// the caller forms X = C Y (FP64), then launch_oa_prepare (Rayleigh
// quotients, residuals R32 = X - Y diag(lambda), clusters, rotated
// clusters), K32 = Y32^H R32 (FP32), launch_oa_eigen (eigenvalues,
// ascending, into w), and for the vectors launch_oa_emat (E32, columns in
// the final order), D32 = Y32 E32 (FP32) and launch_oa_final (Yr, the
// refined vectors normalised).  Columns j >= mw[k] are not refined: they
// keep theta's eigenvalues and form no vector.  lam, rn, perm, perm2,
// cinfo, clist [nb][n], tau, ncl [nb] are scratch; R32 and D32 may share
// storage.
struct OaArgs {
    int n;
    size_t stride;
    const int *nact;               // [nb] active order (padding after); null: n
    const int *mw;                 // [nb] refined columns; null: n
    const double *theta;           // [nb][n] eigenvalues of the FP32 tridiagonal
    double2 *X, *Y;                // rotated in place over the clusters
    float2 *Y32, *R32, *K32, *E32, *D32;
    double2 *Yr;
    double *lam, *rn, *tau, *w;
    int *perm, *perm2, *cinfo, *clist, *ncl;
};
void launch_oa_prepare(const OaArgs &a, int nb, hipStream_t st);
void launch_oa_eigen(const OaArgs &a, int nb, hipStream_t st);
// columns j >= cut[k] are not formed (cut may be null)
void launch_oa_emat(const OaArgs &a, int nb, hipStream_t st, const int *cut = nullptr);
void launch_oa_final(const OaArgs &a, int nb, hipStream_t st, const int *cut = nullptr);
void launch_coldot(int n, const double2 *z, const double2 *x, const double2 *y, size_t stride,
                   double2 *px, double2 *py, int nb, hipStream_t st, const int *cut = nullptr);
void launch_perturb(const PerturbArgs &p, hipStream_t st);
void launch_fermi_sums(const FermiArgs &f, hipStream_t st);
// one workgroup per structure; ef_known: E_F already set by launch_fermi_window_group
void launch_fermi_sums_group(const FermiArgs *f_dev, int nstruct, hipStream_t st, int ef_known = 0);
// E_F of every structure and the cut of each k: the first state at or above
// E_F + 20 kT, moved up to the next spectral gap (no state above it enters
// the Fermi sums; no group of close eigenvalues straddles it).
void launch_fermi_window_group(const FermiArgs *f_dev, int nstruct, hipStream_t st);
// Spin directions of non-collinear structures (diag.f:458-523,
// bande.f:1484-1521): per state j of each item and atom a of its structure
// spev[item][j][a][0..2] = (-2 Re, 2 Im, |z_dn|^2 - |z_up|^2) of conj(z_up) z_dn
// summed over the atom's orbitals (states at or beyond cut[k] zero); then
// per structure the occupation-weighted sum over k and states into
// out[s][a][0..2] (one workgroup per Fermi group, E_F from its out[0]).
struct SpinArgs {
    int n, maxnat, k0;
    size_t stride;
    const StructDev *sd;
    const int *item_s, *nact, *cut;
    const double2 *z;
    double *spev;                  // [items][n][maxnat][3]
};
void launch_spin_expect(const SpinArgs &a, int nb, hipStream_t st);
void launch_spin_sum(const FermiArgs *f_dev, const StructDev *sd, int nstruct, const double *spev, int n,
                     int maxnat, double *out, hipStream_t st);
void launch_copy_int(const int *src, int *dst, int n, hipStream_t st);
// Lowdin projections: S = I on every item; T[i, j] *= sqrt(lam[k][i]);
// proj[k][e] = |Y[k][e]|^2 over the n x n slabs of nb items
void launch_set_identity(int n, double2 *s, size_t stride, int nb, hipStream_t st);
void launch_scale_rows_sqrt(int n, double2 *t, size_t stride, const double *lam, int nb, hipStream_t st);
void launch_abs2(int n, const double2 *y, size_t stride, double *proj, int nb, hipStream_t st);
void launch_copy16(const double2 *src, double2 *dst, size_t n, hipStream_t st);

}  // namespace tbk
