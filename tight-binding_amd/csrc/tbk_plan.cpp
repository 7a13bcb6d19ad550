// tbk_plan.cpp -- device plans: host set-up upload, the per-batch k-point
// pipeline, Fermi + sums, and the C ABI around them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "tbk.h"
#include "tbk_internal.h"
#include "tbk_kernels.h"

namespace {

constexpr int kNB = 32;        // panel width of the factorizations
constexpr int kMaxN = 1280;    // LDS-resident vectors of k_hetrd / k_stein

#define HIP_TRY(x)                                   \
    do {                                             \
        hipError_t e_ = (x);                         \
        if (e_ != hipSuccess) return TBK_ERR_HIP;    \
    } while (0)

template <class T>
int dalloc(T **p, size_t n)
{
    *p = nullptr;
    if (n == 0) return TBK_OK;
    if (hipMalloc((void **)p, n * sizeof(T)) != hipSuccess) { *p = nullptr; return TBK_ERR_NOMEM; }
    return TBK_OK;
}
template <class T>
void dfree(T *&p)
{
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}

// Device buffers of one batch of b k-points at order n.
struct Workspace {
    int n = 0, b = 0;
    double2 *A = nullptr, *S = nullptr, *Y = nullptr, *Z = nullptr, *X = nullptr;
    double2 *Vp = nullptr, *Wp = nullptr, *Tf = nullptr, *px = nullptr, *py = nullptr;
    double2 *VT = nullptr;
    double *d = nullptr, *e = nullptr, *Vt = nullptr, *scratch = nullptr, *w = nullptr;
    int *info = nullptr, *sinfo = nullptr;
    // mixed precision: FP32 copies for the reduction and back-transform
    float2 *A32 = nullptr, *P32 = nullptr, *T32 = nullptr, *VT32 = nullptr, *Y32 = nullptr,
           *X32 = nullptr;
    double *lam = nullptr, *delta = nullptr;
    int *perm = nullptr;
    size_t stride() const { return (size_t)n * n; }
    // WY factors: one kNB x kNB block per panel, ceil(n / kNB) panels
    size_t tstride() const { return (size_t)((n + kNB - 1) / kNB) * kNB * kNB; }
    int alloc(int n_, int b_)
    {
        n = n_; b = b_;
        const size_t nn = (size_t)n * n * b, np = (size_t)n * kNB * b;
        int rc = TBK_OK;
        rc |= dalloc(&A, nn); rc |= dalloc(&S, nn); rc |= dalloc(&Y, nn);
        rc |= dalloc(&Z, nn); rc |= dalloc(&X, nn); rc |= dalloc(&VT, nn);
        rc |= dalloc(&Vp, (size_t)n * 96 * b); rc |= dalloc(&Wp, np); rc |= dalloc(&Tf, tstride() * b);
        rc |= dalloc(&px, (size_t)n * b); rc |= dalloc(&py, (size_t)n * b);
        rc |= dalloc(&d, (size_t)n * b); rc |= dalloc(&e, (size_t)n * b);
        rc |= dalloc(&w, (size_t)n * b);
        rc |= dalloc(&Vt, nn);
        rc |= dalloc(&scratch, (size_t)5 * n * tbk::stein_threads(n) * b);
        rc |= dalloc(&info, (size_t)b); rc |= dalloc(&sinfo, (size_t)b);
        return rc ? TBK_ERR_NOMEM : TBK_OK;
    }
    int alloc_mixed()
    {
        const size_t nn = (size_t)n * n * b;
        int rc = TBK_OK;
        rc |= dalloc(&A32, nn); rc |= dalloc(&VT32, nn); rc |= dalloc(&Y32, nn); rc |= dalloc(&X32, nn);
        rc |= dalloc(&P32, (size_t)n * 96 * b); rc |= dalloc(&T32, tstride() * b);
        rc |= dalloc(&lam, (size_t)n * b); rc |= dalloc(&perm, (size_t)n * b); rc |= dalloc(&delta, (size_t)b);
        return rc ? TBK_ERR_NOMEM : TBK_OK;
    }
    void release()
    {
        dfree(A); dfree(S); dfree(Y); dfree(Z); dfree(X); dfree(VT); dfree(Vp); dfree(Wp); dfree(Tf);
        dfree(px); dfree(py); dfree(d); dfree(e); dfree(w); dfree(Vt); dfree(scratch);
        dfree(info); dfree(sinfo);
        dfree(A32); dfree(P32); dfree(T32); dfree(VT32); dfree(Y32); dfree(X32); dfree(lam); dfree(perm); dfree(delta);
    }
};

tbk::GemmArgs gemm(int m, int n, int k, const double2 *a, int lda, size_t sa, const double2 *b,
                   int ldb, size_t sb, double2 *c, int ldc, size_t sc, int tri = 0,
                   double2 alpha = {1.0, 0.0}, double2 beta = {0.0, 0.0})
{
    tbk::GemmArgs g;
    g.m = m; g.n = n; g.k = k; g.alpha = alpha; g.beta = beta;
    g.a = a; g.lda = lda; g.sa = sa; g.b = b; g.ldb = ldb; g.sb = sb;
    g.c = c; g.ldc = ldc; g.sc = sc; g.tri = tri;
    return g;
}

struct StageTimer {
    bool on = false;
    hipStream_t st = nullptr;
    std::vector<std::string> names;
    std::vector<double> ms;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> pending;
    hipEvent_t cur = nullptr;
    int cur_idx = -1;
    int index(const char *nm)
    {
        for (size_t i = 0; i < names.size(); ++i) if (names[i] == nm) return (int)i;
        names.push_back(nm); ms.push_back(0.0);
        return (int)names.size() - 1;
    }
    void begin(const char *nm)
    {
        if (!on) return;
        hipEvent_t e0;
        (void)hipEventCreate(&e0);
        (void)hipEventRecord(e0, st);
        cur = e0; cur_idx = index(nm);
    }
    void end()
    {
        if (!on || !cur) return;
        hipEvent_t e1;
        (void)hipEventCreate(&e1);
        (void)hipEventRecord(e1, st);
        pending.push_back({cur_idx, {cur, e1}});
        cur = nullptr;
    }
    void collect()
    {
        for (auto &p : pending) {
            float t = 0.f;
            (void)hipEventSynchronize(p.second.second);
            (void)hipEventElapsedTime(&t, p.second.first, p.second.second);
            ms[p.first] += t;
            (void)hipEventDestroy(p.second.first);
            (void)hipEventDestroy(p.second.second);
        }
        pending.clear();
    }
    void reset() { for (auto &m : ms) m = 0.0; }
};

// Blocked tridiagonalisation: per panel one k_hetrd_panel launch and one
// GEMM for the trailing update A[s:, s:] -= [V W] [W V]^H.
void hetrd(Workspace &ws, int nb, hipStream_t st)
{
    const int n = ws.n;
    const size_t ss = ws.stride(), sp = (size_t)n * 96;
    tbk::HetrdArgs h{n, ss, ws.A, ws.Vp, ws.Wp, ws.Tf, ws.VT, ws.d, ws.e, ws.tstride()};
    for (int i0 = 0; i0 < n - 1; i0 += kNB) {
        const int pw = std::min(kNB, n - 1 - i0);
        tbk::launch_hetrd_panel(h, i0, nb, st);
        const int s = i0 + pw, m = n - s;
        if (m > 0)
            tbk::launch_zgemm(gemm(m, m, 2 * kNB, ws.Vp + s, n, sp, ws.Vp + (size_t)kNB * n + s, n, sp,
                                   ws.A + (size_t)s * n + s, n, ss, 4, double2{-1.0, 0.0},
                                   double2{1.0, 0.0}), 0, 1, nb, st);
    }
    tbk::launch_hetrd_finish(h, nb, st);
}

// S = U^H U and U^{-1} in place (lower triangle cleared): panel kernels for
// the diagonal blocks, GEMMs for the O(n^3) updates.
void potrf_trtri(Workspace &ws, int nb, hipStream_t st)
{
    const int n = ws.n;
    const size_t ss = ws.stride();
    for (int j0 = 0; j0 < n; j0 += kNB) {
        const int w = std::min(kNB, n - j0), s0 = j0 + w, m = n - s0;
        tbk::launch_potrf_panel(n, ws.S, ss, j0, ws.info, nb, st);
        if (m > 0) {   // S22 -= U12^H U12
            const double2 *u12 = ws.S + (size_t)s0 * n + j0;
            tbk::launch_zgemm(gemm(m, m, w, u12, n, ss, u12, n, ss, ws.S + (size_t)s0 * n + s0, n, ss,
                                   0, double2{-1.0, 0.0}, double2{1.0, 0.0}), 1, 0, nb, st);
        }
    }
    for (int j0 = 0; j0 < n; j0 += kNB) {
        const int w = std::min(kNB, n - j0);
        if (j0 > 0)    // Y = Uinv[0:j0, 0:j0] U[0:j0, J]
            tbk::launch_zgemm(gemm(j0, w, j0, ws.S, n, ss, ws.S + (size_t)j0 * n, n, ss, ws.Y, n, ss, 1),
                              0, 0, nb, st);
        tbk::launch_trtri_panel(n, ws.S, ss, ws.Y, j0, ws.info, nb, st);
    }
}

tbk::GemmArgsT<float> cgemm(int m, int n, int k, const float2 *a, int lda, size_t sa, const float2 *b,
                            int ldb, size_t sb, float2 *c, int ldc, size_t sc, int tri = 0,
                            float2 alpha = {1.f, 0.f}, float2 beta = {0.f, 0.f})
{
    tbk::GemmArgsT<float> g;
    g.m = m; g.n = n; g.k = k; g.alpha = alpha; g.beta = beta;
    g.a = a; g.lda = lda; g.sa = sa; g.b = b; g.ldb = ldb; g.sb = sb;
    g.c = c; g.ldc = ldc; g.sc = sc; g.tri = tri;
    return g;
}

// Mixed-precision solve: FP64 Cholesky, inverse and reduction; FP32 panel
// tridiagonalisation and back-transform on the FP32 MFMA units; FP64
// bisection/eigenvectors of the tridiagonal; one FP64 Ogita-Aishima step on
// the reduced problem C; FP64 Z = U^{-1} Y.
void solve_batch_mixed(Workspace &ws, int nb, double *w_out, hipStream_t st, StageTimer &tm)
{
    const int n = ws.n;
    const size_t ss = ws.stride(), sp = (size_t)n * 96;
    tm.begin("potrf_trtri");
    potrf_trtri(ws, nb, st);
    tm.end();
    tm.begin("reduce_gemm");
    tbk::launch_zgemm(gemm(n, n, n, ws.A, n, ss, ws.S, n, ss, ws.Y, n, ss, 3), 0, 0, nb, st);
    tbk::launch_zgemm(gemm(n, n, n, ws.S, n, ss, ws.Y, n, ss, ws.A, n, ss, 2), 1, 0, nb, st);
    tm.end();
    tm.begin("hetrd");
    tbk::launch_c64_to_c32(ss * nb, ws.A, ws.A32, st);
    tbk::HetrdArgsT<float> h{n, ss, ws.A32, ws.P32, nullptr, ws.T32, ws.VT32, ws.d, ws.e, ws.tstride()};
    for (int i0 = 0; i0 < n - 1; i0 += kNB) {
        const int pw = std::min(kNB, n - 1 - i0);
        tbk::launch_hetrd_panel(h, i0, nb, st);
        const int s = i0 + pw, m = n - s;
        if (m > 0)
            tbk::launch_cgemm(cgemm(m, m, 2 * kNB, ws.P32 + s, n, sp, ws.P32 + (size_t)kNB * n + s, n, sp,
                                    ws.A32 + (size_t)s * n + s, n, ss, 4, float2{-1.f, 0.f},
                                    float2{1.f, 0.f}), 0, 1, nb, st);
    }
    tbk::launch_hetrd_finish(h, nb, st);
    tm.end();
    tm.begin("stein");
    tbk::SteinArgs sa{n, 1e-4, ws.d, ws.e, ws.w, ws.Vt, ws.scratch, ws.sinfo};
    tbk::launch_stein(sa, nb, st);
    tm.end();
    tm.begin("backtransform");
    tbk::launch_real_to_c32_t(n, ws.Vt, ws.Y32, ss, nb, st);
    for (int i0 = ((n - 2) / kNB) * kNB; i0 >= 0; i0 -= kNB) {
        const int pw = std::min(kNB, n - 1 - i0);
        if (pw <= 0) continue;
        const int r0 = i0 + 1, m = n - r0;
        const float2 *Vsub = ws.A32 + (size_t)i0 * n + r0;
        const float2 *VTsub = ws.VT32 + (size_t)i0 * n + r0;
        tbk::launch_cgemm(cgemm(pw, n, m, VTsub, n, ss, ws.Y32 + r0, n, ss, ws.X32, n, ss), 1, 0, nb, st);
        tbk::launch_cgemm(cgemm(m, n, pw, Vsub, n, ss, ws.X32, n, ss, ws.Y32 + r0, n, ss, 0,
                                float2{-1.f, 0.f}, float2{1.f, 0.f}), 0, 0, nb, st);
    }
    tbk::launch_c32_to_c64(ss * nb, ws.Y32, ws.Y, st);
    tm.end();
    tm.begin("refine");
    // X = C Y, Z = Y^H X (= Y^H C Y), VT = Y^H Y
    tbk::launch_zgemm(gemm(n, n, n, ws.A, n, ss, ws.Y, n, ss, ws.X, n, ss), 0, 0, nb, st);
    tbk::launch_zgemm(gemm(n, n, n, ws.Y, n, ss, ws.X, n, ss, ws.Z, n, ss), 1, 0, nb, st);
    tbk::launch_zgemm(gemm(n, n, n, ws.Y, n, ss, ws.Y, n, ss, ws.VT, n, ss), 1, 0, nb, st);
    // X = (I + E) with sorted columns; VT = Y X; Z = U^{-1} VT
    tbk::launch_oa_update(n, ws.Z, ws.VT, ws.X, ss, ws.lam, ws.perm, ws.delta, w_out, nb, st);
    tbk::launch_zgemm(gemm(n, n, n, ws.Y, n, ss, ws.X, n, ss, ws.VT, n, ss), 0, 0, nb, st);
    tbk::launch_zgemm(gemm(n, n, n, ws.S, n, ss, ws.VT, n, ss, ws.Z, n, ss, 1), 0, 0, nb, st);
    tm.end();
}

// Stages 2-9 of the pipeline on ws.A = H, ws.S = S for nb k-points:
// eigenvalues into w_out[nb*n], eigenvectors into ws.Z.
void solve_batch(Workspace &ws, int nb, double *w_out, hipStream_t st, StageTimer &tm)
{
    const int n = ws.n;
    const size_t ss = ws.stride();
    tm.begin("potrf_trtri");
    potrf_trtri(ws, nb, st);
    tm.end();
    tm.begin("reduce_gemm");
    // Y = H Uinv (Uinv upper) ; A = Uinv^H Y
    tbk::launch_zgemm(gemm(n, n, n, ws.A, n, ss, ws.S, n, ss, ws.Y, n, ss, 3), 0, 0, nb, st);
    tbk::launch_zgemm(gemm(n, n, n, ws.S, n, ss, ws.Y, n, ss, ws.A, n, ss, 2), 1, 0, nb, st);
    tm.end();
    tm.begin("hetrd");
    hetrd(ws, nb, st);
    tm.end();
    tm.begin("stein");
    tbk::SteinArgs sa{n, 1e-4, ws.d, ws.e, w_out, ws.Vt, ws.scratch, ws.sinfo};
    tbk::launch_stein(sa, nb, st);
    tm.end();
    tm.begin("backtransform");
    tbk::launch_real_to_complex_t(n, ws.Vt, ws.Y, ss, nb, st);
    for (int i0 = ((n - 2) / kNB) * kNB; i0 >= 0; i0 -= kNB) {
        const int pw = std::min(kNB, n - 1 - i0);
        if (pw <= 0) continue;
        const int r0 = i0 + 1, m = n - r0;
        const double2 *Vsub = ws.A + (size_t)i0 * n + r0;
        const double2 *VTsub = ws.VT + (size_t)i0 * n + r0;
        // X[0:pw, :] = (V_p T_p^H)^H Y[r0:, :]
        tbk::launch_zgemm(gemm(pw, n, m, VTsub, n, ss, ws.Y + r0, n, ss, ws.X, n, ss), 1, 0, nb, st);
        // Y[r0:, :] -= V_p X[0:pw, :]
        tbk::launch_zgemm(gemm(m, n, pw, Vsub, n, ss, ws.X, n, ss, ws.Y + r0, n, ss, 0,
                               double2{-1.0, 0.0}, double2{1.0, 0.0}), 0, 0, nb, st);
    }
    // Z = Uinv Y
    tbk::launch_zgemm(gemm(n, n, n, ws.S, n, ss, ws.Y, n, ss, ws.Z, n, ss, 1), 0, 0, nb, st);
    tm.end();
}

}  // namespace

struct tbk_plan {
    int device = 0;
    hipStream_t stream = nullptr;
    int natoms = 0, nv = 0, npairs = 0, nk = 0, batch = 0, precision = TBK_FP64;
    double thisvol = 0.0, zfill = 0.0;
    std::vector<double> kpts, weights;
    // structure data on the device
    int *row_atom = nullptr, *row_orb = nullptr, *ksk = nullptr, *nbas = nullptr, *blk_ptr = nullptr;
    double *dvec = nullptr, *eh = nullptr, *es = nullptr, *edh = nullptr, *eds = nullptr;
    double *paros = nullptr, *dparos = nullptr, *kp = nullptr, *wts = nullptr;
    // results
    double *esk = nullptr, *desk = nullptr, *out = nullptr;
    double2 *ph = nullptr;          // Bloch phases of the current batch
    int *ncut = nullptr, *err = nullptr;
    std::vector<int> info_host;
    Workspace ws;
    StageTimer timer;
    int last_mode = 0;
};

static void plan_free(tbk_plan *p)
{
    if (!p) return;
    (void)hipSetDevice(p->device);
    dfree(p->row_atom); dfree(p->row_orb); dfree(p->ksk); dfree(p->nbas); dfree(p->blk_ptr);
    dfree(p->dvec); dfree(p->eh); dfree(p->es); dfree(p->edh); dfree(p->eds);
    dfree(p->paros); dfree(p->dparos); dfree(p->kp); dfree(p->wts);
    dfree(p->esk); dfree(p->desk); dfree(p->out); dfree(p->ncut); dfree(p->err); dfree(p->ph);
    p->ws.release();
    if (p->stream) (void)hipStreamDestroy(p->stream);
    delete p;
}

template <class T>
static int upload(T **dst, const std::vector<T> &src)
{
    int rc = dalloc(dst, std::max<size_t>(src.size(), 1));
    if (rc) return rc;
    if (!src.empty() && hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
        return TBK_ERR_HIP;
    return TBK_OK;
}

extern "C" int tbk_plan_create(int device, const tbk_structure_desc *st, int nk,
                               const double *kpts, const double *weights, int batch,
                               int precision, tbk_plan **out)
{
    if (!st || !out || nk <= 0 || !kpts || !weights || st->natoms <= 0) return TBK_ERR_ARG;
    if (precision != TBK_FP64 && precision != TBK_MIXED) return TBK_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return TBK_ERR_NODEVICE;
    if (device < 0 || device >= ndev) return TBK_ERR_ARG;
    if (hipSetDevice(device) != hipSuccess) return TBK_ERR_HIP;
    const int nat = st->natoms, kinds = st->kinds;
    // ---- host set-up (bit-exact) ----
    tbk::PairList pl;
    const int maxpairs = std::max(4096, 4000 * nat);
    int rc = tbk::search2(nat, st->lat, st->posn, st->kkind, kinds, st->nd, st->dnn, st->screenl,
                          maxpairs, pl);
    if (rc) return rc;
    const int np = pl.n;
    std::vector<double> den((size_t)nat * kinds), paros(4 * nat), dparos(4 * nat);
    std::vector<double> hop(14 * (size_t)np), over(14 * (size_t)np), dhop(14 * (size_t)np),
        dover(14 * (size_t)np);
    tbk::setpar(nat, st->kkind, kinds, st->param, st->nltype, st->jsover, pl, den.data(),
                paros.data(), hop.data(), over.data());
    tbk::setparv(nat, st->kkind, kinds, st->param, st->nltype, st->jsover, st->thisvol, pl,
                 den.data(), dparos.data(), dhop.data(), dover.data());
    std::vector<int> ksk(nat), nbas(nat);
    int nv = 0;
    for (int a = 0; a < nat; ++a) {
        ksk[a] = nv;
        nbas[a] = st->kbas[st->kkind[a] - 1];
        if (nbas[a] < 1 || nbas[a] > 9) return TBK_ERR_ARG;
        nv += nbas[a];
    }
    if (nv < 2 || nv > kMaxN) return TBK_ERR_ARG;
    std::vector<int> row_atom(nv), row_orb(nv);
    for (int a = 0; a < nat; ++a)
        for (int o = 0; o < nbas[a]; ++o) { row_atom[ksk[a] + o] = a; row_orb[ksk[a] + o] = o; }
    std::vector<double> eh(81 * (size_t)np), es(81 * (size_t)np), edh(81 * (size_t)np),
        eds(81 * (size_t)np), dvec(3 * (size_t)np);
    std::vector<int> blk_ptr((size_t)nat * nat + 1, 0);
    for (int p = 0; p < np; ++p) {
        const tbk::Pair &q = pl.pairs[p];
        double v[14];
        const bool unlike = q.kind[0] != q.kind[1];
        tbk::arrange_integrals(&hop[14 * p], q.kind[0], q.kind[1], v);
        tbk::sk_block(q.cosines[0], q.cosines[1], q.cosines[2], v, unlike, &eh[81 * p]);
        tbk::arrange_integrals(&over[14 * p], q.kind[0], q.kind[1], v);
        tbk::sk_block(q.cosines[0], q.cosines[1], q.cosines[2], v, unlike, &es[81 * p]);
        tbk::arrange_integrals(&dhop[14 * p], q.kind[0], q.kind[1], v);
        tbk::sk_block(q.cosines[0], q.cosines[1], q.cosines[2], v, unlike, &edh[81 * p]);
        tbk::arrange_integrals(&dover[14 * p], q.kind[0], q.kind[1], v);
        tbk::sk_block(q.cosines[0], q.cosines[1], q.cosines[2], v, unlike, &eds[81 * p]);
        const int i = q.atom[0], j = q.atom[1];
        for (int c = 0; c < 3; ++c)    // rotate.f:144-146 operand order
            dvec[3 * p + c] = q.dlv[c] - st->posq[3 * j + c] + st->posq[3 * i + c];
        blk_ptr[(size_t)i * nat + j + 1]++;
    }
    for (size_t b = 1; b < blk_ptr.size(); ++b) blk_ptr[b] += blk_ptr[b - 1];
    // ---- device state ----
    tbk_plan *p = new tbk_plan();
    p->device = device;
    p->natoms = nat; p->nv = nv; p->npairs = np; p->nk = nk; p->precision = precision;
    p->thisvol = st->thisvol; p->zfill = st->zfill;
    p->kpts.assign(kpts, kpts + 3 * (size_t)nk);
    p->weights.assign(weights, weights + nk);
    p->batch = (batch <= 0 || batch > nk) ? nk : batch;
    if (hipStreamCreate(&p->stream) != hipSuccess) { delete p; return TBK_ERR_HIP; }
    rc = TBK_OK;
    rc |= upload(&p->row_atom, row_atom); rc |= upload(&p->row_orb, row_orb);
    rc |= upload(&p->ksk, ksk); rc |= upload(&p->nbas, nbas); rc |= upload(&p->blk_ptr, blk_ptr);
    rc |= upload(&p->dvec, dvec); rc |= upload(&p->eh, eh); rc |= upload(&p->es, es);
    rc |= upload(&p->edh, edh); rc |= upload(&p->eds, eds);
    rc |= upload(&p->paros, paros); rc |= upload(&p->dparos, dparos);
    rc |= upload(&p->kp, p->kpts); rc |= upload(&p->wts, p->weights);
    rc |= dalloc(&p->esk, (size_t)nk * nv); rc |= dalloc(&p->desk, (size_t)nk * nv);
    rc |= dalloc(&p->ncut, (size_t)nk); rc |= dalloc(&p->err, 1); rc |= dalloc(&p->out, 8);
    if (!rc) rc = dalloc(&p->ph, (size_t)p->batch * std::max(np, 1));
    if (!rc) rc = p->ws.alloc(nv, p->batch);
    if (!rc && precision == TBK_MIXED && nv > tbk::oa_max_n()) rc = TBK_ERR_ARG;
    if (!rc && precision == TBK_MIXED) rc = p->ws.alloc_mixed();
    if (rc) { plan_free(p); return rc == TBK_ERR_HIP ? TBK_ERR_HIP : TBK_ERR_NOMEM; }
    p->timer.st = p->stream;
    p->info_host.assign(nk, 0);
    *out = p;
    return TBK_OK;
}

static tbk::BuildArgs build_args(tbk_plan *p, int k0, bool deriv)
{
    tbk::BuildArgs a;
    a.n = p->nv; a.natoms = p->natoms; a.onsite_s = deriv ? 0 : 1; a.npairs = p->npairs;
    a.ph = p->ph;
    a.stride = p->ws.stride();
    a.row_atom = p->row_atom; a.row_orb = p->row_orb; a.ksk = p->ksk; a.nbas = p->nbas;
    a.blk_ptr = p->blk_ptr; a.dvec = p->dvec;
    a.eh = deriv ? p->edh : p->eh; a.es = deriv ? p->eds : p->es;
    a.onsite = deriv ? p->dparos : p->paros;
    a.h = p->ws.A; a.s = p->ws.S;
    return a;
}

extern "C" int tbk_plan_launch(tbk_plan *p, int mode, double tkb, double eigcut)
{
    if (!p || mode < 1 || mode > 4) return TBK_ERR_ARG;
    if (p->zfill <= 0.0) return TBK_ERR_ZFILL;
    HIP_TRY(hipSetDevice(p->device));
    const bool lpress = mode == 2 || mode == 4;
    const int n = p->nv;
    const size_t ss = p->ws.stride();
    hipStream_t st = p->stream;
    HIP_TRY(hipMemsetAsync(p->err, 0, sizeof(int), st));
    HIP_TRY(hipMemsetAsync(p->desk, 0, sizeof(double) * (size_t)p->nk * n, st));
    p->last_mode = mode;
    for (int k0 = 0; k0 < p->nk; k0 += p->batch) {
        const int nb = std::min(p->batch, p->nk - k0);
        p->timer.begin("build_hs");
        tbk::launch_phases(p->npairs, p->kp + 3 * (size_t)k0, p->dvec, p->ph, nb, st);
        tbk::launch_build_hs(build_args(p, k0, false), p->natoms, nb, st);
        p->timer.end();
        if (p->precision == TBK_MIXED)
            solve_batch_mixed(p->ws, nb, p->esk + (size_t)k0 * n, st, p->timer);
        else
            solve_batch(p->ws, nb, p->esk + (size_t)k0 * n, st, p->timer);
        HIP_TRY(hipMemcpyAsync(p->info_host.data() + k0, p->ws.info, sizeof(int) * nb,
                               hipMemcpyDeviceToHost, st));
        if (lpress) {
            p->timer.begin("pressure");
            tbk::launch_build_hs(build_args(p, k0, true), p->natoms, nb, st);
            tbk::launch_zgemm(gemm(n, n, n, p->ws.A, n, ss, p->ws.Z, n, ss, p->ws.X, n, ss), 0, 0, nb, st);
            tbk::launch_zgemm(gemm(n, n, n, p->ws.S, n, ss, p->ws.Z, n, ss, p->ws.Y, n, ss), 0, 0, nb, st);
            tbk::launch_coldot(n, p->ws.Z, p->ws.X, p->ws.Y, ss, p->ws.px, p->ws.py, nb, st);
            tbk::PerturbArgs pa;
            pa.n = n; pa.nb = nb; pa.k0 = k0; pa.stride = ss; pa.eigcut = eigcut;
            pa.w = p->esk + (size_t)k0 * n; pa.px = p->ws.px; pa.py = p->ws.py;
            pa.z = p->ws.Z; pa.x = p->ws.X; pa.y = p->ws.Y;
            pa.desk = p->desk + (size_t)k0 * n; pa.ncut = p->ncut; pa.err = p->err;
            tbk::launch_perturb(pa, st);
            p->timer.end();
        }
    }
    p->timer.begin("fermi_sums");
    tbk::FermiArgs f;
    f.nk = p->nk; f.nv = n; f.ld = n; f.mode = mode;
    f.el = 2.0; f.tkb = tkb; f.zfill = p->zfill;
    f.esk = p->esk; f.desk = p->desk; f.w = p->wts; f.ncut = p->ncut; f.out = p->out;
    tbk::launch_fermi_sums(f, st);
    p->timer.end();
    if (hipGetLastError() != hipSuccess) return TBK_ERR_HIP;
    return TBK_OK;
}

extern "C" int tbk_plan_wait(tbk_plan *p, double *efermi, double *earg1, double *earg2)
{
    if (!p) return TBK_ERR_ARG;
    HIP_TRY(hipSetDevice(p->device));
    double o[8];
    int err = 0;
    HIP_TRY(hipMemcpyAsync(o, p->out, sizeof(o), hipMemcpyDeviceToHost, p->stream));
    HIP_TRY(hipMemcpyAsync(&err, p->err, sizeof(int), hipMemcpyDeviceToHost, p->stream));
    HIP_TRY(hipStreamSynchronize(p->stream));
    p->timer.collect();
    for (int k = 0; k < p->nk; ++k)
        if (p->info_host[k] != 0) return TBK_ERR_OVERLAP;
    if (err & 2) return TBK_ERR_DEGEN;
    if (efermi) *efermi = o[0];
    if (earg1) *earg1 = o[1];
    if (earg2) *earg2 = o[2];
    return TBK_OK;
}

extern "C" int tbk_plan_run(tbk_plan *p, int mode, double tkb, double eigcut, double *efermi,
                            double *earg1, double *earg2)
{
    int rc = tbk_plan_launch(p, mode, tkb, eigcut);
    if (rc) return rc;
    return tbk_plan_wait(p, efermi, earg1, earg2);
}

// GEMM autotuner: one mode-4 pass records the plan's GEMM shapes, then each
// shape not yet in the table is timed with every tile (one warm-up and
// `reps` launches on the plan's stream) and the fastest tile is kept for the
// rest of the process.  The timing launches overwrite workspace contents, so
// the plan's results are undefined until its next run.
extern "C" int tbk_plan_autotune(tbk_plan *p, int reps, int *ntuned)
{
    if (!p || reps < 1) return TBK_ERR_ARG;
    HIP_TRY(hipSetDevice(p->device));
    tbk::gemm_record(true);
    int rc = tbk_plan_launch(p, 4, 0.005, 1e30);
    if (!rc) rc = tbk_plan_wait(p, nullptr, nullptr, nullptr);
    tbk::gemm_record(false);
    if (rc) return rc;
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    int count = 0;
    for (const tbk::GemmShape &s : tbk::gemm_recorded()) {
        if (tbk::gemm_tuned(s)) continue;
        int best = 0;
        float bt = 3.4e38f;
        for (int t = 0; t < tbk::kGemmTiles; ++t) {
            tbk::launch_gemm_tile(s, t, p->stream);
            (void)hipEventRecord(e0, p->stream);
            for (int r = 0; r < reps; ++r) tbk::launch_gemm_tile(s, t, p->stream);
            (void)hipEventRecord(e1, p->stream);
            if (hipEventSynchronize(e1) != hipSuccess) { rc = TBK_ERR_HIP; break; }
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            if (ms < bt) { bt = ms; best = t; }
        }
        if (rc) break;
        tbk::gemm_store_tile(s, best);
        ++count;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (ntuned) *ntuned = count;
    if (!rc && hipGetLastError() != hipSuccess) rc = TBK_ERR_HIP;
    return rc;
}

// Tile of a GEMM shape (the autotuned entry, else the cost model); for
// reporting which tiles the plan's launches use.
extern "C" int tbk_gemm_tile(int precision, int opa, int opb, int m, int n, int k, int tri, int nb)
{
    return tbk::gemm_tile(precision == TBK_MIXED ? 1 : 0, opa, opb, m, n, k, tri, nb);
}

extern "C" int tbk_plan_eigenvalues(tbk_plan *p, double *esk, double *desk)
{
    if (!p) return TBK_ERR_ARG;
    HIP_TRY(hipSetDevice(p->device));
    const size_t n = (size_t)p->nk * p->nv;
    if (esk) HIP_TRY(hipMemcpy(esk, p->esk, n * sizeof(double), hipMemcpyDeviceToHost));
    if (desk) HIP_TRY(hipMemcpy(desk, p->desk, n * sizeof(double), hipMemcpyDeviceToHost));
    return TBK_OK;
}

extern "C" int tbk_plan_eigenvectors(tbk_plan *p, int k, double *z)
{
    if (!p || !z || k < 0 || k >= p->nk) return TBK_ERR_ARG;
    // the workspace holds the last batch: k must fall inside it
    const int last0 = ((p->nk - 1) / p->batch) * p->batch;
    if (k < last0) return TBK_ERR_ARG;
    HIP_TRY(hipSetDevice(p->device));
    const size_t ss = p->ws.stride();
    HIP_TRY(hipMemcpy(z, p->ws.Z + (size_t)(k - last0) * ss, ss * sizeof(double2),
                      hipMemcpyDeviceToHost));
    return TBK_OK;
}

extern "C" int tbk_plan_hs(tbk_plan *p, int k, int deriv, double *h, double *s)
{
    if (!p || k < 0 || k >= p->nk) return TBK_ERR_ARG;
    HIP_TRY(hipSetDevice(p->device));
    tbk::launch_phases(p->npairs, p->kp + 3 * (size_t)k, p->dvec, p->ph, 1, p->stream);
    tbk::launch_build_hs(build_args(p, k, deriv != 0), p->natoms, 1, p->stream);
    HIP_TRY(hipStreamSynchronize(p->stream));
    const size_t ss = p->ws.stride();
    if (h) HIP_TRY(hipMemcpy(h, p->ws.A, ss * sizeof(double2), hipMemcpyDeviceToHost));
    if (s) HIP_TRY(hipMemcpy(s, p->ws.S, ss * sizeof(double2), hipMemcpyDeviceToHost));
    return TBK_OK;
}

extern "C" int tbk_plan_nv(const tbk_plan *p) { return p ? p->nv : -1; }
extern "C" int tbk_plan_npairs(const tbk_plan *p) { return p ? p->npairs : -1; }

extern "C" int tbk_plan_destroy(tbk_plan *p)
{
    plan_free(p);
    return TBK_OK;
}

extern "C" int tbk_plan_stage_times(const tbk_plan *p, int cap, const char **names, double *ms,
                                    int *n)
{
    if (!p || !n) return TBK_ERR_ARG;
    int m = std::min<int>(cap, (int)p->timer.names.size());
    for (int i = 0; i < m; ++i) { names[i] = p->timer.names[i].c_str(); ms[i] = p->timer.ms[i]; }
    *n = m;
    return TBK_OK;
}

extern "C" int tbk_plan_set_timing(tbk_plan *p, int on)
{
    if (!p) return TBK_ERR_ARG;
    p->timer.collect();
    p->timer.reset();
    p->timer.on = on != 0;
    return TBK_OK;
}

extern "C" int tbk_zhegv(int device, int n, const double *h, const double *s, double *w, double *z)
{
    if (n < 2 || n > kMaxN || !h || !s || !w) return TBK_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return TBK_ERR_NODEVICE;
    HIP_TRY(hipSetDevice(device));
    Workspace ws;
    if (ws.alloc(n, 1)) { ws.release(); return TBK_ERR_NOMEM; }
    hipStream_t st;
    if (hipStreamCreate(&st) != hipSuccess) { ws.release(); return TBK_ERR_HIP; }
    const size_t ss = ws.stride();
    int rc = TBK_OK, info = 0;
    StageTimer tm;
    if (hipMemcpyAsync(ws.A, h, ss * sizeof(double2), hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(ws.S, s, ss * sizeof(double2), hipMemcpyHostToDevice, st) != hipSuccess) {
        rc = TBK_ERR_HIP;
    } else {
        solve_batch(ws, 1, ws.w, st, tm);
        if (hipMemcpyAsync(w, ws.w, n * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipMemcpyAsync(&info, ws.info, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
            (z && hipMemcpyAsync(z, ws.Z, ss * sizeof(double2), hipMemcpyDeviceToHost, st) != hipSuccess) ||
            hipStreamSynchronize(st) != hipSuccess)
            rc = TBK_ERR_HIP;
        else if (info != 0)
            rc = TBK_ERR_OVERLAP;
    }
    (void)hipStreamDestroy(st);
    ws.release();
    return rc;
}

extern "C" int tbk_fermi(int device, int ldn, int nk, int nv, double el, const double *weight,
                         const double *eigen, double tkb, double zfill, double *efermi)
{
    if (!weight || !eigen || !efermi || nk <= 0 || nv < 2 || ldn < nv) return TBK_ERR_ARG;
    if (zfill <= 0.0) return TBK_ERR_ZFILL;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return TBK_ERR_NODEVICE;
    HIP_TRY(hipSetDevice(device));
    double *de = nullptr, *dw = nullptr, *dout = nullptr;
    int *dn = nullptr;
    int rc = dalloc(&de, (size_t)ldn * nk) | dalloc(&dw, (size_t)nk) | dalloc(&dout, 8) |
             dalloc(&dn, (size_t)nk);
    if (!rc) {
        if (hipMemcpy(de, eigen, sizeof(double) * ldn * nk, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(dw, weight, sizeof(double) * nk, hipMemcpyHostToDevice) != hipSuccess) {
            rc = TBK_ERR_HIP;
        } else {
            tbk::FermiArgs f;
            f.nk = nk; f.nv = nv; f.ld = ldn; f.mode = 3; f.el = el; f.tkb = tkb; f.zfill = zfill;
            f.esk = de; f.desk = de; f.w = dw; f.ncut = dn; f.out = dout;
            tbk::launch_fermi_sums(f, nullptr);
            double o[8];
            if (hipMemcpy(o, dout, sizeof(o), hipMemcpyDeviceToHost) != hipSuccess) rc = TBK_ERR_HIP;
            else *efermi = o[0];
        }
    } else {
        rc = TBK_ERR_NOMEM;
    }
    dfree(de); dfree(dw); dfree(dout); dfree(dn);
    return rc;
}

extern "C" const char *tbk_error_string(int code)
{
    switch (code) {
    case TBK_OK: return "ok";
    case TBK_ERR_ARG: return "invalid argument";
    case TBK_ERR_PAIRS: return "pair list overflow (search2)";
    case TBK_ERR_OVERLAP: return "overlap matrix not positive definite";
    case TBK_ERR_CONVERGE: return "tridiagonal eigensolver did not converge";
    case TBK_ERR_DEGEN: return "degenerate group larger than 9 (perturb)";
    case TBK_ERR_ZFILL: return "zfill <= 0";
    case TBK_ERR_HIP: return "HIP runtime error";
    case TBK_ERR_NOMEM: return "device allocation failed";
    case TBK_ERR_NODEVICE: return "no GPU device";
    default: return "unknown error";
    }
}

// ---------------------------------------------------------------------------
// Diagnostics: run the solver stage by stage on one matrix pair and return
// every intermediate (used by the stage-level parity tests).
// ---------------------------------------------------------------------------
extern "C" int tbk_debug_stages(int device, int n, const double *h, const double *s,
                                double *uinv, double *c, double *refl, double *d, double *e,
                                double *w, double *vt, double *z)
{
    if (n < 2 || n > kMaxN || !h || !s) return TBK_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return TBK_ERR_NODEVICE;
    HIP_TRY(hipSetDevice(device));
    Workspace ws;
    if (ws.alloc(n, 1)) { ws.release(); return TBK_ERR_NOMEM; }
    const size_t ss = ws.stride();
    hipStream_t st = nullptr;
    int rc = TBK_OK;
    auto get = [&](void *dst, const void *src, size_t bytes) {
        if (dst && rc == TBK_OK && hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost) != hipSuccess)
            rc = TBK_ERR_HIP;
    };
    if (hipMemcpy(ws.A, h, ss * sizeof(double2), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(ws.S, s, ss * sizeof(double2), hipMemcpyHostToDevice) != hipSuccess) {
        ws.release();
        return TBK_ERR_HIP;
    }
    potrf_trtri(ws, 1, st);
    (void)hipDeviceSynchronize();
    get(uinv, ws.S, ss * sizeof(double2));
    tbk::launch_zgemm(gemm(n, n, n, ws.A, n, ss, ws.S, n, ss, ws.Y, n, ss, 3), 0, 0, 1, st);
    tbk::launch_zgemm(gemm(n, n, n, ws.S, n, ss, ws.Y, n, ss, ws.A, n, ss, 2), 1, 0, 1, st);
    (void)hipDeviceSynchronize();
    get(c, ws.A, ss * sizeof(double2));
    hetrd(ws, 1, st);
    (void)hipDeviceSynchronize();
    get(refl, ws.A, ss * sizeof(double2));
    get(d, ws.d, n * sizeof(double));
    get(e, ws.e, n * sizeof(double));
    tbk::SteinArgs sa{n, 1e-4, ws.d, ws.e, ws.w, ws.Vt, ws.scratch, ws.sinfo};
    tbk::launch_stein(sa, 1, st);
    (void)hipDeviceSynchronize();
    get(w, ws.w, n * sizeof(double));
    get(vt, ws.Vt, ss * sizeof(double));
    int fails = 0;
    get(&fails, ws.sinfo, sizeof(int));
    if (rc == TBK_OK && fails) rc = -fails;      // inverse-iteration failures
    // full pipeline again from the inputs for Z
    if (hipMemcpy(ws.A, h, ss * sizeof(double2), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(ws.S, s, ss * sizeof(double2), hipMemcpyHostToDevice) != hipSuccess) {
        ws.release();
        return TBK_ERR_HIP;
    }
    StageTimer tm;
    solve_batch(ws, 1, ws.w, st, tm);
    (void)hipDeviceSynchronize();
    get(z, ws.Z, ss * sizeof(double2));
    ws.release();
    return rc;
}

template <class C, class GA>
static int debug_gemm_t(int tile, int m, int n, int k, int opa, int opb, int tri, const void *a, int lda,
                        const void *b, int ldb, void *c, int ldc)
{
    const size_t na = (size_t)lda * (opa ? m : k), nb = (size_t)ldb * (opb ? k : n),
                 nc = (size_t)ldc * n;
    C *da = nullptr, *db = nullptr, *dc = nullptr;
    int rc = dalloc(&da, na) | dalloc(&db, nb) | dalloc(&dc, nc);
    if (!rc) {
        (void)hipMemcpy(da, a, na * sizeof(C), hipMemcpyHostToDevice);
        (void)hipMemcpy(db, b, nb * sizeof(C), hipMemcpyHostToDevice);
        (void)hipMemset(dc, 0, nc * sizeof(C));
        GA g;
        g.m = m; g.n = n; g.k = k; g.tri = tri;
        g.alpha.x = 1; g.alpha.y = 0; g.beta.x = 0; g.beta.y = 0;
        g.a = da; g.lda = lda; g.sa = 0; g.b = db; g.ldb = ldb; g.sb = 0; g.c = dc; g.ldc = ldc; g.sc = 0;
        tbk::GemmShape sh{sizeof(C) == 16 ? 0 : 1, opa, opb, m, n, k, tri, 1, {}, {}};
        if constexpr (sizeof(C) == 16) sh.g64 = g; else sh.g32 = g;
        const int t = tile >= 0 ? tile : tbk::gemm_tile(sh.prec, opa, opb, m, n, k, tri, 1);
        tbk::launch_gemm_tile(sh, t, nullptr);
        if (hipMemcpy(c, dc, nc * sizeof(C), hipMemcpyDeviceToHost) != hipSuccess) rc = TBK_ERR_HIP;
    } else {
        rc = TBK_ERR_NOMEM;
    }
    dfree(da); dfree(db); dfree(dc);
    return rc;
}

extern "C" int tbk_debug_zgemm(int device, int precision, int tile, int m, int n, int k, int opa, int opb,
                               int tri, const void *a, int lda, const void *b, int ldb, void *c, int ldc)
{
    if (tile >= tbk::kGemmTiles || opa < 0 || opa > 1 || opb < 0 || opb > 1) return TBK_ERR_ARG;
    HIP_TRY(hipSetDevice(device));
    if (precision == TBK_MIXED)
        return debug_gemm_t<float2, tbk::GemmArgsT<float>>(tile, m, n, k, opa, opb, tri, a, lda, b, ldb, c, ldc);
    return debug_gemm_t<double2, tbk::GemmArgsT<double>>(tile, m, n, k, opa, opb, tri, a, lda, b, ldb, c, ldc);
}
