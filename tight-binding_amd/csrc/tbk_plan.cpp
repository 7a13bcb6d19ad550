// tbk_plan.cpp -- device plans: host set-up upload, the per-batch k-point
// pipeline, Fermi + sums, and the C ABI around them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#include <thread>
#include <atomic>
#include <mutex>
#include <type_traits>

#include "tbk.h"
#include "tbk_internal.h"
#include "tbk_kernels.h"
#include "tbk_prepared.h"

namespace {

constexpr int kNB = 32;        // panel width of the factorizations
constexpr int kMaxN = 1280;    // LDS-resident vectors of k_hetrd / k_stein
// Eigenvalues closer than kOrtolRel * ||T||_1 have their vectors
// orthogonalised together.  A twisted-factorisation vector of an eigenvalue
// with gap g is accurate to an angle of about eps ||T|| / g, so vectors
// outside a group are orthogonal to about 1e-10 (dstein's 1e-3 serves
// inverse iteration, whose vectors depend on the separation much more).
constexpr double kOrtolRel = 1e-6;
// Tridiagonalisation over 64x64 tiles (several workgroups per item) for
// batches below kWideMaxBatch items of order >= kWideMinN.
constexpr int kWideMinN = 128, kWideMaxBatch = 160;
// Two-stage tridiagonalisation (band width TB) by default from this order
// up, in both methods (at smaller orders the one-stage panels are faster).
constexpr int kTwoStageMinN = 512;
using tbk::TB;
using tbk::TBLD;
// The two-stage reduction keeps its band and its stage-2 reflectors and T
// factors in n x n work matrices: true where they fit (every order >= 40).
inline bool two_stage_fits(int n) { return n >= 40 && (size_t)tbk::band_diamonds(n) * TB * TB <= (size_t)n * n; }

#define HIP_TRY(x)                                   \
    do {                                             \
        hipError_t e_ = (x);                         \
        if (e_ != hipSuccess) return TBK_ERR_HIP;    \
    } while (0)

thread_local size_t g_dev_bytes = 0;     // device bytes allocated by this thread
// Plans allocate stream-ordered (hipMallocAsync / hipFreeAsync on their own
// stream), so creating and destroying a plan never synchronises the device
// while other plans run; the one-shot entry points use plain hipMalloc.
thread_local hipStream_t g_alloc_stream = nullptr;

struct AllocStream {
    hipStream_t prev;
    explicit AllocStream(hipStream_t s) : prev(g_alloc_stream) { g_alloc_stream = s; }
    ~AllocStream() { g_alloc_stream = prev; }
};

template <class T>
int dalloc(T **p, size_t n)
{
    *p = nullptr;
    if (n == 0) return TBK_OK;
    const hipError_t e = g_alloc_stream ? hipMallocAsync((void **)p, n * sizeof(T), g_alloc_stream)
                                        : hipMalloc((void **)p, n * sizeof(T));
    if (e != hipSuccess) { *p = nullptr; return TBK_ERR_NOMEM; }
    g_dev_bytes += n * sizeof(T);
    return TBK_OK;
}
template <class T>
void dfree(T *&p)
{
    if (p) {
        if (g_alloc_stream) (void)hipFreeAsync((void *)p, g_alloc_stream);
        else (void)hipFree((void *)p);
    }
    p = nullptr;
}

// Device buffers of one batch of b k-points at order n.
struct Workspace {
    int n = 0, b = 0;
    double2 *A = nullptr, *S = nullptr, *Y = nullptr, *Z = nullptr, *X = nullptr;
    double2 *Vp = nullptr, *Wp = nullptr, *Tf = nullptr, *px = nullptr, *py = nullptr;
    double2 *VT = nullptr;
    // dH/dV, dS/dV of the pressure modes, built with H and S in one pass
    double2 *dA = nullptr, *dS = nullptr;
    double *d = nullptr, *e = nullptr, *Vt = nullptr, *scratch = nullptr, *w = nullptr;
    int *info = nullptr, *sinfo = nullptr;
    // mixed precision: FP32 copies for the reduction and back-transform
    float2 *A32 = nullptr, *P32 = nullptr, *T32 = nullptr, *VT32 = nullptr, *Y32 = nullptr,
           *X32 = nullptr;
    // the refinement step's scratch (tbk::OaArgs)
    double *lam = nullptr, *rn = nullptr, *tau = nullptr;
    int *perm = nullptr, *perm2 = nullptr, *cinfo = nullptr, *clist = nullptr, *ncl = nullptr;
    // multi-workgroup tridiagonalisation (few items of large order)
    bool wide = false;
    // two-stage tridiagonalisation (band reduction + bulge chasing) instead
    // of the one-stage panels
    bool two_stage = false;
    // resident tridiagonalisation (the matrix held on chip, one workgroup per
    // item) instead of the one-stage panels: the default where the order fits
    bool resident = false;
    double2 *wv = nullptr, *wcol = nullptr, *wcarry = nullptr, *wydot = nullptr, *wyax = nullptr,
            *wudot = nullptr, *wscal = nullptr;
    float2 *wv32 = nullptr, *wcol32 = nullptr, *wcarry32 = nullptr, *wydot32 = nullptr, *wyax32 = nullptr,
           *wudot32 = nullptr, *wscal32 = nullptr;
    int tmax() const { return (n - 1 + 63) / 64; }
    size_t stride() const { return (size_t)n * n; }
    // WY factors: one kNB x kNB block per panel, ceil(n / kNB) panels
    size_t tstride() const { return (size_t)((n + kNB - 1) / kNB) * kNB * kNB; }
    int alloc(int n_, int b_)
    {
        n = n_; b = b_;
        two_stage = n >= kTwoStageMinN && two_stage_fits(n);
        resident = n <= tbk::hetrd_resident_max_n(false);
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
    int alloc_deriv()
    {
        const size_t nn = (size_t)n * n * b;
        const int rc = dalloc(&dA, nn) | dalloc(&dS, nn);
        if (rc) { dfree(dA); dfree(dS); }
        return rc ? TBK_ERR_NOMEM : TBK_OK;
    }
    int alloc_mixed()
    {
        const size_t nn = (size_t)n * n * b;
        resident = n <= tbk::hetrd_resident_max_n(true);
        int rc = TBK_OK;
        rc |= dalloc(&A32, nn); rc |= dalloc(&VT32, nn); rc |= dalloc(&Y32, nn); rc |= dalloc(&X32, nn);
        rc |= dalloc(&P32, (size_t)n * 96 * b); rc |= dalloc(&T32, tstride() * b);
        rc |= dalloc(&lam, (size_t)n * b); rc |= dalloc(&rn, (size_t)n * b); rc |= dalloc(&tau, (size_t)b);
        rc |= dalloc(&perm, (size_t)n * b); rc |= dalloc(&perm2, (size_t)n * b); rc |= dalloc(&cinfo, (size_t)n * b);
        rc |= dalloc(&clist, (size_t)n * b); rc |= dalloc(&ncl, (size_t)b);
        return rc ? TBK_ERR_NOMEM : TBK_OK;
    }
    int alloc_wide(bool fp32)
    {
        wide = true;
        const size_t bn = (size_t)n * b, bt = (size_t)tmax() * n * b;
        int rc = TBK_OK;
        if (!fp32) {
            rc |= dalloc(&wv, bn); rc |= dalloc(&wcol, bn); rc |= dalloc(&wcarry, bn);
            rc |= dalloc(&wydot, bt); rc |= dalloc(&wyax, bt);
            rc |= dalloc(&wudot, (size_t)64 * b); rc |= dalloc(&wscal, (size_t)2 * b);
        } else {
            rc |= dalloc(&wv32, bn); rc |= dalloc(&wcol32, bn); rc |= dalloc(&wcarry32, bn); rc |= dalloc(&wydot32, bt);
            rc |= dalloc(&wyax32, bt); rc |= dalloc(&wudot32, (size_t)64 * b); rc |= dalloc(&wscal32, (size_t)2 * b);
        }
        return rc ? TBK_ERR_NOMEM : TBK_OK;
    }
    // The buffers of alloc(n_, b_) (FP64 path) laid out from base instead,
    // each on a 256-byte boundary; returns the bytes they take (base null:
    // measure only, the pointers stay null).
    size_t carve(unsigned char *base, int n_, int b_)
    {
        n = n_; b = b_;
        two_stage = n >= kTwoStageMinN && two_stage_fits(n);
        resident = n <= tbk::hetrd_resident_max_n(false);
        size_t off = 0;
        auto take = [&](auto *&p, size_t count) {
            typedef std::remove_reference_t<decltype(*p)> T;
            p = base ? reinterpret_cast<T *>(base + off) : nullptr;
            off += (count * sizeof(T) + 255) / 256 * 256;
        };
        const size_t nn = (size_t)n * n * b, np = (size_t)n * kNB * b;
        take(A, nn); take(S, nn); take(Y, nn); take(Z, nn); take(X, nn); take(VT, nn);
        take(Vp, (size_t)n * 96 * b); take(Wp, np); take(Tf, tstride() * b);
        take(px, (size_t)n * b); take(py, (size_t)n * b);
        take(d, (size_t)n * b); take(e, (size_t)n * b); take(w, (size_t)n * b);
        take(Vt, nn);
        take(scratch, (size_t)5 * n * tbk::stein_threads(n) * b);
        take(info, (size_t)b); take(sinfo, (size_t)b);
        return off;
    }
    tbk::WideArgsT<double> wide64() const
    {
        return tbk::WideArgsT<double>{tmax(), wv, wcol, wcarry, wydot, wyax, wudot, wscal};
    }
    tbk::WideArgsT<float> wide32() const
    {
        return tbk::WideArgsT<float>{tmax(), wv32, wcol32, wcarry32, wydot32, wyax32, wudot32, wscal32};
    }
    void release()
    {
        dfree(wv); dfree(wcol); dfree(wcarry); dfree(wydot); dfree(wyax); dfree(wudot); dfree(wscal);
        dfree(wv32); dfree(wcol32); dfree(wcarry32); dfree(wydot32); dfree(wyax32); dfree(wudot32); dfree(wscal32);
        dfree(A); dfree(S); dfree(Y); dfree(Z); dfree(X); dfree(VT); dfree(Vp); dfree(Wp); dfree(Tf);
        dfree(dA); dfree(dS);
        dfree(px); dfree(py); dfree(d); dfree(e); dfree(w); dfree(Vt); dfree(scratch);
        dfree(info); dfree(sinfo);
        dfree(A32); dfree(P32); dfree(T32); dfree(VT32); dfree(Y32); dfree(X32);
        dfree(lam); dfree(rn); dfree(tau); dfree(perm); dfree(perm2); dfree(cinfo); dfree(clist); dfree(ncl);
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

template <class G>
G with_ncol(G g, const int *ncol)
{
    g.ncol = ncol;
    return g;
}

// Occupation window of a run (all null/0 when off): the per-structure Fermi
// arguments from which k_fermi_window sets cut[k], the number of columns
// (eigenvectors) each item needs.
struct Window {
    const tbk::FermiArgs *fargs;
    int nstruct;
    int *cut;
    bool vectors = true;           // false: no eigenvector is formed (energy-only modes)
};

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
void hetrd(Workspace &ws, int nb, hipStream_t st, StageTimer &tm)
{
    const int n = ws.n;
    const size_t ss = ws.stride(), sp = (size_t)n * 96;
    tbk::HetrdArgs h{n, ss, ws.A, ws.Vp, ws.Wp, ws.Tf, ws.VT, ws.d, ws.e, ws.tstride()};
    if (ws.resident) {
        tm.begin("hetrd_resident");
        tbk::launch_hetrd_resident(h, nb, st);
        tm.end();
        return;
    }
    for (int i0 = 0; i0 < n - 1; i0 += kNB) {
        const int pw = std::min(kNB, n - 1 - i0);
        tm.begin("hetrd_panel");
        if (ws.wide) {
            for (int i = 0; i < pw; ++i) tbk::launch_hetrd_wide_column(h, ws.wide64(), i0, i, nb, st);
            tbk::launch_hetrd_wide_vt(h, i0, nb, st);
        } else {
            tbk::launch_hetrd_panel(h, i0, nb, st);
        }
        tm.end();
        const int s = i0 + pw, m = n - s;
        tm.begin("hetrd_update");
        if (m > 0)
            tbk::launch_zgemm(gemm(m, m, 2 * kNB, ws.Vp + s, n, sp, ws.Vp + (size_t)kNB * n + s, n, sp,
                                   ws.A + (size_t)s * n + s, n, ss, 4, double2{-1.0, 0.0},
                                   double2{1.0, 0.0}), 0, 1, nb, st);
        tm.end();
    }
    tm.begin("hetrd_update");
    tbk::launch_hetrd_finish(h, nb, st);
    tm.end();
}

// S = U^H U and U^{-1} in place (lower triangle cleared), recursively over
// the diagonal blocks of panels p0 .. p1-1 (kNB columns each): the leading
// half's factor and inverse, U12 = U11^{-H} S12 (into Y), S22 -= U12^H U12,
// the trailing half's, then (U^{-1})_12 = -U11^{-1} U12 U22^{-1} (through Z);
// a single panel is one panel kernel.  The GEMMs of the top levels have
// inner dimensions of half the order, so the trailing matrix is read and
// written once per level rather than once per panel.  Y and Z are scratch.
void potrf_rec(Workspace &ws, int p0, int p1, int nb, hipStream_t st)
{
    const int n = ws.n;
    const size_t ss = ws.stride();
    if (p1 - p0 == 1) {
        tbk::launch_potrf_panel(n, ws.S, ss, p0 * kNB, ws.info, nb, st);   // S[J, J] <- U_JJ^{-1}
        return;
    }
    const int pm = p0 + (p1 - p0) / 2;
    const int a = p0 * kNB, m = pm * kNB, b = std::min(p1 * kNB, n), n1 = m - a, n2 = b - m;
    const size_t o11 = (size_t)a * n + a, o12 = (size_t)m * n + a, o22 = (size_t)m * n + m;
    potrf_rec(ws, p0, pm, nb, st);
    tbk::launch_zgemm(gemm(n1, n2, n1, ws.S + o11, n, ss, ws.S + o12, n, ss, ws.Y + o12, n, ss, 2), 1, 0, nb, st);
    // on the tiles touching the upper triangle (the factorisation reads only
    // that; the panel kernels clear below)
    tbk::launch_zgemm(gemm(n2, n2, n1, ws.Y + o12, n, ss, ws.Y + o12, n, ss, ws.S + o22, n, ss, 7,
                           double2{-1.0, 0.0}, double2{1.0, 0.0}), 1, 0, nb, st);
    potrf_rec(ws, pm, p1, nb, st);
    tbk::launch_zgemm(gemm(n1, n2, n1, ws.S + o11, n, ss, ws.Y + o12, n, ss, ws.Z + o12, n, ss, 1), 0, 0, nb, st);
    tbk::launch_zgemm(gemm(n1, n2, n2, ws.Z + o12, n, ss, ws.S + o22, n, ss, ws.S + o12, n, ss, 3,
                           double2{-1.0, 0.0}), 0, 0, nb, st);
}

void potrf_trtri(Workspace &ws, int nb, hipStream_t st)
{
    potrf_rec(ws, 0, (ws.n + kNB - 1) / kNB, nb, st);
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

template <class R>
tbk::GemmArgsT<R> gemm_t(int m, int n, int k, const typename tbk::cplx_t<R>::type *a, int lda, size_t sa,
                         const typename tbk::cplx_t<R>::type *b, int ldb, size_t sb, typename tbk::cplx_t<R>::type *c,
                         int ldc, size_t sc, R alpha_re = 1, R beta_re = 0)
{
    tbk::GemmArgsT<R> g;
    g.m = m; g.n = n; g.k = k; g.tri = 0;
    g.alpha.x = alpha_re; g.alpha.y = 0; g.beta.x = beta_re; g.beta.y = 0;
    g.a = a; g.lda = lda; g.sa = sa; g.b = b; g.ldb = ldb; g.sb = sb; g.c = c; g.ldc = ldc; g.sc = sc;
    return g;
}
inline void launch_gemm_any(const tbk::GemmArgsT<double> &g, int opa, int opb, int nb, hipStream_t st)
{
    tbk::launch_zgemm(g, opa, opb, nb, st);
}
inline void launch_gemm_any(const tbk::GemmArgsT<float> &g, int opa, int opb, int nb, hipStream_t st)
{
    tbk::launch_cgemm(g, opa, opb, nb, st);
}

// Panels of a Householder reduction applied to Y (Y <- Q Y, Q = Q_0 Q_1
// ... with Q_p = I - V_p T_p V_p^H), `group` panels at a time.  Within a
// group the panels' V T^H blocks are first merged from the last panel down,
//   VT_p <- VT_p - VT_rest (V_rest^H VT_p),
// which turns [VT_p | VT_rest] into V_c T_c^H of the group's combined WY
// factor; each group is then two GEMMs of inner dimension up to pw group.
// A holds the reflectors (panel p covers columns p pw .. p pw + pw - 1, its
// reflector of column c non-zero from row c + off down), VT the V T^H
// blocks at the same place; M is scratch of 96 x 32 per item (stride sm).
struct QPanels {
    int pw, off, np, group;        // panel width, first row of a reflector below its column, panels, merge
};
constexpr int kBT = 3;
inline QPanels one_stage_panels(int n) { return QPanels{kNB, 1, (n - 2) / kNB + 1, kBT}; }
// stage 1 of the two-stage reduction: panels c0 = p TB while c0 + TB + 1 < n
inline int band_panels(int n) { return n >= TB + 2 ? (n - TB - 2) / TB + 1 : 0; }
inline QPanels band_stage_panels(int n) { return QPanels{TB, TB, band_panels(n), 6}; }
template <class R>
void apply_q(int n, int nb, typename tbk::cplx_t<R>::type *A, typename tbk::cplx_t<R>::type *VT,
             typename tbk::cplx_t<R>::type *Y, typename tbk::cplx_t<R>::type *X, typename tbk::cplx_t<R>::type *M,
             size_t sm, hipStream_t st, const int *ncol, typename tbk::GemmArgsT<R>::CX *ycx, const QPanels &qp)
{
    const size_t ss = (size_t)n * n;
    const int P = qp.np, pw = qp.pw, off = qp.off;
    if (P <= 0) return;
    for (int g0 = ((P - 1) / qp.group) * qp.group; g0 >= 0; g0 -= qp.group) {
        const int g1 = std::min(g0 + qp.group, P);     // panels g0 .. g1-1
        const int c0 = g0 * pw;
        const int wlast = std::min(pw, n - off - (g1 - 1) * pw);
        const int W = (g1 - 1 - g0) * pw + wlast;      // columns of the group
        // Column c of V (and of V T^H, a combination of the reflectors c and
        // later) is zero above row c + off: with rows counted from c0 + off,
        // op(A) of the first product of each pair below is upper triangular
        // (tri 1: k below the row tile skipped) and V's block of the second
        // lower trapezoidal (tri 2: k beyond the row tile skipped)
        for (int p = g1 - 2; p >= g0; --p) {            // merge V T^H blocks
            const int cp = p * pw, cr = cp + pw, wr = c0 + W - cr;
            const int rr0 = cr + off, k = n - rr0;       // rows where V_rest can be non-zero
            tbk::GemmArgsT<R> g1a = gemm_t<R>(wr, pw, k, A + (size_t)cr * n + rr0, n, ss, VT + (size_t)cp * n + rr0,
                                              n, ss, M, wr, sm);
            g1a.tri = 1;
            launch_gemm_any(g1a, 1, 0, nb, st);
            tbk::GemmArgsT<R> g1b = gemm_t<R>(k, pw, wr, VT + (size_t)cr * n + rr0, n, ss, M, wr, sm,
                                              VT + (size_t)cp * n + rr0, n, ss, R(-1), R(1));
            g1b.tri = 2;
            launch_gemm_any(g1b, 0, 0, nb, st);
        }
        const int r0 = c0 + off, m = n - r0;
        // X[0:W, :] = (V_c T_c^H)^H Y[r0:, :];  Y[r0:, :] -= V_c X
        tbk::GemmArgsT<R> gx = with_ncol(gemm_t<R>(W, n, m, VT + (size_t)c0 * n + r0, n, ss, Y + r0, n, ss, X, n, ss),
                                         ncol);
        gx.tri = 1;
        launch_gemm_any(gx, 1, 0, nb, st);
        tbk::GemmArgsT<R> g = with_ncol(gemm_t<R>(m, n, W, A + (size_t)c0 * n + r0, n, ss, X, n, ss, Y + r0, n, ss,
                                                  R(-1), R(1)), ncol);
        g.tri = 2;
        if (g0 == 0 && ycx) {          // the last group: Y rows r0.. also in the other precision
            g.cx = ycx + r0; g.ldcx = n; g.scx = ss;
        }
        launch_gemm_any(g, 0, 0, nb, st);
    }
}

// Stage 1 of the two-stage tridiagonalisation on the batch: A (both
// triangles) to a lower band of width TB, panel by panel (k_band_qr, then
// V T, V T^H and the two-sided trailing update as GEMMs), the band into
// `band` (item stride sb).  P: [V | Y/Z | V | V T] columns (stride sp);
// scr: 2 TB^2 per item (stride sscr).
template <class R>
void band_reduce(int n, int nb, typename tbk::cplx_t<R>::type *A, typename tbk::cplx_t<R>::type *P, size_t sp,
                 typename tbk::cplx_t<R>::type *T, size_t tstride, typename tbk::cplx_t<R>::type *VT,
                 typename tbk::cplx_t<R>::type *band, size_t sb, typename tbk::cplx_t<R>::type *scr, size_t sscr,
                 hipStream_t st)
{
    typedef typename tbk::cplx_t<R>::type C;
    const size_t ss = (size_t)n * n;
    tbk::BandArgsT<R> b{n, ss, A, VT, P, sp, T, tstride, band, sb};
    const int np = band_panels(n);
    for (int p = 0; p < np; ++p) {
        const int c0 = p * TB, m = n - c0 - TB, s = c0 + TB;
        tbk::launch_band_qr(b, p, nb, st);
        C *V = P + s, *Y = P + (size_t)TB * n + s, *W0 = P + (size_t)3 * TB * n + s;
        C *Tp = T + (size_t)p * TB * TB, *M = scr, *X = scr + TB * TB;
        // W0 = V T and, for the back-transform, V T^H
        launch_gemm_any(gemm_t<R>(m, TB, TB, V, n, sp, Tp, TB, tstride, W0, n, sp), 0, 0, nb, st);
        launch_gemm_any(gemm_t<R>(m, TB, TB, V, n, sp, Tp, TB, tstride, VT + (size_t)c0 * n + s, n, ss), 0, 1, nb,
                        st);
        // Y = A22 V T;  Z = Y - 1/2 V (T^H (V^H Y))  (in place)
        launch_gemm_any(gemm_t<R>(m, TB, m, A + (size_t)s * n + s, n, ss, W0, n, sp, Y, n, sp), 0, 0, nb, st);
        launch_gemm_any(gemm_t<R>(TB, TB, m, V, n, sp, Y, n, sp, M, TB, sscr), 1, 0, nb, st);
        launch_gemm_any(gemm_t<R>(TB, TB, TB, Tp, TB, tstride, M, TB, sscr, X, TB, sscr), 1, 0, nb, st);
        launch_gemm_any(gemm_t<R>(m, TB, TB, V, n, sp, X, TB, sscr, Y, n, sp, R(-0.5), R(1)), 0, 0, nb, st);
        // A22 -= V Z^H + Z V^H, both triangles
        tbk::GemmArgsT<R> g = gemm_t<R>(m, m, 2 * TB, V, n, sp, Y, n, sp, A + (size_t)s * n + s, n, ss, R(-1), R(1));
        g.tri = 8;
        launch_gemm_any(g, 0, 1, nb, st);
    }
    tbk::launch_band_tail(b, np * TB, nb, st);
}

// Mixed-precision solve: FP64 Cholesky, inverse and reduction; FP32
// tridiagonalisation and back-transform on the FP32 MFMA units; FP64
// bisection/eigenvectors of the tridiagonal; one FP64 refinement step on the
// reduced problem C (Ogita-Aishima, with clusters rotated exactly and the
// eigenvalues to second order); FP64 Z = U^{-1} Y.
// Bisection tolerance of the mixed path, relative to ||T||_1: the
// tridiagonal of the FP32 reduction carries FP32 errors, and its eigenvalues
// only seed the vectors that the FP64 refinement corrects.
constexpr double kMixedBisectTol = 1e-11;
// FP64 path: dstebz's default absolute tolerance eps ||T||_1 as a floor
// (the reduction to T leaves eigenvalue errors of that order anyway), so
// that eigenvalues near zero stop with the rest.
constexpr double kFp64BisectTol = 2.220446049250313e-16;

void solve_batch_mixed(Workspace &ws, int nb, const int *nact, double *w_out, hipStream_t st, StageTimer &tm,
                       const Window &wnd)
{
    const int n = ws.n;
    const size_t ss = ws.stride(), sp = (size_t)n * 96;
    tm.begin("potrf_trtri");
    potrf_trtri(ws, nb, st);
    tm.end();
    tm.begin("reduce_gemm");
    tbk::launch_zgemm(gemm(n, n, n, ws.A, n, ss, ws.S, n, ss, ws.Y, n, ss, 3), 0, 0, nb, st);
    {   // C in FP64 (for the refinement) and, from the same epilogue, in FP32
        tbk::GemmArgs g = gemm(n, n, n, ws.S, n, ss, ws.Y, n, ss, ws.A, n, ss, 6);
        g.cx = ws.A32; g.ldcx = n; g.scx = ss;
        tbk::launch_zgemm(g, 1, 0, nb, st);
    }
    tm.end();
    if (ws.two_stage) {
        // band in Y32, stage-2 reflectors in X32 (scratch of stage 1 before)
        tm.begin("hetrd_band");
        band_reduce<float>(n, nb, ws.A32, ws.P32, sp, ws.T32, ws.tstride(), ws.VT32, ws.Y32, ss, ws.X32, ss, st);
        tm.end();
        tm.begin("hetrd_bulge");
        tbk::launch_bulge(tbk::BulgeArgsT<float>{n, ws.Y32, ss, ws.X32, ss, ws.d, ws.e, (size_t)n, ws.px, (size_t)n},
                          nb, st);
        tm.end();
    } else if (ws.resident) {
        tbk::HetrdArgsT<float> h{n, ss, ws.A32, ws.P32, nullptr, ws.T32, ws.VT32, ws.d, ws.e, ws.tstride()};
        tm.begin("hetrd_resident");
        tbk::launch_hetrd_resident(h, nb, st);
        tm.end();
    } else {
        tbk::HetrdArgsT<float> h{n, ss, ws.A32, ws.P32, nullptr, ws.T32, ws.VT32, ws.d, ws.e, ws.tstride()};
        for (int i0 = 0; i0 < n - 1; i0 += kNB) {
            const int pw = std::min(kNB, n - 1 - i0);
            tm.begin("hetrd_panel");
            if (ws.wide) {
                for (int i = 0; i < pw; ++i) tbk::launch_hetrd_wide_column(h, ws.wide32(), i0, i, nb, st);
                tbk::launch_hetrd_wide_vt(h, i0, nb, st);
            } else {
                tbk::launch_hetrd_panel(h, i0, nb, st);
            }
            tm.end();
            const int s = i0 + pw, m = n - s;
            tm.begin("hetrd_update");
            if (m > 0)
                tbk::launch_cgemm(cgemm(m, m, 2 * kNB, ws.P32 + s, n, sp, ws.P32 + (size_t)kNB * n + s, n, sp,
                                        ws.A32 + (size_t)s * n + s, n, ss, 4, float2{-1.f, 0.f},
                                        float2{1.f, 0.f}), 0, 1, nb, st);
            tm.end();
        }
        tm.begin("hetrd_update");
        tbk::launch_hetrd_finish(h, nb, st);
        tm.end();
    }
    tm.begin("stein");
    tbk::SteinArgs sa{n, kOrtolRel, nact, ws.d, ws.e, ws.w, ws.Vt, ws.scratch, ws.sinfo};
    sa.abstol_rel = kMixedBisectTol;
    tbk::launch_stein(sa, nb, st);
    tm.end();
    tm.begin("backtransform");
    // Y in FP32, back-transformed; its FP64 copy for the refinement comes
    // from the conversion (row 0, which no reflector touches) and from the
    // last back-transform GEMM's epilogue (rows 1..)
    if (ws.two_stage) {
        // rows of the real tridiagonal's vectors times the phases, then the
        // stage-2 reflectors (T factors in Z's storage, unused until the
        // refinement), FP64 copies of rows 1 .. TB-1 (no stage-1 reflector
        // reaches them), then the stage-1 panels
        tbk::launch_real_to_c32_t(n, ws.Vt, ws.Y32, ss, nb, st, ws.Y, ws.px);
        tbk::launch_q2(tbk::Q2ArgsT<float>{n, ws.Y32, ss, ws.X32, ss, reinterpret_cast<float2 *>(ws.Z), ss, nullptr},
                       nb, st);
        tbk::launch_rows_c32_to_c64(n, 1, std::min(TB, n), ws.Y32, ws.Y, ss, nb, st);
        apply_q<float>(n, nb, ws.A32, ws.VT32, ws.Y32, ws.X32, ws.P32, sp, st, nullptr, ws.Y, band_stage_panels(n));
    } else {
        tbk::launch_real_to_c32_t(n, ws.Vt, ws.Y32, ss, nb, st, ws.Y);
        apply_q<float>(n, nb, ws.A32, ws.VT32, ws.Y32, ws.X32, ws.P32, sp, st, nullptr, ws.Y, one_stage_panels(n));
    }
    tm.end();
    tm.begin("refine");
    // X = C Y in FP64; Rayleigh quotients, residuals (FP32, in X32) and
    // rotated clusters; K = Y^H R in FP32 (in A32, free after the
    // reduction); the eigenvalues into w_out
    tbk::launch_zgemm(gemm(n, n, n, ws.A, n, ss, ws.Y, n, ss, ws.X, n, ss), 0, 0, nb, st);
    const tbk::OaArgs oa{n, ss, nact, nullptr, ws.w, ws.X, ws.Y, ws.Y32, ws.X32, ws.A32, ws.VT32, ws.X32, ws.VT,
                         ws.lam, ws.rn, ws.tau, w_out, ws.perm, ws.perm2, ws.cinfo, ws.clist, ws.ncl};
    tbk::launch_oa_prepare(oa, nb, st);
    tbk::launch_cgemm(cgemm(n, n, n, ws.Y32, n, ss, ws.X32, n, ss, ws.A32, n, ss), 1, 0, nb, st);
    tbk::launch_oa_eigen(oa, nb, st);
    // the eigenvalues are final: occupation window, then only the columns
    // it keeps of E, D = Y E (FP32, in X32), the refined Y' (in VT) and
    // Z = U^{-1} Y'
    if (wnd.cut) tbk::launch_fermi_window_group(wnd.fargs, wnd.nstruct, st);
    if (wnd.vectors) {
        tbk::launch_oa_emat(oa, nb, st, wnd.cut);
        tbk::launch_cgemm(with_ncol(cgemm(n, n, n, ws.Y32, n, ss, ws.VT32, n, ss, ws.X32, n, ss), wnd.cut), 0, 0,
                          nb, st);
        tbk::launch_oa_final(oa, nb, st, wnd.cut);
        tbk::launch_zgemm(with_ncol(gemm(n, n, n, ws.S, n, ss, ws.VT, n, ss, ws.Z, n, ss, 1), wnd.cut), 0, 0, nb, st);
    }
    tm.end();
}

// Stages 2-9 of the pipeline on ws.A = H, ws.S = S for nb k-points:
// eigenvalues into w_out[nb*n], eigenvectors into ws.Z.
void solve_batch(Workspace &ws, int nb, const int *nact, double *w_out, hipStream_t st, StageTimer &tm,
                 const Window &wnd)
{
    const int n = ws.n;
    const size_t ss = ws.stride();
    tm.begin("potrf_trtri");
    potrf_trtri(ws, nb, st);
    tm.end();
    tm.begin("reduce_gemm");
    // Y = H Uinv (Uinv upper) ; A = Uinv^H Y
    tbk::launch_zgemm(gemm(n, n, n, ws.A, n, ss, ws.S, n, ss, ws.Y, n, ss, 3), 0, 0, nb, st);
    tbk::launch_zgemm(gemm(n, n, n, ws.S, n, ss, ws.Y, n, ss, ws.A, n, ss, 6), 1, 0, nb, st);
    tm.end();
    if (ws.two_stage) {
        // band in Y, stage-2 reflectors in Z (both free until the back-transform)
        const size_t sp = (size_t)n * 96;
        tm.begin("hetrd_band");
        band_reduce<double>(n, nb, ws.A, ws.Vp, sp, ws.Tf, ws.tstride(), ws.VT, ws.Y, ss, ws.Wp, (size_t)n * kNB,
                            st);
        tm.end();
        tm.begin("hetrd_bulge");
        tbk::launch_bulge(tbk::BulgeArgsT<double>{n, ws.Y, ss, ws.Z, ss, ws.d, ws.e, (size_t)n, ws.px, (size_t)n},
                          nb, st);
        tm.end();
    } else {
        hetrd(ws, nb, st, tm);
    }
    tm.begin("stein");
    tbk::SteinArgs sa{n, kOrtolRel, nact, ws.d, ws.e, w_out, ws.Vt, ws.scratch, ws.sinfo};
    sa.abstol_rel = kFp64BisectTol;
    tbk::launch_stebz(sa, nb, st);
    // the eigenvalues are final: occupation window, then only the vectors it keeps
    if (wnd.cut) tbk::launch_fermi_window_group(wnd.fargs, wnd.nstruct, st);
    if (!wnd.vectors) { tm.end(); return; }
    sa.cut = wnd.cut;
    tbk::launch_stein_vectors(sa, nb, st);
    tm.end();
    tm.begin("backtransform");
    if (ws.two_stage) {
        // phases, the stage-2 reflectors (T factors in X, scratch of the
        // panels' back-transform afterwards), the stage-1 panels
        tbk::launch_real_to_complex_t(n, ws.Vt, ws.Y, ss, nb, st, ws.px);
        tbk::Q2ArgsT<double> q{n, ws.Y, ss, ws.Z, ss, ws.X, ss};
        q.ncol = wnd.cut;
        tbk::launch_q2(q, nb, st);
        apply_q<double>(n, nb, ws.A, ws.VT, ws.Y, ws.X, ws.Vp, (size_t)n * 96, st, wnd.cut, nullptr,
                        band_stage_panels(n));
    } else {
        tbk::launch_real_to_complex_t(n, ws.Vt, ws.Y, ss, nb, st);
        apply_q<double>(n, nb, ws.A, ws.VT, ws.Y, ws.X, ws.Vp, (size_t)n * 96, st, wnd.cut, nullptr,
                        one_stage_panels(n));
    }
    // Z = Uinv Y
    tbk::launch_zgemm(with_ncol(gemm(n, n, n, ws.S, n, ss, ws.Y, n, ss, ws.Z, n, ss, 1), wnd.cut), 0, 0, nb, st);
    tm.end();
}

}  // namespace

// Host view of one structure of a plan.
struct StructHost {
    // nv: states per k (2 norb for a non-collinear structure), norb: orbitals
    int natoms = 0, nv = 0, npairs = 0, nk = 0, k0 = 0, norb = 0;   // k0: first item
    double zfill = 0.0, thisvol = 0.0;
};

struct tbk_plan {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = true;         // false: borrowed from the caller, not destroyed
    int nstruct = 0, N = 0, nk = 0, batch = 0, precision = TBK_FP64, maxnat = 0, maxpairs = 0;
    std::vector<StructHost> sh;
    std::vector<double> kpts, weights;
    std::vector<int> item_s_host;
    // concatenated set-up arrays of all structures
    int *row_atom = nullptr, *row_orb = nullptr, *ksk = nullptr, *nbas = nullptr, *blk_ptr = nullptr;
    double *dvec = nullptr, *eh = nullptr, *es = nullptr, *edh = nullptr, *eds = nullptr;
    double *paros = nullptr, *dparos = nullptr, *kp = nullptr, *wts = nullptr;
    tbk::StructDev *sdev = nullptr;        // [nstruct]
    tbk::FermiArgs *fargs = nullptr;       // [nfermi]
    // spin: a collinear spin-polarised structure, structure 0 the spin-up
    // and structure 1 the spin-down channel, one Fermi level over both
    bool spin = false;
    int nfermi = 0;                        // Fermi groups: nstruct, or 1 with spin
    // noncol: non-collinear structures (2 norb spinor items, one electron
    // per state); spin_terms their per-orbital terms on the device
    bool noncol = false;
    int maxnorb = 0;
    double *spin_terms = nullptr;
    int *item_s = nullptr, *nact = nullptr; // [nk]
    // results, per item with stride N
    double *esk = nullptr, *desk = nullptr, *out = nullptr;   // out [nstruct][8]
    double2 *ph = nullptr;          // Bloch phases of the current batch
    int *ncut = nullptr, *err = nullptr, *info_all = nullptr;
    int *cut = nullptr;             // [nk] occupation window of the last windowed run
    bool window = true, win_last = false, vec_last = true;
    // pressure modes: H, S, dH/dV, dS/dV in one build pass (ws.dA, ws.dS,
    // allocated on the first pressure run) instead of dH, dS after the
    // solve; off by default: at order 288 the one pass takes 1.5 ms against
    // 2 x 0.6 ms for the two
    bool fused_build = false;
    // spin directions of non-collinear structures (bande.f:1484-1561):
    // spev [nk][N][maxnat][3] per state, spdir [nstruct][maxnat][3]
    bool spdir = false;
    double *spev = nullptr, *spdir_out = nullptr;
    std::vector<int> info_host, err_host, status;
    std::vector<double> out_host;
    Workspace ws;
    // Lowdin projections (QLMT modes 6, 10): a second workspace for the
    // eigendecomposition of S, and |(S^{1/2} z_j)_i|^2 of every item,
    // proj[k][j*N + i]
    Workspace ws2;
    double *proj = nullptr;
    StageTimer timer;
    int last_mode = 0;
    size_t bytes = 0;
    // hipGraph of the launch sequence (replayed while mode, tkb, eigcut are unchanged)
    bool use_graph = false, recording = false;
    hipGraphExec_t gexec = nullptr;
    int g_mode = 0;
    double g_tkb = 0.0, g_eigcut = 0.0;
    bool g_window = true;
    int f_mode = 0;                 // (mode, kT) of the device Fermi arguments
    double f_tkb = 0.0;
};

static void plan_free(tbk_plan *p)
{
    if (!p) return;
    (void)hipSetDevice(p->device);
    AllocStream as(p->stream);
    if (p->gexec) (void)hipGraphExecDestroy(p->gexec);
    dfree(p->row_atom); dfree(p->row_orb); dfree(p->ksk); dfree(p->nbas); dfree(p->blk_ptr);
    dfree(p->dvec); dfree(p->eh); dfree(p->es); dfree(p->edh); dfree(p->eds);
    dfree(p->paros); dfree(p->dparos); dfree(p->kp); dfree(p->wts);
    dfree(p->sdev); dfree(p->fargs); dfree(p->item_s); dfree(p->nact); dfree(p->spin_terms);
    dfree(p->esk); dfree(p->desk); dfree(p->out); dfree(p->ncut); dfree(p->err); dfree(p->info_all);
    dfree(p->cut);
    dfree(p->ph);
    dfree(p->proj);
    dfree(p->spev); dfree(p->spdir_out);
    p->ws.release();
    p->ws2.release();
    if (p->stream) {
        (void)hipStreamSynchronize(p->stream);
        if (p->own_stream) (void)hipStreamDestroy(p->stream);
    }
    delete p;
}

// Allocate and copy on the allocation stream (the caller synchronises it
// before the host vectors go away).
template <class T>
static int upload(T **dst, const std::vector<T> &src)
{
    int rc = dalloc(dst, std::max<size_t>(src.size(), 1));
    if (rc) return rc;
    if (!src.empty() && hipMemcpyAsync(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice,
                                       g_alloc_stream) != hipSuccess)
        return TBK_ERR_HIP;
    return TBK_OK;
}

// Host set-up of one structure (bit-exact search2 / setpar / setparv and the
// SK blocks), appended to the concatenated arrays.
struct SetupArrays {
    std::vector<int> row_atom, row_orb, ksk, nbas, blk_ptr;
    std::vector<double> dvec, eh, es, edh, eds, paros, dparos, spin;
};
struct SetupOffsets { size_t rows, atoms, blk, pairs, spin; bool noncol; };

// Spin channel selection of a structure's set-up: atoms with flip[a] != 0
// take their densities and on-site terms from param_flip.
struct SpinSel {
    const double *param_flip = nullptr; const int *flip = nullptr;
    // non-collinear (lnoncol): the spin-down channel and the atoms' spin
    // directions [3*natoms]
    const double *nc_param_dn = nullptr, *nc_spinat = nullptr;
};

static int setup_structure(const tbk_structure_desc *st, SetupArrays &A, StructHost &h, SetupOffsets &off,
                           const SpinSel &sel = SpinSel())
{
    if (!st || st->natoms <= 0) return TBK_ERR_ARG;
    const int nat = st->natoms, kinds = st->kinds;
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
                paros.data(), hop.data(), over.data(), sel.param_flip, sel.flip);
    tbk::setparv(nat, st->kkind, kinds, st->param, st->nltype, st->jsover, st->thisvol, pl,
                 den.data(), dparos.data(), dhop.data(), dover.data(), sel.param_flip, sel.flip);
    std::vector<int> ksk(nat), nbas(nat);
    int nv = 0;
    for (int a = 0; a < nat; ++a) {
        ksk[a] = nv;
        nbas[a] = st->kbas[st->kkind[a] - 1];
        if (nbas[a] < 1 || nbas[a] > 9) return TBK_ERR_ARG;
        nv += nbas[a];
    }
    const bool noncol = sel.nc_param_dn != nullptr;
    if (nv < 1 || (noncol ? 2 * nv : nv) > kMaxN) return TBK_ERR_ARG;
    if (noncol) {
        // bande.f:859-905 with lnoncol: integrals from channel 1, on-site
        // average and spin terms from both (setpar.f:466-485, 744-795),
        // then the per-orbital spinor terms (spinham.f:110-150)
        std::vector<double> den2(2 * (size_t)nat * kinds), spinos(4 * nat), dspnos(4 * nat);
        tbk::setpar_noncol(nat, st->kkind, kinds, st->param, sel.nc_param_dn, pl, den2.data(), paros.data(),
                           spinos.data());
        tbk::setparv_noncol(nat, st->kkind, kinds, st->param, sel.nc_param_dn, st->thisvol, pl, den2.data(),
                            dparos.data(), dspnos.data());
        std::vector<double> t(6 * (size_t)nv);
        tbk::spinor_terms(nat, nbas.data(), spinos.data(), sel.nc_spinat, &t[0], &t[nv], &t[2 * nv]);
        tbk::spinor_terms(nat, nbas.data(), dspnos.data(), sel.nc_spinat, &t[3 * nv], &t[4 * nv], &t[5 * nv]);
        off.spin = A.spin.size();
        A.spin.insert(A.spin.end(), t.begin(), t.end());
    }
    off.noncol = noncol;
    off.rows = A.row_atom.size(); off.atoms = A.ksk.size(); off.blk = A.blk_ptr.size();
    off.pairs = A.dvec.size() / 3;
    for (int a = 0; a < nat; ++a)
        for (int o = 0; o < nbas[a]; ++o) { A.row_atom.push_back(a); A.row_orb.push_back(o); }
    A.ksk.insert(A.ksk.end(), ksk.begin(), ksk.end());
    A.nbas.insert(A.nbas.end(), nbas.begin(), nbas.end());
    A.paros.insert(A.paros.end(), paros.begin(), paros.end());
    A.dparos.insert(A.dparos.end(), dparos.begin(), dparos.end());
    std::vector<int> blk((size_t)nat * nat + 1, 0);
    const size_t e0 = A.eh.size(), d0 = A.dvec.size();
    A.eh.resize(e0 + 81 * (size_t)np); A.es.resize(e0 + 81 * (size_t)np);
    A.edh.resize(e0 + 81 * (size_t)np); A.eds.resize(e0 + 81 * (size_t)np);
    A.dvec.resize(d0 + 3 * (size_t)np);
    // SK blocks per pair (independent: chunks of pairs on host threads)
    constexpr int kChunk = 256;
    tbk::parallel_for((np + kChunk - 1) / kChunk, [&](int ch) {
        const int pe = std::min(np, (ch + 1) * kChunk);
        for (int p = ch * kChunk; p < pe; ++p) {
            const tbk::Pair &q = pl.pairs[p];
            double v[14];
            const bool unlike = q.kind[0] != q.kind[1];
            const size_t o = e0 + 81 * (size_t)p;
            tbk::arrange_integrals(&hop[14 * p], q.kind[0], q.kind[1], v);
            tbk::sk_block(q.cosines[0], q.cosines[1], q.cosines[2], v, unlike, &A.eh[o]);
            tbk::arrange_integrals(&over[14 * p], q.kind[0], q.kind[1], v);
            tbk::sk_block(q.cosines[0], q.cosines[1], q.cosines[2], v, unlike, &A.es[o]);
            tbk::arrange_integrals(&dhop[14 * p], q.kind[0], q.kind[1], v);
            tbk::sk_block(q.cosines[0], q.cosines[1], q.cosines[2], v, unlike, &A.edh[o]);
            tbk::arrange_integrals(&dover[14 * p], q.kind[0], q.kind[1], v);
            tbk::sk_block(q.cosines[0], q.cosines[1], q.cosines[2], v, unlike, &A.eds[o]);
            const int i = q.atom[0], j = q.atom[1];
            for (int c = 0; c < 3; ++c)    // rotate.f:144-146 operand order
                A.dvec[d0 + 3 * (size_t)p + c] = q.dlv[c] - st->posq[3 * j + c] + st->posq[3 * i + c];
        }
    });
    for (int p = 0; p < np; ++p) blk[(size_t)pl.pairs[p].atom[0] * nat + pl.pairs[p].atom[1] + 1]++;
    for (size_t b = 1; b < blk.size(); ++b) blk[b] += blk[b - 1];
    A.blk_ptr.insert(A.blk_ptr.end(), blk.begin(), blk.end());
    h.natoms = nat; h.norb = nv; h.nv = noncol ? 2 * nv : nv; h.npairs = np;
    h.zfill = st->zfill; h.thisvol = st->thisvol;
    return TBK_OK;
}

// Appends one structure's set-up arrays (built from empty) to the plan's.
static void append_setup(SetupArrays &A, const SetupArrays &L, SetupOffsets &off)
{
    off.rows = A.row_atom.size(); off.atoms = A.ksk.size(); off.blk = A.blk_ptr.size();
    off.pairs = A.dvec.size() / 3;
    off.spin = A.spin.size();
    auto cat = [](auto &dst, const auto &src) { dst.insert(dst.end(), src.begin(), src.end()); };
    cat(A.spin, L.spin);
    cat(A.row_atom, L.row_atom); cat(A.row_orb, L.row_orb); cat(A.ksk, L.ksk); cat(A.nbas, L.nbas);
    cat(A.blk_ptr, L.blk_ptr); cat(A.dvec, L.dvec); cat(A.eh, L.eh); cat(A.es, L.es);
    cat(A.edh, L.edh); cat(A.eds, L.eds); cat(A.paros, L.paros); cat(A.dparos, L.dparos);
}

// Host set-up of one structure, built ahead of the plan that uploads it
// (the request server prepares structures on host threads while the device
// works on earlier groups).
struct tbk_prepared {
    SetupArrays a;
    StructHost sh;
    SetupOffsets off{};
};

int tbk_prepare(const tbk_structure_desc *st, tbk_prepared **out)
{
    if (!out) return TBK_ERR_ARG;
    *out = nullptr;
    tbk_prepared *p = new tbk_prepared();
    const int rc = setup_structure(st, p->a, p->sh, p->off);
    if (rc) { delete p; return rc; }
    *out = p;
    return TBK_OK;
}

void tbk_prepared_free(tbk_prepared *p) { delete p; }

int tbk_prepared_nv(const tbk_prepared *p) { return p ? p->sh.nv : 0; }

size_t tbk_prepared_bytes(const tbk_prepared *p)
{
    if (!p) return 0;
    const SetupArrays &a = p->a;
    return 4 * (a.row_atom.size() + a.row_orb.size() + a.ksk.size() + a.nbas.size() + a.blk_ptr.size()) +
           8 * (a.dvec.size() + a.eh.size() + a.es.size() + a.edh.size() + a.eds.size() + a.paros.size() +
                a.dparos.size() + a.spin.size());
}

// The device part of plan creation from prepared structures: concatenated
// set-up arrays, per-item results and the batch workspace.  stream: the
// plan's stream, borrowed (not destroyed with the plan), or null for a
// stream of its own.  The matrix order is the largest structure's, raised to
// min_order where that fits (a fixed order makes each structure's results
// independent of the others in the plan).
static int plan_create_from(int device, int nstruct, const tbk_prepared *const *ps, const int *nk,
                            const double *kpts, const double *weights, int batch, int precision,
                            hipStream_t stream, int min_order, tbk_plan **out)
{
    if (!ps || !out || !nk || nstruct <= 0 || !kpts || !weights) return TBK_ERR_ARG;
    if (precision != TBK_FP64 && precision != TBK_MIXED) return TBK_ERR_ARG;
    *out = nullptr;
    if (hipSetDevice(device) != hipSuccess) return TBK_ERR_HIP;
    SetupArrays A;
    std::vector<StructHost> sh(nstruct);
    std::vector<SetupOffsets> off(nstruct);
    for (int i = 0; i < nstruct; ++i) {
        if (nk[i] <= 0 || !ps[i]) return TBK_ERR_ARG;
        append_setup(A, ps[i]->a, off[i]);
        off[i].noncol = ps[i]->off.noncol;
        sh[i] = ps[i]->sh;
    }
    int nktot = 0, N = 0, maxnat = 0, maxpairs = 1;
    for (int i = 0; i < nstruct; ++i) {
        sh[i].nk = nk[i]; sh[i].k0 = nktot;
        nktot += nk[i];
        N = std::max(N, sh[i].nv);
        maxnat = std::max(maxnat, sh[i].natoms);
        maxpairs = std::max(maxpairs, sh[i].npairs);
    }
    bool any_noncol = false;
    int maxnorb = 0;
    for (int i = 0; i < nstruct; ++i) {
        any_noncol = any_noncol || off[i].noncol;
        maxnorb = std::max(maxnorb, sh[i].norb);
    }
    if (precision == TBK_MIXED && N > tbk::oa_max_n()) return TBK_ERR_ARG;
    N = std::max(N, std::min(min_order, precision == TBK_MIXED ? std::min(kMaxN, tbk::oa_max_n()) : kMaxN));
    N = std::max(N, 2);           // a one-orbital cell is solved at order 2 (one padding row)
    // ---- device state ----
    tbk_plan *p = new tbk_plan();
    g_dev_bytes = 0;
    p->device = device;
    p->nstruct = nstruct; p->N = N; p->nk = nktot; p->precision = precision;
    p->nfermi = nstruct;
    p->noncol = any_noncol; p->maxnorb = maxnorb;
    p->maxnat = maxnat; p->maxpairs = maxpairs;
    p->sh = sh;
    p->kpts.assign(kpts, kpts + 3 * (size_t)nktot);
    p->weights.assign(weights, weights + nktot);
    p->batch = (batch <= 0 || batch > nktot) ? nktot : batch;
    // non-blocking: plans neither wait for nor hold up work on the null stream
    if (stream) {
        p->stream = stream;
        p->own_stream = false;
    } else if (hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess) {
        delete p;
        return TBK_ERR_HIP;
    }
    AllocStream as(p->stream);
    int rc = TBK_OK;
    rc |= upload(&p->row_atom, A.row_atom); rc |= upload(&p->row_orb, A.row_orb);
    rc |= upload(&p->ksk, A.ksk); rc |= upload(&p->nbas, A.nbas); rc |= upload(&p->blk_ptr, A.blk_ptr);
    rc |= upload(&p->dvec, A.dvec); rc |= upload(&p->eh, A.eh); rc |= upload(&p->es, A.es);
    rc |= upload(&p->edh, A.edh); rc |= upload(&p->eds, A.eds);
    rc |= upload(&p->paros, A.paros); rc |= upload(&p->dparos, A.dparos);
    rc |= upload(&p->spin_terms, A.spin);
    rc |= upload(&p->kp, p->kpts); rc |= upload(&p->wts, p->weights);
    std::vector<int> item_s(nktot), nact(nktot);
    for (int i = 0; i < nstruct; ++i)
        for (int k = 0; k < sh[i].nk; ++k) { item_s[sh[i].k0 + k] = i; nact[sh[i].k0 + k] = sh[i].nv; }
    p->item_s_host = item_s;
    rc |= upload(&p->item_s, item_s); rc |= upload(&p->nact, nact);
    rc |= dalloc(&p->esk, (size_t)nktot * N); rc |= dalloc(&p->desk, (size_t)nktot * N);
    rc |= dalloc(&p->ncut, (size_t)nktot); rc |= dalloc(&p->err, (size_t)nktot);
    rc |= dalloc(&p->cut, (size_t)nktot);
    rc |= dalloc(&p->info_all, (size_t)nktot); rc |= dalloc(&p->out, 8 * (size_t)nstruct);
    if (!rc) {
        std::vector<tbk::StructDev> sd(nstruct);
        for (int i = 0; i < nstruct; ++i) {
            tbk::StructDev &d = sd[i];
            d.natoms = sh[i].natoms; d.nv = sh[i].norb; d.npairs = sh[i].npairs;
            d.spin = off[i].noncol ? p->spin_terms + off[i].spin : nullptr;
            d.row_atom = p->row_atom + off[i].rows; d.row_orb = p->row_orb + off[i].rows;
            d.ksk = p->ksk + off[i].atoms; d.nbas = p->nbas + off[i].atoms;
            d.blk_ptr = p->blk_ptr + off[i].blk;
            d.dvec = p->dvec + 3 * off[i].pairs;
            d.eh = p->eh + 81 * off[i].pairs; d.es = p->es + 81 * off[i].pairs;
            d.edh = p->edh + 81 * off[i].pairs; d.eds = p->eds + 81 * off[i].pairs;
            d.paros = p->paros + 4 * off[i].atoms; d.dparos = p->dparos + 4 * off[i].atoms;
        }
        rc |= upload(&p->sdev, sd);
        std::vector<tbk::FermiArgs> fa(nstruct);
        rc |= upload(&p->fargs, fa);        // filled per run (mode, kT)
        if (hipStreamSynchronize(p->stream) != hipSuccess) rc |= TBK_ERR_HIP;
    }
    if (!rc) rc = dalloc(&p->ph, (size_t)p->batch * maxpairs);
    if (!rc) rc = p->ws.alloc(N, p->batch);
    if (!rc && precision == TBK_MIXED) {
        rc = p->ws.alloc_mixed();
    }
    // Few items of large order: one workgroup per item would leave most CUs
    // idle in the tridiagonalisation, so it runs over tiles instead.
    if (!rc && N >= kWideMinN && p->batch < kWideMaxBatch && !p->ws.resident)
        rc = p->ws.alloc_wide(precision == TBK_MIXED);
    if (!rc && hipStreamSynchronize(p->stream) != hipSuccess) rc = TBK_ERR_HIP;   // uploads done
    if (rc) { plan_free(p); return rc == TBK_ERR_HIP ? TBK_ERR_HIP : TBK_ERR_NOMEM; }
    p->bytes = g_dev_bytes;
    p->timer.st = p->stream;
    p->info_host.assign(nktot, 0);
    p->err_host.assign(nktot, 0);
    p->status.assign(nstruct, TBK_OK);
    p->out_host.assign(8 * (size_t)nstruct, 0.0);
    *out = p;
    return TBK_OK;
}

int tbk_plan_create_prepared(int device, int nstruct, const tbk_prepared *const *ps, const int *nk,
                             const double *kpts, const double *weights, int precision, hipStream_t stream,
                             int order, tbk_plan **out)
{
    return plan_create_from(device, nstruct, ps, nk, kpts, weights, 0, precision, stream, order, out);
}

static int plan_create(int device, int nstruct, const tbk_structure_desc *st, const int *nk, const double *kpts,
                       const double *weights, int batch, int precision, const SpinSel *sel, tbk_plan **out)
{
    if (!st || !out || !nk || nstruct <= 0 || !kpts || !weights) return TBK_ERR_ARG;
    if (precision != TBK_FP64 && precision != TBK_MIXED) return TBK_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return TBK_ERR_NODEVICE;
    if (device < 0 || device >= ndev) return TBK_ERR_ARG;
    if (hipSetDevice(device) != hipSuccess) return TBK_ERR_HIP;
    for (int i = 0; i < nstruct; ++i)
        if (nk[i] <= 0) return TBK_ERR_ARG;
    // host set-up (bit-exact) of every structure on host threads
    std::vector<tbk_prepared> loc(nstruct);
    std::vector<int> rcs(nstruct, TBK_OK);
    tbk::parallel_for(nstruct, [&](int i) {
        rcs[i] = setup_structure(st + i, loc[i].a, loc[i].sh, loc[i].off, sel ? sel[i] : SpinSel());
    });
    std::vector<const tbk_prepared *> ps(nstruct);
    for (int i = 0; i < nstruct; ++i) {
        if (rcs[i]) return rcs[i];
        ps[i] = &loc[i];
    }
    return plan_create_from(device, nstruct, ps.data(), nk, kpts, weights, batch, precision, nullptr, 0, out);
}

extern "C" int tbk_plan_create_group(int device, int nstruct, const tbk_structure_desc *st, const int *nk,
                                     const double *kpts, const double *weights, int batch, int precision,
                                     tbk_plan **out)
{
    return plan_create(device, nstruct, st, nk, kpts, weights, batch, precision, nullptr, out);
}

// bande.f:816-889 with jspins = 2: setpar/setparv once per spin channel
// (bond integrals from that channel, densities and on-site terms of flipped
// atoms from the other), both channels' k-points solved in the same
// launches, one Fermi level with one electron per state (bande.f:1069-1081)
// and the sums over both channels.
extern "C" int tbk_plan_create_spin(int device, const tbk_structure_desc *st, const double *param_dn,
                                    const int *flipos, int nk, const double *kpts, const double *weights,
                                    int batch, int precision, tbk_plan **out)
{
    if (!st || !param_dn || !st->param || nk <= 0 || !kpts || !weights || !out) return TBK_ERR_ARG;
    tbk_structure_desc d[2] = {*st, *st};
    d[1].param = param_dn;
    SpinSel sel[2];
    sel[0].param_flip = param_dn; sel[0].flip = flipos;
    sel[1].param_flip = st->param; sel[1].flip = flipos;
    const int nks[2] = {nk, nk};
    std::vector<double> k2(6 * (size_t)nk), w2(2 * (size_t)nk);
    for (int c = 0; c < 2; ++c) {
        std::copy(kpts, kpts + 3 * (size_t)nk, k2.begin() + 3 * (size_t)nk * c);
        std::copy(weights, weights + nk, w2.begin() + (size_t)nk * c);
    }
    const int rc = plan_create(device, 2, d, nks, k2.data(), w2.data(), batch, precision, sel, out);
    if (rc) return rc;
    (*out)->spin = true;
    (*out)->nfermi = 1;
    return TBK_OK;
}

// bande.f:713-720, 859-905, 945-1036 with lnoncol (jspins = 2): the
// spatial H, S from channel 1's integrals and the channels' average on-site
// terms, doubled into the spinor matrices of order 2 nv with each atom's
// spin direction (spinham.f, addspins.f); one electron per state.
extern "C" int tbk_plan_create_noncol(int device, const tbk_structure_desc *st, const double *param_dn,
                                      const double *spinat, int nk, const double *kpts, const double *weights,
                                      int batch, int precision, tbk_plan **out)
{
    if (!st || !param_dn || !spinat || !st->param || nk <= 0 || !kpts || !weights || !out) return TBK_ERR_ARG;
    SpinSel sel;
    sel.nc_param_dn = param_dn;
    sel.nc_spinat = spinat;
    return plan_create(device, 1, st, &nk, kpts, weights, batch, precision, &sel, out);
}

extern "C" int tbk_plan_magnetization(const tbk_plan *p, double *smag)
{
    if (!p || !smag) return TBK_ERR_ARG;
    *smag = p->spin ? p->out_host[6] : 0.0;
    return TBK_OK;
}

extern "C" int tbk_plan_set_spin_directions(tbk_plan *p, int on)
{
    if (!p || !p->noncol) return TBK_ERR_ARG;
    p->spdir = on != 0;
    if (p->gexec) { (void)hipGraphExecDestroy(p->gexec); p->gexec = nullptr; }
    return TBK_OK;
}

extern "C" int tbk_plan_spin_directions(tbk_plan *p, int s, double *spdir)
{
    if (!p || !spdir || s < 0 || s >= p->nstruct || !p->spdir || !p->spdir_out) return TBK_ERR_ARG;
    HIP_TRY(hipSetDevice(p->device));
    HIP_TRY(hipStreamSynchronize(p->stream));
    const size_t per = (size_t)p->maxnat * 3;
    HIP_TRY(hipMemcpy(spdir, p->spdir_out + (size_t)s * per, (size_t)p->sh[s].natoms * 3 * sizeof(double),
                      hipMemcpyDeviceToHost));
    return TBK_OK;
}

extern "C" int tbk_plan_create(int device, const tbk_structure_desc *st, int nk,
                               const double *kpts, const double *weights, int batch,
                               int precision, tbk_plan **out)
{
    if (!st || nk <= 0) return TBK_ERR_ARG;
    return tbk_plan_create_group(device, 1, st, &nk, kpts, weights, batch, precision, out);
}

// Padding diagonal of H: above every level of a tight-binding structure in
// Ry, and exactly decoupled from the active block.
constexpr double kPadLevel = 1.0e4;

// H, S (deriv false) or dH/dV, dS/dV (deriv true) into ws.A, ws.S; with
// four, all of them, the derivatives into ws.dA, ws.dS.
static tbk::BuildArgs build_args(tbk_plan *p, int k0, bool deriv, bool four = false)
{
    tbk::BuildArgs a;
    a.n = p->N; a.deriv = deriv ? 1 : 0; a.maxpairs = p->maxpairs;
    a.stride = p->ws.stride();
    a.pad = kPadLevel;
    a.sd = p->sdev;
    a.item_s = p->item_s + k0;
    a.ph = p->ph;
    a.h = p->ws.A; a.s = p->ws.S;
    if (four) { a.dh = p->ws.dA; a.ds = p->ws.dS; }
    return a;
}
// The derivative pair of a four-matrix build, for the spinor expansion.
static tbk::BuildArgs deriv_args(tbk_plan *p, int k0)
{
    tbk::BuildArgs a = build_args(p, k0, true);
    a.h = p->ws.dA; a.s = p->ws.dS;
    return a;
}
// Buffers of the four-matrix build, on the first pressure run (a plan
// that cannot hold them builds dH, dS after the solve instead).
static void ensure_deriv_buffers(tbk_plan *p)
{
    if (!p->fused_build || p->ws.dA) return;
    AllocStream as(p->stream);
    const size_t before = g_dev_bytes;
    if (p->ws.alloc_deriv() == TBK_OK) p->bytes += g_dev_bytes - before;
}

// Everything one run enqueues on the plan's stream (captured into a
// hipGraph when graphs are on).
static void enqueue_run(tbk_plan *p, int mode, double eigcut)
{
    const bool lpress = mode == 2 || mode == 4;
    const int n = p->N;
    const size_t ss = p->ws.stride();
    hipStream_t st = p->stream;
    // Occupation window: with every k in one batch, E_F is known once the
    // eigenvalues are, and eigenvectors and derivatives are formed only for
    // the states the Fermi sums use (below E_F + 20 kT).
    // Energy-only modes (1, 3) need no eigenvectors at all.
    // QLMT modes (6, 10) need every eigenvector below eigcut and S^{1/2};
    // 8 and 9 only eigenvalues
    const bool qproj = mode == 6 || mode == 10;
    const bool win = p->window && lpress && p->batch >= p->nk;
    const bool vecs = lpress || qproj || (!p->window && mode <= 4) || p->spdir;
    const Window wnd{win ? p->fargs : nullptr, p->nfermi, win ? p->cut : nullptr, vecs};
    p->win_last = win;
    p->vec_last = vecs;
    // pressure modes: dH/dV, dS/dV from the same pass as H, S when the
    // plan holds their buffers
    const bool fused = lpress && p->fused_build && p->ws.dA;
    const double2 *dH = fused ? p->ws.dA : p->ws.A, *dS = fused ? p->ws.dS : p->ws.S;
    for (int k0 = 0; k0 < p->nk; k0 += p->batch) {
        const int nb = std::min(p->batch, p->nk - k0);
        p->timer.begin("build_hs");
        tbk::launch_phases(p->sdev, p->item_s + k0, p->kp + 3 * (size_t)k0, p->maxpairs, p->ph, nb, st);
        {
            tbk::BuildArgs hb = build_args(p, k0, false, fused);
            hb.s_upper_only = fused ? 0 : 1;
            tbk::launch_build_hs(hb, p->maxnat, nb, st);
        }
        if (p->noncol) {
            tbk::launch_spinor_expand(build_args(p, k0, false), p->maxnorb, nb, st);
            if (fused) tbk::launch_spinor_expand(deriv_args(p, k0), p->maxnorb, nb, st);
        }
        p->timer.end();
        const int *nact = p->nact + k0;
        if (p->precision == TBK_MIXED)
            solve_batch_mixed(p->ws, nb, nact, p->esk + (size_t)k0 * n, st, p->timer, wnd);
        else
            solve_batch(p->ws, nb, nact, p->esk + (size_t)k0 * n, st, p->timer, wnd);
        tbk::launch_copy_int(p->ws.info, p->info_all + k0, nb, st);
        if (p->spdir) {               // each state's spin on each atom (diag.f:458-523)
            tbk::SpinArgs sa{n, p->maxnat, k0, ss, p->sdev, p->item_s + k0, nact, wnd.cut ? wnd.cut + k0 : nullptr,
                             p->ws.Z, p->spev};
            tbk::launch_spin_expect(sa, nb, st);
        }
        if (qproj) {
            // S(k) again, into ws2.A (H into ws2.S, replaced by I), its
            // eigenpairs S = V diag(lambda) V^H, then Y = V (lambda^{1/2} (V^H Z))
            p->timer.begin("projections");
            Workspace &w2 = p->ws2;
            tbk::BuildArgs b = build_args(p, k0, false);
            b.h = w2.S;
            b.s = w2.A;
            tbk::launch_build_hs(b, p->maxnat, nb, st);
            if (p->noncol) tbk::launch_spinor_expand(b, p->maxnorb, nb, st);
            tbk::launch_set_identity(n, w2.S, ss, nb, st);
            StageTimer quiet;
            solve_batch(w2, nb, nact, w2.w, st, quiet, Window{nullptr, 0, nullptr, true});
            tbk::launch_zgemm(gemm(n, n, n, w2.Z, n, ss, p->ws.Z, n, ss, w2.X, n, ss), 1, 0, nb, st);
            tbk::launch_scale_rows_sqrt(n, w2.X, ss, w2.w, nb, st);
            tbk::launch_zgemm(gemm(n, n, n, w2.Z, n, ss, w2.X, n, ss, w2.Y, n, ss), 0, 0, nb, st);
            tbk::launch_abs2(n, w2.Y, ss, p->proj + (size_t)k0 * n * n, nb, st);
            p->timer.end();
        }
        if (lpress) {
            p->timer.begin("pressure");
            if (!fused) {
                // the products below read the lower triangles only
                tbk::BuildArgs db = build_args(p, k0, true);
                db.lower_only = 1;
                tbk::launch_build_hs(db, p->maxnat, nb, st);
                if (p->noncol) tbk::launch_spinor_expand(db, p->maxnorb, nb, st);
            }
            // X = L_dH Z, Y = L_dS Z over the lower triangles (half the
            // products; k_perturb adds the conjugate halves)
            tbk::launch_zgemm(with_ncol(gemm(n, n, n, dH, n, ss, p->ws.Z, n, ss, p->ws.X, n, ss, 9), wnd.cut), 0,
                              0, nb, st);
            tbk::launch_zgemm(with_ncol(gemm(n, n, n, dS, n, ss, p->ws.Z, n, ss, p->ws.Y, n, ss, 9), wnd.cut), 0,
                              0, nb, st);
            tbk::launch_coldot(n, p->ws.Z, p->ws.X, p->ws.Y, ss, p->ws.px, p->ws.py, nb, st, wnd.cut);
            tbk::PerturbArgs pa;
            pa.cut = wnd.cut;
            pa.n = n; pa.nb = nb; pa.k0 = k0; pa.stride = ss; pa.eigcut = eigcut; pa.nact = nact;
            pa.w = p->esk + (size_t)k0 * n; pa.px = p->ws.px; pa.py = p->ws.py;
            pa.z = p->ws.Z; pa.x = p->ws.X; pa.y = p->ws.Y;
            pa.desk = p->desk + (size_t)k0 * n; pa.ncut = p->ncut; pa.err = p->err;
            pa.one_orbital_zero = true;
            tbk::launch_perturb(pa, st);
            p->timer.end();
        }
    }
    p->timer.begin("fermi_sums");
    tbk::launch_fermi_sums_group(p->fargs, p->nfermi, st, win ? 1 : 0);
    if (p->spdir) tbk::launch_spin_sum(p->fargs, p->sdev, p->nfermi, p->spev, n, p->maxnat, p->spdir_out, st);
    p->timer.end();
}

// Modes of bande (bande.f:99-123) a plan runs: 1-4, and the energy-only
// modes that write QLMT (6, 9, 10) or QAPW (8) files, whose sums are those
// of mode 3.  Mode 5 (the unformatted QLMT) is refused: the reference's own
// program aborts on it (a formatted write to the unformatted unit 45,
// bande.f:1609).
static bool plan_mode_ok(int mode) { return (mode >= 1 && mode <= 4) || mode == 6 || mode == 8 || mode == 9 || mode == 10; }

extern "C" int tbk_plan_launch(tbk_plan *p, int mode, double tkb, double eigcut)
{
    if (!p || !plan_mode_ok(mode)) return TBK_ERR_ARG;
    for (const StructHost &h : p->sh)
        if (h.zfill <= 0.0) return TBK_ERR_ZFILL;
    HIP_TRY(hipSetDevice(p->device));
    if ((mode == 6 || mode == 10) && !p->proj) {   // the projection buffers, on first use
        AllocStream as(p->stream);
        const size_t before = g_dev_bytes;
        if (p->ws2.alloc(p->N, p->batch) || dalloc(&p->proj, (size_t)p->nk * p->N * p->N)) {
            p->ws2.release();
            dfree(p->proj);
            return TBK_ERR_NOMEM;
        }
        p->ws2.two_stage = p->ws.two_stage;
        if (p->ws.two_stage || !p->ws.resident) p->ws2.resident = false;
        p->bytes += g_dev_bytes - before;
    }
    if (mode == 2 || mode == 4) ensure_deriv_buffers(p);
    if (p->spdir && !p->spev) {
        AllocStream as(p->stream);
        const size_t before = g_dev_bytes;
        const size_t per = (size_t)p->maxnat * 3;
        if (dalloc(&p->spev, (size_t)p->nk * p->N * per) || dalloc(&p->spdir_out, (size_t)p->nstruct * per)) {
            dfree(p->spev); dfree(p->spdir_out);
            return TBK_ERR_NOMEM;
        }
        p->bytes += g_dev_bytes - before;
    }
    p->last_mode = mode;
    const bool params_changed =
        !(p->g_mode == mode && p->g_tkb == tkb && p->g_eigcut == eigcut && p->g_window == p->window);
    if (p->f_mode != mode || p->f_tkb != tkb) {
        // per-structure Fermi arguments of this (mode, kT)
        std::vector<tbk::FermiArgs> fa(p->nfermi);
        for (int i = 0; i < p->nfermi; ++i) {
            const StructHost &h = p->sh[i];
            tbk::FermiArgs &f = fa[i];
            f.nk = h.nk; f.nv = h.nv; f.ld = p->N; f.mode = mode > 4 ? 3 : mode;
            f.el = 2.0; f.tkb = tkb; f.zfill = h.zfill;
            f.esk = p->esk + (size_t)h.k0 * p->N; f.desk = p->desk + (size_t)h.k0 * p->N;
            f.w = p->wts + h.k0; f.ncut = p->ncut + h.k0; f.out = p->out + 8 * (size_t)i;
            f.cut = p->cut + h.k0;
            f.k0 = h.k0;
            if (p->noncol) f.el = 1.0;          // one electron per spinor state (bande.f:1069)
            if (p->spin) {          // both channels' items follow each other
                f.nk = p->sh[0].nk + p->sh[1].nk;
                f.el = 1.0;
                f.khalf = p->sh[0].nk;
            }
        }
        HIP_TRY(hipMemcpyAsync(p->fargs, fa.data(), fa.size() * sizeof(tbk::FermiArgs),
                               hipMemcpyHostToDevice, p->stream));
        HIP_TRY(hipStreamSynchronize(p->stream));
        p->f_mode = mode; p->f_tkb = tkb;
    }
    const bool graph = p->use_graph && !p->timer.on && !p->recording;
    if (!graph) {
        enqueue_run(p, mode, eigcut);
        if (hipGetLastError() != hipSuccess) return TBK_ERR_HIP;
        p->g_mode = mode; p->g_tkb = tkb; p->g_eigcut = eigcut; p->g_window = p->window;
        if (p->gexec && params_changed) { (void)hipGraphExecDestroy(p->gexec); p->gexec = nullptr; }
        return TBK_OK;
    }
    if (!p->gexec || params_changed) {
        if (p->gexec) { (void)hipGraphExecDestroy(p->gexec); p->gexec = nullptr; }
        hipGraph_t g = nullptr;
        HIP_TRY(hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal));
        enqueue_run(p, mode, eigcut);
        const hipError_t ce = hipStreamEndCapture(p->stream, &g);
        if (ce != hipSuccess || !g) return TBK_ERR_HIP;
        const hipError_t ie = hipGraphInstantiate(&p->gexec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ie != hipSuccess) { p->gexec = nullptr; return TBK_ERR_HIP; }
        p->g_mode = mode; p->g_tkb = tkb; p->g_eigcut = eigcut; p->g_window = p->window;
    }
    HIP_TRY(hipGraphLaunch(p->gexec, p->stream));
    return TBK_OK;
}

extern "C" int tbk_plan_wait(tbk_plan *p, double *efermi, double *earg1, double *earg2)
{
    if (!p) return TBK_ERR_ARG;
    HIP_TRY(hipSetDevice(p->device));
    HIP_TRY(hipMemcpyAsync(p->out_host.data(), p->out, p->out_host.size() * sizeof(double),
                           hipMemcpyDeviceToHost, p->stream));
    HIP_TRY(hipMemcpyAsync(p->info_host.data(), p->info_all, p->nk * sizeof(int), hipMemcpyDeviceToHost,
                           p->stream));
    const bool lpress = p->last_mode == 2 || p->last_mode == 4;
    if (lpress)
        HIP_TRY(hipMemcpyAsync(p->err_host.data(), p->err, p->nk * sizeof(int), hipMemcpyDeviceToHost,
                               p->stream));
    HIP_TRY(hipStreamSynchronize(p->stream));
    p->timer.collect();
    int first = TBK_OK;
    for (int i = 0; i < p->nstruct; ++i) {
        const StructHost &h = p->sh[i];
        int st = TBK_OK;
        for (int k = h.k0; k < h.k0 + h.nk && st == TBK_OK; ++k) {
            if (p->info_host[k] != 0) st = TBK_ERR_OVERLAP;
            else if (lpress && (p->err_host[k] & 2)) st = TBK_ERR_DEGEN;
        }
        p->status[i] = st;
        if (st && !first) first = st;
    }
    if (first) return first;
    if (efermi) *efermi = p->out_host[0];
    if (earg1) *earg1 = p->out_host[1];
    if (earg2) *earg2 = p->out_host[2];
    return TBK_OK;
}

extern "C" int tbk_plan_run(tbk_plan *p, int mode, double tkb, double eigcut, double *efermi,
                            double *earg1, double *earg2)
{
    int rc = tbk_plan_launch(p, mode, tkb, eigcut);
    if (rc) return rc;
    return tbk_plan_wait(p, efermi, earg1, earg2);
}

// TBK_OK when the plan's stream is idle (a launched run has finished), 1
// while it still runs.
extern "C" int tbk_plan_query(const tbk_plan *p)
{
    if (!p) return TBK_ERR_ARG;
    const hipError_t e = hipStreamQuery(p->stream);
    if (e == hipSuccess) return TBK_OK;
    if (e == hipErrorNotReady) return 1;
    return TBK_ERR_HIP;
}

extern "C" int tbk_plan_result(const tbk_plan *p, int s, double *efermi, double *earg1, double *earg2)
{
    if (!p || s < 0 || s >= p->nstruct) return TBK_ERR_ARG;
    if (p->status[s]) return p->status[s];
    if (p->spin) {                  // one result for both channels
        if (p->status[1 - s]) return p->status[1 - s];
        s = 0;
    }
    if (efermi) *efermi = p->out_host[8 * (size_t)s];
    if (earg1) *earg1 = p->out_host[8 * (size_t)s + 1];
    if (earg2) *earg2 = p->out_host[8 * (size_t)s + 2];
    return TBK_OK;
}

// GEMM autotuner: one mode-4 pass records the plan's GEMM shapes (the
// shapes this thread launches during the pass), then each shape not yet in
// the table is timed with every tile (one warm-up and `reps` launches on the
// plan's stream) and the fastest tile is kept for the rest of the process.
// The timing launches overwrite workspace contents, so the plan's results
// are undefined until its next run; launches of other plans meanwhile only
// disturb the timings (a tile choice changes speed, never results).
extern "C" int tbk_plan_autotune(tbk_plan *p, int reps, int *ntuned)
{
    if (!p || reps < 1) return TBK_ERR_ARG;
    HIP_TRY(hipSetDevice(p->device));
    p->recording = true;
    tbk::gemm_record(true);
    int rc = tbk_plan_launch(p, 4, 0.005, 1e30);
    if (!rc) rc = hipStreamSynchronize(p->stream) == hipSuccess ? TBK_OK : TBK_ERR_HIP;
    tbk::gemm_record(false);
    p->recording = false;
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
    if (p->gexec) { (void)hipGraphExecDestroy(p->gexec); p->gexec = nullptr; }   // tiles changed
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

extern "C" int tbk_plan_set_graph(tbk_plan *p, int on)
{
    if (!p) return TBK_ERR_ARG;
    p->use_graph = on != 0;
    if (!p->use_graph && p->gexec) {
        (void)hipSetDevice(p->device);
        (void)hipGraphExecDestroy(p->gexec);
        p->gexec = nullptr;
    }
    return TBK_OK;
}

extern "C" int tbk_plan_structure_eigenvalues(tbk_plan *p, int s, double *esk, double *desk)
{
    if (!p || s < 0 || s >= p->nstruct) return TBK_ERR_ARG;
    HIP_TRY(hipSetDevice(p->device));
    const StructHost &h = p->sh[s];
    const size_t N = p->N;
    // rows of N (padded order) -> rows of nv
    if (esk) HIP_TRY(hipMemcpy2DAsync(esk, h.nv * sizeof(double), p->esk + (size_t)h.k0 * N, N * sizeof(double),
                                      h.nv * sizeof(double), h.nk, hipMemcpyDeviceToHost, p->stream));
    if (desk) HIP_TRY(hipMemcpy2DAsync(desk, h.nv * sizeof(double), p->desk + (size_t)h.k0 * N,
                                       N * sizeof(double), h.nv * sizeof(double), h.nk, hipMemcpyDeviceToHost,
                                       p->stream));
    HIP_TRY(hipStreamSynchronize(p->stream));
    return TBK_OK;
}

extern "C" int tbk_plan_eigenvalues(tbk_plan *p, double *esk, double *desk)
{
    return tbk_plan_structure_eigenvalues(p, 0, esk, desk);
}

// Leading nv x nv block (column stride nv) of an N x N item matrix.
static int copy_leading(double2 *dst, const double2 *src, int N, int nv, hipStream_t st)
{
    HIP_TRY(hipMemcpy2DAsync(dst, nv * sizeof(double2), src, N * sizeof(double2), nv * sizeof(double2), nv,
                             hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return TBK_OK;
}

extern "C" int tbk_plan_eigenvectors(tbk_plan *p, int k, double *z)
{
    if (!p || !z || k < 0 || k >= p->nk) return TBK_ERR_ARG;
    // the workspace holds the last batch: k must fall inside it
    const int last0 = ((p->nk - 1) / p->batch) * p->batch;
    if (k < last0) return TBK_ERR_ARG;
    HIP_TRY(hipSetDevice(p->device));
    HIP_TRY(hipStreamSynchronize(p->stream));
    const size_t ss = p->ws.stride();
    const int nv = p->sh[p->item_s_host[k]].nv;
    if (!p->vec_last) {                      // an energy-only run formed no eigenvector
        for (size_t t = 0; t < (size_t)nv * nv * 2; ++t) z[t] = 0.0;
        return TBK_OK;
    }
    int rc = copy_leading((double2 *)z, p->ws.Z + (size_t)(k - last0) * ss, p->N, nv, p->stream);
    if (rc == TBK_OK && p->win_last) {       // columns beyond the occupation window were not formed
        int c = nv;
        if (hipMemcpy(&c, p->cut + k, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) return TBK_ERR_HIP;
        for (size_t t = (size_t)std::max(0, std::min(c, nv)) * nv * 2; t < (size_t)nv * nv * 2; ++t) z[t] = 0.0;
    }
    return rc;
}

extern "C" int tbk_plan_projections(tbk_plan *p, int k, double eigcut, int *ntot, double *e, double *proj)
{
    if (!p || k < 0 || k >= p->nk || !ntot || !e || !proj) return TBK_ERR_ARG;
    if (!p->proj || !(p->last_mode == 6 || p->last_mode == 10)) return TBK_ERR_ARG;
    HIP_TRY(hipSetDevice(p->device));
    const int n = p->N, nv = p->sh[p->item_s_host[k]].nv;
    std::vector<double> ek(n);
    HIP_TRY(hipMemcpyAsync(ek.data(), p->esk + (size_t)k * n, n * sizeof(double), hipMemcpyDeviceToHost,
                           p->stream));
    HIP_TRY(hipStreamSynchronize(p->stream));
    int m = nv;                                   // diag.f:635-641: states up to eigcut
    for (int j = 0; j < nv; ++j)
        if (ek[j] > eigcut) { m = j; break; }
    *ntot = m;
    for (int j = 0; j < m; ++j) e[j] = ek[j];
    if (m > 0)
        HIP_TRY(hipMemcpy2DAsync(proj, nv * sizeof(double), p->proj + (size_t)k * n * n, n * sizeof(double),
                                 nv * sizeof(double), m, hipMemcpyDeviceToHost, p->stream));
    HIP_TRY(hipStreamSynchronize(p->stream));
    return TBK_OK;
}

extern "C" int tbk_plan_set_tridiag(tbk_plan *p, int stages)
{
    if (!p || stages < 0 || stages > 3) return TBK_ERR_ARG;
    const bool fits_res = p->N <= tbk::hetrd_resident_max_n(p->precision == TBK_MIXED);
    const bool res = stages == 0 ? fits_res : stages == 3;
    const bool two = stages == 0 ? !res && p->N >= kTwoStageMinN && two_stage_fits(p->N) : stages == 2;
    if (two && !two_stage_fits(p->N)) return TBK_ERR_ARG;
    if (res && !fits_res) return TBK_ERR_ARG;
    p->ws.two_stage = two;
    p->ws2.two_stage = two;
    p->ws.resident = res;
    // the projections' eigendecomposition of S runs in FP64
    p->ws2.resident = !two && (stages == 0 || stages == 3) && p->N <= tbk::hetrd_resident_max_n(false);
    if (p->gexec) { (void)hipGraphExecDestroy(p->gexec); p->gexec = nullptr; }
    return TBK_OK;
}

extern "C" int tbk_plan_set_fused_build(tbk_plan *p, int on)
{
    if (!p) return TBK_ERR_ARG;
    p->fused_build = on != 0;
    if (p->gexec) { (void)hipGraphExecDestroy(p->gexec); p->gexec = nullptr; }
    return TBK_OK;
}

extern "C" int tbk_plan_set_window(tbk_plan *p, int on)
{
    if (!p) return TBK_ERR_ARG;
    p->window = on != 0;
    return TBK_OK;
}

extern "C" int tbk_plan_hs(tbk_plan *p, int k, int deriv, double *h, double *s)
{
    if (!p || k < 0 || k >= p->nk) return TBK_ERR_ARG;
    HIP_TRY(hipSetDevice(p->device));
    // the build the plan's runs use: four matrices in one pass when it
    // holds their buffers (pressure runs), else the pair asked for
    ensure_deriv_buffers(p);
    const bool four = p->fused_build && p->ws.dA;
    tbk::launch_phases(p->sdev, p->item_s + k, p->kp + 3 * (size_t)k, p->maxpairs, p->ph, 1, p->stream);
    if (four) {
        tbk::launch_build_hs(build_args(p, k, false, true), p->maxnat, 1, p->stream);
        if (p->noncol) {
            tbk::launch_spinor_expand(build_args(p, k, false), p->maxnorb, 1, p->stream);
            tbk::launch_spinor_expand(deriv_args(p, k), p->maxnorb, 1, p->stream);
        }
    } else {
        tbk::launch_build_hs(build_args(p, k, deriv != 0), p->maxnat, 1, p->stream);
        if (p->noncol) tbk::launch_spinor_expand(build_args(p, k, deriv != 0), p->maxnorb, 1, p->stream);
    }
    HIP_TRY(hipStreamSynchronize(p->stream));
    const int nv = p->sh[p->item_s_host[k]].nv;
    const double2 *hs = four && deriv ? p->ws.dA : p->ws.A, *ss = four && deriv ? p->ws.dS : p->ws.S;
    if (h && copy_leading((double2 *)h, hs, p->N, nv, p->stream)) return TBK_ERR_HIP;
    if (s && copy_leading((double2 *)s, ss, p->N, nv, p->stream)) return TBK_ERR_HIP;
    return TBK_OK;
}

extern "C" int tbk_plan_nv(const tbk_plan *p) { return p ? p->sh[0].nv : -1; }
extern "C" int tbk_plan_npairs(const tbk_plan *p) { return p ? p->sh[0].npairs : -1; }
extern "C" int tbk_plan_order(const tbk_plan *p) { return p ? p->N : -1; }
extern "C" int tbk_plan_nstruct(const tbk_plan *p) { return p ? p->nstruct : -1; }
extern "C" long long tbk_plan_device_bytes(const tbk_plan *p) { return p ? (long long)p->bytes : -1; }

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
        solve_batch(ws, 1, nullptr, ws.w, st, tm, Window{nullptr, 0, nullptr});
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

// perturb.f:1-221 for one k on the device: X = L_dH Z, Y = L_dS Z over the
// lower triangles (GEMMs), the column dots, then the degenerate groups
// (k_coldot, k_perturb).
extern "C" int tbk_perturb(int device, int n, const double *dh, const double *ds, const double *z,
                           const double *eig, double eigcut, int *ntot, double *deig)
{
    if (n < 1 || n > kMaxN || !dh || !ds || !z || !eig || !ntot || !deig) return TBK_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return TBK_ERR_NODEVICE;
    if (device < 0 || device >= ndev) return TBK_ERR_ARG;
    HIP_TRY(hipSetDevice(device));
    Workspace ws;
    if (ws.alloc(n, 1)) { ws.release(); return TBK_ERR_NOMEM; }
    hipStream_t st;
    if (hipStreamCreate(&st) != hipSuccess) { ws.release(); return TBK_ERR_HIP; }
    const size_t ss = ws.stride();
    int rc = TBK_OK, ncut = 0, err = 0;
    if (hipMemcpyAsync(ws.A, dh, ss * sizeof(double2), hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(ws.S, ds, ss * sizeof(double2), hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(ws.Z, z, ss * sizeof(double2), hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(ws.w, eig, n * sizeof(double), hipMemcpyHostToDevice, st) != hipSuccess) {
        rc = TBK_ERR_HIP;
    } else {
        tbk::launch_zgemm(gemm(n, n, n, ws.A, n, ss, ws.Z, n, ss, ws.X, n, ss, 9), 0, 0, 1, st);
        tbk::launch_zgemm(gemm(n, n, n, ws.S, n, ss, ws.Z, n, ss, ws.Y, n, ss, 9), 0, 0, 1, st);
        tbk::launch_coldot(n, ws.Z, ws.X, ws.Y, ss, ws.px, ws.py, 1, st);
        tbk::PerturbArgs pa;
        pa.n = n; pa.nb = 1; pa.k0 = 0; pa.stride = ss; pa.eigcut = eigcut; pa.nact = nullptr;
        pa.w = ws.w; pa.px = ws.px; pa.py = ws.py; pa.z = ws.Z; pa.x = ws.X; pa.y = ws.Y;
        pa.desk = ws.d; pa.ncut = ws.info; pa.err = ws.sinfo;
        tbk::launch_perturb(pa, st);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(deig, ws.d, n * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipMemcpyAsync(&ncut, ws.info, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipMemcpyAsync(&err, ws.sinfo, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            rc = TBK_ERR_HIP;
        else if (err & 2)
            rc = TBK_ERR_DEGEN;
    }
    (void)hipStreamDestroy(st);
    ws.release();
    if (!rc) *ntot = ncut;
    return rc;
}

extern "C" int tbk_fermi(int device, int ldn, int nk, int nv, double el, const double *weight,
                         const double *eigen, double tkb, double zfill, double *efermi)
{
    if (!weight || !eigen || !efermi || nk <= 0 || nv < 1 || ldn < nv) return TBK_ERR_ARG;
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
    tbk::launch_zgemm(gemm(n, n, n, ws.S, n, ss, ws.Y, n, ss, ws.A, n, ss, 6), 1, 0, 1, st);
    (void)hipDeviceSynchronize();
    get(c, ws.A, ss * sizeof(double2));
    {
        StageTimer off;
        hetrd(ws, 1, st, off);
    }
    (void)hipDeviceSynchronize();
    get(refl, ws.A, ss * sizeof(double2));
    get(d, ws.d, n * sizeof(double));
    get(e, ws.e, n * sizeof(double));
    tbk::SteinArgs sa{n, kOrtolRel, nullptr, ws.d, ws.e, ws.w, ws.Vt, ws.scratch, ws.sinfo};
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
    solve_batch(ws, 1, nullptr, ws.w, st, tm, Window{nullptr, 0, nullptr});
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
    if (tri < 0 || tri > 9 || tri == 8 || (tri >= 5 && tri <= 6 && m != n) || (tri == 9 && (opa != 0 || k != m)) ||
        m <= 0 || n <= 0 || k <= 0)
        return TBK_ERR_ARG;
    HIP_TRY(hipSetDevice(device));
    if (precision == TBK_MIXED)
        return debug_gemm_t<float2, tbk::GemmArgsT<float>>(tile, m, n, k, opa, opb, tri, a, lda, b, ldb, c, ldc);
    return debug_gemm_t<double2, tbk::GemmArgsT<double>>(tile, m, n, k, opa, opb, tri, a, lda, b, ldb, c, ldc);
}

// One launch of a plain device-to-device copy kernel over `bytes` bytes
// (reads bytes, writes bytes): the reference point for traffic counters.
extern "C" int tbk_debug_copy(int device, long long bytes)
{
    if (bytes <= 0 || bytes % 16) return TBK_ERR_ARG;
    HIP_TRY(hipSetDevice(device));
    double2 *a = nullptr, *b = nullptr;
    const size_t n = (size_t)bytes / 16;
    int rc = dalloc(&a, n) | dalloc(&b, n);
    if (!rc) {
        (void)hipMemset(a, 0, n * sizeof(double2));
        (void)hipDeviceSynchronize();
        tbk::launch_copy16(a, b, n, nullptr);
        if (hipDeviceSynchronize() != hipSuccess) rc = TBK_ERR_HIP;
    } else {
        rc = TBK_ERR_NOMEM;
    }
    dfree(a); dfree(b);
    return rc;
}

// Average time of `reps` launches of the batched GEMM (nb items of m x n x k,
// operands as tbk_debug_zgemm; contents zero) with tile `tile` (-1: tuned or
// modelled choice): the microbenchmark behind the tile autotuner.
extern "C" int tbk_debug_gemm_time(int device, int precision, int tile, int m, int n, int k, int opa, int opb,
                                   int tri, int nb, int reps, double *ms)
{
    if (m <= 0 || n <= 0 || k <= 0 || nb <= 0 || reps <= 0 || !ms || tile >= tbk::kGemmTiles) return TBK_ERR_ARG;
    HIP_TRY(hipSetDevice(device));
    const size_t lda = opa ? k : m, ldb = opb ? n : k;
    const size_t sa = lda * (opa ? m : k), sb = ldb * (opb ? k : n), sc = (size_t)m * n;
    const size_t esz = precision == TBK_MIXED ? sizeof(float2) : sizeof(double2);
    char *a = nullptr, *b = nullptr, *c = nullptr;
    int rc = dalloc(&a, sa * nb * esz) | dalloc(&b, sb * nb * esz) | dalloc(&c, sc * nb * esz);
    if (rc) { dfree(a); dfree(b); dfree(c); return TBK_ERR_NOMEM; }
    (void)hipMemset(a, 0, sa * nb * esz);
    (void)hipMemset(b, 0, sb * nb * esz);
    tbk::GemmShape sh{precision == TBK_MIXED ? 1 : 0, opa, opb, m, n, k, tri, nb, {}, {}};
    if (precision == TBK_MIXED) {
        sh.g32.m = m; sh.g32.n = n; sh.g32.k = k; sh.g32.tri = tri;
        sh.g32.alpha = float2{1.f, 0.f}; sh.g32.beta = float2{0.f, 0.f};
        sh.g32.a = (const float2 *)a; sh.g32.lda = (int)lda; sh.g32.sa = sa;
        sh.g32.b = (const float2 *)b; sh.g32.ldb = (int)ldb; sh.g32.sb = sb;
        sh.g32.c = (float2 *)c; sh.g32.ldc = m; sh.g32.sc = sc;
    } else {
        sh.g64 = gemm(m, n, k, (const double2 *)a, (int)lda, sa, (const double2 *)b, (int)ldb, sb, (double2 *)c, m,
                      sc, tri);
    }
    const int t = tile >= 0 ? tile : tbk::gemm_tile(sh.prec, opa, opb, m, n, k, tri, nb);
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    tbk::launch_gemm_tile(sh, t, nullptr);
    (void)hipEventRecord(e0, nullptr);
    for (int r = 0; r < reps; ++r) tbk::launch_gemm_tile(sh, t, nullptr);
    (void)hipEventRecord(e1, nullptr);
    float f = 0.f;
    if (hipEventSynchronize(e1) != hipSuccess) rc = TBK_ERR_HIP;
    (void)hipEventElapsedTime(&f, e0, e1);
    *ms = f / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    dfree(a); dfree(b); dfree(c);
    return rc;
}

// ---------------------------------------------------------------------------
// tbk_zhegv_noalloc: tbk_zhegv with its whole device workspace in `scratch`,
// device memory the code object reserves when the library is loaded on a GPU
// (one copy per GPU, held until the process ends), for callers that cannot
// allocate.  No hipMalloc, hipFree or stream creation: host copies and the
// solve run on the null stream.  One call at a time per GPU (a per-device
// mutex serialises this process's threads; each call overwrites the array).
// ---------------------------------------------------------------------------
namespace {
constexpr int kNoAllocMaxN = 512;
constexpr int kNoAllocDevices = 64;
// Workspace::carve(nullptr, n, 1) for n = kNoAllocMaxN, in bytes: six
// complex N x N slabs, the real N x N vectors, the panel and WY buffers, the
// per-thread tridiagonal scratch (5 N ceil64(N) doubles) and the flags.
constexpr size_t noalloc_bytes(size_t n)
{
    const size_t a = 256;
    auto up = [a](size_t x) { return (x + a - 1) / a * a; };
    const size_t nn = n * n, tst = (n + kNB - 1) / kNB * kNB * kNB, sth = (n + 63) / 64 * 64;
    return 6 * up(nn * 16) + up(n * 96 * 16) + up(n * kNB * 16) + up(tst * 16) + 2 * up(n * 16) +
           3 * up(n * 8) + up(nn * 8) + up(5 * n * sth * 8) + 2 * up(sizeof(int));
}
std::mutex g_noalloc_mu[kNoAllocDevices];
unsigned char *g_noalloc_ptr[kNoAllocDevices];    // device address of scratch per GPU
}  // namespace

static __device__ __align__(256) unsigned char scratch[noalloc_bytes(kNoAllocMaxN)];

extern "C" int tbk_zhegv_noalloc(int device, int n, const double *h, const double *s, double *w, double *z)
{
    if (n < 2 || n > kNoAllocMaxN || !h || !s || !w || device < 0 || device >= kNoAllocDevices)
        return TBK_ERR_ARG;
    Workspace ws;
    const size_t need = ws.carve(nullptr, n, 1);
    if (need > sizeof(scratch)) return TBK_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return TBK_ERR_NODEVICE;
    if (device >= ndev) return TBK_ERR_ARG;
    std::lock_guard<std::mutex> lk(g_noalloc_mu[device]);
    HIP_TRY(hipSetDevice(device));
    if (!g_noalloc_ptr[device]) {
        // first call on this GPU: the address of its own copy of the array
        void *ptr = nullptr;
        HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) return TBK_ERR_HIP;
        g_noalloc_ptr[device] = static_cast<unsigned char *>(ptr);
    }
    ws.carve(g_noalloc_ptr[device], n, 1);
    const size_t ss = ws.stride();
    hipStream_t st = nullptr;
    StageTimer tm;
    int info = 0;
    HIP_TRY(hipMemcpy(ws.A, h, ss * sizeof(double2), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(ws.S, s, ss * sizeof(double2), hipMemcpyHostToDevice));
    solve_batch(ws, 1, nullptr, ws.w, st, tm, Window{nullptr, 0, nullptr});
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpy(w, ws.w, n * sizeof(double), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(&info, ws.info, sizeof(int), hipMemcpyDeviceToHost));
    if (z) HIP_TRY(hipMemcpy(z, ws.Z, ss * sizeof(double2), hipMemcpyDeviceToHost));
    return info != 0 ? TBK_ERR_OVERLAP : TBK_OK;
}
