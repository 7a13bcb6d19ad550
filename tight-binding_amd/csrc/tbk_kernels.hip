// tbk_kernels.hip -- CDNA4 (gfx950) kernels of the k-point engine.
//
// Every matrix is complex128 (double2: x = re, y = im), column-major with
// leading dimension N, one N*N slab per k-point of the batch.
//
// Kernels, in pipeline order:
//   k_build_hs      Bloch sums of precomputed Slater-Koster blocks -> H, S
//                   (setham.f + rotate.f; sethamv.f for dH/dV, dS/dV)
//   k_potrf_panel   blocked Cholesky S = U^H U (zpptrf, lapack.f:4466) and
//                   U^{-1}: diagonal blocks factored and inverted in LDS,
//                   panel rows, inverse blocks and updates as GEMMs
//   k_gemm          batched complex GEMM on the FP64 / FP32 MFMA units
//   k_hetrd_panel   blocked Householder tridiagonalisation (zhptrd,
//                   lapack.f:3218) by panels; WY factors for the back-transform
//                   (k_hw_*: the same for few items of large order)
//   k_stebz, k_stein_vec   bisection eigenvalues; eigenvectors by twisted
//                   factorisation (isolated) or inverse iteration (groups)
//   k_oa_*          the refinement step of the mixed method (Ogita-Aishima,
//                   clusters rotated exactly)
//   k_coldot, k_perturb   first-order derivatives (perturb.f:87-221)
//   k_fermi_window_group  E_F and the occupation window of each k
//   k_fermi_sums    Fermi level + band-energy / pressure sums
//                   (fermi.f:163-433, bande.f:1170-1470)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <type_traits>

#include "tbk_kernels.h"

namespace tbk {

// ---------------------------------------------------------------------------
// complex helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ double2 zadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 zsub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 zmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// conj(a) * b
__device__ __forceinline__ double2 zcmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ double2 zconj(double2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ double2 zscale(double s, double2 a) { return make_double2(s * a.x, s * a.y); }
__device__ __forceinline__ double2 zzero() { return make_double2(0.0, 0.0); }
__device__ __forceinline__ float2 zadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 zsub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 zmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 zcmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float2 zconj(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ float2 zscale(float s, float2 a) { return make_float2(s * a.x, s * a.y); }

// complex type of a real type
template <class R> struct cplx;
template <> struct cplx<double> { typedef double2 type; };
template <> struct cplx<float> { typedef float2 type; };
template <class C> __device__ __forceinline__ C czero();
template <> __device__ __forceinline__ double2 czero<double2>() { return make_double2(0.0, 0.0); }
template <> __device__ __forceinline__ float2 czero<float2>() { return make_float2(0.f, 0.f); }
template <class C, class R> __device__ __forceinline__ C cmake(R x, R y);
template <> __device__ __forceinline__ double2 cmake<double2, double>(double x, double y) { return make_double2(x, y); }
template <> __device__ __forceinline__ float2 cmake<float2, float>(float x, float y) { return make_float2(x, y); }

// Block-wide sum of one double (blockDim.x multiple of 64, <= 1024).
__device__ double block_sum(double v, double *red)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[w] = v;
    __syncthreads();
    const int nw = blockDim.x >> 6;
    double s = 0.0;
    for (int i = 0; i < nw; ++i) s += red[i];   // fixed order: deterministic
    __syncthreads();
    return s;
}

// ---------------------------------------------------------------------------
// H(k), S(k) build
// ---------------------------------------------------------------------------
// Bloch phases exp(2 pi i k.d_p) for every (item, pair), d_p = dlv -
// posq_j + posq_i (rotate.f:144-149), written once per batch.
__global__ __launch_bounds__(256) void k_phases(const StructDev *sd, const int *item_s, const double *kpts,
                                                int maxpairs, double2 *ph)
{
#pragma clang fp contract(off)
    const int kb = blockIdx.y;
    const StructDev &S = sd[item_s[kb]];
    const double *ak = kpts + 3 * (size_t)kb;
    const double k0 = ak[0], k1 = ak[1], k2 = ak[2];
    const double tpi = 6.283185307179586;
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < S.npairs; p += gridDim.x * blockDim.x) {
        const double *d = S.dvec + 3 * (size_t)p;
        const double dot = k0 * d[0] + k1 * d[1] + k2 * d[2];
        double sn, cs;
        sincos(tpi * dot, &sn, &cs);
        ph[(size_t)kb * maxpairs + p] = make_double2(cs, sn);
    }
}

// One workgroup per (upper block (ia, ja), ia <= ja; chunk of kBuildItems
// items): three groups of 81 threads, one thread per element (o, b) of the
// 9x9 block with o, the row orbital, fastest, each group taking every third
// item of the chunk.  The element is the Bloch sum over the pairs (ia, ja)
// of the reference's pair list, in list order, of e(o, b) * phase
// (rotate.f:260-326); its conjugate goes to the partner element of the lower
// block (ja, ia) (or of the lower half of a diagonal block).  The
// accumulation order and the absence of FMA contraction reproduce the
// reference's rounding except for the ulp of sin/cos.  When the chunk's
// items share one structure, the block's SK entries and the items' phases
// are staged through LDS 16 pairs at a time and each is loaded once for all
// items.  The last kBuildPadWgs workgroup columns write the padding rows and
// columns nv..n-1.
// NM = 2 builds H, S (deriv = 0) or dH/dV, dS/dV (deriv = 1, sethamv.f);
// NM = 4 builds all four from the same phases in one pass.
constexpr int kBuildPadWgs = 8;
// items per workgroup: 24 for two matrices, 12 for four (the same
// accumulator registers)
template <int NM> constexpr int build_items() { return NM == 4 ? 12 : 24; }

template <int NM>
__global__ __launch_bounds__(256) void k_build_hs(BuildArgs a, int nb)
{
#pragma clang fp contract(off)
    constexpr int kBuildItems = build_items<NM>();
    const int tid = threadIdx.x, N = a.n;
    const int it0 = blockIdx.y * kBuildItems, it1 = min(nb, it0 + kBuildItems);
    const int nup = (int)gridDim.x - kBuildPadWgs;
    const bool dv = NM == 2 && a.deriv;                  // the pair of a 2-matrix build is dH, dS
    double2 *out[NM];
    out[0] = a.h; out[1] = a.s;
    if (NM == 4) { out[2] = a.dh; out[3] = a.ds; }
    if ((int)blockIdx.x >= nup) {
        const int t0 = tid + 256 * ((int)blockIdx.x - nup), ts = 256 * kBuildPadWgs;
        for (int kb = it0; kb < it1; ++kb) {
            const int nv = a.sd[a.item_s[kb]].nv;
            if (nv >= N) continue;
#pragma unroll
            for (int m = 0; m < NM; ++m) {
                const double dg = (m >= 2 || dv) ? 0.0 : (m == 0 ? a.pad : 1.0);
                double2 *M = out[m] + (size_t)kb * a.stride;
                for (int t = t0; t < N * (N - nv); t += ts) {              // columns nv..N-1
                    const int r = t % N, c = nv + t / N;
                    M[(size_t)c * N + r] = make_double2(r == c ? dg : 0.0, 0.0);
                }
                const int np = N - nv;
                for (int t = t0; t < np * nv; t += ts) {                   // rows nv..N-1 of columns < nv
                    const int r = nv + t % np, c = t / np;
                    M[(size_t)c * N + r] = make_double2(0.0, 0.0);
                }
            }
        }
        return;
    }
    const int blk = blockIdx.x;
    int ja = (int)((sqrt(8.0 * blk + 1.0) - 1.0) * 0.5);
    while ((ja + 1) * (ja + 2) / 2 <= blk) ++ja;
    while (ja * (ja + 1) / 2 > blk) --ja;
    const int ia = blk - ja * (ja + 1) / 2;
    const int il = tid / 81, e = tid - 81 * il;          // il < 3 for the 243 working threads
    const int o = e % 9, b = e / 9;
    // SK tables, on-site terms (matrices 0, 2) and diagonal constants (1, 3)
    auto table = [&](const StructDev &SD, int m) -> const double * {
        if (NM == 2) return m == 0 ? (dv ? SD.edh : SD.eh) : (dv ? SD.eds : SD.es);
        return m == 0 ? SD.eh : (m == 1 ? SD.es : (m == 2 ? SD.edh : SD.eds));
    };
    auto diag_value = [&](const StructDev &SD, int m, double acc, int cls) -> double {
        if (m == 0) return acc + (dv ? SD.dparos : SD.paros)[4 * ia + cls];
        if (m == 2) return acc + SD.dparos[4 * ia + cls];
        return acc + ((m == 1 && !dv) ? 1.0 : 0.0);
    };
    // chunk of one structure: the block's SK entries and the items' phases
    // are staged through LDS a few pairs at a time, shared by the three
    // groups (every thread takes part in the loads and barriers)
    bool chunk_same = true;
    for (int kb = it0 + 1; kb < it1; ++kb)
        if (a.item_s[kb] != a.item_s[it0]) chunk_same = false;
    if (chunk_same) {
        constexpr int PC = NM == 4 ? 8 : 16;             // pairs per stage
        // the staged SK entries, and after the sums the LDS transpose of the
        // stores, share one buffer
        constexpr int kSe = NM * PC * 81, kTr = NM * 3 * 81 * 2;
        __shared__ double smem[kSe > kTr ? kSe : kTr];
        double (*se)[PC * 81] = reinterpret_cast<double (*)[PC * 81]>(smem);
        __shared__ double2 sph[kBuildItems][PC];
        const StructDev &SD = a.sd[a.item_s[it0]];
        if (ja >= SD.natoms) return;                     // uniform over the workgroup
        const int kf = it0 + il;
        const bool active = il < 3 && kf < it1 && o < SD.nbas[ia] && b < SD.nbas[ja] && !(ia == ja && b < o);
        const int p0 = SD.blk_ptr[ia * SD.natoms + ja], p1 = SD.blk_ptr[ia * SD.natoms + ja + 1];
        const double *et[NM];
#pragma unroll
        for (int m = 0; m < NM; ++m) et[m] = table(SD, m);
        const int idx = 9 * o + b, nit = it1 - it0;
        constexpr int U = kBuildItems / 3;
        double acc[U][2 * NM];
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int q = 0; q < 2 * NM; ++q) acc[u][q] = 0.0;
        for (int pc = p0; pc < p1; pc += PC) {
            const int np = min(PC, p1 - pc);
            for (int t = tid; t < 81 * np; t += 256)
#pragma unroll
                for (int m = 0; m < NM; ++m) se[m][t] = et[m][81 * (size_t)pc + t];
            for (int t = tid; t < nit * np; t += 256) {
                const int u = t / np, q = t - u * np;
                sph[u][q] = a.ph[(size_t)(it0 + u) * a.maxpairs + pc + q];
            }
            __syncthreads();
            if (active) {
                for (int q = 0; q < np; ++q) {
                    double ev[NM];
#pragma unroll
                    for (int m = 0; m < NM; ++m) ev[m] = se[m][81 * q + idx];
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        const int kb = kf + 3 * u;
                        if (kb < it1) {
                            const double2 z = sph[kb - it0][q];
#pragma unroll
                            for (int m = 0; m < NM; ++m) {
                                acc[u][2 * m] = acc[u][2 * m] + ev[m] * z.x;
                                acc[u][2 * m + 1] = acc[u][2 * m + 1] + ev[m] * z.y;
                            }
                        }
                    }
                }
            }
            __syncthreads();
        }
        // stores, one item at a time: the upper element from its own
        // thread (row orbital fastest: 9-row column segments), the
        // conjugate partner through an LDS transpose of the block so that
        // those stores are 9-row column segments too
        double2 (*tr)[3][81] = reinterpret_cast<double2 (*)[3][81]>(smem);
        const int r0 = SD.ksk[ia], c0 = SD.ksk[ja];
        const int beta = e % 9, omega = e / 9;           // transposed role of this thread
        const bool tr_ok = il < 3 && omega < SD.nbas[ia] && beta < SD.nbas[ja] && (ia < ja || beta > omega);
        const int cls = o == 0 ? 0 : (o <= 3 ? 1 : (o <= 6 ? 2 : 3));
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int kb = kf + 3 * u;
            const bool live = active && kb < it1;
            double2 v[NM];
#pragma unroll
            for (int m = 0; m < NM; ++m) {
                v[m] = make_double2(0.0, 0.0);
                if (live) {
                    v[m] = (ia == ja && b == o) ? make_double2(diag_value(SD, m, acc[u][2 * m], cls), 0.0)
                                                : make_double2(acc[u][2 * m], acc[u][2 * m + 1]);
                    if (!a.lower_only || (ia == ja && b == o))
                        out[m][(size_t)kb * a.stride + (size_t)(c0 + b) * N + r0 + o] = v[m];
                }
                if (il < 3) tr[m][il][e] = v[m];
            }
            __syncthreads();
            if (tr_ok && kb < it1) {
                // partner of upper element (o, b) = (omega, beta): row c0 + beta, column r0 + omega
#pragma unroll
                for (int m = 0; m < NM; ++m) {
                    if (m == 1 && a.s_upper_only) continue;
                    const double2 pv = tr[m][il][omega + 9 * beta];
                    out[m][(size_t)kb * a.stride + (size_t)(r0 + omega) * N + c0 + beta] = make_double2(pv.x, -pv.y);
                }
            }
            __syncthreads();
        }
        return;
    }
    if (il >= 3) return;                                  // no barriers below
    // items it0 + il + 3u of this thread group, one at a time (several
    // structures in the chunk)
    const int cls = o == 0 ? 0 : (o <= 3 ? 1 : (o <= 6 ? 2 : 3));
    for (int kb = it0 + il; kb < it1; kb += 3) {
        const StructDev &SD = a.sd[a.item_s[kb]];
        if (ja >= SD.natoms || o >= SD.nbas[ia] || b >= SD.nbas[ja] || (ia == ja && b < o)) continue;
        const int p0 = SD.blk_ptr[ia * SD.natoms + ja], p1 = SD.blk_ptr[ia * SD.natoms + ja + 1];
        const double *et[NM];
#pragma unroll
        for (int m = 0; m < NM; ++m) et[m] = table(SD, m);
        const int idx = 9 * o + b;
        const double2 *ph = a.ph + (size_t)kb * a.maxpairs;
        double acc[2 * NM];
#pragma unroll
        for (int q = 0; q < 2 * NM; ++q) acc[q] = 0.0;
        for (int p = p0; p < p1; ++p) {
            const double2 z = ph[p];
#pragma unroll
            for (int m = 0; m < NM; ++m) {
                const double ev = et[m][81 * (size_t)p + idx];
                acc[2 * m] = acc[2 * m] + ev * z.x;
                acc[2 * m + 1] = acc[2 * m + 1] + ev * z.y;
            }
        }
        const int r0 = SD.ksk[ia], c0 = SD.ksk[ja];
#pragma unroll
        for (int m = 0; m < NM; ++m) {
            const double2 v = (ia == ja && b == o) ? make_double2(diag_value(SD, m, acc[2 * m], cls), 0.0)
                                                   : make_double2(acc[2 * m], acc[2 * m + 1]);
            double2 *M = out[m] + (size_t)kb * a.stride;
            if (!a.lower_only || (ia == ja && b == o)) M[(size_t)(c0 + b) * N + r0 + o] = v;
            if ((ia < ja || b > o) && !(m == 1 && a.s_upper_only))
                M[(size_t)(r0 + o) * N + c0 + b] = make_double2(v.x, -v.y);   // conjugate partner
        }
    }
}

// ---------------------------------------------------------------------------
// Cholesky S = U^H U and the triangular inverse, one workgroup per k
// ---------------------------------------------------------------------------
constexpr int NB = 32;

// The panel kernel's 256 threads own the elements of the upper triangle of
// the NB x NB block, column by column (element q = c (c + 1) / 2 + r, r <= c),
// q = tid, tid + 256, tid + 512: fixed places, so that no step divides.
constexpr int kPotrfThreads = 256;
constexpr int kPotrfOwn = (NB * (NB + 1) / 2 + kPotrfThreads - 1) / kPotrfThreads;
__device__ __forceinline__ void upper_place(int q, int &r, int &c)
{
    int cc = (int)((sqrtf(8.0f * (float)q + 1.0f) - 1.0f) * 0.5f);
    while (cc * (cc + 1) / 2 > q) --cc;
    while ((cc + 1) * (cc + 2) / 2 <= q) ++cc;
    c = cc;
    r = q - cc * (cc + 1) / 2;
}

// In-place upper Cholesky of a w x w block held in LDS (row stride 33).
// Right-looking on the unscaled rows: step j subtracts
// conj(D[j][r]) D[j][c] / p_j from the upper triangle of the trailing block
// (rows > j only, so row j is read and never written within the step: one
// barrier per column); the pivots p_j are kept and every row is scaled by
// 1 / sqrt(p_j) at the end.
__device__ bool lds_potrf(double2 (*D)[NB + 1], int w, const int (&orow)[kPotrfOwn], const int (&ocol)[kPotrfOwn])
{
    __shared__ double piv[NB];
    __shared__ int bad;
    if (threadIdx.x == 0) bad = 0;
    __syncthreads();
    for (int j = 0; j < w; ++j) {
        const double p = D[j][j].x;
        const bool ok = p > 0.0;
        const double pinv = ok ? 1.0 / p : 1.0;
        if (threadIdx.x == 0) {
            piv[j] = ok ? p : 1.0;
            if (!ok) bad = 1;
        }
#pragma unroll
        for (int s = 0; s < kPotrfOwn; ++s) {
            const int r = orow[s], c = ocol[s];
            if (r > j && c < w) D[r][c] = zsub(D[r][c], zscale(pinv, zcmul(D[j][r], D[j][c])));
        }
        __syncthreads();
    }
#pragma unroll
    for (int s = 0; s < kPotrfOwn; ++s) {
        const int r = orow[s], c = ocol[s];
        if (c >= w) continue;
        if (r < c) D[r][c] = zscale(1.0 / sqrt(piv[r]), D[r][c]);
        else if (r == c) D[r][r] = make_double2(sqrt(piv[r]), 0.0);
    }
    __syncthreads();
    return bad == 0;
}

// Inverse X = U^{-1} of the w x w upper triangular block in LDS into E, by
// rows from the last up: X[i][i] = 1 / U[i][i], X[i][j] = -acc_ij / U[i][i]
// with acc_ij = sum_{k=i+1..j} U[i][k] X[k][j] gathered in the owners'
// registers as the rows k below are finished (one barrier per row).
__device__ void lds_trtri(const double2 (*D)[NB + 1], double2 (*E)[NB + 1], int w, const int (&orow)[kPotrfOwn],
                          const int (&ocol)[kPotrfOwn])
{
    double2 acc[kPotrfOwn];
#pragma unroll
    for (int s = 0; s < kPotrfOwn; ++s) acc[s] = zzero();
    for (int k = w - 1; k >= 0; --k) {
        const double dinv = 1.0 / D[k][k].x;
#pragma unroll
        for (int s = 0; s < kPotrfOwn; ++s)
            if (orow[s] == k && ocol[s] < w)
                E[k][ocol[s]] = ocol[s] == k ? make_double2(dinv, 0.0) : zscale(-dinv, acc[s]);
        __syncthreads();
#pragma unroll
        for (int s = 0; s < kPotrfOwn; ++s) {
            const int r = orow[s], c = ocol[s];
            if (r < k && k <= c && c < w) acc[s] = zadd(acc[s], zmul(D[r][k], E[k][c]));
        }
    }
    __syncthreads();
}

// Cholesky, right-looking by panels of NB columns: this kernel factors the
// diagonal block S11 = U11^H U11 in LDS, inverts U11 there and stores
// U11^{-1} in place of the block.  The factor itself is not needed again:
// the caller forms the panel row U12 = U11^{-H} S12 and, later, the column
// blocks of U^{-1} as GEMMs with U11^{-1}, and the trailing update
// S22 -= U12^H U12.  The strict lower triangle of the panel's columns is
// cleared (the GEMMs read whole tiles).  info[k] = first failing column + 1.
__global__ __launch_bounds__(kPotrfThreads) void k_potrf_panel(int N, double2 *S, size_t stride, int j0, int *info)
{
    __shared__ double2 D[NB][NB + 1], E[NB][NB + 1];
    const int kb = blockIdx.x, tid = threadIdx.x;
    double2 *A = S + (size_t)kb * stride;
    if (j0 == 0) {
        if (tid == 0) info[kb] = 0;
    } else if (info[kb] != 0) {
        return;                                      // uniform per block
    }
    const int w = min(NB, N - j0);
    int orow[kPotrfOwn], ocol[kPotrfOwn];
#pragma unroll
    for (int s = 0; s < kPotrfOwn; ++s) {
        const int q = tid + kPotrfThreads * s;
        if (q < NB * (NB + 1) / 2) upper_place(q, orow[s], ocol[s]);
        else { orow[s] = NB; ocol[s] = NB; }         // no element
    }
    // eight threads per column of the block, each every eighth row
    const int cc = tid >> 3, r8 = tid & 7;
    if (cc < w)
        for (int r = r8; r < w; r += 8) D[r][cc] = A[(size_t)(j0 + cc) * N + j0 + r];
    __syncthreads();
    if (!lds_potrf(D, w, orow, ocol)) {
        if (tid == 0) info[kb] = j0 + 1;
        return;
    }
    lds_trtri(D, E, w, orow, ocol);
    // U11^{-1} into the block, the panel's columns cleared below it
    if (cc < w)
        for (int r = j0 + r8; r < N; r += 8) {
            if (r - j0 <= cc) A[(size_t)(j0 + cc) * N + r] = E[r - j0][cc];
            else A[(size_t)(j0 + cc) * N + r] = zzero();
        }
}

// ---------------------------------------------------------------------------
// Batched complex GEMM on the FP64 matrix cores
// ---------------------------------------------------------------------------
// C = alpha * op(A) * op(B) + beta * C.  BM x BN tile per workgroup of 256
// threads (4 waves in 2 x 2, each (BM/2) x (BN/2) = MI x NJ blocks of the
// 16x16x4 MFMA), K staged through LDS 16 at a time with the next slab
// prefetched into registers while the matrix cores work on the current one.
// A complex product is three real MFMAs (Gauss): with P1 = sum ar*br,
// P2 = sum ai*bi and P3 = sum (ar+ai)(br+bi), re = P1 - P2 and
// im = P3 - P1 - P2.  The sums ar+ai, br+bi are formed from the LDS operands
// just before the MFMAs, so every tile shape rounds identically.
// tri: 1 = op(A) upper triangular (skip k < row tile), 2 = op(A) lower (skip
// k beyond the row tile), 3 = op(B) upper (skip k beyond the column tile),
// 4 = only tiles touching the lower triangle of C are computed; 5 = as 4,
// and C receives the lower triangle of the product and its conjugate
// transpose above the diagonal (a Hermitian C from its lower half; alpha
// real, beta 0); 6 = 5 with op(A) lower triangular as in 2; 7 = only tiles
// touching the upper triangle of C are computed; 9 = op(A) (not transposed)
// is read as L = tril(A, -1) + diag(A) / 2 of a Hermitian A (elements above
// the diagonal as zeros, the diagonal halved; k beyond the row tile
// skipped), so that A = L + L^H and z^H A z = 2 Re(z^H L z).
typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

__device__ __forceinline__ v4d mfma16(double a, double b, v4d c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ v4f mfma16(float a, float b, v4f c)
{
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
// Row of accumulator item q in lane l of a 16x16 tile: the f64 MFMA holds
// rows l/16 + 4q, the f32 MFMA rows 4(l/16) + q (column l % 16 for both).
__device__ __forceinline__ int mfma_row(double, int lane, int q) { return (lane >> 4) + 4 * q; }
__device__ __forceinline__ int mfma_row(float, int lane, int q) { return 4 * (lane >> 4) + q; }

// Workgroups reach the 8 XCDs round-robin by linear id (XCD = id % 8), and
// each XCD has its own L2.  The output tiles of one item read the same A and
// B, so linear ids are renumbered so that every XCD receives a contiguous
// range of logical ids -- whole items -- and an item's operands are fetched
// into one L2 instead of eight.  A bijection for any grid size.
__device__ __forceinline__ unsigned xcd_logical(unsigned lin, unsigned total)
{
    const unsigned x = lin & 7u, q = total >> 3, r = total & 7u;
    return x * q + min(x, r) + (lin >> 3);
}

template <class R, int OPA, int OPB, int BM, int BN>
__global__ __launch_bounds__(256) void k_gemm(GemmArgsT<R> g)
{
    typedef typename cplx<R>::type C;
    typedef R v4 __attribute__((ext_vector_type(4)));
    constexpr int BK = 16, MI = BM / 32, NJ = BN / 32;
    constexpr int LA = BM * BK / 256, LB = BN * BK / 256;
    __shared__ R As[2][BK][BM + 1];  // [re/im][k][m]
    __shared__ R Bs[2][BK][BN + 1];  // [re/im][k][n]
    const unsigned gxy = gridDim.x * gridDim.y;
    const unsigned lg = xcd_logical(blockIdx.x + gridDim.x * blockIdx.y + gxy * blockIdx.z, gxy * gridDim.z);
    const int bm = (int)(lg % gridDim.x) * BM, bn = (int)((lg / gridDim.x) % gridDim.y) * BN, kb = (int)(lg / gxy);
    if ((g.tri == 4 || g.tri == 5 || g.tri == 6 || g.tri == 8) && bm + BM <= bn) return;
    if (g.tri == 7 && bn + BN <= bm) return;            // tile wholly below the diagonal
    if (g.ncol && bn >= g.ncol[kb]) return;             // columns outside the item's window
    const C *A = g.a + (size_t)kb * g.sa;
    const C *B = g.b + (size_t)kb * g.sb;
    C *Cm = g.c + (size_t)kb * g.sc;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int wm = (wv & 1) * (BM / 2), wn = (wv >> 1) * (BN / 2);
    int kbeg = 0, kend = g.k;
    if (g.tri == 1) kbeg = bm;
    if (g.tri == 2 || g.tri == 6 || g.tri == 9) kend = min(g.k, bm + BM);
    if (g.tri == 3) kend = min(g.k, bn + BN);
    kbeg = (kbeg / BK) * BK;
    v4 acc1[MI][NJ], acc2[MI][NJ], acc3[MI][NJ];
#pragma unroll
    for (int i = 0; i < MI; ++i)
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
            acc1[i][j] = v4{0, 0, 0, 0}; acc2[i][j] = v4{0, 0, 0, 0}; acc3[i][j] = v4{0, 0, 0, 0};
        }
    C ra[LA], rb[LB];
    auto load = [&](int k0) {
#pragma unroll
        for (int u = 0; u < LA; ++u) {
            const int t = tid + 256 * u;
            int m, kk;
            if (OPA == 0) { m = t % BM; kk = t / BM; } else { kk = t % BK; m = t / BK; }
            const int gm = bm + m, gk = k0 + kk;
            C v = czero<C>();
            if (gm < g.m && gk < kend) {
                if (OPA == 0) v = A[(size_t)gk * g.lda + gm];
                else { v = A[(size_t)gm * g.lda + gk]; if (OPA == 1) v.y = -v.y; }
                if (OPA == 0 && g.tri == 9 && gk >= gm) {
                    if (gk > gm) v = czero<C>();
                    else { v.x *= R(0.5); v.y *= R(0.5); }
                }
            }
            ra[u] = v;
        }
#pragma unroll
        for (int u = 0; u < LB; ++u) {
            const int t = tid + 256 * u;
            int n, kk;
            if (OPB == 0) { kk = t % BK; n = t / BK; } else { n = t % BN; kk = t / BN; }
            const int gn = bn + n, gk = k0 + kk;
            C v = czero<C>();
            if (gn < g.n && gk < kend) {
                if (OPB == 0) v = B[(size_t)gn * g.ldb + gk];
                else { v = B[(size_t)gk * g.ldb + gn]; if (OPB == 1) v.y = -v.y; }
            }
            rb[u] = v;
        }
    };
    if (kbeg < kend) load(kbeg);
    for (int k0 = kbeg; k0 < kend; k0 += BK) {
        __syncthreads();
#pragma unroll
        for (int u = 0; u < LA; ++u) {
            const int t = tid + 256 * u;
            int m, kk;
            if (OPA == 0) { m = t % BM; kk = t / BM; } else { kk = t % BK; m = t / BK; }
            As[0][kk][m] = ra[u].x; As[1][kk][m] = ra[u].y;
        }
#pragma unroll
        for (int u = 0; u < LB; ++u) {
            const int t = tid + 256 * u;
            int n, kk;
            if (OPB == 0) { kk = t % BK; n = t / BK; } else { n = t % BN; kk = t / BN; }
            Bs[0][kk][n] = rb[u].x; Bs[1][kk][n] = rb[u].y;
        }
        __syncthreads();
        if (k0 + BK < kend) load(k0 + BK);
        // not unrolled: the operands of one k step at a time keep the
        // register count low enough for several workgroups per CU
#pragma unroll 1
        for (int ks = 0; ks < BK; ks += 4) {
            const int kk = ks + (lane >> 4);
            R ar[MI], ai[MI], as[MI], br[NJ], bi[NJ], bs[NJ];
#pragma unroll
            for (int i = 0; i < MI; ++i) {
                ar[i] = As[0][kk][wm + 16 * i + (lane & 15)];
                ai[i] = As[1][kk][wm + 16 * i + (lane & 15)];
                as[i] = ar[i] + ai[i];
            }
#pragma unroll
            for (int j = 0; j < NJ; ++j) {
                br[j] = Bs[0][kk][wn + 16 * j + (lane & 15)];
                bi[j] = Bs[1][kk][wn + 16 * j + (lane & 15)];
                bs[j] = br[j] + bi[j];
            }
#pragma unroll
            for (int i = 0; i < MI; ++i)
#pragma unroll
                for (int j = 0; j < NJ; ++j) {
                    acc1[i][j] = mfma16(ar[i], br[j], acc1[i][j]);
                    acc2[i][j] = mfma16(ai[i], bi[j], acc2[i][j]);
                    acc3[i][j] = mfma16(as[i], bs[j], acc3[i][j]);
                }
        }
    }
    const bool usebeta = g.beta.x != R(0) || g.beta.y != R(0);
#pragma unroll
    for (int i = 0; i < MI; ++i)
#pragma unroll
        for (int j = 0; j < NJ; ++j)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int r = bm + wm + 16 * i + mfma_row(R(0), lane, q);
                const int c = bn + wn + 16 * j + (lane & 15);
                if (r < g.m && c < g.n) {
                    const R p1 = acc1[i][j][q], p2 = acc2[i][j][q];
                    const C v = cmake<C, R>(p1 - p2, acc3[i][j][q] - p1 - p2);
                    C out = zmul(g.alpha, v);
                    C *p = Cm + (size_t)c * g.ldc + r;
                    typedef typename GemmArgsT<R>::CX CX;
                    typedef typename std::conditional<sizeof(R) == 8, float, double>::type RX;
                    CX *Xm = g.cx ? g.cx + (size_t)kb * g.scx : nullptr;
                    if (g.tri == 5 || g.tri == 6 || g.tri == 8) {
                        // Hermitian from the lower half: element (r, c), r >= c,
                        // and its conjugate at (c, r); 16 lanes of one row write
                        // 16 consecutive rows of column r (tri 8: with beta C)
                        if (g.tri == 8 && r >= c && usebeta) out = zadd(out, zmul(g.beta, *p));
                        if (r >= c) {
                            *p = out;
                            if (Xm) Xm[(size_t)c * g.ldcx + r] = cmake<CX, RX>((RX)out.x, (RX)out.y);
                            if (r > c && r < g.n) {
                                Cm[(size_t)r * g.ldc + c] = cmake<C, R>(out.x, -out.y);
                                if (Xm) Xm[(size_t)r * g.ldcx + c] = cmake<CX, RX>((RX)out.x, (RX)-out.y);
                            }
                        }
                        continue;
                    }
                    if (usebeta) out = zadd(out, zmul(g.beta, *p));
                    *p = out;
                    if (Xm) Xm[(size_t)c * g.ldcx + r] = cmake<CX, RX>((RX)out.x, (RX)out.y);
                }
            }
}

// ---------------------------------------------------------------------------
// Householder tridiagonalisation (lower, blocked): one panel per launch
// ---------------------------------------------------------------------------
// k_hetrd_panel reduces columns i0 .. i0+pw-1 of every matrix of the batch
// (one workgroup of 8 waves per k), following zlatrd (delayed update of the
// trailing matrix through V and W).  It leaves
//   P = [V | W | V]   (n x 96, column-major) for the trailing update
//       A[s:, s:] -= V W^H + W V^H, which the caller runs as one GEMM,
//   the reflectors in A (A[j+1, j] = 1, v below, zero at and above j),
//   the panel's upper triangular WY factor T (Q_p = I - V T V^H),
//   d[j], e[j] (the real tridiagonal).
// Only the lower triangle of the trailing matrix is read or updated: the
// Hermitian product y = A_t v is formed column by column from A[r:, r], the
// wave owning column r taking the dot conj(A[r:, r])^T v[r:] (-> y_r) and
// accumulating A[r+1:, r] v_r into its private partial y (registers).  The
// partials are combined in a fixed order, so results are deterministic.
constexpr int HT = 512;            // threads of the panel kernel
constexpr int HW = HT / 64;        // waves

// Sum over the 64 lanes of a wave of K values per lane (K a power of two,
// K <= 32), transposed: at each halving step a lane keeps half of its values
// and adds the partner's copies of them, so K - 1 + 6 - log2 K exchanges
// replace 6 K.  Lane l returns the total of value (l >> (6 - log2 K)).
template <int K> struct ilog2c { static constexpr int v = K <= 1 ? 0 : 1 + ilog2c<K / 2>::v; };
template <> struct ilog2c<1> { static constexpr int v = 0; };
template <class R, int K>
__device__ __forceinline__ R wave_sum_transposed(R (&v)[K], int lane)
{
    constexpr int LK = ilog2c<K>::v;
#pragma unroll
    for (int s = 0; s < LK; ++s) {
        const int off = 32 >> s;
        const bool up = (lane & off) != 0;
        const int cnt = K >> (s + 1);
#pragma unroll
        for (int i = 0; i < cnt; ++i) {
            // the two values as plain registers: otherwise the lane-dependent
            // choice between them becomes an indexed read of the array
            // (a compare/select chain over all K entries)
            R lo = v[i], hi = v[cnt + i];
            asm volatile("" : "+v"(lo), "+v"(hi));
            const R keep = up ? hi : lo;
            const R send = up ? lo : hi;
            v[i] = keep + __shfl_xor(send, off, 64);
        }
    }
#pragma unroll
    for (int off = 32 >> LK; off > 0; off >>= 1) v[0] += __shfl_xor(v[0], off, 64);
    return v[0];
}

__device__ __forceinline__ double2 wave_sum2(double2 v)
{
    for (int off = 32; off > 0; off >>= 1) {
        v.x += __shfl_xor(v.x, off, 64);
        v.y += __shfl_xor(v.y, off, 64);
    }
    return v;
}

// dynamic LDS of k_hetrd_panel: vloc[N], ybuf[N], partial-y slots [HS][N]
// (one per wave up to order 320, else combined in pairs into four)
template <int MAXQ> constexpr int hetrd_slots() { return MAXQ <= 5 ? 8 : 4; }
template <class R, int MAXQ> constexpr size_t hetrd_lds_bytes(int n)
{
    return (size_t)(2 + hetrd_slots<MAXQ>()) * n * 2 * sizeof(R);
}

template <class R, int MAXQ>
__global__ __launch_bounds__(HT, (MAXQ <= 5 ? 4 : 2)) void k_hetrd_panel(HetrdArgsT<R> h, int i0)
{
    typedef typename cplx<R>::type C;
    extern __shared__ double4 hetrd_smem[];
    C *vloc = reinterpret_cast<C *>(hetrd_smem);     // reflector of the current column
    C *ybuf = vloc + h.n;                              // dot part of A_t v
    C *slot = ybuf + h.n;                              // [HS][N] partial sums of A_t v
    constexpr int HS = hetrd_slots<MAXQ>();
    __shared__ C udot[2 * NB];         // W^H v (0..31) and V^H v (32..63)
    __shared__ C uwn[NB], uvn[NB];     // conj(W[j+1, l]), conj(V[j+1, l])
    __shared__ C Ts[NB][NB + 1];       // WY factor T of the panel, Ts[col][row]
    // separate reduction scratch for the norm (1) and for w^H v (4), so that
    // neither phase has to wait for the other's readers
    __shared__ R red1[HW], red4[2 * HW];
    __shared__ C scal[3];              // tau, (beta, 0), scale of the reflector
    const int N = h.n;
    const int kb = blockIdx.x;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    C *A = h.a + (size_t)kb * h.stride;
    C *P = h.vp + (size_t)kb * N * 96;
    C *Tp = h.t + (size_t)kb * h.tstride + (size_t)(i0 / NB) * NB * NB;
    double *d = h.d + (size_t)kb * N;
    double *e = h.e + (size_t)kb * N;
    const int pw = min(NB, N - 1 - i0);
    C *V = P, *W = P + (size_t)NB * N, *V2 = P + (size_t)2 * NB * N;
    for (int t = tid; t < N * 96; t += HT) P[t] = czero<C>();
    for (int t = tid; t < NB * (NB + 1); t += HT) Ts[t / (NB + 1)][t % (NB + 1)] = czero<C>();
    // Rows are owned by threads in a fixed map (r = tid + HT s); carry[s] is
    // the delayed update V[r,:] conj(W[j,:]) + W[r,:] conj(V[j,:]) of the
    // current column j, accumulated while the previous column's w was formed.
    constexpr int RPT = (64 * MAXQ + HT - 1) / HT;   // rows per thread
    C carry[RPT];
#pragma unroll
    for (int q = 0; q < RPT; ++q) carry[q] = czero<C>();
    __syncthreads();
    for (int i = 0; i < pw; ++i) {
        const int j = i0 + i, t0 = j + 1, m = N - t0;
        // (1) column j brought up to date: a_r = A[r, j] - carry_r
        R xs = 0;
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
            const int r = tid + HT * q;
            if (r < j || r >= N) continue;
            C a = zsub(A[(size_t)j * N + r], carry[q]);
            if (r == j) a.y = 0;
            vloc[r] = a;
            if (r >= j + 2) xs += a.x * a.x + a.y * a.y;
        }
        for (int off = 32; off > 0; off >>= 1) xs += __shfl_xor(xs, off, 64);
        if (lane == 0) red1[wv] = xs;
        __syncthreads();
        // (2) reflector (zlarfg) for x = a[j+1:], with the scale 1/(alpha - beta)
        //     of its tail, so that no other thread needs vloc[j + 1]
        if (tid == 0) {
            R xn = 0;
            for (int q = 0; q < HW; ++q) xn += red1[q];
            const C al = vloc[j + 1];
            C tau = czero<C>();
            R beta = al.x;
            if (!(xn == R(0) && al.y == R(0))) {
                beta = -copysign(sqrt(al.x * al.x + al.y * al.y + xn), al.x);
                tau = cmake<C, R>((beta - al.x) / beta, -al.y / beta);
            }
            const C dv = cmake<C, R>(al.x - beta, al.y);
            const R den = dv.x * dv.x + dv.y * dv.y;
            const bool refl = tau.x != R(0) || tau.y != R(0);
            scal[0] = tau; scal[1] = cmake<C, R>(beta, R(0));
            scal[2] = refl ? cmake<C, R>(dv.x / den, -dv.y / den) : cmake<C, R>(R(1), R(0));
            d[j] = (double)vloc[j].x;
            e[j] = (double)beta;
        }
        __syncthreads();
        const C tau = scal[0];
        {
            const C sc = scal[2];
            for (int r = tid; r < N; r += HT) {     // rows owned as in (1)
                C v;
                if (r <= j) v = czero<C>();
                else if (r == j + 1) v = cmake<C, R>(R(1), R(0));
                else v = zmul(vloc[r], sc);
                vloc[r] = v;
                A[(size_t)j * N + r] = v;
                V[(size_t)i * N + r] = v;
                V2[(size_t)i * N + r] = v;
            }
        }
        __syncthreads();
        // (3) the 2i panel dots W[:,l]^H v, V[:,l]^H v over rows >= t0 (one
        //     wave task each), then y = A_t v from the lower triangle (columns
        //     r >= t0); udot[l] = W[:,l]^H v, udot[NB + l] = V[:,l]^H v
        {
            // kTg tasks per wave pass, every row chunk of a task loaded at once
            constexpr int kTg = 1, QT = MAXQ <= 5 ? MAXQ : 5;
            const int nqt = (N + 63) >> 6, qlo = t0 >> 6;
            const int wvu = __builtin_amdgcn_readfirstlane(wv);
            for (int tk = wvu; tk < 2 * i; tk += HW * kTg) {
                const C *colp[kTg];
                C acc[kTg];
#pragma unroll
                for (int g = 0; g < kTg; ++g) {
                    const int task = min(tk + HW * g, 2 * i - 1);
                    colp[g] = task < i ? W + (size_t)task * N : V + (size_t)(task - i) * N;
                    acc[g] = czero<C>();
                }
                for (int q0 = qlo; q0 < nqt; q0 += QT) {
                    C a[kTg][QT];
#pragma unroll
                    for (int u = 0; u < QT; ++u) {
                        const int c = ((q0 + u) << 6) + lane;
#pragma unroll
                        for (int g = 0; g < kTg; ++g)
                            a[g][u] = (tk + HW * g < 2 * i && c >= t0 && c < N) ? colp[g][c] : czero<C>();
                    }
#pragma unroll
                    for (int u = 0; u < QT; ++u) {
                        const int c = ((q0 + u) << 6) + lane;
                        const C vc = (c >= t0 && c < N) ? vloc[c] : czero<C>();
#pragma unroll
                        for (int g = 0; g < kTg; ++g) acc[g] = zadd(acc[g], zcmul(a[g][u], vc));
                    }
                }
                for (int off = 32; off > 0; off >>= 1)
#pragma unroll
                    for (int g = 0; g < kTg; ++g) {
                        acc[g].x += __shfl_xor(acc[g].x, off, 64);
                        acc[g].y += __shfl_xor(acc[g].y, off, 64);
                    }
                if (lane == 0)
#pragma unroll
                    for (int g = 0; g < kTg; ++g) {
                        const int task = tk + HW * g;
                        if (task < 2 * i) udot[task < i ? task : NB + task - i] = acc[g];
                    }
            }
        }
        if (tid < i && t0 < N) {
            uwn[tid] = zconj(W[(size_t)tid * N + t0]);
            uvn[tid] = zconj(V[(size_t)tid * N + t0]);
        }
        C ypart[MAXQ];
#pragma unroll
        for (int q = 0; q < MAXQ; ++q) ypart[q] = czero<C>();
        const int nq = (N + 63) >> 6;
        // columns r = t0 + tb + HW g (g < HG) per wave pass: HG x QC loads in
        // flight per lane and HG independent wave reductions.  Up to order
        // 640 a column's chunks are all loaded at once.  Row chunks of 64
        // that lie wholly above the pass's first column hold no element of
        // the lower triangle and are skipped by a uniform branch (loads
        // first, then the arithmetic, so the loads of a pass stay in flight
        // together): the product's cost follows the triangle, not the square.
        constexpr int HG = sizeof(R) == 8 ? 2 : 4;
        constexpr int QC = sizeof(R) == 8 ? 3 : (MAXQ <= 5 ? MAXQ : (MAXQ <= 10 ? 5 : 4));
        // the wave index is uniform: columns and their base addresses live in
        // scalar registers, each load needs only the lane's row offset
        const int wvu = __builtin_amdgcn_readfirstlane(wv);
        for (int tb = wvu; tb < m; tb += HW * HG) {
            C acc[HG], vr[HG];
            int rg[HG];
#pragma unroll
            for (int g = 0; g < HG; ++g) {
                rg[g] = t0 + tb + HW * g;
                acc[g] = czero<C>();
                vr[g] = rg[g] < N ? vloc[rg[g]] : czero<C>();
            }
            const int qlo = rg[0] >> 6;                // first chunk holding rows >= rg[0]
#pragma unroll
            for (int q0 = 0; q0 < MAXQ; q0 += QC) {
                if (q0 < nq && q0 + QC > qlo) {
                    C a[HG][QC];
#pragma unroll
                    for (int u = 0; u < QC; ++u) {
                        if (q0 + u >= qlo && q0 + u < nq) {
#pragma unroll
                            for (int g = 0; g < HG; ++g) {
                                const C *colg = A + (size_t)rg[g] * N;
                                const int c = ((q0 + u) << 6) + lane;
                                a[g][u] = (c >= rg[g] && c < N) ? colg[c] : czero<C>();
                            }
                        }
                    }
#pragma unroll
                    for (int u = 0; u < QC; ++u) {
                        const int c = ((q0 + u) << 6) + lane;
                        if (q0 + u < MAXQ && q0 + u >= qlo && q0 + u < nq) {
                            const C vc = c < N ? vloc[c] : czero<C>();
#pragma unroll
                            for (int g = 0; g < HG; ++g) {
                                acc[g] = zadd(acc[g], zcmul(a[g][u], vc));
                                if (c > rg[g]) ypart[q0 + u] = zadd(ypart[q0 + u], zmul(a[g][u], vr[g]));
                            }
                        }
                    }
                }
            }
            if constexpr (sizeof(R) == 8) {
                // FP64: the 2 HG sums (re, im of each column) reduced together;
                // lane l holds value (l >> sh): column value / 2, component value % 2
                R vals[2 * HG];
#pragma unroll
                for (int g = 0; g < HG; ++g) { vals[2 * g] = acc[g].x; vals[2 * g + 1] = acc[g].y; }
                const R tot = wave_sum_transposed<R, 2 * HG>(vals, lane);
                constexpr int sh = 6 - ilog2c<2 * HG>::v;
                const int val = lane >> sh, rgl = t0 + tb + HW * (val >> 1);
                if ((lane & ((1 << sh) - 1)) == 0 && rgl < N)
                    reinterpret_cast<R *>(ybuf)[2 * (size_t)rgl + (val & 1)] = tot;
            } else {
                // FP32 (more columns per pass): plain butterflies, which need
                // fewer registers here
                for (int off = 32; off > 0; off >>= 1)
#pragma unroll
                    for (int g = 0; g < HG; ++g) {
                        acc[g].x += __shfl_xor(acc[g].x, off, 64);
                        acc[g].y += __shfl_xor(acc[g].y, off, 64);
                    }
                if (lane == 0)
#pragma unroll
                    for (int g = 0; g < HG; ++g)
                        if (rg[g] < N) ybuf[rg[g]] = acc[g];
            }
        }
        // partials: one slot per wave, or slot[w] = part_w + part_{w+4}
        if (HS == HW || wv >= 4) {
            const int sw = HS == HW ? wv : wv - 4;
#pragma unroll
            for (int q = 0; q < MAXQ; ++q) {
                const int c = (q << 6) + lane;
                if (q < nq && c >= t0 && c < N) slot[(size_t)sw * N + c] = ypart[q];
            }
        }
        __syncthreads();
        if (HS != HW) {
            if (wv < 4) {
#pragma unroll
                for (int q = 0; q < MAXQ; ++q) {
                    const int c = (q << 6) + lane;
                    if (q < nq && c >= t0 && c < N) {
                        const size_t o = (size_t)wv * N + c;
                        slot[o] = zadd(ypart[q], slot[o]);
                    }
                }
            }
            __syncthreads();
        }
        // (4) w = tau (y - V u1 - W u2), u1 = W^H v, u2 = V^H v; then the
        //     correction w += alpha v with alpha = -tau/2 (w^H v).  The same
        //     pass over V[r, 0:i], W[r, 0:i] accumulates the next column's
        //     delayed update into carry.
        R pr = 0, pi = 0;
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
            const int r = tid + HT * q;
            if (r < t0 || r >= N) continue;
            C y = ybuf[r];
#pragma unroll
            for (int sw = 0; sw < HS; ++sw) y = zadd(y, slot[(size_t)sw * N + r]);
            C nx = czero<C>();
            // the row's V and W entries are loaded kWst columns at a time
            // before they are used, so their loads are in flight together
            constexpr int kWst = 2;
            for (int l0 = 0; l0 < i; l0 += kWst) {
                C vv[kWst], ww[kWst];
#pragma unroll
                for (int u = 0; u < kWst; ++u)
                    if (l0 + u < i) { vv[u] = V[(size_t)(l0 + u) * N + r]; ww[u] = W[(size_t)(l0 + u) * N + r]; }
#pragma unroll
                for (int u = 0; u < kWst; ++u) {
                    const int l = l0 + u;
                    if (l < i) {
                        y = zsub(y, zmul(vv[u], udot[l]));
                        y = zsub(y, zmul(ww[u], udot[NB + l]));
                        nx = zadd(nx, zmul(vv[u], uwn[l]));
                        nx = zadd(nx, zmul(ww[u], uvn[l]));
                    }
                }
            }
            carry[q] = nx;
            const C w = zmul(tau, y);
            ybuf[r] = w;
            const C qv = zcmul(w, vloc[r]);
            pr += qv.x; pi += qv.y;
        }
        for (int off = 32; off > 0; off >>= 1) {
            pr += __shfl_xor(pr, off, 64);
            pi += __shfl_xor(pi, off, 64);
        }
        if (lane == 0) { red4[wv] = pr; red4[HW + wv] = pi; }
        __syncthreads();
        R sr = 0, si = 0;
        for (int q = 0; q < HW; ++q) { sr += red4[q]; si += red4[HW + q]; }
        const C alpha = zscale(R(-0.5), zmul(tau, cmake<C, R>(sr, si)));
        // W[j+1, i] (v_{j+1} = 1) for the l = i term of the next column's update
        const C wnext = t0 < N ? zadd(ybuf[t0], alpha) : czero<C>();
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
            const int r = tid + HT * q;
            if (r < t0 || r >= N) continue;
            const C wr = zadd(ybuf[r], zmul(alpha, vloc[r]));
            W[(size_t)i * N + r] = wr;
            carry[q] = zadd(carry[q], zadd(zmul(vloc[r], zconj(wnext)), wr));
        }
        // (5) WY factor: T[0:i, i] = -tau T[0:i, 0:i] (V^H v), T[i, i] = tau
        if (tid < i) {
            C acc = czero<C>();
            for (int l = tid; l < i; ++l) acc = zadd(acc, zmul(Ts[l][tid], udot[NB + l]));
            Ts[i][tid] = zmul(cmake<C, R>(-tau.x, -tau.y), acc);
        }
        if (tid == 0) Ts[i][i] = tau;
        // no barrier here: the next column's first reads of anything written
        // above (Ts, W, udot) follow its own barriers
    }
    __syncthreads();
    for (int t = tid; t < NB * NB; t += HT) Tp[t] = Ts[t / NB][t % NB];
    // V T^H of the panel (rows > i0), for the back-transform
    // Q_p Y = Y - V ((V T^H)^H Y)
    C *VT = h.vt + (size_t)kb * h.stride + (size_t)i0 * N;
    for (int t = tid; t < N * NB; t += HT) {
        const int r = t % N, c = t / N;
        C acc = czero<C>();
        if (c < pw && r > i0)
            for (int l = c; l < pw; ++l) acc = zadd(acc, zmul(V[(size_t)l * N + r], zconj(Ts[l][c])));
        if (i0 + c < N) VT[(size_t)c * N + r] = acc;
    }
}

// ---------------------------------------------------------------------------
// Householder tridiagonalisation, multi-workgroup form for few items of
// large order: per column three launches -- column update and reflector
// (one workgroup per item), the Hermitian product over 64x64 tiles of the
// lower triangle (many workgroups per item, partial sums per tile row and
// tile column), then w, the WY factor and the delayed-update panel (one
// workgroup per item).  Same algorithm and outputs as k_hetrd_panel; the
// partial sums are combined in a fixed order.
// ---------------------------------------------------------------------------
constexpr int WT = 1024;           // threads of the per-item kernels

template <class R>
__device__ __forceinline__ R block_sum_r(R v, R *red)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wv] = v;
    __syncthreads();
    R s = 0;
    for (int q = 0; q < (int)(blockDim.x >> 6); ++q) s += red[q];
    __syncthreads();
    return s;
}

template <class R>
__device__ void hw_prep_body(const HetrdArgsT<R> &h, const WideArgsT<R> &w, int i0, int i)
{
    typedef typename cplx<R>::type C;
    __shared__ R red[WT / 64];
    __shared__ C scal[2];
    const int N = h.n, kb = blockIdx.x, tid = threadIdx.x;
    const int j = i0 + i;
    C *A = h.a + (size_t)kb * h.stride;
    C *P = h.vp + (size_t)kb * N * 96;
    C *V = P, *W = P + (size_t)NB * N, *V2 = P + (size_t)2 * NB * N;
    C *Tp = h.t + (size_t)kb * h.tstride + (size_t)(i0 / NB) * NB * NB;
    C *vv = w.v + (size_t)kb * N;
    C *col = w.col + (size_t)kb * N;
    if (i == 0) {
        for (int t = tid; t < N * 96; t += WT) P[t] = czero<C>();
        for (int t = tid; t < NB * NB; t += WT) Tp[t] = czero<C>();
        __syncthreads();
    }
    // delayed update of column j: carry = V[r,:] conj(W[j,:]) + W[r,:] conj(V[j,:]),
    // accumulated by k_hw_fin of the previous column
    const C *carry = w.carry + (size_t)kb * N;
    R xs = 0;
    for (int r = j + tid; r < N; r += WT) {
        C a = A[(size_t)j * N + r];
        if (i > 0) a = zsub(a, carry[r]);
        if (r == j) a.y = 0;
        col[r] = a;
        if (r >= j + 2) xs += a.x * a.x + a.y * a.y;
    }
    xs = block_sum_r(xs, red);
    if (tid == 0) {
        const C al = col[j + 1];
        C tau = czero<C>();
        R beta = al.x;
        if (!(xs == R(0) && al.y == R(0))) {
            beta = -copysign(sqrt(al.x * al.x + al.y * al.y + xs), al.x);
            tau = cmake<C, R>((beta - al.x) / beta, -al.y / beta);
        }
        scal[0] = tau; scal[1] = cmake<C, R>(beta, R(0));
        w.scal[2 * (size_t)kb] = tau;
        h.d[(size_t)kb * N + j] = (double)col[j].x;
        h.e[(size_t)kb * N + j] = (double)beta;
    }
    __syncthreads();
    const C tau = scal[0];
    const R beta = scal[1].x;
    const C al = col[j + 1];
    const C dv = cmake<C, R>(al.x - beta, al.y);
    const R den = dv.x * dv.x + dv.y * dv.y;
    const bool refl = tau.x != R(0) || tau.y != R(0);
    const C sc = refl ? cmake<C, R>(dv.x / den, -dv.y / den) : cmake<C, R>(R(1), R(0));
    for (int r = tid; r < N; r += WT) {
        C v;
        if (r <= j) v = czero<C>();
        else if (r == j + 1) v = cmake<C, R>(R(1), R(0));
        else v = zmul(col[r], sc);
        A[(size_t)j * N + r] = v;
        V[(size_t)i * N + r] = v;
        V2[(size_t)i * N + r] = v;
        vv[r] = v;
    }
}

// Tile (tr, tc), tr >= tc, of the lower triangle of the trailing matrix
// (rows/columns >= t0): column dots conj(A[r, c]) v_r over the tile's rows
// r >= c into ydot[tr][c]; row sums A[r, c] v_c over the tile's columns
// c < r into yax[tc][r].  The extra workgroup blockIdx.x == ntile forms the
// panel dots W[:, l]^H v and V[:, l]^H v.
template <class R>
__global__ __launch_bounds__(256) void k_hw_hemv(HetrdArgsT<R> h, WideArgsT<R> w, int i0, int i, int ntile)
{
    typedef typename cplx<R>::type C;
    __shared__ C acc_s[4][64];
    const int N = h.n, kb = blockIdx.y, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int j = i0 + i, t0 = j + 1, m = N - t0;
    const int T = (m + 63) >> 6;
    const C *A = h.a + (size_t)kb * h.stride;
    const C *vv = w.v + (size_t)kb * N;
    const size_t tstr = (size_t)w.tmax * N;
    if ((int)blockIdx.x == ntile) {
        if (i == 0) return;
        const C *P = h.vp + (size_t)kb * N * 96;
        const C *V = P, *W = P + (size_t)NB * N;
        for (int task = wv; task < 2 * i; task += 4) {
            const C *colp = task < i ? W + (size_t)task * N : V + (size_t)(task - i) * N;
            C acc = czero<C>();
            for (int c = t0 + lane; c < N; c += 64) acc = zadd(acc, zcmul(colp[c], vv[c]));
            for (int off = 32; off > 0; off >>= 1) {
                acc.x += __shfl_xor(acc.x, off, 64);
                acc.y += __shfl_xor(acc.y, off, 64);
            }
            if (lane == 0) w.udot[(size_t)kb * 2 * NB + (task < i ? task : NB + task - i)] = acc;
        }
        return;
    }
    const int tile = blockIdx.x;
    if (tile >= T * (T + 1) / 2) return;
    int tr = (int)((sqrt(8.0 * tile + 1.0) - 1.0) * 0.5);
    while ((tr + 1) * (tr + 2) / 2 <= tile) ++tr;
    while (tr * (tr + 1) / 2 > tile) --tr;
    const int tc = tile - tr * (tr + 1) / 2;
    const int r = t0 + 64 * tr + lane;
    const C vr = r < N ? vv[r] : czero<C>();
    C acc = czero<C>();
    C *ydot = w.ydot + (size_t)kb * tstr + (size_t)tr * N;
    for (int u = 0; u < 16; ++u) {
        const int c = t0 + 64 * tc + 16 * wv + u;
        if (c >= N) break;
        const C a = r < N ? A[(size_t)c * N + r] : czero<C>();
        C pd = r >= c ? zcmul(a, vr) : czero<C>();
        if (c < r) acc = zadd(acc, zmul(a, vv[c]));
        for (int off = 32; off > 0; off >>= 1) {
            pd.x += __shfl_xor(pd.x, off, 64);
            pd.y += __shfl_xor(pd.y, off, 64);
        }
        if (lane == 0) ydot[c] = pd;
    }
    acc_s[wv][lane] = acc;
    __syncthreads();
    if (wv == 0 && r < N)
        w.yax[(size_t)kb * tstr + (size_t)tc * N + r] =
            zadd(zadd(zadd(acc_s[0][lane], acc_s[1][lane]), acc_s[2][lane]), acc_s[3][lane]);
}

template <class R>
__device__ void hw_fin_body(const HetrdArgsT<R> &h, const WideArgsT<R> &w, int i0, int i)
{
    typedef typename cplx<R>::type C;
    __shared__ C udot[2 * NB];
    __shared__ C uwn[NB], uvn[NB];     // conj(W[j+1, l]), conj(V[j+1, l])
    __shared__ R red[2 * (WT / 64)];
    const int N = h.n, kb = blockIdx.x, tid = threadIdx.x;
    const int j = i0 + i, t0 = j + 1, m = N - t0;
    const int T = (m + 63) >> 6;
    C *P = h.vp + (size_t)kb * N * 96;
    C *V = P, *W = P + (size_t)NB * N;
    C *Tp = h.t + (size_t)kb * h.tstride + (size_t)(i0 / NB) * NB * NB;
    const C *vv = w.v + (size_t)kb * N;
    C *col = w.col + (size_t)kb * N;
    const size_t tstr = (size_t)w.tmax * N;
    const C *ydot = w.ydot + (size_t)kb * tstr, *yax = w.yax + (size_t)kb * tstr;
    if (tid < 2 * NB) udot[tid] = w.udot[(size_t)kb * 2 * NB + tid];
    if (tid < i && t0 < N) { uwn[tid] = zconj(W[(size_t)tid * N + t0]); uvn[tid] = zconj(V[(size_t)tid * N + t0]); }
    const C tau = w.scal[2 * (size_t)kb];
    __syncthreads();
    // w = tau (y - V u1 - W u2); the same pass over V[r, 0:i], W[r, 0:i]
    // accumulates the next column's delayed update into carry
    C *carry = w.carry + (size_t)kb * N;
    R pr = 0, pi = 0;
    for (int r = t0 + tid; r < N; r += WT) {
        const int tr0 = (r - t0) >> 6;
        C y = czero<C>(), nx = czero<C>();
        for (int tr = tr0; tr < T; ++tr) y = zadd(y, ydot[(size_t)tr * N + r]);
        for (int tc = 0; tc <= tr0; ++tc) y = zadd(y, yax[(size_t)tc * N + r]);
#pragma unroll 4
        for (int l = 0; l < i; ++l) {
            const C vl = V[(size_t)l * N + r], wl = W[(size_t)l * N + r];
            y = zsub(y, zmul(vl, udot[l]));
            y = zsub(y, zmul(wl, udot[NB + l]));
            nx = zadd(nx, zmul(vl, uwn[l]));
            nx = zadd(nx, zmul(wl, uvn[l]));
        }
        carry[r] = nx;
        const C wr = zmul(tau, y);
        col[r] = wr;
        const C q = zcmul(wr, vv[r]);
        pr += q.x; pi += q.y;
    }
    const R sr = block_sum_r(pr, red), si = block_sum_r(pi, red + WT / 64);
    const C alpha = zscale(R(-0.5), zmul(tau, cmake<C, R>(sr, si)));
    const C wnext = t0 < N ? zadd(col[t0], alpha) : czero<C>();     // W[j+1, i] (v_{j+1} = 1)
    for (int r = t0 + tid; r < N; r += WT) {
        const C wr = zadd(col[r], zmul(alpha, vv[r]));
        W[(size_t)i * N + r] = wr;
        carry[r] = zadd(carry[r], zadd(zmul(vv[r], zconj(wnext)), wr));
    }
    // WY factor: T[0:i, i] = -tau T[0:i, 0:i] (V^H v), T[i, i] = tau
    if (tid < i) {
        C acc = czero<C>();
        for (int l = tid; l < i; ++l) acc = zadd(acc, zmul(Tp[(size_t)l * NB + tid], udot[NB + l]));
        Tp[(size_t)i * NB + tid] = zmul(cmake<C, R>(-tau.x, -tau.y), acc);
    }
    if (tid == 0) Tp[(size_t)i * NB + i] = tau;
}

template <class R>
__global__ __launch_bounds__(WT) void k_hw_prep(HetrdArgsT<R> h, WideArgsT<R> w, int i0, int i)
{
    hw_prep_body(h, w, i0, i);
}
// w, WY column and carry of column i, then the reflector of column i + 1
// (same workgroup, one launch)
template <class R>
__global__ __launch_bounds__(WT) void k_hw_finprep(HetrdArgsT<R> h, WideArgsT<R> w, int i0, int i, int next)
{
    hw_fin_body(h, w, i0, i);
    if (next) {
        __syncthreads();
        hw_prep_body(h, w, i0, i + 1);
    }
}

// V T^H of the panel (rows > i0), for the back-transform.
template <class R>
__global__ __launch_bounds__(256) void k_hw_vt(HetrdArgsT<R> h, int i0)
{
    typedef typename cplx<R>::type C;
    const int N = h.n, kb = blockIdx.y;
    const int pw = min(NB, N - 1 - i0);
    const C *V = h.vp + (size_t)kb * N * 96;
    const C *Tp = h.t + (size_t)kb * h.tstride + (size_t)(i0 / NB) * NB * NB;
    C *VT = h.vt + (size_t)kb * h.stride + (size_t)i0 * N;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < N * NB; t += gridDim.x * blockDim.x) {
        const int r = t % N, c = t / N;
        C acc = czero<C>();
        if (c < pw && r > i0)
            for (int l = c; l < pw; ++l) acc = zadd(acc, zmul(V[(size_t)l * N + r], zconj(Tp[(size_t)l * NB + c])));
        if (i0 + c < N) VT[(size_t)c * N + r] = acc;
    }
}

// After the last panel's trailing update: d[n-1] and the empty last column.
template <class R>
__global__ void k_hetrd_finish(HetrdArgsT<R> h)
{
    typedef typename cplx<R>::type C;
    const int N = h.n, kb = blockIdx.x;
    C *A = h.a + (size_t)kb * h.stride;
    if (threadIdx.x == 0) {
        h.d[(size_t)kb * N + N - 1] = (double)A[(size_t)(N - 1) * N + N - 1].x;
        h.e[(size_t)kb * N + N - 1] = 0.0;
    }
    __syncthreads();
    for (int r = threadIdx.x; r < N; r += blockDim.x) A[(size_t)(N - 1) * N + r] = czero<C>();
}

// ---------------------------------------------------------------------------
// Tridiagonal eigenpairs: bisection + inverse iteration, one WG per k
// ---------------------------------------------------------------------------
// 1/x from the hardware reciprocal refined by one Newton step (relative
// error a few ulp); x is never 0 here (pivots are clamped to -pivmin).
__device__ __forceinline__ double rcp_nr(double x)
{
    const double r = __builtin_amdgcn_rcp(x);
    return fma(fma(-x, r, 1.0), r, r);
}

// Number of eigenvalues of the tridiagonal below x (Sturm count of the
// LDL^T pivots, dstebz's dlaebz with pivmin clamping).
__device__ int sturm_count(const double *d, const double *e2, int n, double x, double pivmin)
{
    int cnt = 0;
    double q = d[0] - x;
    if (fabs(q) <= pivmin) q = -pivmin;
    cnt += q < 0.0;
    for (int i = 1; i < n; ++i) {
        q = d[i] - x - e2[i - 1] * rcp_nr(q);
        if (fabs(q) <= pivmin) q = -pivmin;
        cnt += q < 0.0;
    }
    return cnt;
}

// Deterministic uniform(-1,1) start vector entry (inverse iteration).
__device__ __forceinline__ double start_entry(int i, int j)
{
    uint32_t x = 0x9E3779B9u * (uint32_t)(i + 1) ^ (0x85EBCA6Bu * (uint32_t)(j + 7));
    x ^= x >> 16; x *= 0x7FEB352Du; x ^= x >> 15; x *= 0x846CA68Bu; x ^= x >> 16;
    return (double)x * (2.0 / 4294967296.0) - 1.0;
}

// Gershgorin interval, 1-norm (dstein's onenrm) and pivmin of the
// tridiagonal held in LDS (d, e2 = e^2, e), reduced over the block.
struct TriBounds { double gl, gu, onenrm, pivmin; };
__device__ TriBounds tri_bounds(const double *d, const double *e, const double *e2, int N, double *red4)
{
    const int tid = threadIdx.x, nt = blockDim.x;
    double lo = DBL_MAX, hi = -DBL_MAX, nrm = 0.0, emax = 0.0;
    for (int i = tid; i < N; i += nt) {
        const double r = fabs(e[i]) + (i > 0 ? fabs(e[i - 1]) : 0.0);
        lo = fmin(lo, d[i] - r); hi = fmax(hi, d[i] + r);
        nrm = fmax(nrm, fabs(d[i]) + r);
        emax = fmax(emax, e2[i]);
    }
    for (int off = 32; off > 0; off >>= 1) {
        lo = fmin(lo, __shfl_xor(lo, off, 64)); hi = fmax(hi, __shfl_xor(hi, off, 64));
        nrm = fmax(nrm, __shfl_xor(nrm, off, 64)); emax = fmax(emax, __shfl_xor(emax, off, 64));
    }
    const int lane = tid & 63, wv = tid >> 6, nw = (nt + 63) >> 6;
    __syncthreads();
    if (lane == 0) { red4[4 * wv] = lo; red4[4 * wv + 1] = hi; red4[4 * wv + 2] = nrm; red4[4 * wv + 3] = emax; }
    __syncthreads();
    lo = red4[0]; hi = red4[1]; nrm = red4[2]; emax = red4[3];
    for (int q = 1; q < nw; ++q) {
        lo = fmin(lo, red4[4 * q]); hi = fmax(hi, red4[4 * q + 1]);
        nrm = fmax(nrm, red4[4 * q + 2]); emax = fmax(emax, red4[4 * q + 3]);
    }
    __syncthreads();
    TriBounds b;
    b.onenrm = nrm;
    b.pivmin = DBL_MIN * fmax(1.0, emax);
    const double eps = DBL_EPSILON * 0.5;                 // dlamch('E')
    b.gl = lo - 2.0 * eps * nrm - b.pivmin;
    b.gu = hi + 2.0 * eps * nrm + b.pivmin;
    return b;
}

// Eigenvalues of the tridiagonal by trisection on Sturm counts, two lanes
// per eigenvalue (neighbouring lanes of a wave): each step counts the
// eigenvalues below the interval's two inner third points, one point per
// lane, both lanes keep the third that holds eigenvalue j (count(a) <= j <
// count(b) throughout), and stop when it is as narrow as dstebz's absolute
// tolerance 2 eps max(|a|,|b|) (or the coarser abstol_rel ||T||_1 when the
// caller refines the eigenpairs).  A step is one Sturm chain per lane, as a
// bisection step is per thread, and shrinks the interval three-fold: 24
// instead of 38 steps at 1e-11, with twice the chains in flight to cover
// the latency of the chain's reciprocal and FMAs.
__global__ __launch_bounds__(256) void k_stebz(SteinArgs s)
{
    __shared__ double d[1280], e[1280], e2[1280];
    __shared__ double red4[64];
    const int LD = s.n, kb = blockIdx.y, tid = threadIdx.x;
    const int N = s.nact ? s.nact[kb] : LD;           // active order (padding after)
    const double *dg = s.d + (size_t)kb * LD;
    const double *eg = s.e + (size_t)kb * LD;
    for (int i = tid; i < N; i += blockDim.x) {
        d[i] = dg[i];
        e[i] = (i < N - 1) ? eg[i] : 0.0;
        e2[i] = e[i] * e[i];
    }
    __syncthreads();
    const TriBounds tb = tri_bounds(d, e, e2, N, red4);
    const double eps = DBL_EPSILON * 0.5;
    if (blockIdx.x == 0 && tid == 0) s.info[kb] = 0;
    // one shared step: the counts at the 256 interior points of a uniform
    // grid over the Gershgorin interval, one per thread, bracket every
    // eigenvalue to 1/257 of it (about five trisection steps)
    __shared__ int gcnt[256];
    const int G = (int)blockDim.x;
    const double gh = (tb.gu - tb.gl) / (double)(G + 1);
    gcnt[tid] = sturm_count(d, e2, N, tb.gl + (double)(tid + 1) * gh, tb.pivmin);
    __syncthreads();
    const int gt = blockIdx.x * blockDim.x + tid;
    const int j = gt >> 1, side = gt & 1;
    if (j >= LD) return;
    if (j >= N) {                                     // decoupled padding
        if (side == 0) s.w[(size_t)kb * LD + j] = dg[j];
        return;
    }
    double a = tb.gl, b = tb.gu;
    {
        // the first grid point whose count exceeds j; the bracket is kept
        // only if the counts around it confirm it (they are monotone but
        // for rounding)
        int lo = 0, hi = G;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (gcnt[mid] >= j + 1) hi = mid; else lo = mid + 1;
        }
        if ((lo == 0 || gcnt[lo - 1] < j + 1) && (lo == G || gcnt[lo] >= j + 1)) {
            if (lo > 0) a = tb.gl + (double)lo * gh;
            if (lo < G) b = tb.gl + (double)(lo + 1) * gh;
        }
    }
    const double atol = s.abstol_rel * tb.onenrm;
    for (int it = 0; it < 200; ++it) {
        const double mid = 0.5 * (a + b);
        const double tol = fmax(2.0 * eps * fmax(fabs(a), fabs(b)) + tb.pivmin, atol);
        if (b - a <= tol || mid <= a || mid >= b) break;
        const double third = (b - a) * (1.0 / 3.0);
        const double x = side == 0 ? a + third : b - third;
        const int cnt = sturm_count(d, e2, N, x, tb.pivmin);
        const double xo = __shfl_xor(x, 1, 64);
        const int co = __shfl_xor(cnt, 1, 64);
        const double x1 = side == 0 ? x : xo, x2 = side == 0 ? xo : x;
        const int c1 = side == 0 ? cnt : co, c2 = side == 0 ? co : cnt;
        if (c1 >= j + 1) b = x1;
        else if (c2 >= j + 1) { a = x1; b = x2; }
        else a = x2;
    }
    if (side == 0) s.w[(size_t)kb * LD + j] = 0.5 * (a + b);
}

// Eigenvectors of the tridiagonal, one thread per eigenvalue (64 per
// workgroup).  Eigenvalues separated from their predecessor by more than
// 1e-10 * onenrm: twisted factorisation of T - lambda I (L D+ L^T from the
// top, U D- U^T from the bottom, twist r at min |gamma_r|, z_r = 1 and the
// two bidiagonal recurrences).  Near-exact degeneracies: inverse iteration
// from a distinct start vector with the shift perturbed away from the
// previous member (dstein).  Vectors of close eigenvalues are
// orthogonalised afterwards by k_stein_orth.
__global__ __launch_bounds__(64) void k_stein_vec(SteinArgs s)
{
    // d, e, e2, w of the active order in dynamic LDS (4 s.n doubles): small
    // enough that many 64-thread workgroups share a CU
    extern __shared__ double stein_smem[];
    double *d = stein_smem, *e = d + s.n, *e2 = e + s.n, *w = e2 + s.n;
    __shared__ double red4[16];
    const int LD = s.n;                               // stride; order of the padded batch
    const int kb = blockIdx.y;
    const int N = s.nact ? s.nact[kb] : LD;           // active order
    const int tid = threadIdx.x;
    const int nthr = gridDim.x * blockDim.x;          // threads per k
    const int gt = blockIdx.x * blockDim.x + tid;
    const double *dg = s.d + (size_t)kb * LD;
    const double *eg = s.e + (size_t)kb * LD;
    double *Vt = s.v + (size_t)kb * LD * LD;          // Vt[i*LD + j] = v_j(i)
    double *scr = s.scratch + (size_t)kb * 5 * LD * nthr;
    for (int i = tid; i < N; i += blockDim.x) {
        d[i] = dg[i];
        e[i] = (i < N - 1) ? eg[i] : 0.0;
        e2[i] = e[i] * e[i];
        w[i] = s.w[(size_t)kb * LD + i];
    }
    __syncthreads();
    const TriBounds tb = tri_bounds(d, e, e2, N, red4);
    const double onenrm = tb.onenrm, pivmin = tb.pivmin;
    const double eps = DBL_EPSILON * 0.5;
    double *wa = scr + 0 * (size_t)N * nthr, *wb = scr + 1 * (size_t)N * nthr;
    double *wc = scr + 2 * (size_t)N * nthr, *wd = scr + 3 * (size_t)N * nthr;
    double *wy = scr + 4 * (size_t)N * nthr;
#define AT(arr, i) arr[(size_t)(i) * nthr + gt]
    const double dtpcrt = sqrt(0.1 / N);
    int fails = 0;
    const int j = gt;
    const double dtol = 1e-10 * onenrm;
    if (s.cut && j < N && j >= s.cut[kb]) {               // outside the window: zero vector
        for (int i = 0; i < N; ++i) Vt[(size_t)i * LD + j] = 0.0;
    } else if (j < N) {
        if (j == 0 || w[j] - w[j - 1] > dtol) {
            // D+ (top-down) into scratch; U_i = e_i / D-_{i+1} (bottom-up)
            // into slot i+1 of the output column; z overwrites both.
            const double lam = w[j];
            double dp = d[0] - lam;
            if (fabs(dp) < pivmin) dp = -pivmin;
            AT(wa, 0) = dp;
            for (int i = 0; i < N - 1; ++i) {
                const double l = e[i] * rcp_nr(dp);
                dp = d[i + 1] - lam - l * e[i];
                if (fabs(dp) < pivmin) dp = -pivmin;
                AT(wa, i + 1) = dp;
            }
            double dm = d[N - 1] - lam;
            if (fabs(dm) < pivmin) dm = -pivmin;
            double gbest = fabs(AT(wa, N - 1) + dm - (d[N - 1] - lam));
            int r = N - 1;
            // unrolled so that the scratch loads of several steps, which do
            // not depend on the recurrence, are in flight together
#pragma unroll 8
            for (int i = N - 2; i >= 0; --i) {
                const double u = e[i] * rcp_nr(dm);
                Vt[(size_t)(i + 1) * LD + j] = u;
                dm = d[i] - lam - u * e[i];
                if (fabs(dm) < pivmin) dm = -pivmin;
                const double g = fabs(AT(wa, i) + dm - (d[i] - lam));
                if (g < gbest) { gbest = g; r = i; }
            }
            double z = 1.0, ss = 1.0;
            Vt[(size_t)r * LD + j] = 1.0;
#pragma unroll 8
            for (int i = r; i < N - 1; ++i) {
                z = -Vt[(size_t)(i + 1) * LD + j] * z;
                Vt[(size_t)(i + 1) * LD + j] = z;
                ss += z * z;
            }
            z = 1.0;
#pragma unroll 8
            for (int i = r - 1; i >= 0; --i) {
                z = -(e[i] * rcp_nr(AT(wa, i))) * z;
                Vt[(size_t)i * LD + j] = z;
                ss += z * z;
            }
            const double scl = 1.0 / sqrt(ss);
#pragma unroll 8
            for (int i = 0; i < N; ++i) Vt[(size_t)i * LD + j] *= scl;
        } else {
            // shift of member j of the degenerate run, perturbed away from
            // the previous member's shift (dstein's pertol)
            int j0 = j;
            while (j0 > 0 && w[j0] - w[j0 - 1] <= dtol) --j0;
            double xjm = w[j0], xj = w[j0];
            for (int jj = j0 + 1; jj <= j; ++jj) {
                xj = w[jj];
                const double pertol = 10.0 * fabs(eps * xj);
                if (xj - xjm < pertol) xj = xjm + pertol;
                xjm = xj;
            }
            // LU of T - xj I with partial pivoting (dlagtf)
            for (int i = 0; i < N; ++i) {
                AT(wa, i) = d[i] - xj;
                AT(wb, i) = e[i];
                AT(wc, i) = e[i];
                AT(wd, i) = 0.0;
            }
            double scale1 = fabs(AT(wa, 0)) + (N > 1 ? fabs(AT(wb, 0)) : 0.0);
            // pivot flags (0 / 1) in wy
            for (int k = 0; k < N - 1; ++k) {
                double scale2 = fabs(AT(wc, k)) + fabs(AT(wa, k + 1));
                if (k < N - 2) scale2 += fabs(AT(wb, k + 1));
                const double ak = AT(wa, k);
                const double piv1 = (ak == 0.0) ? 0.0 : fabs(ak) / scale1;
                const double ck = AT(wc, k);
                if (ck == 0.0) {
                    AT(wy, k) = 0.0;
                    scale1 = scale2;
                    if (k < N - 2) AT(wd, k) = 0.0;
                } else {
                    const double piv2 = fabs(ck) / scale2;
                    if (piv2 <= piv1) {
                        AT(wy, k) = 0.0;
                        scale1 = scale2;
                        const double mult = ck / ak;
                        AT(wc, k) = mult;
                        AT(wa, k + 1) = AT(wa, k + 1) - mult * AT(wb, k);
                        if (k < N - 2) AT(wd, k) = 0.0;
                    } else {
                        AT(wy, k) = 1.0;
                        const double mult = ak / ck;
                        AT(wa, k) = ck;
                        const double temp = AT(wa, k + 1);
                        AT(wa, k + 1) = AT(wb, k) - mult * temp;
                        if (k < N - 2) {
                            AT(wd, k) = AT(wb, k + 1);
                            AT(wb, k + 1) = -mult * AT(wd, k);
                        }
                        AT(wb, k) = temp;
                        AT(wc, k) = mult;
                    }
                }
            }
            // the iterate lives in the output column of Vt
            double tolp = fabs(AT(wa, 0));
            if (N > 1) tolp = fmax(tolp, fmax(fabs(AT(wa, 1)), fabs(AT(wb, 0))));
            for (int k = 2; k < N; ++k)
                tolp = fmax(tolp, fmax(fabs(AT(wa, k)), fmax(fabs(AT(wb, k - 1)), fabs(AT(wd, k - 2)))));
            tolp *= eps;
            if (tolp == 0.0) tolp = eps;
            // start vector in the output column of Vt
            for (int i = 0; i < N; ++i) Vt[(size_t)i * LD + j] = start_entry(i, j);
            int its = 0, nrmchk = 0;
            bool ok = false;
            while (its < 5) {
                ++its;
                double asum = 0.0;
                for (int i = 0; i < N; ++i) asum += fabs(Vt[(size_t)i * LD + j]);
                const double scl = N * onenrm * fmax(eps, fabs(AT(wa, N - 1))) / fmax(asum, DBL_MIN);
                // forward: apply P and L
                double yprev = Vt[j] * scl;
                Vt[j] = yprev;
                for (int k = 1; k < N; ++k) {
                    double yk = Vt[(size_t)k * LD + j] * scl;
                    if (AT(wy, k - 1) == 0.0) {
                        yk = yk - AT(wc, k - 1) * yprev;
                    } else {
                        const double t = yprev;
                        Vt[(size_t)(k - 1) * LD + j] = yk;
                        yk = t - AT(wc, k - 1) * yk;
                    }
                    Vt[(size_t)k * LD + j] = yk;
                    yprev = yk;
                }
                // back substitution with perturbed tiny pivots (dlagts, job -1)
                for (int k = N - 1; k >= 0; --k) {
                    double temp = Vt[(size_t)k * LD + j];
                    if (k <= N - 3)
                        temp = temp - AT(wb, k) * Vt[(size_t)(k + 1) * LD + j] - AT(wd, k) * Vt[(size_t)(k + 2) * LD + j];
                    else if (k == N - 2)
                        temp = temp - AT(wb, k) * Vt[(size_t)(k + 1) * LD + j];
                    double ak = AT(wa, k);
                    double pert = copysign(tolp, ak);
                    for (int g = 0; g < 64; ++g) {
                        const double absak = fabs(ak);
                        if (absak < 1.0) {
                            if (absak < DBL_MIN) {
                                if (absak == 0.0 || fabs(temp) * DBL_MIN > absak) { ak += pert; pert *= 2.0; continue; }
                                temp *= 1.0 / DBL_MIN; ak *= 1.0 / DBL_MIN;
                            } else if (fabs(temp) > absak * (1.0 / DBL_MIN)) {
                                ak += pert; pert *= 2.0; continue;
                            }
                        }
                        break;
                    }
                    Vt[(size_t)k * LD + j] = temp / ak;
                }
                double amax = 0.0;
                for (int i = 0; i < N; ++i) amax = fmax(amax, fabs(Vt[(size_t)i * LD + j]));
                if (amax < dtpcrt) continue;
                ++nrmchk;
                if (nrmchk < 3) continue;
                ok = true;
                break;
            }
            if (!ok) ++fails;
            double ss = 0.0, amax = 0.0;
            int imax = 0;
            for (int i = 0; i < N; ++i) {
                const double v = Vt[(size_t)i * LD + j];
                ss += v * v;
                if (fabs(v) > amax) { amax = fabs(v); imax = i; }
            }
            double scl = 1.0 / sqrt(ss);
            if (Vt[(size_t)imax * LD + j] < 0.0) scl = -scl;
            for (int i = 0; i < N; ++i) Vt[(size_t)i * LD + j] *= scl;
            xjm = xj;
        }
    }
#undef AT
    if (fails) atomicAdd(s.info + kb, fails);
    if (N < LD && blockIdx.x == 0) {       // padding: zero rows, unit columns
        for (int t = tid; t < (LD - N) * LD; t += blockDim.x) {
            const int i = N + t / LD, jj = t % LD;
            Vt[(size_t)i * LD + jj] = (i == jj) ? 1.0 : 0.0;
        }
        for (int t = tid; t < N * (LD - N); t += blockDim.x) {
            const int i = t / (LD - N), jj = N + t % (LD - N);
            Vt[(size_t)i * LD + jj] = 0.0;
        }
    }
}

// Orthogonalisation of the vectors of each group of close eigenvalues
// (consecutive gaps <= ortol_rel * onenrm, as dstein groups them), one
// workgroup per k: classical Gram-Schmidt applied twice per vector, in
// ascending order, against the group's earlier (already orthonormal)
// vectors, with the dot products of one vector formed together.
__global__ __launch_bounds__(256) void k_stein_orth(SteinArgs s)
{
    __shared__ double d[1280], e[1280], e2[1280], w[1280];
    __shared__ double dots[1280];
    __shared__ double red4[16 * 4];
    __shared__ double redn[4];
    const int LD = s.n, kb = blockIdx.x, tid = threadIdx.x;
    const int N = s.nact ? s.nact[kb] : LD;
    const double *dg = s.d + (size_t)kb * LD;
    const double *eg = s.e + (size_t)kb * LD;
    double *Vt = s.v + (size_t)kb * LD * LD;
    for (int i = tid; i < N; i += blockDim.x) {
        d[i] = dg[i];
        e[i] = (i < N - 1) ? eg[i] : 0.0;
        e2[i] = e[i] * e[i];
        w[i] = s.w[(size_t)kb * LD + i];
    }
    __syncthreads();
    const TriBounds tb = tri_bounds(d, e, e2, N, red4);
    const double ortol = s.ortol_rel * tb.onenrm;
    const int Nv = s.cut ? min(N, s.cut[kb]) : N;         // vectors formed
    int j0 = 0;
    while (j0 < Nv) {
        int j1 = j0 + 1;
        while (j1 < Nv && w[j1] - w[j1 - 1] <= ortol) ++j1;
        for (int j = j0 + 1; j < j1; ++j) {
            const int g = j - j0;
            for (int pass = 0; pass < 2; ++pass) {
                for (int l = tid; l < g; l += blockDim.x) {
                    double acc = 0.0;
                    for (int i = 0; i < N; ++i) acc += Vt[(size_t)i * LD + j0 + l] * Vt[(size_t)i * LD + j];
                    dots[l] = acc;
                }
                __syncthreads();
                for (int i = tid; i < N; i += blockDim.x) {
                    const double *row = Vt + (size_t)i * LD + j0;
                    double acc = 0.0;
                    for (int l = 0; l < g; ++l) acc += row[l] * dots[l];
                    Vt[(size_t)i * LD + j] -= acc;
                }
                __syncthreads();
            }
            double ss = 0.0;
            for (int i = tid; i < N; i += blockDim.x) ss += Vt[(size_t)i * LD + j] * Vt[(size_t)i * LD + j];
            for (int off = 32; off > 0; off >>= 1) ss += __shfl_xor(ss, off, 64);
            if ((tid & 63) == 0) redn[tid >> 6] = ss;
            __syncthreads();
            ss = redn[0] + redn[1] + redn[2] + redn[3];
            const double scl = 1.0 / sqrt(ss);
            for (int i = tid; i < N; i += blockDim.x) Vt[(size_t)i * LD + j] *= scl;
            __syncthreads();
        }
        j0 = j1;
    }
}

// Complex copy of the real row-major eigenvectors (Vt[i*N + j] = v_j(i)):
// Y[:, j] = v_j, a transpose through 32 x 32 LDS tiles so that both the
// reads of Vt rows and the writes of Y columns are coalesced.
template <class C>
__global__ __launch_bounds__(256) void k_vt_to_columns(int N, const double *Vt, C *Y, size_t stride,
                                                       double2 *Y64row0, const double2 *ph)
{
    typedef typename std::conditional<sizeof(C) == 16, double, float>::type R;
    __shared__ double tile[32][33];
    const int kb = blockIdx.z, i0 = blockIdx.y * 32, j0 = blockIdx.x * 32;
    const double *V = Vt + (size_t)kb * N * N;
    C *Yk = Y + (size_t)kb * stride;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int r = ty; r < 32; r += 8) {
        const int i = i0 + r, j = j0 + tx;
        tile[r][tx] = (i < N && j < N) ? V[(size_t)i * N + j] : 0.0;
    }
    __syncthreads();
    for (int c = ty; c < 32; c += 8) {
        const int j = j0 + c, i = i0 + tx;
        if (i < N && j < N) {
            const double v = tile[tx][c];
            const double2 f = ph ? ph[(size_t)kb * N + i] : make_double2(1.0, 0.0);
            Yk[(size_t)j * N + i] = cmake<C, R>((R)(v * f.x), (R)(v * f.y));
            if (i == 0 && Y64row0)
                Y64row0[(size_t)kb * stride + (size_t)j * N] = make_double2((double)(R)(v * f.x), (double)(R)(v * f.y));
        }
    }
}

// ---------------------------------------------------------------------------
// Mixed precision: the refinement step
// ---------------------------------------------------------------------------

// One refinement step for the Hermitian eigenproblem of C from the
// approximate eigenvectors Y of the FP32 reduction (Ogita & Aishima's step,
// RefSyEv, extended to complex Hermitian, in residual form, with clusters
// rotated exactly).  Columns j < mw are refined (mw: a provisional
// occupation window, or all columns); the others keep the FP32
// tridiagonal's eigenvalues theta and are not formed.
//  1. X = C Y (FP64 GEMM, by the caller).  Per column the Rayleigh quotient
//     lambda_j = Re(y_j^H x_j) / y_j^H y_j, the residual
//     r_j = x_j - lambda_j y_j (FP64, stored in FP32) and its norm
//     (k_oa_resid).
//  2. tau = kOaRhoR max_j ||r_j||, so that every pair further apart than tau
//     has |y_i^H r_j| / |lambda_j - lambda_i| <= 1 / kOaRhoR; clusters = runs
//     of the ascending lambda whose consecutive gaps are below tau
//     (k_oa_clusters).
//  3. Each cluster of 2 .. kOaMaxCl states is rotated by Rayleigh-Ritz on
//     T_cc = Y_c^H X_c and G_cc = Y_c^H Y_c (FP64 dots): L = I + (I - G_cc) / 2
//     (G_cc^{-1/2} to second order), M = L T_cc L = Q diag Q^H by cyclic
//     Jacobi, P = L Q; Y_c, X_c <- Y_c P, X_c P in FP64, their FP32 copies
//     and residuals rewritten, lambda_c <- eig(M) (k_oa_rr).
//  4. K = Y^H R (FP32 GEMM, by the caller): K_ij = y_i^H r_j =
//     (Y^H C Y)_ij - lambda_j (Y^H Y)_ij, the numerator of the first-order
//     step.  It is of the size of the FP32 reduction's errors, so FP32
//     products carry it to ~1e-7 of itself.
//  5. lambda'_j = lambda_j + sum_{i first order} |K_ij|^2 / (lambda_j - lambda_i),
//     the Rayleigh quotient of the refined vector to second order
//     (k_oa_second), sorted (k_oa_sort).
//  6. E_ij = K_ij / (lambda_j - lambda_i) between states of different
//     clusters further apart than tau; within a run of close states too long
//     to rotate E_ij = -(Y^H Y)_ij / 2 (FP64 dots: the symmetric
//     orthogonalisation); 0 within a rotated cluster, on the diagonal and
//     for pairs closer than tau across the window's edge (k_oa_emat, columns
//     in the final order).
//  7. D = Y E (FP32 GEMM, by the caller); y'_j = y_j + d_j, normalised in
//     FP64 (k_oa_final).
// Only step 1's product and the Rayleigh-Ritz dots run in FP64 matrix
// arithmetic; Ogita & Aishima's form takes Y^H C Y and Y^H Y as two further
// FP64 products of order N^3, whose cancellation the residual form moves
// into the O(N^2) residual.
constexpr int kOaMax = 1536;
constexpr int kOaMaxCl = 16;
// ||r_j|| bounds every |y_i^H r_j| and exceeds the largest by about the
// square root of the number of states r_j spreads over: 10 here puts the
// clusters about where 100 max |y_i^H r_j| would.  Larger values lengthen the
// runs beyond kOaMaxCl, which stay unrotated (at the 108-atom cell, order
// 972: largest eigenvalue error 4.2e-12 Ry at 10, 9.7e-11 at 30, 3.4e-10 at
// 100).
constexpr double kOaRhoR = 10.0;
constexpr int kOaSlots = 8;                  // workgroups per item in the cluster kernel

// Sum over the 64 lanes of a wave; every lane receives the same value
// (each exchange adds the same two operands in both lanes).
__device__ __forceinline__ double wave_allsum(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ int oa_active(const OaArgs &a, int kb) { return a.nact ? a.nact[kb] : a.n; }
// refined active columns
__device__ __forceinline__ int oa_refined(const OaArgs &a, int kb)
{
    return min(oa_active(a, kb), a.mw ? a.mw[kb] : a.n);
}
// states i and c (cinfo entries ci, cc) are coupled at first order
__device__ __forceinline__ bool oa_first_order(int ci, int cc, double li, double lc, double tau)
{
    return (ci & 0xffff) != (cc & 0xffff) && fabs(lc - li) > tau;
}

// One wave per column (four per workgroup).
__global__ __launch_bounds__(256) void k_oa_resid(OaArgs a)
{
    const int N = a.n, kb = blockIdx.y, lane = threadIdx.x & 63;
    const int j = blockIdx.x * 4 + (int)(threadIdx.x >> 6);
    if (j >= N) return;
    double *lam = a.lam + (size_t)kb * N;
    double *rn = a.rn + (size_t)kb * N;
    if (a.mw && j >= a.mw[kb]) {                      // not refined
        if (lane == 0) { lam[j] = a.theta[(size_t)kb * N + j]; rn[j] = 0.0; }
        return;
    }
    // Y in FP64 holds the FP32 values of Y32 until the clusters are rotated
    const double2 *x = a.X + (size_t)kb * a.stride + (size_t)j * N;
    const float2 *y = a.Y32 + (size_t)kb * a.stride + (size_t)j * N;
    float2 *r = a.R32 + (size_t)kb * a.stride + (size_t)j * N;
    double t = 0.0, g = 0.0;
    for (int i = lane; i < N; i += 64) {
        const double2 xv = x[i];
        const double yx = y[i].x, yy = y[i].y;
        t += yx * xv.x + yy * xv.y;
        g += yx * yx + yy * yy;
    }
    t = wave_allsum(t);
    g = wave_allsum(g);
    const double l = t / g;
    double rr = 0.0;
    for (int i = lane; i < N; i += 64) {
        const double2 xv = x[i];
        const double yx = y[i].x, yy = y[i].y;
        const double rx = xv.x - l * yx, ry = xv.y - l * yy;
        r[i] = make_float2((float)rx, (float)ry);
        rr += rx * rx + ry * ry;
    }
    rr = wave_allsum(rr);
    if (lane == 0) { lam[j] = l; rn[j] = sqrt(rr); }
}

// tau, the ascending order of the refined lambda and the clusters.  cinfo:
// bits 0-15 the cluster (the sorted position of its first state; an
// unrefined state is its own, numbered by its column), 16-23 the member's
// place, 24-31 the cluster's size if it is rotated, else 1.
__global__ __launch_bounds__(1024) void k_oa_clusters(OaArgs a)
{
    __shared__ double lsort[kOaMax];
    __shared__ int idx[kOaMax], cinf[kOaMax];
    __shared__ double red[16];
    __shared__ int ncl_s;
    const int N = a.n, kb = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int nw = (int)(blockDim.x >> 6);
    const int Ma = oa_refined(a, kb);
    const double *lam = a.lam + (size_t)kb * N;
    const double *rn = a.rn + (size_t)kb * N;
    double mx = 0.0;
    for (int i = tid; i < N; i += blockDim.x) {
        if (i < Ma) mx = fmax(mx, rn[i]);
        lsort[i] = lam[i];
        idx[i] = i;
        cinf[i] = i | (1 << 24);
    }
    for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
    if (lane == 0) red[wv] = mx;
    __syncthreads();
    if (tid == 0) {
        double m2 = 0.0;
        for (int q = 0; q < nw; ++q) m2 = fmax(m2, red[q]);
        const double tau = kOaRhoR * m2;
        a.tau[kb] = tau;
        // stable insertion sort of the refined part (nearly sorted already)
        for (int i = 1; i < Ma; ++i) {
            const double x = lsort[i];
            const int xi = idx[i];
            int k = i - 1;
            while (k >= 0 && lsort[k] > x) { lsort[k + 1] = lsort[k]; idx[k + 1] = idx[k]; --k; }
            lsort[k + 1] = x; idx[k + 1] = xi;
        }
        // runs with consecutive gaps below tau; those of 2 .. kOaMaxCl
        // states are rotated, longer runs are orthogonalised only
        int ncl = 0, s0 = 0;
        int *clist = a.clist + (size_t)kb * N;
        while (s0 < Ma) {
            int e = s0 + 1;
            while (e < Ma && lsort[e] - lsort[e - 1] < tau) ++e;
            const int m = e - s0;
            const bool rot = m >= 2 && m <= kOaMaxCl;
            if (rot) clist[ncl++] = s0;
            for (int q = 0; q < m; ++q) cinf[idx[s0 + q]] = s0 | (rot ? (q << 16) | (m << 24) : (1 << 24));
            s0 = e;
        }
        ncl_s = ncl;
    }
    __syncthreads();
    if (tid == 0) a.ncl[kb] = ncl_s;
    for (int i = tid; i < N; i += blockDim.x) {
        a.perm[(size_t)kb * N + i] = idx[i];
        a.cinfo[(size_t)kb * N + i] = cinf[i];
    }
}

// Rayleigh-Ritz in each rotated cluster (one wave per cluster, the m x m
// matrices in LDS), then the cluster's columns of Y, X, Y32 and R32 in the
// rotated basis.  Clusters own disjoint columns, so the clusters of an item
// run in parallel.
__global__ __launch_bounds__(64) void k_oa_rr(OaArgs a)
{
    __shared__ double2 Am[kOaMaxCl][kOaMaxCl + 1], Qm[kOaMaxCl][kOaMaxCl + 1], Lm[kOaMaxCl][kOaMaxCl + 1];
    __shared__ double2 Pm[kOaMaxCl][kOaMaxCl + 1];
    __shared__ double ev[kOaMaxCl];
    __shared__ int ord[kOaMaxCl], mem[kOaMaxCl];
    const int N = a.n, kb = blockIdx.y, lane = threadIdx.x;
    const int *perm = a.perm + (size_t)kb * N;
    const int ncl = a.ncl[kb];
    double2 *Xk = a.X + (size_t)kb * a.stride, *Yk = a.Y + (size_t)kb * a.stride;
    float2 *Y32k = a.Y32 + (size_t)kb * a.stride, *R32k = a.R32 + (size_t)kb * a.stride;
    for (int ci = blockIdx.x; ci < ncl; ci += gridDim.x) {
        const int b = a.clist[(size_t)kb * N + ci];
        const int m = (a.cinfo[(size_t)kb * N + perm[b]] >> 24) & 0xff;
        if (lane < m) mem[lane] = perm[b + lane];
        __syncthreads();
        // T_cc = Y_c^H X_c into Am, L = I + (I - Y_c^H Y_c) / 2 into Lm:
        // the upper triangle by dots over the rows, mirrored
        for (int q = 0; q < m; ++q)
            for (int k = q; k < m; ++k) {
                const double2 *yq = Yk + (size_t)mem[q] * N;
                const double2 *xk = Xk + (size_t)mem[k] * N, *yk = Yk + (size_t)mem[k] * N;
                double tr = 0.0, ti = 0.0, gr = 0.0, gi = 0.0;
                for (int i = lane; i < N; i += 64) {
                    const double2 u = yq[i], xv = xk[i], yv = yk[i];
                    tr += u.x * xv.x + u.y * xv.y;
                    ti += u.x * xv.y - u.y * xv.x;
                    gr += u.x * yv.x + u.y * yv.y;
                    gi += u.x * yv.y - u.y * yv.x;
                }
                tr = wave_allsum(tr); ti = wave_allsum(ti);
                gr = wave_allsum(gr); gi = wave_allsum(gi);
                if (lane == 0) {
                    if (q == k) {
                        Am[q][q] = make_double2(tr, 0.0);
                        Lm[q][q] = make_double2(1.5 - 0.5 * gr, 0.0);
                    } else {
                        Am[q][k] = make_double2(tr, ti);
                        Am[k][q] = make_double2(tr, -ti);
                        Lm[q][k] = make_double2(-0.5 * gr, -0.5 * gi);
                        Lm[k][q] = make_double2(-0.5 * gr, 0.5 * gi);
                    }
                }
            }
        __syncthreads();
        // M = L T L: Qm = T L, then Am = L Qm, Hermitian part
        for (int t = lane; t < m * m; t += 64) {
            const int q = t % m, k = t / m;
            double2 acc = make_double2(0.0, 0.0);
            for (int p = 0; p < m; ++p) acc = zadd(acc, zmul(Am[q][p], Lm[p][k]));
            Qm[q][k] = acc;
        }
        __syncthreads();
        double2 mv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int t = lane + 64 * u;
            mv[u] = make_double2(0.0, 0.0);
            if (t < m * m) {
                const int q = t % m, k = t / m;
                for (int p = 0; p < m; ++p) mv[u] = zadd(mv[u], zmul(Lm[q][p], Qm[p][k]));
            }
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int t = lane + 64 * u;
            if (t < m * m) Am[t % m][t / m] = mv[u];
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int t = lane + 64 * u;
            if (t < m * m) {
                const int q = t % m, k = t / m;
                const double2 x = Am[q][k], y = Am[k][q];
                mv[u] = q == k ? make_double2(x.x, 0.0) : make_double2(0.5 * (x.x + y.x), 0.5 * (x.y - y.y));
            }
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int t = lane + 64 * u;
            if (t < m * m) {
                const int q = t % m, k = t / m;
                Am[q][k] = mv[u];
                Qm[q][k] = make_double2(q == k ? 1.0 : 0.0, 0.0);
            }
        }
        __syncthreads();
        // cyclic Jacobi: A <- J^H A J, Q <- Q J with J_pp = J_qq = c,
        // J_pq = s ph, J_qp = -s conj(ph), ph = A_pq / |A_pq|
        for (int sweep = 0; sweep < 32; ++sweep) {
            double off = 0.0, dia = 0.0;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int t = lane + 64 * u;
                if (t < m * m) {
                    const int q = t % m, k = t / m;
                    const double2 x = Am[q][k];
                    if (q == k) dia += x.x * x.x;
                    else off += x.x * x.x + x.y * x.y;
                }
            }
            for (int o = 32; o > 0; o >>= 1) {
                off += __shfl_xor(off, o, 64);
                dia += __shfl_xor(dia, o, 64);
            }
            if (!(off > 1e-34 * dia)) break;
            for (int p = 0; p < m - 1; ++p)
                for (int q = p + 1; q < m; ++q) {
                    const double2 apq = Am[p][q];
                    const double aa = sqrt(apq.x * apq.x + apq.y * apq.y);
                    if (!(aa > 1e-300)) continue;
                    const double app = Am[p][p].x, aqq = Am[q][q].x;
                    const double2 ph = make_double2(apq.x / aa, apq.y / aa);
                    const double th = (aqq - app) / (2.0 * aa);
                    const double t = th == 0.0 ? 1.0 : copysign(1.0, th) / (fabs(th) + sqrt(1.0 + th * th));
                    const double c = 1.0 / sqrt(1.0 + t * t), sn = t * c;
                    const double2 sp = make_double2(sn * ph.x, sn * ph.y);       // s ph
                    const double2 spc = make_double2(sn * ph.x, -sn * ph.y);     // s conj(ph)
                    if (lane < m) {
                        const double2 x = Am[lane][p], y = Am[lane][q];
                        Am[lane][p] = zsub(zscale(c, x), zmul(spc, y));
                        Am[lane][q] = zadd(zmul(sp, x), zscale(c, y));
                    }
                    __syncthreads();
                    if (lane < m) {
                        const double2 x = Am[p][lane], y = Am[q][lane];
                        Am[p][lane] = zsub(zscale(c, x), zmul(sp, y));
                        Am[q][lane] = zadd(zmul(spc, x), zscale(c, y));
                    }
                    __syncthreads();
                    if (lane == 0) {
                        Am[p][q] = make_double2(0.0, 0.0);
                        Am[q][p] = make_double2(0.0, 0.0);
                        Am[p][p].y = 0.0;
                        Am[q][q].y = 0.0;
                    }
                    if (lane < m) {
                        const double2 x = Qm[lane][p], y = Qm[lane][q];
                        Qm[lane][p] = zsub(zscale(c, x), zmul(spc, y));
                        Qm[lane][q] = zadd(zmul(sp, x), zscale(c, y));
                    }
                    __syncthreads();
                }
        }
        if (lane == 0) {
            for (int k = 0; k < m; ++k) { ev[k] = Am[k][k].x; ord[k] = k; }
            for (int k = 1; k < m; ++k) {            // ascending, stable
                const double x = ev[k];
                const int xi = ord[k];
                int j = k - 1;
                while (j >= 0 && ev[j] > x) { ev[j + 1] = ev[j]; ord[j + 1] = ord[j]; --j; }
                ev[j + 1] = x; ord[j + 1] = xi;
            }
        }
        __syncthreads();
        // P = L Q (columns in ascending order of their eigenvalues)
        for (int t = lane; t < m * m; t += 64) {
            const int q = t % m, k = t / m;
            double2 acc = make_double2(0.0, 0.0);
            for (int p = 0; p < m; ++p) acc = zadd(acc, zmul(Lm[q][p], Qm[p][ord[k]]));
            Pm[q][k] = acc;
        }
        __syncthreads();
        for (int i = lane; i < N; i += 64) {
            double2 v[kOaMaxCl], u[kOaMaxCl];
#pragma unroll
            for (int q = 0; q < kOaMaxCl; ++q)
                if (q < m) {
                    v[q] = Yk[(size_t)mem[q] * N + i];
                    u[q] = Xk[(size_t)mem[q] * N + i];
                }
#pragma unroll
            for (int k = 0; k < kOaMaxCl; ++k) {
                if (k >= m) break;
                double2 yh = make_double2(0.0, 0.0), xh = make_double2(0.0, 0.0);
#pragma unroll
                for (int q = 0; q < kOaMaxCl; ++q)
                    if (q < m) {
                        yh = zadd(yh, zmul(v[q], Pm[q][k]));
                        xh = zadd(xh, zmul(u[q], Pm[q][k]));
                    }
                const size_t e = (size_t)mem[k] * N + i;
                Yk[e] = yh;
                Xk[e] = xh;
                Y32k[e] = make_float2((float)yh.x, (float)yh.y);
                R32k[e] = make_float2((float)(xh.x - ev[k] * yh.x), (float)(xh.y - ev[k] * yh.y));
            }
        }
        if (lane < m) a.lam[(size_t)kb * N + mem[lane]] = ev[lane];
        __syncthreads();
    }
}

// Second-order eigenvalues, one wave per refined column: lambda'_c into w
// (unsorted; unrefined columns keep lambda).
__global__ __launch_bounds__(1024) void k_oa_second(OaArgs a)
{
    const int N = a.n, kb = blockIdx.y, lane = threadIdx.x & 63;
    const int c = blockIdx.x * (int)(blockDim.x >> 6) + (int)(threadIdx.x >> 6);
    if (c >= N) return;
    const int Na = oa_active(a, kb), Ma = oa_refined(a, kb);
    const double *lam = a.lam + (size_t)kb * N;
    const double lc = lam[c];
    if (c >= Ma) {
        if (lane == 0) a.w[(size_t)kb * N + c] = lc;
        return;
    }
    const int *cinfo = a.cinfo + (size_t)kb * N;
    const int cc = cinfo[c];
    const double tau = a.tau[kb];
    const float2 *Kc = a.K32 + (size_t)kb * a.stride + (size_t)c * N;
    double corr = 0.0;
    for (int i = lane; i < Na; i += 64) {
        const double li = lam[i];
        if (i == c || !oa_first_order(cinfo[i], cc, li, lc, tau)) continue;
        const float2 k = Kc[i];
        corr += ((double)k.x * k.x + (double)k.y * k.y) / (lc - li);
    }
    corr = wave_allsum(corr);
    if (lane == 0) a.w[(size_t)kb * N + c] = lc + corr;
}

// lambda' sorted (from the order of the first sort, which it nearly keeps):
// the final order perm2 and the eigenvalues w (ascending).
__global__ __launch_bounds__(256) void k_oa_sort(OaArgs a)
{
    __shared__ double ls[kOaMax];
    __shared__ int ix[kOaMax];
    const int N = a.n, kb = blockIdx.x, tid = threadIdx.x;
    const int Na = oa_active(a, kb);
    double *w = a.w + (size_t)kb * N;
    const int *perm = a.perm + (size_t)kb * N;
    for (int s = tid; s < N; s += blockDim.x) {
        const int i = s < Na ? perm[s] : s;
        ix[s] = i;
        ls[s] = w[i];
    }
    __syncthreads();
    if (tid == 0)
        for (int i = 1; i < Na; ++i) {
            const double x = ls[i];
            const int xi = ix[i];
            int k = i - 1;
            while (k >= 0 && ls[k] > x) { ls[k + 1] = ls[k]; ix[k + 1] = ix[k]; --k; }
            ls[k + 1] = x; ix[k + 1] = xi;
        }
    __syncthreads();
    for (int s = tid; s < N; s += blockDim.x) {
        w[s] = ls[s];
        a.perm2[(size_t)kb * N + s] = ix[s];
    }
}

// E (FP32) with columns in the final order: E[:, j] is the step of the
// refined column perm2[j].
__global__ __launch_bounds__(256) void k_oa_emat(OaArgs a, const int *cut)
{
    const int N = a.n, kb = blockIdx.y;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)N * N) return;
    const int i = (int)(t % N), j = (int)(t / N);
    if (cut && j >= cut[kb]) return;                  // column outside the occupation window
    const int c = a.perm2[(size_t)kb * N + j];
    const int Na = oa_active(a, kb), Ma = oa_refined(a, kb);
    float2 e = make_float2(0.f, 0.f);
    if (c < Ma && i < Na && i != c) {
        const int *cinfo = a.cinfo + (size_t)kb * N;
        const double *lam = a.lam + (size_t)kb * N;
        const int ci = cinfo[i], cc = cinfo[c];
        const double li = lam[i], lc = lam[c];
        if (oa_first_order(ci, cc, li, lc, a.tau[kb])) {
            const float2 k = a.K32[(size_t)kb * a.stride + (size_t)c * N + i];
            const double den = lc - li;
            e = make_float2((float)(k.x / den), (float)(k.y / den));
        } else if ((ci & 0xffff) == (cc & 0xffff) && ((cc >> 24) & 0xff) < 2) {
            // a run too long to rotate: -(Y^H Y)_ic / 2
            const double2 *yi = a.Y + (size_t)kb * a.stride + (size_t)i * N;
            const double2 *yc = a.Y + (size_t)kb * a.stride + (size_t)c * N;
            double gr = 0.0, gi = 0.0;
            for (int r = 0; r < N; ++r) {
                const double2 u = yi[r], v = yc[r];
                gr += u.x * v.x + u.y * v.y;
                gi += u.x * v.y - u.y * v.x;
            }
            e = make_float2((float)(-0.5 * gr), (float)(-0.5 * gi));
        }
    }
    a.E32[(size_t)kb * a.stride + t] = e;
}

// y'_j = y_c + d_j (c = perm2[j]), normalised; one wave per column.
__global__ __launch_bounds__(256) void k_oa_final(OaArgs a, const int *cut)
{
    const int N = a.n, kb = blockIdx.y, lane = threadIdx.x & 63;
    const int j = blockIdx.x * 4 + (int)(threadIdx.x >> 6);
    if (j >= N || (cut && j >= cut[kb])) return;
    const int c = a.perm2[(size_t)kb * N + j];
    const double2 *y = a.Y + (size_t)kb * a.stride + (size_t)c * N;
    const float2 *d = a.D32 + (size_t)kb * a.stride + (size_t)j * N;
    double2 *o = a.Yr + (size_t)kb * a.stride + (size_t)j * N;
    double ss = 0.0;
    for (int i = lane; i < N; i += 64) {
        const double vx = y[i].x + (double)d[i].x, vy = y[i].y + (double)d[i].y;
        ss += vx * vx + vy * vy;
    }
    ss = wave_allsum(ss);
    const double sc = 1.0 / sqrt(ss);
    for (int i = lane; i < N; i += 64)
        o[i] = make_double2(sc * (y[i].x + (double)d[i].x), sc * (y[i].y + (double)d[i].y));
}

// ---------------------------------------------------------------------------
// First-order perturbation (perturb.f:87-221)
// ---------------------------------------------------------------------------
// Column dots: px[j] = z_j^H X[:, j], py[j] = z_j^H Y[:, j].
__global__ __launch_bounds__(256) void k_coldot(int N, const double2 *Z, const double2 *X,
                                                const double2 *Y, size_t stride, double2 *px,
                                                double2 *py, const int *cut)
{
    const int kb = blockIdx.y;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int j = blockIdx.x * 4 + wv;
    if (j >= N || (cut && j >= cut[kb])) return;
    const double2 *z = Z + (size_t)kb * stride + (size_t)j * N;
    const double2 *x = X + (size_t)kb * stride + (size_t)j * N;
    const double2 *y = Y + (size_t)kb * stride + (size_t)j * N;
    double ar = 0, ai = 0, br = 0, bi = 0;
    for (int i = lane; i < N; i += 64) {
        const double2 zi = z[i];
        const double2 p = zcmul(zi, x[i]), q = zcmul(zi, y[i]);
        ar += p.x; ai += p.y; br += q.x; bi += q.y;
    }
    for (int off = 32; off > 0; off >>= 1) {
        ar += __shfl_down(ar, off, 64); ai += __shfl_down(ai, off, 64);
        br += __shfl_down(br, off, 64); bi += __shfl_down(bi, off, 64);
    }
    if (lane == 0) {
        px[(size_t)kb * N + j] = make_double2(ar, ai);
        py[(size_t)kb * N + j] = make_double2(br, bi);
    }
}

// Eigenvalues of a small Hermitian matrix (n <= 9) by cyclic Jacobi on its
// real symmetric 2n x 2n embedding; every eigenvalue appears twice.
__device__ void small_heev(int n, const double2 *hd, double *out)
{
    double b[18][18];
    const int m = 2 * n;
    for (int c = 0; c < n; ++c)
        for (int r = 0; r < n; ++r) {
            const double2 v = hd[c * 9 + r];
            b[r][c] = v.x; b[r + n][c + n] = v.x;
            b[r + n][c] = v.y; b[r][c + n] = -v.y;
        }
    for (int sweep = 0; sweep < 60; ++sweep) {
        double off = 0.0;
        for (int r = 0; r < m; ++r) for (int c = 0; c < m; ++c) if (r != c) off += b[r][c] * b[r][c];
        if (off < 1e-300) break;
        for (int p = 0; p < m - 1; ++p)
            for (int q = p + 1; q < m; ++q) {
                const double apq = b[p][q];
                if (apq == 0.0) continue;
                const double th = (b[q][q] - b[p][p]) / (2.0 * apq);
                const double t = copysign(1.0, th) / (fabs(th) + sqrt(th * th + 1.0));
                const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < m; ++k) {
                    const double x = b[k][p], y = b[k][q];
                    b[k][p] = c * x - s * y; b[k][q] = s * x + c * y;
                }
                for (int k = 0; k < m; ++k) {
                    const double x = b[p][k], y = b[q][k];
                    b[p][k] = c * x - s * y; b[q][k] = s * x + c * y;
                }
            }
    }
    double ev[18];
    for (int i = 0; i < m; ++i) ev[i] = b[i][i];
    for (int i = 1; i < m; ++i) {
        const double x = ev[i];
        int k = i - 1;
        while (k >= 0 && ev[k] > x) { ev[k + 1] = ev[k]; --k; }
        ev[k + 1] = x;
    }
    for (int i = 0; i < n; ++i) out[i] = 0.5 * (ev[2 * i] + ev[2 * i + 1]);
}

// One thread per k-point: degenerate groups (|e_j - e_start| <= 1e-8, at
// most 9) get the eigenvalues of <a|dH - e_start dS|b>; isolated states
// take the diagonal element.  ncut counts states up to eigcut.
__global__ void k_perturb(PerturbArgs p)
{
    const int kb = blockIdx.x * blockDim.x + threadIdx.x;
    if (kb >= p.nb) return;
    const int N = p.nact ? p.nact[kb] : p.n;          // active order (padding after)
    const int LD = p.n;
    const double *w = p.w + (size_t)kb * LD;
    const double2 *px = p.px + (size_t)kb * LD, *py = p.py + (size_t)kb * LD;
    double *desk = p.desk + (size_t)kb * LD;
    const double2 *Z = p.z + (size_t)kb * p.stride;
    const double2 *X = p.x + (size_t)kb * p.stride;
    const double2 *Y = p.y + (size_t)kb * p.stride;
    const int wcut = p.cut ? p.cut[kb] : N;           // derivatives formed below wcut
    int start = 0, err = 0, ncut = N;
    while (start < N) {
        if (w[start] > p.eigcut) { ncut = start; break; }
        int num = N - start;
        for (int j = start + 1; j < N; ++j)
            if (fabs(w[j] - w[start]) > 1e-8) { num = j - start; break; }
        if (num > 9) { err = 1; ncut = start; break; }
        const double es = w[start];
        if (start >= wcut) {
            // above the occupation window: never summed; a group cannot
            // straddle wcut (it lies at a gap wider than the grouping tolerance)
            for (int j = start; j < start + num; ++j) desk[j] = 0.0;
        } else if (N == 1 && p.one_orbital_zero) {
            // a one-orbital cell: diag.f:320-349 forms no eigenvector for the
            // 1 x 1 pencil, so the reference's perturb works on a zero vector
            desk[start] = 0.0;
        } else if (num == 1) {
            desk[start] = 2.0 * (px[start].x - es * py[start].x);
        } else {
            double2 hd[81];
            for (int b = 0; b < num; ++b)
                for (int a = 0; a < num; ++a) {
                    const int ja = start + a, jb = start + b;
                    double2 v;
                    if (a == b) {
                        v = make_double2(2.0 * (px[ja].x - es * py[ja].x), 0.0);
                    } else {
                        // z_a^H M z_b = z_a^H (L z_b) + conj(z_b^H (L z_a))
                        double sr = 0, si = 0;
                        for (int i = 0; i < N; ++i) {
                            const double2 za = Z[(size_t)ja * LD + i], zb = Z[(size_t)jb * LD + i];
                            const double2 qb = zsub(X[(size_t)jb * LD + i], zscale(es, Y[(size_t)jb * LD + i]));
                            const double2 qa = zsub(X[(size_t)ja * LD + i], zscale(es, Y[(size_t)ja * LD + i]));
                            const double2 t = zcmul(za, qb), u = zcmul(zb, qa);
                            sr += t.x + u.x; si += t.y - u.y;
                        }
                        v = make_double2(sr, si);
                    }
                    hd[b * 9 + a] = v;
                }
            if (num == 2) {
                const double h1 = hd[0].x, h3 = hd[10].x;
                const double2 h2 = hd[9];
                const double sq = sqrt((h1 - h3) * (h1 - h3) + 4.0 * (h2.x * h2.x + h2.y * h2.y));
                desk[start] = 0.5 * ((h1 + h3) - sq);
                desk[start + 1] = 0.5 * ((h1 + h3) + sq);
            } else {
                small_heev(num, hd, desk + start);
            }
        }
        start += num;
    }
    for (int j = ncut; j < LD; ++j) desk[j] = 0.0;
    p.ncut[p.k0 + kb] = ncut;
    p.err[p.k0 + kb] = err ? 2 : 0;
}

// ---------------------------------------------------------------------------
// Fermi level and sums over all k-points, one workgroup
// ---------------------------------------------------------------------------
// Reductions run over a fixed thread -> k mapping and a fixed tree, so the
// result does not depend on how the k-points were batched.
__device__ void block_min_max(double &mn, double &mx, double *sm, double *sx)
{
    const int tid = threadIdx.x;
    sm[tid] = mn; sx[tid] = mx;
    __syncthreads();
    for (int st = blockDim.x / 2; st > 0; st >>= 1) {
        if (tid < st) { sm[tid] = fmin(sm[tid], sm[tid + st]); sx[tid] = fmax(sx[tid], sx[tid + st]); }
        __syncthreads();
    }
    mn = sm[0]; mx = sx[0];
    __syncthreads();
}

__device__ double block_sum_tree(double v, double *sbuf)
{
    const int tid = threadIdx.x;
    sbuf[tid] = v;
    __syncthreads();
    for (int st = blockDim.x / 2; st > 0; st >>= 1) {
        if (tid < st) sbuf[tid] = sbuf[tid] + sbuf[tid + st];
        __syncthreads();
    }
    const double r = sbuf[0];
    __syncthreads();
    return r;
}

// E_F of one structure (fermi.f:197-300): bracket, then Newton steps
// safeguarded by bisection on the electron count; every thread of the
// workgroup returns the same value.
__device__ double fermi_level(const FermiArgs &f, double *s1, double *s2, int &iters_out)
{
    __shared__ double mu_s;
    const int tid = threadIdx.x;
    const int nk = f.nk, nv = f.nv, ld = f.ld;
    const double el = f.el, tkb = f.tkb, zfill = f.zfill;
    const double nearz = 1e-5, tol = 1e-8, twenty = 20.0;
    const double beta = (tkb < nearz) ? 1.0 / nearz : 1.0 / tkb;
    int nfill = (int)(zfill / el + nearz);
    const bool linst = fabs(zfill / el - (double)nfill) < tol;
    if (nfill < 1) nfill = 1;
    // bracket (fermi.f:197-227).  Levels nfill and nfill + 1 lie above the
    // nv computed ones when the bands are full, as with one electron per
    // state (jspins = 2, el = 1: nfill = zfill); the reference then reads
    // rows of its eigenvalue array that no k-point wrote, zero in its
    // static storage, and so does this search.
    double stlow = -DBL_MAX, sthigh = DBL_MAX, mulow = DBL_MAX, muhigh = -DBL_MAX;
    for (int k = tid; k < nk; k += blockDim.x) {
        const double *ev = f.esk + (size_t)k * ld;
        stlow = fmax(stlow, nfill - 1 < nv ? ev[nfill - 1] : 0.0);
        sthigh = fmin(sthigh, nfill < nv ? ev[nfill] : 0.0);
        mulow = fmin(mulow, ev[0]);
        muhigh = fmax(muhigh, ev[nv - 1]);
    }
    {
        double a = -stlow, b = muhigh;
        block_min_max(a, b, s1, s2);       // max(stlow) = -min(-stlow)
        stlow = -a; muhigh = b;
        double c = mulow, d = -sthigh;
        block_min_max(c, d, s1, s2);
        mulow = c; sthigh = -d;
    }
    double mu;
    bool done = false;
    if (linst && (sthigh - stlow > twenty * tkb)) { mu = 0.5 * (stlow + sthigh); done = true; }
    else mu = 0.5 * (stlow + sthigh);
    int iters = 0;
    while (!done && iters < 1000) {
        ++iters;
        const double mum20 = mu - tkb * twenty, mup20 = mu + tkb * twenty;
        const double btmu = -beta * mu;
        double ns = 0.0, dns = 0.0;
        for (int k = tid; k < nk; k += blockDim.x) {
            const double *ev = f.esk + (size_t)k * ld;
            // the levels of a k are ascending: those below mu - 20 kT (full,
            // 1 each) are counted by bisection, the sum of their ones being
            // exact, and the walk starts at the first one above
            int lo = 0, hi = nv;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (ev[mid] < mum20) lo = mid + 1; else hi = mid;
            }
            double a = (double)lo, b = 0.0;
            for (int i = lo; i < nv; ++i) {
                const double x = ev[i];
                if (x < mum20) a = a + 1.0;
                else if (x < mup20) {
                    const double ex = exp(beta * x + btmu);
                    const double fx = 1.0 / (1.0 + ex);
                    a = a + fx;
                    b = b + ex * fx * fx;
                } else break;
            }
            ns = ns + f.w[k] * a;
            dns = dns + f.w[k] * b;
        }
        double nsum = block_sum_tree(ns, s1);
        double dnsum = block_sum_tree(dns, s2);
        nsum = el * nsum;
        dnsum = beta * el * dnsum;
        if (fabs(nsum - zfill) < tol) break;
        if (nsum < zfill) mulow = mu; else muhigh = mu;
        double mu1;
        if (dnsum < tol) mu1 = 0.5 * (mulow + muhigh);
        else {
            mu1 = mu + (zfill - nsum) / dnsum;
            if (mu1 > muhigh) mu1 = 0.5 * (mu + muhigh);
            else if (mu1 < mulow) mu1 = 0.5 * (mu + mulow);
        }
        if (muhigh - mulow > tol) mu = mu1;
        else break;
    }
    if (tid == 0) mu_s = mu;
    __syncthreads();
    iters_out = iters;
    return mu_s;
}

// ef_known: k_fermi_window_group has already found E_F (out[0], out[5])
// from the same eigenvalues in this run.
__device__ void fermi_sums_body(const FermiArgs &f, int ef_known)
{
    __shared__ double s1[1024], s2[1024];
    const int tid = threadIdx.x;
    const int nk = f.nk, nv = f.nv, ld = f.ld;
    const double el = f.el, tkb = f.tkb;
    const double nearz = 1e-5, twenty = 20.0;
    int iters = ef_known ? (int)f.out[5] : 0;
    const double efermi = ef_known ? f.out[0] : fermi_level(f, s1, s2, iters);
    // sums (bande.f:1170-1321)
    const double tkbi = (tkb < nearz) ? 1.0 / nearz : 1.0 / tkb;
    const double etol = twenty * tkb, efp = efermi + etol, efm = efermi - etol;
    const bool lpress = f.mode == 2 || f.mode == 4;
    const bool lentro = f.mode < 3;
    double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    // esum desum1 desum2 desum3 dfbsum dferm entropy dent1 dent2
    double occ = 0.0;               // spin up minus spin down (bande.f:1342-1402)
    for (int k = tid; k < nk; k += blockDim.x) {
        const double *ev = f.esk + (size_t)k * ld;
        const double *dv = f.desk + (size_t)k * ld;
        const double wt = el * f.w[k];
        if (f.khalf > 0) {
            double c = 0.0;
            for (int i = 0; i < nv; ++i) {
                const double x = ev[i];
                if (x < efm) c = c + f.w[k];
                else if (x < efp) c = c + f.w[k] * (1.0 / (1.0 + exp(tkbi * (x - efermi))));
            }
            occ = k < f.khalf ? occ + c : occ - c;
        }
        const int ni = lpress ? f.ncut[k] : nv;
        for (int i = 0; i < ni; ++i) {
            const double x = ev[i];
            if (!(x < efp)) continue;
            const double ewt = wt * x;
            if (x < efm) {
                acc[0] = acc[0] + ewt;
                if (lpress) acc[1] = acc[1] + wt * dv[i];
            } else {
                const double fbe = 1.0 / (1.0 + exp(tkbi * (x - efermi)));
                acc[0] = acc[0] + ewt * fbe;
                if (lentro) acc[6] = acc[6] - wt * (fbe * log(fbe) + (1.0 - fbe) * log(1.0 - fbe));
                if (lpress) {
                    const double fbm = 1.0 / (1.0 + exp(tkbi * (-(x - efermi))));
                    const double dfbe = -tkbi * fbe * fbm;
                    acc[4] = acc[4] + dfbe * wt;
                    acc[3] = acc[3] + dfbe * ewt;
                    const double dewt = wt * dv[i];
                    acc[1] = acc[1] + dewt * fbe;
                    acc[5] = acc[5] + dfbe * dewt;
                    acc[2] = acc[2] + dfbe * ewt * dv[i];
                    if (lentro) {
                        const double con1 = dfbe * wt * (x - efermi);
                        acc[7] = acc[7] + con1 * dv[i];
                        acc[8] = acc[8] + con1;
                    }
                }
            }
        }
    }
    double tot[9];
    for (int q = 0; q < 9; ++q) tot[q] = block_sum_tree(acc[q], s1);
    const double smag = f.khalf > 0 ? block_sum_tree(occ, s1) : 0.0;
    if (tid == 0) {
        const double esum = tot[0];
        double e1, e2 = 0.0;
        if (lentro) { const double fr = esum - tkb * tot[6]; e1 = 0.5 * (esum + fr); e2 = esum; }
        else e1 = esum;
        if (lpress) {
            const double dfermi = (fabs(tot[4]) > 1e-10) ? tot[5] / tot[4] : 0.0;
            const double press = -(tot[1] + (tot[2] - dfermi * tot[3]));
            if (lentro) { const double pent = dfermi * tot[8] - tot[7]; e2 = 0.5 * (press + (press - pent)); }
            else e2 = press;
        }
        f.out[0] = efermi; f.out[1] = e1; f.out[2] = e2;
        f.out[3] = esum; f.out[4] = tot[6];
        f.out[5] = (double)iters;
        f.out[6] = smag;
    }
}

__global__ __launch_bounds__(1024) void k_fermi_sums(FermiArgs f) { fermi_sums_body(f, 0); }
// one workgroup per structure of a plan
__global__ __launch_bounds__(1024) void k_fermi_sums_group(const FermiArgs *f, int ef_known)
{
    fermi_sums_body(f[blockIdx.x], ef_known);
}

// Occupation window of every k of one structure: the sums above skip every
// state at or above E_F + 20 kT, so eigenvectors and derivatives are needed
// only below the first such state.  The cut is moved up to the next gap
// wider than 3e-6 max|e| (and 1e-8): no group of degenerate states
// (perturb's 1e-8) or of close eigenvalues (the eigenvector
// orthogonalisation's 1e-6 ||T||_1, ||T||_1 <= 3 max|e|) straddles it, so
// every quantity below the cut is formed exactly as without the window.
__global__ __launch_bounds__(1024) void k_fermi_window_group(const FermiArgs *fa)
{
    __shared__ double s1[1024], s2[1024];
    const FermiArgs &f = fa[blockIdx.x];
    int iters = 0;
    const double ef = fermi_level(f, s1, s2, iters);
    if (threadIdx.x == 0) { f.out[0] = ef; f.out[5] = (double)iters; }   // reused by the sums
    const double efp = ef + 20.0 * f.tkb;
    for (int k = threadIdx.x; k < f.nk; k += blockDim.x) {
        const double *ev = f.esk + (size_t)k * f.ld;
        const double gap = fmax(1e-8, 3e-6 * fmax(fabs(ev[0]), fabs(ev[f.nv - 1])));
        int c = 0;
        while (c < f.nv && ev[c] < efp) ++c;
        while (c > 0 && c < f.nv && !(ev[c] - ev[c - 1] > gap)) ++c;
        f.cut[k] = c;
    }
}

// Spin expectation of each state on each atom (diag.f:458-523): one wave
// per state, lanes over atoms.  The reference sums over kbas(kkind(matom))
// orbitals of every atom (diag.f:488), the kind of its last atom slot; this
// sums over each atom's own orbitals, the same where the structure fills
// the reference's atom dimension (for fewer atoms that slot's kind is 0 and
// the bound reads the word before kbas, the COMMON's number of kinds).
__global__ __launch_bounds__(256) void k_spin_expect(SpinArgs a)
{
    const int kb = blockIdx.y, lane = threadIdx.x & 63, N = a.n;
    const int j = blockIdx.x * 4 + (int)(threadIdx.x >> 6);
    if (j >= N) return;
    const StructDev &S = a.sd[a.item_s[kb]];
    const int Na = a.nact ? a.nact[kb] : N;
    const bool on = S.spin && j < Na && (!a.cut || j < a.cut[kb]);
    const double2 *z = a.z + (size_t)kb * a.stride + (size_t)j * N;
    double *sp = a.spev + ((size_t)(a.k0 + kb) * N + j) * a.maxnat * 3;
    for (int at = lane; at < S.natoms; at += 64) {
        double sx = 0.0, sy = 0.0, sz = 0.0;
        if (on) {
            const int b0 = S.ksk[at], nb = S.nbas[at], nv = S.nv;
            for (int ib = b0; ib < b0 + nb; ++ib) {
                const double2 u = z[ib], d = z[ib + nv];
                const double cr = u.x * d.x + u.y * d.y, ci = u.x * d.y - u.y * d.x;   // conj(u) d
                sx = sx - cr - cr;
                sy = sy + (ci + ci);
                sz = sz + ((d.x * d.x + d.y * d.y) - (u.x * u.x + u.y * u.y));
            }
        }
        sp[3 * at] = sx;
        sp[3 * at + 1] = sy;
        sp[3 * at + 2] = sz;
    }
}

// The occupation-weighted sums of bande.f:1501-1521 (states below
// E_F + 20 kT, weight sfac w_k, Fermi-Dirac within 20 kT of E_F), k and
// states in order; one workgroup per structure, threads over (atom, axis).
__global__ __launch_bounds__(256) void k_spin_sum(const FermiArgs *fa, const StructDev *sdv, const double *spev,
                                                  int n, int maxnat, double *out)
{
    const FermiArgs &f = fa[blockIdx.x];
    const StructDev &S = sdv[blockIdx.x];
    const double ef = f.out[0], tkb = f.tkb, nearz = 1e-5;
    const double tkbi = tkb < nearz ? 1.0 / nearz : 1.0 / tkb;
    const double efp = ef + 20.0 * tkb, efm = ef - 20.0 * tkb;
    for (int t = threadIdx.x; t < 3 * S.natoms; t += blockDim.x) {
        const int at = t / 3, l = t % 3;
        double acc = 0.0;
        for (int k = 0; k < f.nk; ++k) {
            const double *ev = f.esk + (size_t)k * f.ld;
            const double wt = f.el * f.w[k];
            const double *sp = spev + (size_t)(f.k0 + k) * n * maxnat * 3 + 3 * at + l;
            for (int i = 0; i < f.nv; ++i) {
                const double e = ev[i];
                if (!(e < efp)) continue;
                const double fbe = e > efm ? wt * (1.0 / (1.0 + exp(tkbi * (e - ef)))) : wt;
                acc = acc + fbe * sp[(size_t)i * maxnat * 3];
            }
        }
        out[((size_t)blockIdx.x * maxnat + at) * 3 + l] = acc;
    }
}

// addspins.f:68-96 on dense matrices: with H, S the spatial matrices in the
// leading nv x nv block of an item of order N = 2 nv (the rest decoupled
// padding from k_build_hs), the spin-down block becomes H - diag(d), S; the
// spin-up diagonal H + diag(d); the mixing entries H(i, nv+i) = m_i and its
// conjugate; the other off-diagonal entries stay zero.  One thread per
// element of the spatial block; the derivative build uses the derivative
// terms.
__global__ __launch_bounds__(256) void k_spinor_expand(BuildArgs a, int nb)
{
    const int kb = blockIdx.y;
    if (kb >= nb) return;
    const StructDev &sd = a.sd[a.item_s[kb]];
    if (!sd.spin) return;
    const int nv = sd.nv, N = a.n;
    if (2 * nv > N) return;
    const double *sp = sd.spin + (a.deriv ? 3 * nv : 0);
    double2 *H = a.h + (size_t)kb * a.stride, *S = a.s + (size_t)kb * a.stride;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nv * nv; t += gridDim.x * blockDim.x) {
        const int r = t % nv, c = t / nv;
        const double2 h = H[(size_t)c * N + r];
        double2 hd = h;
        if (r == c) hd.x = h.x - sp[r];
        H[(size_t)(nv + c) * N + nv + r] = hd;
        S[(size_t)(nv + c) * N + nv + r] = S[(size_t)c * N + r];
        if (r == c) {
            H[(size_t)c * N + r] = make_double2(h.x + sp[r], h.y);
            H[(size_t)(nv + r) * N + r] = make_double2(sp[nv + r], sp[2 * nv + r]);
            H[(size_t)r * N + nv + r] = make_double2(sp[nv + r], -sp[2 * nv + r]);
        }
    }
}

// Plain streaming copy (16-byte elements), the calibration kernel for
// counter-based traffic measurements.
__global__ __launch_bounds__(256) void k_copy16(const double2 *src, double2 *dst, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

__global__ void k_copy_int(const int *src, int *dst, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

// ---------------------------------------------------------------------------
// Lowdin projections (diag.f:526-758): with S = V diag(lambda) V^H, the
// states y_j = S^{1/2} z_j = V diag(lambda^{1/2}) V^H z_j and their weights
// |y_j(i)|^2 on the orbitals (sum_i = 1 as z^H S z = 1).
// ---------------------------------------------------------------------------
// S of every item of the batch set to the identity
__global__ __launch_bounds__(256) void k_set_identity(int n, double2 *s, size_t stride)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)n * n) return;
    const int i = (int)(t % n), j = (int)(t / n);
    s[(size_t)blockIdx.y * stride + t] = make_double2(i == j ? 1.0 : 0.0, 0.0);
}

// T[i, j] *= sqrt(lambda_i)  (lambda: the item's eigenvalues of S)
__global__ __launch_bounds__(256) void k_scale_rows_sqrt(int n, double2 *t, size_t stride, const double *lam)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)n * n) return;
    const int kb = blockIdx.y, i = (int)(e % n);
    const double f = sqrt(lam[(size_t)kb * n + i]);
    double2 &v = t[(size_t)kb * stride + e];
    v.x *= f;
    v.y *= f;
}

// proj[k][j*n + i] = |Y[i, j]|^2 for the items k0 .. k0 + nb - 1
__global__ __launch_bounds__(256) void k_abs2(int n, const double2 *y, size_t stride, double *proj)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)n * n) return;
    const double2 v = y[(size_t)blockIdx.y * stride + e];
    proj[(size_t)blockIdx.y * n * n + e] = v.x * v.x + v.y * v.y;
}

// ---------------------------------------------------------------------------
// Two-stage tridiagonalisation
// ---------------------------------------------------------------------------
// Stage 1 reduces the Hermitian matrix (both triangles stored) to a lower
// band of width TB, TB columns at a time: k_band_qr factors the panel below
// the band, A[c0+TB:, c0:c0+TB] = Q R, and the trailing matrix takes the
// two-sided update Q^H A22 Q as GEMMs (Y = A22 V T, Z = Y - 1/2 V T^H V^H Y,
// A22 -= V Z^H + Z V^H).  Stage 2 (k_bulge) chases the band to tridiagonal
// form, one column per sweep, sweeps pipelined across the waves of a
// workgroup.  Eigenvectors go back through the stage-2 reflectors grouped
// TB at a time (k_q2_tfactors, k_apply_q2) and then the stage-1 panels.

// Sum over the 64 lanes of a wave (every lane gets the total).
template <class R>
__device__ __forceinline__ R wsum(R v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
template <class C>
__device__ __forceinline__ C cshfl(C v, int src)
{
    v.x = __shfl(v.x, src, 64);
    v.y = __shfl(v.y, src, 64);
    return v;
}
template <class C>
__device__ __forceinline__ C cshfl_xor(C v, int off)
{
    v.x = __shfl_xor(v.x, off, 64);
    v.y = __shfl_xor(v.y, off, 64);
    return v;
}
// 1 / z
template <class C>
__device__ __forceinline__ C crecip(C z)
{
    typedef decltype(z.x) R;
    const R d = z.x * z.x + z.y * z.y;
    return cmake<C, R>(z.x / d, -z.y / d);
}

// Householder reflector of a column x held one row per thread: returns tau
// and beta (H^H x = beta e1 with H = I - tau v v^H, v[0] = 1, beta real) and
// the scale 1 / (alpha - beta) of the rows below the first; tau = 0 when the
// column is already of that form.
template <class R, class C>
__device__ __forceinline__ void larfg_coef(C alpha, R xn2, C &tau, R &beta, C &scale)
{
    if (xn2 == R(0) && alpha.y == R(0)) {
        tau = czero<C>();
        beta = alpha.x;
        scale = czero<C>();
        return;
    }
    beta = -copysign(sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xn2), alpha.x);
    tau = cmake<C, R>((beta - alpha.x) / beta, -alpha.y / beta);
    scale = crecip(cmake<C, R>(alpha.x - beta, alpha.y));
}

// ---------------------------------------------------------------------------
// Resident Householder tridiagonalisation (zhetd2's unblocked steps), one
// workgroup of 512 threads per item, for orders up to 16 * NTM: the lower
// triangle of the item's matrix stays in registers for the whole reduction,
// as 16 x 16 tiles (diagonal tiles held in full, both triangles), so a
// column costs register arithmetic and LDS traffic only -- no pass over
// the trailing matrix in memory, no delayed update.
//
// Tiles are numbered from the last tile column leftwards, t = cnt(J + 1) +
// I - J with cnt(x) = (NT - x)(NT - x + 1) / 2 the tiles of columns >= x,
// and wave w holds tiles w, w + 8, ...: the live tiles (columns beyond the
// current one) are always a prefix of that order, so every wave keeps an
// equal share to the end.  Lane l of a wave holds column c = l & 15 and rows
// 4 g .. 4 g + 3 (g = l >> 4) of each of its tiles.
//
// Column j: (U) A -= v w^H + w v^H with the previous column's v, w (exact
// zeros outside the trailing rows, so dead parts of a tile stay untouched
// and add nothing), column j copied out; (R) the reflector, computed by
// every wave from the same LDS copy (identical results), v into LDS and into
// column j of A; (H) y = A v: per tile the row sums A_IJ v_J (reduced over a
// 16-lane DPP row) and, below the diagonal, the column sums A_IJ^H v_I
// (reduced over the four rows by lane-half swaps) into one LDS slot each,
// and v^H A v;
// (W) y_r from the slots of its tile row and tile column in a fixed order,
// w = tau y - |tau|^2 (v^H A v) / 2 v.  Deterministic: no atomics.
// The current 32-column panel's reflectors are kept in LDS: their dots with
// each new v (16 threads per dot) give the WY factor's next column, and at
// the panel's end T and V T^H are written.  Outputs as k_hetrd_panel's: the
// reflectors in A, d, e, T and V T^H of every panel.
// ---------------------------------------------------------------------------
constexpr int RZ_THREADS = 512, RZ_WAVES = RZ_THREADS / 64;

template <int CTRL>
__device__ __forceinline__ float dpp_mov(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double v)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)b, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// Sum over the 16 lanes of a DPP row (c = lane & 15) of 8 values per lane,
// transposed: each halving step keeps half of the values and adds the
// partner's copies of them (partners c ^ 8, c ^ 7, c ^ 3, c ^ 1: row_ror:8,
// row_half_mirror, quad_perm [3,2,1,0], [1,0,3,2]).  Lane c returns the
// total of value (c >> 1) & 7.
template <class R>
__device__ __forceinline__ R row16_sum8(R (&v)[8], int c)
{
    const bool b3 = (c & 8) != 0, b2 = (c & 4) != 0, b1 = (c & 2) != 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        R lo = v[q], hi = v[q + 4];
        v[q] = (b3 ? hi : lo) + dpp_mov<0x128>(b3 ? lo : hi);
    }
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        R lo = v[q], hi = v[q + 2];
        v[q] = (b2 ? hi : lo) + dpp_mov<0x141>(b2 ? lo : hi);
    }
    {
        R lo = v[0], hi = v[1];
        v[0] = (b1 ? hi : lo) + dpp_mov<0x1B>(b1 ? lo : hi);
    }
    return v[0] + dpp_mov<0xB1>(v[0]);
}

__host__ __device__ constexpr int rz_cnt(int nt, int x) { return (nt - x) * (nt - x + 1) / 2; }

// 1 / x: FP32 from the hardware reciprocal refined by one Newton step
// (within an ulp; the mixed method's FP32 reduction is refined in FP64
// afterwards), FP64 divided
__device__ __forceinline__ float rz_rcp(float x)
{
    const float r = __builtin_amdgcn_rcpf(x);
    return fmaf(fmaf(-x, r, 1.0f), r, r);
}
__device__ __forceinline__ double rz_rcp(double x) { return 1.0 / x; }
__device__ __forceinline__ float rz_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ double rz_sqrt(double x) { return sqrt(x); }
// larfg_coef with two reciprocals (1 / beta and 1 / |alpha - beta|^2)
template <class R, class C>
__device__ __forceinline__ void rz_larfg(C alpha, R xn2, C &tau, R &beta, C &scale)
{
    if (xn2 == R(0) && alpha.y == R(0)) {
        tau = czero<C>();
        beta = alpha.x;
        scale = czero<C>();
        return;
    }
    beta = -copysign(rz_sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xn2), alpha.x);
    const R rb = rz_rcp(beta);
    tau = cmake<C, R>((beta - alpha.x) * rb, -alpha.y * rb);
    const R dx = alpha.x - beta, dy = alpha.y, rd = rz_rcp(dx * dx + dy * dy);
    scale = cmake<C, R>(dx * rd, -dy * rd);
}

// v + v(lane ^ 32) and v + v(lane ^ 16) through gfx950's lane-half swaps
// (identical in both lanes of a pair: the addition commutes)
__device__ __forceinline__ float xor32_sum(float v)
{
    const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(v), __float_as_uint(v), false, false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}
__device__ __forceinline__ float xor16_sum(float v)
{
    const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(v), __float_as_uint(v), false, false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}
__device__ __forceinline__ double xor32_sum(double v)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)b, (unsigned)b, false, false);
    const auto hi = __builtin_amdgcn_permlane32_swap((unsigned)(b >> 32), (unsigned)(b >> 32), false, false);
    const double a = __longlong_as_double((long long)(((unsigned long long)hi[0] << 32) | lo[0]));
    const double c = __longlong_as_double((long long)(((unsigned long long)hi[1] << 32) | lo[1]));
    return a + c;
}
__device__ __forceinline__ double xor16_sum(double v)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const auto lo = __builtin_amdgcn_permlane16_swap((unsigned)b, (unsigned)b, false, false);
    const auto hi = __builtin_amdgcn_permlane16_swap((unsigned)(b >> 32), (unsigned)(b >> 32), false, false);
    const double a = __longlong_as_double((long long)(((unsigned long long)hi[0] << 32) | lo[0]));
    const double c = __longlong_as_double((long long)(((unsigned long long)hi[1] << 32) | lo[1]));
    return a + c;
}
// Sum over the 16 lanes of each DPP row (row_ror 8, 4, 2, 1): every lane of
// the row ends with the same value.
template <class R>
__device__ __forceinline__ R row16_sum(R v)
{
    v += dpp_mov<0x128>(v);
    v += dpp_mov<0x124>(v);
    v += dpp_mov<0x122>(v);
    v += dpp_mov<0x121>(v);
    return v;
}
// Sum over the 64 lanes of a wave, the same value in every lane.
template <class R>
__device__ __forceinline__ R wave_sum_fast(R v)
{
    return xor32_sum(xor16_sum(row16_sum(v)));
}

// LDS of the resident reduction: tile dimension NTM, NTM (NTM + 1) / 2 tiles
template <class R, int NTM>
struct ResLds {
    typedef typename cplx<R>::type C;
    static constexpr int LM = NTM * (NTM + 1) / 2;
    C colb[16 * NTM];              // column j after its last update
    C vb[16 * NTM], wb[16 * NTM];  // v and w of the current column
    C slotR[LM][16], slotC[LM][16];
    C Vp[NB][16 * NTM];            // the current panel's reflectors
    C Ts[NB][NB + 1];              // its WY factor, Ts[row][col]
    C Gc[NB];                      // V_l^H v of the current column
    R red[RZ_WAVES];
};

template <class R, int NTM, int MT>
__global__ __launch_bounds__(RZ_THREADS, 1) void k_hetrd_res(HetrdArgsT<R> h)
{
    typedef typename cplx<R>::type C;
    typedef R v4 __attribute__((ext_vector_type(4)));
    typedef ResLds<R, NTM> L_t;
    constexpr int LM = L_t::LM;
    static_assert(LM <= RZ_WAVES * MT, "every tile needs a register slot");
    static_assert(16 * NTM <= RZ_THREADS && 16 * NTM <= 40 * RZ_WAVES, "one row per thread");
    static_assert(sizeof(L_t) <= 160 * 1024, "LDS of one CU");
    __shared__ L_t S;
    const int N = h.n, NT = (N + 15) >> 4;
    const int kb = blockIdx.x, tid = threadIdx.x, lane = tid & 63;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int c = lane & 15, g = lane >> 4;
    // row within its tile of this lane's accumulator item i (the MFMA
    // accumulator layout of the precision; column c for every item)
    auto rowoff = [&](int i) { return mfma_row(R(0), lane, i); };
    C *A = h.a + (size_t)kb * h.stride;
    double *dd = h.d + (size_t)kb * N, *ee = h.e + (size_t)kb * N;
    const C one = cmake<C, R>(R(1), R(0));

    // tiles of this wave: (tI, tJ), tI = -1 where the index is beyond the last
    int tI[MT], tJ[MT];
#pragma unroll
    for (int k = 0; k < MT; ++k) {
        const int t = wv + RZ_WAVES * k;
        int J = NT - 1;
        while (J >= 0 && t >= rz_cnt(NT, J)) --J;
        tJ[k] = J;
        tI[k] = J >= 0 ? J + t - rz_cnt(NT, J + 1) : -1;
    }
    // the lower triangle, real and imaginary parts as MFMA accumulators
    // (the upper triangle of a diagonal tile from the lower one); every
    // load is issued unconditionally from a clamped address
    v4 xr[MT], xi[MT];
#pragma unroll
    for (int k = 0; k < MT; ++k) {
        const int cc = 16 * tJ[k] + c;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int r = 16 * tI[k] + rowoff(i);
            const bool ok = tI[k] >= 0 && r < N && cc < N;
            const size_t off = !ok ? 0 : (r >= cc ? (size_t)cc * N + r : (size_t)r * N + cc);
            const C a = A[off];
            xr[k][i] = ok ? a.x : R(0);
            xi[k][i] = !ok || r == cc ? R(0) : (r > cc ? a.y : -a.y);
        }
    }
    for (int r = tid; r < 16 * NTM; r += RZ_THREADS) {
        S.vb[r] = czero<C>(); S.wb[r] = czero<C>(); S.colb[r] = czero<C>();
    }
    __syncthreads();

    for (int j = 0; j < N; ++j) {
        // (U) the previous column's rank-2 update A -= v w^H + w v^H on tile
        // columns >= j / 16, two MFMAs per tile with K = 4:
        //   Re: [-v.x, -v.y, -w.x, -w.y]_r . [w.x, w.y, v.x, v.y]_c
        //   Im: [-v.y,  v.x, -w.y,  w.x]_r . [w.x, w.y, v.x, v.y]_c
        // (lane l supplies row l % 16 of the first operand and column l % 16
        // of the second, component l / 16); column j into colb
        const int Ju = j >> 4, Pu = rz_cnt(NT, Ju), jc = j & 15;
#pragma unroll
        for (int k = 0; k < MT; ++k) {
            if (wv + RZ_WAVES * k < Pu) {
                // the tile coordinates pass through an opaque copy each
                // column, so that no per-tile lane address is kept live
                // across the column loop (they are recomputed at each use)
                int I = tI[k], J = tJ[k];
                asm volatile("" : "+s"(I), "+s"(J));
                if (j > 0) {
                    const C pv = g < 2 ? S.vb[16 * I + c] : S.wb[16 * I + c];
                    const C qv = g < 2 ? S.wb[16 * J + c] : S.vb[16 * J + c];
                    const R are = (g & 1) ? -pv.y : -pv.x;
                    const R aim = (g & 1) ? pv.x : -pv.y;
                    const R bq = (g & 1) ? qv.y : qv.x;
                    xr[k] = mfma16(are, bq, xr[k]);
                    xi[k] = mfma16(aim, bq, xi[k]);
                }
                if (J == Ju && c == jc) {
#pragma unroll
                    for (int i = 0; i < 4; ++i) S.colb[16 * I + rowoff(i)] = cmake<C, R>(xr[k][i], xi[k][i]);
                }
            }
        }
        __syncthreads();
        if (j == N - 1) {                   // the last diagonal element
            if (tid == 0) { dd[j] = (double)S.colb[j].x; ee[j] = 0.0; }
            for (int r = tid; r < N; r += RZ_THREADS) A[(size_t)j * N + r] = czero<C>();
            break;
        }
        const int pcol = j % NB, i0 = j - pcol;
        // (R) reflector of x = colb[j+1 ..]: H^H x = beta e1, H = I - tau v v^H
        R xs = 0;
#pragma unroll
        for (int q = 0; q < (16 * NTM + 63) / 64; ++q) {
            const int r = j + 2 + lane + 64 * q;
            const C a = S.colb[r < N ? r : 0];
            xs += r < N ? a.x * a.x + a.y * a.y : R(0);
        }
        xs = wave_sum_fast(xs);
        C tau, scale;
        R beta;
        rz_larfg<R, C>(S.colb[j + 1], xs, tau, beta, scale);
        if (tid == 0) { dd[j] = (double)S.colb[j].x; ee[j] = (double)beta; }
        {
            const int r = 40 * wv + lane;
            if (lane < 40 && r < 16 * NT) {
                C v = czero<C>();
                if (r == j + 1) v = one;
                else if (r > j + 1 && r < N) v = zmul(S.colb[r], scale);
                S.vb[r] = v;
                S.Vp[pcol][r] = v;
                if (r < N) A[(size_t)j * N + r] = v;
            }
        }
        __syncthreads();
        // (H) y = A v over the tile columns >= (j + 1) / 16; first the panel
        // dots V_l^H v (l < pcol), 16 threads each over rows > j
        {
            const int l = tid >> 4, part = tid & 15;
            if (l < pcol) {
                C a0 = czero<C>(), a1 = czero<C>();
                int r = j + 1 + part;
                for (; r + 16 < N; r += 32) {
                    a0 = zadd(a0, zcmul(S.Vp[l][r], S.vb[r]));
                    a1 = zadd(a1, zcmul(S.Vp[l][r + 16], S.vb[r + 16]));
                }
                if (r < N) a0 = zadd(a0, zcmul(S.Vp[l][r], S.vb[r]));
                a0 = zadd(a0, a1);
                a0.x = row16_sum(a0.x);
                a0.y = row16_sum(a0.y);
                if (part == 0) S.Gc[l] = a0;
            }
        }
        const int Jh = (j + 1) >> 4, Ph = rz_cnt(NT, Jh);
        R vav = 0;
#pragma unroll
        for (int k = 0; k < MT; ++k) {
            const int t = wv + RZ_WAVES * k;
            if (t < Ph) {
                int I = tI[k], J = tJ[k];
                asm volatile("" : "+s"(I), "+s"(J));
                const C vJ = S.vb[16 * J + c];
                R vals[8];
                C q = czero<C>();
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const C vI = S.vb[16 * I + rowoff(i)];
                    const C x = cmake<C, R>(xr[k][i], xi[k][i]);
                    const C pr = zmul(x, vJ);
                    vals[2 * i] = pr.x;
                    vals[2 * i + 1] = pr.y;
                    q = zadd(q, zcmul(x, vI));                // conj(A) v_I
                }
                // lane c holds component (c >> 1) & 1 of the sum of row
                // rowoff(c >> 2) (in lanes c and c ^ 1); v^H A v takes
                // Re(conj(v_r) (A_IJ v_J)_r) from the even lanes, twice below
                // the diagonal (the tile's mirror)
                const R rs = row16_sum8<R>(vals, c);
                const int rrow = 16 * I + rowoff(c >> 2);
                const R vcomp = reinterpret_cast<const R *>(&S.vb[rrow])[(c >> 1) & 1];
                if ((c & 1) == 0) {
                    reinterpret_cast<R *>(&S.slotR[t][rowoff(c >> 2)])[(c >> 1) & 1] = rs;
                    vav += (I == J ? R(1) : R(2)) * (vcomp * rs);
                }
                if (I != J) {
                    q.x = xor32_sum(xor16_sum(q.x));
                    q.y = xor32_sum(xor16_sum(q.y));
                    if (g == 0) S.slotC[t][c] = q;
                }
            }
        }
        vav = wave_sum_fast(vav);
        if (lane == 0) S.red[wv] = vav;
        __syncthreads();
        // (W) w = tau y - |tau|^2 (v^H A v) / 2 v; the WY column
        // T[0:pcol, pcol] = -tau T[0:pcol, 0:pcol] G, T[pcol, pcol] = tau
        // (16 threads per row)
        {
            R s = 0;
#pragma unroll
            for (int q = 0; q < RZ_WAVES; ++q) s += S.red[q];
            const R hw = R(-0.5) * (tau.x * tau.x + tau.y * tau.y) * s;
            const int r = tid;
            if (r < 16 * NT) {
                C wr = czero<C>();
                if (r > j && r < N) {
                    const int I = r >> 4, a = r & 15;
                    C y = czero<C>();
                    for (int J = Jh; J <= I; ++J) y = zadd(y, S.slotR[rz_cnt(NT, J + 1) + I - J][a]);
                    const int base = rz_cnt(NT, I + 1) - I;
                    for (int I2 = I + 1; I2 < NT; ++I2) y = zadd(y, S.slotC[base + I2][a]);
                    wr = zadd(zmul(tau, y), zscale(hw, S.vb[r]));
                }
                S.wb[r] = wr;
            }
            const int l = tid >> 4, part = tid & 15;
            C acc = czero<C>();
            for (int q = l + part; q < pcol; q += 16) acc = zadd(acc, zmul(S.Ts[l][q], S.Gc[q]));
            acc.x = row16_sum(acc.x);
            acc.y = row16_sum(acc.y);
            if (part == 0) {
                if (l < pcol) S.Ts[l][pcol] = zmul(cmake<C, R>(-tau.x, -tau.y), acc);
                else if (l == pcol) S.Ts[pcol][pcol] = tau;
            }
        }
        __syncthreads();
        if (pcol == NB - 1 || j == N - 2) {
            // the panel is complete: T, and V T^H (rows > i0) on the matrix
            // cores, 16 x 16 output tiles (rows 16 rt, columns 16 ct) with
            // K = pw in steps of 4; conj(T)^T entries outside 0 <= cc <= l < pw
            // are zeros
            const int pw = pcol + 1;
            C *Tp = h.t + (size_t)kb * h.tstride + (size_t)(i0 / NB) * NB * NB;
            for (int q = tid; q < NB * NB; q += RZ_THREADS) {
                const int row = q % NB, col = q / NB;
                Tp[q] = (row <= col && col < pw) ? S.Ts[row][col] : czero<C>();
            }
            C *VT = h.vt + (size_t)kb * h.stride + (size_t)i0 * N;
            const int ntile = NT * 2;
            for (int ot = wv; ot < ntile; ot += RZ_WAVES) {
                const int rt = ot >> 1, ct = ot & 1;
                v4 p1 = {0, 0, 0, 0}, p2 = {0, 0, 0, 0}, p3 = {0, 0, 0, 0};
                for (int k0 = 0; k0 < pw; k0 += 4) {
                    const int l = k0 + g;
                    const C a = l < pw ? S.Vp[l][16 * rt + c] : czero<C>();
                    const int cc = 16 * ct + c;
                    const C t = (l < pw && cc <= l) ? S.Ts[cc][l] : czero<C>();
                    const R br = t.x, bi = -t.y;                 // conj(T[cc][l])
                    p1 = mfma16(a.x, br, p1);
                    p2 = mfma16(a.y, bi, p2);
                    p3 = mfma16(a.x + a.y, br + bi, p3);
                }
                const int cc = 16 * ct + c;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int r = 16 * rt + rowoff(i);
                    if (r < N && i0 + cc < N)
                        VT[(size_t)cc * N + r] = cmake<C, R>(p1[i] - p2[i], p3[i] - p1[i] - p2[i]);
                }
            }
            // (the next column rewrites Vp and Ts only after its first barrier)
        }
    }
}

// ---- stage 1: panel QR ------------------------------------------------------
// One workgroup of NW waves per item; the panel's m = n - c0 - TB rows are
// held in registers, row tid + k NT in slot k (RPT slots).  Column j: the
// norm and the TB dot products of the reflector with the other columns (the
// later columns for the update, the earlier reflectors for T) are each one
// block reduction.  Writes: the reflectors V (unit diagonal, explicit) into
// the panel of A and into P columns 0..TB-1 and 2TB..3TB-1, T, and the band
// columns c0..c0+TB-1 (V T and V T^H follow as GEMMs).
template <class R, int NW, int RPT>
__global__ __launch_bounds__(64 * NW) void k_band_qr(BandArgsT<R> g, int c0, int panel)
{
    typedef typename cplx<R>::type C;
    constexpr int NT = 64 * NW;
    const int n = g.n, m = n - c0 - TB, s0 = c0 + TB, kb = blockIdx.x;
    C *A = g.a + (size_t)kb * g.ss;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    __shared__ R red[NW][2 * TB];
    __shared__ R red2[2 * TB];         // the partial sums combined over the waves
    __shared__ R rnorm[NW];
    __shared__ C alpha_sh;
    __shared__ C Tm[TB][TB + 1];
    for (int e = tid; e < TB * (TB + 1); e += NT) Tm[e / (TB + 1)][e % (TB + 1)] = czero<C>();
    C p[RPT][TB];
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
        const int r = tid + k * NT;
#pragma unroll
        for (int c = 0; c < TB; ++c)
            p[k][c] = r < m ? A[(size_t)(c0 + c) * n + s0 + r] : czero<C>();
    }
    // Columns rotate through the slots: at step j, slot q holds column
    // (q + j) mod TB, so the current column is always slot 0, the later
    // columns slots 1 .. TB-1-j and the finished reflectors the rest (no
    // register array is indexed by j); after TB steps the order is restored.
    const C one = cmake<C, R>(R(1), R(0));
    for (int j = 0; j < TB; ++j) {
        R nrm = 0;
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
            const int r = tid + k * NT;
            const R x2 = p[k][0].x * p[k][0].x + p[k][0].y * p[k][0].y;
            nrm += r > j ? x2 : R(0);
        }
        if (tid == j) alpha_sh = p[0][0];
        nrm = wsum(nrm);
        if (lane == 0) rnorm[wv] = nrm;
        __syncthreads();
        R xn2 = 0;
#pragma unroll
        for (int w = 0; w < NW; ++w) xn2 += rnorm[w];
        const C alpha = alpha_sh;
        C tau = czero<C>(), scale = czero<C>();
        R beta = alpha.x;
        const bool have = j < m - 1;            // rows j and j + 1 exist
        if (have) larfg_coef<R, C>(alpha, xn2, tau, beta, scale);
        // v (rows > j scaled in place; zero above row j, and rows >= m hold
        // zeros throughout), R(j, j) = beta
        C vr[RPT];
        const C sc = have ? scale : one;
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
            const int r = tid + k * NT;
            const C x = p[k][0];
            const C v = r > j ? zmul(x, sc) : (r == j ? one : czero<C>());
            vr[k] = v;
            p[k][0] = r == j ? (have ? cmake<C, R>(beta, R(0)) : alpha) : (r > j ? v : x);
        }
        // partial dots by slot q (column c = (q + j) mod TB): the earlier
        // reflectors (q >= TB - j): conj(V[r, c]) v[r]; the later columns
        // (1 <= q < TB - j): conj(v[r]) P[r, c]; two halves of TB / 2 slots.
        // v is zero above row j, so every row can enter every sum.
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            R part[TB];
#pragma unroll
            for (int qq = 0; qq < TB / 2; ++qq) {
                const int q = h * (TB / 2) + qq;
                C ea = czero<C>(), eb = czero<C>();
#pragma unroll
                for (int k = 0; k < RPT; ++k) {
                    ea = zadd(ea, zcmul(p[k][q], vr[k]));
                    eb = zadd(eb, zcmul(vr[k], p[k][q]));
                }
                const C acc = q >= TB - j ? ea : (q >= 1 ? eb : czero<C>());
                part[2 * qq] = acc.x;
                part[2 * qq + 1] = acc.y;
            }
            const R tot = wave_sum_transposed<R, TB>(part, lane);
            if ((lane & 3) == 0) {
                const int v = lane >> 2, q = h * (TB / 2) + (v >> 1), c = (q + j) & (TB - 1);
                red[wv][2 * c + (v & 1)] = tot;                 // stored by column
            }
        }
        __syncthreads();
        if (NW > 1) {
            if (tid < 2 * TB) {
                R t = 0;
#pragma unroll
                for (int u = 0; u < NW; ++u) t += red[u][tid];
                red2[tid] = t;
            }
            __syncthreads();
        }
        const R *rs = NW > 1 ? red2 : red[0];
        // update of the later columns: P[r, c] -= conj(tau) v[r] w[c] (rows
        // above j have v = 0; the earlier reflectors' slots take f = 0)
        const C ctau = zconj(tau);
#pragma unroll
        for (int q = 1; q < TB; ++q) {
            const int c = (q + j) & (TB - 1);
            const C w = cmake<C, R>(rs[2 * c], rs[2 * c + 1]);
            const C f = q < TB - j ? zmul(ctau, w) : czero<C>();
#pragma unroll
            for (int k = 0; k < RPT; ++k) p[k][q] = zsub(p[k][q], zmul(vr[k], f));
        }
        // column j of T: T[l, j] = -tau sum_{q=l}^{j-1} T[l, q] (V^H v)[q]
        if (tid < j) {
            C acc = czero<C>();
            for (int q = tid; q < j; ++q) acc = zadd(acc, zmul(Tm[tid][q], cmake<C, R>(rs[2 * q], rs[2 * q + 1])));
            Tm[tid][j] = zmul(cmake<C, R>(-tau.x, -tau.y), acc);
        }
        if (tid == j) Tm[j][j] = tau;
        // rotate the slots left by one
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
            const C first = p[k][0];
#pragma unroll
            for (int q = 0; q < TB - 1; ++q) p[k][q] = p[k][q + 1];
            p[k][TB - 1] = first;
        }
        __syncthreads();
    }
    // ---- outputs ----
    C *P = g.p + (size_t)kb * g.sp;
    C *Tg = g.t + (size_t)kb * g.st + (size_t)panel * TB * TB;
    for (int e = tid; e < TB * TB; e += NT) Tg[e] = Tm[e % TB][e / TB];      // column-major
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
        const int r = tid + k * NT;
        if (r >= m) continue;
        const size_t row = (size_t)s0 + r;
#pragma unroll
        for (int c = 0; c < TB; ++c) {
            const C vc = r > c ? p[k][c] : (r == c ? one : czero<C>());   // V[r, c]
            A[(size_t)(c0 + c) * n + row] = vc;
            P[(size_t)c * n + row] = vc;
            P[(size_t)(2 * TB + c) * n + row] = vc;
        }
    }
    // band columns c0 .. c0+TB-1: the diagonal block from A, R, zeros
    C *band = g.band + (size_t)kb * g.sb;
    if (tid < TB) {
#pragma unroll
        for (int c = 0; c < TB; ++c)
            if (c >= tid) band[(size_t)(c0 + c) * TBLD + TB + tid - c] = tid < m ? p[0][c] : czero<C>();
    }
    for (int e = tid; e < TB * TBLD; e += NT) {
        const int c = e / TBLD, dd = e % TBLD;
        if (dd <= TB - 1 - c) band[(size_t)(c0 + c) * TBLD + dd] = A[(size_t)(c0 + c) * n + c0 + c + dd];
        else if (dd > TB) band[(size_t)(c0 + c) * TBLD + dd] = czero<C>();
    }
    // the panel's columns of A and vt above the reflectors are zero for the
    // back-transform, whose merged panels read them from the first
    // reflector row of the group (the diagonal block was read just above)
    __syncthreads();
    C *VT = g.vt + (size_t)kb * g.ss;
    for (int e = tid; e < TB * s0; e += NT) {
        const int c = e / s0, r = e % s0;
        A[(size_t)(c0 + c) * n + r] = czero<C>();
        VT[(size_t)(c0 + c) * n + r] = czero<C>();
    }
}

// Band columns c_end .. n-1 (after the last panel) from the trailing block.
template <class R>
__global__ __launch_bounds__(256) void k_band_tail(BandArgsT<R> g, int c_end)
{
    typedef typename cplx<R>::type C;
    const int n = g.n, kb = blockIdx.x;
    const C *A = g.a + (size_t)kb * g.ss;
    C *band = g.band + (size_t)kb * g.sb;
    for (int e = threadIdx.x; e < (n - c_end) * TBLD; e += blockDim.x) {
        const int c = c_end + e / TBLD, dd = e % TBLD;
        band[(size_t)c * TBLD + dd] = (dd <= TB && c + dd < n) ? A[(size_t)c * n + c + dd] : czero<C>();
    }
}

// ---- stage 2: bulge chasing -------------------------------------------------
// Sweep i annihilates column i below the subdiagonal; its step s works on
// rows r0 = i + 1 + s TB .. r0 + TB - 1: the previous step's reflector is
// applied from the right to the block E = A[r0:, r0-TB:r0] (filling it in
// below the band), a new reflector annihilates E's first column below its
// first row and is applied from the left to E and two-sided to the
// diagonal block D = A[r0:, r0:].  Step (i, s) runs at time t = 2 i + s:
// steps at the same time touch disjoint elements, and every step that
// must precede it ran at an earlier time.  One wave per step; a barrier
// between times.  The band is stored by columns, band[c TBLD + (r - c)],
// r - c < 2 TB (the fill of E reaches r - c = 2 TB - 1).
// Reflectors of the sweeps in flight are kept in LDS slots (i mod NSLOT):
// at most (steps of sweep 0) / 2 + 1 sweeps run at once: 10 at order 288
// (the largest FP64 band kept in LDS), 19 at 576 (FP32), 41 at 1280.

__device__ __forceinline__ int bulge_steps(int n, int i) { return (n - 2 - i) / TB + 1; }

template <class R, class C, int NSLOT>
__device__ __forceinline__ void bulge_step(C *band, int n, int i, int s, C (*slots)[TB], C *v2, const int *dbase,
                                           int lane)
{
    const int a = lane & (TB - 1), cq = lane >> 4;       // row a, columns 4 cq .. 4 cq + 3
    const int r0 = i + 1 + s * TB, L = min(TB, n - r0);
    const C one = cmake<C, R>(R(1), R(0));
    C tau = czero<C>(), va = czero<C>();
    R beta = 0;
    // the diagonal block D = A[r0:, r0:] (Hermitian, from the lower band),
    // loaded first: nothing below writes it before its own update
    C D[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int c = 4 * cq + q;
        C dv = czero<C>();
        if (a < L && c < L)
            dv = c <= a ? band[(size_t)(r0 + c) * TBLD + a - c] : zconj(band[(size_t)(r0 + a) * TBLD + c - a]);
        D[q] = dv;
    }
    if (s == 0) {
        // column i, rows r0 .. r0 + L - 1 (L >= 2)
        const C x = (cq == 0 && a < L) ? band[(size_t)i * TBLD + 1 + a] : czero<C>();
        const R xn2 = wsum((cq == 0 && a >= 1) ? x.x * x.x + x.y * x.y : R(0));
        const C alpha = cshfl(x, 0);
        C scale;
        larfg_coef<R, C>(alpha, xn2, tau, beta, scale);
        const C xa = cshfl(x, a);
        va = a == 0 ? one : (a < L ? zmul(xa, scale) : czero<C>());
        if (cq == 0 && a < L)
            band[(size_t)i * TBLD + 1 + a] = a == 0 ? cmake<C, R>(beta, R(0)) : czero<C>();
    } else {
        const int p0 = r0 - TB;
        const int sl = i % NSLOT;
        const C taup = slots[sl][0];
        C E[4], vp[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int c = 4 * cq + q;
            E[q] = a < L ? band[(size_t)(p0 + c) * TBLD + TB + a - c] : czero<C>();
            vp[q] = c == 0 ? one : slots[sl][c];
        }
        // E <- E Hp: y = E vp, E -= taup y vp^H
        C y = czero<C>();
#pragma unroll
        for (int q = 0; q < 4; ++q) y = zadd(y, zmul(E[q], vp[q]));
        y = zadd(y, cshfl_xor(y, 16));
        y = zadd(y, cshfl_xor(y, 32));
        const C ty = zmul(taup, y);
#pragma unroll
        for (int q = 0; q < 4; ++q) E[q] = zsub(E[q], zmul(ty, zconj(vp[q])));
        if (L >= 2) {
            // reflector of E's first column (lanes cq == 0, slot 0)
            const C x = cq == 0 ? E[0] : czero<C>();
            const R xn2 = wsum((cq == 0 && a >= 1) ? x.x * x.x + x.y * x.y : R(0));
            const C alpha = cshfl(x, 0);
            C scale;
            larfg_coef<R, C>(alpha, xn2, tau, beta, scale);
            const C xa = cshfl(x, a);
            va = a == 0 ? one : (a < L ? zmul(xa, scale) : czero<C>());
            // E[:, 1:] -= conj(tau) v (v^H E[:, 1:])
            const C ct = zconj(tau);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                C w = zcmul(va, E[q]);
                w = zadd(w, cshfl_xor(w, 1));
                w = zadd(w, cshfl_xor(w, 2));
                w = zadd(w, cshfl_xor(w, 4));
                w = zadd(w, cshfl_xor(w, 8));
                const int c = 4 * cq + q;
                if (c == 0) E[q] = a == 0 ? cmake<C, R>(beta, R(0)) : czero<C>();
                else E[q] = zsub(E[q], zmul(zmul(ct, va), w));
            }
        }
        if (a < L) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int c = 4 * cq + q;
                band[(size_t)(p0 + c) * TBLD + TB + a - c] = E[q];
            }
        }
        if (L < 2) return;
    }
    // the reflector, for the next step of this sweep and the back-transform
    const int sl = i % NSLOT;
    const int jb = i / TB;
    if (cq == 0) {
        const C sv = a == 0 ? tau : va;
        slots[sl][a] = sv;
        v2[((size_t)(dbase[jb] + s) * TB + (i - jb * TB)) * TB + a] = sv;
    }
    // D <- H^H D H: p = D v, w = tau p - 1/2 |tau|^2 (v^H D v) v,
    // D -= v w^H + w v^H
    C vc[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) vc[q] = cshfl(va, 4 * cq + q);
    C pp = czero<C>();
#pragma unroll
    for (int q = 0; q < 4; ++q) pp = zadd(pp, zmul(D[q], vc[q]));
    pp = zadd(pp, cshfl_xor(pp, 16));
    pp = zadd(pp, cshfl_xor(pp, 32));
    const C wv = zmul(tau, pp);
    C gam = zcmul(va, wv);
    gam = zadd(gam, cshfl_xor(gam, 1));
    gam = zadd(gam, cshfl_xor(gam, 2));
    gam = zadd(gam, cshfl_xor(gam, 4));
    gam = zadd(gam, cshfl_xor(gam, 8));
    const C h = zscale(R(0.5), zmul(zconj(tau), gam));
    const C wp = zsub(wv, zmul(h, va));
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int c = 4 * cq + q;
        const C wpc = cshfl(wp, c);
        D[q] = zsub(D[q], zadd(zmul(va, zconj(wpc)), zmul(wp, zconj(vc[q]))));
        if (a < L && c <= a) band[(size_t)(r0 + c) * TBLD + a - c] = D[q];
    }
}

// One workgroup per item; dynamic LDS holds the band when LDSB.
template <class R, bool LDSB>
__global__ __launch_bounds__(1024) void k_bulge(BulgeArgsT<R> g)
{
    typedef typename cplx<R>::type C;
    constexpr int NSLOT = LDSB ? (sizeof(R) == 8 ? 16 : 32) : 64;
    extern __shared__ unsigned char bulge_dyn[];
    __shared__ C slots[NSLOT][TB];
    __shared__ int dbase[kMaxBandBlocks + 1];
    const int n = g.n, kb = blockIdx.x;
    const int nw = blockDim.x >> 6, wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    C *gband = g.band + (size_t)kb * g.sb;
    C *band = LDSB ? reinterpret_cast<C *>(bulge_dyn) : gband;
    const int nj = (n - 3) / TB + 1;                     // sweep blocks
    if (threadIdx.x == 0) {
        int b = 0;
        for (int j = 0; j < nj; ++j) { dbase[j] = b; b += bulge_steps(n, j * TB); }
        dbase[nj] = b;
    }
    if (LDSB)
        for (int e = threadIdx.x; e < n * TBLD; e += blockDim.x) band[e] = gband[e];
    __syncthreads();
    C *v2 = g.v2 + (size_t)kb * g.sv;
    for (int t = 0; t <= 2 * n - 6; ++t) {
        const int ihi = min(t >> 1, n - 3);
        for (int i = ihi - wv; i >= 0; i -= nw) {
            const int s = t - 2 * i;
            if (s >= bulge_steps(n, i)) break;
            bulge_step<R, C, NSLOT>(band, n, i, s, slots, v2, dbase, lane);
        }
        __syncthreads();
    }
    // tridiagonal: d, |e|, and the phases that make e real
    double *d = g.d + (size_t)kb * g.sde, *e = g.e + (size_t)kb * g.sde;
    double2 *ph = g.ph + (size_t)kb * g.sph;
    for (int r = threadIdx.x; r < n; r += blockDim.x) {
        d[r] = (double)band[(size_t)r * TBLD].x;
        if (r < n - 1) {
            const C ec = band[(size_t)r * TBLD + 1];
            e[r] = sqrt((double)ec.x * ec.x + (double)ec.y * ec.y);
        } else {
            e[r] = 0.0;
        }
    }
    if (threadIdx.x == 0) {
        double2 f = make_double2(1.0, 0.0);
        ph[0] = f;
        for (int r = 0; r < n - 1; ++r) {
            const C ec = band[(size_t)r * TBLD + 1];
            const double ax = sqrt((double)ec.x * ec.x + (double)ec.y * ec.y);
            if (ax > 0.0) f = zmul(f, make_double2((double)ec.x / ax, (double)ec.y / ax));
            ph[r + 1] = f;
        }
    }
}

// ---- back-transform through the stage-2 reflectors -------------------------
// Diamond (j, s) groups the reflectors of sweeps jTB .. jTB+TB-1 at step s:
// reflector c acts on rows jTB + 1 + sTB + c .. + TB - 1, so their WY form
// I - V T V^H spans 2 TB - 1 rows with V[c + q, c] = v_c[q].  Reflectors
// that do not exist (sweeps beyond n - 3, steps beyond a sweep's last, or a
// last step of one row) are cleared to zero here; T is built by the forward
// recurrence T[:c, c] = -tau_c T[:c, :c] (V^H V)[:c, c].  One wave per
// diamond.
template <class R>
__global__ __launch_bounds__(64) void k_q2_tfactors(Q2ArgsT<R> g)
{
    typedef typename cplx<R>::type C;
    __shared__ C vv[TB][TB];       // vv[c][q] = v_c[q] (q = 0: tau)
    __shared__ C G[TB][TB + 1];
    __shared__ C Tm[TB][TB + 1];
    const int n = g.n, kb = blockIdx.y, dix = blockIdx.x, lane = threadIdx.x;
    const int nj = (n - 3) / TB + 1;
    int j = 0, b = 0;
    while (j < nj && b + bulge_steps(n, j * TB) <= dix) { b += bulge_steps(n, j * TB); ++j; }
    if (j >= nj) return;
    const int s = dix - b;
    C *blk = g.v2 + (size_t)kb * g.sv + (size_t)dix * TB * TB;
    for (int e = lane; e < TB * TB; e += 64) {
        const int c = e / TB, q = e % TB, i = j * TB + c;
        const bool ok = i <= n - 3 && s < bulge_steps(n, i) && min(TB, n - (i + 1 + s * TB)) >= 2 &&
                        q < min(TB, n - (i + 1 + s * TB));
        C x = ok ? blk[e] : czero<C>();
        blk[e] = x;
        vv[c][q] = x;
    }
    __syncthreads();
    // G[l][c] = (V^H V)[l][c], l < c: sum over rows of conj(v_l[r - l]) v_c[r - c]
    for (int e = lane; e < TB * TB; e += 64) {
        const int l = e / TB, c = e % TB;
        if (l >= c) continue;
        C acc = czero<C>();
        for (int r = c; r < l + TB; ++r) {
            const C vl = (r == l) ? cmake<C, R>(R(1), R(0)) : vv[l][r - l];
            const C vcc = (r == c) ? cmake<C, R>(R(1), R(0)) : vv[c][r - c];
            acc = zadd(acc, zcmul(vl, vcc));
        }
        G[l][c] = acc;
    }
    __syncthreads();
    for (int c = 0; c < TB; ++c) {
        const C tc = vv[c][0];
        if (lane < c) {
            C acc = czero<C>();
            for (int q = lane; q < c; ++q) acc = zadd(acc, zmul(Tm[lane][q], G[q][c]));
            Tm[lane][c] = zmul(cmake<C, R>(-tc.x, -tc.y), acc);
        } else if (lane < TB) {
            Tm[lane][c] = lane == c ? tc : czero<C>();
        }
        __syncthreads();
    }
    C *T = g.t2 + (size_t)kb * g.st2 + (size_t)dix * TB * TB;
    for (int e = lane; e < TB * TB; e += 64) T[e] = Tm[e % TB][e / TB];     // column-major
}

// Y <- Q2 Y: diamonds by block j descending, step s ascending; each is
//   W = V^H Y_rows (16 x 16),  W <- T W,  Y_rows -= V W  (32 rows)
// on the matrix cores, complex products as three real ones.  A workgroup
// stages each diamond's V (rows padded to 32) and T in LDS, double-buffered
// (one barrier per diamond); every wave holds q2_tiles<R>() tiles of 16
// columns of Y (FP64: one, with three workgroups per CU, so that more
// workgroups hide the latency of each diamond's chain of products; FP32:
// two, two per CU).  The K index of every product is laid out like the rows of an
// accumulator (slot q of lane group g4 is row mfma_row(q)), so each
// accumulator is the next product's B operand in the same lane, and the 32
// rows of Y a wave holds per tile (two halves of 16) are both B operand and
// update target.  Consecutive diamonds of a block overlap by 16 rows: the
// lower half moves up and one new half is loaded.
constexpr int kQ2Waves = 4;
template <class R> constexpr int q2_tiles() { return sizeof(R) == 8 ? 1 : 2; }
template <class R>
__global__ __launch_bounds__(64 * kQ2Waves, (sizeof(R) == 8 ? 3 : 2)) void k_apply_q2(Q2ArgsT<R> g)
{
    typedef typename cplx<R>::type C;
    typedef R v4 __attribute__((ext_vector_type(4)));
    constexpr int NT = q2_tiles<R>();
    __shared__ C Vs[2][32][TB + 1];
    __shared__ C Ts[2][TB][TB + 1];
    const int n = g.n, kb = blockIdx.y, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int ncols = g.ncol ? min(g.ncol[kb], n) : n;
    const int col0 = blockIdx.x * kQ2Waves * 16 * NT;
    if (col0 >= ncols) return;                           // uniform over the workgroup
    const int wc0 = col0 + wv * 16 * NT;
    const int r16 = lane & 15;
    bool colok[NT];
    C *Y[NT];
#pragma unroll
    for (int u = 0; u < NT; ++u) {
        colok[u] = wc0 + 16 * u + r16 < ncols;
        Y[u] = g.y + (size_t)kb * g.sy + (size_t)min(wc0 + 16 * u + r16, n - 1) * n;
    }
    const C *V2 = g.v2 + (size_t)kb * g.sv;
    const C *T2 = g.t2 + (size_t)kb * g.st2;
    int rm[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) rm[q] = mfma_row(R(0), lane, q);
    const int nj = (n - 3) / TB + 1;
    int dtot = 0;
    for (int j = 0; j < nj; ++j) dtot += bulge_steps(n, j * TB);
    // V[r][c] = v_c[r - c] (1 on the diagonal, 0 outside the staircase): the
    // zeros are written once, each diamond then fills the staircase; thread
    // e = c TB + q (256 threads for the TB^2 entries) moves entry q of
    // reflector c and entry e of T, loaded one diamond ahead
    static_assert(64 * kQ2Waves == TB * TB, "one V and one T entry per thread");
    for (int e = tid; e < 2 * 32 * (TB + 1); e += 64 * kQ2Waves) (&Vs[0][0][0])[e] = czero<C>();
    __syncthreads();
    const int sc = tid / TB, sq = tid % TB;
    auto put = [&](int b, C pv, C pt) {
        Vs[b][sc + sq][sc] = sq == 0 ? cmake<C, R>(R(1), R(0)) : pv;
        Ts[b][sq][sc] = pt;
    };
    // diamonds in application order: j descending, s ascending
    int j = nj - 1, s = 0, d0 = dtot - bulge_steps(n, j * TB), ns = bulge_steps(n, j * TB);
    put(0, V2[(size_t)d0 * TB * TB + tid], T2[(size_t)d0 * TB * TB + tid]);
    C y[NT][2][4];
    int it = 0;
    while (true) {
        const int rlo = j * TB + 1 + s * TB;
        if (s == 0) {                                    // a new block: load its first 32 rows
#pragma unroll
            for (int u = 0; u < NT; ++u)
#pragma unroll
                for (int h = 0; h < 2; ++h)
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const int r = rlo + 16 * h + rm[q];
                        y[u][h][q] = (colok[u] && r < n) ? Y[u][r] : czero<C>();
                    }
        }
        // the next diamond (position in the order): its V and T entries and
        // this block's next 16 rows of Y are loaded now, used after the products
        int jn = j, sn = s + 1, d0n = d0, nsn = ns;
        if (sn == ns) { jn = j - 1; sn = 0; if (jn >= 0) { nsn = bulge_steps(n, jn * TB); d0n = d0 - nsn; } }
        C pv = czero<C>(), pt = czero<C>();
        if (jn >= 0) {
            pv = V2[(size_t)(d0n + sn) * TB * TB + tid];
            pt = T2[(size_t)(d0n + sn) * TB * TB + tid];
        }
        const bool last = s + 1 == ns;
        C ynext[NT][4];
#pragma unroll
        for (int u = 0; u < NT; ++u)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int rn = rlo + 32 + rm[q];
                ynext[u][q] = (!last && colok[u] && rn < n) ? Y[u][rn] : czero<C>();
            }
        __syncthreads();                                 // buffer it & 1 staged; buffer (it+1) & 1 free
        const int b = it & 1;
#pragma unroll
        for (int u = 0; u < NT; ++u) {
            // W = V^H Y
            v4 a1 = {0, 0, 0, 0}, a2 = {0, 0, 0, 0}, a3 = {0, 0, 0, 0};
#pragma unroll
            for (int h = 0; h < 2; ++h)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const C v = Vs[b][16 * h + rm[q]][r16];
                    const R ar = v.x, ai = -v.y;
                    a1 = mfma16(ar, y[u][h][q].x, a1);
                    a2 = mfma16(ai, y[u][h][q].y, a2);
                    a3 = mfma16(ar + ai, y[u][h][q].x + y[u][h][q].y, a3);
                }
            C w[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) w[q] = cmake<C, R>(a1[q] - a2[q], a3[q] - a1[q] - a2[q]);
            // W <- T W
            a1 = v4{0, 0, 0, 0}; a2 = v4{0, 0, 0, 0}; a3 = v4{0, 0, 0, 0};
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const C t = Ts[b][r16][rm[q]];
                a1 = mfma16(t.x, w[q].x, a1);
                a2 = mfma16(t.y, w[q].y, a2);
                a3 = mfma16(t.x + t.y, w[q].x + w[q].y, a3);
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) w[q] = cmake<C, R>(a1[q] - a2[q], a3[q] - a1[q] - a2[q]);
            // Y_rows -= V W
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                a1 = v4{0, 0, 0, 0}; a2 = v4{0, 0, 0, 0}; a3 = v4{0, 0, 0, 0};
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const C v = Vs[b][16 * h + r16][rm[q]];
                    a1 = mfma16(v.x, w[q].x, a1);
                    a2 = mfma16(v.y, w[q].y, a2);
                    a3 = mfma16(v.x + v.y, w[q].x + w[q].y, a3);
                }
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    y[u][h][q] = cmake<C, R>(y[u][h][q].x - (a1[q] - a2[q]), y[u][h][q].y - (a3[q] - a1[q] - a2[q]));
            }
        }
        // rows rlo .. rlo+15 are done for this block; the lower half moves up
        // (or, after the block's last diamond, is stored too)
#pragma unroll
        for (int u = 0; u < NT; ++u)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int r = rlo + rm[q];
                if (colok[u] && r < n) Y[u][r] = y[u][0][q];
                if (last) {
                    if (colok[u] && r + 16 < n) Y[u][r + 16] = y[u][1][q];
                } else {
                    y[u][0][q] = y[u][1][q];
                    y[u][1][q] = ynext[u][q];
                }
            }
        if (jn >= 0) put((it + 1) & 1, pv, pt);
        if (jn < 0) break;
        j = jn; s = sn; d0 = d0n; ns = nsn;
        ++it;
    }
}

// Y64[r, j] = Y32[r, j] for rows r0 .. r1 - 1 of every column.
__global__ __launch_bounds__(256) void k_rows_c32_to_c64(int n, int r0, int r1, const float2 *y32, double2 *y64,
                                                         size_t stride)
{
    const int kb = blockIdx.y, rows = r1 - r0;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < rows * n; e += gridDim.x * blockDim.x) {
        const int r = r0 + e % rows, c = e / rows;
        const float2 v = y32[(size_t)kb * stride + (size_t)c * n + r];
        y64[(size_t)kb * stride + (size_t)c * n + r] = make_double2((double)v.x, (double)v.y);
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
void launch_phases(const StructDev *sd, const int *item_s, const double *kpts, int maxpairs,
                   double2 *ph, int nb, hipStream_t st)
{
    const int gx = std::max(1, std::min((maxpairs + 255) / 256, 64));
    hipLaunchKernelGGL(k_phases, dim3(gx, nb), dim3(256), 0, st, sd, item_s, kpts, maxpairs, ph);
}
void launch_spinor_expand(const BuildArgs &a, int maxnv, int nb, hipStream_t st)
{
    const int gx = std::max(1, std::min((maxnv * maxnv + 255) / 256, 64));
    hipLaunchKernelGGL(k_spinor_expand, dim3(gx, nb), dim3(256), 0, st, a, nb);
}
void launch_build_hs(const BuildArgs &a, int maxnat, int nb, hipStream_t st)
{
    const int nblk = maxnat * (maxnat + 1) / 2 + kBuildPadWgs;     // upper blocks + padding
    if (a.dh)
        hipLaunchKernelGGL(k_build_hs<4>, dim3(nblk, (nb + build_items<4>() - 1) / build_items<4>()), dim3(256), 0,
                           st, a, nb);
    else
        hipLaunchKernelGGL(k_build_hs<2>, dim3(nblk, (nb + build_items<2>() - 1) / build_items<2>()), dim3(256), 0,
                           st, a, nb);
}
void launch_potrf_panel(int n, double2 *s, size_t stride, int j0, int *info, int nb,
                        hipStream_t st)
{
    hipLaunchKernelGGL(k_potrf_panel, dim3(nb), dim3(256), 0, st, n, s, stride, j0, info);
}
constexpr int kTileBM[kGemmTiles] = {64, 96, 128, 32};
constexpr int kTileBN[kGemmTiles] = {64, 96, 64, 32};

template <class R, int OPA, int OPB>
static void launch_gemm_ops(const GemmArgsT<R> &g, int tile, int nb, hipStream_t st)
{
    dim3 grid((g.m + kTileBM[tile] - 1) / kTileBM[tile], (g.n + kTileBN[tile] - 1) / kTileBN[tile], nb);
    switch (tile) {
    case 1: hipLaunchKernelGGL((k_gemm<R, OPA, OPB, 96, 96>), grid, dim3(256), 0, st, g); break;
    case 2: hipLaunchKernelGGL((k_gemm<R, OPA, OPB, 128, 64>), grid, dim3(256), 0, st, g); break;
    case 3: hipLaunchKernelGGL((k_gemm<R, OPA, OPB, 32, 32>), grid, dim3(256), 0, st, g); break;
    default: hipLaunchKernelGGL((k_gemm<R, OPA, OPB, 64, 64>), grid, dim3(256), 0, st, g); break;
    }
}
static void store_args(GemmShape &s, const GemmArgsT<double> &g) { s.g64 = g; }
static void store_args(GemmShape &s, const GemmArgsT<float> &g) { s.g32 = g; }
template <class R>
static void launch_gemm_t(const GemmArgsT<R> &g, int opa, int opb, int nb, hipStream_t st)
{
    const int prec = sizeof(R) == 8 ? 0 : 1;
    GemmShape sh{prec, opa, opb, g.m, g.n, g.k, g.tri, nb, {}, {}};
    store_args(sh, g);
    gemm_note(sh);
    const int tile = gemm_tile(prec, opa, opb, g.m, g.n, g.k, g.tri, nb);
    if (opa == 0 && opb == 0) launch_gemm_ops<R, 0, 0>(g, tile, nb, st);
    else if (opa == 1 && opb == 0) launch_gemm_ops<R, 1, 0>(g, tile, nb, st);
    else if (opa == 0 && opb == 1) launch_gemm_ops<R, 0, 1>(g, tile, nb, st);
    else launch_gemm_ops<R, 1, 1>(g, tile, nb, st);
}
void launch_gemm_tile(const GemmShape &s, int tile, hipStream_t st)
{
    if (s.prec == 0) {
        if (s.opa == 0 && s.opb == 0) launch_gemm_ops<double, 0, 0>(s.g64, tile, s.nb, st);
        else if (s.opa == 1 && s.opb == 0) launch_gemm_ops<double, 1, 0>(s.g64, tile, s.nb, st);
        else if (s.opa == 0 && s.opb == 1) launch_gemm_ops<double, 0, 1>(s.g64, tile, s.nb, st);
        else launch_gemm_ops<double, 1, 1>(s.g64, tile, s.nb, st);
    } else {
        if (s.opa == 0 && s.opb == 0) launch_gemm_ops<float, 0, 0>(s.g32, tile, s.nb, st);
        else if (s.opa == 1 && s.opb == 0) launch_gemm_ops<float, 1, 0>(s.g32, tile, s.nb, st);
        else if (s.opa == 0 && s.opb == 1) launch_gemm_ops<float, 0, 1>(s.g32, tile, s.nb, st);
        else launch_gemm_ops<float, 1, 1>(s.g32, tile, s.nb, st);
    }
}
void launch_zgemm(const GemmArgs &g, int opa, int opb, int nb, hipStream_t st)
{
    launch_gemm_t<double>(g, opa, opb, nb, st);
}
void launch_cgemm(const GemmArgsT<float> &g, int opa, int opb, int nb, hipStream_t st)
{
    launch_gemm_t<float>(g, opa, opb, nb, st);
}
template <class R, int MAXQ>
static void launch_hetrd_panel_q(const HetrdArgsT<R> &h, int i0, int nb, hipStream_t st)
{
    const size_t lds = hetrd_lds_bytes<R, MAXQ>(h.n);
    // once per instantiation (thread-safe initialisation): the dynamic LDS
    // limit for the largest order it serves (MAXQ column slots of 64 rows)
    static const bool limit_set = [] {
        const size_t most = hetrd_lds_bytes<R, MAXQ>(std::min(64 * MAXQ, 1280));
        if (most > 65536)
            (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_hetrd_panel<R, MAXQ>),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)most);
        return true;
    }();
    (void)limit_set;
    hipLaunchKernelGGL((k_hetrd_panel<R, MAXQ>), dim3(nb), dim3(HT), lds, st, h, i0);
}
template <class R>
static void launch_hetrd_panel_t(const HetrdArgsT<R> &h, int i0, int nb, hipStream_t st)
{
    if (h.n <= 320) launch_hetrd_panel_q<R, 5>(h, i0, nb, st);
    else if (h.n <= 640) launch_hetrd_panel_q<R, 10>(h, i0, nb, st);
    else launch_hetrd_panel_q<R, 20>(h, i0, nb, st);
}
template <class R>
static void launch_hetrd_finish_t(const HetrdArgsT<R> &h, int nb, hipStream_t st)
{
    hipLaunchKernelGGL((k_hetrd_finish<R>), dim3(nb), dim3(256), 0, st, h);
}
template <class R>
static void launch_hetrd_wide_column_t(const HetrdArgsT<R> &h, const WideArgsT<R> &w, int i0, int i, int nb,
                                       hipStream_t st)
{
    const int m0 = h.n - i0 - 1;                      // trailing order at the panel's first column
    const int T0 = (m0 + 63) / 64, ntile = T0 * (T0 + 1) / 2;
    const int pw = std::min(NB, h.n - 1 - i0);
    if (i == 0) hipLaunchKernelGGL((k_hw_prep<R>), dim3(nb), dim3(WT), 0, st, h, w, i0, i);
    hipLaunchKernelGGL((k_hw_hemv<R>), dim3(ntile + 1, nb), dim3(256), 0, st, h, w, i0, i, ntile);
    hipLaunchKernelGGL((k_hw_finprep<R>), dim3(nb), dim3(WT), 0, st, h, w, i0, i, i + 1 < pw ? 1 : 0);
}
template <class R>
static void launch_hetrd_wide_vt_t(const HetrdArgsT<R> &h, int i0, int nb, hipStream_t st)
{
    const int blocks = std::min((h.n * NB + 255) / 256, 64);
    hipLaunchKernelGGL((k_hw_vt<R>), dim3(blocks, nb), dim3(256), 0, st, h, i0);
}
void launch_hetrd_wide_column(const HetrdArgs &h, const WideArgsT<double> &w, int i0, int i, int nb,
                              hipStream_t st) { launch_hetrd_wide_column_t(h, w, i0, i, nb, st); }
void launch_hetrd_wide_column(const HetrdArgsT<float> &h, const WideArgsT<float> &w, int i0, int i, int nb,
                              hipStream_t st) { launch_hetrd_wide_column_t(h, w, i0, i, nb, st); }
void launch_hetrd_wide_vt(const HetrdArgs &h, int i0, int nb, hipStream_t st) { launch_hetrd_wide_vt_t(h, i0, nb, st); }
void launch_hetrd_wide_vt(const HetrdArgsT<float> &h, int i0, int nb, hipStream_t st) { launch_hetrd_wide_vt_t(h, i0, nb, st); }
// resident tridiagonalisation: tile dimension NTM and tiles per wave MT for
// the order (FP64 up to 160: the panel's reflectors share the LDS)
template <class R, int NTM, int MT>
static void launch_hetrd_res_q(const HetrdArgsT<R> &h, int nb, hipStream_t st)
{
    hipLaunchKernelGGL((k_hetrd_res<R, NTM, MT>), dim3(nb), dim3(RZ_THREADS), 0, st, h);
}
int hetrd_resident_max_n(bool fp32) { return fp32 ? 288 : 160; }
// (callers keep h.n <= hetrd_resident_max_n: the plans choose the resident
// form only there, and tbk_plan_set_tridiag refuses it beyond)
void launch_hetrd_resident(const HetrdArgsT<float> &h, int nb, hipStream_t st)
{
    if (h.n > hetrd_resident_max_n(true)) return;
    if (h.n <= 128) launch_hetrd_res_q<float, 8, 5>(h, nb, st);
    else if (h.n <= 192) launch_hetrd_res_q<float, 12, 10>(h, nb, st);
    else if (h.n <= 256) launch_hetrd_res_q<float, 16, 17>(h, nb, st);
    else launch_hetrd_res_q<float, 18, 22>(h, nb, st);
}
void launch_hetrd_resident(const HetrdArgs &h, int nb, hipStream_t st)
{
    if (h.n > hetrd_resident_max_n(false)) return;
    if (h.n <= 128) launch_hetrd_res_q<double, 8, 5>(h, nb, st);
    else launch_hetrd_res_q<double, 10, 7>(h, nb, st);
}
void launch_hetrd_panel(const HetrdArgs &h, int i0, int nb, hipStream_t st) { launch_hetrd_panel_t(h, i0, nb, st); }
void launch_hetrd_finish(const HetrdArgs &h, int nb, hipStream_t st) { launch_hetrd_finish_t(h, nb, st); }
void launch_hetrd_panel(const HetrdArgsT<float> &h, int i0, int nb, hipStream_t st) { launch_hetrd_panel_t(h, i0, nb, st); }
void launch_hetrd_finish(const HetrdArgsT<float> &h, int nb, hipStream_t st) { launch_hetrd_finish_t(h, nb, st); }
int stein_threads(int n) { return (n + 63) / 64 * 64; }
void launch_stebz(const SteinArgs &s, int nb, hipStream_t st)
{
    hipLaunchKernelGGL(k_stebz, dim3((2 * s.n + 255) / 256, nb), dim3(256), 0, st, s);
}
void launch_stein_vectors(const SteinArgs &s, int nb, hipStream_t st)
{
    hipLaunchKernelGGL(k_stein_vec, dim3((s.n + 63) / 64, nb), dim3(64), 4 * (size_t)s.n * sizeof(double), st, s);
    hipLaunchKernelGGL(k_stein_orth, dim3(nb), dim3(256), 0, st, s);
}
void launch_stein(const SteinArgs &s, int nb, hipStream_t st)
{
    launch_stebz(s, nb, st);
    launch_stein_vectors(s, nb, st);
}
void launch_real_to_complex_t(int N, const double *Vt, double2 *Y, size_t stride, int nb,
                              hipStream_t st, const double2 *ph)
{
    const int t = (N + 31) / 32;
    hipLaunchKernelGGL((k_vt_to_columns<double2>), dim3(t, t, nb), dim3(256), 0, st, N, Vt, Y, stride,
                       (double2 *)nullptr, ph);
}
void launch_real_to_c32_t(int n, const double *vt, float2 *y, size_t stride, int nb, hipStream_t st,
                          double2 *y64_row0, const double2 *ph)
{
    const int t = (n + 31) / 32;
    hipLaunchKernelGGL((k_vt_to_columns<float2>), dim3(t, t, nb), dim3(256), 0, st, n, vt, y, stride, y64_row0,
                       ph);
}
int band_diamonds(int n)
{
    int b = 0;
    for (int i = 0; i <= n - 3; i += TB) b += (n - 2 - i) / TB + 1;
    return b;
}
// The panel rows live in registers, RPT per thread: FP64 keeps RPT at 2
// (a double2 slot row of TB columns is 64 registers) and adds waves with m.
void launch_band_qr(const BandArgsT<double> &g, int panel, int nb, hipStream_t st)
{
    const int c0 = panel * TB, m = g.n - c0 - TB;
    if (m <= 128) hipLaunchKernelGGL((k_band_qr<double, 1, 2>), dim3(nb), dim3(64), 0, st, g, c0, panel);
    else if (m <= 256) hipLaunchKernelGGL((k_band_qr<double, 2, 2>), dim3(nb), dim3(128), 0, st, g, c0, panel);
    else if (m <= 512) hipLaunchKernelGGL((k_band_qr<double, 4, 2>), dim3(nb), dim3(256), 0, st, g, c0, panel);
    else if (m <= 1024) hipLaunchKernelGGL((k_band_qr<double, 8, 2>), dim3(nb), dim3(512), 0, st, g, c0, panel);
    else hipLaunchKernelGGL((k_band_qr<double, 10, 2>), dim3(nb), dim3(640), 0, st, g, c0, panel);
}
void launch_band_qr(const BandArgsT<float> &g, int panel, int nb, hipStream_t st)
{
    const int c0 = panel * TB, m = g.n - c0 - TB;
    if (m <= 128) hipLaunchKernelGGL((k_band_qr<float, 1, 2>), dim3(nb), dim3(64), 0, st, g, c0, panel);
    else if (m <= 320) hipLaunchKernelGGL((k_band_qr<float, 1, 5>), dim3(nb), dim3(64), 0, st, g, c0, panel);
    else hipLaunchKernelGGL((k_band_qr<float, 4, 5>), dim3(nb), dim3(256), 0, st, g, c0, panel);
}
template <class R>
static void launch_band_tail_t(const BandArgsT<R> &g, int c_end, int nb, hipStream_t st)
{
    hipLaunchKernelGGL((k_band_tail<R>), dim3(nb), dim3(256), 0, st, g, c_end);
}
void launch_band_tail(const BandArgsT<double> &g, int c_end, int nb, hipStream_t st) { launch_band_tail_t(g, c_end, nb, st); }
void launch_band_tail(const BandArgsT<float> &g, int c_end, int nb, hipStream_t st) { launch_band_tail_t(g, c_end, nb, st); }
// LDS of one workgroup (bytes) and what k_bulge's static arrays take at most
constexpr size_t kLdsBytes = 160 * 1024, kBulgeStatic = 10 * 1024;
template <class R>
static void launch_bulge_t(const BulgeArgsT<R> &g, int nb, hipStream_t st)
{
    typedef typename cplx<R>::type C;
    const int n = g.n;
    const int nw = std::min(16, (n - 2) / TB / 2 + 2);        // sweeps in flight at most
    const size_t bytes = (size_t)n * TBLD * sizeof(C);
    if (bytes + kBulgeStatic <= kLdsBytes) {
        static const bool limit_set = [] {
            (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_bulge<R, true>),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)(kLdsBytes - kBulgeStatic));
            return true;
        }();
        (void)limit_set;
        hipLaunchKernelGGL((k_bulge<R, true>), dim3(nb), dim3(64 * nw), bytes, st, g);
    } else {
        hipLaunchKernelGGL((k_bulge<R, false>), dim3(nb), dim3(64 * nw), 0, st, g);
    }
}
void launch_bulge(const BulgeArgsT<double> &g, int nb, hipStream_t st) { launch_bulge_t(g, nb, st); }
void launch_bulge(const BulgeArgsT<float> &g, int nb, hipStream_t st) { launch_bulge_t(g, nb, st); }
template <class R>
static void launch_q2_t(const Q2ArgsT<R> &g, int nb, hipStream_t st)
{
    hipLaunchKernelGGL((k_q2_tfactors<R>), dim3(band_diamonds(g.n), nb), dim3(64), 0, st, g);
    constexpr int cols = 16 * kQ2Waves * q2_tiles<R>();
    hipLaunchKernelGGL((k_apply_q2<R>), dim3((g.n + cols - 1) / cols, nb), dim3(64 * kQ2Waves), 0, st, g);
}
void launch_q2(const Q2ArgsT<double> &g, int nb, hipStream_t st) { launch_q2_t(g, nb, st); }
void launch_q2(const Q2ArgsT<float> &g, int nb, hipStream_t st) { launch_q2_t(g, nb, st); }
void launch_rows_c32_to_c64(int n, int r0, int r1, const float2 *y32, double2 *y64, size_t stride, int nb,
                            hipStream_t st)
{
    if (r1 <= r0) return;
    const int gx = std::max(1, std::min(((r1 - r0) * n + 255) / 256, 64));
    hipLaunchKernelGGL(k_rows_c32_to_c64, dim3(gx, nb), dim3(256), 0, st, n, r0, r1, y32, y64, stride);
}
int oa_max_n() { return kOaMax; }
void launch_oa_prepare(const OaArgs &a, int nb, hipStream_t st)
{
    hipLaunchKernelGGL(k_oa_resid, dim3((a.n + 3) / 4, nb), dim3(256), 0, st, a);
    hipLaunchKernelGGL(k_oa_clusters, dim3(nb), dim3(1024), 0, st, a);
    hipLaunchKernelGGL(k_oa_rr, dim3(kOaSlots, nb), dim3(64), 0, st, a);
}
void launch_oa_eigen(const OaArgs &a, int nb, hipStream_t st)
{
    hipLaunchKernelGGL(k_oa_second, dim3((a.n + 15) / 16, nb), dim3(1024), 0, st, a);
    hipLaunchKernelGGL(k_oa_sort, dim3(nb), dim3(256), 0, st, a);
}
void launch_oa_emat(const OaArgs &a, int nb, hipStream_t st, const int *cut)
{
    hipLaunchKernelGGL(k_oa_emat, dim3((unsigned)(((size_t)a.n * a.n + 255) / 256), nb), dim3(256), 0, st, a,
                       cut);
}
void launch_oa_final(const OaArgs &a, int nb, hipStream_t st, const int *cut)
{
    hipLaunchKernelGGL(k_oa_final, dim3((a.n + 3) / 4, nb), dim3(256), 0, st, a, cut);
}
void launch_coldot(int N, const double2 *Z, const double2 *X, const double2 *Y, size_t stride,
                   double2 *px, double2 *py, int nb, hipStream_t st, const int *cut)
{
    hipLaunchKernelGGL(k_coldot, dim3((N + 3) / 4, nb), dim3(256), 0, st, N, Z, X, Y, stride, px, py, cut);
}
void launch_perturb(const PerturbArgs &p, hipStream_t st)
{
    hipLaunchKernelGGL(k_perturb, dim3((p.nb + 63) / 64), dim3(64), 0, st, p);
}
void launch_fermi_window_group(const FermiArgs *f_dev, int nstruct, hipStream_t st)
{
    hipLaunchKernelGGL(k_fermi_window_group, dim3(nstruct), dim3(1024), 0, st, f_dev);
}
void launch_fermi_sums_group(const FermiArgs *f_dev, int nstruct, hipStream_t st, int ef_known)
{
    hipLaunchKernelGGL(k_fermi_sums_group, dim3(nstruct), dim3(1024), 0, st, f_dev, ef_known);
}
void launch_spin_expect(const SpinArgs &a, int nb, hipStream_t st)
{
    hipLaunchKernelGGL(k_spin_expect, dim3((a.n + 3) / 4, nb), dim3(256), 0, st, a);
}
void launch_spin_sum(const FermiArgs *f_dev, const StructDev *sd, int nstruct, const double *spev, int n,
                     int maxnat, double *out, hipStream_t st)
{
    hipLaunchKernelGGL(k_spin_sum, dim3(nstruct), dim3(256), 0, st, f_dev, sd, spev, n, maxnat, out);
}
void launch_copy16(const double2 *src, double2 *dst, size_t n, hipStream_t st)
{
    hipLaunchKernelGGL(k_copy16, dim3(4096), dim3(256), 0, st, src, dst, n);
}
void launch_set_identity(int n, double2 *s, size_t stride, int nb, hipStream_t st)
{
    hipLaunchKernelGGL(k_set_identity, dim3((unsigned)(((size_t)n * n + 255) / 256), nb), dim3(256), 0, st, n, s,
                       stride);
}
void launch_scale_rows_sqrt(int n, double2 *t, size_t stride, const double *lam, int nb, hipStream_t st)
{
    hipLaunchKernelGGL(k_scale_rows_sqrt, dim3((unsigned)(((size_t)n * n + 255) / 256), nb), dim3(256), 0, st, n,
                       t, stride, lam);
}
void launch_abs2(int n, const double2 *y, size_t stride, double *proj, int nb, hipStream_t st)
{
    hipLaunchKernelGGL(k_abs2, dim3((unsigned)(((size_t)n * n + 255) / 256), nb), dim3(256), 0, st, n, y, stride,
                       proj);
}
void launch_copy_int(const int *src, int *dst, int n, hipStream_t st)
{
    hipLaunchKernelGGL(k_copy_int, dim3((n + 255) / 256), dim3(256), 0, st, src, dst, n);
}
void launch_fermi_sums(const FermiArgs &f, hipStream_t st)
{
    hipLaunchKernelGGL(k_fermi_sums, dim3(1), dim3(1024), 0, st, f);
}

}  // namespace tbk
