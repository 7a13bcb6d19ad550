// tbk_tune.cpp -- GEMM tile choice: a per-process table filled by the
// autotuner (tbk_plan_autotune) with a cost model as the fallback.
#include <cmath>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "tbk_kernels.h"

namespace tbk {
namespace {
typedef std::tuple<int, int, int, int, int, int, int, int> Key;   // prec, opa, opb, m, n, k, tri, nb
std::mutex g_mu;
std::map<Key, int> g_table;
int g_forced = -1;
// shapes recorded by the autotuner's pass: per host thread, since the pass
// launches from the calling thread, so that launches other threads make
// meanwhile (other plans, the request loop's workers) are not recorded with it
thread_local bool g_record = false;
thread_local std::vector<GemmShape> g_recorded;

const int kBM[kGemmTiles] = {64, 96, 128, 32};
const int kBN[kGemmTiles] = {64, 96, 64, 32};
// relative time per padded output element of a well-occupied launch
const double kCost[kGemmTiles] = {1.0, 0.86, 0.92, 1.7};
}  // namespace

int gemm_tile_model(int m, int n, int nb)
{
    int best = 0;
    double bc = 1e300;
    for (int t = 0; t < kGemmTiles; ++t) {
        const double tm = std::ceil((double)m / kBM[t]), tn = std::ceil((double)n / kBN[t]);
        const double wgs = tm * tn * nb;
        double c = tm * kBM[t] * tn * kBN[t] * kCost[t];
        if (wgs < 1024) c *= 1024.0 / wgs;        // fewer than 4 workgroups per CU
        if (c < bc) { bc = c; best = t; }
    }
    return best;
}

int gemm_tile(int prec, int opa, int opb, int m, int n, int k, int tri, int nb)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_forced >= 0) return g_forced;
    auto it = g_table.find(Key(prec, opa, opb, m, n, k, tri, nb));
    if (it != g_table.end()) return it->second;
    return gemm_tile_model(m, n, nb);
}

void gemm_force_tile(int tile)
{
    std::lock_guard<std::mutex> lk(g_mu);
    g_forced = tile;
}

void gemm_store_tile(const GemmShape &s, int tile)
{
    std::lock_guard<std::mutex> lk(g_mu);
    g_table[Key(s.prec, s.opa, s.opb, s.m, s.n, s.k, s.tri, s.nb)] = tile;
}

bool gemm_tuned(const GemmShape &s)
{
    std::lock_guard<std::mutex> lk(g_mu);
    return g_table.count(Key(s.prec, s.opa, s.opb, s.m, s.n, s.k, s.tri, s.nb)) != 0;
}

void gemm_record(bool on)
{
    g_record = on;
    if (on) g_recorded.clear();
}

void gemm_note(const GemmShape &s)
{
    if (!g_record) return;
    for (const GemmShape &r : g_recorded)
        if (r.prec == s.prec && r.opa == s.opa && r.opb == s.opb && r.m == s.m && r.n == s.n &&
            r.k == s.k && r.tri == s.tri && r.nb == s.nb)
            return;
    g_recorded.push_back(s);
}

std::vector<GemmShape> gemm_recorded() { return g_recorded; }

}  // namespace tbk
