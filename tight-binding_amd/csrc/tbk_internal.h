// tbk_internal.h -- host-side types shared by the set-up, the device
// driver and the C ABI.
#pragma once

#include <algorithm>
#include <atomic>
#include <thread>
#include <vector>

namespace tbk {

// f(0) .. f(n-1) on up to 16 host threads (each index once; f writes only
// its own outputs).  A call made from inside another parallel_for runs
// serially, so nested set-up work does not multiply the thread count.
inline thread_local bool t_in_parallel = false;
template <class F>
void parallel_for(int n, F f)
{
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const int nt = t_in_parallel ? 1 : (int)std::min<unsigned>({(unsigned)std::max(n, 0), hw, 16u});
    if (nt <= 1) { for (int i = 0; i < n; ++i) f(i); return; }
    std::atomic<int> next{0};
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t)
        th.emplace_back([&]() {
            t_in_parallel = true;
            for (int i; (i = next.fetch_add(1)) < n;) f(i);
        });
    for (std::thread &x : th) x.join();
}

// One entry of the neighbour list (COMMON /pairs/, /neigh/, search2.f:72-81).
struct Pair {
    double cosines[3];   // tt_list: direction cosines of R_j - R_i
    double dlv[3];       // dlv_list: lattice translation of atom j
    double dist;         // dist_list
    double screen;       // screen_list: smooth cutoff F(R)
    double dscrn;        // dF/dR = dscrn * F
    int kind[2];         // jkind_list (1-based)
    int atom[2];         // jatm_list (0-based here)
};

struct PairList {
    std::vector<Pair> pairs;
    int n = 0;
    double dismin = 0.0;
    void clear() { pairs.clear(); n = 0; dismin = 0.0; }
    void push(const Pair &p) { pairs.push_back(p); n = (int)pairs.size(); }
};

int search2(int natoms, const double *lat, const double *posn, const int *kkind,
            int kinds, const int *nd, const double *dnn, const double *screenl,
            int maxpairs, PairList &out);
// param: the spin channel being set up; atoms with flip[a] != 0 take their
// density decay and on-site terms from param_flip (the other channel).
int setpar(int natoms, const int *kkind, int kinds, const double *param, int nltype,
           int jsover, const PairList &pl, double *den, double *paros, double *hop,
           double *over, const double *param_flip = nullptr, const int *flip = nullptr);
int setparv(int natoms, const int *kkind, int kinds, const double *param, int nltype,
            int jsover, double thisvol, const PairList &pl, const double *den,
            double *dparos, double *dhop, double *dover, const double *param_flip = nullptr,
            const int *flip = nullptr);
// Non-collinear (lnoncol) on-site terms: den2 [2][natoms*kinds], paros and
// spinos [4*natoms]; their volume derivatives; the per-orbital spinor terms.
void setpar_noncol(int natoms, const int *kkind, int kinds, const double *param_up, const double *param_dn,
                   const PairList &pl, double *den2, double *paros, double *spinos);
void setparv_noncol(int natoms, const int *kkind, int kinds, const double *param_up, const double *param_dn,
                    double thisvol, const PairList &pl, const double *den2, double *dparos, double *dspnos);
void spinor_terms(int natoms, const int *nbas, const double *spinos, const double *spinat, double *d,
                  double *mre, double *mim);
void sk_block(double l, double m, double n, const double *p, bool unlike, double *e);
void arrange_integrals(const double *stored, int k1, int k2, double *out);

}  // namespace tbk
