// tbk_multi.cpp -- many structures of uneven orbital count: structures are
// grouped by padded order (the orbital count rounded up to a multiple of
// `granularity`), each group is one batched plan (tbk_plan_create_group),
// and the groups run concurrently, each on its own stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <vector>

#include "tbk.h"

struct tbk_multi {
    std::vector<tbk_plan *> plans;
    std::vector<int> plan_of, index_in;     // structure -> (plan, index in the plan)
    std::vector<std::vector<int>> members;  // plan -> structures
    std::vector<int> status;
};

// Orbital count of a structure descriptor (sum of kbas over its atoms).
static int orbital_count(const tbk_structure_desc &d)
{
    int nv = 0;
    for (int a = 0; a < d.natoms; ++a) {
        const int k = d.kkind[a];
        if (k < 1 || k > d.kinds) return -1;
        nv += d.kbas[k - 1];
    }
    return nv;
}

extern "C" int tbk_multi_create(int device, int nstruct, const tbk_structure_desc *st, const int *nk,
                                const double *kpts, const double *weights, int precision, int granularity,
                                long long max_batch_bytes, tbk_multi **out)
{
    if (!st || !nk || !kpts || !weights || !out || nstruct <= 0 || granularity <= 0) return TBK_ERR_ARG;
    *out = nullptr;
    std::vector<size_t> k0(nstruct + 1, 0);
    std::map<int, std::vector<int>> groups;            // padded order -> structures
    for (int i = 0; i < nstruct; ++i) {
        if (nk[i] <= 0) return TBK_ERR_ARG;
        k0[i + 1] = k0[i] + nk[i];
        const int nv = orbital_count(st[i]);
        if (nv < 1) return TBK_ERR_ARG;
        groups[(nv + granularity - 1) / granularity * granularity].push_back(i);
    }
    tbk_multi *m = new tbk_multi();
    m->plan_of.assign(nstruct, -1);
    m->index_in.assign(nstruct, -1);
    m->status.assign(nstruct, TBK_OK);
    int rc = TBK_OK;
    for (auto &g : groups) {
        const std::vector<int> &mem = g.second;
        std::vector<tbk_structure_desc> d;
        std::vector<int> gk;
        std::vector<double> gp, gw;
        for (int i : mem) {
            d.push_back(st[i]);
            gk.push_back(nk[i]);
            gp.insert(gp.end(), kpts + 3 * k0[i], kpts + 3 * k0[i + 1]);
            gw.insert(gw.end(), weights + k0[i], weights + k0[i + 1]);
        }
        // items per launch bounded by the workspace budget (about 180 N^2
        // bytes per item at order N)
        const long long per_item = 180LL * g.first * g.first;
        int batch = 0;
        if (max_batch_bytes > 0) batch = (int)std::max<long long>(1, max_batch_bytes / per_item);
        tbk_plan *p = nullptr;
        rc = tbk_plan_create_group(device, (int)mem.size(), d.data(), gk.data(), gp.data(), gw.data(), batch,
                                   precision, &p);
        if (rc) break;
        for (size_t j = 0; j < mem.size(); ++j) { m->plan_of[mem[j]] = (int)m->plans.size(); m->index_in[mem[j]] = (int)j; }
        m->plans.push_back(p);
        m->members.push_back(mem);
    }
    if (rc) {
        for (tbk_plan *p : m->plans) tbk_plan_destroy(p);
        delete m;
        return rc;
    }
    *out = m;
    return TBK_OK;
}

extern "C" int tbk_multi_launch(tbk_multi *m, int mode, double tkb, double eigcut)
{
    if (!m) return TBK_ERR_ARG;
    // largest groups first so the long solves start early
    for (size_t i = m->plans.size(); i-- > 0;) {
        const int rc = tbk_plan_launch(m->plans[i], mode, tkb, eigcut);
        if (rc) return rc;
    }
    return TBK_OK;
}

extern "C" int tbk_multi_wait(tbk_multi *m)
{
    if (!m) return TBK_ERR_ARG;
    int first = TBK_OK;
    for (size_t i = 0; i < m->plans.size(); ++i) {
        const int rc = tbk_plan_wait(m->plans[i], nullptr, nullptr, nullptr);
        if (rc == TBK_ERR_HIP || rc == TBK_ERR_ARG) return rc;
        for (size_t j = 0; j < m->members[i].size(); ++j) {
            const int s = m->members[i][j];
            m->status[s] = tbk_plan_result(m->plans[i], (int)j, nullptr, nullptr, nullptr);
            if (m->status[s] && !first) first = m->status[s];
        }
    }
    return first;
}

extern "C" int tbk_multi_run(tbk_multi *m, int mode, double tkb, double eigcut)
{
    const int rc = tbk_multi_launch(m, mode, tkb, eigcut);
    if (rc) return rc;
    return tbk_multi_wait(m);
}

extern "C" int tbk_multi_result(const tbk_multi *m, int s, double *efermi, double *earg1, double *earg2)
{
    if (!m || s < 0 || s >= (int)m->plan_of.size()) return TBK_ERR_ARG;
    return tbk_plan_result(m->plans[m->plan_of[s]], m->index_in[s], efermi, earg1, earg2);
}

extern "C" int tbk_multi_eigenvalues(tbk_multi *m, int s, double *esk, double *desk)
{
    if (!m || s < 0 || s >= (int)m->plan_of.size()) return TBK_ERR_ARG;
    return tbk_plan_structure_eigenvalues(m->plans[m->plan_of[s]], m->index_in[s], esk, desk);
}

extern "C" int tbk_multi_ngroups(const tbk_multi *m) { return m ? (int)m->plans.size() : -1; }

extern "C" int tbk_multi_group_order(const tbk_multi *m, int g)
{
    if (!m || g < 0 || g >= (int)m->plans.size()) return -1;
    return tbk_plan_order(m->plans[g]);
}

extern "C" int tbk_multi_set_graph(tbk_multi *m, int on)
{
    if (!m) return TBK_ERR_ARG;
    for (tbk_plan *p : m->plans) tbk_plan_set_graph(p, on);
    return TBK_OK;
}

extern "C" int tbk_multi_autotune(tbk_multi *m, int reps)
{
    if (!m) return TBK_ERR_ARG;
    for (tbk_plan *p : m->plans) {
        const int rc = tbk_plan_autotune(p, reps, nullptr);
        if (rc && rc != TBK_ERR_OVERLAP && rc != TBK_ERR_DEGEN) return rc;
    }
    return TBK_OK;
}

extern "C" long long tbk_multi_device_bytes(const tbk_multi *m)
{
    if (!m) return -1;
    long long b = 0;
    for (tbk_plan *p : m->plans) b += tbk_plan_device_bytes(p);
    return b;
}

extern "C" int tbk_multi_destroy(tbk_multi *m)
{
    if (!m) return TBK_ERR_ARG;
    for (tbk_plan *p : m->plans) tbk_plan_destroy(p);
    delete m;
    return TBK_OK;
}
