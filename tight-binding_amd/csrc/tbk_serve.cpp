// tbk_serve.cpp -- request loop with continuous batching.
//
// Requests (one structure + k-mesh + mode/kT/eigcut each) are queued by
// tbk_server_submit.  A worker thread admits pending requests whenever the
// device memory they need fits in the budget and an in-flight slot is free;
// the requests admitted together are grouped by padded order (and by mode,
// kT, eigcut) into group plans, each launched on its own stream, so their
// work overlaps on the device with the groups already running.  Finished
// groups are detected with stream queries, their results stored and their
// memory released, which admits the next requests.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <tuple>
#include <vector>

#include "tbk.h"

namespace {

typedef std::chrono::steady_clock Clock;

// Deep copy of a structure descriptor (the caller's arrays may go away).
struct OwnedStructure {
    tbk_structure_desc d{};
    std::vector<double> lat, posn, posq, valence, dnn, screenl, param;
    std::vector<int> kkind, kbas;
    int nv = 0;
    OwnedStructure(const tbk_structure_desc &s)
    {
        const int nat = s.natoms, kinds = s.kinds, npair = kinds * (kinds + 1) / 2;
        lat.assign(s.lat, s.lat + 9);
        posn.assign(s.posn, s.posn + 3 * nat);
        posq.assign(s.posq, s.posq + 3 * nat);
        kkind.assign(s.kkind, s.kkind + nat);
        kbas.assign(s.kbas, s.kbas + kinds);
        valence.assign(s.valence, s.valence + kinds);
        dnn.assign(s.dnn, s.dnn + npair);
        screenl.assign(s.screenl, s.screenl + npair);
        param.assign(s.param, s.param + s.nparam);
        d = s;
        d.lat = lat.data(); d.posn = posn.data(); d.posq = posq.data(); d.kkind = kkind.data();
        d.kbas = kbas.data(); d.valence = valence.data(); d.dnn = dnn.data(); d.screenl = screenl.data();
        d.param = param.data();
        for (int a = 0; a < nat; ++a) {
            const int k = kkind[a];
            nv += (k >= 1 && k <= kinds) ? kbas[k - 1] : 0;
        }
    }
};

struct Request {
    long long id = 0;
    std::unique_ptr<OwnedStructure> st;
    std::vector<double> kpts, weights;
    int mode = 4;
    double tkb = 0.005, eigcut = 10.0;
    long long bytes = 0;
    Clock::time_point t_submit, t_admit, t_done;
    int status = -1;              // -1 pending / running
    double res[3] = {0, 0, 0};
};

struct Group {
    tbk_plan *plan = nullptr;
    std::vector<long long> ids;
    long long bytes = 0;
};

}  // namespace

struct tbk_server {
    int device = 0, precision = TBK_MIXED, granularity = 32, max_inflight = 8;
    long long budget = 0, inflight_bytes = 0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<long long> pending;
    std::map<long long, std::unique_ptr<Request>> reqs;
    std::vector<Group> running;
    long long next_id = 1;
    bool stop = false;
    std::thread worker;
    // statistics
    long long completed = 0, kpoints_done = 0, groups_launched = 0, max_group = 0;
    int last_error = TBK_OK;
};

// Device bytes of a request solved at padded order N (set-up arrays,
// per-k workspace and results; mixed precision adds the FP32 copies).
static long long estimate_bytes(int N, int nk, int natoms, int precision)
{
    const long long per_item = (precision == TBK_MIXED ? 180LL : 148LL) * N * N + 64LL * N + 4096;
    const long long setup = 4000LL * natoms * 2 * 81 * 8 * 2;   // SK blocks (upper bound on pairs)
    return per_item * nk + setup + (16LL << 20);
}

static void launch_admitted(tbk_server *s, std::vector<Request *> &adm)
{
    // group by (padded order, mode, kT, eigcut)
    std::map<std::tuple<int, int, double, double>, std::vector<Request *>> g;
    for (Request *r : adm) {
        const int N = (r->st->nv + s->granularity - 1) / s->granularity * s->granularity;
        g[std::make_tuple(N, r->mode, r->tkb, r->eigcut)].push_back(r);
    }
    for (auto &kv : g) {
        std::vector<Request *> &rs = kv.second;
        std::vector<tbk_structure_desc> d;
        std::vector<int> nk;
        std::vector<double> kp, w;
        for (Request *r : rs) {
            d.push_back(r->st->d);
            nk.push_back((int)r->weights.size());
            kp.insert(kp.end(), r->kpts.begin(), r->kpts.end());
            w.insert(w.end(), r->weights.begin(), r->weights.end());
        }
        Group grp;
        int rc = tbk_plan_create_group(s->device, (int)rs.size(), d.data(), nk.data(), kp.data(), w.data(), 0,
                                       s->precision, &grp.plan);
        if (!rc) rc = tbk_plan_launch(grp.plan, std::get<1>(kv.first), std::get<2>(kv.first),
                                      std::get<3>(kv.first));
        std::lock_guard<std::mutex> lk(s->mu);
        if (rc) {
            if (grp.plan) tbk_plan_destroy(grp.plan);
            for (Request *r : rs) {
                r->status = rc; r->t_done = Clock::now();
                s->inflight_bytes -= r->bytes;
            }
            s->last_error = rc;
            s->cv_done.notify_all();
            continue;
        }
        for (Request *r : rs) { grp.ids.push_back(r->id); grp.bytes += r->bytes; }
        s->groups_launched++;
        s->max_group = std::max<long long>(s->max_group, (long long)rs.size());
        s->running.push_back(grp);
    }
}

static void worker_loop(tbk_server *s)
{
    (void)hipSetDevice(s->device);
    for (;;) {
        std::vector<Request *> adm;
        std::vector<Group> finished;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            if (s->pending.empty() && s->running.empty())
                s->cv_work.wait(lk, [&] { return s->stop || !s->pending.empty(); });
            if (s->stop && s->pending.empty() && s->running.empty()) return;
            // finished groups
            for (size_t i = 0; i < s->running.size();) {
                if (tbk_plan_query(s->running[i].plan) != 1) {     // finished (or failed)
                    finished.push_back(s->running[i]);
                    s->running.erase(s->running.begin() + i);
                } else {
                    ++i;
                }
            }
        }
        for (Group &g : finished) {
            tbk_plan_wait(g.plan, nullptr, nullptr, nullptr);
            std::vector<std::tuple<int, double, double, double>> out;
            for (size_t j = 0; j < g.ids.size(); ++j) {
                double ef = 0, e1 = 0, e2 = 0;
                const int rc = tbk_plan_result(g.plan, (int)j, &ef, &e1, &e2);
                out.emplace_back(rc, ef, e1, e2);
            }
            tbk_plan_destroy(g.plan);
            std::lock_guard<std::mutex> lk(s->mu);
            const auto now = Clock::now();
            for (size_t j = 0; j < g.ids.size(); ++j) {
                Request *r = s->reqs[g.ids[j]].get();
                r->status = std::get<0>(out[j]);
                r->res[0] = std::get<1>(out[j]); r->res[1] = std::get<2>(out[j]); r->res[2] = std::get<3>(out[j]);
                r->t_done = now;
                s->completed++;
                s->kpoints_done += (long long)r->weights.size();
            }
            s->inflight_bytes -= g.bytes;
            s->cv_done.notify_all();
        }
        {
            // admission: first fit over the queue in arrival order, by memory
            // footprint and in-flight slots
            std::lock_guard<std::mutex> lk(s->mu);
            const bool slot_free = (int)s->running.size() < s->max_inflight;
            for (auto it = s->pending.begin(); slot_free && it != s->pending.end();) {
                Request *r = s->reqs[*it].get();
                if (s->inflight_bytes + r->bytes <= s->budget || (s->running.empty() && adm.empty())) {
                    s->inflight_bytes += r->bytes;
                    r->t_admit = Clock::now();
                    adm.push_back(r);
                    it = s->pending.erase(it);
                } else {
                    ++it;
                }
            }
        }
        if (!adm.empty()) launch_admitted(s, adm);
        if (finished.empty() && adm.empty()) std::this_thread::sleep_for(std::chrono::microseconds(200));
    }
}

extern "C" int tbk_server_create(int device, int precision, long long budget_bytes, int max_inflight,
                                 int granularity, tbk_server **out)
{
    if (!out || budget_bytes <= 0 || max_inflight <= 0 || granularity <= 0) return TBK_ERR_ARG;
    if (precision != TBK_FP64 && precision != TBK_MIXED) return TBK_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return TBK_ERR_NODEVICE;
    if (device < 0 || device >= ndev) return TBK_ERR_ARG;
    tbk_server *s = new tbk_server();
    s->device = device; s->precision = precision; s->budget = budget_bytes;
    s->max_inflight = max_inflight; s->granularity = granularity;
    s->worker = std::thread(worker_loop, s);
    *out = s;
    return TBK_OK;
}

extern "C" int tbk_server_submit(tbk_server *s, const tbk_structure_desc *st, int nk, const double *kpts,
                                 const double *weights, int mode, double tkb, double eigcut, long long *id)
{
    if (!s || !st || nk <= 0 || !kpts || !weights || !id || mode < 1 || mode > 4) return TBK_ERR_ARG;
    if (st->natoms <= 0 || !st->kkind || !st->kbas) return TBK_ERR_ARG;
    std::unique_ptr<Request> r(new Request());
    r->st.reset(new OwnedStructure(*st));
    if (r->st->nv < 2 || r->st->nv > 1280) return TBK_ERR_ARG;
    r->kpts.assign(kpts, kpts + 3 * (size_t)nk);
    r->weights.assign(weights, weights + nk);
    r->mode = mode; r->tkb = tkb; r->eigcut = eigcut;
    const int N = (r->st->nv + s->granularity - 1) / s->granularity * s->granularity;
    r->bytes = estimate_bytes(N, nk, st->natoms, s->precision);
    r->t_submit = Clock::now();
    std::lock_guard<std::mutex> lk(s->mu);
    r->id = s->next_id++;
    *id = r->id;
    s->pending.push_back(r->id);
    s->reqs[r->id] = std::move(r);
    s->cv_work.notify_one();
    return TBK_OK;
}

// Result of a request: returns TBK_OK when done (status in *status), 1 when
// still pending or running (timeout_ms elapsed; < 0 waits without limit).
extern "C" int tbk_server_result(tbk_server *s, long long id, int timeout_ms, int *status, double *efermi,
                                 double *earg1, double *earg2, double *latency_ms, double *queue_ms)
{
    if (!s) return TBK_ERR_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return TBK_ERR_ARG;
    Request *r = it->second.get();
    auto ready = [&] { return r->status != -1; };
    if (timeout_ms < 0) s->cv_done.wait(lk, ready);
    else if (!s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready)) return 1;
    if (status) *status = r->status;
    if (efermi) *efermi = r->res[0];
    if (earg1) *earg1 = r->res[1];
    if (earg2) *earg2 = r->res[2];
    if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
    if (queue_ms) *queue_ms = std::chrono::duration<double, std::milli>(r->t_admit - r->t_submit).count();
    return TBK_OK;
}

// Drop a finished request's record.
extern "C" int tbk_server_release(tbk_server *s, long long id)
{
    if (!s) return TBK_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end() || it->second->status == -1) return TBK_ERR_ARG;
    s->reqs.erase(it);
    return TBK_OK;
}

// stats: completed, k-points done, groups launched, largest group,
// pending, running groups, in-flight bytes, last error.
extern "C" int tbk_server_stats(tbk_server *s, long long *stats8)
{
    if (!s || !stats8) return TBK_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    stats8[0] = s->completed; stats8[1] = s->kpoints_done; stats8[2] = s->groups_launched;
    stats8[3] = s->max_group; stats8[4] = (long long)s->pending.size(); stats8[5] = (long long)s->running.size();
    stats8[6] = s->inflight_bytes; stats8[7] = s->last_error;
    return TBK_OK;
}

extern "C" int tbk_server_destroy(tbk_server *s)
{
    if (!s) return TBK_ERR_ARG;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        s->cv_work.notify_all();
    }
    s->worker.join();
    delete s;
    return TBK_OK;
}
