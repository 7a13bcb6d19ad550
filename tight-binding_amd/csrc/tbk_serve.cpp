// tbk_serve.cpp -- request loop with continuous batching.
//
// tbk_server_submit queues a request (one structure + k-mesh + mode/kT/eigcut
// each) and returns.  A pool of host threads builds each request's set-up
// (neighbour list, bond integrals, SK blocks) as soon as it arrives.  An
// admission thread hands prepared requests to in-flight slots whenever a slot
// is free and the device memory they need fits in the budget: the requests
// admitted together are grouped by padded order (and by mode, kT, eigcut),
// one group per slot.  Each slot has a stream and a thread of its own, which
// creates the group's plan (uploading the prepared set-up), launches it,
// sleeps on its completion event, stores the results and frees the slot, so
// groups are created, run and collected concurrently.  Freed device memory
// goes back to the device's stream-ordered pool, which the server keeps
// rather than returning it to the system.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <tuple>
#include <vector>

#include "tbk.h"
#include "tbk_internal.h"
#include "tbk_prepared.h"

namespace {

typedef std::chrono::steady_clock Clock;

double ms_between(Clock::time_point a, Clock::time_point b)
{
    return std::chrono::duration<double, std::milli>(b - a).count();
}

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
    tbk_prepared *prep = nullptr;   // host set-up, until its group is launched
    int N = 0;                      // padded order
    long long bytes = 0;
    Clock::time_point t_submit, t_admit, t_done;
    int status = -1;                // -1 pending / running
    double res[3] = {0, 0, 0};
    ~Request() { tbk_prepared_free(prep); }
};

// An in-flight slot: a stream, the events bracketing its group's work on
// the device, and the thread that runs its groups.
struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;   // ev1: blocking-sync event
    bool busy = false;                         // a group is assigned or running
    std::vector<Request *> job;                // the assigned group, until taken
    std::condition_variable cv;
    std::thread th;
};

struct Group {
    tbk_plan *plan = nullptr;
    std::vector<long long> ids;
    long long bytes = 0;
    double work = 0.0;              // sum of nv^3 nk
};

// The device's default stream-ordered pool keeps up to this many freed
// bytes while servers run (restored when the last one ends).
std::mutex g_pool_mu;
int g_pool_users = 0;
uint64_t g_pool_old_threshold = 0;

}  // namespace

struct tbk_server {
    int device = 0, precision = TBK_MIXED, granularity = 32, max_inflight = 8;
    long long budget = 0, inflight_bytes = 0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done, cv_prep;
    std::deque<long long> prep_q;   // submitted, set-up not built yet
    std::deque<long long> ready;    // prepared, waiting for admission
    int preparing = 0;
    std::map<long long, std::unique_ptr<Request>> reqs;
    std::vector<std::unique_ptr<Slot>> slots;
    long long events = 0;           // bumped whenever admission may have become possible
    hipEvent_t ref = nullptr;       // time origin of the device intervals
    std::vector<std::pair<double, double>> busy;   // [start, end] ms of each group
    long long next_id = 1;
    bool stop = false, stop_slots = false;
    std::vector<std::thread> preps;
    std::thread worker;
    // statistics
    long long completed = 0, kpoints_done = 0, groups_launched = 0, max_group = 0;
    int last_error = TBK_OK;
    double work_done = 0.0, prep_ms = 0.0, launch_ms = 0.0;
    bool any_submit = false;
    Clock::time_point t_first, t_last;
};

// Device bytes of a request solved at padded order N: per-k workspace and
// results (mixed precision adds the FP32 copies and the refinement's
// scratch), its set-up arrays, slack.
static long long estimate_bytes(int N, int nk, size_t setup_bytes, int precision)
{
    const long long per_item = (precision == TBK_MIXED ? 180LL * N * N + 384LL * N : 148LL * N * N + 64LL * N) + 4096;
    return per_item * nk + (long long)setup_bytes + (16LL << 20);
}

static void prep_loop(tbk_server *s)
{
    // the set-up of one structure runs on this thread alone: the pool
    // prepares several requests at once instead
    tbk::t_in_parallel = true;
    for (;;) {
        Request *r = nullptr;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_prep.wait(lk, [&] { return s->stop || !s->prep_q.empty(); });
            if (s->prep_q.empty()) return;          // stopping, nothing left to prepare
            r = s->reqs[s->prep_q.front()].get();
            s->prep_q.pop_front();
            s->preparing++;
        }
        const Clock::time_point t0 = Clock::now();
        tbk_prepared *pp = nullptr;
        const int rc = tbk_prepare(&r->st->d, &pp);
        const Clock::time_point t1 = Clock::now();
        std::lock_guard<std::mutex> lk(s->mu);
        s->preparing--;
        s->prep_ms += ms_between(t0, t1);
        if (rc) {
            r->status = rc;
            r->t_admit = r->t_done = t1;
            s->last_error = rc;
            s->t_last = t1;
            s->cv_done.notify_all();
        } else {
            r->prep = pp;
            r->N = (tbk_prepared_nv(pp) + s->granularity - 1) / s->granularity * s->granularity;
            r->bytes = estimate_bytes(r->N, (int)r->weights.size(), tbk_prepared_bytes(pp), s->precision);
            s->ready.push_back(r->id);
        }
        s->events++;
        s->cv_work.notify_one();
    }
}

// Creates and launches the plan of one group of admitted requests on its
// slot's stream.  Returns the plan's status; the prepared set-up is freed
// either way (the plan holds its own copy).
static int launch_group(tbk_server *s, Slot &sl, std::vector<Request *> &rs, Group &grp)
{
    const Clock::time_point t0 = Clock::now();
    std::vector<const tbk_prepared *> ps;
    std::vector<int> nk;
    std::vector<double> kp, w;
    for (Request *r : rs) {
        ps.push_back(r->prep);
        nk.push_back((int)r->weights.size());
        kp.insert(kp.end(), r->kpts.begin(), r->kpts.end());
        w.insert(w.end(), r->weights.begin(), r->weights.end());
        const double nv = tbk_prepared_nv(r->prep);
        grp.work += nv * nv * nv * (double)r->weights.size();
        grp.ids.push_back(r->id);
        grp.bytes += r->bytes;
    }
    const Request &r0 = *rs[0];
    int rc = tbk_plan_create_prepared(s->device, (int)rs.size(), ps.data(), nk.data(), kp.data(), w.data(),
                                      s->precision, sl.stream, r0.N, &grp.plan);
    if (!rc && hipEventRecord(sl.ev0, sl.stream) != hipSuccess) rc = TBK_ERR_HIP;
    if (!rc) rc = tbk_plan_launch(grp.plan, r0.mode, r0.tkb, r0.eigcut);
    if (!rc && hipEventRecord(sl.ev1, sl.stream) != hipSuccess) rc = TBK_ERR_HIP;
    if (rc && grp.plan) {
        tbk_plan_destroy(grp.plan);
        grp.plan = nullptr;
    }
    const Clock::time_point t1 = Clock::now();
    std::lock_guard<std::mutex> lk(s->mu);
    s->launch_ms += ms_between(t0, t1);
    for (Request *r : rs) {
        tbk_prepared_free(r->prep);
        r->prep = nullptr;
    }
    if (!rc) {
        s->groups_launched++;
        s->max_group = std::max<long long>(s->max_group, (long long)rs.size());
    }
    return rc;
}

// Waits for a launched group, then stores its results in its requests.
static void finish_group(tbk_server *s, Slot &sl, Group &g)
{
    (void)hipEventSynchronize(sl.ev1);              // sleeps (blocking-sync event)
    float t0 = 0.0f, t1 = 0.0f;
    const bool timed = hipEventElapsedTime(&t0, s->ref, sl.ev0) == hipSuccess &&
                       hipEventElapsedTime(&t1, s->ref, sl.ev1) == hipSuccess;
    tbk_plan_wait(g.plan, nullptr, nullptr, nullptr);
    std::vector<std::tuple<int, double, double, double>> out;
    for (size_t j = 0; j < g.ids.size(); ++j) {
        double ef = 0, e1 = 0, e2 = 0;
        const int rc = tbk_plan_result(g.plan, (int)j, &ef, &e1, &e2);
        out.emplace_back(rc, ef, e1, e2);
    }
    tbk_plan_destroy(g.plan);
    g.plan = nullptr;
    std::lock_guard<std::mutex> lk(s->mu);
    const Clock::time_point now = Clock::now();
    for (size_t j = 0; j < g.ids.size(); ++j) {
        Request *r = s->reqs[g.ids[j]].get();
        r->status = std::get<0>(out[j]);
        r->res[0] = std::get<1>(out[j]); r->res[1] = std::get<2>(out[j]); r->res[2] = std::get<3>(out[j]);
        r->t_done = now;
        s->completed++;
        s->kpoints_done += (long long)r->weights.size();
    }
    if (timed) s->busy.emplace_back((double)t0, (double)t1);
    s->work_done += g.work;
    s->t_last = now;
}

static void slot_loop(tbk_server *s, Slot *sl)
{
    (void)hipSetDevice(s->device);
    for (;;) {
        std::vector<Request *> rs;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            sl->cv.wait(lk, [&] { return !sl->job.empty() || s->stop_slots; });
            if (sl->job.empty()) return;
            rs.swap(sl->job);
        }
        Group g;
        const int rc = launch_group(s, *sl, rs, g);
        if (!rc) {
            finish_group(s, *sl, g);
        } else {
            std::lock_guard<std::mutex> lk(s->mu);
            const Clock::time_point now = Clock::now();
            for (Request *r : rs) {
                r->status = rc;
                r->t_done = now;
            }
            s->last_error = rc;
            s->t_last = now;
        }
        std::lock_guard<std::mutex> lk(s->mu);
        s->inflight_bytes -= g.bytes;
        sl->busy = false;
        s->events++;
        s->cv_work.notify_one();
        s->cv_done.notify_all();
    }
}

// Admission: whenever a request has been prepared or a slot freed, the
// prepared requests in arrival order join this round's group of their
// (order, mode, kT, eigcut) or open one on a free slot; the first one that
// does not fit in the memory budget ends the round (so large requests are
// not passed over indefinitely).
static void worker_loop(tbk_server *s)
{
    long long seen = -1;
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        s->cv_work.wait(lk, [&] { return s->events != seen || s->stop; });
        seen = s->events;
        std::vector<Slot *> free_slots;
        int nbusy = 0;
        for (auto &sl : s->slots) {
            if (sl->busy) nbusy++;
            else free_slots.push_back(sl.get());
        }
        if (s->stop && s->prep_q.empty() && s->preparing == 0 && s->ready.empty() && nbusy == 0) return;
        std::vector<std::tuple<int, int, double, double>> keys;
        std::vector<std::vector<Request *>> groups;
        for (auto it = s->ready.begin(); it != s->ready.end();) {
            Request *r = s->reqs[*it].get();
            const auto key = std::make_tuple(r->N, r->mode, r->tkb, r->eigcut);
            const size_t g = std::find(keys.begin(), keys.end(), key) - keys.begin();
            if (g == keys.size() && keys.size() >= free_slots.size()) { ++it; continue; }
            const bool fits = s->inflight_bytes + r->bytes <= s->budget || (nbusy == 0 && groups.empty());
            if (!fits) break;
            s->inflight_bytes += r->bytes;
            r->t_admit = Clock::now();
            if (g == keys.size()) {
                keys.push_back(key);
                groups.emplace_back();
            }
            groups[g].push_back(r);
            it = s->ready.erase(it);
        }
        for (size_t g = 0; g < groups.size(); ++g) {
            Slot *sl = free_slots[g];
            sl->busy = true;
            sl->job.swap(groups[g]);
            sl->cv.notify_one();
        }
        if (s->stop) {
            // while stopping, wake again shortly even without events (the
            // last set-up may end between the check above and the wait)
            s->cv_work.wait_for(lk, std::chrono::milliseconds(2));
            seen = -1;
        }
    }
}

static void server_free(tbk_server *s)
{
    (void)hipSetDevice(s->device);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop_slots = true;
        for (auto &sl : s->slots) sl->cv.notify_all();
    }
    for (auto &sl : s->slots)
        if (sl->th.joinable()) sl->th.join();
    for (auto &sl : s->slots) {
        if (sl->stream) { (void)hipStreamSynchronize(sl->stream); (void)hipStreamDestroy(sl->stream); }
        if (sl->ev0) (void)hipEventDestroy(sl->ev0);
        if (sl->ev1) (void)hipEventDestroy(sl->ev1);
    }
    if (s->ref) (void)hipEventDestroy(s->ref);
    std::lock_guard<std::mutex> lk(g_pool_mu);
    if (--g_pool_users == 0) {
        hipMemPool_t pool = nullptr;
        if (hipDeviceGetDefaultMemPool(&pool, s->device) == hipSuccess) {
            uint64_t thr = g_pool_old_threshold;
            (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr);
            (void)hipMemPoolTrimTo(pool, (size_t)thr);
        }
    }
    delete s;
}

extern "C" int tbk_server_create(int device, int precision, long long budget_bytes, int max_inflight,
                                 int granularity, tbk_server **out)
{
    if (!out || budget_bytes <= 0 || max_inflight <= 0 || max_inflight > 64 || granularity <= 0)
        return TBK_ERR_ARG;
    if (precision != TBK_FP64 && precision != TBK_MIXED) return TBK_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return TBK_ERR_NODEVICE;
    if (device < 0 || device >= ndev) return TBK_ERR_ARG;
    if (hipSetDevice(device) != hipSuccess) return TBK_ERR_HIP;
    tbk_server *s = new tbk_server();
    s->device = device; s->precision = precision; s->budget = budget_bytes;
    s->max_inflight = max_inflight; s->granularity = granularity;
    {
        std::lock_guard<std::mutex> lk(g_pool_mu);
        hipMemPool_t pool = nullptr;
        if (hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess) {
            if (g_pool_users == 0)
                (void)hipMemPoolGetAttribute(pool, hipMemPoolAttrReleaseThreshold, &g_pool_old_threshold);
            uint64_t thr = (uint64_t)budget_bytes, cur = 0;
            if (hipMemPoolGetAttribute(pool, hipMemPoolAttrReleaseThreshold, &cur) == hipSuccess && cur < thr)
                (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr);
        }
        g_pool_users++;
    }
    bool ok = hipEventCreate(&s->ref) == hipSuccess;
    for (int i = 0; i < max_inflight; ++i) {
        s->slots.emplace_back(new Slot());
        Slot &sl = *s->slots.back();
        ok = ok && hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) == hipSuccess &&
             hipEventCreate(&sl.ev0) == hipSuccess &&
             hipEventCreateWithFlags(&sl.ev1, hipEventBlockingSync) == hipSuccess;
    }
    ok = ok && hipEventRecord(s->ref, s->slots[0]->stream) == hipSuccess &&
         hipEventSynchronize(s->ref) == hipSuccess;
    if (!ok) { server_free(s); return TBK_ERR_HIP; }
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const int nprep = (int)std::min<unsigned>(8u, std::max(2u, hw > 2 ? hw - 2 : 1u));
    for (int i = 0; i < nprep; ++i) s->preps.emplace_back(prep_loop, s);
    for (auto &sl : s->slots) sl->th = std::thread(slot_loop, s, sl.get());
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
    if (r->st->nv < 1 || r->st->nv > 1280) return TBK_ERR_ARG;
    r->kpts.assign(kpts, kpts + 3 * (size_t)nk);
    r->weights.assign(weights, weights + nk);
    r->mode = mode; r->tkb = tkb; r->eigcut = eigcut;
    r->t_submit = Clock::now();
    std::lock_guard<std::mutex> lk(s->mu);
    if (!s->any_submit) { s->any_submit = true; s->t_first = r->t_submit; }
    r->id = s->next_id++;
    *id = r->id;
    s->prep_q.push_back(r->id);
    s->reqs[r->id] = std::move(r);
    s->cv_prep.notify_one();
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
    if (latency_ms) *latency_ms = ms_between(r->t_submit, r->t_done);
    if (queue_ms) *queue_ms = ms_between(r->t_submit, r->t_admit);
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
// pending (waiting for set-up or admission), running groups, in-flight
// bytes, last error.
extern "C" int tbk_server_stats(tbk_server *s, long long *stats8)
{
    if (!s || !stats8) return TBK_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    stats8[0] = s->completed; stats8[1] = s->kpoints_done; stats8[2] = s->groups_launched;
    stats8[3] = s->max_group;
    stats8[4] = (long long)(s->prep_q.size() + s->ready.size()) + s->preparing;
    long long nbusy = 0;
    for (const auto &sl : s->slots) nbusy += sl->busy ? 1 : 0;
    stats8[5] = nbusy;
    stats8[6] = s->inflight_bytes; stats8[7] = s->last_error;
    return TBK_OK;
}

// timing: device busy ms (the union over finished groups of the intervals
// from launch to completion on the device), work done (sum of nv^3 nk over
// completed requests), ms from the first submission to the last completion,
// host set-up ms (summed over requests), plan creation + launch ms on the
// worker (summed over groups), set-up threads.
extern "C" int tbk_server_timing(tbk_server *s, double *t6)
{
    if (!s || !t6) return TBK_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    std::vector<std::pair<double, double>> iv = s->busy;
    std::sort(iv.begin(), iv.end());
    double busy = 0.0, a = 0.0, b = -1.0;
    for (const auto &x : iv) {
        if (x.first > b) {
            if (b > a) busy += b - a;
            a = x.first; b = x.second;
        } else {
            b = std::max(b, x.second);
        }
    }
    if (b > a) busy += b - a;
    t6[0] = busy;
    t6[1] = s->work_done;
    t6[2] = s->any_submit ? std::max(0.0, ms_between(s->t_first, s->t_last)) : 0.0;
    t6[3] = s->prep_ms;
    t6[4] = s->launch_ms;
    t6[5] = (double)s->preps.size();
    return TBK_OK;
}

extern "C" int tbk_server_destroy(tbk_server *s)
{
    if (!s) return TBK_ERR_ARG;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        s->events++;
        s->cv_work.notify_all();
        s->cv_prep.notify_all();
    }
    for (std::thread &t : s->preps) t.join();
    s->worker.join();
    server_free(s);
    return TBK_OK;
}
