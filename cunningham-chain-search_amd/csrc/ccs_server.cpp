// Served form: a continuous-batching request loop over the batched kernels.
//
// Requests (class-form search jobs of any depth, width and length) are admitted while
// their resident device footprint fits the memory budget; every slice takes the next
// chunk of every admitted job, grouped by depth profile, runs one persistent batched
// filter launch + one prove launch per group, and hands results back per job.  Requests
// that arrive while slices are on the GPU join the next slice.  Two buffer lanes
// alternate so the GPU runs slice k+1 while the host distributes slice k; within a lane
// the proves run on a stream of their own, behind an event per group, so they overlap the
// next groups' filter launches.
//
// The reference has no serving mode; its closest pattern is the v16 dispatch ring with
// an asynchronous prove pool (v16...cu:2938-3279, :5594-5965).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/ccs.h"

namespace {

constexpr uint32_t kEagerCopy = 1024;   // survivors copied back with every slice

double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

struct Job {
    uint64_t id = 0;
    ccs_request req{};
    int group = -1;
    uint64_t cursor = 0;            // next u offset to hand out
    uint32_t inflight = 0;          // slices launched, not yet collected
    double t_submit = 0, t_admit = 0, t_done = 0;
    std::vector<ccs_job_survivor> surv;   // u_off relative to the request's u0
    std::vector<ccs_chain_result> res;
    bool done = false;
};

struct Entry {          // one job's chunk inside a slice
    uint64_t job;
    uint64_t base;      // u offset of the chunk within the job
};

struct Lane {           // buffers of one in-flight slice for one depth group
    ccs_jobset* js = nullptr;
    ccs_job_survivor* d_surv = nullptr;
    ccs_chain_result* d_res = nullptr;
    uint32_t* d_count = nullptr;
    ccs_job_survivor* h_surv = nullptr;   // pinned
    ccs_chain_result* h_res = nullptr;    // pinned
    uint32_t* h_count = nullptr;          // pinned
    std::vector<ccs_job> jobs;
    std::vector<uint32_t> job_shape;      // group lanes: the depth profile of each job
    std::vector<Entry> entries;
    hipEvent_t filtered = nullptr;        // the group's filter launch is done (prove may start)
    bool busy = false;
};

}  // namespace

struct ccs_server {
    ccs_ctx* ctx = nullptr;
    ccs_server_params p{};
    std::vector<const ccs_shape*> shapes;   // index = depth group
    std::vector<int32_t> depths;
    hipStream_t streams[2] = {nullptr, nullptr};    // per lane: filter launches
    hipStream_t pstreams[2] = {nullptr, nullptr};   // per lane: prove launches and copies
    std::vector<Lane> lanes[2];             // [lane][group]; one group set per lane when grouped
    bool grouped = false;                   // several depth profiles: one launch per slice
    std::mutex mu;
    std::condition_variable cv;
    std::deque<uint64_t> pending;
    std::map<uint64_t, Job> jobs;
    std::vector<uint64_t> active;
    std::vector<ccs_job_done> finished;     // completed since the last poll
    uint64_t next_id = 1;
    uint64_t resident_bytes = 0;
    uint64_t stat_candidates = 0, stat_slices = 0, stat_launches = 0;
    bool stop = false, drain = false;
    int error = 0;
    std::string error_msg;
    std::thread th;
};

static uint64_t job_footprint(const ccs_server_params& p) {
    // device-resident state of an admitted job: its job record in both lanes' slice tables
    // plus its share of the survivor and result slots
    return 2 * (sizeof(ccs_job) + (uint64_t)p.surv_per_job * (sizeof(ccs_job_survivor) + sizeof(ccs_chain_result)));
}

static int server_fail(ccs_server* s, int rc, const char* what) {
    std::lock_guard<std::mutex> g(s->mu);
    s->error = rc;
    s->error_msg = std::string(what) + ": " + ccs_last_error();
    s->stop = true;
    s->cv.notify_all();
    return rc;
}

// collect one lane's results into the jobs; mark finished jobs
static int collect_lane(ccs_server* s, int li) {
    bool any = false;
    for (auto& L : s->lanes[li]) any |= L.busy;
    if (!any) return CCS_OK;
    if (hipStreamSynchronize(s->streams[li]) != hipSuccess || hipStreamSynchronize(s->pstreams[li]) != hipSuccess)
        return CCS_ERR_HIP;
    const double t = now_s();
    std::lock_guard<std::mutex> g(s->mu);
    for (auto& L : s->lanes[li]) {
        if (!L.busy) continue;
        L.busy = false;
        uint32_t n = std::min(*L.h_count, s->p.max_slice_survivors);
        if (n > kEagerCopy) {   // the rest of a large slice's survivors
            if (hipMemcpy(L.h_surv + kEagerCopy, L.d_surv + kEagerCopy, (size_t)(n - kEagerCopy) * sizeof(ccs_job_survivor),
                          hipMemcpyDeviceToHost) != hipSuccess ||
                hipMemcpy(L.h_res + kEagerCopy, L.d_res + kEagerCopy, (size_t)(n - kEagerCopy) * sizeof(ccs_chain_result),
                          hipMemcpyDeviceToHost) != hipSuccess)
                return CCS_ERR_HIP;
        }
        for (uint32_t i = 0; i < n; ++i) {
            const ccs_job_survivor sv = L.h_surv[i];
            const Entry& e = L.entries[sv.job_idx];
            Job& j = s->jobs[e.job];
            ccs_job_survivor out;
            out.job_idx = 0;                                   // per-job list: the job itself
            out.u_off = (uint32_t)(e.base + sv.u_off);         // u offset within the request
            j.surv.push_back(out);
            ccs_chain_result r = L.h_res[i];
            r.survivor_idx = (uint32_t)(j.surv.size() - 1);
            j.res.push_back(r);
        }
        if (*L.h_count > s->p.max_slice_survivors) {
            s->error = CCS_OVERFLOW;
            s->error_msg = "slice survivor buffer overflow";
        }
        for (const Entry& e : L.entries) {
            Job& j = s->jobs[e.job];
            if (--j.inflight == 0 && j.cursor >= j.req.count && !j.done) {
                j.done = true;
                j.t_done = t;
                ccs_job_done d{};
                d.job_id = j.id;
                d.tag = j.req.tag;
                d.candidates = j.req.count;
                d.n_survivors = (uint32_t)j.surv.size();
                uint32_t primes = 0, best = 0;
                for (auto& r : j.res)
                    if (r.flags & CCS_F_PRIME) { ++primes; best = std::max(best, r.chain_len); }
                d.n_primes = primes;
                d.best_len = best;
                d.t_submit = j.t_submit;
                d.t_admit = j.t_admit;
                d.t_done = j.t_done;
                s->finished.push_back(d);
                s->resident_bytes -= job_footprint(s->p);
                s->active.erase(std::find(s->active.begin(), s->active.end(), j.id));
            }
        }
    }
    s->cv.notify_all();
    return CCS_OK;
}

static void server_loop(ccs_server* s) {
    int li = 0;
    for (;;) {
        // admission and slice planning under the lock
        std::vector<std::vector<Entry>> plan(s->shapes.size());
        std::vector<std::vector<ccs_job>> recs(s->shapes.size());
        {
            std::unique_lock<std::mutex> g(s->mu);
            for (;;) {
                while (!s->pending.empty() && s->active.size() < s->p.max_active &&
                       s->resident_bytes + job_footprint(s->p) <= s->p.mem_budget_bytes) {
                    Job& j = s->jobs[s->pending.front()];
                    s->pending.pop_front();
                    j.t_admit = now_s();
                    s->resident_bytes += job_footprint(s->p);
                    s->active.push_back(j.id);
                }
                bool work = false;
                for (uint64_t id : s->active)
                    if (s->jobs[id].cursor < s->jobs[id].req.count) { work = true; break; }
                bool inflight = false;
                for (auto& lv : s->lanes)
                    for (auto& L : lv) inflight |= L.busy;
                if (work || inflight || s->stop) break;
                if (s->drain && s->pending.empty()) { s->drain = false; s->cv.notify_all(); }
                s->cv.wait(g);
            }
            if (s->stop) break;
            // per-job quota: an even share of the slice target, at least one unit
            uint64_t n_work = 0;
            for (uint64_t id : s->active) n_work += s->jobs[id].cursor < s->jobs[id].req.count;
            uint64_t quota = n_work ? s->p.slice_candidates / n_work : 0;
            quota = std::max<uint64_t>(quota, s->p.min_chunk);
            quota = std::min<uint64_t>(quota, s->p.max_chunk);
            size_t planned = 0;
            for (uint64_t id : s->active) {
                Job& j = s->jobs[id];
                if (j.cursor >= j.req.count) continue;
                if (planned >= s->p.max_active) continue;   // (active jobs never exceed it)
                const uint64_t n = std::min<uint64_t>(quota, j.req.count - j.cursor);
                uint64_t u0lo = j.req.u0.lo + j.cursor;
                uint64_t u0hi = j.req.u0.hi + (u0lo < j.req.u0.lo);
                ccs_job r{};
                r.a = j.req.a;
                r.u0.lo = u0lo;
                r.u0.hi = u0hi;
                r.count = n;
                recs[j.group].push_back(r);
                plan[j.group].push_back(Entry{id, j.cursor});
                ++planned;
                j.cursor += n;
                j.inflight++;
                s->stat_candidates += n;
            }
        }
        // launch the slice on lane li: one group launch over every depth profile, or one
        // launch per profile
        bool launched = false;
        if (s->grouped) {   // flatten the plan in profile order (group sets want jobs grouped)
            for (size_t gi = 1; gi < s->shapes.size(); ++gi) {
                plan[0].insert(plan[0].end(), plan[gi].begin(), plan[gi].end());
                recs[0].insert(recs[0].end(), recs[gi].begin(), recs[gi].end());
            }
        }
        for (size_t gi = 0; gi < s->lanes[li].size(); ++gi) {
            if (plan[gi].empty()) continue;
            Lane& L = s->lanes[li][gi];
            L.jobs = recs[gi];
            L.entries = plan[gi];
            int rc;
            if (s->grouped) {
                L.job_shape.clear();
                for (const Entry& e : L.entries) L.job_shape.push_back((uint32_t)s->jobs[e.job].group);
                rc = ccs_groupset_update(s->ctx, L.js, L.jobs.data(), L.job_shape.data(), (uint32_t)L.jobs.size(),
                                         s->streams[li]);
            } else {
                rc = ccs_jobset_update(s->ctx, L.js, L.jobs.data(), (uint32_t)L.jobs.size(), s->streams[li]);
            }
            if (!rc) rc = ccs_jobset_launch(s->ctx, L.js, L.d_surv, s->p.max_slice_survivors, L.d_count, s->streams[li]);
            // the prove of this group runs on the lane's prove stream, so the next group's
            // filter launch does not wait behind the prove's latency-bound tail
            if (!rc && (hipEventRecord(L.filtered, s->streams[li]) != hipSuccess ||
                        hipStreamWaitEvent(s->pstreams[li], L.filtered, 0) != hipSuccess))
                rc = CCS_ERR_HIP;
            if (!rc) {
                ccs_prove_params pp = s->p.prove;
                rc = ccs_prove_jobset_launch(s->ctx, L.js, L.d_surv, L.d_count, s->p.max_slice_survivors, &pp,
                                             L.d_res, s->pstreams[li]);
            }
            if (!rc && hipMemcpyAsync(L.h_count, L.d_count, 4, hipMemcpyDeviceToHost, s->pstreams[li]) != hipSuccess)
                rc = CCS_ERR_HIP;
            // survivors are rare: copy a bounded prefix now, the rest (if any) at collection
            const size_t pre = std::min<uint32_t>(kEagerCopy, s->p.max_slice_survivors);
            if (!rc && hipMemcpyAsync(L.h_surv, L.d_surv, pre * sizeof(ccs_job_survivor),
                                      hipMemcpyDeviceToHost, s->pstreams[li]) != hipSuccess)
                rc = CCS_ERR_HIP;
            if (!rc && hipMemcpyAsync(L.h_res, L.d_res, pre * sizeof(ccs_chain_result),
                                      hipMemcpyDeviceToHost, s->pstreams[li]) != hipSuccess)
                rc = CCS_ERR_HIP;
            if (rc) { server_fail(s, rc, "slice launch"); return; }
            L.busy = true;
            launched = true;
            s->stat_launches++;
        }
        if (launched) s->stat_slices++;
        // while lane li runs, collect the other lane
        if (collect_lane(s, li ^ 1)) { server_fail(s, CCS_ERR_HIP, "collect"); return; }
        li ^= 1;
    }
    collect_lane(s, 0);
    collect_lane(s, 1);
}

extern "C" int ccs_server_create(ccs_ctx* ctx, const ccs_shape* const* shapes, const int32_t* depths,
                                 uint32_t n_shapes, const ccs_server_params* params, ccs_server** out) {
    if (!ctx || !shapes || !depths || !n_shapes || !params || !out) return CCS_ERR_INVALID;
    *out = nullptr;
    ccs_server* s = new ccs_server();
    s->ctx = ctx;
    s->p = *params;
    if (!s->p.max_active) s->p.max_active = 1024;
    if (!s->p.surv_per_job) s->p.surv_per_job = 16;
    if (!s->p.max_slice_survivors) s->p.max_slice_survivors = 1u << 16;
    if (!s->p.slice_candidates) s->p.slice_candidates = 1ull << 35;
    if (!s->p.min_chunk) s->p.min_chunk = 1u << 18;
    if (!s->p.max_chunk) s->p.max_chunk = 1ull << 32;
    if (!s->p.mem_budget_bytes) s->p.mem_budget_bytes = 256ull << 20;
    s->p.max_chunk = std::min<uint64_t>(s->p.max_chunk, 1ull << 32);
    for (uint32_t i = 0; i < n_shapes; ++i) {
        s->shapes.push_back(shapes[i]);
        s->depths.push_back(depths[i]);
    }
    int rc = CCS_OK;
    // several depth profiles with one row layout: one group launch per slice
    s->grouped = n_shapes > 1;
    {
        ccs_tuning t0, t;
        rc = ccs_shape_get_tuning(shapes[0], &t0);
        for (uint32_t i = 1; i < n_shapes && !rc; ++i) {
            rc = ccs_shape_get_tuning(shapes[i], &t);
            if (!rc && t.rows_per_group != t0.rows_per_group) s->grouped = false;
        }
    }
    const uint64_t slice = std::max<uint64_t>(s->p.slice_candidates, s->p.min_chunk);
    const uint64_t per = std::min<uint64_t>(std::max<uint64_t>(s->p.slice_candidates / s->p.max_active, s->p.min_chunk),
                                            s->p.max_chunk);
    // job table sized for max_active jobs, stage-B buffer for a whole slice of any profile
    auto dummy_job = [&](uint64_t count) {
        ccs_job d;
        d.a = 0;
        d.u0.lo = 1;
        d.u0.hi = 0;
        d.count = count;
        return d;
    };
    for (int li = 0; li < 2 && !rc; ++li) {
        if (hipStreamCreateWithFlags(&s->streams[li], hipStreamNonBlocking) != hipSuccess ||
            hipStreamCreateWithFlags(&s->pstreams[li], hipStreamNonBlocking) != hipSuccess)
            rc = CCS_ERR_HIP;
        s->lanes[li].resize(s->grouped ? 1 : n_shapes);
        for (uint32_t gi = 0; gi < s->lanes[li].size() && !rc; ++gi) {
            Lane& L = s->lanes[li][gi];
            if (s->grouped) {
                std::vector<ccs_job> dummy;
                std::vector<uint32_t> dshape;
                for (uint32_t k = 0; k < n_shapes; ++k)
                    for (uint64_t left = slice; left > 0;) {
                        const uint64_t c = std::min<uint64_t>(left, s->p.max_chunk);
                        dummy.push_back(dummy_job(c));
                        dshape.push_back(k);
                        left -= c;
                    }
                while (dummy.size() < s->p.max_active) {
                    dummy.push_back(dummy_job(per));
                    dshape.push_back(n_shapes - 1);
                }
                rc = ccs_groupset_create(ctx, shapes, n_shapes, dummy.data(), dshape.data(), (uint32_t)dummy.size(),
                                         &L.js);
            } else {
                std::vector<ccs_job> dummy(s->p.max_active, dummy_job(per));
                dummy[0].count = std::min<uint64_t>(slice, s->p.max_chunk);
                rc = ccs_jobset_create(ctx, shapes[gi], dummy.data(), (uint32_t)dummy.size(), &L.js);
            }
            if (!rc) rc = ccs_dmalloc(ctx, (void**)&L.d_surv, (uint64_t)s->p.max_slice_survivors * sizeof(ccs_job_survivor));
            if (!rc) rc = ccs_dmalloc(ctx, (void**)&L.d_res, (uint64_t)s->p.max_slice_survivors * sizeof(ccs_chain_result));
            if (!rc) rc = ccs_dmalloc(ctx, (void**)&L.d_count, 64);
            if (!rc && hipHostMalloc((void**)&L.h_surv, (size_t)s->p.max_slice_survivors * sizeof(ccs_job_survivor)) != hipSuccess) rc = CCS_ERR_NOMEM;
            if (!rc && hipHostMalloc((void**)&L.h_res, (size_t)s->p.max_slice_survivors * sizeof(ccs_chain_result)) != hipSuccess) rc = CCS_ERR_NOMEM;
            if (!rc && hipHostMalloc((void**)&L.h_count, 64) != hipSuccess) rc = CCS_ERR_NOMEM;
            if (!rc && hipEventCreateWithFlags(&L.filtered, hipEventDisableTiming) != hipSuccess) rc = CCS_ERR_HIP;
        }
    }
    if (rc) {
        ccs_server_shutdown(s);
        return rc;
    }
    s->th = std::thread(server_loop, s);
    *out = s;
    return CCS_OK;
}

extern "C" int ccs_server_submit(ccs_server* s, const ccs_request* req, uint64_t* job_id) {
    if (!s || !req) return CCS_ERR_INVALID;
    int group = -1;
    for (size_t i = 0; i < s->depths.size(); ++i)
        if (s->depths[i] == req->depth) group = (int)i;
    if (group < 0) return CCS_ERR_INVALID;
    std::lock_guard<std::mutex> g(s->mu);
    if (s->stop) return s->error ? s->error : CCS_ERR_BUSY;
    Job j;
    j.id = s->next_id++;
    j.req = *req;
    j.group = group;
    j.t_submit = now_s();
    if (j.req.count == 0) {
        j.done = true;
        j.t_admit = j.t_done = j.t_submit;
        ccs_job_done d{};
        d.job_id = j.id;
        d.tag = req->tag;
        d.t_submit = d.t_admit = d.t_done = j.t_submit;
        s->finished.push_back(d);
    } else {
        s->pending.push_back(j.id);
    }
    if (job_id) *job_id = j.id;
    s->jobs[j.id] = std::move(j);
    s->cv.notify_all();
    return CCS_OK;
}

extern "C" int ccs_server_poll(ccs_server* s, ccs_job_done* out, uint32_t cap, uint32_t* n) {
    if (!s || !n) return CCS_ERR_INVALID;
    std::lock_guard<std::mutex> g(s->mu);
    const uint32_t k = std::min<uint32_t>(cap, (uint32_t)s->finished.size());
    for (uint32_t i = 0; i < k; ++i) out[i] = s->finished[i];
    s->finished.erase(s->finished.begin(), s->finished.begin() + k);
    *n = k;
    return s->error < 0 ? s->error : CCS_OK;
}

extern "C" int ccs_server_job_results(ccs_server* s, uint64_t job_id, ccs_job_survivor* surv,
                                      ccs_chain_result* res, uint32_t cap, uint32_t* n) {
    if (!s || !n) return CCS_ERR_INVALID;
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->jobs.find(job_id);
    if (it == s->jobs.end() || !it->second.done) return CCS_ERR_INVALID;
    const Job& j = it->second;
    const uint32_t k = std::min<uint32_t>(cap, (uint32_t)j.surv.size());
    for (uint32_t i = 0; i < k; ++i) {
        if (surv) surv[i] = j.surv[i];
        if (res) res[i] = j.res[i];
    }
    *n = (uint32_t)j.surv.size();
    return k < j.surv.size() ? CCS_OVERFLOW : CCS_OK;
}

extern "C" int ccs_server_drain(ccs_server* s) {
    if (!s) return CCS_ERR_INVALID;
    std::unique_lock<std::mutex> g(s->mu);
    s->cv.wait(g, [&] {
        if (s->stop) return true;
        if (!s->pending.empty()) return false;
        for (uint64_t id : s->active) if (!s->jobs[id].done) return false;
        return true;
    });
    return s->error < 0 ? s->error : CCS_OK;
}

extern "C" int ccs_server_stats(ccs_server* s, ccs_server_counters* out) {
    if (!s || !out) return CCS_ERR_INVALID;
    std::lock_guard<std::mutex> g(s->mu);
    out->candidates = s->stat_candidates;
    out->slices = s->stat_slices;
    out->launches = s->stat_launches;
    out->active = (uint32_t)s->active.size();
    out->pending = (uint32_t)s->pending.size();
    out->resident_bytes = s->resident_bytes;
    return CCS_OK;
}

extern "C" int ccs_server_shutdown(ccs_server* s) {
    if (!s) return CCS_ERR_INVALID;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stop = true;
        s->cv.notify_all();
    }
    if (s->th.joinable()) s->th.join();
    for (int li = 0; li < 2; ++li) {
        if (s->streams[li]) (void)hipStreamSynchronize(s->streams[li]);
        if (s->pstreams[li]) (void)hipStreamSynchronize(s->pstreams[li]);
        for (auto& L : s->lanes[li]) {
            if (L.filtered) (void)hipEventDestroy(L.filtered);
            if (L.js) ccs_jobset_free(L.js);
            if (L.d_surv) ccs_dfree(s->ctx, L.d_surv);
            if (L.d_res) ccs_dfree(s->ctx, L.d_res);
            if (L.d_count) ccs_dfree(s->ctx, L.d_count);
            if (L.h_surv) (void)hipHostFree(L.h_surv);
            if (L.h_res) (void)hipHostFree(L.h_res);
            if (L.h_count) (void)hipHostFree(L.h_count);
        }
        if (s->streams[li]) (void)hipStreamDestroy(s->streams[li]);
        if (s->pstreams[li]) (void)hipStreamDestroy(s->pstreams[li]);
    }
    const int err = s->error;
    delete s;
    return err < 0 ? err : CCS_OK;
}
