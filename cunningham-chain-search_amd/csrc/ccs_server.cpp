// Served form: a continuous-batching request loop over the batched kernels.
//
// Requests (class-form search jobs of any depth, width and length) are admitted while a
// slot of the server's device arena is free: the arena holds each admitted job's residue
// state (the phases of its sieve primes at the request's u0, written once by a kernel at
// admission) and is sized by the memory budget.  Every slice takes the next chunk of every
// admitted job and runs, per launch group of depth profiles, one persistent group launch
// (the jobs' sieve phases advanced from their resident state) + one prove launch, and hands
// results back per job.  Depth profiles share a launch group when their sieve kernel
// instances are close (same row layout, stage-B register widths within a factor of two): a
// launch runs at the widest instance of its group, and the shallow profiles' 16-group
// instance slows the deep ones.  Requests that arrive while slices are on the GPU join the
// next slice.  Three buffer lanes rotate: slice k is launched before slice k-2 is collected,
// so the host's collection and planning never leave the GPU without a queued slice (with
// two lanes, a slice could only be launched once the previous one's prove, which runs in
// the tail of the next filter, was collected).  The filter launches of every slice go to
// one stream and the proves (with the result copies) to a second one, behind an event per
// launch group, so a slice's prove overlaps the next slice's filter: with a filter and a
// prove stream per lane, the runtime put one lane's filter stream and another's prove
// stream on one hardware queue (streams share the few hardware queues round-robin), which
// ran every filter behind the previous slice's prove.
//
// The reference has no serving mode; its closest pattern is the v16 dispatch ring with
// an asynchronous prove pool (v16...cu:2938-3279, :5594-5965), whose drained per-variant
// cursors (v16...cu:4947-4955, :3104-3115) are the analogue of the resident job state here.
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
#include "ccs_internal.hpp"

namespace {

using ccs::kNoSlot;

constexpr uint32_t kEagerMin = 4096;         // survivors copied back with every slice, at least
constexpr int kLanes = 3;                     // slices in flight: one collected, two on the GPU
constexpr uint64_t kMaxRequest = 1ull << 48;  // resident phases advance exactly below this offset

double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

struct Job {
    uint64_t id = 0;
    ccs_request req{};
    int group = -1;
    uint32_t slot = kNoSlot;        // arena slot while admitted
    uint64_t cursor = 0;            // next u offset to hand out
    uint32_t inflight = 0;          // slices launched, not yet collected
    double t_submit = 0, t_admit = 0, t_done = 0;
    std::vector<uint64_t> u_off;    // survivors: u offset within the request
    std::vector<ccs_chain_result> res;
    bool done = false, incomplete = false;
};

struct Entry {          // one job's chunk inside a slice
    uint64_t job;
    uint64_t base;      // u offset of the chunk within the job
};

struct Lane {           // buffers of one in-flight slice for one launch group
    ccs_jobset* js = nullptr;
    ccs_job_survivor* d_surv = nullptr;
    ccs_chain_result* d_res = nullptr;
    uint32_t* d_count = nullptr;
    ccs_job_survivor* h_surv = nullptr;   // pinned
    ccs_chain_result* h_res = nullptr;    // pinned
    uint32_t* h_count = nullptr;          // pinned: [0] survivors, [1] overflowed launches so far
    uint32_t seen_overflows = 0;          // h_count[1] at the last collection
    std::vector<ccs_job> jobs;
    std::vector<uint32_t> job_shape;      // the depth profile of each job
    std::vector<uint32_t> slots;          // the arena slot of each job
    std::vector<uint64_t> offs;           // each chunk's u offset from the resident phases
    std::vector<Entry> entries;
    hipEvent_t filtered = nullptr;        // the group's filter launch is done (prove may start)
    hipEvent_t done = nullptr;            // its prove and result copies are done (buffers free)
    uint32_t eager = 0;                   // survivor entries copied back with the slice
    bool busy = false;
};

}  // namespace

struct ccs_server {
    ccs_ctx* ctx = nullptr;
    ccs_server_params p{};
    std::vector<const ccs_shape*> shapes;   // index = depth group
    std::vector<int32_t> depths;
    hipStream_t fstream = nullptr;          // phase inits and filter launches
    hipStream_t pstream = nullptr;          // prove launches and their result copies
    hipStream_t cstream = nullptr;          // copies of a large slice's remaining results
    std::vector<Lane> lanes[kLanes];        // [lane][launch group]
    // launch groups: the depth profiles whose sieve launches share one kernel instance
    // closely enough to run as one group launch (same row layout, stage-B register widths
    // within a factor of two); lg_of[profile], pos_in_lg[profile] = its index in the group
    std::vector<std::vector<const ccs_shape*>> lg_shapes;
    std::vector<std::vector<uint32_t>> lg_profiles;
    std::vector<uint32_t> lg_of, pos_in_lg;
    bool resident = true;                   // phases from the arena (else from u0 per unit)
    // the arena: per slot `stride` words of phases
    uint32_t stride = 0, n_slots = 0;
    uint32_t* d_phases = nullptr;
    uint64_t job_bytes = 0, fixed_bytes = 0;
    std::vector<uint32_t> free_slots;
    std::vector<ccs::PhaseInit> inits;      // admitted since the last slice launch
    ccs::PhaseInit* h_init[kLanes] = {};    // pinned staging per lane
    ccs::PhaseInit* d_init[kLanes] = {};
    std::mutex mu;
    std::condition_variable cv;
    std::deque<uint64_t> pending;
    std::map<uint64_t, Job> jobs;
    std::vector<uint64_t> active;
    std::vector<ccs_job_done> finished;     // completed since the last poll
    uint64_t next_id = 1;
    uint64_t resident_bytes = 0;
    uint64_t stat_candidates = 0, stat_slices = 0, stat_launches = 0;
    uint32_t n_incomplete = 0;
    double busy_s = 0;
    bool stop = false;
    int error = 0;
    std::string error_msg;
    std::thread th;
};

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
    for (auto& L : s->lanes[li])
        if (L.busy && hipEventSynchronize(L.done) != hipSuccess) return CCS_ERR_HIP;
    // survivors lost to a full stage-B buffer (the set's overflow counter, copied back with
    // the survivor count, has grown) or a full slice buffer: the slice's jobs are finished
    // incomplete (and flagged so)
    std::vector<char> lost(s->lanes[li].size(), 0);
    for (size_t gi = 0; gi < s->lanes[li].size(); ++gi) {
        Lane& L = s->lanes[li][gi];
        if (!L.busy) continue;
        lost[gi] = L.h_count[1] != L.seen_overflows || L.h_count[0] > s->p.max_slice_survivors;
        L.seen_overflows = L.h_count[1];
    }
    const double t = now_s();
    std::lock_guard<std::mutex> g(s->mu);
    for (size_t gi = 0; gi < s->lanes[li].size(); ++gi) {
        Lane& L = s->lanes[li][gi];
        if (!L.busy) continue;
        L.busy = false;
        const uint32_t n = std::min(*L.h_count, s->p.max_slice_survivors);
        if (n > L.eager) {   // the rest of an unexpectedly large slice's survivors (a wait that
                             // may last until the next slice's filter launch has finished)
            if (hipMemcpyAsync(L.h_surv + L.eager, L.d_surv + L.eager,
                               (size_t)(n - L.eager) * sizeof(ccs_job_survivor), hipMemcpyDeviceToHost,
                               s->cstream) != hipSuccess ||
                hipMemcpyAsync(L.h_res + L.eager, L.d_res + L.eager,
                               (size_t)(n - L.eager) * sizeof(ccs_chain_result), hipMemcpyDeviceToHost,
                               s->cstream) != hipSuccess ||
                hipStreamSynchronize(s->cstream) != hipSuccess)
                return CCS_ERR_HIP;
        }
        for (uint32_t i = 0; i < n; ++i) {
            const ccs_job_survivor sv = L.h_surv[i];
            const Entry& e = L.entries[sv.job_idx];
            Job& j = s->jobs[e.job];
            j.u_off.push_back(e.base + sv.u_off);   // 64-bit: requests can exceed 2^32
            ccs_chain_result r = L.h_res[i];
            r.survivor_idx = (uint32_t)(j.u_off.size() - 1);
            j.res.push_back(r);
        }
        for (const Entry& e : L.entries) {
            Job& j = s->jobs[e.job];
            if (lost[gi]) j.incomplete = true;
            if (--j.inflight == 0 && j.cursor >= j.req.count && !j.done) {
                j.done = true;
                j.t_done = t;
                ccs_job_done d{};
                d.job_id = j.id;
                d.tag = j.req.tag;
                d.candidates = j.req.count;
                d.n_survivors = (uint32_t)j.u_off.size();
                uint32_t primes = 0, best = 0;
                for (auto& r : j.res)
                    if (r.flags & CCS_F_PRIME) { ++primes; best = std::max(best, r.chain_len); }
                d.n_primes = primes;
                d.best_len = best;
                d.flags = j.incomplete ? CCS_JOB_INCOMPLETE : 0u;
                s->n_incomplete += j.incomplete ? 1u : 0u;
                d.t_submit = j.t_submit;
                d.t_admit = j.t_admit;
                d.t_done = j.t_done;
                s->finished.push_back(d);
                s->resident_bytes -= s->job_bytes;
                s->free_slots.push_back(j.slot);
                j.slot = kNoSlot;
                s->active.erase(std::find(s->active.begin(), s->active.end(), j.id));
            }
        }
    }
    s->cv.notify_all();
    return CCS_OK;
}

// the phase-init kernel for the jobs admitted since the last slice, on the filter stream
// (ahead of every filter launch that reads their phases)
static int launch_inits(ccs_server* s, int li, std::vector<ccs::PhaseInit>& inits) {
    if (inits.empty()) return CCS_OK;
    const uint32_t n = (uint32_t)inits.size();   // at most n_slots: one per free slot
    std::memcpy(s->h_init[li], inits.data(), (size_t)n * sizeof(ccs::PhaseInit));
    if (hipMemcpyAsync(s->d_init[li], s->h_init[li], (size_t)n * sizeof(ccs::PhaseInit), hipMemcpyHostToDevice,
                       s->fstream) != hipSuccess ||
        ccs::launch_phase_init(s->d_init[li], n, s->d_phases, s->stride, s->fstream) != hipSuccess)
        return ccs::set_error(CCS_ERR_HIP, "phase init launch");
    inits.clear();
    return CCS_OK;
}

static void server_loop(ccs_server* s) {
    int li = 0;
    for (;;) {
        // admission and slice planning under the lock
        const size_t n_groups = s->lanes[0].size();
        std::vector<std::vector<Entry>> plan(n_groups);
        std::vector<std::vector<ccs_job>> recs(n_groups);
        std::vector<std::vector<uint32_t>> shp(n_groups), slots(n_groups);
        std::vector<std::vector<uint64_t>> offs(n_groups);
        std::vector<ccs::PhaseInit> inits;
        double t_busy = 0;
        {
            std::unique_lock<std::mutex> g(s->mu);
            for (;;) {
                while (!s->pending.empty() && s->active.size() < s->p.max_active && !s->free_slots.empty()) {
                    Job& j = s->jobs[s->pending.front()];
                    s->pending.pop_front();
                    j.t_admit = now_s();
                    j.slot = s->free_slots.back();
                    s->free_slots.pop_back();
                    s->resident_bytes += s->job_bytes;
                    s->active.push_back(j.id);
                    if (s->resident) {
                        const ccs_shape* sh = s->shapes[j.group];
                        ccs::PhaseInit pi;
                        pi.pmod = ccs::shape_pmod(sh);
                        pi.n = ccs::shape_phase_words(sh);
                        pi.slot = j.slot;
                        pi.a = j.req.a;
                        pi.a_hi = 0;
                        pi.u_lo = j.req.u0.lo;
                        pi.u_hi = j.req.u0.hi;
                        s->inits.push_back(pi);
                    }
                }
                bool work = false;
                for (uint64_t id : s->active)
                    if (s->jobs[id].cursor < s->jobs[id].req.count) { work = true; break; }
                bool inflight = false;
                for (auto& lv : s->lanes)
                    for (auto& L : lv) inflight |= L.busy;
                if (work || inflight || s->stop) break;
                s->cv.notify_all();   // idle: a drain may be waiting
                s->cv.wait(g);
            }
            if (s->stop) break;
            t_busy = now_s();
            inits.swap(s->inits);
            // per-job quota: an even share of the slice target, at least min_chunk; a share
            // above max_chunk goes into the launch as several pieces (rows of the job table,
            // each with its own offset), so the slices stay full when few jobs are left
            uint64_t n_work = 0;
            for (uint64_t id : s->active) n_work += s->jobs[id].cursor < s->jobs[id].req.count;
            uint64_t quota = n_work ? s->p.slice_candidates / n_work : 0;
            quota = std::max<uint64_t>(quota, s->p.min_chunk);
            for (uint64_t id : s->active) {
                Job& j = s->jobs[id];
                if (j.cursor >= j.req.count) continue;
                for (uint64_t share = std::min<uint64_t>(quota, j.req.count - j.cursor); share > 0;) {
                    const uint64_t n = std::min<uint64_t>(share, s->p.max_chunk);
                    const uint64_t u0lo = j.req.u0.lo + j.cursor;
                    ccs_job r{};
                    r.a = j.req.a;
                    r.u0.lo = u0lo;
                    r.u0.hi = j.req.u0.hi + (u0lo < j.req.u0.lo);
                    r.count = n;
                    const size_t gi = s->lg_of[j.group];
                    recs[gi].push_back(r);
                    shp[gi].push_back(s->pos_in_lg[j.group]);
                    slots[gi].push_back(j.slot);
                    offs[gi].push_back(j.cursor);   // the resident phases are those at u0
                    plan[gi].push_back(Entry{id, j.cursor});
                    j.cursor += n;
                    j.inflight++;
                    s->stat_candidates += n;
                    share -= n;
                }
            }
        }
        // group sets want their jobs grouped by shape: order each launch group's jobs by profile
        for (size_t gi = 0; gi < n_groups; ++gi) {
            std::vector<size_t> order(recs[gi].size());
            for (size_t i = 0; i < order.size(); ++i) order[i] = i;
            std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return shp[gi][a] < shp[gi][b]; });
            Lane& L = s->lanes[li][gi];
            L.jobs.clear();
            L.job_shape.clear();
            L.slots.clear();
            L.offs.clear();
            L.entries.clear();
            for (size_t i : order) {
                L.jobs.push_back(recs[gi][i]);
                L.job_shape.push_back(shp[gi][i]);
                L.slots.push_back(slots[gi][i]);
                L.offs.push_back(offs[gi][i]);
                L.entries.push_back(plan[gi][i]);
            }
        }
        // the lane's staging buffer is free: its previous slice was collected
        int rc = launch_inits(s, li, inits);
        if (rc) { server_fail(s, rc, "phase init"); return; }
        bool launched = false;
        for (size_t gi = 0; gi < n_groups; ++gi) {
            Lane& L = s->lanes[li][gi];
            if (L.jobs.empty()) continue;
            const uint32_t nj = (uint32_t)L.jobs.size();
            if (s->resident)
                rc = ccs::groupset_update_resident(s->ctx, L.js, L.jobs.data(), L.job_shape.data(), L.slots.data(),
                                                   L.offs.data(), nj, s->d_phases, s->stride, s->fstream);
            else
                rc = ccs_groupset_update(s->ctx, L.js, L.jobs.data(), L.job_shape.data(), nj, s->fstream);
            if (!rc) rc = ccs_jobset_launch(s->ctx, L.js, L.d_surv, s->p.max_slice_survivors, L.d_count, s->fstream);
            // the prove runs on the prove stream, so the next filter launch does not wait
            // behind the prove's latency-bound tail
            if (!rc && (hipEventRecord(L.filtered, s->fstream) != hipSuccess ||
                        hipStreamWaitEvent(s->pstream, L.filtered, 0) != hipSuccess))
                rc = CCS_ERR_HIP;
            if (!rc) {
                ccs_prove_params pp = s->p.prove;
                rc = ccs_prove_jobset_launch(s->ctx, L.js, L.d_surv, L.d_count, s->p.max_slice_survivors, &pp,
                                             L.d_res, s->pstream);
            }
            if (!rc && (hipMemcpyAsync(L.h_count, L.d_count, 4, hipMemcpyDeviceToHost, s->pstream) != hipSuccess ||
                        hipMemcpyAsync(L.h_count + 1, ccs::jobset_status_words(L.js), 4, hipMemcpyDeviceToHost,
                                       s->pstream) != hipSuccess))
                rc = CCS_ERR_HIP;
            // copy back twice the slice's expected survivors (at least kEagerMin) behind the
            // prove: the rest, if any, is copied at collection, when a copy can wait behind
            // the next slice's filter launch, which holds every CU
            double expect = 0;
            for (size_t i = 0; i < L.jobs.size(); ++i)
                expect += (double)L.jobs[i].count *
                          ccs::shape_final_survival(s->lg_shapes[gi][L.job_shape[i]]);
            L.eager = (uint32_t)std::min<double>(s->p.max_slice_survivors, 2.0 * expect + kEagerMin);
            const size_t pre = L.eager;
            if (!rc && hipMemcpyAsync(L.h_surv, L.d_surv, pre * sizeof(ccs_job_survivor),
                                      hipMemcpyDeviceToHost, s->pstream) != hipSuccess)
                rc = CCS_ERR_HIP;
            if (!rc && hipMemcpyAsync(L.h_res, L.d_res, pre * sizeof(ccs_chain_result),
                                      hipMemcpyDeviceToHost, s->pstream) != hipSuccess)
                rc = CCS_ERR_HIP;
            if (!rc && hipEventRecord(L.done, s->pstream) != hipSuccess) rc = CCS_ERR_HIP;
            if (rc) { server_fail(s, rc, "slice launch"); return; }
            L.busy = true;
            launched = true;
            s->stat_launches++;
        }
        if (launched) s->stat_slices++;
        // while lanes li and li-1 are on the GPU, collect the oldest slice (lane li+1, whose
        // buffers the next slice takes)
        const int next = (li + 1) % kLanes;
        if ((rc = collect_lane(s, next))) { server_fail(s, rc, "collect"); return; }
        {
            std::lock_guard<std::mutex> g(s->mu);
            s->busy_s += now_s() - t_busy;
        }
        li = next;
    }
    for (int k = 1; k <= kLanes; ++k) collect_lane(s, (li + k) % kLanes);
}

static void destroy(ccs_server* s) {
    for (hipStream_t st : {s->fstream, s->pstream, s->cstream})
        if (st) (void)hipStreamSynchronize(st);
    for (int li = 0; li < kLanes; ++li) {
        for (auto& L : s->lanes[li]) {
            if (L.filtered) (void)hipEventDestroy(L.filtered);
            if (L.done) (void)hipEventDestroy(L.done);
            if (L.js) ccs_jobset_free(L.js);
            if (L.d_surv) ccs_dfree(s->ctx, L.d_surv);
            if (L.d_res) ccs_dfree(s->ctx, L.d_res);
            if (L.d_count) ccs_dfree(s->ctx, L.d_count);
            if (L.h_surv) (void)hipHostFree(L.h_surv);
            if (L.h_res) (void)hipHostFree(L.h_res);
            if (L.h_count) (void)hipHostFree(L.h_count);
        }
        if (s->h_init[li]) (void)hipHostFree(s->h_init[li]);
        if (s->d_init[li]) ccs_dfree(s->ctx, s->d_init[li]);
    }
    for (hipStream_t st : {s->fstream, s->pstream, s->cstream})
        if (st) (void)hipStreamDestroy(st);
    if (s->d_phases) ccs_dfree(s->ctx, s->d_phases);
    delete s;
}

extern "C" int ccs_server_create(ccs_ctx* ctx, const ccs_shape* const* shapes, const int32_t* depths,
                                 uint32_t n_shapes, const ccs_server_params* params, ccs_server** out) {
    if (!ctx || !shapes || !depths || !n_shapes || !params || !out)
        return ccs::set_error(CCS_ERR_INVALID, "null argument");
    *out = nullptr;
    ccs_server* s = new ccs_server();
    s->ctx = ctx;
    s->p = *params;
    if (!s->p.max_active) s->p.max_active = 1024;
    if (!s->p.max_slice_survivors) s->p.max_slice_survivors = 1u << 16;
    if (!s->p.slice_candidates) s->p.slice_candidates = 1ull << 35;
    if (!s->p.min_chunk) s->p.min_chunk = 1u << 18;
    if (!s->p.max_chunk) s->p.max_chunk = 1ull << 32;
    if (!s->p.mem_budget_bytes) s->p.mem_budget_bytes = 256ull << 20;
    s->p.max_chunk = std::min<uint64_t>(s->p.max_chunk, 1ull << 32);
    s->p.min_chunk = std::min<uint64_t>(s->p.min_chunk, s->p.max_chunk);   // one piece per job at least
    s->resident = !(s->p.flags & CCS_SERVER_RECOMPUTE_PHASES);
    for (uint32_t i = 0; i < n_shapes; ++i) {
        if (!shapes[i]) { delete s; return ccs::set_error(CCS_ERR_INVALID, "null shape"); }
        s->shapes.push_back(shapes[i]);
        s->depths.push_back(depths[i]);
        s->stride = std::max(s->stride, ccs::shape_phase_words(shapes[i]));
    }
    s->stride = (s->stride + 63) & ~63u;   // slots on 256-byte boundaries
    // one admitted job's device state: its phases in the arena, and its rows (job record and
    // unit prefix) in every lane's slice table
    s->job_bytes = (s->resident ? (uint64_t)s->stride * 4 : 0) + kLanes * (sizeof(ccs::JobRec) + 8);
    const uint64_t fit = s->p.mem_budget_bytes / s->job_bytes;
    s->n_slots = (uint32_t)std::min<uint64_t>(s->p.max_active, fit);
    if (!s->n_slots) {
        delete s;
        return ccs::set_error(CCS_ERR_INVALID, "the memory budget holds no job's resident state");
    }
    s->p.max_active = std::min(s->p.max_active, s->n_slots);
    for (uint32_t k = s->n_slots; k-- > 0;) s->free_slots.push_back(k);
    int rc = CCS_OK;
    // launch groups: profiles sorted by stage-B register width (widest first); a group
    // takes the next profile while it has the group's row layout and (with a ratio set) at
    // least 1/ratio of the group's widest stage-B width.  By default (ratio 0) every profile
    // of a row layout shares one launch per slice: one persistent launch at the widest
    // profile's kernel instance runs the served mix faster than a launch per width class
    // (fewer launch tails per slice)
    {
        std::vector<uint32_t> order(n_shapes), width(n_shapes), rows(n_shapes);
        for (uint32_t i = 0; i < n_shapes && !rc; ++i) {
            ccs_tuning t;
            rc = ccs_shape_get_tuning(shapes[i], &t);
            ccs_shape_info info;
            if (!rc) rc = ccs_shape_query(shapes[i], &info);
            order[i] = i;
            rows[i] = t.rows_per_group;
            width[i] = ccs::sieve_stage_b_width(info.n_tierb);
        }
        std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
            return rows[a] != rows[b] ? rows[a] < rows[b] : width[a] > width[b];
        });
        s->lg_of.assign(n_shapes, 0);
        s->pos_in_lg.assign(n_shapes, 0);
        for (uint32_t k = 0; k < n_shapes && !rc; ++k) {
            const uint32_t i = order[k];
            bool fresh = s->lg_profiles.empty();
            if (!fresh) {
                const uint32_t head = s->lg_profiles.back().front();
                fresh = rows[head] != rows[i] ||
                        (s->p.group_width_ratio && s->p.group_width_ratio * width[i] < width[head]);
            }
            if (fresh) {
                s->lg_profiles.emplace_back();
                s->lg_shapes.emplace_back();
            }
            s->lg_of[i] = (uint32_t)s->lg_profiles.size() - 1;
            s->pos_in_lg[i] = (uint32_t)s->lg_profiles.back().size();
            s->lg_profiles.back().push_back(i);
            s->lg_shapes.back().push_back(shapes[i]);
        }
    }
    if (s->resident && !rc)
        rc = ccs_dmalloc(ctx, (void**)&s->d_phases, (uint64_t)s->n_slots * s->stride * 4);
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
    // created one after the other, so the runtime's round-robin gives them separate
    // hardware queues
    if (hipStreamCreateWithFlags(&s->fstream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&s->pstream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&s->cstream, hipStreamNonBlocking) != hipSuccess)
        rc = ccs::set_error(CCS_ERR_HIP, "server streams");
    for (int li = 0; li < kLanes && !rc; ++li) {
        if (!rc && s->resident) {
            if (hipHostMalloc((void**)&s->h_init[li], (size_t)s->n_slots * sizeof(ccs::PhaseInit)) != hipSuccess)
                rc = ccs::set_error(CCS_ERR_NOMEM, "phase init staging");
            if (!rc) rc = ccs_dmalloc(ctx, (void**)&s->d_init[li], (uint64_t)s->n_slots * sizeof(ccs::PhaseInit));
        }
        s->lanes[li].resize(s->lg_profiles.size());
        for (uint32_t gi = 0; gi < s->lanes[li].size() && !rc; ++gi) {
            Lane& L = s->lanes[li][gi];
            // a group set over the profiles of this launch group
            const uint32_t count = (uint32_t)s->lg_shapes[gi].size();
            std::vector<ccs_job> dummy;
            std::vector<uint32_t> dshape;
            for (uint32_t k = 0; k < count; ++k)
                for (uint64_t left = slice; left > 0;) {
                    const uint64_t c = std::min<uint64_t>(left, s->p.max_chunk);
                    dummy.push_back(dummy_job(c));
                    dshape.push_back(k);
                    left -= c;
                }
            // every active job a piece, plus the extra pieces of shares above max_chunk (at most
            // slice / max_chunk + 1 over the whole slice)
            while (dummy.size() < s->p.max_active + slice / s->p.max_chunk + 1) {
                dummy.push_back(dummy_job(per));
                dshape.push_back(count - 1);
            }
            rc = ccs_groupset_create(ctx, s->lg_shapes[gi].data(), count, dummy.data(), dshape.data(),
                                     (uint32_t)dummy.size(), &L.js);
            if (!rc) rc = ccs_dmalloc(ctx, (void**)&L.d_surv, (uint64_t)s->p.max_slice_survivors * sizeof(ccs_job_survivor));
            if (!rc) rc = ccs_dmalloc(ctx, (void**)&L.d_res, (uint64_t)s->p.max_slice_survivors * sizeof(ccs_chain_result));
            if (!rc) rc = ccs_dmalloc(ctx, (void**)&L.d_count, 64);
            if (!rc && hipHostMalloc((void**)&L.h_surv, (size_t)s->p.max_slice_survivors * sizeof(ccs_job_survivor)) != hipSuccess) rc = CCS_ERR_NOMEM;
            if (!rc && hipHostMalloc((void**)&L.h_res, (size_t)s->p.max_slice_survivors * sizeof(ccs_chain_result)) != hipSuccess) rc = CCS_ERR_NOMEM;
            if (!rc && hipHostMalloc((void**)&L.h_count, 64) != hipSuccess) rc = CCS_ERR_NOMEM;
            if (!rc && (hipEventCreateWithFlags(&L.filtered, hipEventDisableTiming) != hipSuccess ||
                        hipEventCreateWithFlags(&L.done, hipEventDisableTiming) != hipSuccess))
                rc = CCS_ERR_HIP;
            s->fixed_bytes += (uint64_t)s->p.max_slice_survivors * (sizeof(ccs_job_survivor) + sizeof(ccs_chain_result));
        }
    }
    if (rc) {
        destroy(s);
        return rc;
    }
    s->th = std::thread(server_loop, s);
    *out = s;
    return CCS_OK;
}

extern "C" int ccs_server_submit(ccs_server* s, const ccs_request* req, uint64_t* job_id) {
    if (!s || !req) return ccs::set_error(CCS_ERR_INVALID, "null argument");
    int group = -1;
    for (size_t i = 0; i < s->depths.size(); ++i)
        if (s->depths[i] == req->depth) group = (int)i;
    if (group < 0) return ccs::set_error(CCS_ERR_INVALID, "no shape of the request's depth");
    if (req->count >= kMaxRequest) return ccs::set_error(CCS_ERR_INVALID, "request count must be below 2^48");
    ccs_shape_info info;
    if (ccs_shape_query(s->shapes[group], &info) != CCS_OK) return CCS_ERR_INVALID;
    if (info.W_hi || req->a >= info.W) return ccs::set_error(CCS_ERR_INVALID, "request residue a must be < W");
    std::lock_guard<std::mutex> g(s->mu);
    if (s->stop) return s->error < 0 ? s->error : ccs::set_error(CCS_ERR_BUSY, "server is shut down");
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

// CCS_OVERFLOW when a returned job finished incomplete, < 0 when the server failed
extern "C" int ccs_server_poll(ccs_server* s, ccs_job_done* out, uint32_t cap, uint32_t* n) {
    if (!s || !n || (cap && !out)) return ccs::set_error(CCS_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> g(s->mu);
    const uint32_t k = std::min<uint32_t>(cap, (uint32_t)s->finished.size());
    bool incomplete = false;
    for (uint32_t i = 0; i < k; ++i) {
        out[i] = s->finished[i];
        incomplete |= (out[i].flags & CCS_JOB_INCOMPLETE) != 0;
    }
    s->finished.erase(s->finished.begin(), s->finished.begin() + k);
    *n = k;
    if (s->error < 0) return ccs::set_error(s->error, s->error_msg.c_str());
    return incomplete ? CCS_OVERFLOW : CCS_OK;
}

extern "C" int ccs_server_job_results(ccs_server* s, uint64_t job_id, uint64_t* u_off, ccs_chain_result* res,
                                      uint32_t cap, uint32_t* n) {
    if (!s || !n) return ccs::set_error(CCS_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->jobs.find(job_id);
    if (it == s->jobs.end() || !it->second.done) return ccs::set_error(CCS_ERR_INVALID, "no finished job of that id");
    const Job& j = it->second;
    const uint32_t k = std::min<uint32_t>(cap, (uint32_t)j.u_off.size());
    for (uint32_t i = 0; i < k; ++i) {
        if (u_off) u_off[i] = j.u_off[i];
        if (res) res[i] = j.res[i];
    }
    *n = (uint32_t)j.u_off.size();
    return k < j.u_off.size() ? CCS_OVERFLOW : CCS_OK;
}

// CCS_OVERFLOW when some job so far finished incomplete, < 0 when the server failed
extern "C" int ccs_server_drain(ccs_server* s) {
    if (!s) return ccs::set_error(CCS_ERR_INVALID, "null argument");
    std::unique_lock<std::mutex> g(s->mu);
    s->cv.wait(g, [&] {
        if (s->stop) return true;
        if (!s->pending.empty()) return false;
        for (uint64_t id : s->active) if (!s->jobs[id].done) return false;
        return true;
    });
    if (s->error < 0) return ccs::set_error(s->error, s->error_msg.c_str());
    return s->n_incomplete ? CCS_OVERFLOW : CCS_OK;
}

extern "C" int ccs_server_stats(ccs_server* s, ccs_server_counters* out) {
    if (!s || !out) return ccs::set_error(CCS_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> g(s->mu);
    std::memset(out, 0, sizeof(*out));
    out->candidates = s->stat_candidates;
    out->slices = s->stat_slices;
    out->launches = s->stat_launches;
    out->active = (uint32_t)s->active.size();
    out->pending = (uint32_t)s->pending.size();
    out->resident_bytes = s->resident_bytes;
    out->job_bytes = s->job_bytes;
    out->fixed_bytes = s->fixed_bytes;
    out->slots = s->n_slots;
    out->incomplete = s->n_incomplete;
    out->busy_s = s->busy_s;
    return CCS_OK;
}

extern "C" int ccs_server_shutdown(ccs_server* s) {
    if (!s) return ccs::set_error(CCS_ERR_INVALID, "null argument");
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stop = true;
        s->cv.notify_all();
    }
    if (s->th.joinable()) s->th.join();
    const int err = s->error;
    const std::string msg = s->error_msg;
    destroy(s);
    return err < 0 ? ccs::set_error(err, msg.c_str()) : CCS_OK;
}
