// tos_server.cpp -- a served stream of xi jobs with continuous batching on
// one device.
//
// Life of a job:
//   submit     inputs copied; queued for the candidate search
//   search     a host worker enumerates the candidate sightline pairs (the
//              skybin_xi_task search, XiEstimator.cc:121-159); this fixes
//              the job's exact HBM footprint.  With search_on_device the
//              worker only prepares the sky index; the scheduler runs the
//              device count pass (the pair count, hence the footprint) and,
//              once the job is admitted, the fill pass straight into the
//              job's resident pair buffers.  A plate test that rounding
//              leaves undecided sends the job back for the host search.
//   admit      the scheduler admits jobs in arrival order while the resident
//              bytes stay within the budget and fewer than max_inflight jobs
//              are resident; their pixels and pairs are uploaded and stay
//              resident until the job is done
//   launch     a job's work items are cut into slices of bounded estimated
//              device time (kSliceWork, about 20 ms of the whole device);
//              whenever a launch slot (stream) is free, one batched launch
//              of the pair kernel takes the next slice of the resident jobs
//              with items left, least remaining work first, up to
//              kLaunchWork in all -- and, when the first of them has little
//              work left, up to kLeadFactor times that work, so that a
//              small job's launch ends soon after it.  Every launch is
//              short, so a job that arrives while a large one runs starts
//              within a launch or two and runs beside it, and the large
//              job's slices keep filling the device between the others.
//   done       a slot's completion event has fired; the slices' histograms
//              are added into their jobs' host results (counts exactly,
//              sums in FP64 in slice order); a job whose items are all
//              launched and whose slices are all back is done, and its
//              device buffers are queued for release
// The scheduler thread is the only thread that uses the context's HIP
// resources; tos_job_* calls only touch the job table under its mutex.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <thread>

#include "tos_batch.h"

using namespace tos;

namespace tos {
namespace {

constexpr int kSlots = 3;          // concurrent batched launches
constexpr int kMaxBatch = 64;      // job slices per launch
// Work of one slice of a job (item cost units, wavefront steps of 64 pixel
// pairs, FP64 weighted: about 20 ms of the whole device at the full-survey
// FP32C rate of 3.7e10 steps/s), and of one launch.  With fewer resident
// jobs than kLaunchWork / kSliceWork, each job's slice grows to its share
// of kLaunchWork, so that a lone large job is not cut finer than needed.
constexpr double kSliceWork = 7.5e8;
constexpr double kLaunchWork = 2.25e9;
// A launch led by a job with W work left holds at most kLeadFactor W (and
// at least kMinLaunchWork, about 2 ms of the device) in all.
constexpr double kLeadFactor = 4.0;
constexpr double kMinLaunchWork = 7.5e7;
// Relative device time of a pixel pair in FP64 against FP32C (passes over
// 160,000 quasars and 10,000 deg^2: 575 ms against 354 ms).
constexpr double kFp64Cost = 1.6;

enum JobState { J_QUEUED, J_SEARCH, J_PREPARED, J_SEARCHED, J_RESIDENT, J_RUNNING, J_DONE, J_FAILED, J_CANCELLED };

struct ServerJob {
    tos_job_id id = 0;
    JobState state = J_QUEUED;
    int status = TOS_ERR_NOT_READY;
    std::string err;
    bool cancel = false;
    // threads blocked in tos_job_wait on this job; a job released by
    // tos_job_result while some wait is deleted by the last of them
    int waiters = 0;
    bool released = false;
    // copied inputs
    std::vector<int64_t> offsets;
    std::vector<double> ra, dec;
    std::vector<float> value, loglam, weight, distance;
    std::vector<int32_t> plate_of_los, plate_id;
    std::vector<double> plate_ra, plate_dec;
    tos_skybins sky{};
    tos_grid grid{};
    double scale = 0;
    int32_t precision = TOS_FP64;
    // candidate pairs and device residency
    HostPairs hp;
    std::unique_ptr<DeviceSearch> ds;   // device search, between prepare and fill
    bool host_search = false;           // the device search left a plate test undecided
    tos_catalog *cat = nullptr;
    tos_pairs *pairs = nullptr;
    int64_t footprint = 0;
    double work_scale = 1;   // device work per item cost unit (FP64 weighted)
    int32_t next_item = 0;   // first work item not yet launched
    double work_left = 0;    // estimated work of the items not yet launched
    int32_t in_flight = 0;   // slices launched and not yet back
    bool failed = false;     // a slice failed: launch no more, fail when all are back
    // results
    std::vector<uint64_t> cnt;
    std::vector<double> didj, wgt;
    uint64_t used = 0, rechecks = 0, evaluated = 0;
    double kernel_ms = 0;
    tos_job_timing timing{};

    tos_sightlines sightlines() const
    {
        tos_sightlines s{};
        s.n_los = (int64_t)ra.size();
        s.offsets = offsets.data();
        s.ra = ra.data();
        s.dec = dec.data();
        s.value = value.data();
        s.loglam = loglam.data();
        s.weight = weight.data();
        s.distance = distance.data();
        return s;
    }
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    BatchState B;
    std::vector<ServerJob *> jobs;      // the job of each slice in flight
    std::vector<ItemSlice> slices;      // its items
    bool busy = false;
    // buffers the slot outgrew, freed with the graveyard (scheduler thread)
    std::vector<void *> dead_dev, dead_host;
};

int64_t bins_of(const tos_grid &g) { return (int64_t)g.nbins[0] * g.nbins[1] * g.nbins[2]; }

// Device bytes a job keeps resident: pixels (16 B) and sightlines (16 B),
// candidate pairs (partner 4 B + cos 8 B), work items (16 B, at most one
// per 64 primary pixels and partner row plus one per pair step, and the
// 8 B prefix of their costs), and its share of the histogram partials
// (bounded by 64 workgroups' rows); plus the temporaries of the work item
// build (an 8 B cost per item, at most 24 B per row), freed later with the
// job's buffers.
int64_t footprint_of(int64_t n_los, int64_t n_pix, int64_t n_pairs, const tos_grid &g)
{
    const int64_t items = n_pix / 64 + 2 * n_los + n_pairs / 64 + 1;
    return n_pix * 16 + n_los * (16 + 24) + n_pairs * 12 + items * (16 + 8 + 8) + 64 * bins_of(g) * 20;
}

double now_ms(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace
}  // namespace tos

struct tos_server {
    tos_ctx *ctx = nullptr;
    tos_server_opts opts{};
    std::mutex mu;
    std::condition_variable cv_work, cv_sched, cv_done;
    std::deque<ServerJob *> search_q, count_q, admit_q;
    std::map<tos_job_id, ServerJob *> jobs;
    int64_t resident_bytes = 0;      // admitted jobs + buffers awaiting release
    int resident_jobs = 0;
    bool stopping = false;
    int waiting = 0;                 // threads inside tos_job_wait
    std::condition_variable cv_left; // a waiter left tos_job_wait
    tos_job_id next_id = 1;
    std::vector<std::thread> workers;
    std::thread sched;
    Slot slots[kSlots];
    std::chrono::steady_clock::time_point t0;
    // device buffers of finished jobs, freed when no batch is running
    std::vector<void *> graveyard;
    int64_t graveyard_bytes = 0;
};

namespace tos {
namespace {

void search_worker(tos_server *S, int nthreads)
{
    for (;;) {
        ServerJob *J = nullptr;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            S->cv_work.wait(lk, [&] { return S->stopping || !S->search_q.empty(); });
            if (S->search_q.empty()) return;   // stopping
            J = S->search_q.front();
            S->search_q.pop_front();
            if (J->cancel) {
                J->state = J_CANCELLED;
                J->status = TOS_ERR_CANCELLED;
                S->cv_done.notify_all();
                continue;
            }
            J->state = J_SEARCH;
        }
        const tos_sightlines s = J->sightlines();
        tos_skybins sky = J->sky;
        sky.plate_of_los = J->plate_of_los.empty() ? nullptr : J->plate_of_los.data();
        sky.plate_id = J->plate_id.empty() ? nullptr : J->plate_id.data();
        sky.plate_ra = J->plate_ra.empty() ? nullptr : J->plate_ra.data();
        sky.plate_dec = J->plate_dec.empty() ? nullptr : J->plate_dec.data();
        const double max_ang = tos_grid_max_angular_scale(&J->grid, J->scale);
        std::string err;
        const bool on_device = S->opts.search_on_device && !J->host_search &&
                               (sky.kind == TOS_SKY_HEALPIX || sky.kind == TOS_SKY_PLATES);
        int rc;
        if (on_device) {
            J->ds.reset(new (std::nothrow) DeviceSearch());
            rc = J->ds ? J->ds->prepare(s, sky, max_ang, err) : TOS_ERR_NOMEM;
        } else {
            rc = build_pairs_host(s, sky, max_ang, nthreads, J->hp, err);
        }
        {
            std::lock_guard<std::mutex> lk(S->mu);
            if (rc != TOS_OK) {
                J->ds.reset();   // prepare allocates nothing on the device
                J->state = J_FAILED;
                J->status = rc;
                J->err = err;
                S->cv_done.notify_all();
                continue;
            }
            if (on_device) {
                J->state = J_PREPARED;
                S->count_q.push_back(J);
                S->cv_sched.notify_all();
                continue;
            }
            J->footprint = footprint_of(s.n_los, J->offsets.back(), (int64_t)J->hp.other.size(), J->grid);
            J->timing.footprint_bytes = J->footprint;
            J->timing.num_pixel_pairs = J->hp.stats.num_pixel_pairs;
            J->state = J_SEARCHED;
            S->admit_q.push_back(J);
        }
        S->cv_sched.notify_all();
    }
}

void release_graveyard(tos_server *S)
{
    std::vector<void *> g;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        g.swap(S->graveyard);
        S->resident_bytes -= S->graveyard_bytes;
        S->graveyard_bytes = 0;
    }
    for (Slot &sl : S->slots) {
        g.insert(g.end(), sl.dead_dev.begin(), sl.dead_dev.end());
        sl.dead_dev.clear();
        for (void *h : sl.dead_host) (void)hipHostFree(h);
        sl.dead_host.clear();
    }
    for (void *p : g) (void)hipFree(p);
    if (!g.empty()) S->cv_sched.notify_all();
}

void fail_job(tos_server *S, ServerJob *J, int status, const std::string &err)
{
    std::vector<void *> dev;
    if (J->pairs) pairs_detach(J->pairs, dev);
    if (J->cat) catalog_detach(J->cat, dev);
    if (J->ds) J->ds->detach(dev);
    J->ds.reset();
    J->pairs = nullptr;
    J->cat = nullptr;
    std::lock_guard<std::mutex> lk(S->mu);
    S->graveyard.insert(S->graveyard.end(), dev.begin(), dev.end());
    S->graveyard_bytes += J->footprint;
    S->resident_jobs -= 1;
    J->state = J_FAILED;
    J->status = status;
    J->err = err;
    J->timing.status = status;
    S->cv_done.notify_all();
}

// Hand device buffers to the graveyard: they are freed when no batch runs
// (hipFree waits for the whole device).  bytes: what of them the resident
// total already counts.
void bury(tos_server *S, std::vector<void *> &dev, int64_t bytes)
{
    std::lock_guard<std::mutex> lk(S->mu);
    S->graveyard.insert(S->graveyard.end(), dev.begin(), dev.end());
    S->graveyard_bytes += bytes;
    dev.clear();
}

// Count pass of a prepared device search: the exact pair count and so the
// footprint; the search tables stay on the device for the fill pass.
// Returns false when the job left the pipeline (failed, cancelled, or sent
// back for the host search).
bool count_on_device(tos_server *S, ServerJob *J)
{
    tos_ctx *ctx = S->ctx;
    std::vector<void *> dev;
    bool undecided = false, cancel;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        cancel = J->cancel;
    }
    int rc = TOS_OK;
    if (!cancel) rc = J->ds->count(ctx, J->hp, undecided);
    if (cancel || rc != TOS_OK || undecided) {
        J->ds->detach(dev);
        J->ds.reset();
        bury(S, dev, 0);
        std::lock_guard<std::mutex> lk(S->mu);
        // a server that is stopping has no search workers left: cancel
        if (cancel || J->cancel || (undecided && S->stopping)) {
            J->state = J_CANCELLED;
            J->status = TOS_ERR_CANCELLED;
        } else if (rc != TOS_OK) {
            J->state = J_FAILED;
            J->status = rc;
            J->err = ctx->err;
            J->timing.status = rc;
        } else {
            J->host_search = true;
            J->hp = HostPairs();
            J->state = J_QUEUED;
            S->search_q.push_front(J);
            S->cv_work.notify_all();
            return false;
        }
        S->cv_done.notify_all();
        return false;
    }
    std::lock_guard<std::mutex> lk(S->mu);
    J->footprint = footprint_of((int64_t)J->ra.size(), J->offsets.back(), J->ds->total, J->grid) +
                   J->ds->device_bytes();
    J->timing.footprint_bytes = J->footprint;
    J->timing.num_pixel_pairs = J->hp.stats.num_pixel_pairs;
    J->state = J_SEARCHED;
    S->admit_q.push_back(J);
    return true;
}

// Upload an admitted job: catalogue, pairs (host search) or the fill pass
// into the job's pair buffers (device search), work items.
int make_resident(tos_server *S, ServerJob *J)
{
    tos_ctx *ctx = S->ctx;
    const tos_sightlines s = J->sightlines();
    int rc = tos_catalog_upload(ctx, &s, &J->cat);
    if (rc != TOS_OK) return rc;
    std::vector<void *> dev;
    if (J->ds) {
        tos_pairs *p = new (std::nothrow) tos_pairs();
        if (!p) return TOS_ERR_NOMEM;
        J->pairs = p;
        p->h = std::move(J->hp);
        p->cat = J->cat;
        rc = J->ds->fill(ctx, p->other, p->csep);
        p->pairs_on_host = false;
        p->n_rows = p->h.stats.n_rows;
        p->n_pairs = J->ds->total;
        J->ds->detach(dev);
        J->ds.reset();
    } else {
        rc = pairs_from_host(ctx, J->cat, std::move(J->hp), &J->pairs);
    }
    if (rc == TOS_OK) rc = ensure_items(ctx, J->pairs, true, &dev);
    if (rc == TOS_OK) {
        J->work_scale = J->precision == TOS_FP64 ? kFp64Cost : 1.0;
        J->next_item = 0;
        J->work_left = 0;
        for (uint64_t c : J->pairs->item_cost) J->work_left += J->work_scale * (double)c;
        J->in_flight = 0;
        const int64_t nb = bins_of(J->grid);
        J->cnt.assign((size_t)nb, 0);
        J->didj.assign((size_t)nb, 0.0);
        J->wgt.assign((size_t)nb, 0.0);
        J->used = J->rechecks = J->evaluated = 0;
        J->kernel_ms = 0;
    }
    if (rc == TOS_OK && J->grid.type == TOS_OBSERVING) rc = ensure_theta(ctx, J->pairs, J->grid);
    if (rc == TOS_OK) {
        const hipError_t e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) rc = hip_fail(ctx, e, "job upload");
    }
    // the temporaries are part of the footprint counted at admission
    bury(S, dev, 0);
    return rc;
}

// Retire a job whose items are all launched and whose slices are all
// back: done, cancelled or failed; its device buffers go to the graveyard.
void retire_job(tos_server *S, ServerJob *J, int status, const std::string &err)
{
    std::vector<void *> dev;
    if (J->pairs) pairs_detach(J->pairs, dev);
    if (J->cat) catalog_detach(J->cat, dev);
    J->pairs = nullptr;
    J->cat = nullptr;
    const double t = now_ms(S->t0);
    std::lock_guard<std::mutex> lk(S->mu);
    J->timing.done_ms = t;
    S->resident_jobs -= 1;
    S->graveyard_bytes += J->footprint;
    S->graveyard.insert(S->graveyard.end(), dev.begin(), dev.end());
    if (status != TOS_OK) {
        J->state = J_FAILED;
        J->status = status;
        J->err = err;
    } else if (J->cancel) {
        J->state = J_CANCELLED;
        J->status = TOS_ERR_CANCELLED;
    } else {
        J->state = J_DONE;
        J->status = TOS_OK;
    }
    J->timing.status = J->status;
    S->cv_done.notify_all();
}

bool all_launched(const ServerJob *J) { return J->pairs && J->next_item >= (int32_t)J->pairs->n_items; }

// A slot's launch has completed (or failed): add each slice's histogram
// into its job and retire the jobs that are complete.
void finish_slot(tos_server *S, Slot &sl, int status)
{
    const int32_t n = (int32_t)sl.jobs.size();
    float ms = 0;
    if (status == TOS_OK) hipEventElapsedTime(&ms, sl.ev0, sl.ev1);
    std::vector<tos_xi_result> out((size_t)n);
    std::vector<std::vector<uint64_t>> cnt((size_t)n);
    std::vector<std::vector<double>> didj((size_t)n), wgt((size_t)n);
    for (int32_t k = 0; k < n; ++k) {
        const int64_t nb = bins_of(sl.jobs[(size_t)k]->grid);
        cnt[(size_t)k].assign((size_t)nb, 0);
        didj[(size_t)k].assign((size_t)nb, 0.0);
        wgt[(size_t)k].assign((size_t)nb, 0.0);
        std::memset(&out[(size_t)k], 0, sizeof(tos_xi_result));
        out[(size_t)k].num_pairs = cnt[(size_t)k].data();
        out[(size_t)k].didj = didj[(size_t)k].data();
        out[(size_t)k].wgt = wgt[(size_t)k].data();
    }
    if (status == TOS_OK) batch_collect(sl.B, n, ms, out.data());
    const std::string err = status == TOS_OK ? std::string() : S->ctx->err;
    for (int32_t k = 0; k < n; ++k) {
        ServerJob *J = sl.jobs[(size_t)k];
        if (status != TOS_OK) {
            J->failed = true;
        } else {
            for (size_t b = 0; b < cnt[(size_t)k].size(); ++b) {
                J->cnt[b] += cnt[(size_t)k][b];
                J->didj[b] += didj[(size_t)k][b];
                J->wgt[b] += wgt[(size_t)k][b];
            }
            J->used += out[(size_t)k].num_pixel_pairs_used;
            J->rechecks += out[(size_t)k].num_guard_rechecks;
            J->evaluated += out[(size_t)k].num_pairs_evaluated;
            J->kernel_ms += ms;
        }
        J->in_flight -= 1;
        bool cancel;
        {
            std::lock_guard<std::mutex> lk(S->mu);
            cancel = J->cancel;
        }
        if (J->in_flight == 0 && (J->failed || cancel || all_launched(J)))
            retire_job(S, J, J->failed ? (status != TOS_OK ? status : TOS_ERR_DEVICE) : TOS_OK,
                       J->failed ? (err.empty() ? std::string("a slice of the job failed") : err) : std::string());
    }
    sl.jobs.clear();
    sl.slices.clear();
    sl.busy = false;
}

void scheduler(tos_server *S)
{
    tos_ctx *ctx = S->ctx;
    hipSetDevice(ctx->device);
    std::deque<ServerJob *> ready;
    for (;;) {
        bool progress = false;
        // 1. completions
        int busy = 0;
        for (Slot &sl : S->slots) {
            if (!sl.busy) continue;
            const hipError_t q = hipEventQuery(sl.ev1);
            if (q == hipSuccess) {
                finish_slot(S, sl, TOS_OK);
                progress = true;
            } else if (q != hipErrorNotReady) {
                hip_fail(ctx, q, "served batch");
                finish_slot(S, sl, TOS_ERR_DEVICE);
                progress = true;
            } else {
                ++busy;
            }
        }
        // 2. count passes of device searches (their footprints are then
        // known); free finished jobs' buffers when the device is idle
        // (hipFree synchronises), or when admission waits for memory
        for (;;) {
            ServerJob *J = nullptr;
            {
                std::lock_guard<std::mutex> lk(S->mu);
                if (S->count_q.empty()) break;
                J = S->count_q.front();
                S->count_q.pop_front();
            }
            count_on_device(S, J);
            progress = true;
        }
        bool mem_blocked = false;
        std::vector<ServerJob *> admitted;
        {
            std::lock_guard<std::mutex> lk(S->mu);
            while (!S->admit_q.empty()) {
                ServerJob *J = S->admit_q.front();
                if (J->cancel) {
                    S->admit_q.pop_front();
                    if (J->ds) J->ds->detach(S->graveyard);
                    J->ds.reset();
                    J->state = J_CANCELLED;
                    J->status = TOS_ERR_CANCELLED;
                    S->cv_done.notify_all();
                    continue;
                }
                if (J->footprint > S->opts.hbm_budget_bytes) {
                    S->admit_q.pop_front();
                    if (J->ds) J->ds->detach(S->graveyard);
                    J->ds.reset();
                    J->state = J_FAILED;
                    J->status = TOS_ERR_NOMEM;
                    J->err = "job footprint exceeds the server's HBM budget";
                    J->timing.status = TOS_ERR_NOMEM;
                    S->cv_done.notify_all();
                    continue;
                }
                if (S->resident_jobs >= S->opts.max_inflight) break;
                if (S->resident_bytes + J->footprint > S->opts.hbm_budget_bytes) {
                    mem_blocked = true;
                    break;
                }
                S->admit_q.pop_front();
                S->resident_bytes += J->footprint;
                S->resident_jobs += 1;
                J->timing.admit_ms = now_ms(S->t0);
                admitted.push_back(J);
            }
        }
        bool outgrown = false;
        for (Slot &sl : S->slots) outgrown |= !sl.dead_dev.empty() || !sl.dead_host.empty();
        if ((!S->graveyard.empty() || outgrown) && (busy == 0 || mem_blocked)) {
            release_graveyard(S);
            progress = true;
        }
        // 3. make admitted jobs resident
        for (ServerJob *J : admitted) {
            const int rc = make_resident(S, J);
            if (rc != TOS_OK) {
                fail_job(S, J, rc, ctx->err);
            } else {
                std::lock_guard<std::mutex> lk(S->mu);
                J->state = J_RESIDENT;
                ready.push_back(J);
            }
            progress = true;
        }
        // 4. on each free slot, one launch of the next slices of the
        // resident jobs with items left, least remaining work first (jobs
        // with no pair at all are done at once); a job leaves `ready` once
        // all its items are launched
        for (auto it = ready.begin(); it != ready.end();) {
            ServerJob *J = *it;
            bool cancel;
            {
                std::lock_guard<std::mutex> lk(S->mu);
                cancel = J->cancel;
            }
            if (J->pairs && J->pairs->n_items == 0 && J->in_flight == 0) {
                it = ready.erase(it);
                retire_job(S, J, TOS_OK, std::string());
                progress = true;
            } else if ((cancel || J->failed) && J->in_flight == 0) {
                it = ready.erase(it);
                retire_job(S, J, J->failed ? TOS_ERR_DEVICE : TOS_OK, "a slice of the job failed");
                progress = true;
            } else if (cancel || J->failed || all_launched(J)) {
                it = ready.erase(it);   // retired when its last slice is back
            } else {
                ++it;
            }
        }
        for (Slot &sl : S->slots) {
            if (ready.empty()) break;
            if (sl.busy) continue;
            std::vector<ServerJob *> order(ready.begin(), ready.end());
            std::stable_sort(order.begin(), order.end(),
                             [](const ServerJob *x, const ServerJob *y) { return x->work_left < y->work_left; });
            const double cap = std::min(kLaunchWork, std::max(kMinLaunchWork, kLeadFactor * order.front()->work_left));
            // each job's slice: its share of the launch, at least kSliceWork
            // (or the whole launch when that is smaller)
            const double target = std::max(cap / (double)order.size(), std::min(kSliceWork, cap));
            double work = 0;
            const double t_launch = now_ms(S->t0);
            for (ServerJob *J : order) {
                if ((int)sl.jobs.size() >= kMaxBatch || work >= cap) break;
                const double budget = std::min(target, cap - work);
                const std::vector<uint64_t> &cost = J->pairs->item_cost;
                const int32_t n_items = (int32_t)J->pairs->n_items;
                ItemSlice slc{J->next_item, J->next_item};
                double w = 0;
                while (slc.end < n_items && (slc.end == slc.begin || w < budget)) {
                    w += J->work_scale * (double)cost[(size_t)slc.end];
                    ++slc.end;
                }
                sl.jobs.push_back(J);
                sl.slices.push_back(slc);
                work += w;
                J->work_left -= w;
                J->next_item = slc.end;
                J->in_flight += 1;
                if (J->state != J_RUNNING) {
                    std::lock_guard<std::mutex> lk(S->mu);
                    J->state = J_RUNNING;
                    J->timing.start_ms = t_launch;
                }
            }
            ready.erase(std::remove_if(ready.begin(), ready.end(), [](const ServerJob *J) { return all_launched(J); }),
                        ready.end());
            if (sl.jobs.empty()) break;
            std::vector<tos_batch_job> bj;
            for (ServerJob *J : sl.jobs) {
                tos_batch_job b{};
                b.cat = J->cat;
                b.pairs = J->pairs;
                b.grid = J->grid;
                b.precision = J->precision;
                b.accum = TOS_ACC_CORE;
                bj.push_back(b);
            }
            int rc = batch_prepare(ctx, (int32_t)bj.size(), bj.data(), sl.B, sl.slices.data());
            if (rc == TOS_OK) {
                hipEventRecord(sl.ev0, sl.stream);
                rc = batch_enqueue(ctx, sl.B, (int32_t)bj.size(), sl.stream);
                hipEventRecord(sl.ev1, sl.stream);
            }
            sl.busy = true;
            if (rc != TOS_OK) finish_slot(S, sl, rc);
            progress = true;
        }
        // 5. sleep until something changes
        std::unique_lock<std::mutex> lk(S->mu);
        bool any_busy = false;
        for (Slot &sl : S->slots) any_busy |= sl.busy;
        if (S->stopping && S->search_q.empty() && S->count_q.empty() && S->admit_q.empty() && ready.empty() &&
            !any_busy) {
            bool searching = false;
            for (auto &kv : S->jobs) searching |= kv.second->state == J_SEARCH;
            if (!searching) break;
        }
        if (!progress && S->count_q.empty())
            S->cv_sched.wait_for(lk, std::chrono::microseconds(any_busy ? 200 : 2000));
    }
    release_graveyard(S);
}

ServerJob *find_job(tos_server *S, tos_job_id id)
{
    auto it = S->jobs.find(id);
    return it == S->jobs.end() ? nullptr : it->second;
}

bool finished(const ServerJob *J) { return J->state == J_DONE || J->state == J_FAILED || J->state == J_CANCELLED; }

}  // namespace
}  // namespace tos

extern "C" int tos_server_start(tos_ctx *ctx, const tos_server_opts *opts, tos_server **out)
{
    if (!ctx || !opts || !out) return TOS_ERR_INVALID;
    *out = nullptr;
    if (opts->hbm_budget_bytes <= 0 || opts->max_inflight <= 0) {
        set_err(ctx, "tos_server_start: hbm_budget_bytes and max_inflight must be positive");
        return TOS_ERR_INVALID;
    }
    tos_server *S = new (std::nothrow) tos_server();
    if (!S) return TOS_ERR_NOMEM;
    S->ctx = ctx;
    S->opts = *opts;
    S->t0 = std::chrono::steady_clock::now();
    hipSetDevice(ctx->device);
    for (Slot &sl : S->slots) {
        sl.B.defer_dev = &sl.dead_dev;
        sl.B.defer_host = &sl.dead_host;
        if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreate(&sl.ev0) != hipSuccess || hipEventCreate(&sl.ev1) != hipSuccess) {
            set_err(ctx, "tos_server_start: stream or event creation failed");
            for (Slot &t : S->slots) {
                if (t.stream) (void)hipStreamDestroy(t.stream);
                if (t.ev0) (void)hipEventDestroy(t.ev0);
                if (t.ev1) (void)hipEventDestroy(t.ev1);
            }
            delete S;
            return TOS_ERR_DEVICE;
        }
    }
    // two search workers when there are threads to spare; each searches with
    // all the threads, so one large job alone (the usual latency tail) gets
    // the whole host, and two at once share it
    const int nthreads = std::max(1, opts->nthreads);
    const int nworkers = nthreads >= 4 ? 2 : 1;
    for (int w = 0; w < nworkers; ++w) S->workers.emplace_back(search_worker, S, nthreads);
    S->sched = std::thread(scheduler, S);
    *out = S;
    return TOS_OK;
}

extern "C" int tos_server_stop(tos_server *S)
{
    if (!S) return TOS_ERR_INVALID;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stopping = true;
        // jobs not yet searched are cancelled; admitted ones finish
        for (ServerJob *J : S->search_q) {
            J->state = J_CANCELLED;
            J->status = TOS_ERR_CANCELLED;
        }
        S->search_q.clear();
        for (std::deque<ServerJob *> *q : {&S->count_q, &S->admit_q}) {
            for (ServerJob *J : *q) {
                if (J->ds) J->ds->detach(S->graveyard);
                J->ds.reset();
                J->state = J_CANCELLED;
                J->status = TOS_ERR_CANCELLED;
            }
            q->clear();
        }
    }
    S->cv_work.notify_all();
    S->cv_sched.notify_all();
    for (std::thread &t : S->workers) t.join();
    S->sched.join();
    {
        // every job is finished now; wake the waiters and let them leave
        // before the job table goes
        std::unique_lock<std::mutex> lk(S->mu);
        S->cv_done.notify_all();
        S->cv_left.wait(lk, [&] { return S->waiting == 0; });
    }
    for (Slot &sl : S->slots) {
        (void)hipStreamSynchronize(sl.stream);
        sl.B.drop_graph();
        for (DevBuf *d : {&sl.B.args, &sl.B.tasks, &sl.B.red, &sl.B.parts_cnt, &sl.B.parts_sum, &sl.B.res})
            d->release();
        if (sl.B.host_res) (void)hipHostFree(sl.B.host_res);
        (void)hipStreamDestroy(sl.stream);
        (void)hipEventDestroy(sl.ev0);
        (void)hipEventDestroy(sl.ev1);
    }
    for (auto &kv : S->jobs) delete kv.second;   // no waiter is left (see above)
    delete S;
    return TOS_OK;
}

extern "C" int tos_job_footprint(const tos_job_desc *job, int64_t *bytes)
{
    if (!job || !bytes || job->sightlines.n_los < 0 || (job->sightlines.n_los > 0 && !job->sightlines.offsets))
        return TOS_ERR_INVALID;
    const int64_t n = job->sightlines.n_los;
    const int64_t npix = n > 0 ? job->sightlines.offsets[n] : 0;
    // the candidate pairs are only known after the search: this is the
    // footprint without them (the admitted figure is in tos_job_timing)
    *bytes = footprint_of(n, npix, 0, job->grid);
    return TOS_OK;
}

extern "C" int tos_job_submit(tos_server *S, const tos_job_desc *d, tos_job_id *id)
{
    if (!S || !d || !id) return TOS_ERR_INVALID;
    const tos_sightlines &s = d->sightlines;
    if (s.n_los < 0 || (s.n_los > 0 && (!s.offsets || !s.ra || !s.dec || !s.value || !s.loglam || !s.weight ||
                                       !s.distance)))
        return TOS_ERR_INVALID;
    if (!grid_ok(S->ctx, &d->grid)) return TOS_ERR_INVALID;
    if (d->precision != TOS_FP64 && d->precision != TOS_FP32C) return TOS_ERR_INVALID;
    if (d->sky.kind == TOS_SKY_PLATES &&
        (!d->sky.plate_of_los || d->sky.n_plates < 0 ||
         (d->sky.n_plates > 0 && (!d->sky.plate_id || !d->sky.plate_ra || !d->sky.plate_dec)))) {
        set_err(S->ctx, "tos_job_submit: plate sky bins need plate_of_los and, for n_plates > 0, the plate arrays");
        return TOS_ERR_INVALID;
    }
    {
        // served jobs run in batched launches, whose LDS histogram holds the
        // whole grid: a grid too large for it is refused here rather than
        // failing the batch it would join
        GridExact ge;
        grid_exact_consts(d->grid, ge);
        if ((int64_t)xi_lds_bytes(TOS_ACC_CORE, (int)ge.nslots) + 64 > S->ctx->max_lds_block) {
            set_err(S->ctx, "tos_job_submit: grid has too many bins for a served job's LDS histogram");
            return TOS_ERR_INVALID;
        }
    }
    ServerJob *J = new (std::nothrow) ServerJob();
    if (!J) return TOS_ERR_NOMEM;
    const int64_t n = s.n_los;
    const int64_t npix = n > 0 ? s.offsets[n] : 0;
    if (npix < 0) {
        delete J;
        return TOS_ERR_INVALID;
    }
    J->offsets.assign(s.offsets, s.offsets + (n > 0 ? n + 1 : 0));
    if (n == 0) J->offsets.assign(1, 0);
    J->ra.assign(s.ra, s.ra + n);
    J->dec.assign(s.dec, s.dec + n);
    J->value.assign(s.value, s.value + npix);
    J->loglam.assign(s.loglam, s.loglam + npix);
    J->weight.assign(s.weight, s.weight + npix);
    J->distance.assign(s.distance, s.distance + npix);
    J->sky = d->sky;
    if (d->sky.kind == TOS_SKY_PLATES) {
        J->plate_of_los.assign(d->sky.plate_of_los, d->sky.plate_of_los + n);
        const int64_t np = d->sky.n_plates;
        J->plate_id.assign(d->sky.plate_id, d->sky.plate_id + np);
        J->plate_ra.assign(d->sky.plate_ra, d->sky.plate_ra + np);
        J->plate_dec.assign(d->sky.plate_dec, d->sky.plate_dec + np);
    }
    J->grid = d->grid;
    J->scale = d->scale;
    J->precision = d->precision;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        if (S->stopping) {
            delete J;
            return TOS_ERR_CANCELLED;
        }
        J->id = S->next_id++;
        J->timing.submit_ms = now_ms(S->t0);
        J->timing.status = TOS_ERR_NOT_READY;
        S->jobs[J->id] = J;
        S->search_q.push_back(J);
        *id = J->id;
    }
    S->cv_work.notify_one();
    return TOS_OK;
}

extern "C" int tos_job_poll(tos_server *S, tos_job_id id)
{
    if (!S) return TOS_ERR_INVALID;
    std::lock_guard<std::mutex> lk(S->mu);
    ServerJob *J = find_job(S, id);
    if (!J) return TOS_ERR_INVALID;
    return finished(J) ? J->status : TOS_ERR_NOT_READY;
}

extern "C" int tos_job_wait(tos_server *S, tos_job_id id, double timeout_ms)
{
    if (!S) return TOS_ERR_INVALID;
    std::unique_lock<std::mutex> lk(S->mu);
    ServerJob *J = find_job(S, id);
    if (!J) return TOS_ERR_INVALID;
    // J stays valid while this thread is counted among its waiters (see
    // tos_job_result) and the server while it is counted in S->waiting
    ++J->waiters;
    ++S->waiting;
    auto done = [&] { return finished(J); };
    bool ok = true;
    if (timeout_ms < 0)
        S->cv_done.wait(lk, done);
    else
        ok = S->cv_done.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), done);
    const int status = ok ? J->status : TOS_ERR_TIMEOUT;
    --J->waiters;
    if (J->released && J->waiters == 0) delete J;
    if (--S->waiting == 0) S->cv_left.notify_all();
    return status;
}

extern "C" int tos_job_timing_get(tos_server *S, tos_job_id id, tos_job_timing *timing)
{
    if (!S || !timing) return TOS_ERR_INVALID;
    std::lock_guard<std::mutex> lk(S->mu);
    ServerJob *J = find_job(S, id);
    if (!J) return TOS_ERR_INVALID;
    *timing = J->timing;
    return TOS_OK;
}

extern "C" int tos_job_cancel(tos_server *S, tos_job_id id)
{
    if (!S) return TOS_ERR_INVALID;
    std::lock_guard<std::mutex> lk(S->mu);
    ServerJob *J = find_job(S, id);
    if (!J) return TOS_ERR_INVALID;
    if (!finished(J)) J->cancel = true;
    S->cv_sched.notify_all();
    return TOS_OK;
}

extern "C" int tos_job_result(tos_server *S, tos_job_id id, tos_xi_result *out, tos_job_timing *timing)
{
    if (!S) return TOS_ERR_INVALID;
    std::lock_guard<std::mutex> lk(S->mu);
    ServerJob *J = find_job(S, id);
    if (!J) return TOS_ERR_INVALID;
    if (!finished(J)) return TOS_ERR_NOT_READY;
    if (timing) *timing = J->timing;
    const int status = J->status;
    if (status == TOS_OK && out) {
        if (!out->num_pairs || !out->didj || !out->wgt) return TOS_ERR_INVALID;
        std::memcpy(out->num_pairs, J->cnt.data(), J->cnt.size() * sizeof(uint64_t));
        std::memcpy(out->didj, J->didj.data(), J->didj.size() * sizeof(double));
        std::memcpy(out->wgt, J->wgt.data(), J->wgt.size() * sizeof(double));
        out->num_pixel_pairs_used = J->used;
        out->num_guard_rechecks = J->rechecks;
        out->num_pairs_evaluated = J->evaluated;
        out->kernel_ms = J->kernel_ms;
    }
    S->jobs.erase(id);
    if (J->waiters > 0)
        J->released = true;   // the last waiter deletes it
    else
        delete J;
    return status;
}
