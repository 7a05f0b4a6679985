// tos_batch.cpp -- many estimator jobs (mock realisations, sub-catalogues)
// in one launch of the pair kernel, optionally captured in a hipGraph, and
// the launch-shape autotuner.
//
// A batch replaces calling XiEstimator::run (XiEstimator.cc:31-59) once per
// mock.  Every workgroup of the launch serves exactly one job: a job of
// pixel-pair cost C gets ceil(C / T) workgroups (T = total cost over
// kTasksPerSlot workgroups per resident slot, at most kMaxTasks in all, so
// no task costs more than T), inside a job the items are balanced as for a
// single run, and the tasks are issued longest first, so the hardware
// dispatcher, which starts workgroups in index order as slots free up, ends
// the launch on the smallest ones.  Each job keeps its own partial rows,
// counters and outputs, so the results are those of one tos_xi_run per
// job.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <vector>

#include "tos_batch.h"

using namespace tos;

namespace tos {

// Workgroups per segment (a job, or a sky bin's rows): ceil(cost / target)
// with target = total / budget, at least one for a non-empty segment, at
// most its item count.
static int seg_workgroups(double cost, double total, int64_t budget, int64_t n_items)
{
    if (n_items <= 0) return 0;
    const double target = total / (double)std::max<int64_t>(1, budget);
    int64_t n = target > 0 ? (int64_t)std::ceil(cost / target - 1e-9) : 1;
    return (int)std::max<int64_t>(1, std::min<int64_t>(n, n_items));
}

// Longest-processing-time order of tasks[begin, end): descending estimated
// cost, ties in the original order.  Partial rows travel with the task, so
// the order changes nothing in the results.
static void order_longest_first(std::vector<XiWgTask> &tasks, std::vector<uint64_t> &task_cost, size_t begin)
{
    const size_t n = tasks.size() - begin;
    std::vector<size_t> idx(n);
    for (size_t k = 0; k < n; ++k) idx[k] = k;
    std::stable_sort(idx.begin(), idx.end(),
                     [&](size_t a, size_t b) { return task_cost[begin + a] > task_cost[begin + b]; });
    std::vector<XiWgTask> t(n);
    std::vector<uint64_t> c(n);
    for (size_t k = 0; k < n; ++k) {
        t[k] = tasks[begin + idx[k]];
        c[k] = task_cost[begin + idx[k]];
    }
    std::copy(t.begin(), t.end(), tasks.begin() + (std::ptrdiff_t)begin);
    std::copy(c.begin(), c.end(), task_cost.begin() + (std::ptrdiff_t)begin);
}

static uint64_t range_cost(const std::vector<uint64_t> &item_cost, int32_t a, int32_t b)
{
    uint64_t c = 0;
    for (int32_t i = a; i < b; ++i) c += item_cost[(size_t)i];
    return c;
}

// Task budget of a batched launch: many more workgroups than resident
// slots, so that the last tasks to finish are short; capped because every
// task keeps a partial histogram row.
constexpr int64_t kTasksPerSlot = 32;
constexpr int64_t kMaxTasks = 16384;

static int64_t task_budget(const tos_ctx *ctx, int bpc, const TuneEntry *t)
{
    const int64_t per_slot = (t && t->wg_per_slot > 0) ? t->wg_per_slot : kTasksPerSlot;
    return std::min<int64_t>((int64_t)ctx->ncu * bpc * per_slot, kMaxTasks);
}

static void release_state(BatchState *&b)
{
    if (!b) return;
    b->drop_graph();
    for (DevBuf *d : {&b->args, &b->tasks, &b->red, &b->parts_cnt, &b->parts_sum, &b->res}) d->release();
    if (b->host_res) (void)hipHostFree(b->host_res);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    delete b;
    b = nullptr;
}

void batch_release(tos_ctx *ctx)
{
    release_state(ctx->batch);
    release_state(ctx->skybin);
}

namespace {

void put(std::vector<uint64_t> &k, const void *p) { k.push_back((uint64_t)(uintptr_t)p); }
void put(std::vector<uint64_t> &k, double v)
{
    uint64_t u;
    std::memcpy(&u, &v, 8);
    k.push_back(u);
}

std::vector<uint64_t> job_key(int32_t n, const tos_batch_job *jobs)
{
    std::vector<uint64_t> k;
    k.push_back((uint64_t)n);
    for (int32_t j = 0; j < n; ++j) {
        const tos_batch_job &J = jobs[j];
        k.push_back(J.cat->uid);
        k.push_back(J.pairs->uid);
        k.push_back(J.pairs->items_gen);
        k.push_back(((uint64_t)(uint32_t)J.precision << 32) | (uint32_t)J.grid.type);
        for (int a = 0; a < 3; ++a) {
            k.push_back((uint64_t)(uint32_t)J.grid.nbins[a]);
            put(k, J.grid.min[a]);
            put(k, J.grid.width[a]);
        }
    }
    return k;
}

// Per-job device buffers the prepared arguments depend on (a reallocation
// invalidates the prepared batch and its graph).
void add_buffer_key(std::vector<uint64_t> &k, int32_t n, const tos_batch_job *jobs, const tos_ctx *ctx)
{
    for (int32_t j = 0; j < n; ++j) {
        const tos_pairs *p = jobs[j].pairs;
        put(k, p->items.ptr);
        put(k, p->theta.ptr);
        put(k, p->other.ptr);
        put(k, p->csep.ptr);
        put(k, jobs[j].cat->pix.ptr);
    }
    for (const TuneEntry &t : ctx->tuned) k.push_back(((uint64_t)t.block << 32) | (uint32_t)t.blocks_per_cu);
}

}  // namespace

int batch_prepare(tos_ctx *ctx, int32_t n, const tos_batch_job *jobs, BatchState &B, const ItemSlice *slices)
{
    // group jobs by kernel variant, decided from each job's own constants
    std::map<VariantKey, std::vector<int>> by;
    B.nbins.assign((size_t)n, 0);
    std::vector<XiKernelArgs> args((size_t)n);
    for (int32_t j = 0; j < n; ++j) {
        fill_args(jobs[j].cat, jobs[j].pairs, jobs[j].grid, true, args[(size_t)j]);
        bool fixed, simple;
        xi_variant(jobs[j].grid.type, args[(size_t)j].ge, args[(size_t)j].gf, fixed, simple);
        by[VariantKey{jobs[j].grid.type, jobs[j].precision, fixed, simple}].push_back(j);
        B.nbins[(size_t)j] = jobs[j].grid.nbins[0] * jobs[j].grid.nbins[1] * jobs[j].grid.nbins[2];
    }
    std::vector<XiWgTask> tasks;
    std::vector<uint64_t> task_cost;
    std::vector<XiReduceJob> red((size_t)n);
    std::vector<int> job_nwg((size_t)n, 0);
    std::vector<std::vector<int32_t>> job_ranges((size_t)n);
    B.groups.clear();
    // partial-row offsets, output offsets
    size_t cnt_words = 0, sum_words = 0;
    B.out_off.assign((size_t)n, 0);
    B.res_words = 2 * (size_t)n;   // counters first: [rechecks, evaluated] per job
    B.max_fields = 0;
    for (auto &kv : by) {
        BatchGroup G;
        G.key = kv.first;
        G.jobs = kv.second;
        // launch shape: the largest job's LDS, the tuned or default block
        int max_slots = 0;
        for (int j : G.jobs) {
            GridExact ge;
            grid_exact_consts(jobs[j].grid, ge);
            max_slots = std::max(max_slots, (int)ge.nslots);
        }
        G.lds = xi_lds_bytes(TOS_ACC_CORE, max_slots);
        if ((int64_t)G.lds + 64 > ctx->max_lds_block) {
            set_err(ctx, "tos_batch_run: a grid has too many bins for the LDS histogram");
            return TOS_ERR_INVALID;
        }
        const TuneEntry *t = find_tuned(ctx, G.key.coord, G.key.prec, G.key.fixed, G.key.simple);
        G.block = t ? t->block : 512;
        int bpc = t ? t->blocks_per_cu : max_blocks_per_cu(ctx, G.lds, G.block);
        bpc = std::min(bpc, max_blocks_per_cu(ctx, G.lds, G.block));
        const int64_t budget = task_budget(ctx, bpc, t);
        double total = 0;
        std::vector<double> cost(G.jobs.size(), 0.0);
        auto slice_of = [&](int j) {
            return slices ? slices[j] : ItemSlice{0, (int32_t)jobs[j].pairs->n_items};
        };
        for (size_t q = 0; q < G.jobs.size(); ++q) {
            const tos_pairs *p = jobs[G.jobs[q]].pairs;
            const ItemSlice sl = slice_of(G.jobs[q]);
            cost[q] = (double)range_cost(p->item_cost, sl.begin, sl.end);
            total += cost[q];
        }
        G.task_begin = (int)tasks.size();
        for (size_t q = 0; q < G.jobs.size(); ++q) {
            const int j = G.jobs[q];
            const tos_pairs *p = jobs[j].pairs;
            const ItemSlice sl = slice_of(j);
            const int nwg = seg_workgroups(cost[q], total, budget, sl.end - sl.begin);
            job_nwg[(size_t)j] = nwg;
            if (nwg > 0) {
                if (sl.begin == 0 && sl.end == (int32_t)p->n_items) {
                    balance_items(p->item_cost, nwg, job_ranges[(size_t)j]);
                } else {
                    const std::vector<uint64_t> sub(p->item_cost.begin() + sl.begin, p->item_cost.begin() + sl.end);
                    balance_items(sub, nwg, job_ranges[(size_t)j]);
                    for (int32_t &x : job_ranges[(size_t)j]) x += sl.begin;
                }
            }
            for (int w = 0; w < nwg; ++w) {
                const int32_t a = job_ranges[(size_t)j][(size_t)w], b = job_ranges[(size_t)j][(size_t)w + 1];
                tasks.push_back(XiWgTask{j, a, b, w});
                task_cost.push_back(range_cost(p->item_cost, a, b));
            }
        }
        for (size_t t = (size_t)G.task_begin; t < tasks.size(); ++t)
            if (task_cost[t] >= 0xFFFFFFFFull) {
                // a task's u32 partial counts hold at most its items' cost
                set_err(ctx, "tos_batch_run: a workgroup task would count 2^32 pixel pairs; split the job");
                return TOS_ERR_INVALID;
            }
        order_longest_first(tasks, task_cost, (size_t)G.task_begin);
        G.n_wg = (int)tasks.size() - G.task_begin;
        B.groups.push_back(G);
    }
    for (int32_t j = 0; j < n; ++j) {
        const size_t nb = (size_t)B.nbins[(size_t)j];
        B.out_off[(size_t)j] = B.res_words;
        B.res_words += 3 * nb;   // counts, then nb x {didj, wgt}
        B.max_fields = std::max<int64_t>(B.max_fields, (int64_t)nb * 3);
        red[(size_t)j].n_wg = job_nwg[(size_t)j];
        red[(size_t)j].nbins = (int32_t)nb;
        red[(size_t)j].nsum = 2;
        // partial offsets recorded now, pointers filled after allocation
        red[(size_t)j].part_cnt = (const uint32_t *)(uintptr_t)cnt_words;
        red[(size_t)j].part_sum = (const double *)(uintptr_t)sum_words;
        cnt_words += (size_t)job_nwg[(size_t)j] * nb;
        sum_words += (size_t)job_nwg[(size_t)j] * nb * 2;
    }
    // grow a buffer to n bytes, handing the old one to B.defer_dev if set
    auto grow = [&B](DevBuf &b, size_t n) -> int {
        if (n <= b.bytes) return TOS_OK;
        if (B.defer_dev) {
            if (b.ptr) B.defer_dev->push_back(b.ptr);
            b.ptr = nullptr;
            b.bytes = 0;
            n += n / 2;
        }
        return b.reserve(n);
    };
    if (grow(B.parts_cnt, std::max<size_t>(4, cnt_words * 4)) != TOS_OK ||
        grow(B.parts_sum, std::max<size_t>(8, sum_words * 8)) != TOS_OK || grow(B.res, B.res_words * 8) != TOS_OK ||
        grow(B.args, args.size() * sizeof(XiKernelArgs)) != TOS_OK ||
        grow(B.tasks, tasks.size() * sizeof(XiWgTask)) != TOS_OK ||
        grow(B.red, red.size() * sizeof(XiReduceJob)) != TOS_OK) {
        set_err(ctx, "tos_batch_run: device allocation failed for the partials");
        return TOS_ERR_NOMEM;
    }
    if (B.host_res_bytes < B.res_words * 8) {
        size_t want = B.res_words * 8;
        if (B.host_res && B.defer_host) {
            B.defer_host->push_back(B.host_res);
            want += want / 2;
        } else if (B.host_res) {
            hipHostFree(B.host_res);
        }
        B.host_res = nullptr;
        B.host_res_bytes = 0;
        if (hipHostMalloc(&B.host_res, want, hipHostMallocDefault) != hipSuccess) {
            B.host_res = nullptr;
            set_err(ctx, "tos_batch_run: pinned host allocation failed");
            return TOS_ERR_NOMEM;
        }
        B.host_res_bytes = want;
    }
    uint64_t *res = (uint64_t *)B.res.ptr;
    for (int32_t j = 0; j < n; ++j) {
        XiKernelArgs &a = args[(size_t)j];
        const size_t co = (size_t)(uintptr_t)red[(size_t)j].part_cnt;
        const size_t so = (size_t)(uintptr_t)red[(size_t)j].part_sum;
        a.wg_item_ptr = nullptr;
        a.part_cnt = (uint32_t *)B.parts_cnt.ptr + co;
        a.part_sum = (double *)B.parts_sum.ptr + so;
        a.n_recheck = (unsigned long long *)(res + 2 * j);
        red[(size_t)j].part_cnt = a.part_cnt;
        red[(size_t)j].part_sum = a.part_sum;
        red[(size_t)j].out_cnt = res + B.out_off[(size_t)j];
        red[(size_t)j].out_sum = (double *)(res + B.out_off[(size_t)j] + B.nbins[(size_t)j]);
    }
    for (const BatchGroup &G : B.groups) {
        const int e = launch_xi_batch(G.key.coord, G.key.prec, G.key.fixed, G.key.simple, XiBatchArgs{}, 0, G.block,
                                      G.lds, ctx->stream);
        if (e != hipSuccess) return hip_fail(ctx, (hipError_t)e, "batched kernel LDS limit");
    }
    int rc;
    if ((rc = upload(ctx, B.args, args.data(), args.size() * sizeof(XiKernelArgs), "batch job table")) != TOS_OK ||
        (rc = upload(ctx, B.tasks, tasks.data(), tasks.size() * sizeof(XiWgTask), "batch workgroup tasks")) !=
            TOS_OK ||
        (rc = upload(ctx, B.red, red.data(), red.size() * sizeof(XiReduceJob), "batch reduction table")) != TOS_OK)
        return rc;
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "tos_batch_run upload");
    return TOS_OK;
}

// The device work of one batch, in stream order: zero the counters, one
// pair-kernel launch per variant group, the batched reduction, one copy of
// every result into pinned host memory.  No memset nodes (graph-safe).
int batch_enqueue(tos_ctx *ctx, BatchState &B, int32_t n, hipStream_t stream)
{
    int e = launch_zero_u64((unsigned long long *)B.res.ptr, 2 * (int64_t)n, stream);
    if (e != hipSuccess) return hip_fail(ctx, (hipError_t)e, "batch counter clear");
    for (const BatchGroup &G : B.groups) {
        if (G.n_wg == 0) continue;
        XiBatchArgs ba;
        ba.jobs = (const XiKernelArgs *)B.args.ptr;
        ba.wg = (const XiWgTask *)B.tasks.ptr + G.task_begin;
        e = launch_xi_batch(G.key.coord, G.key.prec, G.key.fixed, G.key.simple, ba, G.n_wg, G.block, G.lds,
                            stream);
        if (e != hipSuccess) return hip_fail(ctx, (hipError_t)e, "batched pair kernel launch");
    }
    e = launch_reduce_batch((const XiReduceJob *)B.red.ptr, n, B.max_fields, stream);
    if (e != hipSuccess) return hip_fail(ctx, (hipError_t)e, "batched reduction launch");
    hipError_t he = hipMemcpyAsync(B.host_res, B.res.ptr, B.res_words * 8, hipMemcpyDeviceToHost, stream);
    if (he != hipSuccess) return hip_fail(ctx, he, "batch result copy");
    return TOS_OK;
}

void batch_collect(const BatchState &B, int32_t n, float ms, tos_xi_result *out)
{
    const uint64_t *hr = (const uint64_t *)B.host_res;
    for (int32_t j = 0; j < n; ++j) {
        const int nb = B.nbins[(size_t)j];
        const uint64_t *cnt = hr + B.out_off[(size_t)j];
        const double *sum = (const double *)(cnt + nb);
        uint64_t used = 0;
        for (int b = 0; b < nb; ++b) {
            out[j].num_pairs[b] = cnt[b];
            out[j].didj[b] = sum[2 * b];
            out[j].wgt[b] = sum[2 * b + 1];
            used += cnt[b];
        }
        out[j].num_pixel_pairs_used = used;
        out[j].num_guard_rechecks = hr[2 * j];
        out[j].num_pairs_evaluated = hr[2 * j + 1];
        out[j].kernel_ms = ms;
    }
}

}  // namespace tos

namespace tos {

// Per-sky-bin subsamples of one pass (XiEstimator.h:78, skybin_xis_): the
// sky bins of the candidate rows are the jobs of one batched launch over a
// single catalogue and pair list.  The total is the sum of the subsamples in
// ascending sky-bin key order, the reference's merge (XiEstimator.cc:298-305).
int xi_run_per_skybin(tos_ctx *ctx, const tos_catalog *cat, tos_pairs *p, const tos_grid &g, int prec, int flags,
                      tos_xi_result *out)
{
    int rc;
    if ((rc = ensure_items(ctx, p, !(flags & TOS_XI_NO_TRANSPOSE), nullptr, tuned_step_cap(ctx, g, prec))) != TOS_OK)
        return rc;
    if (g.type == TOS_OBSERVING && (rc = ensure_theta(ctx, p, g)) != TOS_OK) return rc;
    const int64_t nr = p->n_rows;
    const int nbins = g.nbins[0] * g.nbins[1] * g.nbins[2];
    // sky-bin keys in ascending order, and the item segments of each
    // every occupied sky bin has a subsample, empty or not (the reference
    // stores one per task, XiEstimator.cc:155)
    std::vector<int32_t> keys = p->h.skybin_keys;
    keys.insert(keys.end(), p->h.row_skybin.begin(), p->h.row_skybin.end());
    std::sort(keys.begin(), keys.end());
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    if ((int64_t)keys.size() != p->h.stats.n_skybins) {
        set_err(ctx, "tos_xi_run: sky-bin keys of the rows disagree with the pair statistics");
        return TOS_ERR_INTERNAL;
    }
    const int nk = (int)keys.size();
    struct Seg {
        int key;
        int32_t a, b;
    };
    std::vector<Seg> segs;
    for (int64_t r = 0; r < nr; ++r) {
        const int k = (int)(std::lower_bound(keys.begin(), keys.end(), p->h.row_skybin[(size_t)r]) - keys.begin());
        const int32_t a = p->row_item_ptr[(size_t)r], b = p->row_item_ptr[(size_t)r + 1];
        if (a == b) continue;
        if (!segs.empty() && segs.back().key == k && segs.back().b == a) segs.back().b = b;
        else segs.push_back(Seg{k, a, b});
    }
    if (nk == 0) {
        std::memset(out->num_pairs, 0, sizeof(uint64_t) * nbins);
        std::memset(out->didj, 0, sizeof(double) * nbins);
        std::memset(out->wgt, 0, sizeof(double) * nbins);
        out->num_pixel_pairs_used = out->num_guard_rechecks = out->num_pairs_evaluated = 0;
        out->kernel_ms = 0;
        return TOS_OK;
    }
    if (!ctx->skybin) ctx->skybin = new BatchState();
    BatchState &B = *ctx->skybin;
    B.drop_graph();
    B.key.clear();
    XiKernelArgs base;
    fill_args(cat, p, g, !(flags & TOS_XI_NO_WINDOW), base);
    base.interleave = (flags & TOS_XI_NO_INTERLEAVE) ? 0 : 1;
    base.zconst = (flags & TOS_XI_NO_ZCONST) ? 0 : 1;
    base.wg_item_ptr = nullptr;
    bool fixed, simple;
    xi_variant(g.type, base.ge, base.gf, fixed, simple);
    BatchGroup G;
    G.key = VariantKey{g.type, prec, fixed, simple};
    GridExact ge;
    grid_exact_consts(g, ge);
    G.lds = xi_lds_bytes(TOS_ACC_CORE, (int)ge.nslots);
    if ((int64_t)G.lds + 64 > ctx->max_lds_block) {
        set_err(ctx, "grid has too many bins for the LDS histogram");
        return TOS_ERR_INVALID;
    }
    const TuneEntry *t = find_tuned(ctx, g.type, prec, fixed, simple);
    G.block = t ? t->block : 512;
    const int bpc = std::min(t ? t->blocks_per_cu : 64, max_blocks_per_cu(ctx, G.lds, G.block));
    const int64_t budget = task_budget(ctx, bpc, t);
    double total = 0;
    std::vector<double> seg_cost(segs.size(), 0.0);
    for (size_t q = 0; q < segs.size(); ++q) {
        for (int32_t i = segs[q].a; i < segs[q].b; ++i) seg_cost[q] += (double)p->item_cost[(size_t)i];
        total += seg_cost[q];
    }
    std::vector<XiWgTask> tasks;
    std::vector<uint64_t> task_cost;
    std::vector<int> rows((size_t)nk, 0);
    for (size_t q = 0; q < segs.size(); ++q) {
        const Seg &S = segs[q];
        const int nwg = seg_workgroups(seg_cost[q], total, budget, (int64_t)(S.b - S.a));
        std::vector<uint64_t> c(p->item_cost.begin() + S.a, p->item_cost.begin() + S.b);
        std::vector<int32_t> ptr;
        balance_items(c, nwg, ptr);
        for (int w = 0; w < nwg; ++w) {
            tasks.push_back(XiWgTask{S.key, S.a + ptr[(size_t)w], S.a + ptr[(size_t)w + 1], rows[(size_t)S.key]++});
            task_cost.push_back(range_cost(c, ptr[(size_t)w], ptr[(size_t)w + 1]));
        }
    }
    order_longest_first(tasks, task_cost, 0);
    G.task_begin = 0;
    G.n_wg = (int)tasks.size();
    G.jobs.clear();
    B.groups.assign(1, G);
    // partial rows, outputs
    B.nbins.assign((size_t)nk, nbins);
    B.out_off.assign((size_t)nk, 0);
    B.res_words = 2 * (size_t)nk;
    B.max_fields = (int64_t)nbins * 3;
    std::vector<size_t> row_off((size_t)nk, 0);
    size_t total_rows = 0;
    for (int k = 0; k < nk; ++k) {
        row_off[(size_t)k] = total_rows;
        total_rows += (size_t)rows[(size_t)k];
        B.out_off[(size_t)k] = B.res_words;
        B.res_words += 3 * (size_t)nbins;
    }
    if (B.parts_cnt.reserve(std::max<size_t>(4, total_rows * nbins * 4)) != TOS_OK ||
        B.parts_sum.reserve(std::max<size_t>(8, total_rows * nbins * 16)) != TOS_OK ||
        B.res.reserve(B.res_words * 8) != TOS_OK) {
        set_err(ctx, "device allocation failed for the sky-bin partials");
        return TOS_ERR_NOMEM;
    }
    if (B.host_res_bytes < B.res_words * 8) {
        if (B.host_res) (void)hipHostFree(B.host_res);
        B.host_res = nullptr;
        B.host_res_bytes = 0;
        if (hipHostMalloc(&B.host_res, B.res_words * 8, hipHostMallocDefault) != hipSuccess) {
            B.host_res = nullptr;
            set_err(ctx, "pinned host allocation failed for the sky-bin results");
            return TOS_ERR_NOMEM;
        }
        B.host_res_bytes = B.res_words * 8;
    }
    std::vector<XiKernelArgs> args((size_t)nk, base);
    std::vector<XiReduceJob> red((size_t)nk);
    uint64_t *res = (uint64_t *)B.res.ptr;
    for (int k = 0; k < nk; ++k) {
        XiKernelArgs &a = args[(size_t)k];
        a.part_cnt = (uint32_t *)B.parts_cnt.ptr + row_off[(size_t)k] * nbins;
        a.part_sum = (double *)B.parts_sum.ptr + row_off[(size_t)k] * nbins * 2;
        a.n_recheck = (unsigned long long *)(res + 2 * k);
        XiReduceJob &r = red[(size_t)k];
        r.part_cnt = a.part_cnt;
        r.part_sum = a.part_sum;
        r.n_wg = rows[(size_t)k];
        r.nbins = nbins;
        r.nsum = 2;
        r.out_cnt = res + B.out_off[(size_t)k];
        r.out_sum = (double *)(res + B.out_off[(size_t)k] + nbins);
    }
    int e = launch_xi_batch(G.key.coord, G.key.prec, G.key.fixed, G.key.simple, XiBatchArgs{}, 0, G.block, G.lds,
                            ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, (hipError_t)e, "sky-bin kernel LDS limit");
    if ((rc = upload(ctx, B.args, args.data(), args.size() * sizeof(XiKernelArgs), "sky-bin job table")) !=
            TOS_OK ||
        (rc = upload(ctx, B.tasks, tasks.data(), tasks.size() * sizeof(XiWgTask), "sky-bin workgroup tasks")) !=
            TOS_OK ||
        (rc = upload(ctx, B.red, red.data(), red.size() * sizeof(XiReduceJob), "sky-bin reduction table")) != TOS_OK)
        return rc;
    if (!B.ev0 && (hipEventCreate(&B.ev0) != hipSuccess || hipEventCreate(&B.ev1) != hipSuccess)) {
        set_err(ctx, "event creation failed");
        return TOS_ERR_DEVICE;
    }
    hipEventRecord(B.ev0, ctx->stream);
    if (nk > 0 && (rc = batch_enqueue(ctx, B, nk, ctx->stream)) != TOS_OK) return rc;
    hipEventRecord(B.ev1, ctx->stream);
    const hipError_t he = hipStreamSynchronize(ctx->stream);
    if (he != hipSuccess) return hip_fail(ctx, he, "per-sky-bin pass");
    float ms = 0;
    hipEventElapsedTime(&ms, B.ev0, B.ev1);
    // subsamples in key order, and their sum
    std::memset(out->num_pairs, 0, sizeof(uint64_t) * nbins);
    std::memset(out->didj, 0, sizeof(double) * nbins);
    std::memset(out->wgt, 0, sizeof(double) * nbins);
    const uint64_t *hr = (const uint64_t *)B.host_res;
    uint64_t used = 0, rechecks = 0, evaluated = 0;
    for (int k = 0; k < nk; ++k) {
        const uint64_t *cnt = hr + B.out_off[(size_t)k];
        const double *sum = (const double *)(cnt + nbins);
        out->skybin_keys[k] = keys[(size_t)k];
        for (int b = 0; b < nbins; ++b) {
            const size_t o = (size_t)k * nbins + b;
            out->sub_num_pairs[o] = cnt[b];
            out->sub_didj[o] = sum[2 * b];
            out->sub_wgt[o] = sum[2 * b + 1];
            out->num_pairs[b] += cnt[b];
            out->didj[b] += sum[2 * b];
            out->wgt[b] += sum[2 * b + 1];
            used += cnt[b];
        }
        rechecks += hr[2 * k];
        evaluated += hr[2 * k + 1];
    }
    out->num_pixel_pairs_used = used;
    out->num_guard_rechecks = rechecks;
    out->num_pairs_evaluated = evaluated;
    out->kernel_ms = ms;
    return TOS_OK;
}

}  // namespace tos

extern "C" int tos_batch_run(tos_ctx *ctx, int32_t n, const tos_batch_job *jobs, int use_graph, tos_xi_result *out,
                             double *batch_ms)
{
    if (!ctx || n < 0 || (n > 0 && (!jobs || !out))) return TOS_ERR_INVALID;
    if (n == 0) {
        if (batch_ms) *batch_ms = 0;
        return TOS_OK;
    }
    for (int32_t j = 0; j < n; ++j) {
        const tos_batch_job &J = jobs[j];
        if (!J.cat || !J.pairs || J.pairs->cat != J.cat) {
            set_err(ctx, "tos_batch_run: job " + std::to_string(j) + " has no catalogue or a pair list of another");
            return TOS_ERR_INVALID;
        }
        if (!grid_ok(ctx, &J.grid)) return TOS_ERR_INVALID;
        if ((J.precision != TOS_FP64 && J.precision != TOS_FP32C) || J.accum != TOS_ACC_CORE) {
            set_err(ctx, "tos_batch_run: jobs take TOS_FP64 or TOS_FP32C with TOS_ACC_CORE");
            return TOS_ERR_INVALID;
        }
        if (!out[j].num_pairs || !out[j].didj || !out[j].wgt) {
            set_err(ctx, "tos_batch_run: result arrays missing for job " + std::to_string(j));
            return TOS_ERR_INVALID;
        }
    }
    hipSetDevice(ctx->device);
    int rc;
    for (int32_t j = 0; j < n; ++j) {
        tos_pairs *p = const_cast<tos_pairs *>(jobs[j].pairs);
        if ((rc = ensure_items(ctx, p, true, nullptr, tuned_step_cap(ctx, jobs[j].grid, jobs[j].precision))) !=
            TOS_OK)
            return rc;
        if (jobs[j].grid.type == TOS_OBSERVING && (rc = ensure_theta(ctx, p, jobs[j].grid)) != TOS_OK) return rc;
    }
    if (!ctx->batch) {
        ctx->batch = new BatchState();
        if (hipEventCreate(&ctx->batch->ev0) != hipSuccess || hipEventCreate(&ctx->batch->ev1) != hipSuccess) {
            set_err(ctx, "tos_batch_run: event creation failed");
            return TOS_ERR_DEVICE;
        }
    }
    BatchState &B = *ctx->batch;
    std::vector<uint64_t> key = job_key(n, jobs);
    add_buffer_key(key, n, jobs, ctx);
    if (key != B.key) {
        B.drop_graph();
        B.key.clear();
        if ((rc = batch_prepare(ctx, n, jobs, B)) != TOS_OK) return rc;
        B.key = key;
    }
    hipError_t he;
    hipEventRecord(B.ev0, ctx->stream);
    if (use_graph) {
        if (!B.exec) {
            he = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
            if (he != hipSuccess) return hip_fail(ctx, he, "graph capture begin");
            rc = batch_enqueue(ctx, B, n, ctx->stream);
            hipGraph_t g = nullptr;
            he = hipStreamEndCapture(ctx->stream, &g);
            if (rc != TOS_OK) {
                if (g) hipGraphDestroy(g);
                return rc;
            }
            if (he != hipSuccess) return hip_fail(ctx, he, "graph capture end");
            B.graph = g;
            he = hipGraphInstantiate(&B.exec, B.graph, nullptr, nullptr, 0);
            if (he != hipSuccess) {
                B.exec = nullptr;
                return hip_fail(ctx, he, "graph instantiate");
            }
        }
        he = hipGraphLaunch(B.exec, ctx->stream);
        if (he != hipSuccess) return hip_fail(ctx, he, "graph launch");
    } else {
        if ((rc = batch_enqueue(ctx, B, n, ctx->stream)) != TOS_OK) return rc;
    }
    hipEventRecord(B.ev1, ctx->stream);
    he = hipStreamSynchronize(ctx->stream);
    if (he != hipSuccess) return hip_fail(ctx, he, "tos_batch_run");
    float ms = 0;
    hipEventElapsedTime(&ms, B.ev0, B.ev1);
    if (batch_ms) *batch_ms = ms;
    batch_collect(B, n, ms, out);
    return TOS_OK;
}

extern "C" int tos_xi_autotune(tos_ctx *ctx, const tos_catalog *cat, const tos_pairs *pairs, const tos_grid *grid,
                               int32_t precision, int32_t *best_block, int32_t *best_blocks_per_cu, double *best_ms)
{
    if (!ctx || !cat || !pairs || !grid) return TOS_ERR_INVALID;
    if (!grid_ok(ctx, grid)) return TOS_ERR_INVALID;
    if (precision != TOS_FP64 && precision != TOS_FP32C) {
        set_err(ctx, "tos_xi_autotune: unknown precision");
        return TOS_ERR_INVALID;
    }
    const int nb = grid->nbins[0] * grid->nbins[1] * grid->nbins[2];
    std::vector<uint64_t> cnt((size_t)nb);
    std::vector<double> didj((size_t)nb), wgt((size_t)nb);
    GridExact ge;
    grid_exact_consts(*grid, ge);
    const size_t lds = xi_lds_bytes(TOS_ACC_CORE, (int)ge.nslots);
    bool fixed, simple;
    grid_variant(*grid, fixed, simple);
    // the entry the runs below set: they must not pick up an earlier
    // tuning's work-item size
    TuneEntry *entry = nullptr;
    for (TuneEntry &t : ctx->tuned)
        if (t.coord == grid->type && t.prec == precision && t.fixed == fixed && t.simple == simple) entry = &t;
    if (!entry) {
        ctx->tuned.push_back(TuneEntry{grid->type, precision, fixed, simple, 512, 0, 0, 0});
        entry = &ctx->tuned.back();
    }
    entry->step_cap = 0;
    // the fastest of four runs (the first warms caches and code) of one shape
    auto time_shape = [&](int block, int bpc, int per_slot, double &t) -> int {
        tos_xi_opts o;
        std::memset(&o, 0, sizeof(o));
        o.precision = precision;
        o.accum = TOS_ACC_CORE;
        o.block_size = block;
        o.blocks_per_cu = bpc;
        o.wg_per_slot = per_slot;
        tos_xi_result r;
        std::memset(&r, 0, sizeof(r));
        r.num_pairs = cnt.data();
        r.didj = didj.data();
        r.wgt = wgt.data();
        t = 1e300;
        for (int rep = 0; rep < 4; ++rep) {
            const int rc = tos_xi_run(ctx, cat, pairs, grid, &o, &r);
            if (rc != TOS_OK) return rc;
            if (rep > 0) t = std::min(t, r.kernel_ms);
        }
        return TOS_OK;
    };
    // 1. workgroup size x workgroups per CU (as the LDS allows)
    double best = -1;
    int bb = 512, bp = 0, bs = 0, bc = 0;
    for (int block : {256, 512, 1024}) {
        const int cap = max_blocks_per_cu(ctx, lds, block);
        for (int bpc : {1, 2, 4, 8}) {
            if (bpc > cap) continue;
            double t;
            const int rc = time_shape(block, bpc, 0, t);
            if (rc != TOS_OK) return rc;
            if (best < 0 || t < best) {
                best = t;
                bb = block;
                bp = bpc;
            }
        }
    }
    // 2. workgroups per resident slot (how finely the items are cut over
    // workgroups: finer evens out the tail, coarser keeps fewer partials)
    for (int per_slot : {4, 8, 16, 64}) {
        double t;
        const int rc = time_shape(bb, bp, per_slot, t);
        if (rc != TOS_OK) return rc;
        if (t < best) {
            best = t;
            bs = per_slot;
        }
    }
    // 3. work-item size: partner pixels per item (larger items amortise the
    // per-item set-up, smaller ones balance better); the items are rebuilt
    // for each size, and the default's are rebuilt last
    for (int step_cap : {1024, 2048, 8192, 16384}) {
        entry->step_cap = step_cap;
        double t;
        const int rc = time_shape(bb, bp, bs, t);
        if (rc != TOS_OK) {
            entry->step_cap = bc;
            return rc;
        }
        if (t < best) {
            best = t;
            bc = step_cap;
        }
    }
    entry->block = bb;
    entry->blocks_per_cu = bp;
    entry->wg_per_slot = bs;
    entry->step_cap = bc;
    if (best_block) *best_block = bb;
    if (best_blocks_per_cu) *best_blocks_per_cu = bp;
    if (best_ms) *best_ms = best;
    return TOS_OK;
}

extern "C" int tos_xi_tuned_shape(const tos_ctx *ctx, const tos_grid *grid, int32_t precision, int32_t shape[4])
{
    if (!ctx || !grid || !shape) return TOS_ERR_INVALID;
    bool fixed, simple;
    grid_variant(*grid, fixed, simple);
    const TuneEntry *t = find_tuned(ctx, grid->type, precision, fixed, simple);
    shape[0] = t ? t->block : 0;
    shape[1] = t ? t->blocks_per_cu : 0;
    shape[2] = t ? t->wg_per_slot : 0;
    shape[3] = t ? t->step_cap : 0;
    return TOS_OK;
}
