// tos_xi.cpp -- tos_xi_run: the estimator's pixel-pair pass on the GPU
// (XiEstimator::run, XiEstimator.cc:31-59, 121-159, with the per-pair work of
// accumulate_pixel_pairs, :161-296, and the merge of :298-305).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "tos_capi_impl.h"

using namespace tos;

namespace tos {

// Work items of a pair list, built on the device from the resident pair
// list: one count pass, the row offsets on the host, one fill pass; the
// item costs come back for the workgroup balancing.
int ensure_items(tos_ctx *ctx, tos_pairs *p, bool transpose, std::vector<void *> *defer, int32_t step_cap)
{
    if (step_cap <= 0) step_cap = kItemStepCap;
    if (p->items_ready && p->items_step_cap == step_cap && p->items_transposed == transpose) return TOS_OK;
    const tos_catalog *cat = p->cat;
    const int64_t nr = p->n_rows;
    struct Scratch {   // released (or handed to *defer) on every return
        DevBuf b[5];
        std::vector<void *> *defer = nullptr;
        ~Scratch()
        {
            for (DevBuf &x : b) {
                if (defer && x.ptr) {
                    defer->push_back(x.ptr);
                    x.ptr = nullptr;
                    x.bytes = 0;
                }
                x.release();
            }
        }
    } sc;
    sc.defer = defer;
    DevBuf &d_rlos = sc.b[0], &d_rptr = sc.b[1], &d_cnt = sc.b[2], &d_iptr = sc.b[3], &d_cost = sc.b[4];
    int rc;
    if ((rc = upload(ctx, d_rlos, p->h.row_los.data(), (size_t)nr * 4, "item rows")) != TOS_OK ||
        (rc = upload(ctx, d_rptr, p->h.row_ptr.data(), (size_t)(nr + 1) * 8, "item row offsets")) != TOS_OK)
        return rc;
    if (d_cnt.reserve((size_t)std::max<int64_t>(1, nr) * 4) != TOS_OK) {
        set_err(ctx, "device allocation failed for the work items");
        return TOS_ERR_NOMEM;
    }
    ItemArgs a{};
    a.row_los = (const int32_t *)d_rlos.ptr;
    a.row_ptr = (const int64_t *)d_rptr.ptr;
    a.other = (const int32_t *)p->other.ptr;
    a.los = (const LosInfo *)cat->los.ptr;
    a.n_rows = nr;
    a.step_cap = step_cap;
    a.transpose_below = transpose ? kTransposeBelow : 0;
    a.count = (int32_t *)d_cnt.ptr;
    int e = launch_items(a, false, ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, (hipError_t)e, "work item count launch");
    std::vector<int32_t> cnt((size_t)nr);
    hipError_t he = hipSuccess;
    if (nr > 0) he = hipMemcpyAsync(cnt.data(), d_cnt.ptr, (size_t)nr * 4, hipMemcpyDeviceToHost, ctx->stream);
    if (he == hipSuccess) he = hipStreamSynchronize(ctx->stream);
    if (he != hipSuccess) return hip_fail(ctx, he, "work item count");
    p->row_item_ptr.assign((size_t)nr + 1, 0);
    int64_t total = 0;
    for (int64_t r = 0; r < nr; ++r) {
        p->row_item_ptr[(size_t)r] = (int32_t)total;
        total += cnt[(size_t)r];
        if (total >= (1ll << 31)) {
            set_err(ctx, "more than 2^31 work items in one job");
            return TOS_ERR_INVALID;
        }
    }
    p->row_item_ptr[(size_t)nr] = (int32_t)total;
    if ((rc = upload(ctx, d_iptr, p->row_item_ptr.data(), (size_t)(nr + 1) * 4, "item offsets")) != TOS_OK) return rc;
    if (p->items.reserve((size_t)std::max<int64_t>(1, total) * sizeof(XiItem)) != TOS_OK ||
        d_cost.reserve((size_t)std::max<int64_t>(1, total) * 8) != TOS_OK) {
        set_err(ctx, "device allocation failed for the work items");
        return TOS_ERR_NOMEM;
    }
    a.row_item_ptr = (const int32_t *)d_iptr.ptr;
    a.items = (XiItem *)p->items.ptr;
    a.item_cost = (uint64_t *)d_cost.ptr;
    e = launch_items(a, true, ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, (hipError_t)e, "work item fill launch");
    p->item_cost.resize((size_t)total);
    if (total > 0)
        he = hipMemcpyAsync(p->item_cost.data(), d_cost.ptr, (size_t)total * 8, hipMemcpyDeviceToHost, ctx->stream);
    if (he == hipSuccess) he = hipStreamSynchronize(ctx->stream);
    if (he != hipSuccess) return hip_fail(ctx, he, "work item fill");
    {
        std::vector<uint64_t> pre((size_t)std::max<int64_t>(1, total), 0);
        uint64_t acc = 0;
        for (int64_t k = 0; k < total; ++k) pre[(size_t)k] = acc += p->item_cost[(size_t)k];
        if ((rc = upload(ctx, p->cost_prefix, pre.data(), (size_t)total * 8, "item cost prefix")) != TOS_OK) return rc;
        he = hipStreamSynchronize(ctx->stream);
        if (he != hipSuccess) return hip_fail(ctx, he, "item cost prefix");
    }
    p->n_items = total;
    p->items_ready = true;
    ++p->items_gen;
    p->items_step_cap = step_cap;
    p->items_transposed = transpose;
    return TOS_OK;
}

int ensure_host_pairs(tos_ctx *ctx, tos_pairs *p)
{
    if (p->pairs_on_host) return TOS_OK;
    HostPairs &h = p->h;
    h.other.resize((size_t)p->n_pairs);
    h.cos_sep.resize((size_t)p->n_pairs);
    hipError_t he = hipSuccess;
    if (p->n_pairs > 0) {
        he = hipMemcpyAsync(h.other.data(), p->other.ptr, (size_t)p->n_pairs * 4, hipMemcpyDeviceToHost, ctx->stream);
        if (he == hipSuccess)
            he = hipMemcpyAsync(h.cos_sep.data(), p->csep.ptr, (size_t)p->n_pairs * 8, hipMemcpyDeviceToHost,
                                ctx->stream);
    }
    if (he == hipSuccess) he = hipStreamSynchronize(ctx->stream);
    if (he != hipSuccess) return hip_fail(ctx, he, "pair list copy");
    p->pairs_on_host = true;
    return TOS_OK;
}

int ensure_theta(tos_ctx *ctx, tos_pairs *p, const tos_grid &g)
{
    // QuasarGrid: the sightline-pair angle bin (XiEstimator.cc:166,273),
    // computed once per pair with the host libm acos like the reference.
    if (p->theta_n == g.nbins[1] && p->theta_key[0] == g.min[1] && p->theta_key[1] == g.width[1]) return TOS_OK;
    int rc0 = ensure_host_pairs(ctx, p);
    if (rc0 != TOS_OK) return rc0;
    std::vector<uint32_t> tb((size_t)p->n_pairs);
    const double th_min = g.min[1], inv = 1.0 / g.width[1];
    for (int64_t k = 0; k < p->n_pairs; ++k) {
        const double sep = std::acos(p->h.cos_sep[(size_t)k]);
        tb[(size_t)k] = cvt_x86((sep - th_min) * inv);
    }
    int rc = upload(ctx, p->theta, tb.data(), tb.size() * sizeof(uint32_t), "theta bins");
    if (rc != TOS_OK) return rc;
    p->theta_n = g.nbins[1];
    p->theta_key[0] = g.min[1];
    p->theta_key[1] = g.width[1];
    return TOS_OK;
}

void grid_variant(const tos_grid &g, bool &fixed, bool &simple)
{
    // the variant of a typical catalogue (distances 3000-6000 Mpc/h): a key
    // for the tuned launch shapes only; launches decide from the job's own
    // constants
    GridExact ge;
    GridFast gf;
    grid_exact_consts(g, ge);
    grid_fast_consts(g, 3000.0, 6000.0, gf);
    xi_variant(g.type, ge, gf, fixed, simple);
}

const TuneEntry *find_tuned(const tos_ctx *ctx, int coord, int prec, bool fixed, bool simple)
{
    for (const TuneEntry &t : ctx->tuned)
        if (t.coord == coord && t.prec == prec && t.fixed == fixed && t.simple == simple) return &t;
    return nullptr;
}

int32_t tuned_step_cap(const tos_ctx *ctx, const tos_grid &g, int prec)
{
    bool fixed, simple;
    grid_variant(g, fixed, simple);
    const TuneEntry *t = find_tuned(ctx, g.type, prec, fixed, simple);
    return (t && t->step_cap > 0) ? t->step_cap : kItemStepCap;
}

int max_blocks_per_cu(const tos_ctx *ctx, size_t lds, int block)
{
    int bpc = (int)std::max<int64_t>(1, ctx->max_lds_cu / (int64_t)(lds + 64));
    bpc = std::min(bpc, std::max(1, 2048 / block));
    return std::min(bpc, 8);
}

// Workgroups of a single-job launch: kWgPerSlot per resident slot, the
// items balanced over them by cost; at most kMaxWorkgroups and at most
// kMaxPartialBytes of partial histogram rows.
constexpr int kWgPerSlot = 32;
constexpr int64_t kMaxWorkgroups = 16384;
constexpr int64_t kMaxPartialBytes = (int64_t)1 << 30;

int plan_launch(tos_ctx *ctx, const tos_xi_opts *opts, int prec, int accum, const tos_grid &g, size_t n_items,
                XiLaunch &L, uint32_t band_rows)
{
    GridExact ge;
    grid_exact_consts(g, ge);
    const bool banded = band_rows > 0 && band_rows < ge.n[0];
    L.nbins = banded ? (int)(band_rows * ge.rowlen) : (int)ge.ntot;
    L.nsum = xi_nsum(accum);
    L.lds = xi_lds_bytes(accum, banded ? (int)(band_rows * (ge.rowlen + ge.pad)) : (int)ge.nslots);
    if ((int64_t)L.lds + 64 > ctx->max_lds_block) {
        set_err(ctx, "grid has too many bins for the LDS histogram with this accumulator set");
        return TOS_ERR_INVALID;
    }
    const TuneEntry *t = nullptr;
    if (!(opts && opts->block_size > 0) && accum == TOS_ACC_CORE) {
        bool fixed, simple;
        grid_variant(g, fixed, simple);
        t = find_tuned(ctx, g.type, prec, fixed, simple);
    }
    L.block = (opts && opts->block_size > 0) ? opts->block_size : (t ? t->block : 512);
    if (L.block % 64 != 0 || L.block < 64 || L.block > 1024) {
        set_err(ctx, "block_size must be a multiple of 64 in [64, 1024]");
        return TOS_ERR_INVALID;
    }
    int bpc = (opts && opts->blocks_per_cu > 0) ? opts->blocks_per_cu : (t ? t->blocks_per_cu : 0);
    if (bpc == 0) bpc = max_blocks_per_cu(ctx, L.lds, L.block);
    L.bpc = bpc;
    const int per_slot = (opts && opts->wg_per_slot > 0) ? opts->wg_per_slot
                         : (t && t->wg_per_slot > 0)     ? t->wg_per_slot
                                                          : kWgPerSlot;
    int64_t nwg = std::min<int64_t>((int64_t)ctx->ncu * bpc * per_slot, kMaxWorkgroups);
    nwg = std::min<int64_t>(nwg, kMaxPartialBytes / std::max<int64_t>(1, (int64_t)L.nbins * (4 + 8 * L.nsum)));
    nwg = std::min<int64_t>(nwg, (int64_t)std::max<size_t>(1, n_items));
    L.n_wg = (int)std::max<int64_t>(1, nwg);
    L.n_seg = std::min(32, L.n_wg);
    return TOS_OK;
}

uint32_t band_rows_for(const tos_ctx *ctx, int accum, const tos_grid &g)
{
    GridExact ge;
    grid_exact_consts(g, ge);
    const uint32_t rowslots = ge.rowlen + ge.pad;
    auto fits = [&](uint32_t rows) {
        return (int64_t)xi_lds_bytes(accum, (int)(rows * rowslots)) + 64 <= ctx->max_lds_block;
    };
    if (fits(ge.n[0])) return ge.n[0];
    uint32_t lo = 0, hi = ge.n[0];   // the most rows that fit
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo + 1) / 2;
        if (fits(mid))
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

int reserve_scratch(tos_ctx *ctx, const XiLaunch &L, int nbins_total)
{
    const size_t nb = (size_t)L.nbins;
    const size_t nt = (size_t)std::max(nbins_total, L.nbins);
    if (ctx->part_cnt.reserve((size_t)L.n_wg * nb * 4) != TOS_OK ||
        ctx->part_sum.reserve((size_t)L.n_wg * nb * L.nsum * 8) != TOS_OK ||
        ctx->seg_cnt.reserve((size_t)L.n_seg * nb * 8) != TOS_OK ||
        ctx->seg_sum.reserve((size_t)L.n_seg * nb * L.nsum * 8) != TOS_OK ||
        ctx->out_cnt.reserve(nt * 8) != TOS_OK || ctx->out_sum.reserve(nt * L.nsum * 8) != TOS_OK ||
        ctx->recheck.reserve(64) != TOS_OK) {
        set_err(ctx, "device allocation failed for the histogram partials");
        return TOS_ERR_NOMEM;
    }
    return TOS_OK;
}

void fill_args(const tos_catalog *cat, const tos_pairs *p, const tos_grid &g, bool window, XiKernelArgs &a)
{
    std::memset(&a, 0, sizeof(a));
    a.pix = (const float4 *)cat->pix.ptr;
    a.los = (const LosInfo *)cat->los.ptr;
    a.other = (const int32_t *)p->other.ptr;
    a.csep = (const double *)p->csep.ptr;
    a.theta_bin = (const uint32_t *)p->theta.ptr;
    a.items = (const XiItem *)p->items.ptr;
    grid_exact_consts(g, a.ge);
    grid_fast_consts(g, cat->dmin, cat->dmax, a.gf);
    a.nbins = (int32_t)a.ge.ntot;
    a.window = 0;
    a.win_half = 0.0;
    a.perp_cut = -1.0;
    a.interleave = 1;
    a.zconst = 1;
    a.band_lo8 = 0;
    a.band_len8 = 8u * a.ge.nslots;
    a.band_bin0 = 0;
    a.banded = 0;
    if (window) {
        const bool sorted = g.type == TOS_OBSERVING ? cat->l_sorted : cat->d_sorted;
        const double X = grid_window_half(g, cat->dmax, g.type == TOS_OBSERVING ? cat->lmax_abs : cat->dmax);
        if (sorted && X > 0) {
            a.window = g.type == TOS_OBSERVING ? 2 : 1;
            a.win_half = X;
            if (a.window == 1 && cat->dmin > 0) a.perp_cut = grid_perp_cut(g, cat->dmin, cat->dmax);
        }
    }
}

}  // namespace tos

extern "C" int tos_xi_run(tos_ctx *ctx, const tos_catalog *cat, const tos_pairs *pairs_c, const tos_grid *grid,
                          const tos_xi_opts *opts, tos_xi_result *out)
{
    if (!ctx || !cat || !pairs_c || !out) return TOS_ERR_INVALID;
    if (!grid_ok(ctx, grid)) return TOS_ERR_INVALID;
    tos_pairs *p = const_cast<tos_pairs *>(pairs_c);
    if (p->cat != cat) {
        set_err(ctx, "tos_xi_run: the pair list was built for another catalogue");
        return TOS_ERR_INVALID;
    }
    const int prec = opts ? opts->precision : TOS_FP64;
    const int accum = opts ? opts->accum : TOS_ACC_CORE;
    if ((prec != TOS_FP64 && prec != TOS_FP32C) || (accum != TOS_ACC_CORE && accum != TOS_ACC_ALL)) {
        set_err(ctx, "tos_xi_run: unknown precision or accumulator set");
        return TOS_ERR_INVALID;
    }
    if (opts && opts->per_skybin) {
        const bool subs = p->h.stats.n_skybins > 0;
        if (accum != TOS_ACC_CORE || !out->num_pairs || !out->didj || !out->wgt ||
            (subs && (!out->skybin_keys || !out->sub_num_pairs || !out->sub_didj || !out->sub_wgt))) {
            set_err(ctx, "tos_xi_run: per-sky-bin subsamples take TOS_ACC_CORE and every subsample array");
            return TOS_ERR_INVALID;
        }
        hipSetDevice(ctx->device);
        return xi_run_per_skybin(ctx, cat, p, *grid, prec, opts->flags, out);
    }
    if (!out->num_pairs || !out->didj || !out->wgt ||
        (accum == TOS_ACC_ALL && (!out->di || !out->dj || !out->wi || !out->wj || !out->z))) {
        set_err(ctx, "tos_xi_run: result arrays missing");
        return TOS_ERR_INVALID;
    }
    hipSetDevice(ctx->device);
    const tos_grid g = *grid;
    const int nbins = g.nbins[0] * g.nbins[1] * g.nbins[2];
    out->num_pixel_pairs_used = 0;
    out->num_guard_rechecks = 0;
    out->num_pairs_evaluated = 0;
    out->kernel_ms = 0;

    const int flags = opts ? opts->flags : 0;
    int rc = ensure_items(ctx, p, !(flags & TOS_XI_NO_TRANSPOSE), nullptr, tuned_step_cap(ctx, g, prec));
    if (rc != TOS_OK) return rc;
    const size_t n_items = (size_t)p->n_items;
    if (n_items == 0) {
        std::memset(out->num_pairs, 0, sizeof(uint64_t) * nbins);
        std::memset(out->didj, 0, sizeof(double) * nbins);
        std::memset(out->wgt, 0, sizeof(double) * nbins);
        if (accum == TOS_ACC_ALL)
            for (double *a : {out->di, out->dj, out->wi, out->wj, out->z}) std::memset(a, 0, sizeof(double) * nbins);
        return TOS_OK;
    }
    if (g.type == TOS_OBSERVING && (rc = ensure_theta(ctx, p, g)) != TOS_OK) return rc;
    // a grid too large for one LDS histogram runs in bands of axis-0 rows,
    // one full pass over the pairs per band
    const uint32_t rows = band_rows_for(ctx, accum, g);
    if (rows == 0) {
        set_err(ctx, "one axis-0 row of the grid has too many bins for the LDS histogram");
        return TOS_ERR_INVALID;
    }
    GridExact ge;
    grid_exact_consts(g, ge);
    const bool banded = rows < ge.n[0];
    XiLaunch L;
    if ((rc = plan_launch(ctx, opts, prec, accum, g, n_items, L, banded ? rows : 0)) != TOS_OK) return rc;
    if ((rc = reserve_scratch(ctx, L, nbins)) != TOS_OK) return rc;
    std::vector<int32_t> wgp;
    balance_items(p->item_cost, L.n_wg, wgp);
    {
        // a workgroup's u32 partial counts hold at most its items' cost
        // (pixel pairs before the windows)
        uint64_t worst = 0;
        for (int w = 0; w < L.n_wg; ++w) {
            uint64_t c = 0;
            for (int32_t i = wgp[(size_t)w]; i < wgp[(size_t)w + 1]; ++i) c += p->item_cost[(size_t)i];
            worst = std::max(worst, c);
        }
        if (worst >= 0xFFFFFFFFull) {
            set_err(ctx, "tos_xi_run: a workgroup would count 2^32 pixel pairs; split the catalogue");
            return TOS_ERR_INVALID;
        }
    }
    if ((rc = upload(ctx, ctx->wgptr, wgp.data(), wgp.size() * sizeof(int32_t), "workgroup ranges")) != TOS_OK) return rc;

    XiKernelArgs a;
    fill_args(cat, p, g, !(flags & TOS_XI_NO_WINDOW), a);
    a.interleave = (flags & TOS_XI_NO_INTERLEAVE) ? 0 : 1;
    a.zconst = (flags & TOS_XI_NO_ZCONST) ? 0 : 1;
    a.wg_item_ptr = (const int32_t *)ctx->wgptr.ptr;
    a.part_cnt = (uint32_t *)ctx->part_cnt.ptr;
    a.part_sum = (double *)ctx->part_sum.ptr;
    a.n_recheck = (unsigned long long *)ctx->recheck.ptr;
    hipError_t he = hipMemsetAsync(ctx->recheck.ptr, 0, 16, ctx->stream);
    if (he != hipSuccess) return hip_fail(ctx, he, "tos_xi_run memset");
    hipEventRecord(ctx->ev0, ctx->stream);
    const uint32_t rowslots = ge.rowlen + ge.pad;
    for (uint32_t r0 = 0; r0 < ge.n[0]; r0 += rows) {
        const uint32_t r1 = std::min(ge.n[0], r0 + rows);
        if (banded) {
            a.banded = 1;
            a.band_lo8 = 8u * r0 * rowslots;
            a.band_len8 = 8u * (r1 - r0) * rowslots;
            a.band_bin0 = (int32_t)(r0 * ge.rowlen);
            a.nbins = (int32_t)((r1 - r0) * ge.rowlen);
        }
        int e = launch_xi(g.type, prec, accum, a, L.n_wg, L.block, L.lds, ctx->stream);
        if (e != hipSuccess) return hip_fail(ctx, (hipError_t)e, "xi kernel launch");
        if (!banded) hipEventRecord(ctx->ev1, ctx->stream);   // the pair kernel alone
        e = launch_reduce(a.part_cnt, a.part_sum, L.n_wg, a.nbins, L.nsum, (uint64_t *)ctx->out_cnt.ptr + a.band_bin0,
                          (double *)ctx->out_sum.ptr + (size_t)a.band_bin0 * L.nsum, (uint32_t *)ctx->seg_cnt.ptr,
                          (double *)ctx->seg_sum.ptr, L.n_seg, ctx->stream);
        if (e != hipSuccess) return hip_fail(ctx, (hipError_t)e, "reduce launch");
    }
    if (banded) hipEventRecord(ctx->ev1, ctx->stream);   // every band's kernel and reduction
    std::vector<double> sums((size_t)nbins * L.nsum);
    unsigned long long nre[2] = {0, 0};
    hipMemcpyAsync(out->num_pairs, ctx->out_cnt.ptr, (size_t)nbins * 8, hipMemcpyDeviceToHost, ctx->stream);
    hipMemcpyAsync(sums.data(), ctx->out_sum.ptr, sums.size() * 8, hipMemcpyDeviceToHost, ctx->stream);
    hipMemcpyAsync(nre, ctx->recheck.ptr, 16, hipMemcpyDeviceToHost, ctx->stream);
    he = hipStreamSynchronize(ctx->stream);
    if (he != hipSuccess) return hip_fail(ctx, he, "tos_xi_run");
    float ms = 0;
    hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
    out->kernel_ms = ms;
    out->num_guard_rechecks = nre[0];
    out->num_pairs_evaluated = nre[1];
    uint64_t used = 0;
    for (int b = 0; b < nbins; ++b) {
        const double *s = &sums[(size_t)b * L.nsum];
        out->didj[b] = s[0];
        out->wgt[b] = s[1];
        if (accum == TOS_ACC_ALL) {
            out->di[b] = s[2];
            out->dj[b] = s[3];
            out->wi[b] = s[4];
            out->wj[b] = s[5];
            out->z[b] = s[6];
        }
        used += out->num_pairs[b];
    }
    out->num_pixel_pairs_used = used;
    return TOS_OK;
}


