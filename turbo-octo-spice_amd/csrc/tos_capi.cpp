// tos_capi.cpp -- the C-ABI of include/tos_amd.h: context, device-resident
// catalogues and candidate CSRs, the estimator run and the 3D brute force.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "tos_capi_impl.h"
#include "tos_parallel.h"

using namespace tos;

namespace tos {

void set_err(tos_ctx *ctx, const std::string &m)
{
    if (ctx) {
        std::lock_guard<std::mutex> g(ctx->err_mu);
        ctx->err = m;
    }
}

int hip_fail(tos_ctx *ctx, hipError_t e, const char *where)
{
    set_err(ctx, std::string(where) + ": " + hipGetErrorString(e));
    return TOS_ERR_DEVICE;
}

int DevBuf::reserve(size_t n)
{
    if (n <= bytes) return TOS_OK;
    if (ptr) hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
    if (n == 0) return TOS_OK;
    if (hipMalloc(&ptr, n) != hipSuccess) {
        ptr = nullptr;
        return TOS_ERR_NOMEM;
    }
    bytes = n;
    return TOS_OK;
}

void DevBuf::release()
{
    if (ptr) hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

int upload(tos_ctx *ctx, DevBuf &b, const void *src, size_t n, const char *what)
{
    if (b.reserve(n) != TOS_OK) {
        set_err(ctx, std::string("device allocation failed: ") + what);
        return TOS_ERR_NOMEM;
    }
    if (n == 0) return TOS_OK;
    hipError_t e = hipMemcpyAsync(b.ptr, src, n, hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, what);
    return TOS_OK;
}

bool grid_ok(tos_ctx *ctx, const tos_grid *g)
{
    if (!g) {
        set_err(ctx, "grid is NULL");
        return false;
    }
    if (g->type < TOS_POLAR || g->type > TOS_OBSERVING) {
        set_err(ctx, "grid type must be TOS_POLAR, TOS_CARTESIAN or TOS_OBSERVING");
        return false;
    }
    uint64_t tot = 1;
    for (int k = 0; k < 3; ++k) {
        if (g->nbins[k] <= 0 || g->nbins[k] > (1 << 24)) {
            set_err(ctx, "grid axis bin count must be in [1, 2^24]");
            return false;
        }
        if (k < 2 && !(g->width[k] > 0 && std::isfinite(g->width[k]))) {
            set_err(ctx, "grid axis width must be positive and finite");
            return false;
        }
        tot *= (uint64_t)g->nbins[k];
    }
    if (tot >= (1ull << 31)) {
        set_err(ctx, "grid has too many bins");
        return false;
    }
    return true;
}

}  // namespace tos

// ---------------------------------------------------------------------------
// Context
// ---------------------------------------------------------------------------
extern "C" int tos_ctx_create(int device, tos_ctx **out)
{
    if (!out) return TOS_ERR_INVALID;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return TOS_ERR_DEVICE;
    if (device < 0 || device >= ndev) return TOS_ERR_INVALID;
    tos_ctx *c = new (std::nothrow) tos_ctx();
    if (!c) return TOS_ERR_NOMEM;
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipGetDeviceProperties(&c->prop, device) != hipSuccess ||
        hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
        delete c;
        return TOS_ERR_DEVICE;
    }
    c->ncu = c->prop.multiProcessorCount;
    c->max_lds_block = xi_max_lds(device);
    int v = 0;
    c->max_lds_cu = (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, device) == hipSuccess && v > 0)
                        ? v
                        : c->max_lds_block;
    const int rc = noalloc_setup(c);
    if (rc != TOS_OK) {
        hipEventDestroy(c->ev0);
        hipEventDestroy(c->ev1);
        hipStreamDestroy(c->stream);
        delete c;
        return rc;
    }
    *out = c;
    return TOS_OK;
}

extern "C" int tos_ctx_destroy(tos_ctx *ctx)
{
    if (!ctx) return TOS_ERR_INVALID;
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    batch_release(ctx);
    for (DevBuf *b : {&ctx->part_cnt, &ctx->part_sum, &ctx->seg_cnt, &ctx->seg_sum, &ctx->out_cnt, &ctx->out_sum,
                      &ctx->recheck, &ctx->scratch_a, &ctx->scratch_b, &ctx->scratch_c, &ctx->scratch_d, &ctx->theta,
                      &ctx->items, &ctx->wgptr})
        b->release();
    hipEventDestroy(ctx->ev0);
    hipEventDestroy(ctx->ev1);
    for (int k = 0; k < 2; ++k) {
        if (ctx->stage[k]) hipHostFree(ctx->stage[k]);
        if (ctx->stage_ev[k]) hipEventDestroy(ctx->stage_ev[k]);
    }
    hipStreamDestroy(ctx->stream);
    delete ctx;
    return TOS_OK;
}

extern "C" const char *tos_last_error(const tos_ctx *ctx)
{
    if (!ctx) return "no context";
    return ctx->err.c_str();
}

extern "C" int tos_device_info_get(tos_ctx *ctx, tos_device_info *out)
{
    if (!ctx || !out) return TOS_ERR_INVALID;
    std::memset(out, 0, sizeof(*out));
    std::snprintf(out->name, sizeof(out->name), "%s", ctx->prop.name);
    std::snprintf(out->arch, sizeof(out->arch), "%s", ctx->prop.gcnArchName);
    out->compute_units = ctx->prop.multiProcessorCount;
    out->wavefront_size = ctx->prop.warpSize;
    out->lds_per_block = ctx->max_lds_block;
    out->global_mem = (int64_t)ctx->prop.totalGlobalMem;
    out->clock_khz = ctx->prop.clockRate;
    return TOS_OK;
}

extern "C" const char *tos_build_info(void)
{
    return "tos_amd abi 1; gfx950; hip " TOS_STR(HIP_VERSION_MAJOR) "." TOS_STR(HIP_VERSION_MINOR);
}

// ---------------------------------------------------------------------------
// Grids
// ---------------------------------------------------------------------------
extern "C" int tos_grid_parse_axis(const char *spec, tos_grid *g, int axis)
{
    // likely::createBinning "[min:max]*n" (uniform binning)
    if (!spec || !g || axis < 0 || axis > 2) return TOS_ERR_INVALID;
    double lo, hi;
    int n;
    char tail;
    if (std::sscanf(spec, " [ %lf : %lf ] * %d %c", &lo, &hi, &n, &tail) != 3) return TOS_ERR_INVALID;
    if (n <= 0 || !(hi > lo)) return TOS_ERR_INVALID;
    g->nbins[axis] = n;
    g->min[axis] = lo;
    g->width[axis] = (hi - lo) / n;
    g->max[axis] = lo + n * g->width[axis];
    return TOS_OK;
}

extern "C" double tos_grid_max_angular_scale(const tos_grid *g, double scale)
{
    if (!g) return NAN;
    switch (g->type) {
    case TOS_CARTESIAN: return g->max[1] / scale;                                           // CartesianGrid.h:46
    case TOS_POLAR: return g->max[0] * std::sqrt(1.0 - g->min[1] * g->min[1]) / scale;      // PolarGrid.h:46
    default: return g->max[1];                                                              // QuasarGrid.h:45
    }
}

extern "C" int tos_grid_bin_centers(const tos_grid *g, int32_t index, double centers[3])
{
    if (!g || !centers) return TOS_ERR_INVALID;
    const int64_t tot = (int64_t)g->nbins[0] * g->nbins[1] * g->nbins[2];
    if (index < 0 || index >= tot) return TOS_ERR_INVALID;
    // global index = (i0 * n1 + i1) * n2 + i2 (XiEstimator.cc:249,254)
    int64_t r = index;
    const int64_t i2 = r % g->nbins[2];
    r /= g->nbins[2];
    const int64_t i1 = r % g->nbins[1];
    const int64_t i0 = r / g->nbins[1];
    const int64_t ix[3] = {i0, i1, i2};
    for (int k = 0; k < 3; ++k) centers[k] = g->min[k] + (ix[k] + 0.5) * g->width[k];
    return TOS_OK;
}

// ---------------------------------------------------------------------------
// Catalogue
// ---------------------------------------------------------------------------
// Pixels per pinned staging chunk of a catalogue upload (32 MiB).
constexpr int64_t kStagePixels = (int64_t)1 << 21;

extern "C" int tos_catalog_upload(tos_ctx *ctx, const tos_sightlines *s, tos_catalog **out)
{
    if (!ctx || !s || !out) return TOS_ERR_INVALID;
    *out = nullptr;
    if (s->n_los < 0 || s->n_los >= (1ll << 31) || (s->n_los > 0 && !s->offsets)) {
        set_err(ctx, "tos_catalog_upload: bad sightline count or offsets");
        return TOS_ERR_INVALID;
    }
    const int64_t nl = s->n_los;
    const int64_t np = nl > 0 ? s->offsets[nl] : 0;
    if ((nl > 0 && s->offsets[0] != 0) || np < 0 || np >= (1ll << 31)) {
        set_err(ctx, "tos_catalog_upload: offsets must start at 0 and stay below 2^31");
        return TOS_ERR_INVALID;
    }
    if (np > 0 && (!s->value || !s->loglam || !s->weight || !s->distance)) {
        set_err(ctx, "tos_catalog_upload: pixel arrays missing");
        return TOS_ERR_INVALID;
    }
    if (nl > 0 && (!s->ra || !s->dec)) {
        set_err(ctx, "tos_catalog_upload: ra/dec missing");
        return TOS_ERR_INVALID;
    }
    tos_catalog *c = new (std::nothrow) tos_catalog();
    if (!c) return TOS_ERR_NOMEM;
    c->n_los = nl;
    c->n_pix = np;
    c->h_los.resize((size_t)nl);
    c->h_offsets.assign(s->offsets, s->offsets + nl + 1);
    if (nl == 0) c->h_offsets.assign(1, 0);
    c->h_ra.assign(s->ra, s->ra + nl);
    c->h_dec.assign(s->dec, s->dec + nl);
    // offsets first (serially, cheap): the first sightline whose range
    // decreases or leaves [0, np); the sightlines before it have disjoint
    // ranges and are packed in parallel
    int64_t bad_off = nl;
    for (int64_t k = 0; k < nl; ++k) {
        const int64_t a = s->offsets[k], b = s->offsets[k + 1];
        if (b < a || b > np) {
            bad_off = k;
            break;
        }
    }
    struct LosScan {
        float lmin, lmax;
        double dmin, dmax, labs;
        bool d_sorted, l_sorted, finite;
    };
    std::vector<LosScan> scan((size_t)bad_off);
    // one sightline: its pixels into dst (dst[0] is pixel offsets[k]) and
    // its scan
    auto pack = [&](int64_t k, float4 *dst) {
        const int64_t a = s->offsets[k], b = s->offsets[k + 1];
        LosScan r{INFINITY, -INFINITY, INFINITY, 0.0, 0.0, true, true, true};
        for (int64_t p = a; p < b; ++p) {
            if (p > a) {
                if (!(s->distance[p] >= s->distance[p - 1])) r.d_sorted = false;
                if (!(s->loglam[p] >= s->loglam[p - 1])) r.l_sorted = false;
            }
            const float d = s->distance[p], l = s->loglam[p], w = s->weight[p], v = s->value[p];
            // {distance, loglam, weight, f32(weight * value)} (types.h:78)
            dst[p - a] = make_float4(d, l, w, w * v);
            r.lmin = std::min(r.lmin, l);
            r.lmax = std::max(r.lmax, l);
            r.dmin = std::min(r.dmin, (double)std::fabs(d));
            r.dmax = std::max(r.dmax, (double)std::fabs(d));
            r.labs = std::max(r.labs, (double)std::fabs(l));
            r.finite = r.finite && std::isfinite(d) && std::isfinite(l) && std::isfinite(w) && std::isfinite(v);
        }
        scan[(size_t)k] = r;
    };
    // the first failing sightline in order decides the error, as a
    // sequential pass would; k_end: sightlines scanned so far
    double dmin = INFINITY, dmax = 0;
    auto fold = [&](int64_t k0, int64_t k1) {
        for (int64_t k = k0; k < k1; ++k) {
            const LosScan &r = scan[(size_t)k];
            if (!r.finite) return false;
            c->d_sorted = c->d_sorted && r.d_sorted;
            c->l_sorted = c->l_sorted && r.l_sorted;
            dmin = std::min(dmin, r.dmin);
            dmax = std::max(dmax, r.dmax);
            c->lmax_abs = std::max(c->lmax_abs, r.labs);
            const int64_t a = s->offsets[k], b = s->offsets[k + 1];
            c->h_los[(size_t)k] = LosInfo{(int32_t)a, (int32_t)(b - a), r.lmin, r.lmax};
        }
        return true;
    };
    hipSetDevice(ctx->device);
    int rc = TOS_OK;
    bool finite = true;
    if (bad_off == nl) {
        if (c->pix.reserve((size_t)np * sizeof(float4)) != TOS_OK) {
            delete c;
            set_err(ctx, "device allocation failed: catalog pixels");
            return TOS_ERR_NOMEM;
        }
    }
    // Pinned path: sightlines packed in chunks of whole sightlines into two
    // pinned buffers in turn, each copied to the device while the next one
    // is packed.  A sightline longer than a chunk (or a failed pinned
    // allocation) takes the pageable path: the whole catalogue packed into
    // host memory, one copy.
    int64_t longest = 0;
    for (int64_t k = 0; k < bad_off; ++k) longest = std::max(longest, s->offsets[k + 1] - s->offsets[k]);
    bool pinned = bad_off == nl && np > 0 && longest <= kStagePixels;
    for (int k = 0; pinned && k < 2; ++k) {
        if (!ctx->stage[k] && hipHostMalloc(&ctx->stage[k], (size_t)kStagePixels * sizeof(float4)) != hipSuccess)
            ctx->stage[k] = nullptr;
        if (!ctx->stage_ev[k] && hipEventCreateWithFlags(&ctx->stage_ev[k], hipEventDisableTiming) != hipSuccess)
            ctx->stage_ev[k] = nullptr;
        pinned = ctx->stage[k] && ctx->stage_ev[k];
    }
    if (pinned) {
        bool used[2] = {false, false};
        int buf = 0;
        for (int64_t k0 = 0; k0 < nl && rc == TOS_OK && finite;) {
            int64_t k1 = k0;
            while (k1 < nl && s->offsets[k1 + 1] - s->offsets[k0] <= kStagePixels) ++k1;
            if (used[buf]) {
                const hipError_t e = hipEventSynchronize(ctx->stage_ev[buf]);
                if (e != hipSuccess) {
                    rc = hip_fail(ctx, e, "tos_catalog_upload staging");
                    break;
                }
            }
            float4 *st = (float4 *)ctx->stage[buf];
            const int64_t p0 = s->offsets[k0];
            parallel_for(k1 - k0, 256, [&](int64_t q) { pack(k0 + q, st + (s->offsets[k0 + q] - p0)); });
            finite = fold(k0, k1);
            const size_t bytes = (size_t)(s->offsets[k1] - p0) * sizeof(float4);
            hipError_t e = hipMemcpyAsync((float4 *)c->pix.ptr + p0, st, bytes, hipMemcpyHostToDevice, ctx->stream);
            if (e == hipSuccess) e = hipEventRecord(ctx->stage_ev[buf], ctx->stream);
            if (e != hipSuccess) rc = hip_fail(ctx, e, "tos_catalog_upload copy");
            used[buf] = true;
            buf ^= 1;
            k0 = k1;
        }
        // the copies read the staging buffers: finish them before returning
        const hipError_t e = hipStreamSynchronize(ctx->stream);
        if (rc == TOS_OK && e != hipSuccess) rc = hip_fail(ctx, e, "tos_catalog_upload");
    } else {
        // sightlines before bad_off lie within [0, np) (their offsets were
        // checked), so every one of them packs in place
        std::vector<float4> pix((size_t)np);
        parallel_for(bad_off, 1024, [&](int64_t k) { pack(k, pix.data() + s->offsets[k]); });
        finite = fold(0, bad_off);
        if (finite && bad_off == nl)
            rc = upload(ctx, c->pix, pix.data(), pix.size() * sizeof(float4), "catalog pixels");
        if (rc == TOS_OK && finite && bad_off == nl) {
            const hipError_t e = hipStreamSynchronize(ctx->stream);
            if (e != hipSuccess) rc = hip_fail(ctx, e, "tos_catalog_upload");
        }
    }
    if (rc == TOS_OK && !finite) {
        set_err(ctx, "tos_catalog_upload: non-finite pixel (filter as HDF5Delta.cc:74 does)");
        rc = TOS_ERR_INVALID;
    }
    if (rc == TOS_OK && bad_off < nl) {
        set_err(ctx, "tos_catalog_upload: offsets must be non-decreasing");
        rc = TOS_ERR_INVALID;
    }
    if (rc != TOS_OK) {
        c->pix.release();
        delete c;
        return rc;
    }
    c->dmin = np > 0 ? dmin : 0;
    c->dmax = np > 0 ? dmax : 0;
    rc = upload(ctx, c->los, c->h_los.data(), c->h_los.size() * sizeof(LosInfo), "catalog sightlines");
    if (rc == TOS_OK) {
        hipError_t e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) rc = hip_fail(ctx, e, "tos_catalog_upload");
    }
    if (rc != TOS_OK) {
        c->pix.release();
        c->los.release();
        delete c;
        return rc;
    }
    *out = c;
    return TOS_OK;
}

extern "C" int tos_catalog_free(tos_ctx *ctx, tos_catalog *cat)
{
    if (!ctx || !cat) return TOS_ERR_INVALID;
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    cat->pix.release();
    cat->los.release();
    cat->trig.release();
    delete cat;
    return TOS_OK;
}

extern "C" int tos_catalog_info(const tos_catalog *cat, int64_t *n_los, int64_t *n_pix, int64_t *device_bytes)
{
    if (!cat) return TOS_ERR_INVALID;
    if (n_los) *n_los = cat->n_los;
    if (n_pix) *n_pix = cat->n_pix;
    if (device_bytes) *device_bytes = (int64_t)(cat->pix.bytes + cat->los.bytes + cat->trig.bytes);
    return TOS_OK;
}

// ---------------------------------------------------------------------------
// Sky bins and candidate pairs
// ---------------------------------------------------------------------------
extern "C" int tos_healpix_radec2pix(int32_t order, int64_t n, const double *ra, const double *dec, int32_t *out_pix)
{
    if (order < 0 || order > 13 || n < 0 || (n > 0 && (!ra || !dec || !out_pix))) return TOS_ERR_INVALID;
    const int64_t nside = (int64_t)1 << order;
    for (int64_t i = 0; i < n; ++i) out_pix[i] = (int32_t)healpix_radec2pix(nside, ra[i], dec[i]);
    return TOS_OK;
}

static int pairs_finish_upload(tos_ctx *ctx, const tos_catalog *cat, tos_pairs *p)
{
    p->cat = cat;
    p->n_rows = (int64_t)p->h.row_los.size();
    p->n_pairs = (int64_t)p->h.other.size();
    if (p->n_pairs >= (1ll << 31)) {
        set_err(ctx, "more than 2^31 candidate pairs in one job");
        return TOS_ERR_INVALID;
    }
    hipSetDevice(ctx->device);
    int rc = upload(ctx, p->other, p->h.other.data(), p->h.other.size() * sizeof(int32_t), "pair partners");
    if (rc == TOS_OK) rc = upload(ctx, p->csep, p->h.cos_sep.data(), p->h.cos_sep.size() * sizeof(double), "pair cos");
    if (rc == TOS_OK) {
        hipError_t e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) rc = hip_fail(ctx, e, "tos_pairs upload");
    }
    return rc;
}

extern "C" int tos_pairs_build(tos_ctx *ctx, const tos_catalog *cat, const tos_skybins *sky, double max_ang,
                               int nthreads, int on_device, tos_pairs **out)
{
    if (!ctx || !cat || !sky || !out) return TOS_ERR_INVALID;
    *out = nullptr;
    if (!(max_ang >= 0) || !std::isfinite(max_ang)) {
        set_err(ctx, "tos_pairs_build: max_ang must be finite and >= 0");
        return TOS_ERR_INVALID;
    }
    tos_pairs *p = new (std::nothrow) tos_pairs();
    if (!p) return TOS_ERR_NOMEM;
    int rc = TOS_OK;
    bool undecided = false, done = false;
    if (on_device) {
        rc = build_pairs_device(ctx, cat, *sky, max_ang, p, undecided);
        done = rc != TOS_OK || !undecided;
        if (!done) {   // fall back to the host search from a clean slate
            tos_pairs_free(ctx, p);
            p = new (std::nothrow) tos_pairs();
            if (!p) return TOS_ERR_NOMEM;
        }
    }
    if (!done) {
        tos_sightlines s{};
        s.n_los = cat->n_los;
        s.offsets = cat->h_offsets.data();
        s.ra = cat->h_ra.data();
        s.dec = cat->h_dec.data();
        std::string err;
        rc = build_pairs_host(s, *sky, max_ang, nthreads, p->h, err);
        if (rc != TOS_OK) set_err(ctx, err);
        if (rc == TOS_OK) rc = pairs_finish_upload(ctx, cat, p);
    }
    if (rc != TOS_OK) {
        tos_pairs_free(ctx, p);
        return rc;
    }
    *out = p;
    return TOS_OK;
}

extern "C" int tos_pairs_upload(tos_ctx *ctx, const tos_catalog *cat, const tos_pair_csr *csr, tos_pairs **out)
{
    if (!ctx || !cat || !csr || !out) return TOS_ERR_INVALID;
    *out = nullptr;
    const int64_t nr = csr->n_rows;
    if (nr < 0 || (nr > 0 && (!csr->row_los || !csr->row_ptr))) {
        set_err(ctx, "tos_pairs_upload: row arrays missing");
        return TOS_ERR_INVALID;
    }
    const int64_t np = nr > 0 ? csr->row_ptr[nr] : 0;
    if (np < 0 || (np > 0 && (!csr->other || !csr->cos_sep)) || (nr > 0 && csr->row_ptr[0] != 0)) {
        set_err(ctx, "tos_pairs_upload: pair arrays missing or row_ptr[0] != 0");
        return TOS_ERR_INVALID;
    }
    tos_pairs *p = new (std::nothrow) tos_pairs();
    if (!p) return TOS_ERR_NOMEM;
    HostPairs &h = p->h;
    h.row_los.assign(csr->row_los, csr->row_los + nr);
    if (csr->row_skybin) h.row_skybin.assign(csr->row_skybin, csr->row_skybin + nr);
    else h.row_skybin.assign((size_t)nr, 0);
    h.row_ptr.assign(csr->row_ptr, csr->row_ptr + nr + 1);
    if (nr == 0) h.row_ptr.assign(1, 0);
    h.other.assign(csr->other, csr->other + np);
    h.cos_sep.assign(csr->cos_sep, csr->cos_sep + np);
    // validate and fill stats
    std::vector<int32_t> keys(h.row_skybin);
    std::sort(keys.begin(), keys.end());
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    h.stats.n_skybins = (int64_t)keys.size();
    h.skybin_keys = keys;
    h.stats.n_rows = nr;
    for (int64_t r = 0; r < nr; ++r) {
        const int32_t i = h.row_los[(size_t)r];
        if (i < 0 || i >= cat->n_los || h.row_ptr[(size_t)r + 1] < h.row_ptr[(size_t)r]) {
            delete p;
            set_err(ctx, "tos_pairs_upload: bad row");
            return TOS_ERR_INVALID;
        }
        const int64_t ni = cat->h_los[(size_t)i].n;
        h.stats.num_pixels += (uint64_t)ni;
        for (int64_t q = h.row_ptr[(size_t)r]; q < h.row_ptr[(size_t)r + 1]; ++q) {
            const int32_t j = h.other[(size_t)q];
            if (j < 0 || j >= cat->n_los) {
                delete p;
                set_err(ctx, "tos_pairs_upload: partner index out of range");
                return TOS_ERR_INVALID;
            }
            h.stats.num_pixel_pairs += (uint64_t)(ni * cat->h_los[(size_t)j].n);
        }
    }
    h.stats.num_sightline_pairs = (uint64_t)np;
    h.stats.num_sightline_pairs_used = (uint64_t)np;
    int rc = pairs_finish_upload(ctx, cat, p);
    if (rc != TOS_OK) {
        tos_pairs_free(ctx, p);
        return rc;
    }
    *out = p;
    return TOS_OK;
}

extern "C" int tos_pairs_free(tos_ctx *ctx, tos_pairs *p)
{
    if (!ctx || !p) return TOS_ERR_INVALID;
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    p->other.release();
    p->csep.release();
    p->items.release();
    p->cost_prefix.release();
    p->theta.release();
    delete p;
    return TOS_OK;
}

namespace tos {

int pairs_from_host(tos_ctx *ctx, const tos_catalog *cat, HostPairs &&h, tos_pairs **out)
{
    *out = nullptr;
    tos_pairs *p = new (std::nothrow) tos_pairs();
    if (!p) return TOS_ERR_NOMEM;
    p->h = std::move(h);
    const int rc = pairs_finish_upload(ctx, cat, p);
    if (rc != TOS_OK) {
        tos_pairs_free(ctx, p);
        return rc;
    }
    *out = p;
    return TOS_OK;
}

static void detach(DevBuf &b, std::vector<void *> &dev)
{
    if (b.ptr) dev.push_back(b.ptr);
    b.ptr = nullptr;
    b.bytes = 0;
}

void catalog_detach(tos_catalog *cat, std::vector<void *> &dev)
{
    detach(cat->pix, dev);
    detach(cat->los, dev);
    detach(cat->trig, dev);
    delete cat;
}

void pairs_detach(tos_pairs *p, std::vector<void *> &dev)
{
    detach(p->other, dev);
    detach(p->csep, dev);
    detach(p->items, dev);
    detach(p->cost_prefix, dev);
    detach(p->theta, dev);
    delete p;
}

}  // namespace tos

extern "C" int tos_pairs_stats(const tos_pairs *p, tos_pair_stats *out)
{
    if (!p || !out) return TOS_ERR_INVALID;
    *out = p->h.stats;
    return TOS_OK;
}

extern "C" int tos_pairs_download(tos_ctx *ctx, const tos_pairs *p, int32_t *row_los, int32_t *row_skybin,
                                  int64_t *row_ptr, int32_t *other, double *cos_sep)
{
    if (!ctx || !p) return TOS_ERR_INVALID;
    if (other || cos_sep) {
        const int rc = ensure_host_pairs(ctx, const_cast<tos_pairs *>(p));
        if (rc != TOS_OK) return rc;
    }
    const HostPairs &h = p->h;
    if (row_los) std::copy(h.row_los.begin(), h.row_los.end(), row_los);
    if (row_skybin) std::copy(h.row_skybin.begin(), h.row_skybin.end(), row_skybin);
    if (row_ptr) std::copy(h.row_ptr.begin(), h.row_ptr.end(), row_ptr);
    if (other) std::copy(h.other.begin(), h.other.end(), other);
    if (cos_sep) std::copy(h.cos_sep.begin(), h.cos_sep.end(), cos_sep);
    return TOS_OK;
}

// ---------------------------------------------------------------------------
// Cosmology and self-test
// ---------------------------------------------------------------------------
extern "C" int tos_cosmo_distances(double omega_lambda, double omega_matter, int64_t n, const double *z,
                                   double *los_comoving, double *transverse)
{
    if (n < 0 || (n > 0 && !z)) return TOS_ERR_INVALID;
    if (omega_matter == 0) omega_matter = 1 - omega_lambda;   // h5healxi.cc:124
    for (int64_t i = 0; i < n; ++i) {
        if (!(z[i] > -1.0) || !std::isfinite(z[i])) return TOS_ERR_INVALID;
        if (los_comoving) los_comoving[i] = cosmo_los_comoving(omega_lambda, omega_matter, z[i]);
        if (transverse) transverse[i] = cosmo_transverse(omega_lambda, omega_matter, z[i]);
    }
    return TOS_OK;
}

extern "C" int tos_selftest_f64(tos_ctx *ctx, int64_t n, const double *a, const double *b, double *out_sqrt,
                                double *out_div, uint32_t *out_cvt)
{
    if (!ctx || n < 0 || (n > 0 && (!a || !b || !out_sqrt || !out_div || !out_cvt))) return TOS_ERR_INVALID;
    if (n == 0) return TOS_OK;
    hipSetDevice(ctx->device);
    const size_t nb = (size_t)n * sizeof(double);
    int rc;
    if ((rc = upload(ctx, ctx->scratch_a, a, nb, "selftest a")) != TOS_OK) return rc;
    if ((rc = upload(ctx, ctx->scratch_b, b, nb, "selftest b")) != TOS_OK) return rc;
    if (ctx->scratch_c.reserve(2 * nb) != TOS_OK || ctx->scratch_d.reserve((size_t)n * 4) != TOS_OK) {
        set_err(ctx, "selftest: device allocation failed");
        return TOS_ERR_NOMEM;
    }
    double *s = (double *)ctx->scratch_c.ptr;
    double *q = s + n;
    int e = launch_selftest((const double *)ctx->scratch_a.ptr, (const double *)ctx->scratch_b.ptr, s, q,
                            (uint32_t *)ctx->scratch_d.ptr, n, ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, (hipError_t)e, "selftest launch");
    hipMemcpyAsync(out_sqrt, s, nb, hipMemcpyDeviceToHost, ctx->stream);
    hipMemcpyAsync(out_div, q, nb, hipMemcpyDeviceToHost, ctx->stream);
    hipMemcpyAsync(out_cvt, ctx->scratch_d.ptr, (size_t)n * 4, hipMemcpyDeviceToHost, ctx->stream);
    hipError_t he = hipStreamSynchronize(ctx->stream);
    if (he != hipSuccess) return hip_fail(ctx, he, "selftest");
    return TOS_OK;
}

// ---------------------------------------------------------------------------
// Covariance of the sky-bin subsamples (XiEstimator.cc:60-97)
// ---------------------------------------------------------------------------
extern "C" int tos_subsample_covariance(tos_ctx *ctx, int64_t n_samples, int32_t nbins, const double *didj,
                                        const double *wgt, double *cov, double *mean, double *weight_sum)
{
    if (!ctx) return TOS_ERR_INVALID;
    if (n_samples < 0 || nbins <= 0 || nbins > (1 << 16) || (n_samples > 0 && (!didj || !wgt)) || !cov ||
        !weight_sum || n_samples > ((int64_t)1 << 32) / nbins) {
        set_err(ctx, "tos_subsample_covariance: invalid argument");
        return TOS_ERR_INVALID;
    }
    const size_t nb = (size_t)nbins;
    if (n_samples == 0) {
        std::memset(cov, 0, sizeof(double) * nb * nb);
        if (mean) std::memset(mean, 0, sizeof(double) * nb);
        *weight_sum = 0;
        return TOS_OK;
    }
    hipSetDevice(ctx->device);
    struct Scratch {   // released on every return
        DevBuf b[8];
        ~Scratch()
        {
            for (DevBuf &x : b) x.release();
        }
    } sc;
    DevBuf &d_didj = sc.b[0], &d_wgt = sc.b[1], &d_x = sc.b[2], &d_a = sc.b[3], &d_w = sc.b[4], &d_mu = sc.b[5],
           &d_ws = sc.b[6], &d_cov = sc.b[7];
    const size_t kn = (size_t)n_samples * nb;
    int rc;
    if ((rc = upload(ctx, d_didj, didj, kn * 8, "covariance didj")) != TOS_OK ||
        (rc = upload(ctx, d_wgt, wgt, kn * 8, "covariance wgt")) != TOS_OK)
        return rc;
    if (d_x.reserve(kn * 8) != TOS_OK || d_a.reserve(kn * 8) != TOS_OK ||
        d_w.reserve((size_t)n_samples * 8) != TOS_OK || d_mu.reserve(nb * 8) != TOS_OK || d_ws.reserve(8) != TOS_OK ||
        d_cov.reserve(nb * nb * 8) != TOS_OK) {
        set_err(ctx, "tos_subsample_covariance: device allocation failed");
        return TOS_ERR_NOMEM;
    }
    int e = launch_covariance((const double *)d_didj.ptr, (const double *)d_wgt.ptr, n_samples, nbins,
                              (double *)d_x.ptr, (double *)d_a.ptr, (double *)d_w.ptr, (double *)d_mu.ptr,
                              (double *)d_ws.ptr, (double *)d_cov.ptr, ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, (hipError_t)e, "covariance launch");
    hipError_t he = hipMemcpyAsync(cov, d_cov.ptr, nb * nb * 8, hipMemcpyDeviceToHost, ctx->stream);
    if (he == hipSuccess && mean) he = hipMemcpyAsync(mean, d_mu.ptr, nb * 8, hipMemcpyDeviceToHost, ctx->stream);
    if (he == hipSuccess) he = hipMemcpyAsync(weight_sum, d_ws.ptr, 8, hipMemcpyDeviceToHost, ctx->stream);
    if (he == hipSuccess) he = hipStreamSynchronize(ctx->stream);
    if (he != hipSuccess) return hip_fail(ctx, he, "covariance");
    return TOS_OK;
}

// Positive-definiteness and log-determinant of a covariance (the diagnostics
// XiEstimator.cc:78-86 prints): Cholesky factorisation C = L L^T in FP64,
// positive definite iff every pivot is > 0, log|C| = 2 sum log L_jj.  Rows
// are factorised in blocks of 64: the columns left of a block depend only
// on rows above it, so the block's rows are done in parallel; the block's
// own triangle follows row by row.
namespace {
double dot_prefix(const double *a, const double *b, int n)
{
    double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    int k = 0;
    for (; k + 4 <= n; k += 4) {
        s0 += a[k] * b[k];
        s1 += a[k + 1] * b[k + 1];
        s2 += a[k + 2] * b[k + 2];
        s3 += a[k + 3] * b[k + 3];
    }
    for (; k < n; ++k) s0 += a[k] * b[k];
    return (s0 + s1) + (s2 + s3);
}
}  // namespace

extern "C" int tos_covariance_check(int32_t n, const double *cov, int32_t *is_pos_def, double *log_det)
{
    if (n <= 0 || n > (1 << 16) || !cov || !is_pos_def) return TOS_ERR_INVALID;
    const size_t N = (size_t)n;
    std::vector<double> L;   // row-major, lower triangle used
    try {
        L.assign(N * N, 0.0);
    } catch (const std::bad_alloc &) {
        return TOS_ERR_NOMEM;
    }
    for (size_t i = 0; i < N; ++i)
        for (size_t j = 0; j <= i; ++j) L[i * N + j] = cov[i * N + j];
    std::vector<double> diag(N, 0.0);
    constexpr int kRows = 64;
    bool pd = true;
    double half_ld = 0;
    for (int r0 = 0; r0 < n && pd; r0 += kRows) {
        const int r1 = std::min(n, r0 + kRows);
        parallel_for(r1 - r0, 1, [&](int64_t t) {
            double *Li = &L[(size_t)(r0 + t) * N];
            for (int j = 0; j < r0; ++j) Li[j] = (Li[j] - dot_prefix(Li, &L[(size_t)j * N], j)) / diag[j];
        });
        for (int i = r0; i < r1 && pd; ++i) {
            double *Li = &L[(size_t)i * N];
            for (int j = r0; j < i; ++j) Li[j] = (Li[j] - dot_prefix(Li, &L[(size_t)j * N], j)) / diag[j];
            const double d = Li[i] - dot_prefix(Li, Li, i);
            if (!(d > 0) || !std::isfinite(d)) {
                pd = false;
                break;
            }
            diag[i] = std::sqrt(d);
            Li[i] = diag[i];
            half_ld += std::log(diag[i]);
        }
    }
    *is_pos_def = pd ? 1 : 0;
    if (log_det) *log_det = pd ? 2.0 * half_ld : std::nan("");
    return TOS_OK;
}

// ---------------------------------------------------------------------------
// 3D brute force
// ---------------------------------------------------------------------------
extern "C" int tos_brute3d(tos_ctx *ctx, int32_t mode, int64_t n, const double *x, const double *y, const double *z,
                           const double *d, const double *w, double min, double max, int32_t nbins, double *dsum,
                           double *wsum, uint64_t *counts, double *kernel_ms)
{
    if (!ctx) return TOS_ERR_INVALID;
    if ((mode != TOS_BRUTE_FP64 && mode != TOS_BRUTE_FP32) || n < 0 || n >= (1ll << 30) || nbins <= 0 ||
        nbins > (1 << 20) || !(max > min) || (n > 0 && (!x || !y || !z || !d || !w)) || !dsum || !wsum || !counts) {
        set_err(ctx, "tos_brute3d: invalid argument");
        return TOS_ERR_INVALID;
    }
    const int nslot = mode == TOS_BRUTE_FP64 ? nbins : nbins + 2;
    // histogram copies per workgroup: up to 16 while they fit in 48 KiB
    int copies = 16;
    while (copies > 1 && brute3d_lds_bytes(mode, nbins, copies) > 48 * 1024) copies /= 2;
    if ((int64_t)brute3d_lds_bytes(mode, nbins, copies) > ctx->max_lds_block) {
        set_err(ctx, "tos_brute3d: too many bins for one LDS histogram");
        return TOS_ERR_INVALID;
    }
    if (kernel_ms) *kernel_ms = 0;
    if (n < 2) {
        std::memset(dsum, 0, sizeof(double) * nslot);
        std::memset(wsum, 0, sizeof(double) * nslot);
        std::memset(counts, 0, sizeof(uint64_t) * nslot);
        return TOS_OK;
    }
    hipSetDevice(ctx->device);
    // packed points {x, y, z, delta} and weights, all resident (the
    // reference packs 20-byte float AoS tiles per chunk pair and copies
    // them for every tile pair, brute.cu:86-94,155-168)
    std::vector<double> pts((size_t)n * 4);
    for (int64_t i = 0; i < n; ++i) {
        pts[(size_t)i * 4 + 0] = x[i];
        pts[(size_t)i * 4 + 1] = y[i];
        pts[(size_t)i * 4 + 2] = z[i];
        pts[(size_t)i * 4 + 3] = d[i];
    }
    int rc = upload(ctx, ctx->scratch_a, pts.data(), pts.size() * sizeof(double), "brute3d points");
    if (rc == TOS_OK) rc = upload(ctx, ctx->scratch_c, w, (size_t)n * sizeof(double), "brute3d weights");
    if (rc != TOS_OK) return rc;
    // triangular tiles: 64-point i blocks against j runs of `chunk` points,
    // the run length chosen for about 2^16 tiles (at least 256, at most
    // 65536 points per run)
    const double pairs = 0.5 * (double)n * (double)(n - 1);
    int64_t chunk = (int64_t)std::ceil(pairs / 64.0 / 65536.0);
    chunk = std::max<int64_t>(256, std::min<int64_t>(65536, (chunk + 63) / 64 * 64));
    std::vector<int32_t> tiles;
    for (int64_t ib = 0; ib < n; ib += 64) {
        for (int64_t jb = ib + 1; jb < n; jb += chunk) {
            tiles.push_back((int32_t)ib);
            tiles.push_back((int32_t)jb);
        }
    }
    const int32_t n_tiles = (int32_t)(tiles.size() / 2);
    if ((rc = upload(ctx, ctx->scratch_b, tiles.data(), tiles.size() * sizeof(int32_t), "brute3d tiles")) != TOS_OK)
        return rc;
    // workgroups of 8 wavefronts, a grid-stride loop over the tiles; a
    // workgroup's u32 partial counts hold at most its pairs, so there are
    // enough workgroups that none takes 2^32 pairs
    const uint64_t tile_pairs = 64ull * (uint64_t)chunk;
    int64_t n_wg64 = std::max<int64_t>(1, std::min<int64_t>((int64_t)ctx->ncu * 4, (n_tiles + 7) / 8));
    n_wg64 = std::max<int64_t>(n_wg64, (int64_t)((uint64_t)n_tiles * tile_pairs / 0xFFFFFFFFull) + 1);
    n_wg64 = std::min<int64_t>(n_wg64, n_tiles);
    if ((uint64_t)((n_tiles + n_wg64 - 1) / n_wg64) * tile_pairs >= 0xFFFFFFFFull) {
        set_err(ctx, "tos_brute3d: too many points for the per-workgroup pair counts");
        return TOS_ERR_INVALID;
    }
    const int n_wg = (int)n_wg64;
    if (ctx->part_cnt.reserve((size_t)n_wg * nslot * 4) != TOS_OK ||
        ctx->part_sum.reserve((size_t)n_wg * nslot * 16) != TOS_OK || ctx->out_cnt.reserve((size_t)nslot * 8) != TOS_OK ||
        ctx->out_sum.reserve((size_t)nslot * 16) != TOS_OK) {
        set_err(ctx, "tos_brute3d: device allocation failed");
        return TOS_ERR_NOMEM;
    }
    const int n_seg = std::min(32, n_wg);
    if (ctx->seg_cnt.reserve((size_t)n_seg * nslot * 8) != TOS_OK ||
        ctx->seg_sum.reserve((size_t)n_seg * nslot * 16) != TOS_OK) {
        set_err(ctx, "tos_brute3d: device allocation failed");
        return TOS_ERR_NOMEM;
    }
    BruteArgs a{};
    a.pts = (const double4 *)ctx->scratch_a.ptr;
    a.w = (const double *)ctx->scratch_c.ptr;
    a.n = n;
    a.mn = min;
    a.mx = max;
    a.nbins = nbins;
    a.copies = copies;
    a.tiles = (const int32_t *)ctx->scratch_b.ptr;
    a.n_tiles = n_tiles;
    a.chunk = (int32_t)chunk;
    a.part_cnt = (uint32_t *)ctx->part_cnt.ptr;
    a.part_sum = (double *)ctx->part_sum.ptr;
    // a power-of-two bin width (FP64) or spacing (FP32) divides exactly as
    // a multiplication by its inverse; both must hold for the fast form
    const double bw = (max - min) / (double)nbins;
    const float sp = ((float)max - (float)min) / (float)nbins;
    int e2 = 0, e3 = 0;
    const bool pow2 = std::frexp(bw, &e2) == 0.5 && std::frexp(sp, &e3) == 0.5f && std::isfinite(bw) &&
                      std::isfinite(sp);
    a.inv_width = pow2 ? 1.0 / bw : 0.0;
    a.inv_spacing = pow2 ? 1.0f / sp : 0.0f;
    hipEventRecord(ctx->ev0, ctx->stream);
    int e = launch_brute3d(mode, pow2, a, n_wg, ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, (hipError_t)e, "brute3d launch");
    hipEventRecord(ctx->ev1, ctx->stream);
    e = launch_reduce((const uint32_t *)ctx->part_cnt.ptr, (const double *)ctx->part_sum.ptr, n_wg, nslot, 2,
                      (uint64_t *)ctx->out_cnt.ptr, (double *)ctx->out_sum.ptr, (uint32_t *)ctx->seg_cnt.ptr,
                      (double *)ctx->seg_sum.ptr, n_seg, ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, (hipError_t)e, "brute3d reduce");
    std::vector<double> sums((size_t)nslot * 2);
    hipMemcpyAsync(counts, ctx->out_cnt.ptr, (size_t)nslot * 8, hipMemcpyDeviceToHost, ctx->stream);
    hipMemcpyAsync(sums.data(), ctx->out_sum.ptr, (size_t)nslot * 16, hipMemcpyDeviceToHost, ctx->stream);
    hipError_t he = hipStreamSynchronize(ctx->stream);
    if (he != hipSuccess) return hip_fail(ctx, he, "brute3d");
    for (int b = 0; b < nslot; ++b) {
        dsum[b] = sums[(size_t)2 * b];
        wsum[b] = sums[(size_t)2 * b + 1];
    }
    float ms = 0;
    hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
    if (kernel_ms) *kernel_ms = ms;
    return TOS_OK;
}
