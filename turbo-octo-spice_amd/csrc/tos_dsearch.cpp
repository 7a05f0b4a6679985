// tos_dsearch.cpp -- host side of the device candidate search
// (XiEstimator.cc:121-159 on the GPU, kernels in tos_search.hip), in three
// phases so that a caller can size the result before committing memory:
//   prepare  host only: the sky index, the primaries in row order, members
//            and neighbour bins (HEALPix) or plates; no device work
//   count    uploads those tables, runs the count pass, builds the rows,
//            row offsets and statistics on the host: the pair count is known
//   fill     writes the kept partners and cos values into the caller's
//            device buffers (the resident pair list)
// The job server counts before admission and fills after it;
// tos_pairs_build runs the three in a row.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>

#include "tos_capi_impl.h"

namespace tos {

std::array<DevBuf *, 15> DeviceSearch::bufs()
{
    return {&d_trig, &d_los, &d_prim, &d_pcell, &d_cptr, &d_mem, &d_nptr, &d_nbr, &d_ptrig, &d_pbin, &d_und, &d_cnt,
            &d_cons, &d_pp, &d_off};
}

void DeviceSearch::release()
{
    for (DevBuf *b : bufs()) b->release();
}

void DeviceSearch::detach(std::vector<void *> &dev)
{
    for (DevBuf *b : bufs()) {
        if (b->ptr) dev.push_back(b->ptr);
        b->ptr = nullptr;
        b->bytes = 0;
    }
}

int64_t DeviceSearch::device_bytes()
{
    int64_t n = 0;
    for (const DevBuf *b : bufs()) n += (int64_t)b->bytes;
    return n;
}

int DeviceSearch::prepare(const tos_sightlines &s, const tos_skybins &sky, double max_ang, std::string &err)
{
    plates = sky.kind == TOS_SKY_PLATES;
    if (!plates && (sky.kind != TOS_SKY_HEALPIX || sky.order < 0 || sky.order > 13)) {
        err = "tos_pairs_build: HEALPix order must be in [0, 13]";
        return TOS_ERR_INVALID;
    }
    if (plates && (!sky.plate_of_los || (sky.n_plates > 0 && (!sky.plate_id || !sky.plate_ra || !sky.plate_dec)))) {
        err = "tos_pairs_build: plate arrays missing";
        return TOS_ERR_INVALID;
    }
    SkyIndex ix;
    sky_index(s, sky, max_ang, ix);
    n_bins = (int64_t)ix.occupied.size();
    occupied = ix.occupied;
    // plates, in ascending id: trigonometry and occupied-bin index
    const int64_t npl = plates ? (int64_t)ix.plate_order.size() : 0;
    ptrig.assign((size_t)npl * 4, 0.0);
    pbin.assign((size_t)npl, -1);
    for (int64_t q = 0; q < npl; ++q) {
        const Trig &t = ix.plate_trig[(size_t)q];
        ptrig[(size_t)q * 4 + 0] = t.sd;
        ptrig[(size_t)q * 4 + 1] = t.cd;
        ptrig[(size_t)q * 4 + 2] = t.sr;
        ptrig[(size_t)q * 4 + 3] = t.cr;
        auto it = ix.slot.find(sky.plate_id[ix.plate_order[(size_t)q]]);
        if (it != ix.slot.end()) pbin[(size_t)q] = it->second;
    }
    // primaries in row order (bins in first-insertion order, members in
    // sightline order), members and neighbours per bin
    prim.clear();
    prim_cell.clear();
    members.clear();
    nbr.clear();
    cell_ptr.assign(1, 0);
    nbr_ptr.assign(1, 0);
    for (int64_t b = 0; b < n_bins; ++b) {
        for (int32_t i : ix.members[(size_t)b]) {
            prim.push_back(i);
            prim_cell.push_back((int32_t)b);
            members.push_back(i);
        }
        cell_ptr.push_back((int32_t)members.size());
        // neighbour lists exist for HEALPix bins only (plates are tested per
        // primary in the kernel)
        if (!plates) nbr.insert(nbr.end(), ix.nbr[(size_t)b].begin(), ix.nbr[(size_t)b].end());
        nbr_ptr.push_back((int32_t)nbr.size());
    }
    trig.assign((size_t)s.n_los * 4, 0.0);
    los.assign((size_t)s.n_los, LosInfo{});
    for (int64_t i = 0; i < s.n_los; ++i) {
        const Trig &t = ix.trig[(size_t)i];
        trig[(size_t)i * 4 + 0] = t.sd;
        trig[(size_t)i * 4 + 1] = t.cd;
        trig[(size_t)i * 4 + 2] = t.sr;
        trig[(size_t)i * 4 + 3] = t.cr;
        los[(size_t)i].off = (int32_t)s.offsets[i];
        los[(size_t)i].n = (int32_t)(s.offsets[i + 1] - s.offsets[i]);
    }
    args = SearchArgs{};
    args.cos_max = std::cos(max_ang);   // XiEstimator.cc:25
    args.n_prim = (int64_t)prim.size();
    args.plates = plates ? 1 : 0;
    args.n_plate = (int32_t)npl;
    if (plates) {
        // PlateBins.h:78-81: |acos(c) - 1.5 deg| < max_ang, i.e. acos(c) in
        // (a0 - R, a0 + R); acos falls with c, so c in (cos(a0 + R), cos(a0 - R))
        const double deg2rad = std::atan(1.0) * 4 / 180.0;   // constants.h:12-13
        const double a0 = 1.5 * deg2rad, R = max_ang;
        args.has_up = (a0 + R <= std::acos(-1.0)) ? 1 : 0;
        args.c_up = args.has_up ? std::cos(a0 + R) : -2.0;
        args.has_lo = (a0 - R >= 0.0) ? 1 : 0;
        args.c_lo = args.has_lo ? std::cos(a0 - R) : 2.0;
        args.margin = 1e-12;
    }
    return TOS_OK;
}

int DeviceSearch::count(tos_ctx *ctx, HostPairs &h, bool &undecided)
{
    undecided = false;
    const int64_t np = args.n_prim;
    int rc;
    if ((rc = upload(ctx, d_trig, trig.data(), trig.size() * 8, "search trig")) != TOS_OK ||
        (rc = upload(ctx, d_los, los.data(), los.size() * sizeof(LosInfo), "search sightlines")) != TOS_OK ||
        (rc = upload(ctx, d_prim, prim.data(), prim.size() * 4, "search primaries")) != TOS_OK ||
        (rc = upload(ctx, d_pcell, prim_cell.data(), prim_cell.size() * 4, "search primary bins")) != TOS_OK ||
        (rc = upload(ctx, d_cptr, cell_ptr.data(), cell_ptr.size() * 4, "search bin members")) != TOS_OK ||
        (rc = upload(ctx, d_mem, members.data(), members.size() * 4, "search members")) != TOS_OK ||
        (rc = upload(ctx, d_nptr, nbr_ptr.data(), nbr_ptr.size() * 4, "search neighbours")) != TOS_OK ||
        (rc = upload(ctx, d_nbr, nbr.data(), nbr.size() * 4, "search neighbour bins")) != TOS_OK ||
        (rc = upload(ctx, d_ptrig, ptrig.data(), ptrig.size() * 8, "search plate trig")) != TOS_OK ||
        (rc = upload(ctx, d_pbin, pbin.data(), pbin.size() * 4, "search plate bins")) != TOS_OK)
        return rc;
    const int32_t zero = 0;
    if ((rc = upload(ctx, d_und, &zero, 4, "search flag")) != TOS_OK) return rc;
    if (d_cnt.reserve((size_t)std::max<int64_t>(1, np) * 4) != TOS_OK ||
        d_cons.reserve((size_t)std::max<int64_t>(1, np) * 8) != TOS_OK ||
        d_pp.reserve((size_t)std::max<int64_t>(1, np) * 8) != TOS_OK ||
        d_off.reserve((size_t)std::max<int64_t>(1, np) * 8) != TOS_OK) {
        set_err(ctx, "device allocation failed for the candidate search");
        return TOS_ERR_NOMEM;
    }
    SearchArgs &a = args;
    a.trig = (const double4 *)d_trig.ptr;
    a.los = (const LosInfo *)d_los.ptr;
    a.prim = (const int32_t *)d_prim.ptr;
    a.prim_cell = (const int32_t *)d_pcell.ptr;
    a.cell_ptr = (const int32_t *)d_cptr.ptr;
    a.members = (const int32_t *)d_mem.ptr;
    a.nbr_ptr = (const int32_t *)d_nptr.ptr;
    a.nbr = (const int32_t *)d_nbr.ptr;
    a.plate_trig = (const double4 *)d_ptrig.ptr;
    a.plate_bin = (const int32_t *)d_pbin.ptr;
    a.undecided = (int32_t *)d_und.ptr;
    a.count = (int32_t *)d_cnt.ptr;
    a.considered = (uint64_t *)d_cons.ptr;
    a.pixel_pairs = (uint64_t *)d_pp.ptr;
    int e = launch_search(a, false, ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, (hipError_t)e, "candidate count launch");
    std::vector<int32_t> cnt((size_t)np);
    std::vector<uint64_t> cons((size_t)np), pp((size_t)np);
    hipError_t he = hipSuccess;
    if (np > 0) {
        he = hipMemcpyAsync(cnt.data(), d_cnt.ptr, (size_t)np * 4, hipMemcpyDeviceToHost, ctx->stream);
        if (he == hipSuccess)
            he = hipMemcpyAsync(cons.data(), d_cons.ptr, (size_t)np * 8, hipMemcpyDeviceToHost, ctx->stream);
        if (he == hipSuccess)
            he = hipMemcpyAsync(pp.data(), d_pp.ptr, (size_t)np * 8, hipMemcpyDeviceToHost, ctx->stream);
    }
    int32_t und = 0;
    if (he == hipSuccess) he = hipMemcpyAsync(&und, d_und.ptr, 4, hipMemcpyDeviceToHost, ctx->stream);
    if (he == hipSuccess) he = hipStreamSynchronize(ctx->stream);
    if (he != hipSuccess) return hip_fail(ctx, he, "candidate count");
    if (und) {
        // a plate centre within rounding of the annulus edge: leave the
        // decision to the host search (acos itself)
        undecided = true;
        return TOS_OK;
    }
    // rows: primaries with at least one kept partner
    h = HostPairs();
    std::vector<int64_t> off((size_t)np, 0);
    int64_t tot = 0;
    h.row_ptr.assign(1, 0);
    for (int64_t r = 0; r < np; ++r) {
        off[(size_t)r] = tot;
        h.stats.num_sightline_pairs += cons[(size_t)r];
        h.stats.num_pixel_pairs += pp[(size_t)r];
        h.stats.num_pixels += (uint64_t)los[(size_t)prim[(size_t)r]].n;
        if (cnt[(size_t)r] > 0) {
            tot += cnt[(size_t)r];
            h.row_los.push_back(prim[(size_t)r]);
            h.row_skybin.push_back(occupied[(size_t)prim_cell[(size_t)r]]);
            h.row_ptr.push_back(tot);
        }
    }
    if (tot >= (1ll << 31)) {
        set_err(ctx, "more than 2^31 candidate pairs in one job");
        return TOS_ERR_INVALID;
    }
    h.stats.num_sightline_pairs_used = (uint64_t)tot;
    h.stats.n_rows = (int64_t)h.row_los.size();
    h.stats.n_skybins = n_bins;
    h.skybin_keys = occupied;
    std::sort(h.skybin_keys.begin(), h.skybin_keys.end());
    total = tot;
    if ((rc = upload(ctx, d_off, off.data(), off.size() * 8, "search offsets")) != TOS_OK) return rc;
    return TOS_OK;
}

int DeviceSearch::fill(tos_ctx *ctx, DevBuf &other, DevBuf &csep)
{
    if (other.reserve((size_t)std::max<int64_t>(1, total) * 4) != TOS_OK ||
        csep.reserve((size_t)std::max<int64_t>(1, total) * 8) != TOS_OK) {
        set_err(ctx, "device allocation failed for the candidate pairs");
        return TOS_ERR_NOMEM;
    }
    SearchArgs a = args;
    a.offset = (const int64_t *)d_off.ptr;
    a.other = (int32_t *)other.ptr;
    a.csep = (double *)csep.ptr;
    if (total > 0) {
        const int e = launch_search(a, true, ctx->stream);
        if (e != hipSuccess) return hip_fail(ctx, (hipError_t)e, "candidate fill launch");
    }
    const hipError_t he = hipStreamSynchronize(ctx->stream);
    if (he != hipSuccess) return hip_fail(ctx, he, "candidate fill");
    return TOS_OK;
}

// The three phases in a row for tos_pairs_build.  The pair list stays on the
// device: the work items are built there too, and ensure_host_pairs copies
// it back only if the host needs it.
int build_pairs_device(tos_ctx *ctx, const tos_catalog *cat, const tos_skybins &sky, double max_ang, tos_pairs *p,
                       bool &undecided)
{
    undecided = false;
    tos_sightlines s{};
    s.n_los = cat->n_los;
    s.offsets = cat->h_offsets.data();
    s.ra = cat->h_ra.data();
    s.dec = cat->h_dec.data();
    DeviceSearch ds;
    std::string err;
    int rc = ds.prepare(s, sky, max_ang, err);
    if (rc != TOS_OK) {
        set_err(ctx, err);
        return rc;
    }
    rc = ds.count(ctx, p->h, undecided);
    if (rc == TOS_OK && !undecided) rc = ds.fill(ctx, p->other, p->csep);
    ds.release();
    if (rc != TOS_OK || undecided) return rc;
    p->pairs_on_host = false;
    p->cat = cat;
    p->n_rows = p->h.stats.n_rows;
    p->n_pairs = ds.total;
    return TOS_OK;
}

}  // namespace tos
