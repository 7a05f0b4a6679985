// tos_capi_impl.h -- definitions of the opaque C-ABI handles.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <atomic>
#include <mutex>
#include <string>
#include <vector>

#include "tos_internal.h"

#define TOS_STR2(x) #x
#define TOS_STR(x) TOS_STR2(x)

namespace tos {

// Process-wide identity of catalogues and pair lists (an address can be
// reused after a free; an id is not).
inline uint64_t next_uid()
{
    static std::atomic<uint64_t> c{1};
    return c.fetch_add(1);
}

// Grow-only device allocation.
struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    int reserve(size_t n);
    void release();
};

// Launch shape chosen by the autotuner for one kernel variant.
struct TuneEntry {
    int coord, prec;
    bool fixed, simple;
    int block, blocks_per_cu;
    int wg_per_slot;          // workgroups per resident slot (0: default)
    int step_cap;             // partner pixels per work item (0: default)
};

struct BatchState;                 // tos_batch.cpp
void batch_release(tos_ctx *ctx);

}  // namespace tos

struct tos_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipDeviceProp_t prop{};
    int ncu = 0;
    int64_t max_lds_block = 65536;
    int64_t max_lds_cu = 65536;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::mutex err_mu;
    std::string err;
    // reusable device scratch
    tos::DevBuf part_cnt, part_sum, seg_cnt, seg_sum, out_cnt, out_sum, recheck;
    tos::DevBuf scratch_a, scratch_b, scratch_c, scratch_d, theta, items, wgptr;
    // autotuned launch shapes, and the state of the last batched launch
    std::vector<tos::TuneEntry> tuned;
    tos::BatchState *batch = nullptr;
    tos::BatchState *skybin = nullptr;   // per-sky-bin passes of tos_xi_run
    // two pinned staging chunks for catalogue uploads (packed on the host
    // while the other one is copied), allocated at the first upload
    void *stage[2] = {nullptr, nullptr};
    hipEvent_t stage_ev[2] = {nullptr, nullptr};
    // device address of this device's static scratch array of
    // tos_xi_run_noalloc (tos_noalloc.hip), set by tos_ctx_create
    void *noalloc_scratch = nullptr;
    size_t noalloc_bytes = 0;
};

struct tos_catalog {
    uint64_t uid = tos::next_uid();
    int64_t n_los = 0, n_pix = 0;
    double dmin = 0, dmax = 0;
    double lmax_abs = 0;   // largest |loglam|
    bool d_sorted = true, l_sorted = true;   // pixels ascending in every sightline
    tos::DevBuf pix;    // float4 {distance, loglam, weight, weight*value}
    tos::DevBuf los;    // LosInfo
    tos::DevBuf trig;   // double4 {sin dec, cos dec, sin ra, cos ra} (device search)
    std::vector<tos::LosInfo> h_los;
    std::vector<int64_t> h_offsets;
    std::vector<double> h_ra, h_dec;
};

struct tos_pairs {
    uint64_t uid = tos::next_uid();
    const tos_catalog *cat = nullptr;
    int64_t n_rows = 0, n_pairs = 0;
    tos::HostPairs h;
    tos::DevBuf other, csep;
    // false after a device search: h.other / h.cos_sep are copied back only
    // when something on the host needs them (ensure_host_pairs)
    bool pairs_on_host = true;
    // work items, built on first use
    bool items_ready = false;
    int32_t items_step_cap = 0;
    bool items_transposed = false;
    uint64_t items_gen = 0;         // bumped whenever the items are rebuilt
    int64_t n_items = 0;
    std::vector<uint64_t> item_cost;
    std::vector<int32_t> row_item_ptr;   // items of row r: [ptr[r], ptr[r+1])
    tos::DevBuf items;
    // inclusive prefix sums of item_cost on the device (n_items u64), for
    // workgroup ranges cut on the device (tos_xi_run_noalloc)
    tos::DevBuf cost_prefix;
    // Observing-grid theta bins, cached for the last grid
    tos::DevBuf theta;
    double theta_key[3] = {0, 0, 0};
    int32_t theta_n = -1;
};

namespace tos {

// Launch geometry of one pair-kernel launch.
struct XiLaunch {
    int block = 256;
    int bpc = 1;
    int n_wg = 0;
    size_t lds = 0;
    int nsum = 2;
    int nbins = 0;
    int n_seg = 1;
};

// tos_xi.cpp helpers shared with the batched path.
// defer: when given, the temporary device buffers are handed to it instead
// of freed (hipFree waits for the whole device; the job server frees them
// when no batch runs).
// step_cap: largest partner-pixel run of one work item (0: kItemStepCap).
int ensure_items(tos_ctx *ctx, tos_pairs *p, bool transpose, std::vector<void *> *defer = nullptr,
                 int32_t step_cap = 0);
// The tuned work-item size of a grid's kernel variant, or kItemStepCap.
int32_t tuned_step_cap(const tos_ctx *ctx, const tos_grid &g, int prec);
int ensure_theta(tos_ctx *ctx, tos_pairs *p, const tos_grid &g);
void fill_args(const tos_catalog *cat, const tos_pairs *p, const tos_grid &g, bool window, XiKernelArgs &a);
void grid_variant(const tos_grid &g, bool &fixed, bool &simple);
const TuneEntry *find_tuned(const tos_ctx *ctx, int coord, int prec, bool fixed, bool simple);
int max_blocks_per_cu(const tos_ctx *ctx, size_t lds, int block);
// band_rows > 0 (fewer than the grid's axis-0 bins): the launch histograms
// bands of that many axis-0 rows (tos_xi_run on grids too large for one
// LDS histogram); its LDS and partial sizes are a band's.
int plan_launch(tos_ctx *ctx, const tos_xi_opts *opts, int prec, int accum, const tos_grid &g, size_t n_items,
                XiLaunch &L, uint32_t band_rows = 0);
// Axis-0 rows per LDS histogram of a grid with this accumulator set: all
// of them when the grid fits, else the most that fit (0: not even one).
uint32_t band_rows_for(const tos_ctx *ctx, int accum, const tos_grid &g);
// Resolves this device's static scratch array of tos_xi_run_noalloc
// (tos_noalloc.hip); the device must be current.
int noalloc_setup(tos_ctx *ctx);
// tos_xi_run with per_skybin = 1 (tos_batch.cpp).
int xi_run_per_skybin(tos_ctx *ctx, const tos_catalog *cat, tos_pairs *p, const tos_grid &g, int prec, int flags,
                      tos_xi_result *out);

// A pair list from a host candidate search (moved in), uploaded.
int pairs_from_host(tos_ctx *ctx, const tos_catalog *cat, HostPairs &&h, tos_pairs **out);
// Delete the host object and hand its device allocations to the caller
// (to be freed later, when no kernel is running: hipFree synchronises).
void catalog_detach(tos_catalog *cat, std::vector<void *> &dev);
void pairs_detach(tos_pairs *p, std::vector<void *> &dev);

void set_err(tos_ctx *ctx, const std::string &m);
int hip_fail(tos_ctx *ctx, hipError_t e, const char *where);
int upload(tos_ctx *ctx, DevBuf &b, const void *src, size_t n, const char *what);
bool grid_ok(tos_ctx *ctx, const tos_grid *g);
// The device candidate search in phases (tos_dsearch.cpp): prepare (host
// only), count (device; fills the rows and statistics of h and `total`),
// fill (device; the kept partners and cos values into other / csep).
// undecided: a plate test fell within rounding of its bound; nothing was
// built, use the host search.
struct DeviceSearch {
    bool plates = false;
    int64_t n_bins = 0;
    int64_t total = 0;
    std::vector<int32_t> occupied, prim, prim_cell, cell_ptr, members, nbr_ptr, nbr, pbin;
    std::vector<double> trig, ptrig;
    std::vector<LosInfo> los;
    SearchArgs args{};
    DevBuf d_trig, d_los, d_prim, d_pcell, d_cptr, d_mem, d_nptr, d_nbr, d_ptrig, d_pbin, d_und, d_cnt, d_cons,
        d_pp, d_off;
    int prepare(const tos_sightlines &s, const tos_skybins &sky, double max_ang, std::string &err);
    int count(tos_ctx *ctx, HostPairs &h, bool &undecided);
    int fill(tos_ctx *ctx, DevBuf &other, DevBuf &csep);
    std::array<DevBuf *, 15> bufs();
    int64_t device_bytes();
    void release();
    void detach(std::vector<void *> &dev);   // hand the device tables over, unfreed
};

// Candidate search on the device (HEALPix or plates), the three phases in a
// row; undecided as above.
int build_pairs_device(tos_ctx *ctx, const tos_catalog *cat, const tos_skybins &sky, double max_ang, tos_pairs *p,
                       bool &undecided);
// Copies the pair list of a device search back to p->h (no-op otherwise).
int ensure_host_pairs(tos_ctx *ctx, tos_pairs *p);

}  // namespace tos
