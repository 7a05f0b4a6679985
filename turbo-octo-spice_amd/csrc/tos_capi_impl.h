// tos_capi_impl.h -- definitions of the opaque C-ABI handles.
#pragma once

#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <vector>

#include "tos_internal.h"

#define TOS_STR2(x) #x
#define TOS_STR(x) TOS_STR2(x)

namespace tos {

// Grow-only device allocation.
struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    int reserve(size_t n);
    void release();
};

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
};

struct tos_catalog {
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
    const tos_catalog *cat = nullptr;
    int64_t n_rows = 0, n_pairs = 0;
    tos::HostPairs h;
    tos::DevBuf other, csep;
    // work items, built on first use
    bool items_ready = false;
    int32_t items_step_cap = 0;
    bool items_transposed = false;
    std::vector<tos::XiItem> h_items;
    std::vector<uint64_t> item_cost;
    tos::DevBuf items;
    // Observing-grid theta bins, cached for the last grid
    tos::DevBuf theta;
    double theta_key[3] = {0, 0, 0};
    int32_t theta_n = -1;
};

namespace tos {

void set_err(tos_ctx *ctx, const std::string &m);
int hip_fail(tos_ctx *ctx, hipError_t e, const char *where);
int upload(tos_ctx *ctx, DevBuf &b, const void *src, size_t n, const char *what);
bool grid_ok(tos_ctx *ctx, const tos_grid *g);
int build_pairs_device(tos_ctx *ctx, const tos_catalog *cat, int order, double max_ang, tos_pairs *p);

}  // namespace tos
