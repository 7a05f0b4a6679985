// tos_noalloc.hip -- tos_xi_run_noalloc: the estimator's pixel-pair pass
// without any allocation, for callers that cannot allocate device memory
// (code under HIP graph capture, tight serving loops).
//
// Everything the pass needs beyond its inputs and outputs -- the workgroup
// item ranges, the per-workgroup partial histograms, the reduction's
// segment sums, the interleaved result sums and two counters -- lives in
// one file-scope device array, `scratch`, of kScratchBytes per device.
// tos_ctx_create resolves this device's address of it once
// (noalloc_setup); host code only ever uses that address.  All calls on
// one device share the array: calls that overlap on one device race.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "tos_capi_impl.h"

using namespace tos;

namespace tos {
namespace {

constexpr size_t kScratchBytes = (size_t)256 << 20;
constexpr size_t kAlign = 256;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

size_t align_up(size_t x) { return (x + kAlign - 1) & ~(kAlign - 1); }

// Workgroup w takes items [ptr[w], ptr[w + 1]): ptr[w] is the number of
// items whose inclusive cost prefix is at most w / n_wg of the total, so
// the ranges are contiguous and of about equal cost.
__global__ void wg_ranges_kernel(const uint64_t *__restrict__ prefix, int64_t n_items, int32_t n_wg,
                                 int32_t *__restrict__ ptr)
{
    const int32_t w = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
    if (w > n_wg) return;
    if (w == 0) {
        ptr[0] = 0;
        return;
    }
    if (w == n_wg) {
        ptr[w] = (int32_t)n_items;
        return;
    }
    const uint64_t total = prefix[n_items - 1];
    const uint64_t target = (uint64_t)((double)total * (double)w / (double)n_wg);
    int64_t lo = 0, hi = n_items;   // first k with prefix[k] > target
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (prefix[mid] <= target)
            lo = mid + 1;
        else
            hi = mid;
    }
    ptr[w] = (int32_t)lo;
}

// [nbins][2] interleaved sums -> the caller's didj and wgt arrays.
__global__ void split_sums_kernel(const double *__restrict__ sums, int32_t nbins, double *__restrict__ didj,
                                  double *__restrict__ wgt)
{
    const int32_t b = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
    if (b >= nbins) return;
    didj[b] = sums[2 * (size_t)b];
    wgt[b] = sums[2 * (size_t)b + 1];
}

__global__ void zero_outputs_kernel(uint64_t *cnt, double *didj, double *wgt, int32_t nbins)
{
    const int32_t b = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
    if (b >= nbins) return;
    cnt[b] = 0;
    didj[b] = 0.0;
    wgt[b] = 0.0;
}

int fail(tos_ctx *ctx, int code, const char *msg)
{
    set_err(ctx, std::string("tos_xi_run_noalloc: ") + msg);
    return code;
}

}  // namespace

int noalloc_setup(tos_ctx *ctx)
{
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return hip_fail(ctx, e != hipSuccess ? e : hipErrorInvalidSymbol,
                                               "static scratch of tos_xi_run_noalloc");
    ctx->noalloc_scratch = p;
    ctx->noalloc_bytes = kScratchBytes;
    return TOS_OK;
}

}  // namespace tos

extern "C" int tos_xi_run_noalloc(tos_ctx *ctx, const tos_catalog *cat, const tos_pairs *pairs,
                                  const tos_grid *grid, int32_t precision, uint64_t *num_pairs, double *didj,
                                  double *wgt, void *stream)
{
    if (!ctx || !cat || !pairs) return TOS_ERR_INVALID;
    if (!num_pairs || !didj || !wgt) return fail(ctx, TOS_ERR_INVALID, "result arrays missing");
    if (!grid_ok(ctx, grid)) return TOS_ERR_INVALID;
    if (precision != TOS_FP64 && precision != TOS_FP32C) return fail(ctx, TOS_ERR_INVALID, "unknown precision");
    const tos_pairs *p = pairs;
    if (p->cat != cat) return fail(ctx, TOS_ERR_INVALID, "the pair list was built for another catalogue");
    if (!ctx->noalloc_scratch) return fail(ctx, TOS_ERR_DEVICE, "the context has no static scratch");
    if (!p->items_ready || !p->items_transposed)
        return fail(ctx, TOS_ERR_INVALID,
                    "the pair list's work items are not built: run tos_xi_run on it once (without "
                    "TOS_XI_NO_TRANSPOSE) before the allocation-free calls");
    const tos_grid g = *grid;
    if (g.type == TOS_OBSERVING &&
        !(p->theta_n == g.nbins[1] && p->theta_key[0] == g.min[1] && p->theta_key[1] == g.width[1]))
        return fail(ctx, TOS_ERR_INVALID,
                    "the pair list has no theta bins for this Observing grid: run tos_xi_run on it with "
                    "this grid once first");
    hipSetDevice(ctx->device);
    hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
    const int32_t nbins = g.nbins[0] * g.nbins[1] * g.nbins[2];
    const int64_t n_items = p->n_items;
    if (n_items == 0) {
        hipLaunchKernelGGL(zero_outputs_kernel, dim3((nbins + 255) / 256), dim3(256), 0, st, num_pairs, didj, wgt,
                           nbins);
        const hipError_t e = hipGetLastError();
        return e == hipSuccess ? TOS_OK : hip_fail(ctx, e, "tos_xi_run_noalloc");
    }

    XiLaunch L;
    int rc = plan_launch(ctx, nullptr, precision, TOS_ACC_CORE, g, (size_t)n_items, L);
    if (rc != TOS_OK) return rc;
    // scratch layout: counters | workgroup ranges | partial counts | partial
    // sums | segment counts | segment sums | result sums
    const size_t nb = (size_t)nbins;
    auto layout = [&](int n_wg, int n_seg, size_t off[7]) {
        size_t o = 0;
        off[0] = o; o = align_up(o + 16);
        off[1] = o; o = align_up(o + (size_t)(n_wg + 1) * 4);
        off[2] = o; o = align_up(o + (size_t)n_wg * nb * 4);
        off[3] = o; o = align_up(o + (size_t)n_wg * nb * 16);
        off[4] = o; o = align_up(o + (size_t)n_seg * nb * 8);
        off[5] = o; o = align_up(o + (size_t)n_seg * nb * 16);
        off[6] = o; o = align_up(o + nb * 16);
        return o;
    };
    size_t off[7];
    int n_wg = L.n_wg;
    if (layout(1, 1, off) > kScratchBytes)
        return fail(ctx, TOS_ERR_INVALID, "grid has too many bins for the 256 MiB static scratch");
    // as many workgroups as the plan asks for and the scratch holds
    {
        int lo = 1, hi = n_wg;
        while (lo < hi) {
            const int mid = lo + (hi - lo + 1) / 2;
            if (layout(mid, std::min(32, mid), off) <= kScratchBytes)
                lo = mid;
            else
                hi = mid - 1;
        }
        n_wg = lo;
    }
    const int n_seg = std::min(32, n_wg);
    layout(n_wg, n_seg, off);
    // a workgroup's u32 partial counts hold at most its item costs (pixel
    // pairs before the windows): refuse passes that could wrap them
    uint64_t total = 0, max_item = 0;
    for (uint64_t c : p->item_cost) {
        total += c;
        max_item = std::max(max_item, c);
    }
    if (total / (uint64_t)n_wg + 2 * max_item >= (1ull << 32))
        return fail(ctx, TOS_ERR_INVALID, "the pass is too large for one allocation-free launch (a workgroup "
                                          "would count 2^32 pixel pairs)");

    char *base = (char *)ctx->noalloc_scratch;
    unsigned long long *counters = (unsigned long long *)(base + off[0]);
    int32_t *wgptr = (int32_t *)(base + off[1]);
    uint32_t *part_cnt = (uint32_t *)(base + off[2]);
    double *part_sum = (double *)(base + off[3]);
    uint32_t *seg_cnt = (uint32_t *)(base + off[4]);
    double *seg_sum = (double *)(base + off[5]);
    double *sums = (double *)(base + off[6]);

    int e = launch_zero_u64(counters, 2, st);
    if (e != hipSuccess) return hip_fail(ctx, (hipError_t)e, "tos_xi_run_noalloc counters");
    hipLaunchKernelGGL(wg_ranges_kernel, dim3((n_wg + 1 + 255) / 256), dim3(256), 0, st,
                       (const uint64_t *)p->cost_prefix.ptr, n_items, n_wg, wgptr);
    hipError_t he = hipGetLastError();
    if (he != hipSuccess) return hip_fail(ctx, he, "tos_xi_run_noalloc workgroup ranges");
    XiKernelArgs a;
    fill_args(cat, p, g, true, a);
    a.wg_item_ptr = wgptr;
    a.part_cnt = part_cnt;
    a.part_sum = part_sum;
    a.n_recheck = counters;
    e = launch_xi(g.type, precision, TOS_ACC_CORE, a, n_wg, L.block, L.lds, st);
    if (e != hipSuccess) return hip_fail(ctx, (hipError_t)e, "tos_xi_run_noalloc pair kernel");
    e = launch_reduce(part_cnt, part_sum, n_wg, nbins, 2, num_pairs, sums, seg_cnt, seg_sum, n_seg, st);
    if (e != hipSuccess) return hip_fail(ctx, (hipError_t)e, "tos_xi_run_noalloc reduction");
    hipLaunchKernelGGL(split_sums_kernel, dim3((nbins + 255) / 256), dim3(256), 0, st, sums, nbins, didj, wgt);
    he = hipGetLastError();
    return he == hipSuccess ? TOS_OK : hip_fail(ctx, he, "tos_xi_run_noalloc");
}
