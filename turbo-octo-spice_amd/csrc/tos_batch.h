// tos_batch.h -- batched launches of the pair kernel: the prepared state
// of one batch (job table, workgroup tasks, partials, pinned results) and
// the steps shared by tos_batch_run and the job server.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "tos_capi_impl.h"

namespace tos {

// Kernel variant shared by all jobs of one launch.
struct VariantKey {
    int coord, prec;
    bool fixed, simple;
    bool operator<(const VariantKey &o) const
    {
        if (coord != o.coord) return coord < o.coord;
        if (prec != o.prec) return prec < o.prec;
        if (fixed != o.fixed) return fixed < o.fixed;
        return simple < o.simple;
    }
};

struct BatchGroup {
    VariantKey key;
    std::vector<int> jobs;
    int block = 512;
    int n_wg = 0;
    size_t lds = 0;
    int task_begin = 0;     // first XiWgTask of the group
};

struct BatchState {
    // identity of the prepared batch: job inputs and the device buffers the
    // prepared arguments point to
    std::vector<uint64_t> key;
    std::vector<BatchGroup> groups;
    std::vector<int> nbins;            // per job
    std::vector<size_t> out_off;       // per job, in uint64 words of `res`
    size_t res_words = 0;
    int64_t max_fields = 0;
    DevBuf args, tasks, red, parts_cnt, parts_sum, res;
    void *host_res = nullptr;          // pinned copy of `res`
    size_t host_res_bytes = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // When set, buffers that batch_prepare replaces by larger ones are
    // handed over here (device / pinned host) instead of freed, and grow by
    // half again their need: hipFree and hipHostFree wait for the whole
    // device, which a caller with launches in flight on other streams must
    // not do.
    std::vector<void *> *defer_dev = nullptr, *defer_host = nullptr;

    void drop_graph()
    {
        if (exec) hipGraphExecDestroy(exec);
        if (graph) hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
};

// Items [begin, end) of a job's work-item list (a slice of the job).
struct ItemSlice {
    int32_t begin, end;
};

// Plan the batch (variant groups, workgroups per job, item ranges), size
// its buffers and upload the tables; synchronises ctx->stream.  slices:
// when given, job j covers only its items slices[j] (its result is then the
// histogram of those items).
int batch_prepare(tos_ctx *ctx, int32_t n, const tos_batch_job *jobs, BatchState &B,
                  const ItemSlice *slices = nullptr);
// Enqueue the device work of a prepared batch on `stream`.
int batch_enqueue(tos_ctx *ctx, BatchState &B, int32_t n, hipStream_t stream);
// Copy the results of a completed batch out of its pinned buffer.
void batch_collect(const BatchState &B, int32_t n, float ms, tos_xi_result *out);

}  // namespace tos
