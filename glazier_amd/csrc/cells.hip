// cells.hip - the per-label kernels (cell_body.h) over one lattice, and their
// launchers.
#include "cell_body.h"

namespace ghip {

__global__ void fill_i64_kernel(int64_t *p, size_t n) {
    fill_i64_body(p, n, blockIdx.x, gridDim.x);
}

__global__ void fill_u32_kernel(uint32_t *p, size_t n, uint32_t v) {
    fill_u32_body(p, n, v, blockIdx.x, gridDim.x);
}

__global__ void consolidate_kernel(int64_t *a, int64_t *e, const uint32_t *n_labels, int cw,
                                   uint32_t *clear4) {
    consolidate_body(a, e, n_labels, cw, blockIdx.x, gridDim.x);
    // the step's flags start clear, except word 2 (ids ran out), which the
    // host reads and resets once the call is done
    if (clear4 && blockIdx.x == 0 && threadIdx.x < 4 && threadIdx.x != 2) clear4[threadIdx.x] = 0u;
}

__global__ void first_site_kernel(CellArgs a, size_t n_sites) {
    first_site_body(a, n_sites, blockIdx.x, gridDim.x);
}

__global__ void anchor_from_first_kernel(CellArgs a) {
    anchor_from_first_body(a, blockIdx.x, gridDim.x);
}

__global__ __launch_bounds__(kRecountCols) void recount_kernel(CellArgs a, size_t n_sites, int moments_only,
                                                               int gate) {
    recount_body(a, n_sites, moments_only, gate, blockIdx.x, gridDim.x);
}

__global__ void zero_moments_kernel(CellArgs a) {
    zero_moments_body(a, blockIdx.x, gridDim.x);
}

__global__ void recentre_kernel(CellArgs a, int gate) {
    recentre_body(a, gate, blockIdx.x, gridDim.x);
}

__global__ void check_recentre_kernel(CellArgs a) {
    check_recentre_body(a, blockIdx.x, gridDim.x);
}

__global__ void plan_kernel(CellArgs a, int64_t *e) {
    // the moved-site counts the next stage accumulates start at zero
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < a.capacity) a.scratch_u32[i] = 0u;
    plan_body(a, e, blockIdx.x, gridDim.x);
}

__global__ void count_moved_kernel(CellArgs a, size_t n_sites) {
    count_moved_body(a, n_sites, blockIdx.x, gridDim.x);
}

__global__ __launch_bounds__(1024) void assign_daughters_kernel(CellArgs a) {
    assign_daughters_body(a, blockIdx.x, gridDim.x);
}

__global__ void apply_population_kernel(CellArgs a, size_t n_sites) {
    apply_population_body(a, n_sites, blockIdx.x, gridDim.x);
}

__global__ void zero_state_kernel(CellArgs a, int gate) {
    zero_state_body(a, gate, blockIdx.x, gridDim.x);
}


__global__ void compact_move_kernel(CellArgs a, const uint32_t *type_copy, const int32_t *anchor_copy) {
    compact_move_body(a, type_copy, anchor_copy, blockIdx.x, gridDim.x);
}

__global__ void compact_relabel_kernel(CellArgs a, size_t n_sites) {
    compact_relabel_body(a, n_sites, blockIdx.x, gridDim.x);
}


__global__ void copy_u32_kernel(uint32_t *dst, const uint32_t *src, const uint32_t *n_labels, int per) {
    copy_u32_body(dst, src, n_labels, per, blockIdx.x, gridDim.x);
}

// The compaction's staging copies of the label types (one word a label),
// anchors (three) and records (cw int64) in one launch, only on a step whose
// plan compacts.
__global__ void compact_stage_kernel(CellArgs a, uint32_t *type_copy, int32_t *anchor_copy) {
    compact_stage_body(a, type_copy, anchor_copy, blockIdx.x, gridDim.x);
}

__global__ void length_prepass_kernel(const int64_t *a, const int64_t *e, const uint32_t *n_labels, int cw,
                                      const uint32_t *ltype, const DeviceTables *tables, int d3, double *out) {
    length_prepass_body(a, e, n_labels, cw, ltype, tables, d3 != 0, out, blockIdx.x, gridDim.x);
}

// ---------------------------------------------------------------- launchers

static int blocks_for(size_t n, int threads) { return (int)((n + threads - 1) / threads); }

hipError_t launch_length_prepass(const int64_t *a, const int64_t *e, const uint32_t *n_labels, int cw,
                                 const uint32_t *ltype, const DeviceTables *tables, bool d3, double *out,
                                 uint32_t capacity, hipStream_t s) {
    hipLaunchKernelGGL(length_prepass_kernel, dim3(blocks_for(capacity, 256)), dim3(256), 0, s, a, e, n_labels,
                       cw, ltype, tables, d3 ? 1 : 0, out);
    return hipGetLastError();
}

hipError_t launch_fill_i64(int64_t *p, size_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(fill_i64_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t *p, size_t n, uint32_t v, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(fill_u32_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

hipError_t launch_consolidate(int64_t *a, int64_t *e, const uint32_t *n_labels, int cw, size_t cap,
                              hipStream_t s, uint32_t *clear4) {
    int blocks = blocks_for(cap * cw, 256);
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(consolidate_kernel, dim3(blocks), dim3(256), 0, s, a, e, n_labels, cw, clear4);
    return hipGetLastError();
}

static void recount_gated(const CellArgs &a, size_t n_sites, int moments_only, int gate, hipStream_t s) {
    hipLaunchKernelGGL(recount_kernel, dim3(recount_blocks(a.w, a.h, a.d)), dim3(kRecountCols), 0, s, a, n_sites,
                       moments_only, gate);
}

static void zero_state(const CellArgs &a, int gate, hipStream_t s) {
    int blocks = blocks_for((size_t)a.capacity * a.cw, 256);
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(zero_state_kernel, dim3(blocks), dim3(256), 0, s, a, gate);
}

// Full recount of volume, surface and (when kept) moments.
hipError_t launch_recount(const CellArgs &a, size_t n_sites, hipStream_t s) {
    zero_state(a, -1, s);
    recount_gated(a, n_sites, 0, -1, s);
    return hipGetLastError();
}

hipError_t launch_recount_moments(const CellArgs &a, size_t n_sites, hipStream_t s) {
    hipLaunchKernelGGL(zero_moments_kernel, dim3(blocks_for(a.capacity, 256)), dim3(256), 0, s, a);
    recount_gated(a, n_sites, 1, -1, s);
    return hipGetLastError();
}

// Anchors at each label's first site, then one recentring onto the rounded
// centroid. The state's volume words must be current.
hipError_t launch_initial_anchors(const CellArgs &a, size_t n_sites, hipStream_t s) {
    hipError_t e = launch_fill_u32(a.scratch_u32, a.capacity, 0xFFFFFFFFu, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(first_site_kernel, dim3(blocks_for(n_sites, 256)), dim3(256), 0, s, a, n_sites);
    hipLaunchKernelGGL(anchor_from_first_kernel, dim3(blocks_for(a.capacity, 256)), dim3(256), 0, s, a);
    hipLaunchKernelGGL(zero_moments_kernel, dim3(blocks_for(a.capacity, 256)), dim3(256), 0, s, a);
    recount_gated(a, n_sites, 1, -1, s);
    hipLaunchKernelGGL(recentre_kernel, dim3(blocks_for(a.capacity, 256)), dim3(256), 0, s, a, -1);
    recount_gated(a, n_sites, 1, -1, s);
    return hipGetLastError();
}

// End-of-step check: recentre and resum when some label has drifted from its
// anchor. Enqueued unconditionally; the work is gated on the device flag.
hipError_t launch_recentre_check(const CellArgs &a, size_t n_sites, hipStream_t s) {
    hipLaunchKernelGGL(check_recentre_kernel, dim3(blocks_for(a.capacity, 256)), dim3(256), 0, s, a);
    hipLaunchKernelGGL(recentre_kernel, dim3(blocks_for(a.capacity, 256)), dim3(256), 0, s, a, 1);
    recount_gated(a, n_sites, 1, 1, s);
    return hipGetLastError();
}

// Division and death of one step: the plan (which first folds the last
// pass's deltas into the records), the daughter assignment (which also plans
// any compaction), the relabelling pass (which recounts the touched
// labels), the compaction with the records moved along, then the recentring
// the plan's drift check asked for; every stage after the plan gated on the
// device flags so nothing returns to the host.
hipError_t launch_population(const CellArgs &a, int64_t *last_deltas, size_t n_sites, uint32_t *type_copy,
                             int32_t *anchor_copy, hipStream_t s) {
    const int lb = blocks_for(a.capacity, 256);
    const int sb = blocks_for(n_sites, 256);
    hipLaunchKernelGGL(plan_kernel, dim3(lb), dim3(256), 0, s, a, last_deltas);
    hipLaunchKernelGGL(count_moved_kernel, dim3(sb), dim3(256), 0, s, a, n_sites);
    hipLaunchKernelGGL(assign_daughters_kernel, dim3(1), dim3(1024), 0, s, a);
    hipLaunchKernelGGL(apply_population_kernel, dim3(sb), dim3(256), 0, s, a, n_sites);
    {
        int cb = blocks_for((size_t)a.capacity * a.cw, 256);
        if (cb > 4096) cb = 4096;
        hipLaunchKernelGGL(compact_stage_kernel, dim3(cb), dim3(256), 0, s, a, type_copy, anchor_copy);
    }
    hipLaunchKernelGGL(compact_move_kernel, dim3(lb), dim3(256), 0, s, a, type_copy, anchor_copy);
    hipLaunchKernelGGL(compact_relabel_kernel, dim3(sb), dim3(256), 0, s, a, n_sites);
    hipLaunchKernelGGL(recentre_kernel, dim3(lb), dim3(256), 0, s, a, 1);
    recount_gated(a, n_sites, 1, 1, s);
    return hipGetLastError();
}

}  // namespace ghip
