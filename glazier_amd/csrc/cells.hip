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

// ---- the recount of a plane whose records are volume and surface alone
//
// A workgroup of 256 lanes takes kPlaneRows rows of 256 columns. The rows
// with a one-site halo land in LDS by 16-byte loads (a row's interior at word
// 4 of its LDS row, the left halo word at 3, the right one past the
// interior). Each lane walks its column down with the 3 x 3 neighbourhood in
// registers (three LDS reads a row), summing volume and bonds of a run of one
// label; where runs end, the wavefront first sums the runs of adjacent lanes
// that share a label (a cell spans several columns), and only each group's
// first lane adds into the workgroup's LDS table. The table keeps the 16
// records of a 128-byte line in adjacent slots, so its global atomics land on
// few lines per instruction. Integer sums: the result does not depend on
// the order of any of it.
constexpr int kPlaneRows = 16;
constexpr int kPlaneCols = 256;
constexpr int kPlanePitch = kPlaneCols + 8;
constexpr int kPlaneSlots = 256;

bool recount_plane_applies(const CellArgs &a) {
    return a.d == 1 && a.cw == 1 && a.w % kPlaneCols == 0 && a.h % kPlaneRows == 0 && (a.n_off == 4 || a.n_off == 8);
}

// The sum of `v` over the lanes of this lane's group: maximal runs of
// adjacent active lanes holding one key. Returns true on the group's first
// lane, which then holds the group's sum in `sum`.
__device__ __forceinline__ bool wave_group_sum(bool active, uint32_t key, uint32_t v, uint32_t *sum) {
    const int lane = (int)(threadIdx.x & 63);
    const uint32_t pkey = __shfl_up(key, 1, 64);
    const bool pact = __shfl_up(active ? 1 : 0, 1, 64) != 0;
    const bool head = active && (lane == 0 || !pact || pkey != key);
    const unsigned long long heads = __ballot(head);
    const unsigned long long breaks = heads | ~__ballot(active);
    // inclusive prefix sum over the wavefront of the active values
    const uint32_t own = active ? v : 0u;
    uint32_t p = own;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t q = __shfl_up(p, (unsigned)d, 64);
        if (lane >= d) p += q;
    }
    // a group ends before the next break after its head (or at lane 63)
    const unsigned long long later = lane == 63 ? 0ull : (breaks >> (lane + 1));
    const int end = later ? lane + __builtin_ctzll(later) : 63;
    *sum = __shfl(p, end, 64) - (p - own);
    return head;
}

template <int NOFF>
__global__ __launch_bounds__(kPlaneCols) void recount_plane_kernel(CellArgs a) {
    __shared__ __attribute__((aligned(16))) uint32_t rt[(kPlaneRows + 2) * kPlanePitch];
    __shared__ uint32_t keys[kPlaneSlots];
    __shared__ unsigned long long vals[kPlaneSlots];
    const int tid = (int)threadIdx.x;
    const int col_blocks = a.w / kPlaneCols;
    const int x0 = (int)(blockIdx.x % (uint32_t)col_blocks) * kPlaneCols;
    const int y0 = (int)(blockIdx.x / (uint32_t)col_blocks) * kPlaneRows;
    const int W = a.w, H = a.h;
    for (int i = tid; i < kPlaneSlots; i += kPlaneCols) {
        keys[i] = 0xFFFFFFFFu;
        vals[i] = 0ull;
    }
    // rows y0 - 1 .. y0 + kPlaneRows (wrapped): 64 16-byte words a row
    constexpr int kRowsHeld = kPlaneRows + 2;
    constexpr int kQuads = kRowsHeld * (kPlaneCols / 4);
    constexpr int kLoads = (kQuads + kPlaneCols - 1) / kPlaneCols;
    uint4 v[kLoads];
#pragma unroll
    for (int k = 0; k < kLoads; ++k) {
        const int q = k * kPlaneCols + tid;
        const int r = q < kQuads ? q / (kPlaneCols / 4) : kRowsHeld - 1;
        const int c4 = q % (kPlaneCols / 4);
        int gy = y0 - 1 + r;
        gy = gy < 0 ? gy + H : gy >= H ? gy - H : gy;
        v[k] = *reinterpret_cast<const uint4 *>(a.labels + (size_t)gy * (size_t)W + (size_t)(x0 + 4 * c4));
    }
    // the halo columns: lane 2 r + side
    uint32_t hv = 0;
    const bool h_live = tid < 2 * kRowsHeld;
    if (h_live) {
        const int r = tid >> 1;
        int gy = y0 - 1 + r;
        gy = gy < 0 ? gy + H : gy >= H ? gy - H : gy;
        const int gx = (tid & 1) ? (x0 + kPlaneCols == W ? 0 : x0 + kPlaneCols) : (x0 == 0 ? W - 1 : x0 - 1);
        hv = a.labels[(size_t)gy * (size_t)W + (size_t)gx];
    }
#pragma unroll
    for (int k = 0; k < kLoads; ++k) {
        const int q = k * kPlaneCols + tid;
        if (q < kQuads)
            *reinterpret_cast<uint4 *>(&rt[(q / (kPlaneCols / 4)) * kPlanePitch + 4 + 4 * (q % (kPlaneCols / 4))]) =
                v[k];
    }
    if (h_live) rt[(tid >> 1) * kPlanePitch + ((tid & 1) ? 4 + kPlaneCols : 3)] = hv;
    __syncthreads();

    auto insert = [&](uint32_t l, int64_t d) {
        uint32_t h = ((((l >> 4) * 0x9E3779B1u) >> 24) & ~15u) | (l & 15u);
        for (int probe = 0; probe < kPlaneSlots / 16; ++probe) {
            h &= (uint32_t)(kPlaneSlots - 1);
            const uint32_t k = atomicCAS(&keys[h], 0xFFFFFFFFu, l);
            if (k == 0xFFFFFFFFu || k == l) {
                atomicAdd(&vals[h], (unsigned long long)d);
                return;
            }
            h += 16;
        }
        atomicAdd(reinterpret_cast<unsigned long long *>(a.state + l), (unsigned long long)d);
    };
    // a run's volume and bonds ride in one word, volume * 2^16 + bonds: a
    // wavefront's group sums stay below 64 * kPlaneRows sites and 8 times
    // as many bonds, inside 16 bits each
    static_assert(64 * kPlaneRows * 8 < 65536, "a group's bond sum fits 16 bits");
    auto flush = [&](bool active, uint32_t word, uint32_t acc) {
        if (__ballot(active) == 0ull) return;
        uint32_t sum;
        if (wave_group_sum(active, word, acc, &sum)) insert(label_id(word), pack_vs(sum >> 16, sum & 0xFFFFu));
    };

    // the 3 x 3 neighbourhood of this lane's site, rows j - 1 .. j + 1; a
    // run ends where the word changes (a label's word carries its type)
    const int c = 4 + tid;
    uint32_t n00 = rt[c - 1], n01 = rt[c], n02 = rt[c + 1];
    uint32_t n10 = rt[kPlanePitch + c - 1], n11 = rt[kPlanePitch + c], n12 = rt[kPlanePitch + c + 1];
    uint32_t run = n11;
    uint32_t acc = 0;
#pragma unroll 4
    for (int j = 1; j <= kPlaneRows; ++j) {
        const uint32_t *below = &rt[(j + 1) * kPlanePitch + c];
        const uint32_t n20 = below[-1], n21 = below[0], n22 = below[1];
        const uint32_t word = n11;
        uint32_t like = (n01 == word) + (n10 == word) + (n12 == word) + (n21 == word);
        if constexpr (NOFF == 8) like += (n00 == word) + (n02 == word) + (n20 == word) + (n22 == word);
        const bool ends = word != run;
        flush(ends, run, acc);
        if (ends) {
            run = word;
            acc = 0;
        }
        acc += 65536u + (uint32_t)NOFF - like;
        n00 = n10; n01 = n11; n02 = n12;
        n10 = n20; n11 = n21; n12 = n22;
    }
    flush(true, run, acc);
    __syncthreads();
    for (int i = tid; i < kPlaneSlots; i += kPlaneCols) {
        const uint32_t l = keys[i];
        if (l != 0xFFFFFFFFu) atomicAdd(reinterpret_cast<unsigned long long *>(a.state + l), vals[i]);
    }
}

static void recount_gated(const CellArgs &a, size_t n_sites, int moments_only, int gate, hipStream_t s) {
    if (!moments_only && gate < 0 && recount_plane_applies(a)) {
        const uint32_t blocks = (uint32_t)(a.w / kPlaneCols) * (uint32_t)(a.h / kPlaneRows);
        if (a.n_off == 8)
            hipLaunchKernelGGL(recount_plane_kernel<8>, dim3(blocks), dim3(kPlaneCols), 0, s, a);
        else
            hipLaunchKernelGGL(recount_plane_kernel<4>, dim3(blocks), dim3(kPlaneCols), 0, s, a);
        return;
    }
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
