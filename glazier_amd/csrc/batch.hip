// batch.hip - the batched step's kernels: argument preparation from the
// per-replica step counters, and the sweep, field and per-label kernels over
// (replica, local block) work lists.
//
// The kernel bodies are the single-lattice ones (sweep_body.h, field_body.h,
// cell_body.h) called with the local block index, so a replica in a batch
// computes exactly what it computes alone: the same draws, the same
// acceptance, the same per-label integers.
#include "batch.h"
#include "cell_body.h"
#include "field_body.h"
#include "sweep_body.h"

namespace ghip {

namespace {

// The work-list entry holding block b: first[g] <= b < first[g + 1].
__device__ __forceinline__ uint32_t find_entry(const WorkList &w, uint32_t b) {
    uint32_t lo = 0, hi = w.n;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (w.first[mid] <= b) lo = mid;
        else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ uint64_t passes_before(const BatchReplica &R, uint64_t m) {
    return R.pass0 + (m - R.mcs0) * (uint64_t)R.colours;
}

// Index of the current label (or activity) buffer after pass p.
__device__ __forceinline__ int flip_index(const BatchReplica &R, int base, uint64_t p) {
    return R.has_mot ? (int)((base + (p - R.pass0)) & 1) : base;
}

}  // namespace

// ---------------------------------------------------------------- sweep

// One thread per replica: the arguments of pass `phase` of its current step,
// as the single-lattice driver builds them (engine.cpp, run_sweep_pass). The
// record already holds the replica's static arguments (uploaded with the
// replica records); only the per-pass ones are written.
__global__ void batch_prepare_sweep_kernel(BatchArgs b, int phase) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= b.n_reps) return;
    const BatchReplica &R = b.reps[r];
    SweepArgs &a = b.cur_sweep[r];
    if (phase >= R.colours) {
        a.n_off = 0;  // no pass in this phase
        return;
    }
    a.n_off = R.sweep.n_off;
    const uint64_t m = b.mcs[r];
    int order[8];
    colour_order(R.seed, R.replica, m, R.colours, order);
    const int colour = order[phase];
    const uint64_t p = passes_before(R, m) + (uint64_t)phase;
    a.a_cur = R.A[p % 2];
    a.e_prev = R.D[(p + 2) % 3];
    a.f_acc = R.D[p % 3];
    a.a_next = R.A[(p + 1) % 2];
    a.g_zero = R.D[(p + 1) % 3];
    if (R.has_mot) {
        const int li = flip_index(R, R.lab0, p), ai = flip_index(R, R.act0, p);
        a.labels_prev = R.labels[li];
        a.labels = R.labels[1 - li];
        a.act_prev = R.act[ai];
        a.act = R.act[1 - ai];
    }
    for (int k = 0; k < R.n_species; ++k)
        a.fields[k] = R.fbuf[k][(R.field0[k] + (m - R.mcs0) * (uint64_t)R.chunks[k]) & 1];
    a.cx = colour & 1;
    a.cy = (colour >> 1) & 1;
    a.cz = (colour >> 2) & 1;
    a.key = pass_key(R.seed, R.replica, m, (uint32_t)colour);
    a.key2 = second_key(a.key);
}

template <bool D3, bool FULL, bool MOORE2D>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(D3 ? 5 : 4))) void batch_sweep_kernel(BatchArgs b, WorkList w) {
    const uint32_t g = find_entry(w, blockIdx.x);
    const SweepArgs &a = b.cur_sweep[w.members[g]];
    if (a.n_off == 0) return;
    sweep_body<D3, FULL, MOORE2D>(a, (int)(blockIdx.x - w.first[g]));
}

// The length term's per-label `before` of this phase's pass, for every
// replica whose model prices length (launched after the phase's prepare).
__global__ __launch_bounds__(256) void batch_length_kernel(BatchArgs b, WorkList w) {
    const uint32_t g = find_entry(w, blockIdx.x);
    const uint32_t r = w.members[g];
    const SweepArgs &s = b.cur_sweep[r];
    if (s.n_off == 0) return;
    length_prepass_body(s.a_cur, s.e_prev, s.n_labels, s.cell_words, b.reps[r].cell.ltype, s.tables, s.d > 1,
                        const_cast<double *>(s.len_before), (int)(blockIdx.x - w.first[g]),
                        (int)(w.first[g + 1] - w.first[g]));
}

// ---------------------------------------------------------------- fields

// One thread per (replica, species): the arguments of field launch `chunk`
// of this step (engine.cpp, run_fields). Species are independent within a
// step, so chunk c of every species runs in one launch.
__global__ void batch_prepare_field_kernel(BatchArgs b, int chunk) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t S = b.n_species_max;
    if (t >= b.n_reps * S) return;
    const uint32_t r = t / S, k = t % S;
    const BatchReplica &R = b.reps[r];
    FieldArgs f{};
    if ((int)k < R.n_species && chunk < R.chunks[k]) {
        const uint64_t m = b.mcs[r];
        const int cur = (int)((R.field0[k] + (m - R.mcs0) * (uint64_t)R.chunks[k] + (uint64_t)chunk) & 1);
        const uint64_t p_end = passes_before(R, m + 1);
        f.in = R.fbuf[k][cur];
        f.out = R.fbuf[k][1 - cur];
        f.labels = R.labels[flip_index(R, R.lab0, p_end)];
        f.tables = R.sweep.tables;
        f.w = R.sweep.w;
        f.h = R.sweep.h;
        f.d = R.sweep.d;
        f.ix_log2 = R.fx_log2;
        f.iy_log2 = R.fy_log2;
        f.iz = R.fz;
        f.steps = R.chunk_steps[k][chunk];
        f.exchange = chunk == R.chunks[k] - 1 ? 1 : 0;
        f.species = (int)k;
        f.dcoef = R.fd[k];
        f.kcoef = R.fk[k];
        f.tiles_x = R.f_tiles_x;
        f.tiles_y = R.f_tiles_y;
    }
    b.cur_field[t] = f;
}

template <bool D3>
__global__ __launch_bounds__(256) void batch_field_kernel(BatchArgs b, WorkList w) {
    const uint32_t g = find_entry(w, blockIdx.x);
    const FieldArgs &f = b.cur_field[w.members[g]];
    if (f.steps == 0) return;
    const int blk = (int)(blockIdx.x - w.first[g]);
    if (!D3 && f.ix_log2 == 6 && f.iy_log2 == 5 && field_planes_apply(f.w, f.h, false)) {
        // a plane of at least one 64 x 32 tile takes the single-lattice
        // launch's plane body (the dynamic LDS holds its two buffers)
        extern __shared__ float fl[];
        switch (f.steps) {
            case 1: field_body_2d<1>(f, blk, fl, fl + (kFieldIX + 2) * (kFieldIY + 2)); return;
            case 2: field_body_2d<2>(f, blk, fl, fl + (kFieldIX + 4) * (kFieldIY + 4)); return;
            case 3: field_body_2d<3>(f, blk, fl, fl + (kFieldIX + 6) * (kFieldIY + 6)); return;
            case 4: field_body_2d<4>(f, blk, fl, fl + (kFieldIX + 8) * (kFieldIY + 8)); return;
            default: break;
        }
    }
    field_body<D3>(f, blk);
}

// ---------------------------------------------------------------- cells

// One thread per replica: this step's per-label arguments (engine.cpp,
// run_cells) and the activity buffer the decay steps down. As for the
// sweep, the static arguments are already in the record.
__global__ void batch_prepare_cells_kernel(BatchArgs b) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= b.n_reps) return;
    const BatchReplica &R = b.reps[r];
    CellArgs &c = b.cur_cell[r];
    const uint64_t m = b.mcs[r];
    const uint64_t p_end = passes_before(R, m + 1);
    c.state = R.A[p_end % 2];
    c.state_copy = R.A[(p_end + 1) % 2];
    // a lattice that divides or dies alternates two sets of flag words
    if (R.has_pop) {
        c.flags = R.flags_base + 4 * (m & 1);
        c.flags_next = R.flags_base + 4 * ((m + 1) & 1);
    }
    c.labels = R.labels[flip_index(R, R.lab0, p_end)];
    c.key = pass_key(R.seed, R.replica, m, 0xDEADCE11U);
    c.key2 = second_key(c.key);
    b.cur_act[r] = R.act[flip_index(R, R.act0, p_end)];
}

__global__ void batch_advance_kernel(BatchArgs b) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < b.n_reps) b.mcs[r] += 1;
}

template <int OP, int THREADS>
__global__ __launch_bounds__(THREADS) void batch_cell_kernel(BatchArgs b, WorkList w) {
    const uint32_t g = find_entry(w, blockIdx.x);
    const uint32_t r = w.members[g];
    const int lb = (int)(blockIdx.x - w.first[g]);
    const int gb = (int)(w.first[g + 1] - w.first[g]);
    const BatchReplica &R = b.reps[r];
    const CellArgs &c = b.cur_cell[r];
    const size_t n = R.n_sites;
    if constexpr (OP == kOpConsolidate) {
        const uint64_t p = passes_before(R, b.mcs[r] + 1);
        consolidate_body(R.A[p % 2], R.D[(p + 2) % 3], c.n_labels, c.cw, lb, gb);
        // the step's flags start clear, except word 2 (ids ran out), which
        // the host reads and resets after the call
        if (lb == 0 && threadIdx.x < 4 && threadIdx.x != 2) c.flags[threadIdx.x] = 0u;
    } else if constexpr (OP == kOpPlan) {
        // the moved-site counts the next stage accumulates start at zero
        const uint32_t i = lb * blockDim.x + threadIdx.x;
        if (i < c.capacity) c.scratch_u32[i] = 0u;
        const uint64_t p = passes_before(R, b.mcs[r] + 1);
        plan_body(c, R.D[(p + 2) % 3], lb, gb);
    } else if constexpr (OP == kOpCountMoved) {
        count_moved_body(c, n, lb, gb);
    } else if constexpr (OP == kOpAssign) {
        assign_daughters_body(c, lb, gb);
    } else if constexpr (OP == kOpApply) {
        apply_population_body(c, n, lb, gb);
    } else if constexpr (OP == kOpCompactStage) {
        compact_stage_body(c, R.type_copy, R.anchor_copy, lb, gb);
    } else if constexpr (OP == kOpCompactMove) {
        compact_move_body(c, R.type_copy, R.anchor_copy, lb, gb);
    } else if constexpr (OP == kOpCompactRelabel) {
        compact_relabel_body(c, n, lb, gb);
    } else if constexpr (OP == kOpCheckRecentre) {
        check_recentre_body(c, lb, gb);
    } else if constexpr (OP == kOpRecentre1) {
        recentre_body(c, 1, lb, gb);
    } else if constexpr (OP == kOpRecountMoments1) {
        recount_body(c, n, 1, 1, lb, gb);
    } else if constexpr (OP == kOpDecay) {
        // activity_decay_kernel (fields.hip) on this replica's buffer
        float *act = b.cur_act[r];
        const size_t i = (size_t)lb * blockDim.x + threadIdx.x;
        if (i < n) {
            const float v = act[i];
            act[i] = v > 0.0f ? v - 1.0f : 0.0f;
        }
    }
}

// ---------------------------------------------------------------- launchers

namespace {
inline dim3 per_replica_grid(uint32_t n) { return dim3((n + 63) / 64); }
}  // namespace

hipError_t launch_batch_prepare_sweep(const BatchArgs &b, int phase, hipStream_t s) {
    hipLaunchKernelGGL(batch_prepare_sweep_kernel, per_replica_grid(b.n_reps), dim3(64), 0, s, b, phase);
    return hipGetLastError();
}

hipError_t launch_batch_sweep(const BatchArgs &b, bool d3, bool full, bool moore, const WorkList &w,
                              int threads, size_t lds, hipStream_t s) {
    if (w.blocks == 0) return hipSuccess;
#define GHIP_BS(D, F, M) \
    hipLaunchKernelGGL((batch_sweep_kernel<D, F, M>), dim3(w.blocks), dim3(threads), lds, s, b, w)
    if (d3) {
        if (full) GHIP_BS(true, true, false);
        else GHIP_BS(true, false, false);
    } else if (moore) {
        if (full) GHIP_BS(false, true, true);
        else GHIP_BS(false, false, true);
    } else {
        if (full) GHIP_BS(false, true, false);
        else GHIP_BS(false, false, false);
    }
#undef GHIP_BS
    return hipGetLastError();
}

hipError_t launch_batch_length(const BatchArgs &b, const WorkList &w, hipStream_t s) {
    if (w.blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(batch_length_kernel, dim3(w.blocks), dim3(256), 0, s, b, w);
    return hipGetLastError();
}

hipError_t launch_batch_prepare_field(const BatchArgs &b, int chunk, hipStream_t s) {
    const uint32_t n = b.n_reps * b.n_species_max;
    hipLaunchKernelGGL(batch_prepare_field_kernel, per_replica_grid(n), dim3(64), 0, s, b, chunk);
    return hipGetLastError();
}

hipError_t launch_batch_field(const BatchArgs &b, bool d3, const WorkList &w, size_t lds, hipStream_t s) {
    if (w.blocks == 0) return hipSuccess;
    if (d3)
        hipLaunchKernelGGL(batch_field_kernel<true>, dim3(w.blocks), dim3(256), lds, s, b, w);
    else
        hipLaunchKernelGGL(batch_field_kernel<false>, dim3(w.blocks), dim3(256), lds, s, b, w);
    return hipGetLastError();
}

hipError_t launch_batch_prepare_cells(const BatchArgs &b, hipStream_t s) {
    hipLaunchKernelGGL(batch_prepare_cells_kernel, per_replica_grid(b.n_reps), dim3(64), 0, s, b);
    return hipGetLastError();
}

hipError_t launch_batch_advance(const BatchArgs &b, hipStream_t s) {
    hipLaunchKernelGGL(batch_advance_kernel, per_replica_grid(b.n_reps), dim3(64), 0, s, b);
    return hipGetLastError();
}

int batch_cell_threads(int op) { return op == kOpAssign ? 1024 : 256; }

hipError_t launch_batch_cell(int op, const BatchArgs &b, const WorkList &w, hipStream_t s) {
    if (w.blocks == 0) return hipSuccess;
#define GHIP_BC(OPV, T)                                                                       \
    case OPV:                                                                                 \
        hipLaunchKernelGGL((batch_cell_kernel<OPV, T>), dim3(w.blocks), dim3(T), 0, s, b, w); \
        break
    switch (op) {
        GHIP_BC(kOpConsolidate, 256);
        GHIP_BC(kOpPlan, 256);
        GHIP_BC(kOpCountMoved, 256);
        GHIP_BC(kOpAssign, 1024);
        GHIP_BC(kOpApply, 256);
        GHIP_BC(kOpCompactStage, 256);
        GHIP_BC(kOpCompactMove, 256);
        GHIP_BC(kOpCompactRelabel, 256);
        GHIP_BC(kOpCheckRecentre, 256);
        GHIP_BC(kOpRecentre1, 256);
        GHIP_BC(kOpRecountMoments1, 256);
        GHIP_BC(kOpDecay, 256);
        default:
            return hipErrorInvalidValue;
    }
#undef GHIP_BC
    return hipGetLastError();
}

}  // namespace ghip
