// batch_host.cpp - ghip_batch_*: many lattices stepped together, one Monte
// Carlo step of all of them captured in a hipGraph (kernels in batch.hip).
//
// The reference steps one lattice per process (glazier/src/main.rs:137-160);
// its answers are distributional, so ensembles of seeds and parameter sweeps
// are the normal workload. A batch takes lattices created by ghip_create
// (each keeping its buffers, seed and replica index) and launches each stage
// of a step once for all of them, grouped by kernel variant, so that small
// lattices share the chip instead of each leaving it mostly idle.
//
// Tile shapes: the first steps of a batch run the sweep groups with each
// candidate shape in turn (every step is a real step: the draws and
// acceptance do not depend on the tile shape), time them with HIP events,
// keep the fastest per group, and then capture the step.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/glazier_hip.h"
#include "batch.h"
#include "sim.h"

using namespace ghip;

namespace {

struct Tile {
    int tx, ty, tz, threads;
};

// Candidate shapes in targets (a tile spans twice as many sites per axis).
constexpr int kCandidates = 5;
const Tile kTiles2D[kCandidates] = {
    {64, 8, 1, 512}, {32, 8, 1, 256}, {64, 4, 1, 256}, {32, 16, 1, 512}, {16, 16, 1, 256}};
const Tile kTiles3D[kCandidates] = {
    {16, 4, 4, 256}, {8, 8, 4, 256}, {16, 4, 8, 512}, {8, 4, 4, 128}, {16, 8, 4, 512}};

// Interior tiles of the field launches in sites, timed on the same steps as
// the colour passes' candidates (a field step's bits do not depend on its
// tile). The first is the single lattice's own choice.
const Tile kFieldTiles2D[kCandidates] = {
    {64, 32, 1, 256}, {128, 16, 1, 256}, {32, 32, 1, 256}, {64, 64, 1, 256}, {128, 32, 1, 256}};
const Tile kFieldTiles3D[kCandidates] = {
    {16, 8, 8, 256}, {32, 8, 4, 256}, {16, 16, 4, 256}, {8, 8, 8, 256}, {32, 4, 8, 256}};

// Occupancy candidates, timed on the chosen tiles: at most this many
// workgroups of one launch per CU (0: as many as registers and LDS allow).
// A cap is reached by asking for more LDS than the kernel uses, so that one
// more workgroup would not fit in the CU's 160 KiB; every cap stays within
// the 64 KiB a workgroup may take.
constexpr int kOccupancy = 4;
const int kWgsPerCu[kOccupancy] = {0, 4, 3, 2};
constexpr size_t kLdsPerCu = 160 * 1024;

size_t occupancy_lds(size_t lds, int occ) {
    const int wgs = kWgsPerCu[occ];
    if (wgs == 0) return lds;
    const size_t need = (kLdsPerCu / (size_t)(wgs + 1) + 16) & ~(size_t)15;
    return std::max(lds, need);
}

// The candidate clamped to the lattice; spare lanes go to the next axis.
void clamp_tile(const ghip_sim *s, const Tile &t, int *tx, int *ty, int *tz) {
    *tx = std::min(t.tx, pow2_floor(s->W / 2));
    if (!s->d3) {
        *ty = std::min(pow2_floor(t.threads / *tx), pow2_floor(s->H / 2));
        *tz = 1;
    } else {
        *ty = std::min(t.ty, pow2_floor(s->H / 2));
        *tz = std::min(pow2_floor(t.threads / (*tx * *ty)), pow2_floor(s->D / 2));
    }
}

uint32_t cdiv(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// Grid of a per-label launch for one lattice, as cells.hip sizes it.
uint32_t cell_blocks(const ghip_sim *s, int op) {
    const uint64_t words = (uint64_t)s->cap * s->cw;
    switch (op) {
        case kOpConsolidate:
        case kOpCompactStage:
            return std::min<uint32_t>(4096, cdiv(words, 256));
        case kOpAssign:
            return 1;
        case kOpCountMoved:
        case kOpApply:
        case kOpCompactRelabel:
        case kOpDecay:
            return cdiv(s->N, 256);
        case kOpRecountMoments1:
            return recount_blocks(s->W, s->H, s->D);
        default:
            return cdiv(s->cap, 256);
    }
}

bool op_is_population(int op) { return op >= kOpPlan && op <= kOpCompactRelabel; }

// Host form of a work list: offsets into the batch's index pool.
struct ListRef {
    size_t first = 0, members = 0;
    uint32_t n = 0, blocks = 0;
};

struct SweepGroup {
    bool d3 = false, full = false, moore = false;
    int colours = 4;
    std::vector<uint32_t> reps;  // largest first
    int cand = 0;
    float ms[kCandidates] = {};
    int occ = 0;
    float occ_ms[kOccupancy] = {};
    int threads = 0;
    size_t lds = 0;
    ListRef ref;
    WorkList w{};
};

struct FieldLaunch {
    bool d3 = false;
    int chunk = 0;
    size_t lds = 0;  // what the kernel uses; the launch adds the occupancy cap
    ListRef ref;
    WorkList w{};
};

struct CellLaunch {
    int op = 0;
    ListRef ref;
    WorkList w{};
};

}  // namespace

struct ghip_batch {
    std::vector<ghip_sim *> sims;
    hipStream_t stream = nullptr;
    std::vector<BatchReplica> reps;
    BatchArgs args{};
    BatchReplica *d_reps = nullptr;
    uint64_t *d_mcs = nullptr;
    SweepArgs *d_sweep = nullptr;
    CellArgs *d_cell = nullptr;
    FieldArgs *d_field = nullptr;
    float **d_act = nullptr;
    uint32_t *d_pool = nullptr;
    size_t pool_cap = 0;
    std::vector<uint32_t> pool;
    std::vector<SweepGroup> groups;
    std::vector<FieldLaunch> fields;
    std::vector<CellLaunch> cells;
    ListRef length_ref;
    WorkList length{};          // replicas that price length: the per-pass prepass
    int max_colours = 0, max_chunks = 0;
    bool any_cells = false, any_mot = false;
    std::vector<uint64_t> expect_mcs;  // each lattice's step count as the batch left it
    // 0: warm-up; 1..kCandidates: tile candidate tune_step - 1 (colour passes
    // and fields); then occupancy candidate tune_step - kCandidates
    int tune_step = 0;
    bool tuned = false;
    // the field launches' tile and occupancy candidates, per dimension
    int field_cand[2] = {0, 0}, field_occ[2] = {0, 0};
    float field_ms[2][kCandidates] = {}, field_occ_ms[2][kOccupancy] = {};
    bool autotune = true;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<hipEvent_t> events;
    uint64_t graph_steps = 0, eager_steps = 0;

    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }

    ~ghip_batch() {
        if (stream) (void)hipStreamSynchronize(stream);
        drop_graph();
        for (hipEvent_t e : events) (void)hipEventDestroy(e);
        void *bufs[] = {d_reps, d_mcs, d_sweep, d_cell, d_field, d_act, d_pool};
        for (void *p : bufs)
            if (p) (void)hipFree(p);
        if (stream) (void)hipStreamDestroy(stream);
        for (ghip_sim *s : sims) s->batch = nullptr;
    }
};

namespace {

ListRef add_list(ghip_batch *b, const std::vector<uint32_t> &members, const std::vector<uint32_t> &blocks) {
    ListRef r;
    r.n = (uint32_t)members.size();
    r.first = b->pool.size();
    uint32_t acc = 0;
    b->pool.push_back(0);
    for (uint32_t x : blocks) {
        acc += x;
        b->pool.push_back(acc);
    }
    r.blocks = acc;
    r.members = b->pool.size();
    b->pool.insert(b->pool.end(), members.begin(), members.end());
    return r;
}

WorkList materialise(const ghip_batch *b, const ListRef &r) {
    WorkList w{};
    w.first = b->d_pool + r.first;
    w.members = b->d_pool + r.members;
    w.n = r.n;
    w.blocks = r.blocks;
    return w;
}

// The sweep template of lattice s under tile candidate `cand`.
SweepArgs sweep_for(const ghip_sim *s, int cand, uint32_t *blocks) {
    SweepArgs a = s->sweep_template();
    int tx, ty, tz;
    clamp_tile(s, s->d3 ? kTiles3D[cand] : kTiles2D[cand], &tx, &ty, &tz);
    a.tx_log2 = ilog2(tx);
    a.ty_log2 = ilog2(ty);
    a.tz_log2 = ilog2(tz);
    a.tiles_x = cdiv(s->W / 2, tx);
    a.tiles_y = cdiv(s->H / 2, ty);
    a.tiles_z = s->d3 ? cdiv(s->D / 2, tz) : 1;
    *blocks = (uint32_t)a.tiles_x * a.tiles_y * a.tiles_z;
    const uint64_t words = (uint64_t)s->cap * s->cw;
    a.fold_per_block = cdiv(words, *blocks);
    a.xcd_chunk = 0;
    a.reverse = 0;
    return a;
}

// The field tile of lattice s under field candidate `cand`, clamped to the
// lattice, into its replica record; returns the workgroups per species.
uint32_t field_tile(const ghip_sim *s, int cand, BatchReplica *B) {
    const Tile &t = (s->d3 ? kFieldTiles3D : kFieldTiles2D)[cand];
    B->fx_log2 = ilog2(std::min(t.tx, pow2_floor(s->W)));
    B->fy_log2 = ilog2(std::min(t.ty, pow2_floor(s->H)));
    B->fz = s->d3 ? std::min(t.tz, pow2_floor(s->D)) : 1;
    B->f_tiles_x = (s->W + (1 << B->fx_log2) - 1) >> B->fx_log2;
    B->f_tiles_y = (s->H + (1 << B->fy_log2) - 1) >> B->fy_log2;
    const uint32_t tiles_z = s->d3 ? cdiv(s->D, B->fz) : 1;
    return (uint32_t)B->f_tiles_x * B->f_tiles_y * tiles_z;
}

// Rebuild every replica record and work list from the lattices' current
// state and the groups' current tile candidates, and upload them.
ghip_status rebuild(ghip_batch *b) {
    HIP_TRY(hipStreamSynchronize(b->stream));
    b->drop_graph();
    const size_t R = b->sims.size();
    b->pool.clear();
    std::vector<uint32_t> sweep_blocks(R, 0), field_blocks(R, 0);
    for (auto &g : b->groups) {
        g.lds = 0;
        g.threads = (g.d3 ? kTiles3D : kTiles2D)[g.cand].threads;
        for (uint32_t r : g.reps) {
            const ghip_sim *s = b->sims[r];
            BatchReplica &B = b->reps[r];
            B.sweep = sweep_for(s, g.cand, &sweep_blocks[r]);
            g.lds = std::max(g.lds, sweep_lds_bytes(B.sweep, s->d3));
        }
        std::vector<uint32_t> order = g.reps;
        std::stable_sort(order.begin(), order.end(),
                         [&](uint32_t x, uint32_t y) { return sweep_blocks[x] > sweep_blocks[y]; });
        std::vector<uint32_t> blocks;
        for (uint32_t r : order) blocks.push_back(sweep_blocks[r]);
        g.ref = add_list(b, order, blocks);
    }
    for (size_t r = 0; r < R; ++r) {
        const ghip_sim *s = b->sims[r];
        BatchReplica &B = b->reps[r];
        B.cell = const_cast<ghip_sim *>(s)->cell_args();
        for (int i = 0; i < 2; ++i) {
            B.labels[i] = s->labels[i];
            B.act[i] = s->act[i];
            B.A[i] = s->A[i];
        }
        for (int i = 0; i < 3; ++i) B.D[i] = s->Dl[i];
        for (uint32_t k = 0; k < s->m.n_species; ++k)
            for (int i = 0; i < 2; ++i) B.fbuf[k][i] = s->field[k][i];
        B.type_copy = s->type_copy;
        B.anchor_copy = s->anchor_copy;
        B.seed = s->m.seed;
        B.mcs0 = s->mcs;
        B.pass0 = s->pass;
        B.replica = s->m.replica;
        B.n_sites = (uint32_t)s->N;
        B.colours = s->colours;
        B.has_mot = s->has_mot ? 1 : 0;
        B.has_pop = s->has_pop ? 1 : 0;
        B.flags_base = s->flags;
        B.n_species = (int)s->m.n_species;
        B.lab0 = s->lab_cur;
        B.act0 = s->act_cur;
        for (uint32_t k = 0; k < s->m.n_species; ++k) {
            B.field0[k] = s->field_cur[k];
            int left = s->substeps[k], c = 0;
            while (left > 0) {
                const int take = std::min(left, s->f_kmax);
                B.chunk_steps[k][c++] = take;
                left -= take;
            }
            B.chunks[k] = c;
            B.fd[k] = s->fd[k];
            B.fk[k] = s->fk[k];
        }
        field_blocks[r] = field_tile(s, b->field_cand[s->d3 ? 1 : 0], &B);
    }
    // field launches: chunk c of every species of every lattice of one dimension
    b->fields.clear();
    const uint32_t S = b->args.n_species_max;
    for (int c = 0; c < b->max_chunks; ++c)
        for (int d3 = 0; d3 < 2; ++d3) {
            FieldLaunch fl;
            fl.d3 = d3 != 0;
            fl.chunk = c;
            std::vector<uint32_t> members, blocks;
            for (size_t r = 0; r < R; ++r) {
                const ghip_sim *s = b->sims[r];
                if (s->d3 != fl.d3) continue;
                for (uint32_t k = 0; k < s->m.n_species; ++k) {
                    if (b->reps[r].chunks[k] <= c) continue;
                    members.push_back((uint32_t)r * S + k);
                    blocks.push_back(field_blocks[r]);
                    FieldArgs f{};
                    f.ix_log2 = b->reps[r].fx_log2;
                    f.iy_log2 = b->reps[r].fy_log2;
                    f.iz = b->reps[r].fz;
                    f.steps = b->reps[r].chunk_steps[k][c];
                    fl.lds = std::max(fl.lds, field_lds_bytes(f, s->d3));
                }
            }
            if (members.empty()) continue;
            fl.ref = add_list(b, members, blocks);
            b->fields.push_back(fl);
        }
    // the length term's per-pass prepass
    {
        std::vector<uint32_t> members, blocks;
        for (size_t r = 0; r < R; ++r) {
            const ghip_sim *s = b->sims[r];
            if (!(s->sweep_flags & kLength) || !s->d3) continue;  // planes price it inline
            members.push_back((uint32_t)r);
            blocks.push_back(cdiv(s->cap, 256));
        }
        b->length_ref = add_list(b, members, blocks);
    }
    // per-label launches and the activity decay
    b->cells.clear();
    for (int op = 0; op < kNumCellOps; ++op) {
        std::vector<uint32_t> members, blocks;
        for (size_t r = 0; r < R; ++r) {
            const ghip_sim *s = b->sims[r];
            // a lattice that divides or dies checks for drift in its plan
            const bool take = op == kOpDecay           ? s->has_mot
                              : op_is_population(op)   ? s->has_pop
                              : op == kOpConsolidate || op == kOpCheckRecentre
                                  ? s->tracks_moments && !s->has_pop
                                  : s->tracks_moments;
            if (!take) continue;
            members.push_back((uint32_t)r);
            blocks.push_back(cell_blocks(s, op));
        }
        if (members.empty()) continue;
        CellLaunch cl;
        cl.op = op;
        cl.ref = add_list(b, members, blocks);
        b->cells.push_back(cl);
    }
    if (b->pool.size() > b->pool_cap) {
        if (b->d_pool) HIP_TRY(hipFree(b->d_pool));
        b->d_pool = nullptr;
        b->pool_cap = b->pool.size() + b->pool.size() / 2 + 64;
        HIP_TRY(hipMalloc(&b->d_pool, b->pool_cap * 4));
    }
    HIP_TRY(hipMemcpy(b->d_pool, b->pool.data(), b->pool.size() * 4, hipMemcpyHostToDevice));
    for (auto &g : b->groups) g.w = materialise(b, g.ref);
    for (auto &f : b->fields) f.w = materialise(b, f.ref);
    for (auto &c : b->cells) c.w = materialise(b, c.ref);
    b->length = materialise(b, b->length_ref);
    HIP_TRY(hipMemcpy(b->d_reps, b->reps.data(), R * sizeof(BatchReplica), hipMemcpyHostToDevice));
    // the per-launch records start as the static arguments; the prepare
    // kernels write only what changes from pass to pass
    std::vector<SweepArgs> sweep(R);
    std::vector<CellArgs> cell(R);
    for (size_t r = 0; r < R; ++r) {
        sweep[r] = b->reps[r].sweep;
        cell[r] = b->reps[r].cell;
    }
    HIP_TRY(hipMemcpy(b->d_sweep, sweep.data(), R * sizeof(SweepArgs), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(b->d_cell, cell.data(), R * sizeof(CellArgs), hipMemcpyHostToDevice));
    std::vector<uint64_t> mcs(R);
    for (size_t r = 0; r < R; ++r) mcs[r] = b->sims[r]->mcs;
    HIP_TRY(hipMemcpy(b->d_mcs, mcs.data(), R * 8, hipMemcpyHostToDevice));
    b->expect_mcs = mcs;
    return GHIP_OK;
}

// A timed launch: a sweep group (kind 0, its index) or a field launch (kind
// 1, its dimension's index: 0 planes, 1 volumes).
struct Mark {
    int kind, idx;
    hipEvent_t start, stop;
};

// One step of every lattice. With `marks`, events bracket each sweep and
// field launch and are appended to it.
ghip_status enqueue_step(ghip_batch *b, std::vector<Mark> *marks) {
    const BatchArgs &A = b->args;
    hipStream_t s = b->stream;
    size_t ev = 0;
    auto next_event = [&](hipEvent_t *out) -> hipError_t {
        if (ev == b->events.size()) {
            hipEvent_t e;
            hipError_t err = hipEventCreate(&e);
            if (err != hipSuccess) return err;
            b->events.push_back(e);
        }
        *out = b->events[ev++];
        return hipSuccess;
    };
    for (int ph = 0; ph < b->max_colours; ++ph) {
        HIP_TRY(launch_batch_prepare_sweep(A, ph, s));
        HIP_TRY(launch_batch_length(A, b->length, s));
        for (size_t g = 0; g < b->groups.size(); ++g) {
            const SweepGroup &G = b->groups[g];
            if (ph >= G.colours) continue;
            hipEvent_t e0 = nullptr, e1 = nullptr;
            if (marks) {
                HIP_TRY(next_event(&e0));
                HIP_TRY(next_event(&e1));
                HIP_TRY(hipEventRecord(e0, s));
            }
            HIP_TRY(launch_batch_sweep(A, G.d3, G.full, G.moore, G.w, G.threads, occupancy_lds(G.lds, G.occ),
                                       s));
            if (marks) {
                HIP_TRY(hipEventRecord(e1, s));
                marks->push_back({0, (int)g, e0, e1});
            }
        }
    }
    for (int c = 0; c < b->max_chunks; ++c) {
        HIP_TRY(launch_batch_prepare_field(A, c, s));
        for (const auto &f : b->fields) {
            if (f.chunk != c) continue;
            const int d = f.d3 ? 1 : 0;
            hipEvent_t e0 = nullptr, e1 = nullptr;
            if (marks) {
                HIP_TRY(next_event(&e0));
                HIP_TRY(next_event(&e1));
                HIP_TRY(hipEventRecord(e0, s));
            }
            HIP_TRY(launch_batch_field(A, f.d3, f.w, occupancy_lds(f.lds, b->field_occ[d]), s));
            if (marks) {
                HIP_TRY(hipEventRecord(e1, s));
                marks->push_back({1, d, e0, e1});
            }
        }
    }
    if (b->any_cells || b->any_mot) {
        HIP_TRY(launch_batch_prepare_cells(A, s));
        for (const auto &c : b->cells) HIP_TRY(launch_batch_cell(c.op, A, c.w, s));
    }
    HIP_TRY(launch_batch_advance(A, s));
    return GHIP_OK;
}

// The lattices' host-side counters after `steps` batched steps.
void advance_host(ghip_batch *b, uint64_t steps) {
    for (size_t r = 0; r < b->sims.size(); ++r) {
        ghip_sim *s = b->sims[r];
        const BatchReplica &B = b->reps[r];
        s->mcs += steps;
        s->pass += steps * (uint64_t)s->colours;
        s->attempted += steps * (uint64_t)s->N;
        // an even number of passes per step leaves the label and activity
        // buffers where they were; a field flips once per launch
        for (uint32_t k = 0; k < s->m.n_species; ++k)
            if (B.chunks[k] & 1) s->field_cur[k] ^= (int)(steps & 1);
        b->expect_mcs[r] = s->mcs;
    }
}

ghip_status check_capacity(ghip_batch *b) {
    for (ghip_sim *s : b->sims) {
        if (!s->has_pop) continue;
        uint32_t fl[8];
        HIP_TRY(hipMemcpy(fl, s->flags, 32, hipMemcpyDeviceToHost));
        if (fl[2] || fl[6]) {
            HIP_TRY(hipMemset(s->flags + 2, 0, 4));
            HIP_TRY(hipMemset(s->flags + 6, 0, 4));
            return fail(GHIP_DEVICE, "a lattice in the batch reached the label id limit; "
                                     "divisions were dropped");
        }
    }
    return GHIP_OK;
}

}  // namespace

extern "C" {

ghip_status ghip_batch_create(ghip_sim *const *sims, size_t n, ghip_batch **out) {
    if (!sims || !out || n == 0) return fail(GHIP_BAD_ARG, "a batch needs at least one lattice");
    *out = nullptr;
    for (size_t r = 0; r < n; ++r) {
        if (!sims[r]) return fail(GHIP_BAD_ARG, "a lattice handle is null");
        if (sims[r]->batch) return fail(GHIP_BAD_ARG, "a lattice already belongs to a batch");
        for (size_t q = 0; q < r; ++q)
            if (sims[q] == sims[r]) return fail(GHIP_BAD_ARG, "a lattice appears twice in the batch");
        for (uint32_t k = 0; k < sims[r]->m.n_species; ++k)
            if ((sims[r]->substeps[k] + sims[r]->f_kmax - 1) / sims[r]->f_kmax > kMaxChunks)
                return fail(GHIP_UNSUPPORTED, "a species takes more substeps per step than a batch holds");
    }
    ghip_batch *b = new ghip_batch();
    auto cleanup = [&](ghip_status st) {
        delete b;
        return st;
    };
#define BATCH_TRY(expr)                                                                        \
    do {                                                                                       \
        hipError_t _e = (expr);                                                                \
        if (_e != hipSuccess)                                                                  \
            return cleanup(fail(GHIP_DEVICE, std::string(#expr) + ": " + hipGetErrorString(_e))); \
    } while (0)
    b->sims.assign(sims, sims + n);
    b->reps.assign(n, BatchReplica{});
    uint32_t S = 1;
    for (ghip_sim *s : b->sims) {
        S = std::max<uint32_t>(S, s->m.n_species);
        b->max_colours = std::max(b->max_colours, s->colours);
        b->any_cells = b->any_cells || s->tracks_moments;
        b->any_mot = b->any_mot || s->has_mot;
        int chunks = 0;
        for (uint32_t k = 0; k < s->m.n_species; ++k)
            chunks = std::max(chunks, (s->substeps[k] + s->f_kmax - 1) / s->f_kmax);
        b->max_chunks = std::max(b->max_chunks, chunks);
    }
    // sweep groups by kernel variant
    for (size_t r = 0; r < n; ++r) {
        const ghip_sim *s = b->sims[r];
        const bool moore = !s->d3 && s->n_off == 8;
        SweepGroup *g = nullptr;
        for (auto &x : b->groups)
            if (x.d3 == s->d3 && x.full == s->full && x.moore == moore) g = &x;
        if (!g) {
            b->groups.emplace_back();
            g = &b->groups.back();
            g->d3 = s->d3;
            g->full = s->full;
            g->moore = moore;
            g->colours = s->colours;
        }
        g->reps.push_back((uint32_t)r);
    }
    BATCH_TRY(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    BATCH_TRY(hipMalloc(&b->d_reps, n * sizeof(BatchReplica)));
    BATCH_TRY(hipMalloc(&b->d_mcs, n * 8));
    BATCH_TRY(hipMalloc(&b->d_sweep, n * sizeof(SweepArgs)));
    BATCH_TRY(hipMalloc(&b->d_cell, n * sizeof(CellArgs)));
    BATCH_TRY(hipMalloc(&b->d_field, n * S * sizeof(FieldArgs)));
    BATCH_TRY(hipMalloc(&b->d_act, n * sizeof(float *)));
#undef BATCH_TRY
    b->args.reps = b->d_reps;
    b->args.mcs = b->d_mcs;
    b->args.cur_sweep = b->d_sweep;
    b->args.cur_cell = b->d_cell;
    b->args.cur_field = b->d_field;
    b->args.cur_act = b->d_act;
    b->args.n_reps = (uint32_t)n;
    b->args.n_species_max = S;
    for (ghip_sim *s : b->sims) {
        hipError_t e = hipStreamSynchronize(s->stream);
        if (e != hipSuccess) return cleanup(fail(GHIP_DEVICE, hipGetErrorString(e)));
    }
    ghip_status st = rebuild(b);
    if (st != GHIP_OK) return cleanup(st);
    for (ghip_sim *s : b->sims) {
        s->batch = b;
        s->ever_batched = true;  // the batch's passes do not keep pass_flags
    }
    *out = b;
    return GHIP_OK;
}

ghip_status ghip_batch_destroy(ghip_batch *b) {
    delete b;
    return GHIP_OK;
}

ghip_status ghip_batch_set_tiles(ghip_batch *b, int candidate) {
    if (!b) return fail(GHIP_BAD_ARG, "null handle");
    if (candidate < -1 || candidate >= kCandidates) return fail(GHIP_BAD_ARG, "no such tile candidate");
    if (candidate < 0) {
        b->autotune = true;
        b->tuned = false;
        b->tune_step = 0;
        return GHIP_OK;
    }
    b->autotune = false;
    b->tuned = true;
    for (auto &g : b->groups) {
        g.cand = candidate;
        g.occ = 0;
    }
    for (int d = 0; d < 2; ++d) {
        b->field_cand[d] = candidate;
        b->field_occ[d] = 0;
    }
    return rebuild(b);
}

ghip_status ghip_batch_set_occupancy(ghip_batch *b, int candidate) {
    if (!b) return fail(GHIP_BAD_ARG, "null handle");
    if (candidate < 0 || candidate >= kOccupancy) return fail(GHIP_BAD_ARG, "no such occupancy candidate");
    HIP_TRY(hipStreamSynchronize(b->stream));
    b->drop_graph();
    b->autotune = false;
    b->tuned = true;
    for (auto &g : b->groups) g.occ = candidate;
    for (int d = 0; d < 2; ++d) b->field_occ[d] = candidate;
    return GHIP_OK;
}

}  // extern "C"

namespace {

// `steps` steps of every lattice, none of which can run out of label ids.
ghip_status batch_run(ghip_batch *b, uint64_t steps) {
    uint64_t q = 0;
    // tuning: real steps, one candidate each, timed per sweep group and per
    // dimension's field launches; first the tiles (from the warm-up on, with
    // no occupancy cap), then the occupancy caps on the fastest tiles
    while (q < steps && b->autotune && !b->tuned) {
        const int t = b->tune_step;
        const bool tiles = t <= kCandidates;
        const int cand = t == 0 ? 0 : t - 1;
        const int occ = tiles ? 0 : t - kCandidates;
        bool reshape = false;
        for (auto &g : b->groups) {
            if (tiles && g.cand != cand) {
                g.cand = cand;
                reshape = true;
            }
            g.occ = occ;
        }
        for (int d = 0; d < 2; ++d) {
            if (tiles && b->field_cand[d] != cand) {
                b->field_cand[d] = cand;
                reshape = true;
            }
            b->field_occ[d] = occ;
        }
        if (reshape) {
            ghip_status st = rebuild(b);
            if (st != GHIP_OK) return st;
        }
        std::vector<Mark> marks;
        ghip_status st = enqueue_step(b, &marks);
        if (st != GHIP_OK) return st;
        HIP_TRY(hipStreamSynchronize(b->stream));
        if (t > 0) {
            for (auto &g : b->groups) (tiles ? g.ms[cand] : g.occ_ms[occ]) = 0.0f;
            for (int d = 0; d < 2; ++d) (tiles ? b->field_ms[d][cand] : b->field_occ_ms[d][occ]) = 0.0f;
            for (const Mark &mk : marks) {
                float ms = 0.0f;
                HIP_TRY(hipEventElapsedTime(&ms, mk.start, mk.stop));
                if (mk.kind == 0)
                    (tiles ? b->groups[mk.idx].ms[cand] : b->groups[mk.idx].occ_ms[occ]) += ms;
                else
                    (tiles ? b->field_ms[mk.idx][cand] : b->field_occ_ms[mk.idx][occ]) += ms;
            }
        }
        advance_host(b, 1);
        b->eager_steps += 1;
        ++q;
        ++b->tune_step;
        auto fastest = [](const float *ms, int n) {
            int best = 0;
            for (int c = 1; c < n; ++c)
                if (ms[c] < ms[best]) best = c;
            return best;
        };
        if (b->tune_step == kCandidates + 1) {
            // the tiles are chosen; their uncapped time is occupancy candidate 0
            for (auto &g : b->groups) {
                g.cand = fastest(g.ms, kCandidates);
                g.occ_ms[0] = g.ms[g.cand];
            }
            for (int d = 0; d < 2; ++d) {
                b->field_cand[d] = fastest(b->field_ms[d], kCandidates);
                b->field_occ_ms[d][0] = b->field_ms[d][b->field_cand[d]];
            }
            st = rebuild(b);
            if (st != GHIP_OK) return st;
        } else if (b->tune_step == kCandidates + kOccupancy) {
            for (auto &g : b->groups) g.occ = fastest(g.occ_ms, kOccupancy);
            for (int d = 0; d < 2; ++d) b->field_occ[d] = fastest(b->field_occ_ms[d], kOccupancy);
            b->tuned = true;
        }
    }
    if (q < steps) {
        if (!b->exec) {
            HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            ghip_status st = enqueue_step(b, nullptr);
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(b->stream, &g);
            if (st != GHIP_OK) {
                if (g) (void)hipGraphDestroy(g);
                return st;
            }
            HIP_TRY(e);
            b->graph = g;
            HIP_TRY(hipGraphInstantiateWithFlags(&b->exec, b->graph, 0));
        }
        for (uint64_t k = q; k < steps; ++k) HIP_TRY(hipGraphLaunch(b->exec, b->stream));
        HIP_TRY(hipStreamSynchronize(b->stream));
        advance_host(b, steps - q);
        b->graph_steps += steps - q;
    }
    return GHIP_OK;
}

}  // namespace

extern "C" {

ghip_status ghip_batch_step(ghip_batch *b, uint64_t steps) {
    if (!b) return fail(GHIP_BAD_ARG, "null handle");
    for (ghip_sim *s : b->sims)
        if (s->traj) return fail(GHIP_UNSUPPORTED, "a lattice in the batch is recording its trajectory; "
                                                   "step it alone");
    bool stale = false;
    for (size_t r = 0; r < b->sims.size(); ++r) {
        HIP_TRY(hipStreamSynchronize(b->sims[r]->stream));
        if (b->sims[r]->mcs != b->expect_mcs[r]) stale = true;
    }
    if (stale) {
        ghip_status st = rebuild(b);
        if (st != GHIP_OK) return st;
    }
    while (steps > 0) {
        // every lattice that divides gets room for the chunk; one that had
        // to grow moved its buffers, so the work lists and graph are rebuilt
        uint64_t take = steps;
        bool moved = false;
        for (ghip_sim *s : b->sims) {
            uint64_t chunk = 0;
            bool grew = false;
            ghip_status st = ensure_label_room(s, take, &chunk, &grew);
            if (st != GHIP_OK) return st;
            take = std::min(take, chunk);
            moved = moved || grew;
        }
        if (moved) {
            ghip_status st = rebuild(b);
            if (st != GHIP_OK) return st;
        }
        ghip_status st = batch_run(b, take);
        if (st != GHIP_OK) return st;
        steps -= take;
    }
    return check_capacity(b);
}

ghip_status ghip_batch_describe(const ghip_batch *b, char *out, size_t n) {
    if (!b || !out || n == 0) return fail(GHIP_BAD_ARG, "null argument");
    auto list = [](const float *ms, int k) {
        std::string s;
        char buf[32];
        for (int c = 0; c < k; ++c) {
            std::snprintf(buf, sizeof buf, "%s%.4f", c ? ", " : "", ms[c]);
            s += buf;
        }
        return s;
    };
    std::string s = "{\"lattices\": " + std::to_string(b->sims.size()) +
                    ", \"tuned\": " + (b->tuned ? "true" : "false") +
                    ", \"graph_steps\": " + std::to_string(b->graph_steps) +
                    ", \"eager_steps\": " + std::to_string(b->eager_steps) + ", \"groups\": [";
    for (size_t g = 0; g < b->groups.size(); ++g) {
        const SweepGroup &G = b->groups[g];
        const Tile &t = (G.d3 ? kTiles3D : kTiles2D)[G.cand];
        char buf[320];
        std::snprintf(buf, sizeof buf,
                      "%s{\"dims\": %d, \"full\": %s, \"lattices\": %zu, \"tile\": [%d, %d, %d], "
                      "\"threads\": %d, \"blocks_per_pass\": %u, \"wgs_per_cu\": %d, \"lds_bytes\": %zu, "
                      "\"tuning_ms\": [",
                      g ? ", " : "", G.d3 ? 3 : 2, G.full ? "true" : "false", G.reps.size(), t.tx, t.ty,
                      t.tz, t.threads, G.ref.blocks, kWgsPerCu[G.occ], occupancy_lds(G.lds, G.occ));
        s += buf;
        s += list(G.ms, kCandidates) + "], \"occupancy_ms\": [" + list(G.occ_ms, kOccupancy) + "]}";
    }
    s += "], \"fields\": [";
    bool first = true;
    for (int d = 0; d < 2; ++d) {
        bool any = false;
        for (const auto &f : b->fields) any = any || (f.d3 ? 1 : 0) == d;
        if (!any) continue;
        const Tile &t = (d ? kFieldTiles3D : kFieldTiles2D)[b->field_cand[d]];
        char buf[192];
        std::snprintf(buf, sizeof buf, "%s{\"dims\": %d, \"tile\": [%d, %d, %d], \"wgs_per_cu\": %d, \"tuning_ms\": [",
                      first ? "" : ", ", d ? 3 : 2, t.tx, t.ty, t.tz, kWgsPerCu[b->field_occ[d]]);
        s += buf;
        s += list(b->field_ms[d], kCandidates) + "], \"occupancy_ms\": [" +
             list(b->field_occ_ms[d], kOccupancy) + "]}";
        first = false;
    }
    s += "], \"wgs_per_cu_candidates\": [";
    for (int c = 0; c < kOccupancy; ++c) s += (c ? ", " : "") + std::to_string(kWgsPerCu[c]);
    s += "]}";
    if (s.size() + 1 > n) return fail(GHIP_BAD_ARG, "the buffer is too small");
    std::memcpy(out, s.c_str(), s.size() + 1);
    return GHIP_OK;
}

}  // extern "C"
