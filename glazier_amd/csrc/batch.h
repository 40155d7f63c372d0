// batch.h - device records and launchers of the batched step (batch.hip),
// driven by batch_host.cpp.
//
// A batch steps many independent lattices ("replicas": other seeds, other
// parameters, other sizes, 2D and 3D mixed) together. Each launch covers
// every replica that takes part in it: block b of the launch maps to
// (replica r, local block lb) through a prefix array of per-replica block
// counts, so a 64^2 replica's few tiles and a 512^2 replica's hundreds fill
// one grid. Whatever changes between launches (colour order, pass key,
// snapshot rotation, ping-pong buffers, death key) is derived on the device
// from a per-replica step counter, so one Monte Carlo step is a fixed list of
// launches that is captured once in a hipGraph and replayed.
#pragma once

#include "ghip_device.h"

namespace ghip {

constexpr int kMaxChunks = 16;  // field launches per step per species

// Static description of one replica: the arguments of its own launches
// (minus what changes per pass) and the buffers those rotate through.
struct BatchReplica {
    SweepArgs sweep;             // tile shape of the batch, per-pass parts unset
    CellArgs cell;               // per-step parts unset
    uint32_t *labels[2];
    float *act[2];
    float *fbuf[kMaxSpecies][2];
    int64_t *A[2];
    int64_t *D[3];
    uint32_t *type_copy;
    int32_t *anchor_copy;
    uint64_t seed;
    uint64_t mcs0, pass0;        // the step and pass the rotations count from
    uint32_t replica;
    uint32_t n_sites;
    int colours;
    int has_mot;
    int has_pop;
    uint32_t *flags_base;  // both sets of flag words
    int n_species;
    int lab0, act0;              // current label and activity buffers at mcs0
    int field0[kMaxSpecies];     // current buffer of each species at mcs0
    int chunks[kMaxSpecies];     // field launches per step of each species
    int chunk_steps[kMaxSpecies][kMaxChunks];
    float fd[kMaxSpecies], fk[kMaxSpecies];
    int fx_log2, fy_log2, fz, f_tiles_x, f_tiles_y;
};

// Per-batch device pointers: the replica records, the step counters, and the
// per-launch argument records the prepare kernels fill.
struct BatchArgs {
    const BatchReplica *reps;
    uint64_t *mcs;               // per replica: the step being taken
    SweepArgs *cur_sweep;        // per replica, this pass
    CellArgs *cur_cell;          // per replica, this step's cell phase
    FieldArgs *cur_field;        // per (replica, species), this field launch
    float **cur_act;             // per replica: activity after the passes
    uint32_t n_reps;
    uint32_t n_species_max;
};

// A launch's work list: members[g] is a replica (or a replica * species_max +
// species entry for fields) and its blocks are [first[g], first[g + 1]).
struct WorkList {
    const uint32_t *first;
    const uint32_t *members;
    uint32_t n;
    uint32_t blocks;
};

// The per-label launches of a step, in the order the single-lattice driver
// enqueues them: for a lattice that divides or dies, the plan (which folds
// the last pass's deltas, clears the next step's flags and checks for
// drift), daughters (with the compaction plan), relabelling, compaction;
// for one that only keeps moments, consolidation (with the flag clear) and
// the drift check; then the recentring, and the activity decay.
enum BatchCellOp : int {
    kOpConsolidate,
    kOpPlan,
    kOpCountMoved,
    kOpAssign,
    kOpApply,
    kOpCompactStage,
    kOpCompactMove,
    kOpCompactRelabel,
    kOpCheckRecentre,
    kOpRecentre1,
    kOpRecountMoments1,
    kOpDecay,
    kNumCellOps
};

hipError_t launch_batch_prepare_sweep(const BatchArgs &b, int phase, hipStream_t s);
hipError_t launch_batch_sweep(const BatchArgs &b, bool d3, bool full, bool moore, const WorkList &w,
                              int threads, size_t lds, hipStream_t s);
hipError_t launch_batch_length(const BatchArgs &b, const WorkList &w, hipStream_t s);
hipError_t launch_batch_prepare_field(const BatchArgs &b, int chunk, hipStream_t s);
hipError_t launch_batch_field(const BatchArgs &b, bool d3, const WorkList &w, size_t lds, hipStream_t s);
hipError_t launch_batch_prepare_cells(const BatchArgs &b, hipStream_t s);
hipError_t launch_batch_cell(int op, const BatchArgs &b, const WorkList &w, hipStream_t s);
hipError_t launch_batch_advance(const BatchArgs &b, hipStream_t s);
// Threads per workgroup of a per-label launch.
int batch_cell_threads(int op);

}  // namespace ghip
