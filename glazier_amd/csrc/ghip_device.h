// ghip_device.h - device-side data layout of the engine: the model tables a
// kernel stages into LDS, and the argument blocks of every kernel family.
#pragma once

#include "ghip_common.h"

namespace ghip {

constexpr int kMaxTypes = 16;
constexpr int kMaxSpecies = 4;
constexpr int kMaxThr = 1024;
constexpr int kAcceptSlots = 64;

// Model parameters as the kernels read them, in the device type index: type
// 0 is the medium, whose contact row is the reference's cell_type[0] row.
struct DeviceTables {
    float J[kMaxTypes][kMaxTypes];
    float lam_v[kMaxTypes];
    float tgt_v[kMaxTypes];
    float lam_s[kMaxTypes];
    float tgt_s[kMaxTypes];
    float chi[kMaxTypes][kMaxSpecies];
    float max_act[kMaxTypes];
    float coef_act[kMaxTypes];
    float ext[kMaxTypes][3];
    float sec[kMaxTypes][kMaxSpecies];
    float up[kMaxTypes][kMaxSpecies];
    double lam_len[kMaxTypes];
    double tgt_len[kMaxTypes];
    double lam_nem[kMaxTypes];
    double div_vol[kMaxTypes];
    double death[kMaxTypes];
    uint32_t connected[kMaxTypes];
    uint32_t piece2d[8];      // bit m: the planar ring mask m is one piece
    uint32_t thr[kMaxThr];
    uint32_t n_thr;
    float accept_c;
    uint32_t pad[2];
};

// Feature bits of a sweep launch.
enum SweepFlags : uint32_t {
    kIntegral = 1u << 0,
    kConnected = 1u << 1,
    kLength = 1u << 2,
    kChemotaxis = 1u << 3,
    kMotility = 1u << 4,
    kDrift = 1u << 5,
    kMoments = 1u << 6,  // per-label moment sums are kept (length, nematic or division)
    kNematic = 1u << 7,  // the nematic coupling: record words 10, 11 hold the field sums
};

// The nematic field is held per site as two int32 in units of 2^-24, so a
// label's field sum is an exact int64 in any order of accumulation.
constexpr double kNematicScale = 16777216.0;       // 2^24
constexpr double kNematicUnit = 5.9604644775390625e-8;  // 2^-24

// One colour pass over one lattice.
struct SweepArgs {
    uint32_t *labels;            // packed label words, n sites
    const uint32_t *labels_prev; // pass-start copy, read at distance two (motility)
    float *act;                  // activity, written on accept
    const float *act_prev;       // pass-start copy
    const float *fields[kMaxSpecies]; // current buffer of each species
    const int64_t *a_cur;        // per-label base snapshot
    const int64_t *e_prev;       // delta of the previous pass
    int64_t *f_acc;              // delta this pass accumulates
    int64_t *a_next;             // base of the next pass: a_cur + e_prev
    int64_t *g_zero;             // delta buffer of the pass after next, cleared
    const int32_t *anchor;       // 3 per label
    const double *len_before;    // per label: length minus target in this pass's snapshot
    const int32_t *nematic;      // (Q_xx, Q_xy) per site in units of 2^-24, or null
    const DeviceTables *tables;
    const uint32_t *n_labels;    // device label count
    unsigned long long *accepted; // kAcceptSlots counters
    int w, h, d;
    int cx, cy, cz;
    uint32_t key, key2;
    int tx_log2, ty_log2, tz_log2; // tile size in targets
    int halo;
    int n_off;
    uint32_t n_types;
    uint32_t n_thr;
    int cell_words;
    int n_species;
    uint32_t flags;
    uint32_t fold_per_block;
    int xcd_chunk;               // tiles per XCD in the remap, 0 for none
    int reverse;                 // walk the tiles in reverse order (every other pass)
    uint32_t *pass_flags;        // plain pass: four words, word p % 4 nonzero once pass p accepts
                                 // a copy (or unknown); null to take every pass whole
    uint32_t pass_slot;          // this pass's index mod 4
    int merge_deltas;            // plain pass: merge a tile's record changes in LDS first
    uint32_t fold_words;         // plain pass: record words to fold when the label count is fixed
                                 // (no division or death), 0 to read it on the device
    int tiles_x, tiles_y, tiles_z;
    int off[26][3];
};

// Tile of block `blk` in a pass over `blocks` tiles. The hardware deals
// workgroups to the eight XCDs in turn (block b to XCD b mod 8); with
// `xcd_chunk` > 0 each XCD walks a contiguous run of tiles, so the halo rows
// its tiles share stay in its L2. With `reverse` set the run is walked
// backwards (the whole order when there is no run). The engine sets it on
// every other pass: a pass then starts on the tiles the previous pass ended
// with, whose lines are still in each XCD's L2. The draws do not depend on
// the order, so neither does any result.
__device__ __forceinline__ int pass_tile(int blk, int xcd_chunk, int reverse, int blocks) {
    if (xcd_chunk > 0) {
        int run = blk >> 3;
        if (reverse) run = xcd_chunk - 1 - run;
        return (blk & 7) * xcd_chunk + run;
    }
    return reverse ? blocks - 1 - blk : blk;
}

// Scalars of the plain-monolayer pass: the medium and one cell type.
struct PlainParams {
    float j[2][2];          // device contact table
    float lam_v, tgt_v, lam_s, tgt_s;
    const uint32_t *thr;    // integral acceptance table
    uint32_t n_thr;
    uint32_t needs_medium;  // whether a medium neighbour can change a price
    uint32_t tx_magic;      // ceil(2^32 / tiles_x): b / tiles_x = umulhi(b, tx_magic) while
                            // tiles_x * blocks < 2^32 (the tile choice ensures it; 0 when tiles_x is 1)
};

// A small plain lattice stepped whole in one workgroup (sweep_plain.hip):
// `steps` Monte Carlo steps from step `mcs0`, the per-label records read as
// a + e at the first pass and left as a_end (with both delta buffers the
// next pass touches cleared) after the last.
struct PlainResidentArgs {
    uint32_t *labels;             // w x h packed words, read and written back
    const int64_t *a_start, *e_start;
    int64_t *a_end;
    int64_t *e_clear0, *e_clear1;
    const uint32_t *n_labels;
    unsigned long long *accepted; // kAcceptSlots counters
    uint64_t seed, mcs0;
    uint32_t replica, steps;
    int w_log2, h;
    uint32_t capacity;            // label records the lattice may name
};

// A mid-size plain lattice stepped whole by one resident workgroup per tile
// (sweep_plain.hip, a cooperative launch): `steps` Monte Carlo steps (at
// most 64) from pass `pass0`, passes
// separated by a grid barrier instead of launches. Records follow the
// streaming passes' rotation through a and d[pass % 3], so the launch leaves
// the state the next streaming pass expects: a_end holds the records as the
// last pass priced them, d[(end + 2) % 3] that pass's changes, d[end % 3]
// cleared.
constexpr int kCoopMaxBlocks = 64;                   // one wavefront polls every arrival word
constexpr int kCoopArriveStride = 32;                // words between arrival words (one line each)
constexpr int kCoopErrWord = kCoopMaxBlocks * kCoopArriveStride;
constexpr int kCoopSyncWords = kCoopErrWord + 32;
constexpr uint32_t kCoopPatience = 1u << 22;  // a barrier's polls before it gives up, by default
struct PlainCoopArgs {
    uint32_t *labels;             // w x h packed words, current when the launch starts and ends
    const int64_t *a_start;
    int64_t *a_end;               // the same buffer as a_start (passes come in fours)
    int64_t *d[3];                // the pass delta buffers, by pass number mod 3
    const uint32_t *n_labels;
    unsigned long long *accepted; // kAcceptSlots counters
    uint32_t *arrive;             // one epoch word per workgroup, kCoopArriveStride words apart
    uint32_t *err;                // set when a barrier's wait gave up
    uint64_t seed, mcs0, pass0;
    uint32_t replica, steps;
    uint32_t epoch0;              // the arrival epoch the previous launch ended on
    uint32_t patience;            // polls a barrier makes before it gives up (0: gives up at once)
    int w_log2, h;
    int tx_log2, ty_log2;         // tile sides
};

// One launch of the field step: `steps` substeps of one species.
struct FieldArgs {
    const float *in;        // one species, n sites
    float *out;
    const uint32_t *labels; // packed words; type in the top bits
    const DeviceTables *tables;
    int w, h, d;
    int ix_log2, iy_log2, iz; // interior tile: 2^ix x 2^iy x iz sites
    int steps;                // substeps this launch takes
    int exchange;             // fuse secretion/uptake after the last one
    int species;
    float dcoef, kcoef;
    int tiles_x, tiles_y;
};

// The field step of every species of a plane in one launch: sp[0..n) with
// the same lattice, tile grid and substep count.
struct FieldPlanes {
    FieldArgs sp[kMaxSpecies];
    int n;
};

// Per-label bookkeeping launches (recounts, anchors, division/death).
// A recount workgroup covers kRecountRun rows of kRecountCols columns (one
// lane per column), summing runs of one label down a column in registers.
constexpr int kRecountRun = 8;
constexpr int kRecountCols = 256;
inline uint32_t recount_blocks(int w, int h, int d) {
    return (uint32_t)((w + kRecountCols - 1) / kRecountCols) *
           (uint32_t)(((size_t)h * (size_t)d + kRecountRun - 1) / kRecountRun);
}

struct CellArgs {
    uint32_t *labels;      // packed words
    int64_t *state;        // consolidated per-label state
    int32_t *anchor;       // 3 per label
    uint32_t *n_labels;    // device label count
    uint32_t *ltype;       // per label: its packed word (type bits | id)
    uint32_t *scratch_u32; // first sites / moved counts / remap, capacity entries
    uint32_t *daughter;    // parent -> daughter id, capacity entries
    uint8_t *plan;         // capacity entries
    double *geo;           // 6 per label
    const int32_t *nematic; // (Q_xx, Q_xy) per site in units of 2^-24, or null
    uint32_t *flags;       // [0] event, [1] recentre, [2] overflow, [3] compacted count
    uint32_t *flags_next;  // the next step's flag words (cleared by this step's plan), or null
    int64_t *state_copy;   // staging for records moved by a compaction (the other base buffer)
    int64_t *deltas[3];    // the pass delta buffers, cleared for a daughter's fresh id
    const DeviceTables *tables;
    int w, h, d;
    int cw;
    int n_off;
    int off[26][3];
    uint32_t capacity;
    uint32_t key, key2;    // death draw key of this step
};

// Launchers, one per kernel family (sweep.hip, fields.hip, cells.hip).
hipError_t launch_sweep(const SweepArgs &a, bool d3, bool full, int blocks, int threads,
                        hipStream_t stream);
size_t sweep_lds_bytes(const SweepArgs &a, bool d3);
hipError_t launch_sweep_plain2d(const SweepArgs &a, const PlainParams &p, int blocks,
                                hipStream_t stream);
bool plain2d_supported(uint32_t n_thr);
// Tile of the plain pass for a w x h lattice (0 x 0: the pass does not apply).
void plain2d_tile(int w, int h, int *px, int *py);
// Whether a w x h lattice with `capacity` label records fits the resident pass.
bool plain2d_resident_fits(int w, int h, uint32_t capacity);
hipError_t launch_sweep_plain2d_resident(const PlainResidentArgs &r, const PlainParams &p,
                                         hipStream_t stream);
// The tiled resident pass's tile for a w x h lattice with `capacity`
// label records: false when it does not apply. *blocks receives the number
// of workgroups (tiles) of its launch.
bool plain2d_coop_shape(int w, int h, uint32_t capacity, int *tx_log2, int *ty_log2, int *blocks);
// Whether `blocks` workgroups of the tiled resident pass fit the device at
// once (its cooperative launch needs every one resident).
bool plain2d_coop_resident(int blocks, int cus);
// Barriers one launch of `steps` steps passes (the arrival epochs it uses).
inline uint32_t plain2d_coop_barriers(uint32_t steps) { return 4u * steps - 1u; }
hipError_t launch_sweep_plain2d_coop(const PlainCoopArgs &r, const PlainParams &p, int blocks,
                                     hipStream_t stream);
hipError_t launch_field(const FieldArgs &a, bool d3, int blocks, hipStream_t stream);
size_t field_lds_bytes(const FieldArgs &a, bool d3);
// The plane field body's interior tile; launch_field_planes (and the batch's
// field launch) take it on a plane of at least one such tile.
constexpr int kFieldIX = 64, kFieldIY = 32;
__host__ __device__ inline bool field_planes_apply(int w, int h, bool d3) {
    return !d3 && w >= kFieldIX && h >= kFieldIY;
}
hipError_t launch_field_planes(const FieldPlanes &p, hipStream_t stream);
hipError_t launch_activity_decay(float *act, size_t n, hipStream_t stream);
// Each label's length minus its type's target length in the snapshot
// a + e (the length term's `before`, the same for every attempt of a pass).
hipError_t launch_length_prepass(const int64_t *a, const int64_t *e, const uint32_t *n_labels, int cw,
                                 const uint32_t *ltype, const DeviceTables *tables, bool d3, double *out,
                                 uint32_t capacity, hipStream_t stream);
hipError_t launch_fill_i64(int64_t *p, size_t n, hipStream_t s);
hipError_t launch_fill_u32(uint32_t *p, size_t n, uint32_t v, hipStream_t s);
// A[i] += E[i], E[i] = 0 over the live labels' words; with `clear4` it also
// zeroes those four flag words.
hipError_t launch_consolidate(int64_t *a, int64_t *e, const uint32_t *n_labels, int cw, size_t cap,
                              hipStream_t s, uint32_t *clear4 = nullptr);
hipError_t launch_recount(const CellArgs &a, size_t n_sites, hipStream_t s);
// Recount of the moment words (1..cw-1, nematic sums included) about the
// current anchors, volume and surface kept.
hipError_t launch_recount_moments(const CellArgs &a, size_t n_sites, hipStream_t s);
hipError_t launch_initial_anchors(const CellArgs &a, size_t n_sites, hipStream_t s);
hipError_t launch_recentre_check(const CellArgs &a, size_t n_sites, hipStream_t s);
hipError_t launch_population(const CellArgs &a, int64_t *last_deltas, size_t n_sites, uint32_t *type_copy,
                             int32_t *anchor_copy, hipStream_t s);

}  // namespace ghip
