// sim.h - the host-side state of one lattice (struct ghip_sim), shared by the
// single-lattice driver (engine.cpp) and the batch driver (batch_host.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <cstring>
#include <string>
#include <vector>

#include "../../include/glazier_hip.h"
#include "ghip_device.h"

namespace ghip {

// Records `msg` as the calling thread's last error and returns `code`.
ghip_status fail(ghip_status code, const std::string &msg);

inline int ilog2(int v) {
    int r = 0;
    while ((1 << (r + 1)) <= v) ++r;
    return r;
}

inline int pow2_floor(int v) { return v < 1 ? 1 : 1 << ilog2(v); }

struct TimedLaunch {
    int family;
    hipEvent_t start, stop;
};

}  // namespace ghip

#define HIP_TRY(expr)                                                                          \
    do {                                                                                       \
        hipError_t _e = (expr);                                                                \
        if (_e != hipSuccess)                                                                  \
            return ::ghip::fail(GHIP_DEVICE, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

struct ghip_sim {
    ghip_model m{};
    int W = 0, H = 0, D = 0;
    size_t N = 0;
    bool d3 = false;
    int colours = 4;
    int n_off = 0;
    int off[26][3] = {};
    hipStream_t stream = nullptr;

    uint32_t *labels[2] = {nullptr, nullptr};
    int lab_cur = 0;
    float *act[2] = {nullptr, nullptr};
    int act_cur = 0;
    float *field[GHIP_MAX_SPECIES][2] = {};
    float *field_base[2] = {nullptr, nullptr};
    int field_cur[GHIP_MAX_SPECIES] = {};
    int64_t *A[2] = {nullptr, nullptr};
    int64_t *Dl[3] = {nullptr, nullptr, nullptr};
    int cw = 1;
    uint32_t cap = 0;
    int32_t *anchor = nullptr;
    uint32_t *ltype = nullptr;
    uint32_t *scratch = nullptr;
    uint32_t *daughter = nullptr;
    uint8_t *plan = nullptr;
    double *geo = nullptr;
    double *len_before = nullptr;  // per label, when the length term is on
    int32_t *nematic = nullptr;    // (Q_xx, Q_xy) per site in units of 2^-24, with the coupling
    bool has_nem = false;
    uint32_t *flags = nullptr;
    uint32_t *type_copy = nullptr;
    int32_t *anchor_copy = nullptr;
    uint32_t *n_labels_dev = nullptr;
    ghip::DeviceTables *tables = nullptr;
    ghip::DeviceTables host_tables{};
    unsigned long long *accepted = nullptr;

    uint64_t pass = 0;
    uint64_t mcs = 0;
    uint64_t attempted = 0;
    uint32_t sweep_flags = 0;
    bool full = false;
    bool has_pop = false, tracks_moments = false, has_mot = false;
    int substeps[GHIP_MAX_SPECIES] = {};
    float fd[GHIP_MAX_SPECIES] = {}, fk[GHIP_MAX_SPECIES] = {};
    int medium_type = 0;

    int tx_log2 = 6, ty_log2 = 3, tz_log2 = 0, threads = 512;
    int tiles_x = 1, tiles_y = 1, tiles_z = 1, blocks = 1;
    bool plain_physics = false;  // the medium, one cell type, integer contact/volume/surface, Moore plane
    bool plain_fast = false;     // the plain-monolayer pass (sweep_plain.hip) applies
    bool resident = false;       // the whole lattice steps inside one workgroup
    // the whole lattice steps in one resident workgroup per tile, passes
    // separated by grid barriers (sweep_plain.hip)
    bool coop = false;
    int coop_tx_log2 = 0, coop_ty_log2 = 0, coop_blocks = 0;
    uint32_t *coop_sync = nullptr;  // kCoopSyncWords: arrival words (a line each), then the error word
    uint32_t coop_epoch = 0;        // the arrival epoch the last launch ended on
    uint32_t coop_patience = ghip::kCoopPatience;
    // a step is its plain streaming passes and nothing else, so the passes
    // can tell from pass_flags when the last two accepted nothing anywhere
    bool quiet_skip = false;
    bool ever_batched = false;   // stepped by a batch at some point: pass_flags no longer apply
    uint32_t *pass_flags = nullptr;
    int cus = 0;                 // compute units of the device
    int fx_log2 = 6, fy_log2 = 5, fz = 1, f_tiles_x = 1, f_tiles_y = 1, f_blocks = 1, f_kmax = 4;

    bool timing = false;         // every launch bracketed by its own events
    bool span_timing = false;    // one event pair around each ghip_step call
    hipEvent_t span_start = nullptr, span_stop = nullptr;
    std::vector<ghip::TimedLaunch> pending;
    std::vector<ghip::TimedLaunch> spare;
    double ms[5] = {0, 0, 0, 0, 0};       // sweep, fields, moments, population, step
    uint64_t launches[5] = {0, 0, 0, 0, 0};

    ghip_batch *batch = nullptr;  // the batch this lattice is stepped in, if any
    uint32_t grown = 0;           // times the per-label buffers moved to a larger capacity

    // owner-type trajectory: one byte per site per recorded step
    uint8_t *traj = nullptr;
    uint64_t traj_cap = 0, traj_len = 0;

    ~ghip_sim() {
        if (stream) (void)hipStreamSynchronize(stream);
        void *bufs[] = {labels[0], labels[1], act[0], act[1], field_base[0], field_base[1], A[0],
                        A[1], Dl[0], Dl[1], Dl[2], anchor, ltype, scratch, daughter, plan, geo,
                        flags, type_copy, anchor_copy, n_labels_dev, tables, accepted, traj, len_before,
                        nematic, pass_flags, coop_sync};
        for (void *p : bufs)
            if (p) (void)hipFree(p);
        for (auto *list : {&pending, &spare})
            for (auto &t : *list) {
                (void)hipEventDestroy(t.start);
                (void)hipEventDestroy(t.stop);
            }
        if (span_start) (void)hipEventDestroy(span_start);
        if (span_stop) (void)hipEventDestroy(span_stop);
        if (stream) (void)hipStreamDestroy(stream);
    }

    ghip::CellArgs cell_args() {
        ghip::CellArgs c{};
        c.labels = labels[lab_cur];
        c.state = A[pass % 2];
        c.anchor = anchor;
        c.n_labels = n_labels_dev;
        c.ltype = ltype;
        c.scratch_u32 = scratch;
        c.daughter = daughter;
        c.plan = plan;
        c.geo = geo;
        c.nematic = nematic;
        // a lattice that divides or dies alternates two sets of flag words
        // by step, each step's plan clearing the next step's set
        c.flags = has_pop ? flags + 4 * (mcs & 1) : flags;
        c.flags_next = has_pop ? flags + 4 * ((mcs + 1) & 1) : nullptr;
        c.state_copy = A[(pass + 1) % 2];
        for (int i = 0; i < 3; ++i) c.deltas[i] = Dl[i];
        c.tables = tables;
        c.w = W;
        c.h = H;
        c.d = D;
        c.cw = cw;
        c.n_off = n_off;
        std::memcpy(c.off, off, sizeof off);
        c.capacity = cap;
        return c;
    }

    // The parts of a colour pass's arguments that do not change from pass to
    // pass (lattice, tables, counters, tile shape of this lattice's own launch).
    ghip::SweepArgs sweep_template() const {
        ghip::SweepArgs a{};
        a.labels = labels[lab_cur];
        a.labels_prev = a.labels;
        a.act = act[act_cur];
        a.act_prev = a.act;
        a.anchor = anchor;
        a.len_before = len_before;
        a.nematic = nematic;
        a.tables = tables;
        a.n_labels = n_labels_dev;
        a.accepted = accepted;
        a.w = W;
        a.h = H;
        a.d = D;
        a.tx_log2 = tx_log2;
        a.ty_log2 = ty_log2;
        a.tz_log2 = tz_log2;
        a.halo = has_mot ? 2 : 1;
        a.n_off = n_off;
        a.n_types = m.n_types;
        a.n_thr = host_tables.n_thr;
        a.cell_words = cw;
        a.n_species = (int)m.n_species;
        a.flags = sweep_flags;
        a.tiles_x = tiles_x;
        a.tiles_y = tiles_y;
        a.tiles_z = tiles_z;
        std::memcpy(a.off, off, sizeof off);
        return a;
    }

    hipError_t begin_timed(int family, ghip::TimedLaunch *out) {
        if (!timing) return hipSuccess;
        if (spare.empty()) {
            ghip::TimedLaunch t{};
            hipError_t e = hipEventCreate(&t.start);
            if (e != hipSuccess) return e;
            e = hipEventCreate(&t.stop);
            if (e != hipSuccess) return e;
            spare.push_back(t);
        }
        *out = spare.back();
        spare.pop_back();
        out->family = family;
        return hipEventRecord(out->start, stream);
    }

    hipError_t end_timed(ghip::TimedLaunch &t) {
        if (!timing) return hipSuccess;
        hipError_t e = hipEventRecord(t.stop, stream);
        pending.push_back(t);
        return e;
    }

    hipError_t collect_timing() {
        for (auto &t : pending) {
            float elapsed = 0.0f;
            hipError_t e = hipEventElapsedTime(&elapsed, t.start, t.stop);
            if (e != hipSuccess) return e;
            ms[t.family] += elapsed;
            launches[t.family] += 1;
            spare.push_back(t);
        }
        pending.clear();
        return hipSuccess;
    }
};

namespace ghip {
// Room for the divisions of the next `steps` (engine.cpp): grows the
// per-label buffers when needed and gives how many steps may run before the
// next check.
ghip_status ensure_label_room(ghip_sim *s, uint64_t steps, uint64_t *chunk, bool *grew);
}  // namespace ghip
