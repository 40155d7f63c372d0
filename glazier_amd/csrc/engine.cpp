// engine.cpp - host driver of the MI355X Cellular Potts engine and its C ABI
// (include/glazier_hip.h).
//
// Replaces GpuSimulation (glazier-cuda/src/lib.rs:749-1569): it owns the
// device buffers, builds the model tables once, and enqueues each Monte
// Carlo step on one HIP stream: the colour passes, the field step, activity
// decay, and division/death, with no host synchronisation inside a call to
// ghip_step. Kernels are precompiled for gfx950; nothing is compiled at run
// time.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/glazier_hip.h"
#include "blueprint.h"
#include "fit.h"
#include "ghip_device.h"
#include "sim.h"

using namespace ghip;

namespace {

thread_local std::string g_last_error;

bool is_integer(double v) { return v == std::floor(v) && std::fabs(v) < 1048576.0; }

const char *kFamilies[] = {"sweep", "fields", "moments", "population", "step"};
constexpr int kStepFamily = 4;

}  // namespace

namespace ghip {
ghip_status fail(ghip_status code, const std::string &msg) {
    g_last_error = msg;
    return code;
}
}  // namespace ghip

namespace {

// The device tables from the model, in the device type index (type 0 is the
// medium, pricing contact as the reference's cell_type[0]).
void build_tables(ghip_sim *s) {
    const ghip_model &m = s->m;
    DeviceTables &t = s->host_tables;
    std::memset(&t, 0, sizeof t);
    const int nt = (int)m.n_types;
    auto map = [&](int i) { return i == 0 ? s->medium_type : i; };
    bool integral = true;
    for (int a = 0; a < nt; ++a)
        for (int b = 0; b < nt; ++b) {
            double j = m.contact[map(a) * nt + map(b)];
            t.J[a][b] = (float)j;
        }
    for (int a = 0; a < nt * nt; ++a)
        if (!is_integer(m.contact[a])) integral = false;
    bool conn = false, len = false, chem = false, mot = false, drift = false, pop = false, nem = false;
    for (int i = 0; i < nt; ++i) {
        const ghip_cell_type &c = m.types[i];
        t.lam_v[i] = (float)c.lambda_volume;
        t.tgt_v[i] = (float)c.target_volume;
        t.lam_s[i] = (float)c.lambda_surface;
        t.tgt_s[i] = (float)c.target_surface;
        if (!is_integer(c.lambda_volume) || !is_integer(c.target_volume) ||
            !is_integer(c.lambda_surface) || !is_integer(c.target_surface))
            integral = false;
        t.max_act[i] = (float)c.max_activity;
        t.coef_act[i] = c.max_activity > 0.0 ? (float)(c.lambda_activity / c.max_activity) : 0.0f;
        for (int a = 0; a < 3; ++a) t.ext[i][a] = (float)c.external[a];
        for (uint32_t k = 0; k < m.n_species; ++k) {
            t.chi[i][k] = (float)m.chemotaxis[i][k];
            t.sec[i][k] = (float)m.secretion[i][k];
            t.up[i][k] = (float)m.uptake[i][k];
            if (m.chemotaxis[i][k] != 0.0) chem = true;
        }
        t.lam_len[i] = c.lambda_length;
        t.tgt_len[i] = c.target_length;
        t.lam_nem[i] = m.has_nematic ? c.lambda_nematic : 0.0;
        if (m.has_nematic && c.lambda_nematic != 0.0) nem = true;
        t.div_vol[i] = c.division_volume;
        t.death[i] = c.death_rate;
        t.connected[i] = c.connected ? 1u : 0u;
        if (c.connected) conn = true;
        if (c.lambda_length != 0.0) len = true;
        if (c.lambda_activity != 0.0 && c.max_activity > 0.0) mot = true;
        if (c.external[0] != 0.0 || c.external[1] != 0.0 || c.external[2] != 0.0) drift = true;
        if (c.division_volume > 0.0 || c.death_rate > 0.0) pop = true;
    }
    if (len || chem || mot || drift || nem) integral = false;
    const double lim = m.temperature * 24.0 * std::log(2.0);
    if (integral && lim + 1.0 < (double)kMaxThr) {
        t.n_thr = (uint32_t)std::floor(lim) + 1;
        t.thr[0] = 1u << 24;
        for (uint32_t i = 1; i < t.n_thr; ++i)
            t.thr[i] = (uint32_t)std::ceil(16777216.0 * std::exp(-(double)i / m.temperature));
    } else {
        integral = false;
        t.n_thr = 0;
    }
    t.accept_c = (float)(1.4426950408889634 / m.temperature);
    // connectivity: one-piece table of the planar ring (the volume's ring is
    // grown in registers)
    int ppos[8][2], q = 0;
    for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
            if (!dx && !dy) continue;
            ppos[q][0] = dx;
            ppos[q][1] = dy;
            ++q;
        }
    for (uint32_t present = 0; present < 256; ++present) {
        bool one = true;
        if (present && (present & (present - 1))) {
            uint32_t reached = present & (~present + 1u);
            for (;;) {
                uint32_t grown = reached;
                for (int a = 0; a < 8; ++a) {
                    if (!(reached & (1u << a))) continue;
                    for (int b = 0; b < 8; ++b)
                        if ((present & (1u << b)) && std::abs(ppos[a][0] - ppos[b][0]) <= 1 &&
                            std::abs(ppos[a][1] - ppos[b][1]) <= 1)
                            grown |= 1u << b;
                }
                if (grown == reached) break;
                reached = grown;
            }
            one = reached == present;
        }
        if (one) t.piece2d[present >> 5] |= 1u << (present & 31);
    }
    s->sweep_flags = 0;
    if (integral) s->sweep_flags |= kIntegral;
    if (conn) s->sweep_flags |= kConnected;
    if (len) s->sweep_flags |= kLength;
    if (chem) s->sweep_flags |= kChemotaxis;
    if (mot) s->sweep_flags |= kMotility;
    if (drift) s->sweep_flags |= kDrift;
    if (nem) s->sweep_flags |= kNematic;
    s->has_pop = pop;
    s->has_mot = mot;
    s->has_nem = nem;
    // model.rs:127-130: the length term, the nematic term and division read the moments
    s->tracks_moments = len || pop || nem;
    if (s->tracks_moments) s->sweep_flags |= kMoments;
    s->full = conn || len || chem || mot || drift || s->tracks_moments;
}

int substeps_for(const ghip_model &m, uint32_t k) {
    const ghip_species &sp = m.species[k];
    if (sp.substeps > 0) return sp.substeps;
    if (sp.diffusion <= 0.0) return 1;
    double dims = m.depth > 1 ? 3.0 : 2.0;
    double c = std::ceil(sp.diffusion / (1.0 / (2.0 * dims)));
    return c < 1.0 ? 1 : (int)c;
}

// Tile shapes: a 2D pass runs 64 x 8 targets (128 x 16 sites, 512 lanes) and
// shrinks to the lattice; a 3D pass runs 16 x 4 x 4 targets. (Cutting small
// lattices' tiles down to one wave, to give every CU a workgroup, measured
// slower: 44 against 36 us a pass on a 64 x 64 x 32 volume.)
void choose_tiles(ghip_sim *s) {
    if (!s->d3) {
        int tx = std::min(64, pow2_floor(s->W / 2));
        int ty = std::min(pow2_floor(512 / tx), pow2_floor(s->H / 2));
        s->tx_log2 = ilog2(tx);
        s->ty_log2 = ilog2(ty);
        s->tz_log2 = 0;
        s->threads = std::max(64, tx * ty);
        s->tiles_x = (s->W / 2 + tx - 1) / tx;
        s->tiles_y = (s->H / 2 + ty - 1) / ty;
        s->tiles_z = 1;
    } else {
        int tx = std::min(16, pow2_floor(s->W / 2));
        int ty = std::min(4, pow2_floor(s->H / 2));
        int tz = std::min(pow2_floor(256 / (tx * ty)), pow2_floor(s->D / 2));
        s->tx_log2 = ilog2(tx);
        s->ty_log2 = ilog2(ty);
        s->tz_log2 = ilog2(tz);
        s->threads = std::max(64, tx * ty * tz);
        s->tiles_x = (s->W / 2 + tx - 1) / tx;
        s->tiles_y = (s->H / 2 + ty - 1) / ty;
        s->tiles_z = (s->D / 2 + tz - 1) / tz;
    }
    s->threads = ((s->threads + 63) / 64) * 64;
    s->blocks = s->tiles_x * s->tiles_y * s->tiles_z;
    int px = 0, py = 0;
    plain2d_tile(s->W, s->H, &px, &py);
    s->plain_physics = !s->d3 && s->n_off == 8 && s->m.n_types == 2 && (s->sweep_flags & kIntegral) &&
                       !s->full && plain2d_supported(s->host_tables.n_thr);
    s->plain_fast = s->plain_physics && px > 0;
    if (s->plain_fast) {
        s->tiles_x = s->W / px;
        s->tiles_y = s->H / py;
        s->tiles_z = 1;
        s->blocks = s->tiles_x * s->tiles_y;
        s->threads = px == 256 ? 512 : 256;
    }
    // fields: 64 x 32 interior in a plane, 16 x 8 x 8 in a volume
    if (!s->d3) {
        s->fx_log2 = ilog2(std::min(64, pow2_floor(s->W)));
        s->fy_log2 = ilog2(std::min(32, pow2_floor(s->H)));
        s->fz = 1;
        s->f_kmax = 4;
    } else {
        s->fx_log2 = ilog2(std::min(16, pow2_floor(s->W)));
        s->fy_log2 = ilog2(std::min(8, pow2_floor(s->H)));
        s->fz = std::min(8, pow2_floor(s->D));
        s->f_kmax = 2;
    }
    s->f_tiles_x = (s->W + (1 << s->fx_log2) - 1) >> s->fx_log2;
    s->f_tiles_y = (s->H + (1 << s->fy_log2) - 1) >> s->fy_log2;
    int f_tiles_z = s->d3 ? (s->D + s->fz - 1) / s->fz : 1;
    s->f_blocks = s->f_tiles_x * s->f_tiles_y * f_tiles_z;
}

PlainParams plain_params(const ghip_sim *s) {
    PlainParams pp{};
    for (int x = 0; x < 2; ++x)
        for (int y = 0; y < 2; ++y) pp.j[x][y] = s->host_tables.J[x][y];
    pp.lam_v = s->host_tables.lam_v[1];
    pp.tgt_v = s->host_tables.tgt_v[1];
    pp.lam_s = s->host_tables.lam_s[1];
    pp.tgt_s = s->host_tables.tgt_s[1];
    pp.thr = s->tables->thr;
    pp.n_thr = s->host_tables.n_thr;
    // with both contact rows flat across the medium and the cell type, a
    // medium neighbour prices like a cell one and the medium count drops out
    pp.needs_medium = !(pp.j[0][0] == pp.j[0][1] && pp.j[1][0] == pp.j[1][1]);
    // a block's tile row by a multiply (exact while tiles_x * blocks < 2^32,
    // which plain2d_tile ensures); one tile column needs none
    const uint64_t tx = (uint64_t)std::max(1, s->tiles_x);
    pp.tx_magic = tx == 1 ? 0u : (uint32_t)((((uint64_t)1 << 32) + tx - 1) / tx);
    return pp;
}

// Steps a small plain lattice inside one workgroup, at most kResidentChunk
// steps per launch so no launch runs long. Pass and step counters advance as
// the per-pass driver's would; the record rotation lands where it would.
constexpr uint64_t kResidentChunk = 64;

// The nematic field as the device holds it: each value rounded to the nearest
// multiple of 2^-24 (ties away from zero), |value| below 128 so it fits int32.
bool nematic_to_fixed(const double *q, size_t n_sites, std::vector<int32_t> *out, std::string *err) {
    out->resize(2 * n_sites);
    for (size_t i = 0; i < 2 * n_sites; ++i) {
        if (!(std::fabs(q[i]) < 128.0)) {
            *err = "nematic value " + std::to_string(q[i]) + " at site " + std::to_string(i / 2) +
                   " lies outside (-128, 128)";
            return false;
        }
        (*out)[i] = (int32_t)std::llround(q[i] * kNematicScale);
    }
    return true;
}

ghip_status run_resident(ghip_sim *s, uint64_t steps) {
    const PlainParams pp = plain_params(s);
    while (steps > 0) {
        const uint32_t take = (uint32_t)std::min<uint64_t>(steps, kResidentChunk);
        const uint64_t p0 = s->pass, p1 = p0 + 4 * (uint64_t)take;
        PlainResidentArgs r{};
        r.labels = s->labels[s->lab_cur];
        r.a_start = s->A[p0 % 2];
        r.e_start = s->Dl[(p0 + 2) % 3];
        r.a_end = s->A[p1 % 2];
        r.e_clear0 = s->Dl[(p1 + 2) % 3];
        r.e_clear1 = s->Dl[p1 % 3];
        r.n_labels = s->n_labels_dev;
        r.accepted = s->accepted;
        r.seed = s->m.seed;
        r.mcs0 = s->mcs;
        r.replica = s->m.replica;
        r.steps = take;
        r.w_log2 = ilog2(s->W);
        r.h = s->H;
        r.capacity = s->cap;
        TimedLaunch t{};
        HIP_TRY(s->begin_timed(0, &t));
        HIP_TRY(launch_sweep_plain2d_resident(r, pp, s->stream));
        HIP_TRY(s->end_timed(t));
        s->pass = p1;
        s->mcs += take;
        s->attempted += (uint64_t)s->N * take;
        steps -= take;
    }
    return GHIP_OK;
}

// Steps a mid-size plain lattice with one resident workgroup per tile, at
// most kResidentChunk steps per launch. The launch leaves the records in the
// rotation a streaming pass expects (the last pass's changes unfolded), and
// the arrival epoch advances by the barriers it passed.
ghip_status run_coop(ghip_sim *s, uint64_t steps) {
    const PlainParams pp = plain_params(s);
    while (steps > 0) {
        const uint32_t take = (uint32_t)std::min<uint64_t>(steps, kResidentChunk);
        const uint32_t barriers = plain2d_coop_barriers(take);
        if (s->coop_epoch > 0x7FFFFFFFu - barriers) {
            // restart the epochs before the wrap-safe comparison could fail
            HIP_TRY(launch_fill_u32(s->coop_sync, kCoopErrWord, 0u, s->stream));
            s->coop_epoch = 0;
        }
        PlainCoopArgs r{};
        r.labels = s->labels[s->lab_cur];
        r.a_start = s->A[s->pass % 2];
        r.a_end = s->A[s->pass % 2];
        for (int i = 0; i < 3; ++i) r.d[i] = s->Dl[i];
        r.n_labels = s->n_labels_dev;
        r.accepted = s->accepted;
        r.arrive = s->coop_sync;
        r.err = s->coop_sync + kCoopErrWord;
        r.seed = s->m.seed;
        r.mcs0 = s->mcs;
        r.pass0 = s->pass;
        r.replica = s->m.replica;
        r.steps = take;
        r.epoch0 = s->coop_epoch;
        r.patience = s->coop_patience;
        r.w_log2 = ilog2(s->W);
        r.h = s->H;
        r.tx_log2 = s->coop_tx_log2;
        r.ty_log2 = s->coop_ty_log2;
        TimedLaunch t{};
        HIP_TRY(s->begin_timed(0, &t));
        HIP_TRY(launch_sweep_plain2d_coop(r, pp, s->coop_blocks, s->stream));
        HIP_TRY(s->end_timed(t));
        s->coop_epoch += barriers;
        s->pass += 4 * (uint64_t)take;
        s->mcs += take;
        s->attempted += (uint64_t)s->N * take;
        steps -= take;
    }
    return GHIP_OK;
}

ghip_status run_sweep_pass(ghip_sim *s, int colour) {
    SweepArgs a = s->sweep_template();
    const uint64_t p = s->pass;
    const bool mot = s->has_mot;
    if (mot) {
        a.labels_prev = s->labels[s->lab_cur];
        a.labels = s->labels[1 - s->lab_cur];
        a.act_prev = s->act[s->act_cur];
        a.act = s->act[1 - s->act_cur];
    }
    for (uint32_t k = 0; k < s->m.n_species; ++k) a.fields[k] = s->field[k][s->field_cur[k]];
    a.a_cur = s->A[p % 2];
    a.e_prev = s->Dl[(p + 2) % 3];
    a.f_acc = s->Dl[p % 3];
    a.a_next = s->A[(p + 1) % 2];
    a.g_zero = s->Dl[(p + 1) % 3];
    a.cx = colour & 1;
    a.cy = (colour >> 1) & 1;
    a.cz = (colour >> 2) & 1;
    a.key = pass_key(s->m.seed, s->m.replica, s->mcs, (uint32_t)colour);
    a.key2 = second_key(a.key);
    const uint64_t words = (uint64_t)s->cap * s->cw;
    a.fold_per_block = (uint32_t)((words + s->blocks - 1) / s->blocks);
    a.xcd_chunk = (s->blocks % 8 == 0) ? s->blocks / 8 : 0;
    a.reverse = (int)(p & 1);
    if (s->quiet_skip && !s->ever_batched) {
        a.pass_flags = s->pass_flags;
        a.pass_slot = (uint32_t)(p & 3);
    }
    // more tiles than four workgroups on every CU: more than one round
    a.merge_deltas = s->blocks > 4 * s->cus ? 1 : 0;
    // without division or death the label count is fixed: every record is live
    a.fold_words = s->has_pop ? 0u : (uint32_t)words;
    TimedLaunch t{};
    HIP_TRY(s->begin_timed(0, &t));
    if ((s->sweep_flags & kLength) && s->d3)
        HIP_TRY(launch_length_prepass(a.a_cur, a.e_prev, s->n_labels_dev, s->cw, s->ltype, s->tables, s->d3,
                                      s->len_before, s->cap, s->stream));
    if (s->plain_fast) {
        HIP_TRY(launch_sweep_plain2d(a, plain_params(s), s->blocks, s->stream));
    } else {
        HIP_TRY(launch_sweep(a, s->d3, s->full, s->blocks, s->threads, s->stream));
    }
    HIP_TRY(s->end_timed(t));
    s->pass += 1;
    if (mot) {
        s->lab_cur = 1 - s->lab_cur;
        s->act_cur = 1 - s->act_cur;
    }
    return GHIP_OK;
}

// One chunk of `take` substeps of species k (the exchange fused into the
// species' last chunk), reading the current buffer and writing the other.
FieldArgs field_chunk(const ghip_sim *s, uint32_t k, int take, bool last) {
    FieldArgs f{};
    f.in = s->field[k][s->field_cur[k]];
    f.out = s->field[k][1 - s->field_cur[k]];
    f.labels = s->labels[s->lab_cur];
    f.tables = s->tables;
    f.w = s->W;
    f.h = s->H;
    f.d = s->D;
    f.ix_log2 = s->fx_log2;
    f.iy_log2 = s->fy_log2;
    f.iz = s->fz;
    f.steps = take;
    f.exchange = last ? 1 : 0;
    f.species = (int)k;
    f.dcoef = s->fd[k];
    f.kcoef = s->fk[k];
    f.tiles_x = s->f_tiles_x;
    f.tiles_y = s->f_tiles_y;
    return f;
}

ghip_status run_fields(ghip_sim *s) {
    TimedLaunch t{};
    HIP_TRY(s->begin_timed(1, &t));
    const uint32_t S = s->m.n_species;
    int left[kMaxSpecies] = {0, 0, 0, 0};
    for (uint32_t k = 0; k < S; ++k) left[k] = s->substeps[k];
    // planes of at least one 64 x 32 tile: the species whose next chunk has
    // the same length step together in one launch
    const bool planes = field_planes_apply(s->W, s->H, s->d3) && s->fx_log2 == 6 && s->fy_log2 == 5;
    for (;;) {
        int take = 0;
        for (uint32_t k = 0; k < S && take == 0; ++k)
            if (left[k] > 0) take = std::min(left[k], s->f_kmax);
        if (take == 0) break;
        FieldPlanes fp{};
        for (uint32_t k = 0; k < S; ++k) {
            if (left[k] <= 0 || std::min(left[k], s->f_kmax) != take) continue;
            left[k] -= take;
            const FieldArgs f = field_chunk(s, k, take, left[k] == 0);
            if (planes) {
                fp.sp[fp.n++] = f;
            } else {
                HIP_TRY(launch_field(f, s->d3, s->f_blocks, s->stream));
            }
            s->field_cur[k] = 1 - s->field_cur[k];
        }
        if (planes) HIP_TRY(launch_field_planes(fp, s->stream));
    }
    HIP_TRY(s->end_timed(t));
    return GHIP_OK;
}

ghip_status run_cells(ghip_sim *s) {
    // the per-label state is brought current, A[p%2] += D[(p+2)%3]: by the
    // division plan where there is one, else by a consolidation that also
    // clears the step's flags
    const uint64_t p = s->pass;
    if (!s->has_pop)
        HIP_TRY(launch_consolidate(s->A[p % 2], s->Dl[(p + 2) % 3], s->n_labels_dev, s->cw, s->cap,
                                   s->stream, s->flags));
    TimedLaunch t{};
    HIP_TRY(s->begin_timed(3, &t));
    CellArgs c = s->cell_args();
    if (s->has_pop) {
        c.key = pass_key(s->m.seed, s->m.replica, s->mcs, 0xDEADCE11U);
        c.key2 = second_key(c.key);
        // the plan checks for drift; the chain ends with the recentring
        HIP_TRY(launch_population(c, s->Dl[(p + 2) % 3], s->N, s->type_copy, s->anchor_copy, s->stream));
    } else {
        HIP_TRY(launch_recentre_check(c, s->N, s->stream));
    }
    HIP_TRY(s->end_timed(t));
    return GHIP_OK;
}

// Labels one step can add at most: every dividing cell holds at least
// max(4, division volume) sites and divides once, so a step adds at most
// N / that many labels. Zero when no type divides.
uint64_t labels_per_step_bound(const ghip_sim *s) {
    double dv = 0.0;
    for (uint32_t t = 1; t < s->m.n_types; ++t) {
        const double v = s->m.types[t].division_volume;
        if (v > 0.0 && (dv == 0.0 || v < dv)) dv = v;
    }
    if (dv == 0.0) return 0;
    const double least = std::max(4.0, std::ceil(dv));
    return (uint64_t)((double)s->N / least);
}

// Copies `old_bytes` of *p into a new zeroed buffer of `new_bytes`.
hipError_t regrow(void **p, size_t old_bytes, size_t new_bytes) {
    if (!*p) return hipSuccess;
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, new_bytes);
    if (e != hipSuccess) return e;
    e = hipMemset(q, 0, new_bytes);
    if (e == hipSuccess && old_bytes) e = hipMemcpy(q, *p, old_bytes, hipMemcpyDeviceToDevice);
    if (e != hipSuccess) {
        (void)hipFree(q);
        return e;
    }
    (void)hipFree(*p);
    *p = q;
    return hipSuccess;
}

}  // namespace

namespace ghip {

// Makes room for the divisions of the next steps, as the reference grows its
// per-label buffers when daughters outnumber them (glazier-cuda/src/lib.rs:
// 1446-1459). With the stream idle, reads the label count and, when fewer
// than one step's worth of ids is free, moves every per-label buffer to a
// larger capacity (contents kept, the rest zeroed). *chunk receives how many
// of `steps` can run before the next check with no chance of running out.
ghip_status ensure_label_room(ghip_sim *s, uint64_t steps, uint64_t *chunk, bool *grew) {
    *chunk = steps;
    if (grew) *grew = false;
    const uint64_t per_step = labels_per_step_bound(s);
    if (!s->has_pop || per_step == 0 || steps == 0) return GHIP_OK;
    HIP_TRY(hipStreamSynchronize(s->stream));
    uint32_t n = 0;
    HIP_TRY(hipMemcpy(&n, s->n_labels_dev, 4, hipMemcpyDeviceToHost));
    const uint64_t limit = kIdMask - 1;
    if ((uint64_t)s->cap - n < per_step && s->cap < limit) {
        uint64_t want = std::max<uint64_t>(2ull * s->cap, (uint64_t)n + 8 * per_step);
        want = std::min(want, limit);
        const size_t oc = s->cap, nc = (size_t)want, cw = (size_t)s->cw;
        for (int i = 0; i < 2; ++i)
            HIP_TRY(regrow((void **)&s->A[i], oc * cw * 8, nc * cw * 8));
        for (int i = 0; i < 3; ++i)
            HIP_TRY(regrow((void **)&s->Dl[i], oc * cw * 8, nc * cw * 8));
        HIP_TRY(regrow((void **)&s->anchor, oc * 12, nc * 12));
        HIP_TRY(regrow((void **)&s->ltype, oc * 4, nc * 4));
        HIP_TRY(regrow((void **)&s->scratch, oc * 4, nc * 4));
        HIP_TRY(regrow((void **)&s->daughter, oc * 4, nc * 4));
        HIP_TRY(regrow((void **)&s->plan, oc, nc));
        HIP_TRY(regrow((void **)&s->geo, oc * 48, nc * 48));
        HIP_TRY(regrow((void **)&s->len_before, oc * 8, nc * 8));
        HIP_TRY(regrow((void **)&s->type_copy, oc * 4, nc * 4));
        HIP_TRY(regrow((void **)&s->anchor_copy, oc * 12, nc * 12));
        s->cap = (uint32_t)nc;
        s->grown += 1;
        if (grew) *grew = true;
    }
    const uint64_t room = s->cap > n ? (uint64_t)s->cap - n : 0;
    *chunk = std::max<uint64_t>(1, std::min(steps, room / per_step));
    return GHIP_OK;
}

}  // namespace ghip

namespace {

ghip_status read_state(const ghip_sim *s, std::vector<int64_t> &cur, uint32_t &n_labels) {
    HIP_TRY(hipStreamSynchronize(s->stream));
    HIP_TRY(hipMemcpy(&n_labels, s->n_labels_dev, 4, hipMemcpyDeviceToHost));
    const uint64_t p = s->pass;
    size_t words = (size_t)n_labels * s->cw;
    std::vector<int64_t> e(words);
    cur.resize(words);
    if (words) {
        HIP_TRY(hipMemcpy(cur.data(), s->A[p % 2], words * 8, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(e.data(), s->Dl[(p + 2) % 3], words * 8, hipMemcpyDeviceToHost));
    }
    for (size_t i = 0; i < words; ++i) cur[i] += e[i];
    return GHIP_OK;
}

ghip_status read_labels_raw(const ghip_sim *s, std::vector<uint32_t> &out) {
    HIP_TRY(hipStreamSynchronize(s->stream));
    out.resize(s->N);
    HIP_TRY(hipMemcpy(out.data(), s->labels[s->lab_cur], s->N * 4, hipMemcpyDeviceToHost));
    return GHIP_OK;
}

}  // namespace

extern "C" {

const char *ghip_last_error(void) { return g_last_error.c_str(); }

ghip_status ghip_create(const ghip_model *model, const uint32_t *labels, size_t n_sites,
                        const uint8_t *cell_type, size_t n_labels, const double *fields,
                        const double *activity, ghip_sim **out) {
    return ghip_create_nematic(model, labels, n_sites, cell_type, n_labels, fields, activity, nullptr, out);
}

ghip_status ghip_create_nematic(const ghip_model *model, const uint32_t *labels, size_t n_sites,
                                const uint8_t *cell_type, size_t n_labels, const double *fields,
                                const double *activity, const double *nematic, ghip_sim **out) {
    g_last_error.clear();
    if (!model || !labels || !cell_type || !out) return fail(GHIP_BAD_ARG, "a required pointer is null");
    *out = nullptr;
    const ghip_model &m = *model;
    if (m.width == 0 || m.height == 0 || m.depth == 0)
        return fail(GHIP_BAD_ARG, "the lattice has a zero dimension");
    if (m.width % 2 || m.height % 2 || (m.depth > 1 && m.depth % 2)) {
        char buf[160];
        std::snprintf(buf, sizeof buf,
                      "the checkerboard split needs even sides; the lattice is %u by %u by %u",
                      m.width, m.height, m.depth);
        return fail(GHIP_ODD_LATTICE, buf);
    }
    if (m.has_nematic && !nematic)
        return fail(GHIP_BAD_ARG, "the model couples to a nematic field that was not given "
                                  "(ghip_create_nematic takes it)");
    if (m.n_types < 1 || m.n_types > GHIP_MAX_TYPES)
        return fail(GHIP_UNSUPPORTED, "the hip engine takes 1 to 16 cell types");
    if (m.n_species > GHIP_MAX_SPECIES) return fail(GHIP_UNSUPPORTED, "the hip engine takes up to 4 species");
    if (m.neighbour_order < 1 || m.neighbour_order > 3)
        return fail(GHIP_BAD_ARG, "neighbour order is outside the range one to three");
    if (!(m.temperature > 0.0)) return fail(GHIP_BAD_ARG, "the temperature must be positive");
    const size_t N = (size_t)m.width * m.height * m.depth;
    if (n_sites != N) return fail(GHIP_BAD_ARG, "the label field does not match the lattice");
    if (N >= 0xFFFFFFFFull) return fail(GHIP_UNSUPPORTED, "the lattice exceeds 2^32 sites");
    if (n_labels < 1 || n_labels >= (size_t)kIdMask) return fail(GHIP_BAD_ARG, "label count out of range");
    for (size_t l = 1; l < n_labels; ++l)
        if (cell_type[l] < 1 || cell_type[l] >= m.n_types)
            return fail(GHIP_BAD_ARG, "every cell needs a type between 1 and n_types - 1");
    if (cell_type[0] >= m.n_types) return fail(GHIP_BAD_ARG, "the medium's type is out of range");
    for (size_t i = 0; i < N; ++i)
        if (labels[i] >= n_labels) return fail(GHIP_BAD_ARG, "a site holds a label beyond n_labels");

    int dev_count = 0;
    if (hipGetDeviceCount(&dev_count) != hipSuccess || dev_count == 0)
        return fail(GHIP_DEVICE, "no HIP device is available");

    ghip_sim *s = new ghip_sim();
    s->m = m;
    s->W = (int)m.width;
    s->H = (int)m.height;
    s->D = (int)m.depth;
    s->N = N;
    s->d3 = m.depth > 1;
    s->colours = s->d3 ? 8 : 4;
    s->medium_type = cell_type[0];
    // neighbour offsets (glazier-core/src/lattice.rs:69-88)
    {
        static const int o1[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
        static const int o2[18][3] = {{1, 0, 0},  {-1, 0, 0}, {0, 1, 0},   {0, -1, 0},  {0, 0, 1},
                                      {0, 0, -1}, {1, 1, 0},  {1, -1, 0},  {-1, 1, 0},  {-1, -1, 0},
                                      {1, 0, 1},  {1, 0, -1}, {-1, 0, 1},  {-1, 0, -1}, {0, 1, 1},
                                      {0, 1, -1}, {0, -1, 1}, {0, -1, -1}};
        int k = 0;
        if (s->d3 && m.neighbour_order == 1) {
            for (int i = 0; i < 6; ++i, ++k) std::memcpy(s->off[k], o1[i], sizeof o1[i]);
        } else if (s->d3 && m.neighbour_order == 2) {
            for (int i = 0; i < 18; ++i, ++k) std::memcpy(s->off[k], o2[i], sizeof o2[i]);
        } else if (!s->d3 && m.neighbour_order == 1) {
            for (int i = 0; i < 4; ++i, ++k) std::memcpy(s->off[k], o1[i], sizeof o1[i]);
        } else {
            for (int dz = -1; dz <= 1; ++dz)
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dx = -1; dx <= 1; ++dx) {
                        if (!dx && !dy && !dz) continue;
                        if (!s->d3 && dz) continue;
                        s->off[k][0] = dx;
                        s->off[k][1] = dy;
                        s->off[k][2] = dz;
                        ++k;
                    }
        }
        s->n_off = k;
    }
    build_tables(s);
    choose_tiles(s);
    s->cw = s->tracks_moments ? (s->has_nem ? 12 : 10) : 1;
    std::vector<int32_t> nq;
    if (s->has_nem) {
        std::string err;
        if (!nematic_to_fixed(nematic, N, &nq, &err)) {
            delete s;
            return fail(GHIP_BAD_ARG, err);
        }
    }
    uint64_t cap = n_labels;
    if (s->has_pop) cap = std::max<uint64_t>(2 * n_labels + 1024, n_labels + N / 16);
    cap = std::min<uint64_t>(cap, kIdMask - 1);
    s->cap = (uint32_t)cap;
    s->resident = s->plain_physics && m.n_species == 0 && plain2d_resident_fits(s->W, s->H, s->cap);
    s->coop = !s->resident && s->plain_physics && m.n_species == 0 &&
              plain2d_coop_shape(s->W, s->H, s->cap, &s->coop_tx_log2, &s->coop_ty_log2, &s->coop_blocks);
    auto quiet_skip = [&] {
        return s->plain_fast && !s->resident && !s->coop && !s->tracks_moments && !s->has_pop && !s->has_mot &&
               !s->has_nem && m.n_species == 0;
    };
    s->quiet_skip = quiet_skip();
    for (uint32_t k = 0; k < m.n_species; ++k) {
        s->substeps[k] = substeps_for(m, k);
        s->fd[k] = (float)(m.species[k].diffusion / (double)s->substeps[k]);
        s->fk[k] = (float)(m.species[k].decay / (double)s->substeps[k]);
    }

    auto cleanup = [&](ghip_status st) {
        delete s;
        return st;
    };
#define CREATE_TRY(expr)                                                                       \
    do {                                                                                       \
        hipError_t _e = (expr);                                                                \
        if (_e != hipSuccess)                                                                  \
            return cleanup(fail(GHIP_DEVICE, std::string(#expr) + ": " + hipGetErrorString(_e))); \
    } while (0)

    CREATE_TRY(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    {
        int dev = 0;
        CREATE_TRY(hipGetDevice(&dev));
        CREATE_TRY(hipDeviceGetAttribute(&s->cus, hipDeviceAttributeMultiprocessorCount, dev));
    }
    if (s->coop && !plain2d_coop_resident(s->coop_blocks, s->cus)) {
        // the device cannot hold every tile's workgroup at once (fewer CUs
        // than tiles): streaming passes instead
        s->coop = false;
        s->quiet_skip = quiet_skip();
    }
    CREATE_TRY(hipMalloc(&s->labels[0], N * 4));
    if (s->has_mot) CREATE_TRY(hipMalloc(&s->labels[1], N * 4));
    CREATE_TRY(hipMalloc(&s->act[0], N * 4));
    if (s->has_mot) CREATE_TRY(hipMalloc(&s->act[1], N * 4));
    // species-major halves; each species ping-pongs between its two slices
    if (m.n_species) {
        CREATE_TRY(hipMalloc(&s->field_base[0], (size_t)m.n_species * N * 4));
        CREATE_TRY(hipMalloc(&s->field_base[1], (size_t)m.n_species * N * 4));
        for (uint32_t k = 0; k < m.n_species; ++k) {
            s->field[k][0] = s->field_base[0] + k * N;
            s->field[k][1] = s->field_base[1] + k * N;
        }
    }
    const size_t words = (size_t)cap * s->cw;
    for (int i = 0; i < 2; ++i) CREATE_TRY(hipMalloc(&s->A[i], words * 8));
    for (int i = 0; i < 3; ++i) CREATE_TRY(hipMalloc(&s->Dl[i], words * 8));
    CREATE_TRY(hipMalloc(&s->anchor, (size_t)cap * 3 * 4));
    CREATE_TRY(hipMalloc(&s->ltype, (size_t)cap * 4));
    CREATE_TRY(hipMalloc(&s->scratch, (size_t)cap * 4));
    CREATE_TRY(hipMalloc(&s->daughter, (size_t)cap * 4));
    CREATE_TRY(hipMalloc(&s->plan, (size_t)cap));
    CREATE_TRY(hipMalloc(&s->geo, (size_t)cap * 6 * 8));
    if (s->sweep_flags & kLength) CREATE_TRY(hipMalloc(&s->len_before, (size_t)cap * 8));
    if (s->has_nem) {
        CREATE_TRY(hipMalloc(&s->nematic, N * 2 * 4));
        CREATE_TRY(hipMemcpy(s->nematic, nq.data(), N * 2 * 4, hipMemcpyHostToDevice));
    }
    CREATE_TRY(hipMalloc(&s->flags, 32));
    if (s->quiet_skip) {
        // every pass so far counts as having accepted copies (the second
        // record buffer does not hold the first's values yet)
        const uint32_t ones[4] = {1, 1, 1, 1};
        CREATE_TRY(hipMalloc(&s->pass_flags, sizeof ones));
        CREATE_TRY(hipMemcpy(s->pass_flags, ones, sizeof ones, hipMemcpyHostToDevice));
    }
    CREATE_TRY(hipMalloc(&s->type_copy, (size_t)cap * 4));
    CREATE_TRY(hipMalloc(&s->anchor_copy, (size_t)cap * 3 * 4));
    CREATE_TRY(hipMalloc(&s->n_labels_dev, 4));
    CREATE_TRY(hipMalloc(&s->tables, sizeof(DeviceTables)));
    CREATE_TRY(hipMalloc(&s->accepted, kAcceptSlots * 8));
    if (s->coop) {
        CREATE_TRY(hipMalloc(&s->coop_sync, kCoopSyncWords * 4));
        CREATE_TRY(hipMemset(s->coop_sync, 0, kCoopSyncWords * 4));
    }

    // uploads: packed label words, per-label type words, fields, activity
    std::vector<uint32_t> packed(N);
    for (size_t i = 0; i < N; ++i) {
        uint32_t l = labels[i];
        packed[i] = l == 0 ? 0u : (l | ((uint32_t)cell_type[l] << kTypeShift));
    }
    std::vector<uint32_t> lt(cap, 0u);
    for (size_t l = 1; l < n_labels; ++l) lt[l] = (uint32_t)l | ((uint32_t)cell_type[l] << kTypeShift);
    CREATE_TRY(hipMemcpy(s->labels[0], packed.data(), N * 4, hipMemcpyHostToDevice));
    if (s->has_mot) CREATE_TRY(hipMemcpy(s->labels[1], packed.data(), N * 4, hipMemcpyHostToDevice));
    CREATE_TRY(hipMemcpy(s->ltype, lt.data(), (size_t)cap * 4, hipMemcpyHostToDevice));
    {
        std::vector<float> a(N, 0.0f);
        if (activity)
            for (size_t i = 0; i < N; ++i) a[i] = (float)activity[i];
        CREATE_TRY(hipMemcpy(s->act[0], a.data(), N * 4, hipMemcpyHostToDevice));
        if (s->has_mot) CREATE_TRY(hipMemcpy(s->act[1], a.data(), N * 4, hipMemcpyHostToDevice));
    }
    for (uint32_t k = 0; k < m.n_species; ++k) {
        std::vector<float> f(N);
        for (size_t i = 0; i < N; ++i) f[i] = (float)(fields ? fields[k * N + i] : m.species[k].initial);
        CREATE_TRY(hipMemcpy(s->field[k][0], f.data(), N * 4, hipMemcpyHostToDevice));
    }
    uint32_t nl = (uint32_t)n_labels;
    CREATE_TRY(hipMemcpy(s->n_labels_dev, &nl, 4, hipMemcpyHostToDevice));
    CREATE_TRY(hipMemcpy(s->tables, &s->host_tables, sizeof(DeviceTables), hipMemcpyHostToDevice));
    CREATE_TRY(hipMemset(s->accepted, 0, kAcceptSlots * 8));
    CREATE_TRY(hipMemset(s->flags, 0, 32));
    for (int i = 0; i < 3; ++i) CREATE_TRY(hipMemset(s->Dl[i], 0, words * 8));
    CREATE_TRY(hipMemset(s->A[1], 0, words * 8));
    CREATE_TRY(hipMemset(s->anchor, 0, (size_t)cap * 12));
    CREATE_TRY(hipDeviceSynchronize());

    // initial counters and anchors
    CellArgs c = s->cell_args();
    CREATE_TRY(launch_recount(c, N, s->stream));
    if (s->tracks_moments) CREATE_TRY(launch_initial_anchors(c, N, s->stream));
    CREATE_TRY(hipStreamSynchronize(s->stream));
    *out = s;
    return GHIP_OK;
#undef CREATE_TRY
}

ghip_status ghip_destroy(ghip_sim *sim) {
    if (sim && sim->batch) return fail(GHIP_BAD_ARG, "the lattice is in a batch; destroy the batch first");
    delete sim;
    return GHIP_OK;
}

ghip_status ghip_step(ghip_sim *s, uint64_t steps) {
    if (!s) return fail(GHIP_BAD_ARG, "null handle");
    if (s->traj && s->traj_len + steps > s->traj_cap)
        return fail(GHIP_BAD_ARG, "the recording has room for " + std::to_string(s->traj_cap - s->traj_len) +
                                      " more steps");
    const uint64_t pass0 = s->pass;
    if (s->span_timing) {
        if (!s->span_start) HIP_TRY(hipEventCreate(&s->span_start));
        if (!s->span_stop) HIP_TRY(hipEventCreate(&s->span_stop));
        HIP_TRY(hipEventRecord(s->span_start, s->stream));
    }
    const bool coop_run = s->coop && !s->traj && steps > 0;
    if (s->resident && !s->traj) {
        ghip_status st = run_resident(s, steps);
        if (st != GHIP_OK) return st;
        steps = 0;
    } else if (coop_run) {
        ghip_status st = run_coop(s, steps);
        if (st != GHIP_OK) return st;
        steps = 0;
    }
    uint64_t chunk = 0;
    for (uint64_t q = 0; q < steps; ++q) {
        if (chunk == 0) {
            ghip_status st = ensure_label_room(s, steps - q, &chunk, nullptr);
            if (st != GHIP_OK) return st;
        }
        --chunk;
        int order[8];
        colour_order(s->m.seed, s->m.replica, s->mcs, s->colours, order);
        for (int c = 0; c < s->colours; ++c) {
            ghip_status st = run_sweep_pass(s, order[c]);
            if (st != GHIP_OK) return st;
        }
        if (s->traj) {
            // the owners this step's field exchange reads
            HIP_TRY(launch_record_types(s->labels[s->lab_cur], s->traj + s->traj_len * s->N, s->N, s->stream));
            s->traj_len += 1;
        }
        if (s->m.n_species) {
            ghip_status st = run_fields(s);
            if (st != GHIP_OK) return st;
        }
        if (s->has_mot) HIP_TRY(launch_activity_decay(s->act[s->act_cur], s->N, s->stream));
        if (s->tracks_moments) {
            ghip_status st = run_cells(s);
            if (st != GHIP_OK) return st;
        }
        s->mcs += 1;
        s->attempted += s->N;
    }
    if (s->span_timing) HIP_TRY(hipEventRecord(s->span_stop, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    if (s->timing) HIP_TRY(s->collect_timing());
    if (s->span_timing) {
        float elapsed = 0.0f;
        HIP_TRY(hipEventElapsedTime(&elapsed, s->span_start, s->span_stop));
        s->ms[kStepFamily] += elapsed;
        s->launches[kStepFamily] += s->pass - pass0;
    }
    if (coop_run) {
        uint32_t bad = 0;
        HIP_TRY(hipMemcpy(&bad, s->coop_sync + kCoopErrWord, 4, hipMemcpyDeviceToHost));
        if (bad) {
            HIP_TRY(hipMemset(s->coop_sync + kCoopErrWord, 0, 4));
            return fail(GHIP_DEVICE, "a grid barrier of the resident pass gave up waiting; the lattice is not valid");
        }
    }
    if (s->has_pop) {
        // word 2 of either set outlives the per-step clears: set by any
        // step of the call
        uint32_t fl[8];
        HIP_TRY(hipMemcpy(fl, s->flags, 32, hipMemcpyDeviceToHost));
        if (fl[2] || fl[6]) {
            HIP_TRY(hipMemset(s->flags + 2, 0, 4));
            HIP_TRY(hipMemset(s->flags + 6, 0, 4));
            return fail(GHIP_DEVICE, "the label capacity ran out at the id limit; divisions were dropped");
        }
    }
    return GHIP_OK;
}

uint64_t ghip_mcs(const ghip_sim *s) { return s ? s->mcs : 0; }

ghip_status ghip_n_labels(const ghip_sim *s, size_t *out) {
    if (!s || !out) return fail(GHIP_BAD_ARG, "null argument");
    HIP_TRY(hipStreamSynchronize(s->stream));
    uint32_t n = 0;
    HIP_TRY(hipMemcpy(&n, s->n_labels_dev, 4, hipMemcpyDeviceToHost));
    *out = n;
    return GHIP_OK;
}

ghip_status ghip_labels(const ghip_sim *s, uint32_t *out, size_t n) {
    if (!s || !out) return fail(GHIP_BAD_ARG, "null argument");
    if (n != s->N) return fail(GHIP_BAD_ARG, "the buffer does not match the lattice");
    std::vector<uint32_t> raw;
    ghip_status st = read_labels_raw(s, raw);
    if (st != GHIP_OK) return st;
    for (size_t i = 0; i < n; ++i) out[i] = label_id(raw[i]);
    return GHIP_OK;
}

ghip_status ghip_volumes(const ghip_sim *s, uint32_t *out, size_t n) {
    if (!s || !out) return fail(GHIP_BAD_ARG, "null argument");
    std::vector<int64_t> cur;
    uint32_t nl = 0;
    ghip_status st = read_state(s, cur, nl);
    if (st != GHIP_OK) return st;
    if (n != nl) return fail(GHIP_BAD_ARG, "the buffer does not match the label count");
    uint64_t cells = 0;
    for (size_t l = 1; l < nl; ++l) {
        int64_t v = unpack_vol(cur[l * s->cw]);
        out[l] = (uint32_t)(v > 0 ? v : 0);
        cells += out[l];
    }
    out[0] = (uint32_t)(s->N - cells);
    return GHIP_OK;
}

ghip_status ghip_surfaces(const ghip_sim *s, int64_t *out, size_t n) {
    if (!s || !out) return fail(GHIP_BAD_ARG, "null argument");
    std::vector<int64_t> cur;
    uint32_t nl = 0;
    ghip_status st = read_state(s, cur, nl);
    if (st != GHIP_OK) return st;
    if (n != nl) return fail(GHIP_BAD_ARG, "the buffer does not match the label count");
    for (size_t l = 1; l < nl; ++l) out[l] = unpack_surf(cur[l * s->cw]);
    // the medium is not tracked during a pass: count it from the lattice
    std::vector<uint32_t> raw;
    st = read_labels_raw(s, raw);
    if (st != GHIP_OK) return st;
    int64_t s0 = 0;
    const size_t plane = (size_t)s->W * s->H;
    for (size_t i = 0; i < s->N; ++i) {
        if (raw[i] != 0) continue;
        int z = (int)(i / plane), y = (int)((i % plane) / s->W), x = (int)(i % s->W);
        for (int k = 0; k < s->n_off; ++k) {
            int nx = (x + s->off[k][0] + s->W) % s->W, ny = (y + s->off[k][1] + s->H) % s->H;
            int nz = (z + s->off[k][2] + s->D) % s->D;
            size_t nb = ((size_t)nz * s->H + ny) * s->W + nx;
            if (nb != i && raw[nb] != 0) ++s0;
        }
    }
    out[0] = s0;
    return GHIP_OK;
}

ghip_status ghip_cell_types(const ghip_sim *s, uint8_t *out, size_t n) {
    if (!s || !out) return fail(GHIP_BAD_ARG, "null argument");
    size_t nl = 0;
    ghip_status st = ghip_n_labels(s, &nl);
    if (st != GHIP_OK) return st;
    if (n != nl) return fail(GHIP_BAD_ARG, "the buffer does not match the label count");
    std::vector<uint32_t> lt(nl);
    HIP_TRY(hipMemcpy(lt.data(), s->ltype, nl * 4, hipMemcpyDeviceToHost));
    out[0] = (uint8_t)s->medium_type;
    for (size_t l = 1; l < nl; ++l) out[l] = (uint8_t)label_type(lt[l]);
    return GHIP_OK;
}

ghip_status ghip_field_f32(const ghip_sim *s, uint32_t species, float *out, size_t n) {
    if (!s || !out) return fail(GHIP_BAD_ARG, "null argument");
    if (species >= s->m.n_species) return fail(GHIP_BAD_ARG, "no such species");
    if (n != s->N) return fail(GHIP_BAD_ARG, "the buffer does not match the lattice");
    HIP_TRY(hipStreamSynchronize(s->stream));
    HIP_TRY(hipMemcpy(out, s->field[species][s->field_cur[species]], n * 4, hipMemcpyDeviceToHost));
    return GHIP_OK;
}

ghip_status ghip_field(const ghip_sim *s, uint32_t species, double *out, size_t n) {
    std::vector<float> f(n);
    ghip_status st = ghip_field_f32(s, species, f.data(), n);
    if (st != GHIP_OK) return st;
    for (size_t i = 0; i < n; ++i) out[i] = (double)f[i];
    return GHIP_OK;
}

ghip_status ghip_blueprint_parse(const char *text, char *out, size_t cap, size_t *needed) {
    if (!text || !needed) return fail(GHIP_BAD_ARG, "null argument");
    Blueprint bp;
    std::string err;
    if (!blueprint_from_json(text, &bp, &err)) return fail(GHIP_BAD_ARG, err);
    const std::string json = check_json(bp, true);
    *needed = json.size() + 1;
    if (!out) return GHIP_OK;
    if (cap < json.size() + 1) return fail(GHIP_BAD_ARG, "the buffer is shorter than the description");
    std::memcpy(out, json.c_str(), json.size() + 1);
    return GHIP_OK;
}

ghip_status ghip_set_nematic_field(ghip_sim *s, const double *q, size_t n) {
    if (!s || !q) return fail(GHIP_BAD_ARG, "null argument");
    if (n != s->N)
        return fail(GHIP_BAD_ARG, "nematic field has " + std::to_string(n) + " sites, the lattice " +
                                      std::to_string(s->N));
    if (!s->has_nem) return fail(GHIP_BAD_ARG, "the model has no nematic coupling to update");
    std::vector<int32_t> nq;
    std::string err;
    if (!nematic_to_fixed(q, n, &nq, &err)) return fail(GHIP_BAD_ARG, err);
    HIP_TRY(hipStreamSynchronize(s->stream));
    HIP_TRY(hipMemcpy(s->nematic, nq.data(), n * 2 * 4, hipMemcpyHostToDevice));
    // every label's field sum is recounted: the record is brought current
    // (A += the last pass's delta), then its moment words are resummed
    const uint64_t p = s->pass;
    HIP_TRY(launch_consolidate(s->A[p % 2], s->Dl[(p + 2) % 3], s->n_labels_dev, s->cw, s->cap, s->stream));
    HIP_TRY(launch_recount_moments(s->cell_args(), s->N, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return GHIP_OK;
}

ghip_status ghip_set_field(ghip_sim *s, uint32_t species, const double *values, size_t n) {
    if (!s || !values) return fail(GHIP_BAD_ARG, "null argument");
    if (species >= s->m.n_species) return fail(GHIP_BAD_ARG, "no such species");
    if (n != s->N) return fail(GHIP_BAD_ARG, "the field does not match the lattice");
    std::vector<float> f(n);
    for (size_t i = 0; i < n; ++i) f[i] = (float)values[i];
    HIP_TRY(hipStreamSynchronize(s->stream));
    HIP_TRY(hipMemcpy(s->field[species][s->field_cur[species]], f.data(), n * 4, hipMemcpyHostToDevice));
    return GHIP_OK;
}

ghip_status ghip_activity_f32(const ghip_sim *s, float *out, size_t n) {
    if (!s || !out) return fail(GHIP_BAD_ARG, "null argument");
    if (n != s->N) return fail(GHIP_BAD_ARG, "the buffer does not match the lattice");
    HIP_TRY(hipStreamSynchronize(s->stream));
    HIP_TRY(hipMemcpy(out, s->act[s->act_cur], n * 4, hipMemcpyDeviceToHost));
    return GHIP_OK;
}

ghip_status ghip_counters(const ghip_sim *s, uint64_t *attempted, uint64_t *accepted) {
    if (!s || !attempted || !accepted) return fail(GHIP_BAD_ARG, "null argument");
    HIP_TRY(hipStreamSynchronize(s->stream));
    unsigned long long slots[kAcceptSlots];
    HIP_TRY(hipMemcpy(slots, s->accepted, sizeof slots, hipMemcpyDeviceToHost));
    uint64_t total = 0;
    for (int i = 0; i < kAcceptSlots; ++i) total += slots[i];
    *attempted = s->attempted;
    *accepted = total;
    return GHIP_OK;
}

ghip_status ghip_cell_state(const ghip_sim *s, int64_t *out, size_t n, uint32_t *words_per_label) {
    if (!s || !words_per_label) return fail(GHIP_BAD_ARG, "null argument");
    // volume and surface, then every further record word (moments, nematic sums)
    const uint32_t words = (uint32_t)s->cw + 1;
    *words_per_label = words;
    if (!out) return GHIP_OK;
    std::vector<int64_t> cur;
    uint32_t nl = 0;
    ghip_status st = read_state(s, cur, nl);
    if (st != GHIP_OK) return st;
    if (n != (size_t)nl * words) return fail(GHIP_BAD_ARG, "the buffer does not match the label count");
    for (size_t l = 0; l < nl; ++l) {
        const int64_t *c = cur.data() + l * s->cw;
        out[words * l] = unpack_vol(c[0]);
        out[words * l + 1] = unpack_surf(c[0]);
        for (uint32_t q = 2; q < words; ++q) out[words * l + q] = c[q - 1];
    }
    return GHIP_OK;
}

ghip_status ghip_set_barrier_patience(ghip_sim *s, uint32_t polls) {
    if (!s) return fail(GHIP_BAD_ARG, "null handle");
    s->coop_patience = polls;
    return GHIP_OK;
}

ghip_status ghip_set_timing(ghip_sim *s, int on) {
    if (!s) return fail(GHIP_BAD_ARG, "null handle");
    if (on < 0 || on > 2) return fail(GHIP_BAD_ARG, "timing mode is 0, 1 or 2");
    s->timing = on == 1;
    s->span_timing = on == 2;
    return GHIP_OK;
}

ghip_status ghip_kernel_time(const ghip_sim *s, const char *family, double *total_ms, uint64_t *launches) {
    if (!s || !family || !total_ms || !launches) return fail(GHIP_BAD_ARG, "null argument");
    for (int i = 0; i < 5; ++i)
        if (std::strcmp(family, kFamilies[i]) == 0) {
            *total_ms = s->ms[i];
            *launches = s->launches[i];
            return GHIP_OK;
        }
    return fail(GHIP_BAD_ARG, "unknown kernel family");
}

}  // extern "C"
