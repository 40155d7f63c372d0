// sweep_body.h - one colour pass of the checkerboard Metropolis sweep on gfx950.
//
// Device counterpart of cpm_sweep (glazier-cuda/src/lib.rs:436-716) and of
// the serial attempt (glazier-core/src/cpu.rs:432-517), laid out for CDNA4:
//
//  * a workgroup owns a tile of 2TX x 2TY (x 2TZ) sites; it stages the tile's
//    labels plus a Moore halo in LDS with all lanes' loads in flight, so every target
//    reads its neighbourhood from LDS instead of strided global gathers;
//  * one lane per target draws it: TX targets along x per row, so a 64-lane
//    wavefront covers one target row of a 128-site tile row; the targets
//    whose label would change are queued in LDS, and the gathers, pricing
//    and commit run on the queue's dense prefix of lanes, one target each
//    (in a sheet about one target in five is queued, so most wavefronts
//    skip the pricing whole instead of running it for a few lanes);
//  * each target takes ONE 32-bit counter hash (pick from the low bits,
//    uniform from the top 24), so no RNG state is read or written;
//  * per-label volume/surface/moments are read from the pass-start snapshot
//    a_cur + e_prev and accepted copies add exact int64 deltas into f_acc,
//    so the pass is deterministic whatever the wavefront order;
//  * the same launch folds the snapshot for the next pass
//    (a_next = a_cur + e_prev) and clears the delta buffer the pass after
//    next accumulates into, so a pass is one launch with no separate copy.
//
// The pass is HBM/Infinity-Cache bound: it must read every label once
// (4 B per site per colour) and write 4 B per accepted copy.
#pragma once

#include "ghip_device.h"
#include "geometry.h"

namespace ghip {

namespace {

__device__ __forceinline__ int wrap_coord(int v, int span) {
    if (v < 0) v += span;
    if (v >= span) v -= span;
    if (v >= span) v %= span;
    return v;
}

// Moore ring of the connectivity test (connectivity.rs:23-29): ORDER3's order,
// dz then dy then dx from -1 to 1, dropping the z offsets in a plane.
__device__ __forceinline__ void ring_offset(int k, bool d3, int *dx, int *dy, int *dz) {
    int q = k;
    if (d3) {
        q = k + (k >= 13 ? 1 : 0);
        *dz = q / 9 - 1;
        q = q % 9;
    } else {
        q = k + (k >= 4 ? 1 : 0);
        *dz = 0;
    }
    *dy = q / 3 - 1;
    *dx = q % 3 - 1;
}

// Whether the ring positions set in `present` form one piece: the planar
// ring by the host's 256-bit table, the volume's by cube_one_piece
// (ghip_common.h).
__device__ __forceinline__ bool one_piece(uint32_t present, bool d3, const DeviceTables *t) {
    if (present == 0u || (present & (present - 1u)) == 0u) return true;
    if (!d3) return (t->piece2d[present >> 5] >> (present & 31u)) & 1u;
    return cube_one_piece(present);
}

}  // namespace

// FULL adds connectivity, length, chemotaxis, motility and drift behind
// runtime flags; the plain variant prices contact, volume and surface only.
// MOORE2D fixes the eight planar Moore offsets at compile time.
// One workgroup's share of a colour pass: block `blk` of the pass's launch
// over this lattice (it picks the tile, the fold range and the counter slot).
template <bool D3, bool FULL, bool MOORE2D>
__device__ __forceinline__ void sweep_body(const SweepArgs &a, const int blk) {
    extern __shared__ uint32_t lds[];
    const int tid = threadIdx.x;
    const int nthreads = blockDim.x;
    const DeviceTables *tab = a.tables;

    // ---- block -> tile: each XCD on a contiguous run of tiles, walked
    // backwards on every other pass
    const int b = pass_tile(blk, a.xcd_chunk, a.reverse, a.tiles_x * a.tiles_y * (D3 ? a.tiles_z : 1));
    const int tile_x = b % a.tiles_x;
    const int tile_y = (b / a.tiles_x) % a.tiles_y;
    const int tile_z = D3 ? b / (a.tiles_x * a.tiles_y) : 0;
    const int TX = 1 << a.tx_log2, TY = 1 << a.ty_log2, TZ = D3 ? 1 << a.tz_log2 : 1;
    const int halo = a.halo;
    const int LX = 2 * TX + 2 * halo, LY = 2 * TY + 2 * halo, LZ = D3 ? 2 * TZ + 2 * halo : 1;
    const int x0 = tile_x * 2 * TX, y0 = tile_y * 2 * TY, z0 = D3 ? tile_z * 2 * TZ : 0;

    // ---- LDS: contact table, per-type volume/surface, acceptance table,
    // neighbour offsets, then the label tile (and the activity tile).
    const int T = (int)a.n_types;
    float *sJ = reinterpret_cast<float *>(lds);
    float *sP = sJ + T * T;                       // 4 per type
    uint32_t *sThr = reinterpret_cast<uint32_t *>(sP + 4 * T);
    int *sOff = reinterpret_cast<int *>(sThr + a.n_thr); // [0,32) LDS deltas, [32,64) packed xyz
    uint32_t *sLab = reinterpret_cast<uint32_t *>(sOff + 64);
    float *sAct = reinterpret_cast<float *>(sLab + LX * LY * LZ);
    // the queue of targets whose label would change: a count, then lane ids
    uint32_t *sQ = reinterpret_cast<uint32_t *>(sAct) + ((a.flags & kMotility) ? LX * LY * LZ : 0);

    // ---- every load of the prologue goes out before the first wait on one:
    // the labels (and activity) of the tile with its halo, as one flat range
    // over all lanes, up to kStage loads per lane; then this block's share of
    // the snapshot fold (two words per lane now, stored at the end of the
    // pass); then the tables bound for LDS. A wavefront's wait for a load is
    // a wait for every load it issued before, so the chain is one round trip
    // where loading the tables, folding and staging in turn made it three.
    const uint32_t *src_lab = (FULL && (a.flags & kMotility)) ? a.labels_prev : a.labels;
    const bool with_act = FULL && (a.flags & kMotility);
    const int staged = LX * LY * LZ;
    constexpr int kStage = D3 ? 4 : 8;
    auto stage_load = [&](int base, uint32_t (&lv)[kStage], float (&av)[kStage]) {
#pragma unroll
        for (int u = 0; u < kStage; ++u) {
            const int e = base + u * nthreads + tid;
            if (e < staged) {
                const int r = e / LX, c = e - r * LX;
                const int ly = D3 ? r % LY : r, lz = D3 ? r / LY : 0;
                const int gxx = wrap_coord(x0 - halo + c, a.w);
                const int gyy = wrap_coord(y0 - halo + ly, a.h);
                const int gzz = D3 ? wrap_coord(z0 - halo + lz, a.d) : 0;
                const size_t g = ((size_t)gzz * a.h + gyy) * (size_t)a.w + gxx;
                lv[u] = src_lab[g];
                if (with_act) av[u] = a.act_prev[g];
            }
        }
    };
    auto stage_store = [&](int base, const uint32_t (&lv)[kStage], const float (&av)[kStage]) {
#pragma unroll
        for (int u = 0; u < kStage; ++u) {
            const int e = base + u * nthreads + tid;
            if (e < staged) {
                sLab[e] = lv[u];
                if (with_act) sAct[e] = av[u];
            }
        }
    };
    uint32_t lv0[kStage];
    float av0[kStage];
    stage_load(0, lv0, av0);

    // the fold of this block's share of the per-label snapshot for the next
    // pass (a_next = a_cur + e_prev), and the clearing of the delta buffer
    // two passes ahead: nothing in the pass reads either buffer it writes
    const uint64_t words = (uint64_t)(*a.n_labels) * (uint64_t)a.cell_words;
    const uint64_t f_lo = (uint64_t)blk * a.fold_per_block;
    const uint64_t f_hi = f_lo + a.fold_per_block < words ? f_lo + a.fold_per_block : words;
    const uint64_t f0 = f_lo + tid, f1 = f0 + nthreads;
    int64_t fa0 = 0, fe0 = 0, fa1 = 0, fe1 = 0;
    if (f0 < f_hi) {
        fa0 = a.a_cur[f0];
        fe0 = a.e_prev[f0];
    }
    if (f1 < f_hi) {
        fa1 = a.a_cur[f1];
        fe1 = a.e_prev[f1];
    }

    for (int i = tid; i < T * T; i += nthreads) sJ[i] = tab->J[i / T][i % T];
    for (int i = tid; i < T; i += nthreads) {
        sP[4 * i + 0] = tab->lam_v[i];
        sP[4 * i + 1] = tab->tgt_v[i];
        sP[4 * i + 2] = tab->lam_s[i];
        sP[4 * i + 3] = tab->tgt_s[i];
    }
    for (int i = tid; i < (int)a.n_thr; i += nthreads) sThr[i] = tab->thr[i];
    if (tid < a.n_off) {
        sOff[tid] = a.off[tid][0] + a.off[tid][1] * LX + a.off[tid][2] * LX * LY;
        sOff[32 + tid] = (a.off[tid][0] + 1) | ((a.off[tid][1] + 1) << 2) | ((a.off[tid][2] + 1) << 4);
    }
    if (tid == 0) sQ[0] = 0;
    stage_store(0, lv0, av0);
    // a tile larger than kStage loads per lane stages the rest in turn
    for (int base = kStage * nthreads; base < staged; base += kStage * nthreads) {
        uint32_t lv[kStage];
        float av[kStage];
        stage_load(base, lv, av);
        stage_store(base, lv, av);
    }
    const int wave = tid >> 6, lane = tid & 63, nwaves = nthreads >> 6;
    __syncthreads();

    // ---- target t of the tile (lane t's in the drawing): its coordinates,
    // its site in the tile, its draws and the picked source
    struct Target {
        int gx, gy, gz, tl, sl, pdx, pdy, pdz;
        uint32_t site, r24;
        bool active;
    };
    auto target = [&](int t) {
        Target g;
        const int i = t & (TX - 1);
        const int j = (t >> a.tx_log2) & (TY - 1);
        const int k = D3 ? (t >> (a.tx_log2 + a.ty_log2)) : 0;
        g.gx = x0 + 2 * i + a.cx;
        g.gy = y0 + 2 * j + a.cy;
        g.gz = D3 ? z0 + 2 * k + a.cz : 0;
        g.active = t < TX * TY * TZ && g.gx < a.w && g.gy < a.h && (!D3 || g.gz < a.d);
        const int lx = halo + 2 * i + a.cx, ly = halo + 2 * j + a.cy, lz = D3 ? halo + 2 * k + a.cz : 0;
        g.tl = (lz * LY + ly) * LX + lx;
        g.site = (uint32_t)(((size_t)g.gz * a.h + g.gy) * (size_t)a.w + g.gx);
        const uint32_t ph = site_pick(a.key, g.site);
        const uint32_t h = site_draw(a.key2, ph);
        int pick;
        if (MOORE2D) {
            pick = (int)(ph & 7u);
            g.r24 = h >> 8;
            int q = pick + (pick >= 4 ? 1 : 0);
            g.pdy = q / 3 - 1;
            g.pdx = q - 3 * (g.pdy + 1) - 1;
            g.pdz = 0;
        } else {
            const int n = a.n_off;
            if ((n & (n - 1)) == 0) {
                pick = (int)(ph & (uint32_t)(n - 1));
                g.r24 = h >> 8;
            } else {
                pick = (int)(((uint64_t)h * (uint64_t)n) >> 32);
                g.r24 = second_uniform(h);
            }
            const int packed = sOff[32 + pick];
            g.pdx = (packed & 3) - 1;
            g.pdy = ((packed >> 2) & 3) - 1;
            g.pdz = ((packed >> 4) & 3) - 1;
        }
        g.sl = g.tl + g.pdx + g.pdy * LX + g.pdz * LX * LY;
        return g;
    };

    // ---- every lane draws its own target and queues it when its label
    // would change; the pricing below then runs on a dense prefix of the
    // lanes (in a sheet about one target in five is queued), so a wavefront
    // of queued targets replaces several mostly idle ones. What a target
    // decides and commits does not depend on the lane that prices it.
    {
        const Target g = target(tid);
        const bool cand = g.active && sLab[g.tl] != sLab[g.sl];
        const unsigned long long m = __ballot(cand);
        uint32_t base = 0;
        if (lane == 0 && m) base = atomicAdd(&sQ[0], (uint32_t)__builtin_popcountll(m));
        base = __builtin_amdgcn_readfirstlane(base);
        if (cand)
            sQ[1 + base + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u))] =
                (uint32_t)tid;
    }
    __syncthreads();
    const uint32_t nq = sQ[0];

    bool accepted = false;
    int tl_keep = 0;
    uint32_t nw_keep = 0;
    float act_keep = -1.0f;
    if (tid < (int)nq) {
        const Target g = target((int)sQ[1 + tid]);
        const int gx = g.gx, gy = g.gy, gz = g.gz;
        const int tl = g.tl, sl = g.sl, pdx = g.pdx, pdy = g.pdy, pdz = g.pdz;
        const uint32_t site = g.site, r24 = g.r24;
        const uint32_t old = sLab[tl];
        const uint32_t nw = sLab[sl];
        const uint32_t oid = label_id(old), nid = label_id(nw);
        const uint32_t eto = label_type(old), etn = label_type(nw);
        const int cw = a.cell_words;
        // every gather the decision can need goes out before the first wait
        // on one: both records; then, in a plane, for a target whose label
        // would change, the chemotactic fields at target and source with
        // their coefficients, and the moment words, anchors and table entries
        // the length term reads. (A volume's terms, which would hold too many
        // registers, and the nematic term gather once the records have decided
        // the copy is priced.)
        int64_t ra_o = 0, re_o = 0, ra_n = 0, re_n = 0;
        if (oid != 0) {
            ra_o = a.a_cur[(size_t)oid * cw];
            re_o = a.e_prev[(size_t)oid * cw];
        }
        if (nid != 0) {
            ra_n = a.a_cur[(size_t)nid * cw];
            re_n = a.e_prev[(size_t)nid * cw];
        }
        const bool cand = old != nw;
        int sx = 0, sy = 0, sz = 0;
        if (FULL) {
            sx = wrap_coord(gx + pdx, a.w);
            sy = wrap_coord(gy + pdy, a.h);
            sz = D3 ? wrap_coord(gz + pdz, a.d) : 0;
        }
        float chi_p[kMaxSpecies], ft_p[kMaxSpecies], fs_p[kMaxSpecies];
        const bool chem = FULL && !D3 && (a.flags & kChemotaxis) && cand && nid != 0;
        if (chem) {
            const size_t ssite = ((size_t)sz * a.h + sy) * (size_t)a.w + sx;
#pragma unroll
            for (int q = 0; q < kMaxSpecies; ++q) {
                if (q < a.n_species) {
                    chi_p[q] = tab->chi[etn][q];
                    ft_p[q] = a.fields[q][site];
                    fs_p[q] = a.fields[q][ssite];
                }
            }
        }
        // planar length: words sx, sy, sxx, syy, sxy (moments 0, 1, 3, 4, 6)
        constexpr int kPlanarWords[5] = {0, 1, 3, 4, 6};
        int64_t lm_a[2][5], lm_e[2][5];
        int32_t anc_p[2][2];
        double lam_p[2], tgt_p[2];
        const bool plen = FULL && !D3 && (a.flags & kLength) && cand;
        if (plen) {
#pragma unroll
            for (int side = 0; side < 2; ++side) {
                const uint32_t id = side == 0 ? oid : nid;
                const uint32_t ty = side == 0 ? eto : etn;
                lam_p[side] = tab->lam_len[ty];
                tgt_p[side] = tab->tgt_len[ty];
                if (id == 0) continue;
                const size_t base = (size_t)id * cw + 1;
#pragma unroll
                for (int q = 0; q < 5; ++q) {
                    lm_a[side][q] = a.a_cur[base + kPlanarWords[q]];
                    lm_e[side][q] = a.e_prev[base + kPlanarWords[q]];
                }
                anc_p[side][0] = a.anchor[3 * id];
                anc_p[side][1] = a.anchor[3 * id + 1];
            }
        }
        const int64_t pv_old = ra_o + re_o, pv_new = ra_n + re_n;
        bool go = cand && !(oid != 0 && unpack_vol(pv_old) <= 1);

        if (FULL && go && oid != 0 && (a.flags & kConnected) && tab->connected[label_type(old)]) {
            uint32_t present = 0u;
            const int rn = D3 ? 26 : 8;
            for (int q = 0; q < rn; ++q) {
                int dx, dy, dz;
                ring_offset(q, D3, &dx, &dy, &dz);
                if (sLab[tl + dx + dy * LX + dz * LX * LY] == old) present |= 1u << q;
            }
            go = one_piece(present, D3, tab);
        }

        if (go) {
            float cn = 0.0f, co = 0.0f;
            int like_old = 0, like_new = 0;
            const float *rowN = sJ + etn * T;
            const float *rowO = sJ + eto * T;
            if (MOORE2D) {
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    int qq = q + (q >= 4 ? 1 : 0);
                    int dy = qq / 3 - 1, dx = qq % 3 - 1;
                    uint32_t ln = sLab[tl + dx + dy * LX];
                    uint32_t tn = label_type(ln);
                    if (ln != nw) cn = cn + rowN[tn]; else ++like_new;
                    if (ln != old) co = co + rowO[tn]; else ++like_old;
                }
            } else {
                for (int q = 0; q < a.n_off; ++q) {
                    uint32_t ln = sLab[tl + sOff[q]];
                    uint32_t tn = label_type(ln);
                    if (ln != nw) cn = cn + rowN[tn]; else ++like_new;
                    if (ln != old) co = co + rowO[tn]; else ++like_old;
                }
            }
            const int bonds = MOORE2D ? 8 : a.n_off;
            float dh = cn - co;
            const int ds_old = 2 * like_old - bonds;
            const int ds_new = bonds - 2 * like_new;
            if (oid != 0) {
                float lv = sP[4 * eto + 0];
                if (lv != 0.0f) {
                    float v = (float)unpack_vol(pv_old);
                    dh = dh + lv * (1.0f - 2.0f * (v - sP[4 * eto + 1]));
                }
            }
            if (nid != 0) {
                float lv = sP[4 * etn + 0];
                if (lv != 0.0f) {
                    float v = (float)unpack_vol(pv_new);
                    dh = dh + lv * (1.0f + 2.0f * (v - sP[4 * etn + 1]));
                }
            }
            if (oid != 0 && ds_old != 0) {
                float ls = sP[4 * eto + 2];
                if (ls != 0.0f) {
                    float s = (float)unpack_surf(pv_old);
                    float dsf = (float)ds_old;
                    dh = dh + ls * (dsf * (2.0f * (s - sP[4 * eto + 3]) + dsf));
                }
            }
            if (nid != 0 && ds_new != 0) {
                float ls = sP[4 * etn + 2];
                if (ls != 0.0f) {
                    float s = (float)unpack_surf(pv_new);
                    float dsf = (float)ds_new;
                    dh = dh + ls * (dsf * (2.0f * (s - sP[4 * etn + 3]) + dsf));
                }
            }
            if (FULL && !D3 && (a.flags & kLength)) {
                for (int side = 0; side < 2; ++side) {
                    const uint32_t id = side == 0 ? oid : nid;
                    if (id == 0 || lam_p[side] == 0.0) continue;
                    const int sign = side == 0 ? -1 : 1;
                    int64_t m[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, after[9];
#pragma unroll
                    for (int q = 0; q < 5; ++q) m[kPlanarWords[q]] = lm_a[side][q] + lm_e[side][q];
                    const int64_t n = unpack_vol(side == 0 ? pv_old : pv_new);
                    const int64_t ux = fold_int(gx - anc_p[side][0], a.w);
                    const int64_t uy = fold_int(gy - anc_p[side][1], a.h);
                    after[0] = m[0] + sign * ux;
                    after[1] = m[1] + sign * uy;
                    after[3] = m[3] + sign * ux * ux;
                    after[4] = m[4] + sign * uy * uy;
                    after[6] = m[6] + sign * ux * uy;
                    after[2] = after[5] = after[7] = after[8] = 0;
                    // the label's length in the pass's snapshot, from the
                    // words just gathered (the closed form of a plane)
                    const double before = length_of(m, n, false) - tgt_p[side];
                    const double now = length_of(after, n + sign, false) - tgt_p[side];
                    dh = dh + (float)(lam_p[side] * (now * now - before * before));
                }
            }
            if (FULL && D3 && (a.flags & kLength)) {
                const bool three = D3;
                for (int side = 0; side < 2; ++side) {
                    const uint32_t id = side == 0 ? oid : nid;
                    const uint32_t ty = side == 0 ? eto : etn;
                    if (id == 0 || tab->lam_len[ty] == 0.0) continue;
                    const int sign = side == 0 ? -1 : 1;
                    const size_t base = (size_t)id * cw;
                    int64_t m[9], after[9];
                    for (int q = 0; q < 9; ++q) m[q] = a.a_cur[base + 1 + q] + a.e_prev[base + 1 + q];
                    const int64_t n = unpack_vol(side == 0 ? pv_old : pv_new);
                    const int64_t ux = fold_int(gx - a.anchor[3 * id], a.w);
                    const int64_t uy = fold_int(gy - a.anchor[3 * id + 1], a.h);
                    const int64_t uz = fold_int(gz - a.anchor[3 * id + 2], a.d);
                    after[0] = m[0] + sign * ux;
                    after[1] = m[1] + sign * uy;
                    after[2] = m[2] + sign * uz;
                    after[3] = m[3] + sign * ux * ux;
                    after[4] = m[4] + sign * uy * uy;
                    after[5] = m[5] + sign * uz * uz;
                    after[6] = m[6] + sign * ux * uy;
                    after[7] = m[7] + sign * ux * uz;
                    after[8] = m[8] + sign * uy * uz;
                    // the label's length in the pass's snapshot, an iterative
                    // eigensolver in a volume, once per label by the prepass
                    double before = a.len_before[id];
                    double now = length_of(after, n + sign, three) - tab->tgt_len[ty];
                    dh = dh + (float)(tab->lam_len[ty] * (now * now - before * before));
                }
            }
            if (FULL && (a.flags & kNematic)) {
                // -lambda (F' . a' - F . a) (glazier-core/src/cpu.rs:315-332): F the
                // label's field sum (words 10, 11), a its anisotropy, before and
                // after the site moves
                const int64_t q0 = a.nematic[2 * site], q1 = a.nematic[2 * site + 1];
                for (int side = 0; side < 2; ++side) {
                    const uint32_t id = side == 0 ? oid : nid;
                    const uint32_t ty = side == 0 ? eto : etn;
                    if (id == 0 || tab->lam_nem[ty] == 0.0) continue;
                    const int sign = side == 0 ? -1 : 1;
                    const size_t base = (size_t)id * cw;
                    int64_t m[9], after[9];
                    for (int q = 0; q < 9; ++q) m[q] = a.a_cur[base + 1 + q] + a.e_prev[base + 1 + q];
                    const int64_t f0 = a.a_cur[base + 10] + a.e_prev[base + 10];
                    const int64_t f1 = a.a_cur[base + 11] + a.e_prev[base + 11];
                    const int64_t n = unpack_vol(side == 0 ? pv_old : pv_new);
                    const int64_t ux = fold_int(gx - a.anchor[3 * id], a.w);
                    const int64_t uy = fold_int(gy - a.anchor[3 * id + 1], a.h);
                    const int64_t uz = fold_int(gz - a.anchor[3 * id + 2], a.d);
                    after[0] = m[0] + sign * ux;
                    after[1] = m[1] + sign * uy;
                    after[2] = m[2] + sign * uz;
                    after[3] = m[3] + sign * ux * ux;
                    after[4] = m[4] + sign * uy * uy;
                    after[5] = m[5] + sign * uz * uz;
                    after[6] = m[6] + sign * ux * uy;
                    after[7] = m[7] + sign * ux * uz;
                    after[8] = m[8] + sign * uy * uz;
                    double a0, a1, b0, b1;
                    anisotropy(m, n, &a0, &a1);
                    anisotropy(after, n + sign, &b0, &b1);
                    const double before = (double)f0 * a0 + (double)f1 * a1;
                    const double now = (double)(f0 + sign * q0) * b0 + (double)(f1 + sign * q1) * b1;
                    dh = dh + (float)(-tab->lam_nem[ty] * ((now - before) * kNematicUnit));
                }
            }
            if (chem) {
#pragma unroll
                for (int q = 0; q < kMaxSpecies; ++q) {
                    if (q >= a.n_species || chi_p[q] == 0.0f) continue;
                    dh = dh - chi_p[q] * (ft_p[q] - fs_p[q]);
                }
            }
            if (FULL && D3 && (a.flags & kChemotaxis) && nid != 0) {
                const size_t ssite = ((size_t)sz * a.h + sy) * (size_t)a.w + sx;
                for (int q = 0; q < a.n_species; ++q) {
                    const float chi = tab->chi[etn][q];
                    if (chi == 0.0f) continue;
                    dh = dh - chi * (a.fields[q][site] - a.fields[q][ssite]);
                }
            }
            if (FULL && (a.flags & kMotility) && nid != 0 && tab->coef_act[etn] != 0.0f &&
                tab->max_act[etn] > 0.0f) {
                // geometric mean of the activity over the source's label
                // around the source (side 0) and the target's around the
                // target (side 1); a non-positive value zeroes a side. Both
                // sides walk the neighbourhood together, so each step's four
                // LDS reads are in flight at once; the logs add in neighbour
                // order, and a zeroed side's sum is never read.
                const float here0 = sAct[sl], here1 = sAct[tl];
                bool zero0 = !(here0 > 0.0f), zero1 = !(here1 > 0.0f);
                float lsum0 = zero0 ? 0.0f : log2_f32(here0);
                float lsum1 = zero1 ? 0.0f : log2_f32(here1);
                float count0 = 1.0f, count1 = 1.0f;
                for (int q = 0; q < a.n_off; ++q) {
                    const int o = sOff[q];
                    const uint32_t l0 = sLab[sl + o], l1 = sLab[tl + o];
                    const float v0 = sAct[sl + o], v1 = sAct[tl + o];
                    if (l0 == nw) {
                        if (v0 > 0.0f) {
                            lsum0 = lsum0 + log2_f32(v0);
                            count0 = count0 + 1.0f;
                        } else {
                            zero0 = true;
                        }
                    }
                    if (l1 == old) {
                        if (v1 > 0.0f) {
                            lsum1 = lsum1 + log2_f32(v1);
                            count1 = count1 + 1.0f;
                        } else {
                            zero1 = true;
                        }
                    }
                }
                const float gm0 = zero0 ? 0.0f : exp2_f32(lsum0 / count0);
                const float gm1 = zero1 ? 0.0f : exp2_f32(lsum1 / count1);
                dh = dh - tab->coef_act[etn] * (gm0 - gm1);
            }
            if (FULL && (a.flags & kDrift) && nid != 0) {
                float wk = tab->ext[etn][0] * (float)(-pdx) + tab->ext[etn][1] * (float)(-pdy) +
                           tab->ext[etn][2] * (float)(-pdz);
                dh = dh - wk;
            }

            bool acc;
            if (dh <= 0.0f) {
                acc = true;
            } else if (a.flags & kIntegral) {
                acc = dh < (float)a.n_thr && r24 < sThr[(int)dh];
            } else {
                float p = exp2_f32(-dh * tab->accept_c);
                acc = (float)r24 < p * 16777216.0f;
            }

            if (acc) {
                accepted = true;
                if (!(FULL && (a.flags & kMotility))) a.labels[site] = nw;
                if (oid != 0)
                    atomicAdd(reinterpret_cast<unsigned long long *>(a.f_acc + (size_t)oid * cw),
                              (unsigned long long)pack_vs(-1, ds_old));
                if (nid != 0)
                    atomicAdd(reinterpret_cast<unsigned long long *>(a.f_acc + (size_t)nid * cw),
                              (unsigned long long)pack_vs(1, ds_new));
                if (FULL && (a.flags & kMoments)) {
                    for (int side = 0; side < 2; ++side) {
                        const uint32_t id = side == 0 ? oid : nid;
                        if (id == 0) continue;
                        const int64_t sg = side == 0 ? -1 : 1;
                        const int64_t ux = fold_int(gx - a.anchor[3 * id], a.w);
                        const int64_t uy = fold_int(gy - a.anchor[3 * id + 1], a.h);
                        const int64_t uz = fold_int(gz - a.anchor[3 * id + 2], a.d);
                        const int64_t dm[9] = {sg * ux,      sg * uy,      sg * uz,
                                               sg * ux * ux, sg * uy * uy, sg * uz * uz,
                                               sg * ux * uy, sg * ux * uz, sg * uy * uz};
                        unsigned long long *dst =
                            reinterpret_cast<unsigned long long *>(a.f_acc + (size_t)id * cw + 1);
                        for (int q = 0; q < 9; ++q)
                            if (dm[q] != 0) atomicAdd(dst + q, (unsigned long long)dm[q]);
                        if (a.flags & kNematic) {
                            const int64_t d0 = sg * (int64_t)a.nematic[2 * site];
                            const int64_t d1 = sg * (int64_t)a.nematic[2 * site + 1];
                            if (d0 != 0) atomicAdd(dst + 9, (unsigned long long)d0);
                            if (d1 != 0) atomicAdd(dst + 10, (unsigned long long)d1);
                        }
                    }
                }
                if (FULL && (a.flags & kMotility)) {
                    // written through the tile once every lane has priced
                    tl_keep = tl;
                    nw_keep = nw;
                    if (nid != 0 && tab->max_act[etn] > 0.0f) act_keep = tab->max_act[etn];
                }
            }
        }
    }

    // ---- the fold's stores, then the rest of its share
    if (f0 < f_hi) {
        a.a_next[f0] = fa0 + fe0;
        a.g_zero[f0] = 0;
    }
    if (f1 < f_hi) {
        a.a_next[f1] = fa1 + fe1;
        a.g_zero[f1] = 0;
    }
    for (uint64_t i = f1 + nthreads; i < f_hi; i += nthreads) {
        a.a_next[i] = a.a_cur[i] + a.e_prev[i];
        a.g_zero[i] = 0;
    }

    // ---- acceptance count: one atomic per wavefront into a spread slot.
    const unsigned long long ball = __ballot(accepted);
    if (lane == 0 && ball)
        atomicAdd(a.accepted + (blk & (kAcceptSlots - 1)),
                  (unsigned long long)__builtin_popcountll(ball));

    // ---- with motility the pass reads a pass-start copy and writes the
    // whole tile interior into the other buffer, so that buffer is complete.
    if (FULL && (a.flags & kMotility)) {
        __syncthreads();
        if (accepted) {
            sLab[tl_keep] = nw_keep;
            if (act_keep >= 0.0f) sAct[tl_keep] = act_keep;
        }
        __syncthreads();
        const int IX = 2 * TX, IY = 2 * TY, IZ = D3 ? 2 * TZ : 1;
        for (int r = wave; r < IY * IZ; r += nwaves) {
            int iy = r % IY, iz = r / IY;
            int gyy = y0 + iy, gzz = z0 + iz;
            if (gyy >= a.h || (D3 && gzz >= a.d)) continue;
            size_t row = ((size_t)gzz * a.h + gyy) * (size_t)a.w;
            int lrow = ((D3 ? iz + halo : 0) * LY + iy + halo) * LX + halo;
            for (int c = lane; c < IX; c += 64) {
                if (x0 + c >= a.w) continue;
                a.labels[row + x0 + c] = sLab[lrow + c];
                a.act[row + x0 + c] = sAct[lrow + c];
            }
        }
    }
}

}  // namespace ghip
