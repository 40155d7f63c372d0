// sweep_plain.hip - the colour pass specialised for the plain monolayer: a
// plane, the eight Moore neighbours, the medium plus one cell type, and
// integer energies (contact, volume, surface). This is the throughput path.
//
// Same draws, same decisions and the same bookkeeping as sweep_kernel (and so
// as the CPU emulation); only the work is organised for gfx950:
//
//  * a workgroup owns a PX x PY tile (256 x 32 sites with 512 lanes above
//    512^2 where the lattice divides, 128 x 8 up to 256^2, else 128 x 16,
//    both with 256 lanes), the tiles of each XCD in one contiguous run,
//    walked backwards on every other pass so a pass starts on the lines the
//    previous one left in that XCD's L2. Every global read of the
//    workgroup (halo columns, acceptance table, the PY + 2 tile rows as
//    16-byte loads, its share of the snapshot fold, the pass flags) is
//    issued before the first wait, branch-free and in the order it is used.
//    Rows land in an LDS tile of pitch PX + 4 words, the interior at column
//    4 (so each 16-byte LDS write is aligned), the left halo at column 3 and
//    the right halo one past the interior, which is column 0 of the next
//    row. At 256 x 32 that keeps the workgroup under 40 KiB of LDS, four per
//    CU, and 2048 tiles of a 4096^2 lattice fill the chip in two rounds;
//  * phase A draws every target (one hash round: the pick). A wavefront's 64
//    targets share one tile row; the target's label and its picked
//    neighbour come from LDS, and the targets whose label would change are
//    appended to an LDS queue of 16-bit target indices (one reservation per
//    wavefront). The targets that read only the tile's first rows are drawn
//    while the last rows still arrive. In a confluent sheet most targets
//    have old == new and cost nothing more;
//  * phase B prices the queued targets densely: it redraws the pick round
//    (cheaper than holding it in LDS, which would cost residency), issues
//    the per-label gathers together, takes the contact sum from neighbour
//    counts (exact with integer energies) and the uniform's hash round only
//    when the acceptance table decides; the medium count is skipped for
//    tiles that hold no medium. On lattices of several rounds a tile merges
//    its accepted copies' record changes per label in LDS before the global
//    atomics;
//  * a pass whose two predecessors accepted nothing anywhere (the pass
//    flags) reads no deltas and stores no fold.
//
// Larger tiles carry more loads per workgroup lifetime and amortise the
// fixed per-workgroup work; they matter most once copies are accepted.
#include "ghip_device.h"

namespace ghip {

constexpr int kPThr = 256;  // acceptance-table entries held in LDS

namespace {
// Count of neighbours differing from v: each compare's bit added as a
// carry (a compare and an add-with-carry per neighbour).
__device__ __forceinline__ uint32_t unequal8(uint32_t v, uint32_t n0, uint32_t n1, uint32_t n2,
                                             uint32_t n3, uint32_t n4, uint32_t n5, uint32_t n6,
                                             uint32_t n7) {
    uint32_t c = (uint32_t)(n0 != v);
    c += (uint32_t)(n1 != v);
    c += (uint32_t)(n2 != v);
    c += (uint32_t)(n3 != v);
    c += (uint32_t)(n4 != v);
    c += (uint32_t)(n5 != v);
    c += (uint32_t)(n6 != v);
    c += (uint32_t)(n7 != v);
    return c;
}

// The decision on one target whose label would change from `old` to `nw`:
// contact gained minus lost from neighbour counts (exact with integer
// energies), volume and surface against each label's pass-start record
// (volume * 2^32 + surface), then the integral acceptance table. The
// target's two hash rounds (the pick round of `site`, then the uniform's)
// are taken only when the table decides, 0 < dH < its length. Shared by the
// streaming pass and the resident one, so both decide identically.
struct PlainMove {
    bool acc;
    int ds_old, ds_new;
};

// The neighbourhood's part of a decision: neighbours unlike the old label,
// unlike the new one, and (when the medium can change a price) medium ones.
struct PlainCounts {
    int unlike_old, unlike_new, med;
};

__device__ __forceinline__ PlainCounts plain_counts(uint32_t old, uint32_t nw, uint32_t n0, uint32_t n1,
                                                    uint32_t n2, uint32_t n3, uint32_t n4, uint32_t n5,
                                                    uint32_t n6, uint32_t n7, bool has_medium) {
    PlainCounts c;
    c.unlike_old = (int)unequal8(old, n0, n1, n2, n3, n4, n5, n6, n7);
    c.unlike_new = (int)unequal8(nw, n0, n1, n2, n3, n4, n5, n6, n7);
    c.med = has_medium ? 8 - (int)unequal8(0u, n0, n1, n2, n3, n4, n5, n6, n7) : 0;
    return c;
}

// `site_of()` yields the target's site; it is asked only when the acceptance
// table decides.
template <class SiteOf>
__device__ __forceinline__ PlainMove plain_move_counts(const PlainParams &p, const uint32_t *thr_s, uint32_t key,
                                                       uint32_t key2, SiteOf site_of, uint32_t old, uint32_t nw,
                                                       const PlainCounts &c, int64_t pv_old, int64_t pv_new) {
    const uint32_t oid = label_id(old), nid = label_id(nw);
    const int unlike_old = c.unlike_old, unlike_new = c.unlike_new;
    const int like_old = 8 - unlike_old, like_new = 8 - unlike_new;
    // absolute states keep surface in [0, 2^31): volume is the high word
    const int vol_old = (int)(uint32_t)((uint64_t)pv_old >> 32);
    const int vol_new = (int)(uint32_t)((uint64_t)pv_new >> 32);
    // a copy that would empty a cell is refused (no branch before the
    // records are used, so no gather waits behind another)
    const bool emptying = oid != 0 && vol_old <= 1;
    const int etn = nid != 0 ? 1 : 0, eto = oid != 0 ? 1 : 0;
    const int med = c.med;
    // neighbours unlike the new label: 8 - like_new; the medium ones among
    // them are all medium neighbours when the new label is a cell, none
    // when it is the medium (then like_new counts them). So `med` enters
    // only beside a cell label, where flat contact rows cancel it.
    const int med_ne_new = nid != 0 ? med : 0;
    const int cell_ne_new = unlike_new - med_ne_new;
    const int med_ne_old = oid != 0 ? med : 0;
    const int cell_ne_old = unlike_old - med_ne_old;
    const float cn = p.j[etn][0] * (float)med_ne_new + p.j[etn][1] * (float)cell_ne_new;
    const float co = p.j[eto][0] * (float)med_ne_old + p.j[eto][1] * (float)cell_ne_old;
    float dh = cn - co;
    PlainMove mv;
    mv.ds_old = 2 * like_old - 8;
    mv.ds_new = 8 - 2 * like_new;
    if (oid != 0 && p.lam_v != 0.0f) {
        const float vv = (float)vol_old;
        dh = dh + p.lam_v * (1.0f - 2.0f * (vv - p.tgt_v));
    }
    if (nid != 0 && p.lam_v != 0.0f) {
        const float vv = (float)vol_new;
        dh = dh + p.lam_v * (1.0f + 2.0f * (vv - p.tgt_v));
    }
    if (oid != 0 && mv.ds_old != 0 && p.lam_s != 0.0f) {
        const float s = (float)(int32_t)(uint32_t)pv_old;
        const float dsf = (float)mv.ds_old;
        dh = dh + p.lam_s * (dsf * (2.0f * (s - p.tgt_s) + dsf));
    }
    if (nid != 0 && mv.ds_new != 0 && p.lam_s != 0.0f) {
        const float s = (float)(int32_t)(uint32_t)pv_new;
        const float dsf = (float)mv.ds_new;
        dh = dh + p.lam_s * (dsf * (2.0f * (s - p.tgt_s) + dsf));
    }
    bool acc = dh <= 0.0f;
    if (!acc && dh < (float)p.n_thr) acc = (site_draw(key2, site_pick(key, site_of())) >> 8) < thr_s[(int)dh];
    mv.acc = !emptying && acc;
    return mv;
}

__device__ __forceinline__ PlainMove plain_move(const PlainParams &p, const uint32_t *thr_s, uint32_t key,
                                                uint32_t key2, uint32_t site, uint32_t old, uint32_t nw,
                                                uint32_t n0, uint32_t n1, uint32_t n2, uint32_t n3, uint32_t n4,
                                                uint32_t n5, uint32_t n6, uint32_t n7, int64_t pv_old,
                                                int64_t pv_new, bool has_medium) {
    const PlainCounts c = plain_counts(old, nw, n0, n1, n2, n3, n4, n5, n6, n7, has_medium);
    return plain_move_counts(p, thr_s, key, key2, [=] { return site; }, old, nw, c, pv_old, pv_new);
}
}  // namespace

// ---- one tile's colour pass, shared by the streaming and the pipelined
// kernels. The tile is a PX x PY block of sites at (x0, y0), landed in LDS
// with one halo row above and below and one halo column either side: row r
// of the lattice block (r = -1 .. PY) at word (r + 1) * (PX + 4) + 4.
template <int PX, int PY>
struct PlainTile {
    static constexpr int kPitch = PX + 4;
    static constexpr int kRows = PY + 2;
    static constexpr int kWords = kRows * kPitch + 4;
    static constexpr int kTargets = PX * PY / 4;
    static constexpr int kTRow = PX / 2;  // targets per target row
    static_assert(kTRow % 64 == 0, "a wavefront's targets share one row");
    // a queue entry is the target's index in the tile with its pick (the
    // neighbour, 0..7) above it, so pricing takes no second hash round
    static constexpr int kQBits = kTargets == 2048 ? 11 : kTargets == 1024 ? 10 : kTargets == 512 ? 9 : 8;
    static_assert((1 << kQBits) == kTargets && kQBits + 3 <= 16, "16-bit queue entries");

    const uint32_t *tile;
    uint32_t site_base;  // site of the tile's target 0: (y0 + cy) w + x0 + cx
    int tl_base;         // its LDS word: (1 + cy) kPitch + 4 + cx
    uint32_t w2;         // two lattice rows, below 2^24 (the tile choice ensures it)

    // target q of the tile lies on target row j = q / kTRow (uniform in a
    // wavefront), column i = q % kTRow: its LDS word, and its site, which
    // pricing asks for only when it needs it. Row steps are 24-bit
    // multiplies (j < 64).
    __device__ __forceinline__ int target_tl(uint32_t q) const {
        const uint32_t j = q / kTRow, i = q % kTRow;
        return tl_base + (int)(j * (2u * kPitch) + 2u * i);
    }
    __device__ __forceinline__ uint32_t target_site(uint32_t q) const {
        const uint32_t j = q / kTRow, i = q % kTRow;
        return site_base + __umul24(j, w2) + 2u * i;
    }
    // the neighbour pick k (0..7) names, in ORDER3's planar order, (dx, dy)
    // row by row; dx + 1 and dy + 1 are the k-th 2-bit fields of 0x9224 and
    // 0xA940
    __device__ __forceinline__ uint32_t picked(uint32_t k, int tl) const {
        const uint32_t sh = k * 2u;
        const int dxy = (int)(((0xA940u >> sh) & 3u) * (uint32_t)kPitch + ((0x9224u >> sh) & 3u));
        return tile[tl - kPitch - 1 + dxy];
    }

    // Phase A draws every target and queues the ones whose label would
    // change. A lane's targets are q = k NT + tid; bit k of `moves` says
    // target k would change, and bits 3k..3k+2 of `picks` hold its pick.
    // Target k of the lane lies on tile row 2 j + cy, j = (k NT + tid) /
    // kTRow, and reads the LDS rows either side of it; NT is a multiple of
    // kTRow, so the lane's column is fixed and its rows step by NT / kTRow
    // from one k to the next.
    template <int NT, int K0, int K1>
    __device__ __forceinline__ void draw(const SweepArgs &a, uint32_t &moves, uint32_t &picks) const {
        static_assert(NT % kTRow == 0, "a lane keeps its column");
        constexpr int kJStep = NT / kTRow;
        const uint32_t site0 = target_site(threadIdx.x);
        const int tl0 = target_tl(threadIdx.x);
        const uint32_t dsite = (uint32_t)kJStep * w2;
#pragma unroll
        for (int k = K0; k < K1; ++k) {
            const uint32_t site = site0 + (uint32_t)k * dsite;
            const int tl = tl0 + k * 2 * kJStep * kPitch;
            const uint32_t pk = site_pick8(a.key, site);
            moves |= (tile[tl] != picked(pk, tl) ? 1u : 0u) << k;
            picks |= pk << (3 * k);
        }
    }

    // The queue entries of the targets `moves` marks, each with its pick:
    // one reservation per wavefront, its entries ordered by the lane's
    // target k (one target row) and then by lane, so the pricing wavefront
    // that takes 64 consecutive entries writes its accepted labels into one
    // or two target rows.
    template <int NT>
    __device__ __forceinline__ void queue_moves(uint32_t moves, uint32_t picks, uint16_t *queue,
                                                uint32_t *qcount) const {
        constexpr int kPerLane = kTargets / NT;
        static_assert(kTargets % NT == 0 && kPerLane <= 8, "a lane's targets fit the move byte");
        const int tid = threadIdx.x;
        unsigned long long m[kPerLane];
        uint32_t total = 0;
#pragma unroll
        for (int k = 0; k < kPerLane; ++k) {
            m[k] = __ballot((moves >> k) & 1u);
            total += (uint32_t)__builtin_popcountll(m[k]);
        }
        if (total == 0) return;
        auto below = [](unsigned long long x) {
            return __builtin_amdgcn_mbcnt_hi((uint32_t)(x >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)x, 0u));
        };
        uint32_t base = 0;
        if ((tid & 63) == 0) base = atomicAdd(qcount, total);
        base = __builtin_amdgcn_readfirstlane(base);
#pragma unroll
        for (int k = 0; k < kPerLane; ++k) {
            if (moves & (1u << k))
                queue[base + below(m[k])] = (uint16_t)((k * NT + tid) | (int)((picks >> (3 * k)) & 7u) << kQBits);
            base += (uint32_t)__builtin_popcountll(m[k]);
        }
    }

    // Phase B: price the nq queued targets, write the accepted copies'
    // labels, and return this lane's accepted count. Each entry carries its
    // pick, the target's neighbourhood is read from LDS first and the
    // per-label gathers go out together behind it. Records are one word per
    // label here (the plain pass keeps no moments), so a label's record is
    // at its id; with e_zero (the previous pass accepted nothing anywhere,
    // a uniform fact taken outside the loop so the loop body has no branch
    // before its decision) the previous pass's deltas are all zero and not
    // read. The record changes of a lane's first entry (queue entry tid)
    // leave in d_old and d_new as id | (ds + 8) / 2 << 28 (the surface
    // change ds is even, -8..8; the volume changes by -1 and +1), or kEmpty,
    // for merge_deltas when the launch merges; those of later entries (a
    // tile with more than NT queued targets), and all of them when it does
    // not, go straight to the records.
    static constexpr uint32_t kEmpty = 0xFFFFFFFFu;

    template <int NT, bool e_zero>
    __device__ __forceinline__ uint32_t phase_b(const SweepArgs &a, const PlainParams &p, const uint32_t *thr_s,
                                                const uint16_t *queue, uint32_t nq, bool has_medium,
                                                uint32_t &d_old, uint32_t &d_new) const {
        uint32_t my_acc = 0;
        d_old = d_new = kEmpty;
        for (uint32_t qi = threadIdx.x; qi < nq; qi += NT) {
            const uint32_t entry = queue[qi];
            const uint32_t q = entry & (uint32_t)(kTargets - 1);
            const int tl = target_tl(q);
            const uint32_t old = tile[tl];
            const uint32_t nw = picked(entry >> kQBits, tl);
            const uint32_t oid = label_id(old), nid = label_id(nw);
            const uint32_t n0 = tile[tl - kPitch - 1], n1 = tile[tl - kPitch], n2 = tile[tl - kPitch + 1];
            const uint32_t n3 = tile[tl - 1], n4 = tile[tl + 1];
            const uint32_t n5 = tile[tl + kPitch - 1], n6 = tile[tl + kPitch], n7 = tile[tl + kPitch + 1];
            // label 0's record is real memory whose value is never used
            const int64_t ao = a.a_cur[oid], an = a.a_cur[nid];
            int64_t eo = 0, en = 0;
            if (!e_zero) {
                eo = a.e_prev[oid];
                en = a.e_prev[nid];
            }
            const PlainCounts c = plain_counts(old, nw, n0, n1, n2, n3, n4, n5, n6, n7, has_medium);
            const PlainMove mv = plain_move_counts(
                p, thr_s, a.key, a.key2, [&] { return target_site(q); }, old, nw, c, ao + eo, an + en);
            if (!mv.acc) continue;
            a.labels[target_site(q)] = nw;
            if (a.merge_deltas && qi < (uint32_t)NT) {
                if (oid != 0) d_old = oid | (uint32_t)((mv.ds_old + 8) >> 1) << kTypeShift;
                if (nid != 0) d_new = nid | (uint32_t)((mv.ds_new + 8) >> 1) << kTypeShift;
            } else {
                if (oid != 0)
                    atomicAdd(reinterpret_cast<unsigned long long *>(a.f_acc + oid),
                              (unsigned long long)pack_vs(-1, mv.ds_old));
                if (nid != 0)
                    atomicAdd(reinterpret_cast<unsigned long long *>(a.f_acc + nid),
                              (unsigned long long)pack_vs(1, mv.ds_new));
            }
            ++my_acc;
        }
        return my_acc;
    }

    // The record changes of the tile's accepted copies, merged per label
    // before they leave the workgroup: once every queue entry has been read,
    // the queue's LDS holds a small open-addressed table of (label, int64
    // delta); the copies add into it with LDS atomics, and then each label
    // the tile touched costs one global atomic instead of one per copy (a
    // label that finds no slot in a few probes goes straight to its record).
    // The table keeps the 16 records of one 128-byte line in 16 adjacent
    // slots (a line hashes to a group of 16, a taken slot probes the next
    // group), so each wavefront's atomics land on a few lines: a global
    // atomic instruction costs per line it touches, and a tile's labels are
    // mostly consecutive ids. Integer sums in any order: the records are
    // exactly those of one atomic per copy. Called by the whole workgroup.
    static constexpr int kSlots = kTargets >= 2048 ? 256 : kTargets >= 1024 ? 64 : 32;
    static_assert(kSlots * 12 <= kTargets * 2, "the table fits in the queue's bytes");

    template <int NT>
    __device__ __forceinline__ void merge_deltas(const SweepArgs &a, uint16_t *queue, uint32_t d_old,
                                                 uint32_t d_new) const {
        const int tid = threadIdx.x;
        unsigned long long *vals = reinterpret_cast<unsigned long long *>(queue);
        uint32_t *keys = reinterpret_cast<uint32_t *>(vals + kSlots);
        for (int i = tid; i < kSlots; i += NT) {
            keys[i] = kEmpty;
            vals[i] = 0ull;
        }
        __syncthreads();
        auto add = [&](uint32_t packed, int64_t dv) {
            const uint32_t id = label_id(packed);
            const int64_t d = pack_vs(dv, 2 * (int64_t)(packed >> kTypeShift) - 8);
            uint32_t h = ((((id >> 4) * 0x9E3779B1u) >> 24) & ~15u) | (id & 15u);
#pragma unroll
            for (int probe = 0; probe < 8; ++probe) {
                h &= (uint32_t)(kSlots - 1);
                const uint32_t k = atomicCAS(&keys[h], kEmpty, id);
                if (k == kEmpty || k == id) {
                    atomicAdd(&vals[h], (unsigned long long)d);
                    return;
                }
                h += 16;
            }
            atomicAdd(reinterpret_cast<unsigned long long *>(a.f_acc + id), (unsigned long long)d);
        };
        if (d_old != kEmpty) add(d_old, -1);
        if (d_new != kEmpty) add(d_new, 1);
        __syncthreads();
        for (int i = tid; i < kSlots; i += NT) {
            const uint32_t id = keys[i];
            if (id != kEmpty) atomicAdd(reinterpret_cast<unsigned long long *>(a.f_acc + id), vals[i]);
        }
    }
};

template <int PX, int PY, int NT>
__global__ __launch_bounds__(NT) void sweep_plain2d_kernel(SweepArgs a, PlainParams p) {
    using T = PlainTile<PX, PY>;
    constexpr int kPitch = T::kPitch;
    constexpr int kRows = T::kRows;
    constexpr int kC4 = PX / 4;                          // 16-byte loads per row
    constexpr int kLoads = (kRows * kC4 + NT - 1) / NT;  // per lane

    __shared__ __attribute__((aligned(16))) uint32_t tile[T::kWords];
    __shared__ __attribute__((aligned(16))) uint16_t queue[T::kTargets];  // then the delta table
    __shared__ uint32_t thr_s[kPThr];
    __shared__ uint32_t qcount;
    __shared__ uint32_t acc_block;

    const int tid = threadIdx.x;
    const uint32_t b = (uint32_t)pass_tile(blockIdx.x, a.xcd_chunk, a.reverse, a.tiles_x * a.tiles_y);
    const uint32_t tile_y = a.tiles_x == 1 ? b : __umulhi(b, p.tx_magic);
    const uint32_t tile_x = b - tile_y * (uint32_t)a.tiles_x;
    const int x0 = (int)tile_x * PX, y0 = (int)tile_y * PY;
    const int W = a.w, H = a.h;
    // row offsets of the tile's halo rows (each wraps at the lattice's edge)
    // and of the row above the tile (wrapping in unsigned arithmetic when y0
    // is 0; every row it is used for lies below it)
    const uint32_t off_top = (uint32_t)(y0 == 0 ? H - 1 : y0 - 1) * (uint32_t)W;
    const uint32_t off_bottom = (uint32_t)(y0 + PY >= H ? 0 : y0 + PY) * (uint32_t)W;
    const uint32_t off_above = (uint32_t)(y0 - 1) * (uint32_t)W;

    if (tid == 0) {
        qcount = 0;
        acc_block = 0;
    }

    // Passes p - 1 and p - 2 accepted nothing anywhere: the previous pass's
    // deltas are all zero (e_zero), and the fold is the identity as well
    // (fold_idle): the record buffer it would write already holds the
    // snapshot, and the delta buffer it would clear is already clear. Pass p
    // clears the word of pass p + 1, which no pass p reads.
    if (a.pass_flags && blockIdx.x == 0 && tid == 0) a.pass_flags[(a.pass_slot + 1) & 3u] = 0;

    // ---- issue every global read of the workgroup before the first wait, in
    // the order they are used: the halo words, the acceptance table and the
    // pass flags, the tile's rows, the fold's words. Every lane issues every
    // load, with no branch (a lane with nothing to load reads a valid word
    // and ignores it), so each wait covers exactly the loads it needs: a
    // wavefront's wait for a load is a wait for every load issued before it.
    // the fold's extent: from the host when the label count is fixed, else the
    // device count (read only then, so no load ahead of the rows waits on it)
    uint32_t words = a.fold_words;
    if (words == 0) words = *a.n_labels;
    const bool h_live = tid < kRows * 2;
    uint32_t hv;
    {
        // lane 2 r + side reads row r - 1 of the tile, the column left of it
        // (side 0) or right of it (side 1)
        const uint32_t r = h_live ? (uint32_t)tid >> 1 : 0u;
        const uint32_t x_left = (uint32_t)(x0 == 0 ? W - 1 : x0 - 1);
        const uint32_t x_right = (uint32_t)(x0 + PX >= W ? 0 : x0 + PX);
        const uint32_t gx = (tid & 1) ? x_right : x_left;
        const uint32_t off = r == 0 ? off_top : r >= (uint32_t)kRows - 1 ? off_bottom
                                                                         : off_above + __umul24(r, (uint32_t)W);
        hv = a.labels[off + gx];
    }
    // a lane with no halo word stores its load into a spare word past the
    // tile, so the load has an unconditional use and stays where it is
    // issued. Word kRows * kPitch is the last row's right halo; words +1 and
    // +2 are spare.
    constexpr int kSpare = kRows * kPitch + 1;
    const int hpos = h_live ? (tid >> 1) * kPitch + ((tid & 1) ? 4 + PX : 3) : kSpare + (tid & 1);
    const uint32_t tv = p.thr[tid < (int)p.n_thr ? tid : 0];
    // the pass flags of passes p - 1 and p - 2 (read first of the uniform
    // words: the decisions on the fold and the deltas wait on them)
    const uint32_t *flags = a.pass_flags ? a.pass_flags : reinterpret_cast<const uint32_t *>(a.accepted);
    const uint32_t f1 = flags[(a.pass_slot + 3) & 3u], f2 = flags[(a.pass_slot + 2) & 3u];
    __builtin_amdgcn_sched_barrier(1);
    // load k of a lane reads tile row r = k kRowsPerLoad + tid / kC4 (row 0
    // is the halo row above, kRows - 1 the one below, each wrapping at the
    // lattice's edge; rows past the tile reread the last one), column 4 (tid
    // % kC4): one multiply per lane, the rest steps of a row offset
    constexpr int kRowsPerLoad = NT / kC4;
    static_assert(NT % kC4 == 0, "a load instruction covers whole rows");
    const int c4 = tid % kC4, rl = tid / kC4;
    const uint32_t *col = a.labels + (uint32_t)x0 + 4u * (uint32_t)c4;
    const uint32_t off_lane = off_above + __umul24((uint32_t)rl, (uint32_t)W);
    const uint32_t row_step = (uint32_t)kRowsPerLoad * (uint32_t)W;
    uint4 v[kLoads];
    int vdst[kLoads];
#pragma unroll
    for (int k = 0; k < kLoads; ++k) {
        const int r = k * kRowsPerLoad + rl;
        const uint32_t off = r == 0 ? off_top : r >= kRows - 1 ? off_bottom : off_lane + (uint32_t)k * row_step;
        v[k] = *reinterpret_cast<const uint4 *>(col + off);
        vdst[k] = r < kRows ? r * kPitch + 4 + 4 * c4 : -1;
    }
    // the global loads stay in this order (the compiler may not move one
    // past these barriers): a later load scheduled ahead of a row would make
    // that row's wait cover it
    __builtin_amdgcn_sched_barrier(1);
    const uint32_t flo = blockIdx.x * a.fold_per_block;
    const uint32_t fhi = min(flo + a.fold_per_block, words);
    const uint32_t fi = flo + tid;
    const uint32_t fc = fi < fhi ? fi : 0u;
    const int64_t fa = a.a_cur[fc], fe = a.e_prev[fc];

    // ---- land them in two parts. The first kL1 loads hold the tile's first
    // rows; the lane's first kA1 targets read no row beyond them, so they are
    // drawn while the rest of the tile is still arriving.
    constexpr int kL1 = (kLoads + 1) / 2;
    constexpr int kR1 = kL1 * kRowsPerLoad < kRows ? kL1 * kRowsPerLoad : kRows;  // rows in the first part
    constexpr int kTRowsPerDraw = NT / T::kTRow;  // target rows one draw of every lane covers
    constexpr int kPerLane = T::kTargets / NT;
    // draw k covers target rows j < (k + 1) kTRowsPerDraw, and row j reads
    // LDS rows 2 j + cy .. 2 j + cy + 2 (cy <= 1): the last early draw's rows
    // end below kR1 when 2 kA1 kTRowsPerDraw + 1 < kR1
    constexpr int kA0 = (kR1 - 2) / (2 * kTRowsPerDraw);
    constexpr int kA1 = kA0 < kPerLane ? kA0 : kPerLane;
    static_assert(kA1 == 0 || 2 * (kA1 * kTRowsPerDraw - 1) + 3 < kR1, "early draws read landed rows only");
    uint32_t lo = 0xFFFFFFFFu;
    auto land = [&](int k0, int k1) {
#pragma unroll
        for (int k = 0; k < kLoads; ++k) {
            if (k < k0 || k >= k1 || vdst[k] < 0) continue;
            *reinterpret_cast<uint4 *>(tile + vdst[k]) = v[k];
            if (p.needs_medium) lo = min(lo, min(min(v[k].x, v[k].y), min(v[k].z, v[k].w)));
        }
    };
    const int cx = a.cx & 1, cy = a.cy & 1;
    const PlainTile<PX, PY> pt{tile, (uint32_t)(y0 + cy) * (uint32_t)W + (uint32_t)(x0 + cx),
                               (1 + cy) * kPitch + 4 + cx, 2u * (uint32_t)W};
    uint32_t moves = 0, picks = 0;
    land(0, kA1 > 0 ? kL1 : kLoads);
    tile[hpos] = hv;
    lo = min(lo, h_live ? hv : 0xFFFFFFFFu);
    // entries past the table's length are never read; lanes past the LDS
    // table store into a spare word (an unconditional use, as for the halo)
    *(tid < kPThr ? &thr_s[tid] : &tile[kSpare + (tid & 1)]) = tv;
    if constexpr (kA1 > 0) {
        __syncthreads();
        pt.template draw<NT, 0, kA1>(a, moves, picks);
        land(kL1, kLoads);
    }
    const bool e_zero = a.pass_flags && f1 == 0;
    const bool fold_idle = e_zero && f2 == 0;
    // the fold's loads go out with the tile's whatever the flags say (so they
    // wait on nothing); an idle fold only skips its stores
    if (!fold_idle && fi < fhi) {
        a.a_next[fi] = fa + fe;
        a.g_zero[fi] = 0;
    }
    for (uint32_t i = fi + NT; !fold_idle && i < fhi; i += NT) {
        a.a_next[i] = a.a_cur[i] + a.e_prev[i];
        a.g_zero[i] = 0;
    }
    const bool has_medium = __syncthreads_or(p.needs_medium && lo == 0u) != 0;

    pt.template draw<NT, kA1, kPerLane>(a, moves, picks);
    pt.template queue_moves<NT>(moves, picks, queue, &qcount);
    __syncthreads();
    uint32_t d_old, d_new;
    const uint32_t nq = qcount;
    const uint32_t my_acc = e_zero ? pt.template phase_b<NT, true>(a, p, thr_s, queue, nq, has_medium, d_old, d_new)
                                   : pt.template phase_b<NT, false>(a, p, thr_s, queue, nq, has_medium, d_old, d_new);
    if (my_acc) atomicAdd(&acc_block, my_acc);
    __syncthreads();
    // every queue entry is read; a tile with accepted copies merges their
    // record changes (a frozen tile takes no further barrier). The merge pays
    // where a pass is many rounds of workgroups and the atomics' traffic
    // counts; on a lattice of one short round its barriers only lengthen the
    // chain, and the host leaves it off.
    const uint32_t acc = acc_block;
    if (acc != 0 && a.merge_deltas) pt.template merge_deltas<NT>(a, queue, d_old, d_new);
    if (tid == 0 && acc != 0) {
        atomicAdd(a.accepted + (blockIdx.x & (kAcceptSlots - 1)), (unsigned long long)acc);
        if (a.pass_flags) a.pass_flags[a.pass_slot] = 1u;
    }
}

// ---- the resident pass: a small lattice stepped whole inside one workgroup
//
// Below about 128^2 sites a colour pass is a few microseconds of dependent
// latency, and four launches a step cost more than the work. One workgroup
// of 1024 lanes then holds the whole lattice and every label's record in
// LDS and runs `steps` Monte Carlo steps, colour order and pass keys derived
// in the kernel as the host derives them. Each pass keeps the per-colour
// snapshot: the records stay fixed while the pass prices against them, its
// copies accumulate into a second array, and a fold after the pass's last
// barrier adds them in. The lattice is written straight into LDS: a target
// only reads sites of the other colours, which no target of its pass writes.
// On the way out the lattice, the folded records and cleared delta buffers
// leave the global state exactly as the streaming passes would have.
constexpr int kResThreads = 1024;
constexpr int kResSites = 16384;   // lattice words held in LDS (64 KiB)
constexpr int kResLabels = 2048;   // per-label records held in LDS, twice 16 KiB
constexpr int kResMaxSteps = 64;   // steps per launch

__global__ __launch_bounds__(kResThreads) void sweep_plain2d_resident_kernel(PlainResidentArgs r, PlainParams p) {
    __shared__ __attribute__((aligned(16))) uint32_t lat[kResSites];
    __shared__ int64_t rec[kResLabels];
    __shared__ unsigned long long delta[kResLabels];
    __shared__ uint16_t queue[kResSites / 4];
    __shared__ uint32_t thr_s[kPThr];
    __shared__ uint32_t qcount;
    // every pass's colour and keys, derived once per launch
    __shared__ uint32_t pass_keys[2 * 4 * kResMaxSteps];
    __shared__ uint8_t pass_colour[4 * kResMaxSteps];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wl = r.w_log2;
    const int W = 1 << wl, H = r.h;
    const uint32_t quads = ((uint32_t)H << wl) >> 2;  // 16-byte words of the lattice
    const uint32_t nl = *r.n_labels;

    for (uint32_t i = tid; i < quads; i += kResThreads)
        reinterpret_cast<uint4 *>(lat)[i] = reinterpret_cast<const uint4 *>(r.labels)[i];
    for (uint32_t i = tid; i < nl; i += kResThreads) {
        rec[i] = r.a_start[i] + r.e_start[i];
        delta[i] = 0ull;
    }
    if (tid < (int)p.n_thr) thr_s[tid] = p.thr[tid];
    if (tid == 0) qcount = 0;
    if (tid < (int)(4u * r.steps)) {
        int order[4];
        colour_order(r.seed, r.replica, r.mcs0 + (uint32_t)tid / 4, 4, order);
        const int colour = order[tid & 3];
        const uint32_t key = pass_key(r.seed, r.replica, r.mcs0 + (uint32_t)tid / 4, (uint32_t)colour);
        pass_keys[2 * tid] = key;
        pass_keys[2 * tid + 1] = second_key(key);
        pass_colour[tid] = (uint8_t)colour;
    }
    __syncthreads();

    const int tw = wl - 1;                         // log2 of targets per row
    const uint32_t nt = quads;                      // targets per pass: a quarter of the sites
    const uint32_t nt_pad = (nt + kResThreads - 1) / kResThreads * kResThreads;
    uint32_t my_acc = 0;
    for (uint32_t step = 0; step < r.steps; ++step) {
        for (uint32_t c = 0; c < 4; ++c) {
            const uint32_t k = 4 * step + c;
            const int colour = pass_colour[k];
            const int cx = colour & 1, cy = (colour >> 1) & 1;
            const uint32_t key = pass_keys[2 * k], key2 = pass_keys[2 * k + 1];

            // phase A: every target's draw; queue the ones whose label would change
            for (uint32_t q = tid; q < nt_pad; q += kResThreads) {
                bool moves = false;
                uint32_t site = 0;
                if (q < nt) {
                    const int x = 2 * (int)(q & ((1u << tw) - 1u)) + cx;
                    const int y = 2 * (int)(q >> tw) + cy;
                    site = ((uint32_t)y << wl) | (uint32_t)x;
                    const uint32_t sh = (site_pick(key, site) & 7u) * 2u;
                    const int nx = (x + (int)((0x9224u >> sh) & 3u) - 1) & (W - 1);
                    int ny = y + (int)((0xA940u >> sh) & 3u) - 1;
                    ny += ny < 0 ? H : 0;
                    ny -= ny >= H ? H : 0;
                    moves = lat[site] != lat[((uint32_t)ny << wl) | (uint32_t)nx];
                }
                const unsigned long long m = __ballot(moves);
                if (m == 0) continue;
                uint32_t base = 0;
                if (lane == 0) base = atomicAdd(&qcount, (uint32_t)__builtin_popcountll(m));
                base = __builtin_amdgcn_readfirstlane(base);
                if (moves) {
                    const uint32_t pos = base + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                                          __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
                    queue[pos] = (uint16_t)site;
                }
            }
            __syncthreads();

            // phase B: price the queued targets against the pass-start records
            const uint32_t nq = qcount;
            for (uint32_t qi = tid; qi < nq; qi += kResThreads) {
                const uint32_t site = queue[qi];
                const int x = (int)(site & (uint32_t)(W - 1)), y = (int)(site >> wl);
                const uint32_t ph = site_pick(key, site);
                const uint32_t xm = (uint32_t)((x - 1) & (W - 1)), xp = (uint32_t)((x + 1) & (W - 1));
                const uint32_t ym = (uint32_t)(y == 0 ? H - 1 : y - 1) << wl;
                const uint32_t y0 = (uint32_t)y << wl;
                const uint32_t yp = (uint32_t)(y == H - 1 ? 0 : y + 1) << wl;
                const uint32_t n0 = lat[ym | xm], n1 = lat[ym | (uint32_t)x], n2 = lat[ym | xp];
                const uint32_t n3 = lat[y0 | xm], n4 = lat[y0 | xp];
                const uint32_t n5 = lat[yp | xm], n6 = lat[yp | (uint32_t)x], n7 = lat[yp | xp];
                const uint32_t old = lat[site];
                // the picked neighbour, k in ORDER3's planar order, read at
                // its offset (a select chain on k would branch per lane)
                const uint32_t sh = (ph & 7u) * 2u;
                const uint32_t nxp = (uint32_t)((x + (int)((0x9224u >> sh) & 3u) - 1) & (W - 1));
                const uint32_t dy = (0xA940u >> sh) & 3u;
                const uint32_t nw = lat[(dy == 0 ? ym : dy == 1 ? y0 : yp) | nxp];
                const uint32_t oid = label_id(old), nid = label_id(nw);
                const PlainMove mv = plain_move(p, thr_s, key, key2, site, old, nw, n0, n1, n2, n3, n4, n5, n6, n7,
                                                rec[oid], rec[nid], p.needs_medium != 0);
                if (!mv.acc) continue;
                lat[site] = nw;
                if (oid != 0) atomicAdd(&delta[oid], (unsigned long long)pack_vs(-1, mv.ds_old));
                if (nid != 0) atomicAdd(&delta[nid], (unsigned long long)pack_vs(1, mv.ds_new));
                ++my_acc;
            }
            __syncthreads();

            // fold the pass's copies into the records the next pass reads
            for (uint32_t i = tid; i < nl; i += kResThreads) {
                rec[i] += (int64_t)delta[i];
                delta[i] = 0ull;
            }
            if (tid == 0) qcount = 0;
            __syncthreads();
        }
    }

    for (uint32_t i = tid; i < quads; i += kResThreads)
        reinterpret_cast<uint4 *>(r.labels)[i] = reinterpret_cast<const uint4 *>(lat)[i];
    for (uint32_t i = tid; i < nl; i += kResThreads) {
        r.a_end[i] = rec[i];
        r.e_clear0[i] = 0;
        r.e_clear1[i] = 0;
    }
    // per-lane counts summed across the wavefront, one atomic per wavefront
    uint32_t t = my_acc;
    for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off, 64);
    if (lane == 0 && t) atomicAdd(r.accepted + ((tid >> 6) & (kAcceptSlots - 1)), (unsigned long long)t);
}

bool plain2d_resident_fits(int w, int h, uint32_t capacity) {
    const bool pow2 = w >= 4 && (w & (w - 1)) == 0;
    return pow2 && h >= 2 && h % 2 == 0 && (size_t)w * (size_t)h <= (size_t)kResSites &&
           capacity <= (uint32_t)kResLabels;
}

hipError_t launch_sweep_plain2d_resident(const PlainResidentArgs &r, const PlainParams &p, hipStream_t stream) {
    if (r.steps > (uint32_t)kResMaxSteps || !plain2d_resident_fits(1 << r.w_log2, r.h, r.capacity))
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(sweep_plain2d_resident_kernel, dim3(1), dim3(kResThreads), 0, stream, r, p);
    return hipGetLastError();
}

// ---- the tiled resident pass: a mid-size lattice, one resident workgroup
// per tile
//
// Above the single workgroup's reach a streaming pass on a lattice of up to
// a million sites is mostly latency: each pass is a dependent launch around
// a chain of memory round trips.
// These lattices are instead cut into at most 64 tiles of up to 128 x 128
// sites (the count nearest 32, which measured fastest), one 1024-lane
// workgroup per tile, all resident at once (a cooperative launch), and a
// pass ends at a grid barrier instead of a kernel boundary, up to 64 steps
// per launch. Each workgroup keeps its tile, a copy of every label's record,
// the acceptance table and every pass's colour and keys in LDS for the
// whole launch. Per pass it refreshes its one-site halo ring from the
// lattice in memory and folds the previous pass's record changes into its
// copy (all of a lane's loads in flight together), then draws, prices and
// commits exactly as the other passes do (the records fixed while the pass
// prices): a tile of at most one target a lane prices each target in the
// lane that drew it, a larger one queues its changing targets first. It
// writes accepted labels through to memory and adds record changes into the
// pass's delta buffer.
//
// What workgroups share inside the launch is written with agent-scope
// stores and atomics (which go through to the coherent level) and read with
// agent-scope loads, so no L2 is written back or invalidated at a barrier:
// every wave waits for the acknowledgements of its own stores, then its
// workgroup publishes the barrier's epoch in its own arrival word (each on a
// cache line of its own: 32 writers on one line cost about 0.3 us a
// barrier), and wave 0 waits until every workgroup's word has reached that
// epoch. The wait is bounded: a workgroup that gives up sets the error word
// and leaves, the others then give up at their next barrier, and the host
// reports it.
constexpr int kCoopThreads = 1024;
constexpr int kCoopSide = 128;        // largest tile side
constexpr int kCoopPitch = kCoopSide + 2;
constexpr int kCoopLabels = 8192;     // records held in LDS
constexpr int kCoopMaxSteps = 64;     // steps per launch
constexpr int kCoopTiles = 32;        // preferred tile count

namespace {
__device__ __forceinline__ uint32_t load_agent(const uint32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int64_t load_agent(const int64_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void store_agent(uint32_t *p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void store_agent(int64_t *p, int64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void add_agent(int64_t *p, int64_t v) {
    (void)__hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// All workgroups of the launch meet; false when the wait gave up.
__device__ __forceinline__ bool coop_barrier(uint32_t *arrive, uint32_t nwg, uint32_t epoch, uint32_t patience,
                                             uint32_t *fail_s, uint32_t *err) {
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): this wave's stores and atomics acknowledged
    __syncthreads();
    if (threadIdx.x < 64) {
        if (threadIdx.x == 0) store_agent(&arrive[blockIdx.x * kCoopArriveStride], epoch);
        bool met = false;
        for (uint32_t it = 0; it < patience && !met; ++it)
            met = __all(threadIdx.x >= nwg || (int32_t)(load_agent(&arrive[threadIdx.x * kCoopArriveStride]) - epoch) >= 0);
        if (!met && threadIdx.x == 0) {
            *fail_s = 1;
            atomicAdd(err, 1u);
        }
    }
    __syncthreads();
    return *fail_s == 0;
}
}  // namespace

__global__ __launch_bounds__(kCoopThreads) void sweep_plain2d_coop_kernel(PlainCoopArgs r, PlainParams p) {
    __shared__ uint32_t lat[kCoopPitch * kCoopPitch];
    __shared__ int64_t rec[kCoopLabels];
    __shared__ uint16_t queue[kCoopSide * kCoopSide / 4];
    __shared__ uint32_t thr_s[kPThr];
    __shared__ uint32_t qcount, fail_s;
    // every pass's colour and keys, derived once per launch
    __shared__ uint32_t pass_keys[2 * 4 * kCoopMaxSteps];
    __shared__ uint8_t pass_colour[4 * kCoopMaxSteps];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wl = r.w_log2, txl = r.tx_log2, tyl = r.ty_log2;
    const int W = 1 << wl, H = r.h, TX = 1 << txl, TY = 1 << tyl, P = TX + 2;
    const uint32_t nwg = gridDim.x, b = blockIdx.x;
    const int x0 = (int)(b & ((1u << (wl - txl)) - 1u)) << txl;
    const int y0 = (int)(b >> (wl - txl)) << tyl;
    const uint32_t nl = *r.n_labels;
    if (nl > (uint32_t)kCoopLabels) {
        // every workgroup leaves here, before any barrier
        if (b == 0 && tid == 0) atomicAdd(r.err, 1u);
        return;
    }
    // this workgroup's share of the delta buffer a pass clears
    const uint32_t share = (nl + nwg - 1) / nwg;
    const uint32_t lo = min(b * share, nl), hi = min(lo + share, nl);

    // every load of a lane issued before the first wait: the tile in 16-byte
    // quads (at most four a lane), the records (at most eight)
    constexpr int kQuads = kCoopSide * kCoopSide / 4 / kCoopThreads;
    constexpr int kRecs = kCoopLabels / kCoopThreads;
    {
        const int quads = (TX * TY) >> 2, qrow = TX >> 2;
        uint4 v[kQuads];
#pragma unroll
        for (int j = 0; j < kQuads; ++j) {
            const int i = tid + j * kCoopThreads;
            if (i < quads) {
                const int ly = i / qrow, lq = i - ly * qrow;
                v[j] = reinterpret_cast<const uint4 *>(r.labels)[(((uint32_t)(y0 + ly) << wl) |
                                                                  (uint32_t)(x0 + 4 * lq)) >> 2];
            }
        }
        int64_t a[kRecs];
#pragma unroll
        for (int j = 0; j < kRecs; ++j) {
            const uint32_t i = tid + j * kCoopThreads;
            if (i < nl) a[j] = r.a_start[i];
        }
#pragma unroll
        for (int j = 0; j < kQuads; ++j) {
            const int i = tid + j * kCoopThreads;
            if (i < quads) {
                const int ly = i / qrow, lq = i - ly * qrow;
                uint32_t *row = &lat[(ly + 1) * P + 1 + 4 * lq];
                row[0] = v[j].x;
                row[1] = v[j].y;
                row[2] = v[j].z;
                row[3] = v[j].w;
            }
        }
#pragma unroll
        for (int j = 0; j < kRecs; ++j) {
            const uint32_t i = tid + j * kCoopThreads;
            if (i < nl) rec[i] = a[j];
        }
    }
    if (tid < (int)p.n_thr) thr_s[tid] = p.thr[tid];
    if (tid == 0) fail_s = 0;
    const uint32_t passes = 4u * r.steps;
    if (tid < (int)passes) {
        int order[4];
        colour_order(r.seed, r.replica, r.mcs0 + (uint32_t)tid / 4, 4, order);
        const int colour = order[tid & 3];
        const uint32_t key = pass_key(r.seed, r.replica, r.mcs0 + (uint32_t)tid / 4, (uint32_t)colour);
        pass_keys[2 * tid] = key;
        pass_keys[2 * tid + 1] = second_key(key);
        pass_colour[tid] = (uint8_t)colour;
    }

    // the halo ring, one site per lane: the rows above and below (corners
    // included), then the columns left and right
    const int ring = 2 * (TX + 2) + 2 * TY;
    int ring_lds = 0;
    uint32_t ring_site = 0;
    if (tid < ring) {
        int lx, ly;
        if (tid < TX + 2) {
            lx = tid - 1;
            ly = -1;
        } else if (tid < 2 * (TX + 2)) {
            lx = tid - (TX + 2) - 1;
            ly = TY;
        } else if (tid < 2 * (TX + 2) + TY) {
            lx = -1;
            ly = tid - 2 * (TX + 2);
        } else {
            lx = TX;
            ly = tid - 2 * (TX + 2) - TY;
        }
        int gy = y0 + ly;
        gy += gy < 0 ? H : 0;
        gy -= gy >= H ? H : 0;
        ring_site = ((uint32_t)gy << wl) | ((uint32_t)(x0 + lx) & (uint32_t)(W - 1));
        ring_lds = (ly + 1) * P + lx + 1;
    }

    const int tw = txl - 1;                       // log2 of targets per tile row
    const uint32_t nt = (uint32_t)(TX * TY) >> 2;  // targets per pass, a multiple of 64
    uint32_t my_acc = 0;
    uint32_t epoch = r.epoch0;
    uint32_t rot = (uint32_t)(r.pass0 % 3);  // this pass's delta buffer
    __syncthreads();                          // the pass table is read from here on
    for (uint32_t k = 0; k < passes; ++k) {
        const int colour = pass_colour[k];
        const int cx = colour & 1, cy = (colour >> 1) & 1;
        const uint32_t key = pass_keys[2 * k], key2 = pass_keys[2 * k + 1];
        const int64_t *d_prev = r.d[rot == 0 ? 2 : rot - 1];
        int64_t *d_acc = r.d[rot];
        int64_t *d_clear = r.d[rot == 2 ? 0 : rot + 1];
        rot = rot == 2 ? 0 : rot + 1;

        // the halo as the last pass left it, the records as this pass prices them
        {
            const uint32_t halo = tid < ring ? load_agent(&r.labels[ring_site]) : 0u;
            int64_t dv[kRecs];
#pragma unroll
            for (int j = 0; j < kRecs; ++j) {
                const uint32_t i = tid + j * kCoopThreads;
                if (i < nl) dv[j] = load_agent(d_prev + i);
            }
            for (uint32_t i = lo + tid; i < hi; i += kCoopThreads) store_agent(d_clear + i, (int64_t)0);
            if (tid < ring) lat[ring_lds] = halo;
            // LDS adds with no return: no read round trip per record
#pragma unroll
            for (int j = 0; j < kRecs; ++j) {
                const uint32_t i = tid + j * kCoopThreads;
                if (i < nl) (void)__hip_atomic_fetch_add(&rec[i], dv[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            }
        }
        if (tid == 0) qcount = 0;
        __syncthreads();

        if (nt <= (uint32_t)kCoopThreads) {
            // at most one target a lane: each lane prices its own target where
            // its label would change, with no queue
            if (tid < (int)nt) {
                const uint32_t q = (uint32_t)tid;
                const int lx = 2 * (int)(q & ((1u << tw) - 1u)) + cx;
                const int ly = 2 * (int)(q >> tw) + cy;
                const uint32_t site = ((uint32_t)(y0 + ly) << wl) | (uint32_t)(x0 + lx);
                const uint32_t pk = site_pick(key, site) & 7u;
                const uint32_t sh = pk * 2u;
                const int tl = (ly + 1) * P + lx + 1;
                const uint32_t old = lat[tl];
                const uint32_t nw = lat[tl + (int)((0x9224u >> sh) & 3u) - 1 + ((int)((0xA940u >> sh) & 3u) - 1) * P];
                if (old != nw) {
                    const uint32_t n0 = lat[tl - P - 1], n1 = lat[tl - P], n2 = lat[tl - P + 1];
                    const uint32_t n3 = lat[tl - 1], n4 = lat[tl + 1];
                    const uint32_t n5 = lat[tl + P - 1], n6 = lat[tl + P], n7 = lat[tl + P + 1];
                    const uint32_t oid = label_id(old), nid = label_id(nw);
                    const PlainMove mv = plain_move(p, thr_s, key, key2, site, old, nw, n0, n1, n2, n3, n4, n5, n6,
                                                    n7, rec[oid], rec[nid], p.needs_medium != 0);
                    if (mv.acc) {
                        lat[tl] = nw;
                        store_agent(&r.labels[site], nw);
                        if (oid != 0) add_agent(d_acc + oid, pack_vs(-1, mv.ds_old));
                        if (nid != 0) add_agent(d_acc + nid, pack_vs(1, mv.ds_new));
                        ++my_acc;
                    }
                }
            }
        } else {
            // phase A: every target's draw; queue (target, pick) where the label would change
            for (uint32_t q = tid; q < nt; q += kCoopThreads) {
                const int lx = 2 * (int)(q & ((1u << tw) - 1u)) + cx;
                const int ly = 2 * (int)(q >> tw) + cy;
                const uint32_t site = ((uint32_t)(y0 + ly) << wl) | (uint32_t)(x0 + lx);
                const uint32_t pk = site_pick(key, site) & 7u;
                const uint32_t sh = pk * 2u;
                const int tl = (ly + 1) * P + lx + 1;
                const int nb = tl + (int)((0x9224u >> sh) & 3u) - 1 + ((int)((0xA940u >> sh) & 3u) - 1) * P;
                const bool moves = lat[tl] != lat[nb];
                const unsigned long long m = __ballot(moves);
                if (m == 0) continue;
                uint32_t base = 0;
                if (lane == 0) base = atomicAdd(&qcount, (uint32_t)__builtin_popcountll(m));
                base = __builtin_amdgcn_readfirstlane(base);
                if (moves) {
                    const uint32_t pos = base + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                                          __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
                    queue[pos] = (uint16_t)(q | pk << 12);
                }
            }
            __syncthreads();

            // phase B: price the queued targets against the pass-start records
            const uint32_t nq = qcount;
            for (uint32_t qi = tid; qi < nq; qi += kCoopThreads) {
                const uint32_t e = queue[qi];
                const uint32_t q = e & 0xFFFu, pk = e >> 12;
                const int lx = 2 * (int)(q & ((1u << tw) - 1u)) + cx;
                const int ly = 2 * (int)(q >> tw) + cy;
                const uint32_t site = ((uint32_t)(y0 + ly) << wl) | (uint32_t)(x0 + lx);
                const int tl = (ly + 1) * P + lx + 1;
                const uint32_t n0 = lat[tl - P - 1], n1 = lat[tl - P], n2 = lat[tl - P + 1];
                const uint32_t n3 = lat[tl - 1], n4 = lat[tl + 1];
                const uint32_t n5 = lat[tl + P - 1], n6 = lat[tl + P], n7 = lat[tl + P + 1];
                const uint32_t old = lat[tl];
                // the picked neighbour read at its offset (a select chain on the
                // pick would branch per lane)
                const uint32_t sh = pk * 2u;
                const uint32_t nw = lat[tl + (int)((0x9224u >> sh) & 3u) - 1 + ((int)((0xA940u >> sh) & 3u) - 1) * P];
                const uint32_t oid = label_id(old), nid = label_id(nw);
                const PlainMove mv = plain_move(p, thr_s, key, key2, site, old, nw, n0, n1, n2, n3, n4, n5, n6, n7,
                                                rec[oid], rec[nid], p.needs_medium != 0);
                if (!mv.acc) continue;
                lat[tl] = nw;
                store_agent(&r.labels[site], nw);
                if (oid != 0) add_agent(d_acc + oid, pack_vs(-1, mv.ds_old));
                if (nid != 0) add_agent(d_acc + nid, pack_vs(1, mv.ds_new));
                ++my_acc;
            }
        }
        if (k + 1 == passes) break;
        if (!coop_barrier(r.arrive, nwg, ++epoch, r.patience, &fail_s, r.err)) return;
    }

    // the records as the last pass priced them; its changes stay in its delta buffer
    for (uint32_t i = lo + tid; i < hi; i += kCoopThreads) r.a_end[i] = rec[i];
    uint32_t t = my_acc;
    for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off, 64);
    if (lane == 0 && t) atomicAdd(r.accepted + ((b * 16u + (uint32_t)(tid >> 6)) & (kAcceptSlots - 1)),
                                  (unsigned long long)t);
}

bool plain2d_coop_shape(int w, int h, uint32_t capacity, int *tx_log2, int *ty_log2, int *blocks) {
    const bool pow2 = w >= 32 && (w & (w - 1)) == 0;
    if (!pow2 || h < 32 || capacity > (uint32_t)kCoopLabels || (size_t)w * (size_t)h <= (size_t)kResSites)
        return false;
    // the shape whose tile count is nearest kCoopTiles, squarer tiles first
    int best = -1, best_tx = 0, best_ty = 0;
    for (int txl = 7; txl >= 5; --txl)
        for (int tyl = 7; tyl >= 5; --tyl) {
            const int tx = 1 << txl, ty = 1 << tyl;
            if (tx > w || ty > h || h % ty != 0) continue;
            const long n = (long)(w / tx) * (long)(h / ty);
            if (n > kCoopMaxBlocks) continue;
            const int score = (int)(n > kCoopTiles ? n - kCoopTiles : kCoopTiles - n) * 8 + (txl > tyl ? txl - tyl : tyl - txl);
            if (best < 0 || score < best) {
                best = score;
                best_tx = txl;
                best_ty = tyl;
            }
        }
    if (best < 0) return false;
    *tx_log2 = best_tx;
    *ty_log2 = best_ty;
    *blocks = (w >> best_tx) * (h >> best_ty);
    return true;
}

bool plain2d_coop_resident(int blocks, int cus) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, sweep_plain2d_coop_kernel, kCoopThreads, 0) !=
        hipSuccess)
        return false;
    return (long)per_cu * (long)cus >= (long)blocks;
}

hipError_t launch_sweep_plain2d_coop(const PlainCoopArgs &r, const PlainParams &p, int blocks, hipStream_t stream) {
    int txl = 0, tyl = 0, n = 0;
    if (r.steps == 0 || r.steps > (uint32_t)kCoopMaxSteps || !plain2d_coop_shape(1 << r.w_log2, r.h, 0, &txl, &tyl, &n) || txl != r.tx_log2 ||
        tyl != r.ty_log2 || n != blocks)
        return hipErrorInvalidValue;
    PlainCoopArgs a = r;
    PlainParams pp = p;
    void *args[] = {&a, &pp};
    return hipLaunchCooperativeKernel(reinterpret_cast<const void *>(sweep_plain2d_coop_kernel), dim3(blocks),
                                      dim3(kCoopThreads), args, 0, stream);
}

bool plain2d_supported(uint32_t n_thr) { return n_thr <= (uint32_t)kPThr; }

void plain2d_tile(int w, int h, int *px, int *py) {
    // up to 512^2 a pass is a few tiles per CU at most, and the smaller
    // tile's shorter workgroup lifetime is what a pass costs
    if (w % 128 == 0 && h % 8 == 0 && (size_t)w * (size_t)h <= ((size_t)1 << 16)) {
        *px = 128;
        *py = 8;
    } else if (w % 256 == 0 && h % 32 == 0 && (size_t)w * (size_t)h > ((size_t)1 << 18)) {
        *px = 256;
        *py = 32;
    } else if (w % 128 == 0 && h % 16 == 0) {
        *px = 128;
        *py = 16;
    } else {
        *px = *py = 0;
    }
    // the pass's 24-bit row multiplies need two rows below 2^24 sites, and
    // its tile-row reciprocal tiles_x * blocks below 2^32
    if (*px > 0) {
        const uint64_t tx = (uint64_t)(w / *px), blocks = tx * (uint64_t)(h / *py);
        if (w >= (1 << 23) || tx * blocks >= ((uint64_t)1 << 32)) *px = *py = 0;
    }
}

hipError_t launch_sweep_plain2d(const SweepArgs &a, const PlainParams &p, int blocks,
                                hipStream_t stream) {
    int px = 0, py = 0;
    plain2d_tile(a.w, a.h, &px, &py);
    // one record word per label: the pass addresses a label's record by its id
    if (a.cell_words != 1) return hipErrorInvalidValue;
    if (px == 256)
        hipLaunchKernelGGL((sweep_plain2d_kernel<256, 32, 512>), dim3(blocks), dim3(512), 0, stream, a, p);
    else if (px == 128 && py == 8)
        hipLaunchKernelGGL((sweep_plain2d_kernel<128, 8, 256>), dim3(blocks), dim3(256), 0, stream, a, p);
    else if (px == 128)
        hipLaunchKernelGGL((sweep_plain2d_kernel<128, 16, 256>), dim3(blocks), dim3(256), 0, stream, a, p);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

}  // namespace ghip
