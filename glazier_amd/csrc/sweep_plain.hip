// sweep_plain.hip - the colour pass specialised for the plain monolayer: a
// plane, the eight Moore neighbours, the medium plus one cell type, and
// integer energies (contact, volume, surface). This is the throughput path.
//
// Same draws, same decisions and the same bookkeeping as sweep_kernel (and so
// as the CPU emulation); only the work is organised for gfx950, where this
// pass is bound by VALU issue rather than by bytes:
//
//  * a 256-lane workgroup owns 128 x 16 sites. Every global read of the
//    workgroup (18 tile rows as 16-byte loads, the halo columns, the
//    acceptance table and its share of the snapshot fold) is issued before
//    the first wait; rows land in an LDS tile padded to a 136-word pitch
//    (interior at column 4, so each 16-byte LDS write is aligned);
//  * phase A draws every target. A wavefront's 64 targets share one tile row,
//    so row bases stay in scalar registers; the target's label and its picked
//    neighbour come from LDS, and the targets whose label would change are
//    appended to an LDS queue (one LDS atomic per wavefront). In a confluent
//    sheet three targets in four have old == new and cost nothing more;
//  * phase B prices the queued targets densely. With integer energies the
//    contact sum is exact in any order, so it comes from neighbour counts
//    taken as min(a ^ b, 1) sums (no compare-to-mask hazards), and the medium
//    count is skipped for tiles that hold no medium.
#include "ghip_device.h"

namespace ghip {

constexpr int kPX = 128;
constexpr int kPY = 16;
constexpr int kPitch = 136;
constexpr int kRows = kPY + 2;
constexpr int kPTargets = kPX * kPY / 4;
constexpr int kPThr = 256;

namespace {
// Count of neighbours differing from v: sum of min(n ^ v, 1).
__device__ __forceinline__ uint32_t unequal8(uint32_t v, uint32_t n0, uint32_t n1, uint32_t n2,
                                             uint32_t n3, uint32_t n4, uint32_t n5, uint32_t n6,
                                             uint32_t n7) {
    const uint32_t a = min(n0 ^ v, 1u) + min(n1 ^ v, 1u) + min(n2 ^ v, 1u);
    const uint32_t b = min(n3 ^ v, 1u) + min(n4 ^ v, 1u) + min(n5 ^ v, 1u);
    return a + b + min(n6 ^ v, 1u) + min(n7 ^ v, 1u);
}
}  // namespace

__global__ __launch_bounds__(256) void sweep_plain2d_kernel(SweepArgs a, PlainParams p) {
    __shared__ __attribute__((aligned(16))) uint32_t tile[kRows * kPitch];
    __shared__ uint4 queue[kPTargets];  // (tile offset, r24, old label, new label)
    __shared__ uint32_t thr_s[kPThr];
    __shared__ uint32_t qcount;
    __shared__ uint32_t acc_block;

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    int b = blockIdx.x;
    if (a.xcd_chunk > 0) b = (b & 7) * a.xcd_chunk + (b >> 3);
    const int tile_x = b % a.tiles_x;
    const int tile_y = b / a.tiles_x;
    const int x0 = tile_x * kPX, y0 = tile_y * kPY;
    const int W = a.w, H = a.h;

    if (tid == 0) {
        qcount = 0;
        acc_block = 0;
    }

    // ---- issue every global read of the workgroup before the first wait.
    // A wavefront loads two tile rows per step: lanes 0-31 the first, lanes
    // 32-63 the second; the row offsets are scalars chosen per half-wave.
    const int half = lane >> 5;
    const uint32_t col4 = (uint32_t)(lane & 31) * 4u;
    uint4 v[3];
    int vdst[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int r0 = 8 * k + 2 * wave;  // uniform
        vdst[k] = -1;
        if (r0 < kRows) {
            int g0 = y0 - 1 + r0, g1 = g0 + 1;
            g0 += g0 < 0 ? H : 0;
            g0 -= g0 >= H ? H : 0;
            g1 -= g1 >= H ? H : 0;
            const uint32_t row = (uint32_t)(half ? g1 : g0) * (uint32_t)W + (uint32_t)x0;
            v[k] = *reinterpret_cast<const uint4 *>(a.labels + row + col4);
            vdst[k] = (r0 + half) * kPitch + 4 + (int)col4;
        }
    }
    uint32_t hv = 0;
    int hpos = -1;
    if (tid < kRows * 2) {
        const int r = tid >> 1, right = tid & 1;
        int gy = y0 - 1 + r;
        gy += gy < 0 ? H : 0;
        gy -= gy >= H ? H : 0;
        int gx = right ? x0 + kPX : x0 - 1;
        gx += gx < 0 ? W : 0;
        gx -= gx >= W ? W : 0;
        hv = a.labels[(uint32_t)gy * (uint32_t)W + (uint32_t)gx];
        hpos = r * kPitch + (right ? 4 + kPX : 3);
    }
    uint32_t tv = 0;
    if (tid < (int)p.n_thr) tv = p.thr[tid];
    const uint32_t words = *a.n_labels * (uint32_t)a.cell_words;
    const uint32_t flo = blockIdx.x * a.fold_per_block;
    const uint32_t fhi = min(flo + a.fold_per_block, words);
    const uint32_t fi = flo + tid;
    int64_t fa = 0, fe = 0;
    if (fi < fhi) {
        fa = a.a_cur[fi];
        fe = a.e_prev[fi];
    }

    // ---- land them
    uint32_t lo = 0xFFFFFFFFu;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (vdst[k] >= 0) {
            *reinterpret_cast<uint4 *>(tile + vdst[k]) = v[k];
            if (p.needs_medium) lo = min(lo, min(min(v[k].x, v[k].y), min(v[k].z, v[k].w)));
        }
    }
    if (hpos >= 0) {
        tile[hpos] = hv;
        lo = min(lo, hv);
    }
    if (tid < (int)p.n_thr) thr_s[tid] = tv;
    if (fi < fhi) {
        a.a_next[fi] = fa + fe;
        a.g_zero[fi] = 0;
    }
    for (uint32_t i = fi + 256; i < fhi; i += 256) {
        a.a_next[i] = a.a_cur[i] + a.e_prev[i];
        a.g_zero[i] = 0;
    }
    const bool has_medium = __syncthreads_or(p.needs_medium && lo == 0u) != 0;

    // ---- phase A: draw every target, queue the ones whose label would change
#pragma unroll
    for (int k = 0; k < kPTargets / 256; ++k) {
        const int j = wave + 4 * k;  // target row, uniform in the wavefront
        const uint32_t srow = (uint32_t)(y0 + 2 * j + a.cy) * (uint32_t)W + (uint32_t)(x0 + a.cx);
        const int tlrow = (1 + 2 * j + a.cy) * kPitch + 4 + a.cx;
        const uint32_t site = srow + 2u * (uint32_t)lane;
        const int tl = tlrow + 2 * lane;
        const uint32_t h = site_draw(a.key, a.key2, site);
        // pick k in ORDER3's planar order, (dx, dy) row by row: dx + 1 and
        // dy + 1 are the k-th 2-bit fields of 0x9224 and 0xA940
        const uint32_t sh = (h & 7u) * 2u;
        const int dxy = (int)(((0xA940u >> sh) & 3u) * (uint32_t)kPitch + ((0x9224u >> sh) & 3u));
        const uint32_t old = tile[tl];
        const uint32_t nw = tile[tl - kPitch - 1 + dxy];
        const unsigned long long m = __ballot(old != nw);
        if (m == 0) continue;
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(&qcount, (uint32_t)__builtin_popcountll(m));
        base = __builtin_amdgcn_readfirstlane(base);
        if (old != nw) {
            const uint32_t pos = base + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                                  __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
            queue[pos] = make_uint4((uint32_t)tl, h >> 8, old, nw);
        }
    }
    __syncthreads();

    // ---- phase B: price the queued targets
    const uint32_t nq = qcount;
    const uint32_t cw = (uint32_t)a.cell_words;
    uint32_t my_acc = 0;
    for (uint32_t qi = tid; qi < nq; qi += 256) {
        const uint4 e = queue[qi];
        const int tl = (int)e.x;
        const uint32_t r24 = e.y, old = e.z, nw = e.w;
        const uint32_t oid = label_id(old), nid = label_id(nw);
        // the four gathers go out together, before anything waits on them;
        // label 0's record is real memory whose value is never used
        const int64_t ao = a.a_cur[oid * cw], eo = a.e_prev[oid * cw];
        const int64_t an = a.a_cur[nid * cw], en = a.e_prev[nid * cw];
        const uint32_t n0 = tile[tl - kPitch - 1], n1 = tile[tl - kPitch], n2 = tile[tl - kPitch + 1];
        const uint32_t n3 = tile[tl - 1], n4 = tile[tl + 1];
        const uint32_t n5 = tile[tl + kPitch - 1], n6 = tile[tl + kPitch], n7 = tile[tl + kPitch + 1];
        const int unlike_old = (int)unequal8(old, n0, n1, n2, n3, n4, n5, n6, n7);
        const int unlike_new = (int)unequal8(nw, n0, n1, n2, n3, n4, n5, n6, n7);
        const int like_old = 8 - unlike_old, like_new = 8 - unlike_new;
        const int64_t pv_old = ao + eo, pv_new = an + en;
        // absolute states keep surface in [0, 2^31): volume is the high word
        const int vol_old = (int)(uint32_t)((uint64_t)pv_old >> 32);
        const int vol_new = (int)(uint32_t)((uint64_t)pv_new >> 32);
        // a copy that would empty a cell is refused (no branch before the
        // gathers are used, so none of them waits behind another)
        const bool emptying = oid != 0 && vol_old <= 1;
        int med = 0;
        if (has_medium) med = 8 - (int)unequal8(0u, n0, n1, n2, n3, n4, n5, n6, n7);
        // contact gained minus lost, from counts (exact: integer energies)
        const int etn = nid != 0 ? 1 : 0, eto = oid != 0 ? 1 : 0;
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
        const int ds_old = 2 * like_old - 8;
        const int ds_new = 8 - 2 * like_new;
        if (oid != 0 && p.lam_v != 0.0f) {
            const float vv = (float)vol_old;
            dh = dh + p.lam_v * (1.0f - 2.0f * (vv - p.tgt_v));
        }
        if (nid != 0 && p.lam_v != 0.0f) {
            const float vv = (float)vol_new;
            dh = dh + p.lam_v * (1.0f + 2.0f * (vv - p.tgt_v));
        }
        if (oid != 0 && ds_old != 0 && p.lam_s != 0.0f) {
            const float s = (float)(int32_t)(uint32_t)pv_old;
            const float dsf = (float)ds_old;
            dh = dh + p.lam_s * (dsf * (2.0f * (s - p.tgt_s) + dsf));
        }
        if (nid != 0 && ds_new != 0 && p.lam_s != 0.0f) {
            const float s = (float)(int32_t)(uint32_t)pv_new;
            const float dsf = (float)ds_new;
            dh = dh + p.lam_s * (dsf * (2.0f * (s - p.tgt_s) + dsf));
        }
        const bool acc = !emptying && (dh <= 0.0f || (dh < (float)p.n_thr && r24 < thr_s[(int)dh]));
        if (!acc) continue;
        const int ly = tl / kPitch, lx = tl - ly * kPitch;
        a.labels[(uint32_t)(y0 + ly - 1) * (uint32_t)W + (uint32_t)(x0 + lx - 4)] = nw;
        if (oid != 0)
            atomicAdd(reinterpret_cast<unsigned long long *>(a.f_acc + oid * cw),
                      (unsigned long long)pack_vs(-1, ds_old));
        if (nid != 0)
            atomicAdd(reinterpret_cast<unsigned long long *>(a.f_acc + nid * cw),
                      (unsigned long long)pack_vs(1, ds_new));
        ++my_acc;
    }
    if (my_acc) atomicAdd(&acc_block, my_acc);
    __syncthreads();
    if (tid == 0 && acc_block)
        atomicAdd(a.accepted + (blockIdx.x & (kAcceptSlots - 1)), (unsigned long long)acc_block);
}

bool plain2d_supported(uint32_t n_thr) { return n_thr <= (uint32_t)kPThr; }

hipError_t launch_sweep_plain2d(const SweepArgs &a, const PlainParams &p, int blocks,
                                hipStream_t stream) {
    hipLaunchKernelGGL(sweep_plain2d_kernel, dim3(blocks), dim3(256), 0, stream, a, p);
    return hipGetLastError();
}

}  // namespace ghip
