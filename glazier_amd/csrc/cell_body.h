// cell_body.h - per-label bookkeeping on gfx950: recounts of volume, surface
// and integer moment sums, anchors, and division/death with label
// compaction, all on the device with no host round trip.
//
// Device counterpart of recount / cell_circular_sums / cell_centroids /
// cell_sums / cell_moments / apply_population (glazier-cuda/src/lib.rs:66-172,
// 255-310) and of the host decisions in run_population (lib.rs:1362-1461),
// which the reference takes through three blocking device->host copies per
// step. Serial counterparts: count_surface, count_moments, divide_and_die
// (glazier-core/src/cpu.rs:11-59, 620-725).
//
// Per-label state is `cw` int64 words: word 0 packs volume * 2^32 + surface;
// words 10, 11 (with the nematic coupling) are the field sums in units of 2^-24;
// words 1..9 (when moments are kept) are the sums of ux, uy, uz, ux^2, uy^2,
// uz^2, ux*uy, ux*uz, uy*uz over the label's sites, u being the site's
// minimum-image offset from the label's integer anchor. Integer sums are
// exact, so their order of accumulation does not matter.
#pragma once

#include "ghip_device.h"
#include "geometry.h"

namespace ghip {

namespace {
// Coordinates of site i; lattices hold fewer than 2^32 sites, so 32-bit
// division suffices.
__device__ __forceinline__ void coords(size_t i, int w, int h, int *x, int *y, int *z) {
    const uint32_t s = (uint32_t)i, plane = (uint32_t)w * (uint32_t)h;
    const uint32_t zz = s / plane, r = s - zz * plane;
    const uint32_t yy = r / (uint32_t)w;
    *z = (int)zz;
    *y = (int)yy;
    *x = (int)(r - yy * (uint32_t)w);
}

// v taken onto [0, span) when it lies at most one span outside.
__device__ __forceinline__ int wrap1(int v, int span) {
    return v < 0 ? v + span : (v >= span ? v - span : v);
}
}  // namespace

__device__ __forceinline__ void fill_i64_body(int64_t *p, size_t n, const int lb, const int gb) {
    size_t i = (size_t)lb * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

__device__ __forceinline__ void fill_u32_body(uint32_t *p, size_t n, uint32_t v, const int lb, const int gb) {
    size_t i = (size_t)lb * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// a += e; e = 0, over the words of the labels in use.
__device__ __forceinline__ void consolidate_body(int64_t *a, int64_t *e, const uint32_t *n_labels, int cw, const int lb, const int gb) {
    size_t n = (size_t)(*n_labels) * cw;
    for (size_t i = (size_t)lb * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gb * blockDim.x) {
        a[i] += e[i];
        e[i] = 0;
    }
}

// Lowest site index of every label.
__device__ __forceinline__ void first_site_body(const CellArgs &a, size_t n_sites, const int lb, const int gb) {
    size_t i = (size_t)lb * blockDim.x + threadIdx.x;
    if (i >= n_sites) return;
    atomicMin(a.scratch_u32 + label_id(a.labels[i]), (uint32_t)i);
}

__device__ __forceinline__ void anchor_from_first_body(const CellArgs &a, const int lb, const int gb) {
    uint32_t l = lb * blockDim.x + threadIdx.x;
    if (l >= *a.n_labels) return;
    uint32_t f = a.scratch_u32[l];
    int x = 0, y = 0, z = 0;
    if (f != 0xFFFFFFFFu) coords(f, a.w, a.h, &x, &y, &z);
    a.anchor[3 * l] = x;
    a.anchor[3 * l + 1] = y;
    a.anchor[3 * l + 2] = z;
}

// Recount of every label from the lattice into a zeroed state (or, with
// moments_only, of the moment words alone). A workgroup covers kRecountRun
// lattice rows (rows of the flattened (z, y) index) of 256 columns; each lane
// walks its column down those rows, so every load of the workgroup reads
// consecutive sites, and sums the words of a run of equal labels in
// registers (a tiled sheet's cells are taller than a run). Only where its
// label changes does a run's total enter the workgroup's LDS hash table,
// with LDS atomics, and each (label, word) of the table then costs one
// global atomic. A label that finds no free slot goes straight to global
// atomics. Integer sums, so the order does not matter. `gate` >= 0 makes
// the launch a no-op unless flags[gate] is set, so a whole step can be
// enqueued without host checks.
constexpr int kSlots = 256;

__device__ __forceinline__ void recount_body(const CellArgs &a, size_t n_sites, int moments_only, int gate,
                                             const int lb, const int gb) {
    if (gate >= 0 && a.flags[gate] == 0) return;
    __shared__ uint32_t keys[kSlots];
    __shared__ unsigned long long vals[kSlots * 12];
    const int cw = a.cw;
    for (int s = threadIdx.x; s < kSlots; s += blockDim.x) keys[s] = 0xFFFFFFFFu;
    for (int s = threadIdx.x; s < kSlots * cw; s += blockDim.x) vals[s] = 0;
    __syncthreads();
    auto flush = [&](uint32_t l, const int64_t *c) {
        uint32_t slot = hash32(l) & (kSlots - 1);
        int found = -1;
        for (int probe = 0; probe < kSlots; ++probe) {
            uint32_t prev = atomicCAS(&keys[slot], 0xFFFFFFFFu, l);
            if (prev == 0xFFFFFFFFu || prev == l) {
                found = (int)slot;
                break;
            }
            slot = (slot + 1) & (kSlots - 1);
        }
#pragma unroll
        for (int q = 0; q < 12; ++q) {
            if (q >= cw || c[q] == 0) continue;
            if (found >= 0)
                atomicAdd(&vals[found * cw + q], (unsigned long long)c[q]);
            else
                atomicAdd(reinterpret_cast<unsigned long long *>(a.state + (size_t)l * cw + q),
                          (unsigned long long)c[q]);
        }
    };
    const int col_blocks = (a.w + kRecountCols - 1) / kRecountCols;
    const int x0 = (lb % col_blocks) * kRecountCols;
    const int x = x0 + (int)threadIdx.x;
    const uint32_t rows = (uint32_t)a.h * (uint32_t)a.d;
    const uint32_t r0 = (uint32_t)(lb / col_blocks) * kRecountRun;
    // In a plane the bonds come from an LDS copy of the workgroup's rows
    // with one row and one column of halo each side (periodic), loaded with
    // consecutive sites per instruction, instead of eight global reads per
    // site. Slot (j, c) holds row r0 - 1 + j, column x0 - 1 + c.
    __shared__ uint32_t rtile[(kRecountRun + 2) * (kRecountCols + 2)];
    constexpr int kRP = kRecountCols + 2;
    const bool plane_bonds = !moments_only && a.d == 1;
    const int ncols = a.w - x0 < kRecountCols ? a.w - x0 : kRecountCols;
    if (plane_bonds) {
        // lanes 0 .. ncols + 1 each load one tile column: the left halo, the
        // workgroup's columns, the right halo
        const int c = (int)threadIdx.x;
        // the first row held (r0 - 1, wrapped), then a step at a time (a
        // plane can be shorter than the rows held, so the wrap may repeat)
        const int gy0 = (int)((r0 + (uint32_t)a.h - 1u) % (uint32_t)a.h);
        for (int cc = c; cc < ncols + 2; cc += kRecountCols) {
            const int gx = wrap1(x0 - 1 + cc, a.w);
            uint32_t v[kRecountRun + 2];
            int gy = gy0;
#pragma unroll
            for (int j = 0; j < kRecountRun + 2; ++j) {
                v[j] = a.labels[(size_t)gy * (size_t)a.w + (size_t)gx];
                gy = gy + 1 == a.h ? 0 : gy + 1;
            }
#pragma unroll
            for (int j = 0; j < kRecountRun + 2; ++j) rtile[j * kRP + cc] = v[j];
        }
        __syncthreads();
    }
    if (x < a.w) {
        int64_t acc[12];
#pragma unroll
        for (int q = 0; q < 12; ++q) acc[q] = 0;
        uint32_t run_label = 0xFFFFFFFFu;
        // (y, z) of the first row, then stepped
        int y = (int)(r0 % (uint32_t)a.h), z = (int)(r0 / (uint32_t)a.h);
        for (uint32_t r = r0; r < r0 + kRecountRun && r < rows; ++r) {
            const size_t i = (size_t)r * (size_t)a.w + (size_t)x;
            uint32_t word = plane_bonds ? rtile[(int)(r - r0 + 1) * kRP + (int)threadIdx.x + 1] : a.labels[i];
            uint32_t l = label_id(word);
            if (l != run_label) {
                if (run_label != 0xFFFFFFFFu) flush(run_label, acc);
#pragma unroll
                for (int q = 0; q < 12; ++q) acc[q] = 0;
                run_label = l;
            }
            if (plane_bonds) {
                const int j = (int)(r - r0) + 1, c = (int)threadIdx.x + 1;
                int bonds = 0;
                for (int k = 0; k < a.n_off; ++k)
                    if (rtile[(j + a.off[k][1]) * kRP + c + a.off[k][0]] != word) ++bonds;
                acc[0] += pack_vs(1, bonds);
            } else if (!moments_only) {
                // neighbour offsets are within one site on every axis
                int bonds = 0;
                for (int k = 0; k < a.n_off; ++k) {
                    const int nx = wrap1(x + a.off[k][0], a.w), ny = wrap1(y + a.off[k][1], a.h),
                              nz = wrap1(z + a.off[k][2], a.d);
                    const size_t nb = ((size_t)nz * a.h + ny) * (size_t)a.w + nx;
                    if (nb != i && a.labels[nb] != word) ++bonds;
                }
                acc[0] += pack_vs(1, bonds);
            }
            if (cw > 1 && l != 0) {
                int64_t ux = fold_int(x - a.anchor[3 * l], a.w);
                int64_t uy = fold_int(y - a.anchor[3 * l + 1], a.h);
                int64_t uz = fold_int(z - a.anchor[3 * l + 2], a.d);
                acc[1] += ux;
                acc[2] += uy;
                acc[3] += uz;
                acc[4] += ux * ux;
                acc[5] += uy * uy;
                acc[6] += uz * uz;
                acc[7] += ux * uy;
                acc[8] += ux * uz;
                acc[9] += uy * uz;
                if (cw > 10) {
                    acc[10] += a.nematic[2 * i];
                    acc[11] += a.nematic[2 * i + 1];
                }
            }
            if (++y == a.h) {
                y = 0;
                ++z;
            }
        }
        if (run_label != 0xFFFFFFFFu) flush(run_label, acc);
    }
    __syncthreads();
    for (int s = threadIdx.x; s < kSlots * cw; s += blockDim.x) {
        uint32_t l = keys[s / cw];
        if (l == 0xFFFFFFFFu || vals[s] == 0) continue;
        atomicAdd(reinterpret_cast<unsigned long long *>(a.state + (size_t)l * cw + s % cw), vals[s]);
    }
}

// Zero the moment words (1..9) of every label, keeping volume and surface.
__device__ __forceinline__ void zero_moments_body(const CellArgs &a, const int lb, const int gb) {
    uint32_t l = lb * blockDim.x + threadIdx.x;
    if (l >= *a.n_labels) return;
    for (int q = 1; q < a.cw; ++q) a.state[(size_t)l * a.cw + q] = 0;
}

// Move each live label's anchor to its rounded centroid, then clear its sums
// for a moments-only recount. Gated on flags[gate] when gate >= 0.
__device__ __forceinline__ void recentre_body(const CellArgs &a, int gate, const int lb, const int gb) {
    if (gate >= 0 && a.flags[gate] == 0) return;
    uint32_t l = lb * blockDim.x + threadIdx.x;
    if (l >= *a.n_labels) return;
    int64_t *s = a.state + (size_t)l * a.cw;
    int64_t v = unpack_vol(s[0]);
    if (l != 0 && v > 0) {
        const int span[3] = {a.w, a.h, a.d};
        for (int ax = 0; ax < 3; ++ax) {
            double mean = (double)s[1 + ax] / (double)v;
            int64_t shift = (int64_t)floor(mean + 0.5);
            int64_t nv = (int64_t)a.anchor[3 * l + ax] + shift;
            nv %= span[ax];
            if (nv < 0) nv += span[ax];
            a.anchor[3 * l + ax] = (int32_t)nv;
        }
    }
    for (int q = 1; q < a.cw; ++q) s[q] = 0;
}

// Whether a live label's centroid (first sums `s`, volume n > 0) sits more
// than an eighth of a side from its anchor.
__device__ __forceinline__ bool drifted(const CellArgs &a, const int64_t *s, int64_t n) {
    const int span[3] = {a.w, a.h, a.d};
    for (int ax = 0; ax < 3; ++ax) {
        if (span[ax] <= 1) continue;
        int64_t m = s[1 + ax];
        int64_t am = m < 0 ? -m : m;
        if (8 * am > (int64_t)span[ax] * n) return true;
    }
    return false;
}

// Sets flags[1] when some live label has drifted from its anchor.
__device__ __forceinline__ void check_recentre_body(const CellArgs &a, const int lb, const int gb) {
    uint32_t l = lb * blockDim.x + threadIdx.x;
    if (l == 0 || l >= *a.n_labels) return;
    const int64_t *s = a.state + (size_t)l * a.cw;
    int64_t n = unpack_vol(s[0]);
    if (n > 0 && drifted(a, s, n)) a.flags[1] = 1;
}

// ---------------------------------------------------------------- population

// Decide each live label's fate: divide at its division volume (cut through
// the centroid across the major axis), else die with the type's rate by a
// counter draw of (step, label). Sets flags[0] when anything happens, and
// flags[1] when some label has drifted from its anchor (the recentring that
// follows the step's division and death).
// With `e`, the label's record is brought current first (a += e, e = 0:
// the consolidation of the step's last pass, one label per thread), and the
// first threads clear the next step's flag words (word 2 is kept for the
// host).
__device__ __forceinline__ void plan_body(const CellArgs &a, int64_t *e, const int lb, const int gb) {
    uint32_t l = lb * blockDim.x + threadIdx.x;
    if (l < 4 && l != 2 && a.flags_next) a.flags_next[l] = 0u;
    // two rounds of loads: the label count with this label's type, then
    // (for a live id) its record, the last pass's deltas and the type's
    // division volume and death rate together; the record is brought
    // current in registers
    if (l >= a.capacity) return;
    const uint32_t n = *a.n_labels;
    const uint32_t tw = a.ltype[l];
    if (l >= n) return;
    constexpr int kWords = 12;  // the widest record: volume/surface, nine moments, two field sums
    const size_t base = (size_t)l * a.cw;
    int64_t rec[kWords], del[kWords];
#pragma unroll
    for (int q = 0; q < kWords; ++q) {
        rec[q] = 0;
        del[q] = 0;
        if (q < a.cw) {
            rec[q] = a.state[base + q];
            if (e) del[q] = e[base + q];
        }
    }
    const uint32_t t = label_type(tw);
    const DeviceTables *tab = a.tables;
    const double div_vol = tab->div_vol[t], death = tab->death[t];
#pragma unroll
    for (int q = 0; q < kWords; ++q) rec[q] += del[q];
    if (e) {
#pragma unroll
        for (int q = 0; q < kWords; ++q) {
            if (q < a.cw) {
                a.state[base + q] = rec[q];
                e[base + q] = 0;
            }
        }
    }
    a.plan[l] = 0;
    if (l == 0) return;
    int64_t v = unpack_vol(rec[0]);
    if (v <= 0) return;
    if (drifted(a, rec, v)) a.flags[1] = 1;
    if (div_vol > 0.0 && (double)v >= div_vol && v >= 4) {
        a.plan[l] = 1;
        double *g = a.geo + 6 * (size_t)l;
        for (int ax = 0; ax < 3; ++ax) g[ax] = (double)rec[1 + ax] / (double)v;
        cov_largest(rec + 1, v, a.d > 1, g + 3);
        a.flags[0] = 1;
    } else if (death > 0.0) {
        uint32_t hsh = hash32(hash32(l ^ a.key) ^ a.key2);
        if ((double)hsh * (1.0 / 4294967296.0) < death) {
            a.plan[l] = 2;
            a.flags[0] = 1;
        }
    }
}

__device__ __forceinline__ bool beyond_cut(const CellArgs &a, uint32_t l, int x, int y, int z) {
    double ux = (double)fold_int(x - a.anchor[3 * l], a.w);
    double uy = (double)fold_int(y - a.anchor[3 * l + 1], a.h);
    double uz = (double)fold_int(z - a.anchor[3 * l + 2], a.d);
    const double *g = a.geo + 6 * (size_t)l;
    double along = (ux - g[0]) * g[3] + (uy - g[1]) * g[4] + (uz - g[2]) * g[5];
    return along > 0.0;
}

// Sites each planned division would move.
__device__ __forceinline__ void count_moved_body(const CellArgs &a, size_t n_sites, const int lb, const int gb) {
    if (a.flags[0] == 0) return;
    size_t i = (size_t)lb * blockDim.x + threadIdx.x;
    if (i >= n_sites) return;
    uint32_t l = label_id(a.labels[i]);
    if (l == 0 || a.plan[l] != 1) return;
    int x, y, z;
    coords(i, a.w, a.h, &x, &y, &z);
    if (beyond_cut(a, l, x, y, z)) atomicAdd(a.scratch_u32 + l, 1u);
}

// One workgroup: cancel divisions that would move nothing or everything,
// hand out daughter ids in ascending parent order, and copy types and
// anchors to the daughters.
// Rank of this lane's flag among the workgroup's set flags (lanes in
// order), and their count: a ballot per wavefront and one barrier-bounded
// pass over the wavefronts' counts, instead of a scan over every lane. The
// workgroup is a whole number of wavefronts, at most 16.
__device__ __forceinline__ uint32_t block_rank(bool flag, uint32_t *wave_counts, uint32_t *total) {
    const unsigned lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const unsigned long long m = __ballot(flag);
    const uint32_t below = (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) wave_counts[wave] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t before = 0, all = 0;
    const unsigned waves = blockDim.x >> 6;
    for (unsigned w = 0; w < waves; ++w) {
        const uint32_t c = wave_counts[w];
        before += w < wave ? c : 0u;
        all += c;
    }
    __syncthreads();
    *total = all;
    return before + below;
}

// Tail of the daughter assignment (same workgroup, after its loop): when
// more than max(64, (n-1)/4) of the n labels are dead once the step's deaths
// and divisions are applied, renumber the live ones densely in ascending
// order (remap into scratch_u32, whose moved-site counts are spent) and set
// flags[3] to the new count. Live: a daughter (id >= n0), a dividing label,
// and any other label with volume left and no death planned.
__device__ __forceinline__ void plan_compaction(const CellArgs &a, uint32_t n0, uint32_t n, uint32_t *wave_counts) {
    __shared__ uint32_t dead_total;
    if (threadIdx.x == 0) dead_total = 0;
    __syncthreads();
    auto live = [&](uint32_t l) {
        if (l >= n0) return true;
        const uint8_t p = a.plan[l];
        return p == 1 || (p == 0 && unpack_vol(a.state[(size_t)l * a.cw]) > 0);
    };
    for (uint32_t l = 1 + threadIdx.x; l < n; l += blockDim.x)
        if (!live(l)) atomicAdd(&dead_total, 1u);
    __syncthreads();
    uint32_t limit = (n - 1) / 4;
    if (limit < 64) limit = 64;
    if (dead_total <= limit) return;
    uint32_t next = 1;
    for (uint32_t start = 1; start < n; start += blockDim.x) {
        uint32_t l = start + threadIdx.x;
        const bool keep = l < n && live(l);
        uint32_t lived = 0;
        const uint32_t rank = block_rank(keep, wave_counts, &lived);
        if (l < n) a.scratch_u32[l] = keep ? next + rank : 0u;
        next += lived;
    }
    if (threadIdx.x == 0) {
        a.scratch_u32[0] = 0;
        a.flags[3] = next; // the new label count, nonzero
    }
}

__device__ __forceinline__ void assign_daughters_body(const CellArgs &a, const int lb, const int gb) {
    if (a.flags[0] == 0) return;
    __shared__ uint32_t wave_counts[16];
    const uint32_t n0 = *a.n_labels;
    uint32_t base = n0;
    for (uint32_t start = 0; start < n0; start += blockDim.x) {
        uint32_t l = start + threadIdx.x;
        uint32_t take = 0;
        if (l < n0 && l != 0 && a.plan[l] == 1) {
            int64_t v = unpack_vol(a.state[(size_t)l * a.cw]);
            uint32_t moved = a.scratch_u32[l];
            if (moved == 0 || (int64_t)moved == v)
                a.plan[l] = 0;
            else
                take = 1;
        }
        uint32_t taken = 0;
        const uint32_t rank = block_rank(take != 0, wave_counts, &taken);
        if (take) {
            uint32_t id = base + rank;
            if (id >= a.capacity) {
                a.plan[l] = 0;
                a.flags[2] = 1;
            } else {
                a.ltype[id] = (a.ltype[l] & ~kIdMask) | id; // the daughter keeps the type
                a.daughter[l] = id;
                a.plan[id] = 0;
                a.anchor[3 * id] = a.anchor[3 * l];
                a.anchor[3 * id + 1] = a.anchor[3 * l + 1];
                a.anchor[3 * id + 2] = a.anchor[3 * l + 2];
                // a fresh id: its record is recounted from its sites, and no
                // pass delta left from an earlier owner of the id survives
                for (int q = 0; q < a.cw; ++q) {
                    a.state[(size_t)id * a.cw + q] = 0;
                    for (int b = 0; b < 3; ++b) a.deltas[b][(size_t)id * a.cw + q] = 0;
                }
            }
        }
        // a dying label's record clears; a dividing one is recounted
        if (l < n0 && l != 0 && a.plan[l] != 0)
            for (int q = 0; q < a.cw; ++q) a.state[(size_t)l * a.cw + q] = 0;
        base += taken;
    }
    const uint32_t n = base < a.capacity ? base : a.capacity;
    if (threadIdx.x == 0) *a.n_labels = n;
    __syncthreads();  // every moved-site count is read; the plans are final
    plan_compaction(a, n0, n, wave_counts);
}

// A site's label once the step's division and death are applied: medium for
// a dying label, the daughter beyond a dividing label's cut. Applying it to
// a label already applied changes nothing (the medium and a daughter plan
// nothing), so it reads the same whether or not the site was rewritten yet.
__device__ __forceinline__ uint32_t label_after(const CellArgs &a, uint32_t l, int x, int y, int z) {
    if (l == 0) return 0;
    const uint8_t p = a.plan[l];
    if (p == 2) return 0;
    if (p == 1 && beyond_cut(a, l, x, y, z)) return a.daughter[l];
    return l;
}

// Relabel, and recount the labels the step touched: a dying label's sites
// become medium (its record was cleared and nobody else's changes: a
// neighbour's bond to such a site is unlike before and after), a dividing
// label's sites beyond the cut go to its daughter, and every site of the
// parent or the daughter adds its volume, unlike bonds (against the
// neighbours' labels after the step), moment sums about the shared anchor
// and field sums to the record it now belongs to.
__device__ __forceinline__ void apply_population_body(const CellArgs &a, size_t n_sites, const int lb, const int gb) {
    if (a.flags[0] == 0) return;
    size_t i = (size_t)lb * blockDim.x + threadIdx.x;
    if (i >= n_sites) return;
    uint32_t word = a.labels[i];
    uint32_t l = label_id(word);
    if (l == 0) return;
    uint8_t p = a.plan[l];
    if (p == 0) return;
    if (p == 2) {
        a.labels[i] = 0;
        return;
    }
    int x, y, z;
    coords(i, a.w, a.h, &x, &y, &z);
    const uint32_t nl = beyond_cut(a, l, x, y, z) ? a.daughter[l] : l;
    if (nl != l) a.labels[i] = (word & ~kIdMask) | nl;
    int bonds = 0;
    for (int k = 0; k < a.n_off; ++k) {
        const int nx = wrap1(x + a.off[k][0], a.w), ny = wrap1(y + a.off[k][1], a.h),
                  nz = wrap1(z + a.off[k][2], a.d);
        const size_t nb = ((size_t)nz * a.h + ny) * (size_t)a.w + nx;
        if (nb != i && label_after(a, label_id(a.labels[nb]), nx, ny, nz) != nl) ++bonds;
    }
    unsigned long long *rec = reinterpret_cast<unsigned long long *>(a.state + (size_t)nl * a.cw);
    atomicAdd(rec, (unsigned long long)pack_vs(1, bonds));
    if (a.cw > 1) {
        const int64_t ux = fold_int(x - a.anchor[3 * nl], a.w);
        const int64_t uy = fold_int(y - a.anchor[3 * nl + 1], a.h);
        const int64_t uz = fold_int(z - a.anchor[3 * nl + 2], a.d);
        const int64_t c[9] = {ux, uy, uz, ux * ux, uy * uy, uz * uz, ux * uy, ux * uz, uy * uz};
#pragma unroll
        for (int q = 0; q < 9; ++q)
            if (c[q] != 0) atomicAdd(rec + 1 + q, (unsigned long long)c[q]);
        if (a.cw > 10) {
            atomicAdd(rec + 10, (unsigned long long)(int64_t)a.nematic[2 * i]);
            atomicAdd(rec + 11, (unsigned long long)(int64_t)a.nematic[2 * i + 1]);
        }
    }
}

// Zero the per-label state ahead of a full recount, gated like recount.
__device__ __forceinline__ void zero_state_body(const CellArgs &a, int gate, const int lb, const int gb) {
    if (gate >= 0 && a.flags[gate] == 0) return;
    size_t n = (size_t)(*a.n_labels) * a.cw;
    for (size_t i = (size_t)lb * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gb * blockDim.x)
        a.state[i] = 0;
}

// Move types and anchors of live labels to their new ids (ascending, so a
// label only ever moves down and no source is overwritten before it is read
// when done in one thread; the parallel form reads from a staged copy).
__device__ __forceinline__ void compact_move_body(const CellArgs &a, const uint32_t *type_copy, const int32_t *anchor_copy, const int lb, const int gb) {
    if (a.flags[0] == 0 || a.flags[3] == 0) return;
    uint32_t l = lb * blockDim.x + threadIdx.x;
    uint32_t n = *a.n_labels;
    if (l == 0 || l >= n) return;
    uint32_t to = a.scratch_u32[l];
    if (to == 0) return;
    a.ltype[to] = (type_copy[l] & ~kIdMask) | to;
    a.anchor[3 * to] = anchor_copy[3 * l];
    a.anchor[3 * to + 1] = anchor_copy[3 * l + 1];
    a.anchor[3 * to + 2] = anchor_copy[3 * l + 2];
    for (int q = 0; q < a.cw; ++q) a.state[(size_t)to * a.cw + q] = a.state_copy[(size_t)l * a.cw + q];
}

// Relabels the sites with the compaction's remap; the first thread also sets
// the new label count (nothing in this launch reads it).
__device__ __forceinline__ void compact_relabel_body(const CellArgs &a, size_t n_sites, const int lb, const int gb) {
    if (a.flags[0] == 0 || a.flags[3] == 0) return;
    size_t i = (size_t)lb * blockDim.x + threadIdx.x;
    if (i == 0) *a.n_labels = a.flags[3];
    if (i >= n_sites) return;
    uint32_t word = a.labels[i];
    uint32_t l = label_id(word);
    if (l == 0) return;
    a.labels[i] = (word & ~kIdMask) | a.scratch_u32[l];
}

// Length minus target of every label in the snapshot a + e: the length
// term's `before` (sweep_body.h), once per label instead of per attempt.
__device__ __forceinline__ void length_prepass_body(const int64_t *a, const int64_t *e, const uint32_t *n_labels,
                                                    int cw, const uint32_t *ltype, const DeviceTables *tables,
                                                    bool d3, double *out, const int lb, const int gb) {
    const uint32_t l = lb * blockDim.x + threadIdx.x;
    if (l == 0 || l >= *n_labels) return;
    const size_t base = (size_t)l * cw;
    int64_t m[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) m[q] = a[base + 1 + q] + e[base + 1 + q];
    const int64_t n = unpack_vol(a[base] + e[base]);
    const uint32_t ty = label_type(ltype[l]);
    out[l] = length_of(m, n, d3) - tables->tgt_len[ty];
}

// Staging copies for a compaction: types, anchors and records of the labels
// in use, read back by compact_move_body.
__device__ __forceinline__ void compact_stage_body(const CellArgs &a, uint32_t *type_copy, int32_t *anchor_copy,
                                                   const int lb, const int gb) {
    if (a.flags[0] == 0 || a.flags[3] == 0) return;
    const size_t n = *a.n_labels, words = n * (size_t)a.cw, span = words > 3 * n ? words : 3 * n;
    for (size_t i = (size_t)lb * blockDim.x + threadIdx.x; i < span; i += (size_t)gb * blockDim.x) {
        if (i < words) a.state_copy[i] = a.state[i];
        if (i < n) type_copy[i] = a.ltype[i];
        if (i < 3 * n) anchor_copy[i] = a.anchor[i];
    }
}

__device__ __forceinline__ void copy_u32_body(uint32_t *dst, const uint32_t *src, const uint32_t *n_labels, int per, const int lb, const int gb) {
    size_t n = (size_t)(*n_labels) * per;
    for (size_t i = (size_t)lb * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gb * blockDim.x)
        dst[i] = src[i];
}

}  // namespace ghip
