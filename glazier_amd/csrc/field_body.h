// field_body.h - the chemical-field step on gfx950: explicit-Euler diffusion
// with decay, temporally blocked in LDS, with secretion and uptake fused into
// the last substep.
//
// Device counterpart of field_diffuse + field_exchange
// (glazier-cuda/src/lib.rs:314-363, launched per species per substep at
// lib.rs:1147-1187) and of Fields::diffuse/exchange
// (glazier-core/src/field.rs:99-149).
//
// A workgroup loads an interior tile plus a halo as deep as the substeps it
// will take, runs those substeps in LDS (each one shrinks the valid region by
// one site), and writes the interior once. Every site sees the same f32
// operations in the same order as one substep per launch would give:
//
//     sum  = u[x+1] + u[x-1];  sum = sum + u[y+1];  sum = sum + u[y-1]
//     (3D: sum = sum + u[z+1];  sum = sum + u[z-1])
//     u'   = (u + d * (sum - faces * u)) - k * u
//
// then, on owned sites after the species' last substep:
//     v = u + sec[type];  v = v - up[type] * v;  u = max(v, 0)
#pragma once

#include "ghip_device.h"

namespace ghip {

// One workgroup's tile (block `b` of the launch over this lattice and species).
// A lane takes whole columns (x, y) of the tile and walks them along z, so a
// site's LDS index is its column's plus a multiple of the plane (no division
// per site), and the z neighbours of the walk slide through registers.
template <bool D3>
__device__ __forceinline__ void field_body(const FieldArgs &a, const int b) {
    extern __shared__ float fl[];
    const int IX = 1 << a.ix_log2, IY = 1 << a.iy_log2, IZ = D3 ? a.iz : 1;
    const int K = a.steps;
    const int hz = D3 ? K : 0;
    const int LX = IX + 2 * K, LY = IY + 2 * K, LZ = IZ + 2 * hz;
    const int plane = LX * LY;
    const int total = plane * LZ;
    float *u = fl;
    float *v = fl + total;
    const int tx = b % a.tiles_x, ty = (b / a.tiles_x) % a.tiles_y, tz = D3 ? b / (a.tiles_x * a.tiles_y) : 0;
    const int x0 = tx * IX - K, y0 = ty * IY - K, z0 = D3 ? tz * IZ - hz : 0;
    const size_t wh = (size_t)a.w * (size_t)a.h;

    for (int c = threadIdx.x; c < plane; c += blockDim.x) {
        const int ly = c / LX, lx = c - ly * LX;
        const float *col = a.in + (size_t)wrapc(y0 + ly, a.h) * (size_t)a.w + (size_t)wrapc(x0 + lx, a.w);
        int gz = D3 ? wrapc(z0, a.d) : 0;  // stepped with the walk, wrapping at the lattice's depth
#pragma unroll 4
        for (int lz = 0; lz < LZ; ++lz) {
            u[lz * plane + c] = col[(size_t)gz * wh];
            if (D3 && ++gz == a.d) gz = 0;
        }
    }
    __syncthreads();

    const float faces = D3 ? 6.0f : 4.0f;
    for (int s = 1; s <= K; ++s) {
        // valid after s substeps: [s, L - s) on each blocked axis
        const int zlo = D3 ? s : 0, zhi = D3 ? LZ - s : 1;
        for (int c = threadIdx.x; c < plane; c += blockDim.x) {
            const int ly = c / LX, lx = c - ly * LX;
            if (lx < s || lx >= LX - s || ly < s || ly >= LY - s) continue;
            int q = zlo * plane + c;
            float here = u[q];
            float below = D3 ? u[q - plane] : 0.0f;
            for (int lz = zlo; lz < zhi; ++lz, q += plane) {
                float sum = u[q + 1] + u[q - 1];
                sum = sum + u[q + LX];
                sum = sum + u[q - LX];
                float above = 0.0f;
                if (D3) {
                    above = u[q + plane];
                    sum = sum + above;
                    sum = sum + below;
                }
                v[q] = (here + a.dcoef * (sum - faces * here)) - a.kcoef * here;
                below = here;
                here = above;
            }
        }
        __syncthreads();
        float *t = u;
        u = v;
        v = t;
    }

    const int interior = IX * IY;
    for (int c = threadIdx.x; c < interior; c += blockDim.x) {
        const int ix = c & (IX - 1), iy = c >> a.ix_log2;
        const int gx = tx * IX + ix, gy = ty * IY + iy;
        if (gx >= a.w || gy >= a.h) continue;
        for (int iz = 0; iz < IZ; ++iz) {
            const int gz = D3 ? tz * IZ + iz : 0;
            if (D3 && gz >= a.d) break;
            const size_t site = (size_t)gz * wh + (size_t)gy * a.w + gx;
            float val = u[((iz + hz) * LY + iy + K) * LX + ix + K];
            if (a.exchange) {
                uint32_t word = a.labels[site];
                if (label_id(word) != 0) {
                    uint32_t t = label_type(word);
                    float w2 = val + a.tables->sec[t][a.species];
                    w2 = w2 - a.tables->up[t][a.species] * w2;
                    val = w2 < 0.0f ? 0.0f : w2;
                }
            }
            a.out[site] = val;
        }
    }
}

// The same step on a plane with the 64 x 32 interior tile and the substep
// count fixed at compile time, so every LDS coordinate divides by a constant;
// the lattice is at least one tile on each side, so the halo wraps with one
// compare. Same operations in the same order as field_body, so the same bits.

template <int K>
__device__ __forceinline__ void field_body_2d(const FieldArgs &a, const int b, float *u, float *v) {
    constexpr int LX = kFieldIX + 2 * K, LY = kFieldIY + 2 * K, total = LX * LY;
    const int tx = b % a.tiles_x, ty = b / a.tiles_x;
    const int x0 = tx * kFieldIX - K, y0 = ty * kFieldIY - K;

    for (int q = threadIdx.x; q < total; q += blockDim.x) {
        const int lx = q % LX, ly = q / LX;
        int gx = x0 + lx, gy = y0 + ly;
        gx += gx < 0 ? a.w : 0;
        gx -= gx >= a.w ? a.w : 0;
        gy += gy < 0 ? a.h : 0;
        gy -= gy >= a.h ? a.h : 0;
        u[q] = a.in[(size_t)gy * a.w + gx];
    }
    __syncthreads();

#pragma unroll
    for (int s = 1; s <= K; ++s) {
        for (int q = threadIdx.x; q < total; q += blockDim.x) {
            const int lx = q % LX, ly = q / LX;
            if (lx < s || lx >= LX - s || ly < s || ly >= LY - s) continue;
            const float here = u[q];
            float sum = u[q + 1] + u[q - 1];
            sum = sum + u[q + LX];
            sum = sum + u[q - LX];
            v[q] = (here + a.dcoef * (sum - 4.0f * here)) - a.kcoef * here;
        }
        __syncthreads();
        float *t = u;
        u = v;
        v = t;
    }

    for (int q = threadIdx.x; q < kFieldIX * kFieldIY; q += blockDim.x) {
        const int ix = q % kFieldIX, iy = q / kFieldIX;
        const int gx = tx * kFieldIX + ix, gy = ty * kFieldIY + iy;
        if (gx >= a.w || gy >= a.h) continue;
        const size_t site = (size_t)gy * a.w + gx;
        float val = u[(iy + K) * LX + ix + K];
        if (a.exchange) {
            const uint32_t word = a.labels[site];
            if (label_id(word) != 0) {
                const uint32_t t = label_type(word);
                float w2 = val + a.tables->sec[t][a.species];
                w2 = w2 - a.tables->up[t][a.species] * w2;
                val = w2 < 0.0f ? 0.0f : w2;
            }
        }
        a.out[site] = val;
    }
}

}  // namespace ghip
