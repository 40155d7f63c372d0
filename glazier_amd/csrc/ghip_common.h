// ghip_common.h - arithmetic shared by the host driver and the gfx950
// kernels of the Cellular Potts engine: the counter hash, the draw layout,
// the f32 exponential and logarithm, and the periodic folds.
//
// Every function here is a fixed sequence of IEEE f32/f64 or integer
// operations (the library is built with -ffp-contract=off), so a kernel and
// the host code that checks it round identically.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define GHD __host__ __device__ __forceinline__

namespace ghip {

// A 32-bit integer finaliser: two xor-shift-multiply rounds (Wellons,
// "hash-prospector", lowbias32). A bijection on 32 bits.
GHD uint32_t hash32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
}

// Key of one colour pass, folding seed, replica, step and colour in turn.
// The driver computes it once per launch; the kernel only mixes in the site.
GHD uint32_t pass_key(uint64_t seed, uint32_t replica, uint64_t step, uint32_t colour) {
    uint32_t k = hash32((uint32_t)seed ^ 0x243F6A88U);
    k = hash32(k ^ (uint32_t)(seed >> 32));
    k = hash32(k ^ replica);
    k = hash32(k ^ (uint32_t)step);
    k = hash32(k ^ (uint32_t)(step >> 32));
    k = hash32(k ^ colour);
    return k;
}

GHD uint32_t second_key(uint32_t key) { return hash32(key ^ 0x85EBCA6BU); }

// The draws of one target in a pass, two rounds of the hash. The first,
// p = site_pick(key, site), is a bijection of the site for a fixed pass (no
// two targets of a pass share it); its low bits pick the neighbour when the
// neighbour count is a power of two. The second, h = site_draw(key2, p),
// holds the acceptance uniform in its top 24 bits (and the pick for other
// neighbour counts). Asking only which neighbour a target picks takes one
// round.
GHD uint32_t site_pick(uint32_t key, uint32_t site) { return hash32(site ^ key); }
GHD uint32_t site_draw(uint32_t key2, uint32_t p) { return hash32(p ^ key2); }

// The pick among eight neighbours, site_pick(key, site) & 7, with the second
// multiply on 24 bits. The pick reads bits 0-2 and 16-18 of that product,
// and bit k of a product depends only on bits 0..k of its operands, so the
// result is the same. On the device it is one full-rate v_mul_u32_u24 (the
// compiler widens a masked multiply back to a 32-bit one, whose low bits are
// the same, so the instruction is named).
GHD uint32_t site_pick8(uint32_t key, uint32_t site) {
    uint32_t x = site ^ key;
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
#if defined(__HIP_DEVICE_COMPILE__)
    asm("v_mul_u32_u24_e32 %0, %1, %2" : "=v"(x) : "s"(0x846ca68bU & 0xFFFFFFu), "v"(x));
#else
    x = (x & 0xFFFFFFu) * (0x846ca68bU & 0xFFFFFFu);
#endif
    return (x ^ (x >> 16)) & 7u;
}

// Second 24-bit uniform, used when the neighbour count is not a power of two
// and the pick has consumed the whole first draw.
GHD uint32_t second_uniform(uint32_t h) { return hash32(h ^ 0x68E31DA4U) >> 8; }

// Colour order of one step: a Fisher-Yates shuffle driven by a hash chain.
GHD void colour_order(uint64_t seed, uint32_t replica, uint64_t step, int colours, int *order) {
    for (int c = 0; c < colours; ++c) order[c] = c;
    uint32_t h = pass_key(seed, replica, step, 0xC0105EEDU);
    for (int i = colours - 1; i > 0; --i) {
        h = hash32(h + (uint32_t)i);
        int j = (int)(((uint64_t)h * (uint64_t)(i + 1)) >> 32);
        int t = order[i];
        order[i] = order[j];
        order[j] = t;
    }
}

GHD float bits_float(uint32_t u) { return __builtin_bit_cast(float, u); }
GHD uint32_t float_bits(float f) { return __builtin_bit_cast(uint32_t, f); }

// 2^x in f32: x = i + f with i the nearest integer and |f| <= 1/2, then a
// degree-7 Taylor polynomial of e^(f ln 2) in Horner form, scaled by 2^i
// through the exponent field. Zero below -125, so no result is subnormal.
GHD float exp2_f32(float x) {
    if (!(x >= -125.0f)) return 0.0f;
    if (x > 127.0f) x = 127.0f;
    int i = (int)x;
    if ((float)i > x) i -= 1;
    float f = x - (float)i;
    if (f > 0.5f) {
        f = f - 1.0f;
        i += 1;
    }
    float p = 1.5252734e-5f;
    p = p * f + 1.5403530e-4f;
    p = p * f + 1.3333558e-3f;
    p = p * f + 9.6181291e-3f;
    p = p * f + 5.5504109e-2f;
    p = p * f + 2.4022651e-1f;
    p = p * f + 6.9314718e-1f;
    p = p * f + 1.0f;
    if (i > 127) i = 127;
    return p * bits_float((uint32_t)(i + 127) << 23);
}

// log2(x) in f32 for normal x > 0, by the decomposition of Sun's fdlibm log
// routines: x = 2^e (1 + f) with 1 + f in [sqrt(1/2), sqrt(2)), s = f/(2 + f),
// and ln(1 + f) = f - f^2/2 + s (f^2/2 + R(s^2)), R(z) = 2z/3 + 2z^2/5 +
// 2z^3/7 + 2z^4/9 (the atanh series; its first omitted term is below 1e-8
// relative for |s| < 0.172). The leading part f - f^2/2 keeps its top 12
// bits in `hi`, so hi times the 12-bit head of 1/ln 2 is exact and the
// rounding lands on small terms: within 1.5 ulp of log2 wherever the engine
// calls it (tested against the C library).
GHD float log2_f32(float x) {
    if (!(x >= 1.17549435e-38f)) return -__builtin_huge_valf();
    uint32_t b = float_bits(x);
    int e = (int)((b >> 23) & 0xff) - 127;
    float m = bits_float((b & 0x7fffffU) | 0x3f800000U);
    if (m > 1.41421356f) {
        m = m * 0.5f;
        e += 1;
    }
    const float f = m - 1.0f;
    const float s = f / (2.0f + f);
    const float z = s * s;
    float r = 0.22222222f;
    r = r * z + 0.28571429f;
    r = r * z + 0.4f;
    r = r * z + 0.66666667f;
    r = r * z;
    const float hfsq = 0.5f * f * f;
    const float hi = bits_float(float_bits(f - hfsq) & 0xfffff000U);
    const float lo = ((f - hi) - hfsq) + s * (hfsq + r);
    const float inv_ln2_hi = 1.44287109375f, inv_ln2_lo = -1.7605286e-4f;
    return (((lo + hi) * inv_ln2_lo + lo * inv_ln2_hi) + hi * inv_ln2_hi) + (float)e;
}

// The 3 x 3 x 3 cube as 27 bits, bit (dz+1)*9 + (dy+1)*3 + (dx+1): every
// set position with its 26 neighbours. The Moore (Chebyshev) neighbourhood
// is the product of the three axes' {-1, 0, 1}, so the dilation is one shift
// pair per axis, each masked so no bit wraps into the next row or plane.
GHD uint32_t dilate_cube(uint32_t s) {
    constexpr uint32_t kCol0 = 0x1249249u;  // dx = -1
    constexpr uint32_t kCol2 = kCol0 << 2;  // dx = +1
    constexpr uint32_t kRow0 = 0x01C0E07u;  // dy = -1
    constexpr uint32_t kRow2 = kRow0 << 6;  // dy = +1
    constexpr uint32_t kCube = 0x7FFFFFFu;
    s = (s | ((s << 1) & ~kCol0) | ((s >> 1) & ~kCol2)) & kCube;
    s = (s | ((s << 3) & ~kRow0) | ((s >> 3) & ~kRow2)) & kCube;
    return (s | (s << 9) | (s >> 9)) & kCube;
}

// Whether the members of a volume's 26-position Moore ring (bit k for ring
// position k in connectivity.rs:23-29's order) form one 26-connected piece
// (connectivity.rs:43-84): the ring's bits go to the cube's bits around the
// centre (bit 13), and the piece holding the lowest one is grown by
// dilation until it stops, the closure of a walk over the ring's adjacency
// held in registers.
GHD bool cube_one_piece(uint32_t present) {
    if (present == 0u || (present & (present - 1u)) == 0u) return true;
    const uint32_t cube = (present & 0x1FFFu) | ((present & ~0x1FFFu) << 1);
    uint32_t reached = cube & (~cube + 1u);
    for (int round = 0; round < 26; ++round) {
        const uint32_t grown = dilate_cube(reached) & cube;
        if (grown == reached) break;
        reached = grown;
    }
    return reached == cube;
}

// Minimum-image offset of an integer coordinate difference, the integer form
// of Moments::unwrap's fold (glazier-core/src/moments.rs:38-48).
GHD int fold_int(int d, int span) {
    if (span <= 1) return 0;
    if (2 * d > span) return d - span;
    if (-2 * d > span) return d + span;
    return d;
}

// A coordinate taken onto the periodic range [0, span).
GHD int wrapc(int v, int span) {
    v %= span;
    return v < 0 ? v + span : v;
}

// Label words on the device carry the cell type in their top bits, so the
// sweep reads a neighbour's type from the lattice without a gather.
constexpr int kTypeShift = 28;
constexpr uint32_t kIdMask = (1u << kTypeShift) - 1u;
GHD uint32_t label_id(uint32_t word) { return word & kIdMask; }
GHD uint32_t label_type(uint32_t word) { return word >> kTypeShift; }

// Volume and surface of one label share an int64: volume * 2^32 + surface.
// Deltas add exactly in two's complement, so one 64-bit atomic carries both.
GHD int64_t pack_vs(int64_t vol, int64_t surf) { return vol * 4294967296LL + surf; }
GHD int32_t unpack_surf(int64_t p) { return (int32_t)(uint32_t)(uint64_t)p; }
GHD int64_t unpack_vol(int64_t p) { return (p - (int64_t)unpack_surf(p)) >> 32; }

}  // namespace ghip
