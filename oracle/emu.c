/*
 * emu.c - CPU emulation of the device engine's checkerboard schedule.
 *
 * TEST INFRASTRUCTURE ONLY: the bit-exact oracle for the HIP kernels. It is a
 * second, independent statement of the schedule the device runs; nothing in
 * the engine links it.
 *
 * The schedule (the device counterpart of glazier-cuda/src/lib.rs:436-716 and
 * :1055-1144, with the reference's live reads replaced by snapshots):
 *
 *   for each Monte Carlo step s:
 *     colours in an order drawn from (seed, replica, s)
 *     for each colour c: a pass over the sites whose coordinate parities equal
 *       c's bits. Same-colour sites are two apart on every axis, so no pass
 *       reads a site another target of the pass writes, except the motility
 *       term, which reads distance two and so reads a copy taken at pass
 *       start. Per-label volume, surface and moments are read from a
 *       snapshot taken at pass start; accepted copies add to the live values.
 *     then fields (diffusion substeps, then secretion/uptake), activity
 *     decay, and division/death.
 *
 *   Each target takes two rounds of a 32-bit counter hash: p of (pass key,
 *   site), whose low bits pick the neighbour when the neighbour count is a
 *   power of two, and h of (second key, p), whose top 24 bits are the
 *   acceptance uniform (for other neighbour counts h also gives the pick,
 *   and a third round the uniform).
 *
 * Energies are f32 in a fixed order (below); the length term and the
 * division geometry are f64 from exact integer moments. Compiled with
 * -ffp-contract=off so each operation rounds once, as on the device.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define EMU_MAX_THR 1024

/* ---------------------------------------------------------------- hash */

/* A 32-bit integer finaliser: two xor-shift-multiply rounds (Wellons,
 * "hash-prospector", lowbias32). A bijection on 32 bits. */
uint32_t oracle_hash32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
}

/* The key of one colour pass: every part of (seed, replica, step, colour)
 * folded in turn. */
uint32_t oracle_pass_key(uint64_t seed, uint32_t replica, uint64_t step, uint32_t colour) {
    uint32_t k = oracle_hash32((uint32_t)seed ^ 0x243F6A88U);
    k = oracle_hash32(k ^ (uint32_t)(seed >> 32));
    k = oracle_hash32(k ^ replica);
    k = oracle_hash32(k ^ (uint32_t)step);
    k = oracle_hash32(k ^ (uint32_t)(step >> 32));
    k = oracle_hash32(k ^ colour);
    return k;
}

static uint32_t second_key(uint32_t key) { return oracle_hash32(key ^ 0x85EBCA6BU); }

/* The two rounds of one site's draw in one pass: the pick round, and the
 * round after it whose top 24 bits are the acceptance uniform. */
uint32_t oracle_site_pick(uint32_t key, uint32_t site) { return oracle_hash32(site ^ key); }
uint32_t oracle_site_draw(uint32_t key, uint32_t site) {
    return oracle_hash32(oracle_site_pick(key, site) ^ second_key(key));
}

/* Colour order of one step: a Fisher-Yates shuffle driven by one hash chain. */
static void colour_order(uint64_t seed, uint32_t replica, uint64_t step, int colours, int *order) {
    for (int c = 0; c < colours; ++c) order[c] = c;
    uint32_t h = oracle_pass_key(seed, replica, step, 0xC0105EEDU);
    for (int i = colours - 1; i > 0; --i) {
        h = oracle_hash32(h + (uint32_t)i);
        int j = (int)(((uint64_t)h * (uint64_t)(i + 1)) >> 32);
        int t = order[i];
        order[i] = order[j];
        order[j] = t;
    }
}

/* ---------------------------------------------------------------- own exp2/log2 */

static inline float bits_to_float(uint32_t u) {
    float f;
    memcpy(&f, &u, 4);
    return f;
}
static inline uint32_t float_to_bits(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    return u;
}

/* 2^x in f32: x = i + f with i = nearest integer and |f| <= 1/2, then a
 * degree-7 Taylor polynomial of e^(f ln 2) in Horner form, scaled by 2^i
 * through the exponent bits. Zero below -125, so no result is subnormal. */
static float g_exp2(float x) {
    if (!(x >= -125.0f)) return 0.0f;
    if (x > 127.0f) x = 127.0f;
    int i = (int)x;
    if ((float)i > x) i -= 1; /* floor */
    float f = x - (float)i;   /* [0, 1) */
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
    return p * bits_to_float((uint32_t)(i + 127) << 23);
}

/* log2(x) in f32 for normal x > 0 (the engine's log2_f32): x = 2^e (1 + f)
 * with 1 + f in [sqrt(1/2), sqrt(2)), s = f/(2 + f), and, as Sun's fdlibm
 * log routines decompose it, ln(1 + f) = f - f^2/2 + s (f^2/2 + R(s^2)) with
 * R(z) = 2z/3 + 2z^2/5 + 2z^3/7 + 2z^4/9; the head of f - f^2/2 (12 bits)
 * times the 12-bit head of 1/ln 2 is exact. */
static float g_log2(float x) {
    if (!(x >= 1.17549435e-38f)) return -INFINITY;
    uint32_t b = float_to_bits(x);
    int e = (int)((b >> 23) & 0xff) - 127;
    float m = bits_to_float((b & 0x7fffffU) | 0x3f800000U);
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
    const float hi = bits_to_float(float_to_bits(f - hfsq) & 0xfffff000U);
    const float lo = ((f - hi) - hfsq) + s * (hfsq + r);
    const float inv_ln2_hi = 1.44287109375f, inv_ln2_lo = -1.7605286e-4f;
    return (((lo + hi) * inv_ln2_lo + lo * inv_ln2_hi) + hi * inv_ln2_hi) + (float)e;
}

/* The two functions as the oracle's callers see them, for the tests that
 * hold them to the C library. */
float oracle_exp2_f32(float x) { return g_exp2(x); }
float oracle_log2_f32(float x) { return g_log2(x); }

/* ---------------------------------------------------------------- state */

struct emu_sim {
    ghip_model m;
    int64_t w, h, d;
    size_t n;
    int colours;
    int n_off;
    int off[26][3];
    int ring_n;
    int ring[26][3];
    uint32_t *labels;
    uint32_t *prev_labels;
    size_t n_labels, cap;
    uint8_t *ctype;
    int64_t *vol, *surf, *mom; /* mom: 9 words per label */
    int64_t *vsnap, *ssnap, *msnap;
    int32_t *anchor;           /* 3 per label */
    int32_t *nq;               /* nematic field per site, 2 fixed-point values */
    int64_t *nem, *nsnap;      /* nematic sums per label, 2 fixed-point words */
    float *field[GHIP_MAX_SPECIES];
    float *fscratch;
    float *act, *prev_act;
    /* f32 model tables */
    float J[GHIP_MAX_TYPES][GHIP_MAX_TYPES];
    float lam_v[GHIP_MAX_TYPES], tgt_v[GHIP_MAX_TYPES];
    float lam_s[GHIP_MAX_TYPES], tgt_s[GHIP_MAX_TYPES];
    float chi[GHIP_MAX_TYPES][GHIP_MAX_SPECIES];
    float max_act[GHIP_MAX_TYPES], coef_act[GHIP_MAX_TYPES];
    float ext[GHIP_MAX_TYPES][3];
    float sec[GHIP_MAX_TYPES][GHIP_MAX_SPECIES], up[GHIP_MAX_TYPES][GHIP_MAX_SPECIES];
    float fd[GHIP_MAX_SPECIES], fk[GHIP_MAX_SPECIES];
    int substeps[GHIP_MAX_SPECIES];
    int medium_type;
    int integral;
    int n_thr;
    uint32_t thr[EMU_MAX_THR];
    float accept_c;
    int has_conn, has_len, has_chem, has_mot, has_ext, has_pop, has_nem, tracks_moments;
    uint64_t mcs, attempted, accepted;
};

static inline int64_t wrapi(int64_t v, int64_t span) {
    v %= span;
    return v < 0 ? v + span : v;
}

static inline size_t e_index(const emu_sim *e, int64_t x, int64_t y, int64_t z) {
    return (size_t)((wrapi(z, e->d) * e->h + wrapi(y, e->h)) * e->w + wrapi(x, e->w));
}

static inline void e_coords(const emu_sim *e, size_t i, int64_t *x, int64_t *y, int64_t *z) {
    size_t plane = (size_t)(e->w * e->h);
    *z = (int64_t)(i / plane);
    size_t r = i % plane;
    *y = (int64_t)(r / (size_t)e->w);
    *x = (int64_t)(r % (size_t)e->w);
}

/* Type a label prices contact as: cell_type of the label, where the medium
 * keeps the reference's cell_type[0] (glazier-core/src/cpu.rs:203). */
static inline int ctype_of(const emu_sim *e, uint32_t label) {
    return label == 0 ? e->medium_type : e->ctype[label];
}

/* Minimum-image offset of an integer coordinate from an anchor, the integer
 * form of Moments::unwrap's fold (glazier-core/src/moments.rs:38-48). */
static inline int64_t fold_int(int64_t d, int64_t span) {
    if (span <= 1) return 0;
    if (2 * d > span) return d - span;
    if (-2 * d > span) return d + span;
    return d;
}

static void ensure_cap(emu_sim *e, size_t want) {
    if (want <= e->cap) return;
    size_t cap = e->cap ? e->cap : 16;
    while (cap < want) cap *= 2;
    e->ctype = realloc(e->ctype, cap);
    e->vol = realloc(e->vol, cap * 8);
    e->surf = realloc(e->surf, cap * 8);
    e->mom = realloc(e->mom, cap * 9 * 8);
    e->vsnap = realloc(e->vsnap, cap * 8);
    e->ssnap = realloc(e->ssnap, cap * 8);
    e->msnap = realloc(e->msnap, cap * 9 * 8);
    e->anchor = realloc(e->anchor, cap * 3 * 4);
    e->nem = realloc(e->nem, cap * 2 * 8);
    e->nsnap = realloc(e->nsnap, cap * 2 * 8);
    e->cap = cap;
}

static int is_integer(double v) { return v == floor(v) && fabs(v) < 1048576.0; }

/* ---------------------------------------------------------------- counters */

static void recount_volume_surface(emu_sim *e) {
    memset(e->vol, 0, e->n_labels * 8);
    memset(e->surf, 0, e->n_labels * 8);
    for (size_t i = 0; i < e->n; ++i) {
        uint32_t l = e->labels[i];
        e->vol[l] += 1;
        int64_t x, y, z;
        e_coords(e, i, &x, &y, &z);
        for (int k = 0; k < e->n_off; ++k) {
            size_t nb = e_index(e, x + e->off[k][0], y + e->off[k][1], z + e->off[k][2]);
            if (nb != i && e->labels[nb] != l) e->surf[l] += 1;
        }
    }
}

/* Integer moment sums of every label about its anchor, and with the
 * nematic coupling the fixed-point field summed over each label's sites
 * (kept with the moments, and recounted with them). */
static void sum_moments(emu_sim *e) {
    memset(e->mom, 0, e->n_labels * 9 * 8);
    memset(e->nem, 0, e->n_labels * 2 * 8);
    if (e->has_nem)
        for (size_t i = 0; i < e->n; ++i) {
            uint32_t l = e->labels[i];
            if (l == 0) continue;
            e->nem[2 * l] += e->nq[2 * i];
            e->nem[2 * l + 1] += e->nq[2 * i + 1];
        }
    for (size_t i = 0; i < e->n; ++i) {
        uint32_t l = e->labels[i];
        if (l == 0) continue;
        int64_t x, y, z;
        e_coords(e, i, &x, &y, &z);
        int64_t ux = fold_int(x - e->anchor[3 * l], e->w);
        int64_t uy = fold_int(y - e->anchor[3 * l + 1], e->h);
        int64_t uz = fold_int(z - e->anchor[3 * l + 2], e->d);
        int64_t *m = e->mom + 9 * l;
        m[0] += ux;
        m[1] += uy;
        m[2] += uz;
        m[3] += ux * ux;
        m[4] += uy * uy;
        m[5] += uz * uz;
        m[6] += ux * uy;
        m[7] += ux * uz;
        m[8] += uy * uz;
    }
}

/* Move each live label's anchor to its rounded centroid, then resum. */
static void recentre(emu_sim *e) {
    int64_t span[3] = {e->w, e->h, e->d};
    for (size_t l = 1; l < e->n_labels; ++l) {
        if (e->vol[l] <= 0) continue;
        for (int a = 0; a < 3; ++a) {
            double mean = (double)e->mom[9 * l + a] / (double)e->vol[l];
            int64_t shift = (int64_t)floor(mean + 0.5);
            e->anchor[3 * l + a] = (int32_t)wrapi(e->anchor[3 * l + a] + shift, span[a]);
        }
    }
    sum_moments(e);
}

/* Anchors from scratch: each label's first site in lattice order, then one
 * recentring. */
static void anchor_from_first_sites(emu_sim *e) {
    for (size_t l = 0; l < e->n_labels; ++l) {
        e->anchor[3 * l] = e->anchor[3 * l + 1] = e->anchor[3 * l + 2] = -1;
    }
    for (size_t i = 0; i < e->n; ++i) {
        uint32_t l = e->labels[i];
        if (e->anchor[3 * l] >= 0) continue;
        int64_t x, y, z;
        e_coords(e, i, &x, &y, &z);
        e->anchor[3 * l] = (int32_t)x;
        e->anchor[3 * l + 1] = (int32_t)y;
        e->anchor[3 * l + 2] = (int32_t)z;
    }
    for (size_t l = 0; l < e->n_labels; ++l)
        if (e->anchor[3 * l] < 0) e->anchor[3 * l] = e->anchor[3 * l + 1] = e->anchor[3 * l + 2] = 0;
    sum_moments(e);
    recentre(e);
}

/* Whether some live label's centroid sits more than an eighth of a side from
 * its anchor, the point at which the sums are rebuilt about a new one. */
static int needs_recentre(const emu_sim *e) {
    int64_t span[3] = {e->w, e->h, e->d};
    for (size_t l = 1; l < e->n_labels; ++l) {
        int64_t n = e->vol[l];
        if (n <= 0) continue;
        for (int a = 0; a < 3; ++a) {
            if (span[a] <= 1) continue;
            int64_t s = e->mom[9 * l + a];
            int64_t as = s < 0 ? -s : s;
            if (8 * as > span[a] * n) return 1;
        }
    }
    return 0;
}

/* ---------------------------------------------------------------- geometry */

/* Largest eigenvalue and its axis of the scaled covariance C = n S2 - S1 S1^T
 * (exact integers), in f64, for the division plan. The plane uses the closed
 * form; a volume runs three sweeps of cyclic Jacobi rotations (within about
 * 1e-10 of the largest eigenvalue on integer covariances). Lengths in a
 * volume take the closed form largest_eigen3 below. */
static double cov_largest(const int64_t *m, int64_t n, int three_d, double axis[3]) {
    double cxx = (double)(n * m[3] - m[0] * m[0]);
    double cyy = (double)(n * m[4] - m[1] * m[1]);
    double cxy = (double)(n * m[6] - m[0] * m[1]);
    if (!three_d) {
        double half = 0.5 * (cxx - cyy);
        double r = sqrt(half * half + cxy * cxy);
        double lam = 0.5 * (cxx + cyy) + r;
        if (axis) {
            /* eigenvector of the largest eigenvalue */
            double ax, ay;
            if (r == 0.0) {
                ax = 1.0;
                ay = 0.0;
            } else if (half >= 0.0) {
                ax = half + r;
                ay = cxy;
            } else {
                ax = cxy;
                ay = r - half;
            }
            double nrm = sqrt(ax * ax + ay * ay);
            axis[0] = ax / nrm;
            axis[1] = ay / nrm;
            axis[2] = 0.0;
        }
        return lam;
    }
    double czz = (double)(n * m[5] - m[2] * m[2]);
    double cxz = (double)(n * m[7] - m[0] * m[2]);
    double cyz = (double)(n * m[8] - m[1] * m[2]);
    double a[3][3] = {{cxx, cxy, cxz}, {cxy, cyy, cyz}, {cxz, cyz, czz}};
    double v[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    static const int pairs[3][2] = {{0, 1}, {0, 2}, {1, 2}};
    for (int sweep = 0; sweep < 3; ++sweep) {
        for (int q = 0; q < 3; ++q) {
            int p = pairs[q][0], r = pairs[q][1];
            double apq = a[p][r];
            if (apq == 0.0) continue;
            double tau = (a[r][r] - a[p][p]) / (2.0 * apq);
            double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
            double c = 1.0 / sqrt(1.0 + t * t);
            double s = t * c;
            for (int k = 0; k < 3; ++k) {
                double akp = a[k][p], akr = a[k][r];
                a[k][p] = c * akp - s * akr;
                a[k][r] = s * akp + c * akr;
            }
            for (int k = 0; k < 3; ++k) {
                double apk = a[p][k], ark = a[r][k];
                a[p][k] = c * apk - s * ark;
                a[r][k] = s * apk + c * ark;
            }
            for (int k = 0; k < 3; ++k) {
                double vkp = v[k][p], vkr = v[k][r];
                v[k][p] = c * vkp - s * vkr;
                v[k][r] = s * vkp + c * vkr;
            }
        }
    }
    int best = 0;
    if (a[1][1] > a[best][best]) best = 1;
    if (a[2][2] > a[best][best]) best = 2;
    if (axis) {
        axis[0] = v[0][best];
        axis[1] = v[1][best];
        axis[2] = v[2][best];
    }
    return a[best][best] > 0.0 ? a[best][best] : 0.0;
}

/* In-plane anisotropy of the scaled covariance (the reference's
 * Moments::anisotropy, moments.rs:131-141, is scale-free): (C_xx - C_yy)/tr
 * and 2 C_xy/tr with tr = C_xx + C_yy, zero below two sites or at tr = 0
 * (an integer trace is 0 or at least 1, i.e. at least 1/n^2 unscaled). */
static void aniso(const int64_t *m, int64_t n, double *a0, double *a1) {
    *a0 = 0.0;
    *a1 = 0.0;
    if (n < 2) return;
    int64_t cxx = n * m[3] - m[0] * m[0];
    int64_t cyy = n * m[4] - m[1] * m[1];
    int64_t cxy = n * m[6] - m[0] * m[1];
    int64_t tr = cxx + cyy;
    if (tr <= 0) return;
    *a0 = (double)(cxx - cyy) / (double)tr;
    *a1 = 2.0 * (double)cxy / (double)tr;
}

// h(u) = cos(2/3 acos u) for u in [0, 1], analytic there, as a degree-14
// polynomial in t = 2u - 1 (a Chebyshev fit at 200 001 nodes written in
// powers of t; largest error 3.9e-14). cos(acos(r)/3) = h(sqrt((1 + r)/2))
// for r in [-1, 1].
static double third_angle_cos(double u) {
    const double t = 2.0 * u - 1.0;
    double h = -1.7072390069472367e-09;
    h = h * t + 5.577076476677568e-09;
    h = h * t + -1.24175244399134e-08;
    h = h * t + 4.321603925338742e-08;
    h = h * t + -1.6036631021280584e-07;
    h = h * t + 5.662416587052154e-07;
    h = h * t + -2.035473767371784e-06;
    h = h * t + 7.540205193337502e-06;
    h = h * t + -2.892003887918587e-05;
    h = h * t + 0.00011642563899772694;
    h = h * t + -0.0005041246781468507;
    h = h * t + 0.0024663527976783333;
    h = h * t + -0.015509188437110516;
    h = h * t + 0.24740906632332305;
    h = h * t + 0.7660444431189831;
    return h;
}

// Largest eigenvalue of the symmetric 3 x 3 matrix with diagonal (a00, a11,
// a22) and off-diagonal (a01, a02, a12), by the trigonometric closed form
// (O. K. Smith, Comm. ACM 4 (1961) 168): with q = tr/3, p^2 = |A - qI|^2/6
// and r = det((A - qI)/p)/2 in [-1, 1], it is q + 2p cos(acos(r)/3). Plain
// multiplies and adds, one division and two square roots, written the same
// in the emulation.
static double largest_eigen3(double a00, double a11, double a22, double a01, double a02, double a12) {
    const double q = (a00 + a11 + a22) / 3.0;
    const double d0 = a00 - q, d1 = a11 - q, d2 = a22 - q;
    const double p2 = d0 * d0 + d1 * d1 + d2 * d2 + 2.0 * (a01 * a01 + a02 * a02 + a12 * a12);
    if (!(p2 > 0.0)) return q;
    const double p = sqrt(p2 / 6.0);
    const double ip = 1.0 / p;
    const double b00 = d0 * ip, b11 = d1 * ip, b22 = d2 * ip;
    const double b01 = a01 * ip, b02 = a02 * ip, b12 = a12 * ip;
    const double det = b00 * (b11 * b22 - b12 * b12) - b01 * (b01 * b22 - b12 * b02) + b02 * (b01 * b12 - b11 * b02);
    double r = 0.5 * det;
    r = r < -1.0 ? -1.0 : (r > 1.0 ? 1.0 : r);
    return q + 2.0 * p * third_angle_cos(sqrt(0.5 * (1.0 + r)));
}

/* Major axis 4 sqrt(lambda_max(cov)) from integer sums of n sites: the
 * planar closed form in a plane, the 3 x 3 closed form above in a volume
 * (the division axis keeps the Jacobi rotations of cov_largest). */
static double length_of(const int64_t *m, int64_t n, int three_d) {
    if (n < 2) return 0.0;
    double lam;
    if (three_d) {
        lam = largest_eigen3((double)(n * m[3] - m[0] * m[0]), (double)(n * m[4] - m[1] * m[1]),
                             (double)(n * m[5] - m[2] * m[2]), (double)(n * m[6] - m[0] * m[1]),
                             (double)(n * m[7] - m[0] * m[2]), (double)(n * m[8] - m[1] * m[2]));
    } else {
        lam = cov_largest(m, n, 0, NULL);
    }
    if (lam < 0.0) lam = 0.0;
    return 4.0 * sqrt(lam) / (double)n;
}

/* The length above, for the tests: m holds the nine moment sums. */
double emu_length(const int64_t *m, int64_t n, int three_d) { return length_of(m, n, three_d); }

/* ---------------------------------------------------------------- creation */

emu_sim *emu_create(const ghip_model *m, const uint32_t *labels, const uint8_t *cell_type,
                    size_t n_labels, const double *fields, const double *activity) {
    emu_sim *e = calloc(1, sizeof *e);
    e->m = *m;
    e->w = m->width;
    e->h = m->height;
    e->d = m->depth;
    e->n = (size_t)(e->w * e->h * e->d);
    e->colours = e->d > 1 ? 8 : 4;
    int64_t off[26][3];
    /* Neighbour tables: the lattice.rs orders (offsets_for, lattice.rs:69-88). */
    {
        static const int o1[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
        static const int o2[18][3] = {{1, 0, 0},  {-1, 0, 0}, {0, 1, 0},   {0, -1, 0},  {0, 0, 1},
                                      {0, 0, -1}, {1, 1, 0},  {1, -1, 0},  {-1, 1, 0},  {-1, -1, 0},
                                      {1, 0, 1},  {1, 0, -1}, {-1, 0, 1},  {-1, 0, -1}, {0, 1, 1},
                                      {0, 1, -1}, {0, -1, 1}, {0, -1, -1}};
        int k = 0;
        int planar = e->d <= 1;
        if (!planar && m->neighbour_order == 1) {
            for (int i = 0; i < 6; ++i, ++k) memcpy(e->off[k], o1[i], sizeof o1[i]);
        } else if (!planar && m->neighbour_order == 2) {
            for (int i = 0; i < 18; ++i, ++k) memcpy(e->off[k], o2[i], sizeof o2[i]);
        } else if (planar && m->neighbour_order == 1) {
            for (int i = 0; i < 4; ++i, ++k) memcpy(e->off[k], o1[i], sizeof o1[i]);
        } else {
            for (int dz = -1; dz <= 1; ++dz)
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dx = -1; dx <= 1; ++dx) {
                        if (!dx && !dy && !dz) continue;
                        if (planar && dz) continue;
                        e->off[k][0] = dx;
                        e->off[k][1] = dy;
                        e->off[k][2] = dz;
                        ++k;
                    }
        }
        e->n_off = k;
        (void)off;
    }
    e->labels = malloc(e->n * 4);
    memcpy(e->labels, labels, e->n * 4);
    e->n_labels = n_labels;
    ensure_cap(e, n_labels);
    memcpy(e->ctype, cell_type, n_labels);
    e->medium_type = cell_type[0];

    int nt = (int)m->n_types;
    e->integral = 1;
    for (int a = 0; a < nt; ++a)
        for (int b = 0; b < nt; ++b) {
            double j = m->contact[a * nt + b];
            e->J[a][b] = (float)j;
            if (!is_integer(j)) e->integral = 0;
        }
    for (int t = 0; t < nt; ++t) {
        const ghip_cell_type *ct = &m->types[t];
        e->lam_v[t] = (float)ct->lambda_volume;
        e->tgt_v[t] = (float)ct->target_volume;
        e->lam_s[t] = (float)ct->lambda_surface;
        e->tgt_s[t] = (float)ct->target_surface;
        if (!is_integer(ct->lambda_volume) || !is_integer(ct->target_volume) ||
            !is_integer(ct->lambda_surface) || !is_integer(ct->target_surface))
            e->integral = 0;
        if (ct->connected) e->has_conn = 1;
        if (ct->lambda_length != 0.0) e->has_len = 1;
        if (ct->lambda_activity != 0.0 && ct->max_activity > 0.0) e->has_mot = 1;
        if (ct->external[0] != 0.0 || ct->external[1] != 0.0 || ct->external[2] != 0.0) e->has_ext = 1;
        if (ct->division_volume > 0.0 || ct->death_rate > 0.0) e->has_pop = 1;
        e->max_act[t] = (float)ct->max_activity;
        e->coef_act[t] = ct->max_activity > 0.0 ? (float)(ct->lambda_activity / ct->max_activity) : 0.0f;
        for (int a = 0; a < 3; ++a) e->ext[t][a] = (float)ct->external[a];
        for (uint32_t k = 0; k < m->n_species; ++k) {
            e->chi[t][k] = (float)m->chemotaxis[t][k];
            if (m->chemotaxis[t][k] != 0.0) e->has_chem = 1;
            e->sec[t][k] = (float)m->secretion[t][k];
            e->up[t][k] = (float)m->uptake[t][k];
        }
    }
    if (e->has_len || e->has_chem || e->has_mot || e->has_ext) e->integral = 0;
    e->tracks_moments = e->has_len || e->has_pop;
    /* Acceptance. The integral table holds, for dH = i, how many of the 2^24
     * uniforms accept: ceil(2^24 exp(-i/T)). */
    double lim = m->temperature * 24.0 * log(2.0);
    if (e->integral && lim + 1.0 < (double)EMU_MAX_THR) {
        e->n_thr = (int)floor(lim) + 1;
        e->thr[0] = 1u << 24;
        for (int i = 1; i < e->n_thr; ++i)
            e->thr[i] = (uint32_t)ceil(16777216.0 * exp(-(double)i / m->temperature));
    } else {
        e->integral = 0;
    }
    e->accept_c = (float)(1.4426950408889634 / m->temperature);

    if (e->has_conn) {
        int k = 0;
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    if (!dx && !dy && !dz) continue;
                    if (e->d <= 1 && dz) continue;
                    e->ring[k][0] = dx;
                    e->ring[k][1] = dy;
                    e->ring[k][2] = dz;
                    ++k;
                }
        e->ring_n = k;
    }
    for (uint32_t k = 0; k < m->n_species; ++k) {
        e->field[k] = malloc(e->n * 4);
        for (size_t i = 0; i < e->n; ++i)
            e->field[k][i] = (float)(fields ? fields[k * e->n + i] : m->species[k].initial);
        int st = serial_substeps(k, m);
        e->substeps[k] = st;
        e->fd[k] = (float)(m->species[k].diffusion / (double)st);
        e->fk[k] = (float)(m->species[k].decay / (double)st);
    }
    e->fscratch = malloc(e->n * 4);
    e->act = malloc(e->n * 4);
    for (size_t i = 0; i < e->n; ++i) e->act[i] = activity ? (float)activity[i] : 0.0f;
    e->prev_labels = malloc(e->n * 4);
    e->prev_act = malloc(e->n * 4);
    recount_volume_surface(e);
    if (e->tracks_moments)
        anchor_from_first_sites(e);
    else
        memset(e->mom, 0, e->n_labels * 9 * 8);
    memset(e->nem, 0, e->n_labels * 2 * 8);
    return e;
}

void emu_destroy(emu_sim *e) {
    if (!e) return;
    free(e->labels);
    free(e->prev_labels);
    free(e->ctype);
    free(e->vol);
    free(e->surf);
    free(e->mom);
    free(e->vsnap);
    free(e->ssnap);
    free(e->msnap);
    free(e->anchor);
    free(e->nq);
    free(e->nem);
    free(e->nsnap);
    for (int k = 0; k < GHIP_MAX_SPECIES; ++k) free(e->field[k]);
    free(e->fscratch);
    free(e->act);
    free(e->prev_act);
    free(e);
}

/* ---------------------------------------------------------------- connectivity */

/* Whether the ring positions in `present` form one piece, two positions being
 * joined when they are within one step on every axis (connectivity.rs:31-35).
 * A breadth-first search over the set bits. */
int emu_one_piece(uint32_t present, int ring_size) {
    int pos[26][3], k = 0;
    for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                if (!dx && !dy && !dz) continue;
                if (ring_size == 8 && dz) continue;
                pos[k][0] = dx;
                pos[k][1] = dy;
                pos[k][2] = dz;
                ++k;
            }
    if (present == 0 || (present & (present - 1)) == 0) return 1;
    uint32_t reached = present & (~present + 1u);
    uint32_t frontier = reached;
    while (frontier) {
        uint32_t next = 0;
        for (int a = 0; a < ring_size; ++a) {
            if (!(frontier & (1u << a))) continue;
            for (int b = 0; b < ring_size; ++b) {
                if (!(present & (1u << b)) || (reached & (1u << b))) continue;
                if (abs(pos[a][0] - pos[b][0]) <= 1 && abs(pos[a][1] - pos[b][1]) <= 1 &&
                    abs(pos[a][2] - pos[b][2]) <= 1)
                    next |= 1u << b;
            }
        }
        reached |= next;
        frontier = next;
    }
    return reached == present;
}

/* ---------------------------------------------------------------- pricing */

/* Geometric mean of the activity over the sites of `label` around `site`,
 * counting the site itself (motility.rs:56-78), from the pass-start copy. */
static float geo_mean(const emu_sim *e, size_t site, uint32_t label) {
    float here = e->prev_act[site];
    if (!(here > 0.0f)) return 0.0f;
    float lsum = g_log2(here);
    float count = 1.0f;
    int64_t x, y, z;
    e_coords(e, site, &x, &y, &z);
    for (int k = 0; k < e->n_off; ++k) {
        size_t nb = e_index(e, x + e->off[k][0], y + e->off[k][1], z + e->off[k][2]);
        if (nb == site || e->prev_labels[nb] != label) continue;
        float v = e->prev_act[nb];
        if (!(v > 0.0f)) return 0.0f;
        lsum = lsum + g_log2(v);
        count = count + 1.0f;
    }
    return g_exp2(lsum / count);
}

/* The f32 geometric mean of the pricing, over the current lattice and
 * activity (a pass reads them as they stood at its start). */
float emu_activity_mean(emu_sim *e, size_t site, uint32_t label) {
    memcpy(e->prev_labels, e->labels, e->n * 4);
    memcpy(e->prev_act, e->act, e->n * 4);
    return geo_mean(e, site, label);
}

/* Change in the length energy of `label` when the site (x, y, z) joins
 * (sign +1) or leaves (sign -1), from the snapshot sums. */
static double length_delta(const emu_sim *e, uint32_t label, int64_t x, int64_t y, int64_t z,
                           int sign) {
    int t = e->ctype[label];
    const ghip_cell_type *ct = &e->m.types[t];
    int64_t n = e->vsnap[label];
    const int64_t *m = e->msnap + 9 * label;
    int64_t ux = fold_int(x - e->anchor[3 * label], e->w);
    int64_t uy = fold_int(y - e->anchor[3 * label + 1], e->h);
    int64_t uz = fold_int(z - e->anchor[3 * label + 2], e->d);
    int64_t after[9] = {m[0] + sign * ux,      m[1] + sign * uy,      m[2] + sign * uz,
                        m[3] + sign * ux * ux, m[4] + sign * uy * uy, m[5] + sign * uz * uz,
                        m[6] + sign * ux * uy, m[7] + sign * ux * uz, m[8] + sign * uy * uz};
    int three_d = e->d > 1;
    double before = length_of(m, n, three_d) - ct->target_length;
    double now = length_of(after, n + sign, three_d) - ct->target_length;
    return ct->lambda_length * (now * now - before * before);
}

/* Change in the nematic energy -lambda F . a (cpu.rs:315-332) of `label`
 * when the site `site` at (x, y, z) joins (sign +1) or leaves (sign -1): F
 * the fixed-point field sum (scale 2^24), a the anisotropy, both from the
 * pass-start snapshot and after the move. */
static double nematic_delta(const emu_sim *e, uint32_t label, int64_t x, int64_t y, int64_t z,
                            size_t site, int sign) {
    int t = e->ctype[label];
    int64_t n = e->vsnap[label];
    const int64_t *m = e->msnap + 9 * label;
    int64_t ux = fold_int(x - e->anchor[3 * label], e->w);
    int64_t uy = fold_int(y - e->anchor[3 * label + 1], e->h);
    int64_t uz = fold_int(z - e->anchor[3 * label + 2], e->d);
    int64_t after[9] = {m[0] + sign * ux,      m[1] + sign * uy,      m[2] + sign * uz,
                        m[3] + sign * ux * ux, m[4] + sign * uy * uy, m[5] + sign * uz * uz,
                        m[6] + sign * ux * uy, m[7] + sign * ux * uz, m[8] + sign * uy * uz};
    double a0, a1, b0, b1;
    aniso(m, n, &a0, &a1);
    aniso(after, n + sign, &b0, &b1);
    const int64_t *f = e->nsnap + 2 * label;
    int64_t q0 = e->nq[2 * site], q1 = e->nq[2 * site + 1];
    double before = (double)f[0] * a0 + (double)f[1] * a1;
    double now = (double)(f[0] + sign * q0) * b0 + (double)(f[1] + sign * q1) * b1;
    return -e->m.types[t].lambda_nematic * ((now - before) * 5.9604644775390625e-8);
}

typedef struct {
    float dh;
    int ds_old, ds_new;
} price_t;

/* The energy change of copying `src`'s label into target `t`, from the
 * pass-start state. Terms are added in this order: contact (gained minus
 * lost), volume old, volume new, surface old, surface new, length old,
 * length new, nematic old, nematic new, chemotaxis per species, motility,
 * drift. */
static price_t price(const emu_sim *e, size_t t, size_t src, int pick) {
    price_t r;
    uint32_t old = e->labels[t], nw = e->labels[src];
    int eto = ctype_of(e, old), etn = ctype_of(e, nw);
    int64_t x, y, z;
    e_coords(e, t, &x, &y, &z);
    float cn = 0.0f, co = 0.0f;
    int like_old = 0, like_new = 0, bonds = e->n_off;
    for (int k = 0; k < e->n_off; ++k) {
        size_t nb = e_index(e, x + e->off[k][0], y + e->off[k][1], z + e->off[k][2]);
        uint32_t ln = e->labels[nb];
        int tn = ctype_of(e, ln);
        if (ln != nw)
            cn = cn + e->J[etn][tn];
        else
            ++like_new;
        if (ln != old)
            co = co + e->J[eto][tn];
        else
            ++like_old;
    }
    float dh = cn - co;
    r.ds_old = 2 * like_old - bonds;
    r.ds_new = bonds - 2 * like_new;
    int to = e->ctype[old], tw = e->ctype[nw];
    if (old != 0 && e->lam_v[to] != 0.0f) {
        float v = (float)e->vsnap[old];
        dh = dh + e->lam_v[to] * (1.0f - 2.0f * (v - e->tgt_v[to]));
    }
    if (nw != 0 && e->lam_v[tw] != 0.0f) {
        float v = (float)e->vsnap[nw];
        dh = dh + e->lam_v[tw] * (1.0f + 2.0f * (v - e->tgt_v[tw]));
    }
    if (old != 0 && e->lam_s[to] != 0.0f && r.ds_old != 0) {
        float s = (float)e->ssnap[old];
        float dsf = (float)r.ds_old;
        dh = dh + e->lam_s[to] * (dsf * (2.0f * (s - e->tgt_s[to]) + dsf));
    }
    if (nw != 0 && e->lam_s[tw] != 0.0f && r.ds_new != 0) {
        float s = (float)e->ssnap[nw];
        float dsf = (float)r.ds_new;
        dh = dh + e->lam_s[tw] * (dsf * (2.0f * (s - e->tgt_s[tw]) + dsf));
    }
    if (e->has_len) {
        if (old != 0 && e->m.types[to].lambda_length != 0.0)
            dh = dh + (float)length_delta(e, old, x, y, z, -1);
        if (nw != 0 && e->m.types[tw].lambda_length != 0.0)
            dh = dh + (float)length_delta(e, nw, x, y, z, 1);
    }
    if (e->has_nem) {
        if (old != 0 && e->m.types[to].lambda_nematic != 0.0)
            dh = dh + (float)nematic_delta(e, old, x, y, z, t, -1);
        if (nw != 0 && e->m.types[tw].lambda_nematic != 0.0)
            dh = dh + (float)nematic_delta(e, nw, x, y, z, t, 1);
    }
    if (e->has_chem && nw != 0) {
        for (uint32_t k = 0; k < e->m.n_species; ++k) {
            float chi = e->chi[tw][k];
            if (chi == 0.0f) continue;
            dh = dh - chi * (e->field[k][t] - e->field[k][src]);
        }
    }
    if (e->has_mot && nw != 0 && e->coef_act[tw] != 0.0f && e->max_act[tw] > 0.0f) {
        float into = geo_mean(e, src, nw);
        float out_of = geo_mean(e, t, old);
        dh = dh - e->coef_act[tw] * (into - out_of);
    }
    if (e->has_ext && nw != 0) {
        /* target minus source is minus the picked offset */
        float w = e->ext[tw][0] * (float)(-e->off[pick][0]) + e->ext[tw][1] * (float)(-e->off[pick][1]) +
                  e->ext[tw][2] * (float)(-e->off[pick][2]);
        dh = dh - w;
    }
    r.dh = dh;
    return r;
}

double emu_delta_energy(const emu_sim *e, size_t target, size_t source) {
    emu_sim *me = (emu_sim *)e;
    memcpy(me->vsnap, e->vol, e->n_labels * 8);
    memcpy(me->ssnap, e->surf, e->n_labels * 8);
    memcpy(me->msnap, e->mom, e->n_labels * 9 * 8);
    memcpy(me->nsnap, e->nem, e->n_labels * 2 * 8);
    memcpy(me->prev_labels, e->labels, e->n * 4);
    memcpy(me->prev_act, e->act, e->n * 4);
    int pick = 0;
    int64_t tx, ty, tz, sx, sy, sz;
    e_coords(e, target, &tx, &ty, &tz);
    e_coords(e, source, &sx, &sy, &sz);
    for (int k = 0; k < e->n_off; ++k)
        if (e_index(e, tx + e->off[k][0], ty + e->off[k][1], tz + e->off[k][2]) == source) pick = k;
    return (double)price(e, target, source, pick).dh;
}

/* ---------------------------------------------------------------- the sweep */

static int accept_draw(const emu_sim *e, float dh, uint32_t r24) {
    if (dh <= 0.0f) return 1;
    if (e->integral) {
        if (!(dh < (float)e->n_thr)) return 0;
        int i = (int)dh;
        return r24 < e->thr[i];
    }
    float p = g_exp2(-dh * e->accept_c);
    return (float)r24 < p * 16777216.0f;
}

static void one_pass(emu_sim *e, int colour) {
    memcpy(e->vsnap, e->vol, e->n_labels * 8);
    memcpy(e->ssnap, e->surf, e->n_labels * 8);
    if (e->tracks_moments) memcpy(e->msnap, e->mom, e->n_labels * 9 * 8);
    if (e->has_nem) memcpy(e->nsnap, e->nem, e->n_labels * 2 * 8);
    if (e->has_mot) {
        memcpy(e->prev_labels, e->labels, e->n * 4);
        memcpy(e->prev_act, e->act, e->n * 4);
    }
    uint32_t key = oracle_pass_key(e->m.seed, e->m.replica, e->mcs, (uint32_t)colour);
    uint32_t key2 = second_key(key);
    int cx = colour & 1, cy = (colour >> 1) & 1, cz = (colour >> 2) & 1;
    int pow2 = (e->n_off & (e->n_off - 1)) == 0;
    for (int64_t z = cz; z < e->d; z += (e->d > 1 ? 2 : 1)) {
        for (int64_t y = cy; y < e->h; y += 2) {
            for (int64_t x = cx; x < e->w; x += 2) {
                size_t t = e_index(e, x, y, z);
                uint32_t ph = oracle_hash32((uint32_t)t ^ key);
                uint32_t h = oracle_hash32(ph ^ key2);
                int pick;
                uint32_t r24;
                if (pow2) {
                    pick = (int)(ph & (uint32_t)(e->n_off - 1));
                    r24 = h >> 8;
                } else {
                    pick = (int)(((uint64_t)h * (uint64_t)e->n_off) >> 32);
                    r24 = oracle_hash32(h ^ 0x68E31DA4U) >> 8;
                }
                e->attempted += 1;
                size_t src = e_index(e, x + e->off[pick][0], y + e->off[pick][1], z + e->off[pick][2]);
                uint32_t old = e->labels[t], nw = e->labels[src];
                if (old == nw) continue;
                if (old != 0 && e->vsnap[old] <= 1) continue;
                if (old != 0 && e->has_conn && e->m.types[e->ctype[old]].connected) {
                    uint32_t present = 0;
                    for (int k = 0; k < e->ring_n; ++k) {
                        size_t nb = e_index(e, x + e->ring[k][0], y + e->ring[k][1], z + e->ring[k][2]);
                        if (nb != t && e->labels[nb] == old) present |= 1u << k;
                    }
                    if (!emu_one_piece(present, e->ring_n)) continue;
                }
                price_t p = price(e, t, src, pick);
                if (!accept_draw(e, p.dh, r24)) continue;
                /* commit */
                e->labels[t] = nw;
                if (e->has_mot && nw != 0 && e->max_act[e->ctype[nw]] > 0.0f) e->act[t] = e->max_act[e->ctype[nw]];
                if (old != 0) {
                    e->vol[old] -= 1;
                    e->surf[old] += p.ds_old;
                }
                if (nw != 0) {
                    e->vol[nw] += 1;
                    e->surf[nw] += p.ds_new;
                }
                if (e->tracks_moments) {
                    uint32_t ls[2] = {old, nw};
                    int sg[2] = {-1, 1};
                    for (int q = 0; q < 2; ++q) {
                        uint32_t l = ls[q];
                        if (l == 0) continue;
                        int64_t ux = fold_int(x - e->anchor[3 * l], e->w);
                        int64_t uy = fold_int(y - e->anchor[3 * l + 1], e->h);
                        int64_t uz = fold_int(z - e->anchor[3 * l + 2], e->d);
                        int64_t *m = e->mom + 9 * l;
                        int64_t s = sg[q];
                        m[0] += s * ux;
                        m[1] += s * uy;
                        m[2] += s * uz;
                        m[3] += s * ux * ux;
                        m[4] += s * uy * uy;
                        m[5] += s * uz * uz;
                        m[6] += s * ux * uy;
                        m[7] += s * ux * uz;
                        m[8] += s * uy * uz;
                    }
                }
                if (e->has_nem) {
                    if (old != 0) {
                        e->nem[2 * old] -= e->nq[2 * t];
                        e->nem[2 * old + 1] -= e->nq[2 * t + 1];
                    }
                    if (nw != 0) {
                        e->nem[2 * nw] += e->nq[2 * t];
                        e->nem[2 * nw + 1] += e->nq[2 * t + 1];
                    }
                }
                e->accepted += 1;
            }
        }
    }
}

/* ---------------------------------------------------------------- fields */

/* One step of every species: substeps of explicit Euler on the face
 * Laplacian (field.rs:111-130) in f32, sum taken +x, -x, +y, -y(, +z, -z),
 * then secretion and uptake on owned sites (field.rs:133-149). */
void emu_fields_step(emu_sim *e) {
    int three = e->d > 1;
    float faces = three ? 6.0f : 4.0f;
    for (uint32_t k = 0; k < e->m.n_species; ++k) {
        float d = e->fd[k], dk = e->fk[k];
        for (int st = 0; st < e->substeps[k]; ++st) {
            float *u = e->field[k];
            for (size_t i = 0; i < e->n; ++i) {
                int64_t x, y, z;
                e_coords(e, i, &x, &y, &z);
                float sum = u[e_index(e, x + 1, y, z)] + u[e_index(e, x - 1, y, z)];
                sum = sum + u[e_index(e, x, y + 1, z)];
                sum = sum + u[e_index(e, x, y - 1, z)];
                if (three) {
                    sum = sum + u[e_index(e, x, y, z + 1)];
                    sum = sum + u[e_index(e, x, y, z - 1)];
                }
                float here = u[i];
                e->fscratch[i] = (here + d * (sum - faces * here)) - dk * here;
            }
            memcpy(u, e->fscratch, e->n * 4);
        }
    }
    for (uint32_t k = 0; k < e->m.n_species; ++k) {
        float *u = e->field[k];
        for (size_t i = 0; i < e->n; ++i) {
            uint32_t l = e->labels[i];
            if (l == 0) continue;
            int t = e->ctype[l];
            float v = u[i] + e->sec[t][k];
            v = v - e->up[t][k] * v;
            u[i] = v < 0.0f ? 0.0f : v;
        }
    }
}

/* ---------------------------------------------------------------- population */

static uint32_t death_draw(const emu_sim *e, uint32_t label) {
    uint32_t key = oracle_pass_key(e->m.seed, e->m.replica, e->mcs, 0xDEADCE11U);
    return oracle_hash32(oracle_hash32(label ^ key) ^ second_key(key));
}

/* Compaction: live labels renumbered densely in ascending order, each taking
 * its type, anchor and records along. */
static void compact(emu_sim *e) {
    uint32_t *remap = malloc(e->n_labels * 4);
    size_t next = 1;
    remap[0] = 0;
    for (size_t l = 1; l < e->n_labels; ++l) {
        if (e->vol[l] > 0) {
            remap[l] = (uint32_t)next;
            e->ctype[next] = e->ctype[l];
            memmove(e->anchor + 3 * next, e->anchor + 3 * l, 3 * 4);
            e->vol[next] = e->vol[l];
            e->surf[next] = e->surf[l];
            memmove(e->mom + 9 * next, e->mom + 9 * l, 9 * 8);
            memmove(e->nem + 2 * next, e->nem + 2 * l, 2 * 8);
            ++next;
        } else {
            remap[l] = 0;
        }
    }
    for (size_t i = 0; i < e->n; ++i) e->labels[i] = remap[e->labels[i]];
    e->n_labels = next;
    free(remap);
}

/* Which side of its parent's cut a site of a dividing label falls on. */
static int beyond_cut(const emu_sim *e, const double *geo, uint32_t l, int64_t x, int64_t y, int64_t z) {
    double ux = (double)fold_int(x - e->anchor[3 * l], e->w);
    double uy = (double)fold_int(y - e->anchor[3 * l + 1], e->h);
    double uz = (double)fold_int(z - e->anchor[3 * l + 2], e->d);
    const double *g = geo + 6 * l;
    double along = (ux - g[0]) * g[3] + (uy - g[1]) * g[4] + (uz - g[2]) * g[5];
    return along > 0.0;
}

/* Division and death of one step. Only the labels an event touches change
 * their records: a dying label's are cleared (its sites become medium, which
 * leaves every other label's surface as it was: a neighbour's bond to the
 * site was unlike before and stays unlike), and a dividing label and its
 * daughter are recounted from their sites, with the daughter taking the
 * parent's anchor. Other labels' records stay exact, so there is no full
 * recount; compaction moves records with their labels. */
static void population(emu_sim *e) {
    int three = e->d > 1;
    size_t n0 = e->n_labels;
    /* plans: 0 nothing, 1 divide, 2 die */
    uint8_t *plan = calloc(n0, 1);
    double *geo = calloc(n0 * 6, sizeof(double)); /* mean xyz, axis xyz */
    int any = 0;
    for (size_t l = 1; l < n0; ++l) {
        int64_t v = e->vol[l];
        if (v <= 0) continue;
        const ghip_cell_type *ct = &e->m.types[e->ctype[l]];
        if (ct->division_volume > 0.0 && (double)v >= ct->division_volume && v >= 4) {
            plan[l] = 1;
            const int64_t *m = e->mom + 9 * l;
            for (int a = 0; a < 3; ++a) geo[6 * l + a] = (double)m[a] / (double)v;
            cov_largest(m, v, three, geo + 6 * l + 3);
            any = 1;
        } else if (ct->death_rate > 0.0) {
            uint32_t h = death_draw(e, (uint32_t)l);
            if ((double)h * (1.0 / 4294967296.0) < ct->death_rate) {
                plan[l] = 2;
                any = 1;
            }
        }
    }
    if (!any) {
        free(plan);
        free(geo);
        return;
    }
    /* count the sites each division would move */
    int64_t *moved = calloc(n0, 8);
    for (size_t i = 0; i < e->n; ++i) {
        uint32_t l = e->labels[i];
        if (l == 0 || plan[l] != 1) continue;
        int64_t x, y, z;
        e_coords(e, i, &x, &y, &z);
        if (beyond_cut(e, geo, l, x, y, z)) moved[l] += 1;
    }
    uint32_t *daughter = calloc(n0, 4);
    size_t next = n0;
    for (size_t l = 1; l < n0; ++l) {
        if (plan[l] != 1) continue;
        if (moved[l] == 0 || moved[l] == e->vol[l]) {
            plan[l] = 0;
            continue;
        }
        daughter[l] = (uint32_t)next++;
    }
    ensure_cap(e, next);
    /* new label of every site, from its old one (a daughter keeps its id) */
    uint32_t *fresh = malloc(e->n * 4);
    for (size_t i = 0; i < e->n; ++i) {
        uint32_t l = e->labels[i];
        fresh[i] = l;
        if (l == 0 || plan[l] == 0) continue;
        if (plan[l] == 2) {
            fresh[i] = 0;
            continue;
        }
        int64_t x, y, z;
        e_coords(e, i, &x, &y, &z);
        if (beyond_cut(e, geo, l, x, y, z)) fresh[i] = daughter[l];
    }
    for (size_t l = 1; l < n0; ++l) {
        if (plan[l] == 0) continue;
        size_t ids[2] = {l, plan[l] == 1 ? daughter[l] : 0};
        for (int q = 0; q < 2; ++q) {
            size_t k = ids[q];
            if (q == 1 && plan[l] != 1) break;
            e->vol[k] = 0;
            e->surf[k] = 0;
            memset(e->mom + 9 * k, 0, 9 * 8);
            memset(e->nem + 2 * k, 0, 2 * 8);
        }
        if (plan[l] == 1) {
            uint32_t dl = daughter[l];
            e->ctype[dl] = e->ctype[l];
            memcpy(e->anchor + 3 * dl, e->anchor + 3 * l, 3 * 4);
        }
    }
    /* recount the parents and daughters from their sites */
    for (size_t i = 0; i < e->n; ++i) {
        uint32_t old = e->labels[i];
        if (old == 0 || plan[old] != 1) continue;
        uint32_t l = fresh[i];
        int64_t x, y, z;
        e_coords(e, i, &x, &y, &z);
        e->vol[l] += 1;
        for (int k = 0; k < e->n_off; ++k) {
            size_t nb = e_index(e, x + e->off[k][0], y + e->off[k][1], z + e->off[k][2]);
            if (nb != i && fresh[nb] != l) e->surf[l] += 1;
        }
        int64_t ux = fold_int(x - e->anchor[3 * l], e->w);
        int64_t uy = fold_int(y - e->anchor[3 * l + 1], e->h);
        int64_t uz = fold_int(z - e->anchor[3 * l + 2], e->d);
        int64_t *m = e->mom + 9 * l;
        m[0] += ux;
        m[1] += uy;
        m[2] += uz;
        m[3] += ux * ux;
        m[4] += uy * uy;
        m[5] += uz * uz;
        m[6] += ux * uy;
        m[7] += ux * uz;
        m[8] += uy * uz;
        if (e->has_nem) {
            e->nem[2 * l] += e->nq[2 * i];
            e->nem[2 * l + 1] += e->nq[2 * i + 1];
        }
    }
    memcpy(e->labels, fresh, e->n * 4);
    e->n_labels = next;
    free(fresh);
    free(plan);
    free(geo);
    free(moved);
    free(daughter);
    size_t dead = 0;
    for (size_t l = 1; l < e->n_labels; ++l)
        if (e->vol[l] <= 0) ++dead;
    size_t limit = (e->n_labels - 1) / 4;
    if (limit < 64) limit = 64;
    if (dead > limit) compact(e);
}

/* ---------------------------------------------------------------- step */

void emu_step(emu_sim *e, uint64_t steps) {
    for (uint64_t q = 0; q < steps; ++q) {
        int order[8];
        colour_order(e->m.seed, e->m.replica, e->mcs, e->colours, order);
        for (int c = 0; c < e->colours; ++c) one_pass(e, order[c]);
        if (e->m.n_species > 0) emu_fields_step(e);
        if (e->has_mot)
            for (size_t i = 0; i < e->n; ++i) e->act[i] = e->act[i] > 0.0f ? e->act[i] - 1.0f : 0.0f;
        /* the drift check reads the records the step's plan reads, before
         * division and death; the recentring follows them */
        const int drifted = e->tracks_moments && needs_recentre(e);
        if (e->has_pop) population(e);
        if (drifted) recentre(e);
        e->mcs += 1;
    }
}

size_t emu_n_labels(const emu_sim *e) { return e->n_labels; }
uint64_t emu_mcs(const emu_sim *e) { return e->mcs; }
void emu_counters(const emu_sim *e, uint64_t *attempted, uint64_t *accepted) {
    *attempted = e->attempted;
    *accepted = e->accepted;
}
void emu_labels(const emu_sim *e, uint32_t *out) { memcpy(out, e->labels, e->n * 4); }
void emu_volumes(const emu_sim *e, uint32_t *out) {
    /* the medium's count is recomputed, as on the device */
    uint64_t cells = 0;
    for (size_t l = 1; l < e->n_labels; ++l) {
        out[l] = (uint32_t)(e->vol[l] > 0 ? e->vol[l] : 0);
        cells += (uint64_t)out[l];
    }
    out[0] = (uint32_t)(e->n - cells);
}
void emu_surfaces(const emu_sim *e, int64_t *out) {
    memcpy(out, e->surf, e->n_labels * 8);
    int64_t s0 = 0;
    for (size_t i = 0; i < e->n; ++i) {
        if (e->labels[i] != 0) continue;
        int64_t x, y, z;
        e_coords(e, i, &x, &y, &z);
        for (int k = 0; k < e->n_off; ++k) {
            size_t nb = e_index(e, x + e->off[k][0], y + e->off[k][1], z + e->off[k][2]);
            if (nb != i && e->labels[nb] != 0) s0 += 1;
        }
    }
    out[0] = s0;
}
void emu_cell_types(const emu_sim *e, uint8_t *out) { memcpy(out, e->ctype, e->n_labels); }

/* Words per label: volume, surface, then the nine moment sums when kept. */
uint32_t emu_cell_state(const emu_sim *e, int64_t *out) {
    uint32_t words = e->tracks_moments ? (e->has_nem ? 13 : 11) : 2;
    if (!out) return words;
    for (size_t l = 0; l < e->n_labels; ++l) {
        out[words * l] = e->vol[l];
        out[words * l + 1] = e->surf[l];
        if (words >= 11) memcpy(out + words * l + 2, e->mom + 9 * l, 9 * 8);
        if (words == 13) memcpy(out + words * l + 11, e->nem + 2 * l, 2 * 8);
    }
    return words;
}

/* The nematic field, one (Q_xx, Q_xy) pair per site, as the engine holds it:
 * each value rounded to the nearest multiple of 2^-24 (ties away from zero),
 * so per-label sums are exact integers in any order. The coupling is on when
 * a field is given and some type has a lambda_nematic (model.rs:120-124); it
 * keeps the moments (model.rs:127-130), which start from the anchors of the
 * current lattice if they were not kept before. Replacing the field resums. */
int32_t emu_nematic_fixed(double q) { return (int32_t)llround(q * 16777216.0); }

void emu_set_nematic(emu_sim *e, const double *q) {
    free(e->nq);
    e->nq = NULL;
    e->has_nem = 0;
    if (q) {
        e->nq = malloc(e->n * 2 * 4);
        for (size_t i = 0; i < 2 * e->n; ++i) e->nq[i] = emu_nematic_fixed(q[i]);
        for (uint32_t t = 0; t < e->m.n_types; ++t)
            if (e->m.types[t].lambda_nematic != 0.0) e->has_nem = 1;
    }
    if (e->has_nem) e->integral = 0; /* the term is real-valued */
    if (e->has_nem && !e->tracks_moments) {
        e->tracks_moments = 1;
        anchor_from_first_sites(e);
    } else {
        sum_moments(e);
    }
}
void emu_field_f32(const emu_sim *e, uint32_t species, float *out) {
    memcpy(out, e->field[species], e->n * 4);
}
void emu_set_field(emu_sim *e, uint32_t species, const double *values) {
    for (size_t i = 0; i < e->n; ++i) e->field[species][i] = (float)values[i];
}
void emu_activity_f32(const emu_sim *e, float *out) { memcpy(out, e->act, e->n * 4); }
