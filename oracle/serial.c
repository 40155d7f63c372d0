/*
 * serial.c - restatement of glazier-core's serial f64 Metropolis engine.
 *
 * TEST INFRASTRUCTURE ONLY: the distributional oracle and the timed CPU
 * baseline. Nothing in the engine links it.
 *
 * Each function names the reference lines it follows. Arithmetic is f64 in
 * the reference's order, built with -ffp-contract=off so no product is fused
 * into a sum. The nematic term is not restated: the device engine refuses a
 * nematic model (glazier-cuda/src/lib.rs:819-821 refuses rather than drops),
 * so nothing on the device path needs it.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

/* ---------------------------------------------------------------- rng.rs */

/* SplitMix64 step (glazier-core/src/rng.rs:19-25, Vigna's published mixer). */
uint64_t oracle_splitmix64_next(uint64_t *state) {
    *state += 0x9E3779B97F4A7C15ULL;
    uint64_t x = *state;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

/* Xoshiro::seed (rng.rs:17-29): four SplitMix64 outputs fill the state. */
void oracle_xoshiro_seed(uint64_t seed, uint64_t s[4]) {
    uint64_t z = seed;
    for (int i = 0; i < 4; ++i) s[i] = oracle_splitmix64_next(&z);
}

static inline uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

/* xoshiro256++ next (rng.rs:32-45). */
uint64_t oracle_xoshiro_next(uint64_t s[4]) {
    uint64_t result = rotl64(s[0] + s[3], 23) + s[0];
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl64(s[3], 45);
    return result;
}

/* next_f64 (rng.rs:48-50): the top 53 bits scaled into [0, 1). */
double oracle_xoshiro_f64(uint64_t s[4]) {
    return (double)(oracle_xoshiro_next(s) >> 11) * (1.0 / 9007199254740992.0);
}

/* below (rng.rs:53-62): rejection above the largest multiple of n. */
uint64_t oracle_xoshiro_below(uint64_t s[4], uint64_t n) {
    uint64_t zone = UINT64_MAX - (UINT64_MAX % n);
    for (;;) {
        uint64_t v = oracle_xoshiro_next(s);
        if (v < zone) return v % n;
    }
}

/* ---------------------------------------------------------------- moments.rs */

typedef struct {
    double n;
    double s[3];
    double p[6]; /* xx, yy, zz, xy, xz, yz */
    int has_anchor;
    double anchor[3];
} moments_t;

static double fold_about(double v, double a, double span) {
    double out = v;
    if (span > 1.0) {
        if (out - a > span / 2.0)
            out -= span;
        else if (a - out > span / 2.0)
            out += span;
    }
    return out;
}

/* Moments::unwrap (moments.rs:32-52). */
static void mom_unwrap(const moments_t *m, double x, double y, double z, const double ext[3],
                       double out[3]) {
    if (!m->has_anchor) {
        out[0] = x;
        out[1] = y;
        out[2] = z;
        return;
    }
    out[0] = fold_about(x, m->anchor[0], ext[0]);
    out[1] = fold_about(y, m->anchor[1], ext[1]);
    out[2] = fold_about(z, m->anchor[2], ext[2]);
}

/* Moments::with (moments.rs:74-89). */
static moments_t mom_with(const moments_t *m, const double site[3], double sign) {
    double x = site[0], y = site[1], z = site[2];
    moments_t o = *m;
    o.n += sign;
    o.s[0] += sign * x;
    o.s[1] += sign * y;
    o.s[2] += sign * z;
    o.p[0] += sign * x * x;
    o.p[1] += sign * y * y;
    o.p[2] += sign * z * z;
    o.p[3] += sign * x * y;
    o.p[4] += sign * x * z;
    o.p[5] += sign * y * z;
    return o;
}

/* Moments::add (moments.rs:55-60). */
static void mom_add(moments_t *m, const double site[3]) {
    if (!m->has_anchor) {
        m->has_anchor = 1;
        m->anchor[0] = site[0];
        m->anchor[1] = site[1];
        m->anchor[2] = site[2];
    }
    *m = mom_with(m, site, 1.0);
}

/* Moments::remove (moments.rs:63-70). */
static void mom_remove(moments_t *m, const double site[3]) {
    int had = m->has_anchor;
    double a[3] = {m->anchor[0], m->anchor[1], m->anchor[2]};
    *m = mom_with(m, site, -1.0);
    m->has_anchor = had;
    memcpy(m->anchor, a, sizeof a);
    if (m->n <= 0.0) memset(m, 0, sizeof *m);
}

/* Moments::covariance (moments.rs:92-105). */
static void mom_covariance(const moments_t *m, double c[9]) {
    if (m->n <= 0.0) {
        memset(c, 0, 9 * sizeof(double));
        return;
    }
    double mx = m->s[0] / m->n, my = m->s[1] / m->n, mz = m->s[2] / m->n;
    double xx = m->p[0] / m->n - mx * mx;
    double yy = m->p[1] / m->n - my * my;
    double zz = m->p[2] / m->n - mz * mz;
    double xy = m->p[3] / m->n - mx * my;
    double xz = m->p[4] / m->n - mx * mz;
    double yz = m->p[5] / m->n - my * mz;
    double cc[9] = {xx, xy, xz, xy, yy, yz, xz, yz, zz};
    memcpy(c, cc, sizeof cc);
}

/* Moments::anisotropy (moments.rs:131-141): the in-plane anisotropy
 * ((c_xx - c_yy) / tr, 2 c_xy / tr) with tr = c_xx + c_yy, zero for fewer
 * than two sites or a trace at or below 1e-12. */
static void mom_anisotropy(const moments_t *m, double *a0, double *a1) {
    *a0 = 0.0;
    *a1 = 0.0;
    if (m->n < 2.0) return;
    double c[9];
    mom_covariance(m, c);
    double trace = c[0] + c[4];
    if (trace <= 1e-12) return;
    *a0 = (c[0] - c[4]) / trace;
    *a1 = 2.0 * c[1] / trace;
}

/* principal_axis (moments.rs:165-186): twenty rounds of power iteration from
 * (1, 0.7, 0.3), then the Rayleigh quotient, floored at zero. */
void oracle_principal_axis(const double c[9], double axis[3], double *lambda) {
    double v[3] = {1.0, 0.7, 0.3};
    for (int round = 0; round < 20; ++round) {
        double w[3] = {c[0] * v[0] + c[1] * v[1] + c[2] * v[2],
                       c[3] * v[0] + c[4] * v[1] + c[5] * v[2],
                       c[6] * v[0] + c[7] * v[1] + c[8] * v[2]};
        double norm = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
        if (norm < 1e-15) {
            axis[0] = 1.0;
            axis[1] = 0.0;
            axis[2] = 0.0;
            *lambda = 0.0;
            return;
        }
        v[0] = w[0] / norm;
        v[1] = w[1] / norm;
        v[2] = w[2] / norm;
    }
    double w[3] = {c[0] * v[0] + c[1] * v[1] + c[2] * v[2], c[3] * v[0] + c[4] * v[1] + c[5] * v[2],
                   c[6] * v[0] + c[7] * v[1] + c[8] * v[2]};
    double l = v[0] * w[0] + v[1] * w[1] + v[2] * w[2];
    axis[0] = v[0];
    axis[1] = v[1];
    axis[2] = v[2];
    *lambda = l > 0.0 ? l : 0.0;
}

/* Moments::length (moments.rs:152-157). */
static double mom_length(const moments_t *m) {
    if (m->n < 2.0) return 0.0;
    double c[9], axis[3], lambda;
    mom_covariance(m, c);
    oracle_principal_axis(c, axis, &lambda);
    return 4.0 * sqrt(lambda);
}

double oracle_moments_length(const double *xyz, size_t n) {
    moments_t m;
    memset(&m, 0, sizeof m);
    for (size_t i = 0; i < n; ++i) mom_add(&m, xyz + 3 * i);
    return mom_length(&m);
}

/* ---------------------------------------------------------------- lattice.rs */

struct serial_sim {
    ghip_model m;
    int64_t w, h, d;
    size_t n;
    int n_off;
    int64_t off[26][3];
    int ring_n;
    int64_t ring[26][3];
    uint32_t *labels;
    size_t n_labels, cap;
    uint32_t *volume;
    int64_t *surface;
    uint8_t *cell_type;
    moments_t *moments;
    double *nematic;  /* (Q_xx, Q_xy) per site, or NULL */
    double *nem_sum;  /* the field summed over each label's sites, 2 per label */
    double *activity;
    double *fields[GHIP_MAX_SPECIES];
    double *scratch;
    int substeps[GHIP_MAX_SPECIES];
    uint64_t rng[4];
    uint64_t mcs, accepted, attempted;
    int tracks_moments, has_motility, has_population, has_length, has_nematic;
};

/* ORDER3 in lattice.rs:42-62: dz, then dy, then dx from -1 to 1. */
static int order3(int64_t out[26][3], int planar) {
    int k = 0;
    for (int64_t dz = -1; dz <= 1; ++dz)
        for (int64_t dy = -1; dy <= 1; ++dy)
            for (int64_t dx = -1; dx <= 1; ++dx) {
                if (dx == 0 && dy == 0 && dz == 0) continue;
                if (planar && dz != 0) continue;
                out[k][0] = dx;
                out[k][1] = dy;
                out[k][2] = dz;
                ++k;
            }
    return k;
}

/* offsets_for (lattice.rs:69-88). */
static int offsets_for(int order, int64_t depth, int64_t out[26][3]) {
    static const int64_t o1[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0},
                                     {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
    static const int64_t o2[18][3] = {{1, 0, 0},  {-1, 0, 0}, {0, 1, 0},   {0, -1, 0},  {0, 0, 1},
                                      {0, 0, -1}, {1, 1, 0},  {1, -1, 0},  {-1, 1, 0},  {-1, -1, 0},
                                      {1, 0, 1},  {1, 0, -1}, {-1, 0, 1},  {-1, 0, -1}, {0, 1, 1},
                                      {0, 1, -1}, {0, -1, 1}, {0, -1, -1}};
    if (depth > 1) {
        if (order == 1) {
            memcpy(out, o1, sizeof o1);
            return 6;
        }
        if (order == 2) {
            memcpy(out, o2, sizeof o2);
            return 18;
        }
        return order3(out, 0);
    }
    if (order >= 2) return order3(out, 1);
    int k = 0;
    for (int i = 0; i < 6; ++i)
        if (o1[i][2] == 0) {
            memcpy(out[k], o1[i], sizeof o1[i]);
            ++k;
        }
    return k;
}

/* wrap (lattice.rs:96-111). */
static inline int64_t wrap(int64_t v, int64_t span) {
    if (v >= 0) {
        if (v < span) return v;
        if (v < 2 * span) return v - span;
        return v % span;
    }
    if (v >= -span) return v + span;
    int64_t r = v % span;
    return r < 0 ? r + span : r;
}

static inline size_t s_index(const serial_sim *s, int64_t x, int64_t y, int64_t z) {
    return (size_t)((wrap(z, s->d) * s->h + wrap(y, s->h)) * s->w + wrap(x, s->w));
}

static inline void s_coords(const serial_sim *s, size_t index, int64_t *x, int64_t *y,
                            int64_t *z) {
    size_t plane = (size_t)(s->w * s->h);
    *z = (int64_t)(index / plane);
    size_t rest = index % plane;
    *x = (int64_t)(rest % (size_t)s->w);
    *y = (int64_t)(rest / (size_t)s->w);
}

static inline uint8_t type_of(const serial_sim *s, uint32_t label) { return s->cell_type[label]; }

static inline double contact_energy(const serial_sim *s, uint8_t a, uint8_t b) {
    return s->m.contact[a * s->m.n_types + b];
}

/* ---------------------------------------------------------------- cpu.rs */

/* count_surface (cpu.rs:46-59). */
static void count_surface(const serial_sim *s, int64_t *surface, size_t n_labels) {
    memset(surface, 0, n_labels * sizeof(int64_t));
    for (size_t i = 0; i < s->n; ++i) {
        int64_t x, y, z;
        s_coords(s, i, &x, &y, &z);
        uint32_t l = s->labels[i];
        for (int k = 0; k < s->n_off; ++k) {
            size_t nb = s_index(s, x + s->off[k][0], y + s->off[k][1], z + s->off[k][2]);
            if (nb != i && s->labels[nb] != l) surface[l] += 1;
        }
    }
}

static void extent_of(const serial_sim *s, double ext[3]) {
    ext[0] = (double)s->w;
    ext[1] = (double)s->h;
    ext[2] = (double)s->d;
}

/* count_moments (cpu.rs:11-21): sites in index order, so each label's
 * anchor is its first site. */
static void count_moments(const serial_sim *s, moments_t *mom, size_t n_labels) {
    double ext[3];
    extent_of(s, ext);
    memset(mom, 0, n_labels * sizeof(moments_t));
    for (size_t i = 0; i < s->n; ++i) {
        uint32_t l = s->labels[i];
        int64_t x, y, z;
        s_coords(s, i, &x, &y, &z);
        double site[3];
        mom_unwrap(&mom[l], (double)x, (double)y, (double)z, ext, site);
        mom_add(&mom[l], site);
    }
}

static void ensure_capacity(serial_sim *s, size_t want) {
    if (want <= s->cap) return;
    size_t cap = s->cap ? s->cap : 16;
    while (cap < want) cap *= 2;
    s->volume = realloc(s->volume, cap * sizeof(uint32_t));
    s->surface = realloc(s->surface, cap * sizeof(int64_t));
    s->cell_type = realloc(s->cell_type, cap);
    s->moments = realloc(s->moments, cap * sizeof(moments_t));
    s->nem_sum = realloc(s->nem_sum, cap * 2 * sizeof(double));
    s->cap = cap;
}

/* substeps_for (field.rs:155-160), with an explicit count taking precedence. */
int serial_substeps(uint32_t species_index, const ghip_model *m) {
    const ghip_species *sp = &m->species[species_index];
    if (sp->substeps > 0) return sp->substeps;
    if (sp->diffusion <= 0.0) return 1;
    double dims = m->depth > 1 ? 3.0 : 2.0;
    double limit = 1.0 / (2.0 * dims);
    double c = ceil(sp->diffusion / limit);
    return c < 1.0 ? 1 : (int)c;
}

static void count_nematic(const serial_sim *s, double *sums, size_t n_labels);

/* Simulation::assemble (cpu.rs:178-217) from a label field and a type per
 * label (cpu.rs:220-224). */
serial_sim *serial_create(const ghip_model *m, const uint32_t *labels, const uint8_t *cell_type,
                          size_t n_labels, const double *fields) {
    serial_sim *s = calloc(1, sizeof *s);
    s->m = *m;
    s->w = m->width;
    s->h = m->height;
    s->d = m->depth;
    s->n = (size_t)(s->w * s->h * s->d);
    s->n_off = offsets_for((int)m->neighbour_order, s->d, s->off);
    int has_conn = 0, has_len = 0;
    for (uint32_t t = 0; t < m->n_types; ++t) {
        const ghip_cell_type *ct = &m->types[t];
        if (ct->connected) has_conn = 1;
        if (ct->lambda_length != 0.0) has_len = 1;
        if (ct->lambda_activity != 0.0 && ct->max_activity > 0.0) s->has_motility = 1;
        if (ct->division_volume > 0.0 || ct->death_rate > 0.0) s->has_population = 1;
    }
    s->tracks_moments = has_len;
    s->has_length = has_len;
    s->ring_n = has_conn ? order3(s->ring, s->d <= 1) : 0;
    s->labels = malloc(s->n * sizeof(uint32_t));
    memcpy(s->labels, labels, s->n * sizeof(uint32_t));
    s->n_labels = n_labels;
    ensure_capacity(s, n_labels);
    memcpy(s->cell_type, cell_type, n_labels);
    memset(s->volume, 0, n_labels * sizeof(uint32_t));
    for (size_t i = 0; i < s->n; ++i) s->volume[s->labels[i]] += 1;
    count_surface(s, s->surface, n_labels);
    count_moments(s, s->moments, n_labels);
    count_nematic(s, s->nem_sum, n_labels);
    s->activity = calloc(s->n, sizeof(double));
    for (uint32_t k = 0; k < m->n_species; ++k) {
        s->fields[k] = malloc(s->n * sizeof(double));
        for (size_t i = 0; i < s->n; ++i)
            s->fields[k][i] = fields ? fields[k * s->n + i] : m->species[k].initial;
        s->substeps[k] = serial_substeps(k, m);
    }
    s->scratch = calloc(s->n, sizeof(double));
    oracle_xoshiro_seed(m->seed, s->rng);
    return s;
}

void serial_destroy(serial_sim *s) {
    if (!s) return;
    free(s->labels);
    free(s->volume);
    free(s->surface);
    free(s->cell_type);
    free(s->moments);
    free(s->nematic);
    free(s->nem_sum);
    free(s->activity);
    for (int k = 0; k < GHIP_MAX_SPECIES; ++k) free(s->fields[k]);
    free(s->scratch);
    free(s);
}

/* volume_term (cpu.rs:413-425). */
static double volume_term(const serial_sim *s, uint32_t label, int64_t change) {
    if (label == 0) return 0.0;
    const ghip_cell_type *spec = &s->m.types[type_of(s, label)];
    if (spec->lambda_volume == 0.0) return 0.0;
    double v = (double)s->volume[label];
    double after = v + (double)change;
    double a = after - spec->target_volume, b = v - spec->target_volume;
    return spec->lambda_volume * (a * a - b * b);
}

/* surface_term (cpu.rs:399-411). */
static double surface_term(const serial_sim *s, uint32_t label, int64_t change) {
    if (label == 0 || change == 0) return 0.0;
    const ghip_cell_type *spec = &s->m.types[type_of(s, label)];
    if (spec->lambda_surface == 0.0) return 0.0;
    double sv = (double)s->surface[label];
    double after = sv + (double)change;
    double a = after - spec->target_surface, b = sv - spec->target_surface;
    return spec->lambda_surface * (a * a - b * b);
}

/* length_term (cpu.rs:291-305). */
static double length_term(const serial_sim *s, uint32_t label, const double site[3], double sign) {
    if (label == 0) return 0.0;
    const ghip_cell_type *spec = &s->m.types[type_of(s, label)];
    if (spec->lambda_length == 0.0) return 0.0;
    const moments_t *mom = &s->moments[label];
    double ext[3], un[3];
    extent_of(s, ext);
    mom_unwrap(mom, site[0], site[1], site[2], ext, un);
    double before = mom_length(mom);
    moments_t with = mom_with(mom, un, sign);
    double after = mom_length(&with);
    double a = after - spec->target_length, b = before - spec->target_length;
    return spec->lambda_length * (a * a - b * b);
}

/* count_nematic (cpu.rs:24-34): the field summed over each label's sites. */
static void count_nematic(const serial_sim *s, double *sums, size_t n_labels) {
    memset(sums, 0, n_labels * 2 * sizeof(double));
    if (!s->nematic) return;
    for (size_t i = 0; i < s->n; ++i) {
        uint32_t l = s->labels[i];
        sums[2 * l] += s->nematic[2 * i];
        sums[2 * l + 1] += s->nematic[2 * i + 1];
    }
}

/* nematic_term (cpu.rs:315-332): the change in -lambda F . a of `label` when
 * the site at `index` joins (sign +1) or leaves (sign -1). */
static double nematic_term(const serial_sim *s, uint32_t label, size_t index, const double site[3],
                           double sign) {
    if (label == 0) return 0.0;
    const ghip_cell_type *spec = &s->m.types[type_of(s, label)];
    if (spec->lambda_nematic == 0.0) return 0.0;
    const moments_t *mom = &s->moments[label];
    double ext[3], un[3];
    extent_of(s, ext);
    mom_unwrap(mom, site[0], site[1], site[2], ext, un);
    double a0, a1, b0, b1;
    mom_anisotropy(mom, &a0, &a1);
    moments_t with = mom_with(mom, un, sign);
    mom_anisotropy(&with, &b0, &b1);
    const double *f = s->nem_sum + 2 * label;
    const double *q = s->nematic + 2 * index;
    double before = f[0] * a0 + f[1] * a1;
    double after = (f[0] + sign * q[0]) * b0 + (f[1] + sign * q[1]) * b1;
    return -spec->lambda_nematic * (after - before);
}

/* delta_energy (cpu.rs:242-288). */
double serial_delta_energy(const serial_sim *s, size_t target, uint32_t new_label) {
    uint32_t old = s->labels[target];
    if (old == new_label) return 0.0;
    int64_t tx, ty, tz;
    s_coords(s, target, &tx, &ty, &tz);
    uint8_t told = type_of(s, old), tnew = type_of(s, new_label);
    double delta = 0.0;
    int64_t like_old = 0, like_new = 0, bonds = 0;
    for (int k = 0; k < s->n_off; ++k) {
        size_t nb = s_index(s, tx + s->off[k][0], ty + s->off[k][1], tz + s->off[k][2]);
        if (nb == target) continue;
        bonds += 1;
        uint32_t ln = s->labels[nb];
        uint8_t tn = type_of(s, ln);
        if (ln != old)
            delta -= contact_energy(s, told, tn);
        else
            like_old += 1;
        if (ln != new_label)
            delta += contact_energy(s, tnew, tn);
        else
            like_new += 1;
    }
    delta += volume_term(s, old, -1);
    delta += volume_term(s, new_label, 1);
    if (s->tracks_moments) {
        double site[3] = {(double)tx, (double)ty, (double)tz};
        delta += length_term(s, old, site, -1.0);
        delta += length_term(s, new_label, site, 1.0);
        if (s->has_nematic) {
            delta += nematic_term(s, old, target, site, -1.0);
            delta += nematic_term(s, new_label, target, site, 1.0);
        }
    }
    delta += surface_term(s, old, 2 * like_old - bonds);
    delta += surface_term(s, new_label, bonds - 2 * like_new);
    return delta;
}

/* chemotaxis_work (cpu.rs:381-395). */
double serial_chemotaxis_work(const serial_sim *s, uint32_t new_label, size_t target,
                              size_t source) {
    if (new_label == 0 || s->m.n_species == 0) return 0.0;
    uint8_t t = type_of(s, new_label);
    double work = 0.0;
    for (uint32_t k = 0; k < s->m.n_species; ++k) {
        double lambda = s->m.chemotaxis[t][k];
        if (lambda == 0.0) continue;
        work -= lambda * (s->fields[k][target] - s->fields[k][source]);
    }
    return work;
}

/* Activity::neighbourhood_mean (motility.rs:56-78). */
static double neighbourhood_mean(const serial_sim *s, size_t site, uint32_t label) {
    int64_t x, y, z;
    s_coords(s, site, &x, &y, &z);
    double here = s->activity[site] > 0.0 ? s->activity[site] : 0.0;
    double log_sum = log(here);
    double count = 1.0;
    for (int k = 0; k < s->n_off; ++k) {
        size_t nb = s_index(s, x + s->off[k][0], y + s->off[k][1], z + s->off[k][2]);
        if (nb == site || s->labels[nb] != label) continue;
        double value = s->activity[nb];
        if (value <= 0.0) return 0.0;
        log_sum += log(value);
        count += 1.0;
    }
    if (!isfinite(log_sum)) return 0.0;
    return exp(log_sum / count);
}

double serial_activity_mean(const serial_sim *s, size_t site, uint32_t label) {
    return neighbourhood_mean(s, site, label);
}

static double step_fold(int64_t a, int64_t b, double span) {
    double raw = (double)(a - b);
    if (raw > span / 2.0) return raw - span;
    if (raw < -span / 2.0) return raw + span;
    return raw;
}

/* move_work (cpu.rs:341-371). */
double serial_move_work(const serial_sim *s, uint32_t new_label, uint32_t old_label,
                        size_t target, size_t source) {
    double work = 0.0;
    if (new_label != 0) {
        const ghip_cell_type *spec = &s->m.types[type_of(s, new_label)];
        if (spec->lambda_activity != 0.0 && spec->max_activity > 0.0) {
            double into = neighbourhood_mean(s, source, new_label);
            double out_of = neighbourhood_mean(s, target, old_label);
            work -= spec->lambda_activity / spec->max_activity * (into - out_of);
        }
        if (spec->external[0] != 0.0 || spec->external[1] != 0.0 || spec->external[2] != 0.0) {
            int64_t tx, ty, tz, sx, sy, sz;
            s_coords(s, target, &tx, &ty, &tz);
            s_coords(s, source, &sx, &sy, &sz);
            work -= spec->external[0] * step_fold(tx, sx, (double)s->w) +
                    spec->external[1] * step_fold(ty, sy, (double)s->h) +
                    spec->external[2] * step_fold(tz, sz, (double)s->d);
        }
    }
    return work;
}

/* locally_connected (connectivity.rs:43-84). */
int serial_locally_connected(const serial_sim *s, size_t target, uint32_t label) {
    int64_t x, y, z;
    s_coords(s, target, &x, &y, &z);
    int64_t members[26][3];
    int count = 0;
    for (int k = 0; k < s->ring_n; ++k) {
        size_t site = s_index(s, x + s->ring[k][0], y + s->ring[k][1], z + s->ring[k][2]);
        if (site != target && s->labels[site] == label) {
            memcpy(members[count], s->ring[k], sizeof members[count]);
            ++count;
        }
    }
    if (count <= 1) return 1;
    int seen[26] = {0}, stack[26];
    int top = 1, reached = 1;
    stack[0] = 0;
    seen[0] = 1;
    while (top > 0) {
        int here = stack[--top];
        for (int other = 0; other < count; ++other) {
            if (seen[other]) continue;
            int64_t dx = members[here][0] - members[other][0];
            int64_t dy = members[here][1] - members[other][1];
            int64_t dz = members[here][2] - members[other][2];
            if (llabs(dx) <= 1 && llabs(dy) <= 1 && llabs(dz) <= 1) {
                seen[other] = 1;
                stack[top++] = other;
                ++reached;
            }
        }
    }
    return reached == count;
}

/* attempt (cpu.rs:432-517). */
int serial_attempt(serial_sim *s) {
    size_t target = (size_t)oracle_xoshiro_below(s->rng, s->n);
    size_t pick = (size_t)oracle_xoshiro_below(s->rng, (uint64_t)s->n_off);
    int64_t tx, ty, tz;
    s_coords(s, target, &tx, &ty, &tz);
    size_t source = s_index(s, tx + s->off[pick][0], ty + s->off[pick][1], tz + s->off[pick][2]);
    uint32_t old = s->labels[target];
    uint32_t nl = s->labels[source];
    s->attempted += 1;
    if (old == nl) return 0;
    if (old != 0 && s->volume[old] <= 1) return 0;
    if (old != 0 && s->ring_n > 0 && s->m.types[type_of(s, old)].connected &&
        !serial_locally_connected(s, target, old))
        return 0;
    double delta = serial_delta_energy(s, target, nl) + serial_chemotaxis_work(s, nl, target, source) +
                   serial_move_work(s, nl, old, target, source);
    int accept = delta <= 0.0 || oracle_xoshiro_f64(s->rng) < exp(-delta / s->m.temperature);
    if (!accept) return 0;
    int64_t like_old = 0, like_new = 0, bonds = 0;
    for (int k = 0; k < s->n_off; ++k) {
        size_t nb = s_index(s, tx + s->off[k][0], ty + s->off[k][1], tz + s->off[k][2]);
        if (nb == target) continue;
        bonds += 1;
        uint32_t ln = s->labels[nb];
        if (ln == old) like_old += 1;
        if (ln == nl) like_new += 1;
    }
    s->surface[old] += 2 * like_old - bonds;
    s->surface[nl] += bonds - 2 * like_new;
    if (s->tracks_moments) {
        double site[3] = {(double)tx, (double)ty, (double)tz}, ext[3], un[3];
        extent_of(s, ext);
        mom_unwrap(&s->moments[old], site[0], site[1], site[2], ext, un);
        mom_remove(&s->moments[old], un);
        mom_unwrap(&s->moments[nl], site[0], site[1], site[2], ext, un);
        mom_add(&s->moments[nl], un);
    }
    if (s->has_nematic) {
        /* cpu.rs:488-495 */
        const double *q = s->nematic + 2 * target;
        s->nem_sum[2 * old] -= q[0];
        s->nem_sum[2 * old + 1] -= q[1];
        s->nem_sum[2 * nl] += q[0];
        s->nem_sum[2 * nl + 1] += q[1];
    }
    if (nl != 0) {
        const ghip_cell_type *spec = &s->m.types[type_of(s, nl)];
        if (spec->max_activity > 0.0) s->activity[target] = spec->max_activity;
    }
    s->labels[target] = nl;
    s->volume[old] -= 1;
    s->volume[nl] += 1;
    s->accepted += 1;
    return 1;
}

/* Fields::diffuse and one_substep (field.rs:99-130). */
static void fields_diffuse(serial_sim *s) {
    int faces = s->d > 1 ? 6 : 4;
    static const int64_t f[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0},
                                    {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
    for (uint32_t k = 0; k < s->m.n_species; ++k) {
        int steps = s->substeps[k];
        double d = s->m.species[k].diffusion / (double)steps;
        double decay = s->m.species[k].decay / (double)steps;
        double *v = s->fields[k];
        for (int st = 0; st < steps; ++st) {
            for (size_t here = 0; here < s->n; ++here) {
                int64_t x, y, z;
                s_coords(s, here, &x, &y, &z);
                double lap = -(double)faces * v[here];
                for (int q = 0; q < faces; ++q) lap += v[s_index(s, x + f[q][0], y + f[q][1], z + f[q][2])];
                s->scratch[here] = v[here] + d * lap - decay * v[here];
            }
            memcpy(v, s->scratch, s->n * sizeof(double));
        }
    }
}

/* Fields::exchange (field.rs:133-149). */
static void fields_exchange(serial_sim *s) {
    for (uint32_t k = 0; k < s->m.n_species; ++k) {
        double *v = s->fields[k];
        for (size_t i = 0; i < s->n; ++i) {
            uint32_t l = s->labels[i];
            if (l == 0) continue;
            uint8_t t = s->cell_type[l];
            double sec = s->m.secretion[t][k], up = s->m.uptake[t][k];
            v[i] += sec;
            v[i] -= up * v[i];
            if (v[i] < 0.0) v[i] = 0.0;
        }
    }
}

void serial_fields_step(serial_sim *s) {
    if (s->m.n_species == 0) return;
    fields_diffuse(s);
    fields_exchange(s);
}

/* Simulation::kill (cpu.rs:715-725). */
static void kill_label(serial_sim *s, uint32_t label) {
    for (size_t i = 0; i < s->n; ++i)
        if (s->labels[i] == label) s->labels[i] = 0;
    uint32_t gone = s->volume[label];
    s->volume[label] = 0;
    s->volume[0] += gone;
    s->surface[label] = 0;
}

/* Simulation::divide (cpu.rs:679-712) via unwrapped_sites (cpu.rs:659-673). */
static int divide_label(serial_sim *s, uint32_t label) {
    double ext[3];
    extent_of(s, ext);
    moments_t mom;
    memset(&mom, 0, sizeof mom);
    size_t count = 0;
    for (size_t i = 0; i < s->n; ++i)
        if (s->labels[i] == label) ++count;
    if (count < 4) return 0;
    size_t *sites = malloc(count * sizeof(size_t));
    double *un = malloc(3 * count * sizeof(double));
    size_t j = 0;
    for (size_t i = 0; i < s->n; ++i) {
        if (s->labels[i] != label) continue;
        int64_t x, y, z;
        s_coords(s, i, &x, &y, &z);
        mom_unwrap(&mom, (double)x, (double)y, (double)z, ext, un + 3 * j);
        mom_add(&mom, un + 3 * j);
        sites[j++] = i;
    }
    double centre[3] = {mom.s[0] / mom.n, mom.s[1] / mom.n, mom.s[2] / mom.n};
    double c[9], axis[3], lambda;
    mom_covariance(&mom, c);
    oracle_principal_axis(c, axis, &lambda);
    uint32_t daughter = (uint32_t)s->n_labels;
    uint32_t moved = 0;
    for (size_t q = 0; q < count; ++q) {
        double along = (un[3 * q] - centre[0]) * axis[0] + (un[3 * q + 1] - centre[1]) * axis[1] +
                       (un[3 * q + 2] - centre[2]) * axis[2];
        if (along > 0.0) {
            s->labels[sites[q]] = daughter;
            ++moved;
        }
    }
    if (moved == 0 || moved == count) {
        for (size_t q = 0; q < count; ++q) s->labels[sites[q]] = label;
        free(sites);
        free(un);
        return 0;
    }
    free(sites);
    free(un);
    ensure_capacity(s, s->n_labels + 1);
    s->volume[label] -= moved;
    s->volume[s->n_labels] = moved;
    s->surface[s->n_labels] = 0;
    memset(&s->moments[s->n_labels], 0, sizeof(moments_t));
    s->cell_type[s->n_labels] = s->cell_type[label];
    s->n_labels += 1;
    return 1;
}

/* divide_and_die (cpu.rs:620-655). */
void serial_divide_and_die(serial_sim *s, size_t *divided_out, size_t *died_out) {
    size_t divided = 0, died = 0;
    size_t n0 = s->n_labels;
    uint32_t *live = malloc((n0 + 1) * sizeof(uint32_t));
    size_t nl = 0;
    for (uint32_t l = 1; l < n0; ++l)
        if (s->volume[l] > 0) live[nl++] = l;
    for (size_t q = 0; q < nl; ++q) {
        uint32_t l = live[q];
        const ghip_cell_type *spec = &s->m.types[type_of(s, l)];
        if (spec->division_volume > 0.0 && (double)s->volume[l] >= spec->division_volume &&
            divide_label(s, l))
            ++divided;
    }
    free(live);
    live = malloc((s->n_labels + 1) * sizeof(uint32_t));
    nl = 0;
    for (uint32_t l = 1; l < s->n_labels; ++l)
        if (s->volume[l] > 0) live[nl++] = l;
    for (size_t q = 0; q < nl; ++q) {
        uint32_t l = live[q];
        const ghip_cell_type *spec = &s->m.types[type_of(s, l)];
        if (spec->death_rate > 0.0 && oracle_xoshiro_f64(s->rng) < spec->death_rate) {
            kill_label(s, l);
            ++died;
        }
    }
    free(live);
    if (divided > 0 || died > 0) {
        count_surface(s, s->surface, s->n_labels);
        count_moments(s, s->moments, s->n_labels);
        if (s->has_nematic) count_nematic(s, s->nem_sum, s->n_labels); /* cpu.rs:652 */
    }
    if (divided_out) *divided_out = divided;
    if (died_out) *died_out = died;
}

/* step (cpu.rs:522-538). */
void serial_step(serial_sim *s, uint64_t steps) {
    for (uint64_t q = 0; q < steps; ++q) {
        for (size_t i = 0; i < s->n; ++i) serial_attempt(s);
        if (s->m.n_species > 0) {
            fields_diffuse(s);
            fields_exchange(s);
        }
        if (s->has_motility)
            for (size_t i = 0; i < s->n; ++i)
                if (s->activity[i] > 0.0) s->activity[i] -= 1.0;
        if (s->has_population) serial_divide_and_die(s, NULL, NULL);
        s->mcs += 1;
    }
}

/* energy (cpu.rs:545-581). */
double serial_energy(const serial_sim *s) {
    double contact = 0.0;
    for (size_t i = 0; i < s->n; ++i) {
        int64_t x, y, z;
        s_coords(s, i, &x, &y, &z);
        uint32_t l = s->labels[i];
        for (int k = 0; k < s->n_off; ++k) {
            size_t nb = s_index(s, x + s->off[k][0], y + s->off[k][1], z + s->off[k][2]);
            uint32_t ln = s->labels[nb];
            if (ln != l) contact += contact_energy(s, type_of(s, l), type_of(s, ln));
        }
    }
    contact /= 2.0;
    int64_t *counted = malloc(s->n_labels * sizeof(int64_t));
    count_surface(s, counted, s->n_labels);
    double volume = 0.0, surface = 0.0, length = 0.0, nematic = 0.0;
    for (size_t l = 1; l < s->n_labels; ++l) {
        const ghip_cell_type *spec = &s->m.types[s->cell_type[l]];
        double dv = (double)s->volume[l] - spec->target_volume;
        volume += spec->lambda_volume * (dv * dv);
        if (spec->lambda_length != 0.0) {
            double dl = mom_length(&s->moments[l]) - spec->target_length;
            length += spec->lambda_length * (dl * dl);
        }
        double ds = (double)counted[l] - spec->target_surface;
        surface += spec->lambda_surface * (ds * ds);
        if (s->has_nematic && spec->lambda_nematic != 0.0) {
            double a0, a1;
            mom_anisotropy(&s->moments[l], &a0, &a1);
            const double *f = s->nem_sum + 2 * l;
            nematic -= spec->lambda_nematic * (f[0] * a0 + f[1] * a1);
        }
    }
    free(counted);
    return contact + volume + surface + length + nematic;
}

size_t serial_n_labels(const serial_sim *s) { return s->n_labels; }
uint64_t serial_mcs(const serial_sim *s) { return s->mcs; }
void serial_counters(const serial_sim *s, uint64_t *attempted, uint64_t *accepted) {
    *attempted = s->attempted;
    *accepted = s->accepted;
}
void serial_labels(const serial_sim *s, uint32_t *out) { memcpy(out, s->labels, s->n * 4); }
void serial_set_labels(serial_sim *s, const uint32_t *labels) {
    memcpy(s->labels, labels, s->n * 4);
}
void serial_volumes(const serial_sim *s, uint32_t *out) {
    memcpy(out, s->volume, s->n_labels * sizeof(uint32_t));
}
void serial_surfaces(const serial_sim *s, int64_t *out) {
    memcpy(out, s->surface, s->n_labels * sizeof(int64_t));
}
void serial_cell_types(const serial_sim *s, uint8_t *out) { memcpy(out, s->cell_type, s->n_labels); }
void serial_recount_volumes(const serial_sim *s, uint32_t *out) {
    memset(out, 0, s->n_labels * sizeof(uint32_t));
    for (size_t i = 0; i < s->n; ++i) out[s->labels[i]] += 1;
}
void serial_recount_surfaces(const serial_sim *s, int64_t *out) {
    count_surface(s, out, s->n_labels);
}
/* Rebuild every per-label counter from the lattice, as a caller that edits
 * the labels does (e.g. glazier-tests/tests/chemotaxis.rs:72-73). */
void serial_recount(serial_sim *s) {
    serial_recount_volumes(s, s->volume);
    count_surface(s, s->surface, s->n_labels);
    count_moments(s, s->moments, s->n_labels);
    count_nematic(s, s->nem_sum, s->n_labels);
}

/* Model::nematic for this run (2 values per site), or NULL to drop it; the
 * coupling is on when some type has a lambda_nematic (model.rs:120-124),
 * and the moments are kept for it (model.rs:127-130). */
void serial_set_nematic(serial_sim *s, const double *q) {
    free(s->nematic);
    s->nematic = NULL;
    s->has_nematic = 0;
    if (q) {
        s->nematic = malloc(s->n * 2 * sizeof(double));
        memcpy(s->nematic, q, s->n * 2 * sizeof(double));
        for (uint32_t t = 0; t < s->m.n_types; ++t)
            if (s->m.types[t].lambda_nematic != 0.0) s->has_nematic = 1;
    }
    s->tracks_moments = s->has_length || s->has_nematic;
    count_nematic(s, s->nem_sum, s->n_labels);
}
void serial_nematic_sums(const serial_sim *s, double *out) {
    memcpy(out, s->nem_sum, s->n_labels * 2 * sizeof(double));
}
void serial_recounted_nematic(const serial_sim *s, double *out) { count_nematic(s, out, s->n_labels); }
/* Each label's anisotropy (moments.rs:131-141) from moments recounted from
 * the lattice (count_moments, cpu.rs:11-21), 2 values per label. */
void serial_recounted_anisotropy(const serial_sim *s, double *out) {
    moments_t *m = malloc(s->n_labels * sizeof(moments_t));
    count_moments(s, m, s->n_labels);
    for (size_t l = 0; l < s->n_labels; ++l) mom_anisotropy(&m[l], &out[2 * l], &out[2 * l + 1]);
    free(m);
}
void serial_lengths(const serial_sim *s, double *out) {
    for (size_t l = 0; l < s->n_labels; ++l) out[l] = mom_length(&s->moments[l]);
}
void serial_recounted_lengths(const serial_sim *s, double *out) {
    moments_t *m = malloc(s->n_labels * sizeof(moments_t));
    count_moments(s, m, s->n_labels);
    for (size_t l = 0; l < s->n_labels; ++l) out[l] = mom_length(&m[l]);
    free(m);
}
void serial_field(const serial_sim *s, uint32_t species, double *out) {
    memcpy(out, s->fields[species], s->n * sizeof(double));
}
void serial_set_field(serial_sim *s, uint32_t species, const double *values) {
    memcpy(s->fields[species], values, s->n * sizeof(double));
}
void serial_activity(const serial_sim *s, double *out) {
    memcpy(out, s->activity, s->n * sizeof(double));
}
void serial_set_activity(serial_sim *s, const double *values) {
    memcpy(s->activity, values, s->n * sizeof(double));
}
/* Activity::decay (motility.rs:38-45): every remembered site loses one. */
void serial_decay_activity(serial_sim *s) {
    for (size_t i = 0; i < s->n; ++i)
        if (s->activity[i] > 0.0) s->activity[i] -= 1.0;
}

/* A deep copy, for what-if checks (the tests clone a Simulation the same
 * way, e.g. glazier-tests/tests/length.rs:50-64). */
serial_sim *serial_clone(const serial_sim *s) {
    serial_sim *c = malloc(sizeof *c);
    *c = *s;
    c->labels = malloc(s->n * 4);
    memcpy(c->labels, s->labels, s->n * 4);
    c->volume = malloc(s->cap * sizeof(uint32_t));
    memcpy(c->volume, s->volume, s->cap * sizeof(uint32_t));
    c->surface = malloc(s->cap * sizeof(int64_t));
    memcpy(c->surface, s->surface, s->cap * sizeof(int64_t));
    c->cell_type = malloc(s->cap);
    memcpy(c->cell_type, s->cell_type, s->cap);
    c->moments = malloc(s->cap * sizeof(moments_t));
    memcpy(c->moments, s->moments, s->cap * sizeof(moments_t));
    c->nem_sum = malloc(s->cap * 2 * sizeof(double));
    memcpy(c->nem_sum, s->nem_sum, s->cap * 2 * sizeof(double));
    c->nematic = NULL;
    if (s->nematic) {
        c->nematic = malloc(s->n * 2 * sizeof(double));
        memcpy(c->nematic, s->nematic, s->n * 2 * sizeof(double));
    }
    c->activity = malloc(s->n * sizeof(double));
    memcpy(c->activity, s->activity, s->n * sizeof(double));
    for (int k = 0; k < GHIP_MAX_SPECIES; ++k) {
        c->fields[k] = NULL;
        if (s->fields[k]) {
            c->fields[k] = malloc(s->n * sizeof(double));
            memcpy(c->fields[k], s->fields[k], s->n * sizeof(double));
        }
    }
    c->scratch = calloc(s->n, sizeof(double));
    return c;
}

/* The state after `target` takes `new_label`, built as the reference's delta
 * tests build it (length.rs:50-64): label, volumes, and moments moved by one
 * site; the surfaces are left for energy() to recount. */
void serial_apply_copy(serial_sim *s, size_t target, uint32_t new_label) {
    uint32_t old = s->labels[target];
    int64_t tx, ty, tz;
    s_coords(s, target, &tx, &ty, &tz);
    double ext[3], un[3];
    extent_of(s, ext);
    mom_unwrap(&s->moments[old], (double)tx, (double)ty, (double)tz, ext, un);
    mom_remove(&s->moments[old], un);
    mom_unwrap(&s->moments[new_label], (double)tx, (double)ty, (double)tz, ext, un);
    mom_add(&s->moments[new_label], un);
    if (s->nematic) {
        const double *q = s->nematic + 2 * target;
        s->nem_sum[2 * old] -= q[0];
        s->nem_sum[2 * old + 1] -= q[1];
        s->nem_sum[2 * new_label] += q[0];
        s->nem_sum[2 * new_label + 1] += q[1];
    }
    s->labels[target] = new_label;
    s->volume[old] -= 1;
    s->volume[new_label] += 1;
}
