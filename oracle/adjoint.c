/*
 * adjoint.c - f64 replay of the field step over a recorded cell-type
 * trajectory, and its hand-derived adjoint.
 *
 * TEST INFRASTRUCTURE ONLY: the tests use it to check the device adjoint and
 * the benchmark times it as the CPU baseline of the gradient; nothing in the
 * engine links or calls it.
 *
 * The forward step restates Fields::diffuse and Fields::exchange
 * (glazier-core/src/field.rs:99-149) exactly as serial.c does
 * (fields_diffuse, fields_exchange): per species `substeps` explicit-Euler
 * substeps with d = D / substeps and k = decay / substeps,
 *
 *     lap = -faces * v[x] + v[x+1] + v[x-1] + v[y+1] + v[y-1] (+ v[z+1] + v[z-1])
 *     v'  = v + d * lap - k * v
 *
 * then on every owned site (type map entry t != 0)
 *
 *     v += sec[t];  v -= up[t] * v;  v = max(v, 0)
 *
 * with the rates taken from a ghip_field_params record and the owner's type
 * from the step's entry of the trajectory. The label dynamics are held fixed:
 * only the fields depend on the parameters.
 *
 * The adjoint is the transpose of that map. The Laplacian of a periodic
 * lattice is symmetric, so for a substep
 *
 *     lambda  = lambda' + d * lap(lambda') - k * lambda'
 *     dL/dd  += sum_x lambda'(x) * lap(v)(x),   dL/dk -= sum_x lambda'(x) * v(x)
 *
 * and for the exchange, with w = (v + s) - up * (v + s) and pass = !(w < 0),
 *
 *     lambda     = pass ? lambda'' * (1 - up) : 0          (owned sites)
 *     dL/ds[t]  += pass * lambda'' * (1 - up)
 *     dL/dup[t] -= pass * lambda'' * (v + s)
 *
 * and dL/dD = dL/dd / substeps, dL/ddecay = dL/dk / substeps. Where the clamp
 * sits exactly at zero the right derivative is taken.
 */
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

typedef struct {
    int w, h, d, faces;
    size_t n;
} grid_t;

static grid_t grid_of(int w, int h, int d) {
    grid_t g = {w, h, d, d > 1 ? 6 : 4, (size_t)w * h * d};
    return g;
}

static size_t neighbour(const grid_t *g, size_t i, int q) {
    const size_t plane = (size_t)g->w * g->h;
    int z = (int)(i / plane), y = (int)((i % plane) / g->w), x = (int)(i % g->w);
    switch (q) {
        case 0: x = x + 1 == g->w ? 0 : x + 1; break;
        case 1: x = x == 0 ? g->w - 1 : x - 1; break;
        case 2: y = y + 1 == g->h ? 0 : y + 1; break;
        case 3: y = y == 0 ? g->h - 1 : y - 1; break;
        case 4: z = z + 1 == g->d ? 0 : z + 1; break;
        default: z = z == 0 ? g->d - 1 : z - 1; break;
    }
    return ((size_t)z * g->h + y) * g->w + x;
}

static double laplacian(const grid_t *g, const double *v, size_t i) {
    double lap = -(double)g->faces * v[i];
    for (int q = 0; q < g->faces; ++q) lap += v[neighbour(g, i, q)];
    return lap;
}

/* One substep of one species: out = in + d*lap(in) - k*in. */
static void substep(const grid_t *g, double d, double k, const double *in, double *out) {
    for (size_t i = 0; i < g->n; ++i) out[i] = in[i] + d * laplacian(g, in, i) - k * in[i];
}

/* One whole step of every species, from `in` (species-major) into `out`. */
static void step_all(const grid_t *g, uint32_t n_species, const int *substeps,
                     const ghip_field_params *p, const uint8_t *types, const double *in,
                     double *out, double *scratch) {
    for (uint32_t s = 0; s < n_species; ++s) {
        const double *src = in + s * g->n;
        double *dst = out + s * g->n;
        const double d = p->diffusion[s] / (double)substeps[s];
        const double k = p->decay[s] / (double)substeps[s];
        memcpy(dst, src, g->n * sizeof(double));
        for (int st = 0; st < substeps[s]; ++st) {
            substep(g, d, k, dst, scratch);
            memcpy(dst, scratch, g->n * sizeof(double));
        }
        for (size_t i = 0; i < g->n; ++i) {
            const uint8_t t = types[i];
            if (t == 0) continue;
            double v = dst[i];
            v += p->secretion[t][s];
            v -= p->uptake[t][s] * v;
            if (v < 0.0) v = 0.0;
            dst[i] = v;
        }
    }
}

void adj_forward(int w, int h, int d, uint32_t n_species, const int *substeps,
                 const ghip_field_params *p, const uint8_t *types, uint64_t steps,
                 const double *u0, double *out) {
    grid_t g = grid_of(w, h, d);
    const size_t all = g.n * n_species;
    double *cur = malloc(all * sizeof(double)), *nxt = malloc(all * sizeof(double));
    double *scratch = malloc(g.n * sizeof(double));
    memcpy(cur, u0, all * sizeof(double));
    for (uint64_t m = 0; m < steps; ++m) {
        step_all(&g, n_species, substeps, p, types + m * g.n, cur, nxt, scratch);
        double *t = cur;
        cur = nxt;
        nxt = t;
    }
    memcpy(out, cur, all * sizeof(double));
    free(cur);
    free(nxt);
    free(scratch);
}

void adj_gradient(int w, int h, int d, uint32_t n_species, const int *substeps,
                  const ghip_field_params *p, const uint8_t *types, uint64_t steps,
                  const double *u0, const double *grad_out, ghip_field_params *grad,
                  double *grad_u0) {
    grid_t g = grid_of(w, h, d);
    const size_t all = g.n * n_species;
    memset(grad, 0, sizeof *grad);
    /* every step's starting state */
    double *starts = malloc((steps + 1) * all * sizeof(double));
    double *scratch = malloc(g.n * sizeof(double));
    memcpy(starts, u0, all * sizeof(double));
    for (uint64_t m = 0; m < steps; ++m)
        step_all(&g, n_species, substeps, p, types + m * g.n, starts + m * all,
                 starts + (m + 1) * all, scratch);
    int max_sub = 0;
    for (uint32_t s = 0; s < n_species; ++s)
        if (substeps[s] > max_sub) max_sub = substeps[s];
    double *states = malloc((size_t)(max_sub + 1) * g.n * sizeof(double));
    double *lam = malloc(all * sizeof(double)), *lam2 = malloc(g.n * sizeof(double));
    memcpy(lam, grad_out, all * sizeof(double));
    for (uint64_t m = steps; m-- > 0;) {
        const uint8_t *ty = types + m * g.n;
        for (uint32_t s = 0; s < n_species; ++s) {
            const int S = substeps[s];
            const double dd = p->diffusion[s] / (double)S, kk = p->decay[s] / (double)S;
            double *L = lam + s * g.n;
            /* the substep states of this step: states[j] before substep j */
            memcpy(states, starts + m * all + s * g.n, g.n * sizeof(double));
            for (int st = 0; st < S; ++st) substep(&g, dd, kk, states + st * g.n, states + (st + 1) * g.n);
            /* exchange */
            const double *pre = states + (size_t)S * g.n;
            for (size_t i = 0; i < g.n; ++i) {
                const uint8_t t = ty[i];
                if (t == 0) continue;
                const double sec = p->secretion[t][s], up = p->uptake[t][s];
                double v = pre[i] + sec;
                const double wv = v - up * v;
                if (wv < 0.0) {
                    L[i] = 0.0;
                    continue;
                }
                grad->secretion[t][s] += L[i] * (1.0 - up);
                grad->uptake[t][s] -= L[i] * v;
                L[i] = L[i] * (1.0 - up);
            }
            /* substeps, last first */
            double gd = 0.0, gk = 0.0;
            for (int st = S; st-- > 0;) {
                const double *v = states + (size_t)st * g.n;
                for (size_t i = 0; i < g.n; ++i) {
                    gd += L[i] * laplacian(&g, v, i);
                    gk -= L[i] * v[i];
                    lam2[i] = L[i] + dd * laplacian(&g, L, i) - kk * L[i];
                }
                memcpy(L, lam2, g.n * sizeof(double));
            }
            grad->diffusion[s] += gd / (double)S;
            grad->decay[s] += gk / (double)S;
        }
    }
    memcpy(grad_u0, lam, all * sizeof(double));
    free(starts);
    free(scratch);
    free(states);
    free(lam);
    free(lam2);
}
