/*
 * oracle.h - CPU oracles for the MI355X Cellular Potts engine.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the engine links, loads or calls this
 * code; the tests, the smoke check and the CPU baseline of the benchmark do.
 *
 * Three oracles live here:
 *   serial_*  a restatement of glazier-core's serial f64 engine
 *             (glazier-core/src/cpu.rs, rng.rs, field.rs, moments.rs,
 *             connectivity.rs, motility.rs, lattice.rs): the distributional
 *             reference and the timed CPU baseline;
 *   emu_*     a CPU emulation of the device's checkerboard schedule: the same
 *             colour order, counter hash, per-colour snapshots and f32/int
 *             arithmetic as the HIP kernels, the bit-exact reference;
 *   adj_*     an f64 replay of the field step over a recorded trajectory and
 *             its adjoint, the reference for the device gradient.
 */
#ifndef GLAZIER_ORACLE_H
#define GLAZIER_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#include "../include/glazier_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- rng */
uint64_t oracle_splitmix64_next(uint64_t *state);
void oracle_xoshiro_seed(uint64_t seed, uint64_t s[4]);
uint64_t oracle_xoshiro_next(uint64_t s[4]);
double oracle_xoshiro_f64(uint64_t s[4]);
uint64_t oracle_xoshiro_below(uint64_t s[4], uint64_t n);

/* ---------------------------------------------------------------- serial */
typedef struct serial_sim serial_sim;

serial_sim *serial_create(const ghip_model *m, const uint32_t *labels, const uint8_t *cell_type,
                          size_t n_labels, const double *fields);
void serial_destroy(serial_sim *s);
void serial_step(serial_sim *s, uint64_t steps);
int serial_attempt(serial_sim *s);
double serial_delta_energy(const serial_sim *s, size_t target, uint32_t new_label);
double serial_energy(const serial_sim *s);
double serial_chemotaxis_work(const serial_sim *s, uint32_t new_label, size_t target,
                              size_t source);
double serial_move_work(const serial_sim *s, uint32_t new_label, uint32_t old_label,
                        size_t target, size_t source);
void serial_divide_and_die(serial_sim *s, size_t *divided, size_t *died);
size_t serial_n_labels(const serial_sim *s);
uint64_t serial_mcs(const serial_sim *s);
void serial_counters(const serial_sim *s, uint64_t *attempted, uint64_t *accepted);
void serial_labels(const serial_sim *s, uint32_t *out);
void serial_set_labels(serial_sim *s, const uint32_t *labels);
void serial_volumes(const serial_sim *s, uint32_t *out);
void serial_surfaces(const serial_sim *s, int64_t *out);
void serial_cell_types(const serial_sim *s, uint8_t *out);
void serial_recount_volumes(const serial_sim *s, uint32_t *out);
void serial_recount_surfaces(const serial_sim *s, int64_t *out);
void serial_recount(serial_sim *s);
void serial_decay_activity(serial_sim *s);
void serial_set_nematic(serial_sim *s, const double *q);
void serial_nematic_sums(const serial_sim *s, double *out);
void serial_recounted_nematic(const serial_sim *s, double *out);
void serial_recounted_anisotropy(const serial_sim *s, double *out);
void serial_lengths(const serial_sim *s, double *out);
void serial_recounted_lengths(const serial_sim *s, double *out);
void serial_field(const serial_sim *s, uint32_t species, double *out);
void serial_set_field(serial_sim *s, uint32_t species, const double *values);
void serial_activity(const serial_sim *s, double *out);
void serial_set_activity(serial_sim *s, const double *values);
void serial_fields_step(serial_sim *s);
int serial_locally_connected(const serial_sim *s, size_t target, uint32_t label);
int serial_substeps(uint32_t species_index, const ghip_model *m);
serial_sim *serial_clone(const serial_sim *s);
void serial_apply_copy(serial_sim *s, size_t target, uint32_t new_label);

/* Moments of a list of sites (moments.rs), for the known-answer tests. */
double oracle_moments_length(const double *xyz, size_t n);
void oracle_principal_axis(const double c[9], double axis[3], double *lambda);

/* ---------------------------------------------------------------- emulation */
typedef struct emu_sim emu_sim;

uint32_t oracle_hash32(uint32_t x);
/* The emulation's f32 2^x and log2 (the engine's exp2_f32 / log2_f32). */
float oracle_exp2_f32(float x);
float oracle_log2_f32(float x);
/* One byte per ring mask in [lo, hi): bit 0 union-find, bit 1 the
 * emulation's breadth-first one-piece test (ringcheck.c). */
void oracle_ring_table(int ring_size, uint32_t lo, uint32_t hi, uint8_t *out);
uint32_t oracle_pass_key(uint64_t seed, uint32_t replica, uint64_t step, uint32_t colour);
uint32_t oracle_site_pick(uint32_t key, uint32_t site);
uint32_t oracle_site_draw(uint32_t key, uint32_t site);

emu_sim *emu_create(const ghip_model *m, const uint32_t *labels, const uint8_t *cell_type,
                    size_t n_labels, const double *fields, const double *activity);
void emu_destroy(emu_sim *e);
void emu_step(emu_sim *e, uint64_t steps);
size_t emu_n_labels(const emu_sim *e);
uint64_t emu_mcs(const emu_sim *e);
void emu_counters(const emu_sim *e, uint64_t *attempted, uint64_t *accepted);
void emu_labels(const emu_sim *e, uint32_t *out);
void emu_volumes(const emu_sim *e, uint32_t *out);
void emu_surfaces(const emu_sim *e, int64_t *out);
void emu_cell_types(const emu_sim *e, uint8_t *out);
uint32_t emu_cell_state(const emu_sim *e, int64_t *out);
int32_t emu_nematic_fixed(double q);
void emu_set_nematic(emu_sim *e, const double *q);
void emu_field_f32(const emu_sim *e, uint32_t species, float *out);
void emu_set_field(emu_sim *e, uint32_t species, const double *values);
void emu_activity_f32(const emu_sim *e, float *out);
double emu_delta_energy(const emu_sim *e, size_t target, size_t source);
/* Major axis length 4 sqrt(lambda_max) of n sites from their nine moment
 * sums (sx, sy, sz, sxx, syy, szz, sxy, sxz, syz), as the pricing takes it. */
double emu_length(const int64_t *m, int64_t n, int three_d);
void emu_fields_step(emu_sim *e);
int emu_one_piece(uint32_t present, int ring_size);

/* ---------------------------------------------------------------- adjoint */
/* f64 replay of the field step over a recorded type trajectory (`steps`
 * maps of n_sites bytes) and its adjoint (adjoint.c). Fields are
 * species-major n_species * n_sites values. */
void adj_forward(int w, int h, int d, uint32_t n_species, const int *substeps,
                 const ghip_field_params *p, const uint8_t *types, uint64_t steps,
                 const double *u0, double *out);
void adj_gradient(int w, int h, int d, uint32_t n_species, const int *substeps,
                  const ghip_field_params *p, const uint8_t *types, uint64_t steps,
                  const double *u0, const double *grad_out, ghip_field_params *grad,
                  double *grad_u0);

#ifdef __cplusplus
}
#endif

#endif
