/*
 * glazier_hip.h - C ABI of the MI355X (gfx950) Cellular Potts engine.
 *
 * The engine is a drop-in for the device path of glazier: the checkerboard
 * Metropolis sweep, the chemical-field stencil, the per-cell moments and
 * division/death. Every entry point below replaces one member of
 * `GpuSimulation` in glazier-cuda/src/lib.rs, cited beside it, and keeps its
 * argument meaning and its error behaviour.
 *
 * Conventions:
 *   - plain C: pointers and sizes, no exceptions cross the boundary;
 *   - every call returns a ghip_status; the message of the last failure on
 *     the calling thread is ghip_last_error();
 *   - host buffers stay owned by the caller; the handle owns device memory;
 *   - one handle per host thread at a time; distinct handles are independent.
 *
 * Lattice order is x fastest, then y, then z (glazier-core/src/lattice.rs:122).
 */
#ifndef GLAZIER_HIP_H
#define GLAZIER_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GHIP_MAX_TYPES 16
#define GHIP_MAX_SPECIES 4

/* Status codes. GHIP_ODD_LATTICE carries the reference's "even sides"
 * wording (glazier-cuda/src/lib.rs:733-736). */
typedef enum ghip_status {
    GHIP_OK = 0,
    GHIP_ODD_LATTICE = 1,
    GHIP_DEVICE = 2,      /* a HIP call failed */
    GHIP_UNSUPPORTED = 3, /* a term the engine refuses rather than drops */
    GHIP_BAD_ARG = 4      /* a null pointer or a length that does not match */
} ghip_status;

/* One cell type: a flat mirror of CellType (glazier-core/src/model.rs:9-41).
 * Type 0 is the medium and takes no constraint. */
typedef struct ghip_cell_type {
    double target_volume;
    double lambda_volume;
    double target_surface;
    double lambda_surface;
    double division_volume;
    double death_rate;
    double target_length;
    double lambda_length;
    double max_activity;
    double lambda_activity;
    double external[3];
    double lambda_nematic;
    int32_t connected;
    int32_t reserved;
} ghip_cell_type;

/* One diffusing species (glazier-core/src/field.rs:15-25). `substeps` 0 asks
 * for the stability count ceil(D / (1/(2 dims))) (field.rs:155-160); a
 * positive value states the count outright. */
typedef struct ghip_species {
    double diffusion;
    double decay;
    double initial;
    int32_t substeps;
    int32_t reserved;
} ghip_species;

/* The whole model: a flat mirror of Model (glazier-core/src/model.rs:44-76)
 * with fixed-size tables, so it crosses the ABI as one POD. */
typedef struct ghip_model {
    uint32_t width;
    uint32_t height;
    uint32_t depth;           /* 1 is a plane */
    uint32_t neighbour_order; /* 1, 2 or 3 (model.rs:183-188) */
    uint32_t n_types;         /* including the medium */
    uint32_t n_species;
    double temperature;
    uint64_t seed;
    uint32_t replica;         /* stream id mixed into every draw */
    uint32_t has_nematic;     /* a nematic field and a coupling are present */
    double contact[GHIP_MAX_TYPES * GHIP_MAX_TYPES]; /* row-major, n_types^2 used */
    ghip_cell_type types[GHIP_MAX_TYPES];
    ghip_species species[GHIP_MAX_SPECIES];
    double secretion[GHIP_MAX_TYPES][GHIP_MAX_SPECIES];
    double uptake[GHIP_MAX_TYPES][GHIP_MAX_SPECIES];
    double chemotaxis[GHIP_MAX_TYPES][GHIP_MAX_SPECIES];
} ghip_model;

typedef struct ghip_sim ghip_sim;

/* Replaces GpuSimulation::from_cpu (glazier-cuda/src/lib.rs:814-1049).
 * `labels` has n_sites = width*height*depth entries; `cell_type` has one
 * entry per label including the medium at index 0 (the reference keeps
 * cell_type[0] = 1, glazier-core/src/cpu.rs:203, and the medium then prices
 * contact as that type); `fields` is n_species*n_sites values, species-major,
 * or NULL for every species' initial value; `activity` is n_sites values or
 * NULL for none. Fails with GHIP_ODD_LATTICE on an odd side, with
 * GHIP_UNSUPPORTED on a term outside this engine, and with GHIP_BAD_ARG on a
 * model with has_nematic set (its field comes through ghip_create_nematic). */
ghip_status ghip_create(const ghip_model *model, const uint32_t *labels, size_t n_sites,
                        const uint8_t *cell_type, size_t n_labels, const double *fields,
                        const double *activity, ghip_sim **out);

/* ghip_create for a model with the nematic coupling: from_cpu uploads
 * Model::nematic with the rest (glazier-cuda/src/lib.rs:886-896, 1016-1020).
 * `nematic` holds 2 * n_sites values, (Q_xx, Q_xy) per site in lattice order
 * (glazier-core/src/model.rs:70-75); each must lie in (-128, 128). The
 * engine holds them in units of 2^-24 (nearest, ties away from zero), so each
 * cell's field sum is exact in any order; the term prices -lambda F . a
 * (glazier-core/src/cpu.rs:315-332). Ignored when has_nematic is 0. */
ghip_status ghip_create_nematic(const ghip_model *model, const uint32_t *labels, size_t n_sites,
                                const uint8_t *cell_type, size_t n_labels, const double *fields,
                                const double *activity, const double *nematic, ghip_sim **out);

/* Replaces the Python binding's parse (glazier-py/src/lib.rs:18-150): reads a
 * Blueprint description (JSON text, glazier-core/src/blueprint.rs:195-447)
 * with this engine's reader and writes what it took as a JSON object, the
 * keys of `glazier --check` (glazier/src/main.rs:62-96) plus "depth".
 * *needed receives the size with the terminating zero; `out` may be NULL to
 * ask for it. GHIP_BAD_ARG, with the reader's message, on a bad description. */
ghip_status ghip_blueprint_parse(const char *text, char *out, size_t cap, size_t *needed);

/* Replaces GpuSimulation::set_nematic_field (glazier-cuda/src/lib.rs:1199-1221):
 * `q` holds 2 * n values, n = n_sites. Fails when the model has no nematic
 * coupling. Every cell's field sum is recounted before the next step. */
ghip_status ghip_set_nematic_field(ghip_sim *sim, const double *q, size_t n);

/* Frees the handle and its device memory (the Drop of GpuSimulation). */
ghip_status ghip_destroy(ghip_sim *sim);

/* Replaces GpuSimulation::step (glazier-cuda/src/lib.rs:1055-1144): runs
 * `steps` Monte Carlo steps and returns once they are done. GHIP_DEVICE
 * when the label capacity ran out at the id limit, or when a grid barrier
 * of the tiled resident pass (plain planes above 128^2 sites) gave up
 * waiting; the lattice is then not valid. */
ghip_status ghip_step(ghip_sim *sim, uint64_t steps);

/* Replaces the public field GpuSimulation::mcs (lib.rs:805). */
uint64_t ghip_mcs(const ghip_sim *sim);

/* Labels, medium included, that the per-label read-backs below cover. It
 * grows when cells divide and shrinks when labels are compacted. */
ghip_status ghip_n_labels(const ghip_sim *sim, size_t *out);

/* Replaces GpuSimulation::labels (lib.rs:1227-1229): n must be n_sites. */
ghip_status ghip_labels(const ghip_sim *sim, uint32_t *out, size_t n);

/* Replaces GpuSimulation::volumes (lib.rs:1258-1269): n must be n_labels. */
ghip_status ghip_volumes(const ghip_sim *sim, uint32_t *out, size_t n);

/* Replaces GpuSimulation::surfaces (lib.rs:1245-1255): n must be n_labels. */
ghip_status ghip_surfaces(const ghip_sim *sim, int64_t *out, size_t n);

/* Cell type per label, medium included: n must be n_labels. */
ghip_status ghip_cell_types(const ghip_sim *sim, uint8_t *out, size_t n);

/* Replaces GpuSimulation::field (lib.rs:1235-1242): one species as f64. */
ghip_status ghip_field(const ghip_sim *sim, uint32_t species, double *out, size_t n);

/* The same species as the device holds it, f32, for bit-level comparison. */
ghip_status ghip_field_f32(const ghip_sim *sim, uint32_t species, float *out, size_t n);

/* Replaces the fields the reference uploads once (lib.rs:951-966): sets one
 * species from f64 values, n = n_sites. */
ghip_status ghip_set_field(ghip_sim *sim, uint32_t species, const double *values, size_t n);

/* Persistent-motility memory per site, f32 as held (motility.rs:18-22). */
ghip_status ghip_activity_f32(const ghip_sim *sim, float *out, size_t n);

/* Copy attempts made and accepted since creation (cpu.rs:99-102). */
ghip_status ghip_counters(const ghip_sim *sim, uint64_t *attempted, uint64_t *accepted);

/* Per-label bookkeeping as the engine keeps it, for bit-level comparison:
 * `words` int64 per label in the order volume, surface, then the integer
 * moment sums n-free (sx, sy, sz, sxx, syy, szz, sxy, sxz, syz) about each
 * label's anchor when the moments are kept (length, nematic or division),
 * then with the nematic coupling the field sums (Q_xx, Q_xy) in units of
 * 2^-24. Returns the word count per
 * label in *words_per_label; `out` may be NULL to ask for it. */
ghip_status ghip_cell_state(const ghip_sim *sim, int64_t *out, size_t n,
                            uint32_t *words_per_label);

/* Kernel timing over the stream the engine launches on. Mode 1: every sweep
 * pass, field step and moment rebuild is bracketed by HIP events. Mode 2:
 * one event pair brackets each ghip_step call, and family "step" adds up
 * those spans, its launch count being the colour passes they ran (so on a
 * model whose step is its colour passes alone, milliseconds per launch is
 * the mean pass with the dispatch gaps between passes, and no events
 * between them). 0 turns timing off; other modes are GHIP_BAD_ARG.
 * ghip_kernel_time returns the summed milliseconds and the launch count of
 * one family: "sweep", "fields", "moments", "population" or "step". */
ghip_status ghip_set_timing(ghip_sim *sim, int on);

/* No counterpart in the reference. The polls each grid barrier of the tiled
 * resident pass (plain planes above 128^2 sites) makes before it gives up,
 * sets the error ghip_step reports and lets every workgroup of the launch
 * leave (default 2^22, seconds). 0 gives up at the first barrier, so a test
 * of that failure path runs with every kernel exiting normally. */
ghip_status ghip_set_barrier_patience(ghip_sim *sim, uint32_t polls);
ghip_status ghip_kernel_time(const ghip_sim *sim, const char *family, double *total_ms,
                             uint64_t *launches);

/* ---- Batches. No counterpart in the reference, which steps one lattice per
 * process (glazier/src/main.rs:137-160, examples/bench.rs:40-60); an
 * ensemble there is a loop over processes or over GpuSimulation handles.
 *
 * A batch steps several lattices together: each stage of a Monte Carlo step
 * (colour passes, field launches, per-label work) is one launch over every
 * lattice taking part, and after the first steps the whole step is one
 * replayed hipGraph. The lattices keep their own seeds, replica indices,
 * models and sizes (2D and 3D may be mixed) and compute exactly what
 * ghip_step computes for each alone. The handles stay valid: read them with
 * the functions above between batch steps; ghip_step on one of them is
 * allowed and the batch picks up where it left it. Destroy the batch before
 * its lattices (ghip_destroy refuses a lattice that is in a batch). */
typedef struct ghip_batch ghip_batch;

ghip_status ghip_batch_create(ghip_sim *const *sims, size_t n, ghip_batch **out);
ghip_status ghip_batch_destroy(ghip_batch *batch);

/* Runs `steps` Monte Carlo steps of every lattice in the batch. The first
 * nine steps of a new batch are timed: after a warm-up, each of five tile
 * shapes of the colour passes and of the field launches, then each of three
 * occupancy caps (at most 4, 3 or 2 workgroups of a launch per CU) on the
 * fastest tiles; the fastest tile and cap per kernel variant are kept. */
ghip_status ghip_batch_step(ghip_batch *batch, uint64_t steps);

/* Fixes every group to one tile candidate (0..4) with no occupancy cap and
 * skips the timing steps, or with -1 starts the timing over. */
ghip_status ghip_batch_set_tiles(ghip_batch *batch, int candidate);

/* Fixes every launch to one occupancy candidate (0: no cap; 1..3: at most
 * 4, 3 or 2 workgroups per CU) on the current tiles and skips the timing
 * steps. */
ghip_status ghip_batch_set_occupancy(ghip_batch *batch, int candidate);

/* A JSON description of the groups, their chosen tiles and timings. */
ghip_status ghip_batch_describe(const ghip_batch *batch, char *out, size_t n);

/* ---- Field replay and its adjoint. No counterpart in the reference, which
 * has no derivative of any kind; the forward step replayed here is the
 * engine's field step (Fields::diffuse and Fields::exchange,
 * glazier-core/src/field.rs:99-149; field_diffuse and field_exchange,
 * glazier-cuda/src/lib.rs:314-363).
 *
 * While recording is on, every ghip_step stores each step's cell-type map
 * (one byte per site, 0 for the medium) after its colour passes: the owners
 * the field exchange of that step used. A fit takes such a trajectory and
 * replays the field steps over it, holding the labels fixed, as a function
 * of the field parameters and the initial field; its backward pass gives
 * the gradient of a loss on the final fields with respect to both. */
typedef struct ghip_field_params {
    double diffusion[GHIP_MAX_SPECIES];
    double decay[GHIP_MAX_SPECIES];
    double secretion[GHIP_MAX_TYPES][GHIP_MAX_SPECIES];
    double uptake[GHIP_MAX_TYPES][GHIP_MAX_SPECIES];
} ghip_field_params;

/* Starts recording with room for `max_steps` steps (0 stops and frees it). */
ghip_status ghip_record(ghip_sim *sim, uint64_t max_steps);
ghip_status ghip_recorded_steps(const ghip_sim *sim, uint64_t *out);
/* The type map of recorded step `step` (n = n_sites). */
ghip_status ghip_recorded_types(const ghip_sim *sim, uint64_t step, uint8_t *out, size_t n);

typedef struct ghip_fit ghip_fit;

/* A replay of the recorded steps of `sim` (copied: the lattice may go on or
 * be destroyed), keeping the fields of every `checkpoint_every`-th step in
 * the forward pass and recomputing the rest in the backward pass. */
ghip_status ghip_fit_create(const ghip_sim *sim, uint32_t checkpoint_every, ghip_fit **out);
ghip_status ghip_fit_destroy(ghip_fit *fit);

/* The fit's own parameters as the model states them (a starting point). */
ghip_status ghip_fit_model_params(const ghip_fit *fit, ghip_field_params *out);

/* Forward replay. `u0` and `u_final` are device pointers to n_species *
 * n_sites f32 values, species-major; `stream` is the hipStream_t to order
 * the work on (NULL for the fit's own, synchronised before returning). The
 * f32 arithmetic is the engine's, so the replay of a lattice's own
 * parameters reproduces its fields bit for bit. */
ghip_status ghip_fit_forward(ghip_fit *fit, const ghip_field_params *params, const float *u0,
                             float *u_final, void *stream);

/* Backward pass of the last forward: `grad_final` (device, like u_final) is
 * dL/du_final; writes dL/dparams (accumulated in f64) to `grad` on the host
 * and dL/du0 to the device buffer `grad_u0`. Synchronises `stream`. */
ghip_status ghip_fit_backward(ghip_fit *fit, const float *grad_final, ghip_field_params *grad,
                              float *grad_u0, void *stream);

/* ghip_fit_backward that allocates nothing. No counterpart in the reference:
 * it is for callers that cannot allocate (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc). Same arguments, results
 * and errors as ghip_fit_backward, bit for bit; the kernels' partial sums and
 * their reduction live in the library's static device scratch (4 MiB per
 * GPU, held until the process ends) instead of the fit's buffers, and are
 * cleared by a kernel, not a memset. A fit whose pass needs more than the
 * scratch holds (species x (workgroups + 1) x 34 doubles: a two-species
 * plane of up to about 3.9 million sites, a volume of about 2 million)
 * fails with GHIP_UNSUPPORTED. Every call on
 * one GPU shares that GPU's scratch: calls that overlap on one GPU (other
 * streams or host threads) race and must be serialised by the caller. */
ghip_status ghip_fit_backward_noalloc(ghip_fit *fit, const float *grad_final, ghip_field_params *grad,
                                      float *grad_u0, void *stream);

/* Message of the last failure on this thread; empty when none. */
const char *ghip_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* GLAZIER_HIP_H */
