// fit.h - device records and launchers of the field replay and its adjoint
// (adjoint.hip), driven by fit_host.cpp.
#pragma once

#include "ghip_device.h"

namespace ghip {

// The lattice a replay runs on.
struct FitGrid {
    int w, h, d;
    uint32_t n;        // sites
    int d3;
    int n_types;
};

// f32 coefficients of one replay, converted from the f64 parameters exactly
// as the engine converts its model: d = (float)(D / substeps), and so on.
struct FitCoef {
    float d[kMaxSpecies], kd[kMaxSpecies];
    float sec[kMaxTypes][kMaxSpecies], up[kMaxTypes][kMaxSpecies];
    int sub[kMaxSpecies];
};

// Per-species gradient words (f64): [0] dL/dd, [1] dL/dk, then per cell type
// t, [2 + 2t] dL/dsec[t] and [3 + 2t] dL/dup[t]. The partials buffer holds
// these words for every (species, workgroup) of a site launch; the reduced
// gradient holds them per species.
constexpr int kPartWords = 2 + 2 * kMaxTypes;

hipError_t launch_record_types(const uint32_t *labels, uint8_t *out, size_t n, hipStream_t s);
hipError_t launch_fit_substep(const FitGrid &g, const FitCoef &c, uint32_t species, const float *in,
                              float *out, int s, hipStream_t st);
hipError_t launch_fit_exchange(const FitGrid &g, const FitCoef &c, uint32_t species, const uint8_t *types,
                               const float *in, float *out, hipStream_t st);
hipError_t launch_fit_substep_adjoint(const FitGrid &g, const FitCoef &c, uint32_t species,
                                      const float *lam_out, const float *u_before, float *lam_in,
                                      double *part, int s, hipStream_t st);
hipError_t launch_fit_exchange_adjoint(const FitGrid &g, const FitCoef &c, uint32_t species,
                                       const uint8_t *types, const float *u_pre, float *lam, double *part,
                                       hipStream_t st);
// One launch per whole step on a 2D lattice (substeps and exchange, or the
// adjoint of both), for K = the largest substep count when fit_fused_fits.
bool fit_fused_fits(const FitGrid &g, int K);
// `steps` (1..fit_steps_per_launch()) forward steps in one launch: the
// owners of step q at types + q * n; the state after step q at out_base +
// q * out_stride, the last at out_last.
int fit_steps_per_launch();
hipError_t launch_fit_steps_fused(const FitGrid &g, const FitCoef &c, uint32_t species, int K,
                                  const uint8_t *types, int steps, const float *in, float *out_base,
                                  size_t out_stride, float *out_last, hipStream_t st);
hipError_t launch_fit_step_adjoint_fused(const FitGrid &g, const FitCoef &c, uint32_t species, int K,
                                         const uint8_t *types, const float *u_start, const float *lam_out,
                                         float *lam_in, double *part, hipStream_t st);
// Workgroups per species of a per-substep launch and of a whole-step launch,
// and the reduction of the partials of `blocks` workgroups per species.
uint32_t fit_blocks(const FitGrid &g);
uint32_t fit_tiles(const FitGrid &g);
hipError_t launch_fit_reduce(uint32_t species, uint32_t blocks, const double *part, double *grad,
                             hipStream_t st);
// Zeroes n doubles (a kernel, so a captured stream holds no memset node).
hipError_t launch_fit_zero(double *p, size_t n, hipStream_t st);

// The library's static scratch for a backward pass that allocates nothing:
// kFitScratchWords doubles (4 MiB) per device, holding the partial sums and
// their reduction. fit_scratch gives the current device's address, looked
// up once per device.
constexpr size_t kFitScratchWords = (size_t)1 << 19;
hipError_t fit_scratch(double **out);

}  // namespace ghip
