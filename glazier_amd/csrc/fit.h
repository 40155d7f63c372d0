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

// Gradient accumulator layout (f64): [2k] dL/dd_k, [2k + 1] dL/dk_k, then
// from kExchangeBase, per (type t, species k), dL/dsec and dL/dup.
constexpr int kExchangeBase = 2 * kMaxSpecies;
constexpr int kGradWords = kExchangeBase + 2 * kMaxTypes * kMaxSpecies;

hipError_t launch_record_types(const uint32_t *labels, uint8_t *out, size_t n, hipStream_t s);
hipError_t launch_fit_substep(const FitGrid &g, const FitCoef &c, uint32_t species, const float *in,
                              float *out, int s, hipStream_t st);
hipError_t launch_fit_exchange(const FitGrid &g, const FitCoef &c, uint32_t species, const uint8_t *types,
                               const float *in, float *out, hipStream_t st);
hipError_t launch_fit_substep_adjoint(const FitGrid &g, const FitCoef &c, uint32_t species,
                                      const float *lam_out, const float *u_before, float *lam_in,
                                      double *gacc, int s, hipStream_t st);
hipError_t launch_fit_exchange_adjoint(const FitGrid &g, const FitCoef &c, uint32_t species,
                                       const uint8_t *types, const float *u_pre, float *lam, double *gacc,
                                       hipStream_t st);

}  // namespace ghip
