// adjoint.hip - the field step replayed over a recorded owner trajectory,
// and its adjoint, on gfx950.
//
// Forward: one launch per substep over every species (blockIdx.y), the same
// f32 operations in the same order as the engine's field step (field_body.h):
//
//     sum = u[x+1] + u[x-1];  sum = sum + u[y+1];  sum = sum + u[y-1]  (+ z)
//     u'  = (u + d * (sum - faces * u)) - k * u
//
// then the exchange on owned sites: w = u + s; w = w - up * w; u = max(w, 0).
//
// Backward (the transpose; the periodic Laplacian is symmetric):
//
//     lambda = (lambda' + d * (sum' - faces * lambda')) - k * lambda'
//     dL/dd += lambda' * (sum(u) - faces * u),   dL/dk -= lambda' * u
//
// and for the exchange, with pass = !(w < 0),
//
//     lambda = pass ? lambda'' - up * lambda'' : 0
//     dL/ds[t] += pass * lambda'' * (1 - up),   dL/dup[t] -= pass * lambda'' * (u + s)
//
// Adjoint fields are f32 like the fields; every parameter derivative is
// formed and summed in f64, per workgroup first and then with one f64 atomic
// per workgroup and parameter.
#include "fit.h"

namespace ghip {

namespace {

struct Nbr {
    uint32_t xp, xm, yp, ym, zp, zm;
};

__device__ __forceinline__ Nbr neighbours(const FitGrid &g, uint32_t i) {
    const uint32_t plane = (uint32_t)g.w * (uint32_t)g.h;
    const uint32_t z = i / plane, r = i - z * plane;
    const uint32_t y = r / (uint32_t)g.w, x = r - y * (uint32_t)g.w;
    const uint32_t row = i - x, slab = i - r;
    Nbr n;
    n.xp = row + (x + 1 == (uint32_t)g.w ? 0 : x + 1);
    n.xm = row + (x == 0 ? (uint32_t)g.w - 1 : x - 1);
    n.yp = slab + (y + 1 == (uint32_t)g.h ? 0 : y + 1) * (uint32_t)g.w + x;
    n.ym = slab + (y == 0 ? (uint32_t)g.h - 1 : y - 1) * (uint32_t)g.w + x;
    n.zp = ((z + 1 == (uint32_t)g.d ? 0 : z + 1) * plane) + r;
    n.zm = ((z == 0 ? (uint32_t)g.d - 1 : z - 1) * plane) + r;
    return n;
}

template <bool D3>
__device__ __forceinline__ float nsum(const float *u, const Nbr &n) {
    float sum = u[n.xp] + u[n.xm];
    sum = sum + u[n.yp];
    sum = sum + u[n.ym];
    if (D3) {
        sum = sum + u[n.zp];
        sum = sum + u[n.zm];
    }
    return sum;
}

// Sum of a per-lane f64 over the workgroup; the total is valid in thread 0.
__device__ __forceinline__ double block_sum(double v, double *red) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int q = 0; q < (int)(blockDim.x >> 6); ++q) t += red[q];
    __syncthreads();
    return t;
}

}  // namespace

// Owner types of one step: the type bits of each site's label word.
__global__ void record_types_kernel(const uint32_t *labels, uint8_t *out, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (uint8_t)label_type(labels[i]);
}

template <bool D3>
__global__ __launch_bounds__(256) void fit_substep_kernel(FitGrid g, FitCoef c, const float *in,
                                                          float *out, int s) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (i >= g.n) return;
    const float *u = in + (size_t)k * g.n;
    const float here = u[i];
    if (s > c.sub[k]) {
        out[(size_t)k * g.n + i] = here;
        return;
    }
    const Nbr n = neighbours(g, i);
    const float sum = nsum<D3>(u, n);
    const float faces = D3 ? 6.0f : 4.0f;
    out[(size_t)k * g.n + i] = (here + c.d[k] * (sum - faces * here)) - c.kd[k] * here;
}

__global__ __launch_bounds__(256) void fit_exchange_kernel(FitGrid g, FitCoef c, const uint8_t *types,
                                                           const float *in, float *out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (i >= g.n) return;
    float val = in[(size_t)k * g.n + i];
    const uint32_t t = types[i];
    if (t != 0) {
        float w = val + c.sec[t][k];
        w = w - c.up[t][k] * w;
        val = w < 0.0f ? 0.0f : w;
    }
    out[(size_t)k * g.n + i] = val;
}

// Transposed substep s of every species; gacc[2k] += dL/dd_k, gacc[2k+1] += dL/dk_k.
template <bool D3>
__global__ __launch_bounds__(256) void fit_substep_adjoint_kernel(FitGrid g, FitCoef c,
                                                                  const float *lam_out,
                                                                  const float *u_before, float *lam_in,
                                                                  double *gacc, int s) {
    __shared__ double red[4];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (s > c.sub[k]) {
        if (i < g.n) lam_in[(size_t)k * g.n + i] = lam_out[(size_t)k * g.n + i];
        return;
    }
    double gd = 0.0, gk = 0.0;
    if (i < g.n) {
        const float *L = lam_out + (size_t)k * g.n;
        const float *u = u_before + (size_t)k * g.n;
        const Nbr n = neighbours(g, i);
        const float faces = D3 ? 6.0f : 4.0f;
        const float lo = L[i];
        const float lsum = nsum<D3>(L, n);
        lam_in[(size_t)k * g.n + i] = (lo + c.d[k] * (lsum - faces * lo)) - c.kd[k] * lo;
        double usum = (double)u[n.xp] + (double)u[n.xm] + (double)u[n.yp] + (double)u[n.ym];
        if (D3) usum += (double)u[n.zp] + (double)u[n.zm];
        const double uh = (double)u[i];
        gd = (double)lo * (usum - (double)faces * uh);
        gk = -(double)lo * uh;
    }
    gd = block_sum(gd, red);
    gk = block_sum(gk, red);
    if (threadIdx.x == 0) {
        atomicAdd(gacc + 2 * k, gd);
        atomicAdd(gacc + 2 * k + 1, gk);
    }
}

// Transposed exchange, in place on lam; per-type sums go to
// gacc[kExchangeBase + 2 * (t * kMaxSpecies + k) + {0: secretion, 1: uptake}].
__global__ __launch_bounds__(256) void fit_exchange_adjoint_kernel(FitGrid g, FitCoef c, const uint8_t *types,
                                                                   const float *u_pre, float *lam,
                                                                   double *gacc) {
    __shared__ double acc[kMaxTypes][2];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    for (int q = threadIdx.x; q < 2 * kMaxTypes; q += blockDim.x) acc[q >> 1][q & 1] = 0.0;
    __syncthreads();
    if (i < g.n) {
        const uint32_t t = types[i];
        if (t != 0) {
            const size_t at = (size_t)k * g.n + i;
            const float up = c.up[t][k];
            const float v = u_pre[at] + c.sec[t][k];
            const float w = v - up * v;
            const float lo = lam[at];
            if (w < 0.0f) {
                lam[at] = 0.0f;
            } else {
                atomicAdd(&acc[t][0], (double)lo * (1.0 - (double)up));
                atomicAdd(&acc[t][1], -(double)lo * (double)v);
                lam[at] = lo - up * lo;
            }
        }
    }
    __syncthreads();
    for (int q = threadIdx.x; q < 2 * g.n_types; q += blockDim.x) {
        const double val = acc[q >> 1][q & 1];
        if (val != 0.0) atomicAdd(gacc + kExchangeBase + 2 * ((q >> 1) * kMaxSpecies + k) + (q & 1), val);
    }
}

// ---------------------------------------------------------------- launchers

namespace {
inline dim3 site_grid(const FitGrid &g, uint32_t species) { return dim3((g.n + 255) / 256, species); }
}  // namespace

hipError_t launch_record_types(const uint32_t *labels, uint8_t *out, size_t n, hipStream_t s) {
    hipLaunchKernelGGL(record_types_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, labels, out, n);
    return hipGetLastError();
}

hipError_t launch_fit_substep(const FitGrid &g, const FitCoef &c, uint32_t species, const float *in,
                              float *out, int s, hipStream_t st) {
    if (g.d3)
        hipLaunchKernelGGL(fit_substep_kernel<true>, site_grid(g, species), dim3(256), 0, st, g, c, in, out, s);
    else
        hipLaunchKernelGGL(fit_substep_kernel<false>, site_grid(g, species), dim3(256), 0, st, g, c, in, out, s);
    return hipGetLastError();
}

hipError_t launch_fit_exchange(const FitGrid &g, const FitCoef &c, uint32_t species, const uint8_t *types,
                               const float *in, float *out, hipStream_t st) {
    hipLaunchKernelGGL(fit_exchange_kernel, site_grid(g, species), dim3(256), 0, st, g, c, types, in, out);
    return hipGetLastError();
}

hipError_t launch_fit_substep_adjoint(const FitGrid &g, const FitCoef &c, uint32_t species,
                                      const float *lam_out, const float *u_before, float *lam_in,
                                      double *gacc, int s, hipStream_t st) {
    if (g.d3)
        hipLaunchKernelGGL(fit_substep_adjoint_kernel<true>, site_grid(g, species), dim3(256), 0, st, g, c,
                           lam_out, u_before, lam_in, gacc, s);
    else
        hipLaunchKernelGGL(fit_substep_adjoint_kernel<false>, site_grid(g, species), dim3(256), 0, st, g, c,
                           lam_out, u_before, lam_in, gacc, s);
    return hipGetLastError();
}

hipError_t launch_fit_exchange_adjoint(const FitGrid &g, const FitCoef &c, uint32_t species,
                                       const uint8_t *types, const float *u_pre, float *lam, double *gacc,
                                       hipStream_t st) {
    hipLaunchKernelGGL(fit_exchange_adjoint_kernel, site_grid(g, species), dim3(256), 0, st, g, c, types, u_pre,
                       lam, gacc);
    return hipGetLastError();
}

}  // namespace ghip
