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
// formed and summed in f64. Each workgroup adds its sum into its own slot of
// a partials buffer (a launch's grid is the same on every step, so slot b
// collects block b's share of every step, in step order, with no atomics and
// no contention); one reduction over the slots ends the backward pass. The
// gradient is therefore the same on every run.
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

// Transposed substep s of every species. Block (b, k) adds dL/dd_k and
// dL/dk_k into part[(k * blocks + b) * kPartWords + {0, 1}].
template <bool D3>
__global__ __launch_bounds__(256) void fit_substep_adjoint_kernel(FitGrid g, FitCoef c,
                                                                  const float *lam_out,
                                                                  const float *u_before, float *lam_in,
                                                                  double *part, int s) {
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
        double *slot = part + ((size_t)k * gridDim.x + blockIdx.x) * kPartWords;
        slot[0] += gd;
        slot[1] += gk;
    }
}

// Transposed exchange, in place on lam. Block (b, k) adds its per-type sums
// into part[(k * blocks + b) * kPartWords + 2 + 2 * t + {0: secretion, 1: uptake}].
__global__ __launch_bounds__(256) void fit_exchange_adjoint_kernel(FitGrid g, FitCoef c, const uint8_t *types,
                                                                   const float *u_pre, float *lam,
                                                                   double *part) {
    __shared__ double acc[4][kMaxTypes][2];  // per wave
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    uint32_t t = 0;
    double gs = 0.0, gu = 0.0;
    if (i < g.n) {
        t = types[i];
        if (t != 0) {
            const size_t at = (size_t)k * g.n + i;
            const float up = c.up[t][k];
            const float v = u_pre[at] + c.sec[t][k];
            const float w = v - up * v;
            const float lo = lam[at];
            if (w < 0.0f) {
                lam[at] = 0.0f;
            } else {
                gs = (double)lo * (1.0 - (double)up);
                gu = -(double)lo * (double)v;
                lam[at] = lo - up * lo;
            }
        }
    }
    // one wave reduction per cell type (the medium exchanges nothing)
    for (int q = 1; q < g.n_types; ++q) {
        double a = t == (uint32_t)q ? gs : 0.0, b = t == (uint32_t)q ? gu : 0.0;
        for (int off = 32; off > 0; off >>= 1) {
            a += __shfl_xor(a, off, 64);
            b += __shfl_xor(b, off, 64);
        }
        if (lane == 0) {
            acc[wave][q][0] = a;
            acc[wave][q][1] = b;
        }
    }
    __syncthreads();
    double *slot = part + ((size_t)k * gridDim.x + blockIdx.x) * kPartWords + 2;
    const int words = 2 * g.n_types, waves = (int)(blockDim.x >> 6);
    for (int q = threadIdx.x; q < words; q += blockDim.x) {
        if ((q >> 1) == 0) continue;
        double v = 0.0;
        for (int wv = 0; wv < waves; ++wv) v += acc[wv][q >> 1][q & 1];
        slot[q] += v;
    }
}

// Sum of the partials of every block: grad[k * kPartWords + w] for species k
// and word w, one workgroup per (word, species), in block order.
__global__ __launch_bounds__(256) void fit_reduce_kernel(const double *part, uint32_t blocks, double *grad) {
    __shared__ double red[4];
    const int w = blockIdx.x, k = blockIdx.y;
    double v = 0.0;
    for (uint32_t b = threadIdx.x; b < blocks; b += blockDim.x)
        v += part[((size_t)k * blocks + b) * kPartWords + w];
    v = block_sum(v, red);
    if (threadIdx.x == 0) grad[k * kPartWords + w] = v;
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
                                      double *part, int s, hipStream_t st) {
    if (g.d3)
        hipLaunchKernelGGL(fit_substep_adjoint_kernel<true>, site_grid(g, species), dim3(256), 0, st, g, c,
                           lam_out, u_before, lam_in, part, s);
    else
        hipLaunchKernelGGL(fit_substep_adjoint_kernel<false>, site_grid(g, species), dim3(256), 0, st, g, c,
                           lam_out, u_before, lam_in, part, s);
    return hipGetLastError();
}

hipError_t launch_fit_exchange_adjoint(const FitGrid &g, const FitCoef &c, uint32_t species,
                                       const uint8_t *types, const float *u_pre, float *lam, double *part,
                                       hipStream_t st) {
    hipLaunchKernelGGL(fit_exchange_adjoint_kernel, site_grid(g, species), dim3(256), 0, st, g, c, types, u_pre,
                       lam, part);
    return hipGetLastError();
}

uint32_t fit_blocks(const FitGrid &g) { return (g.n + 255) / 256; }

hipError_t launch_fit_reduce(const FitGrid &g, uint32_t species, const double *part, double *grad,
                             hipStream_t st) {
    hipLaunchKernelGGL(fit_reduce_kernel, dim3(kPartWords, species), dim3(256), 0, st, part, fit_blocks(g), grad);
    return hipGetLastError();
}

}  // namespace ghip
