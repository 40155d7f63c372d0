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
#include <mutex>

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

// ------------------------------------------------- one launch per 2D step
//
// A workgroup of kFitThreads lanes owns a kFitTX x kFitTY tile of one
// species (blockIdx.y) and works in LDS on the tile widened by a halo on
// every side, loaded with periodic wrap (a widened site holds the value of
// the lattice site its coordinates wrap to, so the stencil on the widened
// tile is the periodic stencil). A substep is computed on every widened site
// but the outermost ring; the sites it leaves valid are one ring fewer than
// those it read. The forward step widens by the substep count K, runs the
// substeps and then the exchange on the tile. The adjoint step widens by 2K:
// it recomputes the substep states u_0..u_K from the step start, runs the
// transposed exchange on the widened tile (valid where u_K is, at least K
// in from the edge), then the K transposed substeps inwards, ending on the
// tile. Every site's f32 arithmetic is that of the per-substep kernels above
// in the same order, so the fields and the adjoint fields are bit for bit
// theirs; the parameter sums hold the same terms, added per workgroup and
// step instead of per launch. K is a template argument so the region's
// sizes are constants.

namespace {

constexpr int kFitTX = 32, kFitTY = 16, kFitThreads = 512, kFitWaves = kFitThreads / 64;
constexpr int kFitMaxSub = 5;
#ifndef GHIP_FIT_STEPS
#define GHIP_FIT_STEPS 2
#endif
constexpr int kFitStepsPerLaunch = GHIP_FIT_STEPS;  // forward steps per launch
static_assert(kFitTX * kFitTY == kFitThreads, "one tile site per lane");

__device__ __forceinline__ uint32_t wrap(int v, int n) {
    v %= n;
    return (uint32_t)(v < 0 ? v + n : v);
}

// Lattice column of each region column and lattice row offset of each region
// row, for the region whose corner is lattice (x0, y0).
template <int RX, int RY>
__device__ __forceinline__ void wrap_tables(const FitGrid &g, int x0, int y0, uint32_t *colx,
                                            uint32_t *rowoff) {
    for (int q = threadIdx.x; q < RX + RY; q += kFitThreads) {
        if (q < RX) colx[q] = wrap(x0 + q, g.w);
        else rowoff[q - RX] = wrap(y0 + q - RX, g.h) * (uint32_t)g.w;
    }
}

// One substep (or transposed substep: the operator is symmetric) on every
// region site but the outermost ring. The passes are unrolled so a lane
// issues the reads of all its sites before the arithmetic of the first.
template <int RX, int RY>
__device__ __forceinline__ void substep_region(const float *in, float *out, float d, float kd) {
    constexpr int CW = RX - 2, N = CW * (RY - 2), P = (N + kFitThreads - 1) / kFitThreads;
    float here[P], e[P], w[P], n_[P], s_[P];
    int at[P];
#pragma unroll
    for (int p = 0; p < P; ++p) {
        const int j = (int)threadIdx.x + p * kFitThreads;
        const int jj = j < N ? j : N - 1;
        const int cy = jj / CW, cx = jj - cy * CW;
        const int r = (cy + 1) * RX + cx + 1;
        at[p] = j < N ? r : -1;
        here[p] = in[r];
        e[p] = in[r + 1];
        w[p] = in[r - 1];
        s_[p] = in[r + RX];
        n_[p] = in[r - RX];
    }
#pragma unroll
    for (int p = 0; p < P; ++p) {
        float sum = e[p] + w[p];
        sum = sum + s_[p];
        sum = sum + n_[p];
        if (at[p] >= 0) out[at[p]] = (here[p] + d * (sum - 4.0f * here[p])) - kd * here[p];
    }
}

// The lane's own tile site: its region index at halo H and its lattice index
// (inside false past the lattice's edge).
struct TileSite {
    int cx, cy;
    uint32_t i;
    bool inside;
};

__device__ __forceinline__ TileSite tile_site(const FitGrid &g, int tx, int ty) {
    TileSite t;
    t.cy = (int)threadIdx.x / kFitTX;
    t.cx = (int)threadIdx.x - t.cy * kFitTX;
    const int x = tx * kFitTX + t.cx, y = ty * kFitTY + t.cy;
    t.inside = x < g.w && y < g.h;
    t.i = t.inside ? (uint32_t)y * (uint32_t)g.w + (uint32_t)x : 0;
    return t;
}

}  // namespace

// `steps` (at most S) whole steps from `in`, the owners of the first at
// `types` and of the next ones every n sites after. The state after step q
// goes to out_base + q * out_stride, the last one to out_last. The region is
// widened by K * S, so the tile stays valid through S steps; the exchange
// runs over the whole region, the owners of every step staged in LDS.
template <int K, int S>
__global__ __launch_bounds__(kFitThreads) void fit_steps_fused_kernel(FitGrid g, FitCoef c, const uint8_t *types,
                                                                      int steps, const float *in, float *out_base,
                                                                      size_t out_stride, float *out_last) {
    constexpr int HALO = K * S, RX = kFitTX + 2 * HALO, RY = kFitTY + 2 * HALO, R = RX * RY;
    constexpr int PER = (R + kFitThreads - 1) / kFitThreads;
    __shared__ float buf[2][R];
    __shared__ uint8_t own[S][R];
    __shared__ uint32_t colx[RX], rowoff[RY];
    const int k = blockIdx.y;
    const int tiles_x = (g.w + kFitTX - 1) / kFitTX;
    const int tx = (int)blockIdx.x % tiles_x, ty = (int)blockIdx.x / tiles_x;
    wrap_tables<RX, RY>(g, tx * kFitTX - HALO, ty * kFitTY - HALO, colx, rowoff);
    const TileSite ts = tile_site(g, tx, ty);
    __syncthreads();
    const float *u = in + (size_t)k * g.n;
#pragma unroll
    for (int p = 0; p < PER; ++p) {
        const int j = (int)threadIdx.x + p * kFitThreads;
        if (j < R) {
            const int cy = j / RX, cx = j - cy * RX;
            const uint32_t i = rowoff[cy] + colx[cx];
            buf[0][j] = u[i];
#pragma unroll
            for (int q = 0; q < S; ++q)
                if (q < steps) own[q][j] = types[(size_t)q * g.n + i];
        }
    }
    __syncthreads();
    const float d = c.d[k], kd = c.kd[k];
    const int ks = c.sub[k];
    const int r = (ts.cy + HALO) * RX + ts.cx + HALO;
    int a = 0;
    for (int q = 0; q < steps; ++q) {
        for (int s = 1; s <= ks; ++s) {
            substep_region<RX, RY>(buf[a], buf[a ^ 1], d, kd);
            __syncthreads();
            a ^= 1;
        }
#pragma unroll
        for (int p = 0; p < PER; ++p) {
            const int j = (int)threadIdx.x + p * kFitThreads;
            if (j < R) {
                const uint32_t t = own[q][j];
                if (t != 0) {
                    float w = buf[a][j] + c.sec[t][k];
                    w = w - c.up[t][k] * w;
                    buf[a][j] = w < 0.0f ? 0.0f : w;
                }
            }
        }
        __syncthreads();
        if (ts.inside) {
            float *dst = q + 1 == steps ? out_last : out_base + (size_t)q * out_stride;
            dst[(size_t)k * g.n + ts.i] = buf[a][r];
        }
    }
}

// The adjoint of one whole step: lam_out (adjoint of the step's end) into
// lam_in (adjoint of its start, u_start the start). Block (b, k) adds its
// sums into part[(k * blocks + b) * kPartWords + w] as the per-substep
// kernels lay them out.
template <int K>
__global__ __launch_bounds__(kFitThreads) void fit_step_adjoint_fused_kernel(FitGrid g, FitCoef c,
                                                                             const uint8_t *types,
                                                                             const float *u_start,
                                                                             const float *lam_out, float *lam_in,
                                                                             double *part) {
    constexpr int H = 2 * K, RX = kFitTX + 2 * H, RY = kFitTY + 2 * H, R = RX * RY;
    constexpr int PER = (R + kFitThreads - 1) / kFitThreads;
    __shared__ float U[K + 1][R];  // u_s, s = 0..K
    __shared__ float L[2][R];
    __shared__ uint32_t colx[RX], rowoff[RY];
    __shared__ double red[kFitWaves];
    __shared__ double acc[kFitWaves][kMaxTypes][2];
    const int k = blockIdx.y;
    const int tiles_x = (g.w + kFitTX - 1) / kFitTX;
    const int tx = (int)blockIdx.x % tiles_x, ty = (int)blockIdx.x / tiles_x;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    wrap_tables<RX, RY>(g, tx * kFitTX - H, ty * kFitTY - H, colx, rowoff);
    const TileSite ts = tile_site(g, tx, ty);
    const float *Lk = lam_out + (size_t)k * g.n;
    // every global load of the step up front: the lane's tile site, then the
    // step start, the adjoint and the owners over the region
    const uint32_t tt = ts.inside ? types[ts.i] : 0;
    const float lt = ts.inside ? Lk[ts.i] : 0.0f;
    __syncthreads();
    float lk[PER];
    uint32_t tk[PER];
    const float *u = u_start + (size_t)k * g.n;
#pragma unroll
    for (int p = 0; p < PER; ++p) {
        const int j = (int)threadIdx.x + p * kFitThreads;
        lk[p] = 0.0f;
        tk[p] = 0;
        if (j < R) {
            const int cy = j / RX, cx = j - cy * RX;
            const uint32_t i = rowoff[cy] + colx[cx];
            U[0][j] = u[i];
            lk[p] = Lk[i];
            tk[p] = types[i];
        }
    }
    __syncthreads();
    const int ks = c.sub[k];
    const float d = c.d[k], kd = c.kd[k];
    for (int s = 1; s <= ks; ++s) {
        substep_region<RX, RY>(U[s - 1], U[s], d, kd);
        __syncthreads();
    }
    const float *pre = U[ks];
    // transposed exchange over the region
#pragma unroll
    for (int p = 0; p < PER; ++p) {
        const int j = (int)threadIdx.x + p * kFitThreads;
        if (j < R) {
            float lo_v = lk[p];
            const uint32_t t = tk[p];
            if (t != 0) {
                const float up = c.up[t][k];
                const float v = pre[j] + c.sec[t][k];
                const float w = v - up * v;
                lo_v = w < 0.0f ? 0.0f : lo_v - up * lo_v;
            }
            L[0][j] = lo_v;
        }
    }
    // its parameter sums over the tile, one wave reduction per cell type
    const int r = (ts.cy + H) * RX + ts.cx + H;
    {
        double gs = 0.0, gu = 0.0;
        if (tt != 0) {
            const float up = c.up[tt][k];
            const float v = pre[r] + c.sec[tt][k];
            const float w = v - up * v;
            if (!(w < 0.0f)) {
                gs = (double)lt * (1.0 - (double)up);
                gu = -(double)lt * (double)v;
            }
        }
        for (int q = 1; q < g.n_types; ++q) {
            double a = tt == (uint32_t)q ? gs : 0.0, b = tt == (uint32_t)q ? gu : 0.0;
            for (int off = 32; off > 0; off >>= 1) {
                a += __shfl_xor(a, off, 64);
                b += __shfl_xor(b, off, 64);
            }
            if (lane == 0) {
                acc[wave][q][0] = a;
                acc[wave][q][1] = b;
            }
        }
    }
    __syncthreads();
    // transposed substeps ks..1
    double gd = 0.0, gk = 0.0;
    int a = 0;
    for (int s = ks; s >= 1; --s) {
        if (ts.inside) {
            const float *us = U[s - 1];
            const double usum = (double)us[r + 1] + (double)us[r - 1] + (double)us[r + RX] + (double)us[r - RX];
            const double uh = (double)us[r];
            const double lo_v = (double)L[a][r];
            gd += lo_v * (usum - 4.0 * uh);
            gk += -lo_v * uh;
        }
        substep_region<RX, RY>(L[a], L[a ^ 1], d, kd);
        __syncthreads();
        a ^= 1;
    }
    if (ts.inside) lam_in[(size_t)k * g.n + ts.i] = L[a][r];
    gd = block_sum(gd, red);
    gk = block_sum(gk, red);
    double *slot = part + ((size_t)k * gridDim.x + blockIdx.x) * kPartWords;
    if (threadIdx.x == 0) {
        slot[0] += gd;
        slot[1] += gk;
    }
    const int words = 2 * g.n_types;
    for (int q = threadIdx.x; q < words; q += kFitThreads) {
        if ((q >> 1) == 0) continue;
        double v = 0.0;
        for (int wv = 0; wv < kFitWaves; ++wv) v += acc[wv][q >> 1][q & 1];
        slot[2 + q] += v;
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

uint32_t fit_tiles(const FitGrid &g) {
    return (uint32_t)(((g.w + kFitTX - 1) / kFitTX) * ((g.h + kFitTY - 1) / kFitTY));
}

bool fit_fused_fits(const FitGrid &g, int K) { return !g.d3 && K >= 1 && K <= kFitMaxSub; }

namespace {
template <int K>
void launch_fused(const FitGrid &g, const FitCoef &c, uint32_t species, const uint8_t *types, int steps,
                  const float *in, float *out_base, size_t out_stride, float *out_last, hipStream_t st) {
    hipLaunchKernelGGL((fit_steps_fused_kernel<K, kFitStepsPerLaunch>), dim3(fit_tiles(g), species),
                       dim3(kFitThreads), 0, st, g, c, types, steps, in, out_base, out_stride, out_last);
}
template <int K>
void launch_fused_adjoint(const FitGrid &g, const FitCoef &c, uint32_t species, const uint8_t *types,
                          const float *u_start, const float *lam_out, float *lam_in, double *part,
                          hipStream_t st) {
    hipLaunchKernelGGL(fit_step_adjoint_fused_kernel<K>, dim3(fit_tiles(g), species), dim3(kFitThreads), 0, st, g,
                       c, types, u_start, lam_out, lam_in, part);
}
}  // namespace

hipError_t launch_fit_steps_fused(const FitGrid &g, const FitCoef &c, uint32_t species, int K,
                                  const uint8_t *types, int steps, const float *in, float *out_base,
                                  size_t out_stride, float *out_last, hipStream_t st) {
    if (steps < 1 || steps > kFitStepsPerLaunch) return hipErrorInvalidValue;
    switch (K) {
        case 1: launch_fused<1>(g, c, species, types, steps, in, out_base, out_stride, out_last, st); break;
        case 2: launch_fused<2>(g, c, species, types, steps, in, out_base, out_stride, out_last, st); break;
        case 3: launch_fused<3>(g, c, species, types, steps, in, out_base, out_stride, out_last, st); break;
        case 4: launch_fused<4>(g, c, species, types, steps, in, out_base, out_stride, out_last, st); break;
        case 5: launch_fused<5>(g, c, species, types, steps, in, out_base, out_stride, out_last, st); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

int fit_steps_per_launch() { return kFitStepsPerLaunch; }

hipError_t launch_fit_step_adjoint_fused(const FitGrid &g, const FitCoef &c, uint32_t species, int K,
                                         const uint8_t *types, const float *u_start, const float *lam_out,
                                         float *lam_in, double *part, hipStream_t st) {
    switch (K) {
        case 1: launch_fused_adjoint<1>(g, c, species, types, u_start, lam_out, lam_in, part, st); break;
        case 2: launch_fused_adjoint<2>(g, c, species, types, u_start, lam_out, lam_in, part, st); break;
        case 3: launch_fused_adjoint<3>(g, c, species, types, u_start, lam_out, lam_in, part, st); break;
        case 4: launch_fused_adjoint<4>(g, c, species, types, u_start, lam_out, lam_in, part, st); break;
        case 5: launch_fused_adjoint<5>(g, c, species, types, u_start, lam_out, lam_in, part, st); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_fit_reduce(uint32_t species, uint32_t blocks, const double *part, double *grad,
                             hipStream_t st) {
    hipLaunchKernelGGL(fit_reduce_kernel, dim3(kPartWords, species), dim3(256), 0, st, part, blocks, grad);
    return hipGetLastError();
}

__global__ void fit_zero_kernel(double *p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0.0;
}

hipError_t launch_fit_zero(double *p, size_t n, hipStream_t st) {
    if (n == 0) return hipSuccess;
    size_t blocks = (n + 255) / 256;
    if (blocks > 1024) blocks = 1024;
    hipLaunchKernelGGL(fit_zero_kernel, dim3((unsigned)blocks), dim3(256), 0, st, p, n);
    return hipGetLastError();
}

namespace {

// Scratch of ghip_fit_backward_noalloc (partial sums and their reduction).
// Each GPU holds its own copy until the process ends; host code never names
// it directly, only the per-device address fit_scratch looked up.
static __device__ __align__(256) double scratch[kFitScratchWords];

constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_at[kScratchDevices] = {};

}  // namespace

hipError_t fit_scratch(double **out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kScratchDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> hold(g_scratch_mu);
    if (!g_scratch_at[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_at[dev] = static_cast<double *>(p);
    }
    *out = g_scratch_at[dev];
    return hipSuccess;
}

}  // namespace ghip
