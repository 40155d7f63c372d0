// fields.hip - the field step (field_body.h) as a kernel over one lattice and
// species, the activity decay, and their launchers.
#include "field_body.h"

namespace ghip {

template <bool D3>
__global__ __launch_bounds__(256) void field_kernel(FieldArgs a) {
    field_body<D3>(a, blockIdx.x);
}

// Every species of a plane in one launch, K substeps each on 64 x 32 tiles:
// block b takes tile b % tiles of species b / tiles.
template <int K>
__global__ __launch_bounds__(1024) void field_planes_kernel(FieldPlanes p) {
    constexpr int total = (kFieldIX + 2 * K) * (kFieldIY + 2 * K);
    __shared__ float u[total], v[total];
    const int tiles = p.sp[0].tiles_x * p.sp[0].tiles_y;
    const int s = (int)blockIdx.x / tiles;
    field_body_2d<K>(p.sp[s], (int)blockIdx.x - s * tiles, u, v);
}

// Every site's memory steps down by one, floored at zero (activity_decay,
// glazier-cuda/src/lib.rs:426-432; Activity::decay, motility.rs:39-45).
__global__ void activity_decay_kernel(float *act, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float v = act[i];
    act[i] = v > 0.0f ? v - 1.0f : 0.0f;
}

size_t field_lds_bytes(const FieldArgs &a, bool d3) {
    const int IX = 1 << a.ix_log2, IY = 1 << a.iy_log2, IZ = d3 ? a.iz : 1;
    const int K = a.steps;
    const int LX = IX + 2 * K, LY = IY + 2 * K, LZ = IZ + (d3 ? 2 * K : 0);
    return (size_t)LX * LY * LZ * 2 * sizeof(float);
}

hipError_t launch_field(const FieldArgs &a, bool d3, int blocks, hipStream_t stream) {
    size_t lds = field_lds_bytes(a, d3);
    if (d3)
        hipLaunchKernelGGL(field_kernel<true>, dim3(blocks), dim3(256), lds, stream, a);
    else
        hipLaunchKernelGGL(field_kernel<false>, dim3(blocks), dim3(256), lds, stream, a);
    return hipGetLastError();
}

hipError_t launch_field_planes(const FieldPlanes &p, hipStream_t stream) {
    const FieldArgs &f = p.sp[0];
    if (p.n < 1 || p.n > kMaxSpecies || !field_planes_apply(f.w, f.h, false) || f.ix_log2 != 6 ||
        f.iy_log2 != 5)
        return hipErrorInvalidValue;
    const int blocks = f.tiles_x * f.tiles_y * p.n;
    // 1024 lanes: a tile is one workgroup per CU at these sizes, and its
    // substeps are LDS-latency bound (256 lanes: 15.6 us at 512^2, two
    // species, four substeps; 1024 lanes about 5 us less)
    const int th = 1024;
    switch (f.steps) {
        case 1: hipLaunchKernelGGL(field_planes_kernel<1>, dim3(blocks), dim3(th), 0, stream, p); break;
        case 2: hipLaunchKernelGGL(field_planes_kernel<2>, dim3(blocks), dim3(th), 0, stream, p); break;
        case 3: hipLaunchKernelGGL(field_planes_kernel<3>, dim3(blocks), dim3(th), 0, stream, p); break;
        case 4: hipLaunchKernelGGL(field_planes_kernel<4>, dim3(blocks), dim3(th), 0, stream, p); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_activity_decay(float *act, size_t n, hipStream_t stream) {
    int blocks = (int)((n + 255) / 256);
    hipLaunchKernelGGL(activity_decay_kernel, dim3(blocks), dim3(256), 0, stream, act, n);
    return hipGetLastError();
}

}  // namespace ghip
