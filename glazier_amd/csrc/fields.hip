// fields.hip - the field step (field_body.h) as a kernel over one lattice and
// species, the activity decay, and their launchers.
#include "field_body.h"

namespace ghip {

template <bool D3>
__global__ __launch_bounds__(256) void field_kernel(FieldArgs a) {
    field_body<D3>(a, blockIdx.x);
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

hipError_t launch_activity_decay(float *act, size_t n, hipStream_t stream) {
    int blocks = (int)((n + 255) / 256);
    hipLaunchKernelGGL(activity_decay_kernel, dim3(blocks), dim3(256), 0, stream, act, n);
    return hipGetLastError();
}

}  // namespace ghip
