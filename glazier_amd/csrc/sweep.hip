// sweep.hip - the general colour pass (sweep_body.h) as a kernel over one
// lattice, and its launcher.
#include "sweep_body.h"

namespace ghip {

template <bool D3, bool FULL, bool MOORE2D>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(D3 ? 5 : 4))) void sweep_kernel(SweepArgs a) {
    sweep_body<D3, FULL, MOORE2D>(a, blockIdx.x);
}

// Bytes of dynamic LDS a launch with these tile sizes needs.
size_t sweep_lds_bytes(const SweepArgs &a, bool d3) {
    const int TX = 1 << a.tx_log2, TY = 1 << a.ty_log2, TZ = d3 ? 1 << a.tz_log2 : 1;
    const int LX = 2 * TX + 2 * a.halo, LY = 2 * TY + 2 * a.halo,
              LZ = d3 ? 2 * TZ + 2 * a.halo : 1;
    size_t words = (size_t)a.n_types * a.n_types + 4 * (size_t)a.n_types + a.n_thr + 64 +
                   (size_t)LX * LY * LZ;
    if (a.flags & kMotility) words += (size_t)LX * LY * LZ;
    words += 1 + (size_t)TX * TY * TZ;  // the queue of targets whose label would change
    return words * 4;
}

hipError_t launch_sweep(const SweepArgs &a, bool d3, bool full, int blocks, int threads,
                        hipStream_t stream) {
    size_t lds = sweep_lds_bytes(a, d3);
    const bool moore = !d3 && a.n_off == 8;
    if (d3) {
        if (full)
            hipLaunchKernelGGL((sweep_kernel<true, true, false>), dim3(blocks), dim3(threads), lds,
                               stream, a);
        else
            hipLaunchKernelGGL((sweep_kernel<true, false, false>), dim3(blocks), dim3(threads), lds,
                               stream, a);
    } else if (moore) {
        if (full)
            hipLaunchKernelGGL((sweep_kernel<false, true, true>), dim3(blocks), dim3(threads), lds,
                               stream, a);
        else
            hipLaunchKernelGGL((sweep_kernel<false, false, true>), dim3(blocks), dim3(threads), lds,
                               stream, a);
    } else {
        if (full)
            hipLaunchKernelGGL((sweep_kernel<false, true, false>), dim3(blocks), dim3(threads), lds,
                               stream, a);
        else
            hipLaunchKernelGGL((sweep_kernel<false, false, false>), dim3(blocks), dim3(threads), lds,
                               stream, a);
    }
    return hipGetLastError();
}

}  // namespace ghip
