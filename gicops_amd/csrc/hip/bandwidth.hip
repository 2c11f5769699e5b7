// Measured read bandwidth of the device (gx_measure_read_bandwidth): a
// sequential read of a 4 GiB buffer, or reads of random segments of S
// consecutive 128-B lines of it (the access shape of the search's ion-id
// stream: slices of a few lines at random places). A second denominator for
// the search kernel's roofline beside the published HBM peak.
#include "capi_internal.h"

namespace gx {
namespace {

__device__ __forceinline__ uint32_t mix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
}

__global__ void __launch_bounds__(256) seq_read_kernel(const uint4 *__restrict__ a, uint64_t n16, uint32_t *out) {
    uint32_t acc = 0;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n16; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint4 v = a[i];
        acc += v.x + v.y + v.z + v.w;
    }
    if (acc == 0x9e3779b9u) out[0] = acc;  // keeps the loads
}

// each wave reads `segs` segments of S lines at random line indices in
// [0, nlines - S], 8 segments in flight; lane l reads dwords l, l+64, ...
template <int S>
__global__ void __launch_bounds__(256) seg_read_kernel(const uint32_t *__restrict__ a, uint32_t nlines, uint32_t segs,
                                                       uint32_t *out) {
    const uint32_t lane = threadIdx.x & 63, wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    uint32_t acc = 0;
    constexpr int U = 8;
    for (uint32_t k = 0; k < segs; k += U) {
        uint32_t v[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const uint64_t line = mix32(wave * 0x9E3779B9u + (k + (uint32_t)u) * 0x85ebca6bu) % (nlines - (uint32_t)S + 1u);
            uint32_t x = 0;
#pragma unroll
            for (int q = 0; q < (S * 32 + 63) / 64; q++) {
                const uint32_t d = lane + 64u * (uint32_t)q;
                if (d < (uint32_t)S * 32u) x += a[line * 32u + d];
            }
            v[u] = x;
        }
#pragma unroll
        for (int u = 0; u < U; u++) acc += v[u];
    }
    if (acc == 0x9e3779b9u) out[0] = acc;
}

}  // namespace
}  // namespace gx

using namespace gx;

extern "C" int gx_measure_read_bandwidth(gx_ctx *c, int segment_lines, double *gbps) {
    if (!c || !gbps) return GX_ERR_INVLD_PTR;
    *gbps = 0.0;
    if (segment_lines != 0 && segment_lines != 1 && segment_lines != 2 && segment_lines != 4 && segment_lines != 8)
        return param_error(GX_ERR_INVLD_PARAM, "segment_lines (0, 1, 2, 4 or 8)");
    std::lock_guard<std::mutex> lk(c->mu);
    if (hipSetDevice(c->device) != hipSuccess) return GX_ERR_DEVICE;
    const uint64_t bytes = 4ull << 30;
    const uint32_t nlines = (uint32_t)(bytes / 128);
    uint32_t *a = nullptr, *out = nullptr;
    if (dalloc(&a, bytes / 4) || dalloc(&out, 16)) {
        if (a) (void)hipFree(a);
        return GX_ERR_BAD_MEM_ALLOC;
    }
    hipStream_t st = c->stream;
    const uint32_t waves = (uint32_t)std::max(1, c->n_cu) * 32u;  // 32 waves per CU
    const uint32_t segs = 2048;
    float ms = 0.0f;
    int rc = GX_SUCCESS;
    if (hipMemsetAsync(a, 1, bytes, st) != hipSuccess) rc = GX_ERR_DEVICE;
    for (int rep = 0; rep < 2 && rc == GX_SUCCESS; rep++) {  // the first run warms up
        if (hipEventRecord(c->ev0, st) != hipSuccess) rc = GX_ERR_DEVICE;
        switch (segment_lines) {
            case 0:
                hipLaunchKernelGGL(seq_read_kernel, dim3(waves / 4), dim3(256), 0, st, (const uint4 *)a, bytes / 16, out);
                break;
            case 1:
                hipLaunchKernelGGL(seg_read_kernel<1>, dim3(waves / 4), dim3(256), 0, st, a, nlines, segs, out);
                break;
            case 2:
                hipLaunchKernelGGL(seg_read_kernel<2>, dim3(waves / 4), dim3(256), 0, st, a, nlines, segs, out);
                break;
            case 4:
                hipLaunchKernelGGL(seg_read_kernel<4>, dim3(waves / 4), dim3(256), 0, st, a, nlines, segs, out);
                break;
            default:
                hipLaunchKernelGGL(seg_read_kernel<8>, dim3(waves / 4), dim3(256), 0, st, a, nlines, segs, out);
                break;
        }
        if (hipGetLastError() != hipSuccess || hipEventRecord(c->ev1, st) != hipSuccess ||
            hipEventSynchronize(c->ev1) != hipSuccess || hipEventElapsedTime(&ms, c->ev0, c->ev1) != hipSuccess)
            rc = GX_ERR_DEVICE;
    }
    (void)hipFree(a);
    (void)hipFree(out);
    if (rc == GX_SUCCESS && ms > 0.0f) {
        const double b = segment_lines == 0 ? (double)bytes : (double)waves * segs * segment_lines * 128.0;
        *gbps = b / (ms * 1e6);
    }
    return rc;
}
