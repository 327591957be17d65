// Device-copy bandwidth probe: the measured peak the stream-collide kernels are compared with.
//
// Replaces the reference's cp.copyto peak measurement (benchmarks/bench_mlups.py:66-79) by a
// plain grid-stride copy kernel with 16-byte accesses, run on buffers far larger than the
// 256 MB Infinity Cache so that it measures HBM, not cache.

#include "lbm_common.h"

namespace lbm {

__global__ __launch_bounds__(256) void k_copy(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

}  // namespace lbm

using namespace lbm;

extern "C" {

LBM_API int lbm_device_copy(void* dst, const void* src, int64_t bytes, void* stream)
{
    return guarded("lbm_device_copy", [&] {
        require(bytes >= 0 && bytes % 16 == 0, "copy size must be a multiple of 16 bytes");
        require((reinterpret_cast<uintptr_t>(dst) | reinterpret_cast<uintptr_t>(src)) % 16 == 0,
                "copy buffers must be 16-byte aligned");
        const size_t n = (size_t)bytes / 16;
        if (n == 0) return;
        const size_t blocks = std::min<size_t>((n + 255) / 256, 256 * 64);
        k_copy<<<(unsigned)blocks, 256, 0, as_stream(stream)>>>(static_cast<const uint4*>(src), static_cast<uint4*>(dst), n);
        check_launch("copy kernel");
    });
}

}  // extern "C"
