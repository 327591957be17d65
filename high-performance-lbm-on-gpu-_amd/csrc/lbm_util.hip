// Device-copy bandwidth probe: the measured peak the stream-collide kernels are compared with.
//
// Replaces the reference's cp.copyto peak measurement (benchmarks/bench_mlups.py:66-79) by a
// plain copy kernel with 16-byte accesses, run on buffers far larger than the
// 256 MB Infinity Cache so that it measures HBM, not cache.

#include "lbm_common.h"

namespace lbm {

constexpr int COPY_THREADS = 256;
constexpr int COPY_UNROLL = 4;

// Each thread moves COPY_UNROLL 16-byte words, all loads issued before any store;
// consecutive lanes touch consecutive words.
__global__ __launch_bounds__(COPY_THREADS) void k_copy(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n)
{
    const size_t base = (size_t)blockIdx.x * COPY_THREADS * COPY_UNROLL + threadIdx.x;
    uint4 v[COPY_UNROLL];
#pragma unroll
    for (int k = 0; k < COPY_UNROLL; ++k) {
        const size_t i = base + (size_t)k * COPY_THREADS;
        if (i < n) v[k] = src[i];
    }
#pragma unroll
    for (int k = 0; k < COPY_UNROLL; ++k) {
        const size_t i = base + (size_t)k * COPY_THREADS;
        if (i < n) dst[i] = v[k];
    }
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
        const size_t per_block = (size_t)COPY_THREADS * COPY_UNROLL;
        const size_t blocks = (n + per_block - 1) / per_block;
        require(blocks < (1ull << 31), "copy too large");
        k_copy<<<(unsigned)blocks, COPY_THREADS, 0, as_stream(stream)>>>(static_cast<const uint4*>(src), static_cast<uint4*>(dst), n);
        check_launch("copy kernel");
    });
}

}  // extern "C"
