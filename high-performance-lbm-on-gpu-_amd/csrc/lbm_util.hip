// Device-copy bandwidth probe: the measured peak the stream-collide kernels are compared with.
//
// Replaces the reference's cp.copyto peak measurement (benchmarks/bench_mlups.py:66-79) by a
// plain copy kernel with 16-byte accesses, run on buffers far larger than the
// 256 MB Infinity Cache so that it measures HBM, not cache.

#include "lbm_common.h"

namespace lbm {

constexpr int COPY_THREADS = 256;

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// One 16-byte word per lane, nontemporal load and store (the data is not reused):
// the fastest of the copy shapes measured on MI355X (one, four or eight words per
// lane, temporal or nontemporal), so the reported peak is the device's, not the probe's.
__global__ __launch_bounds__(COPY_THREADS) void k_copy(const u32x4* __restrict__ src, u32x4* __restrict__ dst, size_t n)
{
    const size_t i = (size_t)blockIdx.x * COPY_THREADS + threadIdx.x;
    if (i < n) __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
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
        const size_t blocks = (n + COPY_THREADS - 1) / COPY_THREADS;
        require(blocks < (1ull << 31), "copy too large");
        k_copy<<<(unsigned)blocks, COPY_THREADS, 0, as_stream(stream)>>>(static_cast<const u32x4*>(src), static_cast<u32x4*>(dst), n);
        check_launch("copy kernel");
    });
}

}  // extern "C"
