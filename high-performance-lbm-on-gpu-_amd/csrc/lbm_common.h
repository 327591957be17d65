// Shared host-side plumbing of the C ABI: error reporting, type dispatch, grid description.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

#include "lbm_lattice.h"
#include "lbm_physics.h"
#include "../../include/lbm_hip.h"

namespace lbm {

// Last error message of the calling host thread (lbm_last_error()).
void set_error(const std::string& message);

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void require(bool ok, const std::string& message, int code = LBM_ERR_INVALID)
{
    if (!ok) throw Error(code, message);
}

inline void hip_check(hipError_t e, const char* what)
{
    if (e != hipSuccess) throw Error(LBM_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

// Run an entry point body, turning exceptions into codes + lbm_last_error().
template <class F>
int guarded(const char* name, F&& body)
{
    try {
        body();
        return LBM_OK;
    } catch (const Error& e) {
        set_error(std::string(name) + ": " + e.what());
        return e.code;
    } catch (const std::exception& e) {
        set_error(std::string(name) + ": " + e.what());
        return LBM_ERR_INTERNAL;
    }
}

inline void check_launch(const char* what)
{
    hip_check(hipGetLastError(), what);
}

// Call f(L{}, T{}) for the lattice and precision codes of the C ABI.
template <class F>
void dispatch(int lattice, int dtype, F&& f)
{
    require(lattice == LBM_D2Q9 || lattice == LBM_D3Q19, "unknown lattice code");
    require(dtype == LBM_F32 || dtype == LBM_F64, "unknown dtype code");
    if (lattice == LBM_D2Q9) {
        if (dtype == LBM_F32) f(D2Q9{}, float{});
        else f(D2Q9{}, double{});
    } else {
        if (dtype == LBM_F32) f(D3Q19{}, float{});
        else f(D3Q19{}, double{});
    }
}

template <class F>
void dispatch_real(int dtype, F&& f)
{
    require(dtype == LBM_F32 || dtype == LBM_F64, "unknown dtype code");
    if (dtype == LBM_F32) f(float{});
    else f(double{});
}

// Grid of nx * ny * nz nodes, x fastest; n = nx ny nz < 2^31.
struct Grid {
    int nx, ny, nz;
    uint32_t n;
    uint32_t ns;  // elements from one population plane to the next: n, or more in padded pools
};

inline Grid make_grid(int lattice, int nx, int ny, int nz)
{
    require(nx > 0 && ny > 0 && nz > 0, "grid sizes must be positive");
    require(lattice == LBM_D3Q19 || nz == 1, "a D2Q9 grid has nz = 1");
    const uint64_t n = (uint64_t)nx * ny * nz;
    require(n < (1ull << 31), "grids are limited to 2^31 nodes");
    return Grid{nx, ny, nz, (uint32_t)n, (uint32_t)n};
}

// Lanes per workgroup the stream-collide kernels are compiled for: 1024, except float64
// D3Q19, whose 19 double populations (plus the bounce-back reloads next to walls) need
// more than the 128 VGPRs per lane that a 1024-lane workgroup leaves.
constexpr int step_max_threads(int d, size_t real) { return (d == 3 && real == 8) ? 512 : 1024; }

inline int step_max_threads_of(int lattice, int dtype)
{
    return step_max_threads(lattice == LBM_D3Q19 ? 3 : 2, dtype == LBM_F64 ? 8 : 4);
}

inline dim3 block_of(int bx, int by, int max_threads = 1024)
{
    require(bx > 0 && by > 0 && bx * by <= max_threads,
            max_threads == 1024 ? "block must be positive with at most 1024 threads"
                                : "block must be positive with at most 512 threads for float64 D3Q19");
    return dim3(bx, by, 1);
}

inline dim3 grid_of(const Grid& g, dim3 block)
{
    return dim3((g.nx + block.x - 1) / block.x, (g.ny + block.y - 1) / block.y, g.nz);
}

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// A device attribute of the current device, queried once per device.
inline size_t device_attribute(hipDeviceAttribute_t attribute, size_t (&cache)[64])
{
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    if (dev >= 0 && dev < 64 && cache[dev]) return cache[dev];
    int value = 0;
    hip_check(hipDeviceGetAttribute(&value, attribute, dev), "device attribute query");
    if (dev >= 0 && dev < 64) cache[dev] = (size_t)value;
    return (size_t)value;
}

// LDS bytes one workgroup may allocate on the current device (160 KiB on gfx950).
inline size_t lds_limit()
{
    static size_t cache[64];
    return device_attribute(hipDeviceAttributeMaxSharedMemoryPerBlock, cache);
}

// LDS bytes of one compute unit on the current device (160 KiB on gfx950).
inline size_t lds_per_cu()
{
    static size_t cache[64];
    return device_attribute(hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, cache);
}

// Dynamic LDS a launch requests so that at most `groups_per_cu` of its workgroups are
// resident on one compute unit (0: no cap), never less than the `needed` bytes it uses.
// Capping residency this way is how the streaming kernels pick their occupancy.
inline size_t occupancy_lds(int groups_per_cu, size_t needed)
{
    if (groups_per_cu <= 0) return needed;
    const size_t cap = (lds_per_cu() / (size_t)groups_per_cu) & ~(size_t)255;
    return std::max(needed, std::min(cap, lds_limit()));
}

// Allow a kernel more than the default 64 KiB of dynamic LDS before launching it with `bytes`.
template <class K>
inline void allow_lds(K* kernel, size_t bytes)
{
    if (bytes > 64 * 1024)
        hip_check(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes),
                  "hipFuncSetAttribute(LDS)");
}

}  // namespace lbm
