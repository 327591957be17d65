// The momentum-exchange force reduction shared by the sparse kernels (lbm_rules.hip) and
// the pull step that carries the force of the state it reads (lbm_stream.hip).
//
// Force = sum_r weight[r] * value[r] over the affine rows of a probe (d components, value
// rounded to the population precision, then widened). The rows are split into `groups`
// groups of FORCE_THREADS lanes: lane t of group b takes rows b * FORCE_THREADS + t +
// k * groups * FORCE_THREADS in increasing k, then a fixed tree (each 64-lane wavefront's
// shuffle tree, then the same tree over the wavefronts' sums). One group writes the force,
// several write partial sums that a one-group pass adds in group order. The group count
// depends on the row count only, so a given probe always sums in the same order, whichever
// kernel hosts the groups.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "lbm_affine.h"

namespace lbm {

constexpr int FORCE_THREADS = 1024;
constexpr int64_t FORCE_ONE_GROUP_ROWS = 2048;
constexpr int FORCE_MAX_GROUPS = 1024;
constexpr int FORCE_WAVES = FORCE_THREADS / 64;

// Groups of the reduction: one up to FORCE_ONE_GROUP_ROWS rows (a probe of up to 1024 rows
// costs one row per lane), else one per FORCE_THREADS rows up to FORCE_MAX_GROUPS.
inline int force_groups(int64_t n)
{
    if (n <= FORCE_ONE_GROUP_ROWS) return 1;
    return (int)std::min<int64_t>(FORCE_MAX_GROUPS, (n + FORCE_THREADS - 1) / FORCE_THREADS);
}

// A probe's affine rows on the device (term-major, row stride n) with weight[n][d].
template <class T>
struct ForceRows {
    const int64_t* slot;
    const T* coeff;
    const T* constant;
    const double* weight;
    int64_t n;
    int m;
    int d;
};

// Sum over the 64 lanes of a wavefront by a fixed shuffle tree (lane 0 holds the result).
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int offset = 32; offset > 0; offset >>= 1) v += __shfl_down(v, offset);
    return v;
}

// The accumulation of lane `lane` of group `grp` (of `groups`).
template <class T>
__device__ __forceinline__ void force_lane(const T* __restrict__ F, const ForceRows<T>& fr, int groups, int grp,
                                           int lane, double (&acc)[3])
{
    const int64_t stride = (int64_t)groups * FORCE_THREADS;
    for (int64_t r = (int64_t)grp * FORCE_THREADS + lane; r < fr.n; r += stride) {
        const double v = (double)affine_row(F, fr.slot, fr.coeff, fr.constant, fr.n, fr.m, r);
        for (int a = 0; a < fr.d; ++a) acc[a] += fr.weight[r * fr.d + a] * v;
    }
}

// Sum of every lane's acc[a] into part[a][0] (lane 0 reads it back) by the fixed tree.
// force_wave_sums stores each wavefront's sums, force_waves_total adds them in the first
// wavefront; a kernel may put other work between the halves as long as a workgroup
// barrier separates them. Call from all FORCE_THREADS lanes.
__device__ __forceinline__ void force_wave_sums(const double (&acc)[3], double (&part)[3][FORCE_WAVES])
{
    double s[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) s[a] = wave_sum(acc[a]);
    if ((threadIdx.x & 63) == 0)
        for (int a = 0; a < 3; ++a) part[a][threadIdx.x >> 6] = s[a];
}

__device__ __forceinline__ void waves_total(double (&part)[3][FORCE_WAVES], int tid)
{
    const int lane = tid & 63;
    if ((tid >> 6) == 0) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const double v = wave_sum(lane < FORCE_WAVES ? part[a][lane] : 0.0);
            if (lane == 0) part[a][0] = v;
        }
    }
}

__device__ __forceinline__ void force_waves_total(double (&part)[3][FORCE_WAVES]) { waves_total(part, threadIdx.x); }

__device__ __forceinline__ void force_tree(const double (&acc)[3], double (&part)[3][FORCE_WAVES])
{
    force_wave_sums(acc, part);
    __syncthreads();
    force_waves_total(part);
}

// One group of the reduction hosted by a workgroup of `lanes` threads (a multiple of 64; tid
// is the linear thread index): thread tid plays lanes tid, tid + lanes, ... (below
// FORCE_THREADS) of the group in turn. Each virtual lane is played by exactly one thread, a
// wavefront always plays 64 consecutive virtual lanes (lanes and FORCE_THREADS are multiples
// of 64, so its trip count is uniform), and every lane sum and every tree node is the one a
// FORCE_THREADS-lane workgroup computes.
// Writes out[a] (a < d) for a single group, else partial[3 grp + a]. Call from all threads
// of the workgroup; ends with a barrier, so `part` may be reused at once.
template <class T>
__device__ __forceinline__ void force_group_hosted(const T* __restrict__ F, const ForceRows<T>& fr, int groups,
                                                   int grp, int lanes, int tid, double (&part)[3][FORCE_WAVES],
                                                   double* __restrict__ out, double* __restrict__ partial)
{
    for (int lane = tid; lane < FORCE_THREADS; lane += lanes) {
        double acc[3] = {0.0, 0.0, 0.0};
        force_lane(F, fr, groups, grp, lane, acc);
        double s[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) s[a] = wave_sum(acc[a]);
        if ((lane & 63) == 0)
            for (int a = 0; a < 3; ++a) part[a][lane >> 6] = s[a];
    }
    __syncthreads();
    waves_total(part, tid);
    if (tid == 0) {
        if (groups == 1) {
            for (int a = 0; a < fr.d; ++a) out[a] = part[a][0];
        } else {
            for (int a = 0; a < 3; ++a) partial[3 * grp + a] = part[a][0];
        }
    }
    __syncthreads();
}

// out[a] = the groups' partial sums added lane-strided in group order, then the tree. Call
// from all FORCE_THREADS lanes (threadIdx.x) of a workgroup.
__device__ __forceinline__ void force_partials_total(const double* __restrict__ partial, int groups, int d,
                                                     double* __restrict__ out, double (&part)[3][FORCE_WAVES])
{
    double acc[3] = {0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < groups; b += FORCE_THREADS)
        for (int a = 0; a < 3; ++a) acc[a] += partial[3 * b + a];
    force_tree(acc, part);
    if (threadIdx.x == 0)
        for (int a = 0; a < d; ++a) out[a] = part[a][0];
}

}  // namespace lbm
