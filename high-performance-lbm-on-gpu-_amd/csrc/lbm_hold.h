// Keeping loaded values in registers at a chosen point of the program.
//
// Where a load's result is used only on some paths (a node that turns out to be solid, an
// affine term whose coefficient is zero), the compiler moves the load into those paths:
// the loads of one node or one affine row then go out one at a time, each behind a branch
// and a wait, instead of all together. An empty statement that takes the values as
// register inputs, placed right after the loads, makes every path need them there, so the
// loads are issued together and waited for once. The statement emits no instructions.
#pragma once

#include <hip/hip_runtime.h>

namespace lbm {

template <class T>
__device__ __forceinline__ void hold_value(const T& v)
{
    asm volatile("" : : "v"(v));
}

template <class T, int N>
__device__ __forceinline__ void hold_loaded(const T (&v)[N])
{
#pragma unroll
    for (int i = 0; i < N; ++i) hold_value(v[i]);
}

}  // namespace lbm
