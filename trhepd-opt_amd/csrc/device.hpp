// Wave- and workgroup-level helpers shared by the device kernels (included by the .hip
// translation units only).
#pragma once
#include <hip/hip_runtime.h>

namespace trhepd {
namespace {

// Workgroup barrier ordering LDS only: it does not wait for the wave's global loads and
// stores in flight, as a full __syncthreads() does on this target.
__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// the better of two pivot candidates: the larger |re|+|im|, on a tie the smaller row (the
// first maximum, as izamax); the choice is the same whatever the order of combination
__device__ __forceinline__ void pivot_pick(double &best, int &bi, double ob, int oi)
{
    if (ob > best || (ob == best && oi < bi)) {
        best = ob;
        bi = oi;
    }
}

// one DPP exchange of a candidate (the value as two 32-bit halves, and its row)
template <int CTRL>
__device__ __forceinline__ void pivot_dpp(double &best, int &bi)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(best);
    const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)b, CTRL, 0xf, 0xf, false);
    const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTRL, 0xf, 0xf, false);
    const int oi = __builtin_amdgcn_update_dpp(0, bi, CTRL, 0xf, 0xf, false);
    pivot_pick(best, bi, __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo)), oi);
}

// The wave's best candidate, in every lane: within each row of 16 lanes by DPP (lane ^ 1,
// lane ^ 2, then rotations by 4 and 8), then the four rows' results read as scalars.  All
// 64 lanes must be active.
__device__ __forceinline__ void pivot_wave(double &best, int &bi)
{
    pivot_dpp<0xB1>(best, bi);    // quad_perm [1, 0, 3, 2]
    pivot_dpp<0x4E>(best, bi);    // quad_perm [2, 3, 0, 1]
    pivot_dpp<0x124>(best, bi);   // row_ror 4
    pivot_dpp<0x128>(best, bi);   // row_ror 8
    const unsigned long long b = (unsigned long long)__double_as_longlong(best);
    double rb[4];
    int ri[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, 16 * r);
        const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), 16 * r);
        rb[r] = __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
        ri[r] = __builtin_amdgcn_readlane(bi, 16 * r);
    }
    best = rb[0];
    bi = ri[0];
#pragma unroll
    for (int r = 1; r < 4; ++r) pivot_pick(best, bi, rb[r], ri[r]);
}

// a double of lane l (l uniform), from every lane
__device__ __forceinline__ double readlane_f64(double v, int l)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

}  // namespace
}  // namespace trhepd
