// md_reduce.h -- small device helpers shared by the step kernel (step.hip) and
// the persistent small-system kernel (persist.hip): the minimum image, the
// LJ-only pair force factor, and lane-group / wavefront sums with a fixed
// order of additions (every caller gets the same bits for the same inputs).
#pragma once
#include <hip/hip_runtime.h>

namespace {

// minimum-image component, d - L rint(d / L) (src/neighborlist.cu:271-273)
__device__ __forceinline__ float minimg(float d, float L, float Linv) {
    return fmaf(-L, rintf(d * Linv), d);
}

// LJ-only pairs: r^-2 from one reciprocal, no square root.
__device__ __forceinline__ float pair_ff_lj(float rsq, const float4 lj) {
    const float ri2 = __builtin_amdgcn_rcpf(rsq);
    const float ri6 = ri2 * ri2 * ri2;
    return fmaf(lj.x, ri6, -lj.y) * (ri6 * ri2);   // lj1 r^-14 - lj2 r^-8
}

// x from another lane of the row by a DPP pattern (both 32-bit halves).
template <int CTRL>
__device__ __forceinline__ double dpp_d(double x) {
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

template <int CTRL>
__device__ __forceinline__ float dpp_f(float x) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), CTRL, 0xF, 0xF, false));
}

// Sum of x over the G lanes of each aligned lane group (G = 4, 8 or 16), the
// same bits in every lane of the group: pairs and quads by quad permutations,
// then mirrors within 8 and 16 lanes.
template <int G>
__device__ __forceinline__ float group_sum(float x) {
    static_assert(G == 4 || G == 8 || G == 16, "lanes per atom");
    x += dpp_f<0xB1>(x);                          // quad_perm [1, 0, 3, 2]
    x += dpp_f<0x4E>(x);                          // quad_perm [2, 3, 0, 1]
    if constexpr (G >= 8) x += dpp_f<0x141>(x);   // row_half_mirror
    if constexpr (G >= 16) x += dpp_f<0x140>(x);  // row_mirror
    return x;
}

// Sum of x over the 64 lanes of the wavefront (every lane active), the same bits
// in every lane: pairs, quads (quad permutations), halves of 16-lane rows and
// rows (mirrors; addition is commutative, so both partners get the same sum),
// then (row 0 + row 1) + (row 2 + row 3) from lanes 0, 16, 32 and 48.
__device__ __forceinline__ double wave_sum_d(double x) {
    x += dpp_d<0xB1>(x);    // quad_perm [1, 0, 3, 2]
    x += dpp_d<0x4E>(x);    // quad_perm [2, 3, 0, 1]
    x += dpp_d<0x141>(x);   // row_half_mirror
    x += dpp_d<0x140>(x);   // row_mirror
    const long long b = __double_as_longlong(x);
    double r[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)b, 16 * q);
        const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(b >> 32), 16 * q);
        r[q] = __longlong_as_double(((long long)hi << 32) | (long long)lo);
    }
    return (r[0] + r[1]) + (r[2] + r[3]);
}

}  // namespace
