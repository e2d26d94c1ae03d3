// Device helpers shared by the forward and adjoint kernels: the reference's
// rounding of |r|, the smoothing kernels, the XCD-aware block order and the
// FP64 block reduction.
#pragma once

#include <hip/hip_runtime.h>

#include "sph_common.hpp"

namespace sph {

// |r|^2 as the reference's compiled glm::length forms it: fma(x,x, fma(z,z, y*y)).
__device__ __forceinline__ float dRef(float x, float y, float z) {
    return __fmaf_rn(x, x, __fmaf_rn(z, z, __fmul_rn(y, y)));
}
// sqrtf, not __fsqrt_rn: for gfx950 the compiler lowers sqrtf to the
// v_sqrt_f32 estimate plus an FMA correction (correctly rounded, as the
// reference's sqrtss), while __fsqrt_rn becomes the bare estimate (1 ulp).
__device__ __forceinline__ float lenRef(float x, float y, float z) { return sqrtf(dRef(x, y, z)); }
__device__ __forceinline__ float fastSqrt(float v) { return __builtin_amdgcn_sqrtf(v); }
__device__ __forceinline__ float fastRcp(float v) { return __builtin_amdgcn_rcpf(v); }

#ifndef SPH_XCD_REMAP
#define SPH_XCD_REMAP 1
#endif
// Logical block of this workgroup. Workgroups are dispatched round-robin over
// the 8 XCDs, each with its own L2; with the remap every XCD walks one
// contiguous eighth of the (Morton-sorted) particle range, so the neighbours
// it gathers stay in its own L2.
#ifndef SPH_CU_REMAP
#define SPH_CU_REMAP 0
#endif
__device__ __forceinline__ uint32_t logicalBlock() {
#if SPH_XCD_REMAP
    const uint32_t n = gridDim.x, b = blockIdx.x;
    const uint32_t q = n >> 3, r = n & 7u, x = b & 7u, k = b >> 3;
#if SPH_CU_REMAP
    // second level: the k-th workgroup of an XCD is taken to land on CU k mod 32;
    // give each CU a contiguous run of logical blocks, so the workgroups it
    // holds at once are Morton neighbours that share their halos in L1
    const uint32_t nx = q + (x < r ? 1u : 0u);
    const uint32_t R = nx >> 5, rem = nx & 31u, c = k & 31u, rr = k >> 5;
    const uint32_t kk = (rr < R || c < rem) ? c * R + min(c, rem) + rr : k;
    return x * q + min(x, r) + kk;
#else
    return x * q + min(x, r) + k;
#endif
#else
    return blockIdx.x;
#endif
}
// CubicSpline::W (CubicSpline.h:27-50)
__device__ __forceinline__ float cubicW(const SceneConst& S, float rl) {
    const float q = rl * S.invh;
    if (q > 1.0f) return 0.0f;
    if (q <= 0.5f) {
        const float q2 = q * q;
        return S.kW * (6.0f * (q2 * q) - 6.0f * q2 + 1.0f);
    }
    const float f = 1.0f - q;
    return S.kW * (2.0f * (f * f * f));
}
// CubicSpline::gradW (CubicSpline.h:52-74) as gradW = r * factor
// Both branches are evaluated and selected (the same operations as branching
// on q, without divergent branches inside the neighbour loops).
__device__ __forceinline__ float gradFactor(const SceneConst& S, float rl) {
    const float q = rl * S.invh;
    const float inv = fastRcp(rl * S.h);
    const float f = 1.0f - q;
    const float near = S.lW * q * (3.0f * q - 2.0f);
    const float far = S.lW * (-f * f);
    const float g = (q <= 0.5f ? near : far) * inv;
    return (rl > 1.0e-9f && !(q > 1.0f)) ? g : 0.0f;
}
// The same factor from d = |r|^2 with one reciprocal square root: for
// q <= 1/2 the factor lW q (3q - 2) / (h |r|) is linear in |r|, gA |r| + gB,
// and beyond it is gC (h - |r|)^2 / |r|. A self pair (d = 0) gets a finite
// factor (its r = 0 makes the term zero), q > 1 gets 0.
__device__ __forceinline__ float gradFactorSq(const SceneConst& S, float d) {
    const float dm = fmaxf(d, 1.0e-30f);
    const float rinv = __builtin_amdgcn_rsqf(dm);
    const float rl = dm * rinv;
    const float q = rl * S.invh;
    const float t = S.h - rl;
    const float near = __fmaf_rn(S.gA, rl, S.gB);
    const float far = (S.gC * (t * t)) * rinv;
    const float g = q <= 0.5f ? near : far;
    return q > 1.0f ? 0.0f : g;
}
// ViscoK::laplW (ViscoK.h:29-37)
__device__ __forceinline__ float laplW(const SceneConst& S, float rl) {
    return rl <= S.h ? S.laplC * (S.h - rl) : 0.0f;
}
// Cohesion::W (Cohesion.h:20-40). The branch at r = h/2, where the
// reference's form jumps by 2x, is decided exactly on d = |r|^2 in the
// reference's rounding (d > dhalf <=> sqrt_rn(d) > h/2); rl is the length.
__device__ __forceinline__ float cohesionW(const SceneConst& S, float d, float rl) {
    if (d > S.dhalf && d <= S.dmax) {
        const float a = S.h - rl;
        return S.cohC1 * (a * a * a) * (rl * rl * rl);
    }
    if (d > 0.0f && d <= S.dhalf) {
        const float a = S.h - rl;
        return S.cohC1 * 2.0f * (a * a * a) * (rl * rl * rl) - S.cohC2;
    }
    return 0.0f;
}

template <bool MAX>
__device__ __forceinline__ double combine(double a, double b) {
    return MAX ? fmax(a, b) : a + b;
}
// Sum over the L lanes that share one particle: lanes l, l + 64/L, l + 128/L,
// ... of a wave (every one of them gets the same value).
template <uint32_t L>
__device__ __forceinline__ float lppSum(float v) {
#pragma unroll
    for (uint32_t off = 64u / L; off < 64u; off <<= 1) v += __shfl_xor(v, static_cast<int>(off), 64);
    return v;
}

template <bool MAX>
__device__ double blockReduce(double v) {
    __shared__ double sm[kBlock / 64];
    for (int off = 32; off > 0; off >>= 1) v = combine<MAX>(v, __shfl_down(v, off, 64));
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sm[wid] = v;
    __syncthreads();
    double t = sm[0];
    for (int w = 1; w < kBlock / 64; ++w) t = combine<MAX>(t, sm[w]);
    return t;  // valid in every thread
}

}  // namespace sph
