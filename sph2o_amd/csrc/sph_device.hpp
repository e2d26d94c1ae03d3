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

// Rows [n, n rounded up to kPadRows) of every fluid list hold the particle
// itself (kBuildNeighbors writes them; tapes copy them), whose term in the
// Jacobi passes and their adjoints is exactly zero (r = 0). A lane of those
// passes therefore walks whole rounds of kPadRows = L x B rows with no
// per-row test, and since every lane of a wave is at the same round, a row's
// list address is a wave-uniform base plus the lane's fixed 32-bit offset
// (buffer loads through a scalar descriptor: no vector arithmetic per row).
// Pre: the next round's entries are loaded while this round's gathers are in
// flight.
template <uint32_t L, uint32_t B, bool Pre, class G, class C>
__device__ __forceinline__ void walkRows(const uint32_t* list, uint32_t NP, uint32_t i, uint32_t sub, uint32_t n,
                                         G&& gather, C&& consume) {
    static_assert(L * B == kPadRows, "a round covers kPadRows rows");
    const uint32_t rounds = (n + kPadRows - 1u) / kPadRows;
    if (rounds == 0) return;
    // rows r kPadRows + k L .. + L - 1 as one buffer (a scalar descriptor per
    // row group; the lane reads at its fixed offset within it, < L NP 4 bytes)
    const uint32_t vo = (sub * NP + i) * static_cast<uint32_t>(sizeof(uint32_t));
    const uint32_t groupBytes = L * NP * static_cast<uint32_t>(sizeof(uint32_t));
    auto load = [&](uint32_t r, uint32_t k) {
        const uint32_t* rowK = list + (static_cast<size_t>(r) * kPadRows + k * L) * NP;
        const __amdgpu_buffer_rsrc_t rs =
            __builtin_amdgcn_make_buffer_rsrc(const_cast<uint32_t*>(rowK), 0, groupBytes, 0x00020000);
        return __builtin_amdgcn_raw_buffer_load_b32(rs, vo, 0, 0);
    };
    uint32_t j[B];
#pragma unroll
    for (uint32_t k = 0; k < B; ++k) j[k] = load(0, k);
    for (uint32_t r = 0; r < rounds; ++r) {
        decltype(gather(0u)) g[B];
#pragma unroll
        for (uint32_t k = 0; k < B; ++k) g[k] = gather(j[k]);
        if constexpr (Pre) {
            if (r + 1 < rounds) {
#pragma unroll
                for (uint32_t k = 0; k < B; ++k) j[k] = load(r + 1, k);
            }
        }
#pragma unroll
        for (uint32_t k = 0; k < B; ++k) consume(g[k]);
        if constexpr (!Pre) {
            if (r + 1 < rounds) {
#pragma unroll
                for (uint32_t k = 0; k < B; ++k) j[k] = load(r + 1, k);
            }
        }
    }
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
