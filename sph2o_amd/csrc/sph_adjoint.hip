// Reverse-mode (adjoint) passes of one WCSPH step for gfx950.
//
// Forward step (WCSPHSolver::step, WCSPHSolver.cpp:14-78), per fluid particle i
// with r = x_i - x_j, gradW_ij = G(|r|) r and psi_b = rho0 V_b:
//   rho_i = sum_j m W_ij + sum_b psi_b W_ib                          (SPHSolver.cpp:42-89)
//   p_i   = max(0, B ((rho_i / rho0)^gamma_t - 1)),  A_i = p_i / rho_i^2 (WCSPHSolver.cpp:80-95)
//   a_i   = - sum_j m (A_i + A_j) gradW_ij - sum_b psi_b (A_i + p_i / rho0^2) gradW_ib  (:132-188)
//           + (nu / dt) sum_j (m / rho_j)(v_j - v_i) W_ij
//           - (nu_b / dt) sum_b (psi_b / rho_i) v_i W_ib                 (XSPHViscosity.cpp:6-42)
//           + sum_{0<|r|<=h} k_ij (-gam m C(|r|) r / |r| - gam (n_i - n_j)),
//             k_ij = 2 rho0 / (rho_i + rho_j)                           (SurfaceTensionAkinci.cpp:35-70)
//   n_i   = h sum_j (m / rho_j) gradW_ij                                (SurfaceTensionAkinci.cpp:13-33)
//   v'    = v + (a + g) dt,  x' = x + v' dt                             (SPHSolver.cpp:139-168)
//
// The Jacobian of gradW is H(r) u = G u + G2 (r . u) r with G2 = G'(|r|) / |r|.
// Neighbour lists are symmetric (|r_ij| and |r_ji| round identically), so each
// pass gathers, for particle k, the terms of every pair (k, j) seen from both
// ends; there are no atomics and the result is deterministic. Three gathers
// follow the dependency chain: (1) a-bar -> x-bar, v-bar, normal-bar, rho-bar
// and the coefficient terms; (2) normal-bar -> rho-bar, x-bar; (3) rho-bar ->
// x-bar. The neighbour sets are piecewise constant in x and carry no gradient;
// the cohesion kernel's jump at |r| = h/2 (Cohesion.h:37) is a measure-zero set
// and contributes none either.
#include "sph_adjoint.hpp"
#include "sph_device.hpp"

namespace sph {

namespace {

struct F3 {
    float x, y, z;
};
__device__ __forceinline__ F3 f3(const float4& q) { return F3{q.x, q.y, q.z}; }
__device__ __forceinline__ F3 operator-(const F3& a, const F3& b) { return F3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ float dot3(const F3& a, const F3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ void axpy(F3& y, float s, const F3& x) {
    y.x += s * x.x;
    y.y += s * x.y;
    y.z += s * x.z;
}
__device__ __forceinline__ float4 f4(const F3& a) { return make_float4(a.x, a.y, a.z, 0.0f); }

// G2 = G'(|r|) / |r| for the cubic spline's gradW = G r (CubicSpline.h:52-74):
// G = l (3q - 2) / h^2 for q <= 1/2, -l (1 - q)^2 / (h^2 q) for 1/2 < q <= 1.
__device__ __forceinline__ float gradFactor2(const SceneConst& S, float rl) {
    const float q = rl * S.invh;
    if (!(rl > 1.0e-9f) || q > 1.0f) return 0.0f;
    const float h3 = S.h * S.h * S.h;
    if (q <= 0.5f) return 3.0f * S.lW / (h3 * rl);
    return S.lW * (1.0f - q * q) / (h3 * (q * q) * rl);
}

// dC/dr of the cohesion kernel in the reference's form (factor 2 inside h/2,
// the unscaled constant has no derivative); branch decided on d as forward.
__device__ __forceinline__ float cohesionDeriv(const SceneConst& S, float d, float rl) {
    const float a = S.h - rl;
    const float base = 3.0f * S.cohC1 * (a * a) * (rl * rl) * (S.h - 2.0f * rl);
    if (d > S.dhalf && d <= S.dmax) return base;
    if (d > 0.0f && d <= S.dhalf) return 2.0f * base;
    return 0.0f;
}

}  // namespace

#define ADJ_PROLOGUE                              \
    const uint32_t lb = logicalBlock();           \
    const int sIdx = a.blockScene[lb];            \
    const SceneConst& S = a.scenes[sIdx];         \
    const float dt = a.state[sIdx].dt;            \
    const uint32_t i = lb * kBlock + threadIdx.x; \
    const bool valid = i < S.fEnd;

// Adjoint of the symplectic Euler update: v-tilde = v'-bar + dt x'-bar is the
// adjoint of v'; a-bar = dt v-tilde; x-bar and v-bar start from x'-bar and
// v-tilde. Reads the adjoints in original order, writes slot order.
__global__ void __launch_bounds__(kBlock) kAdjSeed(AdjView a) {
    ADJ_PROLOGUE
    if (!valid) return;
    const uint32_t o = S.fBegin + a.id[i];
    const float4 gx = a.adjX[o], gv = a.adjV[o];
    const float vx = gv.x + dt * gx.x, vy = gv.y + dt * gx.y, vz = gv.z + dt * gx.z;
    a.vBar[i] = make_float4(vx, vy, vz, 0.0f);
    a.xBar[i] = make_float4(gx.x, gx.y, gx.z, 0.0f);
    a.aBar[i] = make_float4(dt * vx, dt * vy, dt * vz, 0.0f);
}

// Pass 1: adjoint of the pressure force, XSPH viscosity and the surface-tension
// force given a-bar. Produces x-bar (partial), v-bar (final), normal-bar,
// rho-bar (partial, incl. the pressure chain p(rho) and A = p / rho^2) and the
// per-block sums of a-bar . da/dnu and a-bar . da/dgam.
__global__ void __launch_bounds__(kBlock) kAdjForces(AdjView a) {
    ADJ_PROLOGUE
    double nuTerm = 0.0, stTerm = 0.0;
    if (valid) {
        const size_t NP = a.NP;
        const F3 xi = f3(a.x[i]), vi = f3(a.v[i]), ai = f3(a.aBar[i]), ni = f3(a.n[i]);
        const float rhoi = a.rho[i], pi_ = a.p[i];
        const float Ai = pi_ / (rhoi * rhoi);
        const bool xsph = S.viscMethod == 2, st = S.stOn != 0;
        const float invdt = 1.0f / dt;
        const float cnu = S.visc * invdt;
        const float mi = S.m / rhoi;
        F3 gx{0.0f, 0.0f, 0.0f}, gv{0.0f, 0.0f, 0.0f}, gn{0.0f, 0.0f, 0.0f};
        float gr = 0.0f, gA = 0.0f, nuSum = 0.0f, stSum = 0.0f;
        const uint32_t nf = min(a.cntF[i], a.capF);
        for (uint32_t s = 0; s < nf; ++s) {
            const uint32_t j = a.nbrF[s * NP + i];
            const F3 r = xi - f3(a.x[j]);
            const float rl = fastSqrt(dot3(r, r));
            const float G = gradFactor(S, rl), G2 = gradFactor2(S, rl);
            const F3 aj = f3(a.aBar[j]);
            const F3 da = ai - aj;
            const float rda = dot3(r, da);
            const float rhoj = a.rho[j];
            // pressure: A-bar_k = -m sum (a_k - a_j) . gradW; x-bar_k -= m (A_k + A_j) H (a_k - a_j)
            const float Aj = a.p[j] / (rhoj * rhoj);
            gA -= S.m * G * rda;
            const float cP = -S.m * (Ai + Aj);
            axpy(gx, cP * G, da);
            axpy(gx, cP * G2 * rda, r);
            if (xsph) {
                const F3 dv = f3(a.v[j]) - vi;
                const float W = cubicW(S, rl);
                const float mj = S.m / rhoj;
                axpy(gv, cnu * W * mi, aj);
                axpy(gv, -cnu * W * mj, ai);
                gr += cnu * (S.m / (rhoi * rhoi)) * W * dot3(aj, dv);
                axpy(gx, cnu * G * (mj * dot3(ai, dv) - mi * dot3(aj, dv)), r);
                nuSum += mj * W * dot3(ai, dv);
            }
            if (st) {
                const float d = dRef(r.x, r.y, r.z);
                if (d > 0.0f && d <= S.dmax) {
                    const float rs = fastSqrt(d);
                    const float C = cohesionW(S, d, rs), dC = cohesionDeriv(S, d, rs);
                    const float kij = 2.0f * S.rho0 / (rhoi + rhoj);
                    const float D = C / rs;
                    const float Dp = (dC * rs - C) / (rs * rs * rs);   // D'(r) / r
                    const F3 dn = ni - f3(a.n[j]);
                    F3 sv{0.0f, 0.0f, 0.0f};   // S_kj = -gam m D r - gam (n_k - n_j)
                    axpy(sv, -S.st * S.m * D, r);
                    axpy(sv, -S.st, dn);
                    const float den = rhoi + rhoj;
                    gr += -2.0f * S.rho0 * dot3(da, sv) / (den * den);
                    axpy(gn, -S.st * kij, da);
                    axpy(gx, -S.st * S.m * kij * D, da);
                    axpy(gx, -S.st * S.m * kij * Dp * rda, r);
                    stSum += kij * (-S.m * D * dot3(ai, r) - dot3(ai, dn));
                }
            }
        }
        float gpMirror = 0.0f;
        if (S.akinci) {
            const uint32_t nb = min(a.cntB[i], a.capB);
            const float pm = pi_ / (S.rho0 * S.rho0);
            const float cb = (xsph && S.bviscOn) ? S.bvisc * invdt : 0.0f;
            const float vai = dot3(ai, vi);
            for (uint32_t s = 0; s < nb; ++s) {
                const float4 q = a.xb[a.nbrB[s * NP + i]];
                const F3 r = xi - f3(q);
                const float rl = fastSqrt(dot3(r, r));
                const float G = gradFactor(S, rl), G2 = gradFactor2(S, rl);
                const float psi = q.w;
                const float rai = dot3(r, ai);
                gA -= psi * G * rai;
                gpMirror -= psi * G * rai;
                const float cPb = -psi * (Ai + pm);
                axpy(gx, cPb * G, ai);
                axpy(gx, cPb * G2 * rai, r);
                if (cb != 0.0f) {
                    const float W = cubicW(S, rl);
                    axpy(gv, -cb * (psi / rhoi) * W, ai);
                    gr += cb * (psi / (rhoi * rhoi)) * W * vai;
                    axpy(gx, -cb * (psi / rhoi) * vai * G, r);
                }
            }
            gpMirror /= S.rho0 * S.rho0;
        }
        // A = p / rho^2 and p(rho)
        const float gp = gA / (rhoi * rhoi) + gpMirror;
        gr -= 2.0f * gA * pi_ / (rhoi * rhoi * rhoi);
        if (pi_ > 0.0f) gr += gp * S.wcB * S.gamma * powf(rhoi / S.rho0, S.gamma - 1.0f) / S.rho0;
        const float4 x0 = a.xBar[i], v0 = a.vBar[i];
        a.xBar[i] = make_float4(x0.x + gx.x, x0.y + gx.y, x0.z + gx.z, 0.0f);
        a.vBar[i] = make_float4(v0.x + gv.x, v0.y + gv.y, v0.z + gv.z, 0.0f);
        a.nBar[i] = f4(gn);
        a.rhoBar[i] = gr;
        nuTerm = static_cast<double>(nuSum) * static_cast<double>(invdt);
        stTerm = static_cast<double>(stSum);
    }
    const double t0 = blockReduce<false>(nuTerm);
    const double t1 = blockReduce<false>(stTerm);
    if (threadIdx.x == 0) {
        a.partials[lb] = t0;
        a.partials[a.numBlocks + lb] = t1;
    }
}

// Pass 2 (surface tension only): normal-bar through n_i = h sum (m / rho_j) gradW_ij.
__global__ void __launch_bounds__(kBlock) kAdjNormals(AdjView a) {
    ADJ_PROLOGUE
    if (!valid || !S.stOn) return;
    const size_t NP = a.NP;
    const F3 xi = f3(a.x[i]), nbi = f3(a.nBar[i]);
    const float rhoi = a.rho[i];
    const float hm = S.h * S.m;
    F3 gx{0.0f, 0.0f, 0.0f};
    float gr = 0.0f;
    const uint32_t nf = min(a.cntF[i], a.capF);
    for (uint32_t s = 0; s < nf; ++s) {
        const uint32_t j = a.nbrF[s * NP + i];
        const F3 r = xi - f3(a.x[j]);
        const float rl = fastSqrt(dot3(r, r));
        const float G = gradFactor(S, rl), G2 = gradFactor2(S, rl);
        const F3 nbj = f3(a.nBar[j]);
        const float rhoj = a.rho[j];
        gr += hm / (rhoi * rhoi) * G * dot3(nbj, r);
        F3 u{0.0f, 0.0f, 0.0f};   // n-bar_k / rho_j - n-bar_j / rho_k
        axpy(u, 1.0f / rhoj, nbi);
        axpy(u, -1.0f / rhoi, nbj);
        axpy(gx, hm * G, u);
        axpy(gx, hm * G2 * dot3(r, u), r);
    }
    a.rhoBar[i] += gr;
    const float4 x0 = a.xBar[i];
    a.xBar[i] = make_float4(x0.x + gx.x, x0.y + gx.y, x0.z + gx.z, 0.0f);
}

// Pass 3: rho-bar through the density sum; writes the adjoints of the state
// before the step back in original order.
__global__ void __launch_bounds__(kBlock) kAdjDensity(AdjView a) {
    ADJ_PROLOGUE
    if (!valid) return;
    const size_t NP = a.NP;
    const F3 xi = f3(a.x[i]);
    const float rk = a.rhoBar[i];
    F3 gx = f3(a.xBar[i]);
    const uint32_t nf = min(a.cntF[i], a.capF);
    for (uint32_t s = 0; s < nf; ++s) {
        const uint32_t j = a.nbrF[s * NP + i];
        const F3 r = xi - f3(a.x[j]);
        const float G = gradFactor(S, fastSqrt(dot3(r, r)));
        axpy(gx, S.m * (rk + a.rhoBar[j]) * G, r);
    }
    if (S.akinci) {
        const uint32_t nb = min(a.cntB[i], a.capB);
        for (uint32_t s = 0; s < nb; ++s) {
            const float4 q = a.xb[a.nbrB[s * NP + i]];
            const F3 r = xi - f3(q);
            const float G = gradFactor(S, fastSqrt(dot3(r, r)));
            axpy(gx, q.w * rk * G, r);
        }
    }
    const uint32_t o = S.fBegin + a.id[i];
    a.adjX[o] = f4(gx);
    a.adjV[o] = a.vBar[i];
}

// One workgroup per scene: add the step's coefficient terms (block partials in
// a fixed order) to the scene's FP64 accumulators.
__global__ void __launch_bounds__(kBlock) kAdjFinish(AdjView a) {
    const int s = blockIdx.x;
    const SceneConst& C = a.scenes[s];
    double v0 = 0.0, v1 = 0.0;
    for (uint32_t b = C.blockBegin + threadIdx.x; b < C.blockEnd; b += kBlock) {
        v0 += a.partials[b];
        v1 += a.partials[a.numBlocks + b];
    }
    v0 = blockReduce<false>(v0);
    v1 = blockReduce<false>(v1);
    if (threadIdx.x == 0) {
        a.acc[2 * s] += v0;
        a.acc[2 * s + 1] += v1;
    }
}

#define ADJ_LAUNCH(kernel, grid)                                             \
    do {                                                                     \
        if ((grid) == 0) return hipSuccess;                                  \
        hipLaunchKernelGGL(kernel, dim3(grid), dim3(kBlock), 0, st, a);      \
        return hipGetLastError();                                            \
    } while (0)

hipError_t launchAdjSeed(const AdjView& a, hipStream_t st) { ADJ_LAUNCH(kAdjSeed, a.numBlocks); }
hipError_t launchAdjForces(const AdjView& a, hipStream_t st) { ADJ_LAUNCH(kAdjForces, a.numBlocks); }
hipError_t launchAdjNormals(const AdjView& a, hipStream_t st) { ADJ_LAUNCH(kAdjNormals, a.numBlocks); }
hipError_t launchAdjDensity(const AdjView& a, hipStream_t st) { ADJ_LAUNCH(kAdjDensity, a.numBlocks); }
hipError_t launchAdjFinish(const AdjView& a, hipStream_t st) {
    ADJ_LAUNCH(kAdjFinish, static_cast<uint32_t>(a.numScenes));
}

}  // namespace sph
