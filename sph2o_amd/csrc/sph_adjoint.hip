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
struct XV {
    float4 x, v;
};

// G2 = G'(|r|) / |r| for the cubic spline's gradW = G r (CubicSpline.h:52-74):
// G = l (3q - 2) / h^2 for q <= 1/2, -l (1 - q)^2 / (h^2 q) for 1/2 < q <= 1.
__device__ __forceinline__ float gradFactor2(const SceneConst& S, float rl) {
    const float q = rl * S.invh;
    if (!(rl > 1.0e-9f) || q > 1.0f) return 0.0f;
    const float h3 = S.h * S.h * S.h;
    if (q <= 0.5f) return 3.0f * S.lW / (h3 * rl);
    return S.lW * (1.0f - q * q) / (h3 * (q * q) * rl);
}

// G and G2 together from d = |r|^2 with one reciprocal square root, in the
// forward Jacobi passes' form (gradFactorSq): near, G = gA |r| + gB and
// G2 = gA / |r|; far, G = gC (h - |r|)^2 / |r| and G2 = -gC (h^2 - d) / |r|^3.
// d is clamped to 1e-30 as in the forward, so a self pair (r = 0) gets finite
// factors (the near branch) and a term of exactly zero; q > 1 gets 0.
__device__ __forceinline__ void gradFactors(const SceneConst& S, float d, float& G, float& G2) {
    const float dm = fmaxf(d, 1.0e-30f);
    const float rinv = __builtin_amdgcn_rsqf(dm);
    const float rl = dm * rinv;
    const float q = rl * S.invh;
    const float t = S.h - rl;
    const bool near = q <= 0.5f, in = !(q > 1.0f);
    const float gNear = __fmaf_rn(S.gA, rl, S.gB), g2Near = S.gA * rinv;
    const float gFar = (S.gC * (t * t)) * rinv;
    const float g2Far = (-S.gC * (S.h * S.h - dm)) * (rinv * rinv * rinv);
    G = in ? (near ? gNear : gFar) : 0.0f;
    G2 = in ? (near ? g2Near : g2Far) : 0.0f;
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

// Lanes per particle of the adjoint Jacobi passes (kAdjVelocity, kAdjSource):
// as in the forward passes, the L lanes of a particle split its neighbour list
// by slot residue and their sums are combined across lanes; each lane loads
// kPadRows / L slots per round.
#ifndef SPH_ADJ_LPP
#define SPH_ADJ_LPP 4
#endif
constexpr uint32_t kAdjLpp = SPH_ADJ_LPP;
static_assert(kAdjLpp == 1 || kAdjLpp == 2 || kAdjLpp == 4 || kAdjLpp == 8, "SPH_ADJ_LPP must be 1, 2, 4 or 8");
#define ADJ_PROLOGUE_LANES(L)                                                                      \
    const uint32_t lb = logicalBlock();                                                            \
    const int sIdx = a.blockScene[lb / (L)];                                                       \
    const SceneConst& S = a.scenes[sIdx];                                                          \
    const float dt = a.state[sIdx].dt;                                                             \
    const uint32_t sub = (threadIdx.x & 63u) / (64u / (L));                                        \
    const uint32_t i = lb * (kBlock / (L)) + (threadIdx.x >> 6) * (64u / (L)) + threadIdx.x % (64u / (L)); \
    const bool valid = i < S.fEnd;
__device__ __forceinline__ void lanesSum3(F3& v) {
    v.x = lppSum<kAdjLpp>(v.x);
    v.y = lppSum<kAdjLpp>(v.y);
    v.z = lppSum<kAdjLpp>(v.z);
}
// Lane `sub` of kAdjLpp visits list slots sub, sub + kAdjLpp, ... < n of
// particle i (entries s NP + i): per round it loads B list entries, then
// their gathers, then consumes them in slot order. Slots past n load slot
// n - 1 (a valid entry) and are not consumed, so every pair of the list is
// consumed once, the self pair included, and no padding row is.
template <uint32_t B, class G, class C>
__device__ __forceinline__ void laneWalk(const uint32_t* list, size_t NP, uint32_t i, uint32_t sub, uint32_t n,
                                         G&& gather, C&& consume) {
    for (uint32_t s0 = sub; s0 < n; s0 += B * kAdjLpp) {
        uint32_t j[B];
#pragma unroll
        for (uint32_t k = 0; k < B; ++k) j[k] = list[static_cast<size_t>(min(s0 + k * kAdjLpp, n - 1u)) * NP + i];
        decltype(gather(0u)) g[B];
#pragma unroll
        for (uint32_t k = 0; k < B; ++k) g[k] = gather(j[k]);
#pragma unroll
        for (uint32_t k = 0; k < B; ++k)
            if (s0 + k * kAdjLpp < n) consume(g[k]);
    }
}

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
    a.rhoBar[i] = 0.0f;
}

// Pass 1: adjoint of the pressure force, XSPH viscosity and the surface-tension
// force given a-bar. Produces x-bar (partial), v-bar (final), normal-bar,
// rho-bar (partial, incl. the pressure chain p(rho) and A = p / rho^2) and the
// per-block sums of a-bar . da/dnu and a-bar . da/dgam.
__global__ void __launch_bounds__(kBlock) kAdjForces(AdjView a, int pressure) {
    ADJ_PROLOGUE_LANES(kAdjLpp)
    const size_t NP = a.NP;
    const uint32_t ii = valid ? i : 0u;   // invalid lanes run empty loops up to the lane sums
    const F3 xi = f3(a.x[ii]), vi = f3(a.v[ii]), ai = f3(a.aBar[ii]), ni = f3(a.n[ii]);
    const float rhoi = a.rho[ii], pi_ = pressure ? a.p[ii] : 0.0f;
    const float Ai = pi_ / (rhoi * rhoi);
    const bool xsph = S.viscMethod == 2, st = S.stOn != 0;
    const float invdt = 1.0f / dt;
    const float cnu = S.visc * invdt;
    const float mi = S.m / rhoi;
    F3 gx{0.0f, 0.0f, 0.0f}, gv{0.0f, 0.0f, 0.0f}, gn{0.0f, 0.0f, 0.0f};
    float gr = 0.0f, gA = 0.0f, nuSum = 0.0f, stSum = 0.0f;
    const uint32_t nf = valid ? min(a.cntF[i], a.capF) : 0u;
    struct Nb { float4 x, a, v, n; float rho, p; };
    const float4 z4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    laneWalk<2>(
        a.nbrF, NP, ii, sub, nf,
        [&](uint32_t j) {
            return Nb{a.x[j], a.aBar[j], xsph ? a.v[j] : z4, st ? a.n[j] : z4, a.rho[j], pressure ? a.p[j] : 0.0f};
        },
        [&](const Nb& q) {
            const F3 r = xi - f3(q.x);
            const float rl = fastSqrt(dot3(r, r));
            const float G = gradFactor(S, rl), G2 = gradFactor2(S, rl);
            const F3 aj = f3(q.a);
            const F3 da = ai - aj;
            const float rda = dot3(r, da);
            const float rhoj = q.rho;
            if (pressure) {
                // A-bar_k = -m sum (a_k - a_j) . gradW; x-bar_k -= m (A_k + A_j) H (a_k - a_j)
                const float Aj = q.p / (rhoj * rhoj);
                gA -= S.m * G * rda;
                const float cP = -S.m * (Ai + Aj);
                axpy(gx, cP * G, da);
                axpy(gx, cP * G2 * rda, r);
            }
            if (xsph) {
                const F3 dv = f3(q.v) - vi;
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
                    const F3 dn = ni - f3(q.n);
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
        });
    float gpMirror = 0.0f;
    const float pm = pi_ / (S.rho0 * S.rho0);
    if (valid && S.akinci) {
        const uint32_t nb = min(a.cntB[i], a.capB);
        const float cb = (xsph && S.bviscOn) ? S.bvisc * invdt : 0.0f;
        const float vai = dot3(ai, vi);
        for (uint32_t s = sub; s < nb; s += kAdjLpp) {
            const float4 q = a.xb[a.nbrB[s * NP + i]];
            const F3 r = xi - f3(q);
            const float rl = fastSqrt(dot3(r, r));
            const float G = gradFactor(S, rl), G2 = gradFactor2(S, rl);
            const float psi = q.w;
            const float rai = dot3(r, ai);
            if (pressure) {
                gA -= psi * G * rai;
                gpMirror -= psi * G * rai;
                const float cPb = -psi * (Ai + pm);
                axpy(gx, cPb * G, ai);
                axpy(gx, cPb * G2 * rai, r);
            }
            if (cb != 0.0f) {
                const float W = cubicW(S, rl);
                axpy(gv, -cb * (psi / rhoi) * W, ai);
                gr += cb * (psi / (rhoi * rhoi)) * W * vai;
                axpy(gx, -cb * (psi / rhoi) * vai * G, r);
            }
        }
    }
    lanesSum3(gx);
    lanesSum3(gv);
    lanesSum3(gn);
    gr = lppSum<kAdjLpp>(gr);
    gA = lppSum<kAdjLpp>(gA);
    nuSum = lppSum<kAdjLpp>(nuSum);
    stSum = lppSum<kAdjLpp>(stSum);
    gpMirror = lppSum<kAdjLpp>(gpMirror);
    double nuTerm = 0.0, stTerm = 0.0;
    if (valid && sub == 0) {
        gpMirror /= S.rho0 * S.rho0;
        // A = p / rho^2 and p(rho)
        const float gp = gA / (rhoi * rhoi) + gpMirror;
        gr -= 2.0f * gA * pi_ / (rhoi * rhoi * rhoi);
        if (pi_ > 0.0f) gr += gp * S.wcB * S.gamma * powf(rhoi / S.rho0, S.gamma - 1.0f) / S.rho0;
        const float4 x0 = a.xBar[i], v0 = a.vBar[i];
        a.xBar[i] = make_float4(x0.x + gx.x, x0.y + gx.y, x0.z + gx.z, 0.0f);
        a.vBar[i] = make_float4(v0.x + gv.x, v0.y + gv.y, v0.z + gv.z, 0.0f);
        a.nBar[i] = f4(gn);
        a.rhoBar[i] += gr;
        nuTerm = static_cast<double>(nuSum) * static_cast<double>(invdt);
        stTerm = static_cast<double>(stSum);
    }
    // one partial per launch block (kAdjLpp per particle block): kAdjFinish
    const double t0 = blockReduce<false>(nuTerm);
    const double t1 = blockReduce<false>(stTerm);
    if (threadIdx.x == 0) {
        a.partials[lb] = t0;
        a.partials[a.numBlocks * kAdjLpp + lb] = t1;
    }
}

// Pass 2 (surface tension only): normal-bar through n_i = h sum (m / rho_j) gradW_ij.
__global__ void __launch_bounds__(kBlock) kAdjNormals(AdjView a) {
    ADJ_PROLOGUE_LANES(kAdjLpp)
    if (!S.stOn) return;   // uniform over the launch block (one scene)
    const size_t NP = a.NP;
    const uint32_t ii = valid ? i : 0u;
    const F3 xi = f3(a.x[ii]), nbi = f3(a.nBar[ii]);
    const float rhoi = a.rho[ii];
    const float hm = S.h * S.m;
    F3 gx{0.0f, 0.0f, 0.0f};
    float gr = 0.0f;
    const uint32_t nf = valid ? min(a.cntF[i], a.capF) : 0u;
    struct Nb { float4 x, n; float rho; };
    laneWalk<2>(
        a.nbrF, NP, ii, sub, nf, [&](uint32_t j) { return Nb{a.x[j], a.nBar[j], a.rho[j]}; },
        [&](const Nb& q) {
            const F3 r = xi - f3(q.x);
            const float rl = fastSqrt(dot3(r, r));
            const float G = gradFactor(S, rl), G2 = gradFactor2(S, rl);
            const F3 nbj = f3(q.n);
            const float rhoj = q.rho;
            gr += hm / (rhoi * rhoi) * G * dot3(nbj, r);
            F3 u{0.0f, 0.0f, 0.0f};   // n-bar_k / rho_j - n-bar_j / rho_k
            axpy(u, 1.0f / rhoj, nbi);
            axpy(u, -1.0f / rhoi, nbj);
            axpy(gx, hm * G, u);
            axpy(gx, hm * G2 * dot3(r, u), r);
        });
    lanesSum3(gx);
    gr = lppSum<kAdjLpp>(gr);
    if (!valid || sub != 0) return;
    a.rhoBar[i] += gr;
    const float4 x0 = a.xBar[i];
    a.xBar[i] = make_float4(x0.x + gx.x, x0.y + gx.y, x0.z + gx.z, 0.0f);
}

// Pass 3: rho-bar through the density sum; writes the adjoints of the state
// before the step back in original order.
__global__ void __launch_bounds__(kBlock) kAdjDensity(AdjView a) {
    ADJ_PROLOGUE_LANES(kAdjLpp)
    const size_t NP = a.NP;
    const uint32_t ii = valid ? i : 0u;
    const F3 xi = f3(a.x[ii]);
    const float rk = a.rhoBar[ii];
    F3 gx{0.0f, 0.0f, 0.0f};
    const uint32_t nf = valid ? min(a.cntF[i], a.capF) : 0u;
    struct Nb { float4 x; float rb; };
    laneWalk<4>(a.nbrF, NP, ii, sub, nf, [&](uint32_t j) { return Nb{a.x[j], a.rhoBar[j]}; },
                [&](const Nb& q) {
                    const F3 r = xi - f3(q.x);
                    const float G = gradFactor(S, fastSqrt(dot3(r, r)));
                    axpy(gx, S.m * (rk + q.rb) * G, r);
                });
    if (valid && S.akinci) {
        const uint32_t nb = min(a.cntB[i], a.capB);
        for (uint32_t s = sub; s < nb; s += kAdjLpp) {
            const float4 q = a.xb[a.nbrB[s * NP + i]];
            const F3 r = xi - f3(q);
            const float G = gradFactor(S, fastSqrt(dot3(r, r)));
            axpy(gx, q.w * rk * G, r);
        }
    }
    lanesSum3(gx);
    if (!valid || sub != 0) return;
    const float4 x0 = a.xBar[i];
    const uint32_t o = S.fBegin + a.id[i];
    a.adjX[o] = make_float4(x0.x + gx.x, x0.y + gx.y, x0.z + gx.z, 0.0f);
    a.adjV[o] = a.vBar[i];
}

// One workgroup per scene: add the step's coefficient terms (block partials in
// a fixed order) to the scene's FP64 accumulators.
__global__ void __launch_bounds__(kBlock) kAdjFinish(AdjView a) {
    const int s = blockIdx.x;
    const SceneConst& C = a.scenes[s];
    double v0 = 0.0, v1 = 0.0;
    const uint32_t n = a.numBlocks * kAdjLpp;   // kAdjForces' launch blocks
    for (uint32_t b = C.blockBegin * kAdjLpp + threadIdx.x; b < C.blockEnd * kAdjLpp; b += kBlock) {
        v0 += a.partials[b];
        v1 += a.partials[n + b];
    }
    v0 = blockReduce<false>(v0);
    v1 = blockReduce<false>(v1);
    if (threadIdx.x == 0) {
        a.acc[2 * s] += v0;
        a.acc[2 * s + 1] += v1;
    }
}

// ---------------------------------------------------------------- DFSPH
// Reverse passes of one DFSPHSolver::step (DFSPHSolver.cpp:33-107) with a
// fixed iteration count and a fixed dt. A step has two halves: the first at
// the start positions x_n (forces, predictVelocities, the density solve), the
// second at x_{n+1} = x_n + v dt after the neighbour search (densities, alpha,
// the divergence solve). Each half is reversed in its own slot order; the
// adjoints carried between halves and steps (x, v, the scaled k_ws / kv_ws,
// rho and alpha at the start positions) live in original particle order.
//
// One Jacobi iteration is a source pass (rho* or divergence, from velocity u)
// and a velocity pass:
//   rho*_i = max(rho_i + dt sum_j m gradW_ij.(u_i - u_j) + dt sum_b psi_b gradW_ib.u_i, rho0)   (:226-299)
//   k_i    = (rho*_i - rho0) alpha_i / (dt^2 rho_i),  k_ws += k                             (:435-453)
//   d_i    = [n_i >= 20] max(sum_j m gradW_ij.(u_i - u_j) + sum_b psi_b gradW_ib.u_i, 0)    (:538-621)
//   k_i    = d_i alpha_i / (dt rho_i),  kv_ws += k                                         (:750-768)
//   u'_i   = u_i - dt sum_j m (k_i + k_j) gradW_ij [|k_i+k_j| > eps] - dt sum_b psi_b k_i gradW_ib [|k_i| > eps]
// Warm starts take k = k_ws / dt^2 where rho* > rho0 (:873-884), k = kv_ws / dt
// where d > 0 (:977-989). Gates and clamps are piecewise constant.

// Gather the carried adjoints into the slot order of one half.
//   second half: x-bar, u-bar (= v-bar of the step's result), rho-bar and
//     alpha-bar of the half's densities (from the next step's first half)
//   first half: x-bar, v-bar + dt x-bar (x_{n+1} = x_n + v dt), zero rho-bar
//   final (mode 2): x-bar, rho-bar, alpha-bar at the initial positions
__global__ void __launch_bounds__(kBlock) kAdjHalfStart(AdjView a, int mode) {
    ADJ_PROLOGUE
    if (!valid) return;
    const uint32_t o = S.fBegin + a.id[i];
    const float4 gx = a.adjX[o];
    a.xBar[i] = make_float4(gx.x, gx.y, gx.z, 0.0f);
    if (mode == 0) {
        const float4 gv = a.adjV[o];
        a.vBar[i] = make_float4(gv.x + dt * gx.x, gv.y + dt * gx.y, gv.z + dt * gx.z, a.kkNext ? a.kkNext[i] : 0.0f);
        a.rhoBar[i] = 0.0f;
        a.alphaBar[i] = 0.0f;
    } else {
        if (mode == 1) {
            const float4 gv = a.adjV[o];
            a.vBar[i] = make_float4(gv.x, gv.y, gv.z, a.kkNext ? a.kkNext[i] : 0.0f);
        }
        a.rhoBar[i] = a.adjRho[o];
        a.alphaBar[i] = a.adjAlpha[o];
    }
}

// Scatter a half's results back to original order.
__global__ void __launch_bounds__(kBlock) kAdjHalfEnd(AdjView a, int mode) {
    ADJ_PROLOGUE
    if (!valid) return;
    const uint32_t o = S.fBegin + a.id[i];
    a.adjX[o] = a.xBar[i];
    if (mode != 2) a.adjV[o] = a.vBar[i];
    if (mode == 0) {
        a.adjRho[o] = a.rhoBar[i];
        a.adjAlpha[o] = a.alphaBar[i];
    }
}

// The (u_i, s-bar_i) record of a Jacobi iteration's source pass: the velocity
// its forward source pass read and the source adjoint the velocity pass just
// formed, gathered together by kAdjSource.
__device__ __forceinline__ void storeSU(const AdjView& a, uint32_t i) {
    const float4 u = a.u[i];
    a.su[i] = make_float4(u.x, u.y, u.z, a.sBar[i]);
}

// Adjoint of a velocity pass given u'-bar (vBar, left unchanged: u' = u - ...):
// k-bar, x-bar from the gradient's Jacobian, then the pass kind's chain:
// loop passes add the k_ws / kv_ws accumulation and go through k(rho*, alpha,
// rho) / k(d, alpha, rho) to the source adjoint sBar; warm passes give the
// adjoint of the k_ws / kv_ws that entered the step.
__global__ void __launch_bounds__(kBlock) kAdjVelocity(AdjView a, int mode) {
    ADJ_PROLOGUE_LANES(kAdjLpp)
    const size_t NP = a.NP;
    const uint32_t ii = valid ? i : 0u;   // invalid lanes run empty loops up to the lane sums
    const F3 xi = f3(a.x[ii]), ui = f3(a.vBar[ii]);
    const float ki = a.kk[ii];
    F3 gx{0.0f, 0.0f, 0.0f};
    float kb = 0.0f;
    const uint32_t nf = valid ? min(a.cntF[i], a.capF) : 0u;
    // whole rounds of the self-padded list (a self pair adds exactly zero):
    // per slot k_j, x_j and v-bar_j at once, then the sums in slot order
    // v-bar_j carries this pass's k_j in w (filed by the pass before), so a
    // neighbour costs two gathers
    const uint32_t bytes = static_cast<uint32_t>(NP * sizeof(float4));
    const __amdgpu_buffer_rsrc_t rX = __builtin_amdgcn_make_buffer_rsrc(const_cast<float4*>(a.x), 0, bytes, 0x00020000);
    const __amdgpu_buffer_rsrc_t rV = __builtin_amdgcn_make_buffer_rsrc(a.vBar, 0, bytes, 0x00020000);
    walkRows<kAdjLpp, kPadRows / kAdjLpp, true>(
        a.nbrF, a.NP, ii, sub, nf,
        [&](uint32_t j) {
            const uint32_t off = j * static_cast<uint32_t>(sizeof(float4));
            return XV{__builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(rX, off, 0, 0)),
                      __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(rV, off, 0, 0))};
        },
        [&](const XV& q) {
            // a closed gate (|k_i + k_j| <= eps) adds exact zeros: the same sums
            // as skipping the pair, without a branch per pair
            const float ks = ki + q.v.w;
            const bool on = fabsf(ks) > S.eps;
            const F3 r = xi - f3(q.x);
            float G, G2;
            gradFactors(S, dRef(r.x, r.y, r.z), G, G2);
            const F3 du = ui - f3(q.v);
            const float rdu = dot3(r, du);
            kb -= on ? dt * S.m * G * rdu : 0.0f;
            const float c = on ? -dt * S.m * ks : 0.0f;
            axpy(gx, c * G, du);
            axpy(gx, c * G2 * rdu, r);
        });
    if (valid && S.akinci && fabsf(ki) > S.eps) {
        const uint32_t nb = min(a.cntB[i], a.capB);
        for (uint32_t s = sub; s < nb; s += kAdjLpp) {
            const float4 q = a.xb[a.nbrB[s * NP + i]];
            const F3 r = xi - f3(q);
            float G, G2;
            gradFactors(S, dRef(r.x, r.y, r.z), G, G2);
            const float rui = dot3(r, ui);
            kb -= dt * q.w * G * rui;
            const float c = -dt * q.w * ki;
            axpy(gx, c * G, ui);
            axpy(gx, c * G2 * rui, r);
        }
    }
    lanesSum3(gx);
    kb = lppSum<kAdjLpp>(kb);
    if (!valid || sub != 0) return;
    const float4 x0 = a.xBar[i];
    a.xBar[i] = make_float4(x0.x + gx.x, x0.y + gx.y, x0.z + gx.z, 0.0f);
    const uint32_t o = S.fBegin + a.id[i];
    const float rhoi = a.rho[i], al = a.alpha[i], src = a.aux[i];
    const float dt2 = dt * dt;
    if (mode == kAdjDensityLoop) {
        kb += dt2 * a.adjKws[o];
        const float ex = src - S.rho0;
        a.sBar[i] = src > S.rho0 ? kb * al / (dt2 * rhoi) : 0.0f;
        storeSU(a, i);
        a.alphaBar[i] += kb * ex / (dt2 * rhoi);
        a.rhoBar[i] += -kb * ex * al / (dt2 * rhoi * rhoi) + a.sBar[i];
    } else if (mode == kAdjDivergenceLoop) {
        kb += dt * a.adjKvws[o];
        a.sBar[i] = src > 0.0f ? kb * al / (dt * rhoi) : 0.0f;
        storeSU(a, i);
        a.alphaBar[i] += kb * src / (dt * rhoi);
        a.rhoBar[i] += -kb * src * al / (dt * rhoi * rhoi);
    } else if (mode == kAdjDensityWarm) {
        a.adjKws[o] = src > S.rho0 ? kb / dt2 : 0.0f;
    } else {
        a.adjKvws[o] = src > 0.0f ? kb / dt : 0.0f;
    }
}

// Adjoint of a source pass (rho*: c = dt, divergence: c = 1) given its
// adjoint sBar: into u-bar (vBar) and x-bar, with u the velocity it read.
__global__ void __launch_bounds__(kBlock) kAdjSource(AdjView a, int mode) {
    ADJ_PROLOGUE_LANES(kAdjLpp)
    const size_t NP = a.NP;
    const float c = mode == kAdjDensityLoop ? dt : 1.0f;
    const uint32_t ii = valid ? i : 0u;   // invalid lanes run empty loops up to the lane sums
    const F3 xi = f3(a.x[ii]), ui = f3(a.u[ii]);
    const float si = a.sBar[ii];
    F3 gx{0.0f, 0.0f, 0.0f}, gu{0.0f, 0.0f, 0.0f};
    const uint32_t nf = valid ? min(a.cntF[i], a.capF) : 0u;
    // the (u_j, s-bar_j) records the velocity pass filed: two gathers per neighbour
    const uint32_t bytes = static_cast<uint32_t>(NP * sizeof(float4));
    const __amdgpu_buffer_rsrc_t rX = __builtin_amdgcn_make_buffer_rsrc(const_cast<float4*>(a.x), 0, bytes, 0x00020000);
    const __amdgpu_buffer_rsrc_t rU = __builtin_amdgcn_make_buffer_rsrc(a.su, 0, bytes, 0x00020000);
    walkRows<kAdjLpp, kPadRows / kAdjLpp, true>(
        a.nbrF, a.NP, ii, sub, nf,
        [&](uint32_t j) {
            const uint32_t off = j * static_cast<uint32_t>(sizeof(float4));
            return XV{__builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(rX, off, 0, 0)),
                      __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(rU, off, 0, 0))};
        },
        [&](const XV& q) {
            // s_i + s_j = 0 gives w = 0 and exact zeros: no branch per pair
            const float ss = si + q.v.w;
            const F3 r = xi - f3(q.x);
            float G, G2;
            gradFactors(S, dRef(r.x, r.y, r.z), G, G2);
            const F3 du = ui - f3(q.v);
            const float w = c * S.m * ss;
            axpy(gu, w * G, r);
            axpy(gx, w * G, du);
            axpy(gx, w * G2 * dot3(r, du), r);
        });
    if (valid && S.akinci && si != 0.0f) {
        const uint32_t nb = min(a.cntB[i], a.capB);
        for (uint32_t s = sub; s < nb; s += kAdjLpp) {
            const float4 q = a.xb[a.nbrB[s * NP + i]];
            const F3 r = xi - f3(q);
            float G, G2;
            gradFactors(S, dRef(r.x, r.y, r.z), G, G2);
            const float w = c * q.w * si;
            axpy(gu, w * G, r);
            axpy(gx, w * G, ui);
            axpy(gx, w * G2 * dot3(r, ui), r);
        }
    }
    lanesSum3(gx);
    lanesSum3(gu);
    if (!valid || sub != 0) return;
    const float4 x0 = a.xBar[i], v0 = a.vBar[i];
    a.xBar[i] = make_float4(x0.x + gx.x, x0.y + gx.y, x0.z + gx.z, 0.0f);
    a.vBar[i] = make_float4(v0.x + gu.x, v0.y + gu.y, v0.z + gu.z, a.kkNext ? a.kkNext[i] : 0.0f);
}

// predictVelocities (DFSPHSolver.cpp:180-202): v* = v + (a + g) dt, so
// a-bar = dt v*-bar and v-bar = v*-bar.
__global__ void __launch_bounds__(kBlock) kAdjPredict(AdjView a) {
    ADJ_PROLOGUE
    if (!valid) return;
    const float4 v = a.vBar[i];
    a.aBar[i] = make_float4(dt * v.x, dt * v.y, dt * v.z, 0.0f);
}

// computeDensities + computeAlpha (DFSPHSolver.cpp:109-178), part 1:
// alpha = rho / den (den > eps) with den = |S|^2 + sum_j |m gradW_ij|^2 and
// S = sum_j m gradW_ij + sum_b psi_b gradW_ib. rho-bar += alpha-bar / den;
// den-bar = -alpha-bar rho / den^2; sgBar = (2 den-bar S, den-bar).
__global__ void __launch_bounds__(kBlock) kAdjAlphaA(AdjView a) {
    ADJ_PROLOGUE_LANES(kAdjLpp)
    const size_t NP = a.NP;
    const uint32_t ii = valid ? i : 0u;
    const F3 xi = f3(a.x[ii]);
    F3 sg{0.0f, 0.0f, 0.0f};
    float ssq = 0.0f;
    const uint32_t nf = valid ? min(a.cntF[i], a.capF) : 0u;
    laneWalk<4>(a.nbrF, NP, ii, sub, nf, [&](uint32_t j) { return a.x[j]; }, [&](const float4& xj) {
        const F3 r = xi - f3(xj);
        const float g = S.m * gradFactor(S, fastSqrt(dot3(r, r)));
        axpy(sg, g, r);
        ssq += g * g * dot3(r, r);
    });
    if (valid && S.akinci) {
        const uint32_t nb = min(a.cntB[i], a.capB);
        for (uint32_t s = sub; s < nb; s += kAdjLpp) {
            const float4 q = a.xb[a.nbrB[s * NP + i]];
            const F3 r = xi - f3(q);
            axpy(sg, q.w * gradFactor(S, fastSqrt(dot3(r, r))), r);
        }
    }
    lanesSum3(sg);
    ssq = lppSum<kAdjLpp>(ssq);
    if (!valid || sub != 0) return;
    const float den = dot3(sg, sg) + ssq;
    float dbar = 0.0f;
    if (den > S.eps) {
        const float ab = a.alphaBar[i];
        a.rhoBar[i] += ab / den;
        dbar = -ab * a.rho[i] / (den * den);
    }
    a.sgBar[i] = make_float4(2.0f * dbar * sg.x, 2.0f * dbar * sg.y, 2.0f * dbar * sg.z, dbar);
}

// Part 2: x-bar through S, through sum |m gradW|^2 and through the density sum.
__global__ void __launch_bounds__(kBlock) kAdjAlphaB(AdjView a) {
    ADJ_PROLOGUE_LANES(kAdjLpp)
    const size_t NP = a.NP;
    const uint32_t ii = valid ? i : 0u;
    const F3 xi = f3(a.x[ii]);
    const float4 sgi = a.sgBar[ii];
    const F3 si = f3(sgi);
    const float rk = a.rhoBar[ii];
    F3 gx{0.0f, 0.0f, 0.0f};
    const uint32_t nf = valid ? min(a.cntF[i], a.capF) : 0u;
    struct Nb { float4 x, sg; float rb; };
    laneWalk<2>(
        a.nbrF, NP, ii, sub, nf, [&](uint32_t j) { return Nb{a.x[j], a.sgBar[j], a.rhoBar[j]}; },
        [&](const Nb& q) {
            const F3 r = xi - f3(q.x);
            const float rl = fastSqrt(dot3(r, r));
            const float G = gradFactor(S, rl), G2 = gradFactor2(S, rl);
            const F3 w = si - f3(q.sg);
            axpy(gx, S.m * G, w);
            axpy(gx, S.m * G2 * dot3(r, w), r);
            const float cq = 2.0f * S.m * S.m * (sgi.w + q.sg.w) * G * (G + G2 * dot3(r, r));
            axpy(gx, cq + S.m * (rk + q.rb) * G, r);
        });
    if (valid && S.akinci) {
        const uint32_t nb = min(a.cntB[i], a.capB);
        for (uint32_t s = sub; s < nb; s += kAdjLpp) {
            const float4 q = a.xb[a.nbrB[s * NP + i]];
            const F3 r = xi - f3(q);
            const float rl = fastSqrt(dot3(r, r));
            const float G = gradFactor(S, rl), G2 = gradFactor2(S, rl);
            axpy(gx, q.w * G, si);
            axpy(gx, q.w * G2 * dot3(r, si) + q.w * rk * G, r);
        }
    }
    lanesSum3(gx);
    if (!valid || sub != 0) return;
    const float4 x0 = a.xBar[i];
    a.xBar[i] = make_float4(x0.x + gx.x, x0.y + gx.y, x0.z + gx.z, 0.0f);
}

#define ADJ_LAUNCH(kernel, grid)                                             \
    do {                                                                     \
        if ((grid) == 0) return hipSuccess;                                  \
        hipLaunchKernelGGL(kernel, dim3(grid), dim3(kBlock), 0, st, a);      \
        return hipGetLastError();                                            \
    } while (0)

int adjLanesPerParticle() { return static_cast<int>(kAdjLpp); }
hipError_t launchAdjSeed(const AdjView& a, hipStream_t st) { ADJ_LAUNCH(kAdjSeed, a.numBlocks); }
hipError_t launchAdjForces(const AdjView& a, hipStream_t st) { return launchAdjForces2(a, 1, st); }
hipError_t launchAdjForces2(const AdjView& a, int pressure, hipStream_t st) {
    if (a.numBlocks == 0) return hipSuccess;
    hipLaunchKernelGGL(kAdjForces, dim3(a.numBlocks * kAdjLpp), dim3(kBlock), 0, st, a, pressure);
    return hipGetLastError();
}
hipError_t launchAdjNormals(const AdjView& a, hipStream_t st) { ADJ_LAUNCH(kAdjNormals, a.numBlocks * kAdjLpp); }
hipError_t launchAdjDensity(const AdjView& a, hipStream_t st) { ADJ_LAUNCH(kAdjDensity, a.numBlocks * kAdjLpp); }
hipError_t launchAdjFinish(const AdjView& a, hipStream_t st) {
    ADJ_LAUNCH(kAdjFinish, static_cast<uint32_t>(a.numScenes));
}

#define ADJ_LAUNCH_MODE(kernel, mode)                                            \
    do {                                                                         \
        if (a.numBlocks == 0) return hipSuccess;                                 \
        hipLaunchKernelGGL(kernel, dim3(a.numBlocks), dim3(kBlock), 0, st, a, mode); \
        return hipGetLastError();                                                \
    } while (0)

hipError_t launchAdjHalfStart(const AdjView& a, int mode, hipStream_t st) { ADJ_LAUNCH_MODE(kAdjHalfStart, mode); }
hipError_t launchAdjHalfEnd(const AdjView& a, int mode, hipStream_t st) { ADJ_LAUNCH_MODE(kAdjHalfEnd, mode); }
hipError_t launchAdjVelocity(const AdjView& a, int mode, hipStream_t st) {
    if (a.numBlocks == 0) return hipSuccess;
    hipLaunchKernelGGL(kAdjVelocity, dim3(a.numBlocks * kAdjLpp), dim3(kBlock), 0, st, a, mode);
    return hipGetLastError();
}
hipError_t launchAdjSource(const AdjView& a, int mode, hipStream_t st) {
    if (a.numBlocks == 0) return hipSuccess;
    hipLaunchKernelGGL(kAdjSource, dim3(a.numBlocks * kAdjLpp), dim3(kBlock), 0, st, a, mode);
    return hipGetLastError();
}
hipError_t launchAdjPredict(const AdjView& a, hipStream_t st) { ADJ_LAUNCH(kAdjPredict, a.numBlocks); }
hipError_t launchAdjAlphaA(const AdjView& a, hipStream_t st) { ADJ_LAUNCH(kAdjAlphaA, a.numBlocks * kAdjLpp); }
hipError_t launchAdjAlphaB(const AdjView& a, hipStream_t st) { ADJ_LAUNCH(kAdjAlphaB, a.numBlocks * kAdjLpp); }

}  // namespace sph
