// Reverse-mode (adjoint) passes of the WCSPH step: launchers and the view of
// one recorded step. Gradients flow from the state after a step back to the
// state before it and to the XSPH viscosity and surface-tension coefficients.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "sph_common.hpp"

namespace sph {

struct AdjView {
    const SceneConst* scenes;
    const SceneState* state;   // dt of each scene (fixed in WCSPH)
    const int32_t* blockScene;
    int32_t numScenes;
    uint32_t numBlocks;
    uint32_t NP;
    // the recorded step, in that step's cell-sorted slot order
    const float4* x;     // positions after the step's reorder
    const float4* v;     // velocities at the start of the step
    const float* rho;
    const float* p;
    const float4* n;     // surface-tension normals (w = density)
    const uint32_t* id;  // slot -> particle index within its scene
    const uint32_t* nbrF;
    const uint32_t* cntF;
    const uint32_t* nbrB;
    const uint32_t* cntB;
    uint32_t capF, capB;
    const float4* xb;    // boundary: position, w = rho0 * volume
    // adjoints of the state after the step on entry, before it on exit,
    // indexed fBegin + particle index (original order)
    float4* adjX;
    float4* adjV;
    // per-slot work arrays
    float4* aBar;
    float4* xBar;
    float4* vBar;
    float4* nBar;
    float* rhoBar;
    double* partials;    // 2 * numBlocks * adjLanesPerParticle(): viscosity and surface-tension terms
    double* acc;         // 2 per scene: dL/d(viscosity), dL/d(surface tension)
    // DFSPH: one Jacobi pass of the recorded step
    const float4* u;     // velocity the pass read
    const float* kk;     // its k (density: 1/dt^2 units, divergence: 1/dt)
    const float* kkNext; // k of the velocity pass that runs next (null: none); the
                         // pass before it files it in the w of v-bar
    float4* su;          // (u, s-bar) of the current iteration, filed by kAdjVelocity
    const float* aux;    // its rho* (density) or clamped divergence
    const float* alpha;  // alpha at the step's positions
    float* kBar;         // adjoint of k of the pass
    float* sBar;         // adjoint of the pass's rho* / divergence
    float* alphaBar;
    float4* sgBar;       // density/alpha adjoint: 2 den-bar S (xyz), den-bar (w)
    // original-order adjoints carried between the halves of a DFSPH step
    float* adjKws;       // of the scaled k_ws / kv_ws leaving the step
    float* adjKvws;
    float* adjRho;       // of rho / alpha at the start-of-step positions
    float* adjAlpha;
};

// Launch blocks per particle block of the adjoint neighbour passes (the
// partials of kAdjForces are per launch block)
int adjLanesPerParticle();

// DFSPH pass kinds of the adjoint kernels
enum AdjMode { kAdjDensityLoop = 0, kAdjDivergenceLoop = 1, kAdjDensityWarm = 2, kAdjDivergenceWarm = 3 };

hipError_t launchAdjSeed(const AdjView& a, hipStream_t st);
hipError_t launchAdjForces(const AdjView& a, hipStream_t st);
hipError_t launchAdjNormals(const AdjView& a, hipStream_t st);
hipError_t launchAdjDensity(const AdjView& a, hipStream_t st);
hipError_t launchAdjFinish(const AdjView& a, hipStream_t st);

// DFSPH (fixed iterations, fixed dt)
hipError_t launchAdjForces2(const AdjView& a, int pressure, hipStream_t st);
// mode: 0 first half, 1 second half, 2 the initial densities after the last reversed step
hipError_t launchAdjHalfStart(const AdjView& a, int mode, hipStream_t st);
hipError_t launchAdjHalfEnd(const AdjView& a, int mode, hipStream_t st);
hipError_t launchAdjVelocity(const AdjView& a, int mode, hipStream_t st);
hipError_t launchAdjSource(const AdjView& a, int mode, hipStream_t st);
hipError_t launchAdjPredict(const AdjView& a, hipStream_t st);
hipError_t launchAdjAlphaA(const AdjView& a, hipStream_t st);
hipError_t launchAdjAlphaB(const AdjView& a, hipStream_t st);

}  // namespace sph
