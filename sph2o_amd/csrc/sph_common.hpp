// Shared host/device definitions of the MI355X particle step.
#pragma once

#include <cstdint>

namespace sph {

constexpr int kBlock = 256;  // particles per workgroup (4 wave64)
// DFSPH Jacobi tiles: groups of kTileGroup consecutive slots (one launch block
// of the 4-lanes-per-particle passes) stage the records of at most kTileMax
// distinct neighbours in LDS
constexpr uint32_t kTileGroup = 64;
// Rows [n, n rounded up to kPadRows) of every particle's fluid list hold the
// particle itself (a zero term in the Jacobi passes and their adjoints)
constexpr uint32_t kPadRows = 8;
constexpr uint32_t kTileMax = 512;

// Per-scene constants, uploaded at init / regrid. One scene = one SPH2O
// single-fluid scene with its own grid, particle ranges and coefficients.
constexpr int kMaxCubes = 4;   // cube boundary objects per scene

struct SceneConst {
    // grid: cell key = floor(x * invh), stored relative to the origin cell
    float h, invh;
    int32_t g0x, g0y, g0z;     // origin cell
    int32_t gnx, gny, gnz;     // cells per axis
    uint32_t cellBase;         // first entry of this scene in the cell-range arrays
    uint32_t keyBase;          // first sort key of this scene
    uint32_t deadKey;          // sort key of padding slots (last key of the scene's range)
    int32_t morton;            // 1: Morton-ordered keys, 0: x-fastest linear keys
    // particle slots
    uint32_t fBegin, fEnd;     // fluid slots [fBegin, fEnd)
    uint32_t fPadEnd;          // padded end (multiple of kBlock)
    uint32_t blockBegin, blockEnd;
    uint32_t bBegin, bEnd;     // boundary particles [bBegin, bEnd)
    // physics (reference FP32 values)
    float m, rho0, gx, gy, gz;
    float kW, lW, laplC, cohC1, cohC2;  // CubicSpline k,l; ViscoK laplCoef; Cohesion valueCoef1/2
    // gradW = G(|r|) r with G = gA |r| + gB for |r| <= h/2 and gC (h - |r|)^2 / |r|
    // beyond: 3 lW / h^3, -2 lW / h^2, -lW / h^3 (CubicSpline.h:52-74 rearranged)
    float gA, gB, gC;
    float visc, bvisc, st;
    int32_t viscMethod;        // -1 off, 0 standard, 1 artificial, 2 XSPH
    int32_t stOn, bviscOn, akinci;
    float wcB, gamma;          // WCSPH: stiffness*rho0/gamma, gamma
    float pciDelta;            // PCISPH scaling factor
    float eps;                 // DFSPH gate 1e-5
    float h2c;                 // 0.01*h*h (artificial viscosity)
    // Exact forms of sqrt_rn(d) <= h and sqrt_rn(d) <= h/2 for d = fma(x,x, fma(z,z, y*y)):
    // the largest floats whose correctly rounded square root does not exceed h (h/2).
    float dmax, dhalf;
    // solver control (reference FP32 values)
    float eta, etaV, cfl, radius, minTs, maxTs;
    int32_t minIt, maxIt, minItV, adaptive;
    int32_t nf;
    // Cells outside the dense box hash to ovMask + 1 overflow buckets: sort
    // keys deadKey - (ovMask + 1) .. deadKey - 1, ranges at cellBase + nCells.
    uint32_t ovMask;
    uint32_t nCells;           // dense cells gnx * gny * gnz
    // counting sort: the scene's sort positions ("order indices") start at
    // ordBase: dense cells in key order, then the overflow buckets, then one
    // for the padding slots
    uint32_t ordBase;
    // Akinci adhesion to the boundary (AdhesionForce.cpp:5-32): on if adhOn
    int32_t adhOn;
    float adh;                 // beta
    // PCISPH boundary object (PCISPHBoundaryObject.cpp): on if pciB; normalFct,
    // tangentialFct, the rest radius h/2 of the collision corrections, Poly6's
    // and Spiky's coefficients 315 / (64 pi h^9) and 45 / (pi h^6) and h^2, as
    // Poly6.h / Spiky.h round them
    int32_t pciB;
    float nFct, tFct, restR, poly6C, spikyC, h2;
    // cube boundary objects (CubeBoundaryObject.cpp:10-88), applied in order
    int32_t nCubes;
    float cubeLo[kMaxCubes][3], cubeHi[kMaxCubes][3];
    float cubeNormal[kMaxCubes], cubeTang[kMaxCubes];
};

// Per-scene solver state, owned by the device while a context steps: the
// time step, the Jacobi-loop control and the statistics of the last step.
// The iteration flags are double-buffered by parity: iteration n's passes read
// active[n & 1] while the last block of its error reduction writes
// active[(n + 1) & 1]; active[2] is always 1 (passes outside a loop).
struct SceneState {
    double err;              // error of the solve in progress (PCISPH: carried across steps)
    double errD, errV;       // final density / divergence errors of the last step
    double vmax;             // FP64 CFL maximum of the last step
    float dt, dt2, t;
    int32_t its, itsV;       // iterations of the last density / divergence solve
    int32_t it;              // iterations done in the solve in progress
    int32_t active[3];
    int32_t capped;          // bit 0: density loop capped, bit 1: divergence loop capped
    int32_t pad[2];
};

// Device-side flag slots per scene (int32 each): particles outside the dense
// grid (count), largest fluid / boundary neighbour counts, non-finite
// positions, coordinates beyond the representable cell range.
enum FlagSlot {
    kFlagOutOfGrid = 0,
    kFlagMaxFluidNbr = 1,
    kFlagMaxBoundaryNbr = 2,
    kFlagNonFinite = 3,
    kFlagFar = 4,
    kFlagBigCell = 5,   // a cell held more particles than the counting sort orders in place
    kFlagSlots = 6
};
// Reduction output slots per scene (double each).
enum RedSlot { kRedSum = 0, kRedMax = 1, kRedSlots = 2 };

}  // namespace sph
