// Launch wrappers of the gfx950 kernels (sph_kernels.hip). Every wrapper
// enqueues on `st` and returns hipSuccess or the launch error.
#pragma once

#include <hip/hip_runtime.h>

#include "sph_common.hpp"

namespace sph {

// Device views shared by the launches of one context.
struct DevView {
    const SceneConst* scenes;
    SceneState* state;
    const int32_t* blockScene;  // scene of each fluid block
    int32_t numScenes;
    uint32_t numBlocks;         // fluid blocks (NP / kBlock)
    uint32_t NP;                // padded fluid slots
    // fluid state (sorted slots)
    float4* x;
    float4* v;
    float4* a;
    float4* n;     // surface-tension normals
    float4* xs;    // PCISPH predicted positions
    float4* ap;    // PCISPH pressure acceleration
    float4* xk;    // (x, k) of the current solver pass, read by kVelocity
    float* rho;
    float* alpha;
    float* p;
    float* k;      // k / kv of the current iteration
    float* kws;    // density warm-start accumulator
    float* kvws;   // divergence warm-start accumulator
    float* aux;    // predicted density / divergence error
    uint32_t* id;  // original particle index of each slot
    // boundary (sorted), w = density0 * volume (Akinci) or the particle mass
    // density0 (2r)^3 (PCISPH boundary objects)
    float4* xb;
    float* vb;
    // PCISPH boundary objects (null otherwise): per boundary particle its
    // normal, density and pressure; per fluid slot the predicted velocity and
    // the last iteration's predicted position and velocity (PCISPH only); per
    // cell slot the first step from which the reference's hash table holds the
    // cell (~0: not yet); the index of the step in progress
    float4* bn;
    float* bRho;
    float* bP;
    float4* vs;
    float4* xsP;
    float4* vsP;
    uint32_t* cellFirst;
    uint32_t stepNo;
    uint32_t markNo;   // the step a search's cells exist from (DFSPH's mid-step search: the next)
    // DFSPH: the fluid's normal (FluidObject::n) as the last correctPositions left it,
    // per slot (and the reorder's target)
    float4* nc;
    float4* ncAlt;
    // grid
    uint2* fRange;
    const uint2* bRange;
    uint32_t* cellOf;   // slot -> its cell's fRange index if it is the cell's first slot, else ~0
    // counting sort (per layout): fRange index -> rank within the scene's order
    // indices; order index -> fRange index (~0 for the padding entry); per
    // search: particles per order index, their exclusive prefix sums, and
    // each slot's place within its order index
    const uint32_t* rankOf;
    const uint32_t* ordToCell;
    uint32_t* counts;
    uint32_t* starts;
    uint32_t* li;
    // DFSPH: copy of v and kv_ws as the reorder leaves them (undo of a divergence
    // solve run on lists that turn out invalid); null for other methods
    float4* snapV;
    // recording (null otherwise): the velocity pass also stores its result here
    float4* vRec;
    float* snapKvws;
    // queued WCSPH steps: kDensityPressure's first workgroup moves the last
    // search's flags (flagCount ints at `flags`) to flagLog (flagMode 1: copy,
    // 2: maximum) and clears them for the next search; flagMode 0: untouched
    int32_t* flagLog;
    uint32_t flagCount;
    int32_t flagMode;
    // neighbour lists, column-major: list[slot * NP + i]
    uint32_t* nbrF;
    uint32_t* cntF;
    uint32_t* nbrB;
    uint32_t* cntB;
    uint32_t capF, capB;
    // DFSPH tiles (null elsewhere): per group g of kTileGroup slots,
    // tileSlot[g * kTileMax + t] (t < tileCount[g]) are the fluid slots its
    // lists name; tileCount[g] = ~0 if more than kTileMax. nbrT: the fluid
    // lists as 16-bit tile indices, word [w * NP + i] = rows 8 (w / 4) + w % 4
    // (low half) and that row + 4 (high half)
    uint32_t* tileSlot;
    uint32_t* tileCount;
    uint32_t* nbrT;
    // reductions and flags
    double* partials;   // numBlocks
    int32_t* flags;     // numScenes * kFlagSlots
};

// clearCells: first zero the ranges the last reorder wrote (d.cellOf), so no
// separate fill of the whole range table is needed. counting: keys are order
// indices, counted into d.counts with each slot's place in d.li (counting
// sort); otherwise radix-sort keys with vals = slot.
hipError_t launchKeys(const DevView& d, int dfsphAdvance, int clearCells, int counting, uint32_t* keys,
                      uint32_t* vals, hipStream_t st);
// counting sort, after the exclusive scan of d.counts into d.starts: slot i
// goes to perm[starts[keys[i]] + li[i]] (kScatter); then every order index
// puts its slots back in ascending (stable) order, writes its cell range and
// zeroes its count for the next search (kOrderFix; cells above 256 particles
// raise kFlagBigCell instead and the search is redone by radix sort). occ
// (ordTotal bytes) tracks which indices held slots at the last fix-up; full:
// write every index (the ranges were written by anything else since).
hipError_t launchScatter(const DevView& d, const uint32_t* keys, uint32_t* perm, hipStream_t st);
hipError_t launchOrderFix(const DevView& d, uint32_t* perm, uint32_t ordTotal, uint8_t* occ, int full,
                          hipStream_t st);
hipError_t scanCounts(void* temp, size_t& tempBytes, const uint32_t* counts, uint32_t* starts, uint32_t n,
                      hipStream_t st);
hipError_t launchFillRanges(uint2* r, size_t n, hipStream_t st);
// ranges: record the cell ranges (radix path; the counting sort's kOrderFix
// writes them itself)
hipError_t launchReorder(const DevView& d, const uint32_t* keysSorted, const uint32_t* valsSorted, float4* xAlt,
                         float4* vAlt, float* kwsAlt, float* kvwsAlt, uint32_t* idAlt, int ranges, hipStream_t st);
// lanes per particle: 1, 2, 4 or 8 (the same lists for each)
hipError_t launchBuildNeighbors(const DevView& d, int lanes, hipStream_t st);
// DFSPH tiles of the lists just built (d.tileSlot, d.tileCount, d.nbrT)
hipError_t launchTileLists(const DevView& d, hipStream_t st);
// cube boundary objects of every scene (no-op for scenes without)
hipError_t launchCubeBoundary(const DevView& d, hipStream_t st);
hipError_t launchBoundaryVolume(const DevView& d, int scene, uint32_t bBegin, uint32_t bEnd, hipStream_t st);
// Densities (and pressures) of a PCISPH boundary object's particles: predicted
// = 0 is SPHSolver::computeBoundaryDensities + WCSPHSolver::computeBoundaryPressures
// (cubic spline at current positions, Tait pressure), predicted = 1 is
// PCISPHSolver::predictBoundaryDensities + the boundary part of updatePressure
// (Poly6 at predicted positions, p += delta max(0, rho - rho0)); gated by the
// scene's iteration flag `slot`.
hipError_t launchBoundaryDensity(const DevView& d, int scene, uint32_t bBegin, uint32_t bEnd, int predicted, int slot,
                                 hipStream_t st);

hipError_t launchDensityAlpha(const DevView& d, int withAlpha, hipStream_t st);
hipError_t launchDensityPressure(const DevView& d, hipStream_t st);
hipError_t launchForces(const DevView& d, int pressure, hipStream_t st);
hipError_t launchSurfaceTension(const DevView& d, int vmax, hipStream_t st);
hipError_t launchVmax(const DevView& d, hipStream_t st);
hipError_t launchPredictVelocity(const DevView& d, hipStream_t st);
// slot: iteration flag (0/1 by parity inside a Jacobi loop, 2 outside)
hipError_t launchRhoStar(const DevView& d, int warm, int slot, hipStream_t st);
// fin >= 0: also finish the preceding Jacobi pass's reduction (a FinishKind)
hipError_t launchVelocity(const DevView& d, int slot, int fin, hipStream_t st);
hipError_t launchDivergence(const DevView& d, int warm, int slot, hipStream_t st);
hipError_t launchScaleDivergence(const DevView& d, hipStream_t st);
// Zero n search flags and their accumulator (kernel nodes only: no memset
// in a captured graph).
hipError_t launchFlagClear(int32_t* flags, int32_t* acc, uint32_t n, hipStream_t st);
hipError_t launchIntegrate(const DevView& d, hipStream_t st);

hipError_t launchPciInit(const DevView& d, hipStream_t st);
hipError_t launchPciDensity(const DevView& d, int slot, hipStream_t st);
hipError_t launchPciPressureAcc(const DevView& d, int slot, int fin, hipStream_t st);
hipError_t launchPciIntegrate(const DevView& d, hipStream_t st);

// Reductions: the launches above that reduce (non-warm rho_star / divergence,
// surface_tension with vmax, vmax, pci_density) leave block partials; the
// caller launches kFinish with the matching kind next on the same stream.
enum FinishKind { kFinishDensity = 0, kFinishDivergence = 1, kFinishTimeStep = 2, kFinishPciDensity = 3 };
hipError_t launchFinish(const DevView& d, int kind, int slot, hipStream_t st);

hipError_t launchCellsOut(const DevView& d, int32_t* cells, hipStream_t st);
// count = the scene's active particles; cells (3 * count int32, device) gets
// each particle's cell key at 3 * (original id)
hipError_t launchCellsOutOriginal(const DevView& d, int scene, uint32_t count, int32_t* cells, hipStream_t st);
hipError_t launchCheckFinite(const DevView& d, hipStream_t st);
// Profiling: the next launch made through this file's launchers (one kernel)
// records start / stop at the kernel's own start and end; launchEventsUsed()
// tells whether a launch took them. Per host thread; clear with nullptrs.
void setLaunchEvents(hipEvent_t start, hipEvent_t stop);
bool launchEventsUsed();
// Up to kCopyBatchMax device-to-device copies of whole 16-byte words in one
// launch (the tape of a recorded rollout takes several per step).
constexpr int kCopyBatchMax = 8;
struct CopyBatch {
    const uint4* src[kCopyBatchMax];
    uint4* dst[kCopyBatchMax];
    uint64_t words[kCopyBatchMax];
    int32_t n;
};
hipError_t launchCopyBatch(const CopyBatch& b, hipStream_t st);
// count = the scene's active particles; pos (count float4) and speed (count
// floats) are device buffers indexed by original id, either may be null
hipError_t launchViewerFeed(const DevView& d, int scene, uint32_t count, float4* pos, float* speed, hipStream_t st);

}  // namespace sph

extern "C" int sphKernelsFuseFinish();
// whether the DFSPH Jacobi passes stage LDS tiles (1) or gather from HBM (0)
extern "C" int sphKernelsTiles();
extern "C" int sphKernelsLanesPerParticle();

namespace sph {
// Stable LSD radix sort of (key, value) pairs over the low `bits` key bits
// (rocPRIM's onesweep sort through hipCUB). On return *swapped tells whether
// the sorted data is in the alternate buffers (k1, v1).
hipError_t sortPairs(void* temp, size_t& tempBytes, uint32_t* k0, uint32_t* k1, uint32_t* v0, uint32_t* v1, int n,
                     int bits, hipStream_t st, bool* swapped);
}  // namespace sph
