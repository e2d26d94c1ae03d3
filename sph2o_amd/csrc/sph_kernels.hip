// gfx950 kernels of the SPH particle step (WCSPH / PCISPH / DFSPH with
// Akinci boundaries, viscosity and Akinci surface tension).
//
// Layout: fluid particles of every scene live in one padded slot array,
// sorted by cell key (Morton order inside each scene's grid). Each scene
// owns whole 256-slot blocks, so a workgroup always belongs to one scene and
// reads that scene's constants through the scalar cache. Neighbour lists are
// column-major (list[slot * NP + i]) so the 64 lanes of a wave read one
// contiguous 256-byte row per neighbour slot.
//
// Numerics follow the reference's FP32 formulas (CubicSpline.h:27-74,
// ViscoK.h:29-37, Cohesion.h:20-40 and the per-pass sums cited below). The
// cell key and the support test reproduce what the reference computes under
// its shipped -O3 -ffast-math build: key = floor(x * fl(1/h)) and
// |r| = sqrt(fma(x,x, fma(z,z, y*y))), both with explicitly rounded
// operations so the neighbour sets and the cohesion branch at r = h/2 match
// bit for bit. Inside the smooth kernel sums the hardware sqrt/rcp are used.
// Convergence and CFL reductions are FP64 and deterministic (fixed-order
// block partials, combined per scene by a one-block-per-scene kFinish launch).
#include <hip/hip_runtime.h>

#include <hip/hip_ext.h>

#include "sph_device.hpp"
#include "sph_kernels.hpp"

namespace sph {

// ---------------------------------------------------------------- helpers
// Build-time switch (a measured variant): non-temporal neighbour-list loads.
#ifndef SPH_NT_LISTS
#define SPH_NT_LISTS 0
#endif
// Rows of the fluid neighbour list that kBuildNeighbors stages in LDS (4 B x
// 256 lanes per row) and stores as whole 1 KB rows at the end; entries past
// the staged rows are stored directly. 0 stores every entry directly.
// Lanes per particle of the DFSPH Jacobi passes: the lanes of a wave take
// 64 / L particles and split each one's neighbour list by slot residue, so one
// gather instruction reads the neighbours of fewer cells. Measured on the
// 1M-particle dam break (ms/step, source passes / velocity pass lanes):
// 1/1 2.75, 2/2 2.46, 4/4 2.39, 8/8 2.65, with 4 best for the source passes
// (kRhoStar, kDivergence) and 2 for the velocity pass.
#ifndef SPH_LPP
#define SPH_LPP 4
#endif
#ifndef SPH_LPP_VEL
#define SPH_LPP_VEL 2
#endif
constexpr uint32_t kLpp = SPH_LPP;        // kRhoStar, kDivergence
constexpr uint32_t kLppVel = SPH_LPP_VEL; // kVelocity
static_assert(kLpp == 1 || kLpp == 2 || kLpp == 4 || kLpp == 8, "SPH_LPP must be 1, 2, 4 or 8");
static_assert(kLppVel == 1 || kLppVel == 2 || kLppVel == 4 || kLppVel == 8, "SPH_LPP_VEL must be 1, 2, 4 or 8");
// Lanes per particle of the once-per-step neighbour passes (kDensityAlpha,
// kForces, kSurfaceTension; kVmax follows kSurfaceTension's partial layout).
// Measured on the 1M-particle dam break (batched list loads): 1 lane 99 / 87 /
// 59 us for surface tension / forces / density, 2 lanes 85 / 74 / 56.
#ifndef SPH_LPP_ONCE
#define SPH_LPP_ONCE 2
#endif
constexpr uint32_t kLppOnce = SPH_LPP_ONCE;
static_assert(kLppOnce == 1 || kLppOnce == 2 || kLppOnce == 4 || kLppOnce == 8, "SPH_LPP_ONCE must be 1, 2, 4 or 8");
// launch blocks per particle block of the widest partial-writing launch
extern "C" __host__ int sphKernelsLanesPerParticle() { return static_cast<int>(kLpp > kLppOnce ? kLpp : kLppOnce); }

// List rows the search stages in LDS per workgroup (the rest are stored
// directly). 16 rows leave room for 8 waves per SIMD; on D1M 8 / 16 / 24 / 32
// rows gave 1.720 / 1.681 / 1.690 / 1.712 ms per step (the same lists).
#ifndef SPH_LIST_STAGE
#define SPH_LIST_STAGE 16
#endif
// Velocity and position of slot idx (separate arrays: interleaved (x, v)
// records measured slower once the passes ran several lanes per particle).
#define VEL(dv, idx) (dv).v[idx]
#define POSV(dv, idx) (dv).x[idx]
// Whether the Jacobi-loop reduction rides on the velocity / PCISPH pressure
// acceleration launch (1) or gets a kFinish launch of its own (0).
#ifndef SPH_FUSE_FINISH
#define SPH_FUSE_FINISH 1
#endif
extern "C" __host__ int sphKernelsFuseFinish() { return SPH_FUSE_FINISH; }

static_assert(SPH_LIST_STAGE % kPadRows == 0, "the staged rows end on a round boundary");
__device__ __forceinline__ uint32_t ldList(const uint32_t* p) {
#if SPH_NT_LISTS
    return __builtin_nontemporal_load(p);
#else
    return *p;
#endif
}
// List slots a lane of a neighbour pass loads per round. A slot's gather
// depends on its list load, so one slot per round costs two memory latencies
// per neighbour; a round of B slots loads B list entries, then B gathers, then
// accumulates them in slot order. The additions keep their order for every B;
// where the compiler chooses which product of a sum to fuse into an FMA
// (passes without spelled-out FMAs), results of different B agree to
// rounding, not bit for bit (one binary is deterministic). Measured on the
// 1M-particle dam break: 2.28 ms/step at B = 1, 2.12 at 2, 2.04 at 4.
#ifndef SPH_BATCH
#define SPH_BATCH 4
#endif
constexpr uint32_t kBatch = SPH_BATCH;
// Whether the source passes (kRhoStar, kDivergence) load the next round's
// list entries while a round's gathers are in flight.
#ifndef SPH_PRE_SRC
#define SPH_PRE_SRC 1
#endif
// Lane `sub` of L visits slots sub, sub + L, ... < n of particle i's list
// (row-major, NP entries per row): gather(j) per slot, then consume(j, g) in
// slot order. Slots past n load slot n - 1 (a valid entry) and are not consumed.
// Pre: load the next round's list entries while this round's gathers are in
// flight (B more registers; pays where a lane has several rounds).
template <uint32_t L, uint32_t B, bool Pre = false, class G, class C>
__device__ __forceinline__ void walkList(const uint32_t* list, size_t NP, uint32_t i, uint32_t sub, uint32_t n,
                                         G&& gather, C&& consume) {
    if constexpr (!Pre) {
        for (uint32_t s0 = sub; s0 < n; s0 += B * L) {
            uint32_t j[B];
#pragma unroll
            for (uint32_t k = 0; k < B; ++k) j[k] = ldList(list + static_cast<size_t>(min(s0 + k * L, n - 1u)) * NP + i);
            decltype(gather(0u)) g[B];
#pragma unroll
            for (uint32_t k = 0; k < B; ++k) g[k] = gather(j[k]);
#pragma unroll
            for (uint32_t k = 0; k < B; ++k)
                if (s0 + k * L < n) consume(j[k], g[k]);
        }
        return;
    }
    if (sub >= n) return;
    uint32_t j[B];
#pragma unroll
    for (uint32_t k = 0; k < B; ++k) j[k] = ldList(list + static_cast<size_t>(min(sub + k * L, n - 1u)) * NP + i);
    for (uint32_t s0 = sub; s0 < n; s0 += B * L) {
        decltype(gather(0u)) g[B];
#pragma unroll
        for (uint32_t k = 0; k < B; ++k) g[k] = gather(j[k]);
        const uint32_t s1 = s0 + B * L;
        uint32_t jn[B];
#pragma unroll
        for (uint32_t k = 0; k < B; ++k)
            jn[k] = s1 < n ? ldList(list + static_cast<size_t>(min(s1 + k * L, n - 1u)) * NP + i) : 0u;
#pragma unroll
        for (uint32_t k = 0; k < B; ++k)
            if (s0 + k * L < n) consume(j[k], g[k]);
#pragma unroll
        for (uint32_t k = 0; k < B; ++k) j[k] = jn[k];
    }
}
// The tile of launch block g (logical block = tile group: 64 particles of a
// 4-lanes-per-particle pass): stages A[slot] (and B[slot]) of its tile's
// slots into tA (tB) and returns true; false (nothing staged) when the
// context has no tiles, the block's scene sits this pass out, or the group's
// tile overflowed. Every condition is uniform over the block.
__device__ __forceinline__ bool stageTile(const DevView& d, uint32_t g, bool active, const float4* A, float4* tA,
                                          const float4* B, float4* tB) {
    if (!d.tileCount || !active) return false;
    const uint32_t T = d.tileCount[g];
    if (T > kTileMax) return false;
    const uint32_t* slots = d.tileSlot + static_cast<size_t>(g) * kTileMax;
    const uint32_t bytes = d.NP * static_cast<uint32_t>(sizeof(float4));
    const __amdgpu_buffer_rsrc_t rA = __builtin_amdgcn_make_buffer_rsrc(const_cast<float4*>(A), 0, bytes, 0x00020000);
    const __amdgpu_buffer_rsrc_t rB =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<float4*>(B ? B : A), 0, bytes, 0x00020000);
    for (uint32_t t = threadIdx.x; t < T; t += kBlock) {
        const uint32_t off = slots[t] * static_cast<uint32_t>(sizeof(float4));
        tA[t] = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(rA, off, 0, 0));
        if (B) tB[t] = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(rB, off, 0, 0));
    }
    __syncthreads();
    return true;
}
// Tile-index walk of particle i's list by lane `sub` of 4: word row 4R + sub
// holds rows 8R + sub and 8R + sub + 4, the rows walkRows<4, 2> gives the lane
// in the same order, so the sums match the untiled walk bit for bit.
template <class C>
__device__ __forceinline__ void walkTile(const DevView& d, uint32_t i, uint32_t sub, uint32_t n, C&& consume) {
    const uint32_t rounds = (n + kPadRows - 1u) / kPadRows;
    if (rounds == 0) return;
    const uint32_t NP = d.NP;
    const uint32_t vo = (sub * NP + i) * static_cast<uint32_t>(sizeof(uint32_t));
    const uint32_t groupBytes = 4u * NP * static_cast<uint32_t>(sizeof(uint32_t));
    auto load = [&](uint32_t R) {
        const uint32_t* row = d.nbrT + static_cast<size_t>(R) * 4u * NP;
        const __amdgpu_buffer_rsrc_t rs =
            __builtin_amdgcn_make_buffer_rsrc(const_cast<uint32_t*>(row), 0, groupBytes, 0x00020000);
        return __builtin_amdgcn_raw_buffer_load_b32(rs, vo, 0, 0);
    };
    uint32_t w = load(0);
    for (uint32_t R = 0; R < rounds; ++R) {
        const uint32_t cur = w;
        if (R + 1 < rounds) w = load(R + 1);
        consume(cur & 0xffffu);
        consume(cur >> 16);
    }
}
struct XV { float4 x, v; };
// The Jacobi passes spell out their FMAs (r^2 as dRef, dot products and the
// accumulations): left to the compiler, which product of a sum such as
// a x + b y it fuses can change with the unrolling or the gather layout, and
// the iteration counts follow the last bit.
__device__ __forceinline__ float pairFma(float a, float b, float c) { return __fmaf_rn(a, b, c); }
__device__ __forceinline__ float dot3f(float ax, float ay, float az, float bx, float by, float bz) {
    return __fmaf_rn(az, bz, __fmaf_rn(ay, by, __fmul_rn(ax, bx)));
}
struct XVP { float4 x, v; float p; };
__device__ __forceinline__ int floorCell(float x, float invh) { return static_cast<int>(floorf(__fmul_rn(x, invh))); }

__device__ __forceinline__ uint32_t spread3(uint32_t v) {  // 10 bits -> every third bit
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000ffu;
    v = (v | (v << 8)) & 0x0300f00fu;
    v = (v | (v << 4)) & 0x030c30c3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

// local cell coordinates (clamped into the grid); returns false if clamped
__device__ __forceinline__ bool localCell(const SceneConst& S, float4 p, int& cx, int& cy, int& cz) {
    bool ok = isfinite(p.x) && isfinite(p.y) && isfinite(p.z);
    int x = ok ? floorCell(p.x, S.invh) - S.g0x : 0;
    int y = ok ? floorCell(p.y, S.invh) - S.g0y : 0;
    int z = ok ? floorCell(p.z, S.invh) - S.g0z : 0;
    // guard against int overflow from huge coordinates: floorf of |x*invh| > 2^31 is UB on cast
    if (ok && (fabsf(p.x * S.invh) > 1.0e9f || fabsf(p.y * S.invh) > 1.0e9f || fabsf(p.z * S.invh) > 1.0e9f)) {
        ok = false;
        x = y = z = 0;
    }
    const bool in = ok && x >= 0 && y >= 0 && z >= 0 && x < S.gnx && y < S.gny && z < S.gnz;
    cx = min(max(x, 0), S.gnx - 1);
    cy = min(max(y, 0), S.gny - 1);
    cz = min(max(z, 0), S.gnz - 1);
    return in;
}
__device__ __forceinline__ uint32_t cellLinear(const SceneConst& S, int cx, int cy, int cz) {
    return static_cast<uint32_t>((cz * S.gny + cy) * S.gnx + cx);
}
__device__ __forceinline__ uint32_t cellKey(const SceneConst& S, int cx, int cy, int cz) {
    if (S.morton) return S.keyBase + (spread3(cx) | (spread3(cy) << 1) | (spread3(cz) << 2));
    return S.keyBase + cellLinear(S, cx, cy, cz);
}

// Global cell floor(x * fl(1/h)) of a position. Returns 0, or 1 for a
// non-finite coordinate, 2 for one beyond +-1e8 cells (the cell indices and
// their differences must stay exact in int32).
__device__ __forceinline__ int globalCell(const SceneConst& S, float4 p, int& gx, int& gy, int& gz) {
    gx = gy = gz = 0;
    if (!(isfinite(p.x) && isfinite(p.y) && isfinite(p.z))) return 1;
    const float ax = __fmul_rn(p.x, S.invh), ay = __fmul_rn(p.y, S.invh), az = __fmul_rn(p.z, S.invh);
    if (fabsf(ax) > 1.0e8f || fabsf(ay) > 1.0e8f || fabsf(az) > 1.0e8f) return 2;
    gx = static_cast<int>(floorf(ax));
    gy = static_cast<int>(floorf(ay));
    gz = static_cast<int>(floorf(az));
    return 0;
}
// Overflow bucket of a cell outside the dense box (the reference's unbounded
// hash grid, HashTable.h:13-16, has no box; these buckets give the same sets).
// Cells go in 4 x 4 x 4 blocks: the block is hashed and its 64 cells take 64
// consecutive buckets, so the 27 cells around a particle lie in a few runs of
// the bucket table (their ranges share cache lines, as in the dense box) and
// the particles of neighbouring cells sit in neighbouring slots.
#ifndef SPH_BUCKET_BLOCKS
#define SPH_BUCKET_BLOCKS 1
#endif
__device__ __forceinline__ uint32_t bucketOf(const SceneConst& S, int gx, int gy, int gz) {
#if SPH_BUCKET_BLOCKS
    // >> on a negative int floors (arithmetic shift) and & 3 is then the
    // matching non-negative remainder
    const int bx = gx >> 2, by = gy >> 2, bz = gz >> 2;
    const uint32_t local = static_cast<uint32_t>((gx & 3) | ((gy & 3) << 2) | ((gz & 3) << 4));
#else
    const int bx = gx, by = gy, bz = gz;
    const uint32_t local = 0u;
#endif
    uint32_t h = static_cast<uint32_t>(bx) * 0x9E3779B1u;
    h ^= static_cast<uint32_t>(by) * 0x85EBCA77u + (h << 6) + (h >> 2);
    h ^= static_cast<uint32_t>(bz) * 0xC2B2AE3Du + (h << 6) + (h >> 2);
    h ^= h >> 15;
    h *= 0x2C1B3C6Du;
    h ^= h >> 12;
#if SPH_BUCKET_BLOCKS
    return ((h << 6) | local) & S.ovMask;
#else
    return h & S.ovMask;
#endif
}
// Where the particles of global cell g lie: a dense cell (bucket = false) or
// the overflow bucket it hashes to, which other cells may share (scans then
// keep only candidates whose own cell is g).
__device__ __forceinline__ uint32_t cellSlot(const SceneConst& S, int gx, int gy, int gz, bool& bucket) {
    const int lx = gx - S.g0x, ly = gy - S.g0y, lz = gz - S.g0z;
    bucket = !(lx >= 0 && ly >= 0 && lz >= 0 && lx < S.gnx && ly < S.gny && lz < S.gnz);
    return bucket ? S.cellBase + S.nCells + bucketOf(S, gx, gy, gz) : S.cellBase + cellLinear(S, lx, ly, lz);
}
__device__ __forceinline__ bool inCell(const SceneConst& S, float4 p, int gx, int gy, int gz) {
    return static_cast<int>(floorf(__fmul_rn(p.x, S.invh))) == gx &&
           static_cast<int>(floorf(__fmul_rn(p.y, S.invh))) == gy &&
           static_cast<int>(floorf(__fmul_rn(p.z, S.invh))) == gz;
}

// correctPositions (PCISPHBoundaryObject.cpp:218-315) of one fluid particle
// after its position update: the boundary particles of the 27 cells around
// its new cell that lie within the rest radius of its new position x move it
// out along the weighted average of their normals, weights from the distance
// of `from` to them (x itself for WCSPH and DFSPH, the last predicted
// position for PCISPH). A cell the reference's hash table does not hold yet
// (no particle in it at any search so far) has no neighbour lists: no
// correction, as there; its first access makes it exist from the next search
// on. Returns the unit normal n, or false where the reference's normal is
// normalize(0) (no weight), which fails every later `length(n) > 0` test.
__device__ __forceinline__ bool pciCollide(const DevView& d, const SceneConst& S, float4& x, float4 from, float& nx,
                                           float& ny, float& nz) {
    nx = ny = nz = 0.0f;
    int cx, cy, cz;
    if (globalCell(S, x, cx, cy, cz)) return false;
    bool bucket;
    const uint32_t home = cellSlot(S, cx, cy, cz, bucket);
    if (!bucket) {
        // the lists in use are those of the last search, which (WCSPH, PCISPH:
        // this step's; DFSPH: the previous step's mid-step one) holds the cells
        // marked up to stepNo; a cell looked up now joins at the next search
        const uint32_t f = d.cellFirst[home];
        if (f > d.stepNo + 1u) atomicMin(d.cellFirst + home, d.stepNo + 1u);
        if (f > d.stepNo) return false;
    }
    float ws = 0.0f, w2 = 0.0f;
    for (int q = 0; q < 27; ++q) {
        bool nbucket;
        const uint32_t c = cellSlot(S, cx + q / 9 - 1, cy + (q / 3) % 3 - 1, cz + q % 3 - 1, nbucket);
        if (nbucket) continue;   // no boundary particle lies outside the dense box
        const uint2 br = d.bRange[c];
        for (uint32_t j = br.x; j < br.y; ++j) {
            const float4 xb = d.xb[j];
            if (!(lenRef(x.x - xb.x, x.y - xb.y, x.z - xb.z) < S.restR)) continue;
            const float dl = lenRef(from.x - xb.x, from.y - xb.y, from.z - xb.z);
            const float w = fmaxf(0.0f, (S.restR - dl) / S.restR);
            ws += w;
            w2 += w * (S.restR - dl);
            const float4 n = d.bn[j];
            nx += w * n.x; ny += w * n.y; nz += w * n.z;
        }
    }
    if (!(ws > 0.0f)) {
        nx = ny = nz = 0.0f;
        return false;
    }
    const float q = 1.0f / sqrtf(nx * nx + ny * ny + nz * nz);   // glm::normalize
    nx *= q; ny *= q; nz *= q;
    const float c = 1.0f / ws * w2;
    x.x += c * nx; x.y += c * ny; x.z += c * nz;
    return true;
}
// correctVelocities (PCISPHBoundaryObject.cpp:317-373): with length(n) > 0,
// the velocity `base` (v itself, or PCISPH's last predicted velocity) split
// along n gives v = tangentialFct vt - normalFct vn (n as the fluid's normal
// array holds it: not always a unit vector).
__device__ __forceinline__ void pciReflect(const SceneConst& S, float4& v, float4 base, float nx, float ny, float nz) {
    if (!(sqrtf(nx * nx + ny * ny + nz * nz) > 0.0f)) return;
    const float dn = base.x * nx + base.y * ny + base.z * nz;
    const float vnx = dn * nx, vny = dn * ny, vnz = dn * nz;
    v.x = S.tFct * (base.x - vnx) - S.nFct * vnx;
    v.y = S.tFct * (base.y - vny) - S.nFct * vny;
    v.z = S.tFct * (base.z - vnz) - S.nFct * vnz;
}
// WCSPH and PCISPH: both corrections after the integration.
__device__ __forceinline__ void pciCorrect(const DevView& d, const SceneConst& S, float4& x, float4& v, float4 from,
                                           float4 base) {
    float nx, ny, nz;
    if (pciCollide(d, S, x, from, nx, ny, nz)) pciReflect(S, v, base, nx, ny, nz);
}


// slot: which iteration flag of the scene state gates this launch (0/1 inside a
// Jacobi loop, by iteration parity; 2 outside a loop, always set).
#define SCENE_PROLOGUE(slot)                                  \
    const uint32_t lb = logicalBlock();                       \
    const int sIdx = d.blockScene[lb];                        \
    const SceneConst& S = d.scenes[sIdx];                     \
    const SceneState& T = d.state[sIdx];                      \
    const float dt = T.dt, dt2 = T.dt2;                       \
    const uint32_t i = lb * kBlock + threadIdx.x;             \
    const bool valid = T.active[slot] != 0 && i < S.fEnd;

// Scene prologue of a launch with L lanes per particle (grid = L x the
// particle blocks); `sub` is the lane's share of each neighbour list.
#define SCENE_PROLOGUE_LANES(slot, L)                                                              \
    const uint32_t lb = logicalBlock();                                                            \
    const int sIdx = d.blockScene[lb / (L)];                                                       \
    const SceneConst& S = d.scenes[sIdx];                                                          \
    const SceneState& T = d.state[sIdx];                                                           \
    const float dt = T.dt, dt2 = T.dt2;                                                            \
    const uint32_t sub = (threadIdx.x & 63u) / (64u / (L));                                        \
    const uint32_t i = lb * (kBlock / (L)) + (threadIdx.x >> 6) * (64u / (L)) + threadIdx.x % (64u / (L)); \
    const bool valid = T.active[slot] != 0 && i < S.fEnd;

// ---- FP64 deterministic reductions: block partials + per-scene kFinish ----
// Block partial of a reduction, stored at the block's logical index with a
// plain store. kFinish, launched next on the same stream, combines each
// scene's partials in a fixed order; stream order makes them visible, so no
// block issues a device-scope fence (on gfx950 one writes back and
// invalidates the whole L2 of the block's XCD, which costs the neighbour
// gathers of every other block resident there their cached lines).
template <bool MAX>
__device__ __forceinline__ void storePartial(const DevView& d, double v) {
    const double t = blockReduce<MAX>(v);
    if (threadIdx.x == 0) d.partials[logicalBlock()] = t;
}

// Jacobi-loop control of one scene after the error reduction of iteration
// `slot` (DFSPHSolver.cpp:431 / :746, PCISPHSolver.cpp:374): store the error,
// count the iteration and decide whether the next iteration runs.
__device__ __forceinline__ void loopEpilogue(SceneState& T, int slot, double err, double tol, int minIt, int maxIt) {
    int next = 0;
    if (T.active[slot]) {
        const int it = T.it + 1;
        T.err = err;
        T.it = it;
        next = ((err > tol) || it < minIt) && it < maxIt;
    }
    T.active[slot ^ 1] = next;
}
// Loop start (done by block 0 of the launch that precedes a loop; the other
// blocks of that launch read only dt, dt2 and active[2]).
__device__ __forceinline__ void loopStart(SceneState& T, double err, double tol, int minIt, int maxIt) {
    T.it = 0;
    T.err = err;
    T.active[0] = ((err > tol) || 0 < minIt) && 0 < maxIt;
    T.active[1] = 0;
}

// ---------------------------------------------------------------- grid
// Cell keys of every slot. dfsphAdvance: first x += v*dt and k_ws *= ts^2
// (DFSPHSolver::updatePositions :204-224 and :516-530; k_ws is cleared
// instead when the density loop did not run), and block 0 files the density
// loop's statistics.
__global__ void __launch_bounds__(kBlock) kKeys(DevView d, int dfsphAdvance, int clearCells, int counting,
                                                uint32_t* keys, uint32_t* vals) {
    const uint32_t lb = logicalBlock();
    const int sIdx = d.blockScene[lb];
    const SceneConst& S = d.scenes[sIdx];
    const SceneState& T = d.state[sIdx];
    const uint32_t i = lb * kBlock + threadIdx.x;
    if (clearCells) {   // the first slot of each cell of the last reorder clears its range
        const uint32_t oc = d.cellOf[i];
        if (oc != ~0u) d.fRange[oc] = make_uint2(0u, 0u);
    }
    if (dfsphAdvance && blockIdx.x == 0) {
        for (int s = threadIdx.x; s < d.numScenes; s += kBlock) {
            SceneState& U = d.state[s];
            const SceneConst& C = d.scenes[s];
            U.its = U.it;
            U.errD = U.err;
            U.capped = (U.it >= C.maxIt && U.err > static_cast<double>(C.eta)) ? 1 : 0;
        }
    }
    const bool live = i < S.fEnd;
    uint32_t key = S.deadKey, ord = S.ordBase + S.nCells + S.ovMask + 1;   // padding slots
    bool bucket = false;
    if (live) {
        float4 p = d.x[i];
        if (dfsphAdvance) {
            // w < 0 marks a particle emitted in this step (w is otherwise the
            // density of the last density pass, or 0): the reference emits
            // after updatePositions (DFSPHSolver.cpp:72), so it does not move yet
            if (!(p.w < 0.0f)) {
                const float dt = T.dt;
                const float4 vi = VEL(d, i);
                p.x += vi.x * dt;
                p.y += vi.y * dt;
                p.z += vi.z * dt;
                if (S.pciB) {
                    // correctPositions after updatePositions (DFSPHSolver.cpp:60-68);
                    // its normal is what the next step's correctVelocities reads
                    float nx, ny, nz;
                    pciCollide(d, S, p, p, nx, ny, nz);
                    d.nc[i] = make_float4(nx, ny, nz, 0.0f);
                }
            } else if (S.pciB) {
                d.nc[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);   // emitted after the correction
            }
            p.w = 0.0f;
            d.x[i] = p;
            d.kws[i] *= (T.it > 0 ? T.dt2 : 0.0f);
        }
        int gx, gy, gz;
        const int bad = globalCell(S, p, gx, gy, gz);
        if (bad) atomicOr(&d.flags[sIdx * kFlagSlots + (bad == 1 ? kFlagNonFinite : kFlagFar)], 1);
        const uint32_t c = cellSlot(S, gx, gy, gz, bucket);
        key = bucket ? S.deadKey - (S.ovMask + 1) + (c - S.cellBase - S.nCells)
                     : cellKey(S, gx - S.g0x, gy - S.g0y, gz - S.g0z);
        // the reference's hash table keeps every cell a particle was inserted in
        // (HashTable::clear empties cells but never erases them)
        if (S.pciB && !bucket && !bad && d.cellFirst[c] > d.markNo) atomicMin(d.cellFirst + c, d.markNo);
        if (counting) ord = S.ordBase + d.rankOf[c];
    }
    // particles in overflow buckets, one atomic per wave
    const unsigned long long outside = __ballot(bucket);
    if ((threadIdx.x & 63) == 0 && outside)
        atomicAdd(&d.flags[sIdx * kFlagSlots + kFlagOutOfGrid], static_cast<int>(__popcll(outside)));
    if (!counting) {
        keys[i] = key;
        vals[i] = i;
        return;
    }
    // Counting sort: slots come in the last search's order, so a wave holds
    // runs of equal order indices; the first lane of each run counts the run
    // with one atomic and hands its base to the others. The bases follow the
    // atomics' order; kOrderFix restores ascending slots within each index.
    // Padding slots are the scene's last fPadEnd - fEnd slots in both orders:
    // they keep their relative order (li = i - fEnd) and the first one counts
    // them all, so their order index needs no fix-up.
    keys[i] = ord;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t prev = __shfl_up(ord, 1u, 64);
    const bool head = live && (lane == 0 || prev != ord);
    const unsigned long long hm = __ballot(head);
    const unsigned long long upTo = hm & ((2ull << lane) - 1ull);   // heads at or below this lane
    const uint32_t h = upTo ? 63u - static_cast<uint32_t>(__clzll(upTo)) : lane;
    uint32_t base = 0;
    if (head) {
        // the run ends at the next head or at the first padding lane
        const unsigned long long after = lane == 63u ? 0ull : (hm >> (lane + 1u));
        uint32_t next = after ? lane + 1u + static_cast<uint32_t>(__ffsll(static_cast<long long>(after)) - 1) : 64u;
        const uint32_t liveEnd = S.fEnd > i ? (S.fEnd - i < 64u - lane ? lane + (S.fEnd - i) : 64u) : lane;
        next = min(next, liveEnd);
        base = atomicAdd(&d.counts[ord], next - lane);
    }
    base = __shfl(base, static_cast<int>(h), 64);
    if (live) {
        d.li[i] = base + (lane - h);
    } else {
        d.li[i] = i - S.fEnd;
        if (i == S.fEnd) atomicAdd(&d.counts[ord], S.fPadEnd - S.fEnd);
    }
}
__global__ void kFillRanges(uint2* r, size_t n) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) r[i] = make_uint2(0u, 0u);
}

// Counting sort, step 2: slot i's place in the sorted order.
__global__ void __launch_bounds__(kBlock) kScatter(DevView d, const uint32_t* keys, uint32_t* perm) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;   // over all NP slots
    perm[d.starts[keys[i]] + d.li[i]] = i;
}

// Counting sort, step 3, one thread per order index: ascending slots within
// the index (the order a stable radix sort of the same keys gives), the cell
// range, and a zero count for the next search. occ[o] records whether index
// o held slots at this fix-up; unless `full` (the ranges were written by
// anything else since), an index empty now and then has a zero count and a
// cleared range already and is left alone: most of the dense box is empty.
constexpr uint32_t kFixInRegs = 32, kFixMax = 256;
__global__ void __launch_bounds__(kBlock) kOrderFix(DevView d, uint32_t* perm, uint32_t ordTotal, uint8_t* occ,
                                                    int full) {
    const uint32_t o = blockIdx.x * kBlock + threadIdx.x;
    if (o >= ordTotal) return;
    const uint32_t b = d.starts[o], e = d.starts[o + 1], n = e - b;
    const uint8_t was = occ[o];
    if (!full && n == 0 && !was) return;
    if (n) d.counts[o] = 0u;   // counts[o] == n (starts is its exclusive scan)
    if (full || was != (n != 0)) occ[o] = n != 0 ? 1 : 0;
    const uint32_t cell = d.ordToCell[o];   // ~0: the scene's padding slots, already in order
    if (n > 1 && cell != ~0u) {
        if (n <= kFixInRegs) {
            uint32_t q[kFixInRegs];
            for (uint32_t k = 0; k < n; ++k) q[k] = perm[b + k];
            for (uint32_t k = 1; k < n; ++k) {
                const uint32_t v = q[k];
                uint32_t j = k;
                while (j > 0 && q[j - 1] > v) {
                    q[j] = q[j - 1];
                    --j;
                }
                q[j] = v;
            }
            for (uint32_t k = 0; k < n; ++k) perm[b + k] = q[k];
        } else if (n <= kFixMax) {
            for (uint32_t k = b + 1; k < e; ++k) {
                const uint32_t v = perm[k];
                uint32_t j = k;
                while (j > b && perm[j - 1] > v) {
                    perm[j] = perm[j - 1];
                    --j;
                }
                perm[j] = v;
            }
        } else {
            // a crowded cell: leave it to the radix sort (the engine redoes the search)
            const uint32_t s = b < d.NP ? static_cast<uint32_t>(d.blockScene[b / kBlock]) : 0u;
            atomicOr(&d.flags[s * kFlagSlots + kFlagBigCell], 1);
        }
    }
    if (cell != ~0u) d.fRange[cell] = n ? make_uint2(b, e) : make_uint2(0u, 0u);
}

// Gather the persistent state into sorted order and record cell ranges.
__global__ void __launch_bounds__(kBlock) kReorder(DevView d, const uint32_t* keysSorted, const uint32_t* valsSorted,
                                                   float4* xAlt, float4* vAlt, float* kwsAlt, float* kvwsAlt,
                                                   uint32_t* idAlt, int ranges) {
    const uint32_t lb = logicalBlock();
    const int sIdx = d.blockScene[lb];
    const SceneConst& S = d.scenes[sIdx];
    const uint32_t i = lb * kBlock + threadIdx.x;
    const uint32_t o = valsSorted[i];
    const float4 p = d.x[o];
    xAlt[i] = p;
    vAlt[i] = VEL(d, o);
    const float kv = d.kvws[o];
    kwsAlt[i] = d.kws[o];
    kvwsAlt[i] = kv;
    idAlt[i] = d.id[o];
    if (d.nc) d.ncAlt[i] = d.nc[o];
    if (d.snapV) {
        d.snapV[i] = VEL(d, o);
        d.snapKvws[i] = kv;
    }
    if (!ranges) return;
    uint32_t first = ~0u;
    if (i < S.fEnd) {
        const uint32_t key = keysSorted[i];
        int gx, gy, gz;
        globalCell(S, p, gx, gy, gz);
        bool bucket;
        const uint32_t c = cellSlot(S, gx, gy, gz, bucket);
        if (i == S.fBegin || keysSorted[i - 1] != key) {
            d.fRange[c].x = i;
            first = c;
        }
        if (i + 1 == S.fEnd || keysSorted[i + 1] != key) d.fRange[c].y = i + 1;
    }
    d.cellOf[i] = first;
}

// Candidates each lane of the search loads per round (1 = one at a time).
#ifndef SPH_BUILD_UNROLL
#define SPH_BUILD_UNROLL 4
#endif
constexpr uint32_t kBuildUnroll = SPH_BUILD_UNROLL;

// Within-support neighbour lists from the 27-cell candidates (HashTable.cpp:23-86
// walk order: dx, dy, dz with dz fastest), filtered by |r| <= h with the
// reference's rounding (DFSPHSolver.cpp:575). Self is included. L lanes share
// a particle: each round they test L consecutive candidates of one cell and a
// wave ballot appends the hits in walk order, so the lists are the same for
// every L.
template <uint32_t L>
__device__ __forceinline__ void buildNeighbors(const DevView& d) {
    SCENE_PROLOGUE_LANES(2, L)
    (void)dt;
    (void)dt2;
    constexpr uint32_t P = 64u / L;         // particles per wave
    constexpr uint32_t PB = kBlock / L;     // particles per block
    const uint32_t p = threadIdx.x % P;
    const uint32_t tp = (threadIdx.x >> 6) * P + p;
#if SPH_LIST_STAGE
    // a particle finds its s-th neighbour at its own pace, so direct stores of
    // one wave land in many different rows; staged rows leave as coalesced lines
    __shared__ uint32_t sList[SPH_LIST_STAGE * PB];
    __shared__ uint32_t sCnt[PB];
#endif
    uint32_t nf = 0, nb = 0;
    // lanes p, p + P, ... of the wave hold one particle; `before` counts the
    // hits of its lanes below this one, `total` all of them
    auto place = [&](bool hit, uint32_t& before, uint32_t& total) {
        if constexpr (L == 1) {
            before = 0;
            total = hit ? 1u : 0u;
        } else {
            const unsigned long long m = __ballot(hit);
            before = 0;
            total = 0;
#pragma unroll
            for (uint32_t k = 0; k < L; ++k) {
                const uint32_t bit = static_cast<uint32_t>((m >> (p + k * P)) & 1ull);
                before += k < sub ? bit : 0u;
                total += bit;
            }
        }
    };
    if (valid) {
        // the scene's constants as values: the list stores below could alias
        // the scene table, which would make every candidate reload them
        const SceneConst Sv = S;
        const float4 xi = d.x[i];
        int cx, cy, cz;
        globalCell(Sv, xi, cx, cy, cz);
        const size_t NP = d.NP;
        // walk cell q = 9 (dx + 1) + 3 (dy + 1) + (dz + 1); the ranges of cell
        // q + 1 are loaded before cell q is scanned, and each lane loads U
        // candidates per round before testing them in order, so a list costs
        // a few rounds of load latency instead of one per candidate
        constexpr uint32_t U = kBuildUnroll;
        auto cellAt = [&](int q, int& ax, int& ay, int& az, bool& bucket) {
            ax = cx + q / 9 - 1;
            ay = cy + (q / 3) % 3 - 1;
            az = cz + q % 3 - 1;
            return cellSlot(Sv, ax, ay, az, bucket);
        };
        int axN, ayN, azN;
        bool bucketN;
        uint32_t cN = cellAt(0, axN, ayN, azN, bucketN);
        uint2 frN = d.fRange[cN];
        uint2 brN = Sv.akinci && !bucketN ? d.bRange[cN] : make_uint2(0u, 0u);
        for (int q = 0; q < 27; ++q) {
            const int ax = axN, ay = ayN, az = azN;
            const bool bucket = bucketN;
            const uint2 fr = frN, br = brN;
            if (q + 1 < 27) {
                cN = cellAt(q + 1, axN, ayN, azN, bucketN);
                frN = d.fRange[cN];
                // boundary particles all lie in the dense box
                brN = Sv.akinci && !bucketN ? d.bRange[cN] : make_uint2(0u, 0u);
            }
            for (uint32_t base = fr.x; base < fr.y; base += U * L) {
                float4 xj[U];
#pragma unroll
                for (uint32_t k = 0; k < U; ++k) xj[k] = d.x[min(base + k * L + sub, fr.y - 1u)];
#pragma unroll
                for (uint32_t k = 0; k < U; ++k) {
                    const uint32_t j = base + k * L + sub;
                    const bool hit = j < fr.y && (!bucket || inCell(Sv, xj[k], ax, ay, az)) &&
                                     dRef(xi.x - xj[k].x, xi.y - xj[k].y, xi.z - xj[k].z) <= Sv.dmax;
                    uint32_t before, total;
                    place(hit, before, total);
                    if (hit) {
                        const uint32_t pos = nf + before;
#if SPH_LIST_STAGE
                        if (pos < SPH_LIST_STAGE) sList[pos * PB + tp] = j;
                        else
#endif
                        if (pos < d.capF) d.nbrF[pos * NP + i] = j;
                    }
                    nf += total;
                }
            }
            for (uint32_t base = br.x; base < br.y; base += U * L) {
                float4 xb[U];
#pragma unroll
                for (uint32_t k = 0; k < U; ++k) xb[k] = d.xb[min(base + k * L + sub, br.y - 1u)];
#pragma unroll
                for (uint32_t k = 0; k < U; ++k) {
                    const uint32_t b = base + k * L + sub;
                    const bool hit = b < br.y && dRef(xi.x - xb[k].x, xi.y - xb[k].y, xi.z - xb[k].z) <= Sv.dmax;
                    uint32_t before, total;
                    place(hit, before, total);
                    if (hit && nb + before < d.capB) d.nbrB[(nb + before) * NP + i] = b;
                    nb += total;
                }
            }
        }
        if (sub == 0) {
            d.cntF[i] = nf;
            d.cntB[i] = nb;
            // rows [nf, nf rounded up to kPadRows) hold the particle itself
            // (the staged rows get theirs below); capF is a multiple of kPadRows
            const uint32_t padEnd = min((nf + kPadRows - 1u) / kPadRows * kPadRows, d.capF);
#if SPH_LIST_STAGE
            for (uint32_t r = max(nf, static_cast<uint32_t>(SPH_LIST_STAGE)); r < padEnd; ++r) d.nbrF[r * NP + i] = i;
#else
            for (uint32_t r = nf; r < padEnd; ++r) d.nbrF[r * NP + i] = i;
#endif
        }
    }
#if SPH_LIST_STAGE
    if (sub == 0) sCnt[tp] = nf;
#endif
    // block-wide max, one atomic per block
    int mf = static_cast<int>(nf), mb = static_cast<int>(nb);
    for (int off = 32; off > 0; off >>= 1) {
        mf = max(mf, __shfl_down(mf, off, 64));
        mb = max(mb, __shfl_down(mb, off, 64));
    }
    __shared__ int smf[kBlock / 64], smb[kBlock / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) {
        smf[wid] = mf;
        smb[wid] = mb;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kBlock / 64; ++w) {
            mf = max(mf, smf[w]);
            mb = max(mb, smb[w]);
        }
        atomicMax(&d.flags[sIdx * kFlagSlots + kFlagMaxFluidNbr], mf);
        atomicMax(&d.flags[sIdx * kFlagSlots + kFlagMaxBoundaryNbr], mb);
    }
#if SPH_LIST_STAGE
    // whole rows up to the block's longest list rounded up to kPadRows; rows
    // of a particle from its count to the next multiple of kPadRows hold the
    // particle itself (a zero term in the Jacobi passes), later ones ~0 (never
    // read: every pass stops at min(count, capacity) rounded up)
    int rows = smf[0];
    for (int w = 1; w < kBlock / 64; ++w) rows = max(rows, smf[w]);
    rows = (rows + static_cast<int>(kPadRows) - 1) / static_cast<int>(kPadRows) * static_cast<int>(kPadRows);
    rows = min(rows, min(static_cast<int>(d.capF), SPH_LIST_STAGE));
    const size_t NPs = d.NP;
    const uint32_t first = lb * PB;   // the block's particles: [first, first + PB) < NP
    for (uint32_t e = threadIdx.x; e < static_cast<uint32_t>(rows) * PB; e += kBlock) {
        const uint32_t row = e / PB, t = e % PB;
        const uint32_t c = sCnt[t];
        const uint32_t pad = (c + kPadRows - 1u) / kPadRows * kPadRows;
        d.nbrF[row * NPs + first + t] = row < c ? sList[row * PB + t] : (row < pad ? first + t : ~0u);
    }
#endif
}
// Lanes per particle of the search: 1, 2, 4 or 8, the same lists for each.
// Where a launch fills the chip one lane is fastest (the 1M-particle dam
// break: 1 lane 293 us per build, 2 lanes 336, 4 lanes 479; cells hold ~8
// candidates, so rounds of 2-4 add ballots without gathering any better: one
// lane already walks each cell's contiguous slots). A scene of a few thousand
// particles puts one lane per particle on a few dozen CUs, where more lanes
// finish sooner. The engine picks the count per context by timing each on
// its first searches (Engine::buildLists).
template <uint32_t L>
__global__ void __launch_bounds__(kBlock) kBuildNeighbors(DevView d) { buildNeighbors<L>(d); }

// ---------------------------------------------------------------- DFSPH tiles
// The Jacobi passes of one launch block (kTileGroup = 64 particles, 4 lanes
// each) read the records of a few hundred distinct neighbours, each of them
// several times, gathered one list entry at a time from L2/HBM. A tile is the
// set of distinct fluid slots the group's lists name: the passes stage their
// records in LDS with one wide cooperative load, then walk 16-bit indices
// into the tile. Tiles are rebuilt with the lists (kTileLists, one block per
// group): a block-wide open-addressing hash set in LDS collects the slots,
// a scan over the table numbers them, and the lists are rewritten as tile
// indices. A group with more than kTileMax distinct neighbours is marked
// (count ~0) and its passes gather from HBM as before; the row order of both
// paths is the same, so their sums are the same. Needs 4 lanes per particle
// in all three Jacobi passes (SPH_LPP_VEL=4). Measured on the 1M-particle dam
// break (ms per step; per-launch us of kRhoStar / kDivergence / velocity):
// tiles 1.83 (66.8 / 75.8 / 41.5, and 109 us per tile build, which also
// slows the once-per-step passes that follow it), no tiles with 4 velocity
// lanes 1.74 (71.1 / 80.5 / 41.6), no tiles with 2 velocity lanes 1.71
// (71.1 / 80.4 / 38.4): off by default.
#ifndef SPH_TILES
#define SPH_TILES 0
#endif
extern "C" __host__ int sphKernelsTiles() { return SPH_TILES; }
constexpr uint32_t kTileHash = 2 * kTileMax;   // hash slots (load factor <= 1/2)
static_assert((kTileHash & (kTileHash - 1)) == 0, "hash size is a power of two");
static_assert(!SPH_TILES || (kTileGroup * SPH_LPP == kBlock && SPH_LPP == 4 && SPH_LPP_VEL == 4),
              "a tile group is one launch block of the 4-lanes-per-particle passes");
__device__ __forceinline__ uint32_t tileHash(uint32_t j) { return (j * 0x9E3779B1u) >> 22; }   // 10 bits
static_assert(kTileHash == 1024, "tileHash yields 10 bits");

__global__ void __launch_bounds__(kBlock) kTileLists(DevView d) {
    __shared__ uint32_t key[kTileHash];
    __shared__ uint16_t tix[kTileHash];
    __shared__ uint32_t nUnique, overflow;
    __shared__ uint32_t wsum[kBlock / 64];
    const uint32_t g = blockIdx.x;
    const uint32_t first = g * kTileGroup;
    const SceneConst& S = d.scenes[d.blockScene[first / kBlock]];
    const size_t NP = d.NP;
    for (uint32_t h = threadIdx.x; h < kTileHash; h += kBlock) key[h] = ~0u;
    if (threadIdx.x == 0) {
        nUnique = 0;
        overflow = 0;
    }
    __syncthreads();
    // four threads per particle, rows q, q + 4, ... (self-padding rows name
    // the particle itself, which its own list holds already)
    const uint32_t p = threadIdx.x % kTileGroup, q = threadIdx.x / kTileGroup;
    const uint32_t i = first + p;
    const uint32_t n = i < S.fEnd ? min(d.cntF[i], d.capF) : 0u;
    const uint32_t nPad = (n + kPadRows - 1u) / kPadRows * kPadRows;
    for (uint32_t r = q; r < n; r += 4) {
        const uint32_t j = d.nbrF[r * NP + i];
        uint32_t h = tileHash(j);
        for (uint32_t probe = 0;; ++probe) {
            if (probe == kTileHash) {
                overflow = 1;
                break;
            }
            const uint32_t old = atomicCAS(&key[h], ~0u, j);
            if (old == ~0u) {
                atomicAdd(&nUnique, 1u);
                break;
            }
            if (old == j) break;
            h = (h + 1u) & (kTileHash - 1u);
        }
    }
    __syncthreads();
    const uint32_t T = nUnique;
    if (overflow || T > kTileMax) {
        if (threadIdx.x == 0) d.tileCount[g] = ~0u;
        return;
    }
    // number the occupied slots in table order: thread t owns slots 8t .. 8t + 7
    constexpr uint32_t kPer = kTileHash / kBlock;
    uint32_t own = 0;
#pragma unroll
    for (uint32_t k = 0; k < kPer; ++k) own += key[threadIdx.x * kPer + k] != ~0u ? 1u : 0u;
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    uint32_t incl = own;
#pragma unroll
    for (uint32_t off = 1; off < 64; off <<= 1) {
        const uint32_t v = __shfl_up(incl, off, 64);
        if (lane >= off) incl += v;
    }
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    uint32_t t = incl - own;
    for (uint32_t k = 0; k < w; ++k) t += wsum[k];
    uint32_t* slots = d.tileSlot + static_cast<size_t>(g) * kTileMax;
#pragma unroll
    for (uint32_t k = 0; k < kPer; ++k) {
        const uint32_t h = threadIdx.x * kPer + k;
        const uint32_t j = key[h];
        if (j != ~0u) {
            tix[h] = static_cast<uint16_t>(t);
            slots[t] = j;
            ++t;
        }
    }
    __syncthreads();
    auto lookup = [&](uint32_t j) {
        uint32_t h = tileHash(j);
        while (key[h] != j) h = (h + 1u) & (kTileHash - 1u);   // every j of the group is in the table
        return static_cast<uint32_t>(tix[h]);
    };
    // word w of a particle: rows 8 (w / 4) + w % 4 and that row + 4
    for (uint32_t wd = q; wd < nPad / 2; wd += 4) {
        const uint32_t r0 = (wd >> 2) * kPadRows + (wd & 3u);
        const uint32_t t0 = lookup(d.nbrF[r0 * NP + i]);
        const uint32_t t1 = lookup(d.nbrF[(r0 + 4u) * NP + i]);
        d.nbrT[wd * NP + i] = t0 | (t1 << 16);
    }
    if (threadIdx.x == 0) d.tileCount[g] = T;
}

// AkinciBoundaryObject::computeVolume (AkinciBoundaryObject.cpp:10-34):
// V_b = 1 / sum of W over the boundary candidates of the 27 cells.
__global__ void kBoundaryVolume(DevView d, int scene, uint32_t bBegin, uint32_t bEnd) {
    const SceneConst& S = d.scenes[scene];
    const uint32_t b = bBegin + blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= bEnd) return;
    const float4 xb = d.xb[b];
    int cx, cy, cz;
    localCell(S, xb, cx, cy, cz);
    float delta = 0.0f;
    for (int dx = -1; dx <= 1; ++dx)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dz = -1; dz <= 1; ++dz) {
                const int ax = cx + dx, ay = cy + dy, az = cz + dz;
                if (ax < 0 || ay < 0 || az < 0 || ax >= S.gnx || ay >= S.gny || az >= S.gnz) continue;
                const uint2 br = d.bRange[S.cellBase + cellLinear(S, ax, ay, az)];
                for (uint32_t c = br.x; c < br.y; ++c) {
                    const float4 xc = d.xb[c];
                    delta += cubicW(S, lenRef(xb.x - xc.x, xb.y - xc.y, xb.z - xc.z));
                }
            }
    const float vol = 1.0f / delta;
    d.vb[b] = vol;
    d.xb[b].w = S.rho0 * vol;
}

// ---------------------------------------------------------------- PCISPH boundary objects
// Poly6::W (Poly6.h:31-43): |r| as glm::length rounds it, then squared.
__device__ __forceinline__ float poly6W(const SceneConst& S, float rl) {
    const float r2 = rl * rl;
    if (!(r2 <= S.h2)) return 0.0f;
    const float diff = S.h2 - r2;
    return S.poly6C * diff * diff * diff;
}
// Spiky::gradW (Spiky.h:44-56) as gradW = r * factor.
__device__ __forceinline__ float spikyFactor(const SceneConst& S, float rl) {
    if (!(rl != 0.0f && rl <= S.h)) return 0.0f;
    const float diff = S.h - rl;
    return -(S.spikyC * diff * diff / rl);
}

// One thread per boundary particle of one scene: the fluid particles of its
// 27 cells (forall_fluid_neighbors), then the boundary particles of the same
// cells (forall_boundary_neighbors), SPHSolver.cpp:91-123 (current
// positions, cubic spline; Tait pressure of WCSPHSolver.cpp:97-115) or
// PCISPHSolver.cpp:211-244 (predicted positions, Poly6; pressure update of
// :284-297).
__global__ void kBoundaryDensity(DevView d, int scene, uint32_t bBegin, uint32_t bEnd, int predicted, int slot) {
    const SceneConst& S = d.scenes[scene];
    if (!d.state[scene].active[slot]) return;
    const uint32_t b = bBegin + blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= bEnd) return;
    const float4 xb = d.xb[b];
    int cx, cy, cz;
    globalCell(S, xb, cx, cy, cz);
    float rho = 0.0f;
    for (int q = 0; q < 27; ++q) {
        bool bucket;
        const uint32_t c = cellSlot(S, cx + q / 9 - 1, cy + (q / 3) % 3 - 1, cz + q % 3 - 1, bucket);
        if (bucket) continue;   // boundary cells lie 4 or more cells inside the dense box
        const uint2 fr = d.fRange[c];
        for (uint32_t j = fr.x; j < fr.y; ++j) {
            const float4 xj = predicted ? d.xs[j] : d.x[j];
            const float rl = lenRef(xb.x - xj.x, xb.y - xj.y, xb.z - xj.z);
            rho += S.m * (predicted ? poly6W(S, rl) : cubicW(S, rl));
        }
    }
    for (int q = 0; q < 27; ++q) {
        bool bucket;
        const uint32_t c = cellSlot(S, cx + q / 9 - 1, cy + (q / 3) % 3 - 1, cz + q % 3 - 1, bucket);
        if (bucket) continue;
        const uint2 br = d.bRange[c];
        for (uint32_t j = br.x; j < br.y; ++j) {
            const float4 xc = d.xb[j];
            const float rl = lenRef(xb.x - xc.x, xb.y - xc.y, xb.z - xc.z);
            rho += xc.w * (predicted ? poly6W(S, rl) : cubicW(S, rl));
        }
    }
    d.bRho[b] = rho;
    if (predicted) d.bP[b] += S.pciDelta * fmaxf(0.0f, rho - S.rho0);
    else d.bP[b] = fmaxf(0.0f, S.wcB * (powf(rho / S.rho0, S.gamma) - 1.0f));
}

// correctPredPositions (PCISPHBoundaryObject.cpp:375-428) of slot i: boundary
// particles within the rest radius of its current position x (all of them
// are in its boundary list, which holds those within h) push its predicted
// position p out along the weighted average of their normals.
__device__ __forceinline__ void pciCorrectPredicted(const DevView& d, const SceneConst& S, uint32_t i, float4 x,
                                                    float& px, float& py, float& pz) {
    const uint32_t nb = min(d.cntB[i], d.capB);
    float ws = 0.0f, w2 = 0.0f, nx = 0.0f, ny = 0.0f, nz = 0.0f;
    for (uint32_t s = 0; s < nb; ++s) {
        const uint32_t bi = ldList(d.nbrB + static_cast<size_t>(s) * d.NP + i);
        const float4 xb = d.xb[bi];
        if (!(lenRef(x.x - xb.x, x.y - xb.y, x.z - xb.z) < S.restR)) continue;
        const float dl = lenRef(px - xb.x, py - xb.y, pz - xb.z);
        const float w = fmaxf(0.0f, (S.restR - dl) / S.restR);
        ws += w;
        w2 += w * (S.restR - dl);
        const float4 n = d.bn[bi];
        nx += w * n.x; ny += w * n.y; nz += w * n.z;
    }
    if (ws > 0.0f) {
        const float q = 1.0f / sqrtf(nx * nx + ny * ny + nz * nz);   // glm::normalize
        const float c = 1.0f / ws * w2;
        px += c * (nx * q); py += c * (ny * q); pz += c * (nz * q);
    }
}

// ---------------------------------------------------------------- density
// SPHSolver::computeFluidDensities (SPHSolver.cpp:42-89) fused with
// DFSPHSolver::computeAlpha (DFSPHSolver.cpp:109-178).
// The w of particle i's position record carries a per-particle scalar that
// the next neighbour pass gathers with the position (one record per neighbour
// instead of two or three): the density passes file rho_i there for the force
// pass, and within the PCISPH loop the predicted-density pass files
// p_i / rho_i^2 for the pressure-acceleration pass (the force pass has read
// rho by then, and the next step's density pass files it again).
__device__ __forceinline__ void storeXW(const DevView& d, uint32_t i, float w) {
    reinterpret_cast<float*>(d.x + i)[3] = w;
}
__device__ __forceinline__ void storeRhoInX(const DevView& d, uint32_t i, float rho) { storeXW(d, i, rho); }
__global__ void __launch_bounds__(kBlock) kDensityAlpha(DevView d, int withAlpha) {
    SCENE_PROLOGUE_LANES(2, kLppOnce)
    (void)dt;
    (void)dt2;
    const size_t NP = d.NP;
    float rho = 0.0f, sgx = 0.0f, sgy = 0.0f, sgz = 0.0f, ssq = 0.0f;
    if (valid) {
        const float4 xi = d.x[i];
        const uint32_t nf = min(d.cntF[i], d.capF);
        walkList<kLppOnce, kBatch, true>(d.nbrF, NP, i, sub, nf, [&](uint32_t j) { return d.x[j]; },
                                   [&](uint32_t, const float4& xj) {
                                       const float rx = xi.x - xj.x, ry = xi.y - xj.y, rz = xi.z - xj.z;
                                       const float rl = fastSqrt(rx * rx + ry * ry + rz * rz);
                                       rho += S.m * cubicW(S, rl);
                                       const float g = S.m * gradFactor(S, rl);
                                       const float gx = g * rx, gy = g * ry, gz = g * rz;
                                       sgx += gx; sgy += gy; sgz += gz;
                                       ssq += gx * gx + gy * gy + gz * gz;
                                   });
        const uint32_t nb = S.akinci ? min(d.cntB[i], d.capB) : 0u;
        for (uint32_t s = sub; s < nb; s += kLppOnce) {
            const float4 xb = d.xb[ldList(d.nbrB + s * NP + i)];
            const float rx = xi.x - xb.x, ry = xi.y - xb.y, rz = xi.z - xb.z;
            const float rl = fastSqrt(rx * rx + ry * ry + rz * rz);
            rho += xb.w * cubicW(S, rl);
            const float g = xb.w * gradFactor(S, rl);
            sgx += g * rx; sgy += g * ry; sgz += g * rz;
        }
    }
    rho = lppSum<kLppOnce>(rho);
    sgx = lppSum<kLppOnce>(sgx);
    sgy = lppSum<kLppOnce>(sgy);
    sgz = lppSum<kLppOnce>(sgz);
    ssq = lppSum<kLppOnce>(ssq);
    if (!valid || sub != 0) return;
    d.rho[i] = rho;
    storeRhoInX(d, i, rho);
    if (withAlpha) {
        const float den = sgx * sgx + sgy * sgy + sgz * sgz + ssq;
        d.alpha[i] = den > S.eps ? rho / den : 0.0f;
    }
}

// WCSPH density + Tait pressure (WCSPHSolver.cpp:80-95), with the lanes per
// particle of the other once-per-step passes (kLppOnce).
__global__ void __launch_bounds__(kBlock) kDensityPressure(DevView d) {
    if (d.flagMode && blockIdx.x == 0)
        for (uint32_t k = threadIdx.x; k < d.flagCount; k += kBlock) {
            const int32_t f = d.flags[k];
            d.flagLog[k] = d.flagMode == 2 ? max(d.flagLog[k], f) : f;
            d.flags[k] = 0;
        }
    SCENE_PROLOGUE_LANES(2, kLppOnce)
    (void)dt;
    (void)dt2;
    const size_t NP = d.NP;
    float rho = 0.0f;
    if (valid) {
        const float4 xi = d.x[i];
        const uint32_t nf = min(d.cntF[i], d.capF);
        walkList<kLppOnce, kBatch, true>(d.nbrF, NP, i, sub, nf, [&](uint32_t j) { return d.x[j]; },
                                         [&](uint32_t, const float4& xj) {
                                             const float rx = xi.x - xj.x, ry = xi.y - xj.y, rz = xi.z - xj.z;
                                             rho += S.m * cubicW(S, fastSqrt(rx * rx + ry * ry + rz * rz));
                                         });
        const uint32_t nb = S.akinci ? min(d.cntB[i], d.capB) : 0u;
        for (uint32_t s = sub; s < nb; s += kLppOnce) {
            const float4 xb = d.xb[ldList(d.nbrB + s * NP + i)];
            const float rx = xi.x - xb.x, ry = xi.y - xb.y, rz = xi.z - xb.z;
            rho += xb.w * cubicW(S, fastSqrt(rx * rx + ry * ry + rz * rz));
        }
    }
    rho = lppSum<kLppOnce>(rho);
    if (!valid || sub != 0) return;
    d.rho[i] = rho;
    storeRhoInX(d, i, rho);
    d.p[i] = fmaxf(0.0f, S.wcB * (powf(rho / S.rho0, S.gamma) - 1.0f));
}

// ---------------------------------------------------------------- forces
// One neighbour pass computing, in the reference's accumulation order:
//   WCSPH pressure force (WCSPHSolver.cpp:132-188) when `pressure`,
//   viscosity: standard (StandardViscosity.cpp:5-49), artificial
//     (ArtificialViscosity.cpp:6-53) or XSPH (XSPHViscosity.cpp:6-42),
//   surface-tension normals (SurfaceTensionAkinci.cpp:13-33).
// The acceleration is written fresh (it is zero at this point of every step).
__global__ void __launch_bounds__(kBlock) kForces(DevView d, int pressure) {
    SCENE_PROLOGUE_LANES(2, kLppOnce)
    (void)dt2;
    const size_t NP = d.NP;
    // lanes of one particle take list slots sub, sub + kLppOnce, ...; invalid
    // lanes run empty loops so that every lane reaches the cross-lane sums
    const uint32_t ii = valid ? i : 0u;
    const float4 xi = d.x[ii];
    const float4 vi = VEL(d, ii);
    const float rhoi = d.rho[ii];
    const uint32_t nf = valid ? min(d.cntF[i], d.capF) : 0u;
    const uint32_t nb = (valid && S.akinci) ? min(d.cntB[i], d.capB) : 0u;
    float ax = 0.0f, ay = 0.0f, az = 0.0f;   // pressure sum
    float fx = 0.0f, fy = 0.0f, fz = 0.0f;   // viscosity sum
    float nx = 0.0f, ny = 0.0f, nz = 0.0f;   // normal sum
    const int vm = S.viscMethod;
    const bool normals = S.stOn;
    const bool other = vm >= 0 || normals;
    const float invdt = 1.0f / dt;
    const float pi_ = pressure && valid ? d.p[i] : 0.0f;
    const float dpi = pi_ / (rhoi * rhoi);
    const float dpb = pi_ / (S.rho0 * S.rho0);
    // one walk for the pressure, viscosity and normal sums, each in list order;
    // w of the position record is the density pass's rho_j
    if (pressure || other) {
        auto gather = [&](uint32_t j) {
            return XVP{POSV(d, j), vm >= 0 ? VEL(d, j) : make_float4(0.0f, 0.0f, 0.0f, 0.0f),
                       pressure ? d.p[j] : 0.0f};
        };
        walkList<kLppOnce, kBatch, true>(d.nbrF, NP, i, sub, nf, gather, [&](uint32_t, const XVP& q) {
            const float4 xj = q.x;
            const float rhoj = xj.w;
            const float rx = xi.x - xj.x, ry = xi.y - xj.y, rz = xi.z - xj.z;
            if (pressure) {
                const float g = S.m * (dpi + q.p / (rhoj * rhoj)) * gradFactor(S, fastSqrt(rx * rx + ry * ry + rz * rz));
                ax -= g * rx; ay -= g * ry; az -= g * rz;
            }
            if (!other) return;
            const float r2 = rx * rx + ry * ry + rz * rz;
            const float rl = fastSqrt(r2);
            if (vm >= 0) {
                const float4 vj = q.v;
                if (vm == 2) {  // XSPH: nu (m/rho_j) (v_j - v_i) W / dt
                    const float c = S.visc * (S.m / rhoj) * cubicW(S, rl) * invdt;
                    fx += c * (vj.x - vi.x); fy += c * (vj.y - vi.y); fz += c * (vj.z - vi.z);
                } else if (vm == 1) {  // artificial: nu (m/rho_j) (v_ij . r_ij)/(r^2 + 0.01 h^2) gradW
                    const float dp = (vi.x - vj.x) * rx + (vi.y - vj.y) * ry + (vi.z - vj.z) * rz;
                    const float c = S.visc * (S.m / rhoj) * dp / (r2 + S.h2c) * gradFactor(S, rl);
                    fx += c * rx; fy += c * ry; fz += c * rz;
                } else {  // standard: m (v_j - v_i)/rho_j laplW, scaled by nu/rho_i below
                    const float c = S.m / rhoj * laplW(S, rl);
                    fx += c * (vj.x - vi.x); fy += c * (vj.y - vi.y); fz += c * (vj.z - vi.z);
                }
            }
            if (normals) {
                const float c = (S.m / rhoj) * gradFactor(S, rl);
                nx += c * rx; ny += c * ry; nz += c * rz;
            }
        });
    }
    if (pressure) {
        for (uint32_t s = sub; s < nb; s += kLppOnce) {
            const uint32_t bi = ldList(d.nbrB + s * NP + i);
            const float4 xb = d.xb[bi];
            const float rx = xi.x - xb.x, ry = xi.y - xb.y, rz = xi.z - xb.z;
            // Akinci: rho0 V_b p_i / rho0^2 (WCSPHSolver.cpp:176-186); a PCISPH
            // boundary particle: m_b (p_i / rho_i^2 + p_b / rho_b^2) (:160-172)
            float coef = dpi + dpb;
            if (S.pciB) {
                const float rb = d.bRho[bi];
                coef = dpi + d.bP[bi] / (rb * rb);
            }
            const float g = xb.w * coef * gradFactor(S, fastSqrt(rx * rx + ry * ry + rz * rz));
            ax -= g * rx; ay -= g * ry; az -= g * rz;
        }
        ax = lppSum<kLppOnce>(ax);
        ay = lppSum<kLppOnce>(ay);
        az = lppSum<kLppOnce>(az);
    }
    if (other) {
        fx = lppSum<kLppOnce>(fx);
        fy = lppSum<kLppOnce>(fy);
        fz = lppSum<kLppOnce>(fz);
        if (vm == 0) {
            const float c = S.visc / rhoi;
            fx *= c; fy *= c; fz *= c;
        }
        ax += fx; ay += fy; az += fz;
        if (vm >= 0 && S.bviscOn) {
            float bx = 0.0f, by = 0.0f, bz = 0.0f;
            for (uint32_t s = sub; s < nb; s += kLppOnce) {
                const float4 xb = d.xb[ldList(d.nbrB + s * NP + i)];
                const float rx = xi.x - xb.x, ry = xi.y - xb.y, rz = xi.z - xb.z;
                const float r2 = rx * rx + ry * ry + rz * rz;
                const float rl = fastSqrt(r2);
                if (vm == 2) {
                    const float c = S.bvisc * (xb.w / rhoi) * cubicW(S, rl) * invdt;
                    bx -= c * vi.x; by -= c * vi.y; bz -= c * vi.z;
                } else if (vm == 1) {
                    const float dp = vi.x * rx + vi.y * ry + vi.z * rz;
                    const float c = S.bvisc * (xb.w / rhoi) * dp / (r2 + S.h2c) * gradFactor(S, rl);
                    bx += c * rx; by += c * ry; bz += c * rz;
                } else {
                    const float c = xb.w / rhoi * laplW(S, rl);
                    bx -= c * vi.x; by -= c * vi.y; bz -= c * vi.z;
                }
            }
            bx = lppSum<kLppOnce>(bx);
            by = lppSum<kLppOnce>(by);
            bz = lppSum<kLppOnce>(bz);
            if (vm == 0) {
                const float c = S.bvisc / rhoi;
                bx *= c; by *= c; bz *= c;
            }
            ax += bx; ay += by; az += bz;
        }
        if (normals) {
            nx = lppSum<kLppOnce>(nx);
            ny = lppSum<kLppOnce>(ny);
            nz = lppSum<kLppOnce>(nz);
            if (valid && sub == 0) d.n[i] = make_float4(nx * S.h, ny * S.h, nz * S.h, rhoi);
        }
    }
    if (valid && sub == 0) d.a[i] = make_float4(ax, ay, az, 0.0f);
}

// DFSPHSolver::updateTimeStep (DFSPHSolver.cpp:1090-1162) on the FP64 maximum
// of |v + (a + g) dt|, evaluated in FP32 as the reference does. Runs in
// kFinish, after every block of the reducing launch has used the old dt.
__device__ __forceinline__ void timeStepEpilogue(const DevView& d, int s, double vmax) {
    SceneState& T = d.state[s];
    const SceneConst& S = d.scenes[s];
    T.vmax = vmax;
    if (!S.adaptive || S.nf == 0) return;
    float maxVel = static_cast<float>(vmax);
    if (maxVel < 1.0e-9f) maxVel = 1.0e-9f;
    float h = T.dt;
    const float cflTs = S.cfl * 0.4f * (2.0f * S.radius) / maxVel;
    const int its = max(T.its, T.itsV);
    if (its > 6) h *= 0.9f;
    else if (its < 4) h *= 1.05f;
    h = fminf(h, cflTs);
    h = fmaxf(h, S.minTs);
    h = fminf(h, S.maxTs);
    T.dt = h;
    T.dt2 = h * h;
}

// SurfaceTensionAkinci::step pass 2 (SurfaceTensionAkinci.cpp:35-70).
// With vmax: the CFL maximum and the new time step (DFSPH).
__global__ void __launch_bounds__(kBlock) kSurfaceTension(DevView d, int vmax) {
    SCENE_PROLOGUE_LANES(2, kLppOnce)
    (void)dt2;
    double vm = 0.0;
    const uint32_t ii = valid ? i : 0u;
    float4 ai = d.a[ii];
    if (S.stOn) {
        const size_t NP = d.NP;
        const float4 xi = d.x[ii];
        const float4 ni = d.n[ii];
        const float rhoi = ni.w;
        const uint32_t nf = valid ? min(d.cntF[i], d.capF) : 0u;
        float sx = 0.0f, sy = 0.0f, sz = 0.0f;
        walkList<kLppOnce, kBatch>(d.nbrF, NP, i, sub, nf, [&](uint32_t j) { return XV{d.x[j], d.n[j]}; },
                                   [&](uint32_t, const XV& q) {
                                       const float4 xj = q.x;
                                       const float rx = xi.x - xj.x, ry = xi.y - xj.y, rz = xi.z - xj.z;
                                       const float dd = dRef(rx, ry, rz);
                                       if (dd > 0.0f && dd <= S.dmax) {
                                           const float rl = fastSqrt(dd);
                                           const float4 nj = q.v;   // the normal; w = density
                                           const float kij = 2.0f * S.rho0 / (rhoi + nj.w);
                                           const float c = S.st * S.m * cohesionW(S, dd, rl) / rl;
                                           sx += kij * (-c * rx - S.st * (ni.x - nj.x));
                                           sy += kij * (-c * ry - S.st * (ni.y - nj.y));
                                           sz += kij * (-c * rz - S.st * (ni.z - nj.z));
                                       }
                                   });
        sx = lppSum<kLppOnce>(sx);
        sy = lppSum<kLppOnce>(sy);
        sz = lppSum<kLppOnce>(sz);
        ai.x += sx;
        ai.y += sy;
        ai.z += sz;
        if (valid && sub == 0) d.a[i] = ai;
    }
    // AdhesionForce::step (AdhesionForce.cpp:5-32), the force after surface
    // tension in the reference's order: a -= beta rho0 V_b A(|r|) r / |r| over
    // the boundary neighbours with h/2 < |r| <= h, A = (-4 r^2/h + 6 r - 2h)^(1/4)
    // as Adhesion::W returns it (Adhesion.h:25-35: without its valueCoef).
    if (S.adhOn) {
        const size_t NP = d.NP;
        const float4 xi = d.x[ii];
        const uint32_t nb = valid ? min(d.cntB[i], d.capB) : 0u;
        const float bc = S.adh * S.rho0;
        float ax = 0.0f, ay = 0.0f, az = 0.0f;
        for (uint32_t s = sub; s < nb; s += kLppOnce) {
            const uint32_t b = ldList(d.nbrB + s * NP + i);
            const float4 xb = d.xb[b];
            const float rx = xi.x - xb.x, ry = xi.y - xb.y, rz = xi.z - xb.z;
            const float dd = dRef(rx, ry, rz);
            if (dd > S.dhalf && dd <= S.dmax) {
                const float rl = sqrtf(dd);   // correctly rounded (see lenRef)
                const float val = -4.0f * rl * rl / S.h + 6.0f * rl - 2.0f * S.h;
                const float c = bc * d.vb[b] * powf(val, 0.25f);
                ax -= c * rx / rl;
                ay -= c * ry / rl;
                az -= c * rz / rl;
            }
        }
        ax = lppSum<kLppOnce>(ax);
        ay = lppSum<kLppOnce>(ay);
        az = lppSum<kLppOnce>(az);
        ai.x += ax;
        ai.y += ay;
        ai.z += az;
        if (valid && sub == 0) d.a[i] = ai;
    }
    if (vmax && valid && sub == 0) {
        const float4 vi = VEL(d, i);
        const float cx = vi.x + (ai.x + S.gx) * dt;
        const float cy = vi.y + (ai.y + S.gy) * dt;
        const float cz = vi.z + (ai.z + S.gz) * dt;
        vm = sqrt(static_cast<double>(cx) * cx + static_cast<double>(cy) * cy + static_cast<double>(cz) * cz);
    }
    if (vmax) storePartial<true>(d, vm);
}

__global__ void __launch_bounds__(kBlock) kVmax(DevView d) {
    SCENE_PROLOGUE_LANES(2, kLppOnce)   // same partial layout as kSurfaceTension
    (void)dt2;
    double vm = 0.0;
    if (valid && sub == 0) {
        const float4 vi = VEL(d, i), ai = d.a[i];
        const float cx = vi.x + (ai.x + S.gx) * dt;
        const float cy = vi.y + (ai.y + S.gy) * dt;
        const float cz = vi.z + (ai.z + S.gz) * dt;
        vm = sqrt(static_cast<double>(cx) * cx + static_cast<double>(cy) * cy + static_cast<double>(cz) * cz);
    }
    storePartial<true>(d, vm);
}

// Combine scene s's block partials in a fixed order and file the result (the
// loop control of a Jacobi iteration, or the CFL time step). Called by every
// thread of one block.
__device__ void finishScene(const DevView& d, int s, int kind, int slot) {
    const SceneConst& C = d.scenes[s];
    const bool mx = kind == kFinishTimeStep;
    // one partial per launch block: kLpp (DFSPH source passes) or kLppOnce
    // (vmax of kSurfaceTension / kVmax) launch blocks per particle block
    const uint32_t mul = (kind == kFinishDensity || kind == kFinishDivergence) ? kLpp
                         : kind == kFinishTimeStep                             ? kLppOnce
                                                                               : 1u;
    const uint32_t bBegin = C.blockBegin * mul, bEnd = C.blockEnd * mul;
    // each thread combines partials threadIdx.x, +256, +512, ... in that order;
    // eight loads are in flight before the first is combined
    double acc = 0.0;
    uint32_t b = bBegin + threadIdx.x;
    for (; b + 7u * kBlock < bEnd; b += 8u * kBlock) {
        double q[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) q[u] = d.partials[b + u * kBlock];
#pragma unroll
        for (int u = 0; u < 8; ++u) acc = mx ? fmax(acc, q[u]) : acc + q[u];
    }
    for (; b < bEnd; b += kBlock) {
        const double p = d.partials[b];
        acc = mx ? fmax(acc, p) : acc + p;
    }
    const double tot = mx ? blockReduce<true>(acc) : blockReduce<false>(acc);
    if (threadIdx.x != 0) return;
    SceneState& U = d.state[s];
    const double mean = C.nf ? tot / C.nf : 0.0;
    if (kind == kFinishDensity || kind == kFinishPciDensity)
        loopEpilogue(U, slot, mean, static_cast<double>(C.eta), C.minIt, C.maxIt);
    else if (kind == kFinishDivergence)
        loopEpilogue(U, slot, mean, static_cast<double>(C.etaV / U.dt), C.minItV, C.maxIt);
    else
        timeStepEpilogue(d, s, tot);
}

// The same, done by the first blocks of the launch that follows a Jacobi pass
// (kVelocityFinish, kPciPressureAcc) once their own particles are done: that
// launch reads active[slot] and dt, which the loop epilogue leaves alone (it
// sets active[slot ^ 1], it, err), so the reduction needs no launch of its
// own. kind < 0: nothing to finish.
__device__ __forceinline__ void finishInline(const DevView& d, int kind, int slot) {
    if (kind < 0) return;
    for (int s = blockIdx.x; s < d.numScenes; s += gridDim.x) finishScene(d, s, kind, slot);
}

// DFSPHSolver::predictVelocities (DFSPHSolver.cpp:180-202)
__global__ void __launch_bounds__(kBlock) kPredictVelocity(DevView d) {
    SCENE_PROLOGUE(2)
    if (!valid) return;
    float4 vi = VEL(d, i);
    const float4 ai = d.a[i];
    vi.x += (ai.x + S.gx) * dt;
    vi.y += (ai.y + S.gy) * dt;
    vi.z += (ai.z + S.gz) * dt;
    if (S.pciB) {
        // correctVelocities after predictVelocities (DFSPHSolver.cpp:46-51) with
        // the fluid's normals as they stand: the last correctPositions' (surface
        // tension keeps its normals in an array of its own, SurfaceTensionAkinci.h:15)
        const float4 n = d.nc[i];
        pciReflect(S, vi, vi, n.x, n.y, n.z);
    }
    VEL(d, i) = vi;
    d.a[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// ---------------------------------------------------------------- DFSPH solves
// This lane's share of sum_j m g_ij (v_i - v_j) . r_ij + sum_b rho0 V_b g_ib v_i . r_ib,
// the density change rate of predictDensities (DFSPHSolver.cpp:250-283) and
// computeDivergenceError (:560-606).
template <uint32_t L>
__device__ __forceinline__ float sourceShare(const DevView& d, const SceneConst& S, uint32_t i, uint32_t sub,
                                             float4 xi, float4 vi, uint32_t nf, uint32_t nb, const float4* tX,
                                             const float4* tV) {
    const size_t NP = d.NP;
    float dc = 0.0f;
    auto pair = [&](float4 xj, float4 vj) {
        const float rx = xi.x - xj.x, ry = xi.y - xj.y, rz = xi.z - xj.z;
        const float g = gradFactorSq(S, dRef(rx, ry, rz));
        dc = pairFma(S.m, g * dot3f(vi.x - vj.x, vi.y - vj.y, vi.z - vj.z, rx, ry, rz), dc);
    };
    if (tX) {
        walkTile(d, i, sub, nf, [&](uint32_t t) { pair(tX[t], tV[t]); });
    } else {
        // raw buffer loads: one 32-bit offset per neighbour for both records
        // instead of two 64-bit addresses (num_records bounds every load)
        const uint32_t bytes = static_cast<uint32_t>(NP * sizeof(float4));
        const __amdgpu_buffer_rsrc_t rX = __builtin_amdgcn_make_buffer_rsrc(d.x, 0, bytes, 0x00020000);
        const __amdgpu_buffer_rsrc_t rV = __builtin_amdgcn_make_buffer_rsrc(d.v, 0, bytes, 0x00020000);
        auto gather = [&](uint32_t j) {
            const uint32_t off = j * static_cast<uint32_t>(sizeof(float4));
            return XV{__builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(rX, off, 0, 0)),
                      __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(rV, off, 0, 0))};
        };
        walkRows<L, kPadRows / L, SPH_PRE_SRC != 0>(d.nbrF, d.NP, i, sub, nf, gather,
                                                    [&](const XV& q) { pair(q.x, q.v); });
    }
    for (uint32_t s = sub; s < nb; s += L) {
        const float4 xb = d.xb[ldList(d.nbrB + s * NP + i)];
        const float rx = xi.x - xb.x, ry = xi.y - xb.y, rz = xi.z - xb.z;
        const float g = gradFactorSq(S, dRef(rx, ry, rz));
        dc = pairFma(xb.w, g * dot3f(vi.x, vi.y, vi.z, rx, ry, rz), dc);
    }
    return dc;
}

// DFSPHSolver::predictDensities (DFSPHSolver.cpp:226-299). warm: the
// warm-start rescale of densityWarmStart (:863-886), and block 0 starts the
// density loop; otherwise the k pass of correctDensityError (:435-453) with
// the FP64 average error feeding the loop control.
__global__ void __launch_bounds__(kBlock) kRhoStar(DevView d, int warm, int slot) {
    SCENE_PROLOGUE_LANES(slot, kLpp)
    if (warm && blockIdx.x == 0)
        for (int s = threadIdx.x; s < d.numScenes; s += kBlock) {
            const SceneConst& C = d.scenes[s];
            loopStart(d.state[s], 0.0, static_cast<double>(C.eta), C.minIt, C.maxIt);
        }
    const bool first = T.it == 0;
    __shared__ float4 tX[kTileMax], tV[kTileMax];
    const bool tiled = stageTile(d, lb, T.active[slot] != 0, d.x, tX, d.v, tV);
    double err = 0.0;
    float4 xi = make_float4(0.0f, 0.0f, 0.0f, 0.0f), vi = xi;
    float dc = 0.0f;
    if (valid) {
        xi = d.x[i];
        vi = VEL(d, i);
        dc = sourceShare<kLpp>(d, S, i, sub, xi, vi, min(d.cntF[i], d.capF), S.akinci ? min(d.cntB[i], d.capB) : 0u,
                               tiled ? tX : nullptr, tV);
    }
    dc = lppSum<kLpp>(dc);
    if (valid && sub == 0) {
        const float rhoi = d.rho[i];
        const float pd = fmaxf(rhoi + dc * dt, S.rho0);
        d.aux[i] = pd;
        float kOut;
        if (warm) {
            kOut = pd > S.rho0 ? d.kws[i] / dt2 : 0.0f;
            d.kws[i] = kOut;
            d.k[i] = kOut;
        } else {
            kOut = (pd - S.rho0) * d.alpha[i] / (dt2 * rhoi);
            d.k[i] = kOut;
            d.kws[i] = first ? kOut : d.kws[i] + kOut;
            err = static_cast<double>((pd - S.rho0) / S.rho0);
        }
        d.xk[i] = make_float4(xi.x, xi.y, xi.z, kOut);
    }
    if (!warm) storePartial<false>(d, err);
}

// The velocity correction shared by both solves and both warm starts
// (DFSPHSolver.cpp:455-511, :788-823, :894-941, :998-1045). xk = (x, k) was
// written by the preceding kRhoStar / kDivergence launch, so each neighbour
// costs one 16-byte gather.
// Velocity update of one Jacobi pass, kLppVel lanes per particle; shared by the
// two kernels below.
__device__ __forceinline__ void velocityUpdate(const DevView& d, const SceneConst& S, uint32_t i, uint32_t sub,
                                               bool valid, float dt, const float4* tXK) {
    const size_t NP = d.NP;
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;
    if (valid) {
        const float4 xi = d.xk[i];
        const float ki = xi.w;
        const uint32_t nf = min(d.cntF[i], d.capF);
        const __amdgpu_buffer_rsrc_t rXK =
            __builtin_amdgcn_make_buffer_rsrc(d.xk, 0, static_cast<uint32_t>(NP * sizeof(float4)), 0x00020000);
        auto gatherXK = [&](uint32_t j) {
            return __builtin_bit_cast(
                float4, __builtin_amdgcn_raw_buffer_load_b128(rXK, j * static_cast<uint32_t>(sizeof(float4)), 0, 0));
        };
        // a closed gate (|k_i + k_j| <= eps) adds c = 0: the same sums as
        // skipping the pair, without a branch per pair
        auto pair = [&](const float4& xj) {
            const float ks = ki + xj.w;
            const float rx = xi.x - xj.x, ry = xi.y - xj.y, rz = xi.z - xj.z;
            const float c0 = (S.m * ks) * gradFactorSq(S, dRef(rx, ry, rz));
            const float c = fabsf(ks) > S.eps ? c0 : 0.0f;
            sx = pairFma(c, rx, sx); sy = pairFma(c, ry, sy); sz = pairFma(c, rz, sz);
        };
        if (tXK) walkTile(d, i, sub, nf, [&](uint32_t t) { pair(tXK[t]); });
        else walkRows<kLppVel, kPadRows / kLppVel, true>(d.nbrF, d.NP, i, sub, nf, gatherXK, pair);
        if (S.akinci && fabsf(ki) > S.eps) {
            const uint32_t nb = min(d.cntB[i], d.capB);
            for (uint32_t s = sub; s < nb; s += kLppVel) {
                const float4 xb = d.xb[ldList(d.nbrB + s * NP + i)];
                const float rx = xi.x - xb.x, ry = xi.y - xb.y, rz = xi.z - xb.z;
                const float c = (xb.w * ki) * gradFactorSq(S, dRef(rx, ry, rz));
                sx = pairFma(c, rx, sx); sy = pairFma(c, ry, sy); sz = pairFma(c, rz, sz);
            }
        }
    }
    sx = lppSum<kLppVel>(sx);
    sy = lppSum<kLppVel>(sy);
    sz = lppSum<kLppVel>(sz);
    if (valid && sub == 0) {
        float4 vi = VEL(d, i);
        vi.x -= dt * sx;
        vi.y -= dt * sy;
        vi.z -= dt * sz;
        VEL(d, i) = vi;
        if (d.vRec) d.vRec[i] = vi;   // a recorded rollout's tape
    }
}

__global__ void __launch_bounds__(kBlock) kVelocity(DevView d, int slot) {
    SCENE_PROLOGUE_LANES(slot, kLppVel)
    (void)dt2;
    __shared__ float4 tXK[kTileMax];
    const bool tiled = stageTile(d, lb, T.active[slot] != 0, d.xk, tXK, nullptr, nullptr);
    velocityUpdate(d, S, i, sub, valid, dt, tiled ? tXK : nullptr);
}
// The same, plus the loop reduction of the preceding pass, done by the first
// blocks after their own particles (see finishInline).
__global__ void __launch_bounds__(kBlock) kVelocityFinish(DevView d, int slot, int fin) {
    SCENE_PROLOGUE_LANES(slot, kLppVel)
    (void)dt2;
    __shared__ float4 tXK[kTileMax];
    const bool tiled = stageTile(d, lb, T.active[slot] != 0, d.xk, tXK, nullptr, nullptr);
    velocityUpdate(d, S, i, sub, valid, dt, tiled ? tXK : nullptr);
    finishInline(d, fin, slot);
}

// DFSPHSolver::computeDivergenceError (DFSPHSolver.cpp:538-621); the
// neighbour count n_i (|r| <= h, fluid + boundary, self included) is the
// length of the within-support lists. warm: divergenceWarmStart rescale
// (:968-990), and block 0 starts the divergence loop; otherwise the kv pass of
// correctDivergenceError (:750-768) with the loop control. The loop bound is
// maxIterations, not maxIterationsV (:746).
__global__ void __launch_bounds__(kBlock) kDivergence(DevView d, int warm, int slot) {
    SCENE_PROLOGUE_LANES(slot, kLpp)
    if (warm && blockIdx.x == 0)
        for (int s = threadIdx.x; s < d.numScenes; s += kBlock) {
            const SceneConst& C = d.scenes[s];
            SceneState& U = d.state[s];
            loopStart(U, 0.0, static_cast<double>(C.etaV / U.dt), C.minItV, C.maxIt);
        }
    const bool first = T.it == 0;
    __shared__ float4 tX[kTileMax], tV[kTileMax];
    const bool tiled = stageTile(d, lb, T.active[slot] != 0, d.x, tX, d.v, tV);
    double err = 0.0;
    float4 xi = make_float4(0.0f, 0.0f, 0.0f, 0.0f), vi = xi;
    float de = 0.0f;
    uint32_t cnt = 0;
    if (valid) {
        xi = d.x[i];
        vi = VEL(d, i);
        const uint32_t nfTrue = d.cntF[i];
        const uint32_t nbTrue = S.akinci ? d.cntB[i] : 0u;
        cnt = nfTrue + nbTrue;
        de = sourceShare<kLpp>(d, S, i, sub, xi, vi, min(nfTrue, d.capF), min(nbTrue, d.capB), tiled ? tX : nullptr,
                               tV);
    }
    de = lppSum<kLpp>(de);
    if (valid && sub == 0) {
        de = cnt < 20u ? 0.0f : fmaxf(de, 0.0f);
        d.aux[i] = de;
        float kOut;
        if (warm) {
            kOut = de > 0.0f ? d.kvws[i] / dt : 0.0f;
            d.kvws[i] = kOut;
            d.k[i] = kOut;
        } else {
            kOut = de * d.alpha[i] / (dt * d.rho[i]);
            d.k[i] = kOut;
            d.kvws[i] = first ? kOut : d.kvws[i] + kOut;
            err = static_cast<double>(de / S.rho0);
        }
        d.xk[i] = make_float4(xi.x, xi.y, xi.z, kOut);
    }
    if (!warm) storePartial<false>(d, err);
}

// kv_ws *= ts after the divergence loop (DFSPHSolver.cpp:829-843; cleared if
// the loop did not run); block 0 files the loop's statistics and advances t.
__global__ void __launch_bounds__(kBlock) kScaleDivergence(DevView d) {
    SCENE_PROLOGUE(2)
    if (blockIdx.x == 0)
        for (int s = threadIdx.x; s < d.numScenes; s += kBlock) {
            SceneState& U = d.state[s];
            const SceneConst& C = d.scenes[s];
            U.itsV = U.it;
            U.errV = U.err;
            if (U.it >= C.maxIt && U.err > static_cast<double>(C.etaV / U.dt)) U.capped |= 2;
            U.t = U.t + U.dt;
        }
    if (!valid) return;
    d.kvws[i] *= (T.it > 0 ? dt : 0.0f);
}

// SPHSolver::integrate (SPHSolver.cpp:139-168); block 0 advances t.
__global__ void __launch_bounds__(kBlock) kIntegrate(DevView d) {
    SCENE_PROLOGUE(2)
    if (blockIdx.x == 0)
        for (int s = threadIdx.x; s < d.numScenes; s += kBlock) d.state[s].t = d.state[s].t + d.state[s].dt;
    if (!valid) return;
    float4 ai = d.a[i], vi = VEL(d, i), xi = d.x[i];
    ai.x += S.gx; ai.y += S.gy; ai.z += S.gz;
    vi.x += ai.x * dt; vi.y += ai.y * dt; vi.z += ai.z * dt;
    xi.x += vi.x * dt; xi.y += vi.y * dt; xi.z += vi.z * dt;
    if (S.pciB) pciCorrect(d, S, xi, vi, xi, vi);   // WCSPHSolver.cpp:62-71
    VEL(d, i) = vi;
    d.x[i] = xi;
    d.a[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// ---------------------------------------------------------------- PCISPH
// initPressure (PCISPHSolver.cpp:82-118) + first predictVelocityAndPosition
// (:120-147); block 0 starts the loop with the error carried from the last
// step (PCISPHSolver.h:18).
// CubeBoundaryObject::correctPositionAndVelocity (CubeBoundaryObject.cpp:10-88)
// for each of the scene's cube objects in order: per axis, a position below
// min / above max is clamped, an outward normal velocity scaled by
// -normalFct, and the two other components by tangentialFct.
__global__ void __launch_bounds__(kBlock) kCubeBoundary(DevView d) {
    SCENE_PROLOGUE(2)
    (void)dt;
    (void)dt2;
    if (!valid || S.nCubes == 0) return;
    float4 p = d.x[i], v = VEL(d, i);
    float* pc[3] = {&p.x, &p.y, &p.z};
    float* vc[3] = {&v.x, &v.y, &v.z};
    for (int c = 0; c < S.nCubes; ++c) {
        const float nf = S.cubeNormal[c], tf = S.cubeTang[c];
        for (int a = 0; a < 3; ++a) {
            const int b = (a + 1) % 3, e = (a + 2) % 3;
            if (*pc[a] < S.cubeLo[c][a]) {
                *pc[a] = S.cubeLo[c][a];
                if (*vc[a] < 0.0f) *vc[a] *= -nf;
                *vc[b] *= tf;
                *vc[e] *= tf;
            } else if (*pc[a] > S.cubeHi[c][a]) {
                *pc[a] = S.cubeHi[c][a];
                if (*vc[a] > 0.0f) *vc[a] *= -nf;
                *vc[b] *= tf;
                *vc[e] *= tf;
            }
        }
    }
    d.x[i] = p;
    VEL(d, i) = v;
}

__global__ void __launch_bounds__(kBlock) kPciInit(DevView d) {
    SCENE_PROLOGUE(2)
    if (blockIdx.x == 0)
        for (int s = threadIdx.x; s < d.numScenes; s += kBlock) {
            const SceneConst& C = d.scenes[s];
            SceneState& U = d.state[s];
            loopStart(U, U.err, static_cast<double>(C.eta), C.minIt, C.maxIt);
        }
    if (!valid) return;
    d.p[i] = 0.0f;
    d.ap[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const float4 xi = d.x[i], vi = VEL(d, i), ai = d.a[i];
    const float vx = vi.x + (ai.x + S.gx) * dt;
    const float vy = vi.y + (ai.y + S.gy) * dt;
    const float vz = vi.z + (ai.z + S.gz) * dt;
    float px = xi.x + vx * dt, py = xi.y + vy * dt, pz = xi.z + vz * dt;
    if (S.pciB) {   // PCISPHSolver.cpp:333-341
        pciCorrectPredicted(d, S, i, xi, px, py, pz);
        d.vs[i] = make_float4(vx, vy, vz, 0.0f);
    }
    d.xs[i] = make_float4(px, py, pz, 0.0f);
}

// predictFluidDensities (PCISPHSolver.cpp:162-209): predicted positions,
// candidates of the *current* position's 27 cells (:178); then
// updatePressure (:246-298) with the FP64 average error feeding the loop.
__global__ void __launch_bounds__(kBlock) kPciDensity(DevView d, int slot) {
    SCENE_PROLOGUE(slot)
    double err = 0.0;
    if (valid) {
        const float4 xsi = d.xs[i];
        int cx, cy, cz;
        globalCell(S, d.x[i], cx, cy, cz);
        float rho = 0.0f;
        // as in the search: the next cell's ranges load ahead, U candidates per
        // round; r^2 and the sums are spelled out (dRef, pairFma) so that the
        // unrolled copies round alike
        constexpr uint32_t U = kBuildUnroll;
        auto cellAt = [&](int q, int& ax, int& ay, int& az, bool& bucket) {
            ax = cx + q / 9 - 1;
            ay = cy + (q / 3) % 3 - 1;
            az = cz + q % 3 - 1;
            return cellSlot(S, ax, ay, az, bucket);
        };
        int axN, ayN, azN;
        bool bucketN;
        uint32_t cN = cellAt(0, axN, ayN, azN, bucketN);
        uint2 frN = d.fRange[cN];
        uint2 brN = S.akinci && !bucketN ? d.bRange[cN] : make_uint2(0u, 0u);
        for (int q = 0; q < 27; ++q) {
            const int ax = axN, ay = ayN, az = azN;
            const bool bucket = bucketN;
            const uint2 fr = frN, br = brN;
            if (q + 1 < 27) {
                cN = cellAt(q + 1, axN, ayN, azN, bucketN);
                frN = d.fRange[cN];
                brN = S.akinci && !bucketN ? d.bRange[cN] : make_uint2(0u, 0u);
            }
            if (!bucket) {
                for (uint32_t base = fr.x; base < fr.y; base += U) {
                    float4 xj[U];
#pragma unroll
                    for (uint32_t k = 0; k < U; ++k) xj[k] = d.xs[min(base + k, fr.y - 1u)];
#pragma unroll
                    for (uint32_t k = 0; k < U; ++k)
                        if (base + k < fr.y)
                            rho = pairFma(S.m, cubicW(S, fastSqrt(dRef(xsi.x - xj[k].x, xsi.y - xj[k].y, xsi.z - xj[k].z))),
                                          rho);
                }
            } else {
                // an overflow bucket holds several cells: keep the candidates
                // whose current position lies in this one, and gather only those
                for (uint32_t j = fr.x; j < fr.y; ++j) {
                    if (!inCell(S, d.x[j], ax, ay, az)) continue;
                    const float4 xj = d.xs[j];
                    rho = pairFma(S.m, cubicW(S, fastSqrt(dRef(xsi.x - xj.x, xsi.y - xj.y, xsi.z - xj.z))), rho);
                }
            }
            for (uint32_t base = br.x; base < br.y; base += U) {
                float4 xb[U];
#pragma unroll
                for (uint32_t k = 0; k < U; ++k) xb[k] = d.xb[min(base + k, br.y - 1u)];
#pragma unroll
                for (uint32_t k = 0; k < U; ++k)
                    if (base + k < br.y)
                        rho = pairFma(xb[k].w, cubicW(S, fastSqrt(dRef(xsi.x - xb[k].x, xsi.y - xb[k].y, xsi.z - xb[k].z))), rho);
            }
        }
        d.rho[i] = rho;
        const float e = fmaxf(0.0f, rho - S.rho0);
        const float pi_ = d.p[i] + S.pciDelta * e;
        d.p[i] = pi_;
        storeXW(d, i, pi_ / (rho * rho));
        err = static_cast<double>(e / S.rho0);
    }
    storePartial<false>(d, err);
}

// computePressureAcc (PCISPHSolver.cpp:300-367) at current positions with
// the predicted densities, then the next predictVelocityAndPosition.
__device__ __forceinline__ void pciPressureAcc(const DevView& d, const SceneConst& S, uint32_t i, float dt) {
    const size_t NP = d.NP;
    const float4 xi = d.x[i];
    const float rhoi = d.rho[i];
    const float dpi = d.p[i] / (rhoi * rhoi);
    float ax = 0.0f, ay = 0.0f, az = 0.0f;
    const uint32_t nf = min(d.cntF[i], d.capF);
    // w of the gathered position is p_j / rho_j^2, filed by kPciDensity
    walkList<1, kBatch>(d.nbrF, NP, i, 0u, nf, [&](uint32_t j) { return d.x[j]; },
                        [&](uint32_t, const float4& xj) {
                            const float rx = xi.x - xj.x, ry = xi.y - xj.y, rz = xi.z - xj.z;
                            const float g = S.m * (dpi + xj.w) * gradFactor(S, fastSqrt(rx * rx + ry * ry + rz * rz));
                            ax -= g * rx; ay -= g * ry; az -= g * rz;
                        });
    const uint32_t nb = S.akinci ? min(d.cntB[i], d.capB) : 0u;
    for (uint32_t s = 0; s < nb; ++s) {
        const uint32_t bi = ldList(d.nbrB + s * NP + i);
        const float4 xb = d.xb[bi];
        const float rx = xi.x - xb.x, ry = xi.y - xb.y, rz = xi.z - xb.z;
        float g;
        if (S.pciB) {   // m_b (p_i / rho_i^2 + p_b / rho_b^2) Spiky::gradW (:339-351)
            const float rb = d.bRho[bi];
            g = xb.w * (dpi + d.bP[bi] / (rb * rb)) * spikyFactor(S, lenRef(rx, ry, rz));
        } else {
            g = xb.w * dpi * gradFactor(S, fastSqrt(rx * rx + ry * ry + rz * rz));
        }
        ax -= g * rx; ay -= g * ry; az -= g * rz;
    }
    d.ap[i] = make_float4(ax, ay, az, 0.0f);
    const float4 vi = VEL(d, i), ai = d.a[i];
    const float vx = vi.x + ((ai.x + ax) + S.gx) * dt;
    const float vy = vi.y + ((ai.y + ay) + S.gy) * dt;
    const float vz = vi.z + ((ai.z + az) + S.gz) * dt;
    float px = xi.x + vx * dt, py = xi.y + vy * dt, pz = xi.z + vz * dt;
    if (S.pciB) {
        // keep the prediction this iteration used (the last one, when the loop
        // stops here, is what correctPositions / correctVelocities read)
        d.xsP[i] = d.xs[i];
        d.vsP[i] = d.vs[i];
        pciCorrectPredicted(d, S, i, xi, px, py, pz);
        d.vs[i] = make_float4(vx, vy, vz, 0.0f);
    }
    d.xs[i] = make_float4(px, py, pz, 0.0f);
}
__global__ void __launch_bounds__(kBlock) kPciPressureAcc(DevView d, int slot, int fin) {
    SCENE_PROLOGUE(slot)
    (void)dt2;
    if (valid) pciPressureAcc(d, S, i, dt);
    finishInline(d, fin, slot);
}

// PCISPHSolver::integrate (PCISPHSolver.cpp:459-488); block 0 files the
// loop statistics and advances t.
__global__ void __launch_bounds__(kBlock) kPciIntegrate(DevView d) {
    SCENE_PROLOGUE(2)
    if (blockIdx.x == 0)
        for (int s = threadIdx.x; s < d.numScenes; s += kBlock) {
            SceneState& U = d.state[s];
            const SceneConst& C = d.scenes[s];
            U.its = U.it;
            U.errD = U.err;
            U.capped = (U.it >= C.maxIt && U.err > static_cast<double>(C.eta)) ? 1 : 0;
            U.t = U.t + U.dt;
        }
    if (!valid) return;
    float4 ai = d.a[i], vi = VEL(d, i), xi = d.x[i];
    const float4 pa = d.ap[i];
    ai.x += S.gx + pa.x; ai.y += S.gy + pa.y; ai.z += S.gz + pa.z;
    vi.x += ai.x * dt; vi.y += ai.y * dt; vi.z += ai.z * dt;
    xi.x += vi.x * dt; xi.y += vi.y * dt; xi.z += vi.z * dt;
    if (S.pciB) pciCorrect(d, S, xi, vi, d.xsP[i], d.vsP[i]);   // PCISPHSolver.cpp:64-73
    VEL(d, i) = vi;
    d.x[i] = xi;
    d.a[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// ---------------------------------------------------------------- utilities
__global__ void __launch_bounds__(kBlock) kCellsOut(DevView d, int32_t* cells) {
    const uint32_t lb = logicalBlock();
    const int sIdx = d.blockScene[lb];
    const SceneConst& S = d.scenes[sIdx];
    const uint32_t i = lb * kBlock + threadIdx.x;
    if (i >= S.fEnd) return;
    const float4 p = d.x[i];
    cells[3 * i] = floorCell(p.x, S.invh);
    cells[3 * i + 1] = floorCell(p.y, S.invh);
    cells[3 * i + 2] = floorCell(p.z, S.invh);
}

// The cell keys of one scene's active particles in original order:
// cells[3 * id + a] for the slots [fBegin, fEnd) (ids are 0 .. active - 1).
__global__ void __launch_bounds__(kBlock) kCellsOutOriginal(DevView d, int scene, int32_t* cells) {
    const SceneConst& S = d.scenes[scene];
    const uint32_t i = S.fBegin + blockIdx.x * kBlock + threadIdx.x;
    if (i >= S.fEnd) return;
    const float4 p = d.x[i];
    const uint32_t o = d.id[i];
    cells[3 * o] = floorCell(p.x, S.invh);
    cells[3 * o + 1] = floorCell(p.y, S.invh);
    cells[3 * o + 2] = floorCell(p.z, S.invh);
}

__global__ void __launch_bounds__(kBlock) kCheckFinite(DevView d) {
    const uint32_t lb = logicalBlock();
    const int sIdx = d.blockScene[lb];
    const SceneConst& S = d.scenes[sIdx];
    const uint32_t i = lb * kBlock + threadIdx.x;
    if (i >= S.fEnd) return;
    const float4 p = d.x[i], v = VEL(d, i);
    if (!(isfinite(p.x) && isfinite(p.y) && isfinite(p.z) && isfinite(v.x) && isfinite(v.y) && isfinite(v.z)))
        atomicOr(&d.flags[sIdx * kFlagSlots + kFlagNonFinite], 1);
}

// ---------------------------------------------------------------- reductions
// One block per scene: the scene's block partials of the preceding launch in
// a fixed order (thread t takes blocks t, t+256, ...; then the block tree),
// the same order whatever other scenes share the launch, then the epilogue:
// density / PCISPH loop control (eta), divergence loop control (etaV / dt,
// DFSPHSolver.cpp:746), or the CFL time step.
__global__ void __launch_bounds__(kBlock) kFinish(DevView d, int kind, int slot) {
    finishScene(d, blockIdx.x, kind, slot);
}


// ---------------------------------------------------------------- launchers
// Profiling events for the next launch (set by the engine around one launch):
// hipExtLaunchKernelGGL has the runtime record them at the kernel's own
// start and end, so a launch's time excludes the dispatch around it.
static thread_local hipEvent_t tlEvStart = nullptr, tlEvStop = nullptr;
static thread_local bool tlEvUsed = false;
void setLaunchEvents(hipEvent_t start, hipEvent_t stop) {
    tlEvStart = start;
    tlEvStop = stop;
    tlEvUsed = false;
}
bool launchEventsUsed() { return tlEvUsed; }
#define LAUNCH_SHMEM(kernel, grid, shmem, ...)                                                               \
    do {                                                                                                     \
        if (tlEvStart) {                                                                                     \
            hipExtLaunchKernelGGL(kernel, dim3(grid), dim3(kBlock), shmem, st, tlEvStart, tlEvStop, 0,        \
                                  __VA_ARGS__);                                                              \
            tlEvUsed = true;                                                                                 \
        } else {                                                                                             \
            hipLaunchKernelGGL(kernel, dim3(grid), dim3(kBlock), shmem, st, __VA_ARGS__);                    \
        }                                                                                                    \
        return hipGetLastError();                                                                            \
    } while (0)
#define LAUNCH(kernel, grid, ...) LAUNCH_SHMEM(kernel, grid, 0, __VA_ARGS__)
// Jacobi-pass launches may reserve unused dynamic LDS to cap how many
// workgroups share a CU (and its 32 KB vector L1); 0 leaves occupancy to the
// register count.
#ifndef SPH_JACOBI_LDS
#define SPH_JACOBI_LDS 0
#endif
#define LAUNCH_JACOBI(kernel, grid, ...) LAUNCH_SHMEM(kernel, grid, SPH_JACOBI_LDS, __VA_ARGS__)

hipError_t launchKeys(const DevView& d, int adv, int clearCells, int counting, uint32_t* keys, uint32_t* vals,
                      hipStream_t st) {
    LAUNCH(kKeys, d.numBlocks, d, adv, clearCells, counting, keys, vals);
}
hipError_t launchScatter(const DevView& d, const uint32_t* keys, uint32_t* perm, hipStream_t st) {
    LAUNCH(kScatter, d.numBlocks, d, keys, perm);
}
hipError_t launchOrderFix(const DevView& d, uint32_t* perm, uint32_t ordTotal, uint8_t* occ, int full,
                          hipStream_t st) {
    const uint32_t g = (ordTotal + kBlock - 1) / kBlock;
    if (g == 0) return hipSuccess;
    LAUNCH(kOrderFix, g, d, perm, ordTotal, occ, full);
}
hipError_t launchFillRanges(uint2* r, size_t n, hipStream_t st) {
    const size_t g = (n + kBlock - 1) / kBlock;
    if (g == 0) return hipSuccess;
    LAUNCH(kFillRanges, g, r, n);
}
hipError_t launchReorder(const DevView& d, const uint32_t* ks, const uint32_t* vs, float4* xa, float4* va, float* kwa,
                         float* kva, uint32_t* ida, int ranges, hipStream_t st) {
    LAUNCH(kReorder, d.numBlocks, d, ks, vs, xa, va, kwa, kva, ida, ranges);
}
hipError_t launchBuildNeighbors(const DevView& d, int lanes, hipStream_t st) {
    switch (lanes) {
    case 2: LAUNCH(kBuildNeighbors<2>, d.numBlocks * 2, d);
    case 4: LAUNCH(kBuildNeighbors<4>, d.numBlocks * 4, d);
    case 8: LAUNCH(kBuildNeighbors<8>, d.numBlocks * 8, d);
    default: LAUNCH(kBuildNeighbors<1>, d.numBlocks, d);
    }
}
hipError_t launchTileLists(const DevView& d, hipStream_t st) {
    LAUNCH(kTileLists, d.numBlocks * (kBlock / kTileGroup), d);
}
hipError_t launchBoundaryVolume(const DevView& d, int scene, uint32_t b0, uint32_t b1, hipStream_t st) {
    const uint32_t g = (b1 - b0 + kBlock - 1) / kBlock;
    if (g == 0) return hipSuccess;
    LAUNCH(kBoundaryVolume, g, d, scene, b0, b1);
}
hipError_t launchBoundaryDensity(const DevView& d, int scene, uint32_t b0, uint32_t b1, int predicted, int slot,
                                 hipStream_t st) {
    const uint32_t g = (b1 - b0 + kBlock - 1) / kBlock;
    if (g == 0) return hipSuccess;
    LAUNCH(kBoundaryDensity, g, d, scene, b0, b1, predicted, slot);
}
hipError_t launchDensityAlpha(const DevView& d, int wa, hipStream_t st) { LAUNCH(kDensityAlpha, d.numBlocks * kLppOnce, d, wa); }
hipError_t launchDensityPressure(const DevView& d, hipStream_t st) {
    LAUNCH(kDensityPressure, d.numBlocks * kLppOnce, d);
}
hipError_t launchForces(const DevView& d, int pressure, hipStream_t st) { LAUNCH(kForces, d.numBlocks * kLppOnce, d, pressure); }
hipError_t launchSurfaceTension(const DevView& d, int vmax, hipStream_t st) {
    LAUNCH(kSurfaceTension, d.numBlocks * kLppOnce, d, vmax);
}
hipError_t launchVmax(const DevView& d, hipStream_t st) { LAUNCH(kVmax, d.numBlocks * kLppOnce, d); }
hipError_t launchPredictVelocity(const DevView& d, hipStream_t st) { LAUNCH(kPredictVelocity, d.numBlocks, d); }
hipError_t launchRhoStar(const DevView& d, int warm, int slot, hipStream_t st) {
    LAUNCH_JACOBI(kRhoStar, d.numBlocks * kLpp, d, warm, slot);
}
hipError_t launchVelocity(const DevView& d, int slot, int fin, hipStream_t st) {
    if (fin < 0) LAUNCH_JACOBI(kVelocity, d.numBlocks * kLppVel, d, slot);
    LAUNCH_JACOBI(kVelocityFinish, d.numBlocks * kLppVel, d, slot, fin);
}
hipError_t launchDivergence(const DevView& d, int warm, int slot, hipStream_t st) {
    LAUNCH_JACOBI(kDivergence, d.numBlocks * kLpp, d, warm, slot);
}
hipError_t launchScaleDivergence(const DevView& d, hipStream_t st) { LAUNCH(kScaleDivergence, d.numBlocks, d); }
__global__ void kFlagClear(int32_t* flags, int32_t* acc, uint32_t n) {
    for (uint32_t k = threadIdx.x; k < n; k += blockDim.x) {
        acc[k] = 0;
        flags[k] = 0;
    }
}
hipError_t launchFlagClear(int32_t* flags, int32_t* acc, uint32_t n, hipStream_t st) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(kFlagClear, dim3(1), dim3(256), 0, st, flags, acc, n);
    return hipGetLastError();
}
hipError_t launchIntegrate(const DevView& d, hipStream_t st) { LAUNCH(kIntegrate, d.numBlocks, d); }
hipError_t launchPciInit(const DevView& d, hipStream_t st) { LAUNCH(kPciInit, d.numBlocks, d); }
hipError_t launchPciDensity(const DevView& d, int slot, hipStream_t st) { LAUNCH(kPciDensity, d.numBlocks, d, slot); }
hipError_t launchPciPressureAcc(const DevView& d, int slot, int fin, hipStream_t st) {
    LAUNCH(kPciPressureAcc, d.numBlocks, d, slot, fin);
}
hipError_t launchPciIntegrate(const DevView& d, hipStream_t st) { LAUNCH(kPciIntegrate, d.numBlocks, d); }
hipError_t launchFinish(const DevView& d, int kind, int slot, hipStream_t st) {
    LAUNCH(kFinish, d.numScenes, d, kind, slot);
}
hipError_t launchCellsOut(const DevView& d, int32_t* cells, hipStream_t st) { LAUNCH(kCellsOut, d.numBlocks, d, cells); }
// copy blockIdx.y of the batch, grid-strided over its words
__global__ void __launch_bounds__(kBlock) kCopyBatch(CopyBatch b) {
    const int c = static_cast<int>(blockIdx.y);
    if (c >= b.n) return;
    const uint4* src = b.src[c];
    uint4* dst = b.dst[c];
    const uint64_t n = b.words[c];
    for (uint64_t w = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; w < n;
         w += static_cast<uint64_t>(gridDim.x) * kBlock)
        dst[w] = src[w];
}
hipError_t launchCopyBatch(const CopyBatch& b, hipStream_t st) {
    if (b.n <= 0) return hipSuccess;
    uint64_t most = 0;
    for (int c = 0; c < b.n; ++c) most = b.words[c] > most ? b.words[c] : most;
    if (most == 0) return hipSuccess;
    const uint64_t want = (most + kBlock - 1) / kBlock;
    const uint32_t gx = static_cast<uint32_t>(want < 2048 ? want : 2048);
    hipLaunchKernelGGL(kCopyBatch, dim3(gx, static_cast<uint32_t>(b.n)), dim3(kBlock), 0, st, b);
    return hipGetLastError();
}
hipError_t launchCellsOutOriginal(const DevView& d, int scene, uint32_t count, int32_t* cells, hipStream_t st) {
    const uint32_t g = (count + kBlock - 1) / kBlock;
    if (g == 0) return hipSuccess;
    LAUNCH(kCellsOutOriginal, g, d, scene, cells);
}
hipError_t launchCubeBoundary(const DevView& d, hipStream_t st) { LAUNCH(kCubeBoundary, d.numBlocks, d); }
hipError_t launchCheckFinite(const DevView& d, hipStream_t st) { LAUNCH(kCheckFinite, d.numBlocks, d); }

// The viewer's per-frame data (System.cpp:119-125): the vec4 positions it
// instances (SimulationObject::getPositions, SimulationObject.h:16; w = 1 as
// FluidObject::build sets it, FluidObject.cpp:55) and the scalar field |v|
// (SPHSimulation::updateScalarField, SPHSimulation.cpp:638-657, glm::length in
// the reference's FMA rounding), in original order, for the scene's active
// particles, written straight into the caller's device buffers.
__global__ void __launch_bounds__(kBlock) kViewerFeed(DevView d, int scene, float4* pos, float* speed) {
    const SceneConst& S = d.scenes[scene];
    const uint32_t i = S.fBegin + blockIdx.x * kBlock + threadIdx.x;
    if (i >= S.fEnd) return;
    const uint32_t o = d.id[i];
    if (pos) {
        const float4 p = d.x[i];
        pos[o] = make_float4(p.x, p.y, p.z, 1.0f);
    }
    if (speed) {
        const float4 v = VEL(d, i);
        speed[o] = lenRef(v.x, v.y, v.z);
    }
}
hipError_t launchViewerFeed(const DevView& d, int scene, uint32_t count, float4* pos, float* speed, hipStream_t st) {
    const uint32_t g = (count + kBlock - 1) / kBlock;
    if (g == 0) return hipSuccess;
    LAUNCH(kViewerFeed, g, d, scene, pos, speed);
}

}  // namespace sph

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

namespace sph {
hipError_t scanCounts(void* temp, size_t& tempBytes, const uint32_t* counts, uint32_t* starts, uint32_t n,
                      hipStream_t st) {
    return rocprim::exclusive_scan(temp, tempBytes, counts, starts, 0u, static_cast<size_t>(n), rocprim::plus<uint32_t>(),
                                   st);
}
hipError_t sortPairs(void* temp, size_t& tempBytes, uint32_t* k0, uint32_t* k1, uint32_t* v0, uint32_t* v1, int n,
                     int bits, hipStream_t st, bool* swapped) {
    // rocprim picks a block sort + merge sort up to 2^20 keys and onesweep
    // above; SPH_SORT_ONESWEEP=1 forces onesweep (measured: about 0.5% slower
    // per step on the 1M-particle dam break: fewer but longer launches).
#ifndef SPH_SORT_ONESWEEP
#define SPH_SORT_ONESWEEP 0
#endif
    using Config = rocprim::radix_sort_config<rocprim::default_config, rocprim::default_config,
                                              rocprim::default_config, SPH_SORT_ONESWEEP ? 0 : (1u << 20)>;
    rocprim::double_buffer<uint32_t> kb(k0, k1), vb(v0, v1);
    hipError_t e = rocprim::radix_sort_pairs<Config>(temp, tempBytes, kb, vb, n, 0u, static_cast<unsigned>(bits), st);
    if (swapped) *swapped = (kb.current() == k1);
    return e;
}
}  // namespace sph
