// Host engine of the MI355X particle step and the C ABI of include/sph2o.h.
//
// The engine owns every device buffer of a context, lays scenes out in
// 256-slot blocks, runs the neighbour search (cell keys -> radix sort ->
// reorder + cell ranges -> within-support lists) and drives the per-method
// step exactly in the reference's call order:
//   WCSPH  WCSPHSolver::step   (WCSPHSolver.cpp:14-78)
//   PCISPH PCISPHSolver::step  (PCISPHSolver.cpp:20-80)
//   DFSPH  DFSPHSolver::step   (DFSPHSolver.cpp:33-107)
// The solver loops stay on the host: each iteration launches its two
// neighbour passes, reads back the FP64 per-scene error while the second
// pass runs, and decides per scene whether to iterate again. Scenes of one
// context advance together; a converged scene sits out later iterations.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <limits>
#include <memory>
#include <mutex>
#include <new>
#include <numeric>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/sph2o.h"
#include "scene.hpp"
#include "sph_adjoint.hpp"
#include "sph_kernels.hpp"

using namespace sph;

// Device scratch of the allocation-free entry point sph_dump_cells_noalloc:
// the cell keys of one scene's active particles in original order (3 int32
// per particle), at most kNoAllocMaxParticles particles (12 MiB). Every GPU
// the library runs on holds its own copy until the process ends. In host code
// the name stands for a host-side shadow, not a device address: kernels and
// copies get the address hipGetSymbolAddress returns for the current device.
static constexpr size_t kNoAllocMaxParticles = static_cast<size_t>(1) << 20;
static __device__ __align__(256) int32_t scratch[3 * kNoAllocMaxParticles];

namespace {

// Per-device address of `scratch`, looked up once per device (with that device
// current) when the first context on it is initialised.
std::mutex gScratchMutex;
std::vector<int32_t*> gScratchAddr;
hipError_t scratchAddress(int device, int32_t** out) {
    std::lock_guard<std::mutex> lock(gScratchMutex);
    if (device < 0) return hipErrorInvalidDevice;
    if (gScratchAddr.size() <= static_cast<size_t>(device)) gScratchAddr.resize(static_cast<size_t>(device) + 1, nullptr);
    if (!gScratchAddr[device]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        gScratchAddr[device] = static_cast<int32_t*>(ptr);
    }
    *out = gScratchAddr[device];
    return hipSuccess;
}

struct HipError {
    hipError_t e;
    const char* where;
};
#define HIPCHK(call)                                        \
    do {                                                    \
        hipError_t e_ = (call);                             \
        if (e_ != hipSuccess) throw HipError{e_, #call};    \
    } while (0)

struct SphError {
    int code;
    std::string msg;
};

template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) {
        o.p = nullptr;
        o.n = 0;
    }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) {
            release();
            p = o.p;
            n = o.n;
            o.p = nullptr;
            o.n = 0;
        }
        return *this;
    }
    void alloc(size_t count) {
        release();
        if (count == 0) count = 1;
        HIPCHK(hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T)));
        n = count;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    ~DevBuf() { release(); }
};

enum KernelId {
    K_KEYS, K_SORT, K_REORDER, K_BUILD, K_DENSITY, K_FORCES, K_SURFACE_TENSION, K_VMAX, K_PREDICT, K_RHO_STAR,
    K_VELOCITY, K_DIVERGENCE, K_SCALE, K_INTEGRATE, K_PCI_INIT, K_PCI_DENSITY, K_PCI_PACC, K_PCI_INTEGRATE,
    K_FILL, K_FINISH, K_ADJ_SEED, K_ADJ_FORCES, K_ADJ_NORMALS, K_ADJ_DENSITY, K_ADJ_FINISH, K_ADJ_VELOCITY,
    K_ADJ_SOURCE, K_CUBE, K_TILES, K_TAPE, K_BDENSITY, K_COUNT
};
const char* kKernelNames[K_COUNT] = {"keys", "sort", "reorder", "build_neighbors", "density", "forces",
                                     "surface_tension", "vmax", "predict_velocity", "rho_star", "velocity",
                                     "divergence", "scale", "integrate", "pci_init", "pci_density",
                                     "pci_pressure_acc", "pci_integrate", "fill_ranges", "finish",
                                     "adj_seed", "adj_forces", "adj_normals", "adj_density", "adj_finish",
                                     "adj_velocity", "adj_source", "cube_boundary", "tile_lists", "tape_copy",
                                     "boundary_density"};

// Host mirror of the reference's FP32 cubic-spline gradient (CubicSpline.h:52-74)
// with the compiled reference's length; used for the PCISPH scaling factor.
float hostLen(float x, float y, float z) { return std::sqrt(std::fma(x, x, std::fma(z, z, y * y))); }
void hostGradW(float h, float lW, const float r[3], float out[3]) {
    const float rl = hostLen(r[0], r[1], r[2]);
    const float q = rl / h;
    out[0] = out[1] = out[2] = 0.0f;
    if (rl > 1.0e-9f && q <= 1.0f) {
        float gq[3] = {r[0] / rl, r[1] / rl, r[2] / rl};
        for (int a = 0; a < 3; ++a) gq[a] /= h;
        float f;
        if (q <= 0.5f) f = lW * q * (3.0f * q - 2.0f);
        else {
            const float t = 1.0f - q;
            f = lW * (-t * t);
        }
        for (int a = 0; a < 3; ++a) out[a] = f * gq[a];
    }
}

// Largest float d with sqrt_rn(d) <= t (sqrt_rn is monotonic, so
// sqrt_rn(d) <= t  <=>  d <= sqrtThreshold(t)). std::sqrt on float is the
// correctly rounded IEEE square root.
float sqrtThreshold(float t) {
    float d = t * t;
    while (std::sqrt(d) > t) d = std::nextafter(d, 0.0f);
    while (std::sqrt(std::nextafter(d, INFINITY)) <= t) d = std::nextafter(d, INFINITY);
    return d;
}

struct SceneH {
    sph_scene_params P{};
    std::vector<float> xf, vf, xb;  // original order
    uint32_t nf = 0, nb = 0;
    // FP32 constants as the reference derives them
    float r = 0, h = 0, invh = 0, m = 0, rho0 = 0, minTs = 0, maxTs = 0;
    float dt = 0, t = 0;
    float kW = 0, lW = 0, laplC = 0, cohC1 = 0, cohC2 = 0, pciDelta = 0;
    bool viscOn = false, stOn = false, akinci = false;
    int its = 0, itsV = 0;
    double pciAvg = 0;
    // layout
    uint32_t fBegin = 0, fEnd = 0, fPadEnd = 0, blockBegin = 0, blockEnd = 0, bBegin = 0, bEnd = 0;
    int g0[3] = {0, 0, 0}, gn[3] = {1, 1, 1};
    uint32_t cellBase = 0, nCells = 0, keyBase = 0, keyRange = 0, ordBase = 0;
    uint32_t ovBuckets = 1024;  // overflow buckets for cells outside the dense box (power of two)
    int morton = 0;
    std::vector<uint32_t> bOrder;  // sorted boundary position -> original index
    // emitters: the first na of the nf particles are active (FluidObject's
    // numActiveParticles); the rest wait in the emitters' budgets
    uint32_t na = 0;
    struct Emitter {
        sph_emitter_params P;
        std::vector<float> group;   // local offsets of one emission (xyz)
        int64_t remaining = 0;      // particles left to emit
        float next = 0.0f;          // time of the next emission
    };
    std::vector<Emitter> emitters;
    bool adhOn = false;   // Akinci adhesion (adhesion method 4, coefficient != 0, Akinci boundary)
    struct Cube {
        float lo[3], hi[3], normal, tang;
    };
    std::vector<Cube> cubes;   // cube boundary objects (boundary method 0)
    // PCISPH boundary object (boundary method 1): normals in original order
    bool pciB = false, normalsSet = false;
    std::vector<float> bn;
    float nFct = 0.0f, tFct = 0.0f;
    // the dense box the cell-first marks were last laid out for
    bool markedLayout = false;
    int markG0[3] = {0, 0, 0}, markGn[3] = {0, 0, 0};
    uint32_t markBase = 0;
};

class Engine {
   public:
    int device = 0;
    hipStream_t stream = nullptr;
    bool ownStream = false;
    std::string lastError;
    std::vector<SceneH> scenes;
    bool initialized = false;
    bool fuseFinish = true;    // Jacobi-loop reductions ride on the velocity launch
    int method = -1;

    size_t NP = 0, NB = 0, totalCells = 0;
    uint32_t numBlocks = 0;
    int keyBits = 1;
    uint32_t capF = 64, capB = 48;

    DevBuf<SceneConst> dScenes;
    DevBuf<SceneState> dState;   // scene states, then the flags (one readback copy)
    DevBuf<int32_t> dBlockScene;
    DevBuf<float4> x, xAlt, v, vAlt, a, n, xs, ap, xk, xb;
    // PCISPH boundary objects: boundary normals, densities, pressures; predicted
    // velocities and the last iteration's predictions (PCISPH); cell marks
    DevBuf<float4> bn, vs, xsP, vsP;
    DevBuf<float> bRho, bP;
    DevBuf<uint32_t> cellFirst;
    uint32_t stepNo = 0;   // steps taken since sph_init (the hash table's clock)
    uint32_t markNo = 0;   // the step from which the cells of the search in progress exist
    DevBuf<float4> nc, ncAlt;   // DFSPH: normal of the last correctPositions per slot
    DevBuf<float> rho, alpha, p, k, kws, kwsAlt, kvws, kvwsAlt, aux, vb;
    DevBuf<uint32_t> id, idAlt, keys, keysAlt, vals, valsAlt;
    DevBuf<uint2> fRange, bRange;
    DevBuf<uint32_t> cellOf;     // per slot: the cell range it opened in the last reorder, or ~0
    // counting sort of the keys (the default; radix sort when a cell overflows it)
    DevBuf<uint32_t> rankOf, ordToCell, counts, starts, li;
    DevBuf<uint8_t> occ;         // per order index: held slots at the last counting fix-up
    bool rangesFull = true;      // the next counting fix-up writes every range (occ is stale)
    DevBuf<unsigned char> scanTemp;
    size_t scanTempBytes = 0;
    uint32_t ordTotal = 0;
    bool radixOnce = false;      // the next search uses the radix sort (a cell was too crowded)
    bool rangesDirty = true;     // fRange holds ranges cellOf does not list: clear all of it
    DevBuf<uint32_t> nbrF, cntF, nbrB, cntB;
    // DFSPH: per group of kTileGroup slots the fluid slots its lists name
    // (kTileMax per group), their count (~0: too many, the group gathers from
    // HBM), and the lists as 16-bit indices into that set, two per word
    DevBuf<uint32_t> tileSlot, tileCount, nbrT;
    bool tiles = false;
    DevBuf<double> partials;
    DevBuf<float4> snapV;   // DFSPH: v and kv_ws before the passes that follow a neighbour search
    DevBuf<float> snapKvws;
    int32_t* dFlags = nullptr;   // inside dState's allocation
    DevBuf<unsigned char> sortTemp;
    size_t sortTempBytes = 0;

    // pinned host mirrors of the device state and flags (one allocation)
    SceneState* hState = nullptr;
    int32_t* hFlags = nullptr;
    size_t stateBlockBytes = 0;
    hipEvent_t evStart = nullptr, evStop = nullptr, evSync = nullptr;
    int64_t syncsInStep = 0;
    // Lanes per particle of the search (buildLists): searchLanesFixed > 0 pins
    // the count (sph_set_search_lanes), 0 lets the tuner pick searchLanes.
    int searchLanesFixed = 0, searchLanes = 1;
    struct SearchTuner {
        static constexpr int kCands = 4, kRounds = 3;
        static constexpr int kLanes[kCands] = {1, 2, 4, 8};
        int trial = 0, pendingIdx = 0;
        bool done = false, pending = false;
        uint64_t activeAtPick = 0;   // active particles of all scenes when the count was picked
        float best[kCands] = {std::numeric_limits<float>::infinity(), std::numeric_limits<float>::infinity(),
                              std::numeric_limits<float>::infinity(), std::numeric_limits<float>::infinity()};
        hipEvent_t e0 = nullptr, e1 = nullptr;
    } tuner;
    int32_t* noAllocScratch = nullptr;   // this device's `scratch` (set at init)

    // profiling
    bool profiling = false;
    uint64_t profileMask = ~0ull;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> evPool;
    std::vector<int> evKernel;
    size_t evUsed = 0;
    double kernelMs[K_COUNT] = {0};
    int64_t kernelLaunches[K_COUNT] = {0};
    double lastStepMs = 0;
    int64_t launchesInStep = 0;

    ~Engine() {
        releasePinned();
        for (auto& e : evPool) {
            (void)hipEventDestroy(e.first);
            (void)hipEventDestroy(e.second);
        }
        if (evStart) (void)hipEventDestroy(evStart);
        if (evStop) (void)hipEventDestroy(evStop);
        if (evSync) (void)hipEventDestroy(evSync);
        if (tuner.e0) (void)hipEventDestroy(tuner.e0);
        if (tuner.e1) (void)hipEventDestroy(tuner.e1);
        if (pairExec) (void)hipGraphExecDestroy(pairExec);
        if (evAdjDone) (void)hipEventDestroy(evAdjDone);
        if (evRecDone) (void)hipEventDestroy(evRecDone);
        if (adjStream) (void)hipStreamDestroy(adjStream);
        if (ownStream && stream) (void)hipStreamDestroy(stream);
    }
    void releasePinned() {
        if (hState) (void)hipHostFree(hState);
        hState = nullptr;
        hFlags = nullptr;
        if (hReplayFlags) (void)hipHostFree(hReplayFlags);
        hReplayFlags = nullptr;
        replayFlagCap = 0;
        if (hQueueFlags) (void)hipHostFree(hQueueFlags);
        hQueueFlags = nullptr;
        queueFlagCap = 0;
    }

    // ------------------------------------------------------------ scenes
    int addScene(const sph_scene_params& P, int64_t nf, const float* pos, const float* vel, int64_t nb,
                 const float* bpos) {
        if (initialized) throw SphError{SPH_E_STATE, "scenes must be added before sph_init"};
        if (nf < 0 || nb < 0 || (nf > 0 && !pos) || (nb > 0 && !bpos))
            throw SphError{SPH_E_ARG, "bad particle arrays"};
        if (nf > (1ll << 30)) throw SphError{SPH_E_ARG, "too many fluid particles"};
        if (P.simulation_method < 0 || P.simulation_method > 2)
            throw SphError{SPH_E_ARG, "simulation_method must be 0, 1 or 2"};
        if (P.boundary_method != SPH_BOUNDARY_AKINCI && P.boundary_method != SPH_BOUNDARY_PCISPH &&
            !(P.boundary_method == SPH_BOUNDARY_NONE) && !(nb == 0 && (P.boundary_method == 0)))
            throw SphError{SPH_E_UNSUPPORTED, "only Akinci and PCISPH boundaries are supported"};
        if (P.viscosity_method > 2 || P.viscosity_method < -1)
            throw SphError{SPH_E_ARG, "viscosity_method must be -1, 0, 1 or 2"};
        if (!(P.particle_radius > 0.0f)) throw SphError{SPH_E_ARG, "particle_radius must be > 0"};
        if (!scenes.empty() && scenes[0].P.simulation_method != P.simulation_method)
            throw SphError{SPH_E_UNSUPPORTED, "all scenes of a context must use the same simulation method"};
        SceneH s;
        s.P = P;
        s.nf = static_cast<uint32_t>(nf);
        s.na = s.nf;
        s.nb = static_cast<uint32_t>(nb);
        s.xf.assign(pos, pos + 3 * nf);
        s.vf.assign(3 * nf, 0.0f);
        if (vel) std::copy(vel, vel + 3 * nf, s.vf.begin());
        if (nb) s.xb.assign(bpos, bpos + 3 * nb);
        for (float f : s.xf)
            if (!std::isfinite(f)) throw SphError{SPH_E_NONFINITE, "non-finite fluid position"};
        for (float f : s.xb)
            if (!std::isfinite(f)) throw SphError{SPH_E_NONFINITE, "non-finite boundary position"};
        derive(s);
        scenes.push_back(std::move(s));
        method = P.simulation_method;
        return static_cast<int>(scenes.size()) - 1;
    }

    // FluidObject::addEmitter (FluidObject.cpp:177-193) with the emission
    // group of Emitter::buildSquareGroup / buildCircleGroup (Emitter.cpp:25-70),
    // in the reference's FP32 (and, where it converts, FP64) arithmetic.
    void addEmitter(int si, const sph_emitter_params& E) {
        if (initialized) throw SphError{SPH_E_STATE, "emitters must be added before sph_init"};
        SceneH& s = scene(si);
        if (E.type != 0 && E.type != 1) throw SphError{SPH_E_ARG, "emitter type must be 0 (square) or 1 (circle)"};
        if (E.num_particles < 0 || static_cast<int64_t>(s.nf) + E.num_particles > (1ll << 30))
            throw SphError{SPH_E_ARG, "emitter particle budget out of range"};
        for (float f : {E.position[0], E.position[1], E.position[2], E.rotation[0], E.rotation[1], E.rotation[2],
                        E.rotation[3], E.velocity, E.start_time, E.width, E.height, E.spacing})
            if (!std::isfinite(f)) throw SphError{SPH_E_NONFINITE, "non-finite emitter parameter"};
        SceneH::Emitter e;
        e.P = E;
        e.remaining = E.num_particles;
        e.next = E.start_time;
        const float dist = 2.0f * s.r;
        if (E.type == 1) {
            // disc: the centre, then rings at (i + 1) dist holding floor(2 pi R / dist) particles
            e.group.insert(e.group.end(), {0.0f, 0.0f, 0.0f});
            const uint32_t levels = static_cast<uint32_t>(std::floor((E.width - 0.5f * dist) / dist));
            const float twoPi = static_cast<float>(2.0 * M_PI);
            for (uint32_t i = 0; i < levels; ++i) {
                const float radius = static_cast<float>(i + 1) * dist;
                const float len = twoPi * radius;
                const uint32_t parts = static_cast<uint32_t>(std::floor(len / dist));
                const float inc = twoPi / static_cast<float>(parts);
                for (uint32_t k = 0; k < parts; ++k) {
                    const float ang = static_cast<float>(k) * inc;
                    e.group.push_back(static_cast<float>(radius * std::sin(static_cast<double>(ang))));
                    e.group.push_back(static_cast<float>(radius * std::cos(static_cast<double>(ang))));
                    e.group.push_back(0.0f);
                }
            }
        } else {
            // rectangle of numX x numY particles centred on the emitter
            const uint32_t numX = static_cast<uint32_t>(std::floor(E.width / dist));
            const uint32_t numY = static_cast<uint32_t>(E.height / dist);
            const float realW = static_cast<float>(numX - 1) * dist, realH = static_cast<float>(numY - 1) * dist;
            for (uint32_t i = 0; i < numX; ++i)
                for (uint32_t j = 0; j < numY; ++j) {
                    e.group.push_back(static_cast<float>(realW * 0.5 - static_cast<double>(static_cast<float>(i) * dist)));
                    e.group.push_back(static_cast<float>(realH * 0.5 - static_cast<double>(static_cast<float>(j) * dist)));
                    e.group.push_back(0.0f);
                }
        }
        s.nf += static_cast<uint32_t>(E.num_particles);
        s.xf.resize(3 * static_cast<size_t>(s.nf), 0.0f);
        s.vf.resize(3 * static_cast<size_t>(s.nf), 0.0f);
        s.emitters.push_back(std::move(e));
    }

    // A CubeBoundaryObject (CubeBoundaryObject.cpp, SPHSimulation.cpp:457-474):
    // the fluid is clamped into [lo, hi] at the end of every step.
    void setBoundaryNormals(int si, int64_t nb, const float* normals, float normalFct, float tangFct) {
        if (initialized) throw SphError{SPH_E_STATE, "boundary normals must be set before sph_init"};
        SceneH& s = scene(si);
        if (s.P.boundary_method != SPH_BOUNDARY_PCISPH)
            throw SphError{SPH_E_ARG, "boundary normals belong to boundary method SPH_BOUNDARY_PCISPH"};
        if (nb != static_cast<int64_t>(s.nb) || (nb > 0 && !normals))
            throw SphError{SPH_E_ARG, "one normal per boundary particle"};
        if (!std::isfinite(normalFct) || !std::isfinite(tangFct)) throw SphError{SPH_E_NONFINITE, "non-finite factor"};
        s.bn.assign(normals, normals + 3 * nb);
        for (float q : s.bn)
            if (!std::isfinite(q)) throw SphError{SPH_E_NONFINITE, "non-finite boundary normal"};
        s.nFct = normalFct;
        s.tFct = tangFct;
        s.normalsSet = true;
    }
    void getBoundaryState(int si, float* density, float* pressure) {
        if (!initialized) throw SphError{SPH_E_STATE, "sph_init not called"};
        SceneH& s = scene(si);
        if (!s.pciB) throw SphError{SPH_E_ARG, "the scene has no PCISPH boundary object"};
        HIPCHK(hipStreamSynchronize(stream));
        std::vector<float> r(s.nb), q(s.nb);
        if (s.nb) {
            HIPCHK(hipMemcpy(r.data(), bRho.p + s.bBegin, s.nb * sizeof(float), hipMemcpyDeviceToHost));
            HIPCHK(hipMemcpy(q.data(), bP.p + s.bBegin, s.nb * sizeof(float), hipMemcpyDeviceToHost));
        }
        for (uint32_t k = 0; k < s.nb; ++k) {
            if (density) density[s.bOrder[k]] = r[k];
            if (pressure) pressure[s.bOrder[k]] = q[k];
        }
    }

    void addCube(int si, const float lo[3], const float hi[3], float normalFct, float tangFct) {
        if (initialized) throw SphError{SPH_E_STATE, "cube boundaries must be added before sph_init"};
        SceneH& s = scene(si);
        if (s.P.boundary_method != 0) throw SphError{SPH_E_ARG, "cube boundaries need boundary method 0"};
        if (s.cubes.size() >= static_cast<size_t>(kMaxCubes))
            throw SphError{SPH_E_UNSUPPORTED, "at most 4 cube boundary objects per scene"};
        SceneH::Cube c;
        for (int a = 0; a < 3; ++a) {
            if (!std::isfinite(lo[a]) || !std::isfinite(hi[a])) throw SphError{SPH_E_NONFINITE, "non-finite cube"};
            c.lo[a] = lo[a];
            c.hi[a] = hi[a];
        }
        c.normal = normalFct;
        c.tang = tangFct;
        s.cubes.push_back(c);
    }
    bool anyCube() const {
        for (auto& s : scenes) if (!s.cubes.empty()) return true;
        return false;
    }
    bool anyAdh() const {
        for (auto& s : scenes) if (s.adhOn) return true;
        return false;
    }

    // Rotation of v by the quaternion q = (w, x, y, z), as the reference applies
    // its Quat4r to a point: v + 2 w (u x v) + 2 u x (u x v) with u = (x, y, z).
    static void rotate(const float q[4], float v[3]) {
        const float w = q[0], ux = q[1], uy = q[2], uz = q[3];
        const float cx = uy * v[2] - uz * v[1], cy = uz * v[0] - ux * v[2], cz = ux * v[1] - uy * v[0];
        const float dx = uy * cz - uz * cy, dy = uz * cx - ux * cz, dz = ux * cy - uy * cx;
        v[0] = v[0] + (cx * w + dx) * 2.0f;
        v[1] = v[1] + (cy * w + dy) * 2.0f;
        v[2] = v[2] + (cz * w + dz) * 2.0f;
    }

    // SPHSimulation::emitParticles -> Emitter::emitParticles / emitGroup
    // (Emitter.cpp:72-113) at the scene time t_n of the current step: every
    // emitter whose next time has come and whose budget still holds a whole
    // group places it in the scene's first free slots (the dead-key padding at
    // the end of its sorted segment). dfsph marks the new particles so the
    // keys pass skips their position update (the reference emits after
    // updatePositions). Returns whether any scene grew.
    bool emitAll(bool dfsph) {
        bool any = false;
        for (size_t si = 0; si < scenes.size(); ++si) {
            SceneH& s = scenes[si];
            if (s.emitters.empty()) continue;
            const float t = hState[si].t;
            const float dist = 2.0f * s.r;
            std::vector<float4> nx, nv;
            for (auto& e : s.emitters) {
                if (!(t >= e.next)) continue;
                const int64_t g = static_cast<int64_t>(e.group.size() / 3);
                if (g > e.remaining) continue;
                const float off = t - e.next;
                for (int64_t k = 0; k < g; ++k) {
                    float vel[3] = {0.0f, 0.0f, e.P.velocity};
                    float pos[3];
                    for (int a = 0; a < 3; ++a) pos[a] = e.group[3 * k + a] + off * (vel[a] + s.P.gravity[a] * off);
                    rotate(e.P.rotation, pos);
                    rotate(e.P.rotation, vel);
                    for (int a = 0; a < 3; ++a) pos[a] += e.P.position[a];
                    nx.push_back(make_float4(pos[0], pos[1], pos[2], dfsph ? -1.0f : 0.0f));   // w < 0: emitted this step
                    nv.push_back(make_float4(vel[0], vel[1], vel[2], 0.0f));
                }
                e.remaining -= g;
                e.next += e.P.spacing * 1.105f * dist / e.P.velocity;
            }
            if (nx.empty()) continue;
            const uint32_t cnt = static_cast<uint32_t>(nx.size());
            if (s.na + cnt > s.nf) throw SphError{SPH_E_STATE, "emitter budget exceeded"};
            HIPCHK(hipStreamSynchronize(stream));
            const uint32_t slot = s.fBegin + s.na;
            std::vector<uint32_t> nid(cnt);
            for (uint32_t k = 0; k < cnt; ++k) nid[k] = s.na + k;
            const std::vector<float> zero(cnt, 0.0f);
            HIPCHK(hipMemcpy(x.p + slot, nx.data(), cnt * sizeof(float4), hipMemcpyHostToDevice));
            uploadVel(slot, nv);
            HIPCHK(hipMemcpy(id.p + slot, nid.data(), cnt * sizeof(uint32_t), hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(kws.p + slot, zero.data(), cnt * sizeof(float), hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(kvws.p + slot, zero.data(), cnt * sizeof(float), hipMemcpyHostToDevice));
            s.na += cnt;
            s.fEnd = s.fBegin + s.na;
            any = true;
        }
        if (any) uploadConsts();
        return any;
    }

    // FP32 constants in the reference's evaluation order
    static void derive(SceneH& s) {
        const sph_scene_params& P = s.P;
        s.r = P.particle_radius;
        s.h = 4.0f * s.r;  // SPHSimulation.cpp:206
        s.invh = 1.0f / s.h;
        s.rho0 = P.density0;
        const float diam = 2.0f * s.r;
        s.m = s.rho0 * (diam * diam * diam);  // SPHSimulation.cpp:145-153
        const float pi = static_cast<float>(M_PI);
        const float h = s.h;
        const float h3 = h * h * h;
        s.kW = 8.0f / (pi * h3);
        s.lW = 48.0f / (pi * h3);
        const float h6 = h * h * h * h * h * h;
        s.laplC = 45.0f / (pi * h6);
        const float h9 = h * h * h * h * h * h * h * h * h;
        s.cohC1 = static_cast<float>(32.0 / M_PI) / h9;
        s.cohC2 = h6 / 64.0f;
        const float fpsTs = 1.0f / P.fps;  // TimeHelper.h:57
        s.minTs = fpsTs < P.min_time_step ? fpsTs : P.min_time_step;
        s.maxTs = P.max_time_step;
        s.dt = P.time_step;
        s.t = 0.0f;
        // SPHSimulation.cpp:390-394
        s.viscOn = P.viscosity_method > -1 && P.viscosity != 0.0f;
        s.stOn = P.surface_tension_method == SPH_SURFACE_TENSION_AKINCI && P.surface_tension != 0.0f;
        s.akinci = P.boundary_method == SPH_BOUNDARY_AKINCI && s.nb > 0;
        s.pciB = P.boundary_method == SPH_BOUNDARY_PCISPH && s.nb > 0;
        // FluidObject::setAdhesionForce creates the force for the Akinci method (4) only
        s.adhOn = P.adhesion_method == 4 && P.adhesion != 0.0f && s.akinci;
        if (P.simulation_method == SPH_PCISPH) {  // PCISPHSolver.cpp:403-436 with the initial dt
            const float dist = 2.0f * s.r;
            float sg[3] = {0, 0, 0}, ssq = 0;
            for (int i = -1; i <= 1; i++)
                for (int j = -1; j <= 1; j++)
                    for (int k = -1; k <= 1; k++) {
                        float rij[3] = {static_cast<float>(-i) * dist, static_cast<float>(-j) * dist,
                                        static_cast<float>(-k) * dist};
                        float g[3];
                        hostGradW(s.h, s.lW, rij, g);
                        for (int a = 0; a < 3; ++a) sg[a] += g[a];
                        ssq += (g[0] * g[0] + g[1] * g[1]) + g[2] * g[2];
                    }
            const float mdt = s.m * s.dt / s.rho0;
            const float beta = 2.0f * (mdt * mdt);
            const float dsg = (sg[0] * sg[0] + sg[1] * sg[1]) + sg[2] * sg[2];
            s.pciDelta = -1.0f / (beta * (-dsg - ssq));
        }
    }

    // ------------------------------------------------------------ layout
    static void aabb(const std::vector<float>& p, float lo[3], float hi[3]) {
        for (size_t i = 0; i + 2 < p.size(); i += 3)
            for (int a = 0; a < 3; ++a) {
                lo[a] = std::min(lo[a], p[i + a]);
                hi[a] = std::max(hi[a], p[i + a]);
            }
    }
    void setGrid(SceneH& s, const float lo[3], const float hi[3]) {
        for (int a = 0; a < 3; ++a) {
            const double clo = std::floor(static_cast<double>(lo[a] * s.invh));
            const double chi = std::floor(static_cast<double>(hi[a] * s.invh));
            const double ext = chi - clo + 1;
            const double margin = 4 + std::ceil(0.1 * ext);
            const double g0 = clo - margin, gn = ext + 2 * margin;
            if (g0 < -1.0e9 || g0 + gn > 1.0e9 || gn > (1 << 20))
                throw SphError{SPH_E_OUT_OF_GRID, "particles spread over too many grid cells"};
            s.g0[a] = static_cast<int>(g0);
            s.gn[a] = static_cast<int>(gn);
        }
        const double cells = static_cast<double>(s.gn[0]) * s.gn[1] * s.gn[2];
        if (cells > static_cast<double>(1u << 28)) throw SphError{SPH_E_OUT_OF_GRID, "grid too large (> 2^28 cells)"};
        s.nCells = static_cast<uint32_t>(cells);
        const int maxDim = std::max(s.gn[0], std::max(s.gn[1], s.gn[2]));
        int bits = 0;
        while ((1 << bits) < maxDim) ++bits;
        if (bits <= 10 && 3 * bits <= 30 && (1u << (3 * bits)) <= 4 * s.nCells + 1024) {
            s.morton = 1;
            s.keyRange = (1u << (3 * bits)) + 1;
        } else {
            s.morton = 0;
            s.keyRange = s.nCells + 1;
        }
        s.keyRange += s.ovBuckets;  // dense keys, overflow buckets, dead key
    }
    void cellOfHost(const SceneH& s, const float* p, int c[3]) const {
        for (int a = 0; a < 3; ++a) c[a] = static_cast<int>(std::floor(p[a] * s.invh)) - s.g0[a];
    }
    static uint32_t spread3h(uint32_t v) {
        v &= 0x3ffu;
        v = (v | (v << 16)) & 0x030000ffu;
        v = (v | (v << 8)) & 0x0300f00fu;
        v = (v | (v << 4)) & 0x030c30c3u;
        v = (v | (v << 2)) & 0x09249249u;
        return v;
    }

    // Assign cell/key bases, upload scene constants and the boundary grid.
    void uploadLayout() {
        uint64_t cells = 0, keysTot = 0;
        for (auto& s : scenes) {
            s.cellBase = static_cast<uint32_t>(cells);
            s.keyBase = static_cast<uint32_t>(keysTot);
            cells += s.nCells + s.ovBuckets;
            keysTot += s.keyRange;
            if (cells > (1ull << 30) || keysTot > (1ull << 32) - 1)
                throw SphError{SPH_E_OUT_OF_GRID, "combined grids too large"};
        }
        totalCells = cells;
        keyBits = 1;
        while (keyBits < 32 && (1ull << keyBits) < keysTot) ++keyBits;
        uploadOrderTables();
        uploadConsts();
        uploadBoundaryGrid();
    }

    // Counting-sort tables of the layout: each scene's order indices are its
    // dense cells in key order (Morton or linear), then its overflow buckets,
    // then one entry for its padding slots.
    void uploadOrderTables() {
        std::vector<uint32_t> rank(totalCells ? totalCells : 1, 0u);
        std::vector<uint32_t> o2c;
        uint32_t ob = 0;
        for (auto& sc : scenes) {
            sc.ordBase = ob;
            std::vector<uint32_t> lins(sc.nCells);
            std::iota(lins.begin(), lins.end(), 0u);
            if (sc.morton) {
                std::vector<uint64_t> kl(sc.nCells);
                const uint32_t nx = static_cast<uint32_t>(sc.gn[0]), ny = static_cast<uint32_t>(sc.gn[1]);
                for (uint32_t l = 0; l < sc.nCells; ++l) {
                    const uint32_t cx = l % nx, cy = (l / nx) % ny, cz = l / (nx * ny);
                    const uint64_t key = spread3h(cx) | (spread3h(cy) << 1) | (spread3h(cz) << 2);
                    kl[l] = (key << 32) | l;
                }
                std::sort(kl.begin(), kl.end());
                for (uint32_t r = 0; r < sc.nCells; ++r) lins[r] = static_cast<uint32_t>(kl[r] & 0xffffffffu);
            }
            for (uint32_t r = 0; r < sc.nCells; ++r) {
                rank[sc.cellBase + lins[r]] = r;
                o2c.push_back(sc.cellBase + lins[r]);
            }
            for (uint32_t b = 0; b < sc.ovBuckets; ++b) {
                rank[sc.cellBase + sc.nCells + b] = sc.nCells + b;
                o2c.push_back(sc.cellBase + sc.nCells + b);
            }
            o2c.push_back(~0u);   // padding slots
            ob += sc.nCells + sc.ovBuckets + 1;
        }
        ordTotal = ob;
        rankOf.alloc(rank.size());
        HIPCHK(hipMemcpy(rankOf.p, rank.data(), rank.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
        ordToCell.alloc(o2c.size() ? o2c.size() : 1);
        if (!o2c.empty())
            HIPCHK(hipMemcpy(ordToCell.p, o2c.data(), o2c.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
        counts.alloc(static_cast<size_t>(ordTotal) + 1);
        starts.alloc(static_cast<size_t>(ordTotal) + 1);
        occ.alloc(static_cast<size_t>(ordTotal) + 1);
        rangesFull = true;
        HIPCHK(hipMemset(counts.p, 0, (static_cast<size_t>(ordTotal) + 1) * sizeof(uint32_t)));
        scanTempBytes = 0;
        HIPCHK(scanCounts(nullptr, scanTempBytes, counts.p, starts.p, ordTotal + 1, stream));
        scanTemp.alloc(scanTempBytes + 256);
    }

    // Per-scene constants (grid, layout, coefficients) to the device.
    void uploadConsts() {
        std::vector<SceneConst> sc(scenes.size());
        for (size_t i = 0; i < scenes.size(); ++i) {
            const SceneH& s = scenes[i];
            SceneConst& c = sc[i];
            std::memset(&c, 0, sizeof(c));
            c.h = s.h;
            c.invh = s.invh;
            c.g0x = s.g0[0]; c.g0y = s.g0[1]; c.g0z = s.g0[2];
            c.gnx = s.gn[0]; c.gny = s.gn[1]; c.gnz = s.gn[2];
            c.cellBase = s.cellBase;
            c.keyBase = s.keyBase;
            c.deadKey = s.keyBase + s.keyRange - 1;
            c.morton = s.morton;
            c.fBegin = s.fBegin; c.fEnd = s.fEnd; c.fPadEnd = s.fPadEnd;
            c.blockBegin = s.blockBegin; c.blockEnd = s.blockEnd;
            c.bBegin = s.bBegin; c.bEnd = s.bEnd;
            c.m = s.m; c.rho0 = s.rho0;
            c.gx = s.P.gravity[0]; c.gy = s.P.gravity[1]; c.gz = s.P.gravity[2];
            c.kW = s.kW; c.lW = s.lW; c.laplC = s.laplC; c.cohC1 = s.cohC1; c.cohC2 = s.cohC2;
            {
                const float h3 = s.h * s.h * s.h;
                c.gA = 3.0f * s.lW / h3;
                c.gB = -2.0f * s.lW / (s.h * s.h);
                c.gC = -(s.lW / h3);
            }
            c.visc = s.P.viscosity; c.bvisc = s.P.boundary_viscosity; c.st = s.P.surface_tension;
            c.viscMethod = s.viscOn ? s.P.viscosity_method : -1;
            c.stOn = s.stOn ? 1 : 0;
            // boundary particles in the densities and neighbour lists: Akinci ones
            // and PCISPH boundary objects alike (their w carries the weight)
            c.akinci = (s.akinci || s.pciB) ? 1 : 0;
            c.bviscOn = (s.viscOn && s.P.boundary_viscosity != 0.0f && s.akinci) ? 1 : 0;
            c.pciB = s.pciB ? 1 : 0;
            c.nFct = s.nFct;
            c.tFct = s.tFct;
            c.restR = 0.5f * s.h;   // PCISPHBoundaryObject.cpp:224
            {   // Poly6.h:20-27, Spiky.h:20-27
                const float h = s.h;
                const float h6 = h * h * h * h * h * h, h9 = h6 * h * h * h;
                c.poly6C = 315.0f / (static_cast<float>(64.0 * M_PI) * h9);
                c.spikyC = 45.0f / (static_cast<float>(M_PI) * h6);
                c.h2 = h * h;
            }
            c.gamma = s.P.gamma;
            c.wcB = s.P.gamma != 0.0f ? s.P.stiffness * s.rho0 / s.P.gamma : 0.0f;
            c.pciDelta = s.pciDelta;
            c.eps = 1.0e-5f;
            c.h2c = 0.01f * (s.h * s.h);
            c.dmax = sqrtThreshold(s.h);
            c.dhalf = sqrtThreshold(0.5f * s.h);
            c.eta = s.P.eta;
            c.etaV = s.P.eta_v;
            c.cfl = s.P.cfl_factor;
            c.radius = s.r;
            c.minTs = s.minTs;
            c.maxTs = s.maxTs;
            c.minIt = static_cast<int32_t>(std::min<uint32_t>(s.P.min_iterations, 1u << 30));
            c.maxIt = static_cast<int32_t>(std::min<uint32_t>(s.P.max_iterations, 1u << 30));
            c.minItV = static_cast<int32_t>(std::min<uint32_t>(s.P.min_iterations_v, 1u << 30));
            c.adaptive = (s.P.simulation_method == SPH_DFSPH && s.P.adaptive_time_step) ? 1 : 0;
            c.nf = static_cast<int32_t>(s.na);
            c.ovMask = s.ovBuckets - 1;
            c.nCells = s.nCells;
            c.ordBase = s.ordBase;
            c.adhOn = s.adhOn ? 1 : 0;
            c.adh = s.P.adhesion;
            c.nCubes = static_cast<int32_t>(s.cubes.size());
            for (size_t k = 0; k < s.cubes.size(); ++k) {
                for (int a = 0; a < 3; ++a) {
                    c.cubeLo[k][a] = s.cubes[k].lo[a];
                    c.cubeHi[k][a] = s.cubes[k].hi[a];
                }
                c.cubeNormal[k] = s.cubes[k].normal;
                c.cubeTang[k] = s.cubes[k].tang;
            }
        }
        if (dScenes.n != scenes.size()) dScenes.alloc(scenes.size());
        HIPCHK(hipStreamSynchronize(stream));
        HIPCHK(hipMemcpy(dScenes.p, sc.data(), sc.size() * sizeof(SceneConst), hipMemcpyHostToDevice));
    }

    void uploadBoundaryGrid() {
        // boundary grid: sort boundary particles of each scene by cell key on the host
        std::vector<uint2> br(totalCells ? totalCells : 1, make_uint2(0u, 0u));
        std::vector<float4> xbs(NB ? NB : 1);
        std::vector<float4> bns(anyPciB() ? xbs.size() : 0);
        for (auto& s : scenes) {
            s.bOrder.resize(s.nb);
            std::vector<uint32_t> key(s.nb), lin(s.nb);
            for (uint32_t b = 0; b < s.nb; ++b) {
                int c[3];
                cellOfHost(s, &s.xb[3 * b], c);
                for (int a = 0; a < 3; ++a)
                    if (c[a] < 0 || c[a] >= s.gn[a]) throw SphError{SPH_E_OUT_OF_GRID, "boundary particle outside grid"};
                lin[b] = static_cast<uint32_t>((c[2] * s.gn[1] + c[1]) * s.gn[0] + c[0]);
                key[b] = s.morton ? (spread3h(c[0]) | (spread3h(c[1]) << 1) | (spread3h(c[2]) << 2)) : lin[b];
            }
            std::iota(s.bOrder.begin(), s.bOrder.end(), 0u);
            std::stable_sort(s.bOrder.begin(), s.bOrder.end(), [&](uint32_t A, uint32_t B) { return key[A] < key[B]; });
            for (uint32_t q = 0; q < s.nb; ++q) {
                const uint32_t o = s.bOrder[q];
                const uint32_t g = s.bBegin + q;
                // a PCISPH boundary particle weighs its mass density0 (2r)^3
                // (SPHSimulation.cpp:434-436); Akinci ones get rho0 V_b below
                xbs[g] = make_float4(s.xb[3 * o], s.xb[3 * o + 1], s.xb[3 * o + 2], s.pciB ? s.m : 0.0f);
                if (s.pciB) bns[g] = make_float4(s.bn[3 * o], s.bn[3 * o + 1], s.bn[3 * o + 2], 0.0f);
                uint2& r = br[s.cellBase + lin[o]];
                if (r.y == 0) r.x = g;
                r.y = g + 1;
            }
        }
        bRange.alloc(br.size());
        HIPCHK(hipMemcpy(bRange.p, br.data(), br.size() * sizeof(uint2), hipMemcpyHostToDevice));
        xb.alloc(xbs.size());
        HIPCHK(hipMemcpy(xb.p, xbs.data(), xbs.size() * sizeof(float4), hipMemcpyHostToDevice));
        vb.alloc(xbs.size());
        if (anyPciB()) {
            bn.alloc(bns.size());
            HIPCHK(hipMemcpy(bn.p, bns.data(), bns.size() * sizeof(float4), hipMemcpyHostToDevice));
            bRho.alloc(xbs.size());
            bP.alloc(xbs.size());
            HIPCHK(hipMemset(bP.p, 0, xbs.size() * sizeof(float)));
            layoutCellMarks(br);
        }
        fRange.alloc(totalCells ? totalCells : 1);
        rangesDirty = true;
        rangesFull = true;
        // Akinci volumes (AkinciBoundaryObject::computeVolume), once per layout
        DevView d = view();
        for (size_t i = 0; i < scenes.size(); ++i)
            if (scenes[i].akinci) HIPCHK(launchBoundaryVolume(d, static_cast<int>(i), scenes[i].bBegin, scenes[i].bEnd, stream));
    }

    bool anyPciB() const {
        for (auto& s : scenes)
            if (s.pciB) return true;
        return false;
    }
    // The cells the reference's hash table holds (a cell exists from the first
    // search that inserted a particle into it, or from the search after a
    // collision correction looked it up): per cell slot the first step from
    // which it exists. Boundary cells exist from the start
    // (SPHSolver::insertBoundaryParticles); on a new dense box the marks of
    // the old one move to the same global cells; overflow buckets count as
    // existing (no boundary particle is within their reach).
    void layoutCellMarks(const std::vector<uint2>& br) {
        std::vector<uint32_t> old;
        if (cellFirst.n) {
            old.resize(cellFirst.n);
            HIPCHK(hipMemcpy(old.data(), cellFirst.p, old.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
        }
        std::vector<uint32_t> m(totalCells ? totalCells : 1, 0u);
        for (auto& s : scenes) {
            if (!s.pciB) continue;
            for (uint32_t c = 0; c < s.nCells; ++c) m[s.cellBase + c] = br[s.cellBase + c].y > 0 ? 0u : ~0u;
            if (s.markedLayout) {
                const uint32_t oldCells = static_cast<uint32_t>(s.markGn[0]) * s.markGn[1] * s.markGn[2];
                for (uint32_t c = 0; c < oldCells && s.markBase + c < old.size(); ++c) {
                    const uint32_t f = old[s.markBase + c];
                    if (f == ~0u) continue;
                    const int l[3] = {static_cast<int>(c % s.markGn[0]), static_cast<int>((c / s.markGn[0]) % s.markGn[1]),
                                      static_cast<int>(c / (static_cast<uint32_t>(s.markGn[0]) * s.markGn[1]))};
                    int q[3];
                    bool in = true;
                    for (int a = 0; a < 3; ++a) {
                        q[a] = l[a] + s.markG0[a] - s.g0[a];
                        in = in && q[a] >= 0 && q[a] < s.gn[a];
                    }
                    if (!in) continue;
                    uint32_t& t = m[s.cellBase + static_cast<uint32_t>((q[2] * s.gn[1] + q[1]) * s.gn[0] + q[0])];
                    t = std::min(t, f);
                }
            }
            s.markedLayout = true;
            for (int a = 0; a < 3; ++a) { s.markG0[a] = s.g0[a]; s.markGn[a] = s.gn[a]; }
            s.markBase = s.cellBase;
        }
        cellFirst.alloc(m.size());
        HIPCHK(hipMemcpy(cellFirst.p, m.data(), m.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    }

    void allocNeighborLists() {
        nbrF.alloc(static_cast<size_t>(capF) * NP);
        if (tiles) nbrT.alloc(static_cast<size_t>(capF / 2) * NP);
        nbrB.alloc(static_cast<size_t>(capB) * NP);
    }

    DevView view() {
        DevView d{};
        d.scenes = dScenes.p;
        d.state = dState.p;
        d.blockScene = dBlockScene.p;
        d.numScenes = static_cast<int32_t>(scenes.size());
        d.numBlocks = numBlocks;
        d.NP = static_cast<uint32_t>(NP);
        d.x = x.p; d.v = v.p; d.a = a.p; d.n = n.p; d.xs = xs.p; d.ap = ap.p; d.xk = xk.p;
        d.rho = rho.p; d.alpha = alpha.p; d.p = p.p; d.k = k.p; d.kws = kws.p; d.kvws = kvws.p; d.aux = aux.p;
        d.id = id.p;
        d.xb = xb.p; d.vb = vb.p;
        d.bn = bn.p; d.bRho = bRho.p; d.bP = bP.p; d.vs = vs.p; d.xsP = xsP.p; d.vsP = vsP.p;
        d.cellFirst = cellFirst.p;
        d.stepNo = stepNo;
        d.markNo = markNo;
        d.nc = nc.p;
        d.ncAlt = ncAlt.p;
        d.fRange = fRange.p; d.bRange = bRange.p; d.cellOf = cellOf.p;
        d.rankOf = rankOf.p; d.ordToCell = ordToCell.p; d.counts = counts.p; d.starts = starts.p; d.li = li.p;
        d.snapV = snapV.p; d.snapKvws = snapKvws.p;
        d.vRec = nullptr;
        d.nbrF = nbrF.p; d.cntF = cntF.p; d.nbrB = nbrB.p; d.cntB = cntB.p;
        d.capF = capF; d.capB = capB;
        d.tileSlot = tiles ? tileSlot.p : nullptr;
        d.tileCount = tiles ? tileCount.p : nullptr;
        d.nbrT = tiles ? nbrT.p : nullptr;
        d.partials = partials.p; d.flags = dFlags;
        return d;
    }

    // ------------------------------------------------------------ init
    void init() {
        if (initialized) throw SphError{SPH_E_STATE, "sph_init called twice"};
        if (scenes.empty()) throw SphError{SPH_E_STATE, "no scene added"};
        for (auto& s : scenes)
            if (s.pciB && !s.normalsSet)
                throw SphError{SPH_E_STATE, "a PCISPH boundary scene needs its normals (sph_set_boundary_normals)"};
        HIPCHK(hipSetDevice(device));
        if (!stream) {
            HIPCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
            ownStream = true;
        }
        HIPCHK(scratchAddress(device, &noAllocScratch));
        HIPCHK(hipEventCreate(&evStart));
        HIPCHK(hipEventCreate(&evStop));
        HIPCHK(hipEventCreateWithFlags(&evSync, hipEventDisableTiming));
        // slots; the Jacobi passes address particle records with 32-bit byte
        // offsets, so the padded total must stay below 2^28 (checked before
        // any 32-bit slot index is formed)
        {
            uint64_t total = 0, btotal = 0;
            for (auto& s : scenes) {
                total += (static_cast<uint64_t>(s.nf) + kBlock - 1) / kBlock * kBlock;
                btotal += s.nb;
            }
            if (total >= (static_cast<uint64_t>(1) << 28))
                throw SphError{SPH_E_UNSUPPORTED, "more than 2^28 particle slots in one context"};
            if (btotal >= (static_cast<uint64_t>(1) << 31))
                throw SphError{SPH_E_UNSUPPORTED, "more than 2^31 boundary particles in one context"};
        }
        uint32_t off = 0, boff = 0;
        for (auto& s : scenes) {
            s.fBegin = off;
            s.fEnd = off + s.na;
            s.fPadEnd = off + ((s.nf + kBlock - 1) / kBlock) * kBlock;
            s.blockBegin = s.fBegin / kBlock;
            s.blockEnd = s.fPadEnd / kBlock;
            off = s.fPadEnd;
            s.bBegin = boff;
            s.bEnd = boff + s.nb;
            boff += s.nb;
            float lo[3] = {1e30f, 1e30f, 1e30f}, hi[3] = {-1e30f, -1e30f, -1e30f};
            aabb(std::vector<float>(s.xf.begin(), s.xf.begin() + 3 * static_cast<size_t>(s.na)), lo, hi);
            aabb(s.xb, lo, hi);
            for (auto& e : s.emitters) aabb(std::vector<float>(e.P.position, e.P.position + 3), lo, hi);
            for (auto& c : s.cubes) {
                aabb(std::vector<float>(c.lo, c.lo + 3), lo, hi);
                aabb(std::vector<float>(c.hi, c.hi + 3), lo, hi);
            }
            if (s.na == 0 && s.nb == 0 && s.emitters.empty() && s.cubes.empty()) {
                lo[0] = lo[1] = lo[2] = 0;
                hi[0] = hi[1] = hi[2] = 0;
            }
            setGrid(s, lo, hi);
        }
        NP = off;
        NB = boff;
        numBlocks = static_cast<uint32_t>(NP / kBlock);
        const size_t np = NP ? NP : 1;
        // flags follow the scene states in the same allocation
        const size_t flagStates = (scenes.size() * kFlagSlots * sizeof(int32_t) + sizeof(SceneState) - 1) /
                                  sizeof(SceneState);
        dState.alloc(scenes.size() + flagStates);
        dFlags = reinterpret_cast<int32_t*>(dState.p + scenes.size());
        stateBlockBytes = scenes.size() * sizeof(SceneState) + scenes.size() * kFlagSlots * sizeof(int32_t);
        std::vector<int32_t> bs(numBlocks ? numBlocks : 1, 0);
        for (size_t i = 0; i < scenes.size(); ++i)
            for (uint32_t b = scenes[i].blockBegin; b < scenes[i].blockEnd; ++b) bs[b] = static_cast<int32_t>(i);
        dBlockScene.alloc(bs.size());
        HIPCHK(hipMemcpy(dBlockScene.p, bs.data(), bs.size() * sizeof(int32_t), hipMemcpyHostToDevice));
        for (auto* b : {&x, &xAlt, &v, &vAlt, &a, &n, &xs, &ap, &xk}) b->alloc(np);
        if (method == SPH_PCISPH && anyPciB())
            for (auto* b : {&vs, &xsP, &vsP}) {
                b->alloc(np);
                HIPCHK(hipMemset(b->p, 0, np * sizeof(float4)));
            }
        if (method == SPH_DFSPH && anyPciB())   // FluidObject's normals start at zero
            for (auto* b : {&nc, &ncAlt}) {
                b->alloc(np);
                HIPCHK(hipMemset(b->p, 0, np * sizeof(float4)));
            }
        stepNo = 0;
        markNo = 0;
        fuseFinish = sphKernelsFuseFinish() != 0;
        for (auto* b : {&rho, &alpha, &p, &k, &kws, &kwsAlt, &kvws, &kvwsAlt, &aux}) b->alloc(np);
        for (auto* b : {&id, &idAlt, &keys, &keysAlt, &vals, &valsAlt, &cntF, &cntB, &cellOf, &li}) b->alloc(np);
        HIPCHK(hipMemset(cellOf.p, 0xFF, np * sizeof(uint32_t)));
        tiles = method == SPH_DFSPH && sphKernelsTiles() != 0;
        if (tiles) {
            const size_t groups = np / kTileGroup + 1;
            tileSlot.alloc(groups * kTileMax);
            tileCount.alloc(groups);
        }
        allocNeighborLists();
        // the DFSPH source passes run sphKernelsLanesPerParticle() launch blocks per particle block
        partials.alloc(static_cast<size_t>(numBlocks ? numBlocks : 1) *
                       static_cast<size_t>(std::max(1, sphKernelsLanesPerParticle())));
        if (method == SPH_DFSPH) {
            snapV.alloc(np);
            snapKvws.alloc(np);
        }
        HIPCHK(hipMemset(a.p, 0, np * sizeof(float4)));
        HIPCHK(hipMemset(n.p, 0, np * sizeof(float4)));
        for (auto* b : {&rho, &alpha, &p, &k, &kws, &kvws, &aux}) HIPCHK(hipMemset(b->p, 0, np * sizeof(float)));
        HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&hState), std::max(stateBlockBytes, sizeof(SceneState))));
        std::memset(hState, 0, stateBlockBytes);
        hFlags = reinterpret_cast<int32_t*>(hState + scenes.size());
        HIPCHK(hipMemset(dFlags, 0, scenes.size() * kFlagSlots * sizeof(int32_t)));
        for (size_t i = 0; i < scenes.size(); ++i) {
            SceneState& T = hState[i];
            std::memset(&T, 0, sizeof(T));
            T.dt = scenes[i].dt;
            T.dt2 = scenes[i].dt * scenes[i].dt;
            T.t = 0.0f;
            T.active[2] = 1;
        }
        HIPCHK(hipMemcpy(dState.p, hState, scenes.size() * sizeof(SceneState), hipMemcpyHostToDevice));
        uploadState();
        // sort keys temp storage
        sortTempBytes = 0;
        HIPCHK(sortPairs(nullptr, sortTempBytes, keys.p, keysAlt.p, vals.p, valsAlt.p, static_cast<int>(NP), 32, stream,
                         nullptr));
        sortTemp.alloc(sortTempBytes + 256);
        uploadLayout();
        initialized = true;
        // solver init passes
        if (method == SPH_DFSPH && NP) {  // DFSPHSolver::init (DFSPHSolver.cpp:22-31)
            neighborSearch(false);
            timed(K_DENSITY, [&] { return launchDensityAlpha(view(), 1, stream); });
            boundaryDensities();
        }
        HIPCHK(hipStreamSynchronize(stream));
    }

    // Velocities of slots [first, first + count) in slot order.
    std::vector<float4> downloadVel(uint32_t first, uint32_t count) {
        std::vector<float4> out(count);
        HIPCHK(hipStreamSynchronize(stream));
        if (!count) return out;
        HIPCHK(hipMemcpy(out.data(), v.p + first, count * sizeof(float4), hipMemcpyDeviceToHost));
        return out;
    }
    void uploadVel(uint32_t first, const std::vector<float4>& vals) {
        if (vals.empty()) return;
        HIPCHK(hipMemcpy(v.p + first, vals.data(), vals.size() * sizeof(float4), hipMemcpyHostToDevice));
    }

    void uploadState() {
        const size_t np = NP ? NP : 1;
        std::vector<float4> hx(np, make_float4(0, 0, 0, 0)), hv(np, make_float4(0, 0, 0, 0));
        std::vector<uint32_t> hid(np, 0);
        for (auto& s : scenes)
            for (uint32_t i = 0; i < s.na; ++i) {
                const uint32_t slot = s.fBegin + i;
                hx[slot] = make_float4(s.xf[3 * i], s.xf[3 * i + 1], s.xf[3 * i + 2], 0.0f);
                hv[slot] = make_float4(s.vf[3 * i], s.vf[3 * i + 1], s.vf[3 * i + 2], 0.0f);
                hid[slot] = i;
            }
        // padding slots: park far away from nothing (never read as neighbours)
        HIPCHK(hipMemcpy(x.p, hx.data(), np * sizeof(float4), hipMemcpyHostToDevice));
        uploadVel(0, hv);
        HIPCHK(hipMemcpy(id.p, hid.data(), np * sizeof(uint32_t), hipMemcpyHostToDevice));
    }

    // NP float4 velocities from one buffer to another, on the engine's stream.
    void copyVel(float4* dst, const float4* src) {
        HIPCHK(hipMemcpyAsync(dst, src, NP * sizeof(float4), hipMemcpyDeviceToDevice, stream));
    }
    // Copy the device state block to the pinned mirror and wait for it. The
    // wait polls an event: a blocking wait on a millisecond of queued work
    // sleeps, and its wake-up leaves the GPU idle for tens of microseconds
    // before the next launches arrive.
    void readState() {
        HIPCHK(hipMemcpyAsync(hState, dState.p, stateBlockBytes, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipEventRecord(evSync, stream));
        hipError_t e;
        while ((e = hipEventQuery(evSync)) == hipErrorNotReady) {
        }
        HIPCHK(e);
        ++syncsInStep;
    }

    template <class F>
    void timed(KernelId kid, F&& launch) {
        ++launchesInStep;
        if (!profiling || !((profileMask >> kid) & 1ull)) {
            HIPCHK(launch());
            return;
        }
        if (evUsed == evPool.size()) {
            hipEvent_t e0, e1;
            HIPCHK(hipEventCreate(&e0));
            HIPCHK(hipEventCreate(&e1));
            evPool.emplace_back(e0, e1);
            evKernel.push_back(0);
        }
        auto& pr = evPool[evUsed];
        evKernel[evUsed] = kid;
        ++evUsed;
        // one-kernel launches of sph_kernels.hip take the events into the
        // launch (kernel start / end); the sort (several kernels, rocPRIM's
        // among them) and the adjoint passes are bracketed on the stream
        const bool inLaunch = kid != K_SORT && !(kid >= K_ADJ_SEED && kid <= K_ADJ_SOURCE);
        if (inLaunch) {
            setLaunchEvents(pr.first, pr.second);
            const hipError_t e = launch();
            const bool used = launchEventsUsed();
            setLaunchEvents(nullptr, nullptr);
            HIPCHK(e);
            if (!used) {   // nothing launched: an empty interval
                HIPCHK(hipEventRecord(pr.first, stream));
                HIPCHK(hipEventRecord(pr.second, stream));
            }
            return;
        }
        // the adjoint passes run on adjS (the reverse sweep's second stream)
        const hipStream_t es = (kid >= K_ADJ_SEED && kid <= K_ADJ_SOURCE && adjS) ? adjS : stream;
        HIPCHK(hipEventRecord(pr.first, es));
        HIPCHK(launch());
        HIPCHK(hipEventRecord(pr.second, es));
    }
    // per-scene combine of the preceding launch's block partials, then the
    // loop-control or time-step epilogue
    void finish(const DevView& d, int kind, int slot) {
        timed(K_FINISH, [&] { return launchFinish(d, kind, slot, stream); });
    }
    // Velocity correction of a Jacobi iteration; with fuseFinish it also files
    // the preceding pass's loop reduction of kind `kind`.
    void velocityPass(const DevView& d, int slot, int kind) {
        const int fin = fuseFinish ? kind : -1;
        timed(K_VELOCITY, [&] { return launchVelocity(d, slot, fin, stream); });
    }
    void collectProfile() {
        if (!evUsed) return;
        HIPCHK(hipStreamSynchronize(stream));
        for (size_t i = 0; i < evUsed; ++i) {
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, evPool[i].first, evPool[i].second));
            kernelMs[evKernel[i]] += ms;
            kernelLaunches[evKernel[i]] += 1;
        }
        evUsed = 0;
    }

    // ------------------------------------------------------------ neighbour search
    // keys -> sort -> reorder (+ cell ranges) -> within-support lists.
    // dfsphAdvance fuses the DFSPH position update into the keys kernel.
    // clearFlags false: the flags were zeroed after the last search (by a
    // queued step's density pass)
    void searchOnce(bool dfsphAdvance, bool clearFlags = true) {
        if (clearFlags) HIPCHK(hipMemsetAsync(dFlags, 0, scenes.size() * kFlagSlots * sizeof(int32_t), stream));
        DevView d = view();
        if (!radixOnce) {
            // counting sort: keys = order indices counted per index, then an
            // exclusive scan, the scatter, and the per-index fix-up that also
            // writes every cell range (so no range fill is needed)
            timed(K_KEYS, [&] { return launchKeys(d, dfsphAdvance ? 1 : 0, 0, 1, keys.p, vals.p, stream); });
            timed(K_SORT, [&] {
                hipError_t e = scanCounts(scanTemp.p, scanTempBytes, counts.p, starts.p, ordTotal + 1, stream);
                if (e == hipSuccess) e = launchScatter(d, keys.p, vals.p, stream);
                if (e == hipSuccess) e = launchOrderFix(d, vals.p, ordTotal, occ.p, rangesFull ? 1 : 0, stream);
                return e;
            });
            rangesDirty = true;   // cellOf does not list these ranges: a radix search fills first
            rangesFull = false;
            timed(K_REORDER, [&] {
                return launchReorder(d, keys.p, vals.p, xAlt.p, vAlt.p, kwsAlt.p, kvwsAlt.p, idAlt.p, 0, stream);
            });
        } else {
            radixOnce = false;
            rangesFull = true;    // ranges written here are not in occ
            const bool fill = rangesDirty;
            timed(K_KEYS, [&] { return launchKeys(d, dfsphAdvance ? 1 : 0, fill ? 0 : 1, 0, keys.p, vals.p, stream); });
            if (fill) timed(K_FILL, [&] { return launchFillRanges(fRange.p, totalCells, stream); });
            rangesDirty = false;   // the reorder below lists every range it writes in cellOf
            bool swapped = false;
            timed(K_SORT, [&] {
                return sortPairs(sortTemp.p, sortTempBytes, keys.p, keysAlt.p, vals.p, valsAlt.p, static_cast<int>(NP),
                                 keyBits, stream, &swapped);
            });
            const uint32_t* ks = swapped ? keysAlt.p : keys.p;
            const uint32_t* vs = swapped ? valsAlt.p : vals.p;
            timed(K_REORDER, [&] {
                return launchReorder(d, ks, vs, xAlt.p, vAlt.p, kwsAlt.p, kvwsAlt.p, idAlt.p, 1, stream);
            });
        }
        std::swap(x.p, xAlt.p);
        std::swap(v.p, vAlt.p);
        std::swap(kws.p, kwsAlt.p);
        std::swap(kvws.p, kvwsAlt.p);
        std::swap(id.p, idAlt.p);
        if (nc.p) std::swap(nc.p, ncAlt.p);
        d = view();
        buildLists(d);
        if (tiles) timed(K_TILES, [&] { return launchTileLists(d, stream); });
    }
    // The list build of a search, with the lanes per particle the context
    // uses. Until that count is settled (searchLanesFixed == 0), each of the
    // context's first searches times one candidate with an event pair taken
    // into the launch, in turn 1, 2, 4, 8, three rounds; the fastest time of
    // each is kept and the count with the least is used from then on. Every
    // count builds the same lists, so the tuning leaves the trajectory alone.
    void buildLists(const DevView& d) {
        tuneCollect();
        if (!searchLanesFixed && tuner.done) {
            // emitters grow a scene: tune again once the active particles have
            // more than doubled (or halved) since the count was picked
            const uint64_t na = activeParticles();
            if (na > 2 * tuner.activeAtPick || 2 * na < tuner.activeAtPick) resetTuner();
        }
        if (searchLanesFixed || tuner.done) {
            const int lanes = searchLanesFixed ? searchLanesFixed : searchLanes;
            timed(K_BUILD, [&] { return launchBuildNeighbors(d, lanes, stream); });
            return;
        }
        if (!tuner.e0) {
            HIPCHK(hipEventCreate(&tuner.e0));
            HIPCHK(hipEventCreate(&tuner.e1));
        }
        const int idx = tuner.trial % SearchTuner::kCands;
        ++launchesInStep;
        setLaunchEvents(tuner.e0, tuner.e1);
        const hipError_t e = launchBuildNeighbors(d, SearchTuner::kLanes[idx], stream);
        const bool used = launchEventsUsed();
        setLaunchEvents(nullptr, nullptr);
        HIPCHK(e);
        tuner.pending = used;
        tuner.pendingIdx = idx;
    }
    // File the timing of the last tuning launch. Every search is followed by a
    // state readback on the same stream before the next one starts, so the
    // launch has finished by the time this runs.
    void tuneCollect() {
        if (!tuner.pending) return;
        tuner.pending = false;
        HIPCHK(hipEventSynchronize(tuner.e1));
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, tuner.e0, tuner.e1));
        const int idx = tuner.pendingIdx;
        tuner.best[idx] = std::min(tuner.best[idx], ms);
        if (profiling && ((profileMask >> K_BUILD) & 1ull)) {
            kernelMs[K_BUILD] += ms;
            kernelLaunches[K_BUILD] += 1;
        }
        if (++tuner.trial < SearchTuner::kCands * SearchTuner::kRounds) return;
        int best = 0;
        for (int c = 1; c < SearchTuner::kCands; ++c)
            if (tuner.best[c] < tuner.best[best]) best = c;
        searchLanes = SearchTuner::kLanes[best];
        tuner.done = true;
        tuner.activeAtPick = activeParticles();
    }
    uint64_t activeParticles() const {
        uint64_t n = 0;
        for (const auto& sc : scenes) n += sc.na;
        return n;
    }
    void resetTuner() {
        tuner.trial = 0;
        tuner.done = false;
        tuner.pending = false;
        std::fill(tuner.best, tuner.best + SearchTuner::kCands, std::numeric_limits<float>::infinity());
    }
    void setSearchLanes(int lanes) {
        if (lanes != 0 && lanes != 1 && lanes != 2 && lanes != 4 && lanes != 8)
            throw SphError{SPH_E_ARG, "search lanes must be 0 (tune), 1, 2, 4 or 8"};
        if (stream) HIPCHK(hipStreamSynchronize(stream));
        searchLanesFixed = lanes;
        if (lanes == 0) resetTuner();   // tune afresh
    }
    // Interpret the flags of the last search (hFlags must be current). Returns
    // true if the lists are valid; otherwise grows what overflowed and returns
    // false (the caller repeats the search without the position update).
    bool searchValid() {
        bool layout = false;
        int maxF = 0, maxB = 0;
        for (size_t si = 0; si < scenes.size(); ++si) {
            const int32_t* f = hFlags + si * kFlagSlots;
            if (f[kFlagNonFinite]) throw SphError{SPH_E_NONFINITE, "non-finite particle position"};
            if (f[kFlagBigCell]) radixOnce = true;   // redo this search with the radix sort
            if (f[kFlagFar]) throw SphError{SPH_E_OUT_OF_GRID, "particle beyond 1e8 cells from the origin"};
            maxF = std::max(maxF, f[kFlagMaxFluidNbr]);
            maxB = std::max(maxB, f[kFlagMaxBoundaryNbr]);
            // particles outside the dense box sit in overflow buckets; past half
            // a bucket each, grow the box if that stays affordable, else the buckets
            const uint32_t outside = static_cast<uint32_t>(std::max(0, f[kFlagOutOfGrid]));
            if (outside > scenes[si].ovBuckets / 2) {
                if (!growGrid(si)) growBuckets(scenes[si], outside);
                layout = true;
            }
        }
        if (layout) {
            uploadLayout();
            return false;
        }
        if (static_cast<uint32_t>(maxF) > capF || static_cast<uint32_t>(maxB) > capB) {
            if (static_cast<uint32_t>(maxF) > capF) capF = ((static_cast<uint32_t>(maxF) * 5 / 4) + 7) & ~7u;
            if (static_cast<uint32_t>(maxB) > capB) capB = ((static_cast<uint32_t>(maxB) * 5 / 4) + 7) & ~7u;
            allocNeighborLists();
            return false;
        }
        if (radixOnce) return false;   // a cell was too crowded for the counting sort's in-place fix-up
        return true;
    }
    static void growBuckets(SceneH& s, uint32_t outside) {
        uint32_t b = s.ovBuckets;
        while (b < 4 * outside && b < (1u << 24)) b <<= 1;
        s.keyRange += b - s.ovBuckets;
        s.ovBuckets = b;
    }
    // Search and check at once (one host sync per attempt). Returns the
    // number of attempts the search took.
    int neighborSearch(bool dfsphAdvance) {
        if (!NP) return 0;
        for (int attempt = 0; attempt < 6; ++attempt) {
            searchOnce(dfsphAdvance && attempt == 0);
            readState();
            if (searchValid()) return attempt + 1;
        }
        throw SphError{SPH_E_OUT_OF_GRID, "could not place particles into a grid"};
    }

    // Grow scene si's dense box to the current fluid and boundary extents when
    // the box stays within a cell budget (the buckets take the rest); returns
    // false, leaving the box, otherwise.
    bool growGrid(size_t si) {
        SceneH& s = scenes[si];
        std::vector<float4> hx(s.na);
        if (s.na) HIPCHK(hipMemcpy(hx.data(), x.p + s.fBegin, s.na * sizeof(float4), hipMemcpyDeviceToHost));
        float lo[3] = {1e30f, 1e30f, 1e30f}, hi[3] = {-1e30f, -1e30f, -1e30f};
        for (const float4& q : hx) {
            lo[0] = std::min(lo[0], q.x); lo[1] = std::min(lo[1], q.y); lo[2] = std::min(lo[2], q.z);
            hi[0] = std::max(hi[0], q.x); hi[1] = std::max(hi[1], q.y); hi[2] = std::max(hi[2], q.z);
        }
        aabb(s.xb, lo, hi);
        double cells = 1.0;
        for (int a = 0; a < 3; ++a) {
            const double ext = std::floor(static_cast<double>(hi[a] * s.invh)) - std::floor(static_cast<double>(lo[a] * s.invh)) + 1;
            cells *= ext + 2 * (4 + std::ceil(0.1 * ext));
        }
        if (cells > static_cast<double>(kMaxDenseCells)) return false;
        setGrid(s, lo, hi);
        return true;
    }
    static constexpr uint32_t kMaxDenseCells = 1u << 25;

    // ------------------------------------------------------------ steps
    void stepAll(int nsteps, sph_step_stats* stats) {
        if (!initialized) throw SphError{SPH_E_STATE, "sph_init not called"};
        HIPCHK(hipSetDevice(device));
        ++stateEpoch;
        launchesInStep = 0;
        syncsInStep = 0;
        HIPCHK(hipEventRecord(evStart, stream));
        for (int st = 0; st < nsteps;) {
            if (method == SPH_WCSPH && canQueueWCSPH()) {
                const int k = std::min(kQueuedSteps, nsteps - st);
                if (queueWCSPH(k, stats ? stats + static_cast<size_t>(st) * scenes.size() : nullptr)) {
                    st += k;
                    continue;
                }
                // a search of the chunk failed its check: the chunk was undone;
                // take its steps with a host check each
                for (int j = 0; j < k; ++j, ++st) {
                    stepWCSPH();
                    if (stats)
                        for (size_t s = 0; s < scenes.size(); ++s) stats[st * scenes.size() + s] = lastStats[s];
                }
                continue;
            }
            if (method == SPH_DFSPH) stepDFSPH();
            else if (method == SPH_PCISPH) stepPCISPH();
            else stepWCSPH();
            if (stats)
                for (size_t s = 0; s < scenes.size(); ++s) stats[st * scenes.size() + s] = lastStats[s];
            ++st;
        }
        HIPCHK(hipEventRecord(evStop, stream));
        HIPCHK(hipEventSynchronize(evStop));
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, evStart, evStop));
        lastStepMs = ms;
        collectProfile();
    }
    std::vector<sph_step_stats> lastStats;

    // WCSPH steps queued without a host wait. A WCSPH step's only host
    // decision is the check of its search (lists within capacity, particles
    // within the grid's reach, cells the counting sort can order); where no
    // emitter (host-timed) or PCISPH boundary object (a hash-table history
    // that a redo would not restore) takes part and the search lanes are
    // settled, a chunk of steps is queued at once: the first workgroup of
    // each step's density pass moves the search's flags into a device log
    // (and clears them for the next search), and after the chunk one copy
    // and one wait check them all.
    // If any failed, the chunk is undone (positions, velocities, ids and
    // time from a snapshot taken before it, the same grid) and retaken step
    // by step with the host check, which grows what overflowed; every step
    // that passed its check is the step the checked path takes, bit for bit.
    static constexpr int kQueuedSteps = 32;
    int32_t* hQueueFlags = nullptr;
    size_t queueFlagCap = 0;
    DevBuf<int32_t> queueLog;   // the chunk's search flags, step by step
    bool canQueueWCSPH() const {
        if (!NP || replaying) return false;
        if (!searchLanesFixed && !tuner.done) return false;
        for (const auto& sc : scenes)
            if (!sc.emitters.empty() || sc.pciB) return false;
        return true;
    }
    // The checks of searchValid without its side effects.
    bool flagsPass(const int32_t* flags) const {
        for (size_t si = 0; si < scenes.size(); ++si) {
            const int32_t* f = flags + si * kFlagSlots;
            if (f[kFlagNonFinite] || f[kFlagBigCell] || f[kFlagFar]) return false;
            if (static_cast<uint32_t>(std::max(0, f[kFlagOutOfGrid])) > scenes[si].ovBuckets / 2) return false;
            if (static_cast<uint32_t>(std::max(0, f[kFlagMaxFluidNbr])) > capF) return false;
            if (static_cast<uint32_t>(std::max(0, f[kFlagMaxBoundaryNbr])) > capB) return false;
        }
        return !radixOnce;
    }
    // Two queued WCSPH steps as one HIP graph, replayed for the step pairs of
    // an even chunk: after two searches every swapped slot buffer is back in
    // place, so the same launches serve every pair. The pair's searches file
    // their flags into one max-accumulating array, all the chunk's
    // check needs. The graph is captured again whenever anything its launches
    // depend on (the device view, the search's buffers and path, the force
    // options) differs from what it was captured with; a capture that fails
    // leaves the host state as it was and switches graphs off for the context.
    struct PairKey {
        DevView d;
        const void* ptr[16];
        uint64_t val[12];
        bool operator==(const PairKey& o) const { return std::memcmp(this, &o, sizeof(PairKey)) == 0; }
    };
    hipGraphExec_t pairExec = nullptr;
    PairKey pairKey{};
    int64_t pairLaunches = 0;
    bool graphsOff = false;
    DevBuf<int32_t> flagAcc;
    PairKey currentPairKey() {
        PairKey k;
        std::memset(static_cast<void*>(&k), 0, sizeof(k));
        k.d = view();
        k.d.stepNo = 0;
        k.d.markNo = 0;
        const void* ps[16] = {xAlt.p, vAlt.p, kwsAlt.p, kvwsAlt.p, idAlt.p, keys.p, vals.p, keysAlt.p,
                              valsAlt.p, scanTemp.p, sortTemp.p, occ.p, ncAlt.p, flagAcc.p, dFlags, nullptr};
        std::copy(ps, ps + 16, k.ptr);
        const uint64_t vs[12] = {scanTempBytes, sortTempBytes, ordTotal, static_cast<uint64_t>(keyBits),
                                 radixOnce, rangesDirty, rangesFull, static_cast<uint64_t>(anyST() || anyAdh()),
                                 static_cast<uint64_t>(anyCube()),
                                 static_cast<uint64_t>(searchLanesFixed ? searchLanesFixed : searchLanes), tiles,
                                 totalCells};
        std::copy(vs, vs + 12, k.val);
        return k;
    }
    // A view whose density pass moves the search's flags to `log` (mode 1:
    // copy, 2: maximum) and clears them.
    DevView flagView(const DevView& d, int32_t* log, int mode) const {
        DevView q = d;
        q.flagLog = log;
        q.flagCount = static_cast<uint32_t>(scenes.size() * kFlagSlots);
        q.flagMode = mode;
        return q;
    }
    // The launches of one queued step; flags go to the log slot or, with
    // acc, into the accumulating array.
    void queuedStep(int32_t* logSlot, bool acc) {
        searchOnce(false, false);
        const DevView d = view();
        const DevView dl = flagView(d, acc ? flagAcc.p : logSlot, acc ? 2 : 1);
        timed(K_DENSITY, [&] { return launchDensityPressure(dl, stream); });
        timed(K_FORCES, [&] { return launchForces(d, 1, stream); });
        if (anyST() || anyAdh()) timed(K_SURFACE_TENSION, [&] { return launchSurfaceTension(d, 0, stream); });
        timed(K_INTEGRATE, [&] { return launchIntegrate(d, stream); });
        if (anyCube()) timed(K_CUBE, [&] { return launchCubeBoundary(d, stream); });
    }
    void capturePair() {
        if (pairExec) (void)hipGraphExecDestroy(pairExec);
        pairExec = nullptr;
        const PairKey key = currentPairKey();
        // host state the two searches change, restored if the capture fails
        float4* sx[4] = {x.p, xAlt.p, v.p, vAlt.p};
        float* sf[4] = {kws.p, kwsAlt.p, kvws.p, kvwsAlt.p};
        uint32_t* si[2] = {id.p, idAlt.p};
        float4* sn[2] = {nc.p, ncAlt.p};
        const bool sr[3] = {radixOnce, rangesDirty, rangesFull};
        const int64_t l0 = launchesInStep;
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal);
        bool ok = e == hipSuccess;
        if (ok) {
            try {
                queuedStep(nullptr, true);
                queuedStep(nullptr, true);
            } catch (const SphError&) {
                ok = false;
            }
            e = hipStreamEndCapture(stream, &g);
            ok = ok && e == hipSuccess && g;
        }
        if (ok) {
            e = hipGraphInstantiate(&pairExec, g, nullptr, nullptr, 0);
            ok = e == hipSuccess;
        }
        if (g) (void)hipGraphDestroy(g);
        if (!ok) {
            pairExec = nullptr;
            graphsOff = true;
            (void)hipGetLastError();
            x.p = sx[0]; xAlt.p = sx[1]; v.p = sx[2]; vAlt.p = sx[3];
            kws.p = sf[0]; kwsAlt.p = sf[1]; kvws.p = sf[2]; kvwsAlt.p = sf[3];
            id.p = si[0]; idAlt.p = si[1];
            nc.p = sn[0]; ncAlt.p = sn[1];
            radixOnce = sr[0]; rangesDirty = sr[1]; rangesFull = sr[2];
            launchesInStep = l0;
            return;
        }
        pairKey = key;
        pairLaunches = launchesInStep - l0;
        launchesInStep = l0;
    }
    // With slot0 >= 0 the chunk is steps slot0 .. slot0 + steps - 1 of the
    // segment being recorded (stepWCSPHRecord's tape writes; the lists are
    // taped at capacity, since the longest list is not known on the host).
    bool queueWCSPH(int steps, sph_step_stats* stats, int slot0 = -1) {
        const size_t nf = scenes.size() * kFlagSlots;
        if (queueFlagCap < static_cast<size_t>(steps) * nf) {
            if (hQueueFlags) HIPCHK(hipHostFree(hQueueFlags));
            hQueueFlags = nullptr;
            queueFlagCap = 0;
            HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&hQueueFlags), static_cast<size_t>(kQueuedSteps) * nf * sizeof(int32_t)));
            queueFlagCap = static_cast<size_t>(kQueuedSteps) * nf;
        }
        if (queueLog.n < static_cast<size_t>(kQueuedSteps) * nf) queueLog.alloc(static_cast<size_t>(kQueuedSteps) * nf);
        saveCheckpoint(queueSnap);
        const uint32_t stepNo0 = stepNo, markNo0 = markNo;
        HIPCHK(hipMemsetAsync(dFlags, 0, nf * sizeof(int32_t), stream));
        const bool record = slot0 >= 0;
        bool graphed = false;
        if (!record && steps % 2 == 0 && !graphsOff && !profiling && !radixOnce) {   // counting-sort searches only
            if (flagAcc.n < nf) flagAcc.alloc(nf);
            HIPCHK(launchFlagClear(dFlags, flagAcc.p, static_cast<uint32_t>(nf), stream));
            if (!pairExec || !(currentPairKey() == pairKey)) capturePair();
            if (pairExec) {
                for (int q = 0; q < steps / 2; ++q) HIPCHK(hipGraphLaunch(pairExec, stream));
                launchesInStep += pairLaunches * (steps / 2);
                stepNo += static_cast<uint32_t>(steps);
                markNo = stepNo - 1;
                HIPCHK(hipMemcpyAsync(hQueueFlags, flagAcc.p, nf * sizeof(int32_t), hipMemcpyDeviceToHost, stream));
                graphed = true;
            }
        }
        for (int k = 0; k < steps && !graphed; ++k) {
            int32_t* logSlot = queueLog.p + static_cast<size_t>(k) * nf;
            if (!record) {
                markNo = stepNo;
                queuedStep(logSlot, false);
                ++stepNo;
                continue;
            }
            // stepWCSPHRecord's launches and tape writes
            searchOnce(false, false);
            const DevView d = view();
            const DevView dl = flagView(d, logSlot, 1);
            const size_t o = static_cast<size_t>(slot0 + k) * NP;
            HalfTape& t = tape.h[0];
            recordSearch(t, slot0 + k, capF, capB);
            tapeCopy(t.v.p + o, v.p, NP * sizeof(float4));
            timed(K_DENSITY, [&] { return launchDensityPressure(dl, stream); });
            timed(K_FORCES, [&] { return launchForces(d, 1, stream); });
            if (anyST()) timed(K_SURFACE_TENSION, [&] { return launchSurfaceTension(d, 0, stream); });
            tapeCopy(t.rho.p + o, rho.p, NP * sizeof(float));
            tapeCopy(t.p.p + o, p.p, NP * sizeof(float));
            if (anyST()) tapeCopy(t.n.p + o, n.p, NP * sizeof(float4));
            flushCopies();   // before the integration moves x and v
            timed(K_INTEGRATE, [&] { return launchIntegrate(d, stream); });
        }
        if (!graphed)
            HIPCHK(hipMemcpyAsync(hQueueFlags, queueLog.p, static_cast<size_t>(steps) * nf * sizeof(int32_t),
                                  hipMemcpyDeviceToHost, stream));
        readState();   // waits for the chunk; the state block as the checked path leaves it
        bool ok = !radixOnce;
        for (int k = 0; k < (graphed ? 1 : steps) && ok; ++k) ok = flagsPass(hQueueFlags + static_cast<size_t>(k) * nf);
        if (!ok) {
            stepNo = stepNo0;
            markNo = markNo0;
            restoreCheckpoint(queueSnap);
            return false;
        }
        // the statistics of each step: t advances by dt per step (the float
        // sums kIntegrate forms); the rest is the state block's
        std::vector<float> t0(scenes.size());
        for (size_t s = 0; s < scenes.size(); ++s) t0[s] = queueSnap.t[s];
        for (int k = 0; k < steps; ++k) {
            beginStats();
            for (size_t s = 0; s < scenes.size(); ++s) {
                t0[s] = t0[s] + hState[s].dt;
                sph_step_stats& o = lastStats[s];
                o.dt = hState[s].dt;
                o.t = t0[s];
                o.max_velocity = hState[s].vmax;
                o.iterations = hState[s].its;
                o.avg_density_error = hState[s].errD;
                o.capped_density = hState[s].capped & 1;
            }
            if (stats)
                for (size_t s = 0; s < scenes.size(); ++s) stats[static_cast<size_t>(k) * scenes.size() + s] = lastStats[s];
        }
        for (size_t s = 0; s < scenes.size(); ++s) {
            scenes[s].dt = hState[s].dt;
            scenes[s].t = hState[s].t;
        }
        if (record)   // what the replay of these steps checks (recordedSearch)
            for (int k = 0; k < steps; ++k) {
                const size_t g = static_cast<size_t>(recBase + slot0 + k);
                std::copy(hQueueFlags + static_cast<size_t>(k) * nf, hQueueFlags + static_cast<size_t>(k + 1) * nf,
                          rec.flags.begin() + static_cast<std::ptrdiff_t>(g * nf));
                rec.rowsF[g] = capF;
                rec.rowsB[g] = capB;
                rec.retried[g] = 0;
            }
        return true;
    }

    bool anyVisc() const {
        for (auto& s : scenes) if (s.viscOn) return true;
        return false;
    }
    bool anyST() const {
        for (auto& s : scenes) if (s.stOn) return true;
        return false;
    }
    int maxOver(uint32_t sph_scene_params::*field) const {
        uint32_t m = 0;
        for (auto& s : scenes) m = std::max(m, s.P.*field);
        return static_cast<int>(std::min<uint32_t>(m, 1u << 30));
    }

    // Launch (A, B) iteration pairs of a Jacobi loop in chunks; the device
    // decides per scene whether an iteration runs (flags by parity), the host
    // only checks after each chunk whether any scene still wants one.
    // afterFirstSync runs once, right after the first chunk's sync; if it
    // returns false the loop is abandoned (the caller redoes the phase).
    // With sync false and a first chunk that covers the bound (a loop whose
    // iteration count is fixed), the loop ends without a host wait.
    template <class A, class B, class C>
    bool runLoop(int firstChunk, int bound, A&& passA, B&& passB, C&& afterFirstSync, bool sync = true) {
        int issued = 0;
        int chunk = std::max(1, std::min(firstChunk, bound));
        bool firstSync = true;
        while (true) {
            for (int k = 0; k < chunk && issued < bound; ++k, ++issued) {
                passA(issued & 1);
                passB(issued & 1);
            }
            if (!sync && issued >= bound) return true;
            readState();
            if (firstSync) {
                firstSync = false;
                if (!afterFirstSync()) return false;
            }
            if (issued >= bound) break;
            bool any = false;
            for (size_t s = 0; s < scenes.size(); ++s) any |= hState[s].active[issued & 1] != 0;
            if (!any) break;
            chunk = 2;
        }
        return true;
    }

    void beginStats() {
        lastStats.assign(scenes.size(), sph_step_stats{});
        for (size_t s = 0; s < scenes.size(); ++s) lastStats[s].scene = static_cast<int32_t>(s);
    }
    // Host mirror after a step whose last device-side bookkeeping (t += dt, loop
    // statistics) ran after the last readState: mirror it on the host.
    void fileStats(bool dfsph) {
        for (size_t s = 0; s < scenes.size(); ++s) {
            SceneH& sc = scenes[s];
            SceneState& T = hState[s];
            sph_step_stats& o = lastStats[s];
            sc.dt = T.dt;
            sc.t = T.t;
            o.dt = T.dt;
            o.t = T.t;
            o.max_velocity = T.vmax;
            if (dfsph) {
                o.iterations = T.its;
                o.avg_density_error = T.errD;
                o.capped_density = T.capped & 1;
                o.iterations_v = T.itsV;
                o.avg_divergence_error = T.errV;
                o.capped_divergence = (T.capped >> 1) & 1;
            } else {
                o.iterations = T.its;
                o.avg_density_error = T.errD;
                o.capped_density = T.capped & 1;
            }
        }
    }

    // SPHSolver::computeBoundaryDensities of PCISPH boundary objects at the
    // current positions (DFSPH computes them with the fluid's; none of its
    // passes reads them)
    void boundaryDensities() {
        const DevView d = view();
        for (size_t si = 0; si < scenes.size(); ++si)
            if (scenes[si].pciB)
                timed(K_BDENSITY, [&] {
                    return launchBoundaryDensity(d, static_cast<int>(si), scenes[si].bBegin, scenes[si].bEnd, 0, 2, stream);
                });
    }

    void stepDFSPH() {
        beginStats();
        if (!NP) return advanceEmpty();
        markNo = stepNo + 1;   // the mid-step search follows this step's position update
        DevView d = view();
        // 1-2. non-pressure forces, then the CFL maximum; the new dt is set on the device
        if (anyVisc() || anyST()) timed(K_FORCES, [&] { return launchForces(d, 0, stream); });
        if (anyST() || anyAdh()) timed(K_SURFACE_TENSION, [&] { return launchSurfaceTension(d, 1, stream); });
        else timed(K_VMAX, [&] { return launchVmax(d, stream); });
        finish(d, kFinishTimeStep, 2);
        // 3. predictVelocities
        timed(K_PREDICT, [&] { return launchPredictVelocity(d, stream); });
        // 4. correctDensityError: warm start (and loop start), then the loop
        timed(K_RHO_STAR, [&] { return launchRhoStar(d, 1, 2, stream); });
        timed(K_VELOCITY, [&] { return launchVelocity(d, 2, -1, stream); });
        int first = 1;
        bool fixedIts = true;   // every scene's density loop runs exactly max_iterations
        for (auto& sc : scenes) {
            first = std::max(first, std::max(static_cast<int>(sc.P.min_iterations), sc.its));
            fixedIts = fixedIts && sc.P.min_iterations >= sc.P.max_iterations;
        }
        const int boundD = maxOver(&sph_scene_params::max_iterations);
        runLoop(first, boundD,
                [&](int slot) {
                    timed(K_RHO_STAR, [&] { return launchRhoStar(d, 0, slot, stream); });
                    if (!fuseFinish) finish(d, kFinishDensity, slot);
                },
                [&](int slot) { velocityPass(d, slot, kFinishDensity); },
                [] { return true; }, !(fixedIts && first >= boundD));
        // (the device files the loop's count and error itself, in the keys pass)
        for (size_t s = 0; s < scenes.size(); ++s)
            scenes[s].its = fixedIts ? static_cast<int>(scenes[s].P.max_iterations) : hState[s].it;
        // 5-7. updatePositions fused with the keys (emitted particles are
        // placed after it, DFSPHSolver.cpp:72), then the neighbour search; its
        // validity is checked at the divergence loop's first sync
        emitAll(true);
        searchOnce(true);
        for (int attempt = 0;; ++attempt) {
            d = view();
            // 8-9. computeDensities + computeAlpha
            timed(K_DENSITY, [&] { return launchDensityAlpha(d, 1, stream); });
            if (anyPciB()) boundaryDensities();
            // 10. correctDivergenceError: warm start (and loop start), then the loop
            timed(K_DIVERGENCE, [&] { return launchDivergence(d, 1, 2, stream); });
            timed(K_VELOCITY, [&] { return launchVelocity(d, 2, -1, stream); });
            int firstV = 1;
            for (auto& sc : scenes) firstV = std::max(firstV, std::max(static_cast<int>(sc.P.min_iterations_v), sc.itsV));
            const bool ok = runLoop(
                firstV, maxOver(&sph_scene_params::max_iterations),
                [&](int slot) {
                    timed(K_DIVERGENCE, [&] { return launchDivergence(d, 0, slot, stream); });
                    if (!fuseFinish) finish(d, kFinishDivergence, slot);
                },
                [&](int slot) { velocityPass(d, slot, kFinishDivergence); },
                [&] { return searchValid(); });
            if (ok) break;
            if (attempt >= 5) throw SphError{SPH_E_OUT_OF_GRID, "could not place particles into a grid"};
            // the lists were invalid: undo the divergence passes and search again
            copyVel(v.p, snapV.p);
            HIPCHK(hipMemcpyAsync(kvws.p, snapKvws.p, NP * sizeof(float), hipMemcpyDeviceToDevice, stream));
            searchOnce(false);
        }
        for (size_t s = 0; s < scenes.size(); ++s) scenes[s].itsV = hState[s].it;
        timed(K_SCALE, [&] { return launchScaleDivergence(view(), stream); });
        if (anyCube()) timed(K_CUBE, [&] { return launchCubeBoundary(view(), stream); });   // DFSPHSolver.cpp:92-102
        // kScaleDivergence filed itsV/errV and advanced t on the device; mirror it
        for (size_t s = 0; s < scenes.size(); ++s) {
            SceneState& T = hState[s];
            const SceneConst_host c = hostConst(s);
            T.itsV = T.it;
            T.errV = T.err;
            if (T.it >= c.maxIt && T.err > static_cast<double>(c.etaV / T.dt)) T.capped |= 2;
            T.t = T.t + T.dt;
        }
        ++stepNo;
        fileStats(true);
    }

    void stepWCSPH() {
        beginStats();
        if (!NP) return advanceEmpty();
        markNo = stepNo;
        neighborSearch(false);
        DevView d = view();
        timed(K_DENSITY, [&] { return launchDensityPressure(d, stream); });
        // boundary densities and pressures of PCISPH boundary objects (SPHSolver.cpp:91-123, WCSPHSolver.cpp:97-115)
        for (size_t si = 0; si < scenes.size(); ++si)
            if (scenes[si].pciB)
                timed(K_BDENSITY, [&] {
                    return launchBoundaryDensity(d, static_cast<int>(si), scenes[si].bBegin, scenes[si].bEnd, 0, 2, stream);
                });
        timed(K_FORCES, [&] { return launchForces(d, 1, stream); });
        if (anyST() || anyAdh()) timed(K_SURFACE_TENSION, [&] { return launchSurfaceTension(d, 0, stream); });
        timed(K_INTEGRATE, [&] { return launchIntegrate(d, stream); });   // + PCISPH boundary corrections
        if (anyCube()) timed(K_CUBE, [&] { return launchCubeBoundary(d, stream); });   // WCSPHSolver.cpp:51-61
        emitAll(false);   // WCSPHSolver.cpp:75, before t += dt
        for (size_t s = 0; s < scenes.size(); ++s) hState[s].t = hState[s].t + hState[s].dt;
        ++stepNo;
        fileStats(false);
    }

    void stepPCISPH() {
        beginStats();
        if (!NP) return advanceEmpty();
        markNo = stepNo;
        neighborSearch(false);
        DevView d = view();
        timed(K_DENSITY, [&] { return launchDensityAlpha(d, 0, stream); });
        if (anyVisc() || anyST()) timed(K_FORCES, [&] { return launchForces(d, 0, stream); });
        if (anyST() || anyAdh()) timed(K_SURFACE_TENSION, [&] { return launchSurfaceTension(d, 0, stream); });
        // initPressure zeroes the boundary pressures (PCISPHSolver.cpp:100-117)
        if (anyPciB()) HIPCHK(hipMemsetAsync(bP.p, 0, bP.n * sizeof(float), stream));
        timed(K_PCI_INIT, [&] { return launchPciInit(d, stream); });
        int first = 1;
        for (auto& sc : scenes) first = std::max(first, std::max(static_cast<int>(sc.P.min_iterations), sc.its));
        runLoop(first, maxOver(&sph_scene_params::max_iterations),
                [&](int slot) {
                    timed(K_PCI_DENSITY, [&] { return launchPciDensity(d, slot, stream); });
                    // predictBoundaryDensities + the boundary part of updatePressure (PCISPHSolver.cpp:211-297)
                    for (size_t si = 0; si < scenes.size(); ++si)
                        if (scenes[si].pciB)
                            timed(K_BDENSITY, [&] {
                                return launchBoundaryDensity(d, static_cast<int>(si), scenes[si].bBegin, scenes[si].bEnd,
                                                             1, slot, stream);
                            });
                    if (!fuseFinish) finish(d, kFinishPciDensity, slot);
                },
                [&](int slot) {
                    const int fin = fuseFinish ? kFinishPciDensity : -1;
                    timed(K_PCI_PACC, [&] { return launchPciPressureAcc(d, slot, fin, stream); });
                },
                [] { return true; });
        timed(K_PCI_INTEGRATE, [&] { return launchPciIntegrate(d, stream); });
        if (anyCube()) timed(K_CUBE, [&] { return launchCubeBoundary(d, stream); });   // PCISPHSolver.cpp:53-61
        emitAll(false);   // PCISPHSolver.cpp:77, before t += dt
        for (size_t s = 0; s < scenes.size(); ++s) {
            SceneState& T = hState[s];
            const SceneConst_host c = hostConst(s);
            scenes[s].its = T.it;
            T.its = T.it;
            T.errD = T.err;
            T.capped = (T.it >= c.maxIt && T.err > static_cast<double>(c.eta)) ? 1 : 0;
            T.t = T.t + T.dt;
        }
        ++stepNo;
        fileStats(false);
    }

    // No fluid particle in any scene: only time advances.
    // ------------------------------------------------------------ differentiable rollout (WCSPH)
    // A recorded rollout keeps a checkpoint (slot-order x, v, id, grids, time)
    // at the start of every segment of `every` steps, and a tape of the last
    // recomputed segment: per step the state after its reorder, rho, p,
    // normals, ids and neighbour lists. The reverse sweep walks the segments
    // from the last, recomputes a segment's tape from its checkpoint (the
    // forward step is deterministic: stable radix sort, fixed-order sums, no
    // atomics in values), and runs the adjoint passes step by step backwards.
    struct GridSave {
        int g0[3], gn[3];
        uint32_t nCells, keyRange, ovBuckets;
        int morton;
        bool operator==(const GridSave& o) const {
            return std::equal(g0, g0 + 3, o.g0) && std::equal(gn, gn + 3, o.gn) && nCells == o.nCells &&
                   keyRange == o.keyRange && ovBuckets == o.ovBuckets && morton == o.morton;
        }
    };
    struct Checkpoint {
        DevBuf<float4> x, v;
        DevBuf<uint32_t> id;
        DevBuf<float> kws, kvws;  // DFSPH warm-start state
        std::vector<GridSave> grids;
        std::vector<SceneState> state;
        std::vector<float> t, dt;
    };
    // One half of a recorded step (WCSPH: the whole step). Per step: the
    // state after the half's neighbour search (x, v, ids, lists), rho, p,
    // alpha, normals; DFSPH halves also keep, per Jacobi iteration, the
    // velocity its source pass read (u) and the pass's k and rho* / divergence
    // (aux), the warm-start pass at index nit.
    struct HalfTape {
        DevBuf<float4> x, v, n, u;
        DevBuf<float> rho, p, alpha, k, aux;
        DevBuf<uint32_t> id, cntF, cntB, nbrF, nbrB;
    };
    struct Tape {
        int steps = 0, nit = 0, halves = 0;
        size_t np = 0;
        uint32_t capF = 0, capB = 0;
        HalfTape h[2];
    };
    struct Recording {
        bool valid = false;
        int nsteps = 0, every = 0, segments = 0, tapeSeg = -1;
        uint64_t epoch = 0;
        std::vector<Checkpoint> ckpt;
        Checkpoint last;
        // per recorded step (the one search of a WCSPH step, the second
        // half's of a DFSPH step): the flags of its accepted search
        // (kFlagSlots per scene), the list rows it put on the tape, and
        // whether it took more than one attempt
        std::vector<int32_t> flags;
        std::vector<uint32_t> rowsF, rowsB;
        std::vector<uint8_t> retried;
        // the segment the second tape (tapeRd) holds from the recording (the
        // one before the last), or -1
        int tapeSegRd = -1;
    };
    Recording rec;
    Checkpoint queueSnap;   // the state before a chunk of queued WCSPH steps
    Tape tape;
    // Recorded steps: recBase is the global index of slot 0 of the segment
    // being recorded or replayed; while `replaying`, searches that took one
    // attempt in the recording are not waited for (recordedSearch), and
    // their flags land in hReplayFlags (pinned, replayFlagCap entries) for
    // the first replayChecks slots of the segment starting at replayBase.
    bool replaying = false;
    int recBase = 0, replayBase = 0, replayChecks = 0;
    int32_t* hReplayFlags = nullptr;
    size_t replayFlagCap = 0;
    // The reverse sweep runs two segments at once: the adjoint passes of one
    // segment read its tape (tapeRd) on adjStream while the forward recomputes
    // the segment before it into `tape` on the engine's stream; the tapes swap
    // at each segment boundary. adjTape is the tape adjView / adjPass read and
    // adjS the stream the adjoint passes go to (the engine's stream when the
    // sweep has one segment or no room for the second tape). The adjoint
    // passes read scene constants and dt from their own copies (dScenesAdj,
    // dStateAdj), which the concurrent forward never writes.
    Tape tapeRd;
    const Tape* adjTape = &tape;
    hipStream_t adjStream = nullptr, adjS = nullptr;
    hipEvent_t evAdjDone = nullptr, evRecDone = nullptr;
    DevBuf<SceneConst> dScenesAdj;
    DevBuf<SceneState> dStateAdj;
    uint64_t stateEpoch = 0;  // bumped by every call that changes the state outside a recording
    DevBuf<float4> adjX, adjV, aBar, xBar, vBar, nBar, sgBar, suBar;
    DevBuf<float> rhoBar, alphaBar, kBar, sBar, adjKws, adjKvws, adjRho, adjAlpha;
    DevBuf<double> adjPartials, adjAcc;
    std::vector<std::vector<float>> seedX, seedV;
    std::vector<std::vector<float>> gradX0, gradV0;
    std::vector<double> gradCoef;
    bool gradValid = false;

    void copyD(void* dst, const void* src, size_t bytes) {
        if (bytes) HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, stream));
    }
    // Tape copies gathered into batched launches (whole 16-byte words; every
    // tape array is a multiple of 256 slots of 4- or 16-byte elements).
    CopyBatch pendingCopies{};
    void tapeCopy(void* dst, const void* src, size_t bytes) {
        if (!bytes) return;
        if ((bytes & 15u) || (reinterpret_cast<uintptr_t>(dst) & 15u) || (reinterpret_cast<uintptr_t>(src) & 15u)) {
            copyD(dst, src, bytes);
            return;
        }
        if (pendingCopies.n == kCopyBatchMax) flushCopies();
        const int c = pendingCopies.n++;
        pendingCopies.src[c] = static_cast<const uint4*>(src);
        pendingCopies.dst[c] = static_cast<uint4*>(dst);
        pendingCopies.words[c] = bytes / 16u;
    }
    void flushCopies() {
        if (!pendingCopies.n) return;
        const CopyBatch b = pendingCopies;
        pendingCopies = CopyBatch{};
        timed(K_TAPE, [&] { return launchCopyBatch(b, stream); });
    }
    void saveCheckpoint(Checkpoint& c) {
        if (c.x.n != NP) {
            c.x.alloc(NP);
            c.v.alloc(NP);
            c.id.alloc(NP);
        }
        copyD(c.x.p, x.p, NP * sizeof(float4));
        copyD(c.v.p, v.p, NP * sizeof(float4));
        copyD(c.id.p, id.p, NP * sizeof(uint32_t));
        if (method == SPH_DFSPH) {
            if (c.kws.n != NP) {
                c.kws.alloc(NP);
                c.kvws.alloc(NP);
            }
            copyD(c.kws.p, kws.p, NP * sizeof(float));
            copyD(c.kvws.p, kvws.p, NP * sizeof(float));
        }
        c.grids.resize(scenes.size());
        c.t.resize(scenes.size());
        c.dt.resize(scenes.size());
        for (size_t i = 0; i < scenes.size(); ++i) {
            const SceneH& sc = scenes[i];
            GridSave& g = c.grids[i];
            std::copy(sc.g0, sc.g0 + 3, g.g0);
            std::copy(sc.gn, sc.gn + 3, g.gn);
            g.nCells = sc.nCells;
            g.keyRange = sc.keyRange;
            g.ovBuckets = sc.ovBuckets;
            g.morton = sc.morton;
            c.t[i] = sc.t;
            c.dt[i] = sc.dt;
        }
        c.state.assign(hState, hState + scenes.size());
    }
    void restoreCheckpoint(const Checkpoint& c) {
        HIPCHK(hipStreamSynchronize(stream));
        bool layout = false;
        for (size_t i = 0; i < scenes.size(); ++i) {
            SceneH& sc = scenes[i];
            GridSave cur;
            std::copy(sc.g0, sc.g0 + 3, cur.g0);
            std::copy(sc.gn, sc.gn + 3, cur.gn);
            cur.nCells = sc.nCells;
            cur.keyRange = sc.keyRange;
            cur.ovBuckets = sc.ovBuckets;
            cur.morton = sc.morton;
            const GridSave& g = c.grids[i];
            if (!(cur == g)) {
                std::copy(g.g0, g.g0 + 3, sc.g0);
                std::copy(g.gn, g.gn + 3, sc.gn);
                sc.nCells = g.nCells;
                sc.keyRange = g.keyRange;
                sc.ovBuckets = g.ovBuckets;
                sc.morton = g.morton;
                layout = true;
            }
            sc.t = c.t[i];
            sc.dt = c.dt[i];
        }
        if (layout) uploadLayout();
        copyD(x.p, c.x.p, NP * sizeof(float4));
        copyD(v.p, c.v.p, NP * sizeof(float4));
        copyD(id.p, c.id.p, NP * sizeof(uint32_t));
        std::copy(c.state.begin(), c.state.end(), hState);
        HIPCHK(hipMemcpyAsync(dState.p, hState, scenes.size() * sizeof(SceneState), hipMemcpyHostToDevice, stream));
        if (method == SPH_DFSPH) {
            // The slots are in the order of the keys of x (the last search
            // sorted them), so a fresh search keeps that order (stable sort)
            // and rebuilds the same lists; rho and alpha follow from them.
            copyD(kws.p, c.kws.p, NP * sizeof(float));
            copyD(kvws.p, c.kvws.p, NP * sizeof(float));
            neighborSearch(false);
            HIPCHK(launchDensityAlpha(view(), 1, stream));
        }
        HIPCHK(hipStreamSynchronize(stream));
    }
    // ------------------------------------------------------------ state files
    // The whole context for an exact resume: per scene its grid layout, time,
    // solver state and emitter progress, then the slot arrays in slot order.
    // A DFSPH load re-runs the search and the density/alpha pass on the loaded
    // slots, as a checkpoint restore does (same keys, stable sort: the same
    // lists, densities and alphas as the saved run had).
    static constexpr char kStateMagic[8] = {'S', 'P', 'H', '2', 'O', 'S', 'T', '1'};
    void saveState(const char* path) {
        if (anyPciB())   // the hash table's cell history (cellFirst) is not part of the file format
            throw SphError{SPH_E_UNSUPPORTED, "state files of PCISPH boundary scenes are not supported"};
        if (!initialized) throw SphError{SPH_E_STATE, "sph_init not called"};
        if (!path) throw SphError{SPH_E_ARG, "path is NULL"};
        HIPCHK(hipStreamSynchronize(stream));
        std::ofstream f(path, std::ios::binary);
        if (!f) throw SphError{SPH_E_ARG, std::string("cannot write ") + path};
        auto put = [&](const void* q, size_t n) { f.write(static_cast<const char*>(q), static_cast<std::streamsize>(n)); };
        const uint32_t ns = static_cast<uint32_t>(scenes.size());
        const int32_t m = method;
        const uint64_t np = NP;
        put(kStateMagic, 8);
        put(&ns, 4);
        put(&m, 4);
        put(&np, 8);
        for (size_t i = 0; i < scenes.size(); ++i) {
            const SceneH& sc = scenes[i];
            const uint32_t counts[4] = {sc.nf, sc.na, sc.nb, static_cast<uint32_t>(sc.emitters.size())};
            put(counts, sizeof(counts));
            const int32_t grid[7] = {sc.g0[0], sc.g0[1], sc.g0[2], sc.gn[0], sc.gn[1], sc.gn[2], sc.morton};
            const uint32_t cells[3] = {sc.nCells, sc.keyRange, sc.ovBuckets};
            put(grid, sizeof(grid));
            put(cells, sizeof(cells));
            const float tt[2] = {sc.t, sc.dt};
            const int32_t its[2] = {sc.its, sc.itsV};
            put(tt, sizeof(tt));
            put(its, sizeof(its));
            put(&hState[i], sizeof(SceneState));
            for (const auto& e : sc.emitters) {
                put(&e.remaining, 8);
                put(&e.next, 4);
            }
        }
        auto hx = download(x.p, NP);
        auto hv = downloadVel(0, static_cast<uint32_t>(NP));
        auto hid = download(id.p, NP);
        auto hkw = download(kws.p, NP);
        auto hkv = download(kvws.p, NP);
        put(hx.data(), NP * sizeof(float4));
        put(hv.data(), NP * sizeof(float4));
        put(hid.data(), NP * sizeof(uint32_t));
        put(hkw.data(), NP * sizeof(float));
        put(hkv.data(), NP * sizeof(float));
        if (!f) throw SphError{SPH_E_ARG, std::string("cannot write ") + path};
    }
    // The whole file is parsed into temporaries and validated before any of
    // the context changes: a truncated or inconsistent file leaves it as it
    // was (same host and device state, same next step).
    void loadState(const char* path) {
        if (anyPciB()) throw SphError{SPH_E_UNSUPPORTED, "state files of PCISPH boundary scenes are not supported"};
        if (!initialized) throw SphError{SPH_E_STATE, "sph_init not called"};
        if (!path) throw SphError{SPH_E_ARG, "path is NULL"};
        std::ifstream f(path, std::ios::binary);
        if (!f) throw SphError{SPH_E_ARG, std::string("cannot read ") + path};
        auto get = [&](void* q, size_t n) {
            f.read(static_cast<char*>(q), static_cast<std::streamsize>(n));
            if (!f) throw SphError{SPH_E_ARG, "truncated state file"};
        };
        auto bad = [](const char* what) { return SphError{SPH_E_ARG, std::string("invalid state file: ") + what}; };
        char magic[8];
        uint32_t ns = 0;
        int32_t m = 0;
        uint64_t np = 0;
        get(magic, 8);
        if (std::memcmp(magic, kStateMagic, 8) != 0) throw SphError{SPH_E_ARG, "not a state file"};
        get(&ns, 4);
        get(&m, 4);
        get(&np, 8);
        if (ns != scenes.size() || m != method || np != NP)
            throw SphError{SPH_E_ARG, "state file does not match the context's scenes"};
        struct SceneLoad {
            uint32_t na = 0;
            GridSave g{};
            float t = 0, dt = 0;
            int32_t its = 0, itsV = 0;
            SceneState st{};
            std::vector<std::pair<int64_t, float>> emit;
        };
        std::vector<SceneLoad> L(scenes.size());
        for (size_t i = 0; i < scenes.size(); ++i) {
            const SceneH& sc = scenes[i];
            SceneLoad& q = L[i];
            uint32_t counts[4];
            get(counts, sizeof(counts));
            if (counts[0] != sc.nf || counts[2] != sc.nb || counts[3] != sc.emitters.size() || counts[1] > sc.nf)
                throw SphError{SPH_E_ARG, "state file does not match the context's scenes"};
            q.na = counts[1];
            int32_t grid[7];
            uint32_t cells[3];
            get(grid, sizeof(grid));
            get(cells, sizeof(cells));
            std::copy(grid, grid + 3, q.g.g0);
            std::copy(grid + 3, grid + 6, q.g.gn);
            q.g.morton = grid[6];
            q.g.nCells = cells[0];
            q.g.keyRange = cells[1];
            q.g.ovBuckets = cells[2];
            // the grid must be one setGrid / growBuckets could have made
            uint64_t nc = 1;
            int maxDim = 1;
            for (int a = 0; a < 3; ++a) {
                if (q.g.gn[a] < 1 || q.g.gn[a] > (1 << 20) || q.g.g0[a] < -1000000000 ||
                    static_cast<int64_t>(q.g.g0[a]) + q.g.gn[a] > 1000000000)
                    throw bad("grid extent");
                nc *= static_cast<uint64_t>(q.g.gn[a]);
                maxDim = std::max(maxDim, q.g.gn[a]);
            }
            if (nc != q.g.nCells || nc > (static_cast<uint64_t>(1) << 28)) throw bad("cell count");
            const uint32_t ob = q.g.ovBuckets;
            if (ob == 0 || (ob & (ob - 1)) != 0 || ob > (1u << 24)) throw bad("overflow buckets");
            int bits = 0;
            while ((1 << bits) < maxDim) ++bits;
            uint64_t keys;
            if (q.g.morton == 1) {
                if (bits > 10) throw bad("Morton grid");
                keys = (static_cast<uint64_t>(1) << (3 * bits)) + 1;
            } else if (q.g.morton == 0) {
                keys = q.g.nCells + 1ull;
            } else {
                throw bad("key layout");
            }
            if (keys + ob != q.g.keyRange) throw bad("key range");
            float tt[2];
            int32_t its[2];
            get(tt, sizeof(tt));
            get(its, sizeof(its));
            q.t = tt[0];
            q.dt = tt[1];
            q.its = its[0];
            q.itsV = its[1];
            if (!std::isfinite(q.t) || !std::isfinite(q.dt) || !(q.dt > 0.0f)) throw bad("time");
            get(&q.st, sizeof(SceneState));
            if (!(q.st.dt > 0.0f) || !std::isfinite(q.st.t)) throw bad("scene state");
            int64_t left = 0;
            for (size_t e = 0; e < sc.emitters.size(); ++e) {
                int64_t rem = 0;
                float next = 0;
                get(&rem, 8);
                get(&next, 4);
                if (rem < 0 || rem > sc.emitters[e].P.num_particles || !std::isfinite(next)) throw bad("emitter");
                left += rem;
                q.emit.emplace_back(rem, next);
            }
            if (static_cast<int64_t>(q.na) + left != static_cast<int64_t>(sc.nf)) throw bad("active count");
        }
        std::vector<float4> hx(NP), hv(NP);
        std::vector<uint32_t> hid(NP);
        std::vector<float> hkw(NP), hkv(NP);
        get(hx.data(), NP * sizeof(float4));
        get(hv.data(), NP * sizeof(float4));
        get(hid.data(), NP * sizeof(uint32_t));
        get(hkw.data(), NP * sizeof(float));
        get(hkv.data(), NP * sizeof(float));
        // every scene's active slots hold a permutation of its active ids
        for (size_t i = 0; i < scenes.size(); ++i) {
            const SceneH& sc = scenes[i];
            std::vector<uint8_t> seen(L[i].na, 0);
            for (uint32_t k = 0; k < L[i].na; ++k) {
                const uint32_t o = hid[sc.fBegin + k];
                if (o >= L[i].na || seen[o]) throw bad("particle ids");
                seen[o] = 1;
                const float4 q = hx[sc.fBegin + k];
                if (!std::isfinite(q.x) || !std::isfinite(q.y) || !std::isfinite(q.z)) throw bad("positions");
            }
        }
        // commit
        HIPCHK(hipStreamSynchronize(stream));
        bool layout = false;
        for (size_t i = 0; i < scenes.size(); ++i) {
            SceneH& sc = scenes[i];
            const SceneLoad& q = L[i];
            GridSave cur;
            std::copy(sc.g0, sc.g0 + 3, cur.g0);
            std::copy(sc.gn, sc.gn + 3, cur.gn);
            cur.nCells = sc.nCells; cur.keyRange = sc.keyRange; cur.ovBuckets = sc.ovBuckets; cur.morton = sc.morton;
            if (!(cur == q.g)) {
                std::copy(q.g.g0, q.g.g0 + 3, sc.g0);
                std::copy(q.g.gn, q.g.gn + 3, sc.gn);
                sc.nCells = q.g.nCells; sc.keyRange = q.g.keyRange; sc.ovBuckets = q.g.ovBuckets;
                sc.morton = q.g.morton;
                layout = true;
            }
            sc.t = q.t;
            sc.dt = q.dt;
            sc.its = q.its;
            sc.itsV = q.itsV;
            for (size_t e = 0; e < sc.emitters.size(); ++e) {
                sc.emitters[e].remaining = q.emit[e].first;
                sc.emitters[e].next = q.emit[e].second;
            }
            sc.na = q.na;
            sc.fEnd = sc.fBegin + sc.na;
            hState[i] = q.st;
        }
        if (layout) uploadLayout();
        else uploadConsts();
        rangesDirty = true;
        rangesFull = true;
        HIPCHK(hipMemcpy(x.p, hx.data(), NP * sizeof(float4), hipMemcpyHostToDevice));
        uploadVel(0, hv);
        HIPCHK(hipMemcpy(id.p, hid.data(), NP * sizeof(uint32_t), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(kws.p, hkw.data(), NP * sizeof(float), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(kvws.p, hkv.data(), NP * sizeof(float), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(dState.p, hState, scenes.size() * sizeof(SceneState), hipMemcpyHostToDevice));
        ++stateEpoch;
        if (method == SPH_DFSPH) {
            neighborSearch(false);
            HIPCHK(launchDensityAlpha(view(), 1, stream));
        }
        HIPCHK(hipStreamSynchronize(stream));
    }

    void tapeAlloc(int steps, int nit, int halves) { tapeAllocInto(tape, steps, nit, halves, capF, capB); }
    void tapeAllocInto(Tape& tp, int steps, int nit, int halves, uint32_t cF, uint32_t cB) {
        const size_t k = static_cast<size_t>(steps), np = NP, it = static_cast<size_t>(nit);
        tp.steps = steps;
        tp.nit = nit;
        tp.halves = halves;
        tp.np = np;
        tp.capF = cF;
        tp.capB = cB;
        for (int hf = 0; hf < 2; ++hf) {
            HalfTape& t = tp.h[hf];
            if (hf >= halves) {
                t = HalfTape{};
                continue;
            }
            for (auto* b : {&t.x, &t.v, &t.n}) b->alloc(k * np);
            for (auto* b : {&t.rho, &t.p, &t.alpha}) b->alloc(k * np);
            for (auto* b : {&t.id, &t.cntF, &t.cntB}) b->alloc(k * np);
            t.nbrF.alloc(k * cF * np);
            t.nbrB.alloc(k * cB * np);
            t.u.alloc(k * std::max<size_t>(it, 1) * np);
            t.k.alloc(k * (it + 1) * np);
            t.aux.alloc(k * (it + 1) * np);
        }
    }
    // Device bytes of a tape of these dimensions (tapeAllocInto's arrays).
    static size_t tapeBytes(const Tape& tp) {
        const size_t k = static_cast<size_t>(tp.steps), np = tp.np, it = static_cast<size_t>(tp.nit);
        const size_t perHalf = k * np *
                               (3 * sizeof(float4) + 3 * sizeof(float) + 3 * sizeof(uint32_t) +
                                (static_cast<size_t>(tp.capF) + tp.capB) * sizeof(uint32_t) +
                                std::max<size_t>(it, 1) * sizeof(float4) + 2 * (it + 1) * sizeof(float));
        return perHalf * static_cast<size_t>(tp.halves);
    }
    // The list capacity grew while a segment was recorded: widen the tape's
    // list blocks, keeping the steps already recorded (entry s of particle i
    // sits at s * NP + i within its step's block for any capacity).
    void tapeGrow(int recorded) { tapeWiden(tape, recorded); }
    void tapeWiden(Tape& tp, int recorded) {
        flushCopies();   // no batched copy may target the blocks replaced here
        const size_t k = static_cast<size_t>(tp.steps), np = tp.np;
        auto widen = [&](DevBuf<uint32_t>& buf, uint32_t oldCap, uint32_t newCap) {
            if (newCap <= oldCap) return;
            DevBuf<uint32_t> w;
            w.alloc(k * newCap * np);
            for (int s = 0; s < recorded; ++s)
                copyD(w.p + s * newCap * np, buf.p + s * oldCap * np, oldCap * np * sizeof(uint32_t));
            buf = std::move(w);
        };
        HIPCHK(hipStreamSynchronize(stream));
        for (int hf = 0; hf < tp.halves; ++hf) {
            widen(tp.h[hf].nbrF, tp.capF, capF);
            widen(tp.h[hf].nbrB, tp.capB, capB);
        }
        HIPCHK(hipStreamSynchronize(stream));
        tp.capF = std::max(tp.capF, capF);
        tp.capB = std::max(tp.capB, capB);
    }
    // The list rows some particle of the last search uses: its longest list,
    // from the flags the host read after it, rounded up to whole kPadRows
    // rounds (the self-padding rows the adjoint Jacobi passes walk).
    void listRows(uint32_t& rowsF, uint32_t& rowsB) const {
        rowsF = 0;
        rowsB = 0;
        for (size_t si = 0; si < scenes.size(); ++si) {
            rowsF = std::max(rowsF, static_cast<uint32_t>(std::max(0, hFlags[si * kFlagSlots + kFlagMaxFluidNbr])));
            rowsB = std::max(rowsB, static_cast<uint32_t>(std::max(0, hFlags[si * kFlagSlots + kFlagMaxBoundaryNbr])));
        }
        rowsF = std::min((rowsF + kPadRows - 1u) / kPadRows * kPadRows, capF);   // with the self-padding rows
        rowsB = std::min(rowsB, capB);
    }
    // The search of step `slot` of a recorded segment (global step recBase +
    // slot). The recording checks it on the host (neighborSearch) and files
    // its flags, list rows and attempt count. A replay (the reverse sweep's
    // recompute) runs from the segment's checkpoint with the recording's grid
    // and at least its list capacities, so a step whose recorded search took
    // one attempt searches the same positions the same way: it is queued
    // without a host wait, and its flags are copied to hReplayFlags, which
    // replayCheck compares with the recorded ones once the segment has run.
    // A step whose search took more attempts is checked on the host again.
    void recordedSearch(int slot, bool dfsphAdvance, uint32_t& rowsF, uint32_t& rowsB) {
        const size_t g = static_cast<size_t>(recBase + slot);
        const size_t nf = scenes.size() * kFlagSlots;
        if (replaying && !rec.retried[g]) {
            searchOnce(dfsphAdvance);
            HIPCHK(hipMemcpyAsync(hReplayFlags + static_cast<size_t>(slot) * nf, dFlags, nf * sizeof(int32_t),
                                  hipMemcpyDeviceToHost, stream));
            replayChecks = std::max(replayChecks, slot + 1);
            rowsF = rec.rowsF[g];
            rowsB = rec.rowsB[g];
            return;
        }
        const int attempts = neighborSearch(dfsphAdvance);
        listRows(rowsF, rowsB);
        if (replaying) {
            if (!std::equal(hFlags, hFlags + nf, rec.flags.begin() + static_cast<std::ptrdiff_t>(g * nf)))
                throw SphError{SPH_E_STATE, "a recomputed step's search differs from the recorded one"};
            std::copy(hFlags, hFlags + nf, hReplayFlags + static_cast<size_t>(slot) * nf);
            replayChecks = std::max(replayChecks, slot + 1);
            return;
        }
        std::copy(hFlags, hFlags + nf, rec.flags.begin() + static_cast<std::ptrdiff_t>(g * nf));
        rec.rowsF[g] = rowsF;
        rec.rowsB[g] = rowsB;
        rec.retried[g] = attempts > 1;
    }
    // Start replaying the segment whose first step is global step `base`
    // (checking the flags of the segment replayed before it first).
    void beginReplay(int base) {
        replayCheck();
        const size_t nf = scenes.size() * kFlagSlots;
        const size_t need = static_cast<size_t>(rec.every) * nf;
        if (replayFlagCap < need) {
            if (hReplayFlags) HIPCHK(hipHostFree(hReplayFlags));
            hReplayFlags = nullptr;
            replayFlagCap = 0;
            HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&hReplayFlags), need * sizeof(int32_t)));
            replayFlagCap = need;
        }
        replaying = true;
        recBase = base;
        replayBase = base;
    }
    // Wait for the replayed segment and compare the flags of its searches
    // with the recording's: equal flags (integer counts of the search) say
    // that the recompute took the recorded path.
    void replayCheck() {
        replaying = false;
        if (!replayChecks) return;
        HIPCHK(hipStreamSynchronize(stream));
        const size_t nf = scenes.size() * kFlagSlots;
        const int n = replayChecks;
        replayChecks = 0;
        for (int s = 0; s < n; ++s) {
            const size_t g = static_cast<size_t>(replayBase + s);
            const int32_t* got = hReplayFlags + static_cast<size_t>(s) * nf;
            for (size_t si = 0; si < scenes.size(); ++si)
                if (got[si * kFlagSlots + kFlagNonFinite])
                    throw SphError{SPH_E_NONFINITE, "non-finite particle position"};
            if (!std::equal(got, got + nf, rec.flags.begin() + static_cast<std::ptrdiff_t>(g * nf)))
                throw SphError{SPH_E_STATE, "a recomputed step's search differs from the recorded one"};
        }
    }
    // x, ids and lists of the current search into a half of step `slot`;
    // of the lists only `rowsF` / `rowsB` rows (listRows)
    void recordSearch(HalfTape& t, int slot, uint32_t rowsF, uint32_t rowsB) {
        const size_t np = NP, o = static_cast<size_t>(slot) * np;
        tapeCopy(t.x.p + o, x.p, np * sizeof(float4));
        tapeCopy(t.id.p + o, id.p, np * sizeof(uint32_t));
        tapeCopy(t.cntF.p + o, cntF.p, np * sizeof(uint32_t));
        tapeCopy(t.cntB.p + o, cntB.p, np * sizeof(uint32_t));
        tapeCopy(t.nbrF.p + static_cast<size_t>(slot) * tape.capF * np, nbrF.p, rowsF * np * sizeof(uint32_t));
        tapeCopy(t.nbrB.p + static_cast<size_t>(slot) * tape.capB * np, nbrB.p, rowsB * np * sizeof(uint32_t));
    }
    // A DFSPH pass of iteration `it` (warm start: nit) writes its k and
    // rho* / divergence straight into the tape (no pass reads them back)
    DevView passView(const DevView& d, HalfTape& t, int slot, int it) {
        const size_t o = (static_cast<size_t>(slot) * (tape.nit + 1) + static_cast<size_t>(it)) * NP;
        DevView q = d;
        q.k = t.k.p + o;
        q.aux = t.aux.p + o;
        return q;
    }
    // a velocity pass that also stores its result at `rec` (the velocity the
    // next source pass reads)
    static DevView velView(const DevView& d, float4* rec) {
        DevView q = d;
        q.vRec = rec;
        return q;
    }
    float4* uSlot(HalfTape& t, int slot, int it) {
        return t.u.p + (static_cast<size_t>(slot) * tape.nit + static_cast<size_t>(it)) * NP;
    }
    // The tape block holding the search fields (x, ids, counts, lists) of a
    // half: a DFSPH step's first half runs on the search its predecessor's
    // second half ran on, so for slot > 0 that is recorded once, in the
    // second half of slot - 1 (with the densities and alpha it computed).
    std::pair<const HalfTape*, size_t> searchHalf(const Tape& tp, int slot, int half) const {
        if (method == SPH_DFSPH && half == 0 && slot > 0) return {&tp.h[1], static_cast<size_t>(slot - 1)};
        return {&tp.h[half], static_cast<size_t>(slot)};
    }
    // stepWCSPH with the tape of step `slot` of the current segment.
    void stepWCSPHRecord(int slot) {
        beginStats();
        uint32_t rowsF = 0, rowsB = 0;
        recordedSearch(slot, false, rowsF, rowsB);
        if (capF > tape.capF || capB > tape.capB) tapeGrow(slot);
        DevView d = view();
        HalfTape& t = tape.h[0];
        const size_t np = NP, o = static_cast<size_t>(slot) * np;
        recordSearch(t, slot, rowsF, rowsB);
        tapeCopy(t.v.p + o, v.p, np * sizeof(float4));
        timed(K_DENSITY, [&] { return launchDensityPressure(d, stream); });
        timed(K_FORCES, [&] { return launchForces(d, 1, stream); });
        if (anyST()) timed(K_SURFACE_TENSION, [&] { return launchSurfaceTension(d, 0, stream); });
        tapeCopy(t.rho.p + o, rho.p, np * sizeof(float));
        tapeCopy(t.p.p + o, p.p, np * sizeof(float));
        if (anyST()) tapeCopy(t.n.p + o, n.p, np * sizeof(float4));
        flushCopies();   // before the integration moves x and v
        timed(K_INTEGRATE, [&] { return launchIntegrate(d, stream); });
        if (replaying) return;   // no statistics: the recording filed them
        for (size_t s = 0; s < scenes.size(); ++s) hState[s].t = hState[s].t + hState[s].dt;
        fileStats(false);
    }
    // stepDFSPH with a fixed iteration count (no loop control), recording both
    // halves of step `slot`; the launches and their order are those of
    // stepDFSPH, so the trajectory is the same bit for bit.
    void stepDFSPHRecord(int slot) {
        beginStats();
        const int nit = tape.nit;
        const size_t np = NP, o = static_cast<size_t>(slot) * np;
        if (capF > tape.capF || capB > tape.capB) tapeGrow(slot);
        HalfTape& A = tape.h[0];
        HalfTape& B = tape.h[1];
        DevView d = view();
        // the first half's search fields, rho and alpha are the second half's
        // of slot - 1 (searchHalf), and its start velocity was stored by the
        // last velocity pass of slot - 1; only a segment's first step copies
        if (slot == 0) {
            uint32_t rowsF = 0, rowsB = 0;
            listRows(rowsF, rowsB);   // the flags of the search before this step (checked on the host)
            recordSearch(A, slot, rowsF, rowsB);
            tapeCopy(A.v.p + o, v.p, np * sizeof(float4));
            tapeCopy(A.rho.p + o, rho.p, np * sizeof(float));
            tapeCopy(A.alpha.p + o, alpha.p, np * sizeof(float));
        }
        if (anyVisc() || anyST()) timed(K_FORCES, [&] { return launchForces(d, 0, stream); });
        if (anyST()) timed(K_SURFACE_TENSION, [&] { return launchSurfaceTension(d, 1, stream); });
        else timed(K_VMAX, [&] { return launchVmax(d, stream); });
        finish(d, kFinishTimeStep, 2);
        if (anyST()) tapeCopy(A.n.p + o, n.p, np * sizeof(float4));
        flushCopies();   // before predictVelocities changes v
        timed(K_PREDICT, [&] { return launchPredictVelocity(d, stream); });
        timed(K_RHO_STAR, [&] { return launchRhoStar(passView(d, A, slot, nit), 1, 2, stream); });
        timed(K_VELOCITY, [&] { return launchVelocity(velView(d, uSlot(A, slot, 0)), 2, -1, stream); });
        for (int it = 0; it < nit; ++it) {
            timed(K_RHO_STAR, [&] { return launchRhoStar(passView(d, A, slot, it), 0, it & 1, stream); });
            if (!fuseFinish) finish(d, kFinishDensity, it & 1);
            velocityPass(velView(d, it + 1 < nit ? uSlot(A, slot, it + 1) : nullptr), it & 1, kFinishDensity);
        }
        for (size_t s = 0; s < scenes.size(); ++s) scenes[s].its = nit;   // the loop ran exactly nit iterations
        uint32_t rowsF = 0, rowsB = 0;
        recordedSearch(slot, true, rowsF, rowsB);
        if (capF > tape.capF || capB > tape.capB) tapeGrow(slot + 1);
        d = view();
        recordSearch(B, slot, rowsF, rowsB);
        timed(K_DENSITY, [&] { return launchDensityAlpha(d, 1, stream); });
        tapeCopy(B.rho.p + o, rho.p, np * sizeof(float));
        tapeCopy(B.alpha.p + o, alpha.p, np * sizeof(float));
        timed(K_DIVERGENCE, [&] { return launchDivergence(passView(d, B, slot, nit), 1, 2, stream); });
        timed(K_VELOCITY, [&] { return launchVelocity(velView(d, uSlot(B, slot, 0)), 2, -1, stream); });
        // the last velocity pass leaves the next step's start velocity
        float4* nextV = slot + 1 < tape.steps ? A.v.p + o + np : nullptr;
        for (int it = 0; it < nit; ++it) {
            timed(K_DIVERGENCE, [&] { return launchDivergence(passView(d, B, slot, it), 0, it & 1, stream); });
            if (!fuseFinish) finish(d, kFinishDivergence, it & 1);
            velocityPass(velView(d, it + 1 < nit ? uSlot(B, slot, it + 1) : nextV), it & 1, kFinishDivergence);
        }
        // aux of the context is the last pass's (sph_get_field reads it)
        tapeCopy(aux.p, B.aux.p + (o * (nit + 1) + static_cast<size_t>(nit - 1) * np), np * sizeof(float));
        flushCopies();   // x, lists, rho and alpha stay until the next step's keys pass
        if (replaying) {   // no statistics: the recording filed them
            timed(K_SCALE, [&] { return launchScaleDivergence(view(), stream); });
            return;
        }
        readState();
        for (size_t s = 0; s < scenes.size(); ++s) scenes[s].itsV = hState[s].it;
        timed(K_SCALE, [&] { return launchScaleDivergence(view(), stream); });
        for (size_t s = 0; s < scenes.size(); ++s) {
            SceneState& T = hState[s];
            const SceneConst_host c = hostConst(s);
            T.itsV = T.it;
            T.errV = T.err;
            if (T.it >= c.maxIt && T.err > static_cast<double>(c.etaV / T.dt)) T.capped |= 2;
            T.t = T.t + T.dt;
        }
        fileStats(true);
    }
    void stepRecord(int slot) {
        if (method == SPH_DFSPH) stepDFSPHRecord(slot);
        else stepWCSPHRecord(slot);
    }
    // Iterations per Jacobi solve of a recorded DFSPH rollout (0 for WCSPH).
    int checkDifferentiable() const {
        if (!initialized) throw SphError{SPH_E_STATE, "sph_init not called"};
        if (method == SPH_PCISPH) throw SphError{SPH_E_UNSUPPORTED, "gradients are available for WCSPH and DFSPH"};
        for (auto& sc : scenes) {
            if (!sc.emitters.empty()) throw SphError{SPH_E_UNSUPPORTED, "gradients of scenes with emitters"};
            if (sc.adhOn || !sc.cubes.empty() || sc.pciB)
                throw SphError{SPH_E_UNSUPPORTED, "gradients of scenes with adhesion, cube or PCISPH boundaries"};
        }
        if (!NP) throw SphError{SPH_E_STATE, "no fluid particles"};
        int nit = 0;
        for (auto& sc : scenes) {
            if (sc.viscOn && sc.P.viscosity_method != 2)
                throw SphError{SPH_E_UNSUPPORTED, "gradients support XSPH viscosity (method 2) only"};
            if (method != SPH_DFSPH) continue;
            // DFSPH: both solves run exactly maxIterations (the divergence loop
            // is bounded by maxIterations, DFSPHSolver.cpp:746) and dt is fixed
            const int n = static_cast<int>(sc.P.max_iterations);
            if (n < 1 || sc.P.min_iterations != sc.P.max_iterations || sc.P.min_iterations_v != sc.P.max_iterations)
                throw SphError{SPH_E_UNSUPPORTED,
                               "DFSPH gradients need min_iterations == min_iterations_v == max_iterations >= 1"};
            if (nit && n != nit) throw SphError{SPH_E_UNSUPPORTED, "DFSPH gradients need one iteration count per context"};
            nit = n;
            if (sc.P.adaptive_time_step && !(sc.minTs == sc.maxTs && sc.maxTs == sc.P.time_step))
                throw SphError{SPH_E_UNSUPPORTED, "DFSPH gradients need a fixed time step (min = max = initial)"};
        }
        return nit;
    }
    void forwardRecord(int nsteps, int every, sph_step_stats* stats) {
        const int nit = checkDifferentiable();
        if (nsteps < 1 || every < 1) throw SphError{SPH_E_ARG, "nsteps and checkpoint_every must be >= 1"};
        HIPCHK(hipSetDevice(device));
        every = std::min(every, nsteps);
        rec.valid = false;
        gradValid = false;
        rec.nsteps = nsteps;
        rec.every = every;
        rec.segments = (nsteps + every - 1) / every;
        rec.ckpt.resize(static_cast<size_t>(rec.segments));
        const int halves = method == SPH_DFSPH ? 2 : 1;
        if (tape.steps != every || tape.np != NP || tape.nit != nit || tape.halves != halves || tape.capF < capF ||
            tape.capB < capB)
            tapeAlloc(every, nit, halves);
        // With room for the reverse sweep's second tape, the segments go to
        // the two tapes in turn, so that the recording leaves the last two
        // segments taped and the sweep recomputes one segment fewer.
        rec.tapeSegRd = -1;
        const bool two = rec.segments > 1 && prepareOverlap();
        launchesInStep = 0;
        syncsInStep = 0;
        replaying = false;
        replayChecks = 0;
        rec.flags.assign(static_cast<size_t>(nsteps) * scenes.size() * kFlagSlots, 0);
        rec.rowsF.assign(static_cast<size_t>(nsteps), 0);
        rec.rowsB.assign(static_cast<size_t>(nsteps), 0);
        rec.retried.assign(static_cast<size_t>(nsteps), 1);
        HIPCHK(hipEventRecord(evStart, stream));
        for (int st = 0; st < nsteps;) {
            if (st % every == 0) {
                saveCheckpoint(rec.ckpt[static_cast<size_t>(st / every)]);
                if (two && st > 0) {
                    flushCopies();
                    std::swap(tape, tapeRd);
                }
            }
            recBase = st - st % every;
            // WCSPH: chunks of steps queued without host waits, within the segment
            int n = 1;
            if (method == SPH_WCSPH && canQueueWCSPH() && capF <= tape.capF && capB <= tape.capB) {
                n = std::min(kQueuedSteps, std::min(every - st % every, nsteps - st));
                if (queueWCSPH(n, stats ? stats + static_cast<size_t>(st) * scenes.size() : nullptr, st % every)) {
                    st += n;
                    continue;
                }
            }
            for (int j = 0; j < n; ++j, ++st) {
                stepRecord(st % every);
                if (stats)
                    for (size_t s = 0; s < scenes.size(); ++s) stats[st * scenes.size() + s] = lastStats[s];
            }
        }
        saveCheckpoint(rec.last);
        if (two) {
            // the list capacity may have grown after the segment before the
            // last was taped: widen its tape to the last one's dimensions
            if (tapeRd.capF < tape.capF || tapeRd.capB < tape.capB) tapeWiden(tapeRd, every);
            rec.tapeSegRd = rec.segments - 2;
        }
        HIPCHK(hipEventRecord(evStop, stream));
        HIPCHK(hipEventSynchronize(evStop));
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, evStart, evStop));
        lastStepMs = ms;
        collectProfile();
        rec.tapeSeg = rec.segments - 1;
        rec.epoch = ++stateEpoch;
        rec.valid = true;
        seedX.assign(scenes.size(), {});
        seedV.assign(scenes.size(), {});
    }
    void adjointSeed(int si, const float* gx, const float* gv) {
        SceneH& sc = scene(si);
        if (!rec.valid || rec.epoch != stateEpoch) throw SphError{SPH_E_STATE, "no recorded rollout for this state"};
        seedX[si].assign(3 * sc.nf, 0.0f);
        seedV[si].assign(3 * sc.nf, 0.0f);
        if (gx) std::copy(gx, gx + 3 * sc.nf, seedX[si].begin());
        if (gv) std::copy(gv, gv + 3 * sc.nf, seedV[si].begin());
    }
    AdjView adjView(int slot, int half) {
        AdjView a{};
        a.scenes = dScenesAdj.p;
        a.state = dStateAdj.p;
        a.blockScene = dBlockScene.p;
        a.numScenes = static_cast<int32_t>(scenes.size());
        a.numBlocks = numBlocks;
        a.NP = static_cast<uint32_t>(NP);
        const Tape& tp = *adjTape;
        const HalfTape& t = tp.h[half];
        const size_t o = static_cast<size_t>(slot) * NP;
        const auto sh = searchHalf(tp, slot, half);
        const HalfTape& ts = *sh.first;
        const size_t os = sh.second * NP;
        a.x = ts.x.p + os;
        a.v = t.v.p + o;
        a.rho = ts.rho.p + os;
        a.p = t.p.p + o;
        a.n = t.n.p + o;
        a.alpha = ts.alpha.p + os;
        a.id = ts.id.p + os;
        a.cntF = ts.cntF.p + os;
        a.cntB = ts.cntB.p + os;
        a.nbrF = ts.nbrF.p + sh.second * tp.capF * NP;
        a.nbrB = ts.nbrB.p + sh.second * tp.capB * NP;
        a.capF = tp.capF;
        a.capB = tp.capB;
        a.xb = xb.p;
        a.adjX = adjX.p;
        a.adjV = adjV.p;
        a.aBar = aBar.p;
        a.xBar = xBar.p;
        a.vBar = vBar.p;
        a.nBar = nBar.p;
        a.rhoBar = rhoBar.p;
        a.alphaBar = alphaBar.p;
        a.kBar = kBar.p;
        a.sBar = sBar.p;
        a.sgBar = sgBar.p;
        a.su = suBar.p;
        a.adjKws = adjKws.p;
        a.adjKvws = adjKvws.p;
        a.adjRho = adjRho.p;
        a.adjAlpha = adjAlpha.p;
        a.partials = adjPartials.p;
        a.acc = adjAcc.p;
        return a;
    }
    // k of pass `it` (warm start: it = nit) of a half of step `slot` on the tape
    const float* tapeK(int slot, int half, int it) const {
        const Tape& tp = *adjTape;
        return tp.h[half].k.p + (static_cast<size_t>(slot) * (tp.nit + 1) + it) * NP;
    }
    // point a DFSPH view at iteration `it` of its half (warm start: it = nit);
    // the passes run velocity then source from it = nit - 1 down to 0, then the
    // warm start's velocity pass, so the next velocity pass after the source
    // pass of `it` is it - 1, or the warm start's after it = 0
    void adjPass(AdjView& a, int slot, int half, int it) {
        const Tape& tp = *adjTape;
        const HalfTape& t = tp.h[half];
        const size_t ko = (static_cast<size_t>(slot) * (tp.nit + 1) + it) * NP;
        a.kk = t.k.p + ko;
        a.aux = t.aux.p + ko;
        a.u = it < tp.nit ? t.u.p + (static_cast<size_t>(slot) * tp.nit + it) * NP : nullptr;
        a.kkNext = it < tp.nit ? tapeK(slot, half, it > 0 ? it - 1 : tp.nit) : nullptr;
    }
    void adjointStep(int slot) {
        if (method == SPH_DFSPH) return adjointStepDFSPH(slot);
        const AdjView a = adjView(slot, 0);
        timed(K_ADJ_SEED, [&] { return launchAdjSeed(a, adjS); });
        timed(K_ADJ_FORCES, [&] { return launchAdjForces(a, adjS); });
        if (anyST()) timed(K_ADJ_NORMALS, [&] { return launchAdjNormals(a, adjS); });
        timed(K_ADJ_DENSITY, [&] { return launchAdjDensity(a, adjS); });
        timed(K_ADJ_FINISH, [&] { return launchAdjFinish(a, adjS); });
    }
    // Reverse of one DFSPH step: the second half (divergence solve, densities
    // and alpha at x_{n+1}), then the first half (density solve, prediction,
    // forces) at x_n.
    void adjointStepDFSPH(int slot) {
        const int nit = adjTape->nit;
        AdjView a = adjView(slot, 1);
        a.kkNext = tapeK(slot, 1, nit - 1);   // the half's first velocity pass
        timed(K_ADJ_SEED, [&] { return launchAdjHalfStart(a, 1, adjS); });
        for (int it = nit - 1; it >= 0; --it) {
            adjPass(a, slot, 1, it);
            timed(K_ADJ_VELOCITY, [&] { return launchAdjVelocity(a, kAdjDivergenceLoop, adjS); });
            timed(K_ADJ_SOURCE, [&] { return launchAdjSource(a, kAdjDivergenceLoop, adjS); });
        }
        adjPass(a, slot, 1, nit);
        timed(K_ADJ_VELOCITY, [&] { return launchAdjVelocity(a, kAdjDivergenceWarm, adjS); });
        timed(K_ADJ_DENSITY, [&] { return launchAdjAlphaA(a, adjS); });
        timed(K_ADJ_DENSITY, [&] { return launchAdjAlphaB(a, adjS); });
        timed(K_ADJ_SEED, [&] { return launchAdjHalfEnd(a, 1, adjS); });
        a = adjView(slot, 0);
        a.kkNext = tapeK(slot, 0, nit - 1);
        timed(K_ADJ_SEED, [&] { return launchAdjHalfStart(a, 0, adjS); });
        for (int it = nit - 1; it >= 0; --it) {
            adjPass(a, slot, 0, it);
            timed(K_ADJ_VELOCITY, [&] { return launchAdjVelocity(a, kAdjDensityLoop, adjS); });
            timed(K_ADJ_SOURCE, [&] { return launchAdjSource(a, kAdjDensityLoop, adjS); });
        }
        adjPass(a, slot, 0, nit);
        timed(K_ADJ_VELOCITY, [&] { return launchAdjVelocity(a, kAdjDensityWarm, adjS); });
        timed(K_ADJ_SEED, [&] { return launchAdjPredict(a, adjS); });
        if (anyVisc() || anyST()) timed(K_ADJ_FORCES, [&] { return launchAdjForces2(a, 0, adjS); });
        if (anyST()) timed(K_ADJ_NORMALS, [&] { return launchAdjNormals(a, adjS); });
        timed(K_ADJ_FINISH, [&] { return launchAdjFinish(a, adjS); });
        timed(K_ADJ_SEED, [&] { return launchAdjHalfEnd(a, 0, adjS); });
    }
    // The reverse sweep's second stream, its events and a second tape of the
    // recorded tape's dimensions (in tapeRd); false, for a sweep on one stream,
    // where device memory cannot hold the second tape.
    bool prepareOverlap() {
        if (!adjStream) {
            HIPCHK(hipStreamCreateWithFlags(&adjStream, hipStreamNonBlocking));
            HIPCHK(hipEventCreateWithFlags(&evAdjDone, hipEventDisableTiming));
            HIPCHK(hipEventCreateWithFlags(&evRecDone, hipEventDisableTiming));
        }
        Tape& sp = tapeRd;
        if (sp.steps == tape.steps && sp.np == tape.np && sp.nit == tape.nit && sp.halves == tape.halves &&
            sp.capF == tape.capF && sp.capB == tape.capB)
            return true;
        HIPCHK(hipStreamSynchronize(stream));
        sp = Tape{};
        rec.tapeSegRd = -1;   // whatever the second tape held is gone
        size_t freeB = 0, totalB = 0;
        HIPCHK(hipMemGetInfo(&freeB, &totalB));
        const size_t need = tapeBytes(tape);
        if (need + need / 8 + (static_cast<size_t>(1) << 30) > freeB) return false;
        try {
            tapeAllocInto(sp, tape.steps, tape.nit, tape.halves, tape.capF, tape.capB);
        } catch (const SphError&) {   // an allocation that failed after all: one stream
            sp = Tape{};
            (void)hipGetLastError();
            return false;
        }
        return true;
    }
    void adjointRun() {
        if (!rec.valid || rec.epoch != stateEpoch) throw SphError{SPH_E_STATE, "no recorded rollout for this state"};
        checkDifferentiable();
        HIPCHK(hipSetDevice(device));
        for (auto* b : {&adjX, &adjV, &aBar, &xBar, &vBar, &nBar, &sgBar, &suBar})
            if (b->n != NP) b->alloc(NP);
        for (auto* b : {&rhoBar, &alphaBar, &kBar, &sBar, &adjKws, &adjKvws, &adjRho, &adjAlpha})
            if (b->n != NP) b->alloc(NP);
        adjPartials.alloc(2 * static_cast<size_t>(numBlocks ? numBlocks : 1) * static_cast<size_t>(adjLanesPerParticle()));
        adjAcc.alloc(2 * scenes.size());
        std::vector<float4> hx(NP, make_float4(0, 0, 0, 0)), hv(NP, make_float4(0, 0, 0, 0));
        for (size_t si = 0; si < scenes.size(); ++si) {
            const SceneH& sc = scenes[si];
            for (uint32_t o = 0; o < sc.nf && !seedX[si].empty(); ++o) {
                hx[sc.fBegin + o] = make_float4(seedX[si][3 * o], seedX[si][3 * o + 1], seedX[si][3 * o + 2], 0.0f);
                hv[sc.fBegin + o] = make_float4(seedV[si][3 * o], seedV[si][3 * o + 1], seedV[si][3 * o + 2], 0.0f);
            }
        }
        HIPCHK(hipMemcpy(adjX.p, hx.data(), NP * sizeof(float4), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(adjV.p, hv.data(), NP * sizeof(float4), hipMemcpyHostToDevice));
        HIPCHK(hipMemset(adjAcc.p, 0, 2 * scenes.size() * sizeof(double)));
        for (auto* b : {&adjKws, &adjKvws, &adjRho, &adjAlpha}) HIPCHK(hipMemset(b->p, 0, NP * sizeof(float)));
        HIPCHK(hipEventRecord(evStart, stream));
        launchesInStep = 0;
        // the adjoint passes' own copies of the scene constants and states
        if (dScenesAdj.n != scenes.size()) dScenesAdj.alloc(scenes.size());
        if (dStateAdj.n != scenes.size()) dStateAdj.alloc(scenes.size());
        copyD(dScenesAdj.p, dScenes.p, scenes.size() * sizeof(SceneConst));
        copyD(dStateAdj.p, dState.p, scenes.size() * sizeof(SceneState));
        const bool overlap = rec.segments > 1 && prepareOverlap();
        try {
            sweep(overlap);
        } catch (...) {
            // leave no adjoint work in flight and no tape marked as holding a
            // segment it may not hold
            if (adjStream) (void)hipStreamSynchronize(adjStream);
            (void)hipStreamSynchronize(stream);
            adjS = stream;
            adjTape = &tape;
            rec.tapeSeg = -1;
            rec.tapeSegRd = -1;
            replaying = false;
            replayChecks = 0;
            try {   // back to where the recorded rollout ended, as after a sweep
                restoreCheckpoint(rec.last);
            } catch (...) {
            }
            throw;
        }
        HIPCHK(hipEventRecord(evStop, stream));
        HIPCHK(hipEventSynchronize(evStop));
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, evStart, evStop));
        lastStepMs = ms;
        collectProfile();
        finishAdjoint();
    }
    // The reverse sweep over every segment (see adjTape / adjS above); with
    // `overlap` the adjoint passes run on adjStream beside the recomputes.
    void sweep(bool overlap) {
        if (!overlap) {
            adjS = stream;
            adjTape = &tape;
            for (int seg = rec.segments - 1; seg >= 0; --seg) {
                const int first = seg * rec.every, count = std::min(rec.every, rec.nsteps - first);
                if (rec.tapeSeg != seg) {
                    restoreCheckpoint(rec.ckpt[static_cast<size_t>(seg)]);
                    beginReplay(first);
                    for (int j = 0; j < count; ++j) stepRecord(j);
                    replayCheck();
                    rec.tapeSeg = seg;
                }
                for (int j = count - 1; j >= 0; --j) adjointStep(j);
            }
        } else {
            // Segment seg's adjoint (adjStream, from tapeRd) runs beside the
            // recompute of segment seg - 1 (the engine's stream, into `tape`).
            // The recompute waits until the adjoint of seg + 1 has finished
            // reading `tape`; the adjoint of seg - 1 waits for its recompute.
            // A few adjoint steps are queued before the recompute's restore,
            // whose host waits would otherwise leave adjStream idle.
            const int last = rec.segments - 1;
            // the recording may have left the segment before the last on tapeRd
            const bool keptRd = rec.tapeSeg == last && rec.tapeSegRd == last - 1;
            rec.tapeSegRd = -1;
            if (rec.tapeSeg != last) {   // a second sweep over the same recording
                restoreCheckpoint(rec.ckpt[static_cast<size_t>(last)]);
                beginReplay(last * rec.every);
                for (int j = 0; j < rec.nsteps - last * rec.every; ++j) stepRecord(j);
                flushCopies();
                replayCheck();
                rec.tapeSeg = last;
            }
            adjS = adjStream;
            adjTape = &tapeRd;
            std::swap(tape, tapeRd);   // tapeRd: the last segment
            rec.tapeSeg = -1;          // `tape` holds no segment until a recompute fills it
            HIPCHK(hipEventRecord(evRecDone, stream));   // the set-up above
            HIPCHK(hipStreamWaitEvent(adjStream, evRecDone, 0));
            HIPCHK(hipEventRecord(evAdjDone, adjStream));
            constexpr int kLead = 2;
            for (int seg = rec.segments - 1; seg >= 0; --seg) {
                const int count = std::min(rec.every, rec.nsteps - seg * rec.every);
                // segments before the last are whole; the one before the last
                // needs no recompute if the recording kept it (it is in `tape`
                // after the swap above)
                const int countPrev = seg > 0 && !(keptRd && seg == last) ? rec.every : 0;
                int j = count - 1, k = 0;
                for (int l = 0; l < kLead && j >= 0; ++l) adjointStep(j--);
                if (countPrev) {
                    HIPCHK(hipStreamWaitEvent(stream, evAdjDone, 0));
                    restoreCheckpoint(rec.ckpt[static_cast<size_t>(seg - 1)]);
                    beginReplay((seg - 1) * rec.every);   // checks the segment replayed before
                }
                while (j >= 0 || k < countPrev) {
                    if (j >= 0) adjointStep(j--);
                    if (k < countPrev) stepRecord(k++);
                }
                HIPCHK(hipEventRecord(evAdjDone, adjStream));
                if (seg > 0) {
                    if (countPrev) flushCopies();
                    HIPCHK(hipEventRecord(evRecDone, stream));
                    HIPCHK(hipStreamWaitEvent(adjStream, evRecDone, 0));
                    std::swap(tape, tapeRd);   // tapeRd: segment seg - 1; tape: segment seg
                    rec.tapeSeg = seg;
                }
            }
            replayCheck();
        }
        if (method == SPH_DFSPH) {
            // rho and alpha of the first step came from the initial positions
            const AdjView a = adjView(0, 0);
            timed(K_ADJ_SEED, [&] { return launchAdjHalfStart(a, 2, adjS); });
            timed(K_ADJ_DENSITY, [&] { return launchAdjAlphaA(a, adjS); });
            timed(K_ADJ_DENSITY, [&] { return launchAdjAlphaB(a, adjS); });
            timed(K_ADJ_SEED, [&] { return launchAdjHalfEnd(a, 2, adjS); });
        }
        if (overlap) {
            // the engine's stream continues after the adjoint passes; the tape
            // of segment 0 becomes `tape` again (sph_recorded_field reads it)
            HIPCHK(hipEventRecord(evAdjDone, adjStream));
            HIPCHK(hipStreamWaitEvent(stream, evAdjDone, 0));
            std::swap(tape, tapeRd);
            rec.tapeSeg = 0;
        }
        adjS = stream;
        adjTape = &tape;
    }
    // Gradients to the host (original order), the context back where the
    // recorded rollout ended, and the finiteness check.
    void finishAdjoint() {
        std::vector<float4> hx = download(adjX.p, NP), hv = download(adjV.p, NP);
        gradCoef = download(adjAcc.p, 2 * scenes.size());
        gradX0.assign(scenes.size(), {});
        gradV0.assign(scenes.size(), {});
        for (size_t si = 0; si < scenes.size(); ++si) {
            const SceneH& sc = scenes[si];
            gradX0[si].resize(3 * sc.nf);
            gradV0[si].resize(3 * sc.nf);
            for (uint32_t o = 0; o < sc.nf; ++o) {
                const float4 q = hx[sc.fBegin + o], w = hv[sc.fBegin + o];
                gradX0[si][3 * o] = q.x; gradX0[si][3 * o + 1] = q.y; gradX0[si][3 * o + 2] = q.z;
                gradV0[si][3 * o] = w.x; gradV0[si][3 * o + 1] = w.y; gradV0[si][3 * o + 2] = w.z;
            }
        }
        // leave the context where the recorded rollout ended
        restoreCheckpoint(rec.last);
        gradValid = true;
        // the forward refuses non-finite states (SPH_E_NONFINITE); so does the
        // reverse sweep, after the results are stored for inspection
        int bad = -1;
        for (size_t si = 0; si < scenes.size() && bad < 0; ++si) {
            bool ok = std::isfinite(gradCoef[2 * si]) && std::isfinite(gradCoef[2 * si + 1]);
            for (float f : gradX0[si]) ok = ok && std::isfinite(f);
            for (float f : gradV0[si]) ok = ok && std::isfinite(f);
            if (!ok) bad = static_cast<int>(si);
        }
        if (bad >= 0)
            throw SphError{SPH_E_NONFINITE, "non-finite gradient in scene " + std::to_string(bad) +
                                                " (results kept for sph_adjoint_result)"};
    }
    // Values the recorded rollout held at one step (original order): see
    // sph_recorded_field in include/sph2o.h.
    void recordedField(int si, int step, int field, int pass, float* out) {
        SceneH& sc = scene(si);
        if (!rec.valid || rec.epoch != stateEpoch) throw SphError{SPH_E_STATE, "no recorded rollout for this state"};
        if (step < 0 || step >= rec.nsteps) throw SphError{SPH_E_ARG, "step outside the recorded rollout"};
        if (step / rec.every != rec.tapeSeg)
            throw SphError{SPH_E_STATE, "step not on the tape (record with checkpoint_every >= nsteps)"};
        if (field < 0 || field > 5) throw SphError{SPH_E_ARG, "unknown field"};
        const int half = (field == 2 || field == 3) ? 1 : 0;
        if (half >= tape.halves || (field <= 3 && method != SPH_DFSPH) || (field == 5 && method != SPH_WCSPH))
            throw SphError{SPH_E_ARG, "field not recorded for this method"};
        if (field <= 3 && (pass < 0 || pass > tape.nit)) throw SphError{SPH_E_ARG, "pass outside 0..iterations"};
        const HalfTape& t = tape.h[half];
        const size_t slot = static_cast<size_t>(step % rec.every);
        const auto sh = searchHalf(tape, static_cast<int>(slot), half);
        auto ids = download(sh.first->id.p + sh.second * NP + sc.fBegin, sc.nf);
        if (field == 4) {
            auto hx = download(sh.first->x.p + sh.second * NP + sc.fBegin, sc.nf);
            for (uint32_t i = 0; i < sc.nf; ++i) {
                out[3 * ids[i]] = hx[i].x; out[3 * ids[i] + 1] = hx[i].y; out[3 * ids[i] + 2] = hx[i].z;
            }
            return;
        }
        const float* src = field == 5 ? t.p.p + slot * NP
                                      : ((field == 0 || field == 2) ? t.k.p : t.aux.p) +
                                            (slot * (tape.nit + 1) + static_cast<size_t>(pass)) * NP;
        auto hv = download(src + sc.fBegin, sc.nf);
        for (uint32_t i = 0; i < sc.nf; ++i) out[ids[i]] = hv[i];
    }
    void adjointResult(int si, float* gx0, float* gv0, double* gnu, double* ggam) {
        scene(si);
        if (!gradValid) throw SphError{SPH_E_STATE, "sph_adjoint_run has not completed"};
        if (gx0) std::copy(gradX0[si].begin(), gradX0[si].end(), gx0);
        if (gv0) std::copy(gradV0[si].begin(), gradV0[si].end(), gv0);
        if (gnu) *gnu = gradCoef[2 * si];
        if (ggam) *ggam = gradCoef[2 * si + 1];
    }
    void setCoefficients(int si, float viscosity, float surfaceTension) {
        SceneH& sc = scene(si);
        if (!std::isfinite(viscosity) || !std::isfinite(surfaceTension))
            throw SphError{SPH_E_ARG, "non-finite coefficient"};
        sc.P.viscosity = viscosity;
        sc.P.surface_tension = surfaceTension;
        sc.viscOn = sc.P.viscosity_method > -1 && sc.P.viscosity != 0.0f;
        sc.stOn = sc.P.surface_tension_method == SPH_SURFACE_TENSION_AKINCI && sc.P.surface_tension != 0.0f;
        ++stateEpoch;
        if (initialized) uploadConsts();
    }

    void advanceEmpty() {
        for (size_t s = 0; s < scenes.size(); ++s) hState[s].t = hState[s].t + hState[s].dt;
        HIPCHK(hipMemcpyAsync(dState.p, hState, scenes.size() * sizeof(SceneState), hipMemcpyHostToDevice, stream));
        HIPCHK(hipStreamSynchronize(stream));
        fileStats(method == SPH_DFSPH);
    }

    struct SceneConst_host {
        int maxIt;
        float eta, etaV;
    };
    SceneConst_host hostConst(size_t s) const {
        return SceneConst_host{static_cast<int>(std::min<uint32_t>(scenes[s].P.max_iterations, 1u << 30)),
                               scenes[s].P.eta, scenes[s].P.eta_v};
    }

    // ------------------------------------------------------------ I/O
    template <class T>
    std::vector<T> download(const T* src, size_t count) {
        std::vector<T> out(count);
        HIPCHK(hipStreamSynchronize(stream));
        if (count) HIPCHK(hipMemcpy(out.data(), src, count * sizeof(T), hipMemcpyDeviceToHost));
        return out;
    }
    SceneH& scene(int s) {
        if (s < 0 || static_cast<size_t>(s) >= scenes.size()) throw SphError{SPH_E_ARG, "bad scene index"};
        return scenes[s];
    }

    void getState(int si, float* pos, float* vel, float* dens, float* pres) {
        SceneH& s = scene(si);
        if (!initialized) {
            if (pos) std::copy(s.xf.begin(), s.xf.end(), pos);
            if (vel) std::copy(s.vf.begin(), s.vf.end(), vel);
            if (dens) std::fill(dens, dens + s.nf, 0.0f);
            if (pres) std::fill(pres, pres + s.nf, 0.0f);
            return;
        }
        // active particles from the device; inactive ones keep their initial values
        const uint32_t na = s.na;
        auto ids = download(id.p + s.fBegin, na);
        if (pos || vel) {
            auto hx = pos ? download(x.p + s.fBegin, na) : std::vector<float4>();
            auto hv = vel ? downloadVel(s.fBegin, na) : std::vector<float4>();
            for (uint32_t i = 0; i < na; ++i) {
                const uint32_t o = ids[i];
                if (o >= na) throw SphError{SPH_E_STATE, "corrupt particle ids"};
                if (pos) { pos[3 * o] = hx[i].x; pos[3 * o + 1] = hx[i].y; pos[3 * o + 2] = hx[i].z; }
                if (vel) { vel[3 * o] = hv[i].x; vel[3 * o + 1] = hv[i].y; vel[3 * o + 2] = hv[i].z; }
            }
        }
        if (pos) std::copy(s.xf.begin() + 3 * static_cast<size_t>(na), s.xf.end(), pos + 3 * static_cast<size_t>(na));
        if (vel) std::copy(s.vf.begin() + 3 * static_cast<size_t>(na), s.vf.end(), vel + 3 * static_cast<size_t>(na));
        if (dens) {
            auto hr = download(rho.p + s.fBegin, na);
            for (uint32_t i = 0; i < na; ++i) dens[ids[i]] = hr[i];
            std::fill(dens + na, dens + s.nf, 0.0f);
        }
        if (pres) {
            auto hp = download(p.p + s.fBegin, na);
            for (uint32_t i = 0; i < na; ++i) pres[ids[i]] = hp[i];
            std::fill(pres + na, pres + s.nf, 0.0f);
        }
    }

    void getField(int si, int field, float* out) {
        SceneH& s = scene(si);
        if (!initialized) throw SphError{SPH_E_STATE, "sph_init not called"};
        const uint32_t na = s.na;   // inactive particles read 0
        auto ids = download(id.p + s.fBegin, na);
        if (field == 1 || field == 5) {
            auto hv = download((field == 1 ? a.p : n.p) + s.fBegin, na);
            std::fill(out, out + 3 * static_cast<size_t>(s.nf), 0.0f);
            for (uint32_t i = 0; i < na; ++i) {
                out[3 * ids[i]] = hv[i].x; out[3 * ids[i] + 1] = hv[i].y; out[3 * ids[i] + 2] = hv[i].z;
            }
            return;
        }
        const float* src = field == 0 ? alpha.p : field == 2 ? aux.p : field == 3 ? kws.p : field == 4 ? kvws.p : nullptr;
        if (!src) throw SphError{SPH_E_ARG, "unknown field"};
        auto hv = download(src + s.fBegin, na);
        std::fill(out, out + s.nf, 0.0f);
        for (uint32_t i = 0; i < na; ++i) out[ids[i]] = hv[i];
    }

    // Positions (x, y, z, 1) and |v| of the scene's active particles in
    // original order into device buffers, on the engine's stream. Buffers the
    // device cannot address (pageable host memory) are refused, not faulted on.
    void viewerFeed(int si, float* pos4, float* speed) {
        SceneH& s = scene(si);
        if (!initialized) throw SphError{SPH_E_STATE, "sph_init not called"};
        for (const void* ptr : {static_cast<const void*>(pos4), static_cast<const void*>(speed)}) {
            if (!ptr) continue;
            hipPointerAttribute_t at{};
            if (hipPointerGetAttributes(&at, ptr) != hipSuccess || at.devicePointer == nullptr) {
                (void)hipGetLastError();
                throw SphError{SPH_E_ARG, "viewer feed buffers must be device memory"};
            }
        }
        HIPCHK(hipSetDevice(device));
        HIPCHK(launchViewerFeed(view(), si, s.na, reinterpret_cast<float4*>(pos4), speed, stream));
    }

    void setState(int si, int64_t nn, const float* pos, const float* vel) {
        SceneH& s = scene(si);
        if (!s.emitters.empty()) throw SphError{SPH_E_UNSUPPORTED, "setting the state of a scene with emitters"};
        ++stateEpoch;
        if (nn != static_cast<int64_t>(s.nf)) throw SphError{SPH_E_ARG, "particle count mismatch"};
        if (pos)
            for (int64_t i = 0; i < 3 * nn; ++i)
                if (!std::isfinite(pos[i])) throw SphError{SPH_E_NONFINITE, "non-finite position"};
        if (!initialized) {
            if (pos) std::copy(pos, pos + 3 * nn, s.xf.begin());
            if (vel) std::copy(vel, vel + 3 * nn, s.vf.begin());
            return;
        }
        // The scene restarts from the original particle order (slot fBegin + o
        // holds particle o) and without DFSPH warm-start history (k_ws = kv_ws
        // = 0, as after sph_init): the next step then sorts the same input
        // whatever earlier steps left (the radix sort is stable within a cell),
        // so a state set twice steps bit-identically.
        auto ids = download(id.p + s.fBegin, s.nf);
        auto hx = download(x.p + s.fBegin, s.nf);
        auto hv = downloadVel(s.fBegin, s.nf);
        std::vector<float4> ox(s.nf), ov(s.nf);
        std::vector<float> okw(s.nf, 0.0f), okv(s.nf, 0.0f);
        std::vector<uint32_t> oid(s.nf);
        for (uint32_t i = 0; i < s.nf; ++i) {
            const uint32_t o = ids[i];
            if (o >= s.nf) throw SphError{SPH_E_STATE, "corrupt particle ids"};
            ox[o] = pos ? make_float4(pos[3 * o], pos[3 * o + 1], pos[3 * o + 2], 0.0f) : hx[i];
            ov[o] = vel ? make_float4(vel[3 * o], vel[3 * o + 1], vel[3 * o + 2], 0.0f) : hv[i];
            oid[o] = o;
        }
        hx.swap(ox);
        HIPCHK(hipMemcpy(x.p + s.fBegin, hx.data(), s.nf * sizeof(float4), hipMemcpyHostToDevice));
        uploadVel(s.fBegin, ov);
        HIPCHK(hipMemcpy(id.p + s.fBegin, oid.data(), s.nf * sizeof(uint32_t), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(kws.p + s.fBegin, okw.data(), s.nf * sizeof(float), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(kvws.p + s.fBegin, okv.data(), s.nf * sizeof(float), hipMemcpyHostToDevice));
        if (method == SPH_DFSPH) {  // density and alpha belong to the positions
            neighborSearch(false);
            HIPCHK(launchDensityAlpha(view(), 1, stream));
            HIPCHK(hipStreamSynchronize(stream));
        }
    }

    void dumpCells(int si, int32_t* cells) {
        SceneH& s = scene(si);
        if (!initialized) throw SphError{SPH_E_STATE, "sph_init not called"};
        DevBuf<int32_t> dc;
        dc.alloc(3 * (NP ? NP : 1));
        HIPCHK(launchCellsOut(view(), dc.p, stream));
        auto hc = download(dc.p + 3 * static_cast<size_t>(s.fBegin), 3 * static_cast<size_t>(s.na));
        auto ids = download(id.p + s.fBegin, s.na);
        std::fill(cells, cells + 3 * static_cast<size_t>(s.nf), 0);   // inactive particles: no cell
        for (uint32_t i = 0; i < s.na; ++i)
            for (int a3 = 0; a3 < 3; ++a3) cells[3 * ids[i] + a3] = hc[3 * i + a3];
    }

    // sph_dump_cells without any allocation: a kernel writes the keys in
    // original order into this device's `scratch`, one copy brings them to
    // the caller, the inactive tail is zeroed on the host.
    void dumpCellsNoAlloc(int si, int32_t* cells) {
        SceneH& s = scene(si);
        if (!initialized) throw SphError{SPH_E_STATE, "sph_init not called"};
        if (!cells) throw SphError{SPH_E_ARG, "null cells buffer"};
        if (s.na > kNoAllocMaxParticles)
            throw SphError{SPH_E_CAPACITY, "sph_dump_cells_noalloc handles at most 1048576 active particles per scene; "
                                           "this scene has " + std::to_string(s.na)};
        HIPCHK(hipSetDevice(device));
        const size_t bytes = 3 * static_cast<size_t>(s.na) * sizeof(int32_t);
        HIPCHK(launchCellsOutOriginal(view(), si, s.na, noAllocScratch, stream));
        if (bytes) HIPCHK(hipMemcpyAsync(cells, noAllocScratch, bytes, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        std::fill(cells + 3 * static_cast<size_t>(s.na), cells + 3 * static_cast<size_t>(s.nf), 0);
    }

    int dumpNeighbors(int si, int64_t* off, int32_t* ids_, int64_t cap, int64_t* boff, int32_t* bids, int64_t bcap) {
        SceneH& s = scene(si);
        if (!initialized) throw SphError{SPH_E_STATE, "sph_init not called"};
        neighborSearch(false);
        collectProfile();   // the search's kernels count in the profile like a step's
        auto ids = download(id.p, NP);
        // active particles only; inactive ones get empty lists
        const uint32_t na = s.na;
        auto cf = download(cntF.p + s.fBegin, na);
        auto cb = download(cntB.p + s.fBegin, na);
        auto lf = download(nbrF.p, static_cast<size_t>(capF) * NP);
        auto lb = download(nbrB.p, static_cast<size_t>(capB) * NP);
        std::vector<uint32_t> slotOf(na);
        for (uint32_t i = 0; i < na; ++i) slotOf[ids[s.fBegin + i]] = s.fBegin + i;
        off[0] = 0;
        boff[0] = 0;
        for (uint32_t o = 0; o < s.nf; ++o) {
            const uint32_t sl = o < na ? slotOf[o] - s.fBegin : 0u;
            off[o + 1] = off[o] + (o < na ? cf[sl] : 0u);
            boff[o + 1] = boff[o] + (o < na ? cb[sl] : 0u);
        }
        if (off[s.nf] > cap || boff[s.nf] > bcap) return SPH_E_CAPACITY;
        for (uint32_t o = 0; o < na; ++o) {
            const uint32_t slot = slotOf[o];
            const uint32_t sl = slot - s.fBegin;
            for (uint32_t q = 0; q < cf[sl]; ++q) ids_[off[o] + q] = static_cast<int32_t>(ids[lf[static_cast<size_t>(q) * NP + slot]]);
            for (uint32_t q = 0; q < cb[sl]; ++q)
                bids[boff[o] + q] = static_cast<int32_t>(s.bOrder[lb[static_cast<size_t>(q) * NP + slot] - s.bBegin]);
        }
        return SPH_OK;
    }

    void getBoundary(int si, float* pos, float* vol) {
        SceneH& s = scene(si);
        if (pos) std::copy(s.xb.begin(), s.xb.end(), pos);
        if (vol) {
            if (!initialized) throw SphError{SPH_E_STATE, "sph_init not called"};
            auto hv = download(vb.p + s.bBegin, s.nb);
            for (uint32_t q = 0; q < s.nb; ++q) vol[s.bOrder[q]] = hv[q];
        }
    }
};

}  // namespace

struct sph_ctx {
    Engine eng;
};

namespace {
template <class F>
int guarded(sph_ctx* ctx, F&& f) {
    if (!ctx) return SPH_E_ARG;
    try {
        return f();
    } catch (const SphError& e) {
        ctx->eng.lastError = e.msg;
        return e.code;
    } catch (const HipError& e) {
        ctx->eng.lastError = std::string("HIP error ") + hipGetErrorString(e.e) + " at " + e.where;
        return SPH_E_HIP;
    } catch (const std::bad_alloc&) {
        ctx->eng.lastError = "host out of memory";
        return SPH_E_HIP;
    } catch (...) {
        ctx->eng.lastError = "unknown error";
        return SPH_E_STATE;
    }
}
}  // namespace

extern "C" {

const char* sph_version(void) { return "sph2o-mi355x 0.1"; }

int sph_create(sph_ctx** ctx, int device, void* stream) {
    if (!ctx) return SPH_E_ARG;
    *ctx = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return SPH_E_HIP;
    if (device < 0 || device >= count) return SPH_E_ARG;
    auto* c = new (std::nothrow) sph_ctx();
    if (!c) return SPH_E_HIP;
    c->eng.device = device;
    c->eng.stream = static_cast<hipStream_t>(stream);
    if (hipSetDevice(device) != hipSuccess) {
        delete c;
        return SPH_E_HIP;
    }
    *ctx = c;
    return SPH_OK;
}

int sph_destroy(sph_ctx* ctx) {
    if (!ctx) return SPH_E_ARG;
    (void)hipSetDevice(ctx->eng.device);
    if (ctx->eng.stream) (void)hipStreamSynchronize(ctx->eng.stream);
    delete ctx;
    return SPH_OK;
}

const char* sph_last_error(const sph_ctx* ctx) { return ctx ? ctx->eng.lastError.c_str() : "null context"; }

int sph_add_scene(sph_ctx* ctx, const sph_scene_params* params, int64_t nf, const float* pos, const float* vel,
                  int64_t nb, const float* bpos) {
    return guarded(ctx, [&] {
        if (!params) throw SphError{SPH_E_ARG, "params is NULL"};
        return ctx->eng.addScene(*params, nf, pos, vel, nb, bpos);
    });
}

namespace {
// SPHSimulation::importScene (SPHSimulation.cpp:299-481) minus the object
// factory: JSON -> parameters + sampled fluid and boundary particles.
// The boundary of a scene of boundary method 1: per-particle normals and the
// object's normalFct / tangFct (SPHSimulation.cpp:428-455).
struct PciBoundaryDesc {
    std::vector<float> normals;
    float normalFct = 0.0f, tangFct = 0.0f;
};

void readScene(const char* path, sph_scene_params& params, std::vector<float>& xf, std::vector<float>& xbv,
               std::vector<sph_emitter_params>* emitters = nullptr, std::vector<BlockDesc>* cubes = nullptr,
               PciBoundaryDesc* pci = nullptr) {
    if (!path) throw SphError{SPH_E_ARG, "path is NULL"};
    std::ifstream f(path);
    if (!f) throw SphError{SPH_E_SCENE, std::string("cannot read ") + path};
    std::stringstream ss;
    ss << f.rdbuf();
    SceneDesc desc;
    std::string err;
    const int rc = parseSceneJson(ss.str(), desc, err);
    if (rc != SPH_OK) throw SphError{rc, err};
    if (desc.numGeometries > 0) throw SphError{SPH_E_UNSUPPORTED, "mesh (geometry) boundaries are not supported"};
    if (desc.params.boundary_method != SPH_BOUNDARY_AKINCI && desc.params.boundary_method != SPH_BOUNDARY_CUBE &&
        desc.params.boundary_method != SPH_BOUNDARY_PCISPH && (!desc.boxes.empty() || !desc.spheres.empty()))
        throw SphError{SPH_E_UNSUPPORTED, "only Akinci, PCISPH and cube boundaries are supported"};
    for (auto& b : desc.fluidBlocks) sampleFluidBlock(b, desc.params.particle_radius, xf);
    if (desc.params.boundary_method == SPH_BOUNDARY_AKINCI) {
        for (const int e : desc.shapeOrder) {
            if (e >= 0) sampleBoxSurface(desc.boxes[e], desc.params.particle_radius, xbv);
            else sampleSphereSurface(desc.spheres[-1 - e], desc.params.particle_radius, xbv);
        }
    }
    if (desc.params.boundary_method == SPH_BOUNDARY_PCISPH) {
        // PCISPHBoundaryObject::correctPositions walks the boundary neighbours of
        // every object with its own object's arrays (forsame_boundary_neighbors,
        // SPHSimulation.h:42-47), which is defined for one object only
        if (desc.numBoundaryObjects > 1)
            throw SphError{SPH_E_UNSUPPORTED, "the PCISPH boundary method is defined for one boundary object"};
        PciBoundaryDesc pb;
        if (desc.numBoundaryObjects == 1) {
            pb.normalFct = desc.normalFct[0];
            pb.tangFct = desc.tangFct[0];
        }
        const float supportRadius = 4.0f * desc.params.particle_radius;
        for (const int e : desc.shapeOrder) {
            if (e >= 0) {
                if (!samplePciBox(desc.boxes[e], supportRadius, desc.boxInverted[e] != 0, xbv, pb.normals))
                    throw SphError{SPH_E_UNSUPPORTED,
                                   "PCISPH boundary box under 1 m on an axis: the reference's sampler (integer abs of the "
                                   "extent, PCISPHBoundaryObject.cpp:62-64) samples one particle at 0/0 there"};
            } else {
                samplePciSphere(desc.spheres[-1 - e], desc.params.particle_radius, desc.sphereInverted[-1 - e] != 0, xbv,
                                pb.normals);
            }
        }
        if (pci) *pci = pb;
    }
    params = desc.params;
    if (emitters) *emitters = desc.emitters;
    if (desc.params.boundary_method == SPH_BOUNDARY_CUBE) {
        if (!cubes && !desc.cubes.empty())
            throw SphError{SPH_E_UNSUPPORTED, "the scene has cube boundaries: load it with sph_load_scene_json"};
        if (cubes) *cubes = desc.cubes;
    }
}
}  // namespace

int sph_load_scene_json(sph_ctx* ctx, const char* path) {
    return guarded(ctx, [&] {
        sph_scene_params params;
        std::vector<float> xf, xbv;
        std::vector<sph_emitter_params> em;
        std::vector<BlockDesc> cubes;
        PciBoundaryDesc pb;
        readScene(path, params, xf, xbv, &em, &cubes, &pb);
        const int id = ctx->eng.addScene(params, static_cast<int64_t>(xf.size() / 3), xf.data(), nullptr,
                                         static_cast<int64_t>(xbv.size() / 3), xbv.data());
        if (params.boundary_method == SPH_BOUNDARY_PCISPH)   // SPHSimulation.cpp:428-455
            ctx->eng.setBoundaryNormals(id, static_cast<int64_t>(pb.normals.size() / 3), pb.normals.data(), pb.normalFct,
                                        pb.tangFct);
        for (auto& e : em) ctx->eng.addEmitter(id, e);   // SPHSimulation.cpp:373-385
        for (auto& c : cubes) ctx->eng.addCube(id, c.min, c.max, 0.0f, 0.0f);   // SPHSimulation.cpp:457-474
        return id;
    });
}

int sph_scene_from_json(const char* path, sph_scene_params* params, int64_t* nf, float* pos, int64_t pos_cap,
                        int64_t* nb, float* bpos, int64_t bpos_cap) {
    sph_ctx dummy;  // error text sink
    return guarded(&dummy, [&] {
        if (!params || !nf || !nb) throw SphError{SPH_E_ARG, "NULL output"};
        std::vector<float> xf, xbv;
        std::vector<sph_emitter_params> em;
        readScene(path, *params, xf, xbv, &em);
        if (!em.empty()) throw SphError{SPH_E_UNSUPPORTED, "the scene has emitters: load it with sph_load_scene_json"};
        *nf = static_cast<int64_t>(xf.size() / 3);
        *nb = static_cast<int64_t>(xbv.size() / 3);
        if (*nf > pos_cap || *nb > bpos_cap) return SPH_E_CAPACITY;
        if (pos && !xf.empty()) std::memcpy(pos, xf.data(), xf.size() * sizeof(float));
        if (bpos && !xbv.empty()) std::memcpy(bpos, xbv.data(), xbv.size() * sizeof(float));
        return SPH_OK;
    });
}

int sph_scene_boundary_normals_from_json(const char* path, int64_t* nb, float* normals, int64_t cap, float* normal_fct,
                                         float* tangential_fct) {
    sph_ctx dummy;
    return guarded(&dummy, [&] {
        if (!nb || !normal_fct || !tangential_fct) throw SphError{SPH_E_ARG, "NULL output"};
        sph_scene_params params;
        std::vector<float> xf, xbv;
        std::vector<sph_emitter_params> em;
        PciBoundaryDesc pb;
        readScene(path, params, xf, xbv, &em, nullptr, &pb);
        if (params.boundary_method != SPH_BOUNDARY_PCISPH)
            throw SphError{SPH_E_ARG, "the scene's boundary method is not SPH_BOUNDARY_PCISPH"};
        *nb = static_cast<int64_t>(pb.normals.size() / 3);
        *normal_fct = pb.normalFct;
        *tangential_fct = pb.tangFct;
        if (*nb > cap) return SPH_E_CAPACITY;
        if (normals && *nb) std::memcpy(normals, pb.normals.data(), pb.normals.size() * sizeof(float));
        return SPH_OK;
    });
}

int sph_set_boundary_normals(sph_ctx* ctx, int scene, int64_t nb, const float* normals, float normal_fct,
                             float tangential_fct) {
    return guarded(ctx, [&] {
        ctx->eng.setBoundaryNormals(scene, nb, normals, normal_fct, tangential_fct);
        return SPH_OK;
    });
}

int sph_get_boundary_state(sph_ctx* ctx, int scene, float* density, float* pressure) {
    return guarded(ctx, [&] {
        ctx->eng.getBoundaryState(scene, density, pressure);
        return SPH_OK;
    });
}

int sph_num_scenes(const sph_ctx* ctx) { return ctx ? static_cast<int>(ctx->eng.scenes.size()) : SPH_E_ARG; }

int sph_get_params(const sph_ctx* ctx, int scene, sph_scene_params* out) {
    if (!ctx || !out || scene < 0 || static_cast<size_t>(scene) >= ctx->eng.scenes.size()) return SPH_E_ARG;
    *out = ctx->eng.scenes[scene].P;
    return SPH_OK;
}
int64_t sph_num_fluid(const sph_ctx* ctx, int scene) {
    if (!ctx || scene < 0 || static_cast<size_t>(scene) >= ctx->eng.scenes.size()) return SPH_E_ARG;
    return ctx->eng.scenes[scene].nf;
}
int64_t sph_num_active(const sph_ctx* ctx, int scene) {
    if (!ctx || scene < 0 || static_cast<size_t>(scene) >= ctx->eng.scenes.size()) return SPH_E_ARG;
    return ctx->eng.scenes[scene].na;
}
int sph_add_cube_boundary(sph_ctx* ctx, int scene, const float* min, const float* max, float normal_fct,
                          float tangential_fct) {
    return guarded(ctx, [&] {
        if (!min || !max) throw SphError{SPH_E_ARG, "NULL box"};
        ctx->eng.addCube(scene, min, max, normal_fct, tangential_fct);
        return SPH_OK;
    });
}
int sph_save_state(sph_ctx* ctx, const char* path) {
    return guarded(ctx, [&] {
        ctx->eng.saveState(path);
        return SPH_OK;
    });
}
int sph_load_state(sph_ctx* ctx, const char* path) {
    return guarded(ctx, [&] {
        ctx->eng.loadState(path);
        return SPH_OK;
    });
}
int sph_add_emitter(sph_ctx* ctx, int scene, const sph_emitter_params* emitter) {
    return guarded(ctx, [&] {
        if (!emitter) throw SphError{SPH_E_ARG, "emitter is NULL"};
        ctx->eng.addEmitter(scene, *emitter);
        return SPH_OK;
    });
}
int64_t sph_num_boundary(const sph_ctx* ctx, int scene) {
    if (!ctx || scene < 0 || static_cast<size_t>(scene) >= ctx->eng.scenes.size()) return SPH_E_ARG;
    return ctx->eng.scenes[scene].nb;
}

int sph_init(sph_ctx* ctx) {
    return guarded(ctx, [&] {
        ctx->eng.init();
        return SPH_OK;
    });
}

int sph_step(sph_ctx* ctx, int nsteps, sph_step_stats* stats) {
    return guarded(ctx, [&] {
        if (nsteps < 0) throw SphError{SPH_E_ARG, "nsteps < 0"};
        ctx->eng.stepAll(nsteps, stats);
        return SPH_OK;
    });
}

int sph_synchronize(sph_ctx* ctx) {
    return guarded(ctx, [&] {
        if (ctx->eng.stream) HIPCHK(hipStreamSynchronize(ctx->eng.stream));
        return SPH_OK;
    });
}

int sph_set_fluid_state(sph_ctx* ctx, int scene, int64_t nn, const float* pos, const float* vel) {
    return guarded(ctx, [&] {
        ctx->eng.setState(scene, nn, pos, vel);
        return SPH_OK;
    });
}

int sph_get_fluid_state(sph_ctx* ctx, int scene, float* pos, float* vel, float* density, float* pressure) {
    return guarded(ctx, [&] {
        ctx->eng.getState(scene, pos, vel, density, pressure);
        return SPH_OK;
    });
}

int sph_get_boundary(sph_ctx* ctx, int scene, float* pos, float* volume) {
    return guarded(ctx, [&] {
        ctx->eng.getBoundary(scene, pos, volume);
        return SPH_OK;
    });
}

int sph_get_field(sph_ctx* ctx, int scene, int field, float* out) {
    return guarded(ctx, [&] {
        if (!out) throw SphError{SPH_E_ARG, "out is NULL"};
        ctx->eng.getField(scene, field, out);
        return SPH_OK;
    });
}

int sph_viewer_feed(sph_ctx* ctx, int scene, float* pos4, float* speed) {
    return guarded(ctx, [&] {
        ctx->eng.viewerFeed(scene, pos4, speed);
        return SPH_OK;
    });
}

int sph_get_time(const sph_ctx* ctx, int scene, double* t, double* dt) {
    if (!ctx || scene < 0 || static_cast<size_t>(scene) >= ctx->eng.scenes.size()) return SPH_E_ARG;
    if (t) *t = ctx->eng.scenes[scene].t;
    if (dt) *dt = ctx->eng.scenes[scene].dt;
    return SPH_OK;
}

int sph_dump_cells(sph_ctx* ctx, int scene, int32_t* cells) {
    return guarded(ctx, [&] {
        if (!cells) throw SphError{SPH_E_ARG, "cells is NULL"};
        ctx->eng.dumpCells(scene, cells);
        return SPH_OK;
    });
}

int sph_dump_cells_noalloc(sph_ctx* ctx, int scene, int32_t* cells) {
    return guarded(ctx, [&] {
        ctx->eng.dumpCellsNoAlloc(scene, cells);
        return SPH_OK;
    });
}

int sph_dump_neighbors(sph_ctx* ctx, int scene, int64_t* offsets, int32_t* ids, int64_t cap, int64_t* boffsets,
                       int32_t* bids, int64_t bcap) {
    return guarded(ctx, [&] {
        if (!offsets || !boffsets || (cap > 0 && !ids) || (bcap > 0 && !bids)) throw SphError{SPH_E_ARG, "NULL output"};
        return ctx->eng.dumpNeighbors(scene, offsets, ids, cap, boffsets, bids, bcap);
    });
}

int sph_last_step_timing(const sph_ctx* ctx, double* total_ms, int64_t* launches) {
    if (!ctx) return SPH_E_ARG;
    if (total_ms) *total_ms = ctx->eng.lastStepMs;
    if (launches) *launches = ctx->eng.launchesInStep;
    return SPH_OK;
}

int sph_set_profiling(sph_ctx* ctx, int on) {
    return guarded(ctx, [&] {
        ctx->eng.profiling = on != 0;
        ctx->eng.profileMask = on == 1 || on == -1 ? ~0ull : static_cast<uint64_t>(static_cast<uint32_t>(on));
        if (!on) {
            std::fill(std::begin(ctx->eng.kernelMs), std::end(ctx->eng.kernelMs), 0.0);
            std::fill(std::begin(ctx->eng.kernelLaunches), std::end(ctx->eng.kernelLaunches), 0);
        }
        return SPH_OK;
    });
}

int sph_set_search_lanes(sph_ctx* ctx, int lanes) {
    return guarded(ctx, [&] {
        ctx->eng.setSearchLanes(lanes);
        return SPH_OK;
    });
}

int sph_get_search_lanes(const sph_ctx* ctx, int* lanes, int* tuned, float* trial_ms) {
    if (!ctx) return SPH_E_ARG;
    const auto& e = ctx->eng;
    if (lanes) *lanes = e.searchLanesFixed ? e.searchLanesFixed : e.searchLanes;
    if (tuned) *tuned = e.searchLanesFixed ? 0 : (e.tuner.done ? 1 : 0);
    if (trial_ms)
        for (int c = 0; c < 4; ++c) trial_ms[c] = e.tuner.best[c];
    return SPH_OK;
}

int sph_set_coefficients(sph_ctx* ctx, int scene, float viscosity, float surface_tension) {
    return guarded(ctx, [&] {
        ctx->eng.setCoefficients(scene, viscosity, surface_tension);
        return SPH_OK;
    });
}

int sph_forward_record(sph_ctx* ctx, int nsteps, int checkpoint_every, sph_step_stats* stats) {
    return guarded(ctx, [&] {
        ctx->eng.forwardRecord(nsteps, checkpoint_every, stats);
        return SPH_OK;
    });
}

int sph_adjoint_seed(sph_ctx* ctx, int scene, const float* dL_dx, const float* dL_dv) {
    return guarded(ctx, [&] {
        ctx->eng.adjointSeed(scene, dL_dx, dL_dv);
        return SPH_OK;
    });
}

int sph_adjoint_run(sph_ctx* ctx) {
    return guarded(ctx, [&] {
        ctx->eng.adjointRun();
        return SPH_OK;
    });
}

int sph_adjoint_result(sph_ctx* ctx, int scene, float* dL_dx0, float* dL_dv0, double* dL_dviscosity,
                       double* dL_dsurface_tension) {
    return guarded(ctx, [&] {
        ctx->eng.adjointResult(scene, dL_dx0, dL_dv0, dL_dviscosity, dL_dsurface_tension);
        return SPH_OK;
    });
}

int sph_recorded_field(sph_ctx* ctx, int scene, int step, int field, int pass, float* out) {
    return guarded(ctx, [&] {
        if (!out) throw SphError{SPH_E_ARG, "out is NULL"};
        ctx->eng.recordedField(scene, step, field, pass, out);
        return SPH_OK;
    });
}

int sph_kernel_count(void) { return K_COUNT; }
const char* sph_kernel_name(int kid) { return (kid >= 0 && kid < K_COUNT) ? kKernelNames[kid] : ""; }

int sph_kernel_stats(const sph_ctx* ctx, int kid, double* total_ms, int64_t* launches) {
    if (!ctx || kid < 0 || kid >= K_COUNT) return SPH_E_ARG;
    if (total_ms) *total_ms = ctx->eng.kernelMs[kid];
    if (launches) *launches = ctx->eng.kernelLaunches[kid];
    return SPH_OK;
}

int sph_reset_kernel_stats(sph_ctx* ctx) {
    if (!ctx) return SPH_E_ARG;
    std::fill(std::begin(ctx->eng.kernelMs), std::end(ctx->eng.kernelMs), 0.0);
    std::fill(std::begin(ctx->eng.kernelLaunches), std::end(ctx->eng.kernelLaunches), 0);
    return SPH_OK;
}

}  // extern "C"
