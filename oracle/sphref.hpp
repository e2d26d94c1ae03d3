// sphref — CPU restatement of the SPH2O particle step (TEST INFRASTRUCTURE).
//
// This is the oracle the HIP path is checked against. It is never linked
// into, loaded by, or called from the product library (sph2o_amd/); only
// tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use it.
//
// It restates, from the formulas and call order of the reference
// (josemanuel28v/SPH2O), one step of WCSPH, PCISPH and DFSPH with Akinci
// boundaries, XSPH / artificial / standard viscosity and Akinci surface
// tension. It is templated on the scalar type: Sim<double> is the FP64
// oracle, Sim<float> the FP32 one. Neighbour candidates come from a hash
// grid with cell side h and the 27-cell union around the particle's cell,
// walked in the reference's order (HashTable.cpp:23-86: dx, dy, dz with dz
// fastest, insertion order inside a cell), so per-particle sums run over the
// same candidates in the same order.
//
// Two numeric policies decide the neighbour sets:
//   * RefCompat (used by Sim<float>): the cell key is floor(x * fl(1/h)) and
//     the support test is sqrt(fma(x,x, fma(z,z, y*y))) <= h. This is what
//     the reference's `floor(pos / h)` (HashTable.cpp:5) and
//     `length(ri - rj) <= h` (DFSPHSolver.cpp:575) compute when built with
//     its shipped -O3 -ffast-math flags on an FMA host; both forms are pinned
//     by fixtures recorded from the compiled reference (tests/golden).
//   * IEEE (used by Sim<double>): floor(x / h) and sqrt((x*x + y*y) + z*z).
//
// Parity pins, by function:
//   kernels W / gradW / laplW / cohesion: closed-form KATs (tests).
//   neighbour sets and cell keys: golden fixtures from the compiled
//     reference + brute force (tests).
//   density, alpha, boundary volumes: lattice closed forms + compiled
//     reference (tests).
//   WCSPH / PCISPH / DFSPH steps: compiled reference trajectories (tests).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <unordered_map>
#include <vector>

namespace sphref {

template <class R>
struct V3 {
    R x, y, z;
    V3() : x(0), y(0), z(0) {}
    V3(R a, R b, R c) : x(a), y(b), z(c) {}
    V3 operator+(const V3& o) const { return V3(x + o.x, y + o.y, z + o.z); }
    V3 operator-(const V3& o) const { return V3(x - o.x, y - o.y, z - o.z); }
    V3 operator-() const { return V3(-x, -y, -z); }
    V3 operator*(R s) const { return V3(x * s, y * s, z * s); }
    V3 operator/(R s) const { return V3(x / s, y / s, z / s); }
    V3& operator+=(const V3& o) { x += o.x; y += o.y; z += o.z; return *this; }
    V3& operator-=(const V3& o) { x -= o.x; y -= o.y; z -= o.z; return *this; }
    V3& operator*=(R s) { x *= s; y *= s; z *= s; return *this; }
    V3& operator/=(R s) { x /= s; y /= s; z /= s; return *this; }
};
template <class R> inline V3<R> operator*(R s, const V3<R>& v) { return v * s; }
// glm::dot for vec3: tmp = a*b; (tmp.x + tmp.y) + tmp.z
template <class R> inline R dot(const V3<R>& a, const V3<R>& b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
// glm::length = sqrt(dot(a, a)) (IEEE form; the FP32 instance overrides it
// below with the form the reference's compiled code evaluates).
template <class R> inline R length(const V3<R>& a) { return std::sqrt(dot(a, a)); }
// Under the reference's shipped -O3 -ffast-math flags on an FMA host, GCC
// contracts glm::length(r) into sqrt(fma(x,x, fma(z,z, y*y))); pinned by
// fixtures recorded from the compiled reference. The FP32 oracle uses this
// form, which matters because Cohesion::W jumps by 2x at r = h/2
// (Cohesion.h:32-37 subtracts h^6/64 without its coefficient), so a 1-ulp
// change of |r| at the lattice spacing flips the branch.
template <> inline float length<float>(const V3<float>& a) {
    return std::sqrt(std::fma(a.x, a.x, std::fma(a.z, a.z, a.y * a.y)));
}

struct Params {
    int32_t method;            // 0 WCSPH, 1 PCISPH, 2 DFSPH (SPHSimulation.cpp:21-23)
    int32_t boundary_method;   // 2 Akinci, -1 none
    int32_t viscosity_method;  // -1 none, 0 standard, 1 artificial, 2 XSPH (SPHSimulation.cpp:29-31)
    int32_t st_method;         // -1 none, 3 Akinci 2013 (SPHSimulation.cpp:32)
    int32_t min_it, max_it, min_it_v, max_it_v;
    int32_t adaptive_dt;       // DFSPH: 1 = reference CFL/iteration rule (DFSPHSolver.cpp:1090-1162)
    int32_t pad0;
    double radius, density0, gx, gy, gz;
    double dt0, min_dt, max_dt, fps, cfl;
    double stiffness, gamma;
    double eta, eta_v;
    double viscosity, bviscosity, surface_tension;
    // Akinci adhesion (AdhesionForce.cpp:5-32) and cube boundary objects
    // (CubeBoundaryObject.cpp:10-88, boundary_method 0): lo xyz, hi xyz, normalFct, tangentialFct
    int32_t adhesion_on, n_cubes;
    double adhesion;
    double cubes[4][8];
};

// SplitMix64 finaliser (a fixed bijection of 64-bit words).
inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// Cell key: floor(pos/h) cast to int (HashTable.cpp:5).
struct Key {
    int x, y, z;
    bool operator==(const Key& o) const { return x == o.x && y == o.y && z == o.z; }
};
// HashTable.h:13-16 forms the same sum in int, which overflows (undefined
// behaviour) for cells from about x = 20, y = 7, z = 8 on: inlined into two
// loops the compiler may then hash one key two ways, and a lookup misses a
// cell that holds particles. The sum is taken modulo 2^32 here; the hash only
// places buckets, so results do not depend on it.
struct KeyHash {
    size_t operator()(const Key& k) const {
        const uint32_t s = static_cast<uint32_t>(k.x) * 73856093u + static_cast<uint32_t>(k.y) * 19349663u +
                           static_cast<uint32_t>(k.z) * 83492791u;
        return static_cast<size_t>(s);
    }
};

template <class R>
struct Policy;
template <>
struct Policy<float> {  // RefCompat
    static Key key(const V3<float>& p, float h) {
        const float inv = 1.0f / h;
        return Key{static_cast<int>(std::floor(p.x * inv)), static_cast<int>(std::floor(p.y * inv)),
                   static_cast<int>(std::floor(p.z * inv))};
    }
    static bool within(const V3<float>& r, float h) { return length(r) <= h; }
};
template <>
struct Policy<double> {  // IEEE
    static Key key(const V3<double>& p, double h) {
        return Key{static_cast<int>(std::floor(p.x / h)), static_cast<int>(std::floor(p.y / h)),
                   static_cast<int>(std::floor(p.z / h))};
    }
    static bool within(const V3<double>& r, double h) { return length(r) <= h; }
};

template <class R>
class Sim {
   public:
    using V = V3<R>;
    Params P{};
    // derived constants (SPHSimulation.cpp:143-155, :201-208; Kernels)
    R r = 0, h = 0, m = 0, rho0 = 0, dt = 0, t = 0, minTs = 0, maxTs = 0;
    V g;
    R kW = 0, lW = 0, laplC = 0, cohC1 = 0, cohC2 = 0;
    bool viscOn = false, stOn = false;
    // fluid state
    std::vector<V> x, v, a, n;
    std::vector<R> rho, p;
    // boundary
    std::vector<V> xb;
    std::vector<R> Vb;
    // grid
    struct Cell { std::vector<int> f, b; };
    std::unordered_map<Key, Cell, KeyHash> grid;
    // DFSPH
    std::vector<R> alpha, dErr, rhoStar, k, kv, kws, kvws;
    int its = 0, itsV = 0;
    R avgErr = 0, avgDiv = 0;
    const R eps = static_cast<R>(1e-5);
    // PCISPH
    std::vector<V> xs, vs, ap;
    R delta = 0, pciAvg = 0, pciMax = 0;
    int steps = 0;
    // Fingerprint of the pairs that took the inner (r <= h/2) branch of the
    // cohesion kernel, over all steps: the reference form jumps 2x there, so
    // finite differences are only meaningful between runs with equal prints.
    uint64_t innerHash = 0;
    int64_t innerPairs = 0;
    // Optional WCSPH pressure branch per step (1: p = B((rho/rho0)^g - 1),
    // 0: p = 0) replacing max(0, .) (WCSPHSolver.cpp:93) for steps 1..size():
    // finite differences then follow the branch another run (an FP32 one)
    // took at particles whose density sits within rounding of rho0.
    std::vector<std::vector<uint8_t>> pBranch;
    // Optional per-step FP32 positions (steps 1..size()) that decide each
    // pair's cohesion branch (|r| <= h/2 in the reference's FP32 length)
    // instead of this run's own positions, for the same purpose.
    std::vector<std::vector<V3<float>>> cohPos;
    // Optional DFSPH branches of steps 1..brSteps from another (FP32) run:
    // per step and pass (0..nit-1 the Jacobi iterations, nit the warm start)
    // that run's k and rho* (density solve) / clamped divergence (divergence
    // solve), [step][pass][N]. They decide the rho* >= rho0 and d >= 0 clamps,
    // the warm-start conditions and the |k_i + k_j| > eps, |k_i| > eps gates.
    std::vector<float> brKD, brSD, brKV, brSV;
    int brSteps = 0, brNit = 0;
    const float* branch(const std::vector<float>& b, int pass) const {
        if (steps < 1 || steps > brSteps || b.empty()) return nullptr;
        return b.data() + (static_cast<size_t>(steps - 1) * (brNit + 1) + pass) * x.size();
    }

    // ---- kernels (CubicSpline.h:27-74, ViscoK.h:29-37, Cohesion.h:20-40) ----
    void setKernels() {
        const R pi = static_cast<R>(M_PI);
        const R h3 = h * h * h;
        kW = static_cast<R>(8.0) / (pi * h3);
        lW = static_cast<R>(48.0) / (pi * h3);
        const R h6 = h * h * h * h * h * h;
        laplC = static_cast<R>(45.0) / (pi * h6);
        const R h9 = h * h * h * h * h * h * h * h * h;
        cohC1 = static_cast<R>(32.0 / M_PI) / h9;
        cohC2 = h6 / static_cast<R>(64.0);
    }
    R W(R rl) const {
        const R q = rl / h;
        if (q <= 1) {
            if (q <= static_cast<R>(0.5)) {
                const R q2 = q * q, q3 = q2 * q;
                return kW * (static_cast<R>(6) * q3 - static_cast<R>(6) * q2 + static_cast<R>(1));
            }
            const R f = 1 - q;
            return kW * (static_cast<R>(2) * (f * f * f));
        }
        return 0;
    }
    R W(const V& rv) const { return W(length(rv)); }
    V gradW(const V& rv) const {
        const R rl = length(rv);
        const R q = rl / h;
        if (rl > static_cast<R>(1e-9) && q <= 1) {
            V gq = rv / rl;
            gq /= h;
            if (q <= static_cast<R>(0.5)) return (lW * q * (static_cast<R>(3) * q - static_cast<R>(2))) * gq;
            const R f = 1 - q;
            return (lW * (-f * f)) * gq;
        }
        return V();
    }
    R laplW(const V& rv) const {
        const R mr = length(rv);
        return mr <= h ? laplC * (h - mr) : 0;
    }
    // Cohesion::W at length mr (0 < mr <= h) on a given branch
    R cohesionOn(R mr, bool inner) const {
        const R a1 = h - mr;
        if (inner) return cohC1 * static_cast<R>(2) * (a1 * a1 * a1) * (mr * mr * mr) - cohC2;
        return cohC1 * (a1 * a1 * a1) * (mr * mr * mr);
    }
    R cohesion(const V& rv) const {
        const R mr = length(rv);
        if (2 * mr > h && mr <= h) {
            const R a1 = h - mr;
            return cohC1 * (a1 * a1 * a1) * (mr * mr * mr);
        }
        if (mr > 0 && 2 * mr <= h) {
            const R a1 = h - mr;
            return cohC1 * static_cast<R>(2) * (a1 * a1 * a1) * (mr * mr * mr) - cohC2;
        }
        return 0;
    }

    // ---- grid (HashTable.cpp; SPHSolver.cpp:170-219) ----
    Key keyOf(const V& pos) const { return Policy<R>::key(pos, h); }
    void insertBoundary() {
        for (auto& c : grid) c.second.b.clear();
        for (int i = 0; i < static_cast<int>(xb.size()); ++i) grid[keyOf(xb[i])].b.push_back(i);
    }
    void insertFluid() {
        for (auto& c : grid) c.second.f.clear();
        for (int i = 0; i < static_cast<int>(x.size()); ++i) grid[keyOf(x[i])].f.push_back(i);
    }
    // visit the 27-cell candidates of `key` (dz fastest)
    template <class FF, class FB>
    void forCandidates(const Key& key, FF&& ff, FB&& fb) const {
        for (int dx = -1; dx <= 1; ++dx)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dz = -1; dz <= 1; ++dz) {
                    auto it = grid.find(Key{key.x + dx, key.y + dy, key.z + dz});
                    if (it == grid.end()) continue;
                    for (int j : it->second.f) ff(j);
                }
        for (int dx = -1; dx <= 1; ++dx)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dz = -1; dz <= 1; ++dz) {
                    auto it = grid.find(Key{key.x + dx, key.y + dy, key.z + dz});
                    if (it == grid.end()) continue;
                    for (int b : it->second.b) fb(b);
                }
    }
    template <class FF>
    void forFluidCandidates(const Key& key, FF&& ff) const {
        forCandidates(key, ff, [](int) {});
    }
    template <class FB>
    void forBoundaryCandidates(const Key& key, FB&& fb) const {
        forCandidates(key, [](int) {}, fb);
    }
    bool akinci() const { return P.boundary_method == 2 && !xb.empty(); }

    // ---- setup (SPHSimulation::init, initMasses; SPHSolver::init) ----
    void setup(const Params& prm, int nf, const float* xf, const float* vf, int nb, const float* xbf) {
        P = prm;
        r = static_cast<R>(P.radius);
        h = static_cast<R>(4.0) * r;  // SPHSimulation.cpp:206
        rho0 = static_cast<R>(P.density0);
        const R diam = static_cast<R>(2.0) * r;
        m = rho0 * (diam * diam * diam);  // SPHSimulation.cpp:145-153
        g = V(static_cast<R>(P.gx), static_cast<R>(P.gy), static_cast<R>(P.gz));
        dt = static_cast<R>(P.dt0);
        // TimeHelper.h:57
        const R fpsTs = static_cast<R>(1.0) / static_cast<R>(P.fps);
        minTs = fpsTs < static_cast<R>(P.min_dt) ? fpsTs : static_cast<R>(P.min_dt);
        maxTs = static_cast<R>(P.max_dt);
        // SPHSimulation.cpp:390-394: force objects only for method > -1 and coefficient != 0
        viscOn = P.viscosity_method > -1 && static_cast<R>(P.viscosity) != 0;
        stOn = P.st_method == 3 && static_cast<R>(P.surface_tension) != 0;
        setKernels();
        x.resize(nf); v.resize(nf); a.assign(nf, V()); n.assign(nf, V());
        rho.assign(nf, 0); p.assign(nf, 0);
        for (int i = 0; i < nf; ++i) {
            x[i] = V(xf[3 * i], xf[3 * i + 1], xf[3 * i + 2]);
            v[i] = vf ? V(vf[3 * i], vf[3 * i + 1], vf[3 * i + 2]) : V();
        }
        xb.resize(nb);
        for (int i = 0; i < nb; ++i) xb[i] = V(xbf[3 * i], xbf[3 * i + 1], xbf[3 * i + 2]);
        Vb.assign(nb, 0);
        alpha.assign(nf, 0); dErr.assign(nf, 0); rhoStar.assign(nf, 0);
        k.assign(nf, 0); kv.assign(nf, 0); kws.assign(nf, 0); kvws.assign(nf, 0);
        xs.assign(nf, V()); vs.assign(nf, V()); ap.assign(nf, V());
        t = 0; its = 0; itsV = 0; avgErr = 0; avgDiv = 0; pciAvg = 0; pciMax = 0; steps = 0;
        innerHash = 0; innerPairs = 0;
    }

    // DFSPH state that belongs to the positions (rho, alpha at step start) and
    // the warm-start history, after the positions were replaced
    void refreshDensities() {
        if (P.method != 2) return;
        insertFluid();
        computeDensities();
        computeAlpha();
        std::fill(kws.begin(), kws.end(), static_cast<R>(0));
        std::fill(kvws.begin(), kvws.end(), static_cast<R>(0));
    }
    void init() {
        if (P.method == 1) computeScalingFactor();  // PCISPHSolver.cpp:11-18
        grid.clear();
        insertBoundary();
        if (akinci()) computeVolumes();
        if (P.method == 2) {  // DFSPHSolver.cpp:22-31
            insertFluid();
            computeDensities();
            computeAlpha();
        }
    }

    // AkinciBoundaryObject::computeVolume (AkinciBoundaryObject.cpp:10-34)
    void computeVolumes() {
        const int nb = static_cast<int>(xb.size());
#pragma omp parallel for schedule(static)
        for (int i = 0; i < nb; ++i) {
            R delta_ = 0;
            const V& ri = xb[i];
            forBoundaryCandidates(keyOf(ri), [&](int b) { delta_ += W(ri - xb[b]); });
            Vb[i] = static_cast<R>(1.0) / delta_;
        }
    }

    // SPHSolver::computeFluidDensities (SPHSolver.cpp:42-89)
    void computeDensities() {
        const int nf = static_cast<int>(x.size());
        const bool ak = akinci();
#pragma omp parallel for schedule(static)
        for (int i = 0; i < nf; ++i) {
            const V& ri = x[i];
            R d = 0;
            forCandidates(
                keyOf(ri), [&](int j) { d += m * W(ri - x[j]); },
                [&](int b) { if (ak) d += rho0 * Vb[b] * W(ri - xb[b]); });
            rho[i] = d;
        }
    }

    // DFSPHSolver::computeAlpha (DFSPHSolver.cpp:109-178)
    void computeAlpha() {
        const int nf = static_cast<int>(x.size());
        const bool ak = akinci();
#pragma omp parallel for schedule(static)
        for (int i = 0; i < nf; ++i) {
            const V& ri = x[i];
            V sumGrad;
            R sumSq = 0;
            forCandidates(
                keyOf(ri),
                [&](int j) {
                    const V mg = m * gradW(ri - x[j]);
                    sumGrad += mg;
                    sumSq += dot(mg, mg);
                },
                [&](int b) {
                    if (ak) sumGrad += (rho0 * Vb[b]) * gradW(ri - xb[b]);
                });
            const R den = dot(sumGrad, sumGrad) + sumSq;
            alpha[i] = den > eps ? rho[i] / den : 0;
        }
    }

    // ---- non-pressure forces (SPHSimulation.cpp:256-265, order: viscosity, ST) ----
    void forces() {
        if (viscOn) {
            if (P.viscosity_method == 2) xsph();
            else if (P.viscosity_method == 1) artificial();
            else if (P.viscosity_method == 0) standardVisc();
        }
        if (stOn) surfaceTension();
        if (P.adhesion_on && akinci()) adhesion();
    }
    // AdhesionForce::step (AdhesionForce.cpp:5-32) with Adhesion::W as the
    // reference returns it (Adhesion.h:25-35, no valueCoef factor):
    // a -= beta rho0 V_b A(|r|) r / |r|, A = (-4 r^2/h + 6 r - 2h)^(1/4) on h/2 < |r| <= h
    void adhesion() {
        const int nf = static_cast<int>(x.size());
        const R beta = static_cast<R>(P.adhesion);
#pragma omp parallel for schedule(static)
        for (int i = 0; i < nf; ++i) {
            const V ri = x[i];
            V ai = a[i];
            forBoundaryCandidates(keyOf(ri), [&](int b) {
                const V rib = ri - xb[b];
                const R rl = length(rib);
                if (rl > 0 && 2 * rl > h && rl <= h) {
                    const R val = static_cast<R>(-4.0) * rl * rl / h + static_cast<R>(6.0) * rl - static_cast<R>(2.0) * h;
                    const R w = std::pow(val, static_cast<R>(0.25));
                    ai -= (beta * rho0 * Vb[b] * w) * rib / rl;
                }
            });
            a[i] = ai;
        }
    }
    // CubeBoundaryObject::correctPositionAndVelocity (CubeBoundaryObject.cpp:10-88),
    // every cube object in order, at the end of the step
    void cubeBoundary() {
        const int nf = static_cast<int>(x.size());
        for (int c = 0; c < P.n_cubes; ++c) {
            const double* q = P.cubes[c];
            const R nfac = static_cast<R>(q[6]), tfac = static_cast<R>(q[7]);
            for (int i = 0; i < nf; ++i) {
                R* pc[3] = {&x[i].x, &x[i].y, &x[i].z};
                R* vc[3] = {&v[i].x, &v[i].y, &v[i].z};
                for (int ax = 0; ax < 3; ++ax) {
                    const R lo = static_cast<R>(q[ax]), hi = static_cast<R>(q[3 + ax]);
                    R* o1 = vc[(ax + 1) % 3];
                    R* o2 = vc[(ax + 2) % 3];
                    if (*pc[ax] < lo) {
                        *pc[ax] = lo;
                        if (*vc[ax] < 0) *vc[ax] *= -nfac;
                        *o1 *= tfac;
                        *o2 *= tfac;
                    } else if (*pc[ax] > hi) {
                        *pc[ax] = hi;
                        if (*vc[ax] > 0) *vc[ax] *= -nfac;
                        *o1 *= tfac;
                        *o2 *= tfac;
                    }
                }
            }
        }
    }
    // XSPHViscosity::step (XSPHViscosity.cpp:6-42)
    void xsph() {
        const int nf = static_cast<int>(x.size());
        const R nu = static_cast<R>(P.viscosity), nub = static_cast<R>(P.bviscosity);
        const bool bOn = nub != 0 && akinci();
#pragma omp parallel for schedule(static)
        for (int i = 0; i < nf; ++i) {
            const V ri = x[i], vi = v[i];
            V ai = a[i];
            forFluidCandidates(keyOf(ri), [&](int j) {
                ai += ((nu * (m / rho[j])) * (v[j] - vi)) * W(ri - x[j]) / dt;
            });
            if (bOn)
                forBoundaryCandidates(keyOf(ri), [&](int b) {
                    ai += ((nub * (rho0 * Vb[b] / rho[i])) * (-vi)) * W(ri - xb[b]) / dt;
                });
            a[i] = ai;
        }
    }
    // ArtificialViscosity::step (ArtificialViscosity.cpp:6-53)
    void artificial() {
        const int nf = static_cast<int>(x.size());
        const R nu = static_cast<R>(P.viscosity), nub = static_cast<R>(P.bviscosity);
        const bool bOn = nub != 0 && akinci();
        const R h2 = h * h;
        const R c = static_cast<R>(10.0 * 0.1);
#pragma omp parallel for schedule(static)
        for (int i = 0; i < nf; ++i) {
            const V ri = x[i], vi = v[i];
            V ai = a[i];
            forFluidCandidates(keyOf(ri), [&](int j) {
                const V rij = ri - x[j];
                const R dp = dot(vi - v[j], rij);
                const R sq = dot(rij, rij);
                ai += (c * nu * (m / rho[j]) * dp / (sq + static_cast<R>(0.01) * h2)) * gradW(rij);
            });
            if (bOn)
                forBoundaryCandidates(keyOf(ri), [&](int b) {
                    const V rib = ri - xb[b];
                    const R dp = dot(vi, rib);
                    const R sq = dot(rib, rib);
                    ai += (c * nub * (rho0 * Vb[b] / rho[i]) * dp / (sq + static_cast<R>(0.01) * h2)) * gradW(rib);
                });
            a[i] = ai;
        }
    }
    // StandardViscosityForce::step (StandardViscosity.cpp:5-49)
    void standardVisc() {
        const int nf = static_cast<int>(x.size());
        const R nu = static_cast<R>(P.viscosity), nub = static_cast<R>(P.bviscosity);
        const bool bOn = nub != 0 && akinci();
#pragma omp parallel for schedule(static)
        for (int i = 0; i < nf; ++i) {
            const V ri = x[i], vi = v[i];
            V f;
            forFluidCandidates(keyOf(ri), [&](int j) { f += (m * (v[j] - vi) / rho[j]) * laplW(ri - x[j]); });
            f *= nu / rho[i];
            V ai = a[i] + f;
            if (bOn) {
                V fb;
                forBoundaryCandidates(keyOf(ri), [&](int b) {
                    fb += ((rho0 * Vb[b]) * (-vi) / rho[i]) * laplW(ri - xb[b]);
                });
                fb *= nub / rho[i];
                ai += fb;
            }
            a[i] = ai;
        }
    }
    // SurfaceTensionAkinci::step (SurfaceTensionAkinci.cpp:5-71)
    void surfaceTension() {
        const int nf = static_cast<int>(x.size());
        const R gam = static_cast<R>(P.surface_tension);
#pragma omp parallel for schedule(static)
        for (int i = 0; i < nf; ++i) {
            const V ri = x[i];
            V ni;
            forFluidCandidates(keyOf(ri), [&](int j) { ni += (m / rho[j]) * gradW(ri - x[j]); });
            ni *= h;
            n[i] = ni;
        }
        uint64_t hsh = 0;
        int64_t cnt = 0;
        const uint64_t stepKey = static_cast<uint64_t>(steps) << 42;
        const V3<float>* bp = steps - 1 < static_cast<int>(cohPos.size()) ? cohPos[steps - 1].data() : nullptr;
        const float halfH = 0.5f * static_cast<float>(h);
#pragma omp parallel for schedule(static) reduction(+ : hsh, cnt)
        for (int i = 0; i < nf; ++i) {
            const V ri = x[i], ni = n[i];
            const R di = rho[i];
            V ai = a[i];
            forFluidCandidates(keyOf(ri), [&](int j) {
                const V rij = ri - x[j];
                const R rl = length(rij);
                const R kij = static_cast<R>(2.0) * rho0 / (di + rho[j]);
                if (rl > 0 && rl <= h) {
                    const bool inner = bp ? length(bp[i] - bp[j]) <= halfH : 2 * rl <= h;
                    V acc;
                    acc -= (gam * m * (bp ? cohesionOn(rl, inner) : cohesion(rij))) * rij / rl;
                    acc -= gam * (ni - n[j]);
                    ai += kij * acc;
                    if (inner) {
                        hsh += mix64(stepKey ^ (static_cast<uint64_t>(i) << 21) ^ static_cast<uint64_t>(j));
                        ++cnt;
                    }
                }
            });
            a[i] = ai;
        }
        innerHash += hsh;
        innerPairs += cnt;
    }

    // ---- WCSPH (WCSPHSolver.cpp:14-188; SPHSolver::integrate :139-168) ----
    void stepWCSPH() {
        ++steps;
        insertFluid();
        computeDensities();
        const int nf = static_cast<int>(x.size());
        const R kk = static_cast<R>(P.stiffness), gm = static_cast<R>(P.gamma);
        const int k = steps - 1;
        const uint8_t* mask = k < static_cast<int>(pBranch.size()) ? pBranch[k].data() : nullptr;
#pragma omp parallel for schedule(static)
        for (int i = 0; i < nf; ++i) {
            const R raw = kk * rho0 / gm * (std::pow(rho[i] / rho0, gm) - 1);
            p[i] = mask ? (mask[i] ? raw : static_cast<R>(0)) : std::max(static_cast<R>(0), raw);
        }
        const bool ak = akinci();
#pragma omp parallel for schedule(static)
        for (int i = 0; i < nf; ++i) {
            const V ri = x[i];
            V ai = a[i];
            const R dpi = p[i] / (rho[i] * rho[i]);
            const R dpb = p[i] / (rho0 * rho0);
            forCandidates(
                keyOf(ri), [&](int j) { ai -= (m * (dpi + p[j] / (rho[j] * rho[j]))) * gradW(ri - x[j]); },
                [&](int b) { if (ak) ai -= (rho0 * Vb[b] * (dpi + dpb)) * gradW(ri - xb[b]); });
            a[i] = ai;
        }
        forces();
        integrateEuler(nullptr);
        cubeBoundary();   // WCSPHSolver.cpp:51-61
        t += dt;
    }
    void integrateEuler(const std::vector<V>* extra) {
        const int nf = static_cast<int>(x.size());
#pragma omp parallel for schedule(static)
        for (int i = 0; i < nf; ++i) {
            V ai = a[i];
            if (extra) ai += g + (*extra)[i];  // PCISPHSolver.cpp:478
            else ai += g;                       // SPHSolver.cpp:158
            v[i] += ai * dt;
            x[i] += v[i] * dt;
            a[i] = V();
        }
    }

    // ---- PCISPH (PCISPHSolver.cpp) ----
    void computeScalingFactor() {  // :403-436
        const R dist = static_cast<R>(2.0) * r;
        V sg;
        R ssq = 0;
        for (int i = -1; i <= 1; i++)
            for (int j = -1; j <= 1; j++)
                for (int kk = -1; kk <= 1; kk++) {
                    V rij(static_cast<R>(-i), static_cast<R>(-j), static_cast<R>(-kk));
                    rij *= dist;
                    const V gr = gradW(rij);
                    sg += gr;
                    ssq += dot(gr, gr);
                }
        const R mdt = m * dt / rho0;
        const R beta = static_cast<R>(2.0) * (mdt * mdt);
        delta = -static_cast<R>(1.0) / (beta * (-dot(sg, sg) - ssq));
    }
    void stepPCISPH() {
        ++steps;
        insertFluid();
        computeDensities();
        forces();
        const int nf = static_cast<int>(x.size());
        for (int i = 0; i < nf; ++i) { p[i] = 0; ap[i] = V(); }
        const bool ak = akinci();
        its = 0;
        while ((pciAvg > static_cast<R>(P.eta) || its < P.min_it) && its < P.max_it) {
#pragma omp parallel for schedule(static)
            for (int i = 0; i < nf; ++i) {  // :120-147
                const V acc = a[i] + ap[i] + g;
                vs[i] = v[i] + acc * dt;
                xs[i] = x[i] + vs[i] * dt;
            }
#pragma omp parallel for schedule(static)
            for (int i = 0; i < nf; ++i) {  // :162-209 (cells of the current position)
                const V ri = xs[i];
                R d = 0;
                forCandidates(
                    keyOf(x[i]), [&](int j) { d += m * W(ri - xs[j]); },
                    [&](int b) { if (ak) d += rho0 * Vb[b] * W(ri - xb[b]); });
                rho[i] = d;
            }
            double sum = 0;
            R mx = 0;
            for (int i = 0; i < nf; ++i) {  // :246-298
                const R e = std::max(static_cast<R>(0), rho[i] - rho0);
                p[i] += delta * e;
                mx = std::max(mx, e / rho0);
                sum += static_cast<double>(e / rho0);
            }
            pciAvg = static_cast<R>(sum / nf);
            pciMax = mx;
#pragma omp parallel for schedule(static)
            for (int i = 0; i < nf; ++i) {  // :300-367
                const V ri = x[i];
                V pa;
                const R dpi = p[i] / (rho[i] * rho[i]);
                forCandidates(
                    keyOf(ri), [&](int j) { pa -= (m * (dpi + p[j] / (rho[j] * rho[j]))) * gradW(ri - x[j]); },
                    [&](int b) { if (ak) pa -= (rho0 * Vb[b] * dpi) * gradW(ri - xb[b]); });
                ap[i] = pa;
            }
            ++its;
        }
        integrateEuler(&ap);
        cubeBoundary();   // PCISPHSolver.cpp:53-61
        t += dt;
    }

    // ---- DFSPH (DFSPHSolver.cpp) ----
    double maxVelocityPredicted() const {  // :1096-1113
        const int nf = static_cast<int>(x.size());
        double mv = 0;
        for (int i = 0; i < nf; ++i) {
            const V acc = a[i] + g;
            const V cv = v[i] + acc * dt;
            mv = std::max(mv, static_cast<double>(length(cv)));
        }
        return mv;
    }
    void updateTimeStep() {  // :1090-1162
        R maxVel = static_cast<R>(maxVelocityPredicted());
        if (maxVel < static_cast<R>(1.0e-9)) maxVel = static_cast<R>(1.0e-9);
        R hh = dt;
        const R cflTs = static_cast<R>(P.cfl) * static_cast<R>(0.4) * (static_cast<R>(2.0) * r) / maxVel;
        const int itsMax = std::max(its, itsV);
        if (itsMax > 6) hh *= static_cast<R>(0.9);
        else if (itsMax < 4) hh *= static_cast<R>(1.05);
        hh = std::min(hh, cflTs);
        hh = std::max(hh, minTs);
        hh = std::min(hh, maxTs);
        dt = hh;
    }
    void predictDensities(const float* clamp = nullptr) {  // :226-299
        const int nf = static_cast<int>(x.size());
        const bool ak = akinci();
#pragma omp parallel for schedule(static)
        for (int i = 0; i < nf; ++i) {
            const V ri = x[i], vi = v[i];
            R dc = 0;
            forCandidates(
                keyOf(ri), [&](int j) { dc += m * dot(vi - v[j], gradW(ri - x[j])); },
                [&](int b) { if (ak) dc += rho0 * Vb[b] * dot(vi, gradW(ri - xb[b])); });
            R pd = rho[i] + dc * dt;
            rhoStar[i] = clamp ? (clamp[i] > static_cast<float>(rho0) ? pd : rho0) : std::max(pd, rho0);
        }
        double s = 0;
        for (int i = 0; i < nf; ++i) s += static_cast<double>((rhoStar[i] - rho0) / rho0);
        avgErr = nf ? static_cast<R>(s / nf) : 0;
    }
    // velocity correction shared by both solves and both warm starts
    void velocityPass(const std::vector<R>& kk, const float* gate = nullptr) {
        const int nf = static_cast<int>(x.size());
        const bool ak = akinci();
#pragma omp parallel for schedule(static)
        for (int i = 0; i < nf; ++i) {
            const V ri = x[i];
            const R ki = kk[i];
            V sum;
            forFluidCandidates(keyOf(ri), [&](int j) {
                const R ks = ki + kk[j];
                const bool open = gate ? std::fabs(gate[i] + gate[j]) > static_cast<float>(eps) : std::fabs(ks) > eps;
                if (open) sum += (m * ks) * gradW(ri - x[j]);
            });
            if ((gate ? std::fabs(gate[i]) > static_cast<float>(eps) : std::fabs(ki) > eps) && ak)
                forBoundaryCandidates(keyOf(ri), [&](int b) { sum += (rho0 * Vb[b] * ki) * gradW(ri - xb[b]); });
            v[i] -= dt * sum;
        }
    }
    void correctDensityError() {  // :414-536 with densityWarmStart :851-955
        const int nf = static_cast<int>(x.size());
        const R ts2 = dt * dt;
        predictDensities(branch(brSD, brNit));
        for (int i = 0; i < nf; ++i) kws[i] = rhoStar[i] > rho0 ? kws[i] / ts2 : 0;
        velocityPass(kws, branch(brKD, brNit));
        std::fill(kws.begin(), kws.end(), static_cast<R>(0));
        its = 0;
        avgErr = 0;
        while ((avgErr > static_cast<R>(P.eta) || its < P.min_it) && its < P.max_it) {
            predictDensities(its < brNit ? branch(brSD, its) : nullptr);
            for (int i = 0; i < nf; ++i) {
                k[i] = (rhoStar[i] - rho0) * alpha[i] / (ts2 * rho[i]);
                kws[i] += k[i];
            }
            velocityPass(k, its < brNit ? branch(brKD, its) : nullptr);
            ++its;
        }
        for (int i = 0; i < nf; ++i) kws[i] *= ts2;
    }
    void computeDivergenceError(const float* clamp = nullptr) {  // :538-621
        const int nf = static_cast<int>(x.size());
        const bool ak = akinci();
#pragma omp parallel for schedule(static)
        for (int i = 0; i < nf; ++i) {
            const V ri = x[i], vi = v[i];
            R de = 0;
            int nn = 0;
            forCandidates(
                keyOf(ri),
                [&](int j) {
                    const V rr = ri - x[j];
                    de += m * dot(vi - v[j], gradW(rr));
                    if (Policy<R>::within(rr, h)) nn++;
                },
                [&](int b) {
                    if (!ak) return;
                    const V rr = ri - xb[b];
                    de += rho0 * Vb[b] * dot(vi, gradW(rr));
                    if (Policy<R>::within(rr, h)) nn++;
                });
            dErr[i] = clamp ? (clamp[i] > 0.0f ? de : static_cast<R>(0)) : nn < 20 ? 0 : std::max(de, static_cast<R>(0));
        }
        double s = 0;
        for (int i = 0; i < nf; ++i) s += static_cast<double>(dErr[i] / rho0);
        avgDiv = static_cast<R>(s / nf);
    }
    void correctDivergenceError() {  // :730-849 with divergenceWarmStart :957-1059
        const int nf = static_cast<int>(x.size());
        computeDivergenceError(branch(brSV, brNit));
        for (int i = 0; i < nf; ++i) kvws[i] = dErr[i] > 0 ? kvws[i] / dt : 0;
        velocityPass(kvws, branch(brKV, brNit));
        std::fill(kvws.begin(), kvws.end(), static_cast<R>(0));
        avgDiv = 0;
        itsV = 0;
        // the loop bound is maxIterations, not maxIterationsV (:746)
        while ((avgDiv > static_cast<R>(P.eta_v) / dt || itsV < P.min_it_v) && itsV < P.max_it) {
            computeDivergenceError(itsV < brNit ? branch(brSV, itsV) : nullptr);
            for (int i = 0; i < nf; ++i) {
                kv[i] = dErr[i] * alpha[i] / (dt * rho[i]);
                kvws[i] += kv[i];
            }
            velocityPass(kv, itsV < brNit ? branch(brKV, itsV) : nullptr);
            ++itsV;
        }
        for (int i = 0; i < nf; ++i) kvws[i] *= dt;
    }
    void stepDFSPH() {  // :33-107
        ++steps;
        forces();
        if (P.adaptive_dt) updateTimeStep();
        const int nf = static_cast<int>(x.size());
#pragma omp parallel for schedule(static)
        for (int i = 0; i < nf; ++i) {  // :180-202
            v[i] += (a[i] + g) * dt;
            a[i] = V();
        }
        correctDensityError();
#pragma omp parallel for schedule(static)
        for (int i = 0; i < nf; ++i) x[i] += v[i] * dt;  // :204-224
        insertFluid();
        computeDensities();
        computeAlpha();
        correctDivergenceError();
        cubeBoundary();   // DFSPHSolver.cpp:92-102
        t += dt;
    }

    void step() {
        if (P.method == 0) stepWCSPH();
        else if (P.method == 1) stepPCISPH();
        else stepDFSPH();
    }

    // within-support neighbour sets (27-cell candidates filtered by the policy test)
    void neighborSets(std::vector<int64_t>& off, std::vector<int32_t>& ids, std::vector<int64_t>& boff,
                      std::vector<int32_t>& bids) {
        grid.clear();
        insertBoundary();
        insertFluid();
        const int nf = static_cast<int>(x.size());
        off.assign(nf + 1, 0);
        boff.assign(nf + 1, 0);
        ids.clear();
        bids.clear();
        for (int i = 0; i < nf; ++i) {
            std::vector<int32_t> s, sb;
            const V ri = x[i];
            forCandidates(
                keyOf(ri), [&](int j) { if (Policy<R>::within(ri - x[j], h)) s.push_back(j); },
                [&](int b) { if (Policy<R>::within(ri - xb[b], h)) sb.push_back(b); });
            std::sort(s.begin(), s.end());
            std::sort(sb.begin(), sb.end());
            ids.insert(ids.end(), s.begin(), s.end());
            bids.insert(bids.end(), sb.begin(), sb.end());
            off[i + 1] = static_cast<int64_t>(ids.size());
            boff[i + 1] = static_cast<int64_t>(bids.size());
        }
    }
};

// ---- scene sampling (FluidObject.cpp:30-61, AkinciBoundaryObject.cpp:97-153) ----
// Fluid block lattice: ppedge = floor((max-min)/(2r) + 1e-5), x = (i,j,k)*2r + (min + r), k fastest,
// rounded as the shipped -O3 build rounds it: x as a multiply then an add, y and z as one FMA.
inline void sampleBlock(const float mn[3], const float mx[3], float radius, std::vector<float>& out) {
    const float dist = 2.0f * radius;
    unsigned pp[3];
    for (int a = 0; a < 3; ++a) pp[a] = static_cast<unsigned>(std::floor((mx[a] - mn[a]) / dist + 1e-5f));
    for (unsigned i = 0; i < pp[0]; ++i)
        for (unsigned j = 0; j < pp[1]; ++j)
            for (unsigned kk = 0; kk < pp[2]; ++kk) {
                const unsigned idx[3] = {i, j, kk};
                for (int a = 0; a < 3; ++a) {
                    const float base = mn[a] + radius;
                    const float t = static_cast<float>(idx[a]);
                    out.push_back(a == 0 ? t * dist + base : std::fma(t, dist, base));
                }
            }
}
// Box surface: n = ceil(|max-min|/(2r)) per axis, d = l/n, emit (i,j,k)*d + min with an index on a face.
// The shipped build forms l/(2r) (a vector over one scalar) as l * fl(1/(2r)).
inline void sampleBox(const float mn[3], const float mx[3], float radius, std::vector<float>& out) {
    const float inv = 1.0f / (2.0f * radius);
    float l[3], nn[3], d[3];
    for (int a = 0; a < 3; ++a) {
        l[a] = std::fabs(mn[a] - mx[a]);
        nn[a] = std::ceil(l[a] * inv);
        d[a] = l[a] / nn[a];
    }
    for (unsigned i = 0; i <= nn[0]; i++)
        for (unsigned j = 0; j <= nn[1]; j++)
            for (unsigned kk = 0; kk <= nn[2]; kk++)
                if (i == 0 || i == nn[0] || j == 0 || j == nn[1] || kk == 0 || kk == nn[2]) {
                    const unsigned idx[3] = {i, j, kk};
                    // rounded as the shipped -O3 build rounds it: z as one FMA
                    for (int a = 0; a < 3; ++a) {
                        const float t = static_cast<float>(idx[a]);
                        out.push_back(a == 2 ? std::fma(t, d[a], mn[a]) : t * d[a] + mn[a]);
                    }
                }
}

// Sphere surface (AkinciBoundaryObject.cpp:125-153): N = floor(4 pi R^2 / (pi r^2) 1.5)
// points in m_theta = round(pi / sqrt(4 pi / N)) rings at theta = pi (m + 1/2) / m_theta,
// ring m holding m_phi = round(2 pi sin(theta) / d_phi) points at phi = 2 pi n / m_phi.
// Rounding of the shipped build: both divisions by the ring counts as products with
// the count's float reciprocal, sin/cos in double on the float angles, x and y
// placed with one FMA, z with a multiply and an add.
inline void sampleSphere(const float c[3], float R, float radius, std::vector<float>& out) {
    const unsigned N = static_cast<unsigned>(std::floor(4.0 * M_PI * R * R / (M_PI * radius * radius) * 1.5));
    if (N == 0) return;
    const float area = static_cast<float>(4.0 * M_PI) / static_cast<float>(N);
    const unsigned rings = static_cast<unsigned>(std::round(M_PI / std::sqrt(area)));
    if (rings == 0) return;
    const float dPhi = area / (static_cast<float>(M_PI) / static_cast<float>(rings));
    const float rRings = 1.0f / static_cast<float>(rings);
    for (unsigned m = 0; m < rings; ++m) {
        const float th = static_cast<float>(M_PI * (m + 0.5)) * rRings;
        const double sth = std::sin(static_cast<double>(th));
        const unsigned per = static_cast<unsigned>(std::round(2.0 * M_PI * sth / dPhi));
        if (per == 0) continue;
        const float rPer = 1.0f / static_cast<float>(per);
        for (unsigned n = 0; n < per; ++n) {
            const double ph = static_cast<double>((static_cast<float>(2.0 * M_PI) * static_cast<float>(n)) * rPer);
            const float u[3] = {static_cast<float>(sth * std::cos(ph)), static_cast<float>(sth * std::sin(ph)),
                                static_cast<float>(std::cos(static_cast<double>(th)))};
            for (int a = 0; a < 3; ++a) out.push_back(a < 2 ? std::fma(R, u[a], c[a]) : R * u[a] + c[a]);
        }
    }
}

}  // namespace sphref
