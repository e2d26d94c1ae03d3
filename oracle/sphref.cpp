// C ABI over the sphref oracle (TEST INFRASTRUCTURE; loaded by tests/ and
// bench.py's CPU-baseline leg through ctypes, never by the product).
#include "sphref.hpp"

#include <chrono>
#include <cstring>
#include <type_traits>
#include <memory>

using namespace sphref;

namespace {
struct Handle {
    int bits;
    std::unique_ptr<Sim<float>> f;
    std::unique_ptr<Sim<double>> d;
};
template <class F>
auto with(Handle* h, F&& fn) {
    if (h->bits == 32) return fn(*h->f);
    return fn(*h->d);
}
}  // namespace

extern "C" {

void* sphref_create(int bits, const Params* p, int nf, const float* xf, const float* vf, int nb, const float* xb) {
    auto* h = new Handle{bits, nullptr, nullptr};
    if (bits == 32) {
        h->f.reset(new Sim<float>());
        h->f->setup(*p, nf, xf, vf, nb, xb);
    } else {
        h->d.reset(new Sim<double>());
        h->d->setup(*p, nf, xf, vf, nb, xb);
    }
    return h;
}

void sphref_destroy(void* hp) { delete static_cast<Handle*>(hp); }

void sphref_init(void* hp) {
    with(static_cast<Handle*>(hp), [](auto& s) { s.init(); return 0; });
}

// stats per step: t, dt, iterations, iterationsV, avg density err, avg div err
double sphref_step(void* hp, int nsteps, double* stats) {
    auto t0 = std::chrono::steady_clock::now();
    with(static_cast<Handle*>(hp), [&](auto& s) {
        for (int k = 0; k < nsteps; ++k) {
            s.step();
            if (stats) {
                double* o = stats + 6 * k;
                o[0] = static_cast<double>(s.t);
                o[1] = static_cast<double>(s.dt);
                o[2] = s.its;
                o[3] = s.itsV;
                o[4] = static_cast<double>(s.P.method == 1 ? s.pciAvg : s.avgErr);
                o[5] = static_cast<double>(s.avgDiv);
            }
        }
        return 0;
    });
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// Overwrite positions / velocities (N*3 doubles, either may be NULL) at full
// precision: the FP64 oracle's finite differences perturb them below FP32 ulps.
// DFSPH: rho and alpha follow the new positions and the warm-start values
// restart at zero (as the C ABI's sph_set_fluid_state does).
void sphref_set_state64(void* hp, const double* x, const double* v) {
    with(static_cast<Handle*>(hp), [&](auto& s) {
        using R = decltype(s.x[0].x);
        const size_t n = s.x.size();
        for (size_t i = 0; i < n; ++i) {
            if (x) s.x[i] = {static_cast<R>(x[3 * i]), static_cast<R>(x[3 * i + 1]), static_cast<R>(x[3 * i + 2])};
            if (v) s.v[i] = {static_cast<R>(v[3 * i]), static_cast<R>(v[3 * i + 1]), static_cast<R>(v[3 * i + 2])};
        }
        if (x) s.refreshDensities();
        return 0;
    });
}

// WCSPH pressure branch of the next `nsteps` steps counted from init:
// mask[k * N + i] for step k + 1 (see Sim::pBranch). nsteps = 0 clears it.
void sphref_set_pressure_branch(void* hp, int nsteps, const uint8_t* mask) {
    with(static_cast<Handle*>(hp), [&](auto& s) {
        const size_t n = s.x.size();
        s.pBranch.assign(static_cast<size_t>(nsteps), std::vector<uint8_t>(n));
        for (int k = 0; k < nsteps; ++k) std::copy(mask + k * n, mask + (k + 1) * n, s.pBranch[k].begin());
        return 0;
    });
}

// FP32 positions deciding the cohesion branch of the next `nsteps` steps
// counted from init: xyz[(k * N + i) * 3 + c] for step k + 1 (see Sim::cohPos).
void sphref_set_cohesion_positions(void* hp, int nsteps, const float* xyz) {
    with(static_cast<Handle*>(hp), [&](auto& s) {
        const size_t n = s.x.size();
        s.cohPos.assign(static_cast<size_t>(nsteps), std::vector<V3<float>>(n));
        for (int k = 0; k < nsteps; ++k)
            for (size_t i = 0; i < n; ++i) {
                const float* q = xyz + (k * n + i) * 3;
                s.cohPos[k][i] = V3<float>(q[0], q[1], q[2]);
            }
        return 0;
    });
}

// DFSPH branches of the next `nsteps` steps from another run (see
// Sim::brKD): arrays [nsteps][nit + 1][N] of k and rho* (density solve) and of
// k and the clamped divergence (divergence solve). nsteps = 0 clears them.
void sphref_set_dfsph_branch(void* hp, int nsteps, int nit, const float* kD, const float* sD, const float* kV,
                             const float* sV) {
    with(static_cast<Handle*>(hp), [&](auto& s) {
        const size_t len = static_cast<size_t>(nsteps) * (nit + 1) * s.x.size();
        s.brSteps = nsteps;
        s.brNit = nit;
        s.brKD.assign(kD, kD + (nsteps ? len : 0));
        s.brSD.assign(sD, sD + (nsteps ? len : 0));
        s.brKV.assign(kV, kV + (nsteps ? len : 0));
        s.brSV.assign(sV, sV + (nsteps ? len : 0));
        return 0;
    });
}

// Inner-branch cohesion fingerprint since init: out[0] = hash, out[1] = pairs.
void sphref_cohesion_print(void* hp, uint64_t* out) {
    with(static_cast<Handle*>(hp), [&](auto& s) {
        out[0] = s.innerHash;
        out[1] = static_cast<uint64_t>(s.innerPairs);
        return 0;
    });
}

// Any output pointer may be NULL. Vectors are N*3, scalars N.
void sphref_get(void* hp, double* x, double* v, double* a, double* rho, double* p, double* alpha) {
    with(static_cast<Handle*>(hp), [&](auto& s) {
        const size_t n = s.x.size();
        for (size_t i = 0; i < n; ++i) {
            if (x) { x[3 * i] = s.x[i].x; x[3 * i + 1] = s.x[i].y; x[3 * i + 2] = s.x[i].z; }
            if (v) { v[3 * i] = s.v[i].x; v[3 * i + 1] = s.v[i].y; v[3 * i + 2] = s.v[i].z; }
            if (a) { a[3 * i] = s.a[i].x; a[3 * i + 1] = s.a[i].y; a[3 * i + 2] = s.a[i].z; }
            if (rho) rho[i] = s.rho[i];
            if (p) p[i] = s.p[i];
            if (alpha) alpha[i] = s.alpha[i];
        }
        return 0;
    });
}

void sphref_get_boundary(void* hp, double* vol) {
    with(static_cast<Handle*>(hp), [&](auto& s) {
        for (size_t i = 0; i < s.Vb.size(); ++i) vol[i] = s.Vb[i];
        return 0;
    });
}

// scalars: t, dt, mass, h, pcisph delta
void sphref_scalars(void* hp, double* out) {
    with(static_cast<Handle*>(hp), [&](auto& s) {
        out[0] = s.t; out[1] = s.dt; out[2] = s.m; out[3] = s.h; out[4] = s.delta;
        return 0;
    });
}

void sphref_cell_keys(void* hp, int32_t* keys) {
    with(static_cast<Handle*>(hp), [&](auto& s) {
        for (size_t i = 0; i < s.x.size(); ++i) {
            const Key k = s.keyOf(s.x[i]);
            keys[3 * i] = k.x; keys[3 * i + 1] = k.y; keys[3 * i + 2] = k.z;
        }
        return 0;
    });
}

// Within-support sets at the current positions. Returns the total number of
// fluid pairs; writes at most cap ids (and bcap boundary ids).
int64_t sphref_neighbors(void* hp, int64_t* off, int32_t* ids, int64_t cap, int64_t* boff, int32_t* bids,
                         int64_t bcap, int64_t* btotal) {
    return with(static_cast<Handle*>(hp), [&](auto& s) {
        std::vector<int64_t> o, bo;
        std::vector<int32_t> id, bid;
        s.neighborSets(o, id, bo, bid);
        std::memcpy(off, o.data(), o.size() * sizeof(int64_t));
        std::memcpy(boff, bo.data(), bo.size() * sizeof(int64_t));
        std::memcpy(ids, id.data(), std::min<int64_t>(cap, id.size()) * sizeof(int32_t));
        std::memcpy(bids, bid.data(), std::min<int64_t>(bcap, bid.size()) * sizeof(int32_t));
        *btotal = static_cast<int64_t>(bid.size());
        return static_cast<int64_t>(id.size());
    });
}

// Kernel evaluations for closed-form checks: mode 0 W(r), 1 |gradW| signed
// radial component, 2 laplW, 3 cohesion. r given as distances along x.
void sphref_kernel_eval(int bits, double h, int mode, int n, const double* rr, double* out) {
    Params p{};
    p.radius = h / 4.0;
    p.density0 = 1000;
    p.fps = 60;
    if (bits == 32) {
        Sim<float> s;
        float zero = 0;
        s.setup(p, 0, &zero, nullptr, 0, &zero);
        for (int i = 0; i < n; ++i) {
            V3<float> rv(static_cast<float>(rr[i]), 0.f, 0.f);
            out[i] = mode == 0 ? s.W(rv) : mode == 1 ? s.gradW(rv).x : mode == 2 ? s.laplW(rv) : s.cohesion(rv);
        }
    } else {
        Sim<double> s;
        float zero = 0;
        s.setup(p, 0, &zero, nullptr, 0, &zero);
        for (int i = 0; i < n; ++i) {
            V3<double> rv(rr[i], 0.0, 0.0);
            out[i] = mode == 0 ? s.W(rv) : mode == 1 ? s.gradW(rv).x : mode == 2 ? s.laplW(rv) : s.cohesion(rv);
        }
    }
}

int64_t sphref_sample_block(const float* mn, const float* mx, float radius, float* out, int64_t cap) {
    std::vector<float> v;
    sampleBlock(mn, mx, radius, v);
    std::memcpy(out, v.data(), std::min<int64_t>(cap * 3, v.size()) * sizeof(float));
    return static_cast<int64_t>(v.size() / 3);
}
int64_t sphref_sample_box(const float* mn, const float* mx, float radius, float* out, int64_t cap) {
    std::vector<float> v;
    sampleBox(mn, mx, radius, v);
    std::memcpy(out, v.data(), std::min<int64_t>(cap * 3, v.size()) * sizeof(float));
    return static_cast<int64_t>(v.size() / 3);
}

int64_t sphref_sample_sphere(const float* c, float R, float radius, float* out, int64_t cap) {
    std::vector<float> v;
    sampleSphere(c, R, radius, v);
    std::memcpy(out, v.data(), std::min<int64_t>(cap * 3, v.size()) * sizeof(float));
    return static_cast<int64_t>(v.size() / 3);
}

}  // extern "C"
