// GpuSPHSolver — the reference-side binding of libsph2o.so: an SPHSolver
// (inc/SPH2O/SPHSolver/SPHSolver.h:6-48) whose init() / step() run the
// MI355X step through the C ABI of include/sph2o.h, so SPHSimulation's
// importScene / init / step and the GUI stay as they are. A maintainer adds
// this header next to DFSPHSolver.h and a case in
// SPHSimulation::setSimulationMethod (SPHSimulation.cpp:216-239) that
// constructs GpuSPHSolver(name) instead of a CPU solver.
#ifndef GPU_SPH_SOLVER_H_
#define GPU_SPH_SOLVER_H_

#include <string>
#include <vector>

#include "SPH2O/SPHSimulation.h"
#include "SPH2O/SPHSolver/SPHSolver.h"
#include "sph2o.h"

class GpuSPHSolver : public SPHSolver {
    sph_ctx* ctx = nullptr;
    int scene = -1;
    std::vector<float> pos, vel;

  public:
    explicit GpuSPHSolver(const std::string& scenePath) {
        if (sph_create(&ctx, /*device*/ 0, /*stream*/ nullptr) != SPH_OK) {
            ERROR("GPU context: no HIP device");
            return;
        }
        scene = sph_load_scene_json(ctx, scenePath.c_str());   // the JSON importScene read
        if (scene < 0) ERROR("GPU scene: ", sph_last_error(ctx));
    }
    ~GpuSPHSolver() override { sph_destroy(ctx); }

    // SPHSolver::init: the CPU-side state the simulation built (or edited)
    // goes to the device, then the solver's init pass runs there
    void init() override {
        SPHSimulation* sim = SPHSimulation::getCurrent();
        FluidObject* f = sim->getFluidModel(0);
        const uint n = f->getSize();
        pos.resize(3 * static_cast<size_t>(n));
        vel.resize(3 * static_cast<size_t>(n));
        for (uint i = 0; i < n; ++i) {
            const Vector3r& x = f->getPosition(i);
            const Vector3r& v = f->getVelocity(i);
            pos[3 * i] = x.x; pos[3 * i + 1] = x.y; pos[3 * i + 2] = x.z;
            vel[3 * i] = v.x; vel[3 * i + 1] = v.y; vel[3 * i + 2] = v.z;
        }
        if (sph_num_active(ctx, scene) == static_cast<int64_t>(n))   // no emitters: the state may be edited
            sph_set_fluid_state(ctx, scene, n, pos.data(), vel.data());
        if (sph_init(ctx) != SPH_OK) ERROR("GPU init: ", sph_last_error(ctx));
    }

    // SPHSolver::step: one step on the device; time and time step go back to
    // the simulation (the CPU solvers advance them inside step()), the state
    // to the fluid object for the renderer and the frame writer
    void step() override {
        SPHSimulation* sim = SPHSimulation::getCurrent();
        sph_step_stats st;
        if (sph_step(ctx, 1, &st) != SPH_OK) {
            ERROR("GPU step: ", sph_last_error(ctx));
            return;
        }
        ++steps;
        iterations = static_cast<uint>(st.iterations);
        sumIterations += iterations;
        maxVel = static_cast<Real>(st.max_velocity);
        sim->setTimeStep(static_cast<Real>(st.dt));
        sim->setTime(static_cast<Real>(st.t));
        FluidObject* f = sim->getFluidModel(0);
        sph_get_fluid_state(ctx, scene, pos.data(), vel.data(), nullptr, nullptr);
        for (uint i = 0; i < f->getSize(); ++i) {
            f->getPosition(i) = Vector3r(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
            f->getVelocity(i) = Vector3r(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]);
        }
        f->setNumActiveParticles(static_cast<uint>(sph_num_active(ctx, scene)));   // emitters
    }

    // SPHSolver::resize: the device context sizes its buffers at sph_init
    void resize() override {}
};

#endif
