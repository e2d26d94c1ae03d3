// Headless command-line driver (the -g false path of the reference's
// src/main.cpp:16-91): import a scene JSON, init, and step until endTime or
// a step count, optionally writing frames as <scene>_NNNN.json with the
// fields of SceneLoader::writeFluid (SceneLoader.cpp:211-275).
//
//   sph2o_cli -c scene.json [-n steps] [-o true|false] [-d device]
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/sph2o.h"

static void writeFrame(const std::string& scene, unsigned frame, int64_t n, int64_t active, double t, double dt,
                       const std::vector<float>& x, const std::vector<float>& v) {
    std::string num = std::to_string(frame);
    while (num.size() < 4) num = "0" + num;  // SceneLoader.h:20 MAX_DIGITS
    std::string fn = scene;
    const std::string ext = ".json";
    if (fn.size() >= ext.size() && fn.compare(fn.size() - ext.size(), ext.size(), ext) == 0)
        fn.insert(fn.size() - ext.size(), "_" + num);
    else
        fn += "_" + num + ext;
    FILE* f = std::fopen(fn.c_str(), "w");
    if (!f) {
        std::fprintf(stderr, "Cannot write fluid state\n");
        return;
    }
    std::fprintf(f, "{\"numParticles\":%lld,\"numActiveParticles\":%lld,\"time\":%.9g,\"timeStep\":%.9g,\"positions\":[",
                 static_cast<long long>(n), static_cast<long long>(active), t, dt);
    for (int64_t i = 0; i < n; ++i)
        std::fprintf(f, "%s[%.9g,%.9g,%.9g]", i ? "," : "", x[3 * i], x[3 * i + 1], x[3 * i + 2]);
    std::fprintf(f, "],\"velocities\":[");
    for (int64_t i = 0; i < n; ++i)
        std::fprintf(f, "%s[%.9g,%.9g,%.9g]", i ? "," : "", v[3 * i], v[3 * i + 1], v[3 * i + 2]);
    std::fprintf(f, "]}");
    std::fclose(f);
}

int main(int argc, char** argv) {
    std::string scene;
    long steps = -1;
    bool save = false;
    int device = 0;
    for (int a = 1; a + 1 < argc; a += 2) {
        if (!std::strcmp(argv[a], "-c")) scene = argv[a + 1];
        else if (!std::strcmp(argv[a], "-n")) steps = std::atol(argv[a + 1]);
        else if (!std::strcmp(argv[a], "-o")) save = !std::strcmp(argv[a + 1], "true");
        else if (!std::strcmp(argv[a], "-d")) device = std::atoi(argv[a + 1]);
    }
    if (scene.empty()) {
        std::fprintf(stderr, "usage: %s -c scene.json [-n steps] [-o true|false] [-d device]\n", argv[0]);
        return 2;
    }
    sph_ctx* ctx = nullptr;
    int rc = sph_create(&ctx, device, nullptr);
    if (rc) {
        std::fprintf(stderr, "sph_create failed (%d)\n", rc);
        return 1;
    }
    const int sid = sph_load_scene_json(ctx, scene.c_str());
    if (sid < 0) {
        std::fprintf(stderr, "Cannot import scene: %s\n", sph_last_error(ctx));
        return 1;
    }
    if ((rc = sph_init(ctx))) {
        std::fprintf(stderr, "init failed: %s\n", sph_last_error(ctx));
        return 1;
    }
    sph_scene_params p;
    sph_get_params(ctx, sid, &p);
    const int64_t n = sph_num_fluid(ctx, sid);
    std::vector<float> x(3 * n), v(3 * n);
    double t = 0, dt = 0;
    float lastSave = 0.0f;   // TimeHelper::lastTimeSave, a Real
    unsigned frame = 1;
    long done = 0;
    // SPHSimulation::step gate (SPHSimulation.cpp:163) and TimeHelper::hasToSave (TimeHelper.h:100-110)
    while (static_cast<float>(t) >= p.start_time && static_cast<float>(t) <= p.end_time && (steps < 0 || done < steps)) {
        sph_step_stats st;
        if ((rc = sph_step(ctx, 1, &st))) {
            std::fprintf(stderr, "step failed: %s\n", sph_last_error(ctx));
            return 1;
        }
        ++done;
        sph_get_time(ctx, sid, &t, &dt);
        // time - lastTimeSave in FP32 against 1.0 / fps in FP64, then
        // lastTimeSave = (Real) frame / fps in FP32, as the reference rounds them
        if (static_cast<double>(static_cast<float>(t) - lastSave) >= 1.0 / static_cast<double>(p.fps)) {
            lastSave = static_cast<float>(frame) / p.fps;
            if (save) {
                sph_get_fluid_state(ctx, sid, x.data(), v.data(), nullptr, nullptr);
                writeFrame(scene, frame, n, sph_num_active(ctx, sid), t, dt, x, v);
            }
            ++frame;
        }
    }
    std::printf("steps %ld t %.6f dt %.6g particles %lld\n", done, t, dt, static_cast<long long>(n));
    sph_destroy(ctx);
    return 0;
}
