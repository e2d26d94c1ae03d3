// Scene-JSON reader and particle samplers (host side of the drop-in boundary).
#include "scene.hpp"

#include <cctype>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>

namespace sph {
namespace {

// ---- a small JSON DOM (objects, arrays, numbers, strings, bools, null) ----
struct JVal {
    enum Kind { Null, Bool, Num, Str, Arr, Obj } kind = Null;
    bool b = false;
    double num = 0;
    std::string str;
    std::vector<JVal> arr;
    std::vector<std::pair<std::string, JVal>> obj;
    const JVal* find(const char* k) const {
        if (kind != Obj) return nullptr;
        for (auto& kv : obj)
            if (kv.first == k) return &kv.second;
        return nullptr;
    }
};

struct Parser {
    const char* p;
    const char* end;
    std::string err;
    void ws() {
        while (p < end && std::isspace(static_cast<unsigned char>(*p))) ++p;
    }
    bool fail(const char* m) {
        if (err.empty()) err = m;
        return false;
    }
    bool parse(JVal& v, int depth = 0) {
        if (depth > 64) return fail("JSON nested too deeply");
        ws();
        if (p >= end) return fail("unexpected end of JSON");
        const char c = *p;
        if (c == '{') {
            ++p;
            v.kind = JVal::Obj;
            ws();
            if (p < end && *p == '}') { ++p; return true; }
            while (true) {
                ws();
                std::string key;
                if (!parseString(key)) return false;
                ws();
                if (p >= end || *p != ':') return fail("expected ':' in JSON object");
                ++p;
                JVal child;
                if (!parse(child, depth + 1)) return false;
                v.obj.emplace_back(std::move(key), std::move(child));
                ws();
                if (p < end && *p == ',') { ++p; continue; }
                if (p < end && *p == '}') { ++p; return true; }
                return fail("expected ',' or '}' in JSON object");
            }
        }
        if (c == '[') {
            ++p;
            v.kind = JVal::Arr;
            ws();
            if (p < end && *p == ']') { ++p; return true; }
            while (true) {
                JVal child;
                if (!parse(child, depth + 1)) return false;
                v.arr.push_back(std::move(child));
                ws();
                if (p < end && *p == ',') { ++p; continue; }
                if (p < end && *p == ']') { ++p; return true; }
                return fail("expected ',' or ']' in JSON array");
            }
        }
        if (c == '"') {
            v.kind = JVal::Str;
            return parseString(v.str);
        }
        if (!std::strncmp(p, "true", std::min<size_t>(4, end - p)) && end - p >= 4) { p += 4; v.kind = JVal::Bool; v.b = true; return true; }
        if (!std::strncmp(p, "false", std::min<size_t>(5, end - p)) && end - p >= 5) { p += 5; v.kind = JVal::Bool; return true; }
        if (!std::strncmp(p, "null", std::min<size_t>(4, end - p)) && end - p >= 4) { p += 4; v.kind = JVal::Null; return true; }
        if (c == '-' || (c >= '0' && c <= '9')) {
            std::string tok;
            while (p < end && (std::isdigit(static_cast<unsigned char>(*p)) || *p == '-' || *p == '+' || *p == '.' ||
                               *p == 'e' || *p == 'E'))
                tok.push_back(*p++);
            char* e = nullptr;
            v.num = std::strtod(tok.c_str(), &e);
            if (!e || *e) return fail("bad JSON number");
            v.kind = JVal::Num;
            return true;
        }
        return fail("unexpected character in JSON");
    }
    bool parseString(std::string& s) {
        if (p >= end || *p != '"') return fail("expected JSON string");
        ++p;
        while (p < end && *p != '"') {
            if (*p == '\\') {
                ++p;
                if (p >= end) return fail("bad escape");
                const char e = *p++;
                switch (e) {
                    case 'n': s.push_back('\n'); break;
                    case 't': s.push_back('\t'); break;
                    case 'r': s.push_back('\r'); break;
                    case 'b': s.push_back('\b'); break;
                    case 'f': s.push_back('\f'); break;
                    case 'u': {  // keep ASCII subset only
                        if (end - p < 4) return fail("bad \\u escape");
                        const unsigned cp = static_cast<unsigned>(std::strtoul(std::string(p, 4).c_str(), nullptr, 16));
                        s.push_back(cp < 128 ? static_cast<char>(cp) : '?');
                        p += 4;
                        break;
                    }
                    default: s.push_back(e);
                }
            } else {
                s.push_back(*p++);
            }
        }
        if (p >= end) return fail("unterminated JSON string");
        ++p;
        return true;
    }
};

struct Reader {
    std::string err;
    bool num(const JVal* obj, const char* key, double& out, const std::string& where) {
        const JVal* v = obj ? obj->find(key) : nullptr;
        if (!v) { err = "missing key " + where + "." + key; return false; }
        if (v->kind == JVal::Num) { out = v->num; return true; }
        if (v->kind == JVal::Bool) { out = v->b ? 1 : 0; return true; }
        err = "key " + where + "." + key + " is not a number";
        return false;
    }
    bool f32(const JVal* obj, const char* key, float& out, const std::string& where) {
        double d;
        if (!num(obj, key, d, where)) return false;
        out = static_cast<float>(d);
        return true;
    }
    bool i32(const JVal* obj, const char* key, int32_t& out, const std::string& where) {
        double d;
        if (!num(obj, key, d, where)) return false;
        out = static_cast<int32_t>(d);
        return true;
    }
    bool u32(const JVal* obj, const char* key, uint32_t& out, const std::string& where) {
        double d;
        if (!num(obj, key, d, where)) return false;
        out = static_cast<uint32_t>(d < 0 ? 0 : d);
        return true;
    }
    bool vecN(const JVal* obj, const char* key, float* out, int n, const std::string& where) {
        const JVal* v = obj ? obj->find(key) : nullptr;
        if (!v || v->kind != JVal::Arr || v->arr.size() < static_cast<size_t>(n)) {
            err = "key " + where + "." + key + " must be an array of " + std::to_string(n) + " numbers";
            return false;
        }
        for (int a = 0; a < n; ++a) {
            if (v->arr[a].kind != JVal::Num) { err = "key " + where + "." + key + " must hold numbers"; return false; }
            out[a] = static_cast<float>(v->arr[a].num);
        }
        return true;
    }
    bool vec3(const JVal* obj, const char* key, float out[3], const std::string& where) {
        return vecN(obj, key, out, 3, where);
    }
};

}  // namespace

int parseSceneJson(const std::string& text, SceneDesc& out, std::string& err) {
    JVal root;
    Parser ps{text.data(), text.data() + text.size(), {}};
    if (!ps.parse(root)) { err = "scene JSON: " + ps.err; return SPH_E_SCENE; }
    Reader rd;
    sph_scene_params& P = out.params;
    std::memset(&P, 0, sizeof(P));
    P.adaptive_time_step = 1;
    P.min_iterations = 3; P.max_iterations = 100;   // SPHSolver.cpp:15-16
    P.min_iterations_v = 1; P.max_iterations_v = 100;
    P.eta = 0.01f;                                  // SPHSolver.cpp:18
    P.density0 = 1000.0f;                           // FluidObject.cpp:12
    P.viscosity_method = P.surface_tension_method = P.adhesion_method = -1;

    // Configuration (SceneLoader.cpp:40-78)
    const JVal* C = root.find("Configuration");
    const std::string cw = "Configuration";
    if (!C) { err = "missing key Configuration"; return SPH_E_SCENE; }
    if (!rd.f32(C, "startTime", P.start_time, cw) || !rd.f32(C, "endTime", P.end_time, cw) ||
        !rd.f32(C, "timeStep", P.time_step, cw) || !rd.f32(C, "fps", P.fps, cw) ||
        !rd.f32(C, "minTimeStep", P.min_time_step, cw) || !rd.f32(C, "maxTimeStep", P.max_time_step, cw) ||
        !rd.f32(C, "cflFactor", P.cfl_factor, cw) || !rd.f32(C, "particleRadius", P.particle_radius, cw) ||
        !rd.i32(C, "boundaryMethod", P.boundary_method, cw) || !rd.i32(C, "simulationMethod", P.simulation_method, cw) ||
        !rd.vec3(C, "gravity", P.gravity, cw)) {
        err = rd.err;
        return SPH_E_SCENE;
    }
    if (P.simulation_method == SPH_WCSPH) {
        if (!rd.f32(C, "stiffness", P.stiffness, cw) || !rd.f32(C, "gamma", P.gamma, cw)) { err = rd.err; return SPH_E_SCENE; }
    } else if (P.simulation_method == SPH_PCISPH) {
        if (!rd.f32(C, "eta", P.eta, cw) || !rd.u32(C, "minIterations", P.min_iterations, cw) ||
            !rd.u32(C, "maxIterations", P.max_iterations, cw)) { err = rd.err; return SPH_E_SCENE; }
    } else if (P.simulation_method == SPH_DFSPH) {
        if (!rd.f32(C, "eta", P.eta, cw) || !rd.f32(C, "etaV", P.eta_v, cw) ||
            !rd.u32(C, "minIterations", P.min_iterations, cw) || !rd.u32(C, "maxIterations", P.max_iterations, cw) ||
            !rd.u32(C, "minIterationsV", P.min_iterations_v, cw) ||
            !rd.u32(C, "maxIterationsV", P.max_iterations_v, cw)) { err = rd.err; return SPH_E_SCENE; }
    } else {
        err = "Configuration.simulationMethod must be 0 (WCSPH), 1 (PCISPH) or 2 (DFSPH)";
        return SPH_E_SCENE;
    }

    // Fluid (SceneLoader.cpp:80-146)
    const JVal* F = root.find("Fluid");
    const std::string fw = "Fluid";
    if (!F) { err = "missing key Fluid"; return SPH_E_SCENE; }
    if (!rd.i32(F, "viscosityMethod", P.viscosity_method, fw) ||
        !rd.i32(F, "surfaceTensionMethod", P.surface_tension_method, fw) ||
        !rd.i32(F, "adhesionMethod", P.adhesion_method, fw)) { err = rd.err; return SPH_E_SCENE; }
    const JVal* FM = F->find("FluidModel");
    if (!FM || FM->kind != JVal::Arr) { err = "missing array Fluid.FluidModel"; return SPH_E_SCENE; }
    out.numFluidModels = static_cast<int>(FM->arr.size());
    if (out.numFluidModels != 1) {
        err = "exactly one FluidModel is supported (the reference's neighbour loops ignore pointSetId)";
        return SPH_E_UNSUPPORTED;
    }
    const JVal& m = FM->arr[0];
    const std::string mw = "Fluid.FluidModel[0]";
    if (!rd.f32(&m, "density0", P.density0, mw)) { err = rd.err; return SPH_E_SCENE; }
    if (P.viscosity_method > -1 &&
        (!rd.f32(&m, "viscosity", P.viscosity, mw) || !rd.f32(&m, "boundaryViscosity", P.boundary_viscosity, mw))) {
        err = rd.err;
        return SPH_E_SCENE;
    }
    if (P.surface_tension_method > -1 && !rd.f32(&m, "surfaceTension", P.surface_tension, mw)) { err = rd.err; return SPH_E_SCENE; }
    if (P.adhesion_method > -1 && !rd.f32(&m, "adhesion", P.adhesion, mw)) { err = rd.err; return SPH_E_SCENE; }
    const JVal* blocks = m.find("fluidBlocks");
    if (!blocks || blocks->kind != JVal::Arr) { err = "missing array " + mw + ".fluidBlocks"; return SPH_E_SCENE; }
    for (size_t j = 0; j < blocks->arr.size(); ++j) {
        BlockDesc b;
        const std::string bw = mw + ".fluidBlocks[" + std::to_string(j) + "]";
        if (!rd.vec3(&blocks->arr[j], "min", b.min, bw) || !rd.vec3(&blocks->arr[j], "max", b.max, bw)) { err = rd.err; return SPH_E_SCENE; }
        out.fluidBlocks.push_back(b);
    }
    // emitters (SceneLoader.cpp:107-141): rotation is the quaternion (w, x, y, z)
    if (const JVal* em = m.find("emitters")) {
        if (em->kind != JVal::Arr) { err = mw + ".emitters must be an array"; return SPH_E_SCENE; }
        for (size_t j = 0; j < em->arr.size(); ++j) {
            const JVal* e = &em->arr[j];
            const std::string ew = mw + ".emitters[" + std::to_string(j) + "]";
            sph_emitter_params q{};
            uint32_t n = 0, type = 0;
            if (!rd.vec3(e, "position", q.position, ew) || !rd.vecN(e, "rotation", q.rotation, 4, ew) ||
                !rd.f32(e, "velocity", q.velocity, ew) || !rd.u32(e, "numParticles", n, ew) ||
                !rd.f32(e, "startTime", q.start_time, ew) || !rd.u32(e, "type", type, ew) ||
                !rd.f32(e, "width", q.width, ew) || !rd.f32(e, "height", q.height, ew) ||
                !rd.f32(e, "spacing", q.spacing, ew)) {
                err = rd.err;
                return SPH_E_SCENE;
            }
            q.num_particles = static_cast<int64_t>(n);
            q.type = static_cast<int32_t>(type);
            out.emitters.push_back(q);
        }
        out.numEmitters = static_cast<int>(out.emitters.size());
    }

    // Boundary (SceneLoader.cpp:148-209); absent sphere/geometry arrays are empty.
    if (const JVal* B = root.find("Boundary")) {
        if (B->kind != JVal::Arr) { err = "Boundary must be an array"; return SPH_E_SCENE; }
        const bool pci = P.boundary_method == 1;
        out.numBoundaryObjects = static_cast<int>(B->arr.size());
        for (size_t i = 0; i < B->arr.size(); ++i) {
            const JVal& bo = B->arr[i];
            const std::string bw = "Boundary[" + std::to_string(i) + "]";
            if (pci) {   // read for the PCISPH method only (SceneLoader.cpp:168-172)
                float nf = 0.0f, tf = 0.0f;
                if (!rd.f32(&bo, "normalFct", nf, bw) || !rd.f32(&bo, "tangFct", tf, bw)) { err = rd.err; return SPH_E_SCENE; }
                out.normalFct.push_back(nf);
                out.tangFct.push_back(tf);
            }
            if (const JVal* bx = bo.find("box")) {
                if (bx->kind != JVal::Arr) { err = bw + ".box must be an array"; return SPH_E_SCENE; }
                for (size_t j = 0; j < bx->arr.size(); ++j) {
                    BlockDesc b;
                    const std::string w = bw + ".box[" + std::to_string(j) + "]";
                    if (!rd.vec3(&bx->arr[j], "min", b.min, w) || !rd.vec3(&bx->arr[j], "max", b.max, w)) { err = rd.err; return SPH_E_SCENE; }
                    int32_t inv = 0;
                    if (pci && !rd.i32(&bx->arr[j], "inverted", inv, w)) { err = rd.err; return SPH_E_SCENE; }
                    out.boxInverted.push_back(inv != 0);
                    out.shapeOrder.push_back(static_cast<int>(out.boxes.size()));
                    out.boxes.push_back(b);
                    if (j == 0) out.cubes.push_back(b);   // the cube method keeps one box per object
                }
            }
            if (const JVal* sp = bo.find("sphere")) {
                if (sp->kind != JVal::Arr) { err = bw + ".sphere must be an array"; return SPH_E_SCENE; }
                for (size_t j = 0; j < sp->arr.size(); ++j) {
                    SphereDesc s;
                    const std::string w = bw + ".sphere[" + std::to_string(j) + "]";
                    if (!rd.vec3(&sp->arr[j], "pos", s.pos, w) || !rd.f32(&sp->arr[j], "radius", s.radius, w)) { err = rd.err; return SPH_E_SCENE; }
                    int32_t inv = 0;
                    if (pci && !rd.i32(&sp->arr[j], "inverted", inv, w)) { err = rd.err; return SPH_E_SCENE; }
                    out.sphereInverted.push_back(inv != 0);
                    out.shapeOrder.push_back(-1 - static_cast<int>(out.spheres.size()));
                    out.spheres.push_back(s);
                }
            }
            if (const JVal* ge = bo.find("geometry"))
                out.numGeometries += ge->kind == JVal::Arr ? static_cast<int>(ge->arr.size()) : 0;
        }
    }
    return SPH_OK;
}

void sampleFluidBlock(const BlockDesc& b, float radius, std::vector<float>& xyz) {
    const float dist = 2.0f * radius;
    unsigned pp[3];
    for (int a = 0; a < 3; ++a) pp[a] = static_cast<unsigned>(std::floor((b.max[a] - b.min[a]) / dist + 1e-5f));
    float base[3];
    for (int a = 0; a < 3; ++a) base[a] = b.min[a] + radius;
    // pos = (i, j, k) * dist + (min + r) as the shipped -O3 build rounds it
    // (pinned against its 1M-particle lattice): x with two roundings, y and
    // z contracted into one FMA each
    for (unsigned i = 0; i < pp[0]; ++i)
        for (unsigned j = 0; j < pp[1]; ++j)
            for (unsigned k = 0; k < pp[2]; ++k) {
                xyz.push_back(static_cast<float>(i) * dist + base[0]);
                xyz.push_back(std::fma(static_cast<float>(j), dist, base[1]));
                xyz.push_back(std::fma(static_cast<float>(k), dist, base[2]));
            }
}

void sampleBoxSurface(const BlockDesc& b, float radius, std::vector<float>& xyz) {
    // n = ceil(l / 2r) with the vector-by-scalar division done, as the shipped
    // -O3 -ffast-math build does it, as l times one rounded reciprocal: for
    // l = 0.06, r = 0.01 that is a tie that rounds to 3 (a true division gives
    // 3.0000002, so 4)
    const float inv2r = 1.0f / (2.0f * radius);
    float l[3], n[3], d[3];
    for (int a = 0; a < 3; ++a) {
        l[a] = std::fabs(b.min[a] - b.max[a]);
        n[a] = std::ceil(l[a] * inv2r);
        d[a] = l[a] / n[a];
    }
    for (unsigned i = 0; i <= n[0]; i++)
        for (unsigned j = 0; j <= n[1]; j++)
            for (unsigned k = 0; k <= n[2]; k++)
                if (i == 0 || i == n[0] || j == 0 || j == n[1] || k == 0 || k == n[2]) {
                    // (i d.x, j d.y, k d.z) + min as the shipped -O3 build rounds
                    // it: z contracted into one FMA, x and y two roundings
                    xyz.push_back(static_cast<float>(i) * d[0] + b.min[0]);
                    xyz.push_back(static_cast<float>(j) * d[1] + b.min[1]);
                    xyz.push_back(std::fma(static_cast<float>(k), d[2], b.min[2]));
                }
}

void sampleSphereSurface(const SphereDesc& s, float radius, std::vector<float>& xyz) {
    // Rounded as the shipped -O3 -ffast-math build rounds it (pinned point by
    // point against its samples): the divisions of theta by m_theta and of
    // phi by m_phi as products with the rounded reciprocal of the loop-invariant
    // count; sin / cos in double on the float angles, the product rounded
    // once to float; x and y placed with one FMA each, z with a multiply and
    // an add.
    const double sep = 1.5;
    const unsigned numParts = static_cast<unsigned>(
        std::floor(4.0 * M_PI * s.radius * s.radius / (M_PI * radius * radius) * sep));
    if (numParts == 0) return;
    const float a = static_cast<float>(4.0 * M_PI) / static_cast<float>(numParts);
    const float d = std::sqrt(a);
    const unsigned mTheta = static_cast<unsigned>(std::round(M_PI / d));
    if (mTheta == 0) return;
    const float dTheta = static_cast<float>(M_PI) / static_cast<float>(mTheta);
    const float dPhi = a / dTheta;
    const float invMTheta = 1.0f / static_cast<float>(mTheta);
    for (unsigned m = 0; m < mTheta; ++m) {
        const float theta = static_cast<float>(M_PI * (m + 0.5)) * invMTheta;
        const double st = std::sin(static_cast<double>(theta)), ct = std::cos(static_cast<double>(theta));
        const unsigned mPhi = static_cast<unsigned>(std::round(2.0 * M_PI * st / dPhi));
        if (mPhi == 0) continue;
        const float invMPhi = 1.0f / static_cast<float>(mPhi);
        for (unsigned n = 0; n < mPhi; ++n) {
            const float phi = (static_cast<float>(2.0 * M_PI) * static_cast<float>(n)) * invMPhi;
            const float x = static_cast<float>(st * std::cos(static_cast<double>(phi)));
            const float y = static_cast<float>(st * std::sin(static_cast<double>(phi)));
            const float z = static_cast<float>(ct);
            xyz.push_back(std::fma(s.radius, x, s.pos[0]));
            xyz.push_back(std::fma(s.radius, y, s.pos[1]));
            xyz.push_back(s.radius * z + s.pos[2]);
        }
    }
}

bool samplePciBox(const BlockDesc& b, float supportRadius, bool inverted, std::vector<float>& xyz,
                  std::vector<float>& nrm) {
    // l = abs(int(min - max)) as the shipped build resolves the unqualified abs
    // (the int overload: the extent truncates to whole metres)
    float l[3], d[3];
    unsigned n[3];
    const float spacing = 0.45f * supportRadius;
    for (int a = 0; a < 3; ++a) {
        const int t = static_cast<int>(b.min[a] - b.max[a]);
        l[a] = static_cast<float>(t < 0 ? -t : t);
        if (l[a] == 0.0f) return false;
        n[a] = static_cast<unsigned>(std::ceil(l[a] / spacing));
        d[a] = l[a] / static_cast<float>(n[a]);
    }
    // face normals (+x at i = 0, -x at i = n, +y at j = 0, -y at j = n, +z at
    // k = 0, -z at k = n); an edge point takes normalize(0.5 (a + b)), a corner
    // normalize((a + b + c) / 3), as the reference's table of 26 cases does
    auto unit = [](float x, float y, float z, float* o) {
        const float q = 1.0f / std::sqrt(x * x + y * y + z * z);
        o[0] = x * q; o[1] = y * q; o[2] = z * q;
    };
    const float sgn = inverted ? 1.0f : -1.0f;
    for (unsigned i = 0; i <= n[0]; i++)
        for (unsigned j = 0; j <= n[1]; j++)
            for (unsigned k = 0; k <= n[2]; k++) {
                const float fx = i == 0 ? 1.0f : (i == n[0] ? -1.0f : 0.0f);
                const float fy = j == 0 ? 1.0f : (j == n[1] ? -1.0f : 0.0f);
                const float fz = k == 0 ? 1.0f : (k == n[2] ? -1.0f : 0.0f);
                const int faces = (fx != 0.0f) + (fy != 0.0f) + (fz != 0.0f);
                if (faces == 0) continue;
                // (i dx, j dy, k dz) + min as the shipped -O3 build rounds it: z as one FMA
                xyz.push_back(static_cast<float>(i) * d[0] + b.min[0]);
                xyz.push_back(static_cast<float>(j) * d[1] + b.min[1]);
                xyz.push_back(std::fma(static_cast<float>(k), d[2], b.min[2]));
                float o[3] = {fx, fy, fz};
                if (faces == 2) unit(fx * 0.5f, fy * 0.5f, fz * 0.5f, o);
                else if (faces == 3) unit(fx / 3.0f, fy / 3.0f, fz / 3.0f, o);
                for (int a = 0; a < 3; ++a) nrm.push_back(sgn * o[a]);
            }
    return true;
}

void samplePciSphere(const SphereDesc& s, float radius, bool inverted, std::vector<float>& xyz,
                     std::vector<float>& nrm) {
    // the count in FP32 (PCISPHBoundaryObject.cpp:192), then the points as
    // sampleSphereSurface places them (the same rounding of the shipped build)
    const float sep = 1.5f;
    const float pi = static_cast<float>(M_PI);
    const unsigned numParts = static_cast<unsigned>(
        std::floor(static_cast<float>(4.0 * M_PI) * s.radius * s.radius / (pi * radius * radius) * sep));
    if (numParts == 0) return;
    const float a = static_cast<float>(4.0 * M_PI) / static_cast<float>(numParts);
    const float d = std::sqrt(a);
    const unsigned mTheta = static_cast<unsigned>(std::round(M_PI / d));
    if (mTheta == 0) return;
    const float dTheta = pi / static_cast<float>(mTheta);
    const float dPhi = a / dTheta;
    const float invMTheta = 1.0f / static_cast<float>(mTheta);
    const float sgn = inverted ? 1.0f : -1.0f;
    for (unsigned m = 0; m < mTheta; ++m) {
        const float theta = static_cast<float>(M_PI * (m + 0.5)) * invMTheta;
        const double st = std::sin(static_cast<double>(theta)), ct = std::cos(static_cast<double>(theta));
        const int mPhi = static_cast<int>(std::round(2.0 * M_PI * st / dPhi));
        if (mPhi <= 0) continue;
        const float invMPhi = 1.0f / static_cast<float>(mPhi);
        for (int n = 0; n < mPhi; ++n) {
            const float phi = (static_cast<float>(2.0 * M_PI) * static_cast<float>(n)) * invMPhi;
            const float p[3] = {std::fma(s.radius, static_cast<float>(st * std::cos(static_cast<double>(phi))), s.pos[0]),
                                std::fma(s.radius, static_cast<float>(st * std::sin(static_cast<double>(phi))), s.pos[1]),
                                s.radius * static_cast<float>(ct) + s.pos[2]};
            for (int q = 0; q < 3; ++q) {
                xyz.push_back(p[q]);
                nrm.push_back(sgn * (s.pos[q] - p[q]));
            }
        }
    }
}

}  // namespace sph
