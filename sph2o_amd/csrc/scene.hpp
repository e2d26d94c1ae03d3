// Scene description and the scene-JSON reader of the reference schema
// (SceneLoader.cpp:40-209, SPHStructs.h:11-106), plus the particle samplers
// (FluidObject::build, AkinciBoundaryObject::sampleCube / sampleSphere).
#pragma once

#include <string>
#include <vector>

#include "../../include/sph2o.h"

namespace sph {

struct BlockDesc { float min[3], max[3]; };
struct SphereDesc { float pos[3]; float radius; };

struct SceneDesc {
    sph_scene_params params{};
    std::vector<BlockDesc> fluidBlocks;   // FluidModel[0].fluidBlocks
    std::vector<BlockDesc> boxes;         // Boundary[*].box (all boundary objects)
    std::vector<BlockDesc> cubes;         // Boundary[i].box[0]: the cube objects of boundary method 0
                                          // (SPHSimulation.cpp:457-474)
    std::vector<SphereDesc> spheres;      // Boundary[*].sphere
    // Akinci sampling order (SPHSimulation.cpp:408-424): object by object,
    // each object's boxes, then its spheres; an entry >= 0 is boxes[e], one
    // < 0 is spheres[-1 - e]
    std::vector<int> shapeOrder;
    // boundary method 1 (PCISPHBoundaryObject): per boundary object its
    // normalFct / tangFct, per box and sphere its "inverted" flag
    // (SceneLoader.cpp:168-193)
    std::vector<float> normalFct, tangFct;
    std::vector<int> boxInverted, sphereInverted;
    int numBoundaryObjects = 0;
    std::vector<sph_emitter_params> emitters;  // FluidModel[0].emitters
    int numFluidModels = 0;
    int numEmitters = 0;
    int numGeometries = 0;
};

// Parses `text` (a scene JSON). Returns SPH_OK or SPH_E_SCENE/SPH_E_UNSUPPORTED with `err` set.
int parseSceneJson(const std::string& text, SceneDesc& out, std::string& err);

// FluidObject::build (FluidObject.cpp:30-61): lattice of spacing 2r starting at min + r.
void sampleFluidBlock(const BlockDesc& b, float radius, std::vector<float>& xyz);
// AkinciBoundaryObject::sampleCube (AkinciBoundaryObject.cpp:97-123).
void sampleBoxSurface(const BlockDesc& b, float radius, std::vector<float>& xyz);
// AkinciBoundaryObject::sampleSphere (AkinciBoundaryObject.cpp:125-153).
void sampleSphereSurface(const SphereDesc& s, float radius, std::vector<float>& xyz);
// PCISPHBoundaryObject::addCube / sampleCube (PCISPHBoundaryObject.cpp:6-185):
// points at spacing l/n with n = ceil(l / 0.45h) and l the box extent
// truncated to whole metres (the reference takes it with the integer abs),
// and the face / edge / corner normals, negated unless `inverted`. Returns
// false for an extent under 1 m on any axis (l = 0 there: the reference
// samples one point at 0/0).
bool samplePciBox(const BlockDesc& b, float supportRadius, bool inverted, std::vector<float>& xyz,
                  std::vector<float>& nrm);
// PCISPHBoundaryObject::addSphere / sampleSphere (PCISPHBoundaryObject.cpp:32-216):
// the Akinci sphere points with the point count formed in FP32, normals
// (centre - point) (length R), negated unless `inverted`.
void samplePciSphere(const SphereDesc& s, float radius, bool inverted, std::vector<float>& xyz,
                     std::vector<float>& nrm);

}  // namespace sph
