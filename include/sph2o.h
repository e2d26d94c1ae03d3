/*
 * sph2o.h — C ABI of the MI355X particle step (libsph2o.so).
 *
 * One context owns one or more independent scenes on one GPU and one HIP
 * stream. Every scene is a single-fluid SPH2O scene (the reference's
 * neighbour macros ignore pointSetId, SPHSimulation.h:27-32, so only one
 * fluid per scene is well defined) with Akinci boundary particles. The three
 * solvers of the reference are supported: WCSPH, PCISPH and DFSPH.
 *
 * Conventions (replacing the reference's bool returns, uncaught exceptions
 * and exit(0); SPHSimulation.cpp:299-555, SceneLoader.cpp:11-38):
 *   - every function returns an int status: SPH_OK (0) or a negative code;
 *   - sph_last_error(ctx) gives a message for the last failure;
 *   - host arrays are caller-owned and use the scene's original particle
 *     order (xyz interleaved, float32);
 *   - no function throws across the ABI.
 *
 * Reference interfaces replaced (file:line in josemanuel28v/SPH2O):
 *   sph_load_scene_json   SPHSimulation::importScene, SPHSimulation.cpp:299-481
 *                         + SceneLoader::readConfiguration, SceneLoader.cpp:11-209
 *   sph_init              SPHSimulation::init, SPHSimulation.cpp:117-126
 *                         + {WCSPH,PCISPH,DFSPH}Solver::init
 *   sph_step              SPHSolver::step (WCSPHSolver.cpp:14-78,
 *                         PCISPHSolver.cpp:20-80, DFSPHSolver.cpp:33-107)
 *                         + the time advance of SPHSimulation::step, :157-191
 *   sph_get_fluid_state   FluidObject getters (FluidObject.h:54-63),
 *                         SimulationObject::getPositions (SimulationObject.h:16)
 *   sph_set_fluid_state   FluidObject::getPosition/getVelocity writes
 *                         (SimulationObject.h:17, FluidObject.h:59)
 *   sph_dump_cells        HashTable key floor(pos/h), HashTable.cpp:5
 *   sph_dump_neighbors    HashTable::neighborhoodSearch, HashTable.cpp:23-86,
 *                         filtered by the support test of DFSPHSolver.cpp:575
 *   sph_add_emitter       FluidObject::addEmitter, FluidObject.cpp:177-193
 *                         (+ Emitter, Emitter.cpp:5-123)
 *   sph_num_active        FluidObject::getNumActiveParticles, FluidObject.h:62
 *   sph_add_cube_boundary CubeBoundaryObject, CubeBoundaryObject.cpp:4-88,
 *                         created by SPHSimulation::importScene, :457-474
 */
#ifndef SPH2O_H_
#define SPH2O_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define SPH_OK 0
#define SPH_E_ARG (-1)          /* bad argument (NULL, size mismatch) */
#define SPH_E_SCENE (-2)        /* scene JSON missing/ill-typed key or unreadable */
#define SPH_E_UNSUPPORTED (-3)  /* method/option outside the hot path (e.g. adhesion) */
#define SPH_E_HIP (-4)          /* HIP runtime error */
#define SPH_E_NONFINITE (-5)    /* NaN/Inf found in particle state */
#define SPH_E_OUT_OF_GRID (-6)  /* particles left every representable grid */
#define SPH_E_STATE (-7)        /* call order (e.g. step before init) */
#define SPH_E_CAPACITY (-8)     /* caller buffer too small */

/* method codes of SPHSimulation.cpp:21-33 */
#define SPH_WCSPH 0
#define SPH_PCISPH 1
#define SPH_DFSPH 2
#define SPH_BOUNDARY_NONE (-1)
#define SPH_BOUNDARY_CUBE 0
#define SPH_BOUNDARY_PCISPH 1
#define SPH_BOUNDARY_AKINCI 2
#define SPH_VISCOSITY_STANDARD 0
#define SPH_VISCOSITY_ARTIFICIAL 1
#define SPH_VISCOSITY_XSPH 2
#define SPH_SURFACE_TENSION_AKINCI 3

/* Scene parameters: the fields of SceneInfo / FluidInfo / Fluid
 * (SPHStructs.h:11-64) as the reference stores them (FP32). */
typedef struct sph_scene_params {
    int32_t simulation_method;      /* SPH_WCSPH / SPH_PCISPH / SPH_DFSPH */
    int32_t boundary_method;        /* SPH_BOUNDARY_AKINCI, SPH_BOUNDARY_PCISPH, SPH_BOUNDARY_CUBE or SPH_BOUNDARY_NONE */
    int32_t viscosity_method;       /* -1 or SPH_VISCOSITY_* */
    int32_t surface_tension_method; /* -1 or SPH_SURFACE_TENSION_AKINCI */
    int32_t adhesion_method;        /* -1 or 4 (Akinci adhesion, AdhesionForce.cpp:5-32; acts with Akinci boundaries) */
    uint32_t min_iterations, max_iterations;
    uint32_t min_iterations_v, max_iterations_v;
    int32_t adaptive_time_step;     /* DFSPH: 1 = CFL + iteration rule of the reference (default) */
    float particle_radius, density0;
    float gravity[3];
    float time_step, min_time_step, max_time_step, fps, cfl_factor;
    float start_time, end_time;
    float stiffness, gamma;         /* WCSPH */
    float eta, eta_v;               /* PCISPH / DFSPH tolerances */
    float viscosity, boundary_viscosity, surface_tension, adhesion;
    int32_t reserved[8];
} sph_scene_params;

/* Per-step statistics (one per scene per step). */
typedef struct sph_step_stats {
    double t, dt;
    int32_t iterations;              /* density / pressure solve iterations */
    int32_t iterations_v;            /* DFSPH divergence iterations */
    double avg_density_error;        /* FP64 reduction the solver tested last */
    double avg_divergence_error;     /* FP64 (DFSPH) */
    double max_velocity;             /* FP64 CFL reduction (DFSPH) */
    int32_t capped_density;          /* 1 if the density loop stopped at max_iterations */
    int32_t capped_divergence;       /* 1 if the divergence loop stopped at the bound */
    int32_t scene;
    int32_t reserved;
} sph_step_stats;

/* A particle emitter (EmitterInfo, SPHStructs.h; Emitter.h:14-30). Every
 * emission places one group: type 0 a square of width x height, type 1 a disc
 * of radius `width`, both at spacing 2r in the emitter's local xy plane, moving
 * along its local +z at `velocity` m/s, rotated by the quaternion `rotation`
 * (w, x, y, z) and translated to `position`. The first emission is at
 * start_time; the next follows spacing * 1.105 * 2r / velocity seconds later,
 * while at least one group of the num_particles budget is left. */
typedef struct sph_emitter_params {
    int32_t type;            /* 0 square, 1 circle */
    int32_t reserved0;
    int64_t num_particles;   /* total this emitter may emit */
    float position[3];
    float rotation[4];       /* quaternion w, x, y, z */
    float velocity;
    float start_time, width, height, spacing;
    int32_t reserved[4];
} sph_emitter_params;

typedef struct sph_ctx sph_ctx;

/* Context: device ordinal and a hipStream_t (NULL = a stream the context creates). */
int sph_create(sph_ctx** ctx, int device, void* stream);
int sph_destroy(sph_ctx* ctx);
const char* sph_last_error(const sph_ctx* ctx);
const char* sph_version(void);

/* Scenes. sph_add_scene returns the scene id (>= 0) or a negative status.
 * fluid/boundary positions are N*3 float32 host arrays; vel may be NULL (zero). */
int sph_add_scene(sph_ctx* ctx, const sph_scene_params* params, int64_t nf, const float* pos, const float* vel,
                  int64_t nb, const float* boundary_pos);
/* Reads a scene JSON of the reference schema, samples fluid blocks and
 * boundary boxes/spheres like FluidObject::build / AkinciBoundaryObject, and
 * adds it. Returns the scene id or a negative status. */
int sph_load_scene_json(sph_ctx* ctx, const char* path);
/* The same reader and samplers without a context or a GPU: fills `params`,
 * writes the fluid and boundary particle counts to *nf / *nb and, when the
 * capacities (in particles) allow, their positions (N*3 float32). Returns
 * SPH_E_CAPACITY if pos_cap or bpos_cap is too small (counts still written). */
int sph_scene_from_json(const char* path, sph_scene_params* params, int64_t* nf, float* pos, int64_t pos_cap,
                        int64_t* nb, float* bpos, int64_t bpos_cap);
/* Adds an emitter to a scene before sph_init (the scene JSON's "emitters" are
 * added by sph_load_scene_json). The scene grows by num_particles inactive
 * particles, numbered after the existing ones in emission order; they become
 * active when emitted. */
int sph_add_emitter(sph_ctx* ctx, int scene, const sph_emitter_params* emitter);
/* Adds a CubeBoundaryObject (CubeBoundaryObject.cpp:10-88) to a scene of
 * boundary method SPH_BOUNDARY_CUBE before sph_init: at the end of every step
 * the fluid is clamped into [min, max]; a clamped coordinate's outward
 * velocity is scaled by -normal_fct and the other two components by
 * tangential_fct. Scene JSONs of boundary method 0 get one per Boundary entry
 * (its first box) with both factors 0, as the reference's loader leaves them
 * (SceneLoader.cpp:168-172 reads them only for the PCISPH method). At most 4. */
int sph_add_cube_boundary(sph_ctx* ctx, int scene, const float* min, const float* max, float normal_fct,
                          float tangential_fct);
/* The PCISPH boundary object (boundary method SPH_BOUNDARY_PCISPH,
 * PCISPHBoundaryObject.h:6-48) of a scene added by sph_add_scene, before
 * sph_init: the normal of each of its nb boundary particles (n, N*3 float32,
 * as PCISPHBoundaryObject::addCube / addSphere set them) and the object's
 * normalFct / tangentialFct (SPHSimulation.cpp:428-455). Boundary particles
 * get mass density0 (2r)^3, a density and a pressure, and push the fluid with
 * the pressure term of each solver; after integration the fluid is moved out
 * of the boundary's rest radius h/2 along the averaged normals and its
 * velocity reflected (correctPositions / correctVelocities,
 * PCISPHBoundaryObject.cpp:218-373, at each solver's points: WCSPH and
 * PCISPH after the integration, DFSPH after predictVelocities and after
 * updatePositions), and PCISPH corrects its predicted positions each
 * iteration (correctPredPositions, :375-447). One boundary object per scene;
 * boxes must span at least 1 m per axis (the reference truncates extents to
 * whole metres). sph_load_scene_json sets this from the scene's Boundary
 * entry. State files and gradients of such scenes give SPH_E_UNSUPPORTED. */
int sph_set_boundary_normals(sph_ctx* ctx, int scene, int64_t nb, const float* normals, float normal_fct,
                             float tangential_fct);
/* The PCISPH boundary's normals and factors of a scene JSON without a context
 * or a GPU (the companion of sph_scene_from_json, whose boundary positions
 * they follow); SPH_E_CAPACITY if cap (in particles) is too small. */
int sph_scene_boundary_normals_from_json(const char* path, int64_t* nb, float* normals, int64_t cap, float* normal_fct,
                                         float* tangential_fct);
/* Densities and pressures of a PCISPH boundary object's particles after the
 * last step (PCISPHBoundaryObject::getDensity / getPressure), in the order of
 * sph_get_boundary. */
int sph_get_boundary_state(sph_ctx* ctx, int scene, float* density, float* pressure);
int sph_num_scenes(const sph_ctx* ctx);
int sph_get_params(const sph_ctx* ctx, int scene, sph_scene_params* out);
/* All fluid particles of a scene, inactive emitter particles included
 * (FluidObject::getSize); the state getters return that many. */
int64_t sph_num_fluid(const sph_ctx* ctx, int scene);
/* The active ones (the first sph_num_active in particle order); inactive
 * particles stay at their initial (zero) position and velocity. */
int64_t sph_num_active(const sph_ctx* ctx, int scene);
int64_t sph_num_boundary(const sph_ctx* ctx, int scene);

/* Build device state, boundary volumes, and the solver's init pass. */
int sph_init(sph_ctx* ctx);
/* Advance every scene nsteps. stats: NULL or nsteps*num_scenes entries (step-major). */
int sph_step(sph_ctx* ctx, int nsteps, sph_step_stats* stats);
/* Exact resume. sph_save_state writes the whole initialised context (every
 * scene's particles, time, time step, solver warm-start and loop state,
 * emitter progress) to a binary file; sph_load_state reads it into an
 * initialised context built from the same scenes, after which sph_step
 * continues bit-identically. The reference has no working counterpart:
 * SPHSimulation::importFluidState ends in exit(0) (SPHSimulation.cpp:548) and
 * SceneLoader::writeFluid (SceneLoader.cpp:211-275) keeps only positions and
 * velocities. */
int sph_save_state(sph_ctx* ctx, const char* path);
int sph_load_state(sph_ctx* ctx, const char* path);
int sph_synchronize(sph_ctx* ctx);

/* State I/O in original order; any pointer may be NULL. */
int sph_set_fluid_state(sph_ctx* ctx, int scene, int64_t n, const float* pos, const float* vel);
int sph_get_fluid_state(sph_ctx* ctx, int scene, float* pos, float* vel, float* density, float* pressure);
int sph_get_boundary(sph_ctx* ctx, int scene, float* pos, float* volume);
/* field: 0 alpha (DFSPH), 1 acceleration xyz (N*3), 2 predicted density / error scratch,
 * 3 k_ws, 4 kv_ws, 5 surface normals xyz (N*3) */
int sph_get_field(sph_ctx* ctx, int scene, int field, float* out);
int sph_get_time(const sph_ctx* ctx, int scene, double* t, double* dt);
/* Viewer feed: the data the reference's viewer pulls each frame (System.cpp:
 * 119-125) — the instanced vec4 positions (SimulationObject::getPositions,
 * SimulationObject.h:16, w = 1) and the scalar field |v|
 * (SPHSimulation::updateScalarField, SPHSimulation.cpp:638-657) — for the
 * scene's sph_num_active particles in original order, written by a kernel
 * into device buffers (pos4: 4 floats per particle, speed: 1; either may be
 * NULL), ordered on the context's stream. A GL or Vulkan buffer mapped into
 * HIP is such a buffer. Pointers the device cannot address give SPH_E_ARG. */
int sph_viewer_feed(sph_ctx* ctx, int scene, float* pos4, float* speed);

/* Neighbour search results at the current positions (original ids).
 * cells: N*3 int32 cell keys. neighbours: CSR offsets[N+1] into ids; each
 * row lists the fluid j with |x_i - x_j| <= h among the 27-cell candidates
 * (self included), unsorted; likewise the boundary rows. Returns
 * SPH_E_CAPACITY when cap/bcap is too small (offsets are still written). */
int sph_dump_cells(sph_ctx* ctx, int scene, int32_t* cells);
int sph_dump_neighbors(sph_ctx* ctx, int scene, int64_t* offsets, int32_t* ids, int64_t cap, int64_t* boffsets,
                       int32_t* bids, int64_t bcap);
/* sph_dump_cells without any allocation. No counterpart in the reference: it
 * is for callers that cannot allocate (code under HIP graph capture, tight
 * serving or rendering loops, harnesses that forbid hipMalloc). A kernel
 * writes the keys into a fixed device scratch array that the library holds
 * on every GPU it runs on until the process ends (12 MiB), and one copy on the
 * context's stream brings them to `cells`; no hipMalloc, hipFree or other
 * allocating call is made. Scenes of more than 1048576 active particles give
 * SPH_E_CAPACITY. Calls on one GPU share that scratch array, so calls that
 * overlap on one GPU (from several host threads or contexts) race. */
int sph_dump_cells_noalloc(sph_ctx* ctx, int scene, int32_t* cells);

/* Wall time of the last sph_step call on the GPU timeline (HIP events on the
 * context's stream) and the number of kernel launches it made. */
int sph_last_step_timing(const sph_ctx* ctx, double* total_ms, int64_t* launches);

/* Per-kernel timing with HIP events on the context's stream (no reference
 * counterpart; the reference's only timer is TimeHelper, TimeHelper.h:70-98).
 * on: 0 off; 1 or -1 every kernel; otherwise a bit mask of kernel ids
 * ((on >> id) & 1, ids 1..30). Profiled launches are bracketed by an event
 * pair and the elapsed times are summed per kernel id at the end of each
 * sph_step. */
int sph_set_profiling(sph_ctx* ctx, int on);
int sph_kernel_count(void);
const char* sph_kernel_name(int kernel_id);
int sph_kernel_stats(const sph_ctx* ctx, int kernel_id, double* total_ms, int64_t* launches);
int sph_reset_kernel_stats(sph_ctx* ctx);

/* Lanes per particle of the neighbour search's list build (no reference
 * counterpart: a launch shape of this library's search, which builds the same
 * lists for every value, the lists HashTable::neighborhoodSearch gives,
 * HashTable.cpp:23-86). lanes 1, 2, 4 or 8 pins the count; 0 (the default)
 * lets the context time each count on its first 12 searches (three rounds of
 * 1, 2, 4, 8) and keep the fastest, and time them again whenever the active
 * particles of its scenes (emitters add them) have more than doubled or
 * halved since. */
int sph_set_search_lanes(sph_ctx* ctx, int lanes);
/* The count the search uses; tuned = 1 once the timing rounds have picked it
 * (0 while they run or when the count is pinned); trial_ms (4 floats, may be
 * NULL) the fastest build seen with 1, 2, 4 and 8 lanes, in ms (infinity for
 * a count not timed). */
int sph_get_search_lanes(const sph_ctx* ctx, int* lanes, int* tuned, float* trial_ms);

/* ---- Gradients of WCSPH and DFSPH rollouts. The reference has no
 * derivatives: these entry points have no counterpart there; they serve
 * optimisation / inverse-design callers and the torch.autograd.Function in
 * sph2o_amd/autograd.py. Supported: WCSPH (WCSPHSolver::step,
 * WCSPHSolver.cpp:14-78) and DFSPH (DFSPHSolver::step, DFSPHSolver.cpp:33-107)
 * with a fixed time step and a fixed iteration count (min_iterations ==
 * min_iterations_v == max_iterations), Akinci boundaries, viscosity off or
 * XSPH (XSPHViscosity.cpp:6-42) and Akinci surface tension
 * (SurfaceTensionAkinci.cpp:5-71); other options give SPH_E_UNSUPPORTED. ---- */

/* Sets a scene's viscosity and surface-tension coefficients (the "viscosity"
 * and "surfaceTension" values SceneLoader.cpp:95-110 hands to the force
 * objects). A force acts iff its method is > -1 and its coefficient is
 * non-zero (SPHSimulation.cpp:390-394); gradients are reported for acting
 * forces. Invalidates a recorded rollout. */
int sph_set_coefficients(sph_ctx* ctx, int scene, float viscosity, float surface_tension);

/* Advances every scene nsteps exactly as sph_step does and records the
 * rollout: a checkpoint of the state every checkpoint_every steps, plus a tape
 * (state, densities, pressures, normals, neighbour lists) of one segment at a
 * time, ~(52 + 4 * list capacity) bytes per particle per step of a segment.
 * stats as in sph_step. */
int sph_forward_record(sph_ctx* ctx, int nsteps, int checkpoint_every, sph_step_stats* stats);

/* Loss gradient w.r.t. the state at the end of the recorded rollout, per
 * scene: dL/dx_T and dL/dv_T, N*3 floats in original order (NULL = zero). */
int sph_adjoint_seed(sph_ctx* ctx, int scene, const float* dL_dx, const float* dL_dv);

/* Reverse sweep: recomputes each segment's tape from its checkpoint and runs
 * the adjoint of every step backwards. Afterwards the context holds the state
 * at the end of the rollout again. Returns SPH_E_NONFINITE when a gradient of
 * any scene is NaN or infinite (the sweep has completed and
 * sph_adjoint_result still returns the values). */
int sph_adjoint_run(sph_ctx* ctx);

/* Results of the last sph_adjoint_run for one scene: dL/dx_0 and dL/dv_0
 * (N*3, original order, NULL to skip) w.r.t. the state at the start of the
 * rollout, dL/d(viscosity) and dL/d(surface tension) (NULL to skip). */
int sph_adjoint_result(sph_ctx* ctx, int scene, float* dL_dx0, float* dL_dv0, double* dL_dviscosity,
                       double* dL_dsurface_tension);

/* What the recorded rollout held at one step (0-based within the rollout), in
 * original order, for inspection and for checkers that follow the same
 * branches: field 0 / 1 the k / rho* of DFSPH density-solve pass `pass`
 * (0..iterations-1, `iterations` = the warm start), 2 / 3 the k / clamped
 * divergence of divergence-solve pass `pass`, 4 the positions at the start of
 * the step (N*3), 5 the WCSPH pressure. The step must lie in the segment the
 * tape holds (always, when checkpoint_every >= nsteps). No counterpart in the
 * reference. */
int sph_recorded_field(sph_ctx* ctx, int scene, int step, int field, int pass, float* out);

#ifdef __cplusplus
}
#endif
#endif /* SPH2O_H_ */
