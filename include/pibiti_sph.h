/*
 * pibiti_sph.h - C ABI of the MI355X-native WCSPH step.
 *
 * The reference exposes its device layer through an extern "C" seam
 * (source/CUDA/System.cuh:1-38, source/CUDA/radixsort.cuh:30-32) that the host
 * class cSPH (source/SPH/SPH.h:9-50) drives once per frame from
 * cSPH::Update (source/SPH/SPH_Update.cpp:12-81).  This header is the
 * replacement seam: plain scalars and pointers only (no vector types, no
 * torch or HIP types), every call returns an sph_status and never exits the
 * process.  Host buffers are caller-owned and copied before the call returns.
 *
 * Threading: one context per host thread; calls on a context are serialised
 * on the context's own HIP stream.  sph_step() is asynchronous, every
 * sph_get_*() synchronises the context stream before it returns.
 */
#ifndef PIBITI_SPH_H
#define PIBITI_SPH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPH_ABI_VERSION 2u

/* ------------------------------------------------------------------------ */
/* Status codes.  The reference exits on any CUDA error
 * (source/external/cutil.h:772-829); here every failure is a return value. */
typedef enum sph_status {
    SPH_OK = 0,
    SPH_ERR_INVALID_ARGUMENT = 1, /* null pointer, bad enum, range error      */
    SPH_ERR_INVALID_PARAMS = 2,   /* h<=0, cell<h, N==0, key bits > 32 ...     */
    SPH_ERR_NO_DEVICE = 3,        /* no HIP device visible                     */
    SPH_ERR_OUT_OF_MEMORY = 4,    /* device allocation failed                  */
    SPH_ERR_DEVICE = 5,           /* a HIP runtime call failed                 */
    SPH_ERR_NOT_INITIALISED = 6,  /* step/readback before the state was set    */
    SPH_ERR_UNSUPPORTED = 7,      /* feature not built (e.g. other bnd types)  */
    SPH_ERR_CAPACITY = 8,         /* size above a fixed capacity               */
    SPH_ERR_ORDER = 9             /* a step's particle order could not be built
                                     (results invalid): reported by the next
                                     synchronising call; the state must be
                                     uploaded again before the next step     */
} sph_status;

/* Boundary types: the reference's BndType values (source/CUDA/Params.cuh:23).
 * The soft walls of each are those of boundary() (source/CUDA/System.cu:41-159). */
enum {
    SPH_BND_BOX = 0,
    SPH_BND_CYL_Y = 1,   /* cylinder around y of radius world_max[0] (no x, z walls) */
    SPH_BND_CYL_Z = 2,   /* cylinder around z of radius -world_min[1] (no x, y walls) */
    SPH_BND_CYL_YZ = 3,  /* box plus both cylinders                              */
    SPH_BND_SPHERE = 4,  /* sphere of radius -world_min[1] about the origin        */
    SPH_BND_PUMP_Y = 5   /* box plus the pump frame, outlet and inlet             */
};
/* Boundary effects along z: the reference's BndEff values
 * (source/CUDA/Params.cuh:27).                                               */
enum {
    SPH_BND_EFF_NONE = 0,
    SPH_BND_EFF_WRAP = 1,   /* no z- wall; particles leaving at z- re-enter at z+   */
    SPH_BND_EFF_CYCLE = 2,  /* no z walls; particles wrap both ways                */
    SPH_BND_EFF_WAVE = 3    /* sloped floor and a z- wall moving with the phase     */
};
/* Accelerator shapes: the reference's AccType values (source/CUDA/Params.cuh:38). */
enum { SPH_ACC_OFF = 0, SPH_ACC_BOX = 1, SPH_ACC_CYL_Y = 2, SPH_ACC_CYL_Y_SMOOTH = 3 };
#define SPH_NUM_ACCEL 4

/* Equation of state.  LINEAR is the reference's p = k (rho - rho0)
 * (source/CUDA/Kernel_Cell.cui:195).  TAIT is an opt-in extension:
 * p = (k rho0 / 7) ((rho/rho0)^7 - 1), whose slope at rho0 equals k.        */
enum { SPH_EOS_LINEAR = 0, SPH_EOS_TAIT = 1 };

/* Accumulation mode of the density and force sums.
 * REFERENCE: per-cell partial sums added in (z,y,x) cell order, exactly the
 *   loop structure of computeDensityD/computeForceD
 *   (source/CUDA/Kernel_Cell.cui:142-199, source/CUDA/System.cu:228-250).
 * KAHAN: one compensated FP32 accumulator per sum over the same candidate
 *   order (extension).
 * FAST: integer results exact (keys, order, cell ranges, neighbour sets
 *   with the cap and both predicates) and every float within 1e-5 of the
 *   reference arithmetic and of FP64 (condition-scaled) instead of bit for
 *   bit: the density as one running sum per particle with fused
 *   multiply-adds, the force as one running sum over the pair entries the
 *   density kernel writes (extension; max_par_in_cell <= 64).              */
enum { SPH_PREC_REFERENCE = 0, SPH_PREC_KAHAN = 1, SPH_PREC_FAST = 2 };

/* Arrays exchanged with the host.  POSITION/VELOCITY keep the reference's
 * 0/1 flag values of cSPH::setArray (source/SPH/SPH_Util.cpp:59-71) and are
 * float4 (x,y,z,w) per particle; DENSITY/PRESSURE are float per particle;
 * DVEL is the last step's force velocity increment as float4 (w = 0);
 * TARGET is the window-mean pair-force acceleration of a kept training
 * window (float4, w = 0; sph_learn_get_window only); NEIGHBOURS is the
 * number of neighbours within h the last surrogate step listed for each
 * particle, or after a plain step whose gathers used the pair entries
 * (SPH_PREC_FAST always; the reference and Kahan orders by default while
 * max_par_in_cell <= 64) the number of force pairs (r2 <= force_r2_max)
 * the force kernel used (uint32 per particle, in the float buffer's words);
 * SPH_ERR_UNSUPPORTED after other steps.
 * All arrays are indexed by the ORIGINAL particle index.                   */
enum {
    SPH_ARRAY_POSITION = 0,
    SPH_ARRAY_VELOCITY = 1,
    SPH_ARRAY_DENSITY = 2,
    SPH_ARRAY_PRESSURE = 3,
    SPH_ARRAY_DVEL = 4,
    SPH_ARRAY_TARGET = 5,
    SPH_ARRAY_NEIGHBOURS = 6,
    SPH_ARRAY_COLOR = 7,        /* float4 (r, g, b, 1) of the last step (clr_type != NONE) */
    SPH_ARRAY_DYE = 8           /* float dye intensity per particle          */
};

/* Colour modes: the reference's ClrType values (source/CUDA/Params.cuh:31). */
enum {
    SPH_CLR_DENS = 0, SPH_CLR_ACCEL = 1, SPH_CLR_DENS_ACC = 2, SPH_CLR_VEL = 3,  /* VEL: foam potential */
    SPH_CLR_VEL_ACC = 4, SPH_CLR_VEL_RGB = 5, SPH_CLR_NONE = 6
};
/* Dye shapes (dyeType). */
enum { SPH_DYE_OFF = 0, SPH_DYE_BOX = 1, SPH_DYE_SPHERE = 2 };

/* Context flags (sph_create). */
enum {
    SPH_FLAG_KEEP_DVEL = 1u /* force kernel also stores dv for readback */
};

/* One accelerator (struct Accel, source/CUDA/Params.cuh:40-45): particles
 * inside the shape about `pos` gain acc * dt in the step's force update
 * (computeForceD, source/CUDA/System.cu:379-399).                          */
typedef struct sph_accel {
    float pos[3];
    float size[3];
    float acc[3];
    uint32_t type;              /* SPH_ACC_*                                 */
} sph_accel;

/* ------------------------------------------------------------------------ */
/* User-level scene description: the raw values that Scene holds before
 * Scene::_UpdatePar/_UpdateGrid derive the kernel constants
 * (source/SPH/Scene.cpp:7-66 for the defaults, :76-113 for the derivation). */
typedef struct sph_scene {
    uint32_t num_particles;
    uint32_t max_par_in_cell;   /* reference default 16                      */
    float time_step;
    float gravity[3];
    float global_damping;
    float world_min[3];
    float world_max[3];
    float cell_size;            /* uniform cell edge; must be >= h           */
    float particle_r;           /* the reference's particleR                 */
    float h;
    float rest_density;
    float stiffness;
    float viscosity;
    float min_dens;             /* factor: minDens <- 1/(min_dens*rho0)^2    */
    float min_dist;             /* factor of particle_r                      */
    float particle_mass;        /* > 0: explicit; 0: rho0*4/3*PI*r^3          */
    float dist_bnd_soft;        /* factor of particle_r                      */
    float dist_bnd_hard;        /* factor of particle_r                      */
    float bnd_stiff;
    float bnd_damp;
    float bnd_damp_c;
    uint32_t eos;               /* SPH_EOS_*                                 */
    uint32_t precision;         /* SPH_PREC_*                                */
    float cfl;                  /* 0: fixed dt; > 0: dt = min(time_step, cfl*h/(c+vmax)) */
    /* Boundary shape and z effect (boundary(), source/CUDA/System.cu:41-159),
     * rotor and height map (computeForceD, source/CUDA/System.cu:253-372),
     * and the SimParams values they read (source/CUDA/Params.cuh:98-113).
     * As in the reference, several features share one value.              */
    uint32_t bnd_type;          /* SPH_BND_*                                 */
    uint32_t bnd_eff_z;         /* SPH_BND_EFF_*                             */
    float rot2_angle;           /* r2Angle: wave floor slope (HSlope), height-map
                                   level (Hheight), second propeller's angle
                                   (Rotor2Angle, degrees in the file)        */
    float rot_twist;            /* rTwist: wave amplitude (WaveAmpl), rotor twist */
    float rot_angle;            /* rAngle: wave phase, rotor angle (rad)     */
    float rot_vel;              /* rVel (WaveSpeed, RotorVel): every step first
                                   advances rot_angle by rot_vel * time_step
                                   (source/App/Update.cpp:15-20)             */
    float exit_vel;             /* ExitVel (rVexit)                          */
    float exit_dist;            /* ExitDist (rDexit)                         */
    float pump_ang_out;         /* PumpAngOut                                */
    float h_close;              /* hClose: pump lid (PumpHClose), height-map
                                   amplitude (Hscale)                        */
    float pump_rad_in;          /* PumpRadIn (factor of world_max[0])        */
    float shape_s[6];           /* s1..s6: pump shape; height-map frequencies,
                                   phases (degrees) and threshold (s1..s5)   */
    /* DEM sphere collider (collideSpheres, source/CUDA/Kernel_Cell.cui:78-96).
     * computeForceD applies it in every scene without a rotor: it is the
     * else branch of the rotor test (source/CUDA/System.cu:310,373-375).  */
    uint32_t collider;          /* 1: on (where no rotor runs)               */
    float coll_pos[3];          /* centre (also the rotor's); coll_pos[0] == -77
                                   places it at the world's x-/y- edge
                                   (source/SPH/Scene.cpp:111-112)            */
    float coll_r;               /* ColliderR                                 */
    float coll_spring, coll_damping, coll_shear; /* also the rotor's and height map's */
    sph_accel acc[SPH_NUM_ACCEL];
    /* Rotor (source/CUDA/System.cu:309-372): spheres of radius rot_r laid
     * out in blades about coll_pos and turning with rot_angle.             */
    int32_t rot_type;           /* SPH_ROTOR_*                               */
    int32_t rot_size[3];        /* RotorSizes: blade extents (x, y, z)       */
    int32_t rot_blades;         /* RotorBlades                               */
    float rot_r;                /* colParR: sphere radius, factor of particle_r */
    float rot_spc;              /* colParSpc: sphere spacing, factor of the radius */
    float rot2_dist;            /* Rotor2Dist: propeller pair separation     */
    float rot2_vel;             /* Rotor2Vel: second propeller's speed factor */
    float rot2_twist;           /* r2twist: second propeller's twist sign    */
    /* Height map (source/CUDA/System.cu:253-307): a sheet of spheres of
     * radius rot_r, spacing rot_spc, at height
     * h_close sin(s1 u + s2) sin(s3 w + s4) + rot2_angle above world_min.  */
    int32_t hmap_type;          /* SPH_HMAP_*                                */
    /* Particle colours and dye (computeForceD's colouring,
     * source/CUDA/System.cu:405-546): a per-particle colour for a renderer
     * and a dye intensity that particles pick up in the dye region and lose
     * at dye_fade per second.                                               */
    uint32_t clr_type;          /* SPH_CLR_* (SPH_CLR_NONE: no colour output) */
    int32_t hue;                /* iHue: 1 = (intensity, dye, b) channels     */
    float brightness, contrast;
    int32_t dye_type;           /* SPH_DYE_*                                 */
    int32_t dye_clear;          /* steps left that clear the dye: every step
                                   first counts it down by one, and a step that
                                   sees it above 0 clears the dye (source/App/Update.cpp:23) */
    float dye_fade;             /* intensity lost per second                 */
    float dye_pos[3];
    float dye_size[3];          /* box half sizes; the sphere's radius is dye_size[1] */
} sph_scene;

/* Derived kernel constants: the plain-scalar equivalent of SimParams
 * (source/CUDA/Params.cuh:53-83) restricted to the hot path, plus the
 * values this build derives once on the host.                              */
typedef struct sph_params {
    float time_step;
    uint32_t num_particles;
    uint32_t max_par_in_cell;
    float gravity[3];
    float global_damping;
    uint32_t grid_size[3];
    float cell_size[3];
    uint32_t grid_size_yx;
    uint32_t num_cells;
    float world_min[3];
    float world_max[3];
    float particle_r, h, h2;
    float spiky_kern, lap_kern, poly6_kern;
    float particle_mass, rest_density, stiffness, viscosity;
    float min_dens;             /* 1/(min_dens*rho0)^2                       */
    float min_dist;             /* min_dist * particle_r                     */
    float dist_bnd_hard, dist_bnd_soft;
    float bnd_damp, bnd_stiff, bnd_damp_c;
    uint32_t bnd_type;          /* SPH_BND_*                                 */
    uint32_t bnd_eff_z;         /* SPH_BND_EFF_*                             */
    uint32_t eos;
    uint32_t precision;
    float cfl;
    float sound_speed;          /* sqrt(k): dp/drho at rho0 for both EOS     */
    uint32_t key_bits;          /* bits of the cell key the sort orders      */
    /* Largest float r2 with max(min_dist, sqrtf(r2)) < h, i.e. the force
     * predicate of compForcePair (source/CUDA/Kernel_Cell.cui:214-218)
     * rewritten as one comparison r2 <= force_r2_max; -1 when min_dist >= h. */
    float force_r2_max;
    /* Launch tuning (0 = library default); any legal value gives the same
     * results.  tune[0]: force kernel variant, 1 = fused candidate and pair
     * loop, 16/24/32/48 = neighbour-list entries per lane; tune[1]: gather
     * grid, 0 = one thread per particle, else a persistent grid of tune[1]
     * blocks of 256 threads pulling 256-particle tiles; tune[2]: surrogate
     * steps per host check in sph_step_learned (0 = 4, at most 64);
     * tune[3]: gather variant with tune[1] == 0 (0 = per-lane loads in
     * XCD-ordered tiles, 1 = candidate positions staged in LDS per tile,
     * 2 = neighbour-row cell tables staged in LDS, 3 = per-lane loads in
     * launch order).                                                       */
    uint32_t tune[4];
    /* First global cell key of this scene: 0 for a single-scene context; a
     * batch gives each scene a disjoint key range (set by sph_batch_create). */
    uint32_t cell_base;
    /* Simulated end time (s): the time step is clipped so the scene's clock
     * stops at t_end, after which the scene is frozen.  0 = no end time.    */
    float t_end;
    /* Boundary shape, rotor and height map: the sph_scene values, with
     * rot_r and rot_spc scaled to metres as Scene::_UpdatePar does
     * (rot_r *= particle_r, then rot_spc *= rot_r, source/SPH/Scene.cpp:93). */
    float rot2_angle, rot_twist, rot_angle, rot_vel;
    float exit_vel, exit_dist;
    float pump_ang_out, h_close, pump_rad_in;
    float shape_s[6];
    /* Derived from the fields above whenever the library takes parameters
     * (values set by the caller are replaced):
     *   wave_zmin_shift = rot_twist (1 + sin(rot_angle)), the moving z- wall
     *     of the wave effect (source/CUDA/System.cu:60);
     *   pump_xs, pump_zs: the pump outlet's half width and the height above
     *     which its side walls act (source/CUDA/System.cu:104-117);
     *   world_size_d: worldSizeD, the world less the soft band less r on
     *     each side (source/SPH/Scene.cpp:99-103), the height map's period.
     * sin and cos here and in the per-particle terms of the pump, rotor and
     * height map are correctly rounded float results.                     */
    float wave_zmin_shift, pump_xs, pump_zs;
    float world_size_d[3];
    /* Collider (centre already placed) and accelerators, as in sph_scene. */
    uint32_t collider;
    float coll_pos[3], coll_r, coll_spring, coll_damping, coll_shear;
    sph_accel acc[SPH_NUM_ACCEL];
    int32_t rot_type, rot_size[3], rot_blades;
    float rot_r, rot_spc, rot2_dist, rot2_vel, rot2_twist;
    int32_t hmap_type;
    uint32_t clr_type;
    int32_t hue;
    float brightness, contrast;
    int32_t dye_type, dye_clear;
    float dye_fade, dye_pos[3], dye_size[3];
    /* Derived: SPH_PHYS_* bits of the terms in use.                         */
    uint32_t physics;
    uint32_t reserved[2];
} sph_params;

/* Rotor kinds (RotorType) and height-map planes (HmapType). */
enum { SPH_ROTOR_NONE = 0, SPH_ROTOR_PUMP_Z = 1, SPH_ROTOR_PUMP_Y = 2, SPH_ROTOR_PROPELLER_Z = 3 };
enum { SPH_HMAP_NONE = 0, SPH_HMAP_XZ = 1, SPH_HMAP_YZ = 2 };

/* sph_params::physics bits: the force-update terms in use.  The collider
 * bit is set only when no rotor runs.                                      */
enum { SPH_PHYS_COLLIDER = 1u, SPH_PHYS_ACCEL = 2u, SPH_PHYS_ROTOR = 4u, SPH_PHYS_HMAP = 8u };

/* Global FP64 diagnostics reduced on the device (extension). */
typedef struct sph_stats {
    double mass;                /* sum m                                     */
    double kinetic_energy;      /* sum 1/2 m |v|^2                           */
    double potential_energy;    /* sum -m g.x                                */
    double max_speed;           /* max |v|                                   */
    double dt_cfl;              /* time step the CFL rule picked (== dt when cfl == 0) */
    double sim_time;            /* simulated seconds advanced so far         */
    uint64_t steps;             /* steps advanced so far                     */
    uint32_t cap_cells;         /* cells holding more than max_par_in_cell   */
    uint32_t out_of_grid;       /* particles whose cell had to be clamped    */
    uint32_t max_cell_count;    /* largest particle count of one cell        */
    uint32_t sort_errors;       /* steps since the last sph_get_stats whose incremental order
                                   failed a consistency check and that were re-sorted in full
                                   instead (their results are valid); a failure that could not
                                   be repaired is SPH_ERR_ORDER instead */
    uint32_t full_sort;         /* 1: the last step ordered the slots with the full radix sort,
                                   0: incrementally from the previous order (same order)      */
    uint32_t reserved[5];
} sph_stats;

typedef struct sph_ctx sph_ctx;

/* ------------------------------------------------------------------------ */
/* Host-only helpers (no device needed). */

/* ABI version of the loaded library (compare with SPH_ABI_VERSION). */
uint32_t sph_abi_version(void);

/* Human-readable text of a status code (never NULL). */
const char* sph_strerror(int status);

/* Scene::InitDefault (source/SPH/Scene.cpp:7-66), hot-path fields only. */
int sph_scene_default(sph_scene* scene);

/* Scene::_UpdatePar + Scene::_UpdateGrid (source/SPH/Scene.cpp:76-113).
 * Unlike the reference (which scales minDist/bnd distances in place and so
 * must run once), this is a pure function of the scene.                   */
int sph_params_from_scene(const sph_scene* scene, sph_params* params);

/* Lattice initialiser in the spirit of cSPH::Reset(0)
 * (source/SPH/SPH_Init.cpp:40-76): particle n = (iy*nz + iz)*nx + ix sits at
 * region_min + (i + 1/2)*spacing on each axis, jittered by
 * U(-jitter, +jitter)*spacing from the seeded counter-based generator, with
 * velocity N(0, vel_sigma^2) per axis plus vel_offset.  pos4 and vel4 hold
 * nx*ny*nz float4 each (w = 1 for positions, 0 for velocities).           */
int sph_scene_lattice(uint32_t nx, uint32_t ny, uint32_t nz, const float region_min[3],
                      float spacing, float jitter, float vel_sigma, const float vel_offset[3],
                      uint64_t seed, uint64_t stream, float* pos4, float* vel4);

/* ------------------------------------------------------------------------ */
/* Scene files (host only).  A scene file is the reference's XML scene list
 * (Scenes.xml): a root element whose <Scene> children carry the values of
 * Scene::_FromXML as attributes (source/SPH/Scene_Load.cpp:11-111), with
 * <Emitter> and <Accel> children.                                          */

/* One emitter (class Emitter, source/SPH/Scene.h:7-17). */
typedef struct sph_emitter {
    int32_t size;               /* width in particles (0 = off)              */
    int32_t size2;              /* height (0 = square)                       */
    float pos[3];               /* centre                                    */
    float rot[3];               /* rotation angles about x and y (degrees)   */
    float vel;                  /* emission speed                            */
} sph_emitter;

/* Scene values that are not simulation parameters: the init volume and fill
 * of cSPH::Reset, the drops and emitters of the app's frame update, and the
 * camera.                                                                  */
typedef struct sph_scene_info {
    char title[64];             /* name attribute ("No name" without one)   */
    uint32_t chapter;           /* starts a chapter (always for scene 0)    */
    uint32_t is_default;        /* reserved (0)                              */
    float init_min[3];          /* init volume after Scene::_UpdateGrid moved */
    float init_max[3];          /* it in by (soft band - particle_r)          */
    int32_t init_type;          /* InitType: 1 random, otherwise lattice (9: no shape test) */
    int32_t init_last;          /* InitLast: axis advanced last, 0 x, 1 y, 2 z */
    float spacing;              /* lattice step (m): spacing * particle_r    */
    float drop_r;               /* dropR                                     */
    int32_t rain;               /* rain interval (frames)                    */
    sph_emitter emit[4];
    float cam_pos[3];
    float cam_rot[3];
    uint32_t reserved[4];
} sph_scene_info;

/* cSPH::LoadScenes (source/SPH/SPH_Scenes.cpp:53-83): the number of scenes
 * in the file text (len bytes) and the index of the startup scene (the last
 * one with a default or def attribute, else 0).  A file whose markup breaks
 * off keeps the scenes read before the break, as the reference's reader
 * does.                                                                    */
int sph_xml_scene_count(const char* text, size_t len, uint32_t* count, uint32_t* default_index);

/* Scene `index` of the file: Scene::InitDefault then Scene::_FromXML
 * (source/SPH/Scene.cpp:7-66, source/SPH/Scene_Load.cpp:11-111), as an
 * sph_scene for sph_params_from_scene (raw factors, as the reference keeps
 * them before _UpdatePar) and the scene's other values (info may be NULL).
 * The collider is on (the force update applies it where no rotor runs).  */
int sph_xml_scene(const char* text, size_t len, uint32_t index, sph_scene* scene, sph_scene_info* info);

/* cSPH::Reset(initType) (source/SPH/SPH_Init.cpp:23-76): scene->num_particles
 * positions (float4, w = 1) and zero velocities.  The lattice walk starts at
 * info->init_min (y raised by particle_r), steps by info->spacing with the
 * init_last axis order, keeps the sites inside the boundary shape and
 * outside the collider, and wraps around the volume until every particle
 * has a site.  init_type 1 fills the world at random from `seed` (the
 * reference draws rand()).  SPH_ERR_INVALID_PARAMS if no site of the
 * volume qualifies.                                                        */
int sph_scene_fill(const sph_scene* scene, const sph_scene_info* info, uint64_t seed, float* pos4,
                   float* vel4);

/* Particle injection between steps (host only): the records a drop, rain
 * or an emitter writes through cSPH::setArray; upload them with
 * sph_set_array (positions and velocities) at the caller's write index.   */

/* Largest emitter block (size x size2) sph_scene_emit writes.             */
#define SPH_MAX_EMIT 65536

/* cSPH::Drop (source/SPH/SPH_Init.cpp:82-118): a ball of radius info->drop_r
 * lattice steps (info->spacing) at rest, centred in the world inside the
 * soft band, or with random = 1 at a random place (high over a box or
 * cylinder-Y floor; the reference draws rand(), here the seeded generator,
 * draw number `draw`).  Writes the records that fit below num_particles
 * starting at particle `first` (as the reference stops there): *count
 * records to pos4/vel4 (float4, w = 1 and 0).                             */
int sph_scene_drop(const sph_scene* scene, const sph_scene_info* info, int random, uint64_t seed,
                   uint64_t draw, uint32_t first, float* pos4, float* vel4, uint32_t* count);

/* One frame of emitter `emitter` (App::UpdateEmitter, source/App/Update.cpp:62-92):
 * a size x size2 block of particles one spacing ahead of the emitter, laid
 * out and moving along the emitter's direction (rotations rot[0] about x
 * and -rot[1] about y, degrees, as the reference's GL modelview).  *count
 * records (0 for an emitter that is off, at most SPH_MAX_EMIT).           */
int sph_scene_emit(const sph_scene_info* info, uint32_t emitter, float* pos4, float* vel4, uint32_t* count);

/* ------------------------------------------------------------------------ */
/* Device context (replaces cSPH::_InitMem/_FreeMem + setParameters,
 * source/SPH/SPH_Mem.cpp:11-82, source/CUDA/System.cu:634-637).            */

/* Allocate every device buffer for params->num_particles particles on the
 * given HIP device.  flags: SPH_FLAG_*.                                    */
int sph_create(const sph_params* params, int device, uint32_t flags, sph_ctx** out);

/* Free the context (NULL is accepted). */
int sph_destroy(sph_ctx* ctx);

/* setParameters (source/CUDA/System.cu:634-637).  num_particles and the
 * grid shape must not change; every other constant may.                   */
int sph_set_params(sph_ctx* ctx, const sph_params* params);

/* cSPH::setArray (source/SPH/SPH_Util.cpp:59-71): copy `count` records
 * starting at original index `first` into the device state.  which is
 * SPH_ARRAY_POSITION or SPH_ARRAY_VELOCITY, data is float4 per particle.  */
int sph_set_array(sph_ctx* ctx, int which, const float* data, uint32_t first, uint32_t count);

/* cSPH::getArray (source/SPH/SPH_Util.cpp:44-56) plus density, pressure
 * and dv: all num_particles records in original index order.              */
int sph_get_array(sph_ctx* ctx, int which, float* out);

/* cSPH::Update (source/SPH/SPH_Update.cpp:12-81) run nsteps times:
 * integrate (boundary, gravity, x) -> hash -> stable sort -> reorder with
 * cellStart/cellEnd -> density + EOS -> force + velocity update.          */
int sph_step(sph_ctx* ctx, uint32_t nsteps);

/* Grid of the last step: sorted cell keys and permutation (sorted slot ->
 * original index) as the reference's dParHash after RadixSort
 * (source/SPH/SPH_Update.cpp:56-58), and cellStart/cellEnd per cell
 * (0xFFFFFFFF for an empty cell, as source/CUDA/System.cu:694).  Any
 * pointer may be NULL to skip that array.                                  */
int sph_get_grid(sph_ctx* ctx, uint32_t* sorted_hash, uint32_t* perm, uint32_t* cell_start,
                 uint32_t* cell_end);

/* Device arrays of the current state, for a renderer that reads them in
 * place (the role of the reference's GL buffer interop,
 * source/CUDA/System.cu:556-560): positions, velocities and colours per
 * sorted slot (float4; colour4 is NULL without a colour mode), the original
 * particle index of each slot, the context's HIP stream (hipStream_t) and
 * device.  The pointers stay valid until the next call that steps, uploads
 * or changes parameters; wait for `stream` before reading them.           */
typedef struct sph_device_view {
    const float* pos4;
    const float* vel4;
    const float* colour4;
    const uint32_t* index;
    void* stream;
    uint32_t n;
    int32_t device;
} sph_device_view;

int sph_get_device_view(sph_ctx* ctx, sph_device_view* out);

/* FP64 reductions of the current state plus counters of the last step.   */
int sph_get_stats(sph_ctx* ctx, sph_stats* out);

/* Block until every queued step has finished. */
int sph_synchronize(sph_ctx* ctx);

/* Device-side time (ms) of the kernels of the last sph_step call, measured
 * with HIP events on the context stream: [0] integrate+hash, [1] sort,
 * [2] reorder, [3] density, [4] force, [5] whole call.  Timing is only
 * recorded while enabled.                                                  */
int sph_set_timing(sph_ctx* ctx, int enable);
int sph_get_timing(sph_ctx* ctx, float* ms6);

/* Record nsteps steps (even: the ping-pong buffers end where they began)
 * into a hipGraph, then replay it: each replay advances nsteps steps with one
 * graph launch.  No counterpart in the reference (extension).  Any later
 * parameter change discards the recording.                                 */
int sph_capture(sph_ctx* ctx, uint32_t nsteps);
int sph_replay(sph_ctx* ctx, uint32_t times);

/* ------------------------------------------------------------------------ */
/* Batches: many independent scenes of uneven particle counts advanced by
 * the same launches (extension; the reference runs one scene per process).
 * Scenes are packed one after another; each keeps its own parameters, grid,
 * time step (fixed, CFL or clipped at t_end) and clock, and its results are
 * those of the same scene run alone.  All scenes share one precision mode.
 * Array and stats calls address one scene with scene-local indices.        */
typedef struct sph_batch sph_batch;

int sph_batch_create(const sph_params* scenes, uint32_t nscenes, int device, uint32_t flags,
                     sph_batch** out);
int sph_batch_destroy(sph_batch* batch);
/* Parameters of a scene as the batch holds them (cell_base, key_bits set). */
int sph_batch_get_params(sph_batch* batch, uint32_t scene, sph_params* out);
int sph_batch_set_params(sph_batch* batch, uint32_t scene, const sph_params* params);
int sph_batch_set_array(sph_batch* batch, uint32_t scene, int which, const float* data,
                        uint32_t first, uint32_t count);
int sph_batch_get_array(sph_batch* batch, uint32_t scene, int which, float* out);
/* sph_get_grid of one scene, in scene-local keys, indices and slots: the
 * scene's sorted keys less its first global key, its permutation (sorted
 * slot -> scene-local original index) and its cells' ranges.             */
int sph_batch_get_grid(sph_batch* batch, uint32_t scene, uint32_t* sorted_hash, uint32_t* perm,
                       uint32_t* cell_start, uint32_t* cell_end);
int sph_batch_step(sph_batch* batch, uint32_t nsteps);
int sph_batch_capture(sph_batch* batch, uint32_t nsteps);
int sph_batch_replay(sph_batch* batch, uint32_t times);
int sph_batch_get_stats(sph_batch* batch, uint32_t scene, sph_stats* out);
int sph_batch_synchronize(sph_batch* batch);
/* Launch tuning of every scene (sph_params::tune). */
int sph_batch_set_tune(sph_batch* batch, const uint32_t tune[4]);
/* Time each launch configuration (force list size x gather grid) over a few
 * steps of the live state, restore positions, velocities and clocks, keep
 * the fastest and append it to cache_path (JSON lines; NULL = no cache).  A
 * cached entry for the same device and batch shape is used without timing.
 * tune_out receives the choice, best_ms its time per step.                */
int sph_batch_autotune(sph_batch* batch, const char* cache_path, uint32_t tune_out[4],
                       float* best_ms);
int sph_batch_info(sph_batch* batch, uint32_t* nscenes, uint32_t* total_particles,
                   uint32_t* total_cells, uint32_t* key_bits);

/* ------------------------------------------------------------------------ */
/* The reference's RadixSort(KeyValuePair* pData0, KeyValuePair* pData1,
 * uint elements, uint bits) (source/CUDA/radixsort.cuh:30-32, driver
 * source/CUDA/radixsort_kernel.cu:445-472): a stable LSD sort of n pairs
 * by the low `bits` bits of the key, 8 bits per pass.  pairs holds n
 * (uint32 key, uint32 value) records interleaved, the KeyValuePair layout,
 * and receives the sorted pairs; temp (n records) is the ping-pong buffer.
 * Both are device pointers on the current HIP device; stream is a
 * hipStream_t (NULL: the null stream).  Unlike the reference, the result is
 * in pairs for every key width (the reference leaves it there for its 16-
 * and 32-bit calls), the last pass of a width that is not a multiple of 8
 * takes only the remaining bits (the reference's passes take whole bytes,
 * so it orders by the low 8 ceil(bits / 8) bits; the same for keys below
 * 2^bits and for its 16- and 32-bit calls), and the call waits for the
 * sort: it returns
 * SPH_ERR_ORDER if a device-side wait of the sort hit its bound (pairs then
 * hold no valid order).  Scratch is allocated and freed by the call.      */
#define SPH_SORT_MAX (1u << 27)
int sph_radix_sort(uint32_t* pairs, uint32_t* temp, uint32_t n, uint32_t bits, void* stream);

/* sph_radix_sort without any device allocation (no counterpart in the
 * reference): for callers that cannot allocate, such as tight serving loops
 * and harnesses that forbid hipMalloc.  Its scratch is a fixed array in
 * device memory, one per GPU, kept until the process ends (about 0.8 MB);
 * n above SPH_SORT_NOALLOC_MAX is SPH_ERR_CAPACITY.  Every call on a GPU
 * uses that GPU's array, so calls on one GPU are serialised by a lock in
 * the library and each waits for its sort: they do not overlap within a
 * process.                                                                 */
#define SPH_SORT_NOALLOC_MAX (1u << 20)
int sph_radix_sort_noalloc(uint32_t* pairs, uint32_t* temp, uint32_t n, uint32_t bits, void* stream);

/* ------------------------------------------------------------------------ */
/* Testing hook (no counterpart in the reference): perturbs one device word
 * of the step's order state between steps, so that tests can drive the
 * consistency checks from the host while every kernel runs normally.
 *   SPH_DEBUG_MOVER_COUNT: adds value to cell `index`'s mover count of the
 *     next step (the step then counts that cell wrong);
 *   SPH_DEBUG_CELL_END: adds value to cell `index`'s end in the current
 *     cell ranges (the next step's old ranges);
 *   SPH_DEBUG_ORDER_ERROR: marks the current order as failed, as a step
 *     whose order could not be built does (index and value unused).
 * Synchronous.                                                             */
enum { SPH_DEBUG_MOVER_COUNT = 1, SPH_DEBUG_CELL_END = 2, SPH_DEBUG_ORDER_ERROR = 3 };
int sph_debug_poke(sph_ctx* ctx, int what, uint32_t index, uint32_t value);

/* Testing hook (no counterpart in the reference): makes the library's
 * after-th device allocation from now on fail with SPH_ERR_OUT_OF_MEMORY,
 * as a full device would (after = 1: the next one); 0 disarms it.  It fires
 * once and disarms itself.  Process-wide.                                  */
int sph_debug_fail_alloc(uint32_t after);

/* ------------------------------------------------------------------------ */
/* Self-test of the BF16 matrix-core operand layouts the learned surrogate
 * uses (no counterpart in the reference).  variant 0: A 32x16, B 16x32,
 * C 32x32; variant 1: A 16x32, B 32x16, C 16x16.  A, B row-major bf16 bit
 * patterns; C row-major float.  Synchronous, uses device 0.               */
int sph_mfma_selftest(int variant, const uint16_t* a, const uint16_t* b, float* c);

/* ------------------------------------------------------------------------ */
/* Learned pair-force surrogate (extension; no counterpart in the reference).
 * For every neighbour pair of particle i (|x_i - x_j| < h, j != i, the
 * density kernel's pairs) the pair network sees 16 features
 *   e = [(x_i - x_j) inv_h (3), |x_i - x_j| inv_h, s (rho_i inv_rho0 - 1), 1, 0, 0,
 *        (v_j - v_i) inv_v (3), s (rho_j inv_rho0 - 1), 0 (4)]       (s = rho_scale)
 *   m_ij = relu(W2 relu(W1 e) + b2)       W1: 64 x 16 (column 5 = bias)
 * summed over the neighbours, H_i = sum_j m_ij (m rounded to BF16, the sum
 * accumulated in FP32 on the matrix cores).  With the node features
 *   q_i = [s (rho_i inv_rho0 - 1), p_i inv_p, nnb_i inv_nnb]
 * the acceleration is a_i = a_scale (W3 [agg_scale H_i ; q_i] + b3), W3: 3 x 67.
 * The pair network runs on BF16 matrix cores (BF16 weights and features,
 * FP32 accumulation); the node layer and all sums are FP32.
 * Parameter vector (floats): W1 [64][16] at 0, W2 [64][64] at 1024,
 * b2 [64] at 5120, W3 [3][67] at 5184, b3 [3] at 5385.                     */
#define SPH_MLP_NUM_PARAMS 5388

typedef struct sph_mlp_norm {
    float inv_h;      /* relative position unit (1/m)                       */
    float inv_v;      /* relative velocity unit (s/m)                       */
    float inv_rho0;   /* 1 / rest density                                   */
    float inv_p;      /* pressure unit                                      */
    float inv_nnb;    /* neighbour-count unit                               */
    float agg_scale;  /* factor on the neighbour sum H                      */
    float a_scale;    /* output unit (m/s^2)                                */
    float rho_scale;  /* gain on the density features (0 means 1)          */
} sph_mlp_norm;

typedef struct sph_model sph_model;

/* Weights drawn from the seed (He-normal W1, W2; small W3; zero biases). */
int sph_model_create(int device, const sph_mlp_norm* norm, uint64_t seed, sph_model** out);
int sph_model_destroy(sph_model* model);
int sph_model_get_norm(sph_model* model, sph_mlp_norm* out);
int sph_model_get_weights(sph_model* model, float* out);     /* SPH_MLP_NUM_PARAMS */
int sph_model_set_weights(sph_model* model, const float* in); /* clears Adam state */

/* Training windows (contexts created with SPH_FLAG_KEEP_DVEL).  Called after
 * a step, sph_learn_begin keeps the state the step's force kernel saw:
 * positions, velocities before the force update, densities and cell lists.
 * The target is the sum of the pair-force velocity increments of that step
 * and of every later step followed by sph_learn_add, divided by
 * (steps in the window x time_step): the mean pair-force acceleration over
 * the window.  Loss: mean over particles of |a / a_scale - target / a_scale|^2
 * summed over the three components.                                        */
int sph_learn_begin(sph_ctx* ctx, sph_model* model);
int sph_learn_add(sph_ctx* ctx, sph_model* model);
/* iters Adam steps (beta 0.9/0.999, eps 1e-8) on the kept window;
 * loss_out (may be NULL) receives the loss of the last forward pass.      */
int sph_learn_train(sph_ctx* ctx, sph_model* model, uint32_t iters, float lr, double* loss_out);
/* Loss and gradient (SPH_MLP_NUM_PARAMS floats) on the kept window. */
int sph_learn_gradient(sph_ctx* ctx, sph_model* model, double* loss_out, float* grad_out);
/* Surrogate acceleration a_i on the kept window (float4, original order). */
int sph_learn_predict(sph_ctx* ctx, sph_model* model, float* accel4_out);
/* Arrays of the kept window in original order: POSITION, VELOCITY (before
 * the force update), DENSITY, PRESSURE, TARGET.                            */
int sph_learn_get_window(sph_ctx* ctx, int which, float* out);

int sph_batch_learn_begin(sph_batch* batch, sph_model* model);
int sph_batch_learn_add(sph_batch* batch, sph_model* model);
int sph_batch_learn_train(sph_batch* batch, sph_model* model, uint32_t iters, float lr,
                          double* loss_out);
int sph_batch_learn_gradient(sph_batch* batch, sph_model* model, double* loss_out,
                             float* grad_out);
int sph_batch_learn_get_window(sph_batch* batch, uint32_t scene, int which, float* out);

typedef struct sph_learned_report {
    uint32_t coarse_steps;  /* surrogate steps kept                         */
    uint32_t fallbacks;     /* surrogate steps undone and replaced          */
    uint32_t plain_steps;   /* plain steps run in their place               */
    uint32_t rebuilds;      /* sorts of skin > 0 runs, replays included     */
    double drift_max;       /* largest |D - D_anchor| of a kept step        */
    double drift_last;      /* |D - D_anchor| of the last surrogate step    */
    double anchor;          /* D of the last plain step                     */
    double sim_time;        /* clock after the call                         */
    /* With sph_set_timing on: mean device time (ms) per kept surrogate step
     * of the density-and-neighbour-list kernel, of the surrogate kernel and
     * of the whole step including its share of the sorts and snapshots
     * (HIP events on the context stream), and of one sort with its reorder
     * (skin > 0); else 0.                                                  */
    double density_ms, mlp_ms, coarse_ms;
    double pairs;           /* mean neighbour pairs per kept surrogate step  */
    double rebuild_ms;
    uint32_t early_rebuilds; /* runs cut short: a particle moved skin h      */
    uint32_t reserved0;
} sph_learned_report;

/* Advance ncoarse surrogate steps of factor x time_step each: the plain
 * integrate (coarse step), the density and neighbour lists of the new
 * positions, then v += a dt_coarse from the surrogate.
 *
 * skin = 0: every step sorts, reorders and walks the 27 neighbour cells.
 * skin > 0: the particles are sorted once per tune[2] steps; the steps in
 * between keep that sort's slot order and cell ranges and walk every cell
 * within (1 + skin) h.  A step in which some particle has moved more than
 * skin h since the sort ends the run early: it is taken again after a new
 * sort.  Either way each step lists exactly the pairs within h.
 *
 * D = mean_i |rho_i / rho0 - 1| of each surrogate step is compared with D
 * of the last plain step (the anchor); a step whose drift |D - anchor|
 * exceeds tol (or whose neighbour lists overflow) is undone and replaced
 * by `factor` plain steps, which set a new anchor.  D is read back once per
 * tune[2] steps (the steps are deterministic, so a late check gives the
 * same trajectory as a check after every step).  tol bounds this drift of
 * the mean only, not each particle's error against the plain trajectory:
 * over 1,000 dt of the 1M dam break a 1% tol leaves the surrogate about 2%
 * RMS away from the plain run.  The current state must come from a step.  Single-scene contexts with a fixed time step (cfl = 0,
 * t_end = 0).                                                              */
int sph_step_learned(sph_ctx* ctx, sph_model* model, uint32_t ncoarse, uint32_t factor, float tol,
                     float skin, sph_learned_report* out);

#ifdef __cplusplus
}
#endif

#endif /* PIBITI_SPH_H */
