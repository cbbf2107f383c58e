/* dem_mph.h — C ABI of the MI355X DEM-MPH explicit-step library (libdemmph.so).
 *
 * The reference solver (source/DEM-MPH.cpp) has no library interface: a single
 * `main` reads a .data parameter file and a .grid particle file, initialises
 * file-static arrays and runs the explicit step in a loop, writing .prof and
 * .vtk files.  Each entry point below replaces one piece of that program; the
 * comment on each names the reference lines it stands in for.  All arrays are
 * plain host pointers owned by the caller; device memory is owned by the handle.
 * Every call returns 0 on success or a negative DMPH_E* code, with a message
 * available from dmph_last_error(); nothing here calls exit().
 */
#ifndef DEM_MPH_H
#define DEM_MPH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DMPH_NTYPE 6

enum {
    DMPH_OK = 0,
    DMPH_E_ARG = -1,       /* invalid argument */
    DMPH_E_IO = -2,        /* file could not be opened/read/written */
    DMPH_E_FORMAT = -3,    /* malformed .grid/.data content */
    DMPH_E_DOMAIN = -4,    /* cell grid unusable (fewer than 3 cells on an axis) */
    DMPH_E_DEVICE = -5,    /* HIP runtime error or no device */
    DMPH_E_STATE = -6,     /* call not valid in the handle's current state */
    DMPH_E_ALLOC = -7      /* host memory exhausted */
};

/* Parameters: the keys of the reference's .data file (source/DEM-MPH.cpp:634-660). */
typedef struct dmph_params {
    double dt, dem_dt, output_interval, vtk_output_interval, end_time;
    double radius_ratio_a, radius_ratio_p, radius_ratio_v;
    double density[DMPH_NTYPE], bulk_modulus[DMPH_NTYPE], bulk_viscosity[DMPH_NTYPE];
    double shear_viscosity[DMPH_NTYPE], surface_tension[DMPH_NTYPE];
    double interaction_ratio[DMPH_NTYPE][DMPH_NTYPE];
    double gravity[3];
    double young_modulus[DMPH_NTYPE], poisson_ratio[DMPH_NTYPE];
    double restitution[DMPH_NTYPE], friction[DMPH_NTYPE];
    double wall_center[DMPH_NTYPE][3], wall_velocity[DMPH_NTYPE][3], wall_omega[DMPH_NTYPE][3];
} dmph_params;

/* A particle set in .grid terms (source/DEM-MPH.cpp:681-851): line 1 Time,
 * line 2 N d xmin xmax ymin ymax zmin zmax, then `type x y z vx vy vz` per
 * particle.  `body` (rigid-body id, -1 = none) and `density` (<= 0 = use the
 * type's density) come from two optional trailing columns and may be NULL. */
typedef struct dmph_grid {
    double time;
    int n;
    double spacing;
    double domain_min[3], domain_max[3];
    int *type;        /* [n] */
    double *pos;      /* [n][3] */
    double *vel;      /* [n][3] */
    int *body;        /* [n] or NULL */
    double *density;  /* [n] or NULL */
} dmph_grid;

enum { DMPH_PREC_FP64 = 0, DMPH_PREC_FP32 = 1 };
enum { DMPH_SORT_CANONICAL = 0, DMPH_SORT_REFERENCE = 1 };
enum { DMPH_REBUILD_SKIN = 0, DMPH_REBUILD_EVERY_STEP = 1 };
enum { DMPH_DEBRIS_PARTICLES = 0, DMPH_DEBRIS_RIGID = 1 };

typedef struct dmph_options {
    int dim;            /* 2 or 3; 0 = infer (2 when the z extent holds < 3 cells) */
    int precision;      /* DMPH_PREC_*: FP64 everywhere, or FP32 pairs on cell-relative positions */
    int sort_mode;      /* DMPH_SORT_*: within-cell order (reference = its bitonic network) */
    int rebuild;        /* DMPH_REBUILD_*: reference 0.05 d skin test, or rebin every step */
    int debris;         /* DMPH_DEBRIS_*: loose DEM grains (reference) or rigid bodies */
    int device;         /* HIP device ordinal */
    int compensated;    /* FP32 only: pair sums carried in FP64 accumulators (1, default) or FP32 (0) */
    int graph;          /* 1 (default): each step replays one captured hipGraph; 0: launches one by one */
    int lanes;          /* FP32 only: lanes per particle in the sweeps, 2, 4, 8 or 16 (0: 8; see dmph_autotune) */
    int rigid_precision; /* FP32 only: DMPH_PREC_FP64 (default: FP64 rigid-body sums and integration, the
                            mixed-precision solver) or DMPH_PREC_FP32 (FP32 there too: the plain FP32 solver) */
    int block;          /* FP32 only: threads per workgroup of the pair sweeps, 64, 128 or 256 (0: 256; see dmph_autotune) */
    double cell_scale;  /* FP32 only: cell width = cell_scale x (MaxRadius + 0.1 d), cell_scale in [1, 2]
                           (0 or 1: the reference's cell grid, source/DEM-MPH.cpp:1211-1239); the 3^D stencil
                           still covers the cutoff, rows hold the same pairs, slot order and list-build
                           cost change.  Negative: tuned at creation -- 1, 1.25 and 1.5 are each tried on a
                           handle of their own (autotuned, timed between rebuilds and on rebuild steps, scored
                           as step time + rebuild cost over 20 steps) and the handle is created with the
                           best; dmph_get_info reports it (FP64: the reference's grid) */
    int wall_lanes;     /* FP32 only: lanes per wall row in the fused pass A, 1, 2 or 4 (0: 2; see dmph_autotune) */
} dmph_options;

void dmph_default_options(dmph_options *opt);

/* readDataFile (source/DEM-MPH.cpp:621-679): unknown lines are ignored as the
 * reference does; returns the number of recognised lines (>= 0) or DMPH_E_IO. */
int dmph_params_read(const char *path, dmph_params *out);

/* readGridFile (source/DEM-MPH.cpp:681-851).  Allocates the arrays of *out;
 * release them with dmph_grid_free. */
int dmph_grid_read(const char *path, dmph_grid *out);
void dmph_grid_free(dmph_grid *g);

/* The scene generator (generator/generator.cpp): readfile (:129-183) and the
 * block readers (:187-658) parse a .boid file, genparticle (:660-944) fills
 * *out (time 0, spacing = ParticleDistance or -1, domain = Lower/UpperDomain,
 * body and density NULL) kind by kind -- Cuboid, Cuboid2, Cyboid, Cyboid2,
 * Recboid, Recboid2 -- with the generator's arithmetic.  Release with
 * dmph_grid_free.  DMPH_E_FORMAT on the generator's parse errors and on end of
 * file inside a block (where the generator hangs), DMPH_E_IO when the file
 * cannot be opened; either way *out holds the shapes read before the error,
 * as the generator would write them.  dmph_boid_error() describes the last
 * failure of the calling thread. */
int dmph_boid_generate(const char *path, dmph_grid *out);
const char *dmph_boid_error(void);

/* writefile (generator/generator.cpp:948-974): the generator's .grid layout
 * (Time "%lf", header "%d %e  %e %e %e  %e %e %e", then
 * "%d    %e %e %e  %e %e %e " per particle), which dmph_grid_read and the
 * reference's readGridFile read back. */
int dmph_grid_write(const char *path, const dmph_grid *g);

typedef struct dmph_handle dmph_handle;

/* The initialisation of main (source/DEM-MPH.cpp:378-425): kernel weights and
 * reference densities, per-type coefficients, wall rotations, the cell grid,
 * device upload, first binning and the initial DensityA / GravityCenter /
 * DensityP passes. */
int dmph_create(const dmph_params *p, const dmph_options *opt, const dmph_grid *g, dmph_handle **out);

/* A batch of independent scenes advanced together (no reference counterpart:
 * the reference runs one scene per process).  All grids must have the same
 * spacing and domain; particle counts, fluid columns and bodies may differ.
 * Scene s occupies the original-order indices [sum of n of scenes < s, + n_s)
 * of every field; rigid-body ids are per scene (body b of scene s is body
 * (bodies of scenes < s) + b in dmph_get_bodies).  The handle steps every scene
 * with one set of launches; a scene's particles never interact with another's.
 * With canonical order and rebinning every step, each scene evolves bit for bit
 * as it would in its own handle. */
int dmph_create_batch(const dmph_params *p, const dmph_options *opt, const dmph_grid *grids, int nscenes,
                      dmph_handle **out);

/* `nsteps` passes of the main loop body (source/DEM-MPH.cpp:450-579): fluid
 * convection, walls, periodic wrap, rebuild on the 0.05 d skin test, the MPH
 * passes, drag (twice), buoyancy, gravity, acceleration and the DEM sub-steps.
 * Asynchronous on the handle's stream; no host synchronisation per step.  With
 * options.graph the step is captured once as a hipGraph and replayed (while
 * device timing is enabled the launches are issued one by one). */
int dmph_step(dmph_handle *h, int nsteps);

/* Wait for queued work; returns a pending device error if any. */
int dmph_sync(dmph_handle *h);

/* Simulated time (the reference's FP64 `Time`, source/DEM-MPH.cpp:578). */
int dmph_time(dmph_handle *h, double *time);

/* Per-particle fields, copied to host memory (float fields as double). */
enum {
    DMPH_FIELD_POSITION = 0,       /* double[n][3] */
    DMPH_FIELD_VELOCITY = 1,       /* double[n][3] */
    DMPH_FIELD_FORCE = 2,          /* double[n][3]; not maintained for wall particles (dead in the reference) */
    DMPH_FIELD_TYPE = 3,           /* int[n] */
    DMPH_FIELD_ID = 4,             /* int[n]: original (file) index */
    DMPH_FIELD_DENSITY_A = 5,
    DMPH_FIELD_GRAVITY_CENTER = 6, /* double[n][3] */
    DMPH_FIELD_PRESSURE_A = 7,
    DMPH_FIELD_VOL_STRAIN_P = 8,
    DMPH_FIELD_DIVERGENCE_P = 9,
    DMPH_FIELD_PRESSURE_P = 10,
    DMPH_FIELD_MASS = 11,
    DMPH_FIELD_ALPHA = 12,
    DMPH_FIELD_DRAG_C = 13,
    DMPH_FIELD_KN = 14,
    DMPH_FIELD_ETAN = 15,
    DMPH_FIELD_KT = 16,
    DMPH_FIELD_ETAT = 17,
    DMPH_FIELD_NEIGHBOR_COUNT = 18,       /* int[n], counted at the last rebuild */
    DMPH_FIELD_SOLID_NEIGHBOR_COUNT = 19, /* int[n] */
    DMPH_FIELD_CELL_KEY = 20,             /* int[n]: cell id + block * cells, at the last rebuild */
    DMPH_FIELD_VIRIAL_PRESSURE = 21,      /* valid after dmph_virial */
    DMPH_FIELD_BODY = 22,                 /* int[n] */
    DMPH_FIELD_KAPPA = 23,
    DMPH_FIELD_COUNT = 24
};
enum { DMPH_ORDER_ORIGINAL = 0, DMPH_ORDER_SLOT = 1 };
int dmph_get_field(dmph_handle *h, int field, int order, void *dst);

/* Overwrite DMPH_FIELD_POSITION or DMPH_FIELD_VELOCITY (double[n][3], in the
 * given order) from host memory, e.g. to restart from a saved state (the
 * reference re-reads a .prof as its .grid, source/DEM-MPH.cpp:681-851).
 * Positions are wrapped into the periodic domain and the next step rebins
 * unconditionally.  Members of rigid bodies are re-posed from their body at the
 * next DEM sub-step.  Synchronous. */
int dmph_set_field(dmph_handle *h, int field, int order, const void *src);

/* Positions (absolute, wrapped into the domain) and velocities as FP32 [n][3]
 * in original order, written to DEVICE buffers on the handle's device (either
 * may be NULL): a copy that stays on the GPU, for coupling with other device
 * code (the surrogate records trajectories this way).  No reference
 * counterpart.  Synchronous. */
int dmph_export_state(dmph_handle *h, float *pos, float *vel);

/* Rigid-body state (no reference counterpart): per body X[3], V[3], q[4] (w,x,y,z), L[3], mass. */
int dmph_get_bodies(dmph_handle *h, double *x, double *v, double *q, double *l, double *mass);

typedef struct dmph_info {
    int n, dim, cells, cell_count[3], nbodies, substeps, precision;
    int fluid_begin, fluid_end, solid_begin, solid_end, wall_begin, wall_end;
    double cell_width[3], n0a, n0p, max_radius;
    int64_t rebuilds;   /* binning passes so far (device counter) */
    int nscenes;        /* scenes in the handle (dmph_create_batch), 1 otherwise; cells is per scene */
    int lanes[2];       /* FP32 sweeps 1 and 2: lanes per particle (dmph_autotune); 1 in FP64 */
    int block[2];       /* FP32 sweeps 1 and 2: threads per workgroup (dmph_autotune); 256 in FP64 */
    double cell_scale;  /* cell width over MaxRadius + 0.1 d (1: the reference's grid) */
    int wall_lanes;     /* FP32 pass A: lanes per wall row (dmph_autotune); 1 in FP64 */
} dmph_info;
int dmph_get_info(dmph_handle *h, dmph_info *info);

/* The optional NaN/Inf check of the boundary: counts, on the device, the
 * values of a real field (DMPH_FIELD_* other than the integer ones, every
 * component) that are NaN or infinite, into *count (0: all finite; integer
 * fields count 0).  The reference lets NaN propagate silently (the drag at zero
 * relative velocity, source/DEM-MPH.cpp:2775-2797, divides by Re = 0); this
 * lets a caller stop a run that produced one.  Synchronous. */
int dmph_count_nonfinite(dmph_handle *h, int field, int64_t *count);

/* calculateVirialStressAtParticle (source/DEM-MPH.cpp:2904-3149), the
 * output-only pass run before each VTK write. */
int dmph_virial(dmph_handle *h);

/* Neighbour pairs within MaxRadius + 0.1 d at the last rebuild, in slot order,
 * as the reference's CSR (source/DEM-MPH.cpp:1680-1840).  Call with idx = NULL
 * to get *nnz; ptr has n + 1 entries. */
int dmph_pairs(dmph_handle *h, int64_t *nnz, int64_t *ptr, int *idx);

/* writeProfFile (source/DEM-MPH.cpp:853-877) and writeVtkFile (:879-1005),
 * byte-compatible, particles in slot (sorted) order as the reference writes them. */
int dmph_write_prof(dmph_handle *h, const char *path);
int dmph_write_vtk(dmph_handle *h, const char *path);

/* Device timing of kernel groups with HIP events on the handle's stream (no
 * reference counterpart; the reference's clock() buckets, source/DEM-MPH.cpp:345,
 * measure host CPU time).  Groups: DMPH_TIMING_*. */
enum {
    DMPH_TIMING_PREDICT = 0,   /* convection, walls, wrap, skin test, rebin decision, time advance */
    DMPH_TIMING_REBIN = 1,     /* keys, scan, sort, reorder (early exit on steps without rebinning) */
    DMPH_TIMING_SWEEP1 = 2,    /* densities, divergence, pressures */
    DMPH_TIMING_SWEEP2 = 3,    /* drag, pressure, surface tension, viscosity, buoyancy, gravity forces */
    DMPH_TIMING_ACCEL = 4,     /* fluid velocity update */
    DMPH_TIMING_DEM = 5,       /* DEM coefficients and contact forces (per sub-step) */
    DMPH_TIMING_MOVE = 6,      /* solid update, rigid-body reduction/integration/re-posing (per sub-step) */
    DMPH_TIMING_END = 7,       /* unused: the time accumulator advances in the predict kernel */
    DMPH_TIMING_GROUPS = 8
};
int dmph_set_timing(dmph_handle *h, int enable);
int dmph_get_timing(dmph_handle *h, int group, double *total_ms, int64_t *launches);

/* Autotuner (no reference counterpart): times the FP32 sweeps with 2, 4, 8
 * and 16 lanes per particle and workgroups of 64, 128 and 256 threads on the
 * handle's current scene(s) -- `reps` launches per candidate, re-running the
 * sweeps on the current state (they recompute the same values; forces are
 * restored), so the simulation does not advance -- and keeps the fastest
 * (lanes, workgroup) for each sweep; the step graph is captured again.  The
 * choice changes FP32 summation order only.  FP64 handles keep one thread per
 * particle.  Writes the chosen lanes to lanes[2] if not NULL; dmph_get_info
 * reports lanes and workgroup sizes.  The cell size needs a new handle: it is
 * tuned at creation when dmph_options.cell_scale is negative. */
int dmph_autotune(dmph_handle *h, int reps, int *lanes);

const char *dmph_last_error(const dmph_handle *h);
void dmph_destroy(dmph_handle *h);

#ifdef __cplusplus
}
#endif
#endif
