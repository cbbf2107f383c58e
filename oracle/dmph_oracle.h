/* FP64 CPU oracle for the DEM-MPH explicit step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a restatement of the reference solver's
 * algorithm (source/DEM-MPH.cpp) used to check the HIP product path.  Only the
 * test suite, the smoke check and the CPU-baseline leg of the benchmark may
 * load it; the product library never links or calls it.
 */
#ifndef DMPH_ORACLE_H
#define DMPH_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_NTYPE 6

/* The .data keys of source/DEM-MPH.cpp:634-660. */
typedef struct orc_params {
    double dt, dem_dt, output_interval, vtk_output_interval, end_time;
    double ratio_a, ratio_p, ratio_v;
    double density[ORC_NTYPE], bulk_modulus[ORC_NTYPE], bulk_viscosity[ORC_NTYPE];
    double shear_viscosity[ORC_NTYPE], surface_tension[ORC_NTYPE];
    double interaction[ORC_NTYPE][ORC_NTYPE];
    double gravity[3];
    double young[ORC_NTYPE], poisson[ORC_NTYPE], restitution[ORC_NTYPE], friction[ORC_NTYPE];
    double wall_center[ORC_NTYPE][3], wall_velocity[ORC_NTYPE][3], wall_omega[ORC_NTYPE][3];
} orc_params;

/* Within-cell order after binning:
 *   ORC_SORT_REFERENCE  the reference's bitonic network, equal keys included
 *                       (source/DEM-MPH.cpp:1317-1338);
 *   ORC_SORT_CANONICAL  ascending original particle id inside each cell. */
enum { ORC_SORT_REFERENCE = 0, ORC_SORT_CANONICAL = 1 };
/* Rebuild schedule: the reference's 0.05 d skin test (:1264-1286) or every step. */
enum { ORC_REBUILD_SKIN = 0, ORC_REBUILD_EVERY_STEP = 1 };

typedef struct orc_opts {
    int dim;        /* 2 or 3 */
    int sort_mode;  /* ORC_SORT_* */
    int rebuild;    /* ORC_REBUILD_* */
    int rigid;      /* 1: particles with body id >= 0 move as rigid bodies */
    int threads;    /* OpenMP threads, 0 = runtime default */
} orc_opts;

typedef struct orc orc;

/* Parse a .data file like readDataFile (source/DEM-MPH.cpp:621-679).
 * Returns the number of recognised lines, or -1 if the file cannot be opened. */
int orc_parse_data(const char *path, orc_params *p);

/* Build a state from particle arrays in file order and run the reference's
 * pre-loop sequence (source/DEM-MPH.cpp:378-425): weights, coefficients, walls,
 * cell grid, binning, neighbour list, DensityA, GravityCenter, DensityP.
 * `body` and `density` may be NULL; density <= 0 means "use Density[type]". */
orc *orc_create(const orc_params *p, const orc_opts *o, double time, int n, double spacing,
                const double *domain_min, const double *domain_max, const int *type,
                const double *pos, const double *vel, const int *body, const double *density);

/* Advance `nsteps` explicit steps (the loop body of source/DEM-MPH.cpp:450-579). */
void orc_step(orc *s, int nsteps);

/* Output-only virial stress pass (source/DEM-MPH.cpp:2904-3149). */
void orc_virial(orc *s);

/* Field access by name.  orc_field_info returns the element count and sets
 * *dtype to 0 (float64), 1 (int32) or 2 (int64); -1 for an unknown name. */
int64_t orc_field_info(const orc *s, const char *name, int *dtype);
int orc_get(const orc *s, const char *name, void *dst);

void orc_destroy(orc *s);

#ifdef __cplusplus
}
#endif
#endif
