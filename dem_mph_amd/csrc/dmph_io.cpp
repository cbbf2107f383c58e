// File formats of the reference solver: .data parameters, .grid particles,
// .prof snapshots and legacy-VTK output (source/DEM-MPH.cpp:621-1005).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "dem_mph.h"
#include "dmph_internal.h"

namespace {

bool scan6(const char *buf, const char *key, double *v) {
    char fmt[160];
    snprintf(fmt, sizeof fmt, " %s %%lf %%lf %%lf %%lf %%lf %%lf", key);
    return sscanf(buf, fmt, &v[0], &v[1], &v[2], &v[3], &v[4], &v[5]) == 6;
}

}  // namespace

extern "C" void dmph_params_defaults(dmph_params *p) {
    // the reference's static initialisers (source/DEM-MPH.cpp:61-75, :155)
    memset(p, 0, sizeof *p);
    p->dt = 1.0e100;
    p->dem_dt = 1.0e100;
}

// readDataFile, source/DEM-MPH.cpp:621-679: a chain of sscanf key matches;
// lines starting with '#' are comments and anything unmatched is reported.
extern "C" int dmph_params_read(const char *path, dmph_params *p) {
    FILE *fp = fopen(path, "r");
    if (!fp) return DMPH_E_IO;
    char buf[1024];
    int recognised = 0;
    dmph::g_invalid_data_lines.clear();
    while (fgets(buf, sizeof buf, fp)) {
        if (buf[0] == '#') continue;
        bool ok = false;
        if (sscanf(buf, " Dt %lf", &p->dt) == 1) ok = true;
        else if (sscanf(buf, " DEMDt %lf", &p->dem_dt) == 1) ok = true;
        else if (sscanf(buf, " OutputInterval %lf", &p->output_interval) == 1) ok = true;
        else if (sscanf(buf, " VtkOutputInterval %lf", &p->vtk_output_interval) == 1) ok = true;
        else if (sscanf(buf, " EndTime %lf", &p->end_time) == 1) ok = true;
        else if (sscanf(buf, " RadiusRatioA %lf", &p->radius_ratio_a) == 1) ok = true;
        else if (sscanf(buf, " RadiusRatioP %lf", &p->radius_ratio_p) == 1) ok = true;
        else if (sscanf(buf, " RadiusRatioV %lf", &p->radius_ratio_v) == 1) ok = true;
        else if (scan6(buf, "Density", p->density)) ok = true;
        else if (scan6(buf, "BulkModulus", p->bulk_modulus)) ok = true;
        else if (scan6(buf, "BulkViscosity", p->bulk_viscosity)) ok = true;
        else if (scan6(buf, "ShearViscosity", p->shear_viscosity)) ok = true;
        else if (scan6(buf, "SurfaceTension", p->surface_tension)) ok = true;
        for (int t = 0; t < DMPH_NTYPE && !ok; t++) {
            char k[40];
            snprintf(k, sizeof k, "InteractionRatio(Type%d)", t);
            ok = scan6(buf, k, p->interaction_ratio[t]);
        }
        if (!ok && sscanf(buf, " Gravity %lf %lf %lf", &p->gravity[0], &p->gravity[1], &p->gravity[2]) == 3) ok = true;
        if (!ok) ok = scan6(buf, "YoungModulus", p->young_modulus);
        if (!ok) ok = scan6(buf, "PoissonRatio", p->poisson_ratio);
        if (!ok) ok = scan6(buf, "RestitutionCoefficient", p->restitution);
        if (!ok) ok = scan6(buf, "ViscousFrictionCoefficient", p->friction);
        for (int w = 4; w < 6 && !ok; w++) {
            char fmt[200];
            snprintf(fmt, sizeof fmt, " Wall%d Center %%lf %%lf %%lf Velocity %%lf %%lf %%lf Omega %%lf %%lf %%lf", w);
            ok = sscanf(buf, fmt, &p->wall_center[w][0], &p->wall_center[w][1], &p->wall_center[w][2],
                        &p->wall_velocity[w][0], &p->wall_velocity[w][1], &p->wall_velocity[w][2],
                        &p->wall_omega[w][0], &p->wall_omega[w][1], &p->wall_omega[w][2]) == 9;
        }
        if (ok) recognised++;
        else dmph::g_invalid_data_lines.push_back(buf);
    }
    fclose(fp);
    return recognised;
}

// readGridFile, source/DEM-MPH.cpp:681-851.  Two optional trailing columns per
// particle line (body id, density) are read when present; the reference's
// seven-field sscanf ignores them.  A short file is an error here (the
// reference leaves the missing particles uninitialised).
extern "C" int dmph_grid_read(const char *path, dmph_grid *g) {
    memset(g, 0, sizeof *g);
    FILE *fp = fopen(path, "r");
    if (!fp) return DMPH_E_IO;
    char buf[1024];
    if (!fgets(buf, sizeof buf, fp) || sscanf(buf, "%lf", &g->time) != 1) { fclose(fp); return DMPH_E_FORMAT; }
    if (!fgets(buf, sizeof buf, fp) ||
        sscanf(buf, "%d  %lf  %lf %lf %lf  %lf %lf %lf", &g->n, &g->spacing, &g->domain_min[0], &g->domain_max[0],
               &g->domain_min[1], &g->domain_max[1], &g->domain_min[2], &g->domain_max[2]) != 8 ||
        g->n < 0) {
        fclose(fp);
        return DMPH_E_FORMAT;
    }
    const int n = g->n;
    g->type = (int *)calloc((size_t)n + 1, sizeof(int));
    g->pos = (double *)calloc(3 * (size_t)n + 1, sizeof(double));
    g->vel = (double *)calloc(3 * (size_t)n + 1, sizeof(double));
    std::vector<int> body((size_t)n, -1);
    std::vector<double> dens((size_t)n, -1.0);
    bool extra = false;
    for (int i = 0; i < n; i++) {
        if (!fgets(buf, sizeof buf, fp)) { fclose(fp); dmph_grid_free(g); return DMPH_E_FORMAT; }
        int b = -1;
        double rho = -1.0;
        const int got = sscanf(buf, "%d  %lf %lf %lf  %lf %lf %lf %d %lf", &g->type[i], &g->pos[3 * i], &g->pos[3 * i + 1],
                               &g->pos[3 * i + 2], &g->vel[3 * i], &g->vel[3 * i + 1], &g->vel[3 * i + 2], &b, &rho);
        if (got < 7) { fclose(fp); dmph_grid_free(g); return DMPH_E_FORMAT; }
        if (got >= 8) { extra = true; body[(size_t)i] = b; }
        if (got >= 9) dens[(size_t)i] = rho;
    }
    fclose(fp);
    if (extra) {
        g->body = (int *)malloc(sizeof(int) * ((size_t)n + 1));
        g->density = (double *)malloc(sizeof(double) * ((size_t)n + 1));
        memcpy(g->body, body.data(), sizeof(int) * (size_t)n);
        memcpy(g->density, dens.data(), sizeof(double) * (size_t)n);
    }
    return DMPH_OK;
}

extern "C" void dmph_grid_free(dmph_grid *g) {
    if (!g) return;
    free(g->type); free(g->pos); free(g->vel); free(g->body); free(g->density);
    g->type = nullptr; g->pos = g->vel = nullptr; g->body = nullptr; g->density = nullptr;
}

namespace dmph {

std::vector<std::string> g_invalid_data_lines;

// writeProfFile, source/DEM-MPH.cpp:853-877 (slot order)
int write_prof(const char *path, double time, int n, double d, const double *dmin, const double *dmax, const int *type,
               const double *pos, const double *vel) {
    FILE *fp = fopen(path, "w");
    if (!fp) return DMPH_E_IO;
    fprintf(fp, "%e\n", time);
    fprintf(fp, "%d %e %e %e %e %e %e %e\n", n, d, dmin[0], dmax[0], dmin[1], dmax[1], dmin[2], dmax[2]);
    for (int i = 0; i < n; ++i)
        fprintf(fp, "%d %e %e %e  %e %e %e\n", type[i], pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], vel[3 * i],
                vel[3 * i + 1], vel[3 * i + 2]);
    fflush(fp);
    fclose(fp);
    return DMPH_OK;
}

// writeVtkFile, source/DEM-MPH.cpp:879-1005: the same field list, order and
// number formatting (positions, velocities, VolStrainP and the virial
// pressure pass through float).
int write_vtk(const char *path, const VtkFields &f) {
    FILE *fp = fopen(path, "w");
    if (!fp) return DMPH_E_IO;
    const int n = f.n;
    fprintf(fp, "# vtk DataFile Version 2.0\n");
    fprintf(fp, "Unstructured Grid Example\n");
    fprintf(fp, "ASCII\n");
    fprintf(fp, "DATASET UNSTRUCTURED_GRID\n");
    fprintf(fp, "POINTS %d float\n", n);
    for (int i = 0; i < n; ++i)
        fprintf(fp, "%e %e %e\n", (float)f.pos[3 * i], (float)f.pos[3 * i + 1], (float)f.pos[3 * i + 2]);
    fprintf(fp, "CELLS %d %d\n", n, 2 * n);
    for (int i = 0; i < n; ++i) fprintf(fp, "1 %d ", i);
    fprintf(fp, "\n");
    fprintf(fp, "CELL_TYPES %d\n", n);
    for (int i = 0; i < n; ++i) fprintf(fp, "1 ");
    fprintf(fp, "\n");
    fprintf(fp, "\n");
    fprintf(fp, "POINT_DATA %d\n", n);
    fprintf(fp, "SCALARS label float 1\n");
    fprintf(fp, "LOOKUP_TABLE default\n");
    for (int i = 0; i < n; ++i) fprintf(fp, "%d\n", f.type[i]);
    fprintf(fp, "\n");
    const char *names[4] = {"NormalSpringConstant", "NormalViscousDampingConstant", "TangentialSpringConstant",
                            "TangentialViscousDampingConstant"};
    const double *vals[4] = {f.kn, f.etan, f.kt, f.etat};
    for (int k = 0; k < 4; k++) {
        fprintf(fp, "SCALARS %s float 1\n", names[k]);
        fprintf(fp, "LOOKUP_TABLE default\n");
        for (int i = 0; i < n; ++i) fprintf(fp, "%e\n", vals[k][i]);
    }
    fprintf(fp, "SCALARS neighbor float 1\n");
    fprintf(fp, "LOOKUP_TABLE default\n");
    for (int i = 0; i < n; ++i) fprintf(fp, "%d\n", f.ncount[i]);
    fprintf(fp, "SCALARS volumef float 1\n");
    fprintf(fp, "LOOKUP_TABLE default\n");
    for (int i = 0; i < n; ++i) fprintf(fp, "%e\n", f.alpha[i]);
    fprintf(fp, "SCALARS drag float 1\n");
    fprintf(fp, "LOOKUP_TABLE default\n");
    for (int i = 0; i < n; ++i) fprintf(fp, "%e\n", f.drag[i]);
    fprintf(fp, "\n");
    fprintf(fp, "VECTORS velocity float\n");
    for (int i = 0; i < n; ++i)
        fprintf(fp, "%e %e %e\n", (float)f.vel[3 * i], (float)f.vel[3 * i + 1], (float)f.vel[3 * i + 2]);
    fprintf(fp, "\n");
    fprintf(fp, "SCALARS VolStrainP float 1\n");
    fprintf(fp, "LOOKUP_TABLE default\n");
    for (int i = 0; i < n; ++i) fprintf(fp, "%e\n", (float)f.vs[i]);
    fprintf(fp, "\n");
    fprintf(fp, "SCALARS VirialPressureAtParticle float 1\n");
    fprintf(fp, "LOOKUP_TABLE default\n");
    for (int i = 0; i < n; ++i) fprintf(fp, "%e\n", (float)f.virial[i]);
    fprintf(fp, "\n");
    fflush(fp);
    fclose(fp);
    return DMPH_OK;
}

}  // namespace dmph
