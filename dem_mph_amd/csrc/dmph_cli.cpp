// DEM-MPH command-line driver on the MI355X library.
//
//   DEM-MPH [data] [grid] [prof-format] [vtk-format] [log]
//
// The same positional arguments, defaults, output cadence, file naming
// (sprintf of the step number) and log lines as the reference's main
// (source/DEM-MPH.cpp:350-617).  Arguments after the fifth are ignored, as
// there.  Engine options come from the environment:
//   DMPH_PRECISION=fp64|fp32   (default fp64)
//   DMPH_SORT=reference|canonical (default reference: the reference's particle order)
//   DMPH_REBUILD=skin|every    (default skin)
//   DMPH_DEBRIS=particles|rigid (default particles)
//   DMPH_DIM=2|3               (default: inferred from the z extent)
//   DMPH_DEVICE=<ordinal>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <string>

#include "dem_mph.h"

static FILE *g_log = nullptr;

static void logf(const char *fmt, ...) {
    va_list a, b;
    va_start(a, fmt);
    va_copy(b, a);
    if (g_log) vfprintf(g_log, fmt, a);
    vfprintf(stderr, fmt, b);
    va_end(a);
    va_end(b);
}

static const char *env(const char *k) {
    const char *v = getenv(k);
    return v ? v : "";
}

static void die(const char *what, const char *path, const dmph_handle *h) {
    fprintf(stderr, "%s %s: %s\n", what, path ? path : "", dmph_last_error(h));
    exit(1);
}

int main(int argc, char **argv) {
    char logname[1024] = "sample.log", dataname[1024] = "sample.data", gridname[1024] = "sample.grid";
    char proffmt[1024] = "sample%03d.prof", vtkfmt[1024] = "sample%03d.vtk";
    if (argc > 1) snprintf(dataname, sizeof dataname, "%s", argv[1]);
    if (argc > 2) snprintf(gridname, sizeof gridname, "%s", argv[2]);
    if (argc > 3) snprintf(proffmt, sizeof proffmt, "%s", argv[3]);
    if (argc > 4) snprintf(vtkfmt, sizeof vtkfmt, "%s", argv[4]);
    if (argc > 5) snprintf(logname, sizeof logname, "%s", argv[5]);

    g_log = fopen(logname, "w");
    if (!g_log) fprintf(stderr, "error in open %s\n", logname);
    {
        time_t t = time(nullptr);
        logf("start reading files at %s\n", ctime(&t));
    }
    dmph_params p;
    memset(&p, 0, sizeof p);
    p.dt = 1.0e100;
    p.dem_dt = 1.0e100;
    if (dmph_params_read(dataname, &p) < 0) {
        fprintf(stderr, "File Open Error\nfilename:%s  mode:r\n", dataname);
        return 1;
    }
    dmph_grid grid;
    const int grc = dmph_grid_read(gridname, &grid);
    if (grc == DMPH_E_IO) {
        fprintf(stderr, "File Open Error\nfilename:%s  mode:r\n", gridname);
        return 1;
    }
    if (grc != DMPH_OK) {
        fprintf(stderr, "malformed grid file %s\n", gridname);
        return 1;
    }
    {
        time_t t = time(nullptr);
        logf("start initialization at %s\n", ctime(&t));
    }
    dmph_options opt;
    dmph_default_options(&opt);
    opt.sort_mode = DMPH_SORT_REFERENCE;
    if (!strcmp(env("DMPH_PRECISION"), "fp32")) opt.precision = DMPH_PREC_FP32;
    if (!strcmp(env("DMPH_SORT"), "canonical")) opt.sort_mode = DMPH_SORT_CANONICAL;
    if (!strcmp(env("DMPH_REBUILD"), "every")) opt.rebuild = DMPH_REBUILD_EVERY_STEP;
    if (!strcmp(env("DMPH_DEBRIS"), "rigid")) opt.debris = DMPH_DEBRIS_RIGID;
    if (*env("DMPH_DIM")) opt.dim = atoi(env("DMPH_DIM"));
    if (*env("DMPH_DEVICE")) opt.device = atoi(env("DMPH_DEVICE"));

    dmph_handle *h = nullptr;
    if (dmph_create(&p, &opt, &grid, &h) != DMPH_OK) die("initialisation failed for", gridname, nullptr);
    dmph_info info;
    dmph_get_info(h, &info);
    logf("N0a = %e\n", info.n0a);
    logf("N0p = %e\n", info.n0p);
    logf("line:%d CellCount[DIM]: %d %d %d\n", __LINE__, info.cell_count[0], info.cell_count[1], info.cell_count[2]);
    logf("line:%d TotalCellCount: %d\n", __LINE__, info.cells);
    logf("line:%d CellWidth[DIM]: %e %e %e\n", __LINE__, info.cell_width[0], info.cell_width[1], info.cell_width[2]);
    logf("Using GPU %d\n", opt.device);
    {
        int64_t nnz = 0;
        dmph_pairs(h, &nnz, nullptr, nullptr);
        logf("line:%d, NeighborIndCount = %lld\n", __LINE__, (long long)nnz);
    }
    if (dmph_write_vtk(h, "output.vtk") != DMPH_OK) die("cannot write", "output.vtk", h);
    {
        time_t t = time(nullptr);
        logf("start main roop at %s\n", ctime(&t));
    }

    double Time = grid.time, OutputNext = 0.0, VtkOutputNext = 0.0;
    const double Dt = p.dt;
    int iStep = (int)(Time / Dt);
    const clock_t c0 = clock();
    timespec w0, w1;
    clock_gettime(CLOCK_MONOTONIC, &w0);
    while (Time < p.end_time + 1.0e-5 * Dt) {
        if (Time + 1.0e-5 * Dt >= OutputNext) {
            char fn[1100];
            snprintf(fn, sizeof fn, proffmt, iStep);
            if (dmph_write_prof(h, fn) != DMPH_OK) die("cannot write", fn, h);
            logf("@ Prof Output Time : %e\n", Time);
            OutputNext += p.output_interval;
        }
        if (dmph_step(h, 1) != DMPH_OK) die("step failed", nullptr, h);
        if (Time + 1.0e-5 * Dt >= VtkOutputNext) {
            char fn[1100];
            snprintf(fn, sizeof fn, vtkfmt, iStep);
            if (dmph_virial(h) != DMPH_OK || dmph_write_vtk(h, fn) != DMPH_OK) die("cannot write", fn, h);
            logf("@ Vtk Output Time : %e\n", Time);
            VtkOutputNext += p.vtk_output_interval;
        }
        Time += Dt;
        iStep++;
    }
    if (dmph_sync(h) != DMPH_OK) die("device error", nullptr, h);
    clock_gettime(CLOCK_MONOTONIC, &w1);
    {
        time_t t = time(nullptr);
        logf("end main roop at %s\n", ctime(&t));
        const double wall = (w1.tv_sec - w0.tv_sec) + 1e-9 * (w1.tv_nsec - w0.tv_nsec);
        logf("total:                   %lf [CPU sec]\n", (double)(clock() - c0) / CLOCKS_PER_SEC);
        logf("wall time of the step loop: %lf [sec]\n", wall);
    }
    dmph_destroy(h);
    dmph_grid_free(&grid);
    if (g_log) fclose(g_log);
    return 0;
}
