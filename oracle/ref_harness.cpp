// Reference driver: TEST INFRASTRUCTURE ONLY.
//
// This file is not compiled on its own.  oracle/Makefile streams the
// reference solver's source (source/DEM-MPH.cpp, read where it lies, with its
// `main` renamed and, for the 3D build, the `#define TWO_DIMENSIONAL` line
// dropped) followed by this file into one translation unit, so the reference's
// file-static passes and arrays are reachable here.  The executable it makes
// under oracle/_ref/ runs the reference's own passes on a .data/.grid pair and
// dumps every per-particle array, which is how the oracle and the golden
// fixtures are pinned to the reference.
//
// It never calls the reference's main (which selects an NVIDIA device through
// OpenACC at source/DEM-MPH.cpp:383-384); instead run_step() calls the passes in
// the order of the main loop body, source/DEM-MPH.cpp:450-579.  Arrays that the
// reference leaves uninitialised after malloc (readGridFile, :704-799) are
// zeroed once after loading, which is the convention the oracle and the GPU
// path document for "never written" values.
#undef main

#include <cstdint>
#include <vector>
#include <string>

namespace dmph_ref_harness {

struct Dump {
    FILE* fp;
    void put(const char* name, int code, const void* data, int64_t count, int esize) {
        char nm[48] = {0};
        snprintf(nm, sizeof nm, "%s", name);
        fwrite(nm, 1, sizeof nm, fp);
        int32_t c = code;
        fwrite(&c, 4, 1, fp);
        fwrite(&count, 8, 1, fp);
        if (count) fwrite(data, (size_t)esize, (size_t)count, fp);
    }
    void f64(const char* n, const double* p, int64_t c) { put(n, 0, p, c, 8); }
    void i32(const char* n, const int* p, int64_t c) { put(n, 1, p, c, 4); }
    void i64(const char* n, const long* p, int64_t c) { put(n, 2, p, c, 8); }
};

static int g_dem_substeps = 1;

static void zero_uninitialised() {
    const size_t n = (size_t)ParticleCount;
    memset(Force, 0, n * sizeof(double[DIM]));
    memset(DensityA, 0, n * sizeof(double));
    memset(GravityCenter, 0, n * sizeof(double[DIM]));
    memset(PressureA, 0, n * sizeof(double));
    memset(VolStrainP, 0, n * sizeof(double));
    memset(DivergenceP, 0, n * sizeof(double));
    memset(PressureP, 0, n * sizeof(double));
    memset(VirialPressureAtParticle, 0, n * sizeof(double));
    memset(VirialStressAtParticle, 0, n * sizeof(double[DIM][DIM]));
    memset(Mass, 0, n * sizeof(double));
    memset(Mu, 0, n * sizeof(double));
    memset(Lambda, 0, n * sizeof(double));
    memset(Kappa, 0, n * sizeof(double));
    memset(InertialTensor, 0, n * sizeof(double[DIM]));
    memset(InverseInertialTensor, 0, n * sizeof(double[DIM]));
    memset(SolidAngularVelocity, 0, n * sizeof(double[DIM]));
    memset(Torque, 0, n * sizeof(double[DIM]));
    memset(AngularMomentum, 0, n * sizeof(double[DIM]));
    memset(NormalSpringConstant, 0, n * sizeof(double));
    memset(NormalViscousDampingConstant, 0, n * sizeof(double));
    memset(TangentialSpringConstant, 0, n * sizeof(double));
    memset(TangentialViscousDampingConstant, 0, n * sizeof(double));
    memset(alpha, 0, n * sizeof(double));
    memset(DragC, 0, n * sizeof(double));
    memset(NeighborCount, 0, n * sizeof(int));
    memset(FluidNeighborCount, 0, n * sizeof(int));
    memset(SolidNeighborCount, 0, n * sizeof(int));
    memset(NeighborCalculatedPosition, 0, n * sizeof(double[DIM]));
}

// Load + initialisation, following main at source/DEM-MPH.cpp:363-425.
static void load(char* data, char* grid, const char* log) {
    log_open(log);
    readDataFile(data);
    readGridFile(grid);
    zero_uninitialised();
    initializeWeight();
    initializeFluid();
    initializeWall();
    initializeDomain();
    calculateCellParticle();
    calculateNeighbor();
    calculateDensityA();
    calculateGravityCenter();
    calculateDensityP();
    g_dem_substeps = Dt / DEMDt;   // source/DEM-MPH.cpp:434 (truncating conversion)
}

// One pass of the main loop body, source/DEM-MPH.cpp:450-579, without file
// output and without the time advance.
static void run_step_body() {
    calculateConvection();
    calculateWall();
    calculatePeriodicBoundary();
    resetForce();
    if (neighborCalculation() == 1) {
        freeNeighbor();
        calculateCellParticle();
        calculateNeighbor();
    }
    calculateDensityA();
    calculateGravityCenter();
    calculateDensityP();
    calculateDivergenceP();
    calculatePhysicalCoefficients();
    calculateDragForce();
    calculatePressureP();
    calculatePressureA();
    calculateDiffuseInterface();
    calculateViscosityV();
    calculateDragForce();
    calculateBuoyancyForce();
    calculateGravity();
    calculateAcceleration();
    for (int i = 0; i < g_dem_substeps; i++) {
        calculateNormalDirectionCoefficient();
        calculateTangentialDirectionCoefficient();
        calculateInertialTensor();
        calculateTorque();
        calculateNormalDirectionForce();
        calculateTangentialDirectionForce();
        calculateDEMconvection();
    }
}

static void run_step() {
    run_step_body();
    Time += Dt;
}

// The reference's main loop with its file output, source/DEM-MPH.cpp:419-580:
// the initial output.vtk, then per step the .prof written before the step and
// the virial pass + .vtk after it, each when its next output time is reached,
// file names from sprintf(format, iStep) with iStep = (int)(Time / Dt) at the
// start, and the "@ ... Output Time" log lines.  Everything it calls is the
// reference's own code; only acc_set_device_num and the clock() buckets are
// left out.
static void run_files(char* proffmt, char* vtkfmt) {
    writeVtkFile((char*)"output.vtk");
    int iStep = (int)(Time / Dt);
    while (Time < EndTime + 1.0e-5 * Dt) {
        if (Time + 1.0e-5 * Dt >= OutputNext) {
            char filename[256];
            sprintf(filename, proffmt, iStep);
            writeProfFile(filename);
            log_printf("@ Prof Output Time : %e\n", Time);
            OutputNext += OutputInterval;
        }
        run_step_body();
        if (Time + 1.0e-5 * Dt >= VtkOutputNext) {
            calculateVirialStressAtParticle();
            char filename[256];
            sprintf(filename, vtkfmt, iStep);
            writeVtkFile(filename);
            log_printf("@ Vtk Output Time : %e\n", Time);
            VtkOutputNext += VtkOutputInterval;
        }
        Time += Dt;
        iStep++;
    }
}

static void dump(const char* path, int step, bool virial) {
    if (virial) calculateVirialStressAtParticle();
    Dump d{fopen(path, "wb")};
    const int64_t n = ParticleCount;
    fwrite("DMPHDUMP", 1, 8, d.fp);
    double sc[] = {Time, (double)step, N0a, N0p, Swa, Swg, Swp, Swv, R2g, CofK, MaxRadius,
                   RadiusA, RadiusG, RadiusP, RadiusV, ParticleVolume, ParticleSpacing, Dt, DEMDt,
                   (double)g_dem_substeps};
    d.f64("scalars", sc, (int64_t)(sizeof sc / sizeof sc[0]));
    int ic[] = {ParticleCount, TotalCellCount, CellCount[0], CellCount[1], CellCount[2],
                FluidParticleBegin, FluidParticleEnd, SolidParticleBegin, SolidParticleEnd,
                WallParticleBegin, WallParticleEnd, PowerParticleCount};
    d.i32("iscalars", ic, (int64_t)(sizeof ic / sizeof ic[0]));
    d.f64("CellWidth", CellWidth, 3);
    d.f64("DomainMin", DomainMin, 3);
    d.f64("DomainMax", DomainMax, 3);
    d.f64("CofA", CofA, TYPE_COUNT);
    d.f64("WallCenter", &WallCenter[0][0], WALL_END * DIM);
    d.f64("WallRotation", &WallRotation[0][0][0], WALL_END * DIM * DIM);
    d.i32("ParticleIndex", ParticleIndex, n);
    d.i32("Property", Property, n);
    d.f64("Position", &Position[0][0], 3 * n);
    d.f64("Velocity", &Velocity[0][0], 3 * n);
    d.f64("Force", &Force[0][0], 3 * n);
    d.f64("Mass", Mass, n);
    d.f64("DensityA", DensityA, n);
    d.f64("GravityCenter", &GravityCenter[0][0], 3 * n);
    d.f64("PressureA", PressureA, n);
    d.f64("VolStrainP", VolStrainP, n);
    d.f64("DivergenceP", DivergenceP, n);
    d.f64("PressureP", PressureP, n);
    d.f64("Kappa", Kappa, n);
    d.f64("Lambda", Lambda, n);
    d.f64("Mu", Mu, n);
    d.f64("alpha", alpha, n);
    d.f64("DragC", DragC, n);
    d.f64("NormalSpringConstant", NormalSpringConstant, n);
    d.f64("NormalViscousDampingConstant", NormalViscousDampingConstant, n);
    d.f64("TangentialSpringConstant", TangentialSpringConstant, n);
    d.f64("TangentialViscousDampingConstant", TangentialViscousDampingConstant, n);
    d.f64("VirialPressureAtParticle", VirialPressureAtParticle, n);
    d.f64("VirialStressAtParticle", &VirialStressAtParticle[0][0][0], 9 * n);
    d.f64("NeighborCalculatedPosition", &NeighborCalculatedPosition[0][0], 3 * n);
    d.i32("NeighborCount", NeighborCount, n);
    d.i32("FluidNeighborCount", FluidNeighborCount, n);
    d.i32("SolidNeighborCount", SolidNeighborCount, n);
    d.i64("NeighborPtr", NeighborPtr, n + 1);
    d.i32("NeighborInd", NeighborInd, NeighborIndCount);
    d.i32("CellIndex", CellIndex, n);
    d.i32("CellParticle", CellParticle, n);
    d.i32("CellFluidParticleBegin", CellFluidParticleBegin, TotalCellCount);
    d.i32("CellFluidParticleEnd", CellFluidParticleEnd, TotalCellCount);
    d.i32("CellSolidParticleBegin", CellSolidParticleBegin, TotalCellCount);
    d.i32("CellSolidParticleEnd", CellSolidParticleEnd, TotalCellCount);
    d.i32("CellWallParticleBegin", CellWallParticleBegin, TotalCellCount);
    d.i32("CellWallParticleEnd", CellWallParticleEnd, TotalCellCount);
    fclose(d.fp);
}

}  // namespace dmph_ref_harness

// usage: refNd <data> <grid> <out_prefix> <nsteps> [dump_every] [--time]
//   writes <out_prefix>_<step>.bin after initialisation (step 0), every
//   dump_every steps and after the last step.  With --time it instead times the
//   step loop (wall clock) and prints one line "steps=<k> seconds=<s>".
// usage: refNd --files <data> <grid> <prof_format> <vtk_format> <log>
//   the reference's program as it runs, file output included (run_files).
int main(int argc, char** argv) {
    using namespace dmph_ref_harness;
    if (argc == 7 && !strcmp(argv[1], "--files")) {
        load(argv[2], argv[3], argv[6]);
        run_files(argv[4], argv[5]);
        log_close();
        return 0;
    }
    if (argc < 5) {
        fprintf(stderr, "usage: %s data grid out_prefix nsteps [dump_every] [--time]\n", argv[0]);
        return 2;
    }
    const int nsteps = atoi(argv[4]);
    int every = argc > 5 && argv[5][0] != '-' ? atoi(argv[5]) : 0;
    bool timing = false;
    for (int a = 5; a < argc; a++) if (!strcmp(argv[a], "--time")) timing = true;
    std::string prefix = argv[3];
    std::string logp = prefix + ".log";
    load(argv[1], argv[2], logp.c_str());
    if (timing) {
        double t0 = omp_get_wtime();
        for (int s = 1; s <= nsteps; s++) run_step();
        double t1 = omp_get_wtime();
        printf("steps=%d seconds=%.6f particles=%d threads=%d\n", nsteps, t1 - t0, ParticleCount,
               omp_get_max_threads());
        return 0;
    }
    char path[4096];
    snprintf(path, sizeof path, "%s_%d.bin", prefix.c_str(), 0);
    dump(path, 0, false);
    for (int s = 1; s <= nsteps; s++) {
        run_step();
        if ((every > 0 && s % every == 0) || s == nsteps) {
            snprintf(path, sizeof path, "%s_%d.bin", prefix.c_str(), s);
            dump(path, s, true);
        }
    }
    return 0;
}
