// C-ABI implementation: set-up of the step's constants (in FP64, with the
// reference's own expressions), device memory, and the per-step launch sequence.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <chrono>
#include <memory>
#include <string>
#include <type_traits>
#include <vector>

#include "dem_mph.h"
#include "dmph_device.h"
#include "dmph_internal.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

struct dmph_handle {
    std::string err;
    virtual ~dmph_handle() {}
    virtual int step(int nsteps) = 0;
    virtual int sync() = 0;
    virtual int time(double *t) = 0;
    virtual int get_field(int field, int order, void *dst) = 0;
    virtual int set_field(int field, int order, const void *src) = 0;
    virtual int export_state(float *pos, float *vel) = 0;
    virtual int get_bodies(double *x, double *v, double *q, double *l, double *m) = 0;
    virtual int info(dmph_info *out) = 0;
    virtual int virial() = 0;
    virtual int count_nonfinite(int field, int64_t *count) = 0;
    virtual int pairs(int64_t *nnz, int64_t *ptr, int *idx) = 0;
    virtual int write_prof(const char *path) = 0;
    virtual int write_vtk(const char *path) = 0;
    virtual int set_timing(int on) = 0;
    virtual int autotune(int reps, int *lanes) = 0;
    virtual int get_timing(int group, double *ms, int64_t *launches) = 0;
};

namespace dmph {

#define HCHK(expr)                                                                              \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess) {                                                                 \
            err = std::string(#expr) + " failed: " + hipGetErrorString(e_);                    \
            return DMPH_E_DEVICE;                                                               \
        }                                                                                       \
    } while (0)

static inline double modw(double x, double w) { return x - w * floor(x / w); }
static inline bool is_solid(int t) { return 2 <= t && t < 4; }
static inline bool is_wall(int t) { return 4 <= t && t < 6; }
static inline bool is_fluid(int t) { return 0 <= t && t < 2; }

// Kernel normalisations and reference number densities: initializeWeight,
// source/DEM-MPH.cpp:1007-1125, written with the reference's association.
static void weights(Geo &g, const dmph_params &p) {
    const double d = g.d;
    g.ra = p.radius_ratio_a * d;
    g.rg = p.radius_ratio_a * d;   // RadiusRatioG = RadiusRatioA
    g.rp = p.radius_ratio_p * d;
    g.rv = p.radius_ratio_v * d;
    if (g.dim == 2) {
        g.swa = 1.0 / 2.0 * 2.0 / 15.0 * M_PI / d / d;
        g.swg = 1.0 / 2.0 * 1.0 / 3.0 * M_PI / d / d;
        g.swp = 1.0 / 2.0 * 1.0 / 3.0 * M_PI / d / d;
        g.swv = 1.0 / 2.0 * 1.0 / 3.0 * M_PI / d / d;
        g.r2g = 1.0 / 2.0 * 1.0 / 30.0 * M_PI * g.rg * g.rg / d / d / g.swg;
    } else {
        g.swa = 1.0 / 3.0 * 1.0 / 5.0 * M_PI / d / d / d;
        g.swg = 1.0 / 3.0 * 2.0 / 5.0 * M_PI / d / d / d;
        g.swp = 1.0 / 3.0 * 2.0 / 5.0 * M_PI / d / d / d;
        g.swv = 1.0 / 3.0 * 2.0 / 5.0 * M_PI / d / d / d;
        g.r2g = 1.0 / 3.0 * 4.0 / 105.0 * M_PI * g.rg * g.rg / d / d / d / g.swg;
    }
    for (int which = 0; which < 2; which++) {
        const double R = which == 0 ? g.ra : g.rp;
        const double sw = which == 0 ? g.swa : g.swp;
        const double hd = g.dim == 2 ? R * R : R * R * R;
        const int range = (int)(R / d + 3.0);
        const int zr = g.dim == 2 ? 0 : range;
        double sum = 0.0;
        for (int ix = -range; ix <= range; ++ix)
            for (int iy = -range; iy <= range; ++iy)
                for (int iz = -zr; iz <= zr; ++iz) {
                    if (ix == 0 && iy == 0 && iz == 0) continue;
                    const double x = d * (double)ix, y = d * (double)iy, z = d * (double)iz;
                    const double r2 = g.dim == 2 ? x * x + y * y : x * x + y * y + z * z;
                    if (r2 <= R * R) {
                        const double r = sqrt(r2);
                        sum += which == 0 ? 1.0 / sw * 1.0 / hd * (r / R) * (1.0 - (r / R)) * (1.0 - (r / R))
                                          : 1.0 / sw * 1.0 / hd * ((1.0 - r / R) * (1.0 - r / R));
                    }
                }
        (which == 0 ? g.n0a : g.n0p) = sum;
    }
}

// initializeFluid constants (source/DEM-MPH.cpp:1143-1155) and the per-type tables
static void tables(Geo &g, const dmph_params &p) {
    double integ_n, integ_x;
    if (g.dim == 2) { g.cofk = 0.350778153; integ_n = 0.024679383; integ_x = 0.226126699; }
    else { g.cofk = 0.326976006; integ_n = 0.021425779; integ_x = 0.233977488; }
    for (int t = 0; t < kNType; t++) {
        g.cofa[t] = p.surface_tension[t] / ((g.rg / g.d) * (integ_n + g.cofk * g.cofk * integ_x));
        g.dens[t] = p.density[t];
        g.bulk[t] = p.bulk_modulus[t];
        g.bvisc[t] = p.bulk_viscosity[t];
        g.shear[t] = p.shear_viscosity[t];
        g.young[t] = p.young_modulus[t];
        g.poisson[t] = p.poisson_ratio[t];
        g.logrest[t] = log(p.restitution[t]);
        g.fric[t] = p.friction[t];
    }
    g.st_active = 0;
    for (int a = 0; a < kNType; a++)
        for (int b = 0; b < kNType; b++) {
            g.inter[a * kNType + b] = p.interaction_ratio[a][b];
            if (p.interaction_ratio[a][b] != 0.0) g.st_active = 1;
        }
    for (int k = 0; k < 3; k++) g.grav[k] = p.gravity[k];
}

// initializeWall (source/DEM-MPH.cpp:1166-1205), with theta = |Omega|^2 as written there
static void walls(Geo &g, const dmph_params &p) {
    for (int w = 4; w < 6; w++) {
        const double *om = p.wall_omega[w];
        double nrm[3] = {0.0, 0.0, 0.0}, q[4];
        const double theta = fabs(om[0] * om[0] + om[1] * om[1] + om[2] * om[2]);
        if (theta != 0.0)
            for (int a = 0; a < 3; a++) nrm[a] = om[a] / theta;
        q[0] = nrm[0] * sin(theta * p.dt / 2.0);
        q[1] = nrm[1] * sin(theta * p.dt / 2.0);
        q[2] = nrm[2] * sin(theta * p.dt / 2.0);
        q[3] = cos(theta * p.dt / 2.0);
        double *R = g.wall_rot[w - 4];
        R[0] = q[0] * q[0] - q[1] * q[1] - q[2] * q[2] + q[3] * q[3];
        R[1] = 2.0 * (q[0] * q[1] - q[2] * q[3]);
        R[2] = 2.0 * (q[0] * q[2] + q[1] * q[3]);
        R[3] = 2.0 * (q[0] * q[1] + q[2] * q[3]);
        R[4] = -q[0] * q[0] + q[1] * q[1] - q[2] * q[2] + q[3] * q[3];
        R[5] = 2.0 * (q[1] * q[2] - q[0] * q[3]);
        R[6] = 2.0 * (q[0] * q[2] - q[1] * q[3]);
        R[7] = 2.0 * (q[1] * q[2] + q[0] * q[3]);
        R[8] = -q[0] * q[0] - q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
        for (int a = 0; a < 3; a++) {
            g.wall_omega[w - 4][a] = om[a];
            g.wall_vel[w - 4][a] = p.wall_velocity[w][a];
        }
    }
}

// initializeDomain (source/DEM-MPH.cpp:1207-1262)
// initializeDomain (source/DEM-MPH.cpp:1207-1262); cell_scale > 1 (FP32 only)
// widens the cells beyond MaxRadius + 0.1 d
static int domain(Geo &g, const dmph_grid &gr, double cell_scale, std::string &err) {
    double m = 0.0;
    m = g.ra > m ? g.ra : m;
    m = g.rg > m ? g.rg : m;
    m = g.rp > m ? g.rp : m;
    m = g.rv > m ? g.rv : m;
    g.rmax = m;
    const double margin = 0.1 * g.d;
    double cnt[3];
    for (int a = 0; a < 3; a++) {
        g.dmin[a] = gr.domain_min[a];
        g.dw[a] = gr.domain_max[a] - gr.domain_min[a];
    }
    const double cmin = cell_scale > 1.0 ? cell_scale * (g.rmax + margin) : g.rmax + margin;
    cnt[0] = floor((gr.domain_max[0] - gr.domain_min[0]) / cmin);
    cnt[1] = floor((gr.domain_max[1] - gr.domain_min[1]) / cmin);
    cnt[2] = g.dim == 2 ? 1 : floor((gr.domain_max[2] - gr.domain_min[2]) / cmin);
    for (int a = 0; a < 3; a++) g.cc[a] = (int)cnt[a];
    if (g.cc[0] < 3 || g.cc[1] < 3 || (g.dim == 3 && g.cc[2] < 3)) {
        char b[200];
        snprintf(b, sizeof b, "cell grid %d x %d x %d: the periodic stencil needs at least 3 cells per axis", g.cc[0],
                 g.cc[1], g.cc[2]);
        err = b;
        return DMPH_E_DOMAIN;
    }
    const double cells = cnt[0] * cnt[1] * cnt[2];
    if (cells * 3 * g.nscene > 2.0e9) {
        err = "too many cells";
        return DMPH_E_DOMAIN;
    }
    g.ncell = (int)cells;
    g.T = g.ncell * g.nscene;
    for (int a = 0; a < 3; a++) g.cw[a] = g.dw[a] / g.cc[a];
    for (int a = 0; a < g.dim; a++)
        if ((int)ceil((g.rmax + margin) / g.cw[a]) != 1) {
            err = "stencil wider than one cell (source/DEM-MPH.cpp:1540-1543)";
            return DMPH_E_DOMAIN;
        }
    g.cut2 = (g.rmax + 0.1 * g.d) * (g.rmax + 0.1 * g.d);
    g.skin2 = 0.5 * (0.1 * g.d) * 0.5 * (0.1 * g.d);
    return DMPH_OK;
}

static void inv3(const double A[9], double B[9]) {
    const double det = A[0] * (A[4] * A[8] - A[5] * A[7]) - A[1] * (A[3] * A[8] - A[5] * A[6]) +
                       A[2] * (A[3] * A[7] - A[4] * A[6]);
    if (det == 0.0) { memset(B, 0, 9 * sizeof(double)); return; }
    B[0] = (A[4] * A[8] - A[5] * A[7]) / det; B[1] = (A[2] * A[7] - A[1] * A[8]) / det; B[2] = (A[1] * A[5] - A[2] * A[4]) / det;
    B[3] = (A[5] * A[6] - A[3] * A[8]) / det; B[4] = (A[0] * A[8] - A[2] * A[6]) / det; B[5] = (A[2] * A[3] - A[0] * A[5]) / det;
    B[6] = (A[3] * A[7] - A[4] * A[6]) / det; B[7] = (A[1] * A[6] - A[0] * A[7]) / det; B[8] = (A[0] * A[4] - A[1] * A[3]) / det;
}

template <class R>
class Engine final : public dmph_handle {
  public:
    Geo g{};
    Dev<R> a{};
    dmph_params p{};
    dmph_options o{};
    hipStream_t stream = nullptr;
    // the second branch of a captured FP32 step (solids' pass A, DEM, body update)
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_solid = nullptr, ev_join = nullptr;
    std::vector<void *> allocs;
    int device = 0;
    double spacing = 0;
    double cell_scale = 1.0;
    double dmax[3] = {0, 0, 0};

    ~Engine() override {
        if (stream) (void)hipStreamSynchronize(stream);
        drop_graph();
        for (auto &p_ : pending) { ev_pool.push_back(p_.second.first); ev_pool.push_back(p_.second.second); }
        for (hipEvent_t e : ev_pool) (void)hipEventDestroy(e);
        for (void *q : allocs) (void)hipFree(q);
        for (hipEvent_t e : {ev_fork, ev_solid, ev_join})
            if (e) (void)hipEventDestroy(e);
        if (side) (void)hipStreamDestroy(side);
        if (stream) (void)hipStreamDestroy(stream);
    }

    template <class T>
    int alloc(T *&ptr, size_t count) {
        void *q = nullptr;
        HCHK(hipMalloc(&q, (count ? count : 1) * sizeof(T)));
        allocs.push_back(q);
        HCHK(hipMemsetAsync(q, 0, (count ? count : 1) * sizeof(T), stream));
        ptr = (T *)q;
        return DMPH_OK;
    }

#define TRY(x) do { int rc_ = (x); if (rc_ != DMPH_OK) return rc_; } while (0)

    // scene_of: scene of each particle of a batch (nscene > 1), or null
    int create(const dmph_params *pp, const dmph_options *op, const dmph_grid *gr, const int *scene_of = nullptr,
               int nscene = 1) {
        p = *pp;
        o = *op;
        g.nscene = nscene;
        const int n = gr->n;
        if (n <= 0 || !gr->type || !gr->pos || !gr->vel || gr->spacing <= 0.0) {
            err = "empty or invalid particle set";
            return DMPH_E_ARG;
        }
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
            err = "no HIP device available";
            return DMPH_E_DEVICE;
        }
        device = o.device;
        HCHK(hipSetDevice(device));
        HCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        HCHK(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
        HCHK(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
        HCHK(hipEventCreateWithFlags(&ev_solid, hipEventDisableTiming));
        HCHK(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
        g.n = n;
        g.d = spacing = gr->spacing;
        for (int k = 0; k < 3; k++) dmax[k] = gr->domain_max[k];
        int dim = o.dim;
        if (dim != 2 && dim != 3) {
            // 2D when the z extent cannot hold three cells of MaxRadius + 0.1 d (the shipped 2D case is one layer)
            const double rmax = gr->spacing * std::fmax(std::fmax(p.radius_ratio_a, p.radius_ratio_p), p.radius_ratio_v);
            dim = (gr->domain_max[2] - gr->domain_min[2]) < 3.0 * (rmax + 0.1 * gr->spacing) ? 2 : 3;
        }
        g.dim = dim;
        g.vol = dim == 2 ? g.d * g.d : g.d * g.d * g.d;
        g.dt = p.dt;
        g.dem_dt = p.dem_dt;
        g.substeps = (int)(p.dt / p.dem_dt);   // source/DEM-MPH.cpp:434
        g.compensated = o.compensated;
        g.lanes[0] = g.lanes[1] = (o.lanes == 2 || o.lanes == 4 || o.lanes == 16) ? o.lanes : 8;
        g.lanes_w = (o.wall_lanes == 1 || o.wall_lanes == 4) ? o.wall_lanes : 2;
        g.block[0] = g.block[1] = (o.block == 64 || o.block == 128) ? o.block : 256;
        g.body_f32 = o.rigid_precision == DMPH_PREC_FP32;
        g.fuse_accel = 1;
        g.fused = getenv("DMPH_SWEEPS") && !strcmp(getenv("DMPH_SWEEPS"), "separate") ? 0 : 1;
        g.stage_b = getenv("DMPH_PASS_B") && !strcmp(getenv("DMPH_PASS_B"), "lds") ? 1 : 0;
        g.sort_reference = o.sort_mode == DMPH_SORT_REFERENCE;
        g.rebuild_every = o.rebuild == DMPH_REBUILD_EVERY_STEP;
        weights(g, p);
        tables(g, p);
        walls(g, p);
        const bool fp32 = !std::is_same<R, double>::value;
        cell_scale = o.cell_scale > 1.0 ? o.cell_scale : 1.0;
        if (cell_scale > 1.0 && (!fp32 || cell_scale > 2.0)) {
            err = "cell_scale above 1 needs FP32 and at most 2 (FP64 keeps the reference's cell grid)";
            return DMPH_E_ARG;
        }
        TRY(domain(g, *gr, cell_scale, err));
        int pw = 0;
        while ((n >> pw) != 0) ++pw;
        g.p2_log = pw;
        g.p2 = 1 << pw;
        // block ranges are fixed by the type counts (fluid, solid, wall keys in that order)
        int nf = 0, ns = 0;
        for (int i = 0; i < n; i++) {
            const int t = gr->type[i];
            if (is_solid(t)) ns++;
            else if (is_wall(t)) {}
            else if (is_fluid(t)) nf++;
            else { err = "particle type outside 0..5"; return DMPH_E_FORMAT; }
        }
        g.fb = 0; g.fe = nf; g.sb = nf; g.se = nf + ns; g.wb = nf + ns; g.we = n;
        if (n >= (1 << 26)) {   // neighbour entries pack the slot index in 26 bits
            err = "more than 67,108,863 particles";
            return DMPH_E_ARG;
        }
        {   // neighbour-list capacity: 2 x the lattice count within MaxRadius + 0.1 d, in
            // multiples of 16; FP32 rows split it into a near and a far segment, each
            // sized from its own lattice count
            // near segment (FP32): every pass but the viscosity and surface-tension
            // ones stays within RadiusP; the 0.1 d skin keeps it conservative
            const double rn = (g.st_active ? std::max(g.rp, g.ra) : g.rp) + 0.1 * g.d;
            g.near2 = rn * rn;
            const double rc = (g.rmax + 0.1 * g.d) / g.d, rnd = rn / g.d;
            const int rr = (int)std::max(rc, rnd) + 1, zr = g.dim == 2 ? 0 : rr;
            int cnt = 0, cnn = 0;
            for (int ix = -rr; ix <= rr; ix++)
                for (int iy = -rr; iy <= rr; iy++)
                    for (int iz = -zr; iz <= zr; iz++) {
                        const double q = (double)(ix * ix + iy * iy + iz * iz);
                        cnt += q <= rc * rc;
                        cnn += q <= std::min(rc, rnd) * std::min(rc, rnd);
                    }
            auto up16 = [](double x) { return ((int)ceil(x) + 15) / 16 * 16; };
            // headroom over a lattice: the dam-break fluid piles up against the
            // walls (rows of 142 entries, far segments of 80+, measured over
            // 0.2 s of the C3 scenes); capacity costs memory only, the sweeps
            // read the entries a row holds
            const double kFill = 2.0;
            if (std::is_same<R, double>::value) {
                g.kcap = g.knear = up16(kFill * cnt);
            } else {
                g.knear = up16(kFill * cnn);
                g.kcap = g.knear + up16(kFill * (cnt - cnn));
            }
            if (getenv("DMPH_NEIGHBOR_CAPACITY")) {
                const int total = std::max(32, up16(atoi(getenv("DMPH_NEIGHBOR_CAPACITY"))));
                g.kcap = total;
                g.knear = std::is_same<R, double>::value
                              ? total
                              : std::min(total - 16, std::max(16, up16((double)total * cnn / std::max(cnt, 1))));
            }
            g.ccap = 32;
            // contact list: within d now means within d + 0.1 d at rebinning; a further
            // 0.15 d covers the solids' motion over the DEM sub-steps of one step
            g.con2 = (1.25 * g.d) * (1.25 * g.d);
        }
        g.rigid = o.debris == DMPH_DEBRIS_RIGID && gr->body != nullptr;
        TRY(allocate());
        TRY(upload(*gr, scene_of));
        // first binning + the pre-loop density passes (source/DEM-MPH.cpp:419-425)
        Scalars sc{};
        sc.time = gr->time;
        sc.flag = 1;
        sc.full_sweep = 2;   // the first step's sweep 1 covers every particle
        sc.rebuilds = 1;
        for (int w = 0; w < 2; w++)
            for (int k = 0; k < 3; k++) sc.wall_center[w][k] = p.wall_center[w + 4][k];
        HCHK(hipMemcpyAsync(a.sc, &sc, sizeof sc, hipMemcpyHostToDevice, stream));
        // FP32: the first rows are decided on the input positions (as the FP64
        // path decides them), not on their FP32 offsets; the copy is freed after
        double *x64 = nullptr;
        if (!std::is_same<R, double>::value) {
            HCHK(hipMalloc(&x64, (size_t)n * 3 * sizeof(double)));
            const hipError_t e = hipMemcpyAsync(x64, gr->pos, (size_t)n * 3 * sizeof(double), hipMemcpyHostToDevice,
                                                stream);
            if (e != hipSuccess) { (void)hipFree(x64); HCHK(e); }
            a.x64 = x64;
        }
        launch_rebin<R>(g, a, stream, 0);
        launch_init_density<R>(g, a, stream);
        a.x64 = nullptr;
        const hipError_t e1 = hipGetLastError(), e2 = hipStreamSynchronize(stream);
        if (x64) (void)hipFree(x64);
        HCHK(e1);
        HCHK(e2);
        return DMPH_OK;
    }

    int allocate() {
        const size_t n = g.n, nk = 3 * (size_t)g.T + 1;
        TRY(alloc(a.id, n)); TRY(alloc(a.type, n)); TRY(alloc(a.body, n)); TRY(alloc(a.key, n));
        R **reals[] = {&a.px, &a.py, &a.pz, &a.vx, &a.vy, &a.vz, &a.rho, &a.fx, &a.fy, &a.fz, &a.bx, &a.by, &a.bz,
                       &a.dens_a, &a.gcx, &a.gcy, &a.gcz, &a.vs_p, &a.div_p, &a.press_p, &a.press_a, &a.mass,
                       &a.kappa, &a.alpha, &a.drag_c, &a.kn, &a.etan, &a.kt, &a.etat, &a.virial_p};
        for (R **r : reals) TRY(alloc(*r, n));
        TRY(alloc(a.ncount, n)); TRY(alloc(a.scount, n));
        TRY(alloc(a.nlist, (size_t)g.kcap * n));
        TRY(alloc(a.nnear, n));
        TRY(alloc(a.clist, (size_t)g.ccap * (n - g.sb)));
        TRY(alloc(a.ncontact, n - g.sb));
        TRY(alloc(a.cell_start, nk)); TRY(alloc(a.counts, nk)); TRY(alloc(a.fill, nk));
        TRY(alloc(a.key_new, n)); TRY(alloc(a.order, n)); TRY(alloc(a.perm, n));
        if (g.sort_reference) { TRY(alloc(a.bkey, (size_t)g.p2)); TRY(alloc(a.bval, (size_t)g.p2)); }
        TRY(alloc(a.block_sums, nk / 1024 + 2));
        TRY(alloc(a.wall_counts, (size_t)g.T + 1));
        TRY(alloc(a.t_int, 5 * n)); TRY(alloc(a.t_real, 8 * n));
        TRY(alloc(a.t_dbl, g.rigid ? 3 * n : 1)); TRY(alloc(a.r0, g.rigid ? 3 * n : 1));
        TRY(alloc(a.sc, 1));
        TRY(alloc(a.skin_part, 8 * 16));
        TRY(alloc(a.nonfinite, 1));
        if (!std::is_same<R, double>::value) {
            TRY(alloc(a.p4, n + 4)); TRY(alloc(a.v4, n));   // the list build reads up to 3 slots past a cell
            TRY(alloc(a.wact, n - g.wb));
            TRY(alloc(a.wnfs, n - g.wb));
            TRY(alloc(a.wsq, n - g.wb));
            TRY(alloc(a.wdem, n - g.wb));
            TRY(alloc(a.pp4, n));
        }
        return DMPH_OK;
    }

    // host copies in file order -> slots (slot = file index until the first binning)
    int upload(const dmph_grid &gr, const int *scene_of) {
        const int n = g.n;
        std::vector<int> id(n), type(n), body(n, -1), key(n, 0);
        // the scene rides in the key (scene * cells + cell) from the start: k_keys reads it back
        if (scene_of)
            for (int i = 0; i < n; i++) key[i] = scene_of[i] * g.ncell;
        std::vector<R> px(n), py(n), pz(n), vx(n), vy(n), vz(n), rho(n), mass(n), kappa(n);
        std::vector<double> rhod(n);
        for (int i = 0; i < n; i++) {
            id[i] = i;
            const int t = type[i] = gr.type[i];
            if (g.rigid && is_solid(t) && gr.body[i] >= 0) body[i] = gr.body[i];
            rhod[i] = (gr.density && gr.density[i] > 0.0) ? gr.density[i] : p.density[t];
            rho[i] = (R)rhod[i];
            mass[i] = (R)(rhod[i] * g.vol);            // initializeFluid writes these in file order (:1130-1141)
            kappa[i] = (R)p.bulk_modulus[t];
            vx[i] = (R)gr.vel[3 * i]; vy[i] = (R)gr.vel[3 * i + 1]; vz[i] = (R)gr.vel[3 * i + 2];
            const double *x = gr.pos + 3 * i;
            if (std::is_same<R, double>::value) {
                px[i] = (R)x[0]; py[i] = (R)x[1]; pz[i] = (R)x[2];
            } else {
                int c[3];
                double off[3];
                for (int k = 0; k < 3; k++) {
                    c[k] = ((int)floor((x[k] - g.dmin[k]) / g.cw[k])) % g.cc[k];
                    c[k] = (c[k] % g.cc[k] + g.cc[k]) % g.cc[k];
                    const double o_ = x[k] - g.dmin[k] - c[k] * g.cw[k];
                    off[k] = modw(o_ + 0.5 * g.dw[k], g.dw[k]) - 0.5 * g.dw[k];
                }
                key[i] += (c[0] * g.cc[1] + c[1]) * g.cc[2] + c[2];
                px[i] = (R)off[0]; py[i] = (R)off[1]; pz[i] = (R)off[2];
            }
        }
        auto up = [&](auto *dst, const auto &src) -> int {
            HCHK(hipMemcpyAsync(dst, src.data(), src.size() * sizeof(src[0]), hipMemcpyHostToDevice, stream));
            return DMPH_OK;
        };
        TRY(up(a.id, id)); TRY(up(a.type, type)); TRY(up(a.body, body)); TRY(up(a.key, key));
        TRY(up(a.px, px)); TRY(up(a.py, py)); TRY(up(a.pz, pz)); TRY(up(a.vx, vx)); TRY(up(a.vy, vy)); TRY(up(a.vz, vz));
        TRY(up(a.bx, px)); TRY(up(a.by, py)); TRY(up(a.bz, pz));
        TRY(up(a.rho, rho)); TRY(up(a.mass, mass)); TRY(up(a.kappa, kappa));
        if (g.rigid) TRY(init_bodies(gr, body, rhod));
        HCHK(hipStreamSynchronize(stream));   // host vectors go out of scope
        return DMPH_OK;
    }

    // Rigid bodies (no reference counterpart): mass, centre (minimum image from
    // the first member), body-frame offsets, inertia of point masses plus each
    // particle's sphere 2/5 m (d/2)^2, and L = I omega with omega fitted to the
    // point-mass momenta.
    int init_bodies(const dmph_grid &gr, const std::vector<int> &body, const std::vector<double> &rho) {
        const int n = g.n;
        int nb = 0;
        for (int i = 0; i < n; i++) nb = body[i] >= nb ? body[i] + 1 : nb;
        g.nb = nb;
        if (!nb) { g.rigid = 0; return DMPH_OK; }
        std::vector<double> M(nb, 0.0), X(3 * nb, 0.0), V(3 * nb, 0.0), L(3 * nb, 0.0), Q(4 * nb, 0.0), I(9 * nb, 0.0),
            Ip(9 * nb, 0.0), Iinv(9 * nb, 0.0), Rm(9 * nb, 0.0), W(3 * nb, 0.0), acc(3 * nb, 0.0), r0(3 * (size_t)n, 0.0);
        std::vector<int> first(nb, -1);
        auto mimg = [&](const double *x, const double *y, double *o_) {
            for (int k = 0; k < 3; k++) o_[k] = modw(x[k] - y[k] + 0.5 * g.dw[k], g.dw[k]) - 0.5 * g.dw[k];
        };
        for (int i = 0; i < n; i++) {
            const int b = body[i];
            if (b < 0) continue;
            if (first[b] < 0) first[b] = i;
            const double m = rho[i] * g.vol;
            double dx[3];
            mimg(gr.pos + 3 * i, gr.pos + 3 * first[b], dx);
            M[b] += m;
            for (int k = 0; k < 3; k++) { acc[3 * b + k] += m * dx[k]; V[3 * b + k] += m * gr.vel[3 * i + k]; }
        }
        for (int b = 0; b < nb; b++) {
            Q[4 * b] = 1.0;
            if (first[b] < 0 || M[b] == 0.0) continue;
            for (int k = 0; k < 3; k++) {
                X[3 * b + k] = gr.pos[3 * first[b] + k] + acc[3 * b + k] / M[b];
                V[3 * b + k] /= M[b];
            }
        }
        const double rs = 0.5 * g.d;
        for (int i = 0; i < n; i++) {
            const int b = body[i];
            if (b < 0) continue;
            const double m = rho[i] * g.vol;
            double *r = &r0[3 * (size_t)i];
            mimg(gr.pos + 3 * i, &X[3 * b], r);
            const double rr = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
            for (int x = 0; x < 3; x++)
                for (int y = 0; y < 3; y++) {
                    const double pt = m * ((x == y ? rr : 0.0) - r[x] * r[y]);
                    Ip[9 * b + 3 * x + y] += pt;
                    I[9 * b + 3 * x + y] += pt + (x == y ? 0.4 * m * rs * rs : 0.0);
                }
            const double u[3] = {gr.vel[3 * i] - V[3 * b], gr.vel[3 * i + 1] - V[3 * b + 1], gr.vel[3 * i + 2] - V[3 * b + 2]};
            L[3 * b + 0] += m * (r[1] * u[2] - r[2] * u[1]);
            L[3 * b + 1] += m * (r[2] * u[0] - r[0] * u[2]);
            L[3 * b + 2] += m * (r[0] * u[1] - r[1] * u[0]);
        }
        for (int b = 0; b < nb; b++) {
            inv3(&I[9 * b], &Iinv[9 * b]);
            double ip[9], w[3];
            inv3(&Ip[9 * b], ip);
            bool ok = false;
            for (int k = 0; k < 9; k++) ok |= ip[k] != 0.0;
            if (ok) {
                for (int k = 0; k < 3; k++) w[k] = ip[3 * k] * L[3 * b] + ip[3 * k + 1] * L[3 * b + 1] + ip[3 * k + 2] * L[3 * b + 2];
                for (int k = 0; k < 3; k++)
                    L[3 * b + k] = I[9 * b + 3 * k] * w[0] + I[9 * b + 3 * k + 1] * w[1] + I[9 * b + 3 * k + 2] * w[2];
            }
            for (int k = 0; k < 3; k++) Rm[9 * b + 4 * k] = 1.0;
            for (int k = 0; k < 3; k++)
                W[3 * b + k] = Iinv[9 * b + 3 * k] * L[3 * b] + Iinv[9 * b + 3 * k + 1] * L[3 * b + 1] + Iinv[9 * b + 3 * k + 2] * L[3 * b + 2];
        }
        TRY(alloc(a.b_mass, nb)); TRY(alloc(a.b_x, 3 * nb)); TRY(alloc(a.b_v, 3 * nb)); TRY(alloc(a.b_l, 3 * nb));
        TRY(alloc(a.b_q, 4 * nb)); TRY(alloc(a.b_iinv, 9 * nb));
        TRY(alloc(a.b_w, 3 * nb)); TRY(alloc(a.b_r, 9 * nb));
        // member map: each member's rank in its body (input order), the body's
        // offsets; the rebinning (k_place or k_gather) fills the slots of the ranks
        std::vector<int> mrank(n, -1), moff(nb + 1, 0);
        g.nloose = 0;
        for (int i = 0; i < n; i++) {
            if (body[i] >= 0) mrank[i] = moff[body[i] + 1]++;
            else g.nloose += is_solid(gr.type[i]);
        }
        int most = 0;
        for (int b = 0; b < nb; b++) most = std::max(most, moff[b + 1]);
        for (int b = 0; b < nb; b++) moff[b + 1] += moff[b];
        g.body_tpb = 64;   // one member per thread up to 1024 members
        while (g.body_tpb < most && g.body_tpb < 1024) g.body_tpb *= 2;
        TRY(alloc(a.mrank, n)); TRY(alloc(a.b_moff, nb + 1)); TRY(alloc(a.b_mslot, std::max(moff[nb], 1)));
        HCHK(hipMemcpyAsync(a.mrank, mrank.data(), n * sizeof(int), hipMemcpyHostToDevice, stream));
        HCHK(hipMemcpyAsync(a.b_moff, moff.data(), (nb + 1) * sizeof(int), hipMemcpyHostToDevice, stream));
        HCHK(hipStreamSynchronize(stream));   // the host vectors go out of scope
        body_mass_host = M;
        auto up = [&](double *dst, const std::vector<double> &src) -> int {
            HCHK(hipMemcpyAsync(dst, src.data(), src.size() * sizeof(double), hipMemcpyHostToDevice, stream));
            return DMPH_OK;
        };
        TRY(up(a.b_mass, M)); TRY(up(a.b_x, X)); TRY(up(a.b_v, V)); TRY(up(a.b_l, L)); TRY(up(a.b_q, Q));
        TRY(up(a.b_iinv, Iinv)); TRY(up(a.b_r, Rm)); TRY(up(a.b_w, W)); TRY(up(a.r0, r0));
        HCHK(hipStreamSynchronize(stream));
        return DMPH_OK;
    }
    std::vector<double> body_mass_host;

    // per-group device timing: event pairs around each launch group, resolved lazily
    bool timing = false;
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> pending;
    double t_ms[DMPH_TIMING_GROUPS] = {0};
    int64_t t_n[DMPH_TIMING_GROUPS] = {0};

    hipEvent_t take_event() {
        if (!ev_pool.empty()) { hipEvent_t e = ev_pool.back(); ev_pool.pop_back(); return e; }
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        return e;
    }

    template <class F>
    void timed(int group, F &&launch) {
        if (!timing) { launch(); return; }
        hipEvent_t e0 = take_event(), e1 = take_event();
        if (e0 && e1) (void)hipEventRecord(e0, stream);
        launch();
        if (e0 && e1) {
            (void)hipEventRecord(e1, stream);
            pending.push_back({group, {e0, e1}});
        }
    }

    int resolve_timing() {
        HCHK(hipStreamSynchronize(stream));
        for (auto &p_ : pending) {
            float ms = 0.f;
            HCHK(hipEventElapsedTime(&ms, p_.second.first, p_.second.second));
            t_ms[p_.first] += ms;
            t_n[p_.first] += 1;
            ev_pool.push_back(p_.second.first);
            ev_pool.push_back(p_.second.second);
        }
        pending.clear();
        return DMPH_OK;
    }

    int set_timing(int on) override {
        TRY(resolve_timing());
        timing = on != 0;
        for (int k = 0; k < DMPH_TIMING_GROUPS; k++) { t_ms[k] = 0; t_n[k] = 0; }
        return DMPH_OK;
    }

    int get_timing(int group, double *ms, int64_t *n) override {
        if (group < 0 || group >= DMPH_TIMING_GROUPS) { err = "bad timing group"; return DMPH_E_ARG; }
        TRY(resolve_timing());
        if (ms) *ms = t_ms[group];
        if (n) *n = t_n[group];
        return DMPH_OK;
    }

    // One pass of the loop body: every launch of a step, in order.  No host
    // synchronisation and no memset: the rebin decision is a device flag that
    // the rebin kernels read, so the same sequence serves every step.
    //
    // Captured FP32 steps with solid slots and the fused passes run in two
    // graph branches after the rebin kernels: the fluid and wall slots' pass A
    // then pass B on the main stream; the solid slots' pass A, the DEM
    // sub-steps and the body update beside them.  Pass B reads the solids'
    // (position, PressureP) from their pass A, so it waits for that launch;
    // the rest of the solid branch reads nothing the fluid and wall passes write
    // (the DEM pass takes a fluid or wall contact's Mass from Density x Volume,
    // its force loop reads solid and wall slots only) and writes only solid and
    // wall slots' forces and kinematics, which those passes do not read (they
    // read the packed copies k_predict wrote).  The step ends when both
    // branches have.
    bool capturing = false;
    int enqueue_step() {
        timed(DMPH_TIMING_PREDICT, [&] { launch_predict<R>(g, a, stream); });
        timed(DMPH_TIMING_REBIN, [&] { launch_rebin<R>(g, a, stream, 1); });
        const bool branch = capturing && !std::is_same<R, double>::value && g.fused && !g.st_active &&
                            g.se > g.sb && !(getenv("DMPH_BRANCHES") && !strcmp(getenv("DMPH_BRANCHES"), "serial"));
        if (branch) {
            HCHK(hipEventRecord(ev_fork, stream));
            HCHK(hipStreamWaitEvent(side, ev_fork, 0));
            launch_sweep1<R>(g, a, stream, side);
            HCHK(hipEventRecord(ev_solid, side));
            for (int s = 0; s < g.substeps; s++) {
                launch_dem<R>(g, a, side);
                launch_move<R>(g, a, side);
            }
            HCHK(hipEventRecord(ev_join, side));
            HCHK(hipStreamWaitEvent(stream, ev_solid, 0));
            launch_sweep2<R>(g, a, stream);
            HCHK(hipStreamWaitEvent(stream, ev_join, 0));
            return DMPH_OK;
        }
        timed(DMPH_TIMING_SWEEP1, [&] { launch_sweep1<R>(g, a, stream); });
        timed(DMPH_TIMING_SWEEP2, [&] { launch_sweep2<R>(g, a, stream); });
        // FP32 applies the fluid acceleration inside sweep 2: no group of its own
        if (std::is_same<R, double>::value || !g.fuse_accel)
            timed(DMPH_TIMING_ACCEL, [&] { launch_accel<R>(g, a, stream); });
        for (int s = 0; s < g.substeps; s++) {
            timed(DMPH_TIMING_DEM, [&] { launch_dem<R>(g, a, stream); });
            timed(DMPH_TIMING_MOVE, [&] { launch_move<R>(g, a, stream); });
        }
        return DMPH_OK;
    }

    // The step as a HIP graph, captured on first use and replayed once per step.
    // One step, and kGraphSteps steps in a row: dmph_step(n) replays the long
    // graph while n allows, so consecutive steps run without a graph launch
    // between them (each launch left the GPU idle ~10 us between steps).
    static constexpr int kGraphSteps = 10;
    hipGraph_t graph = nullptr, graph_k = nullptr;
    hipGraphExec_t graph_exec = nullptr, graph_exec_k = nullptr;
    void drop_graph() {
        for (hipGraphExec_t *x : {&graph_exec, &graph_exec_k})
            if (*x) { (void)hipGraphExecDestroy(*x); *x = nullptr; }
        for (hipGraph_t *x : {&graph, &graph_k})
            if (*x) { (void)hipGraphDestroy(*x); *x = nullptr; }
    }

    int capture_steps(int k, hipGraph_t &graph_out, hipGraphExec_t &exec_out) {
        HCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        capturing = true;
        int rc = DMPH_OK;
        for (int s = 0; s < k && rc == DMPH_OK; s++) rc = enqueue_step();
        capturing = false;
        hipGraph_t gr = nullptr;
        const hipError_t e = hipStreamEndCapture(stream, &gr);
        if (rc != DMPH_OK) {
            if (gr) (void)hipGraphDestroy(gr);
            return rc;
        }
        if (e != hipSuccess || !gr) {
            err = std::string("hipStreamEndCapture: ") + hipGetErrorString(e);
            return DMPH_E_DEVICE;
        }
        graph_out = gr;
        HCHK(hipGraphInstantiate(&exec_out, graph_out, nullptr, nullptr, 0));
        return DMPH_OK;
    }

    // FP32 sweeps with 2, 4, 8 and 16 lanes per particle, timed on the current
    // state by re-running them (they recompute the same values; the forces,
    // which the DEM pass adds to, are saved and restored), fastest kept.
    int autotune(int reps, int *lanes) override {
        HCHK(hipSetDevice(device));
        if (std::is_same<R, double>::value) {
            if (lanes) lanes[0] = lanes[1] = 1;
            return DMPH_OK;
        }
        reps = reps < 1 ? 1 : reps;
        const size_t n = g.n;
        g.fuse_accel = 0;     // the re-runs must not advance the fluid velocities
        R *save = a.t_real;   // rebin scratch, free between steps
        HCHK(hipMemcpyAsync(save, a.fx, n * sizeof(R), hipMemcpyDeviceToDevice, stream));
        HCHK(hipMemcpyAsync(save + n, a.fy, n * sizeof(R), hipMemcpyDeviceToDevice, stream));
        HCHK(hipMemcpyAsync(save + 2 * n, a.fz, n * sizeof(R), hipMemcpyDeviceToDevice, stream));
        hipEvent_t e0 = take_event(), e1 = take_event();
        if (!e0 || !e1) { err = "hipEventCreate failed"; return DMPH_E_DEVICE; }
        const int cand[4] = {2, 4, 8, 16}, bcand[3] = {64, 128, 256};
        const int orig[2] = {g.lanes[0], g.lanes[1]}, borig[2] = {g.block[0], g.block[1]};
        int picks[2] = {orig[0], orig[1]}, bpicks[2] = {borig[0], borig[1]};
        // every candidate timed in two rounds, the rounds interleaved, and
        // scored by the faster of its two: one slow round (a clock or cache
        // transient) does not decide a pick
        for (int w = 0; w < 2; w++) {
            for (int c = 0; c < 4; c++)
                for (int bc = 0; bc < 3; bc++) tune_ms[w][c][bc] = 1e30f;
            for (int round = 0; round < 2; round++)
                for (int c = 0; c < 4; c++)
                    for (int bc = 0; bc < 3; bc++) {
                        g.lanes[w] = cand[c];
                        g.block[w] = bcand[bc];
                        auto launch = [&] {
                            if (w == 0) launch_sweep1<R>(g, a, stream);
                            else launch_sweep2<R>(g, a, stream);
                        };
                        launch();
                        HCHK(hipEventRecord(e0, stream));
                        for (int r = 0; r < reps; r++) launch();
                        HCHK(hipEventRecord(e1, stream));
                        HCHK(hipEventSynchronize(e1));
                        float ms = 0.f;
                        HCHK(hipEventElapsedTime(&ms, e0, e1));
                        tune_ms[w][c][bc] = std::min(tune_ms[w][c][bc], ms / reps);
                    }
            float best = 1e30f;
            for (int c = 0; c < 4; c++)
                for (int bc = 0; bc < 3; bc++)
                    if (tune_ms[w][c][bc] < best) { best = tune_ms[w][c][bc]; picks[w] = cand[c]; bpicks[w] = bcand[bc]; }
            g.lanes[w] = orig[w];
            g.block[w] = borig[w];
        }
        // lanes per wall row in pass A (the fused passes' wall rows walk ~30
        // entries between full sweeps, a fluid row ~140), at the lane count and
        // workgroup size just picked for the fluid rows
        const int worig = g.lanes_w;
        int wpick = worig;
        if (g.fused && !g.st_active && g.we > g.wb) {
            const int wcand[3] = {1, 2, 4};
            float wms[3] = {1e30f, 1e30f, 1e30f};
            g.lanes[0] = picks[0];
            g.block[0] = bpicks[0];
            for (int round = 0; round < 2; round++)
                for (int c = 0; c < 3; c++) {
                    g.lanes_w = wcand[c];
                    launch_sweep1<R>(g, a, stream);
                    HCHK(hipEventRecord(e0, stream));
                    for (int r = 0; r < reps; r++) launch_sweep1<R>(g, a, stream);
                    HCHK(hipEventRecord(e1, stream));
                    HCHK(hipEventSynchronize(e1));
                    float ms = 0.f;
                    HCHK(hipEventElapsedTime(&ms, e0, e1));
                    wms[c] = std::min(wms[c], ms);
                }
            int bw = 0;
            for (int c = 1; c < 3; c++)
                if (wms[c] < wms[bw]) bw = c;
            wpick = wcand[bw];
            g.lanes[0] = orig[0];
            g.block[0] = borig[0];
            g.lanes_w = worig;
        }
        // the sweeps' outputs as the state's own lane counts computed them (the
        // candidates sum in other orders), then the forces as the DEM left them
        launch_sweep1<R>(g, a, stream);
        launch_sweep2<R>(g, a, stream);
        g.lanes[0] = picks[0];
        g.lanes[1] = picks[1];
        g.block[0] = bpicks[0];
        g.block[1] = bpicks[1];
        g.lanes_w = wpick;
        ev_pool.push_back(e0);
        ev_pool.push_back(e1);
        HCHK(hipMemcpyAsync(a.fx, save, n * sizeof(R), hipMemcpyDeviceToDevice, stream));
        HCHK(hipMemcpyAsync(a.fy, save + n, n * sizeof(R), hipMemcpyDeviceToDevice, stream));
        HCHK(hipMemcpyAsync(a.fz, save + 2 * n, n * sizeof(R), hipMemcpyDeviceToDevice, stream));
        g.fuse_accel = 1;
        g.fused = getenv("DMPH_SWEEPS") && !strcmp(getenv("DMPH_SWEEPS"), "separate") ? 0 : 1;
        HCHK(hipStreamSynchronize(stream));
        HCHK(hipGetLastError());
        // launch arguments changed: capture the step again at the next call
        drop_graph();
        if (lanes) { lanes[0] = g.lanes[0]; lanes[1] = g.lanes[1]; }
        return DMPH_OK;
    }
    float tune_ms[2][4][3] = {};

    int step(int nsteps) override {
        HCHK(hipSetDevice(device));
        // per-group event timing needs the launches one by one: eager then
        const bool use_graph = o.graph && !timing;
        if (use_graph) {
            // both graphs at the first call, whatever its count, so that a later
            // call of another length replays without capturing (capture only
            // records: nothing runs until a launch)
            if (nsteps > 0 && !graph_exec) TRY(capture_steps(1, graph, graph_exec));
            if (nsteps > 0 && !graph_exec_k) TRY(capture_steps(kGraphSteps, graph_k, graph_exec_k));
            int left = nsteps;
            for (; left >= kGraphSteps; left -= kGraphSteps) HCHK(hipGraphLaunch(graph_exec_k, stream));
            for (; left > 0; left--) HCHK(hipGraphLaunch(graph_exec, stream));
        } else {
            for (int k = 0; k < nsteps; k++) {
                TRY(enqueue_step());
                if (timing && pending.size() > 4096) TRY(resolve_timing());
            }
        }
        HCHK(hipGetLastError());
        return DMPH_OK;
    }

    int sync() override {
        HCHK(hipSetDevice(device));
        HCHK(hipStreamSynchronize(stream));
        return check_overflow();
    }

    int check_overflow() {
        int ov = 0;
        HCHK(hipMemcpy(&ov, &a.sc->list_overflow, sizeof ov, hipMemcpyDeviceToHost));
        if (ov > 0) {
            char b[200];
            snprintf(b, sizeof b, "neighbour list capacity %d (near segment %d) exceeded by a row of %d entries "
                     "(set DMPH_NEIGHBOR_CAPACITY); results after that rebinning are invalid", g.kcap, g.knear, ov);
            err = b;
            return DMPH_E_STATE;
        }
        return DMPH_OK;
    }

    int time(double *t) override {
        Scalars sc;
        HCHK(hipMemcpyAsync(&sc, a.sc, sizeof sc, hipMemcpyDeviceToHost, stream));
        HCHK(hipStreamSynchronize(stream));
        *t = sc.time;
        return DMPH_OK;
    }

    template <class T>
    int download(std::vector<T> &h, const T *d, size_t count) {
        h.resize(count);
        HCHK(hipMemcpyAsync(h.data(), d, count * sizeof(T), hipMemcpyDeviceToHost, stream));
        HCHK(hipStreamSynchronize(stream));
        return DMPH_OK;
    }

    // positions as absolute coordinates (FP32 offsets folded with their cell origin, wrapped into the domain)
    int positions(std::vector<double> &out) {
        const int n = g.n;
        std::vector<R> x, y, z;
        TRY(download(x, a.px, n)); TRY(download(y, a.py, n)); TRY(download(z, a.pz, n));
        out.assign(3 * (size_t)n, 0.0);
        if (std::is_same<R, double>::value) {
            for (int i = 0; i < n; i++) { out[3 * i] = x[i]; out[3 * i + 1] = y[i]; out[3 * i + 2] = z[i]; }
            return DMPH_OK;
        }
        std::vector<int> key;
        TRY(download(key, a.key, n));
        for (int i = 0; i < n; i++) {
            const int cell = key[i] % g.ncell;
            const int c[3] = {cell / (g.cc[1] * g.cc[2]), (cell / g.cc[2]) % g.cc[1], cell % g.cc[2]};
            const R o_[3] = {x[i], y[i], z[i]};
            for (int k = 0; k < 3; k++) {
                const double v = c[k] * g.cw[k] + (double)o_[k];
                out[3 * i + k] = modw(v, g.dw[k]) + g.dmin[k];
            }
        }
        return DMPH_OK;
    }

    int real_field(int field, std::vector<double> &out, int &comps) {
        const int n = g.n;
        comps = 1;
        auto one = [&](const R *src) -> int {
            std::vector<R> h;
            TRY(download(h, src, n));
            out.assign(h.begin(), h.end());
            return DMPH_OK;
        };
        auto three = [&](const R *x, const R *y, const R *z) -> int {
            std::vector<R> hx, hy, hz;
            TRY(download(hx, x, n)); TRY(download(hy, y, n)); TRY(download(hz, z, n));
            out.resize(3 * (size_t)n);
            for (int i = 0; i < n; i++) { out[3 * i] = hx[i]; out[3 * i + 1] = hy[i]; out[3 * i + 2] = hz[i]; }
            comps = 3;
            return DMPH_OK;
        };
        switch (field) {
            case DMPH_FIELD_POSITION: comps = 3; return positions(out);
            case DMPH_FIELD_VELOCITY: return three(a.vx, a.vy, a.vz);
            case DMPH_FIELD_FORCE: return three(a.fx, a.fy, a.fz);
            case DMPH_FIELD_GRAVITY_CENTER: return three(a.gcx, a.gcy, a.gcz);
            case DMPH_FIELD_DENSITY_A: return one(a.dens_a);
            case DMPH_FIELD_PRESSURE_A: return one(a.press_a);
            case DMPH_FIELD_VOL_STRAIN_P: return one(a.vs_p);
            case DMPH_FIELD_DIVERGENCE_P: return one(a.div_p);
            case DMPH_FIELD_PRESSURE_P: return one(a.press_p);
            case DMPH_FIELD_MASS: return one(a.mass);
            case DMPH_FIELD_ALPHA: return one(a.alpha);
            case DMPH_FIELD_DRAG_C: return one(a.drag_c);
            case DMPH_FIELD_KN: return one(a.kn);
            case DMPH_FIELD_ETAN: return one(a.etan);
            case DMPH_FIELD_KT: return one(a.kt);
            case DMPH_FIELD_ETAT: return one(a.etat);
            case DMPH_FIELD_VIRIAL_PRESSURE: return one(a.virial_p);
            case DMPH_FIELD_KAPPA: return one(a.kappa);
            default: err = "unknown field"; return DMPH_E_ARG;
        }
    }

    int get_field(int field, int order, void *dst) override {
        HCHK(hipSetDevice(device));
        if (!dst || (order != DMPH_ORDER_ORIGINAL && order != DMPH_ORDER_SLOT)) { err = "bad argument"; return DMPH_E_ARG; }
        const int n = g.n;
        std::vector<int> id;
        if (order == DMPH_ORDER_ORIGINAL) TRY(download(id, a.id, n));
        const int *ifield = nullptr;
        switch (field) {
            case DMPH_FIELD_TYPE: ifield = a.type; break;
            case DMPH_FIELD_ID: ifield = a.id; break;
            case DMPH_FIELD_NEIGHBOR_COUNT: ifield = a.ncount; break;
            case DMPH_FIELD_SOLID_NEIGHBOR_COUNT: ifield = a.scount; break;
            case DMPH_FIELD_CELL_KEY: ifield = a.key; break;
            case DMPH_FIELD_BODY: ifield = a.body; break;
            default: break;
        }
        if (ifield) {
            std::vector<int> h;
            TRY(download(h, ifield, n));
            int *o_ = (int *)dst;
            for (int s = 0; s < n; s++) o_[order == DMPH_ORDER_ORIGINAL ? id[s] : s] = h[s];
            return DMPH_OK;
        }
        std::vector<double> v;
        int comps = 1;
        TRY(real_field(field, v, comps));
        double *o_ = (double *)dst;
        for (int s = 0; s < n; s++) {
            const int t = order == DMPH_ORDER_ORIGINAL ? id[s] : s;
            for (int k = 0; k < comps; k++) o_[(size_t)comps * t + k] = v[(size_t)comps * s + k];
        }
        return DMPH_OK;
    }

    // Overwrite positions or velocities from host memory; positions are wrapped
    // into the periodic domain and the next step rebins unconditionally.
    int set_field(int field, int order, const void *src) override {
        HCHK(hipSetDevice(device));
        if (!src || (order != DMPH_ORDER_ORIGINAL && order != DMPH_ORDER_SLOT)) { err = "bad argument"; return DMPH_E_ARG; }
        if (field != DMPH_FIELD_POSITION && field != DMPH_FIELD_VELOCITY) {
            err = "only DMPH_FIELD_POSITION and DMPH_FIELD_VELOCITY can be set";
            return DMPH_E_ARG;
        }
        const int n = g.n;
        const double *x = (const double *)src;
        std::vector<int> id, key;
        TRY(download(id, a.id, n));
        std::vector<R> c0(n), c1(n), c2(n);
        if (field == DMPH_FIELD_POSITION && !std::is_same<R, double>::value) TRY(download(key, a.key, n));
        for (int s = 0; s < n; s++) {
            const double *v = x + 3 * (size_t)(order == DMPH_ORDER_ORIGINAL ? id[s] : s);
            double w[3] = {v[0], v[1], v[2]};
            if (field == DMPH_FIELD_POSITION) {
                for (int k = 0; k < 3; k++) w[k] = modw(w[k] - g.dmin[k], g.dw[k]) + g.dmin[k];
                if (!std::is_same<R, double>::value) {   // offset from the origin of the slot's current cell
                    const int cell = key[s] % g.ncell;
                    const int co[3] = {cell / (g.cc[1] * g.cc[2]), (cell / g.cc[2]) % g.cc[1], cell % g.cc[2]};
                    for (int k = 0; k < 3; k++) {
                        const double o_ = w[k] - g.dmin[k] - co[k] * g.cw[k];
                        w[k] = modw(o_ + 0.5 * g.dw[k], g.dw[k]) - 0.5 * g.dw[k];
                    }
                }
            }
            c0[s] = (R)w[0]; c1[s] = (R)w[1]; c2[s] = (R)w[2];
        }
        R *d0 = field == DMPH_FIELD_POSITION ? a.px : a.vx, *d1 = field == DMPH_FIELD_POSITION ? a.py : a.vy,
          *d2 = field == DMPH_FIELD_POSITION ? a.pz : a.vz;
        HCHK(hipMemcpyAsync(d0, c0.data(), n * sizeof(R), hipMemcpyHostToDevice, stream));
        HCHK(hipMemcpyAsync(d1, c1.data(), n * sizeof(R), hipMemcpyHostToDevice, stream));
        HCHK(hipMemcpyAsync(d2, c2.data(), n * sizeof(R), hipMemcpyHostToDevice, stream));
        if (field == DMPH_FIELD_POSITION) {
            const int one = 1;
            HCHK(hipMemcpyAsync(&a.sc->force, &one, sizeof one, hipMemcpyHostToDevice, stream));
        }
        {   // the next step's sweep 1 covers every particle (walls may have been moved or set in motion)
            const int two = 2, zero = 0;
            HCHK(hipMemcpyAsync(&a.sc->full_sweep, &two, sizeof two, hipMemcpyHostToDevice, stream));
            HCHK(hipMemcpyAsync(&a.sc->walls_rest, &zero, sizeof zero, hipMemcpyHostToDevice, stream));
        }
        HCHK(hipStreamSynchronize(stream));   // host vectors go out of scope
        return DMPH_OK;
    }

    int export_state(float *pos, float *vel) override {
        HCHK(hipSetDevice(device));
        launch_export<R>(g, a, pos, vel, stream);
        HCHK(hipGetLastError());
        HCHK(hipStreamSynchronize(stream));
        return DMPH_OK;
    }

    int get_bodies(double *x, double *v, double *q, double *l, double *m) override {
        HCHK(hipSetDevice(device));
        const size_t nb = g.nb;
        if (!nb) return DMPH_OK;
        if (x) HCHK(hipMemcpyAsync(x, a.b_x, 3 * nb * sizeof(double), hipMemcpyDeviceToHost, stream));
        if (v) HCHK(hipMemcpyAsync(v, a.b_v, 3 * nb * sizeof(double), hipMemcpyDeviceToHost, stream));
        if (q) HCHK(hipMemcpyAsync(q, a.b_q, 4 * nb * sizeof(double), hipMemcpyDeviceToHost, stream));
        if (l) HCHK(hipMemcpyAsync(l, a.b_l, 3 * nb * sizeof(double), hipMemcpyDeviceToHost, stream));
        HCHK(hipStreamSynchronize(stream));
        if (m) memcpy(m, body_mass_host.data(), nb * sizeof(double));
        return DMPH_OK;
    }

    int info(dmph_info *out) override {
        memset(out, 0, sizeof *out);
        out->n = g.n; out->dim = g.dim; out->cells = g.ncell; out->nbodies = g.nb; out->substeps = g.substeps;
        out->precision = std::is_same<R, double>::value ? DMPH_PREC_FP64 : DMPH_PREC_FP32;
        for (int k = 0; k < 3; k++) { out->cell_count[k] = g.cc[k]; out->cell_width[k] = g.cw[k]; }
        out->fluid_begin = g.fb; out->fluid_end = g.fe; out->solid_begin = g.sb; out->solid_end = g.se;
        out->wall_begin = g.wb; out->wall_end = g.we;
        out->n0a = g.n0a; out->n0p = g.n0p; out->max_radius = g.rmax;
        out->nscenes = g.nscene;
        out->lanes[0] = std::is_same<R, double>::value ? 1 : g.lanes[0];
        out->lanes[1] = std::is_same<R, double>::value ? 1 : g.lanes[1];
        out->block[0] = std::is_same<R, double>::value ? 256 : g.block[0];
        out->block[1] = std::is_same<R, double>::value ? 256 : g.block[1];
        out->cell_scale = cell_scale;
        out->wall_lanes = std::is_same<R, double>::value ? 1 : g.lanes_w;
        Scalars sc;
        HCHK(hipMemcpyAsync(&sc, a.sc, sizeof sc, hipMemcpyDeviceToHost, stream));
        HCHK(hipStreamSynchronize(stream));
        out->rebuilds = sc.rebuilds;
        return DMPH_OK;
    }

    int count_nonfinite(int field, int64_t *count) override {
        HCHK(hipSetDevice(device));
        const R *p[3] = {nullptr, nullptr, nullptr};
        switch (field) {
            case DMPH_FIELD_POSITION: p[0] = a.px; p[1] = a.py; p[2] = a.pz; break;
            case DMPH_FIELD_VELOCITY: p[0] = a.vx; p[1] = a.vy; p[2] = a.vz; break;
            case DMPH_FIELD_FORCE: p[0] = a.fx; p[1] = a.fy; p[2] = a.fz; break;
            case DMPH_FIELD_GRAVITY_CENTER: p[0] = a.gcx; p[1] = a.gcy; p[2] = a.gcz; break;
            case DMPH_FIELD_DENSITY_A: p[0] = a.dens_a; break;
            case DMPH_FIELD_PRESSURE_A: p[0] = a.press_a; break;
            case DMPH_FIELD_VOL_STRAIN_P: p[0] = a.vs_p; break;
            case DMPH_FIELD_DIVERGENCE_P: p[0] = a.div_p; break;
            case DMPH_FIELD_PRESSURE_P: p[0] = a.press_p; break;
            case DMPH_FIELD_MASS: p[0] = a.mass; break;
            case DMPH_FIELD_ALPHA: p[0] = a.alpha; break;
            case DMPH_FIELD_DRAG_C: p[0] = a.drag_c; break;
            case DMPH_FIELD_KN: p[0] = a.kn; break;
            case DMPH_FIELD_ETAN: p[0] = a.etan; break;
            case DMPH_FIELD_KT: p[0] = a.kt; break;
            case DMPH_FIELD_ETAT: p[0] = a.etat; break;
            case DMPH_FIELD_VIRIAL_PRESSURE: p[0] = a.virial_p; break;
            case DMPH_FIELD_KAPPA: p[0] = a.kappa; break;
            case DMPH_FIELD_TYPE: case DMPH_FIELD_ID: case DMPH_FIELD_NEIGHBOR_COUNT:
            case DMPH_FIELD_SOLID_NEIGHBOR_COUNT: case DMPH_FIELD_CELL_KEY: case DMPH_FIELD_BODY:
                *count = 0;
                return DMPH_OK;
            default: err = "unknown field"; return DMPH_E_ARG;
        }
        HCHK(hipMemsetAsync(a.nonfinite, 0, sizeof(unsigned long long), stream));
        launch_count_nonfinite<R>(p[0], p[1], p[2], g.n, a.nonfinite, stream);
        HCHK(hipGetLastError());
        unsigned long long c = 0;
        HCHK(hipMemcpyAsync(&c, a.nonfinite, sizeof c, hipMemcpyDeviceToHost, stream));
        HCHK(hipStreamSynchronize(stream));
        *count = (int64_t)c;
        return DMPH_OK;
    }

    int virial() override {
        HCHK(hipSetDevice(device));
        launch_virial<R>(g, a, stream);
        HCHK(hipGetLastError());
        return DMPH_OK;
    }

    int pairs(int64_t *nnz, int64_t *ptr, int *idx) override {
        HCHK(hipSetDevice(device));
        TRY(sync());
        const int n = g.n;
        std::vector<int> cnt;
        TRY(download(cnt, a.ncount, n));
        std::vector<int64_t> pt(n + 1, 0);
        for (int i = 0; i < n; i++) pt[i + 1] = pt[i] + cnt[i];
        if (nnz) *nnz = pt[n];
        if (ptr) memcpy(ptr, pt.data(), sizeof(int64_t) * (n + 1));
        if (!idx) return DMPH_OK;
        std::vector<uint32_t> list;
        TRY(download(list, a.nlist, (size_t)g.kcap * n));
        std::vector<int> nn(n);
        if (std::is_same<R, double>::value) {
            for (int i = 0; i < n; i++) nn[i] = cnt[i];
        } else {
            TRY(download(nn, a.nnear, n));
        }
        // FP32 rows: near segment from position 0, far segment from knear
        for (int i = 0; i < n; i++)
            for (int k = 0; k < cnt[i]; k++) {
                const int q = k < nn[i] ? k : g.knear + (k - nn[i]);
                idx[pt[i] + k] = (int)(list[((size_t)(q >> 3) * n + i) * 8 + (q & 7)] >> 6);
            }
        return DMPH_OK;
    }

    int slot_fields(std::vector<int> &type, std::vector<double> &pos, std::vector<double> &vel) {
        TRY(download(type, a.type, g.n));
        TRY(positions(pos));
        int c;
        TRY(real_field(DMPH_FIELD_VELOCITY, vel, c));
        return DMPH_OK;
    }

    int write_prof(const char *path) override {
        std::vector<int> type;
        std::vector<double> pos, vel;
        TRY(slot_fields(type, pos, vel));
        double t;
        TRY(time(&t));
        const double dmin[3] = {g.dmin[0], g.dmin[1], g.dmin[2]};
        const int rc = dmph::write_prof(path, t, g.n, spacing, dmin, dmax, type.data(), pos.data(), vel.data());
        if (rc) err = std::string("cannot write ") + path;
        return rc;
    }

    int write_vtk(const char *path) override {
        std::vector<int> type, nc;
        std::vector<double> pos, vel, kn, en, kt, et, al, dr, vs, vi;
        TRY(slot_fields(type, pos, vel));
        TRY(download(nc, a.ncount, g.n));
        int c;
        TRY(real_field(DMPH_FIELD_KN, kn, c)); TRY(real_field(DMPH_FIELD_ETAN, en, c));
        TRY(real_field(DMPH_FIELD_KT, kt, c)); TRY(real_field(DMPH_FIELD_ETAT, et, c));
        TRY(real_field(DMPH_FIELD_ALPHA, al, c)); TRY(real_field(DMPH_FIELD_DRAG_C, dr, c));
        TRY(real_field(DMPH_FIELD_VOL_STRAIN_P, vs, c)); TRY(real_field(DMPH_FIELD_VIRIAL_PRESSURE, vi, c));
        VtkFields f{g.n, type.data(), nc.data(), pos.data(), vel.data(), kn.data(), en.data(), kt.data(), et.data(),
                    al.data(), dr.data(), vs.data(), vi.data()};
        const int rc = dmph::write_vtk(path, f);
        if (rc) err = std::string("cannot write ") + path;
        return rc;
    }
};

}  // namespace dmph

// ---------------------------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------------------------

static thread_local std::string g_create_error;

extern "C" {

void dmph_default_options(dmph_options *opt) {
    memset(opt, 0, sizeof *opt);
    opt->dim = 0;
    opt->precision = DMPH_PREC_FP64;
    opt->sort_mode = DMPH_SORT_CANONICAL;
    opt->rebuild = DMPH_REBUILD_SKIN;
    opt->debris = DMPH_DEBRIS_PARTICLES;
    opt->device = 0;
    opt->compensated = 1;
    opt->graph = 1;
}

// One handle of the options' precision over grid g (scene_of / nscene: a batch).
static int make_handle(const dmph_params *p, const dmph_options &o, const dmph_grid *g, const int *scene_of,
                       int nscene, std::unique_ptr<dmph_handle> &h) {
    int rc;
    if (o.precision == DMPH_PREC_FP32) {
        auto e = std::make_unique<dmph::Engine<float>>();
        rc = e->create(p, &o, g, scene_of, nscene);
        h = std::move(e);
    } else if (o.precision == DMPH_PREC_FP64) {
        auto e = std::make_unique<dmph::Engine<double>>();
        rc = e->create(p, &o, g, scene_of, nscene);
        h = std::move(e);
    } else {
        g_create_error = "unknown precision";
        return DMPH_E_ARG;
    }
    if (rc != DMPH_OK) g_create_error = h->err;
    return rc;
}

// Cell-size tuner (FP32, dmph_options.cell_scale < 0): each candidate width is
// tried on a handle of its own -- created, lanes and workgroups autotuned, a
// few steps of warm-up -- timed over replayed steps between rebuilds and over
// steps that rebin every step, and scored as the step time plus the extra
// time of a rebuild step spread over kTuneInterval steps (the dam breaks
// rebuild every ~20 steps).  The candidates do not change which pairs a row
// holds, only the cell grid, the slot order and the list-build cost.
static const double kTuneCells[] = {1.0, 1.25, 1.5};
constexpr int kTuneSteps = 20, kTuneRebuilds = 4, kTuneInterval = 20;

static int tune_cells(const dmph_params *p, dmph_options o, const dmph_grid *g, const int *scene_of, int nscene,
                      double *best) {
    using clk = std::chrono::steady_clock;
    auto secs = [](clk::time_point a, clk::time_point b) { return std::chrono::duration<double>(b - a).count(); };
    double best_score = 1e300;
    *best = 1.0;
    for (double c : kTuneCells) {
        o.cell_scale = c;
        auto e = std::make_unique<dmph::Engine<float>>();
        int rc = e->create(p, &o, g, scene_of, nscene);
        if (rc == DMPH_OK) rc = e->step(5);
        if (rc == DMPH_OK) rc = e->sync();
        if (rc == DMPH_OK) rc = e->autotune(2, nullptr);
        if (rc == DMPH_OK) rc = e->step(2);
        if (rc == DMPH_OK) rc = e->sync();
        const auto t0 = clk::now();
        if (rc == DMPH_OK) rc = e->step(kTuneSteps);
        if (rc == DMPH_OK) rc = e->sync();
        const auto t1 = clk::now();
        e->g.rebuild_every = 1;   // the same launches, every step a rebinning step
        e->drop_graph();
        if (rc == DMPH_OK) rc = e->step(1);
        if (rc == DMPH_OK) rc = e->sync();
        const auto t2 = clk::now();
        if (rc == DMPH_OK) rc = e->step(kTuneRebuilds);
        if (rc == DMPH_OK) rc = e->sync();
        const auto t3 = clk::now();
        if (rc != DMPH_OK) {
            g_create_error = "cell-size tuning at cell_scale " + std::to_string(c) + ": " + e->err;
            return rc;
        }
        const double step = secs(t0, t1) / kTuneSteps, rebuild = secs(t2, t3) / kTuneRebuilds;
        const double score = step + std::max(0.0, rebuild - step) / kTuneInterval;
        if (score < best_score) { best_score = score; *best = c; }
    }
    return DMPH_OK;
}

// dmph_create and dmph_create_batch: the handle, with the cell size tuned first when asked
static int create_handle(const dmph_params *p, const dmph_options *opt, const dmph_grid *g, const int *scene_of,
                         int nscene, dmph_handle **out) {
    dmph_options o;
    if (opt) o = *opt;
    else dmph_default_options(&o);
    if (o.cell_scale < 0.0) {
        double c = 1.0;   // FP64 keeps the reference's cell grid
        if (o.precision == DMPH_PREC_FP32) {
            const int rc = tune_cells(p, o, g, scene_of, nscene, &c);
            if (rc != DMPH_OK) return rc;
        }
        o.cell_scale = c;
    }
    std::unique_ptr<dmph_handle> h;
    const int rc = make_handle(p, o, g, scene_of, nscene, h);
    if (rc != DMPH_OK) return rc;
    *out = h.release();
    return DMPH_OK;
}

int dmph_create(const dmph_params *p, const dmph_options *opt, const dmph_grid *g, dmph_handle **out) {
    if (!p || !g || !out) return DMPH_E_ARG;
    *out = nullptr;
    return create_handle(p, opt, g, nullptr, 1, out);
}

int dmph_create_batch(const dmph_params *p, const dmph_options *opt, const dmph_grid *grids, int nscenes,
                      dmph_handle **out) {
    if (!p || !grids || nscenes <= 0 || !out) return DMPH_E_ARG;
    *out = nullptr;
    if (nscenes == 1) return dmph_create(p, opt, grids, out);
    const dmph_grid &g0 = grids[0];
    int64_t n = 0;
    bool bodies = false, dens = false;
    for (int s = 0; s < nscenes; s++) {
        const dmph_grid &gs = grids[s];
        if (gs.n <= 0 || !gs.type || !gs.pos || !gs.vel) { g_create_error = "empty or invalid scene"; return DMPH_E_ARG; }
        if (gs.spacing != g0.spacing || gs.time != g0.time) {
            g_create_error = "scenes of a batch need the same spacing and time";
            return DMPH_E_ARG;
        }
        for (int k = 0; k < 3; k++)
            if (gs.domain_min[k] != g0.domain_min[k] || gs.domain_max[k] != g0.domain_max[k]) {
                g_create_error = "scenes of a batch need the same domain";
                return DMPH_E_ARG;
            }
        n += gs.n;
        bodies = bodies || gs.body;
        dens = dens || gs.density;
    }
    if (n >= (1 << 26)) { g_create_error = "batch larger than 2^26 particles"; return DMPH_E_ARG; }
    // one combined particle set, scene after scene; body ids offset per scene
    std::vector<int> type(n), body(bodies ? n : 0), scene(n);
    std::vector<double> pos(3 * n), vel(3 * n), rho(dens ? n : 0);
    int64_t off = 0;
    int boff = 0;
    for (int s = 0; s < nscenes; s++) {
        const dmph_grid &gs = grids[s];
        int nb = 0;
        for (int i = 0; i < gs.n; i++) {
            type[off + i] = gs.type[i];
            scene[off + i] = s;
            for (int k = 0; k < 3; k++) { pos[3 * (off + i) + k] = gs.pos[3 * i + k]; vel[3 * (off + i) + k] = gs.vel[3 * i + k]; }
            if (bodies) {
                const int b = gs.body ? gs.body[i] : -1;
                body[off + i] = b >= 0 ? b + boff : -1;
                nb = b + 1 > nb ? b + 1 : nb;
            }
            if (dens) rho[off + i] = gs.density ? gs.density[i] : -1.0;
        }
        off += gs.n;
        boff += nb;
    }
    dmph_grid all = g0;
    all.n = (int)n;
    all.type = type.data();
    all.pos = pos.data();
    all.vel = vel.data();
    all.body = bodies ? body.data() : nullptr;
    all.density = dens ? rho.data() : nullptr;
    return create_handle(p, opt, &all, scene.data(), nscenes, out);
}

int dmph_step(dmph_handle *h, int nsteps) { return h && nsteps >= 0 ? h->step(nsteps) : DMPH_E_ARG; }
int dmph_sync(dmph_handle *h) { return h ? h->sync() : DMPH_E_ARG; }
int dmph_export_state(dmph_handle *h, float *pos, float *vel) {
    return h && (pos || vel) ? h->export_state(pos, vel) : DMPH_E_ARG;
}
int dmph_autotune(dmph_handle *h, int reps, int *lanes) { return h ? h->autotune(reps, lanes) : DMPH_E_ARG; }
int dmph_set_field(dmph_handle *h, int field, int order, const void *src) {
    return h ? h->set_field(field, order, src) : DMPH_E_ARG;
}
int dmph_time(dmph_handle *h, double *t) { return h && t ? h->time(t) : DMPH_E_ARG; }
int dmph_get_field(dmph_handle *h, int field, int order, void *dst) {
    return h ? h->get_field(field, order, dst) : DMPH_E_ARG;
}
int dmph_get_bodies(dmph_handle *h, double *x, double *v, double *q, double *l, double *mass) {
    return h ? h->get_bodies(x, v, q, l, mass) : DMPH_E_ARG;
}
int dmph_get_info(dmph_handle *h, dmph_info *info) { return h && info ? h->info(info) : DMPH_E_ARG; }
int dmph_virial(dmph_handle *h) { return h ? h->virial() : DMPH_E_ARG; }
int dmph_count_nonfinite(dmph_handle *h, int field, int64_t *count) {
    return h && count ? h->count_nonfinite(field, count) : DMPH_E_ARG;
}
int dmph_pairs(dmph_handle *h, int64_t *nnz, int64_t *ptr, int *idx) { return h ? h->pairs(nnz, ptr, idx) : DMPH_E_ARG; }
int dmph_write_prof(dmph_handle *h, const char *path) { return h && path ? h->write_prof(path) : DMPH_E_ARG; }
int dmph_write_vtk(dmph_handle *h, const char *path) { return h && path ? h->write_vtk(path) : DMPH_E_ARG; }
int dmph_set_timing(dmph_handle *h, int enable) { return h ? h->set_timing(enable) : DMPH_E_ARG; }
int dmph_get_timing(dmph_handle *h, int group, double *total_ms, int64_t *launches) {
    return h ? h->get_timing(group, total_ms, launches) : DMPH_E_ARG;
}
const char *dmph_last_error(const dmph_handle *h) { return h ? h->err.c_str() : g_create_error.c_str(); }
void dmph_destroy(dmph_handle *h) { delete h; }

}  // extern "C"
