/* FP64 CPU oracle for the DEM-MPH explicit step.
 *
 * TEST INFRASTRUCTURE ONLY: a restatement of the reference algorithm in
 * source/DEM-MPH.cpp, used as the checker for the HIP product path (tests,
 * smoke check, CPU-baseline timing).  The product never links it.
 *
 * Every pass cites the reference lines it restates.  Arithmetic is written with
 * the same association as the reference so that, compiled without fused
 * multiply-add (oracle/Makefile), its FP64 results can be compared with the
 * reference's bit for bit.  Deliberate additions, all off by default:
 *   - ORC_SORT_CANONICAL: ascending original id inside each cell instead of the
 *     reference's bitonic permutation (the GPU path's fast counting sort);
 *   - ORC_REBUILD_EVERY_STEP: rebin each step instead of the 0.05 d skin test;
 *   - rigid bodies: solid particles with a body id move as one rigid body with
 *     FP64 force/torque sums and a quaternion/angular-momentum integrator (the
 *     reference has none; pinned by closed forms in the tests);
 *   - per-particle density (a .grid extra column), default Density[type].
 * Values the reference never initialises (malloc'd, source/DEM-MPH.cpp:704-799)
 * start at zero here.
 */
#include "dmph_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

typedef double vec3[3];

struct orc {
    orc_params p;
    orc_opts o;
    int n, dim, substeps;
    double d, vol, time;
    double dmin[3], dmax[3], dw[3];
    /* kernel constants (source/DEM-MPH.cpp:1007-1162) */
    double ra, rg, rp, rv, rmax, swa, swg, swp, swv, n0a, n0p, r2g, cofk, cofa[ORC_NTYPE];
    double wall_rot[ORC_NTYPE][3][3], wall_center[ORC_NTYPE][3];
    /* cell grid (:1207-1262) */
    int cc[3], ncell;
    double cw[3];
    /* per slot */
    int *id, *type, *body;
    double *rho;
    vec3 *pos, *vel, *force, *gc, *build_pos, *r0;
    double *mass, *dens_a, *press_a, *vs_p, *div_p, *press_p, *kappa, *lambda, *mu;
    double *alpha, *drag_c, *kn, *etan, *kt, *etat, *virial_p;
    double (*virial)[3][3];
    int *ncount, *fcount, *scount, *nind, *key, *cell_start;
    int64_t *nptr, nind_n;
    int fb, fe, sb, se, wb, we;
    /* rigid bodies */
    int nb;
    double *b_mass;
    vec3 *b_x, *b_v, *b_l;
    double (*b_q)[4], (*b_i)[3][3], (*b_iinv)[3][3];
    int64_t rebuilds;
    double scratch_scalars[24];
};

#define MODW(x, w) ((x) - (w) * floor((x) / (w)))   /* source/DEM-MPH.cpp:76 */
#define IS_FLUID(t) (0 <= (t) && (t) < 2)
#define IS_SOLID(t) (2 <= (t) && (t) < 4)
#define IS_WALL(t) (4 <= (t) && (t) < 6)

/* ------------------------------------------------------------------------- */
/* .data parsing, source/DEM-MPH.cpp:621-679                                  */
/* ------------------------------------------------------------------------- */

static int scan6(const char *buf, const char *key, double *v) {
    char fmt[160];
    snprintf(fmt, sizeof fmt, " %s %%lf %%lf %%lf %%lf %%lf %%lf", key);
    return sscanf(buf, fmt, &v[0], &v[1], &v[2], &v[3], &v[4], &v[5]) == 6;
}

int orc_parse_data(const char *path, orc_params *p) {
    FILE *fp = fopen(path, "r");
    if (!fp) return -1;
    char buf[1024];
    int recognised = 0;
    while (fgets(buf, sizeof buf, fp)) {
        int ok = 0;
        if (buf[0] == '#') continue;
        if (sscanf(buf, " Dt %lf", &p->dt) == 1) ok = 1;
        else if (sscanf(buf, " DEMDt %lf", &p->dem_dt) == 1) ok = 1;
        else if (sscanf(buf, " OutputInterval %lf", &p->output_interval) == 1) ok = 1;
        else if (sscanf(buf, " VtkOutputInterval %lf", &p->vtk_output_interval) == 1) ok = 1;
        else if (sscanf(buf, " EndTime %lf", &p->end_time) == 1) ok = 1;
        else if (sscanf(buf, " RadiusRatioA %lf", &p->ratio_a) == 1) ok = 1;
        else if (sscanf(buf, " RadiusRatioP %lf", &p->ratio_p) == 1) ok = 1;
        else if (sscanf(buf, " RadiusRatioV %lf", &p->ratio_v) == 1) ok = 1;
        else if (scan6(buf, "Density", p->density)) ok = 1;
        else if (scan6(buf, "BulkModulus", p->bulk_modulus)) ok = 1;
        else if (scan6(buf, "BulkViscosity", p->bulk_viscosity)) ok = 1;
        else if (scan6(buf, "ShearViscosity", p->shear_viscosity)) ok = 1;
        else if (scan6(buf, "SurfaceTension", p->surface_tension)) ok = 1;
        else {
            for (int t = 0; t < ORC_NTYPE && !ok; t++) {
                char k[40];
                snprintf(k, sizeof k, "InteractionRatio(Type%d)", t);
                if (scan6(buf, k, p->interaction[t])) ok = 1;
            }
            if (!ok && sscanf(buf, " Gravity %lf %lf %lf", &p->gravity[0], &p->gravity[1], &p->gravity[2]) == 3) ok = 1;
            else if (!ok && scan6(buf, "YoungModulus", p->young)) ok = 1;
            else if (!ok && scan6(buf, "PoissonRatio", p->poisson)) ok = 1;
            else if (!ok && scan6(buf, "RestitutionCoefficient", p->restitution)) ok = 1;
            else if (!ok && scan6(buf, "ViscousFrictionCoefficient", p->friction)) ok = 1;
            for (int w = 4; w < 6 && !ok; w++) {
                char fmt[200];
                snprintf(fmt, sizeof fmt,
                         " Wall%d Center %%lf %%lf %%lf Velocity %%lf %%lf %%lf Omega %%lf %%lf %%lf", w);
                if (sscanf(buf, fmt, &p->wall_center[w][0], &p->wall_center[w][1], &p->wall_center[w][2],
                           &p->wall_velocity[w][0], &p->wall_velocity[w][1], &p->wall_velocity[w][2],
                           &p->wall_omega[w][0], &p->wall_omega[w][1], &p->wall_omega[w][2]) == 9)
                    ok = 1;
            }
        }
        recognised += ok;
    }
    fclose(fp);
    return recognised;
}

/* ------------------------------------------------------------------------- */
/* kernel functions, source/DEM-MPH.cpp:272-342                               */
/* ------------------------------------------------------------------------- */

static inline double hdim(const orc *s, double h) { return s->dim == 2 ? h * h : h * h * h; }
static inline double w_a(const orc *s, double r, double h) {
    return 1.0 / s->swa * 1.0 / hdim(s, h) * (r / h) * (1.0 - (r / h)) * (1.0 - (r / h));
}
static inline double dw_a(const orc *s, double r, double h) {
    return 1.0 / s->swa * 1.0 / hdim(s, h) * (1.0 - (r / h)) * (1.0 - 3.0 * (r / h)) * (1.0 / h);
}
static inline double w_q(double sw, double hd, double r, double h) {   /* wg, wp, wv */
    return 1.0 / sw * 1.0 / hd * ((1.0 - r / h) * (1.0 - r / h));
}
static inline double dw_q(double sw, double hd, double r, double h) {  /* dwgdr, dwpdr, dwvdr */
    return 1.0 / sw * 1.0 / hd * (-2.0 / h * (1.0 - r / h));
}
#define W_G(s, r) w_q((s)->swg, hdim(s, (s)->rg), r, (s)->rg)
#define DW_G(s, r) dw_q((s)->swg, hdim(s, (s)->rg), r, (s)->rg)
#define W_P(s, r) w_q((s)->swp, hdim(s, (s)->rp), r, (s)->rp)
#define DW_P(s, r) dw_q((s)->swp, hdim(s, (s)->rp), r, (s)->rp)
#define DW_V(s, r) dw_q((s)->swv, hdim(s, (s)->rv), r, (s)->rv)

/* minimum-image separation x_j - x_i, as in every pair loop of the reference */
static inline void sep(const orc *s, int i, int j, double x[3]) {
    for (int a = 0; a < 3; a++)
        x[a] = MODW(s->pos[j][a] - s->pos[i][a] + 0.5 * s->dw[a], s->dw[a]) - 0.5 * s->dw[a];
}

/* ------------------------------------------------------------------------- */
/* initialisation                                                             */
/* ------------------------------------------------------------------------- */

/* initializeWeight, source/DEM-MPH.cpp:1007-1125 */
static void init_weight(orc *s) {
    const double d = s->d;
    s->ra = s->p.ratio_a * d;
    s->rg = s->p.ratio_a * d; /* RadiusRatioG = RadiusRatioA (:1009) */
    s->rp = s->p.ratio_p * d;
    s->rv = s->p.ratio_v * d;
    if (s->dim == 2) {
        s->swa = 1.0 / 2.0 * 2.0 / 15.0 * M_PI / d / d;
        s->swg = 1.0 / 2.0 * 1.0 / 3.0 * M_PI / d / d;
        s->swp = 1.0 / 2.0 * 1.0 / 3.0 * M_PI / d / d;
        s->swv = 1.0 / 2.0 * 1.0 / 3.0 * M_PI / d / d;
        s->r2g = 1.0 / 2.0 * 1.0 / 30.0 * M_PI * s->rg * s->rg / d / d / s->swg;
    } else {
        s->swa = 1.0 / 3.0 * 1.0 / 5.0 * M_PI / d / d / d;
        s->swg = 1.0 / 3.0 * 2.0 / 5.0 * M_PI / d / d / d;
        s->swp = 1.0 / 3.0 * 2.0 / 5.0 * M_PI / d / d / d;
        s->swv = 1.0 / 3.0 * 2.0 / 5.0 * M_PI / d / d / d;
        s->r2g = 1.0 / 3.0 * 4.0 / 105.0 * M_PI * s->rg * s->rg / d / d / d / s->swg;
    }
    /* reference number densities: inclusive lattice sums over k != 0 (:1032-1120) */
    for (int which = 0; which < 2; which++) {
        const double R = which == 0 ? s->ra : s->rp;
        const int range = (int)(R / d + 3.0);
        const int zr = s->dim == 2 ? 0 : range;
        double sum = 0.0;
        for (int ix = -range; ix <= range; ++ix)
            for (int iy = -range; iy <= range; ++iy)
                for (int iz = -zr; iz <= zr; ++iz) {
                    if (ix == 0 && iy == 0 && iz == 0) continue;
                    const double x = d * (double)ix, y = d * (double)iy, z = d * (double)iz;
                    const double r2 = s->dim == 2 ? x * x + y * y : x * x + y * y + z * z;
                    if (r2 <= R * R) {
                        const double r = sqrt(r2);
                        sum += which == 0 ? w_a(s, r, R) : W_P(s, r);
                    }
                }
        if (which == 0) s->n0a = sum; else s->n0p = sum;
    }
}

/* initializeFluid, source/DEM-MPH.cpp:1128-1162 (per-particle values are recomputed every step) */
static void init_fluid(orc *s) {
    double integ_n, integ_x;
    if (s->dim == 2) { s->cofk = 0.350778153; integ_n = 0.024679383; integ_x = 0.226126699; }
    else { s->cofk = 0.326976006; integ_n = 0.021425779; integ_x = 0.233977488; }
    for (int t = 0; t < ORC_NTYPE; t++)
        s->cofa[t] = s->p.surface_tension[t] / ((s->rg / s->d) * (integ_n + s->cofk * s->cofk * integ_x));
    /* file-order values; the first binning permutes types but not these arrays */
    for (int i = 0; i < s->n; i++) {
        s->mass[i] = s->rho[i] * s->vol;
        s->kappa[i] = s->p.bulk_modulus[s->type[i]];
        s->lambda[i] = s->p.bulk_viscosity[s->type[i]];
        s->mu[i] = s->p.shear_viscosity[s->type[i]];
    }
}

/* initializeWall, source/DEM-MPH.cpp:1166-1205: "rotation" from Omega, with the
 * reference's theta = |Omega|^2 and normal = Omega/theta. */
static void init_wall(orc *s) {
    for (int w = 4; w < 6; w++) {
        const double *om = s->p.wall_omega[w];
        double nrm[3] = {0.0, 0.0, 0.0}, q[4];
        const double theta = fabs(om[0] * om[0] + om[1] * om[1] + om[2] * om[2]);
        if (theta != 0.0)
            for (int a = 0; a < 3; a++) nrm[a] = om[a] / theta;
        q[0] = nrm[0] * sin(theta * s->p.dt / 2.0);
        q[1] = nrm[1] * sin(theta * s->p.dt / 2.0);
        q[2] = nrm[2] * sin(theta * s->p.dt / 2.0);
        q[3] = cos(theta * s->p.dt / 2.0);
        double (*R)[3] = s->wall_rot[w];
        R[0][0] = q[0] * q[0] - q[1] * q[1] - q[2] * q[2] + q[3] * q[3];
        R[0][1] = 2.0 * (q[0] * q[1] - q[2] * q[3]);
        R[0][2] = 2.0 * (q[0] * q[2] + q[1] * q[3]);
        R[1][0] = 2.0 * (q[0] * q[1] + q[2] * q[3]);
        R[1][1] = -q[0] * q[0] + q[1] * q[1] - q[2] * q[2] + q[3] * q[3];
        R[1][2] = 2.0 * (q[1] * q[2] - q[0] * q[3]);
        R[2][0] = 2.0 * (q[0] * q[2] - q[1] * q[3]);
        R[2][1] = 2.0 * (q[1] * q[2] + q[0] * q[3]);
        R[2][2] = -q[0] * q[0] - q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
        for (int a = 0; a < 3; a++) s->wall_center[w][a] = s->p.wall_center[w][a];
    }
}

/* initializeDomain, source/DEM-MPH.cpp:1207-1262 */
static int init_domain(orc *s) {
    double m = 0.0;
    m = s->ra > m ? s->ra : m;
    m = s->rg > m ? s->rg : m;
    m = s->rp > m ? s->rp : m;
    m = s->rv > m ? s->rv : m;
    s->rmax = m;
    const double margin = 0.1 * s->d;
    for (int a = 0; a < 3; a++) s->dw[a] = s->dmax[a] - s->dmin[a];
    double cnt[3];
    cnt[0] = floor((s->dmax[0] - s->dmin[0]) / (s->rmax + margin));
    cnt[1] = floor((s->dmax[1] - s->dmin[1]) / (s->rmax + margin));
    cnt[2] = s->dim == 2 ? 1 : floor((s->dmax[2] - s->dmin[2]) / (s->rmax + margin));
    for (int a = 0; a < 3; a++) s->cc[a] = (int)cnt[a];
    s->ncell = (int)(cnt[0] * cnt[1] * cnt[2]);
    for (int a = 0; a < 3; a++) s->cw[a] = s->dw[a] / s->cc[a];
    if (s->cc[0] < 3 || s->cc[1] < 3 || (s->dim == 3 && s->cc[2] < 3)) return -1;
    /* stencil half-width per axis must be one cell (the assert at :1540-1543) */
    for (int a = 0; a < s->dim; a++)
        if ((int)ceil((s->rmax + margin) / s->cw[a]) != 1) return -1;
    return 0;
}

/* ------------------------------------------------------------------------- */
/* binning: calculateCellParticle, source/DEM-MPH.cpp:1288-1519               */
/* ------------------------------------------------------------------------- */

static inline int wrapc(int c, int n) { return (c % n + n) % n; }

static int cell_key(const orc *s, int i) {
    int c[3];
    for (int a = 0; a < 3; a++) c[a] = ((int)floor((s->pos[i][a] - s->dmin[a]) / s->cw[a])) % s->cc[a];
    int k = wrapc(c[0], s->cc[0]) * s->cc[1] * s->cc[2] + wrapc(c[1], s->cc[1]) * s->cc[2] + wrapc(c[2], s->cc[2]);
    const int t = s->type[i];
    if (IS_SOLID(t)) k += s->ncell;
    else if (IS_WALL(t)) k += 2 * s->ncell;
    return k;
}

static void permute_int(int *a, const int *perm, int n, int *tmp) {
    for (int i = 0; i < n; i++) tmp[i] = a[perm[i]];
    memcpy(a, tmp, (size_t)n * sizeof(int));
}
static void permute_dbl(double *a, const int *perm, int n, double *tmp) {
    for (int i = 0; i < n; i++) tmp[i] = a[perm[i]];
    memcpy(a, tmp, (size_t)n * sizeof(double));
}
static void permute_vec(vec3 *a, const int *perm, int n, vec3 *tmp) {
    for (int i = 0; i < n; i++) memcpy(tmp[i], a[perm[i]], sizeof(vec3));
    memcpy(a, tmp, (size_t)n * sizeof(vec3));
}

static void bin_particles(orc *s) {
    const int n = s->n, T = s->ncell, nk = 3 * T;
    int *key = (int *)malloc(sizeof(int) * (size_t)(n + 1));
    int *perm = (int *)malloc(sizeof(int) * (size_t)(n + 1));
    for (int i = 0; i < n; i++) key[i] = cell_key(s, i);
    if (s->o.sort_mode == ORC_SORT_REFERENCE) {
        /* the reference's bitonic network over a power of two > n (:787-791, :1317-1338) */
        int pw = 0;
        while ((n >> pw) != 0) ++pw;
        const int p2 = 1 << pw;
        int *k = (int *)malloc(sizeof(int) * (size_t)p2), *v = (int *)malloc(sizeof(int) * (size_t)p2);
        for (int i = 0; i < p2; i++) { k[i] = i < n ? key[i] : nk; v[i] = i < n ? i : n; }
        for (int m = 0; m < pw; ++m)
            for (int sub = 0; sub <= m; ++sub) {
                const int dist = 1 << (m - sub);
                for (int i = 0; i < p2; ++i) {
                    const int up = ((i >> m) & 2) == 0;
                    if ((i & dist) == 0 && ((k[i] > k[i | dist]) == up)) {
                        int t = k[i]; k[i] = k[i | dist]; k[i | dist] = t;
                        t = v[i]; v[i] = v[i | dist]; v[i | dist] = t;
                    }
                }
            }
        for (int i = 0; i < n; i++) perm[i] = v[i];
        free(k); free(v);
    } else {
        /* counting sort by key, ties in ascending original id */
        int *cnt = (int *)calloc((size_t)nk + 1, sizeof(int)), *inv = (int *)malloc(sizeof(int) * (size_t)n);
        for (int i = 0; i < n; i++) { cnt[key[i] + 1]++; inv[s->id[i]] = i; }
        for (int k = 0; k < nk; k++) cnt[k + 1] += cnt[k];
        for (int id = 0; id < n; id++) { const int i = inv[id]; perm[cnt[key[i]]++] = i; }
        free(cnt); free(inv);
    }
    /* cell ranges: exclusive scan of per-key counts; equals the transition +
     * up/down-sweep construction of :1341-1452 */
    memset(s->cell_start, 0, sizeof(int) * (size_t)(nk + 1));
    for (int i = 0; i < n; i++) s->cell_start[key[i] + 1]++;
    for (int k = 0; k < nk; k++) s->cell_start[k + 1] += s->cell_start[k];
    for (int i = 0; i < n; i++) s->key[i] = key[perm[i]];
    s->fb = 0; s->fe = s->cell_start[T]; s->sb = s->cell_start[T]; s->se = s->cell_start[2 * T];
    s->wb = s->cell_start[2 * T]; s->we = s->cell_start[3 * T];
    /* gather-permute per-particle data (:1455-1518); derived per-slot arrays stay put */
    int *ti = (int *)malloc(sizeof(int) * (size_t)n);
    double *td = (double *)malloc(sizeof(double) * (size_t)n);
    vec3 *tv = (vec3 *)malloc(sizeof(vec3) * (size_t)n);
    permute_int(s->id, perm, n, ti);
    permute_int(s->type, perm, n, ti);
    permute_int(s->body, perm, n, ti);
    permute_dbl(s->rho, perm, n, td);
    permute_vec(s->pos, perm, n, tv);
    permute_vec(s->vel, perm, n, tv);
    permute_vec(s->r0, perm, n, tv);
    free(ti); free(td); free(tv); free(key); free(perm);
}

/* ------------------------------------------------------------------------- */
/* neighbour list: calculateNeighbor, source/DEM-MPH.cpp:1521-1851            */
/* ------------------------------------------------------------------------- */

/* Visit every j in i's 3^D stencil cells, block by block (fluid, solid, wall),
 * cells in the reference's rotated ascending order (:1553-1575), that lies
 * within MaxRadius + MARGIN.  Writes j into out (if non-NULL); returns the count. */
static int build_row(const orc *s, int i, int *out, int *nfluid, int *nsolid) {
    const int T = s->ncell, c1 = s->cc[1], c2 = s->cc[2];
    const int ic[3] = {(s->key[i] / (c1 * c2)) % T, (s->key[i] / c2) % c1, s->key[i] % c2};
    const int rng[3] = {1, 1, s->dim == 2 ? 0 : 1};
    int js[3][3], jmin[3];
    for (int a = 0; a < 3; a++) {
        for (int k = 0; k < 2 * rng[a] + 1; k++) js[a][k] = wrapc(ic[a] - rng[a] + k, s->cc[a]);
        const int b = 2 * rng[a] - (ic[a] + rng[a]) % s->cc[a];
        jmin[a] = b > 0 ? b : 0;
    }
    const double cut = (s->rmax + 0.1 * s->d) * (s->rmax + 0.1 * s->d);
    int cnt = 0, nf = 0, ns = 0;
    for (int blk = 0; blk < 3; blk++)
        for (int jx = jmin[0]; jx < jmin[0] + 2 * rng[0] + 1; ++jx)
            for (int jy = jmin[1]; jy < jmin[1] + 2 * rng[1] + 1; ++jy)
                for (int jz = jmin[2]; jz < jmin[2] + 2 * rng[2] + 1; ++jz) {
                    const int cx = js[0][jx % (2 * rng[0] + 1)], cy = js[1][jy % (2 * rng[1] + 1)],
                              cz = js[2][jz % (2 * rng[2] + 1)];
                    const int k = blk * T + cx * c1 * c2 + cy * c2 + cz;
                    for (int j = s->cell_start[k]; j < s->cell_start[k + 1]; ++j) {
                        double q[3];
                        sep(s, i, j, q);
                        if (q[0] * q[0] + q[1] * q[1] + q[2] * q[2] <= cut) {
                            if (out) out[cnt] = j;
                            cnt++;
                            if (blk == 0) nf++;
                            if (blk == 1) ns++;
                        }
                    }
                }
    if (nfluid) *nfluid = nf;
    if (nsolid) *nsolid = ns;
    return cnt;
}

static void build_neighbors(orc *s) {
    const int n = s->n;
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < n; i++) s->ncount[i] = build_row(s, i, NULL, &s->fcount[i], &s->scount[i]);
    s->nptr[0] = 0;
    for (int i = 0; i < n; i++) s->nptr[i + 1] = s->nptr[i] + s->ncount[i];
    s->nind_n = s->nptr[n];
    free(s->nind);
    s->nind = (int *)malloc(sizeof(int) * (size_t)(s->nind_n > 0 ? s->nind_n : 1));
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < n; i++) build_row(s, i, s->nind + s->nptr[i], NULL, NULL);
    memcpy(s->build_pos, s->pos, sizeof(vec3) * (size_t)n);
    s->rebuilds++;
}

/* neighborCalculation, source/DEM-MPH.cpp:1264-1286 */
static int needs_rebuild(const orc *s) {
    double mx = 0.0;
#pragma omp parallel for reduction(max : mx)
    for (int i = 0; i < s->n; i++) {
        double dsp[3];
        for (int a = 0; a < 3; a++)
            dsp[a] = MODW(s->pos[i][a] - s->build_pos[i][a] + 0.5 * s->dw[a], s->dw[a]) - 0.5 * s->dw[a];
        const double sh = dsp[0] * dsp[0] + dsp[1] * dsp[1] + dsp[2] * dsp[2];
        if (sh > mx) mx = sh;
    }
    const double margin = 0.1 * s->d;
    return mx > 0.5 * margin * 0.5 * margin;
}

/* ------------------------------------------------------------------------- */
/* pair passes                                                                */
/* ------------------------------------------------------------------------- */

#define FOR_ROW(s, i, j) for (int64_t _q = (s)->nptr[i], _e = (s)->nptr[i] + (s)->ncount[i]; _q < _e; ++_q) \
                         for (int j = (s)->nind[_q], _once = 1; _once; _once = 0)

/* calculateDensityA, source/DEM-MPH.cpp:1956-1984 */
static void pass_density_a(orc *s) {
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < s->n; i++) {
        double sum = 0.0;
        FOR_ROW(s, i, j) {
            if (IS_SOLID(s->type[i]) && IS_WALL(s->type[j])) continue;
            if (i == j) continue;
            const double ratio = s->p.interaction[s->type[i]][s->type[j]];
            double x[3];
            sep(s, i, j, x);
            const double r2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
            if (s->ra * s->ra - r2 >= 0) sum += ratio * w_a(s, sqrt(r2), s->ra);
        }
        s->dens_a[i] = sum;
    }
}

/* calculateGravityCenter, source/DEM-MPH.cpp:1987-2020 */
static void pass_gravity_center(orc *s) {
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < s->n; i++) {
        double sum[3] = {0.0, 0.0, 0.0};
        FOR_ROW(s, i, j) {
            if (IS_SOLID(s->type[i]) && IS_WALL(s->type[j])) continue;
            if (i == j) continue;
            const double ratio = s->p.interaction[s->type[i]][s->type[j]];
            double x[3];
            sep(s, i, j, x);
            const double r2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
            if (s->rg * s->rg - r2 >= 0) {
                const double w = ratio * W_G(s, sqrt(r2));
                for (int a = 0; a < 3; a++) sum[a] += x[a] * w / s->r2g * s->rg;
            }
        }
        for (int a = 0; a < 3; a++) s->gc[i][a] = sum[a];
    }
}

/* calculateDensityP, source/DEM-MPH.cpp:2129-2156 */
static void pass_density_p(orc *s) {
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < s->n; i++) {
        double sum = 0.0;
        FOR_ROW(s, i, j) {
            if (i == j) continue;
            double x[3];
            sep(s, i, j, x);
            const double r2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
            if (s->rp * s->rp - r2 >= 0) sum += W_P(s, sqrt(r2));
        }
        s->vs_p[i] = sum - s->n0p;
    }
}

/* the skip rules shared by DivergenceP, PressureP (:2175-2177, :2228-2230) */
static inline int skip_sw(int ti, int tj) {
    return (IS_SOLID(ti) && IS_WALL(tj)) || (IS_SOLID(tj) && IS_WALL(ti)) || (IS_SOLID(tj) && IS_SOLID(ti));
}

/* calculateDivergenceP, source/DEM-MPH.cpp:2158-2198 */
static void pass_divergence_p(orc *s) {
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < s->n; i++) {
        double sum = 0.0;
        FOR_ROW(s, i, j) {
            if (i == j) continue;
            double x[3];
            sep(s, i, j, x);
            const double r2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
            if (skip_sw(s->type[i], s->type[j])) continue;
            if (s->rp * s->rp - r2 >= 0) {
                const double r = sqrt(r2), dw = DW_P(s, r);
                const double e[3] = {x[0] / r, x[1] / r, x[2] / r};
                double u[3];
                for (int a = 0; a < 3; a++) u[a] = s->vel[j][a] - s->vel[i][a];
                for (int a = 0; a < 3; a++) sum -= u[a] * e[a] * dw;
            }
        }
        s->div_p[i] = sum;
    }
}

/* calculatePhysicalCoefficients, source/DEM-MPH.cpp:1918-1949 */
static void pass_coefficients(orc *s) {
#pragma omp parallel for
    for (int i = 0; i < s->n; i++) {
        const int t = s->type[i];
        s->mass[i] = s->rho[i] * s->vol;
        s->kappa[i] = s->p.bulk_modulus[t];
        if (s->vs_p[i] < 0.0) s->kappa[i] = 0.0;
        s->lambda[i] = s->p.bulk_viscosity[t];
        s->mu[i] = s->p.shear_viscosity[t];
    }
}

/* calculatePressureP, source/DEM-MPH.cpp:2200-2246.  Solid slots keep their
 * previous PressureP (never written: zero), which fluid neighbours do read. */
static void pass_pressure_p(orc *s) {
#pragma omp parallel for
    for (int i = 0; i < s->n; i++) {
        if (IS_SOLID(s->type[i])) continue;
        s->press_p[i] = -s->lambda[i] * s->div_p[i] + s->kappa[i] * s->vs_p[i];
    }
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < s->n; i++) {
        if (IS_SOLID(s->type[i])) continue;
        double f[3] = {0.0, 0.0, 0.0};
        FOR_ROW(s, i, j) {
            if (i == j) continue;
            double x[3];
            sep(s, i, j, x);
            const double r2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
            if (skip_sw(s->type[i], s->type[j])) continue;
            if (s->rp * s->rp - r2 > 0) {
                const double r = sqrt(r2), dw = DW_P(s, r);
                const double g[3] = {dw * x[0] / r, dw * x[1] / r, dw * x[2] / r};
                for (int a = 0; a < 3; a++) f[a] += (s->press_p[i] + s->press_p[j]) * g[a] * s->vol;
            }
        }
        for (int a = 0; a < 3; a++) s->force[i][a] += f[a];
    }
}

static inline int skip_sw2(int ti, int tj) {
    return (IS_SOLID(ti) && IS_WALL(tj)) || (IS_SOLID(tj) && IS_WALL(ti));
}

/* calculatePressureA, source/DEM-MPH.cpp:2022-2072 */
static void pass_pressure_a(orc *s) {
#pragma omp parallel for
    for (int i = 0; i < s->n; i++) {
        s->press_a[i] = s->cofa[s->type[i]] * (s->dens_a[i] - s->n0a) / s->d;
        if (s->n0a <= s->dens_a[i]) s->press_a[i] = 0.0;
    }
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < s->n; i++) {
        double f[3] = {0.0, 0.0, 0.0};
        FOR_ROW(s, i, j) {
            if (i == j) continue;
            const double rij_ = s->p.interaction[s->type[i]][s->type[j]];
            const double rji_ = s->p.interaction[s->type[j]][s->type[i]];
            double x[3];
            sep(s, i, j, x);
            const double r2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
            if (skip_sw2(s->type[i], s->type[j])) continue;
            if (s->ra * s->ra - r2 > 0) {
                const double r = sqrt(r2);
                const double dwij = rij_ * dw_a(s, r, s->ra), dwji = rji_ * dw_a(s, r, s->ra);
                const double e[3] = {x[0] / r, x[1] / r, x[2] / r};
                for (int a = 0; a < 3; a++) f[a] += (s->press_a[i] * dwij + s->press_a[j] * dwji) * e[a] * s->vol;
            }
        }
        for (int a = 0; a < 3; a++) s->force[i][a] += f[a];
    }
}

/* calculateDiffuseInterface, source/DEM-MPH.cpp:2074-2127 (aj uses type of i, :2088) */
static void pass_diffuse_interface(orc *s) {
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < s->n; i++) {
        const double ai = s->cofa[s->type[i]] * (s->cofk) * (s->cofk);
        double f[3] = {0.0, 0.0, 0.0};
        FOR_ROW(s, i, j) {
            if (i == j) continue;
            const double aj = s->cofa[s->type[i]] * (s->cofk) * (s->cofk);
            const double rij_ = s->p.interaction[s->type[i]][s->type[j]];
            const double rji_ = s->p.interaction[s->type[j]][s->type[i]];
            double x[3];
            sep(s, i, j, x);
            const double r2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
            if (skip_sw2(s->type[i], s->type[j])) continue;
            if (s->rg * s->rg - r2 > 0) {
                const double r = sqrt(r2);
                const double wij = rij_ * W_G(s, r), wji = rji_ * W_G(s, r);
                for (int a = 0; a < 3; a++)
                    f[a] -= (aj * s->gc[j][a] * wji - ai * s->gc[i][a] * wij) / s->r2g * s->rg * (s->vol / s->d);
                const double dwij = rij_ * DW_G(s, r), dwji = rji_ * DW_G(s, r);
                const double e[3] = {x[0] / r, x[1] / r, x[2] / r};
                double gr = 0.0;
                for (int a = 0; a < 3; a++) gr += (aj * s->gc[j][a] * dwji - ai * s->gc[i][a] * dwij) * x[a];
                for (int a = 0; a < 3; a++) f[a] -= (gr)*e[a] / s->r2g * s->rg * (s->vol / s->d);
            }
        }
        for (int a = 0; a < 3; a++) s->force[i][a] += f[a];
    }
}

/* calculateViscosityV, source/DEM-MPH.cpp:2248-2293 */
static void pass_viscosity(orc *s) {
    const double cvisc = s->dim == 2 ? 8.0 : 10.0;
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < s->n; i++) {
        double f[3] = {0.0, 0.0, 0.0};
        FOR_ROW(s, i, j) {
            if (i == j) continue;
            double x[3];
            sep(s, i, j, x);
            const double r2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
            if (IS_SOLID(s->type[i]) && IS_WALL(s->type[j])) continue;
            if (s->rv * s->rv - r2 > 0) {
                const double r = sqrt(r2), dwij = -DW_V(s, r);
                const double e[3] = {x[0] / r, x[1] / r, x[2] / r};
                double u[3];
                for (int a = 0; a < 3; a++) u[a] = s->vel[j][a] - s->vel[i][a];
                const double muij = 2.0 * (s->mu[i] * s->mu[j]) / (s->mu[i] + s->mu[j]);
                for (int a = 0; a < 3; a++)
                    f[a] += cvisc * muij * (u[0] * e[0] + u[1] * e[1] + u[2] * e[2]) * e[a] * dwij / r * s->vol;
            }
        }
        for (int a = 0; a < 3; a++) s->force[i][a] += f[a];
    }
}

/* calculateDragForce, source/DEM-MPH.cpp:2740-2807 (applied twice per step) */
static void pass_drag(orc *s) {
    const double d = s->d;
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = s->sb; i < s->se; ++i) {
        double f[3] = {0.0, 0.0, 0.0};
        if (s->ncount[i] > 0) s->alpha[i] = (double)s->scount[i] / (s->ncount[i]);
        else s->alpha[i] = 0.0;
        const double al = s->alpha[i], rho_i = s->rho[i];
        FOR_ROW(s, i, j) {
            if (i == j) continue;
            double x[3];
            sep(s, i, j, x);
            const double r2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
            if (s->rp * s->rp - r2 > 0 && IS_FLUID(s->type[j])) {
                double u[3] = {0.0, 0.0, 0.0};
                for (int a = 0; a < 3; a++) u[a] = s->vel[j][a] - s->vel[i][a];
                const double vs = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
                const double re = al * rho_i * vs * d / (s->mu[j]);
                if (al < 0.8) s->drag_c[i] = (4.0 / 3.0) * (150.0 * (1.0 - al) / (al * re) + 1.75);
                else if (re <= 1000.0 && al >= 0.8)
                    s->drag_c[i] = 24.0 * (1.0 + 0.15 * pow(re, 0.687)) / (al * re) * pow(al, -2.65);
                else s->drag_c[i] = 0.44;
                for (int a = 0; a < 3; a++) f[a] += 0.5 * s->drag_c[i] * rho_i * d * d * d * u[a] * vs;
            }
        }
        for (int a = 0; a < 3; a++) s->force[i][a] += f[a];
    }
}

/* calculateBuoyancyForce, source/DEM-MPH.cpp:2699-2738 */
static void pass_buoyancy(orc *s) {
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = s->sb; i < s->se; ++i) {
        double f[3] = {0.0, 0.0, 0.0};
        FOR_ROW(s, i, j) {
            if (i == j) continue;
            double x[3];
            sep(s, i, j, x);
            const double r2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
            if (s->rp * s->rp - r2 > 0 && IS_FLUID(s->type[j])) {
                const double w = W_P(s, sqrt(r2));
                for (int a = 0; a < 3; a++) f[a] += s->rho[j] * s->vol * s->p.gravity[a] * w;
            }
        }
        for (int a = 0; a < 3; a++) s->force[i][a] -= f[a];
    }
}

/* calculateGravity + calculateAcceleration, source/DEM-MPH.cpp:2811-2843 */
static void pass_gravity_accel(orc *s) {
    for (int i = s->fb; i < s->fe; ++i)
        for (int a = 0; a < 3; a++) s->force[i][a] += s->mass[i] * s->p.gravity[a];
    for (int i = s->sb; i < s->se; ++i)
        for (int a = 0; a < 3; a++) s->force[i][a] += s->mass[i] * s->p.gravity[a];
    for (int i = s->fb; i < s->fe; ++i)
        for (int a = 0; a < 3; a++) s->vel[i][a] += s->force[i][a] / s->mass[i] * s->p.dt;
}

static inline int same_body(const orc *s, int i, int j) {
    return s->o.rigid && s->body[i] >= 0 && s->body[i] == s->body[j];
}

/* calculateNormal/TangentialDirectionCoefficient, source/DEM-MPH.cpp:2298-2400:
 * for every non-fluid i, each neighbour j within d overwrites i's four
 * coefficients, so the last one in row order wins. */
static void pass_dem_coefficients(orc *s) {
    const double d = s->d;
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < s->n; i++) {
        if (IS_FLUID(s->type[i])) continue;
        const int ti = s->type[i];
        const double le = log(s->p.restitution[ti]);
        FOR_ROW(s, i, j) {
            if (i == j || same_body(s, i, j)) continue;
            double x[3];
            sep(s, i, j, x);
            const double r2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
            if (r2 <= d * d) {
                const int tj = s->type[j];
                const double r = sqrt(r2);
                const double yi = s->p.young[ti], yj = s->p.young[tj], pi = s->p.poisson[ti], pj = s->p.poisson[tj];
                const double eeff = (yi * yj) / ((1 - pi * pi) * yj + (1 - pj * pj) * yi);
                const double reff = d * d / (d + d);
                const double ov = fmax(0.0, d - r);
                const double mij = s->mass[i] * s->mass[j] / (s->mass[i] + s->mass[j]);
                s->kn[i] = (4.0 / 3.0) * eeff * sqrt(reff * ov);
                s->etan[i] = -2.0 * sqrt(mij * s->kn[i]) * le / sqrt(M_PI * M_PI + le * le);
                const double kt = 8.0 * yi * yj / (((1 + pi) * (2 - pi) * yj) + ((1 + pj) * (2 - pj) * yi)) *
                                  sqrt(reff * sqrt(ov));
                s->kt[i] = kt;
                const double damp = -2 * le / sqrt(M_PI * M_PI + le * le);
                s->etat[i] = damp * sqrt(mij * kt);
            }
        }
    }
}

/* calculateNormalDirectionForce + calculateTangentialDirectionForce,
 * source/DEM-MPH.cpp:2560-2697: each pass keeps its own per-i sum. */
static void pass_dem_forces(orc *s) {
    const double d = s->d, dt = s->p.dt, ddt = s->p.dem_dt;
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < s->n; i++) {
        if (IS_FLUID(s->type[i])) continue;
        double fn[3] = {0.0, 0.0, 0.0}, ft[3] = {0.0, 0.0, 0.0};
        const double kn = s->kn[i], en = s->etan[i], kt = s->kt[i], et = s->etat[i];
        const double muf = s->p.friction[s->type[i]];
        FOR_ROW(s, i, j) {
            if (i == j || IS_FLUID(s->type[j]) || same_body(s, i, j)) continue;
            double x[3];
            sep(s, i, j, x);
            const double r2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
            if (r2 <= d * d) {
                const double r = sqrt(r2);
                const double e[3] = {x[0] / r, x[1] / r, x[2] / r};
                double v[3], vn[3], vt[3], dtv[3], sn[3], st[3];
                for (int a = 0; a < 3; a++) v[a] = s->vel[j][a] - s->vel[i][a];
                for (int a = 0; a < 3; a++) vn[a] = v[a] * e[a];
                for (int a = 0; a < 3; a++) fn[a] += -kn * (d - r) * e[a] - (en * vn[a] * e[a] * dt);
                for (int a = 0; a < 3; a++) vt[a] = v[a] - vn[a] * e[a];
                for (int a = 0; a < 3; a++) dtv[a] = 0.0 + vt[a] * ddt;
                for (int a = 0; a < 3; a++) sn[a] = -kn * (d - r) * e[a] - (en * vn[a] * e[a] * dt);
                const double afn = muf * sqrt(sn[0] * sn[0] + sn[1] * sn[1] + sn[2] * sn[2]);
                for (int a = 0; a < 3; a++) st[a] = -kt * dtv[a] - (et * vt[a]);
                const double aft = sqrt(st[0] * st[0] + st[1] * st[1] + st[2] * st[2]);
                if (afn > aft) {
                    for (int a = 0; a < 3; a++) ft[a] += -kt * dtv[a] - (et * vt[a]);
                } else {
                    const double avt = sqrt(vt[0] * vt[0] + vt[1] * vt[1] + vt[2] * vt[2]);
                    if (avt != 0)
                        for (int a = 0; a < 3; a++) ft[a] += afn * vt[a] / avt;
                }
            }
        }
        for (int a = 0; a < 3; a++) s->force[i][a] += fn[a];
        for (int a = 0; a < 3; a++) s->force[i][a] += ft[a];
    }
}

/* ------------------------------------------------------------------------- */
/* rigid bodies (no reference counterpart)                                    */
/* ------------------------------------------------------------------------- */

static void quat_mat(const double q[4], double R[3][3]) {   /* q = (w, x, y, z), unit */
    const double w = q[0], x = q[1], y = q[2], z = q[3];
    R[0][0] = 1 - 2 * (y * y + z * z); R[0][1] = 2 * (x * y - w * z);     R[0][2] = 2 * (x * z + w * y);
    R[1][0] = 2 * (x * y + w * z);     R[1][1] = 1 - 2 * (x * x + z * z); R[1][2] = 2 * (y * z - w * x);
    R[2][0] = 2 * (x * z - w * y);     R[2][1] = 2 * (y * z + w * x);     R[2][2] = 1 - 2 * (x * x + y * y);
}
static void mat_vec(double R[3][3], const double v[3], double o[3]) {
    for (int a = 0; a < 3; a++) o[a] = R[a][0] * v[0] + R[a][1] * v[1] + R[a][2] * v[2];
}
static void cross(const double a[3], const double b[3], double o[3]) {
    o[0] = a[1] * b[2] - a[2] * b[1]; o[1] = a[2] * b[0] - a[0] * b[2]; o[2] = a[0] * b[1] - a[1] * b[0];
}
/* omega = R I_body^-1 R^T L */
static void body_omega(const orc *s, int b, double R[3][3], double w[3]) {
    double lb[3], t[3];
    for (int a = 0; a < 3; a++) lb[a] = R[0][a] * s->b_l[b][0] + R[1][a] * s->b_l[b][1] + R[2][a] * s->b_l[b][2];
    for (int a = 0; a < 3; a++) t[a] = s->b_iinv[b][a][0] * lb[0] + s->b_iinv[b][a][1] * lb[1] + s->b_iinv[b][a][2] * lb[2];
    mat_vec(R, t, w);
}
static int inv3(double A[3][3], double B[3][3]) {
    const double det = A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1]) -
                       A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0]) +
                       A[0][2] * (A[1][0] * A[2][1] - A[1][1] * A[2][0]);
    if (det == 0.0) return -1;
    B[0][0] = (A[1][1] * A[2][2] - A[1][2] * A[2][1]) / det;
    B[0][1] = (A[0][2] * A[2][1] - A[0][1] * A[2][2]) / det;
    B[0][2] = (A[0][1] * A[1][2] - A[0][2] * A[1][1]) / det;
    B[1][0] = (A[1][2] * A[2][0] - A[1][0] * A[2][2]) / det;
    B[1][1] = (A[0][0] * A[2][2] - A[0][2] * A[2][0]) / det;
    B[1][2] = (A[0][2] * A[1][0] - A[0][0] * A[1][2]) / det;
    B[2][0] = (A[1][0] * A[2][1] - A[1][1] * A[2][0]) / det;
    B[2][1] = (A[0][1] * A[2][0] - A[0][0] * A[2][1]) / det;
    B[2][2] = (A[0][0] * A[1][1] - A[0][1] * A[1][0]) / det;
    return 0;
}
static void min_image(const orc *s, const double a[3], const double b[3], double o[3]) {
    for (int k = 0; k < 3; k++) o[k] = MODW(a[k] - b[k] + 0.5 * s->dw[k], s->dw[k]) - 0.5 * s->dw[k];
}

/* Body mass, centre, body-frame offsets, inertia (point masses plus each
 * particle's own sphere, 2/5 m (d/2)^2) and momenta from the initial state. */
static void init_bodies(orc *s) {
    int nb = 0;
    for (int i = 0; i < s->n; i++)
        if (s->body[i] >= nb) nb = s->body[i] + 1;
    s->nb = s->o.rigid ? nb : 0;
    if (!s->nb) return;
    nb = s->nb;
    s->b_mass = (double *)calloc((size_t)nb, sizeof(double));
    s->b_x = (vec3 *)calloc((size_t)nb, sizeof(vec3));
    s->b_v = (vec3 *)calloc((size_t)nb, sizeof(vec3));
    s->b_l = (vec3 *)calloc((size_t)nb, sizeof(vec3));
    s->b_q = (double (*)[4])calloc((size_t)nb, sizeof(double[4]));
    s->b_i = (double (*)[3][3])calloc((size_t)nb, sizeof(double[3][3]));
    s->b_iinv = (double (*)[3][3])calloc((size_t)nb, sizeof(double[3][3]));
    int *first = (int *)malloc(sizeof(int) * (size_t)nb);
    vec3 *acc = (vec3 *)calloc((size_t)nb, sizeof(vec3));
    for (int b = 0; b < nb; b++) first[b] = -1;
    for (int i = 0; i < s->n; i++) {
        const int b = s->body[i];
        if (b < 0) continue;
        if (first[b] < 0) first[b] = i;
        const double m = s->rho[i] * s->vol;
        double dx[3];
        min_image(s, s->pos[i], s->pos[first[b]], dx);
        s->b_mass[b] += m;
        for (int a = 0; a < 3; a++) { acc[b][a] += m * dx[a]; s->b_v[b][a] += m * s->vel[i][a]; }
    }
    for (int b = 0; b < nb; b++) {
        s->b_q[b][0] = 1.0;
        if (first[b] < 0 || s->b_mass[b] == 0.0) continue;
        for (int a = 0; a < 3; a++) {
            s->b_x[b][a] = s->pos[first[b]][a] + acc[b][a] / s->b_mass[b];
            s->b_v[b][a] /= s->b_mass[b];
        }
    }
    const double rs = 0.5 * s->d;
    double (*ipt)[3][3] = (double (*)[3][3])calloc((size_t)nb, sizeof(double[3][3]));
    for (int i = 0; i < s->n; i++) {
        const int b = s->body[i];
        if (b < 0) continue;
        const double m = s->rho[i] * s->vol;
        min_image(s, s->pos[i], s->b_x[b], s->r0[i]);
        const double *r = s->r0[i];
        const double rr = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
        for (int a = 0; a < 3; a++)
            for (int c = 0; c < 3; c++) {
                const double pt = m * ((a == c ? rr : 0.0) - r[a] * r[c]);
                ipt[b][a][c] += pt;
                s->b_i[b][a][c] += pt + (a == c ? 0.4 * m * rs * rs : 0.0);
            }
        double u[3], l[3];
        for (int a = 0; a < 3; a++) u[a] = s->vel[i][a] - s->b_v[b][a];
        cross(r, u, l);
        for (int a = 0; a < 3; a++) s->b_l[b][a] += m * l[a];
    }
    /* angular velocity fitted to the point-mass momenta, then L = I omega so a
     * rigid initial motion keeps its omega once the particle spheres are counted */
    for (int b = 0; b < nb; b++) {
        double ip[3][3], w[3];
        inv3(s->b_i[b], s->b_iinv[b]);
        if (inv3(ipt[b], ip) == 0) {
            for (int a = 0; a < 3; a++) w[a] = ip[a][0] * s->b_l[b][0] + ip[a][1] * s->b_l[b][1] + ip[a][2] * s->b_l[b][2];
            for (int a = 0; a < 3; a++) s->b_l[b][a] = s->b_i[b][a][0] * w[0] + s->b_i[b][a][1] * w[1] + s->b_i[b][a][2] * w[2];
        }
    }
    free(ipt); free(first); free(acc);
}

/* Per-body FP64 force/torque sums, then semi-implicit Euler for V, X, L and the
 * exponential-map update of q, then re-posing of the member particles. */
static void integrate_bodies(orc *s, double h) {
    const int nb = s->nb;
    vec3 *F = (vec3 *)calloc((size_t)nb, sizeof(vec3)), *Tq = (vec3 *)calloc((size_t)nb, sizeof(vec3));
    double (*R)[3][3] = (double (*)[3][3])malloc(sizeof(double[3][3]) * (size_t)nb);
    for (int b = 0; b < nb; b++) quat_mat(s->b_q[b], R[b]);
    for (int i = 0; i < s->n; i++) {
        const int b = s->body[i];
        if (b < 0) continue;
        double arm[3], t[3];
        mat_vec(R[b], s->r0[i], arm);
        cross(arm, s->force[i], t);
        for (int a = 0; a < 3; a++) { F[b][a] += s->force[i][a]; Tq[b][a] += t[a]; }
    }
    for (int b = 0; b < nb; b++) {
        if (s->b_mass[b] == 0.0) continue;
        for (int a = 0; a < 3; a++) {
            s->b_v[b][a] += F[b][a] / s->b_mass[b] * h;
            s->b_x[b][a] += s->b_v[b][a] * h;
            s->b_l[b][a] += Tq[b][a] * h;
        }
        double w[3];
        body_omega(s, b, R[b], w);
        const double wn = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
        if (wn > 0.0) {
            const double half = 0.5 * wn * h, sn = sin(half) / wn;
            const double dq[4] = {cos(half), w[0] * sn, w[1] * sn, w[2] * sn};
            const double *q = s->b_q[b];
            double nq[4] = {dq[0] * q[0] - dq[1] * q[1] - dq[2] * q[2] - dq[3] * q[3],
                            dq[0] * q[1] + dq[1] * q[0] + dq[2] * q[3] - dq[3] * q[2],
                            dq[0] * q[2] - dq[1] * q[3] + dq[2] * q[0] + dq[3] * q[1],
                            dq[0] * q[3] + dq[1] * q[2] - dq[2] * q[1] + dq[3] * q[0]};
            const double nn = sqrt(nq[0] * nq[0] + nq[1] * nq[1] + nq[2] * nq[2] + nq[3] * nq[3]);
            for (int a = 0; a < 4; a++) s->b_q[b][a] = nq[a] / nn;
        }
        quat_mat(s->b_q[b], R[b]);
    }
    for (int i = 0; i < s->n; i++) {
        const int b = s->body[i];
        if (b < 0) continue;
        double arm[3], w[3], wr[3];
        mat_vec(R[b], s->r0[i], arm);
        body_omega(s, b, R[b], w);
        cross(w, arm, wr);
        for (int a = 0; a < 3; a++) {
            s->pos[i][a] = s->b_x[b][a] + arm[a];
            s->vel[i][a] = s->b_v[b][a] + wr[a];
        }
    }
    free(F); free(Tq); free(R);
}

/* calculateDEMconvection, source/DEM-MPH.cpp:1872-1891 (Force is not reset between sub-steps) */
static void pass_dem_convection(orc *s) {
    const double h = s->p.dem_dt;
    for (int i = s->sb; i < s->se; ++i) {
        if (s->o.rigid && s->body[i] >= 0) continue;
        for (int a = 0; a < 3; a++) s->vel[i][a] += s->force[i][a] / s->mass[i] * h;
    }
    for (int i = s->sb; i < s->se; ++i) {
        if (s->o.rigid && s->body[i] >= 0) continue;
        for (int a = 0; a < 3; a++) s->pos[i][a] += s->vel[i][a] * h;
    }
    if (s->nb) integrate_bodies(s, h);
}

/* ------------------------------------------------------------------------- */
/* the step                                                                   */
/* ------------------------------------------------------------------------- */

/* calculateWall, source/DEM-MPH.cpp:2846-2900 */
static void pass_wall(orc *s) {
    const double t = s->time, dt = s->p.dt;
    for (int i = s->wb; i < s->we; ++i) for (int a = 0; a < 3; a++) s->force[i][a] = 0.0;
    for (int i = s->wb; i < s->we; ++i) {
        if (t >= 0.2 && 0.4 >= t) {
            const int w = s->type[i];
            const double *c = s->wall_center[w], *om = s->p.wall_omega[w], *V = s->p.wall_velocity[w];
            double (*R)[3] = s->wall_rot[w];
            double r[3] = {s->pos[i][0] - c[0], s->pos[i][1] - c[1], s->pos[i][2] - c[2]};
            /* in place, as the reference does: each row reads the rows already updated */
            r[0] = R[0][0] * r[0] + R[0][1] * r[1] + R[0][2] * r[2];
            r[1] = R[1][0] * r[0] + R[1][1] * r[1] + R[1][2] * r[2];
            r[2] = R[2][0] * r[0] + R[2][1] * r[1] + R[2][2] * r[2];
            s->vel[i][0] = om[1] * r[2] - om[2] * r[1] + V[0];
            s->vel[i][1] = om[2] * r[0] - om[0] * r[2] + V[1];
            s->vel[i][2] = om[0] * r[1] - om[1] * r[0] + V[2];
            for (int a = 0; a < 3; a++) s->pos[i][a] = r[a] + c[a] + V[a] * dt;
        } else {
            for (int a = 0; a < 3; a++) s->vel[i][a] = 0.0;
        }
    }
    /* `if (0.0 < Time < 0.2)` evaluates (0.0 < Time) < 0.2: true only for Time <= 0 */
    if (!(0.0 < t))
        for (int w = 4; w < 6; w++)
            for (int a = 0; a < 3; a++) s->wall_center[w][a] += s->p.wall_velocity[w][a] * dt;
}

static void one_step(orc *s) {
    const double dt = s->p.dt;
    for (int i = s->fb; i < s->fe; ++i)                              /* calculateConvection :1860 */
        for (int a = 0; a < 3; a++) s->pos[i][a] += s->vel[i][a] * dt;
    pass_wall(s);
    for (int i = 0; i < s->n; ++i)                                   /* calculatePeriodicBoundary :3153 */
        for (int a = 0; a < 3; a++) s->pos[i][a] = MODW(s->pos[i][a] - s->dmin[a], s->dw[a]) + s->dmin[a];
    for (int b = 0; b < s->nb; ++b)
        for (int a = 0; a < 3; a++) s->b_x[b][a] = MODW(s->b_x[b][a] - s->dmin[a], s->dw[a]) + s->dmin[a];
    memset(s->force, 0, sizeof(vec3) * (size_t)s->n);                /* resetForce :1894 */
    if (s->o.rebuild == ORC_REBUILD_EVERY_STEP || needs_rebuild(s)) {
        bin_particles(s);
        build_neighbors(s);
    }
    pass_density_a(s);
    pass_gravity_center(s);
    pass_density_p(s);
    pass_divergence_p(s);
    pass_coefficients(s);
    pass_drag(s);
    pass_pressure_p(s);
    pass_pressure_a(s);
    pass_diffuse_interface(s);
    pass_viscosity(s);
    pass_drag(s);
    pass_buoyancy(s);
    pass_gravity_accel(s);
    for (int k = 0; k < s->substeps; k++) {
        pass_dem_coefficients(s);
        pass_dem_forces(s);
        pass_dem_convection(s);
    }
    s->time += dt;
}

void orc_step(orc *s, int nsteps) {
    if (s->o.threads > 0) omp_set_num_threads(s->o.threads);
    for (int k = 0; k < nsteps; k++) one_step(s);
}

/* calculateVirialStressAtParticle, source/DEM-MPH.cpp:2904-3149 */
void orc_virial(orc *s) {
    const double cvisc = s->dim == 2 ? 8.0 : 10.0;
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < s->n; i++) {
        double st[4][3][3];
        memset(st, 0, sizeof st);
        for (int a = 0; a < 3; a++) for (int c = 0; c < 3; c++) s->virial[i][a][c] = 0.0;
        /* four passes, each summed separately then added in order */
        FOR_ROW(s, i, j) {
            if (i == j) continue;
            double x[3];
            sep(s, i, j, x);
            const double r2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
            if (s->rp * s->rp - r2 > 0) {
                const double r = sqrt(r2), dw = DW_P(s, r);
                const double g[3] = {dw * x[0] / r, dw * x[1] / r, dw * x[2] / r};
                double f[3];
                for (int a = 0; a < 3; a++) f[a] = (s->press_p[i]) * g[a] * s->vol;
                for (int a = 0; a < 3; a++) for (int c = 0; c < 3; c++) st[0][a][c] += 1.0 * f[a] * x[c] / s->vol;
            }
        }
        FOR_ROW(s, i, j) {
            if (i == j) continue;
            double x[3];
            sep(s, i, j, x);
            const double r2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
            if (s->ra * s->ra - r2 > 0) {
                const double ratio = s->p.interaction[s->type[i]][s->type[j]];
                const double r = sqrt(r2), dw = ratio * dw_a(s, r, s->ra);
                const double g[3] = {dw * x[0] / r, dw * x[1] / r, dw * x[2] / r};
                double f[3];
                for (int a = 0; a < 3; a++) f[a] = (s->press_a[i]) * g[a] * s->vol;
                for (int a = 0; a < 3; a++) for (int c = 0; c < 3; c++) st[1][a][c] += 1.0 * f[a] * x[c] / s->vol;
            }
        }
        FOR_ROW(s, i, j) {
            if (i == j) continue;
            double x[3];
            sep(s, i, j, x);
            const double r2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
            if (s->rv * s->rv - r2 > 0) {
                const double r = sqrt(r2), dw = -DW_V(s, r);
                const double e[3] = {x[0] / r, x[1] / r, x[2] / r};
                const double v[3] = {s->vel[j][0] - s->vel[i][0], s->vel[j][1] - s->vel[i][1], s->vel[j][2] - s->vel[i][2]};
                const double muij = 2.0 * (s->mu[i] * s->mu[j]) / (s->mu[i] + s->mu[j]);
                double f[3];
                for (int a = 0; a < 3; a++)
                    f[a] = cvisc * muij * (v[0] * e[0] + v[1] * e[1] + v[2] * e[2]) * e[a] * dw / r * s->vol;
                for (int a = 0; a < 3; a++) for (int c = 0; c < 3; c++) st[2][a][c] += 0.5 * f[a] * x[c] / s->vol;
            }
        }
        FOR_ROW(s, i, j) {
            if (i == j) continue;
            double x[3];
            sep(s, i, j, x);
            const double r2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
            if (s->rg * s->rg - r2 > 0) {
                const double aa = s->cofa[s->type[i]] * (s->cofk) * (s->cofk);
                const double ratio = s->p.interaction[s->type[i]][s->type[j]];
                const double r = sqrt(r2), w = ratio * W_G(s, r);
                double f[3];
                for (int a = 0; a < 3; a++) f[a] = -aa * (-s->gc[i][a]) * w / s->r2g * s->rg * (s->vol / s->d);
                for (int a = 0; a < 3; a++) for (int c = 0; c < 3; c++) st[3][a][c] += 1.0 * f[a] * x[c] / s->vol;
            }
            if (s->rg * s->rg - r2 > 0.0) {
                const double aa = s->cofa[s->type[i]] * (s->cofk) * (s->cofk);
                const double ratio = s->p.interaction[s->type[i]][s->type[j]];
                const double r = sqrt(r2), dw = ratio * DW_G(s, r);
                const double g[3] = {dw * x[0] / r, dw * x[1] / r, dw * x[2] / r};
                double gr = 0.0;
                for (int a = 0; a < 3; a++) gr += (-s->gc[i][a]) * x[a];
                double f[3];
                for (int a = 0; a < 3; a++) f[a] = -aa * (gr)*g[a] / s->r2g * s->rg * (s->vol / s->d);
                for (int a = 0; a < 3; a++) for (int c = 0; c < 3; c++) st[3][a][c] += 1.0 * f[a] * x[c] / s->vol;
            }
        }
        for (int k = 0; k < 4; k++)
            for (int a = 0; a < 3; a++) for (int c = 0; c < 3; c++) s->virial[i][a][c] += st[k][a][c];
        if (s->dim == 2) s->virial_p[i] = -1.0 / 2.0 * (s->virial[i][0][0] + s->virial[i][1][1]);
        else s->virial_p[i] = -1.0 / 3.0 * (s->virial[i][0][0] + s->virial[i][1][1] + s->virial[i][2][2]);
    }
}

/* ------------------------------------------------------------------------- */
/* construction and field access                                              */
/* ------------------------------------------------------------------------- */

#define ALLOC(p, cnt) ((p) = calloc((size_t)(cnt) > 0 ? (size_t)(cnt) : 1, sizeof(*(p))))

orc *orc_create(const orc_params *p, const orc_opts *o, double time, int n, double spacing,
                const double *domain_min, const double *domain_max, const int *type,
                const double *pos, const double *vel, const int *body, const double *density) {
    orc *s = (orc *)calloc(1, sizeof(orc));
    s->p = *p;
    s->o = *o;
    if (s->o.threads > 0) omp_set_num_threads(s->o.threads);
    s->n = n; s->dim = o->dim; s->d = spacing; s->time = time;
    s->vol = s->dim == 2 ? spacing * spacing : spacing * spacing * spacing;
    for (int a = 0; a < 3; a++) { s->dmin[a] = domain_min[a]; s->dmax[a] = domain_max[a]; }
    ALLOC(s->id, n); ALLOC(s->type, n); ALLOC(s->body, n); ALLOC(s->rho, n);
    ALLOC(s->pos, n); ALLOC(s->vel, n); ALLOC(s->force, n); ALLOC(s->gc, n); ALLOC(s->build_pos, n); ALLOC(s->r0, n);
    ALLOC(s->mass, n); ALLOC(s->dens_a, n); ALLOC(s->press_a, n); ALLOC(s->vs_p, n); ALLOC(s->div_p, n);
    ALLOC(s->press_p, n); ALLOC(s->kappa, n); ALLOC(s->lambda, n); ALLOC(s->mu, n); ALLOC(s->alpha, n);
    ALLOC(s->drag_c, n); ALLOC(s->kn, n); ALLOC(s->etan, n); ALLOC(s->kt, n); ALLOC(s->etat, n);
    ALLOC(s->virial_p, n); ALLOC(s->virial, n);
    ALLOC(s->ncount, n); ALLOC(s->fcount, n); ALLOC(s->scount, n); ALLOC(s->key, n); ALLOC(s->nptr, n + 1);
    for (int i = 0; i < n; i++) {
        s->id[i] = i;
        s->type[i] = type[i];
        s->body[i] = (body && IS_SOLID(type[i])) ? body[i] : -1;
        s->rho[i] = (density && density[i] > 0.0) ? density[i] : p->density[type[i]];
        for (int a = 0; a < 3; a++) { s->pos[i][a] = pos[3 * i + a]; s->vel[i][a] = vel[3 * i + a]; }
    }
    init_weight(s);
    init_fluid(s);
    init_wall(s);
    if (init_domain(s) != 0) { orc_destroy(s); return NULL; }
    ALLOC(s->cell_start, 3 * s->ncell + 1);
    s->substeps = (int)(s->p.dt / s->p.dem_dt);   /* source/DEM-MPH.cpp:434 */
    init_bodies(s);
    bin_particles(s);
    build_neighbors(s);
    pass_density_a(s);
    pass_gravity_center(s);
    pass_density_p(s);
    return s;
}

void orc_destroy(orc *s) {
    if (!s) return;
    void *ptrs[] = {s->id, s->type, s->body, s->rho, s->pos, s->vel, s->force, s->gc, s->build_pos, s->r0,
                    s->mass, s->dens_a, s->press_a, s->vs_p, s->div_p, s->press_p, s->kappa, s->lambda, s->mu,
                    s->alpha, s->drag_c, s->kn, s->etan, s->kt, s->etat, s->virial_p, s->virial, s->ncount,
                    s->fcount, s->scount, s->nind, s->key, s->cell_start, s->nptr, s->b_mass, s->b_x, s->b_v,
                    s->b_l, s->b_q, s->b_i, s->b_iinv};
    for (size_t k = 0; k < sizeof ptrs / sizeof ptrs[0]; k++) free(ptrs[k]);
    free(s);
}

typedef struct { const char *name; int dtype; } field_desc;

static const void *field_ptr(const orc *s, const char *nm, int *dt, int64_t *cnt) {
    const int64_t n = s->n;
    static const double zero = 0.0;
#define F64(key, ptr, c) if (!strcmp(nm, key)) { *dt = 0; *cnt = (c); return (const void *)(ptr); }
#define I32(key, ptr, c) if (!strcmp(nm, key)) { *dt = 1; *cnt = (c); return (const void *)(ptr); }
#define I64(key, ptr, c) if (!strcmp(nm, key)) { *dt = 2; *cnt = (c); return (const void *)(ptr); }
    I32("id", s->id, n) I32("type", s->type, n) I32("body", s->body, n) F64("rho", s->rho, n)
    F64("pos", s->pos, 3 * n) F64("vel", s->vel, 3 * n) F64("force", s->force, 3 * n)
    F64("mass", s->mass, n) F64("density_a", s->dens_a, n) F64("gravity_center", s->gc, 3 * n)
    F64("pressure_a", s->press_a, n) F64("vol_strain_p", s->vs_p, n) F64("divergence_p", s->div_p, n)
    F64("pressure_p", s->press_p, n) F64("kappa", s->kappa, n) F64("lambda", s->lambda, n) F64("mu", s->mu, n)
    F64("alpha", s->alpha, n) F64("drag_c", s->drag_c, n) F64("kn", s->kn, n) F64("etan", s->etan, n)
    F64("kt", s->kt, n) F64("etat", s->etat, n) F64("virial_pressure", s->virial_p, n)
    F64("virial_stress", s->virial, 9 * n) F64("build_pos", s->build_pos, 3 * n)
    I32("neighbor_count", s->ncount, n) I32("fluid_neighbor_count", s->fcount, n)
    I32("solid_neighbor_count", s->scount, n) I64("neighbor_ptr", s->nptr, n + 1)
    I32("neighbor_ind", s->nind, s->nind_n) I32("key", s->key, n) I32("cell_start", s->cell_start, 3 * s->ncell + 1)
    I32("cell_count", s->cc, 3) F64("cell_width", s->cw, 3) F64("wall_center", s->wall_center, 18)
    F64("wall_rotation", s->wall_rot, 54) F64("cofa", s->cofa, 6)
    F64("body_mass", s->b_mass ? (const void *)s->b_mass : (const void *)&zero, s->nb)
    F64("body_x", s->b_x, 3 * s->nb) F64("body_v", s->b_v, 3 * s->nb) F64("body_l", s->b_l, 3 * s->nb)
    F64("body_q", s->b_q, 4 * s->nb) F64("body_inertia", s->b_i, 9 * s->nb) F64("body_r0", s->r0, 3 * n)
    if (!strcmp(nm, "scalars")) {
        orc *m = (orc *)s;
        double v[] = {s->time, s->n0a, s->n0p, s->swa, s->swg, s->swp, s->swv, s->r2g, s->cofk, s->rmax,
                      s->ra, s->rg, s->rp, s->rv, s->vol, s->d, s->p.dt, s->p.dem_dt, (double)s->substeps,
                      (double)s->rebuilds};
        memcpy(m->scratch_scalars, v, sizeof v);
        *dt = 0; *cnt = (int64_t)(sizeof v / sizeof v[0]);
        return s->scratch_scalars;
    }
    if (!strcmp(nm, "ranges")) {
        orc *m = (orc *)s;
        int *r = (int *)m->scratch_scalars;
        r[0] = s->n; r[1] = s->ncell; r[2] = s->fb; r[3] = s->fe; r[4] = s->sb; r[5] = s->se; r[6] = s->wb;
        r[7] = s->we; r[8] = s->nb;
        *dt = 1; *cnt = 9;
        return r;
    }
#undef F64
#undef I32
#undef I64
    *dt = -1; *cnt = 0;
    return NULL;
}

int64_t orc_field_info(const orc *s, const char *name, int *dtype) {
    int64_t c = 0;
    const void *p = field_ptr(s, name, dtype, &c);
    return p || c == 0 ? (*dtype < 0 ? -1 : c) : -1;
}

int orc_get(const orc *s, const char *name, void *dst) {
    int dt;
    int64_t c;
    const void *p = field_ptr(s, name, &dt, &c);
    if (dt < 0) return -1;
    const size_t es = dt == 1 ? 4 : 8;
    if (c > 0 && p) memcpy(dst, p, es * (size_t)c);
    return 0;
}
