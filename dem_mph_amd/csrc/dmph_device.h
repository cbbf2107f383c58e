// Device-side data layout and launcher declarations for the DEM-MPH step.
//
// Particles live in "slots": structure-of-arrays sorted by the reference's cell
// key (cell id + block * cells, block 0 fluid / 1 solid / 2 wall;
// source/DEM-MPH.cpp:1292-1313).  A batch of scenes that share one domain adds
// the scene as a middle digit: key = block * T + scene * cells + cell with
// T = scenes * cells, so one set of launches advances every scene and a
// stencil never leaves its scene.  Block ranges never change after the first
// binning (the type counts are fixed), so [fb,fe) [sb,se) [wb,we) are host
// constants.  Per-particle state is permuted at every rebinning; the derived
// per-slot arrays (pressures, DEM coefficients, ...) stay where they are, as
// the reference's do.
//
// Positions: with Real = double they are absolute, as in the reference; with
// Real = float they are offsets from the origin of the particle's cell at the
// last binning, so a pair separation is shift*CellWidth + (o_j - o_i) with the
// stencil shift in {-1, 0, 1}: no periodic Mod, and FP32 resolves ~1e-9 m over
// a cell instead of ~1e-7 m over the domain.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dmph {

constexpr int kNType = 6;

// Uniform parameters of one scene (passed by value to every kernel).
struct Geo {
    int n, ncell, dim, nb, substeps;
    int nscene;         // independent scenes sharing the domain (a batch); 1 for a single scene
    int T;              // cells of one block over all scenes: nscene * ncell
    int cc[3];
    int fb, fe, sb, se, wb, we;
    int st_active;      // any non-zero InteractionRatio: surface-tension passes contribute
    int rigid;          // solid particles with body >= 0 move as rigid bodies
    int sort_reference; // within-cell order from the reference's bitonic network
    int rebuild_every;  // rebin every step instead of the skin test
    int stage_b;        // FP32 pass B with the neighbour cells staged in LDS (k_pass_b_lds)
    int p2, p2_log;     // bitonic size (power of two > n) and its log2
    int compensated;    // FP32 mode: FP64 accumulators for the pair sums
    int kcap;           // neighbour-list capacity per particle
    int knear;          // FP32 rows: near segment capacity (far entries start at this position), a multiple of 16
    int lanes[2];       // FP32 sweeps 1 and 2: lanes per particle (2, 4, 8 or 16)
    int lanes_w;        // FP32 pass A: lanes per wall row (1, 2 or 4)
    int block[2];       // FP32 sweeps 1 and 2: threads per workgroup (64, 128 or 256)
    int fuse_accel;     // FP32: the fluid sweep 2 also applies calculateAcceleration (0 while re-running sweeps)
    int fused;          // FP32 without surface tension: passes A and B in place of sweeps 1 and 2
    int nloose;         // solid particles outside rigid bodies (moved by k_move_solids)
    int ccap;           // DEM contact-list capacity per particle
    int body_f32;       // FP32 only: rigid-body sums and integration in FP32 (the plain solver)
    int body_tpb;       // threads per block of the body update: the largest body's members, 64 to 1024
    double cw[3], dmin[3], dw[3];
    double d, vol, dt, dem_dt;
    double ra, rg, rp, rv, rmax, cut2, skin2;
    double near2, con2;  // squared cutoffs of the FP32 near segment and of the DEM contact list
    double swa, swg, swp, swv, n0a, n0p, r2g, cofk;
    double cofa[kNType], dens[kNType], bulk[kNType], bvisc[kNType], shear[kNType];
    double young[kNType], poisson[kNType], logrest[kNType], fric[kNType];
    double inter[kNType * kNType];
    double grav[3];
    double wall_rot[2][9], wall_omega[2][3], wall_vel[2][3];  // wall types 4 and 5
};

// Device scalars shared by all kernels of a step.
struct Scalars {
    double time;
    unsigned long long skin_bits;   // max squared displacement since the last binning (as bits)
    int flag;                       // rebin this step
    int force;                      // rebin unconditionally at the next decision (first binning)
    long long rebuilds;
    int list_overflow;              // longest row that did not fit the neighbour list (0 = none)
    double wall_center[2][3];
    unsigned int ticket[4];         // blocks finished, per kernel that ends in a last-block phase
    int full_sweep;                 // FP32: > 0 -> sweep 1 covers every wall particle (the key pass counts it down)
    int walls_moving;               // this step moved the walls (0.2 <= Time <= 0.4, written by k_predict)
    int walls_dirty;                // walls moved (or the host set positions) since the last rebinning
    int build_walls_dirty;          // ... as of this rebinning: the list build re-walks every wall row
    int walls_rest;                 // every wall's velocity is zero: the last step moved no wall (set by the
                                    // rebin decision) and the host has written no state since (cleared)
};

template <class R>
struct Dev {
    // per-particle state (permuted on rebinning)
    int *id, *type, *body, *key;
    R *px, *py, *pz, *vx, *vy, *vz, *rho;
    double *r0;                    // [3n] body-frame offsets (rigid)
    // per-slot values (not permuted)
    R *fx, *fy, *fz;
    R *bx, *by, *bz;               // positions at the last binning
    R *dens_a, *gcx, *gcy, *gcz, *vs_p, *div_p, *press_p, *press_a, *mass, *kappa, *alpha, *drag_c;
    R *kn, *etan, *kt, *etat, *virial_p;
    int *ncount, *scount;
    uint32_t *nlist;               // [kcap * n] neighbour entries, slot-interleaved
    int *nnear;                    // FP32: entries in the near segment of each row
    uint32_t *clist;               // [ccap * (n - sb)] DEM contact entries of solid and wall slots
    int *ncontact;                 // [n - sb]
    int *cell_start;               // [3*T + 1]
    // rebinning scratch
    int *counts, *fill;            // [3*T + 1]
    int *key_new, *order, *perm;   // [n]
    int *bkey, *bval;              // [p2] bitonic
    int *block_sums;               // scan carries
    int *wall_counts;              // [T] wall cells' counts at the last rebinning that counted the walls
    int *t_int;                    // [5n] gather scratch
    R *t_real;                     // [8n]
    double *t_dbl;                 // [3n]
    // rigid bodies
    double *b_mass, *b_x, *b_v, *b_l, *b_q, *b_iinv, *b_w, *b_r;
    int *mrank;                    // [n] rank of a member within its body (fixed at set-up, moves with the particle)
    int *b_moff, *b_mslot;         // per body: offsets into b_mslot [nb + 1]; member slots [members], by rank
    unsigned long long *nonfinite; // [1] counter of dmph_count_nonfinite
    unsigned long long *skin_part; // [8 x 16] max squared displacement this step, one slot per XCD (bits of a
                                   // non-negative double, 128 B apart), zeroed by the scan after the key pass reads it
    float4 *p4, *v4;               // FP32 sweeps: packed (position, type) and (velocity, PressureP)
    int *wact;                     // FP32, per wall slot: its row at the last rebinning holds a fluid or solid particle
    int *wnfs;                     // FP32, per wall slot: near entries of fluid and solid particles (the row's first ones)
    double *wsq;                   // FP32, per wall slot: its wall neighbours' density sum, kept while walls stand still
    int *wdem;                     // FP32, per wall slot: its last DEM contact is a wall (coefficients kept while static)
    float4 *pp4;                   // FP32 passes A / B: (position, PressureP) written by pass A
    const double *x64;             // FP32, first binning only: the input positions [n][3] in original order
    Scalars *sc;
};

// Launchers (defined in dmph_kernels.inc, instantiated for float and double).
template <class R> void launch_predict(const Geo &g, const Dev<R> &a, hipStream_t s);
// decide = 1: the key pass takes the rebin decision of a step; 0: the flag is already set (first binning)
template <class R> void launch_rebin(const Geo &g, const Dev<R> &a, hipStream_t s, int decide);
// s_solid: the stream of the solids' pass A (FP32 fused passes; null = s)
template <class R> void launch_sweep1(const Geo &g, const Dev<R> &a, hipStream_t s, hipStream_t s_solid = nullptr);
template <class R> void launch_sweep2(const Geo &g, const Dev<R> &a, hipStream_t s);
template <class R> void launch_accel(const Geo &g, const Dev<R> &a, hipStream_t s);
template <class R> void launch_dem(const Geo &g, const Dev<R> &a, hipStream_t s);
template <class R> void launch_move(const Geo &g, const Dev<R> &a, hipStream_t s);
template <class R> void launch_virial(const Geo &g, const Dev<R> &a, hipStream_t s);
template <class R> void launch_init_density(const Geo &g, const Dev<R> &a, hipStream_t s);
template <class R> void launch_export(const Geo &g, const Dev<R> &a, float *pos, float *vel, hipStream_t s);
// count of the non-finite values of x[0..n) (and y, z when not null) added to *count
template <class R>
void launch_count_nonfinite(const R *x, const R *y, const R *z, int n, unsigned long long *count, hipStream_t s);

// FP32 sweeps (dmph_sweep32.inc, in the FP32 object only)
void launch_sweep1_f32(const Geo &g, const Dev<float> &a, hipStream_t s, int init, hipStream_t s_solid = nullptr);
void launch_sweep2_f32(const Geo &g, const Dev<float> &a, hipStream_t s);

}  // namespace dmph
