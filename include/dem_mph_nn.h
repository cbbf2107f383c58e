/* dem_mph_nn.h — C ABI of the message-passing surrogate kernels (libdemmph_nn.so).
 *
 * No reference counterpart: the reference (source/DEM-MPH.cpp) has no learned
 * component.  These are the building blocks of a graph network that predicts
 * particle accelerations at a 10 dt stride, trained on trajectories of the
 * plain solver; dem_mph_amd/surrogate.py drives them.  All pointers are device
 * pointers; `stream` is a hipStream_t.  bf16 values are raw 16-bit storage.
 * Weights of a layer with K inputs and N outputs are [N][K].  Calls return 0,
 * -1 for unsupported shapes, -2 when a capacity is too small, -3 when a graph
 * would overflow int32 row offsets, -4 when a size exceeds a fixed array of the
 * library, -5 on a HIP error.
 */
#ifndef DEM_MPH_NN_H
#define DEM_MPH_NN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* uniform grid for the radius graph: cell width >= radius; per scene cells */
typedef struct NnGrid {
    float lo[3];
    float inv_w;
    int cc[3];
    int ncell;    /* cc[0] * cc[1] * cc[2] */
    int nscene;   /* scenes in the batch (node scene ids < nscene) */
} NnGrid;

/* node features: velocity scale, tank interior box, 1 / radius */
typedef struct NnFeat {
    float vscale;
    float tlo[3], thi[3];
    float inv_r;
} NnFeat;

/* position fold of the inference edge sums.  The linear part of Wc eps_ij,
 * Wc[:, 0:3] (x_j - x_i) / radius, equals Wc3 P_j - Wc3 P_i with
 * P = (x - origin) * inv_r, so it moves into the node terms: the kernel that
 * writes AB adds Wc3 P_j to the B half, and dmph_nn_edge_fwd_f16 subtracts
 * Wc3 P_i from A_i, which leaves one multiply-add per edge (the distance).
 * Exact in real arithmetic; in IEEE half the B and A terms round at about
 * 2^-11 |Wc3 P|, so the origin should be the centre of the nodes' box and the
 * box a few tens of radii across at most. */
typedef struct NnFold {
    const float *pos;   /* node positions [N][3] */
    const float *Wt;    /* [3][128]: Wc[:, 0:3] transposed, of the edge layer that reads the AB being
                           written (dmph_nn_encode, dmph_nn_node; unused by dmph_nn_edge_fwd_f16) */
    float origin[3];
    float inv_r;        /* 1 / radius */
} NnFold;

/* gemm flags */
#define DMPH_NN_RELU 1      /* ReLU after bias and residual */
#define DMPH_NN_F16_OUT 2   /* Y in IEEE half instead of bf16 */

/* Y[:, 0:N) (row stride ldy) = act(X[M][K] W^T + bias [+ R]); K in {32, 64, 128, 256},
 * N % 128 == 0, ldy and ldr multiples of 4; flags DMPH_NN_*; kstep 16
 * (v_mfma_f32_16x16x16_bf16) or 32 (gfx950 16x16x32) */
int dmph_nn_gemm(const uint16_t *X, int ldx, const uint16_t *W, const float *bias, const uint16_t *R, int ldr,
                 uint16_t *Y, int ldy, int M, int K, int N, int flags, int kstep, void *stream);

/* decoder tail for inference: g = relu(X[M][128] W^T + bias) (W [128][128]) rounded
 * to bf16 as dmph_nn_gemm stores it, then Y3 [M][3] = g Wd^T + bd (Wd [3][128]) as
 * dmph_nn_head computes it; g is never stored */
int dmph_nn_gemm_head(const uint16_t *X, int ldx, const uint16_t *W, const float *bias, const float *Wd,
                      const float *bd, float *Y3, int M, int kstep, void *stream);

/* one round's node update at inference (kstep 32), intermediates in registers:
 * u = relu(W1 HS + b1) (HS [M][256] = [h | S], row stride ldhs; W1 [128][256]),
 * Hn = h + W2 u + b2 (bf16, [M][128] at row stride ldhn; W2 [128][128]),
 * AB = Wab Hn (IEEE half [M][256]; Wab [256][128]).  Equal to dmph_nn_gemm three
 * times up to the order of FP32 sums.  fold (NULL: none): B gets + Wc3 P (NnFold).
 * Uses 160 KB of LDS per workgroup. */
int dmph_nn_node(const uint16_t *HS, int ldhs, const uint16_t *W1, const float *b1, const uint16_t *W2, const float *b2,
                 const uint16_t *Wab, uint16_t *Hn, int ldhn, uint16_t *AB, int M, const NnFold *fold, void *stream);

/* the encoder at inference in one kernel: the features of dmph_nn_features,
 * then H0 = We2 relu(We1 x + be1) + be2 (bf16, row stride ldh0; We1 [128][32],
 * We2 [128][128]) and AB = Wab H0 (IEEE half [N][256]; Wab [256][128]); fold as
 * dmph_nn_node */
int dmph_nn_encode(const float *vel_hist, int N, const int *cls, const float *pos, const NnFeat *ft, const uint16_t *We1,
                   const float *be1, const uint16_t *We2, const float *be2, const uint16_t *Wab, uint16_t *H0, int ldh0,
                   uint16_t *AB, const NnFold *fold, void *stream);

/* the last round's node update and the decoder in one kernel at inference:
 * h' as dmph_nn_node (not stored), then Y3 [M][3] = Wd2 relu(Wd1 h' + bd1) + bd2
 * (Wd1 [128][128] bf16, Wd2 [3][128] FP32; the hidden layer rounded to bf16 as
 * dmph_nn_gemm_head rounds it).  128 KB of LDS per workgroup. */
int dmph_nn_node_head(const uint16_t *HS, int ldhs, const uint16_t *W1, const float *b1, const uint16_t *W2,
                      const float *b2, const uint16_t *Wd1, const float *bd1, const float *Wd2, const float *bd2,
                      float *Y3, int M, void *stream);

/* dW[N][K] (+)= dY^T X, db[N] (+)= sum_m dY (db may be NULL); K in {32, 128, 256},
 * N % 128 == 0; `groups` row chunks summed in a fixed order (deterministic);
 * scratch of dmph_nn_dw_scratch bytes */
size_t dmph_nn_dw_scratch(int M, int K, int N, int groups);
int dmph_nn_gemm_dw(const uint16_t *dY, int lddy, const uint16_t *X, int ldx, float *dW, float *db, int M, int K, int N,
                    int groups, float *scratch, int accumulate, void *stream);

/* S_i = sum_{j in row i} relu(A_i + B_j + Wc eps_ij + bc); AB = [A | B] as [N][256],
 * eps_ij = ((x_j - x_i) / radius, |x_j - x_i| / radius), Wc [128][4], bc [128],
 * S row i at S + i * lds (128 values; lds >= 128, even: lets S land in the right
 * half of a [h | S] buffer) */
int dmph_nn_edge_fwd(const uint16_t *AB, const float *pos, const int *rp, const int *ci, const float *Wc,
                     const float *bc, uint16_t *S, int lds, int N, float radius, void *stream);
/* Inference form of edge_fwd.  edge_eps writes each edge's features once per
 * graph as two IEEE-half pairs, eps[q] = ((dx, dy), (dz, |d|)) / radius (row
 * order of ci; eps holds nnz entries); edge_fwd_f16 then sums with AB in IEEE
 * half ([N][256], DMPH_NN_F16_OUT), packed half arithmetic in batches of 16
 * edges added into FP32 sums; S as in edge_fwd (bf16).  With deg (not NULL) the
 * rows are those of dmph_nn_graph_rows: row i at i * row_cap, deg[i] entries
 * (rp unused), edge_eps not needed (the graph writes eps).  With fold (not
 * NULL; rows only) the B half of AB holds + Wc3 P_j (written by a folding
 * dmph_nn_node or dmph_nn_encode with the same origin and inv_r), A_i gets
 * - Wc3 P_i here, and eps is the `dist` array of dmph_nn_graph_rows (one half
 * per entry): one packed multiply-add per edge instead of four.
 * order (NULL or a permutation of [0, N), e.g. the `order` work array of the
 * graph builders, which lists the nodes by cell): the nodes are summed in that
 * order, so the rows of B read at one time are neighbours in space and come
 * from L2; the sums are the same. */
int dmph_nn_edge_eps(const float *pos, const int *rp, const int *ci, uint32_t *eps, int N, float radius, void *stream);
int dmph_nn_edge_fwd_f16(const uint16_t *AB, const int *rp, const int *ci, const uint32_t *eps, const float *Wc,
                         const float *bc, uint16_t *S, int lds, int N, const int *deg, int row_cap,
                         const int *order, const NnFold *fold, void *stream);
/* gradients of edge_fwd given dS [N][128] FP32: dAB [N][256] bf16, dWc, dbc (+)=;
 * the graph must be symmetric; scratch of dmph_nn_edge_scratch bytes */
size_t dmph_nn_edge_scratch(int N);
int dmph_nn_edge_bwd(const uint16_t *AB, const float *pos, const int *rp, const int *ci, const float *Wc,
                     const float *bc, const float *dS, uint16_t *dAB, float *dWc, float *dbc, int N, float radius,
                     float *scratch, int accumulate, void *stream);

/* radius graph (CSR rp[N + 1], ci[cap]) of the nodes at pos [N][3] (scene ids or NULL):
 * rows ascending by cell then node; *nnz_out = edges; returns -2 if cap < nnz (call again),
 * -3 (nothing written, *nnz_out = -1) if the edges would not fit int32 row offsets.
 * eps (NULL or [2 cap] words): also each edge's features as edge_eps writes them.
 * Work arrays: key, order, deg [N], sorted [4N] (cell-sorted positions and ids);
 * counts, start, fill [cells * scenes + 1]; scratch of
 * dmph_nn_graph_scratch(N, cells * scenes) bytes (scan carries). */
size_t dmph_nn_graph_scratch(int N, int ncells_total);
int dmph_nn_graph(const float *pos, const int *scene, int N, const NnGrid *grid, float radius, int *key, int *counts,
                  int *start, int *fill, int *order, int *sorted, int *deg, int *rp, int *ci, uint32_t *eps, int cap,
                  int *nnz_out, int *scratch, void *stream);

/* the same graph in rows of a fixed capacity, one pass (no degree count first):
 * row i at ci + i * row_cap with deg[i] entries, same order as dmph_nn_graph;
 * eps (NULL or [2 N row_cap] words) as edge_eps writes it; dist (NULL or
 * [N row_cap] IEEE halves) the distance alone, |x_j - x_i| / radius, for the
 * folded edge sums (row_cap even).  Returns -2 with *over_out = rows longer than
 * row_cap (their tails unwritten: call again with a larger row_cap or use
 * dmph_nn_graph); same work arrays and scratch. */
int dmph_nn_graph_rows(const float *pos, const int *scene, int N, const NnGrid *grid, float radius, int *key,
                       int *counts, int *start, int *fill, int *order, int *sorted, int *deg, int *ci, uint32_t *eps,
                       uint16_t *dist, int row_cap, int *over_out, int *scratch, void *stream);

/* bounding box of pos [N][3] into out[6] (min x y z, max x y z; device memory);
 * scratch of dmph_nn_bbox_scratch(N) bytes */
size_t dmph_nn_bbox_scratch(int N);
int dmph_nn_bbox(const float *pos, int N, float *out, float *scratch, void *stream);

/* node features X [N][32] bf16 from vel_hist [5][N][3] (newest first), class 0/1/2, positions */
int dmph_nn_features(const float *vel_hist, int N, const int *cls, const float *pos, const NnFeat *ft, uint16_t *X,
                     void *stream);

/* output head: Y [N][3] = H Wd^T + bd (Wd [3][128]); backward with the ReLU mask of H */
int dmph_nn_head(const uint16_t *H, const float *Wd, const float *bd, float *Y, int N, void *stream);
size_t dmph_nn_head_scratch(int N);
int dmph_nn_head_bwd(const uint16_t *H, const float *Wd, const float *dY, uint16_t *dH, float *dWd, float *dbd, int N,
                     float *scratch, void *stream);

/* weighted mean squared error: sum_i mask[i] |Y_i - clip(T_i)|^2 / count with
 * integer weights mask[i] (0 = left out), targets clipped to [-clip, clip];
 * loss (device float), dY */
size_t dmph_nn_mse_scratch(int N);
int dmph_nn_mse(const float *Y, const float *T, const int *mask, int N, int count, float clip, float *dY, float *loss,
                float *scratch, void *stream);

/* dmph_nn_mse with no scratch argument and no allocation of any kind: the
 * per-block loss partials live in a fixed device array inside the library
 * (256 KB per device, held until the process ends), so N is at most
 * DMPH_NN_MSE_STATIC_MAX_N; larger N returns -4 and launches nothing.  No
 * counterpart in the reference: this entry is for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  Every call on one device shares the one
 * array, so calls that overlap on a device (two streams, two threads) race;
 * serialise them.  Same results as dmph_nn_mse, bit for bit. */
#define DMPH_NN_MSE_STATIC_MAX_N (65536 * 256)
int dmph_nn_mse_static(const float *Y, const float *T, const int *mask, int N, int count, float clip, float *dY,
                       float *loss, void *stream);

int dmph_nn_relu_bwd(const uint16_t *Y, int ldy, const uint16_t *dY, int lddy, uint16_t *dX, int lddx, int M, int n,
                     void *stream);
int dmph_nn_f2bf(const float *x, int ldx, uint16_t *y, int ldy, int M, int n, void *stream);
int dmph_nn_bf2f(const uint16_t *x, int ldx, float *y, int ldy, int M, int n, int accumulate, void *stream);

/* One rollout step of stride h from predicted (scaled) accelerations Y [N][3]:
 * fluid (class 0) v += h Y / ascale, debris (class 1) the mean over its body
 * (body ids < nbodies, sums in body_sums [nbodies][4]), walls (class 2) fixed;
 * x += h v; the velocity history [5][N][3] shifts by one (newest first). */
int dmph_nn_rollout_step(float *pos, float *vel_hist, const float *Y, float ascale, const int *cls, const int *body,
                         float *body_sums, int nbodies, float h, int N, void *stream);

/* Adam step `step` (1-based) on FP32 master weights w[rows][cols]; refreshes the
 * bf16 copy wb [rows][cols] and transpose wbt [cols][rows] when not NULL */
int dmph_nn_adam(float *w, const float *g, float *m, float *v, int n, float lr, float b1, float b2, float eps, int step,
                 float wd, uint16_t *wb, uint16_t *wbt, int rows, int cols, void *stream);

#ifdef __cplusplus
}
#endif
#endif
