// Message-passing surrogate for the DEM-MPH step: HIP kernels for gfx950.
//
// Data layout: node-major rows of features, bf16 activations (raw 16-bit
// storage, FP32 arithmetic), FP32 master weights and gradients.  A weight W of
// a layer with K inputs and N outputs is stored [N][K] (row n contiguous over
// k), so the forward GEMM and, with the transposed copy [K][N], the input
// gradient GEMM read their B operand contiguously along the reduction.
//
// Kernels:
//   gemm      Y = act(X W^T + b [+ R]) on bf16 MFMA (16x16x16 or gfx950 16x16x32),
//             FP32 accumulate, one 128-column block of Y per workgroup column
//   gemm_dw   dW = dY^T X and db = sum dY: per-workgroup FP32 partials over a
//             chunk of rows, summed by reduce_rows (deterministic, no atomics)
//   edge_fwd  S_i = sum_j relu(A_i + B_j + Wc eps_ij + bc) over the radius graph
//   edge_bwd  dA, dB (from the symmetric graph: no atomics), dWc, dbc partials
//   graph     uniform-grid counting sort + CSR radius graph, deterministic order
//   features, head, loss, Adam, ReLU backward, bf16 casts
//
// All entry points are extern "C" with plain device pointers and a stream.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <mutex>
#include <type_traits>
#include <stdint.h>

#include "dem_mph_nn.h"

namespace nn {

typedef uint16_t bf16_t;   // storage

__device__ __forceinline__ float bf2f(bf16_t v) { return __uint_as_float(((uint32_t)v) << 16); }
// round to nearest even (NaN stays NaN): gfx950's v_cvt_pk_bf16_f32, two values per instruction
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
typedef _Float16 h2_t __attribute__((ext_vector_type(2)));
typedef float f32x2_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t pk_bf16(float lo, float hi) {
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(f32x2_t{lo, hi}, bf16x2_t));
}
__device__ __forceinline__ bf16_t f2bf(float f) { return __builtin_bit_cast(bf16_t, (__bf16)f); }
// the same rounding in integer arithmetic (one value at a time: fewer live registers)
__device__ __forceinline__ uint32_t pk_bf16_int(float lo, float hi) {
    auto one = [](float f) {
        uint32_t u = __float_as_uint(f);
        if ((u & 0x7f800000u) == 0x7f800000u) return u >> 16 | ((u & 0xffffu) ? 0x40u : 0u);   // inf / nan
        return (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
    };
    return one(lo) | one(hi) << 16;
}

typedef short s4_t __attribute__((ext_vector_type(4)));
typedef __bf16 b8_t __attribute__((ext_vector_type(8)));
typedef float f4_t __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------------------------
// GEMM: Y[M x ldy-strided][n0 .. n0+128) = act(X[M x K] W[N][K]^T + b [+ R])
// Workgroup: 4 waves, 128 rows; each wave 32 rows x 128 columns = 2 x 8 MFMA
// tiles of 16x16.  W's 128-row block sits in LDS (K <= 256: at most 64 KB).
// KS: reduction step of the instruction, 16 (v_mfma_f32_16x16x16_bf16) or 32
// (v_mfma_f32_16x16x32_bf16, gfx950).  Lane l of a 16x16xKS tile holds A row
// l % 16, k = (KS/4) * (l / 16) + [0, KS/4); B column l % 16, same k; D column
// l % 16, rows 4 * (l / 16) + [0, 4).  The tiles are computed transposed
// (A = W, B = X: D rows are output columns), so a lane ends with 4 adjacent
// outputs of one row and stores them as one 8-byte word.
// ---------------------------------------------------------------------------------------------

// Workgroups are handed to the 8 XCDs (each with its own L2) round-robin by
// index.  For gather kernels whose neighbouring nodes share data, this maps the
// launch's G workgroups so that each XCD takes one contiguous range of them.
__device__ __forceinline__ int xcd_block(int G) {
    const int w = blockIdx.x, q = G >> 3, r = G & 7, x = w & 7;
    return x * q + min(x, r) + (w >> 3);
}

constexpr int kCols = 128;

// LDS position of W[r][k]: 16-byte granule k / 8 of row r is stored at granule
// (k / 8) ^ (r & m), m = min(K / 8, 8) - 1, so the rows an MFMA fragment reads
// at one k fall in different banks; the XOR never leaves the row.
__device__ __forceinline__ int wl_at(int r, int k, int K) {
    const int m = (K >= 64 ? 8 : K / 8) - 1;
    return r * K + (((k >> 3) ^ (r & m)) << 3) + (k & 7);
}

// Workgroup shape, measured on the 10M-row surrogate layers (ms per layer,
// K x N = 128 x 256 / 256 x 128 / 128 x 128 + residual): 2 row blocks per wave,
// 4 waves, 1 column block: 3.08 / 3.05 / 2.14; 1, 8, 1: 2.91 / 2.16 / 2.00;
// 1, 8, 2: 2.18 / 2.18 / 2.02 (X read once for both halves of N = 256).
#ifndef DMPH_GEMM_RB
#define DMPH_GEMM_RB 1   // 16-row blocks per wave
#endif
#ifndef DMPH_GEMM_NW
#define DMPH_GEMM_NW 8   // waves per workgroup
#endif
#ifndef DMPH_GEMM_NC
#define DMPH_GEMM_NC 2   // 128-column blocks per workgroup when N = 256 (and K <= 128)
#endif

// HEAD: the layer is the decoder's hidden layer (N = 128, ReLU); instead of
// storing it, each row's bf16-rounded values go through the 3-output head,
// Y3[r] = Wd g_r + bd (FP32), summed across the four lanes that hold the row.
// PF (persistent): a workgroup stages W once and walks row tiles blockIdx.x,
// + gridDim.x, ..., the next tile's X fragments requested before the current
// epilogue (more registers; measured faster for N = 128, K >= 128 and slower
// for N = 256); else one tile per workgroup.
template <int KS, int K, int RB, int NW, int NC, bool HEAD = false, bool PF = false>
__global__ void __launch_bounds__(NW * 64) k_gemm(const bf16_t *__restrict__ X, int ldx, const bf16_t *__restrict__ W,
                                                  const float *__restrict__ bias, const bf16_t *__restrict__ R, int ldr,
                                                  bf16_t *__restrict__ Y, int ldy, int M, int flags,
                                                  const float *__restrict__ Wd = nullptr,
                                                  const float *__restrict__ bd = nullptr, float *__restrict__ Y3 = nullptr) {
    extern __shared__ bf16_t wl[];   // [NC * 128][K], 16-byte granules XOR-swizzled by row (wl_at)
    constexpr int NK = K / KS, CT = 8 * NC;   // k steps, 16-column tiles
    constexpr int kTileRows = NW * RB * 16;
    typedef typename std::conditional<KS == 32, uint4, uint2>::type frag_t;
    const int n0 = blockIdx.y * kCols * NC;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int lr = lane & 15, lk = (lane >> 4) * (KS / 4);
    const int ntiles = (M + kTileRows - 1) / kTileRows;
    const bool relu = flags & DMPH_NN_RELU, half = flags & DMPH_NN_F16_OUT;
    int tile = blockIdx.x;
    // every X fragment of the wave's 16-row blocks (rows past M read as zero);
    // the first tile's are requested before W is staged, so the latencies overlap
    frag_t xf[NK][RB];
    auto load_x = [&](int row0) {
#pragma unroll
        for (int a = 0; a < RB; a++) {
            const int r = row0 + a * 16 + lr;
#pragma unroll
            for (int kk = 0; kk < NK; kk++) {
                xf[kk][a] = frag_t{};
                if (r < M) xf[kk][a] = *reinterpret_cast<const frag_t *>(X + (size_t)r * ldx + kk * KS + lk);
            }
        }
    };
    load_x((tile * NW + wave) * RB * 16);
    for (int t = threadIdx.x; t < CT * 16 * K / 8; t += blockDim.x) {   // 16-byte copies
        const int r = t / (K / 8), c = (t % (K / 8)) * 8;
        *reinterpret_cast<uint4 *>(wl + wl_at(r, c, K)) = *reinterpret_cast<const uint4 *>(W + (size_t)(n0 + r) * K + c);
    }
    __syncthreads();
    while (true) {
        const int row0 = (tile * NW + wave) * RB * 16;
        // the LDS offset is re-derived per tile: else the compiler keeps every
        // W fragment of the block live across tiles in VGPRs
        int lbase = lr;
        asm volatile("" : "+v"(lbase));
        f4_t acc[RB][CT];
#pragma unroll
        for (int a = 0; a < RB; a++)
#pragma unroll
            for (int c = 0; c < CT; c++) acc[a][c] = f4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kk = 0; kk < NK; kk++) {
            const int k = kk * KS;
#pragma unroll
            for (int c = 0; c < CT; c++) {
                const frag_t v = *reinterpret_cast<const frag_t *>(wl + wl_at(c * 16 + lbase, k + lk, K));
#pragma unroll
                for (int a = 0; a < RB; a++) {
                    if constexpr (KS == 16)
                        acc[a][c] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(__builtin_bit_cast(s4_t, v),
                                                                              __builtin_bit_cast(s4_t, xf[kk][a]), acc[a][c], 0, 0, 0);
                    else
                        acc[a][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(b8_t, v),
                                                                            __builtin_bit_cast(b8_t, xf[kk][a]), acc[a][c], 0, 0, 0);
                }
            }
        }
        const int next = tile + (int)gridDim.x;
        if (PF && next < ntiles) load_x((next * NW + wave) * RB * 16);   // the next tile's rows, under this epilogue
        // epilogue: acc[a][c][e] is row row0 + 16a + lr, column n0 + 16c + 4 (lane / 16) + e
        if constexpr (HEAD) {
#pragma unroll
            for (int a = 0; a < RB; a++) {
                const int r = row0 + a * 16 + lr;
                float part[3] = {0.f, 0.f, 0.f};
#pragma unroll
                for (int c = 0; c < CT; c++) {
                    const int col = c * 16 + (lane >> 4) * 4;
                    const float4 b = *reinterpret_cast<const float4 *>(bias + col);
                    const float bb[4] = {b.x, b.y, b.z, b.w};
#pragma unroll
                    for (int o = 0; o < 3; o++) {
                        const float4 w = *reinterpret_cast<const float4 *>(Wd + o * 128 + col);
                        const float ww[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
                        for (int e = 0; e < 4; e++) part[o] += bf2f(f2bf(fmaxf(acc[a][c][e] + bb[e], 0.f))) * ww[e];
                    }
                }
#pragma unroll
                for (int o = 0; o < 3; o++) {
                    part[o] += __shfl_xor(part[o], 16);
                    part[o] += __shfl_xor(part[o], 32);
                }
                if (lane < 16 && r < M)
#pragma unroll
                    for (int o = 0; o < 3; o++) Y3[3 * (size_t)r + o] = part[o] + bd[o];
            }
        } else {
            uint2 rr[RB][CT];   // residual words, all requested before the first is used
#pragma unroll
            for (int a = 0; a < RB; a++)
#pragma unroll
                for (int c = 0; c < CT; c++) {
                    const int r = row0 + a * 16 + lr;
                    rr[a][c] = make_uint2(0, 0);
                    if (R && r < M)
                        rr[a][c] = *reinterpret_cast<const uint2 *>(R + (size_t)r * ldr + n0 + c * 16 + (lane >> 4) * 4);
                }
#pragma unroll
            for (int c = 0; c < CT; c++) {
                const int col = n0 + c * 16 + (lane >> 4) * 4;
                const float4 b = bias ? *reinterpret_cast<const float4 *>(bias + col) : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
                for (int a = 0; a < RB; a++) {
                    const int r = row0 + a * 16 + lr;
                    if (r >= M) continue;
                    float v[4] = {acc[a][c][0] + b.x, acc[a][c][1] + b.y, acc[a][c][2] + b.z, acc[a][c][3] + b.w};
                    if (R) {
                        const uint2 q = rr[a][c];
                        v[0] += bf2f((bf16_t)(q.x & 0xffff)); v[1] += bf2f((bf16_t)(q.x >> 16));
                        v[2] += bf2f((bf16_t)(q.y & 0xffff)); v[3] += bf2f((bf16_t)(q.y >> 16));
                    }
                    if (relu)
#pragma unroll
                        for (int e = 0; e < 4; e++) v[e] = fmaxf(v[e], 0.f);
                    uint2 o;
                    if (half) {
                        o.x = __builtin_bit_cast(uint32_t, h2_t{(_Float16)v[0], (_Float16)v[1]});
                        o.y = __builtin_bit_cast(uint32_t, h2_t{(_Float16)v[2], (_Float16)v[3]});
                    } else {
                        o.x = pk_bf16(v[0], v[1]);
                        o.y = pk_bf16(v[2], v[3]);
                    }
                    *reinterpret_cast<uint2 *>(Y + (size_t)r * ldy + col) = o;
                }
            }
        }
        if (!PF || next >= ntiles) break;
        tile = next;
    }
}

// Row tiles per launch: one workgroup per tile, or (persistent) two resident
// workgroups per CU walking the tiles.
static unsigned gemm_grid_x(int M, int rows, bool persistent) {
    const unsigned tiles = (unsigned)((M + rows - 1) / rows);
    const unsigned g = 2u * 256u;
    return persistent && tiles > g ? g : tiles;
}

// ---------------------------------------------------------------------------------------------
// One round's node update at inference with the intermediates in registers:
//   u  = relu(W1 [h | S] + b1)     W1 [128][256] (the folded node layer)
//   h' = h + W2 u + b2             W2 [128][128]; h' stored bf16 into Hn
//   AB = Wab h'                    Wab [256][128], stored IEEE half (the next round's [A | B])
// All three weights sit in LDS (160 KB, one workgroup of 16 waves per CU, each
// wave 16 rows of a 256-row tile, tiles walked persistently).  An MFMA result
// tile c leaves lane (r, g) with the outputs of staged weight rows 16c + 4g +
// [0, 4) for row r.  The weights are staged with their rows permuted (rho) so
// that these are features 32 (c / 2) + 8g + 4 (c % 2) + [0, 4): tiles 2s and
// 2s + 1 give lane g the 8 adjacent features 32s + 8g + [0, 8).  That is the
// k order the next product's B operand takes at k step s (no column
// permutation), and every store or residual load is 16 bytes of a row per lane,
// 64 contiguous bytes per row over its four lanes.
// ---------------------------------------------------------------------------------------------

constexpr int kNodeWaves = 16;
constexpr size_t kNodeLds = (size_t)(128 * 256 + 128 * 128 + 256 * 128) * sizeof(bf16_t);   // 160 KB

// output feature of staged weight row 16c + m (row permutation of the node kernels)
__device__ __forceinline__ int rho(int c, int m) { return 32 * (c >> 1) + 8 * (m >> 2) + 4 * (c & 1) + (m & 3); }
// first of the 4 features lane group g holds from tile c
__device__ __forceinline__ int rho_base(int c, int g) { return 32 * (c >> 1) + 8 * g + 4 * (c & 1); }

// DEC (the last round): the third product is the decoder's hidden layer
// instead, g = relu(Wab h' + bab) with Wab [128][128], and its bf16-rounded
// values go through the head, Y3 = Wd2 g + bd2 (FP32 [M][3]); h' is not stored.
// FOLD: the B half of AB gets + Wc3 P of the row (NnFold; Wt [3][128] read per
// tile through an opaque offset, so the compiler keeps it out of registers).
template <bool DEC = false, int NWV = DEC ? 8 : kNodeWaves, bool FOLD = false>   // DEC: 8 waves (no spills)
__global__ void __launch_bounds__(NWV * 64) k_node(const bf16_t *__restrict__ HS, int ldhs,
                                                          const bf16_t *__restrict__ W1, const float *__restrict__ b1,
                                                          const bf16_t *__restrict__ W2, const float *__restrict__ b2,
                                                          const bf16_t *__restrict__ Wab, bf16_t *__restrict__ Hn,
                                                          int ldhn, bf16_t *__restrict__ ABo, int M,
                                                          const float *__restrict__ bab = nullptr,
                                                          const float *__restrict__ Wd2 = nullptr,
                                                          const float *__restrict__ bd2 = nullptr,
                                                          float *__restrict__ Y3 = nullptr, NnFold fold = NnFold{}) {
    constexpr int NAB = DEC ? 128 : 256;   // rows of the third weight
    extern __shared__ bf16_t wl[];
    bf16_t *w1 = wl, *w2 = wl + 128 * 256, *wab = w2 + 128 * 128;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int lr = lane & 15, g = lane >> 4, lk = 8 * g;
    const int ntiles = (M + NWV * 16 - 1) / (NWV * 16);
    int tile = blockIdx.x;
    typedef uint4 frag_t;
    frag_t xf[8];
    auto load_x = [&](int row0) {
        const int r = row0 + lr;
#pragma unroll
        for (int kk = 0; kk < 8; kk++) {
            xf[kk] = frag_t{0, 0, 0, 0};
            if (r < M) xf[kk] = *reinterpret_cast<const frag_t *>(HS + (size_t)r * ldhs + kk * 32 + lk);
        }
    };
    load_x((tile * NWV + wave) * 16);
    for (int t = threadIdx.x; t < 128 * 256 / 8; t += blockDim.x) {   // W1, rows in rho order
        const int r = t / 32, c = (t % 32) * 8;
        *reinterpret_cast<uint4 *>(w1 + wl_at(r, c, 256)) =
            *reinterpret_cast<const uint4 *>(W1 + (size_t)rho(r >> 4, r & 15) * 256 + c);
    }
    for (int t = threadIdx.x; t < (128 + NAB) * 128 / 8; t += blockDim.x) {   // W2 and Wab, rows in rho order
        const int r = t / 16, c = (t % 16) * 8;
        const int rr = r < 128 ? r : r - 128, q = rr & 127;   // Wab: rho within each half of 128 rows
        const int src_row = (rr & ~127) + rho(q >> 4, q & 15);
        const bf16_t *src = (r < 128 ? W2 : Wab) + (size_t)src_row * 128;
        *reinterpret_cast<uint4 *>((r < 128 ? w2 : wab) + wl_at(rr, c, 128)) =
            *reinterpret_cast<const uint4 *>(src + c);
    }
    __syncthreads();
    while (true) {
        const int row0 = (tile * NWV + wave) * 16, r = row0 + lr;
        int lb = lr;
        asm volatile("" : "+v"(lb));   // keep the LDS fragment reads inside the tile loop
        f4_t acc[8];
        // u = relu(W1 [h | S] + b1)
#pragma unroll
        for (int c = 0; c < 8; c++) acc[c] = f4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kk = 0; kk < 8; kk++)
#pragma unroll
            for (int c = 0; c < 8; c++) {
                const frag_t v = *reinterpret_cast<const frag_t *>(w1 + wl_at(c * 16 + lb, kk * 32 + lk, 256));
                acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(b8_t, v),
                                                                 __builtin_bit_cast(b8_t, xf[kk]), acc[c], 0, 0, 0);
            }
        uint32_t uf[8][2];   // tile c of u, 4 bf16 = 2 words
#pragma unroll
        for (int c = 0; c < 8; c++) {
            const float4 b = *reinterpret_cast<const float4 *>(b1 + rho_base(c, g));
            const float v0 = fmaxf(acc[c][0] + b.x, 0.f), v1 = fmaxf(acc[c][1] + b.y, 0.f);
            const float v2 = fmaxf(acc[c][2] + b.z, 0.f), v3 = fmaxf(acc[c][3] + b.w, 0.f);
            uf[c][0] = pk_bf16_int(v0, v1);
            uf[c][1] = pk_bf16_int(v2, v3);
        }
        // the residual h in the output layout (16 bytes per tile pair), requested before the second product
        uint2 hr[8];
#pragma unroll
        for (int s2 = 0; s2 < 4; s2++) {
            uint4 v = make_uint4(0, 0, 0, 0);
            if (r < M) v = *reinterpret_cast<const uint4 *>(HS + (size_t)r * ldhs + 32 * s2 + 8 * g);
            hr[2 * s2] = make_uint2(v.x, v.y);
            hr[2 * s2 + 1] = make_uint2(v.z, v.w);
        }
        // h' = h + W2 u + b2
#pragma unroll
        for (int c = 0; c < 8; c++) acc[c] = f4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kk = 0; kk < 4; kk++) {
            const frag_t uk = make_uint4(uf[2 * kk][0], uf[2 * kk][1], uf[2 * kk + 1][0], uf[2 * kk + 1][1]);
#pragma unroll
            for (int c = 0; c < 8; c++) {
                const frag_t v = *reinterpret_cast<const frag_t *>(w2 + wl_at(c * 16 + lb, kk * 32 + lk, 128));
                acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(b8_t, v),
                                                                 __builtin_bit_cast(b8_t, uk), acc[c], 0, 0, 0);
            }
        }
        const int next = tile + (int)gridDim.x;
        if (!DEC && next < ntiles) load_x((next * NWV + wave) * 16);   // the next tile's rows, under what follows
        uint32_t hf[8][2];
#pragma unroll
        for (int c = 0; c < 8; c++) {
            const float4 b = *reinterpret_cast<const float4 *>(b2 + rho_base(c, g));
            const float v0 = acc[c][0] + b.x + bf2f((bf16_t)(hr[c].x & 0xffff));
            const float v1 = acc[c][1] + b.y + bf2f((bf16_t)(hr[c].x >> 16));
            const float v2 = acc[c][2] + b.z + bf2f((bf16_t)(hr[c].y & 0xffff));
            const float v3 = acc[c][3] + b.w + bf2f((bf16_t)(hr[c].y >> 16));
            hf[c][0] = pk_bf16(v0, v1);
            hf[c][1] = pk_bf16(v2, v3);
            if (!DEC && (c & 1) && r < M)   // tiles c - 1 and c: 8 adjacent features
                *reinterpret_cast<uint4 *>(Hn + (size_t)r * ldhn + rho_base(c - 1, g)) =
                    make_uint4(hf[c - 1][0], hf[c - 1][1], hf[c][0], hf[c][1]);
        }
        if constexpr (DEC) {   // g = relu(Wab h' + bab), then the 3-output head
#pragma unroll
            for (int c = 0; c < 8; c++) acc[c] = f4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kk = 0; kk < 4; kk++) {
                const frag_t hk = make_uint4(hf[2 * kk][0], hf[2 * kk][1], hf[2 * kk + 1][0], hf[2 * kk + 1][1]);
#pragma unroll
                for (int c = 0; c < 8; c++) {
                    const frag_t v = *reinterpret_cast<const frag_t *>(wab + wl_at(c * 16 + lb, kk * 32 + lk, 128));
                    acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(b8_t, v),
                                                                     __builtin_bit_cast(b8_t, hk), acc[c], 0, 0, 0);
                }
            }
            float part[3] = {0.f, 0.f, 0.f};
#pragma unroll
            for (int c = 0; c < 8; c++) {
                // (lb - lr) = 0, opaque: the head's weights are re-read per tile, not held in registers
                const int col = rho_base(c, g) + (lb - lr);
                const float4 b = *reinterpret_cast<const float4 *>(bab + col);
                const float bb[4] = {b.x, b.y, b.z, b.w};
#pragma unroll
                for (int o = 0; o < 3; o++) {
                    const float4 w = *reinterpret_cast<const float4 *>(Wd2 + o * 128 + col);
                    const float ww[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
                    for (int e = 0; e < 4; e++) part[o] += bf2f(f2bf(fmaxf(acc[c][e] + bb[e], 0.f))) * ww[e];
                }
            }
#pragma unroll
            for (int o = 0; o < 3; o++) {
                part[o] += __shfl_xor(part[o], 16);
                part[o] += __shfl_xor(part[o], 32);
            }
            if (lane < 16 && r < M)
#pragma unroll
                for (int o = 0; o < 3; o++) Y3[3 * (size_t)r + o] = part[o] + bd2[o];
        }
        // AB = Wab h', in two halves of 128 outputs
        uint2 ab2[8];
        float px = 0.f, py = 0.f, pz = 0.f;
        if (FOLD && r < M) {
            px = (fold.pos[3 * (size_t)r] - fold.origin[0]) * fold.inv_r;
            py = (fold.pos[3 * (size_t)r + 1] - fold.origin[1]) * fold.inv_r;
            pz = (fold.pos[3 * (size_t)r + 2] - fold.origin[2]) * fold.inv_r;
        }
#pragma unroll
        for (int half = 0; half < (DEC ? 0 : 2); half++) {
#pragma unroll
            for (int c = 0; c < 8; c++) acc[c] = f4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kk = 0; kk < 4; kk++) {
                const frag_t hk = make_uint4(hf[2 * kk][0], hf[2 * kk][1], hf[2 * kk + 1][0], hf[2 * kk + 1][1]);
#pragma unroll
                for (int c = 0; c < 8; c++) {
                    const frag_t v =
                        *reinterpret_cast<const frag_t *>(wab + wl_at(half * 128 + c * 16 + lb, kk * 32 + lk, 128));
                    acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(b8_t, v),
                                                                     __builtin_bit_cast(b8_t, hk), acc[c], 0, 0, 0);
                }
            }
            if (r < M)
#pragma unroll
                for (int c = 0; c < 8; c++) {
                    if (FOLD && half == 1) {   // + Wc3 P of the row
                        const int col = rho_base(c, g) + (lb - lr);
                        const float4 wx = *reinterpret_cast<const float4 *>(fold.Wt + col);
                        const float4 wy = *reinterpret_cast<const float4 *>(fold.Wt + 128 + col);
                        const float4 wz = *reinterpret_cast<const float4 *>(fold.Wt + 256 + col);
                        acc[c][0] += wx.x * px + wy.x * py + wz.x * pz;
                        acc[c][1] += wx.y * px + wy.y * py + wz.y * pz;
                        acc[c][2] += wx.z * px + wy.z * py + wz.z * pz;
                        acc[c][3] += wx.w * px + wy.w * py + wz.w * pz;
                    }
                    ab2[c] = make_uint2(__builtin_bit_cast(uint32_t, h2_t{(_Float16)acc[c][0], (_Float16)acc[c][1]}),
                                        __builtin_bit_cast(uint32_t, h2_t{(_Float16)acc[c][2], (_Float16)acc[c][3]}));
                    if (c & 1)   // tiles c - 1 and c: 8 adjacent outputs
                        *reinterpret_cast<uint4 *>(ABo + (size_t)r * 256 + half * 128 + rho_base(c - 1, g)) =
                            make_uint4(ab2[c - 1].x, ab2[c - 1].y, ab2[c].x, ab2[c].y);
                }
        }
        if (next >= ntiles) break;
        tile = next;
        if (DEC) load_x((tile * NWV + wave) * 16);   // (registers: no prefetch under the decoder)
    }
}

// ---------------------------------------------------------------------------------------------
// dW[N][K] = dY^T X, db[N] = sum_m dY: workgroup g sums rows [g*chunk, (g+1)*chunk) in
// 64-row tiles staged transposed in LDS (dYt[n][m], Xt[k][m]) and writes FP32
// partials P[g][N][K] and Pb[g][N]; reduce_rows adds the partials.  The output
// block of a workgroup is 128 (n) x K (k <= 256): wave w owns n rows
// [32w, 32w + 32) as 2 x (K/16) MFMA tiles.
// ---------------------------------------------------------------------------------------------

constexpr int kTileM = 64;

template <int KT>   // K / 16
__global__ void __launch_bounds__(256) k_gemm_dw(const bf16_t *__restrict__ dY, int lddy, const bf16_t *__restrict__ X,
                                                 int ldx, float *__restrict__ P, float *__restrict__ Pb, int M, int chunk) {
    constexpr int K = KT * 16;
    __shared__ bf16_t dyt[kCols][kTileM + 4];
    __shared__ bf16_t xt[K][kTileM + 4];
    const int n0 = blockIdx.y * kCols;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int lr = lane & 15, lk = (lane >> 4) * 4;
    f4_t acc[2][KT];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int c = 0; c < KT; c++) acc[a][c] = f4_t{0.f, 0.f, 0.f, 0.f};
    float bsum = 0.f;   // thread t < 128: column n0 + t
    const int m_beg = blockIdx.x * chunk, m_end = min(M, m_beg + chunk);
    for (int m0 = m_beg; m0 < m_end; m0 += kTileM) {
        __syncthreads();
        for (int t = threadIdx.x; t < kTileM * kCols; t += blockDim.x) {
            const int m = t / kCols, n = t % kCols;
            dyt[n][m] = (m0 + m < m_end) ? dY[(size_t)(m0 + m) * lddy + n0 + n] : (bf16_t)0;
        }
        for (int t = threadIdx.x; t < kTileM * K; t += blockDim.x) {
            const int m = t / K, k = t % K;
            xt[k][m] = (m0 + m < m_end) ? X[(size_t)(m0 + m) * ldx + k] : (bf16_t)0;
        }
        __syncthreads();
        if (threadIdx.x < kCols)
            for (int m = 0; m < kTileM; m++) bsum += bf2f(dyt[threadIdx.x][m]);
#pragma unroll
        for (int ms = 0; ms < kTileM; ms += 16) {
            s4_t af[2];
#pragma unroll
            for (int a = 0; a < 2; a++) {
                const uint2 v = *reinterpret_cast<const uint2 *>(&dyt[wave * 32 + a * 16 + lr][ms + lk]);
                af[a] = s4_t{(short)(v.x & 0xffff), (short)(v.x >> 16), (short)(v.y & 0xffff), (short)(v.y >> 16)};
            }
#pragma unroll
            for (int c = 0; c < KT; c++) {
                const uint2 v = *reinterpret_cast<const uint2 *>(&xt[c * 16 + lr][ms + lk]);
                const s4_t bf = s4_t{(short)(v.x & 0xffff), (short)(v.x >> 16), (short)(v.y & 0xffff), (short)(v.y >> 16)};
#pragma unroll
                for (int a = 0; a < 2; a++) acc[a][c] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(af[a], bf, acc[a][c], 0, 0, 0);
            }
        }
    }
    const size_t Nall = (size_t)gridDim.y * kCols;
    float *pg = P + (size_t)blockIdx.x * Nall * K;
#pragma unroll
    for (int c = 0; c < KT; c++) {
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int n = n0 + wave * 32 + a * 16 + (lane >> 4) * 4 + e;
                const int k = c * 16 + lr;
                pg[(size_t)n * K + k] = acc[a][c][e];
            }
    }
    if (Pb && threadIdx.x < kCols) Pb[(size_t)blockIdx.x * Nall + n0 + threadIdx.x] = bsum;
}

// out[j] (+)= sum_g P[g * stride + j] for short outputs over many partials: one
// block per output, threads take g = t, t + 256, ..., then a fixed tree (deterministic)
__global__ void __launch_bounds__(256) k_reduce_cols(const float *__restrict__ P, int groups, size_t stride,
                                                     float *__restrict__ out, int accumulate) {
    __shared__ float red[256];
    const size_t j = blockIdx.x;
    float s = 0.f;
    for (int g = threadIdx.x; g < groups; g += blockDim.x) s += P[(size_t)g * stride + j];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[j] = accumulate ? out[j] + red[0] : red[0];
}

// out[j] (+)= sum_g P[g * stride + j] for j < len (FP32 partials, in group order)
__global__ void k_reduce_rows(const float *__restrict__ P, int groups, size_t len, size_t stride, float *__restrict__ out,
                              int accumulate) {
    const size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= len) return;
    float s = 0.f;
    for (int g = 0; g < groups; g++) s += P[(size_t)g * stride + j];
    out[j] = accumulate ? out[j] + s : s;
}

// ---------------------------------------------------------------------------------------------
// Edge messages.  One wave per node i; lane l holds features 2l, 2l+1.
//   S_i = sum_{j in row i} relu(A_i + B_j + Wc eps_ij + bc),  eps_ij = ((x_j - x_i)/R, |x_j - x_i|/R)
// AB: [N][256] = [A | B] (bf16), Wc: [128][4] FP32, bc: [128].
// ---------------------------------------------------------------------------------------------

__device__ __forceinline__ void edge_eps(const float *pos, int i, int j, float inv_r, float e[4]) {
    const float dx = pos[3 * j] - pos[3 * i], dy = pos[3 * j + 1] - pos[3 * i + 1], dz = pos[3 * j + 2] - pos[3 * i + 2];
    e[0] = dx * inv_r; e[1] = dy * inv_r; e[2] = dz * inv_r;
    e[3] = sqrtf(dx * dx + dy * dy + dz * dz) * inv_r;
}

// Edges are taken 16 at a time: lanes 0..15 load the batch's neighbour indices
// and edge features, which are broadcast (readlane: uniform, in scalar
// registers), and all 16 B-row loads are in flight before the arithmetic -- one
// memory round trip per 16 edges instead of a dependent chain per edge.  The
// sum still runs over the row in order.
constexpr int kEdgeBatch = 16;

__device__ __forceinline__ int rl(int v, int l) { return __builtin_amdgcn_readlane(v, l); }
__device__ __forceinline__ float rlf(float v, int l) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l)); }

__global__ void __launch_bounds__(256) k_edge_fwd(const bf16_t *__restrict__ AB, const float *__restrict__ pos,
                                                  const int *__restrict__ rp, const int *__restrict__ ci,
                                                  const float *__restrict__ Wc, const float *__restrict__ bc,
                                                  bf16_t *__restrict__ S, int lds, int N, float inv_r) {
    const int i = xcd_block(gridDim.x) * 4 + (threadIdx.x >> 6);
    if (i >= N) return;
    const int lane = threadIdx.x & 63, f = 2 * lane;
    const float a0 = bf2f(AB[(size_t)i * 256 + f]), a1 = bf2f(AB[(size_t)i * 256 + f + 1]);
    const float4 w0 = *reinterpret_cast<const float4 *>(Wc + 4 * f), w1 = *reinterpret_cast<const float4 *>(Wc + 4 * f + 4);
    const float b0 = bc[f], b1 = bc[f + 1];
    const float xi = pos[3 * i], yi = pos[3 * i + 1], zi = pos[3 * i + 2];
    float s0 = 0.f, s1 = 0.f;
    const int q0 = rp[i], q1 = rp[i + 1];
    for (int q = q0; q < q1; q += kEdgeBatch) {
        const int m = min(kEdgeBatch, q1 - q);
        // lane l < m: edge q + l
        int jl = i;
        float e0 = 0.f, e1 = 0.f, e2 = 0.f, e3 = 0.f;
        if (lane < m) {
            jl = ci[q + lane];
            const float dx = pos[3 * jl] - xi, dy = pos[3 * jl + 1] - yi, dz = pos[3 * jl + 2] - zi;
            e0 = dx * inv_r; e1 = dy * inv_r; e2 = dz * inv_r;
            e3 = sqrtf(dx * dx + dy * dy + dz * dz) * inv_r;
        }
        uint32_t bj[kEdgeBatch];
#pragma unroll
        for (int t = 0; t < kEdgeBatch; t++) {
            const int j = rl(jl, t);   // i itself past the row's end: a valid row, never summed
            bj[t] = *reinterpret_cast<const uint32_t *>(AB + (size_t)j * 256 + 128 + f);
        }
#pragma unroll
        for (int t = 0; t < kEdgeBatch; t++) {
            if (t < m) {
                const float x0 = rlf(e0, t), x1 = rlf(e1, t), x2 = rlf(e2, t), x3 = rlf(e3, t);
                const float c0 = a0 + bf2f((bf16_t)(bj[t] & 0xffff)) + b0 + w0.x * x0 + w0.y * x1 + w0.z * x2 + w0.w * x3;
                const float c1 = a1 + bf2f((bf16_t)(bj[t] >> 16)) + b1 + w1.x * x0 + w1.y * x1 + w1.z * x2 + w1.w * x3;
                s0 += fmaxf(c0, 0.f);
                s1 += fmaxf(c1, 0.f);
            }
        }
    }
    *reinterpret_cast<uint32_t *>(S + (size_t)i * lds + f) = pk_bf16(s0, s1);
}


__device__ __forceinline__ uint32_t h2bits(h2_t v) { return __builtin_bit_cast(uint32_t, v); }
__device__ __forceinline__ h2_t bitsh2(uint32_t v) { return __builtin_bit_cast(h2_t, v); }

// one wave per node, lanes over the row's edges: eps[2q] = (dx, dy) / r, eps[2q + 1] = (dz, |d|) / r
__global__ void __launch_bounds__(256) k_edge_eps(const float *__restrict__ pos, const int *__restrict__ rp,
                                                  const int *__restrict__ ci, uint2 *__restrict__ eps, int N,
                                                  float inv_r) {
    const int i = xcd_block(gridDim.x) * 4 + (threadIdx.x >> 6);
    if (i >= N) return;
    const float xi = pos[3 * i], yi = pos[3 * i + 1], zi = pos[3 * i + 2];
    for (int q = rp[i] + (threadIdx.x & 63); q < rp[i + 1]; q += 64) {
        const int j = ci[q];
        const float dx = pos[3 * j] - xi, dy = pos[3 * j + 1] - yi, dz = pos[3 * j + 2] - zi;
        const float d = sqrtf(dx * dx + dy * dy + dz * dz);
        eps[q] = make_uint2(h2bits(h2_t{(_Float16)(dx * inv_r), (_Float16)(dy * inv_r)}),
                            h2bits(h2_t{(_Float16)(dz * inv_r), (_Float16)(d * inv_r)}));
    }
}

// Inference edge sums in packed IEEE half: lane l holds features 2l, 2l + 1 as
// one pair; per edge c = (A_i + bc) + B_j + sum_k Wc[:, k] eps_k is 1 add and 4
// fused multiply-adds on pairs, then max and add into a half partial that is
// added to the FP32 sums every 16 edges.  The node index, row bounds, column
// indices and edge features are wave-uniform (scalar loads); B_j is one 4-byte
// load per lane.
// BUF (rows of AB within 4 GB: N <= 8M nodes): B_j through one buffer
// descriptor per wave, the row offset j * 512 as the load's scalar offset, so
// a gather costs no vector instruction (else a 64-bit address add per edge).
// order (NULL or a permutation of the nodes, e.g. the graph's cell order):
// wave w sums node order[w], so the waves in flight read neighbouring rows of
// B from L2 instead of HBM; the sums do not depend on it.
// FOLD: B_j already holds + Wc3 P_j (P = (x - origin) / r), A_i gets - Wc3 P_i
// here, so Wc3 (x_j - x_i) / r needs no work per edge: 1 add, 1 multiply-add
// with the distance, max and add (4 instead of 7 instructions per edge).
template <bool BUF, bool FOLD>
__global__ void __launch_bounds__(256) k_edge_fwd_h(const h2_t *__restrict__ AB, const int *__restrict__ rp,
                                                    const int *__restrict__ ci, const uint2 *__restrict__ eps,
                                                    const float *__restrict__ Wc, const float *__restrict__ bc,
                                                    bf16_t *__restrict__ S, int lds, int N,
                                                    const int *__restrict__ deg, int row_cap,
                                                    const int *__restrict__ order, NnFold fold) {
    const int w = __builtin_amdgcn_readfirstlane(xcd_block(gridDim.x) * 4 + (threadIdx.x >> 6));
    if (w >= N) return;
    const int i = order ? __builtin_amdgcn_readfirstlane(order[w]) : w;
    const int lane = threadIdx.x & 63, f = 2 * lane;
    const float4 w0 = *reinterpret_cast<const float4 *>(Wc + 4 * f), w1 = *reinterpret_cast<const float4 *>(Wc + 4 * f + 4);
    const h2_t wx = {(_Float16)w0.x, (_Float16)w1.x}, wy = {(_Float16)w0.y, (_Float16)w1.y};
    const h2_t wz = {(_Float16)w0.z, (_Float16)w1.z}, wd = {(_Float16)w0.w, (_Float16)w1.w};
    const h2_t ai = AB[(size_t)i * 128 + lane];
    float a0 = (float)ai.x + bc[f], a1 = (float)ai.y + bc[f + 1];
    if constexpr (FOLD) {
        const float px = (fold.pos[3 * i] - fold.origin[0]) * fold.inv_r;
        const float py = (fold.pos[3 * i + 1] - fold.origin[1]) * fold.inv_r;
        const float pz = (fold.pos[3 * i + 2] - fold.origin[2]) * fold.inv_r;
        a0 -= w0.x * px + w0.y * py + w0.z * pz;
        a1 -= w1.x * px + w1.y * py + w1.z * pz;
    }
    const h2_t a2 = {(_Float16)a0, (_Float16)a1};
    const h2_t zero = {(_Float16)0.f, (_Float16)0.f};
    const uint32_t lane_bytes = 4u * lane;
    const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<char *>(reinterpret_cast<const char *>(AB) + 256), 0, -1, 0x00020000);
    // B_j pair of this lane
    auto b_of = [&](int j) {
        if constexpr (BUF) {
            // the row offset in unsigned arithmetic: up to 4 GB (N * 512 < 2^32 is checked
            // by the launcher), past the int range for j >= 2^22; the bits are what the
            // scalar offset register takes
            return bitsh2(__builtin_amdgcn_raw_buffer_load_b32(rsrc, (int)lane_bytes,
                                                               (int)((uint32_t)j << 9), 0));
        } else {
            const char *row = reinterpret_cast<const char *>(AB) + (size_t)j * 512 + 256;
            return *reinterpret_cast<const h2_t *>(row + lane_bytes);
        }
    };
    float s0 = 0.f, s1 = 0.f;
    int q = deg ? i * row_cap : rp[i];   // rows of a fixed capacity (row_cap entries, deg used) or CSR
    const int q1 = deg ? q + deg[i] : rp[i + 1];
    auto edge = [&](h2_t bj, uint2 e, h2_t acc) {
        const h2_t exy = bitsh2(e.x), ezd = bitsh2(e.y);
        h2_t c = a2 + bj;
        c = __builtin_elementwise_fma(wx, __builtin_shufflevector(exy, exy, 0, 0), c);
        c = __builtin_elementwise_fma(wy, __builtin_shufflevector(exy, exy, 1, 1), c);
        c = __builtin_elementwise_fma(wz, __builtin_shufflevector(ezd, ezd, 0, 0), c);
        c = __builtin_elementwise_fma(wd, __builtin_shufflevector(ezd, ezd, 1, 1), c);
        return acc + __builtin_elementwise_maximum(c, zero);   // NaN-propagating: an overflow stays visible
    };
    // FOLD: eps holds one half distance per entry, read as pairs (rows start at
    // even entries and every batch but the last single edge has an even size)
    auto edge_fold = [&](h2_t bj, h2_t dp, bool odd, h2_t acc) {
        const h2_t c = __builtin_elementwise_fma(wd, odd ? __builtin_shufflevector(dp, dp, 1, 1)
                                                         : __builtin_shufflevector(dp, dp, 0, 0), a2 + bj);
        return acc + __builtin_elementwise_maximum(c, zero);
    };
    // batches of B edges: all rows requested, then the chains (which interleave)
    auto batch = [&](auto nb) {
        constexpr int B = decltype(nb)::value;
        h2_t bj[B];
#pragma unroll
        for (int t = 0; t < B; t++) bj[t] = b_of(ci[q + t]);
        h2_t acc = zero;
        if constexpr (FOLD) {
            const uint32_t *dq = reinterpret_cast<const uint32_t *>(eps) + (q >> 1);
#pragma unroll
            for (int t = 0; t < B; t++) acc = edge_fold(bj[t], bitsh2(dq[t >> 1]), (t & 1) != 0, acc);
        } else {
#pragma unroll
            for (int t = 0; t < B; t++) acc = edge(bj[t], eps[q + t], acc);
        }
        s0 += (float)acc.x;
        s1 += (float)acc.y;
        q += B;
    };
    // (32 edges per round trip measured no faster)
    while (q + kEdgeBatch <= q1) batch(std::integral_constant<int, kEdgeBatch>{});
    // the rest (< 16) in batches of 8, 4, 2, 1: at most four round trips
    if (q + 8 <= q1) batch(std::integral_constant<int, 8>{});
    if (q + 4 <= q1) batch(std::integral_constant<int, 4>{});
    if (q + 2 <= q1) batch(std::integral_constant<int, 2>{});
    if (q < q1) batch(std::integral_constant<int, 1>{});
    *reinterpret_cast<uint32_t *>(S + (size_t)i * lds + f) = pk_bf16(s0, s1);
}

// dS: [N][128] FP32 (gradient of S).  Outputs dAB [N][256] bf16 and per-block
// partials of dWc [128][4] and dbc [128] in Pw[block][128 * 5].
__global__ void __launch_bounds__(256) k_edge_bwd(const bf16_t *__restrict__ AB, const float *__restrict__ pos,
                                                  const int *__restrict__ rp, const int *__restrict__ ci,
                                                  const float *__restrict__ Wc, const float *__restrict__ bc,
                                                  const float *__restrict__ dS, bf16_t *__restrict__ dAB,
                                                  float *__restrict__ Pw, int N, float inv_r) {
    __shared__ float red[4][128 * 5];
    const int wave = threadIdx.x >> 6;
    const int n = xcd_block(gridDim.x) * 4 + wave;
    const int f = 2 * (threadIdx.x & 63);
    const float4 w0 = *reinterpret_cast<const float4 *>(Wc + 4 * f), w1 = *reinterpret_cast<const float4 *>(Wc + 4 * f + 4);
    const float b0 = bc[f], b1 = bc[f + 1];
    float gw[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};   // dWc rows f, f+1 (4 each), dbc f, f+1
    if (n < N) {
        const float an0 = bf2f(AB[(size_t)n * 256 + f]), an1 = bf2f(AB[(size_t)n * 256 + f + 1]);
        const float bn0 = bf2f(AB[(size_t)n * 256 + 128 + f]), bn1 = bf2f(AB[(size_t)n * 256 + 128 + f + 1]);
        const float ds0 = dS[(size_t)n * 128 + f], ds1 = dS[(size_t)n * 128 + f + 1];
        float da0 = 0.f, da1 = 0.f, db0 = 0.f, db1 = 0.f;
        for (int q = rp[n]; q < rp[n + 1]; q++) {
            const int j = ci[q];
            float e[4];
            // n as destination of j -> n: mask with A_n + B_j + c(eps_nj)
            edge_eps(pos, n, j, inv_r, e);
            const uint32_t bj = *reinterpret_cast<const uint32_t *>(AB + (size_t)j * 256 + 128 + f);
            const float c0 = an0 + bf2f((bf16_t)(bj & 0xffff)) + b0 + w0.x * e[0] + w0.y * e[1] + w0.z * e[2] + w0.w * e[3];
            const float c1 = an1 + bf2f((bf16_t)(bj >> 16)) + b1 + w1.x * e[0] + w1.y * e[1] + w1.z * e[2] + w1.w * e[3];
            const float m0 = c0 > 0.f ? ds0 : 0.f, m1 = c1 > 0.f ? ds1 : 0.f;
            da0 += m0; da1 += m1;
            for (int t = 0; t < 4; t++) { gw[t] += m0 * e[t]; gw[4 + t] += m1 * e[t]; }
            gw[8] += m0; gw[9] += m1;
            // n as source of n -> j (the graph is symmetric: j's row holds n): mask with A_j + B_n + c(eps_jn)
            const uint32_t aj = *reinterpret_cast<const uint32_t *>(AB + (size_t)j * 256 + f);
            const float ej0 = -e[0], ej1 = -e[1], ej2 = -e[2];
            const float d0 = bf2f((bf16_t)(aj & 0xffff)) + bn0 + b0 + w0.x * ej0 + w0.y * ej1 + w0.z * ej2 + w0.w * e[3];
            const float d1 = bf2f((bf16_t)(aj >> 16)) + bn1 + b1 + w1.x * ej0 + w1.y * ej1 + w1.z * ej2 + w1.w * e[3];
            const float2 dsj = *reinterpret_cast<const float2 *>(dS + (size_t)j * 128 + f);
            db0 += d0 > 0.f ? dsj.x : 0.f;
            db1 += d1 > 0.f ? dsj.y : 0.f;
        }
        *reinterpret_cast<uint32_t *>(dAB + (size_t)n * 256 + f) = pk_bf16(da0, da1);
        *reinterpret_cast<uint32_t *>(dAB + (size_t)n * 256 + 128 + f) = pk_bf16(db0, db1);
    }
    for (int t = 0; t < 4; t++) { red[wave][4 * f + t] = gw[t]; red[wave][4 * (f + 1) + t] = gw[4 + t]; }
    red[wave][512 + f] = gw[8];
    red[wave][512 + f + 1] = gw[9];
    __syncthreads();
    for (int t = threadIdx.x; t < 640; t += blockDim.x)
        Pw[(size_t)blockIdx.x * 640 + t] = red[0][t] + red[1][t] + red[2][t] + red[3][t];
}

// ---------------------------------------------------------------------------------------------
// Radius graph on a uniform grid (cell width >= R), deterministic CSR order:
// rows list neighbours by ascending cell (x-major) then ascending node index.
// A node's cell key includes its scene (scene * cells + cell), so scenes of a
// batch never connect.
// ---------------------------------------------------------------------------------------------

__device__ __forceinline__ void cell_of(const float *pos, int i, const float *lo, float inv_w, const int *cc, int c[3]) {
    for (int k = 0; k < 3; k++) {
        int v = (int)floorf((pos[3 * i + k] - lo[k]) * inv_w);
        c[k] = v < 0 ? 0 : (v >= cc[k] ? cc[k] - 1 : v);
    }
}

__global__ void k_g_keys(const float *pos, const int *scene, int N, NnGrid gd, int *key, int *counts) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    int c[3];
    cell_of(pos, i, gd.lo, gd.inv_w, gd.cc, c);
    const int k = (scene ? scene[i] : 0) * gd.ncell + (c[0] * gd.cc[1] + c[1]) * gd.cc[2] + c[2];
    key[i] = k;
    atomicAdd(&counts[k], 1);
}

// Exclusive scan of n ints into out[0..n] (out[n] = total) in three passes over
// tiles of 1024: tile sums, one block scanning the sums, then each tile's scan
// from its carry.  The order of every addition is fixed (deterministic).
constexpr int kTile = 1024;

__global__ void __launch_bounds__(256) k_scan_tile_sums(const int *__restrict__ in, int n, int *__restrict__ sums) {
    __shared__ int ws[4];
    const int base = blockIdx.x * kTile + threadIdx.x * 4;
    int v = 0;
    for (int k = 0; k < 4; k++) v += base + k < n ? in[base + k] : 0;
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) sums[blockIdx.x] = ws[0] + ws[1] + ws[2] + ws[3];
}

__global__ void __launch_bounds__(256) k_scan_tiles(const int *__restrict__ in, int *__restrict__ out, int n,
                                                    const int *__restrict__ carry) {
    __shared__ int ws[4];
    const int base = blockIdx.x * kTile + threadIdx.x * 4;
    int v[4];
    for (int k = 0; k < 4; k++) v[k] = base + k < n ? in[base + k] : 0;
    const int loc = v[0] + v[1] + v[2] + v[3];
    int inc = loc;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(inc, o);
        if (lane >= o) inc += y;
    }
    if (lane == 63) ws[wid] = inc;
    __syncthreads();
    int run = carry[blockIdx.x] + inc - loc;
    for (int w = 0; w < wid; w++) run += ws[w];
    for (int k = 0; k < 4; k++) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == blockDim.x - 1) out[n] = run;   // the total
}

// bounding box of pos [N][3]: per-block min / max, then one block over the blocks
__global__ void __launch_bounds__(256) k_bbox_part(const float *__restrict__ pos, int N, float *__restrict__ part) {
    __shared__ float ws[4][6];
    float m[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x)
        for (int k = 0; k < 3; k++) {
            const float x = pos[3 * (size_t)i + k];
            m[k] = fminf(m[k], x);
            m[3 + k] = fmaxf(m[3 + k], x);
        }
    for (int o = 32; o > 0; o >>= 1)
        for (int k = 0; k < 3; k++) {
            m[k] = fminf(m[k], __shfl_xor(m[k], o));
            m[3 + k] = fmaxf(m[3 + k], __shfl_xor(m[3 + k], o));
        }
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < 6; k++) ws[threadIdx.x >> 6][k] = m[k];
    __syncthreads();
    if (threadIdx.x < 6) {
        float r = ws[0][threadIdx.x];
        for (int w = 1; w < 4; w++) r = threadIdx.x < 3 ? fminf(r, ws[w][threadIdx.x]) : fmaxf(r, ws[w][threadIdx.x]);
        part[6 * blockIdx.x + threadIdx.x] = r;
    }
}

__global__ void k_bbox_final(const float *__restrict__ part, int nb, float *__restrict__ out) {
    const int k = threadIdx.x;
    if (k >= 6) return;
    float r = part[k];
    for (int b = 1; b < nb; b++) r = k < 3 ? fminf(r, part[6 * b + k]) : fmaxf(r, part[6 * b + k]);
    out[k] = r;
}

// exclusive scan of n ints into out[0..n] by one 1024-thread block (tiles of 1024)
__global__ void __launch_bounds__(1024) k_scan1(const int *in, int *out, int n) {
    __shared__ int ws[16];
    __shared__ int carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int b = 0; b < n; b += 1024) {
        const int t = b + threadIdx.x;
        const int v = t < n ? in[t] : 0;
        int inc = v;
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(inc, o);
            if (lane >= o) inc += y;
        }
        if (lane == 63) ws[wid] = inc;
        __syncthreads();
        int wo = 0;
        for (int w = 0; w < wid; w++) wo += ws[w];
        const int c0 = carry;
        if (t < n) out[t] = c0 + wo + inc - v;
        __syncthreads();
        if (threadIdx.x == 1023) carry = c0 + wo + inc;
        __syncthreads();
    }
    if (threadIdx.x == 0) out[n] = carry;
}

__global__ void k_g_scatter(const int *key, const int *start, int *fill, int *order, int N) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int k = key[i];
    order[start[k] + atomicAdd(&fill[k], 1)] = i;
}

// 64-bit sum of n non-negative ints into *out (zeroed by the caller)
__global__ void __launch_bounds__(256) k_sum64(const int *__restrict__ v, int n, unsigned long long *out) {
    unsigned long long t = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) t += (unsigned)v[i];
    for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o);
    if ((threadIdx.x & 63) == 0) atomicAdd(out, t);
}

// ascending node index inside each cell; ps[slot] = (x, y, z, node id bits)
__global__ void k_g_rank(const int *key, const int *start, const int *order, const float *pos, float4 *ps, int N) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= N) return;
    const int i = order[p], k = key[i];
    int r = 0;
    for (int q = start[k]; q < start[k + 1]; q++) r += order[q] < i;
    ps[start[k] + r] = make_float4(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], __int_as_float(i));
}

// One wave per node: the 3 x 3 columns of neighbour cells are 9 contiguous runs
// of the cell-sorted positions (the 3 cells along z are adjacent); lanes test
// 64 candidates at a time and the hits of a step are stored contiguously
// (ballot + prefix), so a row is written in whole runs.  FILL writes ci (and,
// when eps is given, the edge features of edge_eps), else the degree.
// SLAB (FILL only): row i starts at i * row_cap instead of rp[i]; entries past
// row_cap are not written, the degree is, and *over counts the rows cut.
// Waves take the nodes in cell order (wave w: the node at ps[w]), so the waves
// in flight on an XCD read the same few columns of cells from its L2; each row
// is the same whichever wave writes it.
template <bool FILL, bool SLAB = false>
__global__ void __launch_bounds__(256) k_g_rows(const float *__restrict__ pos, const int *__restrict__ key,
                                                const int *__restrict__ start, const float4 *__restrict__ ps, int N,
                                                NnGrid gd, float r2, float inv_r, int *__restrict__ deg,
                                                const int *__restrict__ rp, int *__restrict__ ci,
                                                uint2 *__restrict__ eps, int row_cap = 0, int *over = nullptr,
                                                _Float16 *__restrict__ dist = nullptr) {
    const int w = __builtin_amdgcn_readfirstlane(xcd_block(gridDim.x) * 4 + (threadIdx.x >> 6));
    if (w >= N) return;
    const int lane = threadIdx.x & 63;
    const float4 pw = ps[w];
    const int i = __builtin_amdgcn_readfirstlane(__float_as_int(pw.w));
    const int k = key[i];
    const int sbase = k - k % gd.ncell;
    const int cell = k % gd.ncell;
    const int c[3] = {cell / (gd.cc[1] * gd.cc[2]), (cell / gd.cc[2]) % gd.cc[1], cell % gd.cc[2]};
    const float xi = pw.x, yi = pw.y, zi = pw.z;
    const int z0 = max(c[2] - 1, 0), z1 = min(c[2] + 1, gd.cc[2] - 1);
    const int row0 = SLAB ? 0 : (FILL ? rp[i] : 0);
    int out = row0;
    for (int x = max(c[0] - 1, 0); x <= min(c[0] + 1, gd.cc[0] - 1); x++)
        for (int y = max(c[1] - 1, 0); y <= min(c[1] + 1, gd.cc[1] - 1); y++) {
            const int col = sbase + (x * gd.cc[1] + y) * gd.cc[2];
            const int qe = start[col + z1 + 1];
            for (int qq = start[col + z0]; qq < qe; qq += 64) {
                const int q = qq + lane;
                bool hit = false;
                float dx = 0.f, dy = 0.f, dz = 0.f;
                int j = -1;
                if (q < qe) {
                    const float4 p = ps[q];
                    j = __float_as_int(p.w);
                    dx = p.x - xi; dy = p.y - yi; dz = p.z - zi;
                    hit = j != i && !(dx * dx + dy * dy + dz * dz > r2);
                }
                const uint64_t m = __ballot(hit);
                if (FILL && hit) {
                    const int o = out + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                                     __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
                    if (!SLAB || o < row_cap) {   // a cut row is reported, never written past
                        const size_t at = SLAB ? (size_t)i * row_cap + o : (size_t)o;
                        ci[at] = j;
                        const _Float16 dh = (_Float16)(sqrtf(dx * dx + dy * dy + dz * dz) * inv_r);   // one rounding for both
                        if (eps)
                            eps[at] = make_uint2(h2bits(h2_t{(_Float16)(dx * inv_r), (_Float16)(dy * inv_r)}),
                                                 h2bits(h2_t{(_Float16)(dz * inv_r), dh}));
                        if (dist) dist[at] = dh;
                    }
                }
                out += __popcll(m);
            }
        }
    if ((!FILL || SLAB) && lane == 0) {
        deg[i] = out - row0;
        if (SLAB && out > row_cap) atomicAdd(over, 1);
    }
}

// ---------------------------------------------------------------------------------------------
// Node features (bf16 [N][32]): 5 velocities (newest first) x 3 scaled by vscale,
// class one-hot (fluid, debris, wall), 6 wall-face distances clipped to
// [0, 1] in units of R (tank interior box [tlo, thi]), zero padding.
// ---------------------------------------------------------------------------------------------

__global__ void k_features(const float *vel_hist, int N, const int *cls, const float *pos, NnFeat ft, bf16_t *X) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    bf16_t *x = X + (size_t)i * 32;
    for (int h = 0; h < 5; h++)
        for (int k = 0; k < 3; k++) x[3 * h + k] = f2bf(vel_hist[((size_t)h * N + i) * 3 + k] * ft.vscale);
    for (int c = 0; c < 3; c++) x[15 + c] = f2bf(cls[i] == c ? 1.f : 0.f);
    for (int k = 0; k < 3; k++) {
        const float lo = (pos[3 * i + k] - ft.tlo[k]) * ft.inv_r, hi = (ft.thi[k] - pos[3 * i + k]) * ft.inv_r;
        x[18 + 2 * k] = f2bf(fminf(fmaxf(lo, 0.f), 1.f));
        x[19 + 2 * k] = f2bf(fminf(fmaxf(hi, 0.f), 1.f));
    }
    for (int k = 24; k < 32; k++) x[k] = 0;
}

// head: y[n][c] = sum_f H[n][f] Wd[c][f] + bd[c] (c < 3), FP32 out
__global__ void k_head(const bf16_t *H, const float *Wd, const float *bd, float *Yo, int N) {
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= N) return;
    const int f = 2 * (threadIdx.x & 63);
    const uint32_t h = *reinterpret_cast<const uint32_t *>(H + (size_t)i * 128 + f);
    const float h0 = bf2f((bf16_t)(h & 0xffff)), h1 = bf2f((bf16_t)(h >> 16));
    for (int c = 0; c < 3; c++) {
        float v = h0 * Wd[c * 128 + f] + h1 * Wd[c * 128 + f + 1];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
        if ((threadIdx.x & 63) == 0) Yo[3 * i + c] = v + bd[c];
    }
}

// loss = sum_i w_i |y_i - clip(t_i)|^2 / count with integer weights w_i = mask[i]
// (0: not in the loss); dY = 2 w (y - t) / count; per-block loss partials in Pl
__global__ void k_mse(const float *Yo, const float *T, const int *mask, int N, float inv_count, float clip, float *dY,
                      float *Pl) {
    __shared__ float red[256];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.f;
    if (i < N) {
        for (int c = 0; c < 3; c++) {
            const float t = fminf(fmaxf(T[3 * i + c], -clip), clip);
            const float w = (float)mask[i];   // per-node weight (0 = not in the loss)
            const float d = w > 0.f ? Yo[3 * i + c] - t : 0.f;
            l += w * d * d;
            dY[3 * i + c] = 2.f * w * d * inv_count;
        }
    }
    red[threadIdx.x] = l * inv_count;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) Pl[blockIdx.x] = red[0];
}

// head backward: dH[n][f] = (sum_c dY[n][c] Wd[c][f]) [H > 0] (bf16; H = relu(.)),
// partials of dWd [3][128] and dbd [3] per block in Pw[block][387]
__global__ void k_head_bwd(const bf16_t *H, const float *Wd, const float *dY, bf16_t *dH, float *Pw, int N) {
    __shared__ float red[4][387];
    const int wave = threadIdx.x >> 6;
    const int i = blockIdx.x * 4 + wave;
    const int f = 2 * (threadIdx.x & 63);
    float g[6] = {0, 0, 0, 0, 0, 0}, yb[3] = {0, 0, 0};
    if (i < N) {
        const uint32_t h = *reinterpret_cast<const uint32_t *>(H + (size_t)i * 128 + f);
        const float h0 = bf2f((bf16_t)(h & 0xffff)), h1 = bf2f((bf16_t)(h >> 16));
        const float y[3] = {dY[3 * i], dY[3 * i + 1], dY[3 * i + 2]};
        float d0 = y[0] * Wd[f] + y[1] * Wd[128 + f] + y[2] * Wd[256 + f];
        float d1 = y[0] * Wd[f + 1] + y[1] * Wd[128 + f + 1] + y[2] * Wd[256 + f + 1];
        if (!(h0 > 0.f)) d0 = 0.f;
        if (!(h1 > 0.f)) d1 = 0.f;
        *reinterpret_cast<uint32_t *>(dH + (size_t)i * 128 + f) = pk_bf16(d0, d1);
        for (int c = 0; c < 3; c++) { g[2 * c] = y[c] * h0; g[2 * c + 1] = y[c] * h1; yb[c] = y[c]; }
    }
    for (int c = 0; c < 3; c++) { red[wave][c * 128 + f] = g[2 * c]; red[wave][c * 128 + f + 1] = g[2 * c + 1]; }
    if (f == 0)
        for (int c = 0; c < 3; c++) red[wave][384 + c] = yb[c];
    __syncthreads();
    for (int t = threadIdx.x; t < 387; t += blockDim.x)
        Pw[(size_t)blockIdx.x * 387 + t] = red[0][t] + red[1][t] + red[2][t] + red[3][t];
}

// dX = dY * [Y > 0] (bf16, in place allowed)
__global__ void k_relu_bwd(const bf16_t *Y, int ldy, const bf16_t *dY, int lddy, bf16_t *dX, int lddx, int M, int n) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)M * n) return;
    const size_t r = t / n, c = t % n;
    dX[r * lddx + c] = bf2f(Y[r * ldy + c]) > 0.f ? dY[r * lddy + c] : (bf16_t)0;
}

// y = x (FP32 -> bf16) and y += x conversions with strides
__global__ void k_f2bf(const float *x, int ldx, bf16_t *y, int ldy, int M, int n) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)M * n) return;
    const size_t r = t / n, c = t % n;
    y[r * ldy + c] = f2bf(x[r * ldx + c]);
}
__global__ void k_bf2f_acc(const bf16_t *x, int ldx, float *y, int ldy, int M, int n, int accumulate) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)M * n) return;
    const size_t r = t / n, c = t % n;
    const float v = bf2f(x[r * ldx + c]);
    y[r * ldy + c] = accumulate ? y[r * ldy + c] + v : v;
}

// Adam on FP32 master weights; writes the bf16 copy [N][K] and its transpose [K][N] if given
__global__ void k_adam(float *w, const float *g, float *m, float *v, int n, float lr, float b1, float b2, float eps,
                       float c1, float c2, float wd, bf16_t *wb, bf16_t *wbt, int rows, int cols) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const float gi = g[t] + wd * w[t];
    const float mi = b1 * m[t] + (1.f - b1) * gi, vi = b2 * v[t] + (1.f - b2) * gi * gi;
    m[t] = mi;
    v[t] = vi;
    const float wi = w[t] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[t] = wi;
    if (wb) wb[t] = f2bf(wi);
    if (wbt) wbt[(size_t)(t % cols) * rows + t / cols] = f2bf(wi);
}

// ---------------------------------------------------------------------------------------------
// Rollout at stride h: debris bodies take the mean predicted acceleration of their
// members (translation only), fluid particles their own; walls stay.  Velocity
// history [5][N][3] shifts by one (newest first).
// ---------------------------------------------------------------------------------------------

__global__ void k_body_acc(const float *Y, const int *cls, const int *body, int N, float *bsum) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N || cls[i] != 1 || body[i] < 0) return;
    float *b = bsum + 4 * body[i];
    atomicAdd(b, Y[3 * i]);
    atomicAdd(b + 1, Y[3 * i + 1]);
    atomicAdd(b + 2, Y[3 * i + 2]);
    atomicAdd(b + 3, 1.f);
}

__global__ void k_integrate(float *pos, float *vh, const float *Y, float inv_ascale, const int *cls, const int *body,
                            const float *bsum, float h, int N) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N || cls[i] == 2) return;
    float a[3];
    if (cls[i] == 1 && body[i] >= 0) {
        const float *b = bsum + 4 * body[i];
        for (int k = 0; k < 3; k++) a[k] = b[k] / b[3] * inv_ascale;
    } else {
        for (int k = 0; k < 3; k++) a[k] = Y[3 * i + k] * inv_ascale;
    }
    const size_t st = (size_t)N * 3;
    for (int k = 0; k < 3; k++) {
        const float v = vh[3 * (size_t)i + k] + h * a[k];
        for (int t = 4; t > 0; t--) vh[t * st + 3 * (size_t)i + k] = vh[(t - 1) * st + 3 * (size_t)i + k];
        vh[3 * (size_t)i + k] = v;
        pos[3 * i + k] += h * v;
    }
}

// sum of `groups` partial rows: coalesced over outputs when there are many of
// them and few partials, one block per output otherwise
static void reduce_partials(const float *P, int groups, size_t len, size_t stride, float *out, int accumulate,
                            hipStream_t s) {
    if (len >= 4096 && groups <= 1024)
        hipLaunchKernelGGL(k_reduce_rows, dim3((len + 255) / 256), dim3(256), 0, s, P, groups, len, stride, out, accumulate);
    else
        hipLaunchKernelGGL(k_reduce_cols, dim3(len), dim3(256), 0, s, P, groups, stride, out, accumulate);
}

}  // namespace nn

using namespace nn;

static int launch_ok() { return hipGetLastError() == hipSuccess ? 0 : -5; }

// Loss partials of dmph_nn_mse_static: one float per 256-node block of k_mse,
// in a file-scope device array instead of caller or library allocations.
// Each device that runs the library holds its own copy until the process
// ends; every call on one device shares it, so calls that overlap on one
// device race.  k_mse writes every partial it launches before the reduction
// reads it, so nothing needs clearing between calls.
static __device__ __align__(256) float scratch[DMPH_NN_MSE_STATIC_MAX_N / 256];

// Device address of `scratch` per device ordinal, looked up once per device
// with that device current.  In host code `scratch` names a host-side stand-in,
// never the device memory: kernels and copies get these addresses.
static std::mutex g_scratch_mu;
static float *g_scratch_ptr[256];

static int static_scratch(float **out) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 256) return -5;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return -5;
        g_scratch_ptr[dev] = (float *)p;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

extern "C" {

int dmph_nn_gemm(const uint16_t *X, int ldx, const uint16_t *W, const float *bias, const uint16_t *R, int ldr,
                 uint16_t *Y, int ldy, int M, int K, int N, int relu, int kstep, void *stream) {
    if (M <= 0) return 0;
    // K / 8 must be a power of two (the LDS swizzle stays inside a row)
    if ((K != 32 && K != 64 && K != 128 && K != 256) || N % 128 || (kstep != 16 && kstep != 32)) return -1;
    if (ldy % 4 || (R && ldr % 4) || (relu & ~(DMPH_NN_RELU | DMPH_NN_F16_OUT))) return -1;   // 8-byte row words
    hipStream_t s = (hipStream_t)stream;
    constexpr int RB = DMPH_GEMM_RB, NW = DMPH_GEMM_NW, NCW = DMPH_GEMM_NC;
    constexpr int rows = RB * NW * 16;
    const int nc = (N == 256 && K <= 128) ? NCW : 1;   // W block in LDS <= 64 KB
    const size_t lds = (size_t)kCols * nc * K * sizeof(bf16_t);   // <= 64 KB
    // persistent with next-tile prefetch for one column block and K >= 128 (measured: 2.16 -> 1.79 ms
    // at K = 256, 2.07 -> 1.89 at K = 128 on 10M rows; slower for N = 256 and K = 32)
    const bool pers = nc == 1 && K >= 128;
    const dim3 grid(gemm_grid_x(M, rows, pers), N / (kCols * nc));
#define DMPH_GEMM(KS_, K_)                                                                                              \
    do {                                                                                                              \
        if (nc == 2)                                                                                                  \
            hipLaunchKernelGGL((k_gemm<KS_, K_, RB, NW, 2>), grid, dim3(NW * 64), lds, s, X, ldx, W, bias, R, ldr, Y,   \
                               ldy, M, relu);                                                                         \
        else if (pers)                                                                                                \
            hipLaunchKernelGGL((k_gemm<KS_, K_, RB, NW, 1, false, true>), grid, dim3(NW * 64), lds, s, X, ldx, W, bias, \
                               R, ldr, Y, ldy, M, relu);                                                              \
        else                                                                                                          \
            hipLaunchKernelGGL((k_gemm<KS_, K_, RB, NW, 1>), grid, dim3(NW * 64), lds, s, X, ldx, W, bias, R, ldr, Y,   \
                               ldy, M, relu);                                                                         \
    } while (0)
    if (kstep == 32) {
        switch (K) {
            case 32: DMPH_GEMM(32, 32); break;
            case 64: DMPH_GEMM(32, 64); break;
            case 128: DMPH_GEMM(32, 128); break;
            default: DMPH_GEMM(32, 256); break;
        }
    } else {
        switch (K) {
            case 32: DMPH_GEMM(16, 32); break;
            case 64: DMPH_GEMM(16, 64); break;
            case 128: DMPH_GEMM(16, 128); break;
            default: DMPH_GEMM(16, 256); break;
        }
    }
#undef DMPH_GEMM
    return launch_ok();
}

int dmph_nn_gemm_head(const uint16_t *X, int ldx, const uint16_t *W, const float *bias, const float *Wd,
                      const float *bd, float *Y3, int M, int kstep, void *stream) {
    if (M <= 0) return 0;
    if (kstep != 16 && kstep != 32) return -1;
    constexpr int RB = DMPH_GEMM_RB, NW = DMPH_GEMM_NW, rows = RB * NW * 16;
    const size_t lds = (size_t)kCols * 128 * sizeof(bf16_t);
    const dim3 grid(gemm_grid_x(M, rows, true), 1);
    hipStream_t s = (hipStream_t)stream;
    if (kstep == 32)
        hipLaunchKernelGGL((k_gemm<32, 128, RB, NW, 1, true, true>), grid, dim3(NW * 64), lds, s, X, ldx, W, bias,
                           nullptr, 0, nullptr, 0, M, DMPH_NN_RELU, Wd, bd, Y3);
    else
        hipLaunchKernelGGL((k_gemm<16, 128, RB, NW, 1, true, true>), grid, dim3(NW * 64), lds, s, X, ldx, W, bias,
                           nullptr, 0, nullptr, 0, M, DMPH_NN_RELU, Wd, bd, Y3);
    return launch_ok();
}

}  // extern "C"

// The encoder at inference, from the raw node data to the first round's input:
// the 32 node features of k_features (computed per lane: lane (r, g) makes
// features 8g .. 8g + 7 of row r, its MFMA fragment), h1 = relu(We1 x + be1),
// h0 = We2 h1 + be2 (stored bf16 into H0, the left half of [h | S]) and
// AB = Wab h0 (IEEE half), intermediates in registers (as k_node).
template <bool FOLD>   // as k_node (weights staged with rows in rho order)
__global__ void __launch_bounds__(kNodeWaves * 64) k_enc(const float *__restrict__ vel_hist, const int *__restrict__ cls,
                                                         const float *__restrict__ pos, NnFeat ft,
                                                         const bf16_t *__restrict__ We1, const float *__restrict__ be1,
                                                         const bf16_t *__restrict__ We2, const float *__restrict__ be2,
                                                         const bf16_t *__restrict__ Wab, bf16_t *__restrict__ H0,
                                                         int ldh0, bf16_t *__restrict__ ABo, int M, NnFold fold) {
    extern __shared__ bf16_t wl[];
    bf16_t *we1 = wl, *we2 = wl + 128 * 32, *wab = we2 + 128 * 128;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int lr = lane & 15, g = lane >> 4, lk = 8 * g;
    const int ntiles = (M + kNodeWaves * 16 - 1) / (kNodeWaves * 16);
    typedef uint4 frag_t;
    for (int t = threadIdx.x; t < 128 * 32 / 8; t += blockDim.x) {   // We1, rows in rho order
        const int r = t / 4, c = (t % 4) * 8;
        *reinterpret_cast<uint4 *>(we1 + wl_at(r, c, 32)) =
            *reinterpret_cast<const uint4 *>(We1 + (size_t)rho(r >> 4, r & 15) * 32 + c);
    }
    for (int t = threadIdx.x; t < 384 * 128 / 8; t += blockDim.x) {   // We2 and Wab, rows in rho order
        const int r = t / 16, c = (t % 16) * 8;
        const int rr = r < 128 ? r : r - 128, q = rr & 127;
        const bf16_t *src = (r < 128 ? We2 : Wab) + (size_t)((rr & ~127) + rho(q >> 4, q & 15)) * 128;
        *reinterpret_cast<uint4 *>((r < 128 ? we2 : wab) + wl_at(rr, c, 128)) = *reinterpret_cast<const uint4 *>(src + c);
    }
    __syncthreads();
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int r = (tile * kNodeWaves + wave) * 16 + lr;
        int lb = lr;
        asm volatile("" : "+v"(lb));
        // features 8g .. 8g + 7 of row r, as k_features computes them
        uint32_t xw[4] = {0, 0, 0, 0};
        if (r < M) {
            bf16_t x[8];
#pragma unroll
            for (int t = 0; t < 8; t++) {
                const int f = lk + t;
                float v = 0.f;
                if (f < 15) v = vel_hist[((size_t)(f / 3) * M + r) * 3 + f % 3] * ft.vscale;
                else if (f < 18) v = cls[r] == f - 15 ? 1.f : 0.f;
                else if (f < 24) {
                    const int k = (f - 18) >> 1;   // selects, not indexing: the parameter struct stays in registers
                    const float tlo = k == 0 ? ft.tlo[0] : (k == 1 ? ft.tlo[1] : ft.tlo[2]);
                    const float thi = k == 0 ? ft.thi[0] : (k == 1 ? ft.thi[1] : ft.thi[2]);
                    const float d = (f & 1) ? (thi - pos[3 * r + k]) * ft.inv_r : (pos[3 * r + k] - tlo) * ft.inv_r;
                    v = fminf(fmaxf(d, 0.f), 1.f);
                }
                x[t] = f < 24 ? f2bf(v) : (bf16_t)0;
            }
#pragma unroll
            for (int t = 0; t < 4; t++) xw[t] = (uint32_t)x[2 * t] | (uint32_t)x[2 * t + 1] << 16;
        }
        const frag_t xf = make_uint4(xw[0], xw[1], xw[2], xw[3]);
        f4_t acc[8];
#pragma unroll
        for (int c = 0; c < 8; c++) {
            const frag_t v = *reinterpret_cast<const frag_t *>(we1 + wl_at(c * 16 + lb, lk, 32));
            acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(b8_t, v), __builtin_bit_cast(b8_t, xf),
                                                             f4_t{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
        }
        uint32_t hf[8][2];
#pragma unroll
        for (int c = 0; c < 8; c++) {   // h1 = relu(We1 x + be1)
            const float4 b = *reinterpret_cast<const float4 *>(be1 + rho_base(c, g));
            hf[c][0] = pk_bf16_int(fmaxf(acc[c][0] + b.x, 0.f), fmaxf(acc[c][1] + b.y, 0.f));
            hf[c][1] = pk_bf16_int(fmaxf(acc[c][2] + b.z, 0.f), fmaxf(acc[c][3] + b.w, 0.f));
        }
#pragma unroll
        for (int c = 0; c < 8; c++) acc[c] = f4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kk = 0; kk < 4; kk++) {   // h0 = We2 h1 + be2
            const frag_t hk = make_uint4(hf[2 * kk][0], hf[2 * kk][1], hf[2 * kk + 1][0], hf[2 * kk + 1][1]);
#pragma unroll
            for (int c = 0; c < 8; c++) {
                const frag_t v = *reinterpret_cast<const frag_t *>(we2 + wl_at(c * 16 + lb, kk * 32 + lk, 128));
                acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(b8_t, v), __builtin_bit_cast(b8_t, hk),
                                                                 acc[c], 0, 0, 0);
            }
        }
#pragma unroll
        for (int c = 0; c < 8; c++) {
            const float4 b = *reinterpret_cast<const float4 *>(be2 + rho_base(c, g));
            hf[c][0] = pk_bf16(acc[c][0] + b.x, acc[c][1] + b.y);
            hf[c][1] = pk_bf16(acc[c][2] + b.z, acc[c][3] + b.w);
            if ((c & 1) && r < M)
                *reinterpret_cast<uint4 *>(H0 + (size_t)r * ldh0 + rho_base(c - 1, g)) =
                    make_uint4(hf[c - 1][0], hf[c - 1][1], hf[c][0], hf[c][1]);
        }
        float px = 0.f, py = 0.f, pz = 0.f;
        uint2 ab2[8];
#pragma unroll
        for (int half = 0; half < 2; half++) {   // AB = Wab h0
            if (FOLD && half == 1 && r < M) {
                px = (pos[3 * (size_t)r] - fold.origin[0]) * fold.inv_r;
                py = (pos[3 * (size_t)r + 1] - fold.origin[1]) * fold.inv_r;
                pz = (pos[3 * (size_t)r + 2] - fold.origin[2]) * fold.inv_r;
            }
#pragma unroll
            for (int c = 0; c < 8; c++) acc[c] = f4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kk = 0; kk < 4; kk++) {
                const frag_t hk = make_uint4(hf[2 * kk][0], hf[2 * kk][1], hf[2 * kk + 1][0], hf[2 * kk + 1][1]);
#pragma unroll
                for (int c = 0; c < 8; c++) {
                    const frag_t v =
                        *reinterpret_cast<const frag_t *>(wab + wl_at(half * 128 + c * 16 + lb, kk * 32 + lk, 128));
                    acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(b8_t, v),
                                                                     __builtin_bit_cast(b8_t, hk), acc[c], 0, 0, 0);
                }
            }
            if (r < M)
#pragma unroll
                for (int c = 0; c < 8; c++) {
                    if (FOLD && half == 1) {   // + Wc3 P of the row
                        const int col = rho_base(c, g) + (lb - lr);
                        const float4 wx = *reinterpret_cast<const float4 *>(fold.Wt + col);
                        const float4 wy = *reinterpret_cast<const float4 *>(fold.Wt + 128 + col);
                        const float4 wz = *reinterpret_cast<const float4 *>(fold.Wt + 256 + col);
                        acc[c][0] += wx.x * px + wy.x * py + wz.x * pz;
                        acc[c][1] += wx.y * px + wy.y * py + wz.y * pz;
                        acc[c][2] += wx.z * px + wy.z * py + wz.z * pz;
                        acc[c][3] += wx.w * px + wy.w * py + wz.w * pz;
                    }
                    ab2[c] = make_uint2(__builtin_bit_cast(uint32_t, h2_t{(_Float16)acc[c][0], (_Float16)acc[c][1]}),
                                        __builtin_bit_cast(uint32_t, h2_t{(_Float16)acc[c][2], (_Float16)acc[c][3]}));
                    if (c & 1)
                        *reinterpret_cast<uint4 *>(ABo + (size_t)r * 256 + half * 128 + rho_base(c - 1, g)) =
                            make_uint4(ab2[c - 1].x, ab2[c - 1].y, ab2[c].x, ab2[c].y);
                }
        }
    }
}

extern "C" {

int dmph_nn_encode(const float *vel_hist, int N, const int *cls, const float *pos, const NnFeat *ft, const uint16_t *We1,
                   const float *be1, const uint16_t *We2, const float *be2, const uint16_t *Wab, uint16_t *H0, int ldh0,
                   uint16_t *AB, const NnFold *fold, void *stream) {
    if (N <= 0) return 0;
    if (ldh0 % 4 || (fold && !fold->Wt)) return -1;
    constexpr size_t lds = (size_t)(128 * 32 + 128 * 128 + 256 * 128) * sizeof(bf16_t);   // 104 KB
    static bool lds_set[2] = {false, false};
    if (!lds_set[fold ? 1 : 0]) {
        const void *k = fold ? reinterpret_cast<const void *>(&k_enc<true>) : reinterpret_cast<const void *>(&k_enc<false>);
        if (hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess) return -5;
        lds_set[fold ? 1 : 0] = true;
    }
    const unsigned tiles = (unsigned)((N + kNodeWaves * 16 - 1) / (kNodeWaves * 16));
    const unsigned grid = tiles < 256u ? tiles : 256u;
    const NnFold fd = fold ? *fold : NnFold{};
    if (fold)
        hipLaunchKernelGGL(k_enc<true>, dim3(grid), dim3(kNodeWaves * 64), lds, (hipStream_t)stream, vel_hist, cls, pos,
                           *ft, We1, be1, We2, be2, Wab, H0, ldh0, AB, N, fd);
    else
        hipLaunchKernelGGL(k_enc<false>, dim3(grid), dim3(kNodeWaves * 64), lds, (hipStream_t)stream, vel_hist, cls, pos,
                           *ft, We1, be1, We2, be2, Wab, H0, ldh0, AB, N, fd);
    return launch_ok();
}

int dmph_nn_node(const uint16_t *HS, int ldhs, const uint16_t *W1, const float *b1, const uint16_t *W2, const float *b2,
                 const uint16_t *Wab, uint16_t *Hn, int ldhn, uint16_t *AB, int M, const NnFold *fold, void *stream) {
    if (M <= 0) return 0;
    if (ldhs % 8 || ldhn % 4 || (fold && (!fold->pos || !fold->Wt))) return -1;
    static bool lds_set[2] = {false, false};   // the launch needs all 160 KB of LDS
    if (!lds_set[fold ? 1 : 0]) {
        const void *k = fold ? reinterpret_cast<const void *>(&k_node<false, kNodeWaves, true>)
                             : reinterpret_cast<const void *>(&k_node<false, kNodeWaves, false>);
        if (hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kNodeLds) != hipSuccess) return -5;
        lds_set[fold ? 1 : 0] = true;
    }
    const unsigned tiles = (unsigned)((M + kNodeWaves * 16 - 1) / (kNodeWaves * 16));
    const unsigned grid = tiles < 256u ? tiles : 256u;   // one resident workgroup per CU
    const NnFold fd = fold ? *fold : NnFold{};
    if (fold)
        hipLaunchKernelGGL((k_node<false, kNodeWaves, true>), dim3(grid), dim3(kNodeWaves * 64), kNodeLds,
                           (hipStream_t)stream, HS, ldhs, W1, b1, W2, b2, Wab, Hn, ldhn, AB, M, nullptr, nullptr, nullptr,
                           nullptr, fd);
    else
        hipLaunchKernelGGL((k_node<false, kNodeWaves, false>), dim3(grid), dim3(kNodeWaves * 64), kNodeLds,
                           (hipStream_t)stream, HS, ldhs, W1, b1, W2, b2, Wab, Hn, ldhn, AB, M, nullptr, nullptr, nullptr,
                           nullptr, fd);
    return launch_ok();
}

int dmph_nn_node_head(const uint16_t *HS, int ldhs, const uint16_t *W1, const float *b1, const uint16_t *W2,
                      const float *b2, const uint16_t *Wd1, const float *bd1, const float *Wd2, const float *bd2,
                      float *Y3, int M, void *stream) {
    if (M <= 0) return 0;
    if (ldhs % 8) return -1;
    constexpr size_t lds = (size_t)(128 * 256 + 2 * 128 * 128) * sizeof(bf16_t);   // 128 KB
    static bool lds_set = false;
    if (!lds_set) {
        if (hipFuncSetAttribute(reinterpret_cast<const void *>(&k_node<true>), hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)lds) != hipSuccess)
            return -5;
        lds_set = true;
    }
    const unsigned tiles = (unsigned)((M + 8 * 16 - 1) / (8 * 16));
    const unsigned grid = tiles < 256u ? tiles : 256u;
    hipLaunchKernelGGL(k_node<true>, dim3(grid), dim3(8 * 64), lds, (hipStream_t)stream, HS, ldhs, W1, b1, W2,
                       b2, Wd1, nullptr, 0, nullptr, M, bd1, Wd2, bd2, Y3);
    return launch_ok();
}

size_t dmph_nn_dw_scratch(int M, int K, int N, int groups) {
    (void)M;
    return (size_t)groups * (size_t)N * (K + 1) * sizeof(float);
}

int dmph_nn_gemm_dw(const uint16_t *dY, int lddy, const uint16_t *X, int ldx, float *dW, float *db, int M, int K, int N,
                    int groups, float *scratch, int accumulate, void *stream) {
    if ((K != 32 && K != 128 && K != 256) || N % 128 || groups <= 0) return -1;
    hipStream_t s = (hipStream_t)stream;
    if (M <= 0) return 0;
    const int chunk = ((M + groups - 1) / groups + kTileM - 1) / kTileM * kTileM;
    const int g = (M + chunk - 1) / chunk;
    float *P = scratch, *Pb = scratch + (size_t)g * N * K;
    const dim3 grid(g, N / kCols);
    float *pb = db ? Pb : nullptr;
    switch (K) {
        case 32: hipLaunchKernelGGL(k_gemm_dw<2>, grid, dim3(256), 0, s, dY, lddy, X, ldx, P, pb, M, chunk); break;
        case 128: hipLaunchKernelGGL(k_gemm_dw<8>, grid, dim3(256), 0, s, dY, lddy, X, ldx, P, pb, M, chunk); break;
        case 256: hipLaunchKernelGGL(k_gemm_dw<16>, grid, dim3(256), 0, s, dY, lddy, X, ldx, P, pb, M, chunk); break;
        default: return -1;
    }
    const size_t len = (size_t)N * K;
    reduce_partials(P, g, len, len, dW, accumulate, s);
    if (db) reduce_partials(Pb, g, (size_t)N, (size_t)N, db, accumulate, s);
    return launch_ok();
}

int dmph_nn_edge_fwd(const uint16_t *AB, const float *pos, const int *rp, const int *ci, const float *Wc,
                     const float *bc, uint16_t *S, int lds, int N, float radius, void *stream) {
    if (N <= 0) return 0;
    if (lds < 128 || lds % 2) return -1;
    hipLaunchKernelGGL(k_edge_fwd, dim3((N + 3) / 4), dim3(256), 0, (hipStream_t)stream, AB, pos, rp, ci, Wc, bc, S, lds,
                       N, 1.f / radius);
    return launch_ok();
}

int dmph_nn_edge_eps(const float *pos, const int *rp, const int *ci, uint32_t *eps, int N, float radius, void *stream) {
    if (N <= 0) return 0;
    hipLaunchKernelGGL(k_edge_eps, dim3((N + 3) / 4), dim3(256), 0, (hipStream_t)stream, pos, rp, ci,
                       reinterpret_cast<uint2 *>(eps), N, 1.f / radius);
    return launch_ok();
}

int dmph_nn_edge_fwd_f16(const uint16_t *AB, const int *rp, const int *ci, const uint32_t *eps, const float *Wc,
                         const float *bc, uint16_t *S, int lds, int N, const int *deg, int row_cap,
                         const int *order, const NnFold *fold, void *stream) {
    if (N <= 0) return 0;
    if (lds < 128 || lds % 2 || (deg && (row_cap <= 0 || (long long)N * row_cap > 0x7fffffffLL))) return -1;
    if (fold && (!fold->pos || !deg || row_cap % 2)) return -1;   // the fold reads rows' distance pairs
    const bool buf = (size_t)N * 512 <= 0xffffffffull - 1024;   // row offsets fit the 32-bit scalar offset
    const NnFold fd = fold ? *fold : NnFold{};
    auto go = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3((N + 3) / 4), dim3(256), 0, (hipStream_t)stream,
                           reinterpret_cast<const h2_t *>(AB), rp, ci, reinterpret_cast<const uint2 *>(eps), Wc, bc, S,
                           lds, N, deg, row_cap, order, fd);
    };
    if (buf)
        fold ? go(k_edge_fwd_h<true, true>) : go(k_edge_fwd_h<true, false>);
    else
        fold ? go(k_edge_fwd_h<false, true>) : go(k_edge_fwd_h<false, false>);
    return launch_ok();
}

size_t dmph_nn_edge_scratch(int N) { return (size_t)((N + 3) / 4) * 640 * sizeof(float); }

int dmph_nn_edge_bwd(const uint16_t *AB, const float *pos, const int *rp, const int *ci, const float *Wc,
                     const float *bc, const float *dS, uint16_t *dAB, float *dWc, float *dbc, int N, float radius,
                     float *scratch, int accumulate, void *stream) {
    if (N <= 0) return 0;
    hipStream_t s = (hipStream_t)stream;
    const int g = (N + 3) / 4;
    hipLaunchKernelGGL(k_edge_bwd, dim3(g), dim3(256), 0, s, AB, pos, rp, ci, Wc, bc, dS, dAB, scratch, N, 1.f / radius);
    // partial layout per block (640 floats): [128][4] weights, then [128] bias
    reduce_partials(scratch, g, 512, 640, dWc, accumulate, s);
    reduce_partials(scratch + 512, g, 128, 640, dbc, accumulate, s);
    return launch_ok();
}

static void scan_excl(const int *in, int *out, int n, int *scratch, hipStream_t s) {
    const int nt = (n + kTile - 1) / kTile;
    if (nt <= 1) {
        hipLaunchKernelGGL(k_scan1, dim3(1), dim3(1024), 0, s, in, out, n);
        return;
    }
    hipLaunchKernelGGL(k_scan_tile_sums, dim3(nt), dim3(256), 0, s, in, n, scratch);
    hipLaunchKernelGGL(k_scan1, dim3(1), dim3(1024), 0, s, scratch, scratch + nt + 1, nt);   // carries, total at [nt]
    hipLaunchKernelGGL(k_scan_tiles, dim3(nt), dim3(256), 0, s, in, out, n, scratch + nt + 1);
}

// scratch layout: [0, 2) ints the 64-bit edge total (dmph_nn_graph) or [2]
// the cut-row count (dmph_nn_graph_rows), then the scan carries from [4]
size_t dmph_nn_graph_scratch(int N, int ncells_total) {
    const int n = N > ncells_total ? N : ncells_total;
    return (size_t)(2 * ((n + kTile - 1) / kTile) + 6) * sizeof(int);
}

size_t dmph_nn_bbox_scratch(int N) { (void)N; return 256 * 6 * sizeof(float); }

int dmph_nn_bbox(const float *pos, int N, float *out, float *scratch, void *stream) {
    if (N <= 0) return -1;
    const int nb = (N + 255) / 256 < 256 ? (N + 255) / 256 : 256;
    hipLaunchKernelGGL(k_bbox_part, dim3(nb), dim3(256), 0, (hipStream_t)stream, pos, N, scratch);
    hipLaunchKernelGGL(k_bbox_final, dim3(1), dim3(64), 0, (hipStream_t)stream, scratch, nb, out);
    return launch_ok();
}

int dmph_nn_graph(const float *pos, const int *scene, int N, const NnGrid *grid, float radius, int *key, int *counts,
                  int *start, int *fill, int *order, int *sorted, int *deg, int *rp, int *ci, uint32_t *eps, int cap,
                  int *nnz_out,
                  int *scratch, void *stream) {
    hipStream_t s = (hipStream_t)stream;
    const int nk = grid->ncell * (grid->nscene > 0 ? grid->nscene : 1);
    if (hipMemsetAsync(counts, 0, (size_t)(nk + 1) * sizeof(int), s) != hipSuccess) return -5;
    if (hipMemsetAsync(fill, 0, (size_t)(nk + 1) * sizeof(int), s) != hipSuccess) return -5;
    const int b = (N + 255) / 256;
    hipLaunchKernelGGL(k_g_keys, dim3(b), dim3(256), 0, s, pos, scene, N, *grid, key, counts);
    unsigned long long *total = reinterpret_cast<unsigned long long *>(scratch);
    int *carries = scratch + 4;
    if (hipMemsetAsync(total, 0, sizeof(unsigned long long), s) != hipSuccess) return -5;
    scan_excl(counts, start, nk, carries, s);
    hipLaunchKernelGGL(k_g_scatter, dim3(b), dim3(256), 0, s, key, start, fill, order, N);
    float4 *ps = reinterpret_cast<float4 *>(sorted);
    hipLaunchKernelGGL(k_g_rank, dim3(b), dim3(256), 0, s, key, start, order, pos, ps, N);
    const float r2 = radius * radius;
    const int bw = (N + 3) / 4;   // a wave per node
    hipLaunchKernelGGL(k_g_rows<false>, dim3(bw), dim3(256), 0, s, pos, key, start, ps, N, *grid, r2, 1.f / radius,
                       deg, nullptr, nullptr, nullptr);
    scan_excl(deg, rp, N, carries, s);
    hipLaunchKernelGGL(k_sum64, dim3(b < 1024 ? b : 1024), dim3(256), 0, s, deg, N, total);
    int nnz = 0;
    unsigned long long all = 0;
    if (hipMemcpyAsync(&nnz, rp + N, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess) return -5;
    if (hipMemcpyAsync(&all, total, sizeof(all), hipMemcpyDeviceToHost, s) != hipSuccess) return -5;
    if (hipStreamSynchronize(s) != hipSuccess) return -5;
    if (all > 0x7fffffffull) {   // int32 row offsets would wrap: refuse before anything is written
        *nnz_out = -1;
        return -3;
    }
    *nnz_out = nnz;
    if (nnz > cap) return -2;   // caller grows ci and calls again
    hipLaunchKernelGGL(k_g_rows<true>, dim3(bw), dim3(256), 0, s, pos, key, start, ps, N, *grid, r2, 1.f / radius,
                       deg, rp, ci, reinterpret_cast<uint2 *>(eps));
    return launch_ok();
}

int dmph_nn_graph_rows(const float *pos, const int *scene, int N, const NnGrid *grid, float radius, int *key,
                       int *counts, int *start, int *fill, int *order, int *sorted, int *deg, int *ci, uint32_t *eps,
                       uint16_t *dist, int row_cap, int *over_out, int *scratch, void *stream) {
    if (row_cap <= 0 || (dist && row_cap % 2)) return -1;
    hipStream_t s = (hipStream_t)stream;
    const int nk = grid->ncell * (grid->nscene > 0 ? grid->nscene : 1);
    int *over = scratch + 2;   // graph scratch: [2, 3) the cut-row count here
    if (hipMemsetAsync(counts, 0, (size_t)(nk + 1) * sizeof(int), s) != hipSuccess) return -5;
    if (hipMemsetAsync(fill, 0, (size_t)(nk + 1) * sizeof(int), s) != hipSuccess) return -5;
    if (hipMemsetAsync(over, 0, sizeof(int), s) != hipSuccess) return -5;
    const int b = (N + 255) / 256;
    hipLaunchKernelGGL(k_g_keys, dim3(b), dim3(256), 0, s, pos, scene, N, *grid, key, counts);
    scan_excl(counts, start, nk, scratch + 4, s);
    hipLaunchKernelGGL(k_g_scatter, dim3(b), dim3(256), 0, s, key, start, fill, order, N);
    float4 *ps = reinterpret_cast<float4 *>(sorted);
    hipLaunchKernelGGL(k_g_rank, dim3(b), dim3(256), 0, s, key, start, order, pos, ps, N);
    hipLaunchKernelGGL((k_g_rows<true, true>), dim3((N + 3) / 4), dim3(256), 0, s, pos, key, start, ps, N, *grid,
                       radius * radius, 1.f / radius, deg, nullptr, ci, reinterpret_cast<uint2 *>(eps), row_cap, over,
                       reinterpret_cast<_Float16 *>(dist));
    int cut = 0;
    if (hipMemcpyAsync(&cut, over, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess) return -5;
    if (hipStreamSynchronize(s) != hipSuccess) return -5;
    *over_out = cut;
    return cut ? -2 : launch_ok();
}

int dmph_nn_features(const float *vel_hist, int N, const int *cls, const float *pos, const NnFeat *ft, uint16_t *X,
                     void *stream) {
    if (N <= 0) return 0;
    hipLaunchKernelGGL(k_features, dim3((N + 255) / 256), dim3(256), 0, (hipStream_t)stream, vel_hist, N, cls, pos, *ft, X);
    return launch_ok();
}

int dmph_nn_head(const uint16_t *H, const float *Wd, const float *bd, float *Y, int N, void *stream) {
    if (N <= 0) return 0;
    hipLaunchKernelGGL(k_head, dim3((N + 3) / 4), dim3(256), 0, (hipStream_t)stream, H, Wd, bd, Y, N);
    return launch_ok();
}

size_t dmph_nn_head_scratch(int N) { return (size_t)((N + 3) / 4) * 387 * sizeof(float); }

int dmph_nn_head_bwd(const uint16_t *H, const float *Wd, const float *dY, uint16_t *dH, float *dWd, float *dbd, int N,
                     float *scratch, void *stream) {
    if (N <= 0) return 0;
    hipStream_t s = (hipStream_t)stream;
    const int g = (N + 3) / 4;
    hipLaunchKernelGGL(k_head_bwd, dim3(g), dim3(256), 0, s, H, Wd, dY, dH, scratch, N);
    reduce_partials(scratch, g, 384, 387, dWd, 0, s);
    reduce_partials(scratch + 384, g, 3, 387, dbd, 0, s);
    return launch_ok();
}

size_t dmph_nn_mse_scratch(int N) { return (size_t)((N + 255) / 256) * sizeof(float); }

int dmph_nn_mse(const float *Y, const float *T, const int *mask, int N, int count, float clip, float *dY, float *loss,
                float *scratch, void *stream) {
    if (N <= 0 || count <= 0 || !(clip > 0.f)) return -1;
    hipStream_t s = (hipStream_t)stream;
    const int g = (N + 255) / 256;
    hipLaunchKernelGGL(k_mse, dim3(g), dim3(256), 0, s, Y, T, mask, N, 1.f / (float)count, clip, dY, scratch);
    reduce_partials(scratch, g, 1, 1, loss, 0, s);
    return launch_ok();
}

int dmph_nn_mse_static(const float *Y, const float *T, const int *mask, int N, int count, float clip, float *dY,
                       float *loss, void *stream) {
    if (N <= 0 || count <= 0 || !(clip > 0.f)) return -1;
    if (N > DMPH_NN_MSE_STATIC_MAX_N) return -4;   // the partials would not fit the static array
    float *part = nullptr;
    if (static_scratch(&part) != 0) return -5;
    hipStream_t s = (hipStream_t)stream;
    const int g = (N + 255) / 256;
    hipLaunchKernelGGL(k_mse, dim3(g), dim3(256), 0, s, Y, T, mask, N, 1.f / (float)count, clip, dY, part);
    reduce_partials(part, g, 1, 1, loss, 0, s);
    return launch_ok();
}

int dmph_nn_relu_bwd(const uint16_t *Y, int ldy, const uint16_t *dY, int lddy, uint16_t *dX, int lddx, int M, int n,
                     void *stream) {
    const size_t t = (size_t)M * n;
    if (!t) return 0;
    hipLaunchKernelGGL(k_relu_bwd, dim3((t + 255) / 256), dim3(256), 0, (hipStream_t)stream, Y, ldy, dY, lddy, dX, lddx, M, n);
    return launch_ok();
}

int dmph_nn_f2bf(const float *x, int ldx, uint16_t *y, int ldy, int M, int n, void *stream) {
    const size_t t = (size_t)M * n;
    if (!t) return 0;
    hipLaunchKernelGGL(k_f2bf, dim3((t + 255) / 256), dim3(256), 0, (hipStream_t)stream, x, ldx, y, ldy, M, n);
    return launch_ok();
}

int dmph_nn_bf2f(const uint16_t *x, int ldx, float *y, int ldy, int M, int n, int accumulate, void *stream) {
    const size_t t = (size_t)M * n;
    if (!t) return 0;
    hipLaunchKernelGGL(k_bf2f_acc, dim3((t + 255) / 256), dim3(256), 0, (hipStream_t)stream, x, ldx, y, ldy, M, n, accumulate);
    return launch_ok();
}

int dmph_nn_rollout_step(float *pos, float *vel_hist, const float *Y, float ascale, const int *cls, const int *body,
                         float *body_sums, int nbodies, float h, int N, void *stream) {
    if (N <= 0) return 0;
    hipStream_t s = (hipStream_t)stream;
    if (nbodies > 0) {
        if (hipMemsetAsync(body_sums, 0, (size_t)nbodies * 4 * sizeof(float), s) != hipSuccess) return -5;
        hipLaunchKernelGGL(k_body_acc, dim3((N + 255) / 256), dim3(256), 0, s, Y, cls, body, N, body_sums);
    }
    hipLaunchKernelGGL(k_integrate, dim3((N + 255) / 256), dim3(256), 0, s, pos, vel_hist, Y, 1.f / ascale, cls, body,
                       body_sums, h, N);
    return launch_ok();
}

int dmph_nn_adam(float *w, const float *g, float *m, float *v, int n, float lr, float b1, float b2, float eps, int step,
                 float wd, uint16_t *wb, uint16_t *wbt, int rows, int cols, void *stream) {
    if (n <= 0) return 0;
    const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
    hipLaunchKernelGGL(k_adam, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, w, g, m, v, n, lr, b1, b2, eps,
                       c1, c2, wd, wb, wbt, rows, cols);
    return launch_ok();
}

}  // extern "C"
