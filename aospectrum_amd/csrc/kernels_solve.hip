// Tile triangular solves, RHS packing and FP64 refinement residuals.
//
// A right-hand-side block of W columns (W a multiple of 64) is stored per
// matrix as Nt row tiles x W/64 planar 64x64 chunks, rows in factor
// (permuted) order.  Solves run one tile step per launch, with the FP32 MFMA
// tile core of the factorisation:
//   diag   : Y_k <- D_k Y_k with D_k = inv(L_kk), inv(U_kk) or their
//            conjugate transposes (for A^H = U^H L^H, used by conjugate shifts);
//   update : Y_i -= T_ik Y_k for the tiles i coupled to k.
// Every launch writes only tiles that no other block of the same launch reads.
#include "device_common.h"

namespace aosx {

// RHS layout of the solves: per matrix, Nt row tiles x (W / 64) column chunks,
// each chunk a 64 x 64 complex block stored like a factor tile (re plane, im
// plane, row-major), so tile products reuse the MFMA core of the factorisation.
__device__ __forceinline__ int64_t rhs_block(int64_t mat_stride, int m, int tile, int nchunk, int c) {
    return (int64_t)m * mat_stride + ((int64_t)tile * nchunk + c) * TILE_FLOATS;
}

// Right-hand-side sets of a solve launch.  With pairs = F > 0 the launch
// covers 2F sets: set m < F is solved with factorisation m, set m >= F with
// the conjugate transpose of factorisation m - F (the conjugate shift of a
// contour pair, A^H = U^H L^H), so both shifts of a pair take one sweep.  With
// pairs = 0 set m uses factorisation m and the launch's ctrans.  `which`
// arguments name the diagonal inverse of the plain orientation (0: inv(L_kk),
// 1: inv(U_kk)); a conjugate set takes the other one.
struct SolveSet {
    int fac, ct;
};
__device__ __forceinline__ SolveSet solve_set(int m, int pairs, int ctrans) {
    return pairs > 0 ? SolveSet{m % pairs, m >= pairs ? 1 : 0} : SolveSet{m, ctrans};
}

// Stage the conjugate transpose of a stored tile as the A operand.
__device__ __forceinline__ void stage_a_ctrans(const float *__restrict__ g, float *Are, float *Aim, int tid) {
    for (int e = tid; e < TILE2; e += 256) {
        const int r = e / TILE, c = e % TILE;
        Are[c * LDA + r] = g[e];
        Aim[c * LDA + r] = -g[PLANE + e];
    }
}

// Y_k <- D_k Y_k for all chunks; D_k = inv(L_kk) (which 0) or inv(U_kk)
// (which 1), conjugate-transposed when ctrans.  FP32 MFMA.
__global__ void __launch_bounds__(256) k_solve_diag(float *__restrict__ y, int64_t y_mat_stride, int nchunk,
                                                   const float *__restrict__ dinv, int nt, int k, int which,
                                                   int ctrans, int pairs) {
    __shared__ float Are[TILE * LDA], Aim[TILE * LDA];
    __shared__ __attribute__((aligned(16))) float Bre[TILE2], Bim[TILE2];
    const int m = blockIdx.x, tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const SolveSet st = solve_set(m, pairs, ctrans);
    ctrans = st.ct;
    const float *d = dinv + ((int64_t)st.fac * nt + k) * (2 * TILE_FLOATS) + (which ^ st.ct) * TILE_FLOATS;
    if (ctrans) stage_a_ctrans(d, Are, Aim, tid);
    else stage_a(d, Are, Aim, tid);
    for (int c = 0; c < nchunk; ++c) {
        float *yk = y + rhs_block(y_mat_stride, m, k, nchunk, c);
        __syncthreads();
        stage_b(yk, Bre, Bim, tid);
        __syncthreads();
        floatx16 cre = {}, cim = {};
        mma_fp32<false>(Are, Aim, Bre, Bim, cre, cim, wave, lane);
        store_acc(yk, cre, cim, wave, lane);
    }
}

// The same with both operands split into BF16 hi + lo parts (three MFMA
// products, FP32 accumulation; refinement on FP64 residuals removes the
// difference): the BF16 precisions.
// Split RHS slots: the finished Y_k of step k is also kept, split and
// transposed (rows = RHS columns, k = rows of Y_k) in the six-plane format of
// the split panel tiles (re, im, re + im, each hi and lo, in MFMA operand
// blocks), in slot k % YSLOTS of the matrix's slot area: the B operand of the
// trailing update products, read straight from memory; a step group spans at
// most 32 steps, so its slots are distinct.
constexpr int YSLOTS = 64;

__device__ __forceinline__ uint16_t *yslot(uint16_t *ys, int m, int k, int nchunk, int c) {
    return ys + (((int64_t)m * YSLOTS + k % YSLOTS) * nchunk + c) * PANEL_TILE;
}
__device__ __forceinline__ const uint16_t *yslot(const uint16_t *ys, int m, int k, int nchunk, int c) {
    return ys + (((int64_t)m * YSLOTS + k % YSLOTS) * nchunk + c) * PANEL_TILE;
}

// Accumulator block (a, b) of a 64 x 64 tile (rows 32a.., columns 32b..) in
// the 32 x 32 MFMA layout: register r of lane l holds row 32a + 8(r/4) +
// 4(l/32) + r%4, column 32b + l%32.
__device__ __forceinline__ void load_blk(const float *__restrict__ t, floatx16 &cre, floatx16 &cim, int a, int b,
                                         int lane) {
    const int r0 = a * 32 + 4 * (lane >> 5), c = b * 32 + (lane & 31);
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int row = r0 + 8 * (r >> 2) + (r & 3);
        cre[r] = t[row * TILE + c];
        cim[r] = t[PLANE + row * TILE + c];
    }
}
__device__ __forceinline__ void store_blk(float *__restrict__ t, const floatx16 &cre, const floatx16 &cim, int a,
                                          int b, int lane) {
    const int r0 = a * 32 + 4 * (lane >> 5), c = b * 32 + (lane & 31);
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int row = r0 + 8 * (r >> 2) + (r & 3);
        t[row * TILE + c] = cre[r];
        t[PLANE + row * TILE + c] = cim[r];
    }
}

// Writes the wave's 32 x 32 accumulator block (wave / 2, wave % 2) of Y_k into
// its slot: transposed (row = column of Y_k), six planes, panel block order.
// Lane l holds column c = l % 32 of the block and 4 consecutive rows (k
// values) 8q + 4(l/32) .. + 3: 8 bytes per plane.
__device__ __forceinline__ void store_acc_slot(uint16_t *__restrict__ dst, const floatx16 &cre, const floatx16 &cim,
                                               int wave, int lane) {
    const int a = wave & 1, col = lane & 31;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int k0 = (wave >> 1) * 32 + 4 * (lane >> 5) + 8 * q;  // first of the 4 rows of Y_k
        uint32_t w0[6], w1[6];
        split6(cre[4 * q], cim[4 * q], cre[4 * q + 1], cim[4 * q + 1], w0);
        split6(cre[4 * q + 2], cim[4 * q + 2], cre[4 * q + 3], cim[4 * q + 3], w1);
        const int lane_t = ((k0 >> 3) & 1) * 32 + col, off = lane_t * 8 + (k0 & 7);
#pragma unroll
        for (int pl = 0; pl < 6; ++pl)
            *reinterpret_cast<uint2 *>(dst + panel_block(k0 >> 4, pl >> 1, pl & 1, a) * SPLIT_BLOCK + off) =
                make_uint2(w0[pl], w1[pl]);
    }
}

__global__ void __launch_bounds__(256, 2) k_solve_diag_split(float *__restrict__ y, int64_t y_mat_stride, int nchunk,
                                                            const float *__restrict__ dinv, int nt, int k, int which,
                                                            int ctrans, int pairs, uint16_t *__restrict__ ysplit) {
    __shared__ __attribute__((aligned(16))) uint16_t As[4 * SPL], Bs[4 * SPL];
    const int m = blockIdx.x, tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const SolveSet st = solve_set(m, pairs, ctrans);
    ctrans = st.ct;
    const float *d = dinv + ((int64_t)st.fac * nt + k) * (2 * TILE_FLOATS) + (which ^ st.ct) * TILE_FLOATS;
    stage_split(d, As, ctrans != 0, false, ctrans != 0, tid);
    for (int c = 0; c < nchunk; ++c) {
        float *yk = y + rhs_block(y_mat_stride, m, k, nchunk, c);
        __syncthreads();
        stage_split(yk, Bs, true, false, false, tid);
        __syncthreads();
        floatx16 cre = {}, cim = {};
        mma_split<true>(As, Bs, cre, cim, wave, lane);
        store_acc(yk, cre, cim, wave, lane);
        store_acc_slot(yslot(ysplit, m, k, nchunk, c), cre, cim, wave, lane);
    }
}

// Y_i -= T Y_k for i = list[blockIdx.x], all chunks; T is tile (i,k) or,
// with ctrans, the conjugate transpose of tile (k,i).  FP32 MFMA.
__global__ void __launch_bounds__(256) k_solve_update(float *__restrict__ y, int64_t y_mat_stride, int nchunk,
                                                     const float *__restrict__ tiles, int64_t mat_stride,
                                                     const int32_t *__restrict__ tile_id, int nt, int k,
                                                     const int32_t *__restrict__ list, int ctrans, int pairs) {
    __shared__ float Are[TILE * LDA], Aim[TILE * LDA];
    __shared__ __attribute__((aligned(16))) float Bre[TILE2], Bim[TILE2];
    const int m = blockIdx.y, tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const SolveSet st = solve_set(m, pairs, ctrans);
    ctrans = st.ct;
    const int i = list[blockIdx.x];
    const int32_t id = ctrans ? tile_id[(int64_t)k * nt + i] : tile_id[(int64_t)i * nt + k];
    const float *t = tiles + (int64_t)st.fac * mat_stride + (int64_t)id * TILE_FLOATS;
    if (ctrans) stage_a_ctrans(t, Are, Aim, tid);
    else stage_a(t, Are, Aim, tid);
    for (int c = 0; c < nchunk; ++c) {
        const float *yk = y + rhs_block(y_mat_stride, m, k, nchunk, c);
        float *yi = y + rhs_block(y_mat_stride, m, i, nchunk, c);
        __syncthreads();
        stage_b(yk, Bre, Bim, tid);
        floatx16 cre, cim;
        load_acc(yi, cre, cim, wave, lane);
        __syncthreads();
        mma_fp32<true>(Are, Aim, Bre, Bim, cre, cim, wave, lane);
        store_acc(yi, cre, cim, wave, lane);
    }
}

// The same with split operands (BF16 precisions): -T (or -T^H) and Y_k staged
// as split tiles.
__global__ void __launch_bounds__(256, 2) k_solve_update_split(float *__restrict__ y, int64_t y_mat_stride,
                                                              int nchunk, const float *__restrict__ tiles,
                                                              int64_t mat_stride, const int32_t *__restrict__ tile_id,
                                                              int nt, int k, const int32_t *__restrict__ list,
                                                              int ctrans, int pairs,
                                                              const uint16_t *__restrict__ ysplit) {
    __shared__ __attribute__((aligned(16))) uint16_t As[4 * SPL], Bs[4 * SPL];
    const int m = blockIdx.y, tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const SolveSet st = solve_set(m, pairs, ctrans);
    ctrans = st.ct;
    const int i = list[blockIdx.x];
    const int32_t id = ctrans ? tile_id[(int64_t)k * nt + i] : tile_id[(int64_t)i * nt + k];
    const float *t = tiles + (int64_t)st.fac * mat_stride + (int64_t)id * TILE_FLOATS;
    stage_split(t, As, ctrans != 0, true, ctrans != 0, tid);
    for (int c = 0; c < nchunk; ++c) {
        float *yi = y + rhs_block(y_mat_stride, m, i, nchunk, c);
        const u32x32 yv = fetch_panel4(yslot(ysplit, m, k, nchunk, c), tid);
        __syncthreads();
        put_panel4(Bs, yv, tid);
        floatx16 cre, cim;
        load_acc(yi, cre, cim, wave, lane);
        __syncthreads();
        mma_split<true>(As, Bs, cre, cim, wave, lane);
        store_acc(yi, cre, cim, wave, lane);
    }
}

// Trailing update of a step group [g0, g1) of a sweep: Y_i -= sum_k T_ik Y_k
// over the group's steps k with tile T_ik present, for the rows i of the
// group's list (rows the group's steps couple to, outside the group).  T_ik is
// tile (i,k) or, with ctrans, the conjugate transpose of tile (k,i).  A block
// keeps NC chunks of Y_i in the MFMA accumulators across the group's steps,
// so Y_i is read and written once per group instead of once per step.  Grid:
// 8 * ceil(nrows * nmat * ngrp / 8) blocks in one dimension, mapped to items
// (matrix, chunk group, row) so that each XCD takes a contiguous run (rows of
// one matrix share the group's Y_k blocks in its L2).  FP32 MFMA form.
template <int NC>
__global__ void __launch_bounds__(256, 2) k_solve_trail(float *__restrict__ y, int64_t y_mat_stride, int nchunk,
                                                       const float *__restrict__ tiles, int64_t mat_stride,
                                                       const int32_t *__restrict__ tile_id, int nt, int g0, int g1,
                                                       const int32_t *__restrict__ rows, int nrows, int nmat,
                                                       int ctrans, int pairs) {
    __shared__ float Are[TILE * LDA], Aim[TILE * LDA];
    __shared__ __attribute__((aligned(16))) float Bre[TILE2], Bim[TILE2];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int ngrp = (nchunk + NC - 1) / NC;
    const int64_t total = (int64_t)nrows * nmat * ngrp, per = (total + 7) / 8;
    const int64_t item = (int64_t)(blockIdx.x % 8) * per + blockIdx.x / 8;
    if (item >= total) return;
    const int r = (int)(item % nrows);
    const int64_t mc = item / nrows;
    const int m = (int)(mc / ngrp), c0 = (int)(mc % ngrp) * NC;
    const SolveSet st = solve_set(m, pairs, ctrans);
    ctrans = st.ct;
    const int i = rows[r];
    const int nc = min(NC, nchunk - c0);
    // the row's tiles of the group's steps (at most 32), looked up at once:
    // lane j holds the tile id of step g0 + j
    const int G = g1 - g0;
    const int32_t lane_id = lane >= G ? -1
                            : ctrans ? tile_id[(int64_t)(g0 + lane) * nt + i]
                                     : tile_id[(int64_t)i * nt + g0 + lane];
    const uint32_t present = (uint32_t)__ballot(lane_id >= 0);
    auto tile_of = [&](int k) { return __builtin_amdgcn_readlane(lane_id, k - g0); };
    auto next_k = [&](int k) {  // the first step >= k with a tile, else g1
        const int j = k - g0;
        const uint32_t rest = j >= 32 ? 0u : present & (~0u << j);
        return rest ? g0 + __builtin_ctz(rest) : g1;
    };
    floatx16 cre[NC], cim[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c)
        if (c < nc) load_acc(y + rhs_block(y_mat_stride, m, i, nchunk, c0 + c), cre[c], cim[c], wave, lane);
    for (int k = next_k(g0); k < g1; k = next_k(k + 1)) {
        const int32_t id = tile_of(k);
        const float *t = tiles + (int64_t)st.fac * mat_stride + (int64_t)id * TILE_FLOATS;
        __syncthreads();  // the previous products are done with LDS
        if (ctrans) stage_a_ctrans(t, Are, Aim, tid);
        else stage_a(t, Are, Aim, tid);
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            if (c >= nc) break;
            if (c > 0) __syncthreads();
            stage_b(y + rhs_block(y_mat_stride, m, k, nchunk, c0 + c), Bre, Bim, tid);
            __syncthreads();
            mma_fp32<true>(Are, Aim, Bre, Bim, cre[c], cim[c], wave, lane);
        }
    }
#pragma unroll
    for (int c = 0; c < NC; ++c)
        if (c < nc) store_acc(y + rhs_block(y_mat_stride, m, i, nchunk, c0 + c), cre[c], cim[c], wave, lane);
}

template __global__ void k_solve_trail<2>(float *, int64_t, int, const float *, int64_t, const int32_t *, int, int,
                                          int, const int32_t *, int, int, int, int);

// Stages thread tid's 4 x 4 block of an FP32 tile as the A operand P of a
// three-multiplication product in six planes, panel block order (row = M
// index, k = K index): P = -T (transpose = false) or -T^H (transpose and
// conj): per P row 4 consecutive k values, 8 bytes per plane.
__device__ __forceinline__ void stage_panel_block(const TileBlock &blk, uint16_t *S, bool transpose, bool conj,
                                                  int tid) {
    const int br = tid >> 4, bc = tid & 15;
    float re[4][4], im[4][4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const float4 vr = blk.re[q], vi = blk.im[q];
        re[q][0] = vr.x; re[q][1] = vr.y; re[q][2] = vr.z; re[q][3] = vr.w;
        im[q][0] = vi.x; im[q][1] = vi.y; im[q][2] = vi.z; im[q][3] = vi.w;
    }
    const float si = conj ? 1.f : -1.f;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int row = transpose ? 4 * bc + u : 4 * br + u;
        const int k0 = transpose ? 4 * br : 4 * bc;
        float xr[4], xi[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            xr[e] = -(transpose ? re[e][u] : re[u][e]);
            xi[e] = si * (transpose ? im[e][u] : im[u][e]);
        }
        uint32_t w0[6], w1[6];
        split6(xr[0], xi[0], xr[1], xi[1], w0);
        split6(xr[2], xi[2], xr[3], xi[3], w1);
        const int off = (((k0 >> 3) & 1) * 32 + (row & 31)) * 8 + (k0 & 7);
#pragma unroll
        for (int pl = 0; pl < 6; ++pl)
            *reinterpret_cast<uint2 *>(S + panel_block(k0 >> 4, pl >> 1, pl & 1, row >> 5) * SPLIT_BLOCK + off) =
                make_uint2(w0[pl], w1[pl]);
    }
}

// The same trailing update in the BF16 precisions, by the three-
// multiplication complex product: Y_i += P Y_k with P = -T_ik (or -T_ki^H),
//   T1 = Pr Yr, T2 = Pi Yi, T3 = (Pr + Pi)(Yr + Yi),
//   Re Y_i += T1 - T2,  Im Y_i += T3 - T1 - T2,
// each real product as hi*hi + hi*lo + lo*hi (three accumulators per 32 x 32
// block, started at x1 = Re, x2 = 0, x3 = Re + Im).  P is split once per step
// into LDS (six planes, 48 KiB) from the FP32 tile, which is loaded into
// registers a step ahead; Y_k comes split, transposed and in MFMA block order
// from its slot, straight from memory into registers, one or two K quarters
// ahead, so
// a step takes two barriers.  NC = 2: wave w computes the 64 rows x 32 columns
// of column half w % 2 of chunk c0 + w / 2 (two 32 x 32 blocks); NC = 1: the
// 32 x 32 block (w / 2, w % 2) of chunk c0.
template <int NC>
__global__ void __launch_bounds__(256, 2) k_solve_trail_3m(float *__restrict__ y, int64_t y_mat_stride, int nchunk,
                                                          const float *__restrict__ tiles, int64_t mat_stride,
                                                          const int32_t *__restrict__ tile_id, int nt, int g0, int g1,
                                                          const int32_t *__restrict__ rows, int nrows, int nmat,
                                                          int ctrans, int pairs, const uint16_t *__restrict__ ysplit) {
    __shared__ __attribute__((aligned(16))) uint16_t As[PANEL_TILE];
    constexpr int NB = NC == 2 ? 2 : 1;  // 32 x 32 output blocks per wave
    const int tid = threadIdx.x, wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
    const int ngrp = (nchunk + NC - 1) / NC;
    const int64_t total = (int64_t)nrows * nmat * ngrp, per = (total + 7) / 8;
    const int64_t item = (int64_t)(blockIdx.x % 8) * per + blockIdx.x / 8;
    if (item >= total) return;
    const int r = (int)(item % nrows);
    const int64_t mc = item / nrows;
    const int m = (int)(mc / ngrp), c0 = (int)(mc % ngrp) * NC;
    const SolveSet st = solve_set(m, pairs, ctrans);
    const int ct = st.ct;
    const int i = rows[r];
    const int wc = NC == 2 ? wave >> 1 : 0, wa = NC == 2 ? 0 : wave >> 1, wb = wave & 1;
    const int c = c0 + wc;
    const bool active = c < nchunk;  // NC = 2 and an odd chunk count: the last group's second chunk is absent
    const int G = g1 - g0;
    const int32_t lane_id = lane >= G ? -1
                            : ct ? tile_id[(int64_t)(g0 + lane) * nt + i] : tile_id[(int64_t)i * nt + g0 + lane];
    const uint32_t present = (uint32_t)__ballot(lane_id >= 0);
    auto tile_of = [&](int k) { return __builtin_amdgcn_readlane(lane_id, k - g0); };
    auto next_k = [&](int k) {
        const int j = k - g0;
        const uint32_t rest = j >= 32 ? 0u : present & (~0u << j);
        return rest ? g0 + __builtin_ctz(rest) : g1;
    };
    float *Yi = y + rhs_block(y_mat_stride, m, i, nchunk, active ? c : c0);
    floatx16 x[3][NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        if (active) {
            load_blk(Yi, x[0][b], x[2][b], wa + b, wb, lane);
            x[2][b] += x[0][b];
        } else {
            x[0][b] = x[2][b] = floatx16{};
        }
        x[1][b] = floatx16{};
    }
    const float *tm = tiles + (int64_t)st.fac * mat_stride;
    int k = next_k(g0);
    TileBlock tn;
    if (k < g1) tn = tile_block_load(tm + (int64_t)tile_of(k) * TILE_FLOATS, tid);
    // B operands of K quarter q of step kk: blocks (q, component, part, wb)
    // of the slot of Y_kk, chunk c
    // B loads run QA K quarters ahead through a ring of 2 QA register sets:
    // two with NC = 2 (the registers fit beside two waves per SIMD), one with
    // NC = 1 (which keeps three waves per SIMD)
    constexpr int QA = NC == 2 ? 2 : 1;
    u32x4 bq[2 * QA][3][2];
    auto fetch_b = [&](u32x4 (&dst)[3][2], int kk, int q) {
        const uint16_t *sl = yslot(ysplit, m, kk, nchunk, c) + lane * 8;
#pragma unroll
        for (int comp = 0; comp < 3; ++comp)
#pragma unroll
            for (int h = 0; h < 2; ++h)
                dst[comp][h] = *reinterpret_cast<const u32x4 *>(sl + panel_block(q, comp, h, wb) * SPLIT_BLOCK);
    };
    if (k < g1 && active)
#pragma unroll
        for (int q = 0; q < QA; ++q) fetch_b(bq[q], k, q);
    while (k < g1) {
        const int kn = next_k(k + 1);
        lds_barrier();  // the previous step's reads of As are done
        stage_panel_block(tn, As, ct != 0, ct != 0, tid);
        if (kn < g1) tn = tile_block_load(tm + (int64_t)tile_of(kn) * TILE_FLOATS, tid);
        lds_barrier();
        if (active) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                // the B operands QA quarters on (the next step's first ones
                // after this step's last), loading while the products of this
                // quarter and the next QA - 1 run
                if (q + QA < 4) fetch_b(bq[(q + QA) % (2 * QA)], k, q + QA);
                else if (kn < g1) fetch_b(bq[(q + QA) % (2 * QA)], kn, q + QA - 4);
#pragma unroll
                for (int comp = 0; comp < 3; ++comp) {
                    bf16x8 ah[NB], al[NB];
#pragma unroll
                    for (int b = 0; b < NB; ++b) {
                        ah[b] = lds8(As + panel_block(q, comp, 0, wa + b) * SPLIT_BLOCK + lane * 8);
                        al[b] = lds8(As + panel_block(q, comp, 1, wa + b) * SPLIT_BLOCK + lane * 8);
                    }
                    const bf16x8 bh = __builtin_bit_cast(bf16x8, bq[q % (2 * QA)][comp][0]);
                    const bf16x8 bl = __builtin_bit_cast(bf16x8, bq[q % (2 * QA)][comp][1]);
#pragma unroll
                    for (int pr = 0; pr < 3; ++pr)
#pragma unroll
                        for (int b = 0; b < NB; ++b)
                            x[comp][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(pr == 2 ? al[b] : ah[b],
                                                                                 pr == 1 ? bl : bh, x[comp][b], 0, 0,
                                                                                 0);
                }
            }
        }
        k = kn;
    }
    if (active) {
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            const floatx16 re = x[0][b] - x[1][b], im = x[2][b] - x[0][b] - x[1][b];
            store_blk(Yi, re, im, wa + b, wb, lane);
        }
    }
}

template __global__ void k_solve_trail_3m<2>(float *, int64_t, int, const float *, int64_t, const int32_t *, int, int,
                                             int, const int32_t *, int, int, int, int, const uint16_t *);
template __global__ void k_solve_trail_3m<1>(float *, int64_t, int, const float *, int64_t, const int32_t *, int, int,
                                             int, const int32_t *, int, int, int, int, const uint16_t *);

// The steps of one step group [g0, g1) of a sweep in one launch: block
// (matrix m, chunk c) takes the group's steps in sweep order (ascending when
// fwd, descending otherwise) and finishes each as
//   Y_k <- D_k (Y_k - sum_t T_kt Y_t),  t over the group's finished steps,
// with T_kt tile (k,t) or, with ctrans, the conjugate transpose of tile (t,k),
// and D_k as in k_solve_diag.  The sum runs over t in the order in which the
// step-at-a-time sweep (diag launch, then the update launch of the group's
// rows) would have applied the updates to Y_k, each product accumulated into
// the same FP32 accumulators, so the results are the same bit for bit; the
// group's ~2G launches become one.  Y_k stays in the accumulators across its
// updates.  FP32 MFMA form.
__global__ void __launch_bounds__(256) k_solve_group(float *__restrict__ y, int64_t y_mat_stride, int nchunk,
                                                     const float *__restrict__ tiles, int64_t mat_stride,
                                                     const int32_t *__restrict__ tile_id, int nt,
                                                     const float *__restrict__ dinv, int g0, int g1, int fwd,
                                                     int which, int ctrans, int pairs) {
    __shared__ float Are[TILE * LDA], Aim[TILE * LDA];
    __shared__ __attribute__((aligned(16))) float Bre[TILE2], Bim[TILE2];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int m = blockIdx.x / nchunk, c = blockIdx.x % nchunk;
    const SolveSet st = solve_set(m, pairs, ctrans);
    ctrans = st.ct;
    which ^= st.ct;
    const float *tm = tiles + (int64_t)st.fac * mat_stride;
    const int G = g1 - g0;
    auto tile_of = [&](int k, int t) { return ctrans ? tile_id[(int64_t)t * nt + k] : tile_id[(int64_t)k * nt + t]; };
    for (int s = 0; s < G; ++s) {
        const int k = fwd ? g0 + s : g1 - 1 - s;
        float *yk = y + rhs_block(y_mat_stride, m, k, nchunk, c);
        floatx16 cre, cim;
        load_acc(yk, cre, cim, wave, lane);
        for (int u = 0; u < s; ++u) {
            const int t = fwd ? g0 + u : g1 - 1 - u;
            const int32_t id = tile_of(k, t);
            if (id < 0) continue;
            const float *T = tm + (int64_t)id * TILE_FLOATS;
            __syncthreads();  // the previous products are done with LDS
            if (ctrans) stage_a_ctrans(T, Are, Aim, tid);
            else stage_a(T, Are, Aim, tid);
            stage_b(y + rhs_block(y_mat_stride, m, t, nchunk, c), Bre, Bim, tid);
            __syncthreads();
            mma_fp32<true>(Are, Aim, Bre, Bim, cre, cim, wave, lane);
        }
        // D_k times the updated Y_k, read back from memory as the B operand
        store_acc(yk, cre, cim, wave, lane);
        __syncthreads();
        const float *d = dinv + ((int64_t)st.fac * nt + k) * (2 * TILE_FLOATS) + which * TILE_FLOATS;
        if (ctrans) stage_a_ctrans(d, Are, Aim, tid);
        else stage_a(d, Are, Aim, tid);
        stage_b(yk, Bre, Bim, tid);
        __syncthreads();
        cre = floatx16{};
        cim = floatx16{};
        mma_fp32<false>(Are, Aim, Bre, Bim, cre, cim, wave, lane);
        store_acc(yk, cre, cim, wave, lane);
        __syncthreads();  // Y_k is visible to the whole block
    }
}

// Split B operand (planes re_hi, re_lo, im_hi, im_lo; rows = RHS columns,
// columns = RHS rows, stride SLD) of the wave's accumulator block: the layout
// stage_split(.., transpose = true, ..) gives for the same FP32 values.
__device__ __forceinline__ void acc_to_split(uint16_t *S, const floatx16 &cre, const floatx16 &cim, int wave,
                                             int lane) {
    const int col = (wave & 1) * 32 + (lane & 31);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int row = (wave >> 1) * 32 + 4 * (lane >> 5) + 8 * q;
        uint2 rh, rl, ih, il;
        split2(cre[4 * q], cre[4 * q + 1], rh.x, rl.x);
        split2(cre[4 * q + 2], cre[4 * q + 3], rh.y, rl.y);
        split2(cim[4 * q], cim[4 * q + 1], ih.x, il.x);
        split2(cim[4 * q + 2], cim[4 * q + 3], ih.y, il.y);
        const int off = col * SLD + row;
        *reinterpret_cast<uint2 *>(S + off) = rh;
        *reinterpret_cast<uint2 *>(S + SPL + off) = rl;
        *reinterpret_cast<uint2 *>(S + 2 * SPL + off) = ih;
        *reinterpret_cast<uint2 *>(S + 3 * SPL + off) = il;
    }
}

// The same with split operands (BF16 precisions), as a pipeline of
// operations: op (s, u) with u < s is the update of step s by finished step u
// (tile present), op (s, s) the diagonal product of step s.  While one op's
// products run, the next op's A tile (T_kt or D_k, FP32) and, from its slot,
// the split Y_t are already loading (the barriers between ops order LDS only);
// a diagonal op also loads the next step's Y.  The B operand of a diagonal op
// is the accumulator block written straight to LDS as split values, and the
// finished Y_k stays in LDS (Ys) for the next step, whose first update most
// often needs it; the finished Y_k also goes to memory (FP32, and split in its
// slot for the trailing update).
__global__ void __launch_bounds__(256, 1) k_solve_group_split(float *__restrict__ y, int64_t y_mat_stride,
                                                              int nchunk, const float *__restrict__ tiles,
                                                              int64_t mat_stride, const int32_t *__restrict__ tile_id,
                                                              int nt, const float *__restrict__ dinv, int g0, int g1,
                                                              int fwd, int which, int ctrans, int pairs,
                                                              uint16_t *__restrict__ ysplit) {
    __shared__ __attribute__((aligned(16))) uint16_t As[4 * SPL], Bs[4 * SPL], Ys[4 * SPL];
    __shared__ int32_t op_id[32 * 32];  // tile of update (s, u), u < s; -1 when absent
    __shared__ uint32_t op_mask[32];    // bit u of step s: update (s, u) present
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int m = blockIdx.x / nchunk, c = blockIdx.x % nchunk;
    const SolveSet st = solve_set(m, pairs, ctrans);
    ctrans = st.ct;
    which ^= st.ct;
    const float *tm = tiles + (int64_t)st.fac * mat_stride;
    const int G = g1 - g0;  // at most 32
    auto step_of = [&](int s) { return fwd ? g0 + s : g1 - 1 - s; };
    // the group's tile lookups, all at once: wave w takes steps w, w + 4, ..
    for (int s = wave; s < G; s += 4) {
        const int k = step_of(s), t = step_of(lane);
        const int32_t id = lane >= s ? -1 : ctrans ? tile_id[(int64_t)t * nt + k] : tile_id[(int64_t)k * nt + t];
        const uint32_t bits = (uint32_t)__ballot(id >= 0);
        if (lane < 32) op_id[s * 32 + lane] = id;
        if (lane == 0) op_mask[s] = bits;
    }
    __syncthreads();
    // the op after (s, u): the next present update of step s, else its
    // diagonal; after a diagonal, the first op of step s + 1
    auto advance = [&](int &s, int &u) {
        if (u == s) {
            ++s;
            u = -1;
        }
        if (s >= G) return;
        const uint32_t rest = u < 0 ? op_mask[s] : op_mask[s] & (~0u << (u + 1));
        u = rest ? __builtin_ctz(rest) : s;
    };
    auto a_tile = [&](int s, int u) -> const float * {
        if (u == s) return dinv + ((int64_t)st.fac * nt + step_of(s)) * (2 * TILE_FLOATS) + which * TILE_FLOATS;
        return tm + (int64_t)op_id[s * 32 + u] * TILE_FLOATS;
    };
    int s = 0, u = 0;  // step 0 has no updates: its diagonal comes first
    TileBlock an = tile_block_load(a_tile(0, 0), tid);
    u32x32 bn = {};
    floatx16 cre, cim;
    load_acc(y + rhs_block(y_mat_stride, m, step_of(0), nchunk, c), cre, cim, wave, lane);
    while (s < G) {
        const bool diag = u == s;
        int s2 = s, u2 = u;
        advance(s2, u2);
        lds_barrier();  // the previous products are done with LDS
        stage_split_block(an, As, ctrans != 0, !diag, ctrans != 0, tid);
        const uint16_t *B = Bs;
        if (diag) acc_to_split(Bs, cre, cim, wave, lane);
        else if (u == s - 1) B = Ys;  // the step finished last
        else put_panel4(Bs, bn, tid);
        // the next op's operands: finished steps other than the last are in
        // their slots (stored before an earlier full barrier)
        if (s2 < G) {
            an = tile_block_load(a_tile(s2, u2), tid);
            if (u2 < s2 - 1) bn = fetch_panel4(yslot(ysplit, m, step_of(u2), nchunk, c), tid);
        }
        floatx16 nre = {}, nim = {};
        if (diag && s + 1 < G) load_acc(y + rhs_block(y_mat_stride, m, step_of(s + 1), nchunk, c), nre, nim, wave, lane);
        lds_barrier();
        if (diag) {
            cre = floatx16{};
            cim = floatx16{};
        }
        mma_split<true>(As, B, cre, cim, wave, lane);
        if (diag) {
            const int k = step_of(s);
            store_acc(y + rhs_block(y_mat_stride, m, k, nchunk, c), cre, cim, wave, lane);
            store_acc_slot(yslot(ysplit, m, k, nchunk, c), cre, cim, wave, lane);
            // Ys: its last readers (the updates of step s by step s - 1) ran
            // before this op's first barrier
            acc_to_split(Ys, cre, cim, wave, lane);
            __syncthreads();  // the slot stores and Ys are visible to the whole block
            cre = nre;
            cim = nim;
        }
        s = s2;
        u = u2;
    }
}


// Y[m] (tiled planar, W = 64 * nchunk columns) <- src[src_index[m]]
// (row-major n x w_src complex64); columns >= w_src and padded rows are zero.
__global__ void k_pack_rhs(float *__restrict__ y, int64_t y_mat_stride, int nchunk, const float2 *__restrict__ src,
                           int w_src, const int32_t *__restrict__ src_index, const int32_t *__restrict__ perm,
                           int64_t n, int64_t padded) {
    const int m = blockIdx.y;
    const float2 *s = src + (int64_t)src_index[m] * n * w_src;
    const int W = nchunk * TILE;
    const int64_t total = padded * W, stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += stride) {
        const int64_t row = e / W;
        const int col = (int)(e % W);
        float2 v = make_float2(0.f, 0.f);
        const int32_t orb = perm[row];  // -1 on padding rows
        if (orb >= 0 && col < w_src) v = s[(int64_t)orb * w_src + col];
        float *blk = y + rhs_block(y_mat_stride, m, (int)(row / TILE), nchunk, col / TILE);
        const int off = (int)(row % TILE) * TILE + col % TILE;
        blk[off] = v.x;
        blk[PLANE + off] = v.y;
    }
}

// dst[m] (row-major n x w_dst, original order) <- Y[m]; iperm maps an orbital
// to its factor row
__global__ void k_unpack(float2 *__restrict__ dst, int w_dst, const float *__restrict__ y, int64_t y_mat_stride,
                         int nchunk, const int32_t *__restrict__ iperm, int64_t n) {
    const int m = blockIdx.y;
    float2 *d = dst + (int64_t)m * n * w_dst;
    const int64_t total = n * w_dst, stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += stride) {
        const int64_t orb = e / w_dst;
        const int col = (int)(e % w_dst);
        const int64_t row = iperm[orb];
        const float *blk = y + rhs_block(y_mat_stride, m, (int)(row / TILE), nchunk, col / TILE);
        const int off = (int)(row % TILE) * TILE + col % TILE;
        d[e] = make_float2(blk[off], blk[PLANE + off]);
    }
}

// r[m] = b[b_index[m]] - (H_k - z_m S_k) x[m], FP64 accumulation, rounded to
// complex64; sumsq[m] += ||r||^2 and bsq[m] += ||b||^2 (FP64).  One wave per
// row, lanes over columns.
__global__ void __launch_bounds__(256) k_residual_shifted(
    float2 *__restrict__ r, const float2 *__restrict__ b, const int32_t *__restrict__ b_index,
    const float2 *__restrict__ x, int w, const float2 *__restrict__ h, const float2 *__restrict__ s,
    const int32_t *__restrict__ kidx, const float2 *__restrict__ z, const int32_t *__restrict__ indptr,
    const int32_t *__restrict__ indices, int64_t n, int64_t nnz, double *__restrict__ sumsq,
    double *__restrict__ bsq) {
    const int m = blockIdx.y;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const float2 *hk = h + (int64_t)kidx[m] * nnz, *sk = s + (int64_t)kidx[m] * nnz;
    const float2 *xm = x + (int64_t)m * n * w;
    const float2 *bm = b + (int64_t)b_index[m] * n * w;
    float2 *rm = r + (int64_t)m * n * w;
    const double zr = z[m].x, zi = z[m].y;
    double local = 0.0, blocal = 0.0;
    for (int64_t row = (int64_t)blockIdx.x * 4 + wave; row < n; row += (int64_t)gridDim.x * 4) {
        for (int c0 = 0; c0 < w; c0 += 64) {
            const int c = c0 + lane;
            if (c >= w) break;
            double ar = 0.0, ai = 0.0;
            for (int32_t p = indptr[row]; p < indptr[row + 1]; ++p) {
                const float2 hv = hk[p], sv = sk[p];
                const double vr = (double)hv.x - (zr * sv.x - zi * sv.y);
                const double vi = (double)hv.y - (zr * sv.y + zi * sv.x);
                const float2 xv = xm[(int64_t)indices[p] * w + c];
                ar += vr * xv.x - vi * xv.y;
                ai += vr * xv.y + vi * xv.x;
            }
            const float2 bv = bm[row * w + c];
            const double rr = bv.x - ar, ri = bv.y - ai;
            rm[row * w + c] = make_float2((float)rr, (float)ri);
            local += rr * rr + ri * ri;
            blocal += (double)bv.x * bv.x + (double)bv.y * bv.y;
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        local += __shfl_down(local, off);
        blocal += __shfl_down(blocal, off);
    }
    if (lane == 0) {
        atomicAdd(sumsq + m, local);
        atomicAdd(bsq + m, blocal);
    }
}

// x += d, elementwise over `count` complex values.
__global__ void k_axpy(float2 *__restrict__ x, const float2 *__restrict__ d, int64_t count) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < count; e += stride) {
        float2 v = x[e];
        const float2 u = d[e];
        v.x += u.x;
        v.y += u.y;
        x[e] = v;
    }
}

}  // namespace aosx
