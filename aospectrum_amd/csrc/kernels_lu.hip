// Tile LU with static pivoting for batches of shifted matrices A_m = H - z_m S.
//
// Storage: every matrix of a batch owns n_tiles tiles of TILE x TILE complex
// entries, each stored as a row-major real plane followed by a row-major
// imaginary plane (planar layout feeds the FP32 MFMA operands directly).  The
// tile structure comes from the host symbolic analysis; for step k the tiles
// next[k] hold L(i,k) and, by symmetry of the pattern, U(k,i).
//
// Steps are taken in groups of consecutive steps [g0, g1): left-looking
// inside a group, right-looking across groups.  Per step k of a group:
//   crout  : A(i,j) -= sum_{t in [g0, k)} L(i,t) * U(t,j) for the tiles of
//            row and column k ((k,k), (i,k), (k,i), i in next[k])        (MFMA)
//   getrf  : factor the diagonal tile without row exchanges; a pivot below
//            1e-6 * max|tile| is replaced by that bound (static pivoting in the
//            sense of Li & Demmel's GESP, corrected later by refinement);
//            also writes inv(L_kk) (unit lower) and inv(U_kk).
//   panel  : L(i,k) <- L(i,k) * inv(U_kk),  U(k,i) <- inv(L_kk) * U(k,i)   (MFMA)
// and after the group's last step one trailing launch applies
//   A(i,j) -= sum_{k in group} L(i,k) * U(k,j)   for i, j >= g1.
// Both update launches keep the target in the MFMA accumulators across the
// steps they apply, so every target is read and written once per group.
// The complex products use four real FP32 MFMA products (v_mfma_f32_32x32x2_f32),
// i.e. 8 real flops per complex multiply-add.  In the BF16 precisions the
// panel and Schur products split every operand into hi + lo BF16 parts for
// v_mfma_f32_32x32x16_bf16 (three products, FP32 accumulation), the panels
// also write their tiles in that split form for the group's updates, and the
// stored factor stays FP32.
#include "device_common.h"

namespace aosx {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

__global__ void k_fill_zero(float4 *__restrict__ p, int64_t count4) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count4; i += stride)
        p[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// A_m = H_k - z_m S_k scattered into the tile storage of matrix m
// (matrix m uses k point kidx[m] and shift z[m]); padding rows get a unit
// diagonal so their tiles stay nonsingular.
__global__ void k_scatter_shifted(float *__restrict__ tiles, int64_t mat_stride, const float2 *__restrict__ h,
                                  const float2 *__restrict__ s, int64_t nnz, const int32_t *__restrict__ kidx,
                                  const float2 *__restrict__ z, const int32_t *__restrict__ rows,
                                  const int32_t *__restrict__ cols, const int32_t *__restrict__ iperm,
                                  const int32_t *__restrict__ tile_id, int nt,
                                  const int32_t *__restrict__ padding_rows, int n_padding) {
    const int m = blockIdx.y;
    const float2 zm = z[m];
    const float2 *hk = h + (int64_t)kidx[m] * nnz;
    const float2 *sk = s + (int64_t)kidx[m] * nnz;
    float *base = tiles + (int64_t)m * mat_stride;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < nnz; p += stride) {
        const float2 hv = hk[p], sv = sk[p];
        const float re = hv.x - (zm.x * sv.x - zm.y * sv.y);
        const float im = hv.y - (zm.x * sv.y + zm.y * sv.x);
        const int32_t fr = iperm[rows[p]], fc = iperm[cols[p]];
        const int32_t id = tile_id[(int64_t)(fr / TILE) * nt + fc / TILE];
        if (id < 0) continue;  // cannot happen for a closed tile structure
        float *t = base + (int64_t)id * TILE_FLOATS + (fr % TILE) * TILE + fc % TILE;
        t[0] = re;
        t[PLANE] = im;
    }
    if (blockIdx.x == 0) {
        for (int q = threadIdx.x; q < n_padding; q += blockDim.x) {
            const int32_t r = padding_rows[q];
            const int32_t id = tile_id[(int64_t)(r / TILE) * nt + r / TILE];
            base[(int64_t)id * TILE_FLOATS + (r % TILE) * TILE + r % TILE] = 1.0f;
        }
    }
}

// ----------------------------------------------------------------- getrf
__global__ void __launch_bounds__(256) k_getrf_diag(float *__restrict__ tiles, int64_t mat_stride,
                                                   float *__restrict__ dinv, int nt, int k,
                                                   const int32_t *__restrict__ tile_id,
                                                   int *__restrict__ replaced, float pivot_rel) {
    __shared__ float Are[TILE * LDA], Aim[TILE * LDA];
    __shared__ float Lre[TILE * LDA], Lim[TILE * LDA];
    __shared__ float Ure[TILE * LDA], Uim[TILE * LDA];
    __shared__ float red[256];
    const int m = blockIdx.x, tid = threadIdx.x;
    float *t = tile_ptr(tiles, mat_stride, m, tile_id[(int64_t)k * nt + k]);
    float amax = 0.f;
    for (int e = tid; e < TILE2; e += 256) {
        const int r = e / TILE, c = e % TILE;
        const float re = t[e], im = t[PLANE + e];
        Are[r * LDA + c] = re;
        Aim[r * LDA + c] = im;
        amax = fmaxf(amax, re * re + im * im);
    }
    red[tid] = amax;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (tid < w) red[tid] = fmaxf(red[tid], red[tid + w]);
        __syncthreads();
    }
    float tau = pivot_rel * sqrtf(red[0]);
    if (!(tau > 0.f)) tau = 1.0e-30f;
    // Right-looking elimination, one barrier per step: every thread derives
    // the (possibly replaced) pivot from A(kk,kk) itself, updates its share of
    // the trailing block with the unscaled multipliers times 1/pivot, and the
    // scaled column kk is written after the barrier (no later step reads it).
    // Thread (ti, tj) = (tid / 16, tid % 16) owns rows kk+1+ti+16a and
    // columns kk+1+tj+16b.
    const int ti = tid >> 4, tj = tid & 15;
    int local_replaced = 0;
    for (int kk = 0; kk < TILE; ++kk) {
        float pr = Are[kk * LDA + kk], pi = Aim[kk * LDA + kk];
        const float mag = sqrtf(pr * pr + pi * pi);
        const bool rep = !(mag >= tau);
        if (rep) {
            if (mag > 0.f) { pr = pr / mag * tau; pi = pi / mag * tau; }
            else { pr = tau; pi = 0.f; }
        }
        const float d = pr * pr + pi * pi;
        const float ir = pr / d, ii = -pi / d;
        float lr[4], li[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            const int i = kk + 1 + ti + 16 * a;
            lr[a] = li[a] = 0.f;
            if (i < TILE) {
                const float ar = Are[i * LDA + kk], ai = Aim[i * LDA + kk];
                lr[a] = ar * ir - ai * ii;
                li[a] = ar * ii + ai * ir;
            }
        }
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int j = kk + 1 + tj + 16 * b;
            if (j >= TILE) break;
            const float ur = Are[kk * LDA + j], ui = Aim[kk * LDA + j];
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                const int i = kk + 1 + ti + 16 * a;
                if (i < TILE) {
                    Are[i * LDA + j] -= lr[a] * ur - li[a] * ui;
                    Aim[i * LDA + j] -= lr[a] * ui + li[a] * ur;
                }
            }
        }
        __syncthreads();
        if (tj == 0) {
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                const int i = kk + 1 + ti + 16 * a;
                if (i < TILE) {
                    Are[i * LDA + kk] = lr[a];
                    Aim[i * LDA + kk] = li[a];
                }
            }
        }
        if (tid == 0 && rep) {
            Are[kk * LDA + kk] = pr;
            Aim[kk * LDA + kk] = pi;
            ++local_replaced;
        }
    }
    __syncthreads();
    if (tid == 0 && local_replaced) atomicAdd(replaced + m, local_replaced);
    for (int e = tid; e < TILE2; e += 256) {
        const int r = e / TILE, c = e % TILE;
        t[e] = Are[r * LDA + c];
        t[PLANE + e] = Aim[r * LDA + c];
    }
    // inv(L) (unit lower) by columns and inv(U) by rows; four threads share a
    // column (row) and split each inner product, combined across the four
    // adjacent lanes:
    //   X(i,c) = -sum_{j=c}^{i-1} L(i,j) X(j,c),           X(c,c) = 1
    //   Y(r,j) = (delta_rj - sum_{i=r}^{j-1} Y(r,i) U(i,j)) / U(j,j)
    const int col = tid >> 2, part = tid & 3;
    for (int i = 0; i < TILE; ++i) {
        float sr = 0.f, si = 0.f;
        if (i > col) {
            for (int j = col + part; j < i; j += 4) {
                const float ar = Are[i * LDA + j], ai = Aim[i * LDA + j];
                const float xr = Lre[j * LDA + col], xi = Lim[j * LDA + col];
                sr += ar * xr - ai * xi;
                si += ar * xi + ai * xr;
            }
        }
        sr += __shfl_xor(sr, 1);
        si += __shfl_xor(si, 1);
        sr += __shfl_xor(sr, 2);
        si += __shfl_xor(si, 2);
        if (part == 0) {
            Lre[i * LDA + col] = i < col ? 0.f : (i == col ? 1.f : -sr);
            Lim[i * LDA + col] = i <= col ? 0.f : -si;
        }
        __syncthreads();
    }
    const int row = col;
    for (int j = 0; j < TILE; ++j) {
        float sr = 0.f, si = 0.f;
        if (j > row) {
            for (int i = row + part; i < j; i += 4) {
                const float yr = Ure[row * LDA + i], yi = Uim[row * LDA + i];
                const float ur = Are[i * LDA + j], ui = Aim[i * LDA + j];
                sr += yr * ur - yi * ui;
                si += yr * ui + yi * ur;
            }
        }
        sr += __shfl_xor(sr, 1);
        si += __shfl_xor(si, 1);
        sr += __shfl_xor(sr, 2);
        si += __shfl_xor(si, 2);
        if (part == 0) {
            if (j < row) {
                Ure[row * LDA + j] = 0.f;
                Uim[row * LDA + j] = 0.f;
            } else {
                const float nr = (j == row ? 1.f : 0.f) - sr, ni = -si;
                const float dr = Are[j * LDA + j], di = Aim[j * LDA + j];
                const float dd = dr * dr + di * di;
                Ure[row * LDA + j] = (nr * dr + ni * di) / dd;
                Uim[row * LDA + j] = (ni * dr - nr * di) / dd;
            }
        }
        __syncthreads();
    }
    float *dv = dinv + ((int64_t)m * nt + k) * (2 * TILE_FLOATS);
    for (int e = tid; e < TILE2; e += 256) {
        const int r = e / TILE, c = e % TILE;
        dv[e] = Lre[r * LDA + c];
        dv[PLANE + e] = Lim[r * LDA + c];
        dv[2 * PLANE + e] = Ure[r * LDA + c];
        dv[3 * PLANE + e] = Uim[r * LDA + c];
    }
}

// ----------------------------------------------------------- MFMA tile GEMM
// Stage a stored tile into LDS: A operand row-major with stride LDA, B operand
// row-major with stride TILE.
__device__ __forceinline__ void stage_a(const float *__restrict__ g, float *Are, float *Aim, int tid) {
    const float4 *g4 = reinterpret_cast<const float4 *>(g);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int e4 = q * 256 + tid;
        const int r = (e4 * 4) / TILE, c = (e4 * 4) % TILE;
        const float4 vr = g4[e4], vi = g4[PLANE / 4 + e4];
        float *ar = Are + r * LDA + c, *ai = Aim + r * LDA + c;
        ar[0] = vr.x; ar[1] = vr.y; ar[2] = vr.z; ar[3] = vr.w;
        ai[0] = vi.x; ai[1] = vi.y; ai[2] = vi.z; ai[3] = vi.w;
    }
}
__device__ __forceinline__ void stage_b(const float *__restrict__ g, float *Bre, float *Bim, int tid) {
    const float4 *g4 = reinterpret_cast<const float4 *>(g);
    float4 *br = reinterpret_cast<float4 *>(Bre), *bi = reinterpret_cast<float4 *>(Bim);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int e4 = q * 256 + tid;
        br[e4] = g4[e4];
        bi[e4] = g4[PLANE / 4 + e4];
    }
}

// acc += A * B (or acc -= A * B when SUB) for this wave's 32x32 output block.
template <bool SUB>
__device__ __forceinline__ void mma_fp32(const float *Are, const float *Aim, const float *Bre, const float *Bim,
                                         floatx16 &cre, floatx16 &cim, int wave, int lane) {
    const int i = (wave >> 1) * 32 + (lane & 31);
    const int j = (wave & 1) * 32 + (lane & 31);
    const int kh = lane >> 5;
#pragma unroll 8
    for (int k0 = 0; k0 < TILE; k0 += 2) {
        const int k = k0 + kh;
        float ar = Are[i * LDA + k], ai = Aim[i * LDA + k];
        const float br = Bre[k * TILE + j], bi = Bim[k * TILE + j];
        if (SUB) { ar = -ar; ai = -ai; }
        cre = __builtin_amdgcn_mfma_f32_32x32x2f32(ar, br, cre, 0, 0, 0);
        cre = __builtin_amdgcn_mfma_f32_32x32x2f32(-ai, bi, cre, 0, 0, 0);
        cim = __builtin_amdgcn_mfma_f32_32x32x2f32(ar, bi, cim, 0, 0, 0);
        cim = __builtin_amdgcn_mfma_f32_32x32x2f32(ai, br, cim, 0, 0, 0);
    }
}

// Accumulator <-> tile element mapping of v_mfma_f32_32x32x*: register r of
// lane l holds row 8*(r/4) + 4*(l/32) + r%4, column l%32 of the wave's block.
__device__ __forceinline__ void load_acc(const float *__restrict__ t, floatx16 &cre, floatx16 &cim, int wave,
                                         int lane) {
    const int r0 = (wave >> 1) * 32 + 4 * (lane >> 5), c = (wave & 1) * 32 + (lane & 31);
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int row = r0 + 8 * (r >> 2) + (r & 3);
        cre[r] = t[row * TILE + c];
        cim[r] = t[PLANE + row * TILE + c];
    }
}
__device__ __forceinline__ void store_acc(float *__restrict__ t, const floatx16 &cre, const floatx16 &cim,
                                          int wave, int lane) {
    const int r0 = (wave >> 1) * 32 + 4 * (lane >> 5), c = (wave & 1) * 32 + (lane & 31);
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int row = r0 + 8 * (r >> 2) + (r & 3);
        t[row * TILE + c] = cre[r];
        t[PLANE + row * TILE + c] = cim[r];
    }
}

// Split copies of panel tiles for the BF16 trailing update.  A split tile is
// four planes of 64 x 64 BF16 -- re_hi, re_lo, im_hi, im_lo, with
// x ~ hi + lo, hi = bf16(x), lo = bf16(x - hi) (round to nearest even) --
// holding -L(i,k) row-major ([row][col]) or U(k,j) transposed ([col][row]), so
// that both MFMA operands read 8 consecutive k values per lane.  The tile is
// stored in MFMA operand order: 1 KiB blocks, one per (K half, plane, 32-row
// half a, 16-k step within the half), each holding the 32 x 16 operand of one
// v_mfma_f32_32x32x16_bf16 as lane l of the wave reads it -- row 32a + l % 32,
// k values 8 (l / 32) .. + 7 -- in lane order.  A K half of all four planes is
// 16 KiB of consecutive memory, and a wave reads an operand block from LDS
// as one linear, bank-conflict-free 16-byte-per-lane read.
constexpr int SPLIT_TILE = 4 * TILE2;  // BF16 elements per split tile (32 KiB)
constexpr int SPLIT_HPLANE = TILE * 32;      // one plane of one K half (4 blocks)
constexpr int SPLIT_HALF = 4 * SPLIT_HPLANE; // all planes of one K half (16 blocks)
constexpr int SPLIT_BLOCK = 512;             // BF16 per operand block
__device__ __forceinline__ int split_index(int plane, int row, int k) {
    return (k >> 5) * SPLIT_HALF + plane * SPLIT_HPLANE + (row >> 5) * 1024 + ((k >> 4) & 1) * SPLIT_BLOCK +
           ((k >> 3) & 1) * 256 + (row & 31) * 8 + (k & 7);
}

__device__ __forceinline__ void split_pair(float x, uint16_t &hi, uint16_t &lo) {
    const __bf16 h = (__bf16)x;
    const __bf16 l = (__bf16)(x - (float)h);
    hi = __builtin_bit_cast(uint16_t, h);
    lo = __builtin_bit_cast(uint16_t, l);
}

// Writes the tile held in LDS (re/im planes, row stride LDA) as a split tile.
// negate_rows: -T row-major (the L operand); otherwise T transposed (the U
// operand).  Thread tid covers 16 consecutive destination elements of one
// destination row, written as 2 x 16 B per plane.
__device__ __forceinline__ void write_split(const float *Tre, const float *Tim, uint16_t *__restrict__ dst,
                                            bool negate_rows, int tid) {
    const int r = tid >> 2, c0 = (tid & 3) * 16;
    uint16_t rh[16], rl[16], ih[16], il[16];
#pragma unroll
    for (int c = 0; c < 16; ++c) {
        float xr, xi;
        if (negate_rows) {
            xr = -Tre[r * LDA + c0 + c];
            xi = -Tim[r * LDA + c0 + c];
        } else {
            xr = Tre[(c0 + c) * LDA + r];
            xi = Tim[(c0 + c) * LDA + r];
        }
        split_pair(xr, rh[c], rl[c]);
        split_pair(xi, ih[c], il[c]);
    }
    // k = c0 .. c0 + 7 and c0 + 8 .. c0 + 15: two 8-element runs of one block
    const int off = split_index(0, r, c0), off8 = split_index(0, r, c0 + 8);
    auto pack = [](const uint16_t *v, int o) {
        return make_uint4(v[o] | ((uint32_t)v[o + 1] << 16), v[o + 2] | ((uint32_t)v[o + 3] << 16),
                          v[o + 4] | ((uint32_t)v[o + 5] << 16), v[o + 6] | ((uint32_t)v[o + 7] << 16));
    };
    const uint16_t *planes[4] = {rh, rl, ih, il};
#pragma unroll
    for (int pl = 0; pl < 4; ++pl) {
        *reinterpret_cast<uint4 *>(dst + pl * SPLIT_HPLANE + off) = pack(planes[pl], 0);
        *reinterpret_cast<uint4 *>(dst + pl * SPLIT_HPLANE + off8) = pack(planes[pl], 8);
    }
}

// Panel step k: block x < s updates L(next[x], k) <- L * inv(U_kk); block
// x >= s updates U(k, next[x - s]) <- inv(L_kk) * U.  FP32 MFMA (the BF16
// precisions use k_panel_split).
__global__ void __launch_bounds__(256) k_panel(float *__restrict__ tiles, int64_t mat_stride,
                                              const float *__restrict__ dinv, int nt, int k,
                                              const int32_t *__restrict__ next, int s,
                                              const int32_t *__restrict__ tile_id) {
    __shared__ float Are[TILE * LDA], Aim[TILE * LDA];
    __shared__ __attribute__((aligned(16))) float Bre[TILE2], Bim[TILE2];
    const int m = blockIdx.y, tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const bool lower = (int)blockIdx.x < s;
    const int other = next[lower ? blockIdx.x : blockIdx.x - s];
    const int32_t id = lower ? tile_id[(int64_t)other * nt + k] : tile_id[(int64_t)k * nt + other];
    float *t = tile_ptr(tiles, mat_stride, m, id);
    const float *d = dinv + ((int64_t)m * nt + k) * (2 * TILE_FLOATS);
    if (lower) {
        stage_a(t, Are, Aim, tid);                 // L(i,k)
        stage_b(d + TILE_FLOATS, Bre, Bim, tid);   // inv(U_kk)
    } else {
        stage_a(d, Are, Aim, tid);                 // inv(L_kk)
        stage_b(t, Bre, Bim, tid);                 // U(k,j)
    }
    __syncthreads();
    floatx16 cre = {}, cim = {};
    mma_fp32<false>(Are, Aim, Bre, Bim, cre, cim, wave, lane);
    store_acc(t, cre, cim, wave, lane);
}

// Multi-step update: A(i,j) -= sum_t L(i,t) U(t,j) over the steps t in
// [g0, g1) where both tiles exist, for every pair (i, j) of the list (a
// group's trailing targets, or the row/column targets of one step with the
// group's earlier steps).  The target stays in the MFMA accumulators across
// the steps, so it is read and written once rather than once per step.  The
// grid is 8 * ceil(npairs * nmat / 8) blocks in one
// dimension; consecutive blocks go round-robin to the 8 XCDs, so block b works
// on item (b % 8) * per + b / 8 and each XCD (with its own L2) takes one
// contiguous run of the (matrix, pair) items, whose pairs share L rows and U
// columns.  FP32 MFMA (the BF16 precisions use k_schur_group_split).
__global__ void __launch_bounds__(256) k_schur_group(float *__restrict__ tiles, int64_t mat_stride, int nt, int g0,
                                                    const int32_t *__restrict__ pairs,
                                                    const int32_t *__restrict__ pair_mask, int npairs, int nmat,
                                                    const int32_t *__restrict__ tile_id) {
    __shared__ float Are[TILE * LDA], Aim[TILE * LDA];
    __shared__ __attribute__((aligned(16))) float Bre[TILE2], Bim[TILE2];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int64_t total = (int64_t)npairs * nmat, per = (total + 7) / 8;
    const int64_t item = (int64_t)(blockIdx.x % 8) * per + blockIdx.x / 8;
    if (item >= total) return;
    const int m = (int)(item / npairs), q = (int)(item % npairs);
    const int i = pairs[2 * q], j = pairs[2 * q + 1];
    float *C = tile_ptr(tiles, mat_stride, m, tile_id[(int64_t)i * nt + j]);
    floatx16 cre, cim;
    load_acc(C, cre, cim, wave, lane);
    for (uint32_t mask = (uint32_t)pair_mask[q]; mask; mask &= mask - 1) {
        const int t = g0 + __builtin_ctz(mask);
        const int32_t lid = tile_id[(int64_t)i * nt + t], uid = tile_id[(int64_t)t * nt + j];
        __syncthreads();  // the previous step's products are done with LDS
        stage_a(tile_ptr(tiles, mat_stride, m, lid), Are, Aim, tid);
        stage_b(tile_ptr(tiles, mat_stride, m, uid), Bre, Bim, tid);
        __syncthreads();
        mma_fp32<true>(Are, Aim, Bre, Bim, cre, cim, wave, lane);
    }
    store_acc(C, cre, cim, wave, lane);
}

// The BF16 update kernels below run over the matrices listed in `mats` (nmat
// entries): with per-shift precision the matrices that take one BF16 product
// and those that take three are updated by launches of their own.
//
// Multi-step update from split operands (BF16 precisions): the same pairs
// and step loop as k_schur_group, with L and U read from the
// split tiles the panels of the group wrote.  Slot of L(i,t) in the matrix's
// split area: ppos[t * nt + i] (-1 when absent); slot of U(t,j):
// ppos[t * nt + j] + step_s[t].  The next step's two split tiles are loaded
// into registers while the MFMA products of the current step run.
constexpr int SLD = TILE + 8;          // LDS row stride (BF16) of a split plane: 144 B
constexpr int SPL = TILE * SLD;        // LDS elements per split plane

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ bf16x8 lds8(const uint16_t *p) { return *reinterpret_cast<const bf16x8 *>(p); }
__device__ __forceinline__ bf16x8 neg8(bf16x8 v) {
    return __builtin_bit_cast(bf16x8, __builtin_bit_cast(u32x4, v) ^ (u32x4)(0x80008000u));
}

// acc += P * U with P = -L, both split: re += Pr Ur - Pi Ui, im += Pr Ui + Pi Ur,
// each real product as hi*hi + hi*lo + lo*hi (THREE) or hi*hi.
template <bool THREE>
__device__ __forceinline__ void mma_split(const uint16_t *As, const uint16_t *Bs, floatx16 &cre, floatx16 &cim,
                                          int wave, int lane) {
    const int i = (wave >> 1) * 32 + (lane & 31);
    const int j = (wave & 1) * 32 + (lane & 31);
    const int kb = (lane >> 5) * 8;
#pragma unroll
    for (int k0 = 0; k0 < TILE; k0 += 16) {
        const int ao = i * SLD + k0 + kb, bo = j * SLD + k0 + kb;
        const bf16x8 prh = lds8(As + ao), pih = lds8(As + 2 * SPL + ao);
        const bf16x8 urh = lds8(Bs + bo), uih = lds8(Bs + 2 * SPL + bo);
        const bf16x8 npih = neg8(pih);
        cre = __builtin_amdgcn_mfma_f32_32x32x16_bf16(prh, urh, cre, 0, 0, 0);
        cre = __builtin_amdgcn_mfma_f32_32x32x16_bf16(npih, uih, cre, 0, 0, 0);
        cim = __builtin_amdgcn_mfma_f32_32x32x16_bf16(prh, uih, cim, 0, 0, 0);
        cim = __builtin_amdgcn_mfma_f32_32x32x16_bf16(pih, urh, cim, 0, 0, 0);
        if (THREE) {
            const bf16x8 prl = lds8(As + SPL + ao), pil = lds8(As + 3 * SPL + ao);
            const bf16x8 url = lds8(Bs + SPL + bo), uil = lds8(Bs + 3 * SPL + bo);
            const bf16x8 npil = neg8(pil);
            cre = __builtin_amdgcn_mfma_f32_32x32x16_bf16(prh, url, cre, 0, 0, 0);
            cre = __builtin_amdgcn_mfma_f32_32x32x16_bf16(prl, urh, cre, 0, 0, 0);
            cre = __builtin_amdgcn_mfma_f32_32x32x16_bf16(npih, uil, cre, 0, 0, 0);
            cre = __builtin_amdgcn_mfma_f32_32x32x16_bf16(npil, uih, cre, 0, 0, 0);
            cim = __builtin_amdgcn_mfma_f32_32x32x16_bf16(prh, uil, cim, 0, 0, 0);
            cim = __builtin_amdgcn_mfma_f32_32x32x16_bf16(prl, uih, cim, 0, 0, 0);
            cim = __builtin_amdgcn_mfma_f32_32x32x16_bf16(pih, url, cim, 0, 0, 0);
            cim = __builtin_amdgcn_mfma_f32_32x32x16_bf16(pil, urh, cim, 0, 0, 0);
        }
    }
}

// Stages an FP32 planar tile G (64 x 64 complex, row-major) into LDS as a
// split operand S (planes re_hi, re_lo, im_hi, im_lo; row stride SLD) whose
// rows are the MFMA's M or N index and whose columns are its K index:
//   transpose = false: S[r][c] = f(G[r][c])   (G contracted over its columns)
//   transpose = true : S[r][c] = f(G[c][r])   (G contracted over its rows)
// with f = negation when `negate`, complex conjugation when `conj`.  Thread
// tid converts the 4 x 4 block (rows 4*(tid/16).., columns 4*(tid%16)..) and
// writes 4 runs of 4 consecutive K values per plane.
// Thread tid's 4 x 4 block of an FP32 planar tile, as loaded from memory.
struct TileBlock {
    float4 re[4], im[4];
};
__device__ __forceinline__ TileBlock tile_block_load(const float *__restrict__ g, int tid) {
    const int br = tid >> 4, bc = tid & 15;
    const float4 *g4 = reinterpret_cast<const float4 *>(g);
    TileBlock b;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        b.re[q] = g4[(4 * br + q) * (TILE / 4) + bc];
        b.im[q] = g4[PLANE / 4 + (4 * br + q) * (TILE / 4) + bc];
    }
    return b;
}
__device__ __forceinline__ void stage_split_block(const TileBlock &blk, uint16_t *S, bool transpose, bool negate,
                                                  bool conj, int tid) {
    const int br = tid >> 4, bc = tid & 15;
    float re[4][4], im[4][4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const float4 vr = blk.re[q], vi = blk.im[q];
        re[q][0] = vr.x; re[q][1] = vr.y; re[q][2] = vr.z; re[q][3] = vr.w;
        im[q][0] = vi.x; im[q][1] = vi.y; im[q][2] = vi.z; im[q][3] = vi.w;
    }
    const float sr = negate ? -1.f : 1.f, si = (negate != conj) ? -1.f : 1.f;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        // run u: S row and its 4 consecutive K values
        const int row = transpose ? 4 * bc + u : 4 * br + u;
        const int k0 = transpose ? 4 * br : 4 * bc;
        uint16_t rh[4], rl[4], ih[4], il[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const float xr = sr * (transpose ? re[e][u] : re[u][e]);
            const float xi = si * (transpose ? im[e][u] : im[u][e]);
            split_pair(xr, rh[e], rl[e]);
            split_pair(xi, ih[e], il[e]);
        }
        const int off = row * SLD + k0;
        *reinterpret_cast<uint2 *>(S + off) = make_uint2(rh[0] | ((uint32_t)rh[1] << 16), rh[2] | ((uint32_t)rh[3] << 16));
        *reinterpret_cast<uint2 *>(S + SPL + off) =
            make_uint2(rl[0] | ((uint32_t)rl[1] << 16), rl[2] | ((uint32_t)rl[3] << 16));
        *reinterpret_cast<uint2 *>(S + 2 * SPL + off) =
            make_uint2(ih[0] | ((uint32_t)ih[1] << 16), ih[2] | ((uint32_t)ih[3] << 16));
        *reinterpret_cast<uint2 *>(S + 3 * SPL + off) =
            make_uint2(il[0] | ((uint32_t)il[1] << 16), il[2] | ((uint32_t)il[3] << 16));
    }
}
__device__ __forceinline__ void stage_split(const float *__restrict__ g, uint16_t *S, bool transpose, bool negate,
                                            bool conj, int tid) {
    stage_split_block(tile_block_load(g, tid), S, transpose, negate, conj, tid);
}

// Panel step k in the BF16 precisions: the products of k_panel with both
// operands split into BF16 hi + lo parts (three MFMA products, FP32
// accumulation, the accuracy of the Schur updates); writes the FP32 tile and
// its split copy.
__global__ void __launch_bounds__(256, 2) k_panel_split(float *__restrict__ tiles, int64_t mat_stride,
                                                       const float *__restrict__ dinv, int nt, int k,
                                                       const int32_t *__restrict__ next, int s,
                                                       const int32_t *__restrict__ tile_id,
                                                       uint16_t *__restrict__ split, int64_t split_stride,
                                                       int slot0) {
    __shared__ __attribute__((aligned(16))) uint16_t As[4 * SPL], Bs[4 * SPL];
    const int m = blockIdx.y, tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const bool lower = (int)blockIdx.x < s;
    const int other = next[lower ? blockIdx.x : blockIdx.x - s];
    const int32_t id = lower ? tile_id[(int64_t)other * nt + k] : tile_id[(int64_t)k * nt + other];
    float *t = tile_ptr(tiles, mat_stride, m, id);
    const float *d = dinv + ((int64_t)m * nt + k) * (2 * TILE_FLOATS);
    if (lower) {
        stage_split(t, As, false, false, false, tid);               // L(i,k)
        stage_split(d + TILE_FLOATS, Bs, true, false, false, tid);  // inv(U_kk)
    } else {
        stage_split(d, As, false, false, false, tid);               // inv(L_kk)
        stage_split(t, Bs, true, false, false, tid);                // U(k,j)
    }
    __syncthreads();
    floatx16 cre = {}, cim = {};
    mma_split<true>(As, Bs, cre, cim, wave, lane);
    store_acc(t, cre, cim, wave, lane);
    __syncthreads();  // every wave is done reading the operands
    float *Tre = reinterpret_cast<float *>(As), *Tim = Tre + TILE * LDA;  // 33 KiB of the 74 KiB
    const int r0 = (wave >> 1) * 32 + 4 * (lane >> 5), c = (wave & 1) * 32 + (lane & 31);
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int row = r0 + 8 * (r >> 2) + (r & 3);
        Tre[row * LDA + c] = cre[r];
        Tim[row * LDA + c] = cim[r];
    }
    __syncthreads();
    uint16_t *dst = split + ((int64_t)m * split_stride + slot0 + blockIdx.x) * SPLIT_TILE;
    write_split(Tre, Tim, dst, lower, tid);
}

// One split tile = 2048 chunks of 16 B; thread tid moves chunks q * 256 + tid.
// The chunks travel in a register vector (not an array, which the compiler
// may place in scratch memory when it is carried across loop iterations).
typedef uint32_t u32x32 __attribute__((ext_vector_type(32)));

__device__ __forceinline__ u32x32 fetch_split(const uint16_t *__restrict__ src, int tid) {
    const uint4 *s4 = reinterpret_cast<const uint4 *>(src);
    u32x32 r;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const uint4 v = s4[q * 256 + tid];
        r[4 * q] = v.x;
        r[4 * q + 1] = v.y;
        r[4 * q + 2] = v.z;
        r[4 * q + 3] = v.w;
    }
    return r;
}
__device__ __forceinline__ void put_split(uint16_t *dst, const u32x32 r, int tid) {
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int e = q * 256 + tid;  // 8-element chunk of the split tile
        const int plane = (e >> 8) & 3, row = ((e >> 7) & 1) * 32 + (e & 31);
        const int k = (e >> 10) * 32 + ((e >> 6) & 1) * 16 + ((e >> 5) & 1) * 8;
        *reinterpret_cast<uint4 *>(dst + plane * SPL + row * SLD + k) =
            make_uint4(r[4 * q], r[4 * q + 1], r[4 * q + 2], r[4 * q + 3]);
    }
}

template <bool THREE>
__global__ void __launch_bounds__(256, 2) k_schur_group_split(
    float *__restrict__ tiles, int64_t mat_stride, int nt, int g0, const int32_t *__restrict__ pairs,
    const int32_t *__restrict__ pair_mask, int npairs, int nmat, const int32_t *__restrict__ tile_id,
    const uint16_t *__restrict__ split, int64_t split_stride, const int32_t *__restrict__ ppos,
    const int32_t *__restrict__ step_s, const int32_t *__restrict__ mats) {
    __shared__ __attribute__((aligned(16))) uint16_t As[4 * SPL], Bs[4 * SPL];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int64_t total = (int64_t)npairs * nmat, per = (total + 7) / 8;
    const int64_t item = (int64_t)(blockIdx.x % 8) * per + blockIdx.x / 8;
    if (item >= total) return;
    const int m = mats[item / npairs], q = (int)(item % npairs);
    const int i = pairs[2 * q], j = pairs[2 * q + 1];
    float *C = tile_ptr(tiles, mat_stride, m, tile_id[(int64_t)i * nt + j]);
    const uint16_t *sp = split + (int64_t)m * split_stride * SPLIT_TILE;
    // steps of the group with both L(i,t) and U(t,j) present (groups <= 32 steps)
    uint32_t mask = (uint32_t)pair_mask[q];
    floatx16 cre, cim;
    load_acc(C, cre, cim, wave, lane);
    if (mask) {
        int t = g0 + __builtin_ctz(mask);
        mask &= mask - 1;
        u32x32 ra = fetch_split(sp + (int64_t)ppos[(int64_t)t * nt + i] * SPLIT_TILE, tid);
        u32x32 rb = fetch_split(sp + (int64_t)(ppos[(int64_t)t * nt + j] + step_s[t]) * SPLIT_TILE, tid);
        for (;;) {
            __syncthreads();  // the previous step's products are done with LDS
            put_split(As, ra, tid);
            put_split(Bs, rb, tid);
            __syncthreads();
            const bool more = mask != 0;
            if (more) {
                t = g0 + __builtin_ctz(mask);
                mask &= mask - 1;
            }
            // the next step's tiles (the current step's again after the
            // last), their loads spread over this step's products
            ra = fetch_split(sp + (int64_t)ppos[(int64_t)t * nt + i] * SPLIT_TILE, tid);
            rb = fetch_split(sp + (int64_t)(ppos[(int64_t)t * nt + j] + step_s[t]) * SPLIT_TILE, tid);
            mma_split<THREE>(As, Bs, cre, cim, wave, lane);
#pragma unroll
            for (int g = 0; g < 16; ++g) {
                __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);              // 1 VMEM read
                __builtin_amdgcn_sched_group_barrier(0x008, THREE ? 3 : 1, 0);  // MFMA
            }
            if (!more) break;
        }
    }
    store_acc(C, cre, cim, wave, lane);
}

template __global__ void k_schur_group_split<true>(float *, int64_t, int, int, const int32_t *, const int32_t *,
                                                   int, int, const int32_t *, const uint16_t *, int64_t,
                                                   const int32_t *, const int32_t *, const int32_t *);
template __global__ void k_schur_group_split<false>(float *, int64_t, int, int, const int32_t *, const int32_t *,
                                                    int, int, const int32_t *, const uint16_t *, int64_t,
                                                    const int32_t *, const int32_t *, const int32_t *);

// Trailing update of a step group on 2 x 2 blocks of targets (BF16 precisions):
// quad q = (i0, i1, j0, j1) (an index -1 is absent); wave w updates the full
// 64 x 64 target (i_{w/2}, j_{w%2}) over the steps where both of its operands
// exist, from the split tiles L(i0,t), L(i1,t), U(t,j0), U(t,j1) staged once
// per step for all four waves -- half the operand traffic per target update
// of the one-target form.  One block per CU (147 KiB of LDS); the next step's
// four tiles are loaded into registers while the current products run.
__device__ __forceinline__ void mma_split_quad_chunk(const uint16_t *As, const uint16_t *Bs, floatx16 (&cre)[4],
                                                     floatx16 (&cim)[4], int lane, int k0, bool three) {
    const int kb = (lane >> 5) * 8, r = lane & 31;
    bf16x8 prh[2], pih[2], prl[2], pil[2], urh[2], uih[2], url[2], uil[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int ao = (h * 32 + r) * SLD + k0 + kb;
        prh[h] = lds8(As + ao);
        pih[h] = lds8(As + 2 * SPL + ao);
        urh[h] = lds8(Bs + ao);
        uih[h] = lds8(Bs + 2 * SPL + ao);
        if (three) {
            prl[h] = lds8(As + SPL + ao);
            pil[h] = lds8(As + 3 * SPL + ao);
            url[h] = lds8(Bs + SPL + ao);
            uil[h] = lds8(Bs + 3 * SPL + ao);
        }
    }
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        const bf16x8 npih = neg8(pih[a]);
        const bf16x8 npil = three ? neg8(pil[a]) : npih;
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            floatx16 &cr = cre[2 * a + b], &ci = cim[2 * a + b];
            cr = __builtin_amdgcn_mfma_f32_32x32x16_bf16(prh[a], urh[b], cr, 0, 0, 0);
            cr = __builtin_amdgcn_mfma_f32_32x32x16_bf16(npih, uih[b], cr, 0, 0, 0);
            ci = __builtin_amdgcn_mfma_f32_32x32x16_bf16(prh[a], uih[b], ci, 0, 0, 0);
            ci = __builtin_amdgcn_mfma_f32_32x32x16_bf16(pih[a], urh[b], ci, 0, 0, 0);
            if (three) {
                cr = __builtin_amdgcn_mfma_f32_32x32x16_bf16(prh[a], url[b], cr, 0, 0, 0);
                cr = __builtin_amdgcn_mfma_f32_32x32x16_bf16(prl[a], urh[b], cr, 0, 0, 0);
                cr = __builtin_amdgcn_mfma_f32_32x32x16_bf16(npih, uil[b], cr, 0, 0, 0);
                cr = __builtin_amdgcn_mfma_f32_32x32x16_bf16(npil, uih[b], cr, 0, 0, 0);
                ci = __builtin_amdgcn_mfma_f32_32x32x16_bf16(prh[a], uil[b], ci, 0, 0, 0);
                ci = __builtin_amdgcn_mfma_f32_32x32x16_bf16(prl[a], uih[b], ci, 0, 0, 0);
                ci = __builtin_amdgcn_mfma_f32_32x32x16_bf16(pih[a], url[b], ci, 0, 0, 0);
                ci = __builtin_amdgcn_mfma_f32_32x32x16_bf16(pil[a], urh[b], ci, 0, 0, 0);
            }
        }
    }
}

// Accumulators of a full 64 x 64 tile held by one wave: block 2a+b covers rows
// 32a.., columns 32b.. with the 32 x 32 MFMA layout.
__device__ __forceinline__ void load_acc_tile(const float *__restrict__ t, floatx16 (&cre)[4], floatx16 (&cim)[4],
                                              int lane) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int r0 = (q >> 1) * 32 + 4 * (lane >> 5), c = (q & 1) * 32 + (lane & 31);
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int row = r0 + 8 * (r >> 2) + (r & 3);
            cre[q][r] = t[row * TILE + c];
            cim[q][r] = t[PLANE + row * TILE + c];
        }
    }
}
__device__ __forceinline__ void store_acc_tile(float *__restrict__ t, const floatx16 (&cre)[4],
                                               const floatx16 (&cim)[4], int lane) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int r0 = (q >> 1) * 32 + 4 * (lane >> 5), c = (q & 1) * 32 + (lane & 31);
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int row = r0 + 8 * (r >> 2) + (r & 3);
            t[row * TILE + c] = cre[q][r];
            t[PLANE + row * TILE + c] = cim[q][r];
        }
    }
}

template <bool THREE>
__global__ void __launch_bounds__(256, 1) k_schur_quad_split(
    float *__restrict__ tiles, int64_t mat_stride, int nt, int g0, const int32_t *__restrict__ quads,
    const int32_t *__restrict__ quad_mask, int nquads, int nmat, const int32_t *__restrict__ tile_id,
    const uint16_t *__restrict__ split, int64_t split_stride, const int32_t *__restrict__ ppos,
    const int32_t *__restrict__ step_s, const int32_t *__restrict__ mats) {
    __shared__ __attribute__((aligned(16))) uint16_t A0[4 * SPL], A1[4 * SPL], B0[4 * SPL], B1[4 * SPL];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int64_t total = (int64_t)nquads * nmat, per = (total + 7) / 8;
    const int64_t item = (int64_t)(blockIdx.x % 8) * per + blockIdx.x / 8;
    if (item >= total) return;
    const int m = mats[item / nquads], q = (int)(item % nquads);
    const int I[2] = {quads[4 * q], quads[4 * q + 1]}, J[2] = {quads[4 * q + 2], quads[4 * q + 3]};
    const uint16_t *sp = split + (int64_t)m * split_stride * SPLIT_TILE;
    // step masks: rows (L present), columns (U present), this wave's target
    const int4 qm = reinterpret_cast<const int4 *>(quad_mask)[q];
    const uint32_t rm[2] = {(uint32_t)qm.x, (uint32_t)qm.y}, cm[2] = {(uint32_t)qm.z, (uint32_t)qm.w};
    const int ia = wave >> 1, jb = wave & 1;
    uint32_t mine = rm[ia] & cm[jb];
    // steps where some wave has work, and the tiles they need
    uint32_t all = (rm[0] & cm[0]) | (rm[0] & cm[1]) | (rm[1] & cm[0]) | (rm[1] & cm[1]);
    const uint32_t need_r[2] = {rm[0] & (cm[0] | cm[1]), rm[1] & (cm[0] | cm[1])};
    const uint32_t need_c[2] = {cm[0] & (rm[0] | rm[1]), cm[1] & (rm[0] | rm[1])};
    float *C = mine ? tile_ptr(tiles, mat_stride, m, tile_id[(int64_t)I[ia] * nt + J[jb]]) : nullptr;
    floatx16 cre[4], cim[4];
    if (mine) load_acc_tile(C, cre, cim, lane);
    if (all) {
        // Offsets of the four split tiles of step t; an absent tile points at
        // a present one, so that all four loads are issued without branches
        // (they are not staged).
        auto offsets = [&](int t, int64_t (&o)[4]) {
            const uint32_t bit = 1u << (t - g0);
            const int32_t *pp = ppos + (int64_t)t * nt;
            const int64_t s0 = (need_r[0] & bit) ? pp[I[0]] : -1, s1 = (need_r[1] & bit) ? pp[I[1]] : -1;
            const int64_t s2 = (need_c[0] & bit) ? pp[J[0]] + step_s[t] : -1;
            const int64_t s3 = (need_c[1] & bit) ? pp[J[1]] + step_s[t] : -1;
            const int64_t any = s0 >= 0 ? s0 : (s1 >= 0 ? s1 : (s2 >= 0 ? s2 : s3));
            o[0] = (s0 >= 0 ? s0 : any) * SPLIT_TILE;
            o[1] = (s1 >= 0 ? s1 : any) * SPLIT_TILE;
            o[2] = (s2 >= 0 ? s2 : any) * SPLIT_TILE;
            o[3] = (s3 >= 0 ? s3 : any) * SPLIT_TILE;
        };
        int t = g0 + __builtin_ctz(all);
        all &= all - 1;
        int64_t o[4];
        offsets(t, o);
        u32x32 ra0 = fetch_split(sp + o[0], tid), ra1 = fetch_split(sp + o[1], tid);
        u32x32 rb0 = fetch_split(sp + o[2], tid), rb1 = fetch_split(sp + o[3], tid);
        for (;;) {
            const uint32_t bit = 1u << (t - g0);
            __syncthreads();  // the previous step's products are done with LDS
            if (need_r[0] & bit) put_split(A0, ra0, tid);
            if (need_r[1] & bit) put_split(A1, ra1, tid);
            if (need_c[0] & bit) put_split(B0, rb0, tid);
            if (need_c[1] & bit) put_split(B1, rb1, tid);
            __syncthreads();
            const bool more = all != 0;
            const int tcur = t;
            // the next step's tiles (the current step's again after the last)
            if (more) {
                t = g0 + __builtin_ctz(all);
                all &= all - 1;
            }
            offsets(t, o);
            if (mine & (1u << (tcur - g0))) {
                // the next step's loads spread over this step's products: one
                // load per few MFMAs, so the products never wait behind a
                // full memory queue
                ra0 = fetch_split(sp + o[0], tid);
                ra1 = fetch_split(sp + o[1], tid);
                rb0 = fetch_split(sp + o[2], tid);
                rb1 = fetch_split(sp + o[3], tid);
                const uint16_t *As = ia ? A1 : A0, *Bs = jb ? B1 : B0;
#pragma unroll
                for (int k0 = 0; k0 < TILE; k0 += 16) mma_split_quad_chunk(As, Bs, cre, cim, lane, k0, THREE);
#pragma unroll
                for (int g = 0; g < 32; ++g) {
                    __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);          // 1 VMEM read
                    __builtin_amdgcn_sched_group_barrier(0x008, THREE ? 6 : 2, 0);  // MFMA
                }
            } else {
                ra0 = fetch_split(sp + o[0], tid);
                ra1 = fetch_split(sp + o[1], tid);
                rb0 = fetch_split(sp + o[2], tid);
                rb1 = fetch_split(sp + o[3], tid);
            }
            if (!more) break;
        }
    }
    if (mine) store_acc_tile(C, cre, cim, lane);
}

template __global__ void k_schur_quad_split<true>(float *, int64_t, int, int, const int32_t *, const int32_t *, int, int,
                                                 const int32_t *, const uint16_t *, int64_t, const int32_t *, const int32_t *, const int32_t *);
template __global__ void k_schur_quad_split<false>(float *, int64_t, int, int, const int32_t *, const int32_t *, int, int,
                                                  const int32_t *, const uint16_t *, int64_t, const int32_t *, const int32_t *, const int32_t *);


// Trailing update of a step group on 2 x 2 blocks of targets, pipelined by K
// quarters (BF16 precisions): the same quads, step masks and products as
// k_schur_quad_split, each accumulator in the same order, so the results
// agree bit for bit.  Each step's four split tiles pass through LDS in four
// quarters of 16 K columns, double-buffered (2 x 32 KiB), which leaves room
// for two blocks per CU: while one block waits at its barrier or on its
// target tiles, the other one's products keep the matrix cores busy.  The
// quarters move from memory straight into LDS (buffer loads with the LDS
// destination: no registers and no LDS write instructions), one quarter
// ahead.  Split tiles are stored as operand blocks in lane order, so every
// block is one 1 KiB load and one conflict-free 16-byte-per-lane LDS read.
constexpr int QQ_BUF = 4 * 8 * SPLIT_BLOCK;  // four tiles x eight operand blocks (32 KiB)

// wave w moves the 8 operand blocks (plane, row half) of K quarter q of tile w
// into buffer `buf`
__device__ __forceinline__ void qq_issue(uint16_t *buf, const uint16_t *__restrict__ sp, int4 slot, int q, int wave,
                                         int lane) {
    const int s = __builtin_amdgcn_readfirstlane(wave == 0 ? slot.x : wave == 1 ? slot.y : wave == 2 ? slot.z : slot.w);
    // buffer (not flat) loads: the compiler then tracks them as memory loads
    // into LDS only, and the LDS reads of the other buffer wait for none of them
    const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<uint16_t *>(sp + (int64_t)s * SPLIT_TILE), 0, SPLIT_TILE * 2, 0x00020000);
    uint16_t *d = buf + wave * 8 * SPLIT_BLOCK;
    const int first = (q >> 1) * 16 + (q & 1);  // block of (plane 0, row half 0) in quarter q
#pragma unroll
    for (int j = 0; j < 8; ++j)  // j = 2 * plane + row half
        __builtin_amdgcn_raw_ptr_buffer_load_lds(rsrc, reinterpret_cast<void *>(d + j * SPLIT_BLOCK), 16,
                                                 ((first + 2 * j) * SPLIT_BLOCK + lane * 8) * 2, 0, 0, 0);
}

// products of one quarter buffer: A from tile ua, B from tile ub.  `keep` is
// all ones, or zero when this wave has no update in the step: its A operands
// are then zero and the products add nothing (instead of a branch, which would
// keep the loads and the products of the other waves apart).  Each of the 8
// accumulators takes its six products in the order of mma_split_quad_chunk;
// the accumulators take turns, so consecutive MFMAs never depend on each other.
__device__ __forceinline__ bf16x8 lds8_and(const uint16_t *p, uint32_t keep) {
    return __builtin_bit_cast(bf16x8, __builtin_bit_cast(u32x4, lds8(p)) & (u32x4)(keep));
}
// The 16 operand blocks of one quarter buffer, read into registers before
// any product (A from tile ua, masked by `keep`; B from tile ub).
struct QuarterOps {
    bf16x8 prh[2], pih[2], prl[2], pil[2], urh[2], uih[2], url[2], uil[2];
};
__device__ __forceinline__ QuarterOps qq_read(const uint16_t *buf, int ua, int ub, int lane, bool three,
                                              uint32_t keep) {
    auto blk = [&](int u, int pl, int a) { return buf + (u * 8 + pl * 2 + a) * SPLIT_BLOCK + lane * 8; };
    QuarterOps o;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        o.prh[h] = lds8_and(blk(ua, 0, h), keep);
        o.pih[h] = lds8_and(blk(ua, 2, h), keep);
        o.urh[h] = lds8(blk(ub, 0, h));
        o.uih[h] = lds8(blk(ub, 2, h));
        if (three) {
            o.prl[h] = lds8_and(blk(ua, 1, h), keep);
            o.pil[h] = lds8_and(blk(ua, 3, h), keep);
            o.url[h] = lds8(blk(ub, 1, h));
            o.uil[h] = lds8(blk(ub, 3, h));
        }
    }
    return o;
}
// The products of one quarter.  Each of the 8 accumulators takes its six
// products in the order of mma_split_quad_chunk; the accumulators take turns,
// so consecutive MFMAs never depend on each other.
__device__ __forceinline__ void qq_products(const QuarterOps &o, floatx16 (&cre)[4], floatx16 (&cim)[4],
                                            bool three) {
    bf16x8 npih[2], npil[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        npih[h] = neg8(o.pih[h]);
        npil[h] = three ? neg8(o.pil[h]) : npih[h];
    }
    const int np = three ? 6 : 2;
#pragma unroll
    for (int p = 0; p < np; ++p)
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int a = t >> 1, b = t & 1;
            // re += Pr Ur - Pi Ui (hi hi, then hi lo, lo hi), im += Pr Ui + Pi Ur
            const bf16x8 xr[6] = {o.prh[a], npih[a], o.prh[a], o.prl[a], npih[a], npil[a]};
            const bf16x8 yr[6] = {o.urh[b], o.uih[b], o.url[b], o.urh[b], o.uil[b], o.uih[b]};
            const bf16x8 xi[6] = {o.prh[a], o.pih[a], o.prh[a], o.prl[a], o.pih[a], o.pil[a]};
            const bf16x8 yi[6] = {o.uih[b], o.urh[b], o.uil[b], o.uih[b], o.url[b], o.urh[b]};
            cre[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(xr[p], yr[p], cre[t], 0, 0, 0);
            cim[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(xi[p], yi[p], cim[t], 0, 0, 0);
        }
}

// s_waitcnt immediates (gfx9 encoding: vmcnt[3:0] | expcnt << 4 | lgkmcnt << 8
// | vmcnt[5:4] << 14) that wait for one counter only
constexpr int WAIT_VM8 = 0x0F78, WAIT_VM0 = 0x0F70, WAIT_LGKM0 = 0xC07F;

// One quad (matrix m, quad q) of the trailing update, K quarters double-
// buffered through lds0/lds1 (see k_schur_quad_pp).  Called by every wave of
// the workgroup; leaves both buffers idle and slots free on return.
template <bool THREE>
__device__ __forceinline__ void quad_pp_body(int m, int q, float *__restrict__ tiles, int64_t mat_stride, int nt, int g0,
                                             const int32_t *__restrict__ quads, const int32_t *__restrict__ quad_mask,
                                             const int32_t *__restrict__ tile_id, const uint16_t *__restrict__ split,
                                             int64_t split_stride, const int32_t *__restrict__ ppos,
                                             const int32_t *__restrict__ step_s, uint16_t *lds0, uint16_t *lds1,
                                             int4 *slots, int tid, int wave, int lane) {
    const int I[2] = {quads[4 * q], quads[4 * q + 1]}, J[2] = {quads[4 * q + 2], quads[4 * q + 3]};
    const uint16_t *sp = split + (int64_t)m * split_stride * SPLIT_TILE;
    const int4 qm = reinterpret_cast<const int4 *>(quad_mask)[q];
    const uint32_t rm[2] = {(uint32_t)qm.x, (uint32_t)qm.y}, cm[2] = {(uint32_t)qm.z, (uint32_t)qm.w};
    const int ia = wave >> 1, jb = wave & 1;
    const uint32_t mine = rm[ia] & cm[jb];
    const uint32_t all = (rm[0] & cm[0]) | (rm[0] & cm[1]) | (rm[1] & cm[0]) | (rm[1] & cm[1]);
    const int n = __builtin_popcount(all);
    // listed steps: the set bits of `all`, ascending; this wave's work as a
    // mask over the list
    uint32_t work = 0;
    {
        uint32_t rest = all;
        for (int i = 0; rest; ++i) {
            const int b = __builtin_ctz(rest);
            rest &= rest - 1;
            if (mine & (1u << b)) work |= 1u << i;
            if (tid == i) {
                // an absent tile takes a present one's slot: its half is
                // loaded but read by no wave
                const uint32_t bit = 1u << b;
                const int32_t *pp = ppos + (int64_t)(g0 + b) * nt;
                const int ss = step_s[g0 + b];
                const bool any_c = (cm[0] | cm[1]) & bit, any_r = (rm[0] | rm[1]) & bit;
                int s0 = (rm[0] & bit) && any_c ? pp[I[0]] : -1, s1 = (rm[1] & bit) && any_c ? pp[I[1]] : -1;
                int s2 = (cm[0] & bit) && any_r ? pp[J[0]] + ss : -1, s3 = (cm[1] & bit) && any_r ? pp[J[1]] + ss : -1;
                const int any = s0 >= 0 ? s0 : (s1 >= 0 ? s1 : (s2 >= 0 ? s2 : s3));
                slots[i] = make_int4(s0 >= 0 ? s0 : any, s1 >= 0 ? s1 : any, s2 >= 0 ? s2 : any, s3 >= 0 ? s3 : any);
            }
        }
    }
    float *C = mine ? tile_ptr(tiles, mat_stride, m, tile_id[(int64_t)I[ia] * nt + J[jb]]) : nullptr;
    floatx16 cre[4], cim[4];
    if (mine) {
        load_acc_tile(C, cre, cim, lane);
    } else {
#pragma unroll
        for (int b = 0; b < 4; ++b) cre[b] = cim[b] = floatx16{};
    }
    __syncthreads();
    if (n) {
        const int ua = ia, ub = 2 + jb;
        const int nq = 4 * n;  // K quarters of the listed steps, flattened
        // quarter g (step g / 4, K quarter g % 4) goes to buffer g % 2; past
        // the end the last quarter is loaded again, into a buffer no later
        // quarter reads
        auto issue = [&](uint16_t *buf, int g) {
            const int gg = g < nq ? g : nq - 1;
            qq_issue(buf, sp, slots[gg >> 2], gg & 3, wave, lane);
        };
        issue(lds0, 0);
        issue(lds1, 1);
        // One quarter: wait for its loads (the next quarter's stay in flight),
        // barrier (its four tiles are in LDS for every wave), read this wave's
        // operands into registers, barrier (every wave has read: the buffer is
        // free), refill the buffer two quarters ahead, multiply.  The loads
        // thus have two quarters of products to land.  The barriers are the
        // bare instruction: a __syncthreads() fence would drain the loads in
        // flight.
        auto quarter = [&](uint16_t *buf, int g) {
            __builtin_amdgcn_s_waitcnt(WAIT_VM8);
            __builtin_amdgcn_s_barrier();
            const uint32_t keep = (work >> (g >> 2)) & 1u ? 0xffffffffu : 0u;
            const QuarterOps ops = qq_read(buf, ua, ub, lane, THREE, keep);
            __builtin_amdgcn_s_waitcnt(WAIT_LGKM0);
            __builtin_amdgcn_s_barrier();
            issue(buf, g + 2);
            __builtin_amdgcn_sched_barrier(0);  // the loads go out before the products
            qq_products(ops, cre, cim, THREE);
        };
        for (int g = 0; g < nq; g += 2) {
            quarter(lds0, g);
            quarter(lds1, g + 1);
        }
        // the loads issued past the end land before LDS is reused
        __builtin_amdgcn_s_waitcnt(WAIT_VM0);
        __syncthreads();
    }
    if (mine) store_acc_tile(C, cre, cim, lane);
}

template <bool THREE>
__global__ void __launch_bounds__(256, 2) k_schur_quad_pp(
    float *__restrict__ tiles, int64_t mat_stride, int nt, int g0, const int32_t *__restrict__ quads,
    const int32_t *__restrict__ quad_mask, int nquads, int nmat, const int32_t *__restrict__ tile_id,
    const uint16_t *__restrict__ split, int64_t split_stride, const int32_t *__restrict__ ppos,
    const int32_t *__restrict__ step_s, const int32_t *__restrict__ mats) {
    // two buffers as two LDS objects: the compiler then knows that the reads
    // of one do not wait for the direct loads into the other
    __shared__ __attribute__((aligned(1024))) uint16_t lds0[QQ_BUF];
    __shared__ __attribute__((aligned(1024))) uint16_t lds1[QQ_BUF];
    __shared__ int4 slots[32];  // split slots of the four tiles of the i-th listed step
    const int tid = threadIdx.x, wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
    const int64_t total = (int64_t)nquads * nmat, per = (total + 7) / 8;
    const int64_t item = (int64_t)(blockIdx.x % 8) * per + blockIdx.x / 8;
    if (item >= total) return;
    quad_pp_body<THREE>(mats[item / nquads], (int)(item % nquads), tiles, mat_stride, nt, g0, quads, quad_mask,
                        tile_id, split, split_stride, ppos, step_s, lds0, lds1, slots, tid, wave, lane);
}

template __global__ void k_schur_quad_pp<true>(float *, int64_t, int, int, const int32_t *, const int32_t *, int, int,
                                              const int32_t *, const uint16_t *, int64_t, const int32_t *, const int32_t *, const int32_t *);
template __global__ void k_schur_quad_pp<false>(float *, int64_t, int, int, const int32_t *, const int32_t *, int, int,
                                               const int32_t *, const uint16_t *, int64_t, const int32_t *, const int32_t *, const int32_t *);

}  // namespace aosx
