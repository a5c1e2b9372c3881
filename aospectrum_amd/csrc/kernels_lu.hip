// Tile LU with static pivoting for batches of shifted matrices A_m = H - z_m S.
//
// Storage: every matrix of a batch owns n_tiles tiles of TILE x TILE complex
// entries, each stored as a row-major real plane followed by a row-major
// imaginary plane (planar layout feeds the FP32 MFMA operands directly).  The
// tile structure comes from the host symbolic analysis; for step k the tiles
// next[k] hold L(i,k) and, by symmetry of the pattern, U(k,i).
//
// Per step k (right-looking):
//   getrf  : factor the diagonal tile without row exchanges; a pivot below
//            1e-6 * max|tile| is replaced by that bound (static pivoting in the
//            sense of Li & Demmel's GESP, corrected later by refinement);
//            also writes inv(L_kk) (unit lower) and inv(U_kk).
//   panel  : L(i,k) <- L(i,k) * inv(U_kk),  U(k,i) <- inv(L_kk) * U(k,i)   (MFMA)
//   schur  : A(i,j) -= L(i,k) * U(k,j) for i, j in next[k]                  (MFMA)
// The complex products use four real FP32 MFMA products (v_mfma_f32_32x32x2_f32),
// i.e. 8 real flops per complex multiply-add.  The BF16 variants split every
// FP32 operand into hi + lo BF16 parts (v_mfma_f32_32x32x16_bf16) and keep FP32
// accumulation.
#include "device_common.h"

namespace aosx {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

__global__ void k_fill_zero(float4 *__restrict__ p, int64_t count4) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count4; i += stride)
        p[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// A_m = H_k - z_m S_k scattered into the tile storage of matrix m
// (matrix m uses k point kidx[m] and shift z[m]); padded rows get a unit
// diagonal so the trailing tile stays nonsingular.
__global__ void k_scatter_shifted(float *__restrict__ tiles, int64_t mat_stride, const float2 *__restrict__ h,
                                  const float2 *__restrict__ s, int64_t nnz, const int32_t *__restrict__ kidx,
                                  const float2 *__restrict__ z, const int32_t *__restrict__ rows,
                                  const int32_t *__restrict__ cols, const int32_t *__restrict__ iperm,
                                  const int32_t *__restrict__ tile_id, int nt, int64_t n) {
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
        const int64_t padded = (int64_t)nt * TILE;
        for (int64_t r = n + threadIdx.x; r < padded; r += blockDim.x) {
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
    __shared__ float piv[2];
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
    int local_replaced = 0;
    for (int kk = 0; kk < TILE; ++kk) {
        if (tid == 0) {
            float pr = Are[kk * LDA + kk], pi = Aim[kk * LDA + kk];
            const float mag = sqrtf(pr * pr + pi * pi);
            if (!(mag >= tau)) {
                if (mag > 0.f) { pr = pr / mag * tau; pi = pi / mag * tau; }
                else { pr = tau; pi = 0.f; }
                ++local_replaced;
                Are[kk * LDA + kk] = pr;
                Aim[kk * LDA + kk] = pi;
            }
            const float d = pr * pr + pi * pi;
            piv[0] = pr / d;
            piv[1] = -pi / d;
        }
        __syncthreads();
        const float ir = piv[0], ii = piv[1];
        for (int i = kk + 1 + tid; i < TILE; i += 256) {
            const float ar = Are[i * LDA + kk], ai = Aim[i * LDA + kk];
            Are[i * LDA + kk] = ar * ir - ai * ii;
            Aim[i * LDA + kk] = ar * ii + ai * ir;
        }
        __syncthreads();
        const int w = TILE - 1 - kk;
        for (int e = tid; e < w * w; e += 256) {
            const int i = kk + 1 + e / w, j = kk + 1 + e % w;
            const float lr = Are[i * LDA + kk], li = Aim[i * LDA + kk];
            const float ur = Are[kk * LDA + j], ui = Aim[kk * LDA + j];
            Are[i * LDA + j] -= lr * ur - li * ui;
            Aim[i * LDA + j] -= lr * ui + li * ur;
        }
        __syncthreads();
    }
    if (tid == 0 && local_replaced) atomicAdd(replaced + m, local_replaced);
    for (int e = tid; e < TILE2; e += 256) {
        const int r = e / TILE, c = e % TILE;
        t[e] = Are[r * LDA + c];
        t[PLANE + e] = Aim[r * LDA + c];
    }
    // inv(L): thread c < 64 owns column c; inv(U): thread 64 + r owns row r
    if (tid < TILE) {
        const int c = tid;
        for (int i = 0; i < TILE; ++i) {
            if (i < c) { Lre[i * LDA + c] = 0.f; Lim[i * LDA + c] = 0.f; continue; }
            if (i == c) { Lre[i * LDA + c] = 1.f; Lim[i * LDA + c] = 0.f; continue; }
            float sr = 0.f, si = 0.f;
            for (int j = c; j < i; ++j) {
                const float ar = Are[i * LDA + j], ai = Aim[i * LDA + j];
                const float xr = Lre[j * LDA + c], xi = Lim[j * LDA + c];
                sr += ar * xr - ai * xi;
                si += ar * xi + ai * xr;
            }
            Lre[i * LDA + c] = -sr;
            Lim[i * LDA + c] = -si;
        }
    } else if (tid < 2 * TILE) {
        const int r = tid - TILE;
        for (int j = 0; j < TILE; ++j) {
            if (j < r) { Ure[r * LDA + j] = 0.f; Uim[r * LDA + j] = 0.f; continue; }
            float sr = (j == r) ? 1.f : 0.f, si = 0.f;
            for (int i = r; i < j; ++i) {
                const float yr = Ure[r * LDA + i], yi = Uim[r * LDA + i];
                const float ur = Are[i * LDA + j], ui = Aim[i * LDA + j];
                sr -= yr * ur - yi * ui;
                si -= yr * ui + yi * ur;
            }
            const float dr = Are[j * LDA + j], di = Aim[j * LDA + j];
            const float dd = dr * dr + di * di;
            // y_j = s / d
            Ure[r * LDA + j] = (sr * dr + si * di) / dd;
            Uim[r * LDA + j] = (si * dr - sr * di) / dd;
        }
    }
    __syncthreads();
    float *d = dinv + ((int64_t)m * nt + k) * (2 * TILE_FLOATS);
    for (int e = tid; e < TILE2; e += 256) {
        const int r = e / TILE, c = e % TILE;
        d[e] = Lre[r * LDA + c];
        d[PLANE + e] = Lim[r * LDA + c];
        d[2 * PLANE + e] = Ure[r * LDA + c];
        d[3 * PLANE + e] = Uim[r * LDA + c];
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

__device__ __forceinline__ __bf16 to_bf16(float x) { return (__bf16)x; }

// acc -= A * B with every operand split as x = hi + lo (both BF16):
// A*B ~ Ahi*Bhi + Ahi*Blo + Alo*Bhi (three BF16 MFMA products, FP32 accumulate);
// THREE = false keeps only Ahi*Bhi.
template <bool THREE>
__device__ __forceinline__ void mma_bf16_sub(const float *Are, const float *Aim, const float *Bre,
                                             const float *Bim, floatx16 &cre, floatx16 &cim, int wave,
                                             int lane) {
    const int i = (wave >> 1) * 32 + (lane & 31);
    const int j = (wave & 1) * 32 + (lane & 31);
    const int kb = (lane >> 5) * 8;
#pragma unroll 1
    for (int k0 = 0; k0 < TILE; k0 += 16) {
        bf16x8 arh, arl, aih, ail, nrh, nrl, brh, brl, bih, bil, nih, nil;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int k = k0 + kb + e;
            const float ar = Are[i * LDA + k], ai = Aim[i * LDA + k];
            const float br = Bre[k * TILE + j], bi = Bim[k * TILE + j];
            const __bf16 arH = to_bf16(ar), aiH = to_bf16(ai), brH = to_bf16(br), biH = to_bf16(bi);
            arh[e] = arH; aih[e] = aiH; brh[e] = brH; bih[e] = biH;
            nrh[e] = to_bf16(-ar);
            nih[e] = to_bf16(-ai);
            if (THREE) {
                const __bf16 arL = to_bf16(ar - (float)arH), aiL = to_bf16(ai - (float)aiH);
                arl[e] = arL; ail[e] = aiL;
                nrl[e] = to_bf16(-(ar - (float)arH));
                nil[e] = to_bf16(-(ai - (float)aiH));
                brl[e] = to_bf16(br - (float)brH);
                bil[e] = to_bf16(bi - (float)biH);
            }
        }
        // re -= ar*br - ai*bi  ->  (-ar)*br + ai*bi
        // im -= ar*bi + ai*br  ->  (-ar)*bi + (-ai)*br
        cre = __builtin_amdgcn_mfma_f32_32x32x16_bf16(nrh, brh, cre, 0, 0, 0);
        cre = __builtin_amdgcn_mfma_f32_32x32x16_bf16(aih, bih, cre, 0, 0, 0);
        cim = __builtin_amdgcn_mfma_f32_32x32x16_bf16(nrh, bih, cim, 0, 0, 0);
        cim = __builtin_amdgcn_mfma_f32_32x32x16_bf16(nih, brh, cim, 0, 0, 0);
        if (THREE) {
            cre = __builtin_amdgcn_mfma_f32_32x32x16_bf16(nrh, brl, cre, 0, 0, 0);
            cre = __builtin_amdgcn_mfma_f32_32x32x16_bf16(nrl, brh, cre, 0, 0, 0);
            cre = __builtin_amdgcn_mfma_f32_32x32x16_bf16(aih, bil, cre, 0, 0, 0);
            cre = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ail, bih, cre, 0, 0, 0);
            cim = __builtin_amdgcn_mfma_f32_32x32x16_bf16(nrh, bil, cim, 0, 0, 0);
            cim = __builtin_amdgcn_mfma_f32_32x32x16_bf16(nrl, bih, cim, 0, 0, 0);
            cim = __builtin_amdgcn_mfma_f32_32x32x16_bf16(nih, brl, cim, 0, 0, 0);
            cim = __builtin_amdgcn_mfma_f32_32x32x16_bf16(nil, brh, cim, 0, 0, 0);
        }
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

// Panel step k: block x < s updates L(next[x], k) <- L * inv(U_kk); block
// x >= s updates U(k, next[x - s]) <- inv(L_kk) * U.
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

// Schur step k: A(i,j) -= L(i,k) U(k,j) for the s*s pairs of next[k].
// MODE 0: FP32 MFMA; 1: BF16 x3 split; 2: BF16 x1.
template <int MODE>
__global__ void __launch_bounds__(256) k_schur(float *__restrict__ tiles, int64_t mat_stride, int nt, int k,
                                              const int32_t *__restrict__ next, int s,
                                              const int32_t *__restrict__ tile_id) {
    __shared__ float Are[TILE * LDA], Aim[TILE * LDA];
    __shared__ __attribute__((aligned(16))) float Bre[TILE2], Bim[TILE2];
    const int m = blockIdx.y, tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int a = blockIdx.x / s, b = blockIdx.x % s;
    const int i = next[a], j = next[b];
    const float *L = tile_ptr(tiles, mat_stride, m, tile_id[(int64_t)i * nt + k]);
    const float *U = tile_ptr(tiles, mat_stride, m, tile_id[(int64_t)k * nt + j]);
    float *C = tile_ptr(tiles, mat_stride, m, tile_id[(int64_t)i * nt + j]);
    stage_a(L, Are, Aim, tid);
    stage_b(U, Bre, Bim, tid);
    floatx16 cre, cim;
    load_acc(C, cre, cim, wave, lane);
    __syncthreads();
    if (MODE == 0) mma_fp32<true>(Are, Aim, Bre, Bim, cre, cim, wave, lane);
    else mma_bf16_sub<MODE == 1>(Are, Aim, Bre, Bim, cre, cim, wave, lane);
    store_acc(C, cre, cim, wave, lane);
}

template __global__ void k_schur<0>(float *, int64_t, int, int, const int32_t *, int, const int32_t *);
template __global__ void k_schur<1>(float *, int64_t, int, int, const int32_t *, int, const int32_t *);
template __global__ void k_schur<2>(float *, int64_t, int, int, const int32_t *, int, const int32_t *);

}  // namespace aosx
