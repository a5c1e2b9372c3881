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
// stored factor stays FP32.  The trailing update forms each complex product
// from three real ones (re*re, im*im, (re+im)*(re+im)) instead of four.
#include "device_common.h"

namespace aosx {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

__global__ void k_fill_zero(float4 *__restrict__ p, int64_t count4) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count4; i += stride)
        p[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// A_m = H_k - z_m S_k (matrix m uses k point kidx[m] and shift z[m]) written
// tile by tile: workgroup b holds stored tile b in LDS -- zeros, the tile's
// CSR entries (source p, position pos = 64 r + c in the tile) and the unit
// diagonals of padding rows (source -1), so their tiles stay nonsingular --
// and writes it out whole for every matrix of the batch.  The factor storage
// is thus written once, coalesced, with no separate zero fill; the H and S
// values of an entry are loaded once per k point (the matrices of one k point
// are consecutive).
__global__ void __launch_bounds__(256) k_fill_tiles(float *__restrict__ tiles, int64_t mat_stride,
                                                   const float2 *__restrict__ h, const float2 *__restrict__ s,
                                                   int64_t nnz, const int32_t *__restrict__ kidx,
                                                   const float2 *__restrict__ z, int nmat,
                                                   const int32_t *__restrict__ tptr, const int32_t *__restrict__ tsrc,
                                                   const uint16_t *__restrict__ tpos) {
    __shared__ float4 T[TILE_FLOATS / 4];  // re plane, im plane (32 KiB)
    constexpr int EPT = TILE2 / 256;       // entries per thread (a tile holds at most 64 x 64)
    float *Tf = reinterpret_cast<float *>(T);
    const int tid = threadIdx.x;
    const int64_t b = blockIdx.x;
    for (int i = tid; i < TILE_FLOATS / 4; i += 256) T[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    const int e0 = tptr[b], ne = tptr[b + 1] - e0;
    int32_t src[EPT];
    int pos[EPT];
#pragma unroll
    for (int j = 0; j < EPT; ++j) {
        const int e = tid + 256 * j;
        src[j] = e < ne ? tsrc[e0 + e] : -2;
        pos[j] = e < ne ? (int)tpos[e0 + e] : 0;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < EPT; ++j)
        if (src[j] == -1) Tf[pos[j]] = 1.0f;
    float2 hv[EPT], sv[EPT];
    int kprev = -1;
    for (int m = 0; m < nmat; ++m) {
        const int k = kidx[m];
        const float2 zm = z[m];
        if (k != kprev) {
            const float2 *hk = h + (int64_t)k * nnz, *sk = s + (int64_t)k * nnz;
#pragma unroll
            for (int j = 0; j < EPT; ++j) {
                hv[j] = src[j] >= 0 ? hk[src[j]] : make_float2(0.f, 0.f);
                sv[j] = src[j] >= 0 ? sk[src[j]] : make_float2(0.f, 0.f);
            }
            kprev = k;
        }
#pragma unroll
        for (int j = 0; j < EPT; ++j)
            if (src[j] >= 0) {
                Tf[pos[j]] = hv[j].x - (zm.x * sv[j].x - zm.y * sv[j].y);
                Tf[PLANE + pos[j]] = hv[j].y - (zm.x * sv[j].y + zm.y * sv[j].x);
            }
        __syncthreads();
        float4 *dst = reinterpret_cast<float4 *>(tiles + (int64_t)m * mat_stride + b * TILE_FLOATS);
        for (int i = tid; i < TILE_FLOATS / 4; i += 256) dst[i] = T[i];
        __syncthreads();  // the tile is read out before the next matrix's values go in
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
#pragma unroll
    for (int q = 0; q < TILE2 / 256; ++q) {  // unrolled: all loads in flight at once
        const int e = q * 256 + tid;
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
    // columns kk+1+tj+16b.  A step issues all of its LDS reads (pivot, its
    // column and row entries, its targets) before it uses any and computes
    // every new value before it stores any, so that it waits for LDS once:
    // indices past the tile are clamped to the last row/column (in bounds,
    // the values unused).
    const int ti = tid >> 4, tj = tid & 15;
    int local_replaced = 0;
    for (int kk = 0; kk < TILE; ++kk) {
        int ic[4], jc[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            ic[a] = min(kk + 1 + ti + 16 * a, TILE - 1);
            jc[a] = min(kk + 1 + tj + 16 * a, TILE - 1);
        }
        float pr = Are[kk * LDA + kk], pi = Aim[kk * LDA + kk];
        float cr[4], ci[4], ur[4], ui[4], xr[4][4], xi[4][4];
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            cr[a] = Are[ic[a] * LDA + kk];
            ci[a] = Aim[ic[a] * LDA + kk];
            ur[a] = Are[kk * LDA + jc[a]];
            ui[a] = Aim[kk * LDA + jc[a]];
        }
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                xr[a][b] = Are[ic[a] * LDA + jc[b]];
                xi[a][b] = Aim[ic[a] * LDA + jc[b]];
            }
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
            const bool in = kk + 1 + ti + 16 * a < TILE;
            lr[a] = in ? cr[a] * ir - ci[a] * ii : 0.f;
            li[a] = in ? cr[a] * ii + ci[a] * ir : 0.f;
        }
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                xr[a][b] -= lr[a] * ur[b] - li[a] * ui[b];
                xi[a][b] -= lr[a] * ui[b] + li[a] * ur[b];
            }
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                const int i = kk + 1 + ti + 16 * a, j = kk + 1 + tj + 16 * b;
                if (i < TILE && j < TILE) {
                    Are[i * LDA + j] = xr[a][b];
                    Aim[i * LDA + j] = xi[a][b];
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
#pragma unroll
    for (int q = 0; q < TILE2 / 256; ++q) {  // unrolled: all loads in flight at once
        const int e = q * 256 + tid;
        const int r = e / TILE, c = e % TILE;
        t[e] = Are[r * LDA + c];
        t[PLANE + e] = Aim[r * LDA + c];
    }
    // inv(L) (unit lower) by columns and inv(U) by rows; four threads share a
    // column (row) and split each inner product, combined across the four
    // adjacent lanes:
    //   X(i,c) = -sum_{j=c}^{i-1} L(i,j) X(j,c),           X(c,c) = 1
    //   Y(r,j) = (delta_rj - sum_{i=r}^{j-1} Y(r,i) U(i,j)) / U(j,j)
    // Column c of X and row r of Y depend on nothing but themselves and the
    // factored tile, and the four lanes of a column (row) sit in one wave, so
    // the recurrences need no workgroup barrier: LDS operations of one wave
    // complete in issue order.  Iteration i takes row i of X and column i of Y.
    // Both recurrences of iteration i run over the same indices j = col +
    // part + 4t < i (row = col, jj = i); four t are loaded before any is
    // used, with indices clamped to the tile (the extra values unused).
    const int col = tid >> 2, part = tid & 3, row = col;
    for (int i = 0; i < TILE; ++i) {
        float sr = 0.f, si = 0.f, tr = 0.f, ti2 = 0.f;
        const int jj = i;  // column of Y
        for (int t0 = col + part; t0 < i; t0 += 16) {
            float ar[4], ai[4], xr[4], xi[4], yr[4], yi[4], ur[4], ui[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int j = min(t0 + 4 * u, TILE - 1);
                ar[u] = Are[i * LDA + j];
                ai[u] = Aim[i * LDA + j];
                xr[u] = Lre[j * LDA + col];
                xi[u] = Lim[j * LDA + col];
                yr[u] = Ure[row * LDA + j];
                yi[u] = Uim[row * LDA + j];
                ur[u] = Are[j * LDA + jj];
                ui[u] = Aim[j * LDA + jj];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (t0 + 4 * u < i) {
                    sr += ar[u] * xr[u] - ai[u] * xi[u];
                    si += ar[u] * xi[u] + ai[u] * xr[u];
                    tr += yr[u] * ur[u] - yi[u] * ui[u];
                    ti2 += yr[u] * ui[u] + yi[u] * ur[u];
                }
        }
        sr += __shfl_xor(sr, 1);
        si += __shfl_xor(si, 1);
        tr += __shfl_xor(tr, 1);
        ti2 += __shfl_xor(ti2, 1);
        sr += __shfl_xor(sr, 2);
        si += __shfl_xor(si, 2);
        tr += __shfl_xor(tr, 2);
        ti2 += __shfl_xor(ti2, 2);
        if (part == 0) {
            Lre[i * LDA + col] = i < col ? 0.f : (i == col ? 1.f : -sr);
            Lim[i * LDA + col] = i <= col ? 0.f : -si;
            if (jj < row) {
                Ure[row * LDA + jj] = 0.f;
                Uim[row * LDA + jj] = 0.f;
            } else {
                const float nr = (jj == row ? 1.f : 0.f) - tr, ni = -ti2;
                const float dr = Are[jj * LDA + jj], di = Aim[jj * LDA + jj];
                const float dd = dr * dr + di * di;
                Ure[row * LDA + jj] = (nr * dr + ni * di) / dd;
                Uim[row * LDA + jj] = (ni * dr - nr * di) / dd;
            }
        }
        // this iteration's writes land before the wave's reads of the next
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    __syncthreads();  // X and Y are complete before the copy-out below
    float *dv = dinv + ((int64_t)m * nt + k) * (2 * TILE_FLOATS);
#pragma unroll
    for (int q = 0; q < TILE2 / 256; ++q) {  // unrolled: all loads in flight at once
        const int e = q * 256 + tid;
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

constexpr int SPLIT_BLOCK = 512;  // BF16 per MFMA operand block (32 x 16)

typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float floatx2 __attribute__((ext_vector_type(2)));

// BF16 hi / lo parts of two FP32 values, packed (x0 in the low half): hi =
// bf16(x), lo = bf16(x - hi), both rounded to nearest even.  One packed
// conversion per part covers both values.
__device__ __forceinline__ void split2(float x0, float x1, uint32_t &hi, uint32_t &lo) {
    const floatx2 x = {x0, x1};
    const uint32_t h = __builtin_bit_cast(uint32_t, __builtin_convertvector(x, bf16x2));
    const floatx2 hf = {__builtin_bit_cast(float, h << 16), __builtin_bit_cast(float, h & 0xffff0000u)};
    lo = __builtin_bit_cast(uint32_t, __builtin_convertvector(x - hf, bf16x2));
    hi = h;
}
// The six split parts (re, im, re + im; hi, lo) of two complex values, as
// split2 packs them.
__device__ __forceinline__ void split6(float xr0, float xi0, float xr1, float xi1, uint32_t w[6]) {
    split2(xr0, xr1, w[0], w[1]);
    split2(xi0, xi1, w[2], w[3]);
    split2(xr0 + xi0, xr1 + xi1, w[4], w[5]);
}

// Split panel tiles (BF16 precisions).  The panels of a step group keep each
// tile they finish, for the group's Schur updates, as six planes of 64 x 64
// BF16: the real part, the imaginary part and their sum re + im (formed in
// FP32), each as a hi and a lo part, of -L(i,k) row-major ([row][col]) or of
// U(k,j) transposed ([col][row]), so that both MFMA operands read 8
// consecutive k values per lane.  The sum planes serve the three-
// multiplication complex product (Gauss):
//   P U = (Pr Ur - Pi Ui) + i ((Pr + Pi)(Ur + Ui) - Pr Ur - Pi Ui),
// three real products where the textbook form takes four.  The tile is stored
// in MFMA operand order: 1 KiB blocks, block ((q * 3 + c) * 2 + h) * 2 + a
// for K quarter q, component c (re, im, sum), part h (hi, lo) and 32-row half
// a, each holding the 32 x 16 operand of one v_mfma_f32_32x32x16_bf16 in lane
// order (lane l: row 32a + l % 32, k values 16q + 8 (l / 32) .. + 7).  One
// (quarter, component) is 4 KiB of consecutive memory, and a wave reads a
// block from LDS as one linear, bank-conflict-free 16-byte-per-lane read.
constexpr int PANEL_TILE = 6 * TILE2;  // BF16 elements per split panel tile (48 KiB)
__device__ __forceinline__ int panel_block(int q, int c, int h, int a) { return ((q * 3 + c) * 2 + h) * 2 + a; }

// Writes the tile held in LDS (re/im planes, row stride LDA) as a split panel
// tile.  negate_rows: -T row-major (the L operand); otherwise T transposed
// (the U operand).  Thread tid covers the 16 destination elements of one
// destination row in K quarter tid % 4: per plane, 8 k values in each of the
// two lane halves of one block (2 x 16 B).
__device__ __forceinline__ void write_panel(const float *Tre, const float *Tim, uint16_t *__restrict__ dst,
                                            bool negate_rows, int tid) {
    const int r = tid >> 2, q = tid & 3, c0 = 16 * q;
    uint32_t w[8][6];  // [k pair][component * 2 + part]
    auto ld = [&](int c, float &xr, float &xi) {
        if (negate_rows) {
            xr = -Tre[r * LDA + c0 + c];
            xi = -Tim[r * LDA + c0 + c];
        } else {
            xr = Tre[(c0 + c) * LDA + r];
            xi = Tim[(c0 + c) * LDA + r];
        }
    };
#pragma unroll
    for (int c = 0; c < 16; c += 2) {
        float xr0, xi0, xr1, xi1;
        ld(c, xr0, xi0);
        ld(c + 1, xr1, xi1);
        split6(xr0, xi0, xr1, xi1, w[c / 2]);
    }
#pragma unroll
    for (int p = 0; p < 6; ++p) {
        uint16_t *b = dst + panel_block(q, p >> 1, p & 1, r >> 5) * SPLIT_BLOCK + (r & 31) * 8;
        *reinterpret_cast<uint4 *>(b) = make_uint4(w[0][p], w[1][p], w[2][p], w[3][p]);        // k 16q + 0..7
        *reinterpret_cast<uint4 *>(b + 256) = make_uint4(w[4][p], w[5][p], w[6][p], w[7][p]);  // k 16q + 8..15
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
        float xr[4], xi[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            xr[e] = sr * (transpose ? re[e][u] : re[u][e]);
            xi[e] = si * (transpose ? im[e][u] : im[u][e]);
        }
        uint2 rh, rl, ih, il;
        split2(xr[0], xr[1], rh.x, rl.x);
        split2(xr[2], xr[3], rh.y, rl.y);
        split2(xi[0], xi[1], ih.x, il.x);
        split2(xi[2], xi[3], ih.y, il.y);
        const int off = row * SLD + k0;
        *reinterpret_cast<uint2 *>(S + off) = rh;
        *reinterpret_cast<uint2 *>(S + SPL + off) = rl;
        *reinterpret_cast<uint2 *>(S + 2 * SPL + off) = ih;
        *reinterpret_cast<uint2 *>(S + 3 * SPL + off) = il;
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
    {
        // both operand tiles load before either is split: one memory round
        // trip, not two
        const TileBlock ta = tile_block_load(lower ? t : d, tid);                // L(i,k) or inv(L_kk)
        const TileBlock tb = tile_block_load(lower ? d + TILE_FLOATS : t, tid);  // inv(U_kk) or U(k,j)
        stage_split_block(ta, As, false, false, false, tid);
        stage_split_block(tb, Bs, true, false, false, tid);
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
    uint16_t *dst = split + ((int64_t)m * split_stride + slot0 + blockIdx.x) * PANEL_TILE;
    write_panel(Tre, Tim, dst, lower, tid);
}

// Register vector of a thread's share of 32 KiB (8 chunks of 16 B), not an
// array, which the compiler may place in scratch memory when it is carried
// across loop iterations.
typedef uint32_t u32x32 __attribute__((ext_vector_type(32)));

// The re and im planes of a split panel tile (32 KiB of its 48): 2048 chunks
// of 16 B, chunk e being lane e % 64 of the re/im block e / 64, whose bits
// are (K quarter, component, part, row half) from high to low.  Thread tid
// moves chunks q * 256 + tid, so each wave reads whole blocks.
__device__ __forceinline__ u32x32 fetch_panel4(const uint16_t *__restrict__ src, int tid) {
    u32x32 r;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int e = q * 256 + tid, b = e >> 6;
        const uint4 v = *reinterpret_cast<const uint4 *>(
            src + panel_block(b >> 3, (b >> 2) & 1, (b >> 1) & 1, b & 1) * SPLIT_BLOCK + (e & 63) * 8);
        r[4 * q] = v.x;
        r[4 * q + 1] = v.y;
        r[4 * q + 2] = v.z;
        r[4 * q + 3] = v.w;
    }
    return r;
}
// ... into the LDS operand layout of mma_split (planes re_hi, re_lo, im_hi,
// im_lo; row stride SLD)
__device__ __forceinline__ void put_panel4(uint16_t *dst, const u32x32 r, int tid) {
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int e = q * 256 + tid, b = e >> 6, l = e & 63;
        const int plane = ((b >> 2) & 1) * 2 + ((b >> 1) & 1);
        const int row = (b & 1) * 32 + (l & 31), k = (b >> 3) * 16 + (l >> 5) * 8;
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
    const uint16_t *sp = split + (int64_t)m * split_stride * PANEL_TILE;
    // steps of the group with both L(i,t) and U(t,j) present (groups <= 32 steps)
    uint32_t mask = (uint32_t)pair_mask[q];
    floatx16 cre, cim;
    load_acc(C, cre, cim, wave, lane);
    if (mask) {
        int t = g0 + __builtin_ctz(mask);
        mask &= mask - 1;
        u32x32 ra = fetch_panel4(sp + (int64_t)ppos[(int64_t)t * nt + i] * PANEL_TILE, tid);
        u32x32 rb = fetch_panel4(sp + (int64_t)(ppos[(int64_t)t * nt + j] + step_s[t]) * PANEL_TILE, tid);
        for (;;) {
            __syncthreads();  // the previous step's products are done with LDS
            put_panel4(As, ra, tid);
            put_panel4(Bs, rb, tid);
            __syncthreads();
            const bool more = mask != 0;
            if (more) {
                t = g0 + __builtin_ctz(mask);
                mask &= mask - 1;
            }
            // the next step's tiles (the current step's again after the
            // last), their loads spread over this step's products
            ra = fetch_panel4(sp + (int64_t)ppos[(int64_t)t * nt + i] * PANEL_TILE, tid);
            rb = fetch_panel4(sp + (int64_t)(ppos[(int64_t)t * nt + j] + step_s[t]) * PANEL_TILE, tid);
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

// Trailing Schur updates by the three-multiplication complex product (BF16
// precisions).  A(i,j) -= sum_t L(i,t) U(t,j) is applied as C += P U with
// P = -L from the split panel tiles:
//   T1 = Pr Ur,  T2 = Pi Ui,  T3 = (Pr + Pi)(Ur + Ui),
//   Re C += T1 - T2,  Im C += T3 - T1 - T2,
// each real product as hi*hi + hi*lo + lo*hi (THREE) or hi*hi, FP32
// accumulation.  Three accumulators per 32 x 32 block start at x1 = Re C,
// x2 = 0, x3 = Re C + Im C and end as Re C = x1 - x2, Im C = x3 - x1 - x2.
// Accumulator x_c takes, per step and K quarter, its hi*hi, hi*lo and lo*hi
// products in that order in both kernels below, so they agree bit for bit.

// Reference form, one target per workgroup (wave w: the 32 x 32 block
// (w / 2, w % 2)): both split tiles of a step are copied whole into LDS
// (96 KiB, one workgroup per CU).  Runs the trailing pairs of a group over
// the matrices listed in `mats`.
template <bool THREE>
__global__ void __launch_bounds__(256, 1) k_schur_pairs_3m(
    float *__restrict__ tiles, int64_t mat_stride, int nt, int g0, const int32_t *__restrict__ pairs,
    const int32_t *__restrict__ pair_mask, int npairs, int nmat, const int32_t *__restrict__ tile_id,
    const uint16_t *__restrict__ split, int64_t split_stride, const int32_t *__restrict__ ppos,
    const int32_t *__restrict__ step_s, const int32_t *__restrict__ mats) {
    __shared__ __attribute__((aligned(16))) uint16_t As[PANEL_TILE], Bs[PANEL_TILE];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int64_t total = (int64_t)npairs * nmat, per = (total + 7) / 8;
    const int64_t item = (int64_t)(blockIdx.x % 8) * per + blockIdx.x / 8;
    if (item >= total) return;
    const int m = mats[item / npairs], q = (int)(item % npairs);
    const int i = pairs[2 * q], j = pairs[2 * q + 1];
    float *C = tile_ptr(tiles, mat_stride, m, tile_id[(int64_t)i * nt + j]);
    const uint16_t *sp = split + (int64_t)m * split_stride * PANEL_TILE;
    const int a = wave >> 1, b = wave & 1;
    floatx16 x[3];
    load_acc(C, x[0], x[2], wave, lane);
    x[2] += x[0];
    x[1] = floatx16{};
    for (uint32_t mask = (uint32_t)pair_mask[q]; mask; mask &= mask - 1) {
        const int t = g0 + __builtin_ctz(mask);
        const uint4 *la = reinterpret_cast<const uint4 *>(sp + (int64_t)ppos[(int64_t)t * nt + i] * PANEL_TILE);
        const uint4 *ub = reinterpret_cast<const uint4 *>(sp + (int64_t)(ppos[(int64_t)t * nt + j] + step_s[t]) * PANEL_TILE);
        __syncthreads();  // the previous step's products are done with LDS
        for (int e = tid; e < PANEL_TILE / 8; e += 256) {
            reinterpret_cast<uint4 *>(As)[e] = la[e];
            reinterpret_cast<uint4 *>(Bs)[e] = ub[e];
        }
        __syncthreads();
#pragma unroll
        for (int kq = 0; kq < 4; ++kq)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const bf16x8 ah = lds8(As + panel_block(kq, c, 0, a) * SPLIT_BLOCK + lane * 8);
                const bf16x8 bh = lds8(Bs + panel_block(kq, c, 0, b) * SPLIT_BLOCK + lane * 8);
                x[c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bh, x[c], 0, 0, 0);
                if (THREE) {
                    const bf16x8 al = lds8(As + panel_block(kq, c, 1, a) * SPLIT_BLOCK + lane * 8);
                    const bf16x8 bl = lds8(Bs + panel_block(kq, c, 1, b) * SPLIT_BLOCK + lane * 8);
                    x[c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bl, x[c], 0, 0, 0);
                    x[c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bh, x[c], 0, 0, 0);
                }
            }
    }
    const floatx16 cre = x[0] - x[1], cim = x[2] - x[0] - x[1];
    store_acc(C, cre, cim, wave, lane);
}

template __global__ void k_schur_pairs_3m<true>(float *, int64_t, int, int, const int32_t *, const int32_t *, int, int,
                                               const int32_t *, const uint16_t *, int64_t, const int32_t *,
                                               const int32_t *, const int32_t *);
template __global__ void k_schur_pairs_3m<false>(float *, int64_t, int, int, const int32_t *, const int32_t *, int,
                                                int, const int32_t *, const uint16_t *, int64_t, const int32_t *,
                                                const int32_t *, const int32_t *);

// Trailing update of a step group on 2 x 2 blocks of targets (the production
// form).  Quad q = (i0, i1, j0, j1) (an index -1 is absent); wave w owns the
// 64 x 64 target (i_{w/2}, j_{w%2}) over the steps where both of its operands
// exist, with three accumulators per 32 x 32 block (192 registers).  The
// split tiles L(i0,t), L(i1,t), U(t,j0), U(t,j1) of each listed step t pass
// through LDS in stages of one K quarter and one component: the hi and lo
// blocks of both row halves of the four tiles, 16 KiB, of which a wave reads
// its A and B halves, 8 KiB.  Stages cycle through four buffers (64 KiB, two
// workgroups per CU: while one waits at a barrier or on its target tiles,
// the other one's products keep the matrix cores busy).  Each stage opens
// with one barrier, after which the buffer read in the previous stage is
// refilled three stages ahead: the barrier certifies that the stage's loads
// have landed (every wave first waits for its own) and that every wave has
// read that buffer (its reads fed the previous stage's products).  The loads
// go from memory straight into LDS (buffer loads with the LDS destination: no
// registers, no LDS write instructions).  A step's 12 stages (4 quarters x 3
// components) are unrolled, so every stage's buffer and component are known
// at compile time.  Waves without an update in a step multiply zeroed A
// operands instead of branching.
constexpr int G_STAGE = 16 * SPLIT_BLOCK;  // BF16 elements per stage buffer (16 KiB)

// s_waitcnt immediates (gfx9 encoding: vmcnt[3:0] | expcnt << 4 | lgkmcnt << 8
// | vmcnt[5:4] << 14): vmcnt <= 8 or <= 4 together with lgkmcnt = 0, and
// vmcnt = 0 alone
constexpr int WAIT_VM8_LGKM0 = 0x0078, WAIT_VM4_LGKM0 = 0x0074, WAIT_VM0 = 0x0F70;

// wave w moves the blocks (K quarter q, component c) of its tile -- hi of
// both row halves, then lo -- from split slot s into stage buffer `buf`
template <bool THREE>
__device__ __forceinline__ void g_issue(uint16_t *buf, const uint16_t *__restrict__ sp, int s, int q, int c, int wave,
                                        int lane) {
    // buffer (not flat) loads: the compiler then tracks them as memory loads
    // into LDS only, and the LDS reads of the other buffers wait for none of them
    const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<uint16_t *>(sp + (int64_t)s * PANEL_TILE), 0, PANEL_TILE * 2, 0x00020000);
    uint16_t *d = buf + wave * 4 * SPLIT_BLOCK;
    const int first = panel_block(q, c, 0, 0);  // blocks first + 2h + a
#pragma unroll
    for (int j = 0; j < (THREE ? 4 : 2); ++j)
        __builtin_amdgcn_raw_ptr_buffer_load_lds(rsrc, reinterpret_cast<void *>(d + j * SPLIT_BLOCK), 16, lane * 16,
                                                 (first + j) * SPLIT_BLOCK * 2, 0, 0);
}
// the split slot of wave w's tile in a listed step's slot entry
__device__ __forceinline__ int wave_slot(const int4 &slot, int wave) {
    return __builtin_amdgcn_readfirstlane(wave == 0 ? slot.x : wave == 1 ? slot.y : wave == 2 ? slot.z : slot.w);
}

__device__ __forceinline__ bf16x8 lds8_and(const uint16_t *p, uint32_t keep) {
    return __builtin_bit_cast(bf16x8, __builtin_bit_cast(u32x4, lds8(p)) & (u32x4)(keep));
}

// One stage's operands of a wave: A from tile ua (masked by `keep`), B from
// tile ub; [row half]
struct StageOps {
    bf16x8 ah[2], al[2], bh[2], bl[2];
};
template <bool THREE>
__device__ __forceinline__ StageOps g_read(const uint16_t *buf, int ua, int ub, int lane, uint32_t keep) {
    auto blk = [&](int u, int j) { return buf + (u * 4 + j) * SPLIT_BLOCK + lane * 8; };
    StageOps o;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        o.ah[h] = lds8_and(blk(ua, h), keep);
        o.bh[h] = lds8(blk(ub, h));
        if (THREE) {
            o.al[h] = lds8_and(blk(ua, 2 + h), keep);
            o.bl[h] = lds8(blk(ub, 2 + h));
        }
    }
    return o;
}
// The stage's products into the component's four accumulators, which take
// turns so that consecutive MFMAs never depend on each other.
template <bool THREE>
__device__ __forceinline__ void g_products(const StageOps &o, floatx16 (&x)[4]) {
#pragma unroll
    for (int p = 0; p < (THREE ? 3 : 1); ++p)
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int a = t >> 1, b = t & 1;
            const bf16x8 u = p == 2 ? o.al[a] : o.ah[a];
            const bf16x8 v = p == 1 ? o.bl[b] : o.bh[b];
            x[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(u, v, x[t], 0, 0, 0);
        }
}

template <bool THREE>
__global__ void __launch_bounds__(256, 2) k_schur_quad_3m(
    float *__restrict__ tiles, int64_t mat_stride, int nt, int g0, const int32_t *__restrict__ quads,
    const int32_t *__restrict__ quad_mask, int nquads, int nmat, const int32_t *__restrict__ tile_id,
    const uint16_t *__restrict__ split, int64_t split_stride, const int32_t *__restrict__ ppos,
    const int32_t *__restrict__ step_s, const int32_t *__restrict__ mats) {
    // four buffers as four LDS objects: the compiler then knows that the
    // reads of one do not wait for the direct loads into another
    __shared__ __attribute__((aligned(1024))) uint16_t b0[G_STAGE];
    __shared__ __attribute__((aligned(1024))) uint16_t b1[G_STAGE];
    __shared__ __attribute__((aligned(1024))) uint16_t b2[G_STAGE];
    __shared__ __attribute__((aligned(1024))) uint16_t b3[G_STAGE];
    __shared__ int4 slots[32];  // split slots of the four tiles of the i-th listed step
    const int tid = threadIdx.x, wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
    const int64_t total = (int64_t)nquads * nmat, per = (total + 7) / 8;
    const int64_t item = (int64_t)(blockIdx.x % 8) * per + blockIdx.x / 8;
    if (item >= total) return;
    const int m = mats[item / nquads], q = (int)(item % nquads);
    const int I[2] = {quads[4 * q], quads[4 * q + 1]}, J[2] = {quads[4 * q + 2], quads[4 * q + 3]};
    const uint16_t *sp = split + (int64_t)m * split_stride * PANEL_TILE;
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
            const int bt = __builtin_ctz(rest);
            rest &= rest - 1;
            if (mine & (1u << bt)) work |= 1u << i;
            if (tid == i) {
                // an absent tile takes a present one's slot: its blocks are
                // loaded but read by no wave
                const uint32_t bit = 1u << bt;
                const int32_t *pp = ppos + (int64_t)(g0 + bt) * nt;
                const int ss = step_s[g0 + bt];
                const bool any_c = (cm[0] | cm[1]) & bit, any_r = (rm[0] | rm[1]) & bit;
                const int s0 = (rm[0] & bit) && any_c ? pp[I[0]] : -1, s1 = (rm[1] & bit) && any_c ? pp[I[1]] : -1;
                const int s2 = (cm[0] & bit) && any_r ? pp[J[0]] + ss : -1;
                const int s3 = (cm[1] & bit) && any_r ? pp[J[1]] + ss : -1;
                const int any = s0 >= 0 ? s0 : (s1 >= 0 ? s1 : (s2 >= 0 ? s2 : s3));
                slots[i] = make_int4(s0 >= 0 ? s0 : any, s1 >= 0 ? s1 : any, s2 >= 0 ? s2 : any, s3 >= 0 ? s3 : any);
            }
        }
    }
    float *C = mine ? tile_ptr(tiles, mat_stride, m, tile_id[(int64_t)I[ia] * nt + J[jb]]) : nullptr;
    const int ua = ia, ub = 2 + jb;
    uint16_t *const bufs[4] = {b0, b1, b2, b3};
    // stage u (0..11) of a listed step: K quarter u / 3, component u % 3,
    // buffer u % 4 (a step has 12 stages); past the end the last step's
    // stages are loaded again, into buffers no later stage reads
    auto issue = [&](int s, int u) { g_issue<THREE>(bufs[u & 3], sp, s, u / 3, u % 3, wave, lane); };
    __syncthreads();  // the slots are visible
    // the first three stages load while the target tile does
    if (n) {
        const int s0 = wave_slot(slots[0], wave);
        issue(s0, 0);
        issue(s0, 1);
        issue(s0, 2);
    }
    floatx16 x[3][4];
    if (mine) {
        load_acc_tile(C, x[0], x[2], lane);
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            x[2][t] += x[0][t];
            x[1][t] = floatx16{};
        }
    } else {
#pragma unroll
        for (int t = 0; t < 4; ++t) x[0][t] = x[1][t] = x[2][t] = floatx16{};
    }
    if (n) {
        for (int i = 0; i < n; ++i) {
            const uint32_t keep = (work >> i) & 1u ? 0xffffffffu : 0u;
            // this step's and the next step's slots, looked up once
            const int sc = wave_slot(slots[i], wave), sn = wave_slot(slots[i + 1 < n ? i + 1 : n - 1], wave);
#pragma unroll
            for (int u = 0; u < 12; ++u) {
                // this stage's loads land (the next two stages' stay in
                // flight) and this wave's reads of the previous stage are
                // done; the barrier then makes every wave's loads visible and
                // frees the previous stage's buffer for a refill
                __builtin_amdgcn_s_waitcnt(THREE ? WAIT_VM8_LGKM0 : WAIT_VM4_LGKM0);
                __builtin_amdgcn_s_barrier();
                issue(u + 3 < 12 ? sc : sn, (u + 3) % 12);
                const StageOps o = g_read<THREE>(bufs[u & 3], ua, ub, lane, keep);
                __builtin_amdgcn_sched_barrier(0);  // the loads go out before the products
                // the products run at raised wave priority, so that the other
                // wave on the SIMD takes its barrier and reads meanwhile
                __builtin_amdgcn_s_setprio(1);
                g_products<THREE>(o, x[u % 3]);
                __builtin_amdgcn_s_setprio(0);
                __builtin_amdgcn_sched_barrier(0);  // ... which stay ahead of the next barrier
            }
        }
        // the loads issued past the end land before LDS is reused
        __builtin_amdgcn_s_waitcnt(WAIT_VM0);
        __syncthreads();
    }
    if (mine) {
        floatx16 cre[4], cim[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            cre[t] = x[0][t] - x[1][t];
            cim[t] = x[2][t] - x[0][t] - x[1][t];
        }
        store_acc_tile(C, cre, cim, lane);
    }
}

template __global__ void k_schur_quad_3m<true>(float *, int64_t, int, int, const int32_t *, const int32_t *, int, int,
                                              const int32_t *, const uint16_t *, int64_t, const int32_t *,
                                              const int32_t *, const int32_t *);
template __global__ void k_schur_quad_3m<false>(float *, int64_t, int, int, const int32_t *, const int32_t *, int,
                                               int, const int32_t *, const uint16_t *, int64_t, const int32_t *,
                                               const int32_t *, const int32_t *);


}  // namespace aosx
