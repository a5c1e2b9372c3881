// Tile triangular solves, RHS packing and FP64 refinement residuals.
//
// A right-hand-side block of W columns (W a multiple of 32) is stored per
// matrix as Nt tiles of TILE rows x W interleaved complex64 values, rows in
// factor (permuted) order.  Solves run one tile step per launch:
//   diag   : Y_k <- D_k Y_k with D_k = inv(L_kk), inv(U_kk) or their
//            conjugate transposes (for A^H = U^H L^H, used by conjugate shifts);
//   update : Y_i -= T_ik Y_k for the tiles i coupled to k.
// Every launch writes only tiles that no other block of the same launch reads.
#include "device_common.h"

namespace aosx {

// Load a stored tile (planar) into LDS as row-major re/im with stride LDA,
// optionally as its conjugate transpose.
__device__ __forceinline__ void stage_tile(const float *__restrict__ g, float *Tre, float *Tim, bool ctrans,
                                           int tid) {
    for (int e = tid; e < TILE2; e += 256) {
        const int r = e / TILE, c = e % TILE;
        const float re = g[e], im = g[PLANE + e];
        if (ctrans) { Tre[c * LDA + r] = re; Tim[c * LDA + r] = -im; }
        else { Tre[r * LDA + c] = re; Tim[r * LDA + c] = im; }
    }
}

// out[r][c] = sum_j T[r][j] Y[j][c] for this thread's 8 rows (r0 + 8q) and column c.
__device__ __forceinline__ void apply_tile(const float *Tre, const float *Tim, const float2 *Ys, int r0, int c,
                                           float2 *out) {
#pragma unroll
    for (int q = 0; q < 8; ++q) out[q] = make_float2(0.f, 0.f);
    for (int j = 0; j < TILE; ++j) {
        const float2 y = Ys[j * 32 + c];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int r = r0 + 8 * q;
            const float tr = Tre[r * LDA + j], ti = Tim[r * LDA + j];
            out[q].x += tr * y.x - ti * y.y;
            out[q].y += tr * y.y + ti * y.x;
        }
    }
}

// which: 0 inv(L_kk), 1 inv(U_kk); ctrans applies the conjugate transpose.
__global__ void __launch_bounds__(256) k_solve_diag(float2 *__restrict__ y, int64_t y_mat_stride, int W,
                                                   const float *__restrict__ dinv, int nt, int k, int which,
                                                   int ctrans) {
    __shared__ float Tre[TILE * LDA], Tim[TILE * LDA];
    __shared__ float2 Ys[TILE * 32];
    const int m = blockIdx.y, chunk = blockIdx.x, tid = threadIdx.x;
    const float *d = dinv + ((int64_t)m * nt + k) * (2 * TILE_FLOATS) + which * TILE_FLOATS;
    stage_tile(d, Tre, Tim, ctrans != 0, tid);
    float2 *yk = y + (int64_t)m * y_mat_stride + (int64_t)k * TILE * W + chunk * 32;
    for (int e = tid; e < TILE * 32; e += 256) Ys[e] = yk[(e / 32) * W + e % 32];
    __syncthreads();
    const int c = tid & 31, r0 = tid >> 5;
    float2 out[8];
    apply_tile(Tre, Tim, Ys, r0, c, out);
#pragma unroll
    for (int q = 0; q < 8; ++q) yk[(r0 + 8 * q) * W + c] = out[q];
}

// Y_i -= T Y_k for i = list[blockIdx.y]; T is tile (i,k) or, with ctrans, the
// conjugate transpose of tile (k,i).
__global__ void __launch_bounds__(256) k_solve_update(float2 *__restrict__ y, int64_t y_mat_stride, int W,
                                                     const float *__restrict__ tiles, int64_t mat_stride,
                                                     const int32_t *__restrict__ tile_id, int nt, int k,
                                                     const int32_t *__restrict__ list, int ctrans) {
    __shared__ float Tre[TILE * LDA], Tim[TILE * LDA];
    __shared__ float2 Ys[TILE * 32];
    const int m = blockIdx.z, chunk = blockIdx.x, tid = threadIdx.x;
    const int i = list[blockIdx.y];
    const int32_t id = ctrans ? tile_id[(int64_t)k * nt + i] : tile_id[(int64_t)i * nt + k];
    stage_tile(tiles + (int64_t)m * mat_stride + (int64_t)id * TILE_FLOATS, Tre, Tim, ctrans != 0, tid);
    const float2 *yk = y + (int64_t)m * y_mat_stride + (int64_t)k * TILE * W + chunk * 32;
    for (int e = tid; e < TILE * 32; e += 256) Ys[e] = yk[(e / 32) * W + e % 32];
    __syncthreads();
    const int c = tid & 31, r0 = tid >> 5;
    float2 out[8];
    apply_tile(Tre, Tim, Ys, r0, c, out);
    float2 *yi = y + (int64_t)m * y_mat_stride + (int64_t)i * TILE * W + chunk * 32;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        float2 v = yi[(r0 + 8 * q) * W + c];
        v.x -= out[q].x;
        v.y -= out[q].y;
        yi[(r0 + 8 * q) * W + c] = v;
    }
}

// Y[m] (tiled, permuted, W columns) <- src[src_index[m]] (row-major n x w_src);
// columns >= w_src and padded rows become zero.
__global__ void k_pack_rhs(float2 *__restrict__ y, int64_t y_mat_stride, int W, const float2 *__restrict__ src,
                           int w_src, const int32_t *__restrict__ src_index, const int32_t *__restrict__ perm,
                           int64_t n, int64_t padded) {
    const int m = blockIdx.y;
    const float2 *s = src + (int64_t)src_index[m] * n * w_src;
    float2 *ym = y + (int64_t)m * y_mat_stride;
    const int64_t total = padded * W, stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += stride) {
        const int64_t row = e / W;
        const int c = (int)(e % W);
        float2 v = make_float2(0.f, 0.f);
        if (row < n && c < w_src) v = s[(int64_t)perm[row] * w_src + c];
        ym[e] = v;
    }
}

// dst[m] (row-major n x w_dst, original order) <- Y[m]
__global__ void k_unpack(float2 *__restrict__ dst, int w_dst, const float2 *__restrict__ y, int64_t y_mat_stride,
                         int W, const int32_t *__restrict__ perm, int64_t n) {
    const int m = blockIdx.y;
    const float2 *ym = y + (int64_t)m * y_mat_stride;
    float2 *d = dst + (int64_t)m * n * w_dst;
    const int64_t total = n * w_dst, stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += stride) {
        const int64_t row = e / w_dst;
        const int c = (int)(e % w_dst);
        d[(int64_t)perm[row] * w_dst + c] = ym[row * W + c];
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
