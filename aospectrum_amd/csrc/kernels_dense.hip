// Tall-skinny and small dense kernels of the CISS subspace stage.
//
// Tall-skinny operands (n x w, row-major, complex64) come from the shifted
// solves; the projected problems are small (w <= 256) and live in FP64:
//   spmm        Y = A X with A a CSR H(k) or S(k) (complex64, FP32 accumulate)
//   moments     M[:, mL:(m+1)L] += coef[b][m] X_b           (ciss.py:324-333)
//   gram64      G = A^H B in FP64 (two-stage, deterministic order)
//   tall_gemm   Y = X T with T small (FP64), output complex64 or complex128
//   cholesky / trinv / small_gemm for the projected generalized problem
//   eig_residual ||H v - e S v||_2 per column in FP64
#include "device_common.h"

namespace aosx {

// Y[j] = A_k X[j] for the matrices j of one launch: A values a[kidx[j]],
// X at x + xidx[j]*n*w, Y at y + j*n*w.  One wave per row, lanes over columns.
__global__ void __launch_bounds__(256) k_spmm(float2 *__restrict__ y, const float2 *__restrict__ a,
                                             const int32_t *__restrict__ kidx, const float2 *__restrict__ x,
                                             const int32_t *__restrict__ xidx, int w,
                                             const int32_t *__restrict__ indptr,
                                             const int32_t *__restrict__ indices, int64_t n, int64_t nnz) {
    const int j = blockIdx.y;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const float2 *ak = a + (int64_t)kidx[j] * nnz;
    const float2 *xj = x + (int64_t)xidx[j] * n * w;
    float2 *yj = y + (int64_t)j * n * w;
    for (int64_t row = (int64_t)blockIdx.x * 4 + wave; row < n; row += (int64_t)gridDim.x * 4) {
        for (int c = lane; c < w; c += 64) {
            float ar = 0.f, ai = 0.f;
            for (int32_t p = indptr[row]; p < indptr[row + 1]; ++p) {
                const float2 v = ak[p], xv = xj[(int64_t)indices[p] * w + c];
                ar += v.x * xv.x - v.y * xv.y;
                ai += v.x * xv.y + v.y * xv.x;
            }
            yj[row * w + c] = make_float2(ar, ai);
        }
    }
}

// Atom-block CSR x tall-skinny product.  Every orbital row of an atom has
// the same column list (the pattern is a union of full atom blocks), so one
// workgroup takes one atom's rows and reads each needed row of x once for all
// of them -- an atom's worth (13x) less gather traffic than row-wise SpMM.
// Values are staged through LDS in chunks of CH columns; accumulation is FP64.
//   RESIDUAL: out[m] = b[b_index[m]] - (H_k - z_m S_k) x[m], with ||out||^2 and
//             ||b||^2 added to sumsq[m], bsq[m] (FP64 iterative refinement);
//   else    : out[m] = A_k x[x_index[m]] with A = h (H or S).
// The 8 half-waves of the block take NCC = min(4, ceil(w / 32)) (rounded up
// to a power of two) 32-column chunks of x at a time and split the column
// list S = 8 / NCC ways per chunk (stream s of a chunk takes entries s, s + S,
// ...); each accumulates every row of the atom (up to MAXR) for its 32
// columns, so each gathered x element is loaded once per block and feeds all
// of the atom's rows, and the values of a chunk of the column list are staged
// once for all NCC column chunks.  The partial sums of a column chunk's
// streams are added across the two half-waves of a wave by a shuffle and,
// when a chunk spans several waves, through LDS.
// Grid: 8 * ceil(n_atoms * nmat / 8) blocks in one dimension; block b takes
// item (b % 8) * per + b / 8 = (matrix, atom) so that each XCD walks a
// contiguous run of atoms of one matrix, whose gathers share x rows in L2.
// Three waves per SIMD (168 registers; the compiler moves a few values to
// scratch) with the x gathers one batch ahead: on bulk3d 2.40 ms per call,
// against 2.80 ms without the batch ahead (no spill) and 3.24 ms for the
// form that looped over column chunks outside the column list.
template <bool RESIDUAL, int MAXR>
__global__ void __launch_bounds__(256, 3) k_atom_spmm(float2 *__restrict__ out, const float2 *__restrict__ b,
                                                  const int32_t *__restrict__ b_index, const float2 *__restrict__ x,
                                                  const int32_t *__restrict__ x_index, int w,
                                                  const float2 *__restrict__ h, const float2 *__restrict__ s,
                                                  const int32_t *__restrict__ kidx, const float2 *__restrict__ z,
                                                  const int32_t *__restrict__ indptr,
                                                  const int32_t *__restrict__ indices,
                                                  const int32_t *__restrict__ atom_off, int64_t n, int64_t nnz,
                                                  int n_atoms, int nmat, double *__restrict__ sumsq,
                                                  double *__restrict__ bsq) {
    constexpr int CH = 64, XB = 8;  // column-list chunk; x values gathered per batch
    __shared__ double2 vals[CH][MAXR];  // also the cross-wave reduction buffer
    __shared__ int32_t cidx[CH];
    __shared__ int32_t rstart[MAXR];
    const int64_t total = (int64_t)n_atoms * nmat, per = (total + 7) / 8;
    const int64_t item = (int64_t)(blockIdx.x % 8) * per + blockIdx.x / 8;
    if (item >= total) return;
    const int m = (int)(item / n_atoms), a = (int)(item % n_atoms);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int half = lane >> 5, cl = lane & 31, hw = wave * 2 + half;
    const int r0 = atom_off[a], na = atom_off[a + 1] - r0;
    const int p0 = indptr[r0], len = indptr[r0 + 1] - p0;
    if (tid < na) rstart[tid] = indptr[r0 + tid];
    const int nc32 = (w + 31) / 32;
    const int ncc = nc32 >= 3 ? 4 : nc32;  // column chunks at a time: 1, 2 or 4
    const int S = 8 / ncc;                 // streams per column chunk
    const int cs = hw / S, st = hw % S;    // this half-wave's column chunk and stream
    const int wpc = S / 2;                 // waves per column chunk
    const int k = kidx[m];
    const float2 *hk = h + (int64_t)k * nnz;
    const float2 *sk = RESIDUAL ? s + (int64_t)k * nnz : nullptr;
    const double zr = RESIDUAL ? (double)z[m].x : 0.0, zi = RESIDUAL ? (double)z[m].y : 0.0;
    const float2 *xm = x + (int64_t)(RESIDUAL ? m : x_index[m]) * n * w;
    float2 *om = out + (int64_t)m * n * w;
    double local = 0.0, blocal = 0.0;
    for (int cg = 0; cg < w; cg += 32 * ncc) {
        const int c = cg + 32 * cs + cl;
        const bool cv = c < w;
        double ar[MAXR], ai[MAXR];
#pragma unroll
        for (int r = 0; r < MAXR; ++r) ar[r] = ai[r] = 0.0;
        for (int q0 = 0; q0 < len; q0 += CH) {
            const int cnt = min(CH, len - q0);
            __syncthreads();
            for (int e = tid; e < cnt * na; e += 256) {
                const int pp = e / na, r = e % na;
                const int64_t pos = rstart[r] + q0 + pp;
                const float2 hv = hk[pos];
                double vr = hv.x, vi = hv.y;
                if (RESIDUAL) {
                    const float2 sv = sk[pos];
                    vr -= zr * sv.x - zi * sv.y;
                    vi -= zr * sv.y + zi * sv.x;
                }
                vals[pp][r] = make_double2(vr, vi);
            }
            for (int e = tid; e < cnt; e += 256) cidx[e] = indices[p0 + q0 + e];
            __syncthreads();
            // this stream's entries st, st + S, ..., gathered XB at a time,
            // the next batch loading while the current one is multiplied
            const int mine = cnt > st ? (cnt - st + S - 1) / S : 0;
            float2 xv[XB], xn[XB];
            auto gather = [&](float2 (&dst)[XB], int u0) {
#pragma unroll
                for (int u = 0; u < XB; ++u) {
                    const int pp = st + S * (u0 + u);
                    dst[u] = (cv && u0 + u < mine) ? xm[(int64_t)cidx[pp] * w + c] : make_float2(0.f, 0.f);
                }
            };
            gather(xv, 0);
            for (int u0 = 0; u0 < mine; u0 += XB) {
                if (u0 + XB < mine) gather(xn, u0 + XB);
#pragma unroll
                for (int u = 0; u < XB; ++u) {
                    if (u0 + u >= mine) break;
                    const int pp = st + S * (u0 + u);
#pragma unroll
                    for (int r = 0; r < MAXR; ++r) {
                        if (r < na) {
                            const double2 v = vals[pp][r];
                            ar[r] += v.x * xv[u].x - v.y * xv[u].y;
                            ai[r] += v.x * xv[u].y + v.y * xv[u].x;
                        }
                    }
                }
#pragma unroll
                for (int u = 0; u < XB; ++u) xv[u] = xn[u];
            }
        }
        // the two half-waves of a wave hold streams of the same column chunk
#pragma unroll
        for (int r = 0; r < MAXR; ++r) {
            ar[r] += __shfl_down(ar[r], 32);
            ai[r] += __shfl_down(ai[r], 32);
        }
        // the other waves of the chunk into its first wave, through LDS
        // (vals is free once every wave has left the products)
        double2 *red = &vals[0][0];  // [column chunk][MAXR][32]
        const int wc = wave % wpc;   // this wave's place among its chunk's waves
        for (int src = 1; src < wpc; ++src) {
            __syncthreads();
            if (wc == src && half == 0) {
#pragma unroll
                for (int r = 0; r < MAXR; ++r)
                    if (r < na) red[(cs * MAXR + r) * 32 + cl] = make_double2(ar[r], ai[r]);
            }
            __syncthreads();
            if (wc == 0 && half == 0) {
#pragma unroll
                for (int r = 0; r < MAXR; ++r)
                    if (r < na) {
                        ar[r] += red[(cs * MAXR + r) * 32 + cl].x;
                        ai[r] += red[(cs * MAXR + r) * 32 + cl].y;
                    }
            }
        }
        if (wc == 0 && half == 0 && cv) {
#pragma unroll
            for (int r = 0; r < MAXR; ++r) {
                if (r >= na) break;
                const int64_t row = r0 + r;
                if (RESIDUAL) {
                    const float2 bv = b[(int64_t)b_index[m] * n * w + row * w + c];
                    const double rr = bv.x - ar[r], ri = bv.y - ai[r];
                    om[row * w + c] = make_float2((float)rr, (float)ri);
                    local += rr * rr + ri * ri;
                    blocal += (double)bv.x * bv.x + (double)bv.y * bv.y;
                } else {
                    om[row * w + c] = make_float2((float)ar[r], (float)ai[r]);
                }
            }
        }
    }
    if (RESIDUAL && wave % wpc == 0) {
        for (int off = 16; off > 0; off >>= 1) {
            local += __shfl_down(local, off);
            blocal += __shfl_down(blocal, off);
        }
        if (lane == 0) {
            atomicAdd(sumsq + m, local);
            atomicAdd(bsq + m, blocal);
        }
    }
}

// M[row][m*L + c] += sum_b coef[b*msize + m] * X_b[row][c]   (b over nb solutions)
__global__ void k_moments(float2 *__restrict__ M, const float2 *__restrict__ X, int nb, int L, int msize,
                          const float2 *__restrict__ coef, int64_t n) {
    const int64_t total = n * L, stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += stride) {
        const int64_t row = e / L;
        const int c = (int)(e % L);
        for (int m = 0; m < msize; ++m) {
            float2 acc = M[row * (int64_t)(L * msize) + m * L + c];
            for (int b = 0; b < nb; ++b) {
                const float2 xv = X[(int64_t)b * n * L + e];
                const float2 cf = coef[b * msize + m];
                acc.x += cf.x * xv.x - cf.y * xv.y;
                acc.y += cf.x * xv.y + cf.y * xv.x;
            }
            M[row * (int64_t)(L * msize) + m * L + c] = acc;
        }
    }
}

// dst[k][row][dcol0 + c] = src[k][row][scol0 + c] for c < ncols
__global__ void k_copy_cols(float2 *__restrict__ dst, int dw, int dcol0, const float2 *__restrict__ src, int sw,
                            int scol0, int ncols, int64_t n) {
    const int k = blockIdx.y;
    const float2 *s = src + (int64_t)k * n * sw;
    float2 *d = dst + (int64_t)k * n * dw;
    const int64_t total = n * ncols, stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += stride) {
        const int64_t row = e / ncols;
        const int c = (int)(e % ncols);
        d[row * dw + dcol0 + c] = s[row * sw + scol0 + c];
    }
}

// Partial Gram: part[k][chunk][i][j] = sum_{rows of chunk} conj(A[row][i]) B[row][j]
// over a 32x32 output tile (grid.x tiles, grid.y chunks, grid.z problems).
__global__ void __launch_bounds__(256) k_gram_partial(c64 *__restrict__ part, const float2 *__restrict__ A,
                                                     int64_t a_stride, int P, const float2 *__restrict__ B,
                                                     int64_t b_stride, int Q, int64_t n, int rows_per_chunk) {
    __shared__ float2 As[32][33], Bs[32][33];
    const int k = blockIdx.z, chunk = blockIdx.y, tid = threadIdx.x;
    const int tiles_q = (Q + 31) / 32;
    const int i0 = (blockIdx.x / tiles_q) * 32, j0 = (blockIdx.x % tiles_q) * 32;
    const float2 *a = A + k * a_stride;
    const float2 *b = B + k * b_stride;
    const int64_t r_begin = (int64_t)chunk * rows_per_chunk;
    const int64_t r_end = min(n, r_begin + rows_per_chunk);
    const int ti = tid >> 3, tj = (tid & 7) * 4;
    double accr[4] = {0, 0, 0, 0}, acci[4] = {0, 0, 0, 0};
    for (int64_t r0 = r_begin; r0 < r_end; r0 += 32) {
        for (int e = tid; e < 32 * 32; e += 256) {
            const int rr = e >> 5, cc = e & 31;
            const int64_t row = r0 + rr;
            float2 va = make_float2(0.f, 0.f), vb = make_float2(0.f, 0.f);
            if (row < r_end) {
                if (i0 + cc < P) va = a[row * P + i0 + cc];
                if (j0 + cc < Q) vb = b[row * Q + j0 + cc];
            }
            As[rr][cc] = va;
            Bs[rr][cc] = vb;
        }
        __syncthreads();
        for (int rr = 0; rr < 32; ++rr) {
            const float2 av = As[rr][ti];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const float2 bv = Bs[rr][tj + q];
                accr[q] += (double)av.x * bv.x + (double)av.y * bv.y;
                acci[q] += (double)av.x * bv.y - (double)av.y * bv.x;
            }
        }
        __syncthreads();
    }
    c64 *out = part + ((int64_t)k * gridDim.y + chunk) * (int64_t)P * Q;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int i = i0 + ti, j = j0 + tj + q;
        if (i < P && j < Q) out[(int64_t)i * Q + j] = {accr[q], acci[q]};
    }
}

// G[k][i][j] (leading dimension ldg) = sum over chunks of part; hermitian=1
// also symmetrises (G + G^H)/2 (requires P == Q).
__global__ void k_gram_reduce(c64 *__restrict__ G, int ldg, int64_t g_stride, const c64 *__restrict__ part,
                              int P, int Q, int chunks, int hermitian) {
    const int k = blockIdx.y;
    const c64 *pk = part + (int64_t)k * chunks * P * Q;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < P * Q; e += gridDim.x * blockDim.x) {
        const int i = e / Q, j = e % Q;
        c64 s = {0.0, 0.0};
        for (int c = 0; c < chunks; ++c) s = cadd(s, pk[(int64_t)c * P * Q + e]);
        if (hermitian) {
            c64 t = {0.0, 0.0};
            for (int c = 0; c < chunks; ++c) t = cadd(t, pk[(int64_t)c * P * Q + (int64_t)j * Q + i]);
            s = {(s.re + t.re) * 0.5, (s.im - t.im) * 0.5};
        }
        G[k * g_stride + (int64_t)i * ldg + j] = s;
    }
}

// Y[k][row][j] = sum_i X[k][row][i] T[k][i][j]  (T leading dimension ldt)
template <typename Out>
__global__ void __launch_bounds__(256) k_tall_gemm(Out *__restrict__ Y, int64_t y_stride,
                                                  const float2 *__restrict__ X, int64_t x_stride, int P,
                                                  const c64 *__restrict__ T, int ldt, int64_t t_stride, int Q,
                                                  int64_t n) {
    const int k = blockIdx.z;
    const int j = blockIdx.y * 32 + (threadIdx.x & 31);
    const int64_t row = (int64_t)blockIdx.x * 8 + (threadIdx.x >> 5);
    if (row >= n || j >= Q) return;
    const float2 *x = X + k * x_stride + row * P;
    const c64 *t = T + k * t_stride;
    double ar = 0.0, ai = 0.0;
    for (int i = 0; i < P; ++i) {
        const float2 xv = x[i];
        const c64 tv = t[(int64_t)i * ldt + j];
        ar += xv.x * tv.re - xv.y * tv.im;
        ai += xv.x * tv.im + xv.y * tv.re;
    }
    Out o;
    o.x = ar;
    o.y = ai;
    Y[k * y_stride + row * Q + j] = o;
}
template __global__ void k_tall_gemm<float2>(float2 *, int64_t, const float2 *, int64_t, int, const c64 *, int,
                                             int64_t, int, int64_t);
template __global__ void k_tall_gemm<double2>(double2 *, int64_t, const float2 *, int64_t, int, const c64 *,
                                              int, int64_t, int, int64_t);

// Lower Cholesky in place (upper part zeroed); status[k] = 1 when not positive definite.
__global__ void __launch_bounds__(1024) k_cholesky(c64 *__restrict__ A_all, const int32_t *__restrict__ rk,
                                                  int ld, int64_t stride, int *__restrict__ status) {
    __shared__ int bad;
    const int k = blockIdx.x, tid = threadIdx.x, nthr = blockDim.x;
    const int r = rk[k];
    c64 *A = A_all + k * stride;
    if (tid == 0) bad = 0;
    __syncthreads();
    for (int j = 0; j < r; ++j) {
        if (tid == 0) {
            const double d = A[(int64_t)j * ld + j].re;
            if (!(d > 0.0) || !isfinite(d)) bad = 1;
            else A[(int64_t)j * ld + j] = {sqrt(d), 0.0};
        }
        __syncthreads();
        if (bad) break;
        const double djj = A[(int64_t)j * ld + j].re;
        for (int i = j + 1 + tid; i < r; i += nthr) {
            c64 &v = A[(int64_t)i * ld + j];
            v = {v.re / djj, v.im / djj};
        }
        __syncthreads();
        const int m = r - j - 1;
        for (int e = tid; e < m * m; e += nthr) {
            const int i = j + 1 + e / m, l = j + 1 + e % m;
            if (l > i) continue;
            const c64 a = A[(int64_t)i * ld + j], b = A[(int64_t)l * ld + j];
            c64 &v = A[(int64_t)i * ld + l];
            v = csub(v, cmul(a, cconj(b)));
        }
        __syncthreads();
    }
    if (tid == 0) status[k] = bad;
    if (!bad)
        for (int e = tid; e < r * r; e += nthr) {
            const int i = e / r, j = e % r;
            if (j > i) A[(int64_t)i * ld + j] = {0.0, 0.0};
        }
}

// Linv = inv(L) for lower-triangular L; one thread per column.
__global__ void __launch_bounds__(256) k_trinv_lower(c64 *__restrict__ Linv_all, const c64 *__restrict__ L_all,
                                                    const int32_t *__restrict__ rk, int ld, int64_t stride) {
    const int k = blockIdx.y;
    const int r = rk[k];
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= r) return;
    const c64 *L = L_all + k * stride;
    c64 *X = Linv_all + k * stride;
    for (int i = 0; i < c; ++i) X[(int64_t)i * ld + c] = {0.0, 0.0};
    for (int i = c; i < r; ++i) {
        c64 s = {i == c ? 1.0 : 0.0, 0.0};
        for (int j = c; j < i; ++j) s = csub(s, cmul(L[(int64_t)i * ld + j], X[(int64_t)j * ld + c]));
        const c64 d = L[(int64_t)i * ld + i];
        const double dd = cabs2(d);
        X[(int64_t)i * ld + c] = {(s.re * d.re + s.im * d.im) / dd, (s.im * d.re - s.re * d.im) / dd};
    }
}

// C = op(A) op(B), op 0: identity, 1: conjugate transpose; r x r problems
__global__ void k_small_gemm(c64 *__restrict__ C_all, const c64 *__restrict__ A_all, int opa,
                             const c64 *__restrict__ B_all, int opb, const int32_t *__restrict__ rk, int ld,
                             int64_t stride) {
    const int k = blockIdx.z;
    const int r = rk[k];
    const int i = blockIdx.y, j = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= r || j >= r) return;
    const c64 *A = A_all + k * stride, *B = B_all + k * stride;
    c64 s = {0.0, 0.0};
    for (int l = 0; l < r; ++l) {
        const c64 a = opa ? cconj(A[(int64_t)l * ld + i]) : A[(int64_t)i * ld + l];
        const c64 b = opb ? cconj(B[(int64_t)j * ld + l]) : B[(int64_t)l * ld + j];
        s = cadd(s, cmul(a, b));
    }
    C_all[k * stride + (int64_t)i * ld + j] = s;
}

// Out[k][i][j] = In[k][i][cols[k][j]] * scale[k][j] for i < rows, j < ncols[k]
__global__ void k_gather_cols(c64 *__restrict__ Out, int ldo, int64_t o_stride, const c64 *__restrict__ In,
                              int ldi, int64_t i_stride, const int32_t *__restrict__ cols,
                              const double *__restrict__ scale, int cap, const int32_t *__restrict__ ncols,
                              int rows) {
    const int k = blockIdx.y;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    const int nc = ncols[k];
    if (e >= rows * cap) return;
    const int i = e / cap, j = e % cap;
    c64 v = {0.0, 0.0};
    if (j < nc) {
        const c64 x = In[k * i_stride + (int64_t)i * ldi + cols[k * cap + j]];
        const double sc = scale[k * cap + j];
        v = {x.re * sc, x.im * sc};
    }
    Out[k * o_stride + (int64_t)i * ldo + j] = v;
}

// Ghost-filter scores (ciss.py:153-177): score_j = sum_i |C_ij|^2 / sum_i |C_ij|^2 / sigma_i
__global__ void k_ghost_scores(double *__restrict__ score, const c64 *__restrict__ C_all, int ld, int64_t stride,
                               const double *__restrict__ sigma, int sig_ld, const int32_t *__restrict__ rank) {
    const int k = blockIdx.y;
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = rank[k];
    if (j >= r) return;
    const c64 *C = C_all + k * stride;
    double num = 0.0, den = 0.0;
    for (int i = 0; i < r; ++i) {
        const double a = cabs2(C[(int64_t)i * ld + j]);
        num += a;
        den += a / sigma[(int64_t)k * sig_ld + i];
    }
    score[(int64_t)k * sig_ld + j] = den > 0.0 ? num / den : 0.0;
}

// norm2[k][j] += || (H_k - e_kj S_k) v_kj ||^2 over a row range (FP64).
// V: [k][n][cap] complex128; one wave per row, lanes over columns.
__global__ void __launch_bounds__(256) k_eig_residual(double *__restrict__ norm2, const double2 *__restrict__ V,
                                                     int cap, const int32_t *__restrict__ count,
                                                     const double *__restrict__ energies,
                                                     const float2 *__restrict__ h, const float2 *__restrict__ s,
                                                     const int32_t *__restrict__ indptr,
                                                     const int32_t *__restrict__ indices, int64_t n,
                                                     int64_t nnz) {
    const int k = blockIdx.y;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int cnt = count[k];
    const float2 *hk = h + (int64_t)k * nnz, *sk = s + (int64_t)k * nnz;
    const double2 *vk = V + (int64_t)k * n * cap;
    for (int c0 = 0; c0 < cnt; c0 += 64) {
        const int c = c0 + lane;
        double local = 0.0;
        if (c < cnt) {
            const double e = energies[(int64_t)k * cap + c];
            for (int64_t row = (int64_t)blockIdx.x * 4 + wave; row < n; row += (int64_t)gridDim.x * 4) {
                double rr = 0.0, ri = 0.0;
                for (int32_t p = indptr[row]; p < indptr[row + 1]; ++p) {
                    const double vr = (double)hk[p].x - e * sk[p].x;
                    const double vi = (double)hk[p].y - e * sk[p].y;
                    const double2 x = vk[(int64_t)indices[p] * cap + c];
                    rr += vr * x.x - vi * x.y;
                    ri += vr * x.y + vi * x.x;
                }
                local += rr * rr + ri * ri;
            }
            atomicAdd(norm2 + (int64_t)k * cap + c, local);
        }
    }
}

// The same residual on the atom-block structure (every orbital row of an atom
// has the same column list): workgroup (k, atom) stages the atom's H and S
// values of 64 columns at a time in LDS (FP64), and each lane -- one vector
// column c -- gathers V(j, c) once for all of the atom's rows and accumulates
// (H - e_c S) V over them.  The 8 half-waves split the column list as in
// k_atom_spmm and are reduced through LDS; per column one FP64 atomic add of
// the atom's sum of squares.  V is gathered once per atom instead of once per
// row.  Grid: 8 * ceil(n_atoms * nk / 8) blocks; block b takes item
// (b % 8) * per + b / 8 = (k, atom), so each XCD walks a contiguous run of one
// k point's atoms.
template <int MAXR>
__global__ void __launch_bounds__(256) k_eig_residual_atoms(
    double *__restrict__ norm2, const double2 *__restrict__ V, int cap, const int32_t *__restrict__ count,
    const double *__restrict__ energies, const float2 *__restrict__ h, const float2 *__restrict__ s,
    const int32_t *__restrict__ indptr, const int32_t *__restrict__ indices, const int32_t *__restrict__ atom_off,
    int64_t n, int64_t nnz, int n_atoms, int nk) {
    constexpr int CH = 64, NS = 8;
    __shared__ double2 hv[CH][MAXR], sv[CH][MAXR];
    __shared__ double2 red[MAXR][32];
    __shared__ int32_t cidx[CH];
    const int64_t total = (int64_t)n_atoms * nk, per = (total + 7) / 8;
    const int64_t item = (int64_t)(blockIdx.x % 8) * per + blockIdx.x / 8;
    if (item >= total) return;
    const int k = (int)(item / n_atoms), a = (int)(item % n_atoms);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int half = lane >> 5, cl = lane & 31, stream = wave * 2 + half;
    const int r0 = atom_off[a], na = atom_off[a + 1] - r0;
    const int p0 = indptr[r0], len = indptr[r0 + 1] - p0;
    const int cnt = count[k];
    const float2 *hk = h + (int64_t)k * nnz, *sk = s + (int64_t)k * nnz;
    const double2 *vk = V + (int64_t)k * n * cap;
    for (int c0 = 0; c0 < cnt; c0 += 32) {
        const int c = c0 + cl;
        const bool cv = c < cnt;
        const double e = cv ? energies[(int64_t)k * cap + c] : 0.0;
        double ar[MAXR], ai[MAXR];
#pragma unroll
        for (int r = 0; r < MAXR; ++r) ar[r] = ai[r] = 0.0;
        for (int q0 = 0; q0 < len; q0 += CH) {
            const int qc = min(CH, len - q0);
            __syncthreads();
            for (int x = tid; x < qc * na; x += 256) {
                const int pp = x / na, r = x % na;
                const int64_t pos = indptr[r0 + r] + q0 + pp;
                const float2 hvv = hk[pos], svv = sk[pos];
                hv[pp][r] = make_double2(hvv.x, hvv.y);
                sv[pp][r] = make_double2(svv.x, svv.y);
            }
            for (int x = tid; x < qc; x += 256) cidx[x] = indices[p0 + q0 + x];
            __syncthreads();
            double2 xv[CH / NS];
#pragma unroll
            for (int u = 0; u < CH / NS; ++u) {
                const int pp = stream + NS * u;
                xv[u] = (cv && pp < qc) ? vk[(int64_t)cidx[pp] * cap + c] : make_double2(0.0, 0.0);
            }
#pragma unroll
            for (int u = 0; u < CH / NS; ++u) {
                const int pp = stream + NS * u;
                if (pp >= qc) break;
#pragma unroll
                for (int r = 0; r < MAXR; ++r) {
                    if (r < na) {
                        const double2 hh = hv[pp][r], ss = sv[pp][r];
                        const double vr = hh.x - e * ss.x, vi = hh.y - e * ss.y;
                        ar[r] += vr * xv[u].x - vi * xv[u].y;
                        ai[r] += vr * xv[u].y + vi * xv[u].x;
                    }
                }
            }
        }
#pragma unroll
        for (int r = 0; r < MAXR; ++r) {
            ar[r] += __shfl_down(ar[r], 32);
            ai[r] += __shfl_down(ai[r], 32);
        }
        for (int src = 1; src < 4; ++src) {
            __syncthreads();
            if (wave == src && half == 0) {
#pragma unroll
                for (int r = 0; r < MAXR; ++r)
                    if (r < na) red[r][cl] = make_double2(ar[r], ai[r]);
            }
            __syncthreads();
            if (wave == 0 && half == 0) {
#pragma unroll
                for (int r = 0; r < MAXR; ++r)
                    if (r < na) {
                        ar[r] += red[r][cl].x;
                        ai[r] += red[r][cl].y;
                    }
            }
        }
        if (wave == 0 && half == 0 && cv) {
            double local = 0.0;
#pragma unroll
            for (int r = 0; r < MAXR; ++r)
                if (r < na) local += ar[r] * ar[r] + ai[r] * ai[r];
            atomicAdd(norm2 + (int64_t)k * cap + c, local);
        }
    }
}

template __global__ void k_eig_residual_atoms<16>(double *, const double2 *, int, const int32_t *, const double *,
                                                  const float2 *, const float2 *, const int32_t *, const int32_t *,
                                                  const int32_t *, int64_t, int64_t, int, int);
template __global__ void k_eig_residual_atoms<32>(double *, const double2 *, int, const int32_t *, const double *,
                                                  const float2 *, const float2 *, const int32_t *, const int32_t *,
                                                  const int32_t *, int64_t, int64_t, int, int);

}  // namespace aosx
