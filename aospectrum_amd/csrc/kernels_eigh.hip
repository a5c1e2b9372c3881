// Batched Hermitian eigendecomposition of the small FP64 problems of the CISS
// subspace stage (the compressed moment Gram and the projected Rayleigh-Ritz
// matrix, r <= 256): cyclic two-sided Jacobi with round-robin ordering.
//
// Round t of a sweep pairs the rr = r + (r odd) indices into rr/2 disjoint
// pairs (the circle method: index rr-1 stays, the others rotate).  Every pair
// (p, q), p < q, gets the rotation M = [[c, s], [-s e^{-i phi}, c e^{-i phi}]]
// that zeroes A(p, q) of the 2 x 2 block, and A <- M^H A M, V <- V M for all
// pairs at once.  The work is split over two kernels:
//   k_jacobi_rotate   one workgroup per problem.  A is kept as its packed upper
//                     triangle -- in LDS when rr <= 128 (132 KiB), else in a
//                     global scratch -- and each round updates every 2 x 2
//                     pair block B(a, b) <- M_a^H B M_b once (rows and columns
//                     together, each element read and written once): the
//                     diagonal blocks by the threads that compute the
//                     rotations, the off-diagonal ones by the whole
//                     workgroup.  The rotations are logged; the eigenvalues
//                     come out ascending with the rank of every column.
//   k_jacobi_vectors  rows of V are independent under column rotations, so V
//                     is rebuilt from the log one row per wave, held in LDS;
//                     the log is staged 8 rounds at a time, the next batch
//                     loading while one is applied.
#include "device_common.h"

namespace aosx {

constexpr int EIGH_LDS_MAX = 128;  // rr up to which A lives in LDS
constexpr int EIGH_MAX = 256;
constexpr int EIGH_SWEEPS = 40;

struct Rot {
    double c, s, er, ei;  // e^{i phi} = er + i ei
};

// pair a of round t (t < rr - 1, a < rr / 2), p < q:
//   a = 0: (rr - 1, t);  a > 0: ((t + a) mod (rr - 1), (t - a) mod (rr - 1))
__device__ __forceinline__ void rr_pair(int rr, int t, int a, int &p, int &q) {
    const int m = rr - 1;
    int x = m, y = t;
    if (a > 0) {
        x = t + a >= m ? t + a - m : t + a;
        y = t - a < 0 ? t - a + m : t - a;
    }
    p = min(x, y);
    q = max(x, y);
}

__device__ __forceinline__ int packed(int rr, int i, int j) {  // i <= j
    return i * rr - ((i * (i - 1)) >> 1) + (j - i);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

template <bool IN_LDS>
__global__ void __launch_bounds__(1024) k_jacobi_rotate(const c64 *__restrict__ A_all, int ld, int64_t stride,
                                                        const int32_t *__restrict__ rk, c64 *__restrict__ scratch,
                                                        Rot *__restrict__ log, int64_t log_stride,
                                                        int32_t *__restrict__ rounds_out, double *__restrict__ w_all,
                                                        int32_t *__restrict__ rank_all, int max_sweeps, double tol) {
    __shared__ c64 Sl[IN_LDS ? EIGH_LDS_MAX * (EIGH_LDS_MAX + 1) / 2 : 1];
    __shared__ uint8_t blk_a[IN_LDS ? 2080 : 8256], blk_b[IN_LDS ? 2080 : 8256];
    __shared__ Rot rot[EIGH_MAX / 2];
    __shared__ int16_t rp[EIGH_MAX / 2], rq[EIGH_MAX / 2];
    __shared__ double part[2][16];
    __shared__ int done;
    __shared__ double skip2;  // |A(p,q)|^2 below which a pair takes no rotation
    const int k = blockIdx.x, tid = threadIdx.x, nthr = blockDim.x;
    const int r = rk[k];
    const int rr = r + (r & 1), npairs = rr / 2, nrounds = rr - 1;
    const int nblocks = npairs * (npairs - 1) / 2;  // off-diagonal pair blocks
    c64 *S = IN_LDS ? Sl : scratch + (int64_t)k * (EIGH_MAX * (EIGH_MAX + 1) / 2);
    const c64 *A = A_all + k * stride;
    Rot *lg = log + (int64_t)k * log_stride;
    // packed upper triangle; the padding index (odd r) is a zero row/column
    for (int e = tid; e < rr * rr; e += nthr) {
        const int i = e / rr, j = e % rr;
        if (j < i) continue;
        S[packed(rr, i, j)] = (i < r && j < r) ? A[(int64_t)i * ld + j] : c64{0.0, 0.0};
    }
    for (int a = tid; a < npairs; a += nthr) {
        const int base = a * (npairs - 1) - ((a * (a - 1)) >> 1);  // blocks (a', b), a' < a, b > a'
        for (int b = a + 1; b < npairs; ++b) {
            blk_a[base + b - a - 1] = (uint8_t)a;
            blk_b[base + b - a - 1] = (uint8_t)b;
        }
    }
    __syncthreads();
    // element (i, j) of the Hermitian matrix, and its store, through the
    // packed upper triangle (address and conjugation by selects, no branches)
    auto at = [&](int i, int j) -> c64 {
        const c64 v = S[packed(rr, min(i, j), max(i, j))];
        return {v.re, i <= j ? v.im : -v.im};
    };
    auto put = [&](int i, int j, c64 v) { S[packed(rr, min(i, j), max(i, j))] = {v.re, i <= j ? v.im : -v.im}; };
    // pair block B = [[A(pa,pb), A(pa,qb)], [A(qa,pb), A(qa,qb)]] <- M_a^H B M_b
    auto rotate = [](const Rot &ga, const Rot &gb, c64 &b00, c64 &b01, c64 &b10, c64 &b11) {
        // B M_b: col0 = c B0 - s e^{-i phi} B1, col1 = s B0 + c e^{-i phi} B1
        const c64 eb = {gb.er, -gb.ei};
        const c64 x0 = cmul(eb, b01), x1 = cmul(eb, b11);
        const c64 n00 = {gb.c * b00.re - gb.s * x0.re, gb.c * b00.im - gb.s * x0.im};
        const c64 n01 = {gb.s * b00.re + gb.c * x0.re, gb.s * b00.im + gb.c * x0.im};
        const c64 n10 = {gb.c * b10.re - gb.s * x1.re, gb.c * b10.im - gb.s * x1.im};
        const c64 n11 = {gb.s * b10.re + gb.c * x1.re, gb.s * b10.im + gb.c * x1.im};
        // M_a^H X: row0 = c X0 - s e^{i phi} X1, row1 = s X0 + c e^{i phi} X1
        const c64 ea = {ga.er, ga.ei};
        const c64 y0 = cmul(ea, n10), y1 = cmul(ea, n11);
        b00 = {ga.c * n00.re - ga.s * y0.re, ga.c * n00.im - ga.s * y0.im};
        b01 = {ga.c * n01.re - ga.s * y1.re, ga.c * n01.im - ga.s * y1.im};
        b10 = {ga.s * n00.re + ga.c * y0.re, ga.s * n00.im + ga.c * y0.im};
        b11 = {ga.s * n01.re + ga.c * y1.re, ga.s * n01.im + ga.c * y1.im};
    };
    int sweep = 0;
    for (; sweep < max_sweeps && r > 1; ++sweep) {
        double off = 0.0, tot = 0.0;
        for (int e = tid; e < rr * rr; e += nthr) {
            const int i = e / rr, j = e % rr;
            if (j < i) continue;
            const double v = cabs2(S[packed(rr, i, j)]);
            tot += i == j ? v : 2.0 * v;
            off += i == j ? 0.0 : 2.0 * v;
        }
        off = wave_sum(off);
        tot = wave_sum(tot);
        if ((tid & 63) == 0) {
            part[0][tid >> 6] = off;
            part[1][tid >> 6] = tot;
        }
        __syncthreads();
        if (tid == 0) {
            double o = 0.0, s = 0.0;
            for (int q = 0; q < nthr / 64; ++q) {
                o += part[0][q];
                s += part[1][q];
            }
            done = !(o > tol * tol * s);
            // threshold Jacobi: an element with |A(p,q)|^2 < tol^2 ||A||_F^2 / r^2
            // is left alone; once every element is below that, the
            // off-diagonal norm is below the stopping bound, so the skipped
            // rotations never hold the loop open, and the late sweeps update
            // few pair blocks
            skip2 = tol * tol * s / ((double)r * (double)r);
        }
        __syncthreads();
        if (done) break;
        for (int t = 0; t < nrounds; ++t) {
            if (tid < npairs) {
                int p, q;
                rr_pair(rr, t, tid, p, q);
                Rot g = {1.0, 0.0, 1.0, 0.0};
                if (q < r) {
                    const c64 b = S[packed(rr, p, q)];
                    const double mb = sqrt(cabs2(b));
                    const double app = S[packed(rr, p, p)].re, aqq = S[packed(rr, q, q)].re;
                    if (mb > 1e-300 && mb * mb >= skip2 && mb > 1e-18 * sqrt(fabs(app * aqq))) {
                        const double tau = (aqq - app) / (2.0 * mb);
                        const double tt = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
                        g.c = 1.0 / sqrt(1.0 + tt * tt);
                        g.s = tt * g.c;
                        g.er = b.re / mb;
                        g.ei = b.im / mb;
                    }
                }
                rot[tid] = g;
                rp[tid] = (int16_t)p;
                rq[tid] = (int16_t)q;
                lg[((int64_t)sweep * nrounds + t) * npairs + tid] = g;
                // the pair's own diagonal block (p,p), (p,q), (q,q): no other
                // thread touches it this round
                if (g.s != 0.0) {
                    c64 b00 = S[packed(rr, p, p)], b01 = S[packed(rr, p, q)], b10 = cconj(b01),
                        b11 = S[packed(rr, q, q)];
                    rotate(g, g, b00, b01, b10, b11);
                    S[packed(rr, p, p)] = {b00.re, 0.0};
                    S[packed(rr, p, q)] = b01;
                    S[packed(rr, q, q)] = {b11.re, 0.0};
                }
            }
            __syncthreads();
            // the off-diagonal pair blocks, each read and written once
            for (int e = tid; e < nblocks; e += nthr) {
                const int a = blk_a[e], b = blk_b[e];
                const Rot ga = rot[a], gb = rot[b];
                if (ga.s == 0.0 && gb.s == 0.0) continue;
                const int pa = rp[a], qa = rq[a], pb = rp[b], qb = rq[b];
                c64 b00 = at(pa, pb), b01 = at(pa, qb), b10 = at(qa, pb), b11 = at(qa, qb);
                rotate(ga, gb, b00, b01, b10, b11);
                put(pa, pb, b00);
                put(pa, qb, b01);
                put(qa, pb, b10);
                put(qa, qb, b11);
            }
            __syncthreads();
        }
    }
    if (tid == 0) rounds_out[k] = sweep * nrounds;
    // eigenvalues ascending by rank counting (stable); rank of every column
    double *w = w_all + (int64_t)k * ld;
    for (int i = tid; i < r; i += nthr) {
        const double wi = S[packed(rr, i, i)].re;
        int rank = 0;
        for (int j = 0; j < r; ++j) {
            const double wj = S[packed(rr, j, j)].re;
            rank += (wj < wi) || (wj == wi && j < i);
        }
        w[rank] = wi;
        rank_all[(int64_t)k * EIGH_MAX + i] = rank;
    }
}

constexpr int EIGH_ROWS = 4;   // rows of V per workgroup: one per wave
constexpr int EIGH_BATCH = 8;  // logged rounds staged per batch (of 128 pairs at most)

// Vout[k][row][rank[j]] = V[row][j] for rows/columns < r, V = product of the
// logged rotations applied to the identity.  Rows of V are independent under
// column rotations: wave w rebuilds row 4 * blockIdx.x + w in LDS, lane l
// applying the rotations of pairs l and l + 64 of every round.  The pairs of
// a round are disjoint, and the rounds of a row run in order within its wave
// (LDS operations of one wave complete in issue order), so the workgroup
// synchronises once per batch of logged rounds.  Batches alternate between
// two LDS buffers (two objects, so that the compiler sees that reads of one
// never wait for the loads into the other): the next batch moves from memory
// straight into the other buffer while one is applied.
__global__ void __launch_bounds__(256) k_jacobi_vectors(c64 *__restrict__ Vout_all, int ld, int64_t stride,
                                                        const int32_t *__restrict__ rk,
                                                        const Rot *__restrict__ log, int64_t log_stride,
                                                        const int32_t *__restrict__ rounds_all,
                                                        const int32_t *__restrict__ rank_all) {
    __shared__ c64 V[EIGH_ROWS][EIGH_MAX];
    __shared__ __attribute__((aligned(1024))) Rot R0[EIGH_BATCH * (EIGH_MAX / 2)];  // u * npairs + a
    __shared__ __attribute__((aligned(1024))) Rot R1[EIGH_BATCH * (EIGH_MAX / 2)];
    const int k = blockIdx.y, tid = threadIdx.x, wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
    const int r = rk[k];
    const int row0 = blockIdx.x * EIGH_ROWS;
    if (row0 >= r) return;
    const int rr = r + (r & 1), npairs = rr / 2, nrounds = rr - 1;
    const int rounds = rounds_all[k];
    const Rot *lg = log + (int64_t)k * log_stride;
    const int64_t log_bytes = (int64_t)rounds * npairs * (int64_t)sizeof(Rot);
    const int batch_bytes = EIGH_BATCH * npairs * (int)sizeof(Rot);
    // batch b of the log into `dst`, 1 KiB per wave instruction; a buffer
    // load past the log's end returns zeros (those entries are never read)
    auto fetch = [&](int b, Rot *dst) {
        const int64_t off = (int64_t)b * batch_bytes;
        const int bytes = (int)min<int64_t>(batch_bytes, log_bytes - off);
        const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<char *>(reinterpret_cast<const char *>(lg) + off), 0, bytes, 0x00020000);
        for (int c = wave; c * 1024 < bytes; c += EIGH_ROWS)
            __builtin_amdgcn_raw_ptr_buffer_load_lds(rsrc, reinterpret_cast<char *>(dst) + c * 1024, 16, lane * 16,
                                                     c * 1024, 0, 0);
    };
    c64 *v = V[wave];
    int t = 0;  // round within the sweep
    auto apply = [&](int b, const Rot *Rb) {
        const int nb = min(EIGH_BATCH, rounds - b * EIGH_BATCH);
        for (int u = 0; u < nb; ++u) {
            for (int a = lane; a < npairs; a += 64) {
                int p, q;
                rr_pair(rr, t, a, p, q);
                // the rotation and both elements are read together and written
                // back unconditionally (unchanged under an identity rotation):
                // one LDS round trip per round
                const Rot g = Rb[u * npairs + a];
                const c64 xp = v[p], vq = v[q];
                const c64 yq = cmul({g.er, -g.ei}, vq);
                const bool rot = g.s != 0.0;
                const double pr = g.c * xp.re - g.s * yq.re, pi = g.c * xp.im - g.s * yq.im;
                const double qr = g.s * xp.re + g.c * yq.re, qi = g.s * xp.im + g.c * yq.im;
                v[p] = {rot ? pr : xp.re, rot ? pi : xp.im};
                v[q] = {rot ? qr : vq.re, rot ? qi : vq.im};
            }
            // this round's writes land before any lane of the wave reads for the next
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            t = t + 1 == nrounds ? 0 : t + 1;
        }
    };
    // every wave's loads of a batch have landed (each waits for its own)
    // before the barrier that opens the batch
    constexpr int WAIT_VMEM = 0x0F70;  // s_waitcnt vmcnt(0)
    for (int e = tid; e < EIGH_ROWS * rr; e += 256) {
        const int lr = e / rr, j = e % rr;
        V[lr][j] = {row0 + lr == j ? 1.0 : 0.0, 0.0};
    }
    const int nbatch = (rounds + EIGH_BATCH - 1) / EIGH_BATCH;
    if (nbatch > 0) fetch(0, R0);
    __builtin_amdgcn_s_waitcnt(WAIT_VMEM);
    __syncthreads();
    for (int b = 0; b < nbatch; b += 2) {
        if (b + 1 < nbatch) fetch(b + 1, R1);  // R1 was last read in batch b - 1
        apply(b, R0);
        __builtin_amdgcn_s_waitcnt(WAIT_VMEM);
        __syncthreads();
        if (b + 1 >= nbatch) break;
        if (b + 2 < nbatch) fetch(b + 2, R0);
        apply(b + 1, R1);
        __builtin_amdgcn_s_waitcnt(WAIT_VMEM);
        __syncthreads();
    }
    c64 *Vout = Vout_all + k * stride;
    const int32_t *rank = rank_all + (int64_t)k * EIGH_MAX;
    for (int e = tid; e < EIGH_ROWS * r; e += 256) {
        const int lr = e / r, j = e % r;
        if (row0 + lr < r) Vout[(int64_t)(row0 + lr) * ld + rank[j]] = V[lr][j];
    }
}

}  // namespace aosx
