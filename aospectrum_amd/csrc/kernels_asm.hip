// Bloch-sum assembly of H(k), S(k) on the shared CSR pattern.
//
// Restates BlochWorkspace.update (src/aospectrum/sparse.py:199-245): for CSR
// entry p, v_p = sum over its source entries of value[source] * phase[group],
// followed by the Hermitian projection of _project_hermitian (sparse.py:25-36),
// which sets v_p <- (v_p + conj(v_mirror(p))) / 2 (the real part on the
// diagonal).  Both sums of a mirror pair are formed by the thread that writes
// p, so the projection needs no second pass.  Phases are formed in FP64 and
// rounded to complex64; sums are complex64 as in the reference.
#include "device_common.h"

namespace aosx {

// exp(2 pi i R . k) in FP64, rounded to complex64
__device__ __forceinline__ float2 bloch_phase(const int32_t *__restrict__ R, const double *k) {
    const double arg = 2.0 * 3.14159265358979323846 * (R[0] * k[0] + R[1] * k[1] + R[2] * k[2]);
    double sn, cs;
    sincos(arg, &sn, &cs);
    return make_float2((float)cs, (float)sn);
}

// phase[k][g] = exp(2 pi i R_g . k)
__global__ void k_phases(float2 *__restrict__ phase, const int32_t *__restrict__ shifts, int n_groups,
                         const double *__restrict__ kpts, int nk) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n_groups * nk) return;
    const int k = e / n_groups, g = e % n_groups;
    phase[e] = bloch_phase(shifts + 3 * g, kpts + 3 * k);
}

// The same phases for at most KBATCH_MAX k points passed by value (no device
// copy of the k points).
constexpr int KBATCH_MAX = 64;
struct KBatch {
    double k[KBATCH_MAX][3];
};
__global__ void k_phases_byval(float2 *__restrict__ phase, const int32_t *__restrict__ shifts, int n_groups,
                               const KBatch kpts, int nk) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n_groups * nk) return;
    const int k = e / n_groups, g = e % n_groups;
    phase[e] = bloch_phase(shifts + 3 * g, kpts.k[k]);
}

// acc += v * p (complex64)
__device__ __forceinline__ void bloch_term(float2 &acc, float2 v, float2 p) {
    acc.x += v.x * p.x - v.y * p.y;
    acc.y += v.x * p.y + v.y * p.x;
}
__device__ __forceinline__ float2 bloch_sum(const float2 *__restrict__ vals, const int32_t *__restrict__ source,
                                            const int32_t *__restrict__ group, const float2 *__restrict__ ph,
                                            int32_t s0, int32_t s1) {
    float2 acc = make_float2(0.f, 0.f);
    for (int32_t q = s0; q < s1; ++q) bloch_term(acc, vals[source[q]], ph[group[q]]);
    return acc;
}
// (v_p + conj(v_q)) / 2, the real part on the diagonal
__device__ __forceinline__ float2 hermitian_mean(float2 v0, float2 v1, bool diag) {
    return diag ? make_float2(v0.x, 0.f) : make_float2((v0.x + v1.x) * 0.5f, (v0.y - v1.y) * 0.5f);
}

// out_h[k][p], out_s[k][p] for nk k points.  A thread takes entry p for every
// k point: it reads the gather lists and source values of p and of its
// mirror once and forms the nk phase sums from the phase table (nk x groups,
// cache-resident), so that the lists and the gathered values cross HBM once
// per launch instead of once per k point.
__global__ void __launch_bounds__(256) k_assemble(float2 *__restrict__ out_h, float2 *__restrict__ out_s,
                                                 const float2 *__restrict__ hv, const float2 *__restrict__ sv,
                                                 const int32_t *__restrict__ source,
                                                 const int32_t *__restrict__ group,
                                                 const int32_t *__restrict__ starts,
                                                 const int32_t *__restrict__ mirror,
                                                 const float2 *__restrict__ phase, int n_groups, int64_t nnz, int nk) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < nnz; p += stride) {
        const int32_t q = mirror[p];
        const bool diag = q == p;
        const int32_t a0 = starts[p], a1 = starts[p + 1];
        const int32_t b0 = diag ? a0 : starts[q], b1 = diag ? a1 : starts[q + 1];
        if (a1 - a0 == 1 && b1 - b0 == 1) {
            // one source entry each (no duplicate lattice shifts): values and
            // groups in registers across the k points
            const int32_t sa = source[a0], sb = source[b0], ga = group[a0], gb = group[b0];
            const float2 ha = hv[sa], xa = sv[sa], hb = hv[sb], xb = sv[sb];
            for (int k = 0; k < nk; ++k) {
                const float2 pa = phase[(int64_t)k * n_groups + ga], pb = phase[(int64_t)k * n_groups + gb];
                float2 h0 = make_float2(0.f, 0.f), s0 = h0, h1 = h0, s1 = h0;
                bloch_term(h0, ha, pa);
                bloch_term(s0, xa, pa);
                bloch_term(h1, hb, pb);
                bloch_term(s1, xb, pb);
                out_h[(int64_t)k * nnz + p] = hermitian_mean(h0, h1, diag);
                out_s[(int64_t)k * nnz + p] = hermitian_mean(s0, s1, diag);
            }
        } else {
            for (int k = 0; k < nk; ++k) {
                const float2 *ph = phase + (int64_t)k * n_groups;
                const float2 h0 = bloch_sum(hv, source, group, ph, a0, a1);
                const float2 s0 = bloch_sum(sv, source, group, ph, a0, a1);
                const float2 h1 = bloch_sum(hv, source, group, ph, b0, b1);
                const float2 s1 = bloch_sum(sv, source, group, ph, b0, b1);
                out_h[(int64_t)k * nnz + p] = hermitian_mean(h0, h1, diag);
                out_s[(int64_t)k * nnz + p] = hermitian_mean(s0, s1, diag);
            }
        }
    }
}

}  // namespace aosx
