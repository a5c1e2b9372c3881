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

__device__ __forceinline__ float2 bloch_sum(const float2 *__restrict__ vals, const int32_t *__restrict__ source,
                                            const int32_t *__restrict__ group, const float2 *__restrict__ ph,
                                            int32_t s0, int32_t s1) {
    float2 acc = make_float2(0.f, 0.f);
    for (int32_t q = s0; q < s1; ++q) {
        const float2 v = vals[source[q]], p = ph[group[q]];
        acc.x += v.x * p.x - v.y * p.y;
        acc.y += v.x * p.y + v.y * p.x;
    }
    return acc;
}

// out_h[k][p], out_s[k][p] for nk k points (grid.y = k)
__global__ void __launch_bounds__(256) k_assemble(float2 *__restrict__ out_h, float2 *__restrict__ out_s,
                                                 const float2 *__restrict__ hv, const float2 *__restrict__ sv,
                                                 const int32_t *__restrict__ source,
                                                 const int32_t *__restrict__ group,
                                                 const int32_t *__restrict__ starts,
                                                 const int32_t *__restrict__ mirror,
                                                 const float2 *__restrict__ phase, int n_groups, int64_t nnz) {
    const int k = blockIdx.y;
    const float2 *ph = phase + (int64_t)k * n_groups;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < nnz; p += stride) {
        const int32_t q = mirror[p];
        const float2 h0 = bloch_sum(hv, source, group, ph, starts[p], starts[p + 1]);
        const float2 s0 = bloch_sum(sv, source, group, ph, starts[p], starts[p + 1]);
        float2 hres, sres;
        if (q == p) {
            hres = make_float2(h0.x, 0.f);
            sres = make_float2(s0.x, 0.f);
        } else {
            const float2 h1 = bloch_sum(hv, source, group, ph, starts[q], starts[q + 1]);
            const float2 s1 = bloch_sum(sv, source, group, ph, starts[q], starts[q + 1]);
            hres = make_float2((h0.x + h1.x) * 0.5f, (h0.y - h1.y) * 0.5f);
            sres = make_float2((s0.x + s1.x) * 0.5f, (s0.y - s1.y) * 0.5f);
        }
        out_h[(int64_t)k * nnz + p] = hres;
        out_s[(int64_t)k * nnz + p] = sres;
    }
}

}  // namespace aosx
