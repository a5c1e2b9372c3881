// extern "C" entry points of libaosx_hip.so (declared in include/aosx.h).
//
// Request validation mirrors validate_request of the reference's ubatch call
// (native/src/aospectrum_cudss_complex64_ubatch.c:136-241): ABI version and
// struct size first, then dimensions, null pointers and finiteness, each with
// the same status number.  Exceptions never cross the C boundary.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>

#include "bundle.h"

using aosx::Bundle;
using aosx::Error;

struct aosx_bundle {
    std::unique_ptr<Bundle> impl;
};

namespace {

double now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

template <typename R>
void begin(R *res) {
    std::memset(res, 0, sizeof(*res));
    res->abi_version = AOSX_ABI_VERSION;
    res->struct_size = (uint32_t)sizeof(*res);
}

template <typename R>
int fail(R *res, int status, int hip_status, const char *fmt, ...) {
    res->status = status;
    res->hip_status = hip_status;
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(res->error_message, AOSX_ERROR_CAPACITY, fmt, ap);
    va_end(ap);
    return 1;
}

template <typename Q>
bool abi_ok(const Q *q) {
    return q->abi_version == AOSX_ABI_VERSION && q->struct_size == sizeof(Q);
}

template <typename R, typename F>
int guarded(R *res, F &&body) {
    try {
        return body();
    } catch (const Error &e) {
        return fail(res, e.code, e.hip_status, "%s", e.what());
    } catch (const std::bad_alloc &) {
        return fail(res, AOSX_SIZE_OVERFLOW, 0, "host allocation failed");
    } catch (const std::exception &e) {
        return fail(res, AOSX_HIP_ERROR, 0, "%s", e.what());
    }
}

bool finite_c64(const aosx_complex64 *v, int64_t count) {
    for (int64_t i = 0; i < count; ++i)
        if (!std::isfinite(v[i].real) || !std::isfinite(v[i].imag)) return false;
    return true;
}

}  // namespace

extern "C" {

int aosx_bundle_create_v1(const aosx_bundle_request_v1 *q, aosx_bundle_result_v1 *res, aosx_bundle **out) {
    if (!q || !res) return 1;
    begin(res);
    if (out) *out = nullptr;
    if (!abi_ok(q)) return fail(res, AOSX_ABI_MISMATCH, 0, "request ABI mismatch");
    if (!out || !q->atom_offsets || !q->row_atom || !q->col_atom || !q->lattice_shift || !q->block_offsets ||
        !q->h_values || !q->s_values)
        return fail(res, AOSX_NULL_POINTER, 0, "required input or output pointer is null");
    if (q->n_orbitals <= 0 || q->n_orbitals > INT32_MAX || q->n_atoms <= 0 || q->n_blocks <= 0)
        return fail(res, AOSX_INT32_OVERFLOW, 0, "bundle dimensions exceed the int32 CSR ABI");
    if (q->value_kind != 0 && q->value_kind != 1) return fail(res, AOSX_INVALID_BLOCKS, 0, "value_kind must be 0 or 1");
    if (q->tile_size != 0 && q->tile_size != 64) return fail(res, AOSX_INVALID_SETTINGS, 0, "tile_size must be 0 or 64");
    if (q->ordering < 0 || q->ordering > 3) return fail(res, AOSX_INVALID_SETTINGS, 0, "ordering must be 0..3");
    if (q->ordering == AOSX_ORDERING_NESTED_DISSECTION && !q->atom_positions)
        return fail(res, AOSX_NULL_POINTER, 0, "nested dissection needs atom positions");
    return guarded(res, [&]() -> int {
        auto b = std::make_unique<Bundle>();
        aosx::BlockTable bt;
        bt.n_orbitals = q->n_orbitals;
        bt.n_atoms = q->n_atoms;
        bt.n_blocks = q->n_blocks;
        bt.atom_offsets = q->atom_offsets;
        bt.row_atom = q->row_atom;
        bt.col_atom = q->col_atom;
        bt.lattice_shift = q->lattice_shift;
        bt.block_offsets = q->block_offsets;
        const int64_t n_values = q->block_offsets[q->n_blocks];
        if (n_values <= 0 || n_values > INT32_MAX)
            return fail(res, AOSX_INT32_OVERFLOW, 0, "block value count exceeds the int32 index range");
        double t0 = now();
        int code = 0;
        std::string why = aosx::build_pattern(bt, b->pat, code);
        if (!why.empty()) return fail(res, code, 0, "%s", why.c_str());
        if ((int64_t)b->pat.source.size() > INT32_MAX)
            return fail(res, AOSX_INT32_OVERFLOW, 0, "gather list exceeds the int32 index range");
        // non-finite block values are rejected up front (status 35)
        if (q->value_kind == 0) {
            const float *h = static_cast<const float *>(q->h_values), *s = static_cast<const float *>(q->s_values);
            for (int64_t i = 0; i < n_values; ++i)
                if (!std::isfinite(h[i]) || !std::isfinite(s[i]))
                    return fail(res, AOSX_NONFINITE_VALUES, 0, "operator blocks contain non-finite values");
        } else if (!finite_c64(static_cast<const aosx_complex64 *>(q->h_values), n_values) ||
                   !finite_c64(static_cast<const aosx_complex64 *>(q->s_values), n_values)) {
            return fail(res, AOSX_NONFINITE_VALUES, 0, "operator blocks contain non-finite values");
        }
        double t1 = now();
        std::vector<int32_t> order, groups;
        int64_t atom_fill = 0;
        const bool use_nd = q->ordering == AOSX_ORDERING_NESTED_DISSECTION ||
                            (q->ordering == AOSX_ORDERING_AUTO && q->atom_positions);
        if (use_nd) {
            for (int64_t i = 0; i < 3 * q->n_atoms; ++i)
                if (!std::isfinite(q->atom_positions[i]))
                    return fail(res, AOSX_NONFINITE_VALUES, 0, "atom positions are not finite");
            why = aosx::order_atoms_nd(bt, q->atom_positions, order, groups);
            if (why.empty()) atom_fill = aosx::atom_fill(bt, order);
        } else {
            why = aosx::order_atoms(bt, q->ordering == AOSX_ORDERING_NATURAL ? 1 : 0, order, atom_fill);
        }
        if (!why.empty()) return fail(res, AOSX_INVALID_BLOCKS, 0, "%s", why.c_str());
        double t2 = now();
        why = aosx::build_tiles(bt, b->pat, order, groups, 64, b->ts);
        if (!why.empty()) return fail(res, AOSX_INT32_OVERFLOW, 0, "%s", why.c_str());
        if (!groups.empty()) {
            // node-aligned tiles pay off when the nodes span several tiles;
            // on small bundles the padding can cost more than it saves
            aosx::TileStructure packed;
            why = aosx::build_tiles(bt, b->pat, order, {}, 64, packed);
            if (!why.empty()) return fail(res, AOSX_INT32_OVERFLOW, 0, "%s", why.c_str());
            if (packed.factor_flops < b->ts.factor_flops) b->ts = std::move(packed);
        }
        b->ts.atom_fill_entries = atom_fill;
        double t3 = now();
        b->n = q->n_orbitals;
        b->nnz = b->pat.nnz();
        b->n_groups = b->pat.n_groups();
        res->n = b->n;
        res->nnz = b->nnz;
        res->n_phase_shifts = b->n_groups;
        res->n_diagonal_tiles = b->ts.n_diag;
        res->n_factor_tiles = b->ts.n_tiles;
        res->atom_fill_entries = atom_fill;
        res->tile_size = 64;
        res->factor_flops = b->ts.factor_flops;
        res->schur_flops = b->ts.schur_flops;
        res->factor_bytes = (uint64_t)b->ts.n_tiles * 2 * 4 * 64 * 64 + (uint64_t)b->ts.n_diag * 4 * 4 * 64 * 64;
        res->pattern_seconds = t1 - t0;
        res->ordering_seconds = t2 - t1;
        res->symbolic_seconds = t3 - t2;
        if (q->device >= 0) {
            int count = 0;
            if (hipGetDeviceCount(&count) != hipSuccess || q->device >= count) {
                (void)hipGetLastError();
                return fail(res, AOSX_NO_DEVICE, 0, "HIP device %d is not available", q->device);
            }
            b->device = q->device;
            b->upload(q->h_values, q->s_values, q->value_kind, n_values, q->atom_offsets, q->n_atoms);
            res->resident_device_bytes = b->resident_bytes;
        }
        res->upload_seconds = now() - t3;
        aosx_bundle *handle = new aosx_bundle;
        handle->impl = std::move(b);
        *out = handle;
        return 0;
    });
}

int aosx_bundle_pattern_v1(const aosx_bundle *bundle, int32_t *indptr, int32_t *indices, int32_t *permutation,
                           aosx_status_v1 *res) {
    if (!res) return 1;
    begin(res);
    if (!bundle) return fail(res, AOSX_NULL_POINTER, 0, "bundle handle is null");
    const Bundle &b = *bundle->impl;
    if (indptr) std::memcpy(indptr, b.pat.indptr.data(), b.pat.indptr.size() * 4);
    if (indices) std::memcpy(indices, b.pat.indices.data(), b.pat.indices.size() * 4);
    if (permutation) {  // the orbitals in elimination order (padding rows skipped)
        int64_t q = 0;
        for (int32_t v : b.ts.perm)
            if (v >= 0) permutation[q++] = v;
    }
    return 0;
}

int aosx_bundle_destroy_v1(aosx_bundle *bundle) {
    delete bundle;
    return 0;
}

int aosx_bundle_trim_v1(aosx_bundle *bundle, uint64_t *resident_bytes) {
    if (!bundle) return 1;
    Bundle &b = *bundle->impl;
    if (b.device >= 0) {
        if (hipSetDevice(b.device) != hipSuccess || (b.stream && hipStreamSynchronize(b.stream) != hipSuccess)) {
            (void)hipGetLastError();
            return 1;
        }
        b.release_workspace();
        b.peak_bytes = b.resident_bytes;
    }
    if (resident_bytes) *resident_bytes = b.resident_bytes;
    return 0;
}

int aosx_assemble_v1(aosx_bundle *bundle, const aosx_assemble_request_v1 *q, aosx_status_v1 *res) {
    if (!q || !res) return 1;
    begin(res);
    if (!abi_ok(q)) return fail(res, AOSX_ABI_MISMATCH, 0, "request ABI mismatch");
    if (!bundle || !q->k_fractional || !q->h_data || !q->s_data)
        return fail(res, AOSX_NULL_POINTER, 0, "required input or output pointer is null");
    Bundle &b = *bundle->impl;
    if (b.device < 0) return fail(res, AOSX_NO_DEVICE, 0, "bundle was created without a device");
    if (q->n_kpoints <= 0 || q->n_kpoints > 65535) return fail(res, AOSX_BATCH_DIMENSIONS, 0, "invalid k-point count");
    for (int64_t i = 0; i < 3 * q->n_kpoints; ++i)
        if (!std::isfinite(q->k_fractional[i])) return fail(res, AOSX_NONFINITE_VALUES, 0, "k point is not finite");
    return guarded(res, [&]() -> int {
        double t0 = now();
        AOSX_HIP(hipSetDevice(b.device));
        float2 *d_h = b.buf<float2>("H", (size_t)q->n_kpoints * b.nnz);
        float2 *d_s = b.buf<float2>("S", (size_t)q->n_kpoints * b.nnz);
        b.assemble(q->k_fractional, q->n_kpoints, d_h, d_s);
        AOSX_HIP(hipMemcpyAsync(q->h_data, d_h, (size_t)q->n_kpoints * b.nnz * 8, hipMemcpyDeviceToHost, b.stream));
        AOSX_HIP(hipMemcpyAsync(q->s_data, d_s, (size_t)q->n_kpoints * b.nnz * 8, hipMemcpyDeviceToHost, b.stream));
        AOSX_HIP(hipStreamSynchronize(b.stream));
        res->seconds = now() - t0;
        return 0;
    });
}

int aosx_assemble_noalloc_v1(aosx_bundle *bundle, const aosx_assemble_noalloc_request_v1 *q, aosx_status_v1 *res) {
    if (!q || !res) return 1;
    begin(res);
    if (!abi_ok(q)) return fail(res, AOSX_ABI_MISMATCH, 0, "request ABI mismatch");
    if (!bundle || !q->k_fractional || !q->h_out || !q->s_out)
        return fail(res, AOSX_NULL_POINTER, 0, "required input or output pointer is null");
    Bundle &b = *bundle->impl;
    if (b.device < 0) return fail(res, AOSX_NO_DEVICE, 0, "bundle was created without a device");
    if (q->n_kpoints < 1 || q->n_kpoints > 64)
        return fail(res, AOSX_BATCH_DIMENSIONS, 0, "n_kpoints must be in [1, 64] for the allocation-free entry point");
    if ((uint64_t)q->n_kpoints * (uint64_t)b.n_groups * 8u > AOSX_NOALLOC_SCRATCH_BYTES)
        return fail(res, AOSX_BATCH_DIMENSIONS, 0,
                    "n_kpoints x n_phase_shifts x 8 bytes exceeds the %u-byte scratch array",
                    (unsigned)AOSX_NOALLOC_SCRATCH_BYTES);
    for (int64_t i = 0; i < 3 * q->n_kpoints; ++i)
        if (!std::isfinite(q->k_fractional[i])) return fail(res, AOSX_NONFINITE_VALUES, 0, "k point is not finite");
    return guarded(res, [&]() -> int {
        double t0 = now();
        AOSX_HIP(hipSetDevice(b.device));
        // the outputs must be device allocations on the bundle's device large
        // enough for n_kpoints x nnz values
        const size_t need = (size_t)q->n_kpoints * b.nnz * sizeof(aosx_complex64);
        for (void *out : {(void *)q->h_out, (void *)q->s_out}) {
            hipPointerAttribute_t attr;
            if (hipPointerGetAttributes(&attr, out) != hipSuccess || attr.type != hipMemoryTypeDevice ||
                attr.device != b.device) {
                (void)hipGetLastError();
                return fail(res, AOSX_NULL_POINTER, 0, "h_out and s_out must be device memory on device %d",
                            b.device);
            }
            hipDeviceptr_t base = nullptr;
            size_t size = 0;
            if (hipMemGetAddressRange(&base, &size, out) != hipSuccess ||
                (char *)out + need > (char *)base + size) {
                (void)hipGetLastError();
                return fail(res, AOSX_BATCH_DIMENSIONS, 0,
                            "output buffer holds fewer than n_kpoints x nnz complex64 values");
            }
        }
        b.assemble_noalloc(q->k_fractional, (int)q->n_kpoints, reinterpret_cast<float2 *>(q->h_out),
                           reinterpret_cast<float2 *>(q->s_out));
        res->seconds = now() - t0;
        return 0;
    });
}

int aosx_shift_solve_v1(aosx_bundle *bundle, const aosx_shift_request_v1 *q, aosx_shift_result_v1 *res) {
    if (!q || !res) return 1;
    begin(res);
    if (!abi_ok(q)) return fail(res, AOSX_ABI_MISMATCH, 0, "request ABI mismatch");
    if (q->batch_count <= 0 || q->batch_count > 64 || q->rhs_count <= 0 || q->rhs_count > 256)
        return fail(res, AOSX_BATCH_DIMENSIONS, 0, "invalid uniform-batch dimensions");
    if (!bundle || !q->h_data || !q->s_data || !q->shifts || !q->rhs || !q->solution)
        return fail(res, AOSX_NULL_POINTER, 0, "required input or output pointer is null");
    Bundle &b = *bundle->impl;
    if (b.device < 0) return fail(res, AOSX_NO_DEVICE, 0, "bundle was created without a device");
    if (q->precision < 0 || q->precision > 3 || q->refinement_steps < 0 || q->refinement_steps > 16)
        return fail(res, AOSX_INVALID_SETTINGS, 0, "invalid precision or refinement steps");
    const int64_t n = b.n, nnz = b.nnz;
    if (!finite_c64(q->h_data, nnz) || !finite_c64(q->s_data, nnz) || !finite_c64(q->shifts, q->batch_count))
        return fail(res, AOSX_NONFINITE_VALUES, 0, "shifted matrix batch contains non-finite values");
    if (!finite_c64(q->rhs, n * q->rhs_count))
        return fail(res, AOSX_NONFINITE_RHS, 0, "right-hand-side batch contains non-finite values");
    return guarded(res, [&]() -> int {
        double t0 = now();
        AOSX_HIP(hipSetDevice(b.device));
        const int nb = q->batch_count, w = q->rhs_count, W = (w + 31) / 32 * 32;
        float2 *d_h = b.buf<float2>("H", (size_t)nnz);
        float2 *d_s = b.buf<float2>("S", (size_t)nnz);
        AOSX_HIP(hipMemcpyAsync(d_h, q->h_data, nnz * 8, hipMemcpyHostToDevice, b.stream));
        AOSX_HIP(hipMemcpyAsync(d_s, q->s_data, nnz * 8, hipMemcpyHostToDevice, b.stream));
        std::vector<float2> rhs((size_t)n * W, make_float2(0.f, 0.f));
        for (int64_t r = 0; r < n; ++r)
            for (int c = 0; c < w; ++c)
                rhs[(size_t)r * W + c] = make_float2(q->rhs[r * w + c].real, q->rhs[r * w + c].imag);
        float2 *d_rhs = b.buf<float2>("x_rhs", (size_t)n * W);
        AOSX_HIP(hipMemcpyAsync(d_rhs, rhs.data(), rhs.size() * 8, hipMemcpyHostToDevice, b.stream));
        AOSX_HIP(hipStreamSynchronize(b.stream));
        double t1 = now();
        std::vector<int32_t> kidx(nb, 0);
        std::vector<float2> z(nb);
        for (int i = 0; i < nb; ++i) z[i] = make_float2(q->shifts[i].real, q->shifts[i].imag);
        float *tiles = b.buf<float>("tiles", (size_t)nb * b.factor_floats());
        float *dinv = b.buf<float>("dinv", (size_t)nb * b.ts.n_diag * 4 * 64 * 64);
        int *d_repl = b.buf<int>("replaced", (size_t)std::max(nb, 256));
        b.zero(d_repl, (size_t)(nb + 3) / 4 * 16);
        b.factor(nb, kidx, z, q->precision, tiles, dinv, d_repl);
        AOSX_HIP(hipStreamSynchronize(b.stream));
        double t2 = now();
        float2 *d_x = b.buf<float2>("x_out", (size_t)nb * n * W);
        const bool ir = q->precision != AOSX_PRECISION_FP32;
        double rel = 0.0;
        b.solve_full(nb, W, d_rhs, kidx, d_h, d_s, kidx, z, tiles, dinv, false, ir ? q->refinement_steps : 0, d_x,
                     &rel, q->precision);
        AOSX_HIP(hipStreamSynchronize(b.stream));
        double t3 = now();
        std::vector<float2> x((size_t)nb * n * W);
        // copies on the bundle's stream, not the legacy one (which another
        // thread's graph capture would refuse)
        AOSX_HIP(hipMemcpyAsync(x.data(), d_x, x.size() * 8, hipMemcpyDeviceToHost, b.stream));
        AOSX_HIP(hipStreamSynchronize(b.stream));
        for (int m = 0; m < nb; ++m)
            for (int64_t r = 0; r < n; ++r)
                for (int c = 0; c < w; ++c) {
                    const float2 v = x[((size_t)m * n + r) * W + c];
                    q->solution[((size_t)m * n + r) * w + c] = {v.x, v.y};
                }
        std::vector<int> repl(nb);
        AOSX_HIP(hipMemcpyAsync(repl.data(), d_repl, nb * sizeof(int), hipMemcpyDeviceToHost, b.stream));
        AOSX_HIP(hipStreamSynchronize(b.stream));
        for (int v : repl) res->replaced_pivots += v;
        res->refinement_steps = ir ? q->refinement_steps : 0;
        res->relative_residual = rel;
        res->upload_seconds = t1 - t0;
        res->factorization_seconds = t2 - t1;
        res->solve_seconds = t3 - t2;
        res->download_seconds = now() - t3;
        res->total_seconds = now() - t0;
        res->peak_device_bytes = b.peak_bytes;
        return 0;
    });
}

int aosx_band_solve_v1(aosx_bundle *bundle, const aosx_band_request_v1 *q, aosx_band_result_v1 *res) {
    if (!q || !res) return 1;
    begin(res);
    if (!abi_ok(q)) return fail(res, AOSX_ABI_MISMATCH, 0, "request ABI mismatch");
    if (!bundle || !q->probes || !q->energies_ev || !q->energy_counts || !q->maximum_residual_ev ||
        !q->subspace_rank || !q->saturated || (!q->k_fractional && (!q->h_data || !q->s_data)))
        return fail(res, AOSX_NULL_POINTER, 0, "required input or output pointer is null");
    Bundle &b = *bundle->impl;
    if (b.device < 0) return fail(res, AOSX_NO_DEVICE, 0, "bundle was created without a device");
    if (q->n_kpoints <= 0 || q->k_batch <= 0 || q->k_batch > 64)
        return fail(res, AOSX_BATCH_DIMENSIONS, 0, "invalid k-point count or k batch");
    if (q->integration_points < 1 || q->contour_batch_size < 1 ||
        q->integration_points % q->contour_batch_size || q->block_size < 1 || q->moment_size < 1 ||
        q->block_size % 32 || q->block_size * q->moment_size > 256 || q->precision < 0 || q->precision > 4 ||
        q->refinement_steps < 0 || q->refinement_steps > 16 || q->refinement_passes < 0 ||
        q->refinement_passes > 4 || !(q->lower_ev < q->upper_ev) || !std::isfinite(q->lower_ev) ||
        !std::isfinite(q->upper_ev) || !(q->ellipse_vertical_scale > 0.0))
        return fail(res, AOSX_INVALID_SETTINGS, 0,
                    "CISS sizes must be positive, contour_batch_size must divide integration_points, "
                    "block_size must be a multiple of 32 and block_size*moment_size at most 256");
    if (q->precision == AOSX_PRECISION_SHIFT_IR && !(q->far_threshold >= 0.0 && q->far_threshold <= 2.0))
        return fail(res, AOSX_INVALID_SETTINGS, 0, "far_threshold must lie in [0, 2]");
    if (q->intervals)
        for (int64_t k = 0; k < q->n_kpoints; ++k)
            if (!(q->intervals[2 * k] < q->intervals[2 * k + 1]) || !std::isfinite(q->intervals[2 * k]) ||
                !std::isfinite(q->intervals[2 * k + 1]))
                return fail(res, AOSX_INVALID_SETTINGS, 0, "per-k interval %lld requires lower < upper",
                            (long long)k);
    if ((int64_t)q->block_size * q->moment_size >= b.n)
        return fail(res, AOSX_SEARCH_DIMENSION, 0, "CISS search dimension must be smaller than matrix dimension");
    if (q->energy_capacity < q->block_size * q->moment_size)
        return fail(res, AOSX_BATCH_DIMENSIONS, 0, "energy capacity is below block_size*moment_size");
    if (q->k_fractional) {
        for (int64_t i = 0; i < 3 * q->n_kpoints; ++i)
            if (!std::isfinite(q->k_fractional[i])) return fail(res, AOSX_NONFINITE_VALUES, 0, "k point is not finite");
    } else if (!finite_c64(q->h_data, q->n_kpoints * b.nnz) || !finite_c64(q->s_data, q->n_kpoints * b.nnz)) {
        return fail(res, AOSX_NONFINITE_VALUES, 0, "H/S values contain non-finite entries");
    }
    if (q->augmentation && (q->augmentation_width < 32 || q->augmentation_width % 32 || q->augmentation_width > 224))
        return fail(res, AOSX_INVALID_SETTINGS, 0, "augmentation_width must be a multiple of 32 in [32, 224]");
    if (!finite_c64(q->probes, b.n * q->block_size) ||
        (q->augmentation && !finite_c64(q->augmentation, b.n * q->augmentation_width)))
        return fail(res, AOSX_NONFINITE_RHS, 0, "probe block contains non-finite values");
    return guarded(res, [&]() -> int {
        aosx::CissParams p;
        p.nk = q->n_kpoints;
        p.lower = q->lower_ev;
        p.upper = q->upper_ev;
        p.integration_points = q->integration_points;
        p.contour_batch_size = q->contour_batch_size;
        p.block_size = q->block_size;
        p.moment_size = q->moment_size;
        p.vscale = q->ellipse_vertical_scale;
        p.guard = q->endpoint_guard_ev;
        p.rank_tol = q->rank_relative_tolerance;
        p.spurious = q->spurious_relative_threshold;
        p.precision = q->precision;
        p.refinement_steps = q->refinement_steps;
        p.refinement_passes = q->refinement_passes;
        p.k_batch = q->k_batch;
        p.conjugate_pairs = q->conjugate_pairs;
        p.resident_limit = q->resident_limit;
        p.intervals = q->intervals;
        p.want_vectors = q->eigenvectors != nullptr;
        p.far_threshold = q->far_threshold;
        aosx::CissOutput o;
        b.band(p, q->k_fractional, reinterpret_cast<const float2 *>(q->h_data),
               reinterpret_cast<const float2 *>(q->s_data), reinterpret_cast<const float2 *>(q->probes),
               reinterpret_cast<const float2 *>(q->augmentation), q->augmentation_width, o);
        for (int64_t k = 0; k < q->n_kpoints; ++k) {
            const auto &e = o.energies[k];
            q->energy_counts[k] = (int32_t)e.size();
            for (size_t j = 0; j < e.size() && (int)j < q->energy_capacity; ++j)
                q->energies_ev[k * q->energy_capacity + j] = e[j];
            q->maximum_residual_ev[k] = o.max_residual[k];
            q->subspace_rank[k] = o.rank[k];
            q->saturated[k] = o.saturated[k];
            if (q->pass_rank) q->pass_rank[k] = o.pass_rank[k];
            if (q->pass_width) q->pass_width[k] = o.pass_width[k];
            if (q->eigenvectors) {
                const int cap = q->energy_capacity, c = (int)std::min<size_t>(e.size(), (size_t)cap);
                aosx_complex64 *dst = q->eigenvectors + (size_t)k * b.n * cap;
                const std::vector<float2> &v = o.vectors[k];
                for (int64_t r = 0; r < b.n; ++r)
                    for (int j = 0; j < cap; ++j)
                        dst[(size_t)r * cap + j] = j < c ? aosx_complex64{v[(size_t)r * e.size() + j].x,
                                                                         v[(size_t)r * e.size() + j].y}
                                                         : aosx_complex64{0.f, 0.f};
            }
        }
        for (int i = 0; i < AOSX_STAGE_COUNT; ++i) res->stage_seconds[i] = o.stages[i];
        res->replaced_pivots = o.replaced_pivots;
        res->factorizations = o.factorizations;
        res->peak_device_bytes = b.peak_bytes;
        return 0;
    });
}

int aosx_eig_residual_v1(aosx_bundle *bundle, const aosx_residual_request_v1 *q, aosx_status_v1 *res) {
    if (!q || !res) return 1;
    begin(res);
    if (!abi_ok(q)) return fail(res, AOSX_ABI_MISMATCH, 0, "request ABI mismatch");
    if (!bundle || !q->energies_ev || !q->vectors || !q->residual_ev || (!q->k_fractional && (!q->h_data || !q->s_data)))
        return fail(res, AOSX_NULL_POINTER, 0, "required input or output pointer is null");
    Bundle &b = *bundle->impl;
    if (b.device < 0) return fail(res, AOSX_NO_DEVICE, 0, "bundle was created without a device");
    if (q->n_kpoints <= 0 || q->n_kpoints > 64 || q->count < 1 || q->count > 256)
        return fail(res, AOSX_BATCH_DIMENSIONS, 0, "k-point count must be in [1, 64] and count in [1, 256]");
    if (q->k_fractional) {
        for (int64_t i = 0; i < 3 * q->n_kpoints; ++i)
            if (!std::isfinite(q->k_fractional[i])) return fail(res, AOSX_NONFINITE_VALUES, 0, "k point is not finite");
    } else if (!finite_c64(q->h_data, q->n_kpoints * b.nnz) || !finite_c64(q->s_data, q->n_kpoints * b.nnz)) {
        return fail(res, AOSX_NONFINITE_VALUES, 0, "H/S values contain non-finite entries");
    }
    return guarded(res, [&]() -> int {
        const double t0 = now();
        b.eig_residual(q->n_kpoints, q->k_fractional, reinterpret_cast<const float2 *>(q->h_data),
                       reinterpret_cast<const float2 *>(q->s_data), q->count, q->energies_ev, q->vectors,
                       q->residual_ev);
        res->seconds = now() - t0;
        return 0;
    });
}

int aosx_kernel_stats_read_v1(aosx_bundle *bundle, int32_t enable, aosx_kernel_stats_v1 *out) {
    if (!bundle || !out) return 1;
    Bundle &b = *bundle->impl;
    const aosx::KernelStats &k = b.kstats;
    std::memset(out, 0, sizeof(*out));
    out->abi_version = AOSX_ABI_VERSION;
    out->struct_size = (uint32_t)sizeof(*out);
    out->schur_launches = k.schur_launches;
    out->schur_flops = k.schur_flops;
    out->schur_ms = k.schur_ms;
    out->panel_launches = k.panel_launches;
    out->panel_ms = k.panel_ms;
    out->getrf_launches = k.getrf_launches;
    out->getrf_ms = k.getrf_ms;
    out->solve_launches = k.solve_launches;
    out->solve_ms = k.solve_ms;
    out->solve_bytes = k.solve_bytes;
    out->assemble_launches = k.assemble_launches;
    out->assemble_ms = k.assemble_ms;
    out->assemble_bytes = k.assemble_bytes;
    out->inner_launches = k.inner_launches;
    out->inner_flops = k.inner_flops;
    out->inner_ms = k.inner_ms;
    out->schur_bytes = k.schur_bytes;
    b.kstats = aosx::KernelStats();
    b.kstats.enabled = enable != 0;
    return 0;
}

int aosx_bundle_autotune_v1(aosx_bundle *bundle, const aosx_autotune_request_v1 *q, aosx_autotune_result_v1 *res) {
    if (!q || !res) return 1;
    begin(res);
    if (!abi_ok(q)) return fail(res, AOSX_ABI_MISMATCH, 0, "request ABI mismatch");
    if (!bundle) return fail(res, AOSX_NULL_POINTER, 0, "bundle is null");
    Bundle &b = *bundle->impl;
    if (b.device < 0) return fail(res, AOSX_NO_DEVICE, 0, "bundle was created without a device");
    if (q->matrices < 1 || q->matrices > 256) return fail(res, AOSX_BATCH_DIMENSIONS, 0, "matrices must be 1..256");
    if (q->precision < 0 || q->precision > 3 || q->repeats < 1 || q->repeats > 100)
        return fail(res, AOSX_INVALID_SETTINGS, 0, "invalid precision or repeats");
    if (q->n_candidates < 0 || q->n_candidates > AOSX_AUTOTUNE_MAX_CANDIDATES || (q->n_candidates && !q->candidates))
        return fail(res, AOSX_INVALID_SETTINGS, 0, "invalid candidate list");
    std::vector<int> cands;
    for (int i = 0; i < q->n_candidates; ++i) {
        if (q->candidates[i] < 1 || q->candidates[i] > 32)
            return fail(res, AOSX_INVALID_SETTINGS, 0, "group sizes must be 1..32");
        cands.push_back(q->candidates[i]);
    }
    if (cands.empty()) cands = {4, 8, 12, 16, 24, 32};
    return guarded(res, [&]() -> int {
        std::vector<double> ms;
        res->chosen_group = b.autotune(q->matrices, q->precision, cands, q->repeats, ms);
        res->n_candidates = (int32_t)cands.size();
        for (size_t i = 0; i < cands.size(); ++i) {
            res->candidates[i] = cands[i];
            res->factor_ms[i] = ms[i];
        }
        return 0;
    });
}

int aosx_struct_sizes_v1(int64_t *sizes, int32_t capacity) {
    const int64_t all[] = {(int64_t)sizeof(aosx_bundle_request_v1), (int64_t)sizeof(aosx_bundle_result_v1),
                           (int64_t)sizeof(aosx_status_v1),         (int64_t)sizeof(aosx_assemble_request_v1),
                           (int64_t)sizeof(aosx_shift_request_v1),  (int64_t)sizeof(aosx_shift_result_v1),
                           (int64_t)sizeof(aosx_band_request_v1),   (int64_t)sizeof(aosx_band_result_v1),
                           (int64_t)sizeof(aosx_kernel_stats_v1),   (int64_t)sizeof(aosx_autotune_request_v1),
                           (int64_t)sizeof(aosx_autotune_result_v1), (int64_t)sizeof(aosx_residual_request_v1),
                           (int64_t)sizeof(aosx_assemble_noalloc_request_v1)};
    const int count = (int)(sizeof(all) / sizeof(all[0]));
    for (int i = 0; i < count && i < capacity; ++i) sizes[i] = all[i];
    return count;
}

}  // extern "C"
