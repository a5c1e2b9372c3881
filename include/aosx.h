/*
 * aosx.h -- C-ABI of the MI355X-native AOSpectrum Band path (libaosx_hip.so).
 *
 * Conventions follow the reference's shifted-solve ABI
 * (native/include/aospectrum_cudss_complex64_ubatch.h:10-66): every call takes
 * a versioned request struct and a result struct (abi_version, struct_size),
 * returns 0 on success and 1 on failure, and on failure fills `status` with a
 * numbered code and `error_message` with a NUL-terminated explanation.  Caller
 * buffers are borrowed for the duration of one call; device memory is owned
 * by the bundle handle.  Calls on one handle must come from one host thread at
 * a time; distinct handles are independent.
 *
 * Entry points and the reference interface each one replaces:
 *   aosx_bundle_create_v1   -- BlochPattern.from_bundle (src/aospectrum/sparse.py:63-140)
 *                              plus the symbolic analysis the reference repeats
 *                              inside every cudssExecute(ANALYSIS) call
 *                              (native/src/aospectrum_cudss_complex64_ubatch.c:446-474);
 *                              done once per bundle here.
 *   aosx_bundle_pattern_v1  -- BlochPattern.csr_indptr / csr_indices (sparse.py:122-140).
 *   aosx_assemble_v1        -- BlochWorkspace.update (sparse.py:199-245).
 *   aosx_shift_solve_v1     -- aospectrum_cudss_solve_complex64_ubatch_v1
 *                              (…ubatch.h:64-66) behind CudssShiftSession.solve
 *                              (native/src/bindings.cpp:106-203): solves
 *                              (H - z_b S) X_b = RHS for a batch of shifts.
 *   aosx_band_solve_v1      -- CudssCissSolver.solve_interval (src/aospectrum/ciss.py:266-391)
 *                              applied to BlochWorkspace.update for each k point
 *                              (src/aospectrum/band.py:227-253), for a batch of k.
 *   aosx_eig_residual_v1    -- the eigenpair residual of CudssCissSolver.solve_interval
 *                              (src/aospectrum/ciss.py:362-372) for given pairs.
 *   aosx_bundle_destroy_v1  -- CudssShiftSession.close (bindings.cpp:103).
 */
#ifndef AOSX_H
#define AOSX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AOSX_ABI_VERSION 1u
#define AOSX_ERROR_CAPACITY 512u

/* Status codes.  10-36 keep the numbering of the reference's ubatch call
 * (…ubatch.c:136-241); 20 is a HIP runtime error (CUDA error there), 22 a
 * numerical breakdown of the factorisation (cuDSS INFO != 0 there), 40 a fill
 * estimate above the device memory (short memory estimate there). */
enum {
    AOSX_OK = 0,
    AOSX_ABI_MISMATCH = 10,
    AOSX_INT32_OVERFLOW = 11,
    AOSX_BATCH_DIMENSIONS = 12,
    AOSX_NULL_POINTER = 13,
    AOSX_INDPTR_SPAN = 14,
    AOSX_HIP_ERROR = 20,
    AOSX_NUMERICAL_BREAKDOWN = 22,
    AOSX_ROWS_NOT_MONOTONE = 30,
    AOSX_COLUMN_RANGE = 31,
    AOSX_COLUMNS_UNSORTED = 32,
    AOSX_MISSING_DIAGONAL = 33,
    AOSX_SIZE_OVERFLOW = 34,
    AOSX_NONFINITE_VALUES = 35,
    AOSX_NONFINITE_RHS = 36,
    AOSX_MEMORY_ESTIMATE = 40,
    AOSX_INVALID_BLOCKS = 50,
    AOSX_NOT_HERMITIAN = 51,
    AOSX_INVALID_SETTINGS = 52,
    AOSX_NO_DEVICE = 53,
    AOSX_ZERO_RANK = 60,
    AOSX_NO_STABLE_SUBSPACE = 61,
    AOSX_OVERLAP_NOT_PD = 62,
    AOSX_GHOST_FILTER = 63,
    AOSX_SEARCH_DIMENSION = 64
};

/* Arithmetic of the numeric factorisation. */
enum {
    AOSX_PRECISION_FP32 = 0,       /* complex FP32 LU, FP32 MFMA Schur updates        */
    AOSX_PRECISION_FP32_IR = 1,    /* same LU + iterative refinement on FP64 residuals */
    AOSX_PRECISION_BF16X3_IR = 2,  /* Schur updates as three BF16 MFMA products (split
                                      hi/lo operands, FP32 accumulate) + FP64-residual IR */
    AOSX_PRECISION_BF16_IR = 3,    /* Schur updates in one BF16 MFMA product + IR      */
    AOSX_PRECISION_SHIFT_IR = 4    /* per contour point (band solves only): one BF16
                                      product for points far from the real axis
                                      (|sin theta_j| >= far_threshold), three for
                                      the others; FP64-residual IR for all          */
};

/* Fill-reducing ordering of the atom graph. */
enum {
    AOSX_ORDERING_AUTO = 0,             /* nested dissection with positions, else minimum degree */
    AOSX_ORDERING_NATURAL = 1,
    AOSX_ORDERING_MINIMUM_DEGREE = 2,
    AOSX_ORDERING_NESTED_DISSECTION = 3 /* needs atom_positions                            */
};

typedef struct {
    float real;
    float imag;
} aosx_complex64;

typedef struct aosx_bundle aosx_bundle;

/* ------------------------------------------------------------------ bundle */
typedef struct {
    uint32_t abi_version;
    uint32_t struct_size;
    int64_t n_orbitals;
    int64_t n_atoms;
    int64_t n_blocks;
    const int64_t *atom_offsets;   /* [n_atoms + 1], AO range of every atom        */
    const int32_t *row_atom;       /* [n_blocks]                                   */
    const int32_t *col_atom;       /* [n_blocks]                                   */
    const int32_t *lattice_shift;  /* [n_blocks][3]                                */
    const int64_t *block_offsets;  /* [n_blocks + 1], row-major block values       */
    int32_t value_kind;            /* 0: float32 values, 1: complex64 values       */
    int32_t device;                /* HIP device ordinal; -1: host analysis only   */
    const void *h_values;          /* [block_offsets[n_blocks]]                    */
    const void *s_values;
    int32_t tile_size;             /* factor tile edge: 0 (default 64) or 64       */
    int32_t ordering;              /* AOSX_ORDERING_*                              */
    const double *atom_positions;  /* [n_atoms][3] Cartesian, or NULL; steers the
                                      nested-dissection split planes            */
} aosx_bundle_request_v1;

typedef struct {
    uint32_t abi_version;
    uint32_t struct_size;
    int32_t status;
    int32_t hip_status;
    int64_t n;                     /* orbitals                                     */
    int64_t nnz;                   /* entries of the shared H/S CSR pattern        */
    int64_t n_phase_shifts;        /* unique lattice shifts                        */
    int64_t n_diagonal_tiles;      /* tiles along the diagonal (ceil(n / T))       */
    int64_t n_factor_tiles;        /* stored L+U tiles per factorisation           */
    int64_t atom_fill_entries;     /* nnz(L+U) of the atom-level symbolic factor   */
    int64_t tile_size;
    double factor_flops;           /* real flops of one tile LU (8 per complex FMA) */
    double schur_flops;            /* of which Schur-update GEMMs                  */
    uint64_t factor_bytes;         /* device bytes of one factorisation            */
    uint64_t resident_device_bytes;/* bundle-resident device bytes                 */
    double pattern_seconds;
    double ordering_seconds;
    double symbolic_seconds;
    double upload_seconds;
    char error_message[AOSX_ERROR_CAPACITY];
} aosx_bundle_result_v1;

/* Generic result of the calls that return no statistics of their own. */
typedef struct {
    uint32_t abi_version;
    uint32_t struct_size;
    int32_t status;
    int32_t hip_status;
    double seconds;
    char error_message[AOSX_ERROR_CAPACITY];
} aosx_status_v1;

int aosx_bundle_create_v1(const aosx_bundle_request_v1 *request,
                          aosx_bundle_result_v1 *result, aosx_bundle **bundle);

/* indptr [n+1], indices [nnz], permutation [n] (the orbitals in elimination
 * order; the factor also holds padding rows between nested-dissection nodes,
 * which are not listed); any pointer may be NULL. */
int aosx_bundle_pattern_v1(const aosx_bundle *bundle, int32_t *indptr, int32_t *indices,
                           int32_t *permutation, aosx_status_v1 *result);

int aosx_bundle_destroy_v1(aosx_bundle *bundle);

/* Frees the bundle's device workspace (factors, right-hand sides, subspace
 * buffers) while keeping its resident block data and symbolic analysis.  No
 * counterpart in the reference, which frees everything after every call; the
 * serving loop uses it to hand memory from an idle bundle to another job.
 * Writes the bytes still held in *resident_bytes (may be NULL). */
int aosx_bundle_trim_v1(aosx_bundle *bundle, uint64_t *resident_bytes);

/* ---------------------------------------------------------------- assemble */
typedef struct {
    uint32_t abi_version;
    uint32_t struct_size;
    int64_t n_kpoints;
    const double *k_fractional;    /* [n_kpoints][3]                               */
    aosx_complex64 *h_data;        /* [n_kpoints][nnz] out, CSR order              */
    aosx_complex64 *s_data;        /* [n_kpoints][nnz] out                         */
} aosx_assemble_request_v1;

int aosx_assemble_v1(aosx_bundle *bundle, const aosx_assemble_request_v1 *request,
                     aosx_status_v1 *result);

/* ------------------------------------------------- allocation-free assembly
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory -- code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc: the same H(k), S(k) as aosx_assemble_v1, written into
 * caller-owned device buffers, with no device allocation or free.  The k
 * points travel as kernel arguments and the Bloch phases are staged in a
 * static device array of AOSX_NOALLOC_SCRATCH_BYTES per device, so
 * n_kpoints <= 64 and n_kpoints * n_phase_shifts * 8 <= that size; larger
 * requests fail with AOSX_BATCH_DIMENSIONS.  Calls that overlap on one device
 * share the array and race; the call returns after its kernels finished. */
#define AOSX_NOALLOC_SCRATCH_BYTES (4u << 20)
#define AOSX_MAX_DEVICES 64

typedef struct {
    uint32_t abi_version;
    uint32_t struct_size;
    int64_t n_kpoints;             /* 1..64                                        */
    const double *k_fractional;    /* [n_kpoints][3] host                          */
    aosx_complex64 *h_out;         /* [n_kpoints][nnz] device memory of the bundle's */
    aosx_complex64 *s_out;         /* device, CSR order                             */
} aosx_assemble_noalloc_request_v1;

int aosx_assemble_noalloc_v1(aosx_bundle *bundle, const aosx_assemble_noalloc_request_v1 *request,
                             aosx_status_v1 *result);

/* ------------------------------------------------------------- shift solve */
typedef struct {
    uint32_t abi_version;
    uint32_t struct_size;
    int32_t batch_count;           /* shifts, 1..64                                */
    int32_t rhs_count;             /* right-hand sides, 1..256                     */
    const aosx_complex64 *h_data;  /* [nnz] on the bundle's CSR pattern            */
    const aosx_complex64 *s_data;  /* [nnz]                                        */
    const aosx_complex64 *shifts;  /* [batch_count]                                */
    const aosx_complex64 *rhs;     /* [n][rhs_count], row-major (NumPy C order)    */
    aosx_complex64 *solution;      /* [batch_count][n][rhs_count] out              */
    int32_t precision;             /* AOSX_PRECISION_*                             */
    int32_t refinement_steps;      /* IR steps for the *_IR precisions             */
} aosx_shift_request_v1;

typedef struct {
    uint32_t abi_version;
    uint32_t struct_size;
    int32_t status;
    int32_t hip_status;
    int64_t replaced_pivots;       /* static-pivoting perturbations, all shifts    */
    int32_t refinement_steps;
    int32_t reserved;
    double relative_residual;      /* max_b ||A_b X_b - RHS||_F / ||RHS||_F (FP64) */
    double upload_seconds;
    double factorization_seconds;
    double solve_seconds;
    double download_seconds;
    double total_seconds;
    uint64_t peak_device_bytes;
    char error_message[AOSX_ERROR_CAPACITY];
} aosx_shift_result_v1;

int aosx_shift_solve_v1(aosx_bundle *bundle, const aosx_shift_request_v1 *request,
                        aosx_shift_result_v1 *result);

/* -------------------------------------------------------------- band solve */
typedef struct {
    uint32_t abi_version;
    uint32_t struct_size;
    int64_t n_kpoints;
    const double *k_fractional;    /* [n_kpoints][3]; NULL when h_data/s_data given */
    const aosx_complex64 *h_data;  /* [n_kpoints][nnz] explicit matrices, or NULL  */
    const aosx_complex64 *s_data;
    double lower_ev;               /* absolute solve interval                      */
    double upper_ev;
    const double *intervals;       /* [n_kpoints][2] per-k intervals, or NULL      */
    /* CissSettings (src/aospectrum/ciss.py:35-62) */
    int32_t integration_points;
    int32_t contour_batch_size;
    int32_t block_size;
    int32_t moment_size;
    double ellipse_vertical_scale;
    double endpoint_guard_ev;
    double rank_relative_tolerance;
    double spurious_relative_threshold;
    const aosx_complex64 *probes;  /* [n][block_size] orthonormal probe block      */
    const aosx_complex64 *augmentation; /* [n][augmentation_width] independent probe
                                      block joined to the Ritz block of the first
                                      filter pass, or NULL                      */
    int32_t augmentation_width;    /* columns of `augmentation` (multiple of 32)   */
    int32_t reserved_aug;
    /* MI355X-specific execution knobs */
    int32_t precision;             /* AOSX_PRECISION_*                             */
    int32_t refinement_steps;      /* IR steps per solve (*_IR precisions)         */
    int32_t refinement_passes;     /* extra filter passes on the Ritz block        */
    int32_t k_batch;               /* k points factored together                   */
    int32_t conjugate_pairs;       /* 1: solve z_{N-1-j} = conj(z_j) with A^H      */
    int32_t resident_limit;        /* filter passes keep a batch's factorisations  */
                                   /* resident when they fit in device memory and  */
                                   /* number <= this (0: no limit); otherwise each */
                                   /* pass factors every shift group again         */
    /* outputs */
    float *energies_ev;            /* [n_kpoints][capacity], ascending per k       */
    int32_t energy_capacity;       /* >= block_size * moment_size                  */
    int32_t *energy_counts;        /* [n_kpoints]                                  */
    double *maximum_residual_ev;   /* [n_kpoints]                                  */
    int32_t *subspace_rank;        /* [n_kpoints]                                  */
    int32_t *saturated;            /* [n_kpoints], 1 when the search space may be  */
                                   /* too small: moment rank == L*M, or the last   */
                                   /* filter pass reached the pass capacity        */
    aosx_complex64 *eigenvectors;  /* [n_kpoints][n][energy_capacity] out, or NULL: */
                                   /* the selected Ritz vectors (S-normalised,     */
                                   /* orbital order), column j for energies_ev[j]  */
    int32_t *pass_rank;            /* [n_kpoints] out, or NULL: numerical rank of   */
                                   /* the last filter pass's block                 */
    int32_t *pass_width;           /* [n_kpoints] out, or NULL: the pass capacity,  */
                                   /* moment rank + augmentation_width (the later  */
                                   /* passes refilter the same span)               */
    double far_threshold;          /* AOSX_PRECISION_SHIFT_IR: contour point j     */
                                   /* (angle theta_j = 2 pi (j + 1/2) / N) takes   */
                                   /* one BF16 product when |sin theta_j| >= this; */
                                   /* in [0, 2] (above 1: none)                    */
} aosx_band_request_v1;

#define AOSX_STAGE_COUNT 12
/* stage_seconds order: assembly, probe_setup, factorization, solve, refinement,
 * moments, subspace_compression, rayleigh_ritz, residual, filter_passes,
 * host_sync, total */
typedef struct {
    uint32_t abi_version;
    uint32_t struct_size;
    int32_t status;
    int32_t hip_status;
    int64_t replaced_pivots;
    int64_t factorizations;
    double stage_seconds[AOSX_STAGE_COUNT];
    uint64_t peak_device_bytes;
    char error_message[AOSX_ERROR_CAPACITY];
} aosx_band_result_v1;

int aosx_band_solve_v1(aosx_bundle *bundle, const aosx_band_request_v1 *request,
                       aosx_band_result_v1 *result);

/* ------------------------------------------------------- eigenpair residual
 * ||H(k) v_j - e_j S(k) v_j||_2 for caller-given pairs, in FP64 from the
 * complex64 H(k), S(k) of the bundle: the residual of
 * src/aospectrum/ciss.py:362-372 that aosx_band_solve_v1 reports as
 * maximum_residual_ev, exposed so that it can be checked independently. */
typedef struct {
    uint32_t abi_version;
    uint32_t struct_size;
    int64_t n_kpoints;
    const double *k_fractional;    /* [n_kpoints][3]; NULL when h_data/s_data given */
    const aosx_complex64 *h_data;  /* [n_kpoints][nnz] explicit matrices, or NULL  */
    const aosx_complex64 *s_data;
    int32_t count;                 /* pairs per k point, 1..256                    */
    int32_t reserved;
    const double *energies_ev;     /* [n_kpoints][count]                           */
    const double *vectors;         /* [n_kpoints][n][count] complex128 (re, im)    */
    double *residual_ev;           /* [n_kpoints][count] out                       */
} aosx_residual_request_v1;

int aosx_eig_residual_v1(aosx_bundle *bundle, const aosx_residual_request_v1 *request, aosx_status_v1 *result);

/* ------------------------------------------------------- kernel statistics
 * Device-event timings of the hot kernels, accumulated while enabled.  No
 * counterpart in the reference (which only reports wall-clock stages,
 * …ubatch.h:45-51); used to report roofline figures.  Each call returns the
 * totals gathered since the previous call, resets them, and switches event
 * timing on (enable != 0) or off. */
typedef struct {
    uint32_t abi_version;
    uint32_t struct_size;
    int64_t schur_launches;        /* trailing Schur updates of the step groups    */
    double schur_flops;            /* algorithmic real flops of those launches     */
    double schur_ms;
    int64_t panel_launches;
    double panel_ms;
    int64_t getrf_launches;
    double getrf_ms;
    int64_t solve_launches;        /* full forward + backward sweeps               */
    double solve_ms;
    double solve_bytes;            /* algorithmic bytes of the sweeps              */
    int64_t assemble_launches;
    double assemble_ms;
    double assemble_bytes;         /* algorithmic bytes of the Bloch assembly      */
    int64_t inner_launches;        /* Schur updates inside a step group            */
    double inner_flops;
    double inner_ms;
    double schur_bytes;            /* algorithmic bytes of the trailing launches:  */
                                   /* targets read + written once, operands once   */
} aosx_kernel_stats_v1;

int aosx_kernel_stats_read_v1(aosx_bundle *bundle, int32_t enable, aosx_kernel_stats_v1 *stats);

/* ------------------------------------------------------------- autotuning
 * Picks the step-group size of the bundle's tile LU (how many elimination
 * steps a trailing Schur launch accumulates, and the group the solve sweeps
 * use) by timing factorisations of `matrices` shifted copies of H(Gamma),
 * S(Gamma) for each candidate; keeps the fastest for every later call.  No
 * counterpart in the reference (cuDSS chooses its own schedule); the
 * factorisations use the workspace a band solve of that size would use. */
#define AOSX_AUTOTUNE_MAX_CANDIDATES 16
typedef struct {
    uint32_t abi_version;
    uint32_t struct_size;
    int32_t matrices;              /* factorisations per launch to tune for, 1..256 */
    int32_t precision;             /* AOSX_PRECISION_*                             */
    int32_t n_candidates;          /* 0: the built-in set {4, 8, 12, 16, 24, 32}  */
    int32_t repeats;               /* timed factorisations per candidate, >= 1     */
    const int32_t *candidates;     /* [n_candidates] group sizes in 1..32          */
} aosx_autotune_request_v1;

typedef struct {
    uint32_t abi_version;
    uint32_t struct_size;
    int32_t status;
    int32_t hip_status;
    int32_t chosen_group;
    int32_t n_candidates;
    int32_t candidates[AOSX_AUTOTUNE_MAX_CANDIDATES];
    double factor_ms[AOSX_AUTOTUNE_MAX_CANDIDATES];  /* mean per factorisation launch */
    char error_message[AOSX_ERROR_CAPACITY];
} aosx_autotune_result_v1;

int aosx_bundle_autotune_v1(aosx_bundle *bundle, const aosx_autotune_request_v1 *request,
                            aosx_autotune_result_v1 *result);

/* sizeof of every request/result struct, so bindings can verify their
 * layouts, in this order: bundle request, bundle result, status, assemble
 * request, shift request, shift result, band request, band result, kernel
 * stats, autotune request, autotune result, residual request, allocation-
 * free assemble request; returns the number of sizes written. */
int aosx_struct_sizes_v1(int64_t *sizes, int32_t capacity);

#ifdef __cplusplus
}
#endif

#endif
