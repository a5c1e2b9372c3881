// Device-resident state of one operator bundle and the host pipeline API.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/aosx.h"
#include "symbolic.h"

namespace aosx {

struct Error : std::runtime_error {
    int code;
    int hip_status;
    Error(int c, const std::string &m, int h = 0) : std::runtime_error(m), code(c), hip_status(h) {}
};

#define AOSX_HIP(call)                                                                                 \
    do {                                                                                               \
        hipError_t _e = (call);                                                                        \
        if (_e != hipSuccess)                                                                          \
            throw ::aosx::Error(AOSX_HIP_ERROR, std::string(#call) + " failed: " + hipGetErrorString(_e), \
                                (int)_e);                                                              \
    } while (0)

#define AOSX_LAUNCHED() AOSX_HIP(hipGetLastError())

struct KernelStats {
    bool enabled = false;
    int64_t schur_launches = 0;     // trailing (grouped) Schur updates
    double schur_flops = 0.0;
    double schur_ms = 0.0;
    double schur_bytes = 0.0;       // algorithmic bytes of the trailing launches
    int64_t inner_launches = 0;     // Schur updates inside a step group
    double inner_flops = 0.0;
    double inner_ms = 0.0;
    int64_t panel_launches = 0;
    double panel_ms = 0.0;
    int64_t getrf_launches = 0;
    double getrf_ms = 0.0;
    int64_t solve_launches = 0;
    double solve_ms = 0.0;
    double solve_bytes = 0.0;
    int64_t assemble_launches = 0;
    double assemble_ms = 0.0;
    double assemble_bytes = 0.0;
};

// Categories of device-timed launch ranges (aosx_kernel_stats_v1 fields).
enum RangeKind { RANGE_SCHUR = 0, RANGE_INNER, RANGE_PANEL, RANGE_GETRF, RANGE_SOLVE };

// A pair of events around launches on the bundle's stream, with the
// algorithmic work of what they bracket (flops, or bytes for solves).
struct TimedRange {
    hipEvent_t a = nullptr, b = nullptr;
    int kind = 0;
    double work = 0.0;
    double bytes = 0.0;
};

// A captured launch sequence: the executable graph and the timed ranges whose
// event-record nodes it contains (read after every replay).
struct GraphEntry {
    hipGraphExec_t exec = nullptr;
    std::vector<TimedRange> ranges;
};

struct CissParams {
    int64_t nk = 0;
    double lower = 0.0, upper = 0.0;
    int integration_points = 32, contour_batch_size = 1, block_size = 32, moment_size = 4;
    double vscale = 0.1, guard = 1e-5, rank_tol = 1e-5, spurious = 1e-4;
    int precision = AOSX_PRECISION_FP32, refinement_steps = 0, refinement_passes = 0, k_batch = 1;
    int conjugate_pairs = 0;
    int resident_limit = 0;  // 0: keep factorisations resident whenever they fit
    const double *intervals = nullptr;  // [nk][2] per-k absolute intervals, or null
    bool want_vectors = false;          // return the selected Ritz vectors
    double far_threshold = 2.0;         // AOSX_PRECISION_SHIFT_IR: |sin theta_j| >= this -> one BF16 product
};

struct CissOutput {
    std::vector<std::vector<float>> energies;  // per k, ascending
    std::vector<double> max_residual;
    std::vector<int> rank;
    std::vector<int> saturated;
    std::vector<int> pass_rank, pass_width;      // last filter pass (0 without passes)
    std::vector<std::vector<float2>> vectors;    // per k: [n][count] when requested
    double stages[AOSX_STAGE_COUNT] = {0};
    int64_t replaced_pivots = 0;
    int64_t factorizations = 0;
};

struct Bundle {
    int device = -1;
    BlockTable blocks_view;  // valid during create only
    Pattern pat;
    TileStructure ts;
    int64_t n = 0, nnz = 0, n_groups = 0;
    hipStream_t stream = nullptr;
    // resident device data
    float2 *d_hval = nullptr, *d_sval = nullptr;
    int32_t *d_source = nullptr, *d_group = nullptr, *d_starts = nullptr, *d_mirror = nullptr;
    int32_t *d_indices = nullptr, *d_indptr = nullptr, *d_shifts = nullptr;
    int32_t *d_iperm = nullptr, *d_perm = nullptr, *d_tile_id = nullptr, *d_next = nullptr, *d_prev = nullptr;
    // CSR entries of each stored tile (sources, positions; source -1: a padding
    // row's unit diagonal), for k_fill_tiles
    int32_t *d_tptr = nullptr, *d_tsrc = nullptr;
    uint16_t *d_tpos = nullptr;
    int32_t *d_atom_off = nullptr;   // [n_atoms + 1] orbital offsets (atom-block SpMM)
    // Step groups of the factorisation: steps [grp_begin[g], grp_begin[g+1])
    // update the tiles of their own rows and columns step by step; all other
    // targets take the group's updates in one trailing launch over the pair
    // list d_pairs[grp_pair_ptr[g] .. grp_pair_ptr[g+1]) (tile indices i, j).
    int schur_group = 16;
    std::vector<int32_t> grp_begin;
    std::vector<int64_t> grp_pair_ptr;
    std::vector<double> grp_updates;  // tile updates of each trailing launch
    std::vector<double> grp_tiles;    // tiles a trailing launch must move per matrix:
                                      // each target read + written, each operand read
    int32_t *d_pairs = nullptr;
    // the same trailing targets as 2 x 2 blocks (i0, i1, j0, j1; -1 absent)
    std::vector<int64_t> grp_quad_ptr;
    int32_t *d_quads = nullptr;

    // Step masks, bit t - g0 for step t of the group: per pair (trailing and
    // Crout lists alike) the steps with both operands L(i,t), U(t,j); per quad
    // the steps with L(i0,t), L(i1,t), U(t,j0), U(t,j1) (four masks).
    int32_t *d_pair_mask = nullptr, *d_quad_mask = nullptr;
    // BF16 precisions: the panels of a group also write split copies of their
    // tiles; L(i,t) sits in slot ppos[t*nt+i] of the group's split area and
    // U(t,j) in slot ppos[t*nt+j] + step_s[t] (split_tiles slots per matrix).
    int32_t *d_ppos = nullptr, *d_step_s = nullptr;
    int64_t split_tiles = 0;
    // Solve sweeps use the same step groups: rows outside group g that its
    // steps couple to, forward (i >= g1) and backward (i < g0), ascending.
    std::vector<int64_t> fwd_row_ptr, bwd_row_ptr;
    // Per step t: the tiles of row and column t -- (t,t), (i,t), (t,i) for i in
    // next[t] -- which take the updates of the group's earlier steps in one
    // launch before t is factored (d_pairs[crout_ptr[t] .. crout_ptr[t+1])).
    std::vector<int64_t> crout_ptr;
    std::vector<double> crout_updates;  // tile updates of each of those launches
    int32_t *d_fwd_rows = nullptr, *d_bwd_rows = nullptr;
    int64_t n_atoms = 0;
    int max_atom = 0;                // largest atom, in orbitals
    uint64_t resident_bytes = 0;
    uint64_t peak_bytes = 0;
    KernelStats kstats;
    // Launch sequences of factor() and solve() run as HIP graphs, captured on
    // the first call with a given key (sizes, precision, buffer addresses)
    // and replayed afterwards; AOSX_GRAPHS=0 runs them eagerly.
    bool use_graphs = true;
    // trailing update form (AOSX_QUAD_PIPELINE): 0 one target per workgroup
    // (the reference form), 1 2 x 2 quads staged through LDS
    int quad_pipeline = 1;
    // in-group steps of the grouped solves (AOSX_SOLVE_FOLD): 1 one launch per
    // step group, 0 a diagonal and an update launch per step
    int solve_fold = 1;
    std::map<std::vector<int64_t>, GraphEntry> graphs;
    std::vector<TimedRange> *ranges = nullptr;  // collecting ranges of the running sequence
    bool capturing = false;
    int range_begin(int kind);
    void range_end(int id, double work, double bytes = 0.0);
    void flush_ranges(std::vector<TimedRange> &rs, bool destroy);
    void clear_graphs();
    template <typename F>
    void run_sequence(const std::vector<int64_t> &key, F &&body);
    struct Slot {
        void *ptr = nullptr;
        size_t bytes = 0;
    };
    std::map<std::string, Slot> ws;

    ~Bundle();
    void *raw(const std::string &name, size_t bytes);
    template <typename T>
    T *buf(const std::string &name, size_t count) {
        return static_cast<T *>(raw(name, count * sizeof(T)));
    }
    void plan_schur(int group);
    // times `repeats` factorisations of nmat shifted copies of H(Gamma), S(Gamma)
    // per candidate group size, keeps the fastest and returns it
    int autotune(int nmat, int precision, const std::vector<int> &candidates, int repeats,
                 std::vector<double> &factor_ms);
    uint64_t plan_bytes = 0;  // device bytes of the step-group plan
    void release_workspace();
    uint64_t workspace_bytes() const;
    size_t factor_floats() const { return (size_t)ts.n_tiles * 2 * 64 * 64; }

    void zero(void *p, size_t bytes);  // bytes: multiple of 16
    // Y[j] = A[kidx[j]] X[xidx[j]] (n x w complex64) for j < nmat
    void spmm(float2 *y, const float2 *a, const int32_t *d_kidx, const float2 *x, const int32_t *d_xidx, int w,
              int nmat);
    // r[m] = b[b_index[m]] - (H_k - z_m S_k) x[m] in FP64, norms into sumsq/bsq
    void residual(float2 *r, const float2 *b, const int32_t *d_bidx, const float2 *x, int w, const float2 *d_h,
                  const float2 *d_s, const int32_t *d_kidx, const float2 *d_z, int nmat, double *sumsq, double *bsq);
    void upload(const void *h_values, const void *s_values, int value_kind, int64_t n_values,
                const int64_t *atom_offsets, int64_t atoms);
    // pipeline stages
    void assemble(const double *kpts_host, int64_t nk, float2 *d_h, float2 *d_s);
    // Allocation-free assembly into caller-owned device buffers: the phases
    // go to this device's static scratch array (no workspace, no copies).
    void *noalloc_scratch = nullptr;  // device address of that array
    void assemble_noalloc(const double *kpts_host, int nk, float2 *d_h, float2 *d_s);
    // one_product[m] != 0: matrix m takes one BF16 product per update
    // (AOSX_PRECISION_SHIFT_IR), else three
    void residual_norms(double *norm2, const double2 *V, int cap, const int32_t *d_count, const double *d_e,
                        const float2 *d_H, const float2 *d_S, int nk, hipStream_t st);
    void factor(int nmat, const std::vector<int32_t> &kidx, const std::vector<float2> &z, int precision,
                float *tiles, float *dinv, int *d_replaced, const std::vector<uint8_t> *one_product = nullptr);
    // precision selects the MFMA form of the solve products (BF16 split for
    // the BF16 precisions, FP32 otherwise).  pairs = F > 0: nmat = 2F sets,
    // set m >= F solved with the conjugate transpose of factorisation m - F
    // (ctrans ignored); pairs = 0: set m with factorisation m and ctrans.
    void solve(int nmat, int W, float2 *y, float *tiles, float *dinv, bool ctrans, int precision, int pairs = 0);
    void solve_full(int nmat, int W, const float2 *d_rhs, const std::vector<int32_t> &rhs_index,
                    const float2 *d_h, const float2 *d_s, const std::vector<int32_t> &kidx,
                    const std::vector<float2> &zeff, float *tiles, float *dinv, bool ctrans, int ir_steps,
                    float2 *x_out, double *rel_residual, int precision, int pairs = 0);
    // norm[k][j] = ||H_k v_kj - e_kj S_k v_kj||_2 (FP64) for count pairs per k;
    // H_k, S_k assembled at kpts or given explicitly (host arrays)
    void eig_residual(int64_t nk, const double *kpts, const float2 *h_in, const float2 *s_in, int count,
                      const double *energies, const double *vectors, double *norm);
    void band(const CissParams &p, const double *kpts, const float2 *h_in, const float2 *s_in,
              const float2 *probes, const float2 *augmentation, int augmentation_width, CissOutput &out);
};

}  // namespace aosx
