// Host pipeline of the Band path on one GPU: resident bundle data, Bloch
// assembly, batched tile LU of shifted matrices, triangular solves with
// FP64-residual refinement, and the CISS subspace stage.
//
// The CISS steps follow src/aospectrum/ciss.py:266-391: contour geometry
// (ciss.py:81-104), moments -w_j t_j^m X_j (ciss.py:324-333), rank-revealing
// S-orthonormalisation (ciss.py:107-126), Rayleigh-Ritz (ciss.py:129-150), the
// sigma-weighted ghost filter (ciss.py:153-177), the strict interval selection
// and the maximum eigenpair residual (ciss.py:355-372).  Added here:
//   * conjugate pairs: z_{N-1-j} = conj(z_j) and (H - conj(z) S) = (H - z S)^H,
//     so one factorisation serves both shifts (the second through A^H = U^H L^H);
//   * filter passes: the Ritz block of the previous pass, augmented with the
//     (re-)filtered probe block, is filtered again through the resident
//     factorisations (moment 0 only) and projected anew, which squares the
//     out-of-window leak of the contour filter.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>

#include "bundle.h"
#include "kernels_asm.hip"
#include "kernels_dense.hip"
#include "kernels_eigh.hip"
#include "kernels_lu.hip"
#include "kernels_solve.hip"

namespace aosx {

namespace {
double now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
inline int blocks_for(int64_t work, int per, int cap) {
    int64_t b = (work + per - 1) / per;
    return (int)std::max<int64_t>(1, std::min<int64_t>(b, cap));
}
struct EventTimer {
    hipEvent_t a = nullptr, b = nullptr;
    bool on;
    hipStream_t s;
    EventTimer(bool enabled, hipStream_t st) : on(enabled), s(st) {
        if (on) {
            AOSX_HIP(hipEventCreate(&a));
            AOSX_HIP(hipEventCreate(&b));
            AOSX_HIP(hipEventRecord(a, s));
        }
    }
    double stop() {
        if (!on) return 0.0;
        AOSX_HIP(hipEventRecord(b, s));
        AOSX_HIP(hipEventSynchronize(b));
        float ms = 0.f;
        AOSX_HIP(hipEventElapsedTime(&ms, a, b));
        return ms;
    }
    ~EventTimer() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};
}  // namespace

// ---------------------------------------------------- timed ranges, graphs
int Bundle::range_begin(int kind) {
    if (!kstats.enabled || ranges == nullptr) return -1;
    TimedRange r;
    r.kind = kind;
    AOSX_HIP(hipEventCreate(&r.a));
    AOSX_HIP(hipEventCreate(&r.b));
    AOSX_HIP(hipEventRecordWithFlags(r.a, stream, capturing ? hipEventRecordExternal : 0));
    ranges->push_back(r);
    return (int)ranges->size() - 1;
}

void Bundle::range_end(int id, double work, double bytes) {
    if (id < 0) return;
    TimedRange &r = (*ranges)[id];
    AOSX_HIP(hipEventRecordWithFlags(r.b, stream, capturing ? hipEventRecordExternal : 0));
    r.work = work;
    r.bytes = bytes;
}

// Waits for the last range, adds every range to the kernel statistics.
void Bundle::flush_ranges(std::vector<TimedRange> &rs, bool destroy) {
    if (rs.empty()) return;
    AOSX_HIP(hipEventSynchronize(rs.back().b));
    for (TimedRange &r : rs) {
        float ms = 0.f;
        AOSX_HIP(hipEventElapsedTime(&ms, r.a, r.b));
        switch (r.kind) {
        case RANGE_SCHUR:
            kstats.schur_ms += ms;
            kstats.schur_launches++;
            kstats.schur_flops += r.work;
            kstats.schur_bytes += r.bytes;
            break;
        case RANGE_INNER: kstats.inner_ms += ms; kstats.inner_launches++; kstats.inner_flops += r.work; break;
        case RANGE_PANEL: kstats.panel_ms += ms; kstats.panel_launches++; break;
        case RANGE_GETRF: kstats.getrf_ms += ms; kstats.getrf_launches++; break;
        case RANGE_SOLVE: kstats.solve_ms += ms; kstats.solve_launches++; kstats.solve_bytes += r.work; break;
        }
        if (destroy) {
            (void)hipEventDestroy(r.a);
            (void)hipEventDestroy(r.b);
        }
    }
    if (destroy) rs.clear();
}

// Device allocation and release synchronise the device, and a stream capture
// running in another thread meanwhile is invalidated by HIP (a serving
// process runs one bundle per thread).  Captures and every allocation,
// release and graph destruction of all bundles in the process therefore take
// this lock; it is recursive because a capturing thread may re-enter it.
static std::recursive_mutex g_device_lock;
using DeviceLock = std::lock_guard<std::recursive_mutex>;

void Bundle::clear_graphs() {
    DeviceLock lock(g_device_lock);
    for (auto &kv : graphs) {
        if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
        for (TimedRange &r : kv.second.ranges) {
            (void)hipEventDestroy(r.a);
            (void)hipEventDestroy(r.b);
        }
    }
    graphs.clear();
}

// Runs `body` -- kernel launches and event records on `stream` only, no
// allocation, copy or synchronisation -- through the graph cached under
// `key`, capturing it on first use.
template <typename F>
void Bundle::run_sequence(const std::vector<int64_t> &key_in, F &&body) {
    std::vector<int64_t> key = key_in;
    key.push_back(kstats.enabled ? 1 : 0);
    if (!use_graphs) {
        std::vector<TimedRange> rs;
        ranges = &rs;
        try {
            body();
        } catch (...) {
            ranges = nullptr;
            flush_ranges(rs, true);
            throw;
        }
        ranges = nullptr;
        flush_ranges(rs, true);
        return;
    }
    auto it = graphs.find(key);
    if (it == graphs.end()) {
        if (graphs.size() >= 512) clear_graphs();
        DeviceLock lock(g_device_lock);  // no allocation or release anywhere during the capture
        GraphEntry e;
        ranges = &e.ranges;
        capturing = true;
        AOSX_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        hipGraph_t g = nullptr;
        try {
            body();
        } catch (...) {
            (void)hipStreamEndCapture(stream, &g);
            if (g) (void)hipGraphDestroy(g);
            capturing = false;
            ranges = nullptr;
            for (TimedRange &r : e.ranges) {
                (void)hipEventDestroy(r.a);
                (void)hipEventDestroy(r.b);
            }
            throw;
        }
        capturing = false;
        ranges = nullptr;
        AOSX_HIP(hipStreamEndCapture(stream, &g));
        const hipError_t ie = hipGraphInstantiate(&e.exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        AOSX_HIP(ie);
        it = graphs.emplace(key, std::move(e)).first;
    }
    AOSX_HIP(hipGraphLaunch(it->second.exec, stream));
    if (!it->second.ranges.empty()) flush_ranges(it->second.ranges, false);
}

// Scratch of aosx_assemble_noalloc_v1: the Bloch phases of one call,
// n_kpoints x n_phase_shifts complex64.  Each device holds its own copy until
// the process ends; calls that overlap on one device share it and race.
static __device__ __align__(256) unsigned char scratch[AOSX_NOALLOC_SCRATCH_BYTES];

namespace {
std::mutex g_scratch_mutex;
void *g_scratch_address[AOSX_MAX_DEVICES] = {};
}  // namespace

// The device address of `scratch` on `device` (current), looked up once per
// device: in host code the array's name is only a stand-in.
static void *scratch_address(int device) {
    if (device < 0 || device >= AOSX_MAX_DEVICES) throw Error(AOSX_NO_DEVICE, "device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_address[device]) {
        void *ptr = nullptr;
        AOSX_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) throw Error(AOSX_HIP_ERROR, "scratch symbol has no device address");
        g_scratch_address[device] = ptr;
    }
    return g_scratch_address[device];
}

Bundle::~Bundle() {
    if (device >= 0) {
        DeviceLock lock(g_device_lock);
        (void)hipSetDevice(device);
        clear_graphs();
        release_workspace();
        for (void *p : {(void *)d_hval, (void *)d_sval, (void *)d_source, (void *)d_group, (void *)d_starts,
                        (void *)d_mirror, (void *)d_indices, (void *)d_indptr, (void *)d_shifts,
                        (void *)d_iperm, (void *)d_perm, (void *)d_tile_id, (void *)d_next, (void *)d_prev,
                        (void *)d_atom_off, (void *)d_pairs, (void *)d_quads, (void *)d_ppos, (void *)d_step_s,
                        (void *)d_fwd_rows, (void *)d_bwd_rows, (void *)d_pair_mask, (void *)d_quad_mask,
                        (void *)d_tptr, (void *)d_tsrc, (void *)d_tpos})
            if (p) (void)hipFree(p);
        if (stream) (void)hipStreamDestroy(stream);
    }
}

void *Bundle::raw(const std::string &name, size_t bytes) {
    Slot &s = ws[name];
    if (s.bytes >= bytes && s.ptr) return s.ptr;
    DeviceLock lock(g_device_lock);
    if (s.ptr) {
        AOSX_HIP(hipStreamSynchronize(stream));
        AOSX_HIP(hipFree(s.ptr));
        s.ptr = nullptr;
        s.bytes = 0;
    }
    const size_t grain = 2u << 20;
    const size_t want = std::max<size_t>(grain, (bytes + grain - 1) / grain * grain);
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        throw Error(AOSX_MEMORY_ESTIMATE,
                    "device allocation of " + std::to_string(want) + " bytes for '" + name + "' failed: " +
                        hipGetErrorString(e),
                    (int)e);
    }
    s.ptr = p;
    s.bytes = want;
    peak_bytes = std::max<uint64_t>(peak_bytes, resident_bytes + workspace_bytes());
    return p;
}

void Bundle::release_workspace() {
    DeviceLock lock(g_device_lock);
    clear_graphs();
    for (auto &kv : ws)
        if (kv.second.ptr) (void)hipFree(kv.second.ptr);
    ws.clear();
}

uint64_t Bundle::workspace_bytes() const {
    uint64_t t = 0;
    for (auto &kv : ws) t += kv.second.bytes;
    return t;
}

void Bundle::zero(void *p, size_t bytes) {
    const int64_t c4 = (int64_t)(bytes / 16);
    k_fill_zero<<<blocks_for(c4, 256, 8192), 256, 0, stream>>>(static_cast<float4 *>(p), c4);
    AOSX_LAUNCHED();
}

template <typename T>
static T *upload_vec(const std::vector<T> &v, uint64_t &counter) {
    DeviceLock lock(g_device_lock);
    T *p = nullptr;
    const size_t bytes = std::max<size_t>(1, v.size()) * sizeof(T);
    AOSX_HIP(hipMalloc(&p, bytes));
    if (!v.empty()) AOSX_HIP(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    counter += bytes;
    return p;
}

void Bundle::upload(const void *h_values, const void *s_values, int value_kind, int64_t n_values,
                    const int64_t *atom_offsets, int64_t atoms) {
    AOSX_HIP(hipSetDevice(device));
    AOSX_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    std::vector<float2> tmp(n_values);
    auto convert = [&](const void *src) {
        if (value_kind == 0) {
            const float *f = static_cast<const float *>(src);
            for (int64_t i = 0; i < n_values; ++i) tmp[i] = make_float2(f[i], 0.f);
        } else {
            std::memcpy(tmp.data(), src, n_values * sizeof(float2));
        }
    };
    convert(h_values);
    d_hval = upload_vec(tmp, resident_bytes);
    convert(s_values);
    d_sval = upload_vec(tmp, resident_bytes);
    std::vector<int32_t> src32(pat.source.size()), starts32(pat.starts.size());
    for (size_t i = 0; i < src32.size(); ++i) src32[i] = (int32_t)pat.source[i];
    for (size_t i = 0; i < starts32.size(); ++i) starts32[i] = (int32_t)pat.starts[i];
    d_source = upload_vec(src32, resident_bytes);
    d_starts = upload_vec(starts32, resident_bytes);
    d_group = upload_vec(pat.group, resident_bytes);
    d_mirror = upload_vec(pat.mirror, resident_bytes);
    d_indices = upload_vec(pat.indices, resident_bytes);
    d_indptr = upload_vec(pat.indptr, resident_bytes);
    d_shifts = upload_vec(pat.shifts, resident_bytes);
    d_iperm = upload_vec(ts.iperm, resident_bytes);
    d_perm = upload_vec(ts.perm, resident_bytes);
    d_tile_id = upload_vec(ts.tile_id, resident_bytes);
    {
        // entries of every stored tile by counting sort on the tile id
        const int nt = ts.n_diag;
        const int64_t npad = (int64_t)ts.padding_rows.size();
        auto tile_of = [&](int32_t fr, int32_t fc) {
            const int32_t id = ts.tile_id[(int64_t)(fr / TILE) * nt + fc / TILE];
            if (id < 0) throw Error(AOSX_INVALID_BLOCKS, "pattern entry outside the tile structure");
            return id;
        };
        std::vector<int32_t> tptr(ts.n_tiles + 1, 0), tsrc(nnz + npad);
        std::vector<uint16_t> tpos(nnz + npad);
        for (int64_t p = 0; p < nnz; ++p) ++tptr[tile_of(ts.iperm[pat.rows[p]], ts.iperm[pat.indices[p]]) + 1];
        for (int32_t r : ts.padding_rows) ++tptr[tile_of(r, r) + 1];
        for (int64_t t = 0; t < ts.n_tiles; ++t) tptr[t + 1] += tptr[t];
        std::vector<int32_t> cur(tptr.begin(), tptr.end() - 1);
        for (int64_t p = 0; p < nnz; ++p) {
            const int32_t fr = ts.iperm[pat.rows[p]], fc = ts.iperm[pat.indices[p]];
            const int32_t e = cur[tile_of(fr, fc)]++;
            tsrc[e] = (int32_t)p;
            tpos[e] = (uint16_t)((fr % TILE) * TILE + fc % TILE);
        }
        for (int32_t r : ts.padding_rows) {
            const int32_t e = cur[tile_of(r, r)]++;
            tsrc[e] = -1;
            tpos[e] = (uint16_t)((r % TILE) * TILE + r % TILE);
        }
        d_tptr = upload_vec(tptr, resident_bytes);
        d_tsrc = upload_vec(tsrc, resident_bytes);
        d_tpos = upload_vec(tpos, resident_bytes);
    }
    d_next = upload_vec(ts.next_idx, resident_bytes);
    d_prev = upload_vec(ts.prev_idx, resident_bytes);
    n_atoms = atoms;
    std::vector<int32_t> off(atoms + 1);
    max_atom = 0;
    for (int64_t a = 0; a <= atoms; ++a) off[a] = (int32_t)atom_offsets[a];
    for (int64_t a = 0; a < atoms; ++a) max_atom = std::max<int>(max_atom, off[a + 1] - off[a]);
    d_atom_off = upload_vec(off, resident_bytes);
    int group = 16;
    if (const char *env = std::getenv("AOSX_SCHUR_GROUP")) group = std::min(32, std::max(1, std::atoi(env)));
    plan_schur(group);
    if (const char *env = std::getenv("AOSX_GRAPHS")) use_graphs = std::atoi(env) != 0;
    if (const char *env = std::getenv("AOSX_QUAD_PIPELINE")) quad_pipeline = std::min(1, std::max(0, std::atoi(env)));
    if (const char *env = std::getenv("AOSX_SOLVE_FOLD")) solve_fold = std::atoi(env) != 0 ? 1 : 0;
    peak_bytes = resident_bytes;
    noalloc_scratch = scratch_address(device);

}

// Step groups of `group` consecutive elimination steps and the trailing pair
// list of each: (i, j) with i, j >= g1 and i, j in next[t] for some step t of
// the group.  Pairs are listed in 8 x 8 blocks of (i, j) so that the blocks
// resident on one XCD at a time share 8 L rows and 8 U columns.
void Bundle::plan_schur(int group) {
    const int nt = ts.n_diag;
    schur_group = group;
    grp_begin.clear();
    grp_pair_ptr.assign(1, 0);
    grp_updates.clear();
    grp_tiles.clear();
    std::vector<int32_t> pairs, quads, pair_mask, quad_mask;
    grp_quad_ptr.assign(1, 0);

    std::vector<uint8_t> mark((size_t)nt * nt, 0);
    std::vector<int32_t> ppos((size_t)nt * nt, -1), step_s(nt, 0);
    split_tiles = 0;
    for (int g0 = 0; g0 < nt; g0 += group) {
        const int g1 = std::min(nt, g0 + group);
        grp_begin.push_back(g0);
        double updates = 0.0, operands = 0.0;
        int64_t slot = 0;
        std::vector<int32_t> rows;  // tiles >= g1 touched by the group, ascending
        for (int t = g0; t < g1; ++t) {
            const int32_t *nx = ts.next_idx.data() + ts.next_ptr[t];
            const int s = ts.next_ptr[t + 1] - ts.next_ptr[t];
            step_s[t] = s;
            for (int a = 0; a < s; ++a) ppos[(size_t)t * nt + nx[a]] = (int32_t)(slot + a);
            slot += 2 * s;
            int p = 0;
            while (p < s && nx[p] < g1) ++p;
            updates += (double)(s - p) * (s - p);
            operands += 2.0 * (s - p);
            for (int a = p; a < s; ++a) {
                rows.push_back(nx[a]);
                for (int b = p; b < s; ++b) mark[(size_t)nx[a] * nt + nx[b]] = 1;
            }
        }
        std::sort(rows.begin(), rows.end());
        rows.erase(std::unique(rows.begin(), rows.end()), rows.end());
        const int r = (int)rows.size();
        auto marked = [&](int a, int b) { return a < r && b < r && mark[(size_t)rows[a] * nt + rows[b]] != 0; };
        // steps of the group where tile index x has an L (equivalently U) tile
        auto steps_of = [&](int x) {
            uint32_t mk = 0;
            for (int t = g0; x >= 0 && t < g1; ++t)
                if (ppos[(size_t)t * nt + x] >= 0) mk |= 1u << (t - g0);
            return mk;
        };
        for (int bi = 0; bi < r; bi += 8)
            for (int bj = 0; bj < r; bj += 8)
                for (int a = bi; a < std::min(r, bi + 8); ++a)
                    for (int b = bj; b < std::min(r, bj + 8); ++b)
                        if (marked(a, b)) {
                            pairs.push_back(rows[a]);
                            pairs.push_back(rows[b]);
                            pair_mask.push_back((int32_t)(steps_of(rows[a]) & steps_of(rows[b])));
                        }
        // quads: 2 x 2 blocks of rows and columns, in 4 x 4 blocks of quads; -1
        // pads an odd row count.  The rows are paired in order of their step
        // masks (the group's steps with an L tile in that row), so that the
        // two rows -- and columns -- of a quad mostly share their steps: a
        // quad's four waves then have work in the same steps, and few MFMAs
        // run on the zeroed operands of a wave without an update.
        std::vector<int32_t> qrows(rows);
        std::vector<uint32_t> qmask(r);
        for (int a = 0; a < r; ++a) qmask[a] = steps_of(rows[a]);
        {
            std::vector<int> idx(r);
            for (int a = 0; a < r; ++a) idx[a] = a;
            std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return qmask[x] > qmask[y]; });
            for (int a = 0; a < r; ++a) qrows[a] = rows[idx[a]];
        }
        auto qmarked = [&](int a, int b) { return a < r && b < r && mark[(size_t)qrows[a] * nt + qrows[b]] != 0; };
        for (int bi = 0; bi < r; bi += 8)
            for (int bj = 0; bj < r; bj += 8) {
                for (int a = bi; a < std::min(r, bi + 8); a += 2)
                    for (int b = bj; b < std::min(r, bj + 8); b += 2)
                        if (qmarked(a, b) || qmarked(a, b + 1) || qmarked(a + 1, b) || qmarked(a + 1, b + 1)) {
                            const int q4[4] = {qrows[a], a + 1 < r ? qrows[a + 1] : -1, qrows[b],
                                               b + 1 < r ? qrows[b + 1] : -1};
                            for (int e = 0; e < 4; ++e) {
                                quads.push_back(q4[e]);
                                quad_mask.push_back((int32_t)steps_of(q4[e]));
                            }
                        }
            }
        for (int a = 0; a < r; ++a)
            for (int b = 0; b < r; ++b) mark[(size_t)rows[a] * nt + rows[b]] = 0;
        grp_pair_ptr.push_back((int64_t)pairs.size() / 2);
        grp_quad_ptr.push_back((int64_t)quads.size() / 4);
        grp_tiles.push_back(2.0 * (double)(grp_pair_ptr.back() - grp_pair_ptr[grp_pair_ptr.size() - 2]) + operands);
        grp_updates.push_back(updates);
        split_tiles = std::max(split_tiles, slot);
    }
    grp_begin.push_back(nt);
    // row/column targets of each step, appended after the trailing pairs
    crout_ptr.assign(nt + 1, 0);
    crout_updates.assign(nt, 0.0);
    for (size_t g = 0; g + 1 < grp_begin.size(); ++g) {
        const int g0 = grp_begin[g], g1 = grp_begin[g + 1];
        for (int t = g0; t < g1; ++t) {
            crout_ptr[t] = (int64_t)pairs.size() / 2;
            const int32_t *nx = ts.next_idx.data() + ts.next_ptr[t];
            const int s = ts.next_ptr[t + 1] - ts.next_ptr[t];
            auto has = [&](int tp, int i) {  // i in next[tp]
                const int32_t *b = ts.next_idx.data() + ts.next_ptr[tp], *e = ts.next_idx.data() + ts.next_ptr[tp + 1];
                return std::binary_search(b, e, i);
            };
            auto add = [&](int i, int j) {
                int cnt = 0;
                uint32_t mk = 0;
                for (int tp = g0; tp < t; ++tp)
                    if (has(tp, i) && has(tp, j)) {
                        ++cnt;
                        mk |= 1u << (tp - g0);
                    }
                if (cnt == 0) return;
                pairs.push_back(i);
                pairs.push_back(j);
                pair_mask.push_back((int32_t)mk);
                crout_updates[t] += cnt;
            };
            add(t, t);
            for (int a = 0; a < s; ++a) add(nx[a], t);
            for (int a = 0; a < s; ++a) add(t, nx[a]);
        }
    }
    crout_ptr[nt] = (int64_t)pairs.size() / 2;
    if (std::getenv("AOSX_DEBUG")) {
        // listed steps of the quads (every wave runs the MFMAs of each) and
        // the steps where a wave's target has an update
        int64_t listed = 0, work = 0;
        for (size_t q = 0; q < quad_mask.size() / 4; ++q) {
            const uint32_t rm[2] = {(uint32_t)quad_mask[4 * q], (uint32_t)quad_mask[4 * q + 1]};
            const uint32_t cm[2] = {(uint32_t)quad_mask[4 * q + 2], (uint32_t)quad_mask[4 * q + 3]};
            listed += __builtin_popcount((rm[0] | rm[1]) & (cm[0] | cm[1]));
            for (int a = 0; a < 2; ++a)
                for (int b = 0; b < 2; ++b) work += __builtin_popcount(rm[a] & cm[b]);
        }
        // in-group solve operations: a diagonal product per step and an
        // update per tile below the diagonal inside the step's group
        int64_t ingroup = 0;
        for (size_t g = 0; g + 1 < grp_begin.size(); ++g)
            for (int t = grp_begin[g]; t < grp_begin[g + 1]; ++t)
                for (int32_t a = ts.next_ptr[t]; a < ts.next_ptr[t + 1]; ++a) ingroup += ts.next_idx[a] < grp_begin[g + 1];
        std::fprintf(stderr,
                     "AOSX_DEBUG|plan|group=%d|groups=%zu|quads=%zu|listed_steps=%lld|wave_steps=%lld|"
                     "work_fraction=%.4f|ingroup_updates=%lld|steps=%d\n",
                     group, grp_begin.size() - 1, quad_mask.size() / 4, (long long)listed, (long long)work,
                     listed ? (double)work / (4.0 * (double)listed) : 0.0, (long long)ingroup, nt);
    }
    std::vector<int32_t> frows, brows;
    fwd_row_ptr.assign(1, 0);
    bwd_row_ptr.assign(1, 0);
    for (size_t g = 0; g + 1 < grp_begin.size(); ++g) {
        const int g0 = grp_begin[g], g1 = grp_begin[g + 1];
        std::vector<int32_t> f, b;
        for (int k = g0; k < g1; ++k) {
            for (int32_t a = ts.next_ptr[k]; a < ts.next_ptr[k + 1]; ++a)
                if (ts.next_idx[a] >= g1) f.push_back(ts.next_idx[a]);
            for (int32_t a = ts.prev_ptr[k]; a < ts.prev_ptr[k + 1]; ++a)
                if (ts.prev_idx[a] < g0) b.push_back(ts.prev_idx[a]);
        }
        for (auto *v : {&f, &b}) {
            std::sort(v->begin(), v->end());
            v->erase(std::unique(v->begin(), v->end()), v->end());
        }
        frows.insert(frows.end(), f.begin(), f.end());
        brows.insert(brows.end(), b.begin(), b.end());
        fwd_row_ptr.push_back((int64_t)frows.size());
        bwd_row_ptr.push_back((int64_t)brows.size());
    }
    if (std::getenv("AOSX_DEBUG")) {
        // trailing solve work: rows (one block each per matrix and chunk
        // group) and the tiles they take (their steps)
        int64_t tiles_below = 0;
        for (size_t g = 0; g + 1 < grp_begin.size(); ++g)
            for (int k = grp_begin[g]; k < grp_begin[g + 1]; ++k)
                for (int32_t a = ts.next_ptr[k]; a < ts.next_ptr[k + 1]; ++a) tiles_below += ts.next_idx[a] >= grp_begin[g + 1];
        std::fprintf(stderr, "AOSX_DEBUG|solve_plan|fwd_rows=%zu|tiles=%lld|steps_per_row=%.2f\n", frows.size(),
                     (long long)tiles_below, frows.empty() ? 0.0 : (double)tiles_below / (double)frows.size());
    }
    // the launch sequences captured so far follow the old plan
    DeviceLock lock(g_device_lock);
    clear_graphs();
    for (int32_t *p : {d_pairs, d_quads, d_ppos, d_step_s, d_fwd_rows, d_bwd_rows, d_pair_mask, d_quad_mask})
        if (p) AOSX_HIP(hipFree(p));
    uint64_t pb = 0;
    d_quads = upload_vec(quads, pb);
    d_fwd_rows = upload_vec(frows, pb);
    d_bwd_rows = upload_vec(brows, pb);
    d_pairs = upload_vec(pairs, pb);
    d_pair_mask = upload_vec(pair_mask, pb);
    d_quad_mask = upload_vec(quad_mask, pb);
    d_ppos = upload_vec(ppos, pb);
    d_step_s = upload_vec(step_s, pb);
    resident_bytes = resident_bytes - plan_bytes + pb;
    plan_bytes = pb;
}

int Bundle::autotune(int nmat, int precision, const std::vector<int> &candidates, int repeats,
                     std::vector<double> &factor_ms) {
    AOSX_HIP(hipSetDevice(device));
    const double gamma[3] = {0.0, 0.0, 0.0};
    float2 *d_h = buf<float2>("H", (size_t)nnz);
    float2 *d_s = buf<float2>("S", (size_t)nnz);
    struct StatsOff {  // no kernel statistics while tuning; restored on exit
        KernelStats &k;
        bool saved;
        explicit StatsOff(KernelStats &ks) : k(ks), saved(ks.enabled) { k.enabled = false; }
        ~StatsOff() { k.enabled = saved; }
    } stats_off(kstats);
    assemble(gamma, 1, d_h, d_s);
    const size_t ff = factor_floats(), dif = (size_t)ts.n_diag * 2 * TILE_FLOATS;
    float *tiles = buf<float>("tiles", (size_t)nmat * ff);
    float *dinv = buf<float>("dinv", (size_t)nmat * dif);
    int *repl = buf<int>("replaced", (size_t)std::max(nmat, 256));
    // shifts on the imaginary axis: the factorisation's cost does not depend
    // on the values (static pivoting, fixed structure)
    std::vector<int32_t> kidx(nmat, 0);
    std::vector<float2> z(nmat);
    for (int j = 0; j < nmat; ++j) z[j] = make_float2(0.f, 1.f + (float)j);
    factor_ms.assign(candidates.size(), 0.0);
    int best = candidates.empty() ? schur_group : candidates[0];
    double best_ms = 1e300;
    for (size_t c = 0; c < candidates.size(); ++c) {
        plan_schur(candidates[c]);
        factor(nmat, kidx, z, precision, tiles, dinv, repl);  // captures the sequence
        AOSX_HIP(hipStreamSynchronize(stream));
        const double t0 = now();
        for (int r = 0; r < repeats; ++r) factor(nmat, kidx, z, precision, tiles, dinv, repl);
        AOSX_HIP(hipStreamSynchronize(stream));
        factor_ms[c] = (now() - t0) / repeats * 1e3;
        if (factor_ms[c] < best_ms) {
            best_ms = factor_ms[c];
            best = candidates[c];
        }
    }
    plan_schur(best);
    return best;
}

void Bundle::spmm(float2 *y, const float2 *a, const int32_t *d_kidx, const float2 *x, const int32_t *d_xidx, int w,
                  int nmat) {
    if (max_atom <= 32) {
        const unsigned blocks = (unsigned)(8 * ((n_atoms * nmat + 7) / 8));
        auto kern = max_atom <= 16 ? k_atom_spmm<false, 16> : k_atom_spmm<false, 32>;
        kern<<<blocks, 256, 0, stream>>>(y, nullptr, nullptr, x, d_xidx, w, a, nullptr, d_kidx, nullptr, d_indptr,
                                         d_indices, d_atom_off, n, nnz, (int)n_atoms, nmat, nullptr, nullptr);
    } else {
        k_spmm<<<dim3(blocks_for(n, 4, 2048), nmat), 256, 0, stream>>>(y, a, d_kidx, x, d_xidx, w, d_indptr,
                                                                       d_indices, n, nnz);
    }
    AOSX_LAUNCHED();
}

void Bundle::residual(float2 *r, const float2 *b, const int32_t *d_bidx, const float2 *x, int w, const float2 *d_h,
                      const float2 *d_s, const int32_t *d_kidx, const float2 *d_z, int nmat, double *sumsq,
                      double *bsq) {
    if (max_atom <= 32) {
        const unsigned blocks = (unsigned)(8 * ((n_atoms * nmat + 7) / 8));
        auto kern = max_atom <= 16 ? k_atom_spmm<true, 16> : k_atom_spmm<true, 32>;
        kern<<<blocks, 256, 0, stream>>>(r, b, d_bidx, x, nullptr, w, d_h, d_s, d_kidx, d_z, d_indptr, d_indices,
                                         d_atom_off, n, nnz, (int)n_atoms, nmat, sumsq, bsq);
    } else {
        k_residual_shifted<<<dim3(blocks_for(n, 4, 2048), nmat), 256, 0, stream>>>(
            r, b, d_bidx, x, w, d_h, d_s, d_kidx, d_z, d_indptr, d_indices, n, nnz, sumsq, bsq);
    }
    AOSX_LAUNCHED();
}

// ------------------------------------------------------------------ assembly
void Bundle::assemble(const double *kpts_host, int64_t nk, float2 *d_h, float2 *d_s) {
    double *d_k = buf<double>("kpts", 3 * nk);
    AOSX_HIP(hipMemcpyAsync(d_k, kpts_host, 3 * nk * sizeof(double), hipMemcpyHostToDevice, stream));
    float2 *d_ph = buf<float2>("phase", n_groups * nk);
    const int ng = (int)n_groups;
    k_phases<<<blocks_for(ng * nk, 256, 1 << 20), 256, 0, stream>>>(d_ph, d_shifts, ng, d_k, (int)nk);
    AOSX_LAUNCHED();
    EventTimer tm(kstats.enabled, stream);
    k_assemble<<<blocks_for(nnz, 256, 8192), 256, 0, stream>>>(d_h, d_s, d_hval, d_sval, d_source, d_group,
                                                                d_starts, d_mirror, d_ph, ng, nnz, (int)nk);
    AOSX_LAUNCHED();
    if (kstats.enabled) {
        kstats.assemble_ms += tm.stop();
        kstats.assemble_launches += 1;
        // algorithmic bytes: each source entry's H and S values, source and
        // group index are read for its own CSR entry and for the mirror's;
        // each CSR entry reads its mirror and start offset and writes H and S
        // (complex64) for every k point
        const double src = (double)pat.source.size();
        kstats.assemble_bytes += 2.0 * src * (2 * 8.0 + 4.0 + 4.0) + nnz * (4.0 + 4.0) + nk * nnz * 2 * 8.0;
    }
}

// ----------------------------------------------------- allocation-free assembly
void Bundle::assemble_noalloc(const double *kpts_host, int nk, float2 *d_h, float2 *d_s) {
    if (nk < 1 || nk > KBATCH_MAX) throw Error(AOSX_BATCH_DIMENSIONS, "n_kpoints must be in [1, 64]");
    if ((uint64_t)nk * (uint64_t)n_groups * sizeof(float2) > AOSX_NOALLOC_SCRATCH_BYTES)
        throw Error(AOSX_BATCH_DIMENSIONS, "n_kpoints x n_phase_shifts exceeds the scratch array");
    if (!noalloc_scratch) throw Error(AOSX_NO_DEVICE, "bundle has no scratch address");
    KBatch kb{};
    for (int k = 0; k < nk; ++k)
        for (int d = 0; d < 3; ++d) kb.k[k][d] = kpts_host[3 * k + d];
    float2 *phase = static_cast<float2 *>(noalloc_scratch);
    const int ng = (int)n_groups;
    k_phases_byval<<<blocks_for((int64_t)ng * nk, 256, 1 << 20), 256, 0, stream>>>(phase, d_shifts, ng, kb, nk);
    AOSX_LAUNCHED();
    k_assemble<<<blocks_for(nnz, 256, 8192), 256, 0, stream>>>(d_h, d_s, d_hval, d_sval, d_source, d_group,
                                                                d_starts, d_mirror, phase, ng, nnz, nk);
    AOSX_LAUNCHED();
    AOSX_HIP(hipStreamSynchronize(stream));
}

// ------------------------------------------------------------- factorization
void Bundle::factor(int nmat, const std::vector<int32_t> &kidx, const std::vector<float2> &z, int precision,
                    float *tiles, float *dinv, int *d_replaced, const std::vector<uint8_t> *one_product) {
    // the H, S values of the current k batch live in the "H"/"S" workspaces
    const float2 *d_h = static_cast<const float2 *>(ws.at("H").ptr);
    const float2 *d_s = static_cast<const float2 *>(ws.at("S").ptr);
    int32_t *d_kidx = buf<int32_t>("f_kidx", nmat);
    float2 *d_z = buf<float2>("f_z", nmat);
    const bool split_mode = precision == AOSX_PRECISION_BF16X3_IR || precision == AOSX_PRECISION_BF16_IR ||
                            precision == AOSX_PRECISION_SHIFT_IR;
    uint16_t *split = split_mode ? buf<uint16_t>("split", (size_t)nmat * split_tiles * PANEL_TILE) : nullptr;
    AOSX_HIP(hipMemcpyAsync(d_kidx, kidx.data(), nmat * sizeof(int32_t), hipMemcpyHostToDevice, stream));
    AOSX_HIP(hipMemcpyAsync(d_z, z.data(), nmat * sizeof(float2), hipMemcpyHostToDevice, stream));
    // BF16 updates: the matrices taking three products (BF16x3) first, then
    // those taking one (all of them in BF16_IR; the flagged shifts in SHIFT_IR)
    auto one = [&](int m) {
        return precision == AOSX_PRECISION_BF16_IR ||
               (precision == AOSX_PRECISION_SHIFT_IR && one_product && (*one_product)[m] != 0);
    };
    std::vector<int32_t> mats;
    for (int m = 0; m < nmat && split_mode; ++m)
        if (!one(m)) mats.push_back(m);
    const int n3 = (int)mats.size();
    for (int m = 0; m < nmat && split_mode; ++m)
        if (one(m)) mats.push_back(m);
    int32_t *d_mats = nullptr;
    if (split_mode) {
        d_mats = buf<int32_t>("f_mats", nmat);
        AOSX_HIP(hipMemcpyAsync(d_mats, mats.data(), nmat * sizeof(int32_t), hipMemcpyHostToDevice, stream));
    }
    const int n1 = split_mode ? nmat - n3 : 0;
    const int64_t mstride = (int64_t)factor_floats();
    const int nt = ts.n_diag;
    const float pivot_rel = 1.0e-6f;
    const int ng = (int)grp_begin.size() - 1;
    constexpr double TILE_FLOPS = 8.0 * 64.0 * 64.0 * 64.0;  // one complex 64^3 tile update
    // A(i,j) -= sum_{t in [t0, t1)} L(i,t) U(t,j) for the listed pairs, the
    // target held in the accumulators across the steps
    auto multi_update = [&](int t0, int t1, int64_t first, int64_t np) {
        const int64_t items = np * nmat;
        const unsigned blocks = (unsigned)(8 * ((items + 7) / 8));
        const int32_t *pr = d_pairs + 2 * first, *pm = d_pair_mask + first;
        if (!split_mode) {
            k_schur_group<<<blocks, 256, 0, stream>>>(tiles, mstride, nt, t0, pr, pm, (int)np, nmat, d_tile_id);
            AOSX_LAUNCHED();
            return;
        }
        if (n3 > 0) {
            k_schur_group_split<true><<<(unsigned)(8 * ((np * n3 + 7) / 8)), 256, 0, stream>>>(
                tiles, mstride, nt, t0, pr, pm, (int)np, n3, d_tile_id, split, split_tiles, d_ppos, d_step_s, d_mats);
            AOSX_LAUNCHED();
        }
        if (n1 > 0) {
            k_schur_group_split<false><<<(unsigned)(8 * ((np * n1 + 7) / 8)), 256, 0, stream>>>(
                tiles, mstride, nt, t0, pr, pm, (int)np, n1, d_tile_id, split, split_tiles, d_ppos, d_step_s,
                d_mats + n3);
            AOSX_LAUNCHED();
        }
    };
    auto launches = [&]() {
        k_fill_tiles<<<(unsigned)ts.n_tiles, 256, 0, stream>>>(tiles, mstride, d_h, d_s, nnz, d_kidx, d_z, nmat,
                                                                d_tptr, d_tsrc, d_tpos);
        AOSX_LAUNCHED();
        for (int g = 0; g < ng; ++g) {
            const int g0 = grp_begin[g], g1 = grp_begin[g + 1];
            int next_slot = 0;  // split slots are handed out step by step within the group
            for (int k = g0; k < g1; ++k) {
                const int s = ts.next_ptr[k + 1] - ts.next_ptr[k];
                const int slot0 = next_slot;  // split slot of this step's first panel tile
                next_slot += 2 * s;
                const int32_t *nx = d_next + ts.next_ptr[k];
                // row and column k take the updates of the group's earlier steps
                const int64_t nc = crout_ptr[k + 1] - crout_ptr[k];
                if (nc > 0) {
                    const int r = range_begin(RANGE_INNER);
                    multi_update(g0, k, crout_ptr[k], nc);
                    range_end(r, TILE_FLOPS * crout_updates[k] * nmat);
                }
                int r = range_begin(RANGE_GETRF);
                k_getrf_diag<<<nmat, 256, 0, stream>>>(tiles, mstride, dinv, nt, k, d_tile_id, d_replaced,
                                                       pivot_rel);
                AOSX_LAUNCHED();
                range_end(r, 0.0);
                if (s == 0) continue;
                r = range_begin(RANGE_PANEL);
                if (split_mode)
                    k_panel_split<<<dim3(2 * s, nmat), 256, 0, stream>>>(tiles, mstride, dinv, nt, k, nx, s,
                                                                         d_tile_id, split, split_tiles, slot0);
                else
                    k_panel<<<dim3(2 * s, nmat), 256, 0, stream>>>(tiles, mstride, dinv, nt, k, nx, s, d_tile_id);
                AOSX_LAUNCHED();
                range_end(r, 0.0);
            }
            // every target outside the group's rows and columns, once
            const int64_t np = grp_pair_ptr[g + 1] - grp_pair_ptr[g];
            if (np == 0) continue;
            const int r = range_begin(RANGE_SCHUR);
            if (split_mode) {
                // three-multiplication products: 2 x 2 target quads, or the
                // one-target reference form (AOSX_QUAD_PIPELINE=0)
                const int64_t nq = grp_quad_ptr[g + 1] - grp_quad_ptr[g];
                const int32_t *qd = d_quads + 4 * grp_quad_ptr[g], *qm = d_quad_mask + 4 * grp_quad_ptr[g];
                const int32_t *pr = d_pairs + 2 * grp_pair_ptr[g], *pm = d_pair_mask + grp_pair_ptr[g];
                for (const bool three : {true, false}) {
                    const int cnt = three ? n3 : n1;
                    if (cnt == 0) continue;
                    const int32_t *mats = three ? d_mats : d_mats + n3;
                    if (quad_pipeline == 1) {
                        auto quad = three ? k_schur_quad_3m<true> : k_schur_quad_3m<false>;
                        quad<<<(unsigned)(8 * ((nq * cnt + 7) / 8)), 256, 0, stream>>>(
                            tiles, mstride, nt, g0, qd, qm, (int)nq, cnt, d_tile_id, split, split_tiles, d_ppos,
                            d_step_s, mats);
                    } else {
                        auto pairs = three ? k_schur_pairs_3m<true> : k_schur_pairs_3m<false>;
                        pairs<<<(unsigned)(8 * ((np * cnt + 7) / 8)), 256, 0, stream>>>(
                            tiles, mstride, nt, g0, pr, pm, (int)np, cnt, d_tile_id, split, split_tiles, d_ppos,
                            d_step_s, mats);
                    }
                    AOSX_LAUNCHED();
                }
            } else {
                multi_update(g0, g1, grp_pair_ptr[g], np);
            }
            range_end(r, TILE_FLOPS * grp_updates[g] * nmat, grp_tiles[g] * nmat * TILE_FLOATS * 4.0);
        }
    };
    run_sequence({0, nmat, precision, (int64_t)tiles, (int64_t)dinv, (int64_t)d_replaced, (int64_t)split,
                  (int64_t)d_h, (int64_t)d_s, (int64_t)d_kidx, (int64_t)d_z, (int64_t)d_mats, n3},
                 launches);
}

// ------------------------------------------------------------------- solves
// y: planar tiled RHS (Nt x nchunk blocks per matrix), solved in place.
void Bundle::solve(int nmat, int W, float2 *y2, float *tiles, float *dinv, bool ctrans, int precision, int pairs) {
    const bool split = precision == AOSX_PRECISION_BF16X3_IR || precision == AOSX_PRECISION_BF16_IR ||
                       precision == AOSX_PRECISION_SHIFT_IR;
    float *y = reinterpret_cast<float *>(y2);
    const int nt = ts.n_diag, nchunk = (W + 63) / 64;
    const int64_t ystride = (int64_t)nt * nchunk * TILE_FLOATS;
    const int64_t mstride = (int64_t)factor_floats();
    // split RHS slots (BF16 precisions): finished Y_k blocks, split once
    uint16_t *ysplit = split ? buf<uint16_t>("s_ysplit", (size_t)nmat * YSLOTS * nchunk * PANEL_TILE) : nullptr;
    // Y_k <- D_k Y_k (which: 0 inv(L_kk), 1 inv(U_kk) for a plain set; a
    // conjugate set takes the other), and the split copy
    const int ct = ctrans ? 1 : 0;
    auto diag_launch = [&](int k, int which) {
        if (split)
            k_solve_diag_split<<<nmat, 256, 0, stream>>>(y, ystride, nchunk, dinv, nt, k, which, ct, pairs, ysplit);
        else
            k_solve_diag<<<nmat, 256, 0, stream>>>(y, ystride, nchunk, dinv, nt, k, which, ct, pairs);
        AOSX_LAUNCHED();
    };
    // Y_i -= T_ik Y_k for the `count` rows listed at `rows_dev`
    auto update_launch = [&](int k, const int32_t *rows_dev, int count) {
        if (split)
            k_solve_update_split<<<dim3(count, nmat), 256, 0, stream>>>(y, ystride, nchunk, tiles, mstride, d_tile_id,
                                                                         nt, k, rows_dev, ct, pairs, ysplit);
        else
            k_solve_update<<<dim3(count, nmat), 256, 0, stream>>>(y, ystride, nchunk, tiles, mstride, d_tile_id, nt,
                                                                   k, rows_dev, ct, pairs);
        AOSX_LAUNCHED();
    };
    // chunks per trailing block: two, or one for a single-chunk RHS in the
    // BF16 precisions (its four waves then take one 32 x 32 block each)
    const int tnc = split && nchunk == 1 ? 1 : 2;
    const int ngrp = (nchunk + tnc - 1) / tnc;
    const int ng = (int)grp_begin.size() - 1;
    // the steps of group [g0, g1) in one launch (see k_solve_group)
    auto group_launch = [&](int g0, int g1, bool fwd, int which) {
        const unsigned blocks = (unsigned)(nmat * nchunk);
        if (split)
            k_solve_group_split<<<blocks, 256, 0, stream>>>(y, ystride, nchunk, tiles, mstride, d_tile_id, nt, dinv,
                                                            g0, g1, fwd ? 1 : 0, which, ct, pairs, ysplit);
        else
            k_solve_group<<<blocks, 256, 0, stream>>>(y, ystride, nchunk, tiles, mstride, d_tile_id, nt, dinv, g0, g1,
                                                      fwd ? 1 : 0, which, ct, pairs);
        AOSX_LAUNCHED();
    };
    auto trailing = [&](int g0, int g1, const int32_t *rows, int64_t nrows) {
        if (nrows == 0) return;
        const int64_t items = nrows * nmat * ngrp;
        const unsigned blocks = (unsigned)(8 * ((items + 7) / 8));
        if (!split)
            k_solve_trail<2><<<blocks, 256, 0, stream>>>(y, ystride, nchunk, tiles, mstride, d_tile_id, nt, g0, g1,
                                                         rows, (int)nrows, nmat, ct, pairs);
        else if (tnc == 2)
            k_solve_trail_3m<2><<<blocks, 256, 0, stream>>>(y, ystride, nchunk, tiles, mstride, d_tile_id, nt, g0,
                                                            g1, rows, (int)nrows, nmat, ct, pairs, ysplit);
        else
            k_solve_trail_3m<1><<<blocks, 256, 0, stream>>>(y, ystride, nchunk, tiles, mstride, d_tile_id, nt, g0,
                                                            g1, rows, (int)nrows, nmat, ct, pairs, ysplit);
        AOSX_LAUNCHED();
    };
    // Step groups pay off when a launch covers many matrices; with a few, the
    // per-step sweep (more, shorter launches) keeps more of the GPU busy.
    const bool grouped = nmat >= 8;
    // algorithmic bytes: every stored tile is read once per forward +
    // backward pair, every RHS block read and written once per direction
    const double bytes = (double)nmat * ((double)ts.n_tiles * TILE_FLOATS * 4.0 +
                                         4.0 * (double)nt * nchunk * TILE_FLOATS * 4.0);
    auto launches = [&]() {
        const int r = range_begin(RANGE_SOLVE);
        if (!grouped) {
            for (int k = 0; k < nt; ++k) {
                diag_launch(k, 0);
                const int s = ts.next_ptr[k + 1] - ts.next_ptr[k];
                if (s) {
                    update_launch(k, d_next + ts.next_ptr[k], s);
                }
            }
            for (int k = nt - 1; k >= 0; --k) {
                diag_launch(k, 1);
                const int s = ts.prev_ptr[k + 1] - ts.prev_ptr[k];
                if (s) {
                    update_launch(k, d_prev + ts.prev_ptr[k], s);
                }
            }
        }
        // forward: Y_k <- D_k Y_k, then the rows of the group below k, then the
        // group's trailing rows in one launch
        for (int g = 0; grouped && g < ng; ++g) {
            const int g0 = grp_begin[g], g1 = grp_begin[g + 1];
            if (solve_fold) group_launch(g0, g1, true, 0);
            for (int k = g0; !solve_fold && k < g1; ++k) {
                diag_launch(k, 0);
                const int32_t *nx = ts.next_idx.data() + ts.next_ptr[k];
                const int s = ts.next_ptr[k + 1] - ts.next_ptr[k];
                int p = 0;
                while (p < s && nx[p] < g1) ++p;
                if (p) {
                    update_launch(k, d_next + ts.next_ptr[k], p);
                }
            }
            trailing(g0, g1, d_fwd_rows + fwd_row_ptr[g], fwd_row_ptr[g + 1] - fwd_row_ptr[g]);
        }
        // backward: the same groups in reverse, steps descending
        for (int g = ng - 1; grouped && g >= 0; --g) {
            const int g0 = grp_begin[g], g1 = grp_begin[g + 1];
            if (solve_fold) group_launch(g0, g1, false, 1);
            for (int k = g1 - 1; !solve_fold && k >= g0; --k) {
                diag_launch(k, 1);
                const int32_t *pv = ts.prev_idx.data() + ts.prev_ptr[k];
                const int s = ts.prev_ptr[k + 1] - ts.prev_ptr[k];
                int a = s;
                while (a > 0 && pv[a - 1] >= g0) --a;
                if (a < s) {
                    update_launch(k, d_prev + ts.prev_ptr[k] + a, s - a);
                }
            }
            trailing(g0, g1, d_bwd_rows + bwd_row_ptr[g], bwd_row_ptr[g + 1] - bwd_row_ptr[g]);
        }
        range_end(r, bytes);
    };
    run_sequence(
        {1, nmat, W, (int64_t)y, (int64_t)tiles, (int64_t)dinv, ctrans ? 1 : 0, precision, (int64_t)ysplit, pairs},
        launches);
}

void Bundle::solve_full(int nmat, int W, const float2 *d_rhs, const std::vector<int32_t> &rhs_index,
                        const float2 *d_h, const float2 *d_s, const std::vector<int32_t> &kidx,
                        const std::vector<float2> &zeff, float *tiles, float *dinv, bool ctrans, int ir_steps,
                        float2 *x_out, double *rel_residual, int precision, int pairs) {
    const int nt = ts.n_diag, nchunk = (W + 63) / 64;
    const int64_t padded = (int64_t)nt * 64;
    const int64_t ystride = (int64_t)nt * nchunk * TILE_FLOATS;
    float *y = buf<float>("s_y", (size_t)nmat * ystride);
    int32_t *d_ri = buf<int32_t>("s_rhs_index", nmat);
    int32_t *d_id = buf<int32_t>("s_identity", nmat);
    int32_t *d_ki = buf<int32_t>("s_kidx", nmat);
    float2 *d_z = buf<float2>("s_z", nmat);
    std::vector<int32_t> ident(nmat);
    for (int i = 0; i < nmat; ++i) ident[i] = i;
    AOSX_HIP(hipMemcpyAsync(d_ri, rhs_index.data(), nmat * 4, hipMemcpyHostToDevice, stream));
    AOSX_HIP(hipMemcpyAsync(d_id, ident.data(), nmat * 4, hipMemcpyHostToDevice, stream));
    AOSX_HIP(hipMemcpyAsync(d_ki, kidx.data(), nmat * 4, hipMemcpyHostToDevice, stream));
    AOSX_HIP(hipMemcpyAsync(d_z, zeff.data(), nmat * sizeof(float2), hipMemcpyHostToDevice, stream));
    const int pb = blocks_for(padded * nchunk * 64, 256, 4096);
    const int ub = blocks_for(n * W, 256, 4096);
    float2 *y2 = reinterpret_cast<float2 *>(y);
    k_pack_rhs<<<dim3(pb, nmat), 256, 0, stream>>>(y, ystride, nchunk, d_rhs, W, d_ri, d_perm, n, padded);
    AOSX_LAUNCHED();
    solve(nmat, W, y2, tiles, dinv, ctrans, precision, pairs);
    k_unpack<<<dim3(ub, nmat), 256, 0, stream>>>(x_out, W, y, ystride, nchunk, d_iperm, n);
    AOSX_LAUNCHED();
    if (ir_steps <= 0 && !rel_residual) return;
    float2 *r = buf<float2>("s_r", (size_t)nmat * n * W);
    double *d_norm = buf<double>("s_norm", 2 * (size_t)nmat);
    for (int step = 0; step <= ir_steps; ++step) {
        const bool last = step == ir_steps;
        if (last && !rel_residual) break;
        k_fill_zero<<<1, 64, 0, stream>>>(reinterpret_cast<float4 *>(d_norm), (int64_t)nmat);
        AOSX_LAUNCHED();
        residual(r, d_rhs, d_ri, x_out, W, d_h, d_s, d_ki, d_z, nmat, d_norm, d_norm + nmat);
        if (last) {
            std::vector<double> hn(2 * (size_t)nmat);
            AOSX_HIP(hipMemcpyAsync(hn.data(), d_norm, hn.size() * 8, hipMemcpyDeviceToHost, stream));
            AOSX_HIP(hipStreamSynchronize(stream));
            double worst = 0.0;
            for (int i = 0; i < nmat; ++i)
                worst = std::max(worst, std::sqrt(hn[i] / std::max(hn[nmat + i], 1e-300)));
            *rel_residual = worst;
            break;
        }
        k_pack_rhs<<<dim3(pb, nmat), 256, 0, stream>>>(y, ystride, nchunk, r, W, d_id, d_perm, n, padded);
        AOSX_LAUNCHED();
        solve(nmat, W, y2, tiles, dinv, ctrans, precision, pairs);
        k_unpack<<<dim3(ub, nmat), 256, 0, stream>>>(r, W, y, ystride, nchunk, d_iperm, n);
        AOSX_LAUNCHED();
        k_axpy<<<blocks_for((int64_t)nmat * n * W, 256, 8192), 256, 0, stream>>>(x_out, r, (int64_t)nmat * n * W);
        AOSX_LAUNCHED();
    }
}

// -------------------------------------------------------- eigenpair residual
// norm2[k][c] += ||(H_k - e_c S_k) v_c||^2 for the count[k] columns of V[k]
// (FP64): the atom-block form when atoms have at most 32 orbitals, else the
// row form.
void Bundle::residual_norms(double *norm2, const double2 *V, int cap, const int32_t *d_count, const double *d_e,
                            const float2 *d_H, const float2 *d_S, int nk, hipStream_t st) {
    if (max_atom <= 32) {
        const unsigned blocks = (unsigned)(8 * ((n_atoms * nk + 7) / 8));
        auto kern = max_atom <= 16 ? k_eig_residual_atoms<16> : k_eig_residual_atoms<32>;
        kern<<<blocks, 256, 0, st>>>(norm2, V, cap, d_count, d_e, d_H, d_S, d_indptr, d_indices, d_atom_off, n, nnz,
                                     (int)n_atoms, nk);
    } else {
        k_eig_residual<<<dim3(blocks_for(n, 4, 1024), (unsigned)nk), 256, 0, st>>>(norm2, V, cap, d_count, d_e, d_H,
                                                                                  d_S, d_indptr, d_indices, n, nnz);
    }
    AOSX_LAUNCHED();
}

void Bundle::eig_residual(int64_t nk, const double *kpts, const float2 *h_in, const float2 *s_in, int count,
                          const double *energies, const double *vectors, double *norm) {
    AOSX_HIP(hipSetDevice(device));
    float2 *d_H = buf<float2>("H", (size_t)nk * nnz);
    float2 *d_S = buf<float2>("S", (size_t)nk * nnz);
    if (h_in) {
        AOSX_HIP(hipMemcpyAsync(d_H, h_in, (size_t)nk * nnz * 8, hipMemcpyHostToDevice, stream));
        AOSX_HIP(hipMemcpyAsync(d_S, s_in, (size_t)nk * nnz * 8, hipMemcpyHostToDevice, stream));
    } else {
        assemble(kpts, nk, d_H, d_S);
    }
    double2 *d_vec = buf<double2>("r_vec", (size_t)nk * n * count);
    double *d_e = buf<double>("r_energy", (size_t)nk * count);
    int32_t *d_cnt = buf<int32_t>("r_cnt", nk);
    double *d_n2 = buf<double>("r_norm2", (size_t)nk * count + 2);
    std::vector<int32_t> cnt(nk, count);
    AOSX_HIP(hipMemcpyAsync(d_vec, vectors, (size_t)nk * n * count * sizeof(double2), hipMemcpyHostToDevice, stream));
    AOSX_HIP(hipMemcpyAsync(d_e, energies, (size_t)nk * count * 8, hipMemcpyHostToDevice, stream));
    AOSX_HIP(hipMemcpyAsync(d_cnt, cnt.data(), nk * 4, hipMemcpyHostToDevice, stream));
    zero(d_n2, ((size_t)nk * count * 8 + 15) / 16 * 16);
    residual_norms(d_n2, d_vec, count, d_cnt, d_e, d_H, d_S, nk, stream);
    std::vector<double> n2((size_t)nk * count);
    AOSX_HIP(hipMemcpyAsync(n2.data(), d_n2, n2.size() * 8, hipMemcpyDeviceToHost, stream));
    AOSX_HIP(hipStreamSynchronize(stream));
    for (size_t i = 0; i < n2.size(); ++i) norm[i] = std::sqrt(n2[i]);
}

// ---------------------------------------------------------------- CISS band
namespace {
struct Subspace {
    int rmax = 0;                 // padded rank width (multiple of 32)
    std::vector<int> rank;        // per k
    std::vector<std::vector<float>> energies;
    std::vector<double> residual;
    float2 *ritz = nullptr;       // [nk][n][rmax] Ritz block Q C (when requested)
    std::vector<std::vector<float2>> vectors;  // selected Ritz vectors [n][count] (when requested)
};
}  // namespace

// Hermitian eigendecomposition of nk problems of order rk[k] <= rmax at A
// (leading dimension ld, `stride` elements apart): ascending eigenvalues into
// w[k][*], eigenvectors as the columns of Vout (rows and columns < rk[k]).
// A is left unchanged.
// tol: a sweep stops the iteration once the off-diagonal Frobenius norm is
// below tol times the total (the rounding floor of an r x r FP64 matrix is
// about r * 1e-16)
static void eigh_batched(Bundle &B, const c64 *A, int ld, int64_t stride, const int32_t *d_rk, int rmax, int nk,
                         c64 *Vout, double *w, double tol_default) {
    if (rmax > EIGH_MAX) throw Error(AOSX_SEARCH_DIMENSION, "projected problem exceeds 256 columns");
    hipStream_t st = B.stream;
    const int rr = rmax + (rmax & 1);
    const int64_t log_stride = (int64_t)EIGH_SWEEPS * std::max(1, rr - 1) * (rr / 2);
    Rot *log = B.buf<Rot>("e_log", (size_t)nk * log_stride);
    int32_t *rounds = B.buf<int32_t>("e_rounds", nk);
    int32_t *rank = B.buf<int32_t>("e_rank", (size_t)nk * EIGH_MAX);
    const double tol = tol_default;
    if (rr <= EIGH_LDS_MAX) {
        k_jacobi_rotate<true><<<nk, 1024, 0, st>>>(A, ld, stride, d_rk, nullptr, log, log_stride, rounds, w, rank,
                                                   EIGH_SWEEPS, tol);
    } else {
        c64 *scratch = B.buf<c64>("e_packed", (size_t)nk * (EIGH_MAX * (EIGH_MAX + 1) / 2));
        k_jacobi_rotate<false><<<nk, 1024, 0, st>>>(A, ld, stride, d_rk, scratch, log, log_stride, rounds, w, rank,
                                                    EIGH_SWEEPS, tol);
    }
    AOSX_LAUNCHED();
    if (std::getenv("AOSX_DEBUG")) {
        std::vector<int32_t> hr(nk);
        AOSX_HIP(hipMemcpyAsync(hr.data(), rounds, nk * 4, hipMemcpyDeviceToHost, st));
        AOSX_HIP(hipStreamSynchronize(st));
        std::fprintf(stderr, "AOSX_DEBUG|jacobi|rmax=%d|sweeps:", rmax);
        for (int k = 0; k < nk; ++k) std::fprintf(stderr, " %d", hr[k] / std::max(1, rr - 1));
        std::fprintf(stderr, "\n");
    }
    k_jacobi_vectors<<<dim3((rmax + EIGH_ROWS - 1) / EIGH_ROWS, nk), 256, 0, st>>>(Vout, ld, stride, d_rk, log,
                                                                                    log_stride, rounds, rank);
    AOSX_LAUNCHED();
}

static void compress_and_project(Bundle &B, int nk, const float2 *d_M, int P, const float2 *d_H,
                                 const float2 *d_S, const CissParams &p, const double *lower, const double *upper,
                                 Subspace &out, double *stage, const char *ritz_name, bool want_vectors = false) {
    const int64_t n = B.n;
    hipStream_t st = B.stream;
    double t0 = now();
    std::vector<int32_t> iota_k(nk);
    for (int k = 0; k < nk; ++k) iota_k[k] = k;
    int32_t *d_iota = B.buf<int32_t>("c_iota", nk);
    AOSX_HIP(hipMemcpyAsync(d_iota, iota_k.data(), nk * 4, hipMemcpyHostToDevice, st));
    // SM = S M, G = M^H S M (FP64, hermitised)
    float2 *d_SM = B.buf<float2>("c_SM", (size_t)nk * n * P);
    B.spmm(d_SM, d_S, d_iota, d_M, d_iota, P, nk);
    auto gram = [&](c64 *G, int ldg, const float2 *A, const float2 *Bm, int Pc, int64_t stride_a, int herm) {
        const int chunks = (int)std::min<int64_t>(64, std::max<int64_t>(1, n / 256));
        const int rpc = (int)(((n + chunks - 1) / chunks + 31) / 32 * 32);
        c64 *part = B.buf<c64>("c_part", (size_t)nk * chunks * Pc * Pc);
        const int tiles = ((Pc + 31) / 32) * ((Pc + 31) / 32);
        k_gram_partial<<<dim3(tiles, chunks, nk), 256, 0, st>>>(part, A, stride_a, Pc, Bm, stride_a, Pc, n, rpc);
        AOSX_LAUNCHED();
        k_gram_reduce<<<dim3(blocks_for((int64_t)Pc * Pc, 256, 1024), nk), 256, 0, st>>>(
            G, ldg, (int64_t)ldg * ldg, part, Pc, Pc, chunks, herm);
        AOSX_LAUNCHED();
    };
    c64 *d_G = B.buf<c64>("c_G", (size_t)nk * P * P);
    gram(d_G, P, d_M, d_SM, P, (int64_t)n * P, 1);
    // eigh(G)
    c64 *d_U = B.buf<c64>("c_U", (size_t)nk * P * P);
    double *d_w = B.buf<double>("c_w", (size_t)nk * P);
    std::vector<int32_t> rP(nk, P);
    int32_t *d_rP = B.buf<int32_t>("c_rP", nk);
    AOSX_HIP(hipMemcpyAsync(d_rP, rP.data(), nk * 4, hipMemcpyHostToDevice, st));
    // the Gram serves a rank cut at rank_tol (1e-5) and the basis of the
    // retained directions, applied to complex64 data: 1e-10 is ample, and it
    // halves the sweeps of 1e-15, which the rounding floor of a 128 x 128
    // matrix barely admits
    eigh_batched(B, d_G, P, (int64_t)P * P, d_rP, P, nk, d_U, d_w, 1e-10);
    std::vector<double> w((size_t)nk * P);
    AOSX_HIP(hipMemcpyAsync(w.data(), d_w, w.size() * 8, hipMemcpyDeviceToHost, st));
    AOSX_HIP(hipStreamSynchronize(st));
    if (std::getenv("AOSX_DEBUG")) {
        for (int k = 0; k < nk; ++k) {
            const double top = w[(size_t)k * P + P - 1];
            std::fprintf(stderr, "AOSX_DEBUG|gram|k=%d|P=%d|lambda/top:", k, P);
            for (int i = P - 1; i >= 0 && i >= P - 140; i -= 4)
                std::fprintf(stderr, " %d:%.1e", P - 1 - i, w[(size_t)k * P + i] / top);
            std::fprintf(stderr, "\n");
        }
    }
    // rank truncation (ciss.py:107-126)
    out.rank.assign(nk, 0);
    std::vector<std::vector<int32_t>> keep(nk);
    std::vector<std::vector<double>> sig(nk);
    int rmax = 0;
    for (int k = 0; k < nk; ++k) {
        double scale = 0.0;
        for (int i = 0; i < P; ++i) scale = std::max(scale, std::fabs(w[(size_t)k * P + i]));
        if (scale == 0.0 || !std::isfinite(scale)) throw Error(AOSX_ZERO_RANK, "CISS contour moments have zero numerical rank");
        for (int i = 0; i < P; ++i)
            if (w[(size_t)k * P + i] > scale * p.rank_tol) {
                keep[k].push_back(i);
                sig[k].push_back(std::sqrt(w[(size_t)k * P + i]));
            }
        if (keep[k].empty()) throw Error(AOSX_NO_STABLE_SUBSPACE, "CISS contour moments have no positive stable subspace");
        out.rank[k] = (int)keep[k].size();
        rmax = std::max(rmax, out.rank[k]);
    }
    const int R = (rmax + 31) / 32 * 32;
    out.rmax = R;
    std::vector<int32_t> cols((size_t)nk * R, 0), rk(nk);
    std::vector<double> scl((size_t)nk * R, 0.0), sigma((size_t)nk * R, 1.0);
    for (int k = 0; k < nk; ++k) {
        rk[k] = out.rank[k];
        for (int j = 0; j < out.rank[k]; ++j) {
            cols[(size_t)k * R + j] = keep[k][j];
            scl[(size_t)k * R + j] = 1.0 / sig[k][j];
            sigma[(size_t)k * R + j] = sig[k][j];
        }
    }
    int32_t *d_cols = B.buf<int32_t>("c_cols", (size_t)nk * R);
    double *d_scl = B.buf<double>("c_scl", (size_t)nk * R);
    double *d_sigma = B.buf<double>("c_sigma", (size_t)nk * R);
    int32_t *d_rk = B.buf<int32_t>("c_rk", nk);
    AOSX_HIP(hipMemcpyAsync(d_cols, cols.data(), cols.size() * 4, hipMemcpyHostToDevice, st));
    AOSX_HIP(hipMemcpyAsync(d_scl, scl.data(), scl.size() * 8, hipMemcpyHostToDevice, st));
    AOSX_HIP(hipMemcpyAsync(d_sigma, sigma.data(), sigma.size() * 8, hipMemcpyHostToDevice, st));
    AOSX_HIP(hipMemcpyAsync(d_rk, rk.data(), nk * 4, hipMemcpyHostToDevice, st));
    // T = U[:, keep] diag(lambda^-1/2)   (P x R), Q = M T (n x R, complex64)
    c64 *d_T = B.buf<c64>("c_T", (size_t)nk * P * R);
    k_gather_cols<<<dim3(blocks_for((int64_t)P * R, 256, 4096), nk), 256, 0, st>>>(
        d_T, R, (int64_t)P * R, d_U, P, (int64_t)P * P, d_cols, d_scl, R, d_rk, P);
    AOSX_LAUNCHED();
    float2 *d_Q = B.buf<float2>("c_Q", (size_t)nk * n * R);
    k_tall_gemm<float2><<<dim3((unsigned)((n + 7) / 8), R / 32, nk), 256, 0, st>>>(
        d_Q, (int64_t)n * R, d_M, (int64_t)n * P, P, d_T, R, (int64_t)P * R, R, n);
    AOSX_LAUNCHED();
    AOSX_HIP(hipStreamSynchronize(st));
    double t1 = now();
    stage[6] += t1 - t0;
    // Rayleigh-Ritz: Hh = Q^H H Q, Sh = Q^H S Q (FP64, hermitised)
    float2 *d_HQ = B.buf<float2>("c_HQ", (size_t)nk * n * R);
    float2 *d_SQ = B.buf<float2>("c_SQ", (size_t)nk * n * R);
    B.spmm(d_HQ, d_H, d_iota, d_Q, d_iota, R, nk);
    B.spmm(d_SQ, d_S, d_iota, d_Q, d_iota, R, nk);
    const int64_t RR = (int64_t)R * R;
    c64 *d_Hh = B.buf<c64>("c_Hh", (size_t)nk * RR);
    c64 *d_Sh = B.buf<c64>("c_Sh", (size_t)nk * RR);
    gram(d_Hh, R, d_Q, d_HQ, R, (int64_t)n * R, 1);
    gram(d_Sh, R, d_Q, d_SQ, R, (int64_t)n * R, 1);
    // Cholesky Sh = L L^H, W = L^-1 Hh L^-H, eigh(W), C = L^-H Y
    int32_t *d_status = B.buf<int32_t>("c_status", nk);
    k_cholesky<<<nk, 1024, 0, st>>>(d_Sh, d_rk, R, RR, d_status);
    AOSX_LAUNCHED();
    std::vector<int32_t> status(nk);
    AOSX_HIP(hipMemcpyAsync(status.data(), d_status, nk * 4, hipMemcpyDeviceToHost, st));
    AOSX_HIP(hipStreamSynchronize(st));
    for (int k = 0; k < nk; ++k)
        if (status[k]) throw Error(AOSX_OVERLAP_NOT_PD, "CISS reduced overlap is not positive definite");
    c64 *d_Li = B.buf<c64>("c_Li", (size_t)nk * RR);
    c64 *d_Tm = B.buf<c64>("c_Tm", (size_t)nk * RR);
    c64 *d_Wm = B.buf<c64>("c_Wm", (size_t)nk * RR);
    c64 *d_C = B.buf<c64>("c_C", (size_t)nk * RR);
    k_fill_zero<<<blocks_for(nk * RR, 256, 4096), 256, 0, st>>>(reinterpret_cast<float4 *>(d_Li), nk * RR);
    k_fill_zero<<<blocks_for(nk * RR, 256, 4096), 256, 0, st>>>(reinterpret_cast<float4 *>(d_Tm), nk * RR);
    k_fill_zero<<<blocks_for(nk * RR, 256, 4096), 256, 0, st>>>(reinterpret_cast<float4 *>(d_Wm), nk * RR);
    k_fill_zero<<<blocks_for(nk * RR, 256, 4096), 256, 0, st>>>(reinterpret_cast<float4 *>(d_C), nk * RR);
    AOSX_LAUNCHED();
    k_trinv_lower<<<dim3((R + 255) / 256, nk), 256, 0, st>>>(d_Li, d_Sh, d_rk, R, RR);
    AOSX_LAUNCHED();
    const dim3 sg((R + 63) / 64, R, nk);
    k_small_gemm<<<sg, 64, 0, st>>>(d_Tm, d_Li, 0, d_Hh, 0, d_rk, R, RR);
    k_small_gemm<<<sg, 64, 0, st>>>(d_Wm, d_Tm, 0, d_Li, 1, d_rk, R, RR);
    AOSX_LAUNCHED();
    c64 *d_Y = B.buf<c64>("c_Y", (size_t)nk * RR);
    double *d_E = B.buf<double>("c_E", (size_t)nk * R);
    k_fill_zero<<<blocks_for(nk * RR, 256, 4096), 256, 0, st>>>(reinterpret_cast<float4 *>(d_Y), nk * RR);
    AOSX_LAUNCHED();
    eigh_batched(B, d_Wm, R, RR, d_rk, rmax, nk, d_Y, d_E, 1e-13);  // energies to ~1e-12 eV, r <= 256
    k_small_gemm<<<sg, 64, 0, st>>>(d_C, d_Li, 1, d_Y, 0, d_rk, R, RR);
    AOSX_LAUNCHED();
    double *d_score = B.buf<double>("c_score", (size_t)nk * R);
    k_ghost_scores<<<dim3((R + 63) / 64, nk), 64, 0, st>>>(d_score, d_C, R, RR, d_sigma, R, d_rk);
    AOSX_LAUNCHED();
    std::vector<double> E((size_t)nk * R), score((size_t)nk * R);
    AOSX_HIP(hipMemcpyAsync(E.data(), d_E, E.size() * 8, hipMemcpyDeviceToHost, st));
    AOSX_HIP(hipMemcpyAsync(score.data(), d_score, score.size() * 8, hipMemcpyDeviceToHost, st));
    AOSX_HIP(hipStreamSynchronize(st));
    if (ritz_name) {
        // all Ritz vectors Q C (complex64, n x R), zero beyond each k's rank
        out.ritz = B.buf<float2>(ritz_name, (size_t)nk * n * R);
        k_tall_gemm<float2><<<dim3((unsigned)((n + 7) / 8), R / 32, nk), 256, 0, st>>>(
            out.ritz, (int64_t)n * R, d_Q, (int64_t)n * R, R, d_C, R, RR, R, n);
        AOSX_LAUNCHED();
    }
    double t2 = now();
    stage[7] += t2 - t1;
    // ghost filter and strict interval selection (ciss.py:350-361)
    std::vector<std::vector<int32_t>> sel(nk);
    out.energies.assign(nk, {});
    int mmax = 0;
    for (int k = 0; k < nk; ++k) {
        double top = 0.0;
        for (int j = 0; j < out.rank[k]; ++j) top = std::max(top, score[(size_t)k * R + j]);
        if (top == 0.0 || !std::isfinite(top))
            throw Error(AOSX_GHOST_FILTER, "CISS ghost filter produced no positive finite score");
        for (int j = 0; j < out.rank[k]; ++j) {
            const double sc = score[(size_t)k * R + j];
            const float e32 = (float)E[(size_t)k * R + j];
            if (std::isfinite(sc) && sc >= p.spurious * top && (double)e32 > lower[k] && (double)e32 < upper[k]) {
                sel[k].push_back(j);
                out.energies[k].push_back(e32);
            }
        }
        mmax = std::max<int>(mmax, (int)sel[k].size());
    }
    out.residual.assign(nk, 0.0);
    if (mmax > 0) {
        const int Mc = mmax;
        std::vector<int32_t> scols((size_t)nk * Mc, 0), scnt(nk);
        std::vector<double> sone((size_t)nk * Mc, 1.0), senergy((size_t)nk * Mc, 0.0);
        for (int k = 0; k < nk; ++k) {
            scnt[k] = (int32_t)sel[k].size();
            for (size_t j = 0; j < sel[k].size(); ++j) {
                scols[(size_t)k * Mc + j] = sel[k][j];
                senergy[(size_t)k * Mc + j] = (double)out.energies[k][j];
            }
        }
        int32_t *d_scols = B.buf<int32_t>("r_cols", (size_t)nk * Mc);
        int32_t *d_scnt = B.buf<int32_t>("r_cnt", nk);
        double *d_sone = B.buf<double>("r_one", (size_t)nk * Mc);
        double *d_sen = B.buf<double>("r_energy", (size_t)nk * Mc);
        AOSX_HIP(hipMemcpyAsync(d_scols, scols.data(), scols.size() * 4, hipMemcpyHostToDevice, st));
        AOSX_HIP(hipMemcpyAsync(d_scnt, scnt.data(), nk * 4, hipMemcpyHostToDevice, st));
        AOSX_HIP(hipMemcpyAsync(d_sone, sone.data(), sone.size() * 8, hipMemcpyHostToDevice, st));
        AOSX_HIP(hipMemcpyAsync(d_sen, senergy.data(), senergy.size() * 8, hipMemcpyHostToDevice, st));
        c64 *d_Cs = B.buf<c64>("r_Cs", (size_t)nk * R * Mc);
        k_gather_cols<<<dim3(blocks_for((int64_t)R * Mc, 256, 4096), nk), 256, 0, st>>>(
            d_Cs, Mc, (int64_t)R * Mc, d_C, R, RR, d_scols, d_sone, Mc, d_scnt, R);
        AOSX_LAUNCHED();
        double2 *d_vec = B.buf<double2>("r_vec", (size_t)nk * n * Mc);
        k_tall_gemm<double2><<<dim3((unsigned)((n + 7) / 8), (Mc + 31) / 32, nk), 256, 0, st>>>(
            d_vec, (int64_t)n * Mc, d_Q, (int64_t)n * R, R, d_Cs, Mc, (int64_t)R * Mc, Mc, n);
        AOSX_LAUNCHED();
        double *d_n2 = B.buf<double>("r_norm2", (size_t)nk * Mc + 2);
        const int64_t z4 = ((int64_t)nk * Mc * 8 + 15) / 16;
        k_fill_zero<<<blocks_for(z4, 256, 1024), 256, 0, st>>>(reinterpret_cast<float4 *>(d_n2), z4);
        B.residual_norms(d_n2, d_vec, Mc, d_scnt, d_sen, d_H, d_S, nk, st);
        std::vector<double> n2((size_t)nk * Mc);
        AOSX_HIP(hipMemcpyAsync(n2.data(), d_n2, n2.size() * 8, hipMemcpyDeviceToHost, st));
        std::vector<double2> hv;
        if (want_vectors) {
            hv.resize((size_t)nk * n * Mc);
            AOSX_HIP(hipMemcpyAsync(hv.data(), d_vec, hv.size() * sizeof(double2), hipMemcpyDeviceToHost, st));
        }
        AOSX_HIP(hipStreamSynchronize(st));
        for (int k = 0; k < nk; ++k)
            for (int j = 0; j < scnt[k]; ++j)
                out.residual[k] = std::max(out.residual[k], std::sqrt(n2[(size_t)k * Mc + j]));
        if (want_vectors) {
            out.vectors.assign(nk, {});
            for (int k = 0; k < nk; ++k) {
                const int c = scnt[k];
                std::vector<float2> &v = out.vectors[k];
                v.resize((size_t)n * c);
                for (int64_t row = 0; row < n; ++row)
                    for (int j = 0; j < c; ++j) {
                        const double2 x = hv[((size_t)k * n + row) * Mc + j];
                        v[(size_t)row * c + j] = make_float2((float)x.x, (float)x.y);
                    }
            }
        }
    } else if (want_vectors) {
        out.vectors.assign(nk, {});
    }
    stage[8] += now() - t2;
}

void Bundle::band(const CissParams &p, const double *kpts, const float2 *h_in, const float2 *s_in,
                  const float2 *probes, const float2 *augmentation, int augmentation_width, CissOutput &out) {
    AOSX_HIP(hipSetDevice(device));
    const double t_start = now();
    const int N = p.integration_points, L = p.block_size, Ms = p.moment_size, P = L * Ms;
    if (p.conjugate_pairs && (N % 2)) throw Error(AOSX_INVALID_SETTINGS, "conjugate pairs need an even integration_points");
    if (L % 32) throw Error(AOSX_INVALID_SETTINGS, "block_size must be a multiple of 32 on this device path");
    // contour per k point (ciss.py:81-104), FP64 then complex64; every k has
    // the request's interval unless per-k intervals are given (serving batches
    // k points of different jobs into one factorisation launch)
    struct Contour {
        std::vector<std::complex<double>> z;
        std::vector<std::complex<float>> t32, w32;
        double lower, upper;
    };
    auto make_contour = [&](double lower, double upper) {
        Contour c;
        c.lower = lower;
        c.upper = upper;
        const double lo = lower - 2.0 * p.guard, hi = upper + 2.0 * p.guard;
        const double cen = 0.5 * (lo + hi), rh = 0.5 * (hi - lo), rv = rh * p.vscale;
        std::vector<std::complex<double>> tc(N), wc(N);
        c.z.resize(N);
        for (int j = 0; j < N; ++j) {
            const double th = 2.0 * M_PI * (j + 0.5) / N;
            tc[j] = {std::cos(th), p.vscale * std::sin(th)};
            c.z[j] = cen + rh * tc[j];
            wc[j] = std::complex<double>(rv * std::cos(th), rh * std::sin(th)) / (double)N;
        }
        if (p.conjugate_pairs)
            for (int j = N / 2; j < N; ++j) {
                c.z[j] = std::conj(c.z[N - 1 - j]);
                tc[j] = std::conj(tc[N - 1 - j]);
                wc[j] = std::conj(wc[N - 1 - j]);
            }
        c.t32.resize(N);
        c.w32.resize(N);
        for (int j = 0; j < N; ++j) {
            c.t32[j] = {(float)tc[j].real(), (float)tc[j].imag()};
            c.w32[j] = {(float)wc[j].real(), (float)wc[j].imag()};
        }
        return c;
    };
    std::vector<Contour> contours(p.nk);
    for (int64_t gk = 0; gk < p.nk; ++gk)
        contours[gk] = p.intervals ? make_contour(p.intervals[2 * gk], p.intervals[2 * gk + 1])
                                   : make_contour(p.lower, p.upper);
    auto f2 = [](std::complex<double> v) { return make_float2((float)v.real(), (float)v.imag()); };
    // moment coefficient of shift j, moment m for k point gk: -w_j t_j^m in
    // complex64 arithmetic
    auto coef = [&](int64_t gk, int j, int m) {
        const Contour &c = contours[gk];
        std::complex<float> tp(1.f, 0.f);
        for (int q = 0; q < m; ++q) tp *= c.t32[j];
        const std::complex<float> v = -c.w32[j] * tp;
        return make_float2(v.real(), v.imag());
    };
    float2 *d_V = buf<float2>("probes", (size_t)n * L);
    AOSX_HIP(hipMemcpyAsync(d_V, probes, (size_t)n * L * sizeof(float2), hipMemcpyHostToDevice, stream));
    const int64_t nk_total = p.nk;
    out.energies.assign(nk_total, {});
    out.max_residual.assign(nk_total, 0.0);
    out.rank.assign(nk_total, 0);
    out.saturated.assign(nk_total, 0);
    out.pass_rank.assign(nk_total, 0);
    out.pass_width.assign(nk_total, 0);
    if (p.want_vectors) out.vectors.assign(nk_total, {});
    const int K = std::max(1, p.k_batch);
    const int nfac = p.conjugate_pairs ? N / 2 : N;
    const int cb = std::max(1, std::min(p.contour_batch_size, nfac));
    const bool ir = p.precision != AOSX_PRECISION_FP32;
    const int ir_steps = ir ? std::max(0, p.refinement_steps) : 0;
    // refinement schedule: with filter passes ahead, the moment pass and the
    // intermediate passes only need a subspace, not converged solves -- one
    // step there, the full count on the last filter application
    const int ir_first = p.refinement_passes > 0 ? std::min(ir_steps, 1) : ir_steps;
    // AOSX_PROGRESS: one line per factorisation group and pass on stderr
    const bool progress = std::getenv("AOSX_PROGRESS") != nullptr;
    auto report = [&](int64_t kb, int pass, int g, int ng) {
        if (!progress) return;
        std::fprintf(stderr, "[aosx] k %lld pass %d group %d/%d done (%.1f s)\n", (long long)kb, pass, g + 1, ng,
                     now() - t_start);
        std::fflush(stderr);
    };
    out.stages[1] += now() - t_start;
    for (int64_t kb = 0; kb < nk_total; kb += K) {
        const int Kc = (int)std::min<int64_t>(K, nk_total - kb);
        double t0 = now();
        float2 *d_H = buf<float2>("H", (size_t)Kc * nnz);
        float2 *d_S = buf<float2>("S", (size_t)Kc * nnz);
        if (h_in) {
            AOSX_HIP(hipMemcpyAsync(d_H, h_in + kb * nnz, (size_t)Kc * nnz * 8, hipMemcpyHostToDevice, stream));
            AOSX_HIP(hipMemcpyAsync(d_S, s_in + kb * nnz, (size_t)Kc * nnz * 8, hipMemcpyHostToDevice, stream));
        } else {
            assemble(kpts + 3 * kb, Kc, d_H, d_S);
        }
        AOSX_HIP(hipStreamSynchronize(stream));
        double t1 = now();
        out.stages[0] += t1 - t0;
        // RHS = S V per k
        std::vector<int32_t> iota(Kc), zeros(Kc, 0);
        for (int k = 0; k < Kc; ++k) iota[k] = k;
        int32_t *d_iota = buf<int32_t>("b_iota", Kc), *d_zero = buf<int32_t>("b_zero", Kc);
        AOSX_HIP(hipMemcpyAsync(d_iota, iota.data(), Kc * 4, hipMemcpyHostToDevice, stream));
        AOSX_HIP(hipMemcpyAsync(d_zero, zeros.data(), Kc * 4, hipMemcpyHostToDevice, stream));
        float2 *d_rhs = buf<float2>("b_rhs", (size_t)Kc * n * L);
        spmm(d_rhs, d_S, d_iota, d_V, d_zero, L, Kc);
        float2 *d_M = buf<float2>("b_M", (size_t)Kc * n * P);
        const int64_t m4 = (int64_t)Kc * n * P * 8 / 16;
        k_fill_zero<<<blocks_for(m4, 256, 8192), 256, 0, stream>>>(reinterpret_cast<float4 *>(d_M), m4);
        AOSX_LAUNCHED();
        // factorisation groups: all k of the batch x cb shifts
        const int G = Kc * cb;
        const int ngroups = (nfac + cb - 1) / cb;
        const size_t ff = factor_floats();
        const size_t dif = (size_t)ts.n_diag * 2 * 2 * 64 * 64;
        // The filter passes reuse the factorisations when all of the batch's
        // fit (with 8 GiB to spare for vectors and split panels) and the
        // request allows that many; otherwise the factorisations are streamed:
        // each pass factors every shift group again into one group's slots.
        bool resident = p.refinement_passes > 0 && (p.resident_limit <= 0 || Kc * nfac <= p.resident_limit);
        if (resident) {
            size_t free_b = 0, total_b = 0;
            AOSX_HIP(hipMemGetInfo(&free_b, &total_b));
            size_t held = 0;
            for (const char *name : {"tiles", "dinv"}) {
                auto it = ws.find(name);
                if (it != ws.end()) held += it->second.bytes;
            }
            const double need = (double)Kc * nfac * (double)(ff + dif) * 4.0 + 8.0 * (1ull << 30);
            resident = need <= (double)(free_b + held);
        }
        const int slots = resident ? Kc * nfac : G;
        float *d_tiles = buf<float>("tiles", (size_t)slots * ff);
        float *d_dinv = buf<float>("dinv", (size_t)slots * dif);
        struct Group {
            std::vector<int32_t> kidx;
            std::vector<int> shift;
            std::vector<float2> z, zc;
            // with conjugate pairs, both shifts of every pair in one solve:
            // the group's k indices twice, its shifts then their conjugates
            std::vector<int32_t> kidx2;
            std::vector<float2> z2;
            std::vector<uint8_t> one_product;  // per-shift precision: one BF16 product
            int slot0;
        };
        std::vector<Group> groups(ngroups);
        for (int g = 0; g < ngroups; ++g) {
            Group &gr = groups[g];
            const int j0 = g * cb, j1 = std::min(nfac, j0 + cb);
            for (int k = 0; k < Kc; ++k)
                for (int j = j0; j < j1; ++j) {
                    gr.kidx.push_back(k);
                    gr.shift.push_back(j);
                    gr.z.push_back(f2(contours[kb + k].z[j]));
                    gr.zc.push_back(f2(std::conj(contours[kb + k].z[j])));
                    // the contour point's distance from the real axis is
                    // rv |sin theta_j|; the point and its conjugate share it
                    const double th = 2.0 * M_PI * (j + 0.5) / N;
                    gr.one_product.push_back(std::fabs(std::sin(th)) >= p.far_threshold ? 1 : 0);
                }
            gr.slot0 = resident ? g * G : 0;
            gr.kidx2 = gr.kidx;
            gr.kidx2.insert(gr.kidx2.end(), gr.kidx.begin(), gr.kidx.end());
            gr.z2 = gr.z;
            gr.z2.insert(gr.z2.end(), gr.zc.begin(), gr.zc.end());
        }
        const int xsets = p.conjugate_pairs ? 2 : 1;  // solution blocks per factorisation
        float2 *d_X = buf<float2>("b_X", (size_t)xsets * G * n * std::max(L, 32));
        float2 *d_coef = buf<float2>("b_coef", (size_t)G * Ms);
        auto accumulate = [&](const Group &gr, bool conj_side, float2 *dst, int width, int msize, const float2 *X) {
            // dst[k] += sum over the group's shifts of k of coef * X
            const int per_k = (int)gr.kidx.size() / Kc;
            std::vector<float2> cf(gr.kidx.size() * msize);
            for (size_t b = 0; b < gr.kidx.size(); ++b)
                for (int m = 0; m < msize; ++m) {
                    const int64_t gk = kb + gr.kidx[b];
                    cf[b * msize + m] = conj_side ? coef(gk, N - 1 - gr.shift[b], m) : coef(gk, gr.shift[b], m);
                }
            AOSX_HIP(hipMemcpyAsync(d_coef, cf.data(), cf.size() * sizeof(float2), hipMemcpyHostToDevice, stream));
            for (int k = 0; k < Kc; ++k) {
                k_moments<<<blocks_for(n * width, 256, 4096), 256, 0, stream>>>(
                    dst + (size_t)k * n * width * msize, X + (size_t)k * per_k * n * width, per_k, width, msize,
                    d_coef + (size_t)k * per_k * msize, n);
                AOSX_LAUNCHED();
            }
        };
        // pivot-replacement counters, one per resident factorisation slot
        int *d_repl = buf<int>("replaced", (size_t)std::max(slots, 256));
        zero(d_repl, (size_t)(slots + 3) / 4 * 16);
        for (int g = 0; g < ngroups; ++g) {
            Group &gr = groups[g];
            const int gsz = (int)gr.kidx.size();
            float *tl = d_tiles + (size_t)gr.slot0 * ff;
            float *di = d_dinv + (size_t)gr.slot0 * dif;
            double ta = now();
            factor(gsz, gr.kidx, gr.z, p.precision, tl, di, d_repl + gr.slot0, &gr.one_product);
            AOSX_HIP(hipStreamSynchronize(stream));
            double tb = now();
            out.stages[2] += tb - ta;
            out.factorizations += gsz;
            if (p.conjugate_pairs) {
                solve_full(2 * gsz, L, d_rhs, gr.kidx2, d_H, d_S, gr.kidx2, gr.z2, tl, di, false, ir_first, d_X,
                           nullptr, p.precision, gsz);
                accumulate(gr, false, d_M, L, Ms, d_X);
                accumulate(gr, true, d_M, L, Ms, d_X + (size_t)gsz * n * L);
            } else {
                solve_full(gsz, L, d_rhs, gr.kidx, d_H, d_S, gr.kidx, gr.z, tl, di, false, ir_first, d_X, nullptr,
                           p.precision);
                accumulate(gr, false, d_M, L, Ms, d_X);
            }
            AOSX_HIP(hipStreamSynchronize(stream));
            out.stages[3] += now() - tb;
            report(kb, 0, g, ngroups);
        }
        std::vector<int> repl(slots);
        // on this bundle's (non-blocking) stream: a copy on the legacy stream
        // is refused while another thread captures a graph
        AOSX_HIP(hipMemcpyAsync(repl.data(), d_repl, slots * sizeof(int), hipMemcpyDeviceToHost, stream));
        AOSX_HIP(hipStreamSynchronize(stream));
        for (int v : repl) out.replaced_pivots += v;
        // subspace stage
        Subspace sub;
        const char *ritz_names[2] = {"b_ritz0", "b_ritz1"};
        std::vector<double> klo(Kc), khi(Kc);
        for (int k = 0; k < Kc; ++k) {
            klo[k] = contours[kb + k].lower;
            khi[k] = contours[kb + k].upper;
        }
        compress_and_project(*this, Kc, d_M, P, d_H, d_S, p, klo.data(), khi.data(), sub, out.stages,
                             p.refinement_passes > 0 ? ritz_names[0] : nullptr,
                             p.want_vectors && p.refinement_passes == 0);
        std::vector<int> rank0 = sub.rank;
        std::vector<int> prank(Kc, 0), pwidth(Kc, 0);
        // filter passes: the block [Ritz vectors | augmentation] is filtered
        // again through the resident factorisations.  The augmentation starts
        // as a second independent probe block and is itself re-filtered every
        // pass: it supplies in-window directions that the monomial moment
        // basis loses to the rank truncation when the in-window states bunch
        // at one end of the window (their moments become nearly collinear).
        float2 *d_aug = nullptr;
        const int A = augmentation ? augmentation_width : 0;
        if (p.refinement_passes > 0 && augmentation) {
            d_aug = buf<float2>("p_aug", (size_t)Kc * n * A);
            for (int k = 0; k < Kc; ++k)
                AOSX_HIP(hipMemcpyAsync(d_aug + (size_t)k * n * A, augmentation, (size_t)n * A * sizeof(float2),
                                        hipMemcpyHostToDevice, stream));
        }
        for (int pass = 0; pass < p.refinement_passes; ++pass) {
            double tp = now();
            const int Rpad = sub.rmax;  // column stride of the Ritz block (multiple of 32)
            int R = 0;                  // widest rank: the columns that can be nonzero
            for (int k = 0; k < Kc; ++k) R = std::max(R, sub.rank[k]);
            // The block is [Ritz vectors | augmentation].  The augmentation is
            // trimmed (to no fewer than 32 columns) so that the block keeps to
            // 128 columns when the rank allows: the solves then sweep two
            // 64-column chunks instead of three and the projected problems stay
            // within the LDS-resident eigensolver.
            int Ause = 0;
            if (d_aug) {
                Ause = A;
                if (R + A > 128 && R < 128) Ause = std::max(std::min(A, 32), 128 - R);
                Ause = std::max(0, std::min(Ause, 256 - R));
            }
            const bool augment = Ause > 0;
            const int W2 = R + Ause;
            const bool last_pass = pass + 1 == p.refinement_passes;
            const int ir_pass = last_pass ? ir_steps : std::min(ir_steps, 1);
            float2 *d_blk = buf<float2>("p_blk", (size_t)Kc * n * W2);
            k_copy_cols<<<dim3(blocks_for(n * R, 256, 4096), Kc), 256, 0, stream>>>(d_blk, W2, 0, sub.ritz, Rpad, 0, R,
                                                                                   n);
            AOSX_LAUNCHED();
            if (augment) {
                k_copy_cols<<<dim3(blocks_for(n * Ause, 256, 4096), Kc), 256, 0, stream>>>(d_blk, W2, R, d_aug, A, 0,
                                                                                          Ause, n);
                AOSX_LAUNCHED();
            }
            float2 *d_rhs2 = buf<float2>("p_rhs", (size_t)Kc * n * W2);
            spmm(d_rhs2, d_S, d_iota, d_blk, d_iota, W2, Kc);
            float2 *d_Y = buf<float2>("p_Y", (size_t)Kc * n * W2);
            const int64_t y4 = (int64_t)Kc * n * W2 * 8 / 16;
            k_fill_zero<<<blocks_for(y4, 256, 8192), 256, 0, stream>>>(reinterpret_cast<float4 *>(d_Y), y4);
            AOSX_LAUNCHED();
            d_X = buf<float2>("b_X", (size_t)xsets * G * n * W2);
            for (int g = 0; g < ngroups; ++g) {
                Group &gr = groups[g];
                const int gsz = (int)gr.kidx.size();
                float *tl = d_tiles + (size_t)gr.slot0 * ff;
                float *di = d_dinv + (size_t)gr.slot0 * dif;
                if (!resident) {  // streamed: factor this group again
                    const double ta = now();
                    factor(gsz, gr.kidx, gr.z, p.precision, tl, di, d_repl + gr.slot0, &gr.one_product);
                    AOSX_HIP(hipStreamSynchronize(stream));
                    out.stages[2] += now() - ta;
                    out.factorizations += gsz;
                }
                if (p.conjugate_pairs) {
                    solve_full(2 * gsz, W2, d_rhs2, gr.kidx2, d_H, d_S, gr.kidx2, gr.z2, tl, di, false, ir_pass, d_X,
                               nullptr, p.precision, gsz);
                    accumulate(gr, false, d_Y, W2, 1, d_X);
                    accumulate(gr, true, d_Y, W2, 1, d_X + (size_t)gsz * n * W2);
                } else {
                    solve_full(gsz, W2, d_rhs2, gr.kidx, d_H, d_S, gr.kidx, gr.z, tl, di, false, ir_pass, d_X,
                               nullptr, p.precision);
                    accumulate(gr, false, d_Y, W2, 1, d_X);
                }
                if (progress) {
                    AOSX_HIP(hipStreamSynchronize(stream));
                    report(kb, pass + 1, g, ngroups);
                }
            }
            if (augment) {
                k_copy_cols<<<dim3(blocks_for(n * Ause, 256, 4096), Kc), 256, 0, stream>>>(d_aug, A, 0, d_Y, W2, R,
                                                                                          Ause, n);
                AOSX_LAUNCHED();
            }
            AOSX_HIP(hipStreamSynchronize(stream));
            out.stages[9] += now() - tp;
            const bool more = pass + 1 < p.refinement_passes;
            Subspace next;
            compress_and_project(*this, Kc, d_Y, W2, d_H, d_S, p, klo.data(), khi.data(), next, out.stages,
                                 more ? ritz_names[(pass + 1) & 1] : nullptr, p.want_vectors && !more);
            // Capacity of the filtered space: the moment-pass directions plus
            // the fresh augmentation of the first pass.  Later passes add no
            // direction (their augmentation is the filtered first one, which
            // already lies in the previous pass's span), so every pass's rank
            // is at most rank0 + A; a pass that reaches it may miss states.
            for (int k = 0; k < Kc; ++k) {
                if (pass == 0) pwidth[k] = rank0[k] + Ause;
                prank[k] = next.rank[k];
            }
            if (std::getenv("AOSX_DEBUG"))
                for (int k = 0; k < Kc; ++k)
                    std::fprintf(stderr, "AOSX_DEBUG|pass %d|k=%d|W=%d|rank=%d|count=%zu|residual=%.3e\n", pass + 1, k,
                                 W2, next.rank[k], next.energies[k].size(), next.residual[k]);
            sub = next;
        }
        for (int k = 0; k < Kc; ++k) {
            out.energies[kb + k] = sub.energies[k];
            out.max_residual[kb + k] = sub.residual[k];
            out.rank[kb + k] = rank0[k];
            out.pass_rank[kb + k] = prank[k];
            out.pass_width[kb + k] = pwidth[k];
            // the moment space is full, or the last filter pass kept every
            // column of its block: more states may lie in the window than the
            // search space holds (ciss.py:342)
            out.saturated[kb + k] = rank0[k] == L * Ms || (p.refinement_passes > 0 && prank[k] >= pwidth[k]);
            if (p.want_vectors) out.vectors[kb + k] = std::move(sub.vectors[k]);
        }
    }
    out.stages[11] += now() - t_start;
}

}  // namespace aosx
