// Device-side helpers shared by the Band-path kernels (gfx950 / CDNA4).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace aosx {

constexpr int TILE = 64;          // factor tile edge (complex entries)
constexpr int TILE2 = TILE * TILE;
constexpr int PLANE = TILE2;      // floats per plane of a stored tile
constexpr int TILE_FLOATS = 2 * TILE2;  // re plane followed by im plane
constexpr int LDA = TILE + 1;     // LDS row stride of row-major A tiles (odd: no bank conflicts on column reads)

typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

struct c32 {
    float re, im;
};
struct c64 {
    double re, im;
};

__device__ __forceinline__ c64 cmul(c64 a, c64 b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ c64 cmulc(c64 a, c64 b) { return {a.re * b.re + a.im * b.im, a.re * b.im - a.im * b.re}; }  // conj(a)*b
__device__ __forceinline__ c64 cadd(c64 a, c64 b) { return {a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ c64 csub(c64 a, c64 b) { return {a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ c64 cconj(c64 a) { return {a.re, -a.im}; }
__device__ __forceinline__ double cabs2(c64 a) { return a.re * a.re + a.im * a.im; }

// Workgroup barrier that orders LDS accesses only: it waits for the wave's
// own LDS operations (lgkmcnt), not for its global loads, which stay in
// flight across it (a __syncthreads() fence covers global memory too and
// waits for every outstanding load).
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Per-matrix base of the tile storage: matrices are stored back to back.
__device__ __forceinline__ float *tile_ptr(float *tiles, int64_t mat_stride, int mat, int32_t id) {
    return tiles + (int64_t)mat * mat_stride + (int64_t)id * TILE_FLOATS;
}

}  // namespace aosx
