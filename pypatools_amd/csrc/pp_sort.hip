// Cell sort of the particle state: particles that gather from the same map
// cell end up in the same wavefront, so the 12 (FP32) or 24 (FP64) coefficient
// loads of a wave touch a few cache lines instead of 64 (L1 hits instead of L2
// traffic).  Keys are the coefficient-table cell of each particle; the radix
// sort is hipCUB's; permutation and inverse permutation are plain gathers and
// scatters.  `ids` carries each slot's original particle index so results and
// history rows come back in the caller's order.
#include <hipcub/hipcub.hpp>

#include "pp_internal.h"

namespace pp_sortk {

__device__ __forceinline__ uint32_t key_axis(const MapDesc& m, int a, double x, bool periodic) {
    if (m.n[a] <= 1) return 0;
    const double u = (x - m.lo[a]) * m.inv_d[a];
    const double nc = (double)m.nc[a];
    double f = floor(u);
    if (periodic) {
        f = f - floor(f * m.inv_nz) * nc;
        if (f >= nc) f -= nc;
        if (f < 0.0) f += nc;
    }
    f = fmin(fmax(f, 0.0), nc - 1.0);   // outside / NaN -> a valid edge cell
    return (uint32_t)f;
}

// Sub-cell position along one axis, in 2^bits bins (0 if the axis is inactive).
__device__ __forceinline__ uint32_t sub_axis(const MapDesc& m, int a, double x, int bits) {
    if (bits <= 0 || m.n[a] <= 1) return 0;
    const double u = (x - m.lo[a]) * m.inv_d[a];
    double f = (u - floor(u)) * (double)(1u << bits);
    f = fmin(fmax(f, 0.0), (double)((1u << bits) - 1));
    return (uint32_t)f;
}

__global__ void __launch_bounds__(256) k_keys(const MapDesc m, const double* __restrict__ x,
                                              const double* __restrict__ y, const double* __restrict__ z, int64_t n,
                                              const int32_t* __restrict__ group, int group_shift, int refine_axis,
                                              int refine_bits, uint32_t* __restrict__ keys,
                                              uint32_t* __restrict__ slots) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t a = key_axis(m, 0, x[i], false);
    const uint32_t b = key_axis(m, 1, y[i], false);
    const uint32_t c = key_axis(m, 2, z[i], m.periodic != 0);
    uint32_t k = (a * (uint32_t)m.nc[1] + b) * (uint32_t)m.nc[2] + c;
    if (refine_bits > 0) {
        const double pos = refine_axis == 0 ? x[i] : (refine_axis == 1 ? y[i] : z[i]);
        k = (k << refine_bits) | sub_axis(m, refine_axis, pos, refine_bits);
    }
    if (group) k |= (uint32_t)group[i] << group_shift;
    keys[i] = k;
    slots[i] = (uint32_t)i;
}

// The six state arrays and the id array, permuted together in one pass.
struct StateArrays {
    double* a[6];
    int32_t* ids;
};

__global__ void __launch_bounds__(256) k_gather_state(const StateArrays src, const uint32_t* __restrict__ perm,
                                                      int64_t n, const StateArrays dst) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t p = perm[i];
    if (p >= (uint64_t)n) return;
#pragma unroll
    for (int k = 0; k < 6; ++k) dst.a[k][i] = src.a[k][p];
    dst.ids[i] = src.ids[p];
}

__global__ void __launch_bounds__(256) k_scatter_state(const StateArrays src, int64_t n, const StateArrays dst) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t t = src.ids[i];
    if ((uint32_t)t >= (uint64_t)n) return;
#pragma unroll
    for (int k = 0; k < 6; ++k) dst.a[k][t] = src.a[k][i];
}

__global__ void __launch_bounds__(256) k_copy_state(const StateArrays src, int64_t n, const StateArrays dst,
                                                    int with_ids) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
#pragma unroll
    for (int k = 0; k < 6; ++k) dst.a[k][i] = src.a[k][i];
    dst.ids[i] = with_ids ? src.ids[i] : (int32_t)i;
}

__global__ void __launch_bounds__(256) k_iota(int32_t* ids, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) ids[i] = (int32_t)i;
}

}  // namespace pp_sortk

namespace pp_sort {

static inline unsigned grid_for(int64_t n) { return (unsigned)((n + 255) / 256); }

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// workspace: keys_in, keys_out, slots_in, slots_out (4n each), a staging copy of
// the state (6 x 8n + 4n), cub temp
static size_t staging_bytes(int64_t n) { return 6 * align256(8 * (size_t)n) + align256(4 * (size_t)n); }

hipError_t workspace_bytes(int64_t n, size_t* bytes) {
    size_t cub_bytes = 0;
    hipError_t e = hipcub::DeviceRadixSort::SortPairs(nullptr, cub_bytes, (uint32_t*)nullptr, (uint32_t*)nullptr,
                                                      (uint32_t*)nullptr, (uint32_t*)nullptr, (int)n, 0, 32);
    if (e != hipSuccess) return e;
    *bytes = 4 * align256(4 * (size_t)n) + staging_bytes(n) + align256(cub_bytes);
    return hipSuccess;
}

static pp_sortk::StateArrays staging(char* base, int64_t n) {
    pp_sortk::StateArrays t;
    for (int k = 0; k < 6; ++k) t.a[k] = reinterpret_cast<double*>(base + k * align256(8 * (size_t)n));
    t.ids = reinterpret_cast<int32_t*>(base + 6 * align256(8 * (size_t)n));
    return t;
}

static pp_sortk::StateArrays state_of(double* arrs[6], int32_t* ids) {
    pp_sortk::StateArrays t;
    for (int k = 0; k < 6; ++k) t.a[k] = arrs[k];
    t.ids = ids;
    return t;
}

static int key_bits(const MapDesc& m) {
    const uint64_t ncell = (uint64_t)m.nc[0] * m.nc[1] * m.nc[2];
    int b = 1;
    while (b < 32 && ((uint64_t)1 << b) < ncell) ++b;
    return b;
}

hipError_t sort_state(const MapDesc& m, double* arrs[6], int32_t* ids, int64_t n, void* ws, size_t ws_bytes,
                      hipStream_t s, const int32_t* group, int group_shift, int refine_axis, int refine_bits) {
    size_t need = 0;
    hipError_t e = workspace_bytes(n, &need);
    if (e != hipSuccess) return e;
    if (ws_bytes < need) return hipErrorInvalidValue;
    char* w = static_cast<char*>(ws);
    const size_t q = align256(4 * (size_t)n);
    uint32_t* keys_in = reinterpret_cast<uint32_t*>(w);
    uint32_t* keys_out = reinterpret_cast<uint32_t*>(w + q);
    uint32_t* slots_in = reinterpret_cast<uint32_t*>(w + 2 * q);
    uint32_t* perm = reinterpret_cast<uint32_t*>(w + 3 * q);
    const pp_sortk::StateArrays tmp = staging(w + 4 * q, n);
    char* cub_tmp = w + 4 * q + staging_bytes(n);
    size_t cub_bytes = ws_bytes - (4 * q + staging_bytes(n));
    hipLaunchKernelGGL(pp_sortk::k_keys, dim3(grid_for(n)), dim3(256), 0, s, m, arrs[0], arrs[1], arrs[2], n, group,
                       group_shift, refine_axis, refine_bits, keys_in, slots_in);
    const int bits = group ? 32 : (key_bits(m) + refine_bits > 32 ? 32 : key_bits(m) + refine_bits);
    e = hipcub::DeviceRadixSort::SortPairs(cub_tmp, cub_bytes, keys_in, keys_out, slots_in, perm, (int)n, 0, bits,
                                           s);
    if (e != hipSuccess) return e;
    const pp_sortk::StateArrays st = state_of(arrs, ids);
    hipLaunchKernelGGL(pp_sortk::k_gather_state, dim3(grid_for(n)), dim3(256), 0, s, st, perm, n, tmp);
    hipLaunchKernelGGL(pp_sortk::k_copy_state, dim3(grid_for(n)), dim3(256), 0, s, tmp, n, st, 1);
    return hipGetLastError();
}

hipError_t unsort_state(double* arrs[6], int32_t* ids, int64_t n, void* ws, size_t ws_bytes, hipStream_t s) {
    size_t need = 0;
    hipError_t e = workspace_bytes(n, &need);
    if (e != hipSuccess) return e;
    if (ws_bytes < need) return hipErrorInvalidValue;
    const pp_sortk::StateArrays tmp = staging(static_cast<char*>(ws) + 4 * align256(4 * (size_t)n), n);
    const pp_sortk::StateArrays st = state_of(arrs, ids);
    hipLaunchKernelGGL(pp_sortk::k_scatter_state, dim3(grid_for(n)), dim3(256), 0, s, st, n, tmp);
    hipLaunchKernelGGL(pp_sortk::k_copy_state, dim3(grid_for(n)), dim3(256), 0, s, tmp, n, st, 0);
    return hipGetLastError();
}

hipError_t iota(int32_t* ids, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(pp_sortk::k_iota, dim3(grid_for(n)), dim3(256), 0, s, ids, n);
    return hipGetLastError();
}

}  // namespace pp_sort
