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
                                              uint32_t* __restrict__ slots, const uint8_t* __restrict__ alive,
                                              int dead_shift) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    slots[i] = (uint32_t)i;
    if (dead_shift < 0) {   // no map: the key is the dead flag alone
        keys[i] = alive[i] ? 0u : 1u;
        return;
    }
    const uint32_t a = key_axis(m, 0, x[i], false);
    const uint32_t b = key_axis(m, 1, y[i], false);
    const uint32_t c = key_axis(m, 2, z[i], m.periodic != 0);
    uint32_t k = (a * (uint32_t)m.nc[1] + b) * (uint32_t)m.nc[2] + c;
    if (refine_bits > 0) {
        const double pos = refine_axis == 0 ? x[i] : (refine_axis == 1 ? y[i] : z[i]);
        k = (k << refine_bits) | sub_axis(m, refine_axis, pos, refine_bits);
    }
    if (group) k |= (uint32_t)group[i] << group_shift;
    if (alive && !alive[i]) k |= 1u << dead_shift;
    keys[i] = k;
}

// After the partition: alive = 1 on the live prefix, 0 after it, and the length of
// the prefix.  Keys are sorted with the dead flag as the top bit, so exactly one slot
// sits at the boundary (or the first slot is dead, or none is).
__global__ void __launch_bounds__(256) k_live_prefix(const uint32_t* __restrict__ keys, int64_t n, int dead_bit,
                                                     uint8_t* __restrict__ alive, int64_t* __restrict__ n_live) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const bool dead = (keys[i] >> dead_bit) & 1u;
    alive[i] = dead ? 0 : 1;
    if (!dead && (i + 1 == n || ((keys[i + 1] >> dead_bit) & 1u))) *n_live = i + 1;
    if (dead && i == 0) *n_live = 0;
}

// The six state arrays and the id array, permuted together in one pass (and the
// alive mask, when given, on the way back to the original order).
struct StateArrays {
    double* a[6];
    int32_t* ids;
    uint8_t* alive;
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
    if (src.alive) dst.alive[t] = src.alive[i];
}

__global__ void __launch_bounds__(256) k_copy_state(const StateArrays src, int64_t n, const StateArrays dst,
                                                    int with_ids) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
#pragma unroll
    for (int k = 0; k < 6; ++k) dst.a[k][i] = src.a[k][i];
    dst.ids[i] = with_ids ? src.ids[i] : (int32_t)i;
    if (src.alive && dst.alive) dst.alive[i] = src.alive[i];
}

__global__ void __launch_bounds__(256) k_iota(int32_t* ids, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) ids[i] = (int32_t)i;
}

}  // namespace pp_sortk

namespace pp_sort {

static inline unsigned grid_for(int64_t n) { return (unsigned)((n + 255) / 256); }

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// workspace: keys_in, keys_out, slots_in, slots_out (4n each), 256 B of results,
// a staging copy of the state (6 x 8n + 4n ids + n alive flags), cub temp
static size_t staging_bytes(int64_t n) {
    return 6 * align256(8 * (size_t)n) + align256(4 * (size_t)n) + align256((size_t)n);
}
static size_t staging_offset(int64_t n) { return 4 * align256(4 * (size_t)n) + 256; }

static hipError_t cub_bytes_for(int64_t n, size_t* cub_bytes) {
    *cub_bytes = 0;
    return hipcub::DeviceRadixSort::SortPairs(nullptr, *cub_bytes, (uint32_t*)nullptr, (uint32_t*)nullptr,
                                              (uint32_t*)nullptr, (uint32_t*)nullptr, (int)n, 0, 32);
}

hipError_t workspace_bytes(int64_t n, size_t* bytes) {
    size_t cub_bytes = 0;
    hipError_t e = cub_bytes_for(n, &cub_bytes);
    if (e != hipSuccess) return e;
    *bytes = staging_offset(n) + staging_bytes(n) + align256(cub_bytes);
    return hipSuccess;
}

static pp_sortk::StateArrays staging(char* base, int64_t n) {
    pp_sortk::StateArrays t;
    for (int k = 0; k < 6; ++k) t.a[k] = reinterpret_cast<double*>(base + k * align256(8 * (size_t)n));
    t.ids = reinterpret_cast<int32_t*>(base + 6 * align256(8 * (size_t)n));
    t.alive = reinterpret_cast<uint8_t*>(base + 6 * align256(8 * (size_t)n) + align256(4 * (size_t)n));
    return t;
}

static pp_sortk::StateArrays state_of(double* arrs[6], int32_t* ids, uint8_t* alive = nullptr) {
    pp_sortk::StateArrays t;
    for (int k = 0; k < 6; ++k) t.a[k] = arrs[k];
    t.ids = ids;
    t.alive = alive;
    return t;
}

static int key_bits(const MapDesc& m) {
    const uint64_t ncell = (uint64_t)m.nc[0] * m.nc[1] * m.nc[2];
    int b = 1;
    while (b < 32 && ((uint64_t)1 << b) < ncell) ++b;
    return b;
}

// keys (dead flag at dead_shift when alive is given), radix sort, fused permutation of the state
static hipError_t sort_keys_and_permute(const MapDesc* m, double* arrs[6], int32_t* ids, int64_t n, void* ws,
                                        size_t ws_bytes, hipStream_t s, const int32_t* group, int group_shift,
                                        int refine_axis, int refine_bits, const uint8_t* alive, int dead_shift,
                                        int bits) {
    size_t need = 0, cub_bytes = 0;
    hipError_t e = workspace_bytes(n, &need);
    if (e != hipSuccess) return e;
    if (ws_bytes < need) return hipErrorInvalidValue;
    if ((e = cub_bytes_for(n, &cub_bytes)) != hipSuccess) return e;
    char* w = static_cast<char*>(ws);
    const size_t q = align256(4 * (size_t)n);
    uint32_t* keys_in = reinterpret_cast<uint32_t*>(w);
    uint32_t* keys_out = reinterpret_cast<uint32_t*>(w + q);
    uint32_t* slots_in = reinterpret_cast<uint32_t*>(w + 2 * q);
    uint32_t* perm = reinterpret_cast<uint32_t*>(w + 3 * q);
    const pp_sortk::StateArrays tmp = staging(w + staging_offset(n), n);
    char* cub_tmp = w + staging_offset(n) + staging_bytes(n);
    const MapDesc none{};
    hipLaunchKernelGGL(pp_sortk::k_keys, dim3(grid_for(n)), dim3(256), 0, s, m ? *m : none, arrs[0], arrs[1],
                       arrs[2], n, group, group_shift, refine_axis, refine_bits, keys_in, slots_in, alive,
                       dead_shift);
    e = hipcub::DeviceRadixSort::SortPairs(cub_tmp, cub_bytes, keys_in, keys_out, slots_in, perm, (int)n, 0, bits,
                                           s);
    if (e != hipSuccess) return e;
    const pp_sortk::StateArrays st = state_of(arrs, ids);
    hipLaunchKernelGGL(pp_sortk::k_gather_state, dim3(grid_for(n)), dim3(256), 0, s, st, perm, n, tmp);
    hipLaunchKernelGGL(pp_sortk::k_copy_state, dim3(grid_for(n)), dim3(256), 0, s, tmp, n, st, 1);
    return hipGetLastError();
}

hipError_t sort_state(const MapDesc& m, double* arrs[6], int32_t* ids, int64_t n, void* ws, size_t ws_bytes,
                      hipStream_t s, const int32_t* group, int group_shift, int refine_axis, int refine_bits) {
    const int bits = group ? 32 : (key_bits(m) + refine_bits > 32 ? 32 : key_bits(m) + refine_bits);
    return sort_keys_and_permute(&m, arrs, ids, n, ws, ws_bytes, s, group, group_shift, refine_axis, refine_bits,
                                 nullptr, 0, bits);
}

hipError_t compact_state(const MapDesc* m, double* arrs[6], int32_t* ids, uint8_t* alive, int64_t n, void* ws,
                         size_t ws_bytes, hipStream_t s, int refine_axis, int refine_bits, int64_t* n_live_dev) {
    // dead flag above the cell (and sub-cell) bits; the caller checks that it fits in 32 bits
    const int dead_shift = m ? key_bits(*m) + refine_bits : -1;
    const int dead_bit = m ? dead_shift : 0;
    hipError_t e = sort_keys_and_permute(m, arrs, ids, n, ws, ws_bytes, s, nullptr, 0, refine_axis,
                                         m ? refine_bits : 0, alive, dead_shift, dead_bit + 1);
    if (e != hipSuccess) return e;
    const uint32_t* keys_out = reinterpret_cast<const uint32_t*>(static_cast<char*>(ws) + align256(4 * (size_t)n));
    hipLaunchKernelGGL(pp_sortk::k_live_prefix, dim3(grid_for(n)), dim3(256), 0, s, keys_out, n, dead_bit, alive,
                       n_live_dev);
    return hipGetLastError();
}

int key_bits_of(const MapDesc& m) { return key_bits(m); }

hipError_t unsort_state(double* arrs[6], int32_t* ids, int64_t n, void* ws, size_t ws_bytes, hipStream_t s,
                        uint8_t* alive) {
    size_t need = 0;
    hipError_t e = workspace_bytes(n, &need);
    if (e != hipSuccess) return e;
    if (ws_bytes < need) return hipErrorInvalidValue;
    const pp_sortk::StateArrays tmp = staging(static_cast<char*>(ws) + staging_offset(n), n);
    const pp_sortk::StateArrays st = state_of(arrs, ids, alive);
    pp_sortk::StateArrays dst = tmp;
    if (!alive) dst.alive = nullptr;
    hipLaunchKernelGGL(pp_sortk::k_scatter_state, dim3(grid_for(n)), dim3(256), 0, s, st, n, dst);
    hipLaunchKernelGGL(pp_sortk::k_copy_state, dim3(grid_for(n)), dim3(256), 0, s, dst, n, st, 0);
    return hipGetLastError();
}

// the 256-byte result slot of the workspace (live count of compact_state)
int64_t* result_slot(void* ws, int64_t n) {
    return reinterpret_cast<int64_t*>(static_cast<char*>(ws) + 4 * align256(4 * (size_t)n));
}

hipError_t iota(int32_t* ids, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(pp_sortk::k_iota, dim3(grid_for(n)), dim3(256), 0, s, ids, n);
    return hipGetLastError();
}

}  // namespace pp_sort
