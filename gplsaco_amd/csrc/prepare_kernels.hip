// Device-side problem preparation: the O(n^2) steps of the reference's
// PreparedProblem (cpp/src/prepared_problem.cpp:58-126) for many instances
// at once. The candidate rows come from a per-city selection of the
// `needed` nearest other cities in (distance, id) order, the initial tour
// from a greedy nearest-neighbour walk; the initial checklist 2-opt runs in
// the ant kernel's local search (explicit mode, full checklist). Results are
// identical to the host preparation: integer EUC_2D distances, the same
// order keys and tie rules.
#include "faco_common.h"

namespace faco {
namespace {

constexpr uint32_t kKnnWaves = 4;   // cities per workgroup of the selection

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
__device__ __forceinline__ uint32_t uniform(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
}
__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (int offset = 32; offset; offset >>= 1) v += static_cast<uint32_t>(__shfl_xor(static_cast<int>(v), offset, 64));
    return uniform(v);
}
__device__ __forceinline__ uint32_t wave_min(uint32_t v) {
#pragma unroll
    for (int offset = 32; offset; offset >>= 1) {
        const uint32_t o = static_cast<uint32_t>(__shfl_xor(static_cast<int>(v), offset, 64));
        v = o < v ? o : v;
    }
    return uniform(v);
}
__host__ __device__ constexpr uint32_t align16(uint32_t bytes) { return (bytes + 15u) & ~15u; }

// One wave per city a of instance blockIdx.y. The distance d(a, b) of every
// other city goes to LDS (a itself reads 0xffff, above any EUC_2D distance
// of 16-bit coordinates); a binary search finds the distance D of the
// needed-th nearest city; the cities closer than D and the lowest ids at
// distance D are gathered in id order and each key is written at its rank.
__global__ void __launch_bounds__(64 * kKnnWaves) faco_prepare_knn(const uint32_t* coords, const uint32_t* sizes,
                                                                   uint32_t stride, uint32_t needed, uint32_t* keys) {
    extern __shared__ __attribute__((aligned(16))) unsigned char lds[];
    const uint32_t g = blockIdx.y, n = sizes[g];
    uint32_t* s_xy = reinterpret_cast<uint32_t*>(lds);
    const uint32_t* gxy = coords + static_cast<size_t>(g) * stride;
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) s_xy[i] = gxy[i];
    __syncthreads();
    const uint32_t wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const uint32_t a = blockIdx.x * kKnnWaves + wave;
    if (a >= n) return;   // no block-wide barrier follows
    uint16_t* d = reinterpret_cast<uint16_t*>(lds + align16(4u * stride) + wave * align16(2u * stride));
    uint32_t* list = reinterpret_cast<uint32_t*>(lds + align16(4u * stride) + kKnnWaves * align16(2u * stride) +
                                                 wave * align16(4u * needed));
    const uint32_t xa = s_xy[a];
    for (uint32_t b = lane; b < n; b += kWave) d[b] = static_cast<uint16_t>(b == a ? 0xffffu : euc2d(xa, s_xy[b]));
    wave_sync();
    // smallest D with at least `needed` cities at distance <= D (the sentinel
    // 0xffff is never counted, and n - 1 >= needed cities lie below it)
    uint32_t lo = 0, hi = 0xfffeu;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        uint32_t count = 0;
        for (uint32_t b = lane; b < n; b += kWave) count += d[b] <= mid ? 1u : 0u;
        if (wave_sum(count) >= needed) hi = mid;
        else lo = mid + 1;
    }
    const uint32_t D = lo;
    uint32_t below = 0;
    for (uint32_t b = lane; b < n; b += kWave) below += d[b] < D ? 1u : 0u;
    const uint32_t take = needed - wave_sum(below);   // cities at distance D, lowest ids first
    uint32_t filled = 0, ties = 0;
    for (uint32_t base = 0; base < n; base += kWave) {
        const uint32_t b = base + lane;
        const uint32_t db = b < n ? d[b] : 0xffffu;
        const bool tie = b < n && db == D;
        const uint64_t tie_mask = __ballot(tie);
        const bool chosen = b < n && (db < D || (tie && ties + lanes_below(tie_mask) < take));
        const uint64_t chosen_mask = __ballot(chosen);
        if (chosen) list[filled + lanes_below(chosen_mask)] = (db << 16) | b;
        filled += static_cast<uint32_t>(__builtin_popcountll(chosen_mask));
        ties += static_cast<uint32_t>(__builtin_popcountll(tie_mask));
    }
    wave_sync();
    uint32_t* row = keys + (static_cast<size_t>(g) * stride + a) * needed;
    for (uint32_t i = lane; i < needed; i += kWave) {
        const uint32_t key = list[i];
        uint32_t rank = 0;
        for (uint32_t j = 0; j < needed; ++j) rank += list[j] < key ? 1u : 0u;
        row[rank] = key;
    }
}

// One wave per instance: the greedy nearest-neighbour tour from city 0,
// the next city being the unvisited one of smallest (distance, id).
__global__ void __launch_bounds__(kWave) faco_prepare_nn(const uint32_t* coords, const uint32_t* sizes,
                                                         uint32_t stride, uint32_t* tours) {
    extern __shared__ __attribute__((aligned(16))) unsigned char lds[];
    const uint32_t g = blockIdx.x, n = sizes[g], lane = threadIdx.x;
    uint32_t* s_xy = reinterpret_cast<uint32_t*>(lds);
    uint8_t* seen = lds + align16(4u * stride);
    const uint32_t* gxy = coords + static_cast<size_t>(g) * stride;
    for (uint32_t i = lane; i < n; i += kWave) { s_xy[i] = gxy[i]; seen[i] = 0; }
    wave_sync();
    uint32_t* tour = tours + static_cast<size_t>(g) * stride;
    uint32_t current = 0;
    for (uint32_t i = 0; i < n; ++i) {
        if (lane == 0) { tour[i] = current; seen[current] = 1; }
        wave_sync();
        if (i + 1 == n) break;
        const uint32_t xc = s_xy[current];
        uint32_t best = 0xffffffffu;
        for (uint32_t b = lane; b < n; b += kWave) {
            const uint32_t key = (euc2d(xc, s_xy[b]) << 16) | b;
            best = !seen[b] && key < best ? key : best;
        }
        current = wave_min(best) & 0xffffu;   // an unvisited city remains while i + 1 < n
    }
}

}  // namespace

size_t prepare_knn_lds(uint32_t stride, uint32_t needed) {
    return align16(4u * stride) + kKnnWaves * (static_cast<size_t>(align16(2u * stride)) + align16(4u * needed));
}

hipError_t launch_prepare_knn(const uint32_t* coords, const uint32_t* sizes, uint32_t count, uint32_t stride,
                              uint32_t needed, uint32_t* keys, hipStream_t stream) {
    if (!count) return hipSuccess;
    const dim3 grid((stride + kKnnWaves - 1) / kKnnWaves, count);
    hipLaunchKernelGGL(faco_prepare_knn, grid, dim3(64 * kKnnWaves), prepare_knn_lds(stride, needed), stream, coords,
                       sizes, stride, needed, keys);
    return hipGetLastError();
}

hipError_t launch_prepare_nn(const uint32_t* coords, const uint32_t* sizes, uint32_t count, uint32_t stride,
                             uint32_t* tours, hipStream_t stream) {
    if (!count) return hipSuccess;
    hipLaunchKernelGGL(faco_prepare_nn, dim3(count), dim3(kWave), align16(4u * stride) + align16(stride), stream,
                       coords, sizes, stride, tours);
    return hipGetLastError();
}

}  // namespace faco
