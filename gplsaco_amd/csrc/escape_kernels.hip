// Escape primitives of the E3 candidate-graph exit on MI355X: the escape
// random stream, the slot view that swaps about a quarter of a candidate
// row for backup cities (and its distance-ordered LS form), and the per-ant
// cache of out-of-graph edges with whole-group proposals and commits.
// Reference semantics: cuda/escape_device.cuh:8-27,
// cpp/include/gp_faco/escape_slots.hpp:18-71, escape_edges.hpp:13-122,
// edge_constraints.hpp:15-33.
//
// Design: one wave per row or transaction. The stream, the slot view and
// the LS order are the ant kernel's own (escape_device.h): backup
// eligibility (first occurrence, before the first padding entry, not the
// row's own city, not already in the row) is one ballot per test and the
// eligible list is compacted with mbcnt; the LS row order is a rank sort
// over (distance, id) (the keys are distinct, so it equals the reference's
// insertion sort); a cache lookup checks the 64 cache slots in one lane
// each. The draws that pick slots and cities stay sequential, as the stream
// order defines them.
#include "escape_device.h"

#include <algorithm>
#include <stdexcept>
#include <string>
#include <vector>

namespace faco {
void set_last_error(const char* message);

namespace {

constexpr uint32_t kEscCapacity = 64, kEscRow = 20, kEscBackup = 64;
constexpr uint32_t kEscAllowed = 0, kEscInvalid = 1, kEscOutside = 2, kEscCapacityFull = 3;

__device__ __forceinline__ uint32_t lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }
__device__ __forceinline__ uint64_t wave_ballot(bool v) { return __ballot(v); }

struct Edge { uint32_t a, b; };
__device__ __forceinline__ bool same_edge(Edge x, Edge y) { return (x.a == y.a && x.b == y.b) || (x.a == y.b && x.b == y.a); }

__device__ bool graph_has(const uint32_t* offsets, const uint32_t* neighbors, uint32_t dimension, uint32_t a, uint32_t b) {
    if (a >= dimension || b >= dimension || a == b) return false;
    uint32_t first = offsets[a], last = offsets[a + 1];
    const uint32_t end = last;
    while (first < last) {
        const uint32_t middle = first + (last - first) / 2;
        if (neighbors[middle] < b) first = middle + 1; else last = middle;
    }
    return first < end && neighbors[first] == b;
}

// The escape cache held by one wave: lane l < 64 holds slot l.
struct WaveCache {
    uint32_t a, b;        // this lane's slot
    uint32_t size;        // uniform
    uint64_t epoch;       // uniform
    __device__ bool contains(Edge e) const {
        const uint32_t lane = lane_id();
        return wave_ballot(lane < size && same_edge(Edge{a, b}, e)) != 0;
    }
};

struct Proposal { Edge edges[3]; uint32_t count, expected_size, status; uint64_t expected_epoch; };

// prepare_escape_edges: uniform control flow, the cache lookups lane-parallel.
__device__ Proposal wave_prepare(uint32_t count, const Edge* removed, const Edge* added, const uint32_t* offsets,
                                 const uint32_t* neighbors, uint32_t dimension, const WaveCache& cache, bool permit) {
    Proposal p{};
    p.expected_size = cache.size; p.expected_epoch = cache.epoch; p.status = kEscAllowed;
    bool consumed[3] = {false, false, false};
    for (uint32_t i = 0; i < count; ++i) {
        const Edge e = added[i].a < added[i].b ? added[i] : Edge{added[i].b, added[i].a};
        if (e.a >= dimension || e.b >= dimension || e.a == e.b) { p.status = kEscInvalid; return p; }
        bool unchanged = false;
        for (uint32_t j = 0; j < count; ++j)
            if (!consumed[j] && same_edge(e, removed[j])) { consumed[j] = true; unchanged = true; break; }
        if (unchanged || graph_has(offsets, neighbors, dimension, e.a, e.b) || cache.contains(e)) continue;
        if (!permit) { p.status = kEscOutside; return p; }
        bool duplicate = false;
        for (uint32_t j = 0; j < p.count; ++j) duplicate |= same_edge(e, p.edges[j]);
        if (!duplicate) p.edges[p.count++] = e;
    }
    if (cache.size > kEscCapacity || p.count > kEscCapacity - cache.size) p.status = kEscCapacityFull;
    return p;
}

// commit_escape_edges: the whole group or nothing.
__device__ bool wave_commit(WaveCache& cache, const Proposal& p) {
    if (p.status != kEscAllowed || cache.epoch != p.expected_epoch || cache.size != p.expected_size ||
        cache.size > kEscCapacity || p.count > kEscCapacity - cache.size) return false;
    const uint32_t lane = lane_id();
    for (uint32_t i = 0; i < p.count; ++i)
        if (lane == cache.size + i) { cache.a = p.edges[i].a; cache.b = p.edges[i].b; }
    cache.size += p.count;
    return true;
}

__device__ void store_proposal(const Proposal& p, uint32_t* out) {
    out[0] = p.status; out[1] = p.count; out[2] = p.expected_size;
    out[3] = static_cast<uint32_t>(p.expected_epoch); out[4] = static_cast<uint32_t>(p.expected_epoch >> 32);
    for (uint32_t i = 0; i < 3; ++i) { out[5 + 2 * i] = i < p.count ? p.edges[i].a : 0u; out[6 + 2 * i] = i < p.count ? p.edges[i].b : 0u; }
}

struct RowParams {
    uint32_t rows, width, bw, n, ls;
    uint64_t key;
    uint32_t batch;
    const uint32_t *original, *backup, *from, *events, *coords;
    uint32_t *members, *counts;
    uint8_t* replaced;
};

__global__ void __launch_bounds__(64) faco_escape_rows_kernel(const RowParams p) {
    const uint32_t row = blockIdx.x;
    if (row >= p.rows) return;
    __shared__ uint8_t scratch[128];
    __shared__ uint32_t order[32];
    const uint32_t lane = lane_id();
    EscapeStream rng(p.key, p.batch, row, p.events[row], p.ls != 0);
    const uint32_t orig = lane < p.width ? p.original[static_cast<size_t>(row) * p.width + lane] : p.n;
    const uint32_t b = lane < p.bw ? p.backup[static_cast<size_t>(row) * p.bw + lane] : p.n;
    const LaneView v = lane_slot_view(orig, p.width, b, p.bw, p.n, p.from[row], rng, scratch);
    uint32_t member = v.member, source = v.source;
    if (p.coords) lane_sort_row(member, source, v.valid, p.from[row], p.coords, order);
    if (lane < p.width) {
        p.members[static_cast<size_t>(row) * p.width + lane] = member;
        p.replaced[static_cast<size_t>(row) * p.width + lane] = source != kNoSource ? 1 : 0;
    }
    if (lane == 0) { p.counts[2 * row] = v.valid; p.counts[2 * row + 1] = v.replacements; }
}

struct TxParams {
    uint32_t cases, dimension;
    const uint32_t *counts, *removed, *added, *permits, *cache_edges, *cache_sizes, *offsets, *neighbors;
    const uint64_t* cache_epochs;
    uint32_t *proposals, *committed, *cache_after, *size_after;
};

__global__ void __launch_bounds__(64) faco_escape_tx_kernel(const TxParams p) {
    const uint32_t c = blockIdx.x, lane = lane_id();
    if (c >= p.cases) return;
    const uint32_t count = p.counts[c];
    Edge removed[3], added[3];
    for (uint32_t i = 0; i < 3; ++i) {
        removed[i] = Edge{p.removed[6 * c + 2 * i], p.removed[6 * c + 2 * i + 1]};
        added[i] = Edge{p.added[6 * c + 2 * i], p.added[6 * c + 2 * i + 1]};
    }
    WaveCache cache;
    cache.size = p.cache_sizes[c];
    cache.epoch = p.cache_epochs[c];
    cache.a = p.cache_edges[(static_cast<size_t>(c) * kEscCapacity + lane) * 2];
    cache.b = p.cache_edges[(static_cast<size_t>(c) * kEscCapacity + lane) * 2 + 1];
    const Proposal prop = wave_prepare(count, removed, added, p.offsets, p.neighbors, p.dimension, cache, p.permits[c] != 0);
    const bool ok = wave_commit(cache, prop);
    if (lane == 0) {
        store_proposal(prop, p.proposals + 11 * static_cast<size_t>(c));
        p.committed[c] = ok ? 1u : 0u;
        p.size_after[c] = cache.size;
    }
    p.cache_after[(static_cast<size_t>(c) * kEscCapacity + lane) * 2] = cache.a;
    p.cache_after[(static_cast<size_t>(c) * kEscCapacity + lane) * 2 + 1] = cache.b;
}

__global__ void faco_escape_draws_kernel(uint64_t key, uint32_t batch, uint32_t ant, uint32_t event, uint32_t ls,
                                         uint32_t count, uint32_t* out) {
    EscapeStream rng(key, batch, ant, event, ls != 0);
    for (uint32_t i = 0; i < count; ++i) {
        const uint32_t v = rng.draw();
        if (threadIdx.x == 0) out[i] = v;
    }
}

void checked(hipError_t status, const char* what) {
    if (status != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(status));
}
void require(bool condition, const char* message) {
    if (!condition) throw std::invalid_argument(message);
}
template<class F> int guarded(F&& f) {
    try { f(); return FACO_OK; }
    catch (const std::invalid_argument& e) { set_last_error(e.what()); return FACO_INVALID_ARGUMENT; }
    catch (const std::exception& e) { set_last_error(e.what()); return FACO_RUNTIME_ERROR; }
    catch (...) { set_last_error("unknown native error"); return FACO_RUNTIME_ERROR; }
}
template<class T> struct Dev {
    T* ptr = nullptr;
    explicit Dev(size_t n) { if (n) checked(hipMalloc(reinterpret_cast<void**>(&ptr), n * sizeof(T)), "hipMalloc"); }
    ~Dev() { if (ptr) (void)hipFree(ptr); }
    Dev(const Dev&) = delete;
    Dev& operator=(const Dev&) = delete;
    void up(const T* src, size_t n) { if (n) checked(hipMemcpy(ptr, src, n * sizeof(T), hipMemcpyHostToDevice), "upload"); }
    void down(T* dst, size_t n) const { if (n) checked(hipMemcpy(dst, ptr, n * sizeof(T), hipMemcpyDeviceToHost), "download"); }
};

}  // namespace
}  // namespace faco

using namespace faco;

extern "C" {

int faco_diagnostic_escape_rows(uint32_t rows, uint32_t width, const uint32_t* original, uint32_t bw,
                                const uint32_t* backup, uint32_t n, const uint32_t* from, uint64_t key, uint32_t batch,
                                const uint32_t* events, uint32_t ls, const double* xy, uint32_t* out_members,
                                uint8_t* out_replaced, uint32_t* out_counts) {
    return guarded([&] {
        require(rows && original && backup && from && events && out_members && out_replaced && out_counts,
                "null argument or no rows");
        require(width >= 1 && width <= kEscRow && bw <= kEscBackup, "escape rows hold at most 20 slots and 64 backups");
        require(n >= 2 && n <= 65535, "city count out of range");
        for (size_t i = 0; i < static_cast<size_t>(rows) * width; ++i) require(original[i] <= n, "row entry above n");
        for (size_t i = 0; i < static_cast<size_t>(rows) * bw; ++i) require(backup[i] <= n, "backup entry above n");
        for (uint32_t r = 0; r < rows; ++r) require(from[r] < n, "row city out of range");
        std::vector<uint32_t> packed;
        if (xy) {
            packed.resize(n);
            for (uint32_t i = 0; i < n; ++i) {
                const double x = xy[2 * i], y = xy[2 * i + 1];
                require(x >= 0 && y >= 0 && x <= kMaxCoordinate && y <= kMaxCoordinate && x == static_cast<uint32_t>(x) &&
                        y == static_cast<uint32_t>(y), "coordinates must be integers in [0, 32767]");
                packed[i] = pack_xy(static_cast<uint32_t>(x), static_cast<uint32_t>(y));
            }
        }
        Dev<uint32_t> d_orig(static_cast<size_t>(rows) * width), d_back(std::max<size_t>(1, static_cast<size_t>(rows) * bw)),
            d_from(rows), d_events(rows), d_coords(packed.size()), d_members(static_cast<size_t>(rows) * width),
            d_counts(2 * static_cast<size_t>(rows));
        Dev<uint8_t> d_rep(static_cast<size_t>(rows) * width);
        d_orig.up(original, static_cast<size_t>(rows) * width);
        d_back.up(backup, static_cast<size_t>(rows) * bw);
        d_from.up(from, rows); d_events.up(events, rows);
        d_coords.up(packed.data(), packed.size());
        RowParams p{rows, width, bw, n, ls, key, batch, d_orig.ptr, d_back.ptr, d_from.ptr, d_events.ptr,
                    xy ? d_coords.ptr : nullptr, d_members.ptr, d_counts.ptr, d_rep.ptr};
        hipLaunchKernelGGL(faco_escape_rows_kernel, dim3(rows), dim3(64), 0, nullptr, p);
        checked(hipGetLastError(), "escape rows launch");
        checked(hipDeviceSynchronize(), "escape rows");
        d_members.down(out_members, static_cast<size_t>(rows) * width);
        d_rep.down(out_replaced, static_cast<size_t>(rows) * width);
        d_counts.down(out_counts, 2 * static_cast<size_t>(rows));
    });
}

int faco_diagnostic_escape_transactions(uint32_t cases, const uint32_t* counts, const uint32_t* removed,
                                        const uint32_t* added, const uint32_t* permits, const uint32_t* cache_edges,
                                        const uint32_t* cache_sizes, const uint64_t* cache_epochs,
                                        const uint32_t* offsets, const uint32_t* neighbors, uint32_t dimension,
                                        uint32_t* out_proposals, uint32_t* out_committed, uint32_t* out_cache_after,
                                        uint32_t* out_size_after) {
    return guarded([&] {
        require(cases && counts && removed && added && permits && cache_edges && cache_sizes && cache_epochs && offsets &&
                out_proposals && out_committed && out_cache_after && out_size_after, "null argument or no cases");
        require(dimension >= 1, "graph dimension must be positive");
        for (uint32_t c = 0; c < cases; ++c) {
            require(counts[c] >= 1 && counts[c] <= 3, "escape transactions hold 1 to 3 edges");
            require(cache_sizes[c] <= kEscCapacity, "escape cache holds at most 64 edges");
        }
        const uint32_t degree_total = offsets[dimension];
        for (uint32_t a = 0; a < dimension; ++a) require(offsets[a] <= offsets[a + 1], "CSR offsets must not decrease");
        Dev<uint32_t> d_counts(cases), d_rem(6 * static_cast<size_t>(cases)), d_add(6 * static_cast<size_t>(cases)),
            d_perm(cases), d_cache(128 * static_cast<size_t>(cases)), d_sizes(cases), d_off(dimension + 1),
            d_nb(std::max<uint32_t>(degree_total, 1)), d_prop(11 * static_cast<size_t>(cases)), d_comm(cases),
            d_after(128 * static_cast<size_t>(cases)), d_size_after(cases);
        Dev<uint64_t> d_epochs(cases);
        d_counts.up(counts, cases); d_rem.up(removed, 6 * static_cast<size_t>(cases));
        d_add.up(added, 6 * static_cast<size_t>(cases)); d_perm.up(permits, cases);
        d_cache.up(cache_edges, 128 * static_cast<size_t>(cases)); d_sizes.up(cache_sizes, cases);
        d_epochs.up(cache_epochs, cases); d_off.up(offsets, dimension + 1);
        if (degree_total) d_nb.up(neighbors, degree_total);
        TxParams p{cases, dimension, d_counts.ptr, d_rem.ptr, d_add.ptr, d_perm.ptr, d_cache.ptr, d_sizes.ptr, d_off.ptr,
                   d_nb.ptr, d_epochs.ptr, d_prop.ptr, d_comm.ptr, d_after.ptr, d_size_after.ptr};
        hipLaunchKernelGGL(faco_escape_tx_kernel, dim3(cases), dim3(64), 0, nullptr, p);
        checked(hipGetLastError(), "escape transactions launch");
        checked(hipDeviceSynchronize(), "escape transactions");
        d_prop.down(out_proposals, 11 * static_cast<size_t>(cases));
        d_comm.down(out_committed, cases);
        d_after.down(out_cache_after, 128 * static_cast<size_t>(cases));
        d_size_after.down(out_size_after, cases);
    });
}

int faco_diagnostic_escape_draws(uint64_t key, uint32_t batch, uint32_t ant, uint32_t event, uint32_t ls, uint32_t count,
                                 uint32_t* out) {
    return guarded([&] {
        require(out && count >= 1 && count <= 4096, "count must be in 1..4096");
        require(ant < 0x40000000u, "escape ant indices are below 2^30");
        Dev<uint32_t> d(count);
        hipLaunchKernelGGL(faco_escape_draws_kernel, dim3(1), dim3(64), 0, nullptr, key, batch, ant, event, ls, count, d.ptr);
        checked(hipGetLastError(), "escape draws launch");
        checked(hipDeviceSynchronize(), "escape draws");
        d.down(out, count);
    });
}

}  // extern "C"
