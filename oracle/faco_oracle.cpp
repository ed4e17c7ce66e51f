// CPU oracle for the FACO batch engine.
//
// TEST INFRASTRUCTURE ONLY. This file is a plain CPU restatement of the
// reference algorithm, used as the checker for the HIP engine. Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it. The
// product library (gplsaco_amd/libfaco_hip.so) never links or calls it.
//
// Every function names the reference file:line it restates. Distances are the
// TSPLIB EUC_2D rule on integer coordinates (nint of the Euclidean length),
// computed exactly with an integer square root, so every tour length, gain and
// cost below is an exact integer and the reference's FP64 summation order can
// not change a result.
//
// Parity status of each part is listed in the header comment of
// oracle/oracle.py, which loads this library.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

using u32 = std::uint32_t;
using u64 = std::uint64_t;
using i64 = std::int64_t;

thread_local std::string g_error;

// ---------------------------------------------------------------------------
// EUC_2D on integer coordinates. TSPLIB rule: nint(sqrt(dx^2+dy^2)) with
// nint(x) = floor(x + 0.5) (reference python/gp_faco/distance.py:46-49).
// r = isqrt(s); sqrt(s) + 0.5 >= r + 1  <=>  s > r*r + r  (s, r integers).
u32 isqrt_u64(u64 s) {
    u64 r = static_cast<u64>(std::sqrt(static_cast<double>(s)));
    while (r * r > s) --r;
    while ((r + 1) * (r + 1) <= s) ++r;
    return static_cast<u32>(r);
}

i64 euc2d(const std::int32_t* xy, u32 a, u32 b) {
    const i64 dx = static_cast<i64>(xy[2 * a]) - xy[2 * b];
    const i64 dy = static_cast<i64>(xy[2 * a + 1]) - xy[2 * b + 1];
    const u64 s = static_cast<u64>(dx * dx + dy * dy);
    const u64 r = isqrt_u64(s);
    return static_cast<i64>(r + (s > r * r + r ? 1 : 0));
}

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC'11). Counter words (c0,c1,c2,c3), key (k0,k1).
void philox4x32_10(const u32 in[4], const u32 key_in[2], u32 out[4]) {
    u32 c0 = in[0], c1 = in[1], c2 = in[2], c3 = in[3];
    u32 k0 = key_in[0], k1 = key_in[1];
    for (int round = 0; round < 10; ++round) {
        const u64 p0 = static_cast<u64>(0xD2511F53u) * c0;
        const u64 p1 = static_cast<u64>(0xCD9E8D57u) * c2;
        const u32 hi0 = static_cast<u32>(p0 >> 32), lo0 = static_cast<u32>(p0);
        const u32 hi1 = static_cast<u32>(p1 >> 32), lo1 = static_cast<u32>(p1);
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// One random stream per (key, batch, ant, step), following the reference's
// domain split (cuda/faco_choices.cuh:32-38): the 64-bit sub-sequence is
// (batch << 32) | ant and the step selects a 128-bit counter block. The
// stream yields 32-bit words in block order, four per block; the 128-bit
// counter (step, 0, ant, batch) increments with carry.
//
// For statistical checks only, g_alternate_stream replaces every stream by
// an unrelated generator (SplitMix64 over a per-(key, batch, ant, step) base,
// high 32 bits of each output): the same algorithm driven by independent
// uniforms, so its tour-length distribution can be compared with Philox's.
bool g_alternate_stream = false;
u64 mixed(u64 value);

struct Stream {
    u32 ctr[4];
    u32 key[2];
    u32 block[4];
    int index = 4;
    u64 alternate = 0;
    Stream(u64 seed, u32 batch, u32 ant, u32 step) {
        ctr[0] = step; ctr[1] = 0; ctr[2] = ant; ctr[3] = batch;
        key[0] = static_cast<u32>(seed); key[1] = static_cast<u32>(seed >> 32);
        if (g_alternate_stream)
            alternate = mixed(seed ^ mixed((static_cast<u64>(batch) << 32 | ant) ^ mixed(step + 0x632be59bd9b4e019ULL)));
    }
    u32 next() {
        if (g_alternate_stream) return static_cast<u32>(mixed(alternate++) >> 32);
        if (index == 4) {
            philox4x32_10(ctr, key, block);
            index = 0;
            for (int i = 0; i < 4; ++i) if (++ctr[i] != 0) break;
        }
        return block[index++];
    }
    // 53-bit uniform in [0,1) from two words, high word first
    // (cuda/faco_choices.cuh:40-44).
    double uniform53() {
        const u64 hi = next(), lo = next();
        return static_cast<double>(((hi << 32) | lo) >> 11) * 0x1.0p-53;
    }
};

// SplitMix64 finaliser (cuda/batch_engine.cu:27-32).
u64 mixed(u64 value) {
    value += 0x9e3779b97f4a7c15ULL;
    value = (value ^ (value >> 30)) * 0xbf58476d1ce4e5b9ULL;
    value = (value ^ (value >> 27)) * 0x94d049bb133111ebULL;
    return value ^ (value >> 31);
}

// Per-colony key (cuda/batch_engine.cu:762, cuda/fixed_faco.cu:112).
u64 colony_key(u64 instance_key, u64 seed) {
    return mixed(seed ^ mixed(instance_key + 0xd1b54a32d192ed03ULL));
}

// Start node: rejection sampling of one 32-bit word per try over [0,count)
// in the reserved step 0xffffffff (cuda/faco_choices.cuh:71-84).
u32 start_node(u64 key, u32 batch, u32 ant, u32 count) {
    Stream stream(key, batch, ant, 0xffffffffu);
    const u32 threshold = static_cast<u32>(-count) % count;
    u32 value;
    do { value = stream.next(); } while (value < threshold);
    return value % count;
}

// ---------------------------------------------------------------------------
struct Settings {  // layout mirrors cpp/include/gp_faco/fixed_faco_gpu.hpp:10-21
    u32 ants, primary_width, backup_width, ls_width;
    double beta, retention, p_best, epoch_source_probability;
    u64 ls_evaluation_limit, initial_ls_evaluation_limit;
};
static_assert(sizeof(Settings) == 64, "settings layout");

// normalized_settings (cpp/src/prepared_problem.cpp:19-33).
Settings normalized(Settings c, u32 n) {
    if (c.ants == 0) {
        u32 groups = 1;
        while (static_cast<u64>(256) * groups * groups < n) ++groups;
        c.ants = 64 * groups;
    }
    if (!(n >= 3 && n <= 10000 && c.ants > 0 && c.ants <= 4096 && c.primary_width >= 2 &&
          c.ls_width > 0 && std::isfinite(c.beta) && c.beta > 0 && c.retention >= 0 &&
          c.retention < 1 && c.p_best > 0 && c.p_best < 1 && c.epoch_source_probability >= 0 &&
          c.epoch_source_probability <= 1))
        throw std::invalid_argument("invalid FACO settings or size");
    c.primary_width = std::min(c.primary_width, n - 1);
    c.backup_width = std::min(c.backup_width, n - 1 - c.primary_width);
    c.ls_width = std::min(c.ls_width, n - 1);
    return c;
}

struct Problem {
    u32 n = 0;
    const std::int32_t* xy = nullptr;
    Settings s{};
    std::vector<u32> primary, backup, ls;  // row-major [n][width]
    std::vector<double> local_scale;
    double scale_epsilon = 1e-12;
    bool or_opt = false;  // local search adds Or-opt moves (this engine's extension)
    std::vector<float> prior;  // learned heuristic multipliers [n][pw] (this engine's extension), empty: none
    std::vector<u32> cheap, initial;
    i64 cheap_cost = 0, initial_cost = 0;
    // Hard mode: the undirected CSR graph every tour edge must belong to
    // (null: unrestricted)
    const u32* graph_offsets = nullptr;
    const u32* graph_neighbors = nullptr;
    // Escape mode on the same graph: 1 with its escape gates, 2 with them
    // disabled (which must run exactly as Hard mode)
    u32 escape = 0;
    i64 d(u32 a, u32 b) const { return euc2d(xy, a, b); }
    bool hard() const { return graph_offsets != nullptr; }
};

bool graph_has(const u32* offsets, const u32* neighbors, u32 dimension, u32 a, u32 b);

// Array tour with inverse positions: restates CpuTour (cpp/src/faco_cpu.cpp:46-217).
struct Tour {
    std::vector<u32> order, pos;
    i64 cost = 0;
    const Problem* p = nullptr;
    u32 n() const { return static_cast<u32>(order.size()); }
    void assign(const u32* tour, u32 n, const Problem* problem) {
        p = problem;
        order.assign(tour, tour + n);
        pos.assign(n, 0);
        for (u32 i = 0; i < n; ++i) pos[order[i]] = i;
        cost = recomputed();
    }
    i64 recomputed() const {
        i64 sum = 0;
        u32 prev = order.back();
        for (u32 v : order) { sum += p->d(prev, v); prev = v; }
        return sum;
    }
    u32 succ(u32 v) const { return order[(pos[v] + 1) % n()]; }
    u32 pred(u32 v) const { return order[(pos[v] + n() - 1) % n()]; }
    bool has_edge(u32 a, u32 b) const { return succ(a) == b || pred(a) == b; }

    // CpuTour::relocate (cpp/src/faco_cpu.cpp:90-112): move `node` to directly
    // after `target`; only the closed position range between them shifts.
    bool relocate(u32 target, u32 node) {
        if (succ(target) == node) return false;
        const u32 before = pred(node), after = succ(node), target_after = succ(target);
        const u32 pn = pos[node], pt = pos[target];
        if (pt < pn) {
            for (u32 i = pn; i > pt + 1; --i) order[i] = order[i - 1];
            order[pt + 1] = node;
        } else {
            for (u32 i = pn; i < pt; ++i) order[i] = order[i + 1];
            order[pt] = node;
        }
        for (u32 i = std::min(pn, pt); i <= std::max(pn, pt); ++i) pos[order[i]] = i;
        cost += -p->d(before, node) - p->d(node, after) - p->d(target, target_after)
                + p->d(before, after) + p->d(target, node) + p->d(node, target_after);
        return true;
    }

    // CpuTour::flip_without_cost (cpp/src/faco_cpu.cpp:114-135). The
    // complement side with first == 0 walks one extra lap in the reference;
    // the net effect is the plain reversal of [0, last), which is what the
    // straight branch does, so first == 0 is folded into it here.
    void flip(u32 start, u32 end) {
        u32 first = pos[start], last = pos[end];
        if (first > last) std::swap(first, last);
        const u32 N = n(), length = last - first;
        if (length <= N - length || first == 0) {
            std::reverse(order.begin() + first, order.begin() + last);
            for (u32 i = first; i < last; ++i) pos[order[i]] = i;
        } else {
            u32 left = last, right = first - 1;
            for (u32 k = 0; k < (N - length) / 2; ++k) {
                std::swap(order[left], order[right]);
                pos[order[left]] = left; pos[order[right]] = right;
                left = left + 1 == N ? 0 : left + 1;
                right = right == 0 ? N - 1 : right - 1;
            }
        }
    }
};

// Hard-mode moves (edge_constraints.hpp:47-90): every added edge that does
// not cancel a removed one (as a multiset) must be a graph edge.
bool hard_edges_allowed(const Problem& p, u32 count, const u32 (*removed)[2], const u32 (*added)[2]) {
    bool consumed[3] = {false, false, false};
    for (u32 i = 0; i < count; ++i) {
        bool existing = false;
        for (u32 j = 0; j < count; ++j) {
            const bool same = (added[i][0] == removed[j][0] && added[i][1] == removed[j][1]) ||
                              (added[i][0] == removed[j][1] && added[i][1] == removed[j][0]);
            if (!consumed[j] && same) { consumed[j] = true; existing = true; break; }
        }
        if (!existing && !graph_has(p.graph_offsets, p.graph_neighbors, p.n, added[i][0], added[i][1])) return false;
    }
    return true;
}
bool hard_relocation_allowed(const Problem& p, const Tour& t, u32 target, u32 node) {
    if (target >= p.n || node >= p.n || target == node) return false;
    const u32 target_after = t.succ(target);
    if (target_after == node) return true;
    const u32 before = t.pred(node), after = t.succ(node);
    const u32 removed[3][2] = {{before, node}, {node, after}, {target, target_after}};
    const u32 added[3][2] = {{before, after}, {target, node}, {node, target_after}};
    return hard_edges_allowed(p, 3, removed, added);
}
bool hard_two_opt_allowed(const Problem& p, u32 a, u32 a_neighbor, u32 b, u32 b_neighbor) {
    const u32 removed[2][2] = {{a, a_neighbor}, {b, b_neighbor}};
    const u32 added[2][2] = {{a, b}, {a_neighbor, b_neighbor}};
    return hard_edges_allowed(p, 2, removed, added);
}

// ---------------------------------------------------------------------------
// Escape primitives (the E3 candidate-graph exit), restated from
// cpp/include/gp_faco/escape_slots.hpp:18-71, escape_edges.hpp:13-122,
// edge_constraints.hpp:15-33 and cuda/escape_device.cuh:8-27.

constexpr u32 kEscCapacity = 64, kEscLsWidth = 20, kEscBackupWidth = 64;

// EscapeRandom (cuda/escape_device.cuh:8-27) over the shared Philox: a word
// stream per (batch, domain | ant, event) event, domain 0x40000000 for
// construction gates and 0x80000000 for LS rows; counter block
// (block, event, domain | ant, batch). Replaces cuRAND's stream (unpinned).
struct EscStream {
    u32 key[2], ant, batch, event, block = 0, index = 4, words[4] = {0, 0, 0, 0};
    EscStream(u64 k, u32 b, u32 a, u32 e, bool ls) {
        key[0] = static_cast<u32>(k); key[1] = static_cast<u32>(k >> 32);
        batch = b; event = e; ant = (ls ? 0x80000000u : 0x40000000u) | a;
    }
    u32 draw() {
        if (index == 4) {
            const u32 ctr[4] = {block++, event, ant, batch};
            philox4x32_10(ctr, key, words);
            index = 0;
        }
        return words[index++];
    }
    u32 below(u32 count) {   // unbiased: reject words below 2^32 mod count
        const u32 threshold = (0u - count) % count;
        u32 v;
        do { v = draw(); } while (v < threshold);
        return v % count;
    }
};

// The reference test's scripted draws (tests/cpp/test_escape_slots.cpp:13-17).
struct ScriptDraws {
    u32 rounding;
    u32 draw() { return rounding; }
    u32 below(u32 count) {
        if (!count) throw std::invalid_argument("sampling an empty set");
        return count - 1;
    }
};

struct EscRow {
    u32 members[kEscLsWidth] = {};
    std::uint8_t replaced[kEscLsWidth] = {};
    u32 valid = 0, replacements = 0;
};

// escape_slot_view (escape_slots.hpp:20-49).
template<class Random>
EscRow slot_view(const u32* original, u32 width, const u32* backup, u32 bw, u32 n, u32 from, Random& random) {
    if (width > kEscLsWidth || bw > kEscBackupWidth) throw std::invalid_argument("escape row too wide");
    EscRow row;
    for (u32 j = 0; j < width; ++j) { row.members[j] = original[j]; row.valid += original[j] < n ? 1u : 0u; }
    u32 wanted = row.valid / 4;
    if (row.valid % 4) wanted += (random.draw() & 3u) < row.valid % 4 ? 1u : 0u;
    if (!wanted) return row;
    u32 eligible[kEscBackupWidth], count = 0;
    for (u32 j = 0; j < bw; ++j) {
        const u32 node = backup[j];
        if (node >= n) break;
        if (node == from) continue;
        bool duplicate = false;
        for (u32 k = 0; k < row.valid; ++k) duplicate |= original[k] == node;
        for (u32 k = 0; k < count; ++k) duplicate |= eligible[k] == node;
        if (!duplicate) eligible[count++] = node;
    }
    wanted = std::min(wanted, count);
    u32 slots[kEscLsWidth];
    for (u32 j = 0; j < row.valid; ++j) slots[j] = j;
    for (u32 j = 0; j < wanted; ++j) {
        const u32 pick_slot = j + random.below(row.valid - j);
        std::swap(slots[pick_slot], slots[j]);
        const u32 pick_node = j + random.below(count - j);
        std::swap(eligible[pick_node], eligible[j]);
        row.members[slots[j]] = eligible[j];
        row.replaced[slots[j]] = 1;
    }
    row.replacements = wanted;
    return row;
}

// sort_escape_ls_row (escape_slots.hpp:52-68): insertion sort of the valid
// prefix by (distance, id); the replaced flags move with their members.
void sort_row(EscRow& row, const double* distance_of) {
    for (u32 j = 1; j < row.valid; ++j) {
        const u32 node = row.members[j];
        const std::uint8_t flag = row.replaced[j];
        const double d = distance_of[node];
        u32 at = j;
        while (at) {
            const u32 prior = row.members[at - 1];
            const double previous = distance_of[prior];
            if (previous < d || (previous == d && prior < node)) break;
            row.members[at] = prior; row.replaced[at] = row.replaced[at - 1]; --at;
        }
        row.members[at] = node; row.replaced[at] = flag;
    }
}

struct EscEdge { u32 a, b; };
bool same_edge(EscEdge x, EscEdge y) { return (x.a == y.a && x.b == y.b) || (x.a == y.b && x.b == y.a); }

// SparseGraphView::operator() (edge_constraints.hpp:15-28): binary search of b in a's CSR row.
bool graph_has(const u32* offsets, const u32* neighbors, u32 dimension, u32 a, u32 b) {
    if (a >= dimension || b >= dimension || a == b) return false;
    u32 first = offsets[a], last = offsets[a + 1];
    const u32 end = last;
    while (first < last) {
        const u32 middle = first + (last - first) / 2;
        if (neighbors[middle] < b) first = middle + 1; else last = middle;
    }
    return first < end && neighbors[first] == b;
}

struct EscCache { EscEdge edges[kEscCapacity]; u32 size; u64 epoch; };
bool cache_has(const EscCache& c, EscEdge e) {
    for (u32 i = 0; i < c.size; ++i) if (same_edge(c.edges[i], e)) return true;
    return false;
}

enum EscStatus : u32 { kAllowed = 0, kInvalidEdge = 1, kOutsideGraph = 2, kCapacity = 3 };
struct EscProposal { EscEdge edges[3] = {}; u32 count = 0, expected_size = 0; u64 expected_epoch = 0; u32 status = kAllowed; };

// prepare_escape_edges (escape_edges.hpp:45-72).
EscProposal prepare_edges(u32 count, const EscEdge* removed, const EscEdge* added, const u32* offsets,
                          const u32* neighbors, u32 dimension, const EscCache& cache, bool permit) {
    EscProposal p;
    p.expected_size = cache.size; p.expected_epoch = cache.epoch;
    bool consumed[3] = {false, false, false};
    for (u32 i = 0; i < count; ++i) {
        const EscEdge e = added[i].a < added[i].b ? added[i] : EscEdge{added[i].b, added[i].a};
        if (e.a >= dimension || e.b >= dimension || e.a == e.b) { p.status = kInvalidEdge; return p; }
        bool unchanged = false;
        for (u32 j = 0; j < count; ++j)
            if (!consumed[j] && same_edge(e, removed[j])) { consumed[j] = true; unchanged = true; break; }
        if (unchanged || graph_has(offsets, neighbors, dimension, e.a, e.b) || cache_has(cache, e)) continue;
        if (!permit) { p.status = kOutsideGraph; return p; }
        bool duplicate = false;
        for (u32 j = 0; j < p.count; ++j) duplicate |= same_edge(e, p.edges[j]);
        if (!duplicate) p.edges[p.count++] = e;
    }
    if (cache.size > kEscCapacity || p.count > kEscCapacity - cache.size) p.status = kCapacity;
    return p;
}

// commit_escape_edges (escape_edges.hpp:74-84).
bool commit_edges(EscCache& cache, const EscProposal& p) {
    if (p.status != kAllowed || cache.epoch != p.expected_epoch || cache.size != p.expected_size ||
        cache.size > kEscCapacity || p.count > kEscCapacity - cache.size) return false;
    for (u32 i = 0; i < p.count; ++i) cache.edges[cache.size + i] = p.edges[i];
    cache.size += p.count;
    return true;
}

void load_cache(EscCache& c, const u32* edges, u32 size, u64 epoch) {
    if (size > kEscCapacity) throw std::invalid_argument("escape cache holds at most 64 edges");
    c.size = size; c.epoch = epoch;
    for (u32 i = 0; i < kEscCapacity; ++i) c.edges[i] = i < size ? EscEdge{edges[2 * i], edges[2 * i + 1]} : EscEdge{0, 0};
}

void store_proposal(const EscProposal& p, u32* out) {
    // status, count, expected size, expected epoch (low, high), edges [3][2]
    out[0] = p.status; out[1] = p.count; out[2] = p.expected_size;
    out[3] = static_cast<u32>(p.expected_epoch); out[4] = static_cast<u32>(p.expected_epoch >> 32);
    for (u32 i = 0; i < 3; ++i) { out[5 + 2 * i] = p.edges[i].a; out[6 + 2 * i] = p.edges[i].b; }
}


struct LsStats {
    u64 processed = 0, checks = 0, evaluations = 0, accepted = 0, reactivations = 0;
    u64 rejections = 0;  // Hard mode: evaluated moves that would leave the graph (faco_cpu.cpp:180-183)
    u64 capacity_rejections = 0, new_edges = 0;  // Escape mode: full-cache rejections, committed new edges
    u32 limit_reached = 0;
};

// Escape mode's local search context (cuda/escape_device.cuh:111-128,
// cuda/faco_device.cuh:326-336, 435-447): this ant's LS rows (an anchor's row
// is its escape view), their replaced-slot flags, and the ant's cache of
// out-of-graph edges. A move through a replaced slot may add new edges.
struct EscapeLs {
    const u32* rows;
    const std::uint8_t* replaced;
    EscCache* cache;
    std::vector<u32>* events;   // or null: accepted moves as (kind, a, b, slot, cache size, permit)
};

// prepare_escape_two_opt (escape_edges.hpp:117-122).
EscProposal escape_two_opt_proposal(const Problem& p, const EscCache& cache, u32 a, u32 an, u32 b, u32 bn,
                                    bool permit) {
    const EscEdge removed[2] = {{a, an}, {b, bn}};
    const EscEdge added[2] = {{a, b}, {an, bn}};
    return prepare_edges(2, removed, added, p.graph_offsets, p.graph_neighbors, p.n, cache, permit);
}

// prepare_escape_relocation (escape_edges.hpp:101-115): move `node` to
// directly after `target` in the working tour.
EscProposal escape_relocation_proposal(const Problem& p, const Tour& t, const EscCache& cache, u32 target, u32 node,
                                       bool permit) {
    EscProposal q;
    q.expected_size = cache.size; q.expected_epoch = cache.epoch;
    if (target >= p.n || node >= p.n || target == node) { q.status = kInvalidEdge; return q; }
    const u32 target_after = t.succ(target);
    if (target_after == node) return q;
    const u32 before = t.pred(node), after = t.succ(node);
    const EscEdge removed[3] = {{before, node}, {node, after}, {target, target_after}};
    const EscEdge added[3] = {{before, after}, {target, node}, {node, target_after}};
    return prepare_edges(3, removed, added, p.graph_offsets, p.graph_neighbors, p.n, cache, permit);
}

// Checklist 2-opt: restates CpuTour::checklist_two_opt
// (cpp/src/faco_cpu.cpp:148-217). `checklist` is the whole FIFO history; the
// unprocessed part starts at `next`. `queued[v]` marks membership of the
// unprocessed part (the reference's std::find over [next, end)).
//
// With `or_opt` (this engine's extension; the reference has no Or-opt), each
// popped city a also tries Or-opt moves after its 2-opt scan: the segment
// S = a, succ(a), ... of length L = 1, 2, 3 (L + 4 <= n) is cut out between
// p = pred(a) and nx = succ(last of S) and reinserted at the edge (c, d) of
// an LS neighbour b of a, either forward between c = b and d = succ(b)
// (new edges (b, a), (last, d)) or reversed between c = pred(b) and d = b
// (new edges (c, last), (a, b)); c and d must lie outside S, p and nx. The
// neighbours tried are those the predecessor scan of 2-opt passed
// (d(a, b) < d(pred a, a)), in row order, then L, then forward before
// reversed; each counts one move evaluation against the same budget, with
// the same cut-short rule. The best move over 2-opt and Or-opt wins (ties to
// the earlier in that order, so 2-opt first). An Or-opt move is applied as
// two (reversed) or three (forward) 2-opt reversals and reactivates p, nx,
// a, last, c, d.
namespace {
// Remove edges (x1, x2), (y1, y2) and add (x1, y1), (x2, y2), where x2
// follows x1 and y2 follows y1 in one orientation of the tour.
void two_opt_move(Tour& t, u32 x1, u32 x2, u32 y1, u32 y2) {
    if (t.succ(x1) == x2 && t.succ(y1) == y2) t.flip(x2, y2);
    else if (t.succ(x2) == x1 && t.succ(y2) == y1) t.flip(x1, y1);
    else throw std::runtime_error("Or-opt reversal on edges that are not in the tour");
}
}  // namespace

LsStats two_opt(Tour& t, const u32* ls, u32 lw, std::vector<u32>& checklist,
                std::vector<std::uint8_t>& queued, u64 limit, bool or_opt, const EscapeLs* esc = nullptr) {
    LsStats st;
    if (esc) ls = esc->rows;
    const u32 N = t.n();
    size_t next = 0;
    i64 gain_sum = 0;
    while (next < checklist.size() && st.accepted < N) {
        if (st.evaluations >= limit) { st.limit_reached = 1; break; }
        const u32 a = checklist[next++];
        queued[a] = 0;
        const u32 a_next = t.succ(a), a_prev = t.pred(a);
        ++st.processed;
        const i64 to_next = t.p->d(a, a_next), to_prev = t.p->d(a_prev, a);
        i64 best = -1;
        u32 move[4] = {0, 0, 0, 0};
        u32 best_slot = 0, best_kind = 0, best_b = 0, best_nb = 0;
        bool interrupted = false, or_best = false;
        u32 or_move[6] = {0, 0, 0, 0, 0, 0};  // p, a, last, nx, c, d
        u32 or_forward = 0;
        u32 passed[2] = {0, 0};
        for (int kind = 0; kind < 2 && !interrupted; ++kind) {
            const i64 current = kind == 0 ? to_next : to_prev;
            for (u32 j = 0; j < lw; ++j) {
                const u32 b = ls[static_cast<size_t>(a) * lw + j];
                ++st.checks;
                if (b >= N) break;   // right padding of an LS row (faco_device.cuh:320)
                const i64 ab = t.p->d(a, b);
                if (!(current > ab)) break;
                if (st.evaluations == limit) { interrupted = true; break; }
                ++st.evaluations;
                ++passed[kind];
                const u32 nb = kind == 0 ? t.succ(b) : t.pred(b);
                const u32 closing = kind == 0 ? a_next : a_prev;
                // Hard mode: a move that leaves the graph never wins
                // (cuda/faco_device.cuh:383-394); it still counts as evaluated
                const bool identity = nb == a || b == closing;
                bool allowed;
                if (esc) {
                    // Escape mode: graph edges, cached edges, or new edges
                    // through a replaced slot while the cache has room; an
                    // identity move always passes (its added edges cancel)
                    const EscProposal q = escape_two_opt_proposal(*t.p, *esc->cache, a, closing, b, nb,
                                                                  esc->replaced[static_cast<size_t>(a) * lw + j] != 0);
                    st.capacity_rejections += q.status == kCapacity ? 1 : 0;
                    allowed = q.status == kAllowed;
                } else {
                    allowed = identity || !t.p->hard() || hard_two_opt_allowed(*t.p, a, closing, b, nb);
                }
                st.rejections += allowed ? 0 : 1;
                const i64 gain = identity ? 0
                    : allowed ? current + t.p->d(b, nb) - ab - t.p->d(closing, nb) : std::numeric_limits<i64>::min();
                if (gain > best) {
                    best = gain;
                    best_slot = j; best_kind = static_cast<u32>(kind); best_b = b; best_nb = nb;
                    if (kind == 0) { move[0] = a_next; move[1] = nb; move[2] = a; move[3] = b; }
                    else { move[0] = a; move[1] = b; move[2] = a_prev; move[3] = nb; }
                }
            }
        }
        if (or_opt && !interrupted) {
            const u32 p = a_prev;
            for (u32 j = 0; j < passed[1] && !interrupted; ++j) {
                const u32 b = ls[static_cast<size_t>(a) * lw + j];
                u32 seg[3] = {a, 0, 0};
                for (u32 L = 1; L <= 3 && L + 4 <= N && !interrupted; ++L) {
                    if (L > 1) seg[L - 1] = t.succ(seg[L - 2]);
                    const u32 last = seg[L - 1], nx = t.succ(last);
                    auto outside = [&](u32 v) {
                        if (v == p || v == nx) return false;
                        for (u32 k = 0; k < L; ++k) if (seg[k] == v) return false;
                        return true;
                    };
                    const i64 removed = t.p->d(p, a) + t.p->d(last, nx) - t.p->d(p, nx);
                    for (u32 variant = 0; variant < 2 && !interrupted; ++variant) {
                        const u32 forward = variant == 0 ? 1u : 0u;
                        const u32 c = forward ? b : t.pred(b), d = forward ? t.succ(b) : b;
                        if (!outside(c) || !outside(d)) continue;
                        if (st.evaluations == limit) { interrupted = true; break; }
                        ++st.evaluations;
                        const i64 added = forward ? t.p->d(b, a) + t.p->d(last, d) : t.p->d(c, last) + t.p->d(a, b);
                        const i64 gain = removed + t.p->d(c, d) - added;
                        if (gain > best) {
                            best = gain;
                            or_best = true;
                            or_move[0] = p; or_move[1] = a; or_move[2] = last; or_move[3] = nx;
                            or_move[4] = c; or_move[5] = d;
                            or_forward = forward;
                        }
                    }
                }
            }
        }
        if (interrupted) { st.limit_reached = 1; break; }
        if (best > 0) {
            if (or_best) {
                const u32 p = or_move[0], s1 = or_move[1], sl = or_move[2], nx = or_move[3];
                const u32 c = or_move[4], d = or_move[5];
                if (or_forward) {
                    two_opt_move(t, p, s1, c, d);     // p c .. nx sl .. s1 d
                    two_opt_move(t, p, c, nx, sl);    // p nx .. c sl .. s1 d
                    two_opt_move(t, c, sl, s1, d);    // p nx .. c s1 .. sl d
                } else {
                    two_opt_move(t, p, s1, c, d);     // p c .. nx sl .. s1 d
                    two_opt_move(t, p, c, nx, sl);    // p nx .. c sl .. s1 d
                }
                for (u32 v : {p, nx, s1, sl, c, d}) {
                    if (!queued[v]) { queued[v] = 1; checklist.push_back(v); ++st.reactivations; }
                }
            } else {
                if (esc) {
                    // commit the move's new edges to the ant's cache
                    const EscProposal q = escape_two_opt_proposal(
                        *t.p, *esc->cache, a, best_kind == 0 ? a_next : a_prev, best_b, best_nb,
                        esc->replaced[static_cast<size_t>(a) * lw + best_slot] != 0);
                    if (!commit_edges(*esc->cache, q)) throw std::runtime_error("escape 2-opt commit refused");
                    st.new_edges += q.count;
                    if (esc->events)
                        esc->events->insert(esc->events->end(), {best_kind + 1, a, best_b, best_slot, esc->cache->size,
                                                                 esc->replaced[static_cast<size_t>(a) * lw + best_slot]});
                }
                t.flip(move[0], move[1]);
                for (u32 v : move) {
                    if (!queued[v]) { queued[v] = 1; checklist.push_back(v); ++st.reactivations; }
                }
            }
            ++st.accepted;
            gain_sum -= best;
        }
    }
    t.cost += gain_sum;
    return st;
}

// ---------------------------------------------------------------------------
// select_next (cpp/src/faco_cpu.cpp:247-294) for one construction step.
// Returns the chosen node; stage: 0 primary, 1 zero-weight, 2 backup,
// 3 global fallback, 4 exhausted. In Hard mode (cuda/faco_choices.cuh:46-53)
// a city is available only when moving it after `current` in the working
// tour `work` keeps every edge in the graph.
u32 select_next(const Problem& p, u32 current, const double* products,
                const std::vector<std::uint8_t>& visited, double u, int* stage, const Tour* work = nullptr) {
    const u32 pw = p.s.primary_width, bw = p.s.backup_width, n = p.n;
    auto free_city = [&](u32 v) {
        return v < n && !visited[v] && (!p.hard() || hard_relocation_allowed(p, *work, current, v));
    };
    const u32* primary = p.primary.data() + static_cast<size_t>(current) * pw;
    double sum = 0;
    u32 avail[64];
    double prefix[64];
    u32 count = 0;
    for (u32 j = 0; j < pw; ++j) {
        const u32 v = primary[j];
        if (free_city(v)) { sum += products[j]; avail[count] = v; prefix[count] = sum; ++count; }
    }
    if (count) {
        if (sum == 0) {
            u32 k = static_cast<u32>(u * count);
            if (k >= count) k = count - 1;
            *stage = 1;
            return avail[k];
        }
        const double threshold = u * sum;
        *stage = 0;
        for (u32 i = 0; i < count; ++i) if (threshold < prefix[i]) return avail[i];
        return avail[count - 1];
    }
    const u32* backup = p.backup.data() + static_cast<size_t>(current) * bw;
    for (u32 j = 0; j < bw; ++j) {
        if (free_city(backup[j])) { *stage = 2; return backup[j]; }
    }
    u32 chosen = current;
    i64 best = std::numeric_limits<i64>::max();
    for (u32 v = 0; v < n; ++v) {
        if (!free_city(v)) continue;
        const i64 value = p.d(current, v);
        if (value < best) { best = value; chosen = v; }
    }
    *stage = chosen == current ? 4 : 3;
    return chosen;
}

struct ConstructionStats { u32 mne = 0, steps = 0, relocations = 0, exhausted = 0; };

// Escape counters of one ant, in the order of the reference's EscapeStats
// (escape_slots.hpp:70-75).
enum EscCounter : u32 {
    kOpportunities, kGates, kReplacedSlots, kEscapeRelocations, kAnchorNodes, kLsReplacedSlots,
    kOldViewReactivations, kAnchorReactivations, kNewEdges, kConstructionCapacity, kLsCapacity, kEscCounters
};

struct AntResult {
    std::vector<u32> construction_tour, tour, positions, checklist;
    i64 construction_cost = 0, final_cost = 0;
    ConstructionStats c;
    LsStats ls;
    u32 start = 0;
    std::vector<std::uint8_t> footprint;   // Escape mode: cities whose LS row had replaced slots
    u64 esc[kEscCounters] = {};
};

// One ant: FocusedConstruction (cpp/src/faco_cpu.cpp:219-245) driven by a
// chooser, then checklist 2-opt, as in construct_and_search
// (cuda/faco_device.cuh:161-500).
template<class Choose>
AntResult run_ant(const Problem& p, const u32* parent, u32 start, u32 target, u64 limit,
                  Choose choose) {
    AntResult r;
    const u32 n = p.n;
    Tour reference, work;
    reference.assign(parent, n, &p);
    work = reference;
    std::vector<std::uint8_t> visited(n, 0), queued(n, 0);
    std::vector<u32> checklist;
    u32 current = start;
    r.start = start;
    visited[current] = 1;
    while (r.c.mne < target && r.c.steps + 1 < n) {
        const u32 step = r.c.steps + 1;
        const u32 selected = choose(current, step, visited, work);
        if (selected >= n || visited[selected]) { r.c.exhausted = 1; break; }
        const u32 previous = work.pred(selected);
        visited[selected] = 1;
        ++r.c.steps;
        if (work.relocate(current, selected)) ++r.c.relocations;
        if (!reference.has_edge(current, selected)) {
            ++r.c.mne;
            for (u32 v : {current, selected, previous})
                if (!queued[v]) { queued[v] = 1; checklist.push_back(v); }
        }
        current = selected;
    }
    r.construction_tour = work.order;
    r.construction_cost = work.cost;
    r.ls = two_opt(work, p.ls.data(), p.s.ls_width, checklist, queued, limit, p.or_opt);
    r.tour = work.order;
    r.positions = work.pos;
    r.checklist = checklist;
    r.final_cost = work.cost;
    return r;
}

// Host preparation: restates prepare_problem (cpp/src/prepared_problem.cpp:58-126)
// and make_cheap_problem (:41-56) with EUC_2D integer distances.
void prepare(Problem& p) {
    const u32 n = p.n;
    const Settings& c = p.s;
    p.cheap.resize(n);
    for (u32 i = 0; i < n; ++i) p.cheap[i] = i;
    {
        Tour t; t.assign(p.cheap.data(), n, &p); p.cheap_cost = t.cost;
    }
    const u32 scale_width = std::min<u32>(8, n - 1);
    const u32 needed = std::max({c.primary_width + c.backup_width, c.ls_width, scale_width});
    p.primary.assign(static_cast<size_t>(n) * c.primary_width, 0);
    p.backup.assign(static_cast<size_t>(n) * c.backup_width, 0);
    p.ls.assign(static_cast<size_t>(n) * c.ls_width, 0);
    p.local_scale.assign(n, 0);
    // Scale floor: 1e-12 times the larger coordinate range (at least 1)
    // (cpp/src/prepared_problem.cpp:68-73).
    double min_x = p.xy[0], max_x = min_x, min_y = p.xy[1], max_y = min_y;
    for (u32 i = 0; i < n; ++i) {
        min_x = std::min<double>(min_x, p.xy[2 * i]); max_x = std::max<double>(max_x, p.xy[2 * i]);
        min_y = std::min<double>(min_y, p.xy[2 * i + 1]); max_y = std::max<double>(max_y, p.xy[2 * i + 1]);
    }
    p.scale_epsilon = 1e-12 * std::max({max_x - min_x, max_y - min_y, 1.0});
    std::vector<std::pair<i64, u32>> row;
    for (u32 a = 0; a < n; ++a) {
        row.clear();
        for (u32 b = 0; b < n; ++b) if (a != b) row.emplace_back(p.d(a, b), b);
        std::partial_sort(row.begin(), row.begin() + needed, row.end());
        for (u32 j = 0; j < scale_width; ++j) p.local_scale[a] += static_cast<double>(row[j].first) / scale_width;
        for (u32 j = 0; j < c.primary_width; ++j) p.primary[static_cast<size_t>(a) * c.primary_width + j] = row[j].second;
        for (u32 j = 0; j < c.backup_width; ++j)
            p.backup[static_cast<size_t>(a) * c.backup_width + j] = row[c.primary_width + j].second;
        for (u32 j = 0; j < c.ls_width; ++j) p.ls[static_cast<size_t>(a) * c.ls_width + j] = row[j].second;
    }
    std::vector<u32> nn;
    std::vector<std::uint8_t> seen(n, 0);
    u32 current = 0;
    for (u32 i = 0; i < n; ++i) {
        nn.push_back(current); seen[current] = 1;
        u32 best_node = n; i64 best = std::numeric_limits<i64>::max();
        for (u32 b = 0; b < n; ++b) if (!seen[b]) {
            const i64 value = p.d(current, b);
            if (value < best) { best = value; best_node = b; }
        }
        current = best_node;
    }
    Tour t; t.assign(nn.data(), n, &p);
    std::vector<u32> checklist(nn);
    std::vector<std::uint8_t> queued(n, 1);
    two_opt(t, p.ls.data(), c.ls_width, checklist, queued, c.initial_ls_evaluation_limit, false);
    p.initial = t.order;
    p.initial_cost = t.recomputed();
    if (p.cheap_cost <= p.initial_cost) { p.initial = p.cheap; p.initial_cost = p.cheap_cost; }
}

// candidate_trail_limits (cpp/src/faco_cpu.cpp:296-304) with p computed once.
void trail_limits(u32 width, double p_best, double retention, double cost, double* minimum, double* maximum) {
    const double mx = 1.0 / (cost * (1.0 - retention));
    const double p = std::pow(p_best, 1.0 / width);
    *maximum = mx;
    *minimum = std::min(mx, mx * (1.0 - p) / ((width - 1.0) * p));
}

double heuristic(i64 d, double beta) {
    if (d <= 0) return 1.0;
    return beta == 1.0 ? 1.0 / static_cast<double>(d) : 1.0 / std::pow(static_cast<double>(d), beta);
}

// One colony of fixed-iteration FACO: restates FixedFacoGpu::run_iterations
// (cuda/fixed_faco.cu:105-177) with reset_colony / initialize_products /
// reduce_and_select / update_pheromone (cuda/colony_state.cuh:66-229), i.e.
// the same per-batch pipeline the reference test replays on the CPU
// (tests/cpp/test_fixed_faco.cpp:56-125).
struct ColonyRun {
    std::vector<u32> parent, epoch, global;
    i64 epoch_cost = 0, global_cost = 0, parent_cost = 0;
    double minimum = 0, maximum = 0, default_trail = 0, source_uniform = 0;
    u32 iteration_best = 0, source_is_epoch = 0;
    std::vector<double> trails, products, eta;
    u64 construction_steps = 0, ls_evaluations = 0, constraint_rejections = 0;
    // Escape mode: footprints (per city: the LS row of that city had
    // replaced slots in the ant that produced the tour) of the parent, epoch
    // and global tours, and the colony's escape counters
    std::vector<std::uint8_t> parent_fp, epoch_fp, global_fp;
    u64 esc[kEscCounters] = {};
};

void colony_reset(const Problem& p, ColonyRun& r) {
    const u32 n = p.n, pw = p.s.primary_width;
    r.parent = r.epoch = r.global = p.initial;
    r.epoch_cost = r.global_cost = r.parent_cost = p.initial_cost;
    trail_limits(pw, p.s.p_best, p.s.retention, static_cast<double>(r.epoch_cost), &r.minimum, &r.maximum);
    r.default_trail = r.maximum;
    r.trails.assign(static_cast<size_t>(n) * pw, r.maximum);
    r.eta.resize(static_cast<size_t>(n) * pw);
    r.products.resize(static_cast<size_t>(n) * pw);
    for (u32 a = 0; a < n; ++a) for (u32 j = 0; j < pw; ++j) {
        const size_t i = static_cast<size_t>(a) * pw + j;
        r.eta[i] = p.primary[i] < n ? heuristic(p.d(a, p.primary[i]), p.s.beta) : 0.0;   // padding: never available
        if (!p.prior.empty()) r.eta[i] *= static_cast<double>(p.prior[i]);
        r.products[i] = r.trails[i] * r.eta[i];
    }
    r.construction_steps = r.ls_evaluations = r.constraint_rejections = 0;
    r.source_uniform = 0; r.source_is_epoch = 0; r.iteration_best = 0;
    r.parent_fp.assign(n, 0); r.epoch_fp.assign(n, 0); r.global_fp.assign(n, 0);
    std::fill(r.esc, r.esc + kEscCounters, 0ull);
}

// reduce_and_select + update_pheromone for one batch given every ant's result.
void colony_select_and_update(const Problem& p, ColonyRun& r, const std::vector<AntResult>& ants,
                              u64 key, u32 batch) {
    const u32 n = p.n, pw = p.s.primary_width;
    u32 best = 0;
    for (u32 a = 0; a < ants.size(); ++a) {
        r.construction_steps += ants[a].c.steps;
        r.ls_evaluations += ants[a].ls.evaluations;
        r.constraint_rejections += ants[a].ls.rejections;
        for (u32 k = 0; k < kEscCounters; ++k) r.esc[k] += ants[a].esc[k];
        if (ants[a].final_cost < ants[best].final_cost) best = a;
    }
    r.iteration_best = best;
    const i64 cost = ants[best].final_cost;
    const bool improve_epoch = cost < r.epoch_cost, improve_global = cost < r.global_cost;
    if (improve_epoch) { r.epoch_cost = cost; r.epoch = ants[best].tour; }
    if (improve_global) { r.global_cost = cost; r.global = ants[best].tour; }
    trail_limits(pw, p.s.p_best, p.s.retention, static_cast<double>(r.epoch_cost), &r.minimum, &r.maximum);
    Stream coin(key, batch, 0xffffffffu, 0);
    r.source_uniform = coin.uniform53();
    r.source_is_epoch = r.source_uniform < p.s.epoch_source_probability;
    r.parent_cost = r.source_is_epoch ? r.epoch_cost : cost;
    r.parent = r.source_is_epoch ? r.epoch : ants[best].tour;
    if (p.escape) {
        // footprints follow their tours (cuda/colony_state.cuh:201-206)
        const std::vector<std::uint8_t>& value = ants[best].footprint;
        if (improve_epoch) r.epoch_fp = value;
        if (improve_global) r.global_fp = value;
        r.parent_fp = r.source_is_epoch ? r.epoch_fp : value;
    }
    // Evaporate, then deposit once on each directed candidate entry whose
    // target is the parent predecessor or successor (SparsePheromone
    // evaporate/deposit, cpp/src/faco_cpu.cpp:322-343).
    std::vector<u32> position(n);
    for (u32 i = 0; i < n; ++i) position[r.parent[i]] = i;
    const double increment = 1.0 / static_cast<double>(r.parent_cost);
    for (u32 a = 0; a < n; ++a) {
        const u32 succ = r.parent[(position[a] + 1) % n], pred = r.parent[(position[a] + n - 1) % n];
        for (u32 j = 0; j < pw; ++j) {
            const size_t i = static_cast<size_t>(a) * pw + j;
            double v = std::max(r.minimum, r.trails[i] * p.s.retention);
            const u32 to = p.primary[i];
            if (to == succ || to == pred) v = std::min(r.maximum, v + increment);
            r.trails[i] = v;
            r.products[i] = v * r.eta[i];
        }
    }
    r.default_trail = std::max(r.minimum, r.default_trail * p.s.retention);
}

// One Escape-mode ant: construct_and_search with EscapeChoices and
// EscapeEdges (cuda/faco_device.cuh:161-300, 326-336, 435-447;
// cuda/escape_device.cuh:41-128). Each construction step draws the escape
// gate (one word in 16 opens it) from the escape stream of (key, batch, ant,
// step); an open gate swaps about a quarter of the current city's primary
// row for backup cities (slot view). A replaced slot weighs
// default_trail * d^-beta (the registered backup rows never hold primary
// cities, so the reference's pheromone lookup returns the default trail)
// and may add out-of-graph edges while the ant's 64-edge cache has room;
// other moves keep to graph or cached edges. A new-edge relocation that
// moves a city marks five anchors (current, its old successor, the city, its
// old neighbours). After construction every anchor's LS row becomes its
// escape view (escape stream of (key, batch, ant, city) in the LS domain),
// sorted by (distance, city); the cities the parent's footprint names and
// the anchors join the checklist in city order; the 2-opt then follows the
// same permission rule per slot.
// What one Escape ant did, for replay checks (the reference's diagnostic
// capture, cuda/batch_engine.cu:169-176): its accepted moves in order as
// (kind 0 relocation / 1 successor 2-opt / 2 predecessor 2-opt, a, b, slot,
// cache size after the move, permit), its anchors, LS rows, replaced flags
// and cache.
struct EscapeTrace {
    std::vector<u32> events, rows;
    std::vector<std::uint8_t> anchors, replaced;
    EscCache cache{};
};

AntResult escape_ant(const Problem& p, const ColonyRun& r, u64 key, u32 batch, u32 ant, u32 start, u32 target,
                     EscapeTrace* trace = nullptr) {
    AntResult res;
    const u32 n = p.n, pw = p.s.primary_width, bw = p.s.backup_width, lw = p.s.ls_width;
    const bool disabled = p.escape == 2;
    Tour reference, work;
    reference.assign(r.parent.data(), n, &p);
    work = reference;
    std::vector<std::uint8_t> visited(n, 0), queued(n, 0), anchors(n, 0);
    std::vector<u32> checklist;
    EscCache cache{};
    u64* st = res.esc;
    auto available = [&](u32 current, u32 node) {   // available_node with the ant's cache, no new edges
        return node < n && !visited[node] &&
               escape_relocation_proposal(p, work, cache, current, node, false).status == kAllowed;
    };
    u32 current = start;
    res.start = start;
    visited[current] = 1;
    while (res.c.mne < target && res.c.steps + 1 < n) {
        const u32 step = res.c.steps + 1;
        Stream ordinary(key, batch, ant, step);
        const double u = ordinary.uniform53();
        const u32* prow = p.primary.data() + static_cast<size_t>(current) * pw;
        const u32* brow = bw ? p.backup.data() + static_cast<size_t>(current) * bw : nullptr;
        EscRow row;
        for (u32 j = 0; j < pw; ++j) row.members[j] = prow[j];
        ++st[kOpportunities];
        EscStream gate(key, batch, ant, step, false);
        if (!disabled && (gate.draw() & 15u) == 0) {
            ++st[kGates];
            row = slot_view(prow, pw, brow, bw, n, current, gate);
            st[kReplacedSlots] += row.replacements;
        }
        double weights[kEscLsWidth] = {};
        bool open[kEscLsWidth] = {};
        double total = 0;
        u32 count = 0, chosen = n, slot = n;
        bool novel = false;
        for (u32 j = 0; j < pw; ++j) {
            const u32 node = row.members[j];
            if (node >= n || visited[node]) continue;
            const EscProposal q = escape_relocation_proposal(p, work, cache, current, node, row.replaced[j] != 0);
            st[kConstructionCapacity] += q.status == kCapacity ? 1 : 0;
            if (q.status != kAllowed) continue;
            open[j] = true;
            weights[j] = row.replaced[j] ? r.default_trail * heuristic(p.d(current, node), p.s.beta)
                                         : r.products[static_cast<size_t>(current) * pw + j];
            total += weights[j];
            ++count;
            chosen = node; novel = row.replaced[j] != 0; slot = j;
        }
        if (count) {
            if (total > 0) {
                double prefix = 0;
                const double threshold = u * total;
                for (u32 j = 0; j < pw; ++j) if (open[j]) {
                    prefix += weights[j];
                    if (threshold < prefix) { chosen = row.members[j]; novel = row.replaced[j] != 0; slot = j; break; }
                }
            } else {
                u32 k = static_cast<u32>(u * count);
                if (k >= count) k = count - 1;
                for (u32 j = 0; j < pw; ++j) if (open[j] && k-- == 0) {
                    chosen = row.members[j]; novel = row.replaced[j] != 0; slot = j; break;
                }
            }
        } else {
            for (u32 j = 0; j < bw && chosen == n; ++j) if (available(current, brow[j])) chosen = brow[j];
            if (chosen == n) {
                i64 best = std::numeric_limits<i64>::max();
                for (u32 v = 0; v < n; ++v) {
                    if (!available(current, v)) continue;
                    const i64 value = p.d(current, v);
                    if (value < best) { best = value; chosen = v; }
                }
            }
        }
        if (chosen >= n) { res.c.exhausted = 1; break; }
        const EscProposal q = escape_relocation_proposal(p, work, cache, current, chosen, novel);
        if (!commit_edges(cache, q)) throw std::runtime_error("escape relocation commit refused");
        st[kNewEdges] += q.count;
        const u32 previous = work.pred(chosen), after = work.succ(chosen), target_after = work.succ(current);
        visited[chosen] = 1;
        ++res.c.steps;
        if (work.relocate(current, chosen)) {
            ++res.c.relocations;
            if (trace) trace->events.insert(trace->events.end(), {0u, current, chosen, slot, cache.size, novel ? 1u : 0u});
            if (novel) {
                for (u32 v : {current, target_after, chosen, previous, after}) anchors[v] = 1;
                ++st[kEscapeRelocations];
            }
        }
        if (!reference.has_edge(current, chosen)) {
            ++res.c.mne;
            for (u32 v : {current, chosen, previous})
                if (!queued[v]) { queued[v] = 1; checklist.push_back(v); }
        }
        current = chosen;
    }
    res.construction_tour = work.order;
    res.construction_cost = work.cost;
    // build_ls_row (escape_device.cuh:111-126): anchors get their escape views
    std::vector<u32> rows(p.ls);
    std::vector<std::uint8_t> replaced(static_cast<size_t>(n) * lw, 0);
    res.footprint.assign(n, 0);
    std::vector<double> distance_of(n);
    for (u32 node = 0; node < n; ++node) {
        if (!anchors[node] || disabled) continue;
        EscStream random(key, batch, ant, node, true);
        const u32* lrow = p.ls.data() + static_cast<size_t>(node) * lw;
        EscRow view = slot_view(lrow, lw, bw ? p.backup.data() + static_cast<size_t>(node) * bw : nullptr, bw, n, node,
                                random);
        for (u32 v = 0; v < n; ++v) distance_of[v] = static_cast<double>(p.d(node, v));
        sort_row(view, distance_of.data());
        for (u32 j = 0; j < lw; ++j) {
            rows[static_cast<size_t>(node) * lw + j] = view.members[j];
            replaced[static_cast<size_t>(node) * lw + j] = view.replaced[j];
        }
        res.footprint[node] = view.replacements != 0;
    }
    // the parent's footprint and the anchors reactivate their cities
    // (cuda/faco_device.cuh:273-291)
    for (u32 node = 0; node < n; ++node) {
        if (r.parent_fp[node] && !queued[node]) {
            queued[node] = 1; checklist.push_back(node); ++st[kOldViewReactivations];
        }
        if (anchors[node]) {
            ++st[kAnchorNodes];
            if (!queued[node]) { queued[node] = 1; checklist.push_back(node); ++st[kAnchorReactivations]; }
            for (u32 j = 0; j < lw; ++j) st[kLsReplacedSlots] += replaced[static_cast<size_t>(node) * lw + j];
        }
    }
    const EscapeLs esc{rows.data(), replaced.data(), &cache, trace ? &trace->events : nullptr};
    res.ls = two_opt(work, p.ls.data(), lw, checklist, queued, p.s.ls_evaluation_limit, false, &esc);
    if (trace) {
        trace->anchors = anchors; trace->rows = rows; trace->replaced = replaced; trace->cache = cache;
    }
    st[kNewEdges] += res.ls.new_edges;
    st[kLsCapacity] += res.ls.capacity_rejections;
    res.tour = work.order;
    res.positions = work.pos;
    res.checklist = checklist;
    res.final_cost = work.cost;
    return res;
}

AntResult stochastic_ant(const Problem& p, const ColonyRun& r, u64 key, u32 batch, u32 ant, u32 target) {
    const u32 pw = p.s.primary_width;
    const u32 start = start_node(key, batch, ant, p.n);
    if (p.escape) return escape_ant(p, r, key, batch, ant, start, target);
    return run_ant(p, r.parent.data(), start, target, p.s.ls_evaluation_limit,
        [&](u32 current, u32 step, const std::vector<std::uint8_t>& visited, const Tour& work) {
            Stream stream(key, batch, ant, step);
            const double u = stream.uniform53();
            int stage = 0;
            return select_next(p, current, r.products.data() + static_cast<size_t>(current) * pw, visited, u, &stage,
                               &work);
        });
}

// ===========================================================================
// Restart and perturbation control (the GP / baseline controlled loop).
// Restates cpp/include/gp_faco/control_ops.hpp, cuda/control_state.cuh,
// cpp/include/gp_faco/baseline_policy.hpp, cpp/include/gp_faco/factorial_policy.hpp,
// cuda/baseline_policy.cuh, cuda/gp_score.cuh and the per-batch order of
// cuda/batch_engine.cu:880-1060 (features, decision, action, construction,
// selection, pheromone, fingerprints, archive and feedback).

constexpr u32 kArchive = 4, kRegion = 16, kSamples = 64;
constexpr double kBand = 0.02;

struct Ident {
    u64 first = 0, second = 0;
    bool operator==(const Ident& o) const { return first == o.first && second == o.second; }
    bool operator!=(const Ident& o) const { return !(*this == o); }
    bool operator<(const Ident& o) const { return first < o.first || (first == o.first && second < o.second); }
};

// fingerprint_edge / fingerprint (control_ops.hpp:43-56): XOR over the
// undirected edges of two SplitMix64 images of (min << 32 | max).
Ident edge_ident(u32 a, u32 b) {
    if (b < a) std::swap(a, b);
    const u64 e = (static_cast<u64>(a) << 32) | b;
    return {mixed(e ^ 0xb5ad4eceda1ce2a9ULL), mixed(e ^ 0xd6e8feb86659fd93ULL)};
}

Ident tour_ident(const u32* t, u32 n) {
    Ident r;
    for (u32 i = 0; i < n; ++i) {
        const Ident e = edge_ident(t[i], t[i + 1 == n ? 0 : i + 1]);
        r.first ^= e.first; r.second ^= e.second;
    }
    return r;
}

// A tour with positions, integer cost and identity (TourRecord,
// cpp/include/gp_faco/control_cpu.hpp:8-15).
struct Rec {
    std::vector<u32> order, pos;
    i64 cost = 0;
    Ident id;
    std::vector<std::uint8_t> fp;   // Escape mode: the tour's footprint (empty: none)
    u32 n() const { return static_cast<u32>(order.size()); }
    u32 succ(u32 v) const { return order[pos[v] + 1 == n() ? 0 : pos[v] + 1]; }
    u32 pred(u32 v) const { return order[pos[v] == 0 ? n() - 1 : pos[v] - 1]; }
    bool has(u32 a, u32 b) const { return succ(a) == b || pred(a) == b; }
};

// Diagnostic mode of the controlled run: every tour identity reads as zero,
// so duplicate detection rests on the full comparison alone
// (force_fingerprint_collisions, cuda/control_state.cuh:19-64). Set before a
// run, read by its worker threads.
bool g_force_collisions = false;

Rec make_rec(const u32* order, u32 n, i64 cost) {
    Rec r;
    r.order.assign(order, order + n);
    r.pos.assign(n, 0);
    for (u32 i = 0; i < n; ++i) r.pos[order[i]] = i;
    r.cost = cost;
    r.id = g_force_collisions ? Ident{} : tour_ident(order, n);
    return r;
}

// canonical_compare (control_ops.hpp:60-71): both tours read from city 0 in
// the direction of its smaller neighbour.
int canonical_compare(const Rec& a, const Rec& b) {
    const u32 n = a.n();
    const bool fa = a.succ(0) < a.pred(0), fb = b.succ(0) < b.pred(0);
    u32 pa = a.pos[0], pb = b.pos[0];
    for (u32 i = 1; i < n; ++i) {
        pa = fa ? (pa + 1 == n ? 0 : pa + 1) : (pa == 0 ? n - 1 : pa - 1);
        pb = fb ? (pb + 1 == n ? 0 : pb + 1) : (pb == 0 ? n - 1 : pb - 1);
        if (a.order[pa] != b.order[pb]) return a.order[pa] < b.order[pb] ? -1 : 1;
    }
    return 0;
}
bool same_tour(const Rec& a, const Rec& b) { return a.n() == b.n() && canonical_compare(a, b) == 0; }

// ControlRandom (control_ops.hpp:81-89): rejection-sampled SplitMix64 words.
struct ControlRandom {
    u64 key, counter = 0;
    u32 bounded(u32 bound) {
        const u64 threshold = (0ULL - static_cast<u64>(bound)) % bound;
        u64 word;
        do { word = mixed(key + counter++); } while (word < threshold);
        return static_cast<u32>(word % bound);
    }
};

// rank-th city not in existing[0..count) (control_ops.hpp:92-102).
u32 remaining_node(u32 rank, const u32* existing, u32 count) {
    std::vector<u32> sorted(existing, existing + count);
    std::sort(sorted.begin(), sorted.end());
    u32 node = rank;
    for (u32 v : sorted) if (v <= node) ++node;
    return node;
}

// sample_nodes (control_ops.hpp:104-108).
void sample_nodes(u32 n, u64 instance_key, u32* out) {
    const u32 count = std::min(n, kSamples);
    ControlRandom random{mixed(instance_key ^ 0x6a09e667f3bcc909ULL)};
    for (u32 i = 0; i < count; ++i) out[i] = remaining_node(random.bounded(n - i), out, i);
}

bool has_node(const u32* nodes, u32 count, u32 v) {
    for (u32 i = 0; i < count; ++i) if (nodes[i] == v) return true;
    return false;
}

// build_region (control_ops.hpp:120-146): regions 0-1 are tour segments of
// the reference, region 2 a breadth-first candidate neighbourhood of a random
// city, region 3 (and any shortfall) random distinct cities.
void build_region(const Rec& ref, const u32* primary, u32 width, u64 key, u32 batch, u32 mode, u32 region,
                  u32* nodes) {
    const u32 n = ref.n(), count = std::min(n, kRegion);
    ControlRandom random{mixed(key ^ 0xbb67ae8584caa73bULL ^ mixed(static_cast<u64>(batch) ^ 0x3c6ef372fe94f82bULL) ^
                               mixed(static_cast<u64>(mode * 4 + region + 1) ^ 0xa54ff53a5f1d36f1ULL))};
    u32 used = 0;
    if (region < 2) {
        const u32 start = random.bounded(n);
        for (u32 i = 0; i < count; ++i) nodes[i] = ref.order[(start + i) % n];
        return;
    }
    if (region == 2) {
        nodes[used++] = random.bounded(n);
        for (u32 head = 0; head < used && used < count; ++head)
            for (u32 j = 0; j < width && used < count; ++j) {
                const u32 v = primary[static_cast<size_t>(nodes[head]) * width + j];
                if (v < n && !has_node(nodes, used, v)) nodes[used++] = v;
            }
    }
    while (used < count) {
        const u32 v = remaining_node(random.bounded(n - used), nodes, used);
        nodes[used++] = v;
    }
}

// sampled_difference (control_ops.hpp:148-157).
u32 sampled_difference(const Rec& cand, const Rec& active, const u32* samples, u32 count) {
    u32 missing = 0;
    for (u32 i = 0; i < count; ++i) {
        const u32 v = samples[i];
        missing += !active.has(v, cand.succ(v));
        missing += !active.has(v, cand.pred(v));
    }
    return missing;
}

double unit_clip(double v) { return v < 0 ? 0 : v > 1 ? 1 : v; }

struct Feedback {
    float return_rate = 0, ls_work = 0;
    u64 stagnant = 0, epoch_batches = 0, restarts = 0;
};

// update_feedback / restart_feedback (control_ops.hpp:175-190): EMA with
// alpha = 1/16 in FP32, saturating counters.
void update_feedback(Feedback& f, bool improved, double returned_fraction, double work_fraction, u32 ants) {
    if (!ants) return;
    const float returned = static_cast<float>(unit_clip(returned_fraction));
    const float work = static_cast<float>(unit_clip(work_fraction));
    f.return_rate += (returned - f.return_rate) * 0.0625f;
    f.ls_work += (work - f.ls_work) * 0.0625f;
    if (improved) f.stagnant = 0;
    else if (f.stagnant != UINT64_MAX) ++f.stagnant;
    if (f.epoch_batches != UINT64_MAX) ++f.epoch_batches;
}
void restart_feedback(Feedback& f) {
    f.return_rate = f.ls_work = 0;
    f.epoch_batches = 0;
    if (f.restarts != UINT64_MAX) ++f.restarts;
}

struct Control {
    Feedback fb;
    std::vector<Rec> archive;
    i64 tracked_global = 0, tracked_epoch = 0;
    Ident active, epoch_id;
};

// Baseline and factorial policies: same field layout as the reference's
// BaselinePolicy (baseline_policy.hpp:20-27) and FactorialPolicy
// (factorial_policy.hpp:17-20).
struct Baseline {
    u32 kind, mne_level, max_mne_level, region, restart_mode, restart_period;
    double restart_probability;
    u32 stagnation_step, restart_stagnation, restart_cooldown, pad;
};
static_assert(sizeof(Baseline) == 48, "baseline layout");

// baseline_action (baseline_policy.hpp:59-79).
u32 baseline_action(const Baseline& p, const Feedback& f, u32 batch, double u, u32 mask) {
    u32 level = p.mne_level;
    bool restart = false;
    if (p.kind == 0) {
        restart = p.restart_mode == 1 ? batch > 0 && batch % p.restart_period == 0
                                      : p.restart_mode == 2 && u < p.restart_probability;
    } else {
        if (p.stagnation_step) {
            const u64 extra = f.stagnant / p.stagnation_step;
            const u32 room = p.max_mne_level - p.mne_level;
            level += extra < room ? static_cast<u32>(extra) : room;
        }
        restart = p.restart_stagnation && f.stagnant >= p.restart_stagnation && f.epoch_batches >= p.restart_cooldown;
    }
    const u32 keep = p.region * 4 + level, requested = keep + 16;
    return restart && (mask & (1u << requested)) ? requested : keep;
}

// baseline_uniform (cuda/baseline_policy.cuh:10-16): colony slot ant
// 0xffffffff, step 1; only Bernoulli restarts consume it.
double baseline_uniform(const Baseline& p, u64 key, u32 batch) {
    if (p.kind != 0 || p.restart_mode != 2) return 0;
    Stream s(key, batch, 0xffffffffu, 1);
    return s.uniform53();
}

// factorial_mask (factorial_policy.hpp:32-45); variant 0..3 = M00, M10, M01, M11.
u32 factorial_mask(u32 variant, const Baseline& p, const Feedback& f, u32 batch, double u, u32 legal) {
    if (variant == 3) return legal;
    if (variant == 0) return 1u << baseline_action(p, f, batch, u, legal);
    const u32 requested = baseline_action(p, f, batch, u, 0xffffffffu);
    if (variant == 2) {
        const u32 keep = requested % 16;
        return legal & ((1u << keep) | (1u << (keep + 16)));
    }
    const u32 allowed = legal & (requested >= 16 ? 0xffff0000u : 0xffffu);
    return allowed ? allowed : legal & 0xffffu;
}

// Program IR v1 record (cpp/include/gp_faco/program.hpp:10-19) and the
// three-tree controller (:27-30).
struct Program {
    std::uint16_t ir_version, numeric_spec_id, feature_spec_id, length, constants_count;
    std::uint8_t opcode[63], operand[63];
    float constants[63];
};
static_assert(sizeof(Program) == 388, "program layout");
struct Controller {
    u32 kind;
    Program trees[3];
};
static_assert(sizeof(Controller) == 1168, "controller layout");

// apply_operation (cpp/include/gp_faco/program_ops.hpp:13-31), the CPU
// formula for AQ.
float apply_operation(int op, float left, float right) {
    float r = 0.0f;
    switch (op) {
        case 2: r = left + right; break;
        case 3: r = left - right; break;
        case 4: r = left * right; break;
        case 5: r = std::fmin(left, right); break;
        case 6: r = std::fmax(left, right); break;
        case 7: r = std::fabs(left); break;
        case 8: r = left * (1.0f / std::sqrt(1.0f + right * right)); break;
        default: break;
    }
    return std::fmin(8.0f, std::fmax(-8.0f, r));
}

// Stack VM value of one action; features are [12][32] of one colony.
float program_value(const Program& p, const float* features, u32 action) {
    float stack[6] = {0, 0, 0, 0, 0, 0};
    int top = 0;
    for (int i = 0; i < p.length; ++i) {
        const int op = p.opcode[i];
        if (op == 0) stack[top++] = features[p.operand[i] * 32 + action];
        else if (op == 1) stack[top++] = p.constants[p.operand[i]];
        else {
            const float right = op == 7 ? 0.0f : stack[--top];
            const float left = stack[--top];
            stack[top++] = apply_operation(op, left, right);
        }
    }
    return stack[0];
}

// Smallest legal index within 1e-6 of the largest legal value
// (cuda/gp_score.cuh:24-34, cpp/src/program_cpu.cpp:104-113); 32 if none.
u32 choose(const float* values, const bool* legal, u32 width) {
    float maximum = -std::numeric_limits<float>::infinity();
    for (u32 i = 0; i < width; ++i) if (legal[i]) maximum = std::max(maximum, values[i]);
    for (u32 i = 0; i < width; ++i) if (legal[i] && maximum - values[i] <= 1e-6f) return i;
    return 32;
}

// score_actions (cuda/gp_score.cuh:64-102): one tree over the 32 actions.
u32 score_program(const Program& p, const float* features, u32 mask, float* scores) {
    bool legal[32];
    for (u32 a = 0; a < 32; ++a) { scores[a] = program_value(p, features, a); legal[a] = (mask >> a) & 1u; }
    return choose(scores, legal, 32);
}

// score_controllers (cuda/gp_score.cuh:37-62): kind 0 = one tree; kind 1 =
// conditional restart (2-way, stride 16), region (4-way, stride 4), MNE level
// (4-way, stride 1); a block is legal when any action under it is.
u32 score_controller(const Controller& c, const float* features, u32 mask, float* scores) {
    if (!c.kind) return score_program(c.trees[0], features, mask, scores);
    for (u32 a = 0; a < 32; ++a) scores[a] = 0.0f;
    u32 prefix = 0;
    for (u32 role = 0; role < 3; ++role) {
        const u32 width = role == 0 ? 2 : 4, stride = role == 0 ? 16 : role == 1 ? 4 : 1;
        const u32 slot = role == 0 ? 0 : role == 1 ? 2 : 6;
        float values[4];
        bool legal[4];
        for (u32 l = 0; l < width; ++l) {
            const u32 action = prefix + l * stride;
            legal[l] = ((mask >> action) & ((1u << stride) - 1u)) != 0;
            values[l] = program_value(c.trees[role], features, action);
            scores[slot + l] = values[l];
        }
        prefix += stride * choose(values, legal, width);
    }
    return prefix;
}

struct ControlOut {
    u32 archive_size, pad;
    u64 stagnant_batches, epoch_batches, restarts;
    float return_rate, ls_work;
};

struct Decision {
    float features[12 * 32];
    float scores[32];
    u32 mask = 0, alternative = kArchive, action = 0;
    u32 regions[2][4][kRegion];
    u32 region_count = 0;
};

// build_control_features (cuda/control_state.cuh:200-307) for one colony.
void control_features(const Problem& p, const ColonyRun& r, const Control& ctl, const u32* samples, u64 key,
                      u32 batch, double elapsed, u32 experiment_mask, Decision& d) {
    const u32 n = p.n, pw = p.s.primary_width, sample_count = std::min(n, kSamples);
    const Rec active = make_rec(r.parent.data(), n, r.parent_cost);
    u32 alternative = kArchive, largest = 0;
    for (u32 i = 0; i < ctl.archive.size(); ++i) {
        const Rec& cand = ctl.archive[i];
        if (cand.id == ctl.active && same_tour(cand, active)) continue;
        const u32 missing = sampled_difference(cand, active, samples, sample_count);
        bool better_tie = false;
        if (alternative != kArchive && missing == largest) {
            const Rec& cur = ctl.archive[alternative];
            better_tie = cand.cost < cur.cost ||
                (cand.cost == cur.cost && (cand.id < cur.id || (cand.id == cur.id && canonical_compare(cand, cur) < 0)));
        }
        if (alternative == kArchive || missing > largest || better_tie) { alternative = i; largest = missing; }
    }
    const double global = static_cast<double>(r.global_cost);
    float gap[2];
    gap[0] = static_cast<float>(unit_clip((static_cast<double>(r.parent_cost) / global - 1) / kBand));
    gap[1] = alternative < kArchive
        ? static_cast<float>(unit_clip((static_cast<double>(ctl.archive[alternative].cost) / global - 1) / kBand)) : 0;
    const float difference = alternative < kArchive ? static_cast<float>(largest) / (2 * sample_count) : 0;
    d.alternative = alternative;
    d.mask = experiment_mask & (alternative < kArchive ? 0xffffffffu : 0xffffu);
    d.region_count = std::min(n, kRegion);
    const u32 count = d.region_count, archive_size = static_cast<u32>(ctl.archive.size());
    for (u32 mode = 0; mode < 2; ++mode) for (u32 region = 0; region < 4; ++region) {
        u32* nodes = d.regions[mode][region];
        for (u32 i = 0; i < kRegion; ++i) nodes[i] = 0;
        const bool valid = mode == 0 || alternative < kArchive;
        float regional[4] = {0, 0, 0, 0};
        if (valid) {
            const Rec& ref = mode == 0 ? active : ctl.archive[alternative];
            build_region(ref, p.primary.data(), pw, key, batch, mode, region, nodes);
            double excess = 0, disagreement = 0, strength = 0, dispersion = 0;
            for (u32 i = 0; i < count; ++i) {
                const u32 v = nodes[i], prev = ref.pred(v), next = ref.succ(v);
                const double scale = p.local_scale[v] < p.scale_epsilon ? p.scale_epsilon : p.local_scale[v];
                const double near = 0.5 * static_cast<double>(p.d(v, prev)) + 0.5 * static_cast<double>(p.d(v, next));
                excess += unit_clip((near / scale - 1) / 3);
                double agree = 0, pher[2];
                const u32 neighbours[2] = {prev, next};
                double spread = 0;
                for (u32 k = 0; k < 2; ++k) {
                    const u32 w = neighbours[k];
                    for (u32 j = 0; j < archive_size; ++j) agree += !ctl.archive[j].has(v, w);
                    double trail = r.default_trail;
                    for (u32 j = 0; j < pw; ++j)
                        if (p.primary[static_cast<size_t>(v) * pw + j] == w) { trail = r.trails[static_cast<size_t>(v) * pw + j]; break; }
                    pher[k] = r.minimum == r.maximum ? 1 : unit_clip((trail - r.minimum) / (r.maximum - r.minimum));
                    spread += !has_node(nodes, count, w);
                }
                disagreement += agree;
                strength += pher[0]; strength += pher[1];
                dispersion += spread;
            }
            regional[0] = static_cast<float>(excess / count);
            regional[1] = static_cast<float>(disagreement / (2 * count * archive_size));
            regional[2] = static_cast<float>(strength / (2 * count));
            regional[3] = static_cast<float>(dispersion / (2 * count));
        }
        for (u32 level = 0; level < 4; ++level) {
            const u32 action = mode * 16 + region * 4 + level;
            const float values[12] = {static_cast<float>(unit_clip(elapsed)),
                static_cast<float>(ctl.fb.stagnant < 32 ? ctl.fb.stagnant / 32.0 : 1),
                ctl.fb.return_rate, ctl.fb.ls_work, static_cast<float>(mode), (level + 1) * 0.25f,
                gap[mode], mode ? difference : 0.0f, regional[0], regional[1], regional[2], regional[3]};
            for (u32 f = 0; f < 12; ++f) d.features[f * 32 + action] = values[f];
        }
    }
}

// apply_control_action (cuda/control_state.cuh:309-364): action >= 16
// restarts from the archive alternative.
void control_restart(const Problem& p, ColonyRun& r, Control& ctl, u32 alternative, bool initializing) {
    const u32 n = p.n, pw = p.s.primary_width;
    const Rec& alt = ctl.archive[alternative];
    r.epoch_cost = r.parent_cost = alt.cost;
    trail_limits(pw, p.s.p_best, p.s.retention, static_cast<double>(r.epoch_cost), &r.minimum, &r.maximum);
    r.default_trail = r.maximum;
    r.source_uniform = 0; r.source_is_epoch = 0; r.iteration_best = 0;
    ctl.tracked_epoch = r.epoch_cost;
    ctl.active = ctl.epoch_id = alt.id;
    if (!initializing) restart_feedback(ctl.fb);
    r.parent = r.epoch = alt.order;
    if (p.escape) {
        // the restart carries the archive entry's footprint
        // (reference tests/cpp/test_escape_engine.cpp:93-98)
        if (alt.fp.empty()) std::fill(r.parent_fp.begin(), r.parent_fp.end(), 0);
        else r.parent_fp = alt.fp;
        r.epoch_fp = r.parent_fp;
    }
    for (size_t i = 0; i < static_cast<size_t>(n) * pw; ++i) {
        r.trails[i] = r.maximum;
        r.products[i] = r.maximum * r.eta[i];
    }
}

// Archive selection of update_control (cuda/control_state.cuh:127-160):
// slot 0 is the global best `first`; slots 1-3 take, in order, the best
// pool entries by (cost, identity, canonical order, index) that lie within
// the 2% band of the global cost and are not the same tour as a chosen one.
std::vector<u32> archive_select(const std::vector<Rec>& pool, u32 first, i64 global_cost) {
    const double band = static_cast<double>(global_cost) * (1 + kBand);
    std::vector<u32> chosen{first};
    auto duplicate = [&](u32 a, u32 b) { return a == b || (pool[a].id == pool[b].id && same_tour(pool[a], pool[b])); };
    auto better = [&](u32 a, u32 b) {
        if (a == UINT32_MAX) return b;
        if (pool[a].cost != pool[b].cost) return pool[a].cost < pool[b].cost ? a : b;
        if (pool[a].id != pool[b].id) return pool[a].id < pool[b].id ? a : b;
        const int order = canonical_compare(pool[a], pool[b]);
        if (order) return order < 0 ? a : b;
        return std::min(a, b);
    };
    for (u32 slot = 1; slot < kArchive; ++slot) {
        u32 best = UINT32_MAX;
        for (u32 c = 0; c < pool.size(); ++c) {
            if (static_cast<double>(pool[c].cost) > band) continue;
            bool dup = false;
            for (u32 j : chosen) if (duplicate(c, j)) { dup = true; break; }
            if (!dup) best = better(best, c);
        }
        if (best == UINT32_MAX) break;
        chosen.push_back(best);
    }
    return chosen;
}

// update_control (cuda/control_state.cuh:105-198) after selection.
void control_update(const Problem& p, const ColonyRun& r, Control& ctl, const std::vector<AntResult>& ants) {
    const u32 n = p.n, A = static_cast<u32>(ants.size());
    std::vector<Rec> pool = ctl.archive;
    for (const auto& a : ants) {
        pool.push_back(make_rec(a.tour.data(), n, a.final_cost));
        pool.back().fp = a.footprint;   // archive footprints follow their entries (control_state.cuh:177-181)
    }
    const u32 archived = static_cast<u32>(ctl.archive.size());
    const bool improved = r.global_cost < ctl.tracked_global;
    const std::vector<u32> chosen = archive_select(pool, improved ? archived + r.iteration_best : 0u, r.global_cost);
    double returned = 0, work = 0;
    const u64 limit = p.s.ls_evaluation_limit;
    for (u32 a = 0; a < A; ++a) {
        returned += pool[archived + a].id == ctl.active;
        if (limit) work += static_cast<double>(ants[a].ls.evaluations) / static_cast<double>(limit);
    }
    update_feedback(ctl.fb, improved, returned / A, work / A, A);
    const Ident best_id = pool[archived + r.iteration_best].id;
    if (r.epoch_cost < ctl.tracked_epoch) ctl.epoch_id = best_id;
    ctl.active = r.source_is_epoch ? ctl.epoch_id : best_id;
    ctl.tracked_global = r.global_cost;
    ctl.tracked_epoch = r.epoch_cost;
    std::vector<Rec> next;
    for (u32 c : chosen) next.push_back(pool[c]);
    ctl.archive = std::move(next);
}

void control_init(const Problem& p, Control& ctl) {
    ctl = Control{};
    ctl.archive.push_back(make_rec(p.initial.data(), p.n, p.initial_cost));
    ctl.tracked_global = ctl.tracked_epoch = p.initial_cost;
    ctl.active = ctl.epoch_id = ctl.archive[0].id;
}

AntResult controlled_ant(const Problem& p, const ColonyRun& r, u64 key, u32 batch, u32 ant, u32 target,
                         const u32* start_nodes, u32 start_count) {
    const u32 pw = p.s.primary_width;
    const u32 start = start_nodes[start_node(key, batch, ant, start_count)];
    if (p.escape) return escape_ant(p, r, key, batch, ant, start, target);
    return run_ant(p, r.parent.data(), start, target, p.s.ls_evaluation_limit,
        [&](u32 current, u32 step, const std::vector<std::uint8_t>& visited, const Tour& work) {
            Stream stream(key, batch, ant, step);
            const double u = stream.uniform53();
            int stage = 0;
            return select_next(p, current, r.products.data() + static_cast<size_t>(current) * pw, visited, u, &stage,
                               &work);
        });
}

struct Instance {
    Problem p;
    std::vector<std::int32_t> xy;
};

template<class F> int guarded(F f) {
    try { f(); return 0; }
    catch (const std::invalid_argument& e) { g_error = e.what(); return 1; }
    catch (const std::exception& e) { g_error = e.what(); return 2; }
}

void check_coordinates(u32 n, const std::int32_t* xy) {
    if (n < 3 || n > 10000) throw std::invalid_argument("dimension must be in [3, 10000]");
    for (u32 i = 0; i < 2 * n; ++i)
        if (xy[i] < 0 || xy[i] > 32767) throw std::invalid_argument("coordinates must be integers in [0, 32767]");
}

}  // namespace

extern "C" {

const char* oracle_last_error() { return g_error.c_str(); }

std::int64_t oracle_euc2d(std::int32_t x1, std::int32_t y1, std::int32_t x2, std::int32_t y2) {
    const std::int32_t xy[4] = {x1, y1, x2, y2};
    return euc2d(xy, 0, 1);
}

void oracle_philox4x32_10(const u32* counter, const u32* key, u32* out) { philox4x32_10(counter, key, out); }

// 1: every later run draws from the alternate generator; 0: Philox again.
void oracle_set_alternate_stream(int on) { g_alternate_stream = on != 0; }

double oracle_uniform53(u64 key, u32 batch, u32 ant, u32 step) {
    Stream s(key, batch, ant, step);
    return s.uniform53();
}

u32 oracle_start_node(u64 key, u32 batch, u32 ant, u32 count) { return start_node(key, batch, ant, count); }

u64 oracle_colony_key(u64 instance_key, u64 seed) { return colony_key(instance_key, seed); }
u64 oracle_mixed(u64 value) { return mixed(value); }

void oracle_trail_limits(u32 width, double p_best, double retention, double cost, double* minimum, double* maximum) {
    trail_limits(width, p_best, retention, cost, minimum, maximum);
}

// Normalised settings and preparation outputs for one instance.
int oracle_prepare(u32 n, const std::int32_t* xy, const Settings* settings, Settings* normalized_out,
                   u32* primary, u32* backup, u32* ls, double* local_scale,
                   u32* initial, std::int64_t* initial_cost, std::int64_t* cheap_cost) {
    return guarded([&] {
        check_coordinates(n, xy);
        Problem p; p.n = n; p.xy = xy; p.s = normalized(*settings, n);
        prepare(p);
        *normalized_out = p.s;
        std::copy(p.primary.begin(), p.primary.end(), primary);
        std::copy(p.backup.begin(), p.backup.end(), backup);
        std::copy(p.ls.begin(), p.ls.end(), ls);
        std::copy(p.local_scale.begin(), p.local_scale.end(), local_scale);
        std::copy(p.initial.begin(), p.initial.end(), initial);
        *initial_cost = p.initial_cost;
        *cheap_cost = p.cheap_cost;
    });
}

// Explicit-choice construction + checklist 2-opt for one ant (the
// tests/cpp/test_cuda_faco.cpp:31-104 pattern). visit[0] is the start node,
// visit[1..] the chosen nodes in order. `ls` is [n][lw], distance-ordered.
// counters: mne, steps, relocations, exhausted, processed, checks,
// evaluations, accepted, reactivations, limit_reached, checklist_size.
int oracle_construct_ls(u32 n, const std::int32_t* xy, const u32* ls, u32 lw, const u32* parent,
                        const u32* visit, u32 mne_target, u64 limit,
                        u32* construction_tour, u32* tour, u32* positions, u32* checklist,
                        std::int64_t* costs /* construction, final */, u64* counters, u32 local_search) {
    return guarded([&] {
        check_coordinates(n, xy);
        Problem p; p.n = n; p.xy = xy; p.s.ls_width = lw; p.or_opt = local_search != 0;
        p.ls.assign(ls, ls + static_cast<size_t>(n) * lw);
        u32 cursor = 1;
        auto r = run_ant(p, parent, visit[0], mne_target, limit,
            [&](u32, u32, const std::vector<std::uint8_t>&, const Tour&) { return visit[cursor++]; });
        std::copy(r.construction_tour.begin(), r.construction_tour.end(), construction_tour);
        std::copy(r.tour.begin(), r.tour.end(), tour);
        std::copy(r.positions.begin(), r.positions.end(), positions);
        std::copy(r.checklist.begin(), r.checklist.end(), checklist);
        costs[0] = r.construction_cost; costs[1] = r.final_cost;
        const u64 values[11] = {r.c.mne, r.c.steps, r.c.relocations, r.c.exhausted, r.ls.processed,
                                r.ls.checks, r.ls.evaluations, r.ls.accepted, r.ls.reactivations,
                                r.ls.limit_reached, r.checklist.size()};
        std::copy(values, values + 11, counters);
    });
}

// Checklist 2-opt alone from a tour and an explicit checklist (the initial
// local search of preparation, cpp/src/prepared_problem.cpp:111-113).
int oracle_two_opt(u32 n, const std::int32_t* xy, const u32* ls, u32 lw, u32* tour, const u32* checklist_in,
                   u32 checklist_size, u64 limit, std::int64_t* cost, u64* evaluations, u32 local_search) {
    return guarded([&] {
        check_coordinates(n, xy);
        Problem p; p.n = n; p.xy = xy;
        Tour t; t.assign(tour, n, &p);
        std::vector<u32> checklist(checklist_in, checklist_in + checklist_size);
        std::vector<std::uint8_t> queued(n, 0);
        for (u32 v : checklist) queued[v] = 1;
        const LsStats st = two_opt(t, ls, lw, checklist, queued, limit, local_search != 0);
        std::copy(t.order.begin(), t.order.end(), tour);
        *cost = t.recomputed();
        *evaluations = st.evaluations;
    });
}

// The Hard-mode checklist 2-opt (plain 2-opt, graph given as an undirected
// CSR) with its count of rejected moves (faco_cpu.cpp:180-183).
int oracle_two_opt_hard(u32 n, const std::int32_t* xy, const u32* ls, u32 lw, u32* tour, const u32* checklist_in,
                        u32 checklist_size, u64 limit, const u32* graph_offsets, const u32* graph_neighbors,
                        std::int64_t* cost, u64* evaluations, u64* rejections) {
    return guarded([&] {
        check_coordinates(n, xy);
        Problem p; p.n = n; p.xy = xy;
        p.graph_offsets = graph_offsets; p.graph_neighbors = graph_neighbors;
        Tour t; t.assign(tour, n, &p);
        std::vector<u32> checklist(checklist_in, checklist_in + checklist_size);
        std::vector<std::uint8_t> queued(n, 0);
        for (u32 v : checklist) queued[v] = 1;
        const LsStats st = two_opt(t, ls, lw, checklist, queued, limit, false);
        std::copy(t.order.begin(), t.order.end(), tour);
        *cost = t.recomputed();
        *evaluations = st.evaluations;
        *rejections = st.rejections;
    });
}

// One roulette decision (select_next) with explicit inputs.
int oracle_select_next(u32 n, const std::int32_t* xy, u32 pw, u32 bw, const u32* primary_row,
                       const u32* backup_row, const double* products, const std::uint8_t* visited,
                       u32 current, double u, u32* chosen, int* stage) {
    return guarded([&] {
        Problem p; p.n = n; p.xy = xy; p.s.primary_width = pw; p.s.backup_width = bw;
        p.primary.assign(static_cast<size_t>(n) * pw, n);
        p.backup.assign(static_cast<size_t>(n) * bw, n);
        std::copy(primary_row, primary_row + pw, p.primary.begin() + static_cast<size_t>(current) * pw);
        std::copy(backup_row, backup_row + bw, p.backup.begin() + static_cast<size_t>(current) * bw);
        std::vector<std::uint8_t> v(visited, visited + n);
        *chosen = select_next(p, current, products, v, u, stage);
    });
}

// select_next in Hard mode: availability also needs the relocation of the
// candidate after `current` in `tour` to stay inside the CSR graph.
int oracle_select_next_hard(u32 n, const std::int32_t* xy, u32 pw, u32 bw, const u32* primary_row,
                            const u32* backup_row, const double* products, const std::uint8_t* visited, u32 current,
                            double u, const u32* tour, const u32* offsets, const u32* neighbors, u32* chosen,
                            int* stage) {
    return guarded([&] {
        Problem p; p.n = n; p.xy = xy; p.s.primary_width = pw; p.s.backup_width = bw;
        p.graph_offsets = offsets; p.graph_neighbors = neighbors;
        p.primary.assign(static_cast<size_t>(n) * pw, n);
        p.backup.assign(static_cast<size_t>(n) * bw, n);
        std::copy(primary_row, primary_row + pw, p.primary.begin() + static_cast<size_t>(current) * pw);
        std::copy(backup_row, backup_row + bw, p.backup.begin() + static_cast<size_t>(current) * bw);
        Tour t;
        t.assign(tour, n, &p);
        std::vector<std::uint8_t> v(visited, visited + n);
        *chosen = select_next(p, current, products, v, u, stage, &t);
    });
}

// Full fixed-iteration run of `colonies` independent colonies (instance i %
// instances, seed seeds[i]) for `batches` batches. Outputs per colony the
// global tour, global cost and work counters. Optional per-batch trace for
// colony 0: parent after each batch [batches][n] and trails after each batch
// [batches][n*pw]. Threads: OpenMP over (colony, ant) within a batch.
// A graph spec's own candidate rows (prepared_problem.cpp:128-166), right
// padded with n: primary [n][pw], backup [n][bw] and LS [n][lw] one after
// the other, replacing the prepared kNN rows (the initial tour stays).
size_t row_cells(const Problem& p) {
    return static_cast<size_t>(p.n) * (p.s.primary_width + p.s.backup_width + p.s.ls_width);
}
void apply_rows(Problem& p, const u32* rows) {
    const size_t n = p.n, pw = p.s.primary_width, bw = p.s.backup_width, lw = p.s.ls_width;
    p.primary.assign(rows, rows + n * pw);
    p.backup.assign(rows + n * pw, rows + n * (pw + bw));
    p.ls.assign(rows + n * (pw + bw), rows + n * (pw + bw + lw));
}

int oracle_run_faco(u32 n, u32 instances, const std::int32_t* xy_all, const u64* instance_keys,
                    u32 colonies, const u64* seeds, const Settings* settings, u32 batches,
                    u32 mne_target, int threads,
                    u32* global_tours, std::int64_t* global_costs, u64* work /* [colonies][2] */,
                    u32* trace_parent, double* trace_trails, std::int64_t* trace_costs /*[batches][3]*/,
                    double* elapsed_search_seconds, u32 local_search,
                    const float* multipliers /* [instances][n * pw] or null */,
                    const u32* graph_offsets /* [instances][n + 1] or null: Hard mode */,
                    const u32* graph_neighbors /* [instances][neighbor_stride] */, u32 neighbor_stride,
                    u32 timed_from /* elapsed_search_seconds covers batches timed_from.. only */,
                    u64* rejections /* [colonies] or null: Hard-mode 2-opt moves rejected */,
                    const u32* rows /* [instances][n (pw + bw + lw)] or null: a graph spec's own rows */,
                    u32 escape /* 0, or Escape mode on the graph: 1 gates on, 2 gates disabled */,
                    u64* escape_counters /* [colonies][11] or null */) {
    return guarded([&] {
        if (!instances || !colonies) throw std::invalid_argument("empty panel");
        if (graph_offsets && local_search) throw std::invalid_argument("Hard mode takes the plain 2-opt");
        if (escape > 2 || (escape && (!graph_offsets || multipliers)))
            throw std::invalid_argument("Escape mode needs a graph and takes no learned prior");
        std::vector<Problem> problems(instances);
        for (u32 g = 0; g < instances; ++g) {
            check_coordinates(n, xy_all + static_cast<size_t>(g) * 2 * n);
            problems[g].n = n; problems[g].xy = xy_all + static_cast<size_t>(g) * 2 * n;
            problems[g].s = normalized(*settings, n);
            prepare(problems[g]);
            if (rows) apply_rows(problems[g], rows + static_cast<size_t>(g) * row_cells(problems[g]));
            problems[g].or_opt = local_search != 0;
            problems[g].escape = escape;
            if (graph_offsets) {
                problems[g].graph_offsets = graph_offsets + static_cast<size_t>(g) * (n + 1);
                problems[g].graph_neighbors = graph_neighbors + static_cast<size_t>(g) * neighbor_stride;
                for (u32 i = 0; i < n; ++i)
                    if (!graph_has(problems[g].graph_offsets, problems[g].graph_neighbors, n, problems[g].initial[i],
                                   problems[g].initial[(i + 1) % n]))
                        throw std::invalid_argument("the graph must hold the initial tour");
            }
            if (multipliers) {
                const size_t cells = static_cast<size_t>(n) * problems[g].s.primary_width;
                problems[g].prior.assign(multipliers + g * cells, multipliers + (g + 1) * cells);
            }
        }
        const Settings s = problems[0].s;
        std::vector<ColonyRun> runs(colonies);
        std::vector<u64> keys(colonies);
        for (u32 c = 0; c < colonies; ++c) {
            colony_reset(problems[c % instances], runs[c]);
            keys[c] = colony_key(instance_keys[c % instances], seeds[c]);
        }
        std::vector<std::vector<AntResult>> results(colonies, std::vector<AntResult>(s.ants));
#ifdef _OPENMP
        if (threads > 0) omp_set_num_threads(threads);
        double started = omp_get_wtime();
#endif
        for (u32 batch = 0; batch < batches; ++batch) {
#ifdef _OPENMP
            if (batch == timed_from) started = omp_get_wtime();
#endif
            const long long total = static_cast<long long>(colonies) * s.ants;
#pragma omp parallel for schedule(dynamic, 1)
            for (long long i = 0; i < total; ++i) {
                const u32 c = static_cast<u32>(i / s.ants), a = static_cast<u32>(i % s.ants);
                results[c][a] = stochastic_ant(problems[c % instances], runs[c], keys[c], batch, a, mne_target);
            }
            for (u32 c = 0; c < colonies; ++c)
                colony_select_and_update(problems[c % instances], runs[c], results[c], keys[c], batch);
            if (trace_parent) std::copy(runs[0].parent.begin(), runs[0].parent.end(),
                                        trace_parent + static_cast<size_t>(batch) * n);
            if (trace_trails) std::copy(runs[0].trails.begin(), runs[0].trails.end(),
                                        trace_trails + static_cast<size_t>(batch) * n * s.primary_width);
            if (trace_costs) {
                trace_costs[3 * batch] = runs[0].epoch_cost;
                trace_costs[3 * batch + 1] = runs[0].global_cost;
                trace_costs[3 * batch + 2] = runs[0].parent_cost;
            }
        }
#ifdef _OPENMP
        if (elapsed_search_seconds) *elapsed_search_seconds = omp_get_wtime() - started;
#else
        if (elapsed_search_seconds) *elapsed_search_seconds = 0;
#endif
        for (u32 c = 0; c < colonies; ++c) {
            std::copy(runs[c].global.begin(), runs[c].global.end(), global_tours + static_cast<size_t>(c) * n);
            global_costs[c] = runs[c].global_cost;
            work[2 * c] = runs[c].construction_steps;
            work[2 * c + 1] = runs[c].ls_evaluations;
            if (rejections) rejections[c] = runs[c].constraint_rejections;
            if (escape_counters) std::copy(runs[c].esc, runs[c].esc + kEscCounters, escape_counters + kEscCounters * c);
        }
    });
}

// One Escape-mode ant on explicit colony state, with its trace: the parent
// tour, the colony's products [n][pw] and default trail, the parent's
// footprint; rows (pw + bw + lw per city, a graph spec's rows) or null for
// the prepared rows. Outputs: construction and final tour, anchors, LS rows
// [n][lw] and replaced flags, footprint, cache [64][2] and its size, events
// [2n][6] and their count, escape counters [11], ls [4] = evaluations,
// accepted moves, rejections, checklist size; costs [2] = construction, final.
int oracle_escape_ant(u32 n, const std::int32_t* xy, const Settings* settings, const u32* rows,
                      const u32* graph_offsets, const u32* graph_neighbors, const u32* parent, const double* products,
                      double default_trail, const std::uint8_t* parent_fp, u64 key, u32 batch, u32 ant, u32 start,
                      u32 target, u32 disabled, u32* construction_tour, u32* tour, std::uint8_t* anchors,
                      u32* ls_rows, std::uint8_t* replaced, std::uint8_t* footprint, u32* cache_edges,
                      u32* cache_size, u32* events, u32* event_count, u64* counters, u64* ls,
                      std::int64_t* costs) {
    return guarded([&] {
        check_coordinates(n, xy);
        Problem p; p.n = n; p.xy = xy; p.s = normalized(*settings, n);
        prepare(p);
        if (rows) apply_rows(p, rows);
        p.graph_offsets = graph_offsets; p.graph_neighbors = graph_neighbors;
        p.escape = disabled ? 2u : 1u;
        if (start >= n) throw std::invalid_argument("start city out of range");
        ColonyRun r;
        colony_reset(p, r);
        r.parent.assign(parent, parent + n);
        r.products.assign(products, products + static_cast<size_t>(n) * p.s.primary_width);
        r.default_trail = default_trail;
        r.parent_fp.assign(parent_fp, parent_fp + n);
        EscapeTrace trace;
        const AntResult a = escape_ant(p, r, key, batch, ant, start, target, &trace);
        std::copy(a.construction_tour.begin(), a.construction_tour.end(), construction_tour);
        std::copy(a.tour.begin(), a.tour.end(), tour);
        std::copy(trace.anchors.begin(), trace.anchors.end(), anchors);
        std::copy(trace.rows.begin(), trace.rows.end(), ls_rows);
        std::copy(trace.replaced.begin(), trace.replaced.end(), replaced);
        std::copy(a.footprint.begin(), a.footprint.end(), footprint);
        for (u32 i = 0; i < kEscCapacity; ++i) {
            cache_edges[2 * i] = i < trace.cache.size ? trace.cache.edges[i].a : 0u;
            cache_edges[2 * i + 1] = i < trace.cache.size ? trace.cache.edges[i].b : 0u;
        }
        *cache_size = trace.cache.size;
        if (trace.events.size() > static_cast<size_t>(12) * n) throw std::runtime_error("more than 2n moves");
        std::copy(trace.events.begin(), trace.events.end(), events);
        *event_count = static_cast<u32>(trace.events.size() / 6);
        std::copy(a.esc, a.esc + kEscCounters, counters);
        ls[0] = a.ls.evaluations; ls[1] = a.ls.accepted; ls[2] = a.ls.rejections; ls[3] = a.checklist.size();
        costs[0] = a.construction_cost; costs[1] = a.final_cost;
    });
}

// ---------------------------------------------------------------------------
// Control primitives, exported for tests.

void oracle_tour_ident(u32 n, const u32* tour, u64* out) {
    const Ident id = tour_ident(tour, n);
    out[0] = id.first; out[1] = id.second;
}

int oracle_canonical_compare(u32 n, const u32* a, const u32* b) {
    return canonical_compare(make_rec(a, n, 0), make_rec(b, n, 0));
}

void oracle_sample_nodes(u32 n, u64 instance_key, u32* out) { sample_nodes(n, instance_key, out); }

u32 oracle_remaining_node(u32 rank, const u32* existing, u32 count) { return remaining_node(rank, existing, count); }

void oracle_build_region(u32 n, const u32* tour, const u32* primary, u32 width, u64 key, u32 batch, u32 mode,
                         u32 region, u32* nodes) {
    build_region(make_rec(tour, n, 0), primary, width, key, batch, mode, region, nodes);
}

u32 oracle_baseline_action(const Baseline* policy, const Feedback* feedback, u32 batch, double uniform, u32 mask) {
    return baseline_action(*policy, *feedback, batch, uniform, mask);
}

double oracle_baseline_uniform(const Baseline* policy, u64 key, u32 batch) {
    return baseline_uniform(*policy, key, batch);
}

u32 oracle_factorial_mask(u32 variant, const Baseline* policy, const Feedback* feedback, u32 batch, double uniform,
                          u32 legal) {
    return factorial_mask(variant, *policy, *feedback, batch, uniform, legal);
}

void oracle_update_feedback(Feedback* feedback, int improved, double returned, double work, u32 ants) {
    update_feedback(*feedback, improved != 0, returned, work, ants);
}

void oracle_restart_feedback(Feedback* feedback) { restart_feedback(*feedback); }

int oracle_score_controller(const Controller* controller, const float* features /*[12][32]*/, u32 mask,
                            float* scores, u32* action) {
    return guarded([&] { *action = score_controller(*controller, features, mask, scores); });
}

namespace {
i64 tour_length(const std::int32_t* xy, const u32* t, u32 n) {
    i64 s = 0;
    for (u32 i = 0; i < n; ++i) s += euc2d(xy, t[i], t[i + 1 == n ? 0 : i + 1]);
    return s;
}
}  // namespace

// ReferenceArchive::update semantics through archive_select: the pool is
// [global, entries..., completed...] with slot 0 the global tour. Outputs
// the chosen tours [4][n] and their count.
int oracle_archive_update(u32 n, const std::int32_t* xy, const u32* global, u32 entries, const u32* entry_tours,
                          u32 completed, const u32* completed_tours, u32* out_tours, u32* out_count) {
    return guarded([&] {
        std::vector<Rec> pool;
        pool.push_back(make_rec(global, n, tour_length(xy, global, n)));
        for (u32 i = 0; i < entries; ++i) {
            const u32* t = entry_tours + static_cast<size_t>(i) * n;
            pool.push_back(make_rec(t, n, tour_length(xy, t, n)));
        }
        for (u32 i = 0; i < completed; ++i) {
            const u32* t = completed_tours + static_cast<size_t>(i) * n;
            pool.push_back(make_rec(t, n, tour_length(xy, t, n)));
        }
        const auto chosen = archive_select(pool, 0, pool[0].cost);
        for (u32 i = 0; i < chosen.size(); ++i)
            std::copy(pool[chosen[i]].order.begin(), pool[chosen[i]].order.end(), out_tours + static_cast<size_t>(i) * n);
        *out_count = static_cast<u32>(chosen.size());
    });
}

// Feature snapshot of one colony (build_control_features) from an explicit
// state: active (parent) tour, archive tours (slot 0 = global best), sparse
// trails [n][pw] with their default, bounds, feedback and progress.
int oracle_action_snapshot(u32 n, const std::int32_t* xy, u32 pw, const u32* primary, const double* local_scale,
                           double epsilon, const u32* samples, const u32* active, u32 archive_size,
                           const u32* archive_tours, const double* trails, double default_trail, double minimum,
                           double maximum, const Feedback* feedback, double elapsed, u64 key, u32 batch,
                           u32 experiment_mask, float* features, u32* regions, u32* alternative, u32* mask) {
    return guarded([&] {
        Problem p; p.n = n; p.xy = xy;
        p.s.primary_width = pw;
        p.primary.assign(primary, primary + static_cast<size_t>(n) * pw);
        p.local_scale.assign(local_scale, local_scale + n);
        p.scale_epsilon = epsilon;
        Control ctl;
        ctl.fb = *feedback;
        for (u32 i = 0; i < archive_size; ++i) {
            const u32* t = archive_tours + static_cast<size_t>(i) * n;
            ctl.archive.push_back(make_rec(t, n, tour_length(xy, t, n)));
        }
        ColonyRun r;
        r.parent.assign(active, active + n);
        r.parent_cost = tour_length(xy, active, n);
        r.global_cost = ctl.archive[0].cost;
        r.trails.assign(trails, trails + static_cast<size_t>(n) * pw);
        r.default_trail = default_trail; r.minimum = minimum; r.maximum = maximum;
        ctl.active = tour_ident(active, n);
        Decision d;
        control_features(p, r, ctl, samples, key, batch, elapsed, experiment_mask, d);
        std::copy(d.features, d.features + 12 * 32, features);
        std::memcpy(regions, d.regions, sizeof(d.regions));
        *alternative = d.alternative;
        *mask = d.mask;
    });
}

// Controlled FACO run (evaluate_program / population / controller /
// baseline / factorial evaluations, cuda/batch_engine.cu:376-460,492-1192).
// Colony c uses instance c % instances and decision maker c / instances.
// mode: 1 program, 2 controller, 3 baseline, 4 factorial (program + policy).
// Traces (any may be NULL), per batch b and colony c:
//   features [b][12][C][32], scores [b][C][32], masks/actions/alternatives
//   [b][C], regions [b][C][2][4][16], costs [b][C][3] (epoch, global, parent
//   after the batch).
int oracle_run_controlled(u32 n, u32 instances, const std::int32_t* xy_all, const u64* instance_keys,
                          u32 colonies, const u64* seeds, const Settings* settings, u64 evaluation_limit,
                          u32 mode, u32 deciders, const Program* programs, const Controller* controllers,
                          const Baseline* baseline, u32 variant, u32 experiment_mask, int force_collisions,
                          int threads,
                          u32* global_tours, std::int64_t* global_costs, u64* work, ControlOut* states,
                          float* trace_features, float* trace_scores, u32* trace_masks, std::int32_t* trace_actions,
                          u32* trace_alternatives, u32* trace_regions, std::int64_t* trace_costs,
                          double* elapsed_search_seconds, u32 local_search, u32 stop_batch, u32 fork_batch,
                          u64 fork_seed, std::int32_t fork_action,
                          const u32* graph_offsets /* [instances][n + 1] or null: Hard mode */,
                          const u32* graph_neighbors, u32 neighbor_stride,
                          const float* multipliers /* [instances][n * pw] or null */,
                          const double* progress /* [batches]: the progress feature of each batch, or null:
                                                    completed FE / evaluation_limit */,
                          u64* rejections /* [colonies] or null: Hard-mode 2-opt moves rejected */,
                          const u32* rows /* [instances][n (pw + bw + lw)] or null: a graph spec's own rows */,
                          u32 escape /* 0, or Escape mode on the graph: 1 gates on, 2 gates disabled */,
                          u64* escape_counters /* [colonies][11] or null */) {
    // stop_batch (0 = the whole budget) ends the run early with the progress
    // feature still measured against evaluation_limit. At batch fork_batch
    // (UINT32_MAX = none) a paired fork (reference cuda/batch_engine.cu:839-845,
    // 883-923) replaces the decision by fork_action, the decision itself
    // still computed under the colony's keys so far, and from that batch's
    // ants on every colony runs under mixed(key ^ mixed(fork_seed + golden)).
    return guarded([&] {
        if (!instances || !colonies || colonies % instances) throw std::invalid_argument("invalid panel");
        if (mode < 1 || mode > 4) throw std::invalid_argument("unknown control mode");
        if (!(experiment_mask & 0xffffu)) throw std::invalid_argument("experiment mask must keep a hold action");
        if (escape > 2 || (escape && (!graph_offsets || multipliers)))
            throw std::invalid_argument("Escape mode needs a graph and takes no learned prior");
        g_force_collisions = force_collisions != 0;
        struct Reset { ~Reset() { g_force_collisions = false; } } reset_flag;
        std::vector<Problem> problems(instances);
        std::vector<std::vector<u32>> samples(instances, std::vector<u32>(kSamples, 0));
        for (u32 g = 0; g < instances; ++g) {
            check_coordinates(n, xy_all + static_cast<size_t>(g) * 2 * n);
            problems[g].n = n; problems[g].xy = xy_all + static_cast<size_t>(g) * 2 * n;
            problems[g].s = normalized(*settings, n);
            prepare(problems[g]);
            if (rows) apply_rows(problems[g], rows + static_cast<size_t>(g) * row_cells(problems[g]));
            problems[g].or_opt = local_search != 0;
            problems[g].escape = escape;
            if (graph_offsets) {
                if (local_search) throw std::invalid_argument("Hard mode takes the plain 2-opt");
                problems[g].graph_offsets = graph_offsets + static_cast<size_t>(g) * (n + 1);
                problems[g].graph_neighbors = graph_neighbors + static_cast<size_t>(g) * neighbor_stride;
            }
            if (multipliers) {
                const size_t cells = static_cast<size_t>(n) * problems[g].s.primary_width;
                problems[g].prior.assign(multipliers + g * cells, multipliers + (g + 1) * cells);
            }
            sample_nodes(n, instance_keys[g], samples[g].data());
        }
        const Settings s = problems[0].s;
        if (evaluation_limit % s.ants) throw std::invalid_argument("evaluation limit must be whole ant batches");
        const u32 batches = stop_batch ? std::min(stop_batch, static_cast<u32>(evaluation_limit / s.ants))
                                       : static_cast<u32>(evaluation_limit / s.ants);
        std::vector<ColonyRun> runs(colonies);
        std::vector<Control> controls(colonies);
        std::vector<u64> keys(colonies);
        for (u32 c = 0; c < colonies; ++c) {
            const Problem& p = problems[c % instances];
            colony_reset(p, runs[c]);
            control_init(p, controls[c]);
            control_restart(p, runs[c], controls[c], 0, true);
            keys[c] = colony_key(instance_keys[c % instances], seeds[c]);
        }
        std::vector<std::vector<AntResult>> results(colonies, std::vector<AntResult>(s.ants));
        std::vector<Decision> decisions(colonies);
        std::vector<u32> targets(colonies);
#ifdef _OPENMP
        if (threads > 0) omp_set_num_threads(threads);
        const double started = omp_get_wtime();
#endif
        for (u32 batch = 0; batch < batches; ++batch) {
            const double elapsed = progress ? progress[batch]
                                            : static_cast<double>(static_cast<u64>(batch) * s.ants) /
                                                  static_cast<double>(evaluation_limit);
            for (u32 c = 0; c < colonies; ++c) {
                const Problem& p = problems[c % instances];
                Decision& d = decisions[c];
                control_features(p, runs[c], controls[c], samples[c % instances].data(), keys[c], batch, elapsed,
                                 experiment_mask, d);
                const u32 who = c / instances;
                if (mode == 3 || mode == 4) {
                    const double u = baseline_uniform(*baseline, keys[c], batch);
                    if (mode == 3) {
                        d.action = baseline_action(*baseline, controls[c].fb, batch, u, d.mask);
                        std::fill(d.scores, d.scores + 32, 0.0f);
                    } else {
                        d.mask = factorial_mask(variant, *baseline, controls[c].fb, batch, u, d.mask);
                        d.action = score_program(programs[who < deciders ? who : 0], d.features, d.mask, d.scores);
                    }
                } else if (mode == 1) {
                    d.action = score_program(programs[who < deciders ? who : 0], d.features, d.mask, d.scores);
                } else {
                    d.action = score_controller(controllers[who < deciders ? who : 0], d.features, d.mask, d.scores);
                }
                if (batch == fork_batch) {
                    d.action = static_cast<u32>(fork_action);
                    keys[c] = mixed(keys[c] ^ mixed(fork_seed + 0xa0761d6478bd642fULL));
                }
                if (d.action >= 32) throw std::runtime_error("no legal control action");
                targets[c] = 2u << (d.action % 4);
                if (d.action >= 16) control_restart(p, runs[c], controls[c], d.alternative, false);
            }
            const long long total = static_cast<long long>(colonies) * s.ants;
#pragma omp parallel for schedule(dynamic, 1)
            for (long long i = 0; i < total; ++i) {
                const u32 c = static_cast<u32>(i / s.ants), a = static_cast<u32>(i % s.ants);
                const Decision& d = decisions[c];
                results[c][a] = controlled_ant(problems[c % instances], runs[c], keys[c], batch, a, targets[c],
                                               d.regions[d.action / 16][(d.action / 4) % 4], d.region_count);
            }
#pragma omp parallel for schedule(dynamic, 1)
            for (long long c = 0; c < static_cast<long long>(colonies); ++c) {
                const Problem& p = problems[c % instances];
                colony_select_and_update(p, runs[c], results[c], keys[c], batch);
                control_update(p, runs[c], controls[c], results[c]);
            }
            for (u32 c = 0; c < colonies; ++c) {
                const Decision& d = decisions[c];
                const size_t bc = static_cast<size_t>(batch) * colonies + c;
                if (trace_features)
                    for (u32 f = 0; f < 12; ++f)
                        std::copy(d.features + f * 32, d.features + f * 32 + 32,
                                  trace_features + ((static_cast<size_t>(batch) * 12 + f) * colonies + c) * 32);
                if (trace_scores) std::copy(d.scores, d.scores + 32, trace_scores + bc * 32);
                if (trace_masks) trace_masks[bc] = d.mask;
                if (trace_actions) trace_actions[bc] = static_cast<std::int32_t>(d.action);
                if (trace_alternatives) trace_alternatives[bc] = d.alternative;
                if (trace_regions) std::memcpy(trace_regions + bc * 128, d.regions, sizeof(d.regions));
                if (trace_costs) {
                    trace_costs[3 * bc] = runs[c].epoch_cost;
                    trace_costs[3 * bc + 1] = runs[c].global_cost;
                    trace_costs[3 * bc + 2] = runs[c].parent_cost;
                }
            }
        }
#ifdef _OPENMP
        if (elapsed_search_seconds) *elapsed_search_seconds = omp_get_wtime() - started;
#else
        if (elapsed_search_seconds) *elapsed_search_seconds = 0;
#endif
        for (u32 c = 0; c < colonies; ++c) {
            std::copy(runs[c].global.begin(), runs[c].global.end(), global_tours + static_cast<size_t>(c) * n);
            global_costs[c] = runs[c].global_cost;
            work[2 * c] = runs[c].construction_steps;
            work[2 * c + 1] = runs[c].ls_evaluations;
            if (rejections) rejections[c] = runs[c].constraint_rejections;
            if (escape_counters) std::copy(runs[c].esc, runs[c].esc + kEscCounters, escape_counters + kEscCounters * c);
            if (states) {
                const Control& k = controls[c];
                states[c] = ControlOut{static_cast<u32>(k.archive.size()), 0, k.fb.stagnant, k.fb.epoch_batches,
                                       k.fb.restarts, k.fb.return_rate, k.fb.ls_work};
            }
        }
    });
}

// Explicit-state access to one batch's reduce/select + pheromone update
// (colony_select_and_update) and to one restart transaction
// (control_restart), so that their rules can be checked against hand-made
// cases. The instance is prepared as in a run (rows, heuristic); the colony
// state comes in and goes out as plain arrays:
//   costs [3] = epoch, global, parent cost; tours [3][n] = parent, epoch,
//   global tour; trails [n][pw]; default_trail; bounds [2] = min, max (out);
//   products [n][pw] (out).
int oracle_colony_select(u32 n, const std::int32_t* xy, const Settings* settings, u64 key, u32 batch, u32 ants,
                         const u32* ant_tours, const std::int64_t* ant_costs, std::int64_t* costs, u32* tours,
                         double* trails, double* default_trail, double* bounds, double* products,
                         u32* iteration_best, u32* source_is_epoch, double* source_uniform) {
    return guarded([&] {
        check_coordinates(n, xy);
        if (!ants) throw std::invalid_argument("a batch needs at least one ant");
        Problem p;
        p.n = n; p.xy = xy; p.s = normalized(*settings, n);
        prepare(p);
        ColonyRun r;
        colony_reset(p, r);
        const size_t cells = static_cast<size_t>(n) * p.s.primary_width;
        r.epoch_cost = costs[0]; r.global_cost = costs[1]; r.parent_cost = costs[2];
        r.parent.assign(tours, tours + n);
        r.epoch.assign(tours + n, tours + 2 * n);
        r.global.assign(tours + 2 * n, tours + 3 * n);
        r.trails.assign(trails, trails + cells);
        r.default_trail = *default_trail;
        std::vector<AntResult> results(ants);
        for (u32 a = 0; a < ants; ++a) {
            results[a].tour.assign(ant_tours + static_cast<size_t>(a) * n, ant_tours + static_cast<size_t>(a + 1) * n);
            results[a].final_cost = ant_costs[a];
        }
        colony_select_and_update(p, r, results, key, batch);
        costs[0] = r.epoch_cost; costs[1] = r.global_cost; costs[2] = r.parent_cost;
        std::copy(r.parent.begin(), r.parent.end(), tours);
        std::copy(r.epoch.begin(), r.epoch.end(), tours + n);
        std::copy(r.global.begin(), r.global.end(), tours + 2 * n);
        std::copy(r.trails.begin(), r.trails.end(), trails);
        std::copy(r.products.begin(), r.products.end(), products);
        *default_trail = r.default_trail;
        bounds[0] = r.minimum; bounds[1] = r.maximum;
        *iteration_best = r.iteration_best;
        *source_is_epoch = r.source_is_epoch;
        *source_uniform = r.source_uniform;
    });
}

// archive [count][n] tours with costs; identities out [2][2] = active and
// epoch identity after the transaction.
int oracle_control_restart(u32 n, const std::int32_t* xy, const Settings* settings, u32 count, const u32* archive,
                           const std::int64_t* archive_costs, u32 alternative, int initializing, std::int64_t* costs,
                           u32* tours, double* trails, double* default_trail, double* bounds, double* products,
                           Feedback* feedback, u64* identities) {
    return guarded([&] {
        check_coordinates(n, xy);
        if (!count || count > kArchive || alternative >= count) throw std::invalid_argument("archive slot out of range");
        Problem p;
        p.n = n; p.xy = xy; p.s = normalized(*settings, n);
        prepare(p);
        ColonyRun r;
        colony_reset(p, r);
        const size_t cells = static_cast<size_t>(n) * p.s.primary_width;
        r.epoch_cost = costs[0]; r.global_cost = costs[1]; r.parent_cost = costs[2];
        r.parent.assign(tours, tours + n);
        r.epoch.assign(tours + n, tours + 2 * n);
        r.global.assign(tours + 2 * n, tours + 3 * n);
        r.trails.assign(trails, trails + cells);
        r.default_trail = *default_trail;
        Control ctl;
        ctl.fb = *feedback;
        for (u32 i = 0; i < count; ++i)
            ctl.archive.push_back(make_rec(archive + static_cast<size_t>(i) * n, n, archive_costs[i]));
        control_restart(p, r, ctl, alternative, initializing != 0);
        costs[0] = r.epoch_cost; costs[1] = r.global_cost; costs[2] = r.parent_cost;
        std::copy(r.parent.begin(), r.parent.end(), tours);
        std::copy(r.epoch.begin(), r.epoch.end(), tours + n);
        std::copy(r.global.begin(), r.global.end(), tours + 2 * n);
        std::copy(r.trails.begin(), r.trails.end(), trails);
        std::copy(r.products.begin(), r.products.end(), products);
        *default_trail = r.default_trail;
        bounds[0] = r.minimum; bounds[1] = r.maximum;
        *feedback = ctl.fb;
        identities[0] = ctl.active.first; identities[1] = ctl.active.second;
        identities[2] = ctl.epoch_id.first; identities[3] = ctl.epoch_id.second;
    });
}

int oracle_max_threads() {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

}  // extern "C"

// ===========================================================================
// Learned edge prior: candidate-edge features (this engine's extension; the
// reference has no learned prior). FP32 arithmetic in the order written here
// is the definition; the HIP kernel computes the same operations.
namespace {
float prior_clip(float v) { return v < 0.0f ? 0.0f : (v > 4.0f ? 4.0f : v); }
}  // namespace

extern "C" {
// features [n * pw][8] for edge (a, j) = (a, primary[a][j]):
// 0 d/scale(a), 1 j/(pw-1), 2 d/d(a, nearest), 3 mutual (a in b's row),
// 4 rank of a in b's row / pw (1 when absent), 5 d/scale(b), 6 d/mean scale,
// 7 shared row entries / pw; the distance ratios are clipped to [0, 4].
int oracle_prior_features(u32 n, const std::int32_t* xy, u32 pw, const u32* primary, const double* local_scale,
                          double epsilon, float* out) {
    return guarded([&] {
        check_coordinates(n, xy);
        if (pw < 2 || pw > n - 1) throw std::invalid_argument("primary width out of range");
        double mean = 0;
        for (u32 a = 0; a < n; ++a) mean += local_scale[a];
        const float mean_scale = static_cast<float>(std::max(mean / n, epsilon));
        for (u32 a = 0; a < n; ++a) {
            const u32* row = primary + static_cast<size_t>(a) * pw;
            const float sa = static_cast<float>(std::max(local_scale[a], epsilon));
            const float d0 = static_cast<float>(std::max<i64>(euc2d(xy, a, row[0]), 1));
            for (u32 j = 0; j < pw; ++j) {
                const u32 b = row[j];
                const u32* brow = primary + static_cast<size_t>(b) * pw;
                const float sb = static_cast<float>(std::max(local_scale[b], epsilon));
                const float d = static_cast<float>(euc2d(xy, a, b));
                u32 rank = pw, shared = 0;
                for (u32 k = 0; k < pw; ++k) {
                    if (brow[k] == a && rank == pw) rank = k;
                    for (u32 m = 0; m < pw; ++m) shared += brow[k] == row[m] ? 1u : 0u;
                }
                float* f = out + (static_cast<size_t>(a) * pw + j) * 8;
                f[0] = prior_clip(d / sa);
                f[1] = static_cast<float>(j) / static_cast<float>(pw - 1);
                f[2] = prior_clip(d / d0);
                f[3] = rank < pw ? 1.0f : 0.0f;
                f[4] = static_cast<float>(rank) / static_cast<float>(pw);
                f[5] = prior_clip(d / sb);
                f[6] = prior_clip(d / mean_scale);
                f[7] = static_cast<float>(shared) / static_cast<float>(pw);
            }
        }
    });
}
}  // extern "C"

// ---------------------------------------------------------------------------

extern "C" {

// Escape slot view of one row, optionally sorted as an LS row. `script`:
// 0 = the escape stream of (key, batch, ant, event, ls); else scripted draws
// with draw() = script - 1 and below(count) = count - 1. `distance_of` [n]
// (may be null: no sort) gives the sort key of each node. out_counts = valid,
// replacements.
int oracle_escape_slot_view(const u32* original, u32 width, const u32* backup, u32 bw, u32 n, u32 from, u32 script,
                            u64 key, u32 batch, u32 ant, u32 event, u32 ls, const double* distance_of,
                            u32* out_members, std::uint8_t* out_replaced, u32* out_counts) {
    return guarded([&] {
        EscRow row;
        if (script) { ScriptDraws d{script - 1}; row = slot_view(original, width, backup, bw, n, from, d); }
        else { EscStream d(key, batch, ant, event, ls != 0); row = slot_view(original, width, backup, bw, n, from, d); }
        if (distance_of) sort_row(row, distance_of);
        for (u32 j = 0; j < width; ++j) { out_members[j] = row.members[j]; out_replaced[j] = row.replaced[j]; }
        out_counts[0] = row.valid; out_counts[1] = row.replacements;
    });
}

// `count` words of the escape stream of (key, batch, ant, event, ls).
void oracle_escape_draws(u64 key, u32 batch, u32 ant, u32 event, u32 ls, u32 count, u32* out) {
    EscStream d(key, batch, ant, event, ls != 0);
    for (u32 i = 0; i < count; ++i) out[i] = d.draw();
}

// One escape transaction: prepare (removed/added [count][2]) against the
// CSR graph and the cache (edges [size][2], epoch), then commit into a copy
// of the cache. out_proposal [11] (see store_proposal); out_cache_after
// [64][2] + out_size_after; returns the commit result in *committed.
int oracle_escape_transaction(u32 count, const u32* removed, const u32* added, const u32* offsets, const u32* neighbors,
                              u32 dimension, const u32* cache_edges, u32 cache_size, u64 cache_epoch, u32 permit,
                              u32* out_proposal, u32* committed, u32* out_cache_after, u32* out_size_after) {
    return guarded([&] {
        if (count < 1 || count > 3) throw std::invalid_argument("escape transactions hold 1 to 3 edges");
        EscEdge r[3], a[3];
        for (u32 i = 0; i < count; ++i) { r[i] = {removed[2 * i], removed[2 * i + 1]}; a[i] = {added[2 * i], added[2 * i + 1]}; }
        EscCache c;
        load_cache(c, cache_edges, cache_size, cache_epoch);
        const EscProposal p = prepare_edges(count, r, a, offsets, neighbors, dimension, c, permit != 0);
        store_proposal(p, out_proposal);
        *committed = commit_edges(c, p) ? 1u : 0u;
        for (u32 i = 0; i < kEscCapacity; ++i) { out_cache_after[2 * i] = c.edges[i].a; out_cache_after[2 * i + 1] = c.edges[i].b; }
        *out_size_after = c.size;
    });
}

// commit_escape_edges of a stored proposal (out of store_proposal) into a
// copy of a cache: the commit of a proposal prepared against another cache
// state (expired by a move or by a reset).
int oracle_escape_commit(const u32* proposal, const u32* cache_edges, u32 cache_size, u64 cache_epoch, u32* committed,
                         u32* out_cache_after, u32* out_size_after) {
    return guarded([&] {
        EscCache c;
        load_cache(c, cache_edges, cache_size, cache_epoch);
        EscProposal p;
        p.status = proposal[0]; p.count = proposal[1]; p.expected_size = proposal[2];
        p.expected_epoch = static_cast<u64>(proposal[3]) | (static_cast<u64>(proposal[4]) << 32);
        if (p.count > 3) throw std::invalid_argument("a proposal holds at most 3 edges");
        for (u32 i = 0; i < 3; ++i) p.edges[i] = {proposal[5 + 2 * i], proposal[6 + 2 * i]};
        *committed = commit_edges(c, p) ? 1u : 0u;
        for (u32 i = 0; i < kEscCapacity; ++i) { out_cache_after[2 * i] = c.edges[i].a; out_cache_after[2 * i + 1] = c.edges[i].b; }
        *out_size_after = c.size;
    });
}

// Hard mode (edge_constraints.hpp:47-90): a move is allowed when every
// added edge either cancels a removed one (multiset: an edge removed and
// added back is no new edge) or is an edge of the graph.
int oracle_hard_edges_allowed(u32 count, const u32* removed, const u32* added, const u32* offsets,
                              const u32* neighbors, u32 n) {
    if (count > 3) return -1;
    bool consumed[3] = {false, false, false};
    for (u32 i = 0; i < count; ++i) {
        const EscEdge e{added[2 * i], added[2 * i + 1]};
        bool existing = false;
        for (u32 j = 0; j < count; ++j)
            if (!consumed[j] && same_edge(e, EscEdge{removed[2 * j], removed[2 * j + 1]})) {
                consumed[j] = true; existing = true; break;
            }
        if (!existing && !graph_has(offsets, neighbors, n, e.a, e.b)) return 0;
    }
    return 1;
}

// relocation_allowed: `node` moved to follow `target` on an array tour.
int oracle_hard_relocation_allowed(const u32* tour, const u32* positions, u32 n, u32 target, u32 node,
                                   const u32* offsets, const u32* neighbors) {
    if (target >= n || node >= n || target == node) return 0;
    const u32 target_after = tour[(positions[target] + 1) % n];
    if (target_after == node) return 1;
    const u32 before = tour[(positions[node] + n - 1) % n], after = tour[(positions[node] + 1) % n];
    const u32 removed[6] = {before, node, node, after, target, target_after};
    const u32 added[6] = {before, after, target, node, node, target_after};
    return oracle_hard_edges_allowed(3, removed, added, offsets, neighbors, n);
}

// two_opt_allowed: remove (a, a_neighbor), (b, b_neighbor); add (a, b), (a_neighbor, b_neighbor).
int oracle_hard_two_opt_allowed(u32 a, u32 a_neighbor, u32 b, u32 b_neighbor, const u32* offsets,
                                const u32* neighbors, u32 n) {
    const u32 removed[4] = {a, a_neighbor, b, b_neighbor};
    const u32 added[4] = {a, b, a_neighbor, b_neighbor};
    return oracle_hard_edges_allowed(2, removed, added, offsets, neighbors, n);
}

// prepare_escape_relocation (escape_edges.hpp:101-115) on an array tour.
int oracle_escape_relocation(const u32* tour, const u32* positions, u32 n, u32 target, u32 node, const u32* offsets,
                             const u32* neighbors, const u32* cache_edges, u32 cache_size, u64 cache_epoch, u32 permit,
                             u32* out_proposal) {
    return guarded([&] {
        EscCache c;
        load_cache(c, cache_edges, cache_size, cache_epoch);
        EscProposal p;
        p.expected_size = c.size; p.expected_epoch = c.epoch;
        if (target >= n || node >= n || target == node) {
            p.status = kInvalidEdge;
        } else {
            const u32 target_after = tour[(positions[target] + 1) % n];
            if (target_after != node) {
                const u32 before = tour[(positions[node] + n - 1) % n], after = tour[(positions[node] + 1) % n];
                const EscEdge removed[3] = {{before, node}, {node, after}, {target, target_after}};
                const EscEdge added[3] = {{before, after}, {target, node}, {node, target_after}};
                p = prepare_edges(3, removed, added, offsets, neighbors, n, c, permit != 0);
            }
        }
        store_proposal(p, out_proposal);
    });
}

}  // extern "C"
