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
struct Stream {
    u32 ctr[4];
    u32 key[2];
    u32 block[4];
    int index = 4;
    Stream(u64 seed, u32 batch, u32 ant, u32 step) {
        ctr[0] = step; ctr[1] = 0; ctr[2] = ant; ctr[3] = batch;
        key[0] = static_cast<u32>(seed); key[1] = static_cast<u32>(seed >> 32);
    }
    u32 next() {
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
    std::vector<u32> cheap, initial;
    i64 cheap_cost = 0, initial_cost = 0;
    i64 d(u32 a, u32 b) const { return euc2d(xy, a, b); }
};

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

struct LsStats {
    u64 processed = 0, checks = 0, evaluations = 0, accepted = 0, reactivations = 0;
    u32 limit_reached = 0;
};

// Checklist 2-opt: restates CpuTour::checklist_two_opt
// (cpp/src/faco_cpu.cpp:148-217). `checklist` is the whole FIFO history; the
// unprocessed part starts at `next`. `queued[v]` marks membership of the
// unprocessed part (the reference's std::find over [next, end)).
LsStats two_opt(Tour& t, const u32* ls, u32 lw, std::vector<u32>& checklist,
                std::vector<std::uint8_t>& queued, u64 limit) {
    LsStats st;
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
        bool interrupted = false;
        for (int kind = 0; kind < 2 && !interrupted; ++kind) {
            const i64 current = kind == 0 ? to_next : to_prev;
            for (u32 j = 0; j < lw; ++j) {
                const u32 b = ls[static_cast<size_t>(a) * lw + j];
                ++st.checks;
                const i64 ab = t.p->d(a, b);
                if (!(current > ab)) break;
                if (st.evaluations == limit) { interrupted = true; break; }
                ++st.evaluations;
                const u32 nb = kind == 0 ? t.succ(b) : t.pred(b);
                const u32 closing = kind == 0 ? a_next : a_prev;
                const i64 gain = (nb == a || b == closing) ? 0
                    : current + t.p->d(b, nb) - ab - t.p->d(closing, nb);
                if (gain > best) {
                    best = gain;
                    if (kind == 0) { move[0] = a_next; move[1] = nb; move[2] = a; move[3] = b; }
                    else { move[0] = a; move[1] = b; move[2] = a_prev; move[3] = nb; }
                }
            }
        }
        if (interrupted) { st.limit_reached = 1; break; }
        if (best > 0) {
            t.flip(move[0], move[1]);
            for (u32 v : move) {
                if (!queued[v]) { queued[v] = 1; checklist.push_back(v); ++st.reactivations; }
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
// 3 global fallback, 4 exhausted.
u32 select_next(const Problem& p, u32 current, const double* products,
                const std::vector<std::uint8_t>& visited, double u, int* stage) {
    const u32 pw = p.s.primary_width, bw = p.s.backup_width, n = p.n;
    const u32* primary = p.primary.data() + static_cast<size_t>(current) * pw;
    double sum = 0;
    u32 avail[64];
    double prefix[64];
    u32 count = 0;
    for (u32 j = 0; j < pw; ++j) {
        const u32 v = primary[j];
        if (v < n && !visited[v]) { sum += products[j]; avail[count] = v; prefix[count] = sum; ++count; }
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
        if (backup[j] < n && !visited[backup[j]]) { *stage = 2; return backup[j]; }
    }
    u32 chosen = current;
    i64 best = std::numeric_limits<i64>::max();
    for (u32 v = 0; v < n; ++v) {
        if (visited[v]) continue;
        const i64 value = p.d(current, v);
        if (value < best) { best = value; chosen = v; }
    }
    *stage = chosen == current ? 4 : 3;
    return chosen;
}

struct ConstructionStats { u32 mne = 0, steps = 0, relocations = 0, exhausted = 0; };

struct AntResult {
    std::vector<u32> construction_tour, tour, positions, checklist;
    i64 construction_cost = 0, final_cost = 0;
    ConstructionStats c;
    LsStats ls;
    u32 start = 0;
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
        const u32 selected = choose(current, step, visited);
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
    r.ls = two_opt(work, p.ls.data(), p.s.ls_width, checklist, queued, limit);
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
    two_opt(t, p.ls.data(), c.ls_width, checklist, queued, c.initial_ls_evaluation_limit);
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
    u64 construction_steps = 0, ls_evaluations = 0;
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
        r.eta[i] = heuristic(p.d(a, p.primary[i]), p.s.beta);
        r.products[i] = r.trails[i] * r.eta[i];
    }
    r.construction_steps = r.ls_evaluations = 0;
    r.source_uniform = 0; r.source_is_epoch = 0; r.iteration_best = 0;
}

// reduce_and_select + update_pheromone for one batch given every ant's result.
void colony_select_and_update(const Problem& p, ColonyRun& r, const std::vector<AntResult>& ants,
                              u64 key, u32 batch) {
    const u32 n = p.n, pw = p.s.primary_width;
    u32 best = 0;
    for (u32 a = 0; a < ants.size(); ++a) {
        r.construction_steps += ants[a].c.steps;
        r.ls_evaluations += ants[a].ls.evaluations;
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

AntResult stochastic_ant(const Problem& p, const ColonyRun& r, u64 key, u32 batch, u32 ant, u32 target) {
    const u32 pw = p.s.primary_width;
    const u32 start = start_node(key, batch, ant, p.n);
    return run_ant(p, r.parent.data(), start, target, p.s.ls_evaluation_limit,
        [&](u32 current, u32 step, const std::vector<std::uint8_t>& visited) {
            Stream stream(key, batch, ant, step);
            const double u = stream.uniform53();
            int stage = 0;
            return select_next(p, current, r.products.data() + static_cast<size_t>(current) * pw, visited, u, &stage);
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
                        std::int64_t* costs /* construction, final */, u64* counters) {
    return guarded([&] {
        check_coordinates(n, xy);
        Problem p; p.n = n; p.xy = xy; p.s.ls_width = lw;
        p.ls.assign(ls, ls + static_cast<size_t>(n) * lw);
        u32 cursor = 1;
        auto r = run_ant(p, parent, visit[0], mne_target, limit,
            [&](u32, u32, const std::vector<std::uint8_t>&) { return visit[cursor++]; });
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
                   u32 checklist_size, u64 limit, std::int64_t* cost, u64* evaluations) {
    return guarded([&] {
        check_coordinates(n, xy);
        Problem p; p.n = n; p.xy = xy;
        Tour t; t.assign(tour, n, &p);
        std::vector<u32> checklist(checklist_in, checklist_in + checklist_size);
        std::vector<std::uint8_t> queued(n, 0);
        for (u32 v : checklist) queued[v] = 1;
        const LsStats st = two_opt(t, ls, lw, checklist, queued, limit);
        std::copy(t.order.begin(), t.order.end(), tour);
        *cost = t.recomputed();
        *evaluations = st.evaluations;
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

// Full fixed-iteration run of `colonies` independent colonies (instance i %
// instances, seed seeds[i]) for `batches` batches. Outputs per colony the
// global tour, global cost and work counters. Optional per-batch trace for
// colony 0: parent after each batch [batches][n] and trails after each batch
// [batches][n*pw]. Threads: OpenMP over (colony, ant) within a batch.
int oracle_run_faco(u32 n, u32 instances, const std::int32_t* xy_all, const u64* instance_keys,
                    u32 colonies, const u64* seeds, const Settings* settings, u32 batches,
                    u32 mne_target, int threads,
                    u32* global_tours, std::int64_t* global_costs, u64* work /* [colonies][2] */,
                    u32* trace_parent, double* trace_trails, std::int64_t* trace_costs /*[batches][3]*/,
                    double* elapsed_search_seconds) {
    return guarded([&] {
        if (!instances || !colonies) throw std::invalid_argument("empty panel");
        std::vector<Problem> problems(instances);
        for (u32 g = 0; g < instances; ++g) {
            check_coordinates(n, xy_all + static_cast<size_t>(g) * 2 * n);
            problems[g].n = n; problems[g].xy = xy_all + static_cast<size_t>(g) * 2 * n;
            problems[g].s = normalized(*settings, n);
            prepare(problems[g]);
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
        const double started = omp_get_wtime();
#endif
        for (u32 batch = 0; batch < batches; ++batch) {
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
        }
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
