// Host problem preparation (see host_prepare.h). Rows are built from one
// O(n) distance row per city; the candidate order is (distance, city id),
// encoded as one 64-bit key so a plain sort gives the reference's
// std::pair ordering. Cities are processed in parallel host threads.
#include "host_prepare.h"
#include "faco_common.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <limits>
#include <stdexcept>
#include <thread>

namespace faco {
namespace {
using Clock = std::chrono::steady_clock;
double since(Clock::time_point t) { return std::chrono::duration<double>(Clock::now() - t).count(); }
void require(bool condition, const char* message) {
    if (!condition) throw std::invalid_argument(message);
}

// Array tour with inverse positions for the initial checklist 2-opt
// (reference CpuTour::checklist_two_opt, cpp/src/faco_cpu.cpp:148-217).
struct HostTour {
    const Prepared& p;
    std::vector<uint32_t> order, pos;
    uint32_t n;
    HostTour(const Prepared& problem, const std::vector<uint32_t>& tour)
        : p(problem), order(tour), pos(tour.size()), n(static_cast<uint32_t>(tour.size())) {
        for (uint32_t i = 0; i < n; ++i) pos[order[i]] = i;
    }
    int64_t d(uint32_t a, uint32_t b) const { return euc2d(p.packed[a], p.packed[b]); }
    uint32_t succ(uint32_t v) const { return order[pos[v] + 1 == n ? 0 : pos[v] + 1]; }
    uint32_t pred(uint32_t v) const { return order[pos[v] == 0 ? n - 1 : pos[v] - 1]; }
    // Reverse the shorter side between the two positions; with first == 0 the
    // straight side (the array layout of the reference's flip).
    void flip(uint32_t u, uint32_t v) {
        uint32_t first = pos[u], last = pos[v];
        if (first > last) std::swap(first, last);
        const uint32_t length = last - first;
        const bool straight = length <= n - length || first == 0;
        const uint32_t count = straight ? length : n - length, start = straight ? first : last;
        for (uint32_t k = 0; k < count / 2; ++k) {
            uint32_t l = start + k; if (l >= n) l -= n;
            uint32_t r = start + count - 1 - k; if (r >= n) r -= n;
            std::swap(order[l], order[r]);
            pos[order[l]] = l; pos[order[r]] = r;
        }
    }
    void two_opt(uint64_t limit) {
        const uint32_t lw = p.s_lw;
        std::vector<uint32_t> queue(order);  // every city, in tour order
        std::vector<uint8_t> queued(n, 1);
        size_t head = 0;
        uint64_t evaluations = 0, accepted = 0;
        while (head < queue.size() && accepted < n) {
            if (evaluations >= limit) break;
            const uint32_t a = queue[head++];
            queued[a] = 0;
            const uint32_t a_next = succ(a), a_prev = pred(a);
            const int64_t edge[2] = {d(a, a_next), d(a_prev, a)};
            int64_t best = -1;
            uint32_t move[4] = {0, 0, 0, 0};
            bool interrupted = false;
            for (int kind = 0; kind < 2 && !interrupted; ++kind) {
                for (uint32_t j = 0; j < lw; ++j) {
                    const uint32_t b = p.ls[static_cast<size_t>(a) * lw + j];
                    const int64_t ab = d(a, b);
                    if (!(edge[kind] > ab)) break;
                    if (evaluations == limit) { interrupted = true; break; }
                    ++evaluations;
                    const uint32_t nb = kind == 0 ? succ(b) : pred(b);
                    const uint32_t closing = kind == 0 ? a_next : a_prev;
                    const int64_t gain = (nb == a || b == closing) ? 0 : edge[kind] + d(b, nb) - ab - d(closing, nb);
                    if (gain > best) {
                        best = gain;
                        if (kind == 0) { move[0] = a_next; move[1] = nb; move[2] = a; move[3] = b; }
                        else { move[0] = a; move[1] = b; move[2] = a_prev; move[3] = nb; }
                    }
                }
            }
            if (interrupted) break;
            if (best > 0) {
                flip(move[0], move[1]);
                for (uint32_t v : move) if (!queued[v]) { queued[v] = 1; queue.push_back(v); }
                ++accepted;
            }
        }
    }
};

}  // namespace

faco_settings normalized_settings(faco_settings c, uint32_t n) {
    if (c.ants == 0) {
        uint32_t groups = 1;
        while (static_cast<uint64_t>(256) * groups * groups < n) ++groups;
        c.ants = 64 * groups;
    }
    require(n >= 3 && n <= 10000 && c.ants > 0 && c.ants <= 4096 && c.primary_width >= 2 && c.ls_width > 0 &&
            std::isfinite(c.beta) && c.beta > 0 && c.retention >= 0 && c.retention < 1 && c.p_best > 0 &&
            c.p_best < 1 && c.epoch_source_probability >= 0 && c.epoch_source_probability <= 1,
            "invalid FACO settings or dimension");
    c.primary_width = std::min(c.primary_width, n - 1);
    c.backup_width = std::min(c.backup_width, n - 1 - c.primary_width);
    c.ls_width = std::min(c.ls_width, n - 1);
    return c;
}

bool is_permutation(const uint32_t* tour, uint32_t n) {
    std::vector<uint8_t> seen(n, 0);
    for (uint32_t i = 0; i < n; ++i) {
        if (tour[i] >= n || seen[tour[i]]) return false;
        seen[tour[i]] = 1;
    }
    return true;
}

int64_t tour_cost(const Prepared& p, const uint32_t* tour) {
    int64_t sum = 0;
    uint32_t prev = tour[p.n - 1];
    for (uint32_t i = 0; i < p.n; ++i) { sum += euc2d(p.packed[prev], p.packed[tour[i]]); prev = tour[i]; }
    return sum;
}

uint32_t candidate_count(const faco_settings& c, uint32_t n) {
    return std::max({c.primary_width + c.backup_width, c.ls_width, std::min<uint32_t>(8, n - 1)});
}

Prepared prepare_begin(const std::vector<int32_t>& xy, const faco_settings& c) {
    const auto started = Clock::now();
    Prepared p;
    const uint32_t n = static_cast<uint32_t>(xy.size() / 2);
    p.n = n;
    p.xy = xy;
    p.packed.resize(n);
    for (uint32_t i = 0; i < n; ++i) p.packed[i] = pack_xy(static_cast<uint32_t>(xy[2 * i]), static_cast<uint32_t>(xy[2 * i + 1]));
    p.s_lw = c.ls_width;
    p.cheap.resize(n);
    for (uint32_t i = 0; i < n; ++i) p.cheap[i] = i;
    p.cheap_cost = tour_cost(p, p.cheap.data());
    // scale floor: 1e-12 times the larger coordinate range, at least 1
    // (cpp/src/prepared_problem.cpp:68-73)
    double min_x = xy[0], max_x = xy[0], min_y = xy[1], max_y = xy[1];
    for (uint32_t i = 0; i < n; ++i) {
        min_x = std::min<double>(min_x, xy[2 * i]); max_x = std::max<double>(max_x, xy[2 * i]);
        min_y = std::min<double>(min_y, xy[2 * i + 1]); max_y = std::max<double>(max_y, xy[2 * i + 1]);
    }
    p.scale_epsilon = 1e-12 * std::max({max_x - min_x, max_y - min_y, 1.0});
    p.cheap_seconds = since(started);
    return p;
}

void candidate_keys(const Prepared& p, uint32_t needed, uint32_t begin, uint32_t end, uint32_t* keys) {
    std::vector<uint32_t> row(p.n - 1);
    for (uint32_t a = begin; a < end; ++a) {
        uint32_t k = 0;
        for (uint32_t b = 0; b < p.n; ++b)
            if (b != a) row[k++] = (static_cast<uint32_t>(euc2d(p.packed[a], p.packed[b])) << 16) | b;
        std::partial_sort(row.begin(), row.begin() + needed, row.end());
        std::copy(row.begin(), row.begin() + needed, keys + static_cast<size_t>(a) * needed);
    }
}

void prepare_rows(Prepared& p, const faco_settings& c, uint32_t K, uint32_t nbr_stride, const uint32_t* keys,
                  uint32_t needed) {
    const uint32_t n = p.n, pw = c.primary_width, bw = c.backup_width, lw = c.ls_width;
    const uint32_t scale_width = std::min<uint32_t>(8, n - 1);
    p.primary.assign(static_cast<size_t>(n) * pw, 0);
    p.backup.assign(static_cast<size_t>(n) * bw, 0);
    p.ls.assign(static_cast<size_t>(n) * lw, 0);
    p.local_scale.assign(n, 0);
    p.eta.assign(static_cast<size_t>(n) * pw, 0);
    for (uint32_t a = 0; a < n; ++a) {
        const uint32_t* row = keys + static_cast<size_t>(a) * needed;
        double scale = 0;
        for (uint32_t j = 0; j < scale_width; ++j) scale += static_cast<double>(row[j] >> 16) / scale_width;
        p.local_scale[a] = scale;
        for (uint32_t j = 0; j < pw; ++j) {
            p.primary[static_cast<size_t>(a) * pw + j] = row[j] & 0xffffu;
            const double d = static_cast<double>(row[j] >> 16);
            p.eta[static_cast<size_t>(a) * pw + j] = d > 0 ? (c.beta == 1.0 ? 1.0 / d : 1.0 / std::pow(d, c.beta)) : 1.0;
        }
        for (uint32_t j = 0; j < bw; ++j) p.backup[static_cast<size_t>(a) * bw + j] = row[pw + j] & 0xffffu;
        for (uint32_t j = 0; j < lw; ++j) p.ls[static_cast<size_t>(a) * lw + j] = row[j] & 0xffffu;
    }
    // rows are prefixes of one (distance, id) order: primary = first pw,
    // LS = first lw, so one row of K = max(pw, lw) serves both
    p.nbr.assign(nbr_stride, 0);
    p.nbr_dist.assign(nbr_stride, 0);
    for (uint32_t a = 0; a < n; ++a)
        for (uint32_t j = 0; j < K; ++j) {
            const uint32_t key = keys[static_cast<size_t>(a) * needed + j];
            p.nbr[static_cast<size_t>(a) * K + j] = static_cast<uint16_t>(key & 0xffffu);
            p.nbr_dist[static_cast<size_t>(a) * K + j] = static_cast<uint16_t>(key >> 16);
        }
    p.backup16.assign(p.backup.begin(), p.backup.end());
}

std::vector<uint32_t> nearest_neighbour_tour(const Prepared& p) {
    const uint32_t n = p.n;
    std::vector<uint32_t> nn;
    nn.reserve(n);
    std::vector<uint8_t> seen(n, 0);
    uint32_t current = 0;
    for (uint32_t i = 0; i < n; ++i) {
        nn.push_back(current);
        seen[current] = 1;
        uint32_t next = n;
        int64_t best = std::numeric_limits<int64_t>::max();
        for (uint32_t b = 0; b < n; ++b) {
            if (seen[b]) continue;
            const int64_t v = euc2d(p.packed[current], p.packed[b]);
            if (v < best) { best = v; next = b; }
        }
        current = next;
    }
    return nn;
}

void prepare_finish(Prepared& p, std::vector<uint32_t> tour) {
    p.initial = std::move(tour);
    p.initial_cost = tour_cost(p, p.initial.data());
    if (p.cheap_cost <= p.initial_cost) { p.initial = p.cheap; p.initial_cost = p.cheap_cost; }
}

Prepared prepare_problem(const std::vector<int32_t>& xy, const faco_settings& c, uint32_t K, uint32_t nbr_stride) {
    const auto started = Clock::now();
    Prepared p = prepare_begin(xy, c);
    const uint32_t n = p.n, needed = candidate_count(c, n);
    std::vector<uint32_t> keys(static_cast<size_t>(n) * needed);
    const uint32_t threads = std::max(1u, std::min<uint32_t>(std::thread::hardware_concurrency(), n / 64 + 1));
    if (threads == 1) candidate_keys(p, needed, 0, n, keys.data());
    else {
        std::vector<std::thread> pool;
        for (uint32_t t = 0; t < threads; ++t)
            pool.emplace_back(candidate_keys, std::cref(p), needed,
                              static_cast<uint32_t>(static_cast<uint64_t>(n) * t / threads),
                              static_cast<uint32_t>(static_cast<uint64_t>(n) * (t + 1) / threads), keys.data());
        for (auto& th : pool) th.join();
    }
    prepare_rows(p, c, K, nbr_stride, keys.data(), needed);
    const auto t_candidates = Clock::now();
    p.profile[0] = std::chrono::duration<double>(t_candidates - started).count();
    const std::vector<uint32_t> nn = nearest_neighbour_tour(p);
    const auto t_nn = Clock::now();
    p.profile[1] = std::chrono::duration<double>(t_nn - t_candidates).count();
    HostTour tour(p, nn);
    tour.two_opt(c.initial_ls_evaluation_limit);
    const auto t_ls = Clock::now();
    p.profile[2] = std::chrono::duration<double>(t_ls - t_nn).count();
    prepare_finish(p, tour.order);
    p.profile[3] = since(t_ls);
    p.preparation_seconds = since(started);
    return p;
}

// sample_nodes (control_ops.hpp:104-108): min(n, 64) distinct cities from
// a SplitMix64 stream keyed by the instance; the i-th draw is a rank among
// the cities not drawn yet.
std::vector<uint32_t> sample_nodes(uint32_t n, uint64_t instance_key) {
    const uint32_t count = std::min<uint32_t>(n, kSamples);
    std::vector<uint32_t> out(kSamples, 0), sorted;
    uint64_t key = mixed(instance_key ^ 0x6a09e667f3bcc909ULL), counter = 0;
    for (uint32_t i = 0; i < count; ++i) {
        const uint32_t bound = n - i;
        const uint64_t threshold = (0ULL - static_cast<uint64_t>(bound)) % bound;
        uint64_t word;
        do { word = mixed(key + counter++); } while (word < threshold);
        uint32_t v = static_cast<uint32_t>(word % bound);
        for (uint32_t s : sorted) v += s <= v ? 1u : 0u;
        out[i] = v;
        sorted.insert(std::upper_bound(sorted.begin(), sorted.end(), v), v);
    }
    return out;
}

}  // namespace faco
