// Device bodies of the restart and perturbation control of the
// GP-controlled FACO loop: the per-batch decision (features, scoring,
// action, restart transaction) and the archive / feedback update, each run
// by one workgroup for one colony. The per-batch control kernels
// (control_kernels.hip) and the persistent controlled colony kernel
// (faco_kernels.hip, faco_control_run) both call them, so the two launch
// forms compute the same thing.
//
// Semantics: reference cuda/control_state.cuh:105-364 (update_control,
// build_control_features, apply_control_action),
// cpp/include/gp_faco/control_ops.hpp:14-212, baseline_policy.hpp:59-79,
// factorial_policy.hpp:32-45, cuda/baseline_policy.cuh:10-38 and
// cuda/gp_score.cuh:10-102.
//
// The shared-memory state of each body is a struct the caller places: a
// static __shared__ variable in the per-batch kernels, the LDS of the
// (finished) ants in the persistent kernel.
#pragma once
#include "faco_common.h"
#include "gp_program.h"

namespace faco {
namespace {

constexpr double kBand = 0.02;
constexpr uint32_t kNone = 0xffffffffu;

__device__ __forceinline__ double unit_clip(double v) { return v < 0 ? 0 : v > 1 ? 1 : v; }

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
#pragma unroll
    for (int offset = 32; offset; offset >>= 1) v += __shfl_xor(v, offset);
    return v;
}

// Position of the i-th city (i in [0, n)) of a tour read canonically from
// its city-0 position `zero` (bit 31: forward).
__device__ __forceinline__ uint32_t canonical_index(uint32_t zero, uint32_t i, uint32_t n) {
    const uint32_t z = zero & 0x7fffffffu;
    if (zero >> 31) { const uint32_t k = z + i; return k >= n ? k - n : k; }
    const uint32_t k = z + n - i;
    return k >= n ? k - n : k;
}

// canonical_compare (control_ops.hpp:60-71) by one wave: -1, 0 or 1.
__device__ int wave_canonical_compare(const uint32_t* a, uint32_t za, const uint32_t* b, uint32_t zb, uint32_t n,
                                      uint32_t lane) {
    for (uint32_t base = 1; base < n; base += 64) {
        const uint32_t i = base + lane;
        const bool valid = i < n;
        const uint32_t va = valid ? a[canonical_index(za, i, n)] : 0u;
        const uint32_t vb = valid ? b[canonical_index(zb, i, n)] : 0u;
        const uint64_t diff = __ballot(valid && va != vb);
        if (diff) {
            const int first = __builtin_ctzll(diff);
            const uint32_t x = __shfl(va, first), y = __shfl(vb, first);
            return x < y ? -1 : 1;
        }
    }
    return 0;
}

// ControlRandom (control_ops.hpp:81-89).
// The rejection threshold 2^64 mod bound is below bound < 2^32, so a word
// of 2^32 or more always passes and the threshold is computed only for the
// rare smaller words; below 2^16 the remainder is taken in 32-bit steps,
// word = hi 2^32 + lo: ((hi mod b)(2^32 mod b) + lo mod b) mod b.
struct ControlRandom {
    uint64_t key, counter;
    __device__ uint32_t bounded(uint32_t bound) {
        uint64_t word;
        for (;;) {
            word = mixed(key + counter++);
            if (word >> 32) break;
            if (word >= (0ULL - static_cast<uint64_t>(bound)) % bound) break;
        }
        if (bound > 0xffffu) return static_cast<uint32_t>(word % bound);
        const uint32_t hi = static_cast<uint32_t>(word >> 32), lo = static_cast<uint32_t>(word);
        const uint32_t wrap = (0u - bound) % bound;   // 2^32 mod bound
        return ((hi % bound) * wrap + lo % bound) % bound;
    }
};

__device__ __forceinline__ bool has_node(const uint32_t* nodes, uint32_t count, uint32_t v) {
    for (uint32_t i = 0; i < count; ++i) if (nodes[i] == v) return true;
    return false;
}

// Candidate-row reads for the control bodies: `count` entries of a u16 row
// (global memory) in chunks of 16 whose loads are issued together, so a
// search costs one memory round trip per chunk instead of one per entry.
constexpr uint32_t kRowChunk = 16;
__device__ __forceinline__ void row_chunk(const uint16_t* row, uint32_t base, uint32_t count, uint32_t (&v)[kRowChunk]) {
#pragma unroll
    for (uint32_t j = 0; j < kRowChunk; ++j) v[j] = base + j < count ? row[base + j] : 0xffffffffu;
}

// build_region (control_ops.hpp:120-146) by one wave into nodes[kRegion].
// The serial rule's state is spread over the lanes: lane i holds the
// region's i-th city (`mine`) and the i-th smallest of them (`sorted`), so
// a membership test is one ballot, a sorted insertion one ballot and one
// lane shift, and the rank-th remaining city a scan of `used` lane reads.
// The draws are the serial rule's, in its order; every lane runs the same
// control flow.
__device__ void wave_build_region(const uint32_t* order, const uint16_t* nbr, uint32_t K, uint32_t pw, uint32_t n,
                                  uint64_t key, uint32_t batch, uint32_t mode, uint32_t region, uint32_t lane,
                                  uint32_t* nodes) {
    const uint32_t count = n < kRegion ? n : kRegion;
    ControlRandom random{mixed(key ^ 0xbb67ae8584caa73bULL ^ mixed(static_cast<uint64_t>(batch) ^ 0x3c6ef372fe94f82bULL) ^
                               mixed(static_cast<uint64_t>(mode * 4 + region + 1) ^ 0xa54ff53a5f1d36f1ULL)), 0};
    uint32_t mine = 0, sorted = 0xffffffffu, used = 0;
    // lane i's value of a per-lane register (i uniform)
    auto at_lane = [](uint32_t v, uint32_t i) {
        return static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), static_cast<int>(i)));
    };
    auto member = [&](uint32_t v) { return __ballot(lane < used && mine == v) != 0; };
    auto add = [&](uint32_t v) {
        // sorted insertion after every smaller city (the cities are
        // distinct); lanes 0..15 hold the list, so the shift by one lane
        // stays inside the first DPP row
        const uint32_t at = static_cast<uint32_t>(__popcll(__ballot(lane < used && sorted <= v)));
        const uint32_t before = static_cast<uint32_t>(
            __builtin_amdgcn_update_dpp(0, static_cast<int>(sorted), 0x111 /* row_shr:1 */, 0xf, 0xf, false));
        if (lane > at && lane <= used) sorted = before;
        if (lane == at) sorted = v;
        if (lane == used) mine = v;
        ++used;
    };
    auto city = [&](uint32_t i) { return at_lane(mine, i); };
    if (region < 2) {
        const uint32_t start = random.bounded(n);
        const uint32_t k = start + lane;
        if (lane < count) mine = order[k >= n ? k - n : k];
        used = count;
    } else {
        if (region == 2) {
            add(random.bounded(n));
            for (uint32_t head = 0; head < used && used < count; ++head) {
                const uint16_t* row = nbr + static_cast<size_t>(city(head)) * K;
                for (uint32_t base = 0; base < pw && used < count; base += 64) {
                    const uint32_t entry = base + lane < pw ? row[base + lane] : 0xffffffffu;
                    for (uint32_t j = 0; j < 64 && base + j < pw && used < count; ++j) {
                        const uint32_t v = at_lane(entry, j);
                        if (v < n && !member(v)) add(v);
                    }
                }
            }
        }
        while (used < count) {
            // the rank-th city outside the region (control_ops.hpp:92-102)
            uint32_t v = random.bounded(n - used);
            for (uint32_t i = 0; i < used; ++i) v += at_lane(sorted, i) <= v ? 1u : 0u;
            add(v);
        }
    }
    if (lane < kRegion) nodes[lane] = lane < count ? mine : 0u;
}

// baseline_action (baseline_policy.hpp:59-79).
__device__ uint32_t baseline_action(const faco_baseline_policy& p, uint64_t stagnant, uint64_t epoch_batches,
                                    uint32_t batch, double u, uint32_t mask) {
    uint32_t level = p.mne_level;
    bool restart = false;
    if (p.kind == 0) {
        restart = p.restart_mode == 1 ? batch > 0 && batch % p.restart_period == 0
                                      : p.restart_mode == 2 && u < p.restart_probability;
    } else {
        if (p.stagnation_step) {
            const uint64_t extra = stagnant / p.stagnation_step;
            const uint32_t room = p.max_mne_level - p.mne_level;
            level += extra < room ? static_cast<uint32_t>(extra) : room;
        }
        restart = p.restart_stagnation && stagnant >= p.restart_stagnation && epoch_batches >= p.restart_cooldown;
    }
    const uint32_t keep = p.region * 4 + level, requested = keep + 16;
    return restart && (mask & (1u << requested)) ? requested : keep;
}

// factorial_mask (factorial_policy.hpp:32-45).
__device__ uint32_t factorial_mask(uint32_t variant, const faco_baseline_policy& p, uint64_t stagnant,
                                   uint64_t epoch_batches, uint32_t batch, double u, uint32_t legal) {
    if (variant == 3) return legal;
    if (variant == 0) return 1u << baseline_action(p, stagnant, epoch_batches, batch, u, legal);
    const uint32_t requested = baseline_action(p, stagnant, epoch_batches, batch, u, 0xffffffffu);
    if (variant == 2) {
        const uint32_t keep = requested % 16;
        return legal & ((1u << keep) | (1u << (keep + 16)));
    }
    const uint32_t allowed = legal & (requested >= 16 ? 0xffff0000u : 0xffffu);
    return allowed ? allowed : legal & 0xffffu;
}

__device__ __forceinline__ bool same_id(const uint64_t* a, const uint64_t* b) { return a[0] == b[0] && a[1] == b[1]; }
__device__ __forceinline__ bool id_less(const uint64_t* a, const uint64_t* b) {
    return a[0] < b[0] || (a[0] == b[0] && a[1] < b[1]);
}

// Workgroup-shared state of control_decide.
struct DecideShared {
    ControlState c;
    ColonyState s;
    uint32_t alt, mask, action, pad;
    float gap[2], diff, pad2;
    uint32_t nodes[8][kRegion];
    double contrib[128][5];
    float regional[8][4];
    float feat[12 * 32];
    float scores[32];
};

// Workgroup-shared state of control_update; the per-candidate arrays (36
// bytes per candidate, kArchive + ants candidates) come separately.
struct UpdateShared {
    ControlState c;
    ColonyState s;
    uint32_t chosen[kArchive], count, collision, pairs, returned;
    uint64_t best[256];
};

// ---------------------------------------------------------------------------
// build_control_features + scoring + apply_control_action for one colony
// (cuda/control_state.cuh:200-364, cuda/gp_score.cuh, baseline_policy.cuh)
// by the calling workgroup: thread tid of `threads` (at least 128). The
// colony and control state are read from and written to global memory; a
// restart also writes the new parent links to `lds_links` and its
// positions to `lds_ppos` when they are given (the persistent kernel's LDS
// copies).
__device__ __forceinline__ void control_decide(const ControlParams& p, uint32_t colony, uint32_t colonies, uint32_t tid,
                               uint32_t threads, DecideShared& sh, uint32_t* lds_links, uint16_t* lds_ppos) {
    const uint32_t ns = p.n;
    const uint32_t n = p.shapes ? p.shapes[colony].n : ns;
    const uint32_t ants = p.shapes ? p.shapes[colony].ants : p.ants;
    const uint32_t lane = tid & 63u, wave = tid >> 6;
    const uint32_t geom = p.geometry[colony], pw = p.pw, K = p.K;
    ControlState* ctl = p.control + colony;
    ColonyState* st = p.state + colony;
    if (p.shapes && st->batch >= p.shapes[colony].batches) return;   // colony done
    const uint16_t* nbr = p.nbr + static_cast<size_t>(geom) * p.nbr_stride;
    const uint32_t* coords = p.coords + static_cast<size_t>(geom) * ns;
    uint32_t* parent = p.parent + static_cast<size_t>(colony) * ns;
    uint32_t* plinks = p.links + static_cast<size_t>(colony) * ns;
    const double* trails = p.trails + static_cast<size_t>(colony) * ns * pw;

    ControlState& c = sh.c;
    ColonyState& s = sh.s;
    uint32_t &s_alt = sh.alt, &s_mask = sh.mask, &s_action = sh.action;
    float (&s_gap)[2] = sh.gap;
    float& s_diff = sh.diff;
    auto& s_nodes = sh.nodes;
    auto& s_contrib = sh.contrib;
    auto& s_regional = sh.regional;
    float* s_feat = sh.feat;
    float* s_scores = sh.scores;
    if (tid == 0) { c = *ctl; s = *st; }
    __syncthreads();
    const uint32_t* arch_tours = p.archive_tours + (static_cast<size_t>(colony) * 2 + c.parity) * kArchive * ns;
    const uint32_t* arch_links = p.archive_links + (static_cast<size_t>(colony) * 2 + c.parity) * kArchive * ns;
    const uint32_t count = n < kRegion ? n : kRegion;

    // (1) archive alternative: the entry (other than the active tour) that
    // differs most from it on the sampled cities; ties by (cost, identity,
    // canonical order).
    if (wave == 0) {
        const uint32_t* samples = p.samples + static_cast<size_t>(geom) * kSamples;
        const uint32_t sample_count = n < kSamples ? n : kSamples;
        uint32_t alt = kArchive, largest = 0;
        for (uint32_t i = 0; i < c.archive_size; ++i) {
            const uint32_t* entry = arch_tours + static_cast<size_t>(i) * ns;
            if (same_id(c.archive_id[i], c.active_id) &&
                wave_canonical_compare(entry, c.archive_zero[i], parent, s.parent_zero, n, lane) == 0)
                continue;
            const uint32_t* elinks = arch_links + static_cast<size_t>(i) * ns;
            uint32_t miss = 0;
            if (lane < sample_count) {
                const uint32_t v = samples[lane];
                const uint32_t cl = elinks[v], al = plinks[v];
                const uint32_t cs = cl >> 16, cp = cl & 0xffffu, as = al >> 16, ap = al & 0xffffu;
                miss = (as == cs || ap == cs ? 0u : 1u) + (as == cp || ap == cp ? 0u : 1u);
            }
            miss = wave_sum_u32(miss);
            bool better_tie = false;
            if (alt != kArchive && miss == largest) {
                better_tie = c.archive_cost[i] < c.archive_cost[alt];
                if (c.archive_cost[i] == c.archive_cost[alt]) {
                    if (id_less(c.archive_id[i], c.archive_id[alt])) better_tie = true;
                    else if (same_id(c.archive_id[i], c.archive_id[alt]))
                        better_tie = wave_canonical_compare(entry, c.archive_zero[i], arch_tours + static_cast<size_t>(alt) * ns,
                                                            c.archive_zero[alt], n, lane) < 0;
                }
            }
            if (alt == kArchive || miss > largest || better_tie) { alt = i; largest = miss; }
        }
        if (lane == 0) {
            s_alt = alt;
            const double global = static_cast<double>(s.global_cost);
            s_gap[0] = static_cast<float>(unit_clip((static_cast<double>(s.parent_cost) / global - 1) / kBand));
            s_gap[1] = alt < kArchive
                ? static_cast<float>(unit_clip((static_cast<double>(c.archive_cost[alt]) / global - 1) / kBand)) : 0.0f;
            s_diff = alt < kArchive ? static_cast<float>(largest) / (2 * sample_count) : 0.0f;
            s_mask = p.experiment_mask & (alt < kArchive ? 0xffffffffu : 0xffffu);
        }
    }
    __syncthreads();
    const uint32_t alt = s_alt;

    // (2) start regions: one wave per (mode, region), the waves taking
    // them in turn
    for (uint32_t g = wave; g < 8; g += threads / 64) {
        const uint32_t mode = g / 4, region = g % 4;
        if (mode == 0 || alt < kArchive) {
            const uint32_t* order = mode == 0 ? parent : arch_tours + static_cast<size_t>(alt) * ns;
            wave_build_region(order, nbr, K, pw, n, p.keys[colony], s.batch, mode, region, lane, s_nodes[g]);
        } else if (lane < kRegion) {
            s_nodes[g][lane] = 0;
        }
    }
    __syncthreads();

    // (3) regional statistics: one thread per (region, node); the sums run
    // in node order so the FP64 totals match the serial rule.
    if (tid < 128) {
        const uint32_t g = tid / 16, l = tid % 16, mode = g / 4;
        double* v = s_contrib[tid];
        v[0] = v[1] = v[2] = v[3] = v[4] = 0;
        if ((mode == 0 || alt < kArchive) && l < count) {
            const uint32_t node = s_nodes[g][l];
            const uint32_t link = mode == 0 ? plinks[node] : arch_links[static_cast<size_t>(alt) * ns + node];
            const uint32_t prev = link & 0xffffu, next = link >> 16;
            const double scale_raw = p.local_scale[static_cast<size_t>(geom) * ns + node];
            const double eps = p.epsilon[geom];
            const double scale = scale_raw < eps ? eps : scale_raw;
            const double near = 0.5 * static_cast<double>(euc2d(coords[node], coords[prev])) +
                                0.5 * static_cast<double>(euc2d(coords[node], coords[next]));
            v[0] = unit_clip((near / scale - 1) / 3);
            // slot of each neighbour in the node's primary row (pw: none);
            // the row is read in chunks whose loads are issued together
            const uint32_t neighbours[2] = {prev, next};
            uint32_t slot[2] = {pw, pw};
            const uint16_t* row = nbr + static_cast<size_t>(node) * K;
            for (uint32_t base = 0; base < pw && (slot[0] == pw || slot[1] == pw); base += kRowChunk) {
                uint32_t chunk[kRowChunk];
                row_chunk(row, base, pw, chunk);
#pragma unroll
                for (uint32_t j = 0; j < kRowChunk; ++j)
                    for (uint32_t k = 0; k < 2; ++k)
                        if (chunk[j] == neighbours[k] && slot[k] == pw && base + j < pw) slot[k] = base + j;
            }
            for (uint32_t k = 0; k < 2; ++k) {
                const uint32_t w = neighbours[k];
                // integer-valued terms: exact in any order
                for (uint32_t j = 0; j < c.archive_size; ++j) {
                    const uint32_t al = arch_links[static_cast<size_t>(j) * ns + node];
                    v[1] += ((al & 0xffffu) == w || (al >> 16) == w) ? 0.0 : 1.0;
                }
                const double trail = slot[k] < pw ? trails[static_cast<size_t>(node) * pw + slot[k]] : s.default_trail;
                v[2 + k] = s.minimum == s.maximum ? 1.0 : unit_clip((trail - s.minimum) / (s.maximum - s.minimum));
                v[4] += has_node(s_nodes[g], count, w) ? 0.0 : 1.0;
            }
        }
    }
    __syncthreads();
    if (tid < 8) {
        const uint32_t mode = tid / 4;
        float r[4] = {0, 0, 0, 0};
        if (mode == 0 || alt < kArchive) {
            double excess = 0, disagreement = 0, strength = 0, dispersion = 0;
            for (uint32_t i = 0; i < count; ++i) {
                const double* v = s_contrib[tid * 16 + i];
                excess += v[0]; disagreement += v[1];
                strength += v[2]; strength += v[3]; dispersion += v[4];
            }
            r[0] = static_cast<float>(excess / count);
            r[1] = static_cast<float>(disagreement / (2 * count * c.archive_size));
            r[2] = static_cast<float>(strength / (2 * count));
            r[3] = static_cast<float>(dispersion / (2 * count));
        }
        for (uint32_t k = 0; k < 4; ++k) s_regional[tid][k] = r[k];
    }
    __syncthreads();

    // (4) the 12 features of the 32 actions
    if (tid < 32) {
        const uint32_t a = tid, mode = a / 16, region = (a / 4) % 4, level = a % 4;
        const double progress = p.fixed_progress >= 0
            ? p.fixed_progress
            : static_cast<double>(static_cast<uint64_t>(s.batch) * ants) / static_cast<double>(p.evaluation_limit);
        const float* r = s_regional[mode * 4 + region];
        const float values[12] = {static_cast<float>(unit_clip(progress)),
            static_cast<float>(c.stagnant < 32 ? c.stagnant / 32.0 : 1), c.return_rate, c.ls_work,
            static_cast<float>(mode), (level + 1) * 0.25f, s_gap[mode], mode ? s_diff : 0.0f,
            r[0], r[1], r[2], r[3]};
        for (uint32_t f = 0; f < 12; ++f) {
            s_feat[f * 32 + a] = values[f];
            if (p.features) p.features[(static_cast<size_t>(f) * colonies + colony) * 32 + a] = values[f];
        }
    }
    __syncthreads();

    // (5) decision by wave 0 (lanes 32-63 mirror lanes 0-31 and write nothing)
    if (wave == 0) {
        const uint32_t a = lane & 31u;
        const uint32_t decider = colony / p.panel;
        uint32_t mask = s_mask;
        uint32_t action = 0;
        if (p.mode == FACO_CONTROL_BASELINE || p.mode == FACO_CONTROL_FACTORIAL) {
            double u = 0;
            if (p.baseline.kind == 0 && p.baseline.restart_mode == 2) {
                Stream coin(p.keys[colony], s.batch, 0xffffffffu, 1u);
                u = coin.uniform53();
            }
            if (p.mode == FACO_CONTROL_BASELINE) {
                action = baseline_action(p.baseline, c.stagnant, c.epoch_batches, s.batch, u, mask);
                if (lane < 32) s_scores[a] = 0.0f;
            } else {
                mask = factorial_mask(p.variant, p.baseline, c.stagnant, c.epoch_batches, s.batch, u, mask);
            }
        }
        if (p.mode == FACO_CONTROL_PROGRAM || p.mode == FACO_CONTROL_FACTORIAL) {
            const faco_program& prog = p.programs[decider];
            const float value = program_value(prog, s_feat, 1, 0, a);
            if (lane < 32) s_scores[a] = value;
            action = static_cast<uint32_t>(group32_choose(value, (mask >> a) & 1u, a));
        } else if (p.mode == FACO_CONTROL_CONTROLLER) {
            const faco_controller& ctrl = p.controllers[decider];
            action = static_cast<uint32_t>(controller_decide(ctrl, s_feat, 1, 0, mask, a, lane < 32 ? s_scores : nullptr));
        }
        // a state fork replaces the first batch's decision (reference
        // cuda/batch_engine.cu:922-923); the mask stays the policy's
        if (p.forced_action >= 0) action = static_cast<uint32_t>(p.forced_action);
        if (lane == 0) { s_action = action; s_mask = mask; }
    }
    __syncthreads();
    const uint32_t action = s_action;

    // (6) trace outputs and the start region of this batch's ants
    if (tid < 32) {
        if (p.scores) p.scores[static_cast<size_t>(colony) * 32 + tid] = s_scores[tid];
        if (p.regions) for (uint32_t k = 0; k < 4; ++k)
            p.regions[static_cast<size_t>(colony) * 128 + tid * 4 + k] = s_nodes[(tid * 4 + k) / 16][(tid * 4 + k) % 16];
    }
    if (p.views_only) {
        if (tid == 0) {
            if (p.masks) p.masks[colony] = s_mask;
            if (p.actions) p.actions[colony] = static_cast<int32_t>(action);
            if (p.alternatives) p.alternatives[colony] = alt;
        }
        return;
    }
    if (tid < kRegion) ctl->start_nodes[tid] = s_nodes[(action / 16) * 4 + (action / 4) % 4][tid];
    if (tid == 0) {
        if (p.masks) p.masks[colony] = s_mask;
        if (p.actions) p.actions[colony] = static_cast<int32_t>(action);
        if (p.alternatives) p.alternatives[colony] = alt;
        ctl->start_count = count;
        ctl->alternative = alt;
        ctl->action = action;
        ctl->mask = s_mask;
        st->target = 2u << (action % 4);
    }
    if (action < 16) return;
    // (7) restart from the archive alternative: parent = epoch = that tour,
    // bounds from its cost, every trail at the upper bound
    __syncthreads();
    if (tid == 0) {
        s.epoch_cost = s.parent_cost = c.archive_cost[alt];
        bounds(s, pw, p.retention, p.p_best_root);
        s.default_trail = s.maximum;
        s.source_uniform = 0; s.source_is_epoch = 0; s.iteration_best = 0;
        s.parent_zero = c.archive_zero[alt];
        s.target = 2u << (action % 4);
        *st = s;
        ctl->tracked_epoch = s.epoch_cost;
        ctl->active_id[0] = ctl->epoch_id[0] = c.archive_id[alt][0];
        ctl->active_id[1] = ctl->epoch_id[1] = c.archive_id[alt][1];
        ctl->return_rate = 0.0f; ctl->ls_work = 0.0f; ctl->epoch_batches = 0;
        if (c.restarts != ~0ull) ctl->restarts = c.restarts + 1;
    }
    __syncthreads();
    uint32_t* epoch = p.epoch + static_cast<size_t>(colony) * ns;
    const uint32_t* src = arch_tours + static_cast<size_t>(alt) * ns;
    const uint32_t* src_links = arch_links + static_cast<size_t>(alt) * ns;
    for (uint32_t i = tid; i < n; i += threads) {
        parent[i] = epoch[i] = src[i];
        plinks[i] = src_links[i];
        if (lds_links) lds_links[i] = src_links[i];
        if (lds_ppos) lds_ppos[src[i]] = static_cast<uint16_t>(i);
    }
    if (p.archive_fp) {
        // Escape mode: the restart carries the entry's footprint
        // (reference cuda/control_state.cuh:339-340)
        const uint8_t* fp = p.archive_fp + ((static_cast<size_t>(colony) * 2 + c.parity) * kArchive + alt) * ns;
        for (uint32_t i = tid; i < n; i += threads)
            p.parent_fp[static_cast<size_t>(colony) * ns + i] = p.epoch_fp[static_cast<size_t>(colony) * ns + i] = fp[i];
    }
    const uint32_t cells = n * pw;
    double* tr = p.trails + static_cast<size_t>(colony) * ns * pw;
    double* pr = p.products + static_cast<size_t>(colony) * ns * pw;
    const double* eta = p.eta + static_cast<size_t>(geom) * ns * pw;
    const double maximum = s.maximum;
    for (uint32_t i = tid; i < cells; i += threads) {
        tr[i] = maximum;
        pr[i] = maximum * eta[i];
    }
}

// ---------------------------------------------------------------------------
// update_control (cuda/control_state.cuh:105-198) after selection and
// pheromone, by the calling workgroup (thread tid of `threads`): archive of
// the global best plus up to three distinct tours within 2% of it,
// return-rate / LS-work feedback, tracked identities. Candidates: archive
// entries 0..3 then the colony's ants; `lds` holds
// control_candidate_bytes(ants) bytes.
__device__ __forceinline__ void control_update(const ControlParams& p, uint32_t colony, uint32_t tid, uint32_t threads,
                               UpdateShared& sh, unsigned char* lds) {
    // n, A: this colony's cities and ants; ns and p.ants: the strides
    const uint32_t ns = p.n;
    const uint32_t n = p.shapes ? p.shapes[colony].n : ns;
    const uint32_t A = p.shapes ? p.shapes[colony].ants : p.ants;
    const uint32_t lane = tid & 63u, wave = tid >> 6;
    // threads of the selection reductions: a power of two, at most 256
    uint32_t nt = threads < 256u ? threads : 256u;
    while (nt & (nt - 1u)) nt &= nt - 1u;
    const uint32_t M = kArchive + A;
    ControlState* ctl = p.control + colony;
    const ColonyState* st = p.state + colony;
    // a colony past its iteration budget was left idle by this iteration's
    // colony update (which marks it so)
    if (p.shapes && st->pad == kColonyIdle) return;
    const size_t ant_base = static_cast<size_t>(colony) * p.ants;
    ControlState& c = sh.c;
    ColonyState& s = sh.s;
    uint32_t* s_chosen = sh.chosen;
    uint32_t &s_count = sh.count, &s_collision = sh.collision, &s_pairs = sh.pairs;
    uint64_t* s_best = sh.best;
    // per candidate: cost, identity, canonical start, group leader
    int64_t* cost = reinterpret_cast<int64_t*>(lds);
    uint64_t* ident = reinterpret_cast<uint64_t*>(lds + 8u * M);
    uint32_t* zero = reinterpret_cast<uint32_t*>(lds + 24u * M);
    uint32_t* leader = zero + M;
    uint32_t* pairs = leader + M;          // candidates to check against their leader
    uint32_t* parent_tour = pairs + M;     // ants whose tour is the batch's parent tour
    if (tid == 0) { c = *ctl; s = *st; s_collision = 0; s_pairs = 0; sh.returned = 0; }
    __syncthreads();
    const uint32_t parity = c.parity;
    const uint32_t* old_tours = p.archive_tours + (static_cast<size_t>(colony) * 2 + parity) * kArchive * ns;
    uint32_t* new_tours = p.archive_tours + (static_cast<size_t>(colony) * 2 + (parity ^ 1u)) * kArchive * ns;
    uint32_t* new_links = p.archive_links + (static_cast<size_t>(colony) * 2 + (parity ^ 1u)) * kArchive * ns;
    const uint32_t* ant_tours = p.tours + ant_base * ns;
    auto tour_of = [&](uint32_t k) -> const uint32_t* {
        return k < kArchive ? old_tours + static_cast<size_t>(k) * ns : ant_tours + static_cast<size_t>(k - kArchive) * ns;
    };
    const double band = static_cast<double>(s.global_cost) * (1 + kBand);
    const bool improved = s.global_cost < c.tracked_global;
    const uint32_t first = improved ? kArchive + s.iteration_best : 0u;

    // candidates outside the 2% band (or empty archive slots) take cost -1
    for (uint32_t k = tid; k < M; k += threads) {
        int64_t ck;
        uint64_t i0 = 0, i1 = 0;
        uint32_t z = 0;
        if (k < kArchive) {
            ck = k < c.archive_size ? c.archive_cost[k] : -1;
            if (k < c.archive_size) { i0 = c.archive_id[k][0]; i1 = c.archive_id[k][1]; z = c.archive_zero[k]; }
        } else {
            const uint32_t a = k - kArchive;
            ck = p.info[ant_base + a].final_cost;
            if (!p.force_collisions) { i0 = p.ant_identity[2 * (ant_base + a)]; i1 = p.ant_identity[2 * (ant_base + a) + 1]; }
            z = p.ant_zero[ant_base + a];
        }
        if (ck >= 0 && static_cast<double>(ck) > band) ck = -1;
        cost[k] = ck; ident[2 * k] = i0; ident[2 * k + 1] = i1;
        zero[k] = z & ~kParentTour;
        parent_tour[k] = (z & kParentTour) ? 1u : 0u;
    }
    __syncthreads();
    // group leader: the first candidate with the same (cost, identity).
    // Every member is compared with its leader below, except where
    // equality is already known: ants that both hold the batch's parent
    // tour, and a parent-tour ant after the group's first parent-tour
    // member (`rep`), whose own comparison covers it.
    // (The costs are read eight at a time, the loads of a chunk together.)
    for (uint32_t k = tid; k < M; k += threads) {
        uint32_t lead = k, rep = kNone;
        const int64_t ck = cost[k];
        if (ck >= 0) {
            const uint64_t i0 = ident[2 * k], i1 = ident[2 * k + 1];
            bool done = false;
            for (uint32_t base = 0; base < k && !done; base += 8) {
                int64_t cj[8];
#pragma unroll
                for (uint32_t u = 0; u < 8; ++u) cj[u] = base + u < k ? cost[base + u] : -1;
#pragma unroll
                for (uint32_t u = 0; u < 8; ++u) {
                    const uint32_t j = base + u;
                    if (done || cj[u] != ck || ident[2 * j] != i0 || ident[2 * j + 1] != i1) continue;
                    if (lead == k) {
                        lead = j;
                        done = !parent_tour[k] || parent_tour[j];
                    } else if (parent_tour[j]) {
                        rep = j;
                        done = true;
                    }
                }
            }
        }
        leader[k] = lead;
        if (lead != k && (!parent_tour[k] || (!parent_tour[lead] && rep == kNone)))
            pairs[atomicAdd(&s_pairs, 1u)] = k;
    }
    __syncthreads();
    // every group member must be its leader's tour (else: identity collision)
    {
        // (members x positions) flattened over the workgroup, eight load
        // pairs in flight per thread; (4 + ants) * n < 2^32
        constexpr uint32_t kFlight = 8;
        const uint32_t P = s_pairs, span = n - 1, total = P * span, step = threads;
        bool bad = false;
        for (uint32_t base = tid; base < total; base += kFlight * step) {
            uint32_t va[kFlight], vb[kFlight];
#pragma unroll
            for (uint32_t u = 0; u < kFlight; ++u) {
                const uint32_t k = base + u * step;
                va[u] = vb[u] = 0;
                if (k < total) {
                    const uint32_t q = pairs[k / span], i = 1 + k % span, l = leader[q];
                    va[u] = tour_of(q)[canonical_index(zero[q], i, n)];
                    vb[u] = tour_of(l)[canonical_index(zero[l], i, n)];
                }
            }
#pragma unroll
            for (uint32_t u = 0; u < kFlight; ++u) bad |= va[u] != vb[u];
        }
        if (bad) s_collision = 1;
    }
    __syncthreads();

    if (!s_collision) {
        // groups are distinct tours with distinct (cost, identity): the
        // archive is the global best's group, then the three best other
        // groups by (cost, identity), each represented by its leader
        if (tid == 0) { s_chosen[0] = first; s_count = 1; }
        __syncthreads();
        for (uint32_t slot = 1; slot < kArchive; ++slot) {
            // pick the smallest leader by (cost, identity, index) in three
            // passes of a 64-bit min reduction
            uint32_t best = kNone;
            for (uint32_t k = tid; tid < nt && k < M; k += nt) {
                if (cost[k] < 0 || leader[k] != k) continue;
                bool taken = false;
                for (uint32_t j = 0; j < s_count; ++j) taken |= leader[s_chosen[j]] == k;
                if (taken) continue;
                if (best == kNone || cost[k] < cost[best] ||
                    (cost[k] == cost[best] && (ident[2 * k] < ident[2 * best] ||
                     (ident[2 * k] == ident[2 * best] && ident[2 * k + 1] < ident[2 * best + 1]))))
                    best = k;
            }
            // leaders have distinct (cost, identity), so the reduction key
            // can be the candidate's rank in that order: reduce pairwise
            if (tid < nt) s_best[tid] = best;
            __syncthreads();
            for (uint32_t stride = nt / 2; stride; stride >>= 1) {
                if (tid < stride) {
                    const uint32_t x = static_cast<uint32_t>(s_best[tid]), y = static_cast<uint32_t>(s_best[tid + stride]);
                    uint32_t r = x;
                    if (x == kNone) r = y;
                    else if (y != kNone && (cost[y] < cost[x] || (cost[y] == cost[x] &&
                             (ident[2 * y] < ident[2 * x] || (ident[2 * y] == ident[2 * x] && ident[2 * y + 1] < ident[2 * x + 1])))))
                        r = y;
                    s_best[tid] = r;
                }
                __syncthreads();
            }
            const uint32_t winner = static_cast<uint32_t>(s_best[0]);
            __syncthreads();
            if (winner == kNone) break;
            if (tid == 0) s_chosen[s_count++] = winner;
            __syncthreads();
        }
    } else if (wave == 0) {
        // identity collision: the serial rule with canonical comparisons
        auto cmp = [&](uint32_t x, uint32_t y) {
            return wave_canonical_compare(tour_of(x), zero[x], tour_of(y), zero[y], n, lane);
        };
        uint32_t chosen[kArchive] = {first, kNone, kNone, kNone};
        uint32_t count = 1;
        for (uint32_t slot = 1; slot < kArchive; ++slot) {
            uint32_t best = kNone;
            for (uint32_t k = 0; k < M; ++k) {
                if (cost[k] < 0) continue;
                bool dup = false;
                for (uint32_t j = 0; j < count && !dup; ++j) {
                    const uint32_t o = chosen[j];
                    dup = k == o || (ident[2 * k] == ident[2 * o] && ident[2 * k + 1] == ident[2 * o + 1] && cmp(k, o) == 0);
                }
                if (dup) continue;
                if (best == kNone) { best = k; continue; }
                if (cost[k] != cost[best]) { if (cost[k] < cost[best]) best = k; continue; }
                if (ident[2 * k] != ident[2 * best] || ident[2 * k + 1] != ident[2 * best + 1]) {
                    if (ident[2 * k] < ident[2 * best] || (ident[2 * k] == ident[2 * best] && ident[2 * k + 1] < ident[2 * best + 1]))
                        best = k;
                    continue;
                }
                if (cmp(k, best) < 0) best = k;
            }
            if (best == kNone) break;
            chosen[count++] = best;
        }
        if (lane == 0) {
            for (uint32_t j = 0; j < kArchive; ++j) s_chosen[j] = chosen[j];
            s_count = count;
        }
    }
    __syncthreads();

    // write the next archive (tours, links; every slot's positions over the
    // workgroup at once) and stage each ant's LS work term for the feedback
    const uint32_t count = s_count;
    for (uint32_t k = tid; k < count * n; k += threads) {
        const uint32_t slot = k / n, i = k - slot * n;
        const uint32_t* src = tour_of(s_chosen[slot]);
        const uint32_t v = src[i];
        const uint32_t pred = src[i == 0 ? n - 1 : i - 1], succ = src[i + 1 == n ? 0 : i + 1];
        new_tours[static_cast<size_t>(slot) * ns + i] = v;
        new_links[static_cast<size_t>(slot) * ns + v] = pred | (succ << 16);
    }
    if (p.archive_fp) {
        // Escape mode: each entry's footprint moves with it
        // (reference cuda/control_state.cuh:177-189)
        const uint8_t* old_fp = p.archive_fp + (static_cast<size_t>(colony) * 2 + parity) * kArchive * ns;
        uint8_t* new_fp = p.archive_fp + (static_cast<size_t>(colony) * 2 + (parity ^ 1u)) * kArchive * ns;
        for (uint32_t k = tid; k < count * n; k += threads) {
            const uint32_t slot = k / n, i = k - slot * n, e = s_chosen[slot];
            new_fp[static_cast<size_t>(slot) * ns + i] =
                e < kArchive ? old_fp[static_cast<size_t>(e) * ns + i] : p.ant_fp[(ant_base + e - kArchive) * ns + i];
        }
    }
    double* work_terms = reinterpret_cast<double*>(cost);   // the costs are no longer read
    for (uint32_t a = tid; a < A; a += threads) {
        work_terms[a] = p.ls_limit ? static_cast<double>(p.info[ant_base + a].move_evaluations) /
                                         static_cast<double>(p.ls_limit) : 0.0;
        const uint32_t k = kArchive + a;
        if (ident[2 * k] == c.active_id[0] && ident[2 * k + 1] == c.active_id[1]) atomicAdd(&sh.returned, 1u);
    }
    __syncthreads();
    if (tid == 0) {
        ControlState out = c;
        // feedback: ants that returned to the active tour (an integer
        // count, exact in any order) and mean LS work, summed in ant order
        // (cuda/control_state.cuh:166-172)
        const double returned = static_cast<double>(sh.returned);
        double work = 0;
        if (p.ls_limit)
            for (uint32_t a = 0; a < A; ++a) work += work_terms[a];
        if (A) {
            const float r = static_cast<float>(unit_clip(returned / A));
            const float w = static_cast<float>(unit_clip(work / A));
            out.return_rate += (r - out.return_rate) * 0.0625f;
            out.ls_work += (w - out.ls_work) * 0.0625f;
            if (improved) out.stagnant = 0;
            else if (out.stagnant != ~0ull) ++out.stagnant;
            if (out.epoch_batches != ~0ull) ++out.epoch_batches;
        }
        const uint32_t ib = kArchive + s.iteration_best;
        if (s.epoch_cost < c.tracked_epoch) { out.epoch_id[0] = ident[2 * ib]; out.epoch_id[1] = ident[2 * ib + 1]; }
        if (s.source_is_epoch) { out.active_id[0] = out.epoch_id[0]; out.active_id[1] = out.epoch_id[1]; }
        else { out.active_id[0] = ident[2 * ib]; out.active_id[1] = ident[2 * ib + 1]; }
        out.tracked_global = s.global_cost;
        out.tracked_epoch = s.epoch_cost;
        // costs of ants were read before the band filter; reread the raw ones
        for (uint32_t j = 0; j < kArchive; ++j) {
            if (j < count) {
                const uint32_t k = s_chosen[j];
                out.archive_cost[j] = k < kArchive ? c.archive_cost[k] : p.info[ant_base + k - kArchive].final_cost;
                out.archive_id[j][0] = ident[2 * k]; out.archive_id[j][1] = ident[2 * k + 1];
            } else {
                out.archive_cost[j] = 0; out.archive_id[j][0] = out.archive_id[j][1] = 0;
            }
        }
        out.archive_size = count;
        out.parity = parity ^ 1u;
        // canonical starts of the new entries
        for (uint32_t j = 0; j < kArchive; ++j) out.archive_zero[j] = j < count ? zero[s_chosen[j]] : 0u;
        *ctl = out;
    }
}

}  // namespace
}  // namespace faco
