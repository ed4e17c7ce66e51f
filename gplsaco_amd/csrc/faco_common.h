// Shared host/device definitions of the FACO engine: device records, the
// Philox4x32-10 stream, the EUC_2D distance and the SplitMix64 key mix.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/faco_engine.h"

#define FACO_HD __host__ __device__ __forceinline__

namespace faco {

// Largest candidate widths one wave handles: primary/neighbour rows are one
// lane per slot; the 2-opt scan puts successor moves in lanes [0,32) and
// predecessor moves in lanes [32,64).
constexpr uint32_t kMaxPrimary = 64;
constexpr uint32_t kMaxLs = 32;
constexpr uint32_t kMaxBackup = 256;
constexpr uint32_t kWave = 64;
constexpr uint32_t kMaxCoordinate = 32767;

// Per-colony device state (the reference's Colony, cuda/colony_state.cuh:59-64,
// with integer costs and the batch counter kept on the device so one
// captured iteration body can be replayed).
struct ColonyState {
    int64_t global_cost, epoch_cost, parent_cost;
    double minimum, maximum, default_trail, source_uniform;
    uint32_t iteration_best, source_is_epoch;
    uint32_t batch, pad;
    uint64_t work_construction, work_ls, work_constraints;
    uint32_t target;       // MNE target of this batch's ants
    uint32_t parent_zero;  // position of city 0 in the parent | (succ(0) < pred(0)) << 31
    uint64_t best_key;     // smallest (final cost << 32 | ant) of this batch's ants so far; all ones between batches
};

// MMAS bounds from the epoch cost (reference cuda/colony_state.cuh:66-70).
FACO_HD void bounds(ColonyState& s, uint32_t width, double retention, double p) {
    s.maximum = 1.0 / (static_cast<double>(s.epoch_cost) * (1.0 - retention));
    const double lo = s.maximum * (1.0 - p) / ((width - 1.0) * p);
    s.minimum = s.maximum < lo ? s.maximum : lo;
}

// Per-colony control state of the GP / baseline controlled loop (the
// reference's ControllerState, cpp/include/gp_faco/control_ops.hpp:160-173,
// plus this batch's decision). Archive tours live in a two-buffer ring
// ([C][2][4][n] tours and parent-style links); `parity` names the current
// buffer so the update can write the next archive without a scratch copy.
constexpr uint32_t kArchive = 4, kRegion = 16, kSamples = 64;
struct ControlState {
    float return_rate, ls_work;
    uint64_t stagnant, epoch_batches, restarts;
    int64_t tracked_global, tracked_epoch;
    uint64_t active_id[2], epoch_id[2];
    int64_t archive_cost[kArchive];
    uint64_t archive_id[kArchive][2];
    uint32_t archive_zero[kArchive];   // position of city 0 | forward << 31
    uint32_t archive_size, parity;
    uint32_t alternative, action, mask, start_count;
    uint32_t start_nodes[kRegion];     // region the ants of this batch start in
};

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), shared by host and device.
FACO_HD void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int round = 0; round < 10; ++round) {
        const uint64_t p0 = static_cast<uint64_t>(0xD2511F53u) * c[0];
        const uint64_t p1 = static_cast<uint64_t>(0xCD9E8D57u) * c[2];
        const uint32_t n0 = static_cast<uint32_t>(p1 >> 32) ^ c[1] ^ k0;
        const uint32_t n2 = static_cast<uint32_t>(p0 >> 32) ^ c[3] ^ k1;
        c[1] = static_cast<uint32_t>(p1);
        c[3] = static_cast<uint32_t>(p0);
        c[0] = n0;
        c[2] = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

// Stream of 32-bit words for (key, batch, ant, step): counter block
// (step, 0, ant, batch), four words per block, 128-bit counter increment.
// The domain split follows cuda/faco_choices.cuh:32-38 of the reference.
struct Stream {
    uint32_t ctr[4];
    uint32_t block[4];
    uint32_t k0, k1;
    uint32_t index;
    FACO_HD Stream(uint64_t key, uint32_t batch, uint32_t ant, uint32_t step) {
        ctr[0] = step; ctr[1] = 0; ctr[2] = ant; ctr[3] = batch;
        k0 = static_cast<uint32_t>(key); k1 = static_cast<uint32_t>(key >> 32);
        index = 4;
        block[0] = block[1] = block[2] = block[3] = 0;
    }
    FACO_HD uint32_t next() {
        if (index == 4) {
            block[0] = ctr[0]; block[1] = ctr[1]; block[2] = ctr[2]; block[3] = ctr[3];
            philox4x32_10(block, k0, k1);
            index = 0;
            if (++ctr[0] == 0 && ++ctr[1] == 0 && ++ctr[2] == 0) ++ctr[3];
        }
        return block[index++];
    }
    FACO_HD double uniform53() {
        const uint64_t hi = next(), lo = next();
        return static_cast<double>(((hi << 32) | lo) >> 11) * 0x1.0p-53;
    }
};

// First two words of block `step` only: the per-step roulette draw needs
// exactly one block, so the counter never carries.
FACO_HD double step_uniform(uint64_t key, uint32_t batch, uint32_t ant, uint32_t step) {
    uint32_t c[4] = {step, 0u, ant, batch};
    philox4x32_10(c, static_cast<uint32_t>(key), static_cast<uint32_t>(key >> 32));
    const uint64_t hi = c[0], lo = c[1];
    return static_cast<double>(((hi << 32) | lo) >> 11) * 0x1.0p-53;
}

// Rejection-sampled start node over [0,count) in the reserved step
// 0xffffffff (reference cuda/faco_choices.cuh:71-84).
// The first word is accepted unless it falls below (2^32 mod count), so the
// stream is continued only on that rare path; words are taken in the
// stream's order, four per counter block.
FACO_HD uint32_t start_node(uint64_t key, uint32_t batch, uint32_t ant, uint32_t count) {
    const uint32_t threshold = (0u - count) % count;
    const uint32_t k0 = static_cast<uint32_t>(key), k1 = static_cast<uint32_t>(key >> 32);
    uint32_t c[4] = {0xffffffffu, 0u, ant, batch};
    philox4x32_10(c, k0, k1);
    if (c[0] >= threshold) return c[0] % count;
    uint32_t ctr[4] = {0xffffffffu, 0u, ant, batch};
    uint32_t w1 = c[1], w2 = c[2], w3 = c[3];
    for (;;) {
        if (w1 >= threshold) return w1 % count;
        if (w2 >= threshold) return w2 % count;
        if (w3 >= threshold) return w3 % count;
        if (++ctr[0] == 0 && ++ctr[1] == 0 && ++ctr[2] == 0) ++ctr[3];
        uint32_t b[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
        philox4x32_10(b, k0, k1);
        if (b[0] >= threshold) return b[0] % count;
        w1 = b[1]; w2 = b[2]; w3 = b[3];
    }
}

// SplitMix64 finaliser (reference cuda/batch_engine.cu:27-32).
FACO_HD uint64_t mixed(uint64_t value) {
    value += 0x9e3779b97f4a7c15ULL;
    value = (value ^ (value >> 30)) * 0xbf58476d1ce4e5b9ULL;
    value = (value ^ (value >> 27)) * 0x94d049bb133111ebULL;
    return value ^ (value >> 31);
}

FACO_HD uint64_t colony_key(uint64_t instance_key, uint64_t seed) {
    return mixed(seed ^ mixed(instance_key + 0xd1b54a32d192ed03ULL));
}

// Undirected edge identity: two SplitMix64 images of (min << 32 | max); a
// tour's identity is the XOR over its edges (reference
// cpp/include/gp_faco/control_ops.hpp:43-56).
FACO_HD void edge_identity(uint32_t a, uint32_t b, uint64_t& first, uint64_t& second) {
    const uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
    const uint64_t e = (static_cast<uint64_t>(lo) << 32) | hi;
    first = mixed(e ^ 0xb5ad4eceda1ce2a9ULL);
    second = mixed(e ^ 0xd6e8feb86659fd93ULL);
}

// EUC_2D on packed coordinates (x | y << 16), both in [0, 32767]:
// d = nint(sqrt(s)) with s = dx^2 + dy^2 < 2^31, i.e. the r >= 0 with
// r^2 - r < s <= r^2 + r (r = 0 only for s = 0).
// Device: the float root of s is within 0.01 of sqrt(s), so r0 =
// floor(f + 0.5) is within one of d and one comparison each way (both
// against r0) gives d; for r0 = 0 (s = 0) the lower test wraps to false.
// Every product below fits 24 bits of operand (|dx|, |dy| <= 32767 and
// r <= 46341), so the device uses full-rate 24-bit multiplies.
FACO_HD uint32_t euc2d(uint32_t a, uint32_t b) {
    const int dx = static_cast<int>(a & 0xffffu) - static_cast<int>(b & 0xffffu);
    const int dy = static_cast<int>(a >> 16) - static_cast<int>(b >> 16);
#if defined(__HIP_DEVICE_COMPILE__)
    const uint32_t s = static_cast<uint32_t>(__mul24(dx, dx)) + static_cast<uint32_t>(__mul24(dy, dy));
    const uint32_t r = static_cast<uint32_t>(__builtin_amdgcn_sqrtf(static_cast<float>(s)) + 0.5f);
    const uint32_t t = __umul24(r, r);
    return r + ((s > t + r) ? 1u : 0u) - ((t > s + r - 1u) ? 1u : 0u);
#else
    const uint32_t s = static_cast<uint32_t>(dx * dx) + static_cast<uint32_t>(dy * dy);
    uint32_t r = static_cast<uint32_t>(__builtin_sqrtf(static_cast<float>(s)));
    r -= (r * r > s) ? 1u : 0u;
    r += ((r + 1) * (r + 1) <= s) ? 1u : 0u;
    return r + ((s > r * r + r) ? 1u : 0u);
#endif
}

FACO_HD uint32_t pack_xy(uint32_t x, uint32_t y) { return x | (y << 16); }

// ---------------------------------------------------------------------------
// Kernel parameter blocks (passed by value as kernel arguments).
constexpr uint32_t kPhaseSlots = 10;
// Largest instance whose construction runs on the linked-list tour; such
// workgroups also stage the parent positions in LDS. The fallback list
// ranking keeps its entries in registers (kRankedInRegisters / 64 per lane)
// up to kRankedInRegisters cities and works in LDS above that.
constexpr uint32_t kMaxLinked = 2048;
constexpr uint32_t kRankedInRegisters = 1024;
// Default margin of the ant kernel's FP32 roulette fast path, relative to the
// row total (the FP32 and exact FP64 comparisons provably agree outside
// 2^-20.4 of it; tests may only widen it).
constexpr double kRouletteMargin = 0x1.0p-18;

// Per-colony shape of a mixed-size batch: the colony's city count, ants and
// its own iteration budget (a colony past its budget is left unchanged).
// Null in fixed-size engines, where every colony has n and ants of the
// parameter block. In a mixed batch the parameter block's n and ants are the
// strides of every per-city and per-ant array (the batch maxima).
struct ColonyShape { uint32_t n, ants, batches, pad; };
// ColonyState::pad after a colony update that left a finished colony idle
// (0 after every update that ran).
constexpr uint32_t kColonyIdle = 1;
// bit of an ant's canonical start (out_zero / ant_zero): its tour is the
// batch's parent tour (cities < 2^16, so the position needs bits 0..15)
constexpr uint32_t kParentTour = 1u << 30;
// LDS bytes of the archive update's candidate arrays (kArchive + ants
// candidates): cost, identity, then canonical start, group leader, pair
// list and parent-tour flag.
constexpr size_t control_candidate_bytes(uint32_t ants) { return 40u * static_cast<size_t>(kArchive + ants); }

struct AntParams {
    uint32_t nbr_in_lds;                // neighbour rows staged in LDS (else read from HBM)
    uint32_t n, ants, pw, bw, lw, K;   // K = neighbour row width: max(pw, lw), or pw + lw for split rows
    uint32_t ls_offset;                 // LS row within a neighbour row: 0 (prefix of the primary order) or pw
    uint32_t waves;                     // waves per workgroup
    uint32_t chunk;                     // ants of one colony handled by one workgroup
    uint32_t blocks_per_colony;         // workgroups per colony (explicit mode: unused)
    // A launch over part of a mixed batch (its colonies listed from
    // `order`): the whole batch's colony count (0: the grid) and the largest
    // colony n of the launch, which sizes its LDS (0: n)
    uint32_t total_colonies;
    uint32_t lds_n;
    uint32_t nbr_lds_bytes;             // neighbour rows in LDS (16-byte rounded)
    uint32_t nbr_stride;                // u16 elements per geometry (nbr_lds_bytes / 2)
    uint32_t per_wave_bytes;            // per-ant LDS block
    uint32_t xcd_remap;                 // keep colony c's workgroups on XCD c % 8 (its colony update's XCD)
    uint32_t explicit_mode;             // visits[] drive construction (diagnostic)
    uint32_t per_ant_parent;            // parent/links/cost indexed by ant, not colony
    uint32_t log_checklist;             // append every checklist push to diag_checklist
    uint32_t or_opt;                    // local search adds Or-opt segment moves
    uint32_t filter_tours;              // write an ant's tour only while it is the batch's best so far
                                        // (plain FACO reads back the iteration best alone)
    // LS evaluation limit, clamped to 2^32 - 1 by the host: per-ant counts
    // stay far below 2^32 (at most 2 * lw per processed city, fewer than 8n
    // processed cities), so the clamped limit stops exactly where the
    // 64-bit one does
    uint64_t ls_limit;
    double roulette_margin;             // relative margin of the fast roulette (+inf: sequential sums always)
    const uint32_t* coords;             // [G][n] packed x | y << 16
    const uint16_t* nbr;                // [G][n][K] (d, id)-ordered nearest cities
    const uint16_t* nbr_dist;           // [G][n][K] EUC_2D distance of each nbr entry
    const uint16_t* backup;             // [G][n][bw]
    const double* products;             // [C][n][pw] trail * heuristic
    const uint32_t* parent;             // [C][n] (or [A][n])
    const uint32_t* links;              // [C][n] parent predecessor | successor << 16
    const int64_t* parent_cost;         // [A] when per_ant_parent
    ColonyState* state;                 // [C] (parent cost, batch counter, best key of the batch)
    const uint64_t* keys;               // [C]
    const uint32_t* targets;            // [A] MNE targets
    const uint32_t* geometry;           // [C] colony -> geometry, or null (all 0)
    const uint32_t* visits;             // [A][n] explicit choices
    uint32_t* out_tours;                // [A][n]
    faco_ant_info* out_info;            // [A]
    uint32_t* diag_construction;        // [A][n] or null
    uint32_t* diag_positions;           // [A][n] or null
    uint32_t* diag_checklist;           // [A][5n] or null
    uint32_t* trace_selected;           // [A][n] or null
    double* trace_uniform;              // [A][n] or null
    uint64_t* phase_ticks;              // [A][kPhaseSlots] or null: setup, construction, LS (wall clock
                                        // ticks); select, relocate, mne/push, LS scan, LS apply (shader
                                        // clocks); the ant's start and end (absolute wall clock ticks)
    const ControlState* control;        // [C] or null: start cities from the chosen region
    uint64_t* out_identity;             // [A][2] or null: tour identity of each ant
    uint32_t* out_zero;                 // [A] or null: position of city 0 | forward << 31 | kParentTour
    const ColonyShape* shapes;          // [C] or null (fixed size)
    const uint32_t* order;              // [C] or null: colony of each block group (largest work first)
    // explicit mode only (device preparation): block b runs instance b
    // (geometry slot b) of block_n[b] cities, and the checklist starts with
    // every city in tour order instead of the construction's pushes
    const uint32_t* block_n;            // [blocks] or null
    uint32_t full_checklist;
    // Hard mode: undirected CSR graph of every geometry slot, offsets
    // [G][n + 1] and sorted neighbours [G][graph_stride]; null: unrestricted
    const uint32_t* graph_offsets;
    const uint32_t* graph_neighbors;
    uint32_t graph_stride;
    // the same graph as adjacency bits, [G][n][graph_words] (bit b of row a:
    // b is in a's neighbour list), so that a membership test is one load
    const uint32_t* graph_bits;
    uint32_t graph_words;
    // Escape mode (on a Hard graph; null esc_rows: off): per ant its LS
    // rows [A][n][lw] (an anchor's escape view, city | replaced << 15; other
    // rows are not written), footprint [A][n] and counters [A][12] (the
    // reference's EscapeStats order); the parent footprint of every colony
    // [C][n]; d^-beta of every backup slot [G][n][bw]
    uint32_t escape_disabled;           // gates never open (the run equals the Hard run)
    uint16_t* esc_rows;
    uint8_t* esc_footprint;
    uint32_t* esc_stats;
    const uint8_t* parent_fp;
    const double* eta_backup;
};
// Escape counters per ant, in EscapeStats order (escape_slots.hpp:70-75).
constexpr uint32_t kEscStats = 12;
// Per-wave LDS the Escape ant adds: its edge cache (64 packed edges), the
// slot-view scratch (128 B) and the LS order scratch (128 B).
constexpr uint32_t kEscapeWaveBytes = 256u + 128u + 128u;

struct ColonyParams {
    uint32_t n, ants, pw, K, nbr_stride;
    double retention, p_best_root;      // p = p_best^(1/pw), computed on the host
    double epoch_probability;
    const uint16_t* nbr;                // [G][n][K]
    const double* eta;                  // [G][n][pw]
    const uint32_t* geometry;           // [C]
    const uint64_t* keys;               // [C]
    const uint32_t* tours;              // [A][n]
    const faco_ant_info* info;          // [A]
    ColonyState* state;                 // [C]
    uint32_t* parent;                   // [C][n]
    uint32_t* links;                    // [C][n]
    uint32_t* epoch;                    // [C][n]
    uint32_t* global;                   // [C][n]
    double* trails;                     // [C][n][pw]
    double* products;                   // [C][n][pw]
    const ColonyShape* shapes;          // [C] or null (fixed size)
    // Escape mode (null: off): the ants' footprints [A][n] and counters
    // [A][12], the colonies' parent / epoch / global footprints [C][n] and
    // counter totals [C][12]
    const uint8_t* ant_fp;
    const uint32_t* esc_stats;
    uint8_t* parent_fp;
    uint8_t* epoch_fp;
    uint8_t* global_fp;
    uint64_t* esc_totals;
};

struct ResetParams {
    uint32_t n, ants, pw, K, target;
    double retention, p_best_root;
    const uint32_t* geometry;
    const uint32_t* initial;            // [G][n]
    const int64_t* initial_cost;        // [G]
    const double* eta;                  // [G][n][pw]
    ColonyState* state;
    uint32_t* parent; uint32_t* links; uint32_t* epoch; uint32_t* global;
    double* trails; double* products;
    const ColonyShape* shapes;          // [C] or null (fixed size)
};

// Controlled loop (control_kernels.hip). One workgroup per colony.
struct ControlParams {
    uint32_t n, ants, pw, K, nbr_stride;
    uint32_t panel;                     // colonies per decider: decider = colony / panel
    uint32_t mode;                      // FACO_CONTROL_*
    uint32_t variant;                   // factorial variant
    uint32_t experiment_mask;
    uint32_t force_collisions;
    int32_t forced_action;              // >= 0: every colony takes this action (a state fork's first batch)
    uint32_t views_only;                 // decide: write features, scores, masks, actions, alternatives and
                                        // regions only; colony and control state stay unchanged
    uint64_t evaluation_limit;          // FE per colony (progress feature)
    double fixed_progress = -1.0;       // >= 0: the progress feature of this batch (time budgets); else FE-based
    uint64_t ls_limit;                  // LS evaluation limit (LS work feature)
    double retention, p_best_root;
    faco_baseline_policy baseline;
    const uint32_t* coords;             // [G][n]
    const uint16_t* nbr;                // [G][stride] rows, primary = first pw
    const double* eta;                  // [G][n][pw]
    const double* local_scale;          // [G][n]
    const double* epsilon;              // [G]
    const uint32_t* samples;            // [G][64]
    const uint32_t* initial;            // [G][n]
    const int64_t* initial_cost;        // [G]
    const uint32_t* geometry;           // [C]
    const uint64_t* keys;               // [C]
    const faco_program* programs;       // [deciders]
    const faco_controller* controllers; // [deciders]
    ColonyState* state;                 // [C]
    ControlState* control;              // [C]
    uint32_t* parent;                   // [C][n]
    uint32_t* links;                    // [C][n]
    uint32_t* epoch;                    // [C][n]
    double* trails;                     // [C][n][pw]
    double* products;                   // [C][n][pw]
    uint32_t* archive_tours;            // [C][2][4][n]
    uint32_t* archive_links;            // [C][2][4][n]
    const uint32_t* tours;              // [A][n]
    const faco_ant_info* info;          // [A]
    const uint64_t* ant_identity;       // [A][2]
    const uint32_t* ant_zero;           // [A]
    float* features;                    // [12][C][32] or null
    float* scores;                      // [C][32] or null
    uint32_t* masks;                    // [C] or null
    int32_t* actions;                   // [C] or null
    uint32_t* alternatives;             // [C] or null
    uint32_t* regions;                  // [C][2][4][16] or null
    const ColonyShape* shapes;          // [C] or null (fixed size); n and ants above are then strides
    // Escape mode (null archive_fp: off): footprints of the archive
    // [C][2][4][n] (the tours' ring), of the ants [A][n], and the colonies'
    // parent and epoch footprints [C][n], which a restart takes from the
    // archive entry
    uint8_t* archive_fp;
    const uint8_t* ant_fp;
    uint8_t* parent_fp;
    uint8_t* epoch_fp;
};
// Device preparation (prepare_kernels.hip): instances g = 0..count-1 of
// sizes[g] cities at stride `stride` in packed coordinates [count][stride].
// keys [count][stride][needed]: every city's `needed` nearest other cities
// as (distance << 16) | city, ascending; tours [count][stride]: the
// nearest-neighbour tour from city 0 (lowest id on ties).
size_t prepare_knn_lds(uint32_t stride, uint32_t needed);
hipError_t launch_prepare_knn(const uint32_t* coords, const uint32_t* sizes, uint32_t count, uint32_t stride,
                              uint32_t needed, uint32_t* keys, hipStream_t stream);
hipError_t launch_prepare_nn(const uint32_t* coords, const uint32_t* sizes, uint32_t count, uint32_t stride,
                             uint32_t* tours, hipStream_t stream);
hipError_t launch_control_init(const ControlParams& p, uint32_t colonies, hipStream_t stream);
hipError_t launch_control_decide(const ControlParams& p, uint32_t colonies, hipStream_t stream);
hipError_t launch_control_update(const ControlParams& p, uint32_t colonies, hipStream_t stream);
hipError_t launch_score_controller(const faco_controller* controller, const float* features, const uint32_t* masks,
                                   uint32_t colonies, float* scores, int32_t* actions, hipStream_t stream);

// Learned edge prior (prior_kernels.hip).
uint32_t prior_param_count();
uint32_t prior_grad_rows(uint32_t edges);
hipError_t launch_prior_features(uint32_t n, uint32_t pw, const uint32_t* coords, const uint32_t* primary,
                                 const double* local_scale, double epsilon, float mean_scale, float* out,
                                 hipStream_t stream);
// The kernels read the weights from an image in their own LDS layout
// (prior_image_bytes bytes), built from the FP32 parameters by
// launch_prior_image and kept current by every Adam step of
// launch_prior_reduce.
uint32_t prior_image_bytes();
hipError_t launch_prior_image(const float* params, void* image, hipStream_t stream);
hipError_t launch_prior_infer(const void* image, const float* features, uint32_t edges, float* logits,
                              hipStream_t stream);
hipError_t launch_prior_grad(const void* image, const float* features, const uint8_t* labels,
                             const uint32_t* indices, uint32_t offset, uint32_t edges, float pos_weight,
                             float inv_total, float* partial, hipStream_t stream);
hipError_t launch_prior_reduce(const float* partial, uint32_t rows, float* grad, float* params, float* m, float* v,
                               bool adam, float lr, float beta1, float beta2, float eps, uint32_t step,
                               float* loss_sum, void* image, hipStream_t stream);
hipError_t launch_prior_mma_check(const float* A, const float* B, float* C, hipStream_t stream);

// Launch wrappers (defined in faco_kernels.hip).
hipError_t launch_ant_search(const AntParams& p, uint32_t blocks, hipStream_t stream);
int ant_search_blocks_per_cu(const AntParams& p);
hipError_t launch_copy(const void* src, void* dst, size_t bytes, hipStream_t stream);
hipError_t launch_start_check(uint64_t key, uint32_t batch, uint32_t ant, const uint32_t* counts, uint32_t n,
                              uint32_t* out, hipStream_t stream);
hipError_t launch_colony_update(const ColonyParams& p, uint32_t colonies, hipStream_t stream);
// `rounds` batches of plain FACO in one launch, each colony's workgroup
// running its ants and then its colony update (blocks_per_colony must be 1).
hipError_t launch_colony_run(const AntParams& p, const ColonyParams& cp, uint32_t colonies, uint32_t rounds,
                             hipStream_t stream);
hipError_t launch_colony_reset(const ResetParams& p, uint32_t colonies, hipStream_t stream);
// `rounds` batches of GP-controlled FACO in one launch: per colony one
// workgroup runs decision, ants, colony update and archive update batch
// after batch (blocks_per_colony must be 1; no traces, forks or diagnostics).
// `device_ctl` is a copy of `ctl` in device memory (the kernel reads it).
hipError_t launch_control_run(const AntParams& p, const ColonyParams& cp, const ControlParams& ctl,
                              const ControlParams* device_ctl, uint32_t colonies, uint32_t rounds,
                              hipStream_t stream);
uint32_t control_run_lds_bytes(const AntParams& p);
hipError_t launch_distance_check(const int32_t* pairs, uint32_t count, int64_t* out, hipStream_t stream);
hipError_t launch_philox_check(uint64_t key, uint32_t batch, uint32_t ant, uint32_t step, uint32_t words,
                               uint32_t* out, hipStream_t stream);
// LDS bytes of one ant-search workgroup.
inline uint32_t round16_host(uint32_t v) { return (v + 15u) & ~15u; }
// packed coordinates + the colony's parent links + neighbour rows (when
// staged in LDS) + the parent positions of every city (u16, linked-list
// construction only: n <= kMaxLinked)
inline uint32_t ant_shared_bytes(uint32_t n, uint32_t nbr_lds_bytes) {
    return 2u * round16_host(4u * n) + nbr_lds_bytes + (n <= kMaxLinked ? round16_host(2u * n) : 0u);
}
inline uint32_t ant_wave_bytes(uint32_t n) { return 3u * round16_host(2u * n) + round16_host(n); }
// the n that sizes a launch's LDS
inline uint32_t launch_lds_n(const AntParams& p) { return p.lds_n ? p.lds_n : p.n; }

}  // namespace faco
