// HIP/CDNA4 kernels of the FACO batch engine (gfx950).
//
// ant_search: one wave64 per ant. The instance's packed coordinates, the
//   colony's parent links and positions and (when they fit) the
//   distance-ordered neighbour rows sit in LDS once per workgroup; each wave
//   keeps its ant's tour, inverse positions, FIFO checklist and visited/queued
//   flags in LDS (7 bytes per city; during construction the tour is a doubly
//   linked list over the same bytes up to kMaxLinked cities). Scalar FACO
//   state (current city, step, MNE, queue head/tail, cost) is wave-uniform
//   and lives in SGPRs; lanes work on the parallel parts: the 16-slot
//   roulette (ballot availability, an FP32 prefix with an exact FP64
//   fallback), list ranking, relocation shifts and 2-opt flips (64 elements
//   per instruction), and the 2x20 2-opt move scan with a wave-wide (gain,
//   lower index) maximum. Distances are EUC_2D from LDS coordinates computed
//   in registers, except d(a, b) of the 2-opt scan, read from the instance's
//   table of neighbour-row distances (u16, L2-resident).
//   Instantiations: per batch or persistent (the workgroup also runs its
//   colony's update between batches, and in controlled runs the decision and
//   the archive update), unrestricted, Hard or Escape, plain 2-opt or with
//   Or-opt, production, clock-profiled or diagnostic.
//   Semantics: reference cuda/faco_device.cuh:100-501 (construct_and_search),
//   cuda/faco_choices.cuh:86-142 (StochasticChoices::next).
//
// colony_update: one workgroup per colony: work totals, iteration best
//   (lowest ant on ties), epoch/global update, MMAS bounds, epoch-source coin,
//   next parent and its predecessor/successor links, and the fused
//   evaporate/deposit/clamp/product pheromone update over the colony's
//   [n][primary] entries (each entry owned by one thread, no atomics).
//   Semantics: reference cuda/colony_state.cuh:148-229.
#include "control_device.h"
#include "escape_device.h"
#include "faco_common.h"

namespace faco {

// ---------------------------------------------------------------------------
// Wave helpers. Values that are identical in every lane are made explicit
// with readfirstlane so the compiler can keep them in SGPRs.
__device__ __forceinline__ uint32_t uniform(uint32_t v) {
    return static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(v)));
}
__device__ __forceinline__ uint32_t read_lane(uint32_t v, uint32_t lane) {
    return static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), static_cast<int>(lane)));
}
__device__ __forceinline__ double read_lane(double v, uint32_t lane) {
    const uint64_t bits = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = read_lane(static_cast<uint32_t>(bits), lane);
    const uint32_t hi = read_lane(static_cast<uint32_t>(bits >> 32), lane);
    return __builtin_bit_cast(double, (static_cast<uint64_t>(hi) << 32) | lo);
}
__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }
// The value itself, opaque to the optimiser: a loop-invariant condition
// written on it is evaluated where it is used (one scalar compare) instead
// of being hoisted as a 64-bit lane mask that the register allocator then
// spills and reloads inside the loop.
__device__ __forceinline__ uint32_t fresh(uint32_t v) {
    asm volatile("" : "+s"(v));
    return v;
}
// LDS writes by one lane become visible to the other lanes of the same wave
// in program order; this keeps the compiler from moving LDS accesses across
// the hand-off points.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
__device__ __forceinline__ uint32_t ctz64(uint64_t m) { return static_cast<uint32_t>(__builtin_ctzll(m)); }
__device__ __forceinline__ uint32_t last64(uint64_t m) { return 63u - static_cast<uint32_t>(__builtin_clzll(m)); }
__device__ __forceinline__ uint32_t nth_set_bit(uint64_t m, uint32_t k) {
    for (uint32_t i = 0; i < k; ++i) m &= m - 1;
    return ctz64(m);
}
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
#pragma unroll
    for (int offset = 32; offset; offset >>= 1) {
        const uint32_t o = static_cast<uint32_t>(__shfl_xor(static_cast<int>(v), offset, 64));
        v = o > v ? o : v;
    }
    return v;
}
__device__ __forceinline__ uint64_t wave_min_u64(uint64_t v) {
#pragma unroll
    for (int offset = 32; offset; offset >>= 1) {
        const uint64_t o = static_cast<uint64_t>(__shfl_xor(static_cast<unsigned long long>(v), offset, 64));
        v = o < v ? o : v;
    }
    return v;
}

constexpr uint8_t kVisited = 1, kQueued = 2, kAnchor = 4;   // kAnchor: Escape-mode anchor city


__device__ __forceinline__ uint32_t round16(uint32_t v) { return (v + 15u) & ~15u; }

// Wave-wide maximum of a u32 with DPP row operations (quad swaps, half-row
// and row mirrors) and one scalar combine of the four 16-lane rows.
__device__ __forceinline__ uint32_t dpp_max_u32(uint32_t v) {
    // every lane has a source lane in these patterns, so the moves need no
    // old value and fold into the max instructions
    uint32_t o;
    o = static_cast<uint32_t>(__builtin_amdgcn_mov_dpp(static_cast<int>(v), 0xB1, 0xF, 0xF, false)); v = o > v ? o : v;
    o = static_cast<uint32_t>(__builtin_amdgcn_mov_dpp(static_cast<int>(v), 0x4E, 0xF, 0xF, false)); v = o > v ? o : v;
    o = static_cast<uint32_t>(__builtin_amdgcn_mov_dpp(static_cast<int>(v), 0x141, 0xF, 0xF, false)); v = o > v ? o : v;
    o = static_cast<uint32_t>(__builtin_amdgcn_mov_dpp(static_cast<int>(v), 0x140, 0xF, 0xF, false)); v = o > v ? o : v;
    // every lane of a row now holds the row maximum; row broadcasts fold
    // row 0 into row 1, row 2 into row 3, then row 1 into rows 2 and 3, so
    // lane 63 ends with the wave maximum (lanes of rows outside a
    // broadcast's row mask keep their value)
    int x = static_cast<int>(v);
    o = static_cast<uint32_t>(__builtin_amdgcn_update_dpp(x, x, 0x142, 0xA, 0xF, false)); v = o > v ? o : v;
    x = static_cast<int>(v);
    o = static_cast<uint32_t>(__builtin_amdgcn_update_dpp(x, x, 0x143, 0xC, 0xF, false)); v = o > v ? o : v;
    return read_lane(v, 63);
}

// Profile: per-ant phase and sub-phase clocks into p.phase_ticks (the
// production instantiation compiles the counters out).
// Exact in-order running sums over the 16 lanes of each row: after step t,
// lane j holds ((v[j-t] + v[j-t+1]) + ...) + v[j] (left to right), so after
// 15 steps lane j holds the sequential prefix sum v[0] + ... + v[j] of its
// row, bit for bit the left-to-right sum (adding the +0.0 of an
// unavailable slot leaves a non-negative sum unchanged).
__device__ __forceinline__ double row_prefix_sum(double v) {
    double acc = v;
#pragma unroll
    for (int t = 0; t < 15; ++t) {
        const uint64_t bits = __builtin_bit_cast(uint64_t, acc);
        const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(bits), 0x111, 0xF, 0xF, true));
        const uint32_t hi = static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(bits >> 32), 0x111, 0xF, 0xF, true));
        acc = __builtin_bit_cast(double, (static_cast<uint64_t>(hi) << 32) | lo) + v;
    }
    return acc;
}

// Prefix sums over the 16 lanes of each row in 4 doubling steps in FP32 (one
// DPP add per level; lane j ends with v[0] + ... + v[j] in a tree order, not
// left to right): every term passes through at most 4 roundings, so each
// prefix is within 4 * 2^-24 of the exact sum of the (non-negative) terms,
// relative to the row total.
template<int Shift> __device__ __forceinline__ float row_shifted_f32(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x110 + Shift, 0xF, 0xF, true));
}
__device__ __forceinline__ float row_scan4_f32(float v) {
    float acc = v;
    acc = acc + row_shifted_f32<1>(acc);
    acc = acc + row_shifted_f32<2>(acc);
    acc = acc + row_shifted_f32<4>(acc);
    acc = acc + row_shifted_f32<8>(acc);
    return acc;
}
__device__ __forceinline__ float read_lane(float v, uint32_t lane) {
    return __builtin_bit_cast(float, read_lane(__builtin_bit_cast(uint32_t, v), lane));
}

// Colony update of one colony by the calling workgroup (threads tid of
// `threads`, a multiple of 64): work totals, iteration best (lowest ant on
// ties), epoch/global update, MMAS bounds, epoch-source coin, next parent
// and its links, and the fused evaporate/deposit/clamp/product update
// (reference cuda/colony_state.cuh:148-229). `red` holds 4 words per wave,
// `s` is the workgroup's copy of the colony state and `s_links` [n] receives
// the new parent links (LDS); `s_ppos` [n] (LDS, may be null) the parent
// position of every city.
// Footprints: the Escape-mode footprint and counter updates are compiled in
// (the per-batch update kernel and the persistent Escape kernel).
template<bool Footprints = false>
__device__ __forceinline__ void update_colony(const ColonyParams& p, uint32_t colony, uint32_t tid, uint32_t threads,
                                              uint64_t* red, ColonyState& s, uint32_t* s_links, uint16_t* s_ppos) {
    const uint32_t ns = p.n;
    const uint32_t n = p.shapes ? p.shapes[colony].n : ns;
    const uint32_t ants = p.shapes ? p.shapes[colony].ants : p.ants;
    const uint32_t geom = p.geometry ? p.geometry[colony] : 0u;
    ColonyState* st = p.state + colony;
    const faco_ant_info* info = p.info + static_cast<size_t>(colony) * p.ants;
    uint64_t best = ~0ull, wc = 0, wl = 0, wr = 0;
    for (uint32_t a = tid; a < ants; a += threads) {
        // final costs are positive and < 2^31: (cost, ant) orders ties by ant
        const uint64_t k = (static_cast<uint64_t>(info[a].final_cost) << 32) | a;
        best = k < best ? k : best;
        wc += info[a].steps;
        wl += info[a].move_evaluations;
        wr += info[a].constraint_rejections;
    }
#pragma unroll
    for (int offset = 32; offset; offset >>= 1) {
        const uint64_t o = __shfl_xor(best, offset);
        best = o < best ? o : best;
        wc += __shfl_xor(wc, offset);
        wl += __shfl_xor(wl, offset);
        wr += __shfl_xor(wr, offset);
    }
    if ((tid & 63u) == 0) {
        uint64_t* r = red + 4 * (tid >> 6);
        r[0] = best; r[1] = wc; r[2] = wl; r[3] = wr;
    }
    __syncthreads();
    if (tid == 0) {
        uint64_t key = ~0ull, total_c = 0, total_l = 0, total_r = 0;
        for (uint32_t w = 0; w < threads / 64u; ++w) {
            const uint64_t* r = red + 4 * w;
            key = r[0] < key ? r[0] : key;
            total_c += r[1];
            total_l += r[2];
            total_r += r[3];
        }
        s = *st;
        s.work_construction += total_c;
        s.work_ls += total_l;
        s.work_constraints += total_r;
        const uint32_t ib = static_cast<uint32_t>(key);
        const int64_t cost = static_cast<int64_t>(key >> 32);
        s.iteration_best = ib;
        const bool improve_epoch = cost < s.epoch_cost, improve_global = cost < s.global_cost;
        if (improve_epoch) s.epoch_cost = cost;
        if (improve_global) s.global_cost = cost;
        bounds(s, p.pw, p.retention, p.p_best_root);
        Stream coin(p.keys[colony], s.batch, 0xffffffffu, 0u);
        s.source_uniform = coin.uniform53();
        s.source_is_epoch = s.source_uniform < p.epoch_probability ? 1u : 0u;
        s.parent_cost = s.source_is_epoch ? s.epoch_cost : cost;
        s.pad = (improve_epoch ? 1u : 0u) | (improve_global ? 2u : 0u);
        s.best_key = ~0ull;   // the next batch's ants publish into an empty slot
    }
    __syncthreads();
    const uint32_t improve = s.pad;
    const uint32_t* best_tour = p.tours + (static_cast<size_t>(colony) * p.ants + s.iteration_best) * ns;
    uint32_t* epoch = p.epoch + static_cast<size_t>(colony) * ns;
    uint32_t* global = p.global + static_cast<size_t>(colony) * ns;
    uint32_t* parent = p.parent + static_cast<size_t>(colony) * ns;
    uint32_t* links = p.links + static_cast<size_t>(colony) * ns;
    if (improve & 1u) for (uint32_t i = tid; i < n; i += threads) epoch[i] = best_tour[i];
    if (improve & 2u) for (uint32_t i = tid; i < n; i += threads) global[i] = best_tour[i];
    __syncthreads();
    const uint32_t* source = s.source_is_epoch ? epoch : best_tour;
    for (uint32_t i = tid; i < n; i += threads) {
        const uint32_t v = source[i];
        const uint32_t pred = source[i == 0 ? n - 1 : i - 1], succ = source[i + 1 == n ? 0 : i + 1];
        parent[i] = v;
        s_links[v] = pred | (succ << 16);
        links[v] = pred | (succ << 16);
        if (s_ppos) s_ppos[v] = static_cast<uint16_t>(i);
        if (v == 0) s.parent_zero = i | (succ < pred ? 0x80000000u : 0u);
    }
    if (Footprints && p.ant_fp) {
        // Escape mode: footprints follow their tours (reference
        // cuda/colony_state.cuh:201-206); the ants' counters add up per colony
        const uint8_t* best_fp = p.ant_fp + (static_cast<size_t>(colony) * p.ants + s.iteration_best) * ns;
        uint8_t* efp = p.epoch_fp + static_cast<size_t>(colony) * ns;
        uint8_t* gfp = p.global_fp + static_cast<size_t>(colony) * ns;
        uint8_t* pfp = p.parent_fp + static_cast<size_t>(colony) * ns;
        for (uint32_t i = tid; i < n; i += threads) {
            const uint8_t v = best_fp[i];
            if (improve & 1u) efp[i] = v;
            if (improve & 2u) gfp[i] = v;
            pfp[i] = s.source_is_epoch ? efp[i] : v;
        }
        // one wave per counter, its lanes over the ants (callers launch
        // whole waves)
        const uint32_t lane = tid & 63u, waves = threads >> 6;
        for (uint32_t k = tid >> 6; k < kEscStats; k += waves) {
            uint64_t sum = 0;
            for (uint32_t a = lane; a < ants; a += 64u)
                sum += p.esc_stats[(static_cast<size_t>(colony) * p.ants + a) * kEscStats + k];
#pragma unroll
            for (int offset = 32; offset; offset >>= 1)
                sum += static_cast<uint64_t>(__shfl_xor(static_cast<unsigned long long>(sum), offset, 64));
            if (lane == 0) p.esc_totals[static_cast<size_t>(colony) * kEscStats + k] += sum;
        }
    }
    __syncthreads();
    // Evaporate, clamp below, deposit on parent edges, clamp above, refresh
    // products (reference cuda/colony_state.cuh:210-229).
    const uint32_t cells = n * p.pw;
    double* trails = p.trails + static_cast<size_t>(colony) * ns * p.pw;
    double* products = p.products + static_cast<size_t>(colony) * ns * p.pw;
    const double* eta = p.eta + static_cast<size_t>(geom) * ns * p.pw;
    const uint16_t* nbr = p.nbr + static_cast<size_t>(geom) * p.nbr_stride;
    const double minimum = s.minimum, maximum = s.maximum, retention = p.retention;
    const double increment = 1.0 / static_cast<double>(s.parent_cost);
    for (uint32_t i = tid; i < cells; i += threads) {
        const uint32_t from = i / p.pw, j = i - from * p.pw;
        const uint32_t to = nbr[from * p.K + j];
        double v = trails[i] * retention;
        v = v > minimum ? v : minimum;
        const uint32_t link = s_links[from];
        if (to == (link & 0xffffu) || to == (link >> 16)) {
            const double raised = v + increment;
            v = raised < maximum ? raised : maximum;
        }
        trails[i] = v;
        products[i] = v * eta[i];
    }
    __syncthreads();
    if (tid == 0) {
        const double d = s.default_trail * retention;
        s.default_trail = d > minimum ? d : minimum;
        s.batch += 1;
        s.pad = 0;
        *st = s;
    }
}

__global__ void __launch_bounds__(256) faco_colony_update(const ColonyParams p) {
    const uint32_t colony = blockIdx.x, tid = threadIdx.x;
    if (p.shapes && p.state[colony].batch >= p.shapes[colony].batches) {   // colony done: mark it idle
        if (tid == 0) p.state[colony].pad = kColonyIdle;
        return;
    }
    __shared__ uint64_t red[4 * 4];
    __shared__ ColonyState s;
    extern __shared__ uint32_t s_links[];
    update_colony<true>(p, colony, tid, blockDim.x, red, s, s_links, nullptr);
}

// ant_search: the body of the ant kernel. Persistent: the workgroup runs
// `rounds` whole batches of its colony, each followed by the colony update
// (update_colony) in the same workgroup, keeping the colony state in
// `s_state` (LDS); otherwise one batch and the update is a separate kernel.
// Controlled (persistent only): each batch is preceded by the control
// decision and followed by the archive / feedback update of `ctl`, both in
// the same workgroup, on the LDS the ants leave free between batches.
template<bool NbrLds, bool Profile, bool Diag, bool OrOpt, bool Persistent, bool Hard = false,
         bool Controlled = false, bool Escape = false>
__device__ __forceinline__ void ant_search(const AntParams& p, const ColonyParams& cp, uint32_t rounds,
                                           ColonyState* s_state, const ControlParams* ctl = nullptr) {
    static_assert(!Controlled || Persistent, "the fused control runs in the persistent kernel");
    static_assert(!Escape || (Hard && !Controlled && !Diag && !Profile && !OrOpt),
                  "Escape mode runs on a Hard graph with the plain 2-opt");
    // Diagnostic features (explicit visits, per-ant parents, traces and
    // logs) exist only in the Diag instantiation: the production kernel
    // keeps none of their pointers or flags live.
    const bool explicit_mode = Diag && p.explicit_mode;
    const bool per_ant_parent = Diag && p.per_ant_parent;
    uint32_t* const trace_selected = Diag ? p.trace_selected : nullptr;
    double* const trace_uniform = Diag ? p.trace_uniform : nullptr;
    uint32_t* const diag_construction = Diag ? p.diag_construction : nullptr;
    uint32_t* const diag_positions = Diag ? p.diag_positions : nullptr;
    uint32_t* const diag_checklist = Diag ? p.diag_checklist : nullptr;
    extern __shared__ __attribute__((aligned(16))) unsigned char lds[];
    __shared__ uint32_t next_slot;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wave = uniform(threadIdx.x >> 6);
    uint32_t blk = blockIdx.x;
    const uint32_t nblocks = gridDim.x;
    // Workgroups are dispatched round-robin over the 8 XCDs (block b on XCD
    // b % 8), and the per-colony kernels (colony update, reset, control) run
    // colony c on XCD c % 8. The remap keeps every workgroup of colony c on
    // that same XCD, so the colony's pheromone products written by the
    // update stay in the L2 the ant kernel reads them from.
    if (p.xcd_remap && (nblocks & 7u) == 0 && ((nblocks / p.blocks_per_colony) & 7u) == 0) {
        const uint32_t x = blk & 7u, k = blk >> 3;
        blk = (x + 8u * (k / p.blocks_per_colony)) * p.blocks_per_colony + k % p.blocks_per_colony;
    }
    // One workgroup = one chunk of one colony's ants; its waves take ants
    // from the chunk dynamically so a slow ant does not idle the others. In a
    // mixed-size batch the block groups follow `order` (largest work first:
    // the hardware dispatches workgroups in index order, so the longest
    // colonies start first) and each colony has its own n and ants; `ns` is
    // the stride of every per-city array.
    const uint32_t group = explicit_mode ? 0u : blk / p.blocks_per_colony;
    const uint32_t colony = p.order ? p.order[group] : group;
    const uint32_t ns = p.n;
    const uint32_t* const block_n = Diag ? p.block_n : nullptr;
    const uint32_t n = block_n ? block_n[blk] : (p.shapes ? p.shapes[colony].n : ns);
    const uint32_t chunk = p.shapes ? p.shapes[colony].ants / p.blocks_per_colony : p.chunk;
    if (!Persistent && p.shapes && p.state[colony].batch >= p.shapes[colony].batches) return;   // colony done
    const uint32_t chunk_base = (explicit_mode ? blk : blk % p.blocks_per_colony) * chunk;
    const uint32_t geom = block_n ? blk : (p.geometry ? p.geometry[colony] : 0u);

    // Shared data: packed coordinates, the colony's parent links (the
    // parent tour is the same for every ant of the colony in an iteration;
    // the explicit mode's per-ant parents stay in global memory), the
    // neighbour rows and, for linked-list construction (a stride of at
    // most kMaxLinked cities), the parent position of every city.
    const bool with_ppos = ns <= kMaxLinked;
    uint32_t* s_xy = reinterpret_cast<uint32_t*>(lds);
    uint32_t* s_plinks = reinterpret_cast<uint32_t*>(lds + round16(4u * n));
    uint16_t* s_ppos = reinterpret_cast<uint16_t*>(lds + 2u * round16(4u * n) + (NbrLds ? p.nbr_lds_bytes : 0u));
    const uint32_t shared_bytes = 2u * round16(4u * n) + (NbrLds ? p.nbr_lds_bytes : 0u) +
                                  (with_ppos ? round16(2u * n) : 0u);
    const uint16_t* s_nbr;
    {
        const uint32_t* gxy = p.coords + static_cast<size_t>(geom) * ns;
        for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) s_xy[i] = gxy[i];
        if (!explicit_mode) {
            const uint32_t* glinks = p.links + static_cast<size_t>(colony) * ns;
            for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) s_plinks[i] = glinks[i];
            if (with_ppos) {
                const uint32_t* gparent = p.parent + static_cast<size_t>(colony) * ns;
                for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) s_ppos[gparent[i]] = static_cast<uint16_t>(i);
            }
        }
        const uint16_t* g_nbr = p.nbr + static_cast<size_t>(geom) * p.nbr_stride;
        if constexpr (NbrLds) {
            uint16_t* l_nbr = reinterpret_cast<uint16_t*>(lds + 2u * round16(4u * n));
            const uint4* src = reinterpret_cast<const uint4*>(g_nbr);
            uint4* dst = reinterpret_cast<uint4*>(l_nbr);
            const uint32_t chunks = (2u * n * p.K + 15u) / 16u;   // this instance's rows
            for (uint32_t i = threadIdx.x; i < chunks; i += blockDim.x) dst[i] = src[i];
            s_nbr = l_nbr;
        } else {
            s_nbr = g_nbr;
        }
    }
    if (threadIdx.x == 0) {
        next_slot = 0;
        if constexpr (Persistent) *s_state = p.state[colony];
    }
    __syncthreads();

    unsigned char* mine = lds + shared_bytes + wave * p.per_wave_bytes;
    uint16_t* tour = reinterpret_cast<uint16_t*>(mine);
    uint16_t* pos = reinterpret_cast<uint16_t*>(mine + round16(2u * n));
    uint16_t* queue = reinterpret_cast<uint16_t*>(mine + 2u * round16(2u * n));
    uint8_t* flags = reinterpret_cast<uint8_t*>(mine + 3u * round16(2u * n));
    // Escape mode: this ant's cache of out-of-graph edges (min | max << 16)
    // and the scratch of its slot views and LS orders, after its own arrays
    unsigned char* const esc_area = mine + 3u * round16(2u * ns) + round16(ns);
    uint32_t* const esc_cache = reinterpret_cast<uint32_t*>(esc_area);
    uint8_t* const esc_scratch = esc_area + 4u * kEscapeCapacity;
    uint32_t* const esc_order = reinterpret_cast<uint32_t*>(esc_area + 4u * kEscapeCapacity + 128u);

    const uint32_t pw = p.pw, K = p.K, lw = p.lw;
    const uint64_t key = explicit_mode ? 0ull : p.keys[colony];
    const double* products = p.products + static_cast<size_t>(colony) * ns * pw;
    const uint16_t* backup = p.backup + static_cast<size_t>(geom) * ns * p.bw;
    const uint32_t limit = static_cast<uint32_t>(p.ls_limit);   // clamped to 32 bits by the host
    // Hard mode (reference cpp/include/gp_faco/edge_constraints.hpp:15-90):
    // the instance's undirected CSR graph; a move is allowed when every
    // added edge that does not cancel a removed one (as a multiset) is a
    // graph edge
    // (held as adjacency bits: b is a neighbour of a when bit b of row a is
    // set, the rows built from the sorted CSR lists)
    const uint32_t graph_words = p.graph_words;
    const uint32_t* g_bits = Hard ? p.graph_bits + static_cast<size_t>(geom) * ns * graph_words : nullptr;
    auto in_graph = [&](uint32_t a, uint32_t b) -> bool {
        if (a >= n || b >= n || a == b) return false;
        return (g_bits[a * graph_words + (b >> 5)] >> (b & 31u)) & 1u;
    };
    auto edges_allowed = [&](uint32_t count, const uint32_t (&removed)[3][2], const uint32_t (&added)[3][2]) -> bool {
        bool consumed[3] = {false, false, false};
        bool ok = true;
        for (uint32_t i = 0; i < count; ++i) {
            bool existing = false;
            for (uint32_t j = 0; j < count; ++j) {
                const bool same = (added[i][0] == removed[j][0] && added[i][1] == removed[j][1]) ||
                                  (added[i][0] == removed[j][1] && added[i][1] == removed[j][0]);
                if (!existing && !consumed[j] && same) { consumed[j] = true; existing = true; }
            }
            if (!existing && !in_graph(added[i][0], added[i][1])) ok = false;
        }
        return ok;
    };
    const bool in_row = lane < pw;
    const uint64_t row_bits = pw >= kWave ? ~0ull : (1ull << pw) - 1ull;   // ballot(in_row)
    // Product-row loads are issued by every lane (lanes past the row re-read
    // its last slot), not under an exec mask: with no branch around them the
    // compiler counts the loads in flight exactly and waits only for the row
    // it needs, not for the prefetches issued after it.
    const uint32_t row_lane = in_row ? lane : (pw ? pw - 1u : 0u);

    for (uint32_t round = 0;; ++round) {
    if constexpr (Persistent) {
        if (round == rounds || (p.shapes && s_state->batch >= p.shapes[colony].batches)) break;
    }
    if constexpr (Controlled) {
        // this batch's decision: features, scoring, action, restart (which
        // also refreshes the LDS parent links and positions)
        control_decide(*ctl, colony, p.total_colonies ? p.total_colonies : nblocks, threadIdx.x, blockDim.x,
                       *reinterpret_cast<DecideShared*>(lds + shared_bytes),
                       s_plinks, with_ppos ? s_ppos : nullptr);
        __syncthreads();
        if (threadIdx.x == 0) *s_state = p.state[colony];
        __syncthreads();
    }
    const uint32_t batch = explicit_mode ? 0u : (Persistent ? s_state->batch : p.state[colony].batch);
    // Escape mode: the colony's default trail and the backup rows' d^-beta
    const double dtrail = Escape ? p.state[colony].default_trail : 0.0;
    const double* const eta_backup = Escape ? p.eta_backup + static_cast<size_t>(geom) * ns * p.bw : nullptr;
    for (;;) {
        uint32_t slot = 0;
        if (lane == 0) slot = atomicAdd(&next_slot, 1u);
        slot = uniform(slot);
        if (slot >= chunk) break;
        const uint32_t local_ant = chunk_base + slot;                // ant index within the colony
        const uint32_t ant = colony * p.ants + local_ant;            // global ant index
        const uint64_t tick0 = Profile ? wall_clock64() : 0;
        uint64_t sub[5] = {0, 0, 0, 0, 0};  // select, relocate, mne/push, LS scan, LS apply (shader clocks)
        uint64_t mark = 0;

        const size_t parent_base = static_cast<size_t>(per_ant_parent ? ant : colony) * ns;
        const uint32_t* parent = p.parent + parent_base;
        const uint32_t* links = p.links + parent_base;   // explicit mode only; else s_plinks
        // Construction keeps the working tour as a doubly linked list
        // (predecessor | successor << 16 per city, over the tour/pos bytes)
        // plus the city at array position 0; the array itself is rebuilt
        // after construction. Beyond kMaxLinked cities the array is updated
        // in place instead (both give the same array).
        const bool linked = n <= kMaxLinked;
        uint32_t* lnk = reinterpret_cast<uint32_t*>(mine);
        uint32_t head = 0;
        if (linked) {
            if (explicit_mode) {
                for (uint32_t i = lane; i < n; i += kWave) { lnk[i] = links[i]; flags[i] = 0; }
            } else {
                for (uint32_t i = lane; i < n; i += kWave) { lnk[i] = s_plinks[i]; flags[i] = 0; }
            }
            head = uniform(parent[0]);
        } else {
            for (uint32_t i = lane; i < n; i += kWave) {
                const uint32_t v = parent[i];
                tour[i] = static_cast<uint16_t>(v);
                pos[v] = static_cast<uint16_t>(i);
                flags[i] = 0;
            }
        }
        wave_sync();

        const uint32_t target = uniform(explicit_mode ? p.targets[ant]
                                      : (Persistent ? s_state->target : p.state[colony].target));
        const uint32_t* visits = p.visits + static_cast<size_t>(ant) * ns;
        uint32_t* checklist_log = diag_checklist ? diag_checklist + static_cast<size_t>(ant) * 5u * ns : nullptr;
        uint32_t current;
        if (explicit_mode) {
            current = visits[0];
        } else if (p.control) {
            // controlled runs start in the region the colony's action chose
            const ControlState* ctl = p.control + colony;
            current = ctl->start_nodes[start_node(key, batch, local_ant, ctl->start_count)];
        } else {
            current = start_node(key, batch, local_ant, n);
        }
        current = uniform(current);
        const uint32_t start = current;
        if (trace_selected && lane == 0) {
            trace_selected[static_cast<size_t>(ant) * ns] = current;
            trace_uniform[static_cast<size_t>(ant) * ns] = 0.0;
        }
        if (lane == 0) flags[current] = kVisited;
        // The visit order (start, then each selected city) is recorded at the
        // top of the checklist ring, entry i at queue[n - 1 - i], while the
        // construction's pushes fill it from the bottom; an entry is written
        // only while it lies above every push so far, and the record is used
        // only if the two never met (logged + steps < n at the end).
        const bool chain_rec = linked && with_ppos && !explicit_mode && !per_ant_parent;
        if (chain_rec && lane == 0) queue[n - 1] = static_cast<uint16_t>(current);
        const uint32_t record_limit = chain_rec ? n : 0u;   // record step s while logged + s < record_limit
        // array mode: position of `current`; linked mode: its links
        uint32_t pc = linked ? 0u : pos[current];
        uint32_t cur_link = linked ? uniform(lnk[current]) : 0u;
        wave_sync();

        uint32_t steps = 0, mne = 0, relocations = 0, exhausted = 0;
        // FIFO checklist: a circular buffer of n entries (a city is queued at
        // most once at a time); qhead / qtail are its positions, qsize its
        // length and logged the number of pushes so far.
        uint32_t qhead = 0, qtail = 0, qsize = 0, logged = 0;
        const uint64_t tick1 = Profile ? wall_clock64() : 0;

        // Append `node` to the FIFO checklist unless it is already pending
        // (reference append_if_absent with queued flags, cuda/faco_device.cuh:81-92).
        auto push = [&](uint32_t node) -> uint32_t {
            const uint8_t f = flags[node];
            if (f & kQueued) return 0u;
            if (lane == 0) {
                flags[node] = static_cast<uint8_t>(f | kQueued);
                queue[qtail] = static_cast<uint16_t>(node);
                if (checklist_log) checklist_log[logged] = node;
            }
            qtail = qtail + 1 == n ? 0u : qtail + 1;
            ++qsize;
            ++logged;
            wave_sync();
            return 1u;
        };
        // The same for up to four distinct cities at once, pushed in lane
        // order (lane i < count holds the i-th city; a caller passes count
        // lanes of distinct cities): one flag read, then the fresh ones take
        // consecutive queue slots in that order. Equal to pushing them one by
        // one, since pushing one city leaves the others' flags unchanged.
        auto push_distinct = [&](uint32_t node, uint32_t count) -> uint32_t {
            const bool member = lane < count;
            const uint8_t f = flags[node];
            const bool fresh = member && !(f & kQueued);
            const uint64_t fresh_mask = ballot(fresh);
            if (fresh) {
                const uint32_t rank = __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(fresh_mask), 0u);
                const uint32_t slot = qtail + rank;
                flags[node] = static_cast<uint8_t>(f | kQueued);
                queue[slot >= n ? slot - n : slot] = static_cast<uint16_t>(node);
                if (checklist_log) checklist_log[logged + rank] = node;
            }
            const uint32_t pushed = static_cast<uint32_t>(__builtin_popcountll(fresh_mask));
            qtail += pushed;
            qtail = qtail >= n ? qtail - n : qtail;
            qsize += pushed;
            logged += pushed;
            wave_sync();
            return pushed;
        };

        // Hard mode: may city c move to follow `cur`, whose successor in the
        // working tour is cur_after (reference relocation_allowed,
        // edge_constraints.hpp:70-82; availability, faco_choices.cuh:46-53)?
        auto relocation_ok = [&](uint32_t c, uint32_t cur, uint32_t cur_after) -> bool {
            if (c >= n || c == cur) return false;
            if (cur_after == c) return true;
            uint32_t before, after;
            if (linked) {
                const uint32_t l = lnk[c];
                before = l & 0xffffu; after = l >> 16;
            } else {
                const uint32_t pc_c = pos[c];
                before = tour[pc_c == 0 ? n - 1 : pc_c - 1];
                after = tour[pc_c + 1 == n ? 0 : pc_c + 1];
            }
            const uint32_t removed[3][2] = {{before, c}, {c, after}, {cur, cur_after}};
            const uint32_t added[3][2] = {{before, after}, {cur, c}, {c, cur_after}};
            return edges_allowed(3, removed, added);
        };

        // Escape mode (reference escape_edges.hpp:13-122): the ant's cache of
        // out-of-graph edges holds esc_size packed edges (wave-uniform); an
        // edge is known when it is a graph or cache edge. A group of added
        // edges passes when each one cancels a removed edge (as a multiset)
        // or is known, or, with permission, is new; the new ones must fit in
        // the cache. Status 0 allowed, 1 invalid edge, 2 outside the graph,
        // 3 cache full; `novel` receives the new edges.
        uint32_t esc_size = 0;
        uint32_t e_gates = 0, e_replaced = 0, e_relocations = 0, e_anchor_nodes = 0, e_ls_replaced = 0;
        uint32_t e_old_views = 0, e_anchor_pushes = 0, e_new = 0, e_construction_full = 0, e_ls_full = 0;
        auto esc_known = [&](uint32_t x, uint32_t y) -> bool {
            if (in_graph(x, y)) return true;
            const uint32_t packed = x < y ? (x | (y << 16)) : (y | (x << 16));
            bool hit = false;
            for (uint32_t k = 0; k < esc_size; ++k) hit |= esc_cache[k] == packed;
            return hit;
        };
        auto esc_status = [&](uint32_t count, const uint32_t (&removed)[3][2], const uint32_t (&added)[3][2],
                              bool permit, uint32_t (&novel)[3], uint32_t& novel_count) -> uint32_t {
            bool consumed[3] = {false, false, false};
            novel_count = 0;
            for (uint32_t i = 0; i < count; ++i) {
                const uint32_t x = added[i][0], y = added[i][1];
                if (x >= n || y >= n || x == y) return 1u;
                bool unchanged = false;
                for (uint32_t j = 0; j < count; ++j) {
                    const bool same = (x == removed[j][0] && y == removed[j][1]) || (x == removed[j][1] && y == removed[j][0]);
                    if (!unchanged && !consumed[j] && same) { consumed[j] = true; unchanged = true; }
                }
                if (unchanged || esc_known(x, y)) continue;
                if (!permit) return 2u;
                const uint32_t packed = x < y ? (x | (y << 16)) : (y | (x << 16));
                bool duplicate = false;
                for (uint32_t j = 0; j < novel_count; ++j) duplicate |= novel[j] == packed;
                if (!duplicate) novel[novel_count++] = packed;
            }
            return novel_count > kEscapeCapacity - esc_size ? 3u : 0u;
        };
        // moving c to follow `cur` (whose successor is cur_after)
        auto esc_relocation = [&](uint32_t c, uint32_t cur, uint32_t cur_after, bool permit, uint32_t (&novel)[3],
                                  uint32_t& novel_count) -> uint32_t {
            novel_count = 0;
            if (c >= n || c == cur) return 1u;
            if (cur_after == c) return 0u;
            uint32_t before, after;
            if (linked) {
                const uint32_t l = lnk[c];
                before = l & 0xffffu; after = l >> 16;
            } else {
                const uint32_t pc_c = pos[c];
                before = tour[pc_c == 0 ? n - 1 : pc_c - 1];
                after = tour[pc_c + 1 == n ? 0 : pc_c + 1];
            }
            const uint32_t removed[3][2] = {{before, c}, {c, after}, {cur, cur_after}};
            const uint32_t added[3][2] = {{before, after}, {cur, c}, {c, cur_after}};
            return esc_status(3, removed, added, permit, novel, novel_count);
        };
        // a move's new edges join the cache (the values are wave-uniform;
        // a chosen move always fits, its availability checked the room)
        auto esc_commit = [&](const uint32_t (&novel)[3], uint32_t novel_count) {
            if (novel_count > kEscapeCapacity - esc_size) novel_count = 0;
            if (lane == 0) for (uint32_t k = 0; k < novel_count; ++k) esc_cache[esc_size + k] = novel[k];
            esc_size += novel_count;
            e_new += novel_count;
            wave_sync();
        };
        // availability without new edges (backup and nearest-city fallbacks)
        auto reloc_ok = [&](uint32_t c, uint32_t cur, uint32_t cur_after) -> bool {
            if constexpr (Escape) {
                uint32_t novel[3], count;
                return esc_relocation(c, cur, cur_after, false, novel, count) == 0u;
            } else {
                return relocation_ok(c, cur, cur_after);
            }
        };

        // ------------------------------------------------------------ construction
        // The product rows of both parent neighbours of the current city are
        // prefetched at the start of a step: the roulette mostly follows a
        // parent edge, so the next step usually finds its row already
        // loaded. A miss loads the selected city's row right after the
        // selection. The row a step uses is picked only when that step
        // starts (row_src), so the loads' latency overlaps the relocation
        // and the checklist pushes instead of ending the selection with a
        // wait. Results are identical either way.
        uint32_t pf_node0 = n, pf_node1 = n;
        double pf_row0 = 0.0, pf_row1 = 0.0, miss_row = 0.0;
        uint32_t row_src = 2;   // 0: pf_row0, 1: pf_row1, 2: miss_row
        // Inputs of the next step, read as soon as its city is known (right
        // after selection, overlapping the relocation): candidate row,
        // unvisited flags (only the just selected city changes state before
        // the next step, and a city is never its own candidate) and parent
        // links (LDS).
        uint32_t cur_cand = n, cur_link_parent = 0;
        bool cur_free = false;
        // ballot of cur_free, taken where the flags are read (as ballots of
        // single compares: a ballot of a combined condition goes through a
        // vector register)
        uint64_t cur_mask = 0;
        const float margin32 = static_cast<float>(p.roulette_margin);
        if (!explicit_mode) {
            const uint32_t c = s_nbr[current * K + row_lane];
            const uint8_t fl = flags[c < n ? c : 0u];
            cur_cand = in_row ? c : n;
            cur_free = in_row && c < n && !(fl & kVisited);
            cur_mask = ballot(c < n) & ballot(!(fl & kVisited)) & row_bits;
            miss_row = products[static_cast<size_t>(current) * pw + row_lane];
            cur_link_parent = uniform(s_plinks[current]);
        }
        // Roulette draws come in pools of 64 computed lane-parallel on the
        // vector unit: a pool drawn at step s holds the uniform of step s + l
        // in lane l (one Philox block per step, as step_uniform), so the
        // serial scalar Philox of every step becomes one vector Philox per
        // 64 steps.
        double u_pool = 0.0;
        float u_pool32 = 0.0f;
        uint32_t u_left = 0;   // uniforms of the pool not yet used
        // Escape mode: the gates of the same 64 steps, from the first word
        // of each step's escape stream computed in lane l for step s + l
        // (bit l set when that word opens the gate); an open gate rebuilds
        // its step's stream and skips the gate word.
        uint64_t gate_pool = 0;
        while (mne < target && steps + 1 < n) {
            if constexpr (Profile) mark = clock64();
            const uint32_t step = steps + 1;
            uint32_t selected;
            double u = 0.0;
            uint32_t link;
            bool novel = false;   // Escape mode: the move may open new edges
            if (explicit_mode) {
                selected = uniform(visits[step]);
                link = links[current];
            } else {
                const double prod = row_src == 0 ? pf_row0 : (row_src == 1 ? pf_row1 : miss_row);
                link = cur_link_parent;
                pf_node0 = link & 0xffffu;
                pf_node1 = link >> 16;
                pf_row0 = products[static_cast<size_t>(pf_node0) * pw + row_lane];
                pf_row1 = products[static_cast<size_t>(pf_node1) * pw + row_lane];
                if (u_left == 0) {
                    u_left = kWave;
                    // The key passes through an empty asm so the compiler
                    // derives the Philox round keys here, every 64 steps,
                    // instead of keeping them live in scalar registers
                    // across the whole construction loop.
                    uint32_t k_lo = static_cast<uint32_t>(key), k_hi = static_cast<uint32_t>(key >> 32);
                    asm volatile("" : "+s"(k_lo), "+s"(k_hi));
                    u_pool = step_uniform((static_cast<uint64_t>(k_hi) << 32) | k_lo, batch, local_ant, step + lane);
                    u_pool32 = static_cast<float>(u_pool);
                    if constexpr (Escape) {
                        uint32_t w[4] = {0u, step + lane, 0x40000000u | local_ant, batch};
                        philox4x32_10(w, k_lo, k_hi);
                        gate_pool = ballot((w[0] & 15u) == 0u);
                    }
                }
                const uint32_t u_index = kWave - u_left;   // lane of this step's uniform
                --u_left;
                uint32_t cand = cur_cand;
                uint32_t cur_after = 0;
                if constexpr (Hard) cur_after = linked ? (cur_link >> 16) : tour[pc + 1 == n ? 0 : pc + 1];
                double weight = prod;
                bool avail;
                uint32_t source = kNoSource;
                if constexpr (Escape) {
                    // the escape gate (one word in 16 of the ant's escape
                    // stream of this step) opens a slot view of the row:
                    // a replaced slot weighs default_trail * d^-beta (the
                    // backup rows never hold primary cities, so the
                    // reference's pheromone lookup gives the default trail)
                    // and may open new edges (cuda/escape_device.cuh:41-80)
                    bool free_city = cur_free;
                    const bool gated = !p.escape_disabled && ((gate_pool >> u_index) & 1ull);
                    if (gated) {
                        EscapeStream gate(key, batch, local_ant, step, false);
                        gate.draw();
                        ++e_gates;
                        const uint32_t bval = lane < p.bw ? backup[static_cast<size_t>(current) * p.bw + lane] : n;
                        const LaneView view = lane_slot_view(cur_cand, pw, bval, p.bw, n, current, gate, esc_scratch);
                        e_replaced += view.replacements;
                        source = view.source;
                        if (source != kNoSource) {
                            cand = view.member;
                            free_city = !(flags[cand] & kVisited);
                            weight = dtrail * eta_backup[static_cast<size_t>(current) * p.bw + source];
                        }
                    }
                    // with no gate (no lane may open edges) and an empty
                    // cache the test is the Hard one
                    uint32_t status = 1u;
                    if (!gated && esc_size == 0) {
                        status = free_city ? (relocation_ok(cand, current, cur_after) ? 0u : 2u) : 1u;
                    } else {
                        uint32_t novel[3], count;
                        status = free_city ? esc_relocation(cand, current, cur_after, source != kNoSource, novel, count) : 1u;
                    }
                    e_construction_full += static_cast<uint32_t>(__builtin_popcountll(ballot(free_city && status == 3u)));
                    avail = free_city && status == 0u;
                } else {
                    avail = cur_free && (!Hard || relocation_ok(cand, current, cur_after));
                }
                const uint64_t amask = (Hard || Escape) ? ballot(avail) : cur_mask;
                if (amask) {
                    // Roulette of select_next: the first available slot j with
                    // u * total < prefix_j, where total and the prefixes are
                    // the exact left-to-right FP64 sums of the available
                    // weights (T = u * total; last available slot if none).
                    // Fast path (pw <= 16): FP32 doubling-scan prefixes Q_j of
                    // the weights rounded to FP32. With S the exact total:
                    // the sequential FP64 prefixes and T lie within 2^-48 S
                    // of the exact prefixes and u S; each Q_j within
                    // (2^-24 + 4 * 2^-24) S, and T' = fl32(fl32(u) * Q_15)
                    // within 2^-21.3 S of u S; so T - P_j and T' - Q_j differ
                    // by less than 2^-20.4 S. When every available slot has
                    // |T' - Q_j| above the margin (2^-18 Q_15 unless a test
                    // widens it; Q_15 >= S (1 - 2^-21)), each comparison and
                    // so the selection equal the sequential FP64 ones.
                    // Otherwise (about one step in 10^4), or when the FP32
                    // total leaves [2^-100, 2^100] (underflowing or huge
                    // weights), the sequential FP64 sums decide (exact_lane).
                    auto exact_lane = [&]() -> uint32_t {
                        const double uu = read_lane(u_pool, u_index);
                        double acc = 0.0, total = 0.0;
                        if (pw <= 16) {
                            acc = row_prefix_sum(avail ? weight : 0.0);
                            total = read_lane(acc, 15);
                        } else {
                            for (uint64_t m = amask; m; m &= m - 1) {
                                const uint32_t i = ctz64(m);
                                const double v = read_lane(weight, i);
                                acc = (i <= lane) ? acc + v : acc;
                            }
                            total = read_lane(acc, 63);
                        }
                        if (total > 0.0) {
                            const double threshold = uu * total;
                            const uint64_t hit = ballot(avail && threshold < acc);
                            return hit ? ctz64(hit) : last64(amask);
                        }
                        const uint32_t count = static_cast<uint32_t>(__builtin_popcountll(amask));
                        uint32_t k = static_cast<uint32_t>(uu * count);
                        if (k >= count) k = count - 1;
                        return nth_set_bit(amask, k);
                    };
                    uint32_t sel_lane = kWave;
                    if (fresh(pw) <= 16u) {
                        const float w32 = avail ? static_cast<float>(weight) : 0.0f;
                        const float fast = row_scan4_f32(w32);
                        const float fast_total = read_lane(fast, 15);
                        if (fast_total >= 0x1p-100f && fast_total <= 0x1p100f) {
                            const float threshold = read_lane(u_pool32, u_index) * fast_total;
                            const float diff = threshold - fast;
                            const float margin = fast_total * margin32;
                            if (!(ballot(__builtin_fabsf(diff) <= margin) & amask)) {
                                const uint64_t hit = ballot(threshold < fast) & amask;
                                sel_lane = hit ? ctz64(hit) : last64(amask);
                            }
                        }
                    }
                    if (sel_lane == kWave) sel_lane = uniform(exact_lane());
                    selected = read_lane(cand, sel_lane);
                    // Escape mode: the chosen slot also tells whether the move
                    // may open new edges
                    if constexpr (Escape) novel = read_lane(source, sel_lane) != kNoSource;
                } else {
                    selected = n;
                    for (uint32_t base = 0; base < p.bw && selected == n; base += kWave) {
                        const uint32_t j = base + lane;
                        const uint32_t b = j < p.bw ? backup[static_cast<size_t>(current) * p.bw + j] : n;
                        const uint64_t bm = ballot(b < n && !(flags[b] & kVisited) &&
                                                   (!Hard || reloc_ok(b, current, cur_after)));
                        if (bm) selected = uniform(read_lane(b, ctz64(bm)));
                    }
                    if (selected == n) {
                        // Nearest unvisited city, lowest id on ties.
                        const uint32_t cxy = s_xy[current];
                        uint64_t best = ~0ull;
                        for (uint32_t v = lane; v < n; v += kWave) {
                            if (!(flags[v] & kVisited) && (!Hard || reloc_ok(v, current, cur_after))) {
                                const uint64_t kv = (static_cast<uint64_t>(euc2d(cxy, s_xy[v])) << 32) | v;
                                best = kv < best ? kv : best;
                            }
                        }
                        best = wave_min_u64(best);
                        selected = uniform(best == ~0ull ? n : static_cast<uint32_t>(best));
                    }
                }
                selected = uniform(selected);
                if constexpr (Escape) {
                    // the move's new edges join the ant's cache (a move through
                    // a slot that kept its own city was checked without
                    // permission for new edges, so it has none)
                    if (selected < n && novel) {
                        uint32_t novel_edges[3], count;
                        (void)esc_relocation(selected, current, cur_after, novel, novel_edges, count);
                        esc_commit(novel_edges, count);
                    }
                }
                if (selected < n) {
                    const uint32_t c = s_nbr[selected * K + row_lane];
                    const uint8_t fl = flags[c < n ? c : 0u];
                    const bool row_lane_now = lane < fresh(pw);   // = in_row
                    cur_cand = row_lane_now ? c : n;
                    cur_free = row_lane_now && c < n && !(fl & kVisited);
                    cur_mask = ballot(c < n) & ballot(!(fl & kVisited)) & row_bits;
                    cur_link_parent = uniform(s_plinks[selected]);
                    row_src = selected == pf_node0 ? 0u : (selected == pf_node1 ? 1u : 2u);
                    if (row_src == 2u) miss_row = products[static_cast<size_t>(selected) * pw + row_lane];
                }
            }
            if (trace_selected) {
                if (!explicit_mode) u = read_lane(u_pool, kWave - 1u - u_left);
                if (lane == 0) {
                    trace_selected[static_cast<size_t>(ant) * ns + step] = selected;
                    trace_uniform[static_cast<size_t>(ant) * ns + step] = u;
                }
            }
            if constexpr (Profile) { const uint64_t t = clock64(); sub[0] += t - mark; mark = t; }
            if (selected >= n) { exhausted = 1; break; }

            ++steps;
            uint32_t ps = 0, old_prev, after, target_after;
            const bool linked_now = fresh(n) <= kMaxLinked;   // = linked
            if (linked_now) {
                const uint32_t sel_link = uniform(lnk[selected]);
                old_prev = sel_link & 0xffffu;
                after = sel_link >> 16;
                target_after = cur_link >> 16;
            } else {
                ps = uniform(pos[selected]);
                old_prev = uniform(tour[ps == 0 ? n - 1 : ps - 1]);
                after = uniform(tour[ps + 1 == n ? 0 : ps + 1]);
                target_after = uniform(tour[pc + 1 == n ? 0 : pc + 1]);
            }
            uint32_t next_pc = ps;
            const bool moved = target_after != selected;
            const bool record = logged + steps < fresh(record_limit);
            // Lane 0's LDS writes of the step, in one region: the visit
            // record, the visited flag, Escape anchors and, in linked mode,
            // the relinking: unlink `selected` and relink it between
            // `current` and its successor.
            if (lane == 0) {
                if (record) queue[fresh(n) - 1u - steps] = static_cast<uint16_t>(selected);
                flags[selected] = static_cast<uint8_t>(flags[selected] | kVisited);
                if constexpr (Escape) {
                    // a move that may open new edges marks its five endpoints
                    // as anchors (cuda/faco_device.cuh:213-218)
                    if (moved && novel) {
                        flags[current] = static_cast<uint8_t>(flags[current] | kAnchor);
                        flags[target_after] = static_cast<uint8_t>(flags[target_after] | kAnchor);
                        flags[selected] = static_cast<uint8_t>(flags[selected] | kAnchor);
                        flags[old_prev] = static_cast<uint8_t>(flags[old_prev] | kAnchor);
                        flags[after] = static_cast<uint8_t>(flags[after] | kAnchor);
                    }
                }
                if (linked_now && moved) {
                    uint16_t* half = reinterpret_cast<uint16_t*>(lnk);
                    half[2 * old_prev + 1] = static_cast<uint16_t>(after);
                    half[2 * after] = static_cast<uint16_t>(old_prev);
                    half[2 * current + 1] = static_cast<uint16_t>(selected);
                    lnk[selected] = current | (target_after << 16);
                    half[2 * target_after] = static_cast<uint16_t>(selected);
                }
            }
            cur_link = current | (after << 16);   // selected's links when the move is the identity
            if (moved) {
                ++relocations;
                if constexpr (Escape) {
                    if (novel) ++e_relocations;
                }
                if (linked_now) {
                    // the array's first city moves on when it is the one
                    // relocated (the array shift starts at position 0)
                    if (selected == head) head = after;
                    cur_link = current | (target_after << 16);
                }
                // Move `selected` to directly after `current`; only the closed
                // position range between the two shifts by one place. Four
                // 64-element chunks are read before any of them is written.
                else if (pc < ps) {
                    const int32_t lowest = static_cast<int32_t>(pc) + 2;
                    for (int32_t hi = static_cast<int32_t>(ps); hi >= lowest; hi -= 4 * static_cast<int32_t>(kWave)) {
                        uint32_t v[4];
                        int32_t idx[4];
#pragma unroll
                        for (int c = 0; c < 4; ++c) {
                            idx[c] = hi - c * static_cast<int32_t>(kWave) - static_cast<int32_t>(lane);
                            v[c] = idx[c] >= lowest ? tour[idx[c] - 1] : 0u;
                        }
#pragma unroll
                        for (int c = 0; c < 4; ++c)
                            if (idx[c] >= lowest) { tour[idx[c]] = static_cast<uint16_t>(v[c]); pos[v[c]] = static_cast<uint16_t>(idx[c]); }
                    }
                    if (lane == 0) { tour[pc + 1] = static_cast<uint16_t>(selected); pos[selected] = static_cast<uint16_t>(pc + 1); }
                    next_pc = pc + 1;
                } else {
                    for (uint32_t lo = ps; lo < pc; lo += 4 * kWave) {
                        uint32_t v[4], idx[4];
#pragma unroll
                        for (int c = 0; c < 4; ++c) {
                            idx[c] = lo + c * kWave + lane;
                            v[c] = idx[c] < pc ? tour[idx[c] + 1] : 0u;
                        }
#pragma unroll
                        for (int c = 0; c < 4; ++c)
                            if (idx[c] < pc) { tour[idx[c]] = static_cast<uint16_t>(v[c]); pos[v[c]] = static_cast<uint16_t>(idx[c]); }
                    }
                    if (lane == 0) { tour[pc] = static_cast<uint16_t>(selected); pos[selected] = static_cast<uint16_t>(pc); }
                    next_pc = pc;
                }
            }
            pc = next_pc;
            wave_sync();
            if constexpr (Profile) { const uint64_t t = clock64(); sub[1] += t - mark; mark = t; }
            if (selected != (link & 0xffffu) && selected != (link >> 16)) {
                ++mne;
                // old_prev is current itself when selected already followed
                // it in the tour: it is then pushed once, as current
                const uint32_t node = lane == 0 ? current : (lane == 1 ? selected : old_prev);
                push_distinct(node, old_prev == current ? 2u : 3u);
            }
            current = selected;
            if constexpr (Profile) sub[2] += clock64() - mark;
        }

        // Length of the constructed tour: the sum of its edges, one per city
        // and lane (the reference updates it per relocation; the integer sum
        // is the same).
        uint32_t edge_sum = 0;
        if (linked) {
            // Edge lengths from the final links (each city and its predecessor).
            constexpr uint32_t kPer = kRankedInRegisters / kWave;
            const uint32_t chunks = (n + kWave - 1) / kWave;
#pragma unroll
            for (uint32_t k = 0; k < kPer; ++k) {
                const uint32_t v = lane + k * kWave;
                if (k < chunks && v < n) edge_sum += euc2d(s_xy[lnk[v] & 0xffffu], s_xy[v]);
            }
            for (uint32_t v = kRankedInRegisters + lane; v < n; v += kWave) edge_sum += euc2d(s_xy[lnk[v] & 0xffffu], s_xy[v]);
            wave_sync();
            if (chain_rec && logged + steps < n) {
                // The working tour is the visit chain [start, s1, .., sk]
                // followed by the unvisited cities in the parent's cyclic order
                // from start on (a relocation keeps the chain contiguous and
                // the others in their relative order), and the array is that
                // cycle read from `head`. One pass over the parent positions
                // places every city: from head's position when head is
                // unvisited, else from start's, shifted by head's chain index;
                // the whole chain lands where start lies in the parent.
                const uint32_t chain_len = steps + 1;
                uint32_t shift = 0, from;
                if (flags[head] & kVisited) {
                    for (uint32_t i0 = 0; i0 < chain_len; i0 += kWave) {
                        const uint32_t i = i0 + lane;
                        const uint64_t hit = ballot(i < chain_len && queue[n - 1 - (i < chain_len ? i : 0u)] == head);
                        if (hit) { shift = i0 + ctz64(hit); break; }
                    }
                    from = s_ppos[start];
                } else {
                    from = s_ppos[head];
                }
                from = uniform(from);
                uint32_t placed = 0, extra = 0;   // cities placed so far; chain_len once start is passed
                for (uint32_t t0 = 0; t0 < n; t0 += kWave) {
                    const uint32_t t = t0 + lane;
                    const bool valid = t < n;
                    uint32_t at = from + (valid ? t : 0u);
                    at = at >= n ? at - n : at;
                    const uint32_t c = parent[at];
                    const bool free = valid && !(flags[c] & kVisited);
                    const uint64_t fm = ballot(free);
                    const uint64_t sm = ballot(valid && c == start);
                    const uint32_t rank = __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(fm >> 32),
                                          __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(fm), 0u));
                    uint32_t lane_extra = extra;
                    if (sm) {
                        const uint32_t ls = ctz64(sm);
                        const uint32_t below = static_cast<uint32_t>(__builtin_popcountll(fm & ((1ull << ls) - 1ull)));
                        const uint32_t first = placed + below + n - shift;   // chain position of start, shifted
                        for (uint32_t i = lane; i < chain_len; i += kWave) {
                            const uint32_t cc = queue[n - 1 - i];
                            uint32_t d = first + i;
                            d = d >= n ? d - n : d;
                            tour[d] = static_cast<uint16_t>(cc);
                            pos[cc] = static_cast<uint16_t>(d);
                        }
                        lane_extra = lane > ls ? chain_len : extra;
                        extra = chain_len;
                    }
                    if (free) {
                        uint32_t d = placed + rank + lane_extra + n - shift;
                        d = d >= n ? d - n : d;
                        tour[d] = static_cast<uint16_t>(c);
                        pos[c] = static_cast<uint16_t>(d);
                    }
                    placed += static_cast<uint32_t>(__builtin_popcountll(fm));
                }
                wave_sync();
            } else if (n <= kRankedInRegisters) {
                // Array position of every city = its distance from `head`
                // along predecessor links: list ranking by pointer jumping,
                // each lane holding (distance | link << 16) for cities
                // lane + 64k. An entry is one packed word, so every update
                // keeps distance(v) = hops from v to link(v); reading an entry
                // already advanced in this round only jumps further, and the
                // ranks are exact once every link reaches `head`.
                uint32_t e[kPer];
#pragma unroll
                for (uint32_t k = 0; k < kPer; ++k) {
                    const uint32_t v = lane + k * kWave;
                    if (k < chunks && v < n) e[k] = v == head ? (head << 16) : (1u | ((lnk[v] & 0xffffu) << 16));
                }
                wave_sync();
#pragma unroll
                for (uint32_t k = 0; k < kPer; ++k) {
                    const uint32_t v = lane + k * kWave;
                    if (k < chunks && v < n) lnk[v] = e[k];
                }
                wave_sync();
                for (uint32_t round = 0; round < 17; ++round) {
                    bool pending = false;
#pragma unroll
                    for (uint32_t k0 = 0; k0 < kPer; k0 += 8) {
                        if (k0 >= chunks) break;
                        uint32_t t[8];
#pragma unroll
                        for (uint32_t k = k0; k < k0 + 8; ++k) {
                            const uint32_t v = lane + k * kWave;
                            t[k - k0] = (k < chunks && v < n) ? lnk[e[k] >> 16] : 0u;
                        }
#pragma unroll
                        for (uint32_t k = k0; k < k0 + 8; ++k) {
                            const uint32_t v = lane + k * kWave;
                            if (k < chunks && v < n) {
                                e[k] = ((e[k] & 0xffffu) + (t[k - k0] & 0xffffu)) | (t[k - k0] & 0xffff0000u);
                                lnk[v] = e[k];
                                pending |= (e[k] >> 16) != head;
                            }
                        }
                    }
                    wave_sync();
                    if (!ballot(pending)) break;
                }
#pragma unroll
                for (uint32_t k = 0; k < kPer; ++k) {
                    const uint32_t v = lane + k * kWave;
                    if (k < chunks && v < n) {
                        const uint32_t d = e[k] & 0xffffu;
                        tour[d] = static_cast<uint16_t>(v);
                        pos[v] = static_cast<uint16_t>(d);
                    }
                }
                wave_sync();
            } else {
                // The same ranking for n above kRankedInRegisters: pointer
                // jumping in place in LDS (every entry is one packed word, so
                // the argument above holds), then the ranks are staged in the
                // ant's tour row in HBM, which the final tour overwrites,
                // before the array is written over the links.
                for (uint32_t v = lane; v < n; v += kWave) {
                    const uint32_t l = lnk[v];
                    lnk[v] = v == head ? (head << 16) : (1u | ((l & 0xffffu) << 16));
                }
                wave_sync();
                for (uint32_t round = 0; round < 17; ++round) {
                    bool pending = false;
                    for (uint32_t v = lane; v < n; v += kWave) {
                        const uint32_t e = lnk[v];
                        const uint32_t t = lnk[e >> 16];
                        const uint32_t jumped = ((e & 0xffffu) + (t & 0xffffu)) | (t & 0xffff0000u);
                        lnk[v] = jumped;
                        pending |= (jumped >> 16) != head;
                    }
                    wave_sync();
                    if (!ballot(pending)) break;
                }
                uint32_t* const stage = p.out_tours + static_cast<size_t>(ant) * ns;
                for (uint32_t v = lane; v < n; v += kWave) stage[v] = lnk[v] & 0xffffu;
                __threadfence_block();
                wave_sync();
                for (uint32_t v = lane; v < n; v += kWave) {
                    const uint32_t d = stage[v];
                    tour[d] = static_cast<uint16_t>(v);
                    pos[v] = static_cast<uint16_t>(d);
                }
                wave_sync();
            }
        } else {
            for (uint32_t i = lane; i < n; i += kWave) edge_sum += euc2d(s_xy[tour[i]], s_xy[tour[i + 1 == n ? 0 : i + 1]]);
        }
#pragma unroll
        for (int offset = 32; offset; offset >>= 1) edge_sum += __shfl_xor(edge_sum, offset);
        const int64_t cost = uniform(edge_sum);
        if (diag_construction) {
            for (uint32_t i = lane; i < n; i += kWave) diag_construction[static_cast<size_t>(ant) * ns + i] = tour[i];
        }
        const int64_t construction_cost = cost;
        if (Diag && p.full_checklist) {
            // every city in tour order, as the reference's initial 2-opt of
            // the prepared tour starts (cpp/src/prepared_problem.cpp:112-113,
            // cpp/src/faco_cpu.cpp:148-160); the circular queue is full
            for (uint32_t i = lane; i < n; i += kWave) {
                const uint32_t v = tour[i];
                queue[i] = static_cast<uint16_t>(v);
                flags[v] = kQueued;
            }
            qhead = 0; qtail = 0; qsize = n;
            wave_sync();
        }
        // Escape mode: each anchor's LS row becomes its escape view, sorted
        // by (distance, city) (escape stream of the city in the LS domain),
        // and the footprint marks the rows with replaced slots
        // (cuda/escape_device.cuh:111-126); then the parent's footprint and
        // the anchors put their cities on the checklist, in city order
        // (cuda/faco_device.cuh:270-291).
        uint16_t* const my_rows = Escape ? p.esc_rows + static_cast<size_t>(ant) * ns * lw : nullptr;
        if constexpr (Escape) {
            uint8_t* my_fp = p.esc_footprint + static_cast<size_t>(ant) * ns;
            for (uint32_t v0 = 0; v0 < n; v0 += kWave) {
                const uint32_t v = v0 + lane, vi = v < n ? v : 0u;
                const bool anchor = v < n && (flags[vi] & kAnchor);
                uint64_t am = ballot(anchor);
                e_anchor_nodes += static_cast<uint32_t>(__builtin_popcountll(am));
                if (v < n && !anchor) my_fp[v] = 0;
                while (am) {
                    const uint32_t node = v0 + ctz64(am);
                    am &= am - 1;
                    EscapeStream rows_rng(key, batch, local_ant, node, true);
                    const uint32_t orig = lane < lw ? s_nbr[node * K + p.ls_offset + lane] : n;
                    const uint32_t bval = lane < p.bw ? backup[static_cast<size_t>(node) * p.bw + lane] : n;
                    const LaneView view = lane_slot_view(orig, lw, bval, p.bw, n, node, rows_rng, esc_scratch);
                    uint32_t member = view.member, src = view.source;
                    lane_sort_row(member, src, view.valid, node, s_xy, esc_order);
                    if (lane < lw)
                        my_rows[static_cast<size_t>(node) * lw + lane] = static_cast<uint16_t>(
                            lane < view.valid ? (member | (src != kNoSource ? 0x8000u : 0u)) : n);
                    e_ls_replaced += view.replacements;
                    if (lane == 0) my_fp[node] = view.replacements ? 1 : 0;
                }
            }
            const uint8_t* pfp = p.parent_fp + static_cast<size_t>(colony) * ns;
            for (uint32_t v0 = 0; v0 < n; v0 += kWave) {
                const uint32_t v = v0 + lane, vi = v < n ? v : 0u;
                const uint8_t f = v < n ? flags[vi] : 0;
                const bool old_view = v < n && pfp[vi] != 0;
                const bool anchor = (f & kAnchor) != 0, queued = (f & kQueued) != 0;
                const bool fresh = (old_view || anchor) && !queued;
                const uint64_t fm = ballot(fresh);
                e_old_views += static_cast<uint32_t>(__builtin_popcountll(ballot(old_view && !queued)));
                e_anchor_pushes += static_cast<uint32_t>(__builtin_popcountll(ballot(anchor && !old_view && !queued)));
                if (fresh) {
                    const uint32_t rank = __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(fm >> 32),
                                          __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(fm), 0u));
                    flags[v] = static_cast<uint8_t>(f | kQueued);
                    queue[(qtail + rank) % n] = static_cast<uint16_t>(v);
                }
                const uint32_t pushed = static_cast<uint32_t>(__builtin_popcountll(fm));
                qtail = (qtail + pushed) % n;
                qsize += pushed;
                logged += pushed;
                wave_sync();
            }
        }
        const uint64_t tick2 = Profile ? wall_clock64() : 0;

        // ------------------------------------------------------------ checklist 2-opt
        // Lanes [0, lw) scan successor moves (kind 0), lanes [32, 32 + lw)
        // predecessor moves (kind 1); slot j of kind k has tie index k*lw + j.
        const uint32_t kind = lane >> 5, cslot = lane & 31u;
        const uint32_t kind_mask = 0u - kind;   // all ones in kind 1 lanes
        // Lane constants of the branch-free scan: an in-row column for every
        // lane, the step to a neighbour's tour neighbour (+1 successor, -1
        // predecessor) and where a step off either array end lands.
        // (every lane reads inside the LS row, which starts at ls_offset)
        const uint32_t column = p.ls_offset + (cslot < lw ? cslot : lw - 1u);
        const bool scan_lane = cslot < lw;
        const uint32_t off_scan = scan_lane ? 0u : 0xffffffffu;   // all ones outside the scan
        const uint16_t* g_nd = p.nbr_dist + static_cast<size_t>(geom) * p.nbr_stride;
        const uint32_t step_dir = kind ? 0xffffffffu : 1u;
        const uint32_t wrap_to = kind ? n - 1u : 0u;
        const uint32_t tie_index = kind * lw + cslot;
        // Reverse the tour between cities u and v (2-opt flip on the array,
        // reference cpu flip_without_cost): as disjoint pair swaps of the
        // shorter side (the complement side wraps round the array end);
        // first == 0 keeps the straight side, which is the array layout the
        // reference's extra lap produces.
        auto flip_nodes = [&](uint32_t u, uint32_t v) {
            const uint32_t p0 = pos[u], p1 = pos[v];
            const uint32_t first = p0 < p1 ? p0 : p1, last = p0 < p1 ? p1 : p0;
            const uint32_t length = last - first;
            const bool straight = length <= n - length || first == 0;
            const uint32_t count = straight ? length : n - length;
            const uint32_t begin = straight ? first : last;
            for (uint32_t off = lane; off < count / 2; off += kWave) {
                uint32_t left = begin + off; if (left >= n) left -= n;
                uint32_t right = begin + count - 1 - off; if (right >= n) right -= n;
                const uint16_t l = tour[left], r = tour[right];
                tour[left] = r; tour[right] = l;
                pos[r] = static_cast<uint16_t>(left); pos[l] = static_cast<uint16_t>(right);
            }
            wave_sync();
        };
        // Remove edges (x1, x2), (y1, y2), add (x1, y1), (x2, y2), where x2
        // follows x1 and y2 follows y1 in one orientation of the tour.
        auto two_opt_move = [&](uint32_t x1, uint32_t x2, uint32_t y1, uint32_t y2) {
            const uint32_t px = pos[x1];
            const uint32_t succ_x1 = uniform(tour[px + 1 == n ? 0 : px + 1]);
            if (succ_x1 == x2) flip_nodes(x2, y2);
            else flip_nodes(x1, y1);
        };
        uint32_t processed = 0, checks = 0, evaluations = 0, accepted = 0, reactivations = 0, limit_hit = 0;
        uint32_t rejections = 0;
        int64_t gain_sum = 0;
        // The reference's loop condition (queue not empty and fewer than n
        // accepted moves) with the move count tested where it changes.
        while (qsize != 0) {
            if (evaluations >= limit) { limit_hit = 1; break; }
            if constexpr (Profile) mark = clock64();
            const uint32_t a = queue[qhead];
            qhead = qhead + 1 == n ? 0u : qhead + 1;
            --qsize;
            // only the queued bit matters after construction: clear the byte
            // (Escape mode keeps the anchor bit: an anchor reads its escape row)
            bool a_anchor = false;
            if constexpr (Escape) {
                const uint8_t fa = flags[a];
                a_anchor = (fa & kAnchor) != 0;
                wave_sync();
                if (lane == 0) flags[a] = static_cast<uint8_t>(fa & ~kQueued);
            } else {
                if (lane == 0) flags[a] = 0;
            }
            ++processed;
            // every lane reads (in-bounds addresses, no exec masks); lanes
            // outside the scan or past the row end see b = n and count as
            // failing, their other values are not used
            uint32_t row_b = s_nbr[a * K + column];
            uint32_t permit = 0;   // Escape mode: the slot was replaced, its move may open new edges
            if constexpr (Escape) {
                if (a_anchor) {
                    const uint32_t e = my_rows[static_cast<size_t>(a) * lw + (cslot < lw ? cslot : lw - 1u)];
                    row_b = e & 0x7fffu;
                    permit = e >> 15;
                }
            }
            const uint32_t b = min(row_b | off_scan, n);   // row entries are at most n
            const bool real = b < n;
            const uint32_t bi = real ? b : a;
            const uint32_t pa = pos[a];
            const uint32_t xa = s_xy[a];
            const uint32_t xb = s_xy[bi];
            const uint32_t pb = pos[bi];
            const uint32_t a_next = tour[pa + 1 == n ? 0 : pa + 1];
            const uint32_t a_prev = tour[pa == 0 ? n - 1 : pa - 1];
            // kind 0 lanes hold d(a, succ a), kind 1 lanes d(pred a, a)
            const uint32_t closing = (a_prev & kind_mask) | (a_next & ~kind_mask);
            const uint32_t xclosing = s_xy[closing];
            const int32_t current_edge = static_cast<int32_t>(euc2d(xa, xclosing));
            const int32_t dab = (Escape && a_anchor) ? static_cast<int32_t>(euc2d(xa, xb)) : static_cast<int32_t>(g_nd[a * K + column]);
            const bool bad = !real || !(current_edge > dab);
            // First failing slot per kind; lanes past lw count as failing.
            const uint64_t bad_mask = ballot(bad);
            const uint32_t bad0 = static_cast<uint32_t>(bad_mask), bad1 = static_cast<uint32_t>(bad_mask >> 32);
            // lanes past lw always fail, so the first failing slot is at most lw
            // (the extra bit 32 covers lw = 32)
            const uint32_t prefix0 = ctz64(static_cast<uint64_t>(bad0) | (1ull << 32));
            const uint32_t prefix1 = ctz64(static_cast<uint64_t>(bad1) | (1ull << 32));
            // Budget: kind 0 first, then kind 1; the limit check precedes every
            // evaluation and a node cut short commits nothing.
            // (Hard mode: a node cut short still counts the rejections of the
            // moves it evaluated, as the reference's sequential scan does,
            // cuda/faco_device.cuh:316-336; it stops after counting them)
            bool cut = false;
            const uint32_t used0 = min(limit - evaluations, prefix0);
            evaluations += used0;
            checks += min(used0 + 1u, lw);   // the failing slot counts as checked
            uint32_t used1 = 0;
            if (used0 < prefix0) {
                limit_hit = 1;
                if constexpr (!Hard) break;
                cut = true;
            } else {
                used1 = min(limit - evaluations, prefix1);
                evaluations += used1;
                checks += min(used1 + 1u, lw);
                if (used1 < prefix1) {
                    limit_hit = 1;
                    if constexpr (!Hard) break;
                    cut = true;
                }
            }
            if (!cut && used0 + used1 == 0) continue;

            // evaluated slots of this lane's kind (lanes past lw never qualify)
            const bool active = cslot < ((used1 & kind_mask) | (used0 & ~kind_mask));
            uint32_t nidx = pb + step_dir;
            nidx = nidx >= n ? wrap_to : nidx;
            const uint32_t nb = tour[nidx];
            const uint32_t xn = s_xy[nb];
            const int32_t move_gain = current_edge + static_cast<int32_t>(euc2d(xb, xn)) - dab
                                    - static_cast<int32_t>(euc2d(xclosing, xn));
            const int32_t gain = (nb == a || b == closing) ? 0 : move_gain;
            // key: larger gain first, then lower tie index; 0 = no candidate.
            // Hard mode: a 2-opt move that leaves the graph never wins
            // (reference cuda/faco_device.cuh:383-394)
            bool move_allowed = true;
            if constexpr (Hard) {
                uint32_t status = 0;   // Escape mode: 3 = the cache is full
                if (active && nb != a && b != closing) {
                    const uint32_t removed[3][2] = {{a, closing}, {b, nb}, {0u, 0u}};
                    const uint32_t added[3][2] = {{a, b}, {closing, nb}, {0u, 0u}};
                    if constexpr (Escape) {
                        // graph or cache edges, or new ones through a replaced
                        // slot while the cache has room (escape_edges.hpp:117-122)
                        // (a non-anchor row has no replaced slots: with an
                        // empty cache the test is the Hard one)
                        if (!a_anchor && esc_size == 0) {
                            move_allowed = edges_allowed(2, removed, added);
                        } else {
                            uint32_t novel[3], count;
                            status = esc_status(2, removed, added, permit != 0, novel, count);
                            move_allowed = status == 0u;
                        }
                    } else {
                        move_allowed = edges_allowed(2, removed, added);
                    }
                }
                // every evaluated move that would leave the graph counts
                // (reference cpp/src/faco_cpu.cpp:180-183)
                rejections += static_cast<uint32_t>(__builtin_popcountll(ballot(active && !move_allowed)));
                if constexpr (Escape) e_ls_full += static_cast<uint32_t>(__builtin_popcountll(ballot(status == 3u)));
                if (cut) break;
            }
            const uint32_t key_value = active && move_allowed
                ? ((static_cast<uint32_t>(gain + (1 << 20)) << 6) | (63u - tie_index)) : 0u;
            const uint32_t best = uniform(dpp_max_u32(key_value));
            int32_t best_gain = best ? static_cast<int32_t>(best >> 6) - (1 << 20) : -1;

            // Or-opt (this engine's extension): the segment a, succ a, ...
            // of length L = 1..3 reinserted next to a predecessor-scan
            // neighbour b, forward at (b, succ b) or reversed at (pred b, b);
            // one lane per b, six moves per lane in (L, forward first) order.
            uint32_t or_best = 0u;
            uint32_t s2 = 0, s3 = 0, nx3 = 0, or_succ_b = 0;
            if (OrOpt && used1 > 0) {
                s2 = a_next;
                s3 = tour[pa + 2 >= n ? pa + 2 - n : pa + 2];
                nx3 = tour[pa + 3 >= n ? pa + 3 - n : pa + 3];
                const uint32_t xs2 = s_xy[s2], xs3 = s_xy[s3], xnx3 = s_xy[nx3], xp = s_xy[a_prev];
                const int32_t dpa = static_cast<int32_t>(euc2d(xp, xa));
                const int32_t removed[3] = {
                    dpa + static_cast<int32_t>(euc2d(xa, xs2)) - static_cast<int32_t>(euc2d(xp, xs2)),
                    dpa + static_cast<int32_t>(euc2d(xs2, xs3)) - static_cast<int32_t>(euc2d(xp, xs3)),
                    dpa + static_cast<int32_t>(euc2d(xs3, xnx3)) - static_cast<int32_t>(euc2d(xp, xnx3))};
                const uint32_t seg_xy[3] = {xa, xs2, xs3};
                const bool or_lane = active && kind == 1u;
                uint32_t valid_count = 0, lane_key = 0;
                if (or_lane) {
                    or_succ_b = tour[pb + 1 == n ? 0 : pb + 1];
                    const uint32_t pbn = nb;  // predecessor of b (kind 1)
                    const uint32_t xsb = s_xy[or_succ_b], xpb = s_xy[pbn];
                    const int32_t d_bsb = static_cast<int32_t>(euc2d(xb, xsb));
                    const int32_t d_pbb = static_cast<int32_t>(euc2d(xpb, xb));
                    int32_t lane_best = 0;
                    uint32_t lane_k = 0;
                    bool any = false;
                    for (uint32_t L = 1; L <= 3; ++L) {
                        if (L + 4 > n) break;
                        const uint32_t nx = L == 1 ? s2 : (L == 2 ? s3 : nx3);
                        const uint32_t xsl = seg_xy[L - 1];
                        auto outside = [&](uint32_t v) {
                            return v != a_prev && v != nx && v != a && !(L >= 2 && v == s2) && !(L >= 3 && v == s3);
                        };
                        // forward: (b, a) + (last, succ b) replace (b, succ b)
                        if (outside(b) && outside(or_succ_b)) {
                            ++valid_count;
                            const int32_t g = removed[L - 1] + d_bsb - dab - static_cast<int32_t>(euc2d(xsl, xsb));
                            if (!any || g > lane_best) { lane_best = g; lane_k = (L - 1) * 2; any = true; }
                        }
                        // reversed: (pred b, last) + (a, b) replace (pred b, b)
                        if (outside(pbn) && outside(b)) {
                            ++valid_count;
                            const int32_t g = removed[L - 1] + d_pbb - static_cast<int32_t>(euc2d(xpb, xsl)) - dab;
                            if (!any || g > lane_best) { lane_best = g; lane_k = (L - 1) * 2 + 1; any = true; }
                        }
                    }
                    if (any) lane_key = (static_cast<uint32_t>(lane_best + (1 << 20)) << 8) | (255u - (cslot * 8u + lane_k));
                }
                // every Or-opt move counts one evaluation after the 2-opt ones;
                // a node cut short commits nothing
                uint32_t total = valid_count;
#pragma unroll
                for (int offset = 32; offset; offset >>= 1) total += __shfl_xor(total, offset);
                total = uniform(total);
                if (total > limit - evaluations) { evaluations = limit; limit_hit = 1; break; }
                evaluations += total;
                or_best = uniform(dpp_max_u32(lane_key));
            }
            if constexpr (Profile) { const uint64_t t = clock64(); sub[3] += t - mark; mark = t; }
            const int32_t or_gain = or_best ? static_cast<int32_t>(or_best >> 8) - (1 << 20) : -1;
            if (or_best && or_gain > 0 && or_gain > best_gain) {
                const uint32_t index = 255u - (or_best & 255u);
                const uint32_t j = index / 8u, k = index % 8u, L = k / 2 + 1;
                const bool forward = (k & 1u) == 0u;
                const uint32_t bnode = read_lane(b, 32u + j);
                const uint32_t pnode = read_lane(nb, 32u + j), snode = read_lane(or_succ_b, 32u + j);
                const uint32_t sl = L == 1 ? a : (L == 2 ? s2 : s3);
                const uint32_t nx = L == 1 ? s2 : (L == 2 ? s3 : nx3);
                const uint32_t c = forward ? bnode : pnode, d = forward ? snode : bnode;
                // two (reversed) or three (forward) 2-opt reversals
                two_opt_move(a_prev, a, c, d);
                two_opt_move(a_prev, c, nx, sl);
                if (forward) two_opt_move(c, sl, a, d);
                ++accepted;
                gain_sum += or_gain;
                reactivations += push(a_prev);
                reactivations += push(nx);
                reactivations += push(a);
                reactivations += push(sl);
                reactivations += push(c);
                reactivations += push(d);
                if constexpr (Profile) sub[4] += clock64() - mark;
                if (accepted >= n) break;
                continue;
            }
            if (best == 0u) continue;
            if (best_gain <= 0) continue;
            const uint32_t best_index = 63u - (best & 63u);
            const uint32_t best_kind = best_index >= lw ? 1u : 0u;
            const uint32_t best_lane = best_kind * 32u + (best_index - best_kind * lw);
            const uint32_t bnode = read_lane(b, best_lane);
            const uint32_t nnode = read_lane(nb, best_lane);
            uint32_t move[4];
            if (best_kind == 0) { move[0] = a_next; move[1] = nnode; move[2] = a; move[3] = bnode; }
            else { move[0] = a; move[1] = bnode; move[2] = a_prev; move[3] = nnode; }
            if constexpr (Escape) {
                // the move's new edges join the ant's cache (only a replaced
                // slot's move may have any: the others passed without
                // permission for new edges)
                if (read_lane(permit, best_lane) != 0) {
                    const uint32_t closing_node = best_kind == 0 ? a_next : a_prev;
                    const uint32_t removed[3][2] = {{a, closing_node}, {bnode, nnode}, {0u, 0u}};
                    const uint32_t added[3][2] = {{a, bnode}, {closing_node, nnode}, {0u, 0u}};
                    uint32_t novel[3], count;
                    (void)esc_status(2, removed, added, true, novel, count);
                    esc_commit(novel, count);
                }
            }
            flip_nodes(move[0], move[1]);
            ++accepted;
            gain_sum += best_gain;
            // the four endpoints of an improving 2-opt move are distinct
            reactivations += push_distinct(lane == 0 ? move[0] : (lane == 1 ? move[1] : (lane == 2 ? move[2] : move[3])), 4u);
            if constexpr (Profile) sub[4] += clock64() - mark;
            if (accepted >= n) break;
        }

        const uint64_t tick3 = Profile ? wall_clock64() : 0;
        // ------------------------------------------------------------ write back
        const int64_t final_cost = cost - gain_sum;
        // The batch's best ant (smallest (cost, ant) key) always finds its key
        // below every key published before it, so it always writes; an ant
        // whose key is above one already published cannot be the best and
        // skips the write.
        bool write_tour = true;
        if (!Diag && p.filter_tours) {
            const uint64_t my_key = (static_cast<uint64_t>(final_cost) << 32) | local_ant;
            uint64_t before = 0;
            if (lane == 0) before = atomicMin(&p.state[colony].best_key, my_key);
            const uint32_t lo = uniform(static_cast<uint32_t>(before)), hi = uniform(static_cast<uint32_t>(before >> 32));
            write_tour = my_key < ((static_cast<uint64_t>(hi) << 32) | lo);
        }
        uint32_t* out = p.out_tours + static_cast<size_t>(ant) * ns;
        // (the persistent kernel's own update reads the tour back: a plain
        // store keeps it in this CU's caches)
        if (write_tour) {
            if constexpr (Persistent) {
                for (uint32_t i = lane; i < n; i += kWave) out[i] = tour[i];
            } else {
                for (uint32_t i = lane; i < n; i += kWave) __builtin_nontemporal_store(static_cast<uint32_t>(tour[i]), out + i);
            }
        }
        if (diag_positions) {
            for (uint32_t i = lane; i < n; i += kWave) diag_positions[static_cast<size_t>(ant) * ns + i] = pos[i];
        }
        if (p.out_identity) {
            // tour identity: XOR of the edge identities, then the canonical
            // start (position of city 0 and the direction of its smaller
            // neighbour) used to compare tours in the archive update
            uint64_t f = 0, g = 0;
            for (uint32_t i = lane; i < n; i += kWave) {
                uint64_t a, b;
                edge_identity(tour[i], tour[i + 1 == n ? 0 : i + 1], a, b);
                f ^= a; g ^= b;
            }
#pragma unroll
            for (int offset = 32; offset; offset >>= 1) {
                f ^= __shfl_xor(f, offset);
                g ^= __shfl_xor(g, offset);
            }
            // equality with the batch's parent tour as a cycle (every city
            // keeps its two parent neighbours): the archive update takes it
            // as proof that two such ants hold the same tour
            bool moved = false;
            for (uint32_t v = lane; v < n; v += kWave) {
                const uint32_t pv = pos[v];
                const uint32_t succ = tour[pv + 1 == n ? 0 : pv + 1], pred = tour[pv == 0 ? n - 1 : pv - 1];
                const uint32_t link = s_plinks[v], lp = link & 0xffffu, ls = link >> 16;
                moved |= !((pred == lp && succ == ls) || (pred == ls && succ == lp));
            }
            const bool parent_tour = ballot(moved) == 0;
            if (lane == 0) {
                const uint32_t z = pos[0];
                const uint32_t next = tour[z + 1 == n ? 0 : z + 1], prev = tour[z == 0 ? n - 1 : z - 1];
                p.out_identity[2 * static_cast<size_t>(ant)] = f;
                p.out_identity[2 * static_cast<size_t>(ant) + 1] = g;
                p.out_zero[ant] = z | (next < prev ? 0x80000000u : 0u) | (parent_tour ? kParentTour : 0u);
            }
        }
        if (lane == 0) {
            faco_ant_info info;
            info.construction_cost = construction_cost;
            info.final_cost = final_cost;
            info.mne = mne; info.steps = steps; info.nonidentity_relocations = relocations;
            info.legal_exhausted = exhausted;
            info.processed_nodes = processed; info.candidate_checks = checks;
            info.move_evaluations = evaluations; info.accepted_moves = accepted;
            info.reactivations = reactivations; info.evaluation_limit_reached = limit_hit;
            info.checklist_size = logged;
            info.start_node = start;
            info.constraint_rejections = rejections;
            info.pad = 0;
            p.out_info[ant] = info;
            if constexpr (Escape) {
                // EscapeStats order (escape_slots.hpp:70-75); every step
                // (and an exhausted one) was one construction opportunity
                uint32_t* st = p.esc_stats + static_cast<size_t>(ant) * kEscStats;
                st[0] = steps + exhausted; st[1] = e_gates; st[2] = e_replaced; st[3] = e_relocations;
                st[4] = e_anchor_nodes; st[5] = e_ls_replaced; st[6] = e_old_views; st[7] = e_anchor_pushes;
                st[8] = e_new; st[9] = e_construction_full; st[10] = e_ls_full; st[11] = 0;
            }
            if constexpr (Profile) {
                uint64_t* t = p.phase_ticks + kPhaseSlots * static_cast<size_t>(ant);
                t[0] = tick1 - tick0; t[1] = tick2 - tick1; t[2] = tick3 - tick2;
                for (int k = 0; k < 5; ++k) t[3 + k] = sub[k];
                t[8] = tick0; t[9] = wall_clock64();
            }
        }
        wave_sync();
    }
    if constexpr (!Persistent) {
        break;
    } else {
        // every ant of the batch is done: the colony update, in place
        __syncthreads();
        update_colony<Escape>(cp, colony, threadIdx.x, blockDim.x, reinterpret_cast<uint64_t*>(lds + shared_bytes),
                              *s_state, s_plinks, with_ppos ? s_ppos : nullptr);
        if constexpr (Controlled) {
            __syncthreads();
            control_update(*ctl, colony, threadIdx.x, blockDim.x, *reinterpret_cast<UpdateShared*>(lds + shared_bytes),
                           lds + shared_bytes + round16(static_cast<uint32_t>(sizeof(UpdateShared))));
        }
        if (threadIdx.x == 0) next_slot = 0;
        __syncthreads();
    }
    }
}

template<bool NbrLds, bool Profile, bool Diag, bool OrOpt>
__global__ void __launch_bounds__(1024) faco_ant_search(const AntParams p) {
    ant_search<NbrLds, Profile, Diag, OrOpt, false>(p, ColonyParams{}, 1u, nullptr);
}

// Plain FACO over `rounds` batches in one launch: one workgroup per colony
// runs its ants and its colony update batch after batch, so a colony never
// waits for the others between batches.
// Hard mode (graph-constrained construction and 2-opt), per batch.
template<bool NbrLds>
__global__ void __launch_bounds__(1024) faco_ant_search_hard(const AntParams p) {
    ant_search<NbrLds, false, false, false, false, true>(p, ColonyParams{}, 1u, nullptr);
}

// Escape mode (the Hard graph with escape gates, slot views, per-ant edge
// caches and anchor LS views), per batch.
template<bool NbrLds>
__global__ void __launch_bounds__(1024) faco_ant_search_escape(const AntParams p) {
    ant_search<NbrLds, false, false, false, false, true, false, true>(p, ColonyParams{}, 1u, nullptr);
}

template<bool NbrLds, bool OrOpt, bool Hard, bool Escape = false>
__global__ void __launch_bounds__(1024) faco_colony_run(const AntParams p, const ColonyParams cp, uint32_t rounds) {
    __shared__ ColonyState s_state;
    ant_search<NbrLds, false, false, OrOpt, true, Hard, false, Escape>(p, cp, rounds, &s_state);
}

// GP-controlled FACO over `rounds` batches in one launch: one workgroup per
// colony runs decision, ants, colony update and archive update batch after
// batch (the per-batch kernels' order, cuda/batch_engine.cu:880-1060).
// Two instantiations by workgroup size: up to 12 waves (three per SIMD
// leave the ant loop and the control bodies 168 registers) and up to 16
// (128 registers; the control bodies spill some values). The parameters of
// the control bodies are read from device memory (`ctl`).
template<bool NbrLds, bool OrOpt, int MaxThreads>
__global__ void __launch_bounds__(MaxThreads) faco_control_run(const AntParams p, const ColonyParams cp,
                                                                const ControlParams* ctl, uint32_t rounds) {
    __shared__ ColonyState s_state;
    ant_search<NbrLds, false, false, OrOpt, true, false, true>(p, cp, rounds, &s_state, ctl);
}

// ---------------------------------------------------------------------------

// ---------------------------------------------------------------------------
// Colony reset (reference reset_colony + initialize_products,
// cuda/colony_state.cuh:72-110): parent = epoch = global = initial tour,
// bounds from the initial cost, every trail at the upper bound.

__global__ void __launch_bounds__(256) faco_colony_reset(const ResetParams p) {
    const uint32_t colony = blockIdx.x, ns = p.n, tid = threadIdx.x;
    const uint32_t n = p.shapes ? p.shapes[colony].n : ns;
    const uint32_t geom = p.geometry ? p.geometry[colony] : 0u;
    const uint32_t* init = p.initial + static_cast<size_t>(geom) * ns;
    __shared__ ColonyState s;
    if (tid == 0) {
        ColonyState z = {};
        z.global_cost = z.epoch_cost = z.parent_cost = p.initial_cost[geom];
        bounds(z, p.pw, p.retention, p.p_best_root);
        z.default_trail = z.maximum;
        z.target = p.target;
        z.best_key = ~0ull;
        s = z;
    }
    __syncthreads();
    const size_t base = static_cast<size_t>(colony) * ns;
    for (uint32_t i = tid; i < n; i += blockDim.x) {
        const uint32_t v = init[i];
        p.parent[base + i] = p.epoch[base + i] = p.global[base + i] = v;
        const uint32_t pred = init[i == 0 ? n - 1 : i - 1], succ = init[i + 1 == n ? 0 : i + 1];
        p.links[base + v] = pred | (succ << 16);
        if (v == 0) s.parent_zero = i | (succ < pred ? 0x80000000u : 0u);
    }
    __syncthreads();
    if (tid == 0) p.state[colony] = s;
    const uint32_t cells = n * p.pw;
    const double* eta = p.eta + static_cast<size_t>(geom) * ns * p.pw;
    for (uint32_t i = tid; i < cells; i += blockDim.x) {
        p.trails[static_cast<size_t>(colony) * ns * p.pw + i] = s.maximum;
        p.products[static_cast<size_t>(colony) * ns * p.pw + i] = s.maximum * eta[i];
    }
}

// ---------------------------------------------------------------------------
// Diagnostics: device EUC_2D and Philox streams.
__global__ void faco_distance_check(const int32_t* pairs, uint32_t count, int64_t* out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const uint32_t a = pack_xy(static_cast<uint32_t>(pairs[4 * i]), static_cast<uint32_t>(pairs[4 * i + 1]));
    const uint32_t b = pack_xy(static_cast<uint32_t>(pairs[4 * i + 2]), static_cast<uint32_t>(pairs[4 * i + 3]));
    out[i] = euc2d(a, b);
}

__global__ void faco_start_check(uint64_t key, uint32_t batch, uint32_t ant, const uint32_t* counts,
                                 uint32_t n, uint32_t* out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = start_node(key, batch, ant, counts[i]);
}

hipError_t launch_start_check(uint64_t key, uint32_t batch, uint32_t ant, const uint32_t* counts, uint32_t n,
                              uint32_t* out, hipStream_t stream) {
    if (n) hipLaunchKernelGGL(faco_start_check, dim3((n + 63) / 64), dim3(64), 0, stream, key, batch, ant, counts, n, out);
    return hipGetLastError();
}

__global__ void faco_philox_check(uint64_t key, uint32_t batch, uint32_t ant, uint32_t step,
                                             uint32_t words, uint32_t* out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    Stream s(key, batch, ant, step);
    for (uint32_t i = 0; i < words; ++i) out[i] = s.next();
    if (words >= 2) {
        // the single-block fast path must agree with the stream
        const double u = step_uniform(key, batch, ant, step);
        Stream t(key, batch, ant, step);
        if (u != t.uniform53()) out[0] = ~out[0];
    }
}

// Plain streaming copy (16-byte vectors): the calibration kernel whose
// known traffic scales the HBM counters of the engine's kernels.
__global__ void __launch_bounds__(256) faco_copy_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                        size_t count) {
    for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < count;
         i += static_cast<size_t>(gridDim.x) * blockDim.x)
        dst[i] = src[i];
}

hipError_t launch_copy(const void* src, void* dst, size_t bytes, hipStream_t stream) {
    const size_t count = bytes / 16;
    hipLaunchKernelGGL(faco_copy_kernel, dim3(2048), dim3(256), 0, stream, static_cast<const uint4*>(src),
                       static_cast<uint4*>(dst), count);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
int ant_search_blocks_per_cu(const AntParams& p) {
    const uint32_t lds = ant_shared_bytes(launch_lds_n(p), p.nbr_in_lds ? p.nbr_lds_bytes : 0u) + p.waves * p.per_wave_bytes;
    int blocks = 0;
    const hipError_t e = p.nbr_in_lds
        ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, faco_ant_search<true, false, false, false>, 64 * p.waves, lds)
        : hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, faco_ant_search<false, false, false, false>, 64 * p.waves, lds);
    return e == hipSuccess && blocks > 0 ? blocks : 1;
}

hipError_t launch_ant_search(const AntParams& p, uint32_t blocks, hipStream_t stream) {
    const uint32_t lds = ant_shared_bytes(launch_lds_n(p), p.nbr_in_lds ? p.nbr_lds_bytes : 0u) + p.waves * p.per_wave_bytes;
    const dim3 grid(blocks), block(64 * p.waves);
    const bool diag = p.explicit_mode || p.per_ant_parent || p.trace_selected || p.diag_construction ||
                      p.diag_positions || p.diag_checklist;
    // one instantiation per feature set, so the production kernel of plain
    // 2-opt carries neither the Or-opt moves nor the diagnostic outputs
    auto go = [&](auto kernel) { hipLaunchKernelGGL(kernel, grid, block, lds, stream, p); };
    const bool nl = p.nbr_in_lds != 0, oo = p.or_opt != 0;
    if (p.graph_offsets) {
        if (diag || p.phase_ticks || oo) return hipErrorInvalidValue;   // Hard mode: plain 2-opt, no diagnostics
        if (p.esc_rows) nl ? go(faco_ant_search_escape<true>) : go(faco_ant_search_escape<false>);
        else nl ? go(faco_ant_search_hard<true>) : go(faco_ant_search_hard<false>);
        return hipGetLastError();
    }
    if (diag) {
        if (p.phase_ticks) return hipErrorInvalidValue;   // no clock profile of diagnostic runs
        if (nl) oo ? go(faco_ant_search<true, false, true, true>) : go(faco_ant_search<true, false, true, false>);
        else oo ? go(faco_ant_search<false, false, true, true>) : go(faco_ant_search<false, false, true, false>);
    } else if (p.phase_ticks) {
        if (nl) oo ? go(faco_ant_search<true, true, false, true>) : go(faco_ant_search<true, true, false, false>);
        else oo ? go(faco_ant_search<false, true, false, true>) : go(faco_ant_search<false, true, false, false>);
    } else {
        if (nl) oo ? go(faco_ant_search<true, false, false, true>) : go(faco_ant_search<true, false, false, false>);
        else oo ? go(faco_ant_search<false, false, false, true>) : go(faco_ant_search<false, false, false, false>);
    }
    return hipGetLastError();
}

hipError_t launch_colony_run(const AntParams& p, const ColonyParams& cp, uint32_t colonies, uint32_t rounds,
                             hipStream_t stream) {
    if (p.blocks_per_colony != 1 || p.phase_ticks || p.explicit_mode || p.per_ant_parent || p.trace_selected ||
        (p.graph_offsets && p.or_opt) ||
        p.diag_construction || p.diag_positions || p.diag_checklist || p.control || p.out_identity)
        return hipErrorInvalidValue;   // one workgroup per colony, production plain, Hard or Escape FACO only
    const uint32_t lds = ant_shared_bytes(launch_lds_n(p), p.nbr_in_lds ? p.nbr_lds_bytes : 0u) + p.waves * p.per_wave_bytes;
    const dim3 grid(colonies), block(64 * p.waves);
    auto go = [&](auto kernel) { hipLaunchKernelGGL(kernel, grid, block, lds, stream, p, cp, rounds); };
    const bool nl = p.nbr_in_lds != 0, oo = p.or_opt != 0;
    if (p.esc_rows) {        // Escape mode (on the Hard graph): plain 2-opt
        nl ? go(faco_colony_run<true, false, true, true>) : go(faco_colony_run<false, false, true, true>);
    } else if (p.graph_offsets) {   // Hard mode: plain 2-opt
        nl ? go(faco_colony_run<true, false, true>) : go(faco_colony_run<false, false, true>);
    } else if (nl) {
        oo ? go(faco_colony_run<true, true, false>) : go(faco_colony_run<true, false, false>);
    } else {
        oo ? go(faco_colony_run<false, true, false>) : go(faco_colony_run<false, false, false>);
    }
    return hipGetLastError();
}

// LDS of the persistent controlled kernel: the ants' share, or the
// control bodies' if larger (they reuse the ants' per-wave space).
uint32_t control_run_lds_bytes(const AntParams& p) {
    const uint32_t shared = ant_shared_bytes(launch_lds_n(p), p.nbr_in_lds ? p.nbr_lds_bytes : 0u);
    const uint32_t ants = p.waves * p.per_wave_bytes;
    const uint32_t decide = static_cast<uint32_t>(sizeof(DecideShared));
    const uint32_t update = round16_host(static_cast<uint32_t>(sizeof(UpdateShared))) +
                            static_cast<uint32_t>(control_candidate_bytes(p.ants));
    return shared + std::max(ants, std::max(decide, update));
}

hipError_t launch_control_run(const AntParams& p, const ColonyParams& cp, const ControlParams& ctl,
                              const ControlParams* device_ctl, uint32_t colonies, uint32_t rounds,
                              hipStream_t stream) {
    if (p.blocks_per_colony != 1 || p.waves < 2 || p.waves > 16 || p.phase_ticks || p.explicit_mode || p.per_ant_parent ||
        p.trace_selected || p.graph_offsets || p.diag_construction || p.diag_positions || p.diag_checklist ||
        !p.control || !p.out_identity || p.filter_tours || ctl.views_only || ctl.forced_action >= 0)
        return hipErrorInvalidValue;   // one workgroup per colony, production controlled FACO only
    const uint32_t lds = control_run_lds_bytes(p);
    const dim3 grid(colonies), block(64 * p.waves);
    auto go = [&](auto kernel) { hipLaunchKernelGGL(kernel, grid, block, lds, stream, p, cp, device_ctl, rounds); };
    const bool nl = p.nbr_in_lds != 0, oo = p.or_opt != 0;
    if (p.waves <= 12) {
        if (nl) oo ? go(faco_control_run<true, true, 768>) : go(faco_control_run<true, false, 768>);
        else oo ? go(faco_control_run<false, true, 768>) : go(faco_control_run<false, false, 768>);
    } else {
        if (nl) oo ? go(faco_control_run<true, true, 1024>) : go(faco_control_run<true, false, 1024>);
        else oo ? go(faco_control_run<false, true, 1024>) : go(faco_control_run<false, false, 1024>);
    }
    return hipGetLastError();
}

hipError_t launch_colony_update(const ColonyParams& p, uint32_t colonies, hipStream_t stream) {
    hipLaunchKernelGGL(faco_colony_update, dim3(colonies), dim3(256), 4u * p.n, stream, p);
    return hipGetLastError();
}

hipError_t launch_colony_reset(const ResetParams& p, uint32_t colonies, hipStream_t stream) {
    hipLaunchKernelGGL(faco_colony_reset, dim3(colonies), dim3(256), 0, stream, p);
    return hipGetLastError();
}

hipError_t launch_distance_check(const int32_t* pairs, uint32_t count, int64_t* out, hipStream_t stream) {
    if (count) hipLaunchKernelGGL(faco_distance_check, dim3((count + 255) / 256), dim3(256), 0, stream, pairs, count, out);
    return hipGetLastError();
}

hipError_t launch_philox_check(uint64_t key, uint32_t batch, uint32_t ant, uint32_t step, uint32_t words,
                               uint32_t* out, hipStream_t stream) {
    hipLaunchKernelGGL(faco_philox_check, dim3(1), dim3(64), 0, stream, key, batch, ant, step, words, out);
    return hipGetLastError();
}

}  // namespace faco
