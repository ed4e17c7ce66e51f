// Escape-mode device pieces shared by the ant kernel and the diagnostic
// escape kernels: the escape word stream, the slot view of one row held one
// entry per lane, and the (distance, city) order of an LS view.
// Reference semantics: cuda/escape_device.cuh:8-27 (EscapeRandom),
// cpp/include/gp_faco/escape_slots.hpp:18-68 (escape_slot_view,
// sort_escape_ls_row).
#pragma once
#include "faco_common.h"

namespace faco {

constexpr uint32_t kEscapeCapacity = 64;   // out-of-graph edges one ant may add (escape_edges.hpp:14)
constexpr uint8_t kNoSource = 0xff;        // slot view: a slot that kept its own city

// Escape word stream of (key, batch, domain | ant, event): Philox counter
// blocks (block, event, domain | ant, batch), four words per block; domain
// 0x40000000 for construction gates, 0x80000000 for LS rows. Wave-uniform.
struct EscapeStream {
    uint32_t k0, k1, ant, batch, event, block, index, w[4];
    __device__ EscapeStream(uint64_t key, uint32_t b, uint32_t a, uint32_t e, bool ls)
        : k0(static_cast<uint32_t>(key)), k1(static_cast<uint32_t>(key >> 32)),
          ant((ls ? 0x80000000u : 0x40000000u) | a), batch(b), event(e), block(0), index(4), w{0, 0, 0, 0} {}
    __device__ uint32_t draw() {
        if (index == 4) {
            w[0] = block++; w[1] = event; w[2] = ant; w[3] = batch;
            philox4x32_10(w, k0, k1);
            index = 0;
        }
        return w[index++];
    }
    // unbiased draw below `count`: words under 2^32 mod count are rejected
    __device__ uint32_t below(uint32_t count) {
        const uint32_t threshold = (0u - count) % count;
        uint32_t v;
        do { v = draw(); } while (v < threshold);
        return v % count;
    }
};

__device__ __forceinline__ uint32_t esc_lane() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }
__device__ __forceinline__ uint32_t esc_read(uint32_t v, uint32_t lane) {
    return static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), static_cast<int>(lane)));
}
__device__ __forceinline__ void esc_wave_fence() {
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// The slot view of one row (escape_slot_view): lane j < width holds the
// row's entry (n for padding), lane j < bw the backup row's. A quarter of
// the valid slots (the remainder rounded by one draw) take distinct backup
// cities that are not the row's city nor in the row, slots and cities
// picked by sequential partial shuffles in stream order. Each lane learns
// its slot's city and, for a replaced slot, the backup lane it came from
// (kNoSource otherwise). `scratch`: 128 bytes of this wave's LDS.
struct LaneView {
    uint32_t member, source, valid, replacements;
};
__device__ __forceinline__ LaneView lane_slot_view(uint32_t orig, uint32_t width, uint32_t b, uint32_t bw, uint32_t n,
                                                   uint32_t from, EscapeStream& rng, uint8_t* scratch) {
    const uint32_t lane = esc_lane();
    LaneView v{orig, kNoSource, 0u, 0u};
    v.valid = static_cast<uint32_t>(__builtin_popcountll(__ballot(lane < width && orig < n)));
    uint32_t wanted = v.valid / 4;
    if (v.valid % 4) wanted += (rng.draw() & 3u) < v.valid % 4 ? 1u : 0u;
    if (!wanted) return v;
    // eligible backups: before the first padding entry, not the row's city,
    // not among the row's valid members, first occurrence in the backup row
    const uint64_t pad = __ballot(lane < bw && b >= n) | (bw < 64 ? (~0ull << bw) : 0ull);
    const uint32_t first_pad = pad ? static_cast<uint32_t>(__builtin_ctzll(pad)) : 64u;
    bool ok = lane < first_pad && b != from;
    for (uint32_t k = 0; k < v.valid; ++k) ok = ok && esc_read(orig, k) != b;
    for (uint32_t k = 0; k < first_pad; ++k) ok = ok && !(k < lane && esc_read(b, k) == b);
    const uint64_t emask = __ballot(ok);
    const uint32_t count = static_cast<uint32_t>(__builtin_popcountll(emask));
    uint8_t* eligible = scratch;          // [64] backup lanes
    uint8_t* slots = scratch + 64;        // [32]
    uint8_t* pick = scratch + 96;         // [32] backup lane of each slot's new city
    if (ok) eligible[__builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(emask >> 32),
                                               __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(emask), 0u))] =
        static_cast<uint8_t>(lane);
    if (lane < 32) { slots[lane] = static_cast<uint8_t>(lane); pick[lane] = kNoSource; }
    esc_wave_fence();
    if (wanted > count) wanted = count;
    for (uint32_t j = 0; j < wanted; ++j) {
        const uint32_t pick_slot = j + rng.below(v.valid - j);
        const uint32_t pick_node = j + rng.below(count - j);
        if (lane == 0) {
            const uint8_t slot = slots[pick_slot];
            slots[pick_slot] = slots[j]; slots[j] = slot;
            const uint8_t node = eligible[pick_node];
            eligible[pick_node] = eligible[j]; eligible[j] = node;
            pick[slot] = node;
        }
        esc_wave_fence();
    }
    const uint32_t src = lane < width ? pick[lane] : kNoSource;
    const uint32_t from_lane = src == kNoSource ? lane : src;
    const uint32_t moved = static_cast<uint32_t>(__builtin_amdgcn_ds_bpermute(static_cast<int>(from_lane << 2),
                                                                              static_cast<int>(b)));
    esc_wave_fence();   // the scratch is free again for the caller
    v.member = src == kNoSource ? orig : moved;
    v.source = src;
    v.replacements = wanted;
    return v;
}

// The LS form of a view (sort_escape_ls_row): the valid prefix ordered by
// (distance to `from`, city), the replaced flag moving with its city.
// Lanes >= valid keep their entries. `scratch`: 128 bytes of this wave's LDS.
__device__ __forceinline__ void lane_sort_row(uint32_t& member, uint32_t& source, uint32_t valid, uint32_t from,
                                              const uint32_t* s_xy, uint32_t* scratch) {
    const uint32_t lane = esc_lane();
    const bool in = lane < valid;
    const uint32_t d = in ? static_cast<uint32_t>(euc2d(s_xy[from], s_xy[member])) : 0u;
    uint32_t rank = 0;
    for (uint32_t k = 0; k < valid; ++k) {
        const uint32_t dk = esc_read(d, k), mk = esc_read(member, k);
        rank += (dk < d || (dk == d && mk < member)) ? 1u : 0u;
    }
    if (in) scratch[rank] = member | (source << 16);
    esc_wave_fence();
    if (in) {
        const uint32_t e = scratch[lane];
        member = e & 0xffffu;
        source = e >> 16;
    }
    esc_wave_fence();
}

}  // namespace faco
