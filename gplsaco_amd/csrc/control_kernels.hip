// Restart and perturbation control of the GP-controlled FACO loop on
// MI355X: control features, decision, restart transaction, archive and
// feedback update. One workgroup per colony; every kernel reads and writes
// only its own colony's state, so a batch of colonies is independent work.
//
// Semantics: reference cuda/control_state.cuh:19-364 (initialize_control,
// fingerprint_ants, update_control, build_control_features,
// apply_control_action), cpp/include/gp_faco/control_ops.hpp:14-212,
// baseline_policy.hpp:59-79, factorial_policy.hpp:32-45,
// cuda/baseline_policy.cuh:10-38 and cuda/gp_score.cuh:10-102, in the
// per-batch order of cuda/batch_engine.cu:880-1060.
//
// Design (MI355X): features, scoring and the restart transaction are one
// kernel (no host round trip between them); tour fingerprints come from the
// ant kernel's LDS copy of each tour; the archive is a two-buffer ring so
// the update writes the next archive straight from the chosen tours; tours
// are compared in canonical order (from city 0 towards its smaller
// neighbour) 64 positions per wave step with a ballot for the first
// difference. Duplicates are found by grouping candidates with equal (cost,
// identity) and checking each member against its group leader with the
// whole workgroup; only a real identity collision falls back to the serial
#include "control_device.h"

namespace faco {

// ---------------------------------------------------------------------------
// initialize_control (cuda/control_state.cuh:19-43): archive = {initial
// tour}, tracked costs = initial cost, active = epoch identity = its
// identity, feedback cleared.
__global__ void __launch_bounds__(256) faco_control_init(const ControlParams p) {
    // n: this colony's cities; ns: stride of the per-city arrays (they
    // differ only in a mixed-size batch)
    const uint32_t colony = blockIdx.x, ns = p.n, tid = threadIdx.x;
    const uint32_t n = p.shapes ? p.shapes[colony].n : ns;
    const uint32_t geom = p.geometry[colony];
    const uint32_t* init = p.initial + static_cast<size_t>(geom) * ns;
    uint32_t* tours = p.archive_tours + static_cast<size_t>(colony) * 2 * kArchive * ns;
    uint32_t* links = p.archive_links + static_cast<size_t>(colony) * 2 * kArchive * ns;
    __shared__ uint64_t red[2][256];
    __shared__ uint32_t s_zero;
    uint64_t f = 0, g = 0;
    for (uint32_t i = tid; i < n; i += blockDim.x) {
        const uint32_t v = init[i];
        const uint32_t pred = init[i == 0 ? n - 1 : i - 1], succ = init[i + 1 == n ? 0 : i + 1];
        tours[i] = v;
        links[v] = pred | (succ << 16);
        if (v == 0) s_zero = i | (succ < pred ? 0x80000000u : 0u);
        uint64_t a, b;
        edge_identity(v, succ, a, b);
        f ^= a; g ^= b;
    }
    red[0][tid] = f; red[1][tid] = g;
    __syncthreads();
    for (uint32_t stride = blockDim.x / 2; stride; stride >>= 1) {
        if (tid < stride) { red[0][tid] ^= red[0][tid + stride]; red[1][tid] ^= red[1][tid + stride]; }
        __syncthreads();
    }
    if (tid == 0) {
        ControlState c = {};
        c.archive_size = 1;
        c.archive_cost[0] = c.tracked_global = c.tracked_epoch = p.initial_cost[geom];
        c.archive_id[0][0] = p.force_collisions ? 0 : red[0][0];
        c.archive_id[0][1] = p.force_collisions ? 0 : red[1][0];
        c.archive_zero[0] = s_zero;
        c.active_id[0] = c.epoch_id[0] = c.archive_id[0][0];
        c.active_id[1] = c.epoch_id[1] = c.archive_id[0][1];
        c.alternative = kArchive;
        c.start_count = n < kRegion ? n : kRegion;
        p.control[colony] = c;
    }
}

// ---------------------------------------------------------------------------
// Per-batch control kernels: one workgroup per colony.
__global__ void __launch_bounds__(512) faco_control_decide(const ControlParams p) {
    __shared__ DecideShared sh;
    control_decide(p, blockIdx.x, gridDim.x, threadIdx.x, blockDim.x, sh, nullptr, nullptr);
}

__global__ void __launch_bounds__(256) faco_control_update(const ControlParams p) {
    extern __shared__ __attribute__((aligned(16))) unsigned char lds[];
    __shared__ UpdateShared sh;
    control_update(p, blockIdx.x, threadIdx.x, blockDim.x, sh, lds);
}

hipError_t launch_control_init(const ControlParams& p, uint32_t colonies, hipStream_t stream) {
    hipLaunchKernelGGL(faco_control_init, dim3(colonies), dim3(256), 0, stream, p);
    return hipGetLastError();
}

hipError_t launch_control_decide(const ControlParams& p, uint32_t colonies, hipStream_t stream) {
    hipLaunchKernelGGL(faco_control_decide, dim3(colonies), dim3(512), 0, stream, p);
    return hipGetLastError();
}

hipError_t launch_control_update(const ControlParams& p, uint32_t colonies, hipStream_t stream) {
    const size_t lds = control_candidate_bytes(p.ants);
    hipLaunchKernelGGL(faco_control_update, dim3(colonies), dim3(256), lds, stream, p);
    return hipGetLastError();
}

// score_controllers (cuda/gp_score.cuh:37-62): one 32-lane group per colony.
__global__ void __launch_bounds__(256) faco_score_controller_kernel(const faco_controller* controller,
                                                                    const float* features, const uint32_t* masks,
                                                                    uint32_t colonies, float* scores, int32_t* actions) {
    const uint32_t thread = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t colony = thread >> 5, a = thread & 31u;
    const bool valid = colony < colonies;
    const uint32_t cc = valid ? colony : 0u;
    __shared__ float s_scores[8][32];
    float* mine = s_scores[threadIdx.x >> 5];
    const int action = controller_decide(*controller, features, colonies, cc, masks[cc], a, mine);
    __syncthreads();
    if (valid) {
        scores[static_cast<size_t>(colony) * 32 + a] = mine[a];
        if (a == 0) actions[colony] = action;
    }
}

hipError_t launch_score_controller(const faco_controller* controller, const float* features, const uint32_t* masks,
                                   uint32_t colonies, float* scores, int32_t* actions, hipStream_t stream) {
    const uint32_t threads = colonies * 32u;
    hipLaunchKernelGGL(faco_score_controller_kernel, dim3((threads + 255) / 256), dim3(256), 0, stream, controller,
                       features, masks, colonies, scores, actions);
    return hipGetLastError();
}

}  // namespace faco
