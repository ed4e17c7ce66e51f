// GP controller IR v1: validation and the stack interpreter shared by the
// host scorer and the device kernel.
// Semantics: reference cpp/include/gp_faco/program_ops.hpp:13-31,
// cpp/src/program_cpu.cpp:10-116. Every intermediate is clipped to [-8, 8];
// AQ (opcode 8) is left / sqrt(1 + right^2) with IEEE sqrt and division
// (no FMA contraction), the reference CPU's formula.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/faco_engine.h"

namespace faco {

__host__ __device__ __forceinline__ float clip8(float v) { return fminf(8.0f, fmaxf(-8.0f, v)); }

__host__ __device__ __forceinline__ float apply_operation(int op, float left, float right) {
    float r = 0.0f;
    switch (op) {
        case 2: r = left + right; break;
        case 3: r = left - right; break;
        case 4: r = left * right; break;
        case 5: r = fminf(left, right); break;
        case 6: r = fmaxf(left, right); break;
        case 7: r = fabsf(left); break;
        case 8: {
            const float sq = right * right;
            const float den = 1.0f + sq;
            r = left * (1.0f / sqrtf(den));
            break;
        }
        default: break;
    }
    return clip8(r);
}

// Value of `program` for one (colony, action): features are [12][colonies][32].
__host__ __device__ __forceinline__ float program_value(const faco_program& p, const float* features,
                                                       uint32_t colonies, uint32_t colony, uint32_t action) {
    float stack[6] = {0, 0, 0, 0, 0, 0};
    int top = 0;
    for (int i = 0; i < p.length; ++i) {
        const int op = p.opcode[i];
        if (op == 0) {
            stack[top++] = features[(static_cast<size_t>(p.operand[i]) * colonies + colony) * 32 + action];
        } else if (op == 1) {
            stack[top++] = p.constants[p.operand[i]];
        } else {
            // binary operations pop the right operand first
            const float right = op == 7 ? 0.0f : stack[--top];
            const float left = stack[--top];
            stack[top++] = apply_operation(op, left, right);
        }
    }
    return stack[0];
}

// Smallest index of the 32-lane group whose legal value is within 1e-6 of
// the group's largest legal value; 32 when no lane is legal
// (cuda/gp_score.cuh:24-34). Every lane of the group must call it.
__device__ __forceinline__ int group32_choose(float value, bool legal, uint32_t lane32) {
    float maximum = legal ? value : -INFINITY;
#pragma unroll
    for (int offset = 16; offset; offset >>= 1) maximum = fmaxf(maximum, __shfl_xor(maximum, offset, 32));
    int chosen = legal && maximum - value <= 1e-6f ? static_cast<int>(lane32) : 32;
#pragma unroll
    for (int offset = 16; offset; offset >>= 1) {
        const int other = __shfl_xor(chosen, offset, 32);
        chosen = other < chosen ? other : chosen;
    }
    return chosen;
}

// Controller decision of one colony by a 32-lane group (lane32 = action
// slot; cuda/gp_score.cuh:37-62). kind 0 scores the 32 actions with tree 0;
// kind 1 chooses restart (2 blocks of 16), region (4 blocks of 4) and MNE
// level (4 actions) in turn, a block being legal when any of its actions
// is. `scores` (may be null) receives the 32 values (kind 1: zeros, then the
// stage values in slots 0-1, 2-5 and 6-9).
__device__ __forceinline__ int controller_decide(const faco_controller& c, const float* features, uint32_t colonies,
                                                 uint32_t colony, uint32_t mask, uint32_t lane32, float* scores) {
    if (!c.kind) {
        const float v = program_value(c.trees[0], features, colonies, colony, lane32);
        if (scores) scores[lane32] = v;
        return group32_choose(v, (mask >> lane32) & 1u, lane32);
    }
    if (scores) scores[lane32] = 0.0f;
    uint32_t prefix = 0;
    for (uint32_t role = 0; role < 3; ++role) {
        const uint32_t width = role == 0 ? 2u : 4u, stride = role == 0 ? 16u : role == 1 ? 4u : 1u;
        const uint32_t slot = role == 0 ? 0u : role == 1 ? 2u : 6u;
        const uint32_t action = prefix + (lane32 < width ? lane32 : 0u) * stride;
        const bool legal = lane32 < width && ((mask >> action) & ((1u << stride) - 1u)) != 0u;
        const float v = program_value(c.trees[role], features, colonies, colony, action);
        if (scores && lane32 < width) scores[slot + lane32] = v;
        prefix += stride * static_cast<uint32_t>(group32_choose(v, legal, lane32));
    }
    return static_cast<int>(prefix);
}

// Host twin of controller_decide for one colony.
inline int controller_decide_host(const faco_controller& c, const float* features, uint32_t colonies,
                                  uint32_t colony, uint32_t mask, float* scores) {
    auto choose = [](const float* values, const bool* legal, uint32_t width) {
        float maximum = -INFINITY;
        for (uint32_t i = 0; i < width; ++i) if (legal[i]) maximum = values[i] > maximum ? values[i] : maximum;
        for (uint32_t i = 0; i < width; ++i) if (legal[i] && maximum - values[i] <= 1e-6f) return static_cast<int>(i);
        return 32;
    };
    float values[32];
    bool legal[32];
    if (!c.kind) {
        for (uint32_t a = 0; a < 32; ++a) {
            values[a] = scores[a] = program_value(c.trees[0], features, colonies, colony, a);
            legal[a] = (mask >> a) & 1u;
        }
        return choose(values, legal, 32);
    }
    for (uint32_t a = 0; a < 32; ++a) scores[a] = 0.0f;
    uint32_t prefix = 0;
    for (uint32_t role = 0; role < 3; ++role) {
        const uint32_t width = role == 0 ? 2u : 4u, stride = role == 0 ? 16u : role == 1 ? 4u : 1u;
        const uint32_t slot = role == 0 ? 0u : role == 1 ? 2u : 6u;
        for (uint32_t l = 0; l < width; ++l) {
            const uint32_t action = prefix + l * stride;
            legal[l] = ((mask >> action) & ((1u << stride) - 1u)) != 0u;
            values[l] = scores[slot + l] = program_value(c.trees[role], features, colonies, colony, action);
        }
        prefix += stride * static_cast<uint32_t>(choose(values, legal, width));
    }
    return static_cast<int>(prefix);
}

// Throws std::invalid_argument on a malformed program (host only).
void validate_program(const faco_program& p);
void validate_features(const float* features, const uint32_t* masks, uint32_t colonies);
void validate_controller(const faco_controller& c);

hipError_t launch_score(const faco_program* programs, uint32_t replicas, const float* features,
                        const uint32_t* masks, uint32_t colonies, float* scores, int32_t* actions,
                        hipStream_t stream);

}  // namespace faco
