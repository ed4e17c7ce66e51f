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

// Throws std::invalid_argument on a malformed program (host only).
void validate_program(const faco_program& p);
void validate_features(const float* features, const uint32_t* masks, uint32_t colonies);

hipError_t launch_score(const faco_program* programs, uint32_t replicas, const float* features,
                        const uint32_t* masks, uint32_t colonies, float* scores, int32_t* actions,
                        hipStream_t stream);

}  // namespace faco
