// GP controller scoring: host score_cpu and the wave64 device interpreter.
//
// Device layout: one wave64 scores two colonies, lane = (colony half, action);
// the legal-masked maximum and the "smallest legal action within 1e-6 of the
// maximum" rule run as 32-lane butterflies inside each half (reference
// cuda/gp_score.cuh:64-102, tie rule cpp/src/program_cpu.cpp:106-113).
#include "faco_common.h"
#include "gp_program.h"

#include <cmath>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

namespace faco {

void validate_program(const faco_program& p) {
    auto fail = [](const char* m) { throw std::invalid_argument(m); };
    if (p.ir_version != 1 || p.numeric_spec_id != 1 || (p.feature_spec_id != 1 && p.feature_spec_id != 2) ||
        p.length == 0 || p.length > 63 || p.constants_count > 63)
        fail("unknown IR version or program size out of range");
    for (int i = 0; i < p.constants_count; ++i)
        if (!std::isfinite(p.constants[i]) || std::fabs(p.constants[i]) > 2.0f) fail("ERC must be finite and in [-2, 2]");
    int depths[64];
    int top = 0;
    for (int i = 0; i < p.length; ++i) {
        const int op = p.opcode[i], operand = p.operand[i];
        if (op == 0 || op == 1) {
            if ((op == 0 && operand >= 12) || (op == 1 && operand >= p.constants_count)) fail("terminal index out of range");
            depths[top++] = 0;
        } else if (op >= 2 && op <= 8) {
            const int arity = op == 7 ? 1 : 2;
            if (operand != 0 || top < arity) fail("function operand nonzero or stack underflow");
            int depth = 0;
            for (int k = 0; k < arity; ++k) depth = depths[top - 1 - k] > depth ? depths[top - 1 - k] : depth;
            top -= arity;
            if (depth + 1 > 5) fail("tree depth above 5");
            depths[top++] = depth + 1;
        } else {
            fail("unknown opcode");
        }
        if (top > 6) fail("stack peak above 6");
    }
    if (top != 1) fail("final stack is not a single value");
}

// validate_controller (cpp/src/program_cpu.cpp:11-24): every tree is a
// valid feature-spec-2 program, the conditional trees read no feature of a
// decision not yet taken, and the trees hold at most 63 nodes together.
void validate_controller(const faco_controller& c) {
    if (c.kind > 1) throw std::invalid_argument("unknown controller representation");
    unsigned total = 0;
    for (unsigned role = 0; role < (c.kind ? 3u : 1u); ++role) {
        const faco_program& p = c.trees[role];
        validate_program(p);
        total += p.length;
        if (p.feature_spec_id != 2) throw std::invalid_argument("controllers need evaluation-count features (spec 2)");
        for (unsigned i = 0; c.kind && i < p.length; ++i) {
            const unsigned f = p.operand[i];
            if (p.opcode[i] == 0 && ((role == 0 && (f == 5 || f >= 8)) || (role == 1 && f == 5)))
                throw std::invalid_argument("a stage tree reads a feature of an undecided stage");
        }
    }
    if (total > 63) throw std::invalid_argument("controller holds more than 63 nodes");
}

void validate_features(const float* features, const uint32_t* masks, uint32_t colonies) {
    if (!colonies || !features || !masks) throw std::invalid_argument("feature or mask shape invalid");
    for (size_t i = 0; i < static_cast<size_t>(colonies) * 12 * 32; ++i) {
        const float v = features[i];
        if (!std::isfinite(v) || v < 0.0f || v > 1.0f) throw std::invalid_argument("features must be in [0, 1]");
    }
    for (uint32_t c = 0; c < colonies; ++c)
        if (!masks[c]) throw std::invalid_argument("no legal action");
}

__global__ void __launch_bounds__(256) gp_score_kernel(const faco_program* programs, uint32_t replicas,
                                                       const float* features, const uint32_t* masks,
                                                       uint32_t colonies, float* scores, int32_t* actions) {
    const uint32_t thread = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t colony = thread >> 5, action = thread & 31u;
    const bool valid = colony < colonies;
    // every lane of a wave takes part in the butterflies below
    const faco_program& program = programs[valid && replicas ? colony / replicas : 0];
    const float value = valid ? program_value(program, features, colonies, colony, action) : 0.0f;
    const bool legal = valid && ((masks[colony] >> action) & 1u);
    float maximum = legal ? value : -INFINITY;
#pragma unroll
    for (int offset = 16; offset; offset >>= 1) maximum = fmaxf(maximum, __shfl_xor(maximum, offset, 32));
    int chosen = legal && maximum - value <= 1e-6f ? static_cast<int>(action) : 32;
#pragma unroll
    for (int offset = 16; offset; offset >>= 1) {
        const int other = __shfl_xor(chosen, offset, 32);
        chosen = other < chosen ? other : chosen;
    }
    if (valid) {
        scores[static_cast<size_t>(colony) * 32 + action] = value;
        if (action == 0) actions[colony] = chosen;
    }
}

hipError_t launch_score(const faco_program* programs, uint32_t replicas, const float* features, const uint32_t* masks,
                        uint32_t colonies, float* scores, int32_t* actions, hipStream_t stream) {
    const uint32_t threads = colonies * 32u;
    hipLaunchKernelGGL(gp_score_kernel, dim3((threads + 255) / 256), dim3(256), 0, stream, programs, replicas,
                       features, masks, colonies, scores, actions);
    return hipGetLastError();
}

}  // namespace faco

namespace faco { void set_last_error(const char* message); }

namespace {
template<class F> int guarded_score(F&& f) {
    try { f(); return FACO_OK; }
    catch (const std::invalid_argument& e) { faco::set_last_error(e.what()); return FACO_INVALID_ARGUMENT; }
    catch (const std::exception& e) { faco::set_last_error(e.what()); return FACO_RUNTIME_ERROR; }
}
}  // namespace

extern "C" int faco_score_cpu(const faco_program* program, const float* features, const uint32_t* masks,
                              uint32_t colonies, float* scores, int32_t* actions) {
    return guarded_score([&] {
        if (!program || !scores || !actions) throw std::invalid_argument("null argument");
        faco::validate_program(*program);
        faco::validate_features(features, masks, colonies);
        for (uint32_t c = 0; c < colonies; ++c) {
            float maximum = -std::numeric_limits<float>::infinity();
            for (uint32_t a = 0; a < 32; ++a) {
                const float v = faco::program_value(*program, features, colonies, c, a);
                scores[static_cast<size_t>(c) * 32 + a] = v;
                if ((masks[c] >> a) & 1u) maximum = v > maximum ? v : maximum;
            }
            actions[c] = 32;
            for (uint32_t a = 0; a < 32; ++a) {
                if (((masks[c] >> a) & 1u) && maximum - scores[static_cast<size_t>(c) * 32 + a] <= 1e-6f) {
                    actions[c] = static_cast<int32_t>(a);
                    break;
                }
            }
        }
    });
}

extern "C" int faco_score_gpu(const faco_program* program, const float* features, const uint32_t* masks,
                              uint32_t colonies, float* scores, int32_t* actions) {
    return guarded_score([&] {
        if (!program || !scores || !actions) throw std::invalid_argument("null argument");
        faco::validate_program(*program);
        faco::validate_features(features, masks, colonies);
        auto check = [](hipError_t e, const char* what) {
            if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
        };
        faco_program* d_prog = nullptr; float* d_feat = nullptr; uint32_t* d_mask = nullptr;
        float* d_scores = nullptr; int32_t* d_actions = nullptr;
        const size_t fbytes = static_cast<size_t>(colonies) * 12 * 32 * sizeof(float);
        struct Frees { void* p[5]; ~Frees() { for (void* q : p) if (q) (void)hipFree(q); } } frees{};
        check(hipMalloc(reinterpret_cast<void**>(&d_prog), sizeof(faco_program)), "hipMalloc"); frees.p[0] = d_prog;
        check(hipMalloc(reinterpret_cast<void**>(&d_feat), fbytes), "hipMalloc"); frees.p[1] = d_feat;
        check(hipMalloc(reinterpret_cast<void**>(&d_mask), colonies * sizeof(uint32_t)), "hipMalloc"); frees.p[2] = d_mask;
        check(hipMalloc(reinterpret_cast<void**>(&d_scores), colonies * 32 * sizeof(float)), "hipMalloc"); frees.p[3] = d_scores;
        check(hipMalloc(reinterpret_cast<void**>(&d_actions), colonies * sizeof(int32_t)), "hipMalloc"); frees.p[4] = d_actions;
        check(hipMemcpy(d_prog, program, sizeof(faco_program), hipMemcpyHostToDevice), "upload");
        check(hipMemcpy(d_feat, features, fbytes, hipMemcpyHostToDevice), "upload");
        check(hipMemcpy(d_mask, masks, colonies * sizeof(uint32_t), hipMemcpyHostToDevice), "upload");
        check(faco::launch_score(d_prog, 0, d_feat, d_mask, colonies, d_scores, d_actions, nullptr), "score launch");
        check(hipMemcpy(scores, d_scores, colonies * 32 * sizeof(float), hipMemcpyDeviceToHost), "download");
        check(hipMemcpy(actions, d_actions, colonies * sizeof(int32_t), hipMemcpyDeviceToHost), "download");
    });
}

extern "C" int faco_score_controller_cpu(const faco_controller* controller, const float* features,
                                         const uint32_t* masks, uint32_t colonies, float* scores, int32_t* actions) {
    return guarded_score([&] {
        if (!controller || !scores || !actions) throw std::invalid_argument("null argument");
        faco::validate_controller(*controller);
        faco::validate_features(features, masks, colonies);
        for (uint32_t c = 0; c < colonies; ++c)
            actions[c] = faco::controller_decide_host(*controller, features, colonies, c, masks[c],
                                                      scores + static_cast<size_t>(c) * 32);
    });
}

extern "C" int faco_score_controller_gpu(const faco_controller* controller, const float* features,
                                         const uint32_t* masks, uint32_t colonies, float* scores, int32_t* actions) {
    return guarded_score([&] {
        if (!controller || !scores || !actions) throw std::invalid_argument("null argument");
        faco::validate_controller(*controller);
        faco::validate_features(features, masks, colonies);
        auto check = [](hipError_t e, const char* what) {
            if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
        };
        faco_controller* d_ctrl = nullptr; float* d_feat = nullptr; uint32_t* d_mask = nullptr;
        float* d_scores = nullptr; int32_t* d_actions = nullptr;
        const size_t fbytes = static_cast<size_t>(colonies) * 12 * 32 * sizeof(float);
        struct Frees { void* p[5]; ~Frees() { for (void* q : p) if (q) (void)hipFree(q); } } frees{};
        check(hipMalloc(reinterpret_cast<void**>(&d_ctrl), sizeof(faco_controller)), "hipMalloc"); frees.p[0] = d_ctrl;
        check(hipMalloc(reinterpret_cast<void**>(&d_feat), fbytes), "hipMalloc"); frees.p[1] = d_feat;
        check(hipMalloc(reinterpret_cast<void**>(&d_mask), colonies * sizeof(uint32_t)), "hipMalloc"); frees.p[2] = d_mask;
        check(hipMalloc(reinterpret_cast<void**>(&d_scores), colonies * 32 * sizeof(float)), "hipMalloc"); frees.p[3] = d_scores;
        check(hipMalloc(reinterpret_cast<void**>(&d_actions), colonies * sizeof(int32_t)), "hipMalloc"); frees.p[4] = d_actions;
        check(hipMemcpy(d_ctrl, controller, sizeof(faco_controller), hipMemcpyHostToDevice), "upload");
        check(hipMemcpy(d_feat, features, fbytes, hipMemcpyHostToDevice), "upload");
        check(hipMemcpy(d_mask, masks, colonies * sizeof(uint32_t), hipMemcpyHostToDevice), "upload");
        check(faco::launch_score_controller(d_ctrl, d_feat, d_mask, colonies, d_scores, d_actions, nullptr),
              "controller launch");
        check(hipMemcpy(scores, d_scores, colonies * 32 * sizeof(float), hipMemcpyDeviceToHost), "download");
        check(hipMemcpy(actions, d_actions, colonies * sizeof(int32_t), hipMemcpyDeviceToHost), "download");
    });
}
