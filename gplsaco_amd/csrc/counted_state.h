// CountedState: the snapshot container of a controlled run after a whole
// number of batches (reference cpp/include/gp_faco/counted_state.hpp:1-48,
// cpp/src/counted_state.cpp:1-168). The container layout is the reference's
// (little-endian 64-bit fields, the same magic, field order and validation
// rules); the named buffers hold this engine's own device ABI bytes, so a
// snapshot resumes only in an engine of this library.
#pragma once

#include <stdint.h>

#include <string>
#include <utility>
#include <vector>

#include "../../include/faco_engine.h"

namespace faco {

struct StateBuffer {
    std::string name;
    uint64_t element_bytes = 0, planes = 1;
    std::vector<uint8_t> bytes;
};

struct StateIncumbent {
    double cost = 0;
    std::vector<uint32_t> tour;
};

struct CountedState {
    uint32_t dimension = 0, colonies = 0;
    faco_settings settings{};
    uint32_t experiment_mask = 0xffffffffu;
    uint64_t completed_batches = 0, progress_evaluation_limit = 0;
    std::vector<std::pair<uint64_t, uint64_t>> tasks;   // (instance key, seed) per colony
    std::vector<StateIncumbent> incumbents;             // per colony
    std::vector<StateBuffer> buffers;

    // Structure checks of the reference (counted_state.cpp:53-76); throws
    // std::invalid_argument.
    void validate() const;
    std::vector<uint8_t> serialize() const;
    size_t serialized_size() const;
    static CountedState deserialize(const uint8_t* bytes, size_t size);
    // The snapshot of one colony: every buffer's planes cut to that colony's
    // slice (counted_state.cpp:153-165).
    CountedState select_colony(uint32_t colony) const;
    const StateBuffer& buffer(const char* name) const;
};

bool same_settings(const faco_settings& a, const faco_settings& b);

}  // namespace faco
