// Host problem preparation of the FACO engine (the step before the search
// path): settings normalisation, distance-ordered candidate rows, the
// nearest-neighbour tour and its initial checklist 2-opt.
// Semantics: reference cpp/src/prepared_problem.cpp:19-126.
#pragma once

#include <stdint.h>

#include <vector>

#include "../../include/faco_engine.h"

namespace faco {

struct Prepared {
    uint32_t n = 0;
    std::vector<int32_t> xy;           // integer coordinates [n][2]
    std::vector<uint32_t> packed;      // x | y << 16
    std::vector<uint32_t> primary, backup, ls;   // [n][width]
    std::vector<uint16_t> nbr;         // [n][K] rows (first max(pw, lw) nearest), padded to the stride
    std::vector<uint16_t> nbr_dist;    // EUC_2D distance of each nbr entry (same layout)
    std::vector<uint16_t> backup16;    // [n][bw]
    std::vector<double> eta;           // [n][pw] heuristic 1/d^beta (1 when d == 0)
    std::vector<double> local_scale;   // mean of the 8 nearest distances
    double scale_epsilon = 1e-12;      // floor of local_scale in the control features
    std::vector<uint32_t> cheap, initial;
    int64_t cheap_cost = 0, initial_cost = 0;
    double cheap_seconds = 0, preparation_seconds = 0;
    double profile[4] = {0, 0, 0, 0};  // candidates, nearest neighbour, initial LS, finalisation
    uint32_t s_lw = 0;                 // LS row width of `ls`
    std::vector<float> prior;          // [n][pw] heuristic multipliers of a learned edge prior, or empty
};

faco_settings normalized_settings(faco_settings settings, uint32_t n);
Prepared prepare_problem(const std::vector<int32_t>& xy, const faco_settings& settings, uint32_t K,
                         uint32_t nbr_stride);
bool is_permutation(const uint32_t* tour, uint32_t n);
int64_t tour_cost(const Prepared& p, const uint32_t* tour);
// The 64 sampled cities of an instance for the control features ([64],
// first min(n, 64) valid).
std::vector<uint32_t> sample_nodes(uint32_t n, uint64_t instance_key);

}  // namespace faco
