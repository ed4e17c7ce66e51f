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
    // Hard mode: the registered undirected graph as CSR (offsets [n + 1],
    // sorted neighbours), empty for an unrestricted instance
    std::vector<uint32_t> graph_offsets, graph_neighbors;
    // Escape mode: 1/d^beta of every backup slot [n][bw] (0 for padding)
    std::vector<double> eta_backup;
};

faco_settings normalized_settings(faco_settings settings, uint32_t n);
Prepared prepare_problem(const std::vector<int32_t>& xy, const faco_settings& settings, uint32_t K,
                         uint32_t nbr_stride);
// The stages of prepare_problem, shared with the device preparation:
// coordinates, cheap tour and scale floor; the candidate keys of every city
// ((distance << 16) | city of its `needed` nearest others, ascending: the
// reference's (distance, id) order); rows, heuristic and local scales from
// those keys; the nearest-neighbour tour from city 0 (lowest id on ties);
// and the initial tour after its checklist 2-opt (the cheap tour when that
// is no longer).
uint32_t candidate_count(const faco_settings& settings, uint32_t n);
Prepared prepare_begin(const std::vector<int32_t>& xy, const faco_settings& settings);
void candidate_keys(const Prepared& p, uint32_t needed, uint32_t begin, uint32_t end, uint32_t* keys);
void prepare_rows(Prepared& p, const faco_settings& settings, uint32_t K, uint32_t nbr_stride, const uint32_t* keys,
                  uint32_t needed);
std::vector<uint32_t> nearest_neighbour_tour(const Prepared& p);
void prepare_finish(Prepared& p, std::vector<uint32_t> tour);
bool is_permutation(const uint32_t* tour, uint32_t n);
int64_t tour_cost(const Prepared& p, const uint32_t* tour);
// The 64 sampled cities of an instance for the control features ([64],
// first min(n, 64) valid).
std::vector<uint32_t> sample_nodes(uint32_t n, uint64_t instance_key);

}  // namespace faco
