// C-ABI shim over the reference's own CPU core, built from its sources where
// they lie (see oracle/Makefile, target `ref`). TEST INFRASTRUCTURE ONLY: the
// resulting oracle/_ref/libgpfaco_ref.so is used by tests/ to pin the oracle
// restatement against the reference implementation; nothing in the product
// links or loads it, and it is never built on a machine without the
// reference checkout.
//
// The reference classes take an arbitrary DistanceFunction
// (cpp/include/gp_faco/faco_cpu.hpp:13), so the shim hands them the same
// integer EUC_2D distances the oracle uses, as FP64 values.
#include "gp_faco/faco_cpu.hpp"
#include "gp_faco/program.hpp"

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace {
thread_local std::string g_error;
template<class F> int guarded(F f) {
    try { f(); return 0; }
    catch (const std::invalid_argument& e) { g_error = e.what(); return 1; }
    catch (const std::exception& e) { g_error = e.what(); return 2; }
}
using gp_faco::Node;

gp_faco::DistanceFunction matrix_distance(const double* matrix, Node n) {
    return [matrix, n](Node a, Node b) { return matrix[static_cast<std::size_t>(a) * n + b]; };
}

gp_faco::CandidateRows rows_of(const std::uint32_t* flat, Node n, Node width) {
    gp_faco::CandidateRows rows(n);
    for (Node a = 0; a < n; ++a) rows[a].assign(flat + static_cast<std::size_t>(a) * width,
                                                flat + static_cast<std::size_t>(a + 1) * width);
    return rows;
}
}  // namespace

extern "C" {

const char* ref_last_error() { return g_error.c_str(); }

// FocusedConstruction + checklist_two_opt on the reference classes
// (tests/cpp/test_cuda_faco.cpp:78-104 pattern).
int ref_construct_ls(Node n, const double* matrix, const std::uint32_t* ls, Node lw,
                     const std::uint32_t* parent, const std::uint32_t* visit, Node mne_target,
                     std::uint64_t limit, std::uint32_t* construction_tour, std::uint32_t* tour,
                     std::uint32_t* positions, std::uint32_t* checklist, double* costs,
                     std::uint64_t* counters) {
    return guarded([&] {
        const auto distance = matrix_distance(matrix, n);
        gp_faco::CpuTour reference(std::vector<Node>(parent, parent + n), distance);
        gp_faco::FocusedConstruction cpu(reference, visit[0], mne_target);
        Node cursor = 1;
        while (!cpu.done()) cpu.step(visit[cursor++]);
        std::copy(cpu.tour().order().begin(), cpu.tour().order().end(), construction_tour);
        costs[0] = cpu.tour().cost();
        gp_faco::DistanceOrderedCandidates candidates(rows_of(ls, n, lw), distance);
        const auto stats = cpu.tour().checklist_two_opt(candidates, cpu.checklist(), limit);
        std::copy(cpu.tour().order().begin(), cpu.tour().order().end(), tour);
        std::copy(cpu.tour().positions().begin(), cpu.tour().positions().end(), positions);
        std::copy(cpu.checklist().begin(), cpu.checklist().end(), checklist);
        costs[1] = cpu.tour().cost();
        const std::uint64_t values[11] = {cpu.stats().mne, cpu.stats().steps,
            cpu.stats().nonidentity_relocations, cpu.stats().legal_exhausted, stats.processed_nodes,
            stats.candidate_checks, stats.move_evaluations, stats.accepted_moves, stats.reactivations,
            stats.evaluation_limit_reached, cpu.checklist().size()};
        std::copy(values, values + 11, counters);
    });
}

// checklist_two_opt alone from a given tour and checklist (used by problem
// preparation, cpp/src/prepared_problem.cpp:111-113).
int ref_two_opt(Node n, const double* matrix, const std::uint32_t* ls, Node lw, std::uint32_t* tour,
                const std::uint32_t* checklist_in, Node checklist_size, std::uint64_t limit,
                double* cost, std::uint64_t* evaluations) {
    return guarded([&] {
        const auto distance = matrix_distance(matrix, n);
        gp_faco::CpuTour t(std::vector<Node>(tour, tour + n), distance);
        std::vector<Node> checklist(checklist_in, checklist_in + checklist_size);
        gp_faco::DistanceOrderedCandidates candidates(rows_of(ls, n, lw), distance);
        const auto stats = t.checklist_two_opt(candidates, checklist, limit);
        std::copy(t.order().begin(), t.order().end(), tour);
        *cost = t.recomputed_cost();
        *evaluations = stats.move_evaluations;
    });
}

int ref_select_next(Node n, const double* matrix, Node pw, Node bw, const std::uint32_t* primary_row,
                    const std::uint32_t* backup_row, const double* products, const std::uint8_t* visited,
                    Node current, double uniform, std::uint32_t* chosen, int* stage) {
    return guarded([&] {
        const auto distance = matrix_distance(matrix, n);
        const auto result = gp_faco::select_next(current, std::vector<Node>(primary_row, primary_row + pw),
            std::vector<double>(products, products + pw), std::vector<Node>(backup_row, backup_row + bw),
            std::vector<std::uint8_t>(visited, visited + n), distance, uniform);
        *chosen = result.node;
        *stage = static_cast<int>(result.stage);
    });
}

int ref_trail_limits(Node count, double p_best, double retention, double cost, double* minimum,
                     double* maximum) {
    return guarded([&] {
        const auto limits = gp_faco::candidate_trail_limits(count, p_best, retention, cost);
        *minimum = limits.minimum; *maximum = limits.maximum;
    });
}

// SparsePheromone from a uniform start, then `steps` rounds of evaporate +
// symmetric deposit along parents[step] (cpp/src/faco_cpu.cpp:306-343) and the
// product cache (:350-365) after each round, as the reference replay test
// drives it (tests/cpp/test_fixed_faco.cpp:116-125).
int ref_pheromone_run(Node n, const double* matrix, const std::uint32_t* primary, Node pw,
                      double initial, Node steps, const std::uint32_t* parents, double retention,
                      const double* minima, const double* maxima, const double* increments,
                      double beta, double* trails_out, double* defaults_out, double* products_out) {
    return guarded([&] {
        gp_faco::SparsePheromone pheromone(rows_of(primary, n, pw), initial, true);
        const auto cells = static_cast<std::size_t>(n) * pw;
        for (Node s = 0; s < steps; ++s) {
            const std::uint32_t* parent = parents + static_cast<std::size_t>(s) * n;
            pheromone.evaporate(retention, minima[s]);
            for (Node i = 0; i < n; ++i) pheromone.deposit(parent[(i + n - 1) % n], parent[i], increments[s], maxima[s]);
            for (Node a = 0; a < n; ++a) for (Node j = 0; j < pw; ++j)
                trails_out[s * cells + static_cast<std::size_t>(a) * pw + j] =
                    pheromone.get(a, primary[static_cast<std::size_t>(a) * pw + j]);
            defaults_out[s] = pheromone.default_value();
            const auto cache = pheromone.product_cache(matrix_distance(matrix, n), beta);
            std::copy(cache.begin(), cache.end(), products_out + s * cells);
        }
    });
}

// score_cpu (cpp/src/program_cpu.cpp:79-116) on a raw Program record.
int ref_score_cpu(const std::uint8_t* opcode, const std::uint8_t* operand, Node length,
                  const float* constants, Node constants_count, std::uint16_t feature_spec_id,
                  const float* features, const std::uint32_t* masks, Node colonies, float* scores,
                  std::int32_t* actions) {
    return guarded([&] {
        gp_faco::Program p;
        p.feature_spec_id = feature_spec_id;
        p.length = static_cast<std::uint16_t>(length);
        p.constants_count = static_cast<std::uint16_t>(constants_count);
        std::copy(opcode, opcode + length, p.opcode);
        std::copy(operand, operand + length, p.operand);
        std::copy(constants, constants + constants_count, p.constants);
        const auto out = gp_faco::score_cpu(p, std::vector<float>(features, features + 12 * 32 * colonies),
                                            std::vector<std::uint32_t>(masks, masks + colonies));
        std::copy(out.scores.begin(), out.scores.end(), scores);
        std::copy(out.actions.begin(), out.actions.end(), actions);
    });
}

}  // extern "C"
