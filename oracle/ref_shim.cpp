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
#include "gp_faco/baseline_policy.hpp"
#include "gp_faco/control_cpu.hpp"
#include "gp_faco/faco_cpu.hpp"
#include "gp_faco/factorial_policy.hpp"
#include "gp_faco/program.hpp"
#include "gp_faco/sparse_graph.hpp"

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

// Fixed-width rows as the reference's CandidateRows; entries >= n (a graph
// spec's right padding) end a row.
gp_faco::CandidateRows rows_of(const std::uint32_t* flat, Node n, Node width) {
    gp_faco::CandidateRows rows(n);
    for (Node a = 0; a < n; ++a)
        for (Node j = 0; j < width && flat[static_cast<std::size_t>(a) * width + j] < n; ++j)
            rows[a].push_back(flat[static_cast<std::size_t>(a) * width + j]);
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

// Hard mode on the reference's own predicate path: checklist_two_opt with
// its SparseUndirectedGraph (built from the edge list) as EdgeAllowed
// (cpp/src/faco_cpu.cpp:180-183). counters: processed, checks, evaluations,
// accepted, reactivations, constraint rejections, limit reached.
int ref_two_opt_hard(Node n, const double* matrix, const std::uint32_t* ls, Node lw, std::uint32_t* tour,
                     const std::uint32_t* checklist_in, Node checklist_size, std::uint64_t limit,
                     const std::uint32_t* edges, std::uint64_t edge_count, double* cost, std::uint64_t* counters) {
    return guarded([&] {
        const auto distance = matrix_distance(matrix, n);
        std::vector<gp_faco::Edge> list(edge_count);
        for (std::uint64_t i = 0; i < edge_count; ++i) list[i] = {edges[2 * i], edges[2 * i + 1]};
        const gp_faco::SparseUndirectedGraph graph(n, list);
        gp_faco::CpuTour t(std::vector<Node>(tour, tour + n), distance);
        std::vector<Node> checklist(checklist_in, checklist_in + checklist_size);
        gp_faco::DistanceOrderedCandidates candidates(rows_of(ls, n, lw), distance);
        const auto stats = t.checklist_two_opt(candidates, checklist, limit, graph.view());
        std::copy(t.order().begin(), t.order().end(), tour);
        *cost = t.recomputed_cost();
        const std::uint64_t values[7] = {stats.processed_nodes, stats.candidate_checks, stats.move_evaluations,
            stats.accepted_moves, stats.reactivations, stats.constraint_rejections, stats.evaluation_limit_reached};
        std::copy(values, values + 7, counters);
    });
}

// Hard-mode select_next: a city is available only when its relocation after
// `current` keeps `tour` inside the graph, the reference's node predicate
// (cuda/faco_choices.cuh:46-53 over edge_constraints.hpp relocation_allowed).
int ref_select_next_hard(Node n, const double* matrix, Node pw, Node bw, const std::uint32_t* primary_row,
                         const std::uint32_t* backup_row, const double* products, const std::uint8_t* visited,
                         Node current, double uniform, const std::uint32_t* tour, const std::uint32_t* edges,
                         std::uint64_t edge_count, std::uint32_t* chosen, int* stage) {
    return guarded([&] {
        const auto distance = matrix_distance(matrix, n);
        std::vector<gp_faco::Edge> list(edge_count);
        for (std::uint64_t i = 0; i < edge_count; ++i) list[i] = {edges[2 * i], edges[2 * i + 1]};
        const gp_faco::SparseUndirectedGraph graph(n, list);
        std::vector<Node> positions(n);
        for (Node i = 0; i < n; ++i) positions[tour[i]] = i;
        const auto view = graph.view();
        const auto allowed = [&](Node node) {
            return gp_faco::relocation_allowed(tour, positions.data(), n, current, node, view);
        };
        const auto result = gp_faco::select_next(current, std::vector<Node>(primary_row, primary_row + pw),
            std::vector<double>(products, products + pw), std::vector<Node>(backup_row, backup_row + bw),
            std::vector<std::uint8_t>(visited, visited + n), distance, uniform, allowed);
        *chosen = result.node;
        *stage = static_cast<int>(result.stage);
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


// ---------------------------------------------------------------------------
// Control layer (cpp/include/gp_faco/control_ops.hpp, control_cpu.cpp,
// baseline_policy.hpp, factorial_policy.hpp).

void ref_fingerprint(Node n, const std::uint32_t* tour, std::uint64_t* out) {
    std::vector<Node> positions(n);
    for (Node i = 0; i < n; ++i) positions[tour[i]] = i;
    const auto id = gp_faco::fingerprint({tour, positions.data(), n});
    out[0] = id.first; out[1] = id.second;
}

int ref_canonical_compare(Node n, const std::uint32_t* a, const std::uint32_t* b) {
    std::vector<Node> pa(n), pb(n);
    for (Node i = 0; i < n; ++i) { pa[a[i]] = i; pb[b[i]] = i; }
    return gp_faco::canonical_compare({a, pa.data(), n}, {b, pb.data(), n});
}

void ref_sample_nodes(Node n, std::uint64_t key, std::uint32_t* out) { gp_faco::sample_nodes(n, key, out); }

void ref_build_region(Node n, const std::uint32_t* tour, const std::uint32_t* primary, Node width,
                      std::uint64_t key, Node batch, Node mode, Node region, std::uint32_t* nodes) {
    std::vector<Node> positions(n);
    for (Node i = 0; i < n; ++i) positions[tour[i]] = i;
    gp_faco::build_region({tour, positions.data(), n}, primary, width, key, batch, mode, region, nodes);
}

struct RefFeedback { float return_rate, ls_work; std::uint64_t stagnant, epoch_batches, restarts; };

gp_faco::FeedbackState feedback_of(const RefFeedback& f) {
    gp_faco::FeedbackState s;
    s.return_rate = f.return_rate; s.ls_work = f.ls_work; s.stagnant_batches = f.stagnant;
    s.epoch_batches = f.epoch_batches; s.restarts = f.restarts;
    return s;
}

struct RefBaseline {
    std::uint32_t kind, mne_level, max_mne_level, region, restart_mode, restart_period;
    double restart_probability;
    std::uint32_t stagnation_step, restart_stagnation, restart_cooldown, pad;
};

gp_faco::BaselinePolicy policy_of(const RefBaseline& b) {
    gp_faco::BaselinePolicy p;
    p.kind = b.kind ? gp_faco::BaselineKind::Rule : gp_faco::BaselineKind::Static;
    p.mne_level = b.mne_level; p.max_mne_level = b.max_mne_level; p.region = b.region;
    p.restart_mode = b.restart_mode == 0 ? gp_faco::StaticRestart::None :
        b.restart_mode == 1 ? gp_faco::StaticRestart::Periodic : gp_faco::StaticRestart::Bernoulli;
    p.restart_period = b.restart_period; p.restart_probability = b.restart_probability;
    p.stagnation_step = b.stagnation_step; p.restart_stagnation = b.restart_stagnation;
    p.restart_cooldown = b.restart_cooldown;
    return p;
}

int ref_validate_baseline(const RefBaseline* b, std::uint32_t mask) {
    return guarded([&] { gp_faco::validate_baseline(policy_of(*b), mask); });
}

Node ref_baseline_action(const RefBaseline* b, const RefFeedback* f, Node batch, double uniform, std::uint32_t mask) {
    return gp_faco::baseline_action(policy_of(*b), feedback_of(*f), batch, uniform, mask);
}

std::uint32_t ref_factorial_mask(std::uint32_t variant, const RefBaseline* b, const RefFeedback* f, Node batch,
                                 double uniform, std::uint32_t legal) {
    gp_faco::FactorialPolicy p;
    p.variant = static_cast<gp_faco::FactorialVariant>(variant);
    p.baseline = policy_of(*b);
    return gp_faco::factorial_mask(p, feedback_of(*f), batch, uniform, legal);
}

void ref_update_feedback(RefFeedback* f, int improved, double returned, double work, Node ants) {
    auto s = feedback_of(*f);
    gp_faco::update_feedback(s, improved != 0, returned, work, ants);
    *f = {s.return_rate, s.ls_work, s.stagnant_batches, s.epoch_batches, s.restarts};
}

void ref_restart_feedback(RefFeedback* f) {
    auto s = feedback_of(*f);
    gp_faco::restart_feedback(s);
    *f = {s.return_rate, s.ls_work, s.stagnant_batches, s.epoch_batches, s.restarts};
}

namespace {
std::vector<gp_faco::TourRecord> records(const std::uint32_t* tours, Node count, Node n,
                                         const gp_faco::DistanceFunction& distance) {
    std::vector<gp_faco::TourRecord> out;
    for (Node i = 0; i < count; ++i)
        out.emplace_back(std::vector<Node>(tours + static_cast<std::size_t>(i) * n,
                                           tours + static_cast<std::size_t>(i + 1) * n), distance);
    return out;
}
}  // namespace

// ReferenceArchive(global).update(global, entries) then .update(global,
// completed): the archive built from `entries`, then one batch update.
int ref_archive_update(Node n, const double* matrix, const std::uint32_t* global, Node entries,
                       const std::uint32_t* entry_tours, Node completed, const std::uint32_t* completed_tours,
                       std::uint32_t* out_tours, Node* out_count) {
    return guarded([&] {
        const auto distance = matrix_distance(matrix, n);
        const gp_faco::TourRecord g(std::vector<Node>(global, global + n), distance);
        gp_faco::ReferenceArchive archive(g);
        archive.update(g, records(entry_tours, entries, n, distance));
        archive.update(g, records(completed_tours, completed, n, distance));
        const auto& chosen = archive.entries();
        for (Node i = 0; i < chosen.size(); ++i)
            std::copy(chosen[i].order.begin(), chosen[i].order.end(), out_tours + static_cast<std::size_t>(i) * n);
        *out_count = static_cast<Node>(chosen.size());
    });
}

// make_action_snapshot (cpp/src/control_cpu.cpp:75-120) with the archive
// entries built as in ref_archive_update from `archive_tours` (slot 0 =
// global) and a SparsePheromone holding `trails` / `default_trail`.
int ref_action_snapshot(Node n, const double* matrix, Node pw, const std::uint32_t* primary,
                        const double* local_scale, double epsilon, const std::uint32_t* samples, Node sample_count,
                        const std::uint32_t* active, Node archive_size, const std::uint32_t* archive_tours,
                        double pheromone_initial, double retention, double evaporation_floor, Node deposits,
                        const std::uint32_t* deposit_parents, const double* deposit_amounts, double minimum,
                        double maximum, const RefFeedback* feedback, double elapsed, std::uint64_t key, Node batch,
                        std::uint32_t experiment_mask, float* features, std::uint32_t* regions,
                        Node* alternative, std::uint32_t* mask, double* trails_out, double* default_out,
                        std::uint32_t* archive_out, Node* archive_count) {
    return guarded([&] {
        const auto distance = matrix_distance(matrix, n);
        gp_faco::ControlStaticData problem{n, pw, std::vector<Node>(primary, primary + static_cast<std::size_t>(n) * pw),
            std::vector<Node>(samples, samples + sample_count), std::vector<double>(local_scale, local_scale + n), epsilon};
        const auto rows = rows_of(primary, n, pw);
        gp_faco::SparsePheromone pheromone(rows, pheromone_initial, true);
        for (Node s = 0; s < deposits; ++s) {
            const std::uint32_t* parent = deposit_parents + static_cast<std::size_t>(s) * n;
            pheromone.evaporate(retention, evaporation_floor);
            for (Node i = 0; i < n; ++i)
                pheromone.deposit(parent[(i + n - 1) % n], parent[i], deposit_amounts[s], 1e300);
        }
        for (Node a = 0; a < n; ++a) for (Node j = 0; j < pw; ++j)
            trails_out[static_cast<std::size_t>(a) * pw + j] = pheromone.get(a, primary[static_cast<std::size_t>(a) * pw + j]);
        *default_out = pheromone.default_value();
        const auto all = records(archive_tours, archive_size, n, distance);
        gp_faco::ReferenceArchive archive(all[0]);
        archive.update(all[0], std::vector<gp_faco::TourRecord>(all.begin() + 1, all.end()));
        *archive_count = static_cast<Node>(archive.entries().size());
        for (Node i = 0; i < archive.entries().size(); ++i)
            std::copy(archive.entries()[i].order.begin(), archive.entries()[i].order.end(),
                      archive_out + static_cast<std::size_t>(i) * n);
        const gp_faco::TourRecord act(std::vector<Node>(active, active + n), distance);
        gp_faco::FeedbackState fb = feedback_of(*feedback);
        const auto snap = gp_faco::make_action_snapshot(problem, act, archive, fb, pheromone, {minimum, maximum},
                                                        distance, elapsed, key, batch, experiment_mask);
        std::copy(snap.features.begin(), snap.features.end(), features);
        for (Node m = 0; m < 2; ++m) for (Node r = 0; r < 4; ++r)
            std::copy(snap.regions.nodes[m][r], snap.regions.nodes[m][r] + 16, regions + (m * 4 + r) * 16);
        *alternative = snap.alternative_slot;
        *mask = snap.legal_mask;
    });
}

}  // extern "C"
