/*
 * faco_engine.h - C ABI of the MI355X FACO batch engine (libfaco_hip.so).
 *
 * This is the drop-in boundary for the reference's Python->native module
 * gp_faco_ext (reference cpp/src/bindings.cpp:500-1046). Every entry point
 * takes plain pointers and sizes, owns no caller memory and reports errors as
 *   FACO_OK (0), FACO_INVALID_ARGUMENT (1) -> Python ValueError,
 *   FACO_RUNTIME_ERROR (2) -> Python RuntimeError,
 * with a thread-local message from faco_last_error(). This mirrors the
 * reference's std::invalid_argument / std::runtime_error split
 * (cuda/colony_state.cuh:10-15).
 *
 * Distances: TSPLIB EUC_2D, nint(sqrt(dx^2+dy^2)), on integer coordinates in
 * [0, 32767]. Coordinates cross the boundary as FP64 [n][2] like the
 * reference's; non-integral or out-of-range values are rejected.
 */
#ifndef FACO_ENGINE_H
#define FACO_ENGINE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FACO_OK 0
#define FACO_INVALID_ARGUMENT 1
#define FACO_RUNTIME_ERROR 2

/* FixedFacoSettings, same field order and 64-byte layout
 * (reference cpp/include/gp_faco/fixed_faco_gpu.hpp:10-21). ants == 0 means
 * 64 * ceil(sqrt(n) / 16) (reference cpp/src/prepared_problem.cpp:19-33). */
typedef struct faco_settings {
    uint32_t ants;
    uint32_t primary_width;
    uint32_t backup_width;
    uint32_t ls_width;
    double beta;
    double retention;
    double p_best;
    double epoch_source_probability;
    uint64_t ls_evaluation_limit;
    uint64_t initial_ls_evaluation_limit;
} faco_settings;

/* Per-ant diagnostic record, the counters of FacoDiagnosticInfo
 * (reference cpp/include/gp_faco/faco_cuda_diagnostic.hpp:17-25,
 * cpp/include/gp_faco/faco_cpu.hpp:34-42,70-75). Costs are exact integers. */
typedef struct faco_ant_info {
    int64_t construction_cost;
    int64_t final_cost;
    uint32_t mne;
    uint32_t steps;
    uint32_t nonidentity_relocations;
    uint32_t legal_exhausted;
    uint32_t processed_nodes;
    uint32_t candidate_checks;
    uint32_t move_evaluations;
    uint32_t accepted_moves;
    uint32_t reactivations;
    uint32_t evaluation_limit_reached;
    uint32_t checklist_size;
    uint32_t start_node;
    uint32_t constraint_rejections; /* Hard mode: evaluated 2-opt moves that would leave the graph */
    uint32_t pad;
} faco_ant_info;

/* Batch result summary, the scalar fields of the reference's batch_output
 * dict (reference cpp/src/bindings.cpp:277-410). */
#define FACO_STAGE_COUNT 12
typedef struct faco_result {
    uint64_t evaluation_limit_per_colony;
    uint64_t completed_tour_evaluations_per_colony;
    uint64_t total_tour_evaluations;
    uint64_t launched_batches;
    uint64_t completed_batches;
    uint64_t completed_construction_steps;
    uint64_t completed_ls_evaluations;
    uint64_t allocated_device_bytes;
    double actual_seconds;
    double search_seconds;       /* host wall clock of the iteration loop */
    double search_gpu_milliseconds; /* device time of the iteration loop (HIP events) */
    double kernel_milliseconds[FACO_STAGE_COUNT]; /* filled when profiling is on */
    uint64_t completed_constraint_rejections;     /* Hard mode: evaluated 2-opt moves that would leave the graph
                                                     (reference faco_cpu.cpp:180-183) */
    uint64_t persistent_launches; /* launches of the persistent kernels (many batches per colony workgroup) */
} faco_result;

/* Program IR v1 record (reference cpp/include/gp_faco/program.hpp:10-19). */
typedef struct faco_program {
    uint16_t ir_version;
    uint16_t numeric_spec_id;
    uint16_t feature_spec_id;
    uint16_t length;
    uint16_t constants_count;
    uint8_t opcode[63];
    uint8_t operand[63];
    float constants[63];
} faco_program;

/* ControllerProgram (reference cpp/include/gp_faco/program.hpp:27-30),
 * 1168 bytes: kind 0 = one tree scoring the 32 actions; kind 1 = conditional
 * trees choosing restart (2-way), region (4-way) and MNE level (4-way). */
typedef struct faco_controller {
    uint32_t kind;
    faco_program trees[3];
} faco_controller;

/* BaselinePolicy, same 48-byte layout (reference
 * cpp/include/gp_faco/baseline_policy.hpp:20-27). kind: 0 static, 1 rule;
 * restart_mode: 0 none, 1 periodic, 2 bernoulli. */
typedef struct faco_baseline_policy {
    uint32_t kind;
    uint32_t mne_level;
    uint32_t max_mne_level;
    uint32_t region;
    uint32_t restart_mode;
    uint32_t restart_period;
    double restart_probability;
    uint32_t stagnation_step;
    uint32_t restart_stagnation;
    uint32_t restart_cooldown;
    uint32_t pad;
} faco_baseline_policy;

/* Per-colony control summary after a controlled run: the fields of
 * ControllerState the reference exports as control_states
 * (cpp/src/bindings.cpp:394-407; control_ops.hpp:160-173). */
typedef struct faco_control_state {
    uint32_t archive_size;
    uint32_t pad;
    uint64_t stagnant_batches;
    uint64_t epoch_batches;
    uint64_t restarts;
    float return_rate;
    float ls_work;
} faco_control_state;

/* Decision source of a controlled evaluation. */
#define FACO_CONTROL_PROGRAM 1u    /* GP program(s) score the 32 actions */
#define FACO_CONTROL_CONTROLLER 2u /* ControllerProgram(s) */
#define FACO_CONTROL_BASELINE 3u   /* fixed baseline policy */
#define FACO_CONTROL_FACTORIAL 4u  /* baseline-restricted mask, then program(s) */

typedef struct faco_control_spec {
    uint32_t mode;                       /* FACO_CONTROL_* */
    uint32_t deciders;                   /* programs or controllers: 1, or the population size */
    const faco_program* programs;        /* [deciders] for PROGRAM and FACTORIAL */
    const faco_controller* controllers;  /* [deciders] for CONTROLLER */
    faco_baseline_policy baseline;       /* for BASELINE and FACTORIAL */
    uint32_t factorial_variant;          /* 0 M00, 1 M10, 2 M01, 3 M11 */
    uint32_t experiment_mask;            /* legal actions; bits 0..15 must not all be clear */
    uint32_t force_fingerprint_collisions; /* diagnostic: every tour identity reads as zero */
    uint32_t pad;
} faco_control_spec;

/* Decision trace at chosen batches (the reference's decision_iterations,
 * bindings.cpp:721-757, with batch b = iteration - 1): for the k-th listed
 * batch (ascending), features [k][12][C][32], scores [k][C][32], masks,
 * actions and alternatives [k][C], start regions [k][C][2][4][16] and the
 * (epoch, global, parent) costs after that batch [k][C][3]. Any output
 * pointer may be NULL. */
typedef struct faco_control_trace {
    const uint32_t* batches;
    uint32_t count;
    uint32_t pad;
    float* features;
    float* scores;
    uint32_t* masks;
    int32_t* actions;
    uint32_t* alternatives;
    uint32_t* regions;
    int64_t* costs;
} faco_control_trace;

typedef struct faco_prior faco_prior;
typedef struct faco_engine faco_engine;
typedef struct faco_fixed faco_fixed;

/* Thread-local message of the last failing call on this thread. */
const char* faco_last_error(void);

/* Device query (reference cpp/src/bindings.cpp:1031-1044, cuda_device_info). */
int faco_device_info(int* device, char* name, size_t name_size, char* arch, size_t arch_size,
                     int* compute_units, size_t* total_memory, size_t* max_lds_per_block);

/* prepare_common_initial(coordinates, settings) (reference
 * cpp/src/bindings.cpp:522-540): host-only preparation of one instance.
 * Outputs the initial tour [n] and cost, the normalised settings and, when
 * not NULL, the candidate rows primary [n][pw], backup [n][bw], ls [n][lw]
 * and local_scale [n]. Needs no GPU. */
int faco_prepare_common_initial(const double* xy, uint32_t n, const faco_settings* settings,
                                faco_settings* normalized, uint32_t* tour, int64_t* cost,
                                double* cheap_seconds, double* preparation_seconds, uint32_t* primary,
                                uint32_t* backup, uint32_t* ls, double* local_scale);

/* FacoBatchEngine(dimension, colonies, settings, constraint_mode, population_size)
 * (reference cpp/include/gp_faco/batch_engine.hpp:68-70; bindings.cpp:557-573).
 * constraint_mode: 0 unrestricted, 1 Hard (population 1; instances come
 * through faco_engine_register_graph_problem), 2 Escape (the Hard graph
 * with escape gates: reference cuda/escape_device.cuh:29-128,
 * batch_engine.cu:1014-1025; at most 16 primary, 64 backup and 20 LS slots,
 * fewer than 32768 cities, no learned prior). population_size in [1, 128]. */
int faco_engine_create(uint32_t dimension, uint32_t colonies, const faco_settings* settings,
                       uint32_t constraint_mode, uint32_t population_size, faco_engine** out);
/* Mixed-size batch engine: instances of any n up to max_dimension in one
 * panel (all with the same candidate widths), every colony run with its own
 * n and ants (the settings' per-size default when ants = 0) in the same
 * launches; workgroups are dispatched largest work first. No counterpart in
 * the reference, which needs one n per engine (cuda/batch_engine.cu:78-99;
 * python/gp_faco/worker.py:41-53). population_size as for
 * faco_engine_create. Evaluations then need an evaluation limit that is a
 * whole number of every colony's ant batches; each colony stops at its own
 * budget, and each row of `tours` holds the colony's n cities followed by
 * UINT32_MAX. */
int faco_engine_create_mixed(uint32_t max_dimension, uint32_t colonies, const faco_settings* settings,
                             uint32_t population_size, faco_engine** out);
void faco_engine_destroy(faco_engine* engine);
/* The normalised settings the engine runs with. */
int faco_engine_settings(const faco_engine* engine, faco_settings* out);

/* Launch shape of the ant kernel: ants (waves) per workgroup, whether the
 * neighbour rows are staged in LDS, the workgroup's LDS bytes, the
 * workgroups per colony for a run of `colonies` colonies (each workgroup's
 * waves share one chunk of a colony's ants dynamically) and the resident
 * workgroups per CU. No
 * counterpart in the reference (its warps per block is a compile-time
 * choice, CMakeLists.txt:13-19); set_waves(0) restores the automatic shape. */
int faco_engine_launch_shape(const faco_engine* engine, uint32_t colonies, uint32_t* waves, uint32_t* nbr_in_lds,
                             uint32_t* lds_bytes, uint32_t* blocks_per_colony, uint32_t* blocks_per_cu);
/* Force the ant kernel's launch shape: `waves` ants in flight per workgroup
 * (1..16; bit 8 (0x100) keeps the neighbour rows in global memory instead
 * of LDS) and `blocks_per_colony` workgroups per colony (must divide the
 * ants; 0 = automatic). waves = 0 restores the automatic shape. */
int faco_engine_set_waves(faco_engine* engine, uint32_t waves, uint32_t blocks_per_colony);

/* Test hook of the roulette (no counterpart in the reference): the ant
 * kernel decides a roulette step from FP32 tree-order prefix sums when no
 * available prefix lies within margin x row total of the threshold, and from
 * the exact left-to-right FP64 sums otherwise. The default margin is 2^-18,
 * above the 2^-20.4 bound on the two computations' difference; a wider margin
 * (up to +inf: left-to-right sums at every step) must give identical results. */
int faco_engine_set_roulette_margin(faco_engine* engine, double margin);

/* One timed launch shape of faco_engine_autotune. */
typedef struct faco_tune_entry {
    uint32_t waves;                 /* ants in flight per workgroup */
    uint32_t nbr_in_lds;            /* neighbour rows staged in LDS */
    uint32_t blocks_per_colony;     /* workgroups per colony chosen for the batch */
    uint32_t pad;                   /* autotune_control: 1 when the shape runs the persistent controlled kernel */
    double ms_per_iteration;        /* ant kernel + colony update, HIP events */
} faco_tune_entry;

/* Autotuner of the ant kernel's launch shape (no counterpart in the
 * reference, where warps per block is a compile-time choice,
 * CMakeLists.txt:13-19). Times `iterations` eager iterations of plain FACO
 * on the given tasks (any number up to colonies x population_size, so a
 * population engine is tuned at its full shape) for every shape that fits
 * the LDS budget (16, 12, 10,
 * 8, 6, 4, 2 or 1 ants per workgroup, neighbour rows in LDS or in global
 * memory), keeps the fastest as with faco_engine_set_waves and writes the
 * timings to `table` (up to `capacity` rows; `count` receives the number of
 * shapes timed). Later evaluations give identical results for every shape;
 * only their time changes. */
int faco_engine_autotune(faco_engine* engine, const uint64_t* instance_keys, const uint64_t* seeds, uint32_t tasks,
                         uint32_t iterations, faco_tune_entry* table, uint32_t capacity, uint32_t* count);

/* The autotuner on the controlled workload itself: every launch shape that
 * fits is timed on `iterations` batches of the controlled run `spec` over
 * the panel (search_gpu_milliseconds per batch), once as the persistent
 * controlled kernel where the shape allows it (entry pad = 1) and once as
 * graph-replayed per-batch launches; the fastest shape and form are kept
 * for the engine's later controlled runs (faco_engine_set_waves restores
 * the persistent form). No counterpart in the reference (its block shape is
 * a compile-time choice); results never depend on the shape or form. */
int faco_engine_autotune_control(faco_engine* engine, const uint64_t* panel_keys, const uint64_t* panel_seeds,
                                 uint32_t panel, const faco_control_spec* spec, uint32_t iterations,
                                 faco_tune_entry* table, uint32_t capacity, uint32_t* count);

/* Local search of the ants: 0 = the reference's checklist 2-opt (default),
 * 1 = checklist 2-opt + Or-opt. Or-opt has no counterpart in the reference:
 * after a popped city a's 2-opt scan, the segment of 1-3 cities starting at
 * a is tried at the edges next to each LS neighbour b that passed the
 * predecessor scan (d(a,b) < d(pred a, a)), forward at (b, succ b) and
 * reversed at (pred b, b); every such move counts one evaluation, the best
 * move over both operators wins (2-opt on ties) and an Or-opt move is
 * applied as two or three 2-opt reversals. */
int faco_engine_set_local_search(faco_engine* engine, uint32_t mode);

/* register_problem(instance_key, coordinates FP64[n][2])
 * (reference batch_engine.hpp:72, batch_engine.cu:279-295). Host preparation:
 * k-NN rows, NN tour, initial checklist 2-opt. */
int faco_engine_register_problem(faco_engine* engine, uint64_t instance_key, const double* xy,
                                 uint32_t n, double* cheap_seconds, double* preparation_seconds);
/* register_graph_problem(instance_key, coordinates, spec) for a Hard-mode
 * engine (constraint_mode 1; reference batch_engine.cu:296-314,
 * prepared_problem.cpp:128-166, bindings.cpp:574-622): the common
 * preparation plus the candidate graph, edges [edge_count][2] (undirected,
 * no self-loops or repeats, at most n * (primary_width + 1), holding the
 * initial tour), and the spec's own rows: primary [n][primary_width] and
 * LS [n][ls_width] members of the graph (LS rows ordered by distance, then
 * city), backup [n][backup_width] disjoint from the primary row; every row
 * right-padded with n. The spec's initial tour must be the prepared one.
 * In Hard mode every relocation and 2-opt move keeps the tour inside the
 * graph (construction exhausts instead of leaving it); plain 2-opt. */
int faco_engine_register_graph_problem(faco_engine* engine, uint64_t instance_key, const double* xy, uint32_t n,
                                       const uint32_t* edges, uint64_t edge_count, const uint32_t* initial_tour,
                                       const uint32_t* primary, const uint32_t* backup, const uint32_t* ls,
                                       double* cheap_seconds, double* preparation_seconds);
/* Device-side preparation of `count` instances (reference
 * prepared_problem.cpp:58-126, the step register_problem runs on the host):
 * the same k-NN rows, nearest-neighbour tour and initial checklist 2-opt,
 * bit for bit, with the O(n^2) candidate selection, the nearest-neighbour
 * walk and the 2-opt on the GPU. Instance i has key keys[i], sizes[i] cities
 * (sizes NULL: the engine's dimension) and coordinates xy concatenated in
 * order. Keys already registered must carry the same coordinates and are
 * skipped. *seconds (may be NULL): wall time of the whole call; each new
 * instance's preparation_seconds is its share. */
int faco_engine_register_problems_device(faco_engine* engine, const uint64_t* keys, const double* xy,
                                         const uint32_t* sizes, uint32_t count, double* seconds);
/* preparation_profile: candidates_and_scales, nearest_neighbor, initial_ls,
 * finalization seconds (reference bindings.cpp:639-650). */
int faco_engine_preparation_profile(faco_engine* engine, uint64_t instance_key, double out[4]);
/* Prepared data of a registered instance (for host callers that mirror
 * FixedFacoGpu::primary_candidates etc.). Any pointer may be NULL. */
int faco_engine_problem_data(faco_engine* engine, uint64_t instance_key, uint32_t* primary,
                             uint32_t* backup, uint32_t* ls, uint32_t* initial_tour, int64_t* initial_cost);

/* Flags of faco_engine_evaluate_faco. */
#define FACO_RUN_GRAPH 1u   /* replay the iteration body through a captured HIP graph */
#define FACO_RUN_PROFILE 2u /* eager launches with HIP events around each kernel:
                               kernel_milliseconds[0] = construction + LS, [1] = colony update,
                               summed over iterations; [2..4] = mean per-ant setup, construction
                               and local-search milliseconds (device wall clock), summed over
                               iterations; [5..9] = mean per-ant shader megaclocks of roulette
                               selection, relocation, MNE/checklist pushes, 2-opt scan and
                               2-opt apply, summed over iterations; [10] = mean per-ant
                               milliseconds from taking the ant to its write-back and [11] =
                               share of the resident waves' time spent holding an ant, both
                               summed over iterations */
#define FACO_RUN_EVENTS 4u  /* eager launches of the production kernels with HIP events around
                               each: kernel_milliseconds[0] and [1] as for FACO_RUN_PROFILE */
#define FACO_RUN_CONTINUE 8u /* resume the engine's last plain evaluation (same tasks and MNE
                               target, fixed-size engine, no checkpoints) from its last batch
                               instead of resetting the colonies: a run of a batches followed by
                               a continuation of b batches equals one run of a + b batches bit
                               for bit. Work counters and evaluations report the continuation
                               only. No counterpart in the reference (whose plain entry point
                               always starts from the prepared tour); a steady-state window of a
                               long run is measured this way */
#define FACO_RUN_NO_ESCAPE 16u /* Escape-mode engines: no escape gate opens, so the run must equal
                                  the Hard run of the same graph (reference
                                  BatchDiagnosticControls::disable_escape, batch_engine.hpp) */
/* Escape counters per colony, in the reference's EscapeStats order
 * (cpp/include/gp_faco/escape_slots.hpp:70-75; bindings.cpp:376-386):
 * construction_opportunities, construction_gates,
 * construction_replaced_slots, escape_relocations, ls_anchor_nodes,
 * ls_replaced_slots, old_view_reactivations, anchor_reactivations,
 * new_edges, construction_capacity_rejections, ls_capacity_rejections. */
#define FACO_ESCAPE_COUNTERS 11u

/* Curve checkpoints (reference cuda/batch_engine.cu:635-653, the
 * checkpoint_iterations argument of bindings.cpp:701-757,818-853): the
 * global-best tour of every colony after iterations[i] completed batches
 * (0 = after initialisation), written to tours[i][colonies][n]. The
 * iterations are strictly ascending and at most the run's batch count; a
 * mixed-size colony that has finished its own budget keeps its final tour
 * and its row ends in UINT32_MAX padding after its own n. */
typedef struct faco_checkpoints {
    const uint64_t* iterations;
    uint32_t count;
    uint32_t pad;
    uint32_t* tours;
} faco_checkpoints;

/* evaluate_faco_evaluations(keys, seeds, evaluation_limit_per_colony, mne)
 * (reference batch_engine.hpp:87-89, bindings.cpp:701-720): plain FACO on
 * `tasks` colonies for evaluation_limit / ants batches. Outputs per colony
 * the global-best tour [tasks][n] and cost [tasks]; `work` (may be NULL)
 * receives [tasks][2] construction steps and LS evaluations; `checkpoints`
 * (may be NULL) records the curve. */
int faco_engine_evaluate_faco(faco_engine* engine, const uint64_t* instance_keys, const uint64_t* seeds,
                              uint32_t tasks, uint64_t evaluation_limit_per_colony, uint32_t mne_target,
                              uint32_t flags, faco_result* result, uint32_t* tours, int64_t* costs,
                              uint64_t* work, const faco_checkpoints* checkpoints);

/* Controlled FACO: the GP-controlled restart and perturbation loop.
 * One entry point for the reference's evaluate_program_evaluations,
 * evaluate_controller_evaluations, evaluate_population_evaluations,
 * evaluate_baseline_evaluations and evaluate_factorial_evaluations
 * (bindings.cpp:721-879; batch_engine.cu:376-460,492-1192). The panel of
 * `panel` tasks is replicated once per decider: colony c = d * panel + t runs
 * task t under program / controller d (C = panel * deciders colonies, at
 * most the engine's colonies * population_size). Per batch: control features
 * and start regions, decision, restart transaction, construction + 2-opt,
 * selection and pheromone, archive and feedback update
 * (cuda/control_state.cuh). Outputs per colony: global tour [C][n], cost
 * [C], work [C][2] (may be NULL) and control summary [C] (may be NULL).
 * flags: FACO_RUN_GRAPH as for faco_engine_evaluate_faco (ignored while a
 * trace is taken); FACO_RUN_EVENTS (reference profile_events_only,
 * bindings.cpp:721-757): eager launches with HIP events between them,
 * result->kernel_milliseconds [0] construction + LS, [1] colony update
 * (reduction and pheromone), [2] decision (features, regions, scoring, action
 * and restart), [3] archive and feedback update, [4] initialisation, summed
 * over batches; results are unchanged. `checkpoints` (may be NULL) as for
 * faco_engine_evaluate_faco, over the C colonies. */
int faco_engine_evaluate_control(faco_engine* engine, const uint64_t* instance_keys, const uint64_t* seeds,
                                 uint32_t panel, uint64_t evaluation_limit_per_colony,
                                 const faco_control_spec* spec, uint32_t flags, const faco_control_trace* trace,
                                 faco_result* result, uint32_t* tours, int64_t* costs, uint64_t* work,
                                 faco_control_state* states, const faco_checkpoints* checkpoints);

/* ---------------------------------------------------------------------
 * Time-budget evaluations (reference cpp/include/gp_faco/deadline.hpp,
 * cuda/batch_engine.cu:340-362,492-1192 with a budget in seconds;
 * bindings.cpp:662-700 evaluate / evaluate_program and :937-971
 * run_program_diagnostic). The clock runs from the call's start; in
 * preparation mode 0 ("cached_charged") each instance's registered (or
 * frozen, faco_engine_set_preparation_charges) cheap and preparation
 * seconds are charged to it instead of preparing again, in mode 1
 * ("end_to_end") the preparation is redone inside the budget. Every
 * colony's incumbent is first the cheap tour, then the prepared initial
 * tour, then its global best after each whole batch that completes within
 * the budget (a batch still running at the deadline is discarded); a
 * later offer replaces the incumbent only when strictly shorter. */
typedef struct faco_deadline_spec {
    double budget_seconds;         /* finite, >= 0 */
    uint32_t mne_target;           /* plain FACO (program == NULL), > 0 */
    uint32_t preparation_mode;     /* 0 cached_charged, 1 end_to_end */
    const faco_program* program;   /* NULL: plain FACO (evaluate); else program control (evaluate_program),
                                      feature spec 1: the progress feature is elapsed / budget */
    uint32_t experiment_mask;      /* program control: legal actions */
    uint32_t fixed_batches;        /* 0: batches until the deadline; 1..10000: at most this many
                                      (run_program_diagnostic) */
    double fixed_elapsed_ratio;    /* < 0: elapsed / budget; in [0, 1]: this constant progress feature
                                      (needs fixed_batches) */
    uint32_t profile;              /* per-batch profile rows (needs fixed_batches and a fixed ratio) */
    uint32_t pad;
} faco_deadline_spec;

typedef struct faco_deadline_result {
    double budget_seconds;
    double elapsed_seconds;              /* actual + charged */
    double actual_seconds;
    double charged_seconds;
    double last_batch_completed_seconds;
    double overrun_seconds;              /* max(0, elapsed - budget) */
    uint64_t launched_batches, completed_batches, discarded_batches;
    uint64_t completed_construction_steps, completed_ls_evaluations;
    uint64_t allocated_device_bytes;
    uint32_t preparation_completed;
    uint32_t profiled_batches;           /* rows written to `profile` */
    double setup_seconds, upload_seconds, initialization_gpu_milliseconds;
} faco_deadline_result;

/* One batch of a profiled diagnostic run (reference profiling.hpp:26-33):
 * host wall, download and verification seconds; HIP-event milliseconds of
 * the stages in the reference's order (initialization, features_and_regions,
 * gp_scoring, action_and_restart, construction_and_ls, reduction, pheromone,
 * ant_fingerprints, archive_and_feedback; fused stages report under their
 * first name). */
typedef struct faco_batch_profile {
    uint32_t batch, committed;
    double wall_seconds, download_seconds, verification_seconds, collection_seconds;
    double gpu_milliseconds[9];
} faco_batch_profile;

/* tours [tasks][n] and costs, has_incumbent, completed_seconds [tasks]
 * receive the incumbents (a colony without one keeps has_incumbent 0);
 * states [tasks] (may be NULL) the control summary of the last completed
 * batch of a program run; ratios [ratio_capacity] (may be NULL) the
 * progress feature each launched batch decided with; profile [fixed_batches]
 * (may be NULL) the per-batch rows. A fixed-size, unrestricted,
 * population-1 engine (the reference's time-budget path takes no Hard
 * mode and no population). */
int faco_engine_evaluate_deadline(faco_engine* engine, const uint64_t* instance_keys, const uint64_t* seeds,
                                  uint32_t tasks, const faco_deadline_spec* spec, faco_deadline_result* result,
                                  uint32_t* tours, double* costs, uint8_t* has_incumbent, double* completed_seconds,
                                  faco_control_state* states, double* ratios, uint32_t ratio_capacity,
                                  faco_batch_profile* profile);

/* Per-task Hard-mode 2-opt rejections of the last count evaluation
 * (evaluate_faco* or a controlled count run; a continuation counts its own
 * batches), in task order: the per-colony split of
 * faco_result.completed_constraint_rejections (reference
 * cuda/batch_engine.cu:432, colony_work[index][2]). `count` must equal the
 * last evaluation's task count. */
int faco_engine_constraint_rejections(faco_engine* engine, uint64_t* out, uint32_t count);

/* Per-task escape counters of an Escape-mode engine's last count
 * evaluation (plain or controlled), [count][FACO_ESCAPE_COUNTERS] in task
 * order; their sums are the reference's result["escape_counters"]
 * (bindings.cpp:372-388). `count` must equal the last evaluation's task
 * count. reserved_bytes (may be null): the device memory of the Escape
 * buffers (result["reserved_escape_device_bytes"]). */
int faco_engine_escape_counters(faco_engine* engine, uint64_t* out, uint32_t count, uint64_t* reserved_bytes);

/* set_preparation_charges(instance_key, cheap_seconds, preparation_seconds)
 * (reference batch_engine.cu:318-330, bindings.cpp:651-661): frozen
 * charges a cached_charged time-budget evaluation bills for the instance
 * instead of its measured registration times; finite, non-negative and,
 * once set, unchangeable. */
int faco_engine_set_preparation_charges(faco_engine* engine, uint64_t instance_key, double cheap_seconds,
                                        double preparation_seconds);

/* ---------------------------------------------------------------------
 * CountedState snapshots of controlled runs (reference
 * cpp/include/gp_faco/counted_state.hpp, cpp/src/counted_state.cpp,
 * batch_engine.hpp:119-131, bindings.cpp:541-556 and 880-934). The
 * container format is the reference's: little-endian 64-bit fields, magic
 * "GFACOST2" (an older "GFACOST1" file has 8 trailing bytes that are
 * skipped), settings, completed batches, the source progress budget,
 * (instance key, seed) tasks, host incumbents and named buffers
 * (name, element bytes, planes, bytes). The buffers hold this library's
 * device ABI, so a snapshot resumes only in an engine of this library with
 * the same dimension, colony count, settings and registered coordinates. */
typedef struct faco_state faco_state;

/* ForkIntervention (reference counted_state.hpp:40-44): a resumed baseline
 * run whose first decision is forced to action region * 4 + (mne == 2 ? 0
 * : 3) with the snapshot's start regions, under keys re-derived from `seed`.
 * A disabled fork must leave seed 0, region 0 and mne 2. */
typedef struct faco_fork {
    uint64_t seed;
    uint32_t enabled;
    uint32_t region;   /* 0..3 */
    uint32_t mne;      /* 2 or 16 */
    uint32_t pad;
} faco_fork;

/* CountedState.from_bytes: parse and validate (truncated, extended or
 * malformed containers fail with FACO_INVALID_ARGUMENT). */
int faco_state_from_bytes(const uint8_t* bytes, size_t size, faco_state** out);
/* CountedState.to_bytes: with out NULL only *size is set. */
int faco_state_to_bytes(const faco_state* state, uint8_t* out, size_t capacity, size_t* size);
/* CountedState.describe as a JSON object (NUL-terminated; *size counts the
 * NUL): state_spec_id, dimension, colonies, completed batches and
 * evaluations, progress budget, mask, settings, random keys, tasks, per
 * colony members (incumbent cost, feedback, archive size, the four keep
 * regions of the next decision) and the buffer list. */
int faco_state_describe(const faco_state* state, char* out, size_t capacity, size_t* size);
/* CountedState.select_colony: a one-colony snapshot, every buffer cut to the
 * colony's slice of each plane. */
int faco_state_select_colony(const faco_state* state, uint32_t colony, faco_state** out);
void faco_state_destroy(faco_state* state);

/* capture_program_state / capture_baseline_state (reference
 * batch_engine.cu:465-477): a PROGRAM or BASELINE run (one decider) of the
 * `tasks` colonies that stops after capture_after evaluations per colony (a
 * whole number of batches, at most source_limit) with the progress feature
 * measured against source_limit, and returns its state. The outputs are
 * those of faco_engine_evaluate_control for the run up to the capture. */
int faco_engine_capture_control(faco_engine* engine, const uint64_t* instance_keys, const uint64_t* seeds,
                                uint32_t tasks, uint64_t source_limit, uint64_t capture_after,
                                const faco_control_spec* spec, uint32_t flags, const faco_control_trace* trace,
                                faco_result* result, uint32_t* tours, int64_t* costs, uint64_t* work,
                                faco_control_state* states, faco_state** out);
/* continue_program_state / continue_baseline_state (reference
 * batch_engine.cu:479-490, 829-856): restores the snapshot (after checking
 * its buffers, coordinates and costs) and runs `additional` more
 * evaluations per colony, at most what is left of the source budget; the
 * batch numbers, keys and progress feature continue the source run, so a
 * capture plus its continuation equals the uninterrupted run bit for bit.
 * The snapshot's experiment mask overrides spec->experiment_mask. `fork`
 * (NULL = none) applies to baseline continuations. Work counters report the
 * continuation only. */
int faco_engine_continue_control(faco_engine* engine, const faco_state* state, uint64_t additional,
                                 const faco_control_spec* spec, const faco_fork* fork, uint32_t flags,
                                 const faco_control_trace* trace, faco_result* result, uint32_t* tours,
                                 int64_t* costs, uint64_t* work, faco_control_state* states);

/* score_controller_cuda (reference cpp/src/bindings.cpp:1010-1030,
 * cuda/gp_score.cuh:37-62) and its host twin: features FP32 [12][colonies][32],
 * masks [colonies] -> scores [colonies][32] (kind 1: restart, region and
 * level tree values in slots 0-1, 2-5, 6-9) and actions [colonies]. */
int faco_score_controller_cpu(const faco_controller* controller, const float* features, const uint32_t* masks,
                              uint32_t colonies, float* scores, int32_t* actions);
int faco_score_controller_gpu(const faco_controller* controller, const float* features, const uint32_t* masks,
                              uint32_t colonies, float* scores, int32_t* actions);

/* ---------------------------------------------------------------------
 * Learned edge prior. No counterpart in the reference: an MLP 8 -> 64 ->
 * 64 -> 1 (tanh GELU) scores each primary candidate edge (a, b) from eight
 * features (distance over the local scales of a and b, over the nearest
 * distance of a and over the mean scale; candidate rank; mutual-candidate
 * flag and rank; shared candidates). Its sigmoid multiplies the edge's
 * heuristic (eta' = eta * sigmoid(logit)), so pheromone, roulette and
 * local search are unchanged. All products run on bf16 matrix cores with
 * FP32 accumulation; training minimises binary cross-entropy (positives
 * weighted by pos_weight) with Adam on FP32 master weights. */
#define FACO_PRIOR_FEATURES 8
uint32_t faco_prior_param_count(void);  /* 4801: W1[64][8] b1[64] W2[64][64] b2[64] w3[64] b3 */
int faco_prior_create(uint64_t seed, faco_prior** out);  /* Glorot-uniform weights, zero biases */
void faco_prior_destroy(faco_prior* prior);
int faco_prior_get_params(const faco_prior* prior, float* params);
int faco_prior_set_params(faco_prior* prior, const float* params);  /* also clears the Adam state */
/* logits [edges] of features [edges][8] */
int faco_prior_logits(faco_prior* prior, const float* features, uint32_t edges, float* logits);
/* full-batch gradient [param_count] and loss of the weighted BCE (mean over edges) */
int faco_prior_gradients(faco_prior* prior, const float* features, const uint8_t* labels, uint32_t edges,
                         float pos_weight, float* grads, float* loss);
/* faco_prior_gradients without any allocation, for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc). No counterpart in the reference (which
 * has no learned prior). Same kernels and results, bit for bit, as
 * faco_prior_gradients; its gradient partials and staged inputs live in a
 * fixed device array of the library (about 5.2 MB per GPU, held until the
 * process ends), so it takes at most FACO_PRIOR_NOALLOC_EDGES edges
 * (FACO_INVALID_ARGUMENT beyond) and calls no hipMalloc or hipFree. Calls
 * that overlap on one GPU share that array and race: the caller serialises
 * them. */
#define FACO_PRIOR_NOALLOC_EDGES 8192u
int faco_prior_gradients_noalloc(faco_prior* prior, const float* features, const uint8_t* labels, uint32_t edges,
                                 float pos_weight, float* grads, float* loss);
/* `epochs` passes of Adam (lr, betas 0.9/0.999, eps 1e-8) over minibatches of
 * `batch` edges in a shuffled order; mean batch loss per epoch */
int faco_prior_train(faco_prior* prior, const float* features, const uint8_t* labels, uint32_t edges,
                     uint32_t batch, uint32_t epochs, float lr, float pos_weight, uint64_t shuffle_seed,
                     float* epoch_losses);
/* one 16x16x32 bf16 matrix-core product C = A[16][32] B[32][16] (layout check) */
int faco_prior_mma_check(const float* A, const float* B, float* C);
/* edge features [n * primary_width][8] of a registered instance */
int faco_engine_edge_features(faco_engine* engine, uint64_t instance_key, float* features);
/* heuristic multipliers [n][primary_width] of a registered instance
 * (NULL restores plain FACO) */
int faco_engine_set_edge_prior(faco_engine* engine, uint64_t instance_key, const float* multipliers);
/* features -> logits on the device -> sigmoid(logit) ** exponent (FP64,
 * then FP32, floored at 1e-30); sets and returns the multipliers. The
 * committed network is used with exponent 2 (chosen on validation
 * instances). */
int faco_engine_apply_edge_prior(faco_engine* engine, faco_prior* prior, uint64_t instance_key,
                                 float* multipliers, double exponent);

/* FixedFacoGpu(coordinates, instance_key, settings).run_iterations(seed,
 * batches, mne_target) (reference cpp/include/gp_faco/fixed_faco_gpu.hpp:44-60,
 * cuda/fixed_faco.cu:105-177). The optional per-batch trace receives, for
 * batch b: parent after the batch [b][n], trails after the batch
 * [b][n*primary_width], (epoch, global, parent) costs [b][3], iteration best
 * [b], the per-ant records [b][ants] and the selection trace: selected nodes
 * and uniforms [b][ants][n] (index 0 = start node). Pass NULL to skip. */
int faco_fixed_create(const double* xy, uint32_t n, uint64_t instance_key, const faco_settings* settings,
                      faco_fixed** out);
void faco_fixed_destroy(faco_fixed* fixed);
int faco_fixed_set_local_search(faco_fixed* fixed, uint32_t mode);  /* as faco_engine_set_local_search */
int faco_fixed_settings(const faco_fixed* fixed, faco_settings* out);
int faco_fixed_run(faco_fixed* fixed, uint64_t seed, uint32_t batches, uint32_t mne_target,
                   uint32_t* tour, int64_t* cost, int64_t* initial_cost, uint64_t* construction_steps,
                   uint64_t* ls_evaluations, double* solve_seconds,
                   uint32_t* trace_parent, double* trace_trails, int64_t* trace_costs,
                   uint32_t* trace_best, faco_ant_info* trace_ants, uint32_t* trace_selected,
                   double* trace_uniforms, uint32_t* trace_tours);

/* Explicit-choice construction + checklist 2-opt for `ants` independent ants
 * on one instance: the parity harness of cuda_faco_diagnostic (reference
 * cuda/faco_operations.cu:60-132). parents and visits are [ants][n]
 * (visits[a][0] = start node); ls is [n][ls_width], distance ordered.
 * Outputs: construction tours, final tours, final positions [ants][n],
 * checklists [ants][5n] (the first info.checklist_size entries are valid). */
int faco_diagnostic_construct_ls(const double* xy, uint32_t n, const uint32_t* ls, uint32_t ls_width,
                                 uint32_t ants, const uint32_t* parents, const uint32_t* visits,
                                 const uint32_t* mne_targets, uint64_t evaluation_limit,
                                 uint32_t waves_per_block, uint32_t ants_per_block,
                                 uint32_t* construction_tours, uint32_t* tours,
                                 uint32_t* positions, uint32_t* checklists, faco_ant_info* info,
                                 uint32_t local_search);

/* EUC_2D of `count` coordinate pairs on the device (distance kernel check). */
int faco_diagnostic_distances(const int32_t* pairs /* [count][4] */, uint32_t count, int64_t* out);

/* Start cities drawn by the generator's rejection sampler for (key, batch,
 * ant) and each count in counts[n], on the host (on_device = 0, no GPU
 * needed) or on the device (generator check). */
int faco_diagnostic_start_nodes(uint64_t key, uint32_t batch, uint32_t ant, const uint32_t* counts, uint32_t n,
                                uint32_t on_device, uint32_t* out);

/* Streaming copy of `bytes` (multiple of 16) between two device buffers,
 * `repeats` launches of faco_copy_kernel; average launch milliseconds. The
 * known-traffic calibration kernel for the HBM byte counters. */
int faco_diagnostic_copy(uint64_t bytes, uint32_t repeats, double* milliseconds);

/* Philox words of the device generator for (key, batch, ant, step): the
 * first `words` 32-bit outputs of the stream (generator check). */
int faco_diagnostic_philox(uint64_t key, uint32_t batch, uint32_t ant, uint32_t step, uint32_t words,
                           uint32_t* out);

/* score_cpu / score_cuda (reference cpp/include/gp_faco/program.hpp:37-42,
 * cpp/src/program_cpu.cpp:79-116, cuda/gp_score.cu:52-77): features are FP32
 * [12][colonies][32], masks u32 [colonies]; outputs scores [colonies][32] and
 * actions [colonies]. */
int faco_score_cpu(const faco_program* program, const float* features, const uint32_t* masks,
                   uint32_t colonies, float* scores, int32_t* actions);
int faco_score_gpu(const faco_program* program, const float* features, const uint32_t* masks,
                   uint32_t colonies, float* scores, int32_t* actions);

/* Escape primitives of the E3 candidate-graph exit (reference
 * cuda/escape_device.cuh:8-128, cpp/include/gp_faco/escape_slots.hpp:18-71,
 * escape_edges.hpp:13-122), one wave per row or transaction; diagnostic
 * entry points that run them on the device.
 *
 * Rows: for row r (ant index r), the escape stream of (key, batch, r,
 * events[r], ls) builds the slot view of original[r] (width <= 20 slots,
 * valid prefix, padding = n) with backup[r] (bw <= 64, padding = n) for city
 * from[r]; with xy ([n][2] integer coordinates, may be NULL) the view is then
 * ordered by (EUC_2D distance to from[r], id) as an LS row. Outputs members
 * and replaced flags [rows][width] and counts [rows][2] = (valid, replacements).
 * The escape stream: Philox4x32-10 keyed by `key`, counter block
 * (block, event, domain | ant, batch), domain 0x40000000 (construction) or
 * 0x80000000 (LS rows), four words per block. */
int faco_diagnostic_escape_rows(uint32_t rows, uint32_t width, const uint32_t* original, uint32_t bw,
                                const uint32_t* backup, uint32_t n, const uint32_t* from, uint64_t key, uint32_t batch,
                                const uint32_t* events, uint32_t ls, const double* xy, uint32_t* out_members,
                                uint8_t* out_replaced, uint32_t* out_counts);
/* Transactions: case c proposes counts[c] (1..3) added edges against as many
 * removed ones (removed/added [cases][3][2]) for a cache of cache_sizes[c]
 * edges (cache_edges [cases][64][2]) at epoch cache_epochs[c], with the CSR
 * graph (offsets [dimension+1], neighbors sorted per row) and permission
 * permits[c], then commits into the cache. out_proposals [cases][11] =
 * status (0 allowed, 1 invalid edge, 2 outside graph, 3 capacity), count,
 * expected size, expected epoch (low, high), edges [3][2]; out_committed,
 * out_cache_after [cases][64][2] and out_size_after [cases]. */
int faco_diagnostic_escape_transactions(uint32_t cases, const uint32_t* counts, const uint32_t* removed,
                                        const uint32_t* added, const uint32_t* permits, const uint32_t* cache_edges,
                                        const uint32_t* cache_sizes, const uint64_t* cache_epochs,
                                        const uint32_t* offsets, const uint32_t* neighbors, uint32_t dimension,
                                        uint32_t* out_proposals, uint32_t* out_committed, uint32_t* out_cache_after,
                                        uint32_t* out_size_after);
/* `count` (1..4096) words of the escape stream of (key, batch, ant < 2^30, event, ls). */
int faco_diagnostic_escape_draws(uint64_t key, uint32_t batch, uint32_t ant, uint32_t event, uint32_t ls, uint32_t count,
                                 uint32_t* out);

#ifdef __cplusplus
}
#endif

#endif /* FACO_ENGINE_H */
