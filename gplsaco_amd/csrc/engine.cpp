// Host side of the FACO batch engine: problem preparation, device buffers,
// the iteration loop (eager or replayed through a HIP graph), the fixed-FACO
// entry point, diagnostics and the C ABI declared in include/faco_engine.h.
//
// Boundary semantics follow the reference's native module gp_faco_ext
// (cpp/src/bindings.cpp:500-1046) and engine host code
// (cuda/batch_engine.cu:78-1192, cuda/fixed_faco.cu:44-177).
#include "counted_state.h"
#include "faco_common.h"
#include "gp_program.h"
#include "host_prepare.h"
#include "prior.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

namespace faco {
namespace {

thread_local std::string g_error;
using Clock = std::chrono::steady_clock;
double seconds_since(Clock::time_point t) { return std::chrono::duration<double>(Clock::now() - t).count(); }

void require(bool condition, const char* message) {
    if (!condition) throw std::invalid_argument(message);
}
void checked(hipError_t status, const char* what) {
    if (status != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(status));
}

template<class F> int guarded(F&& f) {
    try { f(); return FACO_OK; }
    catch (const std::invalid_argument& e) { g_error = e.what(); return FACO_INVALID_ARGUMENT; }
    catch (const std::exception& e) { g_error = e.what(); return FACO_RUNTIME_ERROR; }
    catch (...) { g_error = "unknown native error"; return FACO_RUNTIME_ERROR; }
}

// Owning device allocation.
template<class T> class DeviceArray {
public:
    DeviceArray() = default;
    DeviceArray(const DeviceArray&) = delete;
    DeviceArray& operator=(const DeviceArray&) = delete;
    ~DeviceArray() { if (ptr_) (void)hipFree(ptr_); }
    void allocate(size_t count) {
        require(ptr_ == nullptr, "device buffer allocated twice");
        count_ = count;
        if (count) checked(hipMalloc(reinterpret_cast<void**>(&ptr_), bytes()), "hipMalloc");
    }
    T* data() const { return ptr_; }
    size_t size() const { return count_; }
    size_t bytes() const { return count_ * sizeof(T); }
    void upload(const T* src, size_t count, hipStream_t s, size_t offset = 0) {
        require(offset + count <= count_, "device upload out of range");
        if (count) checked(hipMemcpyAsync(ptr_ + offset, src, count * sizeof(T), hipMemcpyHostToDevice, s), "upload");
    }
    void download(T* dst, size_t count, hipStream_t s, size_t offset = 0) const {
        require(offset + count <= count_, "device download out of range");
        if (count) checked(hipMemcpyAsync(dst, ptr_ + offset, count * sizeof(T), hipMemcpyDeviceToHost, s), "download");
    }
private:
    T* ptr_ = nullptr;
    size_t count_ = 0;
};

// Device-resident geometry of one engine: per-instance arrays packed by slot.
struct DeviceGeometry {
    DeviceArray<uint32_t> coords;     // [G][n]
    DeviceArray<uint16_t> nbr;        // [G][nbr_stride]
    DeviceArray<uint16_t> nbr_dist;   // [G][nbr_stride] distance of each nbr entry
    DeviceArray<uint16_t> backup;     // [G][n][bw]
    DeviceArray<double> eta;          // [G][n][pw]
    DeviceArray<uint32_t> initial;    // [G][n]
    DeviceArray<int64_t> initial_cost;
};

struct LaunchShape {
    uint32_t waves = 1;
    uint32_t nbr_in_lds = 1;
    uint32_t nbr_lds_bytes = 0;
};

int current_device() { int d = 0; checked(hipGetDevice(&d), "hipGetDevice"); return d; }

size_t max_lds_per_block() {
    int value = 0;
    checked(hipDeviceGetAttribute(&value, hipDeviceAttributeMaxSharedMemoryPerBlock, current_device()),
            "hipDeviceGetAttribute");
    return static_cast<size_t>(value);
}

// Workgroup shape for the ant kernel: the most ants per workgroup that both
// divides the ants of a colony and fits the LDS budget; neighbour rows are
// staged in LDS when that many ants still fit beside them.
LaunchShape choose_shape(uint32_t n, uint32_t K, uint32_t ants, uint32_t forced_waves, uint32_t extra_wave_bytes = 0) {
    LaunchShape shape;
    // the kernels' static LDS (ant slot counter, the persistent kernel's
    // colony state) comes on top of the dynamic share
    constexpr size_t kStaticLds = 256;
    const size_t budget = max_lds_per_block() - kStaticLds;
    shape.nbr_lds_bytes = round16_host(2u * n * K);
    const uint32_t per_wave = ant_wave_bytes(n) + extra_wave_bytes;
    if (forced_waves) {
        // a forced shape: any wave count 1..16 (waves take ants dynamically,
        // so it need not divide the ants); bit 8 keeps neighbour rows global
        const uint32_t w = forced_waves & 0xffu;
        if (w < 1 || w > 16) throw std::invalid_argument("waves per workgroup must be in 1..16");
        for (uint32_t lds_rows : {1u, 0u}) {
            if (lds_rows && (forced_waves & 0x100u)) continue;
            const size_t shared = ant_shared_bytes(n, lds_rows ? shape.nbr_lds_bytes : 0u);
            if (shared + static_cast<size_t>(w) * per_wave <= budget) {
                shape.waves = w; shape.nbr_in_lds = lds_rows;
                return shape;
            }
        }
        throw std::invalid_argument("forced shape exceeds the LDS budget");
    }
    // most ants in flight first (the ant kernel is bound by issue and
    // latency, which more resident waves hide), rows in LDS at equal count
    for (uint32_t w : {16u, 8u, 4u, 2u, 1u}) {
        if (ants % w) continue;
        for (uint32_t lds_rows : {1u, 0u}) {
            const size_t shared = ant_shared_bytes(n, lds_rows ? shape.nbr_lds_bytes : 0u);
            if (shared + static_cast<size_t>(w) * per_wave <= budget) {
                shape.waves = w; shape.nbr_in_lds = lds_rows;
                return shape;
            }
        }
    }
    throw std::invalid_argument("instance too large for the LDS budget of one ant");
}

// The shape with the most ants in flight for instances of up to n cities,
// neighbour rows in LDS where the same count fits beside them (waves are
// taken dynamically, so any count 1..16); waves 0 when none fits.
LaunchShape most_waves_shape(uint32_t n, uint32_t K, uint32_t extra_wave_bytes) {
    LaunchShape shape;
    constexpr size_t kStaticLds = 256;
    const size_t budget = max_lds_per_block() - kStaticLds;
    shape.nbr_lds_bytes = round16_host(2u * n * K);
    const uint32_t per_wave = ant_wave_bytes(n) + extra_wave_bytes;
    for (uint32_t w = 16; w >= 1; --w)
        for (uint32_t lds_rows : {1u, 0u}) {
            const size_t shared = ant_shared_bytes(n, lds_rows ? shape.nbr_lds_bytes : 0u);
            if (shared + static_cast<size_t>(w) * per_wave <= budget) {
                shape.waves = w; shape.nbr_in_lds = lds_rows;
                return shape;
            }
        }
    shape.waves = 0;
    return shape;
}

struct PreparedSlot { uint64_t key; const Prepared* problem; };

}  // namespace

// ---------------------------------------------------------------------------
struct Engine {
    uint32_t n, panel, population, colonies_capacity;
    faco_settings config;
    uint32_t K, nbr_stride, ls_offset = 0;
    int device;
    std::mutex mutex;
    std::map<uint64_t, Prepared> registry;
    hipStream_t stream = nullptr;
    size_t allocated = 0;
    LaunchShape shape;
    uint32_t forced_waves = 0, forced_blocks_per_colony = 0, xcd_remap = 1;
    uint32_t local_search = 0;  // 0: checklist 2-opt, 1: 2-opt + Or-opt
    double roulette_margin = kRouletteMargin;  // fast roulette margin (relative to the row total)
    // Mixed-size batch: instances of any n up to `n` (which is then the
    // stride of every per-city array; config.ants is the ant stride, the
    // ants of the largest n). Each evaluation uploads every colony's shape
    // and the largest-work-first block order.
    bool mixed = false;
    bool hard = false;                     // constraint mode Hard: every tour stays in its instance's graph
    uint32_t graph_stride = 0;             // neighbour capacity per slot: 2 n (primary_width + 1)
    DeviceArray<uint32_t> graph_offsets, graph_neighbors;   // [G][n + 1], [G][graph_stride]
    uint32_t graph_words = 0;              // 32-bit words per adjacency row: ceil(n / 32)
    DeviceArray<uint32_t> graph_bits;      // [G][n][graph_words] adjacency bits of the same graph
    // Escape mode (on the Hard graph): escape gates, slot views, per-ant
    // edge caches and anchor LS views; an evaluation may disable the gates
    // (FACO_RUN_NO_ESCAPE), which must reproduce the Hard run
    bool escape = false, escape_disabled = false;
    uint32_t wave_extra = 0;               // per-wave LDS of the ant kernel beyond the tour arrays
    uint64_t escape_bytes = 0;             // device memory of the Escape buffers
    DeviceArray<uint16_t> esc_rows;        // [A][n][lw] anchors' LS views (city | replaced << 15)
    DeviceArray<uint8_t> esc_fp;           // [A][n] footprints of the ants' tours
    DeviceArray<uint8_t> parent_fp, epoch_fp, global_fp;   // [C][n]
    DeviceArray<uint8_t> archive_fp;       // [C][2][4][n] (the archive ring's layout)
    DeviceArray<uint32_t> esc_stats;       // [A][12]
    DeviceArray<uint64_t> esc_totals;      // [C][12]
    DeviceArray<double> eta_backup;        // [G][n][bw] d^-beta of every backup slot
    std::vector<uint64_t> last_escape;     // [C][11] escape counters of the last count evaluation
    faco_settings base{};                  // settings as given (ants 0 = per-n default)
    DeviceArray<ColonyShape> shapes;       // [C]
    DeviceArray<uint32_t> order;           // [C]
    bool shapes_active = false;
    uint32_t min_ants = 0;                 // smallest ants of the current batch
    // Size classes of a mixed batch on the persistent paths: the colonies of
    // at most small_n cities (listed after the others in `order`) run as a
    // second launch, beside the first on the side stream, with small_shape:
    // more ants in flight than the engine's shape, which is sized for the
    // largest instance, allows. split_small = 0: one launch.
    uint32_t split_big = 0, split_small = 0, small_n = 0;
    LaunchShape small_shape;
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    int compute_units = 1;
    double p_root = 0;
    // device state
    DeviceGeometry geo;
    DeviceArray<uint32_t> geometry, parent, links, epoch, global, tours;
    DeviceArray<uint64_t> keys;
    DeviceArray<ColonyState> state;
    DeviceArray<double> trails, products;
    DeviceArray<faco_ant_info> info;
    std::vector<uint64_t> resident;  // instance key per geometry slot
    // controlled-loop buffers, allocated on the first controlled evaluation
    struct ControlBuffers {
        DeviceArray<ControlState> control;          // [C]
        DeviceArray<uint32_t> archive_tours, archive_links;  // [C][2][4][n]
        DeviceArray<uint64_t> ant_identity;         // [A][2]
        DeviceArray<uint32_t> ant_zero;             // [A]
        DeviceArray<uint32_t> samples;              // [G][64]
        DeviceArray<double> local_scale, epsilon;   // [G][n], [G]
        DeviceArray<faco_program> programs;         // [population]
        DeviceArray<faco_controller> controllers;   // [population]
        DeviceArray<float> features, scores;        // [12][C][32], [C][32]
        DeviceArray<uint32_t> masks, alternatives, regions;  // [C], [C], [C][128]
        DeviceArray<int32_t> actions;               // [C]
        DeviceArray<uint64_t> source_keys;          // [C] keys before a state fork
        DeviceArray<ControlParams> params;          // [1] the persistent controlled kernel's parameters
    };
    std::unique_ptr<ControlBuffers> ctl;
    bool resident_control = false;
    // The last plain evaluation, which FACO_RUN_CONTINUE resumes: its colony
    // keys and MNE target (empty: nothing to resume). Anything that changes
    // what the next batch would compute from the colony state forgets it.
    std::vector<uint64_t> continuable;
    uint32_t continuable_mne = 0;
    // Hard-mode 2-opt rejections of each colony in the last count evaluation
    std::vector<uint64_t> last_rejections;
    // launches of the persistent kernels so far (results report the delta)
    uint64_t persistent_launch_count = 0;
    // controlled runs may use the persistent controlled kernel (the
    // controlled autotuner keeps whichever form was faster)
    bool fuse_control = true;
    // frozen preparation charges per instance key (cheap, preparation
    // seconds) billed by cached_charged time-budget evaluations
    std::map<uint64_t, std::pair<double, double>> charges;
    void forget_run() { continuable.clear(); continuable_mne = 0; }
    // graph cache
    hipGraphExec_t graph = nullptr;
    uint32_t graph_colonies = 0, graph_iterations = 0;
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;

    // split_rows: every neighbour row holds the primary row then the LS row
    // (K = pw + lw, for graph specs whose rows are independent); otherwise
    // both are prefixes of one (distance, id) order (K = max(pw, lw)).
    Engine(uint32_t dimension, uint32_t colonies, faco_settings settings, uint32_t pop, bool mixed_sizes = false,
           bool split_rows = false)
        : n(dimension), panel(colonies), population(pop), colonies_capacity(colonies * pop) {
        mixed = mixed_sizes;
        base = settings;
        require(colonies >= 1 && colonies <= 4096 && pop >= 1 && pop <= 128,
                "panel size must be in 1..4096 and population in 1..128");
        config = normalized_settings(settings, n);
        require(config.primary_width <= kMaxPrimary, "primary_width above 64 is not supported");
        require(config.ls_width <= kMaxLs, "ls_width above 32 is not supported");
        require(config.backup_width <= kMaxBackup, "backup_width above 256 is not supported");
        require(n <= 65535, "dimension above 65535 is not supported");
        K = split_rows ? config.primary_width + config.ls_width : std::max(config.primary_width, config.ls_width);
        ls_offset = split_rows ? config.primary_width : 0u;
        nbr_stride = round16_host(2u * n * K) / 2u;
        device = current_device();
        checked(hipDeviceGetAttribute(&compute_units, hipDeviceAttributeMultiprocessorCount, device), "CU count");
        p_root = std::pow(config.p_best, 1.0 / config.primary_width);
        shape = choose_shape(n, K, config.ants, 0, wave_extra);
        checked(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
        checked(hipEventCreate(&ev_start), "hipEventCreate");
        checked(hipEventCreate(&ev_stop), "hipEventCreate");
        const size_t C = colonies_capacity, A = C * config.ants, G = panel;
        auto alloc = [&](auto& buffer, size_t count) { buffer.allocate(count); allocated += buffer.bytes(); };
        alloc(geo.coords, G * n); alloc(geo.nbr, G * nbr_stride); alloc(geo.nbr_dist, G * nbr_stride); alloc(geo.backup, G * n * std::max(config.backup_width, 1u));
        alloc(geo.eta, G * n * config.primary_width); alloc(geo.initial, G * n); alloc(geo.initial_cost, G);
        alloc(geometry, C); alloc(keys, C); alloc(state, C);
        alloc(parent, C * n); alloc(links, C * n); alloc(epoch, C * n); alloc(global, C * n);
        alloc(trails, C * n * config.primary_width); alloc(products, C * n * config.primary_width);
        alloc(tours, A * n); alloc(info, A);
        if (mixed) { alloc(shapes, C); alloc(order, C); }
    }
    ~Engine() {
        if (graph) (void)hipGraphExecDestroy(graph);
        if (ev_start) (void)hipEventDestroy(ev_start);
        if (ev_stop) (void)hipEventDestroy(ev_stop);
        if (ev_fork) (void)hipEventDestroy(ev_fork);
        if (ev_join) (void)hipEventDestroy(ev_join);
        if (side) (void)hipStreamDestroy(side);
        if (stream) (void)hipStreamDestroy(stream);
    }

    // Escape-mode buffers (constraint mode 2, on top of the Hard graph).
    void enable_escape() {
        escape = true;
        wave_extra = kEscapeWaveBytes;
        shape = choose_shape(n, K, config.ants, 0, wave_extra);
        const size_t C = colonies_capacity, A = C * config.ants, G = panel;
        auto alloc = [&](auto& buffer, size_t count) { buffer.allocate(count); allocated += buffer.bytes(); };
        alloc(esc_rows, A * n * config.ls_width); alloc(esc_fp, A * n);
        alloc(parent_fp, C * n); alloc(epoch_fp, C * n); alloc(global_fp, C * n);
        alloc(archive_fp, C * 2 * kArchive * n);
        alloc(esc_stats, A * kEscStats); alloc(esc_totals, C * kEscStats);
        alloc(eta_backup, G * n * std::max(config.backup_width, 1u));
        escape_bytes = esc_rows.bytes() + esc_fp.bytes() + parent_fp.bytes() + epoch_fp.bytes() + global_fp.bytes() +
                       archive_fp.bytes() + esc_stats.bytes() + esc_totals.bytes() + eta_backup.bytes();
    }
    // the escape counters of the last evaluation's `tasks` colonies
    void read_escape(uint32_t tasks) {
        last_escape.clear();
        if (!escape) return;
        std::vector<uint64_t> totals(static_cast<size_t>(tasks) * kEscStats);
        esc_totals.download(totals.data(), totals.size(), stream);
        checked(hipStreamSynchronize(stream), "escape counters");
        for (uint32_t c = 0; c < tasks; ++c)
            last_escape.insert(last_escape.end(), totals.begin() + static_cast<size_t>(c) * kEscStats,
                               totals.begin() + static_cast<size_t>(c) * kEscStats + FACO_ESCAPE_COUNTERS);
    }
    // every footprint and counter starts at zero with each evaluation
    void reset_escape() {
        if (!escape) return;
        for (auto* b : {&parent_fp, &epoch_fp, &global_fp, &archive_fp})
            checked(hipMemsetAsync(b->data(), 0, b->bytes(), stream), "escape reset");
        checked(hipMemsetAsync(esc_totals.data(), 0, esc_totals.bytes(), stream), "escape reset");
    }

    void ensure_control() {
        if (ctl) return;
        auto b = std::make_unique<ControlBuffers>();
        const size_t C = colonies_capacity, A = C * config.ants, G = panel;
        auto alloc = [&](auto& buffer, size_t count) { buffer.allocate(count); allocated += buffer.bytes(); };
        alloc(b->control, C);
        alloc(b->archive_tours, C * 2 * kArchive * n); alloc(b->archive_links, C * 2 * kArchive * n);
        alloc(b->ant_identity, 2 * A); alloc(b->ant_zero, A);
        alloc(b->samples, G * kSamples); alloc(b->local_scale, G * n); alloc(b->epsilon, G);
        alloc(b->programs, population); alloc(b->controllers, population);
        alloc(b->features, 12 * 32 * C); alloc(b->scores, 32 * C);
        alloc(b->masks, C); alloc(b->alternatives, C); alloc(b->regions, 128 * C); alloc(b->actions, C);
        alloc(b->source_keys, C);
        alloc(b->params, 1);
        ctl = std::move(b);
        resident.clear();  // the control geometry is uploaded with the next geometry upload
    }

    // Ants of an instance of `m` cities (the settings' per-size default
    // when they give ants = 0).
    uint32_t instance_ants(uint32_t m) const { return mixed ? normalized_settings(base, m).ants : config.ants; }

    // Mixed batch: every colony's shape for an evaluation limit per colony
    // (a whole number of its own ant batches), uploaded with the
    // largest-work-first block order. Returns the most iterations any colony
    // runs.
    uint64_t prepare_shapes(const uint64_t* instance_keys, uint32_t tasks, uint64_t limit) {
        std::vector<ColonyShape> host(tasks);
        uint64_t most = 0;
        uint32_t smallest = UINT32_MAX;
        for (uint32_t c = 0; c < tasks; ++c) {
            const uint32_t m = problem(instance_keys[c]).n, a = instance_ants(m);
            require(limit % a == 0 && limit / a <= UINT32_MAX,
                    "evaluation limit must be a whole number of ant batches for every colony");
            host[c] = ColonyShape{m, a, static_cast<uint32_t>(limit / a), 0u};
            most = std::max<uint64_t>(most, limit / a);
            smallest = std::min(smallest, a);
        }
        std::vector<uint32_t> ord(tasks);
        for (uint32_t c = 0; c < tasks; ++c) ord[c] = c;
        std::stable_sort(ord.begin(), ord.end(), [&](uint32_t x, uint32_t y) {
            return static_cast<uint64_t>(host[x].n) * host[x].ants * host[x].batches >
                   static_cast<uint64_t>(host[y].n) * host[y].ants * host[y].batches;
        });
        // Size classes: the small class is bounded by the largest colony
        // size whose workgroup holds the most ants a workgroup can (16), or
        // failing that more than the engine's shape; both classes keep the
        // largest-work-first order.
        split_big = tasks; split_small = 0; small_n = 0;
        if (!hard) {
            std::vector<uint32_t> sizes;
            for (uint32_t c = 0; c < tasks; ++c) sizes.push_back(host[c].n);
            std::sort(sizes.begin(), sizes.end());
            sizes.erase(std::unique(sizes.begin(), sizes.end()), sizes.end());
            for (const uint32_t need : {16u, shape.waves + 1u}) {
                if (need <= shape.waves || small_n) continue;
                for (auto it = sizes.rbegin(); it != sizes.rend(); ++it) {
                    const LaunchShape s = most_waves_shape(*it, K, wave_extra);
                    if (s.waves >= need) { small_n = *it; small_shape = s; break; }
                }
            }
            if (small_n) {
                std::stable_partition(ord.begin(), ord.end(), [&](uint32_t c) { return host[c].n > small_n; });
                for (uint32_t c = 0; c < tasks; ++c) split_small += host[c].n <= small_n ? 1u : 0u;
                split_big = tasks - split_small;
            }
        }
        shapes.upload(host.data(), tasks, stream);
        order.upload(ord.data(), tasks, stream);
        if (smallest != min_ants) drop_graph();   // the workgroup split depends on it
        min_ants = smallest;
        shapes_active = true;
        return most;
    }

    void set_waves(uint32_t waves, uint32_t blocks_per_colony) {
        shape = choose_shape(n, K, config.ants, waves, wave_extra);
        forced_waves = waves;
        forced_blocks_per_colony = blocks_per_colony;
        drop_graph();
    }
    void drop_graph() {
        if (graph) { (void)hipGraphExecDestroy(graph); graph = nullptr; }
        graph_colonies = graph_iterations = 0;
    }
    // The small class's launch parameters from a launch's `ap` (the large
    // class keeps `ap`, with the whole batch's colony count).
    AntParams small_class(const AntParams& ap, uint32_t colonies) const {
        AntParams s = ap;
        s.waves = small_shape.waves;
        s.nbr_in_lds = small_shape.nbr_in_lds;
        s.nbr_lds_bytes = small_shape.nbr_lds_bytes;
        s.per_wave_bytes = ant_wave_bytes(small_n) + wave_extra;
        s.lds_n = small_n;
        s.order = ap.order + split_big;
        s.total_colonies = colonies;
        return s;
    }
    // Persistent launches of a mixed batch of `colonies` colonies by size
    // class: `launch(params, count, stream)` runs one class; the small class
    // runs on the side stream, forked from and joined to the engine stream.
    template<class Launch> void by_class(const AntParams& ap, uint32_t colonies, bool small_fits, Launch&& launch) {
        if (!shapes_active || !split_small || !small_fits) { launch(ap, colonies, stream); return; }
        const AntParams small = small_class(ap, colonies);
        if (!split_big) { launch(small, colonies, stream); return; }
        if (!side) {
            checked(hipStreamCreateWithFlags(&side, hipStreamNonBlocking), "hipStreamCreate");
            checked(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming), "hipEventCreate");
            checked(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming), "hipEventCreate");
        }
        AntParams big = ap;
        big.total_colonies = colonies;
        checked(hipEventRecord(ev_fork, stream), "event");
        checked(hipStreamWaitEvent(side, ev_fork, 0), "stream wait");
        launch(big, split_big, stream);
        launch(small, split_small, side);
        checked(hipEventRecord(ev_join, side), "event");
        checked(hipStreamWaitEvent(stream, ev_join, 0), "stream wait");
    }

    const Prepared& problem(uint64_t key) const {
        const auto it = registry.find(key);
        require(it != registry.end(), "task instance has not been registered");
        return it->second;
    }

    // Upload geometry for the distinct instances of `tasks` (first-appearance
    // order) and the colony -> slot map.
    void upload_geometry(const uint64_t* instance_keys, uint32_t count) {
        std::vector<uint64_t> slots;
        std::vector<uint32_t> map(count);
        for (uint32_t c = 0; c < count; ++c) {
            const auto found = std::find(slots.begin(), slots.end(), instance_keys[c]);
            map[c] = static_cast<uint32_t>(found - slots.begin());
            if (found == slots.end()) slots.push_back(instance_keys[c]);
        }
        geometry.upload(map.data(), count, stream);
        require(slots.size() <= panel, "more distinct instances than the engine's panel");
        if (slots == resident && (!ctl || resident_control)) return;
        const uint32_t bw = config.backup_width, pw = config.primary_width;
        for (uint32_t g = 0; g < slots.size(); ++g) {
            const Prepared& p = problem(slots[g]);
            // counts are the instance's own size (smaller than the stride n
            // only in a mixed-size batch), offsets use the stride
            const uint32_t m = p.n;
            geo.coords.upload(p.packed.data(), m, stream, static_cast<size_t>(g) * n);
            geo.nbr.upload(p.nbr.data(), nbr_stride, stream, static_cast<size_t>(g) * nbr_stride);
            geo.nbr_dist.upload(p.nbr_dist.data(), nbr_stride, stream, static_cast<size_t>(g) * nbr_stride);
            if (bw) geo.backup.upload(p.backup16.data(), static_cast<size_t>(m) * bw, stream, static_cast<size_t>(g) * n * bw);
            if (escape && bw)
                eta_backup.upload(p.eta_backup.data(), static_cast<size_t>(m) * bw, stream, static_cast<size_t>(g) * n * bw);
            if (p.prior.empty()) {
                geo.eta.upload(p.eta.data(), static_cast<size_t>(m) * pw, stream, static_cast<size_t>(g) * n * pw);
            } else {
                // learned edge prior: eta' = eta * multiplier, in FP64
                std::vector<double> scaled(p.eta.size());
                for (size_t i = 0; i < scaled.size(); ++i) scaled[i] = p.eta[i] * static_cast<double>(p.prior[i]);
                geo.eta.upload(scaled.data(), scaled.size(), stream, static_cast<size_t>(g) * n * pw);
                checked(hipStreamSynchronize(stream), "prior upload");
            }
            geo.initial.upload(p.initial.data(), m, stream, static_cast<size_t>(g) * n);
            geo.initial_cost.upload(&p.initial_cost, 1, stream, g);
            if (hard) {
                require(!p.graph_offsets.empty(), "a Hard-mode instance needs its graph (register_graph_problem)");
                graph_offsets.upload(p.graph_offsets.data(), m + 1, stream, static_cast<size_t>(g) * (n + 1));
                graph_neighbors.upload(p.graph_neighbors.data(), p.graph_neighbors.size(), stream,
                                       static_cast<size_t>(g) * graph_stride);
                std::vector<uint32_t> bits(static_cast<size_t>(n) * graph_words, 0u);
                for (uint32_t a = 0; a < m; ++a)
                    for (uint32_t k = p.graph_offsets[a]; k < p.graph_offsets[a + 1]; ++k) {
                        const uint32_t b = p.graph_neighbors[k];
                        bits[static_cast<size_t>(a) * graph_words + (b >> 5)] |= 1u << (b & 31u);
                    }
                graph_bits.upload(bits.data(), bits.size(), stream, static_cast<size_t>(g) * n * graph_words);
                checked(hipStreamSynchronize(stream), "graph upload");
            }
            if (ctl) {
                const auto samples = sample_nodes(m, slots[g]);
                ctl->samples.upload(samples.data(), kSamples, stream, static_cast<size_t>(g) * kSamples);
                ctl->local_scale.upload(p.local_scale.data(), m, stream, static_cast<size_t>(g) * n);
                ctl->epsilon.upload(&p.scale_epsilon, 1, stream, g);
            }
        }
        checked(hipStreamSynchronize(stream), "geometry upload");
        resident = slots;
        resident_control = ctl != nullptr;
    }

    AntParams ant_params(uint32_t colonies) const {
        AntParams p{};
        p.nbr_in_lds = shape.nbr_in_lds;
        p.n = n; p.ants = config.ants; p.pw = config.primary_width; p.bw = config.backup_width;
        p.lw = config.ls_width; p.K = K; p.ls_offset = ls_offset; p.waves = shape.waves;
        p.nbr_lds_bytes = shape.nbr_lds_bytes; p.nbr_stride = nbr_stride;
        p.per_wave_bytes = ant_wave_bytes(n) + wave_extra;
        p.xcd_remap = shapes_active ? 0u : xcd_remap;   // mixed: keep the largest-first order
        // Split each colony over just enough workgroups to fill every CU:
        // fewer, larger chunks let the waves of a workgroup balance uneven
        // ants among themselves.
        uint32_t bpc = 1;
        const uint64_t slots = static_cast<uint64_t>(compute_units) * ant_search_blocks_per_cu(p);
        // (a mixed batch splits every colony the same way: bpc divides its smallest ants)
        const uint32_t ants_split = shapes_active ? min_ants : config.ants;
        while (static_cast<uint64_t>(colonies) * bpc < slots && ants_split % (2 * bpc * shape.waves) == 0) bpc *= 2;
        if (forced_blocks_per_colony && ants_split % forced_blocks_per_colony == 0) bpc = forced_blocks_per_colony;
        p.blocks_per_colony = bpc;
        p.chunk = config.ants / bpc;
        p.ls_limit = std::min<uint64_t>(config.ls_evaluation_limit, UINT32_MAX);   // see AntParams::ls_limit
        p.roulette_margin = roulette_margin;
        p.or_opt = local_search;
        p.filter_tours = 1;   // plain FACO: only the iteration best tour is read back
        p.coords = geo.coords.data(); p.nbr = geo.nbr.data(); p.nbr_dist = geo.nbr_dist.data(); p.backup = geo.backup.data();
        p.products = products.data(); p.parent = parent.data(); p.links = links.data();
        p.state = state.data(); p.keys = keys.data(); p.targets = nullptr;
        p.geometry = geometry.data(); p.out_tours = tours.data(); p.out_info = info.data();
        if (shapes_active) { p.shapes = shapes.data(); p.order = order.data(); }
        if (hard) {
            p.graph_offsets = graph_offsets.data(); p.graph_neighbors = graph_neighbors.data();
            p.graph_stride = graph_stride;
            p.graph_bits = graph_bits.data(); p.graph_words = graph_words;
        }
        if (escape) {
            p.escape_disabled = escape_disabled ? 1u : 0u;
            p.esc_rows = esc_rows.data(); p.esc_footprint = esc_fp.data(); p.esc_stats = esc_stats.data();
            p.parent_fp = parent_fp.data(); p.eta_backup = eta_backup.data();
        }
        (void)colonies;
        return p;
    }
    ColonyParams colony_params() const {
        ColonyParams p{};
        p.n = n; p.ants = config.ants; p.pw = config.primary_width; p.K = K; p.nbr_stride = nbr_stride;
        p.retention = config.retention; p.p_best_root = p_root;
        p.epoch_probability = config.epoch_source_probability;
        p.nbr = geo.nbr.data(); p.eta = geo.eta.data(); p.geometry = geometry.data(); p.keys = keys.data();
        p.tours = tours.data(); p.info = info.data(); p.state = state.data(); p.parent = parent.data();
        p.links = links.data(); p.epoch = epoch.data(); p.global = global.data(); p.trails = trails.data();
        p.products = products.data();
        p.shapes = shapes_active ? shapes.data() : nullptr;
        if (escape) {
            p.ant_fp = esc_fp.data(); p.esc_stats = esc_stats.data();
            p.parent_fp = parent_fp.data(); p.epoch_fp = epoch_fp.data(); p.global_fp = global_fp.data();
            p.esc_totals = esc_totals.data();
        }
        return p;
    }
    void reset(uint32_t colonies, uint32_t mne) {
        ResetParams r{};
        r.n = n; r.ants = config.ants; r.pw = config.primary_width; r.K = K; r.target = mne;
        r.retention = config.retention; r.p_best_root = p_root;
        r.geometry = geometry.data(); r.initial = geo.initial.data(); r.initial_cost = geo.initial_cost.data();
        r.eta = geo.eta.data(); r.state = state.data(); r.parent = parent.data(); r.links = links.data();
        r.epoch = epoch.data(); r.global = global.data(); r.trails = trails.data(); r.products = products.data();
        r.shapes = shapes_active ? shapes.data() : nullptr;
        checked(launch_colony_reset(r, colonies, stream), "colony reset launch");
        reset_escape();
    }
    void iteration(uint32_t colonies, const AntParams& ap, const ColonyParams& cp) {
        const uint32_t blocks = colonies * ap.blocks_per_colony;
        checked(launch_ant_search(ap, blocks, stream), "ant search launch");
        checked(launch_colony_update(cp, colonies, stream), "colony update launch");
    }
    // Eager iterations with HIP events around each kernel on the engine
    // stream; ms[0] += construction+LS, ms[1] += colony update. With
    // `clocks`, the ant kernel's clock-instrumented variant also fills the
    // per-ant phase figures ms[2..9]; without, the production kernel runs.
    void run_profiled(uint32_t colonies, uint64_t iterations, double* ms, bool clocks) {
        AntParams ap = ant_params(colonies);
        const ColonyParams cp = colony_params();
        const size_t ants_total = static_cast<size_t>(colonies) * config.ants;
        DeviceArray<uint64_t> ticks;
        if (clocks) ticks.allocate(kPhaseSlots * ants_total);
        ap.phase_ticks = clocks ? ticks.data() : nullptr;
        std::vector<uint64_t> host(clocks ? kPhaseSlots * ants_total : 0);
        int clock_khz = 0;
        checked(hipDeviceGetAttribute(&clock_khz, hipDeviceAttributeWallClockRate, device), "wall clock rate");
        const double ms_per_tick = clock_khz > 0 ? 1.0 / clock_khz : 1e-5;
        hipEvent_t ev[3];
        for (auto& e : ev) checked(hipEventCreate(&e), "hipEventCreate");
        struct Destroy { hipEvent_t* e; ~Destroy() { for (int i = 0; i < 3; ++i) (void)hipEventDestroy(e[i]); } } d{ev};
        const uint32_t blocks = colonies * ap.blocks_per_colony;
        if (!clocks && persistent(ap)) {
            // the production (persistent) kernel: kernel_milliseconds[0] is the
            // whole launch, ant search and colony update together
            for (uint64_t done = 0; done < iterations;) {
                const uint64_t rounds = std::min(iterations - done, kRoundsPerLaunch);
                checked(hipEventRecord(ev[0], stream), "event");
                by_class(ap, colonies, true, [&](const AntParams& p, uint32_t count, hipStream_t s) {
                    checked(launch_colony_run(p, cp, count, static_cast<uint32_t>(rounds), s), "colony run launch");
                });
                ++persistent_launch_count;
                checked(hipEventRecord(ev[1], stream), "event");
                checked(hipEventSynchronize(ev[1]), "event sync");
                float a = 0;
                checked(hipEventElapsedTime(&a, ev[0], ev[1]), "event time");
                ms[0] += a;
                done += rounds;
            }
            return;
        }
        // waves that can hold an ant at once: resident workgroups x waves
        const double resident_waves = static_cast<double>(std::min<uint64_t>(
            blocks, static_cast<uint64_t>(compute_units) * ant_search_blocks_per_cu(ap))) * ap.waves;
        for (uint64_t i = 0; i < iterations; ++i) {
            checked(hipEventRecord(ev[0], stream), "event");
            checked(launch_ant_search(ap, blocks, stream), "ant search launch");
            checked(hipEventRecord(ev[1], stream), "event");
            checked(launch_colony_update(cp, colonies, stream), "colony update launch");
            checked(hipEventRecord(ev[2], stream), "event");
            checked(hipEventSynchronize(ev[2]), "event sync");
            float a = 0, b = 0;
            checked(hipEventElapsedTime(&a, ev[0], ev[1]), "event time");
            checked(hipEventElapsedTime(&b, ev[1], ev[2]), "event time");
            ms[0] += a; ms[1] += b;
            if (!clocks) continue;
            ticks.download(host.data(), host.size(), stream);
            checked(hipStreamSynchronize(stream), "phase ticks");
            double phase[8] = {0, 0, 0, 0, 0, 0, 0, 0}, held = 0;
            uint64_t first = ~0ull, last = 0;
            for (size_t k = 0; k < ants_total; ++k) {
                const uint64_t* t = host.data() + kPhaseSlots * k;
                for (int j = 0; j < 8; ++j) phase[j] += static_cast<double>(t[j]);
                held += static_cast<double>(t[9] - t[8]);
                first = std::min(first, t[8]);
                last = std::max(last, t[9]);
            }
            // mean per-ant milliseconds of setup, construction and local search
            for (int j = 0; j < 3; ++j) ms[2 + j] += phase[j] * ms_per_tick / ants_total;
            // mean per-ant shader clocks (millions) of the five sub-phases
            for (int j = 3; j < 8; ++j) ms[2 + j] += phase[j] / ants_total * 1e-6;
            // mean per-ant milliseconds from taking the ant to its write-back
            ms[10] += held * ms_per_tick / ants_total;
            // share of the resident waves' time spent holding an ant
            ms[11] += last > first ? held / (resident_waves * static_cast<double>(last - first)) : 0.0;
        }
    }
    // Plain FACO with one workgroup per colony runs as the persistent
    // kernel: up to kRoundsPerLaunch batches per launch, each colony's
    // workgroup doing its own colony update between batches.
    static constexpr uint64_t kRoundsPerLaunch = 1024;
    // (Hard and Escape mode too, with plain 2-opt)
    static bool persistent(const AntParams& ap) {
        return ap.blocks_per_colony == 1 && !(ap.graph_offsets && ap.or_opt);
    }
    // Run `iterations` batches: persistent launches when the shape allows,
    // else per batch an ant launch and a colony-update launch, through a
    // cached HIP graph of 16 batches when use_graph is set.
    void run(uint32_t colonies, uint64_t iterations, bool use_graph) {
        const AntParams ap = ant_params(colonies);
        const ColonyParams cp = colony_params();
        if (persistent(ap)) {
            for (uint64_t done = 0; done < iterations;) {
                const uint64_t rounds = std::min(iterations - done, kRoundsPerLaunch);
                by_class(ap, colonies, true, [&](const AntParams& p, uint32_t count, hipStream_t s) {
                    checked(launch_colony_run(p, cp, count, static_cast<uint32_t>(rounds), s), "colony run launch");
                });
                ++persistent_launch_count;
                done += rounds;
            }
            return;
        }
        constexpr uint32_t kGraphIterations = 16;
        if (use_graph && iterations >= kGraphIterations) {
            if (!graph || graph_colonies != colonies || graph_iterations != kGraphIterations) {
                drop_graph();
                hipGraph_t g = nullptr;
                checked(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "begin capture");
                for (uint32_t i = 0; i < kGraphIterations; ++i) iteration(colonies, ap, cp);
                checked(hipStreamEndCapture(stream, &g), "end capture");
                checked(hipGraphInstantiate(&graph, g, nullptr, nullptr, 0), "graph instantiate");
                (void)hipGraphDestroy(g);
                graph_colonies = colonies; graph_iterations = kGraphIterations;
            }
            const uint64_t rounds = iterations / kGraphIterations;
            for (uint64_t r = 0; r < rounds; ++r) checked(hipGraphLaunch(graph, stream), "graph launch");
            iterations -= rounds * kGraphIterations;
        }
        for (uint64_t i = 0; i < iterations; ++i) iteration(colonies, ap, cp);
    }
};

struct Fixed {
    std::unique_ptr<Engine> engine;
    uint64_t key;
};

void set_last_error(const char* message) { g_error = message; }

}  // namespace faco

using namespace faco;

struct faco_engine { std::unique_ptr<Engine> impl; };
struct faco_fixed { std::unique_ptr<Engine> impl; uint64_t key; };
struct faco_state { faco::CountedState s; };

namespace {

std::vector<int32_t> integer_coordinates(const double* xy, uint32_t n) {
    require(xy != nullptr && n >= 3 && n <= 10000, "coordinates must be FP64 [n,2] with 3 <= n <= 10000");
    std::vector<int32_t> out(2 * static_cast<size_t>(n));
    for (size_t i = 0; i < out.size(); ++i) {
        const double v = xy[i];
        require(std::isfinite(v) && v == std::floor(v) && v >= 0 && v <= kMaxCoordinate,
                "EUC_2D engine needs integer coordinates in [0, 32767]");
        out[i] = static_cast<int32_t>(v);
    }
    return out;
}

void fill_result(faco_result* r, const Engine& e, uint32_t colonies, uint64_t limit, uint64_t batches,
                 uint64_t fe_per_colony, uint64_t steps, uint64_t evals, double actual, double search, float gpu_ms) {
    if (!r) return;
    std::memset(r, 0, sizeof(*r));
    r->evaluation_limit_per_colony = limit;
    r->completed_tour_evaluations_per_colony = fe_per_colony;
    r->total_tour_evaluations = r->completed_tour_evaluations_per_colony * colonies;
    r->launched_batches = r->completed_batches = batches;
    r->completed_construction_steps = steps;
    r->completed_ls_evaluations = evals;
    r->allocated_device_bytes = e.allocated;
    r->actual_seconds = actual;
    r->search_seconds = search;
    r->search_gpu_milliseconds = gpu_ms;
}

// Curve checkpoints (reference cuda/batch_engine.cu:635-653): the run stops
// at each requested batch count, copies every colony's global-best tour on
// the device, and reads all of them back once at the end.
struct Checkpoints {
    const faco_checkpoints* spec = nullptr;
    DeviceArray<uint32_t> tours;
    size_t nodes = 0, cursor = 0;

    Checkpoints(const faco_checkpoints* s, uint64_t batches, size_t colony_nodes) : nodes(colony_nodes) {
        if (!s || !s->count) return;
        require(s->iterations != nullptr && s->tours != nullptr, "checkpoint arrays missing");
        for (uint32_t i = 0; i < s->count; ++i)
            require(s->iterations[i] <= batches && (i == 0 || s->iterations[i] > s->iterations[i - 1]),
                    "checkpoint iterations must be ascending, distinct and inside the budget");
        spec = s;
        tours.allocate(static_cast<size_t>(s->count) * nodes);
    }
    // the next batch count at which the run must stop (or the end)
    uint64_t next(uint64_t batches) const {
        return spec && cursor < spec->count ? spec->iterations[cursor] : batches;
    }
    void record(const uint32_t* global, uint64_t done, hipStream_t stream) {
        if (!spec || cursor >= spec->count || spec->iterations[cursor] != done) return;
        checked(hipMemcpyAsync(tours.data() + cursor * nodes, global, nodes * sizeof(uint32_t),
                               hipMemcpyDeviceToDevice, stream), "checkpoint copy");
        ++cursor;
    }
    // after the run's final synchronisation; `n_of(c)` is colony c's own size
    template<class SizeOf> void read_back(hipStream_t stream, uint32_t stride, SizeOf n_of) {
        if (!spec) return;
        tours.download(spec->tours, tours.size(), stream);
        checked(hipStreamSynchronize(stream), "checkpoint readback");
        for (uint32_t k = 0; k < spec->count; ++k)
            for (size_t c = 0; c < nodes / stride; ++c)
                for (uint32_t i = n_of(static_cast<uint32_t>(c)); i < stride; ++i)
                    spec->tours[k * nodes + c * stride + i] = UINT32_MAX;
    }
};

// Shared evaluation body for the batch engine and the fixed entry point.
// `any_count`: the tuner may run any number of colonies up to the engine's
// capacity (a population engine's whole shape), not just its panel.
void evaluate(Engine& e, const uint64_t* instance_keys, const uint64_t* seeds, uint32_t tasks, uint64_t batches,
              uint32_t mne, uint32_t flags, faco_result* result, uint32_t* tours, int64_t* costs, uint64_t* work,
              uint64_t limit, bool any_count = false, const faco_checkpoints* checkpoints = nullptr) {
    e.last_rejections.clear();
    e.last_escape.clear();
    const uint64_t launches_before = e.persistent_launch_count;
    const bool use_graph = (flags & FACO_RUN_GRAPH) != 0, profile = (flags & FACO_RUN_PROFILE) != 0;
    const bool events = (flags & FACO_RUN_EVENTS) != 0, resume = (flags & FACO_RUN_CONTINUE) != 0;
    require(!(flags & FACO_RUN_NO_ESCAPE) || e.escape, "only an Escape-mode engine can disable its escape");
    require(!resume || !e.escape, "an Escape-mode run is not continued");
    double kernel_ms[FACO_STAGE_COUNT] = {0};
    const auto started = Clock::now();
    std::unique_lock<std::mutex> lock(e.mutex, std::try_to_lock);
    require(lock.owns_lock(), "one evaluation per engine at a time");
    require(current_device() == e.device, "engine must be used on the device it was created on");
    require(tasks == e.panel || (any_count && tasks >= 1 && tasks <= e.colonies_capacity),
            "task count must equal the engine's colonies");
    require(mne > 0, "mne_target must be positive");
    for (uint32_t c = 0; c < tasks; ++c) (void)e.problem(instance_keys[c]);
    e.escape_disabled = (flags & FACO_RUN_NO_ESCAPE) != 0;
    e.upload_geometry(instance_keys, tasks);
    std::vector<uint64_t> keys(tasks);
    for (uint32_t c = 0; c < tasks; ++c) keys[c] = colony_key(instance_keys[c], seeds[c]);
    if (resume)
        require(!e.mixed && !(checkpoints && checkpoints->count) && keys == e.continuable && mne == e.continuable_mne,
                "a continuation resumes the engine's last plain evaluation of the same tasks and MNE target "
                "(fixed-size engines, no checkpoints)");
    e.forget_run();
    e.keys.upload(keys.data(), tasks, e.stream);
    struct Inactive { Engine& e; ~Inactive() { e.shapes_active = false; } } inactive{e};
    if (e.mixed) batches = e.prepare_shapes(instance_keys, tasks, limit);
    Checkpoints marks(checkpoints, batches, static_cast<size_t>(tasks) * e.n);
    // a continuation keeps every colony's state and reports only its own work
    std::vector<ColonyState> before(resume ? tasks : 0);
    if (resume) e.state.download(before.data(), tasks, e.stream);
    else e.reset(tasks, mne);
    marks.record(e.global.data(), 0, e.stream);
    checked(hipStreamSynchronize(e.stream), "initialisation");
    const auto search_started = Clock::now();
    checked(hipEventRecord(e.ev_start, e.stream), "event");
    for (uint64_t done = 0; done < batches;) {
        const uint64_t stop = std::max(marks.next(batches), done + 1);
        if (profile || events) e.run_profiled(tasks, stop - done, kernel_ms, profile);
        else e.run(tasks, stop - done, use_graph);
        done = stop;
        marks.record(e.global.data(), done, e.stream);
    }
    checked(hipEventRecord(e.ev_stop, e.stream), "event");
    checked(hipStreamSynchronize(e.stream), "search");
    const double search = seconds_since(search_started);
    float gpu_ms = 0;
    checked(hipEventElapsedTime(&gpu_ms, e.ev_start, e.ev_stop), "event time");
    std::vector<ColonyState> states(tasks);
    e.state.download(states.data(), tasks, e.stream);
    if (tours) e.global.download(tours, static_cast<size_t>(tasks) * e.n, e.stream);
    checked(hipStreamSynchronize(e.stream), "readback");
    if (tours && e.mixed) {
        // a smaller colony's row ends at its own n: the rest is marked
        for (uint32_t c = 0; c < tasks; ++c)
            for (uint32_t i = e.problem(instance_keys[c]).n; i < e.n; ++i) tours[static_cast<size_t>(c) * e.n + i] = UINT32_MAX;
    }
    marks.read_back(e.stream, e.n, [&](uint32_t c) { return e.problem(instance_keys[c]).n; });
    uint64_t steps = 0, evals = 0, rejections = 0;
    e.last_rejections.assign(tasks, 0);
    for (uint32_t c = 0; c < tasks; ++c) {
        const uint64_t own_steps = states[c].work_construction - (resume ? before[c].work_construction : 0);
        const uint64_t own_evals = states[c].work_ls - (resume ? before[c].work_ls : 0);
        if (costs) costs[c] = states[c].global_cost;
        if (work) { work[2 * c] = own_steps; work[2 * c + 1] = own_evals; }
        steps += own_steps; evals += own_evals;
        e.last_rejections[c] = states[c].work_constraints - (resume ? before[c].work_constraints : 0);
        rejections += e.last_rejections[c];
    }
    e.read_escape(tasks);
    fill_result(result, e, tasks, limit, batches, e.mixed ? limit : batches * e.config.ants, steps, evals,
                seconds_since(started), search, gpu_ms);
    if (result) {
        std::copy(kernel_ms, kernel_ms + FACO_STAGE_COUNT, result->kernel_milliseconds);
        result->completed_constraint_rejections = rejections;
        result->persistent_launches = e.persistent_launch_count - launches_before;
    }
    if (!e.mixed) { e.continuable = std::move(keys); e.continuable_mne = mne; }
}

// validate_baseline (cpp/include/gp_faco/baseline_policy.hpp:36-56).
void validate_baseline(const faco_baseline_policy& p, uint32_t mask) {
    if (p.kind > 1 || p.mne_level > 3 || p.max_mne_level > 3 || p.max_mne_level < p.mne_level || p.region > 3 ||
        !std::isfinite(p.restart_probability) || p.restart_probability < 0 || p.restart_probability > 1 ||
        p.restart_mode > 2)
        throw std::invalid_argument("baseline kind, MNE level, region or probability invalid");
    if (p.kind == 0) {
        if (p.max_mne_level != p.mne_level || p.stagnation_step || p.restart_stagnation || p.restart_cooldown)
            throw std::invalid_argument("a static baseline takes no feedback rule");
        const bool ok = (p.restart_mode == 0 && !p.restart_period && !p.restart_probability) ||
                        (p.restart_mode == 1 && p.restart_period && !p.restart_probability) ||
                        (p.restart_mode == 2 && !p.restart_period);
        if (!ok) throw std::invalid_argument("static restart mode does not match its period / probability");
    } else if (p.restart_mode != 0 || p.restart_period || p.restart_probability ||
               ((p.max_mne_level == p.mne_level) != (p.stagnation_step == 0)) ||
               ((p.restart_stagnation == 0) != (p.restart_cooldown == 0))) {
        throw std::invalid_argument("rule baseline needs a stagnation step and restart cooldown in canonical form");
    }
    uint32_t required = 0;
    for (uint32_t level = p.mne_level; level <= p.max_mne_level; ++level) required |= 1u << (p.region * 4 + level);
    if ((mask & required) != required)
        throw std::invalid_argument("the experiment mask must keep every hold action the baseline may choose");
}

// Controlled evaluation: the panel replicated per decider, per batch
// decide -> ants -> colony update -> control update (reference
// cuda/batch_engine.cu:880-1060).
// A snapshot taken after `capture_after` evaluations per colony, or one
// resumed (with an optional paired fork of its first decision).
struct StateIO {
    CountedState* capture = nullptr;
    uint64_t capture_after = 0;
    const CountedState* resume = nullptr;
    faco_fork fork{0, 0, 0, 2, 0};   // seed, enabled, region, mne
};

// Snapshot of a controlled run (CountedState container). The named buffers
// are every device buffer that carries algorithm state from one batch to the
// next, then the decision views of the next batch (features [12][C][32],
// scores, masks, actions, alternatives, start regions [C][2][4][16]); all
// are colony-major except the feature planes. "xy" (first) holds the
// instances' integer coordinates [C][n][2] for the restore check.
struct StateSlot {
    const char* name;
    uint64_t element, planes;
    size_t count;       // elements over all C colonies
    void* device;
    bool state;         // restored on resume; the views are recomputed
};

std::vector<StateSlot> state_slots(Engine& e, uint32_t C) {
    auto& b = *e.ctl;
    const size_t n = e.n, cells = static_cast<size_t>(C) * n * e.config.primary_width;
    const size_t archive = static_cast<size_t>(C) * 2 * kArchive * n;
    return {{"keys", sizeof(uint64_t), 1, C, e.keys.data(), true},
            {"colony_state", sizeof(ColonyState), 1, C, e.state.data(), true},
            {"control_state", sizeof(ControlState), 1, C, b.control.data(), true},
            {"parent", 4, 1, C * n, e.parent.data(), true},
            {"links", 4, 1, C * n, e.links.data(), true},
            {"epoch", 4, 1, C * n, e.epoch.data(), true},
            {"global", 4, 1, C * n, e.global.data(), true},
            {"trails", 8, 1, cells, e.trails.data(), true},
            {"products", 8, 1, cells, e.products.data(), true},
            {"archive_tours", 4, 1, archive, b.archive_tours.data(), true},
            {"archive_links", 4, 1, archive, b.archive_links.data(), true},
            {"features", 4, 12, static_cast<size_t>(12) * 32 * C, b.features.data(), false},
            {"scores", 4, 1, static_cast<size_t>(32) * C, b.scores.data(), false},
            {"masks", 4, 1, C, b.masks.data(), false},
            {"actions", 4, 1, C, b.actions.data(), false},
            {"alternatives", 4, 1, C, b.alternatives.data(), false},
            {"regions", 4, 1, static_cast<size_t>(128) * C, b.regions.data(), false}};
}

std::vector<uint8_t> state_coordinates(const Engine& e, const std::vector<uint64_t>& inst) {
    std::vector<uint8_t> out;
    for (uint64_t key : inst) {
        const auto& xy = e.problem(key).xy;
        const auto* first = reinterpret_cast<const uint8_t*>(xy.data());
        out.insert(out.end(), first, first + xy.size() * sizeof(int32_t));
    }
    return out;
}

// Restore (reference cuda/batch_engine.cu:231-261 and 829-856): the whole
// buffer set is checked against this engine's layout and the registered
// coordinates before anything is uploaded; the global tours and the host
// incumbents must cost what the state says.
void restore_state(Engine& e, const CountedState& saved, uint32_t C, const std::vector<uint64_t>& inst) {
    const auto slots = state_slots(e, C);
    require(saved.buffers.size() == slots.size() + 1, "snapshot buffer set incomplete or extra");
    const StateBuffer& xy = saved.buffers.front();
    require(xy.name == "xy" && xy.element_bytes == sizeof(int32_t) && xy.planes == 1,
            "snapshot buffer name, layout or ABI does not match the engine");
    require(xy.bytes == state_coordinates(e, inst), "snapshot coordinates differ from the registered instances");
    for (size_t i = 0; i < slots.size(); ++i) {
        const StateBuffer& source = saved.buffers[i + 1];
        require(source.name == slots[i].name && source.element_bytes == slots[i].element &&
                source.planes == slots[i].planes && source.bytes.size() == slots[i].count * slots[i].element,
                "snapshot buffer name, layout or ABI does not match the engine");
    }
    const auto& states = saved.buffer("colony_state").bytes;
    const auto& global = saved.buffer("global").bytes;
    for (uint32_t c = 0; c < C; ++c) {
        const Prepared& p = e.problem(inst[c]);
        ColonyState st;
        std::memcpy(&st, states.data() + static_cast<size_t>(c) * sizeof(ColonyState), sizeof st);
        std::vector<uint32_t> tour(e.n);
        std::memcpy(tour.data(), global.data() + static_cast<size_t>(c) * e.n * 4, static_cast<size_t>(e.n) * 4);
        require(is_permutation(tour.data(), e.n) && tour_cost(p, tour.data()) == st.global_cost,
                "snapshot global tour does not cost what its colony state says");
        const StateIncumbent& item = saved.incumbents[c];
        require(static_cast<double>(tour_cost(p, item.tour.data())) == item.cost,
                "snapshot host incumbent cost changed");
    }
    for (size_t i = 0; i < slots.size(); ++i)
        if (slots[i].state)
            checked(hipMemcpyAsync(slots[i].device, saved.buffers[i + 1].bytes.data(), slots[i].count * slots[i].element,
                                   hipMemcpyHostToDevice, e.stream), "snapshot upload");
}

// Launch parameters of the control kernels for a panel of `panel` tasks
// under `spec`, the progress feature measured against `progress` FE.
ControlParams control_params(Engine& e, uint32_t panel, const faco_control_spec& spec, uint64_t progress) {
    auto& b = *e.ctl;
    ControlParams cp{};
    cp.n = e.n; cp.ants = e.config.ants; cp.pw = e.config.primary_width; cp.K = e.K; cp.nbr_stride = e.nbr_stride;
    cp.panel = panel; cp.mode = spec.mode; cp.variant = spec.factorial_variant;
    cp.experiment_mask = spec.experiment_mask; cp.force_collisions = spec.force_fingerprint_collisions ? 1u : 0u;
    cp.forced_action = -1;
    cp.evaluation_limit = progress; cp.ls_limit = e.config.ls_evaluation_limit;
    cp.retention = e.config.retention; cp.p_best_root = e.p_root; cp.baseline = spec.baseline;
    cp.coords = e.geo.coords.data(); cp.nbr = e.geo.nbr.data(); cp.eta = e.geo.eta.data();
    cp.local_scale = b.local_scale.data(); cp.epsilon = b.epsilon.data(); cp.samples = b.samples.data();
    cp.initial = e.geo.initial.data(); cp.initial_cost = e.geo.initial_cost.data();
    cp.geometry = e.geometry.data(); cp.keys = e.keys.data();
    cp.programs = b.programs.data(); cp.controllers = b.controllers.data();
    cp.state = e.state.data(); cp.control = b.control.data();
    cp.parent = e.parent.data(); cp.links = e.links.data(); cp.epoch = e.epoch.data();
    cp.trails = e.trails.data(); cp.products = e.products.data();
    cp.archive_tours = b.archive_tours.data(); cp.archive_links = b.archive_links.data();
    cp.tours = e.tours.data(); cp.info = e.info.data();
    cp.ant_identity = b.ant_identity.data(); cp.ant_zero = b.ant_zero.data();
    cp.features = b.features.data(); cp.scores = b.scores.data(); cp.masks = b.masks.data();
    cp.actions = b.actions.data(); cp.alternatives = b.alternatives.data(); cp.regions = b.regions.data();
    cp.shapes = e.shapes_active ? e.shapes.data() : nullptr;
    if (e.escape) {
        cp.archive_fp = e.archive_fp.data(); cp.ant_fp = e.esc_fp.data();
        cp.parent_fp = e.parent_fp.data(); cp.epoch_fp = e.epoch_fp.data();
    }
    return cp;
}

// Controlled batches run as one persistent launch when every colony is one
// workgroup of 2..16 waves, the engine is
// unrestricted and the control bodies' LDS fits beside the ants' shared data.
bool fused_control(const AntParams& ap) {
    return ap.blocks_per_colony == 1 && !ap.graph_offsets && ap.waves >= 2 && ap.waves <= 16 &&
           control_run_lds_bytes(ap) + 256u <= max_lds_per_block();
}

// The ant kernel of a controlled run: every ant's tour and identity are
// written (the archive update reads them) and ants start in the region the
// decision chose.
AntParams controlled_ant_params(Engine& e, uint32_t colonies) {
    AntParams ap = e.ant_params(colonies);
    ap.filter_tours = 0;
    ap.control = e.ctl->control.data();
    ap.out_identity = e.ctl->ant_identity.data();
    ap.out_zero = e.ctl->ant_zero.data();
    return ap;
}

void evaluate_control(Engine& e, const uint64_t* panel_keys, const uint64_t* panel_seeds, uint32_t panel,
                      uint64_t limit, const faco_control_spec& spec, uint32_t flags, const faco_control_trace* trace,
                      faco_result* result, uint32_t* tours, int64_t* costs, uint64_t* work,
                      faco_control_state* states, const faco_checkpoints* checkpoints, const StateIO* io = nullptr) {
    e.last_rejections.clear();
    e.last_escape.clear();
    const uint64_t launches_before = e.persistent_launch_count;
    const auto started = Clock::now();
    std::unique_lock<std::mutex> lock(e.mutex, std::try_to_lock);
    require(lock.owns_lock(), "one evaluation per engine at a time");
    require(current_device() == e.device, "engine must be used on the device it was created on");
    e.forget_run();
    require(spec.mode >= FACO_CONTROL_PROGRAM && spec.mode <= FACO_CONTROL_FACTORIAL, "unknown control mode");
    require(!e.hard || !e.local_search, "a Hard-mode engine uses the plain 2-opt");
    require((spec.experiment_mask & 0xffffu) != 0, "the experiment mask must keep at least one hold action");
    require(!(flags & FACO_RUN_NO_ESCAPE) || e.escape, "only an Escape-mode engine can disable its escape");
    e.escape_disabled = (flags & FACO_RUN_NO_ESCAPE) != 0;
    const uint32_t ants = e.config.ants;   // the ant stride in a mixed-size batch
    if (!e.mixed)
        require(limit % ants == 0 && limit / ants <= UINT32_MAX, "evaluation limit must be a whole number of ant batches");
    const bool scored = spec.mode != FACO_CONTROL_BASELINE;
    const uint32_t deciders = scored ? spec.deciders : 1u;
    require(deciders >= 1 && deciders <= e.population, "decider count must be in 1..population_size");
    require(panel >= 1 && panel <= e.panel, "panel size out of range");
    if (spec.mode == FACO_CONTROL_PROGRAM || spec.mode == FACO_CONTROL_FACTORIAL) {
        require(spec.programs != nullptr, "programs missing");
        for (uint32_t d = 0; d < deciders; ++d) {
            validate_program(spec.programs[d]);
            require(spec.programs[d].feature_spec_id == 2, "evaluation-count entry points need feature spec 2");
        }
    }
    if (spec.mode == FACO_CONTROL_CONTROLLER) {
        require(spec.controllers != nullptr, "controllers missing");
        for (uint32_t d = 0; d < deciders; ++d) validate_controller(spec.controllers[d]);
    }
    if (spec.mode == FACO_CONTROL_BASELINE) validate_baseline(spec.baseline, spec.experiment_mask);
    if (spec.mode == FACO_CONTROL_FACTORIAL) {
        require(spec.factorial_variant <= 3, "unknown factorial variant");
        validate_baseline(spec.baseline, 0xffffffffu);
        if (spec.factorial_variant == 0 || spec.factorial_variant == 2)
            validate_baseline(spec.baseline, spec.experiment_mask);
        require(deciders == 1 || spec.factorial_variant == 1 || spec.factorial_variant == 2,
                "population factorial runs take M10 or M01 only");
    }
    const uint32_t C = panel * deciders;
    require(C <= e.colonies_capacity, "colonies exceed the engine's capacity");
    for (uint32_t t = 0; t < panel; ++t) (void)e.problem(panel_keys[t]);
    // state snapshots (reference cuda/batch_engine.cu:536-554): plain
    // program or baseline count runs; the progress feature keeps the
    // source run's budget and a resumed run keeps its batch numbers
    const bool capturing = io && io->capture, resuming = io && io->resume;
    const faco_fork fork = io ? io->fork : faco_fork{0, 0, 0, 2, 0};
    uint64_t progress = limit, run_limit = limit;
    if (capturing || resuming) {
        require(C <= 128, "a state snapshot holds at most 128 colonies");
        require(!e.hard, "state snapshots take unrestricted engines (reference batch_engine.cu:543-548)");
        require(!e.mixed && deciders == 1 && (spec.mode == FACO_CONTROL_PROGRAM || spec.mode == FACO_CONTROL_BASELINE) &&
                !spec.force_fingerprint_collisions && !(flags & FACO_RUN_EVENTS) && !(checkpoints && checkpoints->count),
                "state snapshots take plain program or baseline count runs, without profiling, forced collisions "
                "or checkpoints");
    }
    if (capturing) {
        require(limit > 0 && io->capture_after <= limit && io->capture_after % ants == 0,
                "the capture point must be a whole number of ant batches inside the source budget");
        run_limit = io->capture_after;
    }
    if (resuming) {
        const CountedState& saved = *io->resume;
        saved.validate();
        require(saved.dimension == e.n && saved.colonies == C && same_settings(saved.settings, e.config),
                "snapshot shape or solver settings differ from the engine");
        require(limit <= saved.progress_evaluation_limit - saved.completed_batches * ants,
                "a continuation cannot pass the source progress budget");
        progress = saved.progress_evaluation_limit;
    }
    if (fork.enabled)
        require(resuming && spec.mode == FACO_CONTROL_BASELINE && limit > 0 && fork.region < 4 &&
                (fork.mne == 2 || fork.mne == 16),
                "a paired fork starts from a snapshot, runs a baseline continuation and takes MNE 2 or 16 and a region 0..3");
    else
        require(fork.seed == 0 && fork.region == 0 && fork.mne == 2, "fork fields set without an enabled fork");
    std::vector<uint64_t> keys(C), inst(C);
    for (uint32_t c = 0; c < C; ++c) {
        inst[c] = panel_keys[c % panel];
        keys[c] = colony_key(panel_keys[c % panel], panel_seeds[c % panel]);
    }
    // a mixed-size batch runs until its longest colony is done; the others
    // stop at their own budget
    struct Inactive { Engine& e; ~Inactive() { e.shapes_active = false; } } inactive{e};
    const uint64_t batches = e.mixed ? e.prepare_shapes(inst.data(), C, limit) : run_limit / ants;
    std::vector<uint32_t> traced;
    if (trace && trace->count) {
        require(trace->batches != nullptr, "trace batches missing");
        traced.assign(trace->batches, trace->batches + trace->count);
        for (size_t i = 0; i < traced.size(); ++i)
            require(traced[i] < batches && (i == 0 || traced[i] > traced[i - 1]),
                    "trace batches must be ascending and inside the budget");
    }
    e.ensure_control();
    auto& b = *e.ctl;
    // LDS of the archive update: per candidate cost, identity, start, leader, pair, parent flag
    require(control_candidate_bytes(ants) + 4096u <= max_lds_per_block(), "too many ants for the archive update");
    e.upload_geometry(inst.data(), C);
    if (!resuming) e.keys.upload(keys.data(), C, e.stream);
    if (spec.mode == FACO_CONTROL_PROGRAM || spec.mode == FACO_CONTROL_FACTORIAL)
        b.programs.upload(spec.programs, deciders, e.stream);
    if (spec.mode == FACO_CONTROL_CONTROLLER) b.controllers.upload(spec.controllers, deciders, e.stream);

    ControlParams cp = control_params(e, panel, spec, progress);
    const AntParams ap = controlled_ant_params(e, C);
    const ColonyParams colony = e.colony_params();
    // FACO_RUN_EVENTS: HIP events between the launches of every batch;
    // kernel_ms [0] construction + LS, [1] colony update, [2] decision
    // (features, regions, scoring, action and restart), [3] archive and
    // feedback update, [4] initialisation, all summed over batches
    const bool events = (flags & FACO_RUN_EVENTS) != 0 && traced.empty();
    double kernel_ms[FACO_STAGE_COUNT] = {0};
    hipEvent_t ev[5] = {};
    struct Destroy { hipEvent_t* e; ~Destroy() { for (int i = 0; i < 5; ++i) if (e[i]) (void)hipEventDestroy(e[i]); } } destroy{ev};
    if (events)
        for (auto& x : ev) checked(hipEventCreate(&x), "hipEventCreate");
    auto elapsed = [&](int a, int b) {
        float ms = 0;
        checked(hipEventElapsedTime(&ms, ev[a], ev[b]), "event time");
        return static_cast<double>(ms);
    };
    // a paired fork's first decision takes the forced action, with the start
    // regions drawn under the source keys (reference batch_engine.cu:883-923:
    // that batch keeps the snapshot's decision views); the ants and every
    // later batch use the forked keys
    ControlParams forced = cp;
    forced.forced_action = static_cast<int32_t>(fork.region * 4 + (fork.mne == 2 ? 0 : 3));
    forced.keys = b.source_keys.data();
    uint64_t launched = 0;
    auto iteration = [&]() {
        const bool first_fork = fork.enabled && launched++ == 0;
        if (events) checked(hipEventRecord(ev[0], e.stream), "event");
        checked(launch_control_decide(first_fork ? forced : cp, C, e.stream), "control decide launch");
        if (events) checked(hipEventRecord(ev[1], e.stream), "event");
        checked(launch_ant_search(ap, C * ap.blocks_per_colony, e.stream), "ant search launch");
        if (events) checked(hipEventRecord(ev[2], e.stream), "event");
        checked(launch_colony_update(colony, C, e.stream), "colony update launch");
        if (events) checked(hipEventRecord(ev[3], e.stream), "event");
        checked(launch_control_update(cp, C, e.stream), "control update launch");
        if (!events) return;
        checked(hipEventRecord(ev[4], e.stream), "event");
        checked(hipEventSynchronize(ev[4]), "event sync");
        kernel_ms[2] += elapsed(0, 1); kernel_ms[0] += elapsed(1, 2);
        kernel_ms[1] += elapsed(2, 3); kernel_ms[3] += elapsed(3, 4);
    };

    Checkpoints marks(checkpoints, batches, static_cast<size_t>(C) * e.n);
    if (events) checked(hipEventRecord(ev[0], e.stream), "event");
    std::vector<ColonyState> resumed(resuming ? C : 0);   // work counters of the snapshot
    std::vector<uint64_t> source_keys, forked_keys;
    if (resuming) {
        restore_state(e, *io->resume, C, inst);
        std::memcpy(resumed.data(), io->resume->buffer("colony_state").bytes.data(), C * sizeof(ColonyState));
        if (fork.enabled) {
            // paired forks of one snapshot and seed share their keys
            // (reference batch_engine.cu:839-845)
            source_keys.resize(C);
            std::memcpy(source_keys.data(), io->resume->buffer("keys").bytes.data(), C * sizeof(uint64_t));
            for (uint64_t key : source_keys) forked_keys.push_back(mixed(key ^ mixed(fork.seed + 0xa0761d6478bd642fULL)));
            b.source_keys.upload(source_keys.data(), C, e.stream);
            e.keys.upload(forked_keys.data(), C, e.stream);
        }
    } else {
        e.reset(C, 2u);
        checked(launch_control_init(cp, C, e.stream), "control init launch");
    }
    if (events) {
        checked(hipEventRecord(ev[1], e.stream), "event");
        checked(hipEventSynchronize(ev[1]), "event sync");
        kernel_ms[4] = elapsed(0, 1);
    }
    marks.record(e.global.data(), 0, e.stream);
    checked(hipStreamSynchronize(e.stream), "initialisation");
    const auto search_started = Clock::now();
    checked(hipEventRecord(e.ev_start, e.stream), "event");
    if (!traced.empty()) {
        size_t next = 0;
        std::vector<ColonyState> st(C);
        for (uint64_t batch = 0; batch < batches; ++batch) {
            iteration();
            marks.record(e.global.data(), batch + 1, e.stream);
            if (next < traced.size() && traced[next] == batch) {
                const size_t k = next++;
                if (trace->features) b.features.download(trace->features + k * 12 * 32 * C, 12 * 32 * C, e.stream);
                if (trace->scores) b.scores.download(trace->scores + k * 32 * C, 32 * C, e.stream);
                if (trace->masks) b.masks.download(trace->masks + k * C, C, e.stream);
                if (trace->actions) b.actions.download(trace->actions + k * C, C, e.stream);
                if (trace->alternatives) b.alternatives.download(trace->alternatives + k * C, C, e.stream);
                if (trace->regions) b.regions.download(trace->regions + k * 128 * C, 128 * C, e.stream);
                e.state.download(st.data(), C, e.stream);
                checked(hipStreamSynchronize(e.stream), "trace batch");
                if (trace->costs)
                    for (uint32_t c = 0; c < C; ++c) {
                        trace->costs[(k * C + c) * 3] = st[c].epoch_cost;
                        trace->costs[(k * C + c) * 3 + 1] = st[c].global_cost;
                        trace->costs[(k * C + c) * 3 + 2] = st[c].parent_cost;
                    }
            }
        }
    } else {
        // segments between checkpoints. With one workgroup per colony each
        // segment is one persistent launch (decision, ants, colony update and
        // archive update per batch in the colony's workgroup); otherwise it
        // replays a captured graph of 16 batches where it can and runs the
        // rest eagerly
        constexpr uint32_t kGraphIterations = 16;
        const bool graphs = (flags & FACO_RUN_GRAPH) != 0 && !events;
        const bool fused = graphs && e.fuse_control && fused_control(ap);
        if (fused) b.params.upload(&cp, 1, e.stream);
        hipGraphExec_t exec = nullptr;
        struct ExecGuard { hipGraphExec_t& x; ~ExecGuard() { if (x) (void)hipGraphExecDestroy(x); } } guard{exec};
        uint64_t done = 0;
        if (fork.enabled && batches) { iteration(); done = 1; }   // eager: the graph replays later batches
        while (done < batches) {
            const uint64_t stop = std::max(marks.next(batches), done + 1);
            uint64_t left = stop - done;
            for (; fused && left;) {
                const uint64_t rounds = std::min(left, Engine::kRoundsPerLaunch);
                e.by_class(ap, C, fused_control(e.small_class(ap, C)),
                           [&](const AntParams& p, uint32_t count, hipStream_t s) {
                               checked(launch_control_run(p, colony, cp, b.params.data(), count,
                                                          static_cast<uint32_t>(rounds), s), "control run launch");
                           });
                ++e.persistent_launch_count;
                left -= rounds;
            }
            if (graphs && left >= kGraphIterations) {
                if (!exec) {
                    e.drop_graph();
                    hipGraph_t g = nullptr;
                    checked(hipStreamBeginCapture(e.stream, hipStreamCaptureModeThreadLocal), "begin capture");
                    for (uint32_t i = 0; i < kGraphIterations; ++i) iteration();
                    checked(hipStreamEndCapture(e.stream, &g), "end capture");
                    checked(hipGraphInstantiate(&exec, g, nullptr, nullptr, 0), "graph instantiate");
                    (void)hipGraphDestroy(g);
                }
                for (; left >= kGraphIterations; left -= kGraphIterations)
                    checked(hipGraphLaunch(exec, e.stream), "graph launch");
            }
            for (; left; --left) iteration();
            done = stop;
            marks.record(e.global.data(), done, e.stream);
        }
        checked(hipStreamSynchronize(e.stream), "search");
    }
    checked(hipEventRecord(e.ev_stop, e.stream), "event");
    checked(hipStreamSynchronize(e.stream), "search");
    const double search = seconds_since(search_started);
    float gpu_ms = 0;
    checked(hipEventElapsedTime(&gpu_ms, e.ev_start, e.ev_stop), "event time");
    std::vector<ColonyState> st(C);
    std::vector<ControlState> cs(C);
    e.state.download(st.data(), C, e.stream);
    b.control.download(cs.data(), C, e.stream);
    if (tours) e.global.download(tours, static_cast<size_t>(C) * e.n, e.stream);
    checked(hipStreamSynchronize(e.stream), "readback");
    if (tours && e.mixed)
        for (uint32_t c = 0; c < C; ++c)
            for (uint32_t i = e.problem(inst[c]).n; i < e.n; ++i) tours[static_cast<size_t>(c) * e.n + i] = UINT32_MAX;
    marks.read_back(e.stream, e.n, [&](uint32_t c) { return e.problem(inst[c]).n; });
    uint64_t steps = 0, evals = 0, rejections = 0;
    e.last_rejections.assign(C, 0);
    e.read_escape(C);
    for (uint32_t c = 0; c < C; ++c) {
        e.last_rejections[c] = st[c].work_constraints - (resuming ? resumed[c].work_constraints : 0);
        rejections += e.last_rejections[c];
        // a resumed run reports its own work, not the snapshot's
        const uint64_t own_steps = st[c].work_construction - (resuming ? resumed[c].work_construction : 0);
        const uint64_t own_evals = st[c].work_ls - (resuming ? resumed[c].work_ls : 0);
        if (costs) costs[c] = st[c].global_cost;
        if (work) { work[2 * c] = own_steps; work[2 * c + 1] = own_evals; }
        if (states) {
            faco_control_state& o = states[c];
            o.archive_size = cs[c].archive_size; o.pad = 0;
            o.stagnant_batches = cs[c].stagnant; o.epoch_batches = cs[c].epoch_batches;
            o.restarts = cs[c].restarts; o.return_rate = cs[c].return_rate; o.ls_work = cs[c].ls_work;
        }
        steps += own_steps; evals += own_evals;
    }
    if (capturing) {
        // the next decision's views under the source keys (reference
        // batch_engine.cu:207-229), then every buffer of the state
        ControlParams views = cp;
        views.views_only = 1;
        checked(launch_control_decide(views, C, e.stream), "decision views launch");
        CountedState& out = *io->capture;
        out = CountedState{};
        out.dimension = e.n; out.colonies = C; out.settings = e.config;
        out.experiment_mask = spec.experiment_mask;
        out.completed_batches = batches; out.progress_evaluation_limit = progress;
        for (uint32_t c = 0; c < C; ++c) out.tasks.emplace_back(panel_keys[c], panel_seeds[c]);
        const auto slots = state_slots(e, C);
        out.buffers.resize(slots.size() + 1);
        out.buffers[0].name = "xy"; out.buffers[0].element_bytes = sizeof(int32_t);
        out.buffers[0].bytes = state_coordinates(e, inst);
        for (size_t i = 0; i < slots.size(); ++i) {
            StateBuffer& target = out.buffers[i + 1];
            target.name = slots[i].name; target.element_bytes = slots[i].element; target.planes = slots[i].planes;
            target.bytes.resize(slots[i].count * slots[i].element);
            checked(hipMemcpyAsync(target.bytes.data(), slots[i].device, target.bytes.size(), hipMemcpyDeviceToHost,
                                   e.stream), "snapshot download");
        }
        checked(hipStreamSynchronize(e.stream), "snapshot download");
        const auto& global = out.buffer("global").bytes;
        for (uint32_t c = 0; c < C; ++c) {
            StateIncumbent item;
            item.cost = static_cast<double>(st[c].global_cost);
            item.tour.resize(e.n);
            std::memcpy(item.tour.data(), global.data() + static_cast<size_t>(c) * e.n * 4, static_cast<size_t>(e.n) * 4);
            out.incumbents.push_back(std::move(item));
        }
        out.validate();
    }
    fill_result(result, e, C, limit, batches, e.mixed ? limit : batches * e.config.ants, steps, evals,
                seconds_since(started), search, gpu_ms);
    if (result) {
        std::copy(kernel_ms, kernel_ms + FACO_STAGE_COUNT, result->kernel_milliseconds);
        result->completed_constraint_rejections = rejections;
        result->persistent_launches = e.persistent_launch_count - launches_before;
    }
}

// Time-budget evaluation (reference cpp/include/gp_faco/deadline.hpp and
// the budget path of cuda/batch_engine.cu:492-1192): the clock starts with
// the call; charges (cached_charged preparation) add to it. An incumbent is
// accepted only when offered no later than the budget and strictly shorter
// than the one held.
struct Deadline {
    Clock::time_point started;
    double budget, charged = 0;
    double elapsed() const { return seconds_since(started) + charged; }
    bool can_start() const { return elapsed() < budget; }
    bool expired() const { return elapsed() > budget; }
};

struct Incumbent {
    std::vector<uint32_t> tour;
    double cost = 0, completed = 0;
    bool present = false;
};

void evaluate_deadline(Engine& e, const uint64_t* instance_keys, const uint64_t* seeds, uint32_t tasks,
                       const faco_deadline_spec& spec, faco_deadline_result* result, uint32_t* tours, double* costs,
                       uint8_t* has_incumbent, double* completed_seconds, faco_control_state* states, double* ratios,
                       uint32_t ratio_capacity, faco_batch_profile* profile) {
    Deadline clock{Clock::now(), spec.budget_seconds, 0};
    std::unique_lock<std::mutex> lock(e.mutex, std::try_to_lock);
    require(lock.owns_lock(), "one evaluation per engine at a time");
    require(current_device() == e.device, "engine must be used on the device it was created on");
    // a Hard engine takes evaluation counts only: a cheap tour off the graph
    // must not be published (reference cuda/batch_engine.cu:539-540)
    require(!e.mixed && !e.hard && e.population == 1,
            "time-budget evaluations take a fixed-size, unrestricted engine of population 1");
    require(tasks == e.panel, "task count must equal the engine's colonies");
    require(std::isfinite(spec.budget_seconds) && spec.budget_seconds >= 0, "budget seconds must be finite and >= 0");
    require(spec.preparation_mode <= 1, "unknown preparation mode");
    require(spec.fixed_batches <= 10000, "diagnostic batch count must be in 1..10000");
    const bool fixed_ratio = spec.fixed_elapsed_ratio >= 0;
    require(std::isfinite(spec.fixed_elapsed_ratio) && (!fixed_ratio ||
            (spec.fixed_batches > 0 && spec.fixed_elapsed_ratio <= 1)),
            "a fixed progress feature lies in [0, 1] and needs a fixed batch count");
    require(!spec.profile || (spec.fixed_batches > 0 && fixed_ratio),
            "a profile needs a fixed batch count and a fixed progress feature");
    const bool controlled = spec.program != nullptr;
    if (controlled) {
        validate_program(*spec.program);
        require(spec.program->feature_spec_id == 1, "time-budget programs read progress feature spec 1");
        require((spec.experiment_mask & 0xffffu) != 0, "the experiment mask must keep at least one hold action");
    } else {
        require(spec.mne_target > 0, "mne_target must be positive");
    }
    for (uint32_t c = 0; c < tasks; ++c) (void)e.problem(instance_keys[c]);
    e.forget_run();
    const uint32_t n = e.n;
    std::vector<Incumbent> held(tasks);
    faco_deadline_result r{};
    r.budget_seconds = spec.budget_seconds;
    r.allocated_device_bytes = e.allocated;
    auto offer = [&](const std::vector<uint32_t>& ids, const uint32_t* tour, double cost) {
        const double completed = clock.elapsed();
        if (completed > clock.budget) return;
        for (uint32_t c : ids) {
            Incumbent& h = held[c];
            if (h.present && !(cost < h.cost)) continue;
            h.tour.assign(tour, tour + n);
            h.cost = cost; h.completed = completed; h.present = true;
        }
    };
    if (states) std::fill(states, states + tasks, faco_control_state{});
    auto finish = [&]() {
        r.actual_seconds = seconds_since(clock.started);
        r.charged_seconds = clock.charged;
        r.elapsed_seconds = r.actual_seconds + r.charged_seconds;
        r.overrun_seconds = std::max(0.0, r.elapsed_seconds - clock.budget);
        for (uint32_t c = 0; c < tasks; ++c) {
            const Incumbent& h = held[c];
            if (has_incumbent) has_incumbent[c] = h.present ? 1 : 0;
            if (costs) costs[c] = h.present ? h.cost : 0.0;
            if (completed_seconds) completed_seconds[c] = h.present ? h.completed : 0.0;
            if (tours) {
                if (h.present) std::copy(h.tour.begin(), h.tour.end(), tours + static_cast<size_t>(c) * n);
                else std::fill(tours + static_cast<size_t>(c) * n, tours + static_cast<size_t>(c + 1) * n, UINT32_MAX);
            }
        }
        if (result) *result = r;
    };
    // colonies grouped by instance, in key order
    std::map<uint64_t, std::vector<uint32_t>> groups;
    for (uint32_t c = 0; c < tasks; ++c) groups[instance_keys[c]].push_back(c);
    auto charge_of = [&](uint64_t key, bool cheap) {
        const auto frozen = e.charges.find(key);
        const Prepared& p = e.problem(key);
        if (frozen != e.charges.end()) return cheap ? frozen->second.first : frozen->second.second;
        return cheap ? p.cheap_seconds : p.preparation_seconds;
    };
    // preparation: every instance's cheap tour first, then its prepared
    // initial tour (end_to_end prepares again inside the budget; the result
    // is the registered preparation, which is deterministic)
    const auto setup = Clock::now();
    for (const auto& [key, ids] : groups) {
        if (!clock.can_start()) return finish();
        const Prepared& p = e.problem(key);
        if (spec.preparation_mode == 0) {
            clock.charged += charge_of(key, true);
            offer(ids, p.cheap.data(), static_cast<double>(p.cheap_cost));
        } else {
            const Prepared cheap = prepare_begin(p.xy, e.config);
            offer(ids, cheap.cheap.data(), static_cast<double>(cheap.cheap_cost));
        }
        if (clock.expired()) return finish();
    }
    for (const auto& [key, ids] : groups) {
        if (!clock.can_start()) return finish();
        const Prepared& p = e.problem(key);
        if (spec.preparation_mode == 0) {
            clock.charged += charge_of(key, false);
            offer(ids, p.initial.data(), static_cast<double>(p.initial_cost));
        } else {
            const Prepared fresh = prepare_problem(p.xy, e.config, e.K, e.nbr_stride);
            require(fresh.initial == p.initial, "end-to-end preparation differs from the registered one");
            offer(ids, fresh.initial.data(), static_cast<double>(fresh.initial_cost));
        }
        if (clock.expired()) return finish();
    }
    // upload and initialise the colonies
    if (controlled) e.ensure_control();
    e.upload_geometry(instance_keys, tasks);
    std::vector<uint64_t> keys(tasks);
    for (uint32_t c = 0; c < tasks; ++c) keys[c] = colony_key(instance_keys[c], seeds[c]);
    e.keys.upload(keys.data(), tasks, e.stream);
    r.setup_seconds = seconds_since(setup);
    hipEvent_t ev[6] = {};
    struct Destroy { hipEvent_t* x; ~Destroy() { for (int i = 0; i < 6; ++i) if (x[i]) (void)hipEventDestroy(x[i]); } } destroy{ev};
    if (spec.profile)
        for (auto& x : ev) checked(hipEventCreate(&x), "hipEventCreate");
    auto ms = [&](int a, int b) {
        float v = 0;
        checked(hipEventElapsedTime(&v, ev[a], ev[b]), "event time");
        return static_cast<double>(v);
    };
    faco_control_spec cspec{};
    cspec.mode = FACO_CONTROL_PROGRAM; cspec.deciders = 1; cspec.programs = spec.program;
    cspec.experiment_mask = spec.experiment_mask;
    if (spec.profile) checked(hipEventRecord(ev[0], e.stream), "event");
    ControlParams cp{};
    if (controlled) {
        e.ctl->programs.upload(spec.program, 1, e.stream);
        cp = control_params(e, tasks, cspec, 1);
        e.reset(tasks, 2u);
        checked(launch_control_init(cp, tasks, e.stream), "control init launch");
    } else {
        e.reset(tasks, spec.mne_target);
    }
    if (spec.profile) {
        checked(hipEventRecord(ev[1], e.stream), "event");
        checked(hipEventSynchronize(ev[1]), "event sync");
        r.initialization_gpu_milliseconds = ms(0, 1);
    }
    std::vector<ControlState> cs(controlled ? tasks : 0), committed_cs;
    if (controlled) e.ctl->control.download(cs.data(), tasks, e.stream);
    checked(hipStreamSynchronize(e.stream), "initialisation");
    auto write_states = [&]() {
        if (!states) return;
        for (uint32_t c = 0; c < tasks; ++c) {
            faco_control_state& o = states[c];
            o = faco_control_state{};
            if (committed_cs.empty()) continue;
            const ControlState& k = committed_cs[c];
            o.archive_size = k.archive_size; o.stagnant_batches = k.stagnant; o.epoch_batches = k.epoch_batches;
            o.restarts = k.restarts; o.return_rate = k.return_rate; o.ls_work = k.ls_work;
        }
    };
    if (clock.expired()) { write_states(); return finish(); }
    r.preparation_completed = 1;
    committed_cs = cs;   // the initial control states until a batch completes
    // per batch (eager launches, one synchronisation): decision, ants,
    // colony update, archive update; then the host checks every colony's
    // global tour and cost before offering it
    const AntParams ap = controlled ? controlled_ant_params(e, tasks) : e.ant_params(tasks);
    const ColonyParams colony = e.colony_params();
    std::vector<ColonyState> st(tasks);
    std::vector<uint32_t> global(static_cast<size_t>(tasks) * n);
    for (uint32_t batch = 0; clock.can_start(); ++batch) {
        if (spec.fixed_batches && batch >= spec.fixed_batches) break;
        ++r.launched_batches;
        const auto batch_started = Clock::now();
        faco_batch_profile row{};
        row.batch = batch;
        if (spec.profile) checked(hipEventRecord(ev[0], e.stream), "event");
        if (controlled) {
            const double ratio = fixed_ratio ? spec.fixed_elapsed_ratio : clock.elapsed() / clock.budget;
            if (ratios && batch < ratio_capacity) ratios[batch] = ratio;
            cp.fixed_progress = ratio;
            checked(launch_control_decide(cp, tasks, e.stream), "control decide launch");
        }
        if (spec.profile) checked(hipEventRecord(ev[1], e.stream), "event");
        checked(launch_ant_search(ap, tasks * ap.blocks_per_colony, e.stream), "ant search launch");
        if (spec.profile) checked(hipEventRecord(ev[2], e.stream), "event");
        checked(launch_colony_update(colony, tasks, e.stream), "colony update launch");
        if (spec.profile) checked(hipEventRecord(ev[3], e.stream), "event");
        if (controlled) checked(launch_control_update(cp, tasks, e.stream), "control update launch");
        if (spec.profile) checked(hipEventRecord(ev[4], e.stream), "event");
        checked(hipStreamSynchronize(e.stream), "batch");
        const auto download_started = Clock::now();
        e.state.download(st.data(), tasks, e.stream);
        e.global.download(global.data(), global.size(), e.stream);
        if (controlled) e.ctl->control.download(cs.data(), tasks, e.stream);
        checked(hipStreamSynchronize(e.stream), "batch readback");
        row.download_seconds = seconds_since(download_started);
        const auto verify_started = Clock::now();
        for (uint32_t c = 0; c < tasks; ++c) {
            const uint32_t* t = global.data() + static_cast<size_t>(c) * n;
            if (!is_permutation(t, n) || tour_cost(e.problem(instance_keys[c]), t) != st[c].global_cost)
                throw std::runtime_error("device incumbent is not a tour of the cost its colony reports");
        }
        row.verification_seconds = seconds_since(verify_started);
        if (spec.profile) {
            const auto collection_started = Clock::now();
            row.gpu_milliseconds[1] = ms(0, 1);   // features_and_regions (decision with scoring and restart)
            row.gpu_milliseconds[4] = ms(1, 2);   // construction_and_ls
            row.gpu_milliseconds[5] = ms(2, 3);   // reduction (with the pheromone update)
            row.gpu_milliseconds[8] = ms(3, 4);   // archive_and_feedback (with the fingerprints)
            row.collection_seconds = seconds_since(collection_started);
        }
        const double completed = clock.elapsed();
        r.last_batch_completed_seconds = completed;
        row.wall_seconds = seconds_since(batch_started);
        if (completed > clock.budget) {
            ++r.discarded_batches;
            if (profile) profile[r.profiled_batches++] = row;
            break;
        }
        for (uint32_t c = 0; c < tasks; ++c) offer({c}, global.data() + static_cast<size_t>(c) * n,
                                                   static_cast<double>(st[c].global_cost));
        ++r.completed_batches;
        row.committed = 1;
        if (profile) profile[r.profiled_batches++] = row;
        r.completed_construction_steps = r.completed_ls_evaluations = 0;
        for (uint32_t c = 0; c < tasks; ++c) {
            r.completed_construction_steps += st[c].work_construction;
            r.completed_ls_evaluations += st[c].work_ls;
        }
        if (controlled) committed_cs = cs;
    }
    write_states();
    finish();
}

// Device preparation (the O(n^2) steps of prepare_problem on the GPU):
// candidate keys and nearest-neighbour tours from prepare_kernels.hip, rows
// on the host from the keys, then the initial checklist 2-opt of every
// instance in one explicit-mode launch of the ant kernel (one workgroup per
// instance, the checklist holding every city in tour order). Instances are
// grouped by their candidate count (equal unless an instance has fewer than
// 9 cities). Each instance's preparation time is its share of the total.
double register_problems_device(Engine& e, const uint64_t* keys, const double* xy, const uint32_t* sizes,
                                uint32_t count) {
    const auto started = Clock::now();
    std::unique_lock<std::mutex> lock(e.mutex, std::try_to_lock);
    require(lock.owns_lock(), "registration and evaluation of one engine must not overlap");
    e.forget_run();   // the preparation kernels reuse the ant buffers
    require(current_device() == e.device, "engine must be used on the device it was created on");
    std::map<uint32_t, std::vector<std::pair<uint64_t, Prepared>>> groups;   // candidate count -> instances
    size_t offset = 0;
    for (uint32_t i = 0; i < count; ++i) {
        const uint32_t n = sizes ? sizes[i] : e.n;
        if (e.mixed) {
            require(n >= 3 && n <= e.n, "instance larger than the mixed engine's maximum dimension");
            const faco_settings own = normalized_settings(e.base, n);
            require(own.primary_width == e.config.primary_width && own.backup_width == e.config.backup_width &&
                    own.ls_width == e.config.ls_width,
                    "a mixed-size batch needs the same candidate widths for every instance");
        } else {
            require(n == e.n, "instance dimension does not match the engine");
        }
        auto coords = integer_coordinates(xy + 2 * offset, n);
        offset += n;
        const auto existing = e.registry.find(keys[i]);
        if (existing != e.registry.end()) {
            require(existing->second.xy == coords, "same instance key registered with different coordinates");
            continue;
        }
        bool repeated = false;
        for (const auto& group : groups)
            for (const auto& item : group.second)
                if (item.first == keys[i]) {
                    require(item.second.xy == coords, "same instance key registered with different coordinates");
                    repeated = true;
                }
        if (repeated) continue;
        groups[candidate_count(e.config, n)].emplace_back(keys[i], prepare_begin(coords, e.config));
    }
    const uint32_t K = e.K, lw = e.config.ls_width;
    for (auto& [needed, items] : groups) {
        const uint32_t G = static_cast<uint32_t>(items.size());
        uint32_t stride = 0;
        for (const auto& item : items) stride = std::max(stride, item.second.n);
        require(prepare_knn_lds(stride, needed) <= max_lds_per_block(), "instance too large for device preparation");
        std::vector<uint32_t> packed(static_cast<size_t>(G) * stride, 0), n_of(G);
        for (uint32_t g = 0; g < G; ++g) {
            const Prepared& p = items[g].second;
            n_of[g] = p.n;
            std::copy(p.packed.begin(), p.packed.end(), packed.begin() + static_cast<size_t>(g) * stride);
        }
        const size_t cells = static_cast<size_t>(G) * stride;
        DeviceArray<uint32_t> d_coords, d_sizes, d_keys, d_nn;
        d_coords.allocate(cells); d_sizes.allocate(G); d_keys.allocate(cells * needed); d_nn.allocate(cells);
        d_coords.upload(packed.data(), cells, e.stream);
        d_sizes.upload(n_of.data(), G, e.stream);
        checked(launch_prepare_knn(d_coords.data(), d_sizes.data(), G, stride, needed, d_keys.data(), e.stream),
                "candidate selection launch");
        checked(launch_prepare_nn(d_coords.data(), d_sizes.data(), G, stride, d_nn.data(), e.stream),
                "nearest-neighbour launch");
        std::vector<uint32_t> key_rows(cells * needed), nn(cells);
        d_keys.download(key_rows.data(), cells * needed, e.stream);
        d_nn.download(nn.data(), cells, e.stream);
        checked(hipStreamSynchronize(e.stream), "device preparation");
        // rows on the host, then the initial 2-opt of every instance on the device
        std::vector<uint16_t> nbr(static_cast<size_t>(G) * e.nbr_stride), nbr_dist(static_cast<size_t>(G) * e.nbr_stride);
        std::vector<uint32_t> links(cells, 0), zeros(G, 0);
        std::vector<int64_t> nn_costs(G);
        for (uint32_t g = 0; g < G; ++g) {
            Prepared& p = items[g].second;
            prepare_rows(p, e.config, K, e.nbr_stride, key_rows.data() + static_cast<size_t>(g) * stride * needed, needed);
            std::copy(p.nbr.begin(), p.nbr.end(), nbr.begin() + static_cast<size_t>(g) * e.nbr_stride);
            std::copy(p.nbr_dist.begin(), p.nbr_dist.end(), nbr_dist.begin() + static_cast<size_t>(g) * e.nbr_stride);
            const uint32_t* t = nn.data() + static_cast<size_t>(g) * stride;
            require(is_permutation(t, p.n), "device nearest-neighbour tour is not a permutation");
            for (uint32_t i = 0; i < p.n; ++i)
                links[static_cast<size_t>(g) * stride + t[i]] = t[i == 0 ? p.n - 1 : i - 1] | (t[i + 1 == p.n ? 0 : i + 1] << 16);
            nn_costs[g] = tour_cost(p, t);
        }
        DeviceArray<uint16_t> d_nbr, d_nbr_dist;
        DeviceArray<uint32_t> d_links, d_targets, d_out;
        DeviceArray<int64_t> d_costs;
        DeviceArray<faco_ant_info> d_info;
        d_nbr.allocate(nbr.size()); d_nbr_dist.allocate(nbr_dist.size()); d_links.allocate(cells);
        d_targets.allocate(G); d_out.allocate(cells); d_costs.allocate(G); d_info.allocate(G);
        d_nbr.upload(nbr.data(), nbr.size(), e.stream); d_nbr_dist.upload(nbr_dist.data(), nbr_dist.size(), e.stream);
        d_links.upload(links.data(), cells, e.stream); d_targets.upload(zeros.data(), G, e.stream);
        d_costs.upload(nn_costs.data(), G, e.stream);
        AntParams ap{};
        ap.n = stride; ap.ants = G; ap.pw = 0; ap.bw = 0; ap.lw = lw; ap.K = K; ap.waves = 1; ap.chunk = 1;
        ap.blocks_per_colony = 1; ap.nbr_stride = e.nbr_stride; ap.nbr_in_lds = 1;
        ap.nbr_lds_bytes = round16_host(2u * stride * K); ap.per_wave_bytes = ant_wave_bytes(stride);
        ap.explicit_mode = 1; ap.per_ant_parent = 1; ap.full_checklist = 1; ap.block_n = d_sizes.data();
        ap.ls_limit = std::min<uint64_t>(e.config.initial_ls_evaluation_limit, UINT32_MAX);
        ap.roulette_margin = kRouletteMargin;
        ap.coords = d_coords.data(); ap.nbr = d_nbr.data(); ap.nbr_dist = d_nbr_dist.data();
        ap.parent = d_nn.data(); ap.links = d_links.data(); ap.visits = d_nn.data(); ap.targets = d_targets.data();
        ap.parent_cost = d_costs.data(); ap.out_tours = d_out.data(); ap.out_info = d_info.data();
        if (ant_shared_bytes(stride, ap.nbr_lds_bytes) + ap.per_wave_bytes > max_lds_per_block()) {
            ap.nbr_in_lds = 0;
            require(ant_shared_bytes(stride, 0) + ap.per_wave_bytes <= max_lds_per_block(),
                    "instance too large for the LDS budget");
        }
        checked(launch_ant_search(ap, G, e.stream), "initial local search launch");
        std::vector<uint32_t> improved(cells);
        d_out.download(improved.data(), cells, e.stream);
        checked(hipStreamSynchronize(e.stream), "initial local search");
        for (uint32_t g = 0; g < G; ++g) {
            Prepared& p = items[g].second;
            const uint32_t* t = improved.data() + static_cast<size_t>(g) * stride;
            require(is_permutation(t, p.n), "device 2-opt tour is not a permutation");
            prepare_finish(p, std::vector<uint32_t>(t, t + p.n));
        }
    }
    const double total = seconds_since(started);
    uint32_t fresh = 0;
    for (const auto& group : groups) fresh += static_cast<uint32_t>(group.second.size());
    for (auto& group : groups)
        for (auto& item : group.second) {
            item.second.preparation_seconds = total / fresh;
            e.registry.emplace(item.first, std::move(item.second));
        }
    return total;
}

// JSON view of a snapshot: the fields of the reference's describe()
// (cpp/src/bindings.cpp:422-463), read from this engine's buffer ABI.
struct Json {
    std::string text;
    void raw(const std::string& v) { text += v; }
    void number(uint64_t v) { text += std::to_string(v); }
    void real(double v) {
        char buffer[40];
        std::snprintf(buffer, sizeof buffer, "%.17g", v);
        std::string value(buffer);
        if (value.find_first_of(".en") == std::string::npos) value += ".0";
        text += value;
    }
    void string(const std::string& v) {   // JSON string with escapes (names come from the file)
        text += '"';
        for (const unsigned char c : v) {
            if (c == '"' || c == '\\') { text += '\\'; text += static_cast<char>(c); }
            else if (c < 0x20 || c >= 0x7f) {
                char buffer[8];
                std::snprintf(buffer, sizeof buffer, "\\u%04x", c);
                text += buffer;
            } else {
                text += static_cast<char>(c);
            }
        }
        text += '"';
    }
    void key(const char* name) {
        if (!text.empty() && text.back() != '{' && text.back() != '[') text += ',';
        text += '"'; text += name; text += "\":";
    }
    void item() { if (!text.empty() && text.back() != '[' && text.back() != '{') text += ','; }
};

template<class T> T state_record(const CountedState& s, const char* name, uint32_t colony) {
    const StateBuffer& b = s.buffer(name);
    require(b.element_bytes == sizeof(T) && b.planes == 1 && b.bytes.size() >= (colony + 1ull) * sizeof(T),
            "snapshot field does not match the native ABI");
    T value;
    std::memcpy(&value, b.bytes.data() + static_cast<size_t>(colony) * sizeof(T), sizeof(T));
    return value;
}

std::string describe_state(const CountedState& s) {
    s.validate();
    const StateBuffer& regions = s.buffer("regions");
    require(regions.element_bytes == 4 && regions.bytes.size() == static_cast<size_t>(s.colonies) * 128 * 4,
            "snapshot field does not match the native ABI");
    Json j;
    j.raw("{");
    j.key("state_spec_id"); j.number(2);
    j.key("native_abi_bound"); j.raw("true");
    j.key("dimension"); j.number(s.dimension);
    j.key("colonies"); j.number(s.colonies);
    j.key("completed_batches"); j.number(s.completed_batches);
    j.key("completed_evaluations_per_colony"); j.number(s.completed_batches * s.settings.ants);
    j.key("progress_evaluation_limit"); j.number(s.progress_evaluation_limit);
    j.key("experiment_mask"); j.number(s.experiment_mask);
    const faco_settings& c = s.settings;
    j.key("settings"); j.raw("{");
    j.key("ants"); j.number(c.ants); j.key("primary_width"); j.number(c.primary_width);
    j.key("backup_width"); j.number(c.backup_width); j.key("ls_width"); j.number(c.ls_width);
    j.key("beta"); j.real(c.beta); j.key("retention"); j.real(c.retention); j.key("p_best"); j.real(c.p_best);
    j.key("epoch_source_probability"); j.real(c.epoch_source_probability);
    j.key("ls_evaluation_limit"); j.number(c.ls_evaluation_limit);
    j.key("initial_ls_evaluation_limit"); j.number(c.initial_ls_evaluation_limit);
    j.raw("}");
    j.key("random_keys"); j.raw("[");
    for (uint32_t i = 0; i < s.colonies; ++i) { j.item(); j.number(state_record<uint64_t>(s, "keys", i)); }
    j.raw("]");
    j.key("tasks"); j.raw("[");
    for (const auto& t : s.tasks) {
        j.item(); j.raw("{");
        j.key("instance_key"); j.number(t.first); j.key("seed"); j.number(t.second);
        j.raw("}");
    }
    j.raw("]");
    j.key("members"); j.raw("[");
    const uint32_t count = s.dimension < kRegion ? s.dimension : kRegion;
    for (uint32_t i = 0; i < s.colonies; ++i) {
        const auto control = state_record<ControlState>(s, "control_state", i);
        j.item(); j.raw("{");
        j.key("incumbent_cost"); j.real(s.incumbents[i].cost);
        j.key("return_rate"); j.real(control.return_rate);
        j.key("ls_work"); j.real(control.ls_work);
        j.key("stagnant_batches"); j.number(control.stagnant);
        j.key("epoch_batches"); j.number(control.epoch_batches);
        j.key("restarts"); j.number(control.restarts);
        j.key("archive_size"); j.number(control.archive_size);
        j.key("keep_regions"); j.raw("[");
        for (uint32_t r = 0; r < 4; ++r) {
            j.item(); j.raw("[");
            for (uint32_t k = 0; k < count; ++k) {
                uint32_t node;
                std::memcpy(&node, regions.bytes.data() + ((static_cast<size_t>(i) * 128 + r * 16 + k) * 4), 4);
                j.item(); j.number(node);
            }
            j.raw("]");
        }
        j.raw("]}");
    }
    j.raw("]");
    j.key("buffers"); j.raw("[");
    for (const auto& b : s.buffers) {
        j.item(); j.raw("{");
        j.key("name"); j.string(b.name);
        j.key("bytes"); j.number(b.bytes.size());
        j.key("element_bytes"); j.number(b.element_bytes);
        j.key("planes"); j.number(b.planes);
        j.raw("}");
    }
    j.raw("]}");
    return j.text;
}

void copy_out(const std::vector<uint8_t>& data, uint8_t* out, size_t capacity, size_t* size) {
    if (size) *size = data.size();
    if (!out) return;
    require(capacity >= data.size(), "output buffer too small");
    std::memcpy(out, data.data(), data.size());
}

}  // namespace

extern "C" {

const char* faco_last_error(void) { return g_error.c_str(); }

int faco_device_info(int* device, char* name, size_t name_size, char* arch, size_t arch_size,
                     int* compute_units, size_t* total_memory, size_t* max_lds) {
    return guarded([&] {
        int d = current_device();
        hipDeviceProp_t prop;
        checked(hipGetDeviceProperties(&prop, d), "hipGetDeviceProperties");
        if (device) *device = d;
        if (name && name_size) { std::strncpy(name, prop.name, name_size - 1); name[name_size - 1] = 0; }
        if (arch && arch_size) { std::strncpy(arch, prop.gcnArchName, arch_size - 1); arch[arch_size - 1] = 0; }
        if (compute_units) *compute_units = prop.multiProcessorCount;
        if (total_memory) *total_memory = prop.totalGlobalMem;
        if (max_lds) *max_lds = prop.sharedMemPerBlock;
    });
}

int faco_prepare_common_initial(const double* xy, uint32_t n, const faco_settings* settings,
                                faco_settings* normalized, uint32_t* tour, int64_t* cost,
                                double* cheap_seconds, double* preparation_seconds, uint32_t* primary,
                                uint32_t* backup, uint32_t* ls, double* local_scale) {
    return guarded([&] {
        require(settings != nullptr, "null argument");
        const auto coords = integer_coordinates(xy, n);
        const faco_settings c = normalized_settings(*settings, n);
        const uint32_t K = std::max(c.primary_width, c.ls_width);
        const Prepared p = prepare_problem(coords, c, K, round16_host(2u * n * K) / 2u);
        if (normalized) *normalized = c;
        if (tour) std::copy(p.initial.begin(), p.initial.end(), tour);
        if (cost) *cost = p.initial_cost;
        if (cheap_seconds) *cheap_seconds = p.cheap_seconds;
        if (preparation_seconds) *preparation_seconds = p.preparation_seconds;
        if (primary) std::copy(p.primary.begin(), p.primary.end(), primary);
        if (backup) std::copy(p.backup.begin(), p.backup.end(), backup);
        if (ls) std::copy(p.ls.begin(), p.ls.end(), ls);
        if (local_scale) std::copy(p.local_scale.begin(), p.local_scale.end(), local_scale);
    });
}

int faco_engine_create(uint32_t dimension, uint32_t colonies, const faco_settings* settings,
                       uint32_t constraint_mode, uint32_t population_size, faco_engine** out) {
    return guarded([&] {
        require(out && settings, "null argument");
        require(constraint_mode <= 2, "constraint mode must be 0 (unrestricted), 1 (hard) or 2 (escape)");
        require(constraint_mode == 0 || population_size == 1, "a Hard- or Escape-mode engine holds one population member");
        auto engine = std::make_unique<faco_engine>();
        engine->impl = std::make_unique<Engine>(dimension, colonies, *settings, population_size, false,
                                                constraint_mode != 0);
        Engine& e = *engine->impl;
        if (constraint_mode == 2) {
            // the reference's frozen Escape shape (batch_engine.cu:84-86)
            require(e.config.primary_width <= 16 && e.config.backup_width <= 64 && e.config.ls_width <= 20,
                    "Escape mode takes at most 16 primary, 64 backup and 20 LS slots");
            require(e.n < 32768, "Escape mode takes fewer than 32768 cities");
            e.enable_escape();
        }
        if (constraint_mode != 0) {
            e.hard = true;
            e.graph_stride = 2u * e.n * (e.config.primary_width + 1u);
            e.graph_offsets.allocate(static_cast<size_t>(e.panel) * (e.n + 1));
            e.graph_neighbors.allocate(static_cast<size_t>(e.panel) * e.graph_stride);
            e.graph_words = (e.n + 31u) / 32u;
            e.graph_bits.allocate(static_cast<size_t>(e.panel) * e.n * e.graph_words);
            e.allocated += e.graph_offsets.bytes() + e.graph_neighbors.bytes() + e.graph_bits.bytes();
        }
        *out = engine.release();
    });
}

int faco_engine_create_mixed(uint32_t max_dimension, uint32_t colonies, const faco_settings* settings,
                             uint32_t population_size, faco_engine** out) {
    return guarded([&] {
        require(out && settings, "null argument");
        auto engine = std::make_unique<faco_engine>();
        engine->impl = std::make_unique<Engine>(max_dimension, colonies, *settings, population_size, true);
        *out = engine.release();
    });
}

void faco_engine_destroy(faco_engine* engine) { delete engine; }

int faco_engine_settings(const faco_engine* engine, faco_settings* out) {
    return guarded([&] { require(engine && out, "null argument"); *out = engine->impl->config; });
}

int faco_engine_set_waves(faco_engine* engine, uint32_t waves, uint32_t blocks_per_colony) {
    return guarded([&] {
        require(engine, "null argument");
        engine->impl->set_waves(waves, blocks_per_colony);
        engine->impl->fuse_control = true;   // a forced shape runs the persistent controlled kernel where it can
    });
}

int faco_engine_autotune(faco_engine* engine, const uint64_t* instance_keys, const uint64_t* seeds, uint32_t tasks,
                         uint32_t iterations, faco_tune_entry* table, uint32_t capacity, uint32_t* count) {
    return guarded([&] {
        require(engine && instance_keys && seeds, "null argument");
        require(iterations >= 1 && iterations <= 1000, "autotune iterations must be in 1..1000");
        Engine& e = *engine->impl;
        // An equal evaluation budget for every colony: a whole number of
        // iterations of each colony's ants.
        uint64_t unit = 1;
        for (uint32_t c = 0; c < tasks; ++c) {
            const uint64_t a = e.instance_ants(e.problem(instance_keys[c]).n);
            unit = unit / std::gcd(unit, a) * a;
        }
        const uint64_t limit = unit * iterations;
        std::vector<faco_tune_entry> rows;
        for (uint32_t w : {16u, 12u, 10u, 8u, 6u, 4u, 2u, 1u}) {
            for (uint32_t lds : {1u, 0u}) {
                const uint32_t forced = w | (lds ? 0u : 0x100u);
                try {
                    (void)choose_shape(e.n, e.K, e.config.ants, forced, e.wave_extra);
                } catch (const std::invalid_argument&) {
                    continue;   // does not fit the LDS budget
                }
                e.set_waves(forced, 0);
                faco_result r{};
                evaluate(e, instance_keys, seeds, tasks, limit / e.config.ants, 8, FACO_RUN_EVENTS, &r, nullptr,
                         nullptr, nullptr, limit, true);
                faco_tune_entry row{};
                row.waves = e.shape.waves;
                row.nbr_in_lds = e.shape.nbr_in_lds;
                e.shapes_active = e.mixed;
                if (e.mixed) (void)e.prepare_shapes(instance_keys, tasks, limit);
                row.blocks_per_colony = e.ant_params(tasks).blocks_per_colony;
                e.shapes_active = false;
                row.ms_per_iteration = (r.kernel_milliseconds[0] + r.kernel_milliseconds[1]) /
                                       static_cast<double>(std::max<uint64_t>(r.launched_batches, 1));
                rows.push_back(row);
            }
        }
        require(!rows.empty(), "no launch shape fits the LDS budget");
        size_t best = 0;
        for (size_t i = 1; i < rows.size(); ++i)
            if (rows[i].ms_per_iteration < rows[best].ms_per_iteration) best = i;
        e.set_waves(rows[best].waves | (rows[best].nbr_in_lds ? 0u : 0x100u), 0);
        if (table) std::copy(rows.begin(), rows.begin() + std::min<size_t>(rows.size(), capacity), table);
        if (count) *count = static_cast<uint32_t>(rows.size());
    });
}

int faco_engine_autotune_control(faco_engine* engine, const uint64_t* panel_keys, const uint64_t* panel_seeds,
                                 uint32_t panel, const faco_control_spec* spec, uint32_t iterations,
                                 faco_tune_entry* table, uint32_t capacity, uint32_t* count) {
    return guarded([&] {
        require(engine && panel_keys && panel_seeds && spec, "null argument");
        require(iterations >= 1 && iterations <= 1000, "autotune iterations must be in 1..1000");
        Engine& e = *engine->impl;
        // a whole number of batches of every colony's ants
        uint64_t unit = 1;
        for (uint32_t t = 0; t < panel; ++t) {
            const uint64_t a = e.instance_ants(e.problem(panel_keys[t]).n);
            unit = unit / std::gcd(unit, a) * a;
        }
        const uint64_t limit = unit * iterations;
        const uint32_t deciders = spec->mode == FACO_CONTROL_BASELINE ? 1u : spec->deciders;
        std::vector<faco_tune_entry> rows;
        const uint32_t C = panel * deciders;
        std::vector<uint64_t> inst(C);
        for (uint32_t c = 0; c < C; ++c) inst[c] = panel_keys[c % panel];
        for (uint32_t w : {16u, 12u, 10u, 8u, 6u, 4u, 2u}) {
            for (uint32_t lds : {1u, 0u}) {
                const uint32_t forced = w | (lds ? 0u : 0x100u);
                try {
                    (void)choose_shape(e.n, e.K, e.config.ants, forced, e.wave_extra);
                } catch (const std::invalid_argument&) {
                    continue;   // does not fit the LDS budget
                }
                e.set_waves(forced, 0);
                e.shapes_active = e.mixed;
                if (e.mixed) (void)e.prepare_shapes(inst.data(), C, limit);
                const AntParams ap = e.ant_params(C);   // the launch shape only (no control buffers yet)
                e.shapes_active = false;
                // the persistent controlled kernel where the shape allows it,
                // and the per-batch launches
                for (const bool fuse : {true, false}) {
                    if (fuse && !fused_control(ap)) continue;
                    e.fuse_control = fuse;
                    faco_result r{};
                    evaluate_control(e, panel_keys, panel_seeds, panel, limit, *spec, FACO_RUN_GRAPH, nullptr, &r,
                                     nullptr, nullptr, nullptr, nullptr, nullptr);
                    faco_tune_entry row{};
                    row.waves = e.shape.waves;
                    row.nbr_in_lds = e.shape.nbr_in_lds;
                    row.blocks_per_colony = ap.blocks_per_colony;
                    row.pad = fuse ? 1u : 0u;
                    row.ms_per_iteration = r.search_gpu_milliseconds /
                                           static_cast<double>(std::max<uint64_t>(r.launched_batches, 1));
                    rows.push_back(row);
                }
            }
        }
        require(!rows.empty(), "no launch shape fits the LDS budget");
        size_t best = 0;
        for (size_t i = 1; i < rows.size(); ++i)
            if (rows[i].ms_per_iteration < rows[best].ms_per_iteration) best = i;
        e.set_waves(rows[best].waves | (rows[best].nbr_in_lds ? 0u : 0x100u), 0);
        e.fuse_control = rows[best].pad != 0;
        if (table) std::copy(rows.begin(), rows.begin() + std::min<size_t>(rows.size(), capacity), table);
        if (count) *count = static_cast<uint32_t>(rows.size());
    });
}

int faco_engine_set_roulette_margin(faco_engine* engine, double margin) {
    return guarded([&] {
        require(engine, "null argument");
        require(margin >= kRouletteMargin, "the roulette margin can only be widened (2^-18 or more)");
        engine->impl->roulette_margin = margin;
        engine->impl->drop_graph();
    });
}

int faco_engine_set_local_search(faco_engine* engine, uint32_t mode) {
    return guarded([&] {
        require(engine, "null argument");
        require(mode <= 1, "local search mode must be 0 (2-opt) or 1 (2-opt + Or-opt)");
        require(mode == 0 || !engine->impl->hard, "a Hard-mode engine uses the plain 2-opt");
        engine->impl->local_search = mode;
        engine->impl->forget_run();
        engine->impl->drop_graph();
    });
}

int faco_engine_launch_shape(const faco_engine* engine, uint32_t colonies, uint32_t* waves, uint32_t* nbr_in_lds,
                             uint32_t* lds_bytes, uint32_t* blocks_per_colony, uint32_t* blocks_per_cu) {
    return guarded([&] {
        require(engine, "null argument");
        const auto& e = *engine->impl;
        const AntParams ap = e.ant_params(colonies ? colonies : e.panel);
        if (blocks_per_colony) *blocks_per_colony = ap.blocks_per_colony;
        if (blocks_per_cu) *blocks_per_cu = static_cast<uint32_t>(ant_search_blocks_per_cu(ap));
        if (waves) *waves = e.shape.waves;
        if (nbr_in_lds) *nbr_in_lds = e.shape.nbr_in_lds;
        if (lds_bytes) *lds_bytes = ant_shared_bytes(e.n, e.shape.nbr_in_lds ? e.shape.nbr_lds_bytes : 0u) +
                                    e.shape.waves * (ant_wave_bytes(e.n) + e.wave_extra);
    });
}

int faco_engine_register_problem(faco_engine* engine, uint64_t key, const double* xy, uint32_t n,
                                 double* cheap_seconds, double* preparation_seconds) {
    return guarded([&] {
        require(engine, "null argument");
        Engine& e = *engine->impl;
        std::unique_lock<std::mutex> lock(e.mutex, std::try_to_lock);
        require(lock.owns_lock(), "registration and evaluation of one engine must not overlap");
        require(!e.hard, "a Hard-mode engine registers instances with their graph (register_graph_problem)");
        if (e.mixed) {
            require(n <= e.n, "instance larger than the mixed engine's maximum dimension");
            const faco_settings own = normalized_settings(e.base, n);
            require(own.primary_width == e.config.primary_width && own.backup_width == e.config.backup_width &&
                    own.ls_width == e.config.ls_width,
                    "a mixed-size batch needs the same candidate widths for every instance");
        } else {
            require(n == e.n, "instance dimension does not match the engine");
        }
        auto coords = integer_coordinates(xy, n);
        const auto existing = e.registry.find(key);
        if (existing != e.registry.end()) {
            require(existing->second.xy == coords, "same instance key registered with different coordinates");
        } else {
            e.registry.emplace(key, prepare_problem(coords, e.config, e.K, e.nbr_stride));
        }
        const Prepared& p = e.registry.at(key);
        if (cheap_seconds) *cheap_seconds = p.cheap_seconds;
        if (preparation_seconds) *preparation_seconds = p.preparation_seconds;
    });
}

// register_graph_problem (reference batch_engine.cu:296-314,
// prepared_problem.cpp:128-166): the common preparation (cheap tour, kNN
// rows, initial tour), then the candidate graph and its own rows. The graph
// is undirected, without self-loops or repeated edges, holds the initial
// tour and has at most n (primary_width + 1) edges; the primary and LS rows
// are graph neighbours, LS rows ordered by distance, backup rows
// disjoint from the primary ones; every row is right-padded with n. A Hard
// engine keeps split neighbour rows (primary row, then LS row) for them.
int faco_engine_register_graph_problem(faco_engine* engine, uint64_t key, const double* xy, uint32_t n,
                                       const uint32_t* edges, uint64_t edge_count, const uint32_t* initial_tour,
                                       const uint32_t* primary, const uint32_t* backup, const uint32_t* ls,
                                       double* cheap_seconds, double* preparation_seconds) {
    return guarded([&] {
        require(engine && xy && initial_tour && primary && ls && (edges || !edge_count), "null argument");
        Engine& e = *engine->impl;
        std::unique_lock<std::mutex> lock(e.mutex, std::try_to_lock);
        require(lock.owns_lock(), "registration and evaluation of one engine must not overlap");
        require(e.hard && n == e.n, "graph registration needs a Hard-mode engine of this dimension");
        const uint32_t pw = e.config.primary_width, bw = e.config.backup_width, lw = e.config.ls_width;
        require(bw == 0 || backup, "null argument");
        auto coords = integer_coordinates(xy, n);
        const uint32_t k_prep = std::max(pw, lw);
        Prepared p = prepare_problem(coords, e.config, k_prep, round16_host(2u * n * k_prep) / 2u);
        const auto started = Clock::now();
        require(std::equal(p.initial.begin(), p.initial.end(), initial_tour),
                "the graph spec changed the common initial tour");
        require(edge_count <= static_cast<uint64_t>(n) * (pw + 1u), "the graph exceeds n (primary_width + 1) edges");
        std::vector<std::vector<uint32_t>> adjacency(n);
        for (uint64_t i = 0; i < edge_count; ++i) {
            const uint32_t a = edges[2 * i], b = edges[2 * i + 1];
            require(a < n && b < n && a != b, "graph edge out of range or a self-loop");
            adjacency[a].push_back(b);
            adjacency[b].push_back(a);
        }
        p.graph_offsets.assign(1, 0);
        for (auto& row : adjacency) {
            std::sort(row.begin(), row.end());
            require(std::adjacent_find(row.begin(), row.end()) == row.end(), "graph holds a repeated edge");
            p.graph_neighbors.insert(p.graph_neighbors.end(), row.begin(), row.end());
            p.graph_offsets.push_back(static_cast<uint32_t>(p.graph_neighbors.size()));
        }
        auto member = [&](uint32_t a, uint32_t b) {
            return std::binary_search(p.graph_neighbors.begin() + p.graph_offsets[a],
                                      p.graph_neighbors.begin() + p.graph_offsets[a + 1], b);
        };
        for (uint32_t i = 0; i < n; ++i)
            require(member(p.initial[i], p.initial[(i + 1) % n]), "the graph misses an edge of the initial tour");
        auto d = [&](uint32_t a, uint32_t b) { return euc2d(p.packed[a], p.packed[b]); };
        // fixed-width rows, cities before padding, no repeats or self-loops;
        // primary and LS members in the graph; LS rows by non-decreasing
        // distance (a row ordered by exact distance, then city, stays
        // non-decreasing once distances are rounded, ties in either order)
        auto check_rows = [&](const uint32_t* rows, uint32_t width, bool in_graph, bool ordered) {
            for (uint32_t a = 0; a < n; ++a) {
                const uint32_t* row = rows + static_cast<size_t>(a) * width;
                bool padded = false;
                int64_t last_d = -1;
                for (uint32_t j = 0; j < width; ++j) {
                    const uint32_t b = row[j];
                    if (b == n) { padded = true; continue; }
                    require(!padded && b < n && b != a && std::find(row, row + j, b) == row + j,
                            "graph rows hold an unknown or repeated city, a self-loop or a city after padding");
                    require(!in_graph || member(a, b), "a primary or LS row city is not in the graph");
                    const int64_t db = d(a, b);
                    require(!ordered || last_d <= db, "LS rows must be ordered by distance");
                    last_d = db;
                }
            }
        };
        check_rows(primary, pw, true, false);
        if (bw) check_rows(backup, bw, false, false);
        check_rows(ls, lw, true, true);
        for (uint32_t a = 0; a < n; ++a)
            for (uint32_t j = 0; j < bw; ++j) {
                const uint32_t b = backup[static_cast<size_t>(a) * bw + j];
                require(b == n || std::find(primary + static_cast<size_t>(a) * pw, primary + static_cast<size_t>(a + 1) * pw, b) ==
                                      primary + static_cast<size_t>(a + 1) * pw,
                        "a backup row repeats a primary row city");
            }
        // the spec's rows replace the kNN rows: split neighbour rows, their
        // distances, backup rows and the heuristic of the primary slots
        // (padding slots are never available; their weight is 0)
        p.primary.assign(primary, primary + static_cast<size_t>(n) * pw);
        p.ls.assign(ls, ls + static_cast<size_t>(n) * lw);
        p.backup.assign(static_cast<size_t>(n) * bw, n);
        if (bw) std::copy(backup, backup + static_cast<size_t>(n) * bw, p.backup.begin());
        p.backup16.assign(p.backup.begin(), p.backup.end());
        p.nbr.assign(e.nbr_stride, static_cast<uint16_t>(n));
        p.nbr_dist.assign(e.nbr_stride, 0);
        for (uint32_t a = 0; a < n; ++a) {
            for (uint32_t j = 0; j < pw + lw; ++j) {
                const uint32_t b = j < pw ? primary[static_cast<size_t>(a) * pw + j] : ls[static_cast<size_t>(a) * lw + j - pw];
                p.nbr[static_cast<size_t>(a) * e.K + j] = static_cast<uint16_t>(b);
                p.nbr_dist[static_cast<size_t>(a) * e.K + j] = b < n ? static_cast<uint16_t>(d(a, b)) : 0;
            }
            for (uint32_t j = 0; j < pw; ++j) {
                const uint32_t b = primary[static_cast<size_t>(a) * pw + j];
                const double db = b < n ? static_cast<double>(d(a, b)) : 0.0;
                p.eta[static_cast<size_t>(a) * pw + j] =
                    b >= n ? 0.0 : db > 0 ? (e.config.beta == 1.0 ? 1.0 / db : 1.0 / std::pow(db, e.config.beta)) : 1.0;
            }
        }
        if (e.escape) {
            // the weight of a backup city in a replaced slot is
            // default_trail * d^-beta (reference escape_device.cuh:74-77)
            p.eta_backup.assign(static_cast<size_t>(n) * bw, 0.0);
            for (uint32_t a = 0; a < n; ++a)
                for (uint32_t j = 0; j < bw; ++j) {
                    const uint32_t b = p.backup[static_cast<size_t>(a) * bw + j];
                    const double db = b < n ? static_cast<double>(d(a, b)) : 0.0;
                    p.eta_backup[static_cast<size_t>(a) * bw + j] =
                        b >= n ? 0.0 : db > 0 ? (e.config.beta == 1.0 ? 1.0 / db : 1.0 / std::pow(db, e.config.beta)) : 1.0;
                }
        }
        p.preparation_seconds += seconds_since(started);
        const auto existing = e.registry.find(key);
        if (existing != e.registry.end()) {
            const Prepared& q = existing->second;
            require(q.xy == p.xy && q.graph_offsets == p.graph_offsets && q.graph_neighbors == p.graph_neighbors &&
                    q.primary == p.primary && q.backup == p.backup && q.ls == p.ls,
                    "same instance key registered with different coordinates, graph or rows");
        } else {
            e.registry.emplace(key, std::move(p));
        }
        const Prepared& q = e.registry.at(key);
        if (cheap_seconds) *cheap_seconds = q.cheap_seconds;
        if (preparation_seconds) *preparation_seconds = q.preparation_seconds;
    });
}

int faco_engine_register_problems_device(faco_engine* engine, const uint64_t* keys, const double* xy,
                                         const uint32_t* sizes, uint32_t count, double* seconds) {
    return guarded([&] {
        require(engine && keys && xy, "null argument");
        const double total = register_problems_device(*engine->impl, keys, xy, sizes, count);
        if (seconds) *seconds = total;
    });
}

int faco_engine_edge_features(faco_engine* engine, uint64_t key, float* features) {
    return guarded([&] {
        require(engine && features, "null argument");
        Engine& e = *engine->impl;
        const Prepared& p = e.problem(key);
        const uint32_t n = p.n, pw = e.config.primary_width;
        require(std::find(p.primary.begin(), p.primary.end(), n) == p.primary.end(),
                "edge features need primary rows without padding");
        double mean = 0;
        for (uint32_t a = 0; a < n; ++a) mean += p.local_scale[a];
        const float mean_scale = static_cast<float>(std::max(mean / n, p.scale_epsilon));
        DeviceArray<uint32_t> coords, primary;
        DeviceArray<double> scale;
        DeviceArray<float> out;
        coords.allocate(n); primary.allocate(static_cast<size_t>(n) * pw); scale.allocate(n);
        out.allocate(static_cast<size_t>(n) * pw * 8);
        coords.upload(p.packed.data(), n, e.stream);
        primary.upload(p.primary.data(), static_cast<size_t>(n) * pw, e.stream);
        scale.upload(p.local_scale.data(), n, e.stream);
        checked(launch_prior_features(n, pw, coords.data(), primary.data(), scale.data(), p.scale_epsilon, mean_scale,
                                      out.data(), e.stream), "feature launch");
        out.download(features, out.size(), e.stream);
        checked(hipStreamSynchronize(e.stream), "edge features");
    });
}

int faco_engine_set_edge_prior(faco_engine* engine, uint64_t key, const float* multipliers) {
    return guarded([&] {
        require(engine, "null argument");
        Engine& e = *engine->impl;
        std::unique_lock<std::mutex> lock(e.mutex, std::try_to_lock);
        require(lock.owns_lock(), "the prior of an engine can not change during an evaluation");
        auto it = e.registry.find(key);
        require(it != e.registry.end(), "task instance has not been registered");
        const size_t cells = static_cast<size_t>(it->second.n) * e.config.primary_width;
        require(!multipliers || !e.escape, "an Escape-mode engine takes no learned prior");
        if (multipliers) {
            for (size_t i = 0; i < cells; ++i)
                require(std::isfinite(multipliers[i]) && multipliers[i] > 0.0f, "prior multipliers must be positive");
            it->second.prior.assign(multipliers, multipliers + cells);
        } else {
            it->second.prior.clear();
        }
        e.resident.clear();
        e.forget_run();
    });
}

int faco_engine_apply_edge_prior(faco_engine* engine, faco_prior* prior, uint64_t key, float* multipliers,
                                 double exponent) {
    return guarded([&] {
        require(engine && prior, "null argument");
        require(std::isfinite(exponent) && exponent > 0, "prior exponent must be positive and finite");
        Engine& e = *engine->impl;
        const size_t cells = static_cast<size_t>(e.problem(key).n) * e.config.primary_width;
        std::vector<float> features(cells * 8), logits(cells), mult(cells);
        if (faco_engine_edge_features(engine, key, features.data()) != FACO_OK) throw std::runtime_error(g_error);
        DeviceArray<float> f, s;
        f.allocate(cells * 8); s.allocate(cells);
        f.upload(features.data(), f.size(), e.stream);
        checked(launch_prior_infer(prior_device_image(prior), f.data(), static_cast<uint32_t>(cells), s.data(),
                                   e.stream), "prior inference launch");
        s.download(logits.data(), cells, e.stream);
        checked(hipStreamSynchronize(e.stream), "prior inference");
        for (size_t i = 0; i < cells; ++i) {
            const double sigmoid = 1.0 / (1.0 + std::exp(-static_cast<double>(logits[i])));
            mult[i] = static_cast<float>(exponent == 1.0 ? sigmoid : (exponent == 2.0 ? sigmoid * sigmoid
                                                                                        : std::pow(sigmoid, exponent)));
        }
        for (size_t i = 0; i < cells; ++i) mult[i] = mult[i] > 0.0f ? mult[i] : 1e-30f;
        if (faco_engine_set_edge_prior(engine, key, mult.data()) != FACO_OK) throw std::runtime_error(g_error);
        if (multipliers) std::copy(mult.begin(), mult.end(), multipliers);
    });
}

int faco_engine_preparation_profile(faco_engine* engine, uint64_t key, double out[4]) {
    return guarded([&] {
        require(engine && out, "null argument");
        const Prepared& p = engine->impl->problem(key);
        std::copy(p.profile, p.profile + 4, out);
    });
}

int faco_engine_problem_data(faco_engine* engine, uint64_t key, uint32_t* primary, uint32_t* backup,
                             uint32_t* ls, uint32_t* initial_tour, int64_t* initial_cost) {
    return guarded([&] {
        require(engine, "null argument");
        const Prepared& p = engine->impl->problem(key);
        if (primary) std::copy(p.primary.begin(), p.primary.end(), primary);
        if (backup) std::copy(p.backup.begin(), p.backup.end(), backup);
        if (ls) std::copy(p.ls.begin(), p.ls.end(), ls);
        if (initial_tour) std::copy(p.initial.begin(), p.initial.end(), initial_tour);
        if (initial_cost) *initial_cost = p.initial_cost;
    });
}

int faco_engine_evaluate_faco(faco_engine* engine, const uint64_t* instance_keys, const uint64_t* seeds,
                              uint32_t tasks, uint64_t limit, uint32_t mne, uint32_t flags,
                              faco_result* result, uint32_t* tours, int64_t* costs, uint64_t* work,
                              const faco_checkpoints* checkpoints) {
    return guarded([&] {
        require(engine && instance_keys && seeds, "null argument");
        Engine& e = *engine->impl;
        require(e.population == 1, "population engines evaluate through the population entry point");
        if (!e.mixed)
            require(limit % e.config.ants == 0 && limit / e.config.ants <= UINT32_MAX,
                    "evaluation limit must be a whole number of ant batches");
        evaluate(e, instance_keys, seeds, tasks, limit / e.config.ants, mne, flags, result, tours,
                 costs, work, limit, false, checkpoints);
    });
}

int faco_engine_evaluate_deadline(faco_engine* engine, const uint64_t* instance_keys, const uint64_t* seeds,
                                  uint32_t tasks, const faco_deadline_spec* spec, faco_deadline_result* result,
                                  uint32_t* tours, double* costs, uint8_t* has_incumbent, double* completed_seconds,
                                  faco_control_state* states, double* ratios, uint32_t ratio_capacity,
                                  faco_batch_profile* profile) {
    return guarded([&] {
        require(engine && instance_keys && seeds && spec, "null argument");
        evaluate_deadline(*engine->impl, instance_keys, seeds, tasks, *spec, result, tours, costs, has_incumbent,
                          completed_seconds, states, ratios, ratio_capacity, profile);
    });
}

int faco_engine_constraint_rejections(faco_engine* engine, uint64_t* out, uint32_t count) {
    return guarded([&] {
        require(engine && out, "null argument");
        Engine& e = *engine->impl;
        std::unique_lock<std::mutex> lock(e.mutex, std::try_to_lock);
        require(lock.owns_lock(), "rejections can not be read during an evaluation");
        require(count == e.last_rejections.size(), "one entry per task of the last count evaluation");
        std::copy(e.last_rejections.begin(), e.last_rejections.end(), out);
    });
}

int faco_engine_escape_counters(faco_engine* engine, uint64_t* out, uint32_t count, uint64_t* reserved_bytes) {
    return guarded([&] {
        require(engine && out, "null argument");
        Engine& e = *engine->impl;
        std::unique_lock<std::mutex> lock(e.mutex, std::try_to_lock);
        require(lock.owns_lock(), "escape counters can not be read during an evaluation");
        require(e.escape, "only an Escape-mode engine counts escapes");
        require(static_cast<size_t>(count) * FACO_ESCAPE_COUNTERS == e.last_escape.size(),
                "one entry per task of the last count evaluation");
        std::copy(e.last_escape.begin(), e.last_escape.end(), out);
        if (reserved_bytes) *reserved_bytes = e.escape_bytes;
    });
}

int faco_engine_set_preparation_charges(faco_engine* engine, uint64_t key, double cheap, double preparation) {
    return guarded([&] {
        require(engine, "null argument");
        Engine& e = *engine->impl;
        std::unique_lock<std::mutex> lock(e.mutex, std::try_to_lock);
        require(lock.owns_lock(), "charges of an engine can not change during an evaluation");
        (void)e.problem(key);
        require(std::isfinite(cheap) && cheap >= 0 && std::isfinite(preparation) && preparation >= 0,
                "frozen charges must be finite non-negative numbers");
        const auto [entry, inserted] = e.charges.emplace(key, std::make_pair(cheap, preparation));
        require(inserted || entry->second == std::make_pair(cheap, preparation),
                "frozen preparation charges of one instance can not change");
    });
}

int faco_engine_evaluate_control(faco_engine* engine, const uint64_t* instance_keys, const uint64_t* seeds,
                                 uint32_t panel, uint64_t limit, const faco_control_spec* spec, uint32_t flags,
                                 const faco_control_trace* trace, faco_result* result, uint32_t* tours,
                                 int64_t* costs, uint64_t* work, faco_control_state* states,
                                 const faco_checkpoints* checkpoints) {
    return guarded([&] {
        require(engine && instance_keys && seeds && spec, "null argument");
        evaluate_control(*engine->impl, instance_keys, seeds, panel, limit, *spec, flags, trace, result, tours, costs,
                         work, states, checkpoints);
    });
}

// ---------------------------------------------------------------------------
// CountedState snapshots (reference cpp/src/bindings.cpp:541-556, 880-934).
int faco_state_from_bytes(const uint8_t* bytes, size_t size, faco_state** out) {
    return guarded([&] {
        require(out != nullptr, "null argument");
        auto state = std::make_unique<faco_state>();
        state->s = CountedState::deserialize(bytes, size);
        *out = state.release();
    });
}

int faco_state_to_bytes(const faco_state* state, uint8_t* out, size_t capacity, size_t* size) {
    return guarded([&] {
        require(state != nullptr, "null argument");
        state->s.validate();
        if (!out) { if (size) *size = state->s.serialized_size(); return; }
        copy_out(state->s.serialize(), out, capacity, size);
    });
}

int faco_state_describe(const faco_state* state, char* out, size_t capacity, size_t* size) {
    return guarded([&] {
        require(state != nullptr, "null argument");
        const std::string text = describe_state(state->s);
        if (size) *size = text.size() + 1;
        if (!out) return;
        require(capacity > text.size(), "output buffer too small");
        std::memcpy(out, text.c_str(), text.size() + 1);
    });
}

int faco_state_select_colony(const faco_state* state, uint32_t colony, faco_state** out) {
    return guarded([&] {
        require(state != nullptr && out != nullptr, "null argument");
        auto selected = std::make_unique<faco_state>();
        selected->s = state->s.select_colony(colony);
        *out = selected.release();
    });
}

void faco_state_destroy(faco_state* state) { delete state; }

int faco_engine_capture_control(faco_engine* engine, const uint64_t* instance_keys, const uint64_t* seeds,
                                uint32_t tasks, uint64_t source_limit, uint64_t capture_after,
                                const faco_control_spec* spec, uint32_t flags, const faco_control_trace* trace,
                                faco_result* result, uint32_t* tours, int64_t* costs, uint64_t* work,
                                faco_control_state* states, faco_state** out) {
    return guarded([&] {
        require(engine && instance_keys && seeds && spec && out, "null argument");
        auto state = std::make_unique<faco_state>();
        StateIO io;
        io.capture = &state->s;
        io.capture_after = capture_after;
        evaluate_control(*engine->impl, instance_keys, seeds, tasks, source_limit, *spec, flags, trace, result, tours,
                         costs, work, states, nullptr, &io);
        *out = state.release();
    });
}

int faco_engine_continue_control(faco_engine* engine, const faco_state* state, uint64_t additional,
                                 const faco_control_spec* spec, const faco_fork* fork, uint32_t flags,
                                 const faco_control_trace* trace, faco_result* result, uint32_t* tours,
                                 int64_t* costs, uint64_t* work, faco_control_state* states) {
    return guarded([&] {
        require(engine && state && spec, "null argument");
        const CountedState& saved = state->s;
        saved.validate();
        std::vector<uint64_t> keys, seeds;
        for (const auto& t : saved.tasks) { keys.push_back(t.first); seeds.push_back(t.second); }
        faco_control_spec resumed = *spec;
        resumed.experiment_mask = saved.experiment_mask;   // the snapshot's mask holds for the continuation
        StateIO io;
        io.resume = &saved;
        if (fork) io.fork = *fork;
        evaluate_control(*engine->impl, keys.data(), seeds.data(), saved.colonies, additional, resumed, flags, trace,
                         result, tours, costs, work, states, nullptr, &io);
    });
}

// ---------------------------------------------------------------------------
int faco_fixed_create(const double* xy, uint32_t n, uint64_t key, const faco_settings* settings, faco_fixed** out) {
    return guarded([&] {
        require(out && settings, "null argument");
        auto fixed = std::make_unique<faco_fixed>();
        fixed->impl = std::make_unique<Engine>(n, 1u, *settings, 1u);
        fixed->key = key;
        Engine& e = *fixed->impl;
        auto coords = integer_coordinates(xy, n);
        e.registry.emplace(key, prepare_problem(coords, e.config, e.K, e.nbr_stride));
        *out = fixed.release();
    });
}

void faco_fixed_destroy(faco_fixed* fixed) { delete fixed; }

int faco_fixed_set_local_search(faco_fixed* fixed, uint32_t mode) {
    return guarded([&] {
        require(fixed, "null argument");
        require(mode <= 1, "local search mode must be 0 (2-opt) or 1 (2-opt + Or-opt)");
        fixed->impl->local_search = mode;
        fixed->impl->drop_graph();
    });
}

int faco_fixed_settings(const faco_fixed* fixed, faco_settings* out) {
    return guarded([&] { require(fixed && out, "null argument"); *out = fixed->impl->config; });
}

int faco_fixed_run(faco_fixed* fixed, uint64_t seed, uint32_t batches, uint32_t mne, uint32_t* tour, int64_t* cost,
                   int64_t* initial_cost, uint64_t* construction_steps, uint64_t* ls_evaluations,
                   double* solve_seconds, uint32_t* trace_parent, double* trace_trails, int64_t* trace_costs,
                   uint32_t* trace_best, faco_ant_info* trace_ants, uint32_t* trace_selected,
                   double* trace_uniforms, uint32_t* trace_tours) {
    return guarded([&] {
        require(fixed, "null argument");
        Engine& e = *fixed->impl;
        require(mne > 0, "mne_target must be positive");
        const auto started = Clock::now();
        const bool trace = trace_parent || trace_trails || trace_costs || trace_best || trace_ants ||
                           trace_selected || trace_uniforms || trace_tours;
        const uint64_t keyv = fixed->key;
        if (!trace) {
            uint64_t work[2] = {0, 0};
            int64_t c = 0;
            evaluate(e, &keyv, &seed, 1, batches, mne, FACO_RUN_GRAPH, nullptr, tour, &c, work,
                     static_cast<uint64_t>(batches) * e.config.ants);
            if (cost) *cost = c;
            if (construction_steps) *construction_steps = work[0];
            if (ls_evaluations) *ls_evaluations = work[1];
        } else {
            std::unique_lock<std::mutex> lock(e.mutex, std::try_to_lock);
            require(lock.owns_lock(), "one evaluation per engine at a time");
            e.forget_run();
            const uint32_t n = e.n, ants = e.config.ants, pw = e.config.primary_width;
            e.upload_geometry(&keyv, 1);
            const uint64_t key = colony_key(keyv, seed);
            e.keys.upload(&key, 1, e.stream);
            e.reset(1, mne);
            DeviceArray<uint32_t> selected;
            DeviceArray<double> uniforms;
            selected.allocate(static_cast<size_t>(ants) * n);
            uniforms.allocate(static_cast<size_t>(ants) * n);
            AntParams ap = e.ant_params(1);
            ap.trace_selected = selected.data();
            ap.trace_uniform = uniforms.data();
            const ColonyParams cp = e.colony_params();
            for (uint32_t b = 0; b < batches; ++b) {
                e.iteration(1, ap, cp);
                ColonyState st;
                e.state.download(&st, 1, e.stream);
                if (trace_parent) e.parent.download(trace_parent + static_cast<size_t>(b) * n, n, e.stream);
                if (trace_trails) e.trails.download(trace_trails + static_cast<size_t>(b) * n * pw, static_cast<size_t>(n) * pw, e.stream);
                if (trace_ants) e.info.download(trace_ants + static_cast<size_t>(b) * ants, ants, e.stream);
                if (trace_selected) selected.download(trace_selected + static_cast<size_t>(b) * ants * n, static_cast<size_t>(ants) * n, e.stream);
                if (trace_uniforms) uniforms.download(trace_uniforms + static_cast<size_t>(b) * ants * n, static_cast<size_t>(ants) * n, e.stream);
                if (trace_tours) e.tours.download(trace_tours + static_cast<size_t>(b) * ants * n, static_cast<size_t>(ants) * n, e.stream);
                checked(hipStreamSynchronize(e.stream), "trace batch");
                if (trace_costs) {
                    trace_costs[3 * b] = st.epoch_cost; trace_costs[3 * b + 1] = st.global_cost;
                    trace_costs[3 * b + 2] = st.parent_cost;
                }
                if (trace_best) trace_best[b] = st.iteration_best;
            }
            ColonyState st;
            e.state.download(&st, 1, e.stream);
            if (tour) e.global.download(tour, n, e.stream);
            checked(hipStreamSynchronize(e.stream), "readback");
            if (cost) *cost = st.global_cost;
            if (construction_steps) *construction_steps = st.work_construction;
            if (ls_evaluations) *ls_evaluations = st.work_ls;
        }
        if (initial_cost) *initial_cost = e.problem(keyv).initial_cost;
        if (solve_seconds) *solve_seconds = seconds_since(started);
    });
}

// ---------------------------------------------------------------------------
int faco_diagnostic_construct_ls(const double* xy, uint32_t n, const uint32_t* ls, uint32_t lw, uint32_t ants,
                                 const uint32_t* parents, const uint32_t* visits, const uint32_t* mne_targets,
                                 uint64_t limit, uint32_t waves, uint32_t chunk, uint32_t* construction_tours, uint32_t* tours,
                                 uint32_t* positions, uint32_t* checklists, faco_ant_info* info, uint32_t local_search) {
    return guarded([&] {
        require(ls && parents && visits && mne_targets && tours && info, "null argument");
        require(ants >= 1 && lw >= 1 && lw <= kMaxLs && lw <= n - 1, "invalid ant count or LS width");
        const auto coords = integer_coordinates(xy, n);
        for (uint32_t a = 0; a < ants; ++a) {
            require(is_permutation(parents + static_cast<size_t>(a) * n, n), "parent tour is not a permutation");
            require(mne_targets[a] > 0, "mne_target must be positive");
        }
        for (size_t i = 0; i < static_cast<size_t>(n) * lw; ++i) require(ls[i] < n, "LS candidate out of range");
        if (waves == 0) waves = 1;
        if (chunk == 0) chunk = waves;
        require(waves <= 16 && chunk >= 1 && ants % chunk == 0, "ants per workgroup must divide the ant count");
        hipStream_t stream;
        checked(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "stream");
        struct StreamGuard { hipStream_t s; ~StreamGuard() { (void)hipStreamDestroy(s); } } guard{stream};
        const uint32_t K = lw, stride = round16_host(2u * n * K) / 2u;
        std::vector<uint32_t> packed(n);
        for (uint32_t i = 0; i < n; ++i) packed[i] = pack_xy(coords[2 * i], coords[2 * i + 1]);
        std::vector<uint16_t> nbr(stride, 0), nbr_dist(stride, 0);
        for (size_t i = 0; i < static_cast<size_t>(n) * lw; ++i) {
            nbr[i] = static_cast<uint16_t>(ls[i]);
            nbr_dist[i] = static_cast<uint16_t>(euc2d(packed[i / lw], packed[ls[i]]));
        }
        std::vector<uint32_t> links(static_cast<size_t>(ants) * n);
        std::vector<int64_t> parent_costs(ants);
        for (uint32_t a = 0; a < ants; ++a) {
            const uint32_t* t = parents + static_cast<size_t>(a) * n;
            int64_t c = 0;
            for (uint32_t i = 0; i < n; ++i) {
                const uint32_t prev = t[i == 0 ? n - 1 : i - 1], next = t[i + 1 == n ? 0 : i + 1];
                links[static_cast<size_t>(a) * n + t[i]] = prev | (next << 16);
                c += euc2d(packed[prev], packed[t[i]]);
            }
            parent_costs[a] = c;
        }
        const size_t cells = static_cast<size_t>(ants) * n;
        DeviceArray<uint32_t> d_coords, d_parents, d_links, d_visits, d_targets, d_tours, d_ctours, d_pos, d_check;
        DeviceArray<uint16_t> d_nbr, d_nbr_dist;
        DeviceArray<int64_t> d_costs;
        DeviceArray<faco_ant_info> d_info;
        d_coords.allocate(n); d_nbr.allocate(stride); d_nbr_dist.allocate(stride); d_parents.allocate(cells); d_links.allocate(cells);
        d_visits.allocate(cells); d_targets.allocate(ants); d_tours.allocate(cells); d_ctours.allocate(cells);
        d_pos.allocate(cells); d_check.allocate(cells * 5); d_costs.allocate(ants); d_info.allocate(ants);
        d_coords.upload(packed.data(), n, stream); d_nbr.upload(nbr.data(), stride, stream);
        d_nbr_dist.upload(nbr_dist.data(), stride, stream);
        d_parents.upload(parents, cells, stream); d_links.upload(links.data(), cells, stream);
        d_visits.upload(visits, cells, stream); d_targets.upload(mne_targets, ants, stream);
        d_costs.upload(parent_costs.data(), ants, stream);
        AntParams p{};
        p.nbr_in_lds = 1;
        p.n = n; p.ants = ants; p.pw = 0; p.bw = 0; p.lw = lw; p.K = K; p.waves = waves;
        p.chunk = chunk; p.blocks_per_colony = 1;
        p.nbr_lds_bytes = round16_host(2u * n * K); p.nbr_stride = stride; p.per_wave_bytes = ant_wave_bytes(n);
        p.xcd_remap = 0; p.explicit_mode = 1; p.per_ant_parent = 1; p.log_checklist = 1;
        p.roulette_margin = kRouletteMargin;
        require(local_search <= 1, "local search mode must be 0 or 1");
        p.or_opt = local_search;
        p.ls_limit = std::min<uint64_t>(limit, UINT32_MAX);
        p.coords = d_coords.data(); p.nbr = d_nbr.data(); p.nbr_dist = d_nbr_dist.data();
        p.parent = d_parents.data(); p.links = d_links.data();
        p.parent_cost = d_costs.data(); p.targets = d_targets.data(); p.visits = d_visits.data();
        p.out_tours = d_tours.data(); p.out_info = d_info.data(); p.diag_construction = d_ctours.data();
        p.diag_positions = d_pos.data(); p.diag_checklist = d_check.data();
        const size_t lds = ant_shared_bytes(n, p.nbr_lds_bytes) + static_cast<size_t>(waves) * p.per_wave_bytes;
        if (lds > max_lds_per_block()) {
            p.nbr_in_lds = 0;
            require(ant_shared_bytes(n, 0) + static_cast<size_t>(waves) * p.per_wave_bytes <= max_lds_per_block(),
                    "instance too large for the LDS budget");
        }
        checked(launch_ant_search(p, ants / chunk, stream), "diagnostic launch");
        d_tours.download(tours, cells, stream);
        if (construction_tours) d_ctours.download(construction_tours, cells, stream);
        if (positions) d_pos.download(positions, cells, stream);
        if (checklists) d_check.download(checklists, cells * 5, stream);
        d_info.download(info, ants, stream);
        checked(hipStreamSynchronize(stream), "diagnostic run");
    });
}

int faco_diagnostic_distances(const int32_t* pairs, uint32_t count, int64_t* out) {
    return guarded([&] {
        require(pairs && out, "null argument");
        for (size_t i = 0; i < 4ull * count; ++i)
            require(pairs[i] >= 0 && pairs[i] <= static_cast<int32_t>(kMaxCoordinate), "coordinate out of range");
        DeviceArray<int32_t> d_pairs; DeviceArray<int64_t> d_out;
        d_pairs.allocate(4ull * count); d_out.allocate(count);
        d_pairs.upload(pairs, 4ull * count, nullptr);
        checked(launch_distance_check(d_pairs.data(), count, d_out.data(), nullptr), "distance launch");
        d_out.download(out, count, nullptr);
        checked(hipDeviceSynchronize(), "distance run");
    });
}

int faco_diagnostic_start_nodes(uint64_t key, uint32_t batch, uint32_t ant, const uint32_t* counts, uint32_t n,
                                uint32_t on_device, uint32_t* out) {
    return guarded([&] {
        require(counts && out, "null argument");
        for (uint32_t i = 0; i < n; ++i) require(counts[i] >= 1, "count must be positive");
        if (!on_device) {
            for (uint32_t i = 0; i < n; ++i) out[i] = start_node(key, batch, ant, counts[i]);
            return;
        }
        DeviceArray<uint32_t> d_counts, d_out;
        d_counts.allocate(n); d_out.allocate(n);
        d_counts.upload(counts, n, nullptr);
        checked(launch_start_check(key, batch, ant, d_counts.data(), n, d_out.data(), nullptr), "start launch");
        d_out.download(out, n, nullptr);
        checked(hipDeviceSynchronize(), "start run");
    });
}

int faco_diagnostic_copy(uint64_t bytes, uint32_t repeats, double* milliseconds) {
    return guarded([&] {
        require(bytes >= 16 && bytes % 16 == 0 && bytes <= (1ull << 34) && repeats >= 1, "invalid copy size");
        DeviceArray<uint8_t> a, b;
        a.allocate(bytes); b.allocate(bytes);
        checked(hipMemset(a.data(), 1, bytes), "memset");
        hipEvent_t e0, e1;
        checked(hipEventCreate(&e0), "event"); checked(hipEventCreate(&e1), "event");
        checked(hipEventRecord(e0, nullptr), "event");
        for (uint32_t r = 0; r < repeats; ++r) checked(launch_copy(a.data(), b.data(), bytes, nullptr), "copy launch");
        checked(hipEventRecord(e1, nullptr), "event");
        checked(hipEventSynchronize(e1), "copy");
        float ms = 0;
        checked(hipEventElapsedTime(&ms, e0, e1), "event time");
        (void)hipEventDestroy(e0); (void)hipEventDestroy(e1);
        if (milliseconds) *milliseconds = ms / repeats;
    });
}

int faco_diagnostic_philox(uint64_t key, uint32_t batch, uint32_t ant, uint32_t step, uint32_t words, uint32_t* out) {
    return guarded([&] {
        require(out && words >= 1 && words <= 4096, "invalid word count");
        DeviceArray<uint32_t> d_out; d_out.allocate(words);
        checked(launch_philox_check(key, batch, ant, step, words, d_out.data(), nullptr), "philox launch");
        d_out.download(out, words, nullptr);
        checked(hipDeviceSynchronize(), "philox run");
    });
}

}  // extern "C"
