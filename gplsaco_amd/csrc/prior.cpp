// Host side of the learned edge prior: the faco_prior object (FP32 master
// weights and Adam moments on the device), inference, full-batch gradients
// and the minibatch training loop. Kernels: prior_kernels.hip.
#include "faco_common.h"
#include "prior.h"

#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

namespace faco {
void set_last_error(const char* message);

namespace {

void require(bool condition, const char* message) {
    if (!condition) throw std::invalid_argument(message);
}
void checked(hipError_t status, const char* what) {
    if (status != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(status));
}

template<class F> int guarded(F&& f) {
    try { f(); return FACO_OK; }
    catch (const std::invalid_argument& e) { set_last_error(e.what()); return FACO_INVALID_ARGUMENT; }
    catch (const std::exception& e) { set_last_error(e.what()); return FACO_RUNTIME_ERROR; }
    catch (...) { set_last_error("unknown native error"); return FACO_RUNTIME_ERROR; }
}

template<class T> struct Buffer {
    T* ptr = nullptr;
    size_t count = 0;
    Buffer() = default;
    explicit Buffer(size_t n) { allocate(n); }
    Buffer(const Buffer&) = delete;
    Buffer& operator=(const Buffer&) = delete;
    ~Buffer() { if (ptr) (void)hipFree(ptr); }
    void allocate(size_t n) {
        count = n;
        if (n) checked(hipMalloc(reinterpret_cast<void**>(&ptr), n * sizeof(T)), "hipMalloc");
    }
};

// Pinned host memory (asynchronous uploads read it while the host works on).
template<class T> struct PinnedBuffer {
    T* ptr = nullptr;
    explicit PinnedBuffer(size_t n) {
        if (n) checked(hipHostMalloc(reinterpret_cast<void**>(&ptr), n * sizeof(T), hipHostMallocDefault), "hipHostMalloc");
    }
    PinnedBuffer(const PinnedBuffer&) = delete;
    PinnedBuffer& operator=(const PinnedBuffer&) = delete;
    ~PinnedBuffer() { if (ptr) (void)hipHostFree(ptr); }
};

// A worker thread joined on every exit path.
struct Worker {
    std::thread thread;
    void join() { if (thread.joinable()) thread.join(); }
    ~Worker() { join(); }
};

// SplitMix64 stream for initial weights and epoch shuffles.
struct Mix {
    uint64_t state;
    uint64_t next() { return mixed(state++); }
    double uniform() { return static_cast<double>(next() >> 11) * 0x1.0p-53; }
};

}  // namespace

struct PriorImpl {
    Buffer<float> params, m, v, grad, loss;
    Buffer<uint8_t> image;   // the weights in the kernels' layout, rebuilt with every parameter change
    uint32_t step = 0;
    hipStream_t stream = nullptr;
    PriorImpl() : params(prior_param_count()), m(prior_param_count()), v(prior_param_count()),
                  grad(prior_param_count() + 1), loss(1), image(prior_image_bytes()) {
        checked(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
    }
    ~PriorImpl() { if (stream) (void)hipStreamDestroy(stream); }
    void set(const float* host) {
        const size_t P = prior_param_count();
        for (size_t i = 0; i < P; ++i) require(std::isfinite(host[i]), "prior parameters must be finite");
        checked(hipMemcpyAsync(params.ptr, host, P * sizeof(float), hipMemcpyHostToDevice, stream), "upload");
        checked(hipMemsetAsync(m.ptr, 0, P * sizeof(float), stream), "memset");
        checked(hipMemsetAsync(v.ptr, 0, P * sizeof(float), stream), "memset");
        checked(launch_prior_image(params.ptr, image.ptr, stream), "weight image launch");
        checked(hipStreamSynchronize(stream), "prior set");
        step = 0;
    }
};

const void* prior_device_image(const faco_prior* prior) { return reinterpret_cast<const PriorImpl*>(prior)->image.ptr; }

// Scratch of faco_prior_gradients_noalloc: one minibatch of at most
// FACO_PRIOR_NOALLOC_EDGES edges (the training loop's batch size). Layout in
// floats, every region 256-byte aligned: the per-wave gradient partials
// [rows][P + 1], the reduced gradient and loss [P + 1], the staged features
// [edges][8] and the staged labels (bytes). Each GPU the library runs on
// holds its own copy until the process ends.
namespace noalloc {
constexpr uint32_t kEdges = FACO_PRIOR_NOALLOC_EDGES;
constexpr uint32_t kTileEdges = 32, kWaves = 2;   // prior_grad_rows: one row per wave of 2-wave blocks
constexpr uint32_t kRows = ((kEdges + kTileEdges - 1) / kTileEdges + kWaves - 1) / kWaves * kWaves;
constexpr uint32_t kRow = 4801 + 1;
constexpr size_t align64(size_t floats) { return (floats + 63) / 64 * 64; }
constexpr size_t kPartial = 0;
constexpr size_t kGrad = kPartial + align64(static_cast<size_t>(kRows) * kRow);
constexpr size_t kFeatures = kGrad + align64(kRow);
constexpr size_t kLabels = kFeatures + align64(static_cast<size_t>(kEdges) * 8);
constexpr size_t kFloats = kLabels + align64(kEdges / 4);
constexpr int kMaxDevices = 64;
}  // namespace noalloc

static __device__ __align__(256) float scratch[noalloc::kFloats];

namespace {
// Device address of `scratch` per device, looked up once with that device
// current. In host code the array's name is not a device address: only
// these pointers are passed to kernels and copies.
std::mutex g_scratch_mutex;
float* g_scratch[noalloc::kMaxDevices] = {};

float* device_scratch() {
    int device = 0;
    checked(hipGetDevice(&device), "hipGetDevice");
    require(device >= 0 && device < noalloc::kMaxDevices, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch[device]) {
        void* address = nullptr;
        checked(hipGetSymbolAddress(&address, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        g_scratch[device] = static_cast<float*>(address);
    }
    return g_scratch[device];
}
}  // namespace

}  // namespace faco

using faco::PriorImpl;

extern "C" {

int faco_prior_create(uint64_t seed, faco_prior** out) {
    return faco::guarded([&] {
        faco::require(out != nullptr, "null argument");
        auto impl = std::make_unique<PriorImpl>();
        // Glorot-uniform weights from a SplitMix64 stream, zero biases.
        const uint32_t P = faco::prior_param_count();
        std::vector<float> w(P, 0.0f);
        faco::Mix mix{faco::mixed(seed ^ 0x5851f42d4c957f2dULL)};
        auto fill = [&](uint32_t offset, uint32_t count, double fan_in, double fan_out) {
            const double bound = std::sqrt(6.0 / (fan_in + fan_out));
            for (uint32_t i = 0; i < count; ++i) w[offset + i] = static_cast<float>((2.0 * mix.uniform() - 1.0) * bound);
        };
        fill(0, 512, 8, 64);       // W1
        fill(576, 4096, 64, 64);   // W2
        fill(4736, 64, 64, 1);     // w3
        impl->set(w.data());
        *out = reinterpret_cast<faco_prior*>(impl.release());
    });
}

void faco_prior_destroy(faco_prior* prior) { delete reinterpret_cast<PriorImpl*>(prior); }

uint32_t faco_prior_param_count(void) { return faco::prior_param_count(); }

int faco_prior_get_params(const faco_prior* prior, float* out) {
    return faco::guarded([&] {
        faco::require(prior && out, "null argument");
        const auto* p = reinterpret_cast<const PriorImpl*>(prior);
        faco::checked(hipMemcpy(out, p->params.ptr, faco::prior_param_count() * sizeof(float), hipMemcpyDeviceToHost),
                      "download");
    });
}

int faco_prior_set_params(faco_prior* prior, const float* params) {
    return faco::guarded([&] {
        faco::require(prior && params, "null argument");
        reinterpret_cast<PriorImpl*>(prior)->set(params);
    });
}

int faco_prior_logits(faco_prior* prior, const float* features, uint32_t edges, float* logits) {
    return faco::guarded([&] {
        faco::require(prior && features && logits && edges, "null argument or no edges");
        auto* p = reinterpret_cast<PriorImpl*>(prior);
        faco::Buffer<float> f(static_cast<size_t>(edges) * 8), s(edges);
        faco::checked(hipMemcpyAsync(f.ptr, features, f.count * sizeof(float), hipMemcpyHostToDevice, p->stream), "upload");
        faco::checked(faco::launch_prior_infer(p->image.ptr, f.ptr, edges, s.ptr, p->stream), "prior inference launch");
        faco::checked(hipMemcpyAsync(logits, s.ptr, edges * sizeof(float), hipMemcpyDeviceToHost, p->stream), "download");
        faco::checked(hipStreamSynchronize(p->stream), "prior inference");
    });
}

int faco_prior_gradients(faco_prior* prior, const float* features, const uint8_t* labels, uint32_t edges,
                         float pos_weight, float* grads, float* loss) {
    return faco::guarded([&] {
        faco::require(prior && features && labels && grads && edges, "null argument or no edges");
        faco::require(std::isfinite(pos_weight) && pos_weight > 0, "positive weight must be positive");
        auto* p = reinterpret_cast<PriorImpl*>(prior);
        faco::Buffer<float> f(static_cast<size_t>(edges) * 8);
        faco::Buffer<uint8_t> y(edges);
        const uint32_t rows = faco::prior_grad_rows(edges);
        faco::Buffer<float> partial(static_cast<size_t>(rows) * (faco::prior_param_count() + 1));
        faco::checked(hipMemcpyAsync(f.ptr, features, f.count * sizeof(float), hipMemcpyHostToDevice, p->stream), "upload");
        faco::checked(hipMemcpyAsync(y.ptr, labels, edges, hipMemcpyHostToDevice, p->stream), "upload");
        faco::checked(faco::launch_prior_grad(p->image.ptr, f.ptr, y.ptr, nullptr, 0, edges, pos_weight,
                                              1.0f / static_cast<float>(edges), partial.ptr, p->stream), "gradient launch");
        faco::checked(faco::launch_prior_reduce(partial.ptr, rows, p->grad.ptr, p->params.ptr, p->m.ptr, p->v.ptr, false,
                                                0, 0, 0, 0, 1, nullptr, p->image.ptr, p->stream), "reduce launch");
        std::vector<float> g(faco::prior_param_count() + 1);
        faco::checked(hipMemcpyAsync(g.data(), p->grad.ptr, g.size() * sizeof(float), hipMemcpyDeviceToHost, p->stream),
                      "download");
        faco::checked(hipStreamSynchronize(p->stream), "prior gradients");
        std::memcpy(grads, g.data(), faco::prior_param_count() * sizeof(float));
        if (loss) *loss = g.back();
    });
}

int faco_prior_gradients_noalloc(faco_prior* prior, const float* features, const uint8_t* labels, uint32_t edges,
                                 float pos_weight, float* grads, float* loss) {
    return faco::guarded([&] {
        faco::require(prior && features && labels && grads && edges, "null argument or no edges");
        faco::require(edges <= FACO_PRIOR_NOALLOC_EDGES,
                      "faco_prior_gradients_noalloc takes at most FACO_PRIOR_NOALLOC_EDGES (8192) edges");
        faco::require(std::isfinite(pos_weight) && pos_weight > 0, "positive weight must be positive");
        const uint32_t rows = faco::prior_grad_rows(edges), P = faco::prior_param_count();
        faco::require(rows <= faco::noalloc::kRows && P + 1 == faco::noalloc::kRow, "scratch layout out of date");
        auto* p = reinterpret_cast<PriorImpl*>(prior);
        float* base = faco::device_scratch();
        float* partial = base + faco::noalloc::kPartial;
        float* grad = base + faco::noalloc::kGrad;
        float* staged = base + faco::noalloc::kFeatures;
        auto* y = reinterpret_cast<uint8_t*>(base + faco::noalloc::kLabels);
        // every partial row the reduction reads and every gradient entry is
        // written by this call's kernels, so nothing needs clearing
        faco::checked(hipMemcpyAsync(staged, features, static_cast<size_t>(edges) * 8 * sizeof(float),
                                     hipMemcpyHostToDevice, p->stream), "upload");
        faco::checked(hipMemcpyAsync(y, labels, edges, hipMemcpyHostToDevice, p->stream), "upload");
        faco::checked(faco::launch_prior_grad(p->image.ptr, staged, y, nullptr, 0, edges, pos_weight,
                                              1.0f / static_cast<float>(edges), partial, p->stream), "gradient launch");
        faco::checked(faco::launch_prior_reduce(partial, rows, grad, p->params.ptr, p->m.ptr, p->v.ptr, false,
                                                0, 0, 0, 0, 1, nullptr, p->image.ptr, p->stream), "reduce launch");
        faco::checked(hipMemcpyAsync(grads, grad, P * sizeof(float), hipMemcpyDeviceToHost, p->stream), "download");
        if (loss) faco::checked(hipMemcpyAsync(loss, grad + P, sizeof(float), hipMemcpyDeviceToHost, p->stream), "download");
        faco::checked(hipStreamSynchronize(p->stream), "prior gradients");
    });
}

int faco_prior_train(faco_prior* prior, const float* features, const uint8_t* labels, uint32_t edges, uint32_t batch,
                     uint32_t epochs, float lr, float pos_weight, uint64_t shuffle_seed, float* epoch_losses) {
    return faco::guarded([&] {
        faco::require(prior && features && labels && edges, "null argument or no edges");
        faco::require(batch >= 1 && epochs >= 1, "batch and epochs must be positive");
        faco::require(std::isfinite(lr) && lr > 0 && std::isfinite(pos_weight) && pos_weight > 0,
                      "learning rate and positive weight must be positive");
        auto* p = reinterpret_cast<PriorImpl*>(prior);
        const uint32_t P = faco::prior_param_count();
        faco::Buffer<float> f(static_cast<size_t>(edges) * 8);
        faco::Buffer<uint8_t> y(edges);
        const uint32_t rows = faco::prior_grad_rows(batch < edges ? batch : edges);
        faco::Buffer<float> partial(static_cast<size_t>(rows) * (P + 1));
        faco::checked(hipMemcpyAsync(f.ptr, features, f.count * sizeof(float), hipMemcpyHostToDevice, p->stream), "upload");
        faco::checked(hipMemcpyAsync(y.ptr, labels, edges, hipMemcpyHostToDevice, p->stream), "upload");
        // Fisher-Yates shuffle of epoch e from a SplitMix64 stream of (seed,
        // e). The host draws epoch e + 1's order (pinned buffer, own thread)
        // while the device runs epoch e from the other device buffer.
        auto shuffle = [edges, shuffle_seed](uint32_t e, uint32_t* perm) {
            for (uint32_t i = 0; i < edges; ++i) perm[i] = i;
            faco::Mix mix{faco::mixed(shuffle_seed ^ faco::mixed(e + 0x2545f4914f6cdd1dULL))};
            for (uint32_t i = edges - 1; i > 0; --i) {
                const uint32_t j = static_cast<uint32_t>(mix.next() % (static_cast<uint64_t>(i) + 1));
                std::swap(perm[i], perm[j]);
            }
        };
        faco::Buffer<uint32_t> orders[2] = {faco::Buffer<uint32_t>(edges), faco::Buffer<uint32_t>(edges)};
        faco::PinnedBuffer<uint32_t> perms[2] = {faco::PinnedBuffer<uint32_t>(edges), faco::PinnedBuffer<uint32_t>(edges)};
        shuffle(0, perms[0].ptr);
        faco::Worker next;
        for (uint32_t e = 0; e < epochs; ++e) {
            next.join();
            faco::Buffer<uint32_t>& order = orders[e & 1u];
            faco::checked(hipMemcpyAsync(order.ptr, perms[e & 1u].ptr, edges * sizeof(uint32_t), hipMemcpyHostToDevice,
                                         p->stream), "upload");
            // the pinned buffer of epoch e - 1 is free once the device has
            // taken epoch e - 1's order (the stream synchronised after it)
            if (e + 1 < epochs) next.thread = std::thread(shuffle, e + 1, perms[(e + 1) & 1u].ptr);
            faco::checked(hipMemsetAsync(p->loss.ptr, 0, sizeof(float), p->stream), "memset");
            uint32_t batches = 0;
            for (uint32_t offset = 0; offset < edges; offset += batch, ++batches) {
                const uint32_t count = edges - offset < batch ? edges - offset : batch;
                faco::checked(faco::launch_prior_grad(p->image.ptr, f.ptr, y.ptr, order.ptr, offset, count, pos_weight,
                                                      1.0f / static_cast<float>(count), partial.ptr, p->stream),
                              "gradient launch");
                faco::checked(faco::launch_prior_reduce(partial.ptr, faco::prior_grad_rows(count), p->grad.ptr,
                                                        p->params.ptr, p->m.ptr, p->v.ptr, true, lr, 0.9f, 0.999f, 1e-8f,
                                                        ++p->step, p->loss.ptr, p->image.ptr, p->stream), "adam launch");
            }
            float total = 0;
            faco::checked(hipMemcpyAsync(&total, p->loss.ptr, sizeof(float), hipMemcpyDeviceToHost, p->stream), "download");
            faco::checked(hipStreamSynchronize(p->stream), "training epoch");
            if (epoch_losses) epoch_losses[e] = total / static_cast<float>(batches);
        }
    });
}

int faco_prior_mma_check(const float* A, const float* B, float* C) {
    return faco::guarded([&] {
        faco::require(A && B && C, "null argument");
        faco::Buffer<float> a(16 * 32), b(32 * 16), c(16 * 16);
        faco::checked(hipMemcpy(a.ptr, A, a.count * sizeof(float), hipMemcpyHostToDevice), "upload");
        faco::checked(hipMemcpy(b.ptr, B, b.count * sizeof(float), hipMemcpyHostToDevice), "upload");
        faco::checked(faco::launch_prior_mma_check(a.ptr, b.ptr, c.ptr, nullptr), "mma launch");
        faco::checked(hipMemcpy(C, c.ptr, c.count * sizeof(float), hipMemcpyDeviceToHost), "download");
    });
}

}  // extern "C"
